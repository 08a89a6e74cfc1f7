"""ncarray_amd: MI355X (gfx950) device backend for ncarray's hot path.

Fused elementwise expressions (hipRTC), stencils, axis and full reductions and binning,
behind a C ABI (include/nca_hip.h, libnca_hip.so), a C++ host API (host/ncarray_amd.hh,
libnca_host.so) and the reference's Python array surface:

    import ncarray_amd as nca
    a = nca.NCDevArray([4, 4, 4], nca.DType.float32)
    a.fill(2.0)
    b = a * a + 1.0            # eager: one kernel per operator
    with nca.lazy_mode():
        e = a * b - 0.5 * a    # lazy: one fused kernel when materialised
    c = nca.materialize(e)

The array names are loaded on first use, so that the build step can import this package
before the native modules exist; there is no fallback when they are missing.
"""
__version__ = "0.1.0"

_ARRAY_NAMES = {"DType", "OpCode", "Boundary", "Accum", "NCDevArray", "DevExpr", "Stencil", "NcaError", "NCArray",
                "NCArrayRef", "NCDevArrayRef", "set_eager", "is_eager", "materialize", "lazy_mode", "exp", "log", "sqrt", "abs",
                "init", "from_torch", "from_dlpack", "numpy_dtype"}


def __getattr__(name):
    if name in _ARRAY_NAMES:
        from . import arrays
        return getattr(arrays, name)
    raise AttributeError(f"module 'ncarray_amd' has no attribute {name!r}")
