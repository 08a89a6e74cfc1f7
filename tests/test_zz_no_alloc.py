"""The allocation-free fused stencil entry point (nca_hip_stencil_apply_reduce_noalloc) against
the ordinary one (nca_hip_stencil_apply_reduce), through the C ABI: bit-identical stencil
values, axis-0 column sums and max/sum, on consecutive calls whose second input is smaller (the
fixed scratch array keeps the previous call's contents), on the march and the tile kernels,
at the full 512^3 size, and the error for a kernel that was never loaded."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _c():
    from ncarray_amd import capi, tapes, torch_interop as ti
    return capi, tapes, ti


def _lap(capi, tapes):
    return capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                        [capi.FLOAT32])


def _call(name, st, x, axis0, fuse, stream=None):
    """One fused call (on the torch stream `stream`, or synchronous); returns (lap, col, sum,
    max) as fresh tensors, filled with NaN on torch's current stream before the call."""
    capi, tapes, ti = _c()
    lap = torch.full_like(x, float("nan"))
    col = torch.full(x.shape[1:], float("nan"), dtype=torch.float64, device=x.device) if axis0 else None
    s = torch.full((), float("nan"), dtype=torch.float32, device=x.device)
    m = torch.full((), float("nan"), dtype=torch.float32, device=x.device)
    if stream is not None:
        stream.wait_stream(torch.cuda.current_stream())   # the fills come first
        stream = ti.stream_ptr(stream)
    sa = capi.scalar_array((6.0,), st.scalar_dtypes)
    vx, vl = ti.view(x), ti.view(lap)
    vc = ti.view(col) if axis0 else None
    capi.call(name, st.handle, ctypes.byref(vx), None, 0, sa, 1, ctypes.byref(vl), capi.BOUNDARY_ZERO,
              ctypes.byref(vc) if axis0 else None, fuse, capi.ACCUM_COMPENSATED,
              s.data_ptr() if fuse & capi.FUSE_SUM else None, m.data_ptr() if fuse & capi.FUSE_MAX else None, None,
              stream)
    torch.cuda.synchronize()
    return lap, col, s, m


def _same(a, b):
    return a is None and b is None or torch.equal(a.view(torch.int32) if a.dtype == torch.float32 else
                                                  a.view(torch.int64), b.view(torch.int32) if b.dtype == torch.float32
                                                  else b.view(torch.int64))


@pytest.mark.parametrize("shapes,axis0", [([(64, 96, 128), (20, 40, 64)], True),    # march kernel
                                          ([(48, 70, 62), (9, 11, 14)], False)],    # tile kernel
                         ids=["march", "tile"])
def test_noalloc_matches_ordinary_entry(dev, shapes, axis0):
    capi, tapes, ti = _c()
    st = _lap(capi, tapes)
    fuse = capi.FUSE_SUM | capi.FUSE_MAX
    for k, shape in enumerate(shapes):             # the second call is smaller
        print(f"noalloc case shape={shape} axis0={axis0}", flush=True)
        x = torch.from_numpy(np.random.default_rng(k).standard_normal(shape).astype(np.float32)).to(dev)
        ref = _call("nca_hip_stencil_apply_reduce", st, x, axis0, fuse)
        for rep in range(2):
            got = _call("nca_hip_stencil_apply_reduce_noalloc", st, x, axis0, fuse)
            for a, b in zip(got, ref):
                assert _same(a, b), (shape, rep)
        # and the stencil values against the oracle
        want = O.stencil_f32(x.cpu().numpy(), shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
        assert np.array_equal(got[0].cpu().numpy().view(np.uint32), want.view(np.uint32))


def test_noalloc_on_a_stream_and_full_size(dev):
    """The C1 step (512^3 float32, Laplacian + axis-0 sum + max) on a torch stream, then a
    smaller grid, each bit-identical to the ordinary entry point."""
    capi, tapes, ti = _c()
    st = _lap(capi, tapes)
    side = torch.cuda.Stream()
    for shape in [(512, 512, 512), (128, 64, 256)]:
        print(f"noalloc stream case shape={shape}", flush=True)
        g = torch.Generator(device=dev).manual_seed(0)
        x = torch.randn(shape, generator=g, device=dev, dtype=torch.float32)
        torch.cuda.synchronize()
        ref = _call("nca_hip_stencil_apply_reduce", st, x, True, capi.FUSE_MAX, stream=side)
        got = _call("nca_hip_stencil_apply_reduce_noalloc", st, x, True, capi.FUSE_MAX, stream=side)
        for a, b in zip(got, ref):
            assert _same(a, b), shape
        del x, ref, got
    torch.cuda.empty_cache()


def test_noalloc_needs_a_loaded_kernel(dev):
    capi, tapes, ti = _c()
    st = capi.Stencil([[0, 0, 0], [1, 0, 0]], [capi.pack(capi.OP_LOAD_NCARR, 0), capi.pack(capi.OP_LOAD_NCARR, 1),
                                               capi.pack(capi.OP_SUB)], capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    x = torch.zeros((8, 8, 64), device=dev)
    out = torch.empty_like(x)
    m = torch.empty((), device=dev)
    with pytest.raises(capi.NcaError) as e:
        capi.call("nca_hip_stencil_apply_reduce_noalloc", st.handle, ctypes.byref(ti.view(x)), None, 0, None, 0,
                  ctypes.byref(ti.view(out)), capi.BOUNDARY_ZERO, None, capi.FUSE_MAX, 0, None, m.data_ptr(), None,
                  None)
    assert e.value.code == capi.ERR_INVALID and "not loaded" in str(e.value)
    st.apply_reduce(ti.view(x), ti.view(out), capi.BOUNDARY_ZERO, fuse=capi.FUSE_MAX, max_out=m.data_ptr())
    capi.call("nca_hip_stencil_apply_reduce_noalloc", st.handle, ctypes.byref(ti.view(x)), None, 0, None, 0,
              ctypes.byref(ti.view(out)), capi.BOUNDARY_ZERO, None, capi.FUSE_MAX, 0, None, m.data_ptr(), None, None)
    torch.cuda.synchronize()
    assert float(m) == 0.0
