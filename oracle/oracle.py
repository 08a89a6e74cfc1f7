"""CPU oracle (test infrastructure only): ctypes access to the C restatement of the
reference host semantics (oracle/nca_cpu.c) plus FP64 numpy references.

Reference citations are to the ncarray source tree (src/lib/ncarray/...).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnca_cpu.so")


def build():
    subprocess.run(["make", "-C", HERE], check=True, capture_output=True)


if not os.path.exists(LIB_PATH):
    build()
lib = ctypes.CDLL(LIB_PATH)

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")

lib.nca_cpu_sum_f32.argtypes = [_f32p, ctypes.c_int64]
lib.nca_cpu_sum_f32.restype = ctypes.c_float
lib.nca_cpu_sum_f64.argtypes = [_f64p, ctypes.c_int64]
lib.nca_cpu_sum_f64.restype = ctypes.c_double
lib.nca_cpu_sum_i64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
lib.nca_cpu_sum_i64.restype = ctypes.c_int64
lib.nca_cpu_max.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_int64)]
lib.nca_cpu_max.restype = ctypes.c_double
lib.nca_cpu_min.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_int64)]
lib.nca_cpu_min.restype = ctypes.c_double
lib.nca_cpu_expr.argtypes = [_u32p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int),
                             ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
lib.nca_cpu_expr.restype = None
lib.nca_cpu_expr_host_jit.argtypes = lib.nca_cpu_expr.argtypes
lib.nca_cpu_expr_host_jit.restype = ctypes.c_int
lib.nca_cpu_reduce_axes_f32.argtypes = [_f32p, ctypes.c_int, _i64p, np.ctypeslib.ndpointer(np.intc), ctypes.c_int,
                                        ctypes.c_void_p]
lib.nca_cpu_reduce_axes_f32.restype = None
lib.nca_cpu_stencil_f32.argtypes = [_f32p, _i64p, _f32p, _i64p, ctypes.c_int, _i64p, ctypes.c_int, _u32p,
                                    ctypes.c_int, _f64p, ctypes.c_int]
lib.nca_cpu_stencil_f32.restype = None
lib.nca_cpu_bin_f32.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int64, _i32p, ctypes.c_int, ctypes.c_int, _i64p, _f64p]
lib.nca_cpu_bin_f32.restype = None
lib.nca_cpu_scatter_add_f32.argtypes = [_f32p, ctypes.c_int64, _i32p, _f32p, ctypes.c_int64]
lib.nca_cpu_scatter_add_f32.restype = None
lib.nca_cpu_bin_backward_f32.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, _i32p, ctypes.c_int64, _f32p]
lib.nca_cpu_bin_backward_f32.restype = None
lib.nca_cpu_laplacian_pipeline.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _f32p, _f32p,
                                           ctypes.c_int]
lib.nca_cpu_laplacian_pipeline.restype = ctypes.c_float
lib.nca_cpu_max_threads.restype = ctypes.c_int
lib.nca_cpu_f32_to_bf16.argtypes = [ctypes.c_float]
lib.nca_cpu_f32_to_bf16.restype = ctypes.c_uint16

lib.nca_cpu_expr_sum_mt.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_float), ctypes.c_int64, ctypes.c_int]
lib.nca_cpu_expr_sum_mt.restype = ctypes.c_float
lib.nca_cpu_bin_f32_mt.restype = None
lib.nca_cpu_laplacian_c0.restype = ctypes.c_double

# numpy dtype -> reference DType code
CODE = {np.dtype(np.bool_): 0, np.dtype(np.uint8): 2, np.dtype(np.uint16): 3, np.dtype(np.uint32): 4,
        np.dtype(np.uint64): 5, np.dtype(np.int8): 6, np.dtype(np.int16): 7, np.dtype(np.int32): 8,
        np.dtype(np.int64): 9, np.dtype(np.float32): 10, np.dtype(np.float64): 11}

LOWEST = {10: float(np.finfo(np.float32).min), 11: float(np.finfo(np.float64).min)}
HIGHEST = {10: float(np.finfo(np.float32).max), 11: float(np.finfo(np.float64).max)}


def _lim(code, lo):
    if code in LOWEST:
        return LOWEST[code] if lo else HIGHEST[code]
    info = np.iinfo({0: np.uint8, 2: np.uint8, 3: np.uint16, 4: np.uint32, 6: np.int8, 7: np.int16, 8: np.int32,
                     9: np.int64, 5: np.uint64}[code])
    if code == 0:
        return 0.0 if lo else 1.0
    return float(info.min if lo else info.max)


# ---- reductions -----------------------------------------------------------------------------------
# op_traits<T>::sum_type (op_traits.hh:75, 380-410): the type itself unless promoted
SUM_TYPE = {0: np.uint64, 2: np.uint64, 3: np.uint64, 4: np.uint32, 5: np.uint64, 6: np.int64, 7: np.int64,
            8: np.int32, 9: np.int64}


def sum_fold(x):
    """host::sum_recursive (host/reductions.hh:58-73): sequential fold in sum_type.  Integer
    sums wrap in sum_type (int32 and uint32 at 32 bits, the others at 64)."""
    x = np.ascontiguousarray(x).ravel()
    if x.dtype == np.float32:
        return np.float32(lib.nca_cpu_sum_f32(x, x.size))
    if x.dtype == np.float64:
        return lib.nca_cpu_sum_f64(x, x.size)
    c = CODE[x.dtype]
    wide = np.array([lib.nca_cpu_sum_i64(x.ctypes.data, c, x.size)], np.int64)
    return int(wide.astype(SUM_TYPE[c])[0])


# largest finite bfloat16 (0x7f7f) and float16: the max/min identities of those storage types
BF16_MAX = 3.3895313892515355e38
F16_MAX = 65504.0


def max_fold(x, lowest=None):
    """host::max_recursive + argmax_recursive (host/reductions.hh:130-169): the value by the
    max rule of nca_cpu_max (NaN never selected, -0 below +0), the index of the first element
    strictly above the running value.  `lowest` overrides the dtype's identity (for values of
    a storage type held in a wider array, e.g. bfloat16 as float64)."""
    x = np.ascontiguousarray(x).ravel()
    c = CODE[x.dtype]
    idx = ctypes.c_int64()
    v = lib.nca_cpu_max(x.ctypes.data, c, x.size, _lim(c, True) if lowest is None else lowest, ctypes.byref(idx))
    return v, idx.value


def min_fold(x, highest=None):
    """host::min_recursive + argmin_recursive (host/reductions.hh:171-212)."""
    x = np.ascontiguousarray(x).ravel()
    c = CODE[x.dtype]
    idx = ctypes.c_int64()
    v = lib.nca_cpu_min(x.ctypes.data, c, x.size, _lim(c, False) if highest is None else highest, ctypes.byref(idx))
    return v, idx.value


def reduce_axes_f32(x, axes, op):
    """HostEngine::execute_reduce_axes (engines/hostengine.hh:152-191) for float32 input;
    op 0 sum, 1 max, 2 min, 3 argmax, 4 argmin (linear input index)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    red = np.zeros(x.ndim, dtype=np.intc)
    for a in axes:
        if 0 <= a < x.ndim:
            red[a] = 1
    oshape = tuple(s for d, s in enumerate(x.shape) if not red[d])
    out = np.zeros(oshape, dtype=np.int64 if op >= 3 else np.float32)
    lib.nca_cpu_reduce_axes_f32(x, x.ndim, np.array(x.shape, dtype=np.int64), red, op, out.ctypes.data)
    return out


# ---- expressions ---------------------------------------------------------------------------------------
def expr(instrs, arrays, scalars, eval_dtype, out_dtype, host_jit=False):
    """Tape evaluation with the VM/apply_op semantics (dynamicmvnode.hh:146-251,
    interface.hh:129-181); eval_dtype 10 (float32) or 11 (float64).  host_jit=True restates
    the reference's host x86 JIT for linearizable tapes instead (jit/host/rtcompiler.cc:285-610,
    x86.cc:157-169): the same fold with IEEE division (x/0 = +-inf, 0/0 = NaN)."""
    arrays = [np.ascontiguousarray(a) for a in arrays]
    n = arrays[0].size if arrays else 0
    ptrs = (ctypes.c_void_p * max(1, len(arrays)))(*[a.ctypes.data for a in arrays])
    dts = (ctypes.c_int * max(1, len(arrays)))(*[CODE[a.dtype] for a in arrays])
    out = np.empty(n, dtype=out_dtype)
    sc = np.array(list(scalars) or [0.0], dtype=np.float64)
    args = (np.array(instrs, dtype=np.uint32), len(instrs), ptrs, dts, len(arrays), sc, eval_dtype,
            out.ctypes.data, CODE[np.dtype(out_dtype)], n)
    if host_jit:
        if lib.nca_cpu_expr_host_jit(*args) != 0:
            raise ValueError("the host JIT restatement evaluates float work types only (integer DIV is idiv)")
    else:
        lib.nca_cpu_expr(*args)
    return out.reshape(arrays[0].shape) if arrays else out


# ---- stencil ---------------------------------------------------------------------------------------------
def stencil_f32(src, dshape, offsets, instrs, scalars, zero_boundary=True):
    """dest[c] = f(src[c*step + off_k]) (rtcompiler.hh:339-376), float32 work type."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    dst = np.empty(dshape, dtype=np.float32)
    offs = np.ascontiguousarray(np.array(offsets, dtype=np.int64).ravel())
    sc = np.array(list(scalars) or [0.0], dtype=np.float64)
    lib.nca_cpu_stencil_f32(src, np.array(src.shape, dtype=np.int64), dst, np.array(dshape, dtype=np.int64),
                            src.ndim, offs, len(offsets), np.array(instrs, dtype=np.uint32), len(instrs), sc,
                            1 if zero_boundary else 0)
    return dst


def laplacian_fp64(x):
    """7-point Laplacian with zero Dirichlet in float64 (numpy slicing on np.pad)."""
    p = np.pad(np.asarray(x, dtype=np.float64), 1)
    c = p[1:-1, 1:-1, 1:-1]
    return (p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1]
            + p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:] - 6.0 * c)


def laplacian_pipeline(x, threads=1):
    """CPU baseline: Laplacian (float32, reference op order) -> axis-0 sum -> max."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    lap = np.empty_like(x)
    col = np.empty(x.shape[1:], dtype=np.float32)
    mx = lib.nca_cpu_laplacian_pipeline(x, x.shape[0], x.shape[1], x.shape[2], lap, col, threads)
    return lap, col, np.float32(mx)


# ---- binning ---------------------------------------------------------------------------------------------
def bin_f32(frames, binmap, nb, nonzero=True):
    """scatter_add per frame (kernels.cuh:660-680): exact counts, float64 sums."""
    frames = np.ascontiguousarray(frames, dtype=np.float32)
    F = frames.shape[0]
    P = binmap.size
    counts = np.zeros((F, nb), dtype=np.int64)
    sums = np.zeros((F, nb), dtype=np.float64)
    lib.nca_cpu_bin_f32(frames.reshape(F, P), F, P, np.ascontiguousarray(binmap, dtype=np.int32).ravel(), nb,
                        1 if nonzero else 0, counts, sums)
    return counts, sums


def bin_backward_f32(gsums, binmap):
    gsums = np.ascontiguousarray(gsums, dtype=np.float32)
    F, nb = gsums.shape
    m = np.ascontiguousarray(binmap, dtype=np.int32).ravel()
    out = np.empty((F, m.size), dtype=np.float32)
    lib.nca_cpu_bin_backward_f32(gsums, F, nb, m, m.size, out)
    return out


def scatter_add_f32(dest, idx, src):
    dest = np.array(dest, dtype=np.float32, copy=True)
    lib.nca_cpu_scatter_add_f32(dest, dest.size, np.ascontiguousarray(idx, dtype=np.int32),
                                np.ascontiguousarray(src, dtype=np.float32), src.size)
    return dest


def radial_binmap(ny=1024, nx=1024, cy=512, cx=512, nb=1000):
    """Detector bin map: r = hypot(y-cy, x-cx) in float64, bin = min(floor(r*nb/rmax), nb-1)."""
    y, x = np.mgrid[0:ny, 0:nx]
    r = np.hypot(y - cy, x - cx)
    rmax = np.hypot(cy, cx)
    return np.minimum(np.floor(r * nb / rmax), nb - 1).astype(np.int32)


def bf16_round(x):
    """float32 -> bfloat16 (round to nearest even) -> float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32)
    return (r << 16).view(np.float32).reshape(np.shape(x))


def ulp_diff_f32(a, b):
    """|a - b| in units in the last place of float32 (monotone integer mapping)."""
    ia = np.ascontiguousarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    ib = np.ascontiguousarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def so_offsets(image, image_addr, base_addr, shape, strides, suboffsets):
    """Byte offsets into `image` of every element of an SO view, in C order.

    Restates SOArrayPolicy::advance (layout.hh:677-700) over a byte image of device memory:
    starting from base_addr, for each axis d the address moves by i_d * strides[d]; when
    suboffsets[d] >= 0 it is then replaced by the 8-byte pointer stored there plus
    suboffsets[d].  image_addr is the device address of image[0]; every address the view
    reaches must fall inside the image.
    """
    img = np.frombuffer(np.ascontiguousarray(image).tobytes(), dtype=np.uint8)
    off = np.array(int(base_addr) - int(image_addr), dtype=np.int64)
    for d, n in enumerate(shape):
        off = off[..., None] + np.arange(n, dtype=np.int64) * int(strides[d])
        so = int(suboffsets[d]) if d < len(suboffsets) else -1
        if so >= 0:
            assert off.size == 0 or (off.min() >= 0 and off.max() + 8 <= img.size), "pointer read outside the image"
            ptr = img[off[..., None] + np.arange(8)].copy().view("<u8")[..., 0].astype(np.int64)
            off = ptr - int(image_addr) + so
    return off.reshape(tuple(shape))


def nc_offsets(image, image_addr, base_addr, shape, strides, pointer_axis=-1, ptr_offset=0, row_offset=0):
    """Byte offsets into `image` of every element of an NC view in the device view form, C order.

    Without a pointer axis, element (i_0, ..., i_{n-1}) is at base + sum_d i_d * strides[d].  With
    pointer axis P it is at row + row_offset + sum_{d > P} i_d * strides[d], where row is the
    8-byte pointer at entry i_P * strides[P] + ptr_offset of the table at
    base + sum_{d < P} i_d * strides[d].  This is NCOffsetsPolicy::advance (layout.hh:468-476)
    with the reference's per-axis byte offsets folded in: those of the axes before P (all of
    them without a pointer axis) into base, offsets[P] into ptr_offset (table entries) and those
    of the axes after P into row_offset; strides[P] is 1 for the reference's own views.
    """
    img = np.frombuffer(np.ascontiguousarray(image).tobytes(), dtype=np.uint8)
    off = np.array(int(base_addr) - int(image_addr), dtype=np.int64)
    for d, n in enumerate(shape):
        if d == pointer_axis:
            ent = off[..., None] + (np.arange(n, dtype=np.int64) * int(strides[d]) + int(ptr_offset)) * 8
            assert ent.size == 0 or (ent.min() >= 0 and ent.max() + 8 <= img.size), "pointer read outside the image"
            ptr = img[ent[..., None] + np.arange(8)].copy().view("<u8")[..., 0].astype(np.int64)
            off = ptr - int(image_addr) + int(row_offset)
        else:
            off = off[..., None] + np.arange(n, dtype=np.int64) * int(strides[d])
    return off.reshape(tuple(shape))


def so_gather(image, image_addr, base_addr, dtype, shape, strides, suboffsets):
    """The elements of an SO view (so_offsets) as a C-ordered numpy array."""
    dt = np.dtype(dtype)
    img = np.frombuffer(np.ascontiguousarray(image).tobytes(), dtype=np.uint8)
    off = so_offsets(img, image_addr, base_addr, shape, strides, suboffsets)
    assert off.size == 0 or (off.min() >= 0 and off.max() + dt.itemsize <= img.size), "element outside the image"
    raw = img[off[..., None] + np.arange(dt.itemsize)].copy()
    return raw.view(dt)[..., 0].reshape(tuple(shape))


def so_scatter(image, image_addr, base_addr, values, strides, suboffsets):
    """A copy of `image` with `values` stored at the elements of the SO view of values.shape
    (so_offsets), in C order (a later element overwrites an earlier one at the same address)."""
    vals = np.ascontiguousarray(values)
    img = np.frombuffer(np.ascontiguousarray(image).tobytes(), dtype=np.uint8).copy()
    off = so_offsets(img, image_addr, base_addr, vals.shape, strides, suboffsets).ravel()
    raw = vals.reshape(-1, 1).view(np.uint8).reshape(off.size, vals.dtype.itemsize)
    for i in range(off.size):
        img[off[i]:off[i] + vals.dtype.itemsize] = raw[i]
    return img


class ByteImage:
    """A byte image of one device allocation at address `addr`, for building indirect-view test
    inputs: chunks are placed at aligned offsets and pointers are written as addr + offset, so
    the image copied to that allocation is valid device memory for the views built over it."""

    def __init__(self, size, addr):
        self.buf = np.zeros(int(size), np.uint8)
        self.addr = int(addr)
        self.top = 0

    def alloc(self, nbytes, align=16):
        off = (self.top + align - 1) // align * align
        assert off + nbytes <= self.buf.size, "image too small"
        self.top = off + int(nbytes)
        return off

    def put(self, arr, align=16):
        """Copy `arr` (C order) into the image; returns its offset."""
        raw = np.ascontiguousarray(arr).reshape(-1).view(np.uint8)
        off = self.alloc(raw.size, align)
        self.buf[off:off + raw.size] = raw
        return off

    def put_ptrs(self, offsets, deltas=0):
        """A table of 8-byte pointers to image offsets (plus `deltas` bytes each); returns its offset."""
        vals = np.asarray(offsets, np.int64) + np.asarray(deltas, np.int64) + self.addr
        return self.put(vals.astype("<u8"), align=8)


# ---- CPU baselines (bench.py times these; threads > 1 uses OpenMP) ------------------------------
def expr_sum_mt(instrs, arrays, scalars, eval_dtype, threads=1):
    """(out float32, float32 sum) of a tape over contiguous same-size arrays."""
    arrays = [np.ascontiguousarray(a) for a in arrays]
    n = arrays[0].size
    ptrs = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    dts = (ctypes.c_int * len(arrays))(*[CODE[a.dtype] for a in arrays])
    out = np.empty(n, dtype=np.float32)
    ins = np.array(instrs, dtype=np.uint32)
    sc = np.array(list(scalars) or [0.0], dtype=np.float64)
    s = lib.nca_cpu_expr_sum_mt(ins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(instrs), ptrs, dts,
                                len(arrays), sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), eval_dtype,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, threads)
    return out, np.float32(s)


def bin_f32_mt(frames, binmap, nb, nonzero=True, threads=1):
    frames = np.ascontiguousarray(frames, dtype=np.float32)
    F = frames.shape[0]
    P = binmap.size
    counts = np.zeros((F, nb), dtype=np.int64)
    sums = np.zeros((F, nb), dtype=np.float64)
    lib.nca_cpu_bin_f32_mt(frames.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(F), ctypes.c_int64(P),
                           np.ascontiguousarray(binmap, dtype=np.int32).ctypes.data_as(ctypes.c_void_p), nb,
                           1 if nonzero else 0, counts.ctypes.data_as(ctypes.c_void_p),
                           sums.ctypes.data_as(ctypes.c_void_p), threads)
    return counts, sums


def laplacian_c0(x, y, with_grad=True, threads=1):
    """(loss, dx, dy) of sum(C0(Laplacian(x), y)) (dx, dy None without with_grad)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    a = np.empty_like(x)
    dx = np.empty_like(x) if with_grad else a
    dy = np.empty_like(x) if with_grad else a
    f = lambda t: t.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    loss = lib.nca_cpu_laplacian_c0(f(x), f(y), ctypes.c_int64(x.shape[0]), ctypes.c_int64(x.shape[1]),
                                    ctypes.c_int64(x.shape[2]), f(a), f(dx), f(dy), 1 if with_grad else 0, threads)
    return (loss, dx, dy) if with_grad else (loss, None, None)
