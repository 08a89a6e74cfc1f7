"""Full and axis reductions on the GPU through the C ABI vs the host fold (oracle) and FP64."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _c():
    from ncarray_amd import capi, torch_interop as ti
    return capi, ti


def full(t, op, mode=0):
    capi, ti = _c()
    code = ti.TORCH_TO_CODE[t.dtype]
    out_code = capi.reduce_result_dtype(op, code)
    out = torch.empty((), dtype=ti.CODE_TO_TORCH[out_code], device=t.device)
    capi.call("nca_hip_reduce_full", capi.ctypes.byref(ti.view(t)), op, mode, out.data_ptr(), None)
    return out.cpu()


def axes(t, ax, op, mode=0):
    capi, ti = _c()
    code = ti.TORCH_TO_CODE[t.dtype]
    out_code = capi.reduce_result_dtype(op, code)
    oshape = [s for d, s in enumerate(t.shape) if d not in ax]
    out = torch.empty(oshape, dtype=ti.CODE_TO_TORCH[out_code], device=t.device)
    arr = (capi.ctypes.c_int64 * max(1, len(ax)))(*ax)
    capi.call("nca_hip_reduce_axes", capi.ctypes.byref(ti.view(t)), arr, len(ax), op, mode, capi.ctypes.byref(ti.view(out)), None)
    return out.cpu()


def test_reference_kats(dev):
    capi, ti = _c()
    # test_gpu_reductions.cu:39-55: float iota 256 -> sum 32640, max 255, min 0
    x = torch.arange(256, dtype=torch.float32, device=dev)
    assert full(x, capi.REDUCE_SUM).item() == 32640
    assert full(x, capi.REDUCE_MAX).item() == 255
    assert full(x, capi.REDUCE_MIN).item() == 0
    # test_reductions.cc:26-38: int32 fill 2 x 10 -> sum 20 (int32 sum_type)
    y = torch.full((10,), 2, dtype=torch.int32, device=dev)
    r = full(y, capi.REDUCE_SUM)
    assert r.dtype == torch.int32 and r.item() == 20
    assert full(y, capi.REDUCE_MAX).item() == 2 and full(y, capi.REDUCE_MIN).item() == 2
    # test_reductions.cc:40-88: (iota+1) 2x3 -> axis sums [5,7,9] and [6,15], max/min
    z = torch.arange(1, 7, dtype=torch.int32, device=dev).reshape(2, 3)
    assert axes(z, [0], capi.REDUCE_SUM).tolist() == [5, 7, 9]
    assert axes(z, [1], capi.REDUCE_SUM).tolist() == [6, 15]
    assert axes(z, [0], capi.REDUCE_MAX).tolist() == [4, 5, 6]
    assert axes(z, [1], capi.REDUCE_MIN).tolist() == [1, 4]
    # test_gpu_operations.py:41-67: uint32 0..255 -> 32640 / 255 / 0 (device == host)
    if hasattr(torch, "uint32"):
        u = torch.arange(256, dtype=torch.int64, device=dev).to(torch.uint32)
        assert full(u, capi.REDUCE_SUM).item() == 32640 and full(u, capi.REDUCE_MAX).item() == 255
    # tests/python/test_operations.py:56-72: [[1,2,3],[4,5,6]] -> 21 / 6 / 1, mean([1,2,3]) = 2
    w = torch.tensor([[1, 2, 3], [4, 5, 6]], dtype=torch.int32, device=dev)
    assert full(w, capi.REDUCE_SUM).item() == 21
    m = full(torch.tensor([1.0, 2.0, 3.0], device=dev), capi.REDUCE_MEAN)
    assert m.dtype == torch.float64 and m.item() == 2.0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.int32, torch.int64, torch.int16,
                                   torch.int8, torch.uint8, torch.bool, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("n", [1, 7, 1000, 123457])
def test_extrema_bit_exact(dev, dtype, n):
    """max/min/argmax/argmin == the host fold (first occurrence), for many dtypes/sizes."""
    capi, ti = _c()
    rng = np.random.default_rng(n)
    if dtype == torch.bool:
        x = torch.from_numpy(rng.integers(0, 2, n).astype(np.bool_))
    elif dtype.is_floating_point:
        x = torch.from_numpy(rng.integers(-50, 50, n).astype(np.float32)).to(dtype)  # many ties
    else:
        x = torch.from_numpy(rng.integers(-50 if dtype != torch.uint8 else 0, 50, n)).to(dtype)
    xd = x.to(dev)
    xn = x.float().numpy() if dtype in (torch.bfloat16, torch.float16) else x.numpy()
    vmax, imax = O.max_fold(xn.astype(np.float64) if dtype in (torch.bfloat16, torch.float16) else xn)
    vmin, imin = O.min_fold(xn.astype(np.float64) if dtype in (torch.bfloat16, torch.float16) else xn)
    assert float(full(xd, capi.REDUCE_MAX).float()) == vmax
    assert float(full(xd, capi.REDUCE_MIN).float()) == vmin
    assert full(xd, capi.REDUCE_ARGMAX).item() == imax == int(np.argmax(xn))
    assert full(xd, capi.REDUCE_ARGMIN).item() == imin == int(np.argmin(xn))


def test_nan_and_edge_semantics(dev):
    capi, ti = _c()
    x = torch.tensor([float("nan"), 1.0, 3.0, float("nan"), 3.0, -2.0], device=dev)
    assert full(x, capi.REDUCE_MAX).item() == 3.0       # NaN never selected (strict >)
    assert full(x, capi.REDUCE_ARGMAX).item() == 2      # first occurrence
    assert full(x, capi.REDUCE_ARGMIN).item() == 5
    allnan = torch.full((5,), float("nan"), device=dev)
    assert full(allnan, capi.REDUCE_MAX).item() == float(np.finfo(np.float32).min)  # lowest()
    assert full(allnan, capi.REDUCE_ARGMAX).item() == 0
    neginf = torch.full((4,), -float("inf"), device=dev)
    assert full(neginf, capi.REDUCE_MAX).item() == float(np.finfo(np.float32).min)
    empty = torch.empty(0, device=dev)
    assert full(empty, capi.REDUCE_SUM).item() == 0.0
    assert np.isnan(full(empty, capi.REDUCE_MEAN).item())
    b = torch.tensor([True, True, False], device=dev)
    assert full(b, capi.REDUCE_ALL).item() is False and full(b, capi.REDUCE_ANY).item() is True


@pytest.mark.parametrize("mode", [0, 1])
def test_sum_accuracy(dev, mode):
    """FP32 sums vs FP64: |S - S64| <= c * eps * sum|x| (c=8 plain, 2 compensated)."""
    capi, ti = _c()
    rng = np.random.default_rng(7)
    x = rng.standard_normal(1 << 24).astype(np.float32) + 3.0
    s = full(torch.from_numpy(x).to(dev), capi.REDUCE_SUM, mode).item()
    s64 = x.astype(np.float64).sum()
    bound = (8e-7 if mode == 0 else 1.2e-7) * np.abs(x.astype(np.float64)).sum()
    assert abs(s - s64) <= bound
    # deterministic: same bits twice
    assert full(torch.from_numpy(x).to(dev), capi.REDUCE_SUM, mode).item() == s


def test_integer_sum_types(dev):
    capi, ti = _c()
    x8 = torch.full((1000,), 100, dtype=torch.int8, device=dev)
    r = full(x8, capi.REDUCE_SUM)
    assert r.dtype == torch.int64 and r.item() == 100000     # small ints promote (op_traits.hh:380-410)
    xb = torch.ones(300, dtype=torch.bool, device=dev)
    assert full(xb, capi.REDUCE_SUM).item() == 300
    x32 = torch.full((3,), 2 ** 30, dtype=torch.int32, device=dev)
    assert full(x32, capi.REDUCE_SUM).item() == (3 * 2 ** 30) - 2 ** 32  # int32 sum_type wraps


def test_strided_full(dev):
    capi, ti = _c()
    base = torch.arange(60, dtype=torch.float32, device=dev).reshape(3, 4, 5)
    v = base.permute(2, 0, 1)[:, 1:, ::2]
    vn = v.cpu().numpy()
    assert full(v, capi.REDUCE_SUM).item() == vn.sum()
    assert full(v, capi.REDUCE_ARGMAX).item() == int(np.argmax(vn))


@pytest.mark.parametrize("shape,ax", [((64, 96), [0]), ((64, 96), [1]), ((16, 33, 40), [1]),
                                      ((16, 33, 40), [0, 2]), ((5, 6, 7, 8), [1, 2]), ((7, 1, 9), [0, 1, 2]),
                                      ((256, 4096), [0]), ((3, 5), [7, -1])])
def test_axes_vs_oracle(dev, shape, ax):
    capi, ti = _c()
    rng = np.random.default_rng(len(shape))
    x = rng.integers(-20, 20, shape).astype(np.float32)   # exact sums, ties for argmax
    xd = torch.from_numpy(x).to(dev)
    valid = [a for a in ax if 0 <= a < len(shape)]
    for op, oop in ((capi.REDUCE_SUM, 0), (capi.REDUCE_MAX, 1), (capi.REDUCE_MIN, 2), (capi.REDUCE_ARGMAX, 3),
                    (capi.REDUCE_ARGMIN, 4)):
        got = axes(xd, ax, op).numpy()
        ref = O.reduce_axes_f32(x, ax, oop)
        assert np.array_equal(got, ref), (op, shape, ax)
    # argmax along axes returns the linear input index of the first maximum
    if len(valid) == 1:
        am = axes(xd, valid, capi.REDUCE_ARGMAX).numpy()
        flat = np.arange(x.size).reshape(shape)
        pos = np.argmax(x, axis=valid[0])
        assert np.array_equal(am, np.take_along_axis(flat, np.expand_dims(pos, valid[0]), valid[0]).squeeze(valid[0]))


def test_axis0_sum_512cube(dev):
    """The C1 reduction at full size: sum over axis 0 of 512^3 vs FP64."""
    capi, ti = _c()
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn((512, 512, 512), generator=g, device=dev, dtype=torch.float32)
    xn = x.cpu().numpy()
    ref = xn.sum(axis=0, dtype=np.float64)
    absum = np.abs(xn).sum(axis=0, dtype=np.float64)
    for mode in (0, 1):
        got = axes(x, [0], capi.REDUCE_SUM, mode).numpy().astype(np.float64)
        err = (np.abs(got - ref) / absum).max()
        assert err <= (1e-6 if mode == 0 else 2e-7), err
    mx = full(x, capi.REDUCE_MAX).item()
    assert mx == float(xn.max())
    del x


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16, torch.float64])
def test_compensated_sums_of_infinities(dev, dtype):
    """A compensated sum that meets +inf is +inf and one that meets +inf and -inf is NaN, as the
    plain sum (the compensation term of an infinite sum is inf - inf and must not leak)."""
    capi, ti = _c()
    n = 100003
    base = torch.linspace(-1, 1, n, device=dev, dtype=torch.float32).to(dtype)
    x1 = base.clone()
    x1[777] = float("inf")
    x2 = x1.clone()
    x2[90000] = float("-inf")
    for mode in (capi.ACCUM_PLAIN, capi.ACCUM_COMPENSATED):
        assert full(x1, capi.REDUCE_SUM, mode).item() == float("inf"), mode
        assert np.isnan(full(x2, capi.REDUCE_SUM, mode).item()), mode
        assert np.isnan(full(x1 - x1, capi.REDUCE_SUM, mode).item()), mode  # inf - inf
        # axis sums: columns 0 (inf), 1 (finite), 2 (inf and -inf)
        m = base[:3000].reshape(1000, 3).clone()
        m[10, 0] = float("inf")
        m[20, 2] = float("inf")
        m[30, 2] = float("-inf")
        s = axes(m.contiguous(), [0], capi.REDUCE_SUM, mode).float().numpy()
        assert s[0] == np.inf and np.isfinite(s[1]) and np.isnan(s[2]), (mode, s)


def test_fused_compensated_sums_of_infinities(dev):
    """The same through the fused expression reduction and the march stencil's axis-0 and global sums."""
    from ncarray_amd import tapes
    capi, ti = _c()
    a = torch.linspace(-1, 1, 1 << 16, device=dev)
    a[5] = float("inf")
    b = torch.zeros_like(a)
    out = torch.empty_like(a)
    s = torch.empty((), device=dev)
    P = capi.pack
    tape = capi.Tape([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_ADD)], [capi.FLOAT32] * 2, [],
                     capi.FLOAT32, capi.FLOAT32)
    tape.launch_reduce([ti.view(a), ti.view(b)], ti.view(out), capi.FUSE_SUM, capi.ACCUM_COMPENSATED, s.data_ptr(),
                       None, None)
    torch.cuda.synchronize()
    assert s.item() == float("inf")
    # march stencil (identity tape on one offset): the column through the inf sums to inf, others stay finite
    x = torch.rand((64, 16, 64), device=dev)
    x[7, 3, 5] = float("inf")
    st = capi.Stencil([[0, 0, 0]], [P(capi.OP_LOAD_NCARR, 0)], capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    y = torch.empty_like(x)
    col = torch.empty((16, 64), device=dev)
    tot = torch.empty((), device=dev)
    st.apply_reduce(ti.view(x), ti.view(y), capi.BOUNDARY_ZERO, axis0=ti.view(col), fuse=capi.FUSE_SUM,
                    mode=capi.ACCUM_COMPENSATED, sum_out=tot.data_ptr())
    torch.cuda.synchronize()
    c = col.cpu().numpy()
    assert c[3, 5] == np.inf and np.isfinite(np.delete(c.ravel(), 3 * 64 + 5)).all()
    assert tot.item() == float("inf")


def test_beyond_2_31_elements(dev):
    """64-bit element indexing (the reference's own element index is a 32-bit unsigned,
    array_impl.hh:2101-2179): on 2^31 + 1000 uint8 elements the full reductions, argmax/argmin,
    a fused expression and its fused sum all reach the elements past 2^31."""
    from ncarray_amd import tapes  # noqa: F401
    capi, ti = _c()
    n = (1 << 31) + 1000
    hi = (1 << 31) + 3
    x = torch.zeros(n, dtype=torch.uint8, device=dev)
    x[5] = 3
    x[hi] = 9
    x[-1] = 7
    assert full(x, capi.REDUCE_SUM).item() == 19
    assert full(x, capi.REDUCE_MAX).item() == 9
    assert full(x, capi.REDUCE_ARGMAX).item() == hi
    y = torch.ones(n, dtype=torch.uint8, device=dev)
    y[(1 << 31) + 500] = 0
    assert full(y, capi.REDUCE_ARGMIN).item() == (1 << 31) + 500
    del y
    P = capi.pack
    tape = capi.Tape([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_ADD)], [capi.UINT8] * 2, [],
                     capi.UINT8, capi.UINT8)
    out = torch.empty_like(x)
    s = torch.zeros((), dtype=torch.int64, device=dev)  # uint64 sum, read as int64 (small value)
    tape.launch_reduce([ti.view(x), ti.view(x)], ti.view(out), capi.FUSE_SUM, capi.ACCUM_PLAIN, s.data_ptr(), None,
                       None)
    torch.cuda.synchronize()
    assert s.item() == 38
    assert out[5].item() == 6 and out[hi].item() == 18 and out[-1].item() == 14 and out[hi + 1].item() == 0
