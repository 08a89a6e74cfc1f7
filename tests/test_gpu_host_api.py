"""GPU tests of the C++ host API (tests/cpp/test_host_api) and of the Python array surface.

The first group follows the reference's Python GPU tests (tests/python/test_gpu_operations.py:
16-156) with `import ncarray_amd as nca` in place of `import ncarray as nca`; the rest check
the fused lazy path against the CPU oracle, axis reductions against numpy (bit-exact for
max/min/argmax, FP64-bounded for sums), torch interop and stencils.
"""
import os
import subprocess

import numpy as np
import pytest
import torch

import ncarray_amd as nca
from ncarray_amd import capi, tapes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(autouse=True)
def eager(dev):
    nca.set_eager(True)
    yield
    nca.set_eager(True)


def test_cpp_host_api():
    exe = os.path.join(ROOT, "tests", "cpp", "test_host_api")
    assert os.path.exists(exe), "tests/cpp/test_host_api is built by ncarray_amd/build.py"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-4000:]


# ---- reference Python GPU tests -------------------------------------------------------------------
def test_gpu_binary_arithmetic():
    shape = [4, 4, 4]
    a1 = nca.NCDevArray(shape, nca.DType.float32)
    a2 = nca.NCDevArray(shape, nca.DType.float32)
    a1.fill(2.0)
    a2.fill(4.0)
    s = a1 + a2
    p = a1 * a2
    assert np.allclose(np.ones(shape, dtype=np.float32) * 6.0, s.to_host())
    assert np.allclose(np.ones(shape, dtype=np.float32) * 8.0, p.to_host())


def test_gpu_reductions():
    h = nca.NCArray([256], nca.DType.uint32)
    for i in range(256):
        h[i] = i
    d = h.to_device()
    assert d.sum() == 32640
    assert d.max() == 255
    assert d.min() == 0
    assert d.sum() == h.sum() and d.max() == h.max() and d.min() == h.min()


def test_gpu_copies():
    data = np.random.rand(10, 10).astype(np.float32)
    d = nca.NCArrayRef(data).to_device()
    assert d.dtype == nca.DType.float32
    assert np.array_equal(np.array(d.to_host()), data)
    assert np.array_equal(np.array(d), data)


def test_gpu_inplace_arithmetic():
    d = nca.NCDevArray([100])
    d.fill(10.0)
    d += 5.0
    d *= 2.0
    assert np.allclose(np.array(d), 30.0)


def test_gpu_sliced_assignment():
    d = nca.NCDevArray([10, 10], dtype=nca.DType.float32)
    d.fill(0.0)
    d[1:3, 1:3] = 5.0
    d[7:9, 7:9] = np.ones((2, 2), dtype=np.float32) * 10.0
    h = np.array(d)
    assert h[0, 0] == 0.0
    assert np.allclose(h[1:3, 1:3], 5.0)
    assert np.allclose(h[7:9, 7:9], 10.0)
    assert h.sum() == 4 * 5.0 + 4 * 10.0


def test_gpu_comparisons():
    d = nca.NCDevArray([10], nca.DType.float32)
    d.fill(10.0)
    d[0:5].fill(0.0)
    m = d > 5.0
    assert m.dtype == nca.DType.bool
    h = np.array(m)
    assert not h[0] and h[9]
    d[0:10] = np.linspace(0, 9, 10, dtype=np.float32)
    gt = np.array(d > 5.0)
    ge = np.array(d >= 5.0)
    assert not gt[5] and ge[5]
    assert list(gt) == [False] * 6 + [True] * 4


# ---- fused lazy expression vs the oracle ----------------------------------------------------------------
def test_lazy_c0_is_one_fused_expression():
    from oracle import oracle as O
    rng = np.random.default_rng(11)
    n = 1 << 20
    a = rng.uniform(-1, 1, n).astype(np.float32)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    da, db = nca.NCDevArray.from_host(a), nca.NCDevArray.from_host(b)
    with nca.lazy_mode():
        e = da * db + nca.exp(-(da * da)) - db * 0.5
    assert isinstance(e, nca.DevExpr) and list(e.instrs) == tapes.fused_c0()
    out = np.array(nca.materialize(e))
    ref = O.expr(tapes.fused_c0(), [a, b, a, a, b], [0.5], capi.FLOAT32, np.float32)
    scale = np.abs(a * b).astype(np.float64) + np.exp(-a.astype(np.float64) ** 2) + 0.5 * np.abs(b)
    assert np.all(np.abs(out.astype(np.float64) - ref) <= 2 * np.finfo(np.float32).eps * scale)
    # sum fused into the evaluation (nothing materialised) vs FP64 of the stored values
    s = e.sum()
    s64 = out.astype(np.float64).sum()
    assert abs(s - s64) <= 1e-6 * np.abs(out).astype(np.float64).sum()


def test_eager_and_lazy_agree_bitwise():
    rng = np.random.default_rng(2)
    a = nca.NCDevArray.from_host(rng.standard_normal((64, 33)).astype(np.float32))
    b = nca.NCDevArray.from_host(rng.standard_normal((64, 33)).astype(np.float32))
    eager = a * b + a - b * 3.0
    with nca.lazy_mode():
        lazy = nca.materialize(a * b + a - b * 3.0)
    assert np.array_equal(np.array(eager), np.array(lazy))


# ---- reductions vs numpy ---------------------------------------------------------------------------------
@pytest.mark.parametrize("axes", [[0], [1], [2], [0, 2], [1, 2], [0, 1, 2]])
def test_axis_reductions(axes):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((17, 40, 33)).astype(np.float32)
    x[3, 5, 7] = x.max() + 1  # a tie for argmax below
    x[9, 5, 7] = x[3, 5, 7]
    d = nca.NCDevArray.from_host(x)
    ax = tuple(axes)
    assert np.array_equal(np.array(d.max(axes)), x.max(axis=ax))
    assert np.array_equal(np.array(d.min(axes)), x.min(axis=ax))
    s = np.array(d.sum(axes)).astype(np.float64)
    s64 = x.astype(np.float64).sum(axis=ax)
    bound = 2 * np.finfo(np.float32).eps * np.abs(x).astype(np.float64).sum(axis=ax) + 1e-30
    assert np.all(np.abs(s - s64) <= bound)
    m = np.array(d.mean(axes))
    assert np.allclose(m, x.astype(np.float64).mean(axis=ax), rtol=1e-12, atol=1e-12)
    v = np.array(d.var(axes))
    assert np.allclose(v, x.astype(np.float64).var(axis=ax), rtol=1e-10)
    if len(axes) == 1:
        # argmax stores the C-order linear index into x of the first occurrence
        got = np.array(d.argmax(axes))
        moved = np.moveaxis(x, axes[0], -1)
        k = moved.argmax(axis=-1)
        idx = np.indices(k.shape)
        full = [None] * 3
        j = 0
        for dd in range(3):
            if dd == axes[0]:
                full[dd] = k
            else:
                full[dd] = idx[j]
                j += 1
        want = np.ravel_multi_index(tuple(full), x.shape)
        assert np.array_equal(got, want)


def test_full_reductions_and_var():
    rng = np.random.default_rng(8)
    x = rng.standard_normal(100003).astype(np.float32)
    d = nca.NCDevArray.from_host(x)
    assert d.max() == float(x.max()) and d.min() == float(x.min())
    assert d.argmax() == int(x.argmax()) and d.argmin() == int(x.argmin())
    assert abs(d.sum() - x.astype(np.float64).sum()) <= 2 * np.finfo(np.float32).eps * np.abs(x).sum()
    assert abs(d.mean() - x.astype(np.float64).mean()) <= 1e-12
    assert abs(d.var() - x.astype(np.float64).var()) <= 1e-12 * x.var()
    assert abs(d.var(ddof=1) - x.astype(np.float64).var(ddof=1)) <= 1e-12 * x.var()
    assert abs(d.std() - x.astype(np.float64).std()) <= 1e-12
    b = d > 0.0
    assert b.any() is True and b.all() is False


def test_scatter_add_matches_numpy():
    rng = np.random.default_rng(4)
    idx = rng.integers(-5, 60, 5000).astype(np.int32)
    src = rng.integers(0, 100, 5000).astype(np.float32)  # integers: every order gives the same sum
    dest = nca.NCDevArray([50], nca.DType.float32)
    dest.fill(0.0)
    dest.scatter_add(nca.NCDevArray.from_host(idx), nca.NCDevArray.from_host(src))
    want = np.zeros(50, np.float64)
    ok = (idx >= 0) & (idx < 50)
    np.add.at(want, idx[ok], src[ok])
    assert np.array_equal(np.array(dest), want.astype(np.float32))


# ---- interop, storage types, stencils ---------------------------------------------------------------
def test_torch_zero_copy():
    import torch
    t = torch.arange(24, dtype=torch.float32, device="cuda").reshape(4, 6)
    v = nca.from_torch(t)
    assert v.shape == (4, 6) and v.dtype == nca.DType.float32
    v[1:3, ::2] = -1.0  # writes into the tensor's memory
    torch.cuda.synchronize()
    want = np.arange(24, dtype=np.float32).reshape(4, 6)
    want[1:3, ::2] = -1.0
    assert np.array_equal(t.cpu().numpy(), want)
    assert v.sum() == float(want.sum())
    # and back: an NCDevArray seen by torch without a copy
    back = torch.as_tensor(v, device="cuda")
    assert back.data_ptr() == t.data_ptr()


def test_bfloat16_storage():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(4096).astype(np.float32)
    d = nca.NCDevArray.from_host(x, nca.DType.bfloat16)
    assert d.dtype == nca.DType.bfloat16 and d.itemsize == 2
    import torch
    xb = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(np.array(d), xb)  # round to nearest even, widened on transfer
    s = d.sum()
    assert abs(s - xb.astype(np.float64).sum()) <= 2 * np.finfo(np.float32).eps * np.abs(xb).sum()


def test_python_stencil_matches_capi():
    from oracle import oracle as O
    rng = np.random.default_rng(9)
    x = rng.standard_normal((16, 24, 64)).astype(np.float32)
    st = nca.Stencil(tapes.LAPLACIAN_OFFSETS, lambda v: v[0] + v[1] + v[2] + v[3] + v[4] + v[5] - 6.0 * v[6])
    src = nca.NCDevArray.from_host(x)
    dst = nca.NCDevArray(list(x.shape))
    st.apply(src, dst, nca.Boundary.zero)
    ref = O.stencil_f32(x, x.shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    assert np.array_equal(np.array(dst), ref)
    # the transposed stencil of the Laplacian is the Laplacian (symmetric offsets)
    adj = st.adjoint()
    dst2 = nca.NCDevArray(list(x.shape))
    adj.apply(src, dst2, nca.Boundary.zero)
    assert np.allclose(np.array(dst2), ref, rtol=1e-6, atol=1e-5)


def test_errors_are_raised():
    a = nca.NCDevArray([3, 4])
    b = nca.NCDevArray([5])
    with pytest.raises(nca.NcaError):
        a + b
    with pytest.raises(nca.NcaError):
        a.to_device()
    with pytest.raises(nca.NcaError):
        a[:, ::2].reshape([6])


# ---- DLPack ---------------------------------------------------------------------------------------------
def test_dlpack_from_torch_is_zero_copy():
    t = torch.arange(12, dtype=torch.float32, device="cuda").reshape(3, 4)
    a = nca.from_dlpack(t)
    assert a.shape == (3, 4) and a.data_ptr == t.data_ptr() and a.dtype == nca.DType.float32
    assert np.array_equal(np.array(a), t.cpu().numpy())
    a.fill(5.0)                      # writes land in the torch tensor's memory
    torch.cuda.synchronize()
    assert bool((t == 5.0).all())
    tt = torch.arange(12, dtype=torch.int64, device="cuda").reshape(3, 4).t()   # strides (1, 4) elements
    b = nca.from_dlpack(tt)
    assert b.strides == (8, 32) and np.array_equal(np.array(b), tt.cpu().numpy())


def test_dlpack_import_orders_the_producer_stream():
    """A tensor still being written on a non-default torch stream: from_dlpack passes the
    library's consumer stream (0, ROCm's default) to __dlpack__, so torch orders its pending work
    before the library's next call on the default stream (a 256 MiB fill and multiply, then the
    library's synchronous sum)."""
    side = torch.cuda.Stream()
    n = 1 << 26
    with torch.cuda.stream(side):
        t = torch.empty(n, device="cuda")
        t.fill_(1.0)
        t.mul_(3.0)
        a = nca.from_dlpack(t)
        s = a.sum()
    assert float(s) == 3.0 * n
    torch.cuda.synchronize()


def test_dlpack_to_torch_is_zero_copy():
    a = nca.NCDevArray.from_host(np.arange(24, dtype=np.int32).reshape(2, 3, 4))
    t = torch.from_dlpack(a)
    assert t.is_cuda and t.data_ptr() == a.data_ptr and t.dtype == torch.int32 and tuple(t.shape) == (2, 3, 4)
    assert np.array_equal(t.cpu().numpy(), np.arange(24).reshape(2, 3, 4))
    t += 1                           # torch writes are visible through the array
    torch.cuda.synchronize()
    assert np.array_equal(np.array(a), np.arange(24).reshape(2, 3, 4) + 1)
    v = a[:, 1:, ::2]                # a strided view exports its element strides
    tv = torch.from_dlpack(v)
    assert np.array_equal(tv.cpu().numpy(), (np.arange(24).reshape(2, 3, 4) + 1)[:, 1:, ::2])


@pytest.mark.parametrize("tdt", [torch.bfloat16, torch.float16, torch.bool, torch.uint8, torch.float64])
def test_dlpack_dtypes_round_trip(tdt):
    src = (torch.arange(10, device="cuda") % 3).to(tdt)
    a = nca.from_dlpack(src)
    back = torch.from_dlpack(a)
    assert back.dtype == tdt and back.data_ptr() == src.data_ptr() and torch.equal(back, src)


def test_dlpack_keeps_the_producer_alive():
    t = torch.full((1 << 20,), 7.0, device="cuda")
    ptr = t.data_ptr()
    a = nca.from_dlpack(t)
    del t
    junk = [torch.full((1 << 20,), 1.0, device="cuda") for _ in range(4)]   # would reuse freed blocks
    torch.cuda.synchronize()
    assert a.data_ptr == ptr and float(np.array(a).min()) == 7.0 and all(j.data_ptr() != ptr for j in junk)
    cap = nca.NCDevArray.from_host(np.ones(4, np.float32)).__dlpack__()   # never consumed: freed with it
    del cap
    assert nca.from_dlpack(nca.NCDevArray.from_host(np.ones(3, np.float32)).__dlpack__()).shape == (3,)
