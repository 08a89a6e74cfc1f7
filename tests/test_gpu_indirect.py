"""NC pointer-axis and SO views read in place by the reduction, binning and scatter kernels
(the reference's axes_reduce_kernel, reduce_kernel and scatter_add_kernel take any
ViewArrayLike and walk it with advance: kernels.cuh:166-277, 565-589, 660-680; layout.hh:468-568,
677-778).

Each case is checked against the oracle on the dense equivalent of the view (stacked rows,
oracle.so_gather): max/min/argmax/argmin and counts bit-exact, sums of integer-valued data
exact, other sums bounded against FP64."""
import ctypes

import numpy as np
import pytest
import torch

from ncarray_amd import capi
from ncarray_amd import torch_interop as ti
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _device(dev):
    return dev


def _table(rows):
    return torch.tensor([r.data_ptr() for r in rows], dtype=torch.int64, device="cuda")


def _full(v, code, op, mode=capi.ACCUM_PLAIN):
    out_code = capi.reduce_result_dtype(op, code)
    out = torch.empty((), dtype=ti.CODE_TO_TORCH[out_code], device="cuda")
    capi.call("nca_hip_reduce_full", ctypes.byref(v), op, mode, out.data_ptr(), None)
    return out.cpu()


def _axes(v, shape, code, ax, op, mode=capi.ACCUM_PLAIN):
    out_code = capi.reduce_result_dtype(op, code)
    oshape = [s for d, s in enumerate(shape) if d not in ax]
    out = torch.empty(oshape, dtype=ti.CODE_TO_TORCH[out_code], device="cuda")
    arr = (ctypes.c_int64 * max(1, len(ax)))(*ax)
    capi.call("nca_hip_reduce_axes", ctypes.byref(v), arr, len(ax), op, mode, ctypes.byref(ti.view(out)), None)
    return out.cpu().numpy()


def _nc_rows(rng, R, shape, integer=False):
    if integer:
        rows = [torch.from_numpy(rng.integers(-20, 20, shape).astype(np.float32)).cuda() for _ in range(R)]
    else:
        rows = [torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda() for _ in range(R)]
    return rows


@pytest.mark.parametrize("integer", [True, False], ids=["integer_valued", "normal"])
def test_full_reductions_of_pointer_axis_rows(integer):
    rng = np.random.default_rng(1)
    R, H, W = 7, 33, 130
    rows = _nc_rows(rng, R, (H, W), integer)
    tbl = _table(rows)
    v = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (R, H, W), [1, W * 4, 4], 0)
    X = torch.stack(rows).cpu().numpy()
    for op in (capi.REDUCE_MAX, capi.REDUCE_MIN, capi.REDUCE_ARGMAX, capi.REDUCE_ARGMIN):
        got = _full(v, capi.FLOAT32, op)
        want = (O.max_fold(X) if op in (capi.REDUCE_MAX, capi.REDUCE_ARGMAX) else O.min_fold(X))
        assert got.item() == (want[0] if op in (capi.REDUCE_MAX, capi.REDUCE_MIN) else want[1]), op
    for mode in (capi.ACCUM_PLAIN, capi.ACCUM_COMPENSATED):
        s = _full(v, capi.FLOAT32, capi.REDUCE_SUM, mode).item()
        s64 = X.astype(np.float64).sum()
        if integer:
            assert s == s64
        else:
            assert abs(s - s64) <= 8 * np.finfo(np.float32).eps * np.abs(X).astype(np.float64).sum()
        assert abs(_full(v, capi.FLOAT32, capi.REDUCE_MEAN, mode).item() - s64 / X.size) <= 1e-6 * np.abs(X).mean()
    assert _full(v, capi.FLOAT32, capi.REDUCE_ANY).item() == bool(np.any(X != 0))
    assert _full(v, capi.FLOAT32, capi.REDUCE_ALL).item() == bool(np.all(X != 0))


@pytest.mark.parametrize("ax", [[0], [1], [2], [0, 2], [1, 2], [0, 1, 2]])
def test_axis_reductions_of_pointer_axis_rows(ax):
    """Integer-valued data (exact sums), ties for argmax/argmin; a table offset and a row offset
    (the view starts at row 1 and skips each row's first line)."""
    rng = np.random.default_rng(2 + len(ax))
    R, H, W = 9, 40, 70
    rows = _nc_rows(rng, R, (H, W), integer=True)
    tbl = _table(rows)
    shape = (R - 1, H - 1, W)
    v = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, shape, [1, W * 4, 4], 0, ptr_offset=1, row_offset=W * 4)
    X = np.ascontiguousarray(torch.stack(rows).cpu().numpy()[1:, 1:, :])
    for op, oop in ((capi.REDUCE_SUM, 0), (capi.REDUCE_MAX, 1), (capi.REDUCE_MIN, 2), (capi.REDUCE_ARGMAX, 3),
                    (capi.REDUCE_ARGMIN, 4)):
        got = _axes(v, shape, capi.FLOAT32, ax, op)
        want = O.reduce_axes_f32(X, ax, oop)
        assert np.array_equal(got, want), (op, ax)


def test_axis_reduction_many_outputs_and_nan_signed_zero():
    """Thread-per-output form (many outputs) on rows dense in NaN and signed zeros."""
    rng = np.random.default_rng(3)
    R, H, W = 4, 300, 260
    pool = np.array([np.nan, -0.0, 0.0, -1.0, 1.0], np.float32)
    rows = [torch.from_numpy(rng.choice(pool, (H, W)).astype(np.float32)).cuda() for _ in range(R)]
    tbl = _table(rows)
    v = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (R, H, W), [1, W * 4, 4], 0)
    X = torch.stack(rows).cpu().numpy()
    for op, oop in ((capi.REDUCE_MAX, 1), (capi.REDUCE_MIN, 2), (capi.REDUCE_ARGMAX, 3), (capi.REDUCE_ARGMIN, 4)):
        got = _axes(v, X.shape, capi.FLOAT32, [0], op)
        want = O.reduce_axes_f32(X, [0], oop)
        assert np.array_equal(got.view(np.uint32 if oop <= 2 else np.int64), want.view(np.uint32 if oop <= 2 else np.int64))


def _image(size):
    t = torch.zeros(size, dtype=torch.uint8, device="cuda")
    return t, O.ByteImage(size, t.data_ptr())


def test_reductions_of_a_two_level_so_view():
    rng = np.random.default_rng(4)
    R, H, W, pad = 3, 17, 45, 2
    t, im = _image(1 << 17)
    order = rng.permutation(R * H)
    at = {}
    for r in order:
        i, j = divmod(int(r), H)
        at[(i, j)] = im.put(rng.integers(-9, 9, W + pad).astype(np.float32))
    sub = [im.put_ptrs([at[(i, j)] for j in range(H)], deltas=-8) for i in range(R)]
    base = im.addr + im.put_ptrs(sub, deltas=-16)
    st, so = (8, 8, 4), (16, 8 + pad * 4, -1)
    t.copy_(torch.from_numpy(im.buf))
    torch.cuda.synchronize()
    X = O.so_gather(im.buf, im.addr, base, np.float32, (R, H, W), st, so)
    v = capi.make_so_view(base, capi.FLOAT32, (R, H, W), st, so)
    assert _full(v, capi.FLOAT32, capi.REDUCE_SUM).item() == X.astype(np.float64).sum()
    assert _full(v, capi.FLOAT32, capi.REDUCE_ARGMAX).item() == O.max_fold(X)[1]
    for ax in ([0], [2], [0, 1]):
        for op, oop in ((capi.REDUCE_SUM, 0), (capi.REDUCE_MAX, 1), (capi.REDUCE_ARGMIN, 4)):
            assert np.array_equal(_axes(v, (R, H, W), capi.FLOAT32, ax, op), O.reduce_axes_f32(X, ax, oop)), (ax, op)


# ---- binning over frames behind a pointer table -----------------------------------------------------
@pytest.mark.parametrize("misaligned", [False, True], ids=["aligned", "misaligned_rows"])
def test_binning_frames_behind_pointers(misaligned):
    """Frames allocated one by one (a detector's frame list) binned through a table of frame
    pointers: counts exact, sums within 1 ulp of FP64; a 4-byte row offset makes every frame
    base misaligned for 16-byte loads (the element path), and 1023 x 1021 pixels are not a
    multiple of 4."""
    rng = np.random.default_rng(5)
    F, ny, nx, nb = 12, 1023, 1021, 500
    frames = [rng.exponential(50.0, (ny, nx)).astype(np.float32) for _ in range(F)]
    for f in frames:
        f[rng.random((ny, nx)) < 0.01] = 0.0
    off = 1 if misaligned else 0
    bufs = [torch.from_numpy(np.concatenate([np.zeros(off, np.float32), f.ravel()])).cuda() for f in frames]
    tbl = _table(bufs)
    fv = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (F, ny, nx), [1, nx * 4, 4], 0, row_offset=4 * off)
    bm = O.radial_binmap(ny, nx, ny // 2, nx // 2, nb)
    bmd = torch.from_numpy(bm).cuda()
    counts = torch.empty((F, nb), dtype=torch.int64, device="cuda")
    sums = torch.empty((F, nb), dtype=torch.float32, device="cuda")
    capi.call("nca_hip_bin", ctypes.byref(fv), ctypes.byref(ti.view(bmd)), nb, capi.COUNT_NONZERO,
              ctypes.byref(ti.view(counts)), ctypes.byref(ti.view(sums)), None)
    rc, rs = O.bin_f32(np.stack(frames), bm, nb, nonzero=True)
    assert np.array_equal(counts.cpu().numpy(), rc)
    s = sums.cpu().numpy()
    assert np.all(np.abs(s - rs) <= np.spacing(np.abs(rs).astype(np.float32)))
    # the backward gather into frames behind pointers
    g = torch.from_numpy(rng.standard_normal((F, nb)).astype(np.float32)).cuda()
    outs = [torch.full((ny * nx + off,), np.nan, device="cuda") for _ in range(F)]
    otbl = _table(outs)   # the pointer table must stay allocated while kernels read it
    ov = capi.make_nc_view(otbl.data_ptr(), capi.FLOAT32, (F, ny, nx), [1, nx * 4, 4], 0, row_offset=4 * off)
    capi.call("nca_hip_bin_backward", ctypes.byref(ti.view(g)), ctypes.byref(ti.view(bmd)), ctypes.byref(ov), None)
    want = O.bin_backward_f32(g.cpu().numpy(), bm)
    got = np.stack([o.cpu().numpy()[off:] for o in outs])
    assert np.array_equal(got, want)
    if off:
        assert all(np.isnan(o[0].item()) for o in outs)   # the skipped element untouched


# ---- scatter_add over indirect and strided views -----------------------------------------------------
def test_scatter_add_indirect_and_nd():
    """dest[idx[i]] += src[i] (scatter_add_kernel, kernels.cuh:660-680): dest a 2-D array whose
    rows sit behind a pointer table (the target is dest's C-order linear index), idx and src
    strided 2-D views; integer-valued sources, so any order of the atomics gives the same sums."""
    rng = np.random.default_rng(6)
    R, W = 5, 7
    drows = [torch.zeros(W, device="cuda") for _ in range(R)]
    dtbl = _table(drows)   # the pointer table must stay allocated while kernels read it
    dv = capi.make_nc_view(dtbl.data_ptr(), capi.FLOAT32, (R, W), [1, 4], 0)
    n0, n1 = 60, 50
    ibase = torch.from_numpy(rng.integers(-4, R * W + 4, (n0, 2 * n1)).astype(np.int32)).cuda()
    sbase = torch.from_numpy(rng.integers(0, 40, (2 * n0, n1)).astype(np.float32)).cuda()
    idx, src = ibase[:, ::2], sbase[::2, :]
    capi.call("nca_hip_scatter_add", ctypes.byref(dv), ctypes.byref(ti.view(idx)), ctypes.byref(ti.view(src)), None)
    torch.cuda.synchronize()
    want = O.scatter_add_f32(np.zeros(R * W, np.float32), idx.cpu().numpy().ravel(), src.cpu().numpy().ravel())
    got = torch.stack(drows).cpu().numpy().ravel()
    assert np.array_equal(got, want)
