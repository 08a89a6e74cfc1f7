"""NC pointer-axis arrays: rows allocated separately and reached through a device table of row
pointers (NCOffsetsPolicy::advance, layout.hh:468-480; the NCDevArrayRef list constructor,
pyncarray/utilities.hh:273-314).

The oracle is numpy over the stacked rows.  Elementwise results are bit-exact against the same
expression over a dense copy (the per-element arithmetic is the same emitted code); max, min
and argmax are exact; sums are bounded against FP64.  The indexing cases transcribe the
reference's tests/python/test_indexing.py:14-121 onto device rows.
"""
import ctypes

import numpy as np
import pytest
import torch

import ncarray_amd as nca
from ncarray_amd import capi, tapes
from ncarray_amd import torch_interop as ti

pytestmark = pytest.mark.gpu
EPS32 = np.finfo(np.float32).eps


@pytest.fixture(autouse=True)
def eager(dev):
    nca.set_eager(True)
    yield
    nca.set_eager(True)


def _rows(rng, n, shape, dtype=torch.float32):
    return [torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(dtype).cuda() for _ in range(n)]


def _table(rows):
    return torch.tensor([r.data_ptr() for r in rows], dtype=torch.int64, device="cuda")


def _c0_tape(dt=capi.FLOAT32):
    return capi.Tape(tapes.fused_c0(), [dt] * 5, [capi.FLOAT32], capi.FLOAT32, dt)


# ---- C ABI ---------------------------------------------------------------------------------------
def test_capi_pointer_axis_inputs_and_output():
    rng = np.random.default_rng(10)
    R, H, W = 5, 17, 40
    a_rows, b_rows = _rows(rng, R, (H, W)), _rows(rng, R, (H, W))
    o_rows = [torch.full((H, W), np.nan, device="cuda") for _ in range(R)]
    ta, tb, to = _table(a_rows), _table(b_rows), _table(o_rows)
    st = [1, W * 4, 4]
    va = capi.make_nc_view(ta.data_ptr(), capi.FLOAT32, (R, H, W), st, 0)
    vb = capi.make_nc_view(tb.data_ptr(), capi.FLOAT32, (R, H, W), st, 0)
    vo = capi.make_nc_view(to.data_ptr(), capi.FLOAT32, (R, H, W), st, 0)
    tape = _c0_tape()
    tape.launch([va, vb, va, va, vb], vo, scalars=(0.5,))
    A, B = torch.stack(a_rows), torch.stack(b_rows)
    dense = torch.empty_like(A)
    tape.launch([ti.view(A), ti.view(B), ti.view(A), ti.view(A), ti.view(B)], ti.view(dense), scalars=(0.5,))
    torch.cuda.synchronize()
    for i in range(R):
        assert torch.equal(o_rows[i], dense[i]), i
    a64, b64 = A.double(), B.double()
    want = a64 * b64 + torch.exp(-a64 * a64) - 0.5 * b64
    scale = (a64 * b64).abs() + torch.exp(-a64 * a64) + 0.5 * b64.abs()
    assert bool(((dense.double() - want).abs() <= 3 * EPS32 * scale).all())


def test_capi_table_axes_offsets_and_mixed_operands():
    """Pointer axis 1 with a table axis before it, a table offset and a row byte offset; the
    other operand is an ordinary dense array."""
    rng = np.random.default_rng(11)
    H, W = 9, 33
    rows = _rows(rng, 8, (H, W))
    tbl = _table(rows)
    # element (d0, d1, h, w) = rows[d0*3 + d1 + 2][h, w + 1]
    shape = (2, 3, H, W - 1)
    va = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, shape, [3 * 8, 1, W * 4, 4], 1, ptr_offset=2, row_offset=4)
    gathered = torch.stack([torch.stack([rows[d0 * 3 + d1 + 2][:, 1:] for d1 in range(3)]) for d0 in range(2)])
    b = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda()
    out_nc, out_dense = torch.empty_like(b), torch.empty_like(b)
    tape = _c0_tape()
    tape.launch([va, ti.view(b), va, va, ti.view(b)], ti.view(out_nc), scalars=(0.5,))
    g = gathered.contiguous()
    tape.launch([ti.view(g), ti.view(b), ti.view(g), ti.view(g), ti.view(b)], ti.view(out_dense), scalars=(0.5,))
    torch.cuda.synchronize()
    assert torch.equal(out_nc, out_dense)


@pytest.mark.parametrize("step", [-1, 0, 2])
def test_capi_table_stride(step):
    """strides[P] counts table entries: -1 reverses the rows, 0 repeats one, 2 skips."""
    rng = np.random.default_rng(12)
    R, H, W = 6, 7, 24
    rows = _rows(rng, R, (H, W))
    tbl = _table(rows)
    n = {-1: R, 0: 4, 2: 3}[step]
    first = {-1: R - 1, 0: 2, 2: 1}[step]
    picked = [first + i * step for i in range(n)]
    assert all(0 <= p < R for p in picked)
    v = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (n, H, W), [step, W * 4, 4], 0, ptr_offset=first)
    out = torch.empty((n, H, W), device="cuda")
    tape = capi.Tape([tapes.P(capi.OP_LOAD_NCARR, 0), tapes.P(capi.OP_LOAD_CONST, 0), tapes.P(capi.OP_MUL)],
                     [capi.FLOAT32], [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    tape.launch([v], ti.view(out), scalars=(3.0,))
    torch.cuda.synchronize()
    want = torch.stack([rows[p] for p in picked]) * 3.0
    assert torch.equal(out, want)


def test_capi_fused_reduction_over_rows():
    rng = np.random.default_rng(13)
    R, H, W = 4, 64, 96
    rows = _rows(rng, R, (H, W))
    tbl = _table(rows)
    v = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (R, H, W), [1, W * 4, 4], 0)
    out = capi.make_view(None, capi.FLOAT32, (R, H, W))
    tape = capi.Tape([tapes.P(capi.OP_LOAD_NCARR, 0)], [capi.FLOAT32], [], capi.FLOAT32, capi.FLOAT32)
    s = torch.empty((), dtype=torch.float32, device="cuda")  # the sum of float32 values is float32
    mx = torch.empty((), dtype=torch.float32, device="cuda")
    tape.launch_reduce([v], out, capi.FUSE_SUM | capi.FUSE_MAX, capi.ACCUM_COMPENSATED, s.data_ptr(), mx.data_ptr())
    torch.cuda.synchronize()
    X = torch.stack(rows)
    assert mx.item() == X.max().item()
    assert abs(s.item() - X.double().sum().item()) <= 4 * EPS32 * X.abs().double().sum().item()


def test_capi_bf16_rows():
    rng = np.random.default_rng(14)
    R, H, W = 3, 10, 48
    a_rows, b_rows = _rows(rng, R, (H, W), torch.bfloat16), _rows(rng, R, (H, W), torch.bfloat16)
    ta, tb = _table(a_rows), _table(b_rows)
    st = [1, W * 2, 2]
    va = capi.make_nc_view(ta.data_ptr(), capi.BFLOAT16, (R, H, W), st, 0)
    vb = capi.make_nc_view(tb.data_ptr(), capi.BFLOAT16, (R, H, W), st, 0)
    tape = _c0_tape(capi.BFLOAT16)
    out = torch.empty((R, H, W), dtype=torch.bfloat16, device="cuda")
    tape.launch([va, vb, va, va, vb], ti.view(out), scalars=(0.5,))
    A, B = torch.stack(a_rows), torch.stack(b_rows)
    dense = torch.empty_like(out)
    tape.launch([ti.view(A), ti.view(B), ti.view(A), ti.view(A), ti.view(B)], ti.view(dense), scalars=(0.5,))
    torch.cuda.synchronize()
    assert torch.equal(out, dense)


def test_capi_pointer_views_in_reductions_and_binning_layouts():
    """Reductions read pointer views in place; binning takes frames behind pointers only when the
    pointer axis is the frame axis (a pointer per row inside a frame is refused, not misread)."""
    rows = [torch.full((8,), float(r + 1), device="cuda") for r in range(2)]
    tbl = _table(rows)
    v = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (2, 8), [1, 4], 0)
    slot = torch.empty((), dtype=torch.float32, device="cuda")
    capi.call("nca_hip_reduce_full", ctypes.byref(v), capi.REDUCE_SUM, capi.ACCUM_PLAIN,
              ctypes.c_void_p(slot.data_ptr()), None)
    assert slot.item() == 24.0
    frows = [torch.ones(16, device="cuda") for _ in range(4)]
    ftbl = _table(frows)
    fv = capi.make_nc_view(ftbl.data_ptr(), capi.FLOAT32, (1, 4, 16), [0, 1, 4], 1)   # rows of one frame
    bm = torch.zeros(64, dtype=torch.int32, device="cuda")
    counts = torch.empty((1, 2), dtype=torch.int32, device="cuda")
    with pytest.raises(capi.NcaError) as e:
        capi.call("nca_hip_bin", ctypes.byref(fv), ctypes.byref(ti.view(bm)), 2, capi.COUNT_ALL,
                  ctypes.byref(ti.view(counts)), None, None)
    assert e.value.code == capi.ERR_INVALID


# ---- Python NCDevArrayRef ------------------------------------------------------------------------
def test_reference_indexing_cases():
    arr = torch.arange(10, dtype=torch.int32, device="cuda")
    for ncarr in (nca.NCDevArrayRef(arr), nca.NCDevArrayRef([arr])):
        assert ncarr.shape == (1, 10) and ncarr.pointer_axis == 0
        assert ncarr[0, 0].item() == 0 and ncarr[0, 5].item() == 5 and ncarr[0, -1].item() == 9
        slc = ncarr[0, 2:7]
        assert slc.shape == (5,) and slc.pointer_axis == -1
        assert np.array_equal(np.array(slc), np.arange(10, dtype=np.int32)[2:7])
    a2 = torch.arange(20, dtype=torch.int32, device="cuda").reshape(4, 5)
    for ncarr in (nca.NCDevArrayRef(a2), nca.NCDevArrayRef([a2])):
        assert ncarr[0, 1, 2].item() == 7
        slc = ncarr[0, 1:3, 2:4]
        assert slc.shape == (2, 2)
        assert np.array_equal(np.array(slc), np.arange(20, dtype=np.int32).reshape(4, 5)[1:3, 2:4])
    sq = nca.NCDevArrayRef([torch.ones((1, 5, 1, 3), device="cuda")]).squeeze()
    assert sq.shape == (5, 3) and sq.pointer_axis == -1
    ncarr = nca.NCDevArrayRef([arr])
    ncarr[0, 5] = 42
    assert ncarr[0, 5].item() == 42
    torch.cuda.synchronize()
    assert arr[5].item() == 42  # the reference writes through to the bound memory


def _stacked():
    rng = np.random.default_rng(15)
    rows = _rows(rng, 4, (6, 10))
    ref = nca.NCDevArrayRef(rows)
    return rows, ref, torch.stack(rows).cpu().numpy()


def test_views_of_row_arrays():
    rows, ref, X = _stacked()
    assert ref.shape == (4, 6, 10) and ref.pointer_axis == 0 and not ref.is_contiguous
    assert np.array_equal(np.array(ref), X)
    cases = {
        "row": (ref[1], X[1], -1),
        "tail": (ref[1:], X[1:], 0),
        "reversed": (ref[::-1], X[::-1], 0),
        "stepped": (ref[::2], X[::2], 0),
        "inner": (ref[:, 1:5, ::3], X[:, 1:5, ::3], 0),
        "newaxis": (ref[None], X[None], 1),
        "ellipsis": (ref[..., 4], X[..., 4], 0),
        "row_then_col": (ref[2, :, 7], X[2, :, 7], -1),
        "chained": (ref[1:][::-1][:, 2:, 1:9][1], X[1:][::-1][:, 2:, 1:9][1], -1),
    }
    for name, (v, want, pax) in cases.items():
        assert v.shape == want.shape, name
        assert v.pointer_axis == pax, name
        assert np.array_equal(np.array(v), want), name
    with pytest.raises(nca.NcaError):
        ref.reshape([240])
    assert np.array_equal(np.array(ref.reshape([240], copy=True)), X.reshape(240))
    with pytest.raises(nca.NcaError):
        ref.__cuda_array_interface__
    assert np.array_equal(np.array(ref.dense()), X) and ref.dense().is_contiguous


def test_expressions_and_writes_through_rows():
    rows, ref, X = _stacked()
    d = nca.NCDevArray.from_host(np.full(X.shape, 0.25, np.float32))
    r = ref * 2.0 + d
    assert r.pointer_axis == -1 and np.array_equal(np.array(r), X * np.float32(2.0) + np.float32(0.25))
    with nca.lazy_mode():
        e = ref * ref - ref[::-1]
    assert np.array_equal(np.array(nca.materialize(e)), X * X - X[::-1])
    assert abs(e.sum() - (X.astype(np.float64) ** 2 - X[::-1]).sum()) <= 8 * EPS32 * (X.astype(np.float64) ** 2 + np.abs(X)).sum()
    # writes: in place, sliced assignment, fill
    ref += 1.0
    ref[1:3, 2] = -7.0
    torch.cuda.synchronize()
    want = X + np.float32(1.0)
    want[1:3, 2] = -7.0
    assert np.array_equal(torch.stack(rows).cpu().numpy(), want)
    ref[::-1] = nca.NCDevArray.from_host(X)  # reversed rows take X
    torch.cuda.synchronize()
    assert np.array_equal(torch.stack(rows).cpu().numpy(), X[::-1])
    ref[3].fill(5.0)
    torch.cuda.synchronize()
    assert bool((rows[3] == 5.0).all())


def test_reductions_of_row_arrays():
    rows, ref, X = _stacked()
    X[2, 3, 4] = X[0, 3, 4] = 9.0  # argmax tie along axis 0: first occurrence wins
    ref[...] = nca.NCDevArray.from_host(X)
    X64 = X.astype(np.float64)
    assert ref.max() == X.max() and ref.min() == X.min()
    assert ref.argmax() == int(np.argmax(X)) and ref.argmin() == int(np.argmin(X))
    assert abs(ref.sum() - X64.sum()) <= 4 * EPS32 * np.abs(X64).sum()
    assert abs(ref.mean() - X64.mean()) <= 4 * EPS32 * np.abs(X64).mean()
    assert abs(ref.var() - X64.var()) <= 1e-5 * X64.var()
    assert np.array_equal(np.array(ref.max([0])), X.max(0))
    # axis argmax gives the linear index into the input (the oracle's execute_reduce_axes)
    from oracle import oracle as O
    want = O.reduce_axes_f32(X, [0], 3)
    assert np.array_equal(want, np.argmax(X, 0) * 60 + np.arange(60).reshape(6, 10))
    assert np.array_equal(np.array(ref.argmax([0])), want)
    assert np.array_equal(np.array(ref.min([1, 2])), X.min((1, 2)))
    s0 = np.array(ref.sum([0]))
    assert np.all(np.abs(s0 - X64.sum(0)) <= 4 * EPS32 * np.abs(X64).sum(0))
    v12 = np.array(ref.var([1, 2]))
    assert np.allclose(v12, X64.var((1, 2)), rtol=1e-5)
    assert np.array_equal(np.array(ref[1:].max([2])), X[1:].max(2))


def test_stencil_scatter_and_copies_on_row_arrays():
    rng = np.random.default_rng(16)
    rows = _rows(rng, 6, (12, 40))
    ref = nca.NCDevArrayRef(rows)
    X = torch.stack(rows).cpu().numpy()
    st = nca.Stencil(tapes.LAPLACIAN_OFFSETS, lambda v: v[0] + v[1] + v[2] + v[3] + v[4] + v[5] - 6.0 * v[6])
    dense_src = nca.NCDevArray.from_host(X)
    want = nca.NCDevArray(list(X.shape))
    st.apply(dense_src, want, nca.Boundary.zero)
    got = nca.NCDevArray(list(X.shape))
    st.apply(ref, got, nca.Boundary.zero)
    assert np.array_equal(np.array(got), np.array(want))
    out_rows = [torch.zeros((12, 40), device="cuda") for _ in range(6)]
    out_ref = nca.NCDevArrayRef(out_rows)
    st.apply(ref, out_ref, nca.Boundary.zero)
    torch.cuda.synchronize()
    assert np.array_equal(torch.stack(out_rows).cpu().numpy(), np.array(want))
    # scatter_add into a 1-D row array whose elements are separate 0-d allocations
    # (integer-valued sources: any order gives the same sums)
    tgt = [torch.zeros((), device="cuda") for _ in range(20)]
    tref = nca.NCDevArrayRef(tgt)
    assert tref.shape == (20,) and tref.pointer_axis == 0
    idx = rng.integers(-3, 25, 2000).astype(np.int32)
    src = rng.integers(0, 50, 2000).astype(np.float32)
    tref.scatter_add(nca.NCDevArray.from_host(idx), nca.NCDevArray.from_host(src))
    wantsa = np.zeros(20, np.float64)
    ok = (idx >= 0) & (idx < 20)
    np.add.at(wantsa, idx[ok], src[ok])
    torch.cuda.synchronize()
    assert np.array_equal(torch.stack(tgt).cpu().numpy(), wantsa.astype(np.float32))
    # copies and conversions
    assert np.array_equal(np.array(ref.astype(nca.DType.float64)), X.astype(np.float64))
    assert np.array_equal(ref.to_host(nca.DType.int32), X.astype(np.int32))
    c = ref.copy()
    assert c.pointer_axis == -1 and c.is_contiguous and np.array_equal(np.array(c), X)


def test_bf16_row_arrays():
    rng = np.random.default_rng(17)
    rows = _rows(rng, 3, (8, 64), torch.bfloat16)
    ref = nca.NCDevArrayRef(rows)
    assert ref.dtype == nca.DType.bfloat16
    X = torch.stack(rows).float().cpu().numpy()
    assert np.array_equal(np.array(ref), X)
    assert ref.max() == X.max()
    assert abs(ref.sum() - X.astype(np.float64).sum()) <= 4 * EPS32 * np.abs(X).astype(np.float64).sum()


def test_row_errors():
    a = torch.zeros((3, 4), device="cuda")
    b = torch.zeros((3, 5), device="cuda")
    with pytest.raises(nca.NcaError):
        nca.NCDevArrayRef([a, b])
    with pytest.raises(TypeError):
        nca.NCDevArrayRef([np.zeros(4, np.float32)])
    with pytest.raises(ValueError):
        nca.NCDevArrayRef([a], read_only=True)
    ref = nca.NCDevArrayRef([a, a.clone()])
    with pytest.raises(nca.NcaError):
        ref[2]
    with pytest.raises(nca.NcaError):
        ref.from_rows([ref])
