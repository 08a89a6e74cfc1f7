"""SO views (PEP 3118 strides + suboffsets, SOArrayPolicy::advance, layout.hh:677-700) in the
expression kernels: any number of indirect axes, as operands and as the output.

Inputs are built as one byte image of a device allocation (oracle.ByteImage): pointer tables,
tables of tables and rows placed in a shuffled order.  The oracle resolves the same views over
the host copy of that image (oracle.so_gather / so_scatter).  Elementwise results are bit-exact
against the same expression over the gathered dense arrays (the per-element arithmetic is the
same emitted code); written images are compared byte for byte, which also shows that nothing
outside the view was written; max and min are exact, sums are bounded against FP64.
"""
import ctypes

import numpy as np
import pytest
import torch

from ncarray_amd import capi, tapes
from ncarray_amd import torch_interop as ti
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS32 = np.finfo(np.float32).eps


@pytest.fixture(autouse=True)
def _device(dev):
    return dev


def _image(size):
    t = torch.zeros(size, dtype=torch.uint8, device="cuda")
    return t, O.ByteImage(size, t.data_ptr())


def _upload(t, im):
    t.copy_(torch.from_numpy(im.buf))
    torch.cuda.synchronize()


def _c0_tape(dt=capi.FLOAT32):
    return capi.Tape(tapes.fused_c0(), [dt] * 5, [capi.FLOAT32], capi.FLOAT32, dt)


def _two_level(im, rng, R, H, W, np_dt, pad=2, fill=None, gen=None):
    """A table of R tables of H row pointers; rows hold W values after `pad` values of padding
    (gen(n) draws n values; default standard normal).  Returns (base address, strides,
    suboffsets, rows[i][j] as numpy)."""
    it = np.dtype(np_dt).itemsize
    order = rng.permutation(R * H)
    rows, row_at = {}, {}
    for r in order:
        i, j = divmod(int(r), H)
        if fill is not None:
            v = np.full(W + pad, fill).astype(np_dt)
        elif gen is not None:
            v = gen(W + pad)
        else:
            v = rng.standard_normal(W + pad).astype(np_dt)
        rows[(i, j)] = v[pad:]
        row_at[(i, j)] = im.put(v)
    sub = [im.put_ptrs([row_at[(i, j)] for j in range(H)], deltas=-8) for i in range(R)]
    top = im.put_ptrs(sub, deltas=-16)
    return im.addr + top, (8, 8, it), (16, 8 + pad * it, -1), rows


def test_two_level_operand_matches_dense():
    rng = np.random.default_rng(40)
    R, H, W = 3, 5, 37
    t, im = _image(1 << 16)
    base, st, so, rows = _two_level(im, rng, R, H, W, np.float32)
    _upload(t, im)
    va = capi.make_so_view(base, capi.FLOAT32, (R, H, W), st, so)
    A = O.so_gather(im.buf, im.addr, base, np.float32, (R, H, W), st, so)
    assert np.array_equal(A, np.stack([np.stack([rows[(i, j)] for j in range(H)]) for i in range(R)]))
    b = torch.from_numpy(rng.standard_normal((R, H, W)).astype(np.float32)).cuda()
    Ad = torch.from_numpy(A).cuda()
    out, dense = torch.empty_like(b), torch.empty_like(b)
    tape = _c0_tape()
    tape.launch([va, ti.view(b), va, va, ti.view(b)], ti.view(out), scalars=(0.5,))
    tape.launch([ti.view(Ad), ti.view(b), ti.view(Ad), ti.view(Ad), ti.view(b)], ti.view(dense), scalars=(0.5,))
    torch.cuda.synchronize()
    assert torch.equal(out, dense)
    a64, b64 = Ad.double(), b.double()
    want = a64 * b64 + torch.exp(-a64 * a64) - 0.5 * b64
    scale = (a64 * b64).abs() + torch.exp(-a64 * a64) + 0.5 * b64.abs()
    assert bool(((out.double() - want).abs() <= 3 * EPS32 * scale).all())


def test_two_level_output_written_in_place():
    """The output is an SO view: the values land at the resolved addresses and every other byte
    of the image keeps its value."""
    rng = np.random.default_rng(41)
    R, H, W = 4, 3, 29
    t, im = _image(1 << 16)
    base, st, so, _ = _two_level(im, rng, R, H, W, np.float32, pad=1, fill=np.nan)
    _upload(t, im)
    before = im.buf.copy()
    a = torch.from_numpy(rng.standard_normal((R, H, W)).astype(np.float32)).cuda()
    b = torch.from_numpy(rng.standard_normal((R, H, W)).astype(np.float32)).cuda()
    O.so_offsets(before, im.addr, base, (R, H, W), st, so)  # asserts the view stays inside the image
    vo = capi.make_so_view(base, capi.FLOAT32, (R, H, W), st, so)
    tape = _c0_tape()
    tape.launch([ti.view(a), ti.view(b), ti.view(a), ti.view(a), ti.view(b)], vo, scalars=(0.5,))
    dense = torch.empty_like(a)
    tape.launch([ti.view(a), ti.view(b), ti.view(a), ti.view(a), ti.view(b)], ti.view(dense), scalars=(0.5,))
    torch.cuda.synchronize()
    want = O.so_scatter(before, im.addr, base, dense.cpu().numpy(), st, so)
    assert np.array_equal(t.cpu().numpy(), want)


def test_so_operand_and_output_together():
    """SO in, SO out (both two-level, in one image), plus an NC operand of the same rows."""
    rng = np.random.default_rng(42)
    R, H, W = 2, 6, 64
    t, im = _image(1 << 17)
    bi, st, so, rows = _two_level(im, rng, R, H, W, np.float32)
    bo, sto, soo, _ = _two_level(im, rng, R, H, W, np.float32, pad=3, fill=np.nan)
    # the NC form of sub-table i: pointer axis 1 (table entries); its entries point 8 bytes
    # before each row, and the view skips 2 floats of padding, so the row offset is 16 bytes
    _upload(t, im)
    before = im.buf.copy()
    sub_tables = [int(np.frombuffer(before[bi - im.addr + 8 * i:bi - im.addr + 8 * i + 8].tobytes(), "<u8")[0]) + 16
                  for i in range(R)]
    A = O.so_gather(before, im.addr, bi, np.float32, (R, H, W), st, so)
    O.so_offsets(before, im.addr, bo, (R, H, W), sto, soo)  # asserts the output stays inside the image
    tape = capi.Tape([tapes.P(capi.OP_LOAD_SOARR, 0), tapes.P(capi.OP_LOAD_NCARR, 1), tapes.P(capi.OP_MUL),
                      tapes.P(capi.OP_LOAD_CONST, 0), tapes.P(capi.OP_ADD)],
                     [capi.FLOAT32, capi.FLOAT32], [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    for i in range(R):
        vin = capi.make_so_view(bi + 8 * i, capi.FLOAT32, (1, H, W), st, so)
        vnc = capi.make_nc_view(sub_tables[i], capi.FLOAT32, (1, H, W), [0, 1, 4], 1, ptr_offset=0, row_offset=16)
        vout = capi.make_so_view(bo + 8 * i, capi.FLOAT32, (1, H, W), sto, soo)
        tape.launch([vin, vnc], vout, scalars=(1.25,))
    torch.cuda.synchronize()
    Ad = torch.from_numpy(A).cuda()
    dense = torch.empty_like(Ad)
    tape.launch([ti.view(Ad), ti.view(Ad)], ti.view(dense), scalars=(1.25,))
    torch.cuda.synchronize()
    want = O.so_scatter(before, im.addr, bo, dense.cpu().numpy(), sto, soo)
    assert np.array_equal(t.cpu().numpy(), want)


def test_innermost_axis_indirect_fused_reductions():
    """Every element through its own pointer (suboffset on the last axis), picked in a random
    order with repeats; fused sum/max/min over it."""
    rng = np.random.default_rng(43)
    n, m = 3000, 1000
    t, im = _image(1 << 17)
    vals = rng.standard_normal(m).astype(np.float32)
    voff = im.put(vals)
    pick = rng.integers(0, m, n)
    ptrs = im.put_ptrs(voff + 4 * pick, deltas=-4)
    _upload(t, im)
    v = capi.make_so_view(im.addr + ptrs, capi.FLOAT32, (n,), (8,), (4,))
    X = O.so_gather(im.buf, im.addr, im.addr + ptrs, np.float32, (n,), (8,), (4,))
    assert np.array_equal(X, vals[pick])
    tape = capi.Tape([tapes.P(capi.OP_LOAD_SOARR, 0)], [capi.FLOAT32], [], capi.FLOAT32, capi.FLOAT32)
    out = torch.empty(n, device="cuda")
    s = torch.empty((), dtype=torch.float32, device="cuda")
    mx = torch.empty((), dtype=torch.float32, device="cuda")
    mn = torch.empty((), dtype=torch.float32, device="cuda")
    tape.launch_reduce([v], ti.view(out), capi.FUSE_SUM | capi.FUSE_MAX | capi.FUSE_MIN, capi.ACCUM_COMPENSATED,
                       s.data_ptr(), mx.data_ptr(), mn.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), X)
    assert mx.item() == X.max() and mn.item() == X.min()
    assert abs(s.item() - X.astype(np.float64).sum()) <= 4 * EPS32 * np.abs(X).astype(np.float64).sum()


@pytest.mark.parametrize("tdt", [capi.BFLOAT16, capi.INT32])
def test_two_level_other_dtypes(tdt):
    rng = np.random.default_rng(44)
    R, H, W = 2, 4, 50
    t, im = _image(1 << 16)
    if tdt == capi.BFLOAT16:
        np_dt, tdt_torch = np.uint16, torch.bfloat16
        gen = lambda n: (O.bf16_round(rng.standard_normal(n).astype(np.float32)).view(np.uint32) >> 16).astype(np.uint16)
    else:
        np_dt, tdt_torch = np.int32, torch.int32
        gen = lambda n: rng.integers(-1000, 1000, n).astype(np.int32)
    base, st, so, _ = _two_level(im, rng, R, H, W, np_dt, pad=3, gen=gen)
    _upload(t, im)
    A = O.so_gather(im.buf, im.addr, base, np_dt, (R, H, W), st, so)
    v = capi.make_so_view(base, tdt, (R, H, W), st, so)
    tape = capi.Tape([tapes.P(capi.OP_LOAD_SOARR, 0), tapes.P(capi.OP_LOAD_SOARR, 0), tapes.P(capi.OP_MUL)],
                     [tdt], [], tdt, tdt)
    out = torch.empty((R, H, W), dtype=tdt_torch, device="cuda")
    tape.launch([v], ti.view(out))
    Ad = torch.from_numpy(A.view(np.int16) if tdt == capi.BFLOAT16 else A).cuda()
    if tdt == capi.BFLOAT16:
        Ad = Ad.view(torch.bfloat16)
    dense = torch.empty_like(out)
    tape.launch([ti.view(Ad)], ti.view(dense))
    torch.cuda.synchronize()
    if tdt == capi.BFLOAT16:
        assert torch.equal(out.view(torch.int16), dense.view(torch.int16))
        assert torch.equal(out.float(), (Ad.float() * Ad.float()).bfloat16().float())
    else:
        assert torch.equal(out, dense)
        assert torch.equal(out.cpu(), torch.from_numpy(A * A))


def test_so_without_suboffsets_is_a_plain_view():
    """An SO view whose suboffsets are all < 0 is accepted by every entry point and reads as the
    strided view it is (the reference's SOArray owners, test_reductions.cc:90-155)."""
    x = torch.arange(1, 7, dtype=torch.int32, device="cuda").reshape(2, 3)
    v = capi.make_so_view(x.data_ptr(), capi.INT32, (2, 3), (12, 4), (-1, -1))
    slot = torch.zeros((), dtype=torch.int64, device="cuda")
    capi.call("nca_hip_reduce_full", ctypes.byref(v), capi.REDUCE_SUM, capi.ACCUM_PLAIN,
              ctypes.c_void_p(slot.data_ptr()), None)
    torch.cuda.synchronize()
    assert slot.item() == 21


def test_validation():
    t, im = _image(4096)
    row = im.put(np.zeros(8, np.float32))
    tbl = im.put_ptrs([row, row])
    _upload(t, im)
    slot = torch.empty((), dtype=torch.float32, device="cuda")
    v = capi.make_so_view(im.addr + tbl, capi.FLOAT32, (2, 8), (8, 4), (0, -1))
    # an indirect SO view is read in place by the reductions as well (the sum of two zero rows)
    capi.call("nca_hip_reduce_full", ctypes.byref(v), capi.REDUCE_SUM, capi.ACCUM_PLAIN,
              ctypes.c_void_p(slot.data_ptr()), None)
    assert slot.item() == 0.0
    tape = capi.Tape([tapes.P(capi.OP_LOAD_SOARR, 0)], [capi.FLOAT32], [], capi.FLOAT32, capi.FLOAT32)
    out = torch.empty((2, 8), device="cuda")
    bad = capi.make_so_view(im.addr + tbl, capi.FLOAT32, (2, 8), (8, 4), (0, -1))
    bad.pointer_axis = 0
    with pytest.raises(capi.NcaError) as e:
        tape.launch([bad], ti.view(out))
    assert e.value.code == capi.ERR_INVALID
    bad = capi.make_so_view(im.addr + tbl, capi.FLOAT32, (2, 8), (8, 4), (0, -1))
    bad.layout = 7
    with pytest.raises(capi.NcaError) as e:
        tape.launch([bad], ti.view(out))
    assert e.value.code == capi.ERR_INVALID


# ---- the reference's SOArray known answers, on device SO views -------------------------------------
def test_reference_soarray_ops():
    """test_operations.py:89-97: SOArrayRef([2, 4, 6]) + [1, 2, 3] == [3, 6, 9].  The reference
    binding puts a pointer axis in front (suboffset 0 on axis 0 of shape (1, 3))."""
    t, im = _image(256)
    row = im.put(np.array([2, 4, 6], np.int32))
    tbl = im.put_ptrs([row])
    _upload(t, im)
    so = capi.make_so_view(im.addr + tbl, capi.INT32, (1, 3), (8, 4), (0, -1))
    b = torch.tensor([[1, 2, 3]], dtype=torch.int32, device="cuda")
    out = torch.empty((1, 3), dtype=torch.int32, device="cuda")
    tape = capi.Tape([tapes.P(capi.OP_LOAD_SOARR, 0), tapes.P(capi.OP_LOAD_NCARR, 1), tapes.P(capi.OP_ADD)],
                     [capi.INT32, capi.INT32], [], capi.INT32, capi.INT32)
    tape.launch([so, ti.view(b)], ti.view(out))
    torch.cuda.synchronize()
    assert out.squeeze(0).tolist() == [3, 6, 9]


def test_reference_soarray_slice():
    """test_indexing.py:124-137: soarr[0, 1:3, 0:2] of arange(12).reshape(3, 4) behind a pointer
    axis; slicing the axes after the indirect one moves its suboffset (1 row of 16 bytes)."""
    arr = np.arange(12, dtype=np.int32).reshape(3, 4)
    t, im = _image(512)
    row = im.put(arr)
    tbl = im.put_ptrs([row])
    _upload(t, im)
    v = capi.make_so_view(im.addr + tbl, capi.INT32, (1, 2, 2), (8, 16, 4), (16, -1, -1))
    assert np.array_equal(O.so_gather(im.buf, im.addr, im.addr + tbl, np.int32, (1, 2, 2), (8, 16, 4), (16, -1, -1))[0],
                          arr[1:3, 0:2])
    out = torch.empty((1, 2, 2), dtype=torch.int32, device="cuda")
    tape = capi.Tape([tapes.P(capi.OP_LOAD_SOARR, 0)], [capi.INT32], [], capi.INT32, capi.INT32)
    tape.launch([v], ti.view(out))
    torch.cuda.synchronize()
    assert np.array_equal(out[0].cpu().numpy(), arr[1:3, 0:2])
    assert out[0, 0, 1].item() == 5 and out[0, 1, 0].item() == 8   # soarr[0, 1, 2] == 6 in the unsliced view


def test_reference_soarray_reductions():
    """test_reductions.cc:90-152 on SO views: five 1.5f values sum to 7.5 with max = min = 1.5
    (through one pointer per element, fused into the expression pass), and iota + 1 on 2 x 3
    (an SO view without suboffsets, i.e. the reference's SOArray owner) gives axis-0 sums
    [5, 7, 9], maxima [4, 5, 6], minima [1, 2, 3] and axis-1 sums [6, 15], maxima [3, 6], minima
    [1, 4]."""
    t, im = _image(512)
    vals = im.put(np.full(5, 1.5, np.float32))
    ptrs = im.put_ptrs([vals + 4 * k for k in range(5)])
    _upload(t, im)
    v = capi.make_so_view(im.addr + ptrs, capi.FLOAT32, (5,), (8,), (0,))
    tape = capi.Tape([tapes.P(capi.OP_LOAD_SOARR, 0)], [capi.FLOAT32], [], capi.FLOAT32, capi.FLOAT32)
    s, mx, mn = (torch.empty((), dtype=torch.float32, device="cuda") for _ in range(3))
    tape.launch_reduce([v], capi.make_view(None, capi.FLOAT32, (5,)), capi.FUSE_SUM | capi.FUSE_MAX | capi.FUSE_MIN,
                       capi.ACCUM_PLAIN, s.data_ptr(), mx.data_ptr(), mn.data_ptr())
    torch.cuda.synchronize()
    assert (s.item(), mx.item(), mn.item()) == (7.5, 1.5, 1.5)
    x = (torch.arange(6, dtype=torch.int32, device="cuda") + 1).reshape(2, 3)
    sv = capi.make_so_view(x.data_ptr(), capi.INT32, (2, 3), (12, 4), (-1, -1))
    want = {(0, capi.REDUCE_SUM): [5, 7, 9], (0, capi.REDUCE_MAX): [4, 5, 6], (0, capi.REDUCE_MIN): [1, 2, 3],
            (1, capi.REDUCE_SUM): [6, 15], (1, capi.REDUCE_MAX): [3, 6], (1, capi.REDUCE_MIN): [1, 4]}
    for (axis, op), w in want.items():
        out = torch.empty(len(w), dtype=torch.int32, device="cuda")  # int32 sums stay int32 (sum_type = T)
        ax = (ctypes.c_int64 * 1)(axis)
        capi.call("nca_hip_reduce_axes", ctypes.byref(sv), ax, 1, op, capi.ACCUM_PLAIN, ctypes.byref(ti.view(out)), None)
        torch.cuda.synchronize()
        assert out.tolist() == w, (axis, op)
