"""NC pointer-axis views of every geometry the device view form allows, resolved by the oracle.

The views live in one byte image of a device allocation (oracle.ByteImage): shuffled rows, and
pointer tables placed between them.  oracle.nc_offsets resolves every element of a view over the
host copy of the image; it is pinned against the reference's own NCOffsetsPolicy::advance
(layout.hh:451-480) in tests/test_oracle_ref.py.  Cases: the pointer axis first, in the middle
(outer axes stepping between tables) and last (a table of element pointers); a stepped table
(strides[P] = 2), a table offset (the reference's offsets[P]), a row offset (its byte offsets
after P) and a reversed inner axis.  Each view is checked as an expression operand (bit-exact
against the same tape over the gathered dense array), through full reductions (max/min and
their indices exact, sums bounded against FP64) and axis reductions (max/argmax exact)."""
import ctypes

import numpy as np
import pytest
import torch

from ncarray_amd import capi, tapes
from ncarray_amd import torch_interop as ti
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS32 = float(np.finfo(np.float32).eps)


@pytest.fixture(autouse=True)
def _device(dev):
    return dev


def _image(size):
    t = torch.zeros(size, dtype=torch.uint8, device="cuda")
    return t, O.ByteImage(size, t.data_ptr())


def _rows(im, rng, n, length, pad):
    """n shuffled rows of `length` float32 values, each after `pad` values of other data."""
    at = [0] * n
    for k in rng.permutation(n):
        at[int(k)] = im.put(rng.standard_normal(length + pad).astype(np.float32)) + 4 * pad
    return at


# (shape, pointer axis, table entry step, ptr_offset, row_offset in elements, reversed last axis)
CASES = [
    ((6, 9, 40), 0, 1, 2, 3, False),
    ((5, 7, 33), 0, 2, 1, 0, True),
    ((3, 8, 50), 1, 1, 1, 5, False),
    ((4, 6, 9), 2, 2, 3, 0, False),
]


def _case_view(rng, im, shape, pax, step, poff, roff, rev):
    """Build the tables and rows of one case in `im`; returns (base, strides) of the view."""
    nd = len(shape)
    it = 4
    inner = int(np.prod(shape[pax + 1:])) if pax + 1 < nd else 1
    length = inner + roff + 4
    n_entries = poff + step * (shape[pax] - 1) + 1
    n_tables = int(np.prod(shape[:pax])) if pax else 1
    tables = []
    for _ in range(n_tables):
        rows = _rows(im, rng, n_entries, length, pad=2)
        tables.append(im.put_ptrs(rows))
    tstride = (tables[1] - tables[0]) if n_tables > 1 else 0
    assert all(t == tables[0] + k * tstride for k, t in enumerate(tables)), "tables not evenly spaced"
    strides = [0] * nd
    for d in range(nd):
        if d < pax:
            strides[d] = tstride * int(np.prod(shape[d + 1:pax]))
        elif d == pax:
            strides[d] = step
        else:
            strides[d] = it * int(np.prod(shape[d + 1:]))
    if rev and pax < nd - 1:
        strides[-1] = -it
        roff = roff + shape[-1] - 1  # the row offset then points at each row's last element
    return im.addr + tables[0], strides, roff * it


def _gather(im, base, shape, strides, pax, poff, roff_bytes):
    off = O.nc_offsets(im.buf, im.addr, base, shape, strides, pax, poff, roff_bytes)
    assert off.min() >= 0 and off.max() + 4 <= im.buf.size
    return im.buf[off[..., None] + np.arange(4)].copy().view(np.float32)[..., 0]


def _full(v, op):
    out_code = capi.reduce_result_dtype(op, capi.FLOAT32)
    out = torch.empty((), dtype=ti.CODE_TO_TORCH[out_code], device="cuda")
    capi.call("nca_hip_reduce_full", ctypes.byref(v), op, capi.ACCUM_PLAIN, out.data_ptr(), None)
    return out.cpu().item()


def _axes(v, shape, ax, op):
    out_code = capi.reduce_result_dtype(op, capi.FLOAT32)
    oshape = [s for d, s in enumerate(shape) if d not in ax]
    out = torch.empty(oshape, dtype=ti.CODE_TO_TORCH[out_code], device="cuda")
    arr = (ctypes.c_int64 * len(ax))(*ax)
    capi.call("nca_hip_reduce_axes", ctypes.byref(v), arr, len(ax), op, capi.ACCUM_PLAIN, ctypes.byref(ti.view(out)),
              None)
    return out.cpu().numpy()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_nc_view_geometries_against_resolved_elements(case):
    shape, pax, step, poff, roff, rev = CASES[case]
    print(f"case {case}: shape {shape} pointer axis {pax} step {step} ptr_offset {poff} row_offset {roff} "
          f"reversed {rev}", flush=True)
    rng = np.random.default_rng(70 + case)
    t, im = _image(1 << 20)
    base, strides, roff_bytes = _case_view(rng, im, shape, pax, step, poff, roff, rev)
    t.copy_(torch.from_numpy(im.buf))
    torch.cuda.synchronize()
    X = _gather(im, base, shape, strides, pax, poff, roff_bytes)
    v = capi.make_nc_view(base, capi.FLOAT32, shape, strides, pax, ptr_offset=poff, row_offset=roff_bytes)

    # as an expression operand: bit-exact against the same tape over the dense array
    b = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda()
    Xd = torch.from_numpy(X).cuda()
    out, dense = torch.empty_like(b), torch.empty_like(b)
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    tape.launch([v, ti.view(b), v, v, ti.view(b)], ti.view(out), scalars=(0.5,))
    tape.launch([ti.view(Xd), ti.view(b), ti.view(Xd), ti.view(Xd), ti.view(b)], ti.view(dense), scalars=(0.5,))
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int32), dense.view(torch.int32))

    # full reductions
    mx, imx = O.max_fold(X)
    mn, imn = O.min_fold(X)
    assert _full(v, capi.REDUCE_MAX) == mx and _full(v, capi.REDUCE_ARGMAX) == imx
    assert _full(v, capi.REDUCE_MIN) == mn and _full(v, capi.REDUCE_ARGMIN) == imn
    s64 = X.astype(np.float64).sum()
    assert abs(_full(v, capi.REDUCE_SUM) - s64) <= 8 * EPS32 * np.abs(X).astype(np.float64).sum()

    # axis reductions: max and argmax along the first and the last axis
    for ax in ([0], [len(shape) - 1]):
        assert np.array_equal(_axes(v, shape, ax, capi.REDUCE_MAX), O.reduce_axes_f32(X, ax, 1)), ax
        assert np.array_equal(_axes(v, shape, ax, capi.REDUCE_ARGMAX), O.reduce_axes_f32(X, ax, 3)), ax
