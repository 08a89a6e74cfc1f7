"""max/min/argmax/argmin on data dense in NaN and signed zeros, on every kernel that computes
them: the full reduction, the three axis-reduction kernels (column, row and generic layouts),
the expression kernel's fused max/min, the ragged batch, and the stencils' fused max/min (march
and tile forms).  The rule (the same in the oracle): value max/min never select NaN and order
-0 below +0, so the bits do not depend on the order of combination; argmax/argmin are the host
fold's first occurrence with zeros equal (host/reductions.hh:130-212).  All results bit-exact
against the oracle."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P = lambda op, i=0: ((op & 0xFF) << 24) | (i & 0xFFFFFF)  # noqa: E731  (op_code.hh:174-192)
LOAD = 2


def _c():
    from ncarray_amd import capi, tapes, torch_interop as ti
    return capi, tapes, ti


def _bits(a):
    a = np.asarray(a)
    return a.view({4: np.uint32, 8: np.uint64, 2: np.uint16}[a.dtype.itemsize])


def nan_zero_data(shape, reduced, seed, positive=False):
    """Values from {NaN, -0, +0, -1, -2} (or {NaN, -0, +0, 1, 2}), so max (min) lines end at a
    zero whose sign matters, at a negative (positive) number, or at lowest() (max()) when the
    line is all NaN.  Along the reduced axes some lines have every zero turned into the sign
    the other zero would win over (-0 for max data, +0 for min data: the extremum is then that
    zero) and some are entirely NaN; NaN sits first, in the middle and last."""
    rng = np.random.default_rng(seed)
    pool = np.array([np.nan, -0.0, 0.0, 1.0, 2.0] if positive else [np.nan, -0.0, 0.0, -1.0, -2.0], np.float32)
    x = rng.choice(pool, shape, p=[0.3, 0.25, 0.25, 0.1, 0.1]).astype(np.float32)
    lshape = tuple(1 if d in reduced else s for d, s in enumerate(shape))
    flip = rng.random(lshape) < 0.3
    alln = rng.random(lshape) < 0.05
    x = np.where(flip & (x == 0), np.float32(0.0 if positive else -0.0), x)
    x = np.where(alln, np.float32(np.nan), x).astype(np.float32)
    return x


def _axes(t, ax, op):
    capi, tapes, ti = _c()
    out_code = capi.reduce_result_dtype(op, ti.TORCH_TO_CODE[t.dtype])
    oshape = [s for d, s in enumerate(t.shape) if d not in ax]
    out = torch.empty(oshape, dtype=ti.CODE_TO_TORCH[out_code], device=t.device)
    arr = (capi.ctypes.c_int64 * max(1, len(ax)))(*ax)
    capi.call("nca_hip_reduce_axes", capi.ctypes.byref(ti.view(t)), arr, len(ax), op, 0,
              capi.ctypes.byref(ti.view(out)), None)
    return out.cpu().numpy()


def _full(t, op):
    capi, tapes, ti = _c()
    out_code = capi.reduce_result_dtype(op, ti.TORCH_TO_CODE[t.dtype])
    out = torch.empty((), dtype=ti.CODE_TO_TORCH[out_code], device=t.device)
    capi.call("nca_hip_reduce_full", capi.ctypes.byref(ti.view(t)), op, 0, out.data_ptr(), None)
    return out.cpu()


# (shape, axes): the kernel the layout selects
AXES_CASES = [((96, 64), [0], "columns, split reduced axis + combine"),
              ((24, 65536), [0], "columns, one chunk"),
              ((7, 33, 20), [1], "columns over a middle axis"),
              ((64, 96), [1], "rows (last axis)"),
              ((300, 7), [1], "rows, short"),
              ((8, 37, 5), [1], "generic (inner extent not a vector multiple)"),
              ((6, 7, 8), [0, 2], "generic (reduced axes not adjacent)"),
              ((5, 6, 7), [0, 1, 2], "columns over all axes")]


@pytest.mark.parametrize("shape,ax,kind", AXES_CASES, ids=[c[2] for c in AXES_CASES])
def test_axis_extremes_nan_signed_zero(dev, shape, ax, kind):
    capi, tapes, ti = _c()
    for positive, ops in ((False, ((capi.REDUCE_MAX, 1), (capi.REDUCE_ARGMAX, 3), (capi.REDUCE_ARGMIN, 4))),
                          (True, ((capi.REDUCE_MIN, 2), (capi.REDUCE_ARGMAX, 3), (capi.REDUCE_ARGMIN, 4)))):
        x = nan_zero_data(shape, ax, seed=len(shape) * 7 + shape[-1], positive=positive)
        xd = torch.from_numpy(x).to(dev)
        for op, oop in ops:
            got = _axes(xd, ax, op)
            ref = O.reduce_axes_f32(x, ax, oop)
            if oop <= 2:
                assert np.array_equal(_bits(got), _bits(ref)), (kind, op)
                if ref.size >= 16:   # the zero that loses a mixed tie is an extremum somewhere
                    assert np.any(_bits(ref) == (0x80000000 if not positive else 0))
            else:
                assert np.array_equal(got, ref), (kind, op)


@pytest.mark.parametrize("n", [1, 2, 63, 1000, 65537, 1 << 21])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16])
def test_full_extremes_nan_signed_zero(dev, n, dtype):
    capi, tapes, ti = _c()
    for positive in (False, True):
        for variant in range(3):
            x = nan_zero_data((n,), [0], seed=n + variant, positive=positive)
            if variant == 1:
                x[x == 0] = -0.0          # only -0 zeros
            elif variant == 2:
                x[x == 0] = 0.0           # only +0 zeros
            xd = torch.from_numpy(x).to(dev).to(dtype)
            # the stored values (bf16 keeps NaN, ±0, ±1, ±2) and the dtype's identities
            xs = xd.cpu().numpy() if dtype != torch.bfloat16 else xd.double().cpu().numpy()
            lim = O.BF16_MAX if dtype == torch.bfloat16 else None
            vmax, imax = O.max_fold(xs, None if lim is None else -lim)
            vmin, imin = O.min_fold(xs, lim)
            got_max = _full(xd, capi.REDUCE_MAX).double().numpy()
            got_min = _full(xd, capi.REDUCE_MIN).double().numpy()
            assert _bits(got_max) == _bits(np.float64(vmax)), (n, dtype, positive, variant)
            assert _bits(got_min) == _bits(np.float64(vmin)), (n, dtype, positive, variant)
            assert _full(xd, capi.REDUCE_ARGMAX).item() == imax
            assert _full(xd, capi.REDUCE_ARGMIN).item() == imin


def _copy_tape(capi):
    return capi.Tape([P(LOAD, 0)], [capi.FLOAT32], [], capi.FLOAT32, capi.FLOAT32)


@pytest.mark.parametrize("n", [5, 4099, 1 << 20])
@pytest.mark.parametrize("strided", [False, True])
def test_fused_expression_extremes(dev, n, strided):
    """The expression kernel's fused max/min (contiguous and strided forms) on stored values."""
    capi, tapes, ti = _c()
    tape = _copy_tape(capi)
    for variant in range(3):
        x = nan_zero_data((n,), [0], seed=3 * n + variant, positive=bool(variant & 1))
        if variant == 2:
            x[x == 0] = -0.0
        base = torch.from_numpy(np.repeat(x, 2) if strided else x).to(dev)
        src = base[::2] if strided else base
        out = torch.empty(n, dtype=torch.float32, device=dev)
        mx = torch.empty((), dtype=torch.float32, device=dev)
        mn = torch.empty((), dtype=torch.float32, device=dev)
        tape.launch_reduce([ti.view(src)], ti.view(out), capi.FUSE_MAX | capi.FUSE_MIN, capi.ACCUM_PLAIN,
                           None, mx.data_ptr(), mn.data_ptr())
        assert np.array_equal(_bits(out.cpu().numpy()), _bits(x))
        assert _bits(mx.cpu().numpy()) == _bits(np.float32(O.max_fold(x)[0])), (n, strided, variant)
        assert _bits(mn.cpu().numpy()) == _bits(np.float32(O.min_fold(x)[0])), (n, strided, variant)


def test_batch_extremes(dev):
    """Per-item fused max/min of the ragged batch on NaN / signed-zero items."""
    from ncarray_amd import batch
    capi, tapes, ti = _c()
    tape = _copy_tape(capi)
    sizes = [1, 2, 5, 63, 4097, 70001]
    xs = []
    for i, n in enumerate(sizes):
        x = nan_zero_data((n,), [0], seed=100 + i, positive=bool(i & 1))
        if i % 3 == 2:
            x[x == 0] = -0.0
        xs.append(x)
    ins = [torch.from_numpy(x).to(dev) for x in xs]
    rb = batch.RaggedBatch(tape, [ins], None, fuse=capi.FUSE_MAX | capi.FUSE_MIN, tile=1024, vec=4, waves_per_cu=8)
    rb.launch()
    torch.cuda.synchronize()
    for i, x in enumerate(xs):
        assert _bits(rb.maxs[i].cpu().numpy()) == _bits(np.float32(O.max_fold(x)[0])), i
        assert _bits(rb.mins[i].cpu().numpy()) == _bits(np.float32(O.min_fold(x)[0])), i
    rb.close()


@pytest.mark.parametrize("shape,form", [((16, 8, 64), "march"), ((9, 8, 62), "tile")])
def test_stencil_fused_extremes(dev, shape, form):
    """Fused max/min of an identity stencil (the stored values are the input's, so the data
    decides every tie) against the oracle's fold of the stored output."""
    capi, tapes, ti = _c()
    st = capi.Stencil([[0, 0, 0]], [P(LOAD, 0)], capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    capi.profile_enable(True)
    try:
        capi.profile_read()
        for variant in range(4):
            x = nan_zero_data(shape, [0, 1, 2], seed=variant + shape[-1], positive=bool(variant & 1))
            if variant >= 2:
                x[x == 0] = -0.0 if variant == 2 else 0.0
            xd = torch.from_numpy(x).to(dev)
            out = torch.empty_like(xd)
            mx = torch.empty((), dtype=torch.float32, device=dev)
            mn = torch.empty((), dtype=torch.float32, device=dev)
            st.apply_reduce(ti.view(xd), ti.view(out), capi.BOUNDARY_ZERO, fuse=capi.FUSE_MAX | capi.FUSE_MIN,
                            mode=capi.ACCUM_COMPENSATED, max_out=mx.data_ptr(), min_out=mn.data_ptr())
            stored = out.cpu().numpy()
            assert np.array_equal(_bits(stored), _bits(x))
            assert _bits(mx.cpu().numpy()) == _bits(np.float32(O.max_fold(stored)[0])), variant
            assert _bits(mn.cpu().numpy()) == _bits(np.float32(O.min_fold(stored)[0])), variant
        names = [k for k, _ in capi.profile_read()]
    finally:
        capi.profile_enable(False)
    assert any(form in k for k in names), names
