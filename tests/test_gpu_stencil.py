"""Stencils on the GPU through the C ABI vs the oracle (reference stencil semantics,
float32 op order) and FP64."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _c():
    from ncarray_amd import capi, tapes, torch_interop as ti
    return capi, tapes, ti


def lap_stencil():
    capi, tapes, ti = _c()
    return capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                        [capi.FLOAT32])


@pytest.mark.parametrize("shape", [(1, 1, 4), (5, 7, 12), (33, 17, 64), (16, 70, 132), (64, 64, 64)])
def test_laplacian_bit_exact(dev, shape):
    capi, tapes, ti = _c()
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape).astype(np.float32)
    xd = torch.from_numpy(x).to(dev)
    out = torch.empty_like(xd)
    st = lap_stencil()
    st.apply(ti.view(xd), ti.view(out), capi.BOUNDARY_ZERO, scalars=(6.0,))
    ref = O.stencil_f32(x, shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    f64 = O.laplacian_fp64(x)
    assert np.max(np.abs(out.cpu().numpy() - f64)) <= 8 * np.finfo(np.float32).eps * np.max(np.abs(x)) * 2


def test_laplacian_direct_path_matches(dev):
    """A shape the march kernel does not take (last extent % 4 != 0) runs the tile kernel."""
    capi, tapes, ti = _c()
    shape = (9, 10, 11)
    x = np.random.default_rng(0).standard_normal(shape).astype(np.float32)
    xd = torch.from_numpy(x).to(dev)
    out = torch.empty_like(xd)
    lap_stencil().apply(ti.view(xd), ti.view(out), capi.BOUNDARY_ZERO, scalars=(6.0,))
    ref = O.stencil_f32(x, shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_padded_reference_contract(dev):
    """Reference usage: zero-padded src (n+2)^3 -> dest n^3, offsets 0..2, no bounds checks."""
    capi, tapes, ti = _c()
    n = 20
    x = np.random.default_rng(4).standard_normal((n, n, n)).astype(np.float32)
    xp = np.pad(x, 1)
    offs = [[o[0] + 1, o[1] + 1, o[2] + 1] for o in tapes.LAPLACIAN_OFFSETS]
    st = capi.Stencil(offs, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32])
    src = torch.from_numpy(xp).to(dev)
    out = torch.empty((n, n, n), dtype=torch.float32, device=dev)
    st.apply(ti.view(src), ti.view(out), capi.BOUNDARY_NONE, scalars=(6.0,))
    ref = O.stencil_f32(x, x.shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    assert np.array_equal(out.cpu().numpy(), ref)
    # an out-of-range read is refused (the reference would read past the array)
    bad = torch.empty((n + 2, n + 2, n + 2), dtype=torch.float32, device=dev)
    with pytest.raises(capi.NcaError):
        st.apply(ti.view(src), ti.view(bad), capi.BOUNDARY_NONE, scalars=(6.0,))


def test_pixel_binning_stencil(dev):
    """2x2 window sum x 0.25 with step 2 (examples/benchmark_binning.cu:343-375 pattern)."""
    capi, tapes, ti = _c()
    P = capi.pack
    offs = [[0, 0], [0, 1], [1, 0], [1, 1]]
    ins = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_ADD),
           P(capi.OP_LOAD_NCARR, 3), P(capi.OP_ADD), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MUL)]
    st = capi.Stencil(offs, ins, capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32])
    x = np.random.default_rng(1).standard_normal((64, 96)).astype(np.float32)
    out = torch.empty((32, 48), dtype=torch.float32, device=dev)
    st.apply(ti.view(torch.from_numpy(x).to(dev)), ti.view(out), capi.BOUNDARY_NONE, scalars=(0.25,))
    ref = O.stencil_f32(x, (32, 48), offs, ins, [0.25], False)
    assert np.array_equal(out.cpu().numpy(), ref)
    r = (((x[0::2, 0::2] + x[0::2, 1::2]) + x[1::2, 0::2]) + x[1::2, 1::2]) * np.float32(0.25)
    assert np.array_equal(out.cpu().numpy(), r)


def test_extension_hooks(dev):
    """StencilJITExtensions (jit/device/extensions.hh:20-60): extra params + pro/epilogue."""
    capi, tapes, ti = _c()
    P = capi.pack
    st = capi.Stencil([[0]], [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MUL)], capi.FLOAT32,
                      capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32], extra_params=", unsigned* extra_ptr",
                      prologue="if ((extra_ptr != nullptr) && (*extra_ptr == 42)) { return; }",
                      epilogue="if (threadIdx.x == 0 && blockIdx.x == 0) { *extra_ptr = 42; }")
    x = torch.arange(1000, dtype=torch.float32, device=dev)
    out = torch.zeros(1000, dtype=torch.float32, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    p = ctypes.c_void_p(flag.data_ptr())
    st.apply(ti.view(x), ti.view(out), capi.BOUNDARY_NONE, scalars=(2.0,), ext_args=[ctypes.pointer(p)])
    torch.cuda.synchronize()
    assert flag.item() == 42
    # every element was either computed (2x) or skipped by the prologue (left at 0)
    o, ref = out.cpu(), (x * 2).cpu()
    assert bool(((o == ref) | (o == 0)).all())
    # a second launch sees the flag in every thread: nothing is written
    out2 = torch.full((1000,), -1.0, dtype=torch.float32, device=dev)
    st.apply(ti.view(x), ti.view(out2), capi.BOUNDARY_NONE, scalars=(2.0,), ext_args=[ctypes.pointer(p)])
    assert bool((out2 == -1.0).all())


def test_fused_axis0_and_max(dev):
    """Stencil with the axis-0 sum and global max fused == separate reductions of its output."""
    capi, tapes, ti = _c()
    shape = (96, 40, 132)
    x = np.random.default_rng(8).standard_normal(shape).astype(np.float32)
    xd = torch.from_numpy(x).to(dev)
    st = lap_stencil()
    for mode in (capi.ACCUM_PLAIN, capi.ACCUM_COMPENSATED):
        lap = torch.empty_like(xd)
        col = torch.empty(shape[1:], dtype=torch.float32, device=dev)
        mx = torch.empty((), dtype=torch.float32, device=dev)
        st.apply_reduce(ti.view(xd), ti.view(lap), capi.BOUNDARY_ZERO, axis0=ti.view(col), fuse=capi.FUSE_MAX, mode=mode,
                        max_out=mx.data_ptr(), scalars=(6.0,))
        ref_lap, ref_col, ref_mx = O.laplacian_pipeline(x)
        assert np.array_equal(lap.cpu().numpy(), ref_lap)
        assert mx.item() == ref_mx
        l64 = ref_lap.astype(np.float64)
        err = np.abs(col.cpu().numpy() - l64.sum(0)) / np.maximum(np.abs(l64).sum(0), 1e-30)
        assert err.max() <= 2e-7
        # not materialised: same colsum and max
        col2 = torch.empty_like(col)
        mx2 = torch.empty((), dtype=torch.float32, device=dev)
        st.apply_reduce(ti.view(xd), ti.shape_view(shape, capi.FLOAT32), capi.BOUNDARY_ZERO, axis0=ti.view(col2),
                        fuse=capi.FUSE_MAX, mode=mode, max_out=mx2.data_ptr(), scalars=(6.0,))
        assert torch.equal(col2, col) and mx2.item() == mx.item()


def test_adjoint_dot_product(dev):
    """<L x, g> == <x, L^T g> for the zero-Dirichlet Laplacian (FP64 tolerance)."""
    capi, tapes, ti = _c()
    shape = (24, 20, 36)
    rng = np.random.default_rng(12)
    x = rng.standard_normal(shape).astype(np.float32)
    g = rng.standard_normal(shape).astype(np.float32)
    st = lap_stencil()
    adj = st.adjoint()
    xd, gd = torch.from_numpy(x).to(dev), torch.from_numpy(g).to(dev)
    lx = torch.empty_like(xd)
    ltg = torch.empty_like(gd)
    st.apply(ti.view(xd), ti.view(lx), capi.BOUNDARY_ZERO, scalars=(6.0,))
    adj.apply(ti.view(gd), ti.view(ltg), capi.BOUNDARY_ZERO, scalars=(6.0,))
    lhs = float((lx.double() * gd.double()).sum())
    rhs = float((xd.double() * ltg.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs) + 1.0)
    # symmetric stencil: the adjoint equals the stencil up to float32 reassociation
    assert np.allclose(ltg.cpu().numpy(), O.laplacian_fp64(g), atol=1e-5)


def test_laplacian_512_pipeline_properties(dev):
    """C1 at full size (512^3): fused pipeline == unfused pipeline; a slab matches the
    oracle bit for bit; colsum within the FP64 bound."""
    capi, tapes, ti = _c()
    g = torch.Generator(device=dev).manual_seed(1)
    x = torch.randn((512, 512, 512), generator=g, device=dev, dtype=torch.float32)
    st = lap_stencil()
    lap = torch.empty_like(x)
    col = torch.empty((512, 512), dtype=torch.float32, device=dev)
    mx = torch.empty((), dtype=torch.float32, device=dev)
    st.apply_reduce(ti.view(x), ti.view(lap), capi.BOUNDARY_ZERO, axis0=ti.view(col), fuse=capi.FUSE_MAX,
                    mode=capi.ACCUM_COMPENSATED, max_out=mx.data_ptr(), scalars=(6.0,))
    lap2 = torch.empty_like(x)
    st.apply(ti.view(x), ti.view(lap2), capi.BOUNDARY_ZERO, scalars=(6.0,))
    assert torch.equal(lap, lap2)
    mx2 = torch.empty((), dtype=torch.float32, device=dev)
    capi.call("nca_hip_reduce_full", ctypes.byref(ti.view(lap2)), capi.REDUCE_MAX, 0, mx2.data_ptr(), None)
    assert mx.item() == mx2.item()
    # slab z in [250, 262) against the oracle (needs planes 249..262 of x)
    xs = x[249:263].cpu().numpy()
    ref = O.stencil_f32(xs, xs.shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    assert np.array_equal(lap[250:262].cpu().numpy(), ref[1:13])
    lapn = lap.cpu().numpy()
    ref_col = lapn.sum(axis=0, dtype=np.float64)
    err = np.abs(col.cpu().numpy() - ref_col) / np.abs(lapn).sum(axis=0, dtype=np.float64)
    assert err.max() <= 2e-7
    assert mx.item() == float(lapn.max())


@pytest.mark.parametrize("shape", [(12, 20, 128), (40, 33, 256)])
def test_laplacian_bf16_storage(dev, shape):
    """bf16 in, float32 arithmetic, bf16 out: equals the float32 oracle on the bf16 inputs,
    rounded once to bf16 (round to nearest even); compensated column sums vs FP64."""
    capi, tapes, ti = _c()
    x = np.random.default_rng(21).standard_normal(shape).astype(np.float32)
    xb = torch.from_numpy(x).to(dev).to(torch.bfloat16)
    xr = xb.float().cpu().numpy()
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.BFLOAT16, capi.FLOAT32, capi.BFLOAT16,
                      [capi.FLOAT32])
    lap = torch.empty_like(xb)
    col = torch.empty(shape[1:], dtype=torch.float32, device=dev)
    mx = torch.empty((), dtype=torch.bfloat16, device=dev)
    st.apply_reduce(ti.view(xb), ti.view(lap), capi.BOUNDARY_ZERO, axis0=ti.view(col), fuse=capi.FUSE_MAX,
                    mode=capi.ACCUM_COMPENSATED, max_out=mx.data_ptr(), scalars=(6.0,))
    ref32 = O.stencil_f32(xr, shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    ref = O.bf16_round(ref32)
    got = lap.float().cpu().numpy()
    assert np.array_equal(got, ref)
    assert mx.float().item() == float(got.max())
    g64 = got.astype(np.float64)
    err = np.abs(col.cpu().numpy() - g64.sum(0)) / np.maximum(np.abs(g64).sum(0), 1e-30)
    assert err.max() <= 2e-7


@pytest.mark.parametrize("src", ["float32", "bfloat16"])
@pytest.mark.parametrize("nz", [12, 18, 30, 40])
def test_march_short_z_runs(dev, src, nz):
    """512 x 512 planes are 128 tiles each, so one resident wave of blocks splits z into runs of
    a few planes: every run starts and ends inside the array (its halo planes come from the
    neighbouring runs and it fetches no plane past its last input plane).  Laplacian bit-exact
    against the oracle (BF16: rounded once), max exact, column sums against FP64."""
    capi, tapes, ti = _c()
    shape = (nz, 512, 512)
    x = np.random.default_rng(nz).standard_normal(shape).astype(np.float32)
    code = capi.FLOAT32 if src == "float32" else capi.BFLOAT16
    xd = torch.from_numpy(x).to(dev)
    if src == "bfloat16":
        xd = xd.to(torch.bfloat16)
    xr = xd.float().cpu().numpy()
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), code, capi.FLOAT32, code, [capi.FLOAT32])
    lap = torch.empty_like(xd)
    col = torch.empty(shape[1:], dtype=torch.float32, device=dev)
    mx = torch.empty((), dtype=xd.dtype, device=dev)
    st.apply_reduce(ti.view(xd), ti.view(lap), capi.BOUNDARY_ZERO, axis0=ti.view(col), fuse=capi.FUSE_MAX,
                    mode=capi.ACCUM_COMPENSATED, max_out=mx.data_ptr(), scalars=(6.0,))
    ref = O.stencil_f32(xr, shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    if src == "bfloat16":
        ref = O.bf16_round(ref)
    got = lap.float().cpu().numpy()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert mx.float().item() == float(got.max())
    g64 = got.astype(np.float64)
    err = np.abs(col.cpu().numpy() - g64.sum(0)) / np.maximum(np.abs(g64).sum(0), 1e-30)
    assert err.max() <= 2e-7


def _sum_tape(capi, n):
    P = capi.pack
    t = [P(capi.OP_LOAD_NCARR, 0)]
    for k in range(1, n):
        t += [P(capi.OP_LOAD_NCARR, k), P(capi.OP_ADD)]
    return t


# pure-z offsets of radius up to 3 (register z-queue march) and a z+x diagonal (LDS ring march)
PURE_Z = [[-2, 0, 0], [2, 0, 0], [-1, 0, 0], [0, -2, 0], [0, 2, 1], [0, 0, -3], [0, 0, 3], [0, -1, -1], [0, 0, 0]]
DIAG = [[1, 0, 1], [-1, 1, 0], [0, 0, -1], [0, 0, 0]]


@pytest.mark.parametrize("offs", [PURE_Z, DIAG], ids=["pure_z", "diagonal"])
@pytest.mark.parametrize("src_code", ["float32", "bfloat16"])
@pytest.mark.parametrize("shape", [(3, 37, 96), (11, 50, 128), (40, 17, 64)])
def test_march_general_offsets(dev, offs, src_code, shape):
    """Both march generators against the oracle on ragged tiles (NY not a multiple of the tile
    height, z runs shorter than the unrolled queue), bit-exact: the float32 sums run in the same
    order; bf16 sources are compared with the oracle on the bf16-rounded values."""
    capi, tapes, ti = _c()
    code = getattr(capi, src_code.upper())
    rng = np.random.default_rng(len(offs) + shape[0])
    x = rng.standard_normal(shape).astype(np.float32)
    xd = torch.from_numpy(x).to(dev)
    if src_code == "bfloat16":
        xd = xd.to(torch.bfloat16)
        x = xd.float().cpu().numpy()
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    st = capi.Stencil(offs, _sum_tape(capi, len(offs)), code, capi.FLOAT32, capi.FLOAT32)
    st.apply(ti.view(xd), ti.view(out), capi.BOUNDARY_ZERO)
    ref = O.stencil_f32(x, shape, offs, _sum_tape(capi, len(offs)), [], True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    # fused reductions on the same kernel: axis-0 column sums and the max
    col = torch.empty(shape[1:], dtype=torch.float64, device=dev)
    mx = torch.empty((), dtype=torch.float32, device=dev)
    out2 = torch.empty_like(out)
    st.apply_reduce(ti.view(xd), ti.view(out2), capi.BOUNDARY_ZERO, axis0=ti.view(col), fuse=capi.FUSE_MAX,
                    mode=capi.ACCUM_COMPENSATED, max_out=mx.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(out2, out)
    assert float(mx) == float(ref.max())
    assert np.allclose(col.cpu().numpy(), ref.astype(np.float64).sum(axis=0), rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("dt", ["float32", "bfloat16"])
@pytest.mark.parametrize("march", [True, False], ids=["march", "tile"])
def test_fused_max_min_skip_nan(dev, dt, march):
    """Fused max/min of a stencil output that holds NaN: NaN is never selected and every other
    value is still considered (maximumNumber/minimumNumber), also when the NaN and the extreme value
    sit in the same vector of one thread.  The NaN input at (4, 4, 0) makes the Laplacian NaN at
    (3, 4, 0) but not at (3, 4, 1); input -1000 at (3, 4, 1) puts the maximum +6000 + ... there
    and the minimum, -1000 + ..., at its neighbours."""
    capi, tapes, ti = _c()
    shape = (8, 8, 64) if march else (8, 8, 62)
    x = np.random.default_rng(3).standard_normal(shape).astype(np.float32)
    x[4, 4, 0] = np.nan
    x[3, 4, 1] = -1000.0
    code = capi.FLOAT32 if dt == "float32" else capi.BFLOAT16
    tdt = torch.float32 if dt == "float32" else torch.bfloat16
    xd = torch.from_numpy(x).to(dev).to(tdt)
    xs = xd.float().cpu().numpy()  # the stored input (bf16-rounded)
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), code, capi.FLOAT32, code, [capi.FLOAT32])
    lap = torch.empty_like(xd)
    mx = torch.empty((), dtype=tdt, device=dev)
    mn = torch.empty((), dtype=tdt, device=dev)
    st.apply_reduce(ti.view(xd), ti.view(lap), capi.BOUNDARY_ZERO, fuse=capi.FUSE_MAX | capi.FUSE_MIN,
                    mode=capi.ACCUM_COMPENSATED, max_out=mx.data_ptr(), min_out=mn.data_ptr(), scalars=(6.0,))
    stored = lap.float().cpu().numpy()
    assert np.isnan(stored[3, 4, 0]) and not np.isnan(stored[3, 4, 1])
    ref = O.stencil_f32(xs, shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    if dt == "float32":
        assert np.array_equal(stored, ref, equal_nan=True)
    assert mx.float().item() == np.nanmax(stored) and mx.float().item() > 5000
    assert mn.float().item() == np.nanmin(stored)
