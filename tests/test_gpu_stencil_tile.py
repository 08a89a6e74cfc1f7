"""The tile stencil kernel (every stencil without hooks that the 3-D march does not take: 1-D and
2-D stencils, steps other than 1, strided views, NC pointer-axis and SO sources and
destinations other than tables of C-contiguous planes) and the direct kernel's pointer chains,
against the oracle.

Stencil values are bit-exact against oracle.stencil_f32 (the same float32 operations in tape
order) on the dense equivalent of each view (numpy slicing, stacked rows, oracle.so_gather);
fused sums are bounded against FP64, fused max/min bit-exact.  Reference: Stencil<NDIM>::create
with is_pointer_axis / is_soarr (expression/stencil.hh:76-81) and the pointer-prefix loads of
the generated kernel (jit/device/rtcompiler.hh:247-308)."""
import ctypes

import numpy as np
import pytest
import torch

from ncarray_amd import capi, tapes
from ncarray_amd import torch_interop as ti
from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = capi.pack
LAP2 = [[0, 0], [-1, 0], [1, 0], [0, -1], [0, 1]]


@pytest.fixture(autouse=True)
def _device(dev):
    return dev


def five_point():
    """u1 + u2 + u3 + u4 - s0 * u0 (the 2-D Laplacian with s0 = 4)."""
    return [P(capi.OP_LOAD_NCARR, 1), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 3),
            P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 4), P(capi.OP_ADD), P(capi.OP_LOAD_CONST, 0),
            P(capi.OP_LOAD_NCARR, 0), P(capi.OP_MUL), P(capi.OP_SUB)]


def sum_tape(n):
    t = [P(capi.OP_LOAD_NCARR, 0)]
    for k in range(1, n):
        t += [P(capi.OP_LOAD_NCARR, k), P(capi.OP_ADD)]
    return t


def kernels_of(fn):
    capi.profile_enable(True)
    try:
        capi.profile_read()
        fn()
        torch.cuda.synchronize()
        return [k for k, _ in capi.profile_read()]
    finally:
        capi.profile_enable(False)


def bits(a):
    return np.asarray(a).view(np.uint32)


@pytest.mark.parametrize("shape", [(1, 1), (5, 7), (33, 130), (100, 257), (16, 128), (300, 64)])
def test_2d_laplacian(shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    x = rng.standard_normal(shape).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    out = torch.full_like(xd, np.nan)
    st = capi.Stencil(LAP2, five_point(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32])
    ks = kernels_of(lambda: st.apply(ti.view(xd), ti.view(out), capi.BOUNDARY_ZERO, scalars=(4.0,)))
    assert ks == ["nca_stencil_tile"], ks
    ref = O.stencil_f32(x, shape, LAP2, five_point(), [4.0], True)
    assert np.array_equal(bits(out.cpu().numpy()), bits(ref))


def test_2d_no_boundary_on_padded_source():
    """The reference contract: no bounds handling; a (H+2, W+2) source and a (H, W) destination
    with the offsets shifted by one read only inside the source."""
    rng = np.random.default_rng(1)
    H, W = 70, 133
    x = rng.standard_normal((H + 2, W + 2)).astype(np.float32)
    offs = [[o[0] + 1, o[1] + 1] for o in LAP2]
    st = capi.Stencil(offs, five_point(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32])
    xd = torch.from_numpy(x).cuda()
    out = torch.empty((H, W), device="cuda")
    st.apply(ti.view(xd), ti.view(out), capi.BOUNDARY_NONE, scalars=(4.0,))
    ref = O.stencil_f32(x, (H, W), offs, five_point(), [4.0], False)
    assert np.array_equal(bits(out.cpu().numpy()), bits(ref))


@pytest.mark.parametrize("n", [1, 3, 100, 5000])
def test_1d_three_point(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    st = capi.Stencil([[-1], [0], [1]], sum_tape(3), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    xd = torch.from_numpy(x).cuda()
    out = torch.empty_like(xd)
    ks = kernels_of(lambda: st.apply(ti.view(xd), ti.view(out), capi.BOUNDARY_ZERO))
    assert ks == ["nca_stencil_tile"], ks
    ref = O.stencil_f32(x, (n,), [[-1], [0], [1]], sum_tape(3), [], True)
    assert np.array_equal(bits(out.cpu().numpy()), bits(ref))


@pytest.mark.parametrize("src_shape,window", [((512, 1024), 2), ((96, 130), 2), ((32, 64, 128), 2), ((33, 90), 3)],
                         ids=["2d_512x1024", "2d_ragged", "3d_2x2x2", "2d_3x3_ragged"])
def test_pixel_binning_windows(src_shape, window):
    """Window sums with step = window (pixel binning, examples/benchmark_binning.cu:307-451):
    each output sums window^ndim source pixels; the integer division of the shapes drops a
    ragged remainder, as in the reference."""
    nd = len(src_shape)
    dshape = tuple(s // window for s in src_shape)
    rng = np.random.default_rng(sum(src_shape))
    x = rng.exponential(50.0, src_shape).astype(np.float32)
    offs = [list(o) for o in np.ndindex(*([window] * nd))]
    st = capi.Stencil(offs, sum_tape(len(offs)), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    xd = torch.from_numpy(x).cuda()
    out = torch.empty(dshape, device="cuda")
    s = torch.empty((), dtype=torch.float32, device="cuda")
    m = torch.empty((), dtype=torch.float32, device="cuda")
    ks = kernels_of(lambda: st.apply_reduce(ti.view(xd), ti.view(out), capi.BOUNDARY_NONE,
                                            fuse=capi.FUSE_SUM | capi.FUSE_MAX, mode=capi.ACCUM_COMPENSATED,
                                            sum_out=s.data_ptr(), max_out=m.data_ptr()))
    assert ks[0] == "nca_stencil_tile", ks
    ref = O.stencil_f32(x, dshape, offs, sum_tape(len(offs)), [], False)
    got = out.cpu().numpy()
    assert np.array_equal(bits(got), bits(ref))
    assert float(m) == O.max_fold(got)[0]
    s64 = got.astype(np.float64).sum()
    assert abs(float(s) - s64) <= 2 * np.finfo(np.float32).eps * np.abs(got).astype(np.float64).sum()


@pytest.mark.parametrize("dt", ["float32", "bfloat16"])
def test_fused_reductions_and_storage_types(dt):
    tdt = torch.float32 if dt == "float32" else torch.bfloat16
    code = capi.FLOAT32 if dt == "float32" else capi.BFLOAT16
    rng = np.random.default_rng(5)
    shape = (257, 301)
    xd = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda().to(tdt)
    x = xd.float().cpu().numpy()
    st = capi.Stencil(LAP2, five_point(), code, capi.FLOAT32, code, [capi.FLOAT32])
    out = torch.empty_like(xd)
    s = torch.empty((), dtype=torch.float32, device="cuda")
    mx = torch.empty((), dtype=tdt, device="cuda")
    mn = torch.empty((), dtype=tdt, device="cuda")
    for mode in (capi.ACCUM_PLAIN, capi.ACCUM_COMPENSATED):
        st.apply_reduce(ti.view(xd), ti.view(out), capi.BOUNDARY_ZERO, fuse=capi.FUSE_SUM | capi.FUSE_MAX | capi.FUSE_MIN,
                        mode=mode, sum_out=s.data_ptr(), max_out=mx.data_ptr(), min_out=mn.data_ptr(), scalars=(4.0,))
        got = out.float().cpu().numpy()
        ref = O.stencil_f32(x, shape, LAP2, five_point(), [4.0], True)
        if dt == "float32":
            assert np.array_equal(bits(got), bits(ref))
        else:   # the float32 value rounded once to bfloat16
            assert torch.equal(out, torch.from_numpy(ref).cuda().to(torch.bfloat16))
        assert float(mx.float()) == O.max_fold(got)[0] and float(mn.float()) == O.min_fold(got)[0]
        s64 = got.astype(np.float64).sum()
        c = 8 if mode == capi.ACCUM_PLAIN else 2
        assert abs(float(s) - s64) <= c * np.finfo(np.float32).eps * np.abs(got).astype(np.float64).sum()


def test_strided_views():
    """A stepped, transposed source and a strided destination: the kernel reads the caller's
    strides."""
    rng = np.random.default_rng(6)
    base = torch.from_numpy(rng.standard_normal((90, 140)).astype(np.float32)).cuda()
    src = base[::2, 1::3].t()                 # (47, 45), strides (3, 280) elements
    dbase = torch.full((47, 90), np.nan, device="cuda")
    dst = dbase[:, ::2]
    st = capi.Stencil(LAP2, five_point(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32])
    st.apply(ti.view(src), ti.view(dst), capi.BOUNDARY_ZERO, scalars=(4.0,))
    ref = O.stencil_f32(np.ascontiguousarray(src.cpu().numpy()), tuple(src.shape), LAP2, five_point(), [4.0], True)
    got = dbase.cpu().numpy()
    assert np.array_equal(bits(got[:, ::2]), bits(ref))
    assert np.all(np.isnan(got[:, 1::2]))    # nothing between the destination's columns written


def test_extra_array_in_the_tile_kernel():
    """The Laplacian followed by the C0 expression with an extra pointwise array, on a shape the
    march does not take (last extent not a multiple of 4): the fused value against the oracle's
    Laplacian then expression (exp may differ by 2 ulp of its term)."""
    rng = np.random.default_rng(7)
    shape = (10, 20, 30)
    x = rng.standard_normal(shape).astype(np.float32)
    y = rng.standard_normal(shape).astype(np.float32)
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian_then_c0(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                      [capi.FLOAT32, capi.FLOAT32], extra_dtypes=[capi.FLOAT32])
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    out = torch.empty_like(xd)
    ks = kernels_of(lambda: st.apply(ti.view(xd), ti.view(out), capi.BOUNDARY_ZERO, extras=[ti.view(yd)],
                                     scalars=(6.0, 0.5)))
    assert ks == ["nca_stencil_tile"], ks
    a = O.stencil_f32(x, shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    ref = O.expr(tapes.fused_c0(), [a, y, a, a, y], [0.5], 10, np.float32)
    a64, y64 = a.astype(np.float64), y.astype(np.float64)
    bound = 2 * np.finfo(np.float32).eps * (np.abs(a64 * y64) + np.exp(-a64 * a64) + 0.5 * np.abs(y64))
    assert np.all(np.abs(out.cpu().numpy().astype(np.float64) - ref) <= bound)


# ---- NC pointer axes -----------------------------------------------------------------------------
def _modules(rng, n, shape, dtype=torch.float32):
    return [torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda().to(dtype) for _ in range(n)]


def _table(rows):
    return torch.tensor([r.data_ptr() for r in rows], dtype=torch.int64, device="cuda")


@pytest.mark.parametrize("boundary", [capi.BOUNDARY_ZERO, capi.BOUNDARY_NONE])
@pytest.mark.parametrize("offsets", ["in_module", "across_modules"])
def test_pointer_axis_modules(boundary, offsets):
    """A detector stack of 16 separately allocated 96 x 128 modules behind a table of module
    pointers (pointer axis 0), written into modules reached the same way: a 2-D Laplacian within
    each module, and the 3-D Laplacian whose z offsets read the neighbouring modules."""
    rng = np.random.default_rng(8)
    M, H, W = 16, 96, 128
    mods = _modules(rng, M, (H, W))
    if offsets == "in_module":
        offs = [[0, y, x] for y, x in LAP2]
        tape, sc = five_point(), (4.0,)
    else:
        offs, tape, sc = tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), (6.0,)
    sh = 1 if boundary == capi.BOUNDARY_NONE else 0     # no boundary: interior outputs, offsets shifted
    if sh:
        offs = [[o[0] + (1 if offsets == "across_modules" else 0), o[1] + 1, o[2] + 1] for o in offs]
    dshape = (M - (2 if sh and offsets == "across_modules" else 0), H - 2 * sh, W - 2 * sh)
    outs = [torch.full(dshape[1:], np.nan, device="cuda") for _ in range(dshape[0])]
    tin, tout = _table(mods), _table(outs)
    vin = capi.make_nc_view(tin.data_ptr(), capi.FLOAT32, (M, H, W), [1, W * 4, 4], 0)
    vout = capi.make_nc_view(tout.data_ptr(), capi.FLOAT32, dshape, [1, dshape[2] * 4, 4], 0)
    st = capi.Stencil(offs, tape, capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32], is_pointer_axis=[1, 0, 0])
    ks = kernels_of(lambda: st.apply(vin, vout, boundary, scalars=sc))
    # same-shape source and destination go to the register-queue march over the plane tables,
    # the shifted interior (no boundary) to the tile kernel
    assert ks == (["nca_stencil_march_float32_planessd"] if boundary == capi.BOUNDARY_ZERO else ["nca_stencil_tile"]), ks
    dense = torch.stack(mods).cpu().numpy()
    ref = O.stencil_f32(dense, dshape, offs, tape, list(sc), boundary == capi.BOUNDARY_ZERO)
    got = torch.stack(outs).cpu().numpy()
    assert np.array_equal(bits(got), bits(ref))


def test_pointer_axis_fused_and_bf16():
    """Fused sum/max/min over pointer-axis bf16 modules, with a table offset and a row offset
    (the view starts at module 2 and skips each module's first row)."""
    rng = np.random.default_rng(9)
    M, H, W = 12, 40, 96
    mods = _modules(rng, M, (H, W), torch.bfloat16)
    tbl = _table(mods)
    shape = (M - 2, H - 1, W)
    vin = capi.make_nc_view(tbl.data_ptr(), capi.BFLOAT16, shape, [1, W * 2, 2], 0, ptr_offset=2, row_offset=W * 2)
    dense = torch.stack(mods)[2:, 1:, :].float().cpu().numpy()
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.BFLOAT16, capi.FLOAT32, capi.FLOAT32,
                      [capi.FLOAT32])
    out = torch.empty(shape, device="cuda")
    s = torch.empty((), dtype=torch.float32, device="cuda")
    mx = torch.empty((), dtype=torch.float32, device="cuda")
    mn = torch.empty((), dtype=torch.float32, device="cuda")
    st.apply_reduce(vin, ti.view(out), capi.BOUNDARY_ZERO, fuse=capi.FUSE_SUM | capi.FUSE_MAX | capi.FUSE_MIN,
                    mode=capi.ACCUM_COMPENSATED, sum_out=s.data_ptr(), max_out=mx.data_ptr(), min_out=mn.data_ptr(),
                    scalars=(6.0,))
    ref = O.stencil_f32(dense, shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    got = out.cpu().numpy()
    assert np.array_equal(bits(got), bits(ref))
    assert float(mx) == O.max_fold(got)[0] and float(mn) == O.min_fold(got)[0]
    assert abs(float(s) - got.astype(np.float64).sum()) <= 2 * np.finfo(np.float32).eps * np.abs(got).sum()


def test_declared_layout_is_enforced():
    """A stencil created for pointer axis 0 refuses a dense source, and one created for dense
    sources refuses a pointer-axis source (Stencil::create's contract, stencil.hh:69-72)."""
    rng = np.random.default_rng(10)
    mods = _modules(rng, 4, (8, 64))
    tbl = _table(mods)
    vin = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (4, 8, 64), [1, 256, 4], 0)
    dense = torch.stack(mods)
    out = torch.empty_like(dense)
    ptr_st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                          [capi.FLOAT32], is_pointer_axis=[1, 0, 0])
    dense_st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                            [capi.FLOAT32], is_pointer_axis=[0, 0, 0])
    with pytest.raises(capi.NcaError):
        ptr_st.apply(ti.view(dense), ti.view(out), capi.BOUNDARY_ZERO, scalars=(6.0,))
    with pytest.raises(capi.NcaError):
        dense_st.apply(vin, ti.view(out), capi.BOUNDARY_ZERO, scalars=(6.0,))
    ptr_st.apply(vin, ti.view(out), capi.BOUNDARY_ZERO, scalars=(6.0,))
    ref = torch.empty_like(dense)
    dense_st.apply(ti.view(dense), ti.view(ref), capi.BOUNDARY_ZERO, scalars=(6.0,))
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


# ---- SO views ---------------------------------------------------------------------------------------
def _image(size):
    t = torch.zeros(size, dtype=torch.uint8, device="cuda")
    return t, O.ByteImage(size, t.data_ptr())


def _two_level(im, rng, R, H, W, pad=2):
    """A table of R tables of H row pointers (rows of W float32 after `pad` values), placed in a
    shuffled order; (base, strides, suboffsets) of the (R, H, W) SO view."""
    order = rng.permutation(R * H)
    row_at = {}
    for r in order:
        i, j = divmod(int(r), H)
        row_at[(i, j)] = im.put(rng.standard_normal(W + pad).astype(np.float32))
    sub = [im.put_ptrs([row_at[(i, j)] for j in range(H)], deltas=-8) for i in range(R)]
    top = im.put_ptrs(sub, deltas=-16)
    return im.addr + top, (8, 8, 4), (16, 8 + pad * 4, -1)


def test_so_two_level_source():
    """An SO source with two indirect axes (a table of tables of rows): per-row resolution in
    the tile kernel, bit-exact against the stencil of the gathered array."""
    rng = np.random.default_rng(11)
    R, H, W = 3, 21, 70
    t, im = _image(1 << 17)
    base, st_, so = _two_level(im, rng, R, H, W)
    t.copy_(torch.from_numpy(im.buf))
    torch.cuda.synchronize()
    dense = O.so_gather(im.buf, im.addr, base, np.float32, (R, H, W), st_, so)
    vin = capi.make_so_view(base, capi.FLOAT32, (R, H, W), st_, so)
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                      [capi.FLOAT32], is_pointer_axis=[1, 1, 0], is_soarr=True)
    out = torch.empty((R, H, W), device="cuda")
    ks = kernels_of(lambda: st.apply(vin, ti.view(out), capi.BOUNDARY_ZERO, scalars=(6.0,)))
    assert ks == ["nca_stencil_tile"], ks
    ref = O.stencil_f32(dense, (R, H, W), tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    assert np.array_equal(bits(out.cpu().numpy()), bits(ref))


def test_so_element_pointers_use_the_direct_kernel():
    """An SO source whose last axis is indirect (one pointer per element) goes through the direct
    kernel's pointer chain: bit-exact, with zero boundary."""
    rng = np.random.default_rng(12)
    H, W = 9, 17
    t, im = _image(1 << 16)
    vals = rng.standard_normal((H, W)).astype(np.float32)
    order = rng.permutation(H * W)
    at = {}
    for e in order:
        at[int(e)] = im.put(vals.reshape(-1)[int(e):int(e) + 1], align=4)
    tbl = im.put_ptrs([at[e] for e in range(H * W)])
    t.copy_(torch.from_numpy(im.buf))
    torch.cuda.synchronize()
    vin = capi.make_so_view(im.addr + tbl, capi.FLOAT32, (H, W), (W * 8, 8), (-1, 0))
    assert np.array_equal(O.so_gather(im.buf, im.addr, im.addr + tbl, np.float32, (H, W), (W * 8, 8), (-1, 0)), vals)
    st = capi.Stencil(LAP2, five_point(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32])
    out = torch.empty((H, W), device="cuda")
    ks = kernels_of(lambda: st.apply(vin, ti.view(out), capi.BOUNDARY_ZERO, scalars=(4.0,)))
    assert ks == ["nca_stencil_direct"], ks
    ref = O.stencil_f32(vals, (H, W), LAP2, five_point(), [4.0], True)
    assert np.array_equal(bits(out.cpu().numpy()), bits(ref))
