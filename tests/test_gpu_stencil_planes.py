"""The register-queue march over tables of plane pointers: a 3-D stencil whose source and/or
destination is an NC view with pointer axis 0 or an SO view indirect at axis 0, every plane
C-contiguous (a detector stack of separately allocated modules).  Each plane's base is read from
the table once per plane; the kernel is otherwise the dense march.

Stencil values are bit-exact against oracle.stencil_f32 on the dense stack (the same float32
operations in tape order) and against the dense march on the same data; fused axis-0 sums are
bounded against FP64, max bit-exact.  Reference: Stencil<NDIM>::create with is_pointer_axis /
is_soarr (expression/stencil.hh:76-81) and the pointer-prefix loads of the generated kernel
(jit/device/rtcompiler.hh:247-308)."""
import numpy as np
import pytest
import torch

from ncarray_amd import capi, tapes
from ncarray_amd import torch_interop as ti
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _device(dev):
    return dev


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
    a = np.asarray(a)
    return a.view(np.uint16 if a.dtype.itemsize == 2 else np.uint32)


def planes(rng, n, shape, dtype):
    """n separately allocated planes, allocated in a shuffled order (so that the table, not the
    addresses, gives the stacking order)."""
    vals = [rng.standard_normal(shape).astype(np.float32) for _ in range(n)]
    out = [None] * n
    for i in rng.permutation(n):
        out[int(i)] = torch.from_numpy(vals[int(i)]).cuda().to(dtype)
    return out


def table(rows):
    return torch.tensor([r.data_ptr() for r in rows], dtype=torch.int64, device="cuda")


def nc_planes(tbl, code, shape, item, **kw):
    """An NC view (pointer axis 0) over the plane table `tbl`; the view keeps the table alive."""
    v = capi.make_nc_view(tbl.data_ptr(), code, shape, [1, shape[2] * item, item], 0, **kw)
    v._owner = tbl
    return v


def lap_stencil(code, dst_code, pointer=True):
    return capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), code, capi.FLOAT32, dst_code, [capi.FLOAT32],
                        is_pointer_axis=[1, 0, 0] if pointer else None)


CASES = [(19, 37, 72), (8, 8, 256), (64, 96, 256), (3, 1, 16)]


@pytest.mark.parametrize("shape", CASES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("dst_table", [False, True])
def test_plane_table_march(shape, dtype, dst_table):
    rng = np.random.default_rng(shape[0] * 131 + shape[2] + (7 if dst_table else 0))
    M, H, W = shape
    code = ti.TORCH_TO_CODE[dtype]
    item = torch.tensor([], dtype=dtype).element_size()
    mods = planes(rng, M, (H, W), dtype)
    vin = nc_planes(table(mods), code, shape, item)
    st = lap_stencil(code, code)
    if dst_table:
        outs = [torch.full((H, W), float("nan"), device="cuda").to(dtype) for _ in range(M)]
        vout = nc_planes(table(outs), code, shape, item)
        gather = lambda: torch.stack(outs)  # noqa: E731
    else:
        out = torch.full(shape, float("nan"), device="cuda").to(dtype)
        vout = ti.view(out)
        gather = lambda: out  # noqa: E731
    ks = kernels_of(lambda: st.apply(vin, vout, capi.BOUNDARY_ZERO, scalars=(6.0,)))
    assert len(ks) == 1 and ks[0].startswith("nca_stencil_march_") and ks[0].endswith("_planess" + ("d" if dst_table else "")), ks
    dense = torch.stack(mods)
    ref = O.stencil_f32(dense.float().cpu().numpy(), shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    got = gather().float().cpu().numpy()
    # bf16 storage: the float32 stencil value rounded to nearest even
    assert np.array_equal(bits(got), bits(ref if dtype == torch.float32 else O.bf16_round(ref)))
    # the dense march on the stacked planes gives the same bits
    dd = torch.empty_like(dense)
    lap_stencil(code, code, pointer=False).apply(ti.view(dense), ti.view(dd), capi.BOUNDARY_ZERO, scalars=(6.0,))
    torch.cuda.synchronize()
    assert torch.equal(gather().view(torch.int16 if item == 2 else torch.int32), dd.view(torch.int16 if item == 2 else torch.int32))


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_plane_table_fused_axis0_and_max(dtype):
    """The fused C1 step (Laplacian, axis-0 sum, max) over a plane table: column sums within the
    FP32 compensated bound of FP64, max exact, stencil identical to the dense march's."""
    rng = np.random.default_rng(21)
    M, H, W = 48, 40, 128
    code = ti.TORCH_TO_CODE[dtype]
    item = torch.tensor([], dtype=dtype).element_size()
    mods = planes(rng, M, (H, W), dtype)
    vin = nc_planes(table(mods), code, (M, H, W), item)
    st = lap_stencil(code, code)
    out = torch.empty((M, H, W), device="cuda").to(dtype)
    col = torch.empty((H, W), device="cuda")
    mx = torch.empty((), dtype=dtype, device="cuda")  # the max is stored in the value dtype
    ks = kernels_of(lambda: st.apply_reduce(vin, ti.view(out), capi.BOUNDARY_ZERO, axis0=ti.view(col), fuse=capi.FUSE_MAX,
                                            mode=capi.ACCUM_COMPENSATED, max_out=mx.data_ptr(), scalars=(6.0,)))
    assert ks[0].endswith("_axis0_max_planess"), ks
    dense = torch.stack(mods)
    dd, dcol, dmx = torch.empty_like(dense), torch.empty_like(col), torch.empty_like(mx)
    lap_stencil(code, code, pointer=False).apply_reduce(ti.view(dense), ti.view(dd), capi.BOUNDARY_ZERO,
                                                        axis0=ti.view(dcol), fuse=capi.FUSE_MAX,
                                                        mode=capi.ACCUM_COMPENSATED, max_out=dmx.data_ptr(),
                                                        scalars=(6.0,))
    torch.cuda.synchronize()
    assert torch.equal(out.float(), dd.float()) and torch.equal(col, dcol) and torch.equal(mx, dmx)
    stored = out.double()
    s64 = stored.sum(0)
    assert float((col.double() - s64).abs().max()) <= float(2 * np.finfo(np.float32).eps * stored.abs().sum(0).max())
    assert float(mx.float()) == float(out.float().max())


def test_plane_table_reversed_strided_and_offset():
    """A table walked backwards (stride -1 from the last entry), one walked with stride 2 (every
    other entry), and a row offset (the view starts one row into each plane): the same bits as
    the dense march on the planes in that order."""
    rng = np.random.default_rng(22)
    M, H, W = 20, 33, 64
    mods = planes(rng, 2 * M, (H + 1, W), torch.float32)
    tbl = table(mods)
    st = lap_stencil(capi.FLOAT32, capi.FLOAT32)
    dense_st = lap_stencil(capi.FLOAT32, capi.FLOAT32, pointer=False)
    for strides0, ptr_off, pick in ((-1, 2 * M - 1, list(range(2 * M - 1, M - 1, -1))), (2, 0, list(range(0, 2 * M, 2)))):
        vin = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (M, H, W), [strides0, W * 4, 4], 0, ptr_offset=ptr_off,
                                row_offset=W * 4)
        out = torch.empty((M, H, W), device="cuda")
        ks = kernels_of(lambda: st.apply(vin, ti.view(out), capi.BOUNDARY_ZERO, scalars=(6.0,)))
        assert ks == ["nca_stencil_march_float32_planess"], ks
        dense = torch.stack([mods[i][1:] for i in pick]).contiguous()
        ref = torch.empty_like(dense)
        dense_st.apply(ti.view(dense), ti.view(ref), capi.BOUNDARY_ZERO, scalars=(6.0,))
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (strides0, ptr_off)


def test_plane_table_so_view():
    """An SO view indirect at axis 0 only (suboffsets (32, -1, -1): a pointer per plane plus 32
    bytes, rows contiguous) is the same plane table."""
    rng = np.random.default_rng(23)
    M, H, W = 12, 24, 96
    flat = planes(rng, M, (H * W + 8,), torch.float32)   # 8 leading values skipped by the suboffset
    tbl = table(flat)
    vin = capi.make_so_view(tbl.data_ptr(), capi.FLOAT32, (M, H, W), (8, W * 4, 4), (32, -1, -1))
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                      [capi.FLOAT32], is_pointer_axis=[1, 0, 0], is_soarr=True)
    out = torch.empty((M, H, W), device="cuda")
    ks = kernels_of(lambda: st.apply(vin, ti.view(out), capi.BOUNDARY_ZERO, scalars=(6.0,)))
    assert ks == ["nca_stencil_march_float32_planess"], ks
    dense = torch.stack([m[8:].view(H, W) for m in flat]).contiguous()
    ref = O.stencil_f32(dense.cpu().numpy(), (M, H, W), tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    assert np.array_equal(bits(out.cpu().numpy()), bits(ref))
