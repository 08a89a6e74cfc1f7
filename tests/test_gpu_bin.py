"""Binning by an index map (scatter_add semantics) on the GPU vs the oracle."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _c():
    from ncarray_amd import capi, torch_interop as ti
    return capi, ti


def make_frames(F, ny, nx, seed=0, zero_frac=0.01):
    rng = np.random.default_rng(seed)
    fr = rng.exponential(50.0, (F, ny, nx)).astype(np.float32)
    fr[rng.random((F, ny, nx)) < zero_frac] = 0.0
    return fr


def run_bin(capi, ti, frames_d, map_d, nb, pred, c64=False, s64=False):
    F = frames_d.shape[0]
    counts = torch.empty((F, nb), dtype=torch.int64 if c64 else torch.int32, device=frames_d.device)
    sums = torch.empty((F, nb), dtype=torch.float64 if s64 else torch.float32, device=frames_d.device)
    capi.call("nca_hip_bin", ctypes.byref(ti.view(frames_d)), ctypes.byref(ti.view(map_d)), nb, pred,
              ctypes.byref(ti.view(counts)), ctypes.byref(ti.view(sums)), None)
    return counts.cpu().numpy(), sums.cpu().numpy()


@pytest.mark.parametrize("pred", [0, 1])
def test_radial_bins_small(dev, pred):
    capi, ti = _c()
    F, ny, nx, nb = 8, 64, 64, 100
    fr = make_frames(F, ny, nx, seed=pred)
    bm = O.radial_binmap(ny, nx, 32, 32, nb)
    c, s = run_bin(capi, ti, torch.from_numpy(fr).to(dev), torch.from_numpy(bm).to(dev), nb, pred)
    rc, rs = O.bin_f32(fr, bm, nb, nonzero=bool(pred))
    assert np.array_equal(c, rc)
    for f in range(F):  # np.bincount pins the oracle
        w = fr[f].ravel()
        sel = (w != 0) if pred else np.ones_like(w, bool)
        assert np.array_equal(np.bincount(bm.ravel()[sel], minlength=nb), rc[f])
        assert np.allclose(np.bincount(bm.ravel(), weights=w.astype(np.float64), minlength=nb), rs[f], rtol=1e-12)
    # float64 accumulation: the float32 result is the correctly rounded FP64 sum (<= 1 ulp)
    assert np.all(np.abs(s - rs) <= np.spacing(np.abs(rs).astype(np.float32)))


def test_out_of_range_skipped(dev):
    capi, ti = _c()
    nb = 10
    bm = np.array([-5, 0, 3, 10, 9, 12345, -1, 3] * 32, dtype=np.int32)
    fr = np.arange(2 * bm.size, dtype=np.float32).reshape(2, bm.size) + 1
    c, s = run_bin(capi, ti, torch.from_numpy(fr).to(dev), torch.from_numpy(bm).to(dev), nb, 0, True, True)
    rc, rs = O.bin_f32(fr, bm, nb, nonzero=False)
    assert np.array_equal(c, rc) and np.array_equal(s, rs)


def test_c2_full_size(dev):
    """256 frames of 1024x1024, 1% zeros, 1000 radial bins: counts bit-exact."""
    capi, ti = _c()
    F, ny, nx, nb = 256, 1024, 1024, 1000
    g = torch.Generator(device=dev).manual_seed(3)
    fr = -50.0 * torch.log1p(-torch.rand((F, ny, nx), generator=g, device=dev))
    fr = torch.where(torch.rand((F, ny, nx), generator=g, device=dev) < 0.01, torch.zeros_like(fr), fr).float()
    bm = O.radial_binmap(ny, nx, 512, 512, nb)
    c, s = run_bin(capi, ti, fr, torch.from_numpy(bm).to(dev), nb, 1)
    frn = fr.cpu().numpy()
    rc, rs = O.bin_f32(frn, bm, nb, nonzero=True)
    assert np.array_equal(c, rc)
    assert c.sum() == int((frn != 0).sum())
    assert np.all(np.abs(s - rs) <= np.spacing(np.abs(rs).astype(np.float32)))


def test_backward_gather(dev):
    capi, ti = _c()
    F, ny, nx, nb = 4, 32, 48, 50
    bm = O.radial_binmap(ny, nx, 16, 24, nb)
    bm[0, :5] = -1
    bm[1, :3] = 999
    g = np.random.default_rng(2).standard_normal((F, nb)).astype(np.float32)
    out = torch.empty((F, ny, nx), dtype=torch.float32, device=dev)
    capi.call("nca_hip_bin_backward", ctypes.byref(ti.view(torch.from_numpy(g).to(dev))),
              ctypes.byref(ti.view(torch.from_numpy(bm).to(dev))), ctypes.byref(ti.view(out)), None)
    ref = O.bin_backward_f32(g, bm)
    assert np.array_equal(out.cpu().numpy().reshape(F, -1), ref)


def test_scatter_add_reference(dev):
    """dest[idx[i]] += src[i], out-of-range skipped, accumulating into dest."""
    capi, ti = _c()
    rng = np.random.default_rng(5)
    n, nd = 5000, 37
    idx = rng.integers(-3, nd + 3, n).astype(np.int32)
    src = rng.integers(-8, 8, n).astype(np.float32)      # exact in float32 in any order
    dest0 = rng.integers(0, 4, nd).astype(np.float32)
    d = torch.from_numpy(dest0).to(dev)
    capi.call("nca_hip_scatter_add", ctypes.byref(ti.view(d)), ctypes.byref(ti.view(torch.from_numpy(idx).to(dev))),
              ctypes.byref(ti.view(torch.from_numpy(src).to(dev))), None)
    assert np.array_equal(d.cpu().numpy(), O.scatter_add_f32(dest0, idx, src))


def _bin_outputs(capi, ti, frames_d, map_d, nb, pred, want_counts=True, want_sums=True):
    F = frames_d.shape[0] if frames_d.dim() > 1 else 1
    shape = (F, nb) if frames_d.dim() > 1 else (nb,)
    counts = torch.full(shape, -7, dtype=torch.int64, device=frames_d.device) if want_counts else None
    sums = torch.full(shape, -7.0, dtype=torch.float64, device=frames_d.device) if want_sums else None
    capi.call("nca_hip_bin", ctypes.byref(ti.view(frames_d)), ctypes.byref(ti.view(map_d)), nb, pred,
              ctypes.byref(ti.view(counts)) if want_counts else None,
              ctypes.byref(ti.view(sums)) if want_sums else None, None)
    return (counts.cpu().numpy() if want_counts else None), (sums.cpu().numpy() if want_sums else None)


@pytest.mark.parametrize("pred", [0, 1])
@pytest.mark.parametrize("want", ["counts", "sums", "both"])
def test_counts_or_sums_alone(dev, pred, want):
    """Each output alone equals the same output of the combined call (and the oracle); with
    COUNT_ALL and no sums the kernel counts from the map only."""
    capi, ti = _c()
    F, ny, nx, nb = 21, 40, 52, 37  # F not a multiple of the valid-count groups
    fr = make_frames(F, ny, nx, seed=11, zero_frac=0.05)
    bm = O.radial_binmap(ny, nx, 17, 30, nb)
    bm[3, :7] = -1
    bm[9, 4:9] = nb + 2
    c, s = _bin_outputs(capi, ti, torch.from_numpy(fr).to(dev), torch.from_numpy(bm).to(dev), nb, pred,
                        want in ("counts", "both"), want in ("sums", "both"))
    rc, rs = O.bin_f32(fr, bm, nb, nonzero=bool(pred))
    if c is not None:
        assert np.array_equal(c, rc)
    if s is not None:
        assert np.all(np.abs(s - rs) <= 1e-12 * np.abs(rs) + 1e-300)


@pytest.mark.parametrize("tdt,mdt", [(torch.bfloat16, torch.int32), (torch.float16, torch.int64),
                                     (torch.int32, torch.int32), (torch.float64, torch.int64),
                                     (torch.float32, torch.int64)])
def test_frame_and_map_dtypes(dev, tdt, mdt):
    capi, ti = _c()
    F, ny, nx, nb = 6, 48, 40, 29
    rng = np.random.default_rng(4)
    if tdt == torch.int32:
        frn = rng.integers(-3, 50, (F, ny, nx)).astype(np.int32)
        frd = torch.from_numpy(frn).to(dev)
        vals = frn.astype(np.float64)
    else:
        fr = make_frames(F, ny, nx, seed=5, zero_frac=0.03)
        frd = torch.from_numpy(fr).to(dev).to(tdt)  # float32 -> float64 is exact
        vals = frd.double().cpu().numpy()
    bm = O.radial_binmap(ny, nx, 20, 21, nb)
    c, s = _bin_outputs(capi, ti, frd, torch.from_numpy(bm).to(dev).to(mdt), nb, 1)
    m = bm.ravel()
    for f in range(F):
        w = vals[f].ravel()
        assert np.array_equal(c[f], np.bincount(m[w != 0], minlength=nb))
        ref = np.bincount(m, weights=w, minlength=nb)
        assert np.allclose(s[f], ref, rtol=1e-12, atol=0)


def test_single_frame_ragged_pixels(dev):
    """One 1-D frame whose pixel count is not a multiple of 4 (scalar tail paths)."""
    capi, ti = _c()
    P, nb = 4099 * 3 + 1, 13
    rng = np.random.default_rng(8)
    fr = rng.exponential(2.0, P).astype(np.float32)
    fr[rng.random(P) < 0.1] = 0.0
    bm = rng.integers(-2, nb + 2, P).astype(np.int32)
    c, s = _bin_outputs(capi, ti, torch.from_numpy(fr).to(dev), torch.from_numpy(bm).to(dev), nb, 1)
    rc, rs = O.bin_f32(fr.reshape(1, P), bm, nb, nonzero=True)
    assert np.array_equal(c, rc[0])
    assert np.all(np.abs(s - rs[0]) <= 1e-12 * np.abs(rs[0]))


def test_nan_and_signed_zero_predicate(dev):
    """NaN is a nonzero pixel (counted, and it makes its bin's sum NaN); -0.0 and +0.0 are zeros."""
    capi, ti = _c()
    nb = 4
    bm = np.array([0, 0, 1, 1, 2, 2, 3, 3] * 4, dtype=np.int32)
    fr = np.ones((2, bm.size), dtype=np.float32)
    fr[0, 0] = np.nan
    fr[0, 2] = -0.0
    fr[1, 3] = 0.0
    fr[1, 4] = np.inf
    c, s = _bin_outputs(capi, ti, torch.from_numpy(fr).to(dev), torch.from_numpy(bm).to(dev), nb, 1)
    rc, rs = O.bin_f32(fr, bm, nb, nonzero=True)
    assert np.array_equal(c, rc)
    assert np.array_equal(np.isnan(s), np.isnan(rs))
    fin = ~np.isnan(rs)
    assert np.array_equal(s[fin], rs[fin])


@pytest.mark.parametrize("fpb", ["2", "4"])
def test_frames_per_block_override(dev, monkeypatch, fpb):
    """Blocks that share the map vector over 2 or 4 frames give the same results."""
    capi, ti = _c()
    monkeypatch.setenv("NCA_HIP_BIN_FRAMES", fpb)
    F, ny, nx, nb = 11, 64, 64, 50
    fr = make_frames(F, ny, nx, seed=9)
    bm = O.radial_binmap(ny, nx, 32, 32, nb)
    c, s = _bin_outputs(capi, ti, torch.from_numpy(fr).to(dev), torch.from_numpy(bm).to(dev), nb, 1)
    rc, rs = O.bin_f32(fr, bm, nb, nonzero=True)
    assert np.array_equal(c, rc)
    assert np.all(np.abs(s - rs) <= 1e-12 * np.abs(rs))


def test_frame_offsets_beyond_2_31(dev):
    """2049 uint16 frames of 2^20 pixels (2.15e9 pixels): the last frames start past element 2^31,
    and their counts and sums are exact (closed forms: bin = p mod 10, every pixel 1 except a
    marked block of the last frame)."""
    capi, ti = _c()
    F, P, nb = 2049, 1 << 20, 10
    fr = torch.ones((F, P), dtype=torch.uint16, device=dev)
    fr[F - 1, 1000:1010] = 5          # pixels 1000..1009: one of each bin, +4 each
    fr[F - 2, 7] = 0                  # a zero pixel in bin 7 of frame F-2
    bm = (torch.arange(P, device=dev, dtype=torch.int32) % nb).contiguous()
    counts = torch.empty((F, nb), dtype=torch.int64, device=dev)
    sums = torch.empty((F, nb), dtype=torch.float64, device=dev)
    capi.call("nca_hip_bin", ctypes.byref(ti.view(fr)), ctypes.byref(ti.view(bm)), nb, capi.COUNT_NONZERO,
              ctypes.byref(ti.view(counts)), ctypes.byref(ti.view(sums)), None)
    c, s = counts.cpu().numpy(), sums.cpu().numpy()
    per_bin = np.bincount(np.arange(P) % nb, minlength=nb)
    assert np.array_equal(c[0], per_bin) and np.array_equal(s[0], per_bin.astype(np.float64))
    assert np.array_equal(c[F - 1], per_bin) and np.array_equal(s[F - 1], per_bin + 4.0)
    want = per_bin.copy()
    want[7] -= 1
    assert np.array_equal(c[F - 2], want) and np.array_equal(s[F - 2], want.astype(np.float64))
    assert np.array_equal(c[1:F - 2], np.broadcast_to(per_bin, (F - 3, nb)))
