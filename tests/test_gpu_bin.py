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
