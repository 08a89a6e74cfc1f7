"""Ragged batches in one persistent launch vs per-array launches; hipGraph replay;
autotuner."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _mods():
    from ncarray_amd import batch, capi, tapes, torch_interop as ti
    return batch, capi, tapes, ti


def make_items(dev, sizes, seed=0):
    g = torch.Generator(device=dev).manual_seed(seed)
    a = [torch.rand(int(n), generator=g, device=dev) * 2 - 1 for n in sizes]
    b = [torch.rand(int(n), generator=g, device=dev) * 2 - 1 for n in sizes]
    return a, b


def test_batch_matches_per_array(dev):
    batch, capi, tapes, ti = _mods()
    sizes = [1, 3, 1000, 4097, 65536, 123457, 7, 300000]
    a, b = make_items(dev, sizes)
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    # operands a, b, a, a, b: the batch sees five input lists
    outs = [torch.empty_like(x) for x in a]
    rb = batch.RaggedBatch(tape, [a, b, a, a, b], outs, fuse=capi.FUSE_SUM | capi.FUSE_MAX | capi.FUSE_MIN,
                           tile=4096, vec=4, waves_per_cu=8)
    rb.launch(scalars=(0.5,))
    torch.cuda.synchronize()
    for i, n in enumerate(sizes):
        ref = torch.empty_like(a[i])
        ops = {"a": ti.view(a[i]), "b": ti.view(b[i])}
        tape.launch([ops[k] for k in tapes.C0_OPERANDS], ti.view(ref), scalars=(0.5,))
        assert torch.equal(outs[i], ref)               # same op order: bit-identical
        r = ref.double().cpu().numpy()
        assert rb.maxs[i].item() == r.max() and rb.mins[i].item() == r.min()
        assert abs(rb.sums[i].item() - r.sum()) <= 1e-6 * np.abs(r).sum() + 1e-6
    assert rb.config()["n_tiles"] == sum((n + 4095) // 4096 for n in sizes)
    rb.close()


def test_graph_replays(dev):
    batch, capi, tapes, ti = _mods()
    sizes = batch.ragged_sizes(64, 1e3, 2e5, seed=1)
    a, b = make_items(dev, sizes, seed=1)
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    rb = batch.RaggedBatch(tape, [a, b, a, a, b], None, fuse=capi.FUSE_SUM | capi.FUSE_MAX)
    rb.launch(scalars=(0.5,))
    torch.cuda.synchronize()
    s0, m0 = rb.sums.clone(), rb.maxs.clone()
    rb.capture(scalars=(0.5,))
    for _ in range(3):
        rb.sums.zero_()
        rb.maxs.zero_()
        torch.cuda.synchronize()
        rb.replay()
        torch.cuda.synchronize()
        assert torch.equal(rb.sums, s0) and torch.equal(rb.maxs, m0)   # deterministic, every replay
    rb.close()


def test_autotune_picks_a_config(dev, tmp_path, monkeypatch):
    batch, capi, tapes, ti = _mods()
    monkeypatch.setenv("NCA_HIP_AUTOTUNE_DB", str(tmp_path / "db.txt"))
    sizes = batch.ragged_sizes(32, 1e3, 1e5, seed=2)
    a, b = make_items(dev, sizes, seed=2)
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    rb = batch.RaggedBatch(tape, [a, b, a, a, b], None, fuse=capi.FUSE_SUM | capi.FUSE_MAX)
    rb.launch(scalars=(0.5,))
    torch.cuda.synchronize()
    s0 = rb.sums.clone()
    ms = rb.autotune(scalars=(0.5,), reps=2)
    assert ms > 0
    cfg = rb.config()
    assert (tmp_path / "db.txt").read_text().strip()
    rb.launch(scalars=(0.5,))
    torch.cuda.synchronize()
    assert torch.allclose(rb.sums, s0, rtol=1e-6, atol=1e-5)
    # a new plan for the same workload picks the recorded configuration
    rb2 = batch.RaggedBatch(tape, [a, b, a, a, b], None, fuse=capi.FUSE_SUM | capi.FUSE_MAX)
    c2 = rb2.config()
    assert (c2["tile"], c2["vec"], c2["waves_per_cu"]) == (cfg["tile"], cfg["vec"], cfg["waves_per_cu"])
    rb.close()
    rb2.close()
