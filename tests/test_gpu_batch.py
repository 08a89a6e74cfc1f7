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


# ---- parity of the ragged batch with the oracle (the C0 expression per item, per-item sum, max
# and min); the reference path is one GPUEngine::execute_binary_expression + execute_full_reduce
# per array (gpuengine.hh:90-195, 301-333; kernels.cuh:565-589)

def _c0_oracle(a, b):
    """The C0 expression on host copies by the oracle, and the bound on |device - oracle| per
    element (only exp may differ, by 2 ulp of its term)."""
    from oracle import oracle as O
    from ncarray_amd import tapes
    ref = O.expr(tapes.fused_c0(), [a, b, a, a, b], [0.5], 10, np.float32)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    bound = 2 * float(np.finfo(np.float32).eps) * (np.abs(a64 * b64) + np.exp(-a64 * a64) + 0.5 * np.abs(b64))
    return ref, bound


def _check_item(out, s, mx, mn, a, b, mode, fuse):
    """out: the item's stored output (host float32); s/mx/mn: its fused sum/max/min."""
    from oracle import oracle as O
    ref, bound = _c0_oracle(a, b)
    o64 = out.astype(np.float64)
    assert np.all(np.abs(o64 - ref.astype(np.float64)) <= bound)
    # sums: the device's reduction of its stored values vs FP64 (c eps sum|x|), and vs the oracle's
    s64 = o64.sum()
    tol = (8e-7 if mode == 0 else 1.2e-7) * np.abs(o64).sum() + 1e-30
    if fuse & 1:
        assert abs(float(s) - s64) <= tol, (float(s), s64, tol)
        assert abs(float(s) - ref.astype(np.float64).sum()) <= tol + bound.sum()
    # max/min: bit-exact against the oracle's fold of the stored values
    if fuse & 2:
        assert np.float32(mx).view(np.uint32) == np.float32(O.max_fold(out)[0]).view(np.uint32)
    if fuse & 4:
        assert np.float32(mn).view(np.uint32) == np.float32(O.min_fold(out)[0]).view(np.uint32)


@pytest.mark.parametrize("mode", [0, 1], ids=["plain", "compensated"])
@pytest.mark.parametrize("aligned", [True, False], ids=["aligned", "unaligned"])
def test_batch_vs_oracle(dev, mode, aligned):
    """Ragged items (sizes not multiples of the vector width, single elements) packed in one
    buffer, every item start on a 16-byte boundary or deliberately not (the scalar path)."""
    batch, capi, tapes, ti = _mods()
    sizes = [1, 3, 5, 4097, 123457, 7, 64, 65, 1023, 300001, 2, 16384]
    offs, off = [], 1 if not aligned else 0
    for n in sizes:
        offs.append(off)
        off += n + (0 if not aligned else (-n) % 4) + (1 if not aligned else 0)
    g = torch.Generator(device=dev).manual_seed(7)
    a_all = torch.rand(off, generator=g, device=dev) * 2 - 1
    b_all = torch.rand(off, generator=g, device=dev) * 2 - 1
    o_all = torch.full((off,), float("nan"), device=dev)
    a = [a_all[p:p + n] for p, n in zip(offs, sizes)]
    b = [b_all[p:p + n] for p, n in zip(offs, sizes)]
    o = [o_all[p:p + n] for p, n in zip(offs, sizes)]
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    fuse = capi.FUSE_SUM | capi.FUSE_MAX | capi.FUSE_MIN
    rb = batch.RaggedBatch(tape, [a, b, a, a, b], o, fuse=fuse, mode=mode, tile=4096, vec=4, waves_per_cu=8)
    assert rb.config()["vec"] == (4 if aligned else 1)
    rb.launch(scalars=(0.5,))
    torch.cuda.synchronize()
    an, bn, on = a_all.cpu().numpy(), b_all.cpu().numpy(), o_all.cpu().numpy()
    sums, maxs, mins = rb.sums.cpu().numpy(), rb.maxs.cpu().numpy(), rb.mins.cpu().numpy()
    for i, (p, n) in enumerate(zip(offs, sizes)):
        _check_item(on[p:p + n], sums[i], maxs[i], mins[i], an[p:p + n], bn[p:p + n], mode, fuse)
    # nothing outside the items was written
    mask = np.ones(off, dtype=bool)
    for p, n in zip(offs, sizes):
        mask[p:p + n] = False
    assert np.all(np.isnan(on[mask]))
    # not materialised: the same reductions
    rb2 = batch.RaggedBatch(tape, [a, b, a, a, b], None, fuse=fuse, mode=mode, tile=4096, vec=4, waves_per_cu=8)
    rb2.launch(scalars=(0.5,))
    torch.cuda.synchronize()
    assert torch.equal(rb2.sums, rb.sums) and torch.equal(rb2.maxs, rb.maxs) and torch.equal(rb2.mins, rb.mins)
    rb.close()
    rb2.close()


def test_c4_full_size(dev):
    """The ragged workload at its stated size: 4096 arrays, sizes log-uniform in [1e3, 4e6]
    (seed 0; 1 927 863 588 elements, 7.7 GB per operand), the C0 expression with per-item sum
    and max in one launch.  Size-independent properties: the per-array launch loop (the
    reference's composition) gives bit-equal outputs and max and sums within the FP32 bound;
    a seeded sample of items, the largest and the smallest, checked against the oracle."""
    batch, capi, tapes, ti = _mods()
    sizes = batch.ragged_sizes(4096, 1e3, 4e6, seed=0)
    total = int(sizes.sum())
    assert total == 1927863588
    offs, off = [], 0
    for n in sizes:
        offs.append(off)
        off = -(-(off + int(n)) // 4) * 4
    g = torch.Generator(device=dev).manual_seed(0)
    a_all = torch.rand(off, generator=g, device=dev) * 2 - 1
    b_all = torch.rand(off, generator=g, device=dev) * 2 - 1
    o_all = torch.zeros(off, device=dev)   # the gaps between items stay zero in both outputs
    a = [a_all[p:p + int(n)] for p, n in zip(offs, sizes)]
    b = [b_all[p:p + int(n)] for p, n in zip(offs, sizes)]
    o = [o_all[p:p + int(n)] for p, n in zip(offs, sizes)]
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    fuse = capi.FUSE_SUM | capi.FUSE_MAX
    rb = batch.RaggedBatch(tape, [a, b, a, a, b], o, fuse=fuse)
    rb.launch(scalars=(0.5,))
    torch.cuda.synchronize()
    print(f"c4 batch done: {len(sizes)} items, {total} elements", flush=True)
    # the per-array loop into a second output buffer
    o2_all = torch.zeros(off, device=dev)
    sums = torch.empty(len(sizes), device=dev)
    maxs = torch.empty(len(sizes), device=dev)
    sp, mp = sums.data_ptr(), maxs.data_ptr()
    for i, (p, n) in enumerate(zip(offs, sizes)):
        ins = [ti.view(a[i]), ti.view(b[i]), ti.view(a[i]), ti.view(a[i]), ti.view(b[i])]
        tape.launch_reduce(ins, ti.view(o2_all[p:p + int(n)]), fuse, capi.ACCUM_PLAIN, sp + 4 * i, mp + 4 * i, None,
                           scalars=(0.5,), stream=capi.STREAM_DEFAULT_ASYNC)
    torch.cuda.synchronize()
    assert torch.equal(o_all, o2_all)
    del o2_all
    assert torch.equal(maxs, rb.maxs)
    # both sums within the FP32 bound of the FP64 sum of the stored values
    absum = torch.zeros(len(sizes), dtype=torch.float64, device=dev)
    s64 = torch.zeros(len(sizes), dtype=torch.float64, device=dev)
    for i, (p, n) in enumerate(zip(offs, sizes)):
        seg = o_all[p:p + int(n)].double()
        s64[i] = seg.sum()
        absum[i] = seg.abs().sum()
    tol = 8e-7 * absum + 1e-30
    assert bool(((rb.sums.double() - s64).abs() <= tol).all())
    assert bool(((sums.double() - s64).abs() <= tol).all())
    print("c4 loop and FP64 checks done", flush=True)
    # a seeded sample of items against the oracle
    rng = np.random.default_rng(0)
    pick = sorted(set(rng.choice(len(sizes), 6, replace=False).tolist() + [int(np.argmax(sizes)),
                                                                          int(np.argmin(sizes))]))
    rs, rm = rb.sums.cpu().numpy(), rb.maxs.cpu().numpy()
    for i in pick:
        p, n = offs[i], int(sizes[i])
        _check_item(o_all[p:p + n].cpu().numpy(), rs[i], rm[i], None, a_all[p:p + n].cpu().numpy(),
                    b_all[p:p + n].cpu().numpy(), 0, fuse)
    rb.close()
    del a_all, b_all, o_all
    torch.cuda.empty_cache()
