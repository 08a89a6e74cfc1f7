"""Fused elementwise expressions on the GPU through the C ABI, checked against the CPU
oracle (the reference VM semantics, oracle/nca_cpu.c) and FP64."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _capi():
    from ncarray_amd import capi, tapes, torch_interop as ti
    return capi, tapes, ti


def _c0_inputs(n, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, n).astype(np.float32)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    return a, b


@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 1023, 4097, 1 << 22])
def test_c0_expression_vs_oracle(dev, n):
    capi, tapes, ti = _capi()
    a, b = _c0_inputs(n, seed=n)
    ta, tb = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    ops = {"a": ti.view(ta), "b": ti.view(tb)}
    tape.launch([ops[k] for k in tapes.C0_OPERANDS], ti.view(out), scalars=tapes.C0_SCALARS)
    got = out.cpu().numpy()
    ref = O.expr(tapes.fused_c0(), [a, b, a, a, b], [0.5], 10, np.float32)
    # exp is the only transcendental: the device expf and the host's glibc expf may differ
    # by an ulp of exp(-a*a) in [e^-1, 1]; after the final additions that is an absolute
    # difference of at most 2 ulp of the largest term (outputs near zero can therefore
    # differ by many of their own ulps).  Everything else is evaluated in the same order.
    if n:
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        scale = np.abs(a64 * b64) + np.exp(-a64 * a64) + 0.5 * np.abs(b64)
        eps = float(np.finfo(np.float32).eps)
        assert np.all(np.abs(got.astype(np.float64) - ref) <= 2 * eps * scale)
        if n >= 1000:
            assert np.mean(got == ref) > 0.9   # most elements bit-identical to the oracle
        f64 = a64 * b64 + np.exp(-a64 * a64) - 0.5 * b64
        assert np.all(np.abs(got - f64) <= 3 * eps * scale)


def test_arithmetic_bit_exact(dev):
    """No transcendental: every element bit-identical to the oracle (same op order)."""
    capi, tapes, ti = _capi()
    P = capi.pack
    rng = np.random.default_rng(1)
    n = 100003
    x = rng.standard_normal(n).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    # (x*y + 3) - x/y * 2   (depth 3: VM order in the reference)
    ins = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_MUL), P(capi.OP_LOAD_CONST, 0),
           P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_LOAD_NCARR, 3), P(capi.OP_DIV),
           P(capi.OP_LOAD_CONST, 1), P(capi.OP_MUL), P(capi.OP_SUB)]
    tape = capi.Tape(ins, [capi.FLOAT32] * 4, [capi.FLOAT32, capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    tx, ty = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    tape.launch([ti.view(tx), ti.view(ty), ti.view(tx), ti.view(ty)], ti.view(out), scalars=(3.0, 2.0))
    ref = O.expr(ins, [x, y, x, y], [3.0, 2.0], 10, np.float32)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_reference_kats(dev):
    """test_binary_operations.cc / test_gpu_binary_operations.cu known answers."""
    capi, tapes, ti = _capi()
    P = capi.pack
    add = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_ADD)]
    # int32 2 + 4 = 6
    a = torch.full((4, 4), 2, dtype=torch.int32, device=dev)
    b = torch.full((4, 4), 4, dtype=torch.int32, device=dev)
    o = torch.empty((4, 4), dtype=torch.int32, device=dev)
    capi.Tape(add, [capi.INT32] * 2, [], capi.INT32, capi.INT32).launch([ti.view(a), ti.view(b)], ti.view(o))
    assert (o == 6).all()
    # float32 / float32 promotes to float64 (op_traits truediv_type); 10 / 4 = 2.5
    div = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_DIV)]
    assert capi.op_result_dtype(capi.OP_DIV, capi.FLOAT32) == capi.FLOAT64
    x = torch.full((8,), 10.0, dtype=torch.float32, device=dev)
    y = torch.full((8,), 4.0, dtype=torch.float32, device=dev)
    o64 = torch.empty(8, dtype=torch.float64, device=dev)
    capi.Tape(div, [capi.FLOAT32] * 2, [], capi.FLOAT32, capi.FLOAT64).launch([ti.view(x), ti.view(y)], ti.view(o64))
    assert (o64 == 2.5).all()
    # scalar broadcast 10 + 5
    adds = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_CONST, 0), P(capi.OP_ADD)]
    o2 = torch.empty(8, dtype=torch.float32, device=dev)
    capi.Tape(adds, [capi.FLOAT32], [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32).launch(
        [ti.view(x)], ti.view(o2), scalars=(5.0,))
    assert (o2 == 15.0).all()
    # comparisons -> bool (test_gpu_binary_operations.cu:189-217)
    gt = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_CONST, 0), P(capi.OP_GT)]
    v = torch.tensor([0, 0, 0, 0, 0, 10, 10, 10, 10, 10], dtype=torch.float32, device=dev)
    ob = torch.empty(10, dtype=torch.bool, device=dev)
    capi.Tape(gt, [capi.FLOAT32], [capi.FLOAT32], capi.FLOAT32, capi.BOOL).launch([ti.view(v)], ti.view(ob), scalars=(5.0,))
    assert ob.cpu().tolist() == [False] * 5 + [True] * 5
    # uint32 modulo (test_binary_operations.cc:216-230)
    mod = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MOD)]
    u = torch.arange(0, 20, dtype=torch.int64, device=dev).to(torch.int32)
    om = torch.empty(20, dtype=torch.int32, device=dev)
    capi.Tape(mod, [capi.INT32], [capi.INT32], capi.INT32, capi.INT32).launch([ti.view(u)], ti.view(om), scalars=(7,))
    assert om.cpu().tolist() == [i % 7 for i in range(20)]
    # neg / inc / dec (test_unary_operations.cc:15-37)
    for op, f in ((capi.OP_NEG, lambda q: -q), (capi.OP_INC, lambda q: q + 1), (capi.OP_DEC, lambda q: q - 1)):
        o3 = torch.empty(8, dtype=torch.float32, device=dev)
        capi.Tape([P(capi.OP_LOAD_NCARR, 0), P(op)], [capi.FLOAT32], [], capi.FLOAT32, capi.FLOAT32).launch(
            [ti.view(x)], ti.view(o3))
        assert (o3.cpu() == f(x.cpu())).all()


def test_iota(dev):
    """iota()*2+1 (test_unary_operations.cc:39-74): IDX is the C-order linear index."""
    capi, tapes, ti = _capi()
    P = capi.pack
    ins = [P(capi.OP_IDX), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MUL), P(capi.OP_LOAD_CONST, 1), P(capi.OP_ADD)]
    o = torch.empty((3, 5, 7), dtype=torch.int64, device=dev)
    capi.Tape(ins, [], [capi.INT64, capi.INT64], capi.INT64, capi.INT64).launch([], ti.view(o), scalars=(2, 1))
    assert (o.cpu().ravel() == torch.arange(105) * 2 + 1).all()


@pytest.mark.parametrize("n", [1, 3, 4, 255, 2051, 8 * 64 * 4 * 3 + 5, (1 << 22) + 7])
def test_generators_and_fills_at_chunk_edges(dev, n):
    """Expressions that read no array (an index generator, a constant fill) take the store-only
    form (each wave writes chunks of 8 x 64 vectors, the remainder element by element): every
    element written once and exactly, in float32, int64 and BF16 outputs."""
    capi, tapes, ti = _capi()
    P = capi.pack
    gen = [P(capi.OP_IDX), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MUL), P(capi.OP_LOAD_CONST, 1), P(capi.OP_ADD)]
    o = torch.full((n + 64,), -7.0, dtype=torch.float32, device=dev)
    capi.Tape(gen, [], [capi.FLOAT32, capi.FLOAT32], capi.FLOAT32, capi.FLOAT32).launch([], ti.view(o[:n]), scalars=(0.5, 3.0))
    oi = torch.full((n + 64,), -7, dtype=torch.int64, device=dev)
    capi.Tape(gen, [], [capi.INT64, capi.INT64], capi.INT64, capi.INT64).launch([], ti.view(oi[:n]), scalars=(3, -2))
    ob = torch.full((n + 64,), -7.0, dtype=torch.bfloat16, device=dev)
    capi.Tape([P(capi.OP_LOAD_CONST, 0)], [], [capi.FLOAT32], capi.FLOAT32, capi.BFLOAT16).launch(
        [], ti.view(ob[:n]), scalars=(1.25,))
    torch.cuda.synchronize()
    idx = torch.arange(n, dtype=torch.float64)
    assert torch.equal(o[:n].cpu(), (idx * 0.5 + 3.0).float())
    assert (o[n:].cpu() == -7.0).all()
    assert torch.equal(oi[:n].cpu(), torch.arange(n, dtype=torch.int64) * 3 - 2) and (oi[n:].cpu() == -7).all()
    assert (ob[:n].cpu().float() == 1.25).all() and (ob[n:].cpu().float() == -7.0).all()


def test_div_by_zero_semantics(dev):
    """interface.hh:150-156: finite/0 -> NaN, inf/0 -> inf (kept), nan/0 -> nan."""
    capi, tapes, ti = _capi()
    P = capi.pack
    div = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_DIV)]
    inf = float("inf")
    x = torch.tensor([1.0, -2.0, 0.0, inf, -inf, 6.0, -inf, inf, -0.0], dtype=torch.float32, device=dev)
    y = torch.tensor([0.0, 0.0, 0.0, 0.0, 0.0, 3.0, -0.0, -0.0, -0.0], dtype=torch.float32, device=dev)
    o = torch.empty(9, dtype=torch.float32, device=dev)
    # evaluate in float32 (destination float32, linearizable: work dtype float32)
    capi.Tape(div, [capi.FLOAT32] * 2, [], capi.FLOAT32, capi.FLOAT32).launch([ti.view(x), ti.view(y)], ti.view(o))
    r = o.cpu().numpy()
    assert np.isnan(r[0]) and np.isnan(r[1]) and np.isnan(r[2]) and np.isnan(r[8])
    assert r[3] == np.inf and r[4] == -np.inf and r[5] == 2.0
    assert r[6] == -np.inf and r[7] == np.inf       # an infinite numerator is kept, sign included
    ref = O.expr(div, [x.cpu().numpy(), y.cpu().numpy()], [], 10, np.float32)
    assert np.array_equal(r, ref, equal_nan=True)


def test_strided_and_broadcast(dev):
    capi, tapes, ti = _capi()
    P = capi.pack
    rng = np.random.default_rng(3)
    base = torch.from_numpy(rng.standard_normal((37, 41)).astype(np.float32)).to(dev)
    xt = base.t()                      # transposed (non-contiguous)
    ys = base[:, ::2][:, :20].t()      # strided slice, shape (20, 37)
    xt = xt[:20]                       # (20, 37)
    s = torch.tensor(2.5, dtype=torch.float32, device=dev)  # 0-d broadcast array
    ins = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_MUL), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_ADD)]
    out = torch.empty((20, 37), dtype=torch.float32, device=dev)
    capi.Tape(ins, [capi.FLOAT32] * 3, [], capi.FLOAT32, capi.FLOAT32).launch(
        [ti.view(xt), ti.view(ys), ti.view(s)], ti.view(out))
    ref = (xt.cpu().numpy() * ys.cpu().numpy()) + np.float32(2.5)
    assert np.array_equal(out.cpu().numpy(), ref)
    # strided output
    big = torch.zeros((20, 80), dtype=torch.float32, device=dev)
    capi.Tape(ins, [capi.FLOAT32] * 3, [], capi.FLOAT32, capi.FLOAT32).launch(
        [ti.view(xt), ti.view(ys), ti.view(s)], ti.view(big[:, ::2][:, :37]))
    assert np.array_equal(big[:, ::2][:, :37].cpu().numpy(), ref)
    assert (big[:, 1::2] == 0).all()


def test_fused_reduce(dev):
    capi, tapes, ti = _capi()
    n = (1 << 20) + 7
    a, b = _c0_inputs(n, seed=11)
    ta, tb = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    ops = {"a": ti.view(ta), "b": ti.view(tb)}
    for mode in (capi.ACCUM_PLAIN, capi.ACCUM_COMPENSATED):
        out = torch.empty(n, dtype=torch.float32, device=dev)
        s = torch.empty(1, dtype=torch.float32, device=dev)
        mx = torch.empty(1, dtype=torch.float32, device=dev)
        mn = torch.empty(1, dtype=torch.float32, device=dev)
        tape.launch_reduce([ops[k] for k in tapes.C0_OPERANDS], ti.view(out), capi.FUSE_SUM | capi.FUSE_MAX | capi.FUSE_MIN,
                           mode, s.data_ptr(), mx.data_ptr(), mn.data_ptr(), scalars=(0.5,))
        o = out.cpu().numpy()
        assert mx.item() == o.max() and mn.item() == o.min()
        s64 = o.astype(np.float64).sum()
        # |S - S64| <= 1e-6 * sum|x| (plain float32 running sums) + rounding of the result
        assert abs(s.item() - s64) <= 1e-6 * np.abs(o).sum() + 2 * np.spacing(np.float32(abs(s64)))
        # not materialised: same reductions
        s2 = torch.empty(1, dtype=torch.float32, device=dev)
        mx2 = torch.empty(1, dtype=torch.float32, device=dev)
        tape.launch_reduce([ops[k] for k in tapes.C0_OPERANDS], ti.shape_view([n], capi.FLOAT32),
                           capi.FUSE_SUM | capi.FUSE_MAX, mode, s2.data_ptr(), mx2.data_ptr(), None, scalars=(0.5,))
        assert mx2.item() == mx.item() and s2.item() == s.item()


def test_bf16_storage(dev):
    """BF16 storage, float32 arithmetic; oracle: FP64 of the bf16-rounded inputs."""
    capi, tapes, ti = _capi()
    n = 65536 + 3
    a, b = _c0_inputs(n, seed=5)
    ta = torch.from_numpy(a).to(dev).to(torch.bfloat16)
    tb = torch.from_numpy(b).to(dev).to(torch.bfloat16)
    tape = capi.Tape(tapes.fused_c0(), [capi.BFLOAT16] * 5, [capi.FLOAT32], capi.BFLOAT16, capi.FLOAT32)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    ops = {"a": ti.view(ta), "b": ti.view(tb)}
    tape.launch([ops[k] for k in tapes.C0_OPERANDS], ti.view(out), scalars=(0.5,))
    a64 = ta.float().cpu().numpy().astype(np.float64)
    b64 = tb.float().cpu().numpy().astype(np.float64)
    f64 = a64 * b64 + np.exp(-a64 * a64) - 0.5 * b64
    assert np.max(np.abs(out.cpu().numpy() - f64)) <= 4 * np.finfo(np.float32).eps * np.max(np.abs(f64))
    # bf16 destination: the stored value is the float32 result rounded to bf16
    out16 = torch.empty(n, dtype=torch.bfloat16, device=dev)
    tape2 = capi.Tape(tapes.fused_c0(), [capi.BFLOAT16] * 5, [capi.FLOAT32], capi.BFLOAT16, capi.BFLOAT16)
    tape2.launch([ops[k] for k in tapes.C0_OPERANDS], ti.view(out16), scalars=(0.5,))
    assert torch.equal(out16.cpu(), out.cpu().to(torch.bfloat16))


def test_vjp_c0(dev):
    """d/da = b - 2a exp(-a^2), d/db = a - 0.5, times the upstream gradient."""
    capi, tapes, ti = _capi()
    n = 4099
    a, b = _c0_inputs(n, seed=9)
    ta, tb = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
    g = torch.from_numpy(np.random.default_rng(2).standard_normal(n).astype(np.float32)).to(dev)
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    ga = torch.empty(n, dtype=torch.float32, device=dev)
    gb = torch.empty(n, dtype=torch.float32, device=dev)
    ops = {"a": ti.view(ta), "b": ti.view(tb)}
    ins = [ops[k] for k in tapes.C0_OPERANDS]
    grads = [ti.view(ga), ti.view(gb), ti.view(ga), ti.view(ga), ti.view(gb)]
    tape.vjp(ins, grads, [1, 1, 1, 1, 1], grad_out=ti.view(g), scalars=(0.5,))
    a64, b64, g64 = a.astype(np.float64), b.astype(np.float64), g.cpu().numpy().astype(np.float64)
    ref_a = g64 * (b64 - 2 * a64 * np.exp(-a64 * a64))
    ref_b = g64 * (a64 - 0.5)
    assert np.max(np.abs(ga.cpu().numpy() - ref_a)) < 1e-5
    assert np.max(np.abs(gb.cpu().numpy() - ref_b)) < 1e-5
    # scalar upstream gradient (backward of sum)
    one = torch.ones(1, dtype=torch.float32, device=dev)
    tape.vjp(ins, grads, [1, 1, 1, 1, 1], grad_scalar=one.data_ptr(), scalars=(0.5,))
    assert np.max(np.abs(gb.cpu().numpy() - (a64 - 0.5))) < 1e-6


def test_bf16_store_rounding_matches_torch(dev):
    """float32 -> bfloat16 stores round to nearest even: ties both ways, carries into the
    exponent, subnormals, signed zeros, infinities and overflow, compared bit for bit with
    torch's conversion (NaN only has to stay NaN)."""
    import torch
    from ncarray_amd import capi
    from ncarray_amd import torch_interop as ti
    bits = [0x3F808000, 0x3F818000, 0x3F80FFFF, 0x3F807FFF, 0x3FFF8000, 0x7F7FFFFF, 0x7F7F8000, 0x00008000,
            0x00018000, 0x00000001, 0x807FFFFF, 0x80000000, 0x00000000, 0x7F800000, 0xFF800000, 0x477FE000]
    rng = np.random.default_rng(12)
    rnd = rng.integers(0, 2 ** 32, 4096, dtype=np.uint64).astype(np.uint32)
    rnd = rnd[(rnd & 0x7F800000) != 0x7F800000]  # finite values
    x = np.concatenate([np.array(bits, dtype=np.uint32), rnd]).view(np.float32)
    xd = torch.from_numpy(x.copy()).cuda()
    out = torch.empty(x.shape, dtype=torch.bfloat16, device="cuda")
    tape = capi.Tape([capi.pack(capi.OP_LOAD_NCARR, 0)], [capi.FLOAT32], [], capi.FLOAT32, capi.BFLOAT16)
    tape.launch([ti.view(xd)], ti.view(out))
    want = torch.from_numpy(x.copy()).to(torch.bfloat16)
    assert torch.equal(out.cpu().view(torch.int16), want.view(torch.int16))
    nan = torch.tensor([float("nan")], dtype=torch.float32, device="cuda")
    on = torch.empty(1, dtype=torch.bfloat16, device="cuda")
    tape.launch([ti.view(nan)], ti.view(on))
    assert torch.isnan(on.float()).all()
