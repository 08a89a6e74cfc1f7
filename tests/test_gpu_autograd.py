"""Hand-written adjoints behind torch.autograd.Function: checked with gradcheck (float64),
central finite differences of the FP64 loss (h = 1e-3) and an FP64 reference gradient."""
import ctypes
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _mods():
    from ncarray_amd import autograd as A, capi, tapes
    return A, capi, tapes


def lap64(a):
    p = np.pad(np.asarray(a, dtype=np.float64), 1)
    return (p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1]
            + p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:] - 6.0 * p[1:-1, 1:-1, 1:-1])


def loss64(x, y):
    a = lap64(x)
    b = np.asarray(y, dtype=np.float64)
    return float(np.sum(a * b + np.exp(-a * a) - 0.5 * b))


def grad64(x, y):
    a = lap64(x)
    b = np.asarray(y, dtype=np.float64)
    return lap64(b - 2 * a * np.exp(-a * a)), a - 0.5   # the Laplacian is self-adjoint (zero Dirichlet)


def test_expr_gradcheck(dev):
    A, capi, tapes = _mods()
    tape = A.c0_tape(capi.FLOAT64, capi.FLOAT64, capi.FLOAT64)
    a = torch.randn(40, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.randn(40, dtype=torch.float64, device=dev, requires_grad=True)
    f = lambda a_, b_: A.expr(tape, (0.5,), a_, b_, a_, a_, b_)  # noqa: E731
    assert torch.autograd.gradcheck(f, (a, b), eps=1e-6, atol=1e-7)


def test_stencil_and_sum_gradcheck(dev):
    A, capi, tapes = _mods()
    x = torch.randn((4, 5, 8), dtype=torch.float64, device=dev, requires_grad=True)
    P = capi.pack
    sq = capi.Tape([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_MUL)], [capi.FLOAT64] * 2, [],
                   capi.FLOAT64, capi.FLOAT64)

    def f(t):
        a = A.laplacian(t)
        return A.sum(A.expr(sq, (), a, a))

    assert torch.autograd.gradcheck(f, (x,), eps=1e-6, atol=1e-6)
    assert torch.autograd.gradcheck(lambda t: A.mean(A.laplacian(t)), (x,), eps=1e-6, atol=1e-7)


def test_bin_weighted_sum_backward(dev):
    A, capi, tapes = _mods()
    from oracle import oracle as O
    bm = torch.from_numpy(O.radial_binmap(16, 24, 8, 12, 20)).to(dev)
    fr = torch.rand((3, 16, 24), device=dev, requires_grad=True)
    s = A.bin_weighted_sum(fr, bm, 20)
    g = torch.randn_like(s)
    (s * g).sum().backward()
    ref = g[:, bm.long().reshape(-1)].reshape(3, 16, 24)
    assert torch.equal(fr.grad, ref)


@pytest.mark.parametrize("n", [24, 64])
def test_fused_loss_vs_fp64_and_fd(dev, n):
    A, capi, tapes = _mods()
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n, n)).astype(np.float32)
    y = rng.standard_normal((n, n, n)).astype(np.float32)
    xt = torch.from_numpy(x).to(dev).requires_grad_()
    yt = torch.from_numpy(y).to(dev).requires_grad_()
    loss = A.laplacian_c0_loss(xt, yt)
    L64 = loss64(x, y)
    a = lap64(x)
    mag = float(np.sum(np.abs(a * y) + np.exp(-a * a) + 0.5 * np.abs(y)))
    assert abs(loss.item() - L64) <= 1e-6 * mag
    loss.backward()
    gx64, gy64 = grad64(x, y)
    gx, gy = xt.grad.cpu().numpy(), yt.grad.cpu().numpy()
    assert np.max(np.abs(gx - gx64)) <= 1e-4 * max(1.0, np.max(np.abs(gx64)))
    assert np.max(np.abs(gy - gy64)) <= 1e-5 * max(1.0, np.max(np.abs(gy64)))
    # central finite differences of the FP64 loss along random directions, h = 1e-3
    h = 1e-3
    for k in range(2):
        v = rng.standard_normal(x.shape)
        w = rng.standard_normal(y.shape)
        fd = (loss64(x + h * v, y + h * w) - loss64(x - h * v, y - h * w)) / (2 * h)
        an = float(np.sum(gx.astype(np.float64) * v) + np.sum(gy.astype(np.float64) * w))
        assert abs(fd - an) <= 1e-4 * (abs(fd) + 1.0) + 1e-3 * np.sqrt(x.size), (fd, an)


def test_fused_loss_256(dev):
    """The differentiable pipeline at 256^3 (fp32): gradients vs the FP64 reference."""
    A, capi, tapes = _mods()
    g = torch.Generator(device=dev).manual_seed(5)
    xt = torch.randn((256, 256, 256), generator=g, device=dev).requires_grad_()
    yt = torch.randn((256, 256, 256), generator=g, device=dev).requires_grad_()
    loss = A.laplacian_c0_loss(xt, yt)
    loss.backward()
    x, y = xt.detach().cpu().numpy(), yt.detach().cpu().numpy()
    gx64, gy64 = grad64(x, y)
    assert np.max(np.abs(xt.grad.cpu().numpy() - gx64)) <= 1e-4 * np.max(np.abs(gx64))
    assert np.max(np.abs(yt.grad.cpu().numpy() - gy64)) <= 1e-5 * np.max(np.abs(gy64))
    L64 = loss64(x, y)
    assert abs(loss.item() - L64) <= 1e-6 * abs(L64) + 1.0


def test_fused_loss_bf16(dev):
    """BF16 storage: gradients vs FP64 computed from the bf16-rounded inputs."""
    A, capi, tapes = _mods()
    n = 64
    rng = np.random.default_rng(9)
    xt = torch.from_numpy(rng.standard_normal((n, n, n)).astype(np.float32)).to(dev).to(torch.bfloat16).requires_grad_()
    yt = torch.from_numpy(rng.standard_normal((n, n, n)).astype(np.float32)).to(dev).to(torch.bfloat16).requires_grad_()
    loss = A.laplacian_c0_loss(xt, yt)
    loss.backward()
    x, y = xt.detach().float().cpu().numpy(), yt.detach().float().cpu().numpy()
    gx64, gy64 = grad64(x, y)
    assert xt.grad.dtype == torch.bfloat16
    # bf16 gradients: one rounding to bf16 (rel. 2^-8) of float32 results
    assert np.max(np.abs(xt.grad.float().cpu().numpy() - gx64)) <= 8e-3 * np.max(np.abs(gx64))
    assert np.max(np.abs(yt.grad.float().cpu().numpy() - gy64)) <= 8e-3 * np.max(np.abs(gy64))
    L64 = loss64(x, y)
    assert abs(loss.item() - L64) <= 1e-5 * abs(L64) + 1.0


def test_unfused_composition_matches_fused(dev):
    A, capi, tapes = _mods()
    rng = np.random.default_rng(2)
    x = torch.from_numpy(rng.standard_normal((32, 32, 32)).astype(np.float32)).to(dev)
    y = torch.from_numpy(rng.standard_normal((32, 32, 32)).astype(np.float32)).to(dev)
    x1, y1 = x.clone().requires_grad_(), y.clone().requires_grad_()
    x2, y2 = x.clone().requires_grad_(), y.clone().requires_grad_()
    A.laplacian_c0_loss(x1, y1).backward()
    a = A.laplacian(x2)
    tape = A.c0_tape(capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    A.sum(A.expr(tape, (0.5,), a, y2, a, a, y2)).backward()
    assert torch.allclose(x1.grad, x2.grad, rtol=1e-5, atol=1e-5)
    assert torch.allclose(y1.grad, y2.grad, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_stencil_expr_vjp_matches_unfused(dev, dtype):
    """nca_hip_stencil_expr_vjp (one march pass: Laplacian, derivative expression, two outputs)
    gives the same bits as the Laplacian kernel followed by the expression's vjp kernel, with an
    upstream gradient other than 1."""
    A, capi, tapes = _mods()
    from ncarray_amd import torch_interop as ti
    g = torch.Generator(device=dev).manual_seed(11)
    x = torch.randn((20, 24, 40), generator=g, device=dev).to(dtype)
    y = torch.randn((20, 24, 40), generator=g, device=dev).to(dtype)
    gs = torch.tensor(2.5, dtype=torch.float32, device=dev)
    xc, yc = ti.TORCH_TO_CODE[dtype], ti.TORCH_TO_CODE[dtype]
    lap, _ = A.laplacian_stencil(xc, capi.FLOAT32)
    tape = A.c0_tape(capi.FLOAT32, yc, capi.FLOAT32)
    w = torch.empty(x.shape, dtype=torch.float32, device=dev)
    gy = torch.empty_like(y)
    lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(w), extras=[ti.view(y)], grad_extras=[ti.view(gy)],
                 scalars=(6.0,), expr_scalars=(0.5,), grad_scalar=gs.data_ptr())
    a = torch.empty(x.shape, dtype=torch.float32, device=dev)
    lap.apply(ti.view(x), ti.view(a), capi.BOUNDARY_ZERO, scalars=(6.0,))
    ga, gy2 = torch.empty_like(a), torch.empty_like(y)
    va, vy = ti.view(a), ti.view(y)
    tape.vjp([va, vy, va, va, vy], [ti.view(ga), ti.view(gy2), ti.view(ga), ti.view(ga), ti.view(gy2)], [1] * 5,
             grad_scalar=gs.data_ptr(), scalars=(0.5,))
    torch.cuda.synchronize()
    assert torch.equal(w, ga)
    assert torch.equal(gy, gy2)
    # w alone (no extra gradient asked for), g = 1 when the pointer is null
    w1 = torch.empty_like(w)
    lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(w1), extras=[ti.view(y)], grad_extras=[None],
                 scalars=(6.0,), expr_scalars=(0.5,))
    torch.cuda.synchronize()
    assert torch.equal(w1 * 2.5, ga)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("mode", [0, 1])
def test_stencil_expr_vjp_reduce(dev, dtype, mode):
    """nca_hip_stencil_expr_vjp_reduce: the gradients are the bits of nca_hip_stencil_expr_vjp;
    the fused max and min of the expression's value equal those of the forward kernel's stored
    values (exact); the fused sum, like the forward kernel's fused sum, is within the
    accumulation's bound of the FP64 loss."""
    A, capi, tapes = _mods()
    from ncarray_amd import torch_interop as ti
    g = torch.Generator(device=dev).manual_seed(12)
    x = torch.randn((24, 20, 48), generator=g, device=dev).to(dtype)
    y = torch.randn((24, 20, 48), generator=g, device=dev).to(dtype)
    gs = torch.tensor(0.75, dtype=torch.float32, device=dev)
    xc = ti.TORCH_TO_CODE[dtype]
    lap, _ = A.laplacian_stencil(xc, capi.FLOAT32)
    tape = A.c0_tape(capi.FLOAT32, xc, capi.FLOAT32)
    args = dict(extras=[ti.view(y)], scalars=(6.0,), expr_scalars=(0.5,), grad_scalar=gs.data_ptr())
    w, gy = torch.empty(x.shape, dtype=torch.float32, device=dev), torch.empty_like(y)
    sm, mx, mn = (torch.full((), np.nan, dtype=torch.float32, device=dev) for _ in range(3))
    lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(w), grad_extras=[ti.view(gy)],
                 fuse=capi.FUSE_SUM | capi.FUSE_MAX | capi.FUSE_MIN, mode=mode, sum_out=sm.data_ptr(),
                 max_out=mx.data_ptr(), min_out=mn.data_ptr(), **args)
    w2, gy2 = torch.empty_like(w), torch.empty_like(gy)
    lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(w2), grad_extras=[ti.view(gy2)], **args)
    # the forward: the fused Laplacian-C0 stencil storing its values, and its own fused sum
    fused = A.fused_loss_stencil(xc, xc)
    val = torch.empty(x.shape, dtype=torch.float32, device=dev)
    fsum = torch.empty((), dtype=torch.float32, device=dev)
    fused.apply_reduce(ti.view(x), ti.view(val), capi.BOUNDARY_ZERO, fuse=capi.FUSE_SUM, mode=mode,
                       sum_out=fsum.data_ptr(), extras=[ti.view(y)], scalars=(6.0, 0.5))
    torch.cuda.synchronize()
    assert torch.equal(w, w2) and torch.equal(gy, gy2)
    assert mx.item() == val.max().item() and mn.item() == val.min().item()
    xn, yn = x.float().cpu().numpy(), y.float().cpu().numpy()
    ref = loss64(xn, yn)
    a = lap64(xn)
    mass = float(np.sum(np.abs(a * yn) + np.exp(-a * a) + 0.5 * np.abs(yn)))
    # per-element float32 evaluation (a few ulp) plus the accumulation: Kahan per thread or a
    # float32 running sum over at most NZ values per thread, float64 beyond
    bound = (8 if mode else 8 + x.shape[0]) * np.finfo(np.float32).eps * mass
    assert abs(sm.item() - ref) <= bound
    assert abs(fsum.item() - ref) <= bound


def test_stencil_expr_vjp_reduce_empty_and_rejects(dev):
    """An empty array yields the reductions' identities (a zero sum) from both fused stencil
    entry points; the fused vjp needs a reduction and its output pointers."""
    A, capi, tapes = _mods()
    from ncarray_amd import torch_interop as ti
    x = torch.empty((0, 8, 16), device=dev)
    lap, _ = A.laplacian_stencil(capi.FLOAT32, capi.FLOAT32)
    tape = A.c0_tape(capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    w = torch.empty_like(x)
    sm = torch.full((), np.nan, dtype=torch.float32, device=dev)
    args = dict(extras=[ti.view(x)], scalars=(6.0,), expr_scalars=(0.5,))
    lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(w), fuse=capi.FUSE_SUM, sum_out=sm.data_ptr(), **args)
    fsum = torch.full((), np.nan, dtype=torch.float32, device=dev)
    A.fused_loss_stencil(capi.FLOAT32, capi.FLOAT32).apply_reduce(
        ti.view(x), ti.shape_view(list(x.shape), capi.FLOAT32), capi.BOUNDARY_ZERO, fuse=capi.FUSE_SUM,
        sum_out=fsum.data_ptr(), extras=[ti.view(x)], scalars=(6.0, 0.5))
    torch.cuda.synchronize()
    assert sm.item() == 0.0 and fsum.item() == 0.0
    x = torch.randn((8, 8, 16), device=dev)
    w = torch.empty_like(x)
    args = dict(extras=[ti.view(x)], scalars=(6.0,), expr_scalars=(0.5,))
    with pytest.raises(capi.NcaError, match="output pointer"):
        lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(w), fuse=capi.FUSE_MAX, **args)
    with pytest.raises(capi.NcaError, match="no reduction"):
        capi.call("nca_hip_stencil_expr_vjp_reduce", lap.handle, tape.handle, (ctypes.c_int32 * 5)(-1, 0, -1, -1, 0), 5,
                  ctypes.byref(ti.view(x)), capi.view_array([ti.view(x)]), 1,
                  capi.scalar_array((6.0,), lap.scalar_dtypes), 1, capi.scalar_array((0.5,), tape.scalar_dtypes), 1,
                  None, ctypes.byref(ti.view(w)), capi.view_array([capi.View()]), 0, 0, None, None, None,
                  capi.BOUNDARY_ZERO, None)


def test_value_and_grad_ragged_shape(dev):
    """laplacian_c0_value_and_grad on a shape the vectorised march kernels do not take (last
    extent not a whole number of vectors): the one-pass training kernel agrees with FP64."""
    A, capi, tapes = _mods()
    rng = np.random.default_rng(6)
    x = rng.standard_normal((12, 14, 18)).astype(np.float32)
    y = rng.standard_normal((12, 14, 18)).astype(np.float32)
    xt, yt = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)
    assert not A._march_vjp_ok(xt, yt) and A._train_ok(xt, yt)
    loss, dx, dy = A.laplacian_c0_value_and_grad(xt, yt)
    torch.cuda.synchronize()
    gx64, gy64 = grad64(x, y)
    assert abs(loss.item() - loss64(x, y)) <= 1e-5 * abs(loss64(x, y)) + 1e-3
    assert np.allclose(dx.cpu().numpy(), gx64, rtol=1e-4, atol=1e-4)
    assert np.allclose(dy.cpu().numpy(), gy64, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("method", ["train", "compose"])
def test_fused_loss_backward_ragged_shape(dev, method):
    """A last extent that is not a whole number of vectors: the one-pass training backward (the
    default) and the three-kernel composition agree with the FP64 gradient."""
    A, capi, tapes = _mods()
    rng = np.random.default_rng(4)
    x = rng.standard_normal((16, 20, 18)).astype(np.float32)
    y = rng.standard_normal((16, 20, 18)).astype(np.float32)
    xt, yt = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)
    assert not A._march_vjp_ok(xt, yt)
    dx, dy = A.laplacian_c0_backward(xt, yt, method=method)
    torch.cuda.synchronize()
    gx64, gy64 = grad64(x, y)
    assert np.max(np.abs(dx.cpu().numpy() - gx64)) <= 1e-4 * max(1.0, np.max(np.abs(gx64)))
    assert np.max(np.abs(dy.cpu().numpy() - gy64)) <= 1e-5 * max(1.0, np.max(np.abs(gy64)))
    if method == "train":   # the autograd Function runs the same kernel
        xa, ya = xt.clone().requires_grad_(), yt.clone().requires_grad_()
        A.laplacian_c0_loss(xa, ya).backward()
        assert torch.equal(xa.grad, dx) and torch.equal(ya.grad, dy)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_stencil_expr_train_matches_two_pass(dev, dtype):
    """nca_hip_stencil_expr_train (one pass: w on tiles and rings in LDS, the transposed
    Laplacian from LDS) gives the bits of the two-kernel backward (w through memory) with an
    upstream gradient other than 1; its fused loss is within the accumulation's bound of FP64
    and equals the forward's fused sum to that bound."""
    A, capi, tapes = _mods()
    g = torch.Generator(device=dev).manual_seed(13)
    x = torch.randn((40, 36, 80), generator=g, device=dev).to(dtype)   # several z runs per tile column
    y = torch.randn((40, 36, 80), generator=g, device=dev).to(dtype)
    gs = torch.tensor(1.75, dtype=torch.float32, device=dev)
    dx1, dy1 = A.laplacian_c0_backward(x, y, gs, method="train")
    dx2, dy2 = A.laplacian_c0_backward(x, y, gs, method="vjp")
    torch.cuda.synchronize()
    assert torch.equal(dy1, dy2)
    assert torch.equal(dx1, dx2)
    loss, dx3, dy3 = A.laplacian_c0_value_and_grad(x, y)
    torch.cuda.synchronize()
    xn, yn = x.float().cpu().numpy(), y.float().cpu().numpy()
    a = lap64(xn)
    mass = float(np.sum(np.abs(a * yn) + np.exp(-a * a) + 0.5 * np.abs(yn)))
    assert abs(loss.item() - loss64(xn, yn)) <= 8 * np.finfo(np.float32).eps * mass
    gx64, gy64 = grad64(xn, yn)
    tol = 1e-4 if dtype == torch.float32 else 2e-2
    assert np.max(np.abs(dx3.float().cpu().numpy() - gx64)) <= tol * max(1.0, np.max(np.abs(gx64)))
    assert np.max(np.abs(dy3.float().cpu().numpy() - gy64)) <= tol * max(1.0, np.max(np.abs(gy64)))


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape", [(1, 5, 7), (3, 1, 130), (9, 17, 1), (2, 33, 65), (70, 3, 3), (5, 11, 68)])
def test_stencil_expr_train_edge_shapes(dev, shape, dtype):
    """Single planes, single rows and columns, partial tiles and short z runs (the last shape
    takes the vector form): the training pass equals the three-kernel composition (the same
    arithmetic per element; dy bit-equal, dx within the rounding of its last step)."""
    A, capi, tapes = _mods()
    rng = np.random.default_rng(sum(shape))
    xt = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(dev).to(dtype)
    yt = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(dev).to(dtype)
    x, y = xt.float().cpu().numpy(), yt.float().cpu().numpy()
    dx1, dy1 = A.laplacian_c0_backward(xt, yt, method="train")
    dx2, dy2 = A.laplacian_c0_backward(xt, yt, method="compose")
    torch.cuda.synchronize()
    assert torch.equal(dy1, dy2)
    unit = float(np.finfo(np.float32).eps) if dtype == torch.float32 else 2.0 ** -8
    assert torch.allclose(dx1.float(), dx2.float(), rtol=0, atol=4 * unit * max(1.0, float(dx2.float().abs().max())))
    gx64, _ = grad64(x, y)
    tol = 1e-4 if dtype == torch.float32 else 2e-2
    assert np.max(np.abs(dx1.float().cpu().numpy() - gx64)) <= tol * max(1.0, np.max(np.abs(gx64)))


def test_stencil_expr_train_rejects(dev):
    A, capi, tapes = _mods()
    from ncarray_amd import torch_interop as ti
    x = torch.randn((8, 8, 8), device=dev)
    y = torch.randn((8, 8, 8), device=dev)
    dx = torch.empty_like(x)
    tape = A.c0_tape(capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    args = dict(extras=[ti.view(y)], scalars=(6.0,), expr_scalars=(0.5,))
    wide = capi.Stencil([[0, 0, -2], [0, 0, 0]], [tapes.P(capi.OP_LOAD_NCARR, 0), tapes.P(capi.OP_LOAD_NCARR, 1),
                                                   tapes.P(capi.OP_ADD)], capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    with pytest.raises(capi.NcaError, match="radius"):
        wide.expr_train(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(dx), **args)
    square = capi.Stencil([[0, 0, 0]], [tapes.P(capi.OP_LOAD_NCARR, 0), tapes.P(capi.OP_LOAD_NCARR, 0),
                                        tapes.P(capi.OP_MUL)], capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    with pytest.raises(capi.NcaError, match="not linear"):
        square.expr_train(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(dx), **args)
    lap, _ = A.laplacian_stencil(capi.FLOAT32, capi.FLOAT32)
    with pytest.raises(capi.NcaError, match="zero boundary"):
        lap.expr_train(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(dx), boundary=capi.BOUNDARY_NONE, **args)
    with pytest.raises(capi.NcaError, match="output pointer"):
        lap.expr_train(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(dx), fuse=capi.FUSE_SUM, **args)


def test_stencil_expr_vjp_rejects(dev):
    A, capi, tapes = _mods()
    from ncarray_amd import torch_interop as ti
    x = torch.randn((8, 8, 8), device=dev)
    y = torch.randn((8, 8, 8), device=dev)
    w = torch.empty_like(x)
    lap, _ = A.laplacian_stencil(capi.FLOAT32, capi.FLOAT32)
    tape = A.c0_tape(capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    args = dict(extras=[ti.view(y)], scalars=(6.0,), expr_scalars=(0.5,))
    with pytest.raises(capi.NcaError, match="zero boundary"):
        lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(w), boundary=capi.BOUNDARY_NONE, **args)
    with pytest.raises(capi.NcaError, match="read the stencil value"):
        lap.expr_vjp(tape, [0, 0, 0, 0, 0], ti.view(x), ti.view(w), **args)
    with pytest.raises(capi.NcaError, match="operand code"):
        lap.expr_vjp(tape, [-1, 3, -1, -1, 0], ti.view(x), ti.view(w), **args)
    fused = A.fused_loss_stencil(capi.FLOAT32, capi.FLOAT32)  # has an extra array of its own
    with pytest.raises(capi.NcaError, match="reads only its source"):
        fused.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(w), **args)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape", [(20, 24, 64), (9, 13, 36), (5, 70, 132)])
def test_stencil_expr_train_vector_matches_scalar(dev, dtype, shape, monkeypatch):
    """The 4-point-vector form of the training pass (last extent a multiple of 4, aligned
    arrays; partial tiles in x and y) gives the bits of the scalar form for dx and dy, and a
    loss within the accumulation's bound of it."""
    A, capi, tapes = _mods()
    from ncarray_amd import torch_interop as ti
    g = torch.Generator(device=dev).manual_seed(sum(shape))
    x = torch.randn(shape, generator=g, device=dev).to(dtype)
    y = torch.randn(shape, generator=g, device=dev).to(dtype)
    gs = torch.tensor(0.625, dtype=torch.float32, device=dev)
    lap, _ = A.laplacian_stencil(ti.TORCH_TO_CODE[dtype], capi.FLOAT32)
    tape = A.c0_tape(capi.FLOAT32, ti.TORCH_TO_CODE[dtype], capi.FLOAT32)
    res = []
    for scalar in ("1", "0"):
        monkeypatch.setenv("NCA_HIP_TRAIN_SCALAR", scalar)
        dx, dy = torch.empty_like(x), torch.empty_like(y)
        loss = torch.empty((), dtype=torch.float32, device=dev)
        lap.expr_train(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(dx), extras=[ti.view(y)],
                       grad_extras=[ti.view(dy)], scalars=(6.0,), expr_scalars=(0.5,), grad_scalar=gs.data_ptr(),
                       fuse=capi.FUSE_SUM, mode=capi.ACCUM_COMPENSATED, sum_out=loss.data_ptr())
        res.append((dx, dy, loss))
    torch.cuda.synchronize()
    (dx1, dy1, l1), (dx2, dy2, l2) = res
    assert torch.equal(dx1, dx2) and torch.equal(dy1, dy2)
    xn, yn = x.float().cpu().numpy(), y.float().cpu().numpy()
    a = lap64(xn)
    mass = float(np.sum(np.abs(a * yn) + np.exp(-a * a) + 0.5 * np.abs(yn)))
    assert abs(l1.item() - l2.item()) <= 16 * np.finfo(np.float32).eps * mass
    assert abs(l2.item() - loss64(xn, yn)) <= 8 * np.finfo(np.float32).eps * mass


@pytest.mark.parametrize("tile", ["32x4", "64x16", "128x8", "32x32", "256x2"])
def test_stencil_expr_train_vector_tiles(dev, tile, monkeypatch):
    """The vector form's w vectors (the tile's own, then its ring of rows and columns) at tile
    shapes whose vectors fill different numbers of passes: dx and dy bit-identical to the scalar
    form, on a shape with partial tiles in x and y."""
    A, capi, tapes = _mods()
    from ncarray_amd import torch_interop as ti
    g = torch.Generator(device=dev).manual_seed(77)
    shape = (7, 45, 136)
    x = torch.randn(shape, generator=g, device=dev)
    y = torch.randn(shape, generator=g, device=dev)
    lap, _ = A.laplacian_stencil(capi.FLOAT32, capi.FLOAT32)
    tape = A.c0_tape(capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    monkeypatch.setenv("NCA_HIP_TRAIN_TILE", tile)
    res = []
    for scalar in ("1", "0"):
        monkeypatch.setenv("NCA_HIP_TRAIN_SCALAR", scalar)
        dx, dy = torch.empty_like(x), torch.empty_like(y)
        lap.expr_train(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(dx), extras=[ti.view(y)],
                       grad_extras=[ti.view(dy)], scalars=(6.0,), expr_scalars=(0.5,))
        res.append((dx, dy))
    torch.cuda.synchronize()
    (dx1, dy1), (dx2, dy2) = res
    assert torch.equal(dx1, dx2) and torch.equal(dy1, dy2)


# an asymmetric radius-1 stencil with diagonal offsets and scalar coefficients, and an
# expression of it and two extra arrays: f = a*a*y + a*z (operands a, a, y, a, z)
_DIAG_OFFSETS = [[-1, 1, -1], [1, -1, 1], [0, 1, 1], [1, 0, -1], [0, 0, 0], [-1, -1, 0], [0, -1, -1]]
_DIAG_SCALARS = (0.75, 1.5, -2.0)
_DIAG_COEF = (0.75, 1.0, -1.5, 1.0, -2.0, 1.0, -1.0)


def _diag_tapes(capi, tapes):
    P = tapes.P
    st = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MUL), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_ADD),
          P(capi.OP_LOAD_NCARR, 2), P(capi.OP_LOAD_CONST, 1), P(capi.OP_MUL), P(capi.OP_SUB),
          P(capi.OP_LOAD_NCARR, 3), P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 4), P(capi.OP_LOAD_CONST, 2), P(capi.OP_MUL),
          P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 5), P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 6), P(capi.OP_SUB)]
    stencil = capi.Stencil(_DIAG_OFFSETS, st, capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [capi.FLOAT32] * 3)
    ex = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_MUL), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_MUL),
          P(capi.OP_LOAD_NCARR, 3), P(capi.OP_LOAD_NCARR, 4), P(capi.OP_MUL), P(capi.OP_ADD)]
    expr = capi.Tape(ex, [capi.FLOAT32] * 5, [], capi.FLOAT32, capi.FLOAT32)
    return stencil, expr


def _diag_fp64(x, y, z, g):
    """FP64 reference: a = S(x) (zero outside), loss = sum(a*a*y + a*z), w = g * (2*a*y + z),
    dx = S^T w (w zero outside), dy = g * a*a, dz = g * a."""
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    n = x.shape
    xp = np.pad(x, 1)
    a = np.zeros(n)
    for c, o in zip(_DIAG_COEF, _DIAG_OFFSETS):
        a += c * xp[1 + o[0]:1 + o[0] + n[0], 1 + o[1]:1 + o[1] + n[1], 1 + o[2]:1 + o[2] + n[2]]
    w = g * (2 * a * y + z)
    wp = np.pad(w, 1)
    dx = np.zeros(n)
    for c, o in zip(_DIAG_COEF, _DIAG_OFFSETS):
        dx += c * wp[1 - o[0]:1 - o[0] + n[0], 1 - o[1]:1 - o[1] + n[1], 1 - o[2]:1 - o[2] + n[2]]
    return float(np.sum(a * a * y + a * z)), dx, g * a * a, g * a, a


@pytest.mark.parametrize("shape", [(11, 19, 36), (7, 33, 68), (10, 17, 29)])
def test_stencil_expr_train_general_stencil(dev, shape, monkeypatch):
    """Diagonal offsets in every axis (the vector form's neighbour scalars beside other planes'
    and rows' vectors), an adjoint that is not the stencil itself, two extras with their
    gradients and a loss: scalar and vector forms agree bit for bit and both match FP64."""
    A, capi, tapes = _mods()
    from ncarray_amd import torch_interop as ti
    stencil, expr = _diag_tapes(capi, tapes)
    rng = np.random.default_rng(sum(shape) + 7)
    x, y, z = (rng.standard_normal(shape).astype(np.float32) for _ in range(3))
    xt, yt, zt = (torch.from_numpy(v).to(dev) for v in (x, y, z))
    g = 1.25
    gs = torch.tensor(g, dtype=torch.float32, device=dev)
    forms = ("1", "0") if shape[2] % 4 == 0 else ("1",)
    res = []
    for scalar in forms:
        monkeypatch.setenv("NCA_HIP_TRAIN_SCALAR", scalar)
        dx, dy, dz = torch.empty_like(xt), torch.empty_like(yt), torch.empty_like(zt)
        loss = torch.empty((), dtype=torch.float32, device=dev)
        stencil.expr_train(expr, [-1, -1, 0, -1, 1], ti.view(xt), ti.view(dx), extras=[ti.view(yt), ti.view(zt)],
                           grad_extras=[ti.view(dy), ti.view(dz)], scalars=_DIAG_SCALARS, grad_scalar=gs.data_ptr(),
                           fuse=capi.FUSE_SUM, mode=capi.ACCUM_COMPENSATED, sum_out=loss.data_ptr())
        res.append((dx, dy, dz, loss))
    torch.cuda.synchronize()
    l64, dx64, dy64, dz64, a64 = _diag_fp64(x, y, z, g)
    eps = np.finfo(np.float32).eps
    amax = 6.0 * max(1.0, float(np.max(np.abs(x))))        # |a| bound: sum of |c_k| |x|
    for dx, dy, dz, loss in res:
        assert np.max(np.abs(dx.cpu().numpy() - dx64)) <= 64 * eps * amax * amax * max(1.0, float(np.max(np.abs(y))))
        assert np.max(np.abs(dy.cpu().numpy() - dy64)) <= 32 * eps * amax * amax
        assert np.max(np.abs(dz.cpu().numpy() - dz64)) <= 16 * eps * amax
        mass = float(np.sum(np.abs(a64 * a64 * y) + np.abs(a64 * z)))
        assert abs(loss.item() - l64) <= 64 * eps * mass
    if len(res) == 2:
        for u, v in zip(res[0][:3], res[1][:3]):
            assert torch.equal(u, v)
