"""Differentiation of the hot-path operators behind torch.autograd.Function.

Every forward and backward runs a HIP kernel of libnca_hip through the C ABI:
  * ExprFunction: a fused expression; backward = the derivative expression generated
    symbolically from the tape, times the upstream gradient (one kernel, all inputs).
  * StencilFunction: a stencil linear in its loads; backward = the transposed stencil
    (negated offsets, coefficients d f / d v_k).
  * SumFunction / MeanFunction: full reductions; backward = a broadcast kernel.
  * BinWeightedSumFunction: per-frame weighted bin sums; backward = a gather kernel.
  * LaplacianC0Loss: loss = sum(a*b + exp(-a*a) - 0.5*b) with a = Laplacian(x), b = y, in
    ONE forward kernel (stencil + expression + sum, nothing materialised); backward = one
    march pass that recomputes a and writes dloss/da and dloss/db, then the transposed stencil.
    laplacian_c0_value_and_grad runs the same three kernels without torch's autograd, from the
    calling thread, so a whole step can be captured in a hipGraph.
torch only supplies tensors (device memory) and streams.
"""
import ctypes

import torch

from . import capi, tapes
from . import torch_interop as ti

_cache = {}


def _init(t):
    capi.init(t.device.index or 0)


def _code(t):
    return ti.TORCH_TO_CODE[t.dtype]


def _stream():
    return ti.stream_ptr()


def laplacian_stencil(src_code, dst_code):
    key = ("lap", src_code, dst_code)
    if key not in _cache:
        work = capi.FLOAT64 if capi.FLOAT64 in (src_code, dst_code) else capi.FLOAT32
        st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), src_code, work, dst_code, [work])
        _cache[key] = (st, st.adjoint())
    return _cache[key]


def c0_tape(a_code, b_code, dst_code):
    key = ("c0", a_code, b_code, dst_code)
    if key not in _cache:
        work = a_code if a_code == b_code else capi.FLOAT32
        _cache[key] = capi.Tape(tapes.fused_c0(), [a_code, b_code, a_code, a_code, b_code], [capi.FLOAT32], work,
                                dst_code)
    return _cache[key]


def fused_loss_stencil(x_code, y_code):
    key = ("lapc0", x_code, y_code)
    if key not in _cache:
        _cache[key] = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian_then_c0(), x_code, capi.FLOAT32,
                                   capi.FLOAT32, [capi.FLOAT32, capi.FLOAT32], extra_dtypes=[y_code])
    return _cache[key]


# ---- expressions -------------------------------------------------------------------------------
class ExprFunction(torch.autograd.Function):
    """out = tape(inputs...) for a compiled capi.Tape whose operands are `tensors`."""

    @staticmethod
    def forward(ctx, tape, scalars, *tensors):
        _init(tensors[0])
        ts = [t.contiguous() for t in tensors]
        shape = ts[0].shape
        out = torch.empty(shape, dtype=ti.CODE_TO_TORCH[tape.dest_dtype], device=ts[0].device)
        tape.launch([ti.view(t) for t in ts], ti.view(out), scalars=scalars, stream=_stream())
        ctx.tape, ctx.scalars = tape, scalars
        ctx.save_for_backward(*ts)
        return out

    @staticmethod
    def backward(ctx, g):
        ts = ctx.saved_tensors
        need = ctx.needs_input_grad[2:]
        # operands that are the same tensor share one gradient buffer: the kernel writes the
        # summed derivative at the first occurrence; later occurrences return None (zero)
        bufs, views, wrt, first = {}, [], [], []
        for k, t in enumerate(ts):
            key = t.data_ptr()
            is_first = key not in bufs
            if is_first:
                bufs[key] = torch.empty_like(t) if need[k] else torch.empty(0, device=t.device)
            views.append(ti.view(bufs[key]) if need[k] else ti.view(t))
            wrt.append(bool(need[k]) and is_first)
            first.append(is_first)
        if any(wrt):
            ctx.tape.vjp([ti.view(t) for t in ts], views, wrt, grad_out=ti.view(g.contiguous()), scalars=ctx.scalars,
                         stream=_stream())
        out = [bufs[t.data_ptr()] if (need[k] and first[k]) else None for k, t in enumerate(ts)]
        return (None, None) + tuple(out)


def expr(tape, scalars, *tensors):
    return ExprFunction.apply(tape, tuple(scalars), *tensors)


# ---- stencils ------------------------------------------------------------------------------------
class StencilFunction(torch.autograd.Function):
    """dst = stencil(src) with zero-Dirichlet boundary; backward = transposed stencil."""

    @staticmethod
    def forward(ctx, stencil, adjoint, scalars, x):
        _init(x)
        xc = x.contiguous()
        out = torch.empty(xc.shape, dtype=ti.CODE_TO_TORCH[stencil.dest_dtype], device=x.device)
        stencil.apply(ti.view(xc), ti.view(out), capi.BOUNDARY_ZERO, scalars=scalars, stream=_stream())
        ctx.adjoint, ctx.scalars = adjoint, scalars
        ctx.x_dtype = x.dtype
        return out

    @staticmethod
    def backward(ctx, g):
        gc = g.contiguous()
        if ti.TORCH_TO_CODE[gc.dtype] != ctx.adjoint.src_dtype:
            gc = gc.to(ti.CODE_TO_TORCH[ctx.adjoint.src_dtype])
        gx = torch.empty(gc.shape, dtype=ti.CODE_TO_TORCH[ctx.adjoint.dest_dtype], device=g.device)
        ctx.adjoint.apply(ti.view(gc), ti.view(gx), capi.BOUNDARY_ZERO, scalars=ctx.scalars, stream=_stream())
        return None, None, None, gx


def laplacian(x, out_dtype=None):
    """7-point Laplacian with zero Dirichlet boundary (differentiable)."""
    xc = _code(x)
    dc = xc if out_dtype is None else ti.TORCH_TO_CODE[out_dtype]
    st, adj = laplacian_stencil(xc, dc)
    return StencilFunction.apply(st, adj, tapes.LAPLACIAN_SCALARS, x)


# ---- reductions -----------------------------------------------------------------------------------
def _reduce_full(x, op, mode):
    _init(x)
    xc = x.contiguous()
    code = capi.reduce_result_dtype(op, _code(xc))
    out = torch.empty((), dtype=ti.CODE_TO_TORCH[code], device=x.device)
    capi.call("nca_hip_reduce_full", ctypes.byref(ti.view(xc)), op, mode, out.data_ptr(), _stream())
    return out


_BCAST = [capi.pack(capi.OP_LOAD_NCARR, 0), capi.pack(capi.OP_LOAD_CONST, 0), capi.pack(capi.OP_MUL)]


def _broadcast(g, shape, dtype, scale=1.0):
    """grad_x = g * scale broadcast to `shape` (the backward of sum and mean)."""
    gcode = _code(g)
    dcode = ti.TORCH_TO_CODE[dtype]
    key = ("bcast", gcode, dcode)
    if key not in _cache:
        _cache[key] = capi.Tape(_BCAST, [gcode], [capi.FLOAT64], capi.FLOAT64, dcode)
    out = torch.empty(shape, dtype=dtype, device=g.device)
    gv = capi.make_view(g.data_ptr(), gcode, list(shape), [0] * len(shape))
    _cache[key].launch([gv], ti.view(out), scalars=(scale,), stream=_stream())
    return out


class SumFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, mode):
        ctx.shape, ctx.dtype = x.shape, x.dtype
        return _reduce_full(x, capi.REDUCE_SUM, mode)

    @staticmethod
    def backward(ctx, g):
        return _broadcast(g.contiguous(), ctx.shape, ctx.dtype), None


class MeanFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, mode):
        ctx.shape, ctx.dtype, ctx.n = x.shape, x.dtype, x.numel()
        return _reduce_full(x, capi.REDUCE_MEAN, mode)

    @staticmethod
    def backward(ctx, g):
        return _broadcast(g.contiguous(), ctx.shape, ctx.dtype, 1.0 / max(ctx.n, 1)), None


def sum(x, compensated=True):  # noqa: A001  (mirrors ArrayImpl::sum)
    return SumFunction.apply(x, capi.ACCUM_COMPENSATED if compensated else capi.ACCUM_PLAIN)


def mean(x, compensated=True):
    return MeanFunction.apply(x, capi.ACCUM_COMPENSATED if compensated else capi.ACCUM_PLAIN)


# ---- binning ---------------------------------------------------------------------------------------
def bin_frames(frames, binmap, n_bins, count_nonzero=True, sums_dtype=torch.float32, counts_dtype=torch.int32):
    """(counts, sums) per frame and bin; counts are exact, sums float64-accumulated."""
    _init(frames)
    fr = frames.contiguous()
    F = fr.shape[0] if fr.dim() > 1 else 1
    counts = torch.empty((F, n_bins), dtype=counts_dtype, device=fr.device)
    sums = torch.empty((F, n_bins), dtype=sums_dtype, device=fr.device)
    capi.call("nca_hip_bin", ctypes.byref(ti.view(fr)), ctypes.byref(ti.view(binmap.contiguous())), n_bins,
              capi.COUNT_NONZERO if count_nonzero else capi.COUNT_ALL, ctypes.byref(ti.view(counts)),
              ctypes.byref(ti.view(sums)), _stream())
    return counts, sums


class BinWeightedSumFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, frames, binmap, n_bins):
        _, sums = bin_frames(frames, binmap, n_bins)
        ctx.save_for_backward(binmap)
        ctx.shape, ctx.dtype = frames.shape, frames.dtype
        return sums

    @staticmethod
    def backward(ctx, g):
        (binmap,) = ctx.saved_tensors
        gx = torch.empty(ctx.shape, dtype=ctx.dtype if ctx.dtype.is_floating_point else torch.float32, device=g.device)
        gc = g.contiguous()
        if gc.dtype not in (torch.float32, torch.float64):
            gc = gc.float()
        capi.call("nca_hip_bin_backward", ctypes.byref(ti.view(gc)), ctypes.byref(ti.view(binmap.contiguous())),
                  ctypes.byref(ti.view(gx)), _stream())
        return gx, None, None


def bin_weighted_sum(frames, binmap, n_bins):
    return BinWeightedSumFunction.apply(frames, binmap, n_bins)


# ---- the fused differentiable pipeline ----------------------------------------------------------------
class LaplacianC0Loss(torch.autograd.Function):
    """loss = sum(a*b + exp(-a*a) - 0.5*b), a = Laplacian(x) (zero Dirichlet), b = y."""

    @staticmethod
    def forward(ctx, x, y, mode):
        _init(x)
        xc, yc = x.contiguous(), y.contiguous()
        st = fused_loss_stencil(_code(xc), _code(yc))
        loss = torch.empty((), dtype=torch.float32, device=x.device)
        st.apply_reduce(ti.view(xc), ti.shape_view(list(xc.shape), capi.FLOAT32), capi.BOUNDARY_ZERO,
                        fuse=capi.FUSE_SUM, mode=mode, sum_out=loss.data_ptr(), extras=[ti.view(yc)],
                        scalars=(6.0, 0.5), stream=_stream())
        ctx.save_for_backward(xc, yc)
        return loss

    @staticmethod
    def backward(ctx, g):
        xc, yc = ctx.saved_tensors
        return laplacian_c0_backward(xc, yc, g.contiguous()) + (None,)


def _march_vjp_ok(x, y):
    """The one-pass backward needs what the march kernel needs: 3-D C-contiguous arrays whose
    last extent and base addresses hold whole 16-byte vectors of x's elements."""
    vx = 16 // x.element_size()
    return (x.dim() == 3 and x.is_contiguous() and y.is_contiguous() and x.shape[-1] % vx == 0
            and x.data_ptr() % 16 == 0 and y.data_ptr() % (vx * y.element_size()) == 0)


def _train_ok(x, y):
    """What the one-pass training kernel needs: 3-D C-contiguous x and y of one shape."""
    return x.dim() == 3 and x.is_contiguous() and y.is_contiguous() and tuple(x.shape) == tuple(y.shape)


def _default_method(x, y):
    return "train" if _train_ok(x, y) else ("vjp" if _march_vjp_ok(x, y) else "compose")


def laplacian_c0_backward(x, y, g=None, method=None):
    """(dx, dy) of sum(C0(Laplacian(x), y)) times the 0-d upstream gradient g (None: 1).

    method "train": one pass (nca_hip_stencil_expr_train) stages x with a two-point halo in LDS,
    evaluates w = dC0/da * g on each tile and a one-point ring, writes dy = dC0/db * g and
    applies the transposed Laplacian to w from LDS, so only x, y, dx and dy touch memory (4-point
    vectors when the last extent is a multiple of 4 and the arrays are aligned).  "vjp": the
    two-kernel form (nca_hip_stencil_expr_vjp writes w and dy in one march pass, the
    transposed stencil turns w into dx; the vectorised march kernels' shape and alignment rules
    apply).  "compose": the Laplacian, the derivative expression and the transposed stencil as
    three kernels (any shape).  The default is the fastest measured at 256^3: "train" for 3-D
    C-contiguous arrays (float32 58.3-59.2 us against 45.5 + 32.5 us for "vjp"; bfloat16 55.8 us
    against 34.9 + 26.9 us), else "vjp" where its shape and alignment rules hold, else "compose"."""
    s = _stream()
    xcode, ycode = _code(x), _code(y)
    lap, adj = laplacian_stencil(xcode, capi.FLOAT32)
    tape = c0_tape(capi.FLOAT32, ycode, capi.FLOAT32)
    if g is None:
        gf = None
    else:
        gf = g if g.dtype == torch.float32 else g.float()
    gptr = None if gf is None else gf.data_ptr()
    if method is None:
        method = _default_method(x, y)
    if method == "train":
        dx = torch.empty(x.shape, dtype=x.dtype, device=x.device)
        gy = torch.empty_like(y)
        lap.expr_train(tape, [-1, 0, -1, -1, 0], ti.view(x), ti.view(dx), extras=[ti.view(y)],
                       grad_extras=[ti.view(gy)], scalars=(6.0,), expr_scalars=(0.5,), grad_scalar=gptr, stream=s)
        return dx, gy
    if method == "vjp":
        w = torch.empty(x.shape, dtype=torch.float32, device=x.device)
        gy = torch.empty_like(y)
        vw = ti.view(w)
        lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(x), vw, extras=[ti.view(y)], grad_extras=[ti.view(gy)],
                     scalars=(6.0,), expr_scalars=(0.5,), grad_scalar=gptr, stream=s)
        dx = torch.empty(x.shape, dtype=x.dtype, device=x.device)
        adj.apply(vw, ti.view(dx), capi.BOUNDARY_ZERO, scalars=(6.0,), stream=s)
        return dx, gy
    if method != "compose":
        raise ValueError(f"unknown method {method!r}")
    if gf is None:
        gf = torch.ones((), dtype=torch.float32, device=x.device)
    a = torch.empty(x.shape, dtype=torch.float32, device=x.device)
    va = ti.view(a)
    lap.apply(ti.view(x), va, capi.BOUNDARY_ZERO, scalars=(6.0,), stream=s)
    ga = torch.empty_like(a)
    gy = torch.empty_like(y)
    vy, vga, vgy = ti.view(y), ti.view(ga), ti.view(gy)
    tape.vjp([va, vy, va, va, vy], [vga, vgy, vga, vga, vgy], [1, 1, 1, 1, 1], grad_scalar=gf.data_ptr(),
             scalars=(0.5,), stream=s)
    # the adjoint of the (x dtype -> float32) Laplacian reads float32 and writes x's dtype
    dx = torch.empty(x.shape, dtype=x.dtype, device=x.device)
    adj.apply(vga, ti.view(dx), capi.BOUNDARY_ZERO, scalars=(6.0,), stream=s)
    return dx, gy


def laplacian_c0_loss(x, y, compensated=True):
    return LaplacianC0Loss.apply(x, y, capi.ACCUM_COMPENSATED if compensated else capi.ACCUM_PLAIN)


def laplacian_c0_value_and_grad(x, y, compensated=True, grad_scale=None):
    """(loss, dx, dy) of loss = sum(C0(Laplacian(x), y)) without torch's autograd, all launched
    from the calling thread on its current stream, so a whole training step can be captured in
    one hipGraph.  The loss is fused into the gradient pass (per-thread float32, Kahan when
    compensated, float64 across blocks), on the path laplacian_c0_backward would take: the
    one-pass nca_hip_stencil_expr_train, or nca_hip_stencil_expr_vjp_reduce (loss, dloss/da,
    dy) and the transposed stencil (dx); for arrays neither takes, the fused forward kernel and
    laplacian_c0_backward.  grad_scale: optional 0-d float32 device tensor scaling the gradients
    (the upstream gradient; default 1)."""
    _init(x)
    xc, yc = x.detach().contiguous(), y.detach().contiguous()
    loss = torch.empty((), dtype=torch.float32, device=x.device)
    mode = capi.ACCUM_COMPENSATED if compensated else capi.ACCUM_PLAIN
    gf = None if grad_scale is None else (grad_scale if grad_scale.dtype == torch.float32 else grad_scale.float())
    method = _default_method(xc, yc)
    if method == "vjp":
        s = _stream()
        lap, adj = laplacian_stencil(_code(xc), capi.FLOAT32)
        tape = c0_tape(capi.FLOAT32, _code(yc), capi.FLOAT32)
        w = torch.empty(xc.shape, dtype=torch.float32, device=x.device)
        dy = torch.empty_like(yc)
        vw = ti.view(w)
        lap.expr_vjp(tape, [-1, 0, -1, -1, 0], ti.view(xc), vw, extras=[ti.view(yc)], grad_extras=[ti.view(dy)],
                     scalars=(6.0,), expr_scalars=(0.5,), grad_scalar=None if gf is None else gf.data_ptr(),
                     stream=s, fuse=capi.FUSE_SUM, mode=mode, sum_out=loss.data_ptr())
        dx = torch.empty(xc.shape, dtype=xc.dtype, device=x.device)
        adj.apply(vw, ti.view(dx), capi.BOUNDARY_ZERO, scalars=(6.0,), stream=s)
        return loss, dx, dy
    if method == "train":
        lap, _ = laplacian_stencil(_code(xc), capi.FLOAT32)
        tape = c0_tape(capi.FLOAT32, _code(yc), capi.FLOAT32)
        dx = torch.empty(xc.shape, dtype=xc.dtype, device=x.device)
        dy = torch.empty_like(yc)
        lap.expr_train(tape, [-1, 0, -1, -1, 0], ti.view(xc), ti.view(dx), extras=[ti.view(yc)],
                       grad_extras=[ti.view(dy)], scalars=(6.0,), expr_scalars=(0.5,),
                       grad_scalar=None if gf is None else gf.data_ptr(), stream=_stream(), fuse=capi.FUSE_SUM,
                       mode=mode, sum_out=loss.data_ptr())
        return loss, dx, dy
    st = fused_loss_stencil(_code(xc), _code(yc))
    st.apply_reduce(ti.view(xc), ti.shape_view(list(xc.shape), capi.FLOAT32), capi.BOUNDARY_ZERO,
                    fuse=capi.FUSE_SUM, mode=mode, sum_out=loss.data_ptr(), extras=[ti.view(yc)],
                    scalars=(6.0, 0.5), stream=_stream())
    dx, dy = laplacian_c0_backward(xc, yc, grad_scale)
    return loss, dx, dy
