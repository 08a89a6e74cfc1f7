"""Expression tapes for the standard workloads, written as the reference's operator chain
records them (ExprMVNode::build_node + operators, expression/mvnode.hh:181-459): every
array operand occurrence is its own LOAD, scalars are LOAD_CONST, the tree is in RPN."""
from . import capi

P = capi.pack


def fused_c0():
    """out = a*b + exp(-a*a) - 0.5*b  (operands: a, b, a, a, b; scalar 0.5)."""
    return [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_MUL),
            P(capi.OP_LOAD_NCARR, 2), P(capi.OP_LOAD_NCARR, 3), P(capi.OP_MUL), P(capi.OP_NEG), P(capi.OP_EXP),
            P(capi.OP_ADD),
            P(capi.OP_LOAD_NCARR, 4), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MUL),
            P(capi.OP_SUB)]


C0_OPERANDS = ("a", "b", "a", "a", "b")
C0_SCALARS = (0.5,)

# 7-point Laplacian: x[z-1] + x[z+1] + x[y-1] + x[y+1] + x[x-1] + x[x+1] - 6*x
LAPLACIAN_OFFSETS = [[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1], [0, 0, 0]]
LAPLACIAN_SCALARS = (6.0,)


def laplacian():
    t = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_ADD)]
    for k in range(2, 6):
        t += [P(capi.OP_LOAD_NCARR, k), P(capi.OP_ADD)]
    t += [P(capi.OP_LOAD_CONST, 0), P(capi.OP_LOAD_NCARR, 6), P(capi.OP_MUL), P(capi.OP_SUB)]
    return t


def laplacian_then_c0():
    """fused expr(stencil(x), y): the C0 expression with a := Laplacian(x), b := y.
    Loads 0..6 are the Laplacian offsets, load 7 is the extra pointwise array y."""
    lap = laplacian()
    # a*b + exp(-a*a) - 0.5*b with a = lap (three copies), b = y (LOAD 7)
    t = lap + [P(capi.OP_LOAD_NCARR, 7), P(capi.OP_MUL)]
    t += lap + lap + [P(capi.OP_MUL), P(capi.OP_NEG), P(capi.OP_EXP), P(capi.OP_ADD)]
    t += [P(capi.OP_LOAD_NCARR, 7), P(capi.OP_LOAD_CONST, 1), P(capi.OP_MUL), P(capi.OP_SUB)]
    return t
