"""The C ABI without a GPU: the library loads, exports every declared entry point, applies
the reference's type rules, validates tapes, and every generated kernel compiles for gfx950."""
import os
import re

import pytest

from ncarray_amd import capi, tapes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = capi.pack


def test_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "nca_hip.h")).read()
    declared = set(re.findall(r"\b(nca_hip_\w+)\s*\(", header))
    assert len(declared) >= 35
    for name in declared:
        assert hasattr(capi.lib, name), name
    assert set(capi.EXPORTED) == declared
    assert capi.lib.nca_hip_version().startswith(b"ncarray-hip")


def test_calls_without_init_fail_cleanly():
    if capi._initialised:
        pytest.skip("a device was initialised in this process")
    v = capi.make_view(None, capi.FLOAT32, [4])
    out = capi.ctypes.c_void_p()
    rc = capi.lib.nca_hip_reduce_full(capi.ctypes.byref(v), capi.REDUCE_SUM, 0, None, None)
    assert rc != capi.OK and capi.last_error()


def test_result_dtypes_follow_the_reference():
    # SumTraits::OutT = op_traits<T>::sum_type (op_traits.hh:72-77, 380-410)
    assert capi.reduce_result_dtype(capi.REDUCE_SUM, capi.FLOAT32) == capi.FLOAT32
    assert capi.reduce_result_dtype(capi.REDUCE_SUM, capi.INT8) == capi.INT64
    assert capi.reduce_result_dtype(capi.REDUCE_SUM, capi.UINT16) == capi.UINT64
    assert capi.reduce_result_dtype(capi.REDUCE_SUM, capi.BOOL) == capi.UINT64
    assert capi.reduce_result_dtype(capi.REDUCE_SUM, capi.INT32) == capi.INT32
    assert capi.reduce_result_dtype(capi.REDUCE_SUM, capi.BFLOAT16) == capi.FLOAT32
    assert capi.reduce_result_dtype(capi.REDUCE_ARGMAX, capi.FLOAT32) == capi.INT64
    assert capi.reduce_result_dtype(capi.REDUCE_MAX, capi.BFLOAT16) == capi.BFLOAT16
    assert capi.reduce_result_dtype(capi.REDUCE_MEAN, capi.INT32) == capi.FLOAT64
    # determine_dtype_for_op (op_code.hh:101-161)
    assert capi.op_result_dtype(capi.OP_DIV, capi.FLOAT32) == capi.FLOAT64
    assert capi.op_result_dtype(capi.OP_ADD, capi.UINT8) == capi.UINT64
    assert capi.op_result_dtype(capi.OP_SUB, capi.UINT8) == capi.INT64
    assert capi.op_result_dtype(capi.OP_GT, capi.FLOAT32) == capi.BOOL
    assert capi.op_result_dtype(capi.OP_MUL, capi.INT16) == capi.INT16
    assert capi.op_result_dtype(capi.OP_EXP, capi.INT32) == capi.FLOAT64
    assert capi.op_result_dtype(capi.OP_EXP, capi.FLOAT32) == capi.FLOAT32


@pytest.mark.parametrize("instrs,err", [
    ([P(capi.OP_ADD)], capi.ERR_INVALID),                                     # stack underflow
    ([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 0)], capi.ERR_INVALID),  # two values left
    ([P(capi.OP_LOAD_NCARR, 5)], capi.ERR_INVALID),                           # operand out of range
    ([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_SZOF)], capi.ERR_UNSUPPORTED),       # not implemented by any engine
    ([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 0), P(capi.OP_BAND)], capi.ERR_UNSUPPORTED),  # bitwise on float
])
def test_tape_validation(instrs, err):
    with pytest.raises(capi.NcaError) as e:
        capi.Tape(instrs, [capi.FLOAT32], [], capi.FLOAT32, capi.FLOAT32)
    assert e.value.code == err


def test_unsupported_dtypes_are_reported():
    with pytest.raises(capi.NcaError) as e:
        capi.Tape([P(capi.OP_LOAD_NCARR, 0)], [capi.COMPLEX64], [], capi.COMPLEX64, capi.COMPLEX64)
    assert e.value.code == capi.ERR_UNSUPPORTED


def _compiles(src):
    ok, log = capi.rtc_compile(src)
    assert ok, log[:4000]


@pytest.mark.parametrize("dt", [capi.FLOAT32, capi.FLOAT64, capi.BFLOAT16, capi.FLOAT16])
def test_expression_kernels_compile(dt):
    t = capi.Tape(tapes.fused_c0(), [dt] * 5, [capi.FLOAT32], dt, dt)
    for variant in (0, 1, 2, 3, 4, 5, 6):
        _compiles(t.source(variant, 3))


def test_integer_and_bool_tapes_compile():
    ins = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_DIV), P(capi.OP_LOAD_CONST, 0),
           P(capi.OP_MOD), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_BAND), P(capi.OP_LOAD_CONST, 1), P(capi.OP_LSHFT),
           P(capi.OP_ABS), P(capi.OP_LOAD_NCARR, 3), P(capi.OP_FDIV)]
    for dt in (capi.INT32, capi.INT64, capi.UINT8, capi.UINT64):
        t = capi.Tape(ins, [dt] * 4, [dt, dt], dt, dt)
        _compiles(t.source(0))
        _compiles(t.source(1, 2))
    cmp = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_CONST, 0), P(capi.OP_GT), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_LAND),
           P(capi.OP_LNOT)]
    _compiles(capi.Tape(cmp, [capi.FLOAT32, capi.BOOL], [capi.FLOAT32], capi.FLOAT32, capi.BOOL).source(0))


def test_linearizable_vs_vm_evaluation_dtype():
    """float32 / float32 into a float64 destination: the linearizable tape runs in the work
    dtype (float32), a deeper tape runs in the destination dtype (float64), as the two
    reference paths do (gpuengine.hh:103-191, dynamicmvnode.hh:244-249)."""
    lin = capi.Tape([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_DIV)], [capi.FLOAT32] * 2, [],
                    capi.FLOAT32, capi.FLOAT64)
    assert "const float t" in lin.source(0)
    deep = capi.Tape([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_MUL),
                      P(capi.OP_DIV)], [capi.FLOAT32] * 3, [], capi.FLOAT32, capi.FLOAT64)
    assert "const double t" in deep.source(0)


@pytest.mark.parametrize("dt", [capi.FLOAT32, capi.BFLOAT16])
def test_stencil_kernels_compile(dt):
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), dt, capi.FLOAT32, dt, [capi.FLOAT32])
    for variant in (0, 1, 2, 3, 4, 5, 6):
        _compiles(st.source(variant))
    # compensated column sums: one Kahan step per plane and column
    assert "nca_kahan_add(cs[0][0], cc[0][0]" in st.source(4) and "cc[" not in st.source(1)
    adj = st.adjoint()
    for variant in (0, 2):
        _compiles(adj.source(variant))


@pytest.mark.parametrize("dt", [capi.FLOAT32, capi.BFLOAT16])
def test_stencil_with_extra_array_compiles(dt):
    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian_then_c0(), dt, capi.FLOAT32, capi.FLOAT32,
                      [capi.FLOAT32, capi.FLOAT32], extra_dtypes=[dt])
    for variant in (0, 1, 2):
        _compiles(st.source(variant))


def _sum_tape(n):
    t = [P(capi.OP_LOAD_NCARR, 0)]
    for k in range(1, n):
        t += [P(capi.OP_LOAD_NCARR, k), P(capi.OP_ADD)]
    return t


def test_march_generator_choice_and_radius():
    # pure z offsets (dz only with dy = dx = 0): register z-queue march; radius 2 in z and y,
    # 3 in x, plus an xy diagonal within the plane
    offs = [[-2, 0, 0], [2, 0, 0], [-1, 0, 0], [0, -2, 0], [0, 2, 1], [0, 0, -3], [0, 0, 3], [0, -1, -1], [0, 0, 0]]
    st = capi.Stencil(offs, _sum_tape(len(offs)), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    for variant in (0, 1):
        src = st.source(variant)
        assert "cq[" in src and "lds[2][LH][LW]" in src
        _compiles(src)
    # a z offset combined with an x offset needs other planes' neighbours: LDS ring march
    diag = capi.Stencil([[1, 0, 1], [0, 0, 0]], _sum_tape(2), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    src = diag.source(1)
    assert "lds[NS][LH][LW]" in src
    _compiles(src)


def test_nonlinear_stencil_has_no_adjoint():
    ins = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_MUL)]
    st = capi.Stencil([[0, 0, 0], [1, 0, 0]], ins, capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    with pytest.raises(capi.NcaError):
        st.adjoint()


def test_adjoint_source_negates_offsets():
    st = capi.Stencil([[0, 0, 1], [0, 0, 0]], [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_SUB)],
                      capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    src = st.adjoint().source(2)
    assert "+ (-1)" in src


@pytest.mark.parametrize("dt", [capi.FLOAT32, capi.BFLOAT16])
def test_tile_kernels_compile(dt):
    """The tile kernel for a 2-D 5-point Laplacian, a 1-D three-point stencil (launched as 2-D)
    and the 2x2x2 pixel-binning window, plain and with pointer-read rows and fused reductions."""
    lap2 = [[0, 0], [-1, 0], [1, 0], [0, -1], [0, 1]]
    five = [P(capi.OP_LOAD_NCARR, 1), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 3),
            P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 4), P(capi.OP_ADD), P(capi.OP_LOAD_CONST, 0),
            P(capi.OP_LOAD_NCARR, 0), P(capi.OP_MUL), P(capi.OP_SUB)]
    win = [[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    for st in (capi.Stencil(lap2, five, dt, capi.FLOAT32, dt, [capi.FLOAT32]),
               capi.Stencil([[-1], [0], [1]], _sum_tape(3), dt, capi.FLOAT32, capi.FLOAT32),
               capi.Stencil(win, _sum_tape(8), dt, capi.FLOAT32, capi.FLOAT32)):
        plain, fused = st.source(5), st.source(6)
        _compiles(plain)
        _compiles(fused)
        assert "nca_stencil_tile" in plain and "rp = nca_chase" not in plain
        assert "rp = nca_chase(rp, ssub.v[0])" in fused and "rd = nca_chase(rd, dsub.v[0])" in fused


def test_stencil_rejects_bad_descriptions():
    with pytest.raises(capi.NcaError):
        capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32, [])


@pytest.mark.parametrize("dt", [capi.FLOAT32, capi.BFLOAT16])
def test_training_step_kernel_compiles(dt):
    """The one-pass training kernel (loss, dx via the adjoint from LDS, dy) for the Laplacian
    and the C0 expression: the stencil loads read x's three LDS planes, the adjoint reads w at
    the negated offsets, and the loss is reduced only with fuse."""
    from ncarray_amd import autograd as A
    lap, _ = A.laplacian_stencil(dt, capi.FLOAT32)
    tape = A.c0_tape(capi.FLOAT32, dt, capi.FLOAT32)
    plain = lap.train_source(tape, [-1, 0, -1, -1, 0], [dt], dt)
    fused = lap.train_source(tape, [-1, 0, -1, -1, 0], [dt], dt, fuse=capi.FUSE_SUM | capi.FUSE_MAX,
                             mode=capi.ACCUM_COMPENSATED)
    for src in (plain, fused):
        _compiles(src)
        assert "xs[sm][xb[m] + (0)]" in src and "xs[sp][xb[m] + (0)]" in src   # the z neighbours
        assert "ws[a2][tb[m] + (0)]" in src and "ws[a0][tb[m] + (0)]" in src   # w at z + 1, z - 1
    assert "nca_kahan_add(acc_s" in fused and "acc_s" not in plain and "nca_maxnum(acc_mx" in fused


@pytest.mark.parametrize("dt", [capi.FLOAT32, capi.BFLOAT16])
def test_training_step_vector_kernel_compiles(dt):
    """The 4-point-vector form: x neighbours from aligned LDS vectors and the scalars beside
    them, the adjoint's from the w planes the same way."""
    from ncarray_amd import autograd as A
    lap, _ = A.laplacian_stencil(dt, capi.FLOAT32)
    tape = A.c0_tape(capi.FLOAT32, dt, capi.FLOAT32)
    for fuse in (0, capi.FUSE_SUM | capi.FUSE_MIN):
        src = lap.train_source(tape, [-1, 0, -1, -1, 0], [dt], dt, fuse=fuse, mode=capi.ACCUM_COMPENSATED, vec=True)
        _compiles(src)
        assert "_v4(" in src and "const lvec VX0_0" in src and "const lvec VW1_0" in src and "LW0_0_1" in src
