"""CPU tests of the C++ host API and its Python bindings (no GPU needed).

Lazy expressions are built on placeholder arrays (no device memory), so these tests check
the tape that ExprMVNode's build_node would produce (mvnode.hh:206-253, 289-441), the dtype
promotion rules (interface.hh:91-100, op_code.hh:101-161), the Python scalar binding rule
(dtype.hh:265-273 through pybind11's variant caster) and the view geometry of indexing,
which must equal numpy's for the same index expression.
"""
import numpy as np
import pytest

import ncarray_amd as nca
from ncarray_amd import capi, tapes

P = nca.NCDevArray._placeholder
D = nca.DType


@pytest.fixture(autouse=True)
def lazy():
    with nca.lazy_mode():
        yield


def test_dtype_enum_matches_reference_order():
    names = ["bool", "char", "uint8", "uint16", "uint32", "uint64", "int8", "int16", "int32", "int64", "float32",
             "float64"]
    for code, name in enumerate(names):
        assert int(getattr(D, name)) == code
    assert int(D.complex64) == 13 and int(D.vdouble4) == 21
    assert int(D.bfloat16) == capi.BFLOAT16 and int(D.float16) == capi.FLOAT16


def test_c0_tape_is_build_node_order():
    a, b = P([8]), P([8])
    e = a * b + nca.exp(-(a * a)) - b * 0.5
    assert isinstance(e, nca.DevExpr)
    assert list(e.instrs) == tapes.fused_c0()
    assert e.dtype == D.float32 and e.work_dtype == D.float32
    # a reflected operator puts the left operand first: 0.5 * b -> CONST, LOAD, MUL
    r = 0.5 * b
    assert list(r.instrs) == [capi.pack(capi.OP_LOAD_CONST, 0), capi.pack(capi.OP_LOAD_NCARR, 0),
                              capi.pack(capi.OP_MUL)]


def test_tape_shifts_indices_when_nesting():
    a, b, c = P([4]), P([4]), P([4])
    left = a + 1.0
    right = b * c - 2.0
    e = left * right
    pk = capi.pack
    assert list(e.instrs) == [pk(capi.OP_LOAD_NCARR, 0), pk(capi.OP_LOAD_CONST, 0), pk(capi.OP_ADD),
                              pk(capi.OP_LOAD_NCARR, 1), pk(capi.OP_LOAD_NCARR, 2), pk(capi.OP_MUL),
                              pk(capi.OP_LOAD_CONST, 1), pk(capi.OP_SUB), pk(capi.OP_MUL)]


@pytest.mark.parametrize("lhs,rhs,op,want", [
    (D.float32, D.float32, "add", D.float32),
    (D.float32, D.float32, "truediv", D.float64),   # truediv_type = double
    (D.int32, D.int32, "truediv", D.float64),
    (D.int16, D.int16, "add", D.int64),             # small-int sum promotion
    (D.uint8, D.uint8, "mul", D.uint8),             # MUL keeps the source type (op_code.hh:125-127)
    (D.uint8, D.uint8, "add", D.uint64),
    (D.int32, D.float32, "add", D.float32),         # max enum
    (D.float32, D.float64, "sub", D.float64),
    (D.float32, D.float32, "gt", D.bool),
    (D.int64, D.int64, "eq", D.bool),
])
def test_promotion(lhs, rhs, op, want):
    e = getattr(P([3], lhs), f"__{op}__")(P([3], rhs))
    assert e.dtype == want


def test_python_scalar_binding():
    f = P([3], D.float32)
    assert (f + 0.5).dtype == D.float32          # float -> float32 alternative
    assert (f * 2).dtype == D.float32
    i = P([3], D.int32)
    assert (i + 5).dtype == D.int32              # 5 -> uint8, max enum is int32
    assert (i + 5).work_dtype == D.int32
    u = P([3], D.uint8)
    assert (u + 300).work_dtype == D.uint16       # first unsigned type that holds 300
    assert (u + (-3)).work_dtype == D.int8
    assert (u + 70000).work_dtype == D.uint32
    assert (u + (1 << 40)).work_dtype == D.uint64
    assert (u + True).work_dtype == D.uint8       # bool (0) < uint8 (2)
    assert (u + np.float32(1)).work_dtype == D.float32
    assert (f + np.float64(1)).work_dtype == D.float32  # numpy.float64 is a Python float


def test_exp_of_integers_evaluates_in_float64():
    e = nca.exp(P([3], D.int32))
    assert e.dtype == D.float64 and e.work_dtype == D.float64


def test_broadcast_shapes():
    assert (P([2, 3]) + P([3])).shape == (2, 3)
    assert (P([4, 1]) * P([1, 5])).shape == (4, 5)
    with pytest.raises(nca.NcaError):
        P([2, 3]) + P([4])


@pytest.mark.parametrize("index", [
    (slice(1, None),), (slice(None, None, -1),), (slice(1, 7, 3), slice(None, None, -2)),
    (Ellipsis, 2), (None, 1, slice(2, 4)), (slice(-3, None), Ellipsis, None), (-1,), (slice(5, 2),),
    (slice(-100, 100, 2), 0),
])
def test_view_geometry_matches_numpy(index):
    shape = (7, 6, 5)
    ref = np.empty(shape, dtype=np.float32)
    v = P(list(shape))[index]
    r = ref[index]
    assert v.shape == r.shape
    # numpy reports 0 strides for extent-0/1 axes differently; compare where extent > 1
    assert all(s == t for s, t, n in zip(v.strides, r.strides, r.shape) if n > 1)


def test_index_errors():
    a = P([3, 4])
    with pytest.raises(nca.NcaError):
        a[3]
    with pytest.raises(nca.NcaError):
        a[0, 0, 0]
    with pytest.raises(TypeError):
        a["x"]


def test_reshape_and_squeeze_views():
    a = P([2, 1, 6])
    assert a.reshape([3, 4]).shape == (3, 4)
    assert a.squeeze().shape == (2, 6)
    with pytest.raises(nca.NcaError):
        a[:, :, ::2].reshape([6])  # not contiguous, no copy requested


def test_cuda_array_interface_of_placeholder():
    a = P([3, 4], D.float32)
    ai = a.__cuda_array_interface__
    assert ai["shape"] == (3, 4) and ai["typestr"] == "<f4" and ai["version"] == 3
    assert ai["strides"] is None
    assert a[:, ::2].__cuda_array_interface__["strides"] == (16, 8)


def test_eager_switch():
    assert not nca.is_eager()
    nca.set_eager(True)
    assert nca.is_eager()
    nca.set_eager(False)
    with nca.lazy_mode():
        assert not nca.is_eager()
