"""Checker for the oracle (test infrastructure only): ctypes access to oracle/_ref/libnca_ref.so,
the reference's own header-only host routines compiled from the reference sources where they lie
(`make -C oracle ref`, from oracle/ref_driver.cc): the full reductions of host/reductions.hh,
host::scatter_add_recursive, the expression VM of expression/dynamicmvnode.hh and the views'
own advance (layout.hh).  tests/ use it to pin the C restatement (oracle/nca_cpu.c) and the
view-address restatements of oracle.py against the reference's code.  `lib` is None where the
library is not built (the reference tree is absent).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libnca_ref.so")

lib = ctypes.CDLL(LIB_PATH) if os.path.exists(LIB_PATH) else None

SUM, MAX, MIN, ARGMAX, ARGMIN = range(5)
NC, SO = 0, 1

_i64 = ctypes.POINTER(ctypes.c_int64)
if lib is not None:
    lib.ref_reduce.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, _i64, _i64,
                               _i64, ctypes.c_int, ctypes.POINTER(ctypes.c_double), _i64]
    lib.ref_reduce.restype = ctypes.c_int
    lib.ref_offsets.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, _i64, _i64, _i64, ctypes.c_int,
                                ctypes.c_void_p, _i64]
    lib.ref_offsets.restype = ctypes.c_int
    lib.ref_scatter_add_f32.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int64]
    lib.ref_scatter_add_f32.restype = ctypes.c_int
    lib.ref_vm_f32.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    lib.ref_vm_f32.restype = ctypes.c_int


def _meta(vals, ndim):
    a = (ctypes.c_int64 * max(1, ndim))()
    for d in range(ndim):
        a[d] = int(vals[d])
    return a


def reduce(op, dtype_code, addr, shape, strides, aux=None, layout=NC, pointer_axis=-1):
    """host::<op>_recursive over the view at `addr` (NC: aux = per-axis byte offsets; SO:
    aux = suboffsets).  Returns (float64 value, int64 value) of the Scalar result."""
    nd = len(shape)
    aux = [0] * nd if aux is None else aux
    f, i = ctypes.c_double(), ctypes.c_int64()
    rc = lib.ref_reduce(op, dtype_code, layout, ctypes.c_void_p(int(addr)), nd, _meta(shape, nd), _meta(strides, nd),
                        _meta(aux, nd), pointer_axis, ctypes.byref(f), ctypes.byref(i))
    assert rc == 0, f"ref_reduce returned {rc}"
    return f.value, i.value


def offsets(addr, base_addr, shape, strides, aux=None, layout=NC, pointer_axis=-1):
    """Byte offsets from base_addr of every element of the view, in C order, by the layout's advance."""
    nd = len(shape)
    aux = [0] * nd if aux is None else aux
    out = np.empty(max(1, int(np.prod(shape))), np.int64)
    lib.ref_offsets(layout, ctypes.c_void_p(int(addr)), nd, _meta(shape, nd), _meta(strides, nd), _meta(aux, nd),
                    pointer_axis, ctypes.c_void_p(int(base_addr)), out.ctypes.data_as(_i64))
    return out[:int(np.prod(shape))].reshape(tuple(shape))


def scatter_add_f32(dest, idx, src):
    dest = np.array(dest, dtype=np.float32, copy=True)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    src = np.ascontiguousarray(src, dtype=np.float32)
    lib.ref_scatter_add_f32(dest.ctypes.data, dest.size, idx.ctypes.data, src.ctypes.data, src.size)
    return dest


def vm_f32(instrs, arrays, scalars, so_loads=False):
    """DynamicExprMVNode::eval<float> per element of 1-D float32 operands and float32 scalars
    (LOAD_NCARR tapes; so_loads=True for LOAD_SOARR tapes)."""
    ins = np.ascontiguousarray(instrs, dtype=np.uint32)
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in arrays]
    n = arrs[0].size
    ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
    sc = np.ascontiguousarray(list(scalars) or [0.0], dtype=np.float32)
    out = np.empty(n, np.float32)
    lib.ref_vm_f32(ins.ctypes.data, ins.size, ptrs, len(arrs), sc.ctypes.data, len(scalars), n, int(so_loads), out.ctypes.data)
    return out
