"""Python surface of the device arrays, as the reference's `ncarray` package exposes it.

`NCDevArray`, `DevExpr`, `DType`, `OpCode`, `Stencil`, `set_eager`, `is_eager`,
`materialize` and the math functions come from the native module `_ncarray` (C++ over
libnca_hip.so).  This file adds the host containers that the device API hands data to and
from (NCArray, NCArrayRef: numpy arrays with `to_device`) and `lazy_mode`
(ncarray/__init__.py:193-203 of the reference).  Host arrays are plain numpy storage: the
reference's own host engine (asmjit JIT) is not part of this backend.
"""
import contextlib

import numpy as np

from . import _ncarray
from ._ncarray import (DType, OpCode, Boundary, Accum, NCDevArray, DevExpr, Stencil, NcaError,  # noqa: F401
                       set_eager, is_eager, materialize, exp, log, sqrt, init)

_abs = _ncarray.abs

_NUMPY = {
    DType.bool: np.bool_, DType.char: np.int8, DType.uint8: np.uint8, DType.uint16: np.uint16,
    DType.uint32: np.uint32, DType.uint64: np.uint64, DType.int8: np.int8, DType.int16: np.int16,
    DType.int32: np.int32, DType.int64: np.int64, DType.float32: np.float32, DType.float64: np.float64,
    DType.float16: np.float16,
}


def numpy_dtype(dtype):
    """numpy dtype of a DType (bfloat16 has none: host copies of it are float32)."""
    if dtype == DType.bfloat16:
        return np.dtype(np.float32)
    try:
        return np.dtype(_NUMPY[dtype])
    except KeyError:
        raise TypeError(f"{dtype} has no numpy counterpart") from None


class NCArray(np.ndarray):
    """A host array (NCArray(shape, dtype) of the reference): numpy storage plus to_device()."""

    def __new__(cls, shape, dtype=DType.float32):
        return np.zeros(shape, dtype=numpy_dtype(dtype)).view(cls)

    def to_device(self, dtype=None):
        """Copy to a new NCDevArray (optionally converting to `dtype`)."""
        return NCDevArray.from_host(np.asarray(self), dtype)

    def to_host(self, dtype=None):
        raise RuntimeError("Array is already on host!")

    @property
    def ncdtype(self):
        return _dtype_of(self.dtype)


class NCArrayRef(NCArray):
    """A host array referencing existing numpy memory (NCArrayRef(data) of the reference)."""

    def __new__(cls, data, read_only=False):
        arr = np.asarray(data)
        view = arr.view(cls)
        if read_only:
            view.flags.writeable = False
        return view


def _interface_dtype(obj):
    """The DType of a framework tensor whose __cuda_array_interface__ typestr is ambiguous
    (torch exports bfloat16 as "|V2")."""
    if str(getattr(obj, "dtype", "")) == "torch.bfloat16":
        return DType.bfloat16
    return None


def _device_row(obj):
    if isinstance(obj, NCDevArray):
        return obj
    if hasattr(obj, "__cuda_array_interface__"):
        return NCDevArray.from_cuda_array(obj, _interface_dtype(obj))
    raise TypeError(f"NCDevArrayRef rows must be device arrays (NCDevArray or __cuda_array_interface__), "
                    f"not {type(obj).__name__}")


def NCDevArrayRef(data, read_only=False):  # noqa: N802  (the reference's class name)
    """A device array over separately allocated rows, with a pointer axis in front.

    Mirrors the reference's NCDevArrayRef constructors (pyncarray/ncdevarray_ref.cu:39-52,
    utilities.hh:226-314): a list of equally shaped device arrays becomes one array of shape
    (len(list), *row_shape) whose axis 0 is a pointer axis; a single array gets a pointer axis
    of length 1.  The rows are referenced, not copied, and kept alive by the result.
    """
    if read_only:
        raise ValueError("read-only device references are not supported")
    rows = list(data) if isinstance(data, (list, tuple)) else [data]
    return NCDevArray.from_rows([_device_row(r) for r in rows])


def _dtype_of(np_dtype):
    np_dtype = np.dtype(np_dtype)
    for k, v in _NUMPY.items():
        if np.dtype(v) == np_dtype and k != DType.char:
            return k
    raise TypeError(f"numpy dtype {np_dtype} has no DType")


@contextlib.contextmanager
def lazy_mode():
    """Context manager to delay immediate evaluation of results."""
    old = is_eager()
    set_eager(False)
    try:
        yield
    finally:
        set_eager(old)


def abs(x):  # noqa: A001  (mirrors numpy's name)
    return _abs(x)


def from_torch(tensor):
    """Zero-copy NCDevArray view of a device torch tensor (kept alive by the view)."""
    return NCDevArray.from_cuda_array(tensor, _interface_dtype(tensor))


def from_dlpack(obj):
    """Zero-copy NCDevArray view of a DLPack producer (an object with __dlpack__, such as a torch
    tensor, or a "dltensor" capsule); the producer's memory stays alive while the view does.
    NCDevArray exports itself the same way (__dlpack__ / __dlpack_device__), so
    torch.from_dlpack(array) is a zero-copy tensor of it."""
    return NCDevArray.from_dlpack(obj)


_ncarray._set_host_array_type(NCArray)

__all__ = ["DType", "OpCode", "Boundary", "Accum", "NCDevArray", "DevExpr", "Stencil", "NcaError", "NCArray",
           "NCArrayRef", "NCDevArrayRef", "set_eager", "is_eager", "materialize", "lazy_mode", "exp", "log", "sqrt", "abs", "init",
           "from_torch", "numpy_dtype"]
