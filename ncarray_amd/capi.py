"""ctypes binding of the C ABI declared in include/nca_hip.h (libnca_hip.so).

This module only marshals arguments: every computation happens in the HIP library.
If the library has not been built, importing this module raises ImportError (there is
no CPU fallback).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnca_hip.so")

MAX_NDIM = 10

# status codes
OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_HIP, ERR_COMPILE, ERR_OOM, ERR_NOT_INIT = range(7)

# dtype codes (reference DType order, plus bfloat16/float16)
BOOL, CHAR, UINT8, UINT16, UINT32, UINT64, INT8, INT16, INT32, INT64, FLOAT32, FLOAT64 = range(12)
FLOAT128, COMPLEX64, COMPLEX128, COMPLEX256 = 12, 13, 14, 15
BFLOAT16, FLOAT16 = 22, 23

# opcodes (reference OpCode order, plus unary math)
(OP_NOOP, OP_IDX, OP_LOAD_NCARR, OP_LOAD_SOARR, OP_LOAD_CONST, OP_NEG, OP_INC, OP_DEC, OP_SZOF,
 OP_ADDR, OP_INDR, OP_CAST, OP_LNOT, OP_BNOT, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD, OP_FDIV,
 OP_EQ, OP_NE, OP_LT, OP_LE, OP_GT, OP_GE, OP_LAND, OP_LOR, OP_BAND, OP_BOR, OP_XOR, OP_LSHFT,
 OP_RSHFT) = range(33)
OP_EXP, OP_LOG, OP_SQRT, OP_ABS = 64, 65, 66, 67

REDUCE_SUM, REDUCE_MAX, REDUCE_MIN, REDUCE_ARGMAX, REDUCE_ARGMIN, REDUCE_MEAN, REDUCE_ALL, REDUCE_ANY = range(8)
ACCUM_PLAIN, ACCUM_COMPENSATED = 0, 1
FUSE_SUM, FUSE_MAX, FUSE_MIN = 1, 2, 4
BOUNDARY_NONE, BOUNDARY_ZERO = 0, 1
COUNT_ALL, COUNT_NONZERO = 0, 1
LAYOUT_NC, LAYOUT_SO = 0, 1
# stream argument: the legacy default stream without synchronising (NULL/None synchronises)
STREAM_DEFAULT_ASYNC = 1

ITEMSIZE = {BOOL: 1, CHAR: 1, UINT8: 1, UINT16: 2, UINT32: 4, UINT64: 8, INT8: 1, INT16: 2,
            INT32: 4, INT64: 8, FLOAT32: 4, FLOAT64: 8, BFLOAT16: 2, FLOAT16: 2}
FLOAT_CODES = (FLOAT32, FLOAT64, BFLOAT16, FLOAT16)


def pack(op, idx=0):
    """Instruction word (op << 24) | idx (op_code.hh:174-192)."""
    return ((op & 0xFF) << 24) | (idx & 0xFFFFFF)


class View(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dtype", ctypes.c_int32), ("ndim", ctypes.c_int32),
                ("shape", ctypes.c_int64 * MAX_NDIM), ("strides", ctypes.c_int64 * MAX_NDIM),
                ("pointer_axis", ctypes.c_int32), ("layout", ctypes.c_int32),
                ("ptr_offset", ctypes.c_int64), ("row_offset", ctypes.c_int64),
                ("suboffsets", ctypes.c_int64 * MAX_NDIM)]


class Scalar(ctypes.Union):
    _fields_ = [("i", ctypes.c_int64), ("u", ctypes.c_uint64), ("f", ctypes.c_double)]


class ExprDesc(ctypes.Structure):
    _fields_ = [("instrs", ctypes.POINTER(ctypes.c_uint32)), ("n_instrs", ctypes.c_int32),
                ("n_arrays", ctypes.c_int32), ("arr_dtypes", ctypes.POINTER(ctypes.c_int32)),
                ("n_scalars", ctypes.c_int32), ("work_dtype", ctypes.c_int32),
                ("scalar_dtypes", ctypes.POINTER(ctypes.c_int32)), ("dest_dtype", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class StencilDesc(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("n_offsets", ctypes.c_int32),
                ("offsets", ctypes.POINTER(ctypes.c_int64)), ("src_dtype", ctypes.c_int32),
                ("n_extra", ctypes.c_int32), ("extra_dtypes", ctypes.POINTER(ctypes.c_int32)),
                ("expr", ctypes.POINTER(ExprDesc)), ("extra_params", ctypes.c_char_p),
                ("prologue", ctypes.c_char_p), ("epilogue", ctypes.c_char_p),
                ("is_pointer_axis", ctypes.POINTER(ctypes.c_uint8)), ("is_soarr", ctypes.c_int32)]


class NcaError(RuntimeError):
    """A non-zero status from the C ABI (message from nca_hip_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"[nca status {code}] {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `python ncarray_amd/build.py`")
    return ctypes.CDLL(LIB_PATH)


lib = _load()

_P = ctypes.c_void_p
_VP = ctypes.POINTER(View)
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_SZ = ctypes.c_size_t
_SP = ctypes.POINTER(Scalar)

_SIGS = {
    "nca_hip_init": [ctypes.c_int],
    "nca_hip_shutdown": [],
    "nca_hip_last_error": [ctypes.c_char_p, _SZ],
    "nca_hip_device_info": [ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I64), ctypes.c_char_p, _SZ],
    "nca_hip_malloc": [ctypes.POINTER(_P), _SZ],
    "nca_hip_free": [_P],
    "nca_hip_memcpy_h2d": [_P, _P, _SZ, _P],
    "nca_hip_memcpy_d2h": [_P, _P, _SZ, _P],
    "nca_hip_memcpy_d2d": [_P, _P, _SZ, _P],
    "nca_hip_synchronize": [_P],
    "nca_hip_rtc_compile": [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _SZ],
    "nca_hip_jit_stats": [ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64)],
    "nca_hip_profile_enable": [_I32],
    "nca_hip_profile_read": [ctypes.c_char_p, _SZ],
    "nca_hip_reduce_result_dtype": [_I32, _I32, ctypes.POINTER(_I32)],
    "nca_hip_op_result_dtype": [_I32, _I32, ctypes.POINTER(_I32)],
    "nca_hip_expr_compile": [ctypes.POINTER(ExprDesc), ctypes.POINTER(_U64)],
    "nca_hip_expr_launch": [_U64, _VP, _I32, _SP, _I32, _VP, _P],
    "nca_hip_expr_launch_reduce": [_U64, _VP, _I32, _SP, _I32, _VP, _I32, _I32, _P, _P, _P, _P],
    "nca_hip_expr_vjp": [_U64, _VP, _I32, _SP, _I32, _VP, _P, ctypes.POINTER(_I32), _VP, _P],
    "nca_hip_expr_source": [_U64, _I32, _I32, ctypes.c_char_p, _SZ],
    "nca_hip_reduce_full": [_VP, _I32, _I32, _P, _P],
    "nca_hip_reduce_axes": [_VP, ctypes.POINTER(_I64), _I32, _I32, _I32, _VP, _P],
    "nca_hip_stencil_create": [ctypes.POINTER(StencilDesc), ctypes.POINTER(_U64)],
    "nca_hip_stencil_apply": [_U64, _VP, _VP, _I32, _SP, _I32, _VP, _I32, ctypes.POINTER(_P), _I32, _P],
    "nca_hip_stencil_apply_reduce": [_U64, _VP, _VP, _I32, _SP, _I32, _VP, _I32, _VP, _I32, _I32, _P, _P, _P, _P],
    "nca_hip_stencil_apply_reduce_noalloc": [_U64, _VP, _VP, _I32, _SP, _I32, _VP, _I32, _VP, _I32, _I32, _P, _P, _P,
                                             _P],
    "nca_hip_stencil_adjoint_create": [_U64, ctypes.POINTER(_U64)],
    "nca_hip_stencil_expr_vjp": [_U64, _U64, ctypes.POINTER(_I32), _I32, _VP, _VP, _I32, _SP, _I32, _SP, _I32, _P, _VP,
                                 _VP, _I32, _P],
    "nca_hip_stencil_expr_vjp_reduce": [_U64, _U64, ctypes.POINTER(_I32), _I32, _VP, _VP, _I32, _SP, _I32, _SP, _I32,
                                        _P, _VP, _VP, _I32, _I32, _P, _P, _P, _I32, _P],
    "nca_hip_stencil_expr_train": [_U64, _U64, ctypes.POINTER(_I32), _I32, _VP, _VP, _I32, _SP, _I32, _SP, _I32,
                                   _P, _VP, _VP, _I32, _I32, _P, _P, _P, _I32, _P],
    "nca_hip_stencil_expr_train_source": [_U64, _U64, ctypes.POINTER(_I32), _I32, ctypes.POINTER(_I32), _I32, _I32,
                                          _I32, _I32, _I32, ctypes.c_char_p, _SZ],
    "nca_hip_current_device": [ctypes.POINTER(_I32)],
    "nca_hip_stencil_source": [_U64, _I32, ctypes.c_char_p, _SZ],
    "nca_hip_scatter_add": [_VP, _VP, _VP, _P],
    "nca_hip_bin": [_VP, _VP, _I32, _I32, _VP, _VP, _P],
    "nca_hip_bin_backward": [_VP, _VP, _VP, _P],
    "nca_hip_batch_plan": [_U64, _I32, _I32, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_I64), _I32, _I32,
                           _P, _P, _P, _I32, _I32, _I32, ctypes.POINTER(_U64)],
    "nca_hip_batch_launch": [_U64, _SP, _I32, _P],
    "nca_hip_batch_graph": [_U64, _SP, _I32, ctypes.POINTER(_U64)],
    "nca_hip_graph_launch": [_U64, _P],
    "nca_hip_graph_destroy": [_U64],
    "nca_hip_batch_destroy": [_U64],
    "nca_hip_batch_config": [_U64, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32)],
    "nca_hip_batch_autotune": [_U64, _SP, _I32, _I32, ctypes.POINTER(ctypes.c_float)],
}

EXPORTED = sorted(list(_SIGS) + ["nca_hip_version"])

for _name, _args in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = ctypes.c_int
lib.nca_hip_version.argtypes = []
lib.nca_hip_version.restype = ctypes.c_char_p


def last_error():
    buf = ctypes.create_string_buffer(8192)
    lib.nca_hip_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != OK:
        raise NcaError(rc, last_error())
    return rc


def call(name, *args):
    return check(getattr(lib, name)(*args))


# ---- helpers ------------------------------------------------------------------------------------
def make_view(ptr, dtype, shape, strides_bytes=None):
    """A View over device memory; strides in bytes (C-contiguous when omitted)."""
    shape = list(shape)
    nd = len(shape)
    if nd > MAX_NDIM:
        raise ValueError("too many dimensions")
    if strides_bytes is None:
        strides_bytes = [0] * nd
        acc = ITEMSIZE[dtype]
        for d in range(nd - 1, -1, -1):
            strides_bytes[d] = acc
            acc *= shape[d]
    v = View()
    v.data = ptr
    v.dtype = dtype
    v.ndim = nd
    for d in range(nd):
        v.shape[d] = int(shape[d])
        v.strides[d] = int(strides_bytes[d])
    v.pointer_axis = -1
    v.ptr_offset = 0
    v.row_offset = 0
    return v


def make_nc_view(table_ptr, dtype, shape, strides, pointer_axis, ptr_offset=0, row_offset=0):
    """A View with an NC pointer axis: `table_ptr` is a device table of row pointers;
    strides are bytes except strides[pointer_axis], which counts table entries."""
    v = make_view(table_ptr, dtype, shape, strides)
    v.pointer_axis = pointer_axis
    v.ptr_offset = ptr_offset
    v.row_offset = row_offset
    return v


def make_so_view(ptr, dtype, shape, strides, suboffsets):
    """A View in the PEP 3118 form of SOArrayPolicy (layout.hh:677-700): strides in bytes, and
    for each axis whose suboffset is >= 0 the address after its stride step is replaced by the
    pointer stored there plus that many bytes."""
    v = make_view(ptr, dtype, shape, strides)
    v.layout = LAYOUT_SO
    for d, so in enumerate(suboffsets):
        v.suboffsets[d] = int(so)
    for d in range(len(suboffsets), MAX_NDIM):
        v.suboffsets[d] = -1
    return v


def view_array(views):
    arr = (View * max(1, len(views)))()
    for i, v in enumerate(views):
        arr[i] = v
    return arr


def scalar_array(values, dtypes):
    arr = (Scalar * max(1, len(values)))()
    for i, (v, dt) in enumerate(zip(values, dtypes)):
        if dt in FLOAT_CODES:
            arr[i].f = float(v)
        elif dt in (UINT8, UINT16, UINT32, UINT64, BOOL):
            arr[i].u = int(v) & 0xFFFFFFFFFFFFFFFF
        else:
            arr[i].i = int(v)
    return arr


class Tape:
    """An expression tape (instructions + operand dtypes) compiled once per signature."""

    def __init__(self, instrs, arr_dtypes, scalar_dtypes, work_dtype, dest_dtype):
        self.instrs = list(instrs)
        self.arr_dtypes = list(arr_dtypes)
        self.scalar_dtypes = list(scalar_dtypes)
        self.work_dtype = work_dtype
        self.dest_dtype = dest_dtype
        self._keep = []
        self.desc = self._desc()
        h = _U64()
        call("nca_hip_expr_compile", ctypes.byref(self.desc), ctypes.byref(h))
        self.handle = h.value

    def _desc(self):
        ins = (ctypes.c_uint32 * max(1, len(self.instrs)))(*self.instrs)
        ad = (ctypes.c_int32 * max(1, len(self.arr_dtypes)))(*self.arr_dtypes)
        sd = (ctypes.c_int32 * max(1, len(self.scalar_dtypes)))(*self.scalar_dtypes)
        self._keep += [ins, ad, sd]
        d = ExprDesc()
        d.instrs = ctypes.cast(ins, ctypes.POINTER(ctypes.c_uint32))
        d.n_instrs = len(self.instrs)
        d.n_arrays = len(self.arr_dtypes)
        d.arr_dtypes = ctypes.cast(ad, ctypes.POINTER(ctypes.c_int32))
        d.n_scalars = len(self.scalar_dtypes)
        d.work_dtype = self.work_dtype
        d.scalar_dtypes = ctypes.cast(sd, ctypes.POINTER(ctypes.c_int32))
        d.dest_dtype = self.dest_dtype
        return d

    def source(self, variant=0, ndim=1):
        n = lib.nca_hip_expr_source(self.handle, variant, ndim, None, 0)
        if n < 0:
            check(-n)
        buf = ctypes.create_string_buffer(n + 1)
        lib.nca_hip_expr_source(self.handle, variant, ndim, buf, n + 1)
        return buf.value.decode()

    def launch(self, ins, out, scalars=(), stream=None):
        va = view_array(ins)
        sa = scalar_array(scalars, self.scalar_dtypes)
        call("nca_hip_expr_launch", self.handle, va, len(ins), sa, len(scalars), ctypes.byref(out), stream)

    def launch_reduce(self, ins, out, fuse, mode=ACCUM_PLAIN, sum_out=None, max_out=None, min_out=None,
                      scalars=(), stream=None):
        va = view_array(ins)
        sa = scalar_array(scalars, self.scalar_dtypes)
        call("nca_hip_expr_launch_reduce", self.handle, va, len(ins), sa, len(scalars), ctypes.byref(out), fuse,
             mode, sum_out, max_out, min_out, stream)

    def vjp(self, ins, grads, wrt, grad_out=None, grad_scalar=None, scalars=(), stream=None):
        va = view_array(ins)
        ga = view_array(grads)
        wa = (ctypes.c_int32 * max(1, len(wrt)))(*[1 if w else 0 for w in wrt])
        sa = scalar_array(scalars, self.scalar_dtypes)
        call("nca_hip_expr_vjp", self.handle, va, len(ins), sa, len(scalars),
             ctypes.byref(grad_out) if grad_out is not None else None, grad_scalar, wa, ga, stream)


class Stencil:
    """Stencil<NDIM>: offsets + a tape over the offset loads (and extra pointwise arrays)."""

    def __init__(self, offsets, tape_instrs, src_dtype, work_dtype, dest_dtype, scalar_dtypes=(),
                 extra_dtypes=(), extra_params=None, prologue=None, epilogue=None, handle=None,
                 is_pointer_axis=None, is_soarr=False):
        self.offsets = [list(o) for o in offsets]
        self.ndim = len(self.offsets[0])
        self.src_dtype = src_dtype
        self.extra_dtypes = list(extra_dtypes)
        self.scalar_dtypes = list(scalar_dtypes)
        self.dest_dtype = dest_dtype
        if handle is not None:
            self.handle = handle
            return
        arr_dtypes = [src_dtype] * len(self.offsets) + self.extra_dtypes
        self._tape_desc = Tape.__new__(Tape)
        self._tape_desc.instrs = list(tape_instrs)
        self._tape_desc.arr_dtypes = arr_dtypes
        self._tape_desc.scalar_dtypes = list(scalar_dtypes)
        self._tape_desc.work_dtype = work_dtype
        self._tape_desc.dest_dtype = dest_dtype
        self._tape_desc._keep = []
        ed = self._tape_desc._desc()
        self._ed = ed
        flat = [int(x) for o in self.offsets for x in o]
        offs = (ctypes.c_int64 * len(flat))(*flat)
        exd = (ctypes.c_int32 * max(1, len(self.extra_dtypes)))(*self.extra_dtypes)
        self._keep = [offs, exd]
        sd = StencilDesc()
        sd.ndim = self.ndim
        sd.n_offsets = len(self.offsets)
        sd.offsets = ctypes.cast(offs, ctypes.POINTER(ctypes.c_int64))
        sd.src_dtype = src_dtype
        sd.n_extra = len(self.extra_dtypes)
        sd.extra_dtypes = ctypes.cast(exd, ctypes.POINTER(ctypes.c_int32))
        sd.expr = ctypes.pointer(ed)
        sd.extra_params = extra_params.encode() if extra_params else None
        sd.prologue = prologue.encode() if prologue else None
        sd.epilogue = epilogue.encode() if epilogue else None
        if is_pointer_axis is not None:
            pa = (ctypes.c_uint8 * self.ndim)(*[1 if f else 0 for f in is_pointer_axis])
            self._keep.append(pa)
            sd.is_pointer_axis = ctypes.cast(pa, ctypes.POINTER(ctypes.c_uint8))
        sd.is_soarr = 1 if is_soarr else 0
        h = _U64()
        call("nca_hip_stencil_create", ctypes.byref(sd), ctypes.byref(h))
        self.handle = h.value

    def adjoint(self):
        h = _U64()
        call("nca_hip_stencil_adjoint_create", self.handle, ctypes.byref(h))
        return Stencil([[-x for x in o] for o in self.offsets], None, self.dest_dtype, None, self.src_dtype,
                       self.scalar_dtypes, handle=h.value)

    def expr_vjp(self, tape, operands, src, w, extras=(), grad_extras=(), scalars=(), expr_scalars=(),
                 grad_scalar=None, boundary=BOUNDARY_ZERO, stream=None, fuse=0, mode=ACCUM_PLAIN, sum_out=None,
                 max_out=None, min_out=None):
        """Backward of sum(tape(stencil(src), extras)): writes w = d tape / d a * g and the extras'
        gradients (grad_extras[j] None: not computed) in one pass; `operands[k]` is -1 when tape
        operand k reads the stencil value, j when it reads extras[j]; g = *grad_scalar (device
        pointer, None: 1).  fuse != 0 also reduces the tape's value in the same pass into the
        device scalars sum_out / max_out / min_out."""
        ops = (_I32 * max(1, len(operands)))(*operands)
        ea = view_array(extras)
        ge = [v if v is not None else View() for v in grad_extras] + [View()] * (len(extras) - len(grad_extras))
        ga = view_array(ge)
        sa = scalar_array(scalars, self.scalar_dtypes)
        xa = scalar_array(expr_scalars, tape.scalar_dtypes)
        if fuse:
            call("nca_hip_stencil_expr_vjp_reduce", self.handle, tape.handle, ops, len(operands), ctypes.byref(src),
                 ea, len(extras), sa, len(scalars), xa, len(expr_scalars), grad_scalar, ctypes.byref(w), ga, fuse, mode,
                 sum_out, max_out, min_out, boundary, stream)
        else:
            call("nca_hip_stencil_expr_vjp", self.handle, tape.handle, ops, len(operands), ctypes.byref(src), ea,
                 len(extras), sa, len(scalars), xa, len(expr_scalars), grad_scalar, ctypes.byref(w), ga, boundary,
                 stream)

    def expr_train(self, tape, operands, src, grad_src, extras=(), grad_extras=(), scalars=(), expr_scalars=(),
                   grad_scalar=None, boundary=BOUNDARY_ZERO, stream=None, fuse=0, mode=ACCUM_PLAIN, sum_out=None,
                   max_out=None, min_out=None):
        """One pass of a training step of sum(tape(stencil(src), extras)): grad_src = the stencil's
        adjoint of d tape / d a * g, the extras' gradients (grad_extras[j] None: not computed) and
        with fuse != 0 the reductions of the tape's value into the device scalars sum_out /
        max_out / min_out; `operands` and g as in expr_vjp."""
        ops = (_I32 * max(1, len(operands)))(*operands)
        ea = view_array(extras)
        ge = [v if v is not None else View() for v in grad_extras] + [View()] * (len(extras) - len(grad_extras))
        ga = view_array(ge)
        sa = scalar_array(scalars, self.scalar_dtypes)
        xa = scalar_array(expr_scalars, tape.scalar_dtypes)
        call("nca_hip_stencil_expr_train", self.handle, tape.handle, ops, len(operands), ctypes.byref(src), ea,
             len(extras), sa, len(scalars), xa, len(expr_scalars), grad_scalar, ctypes.byref(grad_src), ga, fuse, mode,
             sum_out, max_out, min_out, boundary, stream)

    def train_source(self, tape, operands, extra_dtypes, grad_src_dtype, fuse=0, mode=ACCUM_PLAIN, vec=False):
        ops = (_I32 * max(1, len(operands)))(*operands)
        xd = (_I32 * max(1, len(extra_dtypes)))(*extra_dtypes)
        args = (self.handle, tape.handle, ops, len(operands), xd, len(extra_dtypes), grad_src_dtype, fuse, mode,
                1 if vec else 0)
        n = lib.nca_hip_stencil_expr_train_source(*args, None, 0)
        if n < 0:
            check(-n)
        buf = ctypes.create_string_buffer(n + 1)
        lib.nca_hip_stencil_expr_train_source(*args, buf, n + 1)
        return buf.value.decode()

    def source(self, variant=0):
        n = lib.nca_hip_stencil_source(self.handle, variant, None, 0)
        if n < 0:
            check(-n)
        buf = ctypes.create_string_buffer(n + 1)
        lib.nca_hip_stencil_source(self.handle, variant, buf, n + 1)
        return buf.value.decode()

    def apply(self, src, dst, boundary=BOUNDARY_ZERO, extras=(), scalars=(), ext_args=(), stream=None):
        ea = view_array(extras)
        sa = scalar_array(scalars, self.scalar_dtypes)
        xa = (ctypes.c_void_p * max(1, len(ext_args)))(*[ctypes.cast(a, ctypes.c_void_p) for a in ext_args])
        call("nca_hip_stencil_apply", self.handle, ctypes.byref(src), ea, len(extras), sa, len(scalars),
             ctypes.byref(dst), boundary, xa, len(ext_args), stream)

    def apply_reduce(self, src, dst, boundary=BOUNDARY_ZERO, axis0=None, fuse=0, mode=ACCUM_PLAIN, sum_out=None,
                     max_out=None, min_out=None, extras=(), scalars=(), stream=None):
        ea = view_array(extras)
        sa = scalar_array(scalars, self.scalar_dtypes)
        call("nca_hip_stencil_apply_reduce", self.handle, ctypes.byref(src), ea, len(extras), sa, len(scalars),
             ctypes.byref(dst), boundary, ctypes.byref(axis0) if axis0 is not None else None, fuse, mode, sum_out,
             max_out, min_out, stream)


def rtc_compile(source, arch="gfx950"):
    """Compile a full generated source with hipRTC (no GPU needed); returns (ok, log)."""
    log = ctypes.create_string_buffer(1 << 16)
    rc = lib.nca_hip_rtc_compile(source.encode(), arch.encode(), log, len(log))
    return rc == OK, log.value.decode(errors="replace")


_initialised = {}


def init(device=0):
    if not _initialised.get(device):
        call("nca_hip_init", device)
        _initialised[device] = True


def jit_stats():
    """(compiles, disk_loads, memory_hits) of the kernel cache in this process."""
    c, d, m = _I64(), _I64(), _I64()
    call("nca_hip_jit_stats", ctypes.byref(c), ctypes.byref(d), ctypes.byref(m))
    return c.value, d.value, m.value


def profile_enable(on=True):
    call("nca_hip_profile_enable", 1 if on else 0)


def profile_read():
    """[(kernel name, ms)] for every library kernel launched since the last read."""
    n = lib.nca_hip_profile_read(None, 0)
    if n < 0:
        check(-n)
    buf = ctypes.create_string_buffer(n + 1)
    n2 = lib.nca_hip_profile_read(buf, len(buf))
    if n2 < 0:
        check(-n2)
    out = []
    for line in buf.value.decode().splitlines():
        name, ms = line.rsplit(" ", 1)
        out.append((name, float(ms)))
    return out


def reduce_result_dtype(op, src_dtype):
    out = _I32()
    call("nca_hip_reduce_result_dtype", op, src_dtype, ctypes.byref(out))
    return out.value


def op_result_dtype(op, work_dtype):
    out = _I32()
    call("nca_hip_op_result_dtype", op, work_dtype, ctypes.byref(out))
    return out.value
