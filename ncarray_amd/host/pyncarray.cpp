// Python bindings of the C++ host API (module ncarray_amd._ncarray).
//
// Mirrors the reference's device-array bindings: NCDevArray and DevExpr with the same
// dunders, reductions, indexing and host transfer methods (pyncarray/binding_builder.hh:
// 162-815, ncdevarray_owner.cu, ncdevarray.cu), DType/OpCode enums, set_eager/is_eager
// and materialize (pyncarray/ncarray.cc:34-46, eager.cc).  In eager mode (the default) a
// binary operator on arrays evaluates at once into a new array (one kernel per operator,
// as in the reference); in lazy mode it returns a DevExpr whose whole tape is evaluated
// by one fused kernel when materialised or assigned.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <limits>

#include "ncarray_amd.hh"

namespace py = pybind11;
using namespace ncarray_amd;

namespace {

bool g_eager = true;
// The host array type that to_host() returns (a numpy subclass).  Held as a raw reference for
// the life of the process: a static py::object would be released after interpreter shutdown.
PyObject* g_host_type = nullptr;

// ---- dtype conversions ------------------------------------------------------------------------
DType from_numpy(const py::dtype& dt) {
  const char kind = dt.kind();
  const ssize_t sz = dt.itemsize();
  if (kind == 'b') return DType::bool_;
  if (kind == 'u') {
    if (sz == 1) return DType::uint8;
    if (sz == 2) return DType::uint16;
    if (sz == 4) return DType::uint32;
    if (sz == 8) return DType::uint64;
  }
  if (kind == 'i') {
    if (sz == 1) return DType::int8;
    if (sz == 2) return DType::int16;
    if (sz == 4) return DType::int32;
    if (sz == 8) return DType::int64;
  }
  if (kind == 'f') {
    if (sz == 2) return DType::float16;
    if (sz == 4) return DType::float32;
    if (sz == 8) return DType::float64;
  }
  throw py::type_error("numpy dtype " + std::string(py::str(dt)) + " has no device counterpart");
}

py::dtype to_numpy(DType d) {
  switch (d) {
    case DType::bool_: return py::dtype("?");
    case DType::char_: return py::dtype("i1");
    case DType::uint8: return py::dtype("u1");
    case DType::uint16: return py::dtype("u2");
    case DType::uint32: return py::dtype("u4");
    case DType::uint64: return py::dtype("u8");
    case DType::int8: return py::dtype("i1");
    case DType::int16: return py::dtype("i2");
    case DType::int32: return py::dtype("i4");
    case DType::int64: return py::dtype("i8");
    case DType::float16: return py::dtype("f2");
    case DType::float32: return py::dtype("f4");
    case DType::float64: return py::dtype("f8");
    case DType::bfloat16: return py::dtype("f4");  // numpy has no bfloat16: widened on transfer
    default: throw py::type_error(std::string("dtype ") + dtype_name(d) + " has no numpy counterpart");
  }
}

// Host dtype used to transfer an array of dtype d (bfloat16 widens to float32).
DType host_transfer_dtype(DType d) { return d == DType::bfloat16 ? DType::float32 : d; }

const char* typestr(DType d) {
  switch (d) {
    case DType::bool_: return "|b1";
    case DType::char_: case DType::int8: return "|i1";
    case DType::uint8: return "|u1";
    case DType::uint16: return "<u2";
    case DType::uint32: return "<u4";
    case DType::uint64: return "<u8";
    case DType::int16: return "<i2";
    case DType::int32: return "<i4";
    case DType::int64: return "<i8";
    case DType::float16: return "<f2";
    case DType::float32: return "<f4";
    case DType::float64: return "<f8";
    default: throw py::attribute_error(std::string("dtype ") + dtype_name(d) + " has no array-interface type string");
  }
}

DType from_typestr(const std::string& t) {
  if (t.size() < 3) throw py::type_error("bad typestr " + t);
  return from_numpy(py::dtype(t));
}

// ---- scalars -----------------------------------------------------------------------------------
// A Python scalar binds to the first alternative of the reference's Scalar variant that
// accepts it (pybind11's variant caster, dtype.hh:265-273): bool -> bool, float -> float
// (float32), int -> the first of uint8..uint64, int8..int64 that holds the value.  numpy
// scalars keep their own dtype, except numpy.float64, which is a Python float.
bool is_scalar(const py::handle& o) {
  return py::isinstance<py::bool_>(o) || py::isinstance<py::int_>(o) || py::isinstance<py::float_>(o) ||
         py::isinstance(o, py::module_::import("numpy").attr("generic"));
}

Scalar to_scalar(const py::handle& o) {
  if (py::isinstance<py::bool_>(o)) return Scalar(o.cast<bool>());
  if (py::isinstance<py::float_>(o)) return Scalar(static_cast<float>(o.cast<double>()));
  if (py::isinstance<py::int_>(o)) {
    // non-negative values: first unsigned type that holds them
    try {
      const unsigned long long u = o.cast<unsigned long long>();
      if (u <= std::numeric_limits<uint8_t>::max()) return Scalar(static_cast<uint8_t>(u));
      if (u <= std::numeric_limits<uint16_t>::max()) return Scalar(static_cast<uint16_t>(u));
      if (u <= std::numeric_limits<uint32_t>::max()) return Scalar(static_cast<uint32_t>(u));
      return Scalar(static_cast<uint64_t>(u));
    } catch (const py::cast_error&) {
    }
    const long long i = o.cast<long long>();
    if (i >= std::numeric_limits<int8_t>::min()) return Scalar(static_cast<int8_t>(i));
    if (i >= std::numeric_limits<int16_t>::min()) return Scalar(static_cast<int16_t>(i));
    if (i >= std::numeric_limits<int32_t>::min()) return Scalar(static_cast<int32_t>(i));
    return Scalar(static_cast<int64_t>(i));
  }
  py::module_ np = py::module_::import("numpy");
  if (py::isinstance(o, np.attr("generic"))) {
    py::array a = np.attr("asarray")(o);
    const DType d = from_numpy(a.dtype());
    return scalar_from_bytes(d == DType::float16 ? DType::float16 : d, a.data());
  }
  throw py::type_error("not a scalar");
}

py::object scalar_to_py(const Scalar& s) {
  return std::visit(
      [](auto v) -> py::object {
        using T = decltype(v);
        if constexpr (std::is_same_v<T, bool>) return py::bool_(v);
        else if constexpr (std::is_floating_point_v<T>) return py::float_(static_cast<double>(v));
        else if constexpr (std::is_unsigned_v<T>) return py::int_(static_cast<unsigned long long>(v));
        else return py::int_(static_cast<long long>(v));
      },
      s);
}

// ---- host transfers ------------------------------------------------------------------------------
NCDevArray from_host(const py::array& arr, std::optional<DType> dtype) {
  const DType hd = from_numpy(arr.dtype());
  std::vector<int64_t> shape(arr.shape(), arr.shape() + arr.ndim());
  std::vector<int64_t> strides(arr.strides(), arr.strides() + arr.ndim());
  NCDevArray d(shape, dtype.value_or(hd));
  d.assign_from_host(arr.data(), hd, shape, strides);
  return d;
}

py::array to_numpy_array(const NCDevArray& a, std::optional<DType> dtype) {
  const DType hd = host_transfer_dtype(dtype.value_or(a.dtype()));
  std::vector<ssize_t> shape(a.shape(), a.shape() + a.ndim());
  py::array out(to_numpy(hd), shape);
  a.copy_into_astype(out.mutable_data(), hd);
  return out;
}

py::object host_array(const NCDevArray& a, std::optional<DType> dtype) {
  py::array out = to_numpy_array(a, dtype);
  if (g_host_type) return out.attr("view")(py::handle(g_host_type));
  return out;
}

// ---- operands ------------------------------------------------------------------------------------
// Resolve a Python operand into something an Expr can hold.  A numpy array is uploaded to a
// temporary device array (kept alive by the expression).
struct PyOperand {
  std::optional<NCDevArray> arr;
  std::optional<Expr> expr;
  std::optional<Scalar> sc;
  Operand op() const {
    if (arr) return Operand(*arr);
    if (expr) return Operand(*expr);
    return Operand(*sc);
  }
};

PyOperand resolve(const py::handle& o) {
  PyOperand r;
  if (py::isinstance<NCDevArray>(o)) r.arr = o.cast<NCDevArray>();
  else if (py::isinstance<Expr>(o)) r.expr = o.cast<Expr>();
  else if (py::isinstance<py::array>(o)) r.arr = from_host(o.cast<py::array>(), std::nullopt);
  else if (is_scalar(o)) r.sc = to_scalar(o);
  else if (py::hasattr(o, "__cuda_array_interface__")) {
    // torch exports bfloat16 with the bare typestr "|V2": name the dtype for it
    py::object as = py::none();
    if (py::hasattr(o, "dtype") && std::string(py::str(o.attr("dtype"))) == "torch.bfloat16") as = py::cast(DType::bfloat16);
    r.arr = py::module_::import("ncarray_amd._ncarray").attr("NCDevArray").attr("from_cuda_array")(o, as).cast<NCDevArray>();
  }
  else throw py::type_error("unsupported operand type " + std::string(py::str(py::type::of(o))));
  return r;
}

py::object finish(Expr&& e) {
  if (g_eager) return py::cast(e.materialize());
  return py::cast(std::move(e));
}

template <typename Self>
py::object binop(const Self& self, const py::object& other, OpCode op, bool reflected) {
  PyOperand o = resolve(other);
  Expr e;
  if (reflected) {
    e.build_node(o.op());
    e.binary(op, Operand(self));
  } else {
    e.build_node(Operand(self));
    e.binary(op, o.op());
  }
  return finish(std::move(e));
}

template <typename Self>
py::object unop(const Self& self, OpCode op) {
  Expr e;
  e.build_node(Operand(self));
  e.unary(op);
  return finish(std::move(e));
}

template <typename Cls, typename Self>
void bind_operators(Cls& cls) {
  struct B {
    const char* name;
    OpCode op;
    bool reflected;
  };
  static const B table[] = {
      {"__add__", OpCode::ADD, false},     {"__radd__", OpCode::ADD, true},   {"__sub__", OpCode::SUB, false},
      {"__rsub__", OpCode::SUB, true},     {"__mul__", OpCode::MUL, false},   {"__rmul__", OpCode::MUL, true},
      {"__truediv__", OpCode::DIV, false}, {"__rtruediv__", OpCode::DIV, true}, {"__mod__", OpCode::MOD, false},
      {"__rmod__", OpCode::MOD, true},     {"__floordiv__", OpCode::FDIV, false}, {"__rfloordiv__", OpCode::FDIV, true},
      {"__eq__", OpCode::EQ, false},       {"__ne__", OpCode::NE, false},     {"__lt__", OpCode::LT, false},
      {"__le__", OpCode::LE, false},       {"__gt__", OpCode::GT, false},     {"__ge__", OpCode::GE, false},
      {"__and__", OpCode::LAND, false},    {"__rand__", OpCode::LAND, true},  {"__or__", OpCode::LOR, false},
      {"__ror__", OpCode::LOR, true},      {"__xor__", OpCode::XOR, false},   {"__rxor__", OpCode::XOR, true},
      {"__lshift__", OpCode::LSHFT, false}, {"__rshift__", OpCode::RSHFT, false},
  };
  for (const B& b : table) {
    const OpCode op = b.op;
    const bool refl = b.reflected;
    cls.def(b.name, [op, refl](const Self& self, const py::object& other) { return binop(self, other, op, refl); },
            py::is_operator());
  }
  cls.def("__neg__", [](const Self& self) { return unop(self, OpCode::NEG); });
  cls.def("__abs__", [](const Self& self) { return unop(self, OpCode::ABS); });
  cls.def("__invert__", [](const Self& self) {
    const DType d = self.dtype();
    return unop(self, d == DType::bool_ ? OpCode::LNOT : OpCode::BNOT);
  });
  cls.def("__pos__", [](const Self& self) { return py::cast(self); });
}

// ---- indexing ------------------------------------------------------------------------------------
std::vector<IndexItem> to_indices(const py::handle& idx) {
  std::vector<IndexItem> out;
  auto one = [&out](const py::handle& it) {
    if (py::isinstance<py::int_>(it) || py::isinstance(it, py::module_::import("numpy").attr("integer")))
      out.emplace_back(static_cast<int64_t>(it.cast<long long>()));
    else if (py::isinstance<py::slice>(it)) {
      py::object s = py::reinterpret_borrow<py::object>(it);
      Slice sl;
      py::object st = s.attr("start"), sp = s.attr("stop"), stp = s.attr("step");
      if (!st.is_none()) sl.start = st.cast<int64_t>();
      if (!sp.is_none()) sl.stop = sp.cast<int64_t>();
      if (!stp.is_none()) sl.step = stp.cast<int64_t>();
      out.emplace_back(sl);
    } else if (it.is(py::ellipsis())) {
      out.emplace_back(Ellipsis{});
    } else if (it.is_none()) {
      out.emplace_back(NewAxis{});
    } else {
      throw py::type_error("Invalid indexing argument!");
    }
  };
  if (py::isinstance<py::tuple>(idx)) {
    for (const py::handle& it : idx.cast<py::tuple>()) one(it);
  } else {
    one(idx);
  }
  return out;
}

void set_into(const NCDevArray& view, const py::object& val) {
  NCDevArray v = view;
  if (py::isinstance<NCDevArray>(val)) v.assign(val.cast<NCDevArray>());
  else if (py::isinstance<Expr>(val)) v = val.cast<Expr>();
  else if (py::isinstance<py::array>(val)) {
    py::array a = val.cast<py::array>();
    std::vector<int64_t> shape(a.shape(), a.shape() + a.ndim());
    std::vector<int64_t> strides(a.strides(), a.strides() + a.ndim());
    v.assign_from_host(a.data(), from_numpy(a.dtype()), shape, strides);
  } else if (is_scalar(val)) {
    v.fill(to_scalar(val));
  } else if (py::isinstance<py::list>(val) || py::isinstance<py::tuple>(val)) {
    set_into(view, py::module_::import("numpy").attr("asarray")(val));
  } else {
    throw py::type_error("Unrecognized type for assignment!");
  }
}

std::vector<int64_t> to_axes(const py::handle& axis) {
  std::vector<int64_t> axes;
  if (py::isinstance<py::int_>(axis)) axes.push_back(axis.cast<int64_t>());
  else
    for (const py::handle& a : axis) axes.push_back(a.cast<int64_t>());
  return axes;
}

py::tuple to_tuple(const int64_t* p, int n) {
  py::tuple t(n);
  for (int i = 0; i < n; ++i) t[i] = py::int_(p[i]);
  return t;
}

// Python stencil: a handle plus its constants (Stencil<NDIM> with NDIM at run time).
struct PyStencil {
  uint64_t handle = 0;
  std::vector<Scalar> scalars;
  int ndim = 0;
  DType dest_dtype = DType::float32;
};

// ---- DLPack --------------------------------------------------------------------------------
// The array-interchange protocol of torch, numpy and others: a PyCapsule named "dltensor" holds a
// DLManagedTensor (the producer's deleter frees it); a consumer renames the capsule
// "used_dltensor" and calls the deleter when it no longer needs the memory.  The C structs are
// the protocol's ABI (version 0.x "DLManagedTensor"), restated here.
namespace dl {
struct Device { int32_t device_type; int32_t device_id; };
struct DataType { uint8_t code; uint8_t bits; uint16_t lanes; };
struct Tensor {
  void* data;
  Device device;
  int32_t ndim;
  DataType dtype;
  int64_t* shape;
  int64_t* strides;  // in elements; NULL: C-contiguous
  uint64_t byte_offset;
};
struct Managed {
  Tensor dl_tensor;
  void* manager_ctx;
  void (*deleter)(Managed*);
};
enum { kInt = 0, kUInt = 1, kFloat = 2, kBfloat = 4, kBool = 6 };
enum { kCUDA = 2, kROCM = 10 };

DataType of(DType d) {
  const uint8_t bits = static_cast<uint8_t>(8 * itemsize(d));
  switch (d) {
    case DType::bool_: return {kBool, 8, 1};
    case DType::char_: case DType::int8: case DType::int16: case DType::int32: case DType::int64: return {kInt, bits, 1};
    case DType::uint8: case DType::uint16: case DType::uint32: case DType::uint64: return {kUInt, bits, 1};
    case DType::float32: case DType::float64: case DType::float16: return {kFloat, bits, 1};
    case DType::bfloat16: return {kBfloat, 16, 1};
    default: throw Error(NCA_ERR_UNSUPPORTED, std::string("no DLPack type for ") + dtype_name(d));
  }
}

DType to_dtype(DataType t) {
  if (t.lanes != 1) throw Error(NCA_ERR_UNSUPPORTED, "vector DLPack types are not supported");
  switch (t.code) {
    case kBool: if (t.bits == 8) return DType::bool_; break;
    case kInt:
      if (t.bits == 8) return DType::int8;
      if (t.bits == 16) return DType::int16;
      if (t.bits == 32) return DType::int32;
      if (t.bits == 64) return DType::int64;
      break;
    case kUInt:
      if (t.bits == 8) return DType::uint8;
      if (t.bits == 16) return DType::uint16;
      if (t.bits == 32) return DType::uint32;
      if (t.bits == 64) return DType::uint64;
      break;
    case kFloat:
      if (t.bits == 16) return DType::float16;
      if (t.bits == 32) return DType::float32;
      if (t.bits == 64) return DType::float64;
      break;
    case kBfloat: if (t.bits == 16) return DType::bfloat16; break;
    default: break;
  }
  throw Error(NCA_ERR_UNSUPPORTED, "unsupported DLPack type (code " + std::to_string(t.code) + ", " +
                                       std::to_string(t.bits) + " bits)");
}

// What an exported capsule owns: the array (keeping its memory alive) and the shape/strides
// arrays the DLTensor points at.
struct Export {
  Managed managed{};
  NCDevArray array;
  std::vector<int64_t> shape, strides;
};

void delete_export(Managed* m) { delete static_cast<Export*>(m->manager_ctx); }

void capsule_destructor(PyObject* cap) {
  // a capsule no consumer took still owns its tensor
  if (PyCapsule_IsValid(cap, "dltensor")) {
    auto* m = static_cast<Managed*>(PyCapsule_GetPointer(cap, "dltensor"));
    if (m && m->deleter) m->deleter(m);
  } else {
    PyErr_Clear();
  }
}

py::capsule to_capsule(const NCDevArray& a) {
  if (a.pointer_axis() >= 0)
    throw Error(NCA_ERR_UNSUPPORTED, "a pointer-axis array has no DLPack form (use dense())");
  auto* e = new Export();
  e->array = a;
  const int64_t item = a.itemsize();
  for (int d = 0; d < a.ndim(); ++d) {
    e->shape.push_back(a.shape(d));
    if (a.strides()[d] % item != 0) {
      delete e;
      throw Error(NCA_ERR_UNSUPPORTED, "DLPack strides count elements; this view's byte strides do not");
    }
    e->strides.push_back(a.strides()[d] / item);
  }
  int32_t dev = 0;
  (void)nca_hip_current_device(&dev);
  Tensor& t = e->managed.dl_tensor;
  t.data = const_cast<void*>(a.data());
  t.device = {kROCM, dev};
  t.ndim = a.ndim();
  t.dtype = of(a.dtype());
  t.shape = e->shape.empty() ? nullptr : e->shape.data();
  t.strides = e->strides.empty() ? nullptr : e->strides.data();
  t.byte_offset = 0;
  e->managed.manager_ctx = e;
  e->managed.deleter = delete_export;
  return py::capsule(&e->managed, "dltensor", &capsule_destructor);
}

NCDevArray from_capsule(py::capsule cap) {
  PyObject* o = cap.ptr();
  if (!PyCapsule_IsValid(o, "dltensor")) throw py::value_error("not an unconsumed DLPack capsule (name \"dltensor\")");
  auto* m = static_cast<Managed*>(PyCapsule_GetPointer(o, "dltensor"));
  const Tensor& t = m->dl_tensor;
  if (t.device.device_type != kROCM && t.device.device_type != kCUDA)
    throw Error(NCA_ERR_UNSUPPORTED, "DLPack tensor is not in GPU memory (device type " +
                                         std::to_string(t.device.device_type) + ")");
  const DType d = to_dtype(t.dtype);
  std::vector<int64_t> shape(t.shape, t.shape + t.ndim), strides(t.ndim);
  int64_t s = itemsize(d);
  for (int k = t.ndim - 1; k >= 0; --k) {
    strides[k] = t.strides ? t.strides[k] * itemsize(d) : s;
    s *= shape[k];
  }
  // the consumer owns it now: rename the capsule and call the producer's deleter when the last
  // view of the memory is gone
  PyCapsule_SetName(o, "used_dltensor");
  std::shared_ptr<void> keep(m, [](void* p) {
    auto* mm = static_cast<Managed*>(p);
    if (mm->deleter) {
      py::gil_scoped_acquire g;
      mm->deleter(mm);
    }
  });
  return NCDevArray::wrap(static_cast<char*>(t.data) + t.byte_offset, shape, strides, d, keep);
}
}  // namespace dl

}  // namespace

PYBIND11_MODULE(_ncarray, m) {
  m.doc() = "MI355X device arrays: fused expressions, stencils, reductions (ncarray device surface)";

  py::enum_<DType>(m, "DType", "Type enumerators (reference order; bfloat16/float16 are storage types)")
      .value("bool", DType::bool_)
      .value("char", DType::char_)
      .value("uint8", DType::uint8)
      .value("uint16", DType::uint16)
      .value("uint32", DType::uint32)
      .value("uint64", DType::uint64)
      .value("int8", DType::int8)
      .value("int16", DType::int16)
      .value("int32", DType::int32)
      .value("int64", DType::int64)
      .value("float32", DType::float32)
      .value("float64", DType::float64)
      .value("complex64", DType::complex64)
      .value("complex128", DType::complex128)
      .value("complex256", DType::complex256)
      .value("vfloat2", DType::vfloat2)
      .value("vfloat3", DType::vfloat3)
      .value("vfloat4", DType::vfloat4)
      .value("vdouble2", DType::vdouble2)
      .value("vdouble3", DType::vdouble3)
      .value("vdouble4", DType::vdouble4)
      .value("bfloat16", DType::bfloat16)
      .value("float16", DType::float16);

  py::enum_<OpCode>(m, "OpCode")
      .value("NOOP", OpCode::NOOP).value("IDX", OpCode::IDX).value("LOAD_NCARR", OpCode::LOAD_NCARR)
      .value("LOAD_SOARR", OpCode::LOAD_SOARR).value("LOAD_CONST", OpCode::LOAD_CONST).value("NEG", OpCode::NEG)
      .value("INC", OpCode::INC).value("DEC", OpCode::DEC).value("SZOF", OpCode::SZOF).value("ADDR", OpCode::ADDR)
      .value("INDR", OpCode::INDR).value("CAST", OpCode::CAST).value("LNOT", OpCode::LNOT).value("BNOT", OpCode::BNOT)
      .value("ADD", OpCode::ADD).value("SUB", OpCode::SUB).value("MUL", OpCode::MUL).value("DIV", OpCode::DIV)
      .value("MOD", OpCode::MOD).value("FDIV", OpCode::FDIV).value("EQ", OpCode::EQ).value("NE", OpCode::NE)
      .value("LT", OpCode::LT).value("LE", OpCode::LE).value("GT", OpCode::GT).value("GE", OpCode::GE)
      .value("LAND", OpCode::LAND).value("LOR", OpCode::LOR).value("BAND", OpCode::BAND).value("BOR", OpCode::BOR)
      .value("XOR", OpCode::XOR).value("LSHFT", OpCode::LSHFT).value("RSHFT", OpCode::RSHFT)
      .value("EXP", OpCode::EXP).value("LOG", OpCode::LOG).value("SQRT", OpCode::SQRT).value("ABS", OpCode::ABS);

  py::enum_<Boundary>(m, "Boundary").value("none", Boundary::none).value("zero", Boundary::zero);
  py::enum_<Accum>(m, "Accum").value("plain", Accum::plain).value("compensated", Accum::compensated);

  py::register_exception<Error>(m, "NcaError", PyExc_RuntimeError);

  m.def("init", &ncarray_amd::init, py::arg("device") = 0, "Select the device for this thread.");
  m.def("set_eager", [](bool e) { g_eager = e; }, py::arg("is_eager"),
        "Toggle whether expressions are eagerly evaluated to arrays or not.");
  m.def("is_eager", []() { return g_eager; }, "Whether expressions are currently being eagerly evaluated.");
  m.def("materialize", [](const py::object& o) -> py::object {
    if (py::isinstance<Expr>(o)) return py::cast(o.cast<Expr>().materialize());
    return o;
  }, py::arg("expr"));
  m.def("_set_host_array_type", [](py::object t) {
    Py_XDECREF(g_host_type);
    g_host_type = t.release().ptr();
  });
  for (auto [name, op] : {std::pair{"exp", OpCode::EXP}, std::pair{"log", OpCode::LOG}, std::pair{"sqrt", OpCode::SQRT},
                          std::pair{"abs", OpCode::ABS}}) {
    const OpCode code = op;
    m.def(name, [code](const py::object& x) {
      PyOperand o = resolve(x);
      Expr e;
      e.build_node(o.op());
      e.unary(code);
      return finish(std::move(e));
    }, py::arg("x"));
  }

  // ---- DevExpr ------------------------------------------------------------------------------
  py::class_<Expr> ex(m, "DevExpr", "A lazy device expression (ExprMVNode<DevTag>)");
  ex.def_property_readonly("dtype", [](const Expr& e) { return e.expr_dtype; })
      .def_property_readonly("work_dtype", [](const Expr& e) { return e.work_dtype; })
      .def_property_readonly("shape", [](const Expr& e) { return to_tuple(e.final_shape.data(), e.ndim()); })
      .def_property_readonly("instrs", [](const Expr& e) { return e.instrs; })
      .def_property_readonly("ndim", &Expr::ndim)
      .def_property_readonly("size", &Expr::size)
      .def("materialize", [](const Expr& e, std::optional<DType> d) { return e.materialize(d); },
           py::arg("dtype") = py::none())
      .def("sum", [](const Expr& e) { return scalar_to_py(e.sum()); },
           "Sum of the evaluated values, fused into the evaluation (nothing is materialised).")
      .def("max", [](const Expr& e) { return scalar_to_py(e.max()); })
      .def("min", [](const Expr& e) { return scalar_to_py(e.min()); })
      .def("to_host", [](const Expr& e, std::optional<DType> d) { return host_array(e.materialize(), d); },
           py::arg("dtype") = py::none())
      .def("__array__", [](const Expr& e, const py::object&, const py::object&) { return to_numpy_array(e.materialize(), std::nullopt); },
           py::arg("dtype") = py::none(), py::arg("copy") = py::none())
      .def("__repr__", [](const Expr& e) {
        return "DevExpr(shape=" + std::string(py::str(to_tuple(e.final_shape.data(), e.ndim()))) +
               ", dtype=" + dtype_name(e.expr_dtype) + ", n_instrs=" + std::to_string(e.instrs.size()) + ")";
      });
  bind_operators<py::class_<Expr>, Expr>(ex);

  // ---- NCDevArray ---------------------------------------------------------------------------
  py::class_<NCDevArray> cls(m, "NCDevArray", "A device array (owner or view) in HBM");
  cls.def(py::init([](const std::vector<int64_t>& shape, DType dtype) { return NCDevArray(shape, dtype); }),
          py::arg("shape"), py::arg("dtype") = DType::float32)
      .def_static("from_host", [](const py::array& a, std::optional<DType> d) { return from_host(a, d); },
                  py::arg("array"), py::arg("dtype") = py::none(), "Copy a host (numpy) array to a new device array.")
      .def_static("from_cuda_array", [](const py::object& obj, std::optional<DType> as) {
        py::dict ai = obj.attr("__cuda_array_interface__");
        py::tuple shp = ai["shape"];
        const std::string ts = ai["typestr"].cast<std::string>();
        DType d;
        if (as) {
          // an explicit dtype names what a bare typestr cannot (e.g. "|V2" for bfloat16)
          if (static_cast<int>(py::dtype(ts).itemsize()) != itemsize(*as))
            throw py::type_error("typestr " + ts + " does not match the size of " + dtype_name(*as));
          d = *as;
        } else {
          if (ts.size() >= 2 && ts[1] == 'V')
            throw py::type_error("typestr " + ts + " does not name a dtype; pass dtype= (e.g. DType.bfloat16)");
          d = from_typestr(ts);
        }
        std::vector<int64_t> shape, strides;
        for (const py::handle& s : shp) shape.push_back(s.cast<int64_t>());
        if (ai.contains("strides") && !ai["strides"].is_none()) {
          for (const py::handle& s : ai["strides"]) strides.push_back(s.cast<int64_t>());
        } else {
          int64_t s = itemsize(d);
          strides.assign(shape.size(), 0);
          for (int k = static_cast<int>(shape.size()) - 1; k >= 0; --k) {
            strides[k] = s;
            s *= shape[k];
          }
        }
        py::tuple data = ai["data"];
        void* ptr = reinterpret_cast<void*>(data[0].cast<uintptr_t>());
        // keep the producer alive while any view of its memory exists
        auto* holder = new py::object(obj);
        std::shared_ptr<void> keep(holder, [](void* p) {
          py::gil_scoped_acquire g;
          delete static_cast<py::object*>(p);
        });
        return NCDevArray::wrap(ptr, shape, strides, d, keep);
      }, py::arg("obj"), py::arg("dtype") = py::none(),
                  "Zero-copy view of an object exposing __cuda_array_interface__ (e.g. a torch tensor).")
      .def_static("_placeholder", [](const std::vector<int64_t>& shape, DType d) {
        std::vector<int64_t> strides(shape.size(), 0);
        int64_t s = itemsize(d);
        for (int k = static_cast<int>(shape.size()) - 1; k >= 0; --k) {
          strides[k] = s;
          s *= shape[k];
        }
        return NCDevArray::wrap(nullptr, shape, strides, d);
      }, py::arg("shape"), py::arg("dtype") = DType::float32,
                  "An array without memory (stencil placeholders; building lazy expressions only).")
      .def("__repr__", &NCDevArray::repr)
      .def_property_readonly("shape", [](const NCDevArray& a) { return to_tuple(a.shape(), a.ndim()); })
      .def_property_readonly("strides", [](const NCDevArray& a) { return to_tuple(a.strides(), a.ndim()); })
      .def("__len__", [](const NCDevArray& a) {
        if (a.ndim() == 0) throw py::type_error("len() of unsized object");
        return a.shape(0);
      })
      .def_property_readonly("size", &NCDevArray::size)
      .def_property_readonly("ndim", &NCDevArray::ndim)
      .def_property_readonly("itemsize", &NCDevArray::itemsize)
      .def_property_readonly("nbytes", &NCDevArray::nbytes)
      .def_property_readonly("dtype", &NCDevArray::dtype)
      .def_property_readonly("is_contiguous", &NCDevArray::is_contiguous)
      .def_property_readonly("data_ptr", [](const NCDevArray& a) { return reinterpret_cast<uintptr_t>(a.data()); })
      .def_static("from_rows", &NCDevArray::from_rows, py::arg("rows"),
                  "A pointer-axis array (axis 0) over separately allocated device rows of one shape, stride "
                  "layout and dtype (NCDevArrayRef of a list); the rows are kept alive.")
      .def_property_readonly("pointer_axis", &NCDevArray::pointer_axis, "The NC pointer axis, or -1.")
      .def("dense", &NCDevArray::dense, "This array, or a C-order copy of a pointer-axis array.")
      .def("__dlpack__", [](const NCDevArray& a, py::object stream, py::object max_version, py::object dl_device,
                            py::object copy) {
        (void)stream;  // the library's calls without a stream have completed when they return
        (void)max_version;
        (void)dl_device;
        if (!copy.is_none() && copy.cast<bool>()) return dl::to_capsule(a.copy());
        return dl::to_capsule(a);
      }, py::arg("stream") = py::none(), py::arg("max_version") = py::none(), py::arg("dl_device") = py::none(),
           py::arg("copy") = py::none(), "A DLPack capsule of this array (zero-copy; the capsule keeps it alive).")
      .def("__dlpack_device__", [](const NCDevArray&) {
        int32_t dev = 0;
        (void)nca_hip_current_device(&dev);
        return py::make_tuple(static_cast<int>(dl::kROCM), dev);
      })
      .def_static("from_dlpack", [](const py::object& obj) {
        if (py::isinstance<py::capsule>(obj)) return dl::from_capsule(obj.cast<py::capsule>());
        if (!py::hasattr(obj, "__dlpack__")) throw py::type_error("object has no __dlpack__ and is not a DLPack capsule");
        // The consumer stream of the DLPack protocol: this library's calls without a stream run on
        // the legacy default stream, which is 0 for a ROCm producer (1 for a CUDA one); the
        // producer then orders its pending work on its current stream before that stream.
        int dev_type = 0;
        if (py::hasattr(obj, "__dlpack_device__")) dev_type = obj.attr("__dlpack_device__")().cast<py::tuple>()[0].cast<int>();
        py::object cap = dev_type == static_cast<int>(dl::kROCM) ? obj.attr("__dlpack__")(py::arg("stream") = 0)
                                                                 : obj.attr("__dlpack__")();
        return dl::from_capsule(cap.cast<py::capsule>());
      }, py::arg("obj"), "Zero-copy view of a DLPack producer (an object with __dlpack__, e.g. a torch tensor, or a "
                         "capsule); the producer's memory stays alive while the view does, and a ROCm producer's "
                         "pending work on its current stream is ordered before the library's default stream.")
      .def_property_readonly("__cuda_array_interface__", [](const NCDevArray& a) {
        if (a.pointer_axis() >= 0)
          throw Error(NCA_ERR_UNSUPPORTED, "a pointer-axis array has no __cuda_array_interface__ (use dense())");
        py::dict d;
        d["shape"] = to_tuple(a.shape(), a.ndim());
        d["typestr"] = typestr(a.dtype());
        d["data"] = py::make_tuple(reinterpret_cast<uintptr_t>(a.data()), false);
        d["strides"] = a.is_contiguous() ? py::object(py::none()) : py::object(to_tuple(a.strides(), a.ndim()));
        d["version"] = 3;
        return d;
      })
      .def("squeeze", &NCDevArray::squeeze, "Collapse and remove all axes of length 1.")
      .def("reshape", [](const NCDevArray& a, const std::vector<int64_t>& shape, bool copy) {
        if (!a.is_contiguous() && !copy) throw Error(NCA_ERR_INVALID, "Array is not contiguous, and copy was not requested!");
        return a.is_contiguous() ? a.reshape(shape) : a.copy_as_shape(shape);
      }, py::arg("shape"), py::arg("copy") = false)
      .def("astype", &NCDevArray::astype, py::arg("dtype"), "Convert to the specified data type (a new array).")
      .def("copy", &NCDevArray::copy)
      .def("view", &NCDevArray::view, "A view sharing this array's memory.")
      .def("__getitem__", [](const NCDevArray& a, const py::object& idx) { return a.view_from_indices(to_indices(idx)); },
           py::keep_alive<0, 1>())
      .def("__setitem__", [](const NCDevArray& a, const py::object& idx, const py::object& val) {
        set_into(a.view_from_indices(to_indices(idx)), val);
      })
      .def("fill", [](NCDevArray& a, const py::object& v) { a.fill(to_scalar(v)); }, "Broadcast a scalar value into an array.")
      .def("assign", [](NCDevArray& a, const py::object& v) { set_into(a, v); }, "Assign (copy) another array or an expression into this one.")
      .def("to_host", [](const NCDevArray& a, std::optional<DType> d) { return host_array(a, d); },
           py::arg("dtype") = py::none(), "Transfer a device array to host.")
      .def("to_device", [](const NCDevArray&, std::optional<DType>) -> py::object {
        throw Error(NCA_ERR_INVALID, "Array is already on device!");
      }, py::arg("dtype") = py::none())
      .def("item", [](const NCDevArray& a) { return scalar_to_py(a.item()); })
      .def("__array__", [](const NCDevArray& a, const py::object& dtype, const py::object&) -> py::object {
        py::array r = to_numpy_array(a, std::nullopt);
        if (!dtype.is_none()) return r.attr("astype")(dtype);
        return r;
      }, py::arg("dtype") = py::none(), py::arg("copy") = py::none())
      .def_property_readonly_static("__array_priority__", [](const py::object&) { return 100.0; })
      .def("__array_ufunc__", [](const NCDevArray&, py::handle ufunc, py::str method, py::args args, py::kwargs kwargs) -> py::object {
        if (method.cast<std::string>() != "__call__") return py::reinterpret_borrow<py::object>(Py_NotImplemented);
        py::list a;
        for (const py::handle& x : args) {
          if (py::isinstance<NCDevArray>(x)) a.append(to_numpy_array(x.cast<NCDevArray>(), std::nullopt));
          else a.append(x);
        }
        return ufunc(*a, **kwargs);
      })
      // full reductions return Python scalars; with `axis` they return device arrays
      .def("sum", [](const NCDevArray& a) { return scalar_to_py(a.sum()); })
      .def("sum", [](const NCDevArray& a, const py::object& axis) { return a.sum(to_axes(axis)); }, py::arg("axis"))
      .def("max", [](const NCDevArray& a) { return scalar_to_py(a.max()); })
      .def("max", [](const NCDevArray& a, const py::object& axis) { return a.max(to_axes(axis)); }, py::arg("axis"))
      .def("min", [](const NCDevArray& a) { return scalar_to_py(a.min()); })
      .def("min", [](const NCDevArray& a, const py::object& axis) { return a.min(to_axes(axis)); }, py::arg("axis"))
      .def("argmax", [](const NCDevArray& a) { return scalar_to_py(a.argmax()); })
      .def("argmax", [](const NCDevArray& a, const py::object& axis) { return a.argmax(to_axes(axis)); }, py::arg("axis"))
      .def("argmin", [](const NCDevArray& a) { return scalar_to_py(a.argmin()); })
      .def("argmin", [](const NCDevArray& a, const py::object& axis) { return a.argmin(to_axes(axis)); }, py::arg("axis"))
      .def("mean", [](const NCDevArray& a) { return scalar_to_py(a.mean()); })
      .def("mean", [](const NCDevArray& a, const py::object& axis) { return a.mean(to_axes(axis)); }, py::arg("axis"))
      .def("var", [](const NCDevArray& a, int64_t ddof) { return scalar_to_py(a.var(ddof)); }, py::arg("ddof") = 0)
      .def("var", [](const NCDevArray& a, const py::object& axis, int64_t ddof) { return a.var(to_axes(axis), ddof); },
           py::arg("axis"), py::arg("ddof") = 0)
      .def("std", [](const NCDevArray& a, int64_t ddof) { return scalar_to_py(a.std(ddof)); }, py::arg("ddof") = 0)
      .def("std", [](const NCDevArray& a, const py::object& axis, int64_t ddof) { return a.std(to_axes(axis), ddof); },
           py::arg("axis"), py::arg("ddof") = 0)
      .def("all", [](const NCDevArray& a) { return scalar_to_py(a.all()); })
      .def("all", [](const NCDevArray& a, const py::object& axis) { return a.all(to_axes(axis)); }, py::arg("axis"))
      .def("any", [](const NCDevArray& a) { return scalar_to_py(a.any()); })
      .def("any", [](const NCDevArray& a, const py::object& axis) { return a.any(to_axes(axis)); }, py::arg("axis"))
      .def("iota", [](const NCDevArray& a) { return finish(a.iota()); },
           "The C-order linear index of each element (an expression in lazy mode).")
      .def("scatter_add", [](NCDevArray& a, const py::object& idx, const py::object& src) {
        PyOperand i = resolve(idx), s = resolve(src);
        if (!i.arr || !s.arr) throw py::type_error("scatter_add needs array indices and sources");
        a.scatter_add(*i.arr, *s.arr);
      }, py::arg("indices"), py::arg("src"));
  bind_operators<py::class_<NCDevArray>, NCDevArray>(cls);
  struct IB {
    const char* name;
    OpCode op;
  };
  static const IB inplace[] = {{"__iadd__", OpCode::ADD}, {"__isub__", OpCode::SUB}, {"__imul__", OpCode::MUL},
                               {"__itruediv__", OpCode::DIV}, {"__imod__", OpCode::MOD}, {"__iand__", OpCode::LAND},
                               {"__ior__", OpCode::LOR}};
  for (const IB& b : inplace) {
    const OpCode op = b.op;
    cls.def(b.name, [op](NCDevArray& self, const py::object& other) {
      PyOperand o = resolve(other);
      Expr e(self);
      e.binary(op, o.op());
      HIPEngine::execute_binary_expression(e, self);
      return self;
    }, py::is_operator());
  }

  // ---- stencils -------------------------------------------------------------------------------
  py::class_<PyStencil>(m, "Stencil", "Stencil<NDIM> (expression/stencil.hh:76-243) with NDIM = len(offsets[0])")
      .def(py::init([](const std::vector<std::vector<int64_t>>& offsets, const py::function& func, DType src_dtype,
                       const std::string& extra_params, const std::string& prologue, const std::string& epilogue,
                       const std::vector<uint8_t>& is_pointer_axis, bool is_soarr) {
        if (offsets.empty()) throw Error(NCA_ERR_INVALID, "a stencil needs at least one offset");
        const int ndim = static_cast<int>(offsets[0].size());
        std::vector<int64_t> flat;
        for (const auto& o : offsets) {
          if (static_cast<int>(o.size()) != ndim) throw Error(NCA_ERR_INVALID, "offsets of different lengths");
          flat.insert(flat.end(), o.begin(), o.end());
        }
        py::list ph;
        const std::vector<int64_t> unit(static_cast<size_t>(ndim), 1);
        for (size_t k = 0; k < offsets.size(); ++k) ph.append(NCDevArray::wrap(nullptr, unit, unit, src_dtype));
        const bool was = g_eager;
        g_eager = false;  // placeholders only build a tape
        py::object r;
        try {
          r = func(ph);
        } catch (...) {
          g_eager = was;
          throw;
        }
        g_eager = was;
        Expr e = py::isinstance<Expr>(r) ? r.cast<Expr>() : Expr(r.cast<NCDevArray>());
        StencilJITExtensions ext{extra_params, prologue, epilogue};
        PyStencil s;
        if (!is_pointer_axis.empty() && static_cast<int>(is_pointer_axis.size()) != ndim)
          throw Error(NCA_ERR_INVALID, "is_pointer_axis needs one flag per axis");
        s.handle = detail::stencil_create(ndim, flat, src_dtype, e, ext, is_pointer_axis, is_soarr);
        s.scalars = e.scalars;
        s.ndim = ndim;
        s.dest_dtype = e.expr_dtype;
        return s;
      }), py::arg("offsets"), py::arg("func"), py::arg("src_dtype") = DType::float32, py::arg("extra_params") = "",
          py::arg("prologue") = "", py::arg("epilogue") = "", py::arg("is_pointer_axis") = std::vector<uint8_t>{},
          py::arg("is_soarr") = false)
      .def_property_readonly("dest_dtype", [](const PyStencil& s) { return s.dest_dtype; })
      .def_property_readonly("handle", [](const PyStencil& s) { return s.handle; })
      .def("apply", [](const PyStencil& s, const NCDevArray& src, const NCDevArray& dest, Boundary b) {
        detail::stencil_apply(s.handle, s.scalars, src, dest, nullptr, b, {});
      }, py::arg("src"), py::arg("dest"), py::arg("boundary") = Boundary::none)
      .def("adjoint", [](const PyStencil& s) {
        PyStencil a = s;
        a.handle = detail::stencil_adjoint(s.handle);
        return a;
      });
}
