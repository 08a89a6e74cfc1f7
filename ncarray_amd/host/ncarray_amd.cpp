// Implementation of the C++ host API (ncarray_amd.hh) over libnca_hip.so.
#include "ncarray_amd.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <sstream>
#include <unordered_map>

#include "ncarray_hip/dtype.hh"

namespace ncarray_amd {

// ---- dtypes and scalars ------------------------------------------------------------------
const char* dtype_name(DType d) { return nca::dtype_info(static_cast<int32_t>(d)).name; }
int itemsize(DType d) { return nca::itemsize(static_cast<int32_t>(d)); }
bool is_float(DType d) { return nca::is_float(static_cast<int32_t>(d)); }
bool is_signed(DType d) { return nca::dtype_info(static_cast<int32_t>(d)).is_signed; }

DType scalar_dtype(const Scalar& s) {
  return std::visit([](auto v) { return dtype_traits<decltype(v)>::value; }, s);
}

static float half_to_float(uint16_t h) {
  const uint32_t sign = (h >> 15) & 1u, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  float v;
  if (e == 0) v = std::ldexp(static_cast<float>(m), -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = std::ldexp(static_cast<float>(m | 0x400u), static_cast<int>(e) - 25);
  return sign ? -v : v;
}

Scalar scalar_from_bytes(DType d, const void* p) {
  switch (d) {
    case DType::bool_: return Scalar(*static_cast<const bool*>(p));
    case DType::char_: return Scalar(*static_cast<const char*>(p));
    case DType::uint8: return Scalar(*static_cast<const uint8_t*>(p));
    case DType::uint16: return Scalar(*static_cast<const uint16_t*>(p));
    case DType::uint32: return Scalar(*static_cast<const uint32_t*>(p));
    case DType::uint64: return Scalar(*static_cast<const uint64_t*>(p));
    case DType::int8: return Scalar(*static_cast<const int8_t*>(p));
    case DType::int16: return Scalar(*static_cast<const int16_t*>(p));
    case DType::int32: return Scalar(*static_cast<const int32_t*>(p));
    case DType::int64: return Scalar(*static_cast<const int64_t*>(p));
    case DType::float32: return Scalar(*static_cast<const float*>(p));
    case DType::float64: return Scalar(*static_cast<const double*>(p));
    case DType::bfloat16: {
      uint32_t u = static_cast<uint32_t>(*static_cast<const uint16_t*>(p)) << 16;
      float f;
      std::memcpy(&f, &u, 4);
      return Scalar(f);
    }
    case DType::float16: return Scalar(half_to_float(*static_cast<const uint16_t*>(p)));
    default: throw Error(NCA_ERR_UNSUPPORTED, std::string("no scalar type for dtype ") + dtype_name(d));
  }
}

double scalar_to_double(const Scalar& s) {
  return std::visit([](auto v) { return static_cast<double>(v); }, s);
}

Scalar scalar_cast(const Scalar& s, DType d) {
  return std::visit(
      [d](auto v) -> Scalar {
        switch (d) {
          case DType::bool_: return Scalar(static_cast<bool>(v));
          case DType::char_: return Scalar(static_cast<char>(v));
          case DType::uint8: return Scalar(static_cast<uint8_t>(v));
          case DType::uint16: return Scalar(static_cast<uint16_t>(v));
          case DType::uint32: return Scalar(static_cast<uint32_t>(v));
          case DType::uint64: return Scalar(static_cast<uint64_t>(v));
          case DType::int8: return Scalar(static_cast<int8_t>(v));
          case DType::int16: return Scalar(static_cast<int16_t>(v));
          case DType::int32: return Scalar(static_cast<int32_t>(v));
          case DType::int64: return Scalar(static_cast<int64_t>(v));
          case DType::float32: return Scalar(static_cast<float>(v));
          case DType::float64: return Scalar(static_cast<double>(v));
          default: throw Error(NCA_ERR_UNSUPPORTED, std::string("no scalar type for dtype ") + dtype_name(d));
        }
      },
      s);
}

std::string scalar_repr(const Scalar& s) {
  std::ostringstream os;
  std::visit(
      [&os](auto v) {
        using T = decltype(v);
        if constexpr (std::is_same_v<T, bool>) os << (v ? "True" : "False");
        else if constexpr (std::is_same_v<T, char> || std::is_same_v<T, int8_t> || std::is_same_v<T, uint8_t>)
          os << static_cast<int>(v);
        else {
          os.precision(std::is_same_v<T, float> ? 9 : 17);
          os << v;
        }
      },
      s);
  return os.str();
}

DType reduce_result_dtype(ReduceOp op, DType src) {
  return static_cast<DType>(nca::reduce_result_dtype(static_cast<int32_t>(op), static_cast<int32_t>(src)));
}

void check(int status, const char* what) {
  if (status == NCA_OK) return;
  char buf[4096];
  buf[0] = 0;
  nca_hip_last_error(buf, sizeof buf);
  throw Error(status, std::string(what) + ": " + buf);
}

namespace {
thread_local int t_device = -1;
void ensure_init() {
  if (t_device < 0) init(0);
}
}  // namespace

void init(int device) {
  check(nca_hip_init(device), "nca_hip_init");
  t_device = device;
}

// ---- storage -------------------------------------------------------------------------------
Storage::Storage(size_t bytes) : owned_(true) {
  ensure_init();
  if (bytes) check(nca_hip_malloc(&ptr_, bytes), "device allocation");
}
Storage::~Storage() {
  if (owned_ && ptr_) (void)nca_hip_free(ptr_);
}

// ---- helpers -------------------------------------------------------------------------------
namespace {

std::vector<int64_t> broadcast_shapes(const std::vector<int64_t>& a, const std::vector<int64_t>& b) {
  const size_t n = std::max(a.size(), b.size());
  std::vector<int64_t> r(n);
  for (size_t i = 0; i < n; ++i) {
    const int64_t x = i < n - a.size() ? 1 : a[i - (n - a.size())];
    const int64_t y = i < n - b.size() ? 1 : b[i - (n - b.size())];
    if (x != y && x != 1 && y != 1) {
      std::ostringstream os;
      os << "operands could not be broadcast together (extents " << x << " and " << y << ")";
      throw Error(NCA_ERR_INVALID, os.str());
    }
    r[i] = x == 1 ? y : x;
  }
  return r;
}

// A view of `v` broadcast to `shape`: leading axes added, extent-1 axes given stride 0.
nca_view broadcast_view(const nca_view& v, const std::vector<int64_t>& shape) {
  const int n = static_cast<int>(shape.size());
  if (v.ndim > n) throw Error(NCA_ERR_INVALID, "operand has more axes than the result");
  nca_view r = v;
  r.ndim = n;
  const int lead = n - v.ndim;
  // a pointer axis keeps its place counted from the right; broadcasting it (extent 1, table
  // stride 0) repeats its one row
  if (v.pointer_axis >= 0) r.pointer_axis = v.pointer_axis + lead;
  for (int d = n - 1; d >= 0; --d) {
    // SO suboffsets move with their axes; added axes are direct
    r.suboffsets[d] = d < lead ? -1 : v.suboffsets[d - lead];
    if (d < lead) {
      r.shape[d] = shape[d];
      r.strides[d] = 0;
      continue;
    }
    const int64_t e = v.shape[d - lead], s = v.strides[d - lead];
    if (e == shape[d]) {
      r.shape[d] = e;
      r.strides[d] = s;
    } else if (e == 1) {
      r.shape[d] = shape[d];
      r.strides[d] = 0;
    } else {
      throw Error(NCA_ERR_INVALID, "operand shape does not broadcast to the result shape");
    }
  }
  return r;
}

struct ExprKey {
  std::string bytes;
  bool operator==(const ExprKey& o) const { return bytes == o.bytes; }
};
struct ExprKeyHash {
  size_t operator()(const ExprKey& k) const { return std::hash<std::string>()(k.bytes); }
};

std::mutex g_expr_mu;
std::unordered_map<ExprKey, uint64_t, ExprKeyHash> g_expr_handles;

template <typename T>
void put(std::string& s, const T& v) {
  s.append(reinterpret_cast<const char*>(&v), sizeof v);
}

// The compiled expression family of a tape (cached per tape signature; the library also
// caches kernels in memory and on disk).
uint64_t expr_handle(const Expr& e, DType dest, std::vector<int32_t>& arr_dt, std::vector<int32_t>& sc_dt) {
  arr_dt.clear();
  for (DType d : e.dtypes) arr_dt.push_back(static_cast<int32_t>(d));
  sc_dt.clear();
  for (const Scalar& s : e.scalars) sc_dt.push_back(static_cast<int32_t>(scalar_dtype(s)));
  ExprKey key;
  put(key.bytes, static_cast<int32_t>(e.work_dtype));
  put(key.bytes, static_cast<int32_t>(dest));
  put(key.bytes, static_cast<int32_t>(e.instrs.size()));
  for (Instruction i : e.instrs) put(key.bytes, i);
  put(key.bytes, static_cast<int32_t>(arr_dt.size()));
  for (int32_t d : arr_dt) put(key.bytes, d);
  for (int32_t d : sc_dt) put(key.bytes, d);
  {
    std::lock_guard<std::mutex> lk(g_expr_mu);
    auto it = g_expr_handles.find(key);
    if (it != g_expr_handles.end()) return it->second;
  }
  nca_expr_desc d{};
  d.instrs = e.instrs.data();
  d.n_instrs = static_cast<int32_t>(e.instrs.size());
  d.n_arrays = static_cast<int32_t>(arr_dt.size());
  d.arr_dtypes = arr_dt.data();
  d.n_scalars = static_cast<int32_t>(sc_dt.size());
  d.work_dtype = static_cast<int32_t>(e.work_dtype);
  d.scalar_dtypes = sc_dt.data();
  d.dest_dtype = static_cast<int32_t>(dest);
  uint64_t h = 0;
  check(nca_hip_expr_compile(&d, &h), "expression compile");
  std::lock_guard<std::mutex> lk(g_expr_mu);
  g_expr_handles.emplace(std::move(key), h);
  return h;
}

// A small device buffer for reduction results (one per thread, process lifetime, like the
// reference's scalar pool, mem_pool.cuh:72-85).
void* result_slot() {
  thread_local void* p = nullptr;
  if (!p) check(nca_hip_malloc(&p, 64), "result buffer");
  return p;
}

DType dt_of(int32_t c) { return static_cast<DType>(c); }

// Copy a strided host array into a C-ordered host buffer.
void compact_host(const char* src, const std::vector<int64_t>& shape, const std::vector<int64_t>& strides,
                  int isz, char* dst) {
  const int nd = static_cast<int>(shape.size());
  int64_t n = 1;
  for (int64_t e : shape) n *= e;
  if (n == 0) return;
  std::vector<int64_t> idx(static_cast<size_t>(nd), 0);
  for (int64_t k = 0; k < n; ++k) {
    int64_t off = 0;
    for (int d = 0; d < nd; ++d) off += idx[d] * strides[d];
    std::memcpy(dst + k * isz, src + off, static_cast<size_t>(isz));
    for (int d = nd - 1; d >= 0; --d) {
      if (++idx[d] < shape[d]) break;
      idx[d] = 0;
    }
  }
}

}  // namespace

namespace detail {
std::vector<nca_scalar> pack_scalars(const std::vector<Scalar>& sc, std::vector<int32_t>* dtypes) {
  std::vector<nca_scalar> out(std::max<size_t>(1, sc.size()));
  for (size_t i = 0; i < sc.size(); ++i) {
    const DType d = scalar_dtype(sc[i]);
    if (dtypes) dtypes->push_back(static_cast<int32_t>(d));
    std::visit(
        [&](auto v) {
          using T = decltype(v);
          if constexpr (std::is_floating_point_v<T>) out[i].f = static_cast<double>(v);
          else if constexpr (std::is_same_v<T, bool> || std::is_unsigned_v<T>) out[i].u = static_cast<uint64_t>(v);
          else out[i].i = static_cast<int64_t>(v);
        },
        sc[i]);
  }
  return out;
}
}  // namespace detail

// ---- Expr ----------------------------------------------------------------------------------
void Expr::build_node(const NCDevArray& a) {
  if (a.empty() && a.data() == nullptr && a.ndim() == 0) throw Error(NCA_ERR_INVALID, "empty array in an expression");
  const std::vector<int64_t> sh = a.shape_vec();
  if (!has_expr_dtype) {
    expr_dtype = work_dtype = a.dtype();
    final_shape = sh;
    has_expr_dtype = true;
  } else {
    final_shape = broadcast_shapes(final_shape, sh);
  }
  views.push_back(a.cview());
  keep.push_back(a.storage());
  dtypes.push_back(a.dtype());
  instrs.push_back(pack_instruction(OpCode::LOAD_NCARR, static_cast<int>(views.size()) - 1));
}

void Expr::build_node(const Expr& e) {
  if (!has_expr_dtype) {
    expr_dtype = work_dtype = e.expr_dtype;
    final_shape = e.final_shape;
    has_expr_dtype = true;
  } else {
    final_shape = broadcast_shapes(final_shape, e.final_shape);
  }
  const int arr_shift = static_cast<int>(views.size()), sc_shift = static_cast<int>(scalars.size());
  views.insert(views.end(), e.views.begin(), e.views.end());
  keep.insert(keep.end(), e.keep.begin(), e.keep.end());
  dtypes.insert(dtypes.end(), e.dtypes.begin(), e.dtypes.end());
  for (Instruction i : e.instrs) {
    const OpCode op = get_op(i);
    if (op == OpCode::LOAD_NCARR || op == OpCode::LOAD_SOARR) instrs.push_back(pack_instruction(op, get_index(i) + arr_shift));
    else if (op == OpCode::LOAD_CONST) instrs.push_back(pack_instruction(op, get_index(i) + sc_shift));
    else instrs.push_back(i);
  }
  scalars.insert(scalars.end(), e.scalars.begin(), e.scalars.end());
}

void Expr::build_node(const Scalar& s) {
  if (!has_expr_dtype) {
    expr_dtype = work_dtype = scalar_dtype(s);
    final_shape.clear();
    has_expr_dtype = true;
  }
  scalars.push_back(s);
  instrs.push_back(pack_instruction(OpCode::LOAD_CONST, static_cast<int>(scalars.size()) - 1));
}

void Expr::build_node(const Operand& o) {
  std::visit(
      [this](auto&& x) {
        using T = std::decay_t<decltype(x)>;
        if constexpr (std::is_same_v<T, Scalar>) build_node(x);
        else build_node(*x);
      },
      o.v);
}

static DType operand_dtype(const Operand& o) {
  return std::visit(
      [](auto&& x) -> DType {
        using T = std::decay_t<decltype(x)>;
        if constexpr (std::is_same_v<T, Scalar>) return scalar_dtype(x);
        else return x->dtype();
      },
      o.v);
}

Expr& Expr::binary(OpCode op, const Operand& right) {
  build_node(right);
  work_dtype = dt_of(nca::promote_expr_types(static_cast<int32_t>(work_dtype), static_cast<int32_t>(operand_dtype(right))));
  expr_dtype = dt_of(nca::op_result_dtype(static_cast<int32_t>(op), static_cast<int32_t>(work_dtype)));
  instrs.push_back(pack_instruction(op, 0));
  return *this;
}

Expr& Expr::unary(OpCode op) {
  // exp/log/sqrt have no integer form: an integer operand is evaluated in float64.
  if ((op == OpCode::EXP || op == OpCode::LOG || op == OpCode::SQRT) && !is_float(work_dtype))
    work_dtype = DType::float64;
  expr_dtype = dt_of(nca::op_result_dtype(static_cast<int32_t>(op), static_cast<int32_t>(work_dtype)));
  instrs.push_back(pack_instruction(op, 0));
  return *this;
}

int64_t Expr::size() const {
  int64_t n = 1;
  for (int64_t e : final_shape) n *= e;
  return n;
}

NCDevArray Expr::materialize(std::optional<DType> as) const {
  NCDevArray r(final_shape, as.value_or(expr_dtype));
  HIPEngine::execute_binary_expression(*this, r);
  return r;
}

namespace {
Scalar expr_reduce(const Expr& e, int fuse, Accum mode) {
  ensure_init();
  std::vector<int32_t> adt, sdt;
  const uint64_t h = expr_handle(e, e.expr_dtype, adt, sdt);
  nca_view out{};
  out.data = nullptr;
  out.dtype = static_cast<int32_t>(e.expr_dtype);
  out.ndim = static_cast<int32_t>(e.final_shape.size());
  out.pointer_axis = -1;
  int64_t s = itemsize(e.expr_dtype);
  for (int d = out.ndim - 1; d >= 0; --d) {
    out.shape[d] = e.final_shape[static_cast<size_t>(d)];
    out.strides[d] = s;
    s *= out.shape[d];
  }
  std::vector<nca_view> ins;
  for (const nca_view& v : e.views) ins.push_back(broadcast_view(v, e.final_shape));
  std::vector<nca_scalar> sc = detail::pack_scalars(e.scalars, nullptr);
  char* slot = static_cast<char*>(result_slot());
  const DType rdt = fuse == NCA_FUSE_SUM ? reduce_result_dtype(ReduceOp::sum, e.expr_dtype) : e.expr_dtype;
  check(nca_hip_expr_launch_reduce(h, ins.data(), static_cast<int32_t>(ins.size()), sc.data(),
                                   static_cast<int32_t>(e.scalars.size()), &out, fuse, static_cast<int32_t>(mode),
                                   fuse == NCA_FUSE_SUM ? slot : nullptr, fuse == NCA_FUSE_MAX ? slot : nullptr,
                                   fuse == NCA_FUSE_MIN ? slot : nullptr, nullptr),
        "fused expression reduction");
  alignas(16) char host[16];
  check(nca_hip_memcpy_d2h(host, slot, static_cast<size_t>(itemsize(rdt)), nullptr), "result copy");
  return scalar_from_bytes(rdt, host);
}
}  // namespace

Scalar Expr::sum(Accum mode) const { return expr_reduce(*this, NCA_FUSE_SUM, mode); }
Scalar Expr::max() const { return expr_reduce(*this, NCA_FUSE_MAX, Accum::plain); }
Scalar Expr::min() const { return expr_reduce(*this, NCA_FUSE_MIN, Accum::plain); }

static Expr bin(OpCode op, const Operand& a, const Operand& b) {
  Expr e;
  e.build_node(a);
  e.binary(op, b);
  return e;
}
static Expr un(OpCode op, const Operand& a) {
  Expr e;
  e.build_node(a);
  e.unary(op);
  return e;
}
Expr operator+(const Operand& a, const Operand& b) { return bin(OpCode::ADD, a, b); }
Expr operator-(const Operand& a, const Operand& b) { return bin(OpCode::SUB, a, b); }
Expr operator*(const Operand& a, const Operand& b) { return bin(OpCode::MUL, a, b); }
Expr operator/(const Operand& a, const Operand& b) { return bin(OpCode::DIV, a, b); }
Expr operator%(const Operand& a, const Operand& b) { return bin(OpCode::MOD, a, b); }
Expr operator==(const Operand& a, const Operand& b) { return bin(OpCode::EQ, a, b); }
Expr operator!=(const Operand& a, const Operand& b) { return bin(OpCode::NE, a, b); }
Expr operator<(const Operand& a, const Operand& b) { return bin(OpCode::LT, a, b); }
Expr operator<=(const Operand& a, const Operand& b) { return bin(OpCode::LE, a, b); }
Expr operator>(const Operand& a, const Operand& b) { return bin(OpCode::GT, a, b); }
Expr operator>=(const Operand& a, const Operand& b) { return bin(OpCode::GE, a, b); }
Expr operator&&(const Operand& a, const Operand& b) { return bin(OpCode::LAND, a, b); }
Expr operator||(const Operand& a, const Operand& b) { return bin(OpCode::LOR, a, b); }
Expr operator&(const Operand& a, const Operand& b) { return bin(OpCode::BAND, a, b); }
Expr operator|(const Operand& a, const Operand& b) { return bin(OpCode::BOR, a, b); }
Expr operator^(const Operand& a, const Operand& b) { return bin(OpCode::XOR, a, b); }
Expr operator<<(const Operand& a, const Operand& b) { return bin(OpCode::LSHFT, a, b); }
Expr operator>>(const Operand& a, const Operand& b) { return bin(OpCode::RSHFT, a, b); }
Expr floordiv(const Operand& a, const Operand& b) { return bin(OpCode::FDIV, a, b); }
Expr operator-(const Operand& a) { return un(OpCode::NEG, a); }
Expr operator!(const Operand& a) { return un(OpCode::LNOT, a); }
Expr operator~(const Operand& a) { return un(OpCode::BNOT, a); }
Expr exp(const Operand& a) { return un(OpCode::EXP, a); }
Expr log(const Operand& a) { return un(OpCode::LOG, a); }
Expr sqrt(const Operand& a) { return un(OpCode::SQRT, a); }
Expr abs(const Operand& a) { return un(OpCode::ABS, a); }

// ---- NCDevArray ------------------------------------------------------------------------------
NCDevArray::NCDevArray(const Shape& shape, DType dtype) : NCDevArray(static_cast<int>(shape.size()), shape.data(), dtype) {}

NCDevArray::NCDevArray(int ndim, const int64_t* shape, DType dtype) : dt_(dtype), nd_(ndim) {
  if (ndim < 0 || ndim > NCA_MAX_NDIM) throw Error(NCA_ERR_INVALID, "ndim out of range");
  if (!nca::is_supported(static_cast<int32_t>(dtype)))
    throw Error(NCA_ERR_UNSUPPORTED, std::string("dtype ") + dtype_name(dtype) + " has no device kernels");
  const int64_t isz = ncarray_amd::itemsize(dtype);
  int64_t s = isz, n = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    if (shape[d] < 0) throw Error(NCA_ERR_INVALID, "negative extent");
    sh_[d] = shape[d];
    st_[d] = s;
    s *= shape[d];
    n *= shape[d];
  }
  storage_ = std::make_shared<Storage>(static_cast<size_t>(n * isz));
  data_ = storage_->ptr();
}

NCDevArray::NCDevArray(const Expr& e) : NCDevArray(e.materialize()) {}

NCDevArray NCDevArray::wrap(void* data, const Shape& shape, const Shape& strides, DType dtype,
                            std::shared_ptr<void> keep) {
  if (shape.size() != strides.size() || shape.size() > NCA_MAX_NDIM) throw Error(NCA_ERR_INVALID, "bad shape/strides");
  NCDevArray a;
  a.dt_ = dtype;
  a.nd_ = static_cast<int>(shape.size());
  for (size_t d = 0; d < shape.size(); ++d) {
    a.sh_[d] = shape[d];
    a.st_[d] = strides[d];
  }
  a.storage_ = std::make_shared<Storage>(data, std::move(keep));
  a.data_ = data;
  return a;
}

NCDevArray NCDevArray::from_row_pointers(const std::vector<void*>& rows, const Shape& row_shape,
                                         const Shape& row_strides, DType dtype, std::shared_ptr<void> keep) {
  ensure_init();
  if (row_shape.size() != row_strides.size() || row_shape.size() + 1 > NCA_MAX_NDIM)
    throw Error(NCA_ERR_INVALID, "bad row shape/strides");
  if (!nca::is_supported(static_cast<int32_t>(dtype)))
    throw Error(NCA_ERR_UNSUPPORTED, std::string("dtype ") + dtype_name(dtype) + " has no device kernels");
  int64_t row_n = 1;
  for (int64_t e : row_shape) row_n *= e;
  for (void* r : rows)
    if (r == nullptr && row_n > 0) throw Error(NCA_ERR_INVALID, "null row pointer");
  NCDevArray a;
  a.dt_ = dtype;
  a.nd_ = static_cast<int>(row_shape.size()) + 1;
  a.sh_[0] = static_cast<int64_t>(rows.size());
  a.st_[0] = 1;
  for (size_t d = 0; d < row_shape.size(); ++d) {
    if (row_shape[d] < 0) throw Error(NCA_ERR_INVALID, "negative extent");
    a.sh_[d + 1] = row_shape[d];
    a.st_[d + 1] = row_strides[d];
  }
  a.storage_ = std::make_shared<Storage>(rows.size() * sizeof(void*));
  a.storage_->hold(std::move(keep));
  if (!rows.empty())
    check(nca_hip_memcpy_h2d(a.storage_->ptr(), rows.data(), rows.size() * sizeof(void*), nullptr), "row table copy");
  a.data_ = a.storage_->ptr();
  a.pax_ = 0;
  a.rows_ = std::make_shared<const std::vector<void*>>(rows);
  return a;
}

NCDevArray NCDevArray::from_rows(const std::vector<NCDevArray>& rows) {
  if (rows.empty()) throw Error(NCA_ERR_INVALID, "from_rows needs at least one row");
  const NCDevArray& r0 = rows[0];
  auto keep = std::make_shared<std::vector<std::shared_ptr<Storage>>>();
  std::vector<void*> ptrs;
  for (const NCDevArray& r : rows) {
    if (r.pax_ >= 0) throw Error(NCA_ERR_INVALID, "rows of a pointer-axis array cannot have a pointer axis");
    if (r.dt_ != r0.dt_ || r.shape_vec() != r0.shape_vec()) throw Error(NCA_ERR_INVALID, "rows differ in shape or dtype");
    if (r.strides_vec() != r0.strides_vec()) throw Error(NCA_ERR_INVALID, "rows differ in stride layout");
    ptrs.push_back(r.data_);
    keep->push_back(r.storage_);
  }
  return from_row_pointers(ptrs, r0.shape_vec(), r0.strides_vec(), r0.dt_, keep);
}

void* NCDevArray::row(int64_t i) const {
  if (pax_ < 0 || !rows_) throw Error(NCA_ERR_INVALID, "not a pointer-axis array");
  if (i < 0 || i >= sh_[pax_]) throw Error(NCA_ERR_INVALID, "Index out of bounds!");
  const int64_t base = (static_cast<const char*>(data_) - static_cast<const char*>(storage_->ptr())) /
                       static_cast<int64_t>(sizeof(void*));
  const int64_t entry = base + i * st_[pax_] + poff_;
  if (entry < 0 || entry >= static_cast<int64_t>(rows_->size())) throw Error(NCA_ERR_INVALID, "row table index out of range");
  return static_cast<char*>((*rows_)[static_cast<size_t>(entry)]) + roff_;
}

NCDevArray NCDevArray::dense() const { return pax_ >= 0 ? copy() : *this; }

NCDevArray& NCDevArray::operator=(const Expr& e) {
  if (empty()) {
    *this = NCDevArray(e.final_shape, e.expr_dtype);
  }
  HIPEngine::execute_binary_expression(e, *this);
  return *this;
}

int64_t NCDevArray::size() const {
  int64_t n = 1;
  for (int d = 0; d < nd_; ++d) n *= sh_[d];
  return n;
}

bool NCDevArray::is_contiguous() const {
  if (pax_ >= 0) return false;
  int64_t s = itemsize();
  for (int d = nd_ - 1; d >= 0; --d) {
    if (sh_[d] != 1 && st_[d] != s) return false;
    s *= sh_[d];
  }
  return true;
}

nca_view NCDevArray::cview() const {
  nca_view v{};
  v.data = data_;
  v.dtype = static_cast<int32_t>(dt_);
  v.ndim = nd_;
  for (int d = 0; d < nd_; ++d) {
    v.shape[d] = sh_[d];
    v.strides[d] = st_[d];
  }
  v.pointer_axis = pax_;
  v.ptr_offset = pax_ >= 0 ? poff_ : 0;
  v.row_offset = pax_ >= 0 ? roff_ : 0;
  return v;
}

std::string NCDevArray::repr() const {
  std::ostringstream os;
  os << "NCDevArray(shape=(";
  for (int d = 0; d < nd_; ++d) os << sh_[d] << (nd_ == 1 ? "," : (d + 1 < nd_ ? ", " : ""));
  os << "), dtype=" << dtype_name(dt_) << ")";
  return os.str();
}

NCDevArray NCDevArray::view_from_indices(const std::vector<IndexItem>& items) const {
  int consumed = 0, n_ell = 0;
  for (const IndexItem& it : items) {
    if (std::holds_alternative<Ellipsis>(it)) ++n_ell;
    else if (!std::holds_alternative<NewAxis>(it)) ++consumed;
  }
  if (n_ell > 1) throw Error(NCA_ERR_INVALID, "an index can only have a single ellipsis");
  if (consumed > nd_) throw Error(NCA_ERR_INVALID, "too many indices for array");
  NCDevArray r = *this;
  r.nd_ = 0;
  r.pax_ = -1;
  // address arithmetic in integers: a geometry-only array has no data pointer
  uintptr_t base = reinterpret_cast<uintptr_t>(data_);
  // pointer axis (out_from_axes_ptr, array_impl.hh:964-1031): axes before it move the table
  // position, an integer index on it dereferences the row, a slice moves the table offset;
  // axes after it move the row offset while the pointer axis is kept
  bool live_ptr = pax_ >= 0;
  int64_t poff = poff_, roff = roff_;
  int src = 0;
  auto push = [&r](int64_t e, int64_t s) {
    if (r.nd_ >= NCA_MAX_NDIM) throw Error(NCA_ERR_INVALID, "too many axes");
    r.sh_[r.nd_] = e;
    r.st_[r.nd_] = s;
    ++r.nd_;
  };
  auto keep_axis = [&](int a) {
    if (a == pax_) r.pax_ = r.nd_;
    push(sh_[a], st_[a]);
  };
  auto move = [&](int a, int64_t bytes) {
    if (live_ptr && a > pax_) roff += bytes;
    else base += static_cast<uintptr_t>(bytes);
  };
  for (const IndexItem& it : items) {
    if (std::holds_alternative<Ellipsis>(it)) {
      const int fill = nd_ - consumed;
      for (int k = 0; k < fill; ++k, ++src) keep_axis(src);
    } else if (std::holds_alternative<NewAxis>(it)) {
      push(1, 0);
    } else if (const int64_t* pi = std::get_if<int64_t>(&it)) {
      int64_t i = *pi;
      if (i < -sh_[src] || i >= sh_[src]) throw Error(NCA_ERR_INVALID, "Index out of bounds!");
      if (i < 0) i += sh_[src];
      if (live_ptr && src == pax_) {
        NCDevArray probe = *this;
        probe.data_ = reinterpret_cast<void*>(base);
        probe.poff_ = poff;
        probe.roff_ = 0;
        base = reinterpret_cast<uintptr_t>(probe.row(i)) + static_cast<uintptr_t>(roff);
        roff = 0;
        live_ptr = false;
      } else {
        move(src, i * st_[src]);
      }
      ++src;
    } else {
      const Slice& s = std::get<Slice>(it);
      const int64_t n = sh_[src];
      const int64_t step = s.step;
      if (step == 0) throw Error(NCA_ERR_INVALID, "slice step cannot be zero");
      // Python's slice.indices()
      int64_t start, stop;
      if (step > 0) {
        start = s.start == Slice::kNone ? 0 : (s.start < 0 ? std::max<int64_t>(0, s.start + n) : std::min(s.start, n));
        stop = s.stop == Slice::kNone ? n : (s.stop < 0 ? std::max<int64_t>(0, s.stop + n) : std::min(s.stop, n));
      } else {
        start = s.start == Slice::kNone ? n - 1
                                        : (s.start < 0 ? std::max<int64_t>(-1, s.start + n) : std::min(s.start, n - 1));
        stop = s.stop == Slice::kNone ? -1 : (s.stop < 0 ? std::max<int64_t>(-1, s.stop + n) : std::min(s.stop, n - 1));
      }
      int64_t len = 0;
      if (step > 0 && stop > start) len = (stop - start + step - 1) / step;
      if (step < 0 && start > stop) len = (start - stop - step - 1) / (-step);
      if (live_ptr && src == pax_) {
        if (len > 0) poff += start * st_[src];
        r.pax_ = r.nd_;
      } else if (len > 0) {
        move(src, start * st_[src]);
      }
      push(len, st_[src] * step);
      ++src;
    }
  }
  for (; src < nd_; ++src) keep_axis(src);
  r.data_ = reinterpret_cast<void*>(base);
  if (live_ptr) {
    r.poff_ = poff;
    r.roff_ = roff;
  } else {
    r.pax_ = -1;
    r.poff_ = r.roff_ = 0;
    r.rows_.reset();
  }
  return r;
}

NCDevArray NCDevArray::reshape(const Shape& shape) const {
  int64_t n = 1;
  for (int64_t e : shape) n *= e;
  if (n != size())
    throw Error(NCA_ERR_INVALID, "Cannot reshape " + std::to_string(size()) + " elements into " + std::to_string(n));
  if (!is_contiguous()) throw Error(NCA_ERR_INVALID, "Array is not contiguous, and copy was not requested!");
  if (shape.size() > NCA_MAX_NDIM) throw Error(NCA_ERR_INVALID, "too many axes");
  NCDevArray r = *this;
  r.nd_ = static_cast<int>(shape.size());
  int64_t s = itemsize();
  for (int d = r.nd_ - 1; d >= 0; --d) {
    r.sh_[d] = shape[static_cast<size_t>(d)];
    r.st_[d] = s;
    s *= r.sh_[d];
  }
  return r;
}

NCDevArray NCDevArray::copy_as_shape(const Shape& shape) const { return copy().reshape(shape); }

NCDevArray NCDevArray::squeeze() const {
  if (pax_ >= 0) {
    // a length-1 pointer axis collapses like any other (array_impl.hh:1146-1172)
    std::vector<IndexItem> items;
    for (int d = 0; d < nd_; ++d) items.push_back(sh_[d] == 1 ? IndexItem(int64_t{0}) : IndexItem(Slice()));
    return view_from_indices(items);
  }
  NCDevArray r = *this;
  r.nd_ = 0;
  for (int d = 0; d < nd_; ++d)
    if (sh_[d] != 1) {
      r.sh_[r.nd_] = sh_[d];
      r.st_[r.nd_] = st_[d];
      ++r.nd_;
    }
  return r;
}

NCDevArray NCDevArray::copy() const { return astype(dt_); }

NCDevArray NCDevArray::astype(DType d) const {
  NCDevArray r(shape_vec(), d);
  r.assign(*this);
  return r;
}

Expr NCDevArray::iota() const {
  Expr e;
  e.expr_dtype = e.work_dtype = dt_;
  e.has_expr_dtype = true;
  e.final_shape = shape_vec();
  e.instrs.push_back(pack_instruction(OpCode::IDX, 0));
  return e;
}

NCDevArray NCDevArray::iota(const Shape& shape, DType dtype) {
  NCDevArray r(shape, dtype);
  r = r.iota();
  return r;
}

void NCDevArray::fill(const Scalar& v) { HIPEngine::execute_fill(*this, v); }
void NCDevArray::assign(const NCDevArray& src) { HIPEngine::execute_assign(*this, src); }

void NCDevArray::assign_from_host(const void* host, DType host_dtype, const Shape& host_shape,
                                  const Shape& host_strides) {
  ensure_init();
  const int isz = ncarray_amd::itemsize(host_dtype);
  int64_t n = 1;
  for (int64_t e : host_shape) n *= e;
  bool c_order = true;
  if (!host_strides.empty()) {
    if (host_strides.size() != host_shape.size()) throw Error(NCA_ERR_INVALID, "host strides/shape mismatch");
    int64_t s = isz;
    for (int d = static_cast<int>(host_shape.size()) - 1; d >= 0; --d) {
      if (host_shape[d] != 1 && host_strides[d] != s) c_order = false;
      s *= host_shape[d];
    }
  }
  std::vector<char> staging;
  const char* src = static_cast<const char*>(host);
  if (!c_order) {
    staging.resize(static_cast<size_t>(n * isz));
    compact_host(src, host_shape, host_strides, isz, staging.data());
    src = staging.data();
  }
  if (host_dtype == dt_ && host_shape == shape_vec() && is_contiguous()) {
    if (n) check(nca_hip_memcpy_h2d(data_, src, static_cast<size_t>(n * isz), nullptr), "host to device copy");
    return;
  }
  NCDevArray tmp(host_shape, host_dtype);
  if (n) check(nca_hip_memcpy_h2d(tmp.data(), src, static_cast<size_t>(n * isz), nullptr), "host to device copy");
  assign(tmp);
}

void NCDevArray::copy_into_astype(void* host, DType host_dtype) const { HIPEngine::execute_copy_into(*this, host, host_dtype); }

Scalar NCDevArray::item() const {
  if (size() != 1) throw Error(NCA_ERR_INVALID, "item() needs an array of one element");
  alignas(16) char buf[16];
  copy_into(buf);
  return scalar_from_bytes(dt_, buf);
}

Scalar NCDevArray::full(ReduceOp op, Accum mode) const {
  if (pax_ >= 0) {
    // sum/max/min read the rows through the fused expression reduction (one pointer read per
    // row); the other reductions follow the row pointers element by element
    if (op == ReduceOp::sum) return Expr(*this).sum(mode);
    if (op == ReduceOp::max) return Expr(*this).max();
    if (op == ReduceOp::min) return Expr(*this).min();
  }
  return HIPEngine::execute_full_reduce(*this, op, mode);
}

NCDevArray NCDevArray::axes_reduce(ReduceOp op, const std::vector<int64_t>& axes, Accum mode) const {
  Shape kept;
  for (int d = 0; d < nd_; ++d)
    if (std::find(axes.begin(), axes.end(), static_cast<int64_t>(d)) == axes.end()) kept.push_back(sh_[d]);
  NCDevArray r(kept, reduce_result_dtype(op, dt_));
  HIPEngine::execute_reduce_axes(*this, axes, op, r, mode);
  return r;
}

Scalar NCDevArray::sum(Accum mode) const { return full(ReduceOp::sum, mode); }
Scalar NCDevArray::max() const { return full(ReduceOp::max, Accum::plain); }
Scalar NCDevArray::min() const { return full(ReduceOp::min, Accum::plain); }
Scalar NCDevArray::argmax() const { return full(ReduceOp::argmax, Accum::plain); }
Scalar NCDevArray::argmin() const { return full(ReduceOp::argmin, Accum::plain); }
Scalar NCDevArray::mean() const { return full(ReduceOp::mean, Accum::compensated); }
Scalar NCDevArray::all() const { return full(ReduceOp::all, Accum::plain); }
Scalar NCDevArray::any() const { return full(ReduceOp::any, Accum::plain); }

// var: host/reductions.hh:96-118 (two passes: mean, then the sum of squared deviations in
// truediv_type, divided by n - ddof).
Scalar NCDevArray::var(int64_t ddof) const {
  const double m = scalar_to_double(mean());
  Expr d = *this - m;
  Expr sq = d * d;
  const double ss = scalar_to_double(sq.sum(Accum::compensated));
  return Scalar(ss / static_cast<double>(size() - ddof));
}
Scalar NCDevArray::std(int64_t ddof) const { return Scalar(std::sqrt(scalar_to_double(var(ddof)))); }

NCDevArray NCDevArray::sum(const std::vector<int64_t>& axes, Accum mode) const { return axes_reduce(ReduceOp::sum, axes, mode); }
NCDevArray NCDevArray::max(const std::vector<int64_t>& axes) const { return axes_reduce(ReduceOp::max, axes, Accum::plain); }
NCDevArray NCDevArray::min(const std::vector<int64_t>& axes) const { return axes_reduce(ReduceOp::min, axes, Accum::plain); }
NCDevArray NCDevArray::argmax(const std::vector<int64_t>& axes) const { return axes_reduce(ReduceOp::argmax, axes, Accum::plain); }
NCDevArray NCDevArray::argmin(const std::vector<int64_t>& axes) const { return axes_reduce(ReduceOp::argmin, axes, Accum::plain); }
NCDevArray NCDevArray::mean(const std::vector<int64_t>& axes) const { return axes_reduce(ReduceOp::mean, axes, Accum::compensated); }
NCDevArray NCDevArray::all(const std::vector<int64_t>& axes) const { return axes_reduce(ReduceOp::all, axes, Accum::plain); }
NCDevArray NCDevArray::any(const std::vector<int64_t>& axes) const { return axes_reduce(ReduceOp::any, axes, Accum::plain); }

NCDevArray NCDevArray::var(const std::vector<int64_t>& axes, int64_t ddof) const {
  NCDevArray m = mean(axes);  // float64, kept axes
  // The mean broadcast back over the reduced axes (stride 0).
  Shape bst(static_cast<size_t>(nd_), 0);
  int64_t count = 1;
  int k = 0;
  for (int d = 0; d < nd_; ++d) {
    if (std::find(axes.begin(), axes.end(), static_cast<int64_t>(d)) == axes.end()) bst[d] = m.strides()[k++];
    else count *= sh_[d];
  }
  NCDevArray mb = wrap(m.data(), shape_vec(), bst, m.dtype(), m.storage());
  NCDevArray dev = NCDevArray(*this - mb);  // float64 deviations
  Expr sq = dev * dev;
  NCDevArray sqa(sq);
  NCDevArray ss = sqa.sum(axes);
  ss /= static_cast<double>(count - ddof);
  return ss;
}
NCDevArray NCDevArray::std(const std::vector<int64_t>& axes, int64_t ddof) const {
  NCDevArray v = var(axes, ddof);
  v = sqrt(v);
  return v;
}

void NCDevArray::scatter_add(const NCDevArray& idx, const NCDevArray& src) {
  HIPEngine::execute_scatter_add(*this, idx, src);
}

NCDevArray& NCDevArray::inplace(OpCode op, const Operand& r) {
  Expr e(*this);
  e.binary(op, r);
  HIPEngine::execute_binary_expression(e, *this);
  return *this;
}

NCDevArray& NCDevArray::inplace_unary(OpCode op) {
  Expr e(*this);
  e.unary(op);
  HIPEngine::execute_binary_expression(e, *this);
  return *this;
}

// ---- HIPEngine -----------------------------------------------------------------------------
void HIPEngine::execute_binary_expression(const Expr& e, const NCDevArray& result, void* stream) {
  ensure_init();
  if (e.instrs.empty()) throw Error(NCA_ERR_INVALID, "empty expression");
  const std::vector<int64_t> rshape = result.shape_vec();
  // operands must broadcast to the result (the reference requires equal shapes)
  if (broadcast_shapes(e.final_shape, rshape) != rshape)
    throw Error(NCA_ERR_INVALID, "expression shape does not match the destination shape");
  std::vector<int32_t> adt, sdt;
  const uint64_t h = expr_handle(e, result.dtype(), adt, sdt);
  std::vector<nca_view> ins;
  ins.reserve(e.views.size());
  for (const nca_view& v : e.views) ins.push_back(broadcast_view(v, rshape));
  std::vector<nca_scalar> sc = detail::pack_scalars(e.scalars, nullptr);
  const nca_view out = result.cview();
  check(nca_hip_expr_launch(h, ins.data(), static_cast<int32_t>(ins.size()), sc.data(),
                            static_cast<int32_t>(e.scalars.size()), &out, stream),
        "expression launch");
}

void HIPEngine::execute_reduce_axes(const NCDevArray& in, const std::vector<int64_t>& axes, ReduceOp op,
                                    const NCDevArray& result, Accum mode, void* stream) {
  ensure_init();
  const nca_view v = in.cview(), o = result.cview();
  check(nca_hip_reduce_axes(&v, axes.data(), static_cast<int32_t>(axes.size()), static_cast<int32_t>(op),
                            static_cast<int32_t>(mode), &o, stream),
        "axis reduction");
}

Scalar HIPEngine::execute_full_reduce(const NCDevArray& in, ReduceOp op, Accum mode) {
  ensure_init();
  const DType rdt = reduce_result_dtype(op, in.dtype());
  void* slot = result_slot();
  const nca_view v = in.cview();
  check(nca_hip_reduce_full(&v, static_cast<int32_t>(op), static_cast<int32_t>(mode), slot, nullptr), "full reduction");
  alignas(16) char host[16];
  check(nca_hip_memcpy_d2h(host, slot, static_cast<size_t>(itemsize(rdt)), nullptr), "result copy");
  return scalar_from_bytes(rdt, host);
}

void HIPEngine::execute_fill(const NCDevArray& arr, const Scalar& v, void* stream) {
  if (arr.size() == 0) return;
  Expr e(v);
  e.final_shape = arr.shape_vec();
  execute_binary_expression(e, arr, stream);
}

void HIPEngine::execute_assign(const NCDevArray& dest, const NCDevArray& src, void* stream) {
  if (dest.size() == 0) return;
  Expr e(src);
  execute_binary_expression(e, dest, stream);
}

void HIPEngine::execute_copy_into(const NCDevArray& src, void* host, DType host_dtype) {
  ensure_init();
  if (src.size() == 0) return;
  if (host_dtype == src.dtype() && src.is_contiguous()) {
    check(nca_hip_memcpy_d2h(host, src.data(), static_cast<size_t>(src.nbytes()), nullptr), "device to host copy");
    return;
  }
  NCDevArray tmp(src.shape_vec(), host_dtype);
  tmp.assign(src);
  check(nca_hip_memcpy_d2h(host, tmp.data(), static_cast<size_t>(tmp.nbytes()), nullptr), "device to host copy");
}

void HIPEngine::execute_scatter_add(const NCDevArray& dest, const NCDevArray& idx, const NCDevArray& src, void* stream) {
  ensure_init();
  // pointer-axis operands are read and written in place (the kernel follows their row pointers)
  const nca_view d = dest.cview(), i = idx.cview(), s = src.cview();
  check(nca_hip_scatter_add(&d, &i, &s, stream), "scatter_add");
}

// ---- stencils ------------------------------------------------------------------------------
namespace detail {
uint64_t stencil_create(int ndim, const std::vector<int64_t>& offsets, DType src, const Expr& e,
                        const StencilJITExtensions& ext, const std::vector<uint8_t>& is_pointer_axis,
                        bool is_soarr) {
  ensure_init();
  const int n_off = static_cast<int>(offsets.size()) / ndim;
  const int n_loads = static_cast<int>(e.views.size());
  if (n_loads > n_off) throw Error(NCA_ERR_INVALID, "the stencil expression loads more values than there are offsets");
  if (n_loads == 0) throw Error(NCA_ERR_INVALID, "the stencil expression loads no source values");
  for (Instruction i : e.instrs)
    if (get_op(i) == OpCode::IDX) throw Error(NCA_ERR_UNSUPPORTED, "IDX inside a stencil expression");
  std::vector<int32_t> adt(static_cast<size_t>(n_loads), static_cast<int32_t>(src)), sdt;
  std::vector<nca_scalar> unused = pack_scalars(e.scalars, &sdt);
  (void)unused;
  nca_expr_desc d{};
  d.instrs = e.instrs.data();
  d.n_instrs = static_cast<int32_t>(e.instrs.size());
  d.n_arrays = n_loads;
  d.arr_dtypes = adt.data();
  d.n_scalars = static_cast<int32_t>(sdt.size());
  d.work_dtype = static_cast<int32_t>(e.work_dtype);
  d.scalar_dtypes = sdt.data();
  d.dest_dtype = static_cast<int32_t>(e.expr_dtype);
  nca_stencil_desc sd{};
  sd.ndim = ndim;
  sd.n_offsets = n_loads;  // the k-th load reads offsets[k]; later offsets are never read
  sd.offsets = offsets.data();
  sd.src_dtype = static_cast<int32_t>(src);
  sd.expr = &d;
  sd.extra_params = ext.extra_params.empty() ? nullptr : ext.extra_params.c_str();
  sd.prologue = ext.prologue.empty() ? nullptr : ext.prologue.c_str();
  sd.epilogue = ext.epilogue.empty() ? nullptr : ext.epilogue.c_str();
  sd.is_pointer_axis = is_pointer_axis.empty() ? nullptr : is_pointer_axis.data();
  sd.is_soarr = is_soarr ? 1 : 0;
  uint64_t h = 0;
  check(nca_hip_stencil_create(&sd, &h), "stencil create");
  return h;
}

void stencil_apply(uint64_t h, const std::vector<Scalar>& scalars, const NCDevArray& src, const NCDevArray& dest,
                   void* stream, Boundary b, const std::vector<void*>& ext_args) {
  ensure_init();
  std::vector<nca_scalar> sc = pack_scalars(scalars, nullptr);
  // pointer-axis sources and destinations are read and written in place by the kernels
  const nca_view s = src.cview(), d = dest.cview();
  check(nca_hip_stencil_apply(h, &s, nullptr, 0, sc.data(), static_cast<int32_t>(scalars.size()), &d,
                              static_cast<int32_t>(b), ext_args.empty() ? nullptr : ext_args.data(),
                              static_cast<int32_t>(ext_args.size()), stream),
        "stencil apply");
}

uint64_t stencil_adjoint(uint64_t h) {
  uint64_t a = 0;
  check(nca_hip_stencil_adjoint_create(h, &a), "stencil adjoint");
  return a;
}
}  // namespace detail

}  // namespace ncarray_amd
