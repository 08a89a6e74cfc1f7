// ncarray_amd.hh -- C++ host API over the device C ABI (include/nca_hip.h).
//
// Mirrors the reference's host-side surface for the device hot path so that C++ callers
// keep their code shape:
//   ArrayImpl<NCOffsetsPolicy, DevOwnerPolicy> / NCDevArray   (array_impl.hh, ncdevarrays.cuh)
//   ExprMVNode<DevTag> + build_node + operators                (expression/mvnode.hh:82-459)
//   GPUEngine static entry points                              (engines/gpuengine.hh:64-638)
//   Stencil<NDIM>::create / apply                              (expression/stencil.hh:76-243)
//   Scalar (a std::variant of the element types)               (dtype.hh:265-273)
// Every computation is a call into libnca_hip.so; there is no host compute path here.
// Failures are thrown as ncarray_amd::Error (a std::runtime_error), as the reference's
// engines throw std::runtime_error.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <variant>
#include <vector>

#include "nca_hip.h"

namespace ncarray_amd {

// ---- types ------------------------------------------------------------------------------
// DType enumerators in the reference's order (dtype.hh:56-80) plus two storage types.
enum class DType : int32_t {
  bool_ = NCA_BOOL, char_ = NCA_CHAR, uint8 = NCA_UINT8, uint16 = NCA_UINT16, uint32 = NCA_UINT32,
  uint64 = NCA_UINT64, int8 = NCA_INT8, int16 = NCA_INT16, int32 = NCA_INT32, int64 = NCA_INT64,
  float32 = NCA_FLOAT32, float64 = NCA_FLOAT64, float128 = NCA_FLOAT128, complex64 = NCA_COMPLEX64,
  complex128 = NCA_COMPLEX128, complex256 = NCA_COMPLEX256, vfloat2 = NCA_VFLOAT2, vfloat3 = NCA_VFLOAT3,
  vfloat4 = NCA_VFLOAT4, vdouble2 = NCA_VDOUBLE2, vdouble3 = NCA_VDOUBLE3, vdouble4 = NCA_VDOUBLE4,
  bfloat16 = NCA_BFLOAT16, float16 = NCA_FLOAT16
};

// OpCode enumerators in the reference's order (op_code.hh:47-87) plus unary math.
enum class OpCode : int32_t {
  NOOP = NCA_OP_NOOP, IDX = NCA_OP_IDX, LOAD_NCARR = NCA_OP_LOAD_NCARR, LOAD_SOARR = NCA_OP_LOAD_SOARR,
  LOAD_CONST = NCA_OP_LOAD_CONST, NEG = NCA_OP_NEG, INC = NCA_OP_INC, DEC = NCA_OP_DEC, SZOF = NCA_OP_SZOF,
  ADDR = NCA_OP_ADDR, INDR = NCA_OP_INDR, CAST = NCA_OP_CAST, LNOT = NCA_OP_LNOT, BNOT = NCA_OP_BNOT,
  ADD = NCA_OP_ADD, SUB = NCA_OP_SUB, MUL = NCA_OP_MUL, DIV = NCA_OP_DIV, MOD = NCA_OP_MOD, FDIV = NCA_OP_FDIV,
  EQ = NCA_OP_EQ, NE = NCA_OP_NE, LT = NCA_OP_LT, LE = NCA_OP_LE, GT = NCA_OP_GT, GE = NCA_OP_GE,
  LAND = NCA_OP_LAND, LOR = NCA_OP_LOR, BAND = NCA_OP_BAND, BOR = NCA_OP_BOR, XOR = NCA_OP_XOR,
  LSHFT = NCA_OP_LSHFT, RSHFT = NCA_OP_RSHFT, EXP = NCA_OP_EXP, LOG = NCA_OP_LOG, SQRT = NCA_OP_SQRT,
  ABS = NCA_OP_ABS
};

enum class ReduceOp : int32_t {
  sum = NCA_REDUCE_SUM, max = NCA_REDUCE_MAX, min = NCA_REDUCE_MIN, argmax = NCA_REDUCE_ARGMAX,
  argmin = NCA_REDUCE_ARGMIN, mean = NCA_REDUCE_MEAN, all = NCA_REDUCE_ALL, any = NCA_REDUCE_ANY
};
enum class Accum : int32_t { plain = NCA_ACCUM_PLAIN, compensated = NCA_ACCUM_COMPENSATED };
enum class Boundary : int32_t { none = NCA_BOUNDARY_NONE, zero = NCA_BOUNDARY_ZERO };

using Instruction = uint32_t;  // (op << 24) | index, op_code.hh:174-192
inline Instruction pack_instruction(OpCode op, int idx) {
  return (static_cast<uint32_t>(op) << 24) | (static_cast<uint32_t>(idx) & 0xFFFFFFu);
}
inline OpCode get_op(Instruction i) { return static_cast<OpCode>(i >> 24); }
inline int get_index(Instruction i) { return static_cast<int>(i & 0xFFFFFFu); }

// The element types with device kernels, in the reference variant's order.
using Scalar = std::variant<bool, char, uint8_t, uint16_t, uint32_t, uint64_t, int8_t, int16_t, int32_t,
                            int64_t, float, double>;

template <typename T> struct dtype_traits;
#define NCA_AMD_DTYPE_TRAIT(T, D) \
  template <> struct dtype_traits<T> { static constexpr DType value = DType::D; };
NCA_AMD_DTYPE_TRAIT(bool, bool_)
NCA_AMD_DTYPE_TRAIT(char, char_)
NCA_AMD_DTYPE_TRAIT(uint8_t, uint8)
NCA_AMD_DTYPE_TRAIT(uint16_t, uint16)
NCA_AMD_DTYPE_TRAIT(uint32_t, uint32)
NCA_AMD_DTYPE_TRAIT(unsigned long, uint64)
NCA_AMD_DTYPE_TRAIT(unsigned long long, uint64)
NCA_AMD_DTYPE_TRAIT(int8_t, int8)
NCA_AMD_DTYPE_TRAIT(int16_t, int16)
NCA_AMD_DTYPE_TRAIT(int32_t, int32)
NCA_AMD_DTYPE_TRAIT(long, int64)
NCA_AMD_DTYPE_TRAIT(long long, int64)
NCA_AMD_DTYPE_TRAIT(float, float32)
NCA_AMD_DTYPE_TRAIT(double, float64)
#undef NCA_AMD_DTYPE_TRAIT

const char* dtype_name(DType d);
int itemsize(DType d);
bool is_float(DType d);
bool is_signed(DType d);
DType scalar_dtype(const Scalar& s);
Scalar scalar_from_bytes(DType d, const void* p);  // one element of dtype d (bf16/fp16 widen to float)
double scalar_to_double(const Scalar& s);
Scalar scalar_cast(const Scalar& s, DType d);      // C conversion of the value to d
std::string scalar_repr(const Scalar& s);
// The result dtype of a reduction op on a source dtype (op_traits sum_type / truediv_type).
DType reduce_result_dtype(ReduceOp op, DType src);

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// Throws Error carrying nca_hip_last_error() when status != NCA_OK.
void check(int status, const char* what);

// Select the device for the calling thread (idempotent).  Every array operation calls it
// for device 0 if the thread has not selected one.
void init(int device = 0);

// ---- indexing (indexing.hh) ---------------------------------------------------------------
struct Slice {
  static constexpr int64_t kNone = INT64_MIN;
  int64_t start = kNone, stop = kNone, step = 1;
  Slice() = default;
  Slice(int64_t start_, int64_t stop_, int64_t step_ = 1) : start(start_), stop(stop_), step(step_) {}
};
struct Ellipsis {};
struct NewAxis {};
using IndexItem = std::variant<int64_t, Slice, Ellipsis, NewAxis>;

// ---- device storage ------------------------------------------------------------------------
// A device allocation (owned: freed with nca_hip_free) or a borrowed pointer kept valid by
// `keep` (e.g. a framework tensor).
class Storage {
 public:
  explicit Storage(size_t bytes);
  Storage(void* borrowed, std::shared_ptr<void> keep) : ptr_(borrowed), keep_(std::move(keep)) {}
  ~Storage();
  Storage(const Storage&) = delete;
  Storage& operator=(const Storage&) = delete;
  void* ptr() const { return ptr_; }
  // Keep `k` alive as long as this storage (the rows behind a row-pointer table).
  void hold(std::shared_ptr<void> k) { keep_ = std::move(k); }

 private:
  void* ptr_ = nullptr;
  bool owned_ = false;
  std::shared_ptr<void> keep_;
};

class Expr;
class NCDevArray;

// An operand of an expression operator: an array, an expression or a scalar.
struct Operand {
  std::variant<const NCDevArray*, const Expr*, Scalar> v;
  Operand(const NCDevArray& a) : v(&a) {}
  Operand(const Expr& e) : v(&e) {}
  Operand(const Scalar& s) : v(s) {}
  template <typename T, typename = std::enable_if_t<std::is_arithmetic_v<T>>>
  Operand(T x) : v(make(x)) {}

 private:
  template <typename T>
  static Scalar make(T x) {
    if constexpr (std::is_same_v<T, long> || std::is_same_v<T, long long>) return Scalar(static_cast<int64_t>(x));
    else if constexpr (std::is_same_v<T, unsigned long> || std::is_same_v<T, unsigned long long>)
      return Scalar(static_cast<uint64_t>(x));
    else if constexpr (std::is_same_v<T, signed char>) return Scalar(static_cast<int8_t>(x));
    else if constexpr (std::is_same_v<T, unsigned char>) return Scalar(static_cast<uint8_t>(x));
    else return Scalar(x);
  }
};

// ---- expression tape (ExprMVNode<DevTag>, mvnode.hh:82-459) --------------------------------
class Expr {
 public:
  std::vector<nca_view> views;                    // operand layouts (layouts[] + data[])
  std::vector<std::shared_ptr<Storage>> keep;     // keeps operand storage alive
  std::vector<DType> dtypes;                      // operand dtypes
  std::vector<Instruction> instrs;                // RPN tape
  std::vector<Scalar> scalars;                    // LOAD_CONST table
  std::vector<int64_t> final_shape;               // shape of the evaluated expression
  DType expr_dtype = DType::float32;              // dtype of the evaluated expression
  DType work_dtype = DType::float32;              // promoted operand dtype
  bool has_expr_dtype = false;

  Expr() = default;
  Expr(const NCDevArray& a) { build_node(a); }
  Expr(const Scalar& s) { build_node(s); }

  void build_node(const NCDevArray& a);
  void build_node(const Expr& e);
  void build_node(const Scalar& s);
  void build_node(const Operand& o);

  // operator<op>(right): appends right's tape and the op (mvnode.hh:289-441).
  Expr& binary(OpCode op, const Operand& right);
  // unary ops (mvnode.hh:263-287) and the math functions added here.
  Expr& unary(OpCode op);

  DType dtype() const { return expr_dtype; }
  const std::vector<int64_t>& shape() const { return final_shape; }
  int ndim() const { return static_cast<int>(final_shape.size()); }
  int64_t size() const;

  // Evaluate into a new array of the expression dtype (or `as`).
  NCDevArray materialize(std::optional<DType> as = std::nullopt) const;
  // Reductions of the evaluated values (in the expression dtype) in the same pass; the
  // expression is not materialised.
  Scalar sum(Accum mode = Accum::compensated) const;
  Scalar max() const;
  Scalar min() const;
};

Expr operator+(const Operand& a, const Operand& b);
Expr operator-(const Operand& a, const Operand& b);
Expr operator*(const Operand& a, const Operand& b);
Expr operator/(const Operand& a, const Operand& b);
Expr operator%(const Operand& a, const Operand& b);
Expr operator==(const Operand& a, const Operand& b);
Expr operator!=(const Operand& a, const Operand& b);
Expr operator<(const Operand& a, const Operand& b);
Expr operator<=(const Operand& a, const Operand& b);
Expr operator>(const Operand& a, const Operand& b);
Expr operator>=(const Operand& a, const Operand& b);
Expr operator&&(const Operand& a, const Operand& b);
Expr operator||(const Operand& a, const Operand& b);
Expr operator&(const Operand& a, const Operand& b);
Expr operator|(const Operand& a, const Operand& b);
Expr operator^(const Operand& a, const Operand& b);
Expr operator<<(const Operand& a, const Operand& b);
Expr operator>>(const Operand& a, const Operand& b);
Expr operator-(const Operand& a);
Expr operator!(const Operand& a);
Expr operator~(const Operand& a);
Expr exp(const Operand& a);
Expr log(const Operand& a);
Expr sqrt(const Operand& a);
Expr abs(const Operand& a);
// Floor division (OpCode::FDIV).
Expr floordiv(const Operand& a, const Operand& b);

// ---- device array (ArrayImpl<NCOffsetsPolicy, DevOwnerPolicy/DevViewPolicy>) ---------------
// Copies of an NCDevArray share storage (a view); copy() makes a new owner.
class NCDevArray {
 public:
  using Shape = std::vector<int64_t>;

  NCDevArray() = default;
  explicit NCDevArray(const Shape& shape, DType dtype = DType::float32);
  NCDevArray(int ndim, const int64_t* shape, DType dtype);
  // NCOwnerFor(expr): a new array holding the evaluated expression.
  NCDevArray(const Expr& e);
  // A view of device memory owned elsewhere (strides in bytes).
  static NCDevArray wrap(void* data, const Shape& shape, const Shape& strides_bytes, DType dtype,
                         std::shared_ptr<void> keep = {});
  // NCDevArrayRef over separately allocated rows (the list constructor, pylist_to_ref,
  // utilities.hh:273-314): axis 0 is a pointer axis over `rows`, which must share one shape,
  // stride layout and dtype.  The row-pointer table is copied to device memory; the rows are
  // kept alive by the result.
  static NCDevArray from_rows(const std::vector<NCDevArray>& rows);
  // The same over raw device row pointers with one row layout (ArrayImpl's vector<void*>
  // constructor, array_impl.hh:388-432, pointer axis 0); `keep` holds the rows alive.
  static NCDevArray from_row_pointers(const std::vector<void*>& rows, const Shape& row_shape,
                                      const Shape& row_strides_bytes, DType dtype, std::shared_ptr<void> keep = {});

  // ArrayImpl::operator=(Expr) (array_operations.hh:42-63): evaluate into this array's
  // storage and dtype.  An empty array is first allocated with the expression's shape/dtype.
  NCDevArray& operator=(const Expr& e);

  // ---- metadata
  int ndim() const { return nd_; }
  const int64_t* shape() const { return sh_.data(); }
  int64_t shape(int d) const { return sh_[static_cast<size_t>(d)]; }
  const int64_t* strides() const { return st_.data(); }  // bytes
  Shape shape_vec() const { return Shape(sh_.begin(), sh_.begin() + nd_); }
  Shape strides_vec() const { return Shape(st_.begin(), st_.begin() + nd_); }
  int64_t size() const;
  int itemsize() const { return ncarray_amd::itemsize(dt_); }
  int64_t nbytes() const { return size() * itemsize(); }
  DType dtype() const { return dt_; }
  bool is_contiguous() const;  // false for pointer-axis arrays (array_impl.hh:1090-1107)
  int pointer_axis() const { return pax_; }
  bool is_pointer_axis(int d) const { return d == pax_; }
  // Row pointer at pointer-axis index i (of a pointer-axis array), offset by the row offset.
  void* row(int64_t i) const;
  // This array, or for a pointer-axis array a new C-order owner with its values (the staged
  // copy that operations without pointer-axis kernels read).
  NCDevArray dense() const;
  bool empty() const { return storage_ == nullptr; }
  void* data() const { return data_; }
  const std::shared_ptr<Storage>& storage() const { return storage_; }
  nca_view cview() const;  // the C-ABI view
  NCDevArray view() const { return *this; }
  std::string repr() const;

  // ---- indexing (array_impl.hh view_from_indices / operator())
  NCDevArray view_from_indices(const std::vector<IndexItem>& idx) const;
  template <typename... I>
  NCDevArray operator()(I... items) const { return view_from_indices({IndexItem(conv(items))...}); }
  NCDevArray reshape(const Shape& shape) const;        // contiguous arrays only (no copy)
  NCDevArray copy_as_shape(const Shape& shape) const;  // copy into a new C-order owner
  NCDevArray squeeze() const;
  NCDevArray copy() const;
  NCDevArray astype(DType d) const;
  Expr iota() const;  // expression of C-order linear indices (array_operations.hh:65-76)
  // A new array holding its C-order linear indices (immediate evaluation).
  static NCDevArray iota(const Shape& shape, DType dtype = DType::int64);

  // ---- data movement (GPUEngine::execute_fill / execute_assign / execute_copy_into)
  void fill(const Scalar& v);
  void assign(const NCDevArray& src);  // broadcast + cast into this view
  void assign(const Expr& e) { *this = e; }
  // Assign from host memory: `host` holds an array of `host_shape` (broadcastable to this
  // shape) with byte strides `host_strides` (empty = C order) and dtype `host_dtype`.
  void assign_from_host(const void* host, DType host_dtype, const Shape& host_shape,
                        const Shape& host_strides = {});
  // copy_into / copy_into_astype: write the elements in C order into host memory.
  void copy_into(void* host) const { copy_into_astype(host, dt_); }
  void copy_into_astype(void* host, DType host_dtype) const;
  template <typename T>
  void copy_into_astype(T* host) const { copy_into_astype(host, dtype_traits<T>::value); }
  Scalar item() const;  // the single element of a size-1 array

  // ---- reductions (array_operations.hh:512-1092)
  Scalar sum(Accum mode = Accum::compensated) const;
  Scalar max() const;
  Scalar min() const;
  Scalar argmax() const;
  Scalar argmin() const;
  Scalar mean() const;
  Scalar var(int64_t ddof = 0) const;
  Scalar std(int64_t ddof = 0) const;
  Scalar all() const;
  Scalar any() const;
  NCDevArray sum(const std::vector<int64_t>& axes, Accum mode = Accum::compensated) const;
  NCDevArray max(const std::vector<int64_t>& axes) const;
  NCDevArray min(const std::vector<int64_t>& axes) const;
  NCDevArray argmax(const std::vector<int64_t>& axes) const;
  NCDevArray argmin(const std::vector<int64_t>& axes) const;
  NCDevArray mean(const std::vector<int64_t>& axes) const;
  NCDevArray var(const std::vector<int64_t>& axes, int64_t ddof = 0) const;
  NCDevArray std(const std::vector<int64_t>& axes, int64_t ddof = 0) const;
  NCDevArray all(const std::vector<int64_t>& axes) const;
  NCDevArray any(const std::vector<int64_t>& axes) const;

  // this[idx[i]] += src[i] for 0 <= idx[i] < size(), others skipped (kernels.cuh:660-680)
  void scatter_add(const NCDevArray& idx, const NCDevArray& src);

  // ---- in-place ops: evaluate `this op right` into this array (array_operations.hh:95-339)
  NCDevArray& operator+=(const Operand& r) { return inplace(OpCode::ADD, r); }
  NCDevArray& operator-=(const Operand& r) { return inplace(OpCode::SUB, r); }
  NCDevArray& operator*=(const Operand& r) { return inplace(OpCode::MUL, r); }
  NCDevArray& operator/=(const Operand& r) { return inplace(OpCode::DIV, r); }
  NCDevArray& operator%=(const Operand& r) { return inplace(OpCode::MOD, r); }
  NCDevArray& operator&=(const Operand& r) { return inplace(OpCode::LAND, r); }
  NCDevArray& operator|=(const Operand& r) { return inplace(OpCode::LOR, r); }
  NCDevArray& operator++() { return inplace_unary(OpCode::INC); }
  NCDevArray& operator--() { return inplace_unary(OpCode::DEC); }

 private:
  static IndexItem conv(const IndexItem& i) { return i; }
  static IndexItem conv(const Slice& s) { return IndexItem(s); }
  template <typename T, typename = std::enable_if_t<std::is_integral_v<T>>>
  static IndexItem conv(T i) { return IndexItem(static_cast<int64_t>(i)); }
  NCDevArray& inplace(OpCode op, const Operand& r);
  NCDevArray& inplace_unary(OpCode op);
  Scalar full(ReduceOp op, Accum mode) const;
  NCDevArray axes_reduce(ReduceOp op, const std::vector<int64_t>& axes, Accum mode) const;

  std::shared_ptr<Storage> storage_;
  void* data_ = nullptr;
  DType dt_ = DType::float32;
  int nd_ = 0;
  std::array<int64_t, NCA_MAX_NDIM> sh_{};
  std::array<int64_t, NCA_MAX_NDIM> st_{};
  // NC pointer axis (NCOffsetsPolicy, layout.hh:451-480): data_ points into the device
  // row-pointer table, st_[pax_] counts table entries, poff_ is the table offset (the
  // reference's m_offsets[P]) and roff_ the byte offset added to each row pointer.
  int pax_ = -1;
  int64_t poff_ = 0;
  int64_t roff_ = 0;
  std::shared_ptr<const std::vector<void*>> rows_;  // host copy of the table
};

// ---- engine (GPUEngine, gpuengine.hh:64-638) ------------------------------------------------
// stream: a hipStream_t as void*; nullptr = legacy default stream, synchronised on return.
struct HIPEngine {
  static void execute_binary_expression(const Expr& e, const NCDevArray& result, void* stream = nullptr);
  static void execute_reduce_axes(const NCDevArray& in, const std::vector<int64_t>& axes, ReduceOp op,
                                  const NCDevArray& result, Accum mode = Accum::compensated,
                                  void* stream = nullptr);
  static Scalar execute_full_reduce(const NCDevArray& in, ReduceOp op, Accum mode = Accum::compensated);
  static void execute_fill(const NCDevArray& arr, const Scalar& v, void* stream = nullptr);
  static void execute_assign(const NCDevArray& dest, const NCDevArray& src, void* stream = nullptr);
  static void execute_copy_into(const NCDevArray& src, void* host, DType host_dtype);
  static void execute_scatter_add(const NCDevArray& dest, const NCDevArray& idx, const NCDevArray& src,
                                  void* stream = nullptr);
};

// ---- stencils (Stencil<NDIM>, expression/stencil.hh:76-243) ---------------------------------
// StencilJITExtensions (jit/device/extensions.hh:68-72): extra kernel parameters, code run
// before the expression (may declare variables) and code run after it (sees `result`,
// the output coordinate `c_dst[]` and `dst_ptr`).
struct StencilJITExtensions {
  std::string extra_params;
  std::string prologue;
  std::string epilogue;
};

template <int NDIM>
class Stencil {
 public:
  using Offset = std::array<int64_t, NDIM>;
  Stencil() = default;

  // op_func receives one placeholder array per offset (in offset order) and returns the
  // expression over them; the k-th load in tape order reads offsets[k]
  // (rtcompiler.hh:211-215).  is_pointer_axis (NDIM flags) and is_soarr declare the sources the
  // stencil serves, as in the reference (stencil.hh:76-81): apply then takes only sources whose
  // pointer axes are the flagged ones; an empty vector accepts any layout.  The kernels read
  // pointer-axis sources and destinations in place (one pointer read per row).
  template <typename SrcT, typename Func>
  static Stencil create(const std::vector<Offset>& offsets, const std::vector<uint8_t>& is_pointer_axis,
                        Func&& op_func, const StencilJITExtensions& ext = {}, bool is_soarr = false) {
    return create(dtype_traits<SrcT>::value, offsets, is_pointer_axis, std::forward<Func>(op_func), ext, is_soarr);
  }
  template <typename Func>
  static Stencil create(DType src_dtype, const std::vector<Offset>& offsets,
                        const std::vector<uint8_t>& is_pointer_axis, Func&& op_func,
                        const StencilJITExtensions& ext = {}, bool is_soarr = false) {
    if (!is_pointer_axis.empty() && is_pointer_axis.size() != static_cast<size_t>(NDIM))
      throw Error(NCA_ERR_INVALID, "is_pointer_axis needs one flag per axis");
    std::vector<NCDevArray> ph;
    const NCDevArray::Shape unit(NDIM, 1);
    for (size_t k = 0; k < offsets.size(); ++k) ph.push_back(NCDevArray::wrap(nullptr, unit, unit, src_dtype));
    Expr e = op_func(ph);
    std::vector<int64_t> flat;
    for (const Offset& o : offsets) flat.insert(flat.end(), o.begin(), o.end());
    Stencil s;
    s.handle_ = create_handle(NDIM, flat, src_dtype, e, ext, is_pointer_axis, is_soarr);
    s.scalars_ = e.scalars;
    s.dest_dtype_ = e.expr_dtype;
    s.has_ext_ = !ext.extra_params.empty();
    return s;
  }

  // dest[c] = f(src[c*step + offsets[k]]) with step = src.shape / dest.shape.
  // With no stream the call synchronises (stencil.hh:236-242).  ext_args point at the
  // values of the extra parameters, in order.
  void apply(const NCDevArray& src, const NCDevArray& dest, std::optional<void*> stream = std::nullopt,
             Boundary boundary = Boundary::none, const std::vector<void*>& ext_args = {}) const {
    apply_handle(handle_, scalars_, src, dest, stream ? *stream : nullptr, boundary, ext_args);
  }
  // The transposed stencil of a stencil linear in its loads.
  Stencil adjoint() const {
    Stencil a = *this;
    a.handle_ = adjoint_handle(handle_);
    return a;
  }
  DType dest_dtype() const { return dest_dtype_; }
  uint64_t handle() const { return handle_; }

 private:
  static uint64_t create_handle(int ndim, const std::vector<int64_t>& offsets, DType src, const Expr& e,
                                const StencilJITExtensions& ext, const std::vector<uint8_t>& is_pointer_axis,
                                bool is_soarr);
  static void apply_handle(uint64_t h, const std::vector<Scalar>& scalars, const NCDevArray& src,
                           const NCDevArray& dest, void* stream, Boundary b, const std::vector<void*>& ext_args);
  static uint64_t adjoint_handle(uint64_t h);

  uint64_t handle_ = 0;
  std::vector<Scalar> scalars_;
  DType dest_dtype_ = DType::float32;
  bool has_ext_ = false;
};

namespace detail {
uint64_t stencil_create(int ndim, const std::vector<int64_t>& offsets, DType src, const Expr& e,
                        const StencilJITExtensions& ext, const std::vector<uint8_t>& is_pointer_axis = {},
                        bool is_soarr = false);
void stencil_apply(uint64_t h, const std::vector<Scalar>& scalars, const NCDevArray& src, const NCDevArray& dest,
                   void* stream, Boundary b, const std::vector<void*>& ext_args);
uint64_t stencil_adjoint(uint64_t h);
std::vector<nca_scalar> pack_scalars(const std::vector<Scalar>& s, std::vector<int32_t>* dtypes);
}  // namespace detail

template <int NDIM>
uint64_t Stencil<NDIM>::create_handle(int ndim, const std::vector<int64_t>& offsets, DType src, const Expr& e,
                                      const StencilJITExtensions& ext, const std::vector<uint8_t>& is_pointer_axis,
                                      bool is_soarr) {
  return detail::stencil_create(ndim, offsets, src, e, ext, is_pointer_axis, is_soarr);
}
template <int NDIM>
void Stencil<NDIM>::apply_handle(uint64_t h, const std::vector<Scalar>& scalars, const NCDevArray& src,
                                 const NCDevArray& dest, void* stream, Boundary b,
                                 const std::vector<void*>& ext_args) {
  detail::stencil_apply(h, scalars, src, dest, stream, b, ext_args);
}
template <int NDIM>
uint64_t Stencil<NDIM>::adjoint_handle(uint64_t h) {
  return detail::stencil_adjoint(h);
}

}  // namespace ncarray_amd
