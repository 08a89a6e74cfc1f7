// Host-side dtype table and type-promotion rules shared by libnca_hip and the C++
// array classes.  Codes and rules follow the reference type system:
//   DType enum            src/lib/ncarray/dtype.hh:56-80
//   promote_expr_types    src/lib/ncarray/expression/interface.hh:91-100 (larger enum wins)
//   determine_dtype_for_op src/lib/ncarray/op_code.hh:101-161
//   sum_type / diff_type  src/lib/ncarray/op_traits.hh:72-77, 380-410
//   truediv_type = double op_traits.hh:76
// Added here: bfloat16 and float16 storage types.  Arithmetic on them runs in float32
// (their "compute type"); for promotion they rank between int64 and float32.
#ifndef NCARRAY_HIP_DTYPE_HH
#define NCARRAY_HIP_DTYPE_HH

#include <cstdint>
#include <stdexcept>
#include <string>

#include "nca_hip.h"

namespace nca {

enum class DType : int32_t {
  bool_ = NCA_BOOL, char_ = NCA_CHAR, uint8 = NCA_UINT8, uint16 = NCA_UINT16,
  uint32 = NCA_UINT32, uint64 = NCA_UINT64, int8 = NCA_INT8, int16 = NCA_INT16,
  int32 = NCA_INT32, int64 = NCA_INT64, float32 = NCA_FLOAT32, float64 = NCA_FLOAT64,
  float128 = NCA_FLOAT128, complex64 = NCA_COMPLEX64, complex128 = NCA_COMPLEX128,
  complex256 = NCA_COMPLEX256, vfloat2 = NCA_VFLOAT2, vfloat3 = NCA_VFLOAT3,
  vfloat4 = NCA_VFLOAT4, vdouble2 = NCA_VDOUBLE2, vdouble3 = NCA_VDOUBLE3,
  vdouble4 = NCA_VDOUBLE4, bfloat16 = NCA_BFLOAT16, float16 = NCA_FLOAT16
};

struct DTypeInfo {
  const char* name;    // reference-style name
  const char* ctype;   // type spelled in generated device code
  int size;            // bytes
  bool is_float;       // floating point (incl. bf16/fp16)
  bool is_signed;      // signed integer or float
  bool supported;      // has device kernels here
};

inline const DTypeInfo& dtype_info(int32_t code) {
  static const DTypeInfo table[] = {
      {"bool", "bool", 1, false, false, true},
      {"char", "signed char", 1, false, true, true},
      {"uint8", "unsigned char", 1, false, false, true},
      {"uint16", "unsigned short", 2, false, false, true},
      {"uint32", "unsigned int", 4, false, false, true},
      {"uint64", "unsigned long long", 8, false, false, true},
      {"int8", "signed char", 1, false, true, true},
      {"int16", "short", 2, false, true, true},
      {"int32", "int", 4, false, true, true},
      {"int64", "long long", 8, false, true, true},
      {"float32", "float", 4, true, true, true},
      {"float64", "double", 8, true, true, true},
      {"float128", "long double", 16, true, true, false},
      {"complex64", "", 8, true, true, false},
      {"complex128", "", 16, true, true, false},
      {"complex256", "", 32, true, true, false},
      {"vfloat2", "", 8, true, true, false},
      {"vfloat3", "", 12, true, true, false},
      {"vfloat4", "", 16, true, true, false},
      {"vdouble2", "", 16, true, true, false},
      {"vdouble3", "", 24, true, true, false},
      {"vdouble4", "", 32, true, true, false},
      {"bfloat16", "nca_bf16", 2, true, true, true},
      {"float16", "nca_f16", 2, true, true, true},
  };
  if (code < 0 || code > NCA_FLOAT16) throw std::invalid_argument("unknown dtype code " + std::to_string(code));
  return table[code];
}

inline int32_t itemsize(int32_t code) { return dtype_info(code).size; }
inline bool is_float(int32_t code) { return dtype_info(code).is_float; }
inline bool is_storage_float(int32_t code) { return code == NCA_BFLOAT16 || code == NCA_FLOAT16; }
inline bool is_integral(int32_t code) { return code <= NCA_INT64; }  // includes bool/char
inline bool is_supported(int32_t code) {
  return code >= 0 && code <= NCA_FLOAT16 && dtype_info(code).supported;
}

// The dtype arithmetic on `code` is carried out in.
inline int32_t compute_dtype(int32_t code) { return is_storage_float(code) ? NCA_FLOAT32 : code; }

// Promotion rank: the reference promotes to the numerically larger DType enumerator
// (interface.hh:91-100).  bf16/fp16 slot in between int64 and float32.
inline double promote_rank(int32_t code) {
  if (code == NCA_BFLOAT16) return 9.25;
  if (code == NCA_FLOAT16) return 9.5;
  return static_cast<double>(code);
}

inline int32_t promote_expr_types(int32_t a, int32_t b) {
  if (a == b) return a;
  if (is_storage_float(a) && is_storage_float(b)) return NCA_FLOAT32;  // bf16 op fp16
  return promote_rank(a) > promote_rank(b) ? a : b;
}

// op_traits<T>::sum_type (op_traits.hh:72-77, 380-410).  Elementwise ADD on the
// storage floats keeps the storage type; reductions use reduce_sum_dtype below.
inline int32_t sum_dtype(int32_t t) {
  switch (t) {
    case NCA_INT8: case NCA_INT16: return NCA_INT64;
    case NCA_UINT8: case NCA_UINT16: case NCA_BOOL: return NCA_UINT64;
    default: return t;
  }
}
inline int32_t diff_dtype(int32_t t) {
  switch (t) {
    case NCA_INT8: case NCA_INT16: case NCA_UINT8: case NCA_UINT16: case NCA_BOOL:
      return NCA_INT64;
    default: return t;
  }
}

// Result dtype of a reduction (SumTraits::OutT etc., reductions.hh:109-341).  Sums of
// bf16/fp16 accumulate and return float32.  MEAN returns truediv_type (float64).
inline int32_t reduce_result_dtype(int32_t op, int32_t src) {
  switch (op) {
    case NCA_REDUCE_SUM: return is_storage_float(src) ? NCA_FLOAT32 : sum_dtype(src);
    case NCA_REDUCE_MAX: case NCA_REDUCE_MIN: return src;
    case NCA_REDUCE_ARGMAX: case NCA_REDUCE_ARGMIN: return NCA_INT64;
    case NCA_REDUCE_MEAN: return NCA_FLOAT64;
    case NCA_REDUCE_ALL: case NCA_REDUCE_ANY: return NCA_BOOL;
    default: throw std::invalid_argument("unknown reduction op " + std::to_string(op));
  }
}

inline bool is_unary_op(int32_t op) {
  return (op >= NCA_OP_NEG && op <= NCA_OP_BNOT) || (op >= NCA_OP_EXP && op <= NCA_OP_ABS);
}
inline bool is_binary_op(int32_t op) { return op >= NCA_OP_ADD && op <= NCA_OP_RSHFT; }
inline bool is_load_op(int32_t op) {
  return op == NCA_OP_IDX || op == NCA_OP_LOAD_NCARR || op == NCA_OP_LOAD_SOARR ||
         op == NCA_OP_LOAD_CONST;
}

// determine_dtype_for_op (op_code.hh:101-161), extended with the unary math ops: they
// keep floating types and send integers to float64 (as true division does).
inline int32_t op_result_dtype(int32_t op, int32_t t) {
  switch (op) {
    case NCA_OP_NEG: case NCA_OP_INC: case NCA_OP_DEC: case NCA_OP_BNOT: return t;
    case NCA_OP_LNOT: return NCA_BOOL;
    case NCA_OP_ADD: return is_storage_float(t) ? t : sum_dtype(t);
    case NCA_OP_SUB: return is_storage_float(t) ? t : diff_dtype(t);
    case NCA_OP_MUL: case NCA_OP_MOD: return t;
    case NCA_OP_DIV: return NCA_FLOAT64;  // truediv_type = double for every scalar type
    case NCA_OP_EQ: case NCA_OP_NE: case NCA_OP_LT: case NCA_OP_LE: case NCA_OP_GT:
    case NCA_OP_GE: case NCA_OP_LAND: case NCA_OP_LOR: return NCA_BOOL;
    case NCA_OP_EXP: case NCA_OP_LOG: case NCA_OP_SQRT:
      return is_float(t) ? t : NCA_FLOAT64;
    case NCA_OP_ABS: return t;
    default: return t;
  }
}

inline const char* op_name(int32_t op) {
  switch (op) {
    case NCA_OP_NOOP: return "NOOP"; case NCA_OP_IDX: return "IDX";
    case NCA_OP_LOAD_NCARR: return "LOAD_NCARR"; case NCA_OP_LOAD_SOARR: return "LOAD_SOARR";
    case NCA_OP_LOAD_CONST: return "LOAD_CONST"; case NCA_OP_NEG: return "NEG";
    case NCA_OP_INC: return "INC"; case NCA_OP_DEC: return "DEC"; case NCA_OP_SZOF: return "SZOF";
    case NCA_OP_ADDR: return "ADDR"; case NCA_OP_INDR: return "INDR"; case NCA_OP_CAST: return "CAST";
    case NCA_OP_LNOT: return "LNOT"; case NCA_OP_BNOT: return "BNOT"; case NCA_OP_ADD: return "ADD";
    case NCA_OP_SUB: return "SUB"; case NCA_OP_MUL: return "MUL"; case NCA_OP_DIV: return "DIV";
    case NCA_OP_MOD: return "MOD"; case NCA_OP_FDIV: return "FDIV"; case NCA_OP_EQ: return "EQ";
    case NCA_OP_NE: return "NE"; case NCA_OP_LT: return "LT"; case NCA_OP_LE: return "LE";
    case NCA_OP_GT: return "GT"; case NCA_OP_GE: return "GE"; case NCA_OP_LAND: return "LAND";
    case NCA_OP_LOR: return "LOR"; case NCA_OP_BAND: return "BAND"; case NCA_OP_BOR: return "BOR";
    case NCA_OP_XOR: return "XOR"; case NCA_OP_LSHFT: return "LSHFT"; case NCA_OP_RSHFT: return "RSHFT";
    case NCA_OP_EXP: return "EXP"; case NCA_OP_LOG: return "LOG"; case NCA_OP_SQRT: return "SQRT";
    case NCA_OP_ABS: return "ABS";
    default: return "?";
  }
}

inline uint32_t pack_instruction(int32_t op, int32_t idx) {
  return (static_cast<uint32_t>(op) << 24) | (static_cast<uint32_t>(idx) & 0xFFFFFFu);
}
inline int32_t instr_op(uint32_t ins) { return static_cast<int32_t>(ins >> 24); }
inline int32_t instr_index(uint32_t ins) { return static_cast<int32_t>(ins & 0xFFFFFFu); }

}  // namespace nca

#endif  // NCARRAY_HIP_DTYPE_HH
