/*
 * nca_cpu.c -- CPU ORACLE (test infrastructure only; never linked into the product).
 *
 * A plain-C restatement of the reference's host-path semantics for the hot path, used
 * by tests/ to check the HIP library and by bench.py as the CPU baseline ("port").
 * Each function cites the reference file:line it follows (paths relative to the ncarray
 * source tree, src/lib/ncarray/).  Built with -O3 -ffp-contract=off so float32 arithmetic
 * is evaluated one IEEE operation at a time, as the reference's scalar SSE code does.
 *
 * Parity pins (see tests/test_oracle.py, and tests/test_oracle_ref.py against the reference's
 * own host routines compiled into oracle/_ref): the reference's own known-answer tests
 * (test_reductions.cc, test_gpu_reductions.cu, test_binary_operations.cc,
 * test_unary_operations.cc, tests/python/test_operations.py), closed forms
 * (sum of iota, Laplacian of polynomials), numpy (argmax first occurrence, bincount,
 * FP64 sums, slicing stencils) and brute force on tiny inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* dtype codes = reference DType order (dtype.hh:56-80) + bf16/fp16 */
enum { D_BOOL = 0, D_CHAR, D_U8, D_U16, D_U32, D_U64, D_I8, D_I16, D_I32, D_I64, D_F32, D_F64,
       D_BF16 = 22, D_F16 = 23 };
/* opcodes = reference OpCode order (op_code.hh:47-87) + unary math */
enum { OP_NOOP = 0, OP_IDX, OP_LOAD_NCARR, OP_LOAD_SOARR, OP_LOAD_CONST, OP_NEG, OP_INC, OP_DEC,
       OP_SZOF, OP_ADDR, OP_INDR, OP_CAST, OP_LNOT, OP_BNOT, OP_ADD, OP_SUB, OP_MUL, OP_DIV,
       OP_MOD, OP_FDIV, OP_EQ, OP_NE, OP_LT, OP_LE, OP_GT, OP_GE, OP_LAND, OP_LOR,
       OP_EXP = 64, OP_LOG, OP_SQRT, OP_ABS };

static float bf16_to_f32(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
/* round to nearest even; NaN stays NaN */
uint16_t nca_cpu_f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

/* value of element i of a typed array, as double (exact for every type but 64-bit ints
 * beyond 2^53, which the tests avoid) */
static double load_d(const void* p, int dt, int64_t i) {
  switch (dt) {
    case D_BOOL: return ((const uint8_t*)p)[i] != 0;
    case D_CHAR: case D_I8: return ((const int8_t*)p)[i];
    case D_U8: return ((const uint8_t*)p)[i];
    case D_U16: return ((const uint16_t*)p)[i];
    case D_I16: return ((const int16_t*)p)[i];
    case D_U32: return ((const uint32_t*)p)[i];
    case D_I32: return ((const int32_t*)p)[i];
    case D_U64: return (double)((const uint64_t*)p)[i];
    case D_I64: return (double)((const int64_t*)p)[i];
    case D_F32: return ((const float*)p)[i];
    case D_F64: return ((const double*)p)[i];
    case D_BF16: return bf16_to_f32(((const uint16_t*)p)[i]);
    default: return NAN;
  }
}

/* value of element i of an integer-typed (or bool) array, exact */
static int64_t load_i(const void* p, int dt, int64_t i) {
  switch (dt) {
    case D_BOOL: return ((const uint8_t*)p)[i] != 0;
    case D_CHAR: case D_I8: return ((const int8_t*)p)[i];
    case D_U8: return ((const uint8_t*)p)[i];
    case D_U16: return ((const uint16_t*)p)[i];
    case D_I16: return ((const int16_t*)p)[i];
    case D_U32: return ((const uint32_t*)p)[i];
    case D_I32: return ((const int32_t*)p)[i];
    case D_U64: return (int64_t)((const uint64_t*)p)[i];
    case D_I64: return ((const int64_t*)p)[i];
    default: return 0;
  }
}

/* ---- expression evaluation ---------------------------------------------------------
 * The reference VM (expression/dynamicmvnode.hh:146-251) evaluated with the operation
 * semantics of ExprOpInterface::apply_op (expression/interface.hh:129-181): loads are
 * cast to the evaluation type, unary ops rewrite the top of the stack, binary ops pop
 * two; DIV by zero gives NaN for a finite numerator and leaves a non-finite one as is.
 * EXP/LOG/SQRT/ABS call libm.  `eval_dtype` is F32 or F64 (evaluation in float or
 * double).  `raw_div` evaluates DIV as plain division (the stencil source generator,
 * jit/device/rtcompiler.hh:224-234). */
#define STACK_MAX 64

static float apply_f(int op, float a, float b, int raw_div) {
  switch (op) {
    case OP_NEG: return -a;
    case OP_INC: return a + 1.0f;
    case OP_DEC: return a - 1.0f;
    case OP_LNOT: return (float)(a == 0.0f);
    case OP_ADD: return a + b;
    case OP_SUB: return a - b;
    case OP_MUL: return a * b;
    case OP_DIV:
      if (!raw_div && b == 0.0f) return isfinite(a) ? NAN : a;
      return a / b;
    case OP_MOD: return fmodf(a, b);
    case OP_FDIV: return floorf(a / b);
    case OP_EQ: return (float)(a == b);
    case OP_NE: return (float)(a != b);
    case OP_LT: return (float)(a < b);
    case OP_LE: return (float)(a <= b);
    case OP_GT: return (float)(a > b);
    case OP_GE: return (float)(a >= b);
    case OP_LAND: return (float)(a != 0.0f && b != 0.0f);
    case OP_LOR: return (float)(a != 0.0f || b != 0.0f);
    case OP_EXP: return expf(a);
    case OP_LOG: return logf(a);
    case OP_SQRT: return sqrtf(a);
    case OP_ABS: return fabsf(a);
    default: return NAN;
  }
}

static double apply_d(int op, double a, double b, int raw_div) {
  switch (op) {
    case OP_NEG: return -a;
    case OP_INC: return a + 1.0;
    case OP_DEC: return a - 1.0;
    case OP_LNOT: return (double)(a == 0.0);
    case OP_ADD: return a + b;
    case OP_SUB: return a - b;
    case OP_MUL: return a * b;
    case OP_DIV:
      if (!raw_div && b == 0.0) return isfinite(a) ? NAN : a;
      return a / b;
    case OP_MOD: return fmod(a, b);
    case OP_FDIV: return floor(a / b);
    case OP_EQ: return (double)(a == b);
    case OP_NE: return (double)(a != b);
    case OP_LT: return (double)(a < b);
    case OP_LE: return (double)(a <= b);
    case OP_GT: return (double)(a > b);
    case OP_GE: return (double)(a >= b);
    case OP_LAND: return (double)(a != 0.0 && b != 0.0);
    case OP_LOR: return (double)(a != 0.0 || b != 0.0);
    case OP_EXP: return exp(a);
    case OP_LOG: return log(a);
    case OP_SQRT: return sqrt(a);
    case OP_ABS: return fabs(a);
    default: return NAN;
  }
}

static int is_binary(int op) { return op >= OP_ADD && op <= 32; }

/* Evaluate one element.  loads[k] is the value of array operand k at this element
 * (already a double holding the exact stored value), scalars[k] likewise. */
static double eval_one(const uint32_t* ins, int n_ins, const double* loads, const double* scalars, double idx,
                       int eval_dtype, int raw_div) {
  if (eval_dtype == D_F32) {
    float st[STACK_MAX];
    int top = -1;
    for (int i = 0; i < n_ins; ++i) {
      int op = (int)(ins[i] >> 24), k = (int)(ins[i] & 0xFFFFFF);
      if (op == OP_LOAD_NCARR || op == OP_LOAD_SOARR) st[++top] = (float)loads[k];
      else if (op == OP_LOAD_CONST) st[++top] = (float)scalars[k];
      else if (op == OP_IDX) st[++top] = (float)idx;
      else if (is_binary(op)) { float b = st[top--]; float a = st[top--]; st[++top] = apply_f(op, a, b, raw_div); }
      else st[top] = apply_f(op, st[top], 0.0f, raw_div);
    }
    return st[0];
  } else {
    double st[STACK_MAX];
    int top = -1;
    for (int i = 0; i < n_ins; ++i) {
      int op = (int)(ins[i] >> 24), k = (int)(ins[i] & 0xFFFFFF);
      if (op == OP_LOAD_NCARR || op == OP_LOAD_SOARR) st[++top] = loads[k];
      else if (op == OP_LOAD_CONST) st[++top] = scalars[k];
      else if (op == OP_IDX) st[++top] = idx;
      else if (is_binary(op)) { double b = st[top--]; double a = st[top--]; st[++top] = apply_d(op, a, b, raw_div); }
      else st[top] = apply_d(op, st[top], 0.0, raw_div);
    }
    return st[0];
  }
}

static void store_d(void* p, int dt, int64_t i, double v) {
  switch (dt) {
    case D_BOOL: ((uint8_t*)p)[i] = (v != 0.0); break;
    case D_I32: ((int32_t*)p)[i] = (int32_t)v; break;
    case D_I64: ((int64_t*)p)[i] = (int64_t)v; break;
    case D_F32: ((float*)p)[i] = (float)v; break;
    case D_F64: ((double*)p)[i] = v; break;
    case D_BF16: ((uint16_t*)p)[i] = nca_cpu_f32_to_bf16((float)v); break;
    default: break;
  }
}

static void expr_eval(const uint32_t* ins, int n_ins, const void* const* arrays, const int* arr_dtypes, int n_arrays,
                      const double* scalars, int eval_dtype, void* out, int out_dtype, int64_t n, int raw_div) {
  double loads[64];
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < n_arrays; ++k) loads[k] = load_d(arrays[k], arr_dtypes[k], i);
    double v = eval_one(ins, n_ins, loads, scalars, (double)i, eval_dtype, raw_div);
    store_d(out, out_dtype, i, v);
  }
}

/* out[i] = tape(arrays[k][i], scalars) for contiguous same-size arrays. */
void nca_cpu_expr(const uint32_t* ins, int n_ins, const void* const* arrays, const int* arr_dtypes, int n_arrays,
                  const double* scalars, int eval_dtype, void* out, int out_dtype, int64_t n) {
  expr_eval(ins, n_ins, arrays, arr_dtypes, n_arrays, scalars, eval_dtype, out, out_dtype, n, 0);
}

/* The reference's host x86 JIT for linearizable tapes (jit/host/rtcompiler.cc:285-610): a
 * left fold per element with one scalar SSE instruction per operation in the work type,
 * DIV as divss/divsd (jit/host/x86.cc:157-169), i.e. IEEE division: x/0 is +-inf for a
 * finite nonzero x and NaN for 0/0, where the VM and the device kernels give NaN for any
 * finite x, and +-inf/-0 flips the sign, where they return the numerator unchanged
 * (interface.hh:150-156).  For a linearizable tape (RPN depth at most 2) the left
 * fold and the stack evaluation coincide.  Float work types only: the JIT's integer DIV is
 * idiv, which traps on a zero divisor; returns -1 (nothing written) for an integer work
 * type. */
int nca_cpu_expr_host_jit(const uint32_t* ins, int n_ins, const void* const* arrays, const int* arr_dtypes,
                          int n_arrays, const double* scalars, int eval_dtype, void* out, int out_dtype, int64_t n) {
  if (eval_dtype != D_F32 && eval_dtype != D_F64) return -1;
  expr_eval(ins, n_ins, arrays, arr_dtypes, n_arrays, scalars, eval_dtype, out, out_dtype, n, 1);
  return 0;
}

/* ---- full reductions (host/reductions.hh) -------------------------------------------- */
/* sum_recursive (host/reductions.hh:58-73): a sequential C-order left fold in sum_type.
 * For float32 that is float arithmetic. */
float nca_cpu_sum_f32(const float* x, int64_t n) {
  float acc = 0.0f;
  for (int64_t i = 0; i < n; ++i) acc += x[i];
  return acc;
}
double nca_cpu_sum_f64(const double* x, int64_t n) {
  double acc = 0.0;
  for (int64_t i = 0; i < n; ++i) acc += x[i];
  return acc;
}
int64_t nca_cpu_sum_i64(const void* x, int dt, int64_t n) {
  /* integer sums wrap modulo 2^64 (then the caller narrows to sum_type) */
  uint64_t acc = 0;
  for (int64_t i = 0; i < n; ++i) acc += (uint64_t)load_i(x, dt, i);
  return (int64_t)acc;
}

/* Value max/min (the rule this backend documents for every max/min it computes): NaN is
 * never selected (the host fold's strict comparison never takes it, host/reductions.hh:
 * 130-146) and -0 orders below +0 (IEEE 754-2019 maximumNumber/minimumNumber), so the
 * result does not depend on the order of combination.  The host fold returns the same
 * value except for the sign of a zero extremum when both zeros occur (it keeps the first).
 * The reference's device traits (reductions.hh:160-166, 262-268) keep the later operand on
 * ties and combine through atomics in no fixed order. */
static double max_num(double a, double b) {
  if (isnan(b)) return a;
  if (isnan(a)) return b;
  if (a == b) return signbit(a) ? b : a;
  return a > b ? a : b;
}
static double min_num(double a, double b) {
  if (isnan(b)) return a;
  if (isnan(a)) return b;
  if (a == b) return signbit(a) ? a : b;
  return a < b ? a : b;
}

/* max_recursive / argmax_recursive (host/reductions.hh:130-169): the running value starts
 * at lowest(); the index is the C-order position of the first element strictly greater
 * than the running value (zeros compare equal, NaN never taken), 0 if none was ever taken.
 * The value follows max_num from lowest(). */
/* 64-bit integers compare exactly (unsigned for uint64); the value returns as a double */
static double int_extreme(const void* x, int dt, int64_t n, int want_max, int64_t* argidx) {
  int64_t idx = 0, have = 0;
  int64_t best = 0;
  uint64_t ubest = 0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t v = load_i(x, dt, i);
    int take;
    if (dt == D_U64) take = !have || (want_max ? (uint64_t)v > ubest : (uint64_t)v < ubest);
    else take = !have || (want_max ? v > best : v < best);
    if (take) { best = v; ubest = (uint64_t)v; idx = i; have = 1; }
  }
  if (argidx) *argidx = idx;
  return dt == D_U64 ? (double)ubest : (double)best;
}

double nca_cpu_max(const void* x, int dt, int64_t n, double lowest, int64_t* argidx) {
  if ((dt == D_I64 || dt == D_U64) && n > 0) return int_extreme(x, dt, n, 1, argidx);
  double acc = lowest, best = lowest;
  int64_t idx = 0;
  for (int64_t i = 0; i < n; ++i) {
    double v = load_d(x, dt, i);
    acc = max_num(acc, v);
    if (v > best) { best = v; idx = i; }
  }
  if (argidx) *argidx = idx;
  return acc;
}
/* min_recursive / argmin_recursive (host/reductions.hh:171-212) */
double nca_cpu_min(const void* x, int dt, int64_t n, double highest, int64_t* argidx) {
  if ((dt == D_I64 || dt == D_U64) && n > 0) return int_extreme(x, dt, n, 0, argidx);
  double acc = highest, best = highest;
  int64_t idx = 0;
  for (int64_t i = 0; i < n; ++i) {
    double v = load_d(x, dt, i);
    acc = min_num(acc, v);
    if (v < best) { best = v; idx = i; }
  }
  if (argidx) *argidx = idx;
  return acc;
}

/* ---- axis reductions (HostEngine::execute_reduce_axes, engines/hostengine.hh:152-191) --
 * For every input index i in C order: j = output index from build_reduction_params
 * (custom_types.hh:132-176: reduced axes have output stride 0), then
 * res[j] = reduce(res[j], x[i]) with res filled by the traits' fill() value.
 * op: 0 sum (float32 accumulate), 1 max, 2 min (max_num / min_num from lowest() / max(), the
 * fill values of reductions.hh:144-283), 3 argmax, 4 argmin (the host fold's strict
 * comparison: first occurrence, NaN never taken; returns the C-order linear input index). */
void nca_cpu_reduce_axes_f32(const float* x, int ndim, const int64_t* shape, const int* reduced, int op,
                             void* out) {
  int64_t ostride[16];
  int64_t acc = 1, n = 1, nout = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    ostride[d] = reduced[d] ? 0 : acc;
    if (!reduced[d]) acc *= shape[d];
    n *= shape[d];
  }
  nout = acc;
  float* of = (float*)out;
  int64_t* oi = (int64_t*)out;
  float* best = NULL;
  if (op == 0) for (int64_t j = 0; j < nout; ++j) of[j] = 0.0f;
  if (op == 1) for (int64_t j = 0; j < nout; ++j) of[j] = -3.402823466e+38f;
  if (op == 2) for (int64_t j = 0; j < nout; ++j) of[j] = 3.402823466e+38f;
  if (op >= 3) {
    best = (float*)malloc(sizeof(float) * (size_t)(nout > 0 ? nout : 1));
    for (int64_t j = 0; j < nout; ++j) { best[j] = op == 3 ? -3.402823466e+38f : 3.402823466e+38f; oi[j] = 0; }
  }
  for (int64_t i = 0; i < n; ++i) {
    int64_t t = i, j = 0;
    for (int d = ndim - 1; d >= 0; --d) {
      int64_t c = t % shape[d];
      t /= shape[d];
      j += c * ostride[d];
    }
    float v = x[i];
    if (op == 0) of[j] = of[j] + v;
    else if (op == 1) of[j] = (float)max_num(of[j], v);
    else if (op == 2) of[j] = (float)min_num(of[j], v);
    else if (op == 3) { if (v > best[j]) { best[j] = v; oi[j] = i; } }
    else { if (v < best[j]) { best[j] = v; oi[j] = i; } }
  }
  free(best);
}

/* ---- stencil (rtcompiler.hh:339-376: dest[c] = f(src[c*step + off_k])) ----------------
 * C-order src/dst of rank ndim; step = src.shape / dst.shape (integer division).
 * zero_boundary: a read outside src gives 0 (zero-Dirichlet); otherwise the caller
 * guarantees in-range reads.  The tape is evaluated in float32 (work type) with plain
 * division, loads k < n_off from src at offsets[k]. */
void nca_cpu_stencil_f32(const float* src, const int64_t* sshape, float* dst, const int64_t* dshape, int ndim,
                         const int64_t* offsets, int n_off, const uint32_t* ins, int n_ins, const double* scalars,
                         int zero_boundary) {
  int64_t n = 1, sstr[16], step[16];
  int64_t acc = 1;
  for (int d = ndim - 1; d >= 0; --d) { sstr[d] = acc; acc *= sshape[d]; }
  for (int d = 0; d < ndim; ++d) { n *= dshape[d]; step[d] = sshape[d] / dshape[d]; }
  double loads[64];
  for (int64_t i = 0; i < n; ++i) {
    int64_t c[16], t = i;
    for (int d = ndim - 1; d >= 0; --d) { c[d] = t % dshape[d]; t /= dshape[d]; }
    for (int k = 0; k < n_off; ++k) {
      int64_t off = 0;
      int inb = 1;
      for (int d = 0; d < ndim; ++d) {
        int64_t s = c[d] * step[d] + offsets[k * ndim + d];
        if (s < 0 || s >= sshape[d]) inb = 0;
        off += s * sstr[d];
      }
      loads[k] = (inb || !zero_boundary) ? (inb ? src[off] : 0.0f) : 0.0f;
    }
    dst[i] = (float)eval_one(ins, n_ins, loads, scalars, (double)i, D_F32, 1);
  }
}

/* ---- binning (scatter_add: kernels.cuh:660-680, host/elementwise.hh:482-505) ----------
 * Per frame f: counts[f][b] = number of pixels p with map[p] == b (only value != 0 when
 * nonzero is set); sums[f][b] = sum of frame values in float64 (the oracle).  Indices
 * outside [0, nb) are skipped. */
void nca_cpu_bin_f32(const float* frames, int64_t F, int64_t P, const int32_t* map, int nb, int nonzero,
                     int64_t* counts, double* sums) {
  for (int64_t f = 0; f < F; ++f) {
    int64_t* cf = counts + f * nb;
    double* sf = sums + f * nb;
    for (int b = 0; b < nb; ++b) { cf[b] = 0; sf[b] = 0.0; }
    const float* fr = frames + f * P;
    for (int64_t p = 0; p < P; ++p) {
      int32_t b = map[p];
      if (b < 0 || b >= nb) continue;
      float v = fr[p];
      sf[b] += (double)v;
      if (!nonzero || v != 0.0f) cf[b] += 1;
    }
  }
}

/* reference scatter_add on a float32 destination: dest[idx[i]] += src[i] in float, in
 * index order (host/elementwise.hh:482-505 for a 1-D destination) */
void nca_cpu_scatter_add_f32(float* dest, int64_t nd, const int32_t* idx, const float* src, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (idx[i] >= 0 && idx[i] < nd) dest[idx[i]] += src[i];
}

/* gather backward of the weighted sums: g_frames[f][p] = g_sums[f][map[p]] (0 outside) */
void nca_cpu_bin_backward_f32(const float* gsums, int64_t F, int nb, const int32_t* map, int64_t P, float* gout) {
  for (int64_t f = 0; f < F; ++f)
    for (int64_t p = 0; p < P; ++p) {
      int32_t b = map[p];
      gout[f * P + p] = (b >= 0 && b < nb) ? gsums[f * nb + b] : 0.0f;
    }
}

/* ---- CPU baseline: the 3-D Laplacian pipeline ------------------------------------------
 * lap = 7-point Laplacian with zero Dirichlet (x[z-1]+x[z+1]+x[y-1]+x[y+1]+x[x-1]+x[x+1]
 * - 6*x, summed in that order in float32), then colsum = sum over axis 0 (float32 fold
 * in z order, as HostEngine::execute_reduce_axes does) and mx = max(lap) (max_num).
 * threads > 1 splits the z planes (OpenMP, as the reference VM path,
 * engines/hostengine.hh:97-118); the fold order per output is unchanged.  Returns mx. */
float nca_cpu_laplacian_pipeline(const float* x, int64_t n0, int64_t n1, int64_t n2, float* lap, float* colsum,
                                 int threads) {
  const int64_t plane = n1 * n2;
  float* pmx = (float*)malloc(sizeof(float) * (size_t)(n0 > 0 ? n0 : 1));
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t z = 0; z < n0; ++z) {
    float lmx = -3.402823466e+38f;
    for (int64_t y = 0; y < n1; ++y) {
      const float* row = x + z * plane + y * n2;
      float* out = lap + z * plane + y * n2;
      for (int64_t xx = 0; xx < n2; ++xx) {
        float zm = z > 0 ? row[xx - plane] : 0.0f;
        float zp = z + 1 < n0 ? row[xx + plane] : 0.0f;
        float ym = y > 0 ? row[xx - n2] : 0.0f;
        float yp = y + 1 < n1 ? row[xx + n2] : 0.0f;
        float xm = xx > 0 ? row[xx - 1] : 0.0f;
        float xp = xx + 1 < n2 ? row[xx + 1] : 0.0f;
        float v = zm + zp;
        v = v + ym;
        v = v + yp;
        v = v + xm;
        v = v + xp;
        v = v - 6.0f * row[xx];
        out[xx] = v;
        lmx = (float)max_num(lmx, v);
      }
    }
    pmx[z] = lmx;
  }
  float mx = -3.402823466e+38f;
  for (int64_t z = 0; z < n0; ++z) mx = (float)max_num(mx, pmx[z]);
  free(pmx);
  for (int64_t k = 0; k < plane; ++k) colsum[k] = 0.0f;
  for (int64_t z = 0; z < n0; ++z) {
    const float* l = lap + z * plane;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
    for (int64_t k = 0; k < plane; ++k) colsum[k] += l[k];
  }
  return mx;
}

int nca_cpu_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---- CPU baselines of the other workloads (timed by bench.py) ------------------------------
 * The same restatements, with threads > 1 splitting the work over OpenMP threads as the
 * reference VM path does (engines/hostengine.hh:97-118); per-output arithmetic is unchanged. */

/* Tape evaluation (nca_cpu_expr's VM) into a float32 out, then the float32 fold of out
 * (host/reductions.hh:58-73); threads > 1: contiguous chunks per thread, each folded in order,
 * the chunk sums added in chunk order.  Returns the sum. */
float nca_cpu_expr_sum_mt(const uint32_t* ins, int n_ins, const void* const* arrays, const int* arr_dtypes,
                          int n_arrays, const double* scalars, int eval_dtype, float* out, int64_t n, int threads) {
  if (threads < 1) threads = 1;
  float part[256];
  if (threads > 256) threads = 256;
  const int64_t chunk = (n + threads - 1) / threads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int t = 0; t < threads; ++t) {
    double loads[64];
    float acc = 0.0f;
    const int64_t i0 = t * chunk, i1 = (t + 1) * chunk < n ? (t + 1) * chunk : n;
    for (int64_t i = i0; i < i1; ++i) {
      for (int k = 0; k < n_arrays; ++k) loads[k] = load_d(arrays[k], arr_dtypes[k], i);
      const float v = (float)eval_one(ins, n_ins, loads, scalars, (double)i, eval_dtype, 0);
      out[i] = v;
      acc += v;
    }
    part[t] = acc;
  }
  float s = 0.0f;
  for (int t = 0; t < threads; ++t) s += part[t];
  return s;
}

/* nca_cpu_bin_f32 with the frames split over threads. */
void nca_cpu_bin_f32_mt(const float* frames, int64_t F, int64_t P, const int32_t* map, int nb, int nonzero,
                        int64_t* counts, double* sums, int threads) {
  if (threads < 1) threads = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t f = 0; f < F; ++f) nca_cpu_bin_f32(frames + f * P, 1, P, map, nb, nonzero, counts + f * nb, sums + f * nb);
}

/* 7-point Laplacian with zero Dirichlet in the reference op order (as the pipeline above). */
static void lap3d(const float* x, float* out, int64_t n0, int64_t n1, int64_t n2, int threads) {
  const int64_t plane = n1 * n2;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t z = 0; z < n0; ++z)
    for (int64_t y = 0; y < n1; ++y) {
      const float* row = x + z * plane + y * n2;
      float* o = out + z * plane + y * n2;
      for (int64_t xx = 0; xx < n2; ++xx) {
        float v = (z > 0 ? row[xx - plane] : 0.0f) + (z + 1 < n0 ? row[xx + plane] : 0.0f);
        v = v + (y > 0 ? row[xx - n2] : 0.0f);
        v = v + (y + 1 < n1 ? row[xx + n2] : 0.0f);
        v = v + (xx > 0 ? row[xx - 1] : 0.0f);
        v = v + (xx + 1 < n2 ? row[xx + 1] : 0.0f);
        o[xx] = v - 6.0f * row[xx];
      }
    }
}

/* The differentiable C3 pipeline: a = Laplacian(x), loss = sum(a*y + exp(-a*a) - 0.5*y) (float32
 * per point in tape order, float64 sum); with with_grad, dy = a - 0.5 and dx = Laplacian(w) with
 * w = y - 2*a*exp(-a*a) (the Laplacian with zero boundary is its own transpose).  Returns the
 * loss. */
double nca_cpu_laplacian_c0(const float* x, const float* y, int64_t n0, int64_t n1, int64_t n2, float* a, float* dx,
                            float* dy, int with_grad, int threads) {
  if (threads < 1) threads = 1;
  const int64_t n = n0 * n1 * n2;
  lap3d(x, a, n0, n1, n2, threads);
  double loss = 0.0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : loss)
#endif
  for (int64_t i = 0; i < n; ++i) {
    const float av = a[i], yv = y[i];
    const float e = expf(-av * av);
    const float v = av * yv + e - 0.5f * yv;
    loss += (double)v;
    if (with_grad) {
      dy[i] = av - 0.5f;
      a[i] = yv - 2.0f * av * e;   /* w, in place */
    }
  }
  if (with_grad) lap3d(a, dx, n0, n1, n2, threads);
  return loss;
}
