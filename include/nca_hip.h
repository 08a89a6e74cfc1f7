/*
 * nca_hip.h -- C ABI of the MI355X (gfx950) device backend for ncarray's hot path:
 * fused elementwise expressions, stencils, axis and full reductions, and binning.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and returns an
 * `int` status (NCA_OK == 0).  On failure, nca_hip_last_error() returns a thread-local
 * message.  `stream` is a hipStream_t passed as `void*`; NULL means "the legacy default
 * stream, synchronise before returning" (the reference synchronises after every engine
 * call, gpuengine.hh:162,194,287,328).  A non-NULL stream is never synchronised, so calls
 * can be captured into a hipGraph (same contract as Stencil::apply, stencil.hh:236-242).
 *
 * Reference interfaces replaced (paths relative to the ncarray source tree):
 *   GPUEngine (src/lib/ncarray/engines/gpuengine.hh:64-638), the device
 *   RuntimeCompiler (src/lib/ncarray/jit/device/rtcompiler.hh:94-161) and
 *   Stencil<NDIM> (src/lib/ncarray/expression/stencil.hh:59-255).
 */
#ifndef NCA_HIP_H
#define NCA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NCA_MAX_NDIM 10 /* NCARRAY_MAX_NDIM */

/* Stream argument meaning "the legacy default stream, asynchronously": work is ordered with
 * everything else on the default stream (e.g. a framework's default stream) and the call
 * returns without synchronising.  NULL keeps the reference's synchronous behaviour. */
#define NCA_STREAM_DEFAULT_ASYNC ((void*)(uintptr_t)1)

/* ---- status codes ------------------------------------------------------------- */
enum {
  NCA_OK = 0,
  NCA_ERR_INVALID = 1,     /* bad argument: shape mismatch, null pointer, bad dtype */
  NCA_ERR_UNSUPPORTED = 2, /* valid in the reference type system, not built here */
  NCA_ERR_HIP = 3,         /* a HIP runtime call failed */
  NCA_ERR_COMPILE = 4,     /* hipRTC compilation failed (log in the error message) */
  NCA_ERR_OOM = 5,         /* device allocation failed */
  NCA_ERR_NOT_INIT = 6     /* nca_hip_init was not called */
};

/* ---- dtypes -------------------------------------------------------------------
 * Codes 0..21 are the reference's DType enumerators in order (dtype.hh:56-80).
 * 22/23 are storage-only types added here: arithmetic on them runs in float32. */
enum {
  NCA_BOOL = 0, NCA_CHAR = 1, NCA_UINT8 = 2, NCA_UINT16 = 3, NCA_UINT32 = 4,
  NCA_UINT64 = 5, NCA_INT8 = 6, NCA_INT16 = 7, NCA_INT32 = 8, NCA_INT64 = 9,
  NCA_FLOAT32 = 10, NCA_FLOAT64 = 11, NCA_FLOAT128 = 12, NCA_COMPLEX64 = 13,
  NCA_COMPLEX128 = 14, NCA_COMPLEX256 = 15, NCA_VFLOAT2 = 16, NCA_VFLOAT3 = 17,
  NCA_VFLOAT4 = 18, NCA_VDOUBLE2 = 19, NCA_VDOUBLE3 = 20, NCA_VDOUBLE4 = 21,
  NCA_BFLOAT16 = 22, NCA_FLOAT16 = 23
};

/* ---- opcodes ------------------------------------------------------------------
 * 0..32 are the reference OpCode enumerators in order (op_code.hh:47-87); an
 * instruction packs (op << 24) | (operand index & 0xFFFFFF) (op_code.hh:174-192).
 * 64.. are unary math ops added here (the reference has no exponential). */
enum {
  NCA_OP_NOOP = 0, NCA_OP_IDX = 1, NCA_OP_LOAD_NCARR = 2, NCA_OP_LOAD_SOARR = 3,
  NCA_OP_LOAD_CONST = 4, NCA_OP_NEG = 5, NCA_OP_INC = 6, NCA_OP_DEC = 7,
  NCA_OP_SZOF = 8, NCA_OP_ADDR = 9, NCA_OP_INDR = 10, NCA_OP_CAST = 11,
  NCA_OP_LNOT = 12, NCA_OP_BNOT = 13, NCA_OP_ADD = 14, NCA_OP_SUB = 15,
  NCA_OP_MUL = 16, NCA_OP_DIV = 17, NCA_OP_MOD = 18, NCA_OP_FDIV = 19,
  NCA_OP_EQ = 20, NCA_OP_NE = 21, NCA_OP_LT = 22, NCA_OP_LE = 23, NCA_OP_GT = 24,
  NCA_OP_GE = 25, NCA_OP_LAND = 26, NCA_OP_LOR = 27, NCA_OP_BAND = 28,
  NCA_OP_BOR = 29, NCA_OP_XOR = 30, NCA_OP_LSHFT = 31, NCA_OP_RSHFT = 32,
  NCA_OP_EXP = 64, NCA_OP_LOG = 65, NCA_OP_SQRT = 66, NCA_OP_ABS = 67
};

/* ---- reductions --------------------------------------------------------------- */
enum {
  NCA_REDUCE_SUM = 0, NCA_REDUCE_MAX = 1, NCA_REDUCE_MIN = 2,
  NCA_REDUCE_ARGMAX = 3, NCA_REDUCE_ARGMIN = 4, NCA_REDUCE_MEAN = 5,
  NCA_REDUCE_ALL = 6, NCA_REDUCE_ANY = 7
};
/* Accumulation of floating-point sums.  Both are deterministic (no float atomics):
 * PLAIN keeps per-thread float32 running sums and a pairwise in-block tree;
 * COMPENSATED keeps Neumaier (sum, compensation) pairs per thread and merges them
 * with error-free two-sums.  Either way the cross-block combine is float64. */
enum { NCA_ACCUM_PLAIN = 0, NCA_ACCUM_COMPENSATED = 1 };
/* bit flags for fused reductions that ride on an expression or stencil launch */
enum { NCA_FUSE_SUM = 1, NCA_FUSE_MAX = 2, NCA_FUSE_MIN = 4 };

/* ---- stencil boundary handling ------------------------------------------------
 * NONE: the reference's contract (rtcompiler.hh:339-376 does no bounds checks):
 *       every access must be in range; the host checks this and returns
 *       NCA_ERR_INVALID instead of reading out of bounds.
 * ZERO: zero-Dirichlet: a source read outside the array reads 0. */
enum { NCA_BOUNDARY_NONE = 0, NCA_BOUNDARY_ZERO = 1 };

/* ---- view layouts --------------------------------------------------------------
 * NC: strided, with an optional NC pointer axis (NCOffsetsPolicy).
 * SO: PEP 3118 strides and suboffsets (SOArrayPolicy). */
enum { NCA_LAYOUT_NC = 0, NCA_LAYOUT_SO = 1 };

/* ---- bin-count predicates ------------------------------------------------------ */
enum { NCA_COUNT_ALL = 0, NCA_COUNT_NONZERO = 1 };

/* A strided device array view.  `data` addresses element [0,...,0]; per-axis byte
 * offsets of the reference layout (layout.hh:468-476) are folded into it.  Strides are
 * in bytes (layout.hh m_strides).  ndim == 0 is a scalar.
 * pointer_axis >= 0 marks an NC pointer axis P (NCOffsetsPolicy::advance, layout.hh:
 * 468-480): element (i_0, ..., i_{n-1}) is at
 *     row + row_offset + sum_{d > P} i_d * strides[d],
 *     row = ((char**)(data + sum_{d < P} i_d * strides[d]))[i_P * strides[P] + ptr_offset],
 * i.e. data is a device table of row pointers, strides[P] counts table entries (1 for the
 * reference's layouts, whose advance ignores it; 0 broadcasts one row; k > 1 is a stepped
 * slice), ptr_offset is the reference's m_offsets[P] and row_offset the sum of its byte
 * offsets of the axes after P.
 * layout == NCA_LAYOUT_SO (pointer_axis must be -1) reads the PEP 3118 form of
 * SOArrayPolicy::advance (layout.hh:677-700): starting from data, for d = 0 .. ndim-1
 *     p += i_d * strides[d];  if (suboffsets[d] >= 0) p = *(char**)p + suboffsets[d];
 * so any number of axes may be indirect.  suboffsets is read only for this layout; an SO
 * view with no suboffset >= 0 is a plain strided view and is accepted everywhere.
 * Views with an NC pointer axis or an SO suboffset >= 0 are read (and written) in place by the
 * expression entry points, stencil apply, full and axis reductions (input), scatter_add and
 * binning (frames through a table of frame pointers); the other entry points (the training
 * kernels, ragged batches) take strided views. */
typedef struct nca_view {
  void* data;
  int32_t dtype;
  int32_t ndim;
  int64_t shape[NCA_MAX_NDIM];
  int64_t strides[NCA_MAX_NDIM];
  int32_t pointer_axis;
  int32_t layout; /* NCA_LAYOUT_NC (0) or NCA_LAYOUT_SO */
  int64_t ptr_offset;
  int64_t row_offset;
  int64_t suboffsets[NCA_MAX_NDIM]; /* NCA_LAYOUT_SO only; < 0: a direct axis */
} nca_view;

/* A scalar constant.  The expression's dtype tables say which member is meant. */
typedef union nca_scalar {
  int64_t i;
  uint64_t u;
  double f;
} nca_scalar;

/* An expression tape: ExprMVNode's instrs/dtypes/scalars (mvnode.hh:82-98).
 * LOAD_NCARR k reads array operand k; LOAD_CONST k reads scalar k; IDX is the C-order
 * linear index of the output element.  work_dtype is ExprMVNode::work_dtype; dest_dtype
 * is the dtype of the array the expression is assigned to. */
typedef struct nca_expr_desc {
  const uint32_t* instrs;
  int32_t n_instrs;
  int32_t n_arrays;
  const int32_t* arr_dtypes;
  int32_t n_scalars;
  int32_t work_dtype;
  const int32_t* scalar_dtypes;
  int32_t dest_dtype;
  int32_t reserved;
} nca_expr_desc;

/* A stencil: `dest[c] = f(src[c*step + offsets[k]] for k, extra_j[c])` with
 * step = src.shape / dest.shape (integer division, rtcompiler.hh:363-368).
 * In expr, LOAD_NCARR k < n_offsets reads src at offsets[k]; LOAD_NCARR
 * n_offsets + j reads extra pointwise array j at the output coordinate (dest shape).
 * extra_params / prologue / epilogue are StencilJITExtensions strings
 * (jit/device/extensions.hh:68-72); any of them may be NULL.
 * is_pointer_axis (ndim flags) and is_soarr are Stencil<NDIM>::create's (stencil.hh:76-81):
 * when is_pointer_axis is given, apply accepts only sources whose indirect axes are exactly
 * the flagged ones (an NC pointer axis, or with is_soarr the SO axes with suboffset >= 0);
 * NULL accepts any layout.  Sources and destinations may be NC pointer-axis or SO views
 * either way (the row pointers are read in the kernel). */
typedef struct nca_stencil_desc {
  int32_t ndim;
  int32_t n_offsets;
  const int64_t* offsets; /* n_offsets * ndim, row-major */
  int32_t src_dtype;
  int32_t n_extra;
  const int32_t* extra_dtypes;
  const nca_expr_desc* expr;
  const char* extra_params;
  const char* prologue;
  const char* epilogue;
  const uint8_t* is_pointer_axis; /* ndim flags, or NULL */
  int32_t is_soarr;
} nca_stencil_desc;

/* ================================================================================ */
/* runtime                                                                          */
/* ================================================================================ */

/* Select `device` for the calling thread and create its state (replaces the implicit
 * CUDA context of jit_utils.cc:28-46).  Idempotent. */
int nca_hip_init(int device);
/* Release cached kernels, workspaces and the JIT memory cache. */
int nca_hip_shutdown(void);
/* Copy the calling thread's last error message into buf (always NUL-terminated).
 * Returns the message length. */
int nca_hip_last_error(char* buf, size_t n);
/* Library version string, e.g. "ncarray-hip 0.1.0 gfx950". */
const char* nca_hip_version(void);
/* Device properties the kernels size themselves by. */
int nca_hip_device_info(int32_t* n_cus, int32_t* lds_per_block, int64_t* total_mem,
                        char* arch_name, size_t arch_len);
/* The device this thread's calls run on (nca_hip_init's, or the one a worker thread adopted):
 * the device id of DLPack exports and similar descriptions of the library's memory. */
int nca_hip_current_device(int32_t* device);

/* Device memory for DevOwnerPolicy (storage.hh:176-186, 242-252).  Like the reference's
 * cudaMallocAsync on one stream, blocks come from the device's default memory pool,
 * stream-ordered on the legacy default stream: work on that stream (and on streams that
 * synchronise with it) may use the memory at once; a caller that uses it on a non-blocking
 * stream orders that stream after the allocation, and finishes that stream's work before
 * nca_hip_free, which returns the block to the pool in the default stream's order (no device
 * synchronisation).  Without a pool these are hipMalloc / hipFree. */
int nca_hip_malloc(void** ptr, size_t bytes);
int nca_hip_free(void* ptr);
int nca_hip_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int nca_hip_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int nca_hip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int nca_hip_synchronize(void* stream);

/* Kernel timing: when enabled, every kernel the library launches is bracketed by a pair
 * of HIP events on its stream.  nca_hip_profile_read synchronises on the recorded events,
 * writes one "kernel_name milliseconds" line per launch since the last read into buf,
 * clears the record, and returns the text length (negative status on error). */
int nca_hip_profile_enable(int32_t on);
int nca_hip_profile_read(char* buf, size_t n);

/* Counters of the kernel cache since the process started: sources compiled with hipRTC,
 * code objects loaded from the on-disk cache, and lookups served from memory.  The disk cache
 * (NCA_HIP_CACHE_DIR, else $XDG_CACHE_HOME/ncarray_amd, else ~/.cache/ncarray_amd) stores
 * each code object under the SHA-256 of its source and compiler options with that key, the
 * code's length and its SHA-256 in a header; a truncated, damaged or foreign file is compiled
 * again and rewritten (the reference's cubin cache, rtcompiler.cc:74-86). */
int nca_hip_jit_stats(int64_t* compiles, int64_t* disk_loads, int64_t* memory_hits);

/* Compile a complete generated source (as returned by nca_hip_expr_source or
 * nca_hip_stencil_source) with hipRTC for `arch` (NULL: gfx950) without loading it.
 * Needs no GPU.  On failure the compiler log is copied into log. */
int nca_hip_rtc_compile(const char* source, const char* arch, char* log, size_t n);

/* Result dtype of a reduction over `src_dtype` (SumTraits/MaxTraits/ArgmaxTraits...
 * OutT, reductions.hh:109-341; sum_type from op_traits.hh:72-410). */
int nca_hip_reduce_result_dtype(int32_t op, int32_t src_dtype, int32_t* out_dtype);
/* Result dtype of applying `op` to `work_dtype` (determine_dtype_for_op,
 * op_code.hh:101-161, extended to EXP/LOG/SQRT/ABS and bf16/fp16). */
int nca_hip_op_result_dtype(int32_t op, int32_t work_dtype, int32_t* out_dtype);

/* ================================================================================ */
/* fused elementwise expressions (hipRTC)                                           */
/* ================================================================================ */

/* Validate a tape and return a handle to its kernel family (replaces
 * RuntimeCompiler::get_expr_kernel / get_dynamic_vm_kernel, rtcompiler.cc:49-89,
 * 341-366).  Kernels are generated and compiled lazily per layout class, cached in
 * memory and on disk (NCA_HIP_CACHE_DIR).  The output shape is a launch argument. */
int nca_hip_expr_compile(const nca_expr_desc* desc, uint64_t* handle);
/* out[c] = expr(ins..., scalars...) for every c in out's shape (the operand views
 * must have out's shape, or broadcast to it with zero strides)
 * (GPUEngine::execute_binary_expression, gpuengine.hh:90-195). */
int nca_hip_expr_launch(uint64_t handle, const nca_view* ins, int32_t n_in,
                        const nca_scalar* scalars, int32_t n_scalars, const nca_view* out,
                        void* stream);
/* As nca_hip_expr_launch, plus reductions of the values stored into out (in out's
 * dtype) in the same pass.  out->data may be NULL: the expression is then not
 * materialised and out only supplies shape and dtype.  fuse is NCA_FUSE_* flags;
 * sum_out (dtype = sum result dtype of out's dtype), max_out and min_out (out's dtype)
 * are device pointers to one element each, may be NULL when not requested. */
int nca_hip_expr_launch_reduce(uint64_t handle, const nca_view* ins, int32_t n_in,
                               const nca_scalar* scalars, int32_t n_scalars,
                               const nca_view* out, int32_t fuse, int32_t accum_mode,
                               void* sum_out, void* max_out, void* min_out, void* stream);
/* Reverse mode: for each array operand k with wrt[k] != 0, write
 * grads[k][c] = g(c) * d expr / d operand_k (c).  g is either the view grad_out
 * (out's shape) or, when grad_out is NULL, the single device value *grad_scalar
 * (broadcast: the backward of sum()).  `grads` has n_in entries; entries with
 * wrt[k] == 0 are ignored.  The derivative tape is generated symbolically.
 * Comparisons, logical and bit ops have derivative 0; DIV differentiates a/b. */
int nca_hip_expr_vjp(uint64_t handle, const nca_view* ins, int32_t n_in,
                     const nca_scalar* scalars, int32_t n_scalars, const nca_view* grad_out,
                     const void* grad_scalar, const int32_t* wrt, const nca_view* grads,
                     void* stream);
/* The generated HIP source of a compiled variant (for inspection and tests);
 * variant 0 = contiguous, 1 = strided, 2 = fused-reduce, 3 = vjp, 4 = ragged batch,
 * 5 = tiled transpose (strided operands staged through LDS along axis 0),
 * 6 = rows (one wave per row of the last axis).
 * Does not need a GPU.  Returns the full length; copies at most n-1 bytes. */
int nca_hip_expr_source(uint64_t handle, int32_t variant, int32_t ndim, char* buf, size_t n);

/* ================================================================================ */
/* reductions                                                                       */
/* ================================================================================ */

/* Full reduction to one element written to the device pointer out_dev
 * (GPUEngine::execute_full_reduce, gpuengine.hh:301-333).  argmax/argmin return the
 * int64 C-order linear index of the first occurrence (host/reductions.hh:148-169: strict
 * comparisons against the running value, zeros equal, NaN never taken).  max/min return
 * the IEEE 754-2019 maximumNumber/minimumNumber of the values and lowest()/max() of the
 * dtype: NaN never selected and -0 below +0, so every order of combination gives the same
 * bits (the host fold differs only in the sign of a zero extremum when both zeros occur). */
int nca_hip_reduce_full(const nca_view* in, int32_t op, int32_t accum_mode, void* out_dev,
                        void* stream);
/* Reduction along `axes` into out (GPUEngine::execute_reduce_axes,
 * gpuengine.hh:222-288; build_reduction_params, custom_types.hh:132-176).  Axes
 * outside [0, ndim) are ignored, as in the reference.  out has the kept axes in order.
 * argmax/argmin store the C-order linear index into `in` of the first occurrence. */
int nca_hip_reduce_axes(const nca_view* in, const int64_t* axes, int32_t n_axes,
                        int32_t op, int32_t accum_mode, const nca_view* out, void* stream);

/* ================================================================================ */
/* stencils (hipRTC)                                                                */
/* ================================================================================ */

/* Stencil<NDIM>::create (stencil.hh:76-156). */
int nca_hip_stencil_create(const nca_stencil_desc* desc, uint64_t* handle);
/* Stencil<NDIM>::apply (stencil.hh:181-243).  ext_args are pointers to the values of
 * the extra kernel parameters declared in extra_params, in order. */
int nca_hip_stencil_apply(uint64_t handle, const nca_view* src, const nca_view* extras,
                          int32_t n_extra, const nca_scalar* scalars, int32_t n_scalars,
                          const nca_view* dst, int32_t boundary, void* const* ext_args,
                          int32_t n_ext_args, void* stream);
/* Stencil apply with reductions of the stored values fused into the same pass:
 * axis0_sum (dst shape without axis 0; sum result dtype) and/or the scalars
 * sum_out / max_out / min_out (device pointers).  dst->data may be NULL (not
 * materialised).  Requires step 1 along every axis. */
int nca_hip_stencil_apply_reduce(uint64_t handle, const nca_view* src,
                                 const nca_view* extras, int32_t n_extra,
                                 const nca_scalar* scalars, int32_t n_scalars,
                                 const nca_view* dst, int32_t boundary,
                                 const nca_view* axis0_sum, int32_t fuse,
                                 int32_t accum_mode, void* sum_out, void* max_out,
                                 void* min_out, void* stream);
/* nca_hip_stencil_apply_reduce that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under hipGraph capture, tight
 * serving loops, harnesses that forbid allocation).  The block partials and z-run column
 * partials live in a fixed scratch array of 32 MiB per device (4 MiB of block partials, 28 MiB
 * of column partials), held from nca_hip_init until the process ends; the caller still passes
 * the inputs and outputs.  A launch that needs more block partials than that fails with
 * NCA_ERR_INVALID; the fused axis-0 sum uses fewer z runs when its column partials would not
 * fit.  Kernels are not compiled or loaded here (loading a code object allocates): the first
 * call with a given stencil, layout and fusion must be an ordinary nca_hip_stencil_apply_reduce,
 * else NCA_ERR_INVALID.  Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (other streams, other threads) race and must be serialised by the caller. */
int nca_hip_stencil_apply_reduce_noalloc(uint64_t handle, const nca_view* src,
                                         const nca_view* extras, int32_t n_extra,
                                         const nca_scalar* scalars, int32_t n_scalars,
                                         const nca_view* dst, int32_t boundary,
                                         const nca_view* axis0_sum, int32_t fuse,
                                         int32_t accum_mode, void* sum_out, void* max_out,
                                         void* min_out, void* stream);
/* The adjoint (transpose) of a stencil that is linear in its source loads:
 * grad_src[p] = sum_k c_k * grad_dst[p - offsets[k]], with c_k = d f / d v_k derived
 * symbolically (they may depend on the scalars, not on the loads).  Unit step only;
 * boundary ZERO treats grad_dst outside its shape as 0, NONE requires in-range. */
int nca_hip_stencil_adjoint_create(uint64_t stencil_handle, uint64_t* handle);
/* Backward of loss = sum(expr(stencil(src), extras...)) in one march pass over src (no
 * counterpart in the reference, whose expressions and stencils have no derivatives; it
 * backs the fused stencil-expression loss of the autograd layer).  At every point the
 * pass computes a = stencil(src) and writes w = d expr / d a * g and, for every
 * grad_extras[j] whose data is not NULL, d expr / d extra_j * g; g is read from the
 * device pointer grad_scalar (the expression's compute type; NULL means 1).
 * expr_operand[k] names what expression operand k reads: -1 the stencil value, j >= 0
 * extra j.  For a stencil linear in its loads the gradient of src is the stencil's
 * adjoint (nca_hip_stencil_adjoint_create) applied to w.  Requires a stencil without
 * extras or hooks whose z offsets have no x/y part, 3-D C-contiguous arrays of one
 * shape with the last extent a multiple of the vector width, and boundary ZERO. */
int nca_hip_stencil_expr_vjp(uint64_t stencil_handle, uint64_t expr_handle,
                             const int32_t* expr_operand, int32_t n_operand,
                             const nca_view* src, const nca_view* extras, int32_t n_extra,
                             const nca_scalar* stencil_scalars, int32_t n_stencil_scalars,
                             const nca_scalar* expr_scalars, int32_t n_expr_scalars,
                             const void* grad_scalar, const nca_view* w,
                             const nca_view* grad_extras, int32_t boundary, void* stream);
/* nca_hip_stencil_expr_vjp that in the same pass also reduces the expression's value
 * expr(stencil(src), extras...) (the loss itself; no counterpart in the reference): fuse
 * (NCA_FUSE_*, not 0) and accum_mode as in nca_hip_stencil_apply_reduce, results in the
 * device scalars sum_out / max_out / min_out typed as a reduction of the expression's
 * dest dtype.  One pass then yields the loss and the gradients of a training step. */
int nca_hip_stencil_expr_vjp_reduce(uint64_t stencil_handle, uint64_t expr_handle,
                                    const int32_t* expr_operand, int32_t n_operand,
                                    const nca_view* src, const nca_view* extras, int32_t n_extra,
                                    const nca_scalar* stencil_scalars, int32_t n_stencil_scalars,
                                    const nca_scalar* expr_scalars, int32_t n_expr_scalars,
                                    const void* grad_scalar, const nca_view* w,
                                    const nca_view* grad_extras, int32_t fuse, int32_t accum_mode,
                                    void* sum_out, void* max_out, void* min_out, int32_t boundary,
                                    void* stream);
/* A whole training step of loss = sum(expr(stencil(src), extras)) in one pass over src and
 * the extras (no counterpart in the reference): grad_src = stencil^T(d expr / d a * g), the
 * extras' gradients (grad_extras[j] with data) and, with fuse != 0, the reductions of the
 * expression's value (as nca_hip_stencil_expr_vjp_reduce).  The pass stages src with a
 * two-point halo in LDS, evaluates d expr / d a on each tile and a one-point ring around it,
 * and applies the stencil's adjoint from the last three planes of it kept in LDS, so
 * nothing but the inputs and the gradients reaches memory.  Requires a linear 3-D stencil
 * without extras or hooks, of radius <= 1 on every axis, boundary ZERO, and C-contiguous
 * arrays of one shape (any extents, no alignment requirement). */
int nca_hip_stencil_expr_train(uint64_t stencil_handle, uint64_t expr_handle,
                               const int32_t* expr_operand, int32_t n_operand,
                               const nca_view* src, const nca_view* extras, int32_t n_extra,
                               const nca_scalar* stencil_scalars, int32_t n_stencil_scalars,
                               const nca_scalar* expr_scalars, int32_t n_expr_scalars,
                               const void* grad_scalar, const nca_view* grad_src,
                               const nca_view* grad_extras, int32_t fuse, int32_t accum_mode,
                               void* sum_out, void* max_out, void* min_out, int32_t boundary,
                               void* stream);
/* The generated source of that pass for the given operand codes, extra dtypes, grad_src
 * dtype and reductions (every extra's gradient written), for inspection and
 * nca_hip_rtc_compile; vec != 0 gives the 4-point-vector form the pass uses when the last
 * extent is a multiple of 4 and every array is aligned to whole vectors.  Returns the text
 * length (negative status on error). */
int nca_hip_stencil_expr_train_source(uint64_t stencil_handle, uint64_t expr_handle,
                                      const int32_t* expr_operand, int32_t n_operand,
                                      const int32_t* extra_dtypes, int32_t n_extra,
                                      int32_t grad_src_dtype, int32_t fuse, int32_t accum_mode,
                                      int32_t vec, char* buf, size_t n);
/* The generated source of a stencil kernel (for inspection and nca_hip_rtc_compile):
 * variant 0 the march kernel, 1 the march with the fused axis-0 sum and max, 4 the same
 * with compensated sums, 2 / 3 the direct kernel without / with the zero boundary, 5 the
 * tile kernel at unit steps, 6 the tile kernel with fused sum/max/min and the first axis
 * read through a table of row pointers.
 * Returns the text length (negative status on error); writes at most n bytes. */
int nca_hip_stencil_source(uint64_t handle, int32_t variant, char* buf, size_t n);

/* ================================================================================ */
/* binning (scatter_add)                                                            */
/* ================================================================================ */

/* dest[idx[i]] += src[i] for 0 <= idx[i] < dest.size, other indices skipped
 * (scatter_add_kernel, kernels.cuh:660-680; host::scatter_add_recursive,
 * host/elementwise.hh:482-505).  i runs over src and idx (same size) in C order; idx[i] is a
 * C-order linear index into dest.  Any layouts, including NC pointer-axis and SO views. */
int nca_hip_scatter_add(const nca_view* dest, const nca_view* idx, const nca_view* src,
                        void* stream);
/* Per-frame binning by an index map: frames is [F, P...] (P = map size), binmap is an
 * int32 view of P elements.  counts [F, n_bins] (int32 or int64) receives the number
 * of pixels per bin (pred NCA_COUNT_NONZERO: only pixels with value != 0), sums
 * [F, n_bins] (float32 or float64) the sum of pixel values per bin.  Either output may
 * be NULL.  Out-of-range bins are skipped.  Counts are exact; sums are accumulated in
 * float64 inside each block (atomics in hardware order) and merged across blocks in float64 in
 * a fixed order.  frames may be C-contiguous or reached through a table of frame pointers (an
 * NC pointer axis 0, or an SO view whose only indirect axis is 0) with contiguous pixels. */
int nca_hip_bin(const nca_view* frames, const nca_view* binmap, int32_t n_bins,
                int32_t count_pred, const nca_view* counts, const nca_view* sums,
                void* stream);
/* Backward of the weighted sums: grad_frames[f, p] = grad_sums[f, binmap[p]], or 0
 * where binmap[p] is out of range (a gather). */
int nca_hip_bin_backward(const nca_view* grad_sums, const nca_view* binmap,
                         const nca_view* grad_frames, void* stream);

/* ================================================================================ */
/* ragged batches: one persistent launch over many uneven 1-D arrays                 */
/* ================================================================================ */

/* Plan a batch of n_items contiguous 1-D problems sharing one expression: item i has
 * n_in input pointers in_ptrs[i*n_in + k] (host array of device pointers), an output
 * out_ptrs[i] (may be NULL: not materialised) and sizes[i] elements.  fuse selects
 * per-item sum/max/min (NCA_FUSE_*), written to sums/maxs/mins[i] (device arrays).
 * tile_elems / vec / waves_per_cu == 0 pick the autotuned or default values. */
int nca_hip_batch_plan(uint64_t expr_handle, int32_t n_items, int32_t n_in,
                       const void* const* in_ptrs, void* const* out_ptrs,
                       const int64_t* sizes, int32_t fuse, int32_t accum_mode,
                       void* sums, void* maxs, void* mins, int32_t tile_elems,
                       int32_t vec, int32_t waves_per_cu, uint64_t* plan);
int nca_hip_batch_launch(uint64_t plan, const nca_scalar* scalars, int32_t n_scalars,
                         void* stream);
/* Capture the plan's launches into a hipGraph (instantiated) bound to `scalars`. */
int nca_hip_batch_graph(uint64_t plan, const nca_scalar* scalars, int32_t n_scalars,
                        uint64_t* graph);
int nca_hip_graph_launch(uint64_t graph, void* stream);
int nca_hip_graph_destroy(uint64_t graph);
int nca_hip_batch_destroy(uint64_t plan);
/* The tile configuration a plan runs with. */
int nca_hip_batch_config(uint64_t plan, int32_t* tile_elems, int32_t* vec,
                         int32_t* waves_per_cu, int32_t* n_tiles);
/* Sweep tile size x vector width x waves/CU for this plan, time each with HIP events
 * (reps launches each), keep the fastest, and record it in the autotune database
 * (NCA_HIP_AUTOTUNE_DB, keyed by expression signature and size histogram). */
int nca_hip_batch_autotune(uint64_t plan, const nca_scalar* scalars, int32_t n_scalars,
                           int32_t reps, float* best_ms);

#ifdef __cplusplus
}
#endif

#endif /* NCA_HIP_H */
