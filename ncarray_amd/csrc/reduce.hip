// Full and axis reductions for gfx950.  Replaces the reference's reduce_kernel
// (kernels.cuh:565-589, CUB BlockReduce + one host-memory atomic per block) and
// axes_reduce_kernel (kernels.cuh:166-277, one float atomic per input element) with
// deterministic, atomics-free two-level reductions:
//   * full: a grid of at most 4 blocks/CU streams the input with 16-byte loads, each
//     thread folds its elements, 64-lane butterflies + LDS combine a block, every block
//     writes one partial, and a one-block kernel combines the partials in block order
//     (float64 for floating-point sums).  No grid-wide fences: on gfx950 a device-scope
//     fence writes back the issuing XCD's L2, which costs more than a second launch.
//   * axes: the input is viewed as [outer kept][reduced][inner kept] when it can be;
//     "column" kernels give each thread a 16-byte group of inner outputs and march the
//     reduced axis (split into S chunks with a fixed-order combine when the outputs alone
//     do not fill the GPU); "row" kernels give each output row to one wavefront; any
//     other layout uses a generic kernel (one thread per output, unravelled strides).
// Semantics: sums in sum_type (host/reductions.hh:35-212); max/min are IEEE 754-2019
// maximumNumber/minimumNumber from lowest()/max() (NaN never selected as in the host fold,
// -0 below +0, so the bits do not depend on the combination order); argmax/argmin follow the
// host fold exactly: the first occurrence (strict comparison, zeros equal, NaN never taken)
// as a C-order linear index into the input, 0 when nothing was selected.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "nca_device.hh"
#include "nca_internal.hh"

namespace nca {

// ---- per-dtype traits ----------------------------------------------------------------------
// T: storage type, V: value/comparison type, S: sum accumulator, SO: sum output type
template <int DT> struct RT;
#define NCA_RT(DT, T_, V_, S_, SO_, FP_)                                          \
  template <> struct RT<DT> {                                                     \
    typedef T_ T; typedef V_ V; typedef S_ S; typedef SO_ SO;                     \
    static constexpr bool fp = FP_;                                               \
  };
NCA_RT(NCA_BOOL, bool, bool, unsigned long long, unsigned long long, false)
NCA_RT(NCA_CHAR, signed char, signed char, long long, signed char, false)
NCA_RT(NCA_UINT8, unsigned char, unsigned char, unsigned long long, unsigned long long, false)
NCA_RT(NCA_UINT16, unsigned short, unsigned short, unsigned long long, unsigned long long, false)
NCA_RT(NCA_UINT32, unsigned int, unsigned int, unsigned long long, unsigned int, false)
NCA_RT(NCA_UINT64, unsigned long long, unsigned long long, unsigned long long, unsigned long long, false)
NCA_RT(NCA_INT8, signed char, signed char, long long, long long, false)
NCA_RT(NCA_INT16, short, short, long long, long long, false)
NCA_RT(NCA_INT32, int, int, long long, int, false)
NCA_RT(NCA_INT64, long long, long long, long long, long long, false)
NCA_RT(NCA_FLOAT32, float, float, float, float, true)
NCA_RT(NCA_FLOAT64, double, double, double, double, true)
NCA_RT(NCA_BFLOAT16, nca_bf16, float, float, float, true)
NCA_RT(NCA_FLOAT16, nca_f16, float, float, float, true)
#undef NCA_RT

// A 16-byte partial: value (or sum) + index.
struct Slot {
  union { double d; long long i; unsigned long long u; float f; char raw[8]; } a;
  long long idx;
};

// ---- accumulators -------------------------------------------------------------------------------
// One accumulator per (dtype, op); fold(v, i) per element, wave() + block combine, and
// slot conversion.  All combine steps are order-fixed.
template <int DT, int OP, bool COMP> struct Acc;

// SUM / MEAN
template <int DT, bool COMP> struct AccSum {
  typedef typename RT<DT>::V V;
  typedef typename RT<DT>::S S;
  S s, c;
  __device__ void init() { s = S(0); c = S(0); }
  __device__ void fold(V v, nca_i64) {
    if constexpr (RT<DT>::fp) {
      if constexpr (COMP) nca_neumaier_add(s, c, static_cast<S>(v));
      else s += static_cast<S>(v);
    } else {
      s += static_cast<S>(v);
    }
  }
  __device__ void wave() {
    if constexpr (RT<DT>::fp && COMP) nca_wave_comp(s, c);
    else s = nca_wave_sum(s);
  }
  __device__ void merge(const AccSum& o) {
    if constexpr (RT<DT>::fp && COMP) nca_comp_merge(s, c, o.s, o.c);
    else s += o.s;
  }
  __device__ void to_slot(Slot& sl) const {
    if constexpr (RT<DT>::fp) sl.a.d = nca_comp_value(s, c);
    else sl.a.u = static_cast<unsigned long long>(s);
    sl.idx = 0;
  }
};
template <int DT, bool COMP> struct Acc<DT, NCA_REDUCE_SUM, COMP> : AccSum<DT, COMP> {};
template <int DT, bool COMP> struct Acc<DT, NCA_REDUCE_MEAN, COMP> : AccSum<DT, COMP> {};

template <int DT, bool IsMax> struct AccExt {
  typedef typename RT<DT>::V V;
  V m;
  __device__ void init() { m = IsMax ? nca_ident<typename RT<DT>::T, V>::lowest() : nca_ident<typename RT<DT>::T, V>::max(); }
  __device__ void fold(V v, nca_i64) { m = IsMax ? nca_maxop(m, v) : nca_minop(m, v); }
  __device__ void wave() { m = IsMax ? nca_wave_max(m) : nca_wave_min(m); }
  __device__ void merge(const AccExt& o) { m = IsMax ? nca_maxop(m, o.m) : nca_minop(m, o.m); }
  __device__ void to_slot(Slot& sl) const { __builtin_memcpy(sl.a.raw, &m, sizeof(V)); sl.idx = 0; }
};
template <int DT, bool COMP> struct Acc<DT, NCA_REDUCE_MAX, COMP> : AccExt<DT, true> {};
template <int DT, bool COMP> struct Acc<DT, NCA_REDUCE_MIN, COMP> : AccExt<DT, false> {};

template <int DT, bool IsMax> struct AccArg {
  typedef typename RT<DT>::V V;
  V m;
  nca_i64 i;
  __device__ void init() { m = IsMax ? nca_ident<typename RT<DT>::T, V>::lowest() : nca_ident<typename RT<DT>::T, V>::max(); i = -1; }
  // the host fold: replace only when strictly better than the running value
  __device__ void fold(V v, nca_i64 idx) {
    if (IsMax ? (v > m) : (v < m)) { m = v; i = idx; }
  }
  __device__ void wave() { nca_wave_arg<V, IsMax>(m, i); }
  __device__ void merge(const AccArg& o) { nca_arg_merge<V, IsMax>(m, i, o.m, o.i); }
  __device__ void to_slot(Slot& sl) const { __builtin_memcpy(sl.a.raw, &m, sizeof(V)); sl.idx = i; }
};
template <int DT, bool COMP> struct Acc<DT, NCA_REDUCE_ARGMAX, COMP> : AccArg<DT, true> {};
template <int DT, bool COMP> struct Acc<DT, NCA_REDUCE_ARGMIN, COMP> : AccArg<DT, false> {};

template <int DT, bool IsAll> struct AccLog {
  typedef typename RT<DT>::V V;
  int b;
  __device__ void init() { b = IsAll ? 1 : 0; }
  __device__ void fold(V v, nca_i64) {
    int t = (v != V(0)) ? 1 : 0;
    b = IsAll ? (b & t) : (b | t);
  }
  __device__ void wave() {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      int o = __shfl_xor(b, m, 64);
      b = IsAll ? (b & o) : (b | o);
    }
  }
  __device__ void merge(const AccLog& o) { b = IsAll ? (b & o.b) : (b | o.b); }
  __device__ void to_slot(Slot& sl) const { sl.a.i = b; sl.idx = 0; }
};
template <int DT, bool COMP> struct Acc<DT, NCA_REDUCE_ALL, COMP> : AccLog<DT, true> {};
template <int DT, bool COMP> struct Acc<DT, NCA_REDUCE_ANY, COMP> : AccLog<DT, false> {};

// Combine per-wave accumulators through LDS; thread 0 returns the block result.
template <typename A>
__device__ void block_combine(A& a, A* lds) {
  a.wave();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) lds[wid] = a;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < nw; ++w) a.merge(lds[w]);
}

// ---- final combine of slots -----------------------------------------------------------------------
// Combines `n` slots (in index order, per-thread strided then butterflies: a fixed order)
// and writes the reduction result as the op's result dtype.
template <int DT, int OP>
__device__ void combine_slots_and_store(const Slot* slots, int n, nca_i64 count, void* out) {
  typedef typename RT<DT>::V V;
  typedef typename RT<DT>::S S;
  __shared__ Slot lds[16];
  const int tid = threadIdx.x;
  Slot acc;
  // per-thread fold of slots tid, tid+bs, ...
  if constexpr (OP == NCA_REDUCE_SUM || OP == NCA_REDUCE_MEAN) {
    if constexpr (RT<DT>::fp) {
      double s = 0.0;
      for (int k = tid; k < n; k += blockDim.x) s += slots[k].a.d;
      s = nca_wave_sum(s);
      if ((tid & 63) == 0) lds[tid >> 6].a.d = s;
      __syncthreads();
      if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += lds[w].a.d;
        if constexpr (OP == NCA_REDUCE_MEAN) *reinterpret_cast<double*>(out) = count > 0 ? t / (double)count : __builtin_nan("");
        else *reinterpret_cast<typename RT<DT>::SO*>(out) = static_cast<typename RT<DT>::SO>(t);
      }
    } else {
      S s = 0;
      for (int k = tid; k < n; k += blockDim.x) s += static_cast<S>(slots[k].a.u);
      s = nca_wave_sum(s);
      if ((tid & 63) == 0) lds[tid >> 6].a.u = static_cast<unsigned long long>(s);
      __syncthreads();
      if (tid == 0) {
        S t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += static_cast<S>(lds[w].a.u);
        if constexpr (OP == NCA_REDUCE_MEAN) *reinterpret_cast<double*>(out) = count > 0 ? (double)t / (double)count : __builtin_nan("");
        else *reinterpret_cast<typename RT<DT>::SO*>(out) = static_cast<typename RT<DT>::SO>(t);
      }
    }
  } else if constexpr (OP == NCA_REDUCE_MAX || OP == NCA_REDUCE_MIN) {
    constexpr bool mx = OP == NCA_REDUCE_MAX;
    V m = mx ? nca_ident<typename RT<DT>::T, V>::lowest() : nca_ident<typename RT<DT>::T, V>::max();
    for (int k = tid; k < n; k += blockDim.x) {
      V v;
      __builtin_memcpy(&v, slots[k].a.raw, sizeof(V));
      m = mx ? nca_maxop(m, v) : nca_minop(m, v);
    }
    m = mx ? nca_wave_max(m) : nca_wave_min(m);
    if ((tid & 63) == 0) __builtin_memcpy(lds[tid >> 6].a.raw, &m, sizeof(V));
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
        V v;
        __builtin_memcpy(&v, lds[w].a.raw, sizeof(V));
        m = mx ? nca_maxop(m, v) : nca_minop(m, v);
      }
      nca_st(reinterpret_cast<typename RT<DT>::T*>(out), m);
    }
  } else if constexpr (OP == NCA_REDUCE_ARGMAX || OP == NCA_REDUCE_ARGMIN) {
    constexpr bool mx = OP == NCA_REDUCE_ARGMAX;
    V m = mx ? nca_ident<typename RT<DT>::T, V>::lowest() : nca_ident<typename RT<DT>::T, V>::max();
    nca_i64 i = -1;
    for (int k = tid; k < n; k += blockDim.x) {
      V v;
      __builtin_memcpy(&v, slots[k].a.raw, sizeof(V));
      nca_arg_merge<V, mx>(m, i, v, slots[k].idx);
    }
    nca_wave_arg<V, mx>(m, i);
    if ((tid & 63) == 0) { __builtin_memcpy(lds[tid >> 6].a.raw, &m, sizeof(V)); lds[tid >> 6].idx = i; }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
        V v;
        __builtin_memcpy(&v, lds[w].a.raw, sizeof(V));
        nca_arg_merge<V, mx>(m, i, v, lds[w].idx);
      }
      *reinterpret_cast<long long*>(out) = i < 0 ? 0 : i;
    }
  } else {
    constexpr bool all = OP == NCA_REDUCE_ALL;
    int b = all ? 1 : 0;
    for (int k = tid; k < n; k += blockDim.x) b = all ? (b & (int)slots[k].a.i) : (b | (int)slots[k].a.i);
#pragma unroll
    for (int mm = 32; mm > 0; mm >>= 1) {
      int o = __shfl_xor(b, mm, 64);
      b = all ? (b & o) : (b | o);
    }
    if ((tid & 63) == 0) lds[tid >> 6].a.i = b;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = all ? (b & (int)lds[w].a.i) : (b | (int)lds[w].a.i);
      *reinterpret_cast<bool*>(out) = b != 0;
    }
  }
}

// ---- full reduction kernel ---------------------------------------------------------------------------
struct NdDesc {
  nca_i64 shape[NCA_MAX_NDIM];
  nca_i64 strides[NCA_MAX_NDIM];  // bytes (the dereference chain's strides for an indirect view)
  nca_i64 subs[NCA_MAX_NDIM];     // indirect views: >= 0 after an axis whose step ends in a pointer read
  int ndim;
  int ind;                        // an indirect view (NC pointer axis or SO suboffsets)
};

// Address of the element with C-order linear index idx.  Indirect views walk their chain
// (SOArrayPolicy::advance, layout.hh:677-700; an NC pointer axis is one such step).
template <int DT>
__device__ __forceinline__ const typename RT<DT>::T* nd_ptr(const char* base, const NdDesc& d, nca_i64 idx) {
  if (d.ind) {
    nca_i64 c[NCA_MAX_NDIM];
    for (int k = d.ndim - 1; k >= 0; --k) {
      c[k] = idx % d.shape[k];
      idx /= d.shape[k];
    }
    const char* p = base;
    for (int k = 0; k < d.ndim; ++k) {
      p += c[k] * d.strides[k];
      if (d.subs[k] >= 0) p = nca_chase(p, d.subs[k]);
    }
    return reinterpret_cast<const typename RT<DT>::T*>(p);
  }
  nca_i64 off = 0;
  for (int k = d.ndim - 1; k >= 0; --k) {
    nca_i64 c = idx % d.shape[k];
    idx /= d.shape[k];
    off += c * d.strides[k];
  }
  return reinterpret_cast<const typename RT<DT>::T*>(base + off);
}

// CONTIG: x is contiguous and 16-byte aligned (vector path), else generic ND path.
template <int DT, int OP, bool COMP, bool CONTIG>
__global__ __launch_bounds__(256) void k_reduce_full(const char* __restrict__ base, NdDesc d, nca_i64 n,
                                                     Slot* __restrict__ partials, unsigned* __restrict__ counter,
                                                     void* __restrict__ out) {
  typedef typename RT<DT>::T T;
  typedef typename RT<DT>::V V;
  Acc<DT, OP, COMP> a;
  a.init();
  const nca_i64 tid = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x;
  const nca_i64 gs = (nca_i64)gridDim.x * blockDim.x;
  if constexpr (CONTIG) {
    constexpr int E = (16 / sizeof(T)) > 0 ? (16 / sizeof(T)) : 1;
    const T* x = reinterpret_cast<const T*>(base);
    const nca_i64 nv = n / E;
    typedef nca_vec<T, E> VT;
    nca_i64 i = tid;
    // two vectors in flight per iteration; streaming loads (every element is read once)
    for (; i + gs < nv; i += 2 * gs) {
      const VT v0 = nca_load_nt(reinterpret_cast<const VT*>(x) + i);
      const VT v1 = nca_load_nt(reinterpret_cast<const VT*>(x) + i + gs);
#pragma unroll
      for (int j = 0; j < E; ++j) a.fold(nca_cvt<V, T>::from(v0.v[j]), i * E + j);
#pragma unroll
      for (int j = 0; j < E; ++j) a.fold(nca_cvt<V, T>::from(v1.v[j]), (i + gs) * E + j);
    }
    for (; i < nv; i += gs) {
      const VT v0 = nca_load_nt(reinterpret_cast<const VT*>(x) + i);
#pragma unroll
      for (int j = 0; j < E; ++j) a.fold(nca_cvt<V, T>::from(v0.v[j]), i * E + j);
    }
    for (nca_i64 k = nv * E + tid; k < n; k += gs) a.fold(nca_cvt<V, T>::from(x[k]), k);
  } else {
    for (nca_i64 k = tid; k < n; k += gs) a.fold(nca_ld<V>(nd_ptr<DT>(base, d, k)), k);
  }
  __shared__ Acc<DT, OP, COMP> lds[16];
  block_combine(a, lds);
  if (threadIdx.x == 0) a.to_slot(partials[blockIdx.x]);
}

// One block folds the per-block partials of k_reduce_full in block order.
template <int DT, int OP>
__global__ __launch_bounds__(256) void k_reduce_full_combine(const Slot* __restrict__ partials, int nblocks, nca_i64 n,
                                                             void* __restrict__ out) {
  combine_slots_and_store<DT, OP>(partials, nblocks, n, out);
}

// ---- axis reductions -------------------------------------------------------------------------------------
// Column kernel: input [KO][R][KI] contiguous; thread owns E consecutive inner columns
// (a 16-byte vector) of one outer index and folds rows r in [r0, r1) of chunk
// blockIdx.y.  With S == 1 it stores the result, else it writes one slot per column.
template <int DT, int OP, bool COMP>
__global__ __launch_bounds__(256) void k_reduce_cols(const typename RT<DT>::T* __restrict__ x, nca_i64 KO, nca_i64 R,
                                                     nca_i64 KI, int S, Slot* __restrict__ slots,
                                                     void* __restrict__ out, nca_i64 count) {
  typedef typename RT<DT>::T T;
  typedef typename RT<DT>::V V;
  constexpr int E = (16 / sizeof(T)) > 0 ? (16 / sizeof(T)) : 1;
  typedef nca_vec<T, E> VT;
  const nca_i64 groups_per_row = KI / E;  // host guarantees KI % E == 0 and alignment
  const nca_i64 g = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= KO * groups_per_row) return;
  const nca_i64 ko = g / groups_per_row;
  const nca_i64 ki0 = (g - ko * groups_per_row) * E;
  const int s = blockIdx.y;
  const nca_i64 rows_per = (R + S - 1) / S;
  const nca_i64 r0 = s * rows_per;
  const nca_i64 r1 = r0 + rows_per < R ? r0 + rows_per : R;
  Acc<DT, OP, COMP> a[E];
#pragma unroll
  for (int j = 0; j < E; ++j) a[j].init();
  const T* base = x + ko * R * KI + ki0;
  nca_i64 r = r0;
  for (; r + 3 < r1; r += 4) {
    const VT v0 = nca_load_nt(reinterpret_cast<const VT*>(base + (r + 0) * KI));
    const VT v1 = nca_load_nt(reinterpret_cast<const VT*>(base + (r + 1) * KI));
    const VT v2 = nca_load_nt(reinterpret_cast<const VT*>(base + (r + 2) * KI));
    const VT v3 = nca_load_nt(reinterpret_cast<const VT*>(base + (r + 3) * KI));
    const nca_i64 lin = (ko * R + r) * KI + ki0;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      a[j].fold(nca_cvt<V, T>::from(v0.v[j]), lin + j);
      a[j].fold(nca_cvt<V, T>::from(v1.v[j]), lin + KI + j);
      a[j].fold(nca_cvt<V, T>::from(v2.v[j]), lin + 2 * KI + j);
      a[j].fold(nca_cvt<V, T>::from(v3.v[j]), lin + 3 * KI + j);
    }
  }
  for (; r < r1; ++r) {
    const VT v0 = nca_load_nt(reinterpret_cast<const VT*>(base + r * KI));
    const nca_i64 lin = (ko * R + r) * KI + ki0;
#pragma unroll
    for (int j = 0; j < E; ++j) a[j].fold(nca_cvt<V, T>::from(v0.v[j]), lin + j);
  }
  const nca_i64 K = KO * KI;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const nca_i64 o = ko * KI + ki0 + j;
    Slot sl;
    a[j].to_slot(sl);
    if (S > 1) {
      slots[(nca_i64)s * K + o] = sl;
    } else {
      // single chunk: store directly (same conversions as the slot combine)
      if constexpr (OP == NCA_REDUCE_SUM) {
        if constexpr (RT<DT>::fp) reinterpret_cast<typename RT<DT>::SO*>(out)[o] = static_cast<typename RT<DT>::SO>(sl.a.d);
        else reinterpret_cast<typename RT<DT>::SO*>(out)[o] = static_cast<typename RT<DT>::SO>(static_cast<typename RT<DT>::S>(sl.a.u));
      } else if constexpr (OP == NCA_REDUCE_MEAN) {
        double t = RT<DT>::fp ? sl.a.d : (double)static_cast<typename RT<DT>::S>(sl.a.u);
        reinterpret_cast<double*>(out)[o] = count > 0 ? t / (double)count : __builtin_nan("");
      } else if constexpr (OP == NCA_REDUCE_MAX || OP == NCA_REDUCE_MIN) {
        V v;
        __builtin_memcpy(&v, sl.a.raw, sizeof(V));
        nca_st(reinterpret_cast<T*>(out) + o, v);
      } else if constexpr (OP == NCA_REDUCE_ARGMAX || OP == NCA_REDUCE_ARGMIN) {
        reinterpret_cast<long long*>(out)[o] = sl.idx < 0 ? 0 : sl.idx;
      } else {
        reinterpret_cast<bool*>(out)[o] = sl.a.i != 0;
      }
    }
  }
}

// Combine S column slots per output in chunk order and store.
template <int DT, int OP>
__global__ __launch_bounds__(256) void k_combine_cols(const Slot* __restrict__ slots, int S, nca_i64 K,
                                                      void* __restrict__ out, nca_i64 count) {
  typedef typename RT<DT>::V V;
  typedef typename RT<DT>::S Sx;
  const nca_i64 o = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= K) return;
  if constexpr (OP == NCA_REDUCE_SUM || OP == NCA_REDUCE_MEAN) {
    if constexpr (RT<DT>::fp) {
      double t = 0.0;
      for (int s = 0; s < S; ++s) t += slots[(nca_i64)s * K + o].a.d;
      if constexpr (OP == NCA_REDUCE_MEAN) reinterpret_cast<double*>(out)[o] = count > 0 ? t / (double)count : __builtin_nan("");
      else reinterpret_cast<typename RT<DT>::SO*>(out)[o] = static_cast<typename RT<DT>::SO>(t);
    } else {
      Sx t = 0;
      for (int s = 0; s < S; ++s) t += static_cast<Sx>(slots[(nca_i64)s * K + o].a.u);
      if constexpr (OP == NCA_REDUCE_MEAN) reinterpret_cast<double*>(out)[o] = count > 0 ? (double)t / (double)count : __builtin_nan("");
      else reinterpret_cast<typename RT<DT>::SO*>(out)[o] = static_cast<typename RT<DT>::SO>(t);
    }
  } else if constexpr (OP == NCA_REDUCE_MAX || OP == NCA_REDUCE_MIN) {
    constexpr bool mx = OP == NCA_REDUCE_MAX;
    V m = mx ? nca_ident<typename RT<DT>::T, V>::lowest() : nca_ident<typename RT<DT>::T, V>::max();
    for (int s = 0; s < S; ++s) {
      V v;
      __builtin_memcpy(&v, slots[(nca_i64)s * K + o].a.raw, sizeof(V));
      m = mx ? nca_maxop(m, v) : nca_minop(m, v);
    }
    nca_st(reinterpret_cast<typename RT<DT>::T*>(out) + o, m);
  } else if constexpr (OP == NCA_REDUCE_ARGMAX || OP == NCA_REDUCE_ARGMIN) {
    constexpr bool mx = OP == NCA_REDUCE_ARGMAX;
    V m = mx ? nca_ident<typename RT<DT>::T, V>::lowest() : nca_ident<typename RT<DT>::T, V>::max();
    nca_i64 i = -1;
    for (int s = 0; s < S; ++s) {
      V v;
      __builtin_memcpy(&v, slots[(nca_i64)s * K + o].a.raw, sizeof(V));
      nca_arg_merge<V, mx>(m, i, v, slots[(nca_i64)s * K + o].idx);
    }
    reinterpret_cast<long long*>(out)[o] = i < 0 ? 0 : i;
  } else {
    constexpr bool all = OP == NCA_REDUCE_ALL;
    int b = all ? 1 : 0;
    for (int s = 0; s < S; ++s) b = all ? (b & (int)slots[(nca_i64)s * K + o].a.i) : (b | (int)slots[(nca_i64)s * K + o].a.i);
    reinterpret_cast<bool*>(out)[o] = b != 0;
  }
}

// Row kernel: input [KO][R] (rows contiguous); one wavefront per output row.
template <int DT, int OP, bool COMP>
__global__ __launch_bounds__(256) void k_reduce_rows(const typename RT<DT>::T* __restrict__ x, nca_i64 KO, nca_i64 R,
                                                     void* __restrict__ out) {
  typedef typename RT<DT>::T T;
  typedef typename RT<DT>::V V;
  const int lane = threadIdx.x & 63;
  const nca_i64 row = (nca_i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= KO) return;  // whole wavefronts exit together
  Acc<DT, OP, COMP> a;
  a.init();
  const T* p = x + row * R;
  for (nca_i64 r = lane; r < R; r += 64) a.fold(nca_cvt<V, T>::from(p[r]), row * R + r);
  a.wave();
  if (lane == 0) {
    Slot sl;
    a.to_slot(sl);
    if constexpr (OP == NCA_REDUCE_SUM) {
      if constexpr (RT<DT>::fp) reinterpret_cast<typename RT<DT>::SO*>(out)[row] = static_cast<typename RT<DT>::SO>(sl.a.d);
      else reinterpret_cast<typename RT<DT>::SO*>(out)[row] = static_cast<typename RT<DT>::SO>(static_cast<typename RT<DT>::S>(sl.a.u));
    } else if constexpr (OP == NCA_REDUCE_MEAN) {
      double t = RT<DT>::fp ? sl.a.d : (double)static_cast<typename RT<DT>::S>(sl.a.u);
      reinterpret_cast<double*>(out)[row] = R > 0 ? t / (double)R : __builtin_nan("");
    } else if constexpr (OP == NCA_REDUCE_MAX || OP == NCA_REDUCE_MIN) {
      V v;
      __builtin_memcpy(&v, sl.a.raw, sizeof(V));
      nca_st(reinterpret_cast<T*>(out) + row, v);
    } else if constexpr (OP == NCA_REDUCE_ARGMAX || OP == NCA_REDUCE_ARGMIN) {
      reinterpret_cast<long long*>(out)[row] = sl.idx < 0 ? 0 : sl.idx;
    } else {
      reinterpret_cast<bool*>(out)[row] = sl.a.i != 0;
    }
  }
}

// Generic kernel: one thread per output; kept and reduced axes given separately with
// byte strides and the logical C-order strides (for argmax indices).
struct AxesDesc {
  nca_i64 kshape[NCA_MAX_NDIM], kstride[NCA_MAX_NDIM], klin[NCA_MAX_NDIM];
  nca_i64 rshape[NCA_MAX_NDIM], rstride[NCA_MAX_NDIM], rlin[NCA_MAX_NDIM];
  nca_i64 ostride[NCA_MAX_NDIM];  // output byte strides per kept axis
  int nk, nr;
};

template <int DT, int OP>
__device__ void store_axis_result(const Slot& sl, char* po, nca_i64 nred);

template <int DT, int OP, bool COMP>
__global__ __launch_bounds__(256) void k_reduce_generic(const char* __restrict__ base, AxesDesc d, nca_i64 nout,
                                                        nca_i64 nred, char* __restrict__ out) {
  typedef typename RT<DT>::T T;
  typedef typename RT<DT>::V V;
  const nca_i64 o = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  nca_i64 rem = o, koff = 0, klin = 0, ooff = 0;
  for (int k = d.nk - 1; k >= 0; --k) {
    nca_i64 c = rem % d.kshape[k];
    rem /= d.kshape[k];
    koff += c * d.kstride[k];
    klin += c * d.klin[k];
    ooff += c * d.ostride[k];
  }
  Acc<DT, OP, COMP> a;
  a.init();
  for (nca_i64 r = 0; r < nred; ++r) {
    nca_i64 rr = r, roff = 0, rlin = 0;
    for (int k = d.nr - 1; k >= 0; --k) {
      nca_i64 c = rr % d.rshape[k];
      rr /= d.rshape[k];
      roff += c * d.rstride[k];
      rlin += c * d.rlin[k];
    }
    a.fold(nca_ld<V>(reinterpret_cast<const T*>(base + koff + roff)), klin + rlin);
  }
  Slot sl;
  a.to_slot(sl);
  store_axis_result<DT, OP>(sl, out + ooff, nred);
}

// Store one axis-reduction result (a folded accumulator in slot form) at po.
template <int DT, int OP>
__device__ void store_axis_result(const Slot& sl, char* po, nca_i64 nred) {
  typedef typename RT<DT>::T T;
  typedef typename RT<DT>::V V;
  if constexpr (OP == NCA_REDUCE_SUM) {
    if constexpr (RT<DT>::fp) *reinterpret_cast<typename RT<DT>::SO*>(po) = static_cast<typename RT<DT>::SO>(sl.a.d);
    else *reinterpret_cast<typename RT<DT>::SO*>(po) = static_cast<typename RT<DT>::SO>(static_cast<typename RT<DT>::S>(sl.a.u));
  } else if constexpr (OP == NCA_REDUCE_MEAN) {
    double t = RT<DT>::fp ? sl.a.d : (double)static_cast<typename RT<DT>::S>(sl.a.u);
    *reinterpret_cast<double*>(po) = nred > 0 ? t / (double)nred : __builtin_nan("");
  } else if constexpr (OP == NCA_REDUCE_MAX || OP == NCA_REDUCE_MIN) {
    V v;
    __builtin_memcpy(&v, sl.a.raw, sizeof(V));
    nca_st(reinterpret_cast<T*>(po), v);
  } else if constexpr (OP == NCA_REDUCE_ARGMAX || OP == NCA_REDUCE_ARGMIN) {
    *reinterpret_cast<long long*>(po) = sl.idx < 0 ? 0 : sl.idx;
  } else {
    *reinterpret_cast<bool*>(po) = sl.a.i != 0;
  }
}

// Axis reductions of an indirect view (NC pointer axis or SO suboffsets): the element at full
// coordinate c is reached through the view's dereference chain.  G threads per output (1, or a
// block of 256 for few, long reductions); each folds reduced elements r = t, t + G, ... in C
// order, so a fold meets the linear input indices in increasing order (argmax/argmin keep the
// first occurrence), then the block combines its threads' accumulators.
struct ChainAxesDesc {
  nca_i64 shape[NCA_MAX_NDIM];
  nca_i64 strides[NCA_MAX_NDIM];  // chain byte strides
  nca_i64 subs[NCA_MAX_NDIM];     // >= 0: pointer read after this axis's step
  nca_i64 lin[NCA_MAX_NDIM];      // C-order linear stride of each axis
  nca_i64 ostride[NCA_MAX_NDIM];  // output byte stride of each kept axis
  int red[NCA_MAX_NDIM];          // 1: reduced axis
  int ndim;
};

template <int DT, int OP, bool COMP, int G>
__global__ __launch_bounds__(256) void k_reduce_axes_chain(const char* __restrict__ base, ChainAxesDesc d, nca_i64 nout,
                                                           nca_i64 nred, char* __restrict__ out) {
  typedef typename RT<DT>::T T;
  typedef typename RT<DT>::V V;
  const nca_i64 o = G == 1 ? (nca_i64)blockIdx.x * blockDim.x + threadIdx.x : (nca_i64)blockIdx.x;
  const int t = G == 1 ? 0 : static_cast<int>(threadIdx.x);
  if (o >= nout) return;  // G == 1: whole threads; G == 256: whole blocks
  nca_i64 kc[NCA_MAX_NDIM];
  nca_i64 ooff = 0;
  {
    nca_i64 rem = o;
    for (int k = d.ndim - 1; k >= 0; --k) {
      if (d.red[k]) continue;
      kc[k] = rem % d.shape[k];
      rem /= d.shape[k];
      ooff += kc[k] * d.ostride[k];
    }
  }
  Acc<DT, OP, COMP> a;
  a.init();
  for (nca_i64 r = t; r < nred; r += G) {
    nca_i64 c[NCA_MAX_NDIM];
    nca_i64 rem = r;
    for (int k = d.ndim - 1; k >= 0; --k) {
      if (d.red[k]) {
        c[k] = rem % d.shape[k];
        rem /= d.shape[k];
      } else {
        c[k] = kc[k];
      }
    }
    const char* p = base;
    nca_i64 lin = 0;
    for (int k = 0; k < d.ndim; ++k) {
      p += c[k] * d.strides[k];
      if (d.subs[k] >= 0) p = nca_chase(p, d.subs[k]);
      lin += c[k] * d.lin[k];
    }
    a.fold(nca_ld<V>(reinterpret_cast<const T*>(p)), lin);
  }
  Slot sl;
  if constexpr (G == 1) {
    a.to_slot(sl);
    store_axis_result<DT, OP>(sl, out + ooff, nred);
  } else {
    __shared__ Acc<DT, OP, COMP> lds[16];
    block_combine(a, lds);
    if (threadIdx.x == 0) {
      a.to_slot(sl);
      store_axis_result<DT, OP>(sl, out + ooff, nred);
    }
  }
}

// ---- finalize kernels for fused reductions (expressions, stencils) ------------------------
// Fused kernels write one 8-byte slot per block: float64 sums for floating values, the
// 64-bit integer sum otherwise, and the value (compute type) for max/min.
template <int DT, int KIND>
__device__ void finalize_scalar_body(const void* __restrict__ partials, int n, void* __restrict__ out) {
  typedef typename RT<DT>::V V;
  const int tid = threadIdx.x;
  __shared__ double ld[16];
  __shared__ unsigned long long lu[16];
  __shared__ V lv[16];
  if constexpr (KIND == NCA_FUSE_SUM) {
    if constexpr (RT<DT>::fp) {
      double s = 0.0;
      for (int k = tid; k < n; k += blockDim.x) s += reinterpret_cast<const double*>(partials)[k];
      s = nca_wave_sum(s);
      if ((tid & 63) == 0) ld[tid >> 6] = s;
      __syncthreads();
      if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += ld[w];
        *reinterpret_cast<typename RT<DT>::SO*>(out) = static_cast<typename RT<DT>::SO>(t);
      }
    } else {
      typedef typename RT<DT>::S S;
      S s = 0;
      for (int k = tid; k < n; k += blockDim.x) s += reinterpret_cast<const S*>(partials)[k];
      s = nca_wave_sum(s);
      if ((tid & 63) == 0) lu[tid >> 6] = static_cast<unsigned long long>(s);
      __syncthreads();
      if (tid == 0) {
        S t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += static_cast<S>(lu[w]);
        *reinterpret_cast<typename RT<DT>::SO*>(out) = static_cast<typename RT<DT>::SO>(t);
      }
    }
  } else {
    constexpr bool mx = KIND == NCA_FUSE_MAX;
    V m = mx ? nca_ident<typename RT<DT>::T, V>::lowest() : nca_ident<typename RT<DT>::T, V>::max();
    for (int k = tid; k < n; k += blockDim.x) {
      V v;
      __builtin_memcpy(&v, reinterpret_cast<const char*>(partials) + 8 * (size_t)k, sizeof(V));
      m = mx ? nca_maxop(m, v) : nca_minop(m, v);
    }
    m = mx ? nca_wave_max(m) : nca_wave_min(m);
    if ((tid & 63) == 0) lv[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = mx ? nca_maxop(m, lv[w]) : nca_minop(m, lv[w]);
      nca_st(reinterpret_cast<typename RT<DT>::T*>(out), m);
    }
  }
}

template <int DT, int KIND>
__global__ __launch_bounds__(256) void k_finalize_scalar(const void* __restrict__ partials, int n, void* __restrict__ out) {
  finalize_scalar_body<DT, KIND>(partials, n, out);
}

// Column sums from z-run partials (blocks [0, ceil(K/256))) and the global reductions
// (the last block): one launch for everything a fused stencil/expression leaves over.
template <int DT>
__global__ __launch_bounds__(256) void k_finalize_fused(const double* __restrict__ colpart, int nzc, nca_i64 K,
                                                        void* __restrict__ colout, int col_f64, int fuse,
                                                        const void* __restrict__ psum, const void* __restrict__ pmax,
                                                        const void* __restrict__ pmin, int nblocks, void* sum_out,
                                                        void* max_out, void* min_out) {
  const nca_i64 ncb = (K + 255) / 256;
  if ((nca_i64)blockIdx.x < ncb) {
    const nca_i64 k = (nca_i64)blockIdx.x * 256 + threadIdx.x;
    if (k < K) {
      double t = 0.0;
      for (int q = 0; q < nzc; ++q) t += colpart[(nca_i64)q * K + k];
      if (col_f64) reinterpret_cast<double*>(colout)[k] = t;
      else reinterpret_cast<float*>(colout)[k] = static_cast<float>(t);
    }
    return;
  }
  if ((fuse & NCA_FUSE_SUM) && sum_out) finalize_scalar_body<DT, NCA_FUSE_SUM>(psum, nblocks, sum_out);
  if ((fuse & NCA_FUSE_MAX) && max_out) finalize_scalar_body<DT, NCA_FUSE_MAX>(pmax, nblocks, max_out);
  if ((fuse & NCA_FUSE_MIN) && min_out) finalize_scalar_body<DT, NCA_FUSE_MIN>(pmin, nblocks, min_out);
}

template <typename OutT>
__global__ __launch_bounds__(256) void k_finalize_columns(const double* __restrict__ p, int S, nca_i64 K,
                                                          OutT* __restrict__ out) {
  const nca_i64 k = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double t = 0.0;
  for (int s = 0; s < S; ++s) t += p[(nca_i64)s * K + k];
  nca_st(out + k, t);
}

// ---- host dispatch -------------------------------------------------------------------------------------
template <template <int> class F, typename... Args>
static void dispatch_dtype(int32_t dt, Args&&... args) {
  switch (dt) {
    case NCA_BOOL: F<NCA_BOOL>::run(args...); break;
    case NCA_CHAR: F<NCA_CHAR>::run(args...); break;
    case NCA_UINT8: F<NCA_UINT8>::run(args...); break;
    case NCA_UINT16: F<NCA_UINT16>::run(args...); break;
    case NCA_UINT32: F<NCA_UINT32>::run(args...); break;
    case NCA_UINT64: F<NCA_UINT64>::run(args...); break;
    case NCA_INT8: F<NCA_INT8>::run(args...); break;
    case NCA_INT16: F<NCA_INT16>::run(args...); break;
    case NCA_INT32: F<NCA_INT32>::run(args...); break;
    case NCA_INT64: F<NCA_INT64>::run(args...); break;
    case NCA_FLOAT32: F<NCA_FLOAT32>::run(args...); break;
    case NCA_FLOAT64: F<NCA_FLOAT64>::run(args...); break;
    case NCA_BFLOAT16: F<NCA_BFLOAT16>::run(args...); break;
    case NCA_FLOAT16: F<NCA_FLOAT16>::run(args...); break;
    default: NCA_THROW(NCA_ERR_UNSUPPORTED, "reduction over an unsupported dtype");
  }
}

template <int DT, int OP, typename F>
static void with_comp(int mode, F&& f) {
  if (RT<DT>::fp && mode == NCA_ACCUM_COMPENSATED && (OP == NCA_REDUCE_SUM || OP == NCA_REDUCE_MEAN))
    f(std::true_type{});
  else
    f(std::false_type{});
}

template <typename F>
static void dispatch_op(int op, F&& f) {
  switch (op) {
    case NCA_REDUCE_SUM: f(std::integral_constant<int, NCA_REDUCE_SUM>{}); break;
    case NCA_REDUCE_MAX: f(std::integral_constant<int, NCA_REDUCE_MAX>{}); break;
    case NCA_REDUCE_MIN: f(std::integral_constant<int, NCA_REDUCE_MIN>{}); break;
    case NCA_REDUCE_ARGMAX: f(std::integral_constant<int, NCA_REDUCE_ARGMAX>{}); break;
    case NCA_REDUCE_ARGMIN: f(std::integral_constant<int, NCA_REDUCE_ARGMIN>{}); break;
    case NCA_REDUCE_MEAN: f(std::integral_constant<int, NCA_REDUCE_MEAN>{}); break;
    case NCA_REDUCE_ALL: f(std::integral_constant<int, NCA_REDUCE_ALL>{}); break;
    case NCA_REDUCE_ANY: f(std::integral_constant<int, NCA_REDUCE_ANY>{}); break;
    default: NCA_THROW(NCA_ERR_INVALID, "unknown reduction op " + std::to_string(op));
  }
}

struct FullArgs {
  const nca_view* in;
  int op, mode;
  void* out;
  hipStream_t s;
};

template <int DT> struct RunFull {
  static void run(const FullArgs& a) {
    typedef typename RT<DT>::T T;
    DeviceState& dev = current_device();
    NdDesc d{};
    const int64_t n = view_size(*a.in);
    const char* base = static_cast<const char*>(a.in->data);
    bool contig = false;
    if (view_indirect(*a.in)) {
      // the view's dereference chain, axis by axis (no axes merged across a pointer read)
      const Chain ch = chain_of(*a.in);
      d.ind = 1;
      d.ndim = a.in->ndim;
      for (int k = 0; k < d.ndim; ++k) { d.shape[k] = a.in->shape[k]; d.strides[k] = ch.strides[k]; d.subs[k] = -1; }
      for (size_t i = 0; i < ch.axes.size(); ++i) d.subs[ch.axes[i]] = ch.subs[i];
      base = reinterpret_cast<const char*>(ch.base);
    } else {
      int64_t cs[NCA_MAX_NDIM], cst[NCA_MAX_NDIM];
      d.ndim = coalesce(a.in->shape, a.in->strides, a.in->ndim, cs, cst);
      for (int k = 0; k < d.ndim; ++k) { d.shape[k] = cs[k]; d.strides[k] = cst[k]; }
      contig = (d.ndim == 0) || (d.ndim == 1 && d.strides[0] == (int64_t)sizeof(T));
      if (d.ndim == 0) { d.ndim = 1; d.shape[0] = n; d.strides[0] = sizeof(T); }
    }
    contig = contig && (reinterpret_cast<uintptr_t>(a.in->data) % 16 == 0);
    constexpr int E = (16 / sizeof(T)) > 0 ? (16 / sizeof(T)) : 1;
    int64_t items = contig ? std::max<int64_t>(1, n / E / 2) : std::max<int64_t>(1, n);
    int blocks = static_cast<int>(std::min<int64_t>((items + 255) / 256, (int64_t)dev.n_cus * 4));
    if (blocks < 1) blocks = 1;
    Workspace& w = workspace(a.s, sizeof(Slot) * (size_t)blocks);
    dispatch_op(a.op, [&](auto opc) {
      constexpr int OP = decltype(opc)::value;
      with_comp<DT, OP>(a.mode, [&](auto cc) {
        constexpr bool COMP = decltype(cc)::value;
        {
          ProfScope ps("nca_reduce_full", a.s);
          if (contig)
            hipLaunchKernelGGL((k_reduce_full<DT, OP, COMP, true>), dim3(blocks), dim3(256), 0, a.s,
                               base, d, (nca_i64)n, (Slot*)w.partials, w.counters, a.out);
          else
            hipLaunchKernelGGL((k_reduce_full<DT, OP, COMP, false>), dim3(blocks), dim3(256), 0, a.s,
                               base, d, (nca_i64)n, (Slot*)w.partials, w.counters, a.out);
        }
        ProfScope ps2("nca_reduce_full_combine", a.s);
        hipLaunchKernelGGL((k_reduce_full_combine<DT, OP>), dim3(1), dim3(256), 0, a.s, (const Slot*)w.partials,
                           blocks, (nca_i64)n, a.out);
      });
    });
  }
};

struct AxesArgs {
  const nca_view* in;
  const nca_view* out;
  std::vector<char> reduced;
  int op, mode;
  hipStream_t s;
};

template <int DT> struct RunAxes {
  // an indirect input: k_reduce_axes_chain, a block per output when outputs are few
  static void run_chain(const AxesArgs& a, int64_t nout, int64_t nred) {
    DeviceState& dev = current_device();
    const nca_view& in = *a.in;
    const nca_view& out = *a.out;
    const Chain ch = chain_of(in);
    ChainAxesDesc d{};
    d.ndim = in.ndim;
    int64_t lin = 1;
    for (int k = in.ndim - 1; k >= 0; --k) { d.lin[k] = lin; lin *= in.shape[k]; }
    int oi = 0;
    for (int k = 0; k < in.ndim; ++k) {
      d.shape[k] = in.shape[k];
      d.strides[k] = ch.strides[k];
      d.subs[k] = -1;
      d.red[k] = a.reduced[k] ? 1 : 0;
      d.ostride[k] = a.reduced[k] ? 0 : out.strides[oi++];
    }
    for (size_t i = 0; i < ch.axes.size(); ++i) d.subs[ch.axes[i]] = ch.subs[i];
    const char* base = reinterpret_cast<const char*>(ch.base);
    dispatch_op(a.op, [&](auto opc) {
      constexpr int OP = decltype(opc)::value;
      with_comp<DT, OP>(a.mode, [&](auto cc) {
        constexpr bool COMP = decltype(cc)::value;
        if (nout == 0) return;
        ProfScope ps("nca_reduce_axes", a.s);
        if (nout < static_cast<int64_t>(dev.n_cus) * 256 && nred >= 256) {
          NCA_REQUIRE(nout < (int64_t(1) << 31), "too many outputs");
          hipLaunchKernelGGL((k_reduce_axes_chain<DT, OP, COMP, 256>), dim3(static_cast<unsigned>(nout)), dim3(256), 0, a.s,
                             base, d, (nca_i64)nout, (nca_i64)nred, (char*)out.data);
        } else {
          hipLaunchKernelGGL((k_reduce_axes_chain<DT, OP, COMP, 1>), dim3(static_cast<unsigned>((nout + 255) / 256)),
                             dim3(256), 0, a.s, base, d, (nca_i64)nout, (nca_i64)nred, (char*)out.data);
        }
      });
    });
  }

  static void run(const AxesArgs& a) {
    typedef typename RT<DT>::T T;
    DeviceState& dev = current_device();
    const nca_view& in = *a.in;
    const nca_view& out = *a.out;
    const int nd = in.ndim;
    int64_t nout = 1, nred = 1;
    for (int d = 0; d < nd; ++d) (a.reduced[d] ? nred : nout) *= in.shape[d];
    if (view_indirect(in)) {
      run_chain(a, nout, nred);
      return;
    }
    // Pattern [KO][R][KI] on a contiguous input and output?
    bool fast = is_c_contiguous(in) && is_c_contiguous(out);
    int first_red = -1, last_red = -1;
    for (int d = 0; d < nd; ++d)
      if (a.reduced[d] && in.shape[d] != 1) { if (first_red < 0) first_red = d; last_red = d; }
    for (int d = first_red; fast && d >= 0 && d <= last_red; ++d)
      if (!a.reduced[d] && in.shape[d] != 1) fast = false;  // reduced axes must be one block
    int64_t KO = 1, R = 1, KI = 1;
    if (first_red >= 0) {
      for (int d = 0; d < first_red; ++d) KO *= in.shape[d];
      for (int d = first_red; d <= last_red; ++d) R *= in.shape[d];
      for (int d = last_red + 1; d < nd; ++d) KI *= in.shape[d];
    } else {
      KO = nout;
    }
    constexpr int E = (16 / sizeof(T)) > 0 ? (16 / sizeof(T)) : 1;
    const int64_t count = nred;
    dispatch_op(a.op, [&](auto opc) {
      constexpr int OP = decltype(opc)::value;
      with_comp<DT, OP>(a.mode, [&](auto cc) {
        constexpr bool COMP = decltype(cc)::value;
        if (nout == 0) return;
        ProfScope ps("nca_reduce_axes", a.s);
        if (fast && KI > 1 && KI % E == 0 && reinterpret_cast<uintptr_t>(in.data) % 16 == 0 && R > 0) {
          const int64_t threads = KO * (KI / E);
          int S = 1;
          const int64_t target = (int64_t)dev.n_cus * 1024;
          while (S < 64 && threads * S < target && R / (S * 2) >= 16) S *= 2;
          Workspace& w = workspace(a.s, 1, S > 1 ? sizeof(Slot) * (size_t)S * (size_t)(KO * KI) : 0);
          dim3 grid((unsigned)((threads + 255) / 256), (unsigned)S);
          hipLaunchKernelGGL((k_reduce_cols<DT, OP, COMP>), grid, dim3(256), 0, a.s, (const T*)in.data,
                             (nca_i64)KO, (nca_i64)R, (nca_i64)KI, S, (Slot*)w.big, out.data, (nca_i64)count);
          if (S > 1)
            hipLaunchKernelGGL((k_combine_cols<DT, OP>), dim3((unsigned)((KO * KI + 255) / 256)), dim3(256), 0, a.s,
                               (const Slot*)w.big, S, (nca_i64)(KO * KI), out.data, (nca_i64)count);
        } else if (fast && KI == 1 && R > 0) {
          const int rows_per_block = 4;
          hipLaunchKernelGGL((k_reduce_rows<DT, OP, COMP>), dim3((unsigned)((KO + rows_per_block - 1) / rows_per_block)),
                             dim3(256), 0, a.s, (const T*)in.data, (nca_i64)KO, (nca_i64)R, out.data);
        } else {
          AxesDesc d{};
          int64_t lin = 1;
          std::vector<int64_t> linstride(nd);
          for (int k = nd - 1; k >= 0; --k) { linstride[k] = lin; lin *= in.shape[k]; }
          int oi = 0;
          for (int k = 0; k < nd; ++k) {
            if (a.reduced[k]) {
              d.rshape[d.nr] = in.shape[k]; d.rstride[d.nr] = in.strides[k]; d.rlin[d.nr] = linstride[k]; ++d.nr;
            } else {
              d.kshape[d.nk] = in.shape[k]; d.kstride[d.nk] = in.strides[k]; d.klin[d.nk] = linstride[k];
              d.ostride[d.nk] = out.strides[oi++]; ++d.nk;
            }
          }
          hipLaunchKernelGGL((k_reduce_generic<DT, OP, COMP>), dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, a.s,
                             (const char*)in.data, d, (nca_i64)nout, (nca_i64)nred, (char*)out.data);
        }
      });
    });
  }
};

template <int DT> struct RunFinalizeScalar {
  static void run(int kind, const void* p, int n, void* out, hipStream_t s) {
    ProfScope ps("nca_finalize_scalar", s);
    if (kind == NCA_FUSE_SUM)
      hipLaunchKernelGGL((k_finalize_scalar<DT, NCA_FUSE_SUM>), dim3(1), dim3(256), 0, s, p, n, out);
    else if (kind == NCA_FUSE_MAX)
      hipLaunchKernelGGL((k_finalize_scalar<DT, NCA_FUSE_MAX>), dim3(1), dim3(256), 0, s, p, n, out);
    else
      hipLaunchKernelGGL((k_finalize_scalar<DT, NCA_FUSE_MIN>), dim3(1), dim3(256), 0, s, p, n, out);
  }
};

void launch_finalize_scalar(int kind, int32_t value_dtype, const void* partials, int nblocks, void* out,
                            hipStream_t s) {
  dispatch_dtype<RunFinalizeScalar>(value_dtype, kind, partials, nblocks, out, s);
}

template <int DT> struct RunFinalizeFused {
  static void run(const FinalizeArgs& a, hipStream_t s) {
    const int64_t ncb = (a.K + 255) / 256;
    const bool scal = a.fuse != 0;
    const unsigned grid = static_cast<unsigned>(ncb + (scal ? 1 : 0));
    if (grid == 0) return;
    ProfScope ps("nca_finalize_fused", s);
    hipLaunchKernelGGL((k_finalize_fused<DT>), dim3(grid), dim3(256), 0, s, a.colpart, a.nzc, (nca_i64)a.K, a.colout,
                       a.col_dtype == NCA_FLOAT64 ? 1 : 0, a.fuse, a.psum, a.pmax, a.pmin, a.nblocks, a.sum_out,
                       a.max_out, a.min_out);
  }
};

void launch_finalize_fused(const FinalizeArgs& a, hipStream_t s) {
  NCA_REQUIRE(a.K == 0 || a.col_dtype == NCA_FLOAT32 || a.col_dtype == NCA_FLOAT64,
              "column sums are written as float32 or float64");
  dispatch_dtype<RunFinalizeFused>(a.value_dtype, a, s);
}

void launch_finalize_columns(const double* partials, int S, int64_t K, void* out, int32_t out_dtype,
                             hipStream_t s) {
  dim3 grid((unsigned)((K + 255) / 256));
  switch (out_dtype) {
    case NCA_FLOAT32: hipLaunchKernelGGL(k_finalize_columns<float>, grid, dim3(256), 0, s, partials, S, (nca_i64)K, (float*)out); break;
    case NCA_FLOAT64: hipLaunchKernelGGL(k_finalize_columns<double>, grid, dim3(256), 0, s, partials, S, (nca_i64)K, (double*)out); break;
    default: NCA_THROW(NCA_ERR_UNSUPPORTED, "column sums are written as float32 or float64");
  }
}

}  // namespace nca

using namespace nca;

extern "C" {

int nca_hip_reduce_full(const nca_view* in, int32_t op, int32_t accum_mode, void* out_dev, void* stream) {
  return api_call([&] {
    NCA_REQUIRE(in != nullptr && out_dev != nullptr, "null argument");
    validate_view_nc(*in, "in");
    (void)reduce_result_dtype(op, in->dtype);
    FullArgs a{in, op, accum_mode, out_dev, as_stream(stream)};
    dispatch_dtype<RunFull>(in->dtype, a);
    check_launch("reduce_full");
    finish_call(stream);
  });
}

int nca_hip_reduce_axes(const nca_view* in, const int64_t* axes, int32_t n_axes, int32_t op, int32_t accum_mode,
                        const nca_view* out, void* stream) {
  return api_call([&] {
    NCA_REQUIRE(in != nullptr && out != nullptr, "null argument");
    NCA_REQUIRE(n_axes == 0 || axes != nullptr, "null axes");
    validate_view_nc(*in, "in");
    validate_view(*out, "out");
    AxesArgs a{in, out, std::vector<char>(in->ndim, 0), op, accum_mode, as_stream(stream)};
    for (int i = 0; i < n_axes; ++i)
      if (axes[i] >= 0 && axes[i] < in->ndim) a.reduced[axes[i]] = 1;  // others ignored
    // output: kept axes in order
    int k = 0;
    for (int d = 0; d < in->ndim; ++d) {
      if (a.reduced[d]) continue;
      NCA_REQUIRE(k < out->ndim && out->shape[k] == in->shape[d], "out shape must be the kept axes of in");
      ++k;
    }
    NCA_REQUIRE(k == out->ndim, "out has extra axes");
    NCA_REQUIRE(out->dtype == reduce_result_dtype(op, in->dtype),
                std::string("out dtype must be ") + dtype_info(reduce_result_dtype(op, in->dtype)).name);
    dispatch_dtype<RunAxes>(in->dtype, a);
    check_launch("reduce_axes");
    finish_call(stream);
  });
}

}  // extern "C"
