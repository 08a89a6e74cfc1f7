// Device-side helpers shared by the ahead-of-time kernels (compiled by hipcc) and by the
// kernels generated at run time (this file's text is prepended to every hipRTC source,
// so it must not include any header).  gfx950: 64-lane wavefronts.
#ifndef NCA_DEVICE_HH
#define NCA_DEVICE_HH

typedef long long nca_i64;
typedef unsigned long long nca_u64;

#define NCA_WAVE 64

// ---- storage-only float types ------------------------------------------------------
// bfloat16: the top 16 bits of an IEEE binary32.  float16: IEEE binary16.
struct nca_bf16 { unsigned short bits; };
struct nca_f16 { unsigned short bits; };

__device__ __forceinline__ float nca_bf16_to_f32(unsigned short b) {
  return __uint_as_float(static_cast<unsigned>(b) << 16);
}
// Round to nearest, ties to even; NaN stays a (quiet) NaN.  gfx950 converts in hardware
// (v_cvt_pk_bf16_f32, two values per instruction).
__device__ __forceinline__ unsigned short nca_f32_to_bf16(float f) {
  return __builtin_bit_cast(unsigned short, static_cast<__bf16>(f));
}
__device__ __forceinline__ float nca_f16_to_f32(unsigned short b) {
  _Float16 h;
  __builtin_memcpy(&h, &b, 2);
  return static_cast<float>(h);
}
__device__ __forceinline__ unsigned short nca_f32_to_f16(float f) {
  _Float16 h = static_cast<_Float16>(f);
  unsigned short b;
  __builtin_memcpy(&b, &h, 2);
  return b;
}

// Load one element of storage type T as compute type W (static_cast semantics, as the
// reference's op_traits<T>::cast, op_traits.hh:93-117).
template <typename W, typename T> struct nca_cvt {
  __device__ __forceinline__ static W from(T v) { return static_cast<W>(v); }
};
template <typename W> struct nca_cvt<W, nca_bf16> {
  __device__ __forceinline__ static W from(nca_bf16 v) { return static_cast<W>(nca_bf16_to_f32(v.bits)); }
};
template <typename W> struct nca_cvt<W, nca_f16> {
  __device__ __forceinline__ static W from(nca_f16 v) { return static_cast<W>(nca_f16_to_f32(v.bits)); }
};
// Store compute value W into storage type T.
template <typename T, typename W> struct nca_store_cvt {
  __device__ __forceinline__ static T to(W v) { return static_cast<T>(v); }
};
template <typename W> struct nca_store_cvt<nca_bf16, W> {
  __device__ __forceinline__ static nca_bf16 to(W v) {
    nca_bf16 r; r.bits = nca_f32_to_bf16(static_cast<float>(v)); return r;
  }
};
template <typename W> struct nca_store_cvt<nca_f16, W> {
  __device__ __forceinline__ static nca_f16 to(W v) {
    nca_f16 r; r.bits = nca_f32_to_f16(static_cast<float>(v)); return r;
  }
};

template <typename W, typename T> __device__ __forceinline__ W nca_ld(const T* p) {
  return nca_cvt<W, T>::from(*p);
}
template <typename T, typename W> __device__ __forceinline__ void nca_st(T* p, W v) {
  *p = nca_store_cvt<T, W>::to(v);
}

// One dereference step of an indirect view (SOArrayPolicy::advance, layout.hh:677-700): the
// pointer stored at p, plus a byte offset.
__device__ __forceinline__ const char* nca_chase(const char* p, nca_i64 sub) {
  return *reinterpret_cast<const char* const*>(p) + sub;
}
__device__ __forceinline__ char* nca_chase(char* p, nca_i64 sub) {
  return *reinterpret_cast<char* const*>(p) + sub;
}

// A view addressed by C-order linear index: strides in bytes and, for indirect views, a pointer
// read after the step of each axis whose sub is >= 0 (SOArrayPolicy::advance, layout.hh:677-700;
// an NC pointer axis is one such step).  Up to 10 axes (NCARRAY_MAX_NDIM).
struct nca_ndview {
  const char* base;
  nca_i64 shape[10];
  nca_i64 strides[10];
  nca_i64 subs[10];
  int ndim;
  int ind;
};
__device__ __forceinline__ const char* nca_nd_addr(const nca_ndview& v, nca_i64 lin) {
  nca_i64 c[10];
  for (int k = v.ndim - 1; k >= 0; --k) {
    c[k] = lin % v.shape[k];
    lin /= v.shape[k];
  }
  const char* p = v.base;
  for (int k = 0; k < v.ndim; ++k) {
    p += c[k] * v.strides[k];
    if (v.ind && v.subs[k] >= 0) p = nca_chase(p, v.subs[k]);
  }
  return p;
}

// ---- numeric limits ------------------------------------------------------------------
template <typename T> struct nca_lim;
template <> struct nca_lim<bool> { __device__ static bool lowest() { return false; } __device__ static bool max() { return true; } };
template <> struct nca_lim<signed char> { __device__ static signed char lowest() { return -128; } __device__ static signed char max() { return 127; } };
template <> struct nca_lim<unsigned char> { __device__ static unsigned char lowest() { return 0; } __device__ static unsigned char max() { return 255; } };
template <> struct nca_lim<short> { __device__ static short lowest() { return -32768; } __device__ static short max() { return 32767; } };
template <> struct nca_lim<unsigned short> { __device__ static unsigned short lowest() { return 0; } __device__ static unsigned short max() { return 65535; } };
template <> struct nca_lim<int> { __device__ static int lowest() { return -2147483647 - 1; } __device__ static int max() { return 2147483647; } };
template <> struct nca_lim<unsigned int> { __device__ static unsigned int lowest() { return 0u; } __device__ static unsigned int max() { return 4294967295u; } };
template <> struct nca_lim<long long> { __device__ static long long lowest() { return -9223372036854775807LL - 1; } __device__ static long long max() { return 9223372036854775807LL; } };
template <> struct nca_lim<unsigned long long> { __device__ static unsigned long long lowest() { return 0ull; } __device__ static unsigned long long max() { return 18446744073709551615ull; } };
template <> struct nca_lim<float> { __device__ static float lowest() { return -3.402823466e+38f; } __device__ static float max() { return 3.402823466e+38f; } };
template <> struct nca_lim<double> { __device__ static double lowest() { return -1.7976931348623157e+308; } __device__ static double max() { return 1.7976931348623157e+308; } };
// Identities of max/min over values stored as T and compared in compute type W: T's own
// lowest() / max() (the fill values of the reduction traits, reductions.hh:144-283); for the
// 2-byte float storage types the largest finite values of those formats (bfloat16 0xff7f /
// 0x7f7f, float16 -+65504), which a float32 identity would round to -+inf when stored.
template <typename T, typename W> struct nca_ident {
  __device__ static W lowest() { return static_cast<W>(nca_lim<T>::lowest()); }
  __device__ static W max() { return static_cast<W>(nca_lim<T>::max()); }
};
template <typename W> struct nca_ident<nca_bf16, W> {
  __device__ static W lowest() { return static_cast<W>(-3.38953138925153547590470800371487867e+38f); }
  __device__ static W max() { return static_cast<W>(3.38953138925153547590470800371487867e+38f); }
};
template <typename W> struct nca_ident<nca_f16, W> {
  __device__ static W lowest() { return static_cast<W>(-65504.0f); }
  __device__ static W max() { return static_cast<W>(65504.0f); }
};

// ---- compensated summation ----------------------------------------------------------
// Neumaier's variant of Kahan summation: s is the running sum, c the running
// compensation; the represented value is s + c.
template <typename F>
__device__ __forceinline__ void nca_neumaier_add(F& s, F& c, F x) {
  F t = s + x;
  F ax = x < F(0) ? -x : x;
  F as = s < F(0) ? -s : s;
  F e = (as >= ax) ? ((s - t) + x) : ((x - t) + s);
  c += e;
  s = t;
}
// Kahan summation with the compensation kept as the correction to add (the represented
// value is s + c, as for nca_neumaier_add): fewer operations than Neumaier's 7, the same
// first-order bound (2 eps sum|x|).  Used where the accumulation is on the critical ALU
// path (the march stencil over 2-byte storage).
// Once the running sum is infinite or NaN the correction (inf - inf) would be NaN and, fed
// back through y, turn the sum itself into NaN; the correction is then dropped, so the sum
// follows a plain sum's IEEE value (two operations more).
template <typename F>
__device__ __forceinline__ void nca_kahan_add(F& s, F& c, F x) {
  const F y = x + c;
  const F t = s + y;
  const F e = y - (t - s);
  c = __builtin_isfinite(t) ? e : F(0);
  s = t;
}
// The float64 value of a compensated pair.  When the running sum is infinite or NaN the
// compensation is meaningless (it holds inf - inf) and the value is the sum itself, as a plain
// sum gives: a sum that meets +inf is +inf, one that meets +inf and -inf is NaN.
template <typename F>
__device__ __forceinline__ double nca_comp_value(F s, F c) {
  const double ds = static_cast<double>(s);
  return __builtin_isfinite(ds) ? ds + static_cast<double>(c) : ds;
}
// Knuth's TwoSum: s + e == a + b exactly.
template <typename F>
__device__ __forceinline__ void nca_two_sum(F a, F b, F& s, F& e) {
  s = a + b;
  F bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
// Merge compensated pair (s2, c2) into (s1, c1).
template <typename F>
__device__ __forceinline__ void nca_comp_merge(F& s1, F& c1, F s2, F c2) {
  F s, e;
  nca_two_sum(s1, s2, s, e);
  s1 = s;
  c1 = c1 + c2 + e;
}

// ---- wave / block reductions (64-lane butterflies; every lane ends with the result) --
template <typename T>
__device__ __forceinline__ T nca_shfl_xor(T v, int m) {
  if constexpr (sizeof(T) < 4) return static_cast<T>(__shfl_xor(static_cast<int>(v), m, NCA_WAVE));
  else return __shfl_xor(v, m, NCA_WAVE);
}
__device__ __forceinline__ bool nca_shfl_xor(bool v, int m) { return __shfl_xor(static_cast<int>(v), m, NCA_WAVE) != 0; }

template <typename T>
__device__ __forceinline__ T nca_wave_sum(T v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = v + nca_shfl_xor(v, m);
  return v;
}
template <typename F>
__device__ __forceinline__ void nca_wave_comp(F& s, F& c) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    F s2 = nca_shfl_xor(s, m), c2 = nca_shfl_xor(c, m);
    nca_comp_merge(s, c, s2, c2);
  }
}
// max/min of values.  Integers: the larger (smaller) one.  Floating types: IEEE 754-2019
// maximumNumber / minimumNumber: a NaN operand never wins over a number (the host fold's
// strict comparison never selects NaN, host/reductions.hh:130-146), and -0 orders below +0,
// so the result is the same for every order of combination (a fold, a butterfly or a tree
// of partials all give the same bits).  One v_max_f32 / v_min_f32 on gfx950.
template <typename T> __device__ __forceinline__ T nca_maxop(T a, T b) { return (b > a) ? b : a; }
template <typename T> __device__ __forceinline__ T nca_add2(T a, T b) { return a + b; }
template <typename T> __device__ __forceinline__ T nca_minop(T a, T b) { return (b < a) ? b : a; }
template <> __device__ __forceinline__ float nca_maxop<float>(float a, float b) { return __builtin_fmaximum_numf(a, b); }
template <> __device__ __forceinline__ float nca_minop<float>(float a, float b) { return __builtin_fminimum_numf(a, b); }
template <> __device__ __forceinline__ double nca_maxop<double>(double a, double b) { return __builtin_fmaximum_num(a, b); }
template <> __device__ __forceinline__ double nca_minop<double>(double a, double b) { return __builtin_fminimum_num(a, b); }
// The same operation under the name the pairwise trees of the march stencils use.
template <typename T> __device__ __forceinline__ T nca_maxnum(T a, T b) { return nca_maxop(a, b); }
template <typename T> __device__ __forceinline__ T nca_minnum(T a, T b) { return nca_minop(a, b); }
template <typename T>
__device__ __forceinline__ T nca_wave_max(T v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = nca_maxop(v, nca_shfl_xor(v, m));
  return v;
}
template <typename T>
__device__ __forceinline__ T nca_wave_min(T v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = nca_minop(v, nca_shfl_xor(v, m));
  return v;
}
// (value, index) pairs for argmax/argmin: idx < 0 means "nothing selected yet"; on
// equal values the lower index wins (first occurrence, host/reductions.hh:148-169).
template <typename T, bool IsMax>
__device__ __forceinline__ void nca_arg_merge(T& v, nca_i64& i, T v2, nca_i64 i2) {
  if (i2 < 0) return;
  bool take;
  if (i < 0) take = true;
  else if (IsMax ? (v2 > v) : (v2 < v)) take = true;
  else if (IsMax ? (v > v2) : (v < v2)) take = false;
  else take = i2 < i;
  if (take) { v = v2; i = i2; }
}
template <typename T, bool IsMax>
__device__ __forceinline__ void nca_wave_arg(T& v, nca_i64& i) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    T v2 = nca_shfl_xor(v, m);
    nca_i64 i2 = nca_shfl_xor(i, m);
    nca_arg_merge<T, IsMax>(v, i, v2, i2);
  }
}

// Zeros for loads that fall outside an array: a kernel reads them instead of zeroing a loaded
// value afterwards (a select on the loaded data would make the load's wave wait for it at once).
[[maybe_unused]] static __device__ __attribute__((aligned(16))) unsigned int nca_zero_words[16];
// Where lanes outside an array store a vector (every wave then issues the same stores): written,
// never read.
[[maybe_unused]] static __device__ __attribute__((aligned(16))) unsigned int nca_sink_words[16];

// 128-bit vector of N elements of T (N * sizeof(T) == 16 or smaller power of two).
template <typename T, int N>
struct __attribute__((aligned(sizeof(T) * N))) nca_vec { T v[N]; };

typedef unsigned int nca_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int nca_u32x2 __attribute__((ext_vector_type(2)));

// Streaming (non-temporal) store of a whole vector: the data is not re-read by this
// kernel, so it should not displace cached inputs.
template <typename T, int N>
__device__ __forceinline__ void nca_store_nt(nca_vec<T, N>* p, const nca_vec<T, N>& v) {
  constexpr int B = static_cast<int>(sizeof(T)) * N;
  if constexpr (B == 16) {
    nca_u32x4 w;
    __builtin_memcpy(&w, &v, 16);
    __builtin_nontemporal_store(w, reinterpret_cast<nca_u32x4*>(p));
  } else if constexpr (B == 32) {
    nca_u32x4 w0, w1;
    __builtin_memcpy(&w0, &v, 16);
    __builtin_memcpy(&w1, reinterpret_cast<const char*>(&v) + 16, 16);
    __builtin_nontemporal_store(w0, reinterpret_cast<nca_u32x4*>(p));
    __builtin_nontemporal_store(w1, reinterpret_cast<nca_u32x4*>(p) + 1);
  } else if constexpr (B == 8) {
    nca_u32x2 w;
    __builtin_memcpy(&w, &v, 8);
    __builtin_nontemporal_store(w, reinterpret_cast<nca_u32x2*>(p));
  } else if constexpr (B == 4) {
    unsigned w;
    __builtin_memcpy(&w, &v, 4);
    __builtin_nontemporal_store(w, reinterpret_cast<unsigned*>(p));
  } else {
    *p = v;
  }
}

// Streaming (non-temporal) load of one value of any 1-, 2-, 4-, 8- or 16-byte type (storage
// types such as nca_bf16 included), for data this kernel reads once.
template <typename T>
__device__ __forceinline__ T nca_ldnt(const T* p) {
  T v;
  if constexpr (sizeof(T) == 16) {
    const nca_u32x4 w = __builtin_nontemporal_load(reinterpret_cast<const nca_u32x4*>(p));
    __builtin_memcpy(&v, &w, 16);
  } else if constexpr (sizeof(T) == 8) {
    const unsigned long long w = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(p));
    __builtin_memcpy(&v, &w, 8);
  } else if constexpr (sizeof(T) == 4) {
    const unsigned w = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(p));
    __builtin_memcpy(&v, &w, 4);
  } else if constexpr (sizeof(T) == 2) {
    const unsigned short w = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p));
    __builtin_memcpy(&v, &w, 2);
  } else if constexpr (sizeof(T) == 1) {
    const unsigned char w = __builtin_nontemporal_load(reinterpret_cast<const unsigned char*>(p));
    __builtin_memcpy(&v, &w, 1);
  } else {
    v = *p;
  }
  return v;
}

// Streaming (non-temporal) load of a whole vector: for data this kernel reads once.
template <typename T, int N>
__device__ __forceinline__ nca_vec<T, N> nca_load_nt(const nca_vec<T, N>* p) {
  constexpr int B = static_cast<int>(sizeof(T)) * N;
  nca_vec<T, N> v;
  if constexpr (B == 16) {
    const nca_u32x4 w = __builtin_nontemporal_load(reinterpret_cast<const nca_u32x4*>(p));
    __builtin_memcpy(&v, &w, 16);
  } else if constexpr (B == 32) {
    const nca_u32x4 w0 = __builtin_nontemporal_load(reinterpret_cast<const nca_u32x4*>(p));
    const nca_u32x4 w1 = __builtin_nontemporal_load(reinterpret_cast<const nca_u32x4*>(p) + 1);
    __builtin_memcpy(&v, &w0, 16);
    __builtin_memcpy(reinterpret_cast<char*>(&v) + 16, &w1, 16);
  } else if constexpr (B == 8) {
    const nca_u32x2 w = __builtin_nontemporal_load(reinterpret_cast<const nca_u32x2*>(p));
    __builtin_memcpy(&v, &w, 8);
  } else {
    v = *p;
  }
  return v;
}

#endif  // NCA_DEVICE_HH
