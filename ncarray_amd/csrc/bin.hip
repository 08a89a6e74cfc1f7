// Binning by an index map: the reference's scatter_add (kernels.cuh:660-680,
// host/elementwise.hh:482-505) -- dest[idx[i]] += src[i] for 0 <= idx[i] < dest.size,
// other indices skipped -- rebuilt as per-frame LDS-privatised histograms.
//
// nca_hip_bin: grid (chunks, frame groups).  A block zeroes LDS tables (float64 sum and
// zero-pixel count per bin and frame), streams its pixel chunk with 16-byte loads of pixels
// (non-temporal) and bins, and adds every pixel to its bin's sum with one LDS atomic.  The
// valid pixels of each bin are counted from the map alone, once per call, spread over the
// blocks of the first frame groups.  Blocks write their tables to partial slots; a combine
// kernel adds the chunk partials of each (frame, bin) in chunk order.  Counts are exact.  The
// float64 LDS atomics of a block run in hardware order, so a float64 sum can differ in its last
// bits from run to run (exact for integer-valued pixels below 2^53); a float32 sum is that
// float64 value rounded once, within one ulp of the FP64 sum and equal between runs unless the
// float64 values of two runs straddle a float32 rounding boundary.
//
// nca_hip_bin_backward: grad_frames[f, p] = grad_sums[f, map[p]] (0 out of range): the
// gradient row of each frame is staged in LDS and gathered.
#include <cstdlib>
#include <type_traits>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nca_device.hh"
#include "nca_internal.hh"

namespace nca {

// Counting modes of the histogram kernel.
enum { BIN_CNT_NONE = 0, BIN_CNT_ALL = 1, BIN_CNT_NONZERO = 2 };

// Where frame f starts: dense frames at base + f * fstride; frames reached through a table of
// pointers (an NC pointer axis 0 or an SO view indirect at axis 0, each frame's pixels
// contiguous) at the pointer stored at base + f * fstride, plus sub bytes.
struct FrameAddr {
  const char* base;
  nca_i64 fstride;
  nca_i64 sub;
  int ind;
};
template <typename T>
__device__ __forceinline__ T* frame_ptr(const FrameAddr& a, nca_i64 f) {
  const char* p = a.base + f * a.fstride;
  if (a.ind) p = nca_chase(p, a.sub);
  return reinterpret_cast<T*>(const_cast<char*>(p));
}

// One block: a chunk of pixels of FB consecutive frames (the map is read once for all FB).
// Every pixel makes one float64 LDS atomic and no branch: a nonzero pixel adds its value to
// its bin's sum, a zero pixel adds 1.0 to its bin's zero count (a float64 table: exact up to
// 2^53; leaving +-0 out of a sum that starts at +0 does not change it), and out-of-range bins
// go to a dummy slot nb.  Counting is split so that no per-pixel count atomic is needed: zero
// pixels are counted per frame as above, and the pixels of each valid bin are counted once per call
// from the map alone -- the blocks of the first S frame groups each count one S-th of their
// chunk (S up to 256, so the extra work is spread evenly); k_bin_valid_total adds those rows
// and the combine kernel forms count_all = valid and count_nonzero = valid - zeros.
// A radial detector map changes bin about once per pixel along a row, so folding runs of
// equal bins does not pay; the atomics run at LDS rate while the frame loads (non-temporal,
// read once) stay in flight.
template <typename T, typename I, int FB, bool SUMS, int CNT>
__global__ __launch_bounds__(1024) void k_bin(FrameAddr fa, const I* __restrict__ map, nca_i64 P,
                                              int nb, nca_i64 chunk, nca_i64 F, int S, nca_i64 slice, int R,
                                              double* __restrict__ psum, unsigned* __restrict__ pzero,
                                              unsigned* __restrict__ pvalid) {
  extern __shared__ unsigned char lds_raw[];
  const int nb1 = nb + 1;
  // zero counts share the sums' float64 atomic when both are wanted (ZD), else uint32
  constexpr bool ZD = SUMS && CNT == BIN_CNT_NONZERO;
  constexpr size_t ZB = ZD ? sizeof(double) : sizeof(unsigned);
  // R replicas of the frame histograms; thread t adds into replica t % R (neighbouring pixels
  // of a radial map often share a bin, and their lanes then no longer meet on one address)
  double* hs = reinterpret_cast<double*>(lds_raw);                                       // [R][FB][nb1] sums
  unsigned char* zraw = lds_raw + (SUMS ? sizeof(double) * R * FB * nb1 : 0);
  double* hzd = reinterpret_cast<double*>(zraw);                                         // [FB][nb1] zeros (ZD)
  unsigned* hz = reinterpret_cast<unsigned*>(zraw);                                      // [FB][nb1] zeros (!ZD)
  unsigned* hv = reinterpret_cast<unsigned*>(zraw + (CNT == BIN_CNT_NONZERO ? ZB * R * FB * nb1 : 0));  // [nb1] valid
  const bool count_valid = CNT != BIN_CNT_NONE && static_cast<int>(blockIdx.y) < S;
  for (int b = threadIdx.x; b < R * FB * nb1; b += blockDim.x) {
    if (SUMS) hs[b] = 0.0;
    if (ZD) hzd[b] = 0.0;
    else if (CNT == BIN_CNT_NONZERO) hz[b] = 0u;
  }
  if (count_valid)
    for (int b = threadIdx.x; b < nb1; b += blockDim.x) hv[b] = 0u;
  __syncthreads();
  const nca_i64 f0 = (nca_i64)blockIdx.y * FB;
  const int nf = (F - f0 < FB) ? static_cast<int>(F - f0) : FB;
  const int c = blockIdx.x;
  const nca_i64 p0 = (nca_i64)c * chunk;
  const nca_i64 p1 = (p0 + chunk < P) ? p0 + chunk : P;
  auto slot = [&](I m) { return (m < 0 || m >= nb) ? nb : static_cast<int>(m); };
  const int rbase = (static_cast<int>(threadIdx.x) % R) * FB * nb1;
  auto add = [&](int fb, int b, double v) {
    b += rbase;
    if (ZD) {
      const bool z = v == 0.0;
      atomicAdd(z ? &hzd[fb * nb1 + b] : &hs[fb * nb1 + b], z ? 1.0 : v);
    } else {
      if (SUMS) atomicAdd(&hs[fb * nb1 + b], v);
      if (CNT == BIN_CNT_NONZERO && v == 0.0) atomicAdd(&hz[fb * nb1 + b], 1u);
    }
  };
  if (SUMS || CNT == BIN_CNT_NONZERO) {
    // the block's frames; a frame reached through a pointer that is not vector-aligned sends the
    // block's pixels through the element loop
    const T* fp[FB];
    bool aligned = true;
#pragma unroll
    for (int fb = 0; fb < FB; ++fb) {
      fp[fb] = frame_ptr<const T>(fa, f0 + (fb < nf ? fb : 0));
      aligned = aligned && (reinterpret_cast<unsigned long long>(fp[fb]) % (4 * sizeof(T))) == 0;
    }
    // vector body: 4 pixels per vector (host guarantees p0 % 4 == 0 and aligned dense frames); U
    // vectors per frame are loaded before any is folded in, so U * FB loads stay in flight
    constexpr int U = (FB >= 4) ? 1 : (4 / FB);
    const nca_i64 nvec = aligned ? (p1 - p0) / 4 : 0;
    for (nca_i64 v0 = threadIdx.x; v0 < nvec; v0 += (nca_i64)U * blockDim.x) {
      nca_vec<I, 4> m[U];
      nca_vec<T, 4> x[U][FB];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const nca_i64 v = v0 + (nca_i64)u * blockDim.x;
        const nca_i64 p = p0 + 4 * (v < nvec ? v : 0);
        m[u] = *reinterpret_cast<const nca_vec<I, 4>*>(map + p);
#pragma unroll
        for (int fb = 0; fb < FB; ++fb)
          x[u][fb] = nca_load_nt(reinterpret_cast<const nca_vec<T, 4>*>(fp[fb] + p));
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool in = v0 + (nca_i64)u * blockDim.x < nvec;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int b = in ? slot(m[u].v[j]) : nb;
#pragma unroll
          for (int fb = 0; fb < FB; ++fb)
            if (fb < nf) add(fb, b, nca_cvt<double, T>::from(x[u][fb].v[j]));
        }
      }
    }
    for (nca_i64 p = p0 + 4 * nvec + threadIdx.x; p < p1; p += blockDim.x) {
      const int b = slot(map[p]);
      for (int fb = 0; fb < nf; ++fb) add(fb, b, nca_cvt<double, T>::from(fp[fb][p]));
    }
  }
  if (count_valid) {
    // valid pixels per bin of sub-slice blockIdx.y of this chunk (slice % 4 == 0)
    const nca_i64 q0 = p0 + (nca_i64)blockIdx.y * slice;
    const nca_i64 q1 = (q0 + slice < p1) ? q0 + slice : p1;
    const nca_i64 nq = (q1 > q0) ? (q1 - q0) / 4 : 0;
    for (nca_i64 v = threadIdx.x; v < nq; v += blockDim.x) {
      const nca_vec<I, 4> m = *reinterpret_cast<const nca_vec<I, 4>*>(map + q0 + 4 * v);
#pragma unroll
      for (int j = 0; j < 4; ++j) atomicAdd(&hv[slot(m.v[j])], 1u);
    }
    for (nca_i64 p = q0 + 4 * nq + threadIdx.x; p < q1; p += blockDim.x) atomicAdd(&hv[slot(map[p])], 1u);
  }
  __syncthreads();
  // partial slots of (frame, chunk) and of (chunk, sub-slice)
  for (int fb = 0; fb < nf; ++fb) {
    const nca_i64 base = ((f0 + fb) * gridDim.x + c) * (nca_i64)nb;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      // the replicas in a fixed order
      double ss = 0.0, zz = 0.0;
      unsigned zu = 0u;
      for (int r = 0; r < R; ++r) {
        const int q = (r * FB + fb) * nb1 + b;
        if (SUMS) ss += hs[q];
        if (ZD) zz += hzd[q];
        else if (CNT == BIN_CNT_NONZERO) zu += hz[q];
      }
      if (SUMS) psum[base + b] = ss;
      if (ZD) pzero[base + b] = static_cast<unsigned>(zz);
      else if (CNT == BIN_CNT_NONZERO) pzero[base + b] = zu;
    }
  }
  if (count_valid) {
    const nca_i64 base = ((nca_i64)c * S + blockIdx.y) * nb;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) pvalid[base + b] = hv[b];
  }
}

// valid[b] = sum over the `rows` valid-count rows (chunk x sub-slice) of the histogram blocks.
// Block: 16 bins x 16 row lanes; a wave reads 4 rows of 16 consecutive bins.
__global__ __launch_bounds__(256) void k_bin_valid_total(const unsigned* __restrict__ pvalid, int rows, int nb,
                                                         unsigned long long* __restrict__ valid) {
  __shared__ unsigned long long part[16][17];
  const int j = threadIdx.x & 15, lane = threadIdx.x >> 4;
  const int b = blockIdx.x * 16 + j;
  unsigned long long t = 0;
  if (b < nb) {
    // 8 independent loads in flight per thread
    int r = lane;
    for (; r + 7 * 16 < rows; r += 8 * 16) {
      unsigned q[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) q[u] = pvalid[(nca_i64)(r + 16 * u) * nb + b];
#pragma unroll
      for (int u = 0; u < 8; ++u) t += q[u];
    }
    for (; r < rows; r += 16) t += pvalid[(nca_i64)r * nb + b];
  }
  part[lane][j] = t;
  __syncthreads();
  if (lane == 0 && b < nb) {
    unsigned long long s = 0;
    for (int q = 0; q < 16; ++q) s += part[q][j];
    valid[b] = s;
  }
}

// counts[f, b] = valid[b] (- zeros[f, b]); sums[f, b] = sum of the chunk partials in chunk order
// (deterministic whatever the block schedule).
template <typename CO, typename SO>
__global__ __launch_bounds__(256) void k_bin_combine(const double* __restrict__ psum, const unsigned* __restrict__ pzero,
                                                     const unsigned long long* __restrict__ valid, int chunks, int nb,
                                                     nca_i64 F, CO* __restrict__ counts, nca_i64 cst0, nca_i64 cst1,
                                                     SO* __restrict__ sums, nca_i64 sst0, nca_i64 sst1) {
  const nca_i64 k = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= F * nb) return;
  const nca_i64 f = k / nb;
  const int b = static_cast<int>(k - f * nb);
  if (counts) {
    long long cnt = valid ? static_cast<long long>(valid[b]) : 0;
    if (pzero)
      for (int c = 0; c < chunks; ++c) cnt -= pzero[(f * chunks + c) * (nca_i64)nb + b];
    counts[f * cst0 + b * cst1] = static_cast<CO>(cnt);
  }
  if (sums) {
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += psum[(f * chunks + c) * (nca_i64)nb + b];
    nca_st(&sums[f * sst0 + b * sst1], s);
  }
}

template <typename G, typename I, typename O, int FB>
__global__ __launch_bounds__(256) void k_bin_backward(const G* __restrict__ g, nca_i64 gst0, nca_i64 gst1,
                                                      const I* __restrict__ map, nca_i64 P, int nb,
                                                      FrameAddr oa, nca_i64 chunk, nca_i64 F) {
  // one block: a chunk of pixels of FB consecutive frames; each map vector serves FB rows
  extern __shared__ unsigned char lds_raw[];
  float* gl = reinterpret_cast<float*>(lds_raw);  // [FB][nb]
  const nca_i64 f0 = (nca_i64)blockIdx.y * FB;
  const int nf = (F - f0 < FB) ? static_cast<int>(F - f0) : FB;
  for (int k = threadIdx.x; k < nf * nb; k += blockDim.x) {
    const int fb = k / nb, b = k - fb * nb;
    gl[k] = nca_ld<float>(g + (f0 + fb) * gst0 + b * gst1);
  }
  __syncthreads();
  const nca_i64 p0 = (nca_i64)blockIdx.x * chunk;
  const nca_i64 p1 = (p0 + chunk < P) ? p0 + chunk : P;
  O* op[FB];
  bool aligned = true;
#pragma unroll
  for (int fb = 0; fb < FB; ++fb) {
    op[fb] = frame_ptr<O>(oa, f0 + (fb < nf ? fb : 0));
    aligned = aligned && (reinterpret_cast<unsigned long long>(op[fb]) % (4 * sizeof(O))) == 0;
  }
  const nca_i64 nvec = aligned ? (p1 - p0) / 4 : 0;
  for (nca_i64 v = threadIdx.x; v < nvec; v += blockDim.x) {
    const nca_i64 p = p0 + 4 * v;
    const nca_vec<I, 4> m = *reinterpret_cast<const nca_vec<I, 4>*>(map + p);
    int bins[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) bins[j] = (m.v[j] < 0 || m.v[j] >= nb) ? -1 : static_cast<int>(m.v[j]);
#pragma unroll
    for (int fb = 0; fb < FB; ++fb) {
      if (fb >= nf) break;
      nca_vec<O, 4> r;
#pragma unroll
      for (int j = 0; j < 4; ++j) r.v[j] = nca_store_cvt<O, float>::to(bins[j] >= 0 ? gl[fb * nb + bins[j]] : 0.0f);
      nca_store_nt(reinterpret_cast<nca_vec<O, 4>*>(op[fb] + p), r);
    }
  }
  for (nca_i64 p = p0 + 4 * nvec + threadIdx.x; p < p1; p += blockDim.x) {
    const I bi = map[p];
    for (int fb = 0; fb < nf; ++fb)
      nca_st(op[fb] + p, (bi >= 0 && bi < nb) ? gl[fb * nb + static_cast<int>(bi)] : 0.0f);
  }
}

// scatter_add with the reference's semantics (scatter_add_kernel, kernels.cuh:660-680): every
// element i adds src[i] to dest[idx[i]] with one global atomic when 0 <= idx[i] < dest.size.
// Contiguous operands.
template <typename D, typename I, typename S>
__global__ __launch_bounds__(256) void k_scatter_add(D* __restrict__ dest, nca_i64 nd, const I* __restrict__ idx,
                                                     const S* __restrict__ src, nca_i64 n) {
  const nca_i64 gs = (nca_i64)gridDim.x * blockDim.x;
  for (nca_i64 i = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gs) {
    const I k = idx[i];
    if (k >= 0 && static_cast<nca_i64>(k) < nd) atomicAdd(&dest[k], nca_cvt<D, S>::from(src[i]));
  }
}

// The same over any layouts: idx and src by C-order linear index i, dest by the linear index
// idx[i] (static_index, array_impl.hh:2101-2179), each through its strides and, for indirect
// views, its pointer chain.
template <typename D, typename I, typename S>
__global__ __launch_bounds__(256) void k_scatter_add_nd(nca_ndview dest, nca_i64 nd, nca_ndview idx, nca_ndview src,
                                                        nca_i64 n) {
  const nca_i64 gs = (nca_i64)gridDim.x * blockDim.x;
  for (nca_i64 i = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gs) {
    const I k = *reinterpret_cast<const I*>(nca_nd_addr(idx, i));
    if (k >= 0 && static_cast<nca_i64>(k) < nd) {
      D* dp = reinterpret_cast<D*>(const_cast<char*>(nca_nd_addr(dest, static_cast<nca_i64>(k))));
      atomicAdd(dp, nca_cvt<D, S>::from(*reinterpret_cast<const S*>(nca_nd_addr(src, i))));
    }
  }
}

// ---- host side ---------------------------------------------------------------------------------------------
static nca_ndview ndview_of(const nca_view& v) {
  nca_ndview r{};
  const Chain ch = chain_of(v);
  r.base = reinterpret_cast<const char*>(ch.base);
  r.ndim = v.ndim;
  for (int d = 0; d < v.ndim; ++d) { r.shape[d] = v.shape[d]; r.strides[d] = ch.strides[d]; r.subs[d] = -1; }
  for (size_t i = 0; i < ch.axes.size(); ++i) r.subs[ch.axes[i]] = ch.subs[i];
  r.ind = ch.axes.empty() ? 0 : 1;
  if (r.ndim == 0) { r.ndim = 1; r.shape[0] = 1; r.strides[0] = 0; r.subs[0] = -1; }
  return r;
}

// Frames per block of the binning kernels: the largest power of two <= cap whose LDS tables fit
// in 64 KiB (and the device's per-block limit), at most the number of frames.  Measured on
// MI355X (256 frames of 1024^2, 1000 bins): the gather backward runs 1.4x faster with 4 frames
// per block (the bin map is read once per 4 output rows); the histogram forward is fastest with
// 1 (more frames per block means more LDS per block and more partial histograms to combine).
// NCA_HIP_BIN_FRAMES (1, 2, 4 or 8, at most max_cap) overrides the default, for tuning.
static int frames_per_block(size_t per_frame, size_t lds_limit, int64_t F, int cap, int max_cap) {
  if (const char* e = std::getenv("NCA_HIP_BIN_FRAMES")) {
    const int v = std::atoi(e);
    if (v == 1 || v == 2 || v == 4 || v == 8) cap = std::min(max_cap, v);
  }
  const size_t limit = std::min<size_t>(lds_limit, 65536);
  int fb = 1;
  while (fb * 2 <= cap && per_frame * fb * 2 <= limit && fb * 2 <= F) fb *= 2;
  return fb;
}
template <typename F>
static void dispatch_frame_type(int32_t dt, F&& f) {
  switch (dt) {
    case NCA_FLOAT32: f((float*)nullptr); break;
    case NCA_FLOAT64: f((double*)nullptr); break;
    case NCA_BFLOAT16: f((nca_bf16*)nullptr); break;
    case NCA_FLOAT16: f((nca_f16*)nullptr); break;
    case NCA_INT32: f((int*)nullptr); break;
    case NCA_UINT16: f((unsigned short*)nullptr); break;
    case NCA_UINT32: f((unsigned int*)nullptr); break;
    default: NCA_THROW(NCA_ERR_UNSUPPORTED, "binning: unsupported frame dtype");
  }
}

template <typename F>
static void dispatch_map_type(int32_t dt, F&& f) {
  switch (dt) {
    case NCA_INT32: f((int*)nullptr); break;
    case NCA_INT64: f((long long*)nullptr); break;
    default: NCA_THROW(NCA_ERR_UNSUPPORTED, "binning: the bin map must be int32 or int64");
  }
}

// Frames [F, pixels...]: C-contiguous, or reached through a table of frame pointers (an NC
// pointer axis 0, or an SO view whose only indirect axis is 0) with each frame's pixels
// contiguous.  Sets F and the frame addressing; returns the pixels per frame.
static int64_t frames_layout(const nca_view& frames, const nca_view& map, int64_t& F, FrameAddr& fa) {
  const int64_t P = view_size(map);
  NCA_REQUIRE(frames.ndim >= 1, "frames must have a frame axis");
  F = frames.shape[0];
  int64_t pf = 1;
  for (int d = 1; d < frames.ndim; ++d) pf *= frames.shape[d];
  if (frames.ndim == 1) { F = 1; pf = frames.shape[0]; }
  NCA_REQUIRE(pf == P, "each frame must have as many pixels as the bin map");
  NCA_REQUIRE(is_c_contiguous(map), "the bin map must be C-contiguous");
  const int64_t item = itemsize(frames.dtype);
  if (view_indirect(frames)) {
    const Chain ch = chain_of(frames);
    NCA_REQUIRE(frames.ndim >= 2 && ch.axes.size() == 1 && ch.axes[0] == 0,
                "frames reached through pointers must have their pointers on the frame axis only");
    int64_t expect = item;
    for (int d = frames.ndim - 1; d >= 1; --d) {
      NCA_REQUIRE(frames.shape[d] == 1 || ch.strides[d] == expect, "the pixels of each frame must be contiguous");
      expect *= frames.shape[d];
    }
    fa = FrameAddr{reinterpret_cast<const char*>(ch.base), ch.strides[0], ch.subs[0], 1};
  } else {
    NCA_REQUIRE(is_c_contiguous(frames), "frames must be C-contiguous (or a table of contiguous frames)");
    fa = FrameAddr{static_cast<const char*>(frames.data), P * item, 0, 0};
  }
  return P;
}

static void out2d(const nca_view* v, int64_t F, int nb, const char* what, int64_t& st0, int64_t& st1) {
  if (!v) return;
  validate_view(*v, what);
  int64_t item = itemsize(v->dtype);
  if (v->ndim == 2) {
    NCA_REQUIRE(v->shape[0] == F && v->shape[1] == nb, std::string(what) + " must be [frames, n_bins]");
    st0 = v->strides[0] / item;
    st1 = v->strides[1] / item;
    NCA_REQUIRE(v->strides[0] % item == 0 && v->strides[1] % item == 0, std::string(what) + " strides not element-aligned");
  } else {
    NCA_REQUIRE(v->ndim == 1 && F == 1 && v->shape[0] == nb, std::string(what) + " must be [frames, n_bins]");
    st0 = 0;
    st1 = v->strides[0] / item;
  }
}

}  // namespace nca

using namespace nca;

extern "C" {

int nca_hip_bin(const nca_view* frames, const nca_view* binmap, int32_t n_bins, int32_t count_pred,
                const nca_view* counts, const nca_view* sums, void* stream) {
  return api_call([&] {
    NCA_REQUIRE(frames && binmap, "null frames/binmap");
    NCA_REQUIRE(counts || sums, "nothing to compute");
    NCA_REQUIRE(n_bins > 0, "n_bins must be positive");
    NCA_REQUIRE(count_pred == NCA_COUNT_ALL || count_pred == NCA_COUNT_NONZERO, "bad count predicate");
    validate_view_nc(*frames, "frames");
    validate_view(*binmap, "binmap");
    int64_t F = 0;
    FrameAddr fa{};
    const int64_t P = frames_layout(*frames, *binmap, F, fa);
    int64_t cst0 = 0, cst1 = 0, sst0 = 0, sst1 = 0;
    out2d(counts, F, n_bins, "counts", cst0, cst1);
    out2d(sums, F, n_bins, "sums", sst0, sst1);
    if (counts) NCA_REQUIRE(counts->dtype == NCA_INT32 || counts->dtype == NCA_INT64, "counts must be int32 or int64");
    if (sums) NCA_REQUIRE(sums->dtype == NCA_FLOAT32 || sums->dtype == NCA_FLOAT64, "sums must be float32 or float64");
    hipStream_t s = as_stream(stream);
    DeviceState& dev = current_device();
    const int cnt_mode = !counts ? BIN_CNT_NONE : (count_pred == NCA_COUNT_NONZERO ? BIN_CNT_NONZERO : BIN_CNT_ALL);
    const bool do_sums = sums != nullptr;
    const size_t nb1 = static_cast<size_t>(n_bins) + 1;
    // LDS per frame of a block (sums, zero counts) and per block (valid-pixel counts)
    const size_t zero_bytes = cnt_mode == BIN_CNT_NONZERO ? (do_sums ? sizeof(double) : sizeof(unsigned)) : 0;
    const size_t per_frame = ((do_sums ? sizeof(double) : 0) + zero_bytes) * nb1;
    const size_t per_block = (cnt_mode != BIN_CNT_NONE ? sizeof(unsigned) : 0) * nb1;
    NCA_REQUIRE(per_frame + per_block <= static_cast<size_t>(dev.lds_per_block), "too many bins for an LDS histogram");
    const int fb = frames_per_block(std::max<size_t>(per_frame, 1), static_cast<size_t>(dev.lds_per_block) - per_block, F,
                                    1, 4);
    // two replicas of the frame histograms when they fit (measured 2.7% faster on radial maps)
    const size_t lds_cap = std::min<size_t>(static_cast<size_t>(dev.lds_per_block), 65536);
    const int R = per_frame * fb * 2 + per_block <= lds_cap ? 2 : 1;
    const size_t lds = per_frame * fb * R + per_block;
    const int64_t groups = (F + fb - 1) / fb;
    // chunks per frame group: about 2 blocks of 1024 threads per CU overall, chunk a multiple of 4 pixels
    int64_t chunks = std::max<int64_t>(1, (static_cast<int64_t>(dev.n_cus) * 2 + groups - 1) / std::max<int64_t>(groups, 1));
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, P / 4096));
    int64_t chunk = (P + chunks - 1) / chunks;
    chunk = (chunk + 3) / 4 * 4;
    chunks = (P + chunk - 1) / chunk;
    NCA_REQUIRE(F < 65536, "at most 65535 frames per call");
    // valid-pixel counts: the blocks of frame groups [0, S) each count one S-th of their chunk
    const int S = static_cast<int>(std::min<int64_t>(256, std::max<int64_t>(groups, 1)));
    const int64_t slice = ((chunk + S - 1) / S + 3) / 4 * 4;
    const size_t slots = static_cast<size_t>(F) * chunks * n_bins;
    const size_t vslots = static_cast<size_t>(chunks) * S * n_bins;
    Workspace& w = workspace(s, 1, slots * (sizeof(double) + sizeof(unsigned)) + vslots * sizeof(unsigned) +
                                       static_cast<size_t>(n_bins) * sizeof(unsigned long long) + 8);
    double* psum = static_cast<double*>(w.big);
    unsigned* pzero = reinterpret_cast<unsigned*>(psum + slots);
    unsigned* pvalid = pzero + slots;
    // 8-byte aligned after the 4-byte tables
    unsigned long long* valid =
        reinterpret_cast<unsigned long long*>((reinterpret_cast<uintptr_t>(pvalid + vslots) + 7) & ~uintptr_t(7));
    const bool work = P > 0 && F > 0;
    dispatch_frame_type(frames->dtype, [&](auto* tp) {
      typedef std::remove_pointer_t<decltype(tp)> T;
      dispatch_map_type(binmap->dtype, [&](auto* ip) {
        typedef std::remove_pointer_t<decltype(ip)> I;
        NCA_REQUIRE((fa.ind || reinterpret_cast<uintptr_t>(frames->data) % (4 * sizeof(T)) == 0) &&
                        reinterpret_cast<uintptr_t>(binmap->data) % (4 * sizeof(I)) == 0 &&
                        (P % 4 == 0 || F == 1 || fa.ind),
                    "binning needs 4-element aligned frames and map (and pixels per frame % 4 == 0)");
        if (!work) return;
        const dim3 grid(static_cast<unsigned>(chunks), static_cast<unsigned>(groups));
        auto go = [&](auto fbc, auto sc, auto cc) {
          constexpr int FBC = decltype(fbc)::value;
          constexpr bool SC = decltype(sc)::value;
          constexpr int CC = decltype(cc)::value;
          ProfScope ps("nca_bin", s);
          hipLaunchKernelGGL((k_bin<T, I, FBC, SC, CC>), grid, dim3(1024), lds, s, fa,
                             (const I*)binmap->data, (nca_i64)P, n_bins, (nca_i64)chunk, (nca_i64)F, S, (nca_i64)slice, R,
                             psum, pzero, pvalid);
        };
        auto with_cnt = [&](auto fbc, auto sc) {
          if (cnt_mode == BIN_CNT_NONZERO) go(fbc, sc, std::integral_constant<int, BIN_CNT_NONZERO>{});
          else if (cnt_mode == BIN_CNT_ALL) go(fbc, sc, std::integral_constant<int, BIN_CNT_ALL>{});
          else go(fbc, sc, std::integral_constant<int, BIN_CNT_NONE>{});
        };
        auto with_sums = [&](auto fbc) {
          if (do_sums) with_cnt(fbc, std::true_type{});
          else with_cnt(fbc, std::false_type{});
        };
        if (fb == 4) with_sums(std::integral_constant<int, 4>{});
        else if (fb == 2) with_sums(std::integral_constant<int, 2>{});
        else with_sums(std::integral_constant<int, 1>{});
      });
    });
    // empty frames or map: zero chunks and no valid rows, so the combine writes zero counts and sums
    const int nchunks = work ? static_cast<int>(chunks) : 0;
    const int rows = work ? static_cast<int>(chunks) * S : 0;
    if (cnt_mode != BIN_CNT_NONE) {
      ProfScope ps("nca_bin_valid_total", s);
      hipLaunchKernelGGL(k_bin_valid_total, dim3(static_cast<unsigned>((n_bins + 15) / 16)), dim3(256), 0, s, pvalid, rows,
                         n_bins, valid);
    }
    dim3 g2(static_cast<unsigned>((F * n_bins + 255) / 256));
    auto launch_combine = [&](auto* cp, auto* sp) {
      typedef std::remove_pointer_t<decltype(cp)> CO;
      typedef std::remove_pointer_t<decltype(sp)> SO;
      if (g2.x == 0) return;
      ProfScope ps("nca_bin_combine", s);
      hipLaunchKernelGGL((k_bin_combine<CO, SO>), g2, dim3(256), 0, s, psum, cnt_mode == BIN_CNT_NONZERO ? pzero : nullptr,
                         cnt_mode != BIN_CNT_NONE ? valid : nullptr, nchunks, n_bins, (nca_i64)F,
                         counts ? (CO*)counts->data : nullptr,
                         (nca_i64)cst0, (nca_i64)cst1, sums ? (SO*)sums->data : nullptr, (nca_i64)sst0, (nca_i64)sst1);
    };
    const bool c64 = counts && counts->dtype == NCA_INT64;
    const bool s64 = sums && sums->dtype == NCA_FLOAT64;
    if (c64 && s64) launch_combine((long long*)nullptr, (double*)nullptr);
    else if (c64) launch_combine((long long*)nullptr, (float*)nullptr);
    else if (s64) launch_combine((int*)nullptr, (double*)nullptr);
    else launch_combine((int*)nullptr, (float*)nullptr);
    check_launch("bin");
    finish_call(stream);
  });
}

int nca_hip_bin_backward(const nca_view* grad_sums, const nca_view* binmap, const nca_view* grad_frames, void* stream) {
  return api_call([&] {
    NCA_REQUIRE(grad_sums && binmap && grad_frames, "null argument");
    validate_view(*grad_sums, "grad_sums");
    validate_view(*binmap, "binmap");
    validate_view_nc(*grad_frames, "grad_frames");
    int64_t F = 0;
    FrameAddr oa{};
    const int64_t P = frames_layout(*grad_frames, *binmap, F, oa);
    NCA_REQUIRE(grad_sums->ndim >= 1, "grad_sums must be [frames, n_bins]");
    const int nb = static_cast<int>(grad_sums->shape[grad_sums->ndim - 1]);
    int64_t gst0 = 0, gst1 = 0;
    out2d(grad_sums, F, nb, "grad_sums", gst0, gst1);
    hipStream_t s = as_stream(stream);
    DeviceState& dev = current_device();
    const size_t per_frame = sizeof(float) * static_cast<size_t>(nb);
    NCA_REQUIRE(per_frame <= static_cast<size_t>(dev.lds_per_block), "too many bins");
    const int fb = frames_per_block(per_frame, static_cast<size_t>(dev.lds_per_block), F, 4, 8);
    const size_t lds = per_frame * fb;
    const int64_t groups = (F + fb - 1) / fb;
    int64_t chunks = std::max<int64_t>(1, (static_cast<int64_t>(dev.n_cus) * 8 + groups - 1) / std::max<int64_t>(groups, 1));
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, P / 4096));
    int64_t chunk = ((P + chunks - 1) / chunks + 3) / 4 * 4;
    chunks = (P + chunk - 1) / chunk;
    if (P == 0 || F == 0) return;
    NCA_REQUIRE(F < 65536, "at most 65535 frames per call");
    auto launch = [&](auto* gp, auto* ip, auto* op) {
      typedef std::remove_pointer_t<decltype(gp)> G;
      typedef std::remove_pointer_t<decltype(ip)> I;
      typedef std::remove_pointer_t<decltype(op)> O;
      NCA_REQUIRE((oa.ind || reinterpret_cast<uintptr_t>(grad_frames->data) % (4 * sizeof(O)) == 0) &&
                      reinterpret_cast<uintptr_t>(binmap->data) % (4 * sizeof(I)) == 0 &&
                      (P % 4 == 0 || F == 1 || oa.ind),
                  "bin backward needs 4-element aligned outputs and map");
      const dim3 grid(static_cast<unsigned>(chunks), static_cast<unsigned>(groups));
      auto go = [&](auto fbc) {
        constexpr int FBC = decltype(fbc)::value;
        ProfScope ps("nca_bin_backward", s);
        hipLaunchKernelGGL((k_bin_backward<G, I, O, FBC>), grid, dim3(256), lds, s, (const G*)grad_sums->data,
                           (nca_i64)gst0, (nca_i64)gst1, (const I*)binmap->data, (nca_i64)P, nb, oa,
                           (nca_i64)chunk, (nca_i64)F);
      };
      if (fb == 8) go(std::integral_constant<int, 8>{});
      else if (fb == 4) go(std::integral_constant<int, 4>{});
      else if (fb == 2) go(std::integral_constant<int, 2>{});
      else go(std::integral_constant<int, 1>{});
    };
    auto with_out = [&](auto* gp, auto* ip) {
      switch (grad_frames->dtype) {
        case NCA_FLOAT32: launch(gp, ip, (float*)nullptr); break;
        case NCA_FLOAT64: launch(gp, ip, (double*)nullptr); break;
        case NCA_BFLOAT16: launch(gp, ip, (nca_bf16*)nullptr); break;
        case NCA_FLOAT16: launch(gp, ip, (nca_f16*)nullptr); break;
        default: NCA_THROW(NCA_ERR_UNSUPPORTED, "bin backward: gradient dtype must be floating point");
      }
    };
    auto with_map = [&](auto* gp) {
      dispatch_map_type(binmap->dtype, [&](auto* ip) { with_out(gp, ip); });
    };
    switch (grad_sums->dtype) {
      case NCA_FLOAT32: with_map((float*)nullptr); break;
      case NCA_FLOAT64: with_map((double*)nullptr); break;
      default: NCA_THROW(NCA_ERR_UNSUPPORTED, "grad_sums must be float32 or float64");
    }
    check_launch("bin_backward");
    finish_call(stream);
  });
}

int nca_hip_scatter_add(const nca_view* dest, const nca_view* idx, const nca_view* src, void* stream) {
  return api_call([&] {
    NCA_REQUIRE(dest && idx && src, "null argument");
    validate_view_nc(*dest, "dest");
    validate_view_nc(*idx, "idx");
    validate_view_nc(*src, "src");
    NCA_REQUIRE(view_size(*idx) == view_size(*src), "idx and src must have the same size");
    const bool contig = dest->ndim == 1 && is_c_contiguous(*dest) && is_c_contiguous(*idx) && is_c_contiguous(*src);
    const int64_t n = view_size(*src), nd = view_size(*dest);
    hipStream_t s = as_stream(stream);
    DeviceState& dev = current_device();
    if (n == 0) return;
    int blocks = static_cast<int>(std::min<int64_t>((n + 255) / 256, dev.n_cus * 16));
    auto launch = [&](auto* dp, auto* ip, auto* sp) {
      typedef std::remove_pointer_t<decltype(dp)> D;
      typedef std::remove_pointer_t<decltype(ip)> I;
      typedef std::remove_pointer_t<decltype(sp)> S;
      ProfScope ps("nca_scatter_add", s);
      if (contig)
        hipLaunchKernelGGL((k_scatter_add<D, I, S>), dim3(blocks), dim3(256), 0, s, (D*)dest->data, (nca_i64)nd,
                           (const I*)idx->data, (const S*)src->data, (nca_i64)n);
      else
        hipLaunchKernelGGL((k_scatter_add_nd<D, I, S>), dim3(blocks), dim3(256), 0, s, ndview_of(*dest), (nca_i64)nd,
                           ndview_of(*idx), ndview_of(*src), (nca_i64)n);
    };
    auto with_src = [&](auto* dp, auto* ip) {
      switch (src->dtype) {
        case NCA_FLOAT32: launch(dp, ip, (float*)nullptr); break;
        case NCA_FLOAT64: launch(dp, ip, (double*)nullptr); break;
        case NCA_INT32: launch(dp, ip, (int*)nullptr); break;
        case NCA_UINT16: launch(dp, ip, (unsigned short*)nullptr); break;
        case NCA_BFLOAT16: launch(dp, ip, (nca_bf16*)nullptr); break;
        default: NCA_THROW(NCA_ERR_UNSUPPORTED, "scatter_add: unsupported src dtype");
      }
    };
    auto with_idx = [&](auto* dp) { dispatch_map_type(idx->dtype, [&](auto* ip) { with_src(dp, ip); }); };
    switch (dest->dtype) {
      case NCA_FLOAT32: with_idx((float*)nullptr); break;
      case NCA_FLOAT64: with_idx((double*)nullptr); break;
      case NCA_INT32: with_idx((int*)nullptr); break;
      case NCA_UINT32: with_idx((unsigned int*)nullptr); break;
      case NCA_UINT64: with_idx((unsigned long long*)nullptr); break;
      default: NCA_THROW(NCA_ERR_UNSUPPORTED, "scatter_add: unsupported dest dtype");
    }
    check_launch("scatter_add");
    finish_call(stream);
  });
}

}  // extern "C"
