// Host-side internals of libnca_hip: error reporting, per-device state, scratch
// workspaces, view analysis and the hipRTC kernel cache.
#ifndef NCA_INTERNAL_HH
#define NCA_INTERNAL_HH

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "nca_hip.h"
#include "ncarray_hip/dtype.hh"

namespace nca {

// ---- errors ------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string& msg);

#define NCA_THROW(code, msg) throw ::nca::Error((code), (msg))
#define NCA_REQUIRE(cond, msg) \
  do { if (!(cond)) NCA_THROW(NCA_ERR_INVALID, (msg)); } while (0)
#define HIP_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t nca_e_ = (expr);                                                      \
    if (nca_e_ != hipSuccess)                                                        \
      NCA_THROW(nca_e_ == hipErrorOutOfMemory ? NCA_ERR_OOM : NCA_ERR_HIP,           \
                std::string(#expr) + " failed: " + hipGetErrorString(nca_e_));       \
  } while (0)

// Wrap an API body: converts exceptions to status codes + thread-local message.
template <typename F>
int api_call(F&& f) {
  try {
    f();
    return NCA_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return NCA_ERR_OOM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return NCA_ERR_INVALID;
  }
}

// ---- device state --------------------------------------------------------------------
// Scratch memory for one (device, stream) pair.  Kernels of one call run in stream
// order, so one workspace per stream is race-free; calls that overlap on different
// streams get different workspaces.
constexpr int kWorkspaceCounters = 1 << 16;

struct Workspace {
  void* partials = nullptr;      // per-block partial results
  size_t partials_bytes = 0;
  unsigned* counters = nullptr;  // zeroed once; kernels re-arm them
  int n_counters = 0;
  void* big = nullptr;           // column partials / binning accumulators (grows)
  size_t big_bytes = 0;
};

struct DeviceState {
  int id = 0;
  int n_cus = 256;
  int lds_per_block = 65536;
  int max_threads = 1024;
  int64_t total_mem = 0;
  std::string arch = "gfx950";
  std::mutex mu;
  std::map<hipStream_t, std::unique_ptr<Workspace>> workspaces;
  // the device's default stream-ordered memory pool (nca_hip_malloc / nca_hip_free); null when
  // the runtime has none, then they fall back to hipMalloc / hipFree
  hipMemPool_t pool = nullptr;
  // this device's copy of the library's fixed scratch array (the entry points that allocate
  // nothing), looked up once when the device is initialised
  void* fixed_scratch = nullptr;
};

// The fixed scratch array of the allocation-free entry points: kFixedScratchBytes per device,
// the first kFixedPartialBytes for per-block partials, the rest for column partials.
constexpr size_t kFixedScratchBytes = size_t(32) << 20;
constexpr size_t kFixedPartialBytes = size_t(4) << 20;

DeviceState& current_device();           // throws NCA_ERR_NOT_INIT when not initialised
bool device_available();                  // a device was initialised on this thread
Workspace& workspace(hipStream_t s, size_t partial_bytes, size_t big_bytes = 0);
void finish_call(void* stream);           // synchronise when the C-ABI stream argument is NULL
void check_launch(const char* what);

// The HIP stream of a C-ABI stream argument (NCA_STREAM_DEFAULT_ASYNC is the null stream).
inline hipStream_t as_stream(void* s) {
  return s == NCA_STREAM_DEFAULT_ASYNC ? nullptr : reinterpret_cast<hipStream_t>(s);
}

// ---- views ---------------------------------------------------------------------------
struct ViewDesc {
  const nca_view* v;
  int64_t size() const {
    int64_t n = 1;
    for (int d = 0; d < v->ndim; ++d) n *= v->shape[d];
    return n;
  }
};
int64_t view_size(const nca_view& v);
void validate_view(const nca_view& v, const char* what, bool allow_null_data = false);
// As validate_view, but also accepts indirect views (expression operands and outputs).
void validate_view_nc(const nca_view& v, const char* what, bool allow_null_data = false);
// True when reaching an element reads a stored pointer: an NC pointer axis, or an SO axis
// whose suboffset is >= 0.
bool view_indirect(const nca_view& v);
// The dereference chain of a view: start address, byte strides, indirect axes and the byte
// offset added after each pointer read.  An NC pointer axis P (NCOffsetsPolicy::advance,
// layout.hh:468-480) is one step: the start moves by ptr_offset table entries, strides[P]
// (table entries) becomes bytes and row_offset is the offset after the read.  An SO view
// (SOArrayPolicy::advance, layout.hh:677-700) has a step at each axis with suboffset >= 0.
struct Chain {
  uintptr_t base = 0;
  std::array<int64_t, NCA_MAX_NDIM> strides{};
  std::vector<int> axes;
  std::vector<int64_t> subs;
};
Chain chain_of(const nca_view& v);
bool is_c_contiguous(const nca_view& v);
bool same_shape(const nca_view& a, const nca_view& b);
// True when every stride is zero (a broadcast scalar).
bool is_broadcast_scalar(const nca_view& v);
// Shape/strides after merging adjacent axes that are contiguous in memory with each
// other (C order); size-1 axes are dropped.  Returns ndim.
int coalesce(const int64_t* shape, const int64_t* strides, int ndim, int64_t* oshape,
             int64_t* ostrides);

// ---- JIT -----------------------------------------------------------------------------
struct JitKernel {
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
};
// Compile `source` (prelude prepended) with hipRTC for the current device's
// architecture, load it, and return kernel `name`.  Cached in memory and on disk.
hipFunction_t jit_get(const std::string& source, const std::string& name);
const std::string& device_prelude();
std::string sha256_hex(const std::string& data);
std::string hiprtc_options_string();

// Launch a generated kernel whose parameters are given as pointers to values.
void jit_launch(hipFunction_t fn, dim3 grid, dim3 block, size_t shmem, hipStream_t s,
                std::vector<void*>& params);

// ---- profiling -------------------------------------------------------------------------
// When enabled (nca_hip_profile_enable), each library kernel launch is bracketed by a pair
// of HIP events on its stream; nca_hip_profile_read reports the durations.
struct ProfScope {
  ProfScope(const char* name, hipStream_t s);
  ~ProfScope();
  const char* name;
  hipStream_t stream;
  void* start = nullptr;
};

// ---- AOT kernels (reduce.hip) used by the JIT paths --------------------------------------
// One launch that finishes fused reductions: column sums from float64 z-run partials
// (colpart[nzc][K] -> colout, when K > 0) and the global sum/max/min from per-block
// partials (8-byte slots: float64 sums, compute-type extrema).
struct FinalizeArgs {
  const double* colpart = nullptr;
  int nzc = 0;
  int64_t K = 0;
  void* colout = nullptr;
  int32_t col_dtype = NCA_FLOAT64;
  int32_t value_dtype = NCA_FLOAT32;
  int fuse = 0;
  const void* psum = nullptr;
  const void* pmax = nullptr;
  const void* pmin = nullptr;
  int nblocks = 0;
  void* sum_out = nullptr;
  void* max_out = nullptr;
  void* min_out = nullptr;
};
void launch_finalize_fused(const FinalizeArgs& a, hipStream_t s);
// Combine `nblocks` per-block partials of a fused scalar reduction into *out.
// kind: NCA_FUSE_SUM / MAX / MIN; dtype: dtype of the reduced values (sum partials
// are float64 for floating values, the integer sum type otherwise).
void launch_finalize_scalar(int kind, int32_t value_dtype, const void* partials,
                            int nblocks, void* out, hipStream_t s);
// out[k] = sum_{j<S} partials[j*K + k] (float64 partials) cast to out dtype.
void launch_finalize_columns(const double* partials, int S, int64_t K, void* out,
                             int32_t out_dtype, hipStream_t s);

}  // namespace nca

#endif  // NCA_INTERNAL_HH
