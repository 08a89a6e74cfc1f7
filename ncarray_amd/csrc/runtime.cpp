// libnca_hip runtime: initialisation, errors, device memory, scratch workspaces and view
// helpers.  Replaces the CUDA context / DevOwnerPolicy / CircularDevicePool plumbing of
// the reference (jit_utils.cc:28-46, storage.hh:176-186, 242-252, mem_pool.cuh:47-123).
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "nca_internal.hh"

// The fixed scratch array of the allocation-free entry points (nca_hip_*_noalloc).  Every GPU
// the library runs on holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for that device (DeviceState::fixed_scratch).
static __device__ __align__(256) char scratch[nca::kFixedScratchBytes];

namespace nca {

namespace {
thread_local std::string t_last_error;
thread_local int t_device = -1;
std::mutex g_devices_mu;
std::map<int, std::unique_ptr<DeviceState>> g_devices;
}  // namespace

void set_last_error(const std::string& msg) { t_last_error = msg; }

bool device_available() { return t_device >= 0; }

DeviceState& current_device() {
  if (t_device < 0) {
    // A thread that did not call nca_hip_init (a worker thread of the caller, e.g. an
    // autograd engine thread) adopts its current HIP device once the library has been
    // initialised for that device by any thread.
    int d = -1;
    if (hipGetDevice(&d) == hipSuccess) {
      std::lock_guard<std::mutex> lk(g_devices_mu);
      if (g_devices.count(d)) t_device = d;
    }
    if (t_device < 0) NCA_THROW(NCA_ERR_NOT_INIT, "nca_hip_init() was not called for this thread's device");
  }
  std::lock_guard<std::mutex> lk(g_devices_mu);
  auto it = g_devices.find(t_device);
  if (it == g_devices.end()) NCA_THROW(NCA_ERR_NOT_INIT, "device state missing");
  return *it->second;
}

static void init_device(int device) {
  int count = 0;
  HIP_CHECK(hipGetDeviceCount(&count));
  NCA_REQUIRE(device >= 0 && device < count,
              "device " + std::to_string(device) + " out of range (" + std::to_string(count) + " devices)");
  HIP_CHECK(hipSetDevice(device));
  std::lock_guard<std::mutex> lk(g_devices_mu);
  if (g_devices.count(device) == 0) {
    auto st = std::make_unique<DeviceState>();
    hipDeviceProp_t p;
    HIP_CHECK(hipGetDeviceProperties(&p, device));
    st->id = device;
    st->n_cus = p.multiProcessorCount;
    st->lds_per_block = static_cast<int>(p.sharedMemPerBlock);
    st->max_threads = p.maxThreadsPerBlock;
    st->total_mem = static_cast<int64_t>(p.totalGlobalMem);
    std::string arch = p.gcnArchName;
    auto colon = arch.find(':');
    st->arch = colon == std::string::npos ? arch : arch.substr(0, colon);
    // Owner allocations come from the device's default memory pool, stream-ordered on the
    // legacy default stream (the reference allocates with cudaMallocAsync on one stream,
    // storage.hh:176-186).  Freed blocks stay in the pool (no release threshold) so the next
    // allocation reuses them without a device synchronisation.
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
      uint64_t keep = UINT64_MAX;
      if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess) st->pool = pool;
    }
    // this device's copy of the fixed scratch array (the device is current here)
    void* fixed = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&fixed, HIP_SYMBOL(scratch)));
    NCA_REQUIRE(fixed != nullptr, "no device address for the fixed scratch array");
    st->fixed_scratch = fixed;
    (void)hipGetLastError();
    g_devices[device] = std::move(st);
  }
  t_device = device;
}

Workspace& workspace(hipStream_t s, size_t partial_bytes, size_t big_bytes) {
  DeviceState& dev = current_device();
  std::lock_guard<std::mutex> lk(dev.mu);
  auto& slot = dev.workspaces[s];
  if (!slot) slot = std::make_unique<Workspace>();
  Workspace& w = *slot;
  // Growing allocates and synchronises, which a stream capture cannot hold: refuse it when this
  // stream is capturing, and for the legacy stream also when the query itself fails (the legacy
  // stream cannot be used while another stream captures in global mode).
  auto capturing = [&]() {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    const hipError_t e = hipStreamIsCapturing(s, &st);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return true;
    }
    return st != hipStreamCaptureStatusNone;
  };
  const size_t min_partials = size_t(1) << 20;
  if (partial_bytes < min_partials) partial_bytes = min_partials;
  if (w.partials_bytes < partial_bytes || w.counters == nullptr ||
      (big_bytes > 0 && w.big_bytes < big_bytes)) {
    if (capturing())
      NCA_THROW(NCA_ERR_INVALID,
                "workspace of this stream must grow during hipGraph capture; workspaces are per stream: run the call once on the capturing stream before capturing");
    // Grow: the device must be idle for this stream before old buffers are freed.
    HIP_CHECK(hipStreamSynchronize(s));
  }
  if (w.partials_bytes < partial_bytes) {
    if (w.partials) HIP_CHECK(hipFree(w.partials));
    w.partials = nullptr;
    HIP_CHECK(hipMalloc(&w.partials, partial_bytes));
    w.partials_bytes = partial_bytes;
  }
  if (w.counters == nullptr) {
    w.n_counters = kWorkspaceCounters;
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&w.counters), w.n_counters * sizeof(unsigned)));
    HIP_CHECK(hipMemset(w.counters, 0, w.n_counters * sizeof(unsigned)));
    HIP_CHECK(hipDeviceSynchronize());
  }
  if (big_bytes > 0 && w.big_bytes < big_bytes) {
    if (w.big) HIP_CHECK(hipFree(w.big));
    w.big = nullptr;
    HIP_CHECK(hipMalloc(&w.big, big_bytes));
    HIP_CHECK(hipMemset(w.big, 0, big_bytes));
    HIP_CHECK(hipDeviceSynchronize());
    w.big_bytes = big_bytes;
  }
  return w;
}

void finish_call(void* stream) {
  if (stream == nullptr) HIP_CHECK(hipStreamSynchronize(nullptr));
}

// ---- profiling ---------------------------------------------------------------------------
namespace {
std::mutex g_prof_mu;
bool g_prof_on = false;
struct ProfRecord {
  std::string name;
  hipEvent_t a, b;
};
std::vector<ProfRecord> g_prof;
}  // namespace

ProfScope::ProfScope(const char* n, hipStream_t s) : name(n), stream(s) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_on) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  if (hipEventRecord(e, s) != hipSuccess) { (void)hipEventDestroy(e); return; }
  start = e;
}

ProfScope::~ProfScope() {
  if (!start) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, stream);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.push_back({name, static_cast<hipEvent_t>(start), e});
}

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) NCA_THROW(NCA_ERR_HIP, std::string(what) + " launch failed: " + hipGetErrorString(e));
}

// ---- views ---------------------------------------------------------------------------
int64_t view_size(const nca_view& v) {
  int64_t n = 1;
  for (int d = 0; d < v.ndim; ++d) n *= v.shape[d];
  return n;
}

bool view_indirect(const nca_view& v) {
  if (v.pointer_axis >= 0) return true;
  if (v.layout != NCA_LAYOUT_SO) return false;
  for (int d = 0; d < v.ndim && d < NCA_MAX_NDIM; ++d)
    if (v.suboffsets[d] >= 0) return true;
  return false;
}

void validate_view_nc(const nca_view& v, const char* what, bool allow_null_data) {
  if (!view_indirect(v)) {
    validate_view(v, what, allow_null_data);
    return;
  }
  nca_view plain = v;
  plain.pointer_axis = -1;
  plain.layout = NCA_LAYOUT_NC;
  validate_view(plain, what, allow_null_data);
  if (v.layout == NCA_LAYOUT_SO) {
    NCA_REQUIRE(v.pointer_axis < 0, std::string(what) + ": an SO view has no NC pointer axis (pointer_axis must be -1)");
    return;
  }
  NCA_REQUIRE(v.pointer_axis < v.ndim, std::string(what) + ": pointer axis out of range");
  NCA_REQUIRE(v.ptr_offset >= 0, std::string(what) + ": negative pointer offset");
  const int64_t n = v.shape[v.pointer_axis];
  NCA_REQUIRE(n == 0 || v.ptr_offset + std::min<int64_t>(0, (n - 1) * v.strides[v.pointer_axis]) >= 0,
              std::string(what) + ": pointer-axis table index below zero");
}

void validate_view(const nca_view& v, const char* what, bool allow_null_data) {
  NCA_REQUIRE(v.ndim >= 0 && v.ndim <= NCA_MAX_NDIM, std::string(what) + ": ndim out of range");
  NCA_REQUIRE(v.layout == NCA_LAYOUT_NC || v.layout == NCA_LAYOUT_SO,
              std::string(what) + ": unknown layout " + std::to_string(v.layout));
  if (!is_supported(v.dtype))
    NCA_THROW(NCA_ERR_UNSUPPORTED,
              std::string(what) + ": dtype " +
                  (v.dtype >= 0 && v.dtype <= NCA_FLOAT16 ? dtype_info(v.dtype).name : "?") +
                  " has no device kernels");
  for (int d = 0; d < v.ndim; ++d)
    NCA_REQUIRE(v.shape[d] >= 0, std::string(what) + ": negative extent");
  if (v.pointer_axis >= 0)
    NCA_THROW(NCA_ERR_UNSUPPORTED, std::string(what) +
                                       ": this entry point does not take NC pointer-axis views (the expression "
                                       "entry points do)");
  if (view_indirect(v))
    NCA_THROW(NCA_ERR_UNSUPPORTED, std::string(what) +
                                       ": this entry point does not take SO views with suboffsets (the expression "
                                       "entry points do)");
  if (!allow_null_data && view_size(v) > 0)
    NCA_REQUIRE(v.data != nullptr, std::string(what) + ": null data pointer");
}

Chain chain_of(const nca_view& v) {
  Chain c;
  c.base = reinterpret_cast<uintptr_t>(v.data);
  for (int d = 0; d < v.ndim; ++d) c.strides[d] = v.strides[d];
  if (v.pointer_axis >= 0) {
    const int P = v.pointer_axis;
    c.base += static_cast<uintptr_t>(v.ptr_offset * static_cast<int64_t>(sizeof(void*)));
    c.strides[P] = v.strides[P] * static_cast<int64_t>(sizeof(void*));
    c.axes.push_back(P);
    c.subs.push_back(v.row_offset);
  } else if (v.layout == NCA_LAYOUT_SO) {
    for (int d = 0; d < v.ndim; ++d)
      if (v.suboffsets[d] >= 0) {
        c.axes.push_back(d);
        c.subs.push_back(v.suboffsets[d]);
      }
  }
  return c;
}

bool is_c_contiguous(const nca_view& v) {
  if (view_indirect(v)) return false;
  int64_t expect = itemsize(v.dtype);
  for (int d = v.ndim - 1; d >= 0; --d) {
    if (v.shape[d] == 1) continue;
    if (v.strides[d] != expect) return false;
    expect *= v.shape[d];
  }
  return true;
}

bool same_shape(const nca_view& a, const nca_view& b) {
  if (a.ndim != b.ndim) return false;
  for (int d = 0; d < a.ndim; ++d)
    if (a.shape[d] != b.shape[d]) return false;
  return true;
}

bool is_broadcast_scalar(const nca_view& v) {
  for (int d = 0; d < v.ndim; ++d)
    if (v.shape[d] != 1 && v.strides[d] != 0) return false;
  return true;
}

int coalesce(const int64_t* shape, const int64_t* strides, int ndim, int64_t* oshape,
             int64_t* ostrides) {
  int n = 0;
  for (int d = 0; d < ndim; ++d) {
    if (shape[d] == 1) continue;
    if (n > 0 && ostrides[n - 1] == strides[d] * shape[d]) {
      oshape[n - 1] *= shape[d];
      ostrides[n - 1] = strides[d];
    } else {
      oshape[n] = shape[d];
      ostrides[n] = strides[d];
      ++n;
    }
  }
  return n;
}

}  // namespace nca

using namespace nca;

extern "C" {

int nca_hip_init(int device) {
  return api_call([&] { init_device(device); });
}

int nca_hip_shutdown(void) {
  return api_call([&] {
    std::lock_guard<std::mutex> lk(g_devices_mu);
    for (auto& kv : g_devices) {
      HIP_CHECK(hipSetDevice(kv.first));
      HIP_CHECK(hipDeviceSynchronize());
      for (auto& w : kv.second->workspaces) {
        if (w.second->partials) (void)hipFree(w.second->partials);
        if (w.second->counters) (void)hipFree(w.second->counters);
        if (w.second->big) (void)hipFree(w.second->big);
      }
      kv.second->workspaces.clear();
      if (kv.second->pool) (void)hipMemPoolTrimTo(kv.second->pool, 0);
    }
  });
}

int nca_hip_last_error(char* buf, size_t n) {
  if (buf && n > 0) {
    size_t k = t_last_error.size() < n - 1 ? t_last_error.size() : n - 1;
    std::memcpy(buf, t_last_error.data(), k);
    buf[k] = '\0';
  }
  return static_cast<int>(t_last_error.size());
}

const char* nca_hip_version(void) { return "ncarray-hip 0.1.0 gfx950"; }

int nca_hip_current_device(int32_t* device) {
  return api_call([&] {
    NCA_REQUIRE(device != nullptr, "null argument");
    *device = current_device().id;
  });
}

int nca_hip_device_info(int32_t* n_cus, int32_t* lds_per_block, int64_t* total_mem,
                        char* arch_name, size_t arch_len) {
  return api_call([&] {
    DeviceState& d = current_device();
    if (n_cus) *n_cus = d.n_cus;
    if (lds_per_block) *lds_per_block = d.lds_per_block;
    if (total_mem) *total_mem = d.total_mem;
    if (arch_name && arch_len > 0) {
      std::snprintf(arch_name, arch_len, "%s", d.arch.c_str());
    }
  });
}

int nca_hip_malloc(void** ptr, size_t bytes) {
  return api_call([&] {
    NCA_REQUIRE(ptr != nullptr, "nca_hip_malloc: null out pointer");
    DeviceState& d = current_device();
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    if (d.pool) HIP_CHECK(hipMallocFromPoolAsync(ptr, bytes, d.pool, nullptr));
    else HIP_CHECK(hipMalloc(ptr, bytes));
  });
}

int nca_hip_free(void* ptr) {
  return api_call([&] {
    if (!ptr) return;
    DeviceState& d = current_device();
    if (d.pool) HIP_CHECK(hipFreeAsync(ptr, nullptr));
    else HIP_CHECK(hipFree(ptr));
  });
}

int nca_hip_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  return api_call([&] {
    if (bytes == 0) return;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    finish_call(stream);
  });
}

int nca_hip_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  return api_call([&] {
    if (bytes == 0) return;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    finish_call(stream);
  });
}

int nca_hip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  return api_call([&] {
    if (bytes == 0) return;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    finish_call(stream);
  });
}

int nca_hip_synchronize(void* stream) {
  return api_call([&] { HIP_CHECK(hipStreamSynchronize(as_stream(stream))); });
}

int nca_hip_profile_enable(int32_t on) {
  return api_call([&] {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on != 0;
  });
}

int nca_hip_profile_read(char* buf, size_t n) {
  // Completed records are turned into text and kept until a buffer large enough to hold
  // all of it is supplied; a NULL or short buffer only reports the needed length.
  static std::string pending;
  int rc = api_call([&] {
    std::vector<ProfRecord> recs;
    {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      recs.swap(g_prof);
    }
    for (auto& r : recs) {
      HIP_CHECK(hipEventSynchronize(r.b));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
      char line[160];
      std::snprintf(line, sizeof line, "%s %.6f\n", r.name.c_str(), ms);
      std::lock_guard<std::mutex> lk(g_prof_mu);
      pending += line;
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
  });
  if (rc != NCA_OK) return -rc;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  const int len = static_cast<int>(pending.size());
  if (buf && n > pending.size()) {
    std::memcpy(buf, pending.data(), pending.size());
    buf[pending.size()] = '\0';
    pending.clear();
  }
  return len;
}

int nca_hip_reduce_result_dtype(int32_t op, int32_t src_dtype, int32_t* out_dtype) {
  return api_call([&] {
    NCA_REQUIRE(out_dtype != nullptr, "null out_dtype");
    *out_dtype = reduce_result_dtype(op, src_dtype);
  });
}

int nca_hip_op_result_dtype(int32_t op, int32_t work_dtype, int32_t* out_dtype) {
  return api_call([&] {
    NCA_REQUIRE(out_dtype != nullptr, "null out_dtype");
    (void)dtype_info(work_dtype);
    *out_dtype = op_result_dtype(op, work_dtype);
  });
}

}  // extern "C"
