// hipRTC kernel cache.  Replaces device::RuntimeCompiler's NVRTC compile and cubin cache
// (rtcompiler.cc:49-89, 445-510; jit_utils.cc:20-46; path_utils.cc:31-52, 131-175).
//
// Key = SHA-256(prelude + source + options + arch + hipRTC version).
// Memory cache: key -> loaded module per device, (key, kernel name) -> function.  Disk cache: <dir>/<key>.hsaco with a
// header (magic, version, payload length, SHA-256 of the payload) that is checked before
// the code object is loaded; files are written to a temporary name and renamed, so a
// concurrent reader never sees a partial file.  A corrupt or truncated file is ignored
// and recompiled.
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "nca_internal.hh"

namespace nca {

// ---- SHA-256 (FIPS 180-4) -------------------------------------------------------------
namespace {
inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
const uint32_t kSha256K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
}  // namespace

std::string sha256_hex(const std::string& data) {
  uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                   0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  std::string msg = data;
  uint64_t bitlen = static_cast<uint64_t>(data.size()) * 8;
  msg.push_back(static_cast<char>(0x80));
  while (msg.size() % 64 != 56) msg.push_back('\0');
  for (int i = 7; i >= 0; --i) msg.push_back(static_cast<char>((bitlen >> (8 * i)) & 0xff));
  for (size_t off = 0; off < msg.size(); off += 64) {
    uint32_t w[64];
    for (int i = 0; i < 16; ++i) {
      const unsigned char* p = reinterpret_cast<const unsigned char*>(msg.data() + off + 4 * i);
      w[i] = (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3];
    }
    for (int i = 16; i < 64; ++i) {
      uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
      uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
      w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; ++i) {
      uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
      uint32_t ch = (e & f) ^ (~e & g);
      uint32_t t1 = hh + S1 + ch + kSha256K[i] + w[i];
      uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
      uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
      uint32_t t2 = S0 + mj;
      hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
  }
  char out[65];
  for (int i = 0; i < 8; ++i) std::snprintf(out + 8 * i, 9, "%08x", h[i]);
  return std::string(out, 64);
}

// ---- prelude ----------------------------------------------------------------------------
static const char kPrelude[] = {
#embed "nca_device.hh"
    , 0};

const std::string& device_prelude() {
  static const std::string s(kPrelude);
  return s;
}

// Options for every generated kernel.  -ffp-contract=off keeps a*b+c as two rounded
// operations (as the CPU reference evaluates them); packed fp32 ops are disabled.
static std::vector<std::string> rtc_options(const std::string& arch) {
  return {"--offload-arch=" + arch, "-O3", "-std=c++17", "-ffp-contract=off",
          "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

std::string hiprtc_options_string() {
  std::string s;
  for (auto& o : rtc_options("gfx950")) s += o + " ";
  return s;
}

static std::string cache_dir() {
  const char* env = std::getenv("NCA_HIP_CACHE_DIR");
  std::string dir;
  if (env && *env) {
    dir = env;
  } else if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) {
    dir = std::string(x) + "/ncarray_amd";
  } else if (const char* h = std::getenv("HOME"); h && *h) {
    dir = std::string(h) + "/.cache/ncarray_amd";
  } else {
    dir = "/tmp/ncarray_amd-" + std::to_string(getuid());
  }
  // Create the directory chain; failures leave the disk cache disabled.
  std::string partial;
  std::stringstream ss(dir);
  std::string item;
  if (!dir.empty() && dir[0] == '/') partial = "/";
  while (std::getline(ss, item, '/')) {
    if (item.empty()) continue;
    partial += item + "/";
    mkdir(partial.c_str(), 0755);
  }
  return dir;
}

// A cached code object: magic, the SHA-256 key of the source it was compiled from (so a file
// copied in under another name is refused), the code length, the SHA-256 of the code (a
// truncated or damaged file is refused), the code.
static const char kMagic[8] = {'N', 'C', 'A', 'H', 'S', 'C', 'O', '2'};

static bool read_cached(const std::string& path, const std::string& key, std::string& code) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  char magic[8];
  char fkey[64];
  uint64_t len = 0;
  char digest[64];
  if (!f.read(magic, 8) || std::memcmp(magic, kMagic, 8) != 0) return false;
  if (!f.read(fkey, 64) || key.size() != 64 || std::memcmp(fkey, key.data(), 64) != 0) return false;
  if (!f.read(reinterpret_cast<char*>(&len), 8) || len == 0 || len > (uint64_t(1) << 30)) return false;
  if (!f.read(digest, 64)) return false;
  code.resize(len);
  if (!f.read(&code[0], static_cast<std::streamsize>(len))) return false;
  return sha256_hex(code) == std::string(digest, 64);
}

static void write_cached(const std::string& path, const std::string& key, const std::string& code) {
  std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    uint64_t len = code.size();
    std::string digest = sha256_hex(code);
    f.write(kMagic, 8);
    f.write(key.data(), 64);
    f.write(reinterpret_cast<const char*>(&len), 8);
    f.write(digest.data(), 64);
    f.write(code.data(), static_cast<std::streamsize>(code.size()));
    if (!f) { std::remove(tmp.c_str()); return; }
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

namespace {
std::atomic<int64_t> g_stat_compiles{0}, g_stat_disk{0}, g_stat_memory{0};
}  // namespace

#define RTC_CHECK(expr)                                                                   \
  do {                                                                                    \
    hiprtcResult r_ = (expr);                                                             \
    if (r_ != HIPRTC_SUCCESS)                                                             \
      NCA_THROW(NCA_ERR_COMPILE, std::string(#expr) + ": " + hiprtcGetErrorString(r_));   \
  } while (0)

static std::string compile_code_object(const std::string& full_src, const std::string& arch) {
  hiprtcProgram prog;
  RTC_CHECK(hiprtcCreateProgram(&prog, full_src.c_str(), "nca_kernel.hip", 0, nullptr, nullptr));
  auto opts = rtc_options(arch);
  std::vector<const char*> copts;
  for (auto& o : opts) copts.push_back(o.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, static_cast<int>(copts.size()), copts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    NCA_THROW(NCA_ERR_COMPILE, "hipRTC compilation failed:\n" + log);
  }
  size_t sz = 0;
  RTC_CHECK(hiprtcGetCodeSize(prog, &sz));
  std::string code(sz, '\0');
  RTC_CHECK(hiprtcGetCode(prog, &code[0]));
  hiprtcDestroyProgram(&prog);
  return code;
}

namespace {
std::mutex g_jit_mu;
std::unordered_map<std::string, hipModule_t> g_modules;   // device id + source key
std::unordered_map<std::string, hipFunction_t> g_funcs;   // device id + source key + name
}  // namespace

hipFunction_t jit_get(const std::string& source, const std::string& name) {
  DeviceState& dev = current_device();
  int maj = 0, min = 0;
  hiprtcVersion(&maj, &min);
  std::string full = device_prelude() + "\n" + source;
  std::string key = sha256_hex(full + "\n//opts:" + hiprtc_options_string() + "\n//arch:" + dev.arch +
                               "\n//hiprtc:" + std::to_string(maj) + "." + std::to_string(min));
  std::string mkey = std::to_string(dev.id) + ":" + key;
  std::lock_guard<std::mutex> lk(g_jit_mu);
  auto fit = g_funcs.find(mkey + ":" + name);
  if (fit != g_funcs.end()) {
    ++g_stat_memory;
    return fit->second;
  }
  hipModule_t mod = nullptr;
  auto mit = g_modules.find(mkey);
  if (mit != g_modules.end()) {
    mod = mit->second;
    ++g_stat_memory;
  } else {
    std::string code;
    std::string dir = cache_dir();
    std::string path = dir + "/" + key + ".hsaco";
    bool from_disk = read_cached(path, key, code);
    hipError_t e = hipErrorInvalidImage;
    hipFunction_t probe = nullptr;
    if (from_disk) {
      e = hipModuleLoadData(&mod, code.data());
      if (e == hipSuccess && hipModuleGetFunction(&probe, mod, name.c_str()) != hipSuccess) {
        // a file that loads but lacks the kernel: not this source's code
        (void)hipModuleUnload(mod);
        mod = nullptr;
        e = hipErrorNotFound;
      }
      (void)hipGetLastError();
    }
    if (e == hipSuccess) {
      ++g_stat_disk;
    } else {
      // no cached file, or one that is truncated, damaged, foreign or stale: compile from source
      code = compile_code_object(full, dev.arch);
      ++g_stat_compiles;
      write_cached(path, key, code);
      e = hipModuleLoadData(&mod, code.data());
    }
    if (e != hipSuccess) NCA_THROW(NCA_ERR_HIP, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    g_modules[mkey] = mod;
  }
  hipFunction_t fn = nullptr;
  HIP_CHECK(hipModuleGetFunction(&fn, mod, name.c_str()));
  g_funcs[mkey + ":" + name] = fn;
  return fn;
}

void jit_launch(hipFunction_t fn, dim3 grid, dim3 block, size_t shmem, hipStream_t s,
                std::vector<void*>& params) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  HIP_CHECK(hipModuleLaunchKernel(fn, grid.x, grid.y, grid.z, block.x, block.y, block.z,
                                  static_cast<unsigned>(shmem), s, params.data(), nullptr));
}

}  // namespace nca

extern "C" int nca_hip_jit_stats(int64_t* compiles, int64_t* disk_loads, int64_t* memory_hits) {
  return nca::api_call([&] {
    NCA_REQUIRE(compiles && disk_loads && memory_hits, "null argument");
    *compiles = nca::g_stat_compiles.load();
    *disk_loads = nca::g_stat_disk.load();
    *memory_hits = nca::g_stat_memory.load();
  });
}

extern "C" int nca_hip_rtc_compile(const char* source, const char* arch, char* log, size_t n) {
  std::string msg;
  int rc = nca::api_call([&] {
    NCA_REQUIRE(source != nullptr, "null source");
    std::string a = (arch && *arch) ? arch : "gfx950";
    std::string src = source;
    // sources from nca_hip_*_source already carry the prelude
    (void)nca::compile_code_object(src, a);
  });
  if (rc != NCA_OK) {
    char buf[1] = {0};
    int len = nca_hip_last_error(buf, 1);
    msg.resize(static_cast<size_t>(len) + 1);
    nca_hip_last_error(&msg[0], msg.size());
  }
  if (log && n > 0) {
    size_t k = msg.size() < n - 1 ? msg.size() : n - 1;
    std::memcpy(log, msg.data(), k);
    log[k] = '\0';
  }
  return rc;
}

