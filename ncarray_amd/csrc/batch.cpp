// Ragged batches: many uneven 1-D arrays sharing one expression, evaluated by one
// persistent launch.  The reference has no counterpart (its NC layout cannot express
// unequal sub-arrays); this is the batching addition.
//
// Plan: every item is cut into tiles of `tile` elements; a tile table (item, start)
// is uploaded once.  Kernel: a grid of (waves_per_cu / 4) blocks per CU walks the tiles
// round-robin (static, so results do not depend on scheduling), evaluates the expression
// with V-element vector loads, and reduces each tile's stored values (sum / max / min) to
// a tile partial.  A combine kernel folds each item's tile partials in tile order.
// The two launches can be captured once into a hipGraph and replayed.  An autotuner
// sweeps (tile, V, waves/CU), times each with HIP events and records the winner in a
// small text database keyed by the expression signature and the size distribution.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "expr_internal.hh"
#include "nca_internal.hh"
#include "tape.hh"

namespace nca {

struct BatchPlan {
  uint64_t expr = 0;
  int n_items = 0, n_in = 0;
  int fuse = 0, mode = NCA_ACCUM_PLAIN;
  bool with_out = true;
  std::vector<const void*> in_ptrs;
  std::vector<void*> out_ptrs;
  std::vector<int64_t> sizes;
  void* sums = nullptr;
  void* maxs = nullptr;
  void* mins = nullptr;
  int tile = 0, vec = 0, wpc = 0;
  int n_tiles = 0;
  int aligned = 1;  // every pointer 16-byte aligned
  std::vector<int> alias;  // operand k reads the same arrays as operand alias[k] in every item
  // device tables
  void* d_items = nullptr;      // per item: n_in in ptrs, out ptr, n
  int* d_tile_item = nullptr;
  int64_t* d_tile_start = nullptr;
  int* d_item_first = nullptr;  // first tile of each item
  void* d_part = nullptr;       // per tile: sum (8 B), max (8 B), min (8 B)
  int dev = 0;
};

namespace {
std::mutex g_batch_mu;
std::unordered_map<uint64_t, std::unique_ptr<BatchPlan>> g_plans;
std::unordered_map<uint64_t, hipGraphExec_t> g_graphs;
std::unordered_map<uint64_t, hipGraph_t> g_graph_defs;
uint64_t g_next_plan = 1ull << 48;
}  // namespace

static BatchPlan& plan_of(uint64_t h) {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  auto it = g_plans.find(h);
  NCA_REQUIRE(it != g_plans.end(), "unknown batch plan");
  return *it->second;
}

std::string gen_batch_source(const Tape& t, int fuse, int mode, int vec, bool with_out, const std::vector<int>& alias_in) {
  const int n_in = static_cast<int>(t.arr_dtypes.size());
  const std::string C = ctype_of(t.compute_dtype);
  const std::string D = ctype_of(t.dest_dtype);
  const int32_t rcd = compute_dtype(t.dest_dtype);
  const std::string RC = ctype_of(rcd);
  const bool fsum = is_float(rcd);
  const std::string SA = fsum ? RC : (dtype_info(rcd).is_signed ? "long long" : "unsigned long long");
  const bool comp = fsum && mode == NCA_ACCUM_COMPENSATED;
  // operands that alias an earlier one (the same arrays in every item) are loaded once
  std::vector<int> alias(n_in);
  for (int k = 0; k < n_in; ++k) alias[k] = (static_cast<int>(alias_in.size()) == n_in) ? alias_in[k] : k;
  std::vector<std::string> lnames, snames, roots;
  for (int k = 0; k < n_in; ++k) lnames.push_back("u" + std::to_string(k));
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) snames.push_back("s" + std::to_string(i));
  std::string body = emit_nodes(t, {t.root}, alias, lnames, snames, "lin", "t", roots);
  std::ostringstream os;
  os << "struct nca_item { const void* in[" << (n_in > 0 ? n_in : 1) << "]; void* out; nca_i64 n; };\n";
  os << "// batch of tape: " << t.signature << "\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void nca_batch(const nca_item* __restrict__ items, "
        "const int* __restrict__ tile_item, const nca_i64* __restrict__ tile_start, int n_tiles, int tile, "
        "char* __restrict__ part";
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  os << ") {\n";
  os << "  __shared__ double red_sd[4];\n  __shared__ " << SA << " red_si[4];\n  __shared__ " << RC << " red_mx[4];\n  __shared__ "
     << RC << " red_mn[4];\n";
  os << "  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;\n";
  os << "  constexpr int V = " << vec << ";\n";
  os << "  for (int tt = blockIdx.x; tt < n_tiles; tt += gridDim.x) {\n";
  os << "    const int it = tile_item[tt];\n    const nca_i64 st0 = tile_start[tt];\n";
  os << "    const nca_i64 n_it = items[it].n;\n";
  os << "    const nca_i64 len = (n_it - st0 < (nca_i64)tile) ? n_it - st0 : (nca_i64)tile;\n";
  for (int k = 0; k < n_in; ++k)
    if (alias[k] == k)
      os << "    const " << ctype_of(t.arr_dtypes[k]) << "* __restrict__ p" << k << " = reinterpret_cast<const "
         << ctype_of(t.arr_dtypes[k]) << "*>(items[it].in[" << k << "]) + st0;\n";
  os << "    " << D << "* __restrict__ out = items[it].out ? reinterpret_cast<" << D << "*>(items[it].out) + st0 : nullptr;\n";
  if (fuse & NCA_FUSE_SUM) {
    os << "    " << SA << " acc_s = 0;\n";
    if (comp) os << "    " << SA << " acc_c = 0;\n";
  }
  if (fuse & NCA_FUSE_MAX) os << "    " << RC << " acc_mx = nca_ident<" << D << ", " << RC << ">::lowest();\n";
  if (fuse & NCA_FUSE_MIN) os << "    " << RC << " acc_mn = nca_ident<" << D << ", " << RC << ">::max();\n";
  std::ostringstream upd;
  upd << "        { const " << RC << " rv = nca_cvt<" << RC << ", " << D << ">::from(st);\n";
  if (fuse & NCA_FUSE_SUM) upd << (comp ? "          nca_neumaier_add(acc_s, acc_c, (" + SA + ")rv);\n" : "          acc_s += (" + SA + ")rv;\n");
  if (fuse & NCA_FUSE_MAX) upd << "          acc_mx = nca_maxop(acc_mx, rv);\n";
  if (fuse & NCA_FUSE_MIN) upd << "          acc_mn = nca_minop(acc_mn, rv);\n";
  upd << "        }\n";
  // vector body: two vectors per input in flight per thread (both issued before either is
  // used), streaming loads and stores (every element is read and written once)
  auto emit_vec = [&](const std::string& sfx, const std::string& idx) {
    os << "      {\n      nca_vec<" << D << ", V> r;\n#pragma unroll\n      for (int j = 0; j < V; ++j) {\n";
    for (int k = 0; k < n_in; ++k)
      if (alias[k] == k)
        os << "        const " << C << " u" << k << " = nca_cvt<" << C << ", " << ctype_of(t.arr_dtypes[k]) << ">::from(x" << k
           << sfx << ".v[j]);\n";
    if (t.uses_idx) os << "        const nca_i64 lin = st0 + " << idx << " * V + j;\n";
    os << body;
    os << "        const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
    os << "        r.v[j] = st;\n" << upd.str() << "      }\n";
    if (with_out)
      os << "      if (out) nca_store_nt(reinterpret_cast<nca_vec<" << D << ", V>*>(out + " << idx << " * V), r);\n";
    os << "      }\n";
  };
  auto emit_loads = [&](const std::string& sfx, const std::string& idx) {
    for (int k = 0; k < n_in; ++k)
      if (alias[k] == k)
        os << "      const nca_vec<" << ctype_of(t.arr_dtypes[k]) << ", V> x" << k << sfx << " = nca_load_nt(reinterpret_cast<const nca_vec<"
           << ctype_of(t.arr_dtypes[k]) << ", V>*>(p" << k << " + " << idx << " * V));\n";
  };
  os << "    const nca_i64 nv = len / V;\n";
  os << "    nca_i64 i = threadIdx.x;\n";
  os << "    for (; i + (nca_i64)blockDim.x < nv; i += 2 * (nca_i64)blockDim.x) {\n";
  os << "      const nca_i64 i1 = i + (nca_i64)blockDim.x;\n";
  emit_loads("a", "i");
  emit_loads("b", "i1");
  emit_vec("a", "i");
  emit_vec("b", "i1");
  os << "    }\n";
  os << "    if (i < nv) {\n";
  emit_loads("a", "i");
  emit_vec("a", "i");
  os << "    }\n";
  os << "    for (nca_i64 i = nv * V + threadIdx.x; i < len; i += blockDim.x) {\n";
  for (int k = 0; k < n_in; ++k)
    if (alias[k] == k) os << "        const " << C << " u" << k << " = nca_ld<" << C << ">(p" << k << " + i);\n";
  if (t.uses_idx) os << "        const nca_i64 lin = st0 + i;\n";
  os << body;
  os << "        const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
  if (with_out) os << "        if (out) out[i] = st;\n";
  os << upd.str() << "    }\n";
  // tile partial
  if (fuse & NCA_FUSE_SUM) {
    if (fsum) {
      if (comp) os << "    nca_wave_comp(acc_s, acc_c);\n    if (lane == 0) red_sd[wid] = nca_comp_value(acc_s, acc_c);\n";
      else os << "    acc_s = nca_wave_sum(acc_s);\n    if (lane == 0) red_sd[wid] = (double)acc_s;\n";
    } else {
      os << "    acc_s = nca_wave_sum(acc_s);\n    if (lane == 0) red_si[wid] = acc_s;\n";
    }
  }
  if (fuse & NCA_FUSE_MAX) os << "    acc_mx = nca_wave_max(acc_mx);\n    if (lane == 0) red_mx[wid] = acc_mx;\n";
  if (fuse & NCA_FUSE_MIN) os << "    acc_mn = nca_wave_min(acc_mn);\n    if (lane == 0) red_mn[wid] = acc_mn;\n";
  os << "    __syncthreads();\n    if (threadIdx.x == 0) {\n      char* slot = part + 24 * (size_t)tt;\n";
  if (fuse & NCA_FUSE_SUM) {
    if (fsum) os << "      *(double*)slot = red_sd[0] + red_sd[1] + red_sd[2] + red_sd[3];\n";
    else os << "      *(" << SA << "*)slot = red_si[0] + red_si[1] + red_si[2] + red_si[3];\n";
  }
  if (fuse & NCA_FUSE_MAX)
    os << "      " << RC << " m = red_mx[0];\n      for (int w = 1; w < 4; ++w) m = nca_maxop(m, red_mx[w]);\n      *(" << RC << "*)(slot + 8) = m;\n";
  if (fuse & NCA_FUSE_MIN)
    os << "      " << RC << " m2 = red_mn[0];\n      for (int w = 1; w < 4; ++w) m2 = nca_minop(m2, red_mn[w]);\n      *(" << RC << "*)(slot + 16) = m2;\n";
  os << "    }\n    __syncthreads();\n  }\n}\n";
  // one wave per item: lane l folds the item's tile slots l, l + 64, ... in order, then a fixed
  // butterfly combines the lanes: deterministic for a given plan
  const int32_t sum_dt = reduce_result_dtype(NCA_REDUCE_SUM, t.dest_dtype);
  os << "extern \"C\" __global__ __launch_bounds__(256) void nca_batch_combine(const int* __restrict__ first, int n_items, "
        "int n_tiles, const char* __restrict__ part, " << ctype_of(sum_dt) << "* __restrict__ sums, " << D
     << "* __restrict__ maxs, " << D << "* __restrict__ mins) {\n"
     << "  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;\n  if (i >= n_items) return;\n"
     << "  const int t0 = first[i], t1 = (i + 1 < n_items) ? first[i + 1] : n_tiles;\n";
  if (fuse & NCA_FUSE_SUM) {
    const std::string ST = fsum ? std::string("double") : SA;
    os << "  " << ST << " s = 0;\n  for (int t = t0 + lane; t < t1; t += 64) s += *(const " << ST
       << "*)(part + 24 * (size_t)t);\n  s = nca_wave_sum(s);\n";
    if (fsum) os << "  if (lane == 0) nca_st(sums + i, s);\n";
    else os << "  if (lane == 0) sums[i] = (" << ctype_of(sum_dt) << ")s;\n";
  }
  if (fuse & NCA_FUSE_MAX)
    os << "  " << RC << " m = nca_ident<" << D << ", " << RC << ">::lowest();\n  for (int t = t0 + lane; t < t1; t += 64) m = nca_maxop(m, *(const "
       << RC << "*)(part + 24 * (size_t)t + 8));\n  m = nca_wave_max(m);\n  if (lane == 0) nca_st(maxs + i, m);\n";
  if (fuse & NCA_FUSE_MIN)
    os << "  " << RC << " m2 = nca_ident<" << D << ", " << RC << ">::max();\n  for (int t = t0 + lane; t < t1; t += 64) m2 = nca_minop(m2, *(const "
       << RC << "*)(part + 24 * (size_t)t + 16));\n  m2 = nca_wave_min(m2);\n  if (lane == 0) nca_st(mins + i, m2);\n";
  os << "}\n";
  return os.str();
}

static void free_plan_tables(BatchPlan& p) {
  if (p.d_items) (void)hipFree(p.d_items);
  if (p.d_tile_item) (void)hipFree(p.d_tile_item);
  if (p.d_tile_start) (void)hipFree(p.d_tile_start);
  if (p.d_item_first) (void)hipFree(p.d_item_first);
  if (p.d_part) (void)hipFree(p.d_part);
  p.d_items = p.d_tile_item = nullptr;
  p.d_tile_start = nullptr;
  p.d_item_first = nullptr;
  p.d_part = nullptr;
}

// (Re)build the tile tables for p.tile.
static void build_tables(BatchPlan& p) {
  free_plan_tables(p);
  std::vector<int> tile_item, first(p.n_items);
  std::vector<int64_t> tile_start;
  for (int i = 0; i < p.n_items; ++i) {
    first[i] = static_cast<int>(tile_item.size());
    for (int64_t s = 0; s < p.sizes[i]; s += p.tile) {
      tile_item.push_back(i);
      tile_start.push_back(s);
    }
    NCA_REQUIRE(tile_item.size() < (size_t(1) << 31), "too many tiles");
  }
  p.n_tiles = static_cast<int>(tile_item.size());
  const size_t item_bytes = sizeof(void*) * (std::max(p.n_in, 1) + 1) + sizeof(int64_t);
  std::vector<unsigned char> items(item_bytes * std::max(p.n_items, 1), 0);
  for (int i = 0; i < p.n_items; ++i) {
    unsigned char* q = items.data() + item_bytes * i;
    for (int k = 0; k < p.n_in; ++k) std::memcpy(q + sizeof(void*) * k, &p.in_ptrs[static_cast<size_t>(i) * p.n_in + k], sizeof(void*));
    void* o = p.with_out ? p.out_ptrs[i] : nullptr;
    std::memcpy(q + sizeof(void*) * std::max(p.n_in, 1), &o, sizeof(void*));
    std::memcpy(q + sizeof(void*) * (std::max(p.n_in, 1) + 1), &p.sizes[i], sizeof(int64_t));
  }
  HIP_CHECK(hipMalloc(&p.d_items, items.size()));
  HIP_CHECK(hipMemcpy(p.d_items, items.data(), items.size(), hipMemcpyHostToDevice));
  const size_t nt = std::max(p.n_tiles, 1);
  HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p.d_tile_item), nt * sizeof(int)));
  HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p.d_tile_start), nt * sizeof(int64_t)));
  HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p.d_item_first), std::max(p.n_items, 1) * sizeof(int)));
  HIP_CHECK(hipMalloc(&p.d_part, nt * 24));
  if (p.n_tiles) {
    HIP_CHECK(hipMemcpy(p.d_tile_item, tile_item.data(), tile_item.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(p.d_tile_start, tile_start.data(), tile_start.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  }
  if (p.n_items) HIP_CHECK(hipMemcpy(p.d_item_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
}

static std::string batch_db_key(const BatchPlan& p, const Tape& t) {
  // size distribution: item count and total elements to 2 significant decimal digits
  double total = 0;
  for (auto s : p.sizes) total += static_cast<double>(s);
  char buf[128];
  std::snprintf(buf, sizeof buf, "|n%d|tot%.1e|f%d|m%d|o%d", p.n_items, total, p.fuse, p.mode, p.with_out ? 1 : 0);
  return sha256_hex(t.signature + buf).substr(0, 32);
}

static std::string autotune_db_path() {
  const char* e = std::getenv("NCA_HIP_AUTOTUNE_DB");
  if (e && *e) return e;
  const char* h = std::getenv("HOME");
  return std::string(h && *h ? h : "/tmp") + "/.cache/ncarray_amd/autotune.txt";
}

static bool db_lookup(const std::string& key, int& tile, int& vec, int& wpc) {
  std::ifstream f(autotune_db_path());
  std::string k;
  int a, b, c;
  double ms;
  bool found = false;
  while (f >> k >> a >> b >> c >> ms)
    if (k == key) { tile = a; vec = b; wpc = c; found = true; }  // last entry wins
  return found;
}

static void db_store(const std::string& key, int tile, int vec, int wpc, double ms) {
  std::ofstream f(autotune_db_path(), std::ios::app);
  if (f) f << key << " " << tile << " " << vec << " " << wpc << " " << ms << "\n";
}

static void batch_launch_impl(BatchPlan& p, const nca_scalar* scalars, int n_scalars, hipStream_t s) {
  ExprProgram& ep = expr_program(p.expr);
  const Tape& t = ep.tape;
  ScalarSlots sc = make_scalar_slots(t, scalars, n_scalars);
  const int vec = p.aligned ? p.vec : 1;
  std::string key = "batch|" + std::to_string(vec) + "|" + std::to_string(p.fuse) + "|" + std::to_string(p.mode) + "|" +
                    std::to_string(p.with_out) + "|a";
  for (int a : p.alias) key += std::to_string(a) + ",";
  hipFunction_t fn = program_kernel(ep, key, [&]() { return gen_batch_source(t, p.fuse, p.mode, vec, p.with_out, p.alias); },
                                    "nca_batch");
  hipFunction_t comb = program_kernel(ep, key + "|combine",
                                      [&]() { return gen_batch_source(t, p.fuse, p.mode, vec, p.with_out, p.alias); },
                                      "nca_batch_combine");
  DeviceState& dev = current_device();
  int blocks = std::max(1, dev.n_cus * p.wpc / 4);
  if (blocks > p.n_tiles) blocks = std::max(1, p.n_tiles);
  std::vector<void*> params;
  void* items = p.d_items;
  void* ti = p.d_tile_item;
  void* ts = p.d_tile_start;
  int nt = p.n_tiles, tile = p.tile;
  void* part = p.d_part;
  params.push_back(&items);
  params.push_back(&ti);
  params.push_back(&ts);
  params.push_back(&nt);
  params.push_back(&tile);
  params.push_back(&part);
  for (auto& x : sc.slots) params.push_back(&x);
  if (p.n_tiles > 0) {
    ProfScope ps("nca_batch", s);
    jit_launch(fn, dim3(blocks), dim3(256), 0, s, params);
  }
  if (p.fuse && p.n_items > 0) {
    std::vector<void*> cp;
    void* first = p.d_item_first;
    int ni = p.n_items;
    void* sums = p.sums;
    void* maxs = p.maxs;
    void* mins = p.mins;
    cp.push_back(&first);
    cp.push_back(&ni);
    cp.push_back(&nt);
    cp.push_back(&part);
    cp.push_back(&sums);
    cp.push_back(&maxs);
    cp.push_back(&mins);
    ProfScope ps("nca_batch_combine", s);
    jit_launch(comb, dim3((p.n_items + 3) / 4), dim3(256), 0, s, cp);  // one wave per item
  }
}

}  // namespace nca

using namespace nca;

extern "C" {

int nca_hip_batch_plan(uint64_t expr_handle, int32_t n_items, int32_t n_in, const void* const* in_ptrs,
                       void* const* out_ptrs, const int64_t* sizes, int32_t fuse, int32_t accum_mode, void* sums,
                       void* maxs, void* mins, int32_t tile_elems, int32_t vec, int32_t waves_per_cu, uint64_t* plan) {
  return api_call([&] {
    NCA_REQUIRE(plan != nullptr && sizes != nullptr, "null argument");
    ExprProgram& ep = expr_program(expr_handle);
    const Tape& t = ep.tape;
    NCA_REQUIRE(n_items >= 0, "negative item count");
    NCA_REQUIRE(n_in == static_cast<int>(t.arr_dtypes.size()), "n_in must equal the expression's array count");
    NCA_REQUIRE(n_in == 0 || in_ptrs != nullptr || n_items == 0, "null in_ptrs");
    NCA_REQUIRE((fuse & ~(NCA_FUSE_SUM | NCA_FUSE_MAX | NCA_FUSE_MIN)) == 0, "bad fuse flags");
    NCA_REQUIRE(!(fuse & NCA_FUSE_SUM) || sums, "sum requested without sums");
    NCA_REQUIRE(!(fuse & NCA_FUSE_MAX) || maxs, "max requested without maxs");
    NCA_REQUIRE(!(fuse & NCA_FUSE_MIN) || mins, "min requested without mins");
    NCA_REQUIRE(out_ptrs != nullptr || fuse != 0, "nothing to compute: no outputs and no reductions");
    DeviceState& dev = current_device();
    auto p = std::make_unique<BatchPlan>();
    p->expr = expr_handle;
    p->n_items = n_items;
    p->n_in = n_in;
    p->fuse = fuse;
    p->mode = accum_mode;
    p->with_out = out_ptrs != nullptr;
    p->in_ptrs.assign(in_ptrs, in_ptrs + static_cast<size_t>(n_items) * n_in);
    if (out_ptrs) p->out_ptrs.assign(out_ptrs, out_ptrs + n_items);
    p->sizes.assign(sizes, sizes + n_items);
    p->sums = sums;
    p->maxs = maxs;
    p->mins = mins;
    p->dev = dev.id;
    for (auto v : p->sizes) NCA_REQUIRE(v >= 0, "negative item size");
    for (auto ptr : p->in_ptrs) if (reinterpret_cast<uintptr_t>(ptr) % 16) p->aligned = 0;
    p->alias.resize(n_in);
    {
      const Tape& tp = expr_program(expr_handle).tape;
      for (int k = 0; k < n_in; ++k) {
        p->alias[k] = k;
        for (int j = 0; j < k && p->alias[k] == k; ++j) {
          if (p->alias[j] != j || tp.arr_dtypes[j] != tp.arr_dtypes[k]) continue;
          bool same = true;
          for (int i = 0; i < n_items && same; ++i)
            same = p->in_ptrs[static_cast<size_t>(i) * n_in + j] == p->in_ptrs[static_cast<size_t>(i) * n_in + k];
          if (same) p->alias[k] = j;
        }
      }
    }
    for (auto ptr : p->out_ptrs) if (reinterpret_cast<uintptr_t>(ptr) % 16) p->aligned = 0;
    int dt = 1, dv = 0, dw = 0;
    // defaults measured on MI355X for the 4096-array C0 batch: 8192-element tiles, 8 waves per CU
    int tt = 8192, tv = 4, tw = 8;
    if (db_lookup(batch_db_key(*p, t), dt, dv, dw)) { tt = dt; tv = dv; tw = dw; }
    p->tile = tile_elems > 0 ? tile_elems : tt;
    p->vec = vec > 0 ? vec : tv;
    p->wpc = waves_per_cu > 0 ? waves_per_cu : tw;
    NCA_REQUIRE(p->vec == 1 || p->vec == 2 || p->vec == 4 || p->vec == 8, "vec must be 1, 2, 4 or 8");
    NCA_REQUIRE(p->tile % (p->vec * 256) == 0 || p->tile < p->vec * 256, "tile must be a multiple of 256 * vec");
    NCA_REQUIRE(p->tile % p->vec == 0, "tile must be a multiple of vec");
    build_tables(*p);
    std::lock_guard<std::mutex> lk(g_batch_mu);
    uint64_t h = g_next_plan++;
    g_plans[h] = std::move(p);
    *plan = h;
  });
}

int nca_hip_batch_launch(uint64_t plan, const nca_scalar* scalars, int32_t n_scalars, void* stream) {
  return api_call([&] {
    BatchPlan& p = plan_of(plan);
    batch_launch_impl(p, scalars, n_scalars, as_stream(stream));
    check_launch("batch");
    finish_call(stream);
  });
}

int nca_hip_batch_graph(uint64_t plan, const nca_scalar* scalars, int32_t n_scalars, uint64_t* graph) {
  return api_call([&] {
    NCA_REQUIRE(graph != nullptr, "null graph");
    BatchPlan& p = plan_of(plan);
    // compile outside the capture
    batch_launch_impl(p, scalars, n_scalars, nullptr);
    HIP_CHECK(hipDeviceSynchronize());
    hipStream_t cs;
    HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
      batch_launch_impl(p, scalars, n_scalars, cs);
    } catch (...) {
      hipGraph_t g;
      (void)hipStreamEndCapture(cs, &g);
      (void)hipStreamDestroy(cs);
      throw;
    }
    hipGraph_t g;
    HIP_CHECK(hipStreamEndCapture(cs, &g));
    hipGraphExec_t ge;
    HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    HIP_CHECK(hipStreamDestroy(cs));
    std::lock_guard<std::mutex> lk(g_batch_mu);
    uint64_t h = g_next_plan++;
    g_graphs[h] = ge;
    g_graph_defs[h] = g;
    *graph = h;
  });
}

int nca_hip_graph_launch(uint64_t graph, void* stream) {
  return api_call([&] {
    hipGraphExec_t ge;
    {
      std::lock_guard<std::mutex> lk(g_batch_mu);
      auto it = g_graphs.find(graph);
      NCA_REQUIRE(it != g_graphs.end(), "unknown graph");
      ge = it->second;
    }
    HIP_CHECK(hipGraphLaunch(ge, as_stream(stream)));
    finish_call(stream);
  });
}

int nca_hip_graph_destroy(uint64_t graph) {
  return api_call([&] {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto it = g_graphs.find(graph);
    NCA_REQUIRE(it != g_graphs.end(), "unknown graph");
    HIP_CHECK(hipGraphExecDestroy(it->second));
    HIP_CHECK(hipGraphDestroy(g_graph_defs[graph]));
    g_graphs.erase(it);
    g_graph_defs.erase(graph);
  });
}

int nca_hip_batch_destroy(uint64_t plan) {
  return api_call([&] {
    std::unique_ptr<BatchPlan> p;
    {
      std::lock_guard<std::mutex> lk(g_batch_mu);
      auto it = g_plans.find(plan);
      NCA_REQUIRE(it != g_plans.end(), "unknown batch plan");
      p = std::move(it->second);
      g_plans.erase(it);
    }
    HIP_CHECK(hipDeviceSynchronize());
    free_plan_tables(*p);
  });
}

int nca_hip_batch_config(uint64_t plan, int32_t* tile_elems, int32_t* vec, int32_t* waves_per_cu, int32_t* n_tiles) {
  return api_call([&] {
    BatchPlan& p = plan_of(plan);
    if (tile_elems) *tile_elems = p.tile;
    if (vec) *vec = p.aligned ? p.vec : 1;
    if (waves_per_cu) *waves_per_cu = p.wpc;
    if (n_tiles) *n_tiles = p.n_tiles;
  });
}

int nca_hip_batch_autotune(uint64_t plan, const nca_scalar* scalars, int32_t n_scalars, int32_t reps, float* best_ms) {
  return api_call([&] {
    BatchPlan& p = plan_of(plan);
    ExprProgram& ep = expr_program(p.expr);
    if (reps < 1) reps = 3;
    const int tiles[] = {4096, 8192, 16384, 32768, 65536, 131072};
    const int vecs[] = {2, 4, 8};
    const int wpcs[] = {8, 16, 32};
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    int bt = p.tile, bv = p.vec, bw = p.wpc;
    for (int tl : tiles)
      for (int v : vecs) {
        if (tl % (v * 256) != 0) continue;
        p.tile = tl;
        p.vec = v;
        build_tables(p);
        for (int w : wpcs) {
          p.wpc = w;
          batch_launch_impl(p, scalars, n_scalars, nullptr);  // warm-up (and compile)
          HIP_CHECK(hipDeviceSynchronize());
          HIP_CHECK(hipEventRecord(e0, nullptr));
          for (int r = 0; r < reps; ++r) batch_launch_impl(p, scalars, n_scalars, nullptr);
          HIP_CHECK(hipEventRecord(e1, nullptr));
          HIP_CHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          ms /= reps;
          if (ms < best) { best = ms; bt = tl; bv = v; bw = w; }
        }
      }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    p.tile = bt;
    p.vec = bv;
    p.wpc = bw;
    build_tables(p);
    db_store(batch_db_key(p, ep.tape), bt, bv, bw, best);
    if (best_ms) *best_ms = best;
  });
}

}  // extern "C"
