// Stencils: Stencil<NDIM>::create/apply (stencil.hh:59-255) and the NVRTC stencil
// source generator (rtcompiler.hh:187-385), rebuilt for gfx950 with hipRTC.
//
// dest[c] = f(src[c*step + off_k] for each offset k, extra_j[c]) with
// step = src.shape / dest.shape (integer division, rtcompiler.hh:363-368).
//
// Two generated kernels:
//  * march (3-D, step 1, |offsets| <= 3, C-contiguous, 16-byte aligned rows): a block owns
//    a 64 x 16 (x, y) tile and a run of z planes.  Each plane of the tile plus its halo is
//    staged once in an LDS ring of 2*Rz+2 planes (one barrier per plane), the next plane
//    is prefetched into registers while the current one is computed, and every thread
//    produces 4 consecutive x outputs (16-byte stores).  Reductions of the stored values
//    (sum along axis 0, global sum/max/min) ride along in registers.  Blocks are mapped so
//    that neighbouring tiles land on the same XCD (shared L2 for halos).
//  * direct (any rank, any step, pro/epilogue hooks): one output per thread, one
//    predicated load per offset, the reference's launch shape (ceil(n/256) x 256).
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <sstream>
#include <unordered_map>

#include "expr_internal.hh"
#include "nca_internal.hh"
#include "tape.hh"

namespace nca {

struct StencilProgram {
  int ndim = 0;
  int n_off = 0;
  std::vector<std::array<int64_t, NCA_MAX_NDIM>> offsets;
  int32_t src_dtype = NCA_FLOAT32;
  std::vector<int32_t> extra_dtypes;
  Tape tape;
  std::string extra_params, prologue, epilogue;
  bool has_ext = false;
  // Stencil<NDIM>::create's is_pointer_axis / is_soarr, when given: the source layout it serves
  bool layout_fixed = false;
  std::array<uint8_t, NCA_MAX_NDIM> pointer_axes{};
  bool is_soarr = false;
  uint64_t forward = 0;  // for adjoints: the forward stencil
  std::mutex mu;
  std::unordered_map<std::string, hipFunction_t> kernels;
  std::unordered_map<std::string, std::string> sources;
};

namespace {
std::mutex g_st_mu;
std::unordered_map<uint64_t, std::unique_ptr<StencilProgram>> g_stencils;
uint64_t g_next_stencil = 1ull << 40;
}  // namespace

static StencilProgram& stencil_program(uint64_t h) {
  std::lock_guard<std::mutex> lk(g_st_mu);
  auto it = g_stencils.find(h);
  NCA_REQUIRE(it != g_stencils.end(), "unknown stencil handle " + std::to_string(h));
  return *it->second;
}

static uint64_t stencil_register(std::unique_ptr<StencilProgram> p) {
  std::lock_guard<std::mutex> lk(g_st_mu);
  uint64_t h = g_next_stencil++;
  g_stencils[h] = std::move(p);
  return h;
}

// The reference's stencil source evaluates the tape with plain C++ operators in work_t
// (rtcompiler.hh:204-234): division has no zero-divisor rule here.
static Tape stencil_tape(const nca_expr_desc& d) {
  Tape t = Tape::from_desc(d);
  t.eval_dtype = t.work_dtype;
  t.compute_dtype = compute_dtype(t.work_dtype);
  for (auto& n : t.nodes)
    if (n.op == NCA_OP_DIV) n.op = NODE_DIVRAW;
  return t;
}

// elements per thread along x in the march kernel: one 16-byte load of the source
static int march_vx(int32_t src_dtype) { return itemsize(src_dtype) == 2 ? 8 : 4; }

struct MarchGeom {
  int rz = 0, ry = 0, rx = 0;
};

static MarchGeom geom_of(const StencilProgram& p) {
  MarchGeom g;
  for (auto& o : p.offsets) {
    g.rz = std::max<int>(g.rz, static_cast<int>(std::llabs(o[0])));
    g.ry = std::max<int>(g.ry, static_cast<int>(std::llabs(o[1])));
    g.rx = std::max<int>(g.rx, static_cast<int>(std::llabs(o[2])));
  }
  return g;
}

static std::vector<std::string> names_u(int n) {
  std::vector<std::string> v;
  for (int k = 0; k < n; ++k) v.push_back("u" + std::to_string(k));
  return v;
}
static std::vector<std::string> names_s(int n) {
  std::vector<std::string> v;
  for (int k = 0; k < n; ++k) v.push_back("s" + std::to_string(k));
  return v;
}

struct StencilFuse {
  bool axis0 = false;
  int fuse = 0;  // NCA_FUSE_* of global reductions
  int mode = NCA_ACCUM_PLAIN;
  bool with_dst = true;
  int32_t axis0_dtype = NCA_FLOAT32;
  // the register-queue march over plane tables: src (dst) points to a table of plane pointers
  // (an NC pointer axis 0 or an SO view indirect at axis 0) whose planes are C-contiguous
  bool src_tab = false;
  bool dst_tab = false;
};

// Kernel symbol of a march variant: storage types and fused reductions are part of the name, so
// kernel traces and the library's profiler tell the variants apart.
static std::string march_name(const StencilProgram& p, const StencilFuse& fz, int32_t dst_dtype) {
  std::string n = std::string("nca_stencil_march_") + dtype_info(p.src_dtype).name;
  if (dst_dtype != p.src_dtype) n += std::string("_") + dtype_info(dst_dtype).name;
  if (fz.axis0) n += "_axis0";
  if (fz.fuse & NCA_FUSE_SUM) n += "_sum";
  if (fz.fuse & NCA_FUSE_MAX) n += "_max";
  if (fz.fuse & NCA_FUSE_MIN) n += "_min";
  if (!fz.with_dst) n += "_nodst";
  if (!p.extra_dtypes.empty()) n += "_x" + std::to_string(p.extra_dtypes.size());
  if (fz.src_tab || fz.dst_tab) n += std::string("_planes") + (fz.src_tab ? "s" : "") + (fz.dst_tab ? "d" : "");
  return n;
}

// ---- march kernel ---------------------------------------------------------------------------------
static std::string gen_march(const StencilProgram& p, const StencilFuse& fz, int32_t dst_dtype) {
  const Tape& t = p.tape;
  const MarchGeom g = geom_of(p);
  const std::string C = ctype_of(t.compute_dtype);
  const std::string S = ctype_of(p.src_dtype);
  const std::string D = ctype_of(dst_dtype);
  const int n_arr = static_cast<int>(t.arr_dtypes.size());
  std::vector<int> alias(n_arr);
  for (int k = 0; k < n_arr; ++k) alias[k] = k;
  std::vector<std::string> roots;
  std::string body = emit_nodes(t, {t.root}, alias, names_u(n_arr), names_s(static_cast<int>(t.scalar_dtypes.size())),
                                "lin", "t", roots);
  const int32_t rcd = compute_dtype(dst_dtype);
  const std::string RC = ctype_of(rcd);
  const bool comp = fz.mode == NCA_ACCUM_COMPENSATED;
  std::ostringstream os;
  os << "#define RZ " << g.rz << "\n#define RY " << g.ry << "\n#define RX " << g.rx << "\n"
     << "#define VX " << march_vx(p.src_dtype) << "\n#define TX (16 * VX)\n#define TY 16\n#define NS (2 * RZ + 2)\n#define LW (TX + 2 * RX)\n#define LH (TY + 2 * RY)\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void " << march_name(p, fz, dst_dtype) << "(const " << S << "* __restrict__ src";
  for (size_t j = 0; j < p.extra_dtypes.size(); ++j) os << ", const " << ctype_of(p.extra_dtypes[j]) << "* __restrict__ e" << j;
  os << ", " << D << "* __restrict__ dst, nca_i64 NZ, nca_i64 NY, nca_i64 NX, int ntx, int nty, int zlen"
     << ", double* __restrict__ colpart, void* __restrict__ colout, void* __restrict__ psum, void* __restrict__ pmax"
     << ", void* __restrict__ pmin";
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  os << ") {\n";
  os << "  __shared__ " << C << " lds[NS][LH][LW];\n";
  // XCD-aware tile order: consecutive tiles on one XCD (blocks are dealt round-robin)
  os << "  const unsigned nb = gridDim.x, b = blockIdx.x;\n"
     << "  const unsigned tix = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;\n"
     << "  const int txi = tix % ntx, tyi = (tix / ntx) % nty, tzi = tix / (ntx * nty);\n"
     << "  const nca_i64 x0 = (nca_i64)txi * TX, y0 = (nca_i64)tyi * TY, z0 = (nca_i64)tzi * zlen;\n"
     << "  const nca_i64 z1 = (z0 + zlen < NZ) ? z0 + zlen : NZ;\n"
     << "  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;\n"
     << "  const nca_i64 gx = x0 + VX * tx, gy = y0 + ty;\n"
     << "  const bool valid = (gx < NX) && (gy < NY);\n"
     << "  const nca_i64 plane = NY * NX;\n  const int NZi = (int)NZ;\n";
  // per-thread staging registers: two planes in flight (sets A and B)
  os << "  " << C << " riA[VX], rhA[VX], rcA, riB[VX], rhB[VX], rcB;\n";
  os << "  const int hid = threadIdx.x;\n"
     << "  const int nhr = 2 * RY * 16;\n"  // halo-row vectors
     << "  const int nhc = 2 * RX * LH;\n";  // halo-column scalars
  // per-thread constants of the plane loads: 32-bit offsets within a plane (the host
  // requires NY*NX < 2^31) and the in-range tests that do not depend on z
  os << "  const unsigned ioff = (unsigned)(gy * NX + gx);\n"
     << "  const bool ifull = (gy < NY) && (gx + VX - 1 < NX);\n"
     << "  bool hrin = false, hrfull = false;\n  unsigned hroff = 0u;\n  nca_i64 hrx = 0;\n"
     << "  if (hid < nhr) {\n"
     << "    const int hr = hid >> 4, hx = hid & 15;\n"
     << "    const nca_i64 yy = y0 + ((hr < RY) ? (hr - RY) : (TY + hr - RY));\n"
     << "    hrx = x0 + VX * hx;\n"
     << "    hrin = (yy >= 0) && (yy < NY);\n"
     << "    hrfull = hrin && (hrx + VX - 1 < NX);\n"
     << "    hroff = hrin ? (unsigned)(yy * NX + hrx) : 0u;\n"
     << "  }\n"
     << "  bool hcin = false;\n  unsigned hcoff = 0u;\n"
     << "  if (hid >= nhr && hid < nhr + nhc) {\n"
     << "    const int q = hid - nhr, hc = q / LH, hrow = q - hc * LH;\n"
     << "    const nca_i64 xx = x0 + ((hc < RX) ? (hc - RX) : (TX + hc - RX));\n"
     << "    const nca_i64 yy = y0 + hrow - RY;\n"
     << "    hcin = (yy >= 0) && (yy < NY) && (xx >= 0) && (xx < NX);\n"
     << "    hcoff = hcin ? (unsigned)(yy * NX + xx) : 0u;\n"
     << "  }\n";
  // load plane z into registers (zero outside the array)
  // Loads are branch-free where possible: an out-of-range load reads a safe address
  // (the first source element) and its value is replaced by zero.  Rows are all-or-nothing
  // because NX is a multiple of VX.
  os << "  auto gload = [&](int z, " << C << " (&ri)[VX], " << C << " (&rh)[VX], " << C << "& rc) {\n"
     << "    const bool zin = (unsigned)z < (unsigned)NZi;\n"
     << "    const " << S << "* __restrict__ pl = src + (size_t)(zin ? z : 0) * plane;\n"
     << "    {\n"
     << "      const bool ok = zin && ifull;\n"
     << "      const nca_vec<" << S << ", VX> v = *reinterpret_cast<const nca_vec<" << S << ", VX>*>(ok ? pl + ioff : src);\n"
     << "#pragma unroll\n      for (int j = 0; j < VX; ++j) ri[j] = ok ? nca_cvt<" << C << ", " << S << ">::from(v.v[j]) : (" << C << ")0;\n"
     << "    }\n"
     << "    if (hid < nhr) {\n"
     << "      const bool ok = zin && hrfull;\n"
     << "      const nca_vec<" << S << ", VX> v = *reinterpret_cast<const nca_vec<" << S << ", VX>*>(ok ? pl + hroff : src);\n"
     << "#pragma unroll\n      for (int j = 0; j < VX; ++j) rh[j] = ok ? nca_cvt<" << C << ", " << S << ">::from(v.v[j]) : (" << C << ")0;\n"
     << "    }\n"
     << "    if (hid >= nhr && hid < nhr + nhc) {\n"
     << "      const bool ok = zin && hcin;\n"
     << "      const " << C << " v = nca_ld<" << C << ">(ok ? pl + hcoff : src);\n"
     << "      rc = ok ? v : (" << C << ")0;\n"
     << "    }\n"
     << "  };\n";
  os << "  auto lstore = [&](int slot, const " << C << " (&ri)[VX], const " << C << " (&rh)[VX], const " << C << "& rc) {\n"
     << "#pragma unroll\n    for (int j = 0; j < VX; ++j) lds[slot][ty + RY][RX + VX * tx + j] = ri[j];\n"
     << "    if (hid < nhr) {\n"
     << "      const int hr = hid >> 4, hx = hid & 15;\n"
     << "      const int row = (hr < RY) ? hr : (TY + hr);\n"
     << "#pragma unroll\n      for (int j = 0; j < VX; ++j) lds[slot][row][RX + VX * hx + j] = rh[j];\n"
     << "    }\n"
     << "    if (hid >= nhr && hid < nhr + nhc) {\n"
     << "      const int q = hid - nhr, hc = q / LH, hrow = q - hc * LH;\n"
     << "      const int col = (hc < RX) ? hc : (TX + hc);\n"
     << "      lds[slot][hrow][col] = rc;\n"
     << "    }\n"
     << "  };\n";
  // ring slots: plane z0 - RZ + k lives in slot k % NS; slots advance by one per plane
  os << "  auto wrap = [](int s) { return s < 0 ? s + NS : (s >= NS ? s - NS : s); };\n";
  // accumulators
  if (fz.axis0) {
    os << "  float cs[VX] = {};\n";
    if (comp) os << "  float cc[VX] = {};\n";
  }
  if (fz.fuse & NCA_FUSE_SUM) {
    os << "  " << RC << " acc_s = 0;\n";
    if (comp) os << "  " << RC << " acc_c = 0;\n";
  }
  if (fz.fuse & NCA_FUSE_MAX) os << "  " << RC << " acc_mx = nca_ident<" << ctype_of(dst_dtype) << ", " << RC << ">::lowest();\n";
  if (fz.fuse & NCA_FUSE_MIN) os << "  " << RC << " acc_mn = nca_ident<" << ctype_of(dst_dtype) << ", " << RC << ">::max();\n";
  // prologue: planes z0-RZ .. z0+RZ-1 straight to LDS, planes z0+RZ and z0+RZ+1 in flight
  os << "  const int zb = (int)z0, ze = (int)z1;\n";
  os << "#pragma unroll\n  for (int k = 0; k < 2 * RZ; ++k) { gload(zb - RZ + k, riA, rhA, rcA); lstore(k, riA, rhA, rcA); }\n";
  os << "  gload(zb + RZ, riA, rhA, rcA);\n  gload(zb + RZ + 1, riB, rhB, rcB);\n";
  os << "  auto compute = [&](int z, int sz) {\n";
  // extras at the output coordinate
  for (size_t j = 0; j < p.extra_dtypes.size(); ++j) {
    const std::string E = ctype_of(p.extra_dtypes[j]);
    os << "    " << C << " ex" << j << "[VX];\n"
       << "    if (valid) { const nca_vec<" << E << ", VX> v = *reinterpret_cast<const nca_vec<" << E << ", VX>*>(e" << j
       << " + (size_t)z * plane + ioff);\n"
       << "#pragma unroll\n      for (int j = 0; j < VX; ++j) ex" << j << "[j] = nca_cvt<" << C << ", " << E << ">::from(v.v[j]); }\n"
       << "    else {\n#pragma unroll\n      for (int j = 0; j < VX; ++j) ex" << j << "[j] = (" << C << ")0; }\n";
  }
  // slots for each z offset
  std::vector<int64_t> dzs;
  for (auto& o : p.offsets)
    if (std::find(dzs.begin(), dzs.end(), o[0]) == dzs.end()) dzs.push_back(o[0]);
  for (int64_t dz : dzs)
    os << "    const int sl" << (dz < 0 ? "m" : "p") << std::llabs(dz) << " = wrap(sz + (" << dz << "));\n";
  os << "    nca_vec<" << D << ", VX> r;\n";
  os << "#pragma unroll\n    for (int j = 0; j < VX; ++j) {\n";
  for (int k = 0; k < p.n_off; ++k) {
    const auto& o = p.offsets[k];
    os << "      const " << C << " u" << k << " = lds[sl" << (o[0] < 0 ? "m" : "p") << std::llabs(o[0]) << "][ty + RY + (" << o[1]
       << ")][RX + VX * tx + j + (" << o[2] << ")];\n";
  }
  for (size_t j = 0; j < p.extra_dtypes.size(); ++j)
    os << "      const " << C << " u" << (p.n_off + j) << " = ex" << j << "[j];\n";
  if (t.uses_idx) os << "      const nca_i64 lin = (nca_i64)z * plane + ioff + j;\n";
  os << body;
  os << "      r.v[j] = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n"
     << "    }\n";
  // reductions of the stored values after the element loop (short live ranges)
  if (fz.axis0 || fz.fuse) {
    const int vx = march_vx(p.src_dtype);
    os << "    if (valid) {\n";
    for (int j = 0; j < vx; ++j)
      os << "      const " << RC << " rv" << j << " = nca_cvt<" << RC << ", " << D << ">::from(r.v[" << j << "]);\n";
    if (fz.axis0)
      for (int j = 0; j < vx; ++j) {
        if (comp) os << "      nca_kahan_add(cs[" << j << "], cc[" << j << "], (float)rv" << j << ");\n";
        else os << "      cs[" << j << "] += (float)rv" << j << ";\n";
      }
    // pairwise trees over the vx values
    auto tree = [&](const std::string& f) {
      std::vector<std::string> v;
      for (int j = 0; j < vx; ++j) v.push_back("rv" + std::to_string(j));
      while (v.size() > 1) {
        std::vector<std::string> w;
        for (size_t k = 0; k + 1 < v.size(); k += 2) w.push_back(f + "(" + v[k] + ", " + v[k + 1] + ")");
        if (v.size() % 2) w.push_back(v.back());
        v = w;
      }
      return v[0];
    };
    if (fz.fuse & NCA_FUSE_SUM) {
      os << "      const " << RC << " qs = " << tree("nca_add2") << ";\n";
      if (comp) os << "      nca_kahan_add(acc_s, acc_c, qs);\n";
      else os << "      acc_s += qs;\n";
    }
    if (fz.fuse & NCA_FUSE_MAX) os << "      acc_mx = nca_maxop(acc_mx, " << tree("nca_maxnum") << ");\n";
    if (fz.fuse & NCA_FUSE_MIN) os << "      acc_mn = nca_minop(acc_mn, " << tree("nca_minnum") << ");\n";
    os << "    }\n";
  }
  if (fz.with_dst)
    os << "    if (valid) nca_store_nt(reinterpret_cast<nca_vec<" << D << ", VX>*>(dst + (size_t)z * plane + ioff), r);\n";
  os << "  };\n";
  // one barrier per plane; the register set just stored is refilled two planes ahead
  // sz: slot of plane z (= (z - z0 + RZ) % NS); the plane stored at iteration z is z + RZ
  os << "  int sz = RZ;\n"
     << "  for (int z = zb; z < ze; z += 2) {\n"
     << "    lstore(wrap(sz + RZ), riA, rhA, rcA);\n"
     << "    __syncthreads();\n"
     << "    if (z + 2 < ze) gload(z + 2 + RZ, riA, rhA, rcA);\n"
     << "    compute(z, sz);\n"
     << "    sz = wrap(sz + 1);\n"
     << "    if (z + 1 < ze) {\n"
     << "      lstore(wrap(sz + RZ), riB, rhB, rcB);\n"
     << "      __syncthreads();\n"
     << "      if (z + 3 < ze) gload(z + 3 + RZ, riB, rhB, rcB);\n"
     << "      compute(z + 1, sz);\n"
     << "      sz = wrap(sz + 1);\n"
     << "    }\n"
     << "  }\n";
  // epilogue: column sums (one run per tile: final values; several runs: float64
  // partials per run, added in z order by the combine kernel) and per-block partials of
  // the global reductions (combined in block order by the same kernel).
  if (fz.axis0) {
    const std::string CT = ctype_of(fz.axis0_dtype);
    const std::string cv = comp ? "nca_comp_value(cs[j], cc[j])" : "(double)cs[j]";
    os << "  if (valid) {\n"
       << "    if (colpart) {\n#pragma unroll\n      for (int j = 0; j < VX; ++j) colpart[((nca_i64)tzi * NY + gy) * NX + gx + j] = "
       << cv << ";\n"
       << "    } else {\n#pragma unroll\n      for (int j = 0; j < VX; ++j) nca_st(reinterpret_cast<" << CT
       << "*>(colout) + gy * NX + gx + j, " << cv << ");\n"
       << "    }\n  }\n";
  }
  if (fz.fuse) {
    os << "  __shared__ double red_sd[4];\n  __shared__ " << RC << " red_mx[4];\n  __shared__ " << RC << " red_mn[4];\n"
       << "  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;\n";
    if (fz.fuse & NCA_FUSE_SUM) {
      if (comp) os << "  nca_wave_comp(acc_s, acc_c);\n  if (lane == 0) red_sd[wid] = nca_comp_value(acc_s, acc_c);\n";
      else os << "  acc_s = nca_wave_sum(acc_s);\n  if (lane == 0) red_sd[wid] = (double)acc_s;\n";
    }
    if (fz.fuse & NCA_FUSE_MAX) os << "  acc_mx = nca_wave_max(acc_mx);\n  if (lane == 0) red_mx[wid] = acc_mx;\n";
    if (fz.fuse & NCA_FUSE_MIN) os << "  acc_mn = nca_wave_min(acc_mn);\n  if (lane == 0) red_mn[wid] = acc_mn;\n";
    os << "  __syncthreads();\n  if (threadIdx.x == 0) {\n";
    if (fz.fuse & NCA_FUSE_SUM) os << "    ((double*)psum)[b] = red_sd[0] + red_sd[1] + red_sd[2] + red_sd[3];\n";
    if (fz.fuse & NCA_FUSE_MAX)
      os << "    " << RC << " m = red_mx[0];\n    for (int w = 1; w < 4; ++w) m = nca_maxop(m, red_mx[w]);\n    *(" << RC
         << "*)((char*)pmax + 8 * (size_t)b) = m;\n";
    if (fz.fuse & NCA_FUSE_MIN)
      os << "    " << RC << " m2 = red_mn[0];\n    for (int w = 1; w < 4; ++w) m2 = nca_minop(m2, red_mn[w]);\n    *(" << RC
         << "*)((char*)pmin + 8 * (size_t)b) = m2;\n";
    os << "  }\n";
  }
  os << "}\n";
  return os.str();
}

// ---- march kernel with a register z-queue -----------------------------------------------------------
// For stencils whose z-offsets are pure (dz != 0 only with dy = dx = 0), e.g. the 7-point
// Laplacian: each thread keeps its own columns of planes z-RZ .. z+RZ+1 in registers (a queue
// unrolled so that every slot index is a constant), and LDS holds only the current plane
// (double-buffered, one barrier per plane) for the x/y neighbours.  The LDS interior starts at a
// 16-byte aligned column, so a row of VX values is one vector read; only the elements an
// x-offset shifts across the thread's VX columns are scalar reads, and x-offsets within a row
// are taken from registers.  Each thread computes RPT rows (ty + 16 r).
struct MarchPlan {
  bool queue = false;
  int vx = 4;
  int rpt = 1;
  int txt = 16;  // threads across x (256 / txt thread rows)
  int pd = 1;    // register-queue march: planes loaded ahead of the one being stored
  bool defer = false;  // register-queue march: loads kept raw (zeros read from a zero buffer) and
                       // converted at first use; every wave issues the same loads and stores
};

// NCA_HIP_MARCH_TILE="rows,vx[,tx]" overrides the tile shape of the register-queue march (rows
// per thread 1 or 2; x elements per thread 4 or 8; threads across x 16, 32, 64 or 128), for tuning.
static bool march_override(int& rpt, int& vx, int& txt) {
  const char* e = std::getenv("NCA_HIP_MARCH_TILE");
  if (!e) return false;
  int a = 0, b = 0, c = txt;
  const int n = std::sscanf(e, "%d,%d,%d", &a, &b, &c);
  if (n < 2 || (a != 1 && a != 2) || (b != 4 && b != 8) || (c != 16 && c != 32 && c != 64 && c != 128)) return false;
  rpt = a;
  vx = b;
  txt = c;
  return true;
}

// Tile shape of the register-queue march: 256 elements across and 8 rows (64 threads x 4
// elements x 2 rows for 4-byte sources, 32 x 8 x 1 for 2-byte ones).  Measured on MI355X for the
// 512^3 Laplacian: wide tiles keep the halo-column reads (a 128-byte line per row for one
// element) small against the interior (HBM reads 1.04x the array instead of 1.24x with 64x32
// tiles); narrower arrays get narrower tiles so that threads are not idle.
// nx, ny, nz: the destination's extents (large when unknown); n_cus: the device's CUs.
static MarchPlan march_plan(const StencilProgram& p, int64_t nx = int64_t(1) << 40, int64_t ny = int64_t(1) << 40,
                            int64_t nz = int64_t(1) << 40, int n_cus = 256) {
  MarchPlan m;
  m.vx = march_vx(p.src_dtype);
  if (p.ndim != 3) return m;
  const MarchGeom g = geom_of(p);
  bool pure_z = true;
  for (auto& o : p.offsets)
    if (o[0] != 0 && (o[1] != 0 || o[2] != 0)) pure_z = false;
  auto halo_threads = [&](int rpt) {
    const int tyt = 256 / m.txt;
    return 2 * g.ry * m.txt + 2 * g.rx * (tyt * rpt + 2 * g.ry);
  };
  m.queue = pure_z && g.rx <= 4;
  m.rpt = (m.queue && itemsize(p.src_dtype) >= 4) ? 2 : 1;
  if (m.queue) {
    m.txt = 256 / m.vx;
    while (m.txt > 16 && static_cast<int64_t>(m.txt / 2) * m.vx >= nx) m.txt /= 2;
    while (m.txt > 16 && halo_threads(m.rpt) > 256) m.txt /= 2;
  }
  if (m.queue) {
    int rpt = m.rpt, vx = m.vx, txt = m.txt;
    if (march_override(rpt, vx, txt) && (vx * itemsize(p.src_dtype) == 16 || vx * itemsize(p.src_dtype) == 32)) {
      m.rpt = rpt;
      m.vx = vx;
      m.txt = txt;
    }
  }
  if (halo_threads(m.rpt) > 256) m.rpt = 1;
  if (halo_threads(m.rpt) > 256) m.queue = false;
  if (!m.queue) {
    m.rpt = 1;
    m.vx = march_vx(p.src_dtype);
    m.txt = 16;
  }
  // 2-byte sources use deferred loads: the plane loads stay raw until first used and every
  // wave issues the same loads and stores (zero buffer, sink), so the compiler's wait counts are
  // exact and the prefetched planes stay in flight.  Measured on MI355X, 512^3 BF16 Laplacian
  // with the axis-0 sum and max: 126.6 -> 108.0 us with two planes ahead (the kernel's immediate
  // conversions waited for the newest loads in the halo waves, and the per-plane barrier made
  // the whole block wait with them).  4-byte sources keep immediate conversions: with deferred
  // loads the float32 kernel read HBM with more requests in flight and ran slower (194.3 ->
  // 196.3 us with one plane ahead, 215.4 with two).
  m.defer = m.queue && itemsize(p.src_dtype) == 2;
  // Two planes in flight when the z runs are long (4-byte sources: measured 2-3% faster at
  // 512^3 float32, whose runs are 64 planes; at 256^3 the runs are 8 planes and the extra
  // planes read past each run's end cost more than the deeper prefetch gains; 2-byte sources
  // with deferred loads: as fast as one plane ahead with the same registers).  The run length is
  // estimated with 4 resident blocks per CU.  NCA_HIP_MARCH_PREFETCH = 1 or 2 overrides it.
  if (m.queue && (itemsize(p.src_dtype) == 4 || m.defer) && p.ndim == 3) {
    const int64_t tx = static_cast<int64_t>(m.txt) * m.vx, ty = static_cast<int64_t>(256 / m.txt) * m.rpt;
    const double tiles = static_cast<double>((std::min<int64_t>(nx, int64_t(1) << 20) + tx - 1) / tx) *
                         static_cast<double>((std::min<int64_t>(ny, int64_t(1) << 20) + ty - 1) / ty);
    const double run = static_cast<double>(std::min<int64_t>(nz, int64_t(1) << 20)) * tiles / (4.0 * n_cus);
    m.pd = run >= 32.0 ? 2 : 1;
  }
  if (const char* e = std::getenv("NCA_HIP_MARCH_PREFETCH")) {
    const int v = std::atoi(e);
    if (m.queue && (v == 1 || v == 2)) m.pd = v;
  }
  return m;
}

// Backward of loss = sum(expr(stencil(src), extras)) in one march pass (vjp mode of the queue
// march): the stencil value a is computed as in the forward, then the expression's symbolic
// derivatives with respect to a and to the extras are evaluated at a and stored, each scaled
// by the upstream gradient g read from a device pointer (1 when the pointer is null).
struct MarchVjp {
  Tape expr;                        // the pointwise expression (derivative nodes are appended)
  std::vector<int> operand;         // expression operand k reads: -1 the stencil value, j >= 0 extra j
  std::vector<int32_t> extra_dtypes;
  std::vector<int> roots;           // derivative roots, one per output (-1: exactly zero)
  std::vector<int32_t> out_dtypes;  // output m's dtype (output 0: d/da; then the extras asked for)
  std::vector<int> alias;           // expression operand k -> first operand reading the same value
  int32_t value_dtype = NCA_FLOAT32; // the expression's dest dtype: the type of the values that fused
                                     // reductions (StencilFuse::fuse) sum, in the forward's way
};

static std::string gen_march_q(const StencilProgram& p, const StencilFuse& fz, int32_t dst_dtype, const MarchPlan& mp,
                               const MarchVjp* vj = nullptr) {
  const Tape& t = p.tape;
  const MarchGeom g = geom_of(p);
  const std::string C = ctype_of(t.compute_dtype);
  const std::string S = ctype_of(p.src_dtype);
  const std::string D = ctype_of(dst_dtype);
  const int n_arr = static_cast<int>(t.arr_dtypes.size());
  std::vector<int> alias(n_arr);
  for (int k = 0; k < n_arr; ++k) alias[k] = k;
  std::vector<std::string> roots;
  std::string body = emit_nodes(t, {t.root}, alias, names_u(n_arr), names_s(static_cast<int>(t.scalar_dtypes.size())),
                                "lin", "t", roots);
  // values that fused reductions read: the stencil's output (dst dtype), or in vjp mode the
  // expression's own value (its dest dtype), held in res / resp
  const int32_t vdt = vj ? vj->value_dtype : dst_dtype;
  const std::string VD = ctype_of(vdt);
  const int32_t rcd = compute_dtype(vdt);
  const std::string RC = ctype_of(rcd);
  const bool comp = fz.mode == NCA_ACCUM_COMPENSATED;
  // vjp mode: the expression's derivative code, its loads renamed to the stencil value / extras
  const std::vector<int32_t>& xdt = vj ? vj->extra_dtypes : p.extra_dtypes;
  std::string vbody, pname;
  std::vector<std::string> vnames;
  if (vj) {
    std::vector<std::string> en(vj->operand.size()), rn(vj->expr.scalar_dtypes.size());
    for (size_t k = 0; k < en.size(); ++k)
      en[k] = vj->operand[k] < 0 ? roots[0] : "u" + std::to_string(p.n_off + vj->operand[k]);
    for (size_t i = 0; i < rn.size(); ++i) rn[i] = "r" + std::to_string(i);
    std::vector<int> live;
    for (int r : vj->roots)
      if (r >= 0) live.push_back(r);
    if (fz.fuse) live.push_back(vj->expr.root);
    std::vector<std::string> ln;
    vbody = emit_nodes(vj->expr, live, vj->alias, en, rn, "lin", "q", ln);
    size_t li = 0;
    for (int r : vj->roots) vnames.push_back(r >= 0 ? ln[li++] : "(" + C + ")0");
    if (fz.fuse) pname = ln[li];
  }
  const std::string rsrc = vj ? "resp" : "res";
  const int VX = mp.vx, RPT = mp.rpt, RZ = g.rz, RY = g.ry, RX = g.rx;
  // register queue: planes z-RZ .. z+RZ and PD planes loading ahead
  const int PD = mp.pd;
  const int Q = 2 * RZ + 1 + PD;
  const int TXT = mp.txt, TYT = 256 / mp.txt;
  const int TX = TXT * VX, TY = TYT * RPT, LH = TY + 2 * RY, PADL = 4, LW = PADL + TX + 4;
  auto I = [](long long v) { return std::to_string(v); };
  // deferred loads: every global load of a plane lands in a raw queue slot (out-of-range rows and
  // columns read zeros from a zero buffer instead of being zeroed after the load) and is
  // converted where it is first used, steps later, so no instruction waits for a load in the step
  // that issued it (a use right after a load, in a branch of its own, waits for every load in
  // flight, the prefetched planes included)
  const bool dfr = mp.defer;
  std::ostringstream os;
  os << "#define VX " << VX << "\n#define TX " << TX << "\n#define TY " << TY << "\n#define LH " << LH << "\n#define LW " << LW
     << "\n#define PADL " << PADL << "\n#define RY " << RY << "\n#define RX " << RX << "\n";
  const std::string kname = vj ? std::string("nca_stencil_march_vjp_") + dtype_info(p.src_dtype).name + (fz.fuse ? "_sum" : "")
                               : march_name(p, fz, dst_dtype);
  os << "extern \"C\" __global__ __launch_bounds__(256) void " << kname << "(const " << S << "* __restrict__ src";
  for (size_t j = 0; j < xdt.size(); ++j) os << ", const " << ctype_of(xdt[j]) << "* __restrict__ e" << j;
  if (vj) {
    for (size_t m = 0; m < vj->out_dtypes.size(); ++m) os << ", " << ctype_of(vj->out_dtypes[m]) << "* __restrict__ o" << m;
    os << ", nca_i64 NZ, nca_i64 NY, nca_i64 NX, int ntx, int nty, int zlen";
    if (fz.fuse) os << ", void* __restrict__ psum, void* __restrict__ pmax, void* __restrict__ pmin";
  } else {
    os << ", " << D << "* __restrict__ dst, nca_i64 NZ, nca_i64 NY, nca_i64 NX, int ntx, int nty, int zlen"
       << ", double* __restrict__ colpart, void* __restrict__ colout, void* __restrict__ psum, void* __restrict__ pmax"
       << ", void* __restrict__ pmin";
    // plane tables: plane z of src starts at *(src + z * stab) + ssub (bytes), of dst likewise
    if (fz.src_tab || fz.dst_tab) os << ", nca_i64 stab, nca_i64 ssub, nca_i64 dtab, nca_i64 dsub";
  }
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  if (vj) {
    for (size_t i = 0; i < vj->expr.scalar_dtypes.size(); ++i) os << ", " << C << " r" << i;
    os << ", const " << C << "* __restrict__ gin";
  }
  os << ") {\n";
  if (vj) os << "  const " << C << " gsc = gin ? *gin : (" << C << ")1;\n";
  const bool stab = !vj && fz.src_tab, dtab = !vj && fz.dst_tab;
  if (stab)
    os << "  auto splane = [&](int pz) { return reinterpret_cast<const " << S
       << "*>(nca_chase(reinterpret_cast<const char*>(src) + (nca_i64)pz * stab, ssub)); };\n";
  if (dtab)
    os << "  auto dplane = [&](int pz) { return reinterpret_cast<" << D
       << "*>(nca_chase(reinterpret_cast<char*>(dst) + (nca_i64)pz * dtab, dsub)); };\n";
  // LDS rows are 16-byte aligned (not aligned to a whole VX vector when VX * sizeof(C) > 16)
  os << "  struct __attribute__((aligned(16))) lvec { " << C << " v[VX]; };\n";
  os << "  __shared__ __attribute__((aligned(16))) " << C << " lds[2][LH][LW];\n";
  os << "  const unsigned nb = gridDim.x, b = blockIdx.x;\n"
     << "  const unsigned tix = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;\n"
     << "  const int txi = tix % ntx, tyi = (tix / ntx) % nty, tzi = tix / (ntx * nty);\n"
     << "  const nca_i64 x0 = (nca_i64)txi * TX, y0 = (nca_i64)tyi * TY, z0 = (nca_i64)tzi * zlen;\n"
     << "  const nca_i64 z1 = (z0 + zlen < NZ) ? z0 + zlen : NZ;\n"
     << "  const int tx = threadIdx.x % " << TXT << ", ty = threadIdx.x / " << TXT << ";\n"
     << "  const nca_i64 gx = x0 + VX * tx;\n"
     << "  const size_t plane = (size_t)NY * (size_t)NX;\n  const int NZi = (int)NZ;\n"
     << "  const int hid = threadIdx.x;\n";
  for (int r = 0; r < RPT; ++r)
    os << "  const nca_i64 gy" << r << " = y0 + ty + " << TYT * r << ";\n"
       << "  const bool valid" << r << " = (gx < NX) && (gy" << r << " < NY);\n"
       << "  const unsigned ioff" << r << " = valid" << r << " ? (unsigned)(gy" << r << " * NX + gx) : 0u;\n";
  const int nhr = 2 * RY * TXT, nhc = 2 * RX * LH;
  // halo roles: halo-row vectors (threads [0, nhr)) and halo-column scalars ([nhr, nhr + nhc))
  os << "  bool hrok = false;\n  unsigned hroff = 0u;\n  int hrrow = 0, hrcol = 0;\n";
  if (nhr > 0)
    os << "  if (hid < " << nhr << ") {\n"
       << "    const int hr = hid / " << TXT << ", hx = hid % " << TXT << ";\n"
       << "    const nca_i64 yy = y0 + ((hr < RY) ? (hr - RY) : (TY + hr - RY));\n"
       << "    const nca_i64 xx = x0 + VX * hx;\n"
       << "    hrok = (yy >= 0) && (yy < NY) && (xx < NX);\n"
       << "    hroff = hrok ? (unsigned)(yy * NX + xx) : 0u;\n"
       << "    hrrow = (hr < RY) ? hr : (TY + hr);\n    hrcol = PADL + VX * hx;\n"
       << "  }\n";
  os << "  bool hcok = false;\n  unsigned hcoff = 0u;\n  int hcrow = 0, hccol = 0;\n";
  if (nhc > 0)
    os << "  if (hid >= " << nhr << " && hid < " << nhr + nhc << ") {\n"
       << "    const int q = hid - " << nhr << ", hc = q / LH, hrow = q - hc * LH;\n"
       << "    const nca_i64 xx = x0 + ((hc < RX) ? (hc - RX) : (TX + hc - RX));\n"
       << "    const nca_i64 yy = y0 + hrow - RY;\n"
       << "    hcok = (yy >= 0) && (yy < NY) && (xx >= 0) && (xx < NX);\n"
       << "    hcoff = hcok ? (unsigned)(yy * NX + xx) : 0u;\n"
       << "    hcrow = hrow;\n    hccol = (hc < RX) ? (PADL - RX + hc) : (PADL + TX + hc - RX);\n"
       << "  }\n";
  os << "  " << C << " cq[" << Q << "][" << RPT << "][VX];\n";
  if (dfr) os << "  nca_vec<" << S << ", VX> rq[" << Q << "][" << RPT << "];\n";
  if (nhr > 0) {
    if (dfr) os << "  nca_vec<" << S << ", VX> hq[" << Q << "];\n";
    else os << "  " << C << " hq[" << Q << "][VX];\n";
  }
  if (nhc > 0) os << "  " << (dfr ? S : C) << " hcq[" << Q << "];\n";
  // accumulators
  if (fz.axis0) {
    os << "  float cs[" << RPT << "][VX] = {};\n";
    if (comp) os << "  float cc[" << RPT << "][VX] = {};\n";
  }
  if (fz.fuse & NCA_FUSE_SUM) {
    os << "  " << RC << " acc_s = 0;\n";
    if (comp) os << "  " << RC << " acc_c = 0;\n";
  }
  if (fz.fuse & NCA_FUSE_MAX) os << "  " << RC << " acc_mx = nca_ident<" << VD << ", " << RC << ">::lowest();\n";
  if (fz.fuse & NCA_FUSE_MIN) os << "  " << RC << " acc_mn = nca_ident<" << VD << ", " << RC << ">::max();\n";
  const std::string zero = "(" + C + ")0";
  // plane pz (an int expression) into register slot q; the plane starts at element offset off
  // one row vector of VX source values into dst[0..VX) (compute type), zero when !ok.  2-byte
  // storage types are read as four 32-bit words (zeroed per word) and unpacked by shifts.
  const bool two_byte = itemsize(p.src_dtype) == 2;
  // the address a lane outside the array reads (its value is discarded): the source, or the zero
  // buffer when src is a table of plane pointers
  const std::string fbk = stab ? "reinterpret_cast<const " + S + "*>(nca_zero_words)" : std::string("src");
  auto load_row = [&](const std::string& ok, const std::string& addr, const std::string& dst) {
    if (two_byte) {
      const bool bf = p.src_dtype == NCA_BFLOAT16;
      os << "        nca_u32x4 w = *reinterpret_cast<const nca_u32x4*>(" << ok << " ? " << addr << " : " << fbk << ");\n"
         << "        w = " << ok << " ? w : (nca_u32x4)(0u);\n";
      for (int i = 0; i < VX / 2; ++i) {
        if (bf) {
          os << "        " << dst << "[" << 2 * i << "] = (" << C << ")__uint_as_float(w[" << i << "] << 16);\n"
             << "        " << dst << "[" << 2 * i + 1 << "] = (" << C << ")__uint_as_float(w[" << i << "] & 0xffff0000u);\n";
        } else {
          os << "        " << dst << "[" << 2 * i << "] = (" << C << ")nca_f16_to_f32((unsigned short)(w[" << i << "] & 0xffffu));\n"
             << "        " << dst << "[" << 2 * i + 1 << "] = (" << C << ")nca_f16_to_f32((unsigned short)(w[" << i << "] >> 16));\n";
        }
      }
    } else {
      os << "        const nca_vec<" << S << ", VX> v = *reinterpret_cast<const nca_vec<" << S << ", VX>*>(" << ok << " ? " << addr
         << " : " << fbk << ");\n";
      for (int j = 0; j < VX; ++j)
        os << "        " << dst << "[" << j << "] = " << ok << " ? nca_cvt<" << C << ", " << S << ">::from(v.v[" << j << "]) : "
           << zero << ";\n";
    }
  };
  // a raw row vector (nca_vec<S, VX>) `v` converted into dst[0..VX) (compute type)
  auto conv_row = [&](const std::string& v, const std::string& dst) {
    if (two_byte) {
      os << "        { nca_u32x" << VX / 2 << " rw;\n        __builtin_memcpy(&rw, &" << v << ", sizeof(rw));\n";
      for (int i = 0; i < VX / 2; ++i) {
        if (p.src_dtype == NCA_BFLOAT16)
          os << "        " << dst << "[" << 2 * i << "] = (" << C << ")__uint_as_float(rw[" << i << "] << 16);\n"
             << "        " << dst << "[" << 2 * i + 1 << "] = (" << C << ")__uint_as_float(rw[" << i << "] & 0xffff0000u);\n";
        else
          os << "        " << dst << "[" << 2 * i << "] = (" << C << ")nca_f16_to_f32((unsigned short)(rw[" << i << "] & 0xffffu));\n"
             << "        " << dst << "[" << 2 * i + 1 << "] = (" << C << ")nca_f16_to_f32((unsigned short)(rw[" << i << "] >> 16));\n";
      }
      os << "        }\n";
    } else {
      for (int j = 0; j < VX; ++j)
        os << "        " << dst << "[" << j << "] = nca_cvt<" << C << ", " << S << ">::from(" << v << ".v[" << j << "]);\n";
    }
  };
  // raw queue slot q -> converted slot q (deferred loads)
  auto unpack = [&](int q) {
    for (int r = 0; r < RPT; ++r) conv_row("rq[" + I(q) + "][" + I(r) + "]", "cq[" + I(q) + "][" + I(r) + "]");
  };
  const std::string zsrc = "reinterpret_cast<const " + S + "*>(nca_zero_words)";
  // planes past the run's last input plane (ze - 1 + RZ) are not loaded: the steps that would
  // prefetch them read the zero buffer or the array's first vector (cached) instead
  auto gload = [&](int q, const std::string& pz, const std::string& off) {
    os << "    {\n      const bool zin = (unsigned)(" << pz << ") < (unsigned)NZi && (" << pz << ") < ze + " << RZ << ";\n";
    if (stab)
      os << "      const " << S << "* __restrict__ pl = zin ? splane(" << pz << ") : reinterpret_cast<const " << S
         << "*>(nca_zero_words);\n";
    else
      os << "      const " << S << "* __restrict__ pl = src + (zin ? (" << off << ") : (size_t)0);\n";
    if (dfr) {
      const std::string VT = "nca_vec<" + S + ", VX>";
      for (int r = 0; r < RPT; ++r)
        os << "      rq[" << q << "][" << r << "] = *reinterpret_cast<const " << VT << "*>((zin && valid" << r << ") ? pl + ioff"
           << r << " : " << zsrc << ");\n";
      // every wave issues the same loads (lanes without halo work read zeros): the compiler's
      // wait counts are then the same on every path, so no wave waits for a newer load
      if (nhr > 0)
        os << "      hq[" << q << "] = *reinterpret_cast<const " << VT << "*>((zin && hrok) ? pl + hroff : " << zsrc << ");\n";
      if (nhc > 0) os << "      hcq[" << q << "] = *((zin && hcok) ? pl + hcoff : " << zsrc << ");\n";
      os << "    }\n";
      return;
    }
    for (int r = 0; r < RPT; ++r) {
      os << "      { const bool ok = zin && valid" << r << ";\n";
      load_row("ok", "pl + ioff" + I(r), "cq[" + I(q) + "][" + I(r) + "]");
      os << "      }\n";
    }
    if (nhr > 0) {
      os << "      if (hid < " << nhr << ") { const bool ok = zin && hrok;\n";
      load_row("ok", "pl + hroff", "hq[" + I(q) + "]");
      os << "      }\n";
    }
    if (nhc > 0)
      os << "      if (hid >= " << nhr << " && hid < " << nhr + nhc << ") { const bool ok = zin && hcok;\n"
         << "        const " << C << " v = nca_ld<" << C << ">(ok ? pl + hcoff : " << fbk << ");\n"
         << "        hcq[" << q << "] = ok ? v : " << zero << "; }\n";
    os << "    }\n";
  };
  auto lstore = [&](int q, const std::string& buf) {
    for (int r = 0; r < RPT; ++r) {
      os << "    { lvec w;\n";
      for (int j = 0; j < VX; ++j) os << "      w.v[" << j << "] = cq[" << q << "][" << r << "][" << j << "];\n";
      os << "      *reinterpret_cast<lvec*>(&lds[" << buf << "][RY + ty + " << TYT * r << "][PADL + VX * tx]) = w; }\n";
    }
    if (nhr > 0) {
      os << "    if (hid < " << nhr << ") { lvec w;\n";
      if (dfr) {
        conv_row("hq[" + I(q) + "]", "w.v");
      } else {
        for (int j = 0; j < VX; ++j) os << "      w.v[" << j << "] = hq[" << q << "][" << j << "];\n";
      }
      os << "      *reinterpret_cast<lvec*>(&lds[" << buf << "][hrrow][hrcol]) = w; }\n";
    }
    if (nhc > 0)
      os << "    if (hid >= " << nhr << " && hid < " << nhr + nhc << ") lds[" << buf << "][hcrow][hccol] = "
         << (dfr ? "nca_cvt<" + C + ", " + S + ">::from(hcq[" + I(q) + "])" : "hcq[" + I(q) + "]") << ";\n";
  };
  // rows (dy) read from LDS, and how far x-offsets reach left/right of the VX columns
  std::map<int64_t, std::pair<int, int>> reach;  // dy -> (left, right) element counts
  std::set<int64_t> vec_rows;                     // dy != 0 rows read as vectors
  for (auto& o : p.offsets) {
    if (o[0] != 0) continue;
    auto& rr = reach[o[1]];
    if (o[2] < 0) rr.first = std::max<int>(rr.first, static_cast<int>(-o[2]));
    if (o[2] > 0) rr.second = std::max<int>(rr.second, static_cast<int>(o[2]));
    if (o[1] != 0) vec_rows.insert(o[1]);
  }
  auto nm = [](int64_t v) { return v < 0 ? "m" + std::to_string(-v) : std::to_string(v); };
  auto compute = [&](int k, const std::string& buf) {
    // slot of plane z + k + dz
    auto slot = [&](int64_t dz) { return static_cast<int>(((k + RZ + dz) % Q + Q) % Q); };
    // where plane z + k of the destination starts: dst + soff, or the plane table's pointer
    const std::string dbase = dtab ? "dpl" + I(k) : std::string("dst + soff");
    if (dtab && fz.with_dst) os << "    " << D << "* const dpl" << k << " = dplane(z + " << k << ");\n";
    for (int r = 0; r < RPT; ++r) {
      os << "    {\n";
      for (int64_t dy : vec_rows)
        os << "      const lvec V" << nm(dy) << " = *reinterpret_cast<const lvec*>(&lds[" << buf << "][RY + ty + " << TYT * r
           << " + (" << dy << ")][PADL + VX * tx]);\n";
      for (auto& kv : reach) {
        for (int m = 1; m <= kv.second.first; ++m)
          os << "      const " << C << " L" << nm(kv.first) << "_" << m << " = lds[" << buf << "][RY + ty + " << TYT * r << " + ("
             << kv.first << ")][PADL + VX * tx - " << m << "];\n";
        for (int m = 1; m <= kv.second.second; ++m)
          os << "      const " << C << " R" << nm(kv.first) << "_" << m << " = lds[" << buf << "][RY + ty + " << TYT * r << " + ("
             << kv.first << ")][PADL + VX * tx + VX - 1 + " << m << "];\n";
      }
      for (size_t jx = 0; jx < xdt.size(); ++jx) {
        const std::string E = ctype_of(xdt[jx]);
        if (dfr)
          os << "      const nca_vec<" << E << ", VX> X" << jx << " = *reinterpret_cast<const nca_vec<" << E << ", VX>*>(valid" << r
             << " ? e" << jx << " + soff + ioff" << r << " : reinterpret_cast<const " << E << "*>(nca_zero_words));\n";
        else
          os << "      nca_vec<" << E << ", VX> X" << jx << " = {};\n"
             << "      if (valid" << r << ") X" << jx << " = *reinterpret_cast<const nca_vec<" << E << ", VX>*>(e" << jx
             << " + soff + ioff" << r << ");\n";
      }
      if (vj) {
        for (size_t m = 0; m < vj->out_dtypes.size(); ++m)
          os << "      nca_vec<" << ctype_of(vj->out_dtypes[m]) << ", VX> res" << m << ";\n";
        if (fz.fuse) os << "      nca_vec<" << VD << ", VX> resp;\n";
      } else
        os << "      nca_vec<" << D << ", VX> res;\n";
      for (int j = 0; j < VX; ++j) {
        os << "      {\n";
        for (int kk = 0; kk < p.n_off; ++kk) {
          const auto& o = p.offsets[kk];
          std::string v;
          if (o[0] != 0) {
            v = "cq[" + I(slot(o[0])) + "][" + I(r) + "][" + I(j) + "]";
          } else {
            const int64_t idx = j + o[2];
            if (idx >= 0 && idx < VX) v = o[1] == 0 ? "cq[" + I(slot(0)) + "][" + I(r) + "][" + I(idx) + "]"
                                                    : "V" + nm(o[1]) + ".v[" + I(idx) + "]";
            else if (idx < 0) v = "L" + nm(o[1]) + "_" + I(-idx);
            else v = "R" + nm(o[1]) + "_" + I(idx - VX + 1);
          }
          os << "        const " << C << " u" << kk << " = " << v << ";\n";
        }
        for (size_t jx = 0; jx < xdt.size(); ++jx)
          os << "        const " << C << " u" << (p.n_off + jx) << " = nca_cvt<" << C << ", " << ctype_of(xdt[jx])
             << ">::from(X" << jx << ".v[" << j << "]);\n";
        if (t.uses_idx || (vj && vj->expr.uses_idx))
          os << "        const nca_i64 lin = (nca_i64)soff + ioff" << r << " + " << j << ";\n";
        os << body;
        if (vj) {
          os << vbody;
          for (size_t m = 0; m < vj->out_dtypes.size(); ++m)
            os << "        res" << m << ".v[" << j << "] = nca_store_cvt<" << ctype_of(vj->out_dtypes[m]) << ", " << C
               << ">::to((" << C << ")(gsc * " << vnames[m] << "));\n";
          if (fz.fuse) os << "        resp.v[" << j << "] = nca_store_cvt<" << VD << ", " << C << ">::to(" << pname << ");\n";
        } else {
          os << "        res.v[" << j << "] = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
        }
        os << "      }\n";
      }
      if (vj)
        for (size_t m = 0; m < vj->out_dtypes.size(); ++m)
          os << "      if (valid" << r << ") nca_store_nt(reinterpret_cast<nca_vec<" << ctype_of(vj->out_dtypes[m])
             << ", VX>*>(o" << m << " + soff + ioff" << r << "), res" << m << ");\n";
      if (fz.axis0 || fz.fuse) {
        os << "      if (valid" << r << ") {\n";
        for (int j = 0; j < VX; ++j)
          os << "        const " << RC << " rv" << j << " = nca_cvt<" << RC << ", " << VD << ">::from(" << rsrc << ".v[" << j
             << "]);\n";
        if (fz.axis0)
          for (int j = 0; j < VX; ++j) {
            const std::string rj = "[" + I(r) + "][" + I(j) + "]";
            if (comp) os << "        nca_kahan_add(cs" << rj << ", cc" << rj << ", (float)rv" << j << ");\n";
            else os << "        cs" << rj << " += (float)rv" << j << ";\n";
          }
        auto tree = [&](const std::string& f) {
          std::vector<std::string> v;
          for (int j = 0; j < VX; ++j) v.push_back("rv" + std::to_string(j));
          while (v.size() > 1) {
            std::vector<std::string> w;
            for (size_t q = 0; q + 1 < v.size(); q += 2) w.push_back(f + "(" + v[q] + ", " + v[q + 1] + ")");
            if (v.size() % 2) w.push_back(v.back());
            v = w;
          }
          return v[0];
        };
        if (fz.fuse & NCA_FUSE_SUM) {
          os << "        const " << RC << " qs = " << tree("nca_add2") << ";\n";
          if (comp) os << "        nca_kahan_add(acc_s, acc_c, qs);\n";
          else os << "        acc_s += qs;\n";
        }
        if (fz.fuse & NCA_FUSE_MAX) os << "        acc_mx = nca_maxop(acc_mx, " << tree("nca_maxnum") << ");\n";
        if (fz.fuse & NCA_FUSE_MIN) os << "        acc_mn = nca_minop(acc_mn, " << tree("nca_minnum") << ");\n";
        os << "      }\n";
      }
      // deferred mode: lanes outside the array store into a sink (every wave issues the same stores)
      if (fz.with_dst && dfr)
        os << "      nca_store_nt(reinterpret_cast<nca_vec<" << D << ", VX>*>(valid" << r << " ? " << dbase << " + ioff" << r
           << " : reinterpret_cast<" << D << "*>(nca_sink_words)), res);\n";
      else if (fz.with_dst)
        os << "      if (valid" << r << ") nca_store_nt(reinterpret_cast<nca_vec<" << D << ", VX>*>(" << dbase << " + ioff" << r
           << "), res);\n";
      os << "    }\n";
    }
  };
  // prologue: planes zb-RZ .. zb+RZ+PD-1 into slots 0 .. 2RZ+PD-1
  os << "  const int zb = (int)z0, ze = (int)z1;\n";
  for (int m = 0; m <= 2 * RZ + PD - 1; ++m) {
    const std::string pz = "zb - " + I(RZ) + " + " + I(m);
    gload(m, pz, "(size_t)(" + pz + ") * plane");
  }
  // deferred loads: planes zb-RZ .. zb+RZ-1 converted now, plane z+RZ at the start of step z
  if (dfr)
    for (int m = 0; m < 2 * RZ; ++m) unpack(m);
  // steady state: step k of an unrolled group handles plane z = z' + k; it stores plane z
  // (slot (k + RZ) % Q) into the LDS buffer of its parity, loads plane z + RZ + PD into the
  // slot just freed ((k - 1) mod Q) and computes plane z.  With Q even the parity is k's.
  os << "  size_t loff = (size_t)(zb + " << RZ + PD << ") * plane;\n"
     << "  size_t soff = (size_t)zb * plane;\n"
     << "  for (int z = zb; z < ze; z += " << Q << ") {\n";
  for (int k = 0; k < Q; ++k) {
    if (k > 0) os << "    if (z + " << k << " >= ze) break;\n";
    const std::string pb = Q % 2 == 0 ? I(k & 1) : "((z - zb + " + I(k) + ") & 1)";
    if (dfr) unpack((k + 2 * RZ) % Q);
    lstore((k + RZ) % Q, pb);
    os << "    __syncthreads();\n";
    gload((k + Q - 1) % Q, "z + " + I(k + RZ + PD), "loff");
    os << "    loff += plane;\n";
    compute(k, pb);
    os << "    soff += plane;\n";
  }
  os << "  }\n";
  // epilogue (as in gen_march)
  if (fz.axis0) {
    const std::string CT = ctype_of(fz.axis0_dtype);
    for (int r = 0; r < RPT; ++r) {
      const std::string cv = comp ? "nca_comp_value(cs[" + I(r) + "][j], cc[" + I(r) + "][j])" : "(double)cs[" + I(r) + "][j]";
      os << "  if (valid" << r << ") {\n"
         << "    if (colpart) {\n#pragma unroll\n      for (int j = 0; j < VX; ++j) colpart[((nca_i64)tzi * NY + gy" << r
         << ") * NX + gx + j] = " << cv << ";\n"
         << "    } else {\n#pragma unroll\n      for (int j = 0; j < VX; ++j) nca_st(reinterpret_cast<" << CT
         << "*>(colout) + gy" << r << " * NX + gx + j, " << cv << ");\n"
         << "    }\n  }\n";
    }
  }
  if (fz.fuse) {
    os << "  __shared__ double red_sd[4];\n  __shared__ " << RC << " red_mx[4];\n  __shared__ " << RC << " red_mn[4];\n"
       << "  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;\n";
    if (fz.fuse & NCA_FUSE_SUM) {
      if (comp) os << "  nca_wave_comp(acc_s, acc_c);\n  if (lane == 0) red_sd[wid] = nca_comp_value(acc_s, acc_c);\n";
      else os << "  acc_s = nca_wave_sum(acc_s);\n  if (lane == 0) red_sd[wid] = (double)acc_s;\n";
    }
    if (fz.fuse & NCA_FUSE_MAX) os << "  acc_mx = nca_wave_max(acc_mx);\n  if (lane == 0) red_mx[wid] = acc_mx;\n";
    if (fz.fuse & NCA_FUSE_MIN) os << "  acc_mn = nca_wave_min(acc_mn);\n  if (lane == 0) red_mn[wid] = acc_mn;\n";
    os << "  __syncthreads();\n  if (threadIdx.x == 0) {\n";
    if (fz.fuse & NCA_FUSE_SUM) os << "    ((double*)psum)[b] = red_sd[0] + red_sd[1] + red_sd[2] + red_sd[3];\n";
    if (fz.fuse & NCA_FUSE_MAX)
      os << "    " << RC << " m = red_mx[0];\n    for (int w = 1; w < 4; ++w) m = nca_maxop(m, red_mx[w]);\n    *(" << RC
         << "*)((char*)pmax + 8 * (size_t)b) = m;\n";
    if (fz.fuse & NCA_FUSE_MIN)
      os << "    " << RC << " m2 = red_mn[0];\n    for (int w = 1; w < 4; ++w) m2 = nca_minop(m2, red_mn[w]);\n    *(" << RC
         << "*)((char*)pmin + 8 * (size_t)b) = m2;\n";
    os << "  }\n";
  }
  os << "}\n";
  return os.str();
}

// ---- training-step march -----------------------------------------------------------------------
// One pass of a whole training step for loss = sum(expr(S(x), extras)) with S linear and of
// radius <= 1 on every axis (zero boundary).  A block owns a TX x TY tile over a run of z
// planes.  Per plane z it stages x with a two-point halo in LDS (three planes, the next one
// prefetched into registers), evaluates w = g * d expr / d a on the tile and a one-point ring
// around it (zero outside the array), and from the last three w planes in LDS applies the
// adjoint of S to give dx at plane z - 1.  The extras' gradients and the reductions of the
// expression's value are taken at the tile's own points of its own planes.  x and the extras
// are read once and dx and the extras' gradients written once: nothing else reaches HBM.
static std::string gen_march_train(const StencilProgram& p, const Tape& adj, const MarchVjp& vj, const StencilFuse& fz,
                                   int32_t dx_dtype, int TXs, int TYs) {
  const Tape& t = p.tape;
  const std::string C = ctype_of(t.compute_dtype);
  const std::string S = ctype_of(p.src_dtype);
  const std::string DX = ctype_of(dx_dtype);
  const std::string VD = ctype_of(vj.value_dtype);
  const int32_t rcd = compute_dtype(vj.value_dtype);
  const std::string RC = ctype_of(rcd);
  const bool comp = fz.mode == NCA_ACCUM_COMPENSATED;
  const int n_off = p.n_off;
  const int n_x = static_cast<int>(vj.extra_dtypes.size());
  auto I = [](long long v) { return std::to_string(v); };
  // stencil value a at a ring point, the derivative roots and (with reductions) the value
  std::vector<int> alias(t.arr_dtypes.size());
  for (size_t k = 0; k < alias.size(); ++k) alias[k] = static_cast<int>(k);
  std::vector<std::string> sroots;
  const std::string sbody = emit_nodes(t, {t.root}, alias, names_u(static_cast<int>(t.arr_dtypes.size())),
                                       names_s(static_cast<int>(t.scalar_dtypes.size())), "lin", "t", sroots);
  std::vector<std::string> en(vj.operand.size()), rn(vj.expr.scalar_dtypes.size());
  for (size_t k = 0; k < en.size(); ++k) en[k] = vj.operand[k] < 0 ? sroots[0] : "u" + I(n_off + vj.operand[k]);
  for (size_t i = 0; i < rn.size(); ++i) rn[i] = "r" + I(i);
  std::vector<int> live;
  for (int r : vj.roots)
    if (r >= 0) live.push_back(r);
  if (fz.fuse) live.push_back(vj.expr.root);
  std::vector<std::string> ln;
  const std::string vbody = emit_nodes(vj.expr, live, vj.alias, en, rn, "lin", "q", ln);
  std::vector<std::string> vnames;
  size_t li = 0;
  for (int r : vj.roots) vnames.push_back(r >= 0 ? ln[li++] : "(" + C + ")0");
  const std::string pname = fz.fuse ? ln[li] : "";
  // the adjoint: LOAD k reads w at the negated offset k
  std::vector<std::string> vn;
  for (int k = 0; k < n_off; ++k) vn.push_back("v" + I(k));
  std::vector<int> aalias(adj.arr_dtypes.size());
  for (size_t k = 0; k < aalias.size(); ++k) aalias[k] = static_cast<int>(k);
  std::vector<std::string> aroots;
  const std::string abody = emit_nodes(adj, {adj.root}, aalias, vn, names_s(static_cast<int>(adj.scalar_dtypes.size())),
                                       "lin", "a", aroots);
  const std::string kname = std::string("nca_stencil_train_") + dtype_info(p.src_dtype).name + (fz.fuse ? "_sum" : "");
  std::ostringstream os;
  os << "#define TX " << TXs << "\n#define TY " << TYs << "\n#define XW (TX + 4)\n#define XH (TY + 4)\n"
     << "#define RW (TX + 2)\n#define RH (TY + 2)\n#define NXL ((XW * XH + 255) / 256)\n#define NRL ((RW * RH + 255) / 256)\n"
     << "#define NTL ((TX * TY + 255) / 256)\n";
  os << "__device__ __forceinline__ int nca_mod3(int z) { const int r = z % 3; return r < 0 ? r + 3 : r; }\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void " << kname << "(const " << S << "* __restrict__ src";
  for (int j = 0; j < n_x; ++j) os << ", const " << ctype_of(vj.extra_dtypes[j]) << "* __restrict__ e" << j;
  os << ", " << DX << "* __restrict__ dxo";
  for (size_t m = 1; m < vj.out_dtypes.size(); ++m) os << ", " << ctype_of(vj.out_dtypes[m]) << "* __restrict__ o" << m;
  os << ", nca_i64 NZ, nca_i64 NY, nca_i64 NX, int ntx, int nty, int zlen";
  if (fz.fuse) os << ", void* __restrict__ psum, void* __restrict__ pmax, void* __restrict__ pmin";
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  for (size_t i = 0; i < vj.expr.scalar_dtypes.size(); ++i) os << ", " << C << " r" << i;
  os << ", const " << C << "* __restrict__ gin) {\n";
  os << "  const " << C << " gsc = gin ? *gin : (" << C << ")1;\n";
  // LDS planes, flat: x with a two-point halo, w and the extras on the ring region
  os << "  __shared__ " << C << " xs[3][XH * XW];\n  __shared__ " << C << " ws[3][RH * RW];\n";
  os << "  const unsigned nb = gridDim.x, b = blockIdx.x;\n"
     << "  const unsigned tix = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;\n"
     << "  const int txi = tix % ntx, tyi = (tix / ntx) % nty, tzi = tix / (ntx * nty);\n"
     << "  const int x0 = txi * TX, y0 = tyi * TY, NYi = (int)NY, NXi = (int)NX;\n"
     << "  const int zb = (int)((nca_i64)tzi * zlen), ze = (int)((zb + (nca_i64)zlen < NZ) ? zb + zlen : NZ);\n"
     << "  const size_t plane = (size_t)NY * (size_t)NX;\n  const int NZi = (int)NZ;\n  const int tid = threadIdx.x;\n";
  // per-thread positions, the same on every plane (element offsets within a plane fit in 32
  // bits; -1: outside the array)
  os << "  int xo[NXL];\n#pragma unroll\n  for (int m = 0; m < NXL; ++m) {\n"
     << "    const int i = tid + 256 * m, ly = i / XW, lx = i - ly * XW, gy = y0 - 2 + ly, gx = x0 - 2 + lx;\n"
     << "    xo[m] = (i < XW * XH && gy >= 0 && gy < NYi && gx >= 0 && gx < NXi) ? gy * NXi + gx : -1;\n  }\n";
  os << "  int go[NRL], xb[NRL];\n  unsigned own = 0u;\n#pragma unroll\n  for (int m = 0; m < NRL; ++m) {\n"
     << "    const int i = tid + 256 * m, ly = i / RW, lx = i - ly * RW, gy = y0 - 1 + ly, gx = x0 - 1 + lx;\n"
     << "    const bool in = i < RW * RH && gy >= 0 && gy < NYi && gx >= 0 && gx < NXi;\n"
     << "    go[m] = in ? gy * NXi + gx : -1;\n    xb[m] = (ly + 1) * XW + lx + 1;\n"
     << "    if (in && ly >= 1 && ly <= TY && lx >= 1 && lx <= TX) own |= 1u << m;\n  }\n";
  os << "  int to[NTL], tb[NTL];\n#pragma unroll\n  for (int m = 0; m < NTL; ++m) {\n"
     << "    const int i = tid + 256 * m, ty = i / TX, tx = i - ty * TX, gy = y0 + ty, gx = x0 + tx;\n"
     << "    to[m] = (i < TX * TY && gy < NYi && gx < NXi) ? gy * NXi + gx : -1;\n    tb[m] = (ty + 1) * RW + tx + 1;\n  }\n";
  os << "  " << C << " xr[NXL];\n";
  // the extras at this thread's ring positions: the plane being evaluated (ec) and the next (er)
  for (int j = 0; j < n_x; ++j) os << "  " << C << " er" << j << "[NRL], ec" << j << "[NRL];\n";
  os << "  auto load_x = [&](int z) {\n    const bool zin = (unsigned)z < (unsigned)NZi;\n"
     << "    const " << S << "* __restrict__ pl = src + (zin ? (size_t)z * plane : (size_t)0);\n"
     << "#pragma unroll\n    for (int m = 0; m < NXL; ++m)\n"
     << "      xr[m] = (zin && xo[m] >= 0) ? nca_cvt<" << C << ", " << S << ">::from(pl[xo[m]]) : (" << C << ")0;\n  };\n";
  os << "  auto store_x = [&](int slot) {\n#pragma unroll\n    for (int m = 0; m < NXL; ++m) {\n"
     << "      const int i = tid + 256 * m;\n      if (i < XW * XH) xs[slot][i] = xr[m];\n    }\n  };\n";
  os << "  auto load_e = [&](int z) {\n    const bool zin = (unsigned)z < (unsigned)NZi;\n";
  for (int j = 0; j < n_x; ++j) {
    const std::string E = ctype_of(vj.extra_dtypes[j]);
    os << "    { const " << E << "* __restrict__ pl = e" << j << " + (zin ? (size_t)z * plane : (size_t)0);\n"
       << "#pragma unroll\n      for (int m = 0; m < NRL; ++m)\n"
       << "        er" << j << "[m] = (zin && go[m] >= 0) ? nca_cvt<" << C << ", " << E << ">::from(pl[go[m]]) : (" << C
       << ")0; }\n";
  }
  os << "  };\n";
  os << "  auto store_e = [&]() {\n#pragma unroll\n    for (int m = 0; m < NRL; ++m) {\n";
  for (int j = 0; j < n_x; ++j) os << "      ec" << j << "[m] = er" << j << "[m];\n";
  os << "    }\n  };\n";
  if (fz.fuse & NCA_FUSE_SUM) {
    os << "  " << RC << " acc_s = 0;\n";
    if (comp) os << "  " << RC << " acc_c = 0;\n";
  }
  if (fz.fuse & NCA_FUSE_MAX) os << "  " << RC << " acc_mx = nca_ident<" << VD << ", " << RC << ">::lowest();\n";
  if (fz.fuse & NCA_FUSE_MIN) os << "  " << RC << " acc_mn = nca_ident<" << VD << ", " << RC << ">::max();\n";
  // prologue: x planes zb-2, zb-1 in LDS, x(zb) and the extras' plane zb-1 in registers
  os << "  load_x(zb - 2);\n  store_x(nca_mod3(zb - 2));\n  load_x(zb - 1);\n  store_x(nca_mod3(zb - 1));\n"
     << "  load_x(zb);\n  load_e(zb - 1);\n";
  os << "  for (int zw = zb - 1; zw <= ze; ++zw) {\n"
     << "    store_x(nca_mod3(zw + 1));\n    store_e();\n"
     << "    if (zw + 2 <= ze + 1) load_x(zw + 2);\n    if (zw + 1 <= ze) load_e(zw + 1);\n"
     << "    __syncthreads();\n";
  // w on the ring region
  os << "    const bool own_z = zw >= zb && zw < ze, zin = (unsigned)zw < (unsigned)NZi;\n"
     << "    const int sw = nca_mod3(zw), sm = nca_mod3(zw - 1), sp = nca_mod3(zw + 1);\n"
     << "    const size_t zoff = (size_t)(zw > 0 ? zw : 0) * plane;\n"
     << "#pragma unroll\n    for (int m = 0; m < NRL; ++m) {\n"
     << "      const int i = tid + 256 * m;\n"
     << "      if (i < RW * RH) {\n"
     << "        " << C << " wv = (" << C << ")0;\n"
     << "        if (zin && go[m] >= 0) {\n";
  for (int k = 0; k < n_off; ++k) {
    const auto& o = p.offsets[k];
    const std::string sl = o[0] < 0 ? "sm" : (o[0] > 0 ? "sp" : "sw");
    os << "          const " << C << " u" << k << " = xs[" << sl << "][xb[m] + (" << o[1] * (TXs + 4) + o[2] << ")];\n";
  }
  for (int j = 0; j < n_x; ++j) os << "          const " << C << " u" << (n_off + j) << " = ec" << j << "[m];\n";
  os << sbody << vbody;
  os << "          wv = (" << C << ")(gsc * " << vnames[0] << ");\n";
  const bool own_work = vj.out_dtypes.size() > 1 || fz.fuse;
  if (own_work) {
    os << "          if (own_z && ((own >> m) & 1u)) {\n"
       << "            const size_t off = zoff + (size_t)go[m];\n";
    for (size_t m = 1; m < vj.out_dtypes.size(); ++m)
      os << "            o" << m << "[off] = nca_store_cvt<" << ctype_of(vj.out_dtypes[m]) << ", " << C << ">::to((" << C
         << ")(gsc * " << vnames[m] << "));\n";
    if (fz.fuse) {
      os << "            const " << RC << " rv = nca_cvt<" << RC << ", " << VD << ">::from(nca_store_cvt<" << VD << ", " << C
         << ">::to(" << pname << "));\n";
      if (fz.fuse & NCA_FUSE_SUM) {
        if (comp) os << "            nca_kahan_add(acc_s, acc_c, rv);\n";
        else os << "            acc_s += rv;\n";
      }
      if (fz.fuse & NCA_FUSE_MAX) os << "            acc_mx = nca_maxnum(acc_mx, rv);\n";
      if (fz.fuse & NCA_FUSE_MIN) os << "            acc_mn = nca_minnum(acc_mn, rv);\n";
    }
    os << "          }\n";
  }
  os << "        }\n        ws[sw][i] = wv;\n      }\n    }\n    __syncthreads();\n";
  // dx at plane zw - 1 from w planes zw - 2 .. zw
  os << "    if (zw - 1 >= zb) {\n"
     << "      const int a0 = nca_mod3(zw - 2), a1 = sm, a2 = sw;\n"
     << "      const size_t doff = (size_t)(zw - 1) * plane;\n"
     << "#pragma unroll\n      for (int m = 0; m < NTL; ++m) {\n"
     << "        if (to[m] >= 0) {\n";
  for (int k = 0; k < n_off; ++k) {
    const auto& o = p.offsets[k];  // the adjoint reads w at -o
    const int dz = static_cast<int>(-o[0]);
    const std::string sl = dz < 0 ? "a0" : (dz > 0 ? "a2" : "a1");
    os << "          const " << C << " v" << k << " = ws[" << sl << "][tb[m] + (" << -o[1] * (TXs + 2) - o[2] << ")];\n";
  }
  os << abody;
  os << "          dxo[doff + (size_t)to[m]] = nca_store_cvt<" << DX << ", " << C << ">::to(" << aroots[0] << ");\n";
  os << "        }\n      }\n    }\n  }\n";
  if (fz.fuse) {
    os << "  __shared__ double red_sd[4];\n  __shared__ " << RC << " red_mx[4];\n  __shared__ " << RC << " red_mn[4];\n"
       << "  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;\n";
    if (fz.fuse & NCA_FUSE_SUM) {
      if (comp) os << "  nca_wave_comp(acc_s, acc_c);\n  if (lane == 0) red_sd[wid] = nca_comp_value(acc_s, acc_c);\n";
      else os << "  acc_s = nca_wave_sum(acc_s);\n  if (lane == 0) red_sd[wid] = (double)acc_s;\n";
    }
    if (fz.fuse & NCA_FUSE_MAX) os << "  acc_mx = nca_wave_max(acc_mx);\n  if (lane == 0) red_mx[wid] = acc_mx;\n";
    if (fz.fuse & NCA_FUSE_MIN) os << "  acc_mn = nca_wave_min(acc_mn);\n  if (lane == 0) red_mn[wid] = acc_mn;\n";
    os << "  __syncthreads();\n  if (threadIdx.x == 0) {\n";
    if (fz.fuse & NCA_FUSE_SUM) os << "    ((double*)psum)[b] = red_sd[0] + red_sd[1] + red_sd[2] + red_sd[3];\n";
    if (fz.fuse & NCA_FUSE_MAX)
      os << "    " << RC << " m = red_mx[0];\n    for (int w = 1; w < 4; ++w) m = nca_maxop(m, red_mx[w]);\n    *(" << RC
         << "*)((char*)pmax + 8 * (size_t)b) = m;\n";
    if (fz.fuse & NCA_FUSE_MIN)
      os << "    " << RC << " m2 = red_mn[0];\n    for (int w = 1; w < 4; ++w) m2 = nca_minop(m2, red_mn[w]);\n    *(" << RC
         << "*)((char*)pmin + 8 * (size_t)b) = m2;\n";
    os << "  }\n";
  }
  os << "}\n";
  return os.str();
}

// The training-step march with 4-point vectors (last extent a multiple of 4, arrays aligned to
// whole vectors): each thread holds aligned 4-point vectors of x, the extras, w, dx and the
// gradients.  w is evaluated on the columns x0-4 .. x0+TX+3 (aligned; the ring needs x0-1 ..
// x0+TX) from x staged on x0-8 .. x0+TX+7; neighbours come from one aligned LDS vector per
// (plane, row) plus the scalars left and right of it, as in the queue march.  The arithmetic
// per element is gen_march_train's.
static std::string gen_march_train_vec(const StencilProgram& p, const Tape& adj, const MarchVjp& vj, const StencilFuse& fz,
                                       int32_t dx_dtype, int TXs, int TYs) {
  const Tape& t = p.tape;
  const std::string C = ctype_of(t.compute_dtype);
  const std::string S = ctype_of(p.src_dtype);
  const std::string DX = ctype_of(dx_dtype);
  const std::string VD = ctype_of(vj.value_dtype);
  const int32_t rcd = compute_dtype(vj.value_dtype);
  const std::string RC = ctype_of(rcd);
  const bool comp = fz.mode == NCA_ACCUM_COMPENSATED;
  const int n_off = p.n_off;
  const int n_x = static_cast<int>(vj.extra_dtypes.size());
  auto I = [](long long v) { return std::to_string(v); };
  std::vector<int> alias(t.arr_dtypes.size());
  for (size_t k = 0; k < alias.size(); ++k) alias[k] = static_cast<int>(k);
  std::vector<std::string> sroots;
  const std::string sbody = emit_nodes(t, {t.root}, alias, names_u(static_cast<int>(t.arr_dtypes.size())),
                                       names_s(static_cast<int>(t.scalar_dtypes.size())), "lin", "t", sroots);
  std::vector<std::string> en(vj.operand.size()), rn(vj.expr.scalar_dtypes.size());
  for (size_t k = 0; k < en.size(); ++k) en[k] = vj.operand[k] < 0 ? sroots[0] : "u" + I(n_off + vj.operand[k]);
  for (size_t i = 0; i < rn.size(); ++i) rn[i] = "r" + I(i);
  std::vector<int> live;
  for (int r : vj.roots)
    if (r >= 0) live.push_back(r);
  if (fz.fuse) live.push_back(vj.expr.root);
  std::vector<std::string> ln;
  const std::string vbody = emit_nodes(vj.expr, live, vj.alias, en, rn, "lin", "q", ln);
  std::vector<std::string> vnames;
  size_t li = 0;
  for (int r : vj.roots) vnames.push_back(r >= 0 ? ln[li++] : "(" + C + ")0");
  const std::string pname = fz.fuse ? ln[li] : "";
  std::vector<std::string> vn;
  for (int k = 0; k < n_off; ++k) vn.push_back("v" + I(k));
  std::vector<int> aalias(adj.arr_dtypes.size());
  for (size_t k = 0; k < aalias.size(); ++k) aalias[k] = static_cast<int>(k);
  std::vector<std::string> aroots;
  const std::string abody = emit_nodes(adj, {adj.root}, aalias, vn, names_s(static_cast<int>(adj.scalar_dtypes.size())),
                                       "lin", "a", aroots);
  const int XW = TXs + 16, XH = TYs + 4, RWV = TXs + 8, RH = TYs + 2;
  const int XV = XW / 4, WV = RWV / 4, TV = TXs / 4;
  // w vectors: the tile's own TV x TY first, then the one-point ring around it.  Without an
  // offset that moves in both y and x, the transposed stencil reads no corner of the ring, so
  // the ring is its two rows and two columns of vectors only (64 x 16 tiles: 256 + 64 vectors,
  // one pass of every wave and one more of one wave, instead of a sixth partly filled pass)
  bool diag = false;
  for (auto& o : p.offsets)
    if (o[1] != 0 && o[2] != 0) diag = true;
  const std::string kname =
      std::string("nca_stencil_train_") + dtype_info(p.src_dtype).name + (fz.fuse ? "_sum" : "") + "_v4";
  std::ostringstream os;
  os << "#define TX " << TXs << "\n#define TY " << TYs << "\n#define XW " << XW << "\n#define XH " << XH << "\n#define RWV "
     << RWV << "\n#define RH " << RH << "\n#define XV " << XV << "\n#define WV " << WV << "\n#define TV " << TV << "\n"
     << "#define NWI " << (diag ? "(WV * RH)" : "(TV * TY + 2 * TV + 2 * TY)") << "\n"
     << "#define NXL ((XV * XH + 255) / 256)\n#define NWL ((NWI + 255) / 256)\n#define NTL ((TV * TY + 255) / 256)\n";
  os << "__device__ __forceinline__ int nca_mod3(int z) { const int r = z % 3; return r < 0 ? r + 3 : r; }\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void " << kname << "(const " << S << "* __restrict__ src";
  for (int j = 0; j < n_x; ++j) os << ", const " << ctype_of(vj.extra_dtypes[j]) << "* __restrict__ e" << j;
  os << ", " << DX << "* __restrict__ dxo";
  for (size_t m = 1; m < vj.out_dtypes.size(); ++m) os << ", " << ctype_of(vj.out_dtypes[m]) << "* __restrict__ o" << m;
  os << ", nca_i64 NZ, nca_i64 NY, nca_i64 NX, int ntx, int nty, int zlen";
  if (fz.fuse) os << ", void* __restrict__ psum, void* __restrict__ pmax, void* __restrict__ pmin";
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  for (size_t i = 0; i < vj.expr.scalar_dtypes.size(); ++i) os << ", " << C << " r" << i;
  os << ", const " << C << "* __restrict__ gin) {\n";
  os << "  const " << C << " gsc = gin ? *gin : (" << C << ")1;\n";
  os << "  struct __attribute__((aligned(16))) lvec { " << C << " v[4]; };\n";
  os << "  __shared__ __attribute__((aligned(16))) " << C << " xs[3][XH * XW];\n"
     << "  __shared__ __attribute__((aligned(16))) " << C << " ws[3][RH * RWV];\n";

  os << "  const unsigned nb = gridDim.x, b = blockIdx.x;\n"
     << "  const unsigned tix = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;\n"
     << "  const int txi = tix % ntx, tyi = (tix / ntx) % nty, tzi = tix / (ntx * nty);\n"
     << "  const int x0 = txi * TX, y0 = tyi * TY, NYi = (int)NY, NXi = (int)NX;\n"
     << "  const int zb = (int)((nca_i64)tzi * zlen), ze = (int)((zb + (nca_i64)zlen < NZ) ? zb + zlen : NZ);\n"
     << "  const size_t plane = (size_t)NY * (size_t)NX;\n  const int NZi = (int)NZ;\n  const int tid = threadIdx.x;\n";
  // per-thread vector positions (element offsets in a plane; -1: outside the array, where a
  // whole vector is in or out because NX and the column origins are multiples of 4)
  os << "  int xo[NXL];\n#pragma unroll\n  for (int m = 0; m < NXL; ++m) {\n"
     << "    const int i = tid + 256 * m, r = i / XV, c = i - r * XV, gy = y0 - 2 + r, gx = x0 - 8 + 4 * c;\n"
     << "    xo[m] = (i < XV * XH && gy >= 0 && gy < NYi && gx >= 0 && gx < NXi) ? gy * NXi + gx : -1;\n  }\n";
  os << "  int go[NWL], xb[NWL], wb[NWL];\n  unsigned own = 0u;\n#pragma unroll\n  for (int m = 0; m < NWL; ++m) {\n"
     << "    const int i = tid + 256 * m;\n";
  if (diag) {
    os << "    const int r = i / WV, c = i - r * WV;\n";
  } else {
    os << "    int r, c;\n"
       << "    if (i < TV * TY) { r = 1 + i / TV; c = 1 + i % TV; }\n"
       << "    else if (i < TV * TY + 2 * TV) { const int k = i - TV * TY; r = k < TV ? 0 : TY + 1; c = 1 + (k < TV ? k : k - TV); }\n"
       << "    else { const int k = i - TV * TY - 2 * TV; r = 1 + (k < TY ? k : k - TY); c = k < TY ? 0 : TV + 1; }\n";
  }
  os << "    const int gy = y0 - 1 + r, gx = x0 - 4 + 4 * c;\n"
     << "    const bool in = i < NWI && gy >= 0 && gy < NYi && gx >= 0 && gx < NXi;\n"
     << "    go[m] = in ? gy * NXi + gx : -1;\n    xb[m] = (r + 1) * XW + 4 + 4 * c;\n    wb[m] = r * RWV + 4 * c;\n"
     << "    if (in && r >= 1 && r <= TY && c >= 1 && c <= TV) own |= 1u << m;\n  }\n";
  os << "  int to[NTL], tb[NTL];\n#pragma unroll\n  for (int m = 0; m < NTL; ++m) {\n"
     << "    const int i = tid + 256 * m, r = i / TV, c = i - r * TV, gy = y0 + r, gx = x0 + 4 * c;\n"
     << "    to[m] = (i < TV * TY && gy < NYi && gx < NXi) ? gy * NXi + gx : -1;\n    tb[m] = (r + 1) * RWV + 4 + 4 * c;\n  }\n";
  os << "  nca_vec<" << S << ", 4> xr[NXL];\n";
  // the extras at this thread's w vectors: the plane being evaluated (ec) and the next (er)
  for (int j = 0; j < n_x; ++j)
    os << "  nca_vec<" << ctype_of(vj.extra_dtypes[j]) << ", 4> er" << j << "[NWL], ec" << j << "[NWL];\n";
  os << "  auto load_x = [&](int z) {\n    const bool zin = (unsigned)z < (unsigned)NZi;\n"
     << "    const " << S << "* __restrict__ pl = src + (zin ? (size_t)z * plane : (size_t)0);\n"
     << "#pragma unroll\n    for (int m = 0; m < NXL; ++m) {\n"
     << "      const bool ok = zin && xo[m] >= 0;\n"
     << "      xr[m] = *reinterpret_cast<const nca_vec<" << S << ", 4>*>(pl + (ok ? xo[m] : 0));\n"
     << "      if (!ok) xr[m] = nca_vec<" << S << ", 4>{};\n    }\n  };\n";
  os << "  auto store_x = [&](int slot) {\n#pragma unroll\n    for (int m = 0; m < NXL; ++m) {\n"
     << "      const int i = tid + 256 * m;\n      if (i < XV * XH) {\n        const int r = i / XV, c = i - r * XV;\n"
     << "        lvec w;\n#pragma unroll\n        for (int j = 0; j < 4; ++j) w.v[j] = nca_cvt<" << C << ", " << S
     << ">::from(xr[m].v[j]);\n"
     << "        *reinterpret_cast<lvec*>(&xs[slot][r * XW + 4 * c]) = w;\n      }\n    }\n  };\n";
  os << "  auto load_e = [&](int z) {\n    const bool zin = (unsigned)z < (unsigned)NZi;\n";
  for (int j = 0; j < n_x; ++j) {
    const std::string E = ctype_of(vj.extra_dtypes[j]);
    os << "    { const " << E << "* __restrict__ pl = e" << j << " + (zin ? (size_t)z * plane : (size_t)0);\n"
       << "#pragma unroll\n      for (int m = 0; m < NWL; ++m) {\n        const bool ok = zin && go[m] >= 0;\n"
       << "        er" << j << "[m] = *reinterpret_cast<const nca_vec<" << E << ", 4>*>(pl + (ok ? go[m] : 0));\n"
       << "        if (!ok) er" << j << "[m] = nca_vec<" << E << ", 4>{};\n      } }\n";
  }
  os << "  };\n";
  os << "  auto store_e = [&]() {\n#pragma unroll\n    for (int m = 0; m < NWL; ++m) {\n";
  for (int j = 0; j < n_x; ++j) os << "      ec" << j << "[m] = er" << j << "[m];\n";
  os << "    }\n  };\n";
  if (fz.fuse & NCA_FUSE_SUM) {
    os << "  " << RC << " acc_s = 0;\n";
    if (comp) os << "  " << RC << " acc_c = 0;\n";
  }
  if (fz.fuse & NCA_FUSE_MAX) os << "  " << RC << " acc_mx = nca_ident<" << VD << ", " << RC << ">::lowest();\n";
  if (fz.fuse & NCA_FUSE_MIN) os << "  " << RC << " acc_mn = nca_ident<" << VD << ", " << RC << ">::max();\n";
  // neighbour reads around an aligned 4-vector at LDS index `base` of plane arrays `arr[slot]`
  // with row pitch `pitch`: per (dz, dy) the aligned vector V and the scalars L / R beside it
  struct Nb { int dz, dy, left = 0, right = 0; };
  auto neighbours = [&](const std::vector<std::array<int64_t, NCA_MAX_NDIM>>& offs, int sign) {
    std::map<std::pair<int, int>, Nb> m;
    for (auto& o : offs) {
      const int dz = sign * static_cast<int>(o[0]), dy = sign * static_cast<int>(o[1]), dx = sign * static_cast<int>(o[2]);
      auto& e = m[{dz, dy}];
      e.dz = dz;
      e.dy = dy;
      if (dx < 0) e.left = std::max(e.left, -dx);
      if (dx > 0) e.right = std::max(e.right, dx);
    }
    return m;
  };
  auto nm = [&](int v) { return v < 0 ? "m" + I(-v) : I(v); };
  auto emit_reads = [&](const std::map<std::pair<int, int>, Nb>& nbs, const std::string& arr,
                        const std::string& slot_m, const std::string& slot_0, const std::string& slot_p,
                        const std::string& base, int pitch, const std::string& tag) {
    for (auto& kv : nbs) {
      const Nb& e = kv.second;
      const std::string sl = e.dz < 0 ? slot_m : (e.dz > 0 ? slot_p : slot_0);
      const std::string at = arr + "[" + sl + "][" + base + " + (" + I(static_cast<long long>(e.dy) * pitch) + ")";
      const std::string id = tag + nm(e.dz) + "_" + nm(e.dy);
      os << "          const lvec V" << id << " = *reinterpret_cast<const lvec*>(&" << at << "]);\n";
      for (int q = 1; q <= e.left; ++q) os << "          const " << C << " L" << id << "_" << q << " = " << at << " - " << q << "];\n";
      for (int q = 1; q <= e.right; ++q) os << "          const " << C << " R" << id << "_" << q << " = " << at << " + " << 3 + q << "];\n";
    }
  };
  auto operand = [&](const std::string& tag, int dz, int dy, int dx, int j) {
    const std::string id = tag + nm(dz) + "_" + nm(dy);
    const int idx = j + dx;
    if (idx >= 0 && idx < 4) return "V" + id + ".v[" + I(idx) + "]";
    if (idx < 0) return "L" + id + "_" + I(-idx);
    return "R" + id + "_" + I(idx - 3);
  };
  const auto xnb = neighbours(p.offsets, 1), wnb = neighbours(p.offsets, -1);
  os << "  load_x(zb - 2);\n  store_x(nca_mod3(zb - 2));\n  load_x(zb - 1);\n  store_x(nca_mod3(zb - 1));\n"
     << "  load_x(zb);\n  load_e(zb - 1);\n";
  os << "  for (int zw = zb - 1; zw <= ze; ++zw) {\n"
     << "    store_x(nca_mod3(zw + 1));\n    store_e();\n"
     << "    if (zw + 2 <= ze + 1) load_x(zw + 2);\n    if (zw + 1 <= ze) load_e(zw + 1);\n"
     << "    __syncthreads();\n"
     << "    const bool own_z = zw >= zb && zw < ze, zin = (unsigned)zw < (unsigned)NZi;\n"
     << "    const int sw = nca_mod3(zw), sm = nca_mod3(zw - 1), sp = nca_mod3(zw + 1);\n"
     << "    const size_t zoff = (size_t)(zw > 0 ? zw : 0) * plane;\n"
     << "#pragma unroll\n    for (int m = 0; m < NWL; ++m) {\n"
     << "      const int i = tid + 256 * m;\n      if (i < NWI) {\n"
     << "        lvec wv = {};\n"
     << "        if (zin && go[m] >= 0) {\n";
  emit_reads(xnb, "xs", "sm", "sw", "sp", "xb[m]", XW, "X");
  for (int j = 0; j < n_x; ++j) os << "          const nca_vec<" << ctype_of(vj.extra_dtypes[j]) << ", 4> E" << j << " = ec" << j << "[m];\n";
  for (size_t m = 1; m < vj.out_dtypes.size(); ++m)
    os << "          nca_vec<" << ctype_of(vj.out_dtypes[m]) << ", 4> g" << m << ";\n";
  if (fz.fuse) os << "          " << RC << " rv[4];\n";
  for (int j = 0; j < 4; ++j) {
    os << "          {\n";
    for (int k = 0; k < n_off; ++k) {
      const auto& o = p.offsets[k];
      os << "            const " << C << " u" << k << " = "
         << operand("X", static_cast<int>(o[0]), static_cast<int>(o[1]), static_cast<int>(o[2]), j) << ";\n";
    }
    for (int jx = 0; jx < n_x; ++jx)
      os << "            const " << C << " u" << (n_off + jx) << " = nca_cvt<" << C << ", " << ctype_of(vj.extra_dtypes[jx])
         << ">::from(E" << jx << ".v[" << j << "]);\n";
    os << sbody << vbody;
    os << "            wv.v[" << j << "] = (" << C << ")(gsc * " << vnames[0] << ");\n";
    for (size_t m = 1; m < vj.out_dtypes.size(); ++m)
      os << "            g" << m << ".v[" << j << "] = nca_store_cvt<" << ctype_of(vj.out_dtypes[m]) << ", " << C << ">::to(("
         << C << ")(gsc * " << vnames[m] << "));\n";
    if (fz.fuse)
      os << "            rv[" << j << "] = nca_cvt<" << RC << ", " << VD << ">::from(nca_store_cvt<" << VD << ", " << C
         << ">::to(" << pname << "));\n";
    os << "          }\n";
  }
  const bool own_work = vj.out_dtypes.size() > 1 || fz.fuse;
  if (own_work) {
    os << "          if (own_z && ((own >> m) & 1u)) {\n            const size_t off = zoff + (size_t)go[m];\n";
    for (size_t m = 1; m < vj.out_dtypes.size(); ++m)
      os << "            *reinterpret_cast<nca_vec<" << ctype_of(vj.out_dtypes[m]) << ", 4>*>(o" << m << " + off) = g" << m
         << ";\n";
    if (fz.fuse & NCA_FUSE_SUM) {
      os << "            const " << RC << " qs = nca_add2(nca_add2(rv[0], rv[1]), nca_add2(rv[2], rv[3]));\n";
      if (comp) os << "            nca_kahan_add(acc_s, acc_c, qs);\n";
      else os << "            acc_s += qs;\n";
    }
    if (fz.fuse & NCA_FUSE_MAX)
      os << "            acc_mx = nca_maxop(acc_mx, nca_maxnum(nca_maxnum(rv[0], rv[1]), nca_maxnum(rv[2], rv[3])));\n";
    if (fz.fuse & NCA_FUSE_MIN)
      os << "            acc_mn = nca_minop(acc_mn, nca_minnum(nca_minnum(rv[0], rv[1]), nca_minnum(rv[2], rv[3])));\n";
    os << "          }\n";
  }
  os << "        }\n        *reinterpret_cast<lvec*>(&ws[sw][wb[m]]) = wv;\n      }\n    }\n    __syncthreads();\n";
  os << "    if (zw - 1 >= zb) {\n"
     << "      const int a0 = nca_mod3(zw - 2), a1 = sm, a2 = sw;\n"
     << "      const size_t doff = (size_t)(zw - 1) * plane;\n"
     << "#pragma unroll\n      for (int m = 0; m < NTL; ++m) {\n"
     << "        if (to[m] >= 0) {\n";
  emit_reads(wnb, "ws", "a0", "a1", "a2", "tb[m]", RWV, "W");
  os << "          nca_vec<" << DX << ", 4> dv;\n";
  for (int j = 0; j < 4; ++j) {
    os << "          {\n";
    for (int k = 0; k < n_off; ++k) {
      const auto& o = p.offsets[k];
      os << "            const " << C << " v" << k << " = "
         << operand("W", -static_cast<int>(o[0]), -static_cast<int>(o[1]), -static_cast<int>(o[2]), j) << ";\n";
    }
    os << abody;
    os << "            dv.v[" << j << "] = nca_store_cvt<" << DX << ", " << C << ">::to(" << aroots[0] << ");\n";
    os << "          }\n";
  }
  os << "          *reinterpret_cast<nca_vec<" << DX << ", 4>*>(dxo + doff + (size_t)to[m]) = dv;\n";
  os << "        }\n      }\n    }\n  }\n";
  if (fz.fuse) {
    os << "  __shared__ double red_sd[4];\n  __shared__ " << RC << " red_mx[4];\n  __shared__ " << RC << " red_mn[4];\n"
       << "  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;\n";
    if (fz.fuse & NCA_FUSE_SUM) {
      if (comp) os << "  nca_wave_comp(acc_s, acc_c);\n  if (lane == 0) red_sd[wid] = nca_comp_value(acc_s, acc_c);\n";
      else os << "  acc_s = nca_wave_sum(acc_s);\n  if (lane == 0) red_sd[wid] = (double)acc_s;\n";
    }
    if (fz.fuse & NCA_FUSE_MAX) os << "  acc_mx = nca_wave_max(acc_mx);\n  if (lane == 0) red_mx[wid] = acc_mx;\n";
    if (fz.fuse & NCA_FUSE_MIN) os << "  acc_mn = nca_wave_min(acc_mn);\n  if (lane == 0) red_mn[wid] = acc_mn;\n";
    os << "  __syncthreads();\n  if (threadIdx.x == 0) {\n";
    if (fz.fuse & NCA_FUSE_SUM) os << "    ((double*)psum)[b] = red_sd[0] + red_sd[1] + red_sd[2] + red_sd[3];\n";
    if (fz.fuse & NCA_FUSE_MAX)
      os << "    " << RC << " m = red_mx[0];\n    for (int w = 1; w < 4; ++w) m = nca_maxop(m, red_mx[w]);\n    *(" << RC
         << "*)((char*)pmax + 8 * (size_t)b) = m;\n";
    if (fz.fuse & NCA_FUSE_MIN)
      os << "    " << RC << " m2 = red_mn[0];\n    for (int w = 1; w < 4; ++w) m2 = nca_minop(m2, red_mn[w]);\n    *(" << RC
         << "*)((char*)pmin + 8 * (size_t)b) = m2;\n";
    os << "  }\n";
  }
  os << "}\n";
  return os.str();
}

// ---- direct kernel ----------------------------------------------------------------------------------------
// sind / dind: the indirect axes of the source's and the destination's dereference chains (a
// pointer read after the step of that axis; the kernel's strides are the chains' byte strides).
static std::string gen_direct(const StencilProgram& p, int32_t dst_dtype, bool zero_boundary, int fuse, int mode,
                              const std::vector<int>& sind = {}, const std::vector<int>& dind = {}) {
  const Tape& t = p.tape;
  const int nd = p.ndim;
  const std::string C = ctype_of(t.compute_dtype);
  const std::string S = ctype_of(p.src_dtype);
  const std::string D = ctype_of(dst_dtype);
  const int n_arr = static_cast<int>(t.arr_dtypes.size());
  std::vector<int> alias(n_arr);
  for (int k = 0; k < n_arr; ++k) alias[k] = k;
  std::vector<std::string> roots;
  std::string body = emit_nodes(t, {t.root}, alias, names_u(n_arr), names_s(static_cast<int>(t.scalar_dtypes.size())),
                                "lin", "t", roots);
  const int32_t rcd = compute_dtype(dst_dtype);
  const std::string RC = ctype_of(rcd);
  const bool comp = mode == NCA_ACCUM_COMPENSATED;
  std::ostringstream os;
  os << "#define NDIM_VAL " << nd << "\nstruct nca_dims { nca_i64 v[NDIM_VAL]; };\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void nca_stencil_direct(const void* src_data, nca_dims sshape, nca_dims sstr";
  if (!sind.empty()) os << ", nca_dims ssub";
  for (size_t j = 0; j < p.extra_dtypes.size(); ++j) os << ", const char* __restrict__ e" << j << ", nca_dims est" << j;
  os << ", void* dest_data, nca_dims dshape, nca_dims dstr";
  if (!dind.empty()) os << ", nca_dims dsub";
  os << ", nca_dims step, nca_i64 n";
  if (fuse) os << ", void* __restrict__ psum, void* __restrict__ pmax, void* __restrict__ pmin";
  os << p.extra_params;
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  os << ") {\n";
  if (fuse) {
    if (fuse & NCA_FUSE_SUM) { os << "  " << RC << " acc_s = 0;\n"; if (comp) os << "  " << RC << " acc_c = 0;\n"; }
    if (fuse & NCA_FUSE_MAX) os << "  " << RC << " acc_mx = nca_ident<" << ctype_of(dst_dtype) << ", " << RC << ">::lowest();\n";
    if (fuse & NCA_FUSE_MIN) os << "  " << RC << " acc_mn = nca_ident<" << ctype_of(dst_dtype) << ", " << RC << ">::max();\n";
  }
  // with hooks: one output per thread (a `return` in the prologue ends the thread, as in
  // the reference); without: grid-stride loop
  if (p.has_ext) {
    os << "  const nca_i64 b_first = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x;\n"
       << "  for (nca_i64 b_idx = b_first; b_idx < n && b_idx == b_first; ++b_idx) {\n";
  } else {
    os << "  const nca_i64 gstride = (nca_i64)gridDim.x * blockDim.x;\n"
       << "  for (nca_i64 b_idx = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; b_idx < n; b_idx += gstride) {\n";
  }
  os << p.prologue << "\n";
  os << "    nca_i64 c_out[NDIM_VAL], c_src[NDIM_VAL];\n"
     << "    { nca_i64 rem = b_idx;\n#pragma unroll\n      for (int d = NDIM_VAL - 1; d >= 0; --d) { c_out[d] = rem % dshape.v[d]; rem /= dshape.v[d]; c_src[d] = c_out[d] * step.v[d]; } }\n";
  auto is_in = [](const std::vector<int>& v, int a) { return std::find(v.begin(), v.end(), a) != v.end(); };
  for (int k = 0; k < p.n_off; ++k) {
    os << "    " << C << " u" << k << ";\n    {\n      const char* pk = reinterpret_cast<const char*>(src_data); bool inb = true;\n";
    for (int d = 0; d < nd; ++d) {
      os << "      { const nca_i64 c = c_src[" << d << "] + (" << p.offsets[k][d] << "); ";
      if (zero_boundary) os << "inb = inb && (c >= 0) && (c < sshape.v[" << d << "]); ";
      os << "pk += c * sstr.v[" << d << "];";
      if (is_in(sind, d)) os << " if (inb) pk = nca_chase(pk, ssub.v[" << d << "]);";
      os << " }\n";
    }
    os << "      u" << k << " = inb ? nca_ld<" << C << ">(reinterpret_cast<const " << S
       << "*>(pk)) : (" << C << ")0;\n    }\n";
  }
  for (size_t j = 0; j < p.extra_dtypes.size(); ++j) {
    os << "    " << C << " u" << (p.n_off + j) << ";\n    { nca_i64 off = 0;\n";
    for (int d = 0; d < nd; ++d) os << "      off += c_out[" << d << "] * est" << j << ".v[" << d << "];\n";
    os << "      u" << (p.n_off + j) << " = nca_ld<" << C << ">(reinterpret_cast<const " << ctype_of(p.extra_dtypes[j]) << "*>(e" << j
       << " + off)); }\n";
  }
  if (t.uses_idx) os << "    const nca_i64 lin = b_idx;\n";
  os << body;
  os << "    const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
  os << "    if (dest_data) { char* pd = reinterpret_cast<char*>(dest_data);\n";
  for (int d = 0; d < nd; ++d) {
    os << "      pd += c_out[" << d << "] * dstr.v[" << d << "];";
    if (is_in(dind, d)) os << " pd = nca_chase(pd, dsub.v[" << d << "]);";
    os << "\n";
  }
  os << "      *reinterpret_cast<" << D << "*>(pd) = st; }\n";
  if (fuse) {
    os << "    { const " << RC << " rv = nca_cvt<" << RC << ", " << D << ">::from(st);\n";
    if (fuse & NCA_FUSE_SUM) os << (comp ? "      nca_neumaier_add(acc_s, acc_c, rv);\n" : "      acc_s += rv;\n");
    if (fuse & NCA_FUSE_MAX) os << "      acc_mx = nca_maxop(acc_mx, rv);\n";
    if (fuse & NCA_FUSE_MIN) os << "      acc_mn = nca_minop(acc_mn, rv);\n";
    os << "    }\n";
  }
  os << p.epilogue << "\n";
  os << "  }\n";
  if (fuse) {
    os << "  __shared__ double red_sd[4];\n  __shared__ " << RC << " red_mx[4];\n  __shared__ " << RC << " red_mn[4];\n"
       << "  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;\n";
    if (fuse & NCA_FUSE_SUM) {
      if (comp) os << "  nca_wave_comp(acc_s, acc_c);\n  if (lane == 0) red_sd[wid] = nca_comp_value(acc_s, acc_c);\n";
      else os << "  acc_s = nca_wave_sum(acc_s);\n  if (lane == 0) red_sd[wid] = (double)acc_s;\n";
    }
    if (fuse & NCA_FUSE_MAX) os << "  acc_mx = nca_wave_max(acc_mx);\n  if (lane == 0) red_mx[wid] = acc_mx;\n";
    if (fuse & NCA_FUSE_MIN) os << "  acc_mn = nca_wave_min(acc_mn);\n  if (lane == 0) red_mn[wid] = acc_mn;\n";
    os << "  __syncthreads();\n  if (threadIdx.x == 0) {\n";
    if (fuse & NCA_FUSE_SUM) os << "    ((double*)psum)[blockIdx.x] = red_sd[0] + red_sd[1] + red_sd[2] + red_sd[3];\n";
    if (fuse & NCA_FUSE_MAX)
      os << "    " << RC << " m = red_mx[0];\n    for (int w = 1; w < 4; ++w) m = nca_maxop(m, red_mx[w]);\n    *(" << RC
         << "*)((char*)pmax + 8 * (size_t)blockIdx.x) = m;\n";
    if (fuse & NCA_FUSE_MIN)
      os << "    " << RC << " m2 = red_mn[0];\n    for (int w = 1; w < 4; ++w) m2 = nca_minop(m2, red_mn[w]);\n    *(" << RC
         << "*)((char*)pmin + 8 * (size_t)blockIdx.x) = m2;\n";
    os << "  }\n";
  }
  os << "}\n";
  return os.str();
}

// ---- tile kernel ---------------------------------------------------------------------------------------
// Any rank >= 2 (a 1-D stencil is launched as 2-D with a leading axis of extent 1), any integer
// steps, strided, NC pointer-axis and SO sources and destinations whose last axis is direct.  A
// block owns a TY x TX tile of the output's last two axes at one coordinate of the outer axes
// and walks a run of such tiles (XCD-aware order).  The offsets are grouped by their outer-axis
// part (one group for a 2-D stencil); each group's source region (the tile's rows and columns
// times the steps, plus the group's halo) is staged in LDS once per tile, in two passes: the
// address of every staged row is resolved once (the pointer chain of NC / SO views is followed
// once per row, as the reference resolves one pointer per offset prefix, rtcompiler.hh:247-308)
// into an LDS table, then the block loads the region as 16-byte vectors (rows whose start is not
// 16-byte aligned, and a strided last axis, element by element).  Reads outside the source give
// 0 (the zero boundary; with no boundary handling the host has checked that every read a
// result uses is in range).  Outputs: one wave per output row, lanes along x, PX outputs per
// lane; fused sum/max/min as in the direct kernel.
struct TileSpec {
  int nd = 2;                                       // rank as launched (>= 2)
  std::vector<std::array<int64_t, NCA_MAX_NDIM>> offsets;  // as launched
  int stepy = 1, stepx = 1;                         // steps of the last two axes (compiled in)
  int px = 2, py = 4;                               // TX = 64 px, TY = 4 py
  int vw = 1;                                       // elements per staging load (16 bytes; 1: strided x)
  bool x32 = false, d32 = false;                    // 32-bit offsets along the last axis (source / dest)
  std::vector<int> src_ind, dst_ind;                // indirect axes (pointer read after the step)
  int fuse = 0, mode = NCA_ACCUM_PLAIN;
  bool with_dst = true;
  int32_t dst_dtype = NCA_FLOAT32;
  // derived: groups of offsets sharing their outer-axis part
  struct Group {
    std::vector<int64_t> outer;
    int64_t miny = 0, maxy = 0, minx = 0, maxx = 0;
    int64_t x0 = 0;                 // first staged column relative to tile x * stepx (a multiple of vw)
    int64_t ry = 0, rx = 0, off = 0, rb = 0;
  };
  std::vector<Group> groups;
  std::vector<int> group_of;
  int64_t lds_elems = 0, rows = 0;
  int TX() const { return 64 * px; }
  int TY() const { return 4 * py; }
  void derive() {
    groups.clear();
    group_of.clear();
    for (auto& o : offsets) {
      std::vector<int64_t> outer(o.begin(), o.begin() + (nd - 2));
      int g = -1;
      for (size_t i = 0; i < groups.size(); ++i)
        if (groups[i].outer == outer) g = static_cast<int>(i);
      if (g < 0) {
        Group gr;
        gr.outer = outer;
        gr.miny = gr.maxy = o[nd - 2];
        gr.minx = gr.maxx = o[nd - 1];
        groups.push_back(gr);
        g = static_cast<int>(groups.size()) - 1;
      }
      Group& gr = groups[g];
      gr.miny = std::min(gr.miny, o[nd - 2]);
      gr.maxy = std::max(gr.maxy, o[nd - 2]);
      gr.minx = std::min(gr.minx, o[nd - 1]);
      gr.maxx = std::max(gr.maxx, o[nd - 1]);
      group_of.push_back(g);
    }
    lds_elems = 0;
    rows = 0;
    for (auto& gr : groups) {
      gr.x0 = gr.minx >= 0 ? gr.minx / vw * vw : -((-gr.minx + vw - 1) / vw) * vw;
      gr.ry = static_cast<int64_t>(TY() - 1) * stepy + (gr.maxy - gr.miny) + 1;
      const int64_t span = static_cast<int64_t>(TX() - 1) * stepx + gr.maxx - gr.x0 + 1;
      gr.rx = (span + vw - 1) / vw * vw;
      gr.off = lds_elems;
      gr.rb = rows;
      lds_elems += gr.ry * gr.rx;
      rows += gr.ry;
    }
  }
  std::string key() const {
    std::ostringstream k;
    k << "t|" << nd << "|" << stepy << "," << stepx << "|" << px << "," << py << "|" << vw << x32 << d32 << "|s";
    for (int a : src_ind) k << a << ",";
    k << "|d";
    for (int a : dst_ind) k << a << ",";
    k << "|" << fuse << "," << mode << "," << with_dst << "," << dst_dtype << "|";
    for (auto& o : offsets)
      for (int d = 0; d < nd; ++d) k << o[d] << (d + 1 < nd ? "," : ";");
    return k.str();
  }
};

static std::string gen_tile(const StencilProgram& p, const TileSpec& ts) {
  const Tape& t = p.tape;
  const int nd = ts.nd;
  const std::string C = ctype_of(t.compute_dtype);
  const std::string S = ctype_of(p.src_dtype);
  const std::string D = ctype_of(ts.dst_dtype);
  const int n_arr = static_cast<int>(t.arr_dtypes.size());
  std::vector<int> alias(n_arr);
  for (int k = 0; k < n_arr; ++k) alias[k] = k;
  std::vector<std::string> roots;
  std::string body = emit_nodes(t, {t.root}, alias, names_u(n_arr), names_s(static_cast<int>(t.scalar_dtypes.size())),
                                "lin", "t", roots);
  const int32_t rcd = compute_dtype(ts.dst_dtype);
  const std::string RC = ctype_of(rcd);
  const bool comp = ts.mode == NCA_ACCUM_COMPENSATED;
  const int fuse = ts.fuse;
  const int TX = ts.TX(), TY = ts.TY();
  const int X = nd - 1, Y = nd - 2;
  const std::string XI = ts.x32 ? "int" : "nca_i64", DI = ts.d32 ? "int" : "nca_i64";
  auto is_in = [](const std::vector<int>& v, int a) { return std::find(v.begin(), v.end(), a) != v.end(); };
  std::ostringstream os;
  os << "#define NDIM_VAL " << nd << "\nstruct nca_dims { nca_i64 v[NDIM_VAL]; };\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void nca_stencil_tile(const char* __restrict__ src, nca_dims sshape, "
        "nca_dims sstr, nca_dims ssub";
  for (size_t j = 0; j < p.extra_dtypes.size(); ++j) os << ", const char* __restrict__ e" << j << ", nca_dims est" << j;
  os << ", char* __restrict__ dst, nca_dims dshape, nca_dims dstr, nca_dims dsub, nca_dims step, nca_i64 ntiles, "
        "nca_i64 ntx, nca_i64 nty, int xcd_order";
  if (fuse) os << ", void* __restrict__ psum, void* __restrict__ pmax, void* __restrict__ pmin";
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  os << ") {\n";
  os << "  __shared__ __attribute__((aligned(16))) " << C << " lds[" << std::max<int64_t>(ts.lds_elems, 1) << "];\n";
  os << "  __shared__ const char* rowp[" << std::max<int64_t>(ts.rows, 1) << "];\n";
  if (fuse) {
    if (fuse & NCA_FUSE_SUM) { os << "  " << RC << " acc_s = 0;\n"; if (comp) os << "  " << RC << " acc_c = 0;\n"; }
    if (fuse & NCA_FUSE_MAX) os << "  " << RC << " acc_mx = nca_ident<" << D << ", " << RC << ">::lowest();\n";
    if (fuse & NCA_FUSE_MIN) os << "  " << RC << " acc_mn = nca_ident<" << D << ", " << RC << ">::max();\n";
  }
  os << "  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;\n";
  os << "  const " << XI << " sx_n = (" << XI << ")sshape.v[" << X << "], sx_s = (" << XI << ")sstr.v[" << X << "];\n";
  os << "  const " << DI << " dx_n = (" << DI << ")dshape.v[" << X << "], dx_s = (" << DI << ")dstr.v[" << X << "];\n";
  // XCD-aware order: block b runs on XCD b % 8; each XCD walks its own contiguous range of tiles,
  // so the halos neighbouring tiles share stay in one L2
  os << "  nca_i64 t_first, t_end, t_step;\n"
     << "  if (xcd_order) {\n"
     << "    const int xcd = blockIdx.x & 7;\n"
     << "    const nca_i64 per = (ntiles + 7) >> 3;\n"
     << "    t_first = (nca_i64)xcd * per + (blockIdx.x >> 3);\n"
     << "    t_end = ((nca_i64)xcd + 1) * per < ntiles ? ((nca_i64)xcd + 1) * per : ntiles;\n"
     << "    t_step = (gridDim.x - xcd + 7) >> 3;\n"
     << "  } else { t_first = blockIdx.x; t_end = ntiles; t_step = gridDim.x; }\n";
  os << "  for (nca_i64 tile = t_first; tile < t_end; tile += t_step) {\n";
  os << "    const nca_i64 txi = tile % ntx, rq = tile / ntx, tyi = rq % nty, oi = rq / nty;\n";
  os << "    nca_i64 co[NDIM_VAL];\n";
  os << "    { nca_i64 rem = oi;\n";
  for (int d = nd - 3; d >= 0; --d) os << "      co[" << d << "] = rem % dshape.v[" << d << "]; rem /= dshape.v[" << d << "];\n";
  os << "      (void)rem;\n    }\n    co[" << Y << "] = tyi * " << TY << "; co[" << X << "] = txi * " << TX << ";\n";
  // pass 1: the address of each staged row (x = 0), or null outside the source
  for (size_t g = 0; g < ts.groups.size(); ++g) {
    const auto& gr = ts.groups[g];
    os << "    for (int rr = threadIdx.x; rr < " << gr.ry << "; rr += 256) {\n"
       << "      const char* rp = src; bool ok = true;\n";
    for (int d = 0; d < X; ++d) {
      if (d < Y)
        os << "      { const nca_i64 c = co[" << d << "] * step.v[" << d << "] + (" << gr.outer[d] << ");";
      else
        os << "      { const nca_i64 c = co[" << d << "] * " << ts.stepy << " + (" << gr.miny << ") + rr;";
      os << " ok = ok && c >= 0 && c < sshape.v[" << d << "]; rp += c * sstr.v[" << d << "];";
      if (is_in(ts.src_ind, d)) os << " if (ok) rp = nca_chase(rp, ssub.v[" << d << "]);";
      os << " }\n";
    }
    os << "      rowp[" << gr.rb << " + rr] = ok ? rp : nullptr;\n    }\n";
  }
  os << "    __syncthreads();\n";
  // pass 2: the regions, 16-byte vectors where the row allows
  const int vw = ts.vw;
  for (size_t g = 0; g < ts.groups.size(); ++g) {
    const auto& gr = ts.groups[g];
    const int64_t nv = gr.rx / vw;
    os << "    {\n      const " << XI << " xa = (" << XI << ")(co[" << X << "] * " << ts.stepx << ") + (" << gr.x0 << ");\n"
       << "      for (int e = threadIdx.x; e < " << gr.ry * nv << "; e += 256) {\n"
       << "        const int r = e / " << nv << ", v = e - r * " << nv << ";\n"
       << "        const char* rp = rowp[" << gr.rb << " + r];\n"
       << "        const " << XI << " xs = xa + v * " << vw << ";\n"
       << "        " << C << "* ld = &lds[" << gr.off << " + r * " << gr.rx << " + v * " << vw << "];\n";
    if (vw > 1) {
      os << "        if (rp != nullptr && xs >= 0 && xs + " << vw << " <= sx_n && (reinterpret_cast<unsigned long long>(rp) & 15) == 0) {\n"
         << "          const nca_vec<" << S << ", " << vw << "> q = *reinterpret_cast<const nca_vec<" << S << ", " << vw
         << ">*>(rp + xs * (" << XI << ")sizeof(" << S << "));\n"
         << "          nca_vec<" << C << ", " << vw << "> w;\n"
         << "#pragma unroll\n          for (int k = 0; k < " << vw << "; ++k) w.v[k] = nca_cvt<" << C << ", " << S << ">::from(q.v[k]);\n"
         << "          *reinterpret_cast<nca_vec<" << C << ", " << vw << ">*>(ld) = w;\n"
         << "        } else {\n"
         << "#pragma unroll\n          for (int k = 0; k < " << vw << "; ++k) {\n"
         << "            const " << XI << " xk = xs + k;\n"
         << "            ld[k] = (rp != nullptr && xk >= 0 && xk < sx_n) ? nca_ld<" << C << ">(reinterpret_cast<const " << S
         << "*>(rp + xk * sx_s)) : (" << C << ")0;\n"
         << "          }\n        }\n";
    } else {
      os << "        ld[0] = (rp != nullptr && xs >= 0 && xs < sx_n) ? nca_ld<" << C << ">(reinterpret_cast<const " << S
         << "*>(rp + xs * sx_s)) : (" << C << ")0;\n";
    }
    os << "      }\n    }\n";
  }
  os << "    __syncthreads();\n";
  // outputs
  os << "#pragma unroll\n    for (int i = 0; i < " << ts.py << "; ++i) {\n"
     << "      const int ly = wid + 4 * i;\n"
     << "      const nca_i64 yo = co[" << Y << "] + ly;\n"
     << "      if (yo < dshape.v[" << Y << "]) {\n";
  if (ts.with_dst) {
    os << "        char* rd = dst;\n";
    for (int d = 0; d < X; ++d) {
      os << "        rd += " << (d < Y ? "co[" + std::to_string(d) + "]" : std::string("yo")) << " * dstr.v[" << d << "];";
      if (is_in(ts.dst_ind, d)) os << " rd = nca_chase(rd, dsub.v[" << d << "]);";
      os << "\n";
    }
  }
  for (size_t j = 0; j < p.extra_dtypes.size(); ++j) {
    os << "        const char* re" << j << " = e" << j << ";\n";
    for (int d = 0; d < X; ++d)
      os << "        re" << j << " += " << (d < Y ? "co[" + std::to_string(d) + "]" : std::string("yo")) << " * est" << j
         << ".v[" << d << "];\n";
  }
  if (t.uses_idx) os << "        const nca_i64 lrow = (oi * dshape.v[" << Y << "] + yo) * dshape.v[" << X << "];\n";
  os << "        const " << DI << " xt = (" << DI << ")co[" << X << "];\n";
  os << "#pragma unroll\n        for (int jx = 0; jx < " << ts.px << "; ++jx) {\n"
     << "          const int lx = lane + 64 * jx;\n"
     << "          const " << DI << " xo = xt + lx;\n"
     << "          if (xo < dx_n) {\n";
  for (int k = 0; k < p.n_off; ++k) {
    const auto& gr = ts.groups[ts.group_of[k]];
    const int64_t oy = ts.offsets[k][Y] - gr.miny, ox = ts.offsets[k][X] - gr.x0;
    os << "            const " << C << " u" << k << " = lds[" << gr.off << " + (ly * " << ts.stepy << " + " << oy << ") * "
       << gr.rx << " + lx * " << ts.stepx << " + " << ox << "];\n";
  }
  for (size_t j = 0; j < p.extra_dtypes.size(); ++j)
    os << "            const " << C << " u" << (p.n_off + j) << " = nca_ld<" << C << ">(reinterpret_cast<const "
       << ctype_of(p.extra_dtypes[j]) << "*>(re" << j << " + (nca_i64)xo * est" << j << ".v[" << X << "]));\n";
  if (t.uses_idx) os << "            const nca_i64 lin = lrow + xo;\n";
  os << body;
  os << "            const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
  if (ts.with_dst) os << "            *reinterpret_cast<" << D << "*>(rd + xo * dx_s) = st;\n";
  if (fuse) {
    os << "            { const " << RC << " rv = nca_cvt<" << RC << ", " << D << ">::from(st);\n";
    if (fuse & NCA_FUSE_SUM) os << (comp ? "              nca_neumaier_add(acc_s, acc_c, rv);\n" : "              acc_s += rv;\n");
    if (fuse & NCA_FUSE_MAX) os << "              acc_mx = nca_maxop(acc_mx, rv);\n";
    if (fuse & NCA_FUSE_MIN) os << "              acc_mn = nca_minop(acc_mn, rv);\n";
    os << "            }\n";
  }
  os << "          }\n        }\n      }\n    }\n    __syncthreads();\n  }\n";
  if (fuse) {
    os << "  __shared__ double red_sd[4];\n  __shared__ " << RC << " red_mx[4];\n  __shared__ " << RC << " red_mn[4];\n";
    if (fuse & NCA_FUSE_SUM) {
      if (comp) os << "  nca_wave_comp(acc_s, acc_c);\n  if (lane == 0) red_sd[wid] = nca_comp_value(acc_s, acc_c);\n";
      else os << "  acc_s = nca_wave_sum(acc_s);\n  if (lane == 0) red_sd[wid] = (double)acc_s;\n";
    }
    if (fuse & NCA_FUSE_MAX) os << "  acc_mx = nca_wave_max(acc_mx);\n  if (lane == 0) red_mx[wid] = acc_mx;\n";
    if (fuse & NCA_FUSE_MIN) os << "  acc_mn = nca_wave_min(acc_mn);\n  if (lane == 0) red_mn[wid] = acc_mn;\n";
    os << "  __syncthreads();\n  if (threadIdx.x == 0) {\n";
    if (fuse & NCA_FUSE_SUM) os << "    ((double*)psum)[blockIdx.x] = red_sd[0] + red_sd[1] + red_sd[2] + red_sd[3];\n";
    if (fuse & NCA_FUSE_MAX)
      os << "    " << RC << " m = red_mx[0];\n    for (int w = 1; w < 4; ++w) m = nca_maxop(m, red_mx[w]);\n    *(" << RC
         << "*)((char*)pmax + 8 * (size_t)blockIdx.x) = m;\n";
    if (fuse & NCA_FUSE_MIN)
      os << "    " << RC << " m2 = red_mn[0];\n    for (int w = 1; w < 4; ++w) m2 = nca_minop(m2, red_mn[w]);\n    *(" << RC
         << "*)((char*)pmin + 8 * (size_t)blockIdx.x) = m2;\n";
    os << "  }\n";
  }
  os << "}\n";
  return os.str();
}

// The kernel of this program for `key`, generated and loaded on first use.  loaded_only: return
// only a kernel already loaded (loading a code object allocates device memory, which the
// allocation-free entry points must not do) and fail otherwise.
static hipFunction_t stencil_kernel(StencilProgram& p, const std::string& key, const std::function<std::string()>& gen,
                                    const std::string& kname, bool loaded_only = false) {
  {
    std::lock_guard<std::mutex> lk(p.mu);
    auto it = p.kernels.find(key);
    if (it != p.kernels.end()) return it->second;
  }
  NCA_REQUIRE(!loaded_only, "kernel " + kname + " of this stencil is not loaded yet: the allocation-free entry point "
                                "only launches loaded kernels (make one nca_hip_stencil_apply_reduce call with the same "
                                "arguments first)");
  std::string src = gen();
  hipFunction_t fn = jit_get(src, kname);
  std::lock_guard<std::mutex> lk(p.mu);
  p.kernels[key] = fn;
  p.sources[key] = src;
  return fn;
}

// Resident 256-thread blocks per CU of a loaded kernel (its registers and LDS), queried once per
// kernel: the march grids are one resident wave of blocks.  NCA_HIP_MARCH_BLOCKS_PER_CU (1..16)
// overrides the query, for tuning.
static int blocks_per_cu(hipFunction_t fn) {
  if (const char* e = std::getenv("NCA_HIP_MARCH_BLOCKS_PER_CU")) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 16) return v;
  }
  static std::mutex mu;
  static std::unordered_map<hipFunction_t, int> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(fn);
    if (it != cache.end()) return it->second;
  }
  int per_cu = 0;
  HIP_CHECK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0));
  if (per_cu < 1) per_cu = 1;
  std::lock_guard<std::mutex> lk(mu);
  cache[fn] = per_cu;
  return per_cu;
}

// Tile kernel launch (see gen_tile).  A 1-D problem runs as 2-D with a leading axis of extent 1.
static void launch_tile(StencilProgram& p, const nca_view& src, const Chain& sch, const nca_view* extras, int n_extra,
                        const nca_view& dst, const Chain& dch, bool with_dst, const int64_t* step, int64_t n, int fuse,
                        int mode, void* sum_out, void* max_out, void* min_out, ScalarSlots& sc, hipStream_t s,
                        void* fixed) {
  DeviceState& dev = current_device();
  const int lift = p.ndim == 1 ? 1 : 0;  // a leading axis of extent 1
  const int nd = p.ndim + lift;
  TileSpec ts;
  ts.nd = nd;
  for (auto& o : p.offsets) {
    std::array<int64_t, NCA_MAX_NDIM> q{};
    for (int d = 0; d < p.ndim; ++d) q[d + lift] = o[d];
    ts.offsets.push_back(q);
  }
  std::array<int64_t, NCA_MAX_NDIM> sshape{}, sstr{}, ssub{}, dshape{}, dstr{}, dsub{}, stp{};
  for (int d = 0; d < nd; ++d) { sshape[d] = dshape[d] = 1; stp[d] = 1; }
  for (int d = 0; d < p.ndim; ++d) {
    sshape[d + lift] = src.shape[d];
    sstr[d + lift] = sch.strides[d];
    dshape[d + lift] = dst.shape[d];
    dstr[d + lift] = with_dst ? dch.strides[d] : 0;
    stp[d + lift] = step[d];
  }
  for (size_t i = 0; i < sch.axes.size(); ++i) {
    ts.src_ind.push_back(sch.axes[i] + lift);
    ssub[sch.axes[i] + lift] = sch.subs[i];
  }
  if (with_dst)
    for (size_t i = 0; i < dch.axes.size(); ++i) {
      ts.dst_ind.push_back(dch.axes[i] + lift);
      dsub[dch.axes[i] + lift] = dch.subs[i];
    }
  NCA_REQUIRE(stp[nd - 2] <= 64 && stp[nd - 1] <= 64, "stencil steps above 64 along the last two axes");
  ts.stepy = static_cast<int>(stp[nd - 2]);
  ts.stepx = static_cast<int>(stp[nd - 1]);
  ts.fuse = fuse;
  ts.mode = mode;
  ts.with_dst = with_dst;
  ts.dst_dtype = dst.dtype;
  // staging loads of 16 bytes when the source's last axis is contiguous
  const int64_t sisz = itemsize(src.dtype);
  ts.vw = (sstr[nd - 1] == sisz && 16 % sisz == 0) ? static_cast<int>(16 / sisz) : 1;
  // 32-bit offsets along the last axis when every position a tile touches fits
  int64_t maxoff = 0;
  for (auto& o : ts.offsets) maxoff = std::max<int64_t>(maxoff, std::llabs(o[nd - 1]));
  const int64_t xreach = (dshape[nd - 1] + 2 * 64 * 8) * stp[nd - 1] + maxoff + 64;
  ts.x32 = xreach * std::max<int64_t>(std::llabs(sstr[nd - 1]), sisz) < (int64_t(1) << 31);
  ts.d32 = (dshape[nd - 1] + 64 * 8) * std::max<int64_t>(std::llabs(dstr[nd - 1]), 16) < (int64_t(1) << 31);
  // tile: TX = 64 px, TY = 4 py, the first of these whose LDS regions fit 24 KiB (six blocks of
  // 256 threads per CU; NCA_HIP_STENCIL_TILE = "px,py" puts another first, for tuning).  Measured
  // on 16384^2 float32 5-point Laplacians: 4x4 429 us, 2x8 440, 2x4 448, 2x2 518; on 2x2 pixel
  // binning of 256 x 1024^2 float32 (16 KiB at 2x2): 2x2 257 us, 2x4 (32 KiB) 283.
  const size_t csz = itemsize(p.tape.compute_dtype);
  std::vector<std::pair<int, int>> cand = {{4, 4}, {2, 8}, {2, 4}, {2, 2}, {1, 4}, {1, 2}, {1, 1}};
  if (lift) cand = {{4, 1}, {2, 1}, {1, 1}};  // a 1-D problem has one row per tile
  bool overridden = false;
  if (const char* e = std::getenv("NCA_HIP_STENCIL_TILE")) {
    int a = 0, b = 0;
    if (std::sscanf(e, "%d,%d", &a, &b) == 2 && a >= 1 && a <= 8 && b >= 1 && b <= 16) {
      cand.insert(cand.begin(), {a, b});
      overridden = true;
    }
  }
  bool fits = false;
  for (size_t i = 0; i < cand.size() && !fits; ++i) {
    ts.px = cand[i].first;
    ts.py = cand[i].second;
    ts.derive();
    const size_t budget = (overridden && i == 0) ? 64 * 1024 : 24 * 1024;
    fits = static_cast<size_t>(ts.lds_elems) * csz + static_cast<size_t>(ts.rows) * 8 <= budget;
  }
  NCA_REQUIRE(fits, "stencil offsets too wide for the tile kernel's LDS");
  const int64_t ntx = (dshape[nd - 1] + ts.TX() - 1) / ts.TX(), nty = (dshape[nd - 2] + ts.TY() - 1) / ts.TY();
  int64_t nouter = 1;
  for (int d = 0; d < nd - 2; ++d) nouter *= dshape[d];
  const int64_t ntiles = nouter * nty * ntx;
  hipFunction_t fn = stencil_kernel(p, ts.key(), [&]() { return gen_tile(p, ts); }, "nca_stencil_tile", fixed != nullptr);
  const int per_cu = blocks_per_cu(fn);
  const int64_t blocks64 = std::min<int64_t>(ntiles, static_cast<int64_t>(dev.n_cus) * per_cu);
  const int blocks = static_cast<int>(std::max<int64_t>(blocks64, 1));
  int xcd_order = blocks >= 8 ? 1 : 0;
  Workspace fw;
  if (fixed) {
    NCA_REQUIRE(static_cast<size_t>(blocks) * 24 <= kFixedPartialBytes, "too many blocks for the fixed scratch array");
    fw.partials = fixed;
  }
  Workspace& w = fixed ? fw : workspace(s, static_cast<size_t>(blocks) * 24);
  std::deque<const void*> ptrs;
  std::deque<std::array<int64_t, NCA_MAX_NDIM>> dims;
  std::vector<void*> params;
  auto push_dims = [&](const std::array<int64_t, NCA_MAX_NDIM>& v) {
    dims.push_back(v);
    params.push_back(dims.back().data());
  };
  ptrs.push_back(reinterpret_cast<const void*>(sch.base));
  params.push_back(&ptrs.back());
  push_dims(sshape);
  push_dims(sstr);
  push_dims(ssub);
  for (int j = 0; j < n_extra; ++j) {
    std::array<int64_t, NCA_MAX_NDIM> es{};
    for (int d = 0; d < p.ndim; ++d) es[d + lift] = extras[j].strides[d];
    ptrs.push_back(extras[j].data);
    params.push_back(&ptrs.back());
    push_dims(es);
  }
  ptrs.push_back(with_dst ? reinterpret_cast<const void*>(dch.base) : nullptr);
  params.push_back(&ptrs.back());
  push_dims(dshape);
  push_dims(dstr);
  push_dims(dsub);
  push_dims(stp);
  int64_t a_ntiles = ntiles, a_ntx = ntx, a_nty = nty;
  params.push_back(&a_ntiles);
  params.push_back(&a_ntx);
  params.push_back(&a_nty);
  params.push_back(&xcd_order);
  void* psum = w.partials;
  void* pmax = static_cast<char*>(w.partials) + static_cast<size_t>(blocks) * 8;
  void* pmin = static_cast<char*>(w.partials) + static_cast<size_t>(blocks) * 16;
  if (fuse) {
    params.push_back(&psum);
    params.push_back(&pmax);
    params.push_back(&pmin);
  }
  for (auto& x : sc.slots) params.push_back(&x);
  (void)n;
  {
    ProfScope ps("nca_stencil_tile", s);
    jit_launch(fn, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, params);
  }
  if (fuse) {
    FinalizeArgs fa;
    fa.value_dtype = dst.dtype;
    fa.fuse = fuse;
    fa.psum = psum;
    fa.pmax = pmax;
    fa.pmin = pmin;
    fa.nblocks = blocks;
    fa.sum_out = sum_out;
    fa.max_out = max_out;
    fa.min_out = min_out;
    launch_finalize_fused(fa, s);
  }
}

// ---- apply ------------------------------------------------------------------------------------------------
// fixed: NULL, or this device's fixed scratch array (kFixedScratchBytes) used instead of the
// per-stream workspace; the call then allocates nothing (see nca_hip_stencil_apply_reduce_noalloc).
static void stencil_apply_impl(uint64_t handle, const nca_view* src, const nca_view* extras, int n_extra,
                               const nca_scalar* scalars, int n_scalars, const nca_view* dst, int boundary,
                               void* const* ext_args, int n_ext_args, const nca_view* axis0, int fuse, int mode,
                               void* sum_out, void* max_out, void* min_out, hipStream_t s, void* fixed = nullptr) {
  StencilProgram& p = stencil_program(handle);
  const Tape& t = p.tape;
  NCA_REQUIRE(src && dst, "null src/dst");
  validate_view_nc(*src, "src");
  const bool with_dst = dst->data != nullptr;
  validate_view_nc(*dst, "dst", true);
  if (p.layout_fixed) {
    // Stencil<NDIM>::create's is_pointer_axis / is_soarr (stencil.hh:76-81): the kernel serves
    // sources with these indirect axes only
    const Chain srcc = chain_of(*src);
    std::vector<int> want;
    for (int d = 0; d < p.ndim; ++d)
      if (p.pointer_axes[d]) want.push_back(d);
    NCA_REQUIRE(srcc.axes == want && (src->layout == NCA_LAYOUT_SO) == p.is_soarr,
                "src layout differs from the pointer axes / SO flag the stencil was created for");
  }
  NCA_REQUIRE(src->ndim == p.ndim && dst->ndim == p.ndim, "src/dst rank differs from the stencil's NDIM");
  NCA_REQUIRE(src->dtype == p.src_dtype, std::string("src dtype must be ") + dtype_info(p.src_dtype).name);
  NCA_REQUIRE(dst->dtype == t.dest_dtype, std::string("dst dtype must be ") + dtype_info(t.dest_dtype).name);
  NCA_REQUIRE(n_extra == static_cast<int>(p.extra_dtypes.size()), "wrong number of extra arrays");
  NCA_REQUIRE(boundary == NCA_BOUNDARY_NONE || boundary == NCA_BOUNDARY_ZERO, "bad boundary mode");
  for (int j = 0; j < n_extra; ++j) {
    validate_view(extras[j], "extra");
    NCA_REQUIRE(extras[j].dtype == p.extra_dtypes[j], "extra dtype mismatch");
    NCA_REQUIRE(same_shape(extras[j], *dst), "extra arrays must have the destination shape");
  }
  ScalarSlots sc = make_scalar_slots(t, scalars, n_scalars);
  int64_t step[NCA_MAX_NDIM];
  for (int d = 0; d < p.ndim; ++d) {
    NCA_REQUIRE(dst->shape[d] > 0 || view_size(*dst) == 0, "bad dst shape");
    step[d] = dst->shape[d] > 0 ? src->shape[d] / dst->shape[d] : 1;
  }
  const int64_t n = view_size(*dst);
  if (boundary == NCA_BOUNDARY_NONE && n > 0) {
    for (int d = 0; d < p.ndim; ++d)
      for (auto& o : p.offsets) {
        const int64_t lo = o[d], hi = (dst->shape[d] - 1) * step[d] + o[d];
        NCA_REQUIRE(lo >= 0 && hi < src->shape[d],
                    "stencil reads outside the source along axis " + std::to_string(d) +
                        " (use zero boundary or pad the source)");
      }
  }
  NCA_REQUIRE(p.has_ext || n_ext_args == 0, "extension arguments given to a stencil without extra_params");
  DeviceState& dev = current_device();
  bool unit = true;
  for (int d = 0; d < p.ndim; ++d) unit = unit && step[d] == 1;
  MarchGeom g = p.ndim == 3 ? geom_of(p) : MarchGeom{};
  const MarchPlan mplan = p.ndim == 3 ? march_plan(p, dst->shape[2], dst->shape[1], dst->shape[0], current_device().n_cus)
                                      : march_plan(p, 0);
  const int vx = mplan.vx;
  // a 3-D view the march reads (writes) plane by plane: C-contiguous with aligned vectors, or (for
  // the register-queue march) a table of plane pointers -- axis 0 the only indirect axis, each
  // plane C-contiguous, the byte offset after the pointer a multiple of the vector size (the
  // plane pointers themselves are the caller's allocations; an unaligned one is still read
  // correctly, global vector accesses need no alignment on this hardware)
  auto plane_table = [&](const nca_view& v, Chain* c) {
    if (v.ndim != 3 || !view_indirect(v) || !mplan.queue) return false;
    *c = chain_of(v);
    const int64_t it = itemsize(v.dtype);
    return c->axes.size() == 1 && c->axes[0] == 0 && c->strides[2] == it && (v.shape[1] == 1 || c->strides[1] == v.shape[2] * it) &&
           c->subs[0] % (vx * it) == 0;
  };
  auto dense_ok = [&](const nca_view& v) {
    return is_c_contiguous(v) && reinterpret_cast<uintptr_t>(v.data) % (vx * itemsize(v.dtype)) == 0;
  };
  Chain stab_c, dtab_c;
  const bool src_tab = !dense_ok(*src) && plane_table(*src, &stab_c);
  const bool dst_tab = with_dst && !dense_ok(*dst) && plane_table(*dst, &dtab_c);
  bool march = p.ndim == 3 && unit && !p.has_ext && g.rz <= 3 && g.ry <= 3 && g.rx <= 3 &&
               dst->shape[1] * dst->shape[2] < (int64_t(1) << 31) && same_shape(*src, *dst) &&
               (itemsize(src->dtype) == 4 || itemsize(src->dtype) == 2) && dst->shape[2] % vx == 0 &&
               (dense_ok(*src) || src_tab) && (!with_dst || dense_ok(*dst) || dst_tab);
  for (int j = 0; j < n_extra && march; ++j)
    march = is_c_contiguous(extras[j]) && reinterpret_cast<uintptr_t>(extras[j].data) % (vx * itemsize(extras[j].dtype)) == 0;
  if (axis0) {
    NCA_REQUIRE(march, "the fused axis-0 sum needs the march kernel: 3-D, step 1, C-contiguous, last extent a multiple of 16 bytes");
    validate_view(*axis0, "axis0_sum");
    NCA_REQUIRE(axis0->ndim == 2 && axis0->shape[0] == dst->shape[1] && axis0->shape[1] == dst->shape[2] &&
                    is_c_contiguous(*axis0),
                "axis0_sum must be C-contiguous with shape dst.shape[1:]");
    NCA_REQUIRE(axis0->dtype == NCA_FLOAT32 || axis0->dtype == NCA_FLOAT64, "axis0_sum dtype must be float32 or float64");
  }
  if (n == 0) {
    // empty: the reductions' identities (zero sums and column sums, the extrema's identities)
    if (fuse || axis0) {
      FinalizeArgs fa;
      fa.K = axis0 ? view_size(*axis0) : 0;
      fa.colout = axis0 ? axis0->data : nullptr;
      fa.col_dtype = axis0 ? axis0->dtype : NCA_FLOAT64;
      fa.value_dtype = dst->dtype;
      fa.fuse = fuse;
      fa.sum_out = sum_out;
      fa.max_out = max_out;
      fa.min_out = min_out;
      launch_finalize_fused(fa, s);
    }
    return;
  }
  if (march) {
    const MarchPlan& mp = mplan;
    const int64_t NZ = dst->shape[0], NY = dst->shape[1], NX = dst->shape[2];
    const int txw = mp.txt * vx, tyw = (256 / mp.txt) * mp.rpt;
    const int ntx = static_cast<int>((NX + txw - 1) / txw), nty = static_cast<int>((NY + tyw - 1) / tyw);
    const int64_t tiles = static_cast<int64_t>(ntx) * nty;
    StencilFuse fz;
    fz.axis0 = axis0 != nullptr;
    fz.fuse = fuse;
    fz.mode = mode;
    fz.with_dst = with_dst;
    fz.axis0_dtype = axis0 ? axis0->dtype : NCA_FLOAT32;
    fz.src_tab = src_tab;
    fz.dst_tab = dst_tab;
    std::string key = "m|" + std::to_string(fz.axis0) + std::to_string(fz.axis0_dtype) + "|" + std::to_string(fuse) + "|" +
                      std::to_string(mode) + "|" + std::to_string(with_dst) + "|" + std::to_string(mp.queue) + "," +
                      std::to_string(mp.vx) + "," + std::to_string(mp.rpt) + "," + std::to_string(mp.txt) + "," + std::to_string(mp.pd) +
                      "|t" + std::to_string(src_tab) + std::to_string(dst_tab);
    const std::string kname = march_name(p, fz, dst->dtype);
    hipFunction_t fn = stencil_kernel(
        p, key, [&]() { return mp.queue ? gen_march_q(p, fz, dst->dtype, mp) : gen_march(p, fz, dst->dtype); }, kname,
        fixed != nullptr);
    // One resident wave of blocks: (tiles x z runs) = CUs x blocks per CU for this kernel's
    // registers and LDS, so no block waits for another to finish (no tail), and runs are
    // as long as possible (fewer halo planes).
    const int per_cu = blocks_per_cu(fn);
    const int64_t resident = static_cast<int64_t>(dev.n_cus) * per_cu;
    int64_t nzc = std::max<int64_t>(1, std::min<int64_t>(resident / std::max<int64_t>(tiles, 1), NZ));
    if (fixed && axis0) {
      // the column partials of the z runs must fit the fixed array: fewer, longer runs
      const int64_t fit = static_cast<int64_t>((kFixedScratchBytes - kFixedPartialBytes) / (sizeof(double) * NY * NX));
      nzc = std::max<int64_t>(1, std::min<int64_t>(nzc, fit));
    }
    int zlen = static_cast<int>((NZ + nzc - 1) / nzc);
    nzc = (NZ + zlen - 1) / zlen;
    const int64_t blocks = tiles * nzc;
    NCA_REQUIRE(blocks < (int64_t(1) << 31), "grid too large");
    size_t colbytes = (axis0 && nzc > 1) ? static_cast<size_t>(nzc * NY * NX) * sizeof(double) : 0;
    Workspace fw;
    if (fixed) {
      NCA_REQUIRE(static_cast<size_t>(blocks) * 24 <= kFixedPartialBytes,
                  "this launch needs " + std::to_string(blocks * 24) + " bytes of block partials, more than the " +
                      std::to_string(kFixedPartialBytes) + " the allocation-free entry point has");
      fw.partials = fixed;
      fw.big = static_cast<char*>(fixed) + kFixedPartialBytes;
    }
    Workspace& w = fixed ? fw : workspace(s, static_cast<size_t>(blocks) * 24, colbytes);
    std::deque<const void*> ptrs;
    std::vector<void*> params;
    ptrs.push_back(src_tab ? reinterpret_cast<const void*>(stab_c.base) : src->data);
    params.push_back(&ptrs.back());
    for (int j = 0; j < n_extra; ++j) { ptrs.push_back(extras[j].data); params.push_back(&ptrs.back()); }
    ptrs.push_back(!with_dst ? nullptr : dst_tab ? reinterpret_cast<const void*>(dtab_c.base) : dst->data);
    params.push_back(&ptrs.back());
    int64_t dims[3] = {NZ, NY, NX};
    params.push_back(&dims[0]);
    params.push_back(&dims[1]);
    params.push_back(&dims[2]);
    int ints[3] = {ntx, nty, zlen};
    params.push_back(&ints[0]);
    params.push_back(&ints[1]);
    params.push_back(&ints[2]);
    void* colpart = (axis0 && nzc > 1) ? w.big : nullptr;
    void* colout = axis0 ? axis0->data : nullptr;
    void* psum = w.partials;
    void* pmax = static_cast<char*>(w.partials) + static_cast<size_t>(blocks) * 8;
    void* pmin = static_cast<char*>(w.partials) + static_cast<size_t>(blocks) * 16;
    params.push_back(&colpart);
    params.push_back(&colout);
    params.push_back(&psum);
    params.push_back(&pmax);
    params.push_back(&pmin);
    int64_t tabs[4] = {src_tab ? stab_c.strides[0] : 0, src_tab ? stab_c.subs[0] : 0, dst_tab ? dtab_c.strides[0] : 0,
                       dst_tab ? dtab_c.subs[0] : 0};
    if (src_tab || dst_tab)
      for (auto& v : tabs) params.push_back(&v);
    for (auto& x : sc.slots) params.push_back(&x);
    {
      ProfScope ps(kname.c_str(), s);
      jit_launch(fn, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, params);
    }
    if (colpart || fuse) {
      FinalizeArgs fa;
      fa.colpart = static_cast<const double*>(colpart);
      fa.nzc = static_cast<int>(nzc);
      fa.K = colpart ? NY * NX : 0;
      fa.colout = colout;
      fa.col_dtype = axis0 ? axis0->dtype : NCA_FLOAT64;
      fa.value_dtype = dst->dtype;
      fa.fuse = fuse;
      fa.psum = psum;
      fa.pmax = pmax;
      fa.pmin = pmin;
      fa.nblocks = static_cast<int>(blocks);
      fa.sum_out = sum_out;
      fa.max_out = max_out;
      fa.min_out = min_out;
      launch_finalize_fused(fa, s);
    }
    return;
  }
  // tile: every stencil without hooks whose source and destination read their last axis directly
  {
    const Chain sch = chain_of(*src);
    const Chain dch = with_dst ? chain_of(*dst) : Chain{};
    const int last = p.ndim - 1;
    const bool direct_x = std::find(sch.axes.begin(), sch.axes.end(), last) == sch.axes.end() &&
                          std::find(dch.axes.begin(), dch.axes.end(), last) == dch.axes.end();
    if (!p.has_ext && direct_x) {
      launch_tile(p, *src, sch, extras, n_extra, *dst, dch, with_dst, step, n, fuse, mode, sum_out, max_out, min_out, sc,
                  s, fixed);
      return;
    }
  }
  // direct
  const bool zero = boundary == NCA_BOUNDARY_ZERO;
  const Chain sch = chain_of(*src);
  const Chain dch = with_dst ? chain_of(*dst) : Chain{};
  std::string key = "d|" + std::to_string(zero) + "|" + std::to_string(fuse) + "|" + std::to_string(mode) + "|s";
  for (int a : sch.axes) key += std::to_string(a) + ",";
  key += "|d";
  for (int a : dch.axes) key += std::to_string(a) + ",";
  hipFunction_t fn = stencil_kernel(
      p, key, [&]() { return gen_direct(p, dst->dtype, zero, fuse, mode, sch.axes, dch.axes); }, "nca_stencil_direct",
      fixed != nullptr);
  int blocks;
  if (p.has_ext) {
    NCA_REQUIRE((n + 255) / 256 < (int64_t(1) << 31), "grid too large");
    blocks = static_cast<int>((n + 255) / 256);
    NCA_REQUIRE(!fuse, "fused reductions are not available with extension hooks");
  } else {
    blocks = grid_blocks(n, 256, fuse ? dev.n_cus * 4 : dev.n_cus * 16);
  }
  Workspace fw;
  if (fixed) {
    NCA_REQUIRE(static_cast<size_t>(blocks) * 24 <= kFixedPartialBytes, "too many blocks for the fixed scratch array");
    fw.partials = fixed;
  }
  Workspace& w = fixed ? fw : workspace(s, static_cast<size_t>(blocks) * 24);
  std::deque<const void*> ptrs;
  std::deque<std::array<int64_t, NCA_MAX_NDIM>> dims;
  std::vector<void*> params;
  auto push_dims = [&](const int64_t* v) {
    std::array<int64_t, NCA_MAX_NDIM> a{};
    for (int d = 0; d < p.ndim; ++d) a[d] = v[d];
    dims.push_back(a);
    params.push_back(dims.back().data());
  };
  auto subs_of = [&](const Chain& c) {
    int64_t v[NCA_MAX_NDIM] = {};
    for (size_t i = 0; i < c.axes.size(); ++i) v[c.axes[i]] = c.subs[i];
    push_dims(v);
  };
  ptrs.push_back(reinterpret_cast<const void*>(sch.base));
  params.push_back(&ptrs.back());
  push_dims(src->shape);
  push_dims(sch.strides.data());
  if (!sch.axes.empty()) subs_of(sch);
  for (int j = 0; j < n_extra; ++j) {
    ptrs.push_back(extras[j].data);
    params.push_back(&ptrs.back());
    push_dims(extras[j].strides);
  }
  ptrs.push_back(with_dst ? reinterpret_cast<const void*>(dch.base) : nullptr);
  params.push_back(&ptrs.back());
  push_dims(dst->shape);
  push_dims(with_dst ? dch.strides.data() : dst->strides);
  if (!dch.axes.empty()) subs_of(dch);
  push_dims(step);
  int64_t nn = n;
  params.push_back(&nn);
  void* psum = w.partials;
  void* pmax = static_cast<char*>(w.partials) + static_cast<size_t>(blocks) * 8;
  void* pmin = static_cast<char*>(w.partials) + static_cast<size_t>(blocks) * 16;
  if (fuse) {
    params.push_back(&psum);
    params.push_back(&pmax);
    params.push_back(&pmin);
  }
  for (int k = 0; k < n_ext_args; ++k) params.push_back(ext_args[k]);
  for (auto& x : sc.slots) params.push_back(&x);
  {
    ProfScope ps("nca_stencil_direct", s);
    jit_launch(fn, dim3(blocks), dim3(256), 0, s, params);
  }
  if (fuse) {
    FinalizeArgs fa;
    fa.value_dtype = dst->dtype;
    fa.fuse = fuse;
    fa.psum = psum;
    fa.pmax = pmax;
    fa.pmin = pmin;
    fa.nblocks = blocks;
    fa.sum_out = sum_out;
    fa.max_out = max_out;
    fa.min_out = min_out;
    launch_finalize_fused(fa, s);
  }
}

}  // namespace nca

using namespace nca;

extern "C" {

int nca_hip_stencil_create(const nca_stencil_desc* desc, uint64_t* handle) {
  return api_call([&] {
    NCA_REQUIRE(desc && handle && desc->expr, "null argument");
    NCA_REQUIRE(desc->ndim >= 1 && desc->ndim <= NCA_MAX_NDIM, "stencil NDIM out of range");
    NCA_REQUIRE(desc->n_offsets >= 1 && desc->offsets, "stencil needs at least one offset");
    NCA_REQUIRE(desc->n_extra >= 0 && (desc->n_extra == 0 || desc->extra_dtypes), "bad extra arrays");
    auto p = std::make_unique<StencilProgram>();
    p->ndim = desc->ndim;
    p->n_off = desc->n_offsets;
    p->src_dtype = desc->src_dtype;
    NCA_REQUIRE(is_supported(p->src_dtype), "unsupported source dtype");
    for (int k = 0; k < desc->n_offsets; ++k) {
      std::array<int64_t, NCA_MAX_NDIM> o{};
      for (int d = 0; d < desc->ndim; ++d) o[d] = desc->offsets[k * desc->ndim + d];
      p->offsets.push_back(o);
    }
    p->extra_dtypes.assign(desc->extra_dtypes, desc->extra_dtypes + desc->n_extra);
    NCA_REQUIRE(desc->expr->n_arrays == desc->n_offsets + desc->n_extra,
                "the expression must load n_offsets + n_extra arrays");
    for (int k = 0; k < desc->expr->n_arrays; ++k) {
      int32_t want = k < desc->n_offsets ? desc->src_dtype : desc->extra_dtypes[k - desc->n_offsets];
      NCA_REQUIRE(desc->expr->arr_dtypes[k] == want, "expression array dtypes disagree with src/extra dtypes");
    }
    p->tape = stencil_tape(*desc->expr);
    p->extra_params = desc->extra_params ? desc->extra_params : "";
    p->prologue = desc->prologue ? desc->prologue : "";
    p->epilogue = desc->epilogue ? desc->epilogue : "";
    p->has_ext = !p->extra_params.empty() || !p->prologue.empty() || !p->epilogue.empty();
    if (desc->is_pointer_axis) {
      p->layout_fixed = true;
      int n_ptr = 0;
      for (int d = 0; d < desc->ndim; ++d) {
        p->pointer_axes[d] = desc->is_pointer_axis[d] ? 1 : 0;
        n_ptr += p->pointer_axes[d];
      }
      p->is_soarr = desc->is_soarr != 0;
      NCA_REQUIRE(p->is_soarr || n_ptr <= 1, "an NC source has at most one pointer axis (use is_soarr for more)");
    } else {
      NCA_REQUIRE(desc->is_soarr == 0, "is_soarr needs is_pointer_axis");
    }
    *handle = stencil_register(std::move(p));
  });
}

int nca_hip_stencil_apply(uint64_t handle, const nca_view* src, const nca_view* extras, int32_t n_extra,
                          const nca_scalar* scalars, int32_t n_scalars, const nca_view* dst, int32_t boundary,
                          void* const* ext_args, int32_t n_ext_args, void* stream) {
  return api_call([&] {
    NCA_REQUIRE(dst && dst->data, "stencil_apply needs a destination");
    stencil_apply_impl(handle, src, extras, n_extra, scalars, n_scalars, dst, boundary, ext_args, n_ext_args, nullptr,
                       0, NCA_ACCUM_PLAIN, nullptr, nullptr, nullptr, as_stream(stream));
    check_launch("stencil");
    finish_call(stream);
  });
}

int nca_hip_stencil_apply_reduce(uint64_t handle, const nca_view* src, const nca_view* extras, int32_t n_extra,
                                 const nca_scalar* scalars, int32_t n_scalars, const nca_view* dst, int32_t boundary,
                                 const nca_view* axis0_sum, int32_t fuse, int32_t accum_mode, void* sum_out,
                                 void* max_out, void* min_out, void* stream) {
  return api_call([&] {
    NCA_REQUIRE((fuse & ~(NCA_FUSE_SUM | NCA_FUSE_MAX | NCA_FUSE_MIN)) == 0, "bad fuse flags");
    NCA_REQUIRE(!(fuse & NCA_FUSE_SUM) || sum_out, "sum requested without sum_out");
    NCA_REQUIRE(!(fuse & NCA_FUSE_MAX) || max_out, "max requested without max_out");
    NCA_REQUIRE(!(fuse & NCA_FUSE_MIN) || min_out, "min requested without min_out");
    stencil_apply_impl(handle, src, extras, n_extra, scalars, n_scalars, dst, boundary, nullptr, 0, axis0_sum, fuse,
                       accum_mode, sum_out, max_out, min_out, as_stream(stream));
    check_launch("fused stencil");
    finish_call(stream);
  });
}

int nca_hip_stencil_apply_reduce_noalloc(uint64_t handle, const nca_view* src, const nca_view* extras, int32_t n_extra,
                                         const nca_scalar* scalars, int32_t n_scalars, const nca_view* dst,
                                         int32_t boundary, const nca_view* axis0_sum, int32_t fuse, int32_t accum_mode,
                                         void* sum_out, void* max_out, void* min_out, void* stream) {
  return api_call([&] {
    NCA_REQUIRE((fuse & ~(NCA_FUSE_SUM | NCA_FUSE_MAX | NCA_FUSE_MIN)) == 0, "bad fuse flags");
    NCA_REQUIRE(!(fuse & NCA_FUSE_SUM) || sum_out, "sum requested without sum_out");
    NCA_REQUIRE(!(fuse & NCA_FUSE_MAX) || max_out, "max requested without max_out");
    NCA_REQUIRE(!(fuse & NCA_FUSE_MIN) || min_out, "min requested without min_out");
    DeviceState& dev = current_device();
    NCA_REQUIRE(dev.fixed_scratch != nullptr, "the fixed scratch array of this device was not found at init");
    stencil_apply_impl(handle, src, extras, n_extra, scalars, n_scalars, dst, boundary, nullptr, 0, axis0_sum, fuse,
                       accum_mode, sum_out, max_out, min_out, as_stream(stream), dev.fixed_scratch);
    check_launch("fused stencil (allocation-free)");
    finish_call(stream);
  });
}

}  // extern "C"

namespace nca {
// The adjoint (transpose) tape of a stencil linear in its source loads: sum_k c_k * LOAD k with
// c_k = d f / d v_k derived symbolically (constants or functions of the scalars, whose slots it
// shares with the forward tape); LOAD k reads the gradient at the negated offset k.
static Tape adjoint_tape(const StencilProgram& f) {
  NCA_REQUIRE(!f.has_ext, "cannot differentiate a stencil with extension hooks");
  Tape base = f.tape;
  std::vector<int> alias(base.arr_dtypes.size());
  for (size_t k = 0; k < alias.size(); ++k) alias[k] = static_cast<int>(k);
  // c_k = d f / d v_k must not depend on any load (linear stencil)
  std::vector<int> coef;
  for (int k = 0; k < f.n_off; ++k) {
    int dk = differentiate(base, base.root, k, alias);
    NCA_REQUIRE(!depends_on_loads(base, dk),
                "the stencil is not linear in its source loads; its adjoint is not a stencil");
    coef.push_back(dk);
  }
  Tape t = base;
  t.arr_dtypes.assign(f.n_off, f.tape.dest_dtype);
  int acc = -1;
  for (int k = 0; k < f.n_off; ++k) {
    if (coef[k] < 0) continue;
    Node ld;
    ld.op = NCA_OP_LOAD_NCARR;
    ld.index = k;
    int term = t.bin(NCA_OP_MUL, coef[k], t.add(ld));
    acc = acc < 0 ? term : t.bin(NCA_OP_ADD, acc, term);
  }
  NCA_REQUIRE(acc >= 0, "the stencil does not depend on its source");
  t.root = acc;
  t.uses_idx = false;
  t.dest_dtype = f.src_dtype;
  t.signature = "adjoint(" + f.tape.signature + ")";
  return t;
}
}  // namespace nca

namespace nca {
// One march pass of the stencil-expression vjp; with fuse != 0 the pass also reduces the
// expression's value (sum / max / min as the fused forward reductions do).
static void stencil_expr_vjp_impl(uint64_t stencil_handle, uint64_t expr_handle, const int32_t* expr_operand,
                                  int32_t n_operand, const nca_view* src, const nca_view* extras, int32_t n_extra,
                                  const nca_scalar* stencil_scalars, int32_t n_stencil_scalars,
                                  const nca_scalar* expr_scalars, int32_t n_expr_scalars, const void* grad_scalar,
                                  const nca_view* w, const nca_view* grad_extras, int32_t fuse, int32_t mode,
                                  void* sum_out, void* max_out, void* min_out, int32_t boundary, void* stream) {
  {
    StencilProgram& p = stencil_program(stencil_handle);
    const Tape& et = expr_program(expr_handle).tape;
    NCA_REQUIRE(src && w && (n_operand == 0 || expr_operand), "null argument");
    NCA_REQUIRE(p.extra_dtypes.empty() && !p.has_ext, "the stencil of a stencil-expression vjp reads only its source");
    NCA_REQUIRE(p.ndim == 3, "the stencil-expression vjp is 3-D");
    NCA_REQUIRE(boundary == NCA_BOUNDARY_ZERO, "the stencil-expression vjp needs the zero boundary");
    NCA_REQUIRE(n_operand == static_cast<int32_t>(et.arr_dtypes.size()), "one operand code per expression operand");
    NCA_REQUIRE(n_extra >= 0 && n_extra <= 8 && (n_extra == 0 || extras), "0..8 extra arrays");
    NCA_REQUIRE(et.compute_dtype == p.tape.compute_dtype, "the expression and the stencil must compute in one dtype");
    validate_view(*src, "src");
    validate_view(*w, "w");
    NCA_REQUIRE(src->dtype == p.src_dtype, std::string("src dtype must be ") + dtype_info(p.src_dtype).name);
    NCA_REQUIRE(src->ndim == 3 && same_shape(*src, *w) && is_c_contiguous(*src) && is_c_contiguous(*w),
                "src and w must be C-contiguous 3-D arrays of one shape");
    NCA_REQUIRE(is_float(w->dtype), "w must be floating point");
    NCA_REQUIRE((fuse & ~(NCA_FUSE_SUM | NCA_FUSE_MAX | NCA_FUSE_MIN)) == 0, "unknown fuse flags");
    NCA_REQUIRE(mode == NCA_ACCUM_PLAIN || mode == NCA_ACCUM_COMPENSATED, "unknown accumulation mode");
    NCA_REQUIRE((!(fuse & NCA_FUSE_SUM) || sum_out) && (!(fuse & NCA_FUSE_MAX) || max_out) &&
                    (!(fuse & NCA_FUSE_MIN) || min_out),
                "a fused reduction needs its output pointer");
    MarchVjp vj;
    vj.expr = et;
    vj.value_dtype = et.dest_dtype;
    vj.operand.assign(expr_operand, expr_operand + n_operand);
    int first_a = -1;
    for (int k = 0; k < n_operand; ++k) {
      NCA_REQUIRE(vj.operand[k] >= -1 && vj.operand[k] < n_extra, "expression operand code out of range");
      if (vj.operand[k] < 0 && first_a < 0) first_a = k;
    }
    NCA_REQUIRE(first_a >= 0, "the expression must read the stencil value");
    for (int j = 0; j < n_extra; ++j) {
      validate_view(extras[j], "extra");
      NCA_REQUIRE(same_shape(extras[j], *w) && is_c_contiguous(extras[j]), "extras must be C-contiguous with w's shape");
      vj.extra_dtypes.push_back(extras[j].dtype);
    }
    for (int k = 0; k < n_operand; ++k)
      if (vj.operand[k] >= 0)
        NCA_REQUIRE(et.arr_dtypes[k] == extras[vj.operand[k]].dtype, "extra dtype differs from the expression operand's");
    vj.alias.resize(n_operand);
    for (int k = 0; k < n_operand; ++k) {
      vj.alias[k] = k;
      for (int q = 0; q < k; ++q)
        if (vj.operand[q] == vj.operand[k]) { vj.alias[k] = q; break; }
    }
    // outputs: w = d expr / d a, then d expr / d extra_j for every grad_extras[j] with data
    std::vector<void*> outs{w->data};
    vj.roots.push_back(differentiate(vj.expr, vj.expr.root, vj.alias[first_a], vj.alias));
    vj.out_dtypes.push_back(w->dtype);
    for (int j = 0; j < n_extra && grad_extras; ++j) {
      if (!grad_extras[j].data) continue;
      validate_view(grad_extras[j], "grad_extra");
      NCA_REQUIRE(same_shape(grad_extras[j], *w) && is_c_contiguous(grad_extras[j]) && is_float(grad_extras[j].dtype),
                  "extra gradients must be C-contiguous floating-point arrays of w's shape");
      int var = -1;
      for (int k = 0; k < n_operand && var < 0; ++k)
        if (vj.operand[k] == j) var = vj.alias[k];
      vj.roots.push_back(var >= 0 ? differentiate(vj.expr, vj.expr.root, var, vj.alias) : -1);
      vj.out_dtypes.push_back(grad_extras[j].dtype);
      outs.push_back(grad_extras[j].data);
    }
    const int64_t NZ = w->shape[0], NY = w->shape[1], NX = w->shape[2];
    MarchPlan mp = march_plan(p, NX, NY, NZ, current_device().n_cus);
    if (mp.defer) {
      // the backward march keeps immediate conversions and one plane ahead (at 256^3 BF16 the
      // deferred form made the loss-and-gradient step 5% slower)
      mp.defer = false;
      mp.pd = 1;
    }
    const MarchGeom g = geom_of(p);
    NCA_REQUIRE(mp.queue && g.rz <= 3 && g.ry <= 3 && g.rx <= 3,
                "the stencil-expression vjp needs a stencil whose z offsets have no x or y part (radius <= 3)");
    NCA_REQUIRE(NY * NX < (int64_t(1) << 31), "plane too large");
    const int vx = mp.vx;
    auto aligned = [&](const void* ptr, int32_t dt) { return reinterpret_cast<uintptr_t>(ptr) % (vx * itemsize(dt)) == 0; };
    bool ok = NX % vx == 0 && aligned(src->data, src->dtype);
    for (int j = 0; j < n_extra; ++j) ok = ok && aligned(extras[j].data, extras[j].dtype);
    for (size_t m = 0; m < outs.size(); ++m) ok = ok && aligned(outs[m], vj.out_dtypes[m]);
    NCA_REQUIRE(ok, "the stencil-expression vjp needs the last extent a multiple of " + std::to_string(vx) +
                        " elements and arrays aligned to whole vectors");
    const int64_t n = view_size(*w);
    hipStream_t s = as_stream(stream);
    if (n == 0) {
      if (fuse) {
        FinalizeArgs fa;
        fa.value_dtype = et.dest_dtype;
        fa.fuse = fuse;
        fa.sum_out = sum_out;
        fa.max_out = max_out;
        fa.min_out = min_out;
        launch_finalize_fused(fa, s);
      }
      finish_call(stream);
      return;
    }
    DeviceState& dev = current_device();
    std::string key = "vjp|" + et.signature + "|";
    for (int o : vj.operand) key += std::to_string(o) + ",";
    key += "|";
    for (auto d : vj.extra_dtypes) key += std::to_string(d) + ",";
    key += "|";
    for (size_t m = 0; m < vj.roots.size(); ++m) key += std::to_string(vj.out_dtypes[m]) + (vj.roots[m] < 0 ? "z," : ",");
    key += "|" + std::to_string(mp.vx) + "," + std::to_string(mp.rpt) + "," + std::to_string(mp.txt) + "," +
           std::to_string(mp.pd);
    key += "|f" + std::to_string(fuse) + "," + std::to_string(fuse ? mode : 0);
    const std::string kname =
        std::string("nca_stencil_march_vjp_") + dtype_info(p.src_dtype).name + (fuse ? "_sum" : "");
    StencilFuse fz;
    fz.with_dst = false;
    fz.fuse = fuse;
    fz.mode = mode;
    hipFunction_t fn = stencil_kernel(p, key, [&]() { return gen_march_q(p, fz, w->dtype, mp, &vj); }, kname);
    const int txw = mp.txt * vx, tyw = (256 / mp.txt) * mp.rpt;
    const int ntx = static_cast<int>((NX + txw - 1) / txw), nty = static_cast<int>((NY + tyw - 1) / tyw);
    const int64_t tiles = static_cast<int64_t>(ntx) * nty;
    const int per_cu = blocks_per_cu(fn);
    const int64_t resident = static_cast<int64_t>(dev.n_cus) * per_cu;
    int64_t nzc = std::max<int64_t>(1, std::min<int64_t>(resident / std::max<int64_t>(tiles, 1), NZ));
    const int zlen = static_cast<int>((NZ + nzc - 1) / nzc);
    nzc = (NZ + zlen - 1) / zlen;
    const int64_t blocks = tiles * nzc;
    NCA_REQUIRE(blocks < (int64_t(1) << 31), "grid too large");
    ScalarSlots ss = make_scalar_slots(p.tape, stencil_scalars, n_stencil_scalars);
    ScalarSlots es = make_scalar_slots(et, expr_scalars, n_expr_scalars);
    std::deque<const void*> ptrs;
    std::vector<void*> params;
    ptrs.push_back(src->data);
    params.push_back(&ptrs.back());
    for (int j = 0; j < n_extra; ++j) { ptrs.push_back(extras[j].data); params.push_back(&ptrs.back()); }
    for (void* o : outs) { ptrs.push_back(o); params.push_back(&ptrs.back()); }
    int64_t dims[3] = {NZ, NY, NX};
    for (auto& d : dims) params.push_back(&d);
    int ints[3] = {ntx, nty, zlen};
    for (auto& i : ints) params.push_back(&i);
    void* psum = nullptr;
    void* pmax = nullptr;
    void* pmin = nullptr;
    if (fuse) {
      Workspace& ws = workspace(s, static_cast<size_t>(blocks) * 24);
      psum = ws.partials;
      pmax = static_cast<char*>(ws.partials) + static_cast<size_t>(blocks) * 8;
      pmin = static_cast<char*>(ws.partials) + static_cast<size_t>(blocks) * 16;
      params.push_back(&psum);
      params.push_back(&pmax);
      params.push_back(&pmin);
    }
    for (auto& x : ss.slots) params.push_back(&x);
    for (auto& x : es.slots) params.push_back(&x);
    ptrs.push_back(grad_scalar);
    params.push_back(&ptrs.back());
    {
      ProfScope ps(kname.c_str(), s);
      jit_launch(fn, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, params);
    }
    check_launch("stencil_expr_vjp");
    if (fuse) {
      FinalizeArgs fa;
      fa.value_dtype = et.dest_dtype;
      fa.fuse = fuse;
      fa.psum = psum;
      fa.pmax = pmax;
      fa.pmin = pmin;
      fa.nblocks = static_cast<int>(blocks);
      fa.sum_out = sum_out;
      fa.max_out = max_out;
      fa.min_out = min_out;
      launch_finalize_fused(fa, s);
    }
    finish_call(stream);
  }
}

// Tile of the training-step march: TX x TY points per block, the fastest measured at 256^3:
// 64 x 16 for 4-byte sources (58-59 us; 32-128 x 8-16 within 2%), 64 x 8 for 2-byte sources
// (vector form 55.8 us against 56.3 at 64 x 16 and 62.6 at 128 x 16, whose 62 KiB of LDS allow
// two blocks per CU); NCA_HIP_TRAIN_TILE = "TXxTY" overrides it, for tuning.
static void train_tile(int32_t src_dtype, bool vec, int& tx, int& ty) {
  const bool two_byte = itemsize(src_dtype) == 2;
  tx = 64;
  ty = two_byte ? 8 : 16;
  if (const char* e = std::getenv("NCA_HIP_TRAIN_TILE")) {
    int a = 0, b = 0;
    if (std::sscanf(e, "%dx%d", &a, &b) == 2 && (a == 32 || a == 64 || a == 128 || a == 256) &&
        (b == 2 || b == 4 || b == 8 || b == 16 || b == 32)) {
      // the three x planes and the three w planes (4-byte compute type) must fit the 64 KiB
      // of LDS one block may hold
      const int64_t xw = vec ? a + 16 : a + 4, rw = vec ? a + 8 : a + 2;
      const int64_t lds = 4 * (3 * (b + 4) * xw + 3 * (b + 2) * rw);
      if (lds <= 60 * 1024) {
        tx = a;
        ty = b;
      }
    }
  }
}

// The whole training step of loss = sum(expr(stencil(src), extras)) in one pass
// (gen_march_train): dx = stencil^T(d expr / d a * g), the extras' gradients and optionally the
// reductions of the expression's value.
static void stencil_expr_train_impl(uint64_t stencil_handle, uint64_t expr_handle, const int32_t* expr_operand,
                                    int32_t n_operand, const nca_view* src, const nca_view* extras, int32_t n_extra,
                                    const nca_scalar* stencil_scalars, int32_t n_stencil_scalars,
                                    const nca_scalar* expr_scalars, int32_t n_expr_scalars, const void* grad_scalar,
                                    const nca_view* grad_src, const nca_view* grad_extras, int32_t fuse, int32_t mode,
                                    void* sum_out, void* max_out, void* min_out, int32_t boundary, void* stream) {
  StencilProgram& p = stencil_program(stencil_handle);
  const Tape& et = expr_program(expr_handle).tape;
  NCA_REQUIRE(src && grad_src && (n_operand == 0 || expr_operand), "null argument");
  NCA_REQUIRE(p.extra_dtypes.empty() && !p.has_ext, "the stencil of a stencil-expression step reads only its source");
  NCA_REQUIRE(p.ndim == 3, "the stencil-expression step is 3-D");
  NCA_REQUIRE(boundary == NCA_BOUNDARY_ZERO, "the stencil-expression step needs the zero boundary");
  for (auto& o : p.offsets)
    NCA_REQUIRE(o[0] >= -1 && o[0] <= 1 && o[1] >= -1 && o[1] <= 1 && o[2] >= -1 && o[2] <= 1,
                "the stencil-expression step needs a stencil of radius <= 1 on every axis");
  NCA_REQUIRE(!p.tape.uses_idx && !et.uses_idx, "the stencil-expression step does not evaluate IDX");
  NCA_REQUIRE(n_operand == static_cast<int32_t>(et.arr_dtypes.size()), "one operand code per expression operand");
  NCA_REQUIRE(n_extra >= 0 && n_extra <= 8 && (n_extra == 0 || extras), "0..8 extra arrays");
  NCA_REQUIRE(et.compute_dtype == p.tape.compute_dtype, "the expression and the stencil must compute in one dtype");
  NCA_REQUIRE((fuse & ~(NCA_FUSE_SUM | NCA_FUSE_MAX | NCA_FUSE_MIN)) == 0, "unknown fuse flags");
  NCA_REQUIRE(mode == NCA_ACCUM_PLAIN || mode == NCA_ACCUM_COMPENSATED, "unknown accumulation mode");
  NCA_REQUIRE((!(fuse & NCA_FUSE_SUM) || sum_out) && (!(fuse & NCA_FUSE_MAX) || max_out) &&
                  (!(fuse & NCA_FUSE_MIN) || min_out),
              "a fused reduction needs its output pointer");
  validate_view(*src, "src");
  validate_view(*grad_src, "grad_src");
  NCA_REQUIRE(src->dtype == p.src_dtype, std::string("src dtype must be ") + dtype_info(p.src_dtype).name);
  NCA_REQUIRE(src->ndim == 3 && same_shape(*src, *grad_src) && is_c_contiguous(*src) && is_c_contiguous(*grad_src),
              "src and grad_src must be C-contiguous 3-D arrays of one shape");
  NCA_REQUIRE(is_float(grad_src->dtype), "grad_src must be floating point");
  const Tape adj = adjoint_tape(p);
  MarchVjp vj;
  vj.expr = et;
  vj.value_dtype = et.dest_dtype;
  vj.operand.assign(expr_operand, expr_operand + n_operand);
  int first_a = -1;
  for (int k = 0; k < n_operand; ++k) {
    NCA_REQUIRE(vj.operand[k] >= -1 && vj.operand[k] < n_extra, "expression operand code out of range");
    if (vj.operand[k] < 0 && first_a < 0) first_a = k;
  }
  NCA_REQUIRE(first_a >= 0, "the expression must read the stencil value");
  for (int j = 0; j < n_extra; ++j) {
    validate_view(extras[j], "extra");
    NCA_REQUIRE(same_shape(extras[j], *src) && is_c_contiguous(extras[j]), "extras must be C-contiguous with src's shape");
    vj.extra_dtypes.push_back(extras[j].dtype);
  }
  for (int k = 0; k < n_operand; ++k)
    if (vj.operand[k] >= 0)
      NCA_REQUIRE(et.arr_dtypes[k] == extras[vj.operand[k]].dtype, "extra dtype differs from the expression operand's");
  vj.alias.resize(n_operand);
  for (int k = 0; k < n_operand; ++k) {
    vj.alias[k] = k;
    for (int q = 0; q < k; ++q)
      if (vj.operand[q] == vj.operand[k]) { vj.alias[k] = q; break; }
  }
  // root 0: d expr / d a (stays on chip); then d expr / d extra_j for every grad_extras[j] with data
  std::vector<void*> outs;
  vj.roots.push_back(differentiate(vj.expr, vj.expr.root, vj.alias[first_a], vj.alias));
  vj.out_dtypes.push_back(p.tape.compute_dtype);
  for (int j = 0; j < n_extra && grad_extras; ++j) {
    if (!grad_extras[j].data) continue;
    validate_view(grad_extras[j], "grad_extra");
    NCA_REQUIRE(same_shape(grad_extras[j], *src) && is_c_contiguous(grad_extras[j]) && is_float(grad_extras[j].dtype),
                "extra gradients must be C-contiguous floating-point arrays of src's shape");
    int var = -1;
    for (int k = 0; k < n_operand && var < 0; ++k)
      if (vj.operand[k] == j) var = vj.alias[k];
    vj.roots.push_back(var >= 0 ? differentiate(vj.expr, vj.expr.root, var, vj.alias) : -1);
    vj.out_dtypes.push_back(grad_extras[j].dtype);
    outs.push_back(grad_extras[j].data);
  }
  const int64_t NZ = src->shape[0], NY = src->shape[1], NX = src->shape[2];
  NCA_REQUIRE(NZ < (int64_t(1) << 30), "too many planes");
  NCA_REQUIRE(NY * NX < (int64_t(1) << 31), "plane too large");
  const int64_t n = view_size(*src);
  hipStream_t s = as_stream(stream);
  if (n == 0) {
    if (fuse) {
      FinalizeArgs fa;
      fa.value_dtype = et.dest_dtype;
      fa.fuse = fuse;
      fa.sum_out = sum_out;
      fa.max_out = max_out;
      fa.min_out = min_out;
      launch_finalize_fused(fa, s);
    }
    finish_call(stream);
    return;
  }
  // 4-point vectors when the last extent is a multiple of 4 and every array is aligned to them
  auto aligned4 = [](const void* ptr, int32_t dt) { return reinterpret_cast<uintptr_t>(ptr) % (4 * itemsize(dt)) == 0; };
  bool vec = NX % 4 == 0 && aligned4(src->data, src->dtype) && aligned4(grad_src->data, grad_src->dtype);
  for (int j = 0; j < n_extra; ++j) vec = vec && aligned4(extras[j].data, extras[j].dtype);
  for (size_t m = 0; m < outs.size(); ++m) vec = vec && aligned4(outs[m], vj.out_dtypes[m + 1]);
  if (const char* e = std::getenv("NCA_HIP_TRAIN_SCALAR")) vec = vec && std::atoi(e) == 0;
  int TXs = 64, TYs = 16;
  train_tile(p.src_dtype, vec, TXs, TYs);
  DeviceState& dev = current_device();
  std::string key = "train|" + et.signature + "|";
  for (int o : vj.operand) key += std::to_string(o) + ",";
  key += "|";
  for (auto d : vj.extra_dtypes) key += std::to_string(d) + ",";
  key += "|";
  for (size_t m = 0; m < vj.roots.size(); ++m) key += std::to_string(vj.out_dtypes[m]) + (vj.roots[m] < 0 ? "z," : ",");
  key += "|dx" + std::to_string(grad_src->dtype) + "|f" + std::to_string(fuse) + "," + std::to_string(fuse ? mode : 0) +
         "|t" + std::to_string(TXs) + "x" + std::to_string(TYs) + (vec ? "|v4" : "");
  const std::string kname =
      std::string("nca_stencil_train_") + dtype_info(p.src_dtype).name + (fuse ? "_sum" : "") + (vec ? "_v4" : "");
  StencilFuse fz;
  fz.with_dst = false;
  fz.fuse = fuse;
  fz.mode = mode;
  hipFunction_t fn = stencil_kernel(
      p, key,
      [&]() {
        return vec ? gen_march_train_vec(p, adj, vj, fz, grad_src->dtype, TXs, TYs)
                   : gen_march_train(p, adj, vj, fz, grad_src->dtype, TXs, TYs);
      },
      kname);
  const int ntx = static_cast<int>((NX + TXs - 1) / TXs), nty = static_cast<int>((NY + TYs - 1) / TYs);
  const int64_t tiles = static_cast<int64_t>(ntx) * nty;
  const int64_t resident = static_cast<int64_t>(dev.n_cus) * blocks_per_cu(fn);
  int64_t nzc = std::max<int64_t>(1, std::min<int64_t>(resident / std::max<int64_t>(tiles, 1), NZ));
  const int zlen = static_cast<int>((NZ + nzc - 1) / nzc);
  nzc = (NZ + zlen - 1) / zlen;
  const int64_t blocks = tiles * nzc;
  NCA_REQUIRE(blocks < (int64_t(1) << 31), "grid too large");
  ScalarSlots ss = make_scalar_slots(p.tape, stencil_scalars, n_stencil_scalars);
  ScalarSlots es = make_scalar_slots(et, expr_scalars, n_expr_scalars);
  std::deque<const void*> ptrs;
  std::vector<void*> params;
  ptrs.push_back(src->data);
  params.push_back(&ptrs.back());
  for (int j = 0; j < n_extra; ++j) { ptrs.push_back(extras[j].data); params.push_back(&ptrs.back()); }
  ptrs.push_back(grad_src->data);
  params.push_back(&ptrs.back());
  for (void* o : outs) { ptrs.push_back(o); params.push_back(&ptrs.back()); }
  int64_t dims[3] = {NZ, NY, NX};
  for (auto& d : dims) params.push_back(&d);
  int ints[3] = {ntx, nty, zlen};
  for (auto& i : ints) params.push_back(&i);
  void* psum = nullptr;
  void* pmax = nullptr;
  void* pmin = nullptr;
  if (fuse) {
    Workspace& ws = workspace(s, static_cast<size_t>(blocks) * 24);
    psum = ws.partials;
    pmax = static_cast<char*>(ws.partials) + static_cast<size_t>(blocks) * 8;
    pmin = static_cast<char*>(ws.partials) + static_cast<size_t>(blocks) * 16;
    params.push_back(&psum);
    params.push_back(&pmax);
    params.push_back(&pmin);
  }
  for (auto& x : ss.slots) params.push_back(&x);
  for (auto& x : es.slots) params.push_back(&x);
  ptrs.push_back(grad_scalar);
  params.push_back(&ptrs.back());
  {
    ProfScope ps(kname.c_str(), s);
    jit_launch(fn, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, params);
  }
  check_launch("stencil_expr_train");
  if (fuse) {
    FinalizeArgs fa;
    fa.value_dtype = et.dest_dtype;
    fa.fuse = fuse;
    fa.psum = psum;
    fa.pmax = pmax;
    fa.pmin = pmin;
    fa.nblocks = static_cast<int>(blocks);
    fa.sum_out = sum_out;
    fa.max_out = max_out;
    fa.min_out = min_out;
    launch_finalize_fused(fa, s);
  }
  finish_call(stream);
}
}  // namespace nca

extern "C" {

int nca_hip_stencil_expr_vjp(uint64_t stencil_handle, uint64_t expr_handle, const int32_t* expr_operand,
                             int32_t n_operand, const nca_view* src, const nca_view* extras, int32_t n_extra,
                             const nca_scalar* stencil_scalars, int32_t n_stencil_scalars, const nca_scalar* expr_scalars,
                             int32_t n_expr_scalars, const void* grad_scalar, const nca_view* w,
                             const nca_view* grad_extras, int32_t boundary, void* stream) {
  return api_call([&] {
    stencil_expr_vjp_impl(stencil_handle, expr_handle, expr_operand, n_operand, src, extras, n_extra, stencil_scalars,
                          n_stencil_scalars, expr_scalars, n_expr_scalars, grad_scalar, w, grad_extras, 0,
                          NCA_ACCUM_PLAIN, nullptr, nullptr, nullptr, boundary, stream);
  });
}

int nca_hip_stencil_expr_train(uint64_t stencil_handle, uint64_t expr_handle, const int32_t* expr_operand,
                               int32_t n_operand, const nca_view* src, const nca_view* extras, int32_t n_extra,
                               const nca_scalar* stencil_scalars, int32_t n_stencil_scalars,
                               const nca_scalar* expr_scalars, int32_t n_expr_scalars, const void* grad_scalar,
                               const nca_view* grad_src, const nca_view* grad_extras, int32_t fuse, int32_t mode,
                               void* sum_out, void* max_out, void* min_out, int32_t boundary, void* stream) {
  return api_call([&] {
    stencil_expr_train_impl(stencil_handle, expr_handle, expr_operand, n_operand, src, extras, n_extra,
                            stencil_scalars, n_stencil_scalars, expr_scalars, n_expr_scalars, grad_scalar, grad_src,
                            grad_extras, fuse, mode, sum_out, max_out, min_out, boundary, stream);
  });
}

int nca_hip_stencil_expr_train_source(uint64_t stencil_handle, uint64_t expr_handle, const int32_t* expr_operand,
                                      int32_t n_operand, const int32_t* extra_dtypes, int32_t n_extra,
                                      int32_t grad_src_dtype, int32_t fuse, int32_t accum_mode, int32_t vec, char* buf,
                                      size_t n) {
  std::string src;
  int rc = api_call([&] {
    StencilProgram& p = stencil_program(stencil_handle);
    const Tape& et = expr_program(expr_handle).tape;
    NCA_REQUIRE(p.ndim == 3 && p.extra_dtypes.empty() && !p.has_ext, "a 3-D stencil that reads only its source");
    NCA_REQUIRE(n_operand == static_cast<int32_t>(et.arr_dtypes.size()) && (n_operand == 0 || expr_operand),
                "one operand code per expression operand");
    NCA_REQUIRE(n_extra >= 0 && n_extra <= 8 && (n_extra == 0 || extra_dtypes), "0..8 extra dtypes");
    MarchVjp vj;
    vj.expr = et;
    vj.value_dtype = et.dest_dtype;
    vj.operand.assign(expr_operand, expr_operand + n_operand);
    vj.extra_dtypes.assign(extra_dtypes, extra_dtypes + n_extra);
    vj.alias.resize(n_operand);
    int first_a = -1;
    for (int k = 0; k < n_operand; ++k) {
      NCA_REQUIRE(vj.operand[k] >= -1 && vj.operand[k] < n_extra, "expression operand code out of range");
      if (vj.operand[k] < 0 && first_a < 0) first_a = k;
      vj.alias[k] = k;
      for (int q = 0; q < k; ++q)
        if (vj.operand[q] == vj.operand[k]) { vj.alias[k] = q; break; }
    }
    NCA_REQUIRE(first_a >= 0, "the expression must read the stencil value");
    vj.roots.push_back(differentiate(vj.expr, vj.expr.root, vj.alias[first_a], vj.alias));
    vj.out_dtypes.push_back(p.tape.compute_dtype);
    for (int j = 0; j < n_extra; ++j) {
      int var = -1;
      for (int k = 0; k < n_operand && var < 0; ++k)
        if (vj.operand[k] == j) var = vj.alias[k];
      vj.roots.push_back(var >= 0 ? differentiate(vj.expr, vj.expr.root, var, vj.alias) : -1);
      vj.out_dtypes.push_back(extra_dtypes[j]);
    }
    StencilFuse fz;
    fz.with_dst = false;
    fz.fuse = fuse;
    fz.mode = accum_mode;
    int TXs = 64, TYs = 16;
    train_tile(p.src_dtype, vec != 0, TXs, TYs);
    const Tape adj = adjoint_tape(p);
    src = device_prelude() + "\n" +
          (vec ? gen_march_train_vec(p, adj, vj, fz, grad_src_dtype, TXs, TYs)
               : gen_march_train(p, adj, vj, fz, grad_src_dtype, TXs, TYs));
  });
  if (rc != NCA_OK) return -rc;
  if (buf && n > 0) {
    size_t k = src.size() < n - 1 ? src.size() : n - 1;
    std::memcpy(buf, src.data(), k);
    buf[k] = '\0';
  }
  return static_cast<int>(src.size());
}

int nca_hip_stencil_expr_vjp_reduce(uint64_t stencil_handle, uint64_t expr_handle, const int32_t* expr_operand,
                                    int32_t n_operand, const nca_view* src, const nca_view* extras, int32_t n_extra,
                                    const nca_scalar* stencil_scalars, int32_t n_stencil_scalars,
                                    const nca_scalar* expr_scalars, int32_t n_expr_scalars, const void* grad_scalar,
                                    const nca_view* w, const nca_view* grad_extras, int32_t fuse, int32_t mode,
                                    void* sum_out, void* max_out, void* min_out, int32_t boundary, void* stream) {
  return api_call([&] {
    NCA_REQUIRE(fuse != 0, "no reduction requested (nca_hip_stencil_expr_vjp computes the gradients alone)");
    stencil_expr_vjp_impl(stencil_handle, expr_handle, expr_operand, n_operand, src, extras, n_extra, stencil_scalars,
                          n_stencil_scalars, expr_scalars, n_expr_scalars, grad_scalar, w, grad_extras, fuse, mode,
                          sum_out, max_out, min_out, boundary, stream);
  });
}

int nca_hip_stencil_adjoint_create(uint64_t stencil_handle, uint64_t* handle) {
  return api_call([&] {
    NCA_REQUIRE(handle != nullptr, "null handle");
    StencilProgram& f = stencil_program(stencil_handle);
    auto p = std::make_unique<StencilProgram>();
    p->ndim = f.ndim;
    p->n_off = f.n_off;
    p->src_dtype = f.tape.dest_dtype;  // the adjoint reads gradients of the destination
    p->forward = stencil_handle;
    p->tape = adjoint_tape(f);
    for (auto& o : f.offsets) {
      std::array<int64_t, NCA_MAX_NDIM> no{};
      for (int d = 0; d < f.ndim; ++d) no[d] = -o[d];
      p->offsets.push_back(no);
    }
    *handle = stencil_register(std::move(p));
  });
}

int nca_hip_stencil_source(uint64_t handle, int32_t variant, char* buf, size_t n) {
  std::string src;
  int rc = api_call([&] {
    StencilProgram& p = stencil_program(handle);
    if (variant == 0 || variant == 1 || variant == 4) {
      NCA_REQUIRE(p.ndim == 3, "the march kernel is 3-D");
      StencilFuse fz;
      if (variant != 0) {
        fz.axis0 = true;
        fz.fuse = NCA_FUSE_MAX;
      }
      if (variant == 4) fz.mode = NCA_ACCUM_COMPENSATED;
      const MarchPlan mp = march_plan(p);
      src = mp.queue ? gen_march_q(p, fz, p.tape.dest_dtype, mp) : gen_march(p, fz, p.tape.dest_dtype);
    } else if (variant == 5 || variant == 6) {
      // the tile kernel at unit steps: 5 plain; 6 with fused sum/max/min (compensated) and the
      // first axis read through a table of pointers in the source and the destination
      const int lift = p.ndim == 1 ? 1 : 0;
      TileSpec ts;
      ts.nd = p.ndim + lift;
      for (auto& o : p.offsets) {
        std::array<int64_t, NCA_MAX_NDIM> q{};
        for (int d = 0; d < p.ndim; ++d) q[d + lift] = o[d];
        ts.offsets.push_back(q);
      }
      ts.dst_dtype = p.tape.dest_dtype;
      ts.vw = static_cast<int>(16 / itemsize(p.src_dtype));
      if (variant == 6) {
        ts.fuse = NCA_FUSE_SUM | NCA_FUSE_MAX | NCA_FUSE_MIN;
        ts.mode = NCA_ACCUM_COMPENSATED;
        ts.src_ind = {0};
        ts.dst_ind = {0};
      }
      ts.derive();
      src = gen_tile(p, ts);
    } else {
      src = gen_direct(p, p.tape.dest_dtype, variant == 3, 0, 0);
    }
    src = device_prelude() + "\n" + src;
  });
  if (rc != NCA_OK) return -rc;
  if (buf && n > 0) {
    size_t k = src.size() < n - 1 ? src.size() : n - 1;
    std::memcpy(buf, src.data(), k);
    buf[k] = '\0';
  }
  return static_cast<int>(src.size());
}

}  // extern "C"
