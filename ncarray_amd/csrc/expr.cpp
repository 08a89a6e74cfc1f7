// Fused elementwise expressions: hipRTC code generation and launch.  Replaces
// RuntimeCompiler::get_expr_kernel / get_expression_kernel_str / get_dynamic_vm_kernel
// (rtcompiler.cc:49-89, 213-366) and GPUEngine::execute_binary_expression
// (gpuengine.hh:90-195).
//
// Every tape, of any depth, becomes straight-line code (no stack VM).  The output shape
// is a launch argument, not baked into the source, so one compiled kernel serves every
// shape.  Variants:
//   contiguous: 16-byte vector loads/stores, grid-stride loop, scalar tail;
//   strided:    per-element unravel of the output index over the (jointly coalesced)
//               shape, byte strides per operand;
//   fused reduce: the contiguous/strided body plus per-thread sum/max/min of the stored
//               values, wave butterflies, per-block partials, then a combine kernel;
//   vjp:        derivative tapes (symbolic) times the upstream gradient.
#include <cstdio>
#include <sstream>
#include <tuple>
#include <unordered_map>

#include "expr_internal.hh"
#include "nca_internal.hh"
#include "tape.hh"

namespace nca {

namespace {
std::mutex g_progs_mu;
std::unordered_map<uint64_t, std::unique_ptr<ExprProgram>> g_progs;
uint64_t g_next_handle = 1;
}  // namespace

ExprProgram& expr_program(uint64_t h) {
  std::lock_guard<std::mutex> lk(g_progs_mu);
  auto it = g_progs.find(h);
  NCA_REQUIRE(it != g_progs.end(), "unknown expression handle " + std::to_string(h));
  return *it->second;
}

uint64_t expr_register(Tape&& t) {
  auto p = std::make_unique<ExprProgram>();
  p->tape = std::move(t);
  std::lock_guard<std::mutex> lk(g_progs_mu);
  uint64_t h = g_next_handle++;
  g_progs[h] = std::move(p);
  return h;
}

// ---- scalars -------------------------------------------------------------------------------
ScalarSlots make_scalar_slots(const Tape& t, const nca_scalar* s, int n) {
  NCA_REQUIRE(n == static_cast<int>(t.scalar_dtypes.size()),
              "expected " + std::to_string(t.scalar_dtypes.size()) + " scalars, got " + std::to_string(n));
  NCA_REQUIRE(n == 0 || s != nullptr, "null scalars");
  ScalarSlots out;
  out.slots.resize(n);
  const int32_t cd = t.compute_dtype;
  for (int i = 0; i < n; ++i) {
    int32_t dt = t.scalar_dtypes[i];
    // value as the widest type of its kind, then static_cast to the compute type
    // (op_traits<T>::cast, op_traits.hh:93-117)
    long double v;
    bool is_u = !is_float(dt) && !dtype_info(dt).is_signed;
    if (is_float(dt)) v = static_cast<long double>(s[i].f);
    else if (is_u) v = static_cast<long double>(s[i].u);
    else v = static_cast<long double>(s[i].i);
    uint64_t bits = 0;
    auto put = [&](auto x) { std::memcpy(&bits, &x, sizeof(x)); };
    switch (cd) {
      case NCA_BOOL: put(static_cast<bool>(is_float(dt) ? s[i].f != 0.0 : (is_u ? s[i].u != 0 : s[i].i != 0))); break;
      case NCA_CHAR: case NCA_INT8: put(static_cast<int8_t>(is_float(dt) ? static_cast<int64_t>(s[i].f) : s[i].i)); break;
      case NCA_UINT8: put(static_cast<uint8_t>(is_float(dt) ? static_cast<uint64_t>(s[i].f) : s[i].u)); break;
      case NCA_INT16: put(static_cast<int16_t>(is_float(dt) ? static_cast<int64_t>(s[i].f) : s[i].i)); break;
      case NCA_UINT16: put(static_cast<uint16_t>(is_float(dt) ? static_cast<uint64_t>(s[i].f) : s[i].u)); break;
      case NCA_INT32: put(static_cast<int32_t>(is_float(dt) ? static_cast<int64_t>(s[i].f) : s[i].i)); break;
      case NCA_UINT32: put(static_cast<uint32_t>(is_float(dt) ? static_cast<uint64_t>(s[i].f) : s[i].u)); break;
      case NCA_INT64: put(static_cast<int64_t>(is_float(dt) ? static_cast<int64_t>(s[i].f) : s[i].i)); break;
      case NCA_UINT64: put(static_cast<uint64_t>(is_float(dt) ? static_cast<uint64_t>(s[i].f) : s[i].u)); break;
      case NCA_FLOAT32: put(static_cast<float>(v)); break;
      case NCA_FLOAT64: put(static_cast<double>(v)); break;
      default: NCA_THROW(NCA_ERR_UNSUPPORTED, "scalar compute dtype");
    }
    out.slots[i] = bits;
  }
  return out;
}

// ---- operand analysis ----------------------------------------------------------------------
static bool views_alias(const nca_view& a, const nca_view& b) {
  if (a.dtype != b.dtype || a.ndim != b.ndim) return false;
  const Chain ca = chain_of(a), cb = chain_of(b);
  if (ca.base != cb.base || ca.axes != cb.axes || ca.subs != cb.subs) return false;
  for (int d = 0; d < a.ndim; ++d)
    if (a.shape[d] != b.shape[d] || (a.shape[d] != 1 && ca.strides[d] != cb.strides[d])) return false;
  return true;
}

Operands analyse_operands(const Tape& t, const nca_view* ins, int n_in, const nca_view& out,
                          bool out_may_be_null) {
  NCA_REQUIRE(n_in == static_cast<int>(t.arr_dtypes.size()),
              "expected " + std::to_string(t.arr_dtypes.size()) + " array operands, got " + std::to_string(n_in));
  NCA_REQUIRE(n_in == 0 || ins != nullptr, "null operand views");
  validate_view_nc(out, "out", out_may_be_null);
  Operands ops;
  ops.n = view_size(out);
  ops.alias.resize(n_in);
  ops.bcast.assign(n_in, 0);
  ops.base.assign(n_in, nullptr);
  ops.ind.assign(n_in, {});
  ops.isub.assign(n_in, {});
  ops.baseo = out.data;
  for (int k = 0; k < n_in; ++k) {
    const nca_view& v = ins[k];
    validate_view_nc(v, "operand");
    NCA_REQUIRE(v.dtype == t.arr_dtypes[k], "operand " + std::to_string(k) + " dtype " +
                                                dtype_info(v.dtype).name + " does not match the tape's " +
                                                dtype_info(t.arr_dtypes[k]).name);
    ops.base[k] = v.data;
    if (!view_indirect(v) && (v.ndim == 0 || (is_broadcast_scalar(v) && view_size(v) <= 1) ||
                               (is_broadcast_scalar(v) && v.ndim == out.ndim && same_shape(v, out) && ops.n > 1))) {
      ops.bcast[k] = 1;
    } else {
      NCA_REQUIRE(same_shape(v, out), "operand " + std::to_string(k) + " shape does not match the output shape");
    }
    ops.alias[k] = k;
    for (int j = 0; j < k; ++j)
      if (ops.alias[j] == j && ops.bcast[j] == ops.bcast[k] && views_alias(ins[j], v)) {
        ops.alias[k] = j;
        break;
      }
  }
  for (int k = 0; k < n_in; ++k)
    if (ops.alias[k] == k) ops.unique.push_back(k);
  // contiguous kind: output and every non-broadcast operand C-contiguous
  bool contig = out.data == nullptr ? true : is_c_contiguous(out);
  int max_item = out.data ? itemsize(out.dtype) : 1;
  for (int k : ops.unique) {
    if (ops.bcast[k]) continue;
    contig = contig && is_c_contiguous(ins[k]);
    max_item = std::max(max_item, itemsize(ins[k].dtype));
  }
  ops.contiguous = contig;
  int vec = 16 / max_item;
  if (vec < 1) vec = 1;
  auto aligned = [&](const void* p, int item) {
    return (reinterpret_cast<uintptr_t>(p) % static_cast<uintptr_t>(vec * item)) == 0;
  };
  if (out.data && !aligned(out.data, itemsize(out.dtype))) vec = 1;
  for (int k : ops.unique)
    if (!ops.bcast[k] && !aligned(ins[k].data, itemsize(ins[k].dtype))) vec = 1;
  ops.vec = contig ? vec : 1;
  // Indirect views (NC pointer axis, SO suboffsets): strided kind over the output's own axes
  // (no coalescing across an indirect axis)
  bool nc = out.data && view_indirect(out);
  for (int k : ops.unique)
    if (!ops.bcast[k] && view_indirect(ins[k])) nc = true;
  if (nc) {
    ops.nc = true;
    ops.contiguous = false;
    ops.vec = 1;
    const int nd = std::max(out.ndim, 1);
    ops.sndim = nd;
    ops.sshape.assign(nd, 1);
    ops.sout.assign(nd, 0);
    for (int d = 0; d < out.ndim; ++d) ops.sshape[d] = out.shape[d];
    if (out.data) {
      const Chain c = chain_of(out);
      for (int d = 0; d < out.ndim; ++d) ops.sout[d] = c.strides[d];
      ops.baseo = reinterpret_cast<const void*>(c.base);
      ops.indo = c.axes;
      ops.isubo = c.subs;
    }
    ops.sstrides.assign(n_in, std::vector<int64_t>(nd, 0));
    for (int k : ops.unique) {
      if (ops.bcast[k]) continue;
      const Chain c = chain_of(ins[k]);
      for (int d = 0; d < ins[k].ndim; ++d) ops.sstrides[k][d] = c.strides[d];
      ops.base[k] = reinterpret_cast<const void*>(c.base);
      ops.ind[k] = c.axes;
      ops.isub[k] = c.subs;
    }
    return ops;
  }
  // strided: jointly coalesce output + non-broadcast unique operands
  if (!contig) {
    int nd = out.ndim;
    std::vector<std::array<int64_t, NCA_MAX_NDIM>> st;
    std::array<int64_t, NCA_MAX_NDIM> sh{};
    for (int d = 0; d < nd; ++d) sh[d] = out.shape[d];
    std::vector<int> members;
    for (int k : ops.unique) if (!ops.bcast[k]) members.push_back(k);
    std::array<int64_t, NCA_MAX_NDIM> so{};
    for (int d = 0; d < nd; ++d) so[d] = out.data ? out.strides[d] : 0;
    st.resize(members.size());
    for (size_t m = 0; m < members.size(); ++m)
      for (int d = 0; d < nd; ++d) st[m][d] = ins[members[m]].strides[d];
    // drop size-1 axes, then merge axes d-1,d when every view is contiguous across them
    std::vector<int> axes;
    for (int d = 0; d < nd; ++d) if (sh[d] != 1) axes.push_back(d);
    std::vector<int64_t> csh;
    std::vector<int64_t> cso;
    std::vector<std::vector<int64_t>> cst(members.size());
    for (int d : axes) {
      bool merge = !csh.empty();
      if (merge) {
        if (cso.back() != so[d] * sh[d]) merge = false;
        for (size_t m = 0; m < members.size() && merge; ++m)
          if (cst[m].back() != st[m][d] * sh[d]) merge = false;
      }
      if (merge) {
        csh.back() *= sh[d];
        cso.back() = so[d];
        for (size_t m = 0; m < members.size(); ++m) cst[m].back() = st[m][d];
      } else {
        csh.push_back(sh[d]);
        cso.push_back(so[d]);
        for (size_t m = 0; m < members.size(); ++m) cst[m].push_back(st[m][d]);
      }
    }
    if (csh.empty()) { csh.push_back(1); cso.push_back(0); for (auto& c : cst) c.push_back(0); }
    ops.sndim = static_cast<int>(csh.size());
    ops.sshape = csh;
    ops.sout = cso;
    ops.sstrides.assign(n_in, std::vector<int64_t>(ops.sndim, 0));
    for (size_t m = 0; m < members.size(); ++m) ops.sstrides[members[m]] = cst[m];
  }
  return ops;
}

// ---- code generation -------------------------------------------------------------------------
static std::string sig_operands(const Tape& t, const Operands& ops) {
  std::ostringstream os;
  for (size_t k = 0; k < ops.alias.size(); ++k) os << ops.alias[k] << (ops.bcast[k] ? "b" : "") << ",";
  if (ops.nc) {
    auto axes = [&](const std::vector<int>& a) {
      os << "[";
      for (int x : a) os << x << ".";
      os << "]";
    };
    os << "nc";
    for (int k : ops.unique) axes(ops.ind[k]);
    os << "o";
    axes(ops.indo);
  }
  return os.str();
}

// Parameters: unique operand pointers, output, n, scalars.
static std::string param_list(const Tape& t, const Operands& ops, bool strided, const std::string& out_t,
                              bool with_out) {
  std::ostringstream os;
  const std::string C = ctype_of(t.compute_dtype);
  bool first = true;
  auto sep = [&]() { if (!first) os << ", "; first = false; };
  for (int k : ops.unique) {
    sep();
    if (strided && !ops.bcast[k]) os << "const char* __restrict__ p" << k;
    else os << "const " << ctype_of(t.arr_dtypes[k]) << "* __restrict__ p" << k;
  }
  if (with_out) {
    sep();
    if (strided) os << "char* __restrict__ outp";
    else os << out_t << "* __restrict__ out";
  }
  if (strided) {
    sep();
    os << "nca_dims shp";
    for (int k : ops.unique) if (!ops.bcast[k]) os << ", nca_dims st" << k;
    if (with_out) os << ", nca_dims sto";
    // indirect views: the byte offset added after each pointer read
    for (int k : ops.unique)
      if (!ops.bcast[k] && ops.nc)
        for (size_t j = 0; j < ops.ind[k].size(); ++j) os << ", nca_i64 q" << k << "_" << j;
    if (with_out && ops.nc)
      for (size_t j = 0; j < ops.indo.size(); ++j) os << ", nca_i64 qo_" << j;
  }
  sep();
  os << "nca_i64 n";
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  return os.str();
}

static std::vector<std::string> load_names(const Tape& t) {
  std::vector<std::string> v;
  for (size_t k = 0; k < t.arr_dtypes.size(); ++k) v.push_back("u" + std::to_string(k));
  return v;
}
static std::vector<std::string> scalar_names(const Tape& t) {
  std::vector<std::string> v;
  for (size_t k = 0; k < t.scalar_dtypes.size(); ++k) v.push_back("s" + std::to_string(k));
  return v;
}

// Broadcast operands are loaded once per thread before the loop.
static std::string bcast_loads(const Tape& t, const Operands& ops) {
  std::ostringstream os;
  const std::string C = ctype_of(t.compute_dtype);
  for (int k : ops.unique)
    if (ops.bcast[k])
      os << "  const " << C << " u" << k << " = nca_ld<" << C << ">(p" << k << ");\n";
  return os.str();
}

static std::string strided_prelude(int nd) {
  std::ostringstream os;
  os << "#define NCA_ND " << nd << "\nstruct nca_dims { nca_i64 v[NCA_ND]; };\n";
  return os.str();
}

// Body that computes offsets of element `idx` for the strided variant.  A plain view
// accumulates one byte offset (o<k>, oo).  An indirect view with steps at axes a_0 < ... < a_{m-1}
// accumulates m+1 segment offsets: o<k>_0 over axes 0..a_0, o<k>_j over a_{j-1}+1..a_j and
// o<k>_m over the axes after a_{m-1}; the address then alternates segment additions and pointer
// reads (operand_address).
static void segment_update(std::ostringstream& os, const std::vector<int>& a, const std::string& o,
                           const std::string& st) {
  os << "        ";
  for (size_t j = 0; j < a.size(); ++j)
    os << "if (d <= " << a[j] << ") " << o << "_" << j << " += c * " << st << ".v[d];\n        else ";
  os << o << "_" << a.size() << " += c * " << st << ".v[d];\n";
}

static std::string strided_offsets(const Tape& t, const Operands& ops, bool with_out,
                                   const std::string& var = "idx", bool skip_last = false) {
  std::ostringstream os;
  auto indirect = [&](int k) { return ops.nc && !ops.ind[k].empty(); };
  const bool indo = with_out && ops.nc && !ops.indo.empty();
  auto decl = [&](const std::string& o, size_t m) {
    os << "      nca_i64 " << o << "_0 = 0";
    for (size_t j = 1; j <= m; ++j) os << ", " << o << "_" << j << " = 0";
    os << ";\n";
  };
  os << "      nca_i64 rem = " << var << ";\n";
  for (int k : ops.unique)
    if (!ops.bcast[k]) {
      if (indirect(k)) decl("o" + std::to_string(k), ops.ind[k].size());
      else os << "      nca_i64 o" << k << " = 0;\n";
    }
  if (with_out) {
    if (indo) decl("oo", ops.indo.size());
    else os << "      nca_i64 oo = 0;\n";
  }
  // the index is below the product of the extents it is unravelled over, so what remains at
  // axis 0 is its coordinate (no divide there; d is a constant after unrolling)
  os << "#pragma unroll\n      for (int d = NCA_ND - " << (skip_last ? 2 : 1) << "; d >= 0; --d) {\n"
     << "        const nca_i64 c = d == 0 ? rem : rem % shp.v[d];\n        if (d > 0) rem /= shp.v[d];\n";
  for (int k : ops.unique) {
    if (ops.bcast[k]) continue;
    if (indirect(k)) segment_update(os, ops.ind[k], "o" + std::to_string(k), "st" + std::to_string(k));
    else os << "        o" << k << " += c * st" << k << ".v[d];\n";
  }
  if (with_out) {
    if (indo) segment_update(os, ops.indo, "oo", "sto");
    else os << "        oo += c * sto.v[d];\n";
  }
  os << "      }\n";
  return os.str();
}

// Address of an element: start + segment 0, then per step a pointer read plus its byte offset
// and the next segment.
static std::string chain_address(const std::string& p, const std::string& o, const std::string& q, size_t m) {
  std::string a = "(" + p + " + " + o + "_0)";
  for (size_t j = 0; j < m; ++j)
    a = "(nca_chase(" + a + ", " + q + std::to_string(j) + ") + " + o + "_" + std::to_string(j + 1) + ")";
  return a;
}

// Address expression of operand k (strided kind).
static std::string operand_address(const Operands& ops, int k) {
  const std::string ks = std::to_string(k);
  if (ops.nc && !ops.ind[k].empty()) return chain_address("p" + ks, "o" + ks, "q" + ks + "_", ops.ind[k].size());
  return "(p" + ks + " + o" + ks + ")";
}

struct ReduceSpec {
  int fuse = 0;
  int mode = NCA_ACCUM_PLAIN;
};

// Reduction accumulator declarations/updates/finalisation for fused kernels.
// The value reduced is `st` (of destination type D) converted to RC.
struct ReduceCode {
  std::string decl, update, finish;
};

static ReduceCode reduce_code(int32_t dest, const ReduceSpec& rs) {
  ReduceCode rc;
  if (!rs.fuse) return rc;
  const int32_t rcd = compute_dtype(dest);
  const std::string RC = ctype_of(rcd);
  const bool fsum = is_float(rcd);
  const std::string SA = fsum ? RC : (dtype_info(rcd).is_signed ? "long long" : "unsigned long long");
  std::ostringstream d, u, f;
  if (rs.fuse & NCA_FUSE_SUM) {
    d << "  " << SA << " acc_s = 0;\n";
    if (fsum && rs.mode == NCA_ACCUM_COMPENSATED) d << "  " << SA << " acc_c = 0;\n";
  }
  if (rs.fuse & NCA_FUSE_MAX) d << "  " << RC << " acc_mx = nca_ident<" << ctype_of(dest) << ", " << RC << ">::lowest();\n";
  if (rs.fuse & NCA_FUSE_MIN) d << "  " << RC << " acc_mn = nca_ident<" << ctype_of(dest) << ", " << RC << ">::max();\n";
  u << "      { const " << RC << " rv = nca_cvt<" << RC << ", " << ctype_of(dest) << ">::from(st);\n";
  if (rs.fuse & NCA_FUSE_SUM) {
    if (fsum && rs.mode == NCA_ACCUM_COMPENSATED) u << "        nca_neumaier_add(acc_s, acc_c, (" << SA << ")rv);\n";
    else u << "        acc_s += (" << SA << ")rv;\n";
  }
  if (rs.fuse & NCA_FUSE_MAX) u << "        acc_mx = nca_maxop(acc_mx, rv);\n";
  if (rs.fuse & NCA_FUSE_MIN) u << "        acc_mn = nca_minop(acc_mn, rv);\n";
  u << "      }\n";
  // block combine: waves -> LDS -> thread 0 -> partial slot of this block
  f << "  __shared__ double red_sd[16];\n  __shared__ " << SA << " red_si[16];\n  __shared__ " << RC
    << " red_mx[16];\n  __shared__ " << RC << " red_mn[16];\n"
    << "  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;\n";
  if (rs.fuse & NCA_FUSE_SUM) {
    if (fsum) {
      if (rs.mode == NCA_ACCUM_COMPENSATED) f << "  nca_wave_comp(acc_s, acc_c);\n  if (lane == 0) red_sd[wid] = nca_comp_value(acc_s, acc_c);\n";
      else f << "  acc_s = nca_wave_sum(acc_s);\n  if (lane == 0) red_sd[wid] = (double)acc_s;\n";
    } else {
      f << "  acc_s = nca_wave_sum(acc_s);\n  if (lane == 0) red_si[wid] = acc_s;\n";
    }
  }
  if (rs.fuse & NCA_FUSE_MAX) f << "  acc_mx = nca_wave_max(acc_mx);\n  if (lane == 0) red_mx[wid] = acc_mx;\n";
  if (rs.fuse & NCA_FUSE_MIN) f << "  acc_mn = nca_wave_min(acc_mn);\n  if (lane == 0) red_mn[wid] = acc_mn;\n";
  f << "  __syncthreads();\n  if (threadIdx.x == 0) {\n";
  if (rs.fuse & NCA_FUSE_SUM) {
    if (fsum) f << "    double ts = 0.0;\n    for (int w = 0; w < nw; ++w) ts += red_sd[w];\n    ((double*)psum)[blockIdx.x] = ts;\n";
    else f << "    " << SA << " ts = 0;\n    for (int w = 0; w < nw; ++w) ts += red_si[w];\n    ((" << SA << "*)psum)[blockIdx.x] = ts;\n";
  }
  if (rs.fuse & NCA_FUSE_MAX) f << "    " << RC << " tm = red_mx[0];\n    for (int w = 1; w < nw; ++w) tm = nca_maxop(tm, red_mx[w]);\n    *(" << RC << "*)((char*)pmax + 8 * (size_t)blockIdx.x) = tm;\n";
  if (rs.fuse & NCA_FUSE_MIN) f << "    " << RC << " tn = red_mn[0];\n    for (int w = 1; w < nw; ++w) tn = nca_minop(tn, red_mn[w]);\n    *(" << RC << "*)((char*)pmin + 8 * (size_t)blockIdx.x) = tn;\n";
  f << "  }\n";
  rc.decl = d.str();
  rc.update = u.str();
  rc.finish = f.str();
  return rc;
}

std::string gen_elementwise(const Tape& t, const Operands& ops, bool strided, int vec, int32_t out_dtype,
                            bool with_out, int fuse, int mode, const std::string& kname) {
  const std::string C = ctype_of(t.compute_dtype);
  const std::string D = ctype_of(out_dtype);
  std::vector<std::string> roots;
  std::string body = emit_nodes(t, {t.root}, ops.alias, load_names(t), scalar_names(t), "lin", "t", roots);
  ReduceSpec rs{fuse, mode};
  ReduceCode rc = reduce_code(out_dtype, rs);
  std::ostringstream os;
  if (strided) os << strided_prelude(ops.sndim);
  os << "// tape: " << t.signature << (t.linearizable ? " (linearizable)" : " (vm order)") << "\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void " << kname << "("
     << param_list(t, ops, strided, D, true);
  if (fuse) os << ", void* __restrict__ psum, void* __restrict__ pmax, void* __restrict__ pmin";
  os << ") {\n";
  os << bcast_loads(t, ops);
  os << rc.decl;
  os << "  const nca_i64 gstride = (nca_i64)gridDim.x * blockDim.x;\n";
  bool reads_array = false;
  for (int k : ops.unique)
    if (!ops.bcast[k]) reads_array = true;
  if (!strided && !reads_array && !fuse && with_out) {
    // An expression that reads no array (a fill, an index generator) is a pure store stream.
    // Stores need many in flight per lane to reach the write bandwidth (measured on MI355X for
    // 512 MiB: one 16-byte store per lane per step 4.6 TB/s, eight about 5.5 TB/s), so each wave
    // writes a contiguous chunk of 8 x 64 vectors per step, streaming (nontemporal) stores.
    os << "  constexpr int V = " << vec << ";\n  constexpr int U = 8;\n";
    os << "  const nca_i64 nv = n / V;\n  const int lane = threadIdx.x & 63;\n";
    os << "  const nca_i64 nwv = (nca_i64)gridDim.x * (blockDim.x >> 6);\n";
    os << "  const nca_i64 nch = (nv + U * 64 - 1) / (U * 64);\n";
    os << "  for (nca_i64 c = (nca_i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < nch; c += nwv) {\n";
    os << "#pragma unroll\n    for (int u = 0; u < U; ++u) {\n";
    os << "      const nca_i64 i = (c * U + u) * 64 + lane;\n      if (i < nv) {\n";
    os << "        nca_vec<" << D << ", V> r;\n";
    os << "#pragma unroll\n        for (int j = 0; j < V; ++j) {\n";
    if (t.uses_idx) os << "      const nca_i64 lin = i * V + j;\n";
    os << body;
    os << "      const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
    os << "          r.v[j] = st;\n        }\n";
    os << "        if (out) nca_store_nt(reinterpret_cast<nca_vec<" << D << ", V>*>(out + i * V), r);\n";
    os << "      }\n    }\n  }\n";
    os << "  for (nca_i64 i = nv * V + (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {\n";
    if (t.uses_idx) os << "      const nca_i64 lin = i;\n";
    os << body;
    os << "      const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
    os << "      if (out) out[i] = st;\n";
    os << "  }\n";
  } else if (!strided) {
    // streaming (nontemporal) loads: each operand vector is read once here; measured on MI355X,
    // 512 MiB float32 -> BF16 cast 0.137 -> 0.121 ms, plain copy 0.185 -> 0.180 ms, the C0
    // expression on cache-resident operands 11.0 -> 10.4 us (stores stay cached: an output is
    // usually the next operator's input)
    os << "  constexpr int V = " << vec << ";\n";
    os << "  const nca_i64 nv = n / V;\n";
    os << "  for (nca_i64 i = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gstride) {\n";
    for (int k : ops.unique)
      if (!ops.bcast[k])
        os << "    const nca_vec<" << ctype_of(t.arr_dtypes[k]) << ", V> x" << k << " = nca_load_nt(reinterpret_cast<const nca_vec<"
           << ctype_of(t.arr_dtypes[k]) << ", V>*>(p" << k << " + i * V));\n";
    os << "    nca_vec<" << D << ", V> r;\n";
    os << "#pragma unroll\n    for (int j = 0; j < V; ++j) {\n";
    for (int k : ops.unique)
      if (!ops.bcast[k])
        os << "      const " << C << " u" << k << " = nca_cvt<" << C << ", " << ctype_of(t.arr_dtypes[k]) << ">::from(x" << k
           << ".v[j]);\n";
    if (t.uses_idx) os << "      const nca_i64 lin = i * V + j;\n";
    os << body;
    os << "      const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
    os << "      r.v[j] = st;\n";
    os << rc.update;
    os << "    }\n";
    if (with_out) os << "    if (out) *reinterpret_cast<nca_vec<" << D << ", V>*>(out + i * V) = r;\n";
    os << "  }\n";
    // scalar tail
    os << "  for (nca_i64 i = nv * V + (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {\n";
    for (int k : ops.unique)
      if (!ops.bcast[k]) os << "      const " << C << " u" << k << " = nca_ld<" << C << ">(p" << k << " + i);\n";
    if (t.uses_idx) os << "      const nca_i64 lin = i;\n";
    os << body;
    os << "      const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
    os << "      if (out) out[i] = st;\n";
    os << rc.update;
    os << "  }\n";
  } else {
    os << "  for (nca_i64 idx = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gstride) {\n";
    os << strided_offsets(t, ops, true);
    for (int k : ops.unique)
      if (!ops.bcast[k])
        os << "      const " << C << " u" << k << " = nca_ld<" << C << ">(reinterpret_cast<const "
           << ctype_of(t.arr_dtypes[k]) << "*>" << operand_address(ops, k) << ");\n";
    if (t.uses_idx) os << "      const nca_i64 lin = idx;\n";
    os << body;
    os << "      const " << D << " st = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n";
    if (ops.nc && !ops.indo.empty())
      os << "      if (outp) *reinterpret_cast<" << D << "*>" << chain_address("outp", "oo", "qo_", ops.indo.size())
         << " = st;\n";
    else
      os << "      if (outp) *reinterpret_cast<" << D << "*>(outp + oo) = st;\n";
    os << rc.update;
    os << "  }\n";
  }
  os << rc.finish;
  os << "}\n";
  return os.str();
}

// ---- launch helpers -------------------------------------------------------------------------------
int grid_blocks(int64_t work_items, int threads, int max_blocks) {
  int64_t b = (work_items + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return static_cast<int>(b);
}

struct LaunchArgs {
  std::deque<const void*> ptrs;             // storage for pointer params (stable addresses)
  std::deque<std::array<int64_t, NCA_MAX_NDIM>> dims;
  std::deque<int64_t> scal;                 // byte offsets after the pointer reads of indirect views
  int64_t n = 0;
  std::vector<void*> params;
};

static void push_common(LaunchArgs& la, const Tape& t, const Operands& ops, const nca_view* ins,
                        const nca_view& out, bool strided, bool with_out, ScalarSlots& sc) {
  la.params.clear();
  for (int k : ops.unique) {
    la.ptrs.push_back(ops.base[k]);
    la.params.push_back(&la.ptrs.back());
  }
  if (with_out) {
    la.ptrs.push_back(ops.baseo);
    la.params.push_back(&la.ptrs.back());
  }
  if (strided) {
    std::array<int64_t, NCA_MAX_NDIM> a{};
    for (int d = 0; d < ops.sndim; ++d) a[d] = ops.sshape[d];
    la.dims.push_back(a);
    la.params.push_back(la.dims.back().data());
    for (int k : ops.unique)
      if (!ops.bcast[k]) {
        std::array<int64_t, NCA_MAX_NDIM> s{};
        for (int d = 0; d < ops.sndim; ++d) s[d] = ops.sstrides[k][d];
        la.dims.push_back(s);
        la.params.push_back(la.dims.back().data());
      }
    if (with_out) {
      std::array<int64_t, NCA_MAX_NDIM> s{};
      for (int d = 0; d < ops.sndim; ++d) s[d] = ops.sout[d];
      la.dims.push_back(s);
      la.params.push_back(la.dims.back().data());
    }
    if (ops.nc) {
      for (int k : ops.unique)
        if (!ops.bcast[k])
          for (int64_t q : ops.isub[k]) {
            la.scal.push_back(q);
            la.params.push_back(&la.scal.back());
          }
      if (with_out)
        for (int64_t q : ops.isubo) {
          la.scal.push_back(q);
          la.params.push_back(&la.scal.back());
        }
    }
  }
  la.n = ops.n;
  la.params.push_back(&la.n);
  for (auto& s : sc.slots) la.params.push_back(&s);
}

hipFunction_t program_kernel(ExprProgram& p, const std::string& key,
                             const std::function<std::string()>& gen, const std::string& kname) {
  {
    std::lock_guard<std::mutex> lk(p.mu);
    auto it = p.kernels.find(key);
    if (it != p.kernels.end()) return it->second;
  }
  std::string src = gen();
  hipFunction_t fn = jit_get(src, kname);
  std::lock_guard<std::mutex> lk(p.mu);
  p.kernels[key] = fn;
  p.sources[key] = src;
  return fn;
}

// ---- tiled transpose form ------------------------------------------------------------------------
// A strided expression whose output runs unit-stride along its last coalesced axis L while an
// operand runs unit-stride along another axis A (a permuted copy such as x.transpose(0, 2) into a
// dense array) reads that operand 4 bytes per 64-lane row at a large stride in the plain strided
// form.  The tiled form stages TA x TL (A, L) tiles of every such operand through LDS, reading
// them coalesced along A, and then computes and stores coalesced along L.  Operands already
// unit-stride along L are read directly in the second phase.
struct TransposePlan {
  int axis = -1;          // A, or -1: another form
  int ta = 64, tl = 64;   // tile extents along A and L (powers of two, 16..256)
  bool pair = false;      // 2-byte data moved as 4-byte pairs (see transpose_plan)
  std::vector<char> lds;  // operand k is staged through LDS
};

// Tile extents: 64 x 64 unless NCA_HIP_TRANSPOSE_TILE = "<TA>x<TL>" names two powers of two in
// 16..256.
static std::pair<int, int> transpose_tile() {
  static const std::pair<int, int> tile = [] {
    std::pair<int, int> d{64, 64};
    const char* e = std::getenv("NCA_HIP_TRANSPOSE_TILE");
    int a = 0, l = 0;
    if (!e || std::sscanf(e, "%dx%d", &a, &l) != 2) return d;
    auto ok = [](int v) { return v >= 16 && v <= 256 && (v & (v - 1)) == 0; };
    return ok(a) && ok(l) ? std::pair<int, int>{a, l} : d;
  }();
  return tile;
}

static TransposePlan transpose_plan(const Tape& t, const Operands& ops, const nca_view& out, int fuse) {
  TransposePlan tp;
  if (ops.contiguous || fuse || ops.nc || t.uses_idx || ops.sndim < 2) return tp;
  const int L = ops.sndim - 1;
  if (ops.sout[L] != itemsize(out.dtype) || ops.sshape[L] < 16) return tp;
  for (int k : ops.unique) {
    if (ops.bcast[k]) continue;
    const int64_t it = itemsize(t.arr_dtypes[k]);
    if (ops.sstrides[k][L] == it) continue;
    for (int a = 0; a < L && tp.axis < 0; ++a)
      if (ops.sstrides[k][a] == it && ops.sshape[a] >= 16) tp.axis = a;
    if (tp.axis >= 0) break;
  }
  if (tp.axis < 0) return tp;
  std::tie(tp.ta, tp.tl) = transpose_tile();
  tp.lds.assign(t.arr_dtypes.size(), 0);
  int64_t lds_bytes = 0;
  for (int k : ops.unique) {
    if (ops.bcast[k]) continue;
    const int64_t it = itemsize(t.arr_dtypes[k]);
    const int64_t b = it * tp.tl * (tp.ta + 1);
    if (ops.sstrides[k][tp.axis] == it && ops.sstrides[k][L] != it && lds_bytes + b <= (int64_t(64) << 10)) {
      tp.lds[k] = 1;
      lds_bytes += b;
    }
  }
  if (lds_bytes == 0) tp.axis = -1;  // no tile fits: another form
  // 2-byte output and staged operands with 4-byte aligned bases and strides (the unit axis
  // aside): each lane loads two A-neighbours and stores two L-neighbours as one 4-byte word
  auto al4 = [](const void* p) { return reinterpret_cast<uintptr_t>(p) % 4 == 0; };
  tp.pair = tp.axis >= 0 && tp.ta >= 32 && tp.tl >= 32 && itemsize(out.dtype) == 2 && al4(ops.baseo);
  for (int d = 0; d < ops.sndim && tp.pair; ++d)
    if (d != L && ops.sout[d] % 4 != 0) tp.pair = false;
  for (int k : ops.unique) {
    if (!tp.pair || ops.bcast[k] || !tp.lds[k]) continue;
    if (itemsize(t.arr_dtypes[k]) != 2 || !al4(ops.base[k])) tp.pair = false;
    for (int d = 0; d < ops.sndim && tp.pair; ++d)
      if (d != tp.axis && ops.sstrides[k][d] % 4 != 0) tp.pair = false;
  }
  return tp;
}

// Kernel of the tiled form; same parameters as the strided form.  Blocks of 256 threads walk the
// tiles (L-tile fastest, then A-tile, then the other axes in row-major order).  Staging: thread
// tid reads A position tid % TA of rows tid / TA, + 256 / TA, ...; evaluation: it takes L
// position tid % TL of A rows tid / TL, + 256 / TL, ...  The LDS rows hold TA + 1 values, so a
// column read by consecutive lanes falls on distinct banks.
static std::string gen_transpose(const Tape& t, const Operands& ops, const TransposePlan& tp, int32_t out_dtype,
                                 const std::string& kname) {
  const std::string C = ctype_of(t.compute_dtype);
  const std::string D = ctype_of(out_dtype);
  std::vector<std::string> roots;
  std::string body = emit_nodes(t, {t.root}, ops.alias, load_names(t), scalar_names(t), "lin", "t", roots);
  std::ostringstream os;
  os << strided_prelude(ops.sndim);
  os << "// tape: " << t.signature << " (tiled transpose, axis " << tp.axis << ")\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void " << kname << "("
     << param_list(t, ops, true, D, true) << ") {\n";
  os << "  constexpr int TA = " << tp.ta << ", TL = " << tp.tl << ", A = " << tp.axis << ", L = NCA_ND - 1;\n";
  for (int k : ops.unique)
    if (!ops.bcast[k] && tp.lds[k]) os << "  __shared__ " << ctype_of(t.arr_dtypes[k]) << " sk" << k << "[TL][TA + 1];\n";
  os << bcast_loads(t, ops);
  os << "  const nca_i64 nA = shp.v[A], nL = shp.v[L];\n"
     << "  const nca_i64 tA = (nA + TA - 1) / TA, tL = (nL + TL - 1) / TL;\n"
     << "  const nca_i64 ntiles = (n / (nA * nL)) * tA * tL;\n"
     << (tp.pair ? "  const int sx = (threadIdx.x % (TA / 2)) * 2, sy = threadIdx.x / (TA / 2), ex = (threadIdx.x % (TL / 2)) * 2, "
                   "ey = threadIdx.x / (TL / 2);\n"
                 : "  const int sx = threadIdx.x % TA, sy = threadIdx.x / TA, ex = threadIdx.x % TL, ey = threadIdx.x / TL;\n");
  // a tile's origin and operand offsets (locate: na0, nl0, no<k>, noo) and its staged operands'
  // loads (fetch)
  for (int k : ops.unique)
    if (!ops.bcast[k]) os << "  nca_i64 no" << k << " = 0;\n";
  os << "  nca_i64 na0 = 0, nl0 = 0, noo = 0;\n"
     << "  auto locate = [&](nca_i64 tile) {\n"
     << "    nca_i64 rem = tile;\n"
     << "    nl0 = (rem % tL) * TL;\n    rem /= tL;\n"
     << "    na0 = (rem % tA) * TA;\n    rem /= tA;\n";
  for (int k : ops.unique)
    if (!ops.bcast[k]) os << "    no" << k << " = 0;\n";
  os << "    noo = 0;\n";
  os << "#pragma unroll\n    for (int d = NCA_ND - 2; d >= 0; --d) {\n      if (d == A) continue;\n"
     << "      const nca_i64 c = rem % shp.v[d];\n      rem /= shp.v[d];\n";
  for (int k : ops.unique)
    if (!ops.bcast[k]) os << "      no" << k << " += c * st" << k << ".v[d];\n";
  os << "      noo += c * sto.v[d];\n    }\n  };\n";
  // stage the operands that are unit-stride along A (sx runs along A).  Every thread issues all
  // its loads of a tile before any LDS store (a fixed count, unrolled), and lanes past the array's
  // edge read the nearest element inside it (never used), so the loads are in flight together and
  // no branch separates a load from its use.
  const int step1 = tp.pair ? 512 / tp.ta : 256 / tp.ta, R1 = tp.tl / step1;
  const int step2 = tp.pair ? 512 / tp.tl : 256 / tp.tl, R2 = tp.ta / step2;
  for (int k : ops.unique) {
    if (ops.bcast[k] || !tp.lds[k]) continue;
    const std::string Tk = ctype_of(t.arr_dtypes[k]);
    const std::string VT = tp.pair ? "nca_vec<" + Tk + ", 2>" : Tk;
    os << "  " << VT << " g" << k << "[" << R1 << "]" << (tp.pair ? " = {}" : "") << ";\n";
  }
  auto fetch = [&](const std::string& ind) {
    os << ind << "{\n" << ind << "  const nca_i64 sa = na0 + sx < nA ? na0 + sx : nA - 1;\n";
    for (int k : ops.unique) {
      if (ops.bcast[k] || !tp.lds[k]) continue;
      const std::string Tk = ctype_of(t.arr_dtypes[k]), K = std::to_string(k);
      const std::string VT = tp.pair ? "nca_vec<" + Tk + ", 2>" : Tk;
      os << "#pragma unroll\n" << ind << "  for (int r = 0; r < " << R1 << "; ++r) {\n"
         << ind << "    const nca_i64 l = (nl0 + sy + r * " << step1 << " < nL) ? nl0 + sy + r * " << step1 << " : nL - 1;\n"
         << ind << "    const char* ad = p" << K << " + no" << K << " + sa * st" << K << ".v[A] + l * st" << K << ".v[L];\n";
      if (tp.pair)
        os << ind << "    if (sa + 1 < nA) g" << K << "[r] = *reinterpret_cast<const " << VT << "*>(ad);\n"
           << ind << "    else g" << K << "[r].v[0] = *reinterpret_cast<const " << Tk << "*>(ad);\n";
      else
        os << ind << "    g" << K << "[r] = *reinterpret_cast<const " << Tk << "*>(ad);\n";
      os << ind << "  }\n";
    }
    os << ind << "}\n";
  };
  // compute and store one tile (ex runs along L).  value(dl) evaluates the tape at L offset dl;
  // operands read directly (not staged) are read at the clamped position (ac, lc*) inside the
  // array, and lanes past the edge store into a sink, so every lane issues the same loads and
  // stores.
  auto value = [&](const std::string& dl, const std::string& lcl, const std::string& dst) {
    os << "        {\n";
    for (int k : ops.unique) {
      if (ops.bcast[k]) continue;
      const std::string Tk = ctype_of(t.arr_dtypes[k]);
      if (tp.lds[k])
        os << "        const " << C << " u" << k << " = nca_cvt<" << C << ", " << Tk << ">::from(sk" << k << "[ex" << dl
           << "][j]);\n";
      else
        os << "        const " << C << " u" << k << " = nca_ld<" << C << ">(reinterpret_cast<const " << Tk << "*>(p" << k
           << " + o" << k << " + ac * st" << k << ".v[A] + " << lcl << " * st" << k << ".v[L]));\n";
    }
    os << body;
    os << "        " << dst << " = nca_store_cvt<" << D << ", " << C << ">::to(" << roots[0] << ");\n        }\n";
  };
  auto phase2 = [&]() {
    os << "#pragma unroll\n    for (int r = 0; r < " << R2 << "; ++r) {\n"
       << "      const int j = ey + r * " << step2 << ";\n"
       << "      const nca_i64 a = a0 + j, l = l0 + ex;\n"
       << "      const bool ok = a < nA && l < nL;\n"
       << "      const nca_i64 ac = a < nA ? a : nA - 1, lc = l < nL ? l : nL - 1;\n"
       << "      char* q = ok ? outp + oo + a * sto.v[A] + l * sto.v[L] : reinterpret_cast<char*>(nca_sink_words);\n";
    if (tp.pair) {
      os << "      const nca_i64 lc1 = l + 1 < nL ? l + 1 : nL - 1;\n"
         << "      " << D << " r0, r1;\n";
      value("", "lc", "r0");
      value(" + 1", "lc1", "r1");
      os << "      if (!ok || l + 1 < nL) {\n"
         << "        nca_vec<" << D << ", 2> w;\n        w.v[0] = r0;\n        w.v[1] = r1;\n"
         << "        *reinterpret_cast<nca_vec<" << D << ", 2>*>(q) = w;\n"
         << "      } else {\n        *reinterpret_cast<" << D << "*>(q) = r0;\n      }\n";
    } else {
      os << "      " << D << " r0;\n";
      value("", "lc", "r0");
      os << "      *reinterpret_cast<" << D << "*>(q) = r0;\n";
    }
    os << "    }\n";
  };
  // one tile: the registers the fetch filled go to LDS, then the tile is computed and stored
  auto tile_body = [&]() {
    os << "    {\n    const nca_i64 a0 = na0, l0 = nl0, oo = noo;\n";
    for (int k : ops.unique)
      if (!ops.bcast[k]) os << "    const nca_i64 o" << k << " = no" << k << ";\n";
    for (int k : ops.unique) {
      if (ops.bcast[k] || !tp.lds[k]) continue;
      const std::string K = std::to_string(k);
      os << "#pragma unroll\n    for (int r = 0; r < " << R1 << "; ++r) {\n";
      if (tp.pair)
        os << "      sk" << K << "[sy + r * " << step1 << "][sx] = g" << K << "[r].v[0];\n"
           << "      sk" << K << "[sy + r * " << step1 << "][sx + 1] = g" << K << "[r].v[1];\n";
      else
        os << "      sk" << K << "[sy + r * " << step1 << "][sx] = g" << K << "[r];\n";
      os << "    }\n";
    }
    os << "    __syncthreads();\n";
    phase2();
    os << "    __syncthreads();\n    }\n";
  };
  os << "  for (nca_i64 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {\n    locate(tile);\n";
  fetch("    ");
  tile_body();
  os << "  }\n}\n";
  return os.str();
}

// ---- row form ------------------------------------------------------------------------------------
// The strided form unravels every element index over all axes (a 64-bit divide per axis) and, for
// indirect views, follows the pointer chain per element.  When the last coalesced axis L holds a
// wave's worth of elements and every pointer step lies before L, one wave takes one row: the row
// index is unravelled and each chain followed once per row, and the lanes step along L, two
// elements per lane per step.
static bool row_form_ok(const Operands& ops, int fuse) {
  if (ops.contiguous || fuse || ops.sndim < 2) return false;
  const int L = ops.sndim - 1;
  if (ops.sshape[L] < 64) return false;
  if (ops.nc) {
    for (int k : ops.unique)
      if (!ops.bcast[k] && !ops.ind[k].empty() && ops.ind[k].back() >= L) return false;
    if (!ops.indo.empty() && ops.indo.back() >= L) return false;
  }
  return true;
}

static std::string gen_rows(const Tape& t, const Operands& ops, int32_t out_dtype, const std::string& kname) {
  const std::string C = ctype_of(t.compute_dtype);
  const std::string D = ctype_of(out_dtype);
  std::vector<std::string> roots;
  std::string body = emit_nodes(t, {t.root}, ops.alias, load_names(t), scalar_names(t), "lin", "t", roots);
  std::ostringstream os;
  os << strided_prelude(ops.sndim);
  os << "// tape: " << t.signature << " (rows)\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void " << kname << "("
     << param_list(t, ops, true, D, true) << ") {\n";
  os << bcast_loads(t, ops);
  os << "  constexpr int L = NCA_ND - 1, U = 2;\n"
     << "  const nca_i64 nL = shp.v[L], nrows = n / nL;\n"
     << "  const int lane = threadIdx.x & 63;\n"
     << "  const nca_i64 nwaves = ((nca_i64)gridDim.x * blockDim.x) >> 6;\n"
     << "  for (nca_i64 row = ((nca_i64)blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < nrows; row += nwaves) {\n";
  os << strided_offsets(t, ops, true, "row", true);
  for (int k : ops.unique)
    if (!ops.bcast[k])
      os << "      const char* rb" << k << " = " << operand_address(ops, k) << ";\n";
  if (ops.nc && !ops.indo.empty())
    os << "      char* rbo = (char*)" << chain_address("outp", "oo", "qo_", ops.indo.size()) << ";\n";
  else
    os << "      char* rbo = outp + oo;\n";
  os << "      for (nca_i64 l0 = lane; l0 < nL; l0 += 64 * U) {\n";
  for (int k : ops.unique)
    if (!ops.bcast[k]) os << "        " << ctype_of(t.arr_dtypes[k]) << " x" << k << "[U];\n";
  os << "#pragma unroll\n        for (int q = 0; q < U; ++q) {\n          const nca_i64 l = l0 + 64 * q;\n"
     << "          if (l < nL) {\n";
  for (int k : ops.unique)
    if (!ops.bcast[k])
      // streaming loads: each element is read once (measured: BF16 plane gathers of 512^3
      // 0.142 -> 0.137 ms; the tiled transpose form measured up to 2% slower with them and keeps
      // cached loads)
      os << "            x" << k << "[q] = nca_ldnt(reinterpret_cast<const " << ctype_of(t.arr_dtypes[k]) << "*>(rb" << k
         << " + l * st" << k << ".v[L]));\n";
  os << "          }\n        }\n";
  os << "#pragma unroll\n        for (int q = 0; q < U; ++q) {\n          const nca_i64 l = l0 + 64 * q;\n"
     << "          if (l < nL) {\n";
  for (int k : ops.unique)
    if (!ops.bcast[k])
      os << "            const " << C << " u" << k << " = nca_cvt<" << C << ", " << ctype_of(t.arr_dtypes[k])
         << ">::from(x" << k << "[q]);\n";
  if (t.uses_idx) os << "            const nca_i64 lin = row * nL + l;\n";
  os << body;
  os << "            *reinterpret_cast<" << D << "*>(rbo + l * sto.v[L]) = nca_store_cvt<" << D << ", " << C << ">::to("
     << roots[0] << ");\n";
  os << "          }\n        }\n      }\n  }\n}\n";
  return os.str();
}

void expr_launch_impl(uint64_t handle, const nca_view* ins, int n_in, const nca_scalar* scalars,
                      int n_scalars, const nca_view* outp, int fuse, int mode, void* sum_out,
                      void* max_out, void* min_out, hipStream_t s) {
  NCA_REQUIRE(outp != nullptr, "null out view");
  ExprProgram& p = expr_program(handle);
  const Tape& t = p.tape;
  const nca_view& out = *outp;
  NCA_REQUIRE(out.dtype == t.dest_dtype, std::string("out dtype ") + dtype_info(out.dtype).name +
                                             " does not match the tape's destination dtype " +
                                             dtype_info(t.dest_dtype).name);
  Operands ops = analyse_operands(t, ins, n_in, out, fuse != 0);
  ScalarSlots sc = make_scalar_slots(t, scalars, n_scalars);
  DeviceState& dev = current_device();
  const bool strided = !ops.contiguous;
  const bool with_out = true;
  const TransposePlan tp = transpose_plan(t, ops, out, fuse);
  if (tp.axis >= 0 && ops.n > 0) {
    std::string key = "t" + std::to_string(ops.sndim) + "a" + std::to_string(tp.axis) + "x" + std::to_string(tp.ta) + "x" +
                      std::to_string(tp.tl) + (tp.pair ? "p" : "") + "|" + sig_operands(t, ops) + "|";
    for (int k : ops.unique) key += tp.lds[k] ? '1' : '0';
    const std::string kname = "nca_expr_transpose";
    hipFunction_t fn = program_kernel(p, key, [&]() { return gen_transpose(t, ops, tp, out.dtype, kname); }, kname);
    LaunchArgs la;
    push_common(la, t, ops, ins, out, true, with_out, sc);
    const int64_t nA = ops.sshape[tp.axis], nL = ops.sshape[ops.sndim - 1];
    const int64_t tiles = (ops.n / (nA * nL)) * ((nA + tp.ta - 1) / tp.ta) * ((nL + tp.tl - 1) / tp.tl);
    const int blocks = static_cast<int>(std::min<int64_t>(tiles, int64_t(dev.n_cus) * 8));
    ProfScope ps("nca_expr_transpose", s);
    jit_launch(fn, dim3(blocks), dim3(256), 0, s, la.params);
    return;
  }
  if (row_form_ok(ops, fuse) && ops.n > 0) {
    const std::string key = "r" + std::to_string(ops.sndim) + "|" + sig_operands(t, ops);
    const std::string kname = "nca_expr_rows";
    hipFunction_t fn = program_kernel(p, key, [&]() { return gen_rows(t, ops, out.dtype, kname); }, kname);
    LaunchArgs la;
    push_common(la, t, ops, ins, out, true, with_out, sc);
    const int64_t rows = ops.n / ops.sshape[ops.sndim - 1];
    const int blocks = grid_blocks(rows * 64, 256, dev.n_cus * 16);
    ProfScope ps("nca_expr_rows", s);
    jit_launch(fn, dim3(blocks), dim3(256), 0, s, la.params);
    return;
  }
  std::string key = std::string(strided ? "s" : "c") + std::to_string(strided ? ops.sndim : ops.vec) + "|" +
                    sig_operands(t, ops) + "|f" + std::to_string(fuse) + "m" + std::to_string(mode);
  const std::string kname = fuse ? "nca_expr_reduce" : (strided ? "nca_expr_strided" : "nca_expr_contig");
  hipFunction_t fn = program_kernel(
      p, key,
      [&]() { return gen_elementwise(t, ops, strided, ops.vec, out.dtype, with_out, fuse, mode, kname); },
      kname);
  LaunchArgs la;
  push_common(la, t, ops, ins, out, strided, with_out, sc);
  const int64_t items = strided ? ops.n : (ops.n / ops.vec > 0 ? ops.n / ops.vec : 1);
  // The contiguous form issues one 16-byte vector per thread per step: 4 blocks per CU stream a
  // 512 MiB float32 copy in 185 us against 232 us with 16.  The strided form (a 64-bit unravel
  // per element) needs the 16.  NCA_HIP_CONTIG_BLOCKS_PER_CU overrides the contiguous cap.
  static const int contig_per_cu = [] {
    const char* e = std::getenv("NCA_HIP_CONTIG_BLOCKS_PER_CU");
    const int v = e ? std::atoi(e) : 4;
    return v >= 1 && v <= 64 ? v : 4;
  }();
  bool reads_array = false;
  for (int k : ops.unique)
    if (!ops.bcast[k]) reads_array = true;
  // the store-only form (no array read): two blocks per CU, as for the fastest measured writes
  const int max_blocks =
      fuse ? dev.n_cus * 4 : dev.n_cus * (strided ? 16 : (!reads_array && with_out ? 2 : contig_per_cu));
  const int blocks = grid_blocks(items, 256, max_blocks);
  if (fuse) {
    Workspace& w = workspace(s, static_cast<size_t>(blocks) * 8 * 3);
    void* psum = w.partials;
    void* pmax = static_cast<char*>(w.partials) + static_cast<size_t>(blocks) * 8;
    void* pmin = static_cast<char*>(w.partials) + static_cast<size_t>(blocks) * 16;
    la.params.push_back(&psum);
    la.params.push_back(&pmax);
    la.params.push_back(&pmin);
    {
      ProfScope ps("nca_expr_reduce", s);
      jit_launch(fn, dim3(blocks), dim3(256), 0, s, la.params);
    }
    FinalizeArgs fa;
    fa.value_dtype = out.dtype;
    fa.fuse = fuse;
    fa.psum = psum;
    fa.pmax = pmax;
    fa.pmin = pmin;
    fa.nblocks = blocks;
    fa.sum_out = sum_out;
    fa.max_out = max_out;
    fa.min_out = min_out;
    launch_finalize_fused(fa, s);
  } else {
    if (ops.n == 0) return;
    ProfScope ps(strided ? "nca_expr_strided" : "nca_expr_contig", s);
    jit_launch(fn, dim3(blocks), dim3(256), 0, s, la.params);
  }
}

// ---- VJP ---------------------------------------------------------------------------------------------
std::string gen_vjp(const Tape& base, const Operands& ops, int vec, const std::vector<int>& wrt_vars,
                    const std::vector<int32_t>& grad_dtypes, bool g_is_view, int32_t g_dtype,
                    const std::string& kname) {
  Tape t = base;  // derivative nodes are appended to a copy
  std::vector<int> droots;
  for (int var : wrt_vars) droots.push_back(differentiate(t, t.root, var, ops.alias));
  const std::string C = ctype_of(t.compute_dtype);
  std::vector<std::string> names;
  std::string body = emit_nodes(t, droots, ops.alias, load_names(t), scalar_names(t), "lin", "t", names);
  std::ostringstream os;
  os << "// vjp of tape: " << t.signature << "\n";
  os << "extern \"C\" __global__ __launch_bounds__(256) void " << kname << "(";
  bool first = true;
  auto sep = [&]() { if (!first) os << ", "; first = false; };
  for (int k : ops.unique) { sep(); os << "const " << ctype_of(t.arr_dtypes[k]) << "* __restrict__ p" << k; }
  sep();
  os << "const " << ctype_of(g_dtype) << "* __restrict__ gin";
  for (size_t m = 0; m < wrt_vars.size(); ++m) os << ", " << ctype_of(grad_dtypes[m]) << "* __restrict__ g" << m;
  os << ", nca_i64 n";
  for (size_t i = 0; i < t.scalar_dtypes.size(); ++i) os << ", " << C << " s" << i;
  os << ") {\n";
  os << bcast_loads(t, ops);
  if (!g_is_view) os << "  const " << C << " gsc = nca_ld<" << C << ">(gin);\n";
  os << "  const nca_i64 gstride = (nca_i64)gridDim.x * blockDim.x;\n";
  os << "  constexpr int V = " << vec << ";\n  const nca_i64 nv = n / V;\n";
  os << "  for (nca_i64 i = (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gstride) {\n";
  for (int k : ops.unique)
    if (!ops.bcast[k])
      os << "    const nca_vec<" << ctype_of(t.arr_dtypes[k]) << ", V> x" << k << " = nca_load_nt(reinterpret_cast<const nca_vec<"
         << ctype_of(t.arr_dtypes[k]) << ", V>*>(p" << k << " + i * V));\n";
  if (g_is_view)
    os << "    const nca_vec<" << ctype_of(g_dtype) << ", V> xg = *reinterpret_cast<const nca_vec<" << ctype_of(g_dtype)
       << ", V>*>(gin + i * V);\n";
  for (size_t m = 0; m < wrt_vars.size(); ++m) os << "    nca_vec<" << ctype_of(grad_dtypes[m]) << ", V> r" << m << ";\n";
  os << "#pragma unroll\n    for (int j = 0; j < V; ++j) {\n";
  for (int k : ops.unique)
    if (!ops.bcast[k])
      os << "      const " << C << " u" << k << " = nca_cvt<" << C << ", " << ctype_of(t.arr_dtypes[k]) << ">::from(x" << k << ".v[j]);\n";
  if (t.uses_idx) os << "      const nca_i64 lin = i * V + j;\n";
  os << "      const " << C << " g = " << (g_is_view ? "nca_cvt<" + C + ", " + ctype_of(g_dtype) + ">::from(xg.v[j])" : "gsc") << ";\n";
  os << body;
  for (size_t m = 0; m < wrt_vars.size(); ++m)
    os << "      r" << m << ".v[j] = nca_store_cvt<" << ctype_of(grad_dtypes[m]) << ", " << C << ">::to((" << C << ")(g * " << names[m] << "));\n";
  os << "    }\n";
  for (size_t m = 0; m < wrt_vars.size(); ++m)
    os << "    *reinterpret_cast<nca_vec<" << ctype_of(grad_dtypes[m]) << ", V>*>(g" << m << " + i * V) = r" << m << ";\n";
  os << "  }\n";
  os << "  for (nca_i64 i = nv * V + (nca_i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {\n";
  for (int k : ops.unique)
    if (!ops.bcast[k]) os << "      const " << C << " u" << k << " = nca_ld<" << C << ">(p" << k << " + i);\n";
  if (t.uses_idx) os << "      const nca_i64 lin = i;\n";
  os << "      const " << C << " g = " << (g_is_view ? "nca_ld<" + C + ">(gin + i)" : "gsc") << ";\n";
  os << body;
  for (size_t m = 0; m < wrt_vars.size(); ++m)
    os << "      nca_st(g" << m << " + i, (" << C << ")(g * " << names[m] << "));\n";
  os << "  }\n}\n";
  return os.str();
}

}  // namespace nca

using namespace nca;

extern "C" {

int nca_hip_expr_compile(const nca_expr_desc* desc, uint64_t* handle) {
  return api_call([&] {
    NCA_REQUIRE(desc != nullptr && handle != nullptr, "null argument");
    Tape t = Tape::from_desc(*desc);
    *handle = expr_register(std::move(t));
  });
}

int nca_hip_expr_launch(uint64_t handle, const nca_view* ins, int32_t n_in, const nca_scalar* scalars,
                        int32_t n_scalars, const nca_view* out, void* stream) {
  return api_call([&] {
    expr_launch_impl(handle, ins, n_in, scalars, n_scalars, out, 0, NCA_ACCUM_PLAIN, nullptr, nullptr, nullptr,
                     as_stream(stream));
    check_launch("expression");
    finish_call(stream);
  });
}

int nca_hip_expr_launch_reduce(uint64_t handle, const nca_view* ins, int32_t n_in, const nca_scalar* scalars,
                               int32_t n_scalars, const nca_view* out, int32_t fuse, int32_t accum_mode,
                               void* sum_out, void* max_out, void* min_out, void* stream) {
  return api_call([&] {
    NCA_REQUIRE((fuse & ~(NCA_FUSE_SUM | NCA_FUSE_MAX | NCA_FUSE_MIN)) == 0 && fuse != 0, "bad fuse flags");
    NCA_REQUIRE(!(fuse & NCA_FUSE_SUM) || sum_out, "sum requested without sum_out");
    NCA_REQUIRE(!(fuse & NCA_FUSE_MAX) || max_out, "max requested without max_out");
    NCA_REQUIRE(!(fuse & NCA_FUSE_MIN) || min_out, "min requested without min_out");
    expr_launch_impl(handle, ins, n_in, scalars, n_scalars, out, fuse, accum_mode, sum_out, max_out, min_out,
                     as_stream(stream));
    check_launch("fused expression");
    finish_call(stream);
  });
}

int nca_hip_expr_vjp(uint64_t handle, const nca_view* ins, int32_t n_in, const nca_scalar* scalars,
                     int32_t n_scalars, const nca_view* grad_out, const void* grad_scalar, const int32_t* wrt,
                     const nca_view* grads, void* stream) {
  return api_call([&] {
    ExprProgram& p = expr_program(handle);
    const Tape& t = p.tape;
    NCA_REQUIRE(wrt != nullptr && grads != nullptr, "null wrt/grads");
    NCA_REQUIRE(is_float(t.compute_dtype), "derivatives need a floating-point evaluation dtype");
    NCA_REQUIRE(grad_out != nullptr || grad_scalar != nullptr, "need grad_out or grad_scalar");
    // the output shape comes from the gradient view, or from the first non-broadcast operand
    nca_view shape_view{};
    if (grad_out) {
      shape_view = *grad_out;
    } else {
      bool found = false;
      for (int k = 0; k < n_in && !found; ++k)
        if (wrt[k]) { shape_view = grads[k]; found = true; }
      NCA_REQUIRE(found, "no gradient requested");
    }
    nca_view sv = shape_view;
    sv.dtype = t.dest_dtype;
    Operands ops = analyse_operands(t, ins, n_in, sv, true);
    NCA_REQUIRE(ops.contiguous, "vjp needs C-contiguous operands");
    if (grad_out) {
      validate_view(*grad_out, "grad_out");
      NCA_REQUIRE(is_c_contiguous(*grad_out), "grad_out must be C-contiguous");
    }
    std::vector<int> vars;
    std::vector<int32_t> gdt;
    std::vector<const void*> gptr;
    int vec = ops.vec;
    for (int k = 0; k < n_in; ++k) {
      if (!wrt[k] || ops.alias[k] != k) continue;
      NCA_REQUIRE(!ops.bcast[k], "cannot differentiate with respect to a broadcast operand");
      const nca_view& g = grads[k];
      validate_view(g, "grad");
      NCA_REQUIRE(same_shape(g, sv) && is_c_contiguous(g), "grad view must match the operand shape, C-contiguous");
      NCA_REQUIRE(is_float(g.dtype), "gradients must be floating point");
      vars.push_back(k);
      gdt.push_back(g.dtype);
      gptr.push_back(g.data);
      if (reinterpret_cast<uintptr_t>(g.data) % (vec * itemsize(g.dtype))) vec = 1;
    }
    NCA_REQUIRE(!vars.empty(), "no gradient requested");
    const int32_t g_dtype = grad_out ? grad_out->dtype : t.dest_dtype;
    if (grad_out && reinterpret_cast<uintptr_t>(grad_out->data) % (vec * itemsize(g_dtype))) vec = 1;
    ScalarSlots sc = make_scalar_slots(t, scalars, n_scalars);
    std::ostringstream key;
    key << "vjp" << vec << "|" << sig_operands(t, ops) << "|g" << (grad_out ? 1 : 0) << g_dtype << "|";
    for (size_t m = 0; m < vars.size(); ++m) key << vars[m] << ":" << gdt[m] << ",";
    hipFunction_t fn = program_kernel(
        p, key.str(), [&]() { return gen_vjp(t, ops, vec, vars, gdt, grad_out != nullptr, g_dtype, "nca_expr_vjp"); },
        "nca_expr_vjp");
    std::deque<const void*> ptrs;
    std::vector<void*> params;
    for (int k : ops.unique) { ptrs.push_back(ins[k].data); params.push_back(&ptrs.back()); }
    ptrs.push_back(grad_out ? grad_out->data : grad_scalar);
    params.push_back(&ptrs.back());
    for (auto gp : gptr) { ptrs.push_back(gp); params.push_back(&ptrs.back()); }
    int64_t n = ops.n;
    params.push_back(&n);
    for (auto& s : sc.slots) params.push_back(&s);
    DeviceState& dev = current_device();
    int blocks = grid_blocks(n / vec > 0 ? n / vec : 1, 256, dev.n_cus * 16);
    if (n > 0) {
      ProfScope ps("nca_expr_vjp", as_stream(stream));
      jit_launch(fn, dim3(blocks), dim3(256), 0, as_stream(stream), params);
    }
    check_launch("vjp");
    finish_call(stream);
  });
}

int nca_hip_expr_source(uint64_t handle, int32_t variant, int32_t ndim, char* buf, size_t n) {
  std::string src;
  int rc = api_call([&] {
    ExprProgram& p = expr_program(handle);
    const Tape& t = p.tape;
    // a representative operand set: every operand distinct, contiguous
    Operands ops;
    ops.alias.resize(t.arr_dtypes.size());
    ops.bcast.assign(t.arr_dtypes.size(), 0);
    for (size_t k = 0; k < t.arr_dtypes.size(); ++k) { ops.alias[k] = static_cast<int>(k); ops.unique.push_back(static_cast<int>(k)); }
    int max_item = itemsize(t.dest_dtype);
    for (int32_t dt : t.arr_dtypes) max_item = std::max(max_item, itemsize(dt));
    ops.vec = std::max(1, 16 / max_item);
    ops.sndim = ndim > 0 ? ndim : 1;
    switch (variant) {
      case 0: src = gen_elementwise(t, ops, false, ops.vec, t.dest_dtype, true, 0, 0, "nca_expr_contig"); break;
      case 1: src = gen_elementwise(t, ops, true, 1, t.dest_dtype, true, 0, 0, "nca_expr_strided"); break;
      case 2: src = gen_elementwise(t, ops, false, ops.vec, t.dest_dtype, true, NCA_FUSE_SUM | NCA_FUSE_MAX, NCA_ACCUM_COMPENSATED, "nca_expr_reduce"); break;
      case 3: {
        NCA_REQUIRE(is_float(t.compute_dtype), "derivatives need a floating-point evaluation dtype");
        std::vector<int> vars;
        std::vector<int32_t> gdt;
        for (size_t k = 0; k < t.arr_dtypes.size(); ++k) { vars.push_back(static_cast<int>(k)); gdt.push_back(NCA_FLOAT32); }
        src = gen_vjp(t, ops, ops.vec, vars, gdt, false, t.dest_dtype, "nca_expr_vjp");
        break;
      }
      case 4: src = gen_batch_source(t, NCA_FUSE_SUM | NCA_FUSE_MAX, NCA_ACCUM_PLAIN, 4, true); break;
      case 5: {
        NCA_REQUIRE(ops.sndim >= 2, "the tiled transpose form needs ndim >= 2");
        TransposePlan tp;
        tp.axis = 0;
        tp.lds.assign(t.arr_dtypes.size(), 0);
        int64_t lds_bytes = 0;
        for (size_t k = 0; k < t.arr_dtypes.size(); ++k) {
          const int64_t b = itemsize(t.arr_dtypes[k]) * tp.tl * (tp.ta + 1);
          if (lds_bytes + b <= (int64_t(64) << 10)) { tp.lds[k] = 1; lds_bytes += b; }
        }
        tp.pair = itemsize(t.dest_dtype) == 2;
        for (size_t k = 0; k < t.arr_dtypes.size(); ++k)
          if (tp.lds[k] && itemsize(t.arr_dtypes[k]) != 2) tp.pair = false;
        src = gen_transpose(t, ops, tp, t.dest_dtype, "nca_expr_transpose");
        break;
      }
      case 6:
        NCA_REQUIRE(ops.sndim >= 2, "the row form needs ndim >= 2");
        src = gen_rows(t, ops, t.dest_dtype, "nca_expr_rows");
        break;
      default: NCA_THROW(NCA_ERR_INVALID, "unknown variant");
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
