// Checker for the oracle (test infrastructure only): a C interface to the reference's own
// header-only host routines, compiled from the reference sources where they lie
// (src/lib/ncarray, see oracle/Makefile target `ref`) into oracle/_ref/libnca_ref.so.  Nothing
// in the product links or loads this; only tests/test_oracle_ref.py does, to pin oracle/nca_cpu.c
// against the reference's code itself:
//   * full reductions: host::{sum,max,min,argmax,argmin}_recursive
//     (host/reductions.hh:35-212) over strided NC views, NC views with a pointer axis
//     (NCOffsetsPolicy::advance, layout.hh:451-480) and SO views (SOArrayPolicy::advance,
//     layout.hh:677-700);
//   * host::scatter_add_recursive (host/elementwise.hh:482-505);
//   * the expression VM: DynamicExprMVNode::eval (expression/dynamicmvnode.hh:146-251) built as
//     HostEngine::execute_binary_expression builds it for tapes it cannot linearize
//     (engines/hostengine.hh:74-81: bytes_for_dynamic_vm + get_dynamic_mv_node).
// The host engine itself (engines/hostengine.hh) is not included: it pulls in the asmjit JIT,
// which is not available here.
#include "ncarray/array_impl.hh"
#include "ncarray/expression/dynamicmvnode.hh"
#include "ncarray/host/elementwise.hh"
#include "ncarray/host/reductions.hh"

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <variant>
#include <vector>

namespace {

using ncarray::Scalar;

// Scalar (a std::variant) -> float64 value and, for integer results, the exact int64 value.
void scalar_out(const Scalar& s, double* f, int64_t* i) {
  std::visit(
      [&](auto&& v) {
        using V = std::decay_t<decltype(v)>;
        if constexpr (std::is_arithmetic_v<V>) {
          *f = static_cast<double>(v);
          *i = std::is_integral_v<V> ? static_cast<int64_t>(v) : 0;
        } else {
          *f = 0.0;
          *i = 0;
        }
      },
      s);
}

ncarray::Metadata meta(const int64_t* v, int n) {
  ncarray::Metadata m;
  std::vector<ssize_t> tmp(v, v + n);
  m.set(tmp.data(), n);
  return m;
}

template <typename T, typename A>
int reduce_typed(int op, const A& a, double* f, int64_t* i) {
  namespace h = ncarray::host;
  switch (op) {
    case 0: scalar_out(h::sum_recursive<T>(a), f, i); return 0;
    case 1: scalar_out(h::max_recursive<T>(a), f, i); return 0;
    case 2: scalar_out(h::min_recursive<T>(a), f, i); return 0;
    case 3: scalar_out(h::argmax_recursive<T>(a), f, i); return 0;
    case 4: scalar_out(h::argmin_recursive<T>(a), f, i); return 0;
    default: return -1;
  }
}

// C-order walk of a view by the layout's own advance (the traversal of
// host/reductions.hh:36-54), writing each element's byte offset from `base`.
template <typename A>
void offsets_walk(const A& a, const void* p, ssize_t axis, const std::uint8_t* base, int64_t*& out) {
  for (ssize_t i = 0; i < a.shape()[axis]; ++i) {
    const void* q = a.advance(p, axis, i);
    if (axis == static_cast<ssize_t>(a.ndim()) - 1) {
      *out++ = reinterpret_cast<const std::uint8_t*>(q) - base;
    } else {
      offsets_walk(a, q, axis + 1, base, out);
    }
  }
}

template <typename A>
int reduce_dispatch(int op, int dtype, const A& a, double* f, int64_t* i) {
  using D = ncarray::DType;
  switch (static_cast<D>(dtype)) {
    case D::bool_: return reduce_typed<bool>(op, a, f, i);
    case D::uint8: return reduce_typed<std::uint8_t>(op, a, f, i);
    case D::int16: return reduce_typed<std::int16_t>(op, a, f, i);
    case D::int32: return reduce_typed<std::int32_t>(op, a, f, i);
    case D::int64: return reduce_typed<std::int64_t>(op, a, f, i);
    case D::uint32: return reduce_typed<std::uint32_t>(op, a, f, i);
    case D::float32: return reduce_typed<float>(op, a, f, i);
    case D::float64: return reduce_typed<double>(op, a, f, i);
    default: return -2;
  }
}

}  // namespace

extern "C" {

// Full reduction `op` (0 sum, 1 max, 2 min, 3 argmax, 4 argmin) of a host view.
// layout 0: NC (strides in bytes; `aux` = per-axis byte offsets; pointer_axis >= 0: data is a
// table of row pointers for that axis); layout 1: SO (`aux` = suboffsets, < 0 for a direct axis).
// The result goes to *f (float64) and, for integer results (integer sums, arg indices), *i.
int ref_reduce(int op, int dtype, int layout, void* data, int ndim, const int64_t* shape, const int64_t* strides,
               const int64_t* aux, int pointer_axis, double* f, int64_t* i) {
  const auto dt = static_cast<ncarray::DType>(dtype);
  if (layout == 0) {
    ncarray::ArrayImpl<ncarray::NCOffsetsPolicy, ncarray::ViewPolicy> a(data, meta(shape, ndim), meta(strides, ndim),
                                                                        meta(aux, ndim), dt, pointer_axis, true);
    return reduce_dispatch(op, dtype, a, f, i);
  }
  ncarray::ArrayImpl<ncarray::SOArrayPolicy, ncarray::ViewPolicy> a(data, meta(shape, ndim), meta(strides, ndim),
                                                                    meta(aux, ndim), dt, -1, true);
  return reduce_dispatch(op, dtype, a, f, i);
}

// The byte offset from `base` of every element of the view (arguments as ref_reduce), in C order.
int ref_offsets(int layout, void* data, int ndim, const int64_t* shape, const int64_t* strides, const int64_t* aux,
                int pointer_axis, const void* base, int64_t* out) {
  const auto* b = static_cast<const std::uint8_t*>(base);
  if (layout == 0) {
    ncarray::ArrayImpl<ncarray::NCOffsetsPolicy, ncarray::ViewPolicy> a(data, meta(shape, ndim), meta(strides, ndim),
                                                                        meta(aux, ndim), ncarray::DType::uint8,
                                                                        pointer_axis, true);
    if (a.size() > 0) offsets_walk(a, data, 0, b, out);
    return 0;
  }
  ncarray::ArrayImpl<ncarray::SOArrayPolicy, ncarray::ViewPolicy> a(data, meta(shape, ndim), meta(strides, ndim),
                                                                    meta(aux, ndim), ncarray::DType::uint8, -1, true);
  if (a.size() > 0) offsets_walk(a, data, 0, b, out);
  return 0;
}

// dest[idx[k]] += src[k] for the n source elements, skipping indices outside [0, n_dest).
int ref_scatter_add_f32(float* dest, int64_t n_dest, const int32_t* idx, const float* src, int64_t n) {
  using V = ncarray::ArrayImpl<ncarray::NCOffsetsPolicy, ncarray::ViewPolicy>;
  const ssize_t sd[1] = {n_dest}, sn[1] = {n}, s4[1] = {4};
  V d(dest, 1, sd, s4, ncarray::DType::float32, -1, false);
  V x(const_cast<int32_t*>(idx), 1, sn, s4, ncarray::DType::int32, -1, true);
  V s(const_cast<float*>(src), 1, sn, s4, ncarray::DType::float32, -1, true);
  ncarray::host::scatter_add_recursive<float, std::int32_t, float>(d, x, s);
  return 0;
}

// The reference's expression VM on n elements: float32 1-D contiguous operands `arrays`
// (n_arrays of them), float32 scalars, work and result dtype float32.  Operand layouts are
// stored as ExprMVNode::build_node stores them (expression/mvnode.hh:206-217): an SO view's
// policy as is (for LOAD_SOARR tapes, so_loads != 0), an NC view's policy reinterpreted
// (for LOAD_NCARR tapes).
int ref_vm_f32(const uint32_t* instrs, int n_instrs, const float* const* arrays, int n_arrays, const float* scalars,
               int n_scalars, int64_t n, int so_loads, float* out) {
  using SOView = ncarray::ArrayImpl<ncarray::SOArrayPolicy, ncarray::ViewPolicy>;
  using NCView = ncarray::ArrayImpl<ncarray::NCOffsetsPolicy, ncarray::ViewPolicy>;
  ncarray::ExprMVNode<ncarray::HostTag> node;
  const ssize_t sh[1] = {n}, st[1] = {4};
  for (int k = 0; k < n_arrays; ++k) {
    void* p = const_cast<float*>(arrays[k]);
    if (so_loads) {
      SOView v(p, 1, sh, st, ncarray::DType::float32, -1, true);
      node.layouts.push_back(static_cast<const ncarray::SOArrayPolicy&>(v));
    } else {
      NCView v(p, 1, sh, st, ncarray::DType::float32, -1, true);
      node.layouts.push_back(reinterpret_cast<const ncarray::SOArrayPolicy&>(v));
      node.soarray = false;
    }
    node.dtypes.push_back(ncarray::DType::float32);
    node.data.push_back(p);
  }
  for (int k = 0; k < n_instrs; ++k) node.instrs.push_back(static_cast<ncarray::Instruction>(instrs[k]));
  for (int k = 0; k < n_scalars; ++k) node.scalars.push_back(Scalar{scalars[k]});
  node.final_shape.set(sh, 1);
  node.final_size = n;
  node.expr_dtype = ncarray::DType::float32;
  node.work_dtype = ncarray::DType::float32;
  node.has_expr_dtype = true;
  const size_t bytes = ncarray::bytes_for_dynamic_vm(node);
  std::vector<std::uint8_t> raw(bytes + 64);
  auto* buf = reinterpret_cast<std::uint8_t*>((reinterpret_cast<uintptr_t>(raw.data()) + 63) & ~uintptr_t(63));
  auto vm = ncarray::get_dynamic_mv_node(node, buf);
  for (int64_t e = 0; e < n; ++e) {
    ncarray::StaticCoords<1, std::uint32_t> c;
    c[0] = static_cast<std::uint32_t>(e);
    out[e] = vm.template eval<float>(c);
  }
  return 0;
}

}  // extern "C"
