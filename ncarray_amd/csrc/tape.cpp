#include "tape.hh"

#include <cstdio>
#include <functional>
#include <sstream>

#include "nca_internal.hh"

namespace nca {

std::string ctype_of(int32_t dtype) { return dtype_info(dtype).ctype; }

// can_linearize (staticmvnode.hh:188-239): <= 16 arrays, <= 32 instructions, one array
// dtype, one layout kind, RPN depth never above 2, final depth 1.
static bool can_linearize(const Tape& t) {
  if (t.arr_dtypes.size() > 16 || t.instrs.size() > 32) return false;
  int depth = 0;
  bool seen_arr = false, so = false;
  int32_t ref = -1;
  for (uint32_t ins : t.instrs) {
    int32_t op = instr_op(ins);
    if (op == NCA_OP_LOAD_NCARR || op == NCA_OP_LOAD_SOARR) {
      ++depth;
      int32_t k = instr_index(ins);
      if (!seen_arr) {
        seen_arr = true;
        ref = t.arr_dtypes[k];
        so = (op == NCA_OP_LOAD_SOARR);
      } else if (op == NCA_OP_LOAD_SOARR && !so) {
        return false;
      }
      if (t.arr_dtypes[k] != ref) return false;
    } else if (op == NCA_OP_LOAD_CONST || op == NCA_OP_IDX) {
      ++depth;
    } else if (is_binary_op(op)) {
      if (depth < 2) return false;
      --depth;
    } else {
      if (depth < 1) return false;
    }
    if (depth > 2) return false;
  }
  return depth == 1;
}

static bool op_supported(int32_t op) {
  switch (op) {
    case NCA_OP_SZOF: case NCA_OP_ADDR: case NCA_OP_INDR: case NCA_OP_CAST: case NCA_OP_NOOP:
      return false;
    default:
      return is_unary_op(op) || is_binary_op(op) || is_load_op(op);
  }
}

Tape Tape::from_desc(const nca_expr_desc& d) {
  Tape t;
  NCA_REQUIRE(d.n_instrs > 0 && d.instrs != nullptr, "expression tape is empty");
  NCA_REQUIRE(d.n_arrays >= 0 && d.n_scalars >= 0, "negative operand counts");
  NCA_REQUIRE(d.n_arrays == 0 || d.arr_dtypes != nullptr, "missing array dtypes");
  NCA_REQUIRE(d.n_scalars == 0 || d.scalar_dtypes != nullptr, "missing scalar dtypes");
  t.instrs.assign(d.instrs, d.instrs + d.n_instrs);
  t.arr_dtypes.assign(d.arr_dtypes, d.arr_dtypes + d.n_arrays);
  t.scalar_dtypes.assign(d.scalar_dtypes, d.scalar_dtypes + d.n_scalars);
  t.work_dtype = d.work_dtype;
  t.dest_dtype = d.dest_dtype;
  for (int32_t dt : t.arr_dtypes)
    if (!is_supported(dt)) NCA_THROW(NCA_ERR_UNSUPPORTED, std::string("array dtype ") + dtype_info(dt).name + " unsupported");
  for (int32_t dt : t.scalar_dtypes)
    if (!is_supported(dt)) NCA_THROW(NCA_ERR_UNSUPPORTED, std::string("scalar dtype ") + dtype_info(dt).name + " unsupported");
  if (!is_supported(t.work_dtype)) NCA_THROW(NCA_ERR_UNSUPPORTED, "work dtype unsupported");
  if (!is_supported(t.dest_dtype)) NCA_THROW(NCA_ERR_UNSUPPORTED, "dest dtype unsupported");

  std::vector<int> stack;
  for (uint32_t ins : t.instrs) {
    int32_t op = instr_op(ins), idx = instr_index(ins);
    if (!op_supported(op))
      NCA_THROW(NCA_ERR_UNSUPPORTED, std::string("opcode ") + op_name(op) + " is not supported");
    Node n;
    n.op = op;
    if (op == NCA_OP_LOAD_NCARR || op == NCA_OP_LOAD_SOARR) {
      NCA_REQUIRE(idx < d.n_arrays, "LOAD index out of range");
      n.op = NCA_OP_LOAD_NCARR;
      n.index = idx;
    } else if (op == NCA_OP_LOAD_CONST) {
      NCA_REQUIRE(idx < d.n_scalars, "LOAD_CONST index out of range");
      n.index = idx;
    } else if (op == NCA_OP_IDX) {
      t.uses_idx = true;
    } else if (is_unary_op(op)) {
      NCA_REQUIRE(!stack.empty(), std::string("unary ") + op_name(op) + " on an empty stack");
      n.a = stack.back();
      stack.pop_back();
    } else {
      NCA_REQUIRE(stack.size() >= 2, std::string("binary ") + op_name(op) + " needs two operands");
      n.b = stack.back();
      stack.pop_back();
      n.a = stack.back();
      stack.pop_back();
    }
    stack.push_back(t.add(n));
  }
  NCA_REQUIRE(stack.size() == 1, "expression tape leaves " + std::to_string(stack.size()) + " values on the stack");
  t.root = stack.back();

  // Evaluation dtype, as the reference dispatches it: tapes that pass can_linearize run
  // the static JIT kernel in work_dtype (rtcompiler.cc:213-339, gpuengine.hh:103-160);
  // the others run the stack VM in the destination dtype, or in work_dtype when the
  // destination is bool (dynamicmvnode.hh:244-249).
  t.linearizable = can_linearize(t);
  if (t.linearizable || t.dest_dtype == NCA_BOOL)
    t.eval_dtype = t.work_dtype;
  else
    t.eval_dtype = t.dest_dtype;
  t.compute_dtype = ::nca::compute_dtype(t.eval_dtype);
  bool fp = is_float(t.compute_dtype);
  for (uint32_t ins : t.instrs) {
    int32_t op = instr_op(ins);
    bool bit = op == NCA_OP_BNOT || op == NCA_OP_BAND || op == NCA_OP_BOR || op == NCA_OP_XOR ||
               op == NCA_OP_LSHFT || op == NCA_OP_RSHFT;
    if (bit && fp)
      NCA_THROW(NCA_ERR_UNSUPPORTED, std::string(op_name(op)) + " on a floating-point evaluation dtype");
  }

  std::ostringstream sig;
  for (uint32_t ins : t.instrs) sig << std::hex << ins << ",";
  sig << "|a";
  for (int32_t dt : t.arr_dtypes) sig << std::dec << dt << ",";
  sig << "|s";
  for (int32_t dt : t.scalar_dtypes) sig << dt << ",";
  sig << "|w" << t.work_dtype << "|d" << t.dest_dtype;
  t.signature = sig.str();
  return t;
}

bool depends_on_loads(const Tape& t, int n) {
  if (n < 0) return false;
  const Node& x = t.nodes[n];
  if (x.op == NCA_OP_LOAD_NCARR || x.op == NCA_OP_IDX) return true;
  return depends_on_loads(t, x.a) || depends_on_loads(t, x.b);
}

// ---- symbolic differentiation ---------------------------------------------------------
static int d_add(Tape& t, int x, int y) {
  if (x < 0) return y;
  if (y < 0) return x;
  return t.bin(NCA_OP_ADD, x, y);
}
static int d_sub(Tape& t, int x, int y) {
  if (y < 0) return x;
  if (x < 0) return t.un(NCA_OP_NEG, y);
  return t.bin(NCA_OP_SUB, x, y);
}
static int d_mul(Tape& t, int x, int y) {  // x * y with x, y possibly zero (-1)
  if (x < 0 || y < 0) return -1;
  return t.bin(NCA_OP_MUL, x, y);
}

int differentiate(Tape& t, int n, int var, const std::vector<int>& alias) {
  Node x = t.nodes[n];  // copy: t.nodes may grow
  switch (x.op) {
    case NCA_OP_LOAD_NCARR:
      return alias[x.index] == var ? t.lit(1.0) : -1;
    case NCA_OP_LOAD_CONST: case NCA_OP_IDX: case NODE_LIT: case NODE_SIGN: case NODE_TRUNC:
      return -1;
    case NCA_OP_NEG: {
      int da = differentiate(t, x.a, var, alias);
      return da < 0 ? -1 : t.un(NCA_OP_NEG, da);
    }
    case NCA_OP_INC: case NCA_OP_DEC:
      return differentiate(t, x.a, var, alias);
    case NCA_OP_EXP: {
      int da = differentiate(t, x.a, var, alias);
      return d_mul(t, n, da);
    }
    case NCA_OP_LOG: {
      int da = differentiate(t, x.a, var, alias);
      return da < 0 ? -1 : t.bin(NODE_DIVRAW, da, x.a);
    }
    case NCA_OP_SQRT: {
      int da = differentiate(t, x.a, var, alias);
      if (da < 0) return -1;
      int two_s = t.bin(NCA_OP_MUL, t.lit(2.0), n);
      return t.bin(NODE_DIVRAW, da, two_s);
    }
    case NCA_OP_ABS: {
      int da = differentiate(t, x.a, var, alias);
      return d_mul(t, t.un(NODE_SIGN, x.a), da);
    }
    case NCA_OP_ADD:
      return d_add(t, differentiate(t, x.a, var, alias), differentiate(t, x.b, var, alias));
    case NCA_OP_SUB:
      return d_sub(t, differentiate(t, x.a, var, alias), differentiate(t, x.b, var, alias));
    case NCA_OP_MUL: {
      int da = differentiate(t, x.a, var, alias);
      int db = differentiate(t, x.b, var, alias);
      return d_add(t, d_mul(t, da, x.b), d_mul(t, x.a, db));
    }
    case NCA_OP_DIV: case NODE_DIVRAW: {
      // d(a/b) = da/b - (a/b) * db / b
      int da = differentiate(t, x.a, var, alias);
      int db = differentiate(t, x.b, var, alias);
      int first = da < 0 ? -1 : t.bin(NODE_DIVRAW, da, x.b);
      int second = -1;
      if (db >= 0) {
        int q = t.bin(NODE_DIVRAW, x.a, x.b);
        second = t.bin(NODE_DIVRAW, t.bin(NCA_OP_MUL, q, db), x.b);
      }
      return d_sub(t, first, second);
    }
    case NCA_OP_MOD: {
      // fmod(a, b) = a - trunc(a/b) * b, with trunc piecewise constant
      int da = differentiate(t, x.a, var, alias);
      int db = differentiate(t, x.b, var, alias);
      int second = db < 0 ? -1 : t.bin(NCA_OP_MUL, t.un(NODE_TRUNC, t.bin(NODE_DIVRAW, x.a, x.b)), db);
      return d_sub(t, da, second);
    }
    default:
      // comparisons, logical and bit operations, FDIV: piecewise constant
      return -1;
  }
}

// ---- code emission ---------------------------------------------------------------------
static std::string lit_text(double v, const std::string& C, bool fp) {
  char buf[64];
  if (fp)
    std::snprintf(buf, sizeof buf, "((%s)%.17g)", C.c_str(), v);
  else
    std::snprintf(buf, sizeof buf, "((%s)%lld)", C.c_str(), static_cast<long long>(v));
  return buf;
}

std::string emit_nodes(const Tape& t, const std::vector<int>& roots, const std::vector<int>& alias,
                       const std::vector<std::string>& load_names,
                       const std::vector<std::string>& scalar_names, const std::string& idx_name,
                       const std::string& prefix, std::vector<std::string>& names) {
  const int32_t cd = t.compute_dtype;
  const std::string C = ctype_of(cd);
  const bool fp = is_float(cd);
  const bool is_f32 = cd == NCA_FLOAT32;
  const bool is_bool = cd == NCA_BOOL;
  const bool is_signed_int = !fp && dtype_info(cd).is_signed;
  const int bits = 8 * itemsize(cd);
  std::vector<char> done(t.nodes.size(), 0);
  std::ostringstream os;
  auto var = [&](int n) { return prefix + std::to_string(n); };

  std::function<void(int)> visit = [&](int n) {
    if (n < 0 || done[n]) return;
    const Node& x = t.nodes[n];
    visit(x.a);
    visit(x.b);
    done[n] = 1;
    std::string A = x.a >= 0 ? var(x.a) : "", B = x.b >= 0 ? var(x.b) : "";
    std::string e;
    auto fn1 = [&](const char* f32, const char* f64) {
      if (fp) return std::string(is_f32 ? f32 : f64) + "(" + A + ")";
      return "(" + C + ")" + f64 + "((double)" + A + ")";
    };
    switch (x.op) {
      case NCA_OP_LOAD_NCARR: e = load_names[alias[x.index]]; break;
      case NCA_OP_LOAD_CONST: e = scalar_names[x.index]; break;
      case NCA_OP_IDX: e = "(" + C + ")(" + idx_name + ")"; break;
      case NODE_LIT: e = lit_text(x.lit, C, fp); break;
      case NCA_OP_NEG: e = is_bool ? A : "(" + C + ")(-" + A + ")"; break;
      case NCA_OP_INC: e = "(" + C + ")(" + A + " + (" + C + ")1)"; break;
      case NCA_OP_DEC: e = "(" + C + ")(" + A + " - (" + C + ")1)"; break;
      case NCA_OP_LNOT: e = "(" + C + ")(" + A + " == (" + C + ")0)"; break;
      case NCA_OP_BNOT: e = is_bool ? "(!" + A + ")" : "(" + C + ")(~" + A + ")"; break;
      case NCA_OP_ADD: e = "(" + C + ")(" + A + " + " + B + ")"; break;
      case NCA_OP_SUB: e = "(" + C + ")(" + A + " - " + B + ")"; break;
      case NCA_OP_MUL: e = is_bool ? "(" + A + " && " + B + ")" : "(" + C + ")(" + A + " * " + B + ")"; break;
      case NCA_OP_DIV:
        // interface.hh:150-156: a zero divisor gives NaN for a finite numerator and
        // leaves a non-finite one unchanged.  Integer evaluation: NaN converts to 0.
        if (fp) {
          std::string nan = is_f32 ? "__builtin_nanf(\"\")" : "__builtin_nan(\"\")";
          e = "((" + B + " == (" + C + ")0) ? (__builtin_isfinite(" + A + ") ? " + nan + " : " + A +
              ") : (" + C + ")(" + A + " / " + B + "))";
        } else if (is_bool) {
          e = "((!" + B + ") ? true : (bool)(" + A + "))";
        } else if (is_signed_int) {
          e = "((" + B + " == (" + C + ")0) ? (" + C + ")0 : (" + B + " == (" + C + ")-1) ? (" + C +
              ")(0ull - (unsigned long long)" + A + ") : (" + C + ")(" + A + " / " + B + "))";
        } else {
          e = "((" + B + " == (" + C + ")0) ? (" + C + ")0 : (" + C + ")(" + A + " / " + B + "))";
        }
        break;
      case NODE_DIVRAW: e = "(" + C + ")(" + A + " / " + B + ")"; break;
      case NCA_OP_MOD:
        if (fp) e = std::string(is_f32 ? "fmodf(" : "fmod(") + A + ", " + B + ")";
        else if (is_bool) e = "false";
        else if (is_signed_int)
          e = "((" + B + " == (" + C + ")0 || " + B + " == (" + C + ")-1) ? (" + C + ")0 : (" + C + ")(" + A + " % " + B + "))";
        else
          e = "((" + B + " == (" + C + ")0) ? (" + C + ")0 : (" + C + ")(" + A + " % " + B + "))";
        break;
      case NCA_OP_FDIV:
        if (fp) e = std::string(is_f32 ? "floorf(" : "floor(") + A + " / " + B + ")";
        else if (is_bool) e = "((bool)(" + A + " && " + B + "))";
        else if (is_signed_int)
          e = "((" + B + " == (" + C + ")0) ? (" + C + ")0 : (" + B + " == (" + C + ")-1) ? (" + C +
              ")(0ull - (unsigned long long)" + A + ") : (" + C + ")((" + A + " / " + B + ") - (((" + A + " % " + B +
              ") != 0) && ((" + A + " < 0) != (" + B + " < 0)))))";
        else
          e = "((" + B + " == (" + C + ")0) ? (" + C + ")0 : (" + C + ")(" + A + " / " + B + "))";
        break;
      case NCA_OP_EQ: e = "(" + C + ")(" + A + " == " + B + ")"; break;
      case NCA_OP_NE: e = "(" + C + ")(" + A + " != " + B + ")"; break;
      case NCA_OP_LT: e = "(" + C + ")(" + A + " < " + B + ")"; break;
      case NCA_OP_LE: e = "(" + C + ")(" + A + " <= " + B + ")"; break;
      case NCA_OP_GT: e = "(" + C + ")(" + A + " > " + B + ")"; break;
      case NCA_OP_GE: e = "(" + C + ")(" + A + " >= " + B + ")"; break;
      case NCA_OP_LAND: e = "(" + C + ")((" + A + " != (" + C + ")0) && (" + B + " != (" + C + ")0))"; break;
      case NCA_OP_LOR: e = "(" + C + ")((" + A + " != (" + C + ")0) || (" + B + " != (" + C + ")0))"; break;
      case NCA_OP_BAND: e = "(" + C + ")(" + A + " & " + B + ")"; break;
      case NCA_OP_BOR: e = "(" + C + ")(" + A + " | " + B + ")"; break;
      case NCA_OP_XOR: e = "(" + C + ")(" + A + " ^ " + B + ")"; break;
      case NCA_OP_LSHFT:
        e = is_bool ? A : "(" + C + ")(" + A + " << ((unsigned)" + B + " & " + std::to_string(bits - 1) + "u))";
        break;
      case NCA_OP_RSHFT:
        e = is_bool ? A : "(" + C + ")(" + A + " >> ((unsigned)" + B + " & " + std::to_string(bits - 1) + "u))";
        break;
      case NCA_OP_EXP: e = fn1("expf", "exp"); break;
      case NCA_OP_LOG: e = fn1("logf", "log"); break;
      case NCA_OP_SQRT: e = fn1("sqrtf", "sqrt"); break;
      case NCA_OP_ABS:
        if (fp) e = std::string(is_f32 ? "fabsf(" : "fabs(") + A + ")";
        else if (is_signed_int) e = "((" + A + " < (" + C + ")0) ? (" + C + ")(0ull - (unsigned long long)" + A + ") : " + A + ")";
        else e = A;
        break;
      case NODE_SIGN: e = "(" + C + ")((" + A + " > (" + C + ")0) - (" + A + " < (" + C + ")0))"; break;
      case NODE_TRUNC: e = fp ? std::string(is_f32 ? "truncf(" : "trunc(") + A + ")" : A; break;
      default:
        NCA_THROW(NCA_ERR_UNSUPPORTED, std::string("cannot emit opcode ") + op_name(x.op));
    }
    os << "      const " << C << " " << var(n) << " = " << e << ";\n";
  };
  names.clear();
  for (int r : roots) {
    if (r < 0) {
      names.push_back(lit_text(0.0, C, fp));
    } else {
      visit(r);
      names.push_back(var(r));
    }
  }
  return os.str();
}

}  // namespace nca
