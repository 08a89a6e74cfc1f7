// Expression tapes: the RPN instruction list of the reference's ExprMVNode
// (mvnode.hh:82-459) turned into an expression DAG, with the reference's evaluation
// semantics (interface.hh:129-181, staticmvnode.hh:142-181, dynamicmvnode.hh:146-251),
// symbolic differentiation, and emission of straight-line device code.
#ifndef NCA_TAPE_HH
#define NCA_TAPE_HH

#include <cstdint>
#include <string>
#include <vector>

#include "nca_hip.h"

namespace nca {

// Internal node kinds beyond the reference opcodes (used by derivative trees).
enum : int32_t {
  NODE_LIT = 1000,     // literal value `lit`
  NODE_DIVRAW = 1001,  // a / b without the reference's zero-divisor rule
  NODE_SIGN = 1002,    // sign(a) in {-1, 0, 1}
  NODE_TRUNC = 1003    // trunc(a)
};

struct Node {
  int32_t op = NCA_OP_NOOP;
  int32_t a = -1;      // first child
  int32_t b = -1;      // second child
  int32_t index = -1;  // LOAD_*: array operand; LOAD_CONST: scalar
  double lit = 0.0;    // NODE_LIT
};

struct Tape {
  std::vector<uint32_t> instrs;
  std::vector<int32_t> arr_dtypes;
  std::vector<int32_t> scalar_dtypes;
  int32_t work_dtype = NCA_FLOAT32;
  int32_t dest_dtype = NCA_FLOAT32;

  // derived
  std::vector<Node> nodes;
  int32_t root = -1;
  bool linearizable = false;  // can_linearize (staticmvnode.hh:188-239)
  int32_t eval_dtype = NCA_FLOAT32;     // dtype the reference evaluates in
  int32_t compute_dtype = NCA_FLOAT32;  // eval_dtype, with bf16/fp16 -> float32
  bool uses_idx = false;
  std::string signature;  // canonical text of the tape, part of kernel cache keys

  static Tape from_desc(const nca_expr_desc& d);
  int add(const Node& n) { nodes.push_back(n); return static_cast<int>(nodes.size()) - 1; }
  int lit(double v) { Node n; n.op = NODE_LIT; n.lit = v; return add(n); }
  int un(int32_t op, int x) { Node n; n.op = op; n.a = x; return add(n); }
  int bin(int32_t op, int x, int y) { Node n; n.op = op; n.a = x; n.b = y; return add(n); }
};

// Derivative of node `n` with respect to the value variable `var`, where LOAD k
// refers to variable alias[k].  Returns a node index, or -1 for an exact zero.
int differentiate(Tape& t, int n, int var, const std::vector<int>& alias);
// True when the subtree references any LOAD/IDX node.
bool depends_on_loads(const Tape& t, int n);

// Emits SSA statements `C tN = ...;` for all nodes reachable from `roots`, in
// dependency order.  Operand k's value is the C expression `load_names[alias[k]]`;
// scalar i is `scalar_names[i]`; IDX is `idx_name`.  Returns code; `names` gets the
// variable holding each root.
std::string emit_nodes(const Tape& t, const std::vector<int>& roots,
                       const std::vector<int>& alias,
                       const std::vector<std::string>& load_names,
                       const std::vector<std::string>& scalar_names,
                       const std::string& idx_name, const std::string& prefix,
                       std::vector<std::string>& names);

std::string ctype_of(int32_t dtype);

}  // namespace nca

#endif  // NCA_TAPE_HH
