// Shared declarations of the expression, stencil and batch code generators.
#ifndef NCA_EXPR_INTERNAL_HH
#define NCA_EXPR_INTERNAL_HH

#include <deque>
#include <functional>
#include <unordered_map>

#include "nca_internal.hh"
#include "tape.hh"

namespace nca {

struct ExprProgram {
  Tape tape;
  std::mutex mu;
  std::unordered_map<std::string, hipFunction_t> kernels;  // variant key -> kernel
  std::unordered_map<std::string, std::string> sources;    // variant key -> source
};

ExprProgram& expr_program(uint64_t h);
uint64_t expr_register(Tape&& t);

// Scalars converted to the compute type, one 8-byte slot each (kernel parameters).
struct ScalarSlots {
  std::vector<uint64_t> slots;
};
ScalarSlots make_scalar_slots(const Tape& t, const nca_scalar* s, int n);

// How the operands of one launch look: aliasing, broadcasting, contiguity.
struct Operands {
  int64_t n = 0;                 // output elements
  std::vector<int> alias;        // operand k reads the same values as operand alias[k]
  std::vector<char> bcast;       // operand k is one value broadcast to every element
  std::vector<int> unique;       // operands with alias[k] == k
  bool contiguous = true;
  int vec = 1;
  int sndim = 1;                 // strided kind: coalesced shape and byte strides
  std::vector<int64_t> sshape;
  std::vector<int64_t> sout;
  std::vector<std::vector<int64_t>> sstrides;
  // Start addresses passed to the kernel: the view's data, or for an indirect view the start
  // of its dereference chain.
  std::vector<const void*> base;
  const void* baseo = nullptr;
  // Indirect views (strided kind only, axes not coalesced) in the form of SOArrayPolicy::advance
  // (layout.hh:677-700): after the stride step of axis ind[k][j] the address becomes the
  // pointer stored there plus isub[k][j] bytes.  NC pointer-axis views enter rewritten to that
  // form (one step at the pointer axis); indo/isubo describe the output.
  bool nc = false;
  std::vector<std::vector<int>> ind;
  std::vector<std::vector<int64_t>> isub;
  std::vector<int> indo;
  std::vector<int64_t> isubo;
};
Operands analyse_operands(const Tape& t, const nca_view* ins, int n_in, const nca_view& out,
                          bool out_may_be_null);

int grid_blocks(int64_t work_items, int threads, int max_blocks);
hipFunction_t program_kernel(ExprProgram& p, const std::string& key,
                             const std::function<std::string()>& gen, const std::string& kname);

// Ragged-batch kernel source (batch.cpp); alias[k] names the operand whose arrays operand k
// reads (empty: none alias).
std::string gen_batch_source(const Tape& t, int fuse, int mode, int vec, bool with_out,
                             const std::vector<int>& alias = {});

}  // namespace nca

#endif  // NCA_EXPR_INTERNAL_HH
