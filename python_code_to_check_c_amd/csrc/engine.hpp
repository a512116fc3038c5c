// Tensor-expression engine: the residual builders describe each reference
// einsum as a node; identical nodes are shared (which hoists everything that
// the four RK4 chains and their stages have in common); the node list is then
// planned into one device arena by liveness and lowered to kernel launches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace tdcc {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);
#define TDCC_HIP(x) ::tdcc::hip_check((x), #x)

enum class Cls : uint8_t { O = 0, V = 1, N = 2 };  // occupied, virtual, all orbitals

struct Shape {
  int rank = 0;
  Cls cls[4] = {Cls::O, Cls::O, Cls::O, Cls::O};
};

struct Dev {  // device data of a tensor: planes + element strides
  double* re = nullptr;
  double* im = nullptr;
  int64_t st[4] = {0, 0, 0, 0};
};

enum class Op { External, View, Contract, Lincomb, Field, Finalize, Trace, Assign, Ladder, FieldVjp, Seed, TraceVjp };

struct Term {
  int src;
  int perm[4];  // output axis d reads source axis perm[d]
  double cr, ci;
  // product term (outer product of two tensors folded in at planning): the
  // value is c * src * src2, output axis d reading axis perm[d] of src and
  // perm2[d] of src2 (-1: not an axis of that factor)
  int src2 = -1;
  int perm2[4] = {-1, -1, -1, -1};
};

struct Node {
  Op op;
  Shape shape;
  bool cplx = false;
  std::string key;
  // Contract
  std::string spec;
  int a = -1, b = -1;
  double ar = 1.0, ai = 0.0;
  // Lincomb
  std::vector<Term> terms;
  // View
  int base = -1;
  int64_t voff = 0;
  int64_t vst[4] = {0, 0, 0, 0};
  // Field
  int stage = -1;
  // Field / Trace of a batched program: trajectory index into Runtime::batch
  // (-1: the runtime's single-trajectory parameters)
  int traj = -1;
  // External
  Dev ext;
  // Finalize: inputs k[4], target external `a`, base external `b`
  int k[4] = {-1, -1, -1, -1};
  int lambda_sign = 0;
  // Trace: inputs dij, l1, dai, dab in k[0..3]
  // Assign: src in a, destination external in b
  // planning
  int last_use = -1;
  int64_t arena_off = -1;
  int64_t scratch_off = -1;  // split-K partials
  int64_t scratch_len = 0;
  // stays FP64 when the runtime runs the large contractions in split precision
  bool fp64 = false;
  // Lincomb: materialized (never folded into the linear combinations reading it)
  bool keep = false;
  // Ladder: only the i < j rows are computed (input antisymmetric in (i, j))
  bool ij_pairs = false;
  // Contract: two summed letters p, q of which only the terms p < q are
  // summed (empty: all), and the factor the launch applies then
  std::string kpair;
  double kfac = 1.0;
  // Contract: letters of an antisymmetric output pair among the rows (M) and
  // among the columns (N) of a GEMM-path launch, computed for index(p) <=
  // index(q) only and mirrored by the epilogue (empty: none)
  std::string mpair, npair;
  // antisymmetric axis pairs: bit anti_bit(d1, d2) set when the tensor
  // changes sign under exchanging axes d1 and d2
  uint8_t anti = 0;
  // unchanged under exchanging axes 0, 1 and axes 2, 3 together (the outer
  // product 'ia,jb->ijab' of a tensor with itself)
  bool xsym = false;
  // stacked trajectories (Graph::set_stack): the node holds one value per
  // trajectory z, trajectory z at element offset z * zs from trajectory 0
  // (zdep false, zs 0: one value shared by all trajectories)
  bool zdep = false;
  int64_t zs = 0;
  // Contract: the launches run once per trajectory (both operands per
  // trajectory on a dense-GEMM path) instead of once over the stacked rows
  bool zloop = false;
};

struct Handle {
  int id = -1;
};

class Graph;

// Runtime parameters shared by a graph's kernels (owned by the context).
struct Runtime {
  int o = 0, v = 0;
  hipStream_t stream = nullptr;
  const double* F = nullptr;        // (n,n) real Fock
  const double* mu = nullptr;       // (3,n,n) real dipoles
  int pol = 2;
  double negA = 0.0;
  const double* field_table = nullptr;  // [steps][3][2]
  int64_t* counter = nullptr;
  double* mu_out = nullptr;         // [steps][3][2]
  double dt6 = 0.0, round_scale = 1e15;
  // per-trajectory field and dipole parameters of batched programs
  struct Traj {
    int pol = 2;
    double negA = 0.0;
    const double* field_table = nullptr;  // [steps][3][2]
    double* mu_out = nullptr;             // [steps][3][2]
    bool operator==(const Traj& o) const {
      return pol == o.pol && negA == o.negA && field_table == o.field_table && mu_out == o.mu_out;
    }
  };
  std::vector<Traj> batch;
  // adjoint programs: per-step loss weights [steps], stage times [steps][3],
  // gradient accumulators g[4] = dL/dA, dL/dw0 (complex) and dL/dmu_pol (n,n)
  const double* wtab = nullptr;
  const double* ttab = nullptr;
  double* g = nullptr;
  double* gmu_re = nullptr;
  double* gmu_im = nullptr;
  // packed antisymmetric <ab||ef>: vpack[0][q(ab)][q(ef)] = <ab||ef>,
  // vpack[1][q(ab)][q(ef)] = <ef||ab> (P = v(v-1)/2 pairs a < b)
  const double* vpack[2] = {nullptr, nullptr};
  // Split precision: 0 = FP64 everywhere; 2..kOzMaxSlices = the dense-GEMM
  // contractions and the ladders run on int8 digit planes with this many
  // digits per operand (everything else stays FP64).
  int split_digits = 0;
  // digit planes of vpack[k]: split_digits planes of P x vpack_Kp bytes,
  // vpack_plane bytes apart, and the P row scales
  const int8_t* vpack8[2] = {nullptr, nullptr};
  const double* vpack8_scale[2] = {nullptr, nullptr};
  int64_t vpack_Kp = 0, vpack_plane = 0;
};

class Graph {
 public:
  Graph(int o, int v, const Runtime* rt);
  ~Graph();
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;

  int64_t extent(Cls c) const { return c == Cls::O ? o_ : (c == Cls::V ? v_ : o_ + v_); }
  int64_t size(const Shape& s) const;
  const Node& node(int id) const { return nodes_[id]; }
  int count() const { return (int)nodes_.size(); }

  // ---- construction ----
  int external(const Shape& s, bool cplx, const Dev& d, const std::string& name);
  int view(int base, const Shape& s, int64_t off, const int64_t st[4]);
  int block(int base_n4, const char* classes);       // TEI block, e.g. "ovvv"
  int fblock(int base_nn, const char* classes);      // (n,n) block, e.g. "ov"
  int contract(const std::string& spec, int a, int b, double ar = 1.0, double ai = 0.0);
  // Declares an external antisymmetric under the given axis pairs (bits of
  // anti_bit); views of integral blocks, linear combinations, contractions
  // and ladders derive theirs from their inputs.
  void set_anti(int id, uint8_t mask) { nodes_[id].anti = mask; }
  // A contraction summing over two letters in which both operands are
  // antisymmetric has a summand symmetric under exchanging them: it is then
  // lowered as twice the sum over the pairs p < q.  Only while enabled.
  void set_k_pairs(bool on) { k_pairs_ = on; }
  bool k_pairs() const { return k_pairs_; }
  int contract3(const std::string& spec, int a, int b, int c, double ar = 1.0);
  int lincomb(std::vector<Term> terms);
  // a linear combination kept as its own buffer (a layout change read by several
  // linear combinations, not folded into them)
  int materialize(std::vector<Term> terms);
  int field(int stage);
  // Field and Trace nodes created from now on belong to trajectory b of a
  // batched program (-1: not batched).
  void set_traj(int b) { traj_ = b; }
  // Stacked trajectories: the graph describes one trajectory and runs it for
  // nz of them at once.  Externals declared per trajectory (set_zstride) and
  // Field nodes (one field per trajectory, from Runtime::batch) carry a
  // trajectory axis, and so does every node computed from them; nodes of
  // integrals alone are shared.  Contractions with one per-trajectory operand
  // stack the trajectories into that operand's free rows (one GEMM against
  // the shared operand); linear combinations, GETT batches and the rest take
  // the trajectory axis in their launch, or run once per trajectory.
  void set_stack(int nz) { nz_ = nz < 1 ? 1 : nz; }
  int stack() const { return nz_; }
  void set_zstride(int ext_id, int64_t zs) { nodes_[ext_id].zs = zs; }
  // Particle-particle ladder over the antisymmetric integral block:
  //   ket == 0: out_ijab = sum_ef <ab||ef> u_ijef   ('abef,ijef->ijab')
  //   ket == 1: out_ijab = sum_ef <ef||ab> u_ijef   ('efab,ijef->ijab')
  // computed over pairs e < f and a < b (exact for antisymmetrized integrals).
  int ladder(int u, int ket);
  // Ladders created from now on compute only the i < j rows and mirror them
  // (valid when every ladder input is antisymmetric in its first two indices,
  // as the amplitudes of the propagation are).
  void set_ij_pairs(bool on) { ij_pairs_ = on; }
  int ones_view(const Shape& s);                     // all-ones tensor (stride 0)
  // While an Fp64Scope is open, every node created or shared is pinned to
  // FP64 arithmetic regardless of the runtime's split precision.
  struct Fp64Scope {
    Graph& g;
    explicit Fp64Scope(Graph& gr) : g(gr) { ++g.fp64_depth_; }
    ~Fp64Scope() { --g.fp64_depth_; }
  };
  void finalize_rk(int y_ext, int base_ext, const int k[4], bool lambda);

  // ---- reverse mode ----
  // Adjoints of the nodes built so far (holomorphic maps: transposed
  // contractions and permutations, complex coefficients as they are).
  // seeds: (node, term over an adjoint node) pairs.  Returns the summed
  // adjoint of every non-constant External reached, and the adjoints of the
  // views of Field nodes (the field's blocks).  Rounding does not appear in
  // the graphs differentiated here.
  struct VjpResult {
    std::map<int, int> ext;                        // external id -> adjoint node
    std::vector<std::pair<int, int>> field_views;  // (view of a Field node, adjoint node)
  };
  VjpResult vjp(const std::vector<std::pair<int, Term>>& seeds);
  // Field adjoint of one view of a stage field into the runtime's gradients.
  void field_vjp(int adj, int view);
  // w mu_pol restricted to a (row class, col class) block, w from the
  // runtime's weight table.
  int seed(Cls rows, Cls cols);
  void trace_vjp(int dij, int l1, int dai, int dab);
  void trace(int dij, int l1, int dai, int dab);
  void assign(int dst_ext, int src);

  // ---- lowering ----
  void plan();                  // liveness + arena layout + offset tables
  int64_t arena_bytes() const { return arena_bytes_; }
  void bind(double* arena);     // resolve device pointers, build launch records
  // Autotuner of the digit GEMM launches (tile order, persistent grid, K
  // splits up to the planned count), run at bind on the bound operands; on by
  // default for stacked programs, TDCC_AUTOTUNE=1 / 0 forces it on / off.
  // Choices are kept per launch shape for the process and appended to the
  // file TDCC_AUTOTUNE_FILE names.
  void autotune(bool k_splits);
  void run(hipStream_t s) const;
  // Step programs: the externals holding the Lambda amplitudes.  The nodes
  // that depend on them, and those only they use, run on a second stream
  // concurrently with the T side (the Lambda chains read the T amplitudes of
  // the step's start, which nothing writes before the finalize), joined before
  // the finalize; TDCC_STREAMS=1 keeps one stream.
  void set_lambda_externals(std::vector<int> ext) { lam_ext_ = std::move(ext); }

  // One eager pass with a HIP event pair around every kernel launch.
  struct LaunchRecord {
    std::string name;   // kernel name incl. template variant
    double flops;       // algorithmic real flops of the launch (contractions)
    double bytes;       // algorithmic bytes moved (elementwise kernels)
    float ms;
    std::string detail; // contraction string and M, N, K, batch, splits
  };
  void run_profiled(hipStream_t s, std::vector<LaunchRecord>& out) const;

  // statistics
  int64_t contract_flops() const;  // executed real flops of all contractions

 private:
  int add(Node n);
  Dev resolve(int id, int z = 0) const;
  void mark_stack();               // zdep / zs of every live node (plan)
  int64_t zbytes(const Node& n) const;  // bytes of one trajectory's buffer
  void dump_nodes() const;  // TDCC_DUMP_NODES=1 at plan
  void lower_contract(int id);
  void lower_lincomb(int id);

  int o_, v_;
  const Runtime* rt_;
  std::vector<Node> nodes_;
  std::map<std::string, int> cse_;
  int64_t arena_bytes_ = 0;
  double* arena_ = nullptr;

  struct Launch;
  std::vector<std::unique_ptr<Launch>> launches_;
  std::vector<int64_t> host_tables_;
  int64_t* dev_tables_ = nullptr;
  std::vector<char> live_;
  int ones_id_ = -1;
  int fp64_depth_ = 0;
  int traj_ = -1;
  int nz_ = 1;
  bool ij_pairs_ = false;
  bool k_pairs_ = false;
  // digit planes of static operands (split precision), filled at bind
  bool is_static(int id) const;
  void simplify_split_tables(SplitArgs& a, int64_t rtab_off, int64_t ktab_off) const;
  bool linear_table(int64_t tab, int64_t K, int64_t& step, int64_t& first) const;
  struct DigitCache {
    int64_t digits_off = 0, scale_off = 0;
    bool filled = false;
  };
  std::vector<DigitCache> cache_;
  std::map<std::string, int> cache_index_;
  // digit planes of an operand computed by the program that several split
  // launches read in the same layout: cut by the first launch into an arena
  // block that stays allocated until the last one
  struct DigitShare {
    int64_t off = 0, bytes = 0;
    int first = -1, last = -1;
  };
  std::vector<DigitShare> shares_;
  struct DynUse {
    int node, side, src;
    int64_t bytes;
  };
  std::map<std::string, std::vector<DynUse>> dyn_uses_;  // layout key -> launches in node order
  void plan_shares();
  void inline_lincombs(std::vector<char>& live);
  void fold_outer_products(std::vector<char>& live);
  int64_t cache_bytes_ = 0;
  char* cache_buf_ = nullptr;
  double* ones_buf_ = nullptr;
  // two-stream step (set_lambda_externals): side of every live node (0: T
  // side, stream of run(); 1: Lambda side, second stream; 2: finalize and
  // after, on the first stream once the second has joined), and whether a T
  // side node is read by the Lambda side (its buffer is then never reused
  // within the step)
  std::vector<int> lam_ext_;
  std::vector<int> side_;
  std::vector<char> cross_;
  void compute_sides();
  // execution order of the nodes (a topological order; node order in a
  // one-stream program; in a two-stream step the crossing nodes and their
  // ancestors first, then the rest up to the first finalize, then the rest)
  // and each node's position in it
  std::vector<int> order_, pos_;
  void plan_streams();  // at bind, from the sides
  // launch schedule: stream of each launch, events to wait on before it,
  // event to record after it, index of the first launch after the join
  std::vector<int> lstream_, lrecord_;
  std::vector<std::vector<int>> lwait_;
  int first_post_ = -1;
  hipStream_t sb_ = nullptr;
  std::vector<hipEvent_t> events_;
  hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr;
};

constexpr uint8_t anti_bit(int d1, int d2) {
  return (uint8_t)(1u << (d1 < d2 ? (d1 == 0 ? d2 - 1 : (d1 == 1 ? d2 + 1 : 5)) : (d2 == 0 ? d1 - 1 : (d2 == 1 ? d1 + 1 : 5))));
}

// Elementwise helpers for the builders.
Term tm(int src, double c = 1.0);                           // identity permutation
Term tmp(int src, std::initializer_list<int> perm, double c = 1.0);

}  // namespace tdcc
