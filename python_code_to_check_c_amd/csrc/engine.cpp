// Tensor-expression engine: construction with structural sharing, planning
// (dead-node removal, liveness, arena layout, GETT offset tables) and lowering
// to kernel launches.
#include "engine.hpp"

#include <algorithm>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <set>
#include <sstream>

namespace tdcc {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw Error(std::string("HIP error: ") + hipGetErrorString(e) + " in " + what);
  }
}

static std::string hexf(double x) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", x);
  return buf;
}

Term tm(int src, double c) { return Term{src, {0, 1, 2, 3}, c, 0.0}; }

Term tmp(int src, std::initializer_list<int> perm, double c) {
  Term t{src, {0, 1, 2, 3}, c, 0.0};
  int d = 0;
  for (int p : perm) t.perm[d++] = p;
  return t;
}

struct LadderArgs {
  const double* u_re = nullptr;
  const double* u_im = nullptr;
  double* out_re = nullptr;
  double* out_im = nullptr;
  double* up = nullptr;   // (2 rows) x P packed input
  double* rp = nullptr;   // (2 rows) x P packed result
  const double* vp = nullptr;
  int o = 0;
  bool ij_pairs = false;  // rows are the pairs i < j
  int splits = 1;         // K chunks of the GEMM (partial results summed before the unpack)
  int64_t k_chunk = 0;
  int64_t rows = 0;        // rows of one trajectory
  int nz = 1;              // stacked trajectories: rows of trajectory z at z * 2 rows of up / rp
  int64_t u_zs = 0, out_zs = 0;
  int v = 0;
  int64_t P = 0;
  // split precision (S > 0): digits of up, and the digit GEMM against vpack8
  int S = 0;
  SplitArgs su{};
  OzGemmArgs oz{};
};

// A contraction lowered to dense GEMM form: gather A' (planes stacked along
// rows, pa*M x K), gather B' (pb*N x K when bt, else K x pb*N), one DGEMM
// (optionally split over K) into Z (pa*M x pb*N per split), scatter into C.
// In split precision (S > 0) the gathers become digit splits of A' and of
// B' (always pb*N x K), and the DGEMM the int8 digit GEMM.
struct GemmPath {
  bool on = false;
  bool bt = false;
  int pa = 1, pb = 1;
  int splits = 1;
  int64_t k_chunk = 0;
  int64_t a_off = 0, b_off = 0, z_off = 0;  // element offsets inside the node scratch
  int64_t ldb = 0;
  // split precision: digits, split orientation, padded K, byte offsets
  int S = 0;
  bool a_cols = false, b_cols = false;
  int64_t Kp = 0, a_plane = 0, b_plane = 0;
  bool dil = false;  // digit-interleaved rows for both operands (short K)
  // scales per K split of A' / B' (rows too long for a single-pass split)
  bool a_chunked = false, b_chunked = false;
  int64_t a8_off = 0, b8_off = 0, sa_off = 0, sb_off = 0, pa_off = 0, pb_off = 0, zb_off = 0;
  // operands that never change while the program is bound (integral blocks)
  // are split once per bind into the graph's digit cache: entry index or -1
  int a_cache = -1, b_cache = -1;
  // operands computed by the program whose digits several launches share:
  // entry of the graph's share table (or -1), and whether this launch cuts them
  int a_share = -1, b_share = -1;
  bool a_split = true, b_split = true;
  // one split: the digit GEMM scatters into C itself (planes interleaved by row)
  bool fused = false;
  // complex x complex from three real products (3M): digit rows [re; im;
  // re + im] of each operand (3 M and 3 N rows, planes not interleaved); one
  // launch computes P1 = A_re B_re^T and P2 = A_im B_im^T into the node
  // scratch at p12_off, a second P3 = A_s B_s^T with the combining epilogue
  bool m3 = false;
  int64_t p12_off = 0;
  // dot-product kernel (FP64) instead of the GEMM: partials in the node scratch
  bool dot = false;
};

struct Graph::Launch {
  enum Kind { GETT, LIN, FIELD, FIN, TRACE, LADDER, FVJP, SEED, TVJP } kind;
  int node = -1;
  LadderArgs lad{};
  GemmPath gp{};
  Gather2dArgs ga{}, gb{};
  DgemmArgs dg{};
  SplitArgs sa{}, sb{};
  OzGemmArgs oz{}, oz2{};  // oz2: the 3M path's second launch
  GemmScatterArgs sc{};
  DotArgs dt{};
  // skinny path: one operand with at most kSkinnyMax rows (small_is_a: the A side)
  bool skinny = false, small_is_a = true;
  SkinnyArgs sk{};
  GettArgs g{};
  bool ac = false, bc = false, akc = false, bkc = false, c_cplx = false;
  int64_t tab[11] = {0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1};  // am ak bn bk cm cn ab bb cb cm2 cn2 offsets into tables
  LincombArgs l{};
  FieldArgs f{};
  FinalizeArgs fin{};
  TraceArgs tr{};
  FieldVjpArgs fv{};
  SeedArgs sd{};
  TraceVjpArgs tv{};
};

Graph::Graph(int o, int v, const Runtime* rt) : o_(o), v_(v), rt_(rt) {}

Graph::~Graph() {
  if (dev_tables_) (void)hipFree(dev_tables_);
  if (ones_buf_) (void)hipFree(ones_buf_);
  if (cache_buf_) (void)hipFree(cache_buf_);
  for (hipEvent_t e : events_) (void)hipEventDestroy(e);
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  if (join_ev_) (void)hipEventDestroy(join_ev_);
  if (sb_) (void)hipStreamDestroy(sb_);
}

int64_t Graph::size(const Shape& s) const {
  int64_t n = 1;
  for (int d = 0; d < s.rank; ++d) n *= extent(s.cls[d]);
  return n;
}

static void contiguous_strides(const Graph& g, const Shape& s, int64_t st[4]) {
  int64_t acc = 1;
  for (int d = s.rank - 1; d >= 0; --d) {
    st[d] = acc;
    acc *= g.extent(s.cls[d]);
  }
  for (int d = s.rank; d < 4; ++d) st[d] = 0;
}

int Graph::add(Node n) {
  if (!n.key.empty()) {
    auto it = cse_.find(n.key);
    if (it != cse_.end()) {
      if (fp64_depth_ > 0) nodes_[it->second].fp64 = true;
      return it->second;
    }
  }
  n.fp64 = fp64_depth_ > 0;
  const int id = (int)nodes_.size();
  if (n.key.empty()) n.key = "#" + std::to_string(id);
  cse_[n.key] = id;
  nodes_.push_back(std::move(n));
  return id;
}

int Graph::external(const Shape& s, bool cplx, const Dev& d, const std::string& name) {
  Node n;
  n.op = Op::External;
  n.shape = s;
  n.cplx = cplx;
  n.ext = d;
  n.key = "X|" + name + "|" + std::to_string(nodes_.size());
  return add(std::move(n));
}

int Graph::view(int base, const Shape& s, int64_t off, const int64_t st[4]) {
  Node n;
  n.op = Op::View;
  n.shape = s;
  n.cplx = nodes_[base].cplx;
  n.base = base;
  n.voff = off;
  std::ostringstream k;
  k << "V|" << base << "|" << off;
  for (int d = 0; d < 4; ++d) {
    n.vst[d] = st[d];
    k << "|" << st[d];
  }
  n.key = k.str();
  return add(std::move(n));
}

static Cls cls_of(char c) {
  if (c == 'o') return Cls::O;
  if (c == 'v') return Cls::V;
  throw Error(std::string("bad orbital class ") + c);
}

int Graph::block(int base, const char* classes) {
  const int64_t n = o_ + v_;
  const int64_t st[4] = {n * n * n, n * n, n, 1};
  Shape s;
  s.rank = 4;
  int64_t off = 0;
  for (int d = 0; d < 4; ++d) {
    s.cls[d] = cls_of(classes[d]);
    if (s.cls[d] == Cls::V) off += o_ * st[d];
  }
  const int id = view(base, s, off, st);
  // the base's antisymmetric axis pairs that stay inside the block
  uint8_t m = 0;
  if (nodes_[base].shape.rank == 4)
    for (int d1 = 0; d1 < 4; ++d1)
      for (int d2 = d1 + 1; d2 < 4; ++d2)
        if ((nodes_[base].anti & anti_bit(d1, d2)) && s.cls[d1] == s.cls[d2]) m |= anti_bit(d1, d2);
  nodes_[id].anti = m;
  return id;
}

int Graph::fblock(int base, const char* classes) {
  const int64_t n = o_ + v_;
  const int64_t st[4] = {n, 1, 0, 0};
  Shape s;
  s.rank = 2;
  int64_t off = 0;
  for (int d = 0; d < 2; ++d) {
    s.cls[d] = cls_of(classes[d]);
    if (s.cls[d] == Cls::V) off += o_ * st[d];
  }
  return view(base, s, off, st);
}

// ---------------------------------------------------------------------------
static void parse_spec(const std::string& spec, std::vector<std::string>& ins, std::string& out) {
  const auto arrow = spec.find("->");
  if (arrow == std::string::npos) throw Error("einsum spec without '->': " + spec);
  out = spec.substr(arrow + 2);
  std::string lhs = spec.substr(0, arrow);
  ins.clear();
  std::string cur;
  for (char c : lhs) {
    if (c == ',') {
      ins.push_back(cur);
      cur.clear();
    } else if (c != ' ') {
      cur.push_back(c);
    }
  }
  ins.push_back(cur);
}

static bool has(const std::string& s, char c) { return s.find(c) != std::string::npos; }

// Contractions with at least this many flops (and GEMM-sized groups) take the
// dense-GEMM path; TDCC_GEMM_MIN_FLOPS overrides it (0: every batch-free one).
static double gemm_min_flops() {
  const char* e = std::getenv("TDCC_GEMM_MIN_FLOPS");
  return e ? std::atof(e) : 2e9;
}

// TDCC_SPLIT_MULTIPASS=1 makes every digit split use the three-pass kernels.
static bool split_multipass() {
  const char* e = std::getenv("TDCC_SPLIT_MULTIPASS");
  return e && std::atoi(e) != 0;
}

// Elements of an integral operand above which a skinny contraction runs on
// the split GEMM path (TDCC_STATIC_STREAM_MIN; negative: never).
static int64_t static_stream_min() {
  const char* e = std::getenv("TDCC_STATIC_STREAM_MIN");
  return e ? std::atoll(e) : ((int64_t)1 << 23);
}

// Summed length below which a GEMM-path launch computes every column of an
// output antisymmetric in two column letters instead of their ordered pairs:
// with few K steps the launch is bound by its stores, and the mirror image of
// a column pair lands one column stride apart per lane (TDCC_NPAIR_MIN_K).
static int64_t npair_min_k() {
  const char* e = std::getenv("TDCC_NPAIR_MIN_K");
  return e ? std::atoll(e) : 1024;
}

// Antisymmetric linear combinations summed at their canonical elements and
// mirrored (TDCC_LINCOMB_ANTI=0: every element summed).
static bool lincomb_anti_on() {
  const char* e = std::getenv("TDCC_LINCOMB_ANTI");
  return !e || std::atoi(e) != 0;
}

// Summed length from which dot-product shapes take the dot kernel
// (TDCC_DOT_MIN_K; negative: never).
static int64_t dot_min_k() {
  const char* e = std::getenv("TDCC_DOT_MIN_K");
  return e ? std::atoll(e) : 8192;
}

// TDCC_GETT_LINEAR_K=0: GETT launches read every K offset from its table.
static bool gett_linear_k_on() {
  const char* e = std::getenv("TDCC_GETT_LINEAR_K");
  return !e || std::atoi(e) != 0;
}

// Skinny contractions with at least this much big-operand work (rows x K) take
// the VALU streaming kernel; TDCC_SKINNY_MIN overrides it (0: all skinny
// shapes, a negative value: none).
static double skinny_min_work() {
  const char* e = std::getenv("TDCC_SKINNY_MIN");
  const double v = e ? std::atof(e) : 65536.0;
  return v < 0 ? 1e300 : v;
}

// TDCC_SPLIT_CHUNKED=1: scales per K split for every row-oriented operand of a
// split-K digit GEMM with K > 256 (by default only for rows over 16384).
static bool split_chunked_always() {
  const char* e = std::getenv("TDCC_SPLIT_CHUNKED");
  return e && std::atoi(e) != 0;
}

// Elements per block (run_rows x run) from which a k-major skinny contraction
// with short runs takes the short-run kernels; TDCC_SKINNY_RUNS_MIN
// overrides it (0: every such shape, a negative value: none).
static int64_t skinny_runs_min() {
  const char* e = std::getenv("TDCC_SKINNY_RUNS_MIN");
  const int64_t v = e ? std::atoll(e) : 512;
  return v < 0 ? INT64_MAX : v;
}

// K splits of a digit GEMM whose tile grid fills few rounds of the CUs: the
// split count (1 .. 64, chunks a multiple of 64) that minimizes a cost model
// in units of one tile's 32-wide K step: rounds x (K steps per split + 3 for
// a tile's prologue and epilogue), plus, with splits, writing and summing the
// partial products (about one K step per 1.1 us at 4 TB/s).  The tail of a
// grid just over a whole number of rounds is what it avoids (e.g. 175 tiles
// x 3 splits = 2.05 rounds).  TDCC_SPLIT_MODEL=0 keeps the fixed rule.
constexpr int64_t kModelCUs = 256;
static bool split_model_on() {
  const char* e = std::getenv("TDCC_SPLIT_MODEL");
  return !e || std::atoi(e) != 0;
}
static int64_t model_splits(int64_t tiles, int64_t Kp, int64_t R, int64_t NB, int64_t planned) {
  auto cost = [&](int64_t s) {
    int64_t chunk = (Kp + s - 1) / s;
    chunk = (chunk + 63) / 64 * 64;
    const int64_t sp = (Kp + chunk - 1) / chunk;
    const double rounds = (double)((tiles * sp + kModelCUs - 1) / kModelCUs);
    double t = rounds * ((double)chunk / 32.0 + 3.0);
    if (sp > 1) t += (double)(sp + 1) * R * NB * 8.0 / 4e12 / 1.1e-6;
    return t;
  };
  int64_t best = planned;
  double bc = cost(planned);
  for (int64_t s = 1; s <= 64 && s * 64 <= Kp; ++s) {
    const double c = cost(s);
    if (c < bc * 0.97) {  // a clear gain only
      bc = c;
      best = s;
    }
  }
  return best;
}

// Complex x complex digit GEMMs from three real products: large ones by
// default; TDCC_M3=0 keeps the four-product GEMM on interleaved re/im rows,
// 2 takes every complex x complex single-split digit GEMM (a test mode).
static int m3_mode() {
  const char* e = std::getenv("TDCC_M3");
  return e ? std::atoi(e) : 1;
}

// Digits a digit-GEMM launch multiplies with (at most the split's S):
// TDCC_DIGITS_LOW lists launches (einsum specs, "ladder", "all") that use
// TDCC_DIGITS_LOW_S digits (default S - 1) of their operands' S-digit planes.
static int launch_digits(const std::string& spec, bool ladder, int S) {
  const char* e = std::getenv("TDCC_DIGITS_LOW");
  if (!e || !*e || S <= 2) return S;
  bool hit = false;
  std::string cur;
  for (const char* c = e;; ++c) {
    if (*c == ';' || *c == 0) {
      hit = hit || cur == "all" || (ladder ? cur == "ladder" : cur == spec);
      cur.clear();
      if (*c == 0) break;
    } else {
      cur.push_back(*c);
    }
  }
  if (!hit) return S;
  const char* es = std::getenv("TDCC_DIGITS_LOW_S");
  const int s = es ? std::atoi(es) : S - 1;
  return std::max(2, std::min(S, s));
}

int Graph::contract(const std::string& spec, int a, int b, double ar, double ai) {
  std::vector<std::string> ins;
  std::string lc;
  parse_spec(spec, ins, lc);
  if (ins.size() != 2) throw Error("contract needs two operands: " + spec);
  std::string la = ins[0], lb = ins[1];
  const Node& A = nodes_[a];
  const Node& B = nodes_[b];
  if ((int)la.size() != A.shape.rank || (int)lb.size() != B.shape.rank)
    throw Error("rank mismatch in " + spec);
  std::map<char, Cls> lcls;
  auto note = [&](const std::string& ls, const Shape& s) {
    for (int d = 0; d < (int)ls.size(); ++d) {
      if (std::count(ls.begin(), ls.end(), ls[d]) != 1) throw Error("repeated index in operand: " + spec);
      auto it = lcls.find(ls[d]);
      if (it != lcls.end() && extent(it->second) != extent(s.cls[d]))
        throw Error("extent mismatch for index in " + spec);
      lcls[ls[d]] = s.cls[d];
    }
  };
  note(la, A.shape);
  note(lb, B.shape);
  // Indices that occur in one operand only and not in the output are summed
  // inside that operand first (contraction with a ones tensor).
  auto pre_reduce = [&](int x, const std::string& lx, const std::string& ly) -> std::pair<int, std::string> {
    std::string keep, drop;
    for (char c : lx) (has(ly, c) || has(lc, c) ? keep : drop).push_back(c);
    if (drop.empty()) return {x, lx};
    Shape os;
    os.rank = (int)drop.size();
    for (int d = 0; d < os.rank; ++d) os.cls[d] = lcls[drop[d]];
    const int ones = ones_view(os);
    return {contract(lx + "," + drop + "->" + keep, x, ones), keep};
  };
  auto ra = pre_reduce(a, la, lb);
  auto rb = pre_reduce(b, lb, la);
  a = ra.first;
  la = ra.second;
  b = rb.first;
  lb = rb.second;
  const std::string nspec = la + "," + lb + "->" + lc;

  Node n;
  n.op = Op::Contract;
  n.shape.rank = (int)lc.size();
  for (int d = 0; d < n.shape.rank; ++d) {
    if (!lcls.count(lc[d])) throw Error("output index not in inputs: " + spec);
    n.shape.cls[d] = lcls[lc[d]];
  }
  n.cplx = nodes_[a].cplx || nodes_[b].cplx || ai != 0.0;
  n.spec = nspec;
  n.a = a;
  n.b = b;
  n.ar = ar;
  n.ai = ai;
  n.xsym = a == b && la.size() == 2 && lb.size() == 2 && lc.size() == 4 && lc[0] == la[0] && lc[1] == lb[0] &&
           lc[2] == la[1] && lc[3] == lb[1] && ai == 0.0;
  // output axis pairs taken from one antisymmetric pair of an operand
  for (int d1 = 0; d1 < n.shape.rank; ++d1)
    for (int d2 = d1 + 1; d2 < n.shape.rank; ++d2)
      for (int side = 0; side < 2; ++side) {
        const std::string& lx = side ? lb : la;
        const size_t x1 = lx.find(lc[d1]), x2 = lx.find(lc[d2]);
        if (x1 != std::string::npos && x2 != std::string::npos &&
            (nodes_[side ? b : a].anti & anti_bit((int)x1, (int)x2)))
          n.anti |= anti_bit(d1, d2);
      }
  n.key = "C|" + nspec + "|" + std::to_string(a) + "|" + std::to_string(b) + "|" + hexf(ar) + "|" + hexf(ai);
  return add(std::move(n));
}

int Graph::ones_view(const Shape& s) {
  if (ones_id_ < 0) {
    TDCC_HIP(hipMalloc(&ones_buf_, sizeof(double)));
    const double one = 1.0;
    TDCC_HIP(hipMemcpy(ones_buf_, &one, sizeof(double), hipMemcpyHostToDevice));
    Shape s0;
    Dev d;
    d.re = ones_buf_;
    ones_id_ = external(s0, false, d, "ones");
  }
  const int64_t st[4] = {0, 0, 0, 0};
  return view(ones_id_, s, 0, st);
}

int Graph::contract3(const std::string& spec, int a, int b, int c, double ar) {
  std::vector<std::string> ins;
  std::string lo;
  parse_spec(spec, ins, lo);
  if (ins.size() != 3) throw Error("contract3 needs three operands: " + spec);
  const int ids[3] = {a, b, c};
  std::map<char, int64_t> ext;
  for (int t = 0; t < 3; ++t)
    for (int d = 0; d < (int)ins[t].size(); ++d) ext[ins[t][d]] = extent(nodes_[ids[t]].shape.cls[d]);
  double best = -1;
  int bx = 0, by = 1, bz = 2;
  std::string bmid;
  const int pairs[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
  for (auto& p : pairs) {
    const std::string &x = ins[p[0]], &y = ins[p[1]], &z = ins[p[2]];
    std::string uni = x;
    for (char ch : y)
      if (!has(uni, ch)) uni.push_back(ch);
    std::string mid;
    for (char ch : uni)
      if (has(z, ch) || has(lo, ch)) mid.push_back(ch);
    double c1 = 1, c2 = 1;
    for (char ch : uni) c1 *= (double)ext[ch];
    std::string uni2 = mid;
    for (char ch : z)
      if (!has(uni2, ch)) uni2.push_back(ch);
    for (char ch : uni2) c2 *= (double)ext[ch];
    if (best < 0 || c1 + c2 < best) {
      best = c1 + c2;
      bx = p[0];
      by = p[1];
      bz = p[2];
      bmid = mid;
    }
  }
  const int t = contract(ins[bx] + "," + ins[by] + "->" + bmid, ids[bx], ids[by]);
  return contract(bmid + "," + ins[bz] + "->" + lo, t, ids[bz], ar);
}

int Graph::lincomb(std::vector<Term> terms) {
  if (terms.empty()) throw Error("empty linear combination");
  // Inline sources that are themselves linear combinations while the total
  // stays within one kernel.
  std::vector<Term> flat;
  for (const Term& t : terms) {
    const Node& s = nodes_[t.src];
    if (s.op == Op::Lincomb && !s.keep && flat.size() + s.terms.size() + (terms.size() - 1) <= (size_t)kInlineTerms) {
      for (const Term& u : s.terms) {
        Term c;
        c.src = u.src;
        for (int d = 0; d < 4; ++d) c.perm[d] = u.perm[t.perm[d]];
        c.cr = t.cr * u.cr - t.ci * u.ci;
        c.ci = t.cr * u.ci + t.ci * u.cr;
        flat.push_back(c);
      }
    } else {
      flat.push_back(t);
    }
  }
  Node n;
  n.op = Op::Lincomb;
  const Node& s0 = nodes_[flat[0].src];
  n.shape.rank = s0.shape.rank;
  for (int d = 0; d < n.shape.rank; ++d) n.shape.cls[d] = s0.shape.cls[flat[0].perm[d]];
  std::ostringstream k;
  k << "L";
  for (const Term& t : flat) {
    const Node& s = nodes_[t.src];
    if (s.shape.rank != n.shape.rank) throw Error("rank mismatch in linear combination");
    for (int d = 0; d < n.shape.rank; ++d)
      if (extent(s.shape.cls[t.perm[d]]) != extent(n.shape.cls[d])) throw Error("shape mismatch in linear combination");
    n.cplx = n.cplx || s.cplx || t.ci != 0.0;
    k << "|" << t.src << ":" << t.perm[0] << t.perm[1] << t.perm[2] << t.perm[3] << ":" << hexf(t.cr) << ":" << hexf(t.ci);
  }
  // an axis pair is antisymmetric when every term either reads an
  // antisymmetric pair of its source there or is cancelled in sign by a
  // partner term reading the same source with the two axes exchanged
  for (int d1 = 0; d1 < n.shape.rank; ++d1)
    for (int d2 = d1 + 1; d2 < n.shape.rank; ++d2) {
      bool ok = true;
      for (size_t t = 0; t < flat.size() && ok; ++t) {
        const Term& T = flat[t];
        if (nodes_[T.src].anti & anti_bit(T.perm[d1], T.perm[d2])) continue;
        // the reading of T with output axes d1, d2 exchanged (and, for a
        // source unchanged under exchanging its axes 0, 1 and 2, 3 together,
        // the same reading with that exchange applied)
        int want[4], alt[4];
        for (int d = 0; d < n.shape.rank; ++d) {
          want[d] = T.perm[d == d1 ? d2 : (d == d2 ? d1 : d)];
          alt[d] = want[d] ^ 1;
        }
        const bool xs = nodes_[T.src].xsym;
        bool partner = false;
        for (size_t u = 0; u < flat.size() && !partner; ++u) {
          const Term& U = flat[u];
          if (u == t || U.src != T.src || U.cr != -T.cr || U.ci != -T.ci) continue;
          bool same = true, same_alt = xs;
          for (int d = 0; d < n.shape.rank; ++d) {
            same = same && U.perm[d] == want[d];
            same_alt = same_alt && U.perm[d] == alt[d];
          }
          partner = same || same_alt;
        }
        ok = partner;
      }
      if (ok) n.anti |= anti_bit(d1, d2);
    }
  n.terms = std::move(flat);
  n.key = k.str();
  return add(std::move(n));
}

int Graph::materialize(std::vector<Term> terms) {
  const int x = lincomb(std::move(terms));
  if (nodes_[x].op != Op::Lincomb) return x;
  if (nodes_[x].keep) return x;
  Node n = nodes_[x];  // a node of its own (the plain one may be folded elsewhere)
  n.keep = true;
  n.key += "|keep";
  return add(std::move(n));
}

int Graph::field(int stage) {
  Node n;
  n.op = Op::Field;
  n.shape.rank = 2;
  n.shape.cls[0] = n.shape.cls[1] = Cls::N;
  n.cplx = true;
  n.stage = stage;
  n.traj = traj_;
  n.key = "F|" + std::to_string(stage) + "|" + std::to_string(traj_);
  return add(std::move(n));
}

int Graph::ladder(int u, int ket) {
  const Node& U = nodes_[u];
  if (U.shape.rank != 4 || U.shape.cls[0] != Cls::O || U.shape.cls[1] != Cls::O || U.shape.cls[2] != Cls::V ||
      U.shape.cls[3] != Cls::V)
    throw Error("ladder input must be (o,o,v,v)");
  Node n;
  n.op = Op::Ladder;
  n.shape = U.shape;
  n.cplx = true;
  n.a = u;
  n.stage = ket;
  n.ij_pairs = ij_pairs_;
  n.anti = (uint8_t)(anti_bit(2, 3) | (U.anti & anti_bit(0, 1)));
  n.key = "LAD|" + std::to_string(u) + "|" + std::to_string(ket) + "|" + (ij_pairs_ ? "p" : "f");
  return add(std::move(n));
}

void Graph::finalize_rk(int y_ext, int base_ext, const int k[4], bool lambda) {
  Node n;
  n.op = Op::Finalize;
  n.shape = nodes_[y_ext].shape;
  n.a = y_ext;
  n.b = base_ext;
  for (int s = 0; s < 4; ++s) n.k[s] = k[s];
  n.lambda_sign = lambda ? 1 : 0;
  add(std::move(n));
}

// ---------------------------------------------------------------------------
// Reverse mode.
Graph::VjpResult Graph::vjp(const std::vector<std::pair<int, Term>>& seeds) {
  const int top = (int)nodes_.size();
  // nodes that depend on a non-constant input (state externals or fields)
  std::vector<char> var(top, 0);
  for (int i = 0; i < top; ++i) {
    const Node& n = nodes_[i];
    switch (n.op) {
      case Op::External: var[i] = !is_static(i); break;
      case Op::View: var[i] = var[n.base]; break;
      case Op::Field: var[i] = 1; break;
      case Op::Contract: var[i] = var[n.a] || var[n.b]; break;
      case Op::Lincomb:
        for (const Term& t : n.terms) var[i] = var[i] || var[t.src];
        break;
      case Op::Ladder: var[i] = var[n.a]; break;
      default: var[i] = 0; break;
    }
  }
  std::map<int, std::vector<Term>> acc;
  for (const auto& s : seeds) {
    if (s.first < 0 || s.first >= top) throw Error("vjp seed out of range");
    if (var[s.first]) acc[s.first].push_back(s.second);
  }
  auto letters_of = [](const std::string& spec, std::string& la, std::string& lb, std::string& lc) {
    const auto comma = spec.find(','), arrow = spec.find("->");
    la = spec.substr(0, comma);
    lb = spec.substr(comma + 1, arrow - comma - 1);
    lc = spec.substr(arrow + 2);
  };
  VjpResult res;
  for (int id = top - 1; id >= 0; --id) {
    auto it = acc.find(id);
    if (it == acc.end()) continue;
    std::vector<Term> terms = std::move(it->second);
    acc.erase(it);
    const Node n = nodes_[id];  // copy: adding nodes may move the vector
    const bool ident = terms.size() == 1 && terms[0].cr == 1.0 && terms[0].ci == 0.0 &&
                       terms[0].perm[0] == 0 && terms[0].perm[1] == 1 && terms[0].perm[2] == 2 && terms[0].perm[3] == 3;
    const int bar = ident ? terms[0].src : lincomb(terms);
    switch (n.op) {
      case Op::External:
        res.ext[id] = bar;
        break;
      case Op::View:
        if (nodes_[n.base].op == Op::Field) res.field_views.push_back({id, bar});
        else throw Error("vjp: view of a non-constant, non-field tensor");
        break;
      case Op::Lincomb:
        for (const Term& t : n.terms) {
          if (!var[t.src]) continue;
          Term b{bar, {0, 1, 2, 3}, t.cr, t.ci};
          for (int d = 0; d < n.shape.rank; ++d) b.perm[t.perm[d]] = d;
          acc[t.src].push_back(b);
        }
        break;
      case Op::Contract: {
        if (!n.kpair.empty()) throw Error("vjp: pair-restricted contraction");
        std::string la, lb, lc;
        letters_of(n.spec, la, lb, lc);
        if (var[n.a]) acc[n.a].push_back(tm(contract(lc + "," + lb + "->" + la, bar, n.b, n.ar, n.ai)));
        if (var[n.b]) acc[n.b].push_back(tm(contract(la + "," + lc + "->" + lb, n.a, bar, n.ar, n.ai)));
        break;
      }
      case Op::Ladder:
        // out = sum_ef <ab||ef> u_ijef (ket 0) has the transpose
        // ubar_ijef = sum_ab <ab||ef> outbar_ijab, the ket-1 ladder (and back)
        {
          // the adjoint need not be antisymmetric in (i, j): full rows
          const bool keep = ij_pairs_;
          ij_pairs_ = false;
          acc[n.a].push_back(tm(ladder(bar, 1 - n.stage)));
          ij_pairs_ = keep;
        }
        break;
      case Op::Field:
        throw Error("vjp: a stage field used other than through its blocks");
      default:
        throw Error("vjp: node type without an adjoint");
    }
  }
  return res;
}

void Graph::field_vjp(int adj, int view) {
  const Node& v = nodes_[view];
  if (v.op != Op::View || nodes_[v.base].op != Op::Field || v.shape.rank != 2)
    throw Error("field_vjp needs a block view of a stage field");
  Node n;
  n.op = Op::FieldVjp;
  n.a = adj;
  n.b = view;
  n.stage = nodes_[v.base].stage;
  n.traj = nodes_[v.base].traj;
  add(std::move(n));
}

int Graph::seed(Cls rows, Cls cols) {
  Node n;
  n.op = Op::Seed;
  n.shape.rank = 2;
  n.shape.cls[0] = rows;
  n.shape.cls[1] = cols;
  n.cplx = true;
  return add(std::move(n));
}

void Graph::trace_vjp(int dij, int l1, int dai, int dab) {
  Node n;
  n.op = Op::TraceVjp;
  n.k[0] = dij;
  n.k[1] = l1;
  n.k[2] = dai;
  n.k[3] = dab;
  add(std::move(n));
}

void Graph::trace(int dij, int l1, int dai, int dab) {
  Node n;
  n.op = Op::Trace;
  n.traj = traj_;
  n.k[0] = dij;
  n.k[1] = l1;
  n.k[2] = dai;
  n.k[3] = dab;
  add(std::move(n));
}

void Graph::assign(int dst, int src) {
  Node n;
  n.op = Op::Assign;
  n.shape = nodes_[dst].shape;
  n.a = src;
  n.b = dst;
  add(std::move(n));
}

// ---------------------------------------------------------------------------
static bool owns_buffer(Op op) {
  return op == Op::Contract || op == Op::Lincomb || op == Op::Field || op == Op::Ladder || op == Op::Seed;
}

static void inputs_of(const Node& n, std::vector<int>& in) {
  in.clear();
  switch (n.op) {
    case Op::Contract: in = {n.a, n.b}; break;
    case Op::Lincomb:
      for (const Term& t : n.terms) {
        in.push_back(t.src);
        if (t.src2 >= 0) in.push_back(t.src2);
      }
      break;
    case Op::View: in = {n.base}; break;
    case Op::Finalize: in = {n.k[0], n.k[1], n.k[2], n.k[3], n.a, n.b}; break;
    case Op::Trace: in = {n.k[0], n.k[1], n.k[2], n.k[3]}; break;
    case Op::Assign: in = {n.a, n.b}; break;
    case Op::Ladder: in = {n.a}; break;
    case Op::FieldVjp: in = {n.a}; break;
    case Op::TraceVjp: in = {n.k[0], n.k[1], n.k[2], n.k[3]}; break;
    default: break;
  }
}

// Ladder scratch: packed input up (R2 = 2 rows x P doubles; rows = o^2, or
// o (o - 1) / 2 with ij pairs, times the trajectories of a stacked program)
// and the result rp (one R2 x P partial per K chunk of the digit GEMM), then,
// in split precision, the digit planes of up, its row scales and the split
// workspace.
static int64_t al256(int64_t x) { return (x + 255) / 256 * 256; }
static int64_t ladder_rows(int o, bool pr) { return pr ? (int64_t)o * (o - 1) / 2 : (int64_t)o * o; }
static int64_t ladder_rows2(int o, bool pr) { return 2 * ladder_rows(o, pr); }
static int64_t ladder_P(int v) { return (int64_t)v * (v - 1) / 2; }
static int64_t ladder_Kp(int v) { return (ladder_P(v) + 63) / 64 * 64; }
static int64_t ladder_plane(int64_t R2, int v) { return al256(R2 * ladder_Kp(v)); }
// K chunk of the ladder GEMM (0: one chunk): a grid of fewer than two rounds
// of 128 x 128 tiles over the 256 CUs is cut along K so the last round is
// not mostly idle.  Digit GEMM: multiples of 64 within the exact int32
// range; FP64 GEMM: multiples of 16 of K = P.
static int64_t ladder_k_chunk(int64_t R2, int v, int S) {
  const int64_t gran = S > 0 ? 64 : 16;
  const int64_t K = S > 0 ? ladder_Kp(v) : ladder_P(v), steps = K / gran;
  const int64_t tiles = ((R2 + 127) / 128) * ((ladder_P(v) + 127) / 128);
  int64_t sp = tiles >= 512 ? 1 : std::min<int64_t>(8, (1280 + tiles - 1) / tiles);
  sp = std::max<int64_t>(1, std::min<int64_t>(sp, steps / 2));
  int64_t chunk = (steps + sp - 1) / sp * gran;
  if (S > 0) chunk = std::min<int64_t>(chunk, ozgemm_max_k(S));
  return chunk >= K ? 0 : chunk;
}
static int64_t ladder_splits(int64_t R2, int v, int S) {
  const int64_t c = ladder_k_chunk(R2, v, S);
  return c > 0 ? ((S > 0 ? ladder_Kp(v) : ladder_P(v)) + c - 1) / c : 1;
}
static int64_t ladder_digits_off(int64_t R2, int v, int S) {
  return al256((1 + ladder_splits(R2, v, S)) * R2 * ladder_P(v) * (int64_t)sizeof(double));
}
static int64_t ladder_scale_off(int64_t R2, int v, int S) {
  return ladder_digits_off(R2, v, S) + S * ladder_plane(R2, v);
}
static int64_t ladder_pmax_off(int64_t R2, int v, int S) {
  return ladder_scale_off(R2, v, S) + al256(R2 * (int64_t)sizeof(double));
}
static int64_t ladder_scratch_bytes(int64_t R2, int v, int S) {
  if (S <= 0) return ladder_digits_off(R2, v, S);
  return ladder_pmax_off(R2, v, S) + al256(ozaki_split_workspace(R2, ladder_Kp(v), false));
}

// A split's offset tables with a constant step (rows r * ldx + c, or k + c)
// are replaced by that step, their constant folded into the source planes, so
// the split kernels read rows without table lookups (and contiguous rows with
// 16-byte loads).
void Graph::simplify_split_tables(SplitArgs& a, int64_t rtab_off, int64_t ktab_off) const {
  const int64_t* rt = host_tables_.data() + rtab_off;
  const int64_t* kt = host_tables_.data() + ktab_off;
  const int64_t R = a.Rb, K = a.K;
  if (R < 1 || K < 1) return;
  bool kc = true;
  for (int64_t k = 0; kc && k < K; ++k) kc = kt[k] == kt[0] + k;
  const int64_t step = R > 1 ? rt[1] - rt[0] : 0;
  bool ra = true;
  for (int64_t r = 0; ra && r < R; ++r) ra = rt[r] == rt[0] + r * step;
  int64_t shift = 0;
  if (kc) {
    shift += kt[0];
    a.ktab = nullptr;
  }
  if (ra) {
    shift += rt[0];
    a.rtab = nullptr;
    a.ldx = step;
  }
  if (shift != 0) {
    a.src_re += shift;
    if (a.src_im) a.src_im += shift;
  }
}

bool Graph::is_static(int id) const {
  const Node& n = nodes_[id];
  if (n.op == Op::View) return is_static(n.base);
  return n.op == Op::External && (n.key.rfind("X|tei|", 0) == 0 || n.key.rfind("X|ones|", 0) == 0);
}

static bool two_streams_on() {
  const char* e = std::getenv("TDCC_STREAMS");
  return !e || std::atoi(e) != 1;
}

static bool is_post_op(Op op) {
  return op == Op::Finalize || op == Op::Trace || op == Op::Assign || op == Op::FieldVjp || op == Op::TraceVjp;
}

// Sides of a two-stream step (see set_lambda_externals): a node depending on
// a Lambda external is on the Lambda side; a node that does not is on the
// Lambda side when the Lambda side reads it and the T side does not (the
// finalize and what follows it do not count), else on the T side.  A T side
// node the Lambda side reads is a crossing.
static bool cross_first_on() {
  const char* e = std::getenv("TDCC_CROSS_FIRST");
  return !e || std::atoi(e) != 0;
}

void Graph::compute_sides() {
  side_.clear();
  cross_.clear();
  order_.resize(nodes_.size());
  pos_.resize(nodes_.size());
  for (int i = 0; i < (int)nodes_.size(); ++i) order_[i] = pos_[i] = i;
  bool have_fin = false;
  for (int i = 0; i < (int)nodes_.size(); ++i) have_fin = have_fin || (live_[i] && nodes_[i].op == Op::Finalize);
  if (lam_ext_.empty() || !have_fin || !two_streams_on()) return;
  const int N = (int)nodes_.size();
  auto root = [&](int x) {
    while (nodes_[x].op == Op::View) x = nodes_[x].base;
    return x;
  };
  std::vector<char> lam(N, 0);
  for (int x : lam_ext_) lam[x] = 1;
  std::vector<int> in;
  for (int i = 0; i < N; ++i) {
    const Node& n = nodes_[i];
    if (n.op == Op::External) continue;
    inputs_of(n, in);
    for (int x : in) lam[i] = lam[i] || lam[x];
  }
  side_.assign(N, 0);
  cross_.assign(N, 0);
  std::vector<char> has_user(N, 0), all_lam(N, 1);
  for (int i = N - 1; i >= 0; --i) {
    const Node& n = nodes_[i];
    if (!live_[i] || n.op == Op::View || n.op == Op::External) continue;
    if (is_post_op(n.op)) {
      side_[i] = 2;
      continue;
    }
    side_[i] = lam[i] ? 1 : ((has_user[i] && all_lam[i]) ? 1 : 0);
    inputs_of(n, in);
    for (int x : in) {
      const int r = root(x);
      has_user[r] = 1;
      if (side_[i] != 1) all_lam[r] = 0;
    }
  }
  // nodes after the first finalize run on the first stream after the join
  int first_fin = N;
  for (int i = 0; i < N; ++i)
    if (live_[i] && side_[i] == 2) {
      first_fin = i;
      break;
    }
  for (int i = first_fin; i < N; ++i)
    if (side_[i] == 1) side_[i] = 0;
  for (int i = 0; i < N; ++i) {
    if (!live_[i] || side_[i] != 1) continue;
    inputs_of(nodes_[i], in);
    for (int x : in) {
      const int r = root(x);
      if (nodes_[r].op != Op::External && side_[r] == 0) cross_[r] = 1;
    }
  }
  // The T side computes what the Lambda side reads from it (the crossing
  // nodes and everything they are computed from) before its own remaining
  // work, so that the Lambda side is not held up by T-side work it does not
  // need (e.g. a field-dependent intermediate of the step-start t first built
  // for a late T stage); TDCC_CROSS_FIRST=0 keeps node order.
  if (!cross_first_on()) return;
  std::vector<char> cl(N, 0);
  for (int i = N - 1; i >= 0; --i) {
    if (!live_[i]) continue;
    if (cross_[i]) cl[i] = 1;
    if (!cl[i]) continue;
    inputs_of(nodes_[i], in);
    for (int x : in) cl[x] = 1;
  }
  order_.clear();
  for (int pass = 0; pass < 3; ++pass)
    for (int i = 0; i < N; ++i) {
      const int cls = i >= first_fin ? 2 : (cl[i] ? 0 : 1);
      if (cls == pass) order_.push_back(i);
    }
  for (int p = 0; p < N; ++p) pos_[order_[p]] = p;
}

void Graph::plan() {
  const int N = (int)nodes_.size();
  cache_.clear();
  cache_index_.clear();
  cache_bytes_ = 0;
  // dead-node removal: keep what side-effect nodes reach
  std::vector<char> live(N, 0);
  std::vector<int> in;
  for (int i = N - 1; i >= 0; --i) {
    const Op op = nodes_[i].op;
    if (op == Op::Finalize || op == Op::Trace || op == Op::Assign || op == Op::FieldVjp || op == Op::TraceVjp)
      live[i] = 1;
    if (!live[i]) continue;
    inputs_of(nodes_[i], in);
    for (int x : in) live[x] = 1;
  }
  fold_outer_products(live);
  inline_lincombs(live);
  live_ = live;
  mark_stack();
  compute_sides();
  // liveness: uses of views count as uses of their base
  auto root = [&](int x) {
    while (nodes_[x].op == Op::View) x = nodes_[x].base;
    return x;
  };
  // (positions in the execution order)
  for (auto& n : nodes_) n.last_use = -1;
  for (int p = 0; p < N; ++p) {
    const int i = order_[p];
    if (!live[i] || nodes_[i].op == Op::View) continue;
    inputs_of(nodes_[i], in);
    for (int x : in) {
      const int r = root(x);
      nodes_[r].last_use = std::max(nodes_[r].last_use, p);
    }
  }
  // arena layout: first fit over a coalescing free list per region (one
  // region in a one-stream program; in a two-stream step one per stream, each
  // planned over that stream's launch order, and one never reused within the
  // step for the T side buffers the Lambda side reads)
  std::map<int64_t, int64_t> free_lists[3];  // offset -> bytes
  int64_t tops[3] = {0, 0, 0};
  int cur_region = 0;
  auto alloc = [&](int64_t bytes) -> int64_t {
    std::map<int64_t, int64_t>& free_blocks = free_lists[cur_region];
    bytes = (bytes + 255) & ~int64_t(255);
    for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
      if (it->second >= bytes) {
        const int64_t off = it->first;
        const int64_t rest = it->second - bytes;
        free_blocks.erase(it);
        if (rest > 0) free_blocks[off + bytes] = rest;
        return off;
      }
    }
    const int64_t off = tops[cur_region];
    tops[cur_region] += bytes;
    return off;
  };
  auto release = [&](int64_t off, int64_t bytes) {
    if (cur_region == 2) return;  // crossing buffers live through the step
    std::map<int64_t, int64_t>& free_blocks = free_lists[cur_region];
    bytes = (bytes + 255) & ~int64_t(255);
    auto it = free_blocks.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_blocks.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_blocks.erase(nx);
    }
    if (it != free_blocks.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_blocks.erase(it);
      }
    }
  };
  std::vector<std::vector<int>> ends(N);  // by position
  for (int i = 0; i < N; ++i)
    if (live[i] && owns_buffer(nodes_[i].op)) ends[std::max(nodes_[i].last_use, pos_[i])].push_back(i);
  host_tables_.clear();
  launches_.clear();
  dyn_uses_.clear();
  shares_.clear();
  for (int p = 0; p < N; ++p) {
    const int i = order_[p];
    if (!live[i]) continue;
    Node& n = nodes_[i];
    if (n.op == Op::Ladder)
      n.scratch_len = ladder_scratch_bytes(ladder_rows2(o_, n.ij_pairs) * (n.zdep ? nz_ : 1), v_, rt_->split_digits);
    if (n.op == Op::Contract) lower_contract(i);
  }
  plan_shares();
  std::vector<std::vector<int>> share_at(N), share_end(N);
  for (int k = 0; k < (int)shares_.size(); ++k) {
    share_at[shares_[k].first].push_back(k);
    share_end[shares_[k].last].push_back(k);
  }
  const bool two = !side_.empty();
  auto side_region = [&](int i) { return two && side_[i] == 1 ? 1 : 0; };
  auto buf_region = [&](int i) { return two && cross_[i] ? 2 : side_region(i); };
  std::vector<int> node_region(N, 0), share_region(shares_.size(), 0);
  for (int p = 0; p < N; ++p) {
    const int i = order_[p];
    if (!live[i]) continue;
    Node& n = nodes_[i];
    if (owns_buffer(n.op)) {
      cur_region = node_region[i] = buf_region(i);
      n.arena_off = alloc(zbytes(n) * (n.zdep ? nz_ : 1));
    }
    cur_region = side_region(i);
    for (int k : share_at[i]) {
      share_region[k] = cur_region;
      shares_[k].off = alloc(shares_[k].bytes);
    }
    if ((n.op == Op::Contract || n.op == Op::Ladder) && n.scratch_len > 0) {
      n.scratch_off = alloc(n.scratch_len);
      release(n.scratch_off, n.scratch_len);
      node_region[i] |= cur_region << 4;  // scratch region
    }
    for (int k : share_end[i]) {
      cur_region = share_region[k];
      release(shares_[k].off, shares_[k].bytes);
    }
    for (int x : ends[p]) {
      cur_region = node_region[x] & 15;
      release(nodes_[x].arena_off, zbytes(nodes_[x]) * (nodes_[x].zdep ? nz_ : 1));
    }
  }
  // regions laid out one after another
  const int64_t base_of[3] = {0, tops[0], tops[0] + tops[1]};
  for (int i = 0; i < N; ++i) {
    if (!live[i]) continue;
    Node& n = nodes_[i];
    if (owns_buffer(n.op)) n.arena_off += base_of[node_region[i] & 15];
    if ((n.op == Op::Contract || n.op == Op::Ladder) && n.scratch_len > 0) n.scratch_off += base_of[node_region[i] >> 4];
  }
  for (size_t k = 0; k < shares_.size(); ++k) shares_[k].off += base_of[share_region[k]];
  arena_bytes_ = tops[0] + tops[1] + tops[2];
  if (const char* d = std::getenv("TDCC_DUMP_NODES"); d && d[0] == '1') dump_nodes();
  if (dev_tables_) {
    (void)hipFree(dev_tables_);
    dev_tables_ = nullptr;
  }
  if (!host_tables_.empty()) {
    TDCC_HIP(hipMalloc(&dev_tables_, host_tables_.size() * sizeof(int64_t)));
    TDCC_HIP(hipMemcpy(dev_tables_, host_tables_.data(), host_tables_.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  }
  if (cache_buf_) {
    (void)hipFree(cache_buf_);
    cache_buf_ = nullptr;
  }
  if (cache_bytes_ > 0) TDCC_HIP(hipMalloc(&cache_buf_, cache_bytes_));
}

// A term of a linear combination that reads an outer product of two rank-2
// tensors (a contraction 'ia,jb->ijab' with nothing summed, such as t1 t1), or
// a rank-4 tensor scaled element-wise by a small one ('ijab,b->ijab'), is
// replaced by a product term reading the two factors, so the product is
// computed where it is added instead of being written and read back; the
// contraction is dropped when nothing else reads it.  TDCC_FOLD_OUTER=0
// disables it.
void Graph::fold_outer_products(std::vector<char>& live) {
  if (const char* e = std::getenv("TDCC_FOLD_OUTER"); e && std::atoi(e) == 0) return;
  const int N = (int)nodes_.size();
  std::vector<int> users(N, 0), in;
  auto unique_inputs = [&](const Node& n, std::vector<int>& v) {
    inputs_of(n, v);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  for (int i = 0; i < N; ++i) {
    if (!live[i]) continue;
    unique_inputs(nodes_[i], in);
    for (int x : in) ++users[x];
  }
  // a rank-4 contraction with no summed letter: the outer product of two
  // matrices, or a rank-4 tensor scaled element-wise by a vector or matrix
  // over some of its letters (e.g. x_ijab f_b after the sum over e of
  // x_ijab F_eb), read as x * y at every element
  auto outer = [&](int c, std::string& la, std::string& lb, std::string& lc) {
    const Node& C = nodes_[c];
    if (C.op != Op::Contract || C.shape.rank != 4 || !C.kpair.empty()) return false;
    std::vector<std::string> ins;
    parse_spec(C.spec, ins, lc);
    if (ins.size() != 2) return false;
    la = ins[0];
    lb = ins[1];
    if (lc.size() != 4) return false;
    bool shared = false;
    for (char x : la) shared = shared || lb.find(x) != std::string::npos;
    for (const std::string* l : {&la, &lb})
      for (char x : *l)
        if (lc.find(x) == std::string::npos) return false;
    if (!shared) return la.size() == 2 && lb.size() == 2;
    return (la.size() == 4 && lb.size() <= 2) || (lb.size() == 4 && la.size() <= 2);
  };
  std::vector<int> before, after;
  for (int i = 0; i < N; ++i) {
    if (!live[i] || nodes_[i].op != Op::Lincomb || nodes_[i].shape.rank != 4) continue;
    Node& L = nodes_[i];
    unique_inputs(L, before);
    bool any = false;
    for (Term& T : L.terms) {
      std::string la, lb, lc;
      if (T.src2 >= 0 || !outer(T.src, la, lb, lc)) continue;
      const Node& C = nodes_[T.src];
      Term U;
      U.src = C.a;
      U.src2 = C.b;
      for (int d = 0; d < 4; ++d) {
        const char ch = lc[T.perm[d]];
        const size_t pa = la.find(ch), pb = lb.find(ch);
        U.perm[d] = pa == std::string::npos ? -1 : (int)pa;
        U.perm2[d] = pb == std::string::npos ? -1 : (int)pb;
      }
      U.cr = T.cr * C.ar - T.ci * C.ai;
      U.ci = T.cr * C.ai + T.ci * C.ar;
      T = U;
      any = true;
    }
    if (!any) continue;
    unique_inputs(L, after);
    for (int x : before)
      if (!std::binary_search(after.begin(), after.end(), x)) --users[x];
    for (int x : after)
      if (!std::binary_search(before.begin(), before.end(), x)) ++users[x];
  }
  for (int c = 0; c < N; ++c) {
    std::string la, lb, lc;
    if (live[c] && users[c] == 0 && outer(c, la, lb, lc)) live[c] = 0;
  }
}

// A linear combination S whose only consumer is another linear combination L
// is folded into L (its terms composed into L's, equal source readings merged)
// while L stays within one launch, so S is neither written nor read back.
// For an antisymmetric L (summed at its canonical elements, where the terms
// reading one source through the antisymmetrizer's permutations cover it once)
// every reading of S is folded; otherwise only an S that L reads once.
// TDCC_INLINE_LINCOMB=0 disables it.
void Graph::inline_lincombs(std::vector<char>& live) {
  if (const char* e = std::getenv("TDCC_INLINE_LINCOMB"); e && std::atoi(e) == 0) return;
  const int N = (int)nodes_.size();
  std::vector<int> users(N, 0), user_of(N, -1);
  std::vector<int> in;
  for (int i = 0; i < N; ++i) {
    if (!live[i]) continue;
    inputs_of(nodes_[i], in);
    std::sort(in.begin(), in.end());
    in.erase(std::unique(in.begin(), in.end()), in.end());
    for (int x : in) {
      ++users[x];
      user_of[x] = i;
    }
  }
  for (int i = 0; i < N; ++i) {
    if (!live[i] || nodes_[i].op != Op::Lincomb) continue;
    for (bool changed = true; changed;) {
      changed = false;
      Node& L = nodes_[i];
      const bool anti = L.shape.rank == 4 && (L.anti & (anti_bit(0, 1) | anti_bit(2, 3)));
      for (size_t t = 0; t < L.terms.size() && !changed; ++t) {
        const int s = L.terms[t].src;
        const Node& S = nodes_[s];
        if (S.op != Op::Lincomb || S.keep || users[s] != 1 || user_of[s] != i) continue;
        // (a factor of a product term stays a tensor of its own)
        int refs = 0;
        bool factor = false;
        for (const Term& T : L.terms) {
          refs += T.src == s;
          factor = factor || (T.src2 >= 0 && (T.src == s || T.src2 == s));
        }
        if (factor || (!anti && refs > 1)) continue;
        std::vector<Term> nt;
        auto put = [&](const Term& c) {
          for (Term& u : nt)
            if (u.src == c.src && u.src2 == c.src2 && std::equal(u.perm, u.perm + 4, c.perm) &&
                std::equal(u.perm2, u.perm2 + 4, c.perm2)) {
              u.cr += c.cr;
              u.ci += c.ci;
              return;
            }
          nt.push_back(c);
        };
        for (const Term& T : L.terms) {
          if (T.src != s) {
            put(T);
            continue;
          }
          for (const Term& u : S.terms) {
            Term c;
            c.src = u.src;
            c.src2 = u.src2;
            for (int d = 0; d < 4; ++d) {
              c.perm[d] = u.perm[T.perm[d]];
              c.perm2[d] = u.src2 >= 0 ? u.perm2[T.perm[d]] : -1;
            }
            c.cr = T.cr * u.cr - T.ci * u.ci;
            c.ci = T.cr * u.ci + T.ci * u.cr;
            put(c);
          }
        }
        if (nt.size() > (size_t)kMaxTerms) continue;
        // S disappears and L reads its sources: a source that L read already
        // loses one user, any other keeps its count with L as the reader
        std::vector<int> lin, sin;
        inputs_of(L, lin);
        std::sort(lin.begin(), lin.end());
        inputs_of(S, sin);
        std::sort(sin.begin(), sin.end());
        sin.erase(std::unique(sin.begin(), sin.end()), sin.end());
        L.terms = std::move(nt);
        for (int x : sin) {
          if (std::binary_search(lin.begin(), lin.end(), x)) --users[x];
          user_of[x] = i;
        }
        live[s] = 0;
        users[s] = 0;
        changed = true;
      }
    }
  }
}

// Groups the launches that split one computed operand in one layout into
// shares: consecutive uses with no write of the operand's external in
// between (a finalize or assign) share one set of digit planes, cut by the
// first launch of the group.  TDCC_SHARE_DIGITS=0 disables it.
void Graph::plan_shares() {
  if (const char* e = std::getenv("TDCC_SHARE_DIGITS"); e && std::atoi(e) == 0) return;
  std::map<int, std::vector<int>> writes;  // external -> nodes writing it, ascending
  for (int i = 0; i < (int)nodes_.size(); ++i) {
    if (!live_[i]) continue;
    const Node& n = nodes_[i];
    if (n.op == Op::Finalize) {
      writes[n.a].push_back(i);
      writes[n.b].push_back(i);
    } else if (n.op == Op::Assign) {
      writes[n.b].push_back(i);
    }
  }
  std::map<int, Launch*> by_node;
  for (auto& L : launches_) by_node[L->node] = L.get();
  auto root = [&](int x) {
    while (nodes_[x].op == Op::View) x = nodes_[x].base;
    return x;
  };
  auto written_between = [&](int src, int lo, int hi) {  // nodes, compared by position
    auto it = writes.find(root(src));
    if (it == writes.end()) return false;
    for (int w : it->second)
      if (pos_[w] > pos_[lo] && pos_[w] < pos_[hi]) return true;
    return false;
  };
  for (auto& [key, all_uses] : dyn_uses_) {
   // a two-stream step shares digit planes only within one side
   for (int sd = 0; sd < (side_.empty() ? 1 : 2); ++sd) {
    std::vector<DynUse> uses;
    for (const DynUse& u : all_uses)
      if (side_.empty() || (side_[u.node] == 1 ? 1 : 0) == sd) uses.push_back(u);
    std::stable_sort(uses.begin(), uses.end(),
                     [&](const DynUse& x, const DynUse& y) { return pos_[x.node] < pos_[y.node]; });
    size_t g0 = 0;
    while (g0 < uses.size()) {
      size_t g1 = g0 + 1;
      while (g1 < uses.size() && !written_between(uses[g0].src, uses[g1 - 1].node, uses[g1].node)) ++g1;
      if (g1 - g0 >= 2) {
        DigitShare sh;
        sh.bytes = uses[g0].bytes;
        sh.first = uses[g0].node;
        sh.last = uses[g1 - 1].node;
        const int idx = (int)shares_.size();
        shares_.push_back(sh);
        for (size_t u = g0; u < g1; ++u) {
          GemmPath& gp = by_node.at(uses[u].node)->gp;
          (uses[u].side == 0 ? gp.a_share : gp.b_share) = idx;
          (uses[u].side == 0 ? gp.a_split : gp.b_split) = u == g0;
        }
      }
      g0 = g1;
    }
   }
  }
}

// One line per live node on stderr: id, op, shape classes, planes, last use
// and what the node computes (contraction string, lincomb terms with their
// permutations and coefficients).
void Graph::dump_nodes() const {
  static const char* ops[] = {"ext", "view", "contract", "lincomb", "field", "finalize", "trace", "assign",
                              "ladder", "field_vjp", "seed", "trace_vjp"};
  for (int i = 0; i < (int)nodes_.size(); ++i) {
    if (!live_[i]) continue;
    const Node& n = nodes_[i];
    std::string cls;
    for (int k = 0; k < n.shape.rank; ++k) cls += "ovn"[(int)n.shape.cls[k]];
    std::fprintf(stderr, "node %d %s [%s] %s last=%d anti=%d%s", i, ops[(int)n.op], cls.c_str(), n.cplx ? "c" : "r",
                 n.last_use, (int)n.anti, n.xsym ? "x" : "");
    if (n.op == Op::Contract)
      std::fprintf(stderr, " %s a=%d b=%d%s%s", n.spec.c_str(), n.a, n.b, n.kpair.empty() ? "" : " pairs=", n.kpair.c_str());
    if (n.op == Op::View) std::fprintf(stderr, " base=%d", n.base);
    if (n.op == Op::Ladder) std::fprintf(stderr, " u=%d", n.a);
    if (n.op == Op::Lincomb)
      for (const Term& t : n.terms) {
        std::fprintf(stderr, " %d(%d%d%d%d)", t.src, t.perm[0], t.perm[1], t.perm[2], t.perm[3]);
        if (t.src2 >= 0) std::fprintf(stderr, "x%d(%d%d%d%d)", t.src2, t.perm2[0], t.perm2[1], t.perm2[2], t.perm2[3]);
        std::fprintf(stderr, "*(%g,%g)", t.cr, t.ci);
      }
    if (n.op == Op::External) std::fprintf(stderr, " %s", n.key.c_str());
    std::fprintf(stderr, "\n");
  }
}

Dev Graph::resolve(int id, int z) const {
  const Node& n = nodes_[id];
  Dev d;
  const int64_t zo = n.zdep ? (int64_t)z * n.zs : 0;
  switch (n.op) {
    case Op::External:
      d = n.ext;
      if (zo) {
        d.re += zo;
        if (d.im) d.im += zo;
      }
      return d;
    case Op::View: {
      Dev b = resolve(n.base, z);
      d.re = b.re + n.voff;
      d.im = b.im ? b.im + n.voff : nullptr;
      for (int k = 0; k < 4; ++k) d.st[k] = n.vst[k];
      return d;
    }
    default: {
      if (n.arena_off < 0) throw Error("node has no buffer");
      d.re = arena_ + n.arena_off / (int64_t)sizeof(double) + zo;
      d.im = n.cplx ? d.re + size(n.shape) : nullptr;
      contiguous_strides(*this, n.shape, d.st);
      return d;
    }
  }
}

int64_t Graph::zbytes(const Node& n) const {
  return (size(n.shape) * (n.cplx ? 2 : 1) * (int64_t)sizeof(double) + 255) & ~int64_t(255);
}

// Which live nodes carry a trajectory axis: per-trajectory externals (zs set
// by the caller), Field nodes (one field per trajectory) and everything
// computed from them; their z-stride is the external's, the view base's, or
// one trajectory's (256-byte aligned) buffer.
void Graph::mark_stack() {
  std::vector<int> in;
  for (int i = 0; i < (int)nodes_.size(); ++i) {
    Node& n = nodes_[i];
    if (nz_ <= 1) {
      n.zdep = false;
      if (n.op != Op::External) n.zs = 0;
      continue;
    }
    switch (n.op) {
      case Op::External: n.zdep = n.zs > 0; break;
      case Op::View:
        n.zdep = nodes_[n.base].zdep;
        n.zs = nodes_[n.base].zs;
        break;
      case Op::Field: n.zdep = true; break;
      default: {
        inputs_of(n, in);
        n.zdep = false;
        for (int x : in) n.zdep = n.zdep || nodes_[x].zdep;
      }
    }
    if (n.op != Op::External && n.op != Op::View) n.zs = n.zdep ? zbytes(n) / (int64_t)sizeof(double) : 0;
    if (n.zdep && (n.op == Op::FieldVjp || n.op == Op::Seed || n.op == Op::TraceVjp))
      throw Error("adjoints of stacked trajectories are not supported");
  }
}

// Offset tables of one index group: row-major enumeration of `letters` with
// the given strides.
// Offsets of every combination of the letters (first letter slowest); with
// a pair "pq" of them, only the combinations with index(p) < index(q), or
// index(p) <= index(q) with `diag`.  With `mirror` the offsets written are
// those of the same combinations with p and q exchanged (-1 where equal).
static void group_table(const std::string& letters, const std::map<char, int64_t>& ext,
                        const std::map<char, int64_t>& st, std::vector<int64_t>& out,
                        const std::string& pair = "", bool diag = false, bool mirror = false) {
  int64_t total = 1;
  for (char c : letters) total *= ext.at(c);
  const int dp = pair.empty() ? -1 : (int)letters.find(pair[0]);
  const int dq = pair.empty() ? -1 : (int)letters.find(pair[1]);
  out.clear();
  out.reserve(total);
  for (int64_t idx = 0; idx < total; ++idx) {
    int64_t rem = idx, off = 0, ip = 0, iq = 0;
    for (int d = (int)letters.size() - 1; d >= 0; --d) {
      const int64_t e = ext.at(letters[d]);
      const int64_t i = rem % e;
      rem /= e;
      off += i * st.at(letters[d]);
      if (d == dp) ip = i;
      if (d == dq) iq = i;
    }
    if (dp >= 0 && dq >= 0 && (diag ? ip > iq : ip >= iq)) continue;
    if (mirror) {
      const int64_t sp = st.at(letters[dp]), sq = st.at(letters[dq]);
      off = ip == iq ? -1 : off - ip * sp - iq * sq + iq * sp + ip * sq;
    }
    out.push_back(off);
  }
}

void Graph::lower_contract(int id) {
  Node& n = nodes_[id];
  std::vector<std::string> ins;
  std::string lc;
  parse_spec(n.spec, ins, lc);
  std::string la = ins[0], lb = ins[1];
  int a = n.a, b = n.b;
  // strides are known before binding (external/view strides are fixed,
  // owned buffers are contiguous)
  auto strides_of = [&](int x, int64_t st[4]) {
    const Node& m = nodes_[x];
    if (m.op == Op::External) {
      for (int d = 0; d < 4; ++d) st[d] = m.ext.st[d];
    } else if (m.op == Op::View) {
      for (int d = 0; d < 4; ++d) st[d] = m.vst[d];
    } else {
      contiguous_strides(*this, m.shape, st);
    }
  };
  int64_t sa4[4], sb4[4], sc4[4];
  strides_of(a, sa4);
  strides_of(b, sb4);
  contiguous_strides(*this, n.shape, sc4);
  std::map<char, int64_t> ext, sa, sb, sc;
  for (int d = 0; d < (int)la.size(); ++d) {
    ext[la[d]] = extent(nodes_[a].shape.cls[d]);
    sa[la[d]] = sa4[d];
  }
  for (int d = 0; d < (int)lb.size(); ++d) {
    ext[lb[d]] = extent(nodes_[b].shape.cls[d]);
    sb[lb[d]] = sb4[d];
  }
  for (int d = 0; d < (int)lc.size(); ++d) sc[lc[d]] = sc4[d];
  // put the output's fastest index on the N side when it is not a batch index
  if (!lc.empty()) {
    const char inner = lc.back();
    if (has(la, inner) && !has(lb, inner)) {
      std::swap(la, lb);
      std::swap(a, b);
      std::swap(sa, sb);
    }
  }
  // Stacked trajectories: the trajectory axis is the letter 'Z' of extent nz
  // with each operand's z-stride.  With one per-trajectory operand it joins
  // that operand's free letters, so the trajectories stack into its rows (or
  // columns) of one launch against the shared operand; with two it is a batch
  // letter (GETT batch) unless the one-trajectory contraction takes a
  // dense-GEMM, dot or skinny route, which then runs once per trajectory.
  const bool za = nodes_[a].zdep, zb = nodes_[b].zdep;
  n.zloop = false;
  bool zbatch = false;
  if (nz_ > 1 && (za || zb)) {
    if (za && zb) {
      zbatch = true;  // decided below, once the one-trajectory route is known
    } else {
      (za ? la : lb).push_back('Z');
      lc.push_back('Z');
      ext['Z'] = nz_;
      (za ? sa : sb)['Z'] = nodes_[a].zdep ? nodes_[a].zs : nodes_[b].zs;
      sc['Z'] = n.zs;
    }
  }
  std::string gb, gm, gn, gk;
  auto groups = [&] {
    gb.clear();
    gm.clear();
    gn.clear();
    gk.clear();
    for (char c : lc) {
      if (has(la, c) && has(lb, c)) gb.push_back(c);
      else if (has(la, c)) gm.push_back(c);
      else gn.push_back(c);
    }
    for (char c : la)
      if (has(lb, c) && !has(lc, c)) gk.push_back(c);
  };
  groups();
  auto gsize = [&](const std::string& ls) {
    int64_t s = 1;
    for (char c : ls) s *= ext[c];
    return s;
  };
  // Large contractions run as dense GEMMs on the DGEMM kernel; the complex
  // operand (if only one is complex) goes on the row side, planes stacked.
  const double mul0 = (nodes_[a].cplx && nodes_[b].cplx) ? 8 : ((nodes_[a].cplx || nodes_[b].cplx) ? 4 : 2);
  // two summed letters over which both operands are antisymmetric: the
  // summand is symmetric in them, so the sum is twice the sum over p < q
  n.kpair.clear();
  n.kfac = 1.0;
  if (k_pairs_ && gb.empty())
    for (size_t x = 0; x < gk.size() && n.kpair.empty(); ++x)
      for (size_t y = x + 1; y < gk.size() && n.kpair.empty(); ++y) {
        const char p = gk[x], q = gk[y];
        if (ext[p] != ext[q]) continue;
        if ((nodes_[a].anti & anti_bit((int)la.find(p), (int)la.find(q))) &&
            (nodes_[b].anti & anti_bit((int)lb.find(p), (int)lb.find(q)))) {
          n.kpair = std::string{p, q};
          n.kfac = 2.0;
        }
      }
  const double kpair_frac = n.kpair.empty() ? 1.0 : (ext[n.kpair[0]] - 1.0) / (2.0 * ext[n.kpair[0]]);
  const double gflops = mul0 * gsize(gm) * (double)gsize(gn) * gsize(gk) * kpair_frac;
  const double min_flops = gemm_min_flops();
  // dot-product shapes (both free groups small, K long) also stream well
  // through the GEMM path's split-K
  const bool dot_shape = gsize(gm) >= 16 && gsize(gn) >= 16 && gsize(gk) >= 8192 && gflops >= 2.5e8;
  // In split precision an integral block of at least 64 MB against a few
  // amplitude rows also goes through the GEMM path: its digit planes are cut
  // once per bind, so each step streams S bytes per element from the digit
  // cache instead of 8 through the skinny or strided kernels.
  const int64_t stream_min = static_stream_min();
  const bool static_stream =
      rt_->split_digits > 0 && !n.fp64 && min_flops > 0.0 && stream_min >= 0 && gsize(gk) >= 64 &&
      ((is_static(a) && gsize(gm) * gsize(gk) >= stream_min && gsize(gn) >= 4) ||
       (is_static(b) && gsize(gn) * gsize(gk) >= stream_min && gsize(gm) >= 4));
  const bool use_gemm = gb.empty() && (min_flops == 0.0 ||
                                       (gflops >= min_flops && gsize(gm) >= 64 && gsize(gn) >= 64 && gsize(gk) >= 32) ||
                                       (dot_shape && min_flops <= 2e9) || static_stream);
  if (use_gemm && !nodes_[a].cplx && nodes_[b].cplx) {
    std::swap(la, lb);
    std::swap(a, b);
    std::swap(sa, sb);
    groups();
  }
  auto innermost = [&](const std::string& ls, const std::map<char, int64_t>& st) -> char {
    char best = 0;
    int64_t bs = -1;
    for (char c : ls) {
      const int64_t s = st.at(c);
      if (s > 0 && (bs < 0 || s < bs)) {
        bs = s;
        best = c;
      }
    }
    return best;
  };
  const char ia = innermost(la, sa), ib = innermost(lb, sb);
  const bool akc = ia && has(gk, ia);
  const bool bkc = ib && has(gk, ib);
  // Dot products (long summed group, few rows) run in FP64 on the dot kernel, which streams
  // both operands once through the offset tables: up to 64 A rows (planes
  // counted) and 512 B rows whatever the precision, HBM-bound shapes on which
  // neither the 128 x 128 GEMM tiles nor digit splits pay; up to 512 rows on
  // both sides for a contraction computed in FP64 anyway (at 320 x 160 rows
  // it beats the gather + DGEMM path, not the digit GEMM).
  const int64_t dot_k = dot_min_k();
  const int64_t a_rows = (nodes_[a].cplx ? 2 : 1) * gsize(gm), b_rows = (nodes_[b].cplx ? 2 : 1) * gsize(gn);
  const bool in_fp64 = rt_->split_digits == 0 || n.fp64;
  // (an operand contiguous along K streams along the summed letters, one
  // contiguous along its rows along them; two contiguous along different
  // summed letters cannot both stream)
  const bool use_dot = use_gemm && dot_k >= 0 && !(akc && bkc && ia != ib) && b_rows <= kDotMaxRows &&
                       (a_rows <= 64 || (in_fp64 && a_rows <= kDotMaxRows)) &&
                       (double)gsize(gk) * kpair_frac >= (double)dot_k;
  if (zbatch) {
    const bool would_skinny = !use_gemm && std::min(gsize(gm), gsize(gn)) <= kSkinnyMax &&
                              (double)std::max(gsize(gm), gsize(gn)) * gsize(gk) >= skinny_min_work();
    if (use_gemm || use_dot || would_skinny) {
      n.zloop = true;
    } else {
      la.push_back('Z');
      lb.push_back('Z');
      lc.push_back('Z');
      ext['Z'] = nz_;
      sa['Z'] = nodes_[a].zs;
      sb['Z'] = nodes_[b].zs;
      sc['Z'] = n.zs;
      groups();
    }
  }
  auto by_stride = [&](std::string g, const std::map<char, int64_t>& st) {
    std::stable_sort(g.begin(), g.end(), [&](char x, char y) { return st.at(x) > st.at(y); });
    return g;
  };
  // Output-bound strided contractions (a short summed group, an output much
  // larger than either operand) keep the output's letter order on both free
  // sides, so that their stores run along the output's memory; otherwise a
  // side not contiguous along K follows its operand's memory order.
  const bool out_bound = !use_gemm && gb.empty() && gsize(gk) <= 32 &&
                         (double)gsize(gm) * gsize(gn) >= 4.0 * (double)gsize(gk) * (gsize(gm) + gsize(gn));
  gm = (akc || out_bound) ? gm : by_stride(gm, sa);
  gn = (bkc || out_bound) ? gn : by_stride(gn, sb);
  // On the GEMM path an output antisymmetric in two letters of the M (or N)
  // group is computed for their ordered combinations (diagonal included) and
  // the epilogue stores each element's mirror images as well.
  std::string& mpair = n.mpair;
  std::string& npair = n.npair;
  mpair.clear();
  npair.clear();
  if (k_pairs_ && use_gemm && !use_dot) {
    auto find_pair = [&](const std::string& grp, std::string& pr) {
      for (size_t x = 0; x < grp.size() && pr.empty(); ++x)
        for (size_t y = x + 1; y < grp.size() && pr.empty(); ++y)
          if (grp[x] != 'Z' && grp[y] != 'Z' && ext[grp[x]] == ext[grp[y]] &&
              (n.anti & anti_bit((int)lc.find(grp[x]), (int)lc.find(grp[y]))))
            pr = std::string{grp[x], grp[y]};
    };
    find_pair(gm, mpair);
    if ((double)gsize(gk) * kpair_frac >= (double)npair_min_k()) find_pair(gn, npair);
  }
  // summed letters in the memory order of an operand whose contiguous index
  // is summed; when both are, the per-step digit split of the larger one
  // follows its own memory order (integral blocks are split once per bind),
  // and a skinny contraction streams its big operand in its memory order
  const double a_cut = is_static(a) ? 0.0 : (double)gsize(gm), b_cut = is_static(b) ? 0.0 : (double)gsize(gn);
  const bool skinny_shape = !use_gemm && !use_dot && std::min(gsize(gm), gsize(gn)) <= kSkinnyMax;
  const bool k_by_b = bkc && (!akc || (rt_->split_digits > 0 && !n.fp64 && use_gemm && !use_dot && b_cut > a_cut) ||
                              (skinny_shape && gsize(gn) > gsize(gm)));
  gk = k_by_b ? by_stride(gk, sb) : by_stride(gk, sa);

  auto push = [&](const std::vector<int64_t>& t) {
    // every table starts on a 32-byte boundary, so that kernels can read four
    // consecutive offsets with two 16-byte loads
    while (host_tables_.size() % 4) host_tables_.push_back(0);
    const int64_t off = (int64_t)host_tables_.size();
    host_tables_.insert(host_tables_.end(), t.begin(), t.end());
    return off;
  };
  std::vector<int64_t> t;
  auto L = std::make_unique<Launch>();
  L->kind = Launch::GETT;
  group_table(gm, ext, sa, t, mpair, true);
  L->tab[0] = push(t);
  const int64_t M = (int64_t)t.size();
  group_table(gk, ext, sa, t, n.kpair);
  L->tab[1] = push(t);
  const int64_t K = (int64_t)t.size();
  group_table(gn, ext, sb, t, npair, true);
  L->tab[2] = push(t);
  const int64_t Nn = (int64_t)t.size();
  group_table(gk, ext, sb, t, n.kpair);
  L->tab[3] = push(t);
  group_table(gm, ext, sc, t, mpair, true);
  L->tab[4] = push(t);
  group_table(gn, ext, sc, t, npair, true);
  L->tab[5] = push(t);
  if (!mpair.empty()) {
    group_table(gm, ext, sc, t, mpair, true, true);
    L->tab[9] = push(t);
  }
  if (!npair.empty()) {
    group_table(gn, ext, sc, t, npair, true, true);
    L->tab[10] = push(t);
  }
  int64_t Bt = 1;
  if (!gb.empty()) {
    group_table(gb, ext, sa, t);
    L->tab[6] = push(t);
    Bt = (int64_t)t.size();
    group_table(gb, ext, sb, t);
    L->tab[7] = push(t);
    group_table(gb, ext, sc, t);
    L->tab[8] = push(t);
  }
  if (M > INT32_MAX || Nn > INT32_MAX || K > INT32_MAX || Bt > 65535) throw Error("contraction too large: " + n.spec);
  GettArgs& g = L->g;
  g.M = (int)M;
  g.N = (int)Nn;
  g.K = (int)K;
  g.batch = (int)Bt;
  g.tiles_m = (int)((M + 63) / 64);
  g.tiles_n = (int)((Nn + 63) / 64);
  g.alpha_re = n.ar * n.kfac;
  g.alpha_im = n.ai * n.kfac;
  g.accumulate = 0;
  const int64_t tiles = (int64_t)g.tiles_m * g.tiles_n * Bt;
  int splits = 1;
  // Few tiles: split K so that each block runs a couple of 16-wide steps.
  // Such launches are bound by the chain of dependent global reads per step
  // (offset tables, operands), not by bandwidth or MFMA rate; at most 64
  // partial products keep the ordered reduction cheap.
  if (tiles < 256 && K >= 64) {
    splits = (int)std::min<int64_t>((K + 31) / 32, std::min<int64_t>(64, std::max<int64_t>(1, 1024 / tiles)));
  }
  int64_t chunk = (K + splits - 1) / splits;
  chunk = (chunk + 15) / 16 * 16;
  if (chunk < 16) chunk = 16;
  splits = (int)((K + chunk - 1) / chunk);
  if (splits < 1) splits = 1;
  g.splits = splits;
  g.k_chunk = (int)chunk;
  L->ac = nodes_[a].cplx;
  L->bc = nodes_[b].cplx;
  L->akc = akc;
  L->bkc = bkc;
  L->c_cplx = n.cplx;
  n.a = a;  // orientation used by the launch
  n.b = b;
  n.scratch_len = splits > 1 ? (int64_t)splits * Bt * M * Nn * (L->ac || L->bc ? 2 : 1) * (int64_t)sizeof(double) : 0;
  // Skinny shapes (one side at most kSkinnyMax rows) stream the big operand
  // once on the VALU kernel instead of padding 64 x 64 MFMA tiles.
  // The big operand must be streamed along its rows (contiguous index not in
  // K), except for matrix-vector shapes (at most 4 small rows).
  const bool small_a = M <= Nn;
  const bool big_kmajor = small_a ? bkc : akc;
  if (!use_gemm && Bt == 1 && std::min(M, Nn) <= kSkinnyMax && (double)std::max(M, Nn) * K >= skinny_min_work() &&
      (!big_kmajor || std::min(M, Nn) <= 4)) {
    L->skinny = true;
    L->small_is_a = small_a;
    SkinnyArgs& k = L->sk;
    k.I = (int)(L->small_is_a ? M : Nn);
    k.J = L->small_is_a ? Nn : M;
    k.K = K;
    k.g_kmajor = L->small_is_a ? bkc : akc;
    const int64_t per = k.g_kmajor ? 4 : 256;
    const int64_t blocks = (k.J + per - 1) / per;
    int64_t sp = std::max<int64_t>(1, std::min<int64_t>((1024 + blocks - 1) / blocks, (K + 255) / 256));
    int64_t ch = (K + sp - 1) / sp;
    ch = (ch + 63) / 64 * 64;
    sp = (K + ch - 1) / ch;
    k.splits = (int)sp;
    k.k_chunk = ch;
    k.alpha_re = n.ar * n.kfac;
    k.alpha_im = n.ai * n.kfac;
    // k-major big operand whose summed index runs only a few elements: the
    // short-run kernel pair when its blocks of run_rows x run elements are
    // contiguous (checked on the offset tables)
    if (k.g_kmajor && k.I <= 4) {
      const int64_t* gr = host_tables_.data() + (small_a ? L->tab[2] : L->tab[0]);
      const int64_t* gkt = host_tables_.data() + (small_a ? L->tab[3] : L->tab[1]);
      int64_t R = 1;
      while (R < K && gkt[R] == gkt[0] + R) ++R;
      bool ok = R >= 2 && R <= 32 && K % R == 0 && K / R >= 2;
      for (int64_t q = 0; ok && q < K; ++q) ok = gkt[q] == gkt[q - q % R] + q % R;
      int64_t Jl = 1;
      while (ok && Jl < k.J && gr[Jl] == gr[0] + Jl * R) ++Jl;
      ok = ok && k.J % Jl == 0 && Jl * (R + 1) <= kSkinnyRunsMax && Jl * R >= skinny_runs_min() &&
           k.J / Jl <= 65535;
      for (int64_t j = 0; ok && j < k.J; ++j) ok = gr[j] == gr[j - j % Jl] + (j % Jl) * R;
      if (ok) {
        k.run = (int)R;
        k.run_rows = (int)Jl;
        k.splits = (int)(K / R);
        k.k_chunk = R;
        sp = K / R;
      }
    }
    n.scratch_len = sp > 1 ? sp * (L->ac || L->bc ? 2 : 1) * (int64_t)k.I * k.J * (int64_t)sizeof(double) : 0;
  }
  if (use_dot) {
    GemmPath& gp = L->gp;
    gp.on = true;
    gp.dot = true;
    gp.pa = L->ac ? 2 : 1;
    gp.pb = L->bc ? 2 : 1;
    // about 512 blocks (two per CU, one round; TDCC_DOT_BLOCKS) over the row
    // and column tiles and K ranges of 32-multiples: 12 % faster than 1024,
    // 25 % faster than 256 or 2048 on this step's dot products
    const int64_t ntn = ((gp.pb * Nn + 63) / 64) * ((gp.pa * M + 63) / 64);
    const char* dbe = std::getenv("TDCC_DOT_BLOCKS");
    const int64_t dblocks = dbe ? std::max<int64_t>(1, std::atoll(dbe)) : 512;
    int64_t sp = std::max<int64_t>(1, std::min<int64_t>((K + 255) / 256, (dblocks + ntn - 1) / ntn));
    int64_t chunk = (K + sp - 1) / sp;
    chunk = (chunk + 31) / 32 * 32;
    sp = (K + chunk - 1) / chunk;
    gp.splits = (int)sp;
    gp.k_chunk = chunk;
    n.scratch_len = sp * gp.pa * M * gp.pb * Nn * (int64_t)sizeof(double);
    L->node = id;
    launches_.push_back(std::move(L));
    return;
  }
  if (use_gemm) {
    GemmPath& gp = L->gp;
    gp.on = true;
    gp.pa = L->ac ? 2 : 1;
    gp.pb = L->bc ? 2 : 1;
    gp.bt = bkc;
    const int64_t R = gp.pa * M, NB = gp.pb * Nn;
    if (rt_->split_digits > 0 && !n.fp64) {
      // digit planes of A' (R x Kp) and B' (NB x Kp); each operand is split
      // along its contiguous index group (k, or its rows)
      const int S = rt_->split_digits;
      auto al256 = [](int64_t x) { return (x + 255) / 256 * 256; };
      gp.S = S;
      gp.a_cols = !akc;
      gp.b_cols = !bkc;
      gp.Kp = (K + 63) / 64 * 64;
      // 3M: complex x complex, long enough K for the fused single-split path
      // of the LDS-direct kernel, and enough work for the two launches to pay
      gp.m3 = L->ac && L->bc && m3_mode() > 0 && ozgemm_interleaved_ok(S) &&
              (m3_mode() == 2 || (gp.Kp > 256 && M >= 512 && Nn >= 512 && (double)M * Nn * K >= 1e9));
      const int64_t Rx = gp.m3 ? 3 * M : R, NBx = gp.m3 ? 3 * Nn : NB;  // digit rows
      gp.a_plane = al256(Rx * gp.Kp);
      gp.b_plane = al256(NBx * gp.Kp);
      const int64_t bn = S <= 5 ? 128 : 64;
      const int64_t tiles = ((R + 127) / 128) * ((NB + bn - 1) / bn);
      int64_t sp = 1;
      if (tiles < 240 && K >= 2048) sp = std::min<int64_t>((K + 1023) / 1024, (480 + tiles - 1) / tiles);
      if (split_model_on() && tiles < 4 * kModelCUs && K >= 2048) sp = model_splits(tiles, gp.Kp, R, NB, sp);
      if (const char* fs = std::getenv("TDCC_GEMM_SPLITS")) sp = std::max<int64_t>(1, std::min<int64_t>(std::atol(fs), gp.Kp / 64));
      const int64_t maxk = ozgemm_max_k(S);
      sp = std::max<int64_t>(sp, (gp.Kp + maxk - 1) / maxk);
      int64_t chunk = (gp.Kp + sp - 1) / sp;
      chunk = (chunk + 63) / 64 * 64;
      sp = (gp.Kp + chunk - 1) / chunk;
      gp.splits = (int)sp;
      gp.k_chunk = sp > 1 ? chunk : 0;
      // short K (at most 8 K steps of 32): a digit row's K block keeps its S
      // digits side by side, so a K step reads whole lines of each row (the
      // planar layout reads a quarter line per row and step, and at short K
      // the lines leave L2 before the next steps use the rest)
      if (sp > 1) gp.m3 = false;  // partial products go through the scatter kernel
      gp.dil = gp.Kp <= 256 && ozgemm_interleaved_ok(S) && !gp.m3;
      if (!gp.m3) {
        gp.a_plane = al256(R * gp.Kp);
        gp.b_plane = al256(NB * gp.Kp);
      }
      // a row-oriented operand whose rows are too long for the single-pass
      // split (over 16384) gets one scale per row and K split, so that each
      // chunk is split in one pass instead of three (max, scale, digits)
      auto chunked = [&](bool cols) {
        return sp > 1 && !cols && chunk <= 16384 && gp.Kp > 256 && (gp.Kp > 16384 || split_chunked_always());
      };
      gp.a_chunked = chunked(gp.a_cols);
      gp.b_chunked = chunked(gp.b_cols);
      const int64_t a_scales = (gp.m3 ? 3 * M : R) * (gp.a_chunked ? sp : 1),
                    b_scales = (gp.m3 ? 3 * Nn : NB) * (gp.b_chunked ? sp : 1);
      gp.a8_off = 0;
      gp.b8_off = S * gp.a_plane;
      gp.sa_off = gp.b8_off + S * gp.b_plane;
      gp.sb_off = gp.sa_off + al256(8 * a_scales);
      gp.pa_off = gp.sb_off + al256(8 * b_scales);
      gp.pb_off = gp.pa_off + al256(ozaki_split_workspace(gp.m3 ? 3 * M : R, gp.Kp, gp.a_cols));
      gp.zb_off = gp.pb_off + al256(ozaki_split_workspace(gp.m3 ? 3 * Nn : NB, gp.Kp, gp.b_cols));
      gp.fused = sp == 1;
      gp.p12_off = gp.zb_off;
      n.scratch_len = gp.zb_off + (gp.m3 ? al256(8 * 2 * M * Nn) : (gp.fused ? 0 : al256(8 * sp * R * NB)));
      // layout of an operand's digit planes: its value, its index labels, the
      // row and summed letters in order, pair restrictions, orientation, rows
      // (planes included), digits, plane interleaving
      auto layout_key = [&](int x, const std::string& labels, const std::string& rows, const std::string& rpair,
                            bool cols, int64_t nrows, bool chk) {
        return nodes_[x].key + "|" + labels + "|" + rows + "|" + rpair + "|" + gk + "|" + n.kpair + "|" +
               (cols ? "c" : "r") + "|" + std::to_string(nrows) + "|" + std::to_string(S) + "|" +
               (gp.m3 ? "3" : (gp.fused ? "i" : "s")) + (gp.dil ? "|d" : "") +
               (chk ? "|k" + std::to_string(chunk) : "");
      };
      auto cache_entry = [&](int x, const std::string& key, int64_t nscales, int64_t plane) {
        if (!is_static(x)) return -1;
        auto it = cache_index_.find(key);
        if (it != cache_index_.end()) return it->second;
        DigitCache e;
        e.digits_off = cache_bytes_;
        e.scale_off = cache_bytes_ + S * plane;
        cache_bytes_ = e.scale_off + al256(8 * nscales);
        cache_.push_back(e);
        const int idx = (int)cache_.size() - 1;
        cache_index_[key] = idx;
        return idx;
      };
      const std::string akey = layout_key(a, la, gm, mpair, gp.a_cols, gp.m3 ? 3 * M : R, gp.a_chunked);
      const std::string bkey = layout_key(b, lb, gn, npair, gp.b_cols, gp.m3 ? 3 * Nn : NB, gp.b_chunked);
      gp.a_cache = cache_entry(a, akey, a_scales, gp.a_plane);
      gp.b_cache = cache_entry(b, bkey, b_scales, gp.b_plane);
      if (gp.a_cache < 0 && !n.zloop) dyn_uses_[akey].push_back({id, 0, a, S * gp.a_plane + al256(8 * a_scales)});
      if (gp.b_cache < 0 && bkey != akey && !n.zloop)
        dyn_uses_[bkey].push_back({id, 1, b, S * gp.b_plane + al256(8 * b_scales)});
      L->node = id;
      launches_.push_back(std::move(L));
      return;
    }
    const int64_t lda = K + (K & 1);
    gp.ldb = gp.bt ? lda : NB + (NB & 1);
    const int64_t tiles = ((R + 127) / 128) * ((NB + 127) / 128);
    int64_t sp = 1;
    if (tiles < 240 && K >= 2048) sp = std::min<int64_t>((K + 1023) / 1024, (480 + tiles - 1) / tiles);
    if (const char* fs = std::getenv("TDCC_GEMM_SPLITS")) sp = std::max<int64_t>(1, std::min<int64_t>(std::atol(fs), (K + 15) / 16));
    int64_t chunk = (K + sp - 1) / sp;
    chunk = (chunk + 15) / 16 * 16;
    sp = (K + chunk - 1) / chunk;
    gp.splits = (int)std::max<int64_t>(sp, 1);
    gp.k_chunk = gp.splits > 1 ? chunk : 0;
    auto al = [](int64_t x) { return (x + 31) / 32 * 32; };  // 256-byte alignment in doubles
    gp.a_off = 0;
    gp.b_off = al(R * lda);
    gp.z_off = gp.b_off + al(gp.bt ? NB * lda : K * gp.ldb);
    n.scratch_len = (gp.z_off + al((int64_t)gp.splits * R * NB)) * (int64_t)sizeof(double);
  }
  L->node = id;
  launches_.push_back(std::move(L));
}

// Whether the K offsets at `tab` (K entries) step evenly: offset k is
// first + k * step, step >= 0.
bool Graph::linear_table(int64_t tab, int64_t K, int64_t& step, int64_t& first) const {
  if (K < 1) return false;
  const int64_t* t = host_tables_.data() + tab;
  first = t[0];
  step = K > 1 ? t[1] - t[0] : 0;
  if (step < 0) return false;
  for (int64_t k = 0; k < K; ++k)
    if (t[k] != first + k * step) return false;
  return true;
}

void Graph::bind(double* arena) {
  arena_ = arena;
  for (auto& e : cache_) e.filled = false;
  bool cache_launched = false;
  // rebuild every launch record with device pointers
  std::vector<std::unique_ptr<Launch>> gett;
  for (auto& L : launches_)
    if (L->kind == Launch::GETT) gett.push_back(std::move(L));
  launches_.clear();
  size_t gi = 0;
  const int N = (int)nodes_.size();
  const int64_t n2 = (int64_t)(o_ + v_) * (o_ + v_);
  for (int p = 0; p < N; ++p) {
    const int i = order_[p];
    if (!live_[i]) continue;
    const Node& n = nodes_[i];
    const size_t nl0 = launches_.size();
    switch (n.op) {
      case Op::Contract: {
        auto L0 = std::move(gett.at(gi++));
        if (L0->node != i) throw Error("internal: contraction launch order");
        const int zc = n.zloop ? nz_ : 1;  // once per trajectory, or once over the stacked rows
        for (int z = 0; z < zc; ++z) {
        auto L = z + 1 == zc ? std::move(L0) : std::make_unique<Launch>(*L0);
        GettArgs& g = L->g;
        const Dev A = resolve(n.a, z), B = resolve(n.b, z), C = resolve(i, z);
        g.a_re = A.re;
        g.a_im = L->ac ? A.im : nullptr;
        g.b_re = B.re;
        g.b_im = L->bc ? B.im : nullptr;
        g.c_re = C.re;
        g.c_im = n.cplx ? C.im : nullptr;
        g.am = dev_tables_ + L->tab[0];
        g.ak = dev_tables_ + L->tab[1];
        g.bn = dev_tables_ + L->tab[2];
        g.bk = dev_tables_ + L->tab[3];
        g.cm = dev_tables_ + L->tab[4];
        g.cn = dev_tables_ + L->tab[5];
        g.ab = L->tab[6] >= 0 ? dev_tables_ + L->tab[6] : nullptr;
        g.bb = L->tab[7] >= 0 ? dev_tables_ + L->tab[7] : nullptr;
        g.cb = L->tab[8] >= 0 ? dev_tables_ + L->tab[8] : nullptr;
        g.partial = n.scratch_len > 0 ? arena_ + n.scratch_off / (int64_t)sizeof(double) : nullptr;
        if (L->skinny) {
          SkinnyArgs& k = L->sk;
          const bool sa_ = L->small_is_a;
          k.s_re = sa_ ? g.a_re : g.b_re;
          k.s_im = sa_ ? g.a_im : g.b_im;
          k.g_re = sa_ ? g.b_re : g.a_re;
          k.g_im = sa_ ? g.b_im : g.a_im;
          k.sr = sa_ ? g.am : g.bn;
          k.sk = sa_ ? g.ak : g.bk;
          k.gr = sa_ ? g.bn : g.am;
          k.gk = sa_ ? g.bk : g.ak;
          k.cs = sa_ ? g.cm : g.cn;
          k.cb = sa_ ? g.cn : g.cm;
          k.c_re = C.re;
          k.c_im = n.cplx ? C.im : nullptr;
          k.partial = g.partial;
          if ((k.s_im || k.g_im) && !k.c_im) throw Error("internal: complex product into a real output");
          if ((k.splits > 1 || k.run > 0) && !k.partial) throw Error("internal: skinny partial sums without scratch");
          if (k.run > 0 && (k.splits != k.K / k.run || k.splits < 2 || k.run > 32 || k.I > 4 ||
                            (int64_t)k.run_rows * (k.run + 1) > kSkinnyRunsMax || k.J % k.run_rows != 0))
            throw Error("internal: inconsistent short-run skinny launch");
        }
        if (L->gp.dot) {
          const GemmPath& gp = L->gp;
          DotArgs& d = L->dt;
          d.a_re = g.a_re;
          d.a_im = g.a_im;
          d.b_re = g.b_re;
          d.b_im = g.b_im;
          d.am = g.am;
          d.ak = g.ak;
          d.bn = g.bn;
          d.bk = g.bk;
          d.M = g.M;
          d.N = g.N;
          d.K = g.K;
          d.k_chunk = gp.k_chunk;
          d.splits = gp.splits;
          d.a_kc = L->akc ? 1 : 0;
          d.b_kc = L->bkc ? 1 : 0;
          d.Z = arena_ + n.scratch_off / (int64_t)sizeof(double);
          GemmScatterArgs& sc = L->sc;
          sc.Z = d.Z;
          sc.ldz = gp.pb * (int64_t)g.N;
          sc.splits = 1;
          sc.z_stride = 0;
          sc.M = g.M;
          sc.N = g.N;
          sc.kind = (L->ac && L->bc) ? 2 : (L->ac ? 1 : 0);
          if (!L->ac && L->bc) throw Error("internal: complex operand not on the row side");
          sc.cm = g.cm;
          sc.cn = g.cn;
          sc.cm2 = nullptr;
          sc.cn2 = nullptr;
          sc.out_re = C.re;
          sc.out_im = n.cplx ? C.im : nullptr;
          sc.alpha_re = n.ar * n.kfac;
          sc.alpha_im = n.ai * n.kfac;
          sc.accumulate = 0;
        } else if (L->gp.on && L->gp.S > 0) {
          const GemmPath& gp = L->gp;
          char* scr = reinterpret_cast<char*>(arena_) + n.scratch_off;
          SplitArgs& sa = L->sa;  // A' = [A_re; A_im] rows x K
          sa.src_re = A.re;
          sa.src_im = L->ac ? A.im : nullptr;
          sa.rtab = g.am;
          sa.ktab = g.ak;
          sa.Rb = g.M;
          sa.planes = gp.m3 ? 3 : gp.pa;
          sa.K = g.K;
          sa.Kp = gp.Kp;
          sa.S = gp.S;
          sa.D = reinterpret_cast<int8_t*>(scr + gp.a8_off);
          sa.plane = gp.a_plane;
          sa.scale = reinterpret_cast<double*>(scr + gp.sa_off);
          sa.pmax = reinterpret_cast<double*>(scr + gp.pa_off);
          sa.multipass = split_multipass();
          sa.interleave = gp.fused && !gp.m3;
          sa.dil = gp.dil;
          sa.sk_chunk = gp.a_chunked ? gp.k_chunk : 0;
          SplitArgs& sb = L->sb;  // B' = [B_re; B_im] (pb*N) x K
          sb.src_re = B.re;
          sb.src_im = L->bc ? B.im : nullptr;
          sb.rtab = g.bn;
          sb.ktab = g.bk;
          sb.Rb = g.N;
          sb.planes = gp.m3 ? 3 : gp.pb;
          sb.K = g.K;
          sb.Kp = gp.Kp;
          sb.S = gp.S;
          sb.D = reinterpret_cast<int8_t*>(scr + gp.b8_off);
          sb.plane = gp.b_plane;
          sb.scale = reinterpret_cast<double*>(scr + gp.sb_off);
          sb.pmax = reinterpret_cast<double*>(scr + gp.pb_off);
          sb.multipass = split_multipass();
          sb.interleave = gp.fused && !gp.m3;
          sb.dil = gp.dil;
          sb.sk_chunk = gp.b_chunked ? gp.k_chunk : 0;
          simplify_split_tables(sa, L->tab[0], L->tab[1]);
          simplify_split_tables(sb, L->tab[2], L->tab[3]);
          // static operands: digits live in the cache, split here once per bind
          // (the node's scratch serves as the split workspace)
          // shared digits of computed operands live in their arena block
          for (auto [idx, sp, plane] : {std::make_tuple(gp.a_share, &sa, gp.a_plane),
                                        std::make_tuple(gp.b_share, &sb, gp.b_plane)}) {
            if (idx < 0) continue;
            char* blk = reinterpret_cast<char*>(arena_) + shares_[idx].off;
            sp->D = reinterpret_cast<int8_t*>(blk);
            sp->scale = reinterpret_cast<double*>(blk + gp.S * plane);
          }
          for (auto [idx, sp, cols] : {std::make_tuple(gp.a_cache, &sa, gp.a_cols),
                                       std::make_tuple(gp.b_cache, &sb, gp.b_cols)}) {
            if (idx < 0) continue;
            DigitCache& e = cache_[idx];
            sp->D = reinterpret_cast<int8_t*>(cache_buf_ + e.digits_off);
            sp->scale = reinterpret_cast<double*>(cache_buf_ + e.scale_off);
            if (!e.filled) {
              launch_ozaki_split(*sp, cols, rt_->stream);
              e.filled = true;
              cache_launched = true;
            }
          }
          OzGemmArgs& oz = L->oz;
          oz.A = sa.D;
          oz.B = sb.D;
          oz.a_plane = gp.a_plane;
          oz.b_plane = gp.b_plane;
          oz.dil = gp.dil;
          oz.sa = sa.scale;
          oz.sb = sb.scale;
          oz.sa_split_stride = gp.a_chunked ? (int64_t)gp.pa * g.M : 0;
          oz.sb_split_stride = gp.b_chunked ? (int64_t)gp.pb * g.N : 0;
          oz.C = reinterpret_cast<double*>(scr + gp.zb_off);
          oz.M = (int)(gp.pa * g.M);
          oz.N = (int)(gp.pb * g.N);
          oz.ldc = oz.N;
          oz.Kp = (int)gp.Kp;
          oz.S = launch_digits(n.spec, false, gp.S);
          if (gp.dil && !ozgemm_interleaved_ok(oz.S)) oz.S = gp.S;
          oz.Sl = gp.S;
          oz.k_chunk = (int)gp.k_chunk;
          oz.splits = gp.splits;
          oz.c_split_stride = (int64_t)oz.M * oz.N;
          if (gp.m3) {
            // launch 1: P1, P2 (digit rows [0, M) and [M, 2M) of A against the
            // same of B) into the node scratch; launch 2: P3 (rows [2M, 3M))
            // and the complex result through the output tables
            oz.M = g.M;
            oz.N = g.N;
            oz.ldc = g.N;
            oz.C = reinterpret_cast<double*>(scr + gp.p12_off);
            oz.epi = -1;
            oz.nb = 2;
            oz.a_zstride = (int64_t)g.M * gp.Kp;
            oz.b_zstride = (int64_t)g.N * gp.Kp;
            oz.sa_zstride = g.M;
            oz.sb_zstride = g.N;
            oz.c_zstride = (int64_t)g.M * g.N;
            OzGemmArgs& o2 = L->oz2;
            o2 = oz;
            o2.nb = 1;
            o2.A = oz.A + 2 * oz.a_zstride;
            o2.B = oz.B + 2 * oz.b_zstride;
            o2.sa = oz.sa + 2 * oz.sa_zstride;
            o2.sb = oz.sb + 2 * oz.sb_zstride;
            o2.C = nullptr;
            o2.epi = 3;
            o2.p12 = oz.C;
            o2.p12_ld = g.N;
            o2.p12_stride = (int64_t)g.M * g.N;
            o2.cm = g.cm;
            o2.cn = g.cn;
            o2.cm2 = L->tab[9] >= 0 ? dev_tables_ + L->tab[9] : nullptr;
            o2.cn2 = L->tab[10] >= 0 ? dev_tables_ + L->tab[10] : nullptr;
            o2.out_re = C.re;
            o2.out_im = C.im;
            o2.alpha_re = n.ar * n.kfac;
            o2.alpha_im = n.ai * n.kfac;
          } else if (gp.fused) {
            oz.C = nullptr;
            oz.epi = (L->ac && L->bc) ? 2 : (L->ac ? 1 : 0);
            oz.cm = g.cm;
            oz.cn = g.cn;
            oz.cm2 = L->tab[9] >= 0 ? dev_tables_ + L->tab[9] : nullptr;
            oz.cn2 = L->tab[10] >= 0 ? dev_tables_ + L->tab[10] : nullptr;
            oz.out_re = C.re;
            oz.out_im = n.cplx ? C.im : nullptr;
            oz.alpha_re = n.ar * n.kfac;
            oz.alpha_im = n.ai * n.kfac;
            if (!L->ac && L->bc) throw Error("internal: complex operand not on the row side");
          }
          GemmScatterArgs& sc = L->sc;
          sc.Z = oz.C;
          sc.ldz = oz.ldc;
          sc.splits = gp.splits;
          sc.z_stride = oz.c_split_stride;
          sc.M = g.M;
          sc.N = g.N;
          sc.kind = (L->ac && L->bc) ? 2 : (L->ac ? 1 : 0);
          if (!L->ac && L->bc) throw Error("internal: complex operand not on the row side");
          sc.cm = g.cm;
          sc.cn = g.cn;
          sc.cm2 = L->tab[9] >= 0 ? dev_tables_ + L->tab[9] : nullptr;
          sc.cn2 = L->tab[10] >= 0 ? dev_tables_ + L->tab[10] : nullptr;
          sc.out_re = C.re;
          sc.out_im = n.cplx ? C.im : nullptr;
          sc.alpha_re = n.ar * n.kfac;
          sc.alpha_im = n.ai * n.kfac;
          sc.accumulate = 0;
        } else if (L->gp.on) {
          const GemmPath& gp = L->gp;
          double* scr = arena_ + n.scratch_off / (int64_t)sizeof(double);
          const int64_t M = g.M, Nn = g.N, K = g.K;
          const int64_t lda = K + (K & 1);
          Gather2dArgs& ga = L->ga;  // A' = [A_re; A_im] rows x K
          ga.src_re = A.re;
          ga.src_im = L->ac ? A.im : nullptr;
          ga.rtab = g.am;
          ga.ctab = g.ak;
          ga.Rb = M;
          ga.Cb = K;
          ga.planes = gp.pa;
          ga.planes_along_rows = 1;
          ga.dst = scr + gp.a_off;
          ga.ld = lda;
          Gather2dArgs& gbb = L->gb;
          gbb.src_re = B.re;
          gbb.src_im = L->bc ? B.im : nullptr;
          gbb.planes = gp.pb;
          gbb.dst = scr + gp.b_off;
          gbb.ld = gp.ldb;
          if (gp.bt) {  // B' = [B_re; B_im] (pb*N) x K
            gbb.rtab = g.bn;
            gbb.ctab = g.bk;
            gbb.Rb = Nn;
            gbb.Cb = K;
            gbb.planes_along_rows = 1;
          } else {      // B' = [B_re | B_im] K x (pb*N)
            gbb.rtab = g.bk;
            gbb.ctab = g.bn;
            gbb.Rb = K;
            gbb.Cb = Nn;
            gbb.planes_along_rows = 0;
          }
          DgemmArgs& dg = L->dg;
          dg.A = ga.dst;
          dg.B = gbb.dst;
          dg.C = scr + gp.z_off;
          dg.M = (int)(gp.pa * M);
          dg.N = (int)(gp.pb * Nn);
          dg.K = (int)K;
          dg.lda = lda;
          dg.ldb = gp.ldb;
          dg.ldc = gp.pb * Nn;
          dg.alpha = 1.0;
          dg.beta = 0.0;
          dg.k_chunk = (int)gp.k_chunk;
          dg.splits = gp.splits;
          dg.c_split_stride = (int64_t)dg.M * dg.ldc;
          GemmScatterArgs& sc = L->sc;
          sc.Z = dg.C;
          sc.ldz = dg.ldc;
          sc.splits = gp.splits;
          sc.z_stride = dg.c_split_stride;
          sc.M = (int)M;
          sc.N = (int)Nn;
          sc.kind = (L->ac && L->bc) ? 2 : (L->ac ? 1 : 0);
          if (!L->ac && L->bc) throw Error("internal: complex operand not on the row side");
          sc.cm = g.cm;
          sc.cn = g.cn;
          sc.cm2 = L->tab[9] >= 0 ? dev_tables_ + L->tab[9] : nullptr;
          sc.cn2 = L->tab[10] >= 0 ? dev_tables_ + L->tab[10] : nullptr;
          sc.out_re = C.re;
          sc.out_im = n.cplx ? C.im : nullptr;
          sc.alpha_re = n.ar * n.kfac;
          sc.alpha_im = n.ai * n.kfac;
          sc.accumulate = 0;
        }
        // a plain GETT launch whose K offsets step evenly reads them as
        // k * step, the first offset folded into the operand pointers
        g.ak_step = g.bk_step = -1;
        if (!L->skinny && !L->gp.on && gett_linear_k_on()) {
          int64_t st = 0, k0 = 0;
          if (linear_table(L->tab[1], g.K, st, k0)) {
            g.ak_step = st;
            g.a_re += k0;
            if (g.a_im) g.a_im += k0;
          }
          if (linear_table(L->tab[3], g.K, st, k0)) {
            g.bk_step = st;
            g.b_re += k0;
            if (g.b_im) g.b_im += k0;
          }
        }
        launches_.push_back(std::move(L));
        }
        break;
      }
      case Op::Lincomb:
      case Op::Assign: {
        const Dev O = n.op == Op::Assign ? resolve(n.b) : resolve(i);
        std::vector<Term> terms = n.op == Op::Assign ? std::vector<Term>{tm(n.a)} : n.terms;
        const bool out_cplx = n.op == Op::Assign ? nodes_[n.b].cplx : n.cplx;
        for (size_t t0 = 0; t0 < terms.size(); t0 += kMaxTerms) {
          auto L = std::make_unique<Launch>();
          L->kind = Launch::LIN;
          L->node = i;
          LincombArgs& l = L->l;
          l.out_re = O.re;
          l.out_im = out_cplx ? O.im : nullptr;
          l.rank = n.shape.rank;
          l.total = size(n.shape);
          for (int d = 0; d < 4; ++d) {
            l.ost[d] = O.st[d];
            l.dims[d] = d < n.shape.rank ? extent(n.shape.cls[d]) : 1;
          }
          l.accumulate = t0 > 0;
          l.nterms = (int)std::min<size_t>(kMaxTerms, terms.size() - t0);
          const Node& on = n.op == Op::Assign ? nodes_[n.b] : n;
          l.nz = on.zdep ? nz_ : 1;
          l.ozs = on.zdep ? on.zs : 0;
          // an antisymmetric result is summed at its canonical elements and
          // mirrored (one launch only)
          l.anti = 0;
          if (n.op == Op::Lincomb && n.shape.rank == 4 && terms.size() <= (size_t)kMaxTerms && lincomb_anti_on())
            l.anti = ((n.anti & anti_bit(0, 1)) ? 1 : 0) | ((n.anti & anti_bit(2, 3)) ? 2 : 0);
          for (int k = 0; k < l.nterms; ++k) {
            const Term& T = terms[t0 + k];
            const Dev S = resolve(T.src);
            l.t[k].re = S.re;
            l.t[k].im = nodes_[T.src].cplx ? S.im : nullptr;
            for (int d = 0; d < 4; ++d) l.t[k].st[d] = d < n.shape.rank && T.perm[d] >= 0 ? S.st[T.perm[d]] : 0;
            l.t[k].cr = T.cr;
            l.t[k].ci = T.ci;
            l.t[k].zs = nodes_[T.src].zdep ? nodes_[T.src].zs : 0;
            l.t[k].zs2 = (T.src2 >= 0 && nodes_[T.src2].zdep) ? nodes_[T.src2].zs : 0;
            l.t[k].re2 = nullptr;
            l.t[k].im2 = nullptr;
            for (int d = 0; d < 4; ++d) l.t[k].st2[d] = 0;
            if (T.src2 >= 0) {
              const Dev S2 = resolve(T.src2);
              l.t[k].re2 = S2.re;
              l.t[k].im2 = nodes_[T.src2].cplx ? S2.im : nullptr;
              for (int d = 0; d < 4; ++d) l.t[k].st2[d] = d < n.shape.rank && T.perm2[d] >= 0 ? S2.st[T.perm2[d]] : 0;
            }
          }
          launches_.push_back(std::move(L));
        }
        break;
      }
      case Op::Ladder: {
        // stacked trajectories: one GEMM over the packed rows of every
        // trajectory (the packing and unpacking run once per trajectory)
        const int zn = n.zdep ? nz_ : 1;
        {
        const int z = 0;
        auto L = std::make_unique<Launch>();
        L->kind = Launch::LADDER;
        L->node = i;
        const Dev U = resolve(n.a, z), O = resolve(i, z);
        if (!nodes_[n.a].cplx) throw Error("internal: real ladder input");
        LadderArgs& a = L->lad;
        a.u_re = U.re;
        a.u_im = U.im;
        a.out_re = O.re;
        a.out_im = O.im;
        a.o = o_;
        a.ij_pairs = n.ij_pairs;
        a.rows = ladder_rows(o_, n.ij_pairs);
        a.nz = zn;
        a.u_zs = nodes_[n.a].zdep ? nodes_[n.a].zs : 0;
        a.out_zs = n.zdep ? n.zs : 0;
        const int64_t R2 = 2 * a.rows * zn;
        a.v = v_;
        a.P = (int64_t)v_ * (v_ - 1) / 2;
        a.up = arena_ + n.scratch_off / (int64_t)sizeof(double);
        a.rp = a.up + R2 * a.P;
        a.vp = rt_->vpack[n.stage];
        a.k_chunk = ladder_k_chunk(R2, v_, rt_->split_digits);
        a.splits = (int)ladder_splits(R2, v_, rt_->split_digits);
        if (!a.vp) throw Error("packed <ab||ef> not built");
        a.S = rt_->split_digits;
        if (a.S > 0) {
          if (!rt_->vpack8[n.stage]) throw Error("digit planes of the packed <ab||ef> not built");
          char* scr = reinterpret_cast<char*>(arena_) + n.scratch_off;
          const int64_t Kp = ladder_Kp(v_), plane = ladder_plane(R2, v_);
          if (Kp != rt_->vpack_Kp) throw Error("internal: ladder digit row length");
          SplitArgs& su = a.su;  // rows of up, dense
          su.src_re = a.up;
          su.ldx = a.P;
          su.Rb = R2;
          su.planes = 1;
          su.K = a.P;
          su.Kp = Kp;
          su.S = a.S;
          su.D = reinterpret_cast<int8_t*>(scr + ladder_digits_off(R2, v_, a.S));
          su.plane = plane;
          su.scale = reinterpret_cast<double*>(scr + ladder_scale_off(R2, v_, a.S));
          su.pmax = reinterpret_cast<double*>(scr + ladder_pmax_off(R2, v_, a.S));
          su.multipass = split_multipass();
          OzGemmArgs& oz = a.oz;
          oz.A = su.D;
          oz.B = rt_->vpack8[n.stage];
          oz.a_plane = plane;
          oz.b_plane = rt_->vpack_plane;
          oz.sa = su.scale;
          oz.sb = rt_->vpack8_scale[n.stage];
          oz.C = a.rp;
          oz.ldc = a.P;
          oz.M = (int)R2;
          oz.N = (int)a.P;
          oz.Kp = (int)Kp;
          oz.S = launch_digits("", true, a.S);
          oz.Sl = a.S;
          const int64_t chunk = ladder_k_chunk(R2, v_, a.S);
          if (chunk == 0 && Kp > ozgemm_max_k(a.S)) throw Error("ladder K beyond the exact int32 range of the digit GEMM");
          oz.k_chunk = (int)chunk;
          oz.splits = (int)ladder_splits(R2, v_, a.S);
          oz.c_split_stride = R2 * a.P;
          a.splits = oz.splits;
        }
        launches_.push_back(std::move(L));
        }
        break;
      }
      case Op::Field: {
        for (int z = 0; z < (n.zdep ? nz_ : 1); ++z) {
        auto L = std::make_unique<Launch>();
        L->kind = Launch::FIELD;
        const Dev O = resolve(i, z);
        if (n.traj >= (int)rt_->batch.size()) throw Error("internal: trajectory without parameters");
        if (n.zdep && (int)rt_->batch.size() < nz_) throw Error("internal: stacked trajectories without parameters");
        const int tj = n.zdep ? z : n.traj;
        const bool bt = tj >= 0;
        L->f.F = rt_->F;
        L->f.mu_pol = rt_->mu + (int64_t)(bt ? rt_->batch[tj].pol : rt_->pol) * n2;
        L->f.out_re = O.re;
        L->f.out_im = O.im;
        L->f.n2 = n2;
        L->f.negA = bt ? rt_->batch[tj].negA : rt_->negA;
        L->f.table = bt ? rt_->batch[tj].field_table : rt_->field_table;
        L->f.counter = rt_->counter;
        L->f.stage = n.stage;
        launches_.push_back(std::move(L));
        }
        break;
      }
      case Op::Finalize: {
        for (int z = 0; z < (n.zdep ? nz_ : 1); ++z) {
        auto L = std::make_unique<Launch>();
        L->kind = Launch::FIN;
        const Dev Y = resolve(n.a, z);
        const Dev Bs = resolve(n.b, z);
        L->fin.y_re = Y.re;
        L->fin.y_im = Y.im;
        L->fin.base_re = Bs.re;
        L->fin.base_im = Bs.im;
        for (int s = 0; s < 4; ++s) {
          const Dev K = resolve(n.k[s], z);
          L->fin.k_re[s] = K.re;
          L->fin.k_im[s] = K.im;
          if (!nodes_[n.k[s]].cplx) throw Error("internal: real stage derivative");
        }
        L->fin.n = size(n.shape);
        L->fin.dt6 = rt_->dt6;
        L->fin.scale = rt_->round_scale;
        L->fin.lambda_sign = n.lambda_sign;
        launches_.push_back(std::move(L));
        }
        break;
      }
      case Op::Trace: {
        for (int z = 0; z < (n.zdep ? nz_ : 1); ++z) {
        auto L = std::make_unique<Launch>();
        L->kind = Launch::TRACE;
        const Dev d0 = resolve(n.k[0], z), d1 = resolve(n.k[1], z), d2 = resolve(n.k[2], z), d3 = resolve(n.k[3], z);
        L->tr.dij_re = d0.re;
        L->tr.dij_im = d0.im;
        L->tr.l1_re = d1.re;
        L->tr.l1_im = d1.im;
        L->tr.dai_re = d2.re;
        L->tr.dai_im = d2.im;
        L->tr.dab_re = d3.re;
        L->tr.dab_im = d3.im;
        L->tr.mu = rt_->mu;
        L->tr.o = o_;
        L->tr.v = v_;
        if (n.traj >= (int)rt_->batch.size()) throw Error("internal: trajectory without parameters");
        if (n.zdep && (int)rt_->batch.size() < nz_) throw Error("internal: stacked trajectories without parameters");
        const int tj = n.zdep ? z : n.traj;
        L->tr.out = tj >= 0 ? rt_->batch[tj].mu_out : rt_->mu_out;
        L->tr.counter = rt_->counter;
        launches_.push_back(std::move(L));
        }
        break;
      }
      case Op::FieldVjp: {
        auto L = std::make_unique<Launch>();
        L->kind = Launch::FVJP;
        L->node = i;
        if (!rt_->g || !rt_->ttab) throw Error("internal: field adjoint without gradient buffers");
        if (n.traj >= 0) throw Error("adjoint of batched programs is not supported");
        const Node& v = nodes_[n.b];
        const Dev A = resolve(n.a);
        FieldVjpArgs& f = L->fv;
        f.ab_re = A.re;
        f.ab_im = A.im;
        f.rows = (int)extent(v.shape.cls[0]);
        f.cols = (int)extent(v.shape.cls[1]);
        f.p0 = v.shape.cls[0] == Cls::V ? o_ : 0;
        f.q0 = v.shape.cls[1] == Cls::V ? o_ : 0;
        f.n = o_ + v_;
        f.mu_pol = rt_->mu + (int64_t)rt_->pol * n2;
        f.negA = rt_->negA;
        f.table = rt_->field_table;
        f.ttab = rt_->ttab;
        f.counter = rt_->counter;
        f.stage = n.stage;
        f.g = rt_->g;
        f.gmu_re = rt_->gmu_re;
        f.gmu_im = rt_->gmu_im;
        if (!nodes_[n.a].cplx) throw Error("internal: real field adjoint");
        launches_.push_back(std::move(L));
        break;
      }
      case Op::Seed: {
        auto L = std::make_unique<Launch>();
        L->kind = Launch::SEED;
        L->node = i;
        if (!rt_->wtab) throw Error("internal: seed without a weight table");
        const Dev O = resolve(i);
        SeedArgs& d = L->sd;
        d.out_re = O.re;
        d.out_im = O.im;
        d.rows = (int)extent(n.shape.cls[0]);
        d.cols = (int)extent(n.shape.cls[1]);
        d.p0 = n.shape.cls[0] == Cls::V ? o_ : 0;
        d.q0 = n.shape.cls[1] == Cls::V ? o_ : 0;
        d.n = o_ + v_;
        d.mu_pol = rt_->mu + (int64_t)rt_->pol * n2;
        d.wtab = rt_->wtab;
        d.counter = rt_->counter;
        launches_.push_back(std::move(L));
        break;
      }
      case Op::TraceVjp: {
        auto L = std::make_unique<Launch>();
        L->kind = Launch::TVJP;
        L->node = i;
        if (!rt_->gmu_re || !rt_->wtab) throw Error("internal: read-out adjoint without gradient buffers");
        const Dev d0 = resolve(n.k[0]), d1 = resolve(n.k[1]), d2 = resolve(n.k[2]), d3 = resolve(n.k[3]);
        TraceVjpArgs& t = L->tv;
        t.dij_re = d0.re;
        t.dij_im = d0.im;
        t.l1_re = d1.re;
        t.l1_im = d1.im;
        t.dai_re = d2.re;
        t.dai_im = d2.im;
        t.dab_re = d3.re;
        t.dab_im = d3.im;
        t.o = o_;
        t.v = v_;
        t.wtab = rt_->wtab;
        t.counter = rt_->counter;
        t.gmu_re = rt_->gmu_re;
        t.gmu_im = rt_->gmu_im;
        launches_.push_back(std::move(L));
        break;
      }
      default: break;
    }
    for (size_t j = nl0; j < launches_.size(); ++j) launches_[j]->node = i;
  }
  if (cache_launched) {
    TDCC_HIP(hipGetLastError());
    TDCC_HIP(hipStreamSynchronize(rt_->stream));
  }
  // 1: tile orders and persistent grids, which leave every output bit for bit
  // unchanged; 2: also fewer K splits (they reorder the FP64 sum of the
  // partials).  Default 1 for batched programs, off otherwise.
  const char* at = std::getenv("TDCC_AUTOTUNE");
  const int mode = at ? std::atoi(at) : (nz_ > 1 ? 1 : 0);
  if (mode > 0) autotune(mode >= 2);
  plan_streams();
}

// Two-stream step: the stream of every launch, the crossing events each
// Lambda side launch waits for first, and the event a T side launch records
// once the last launch of a crossing node is queued.
void Graph::plan_streams() {
  const size_t NL = launches_.size();
  lstream_.assign(NL, 0);
  lrecord_.assign(NL, -1);
  lwait_.assign(NL, {});
  first_post_ = -1;
  if (side_.empty()) return;
  auto root = [&](int x) {
    while (nodes_[x].op == Op::View) x = nodes_[x].base;
    return x;
  };
  std::map<int, int> event_of;  // crossing node -> event
  std::map<int, size_t> last_launch;
  for (size_t j = 0; j < NL; ++j) {
    const int nd = launches_[j]->node;
    if (nd < 0) throw Error("internal: launch without node");
    lstream_[j] = side_[nd] == 1 ? 1 : 0;
    if (first_post_ < 0 && side_[nd] == 2) first_post_ = (int)j;
    if (cross_[nd]) {
      if (!event_of.count(nd)) event_of.emplace(nd, (int)event_of.size());
      last_launch[nd] = j;
    }
  }
  for (auto& [nd, j] : last_launch) lrecord_[j] = event_of.at(nd);
  std::vector<char> waited(event_of.size(), 0);
  std::vector<int> in;
  for (size_t j = 0; j < NL; ++j) {
    if (!lstream_[j]) continue;
    inputs_of(nodes_[launches_[j]->node], in);
    for (int x : in) {
      auto it = event_of.find(root(x));
      if (it == event_of.end() || waited[it->second]) continue;
      if (last_launch.at(it->first) > j) throw Error("internal: crossing read before it is written");
      waited[it->second] = 1;
      lwait_[j].push_back(it->second);
    }
  }
  if (!sb_) TDCC_HIP(hipStreamCreateWithFlags(&sb_, hipStreamNonBlocking));
  if (!fork_ev_) TDCC_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
  if (!join_ev_) TDCC_HIP(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming));
  while (events_.size() < event_of.size()) {
    hipEvent_t e = nullptr;
    TDCC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    events_.push_back(e);
  }
}

namespace {
struct TuneChoice {
  int group = 0, pb = -1, k_chunk = 0, splits = 1;
  float ms = 0.f;
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune;  // launch shape -> choice (this process)
}  // namespace

void Graph::autotune(bool k_splits) {
  hipStream_t s = rt_->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TDCC_HIP(hipEventCreate(&e0));
  TDCC_HIP(hipEventCreate(&e1));
  std::string report;
  try {
    // the digit-GEMM launches (two for a 3M contraction)
    struct Target {
      OzGemmArgs* oz;
      GemmScatterArgs* sc;
      int* lsplits;
    };
    std::vector<Target> targets;
    for (auto& L : launches_) {
      if (L->kind == Launch::GETT && L->gp.on && L->gp.S > 0 && !L->gp.dot && L->g.M > 0 && L->g.N > 0) {
        targets.push_back({&L->oz, L->gp.fused ? nullptr : &L->sc, nullptr});
        if (L->gp.m3) targets.push_back({&L->oz2, nullptr, nullptr});
      } else if (L->kind == Launch::LADDER && L->lad.S > 0) {
        targets.push_back({&L->lad.oz, nullptr, &L->lad.splits});
      }
    }
    for (const Target& tg : targets) {
      OzGemmArgs* oz = tg.oz;
      GemmScatterArgs* sc = tg.sc;
      int* lsplits = tg.lsplits;
      char key[160];
      std::snprintf(key, sizeof key, "S=%d M=%d N=%d Kp=%d chunk=%d splits=%d epi=%d%s%s", oz->S, oz->M, oz->N,
                    oz->Kp, oz->k_chunk, oz->splits, oz->epi, oz->nb > 1 ? " nb=2" : "", k_splits ? " ksplits" : "");
      TuneChoice best;
      bool have = false;
      {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
          best = it->second;
          have = true;
        }
      }
      if (!have) {
        std::vector<TuneChoice> cand;
        std::vector<std::pair<int, int>> ks = {{oz->k_chunk, oz->splits}};
        // fewer K chunks (the partial buffers hold the planned count); not with
        // scales per K split, which are cut at the planned chunks
        if (k_splits && oz->splits > 2 && !oz->sa_split_stride && !oz->sb_split_stride)
          for (int sp = oz->splits / 2; sp >= 2; sp /= 2) {
            int64_t ch = (oz->Kp + sp - 1) / sp;
            ch = (ch + 63) / 64 * 64;
            if (ch > ozgemm_max_k(oz->S)) break;
            ks.push_back({(int)ch, (int)((oz->Kp + ch - 1) / ch)});
          }
        for (auto [ch, sp] : ks)
          for (int grp : {0, 1, 4, 8})
            for (int pb : {-1, 0, 256, 512}) {
              if (sp > 1 && pb > 0) continue;  // persistent blocks only without K splits
              TuneChoice c;
              c.group = grp;
              c.pb = pb;
              c.k_chunk = ch;
              c.splits = sp;
              cand.push_back(c);
            }
        OzGemmArgs base = *oz;
        for (auto& c : cand) {
          OzGemmArgs t = base;
          t.tune_group = c.group;
          t.tune_pb = c.pb;
          t.k_chunk = c.k_chunk;
          t.splits = c.splits;
          launch_ozgemm(t, s);  // warm
          TDCC_HIP(hipEventRecord(e0, s));
          for (int r = 0; r < 3; ++r) launch_ozgemm(t, s);
          TDCC_HIP(hipEventRecord(e1, s));
          TDCC_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          TDCC_HIP(hipEventElapsedTime(&ms, e0, e1));
          c.ms = ms / 3.f;
          if (!have || c.ms < best.ms) {
            best = c;
            have = true;
          }
        }
        TDCC_HIP(hipGetLastError());
        {
          std::lock_guard<std::mutex> lk(g_tune_mu);
          g_tune[key] = best;
        }
        char line[320];
        std::snprintf(line, sizeof line, "%s\tgroup=%d\tpersistent=%d\tk_chunk=%d\tsplits=%d\tms=%.4f\tdefault_ms=%.4f\n",
                      key, best.group, best.pb, best.k_chunk, best.splits, best.ms, cand.front().ms);
        report += line;
      }
      oz->tune_group = best.group;
      oz->tune_pb = best.pb;
      oz->k_chunk = best.k_chunk;
      oz->splits = best.splits;
      if (sc) sc->splits = best.splits;
      if (lsplits) *lsplits = best.splits;
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  TDCC_HIP(hipEventDestroy(e0));
  TDCC_HIP(hipEventDestroy(e1));
  if (const char* f = std::getenv("TDCC_AUTOTUNE_FILE"); f && !report.empty()) {
    if (FILE* fp = std::fopen(f, "a")) {
      std::fputs(report.c_str(), fp);
      std::fclose(fp);
    }
  }
}

static DgemmArgs ladder_gemm(const LadderArgs& a) {
  DgemmArgs g{};
  g.A = a.up;
  g.B = a.vp;
  g.C = a.rp;
  g.M = (int)(2 * a.rows * a.nz);
  g.N = (int)a.P;
  g.K = (int)a.P;
  g.lda = g.ldb = g.ldc = a.P;
  g.alpha = 1.0;
  g.beta = 0.0;
  g.k_chunk = (int)a.k_chunk;
  g.splits = a.S > 0 ? 1 : a.splits;
  g.c_split_stride = 2 * a.rows * a.nz * a.P;
  return g;
}

// Packing and unpacking of each trajectory's rows (one trajectory unless stacked).
static void ladder_pack(const LadderArgs& a, hipStream_t s) {
  for (int z = 0; z < a.nz; ++z)
    launch_pack_pairs(a.u_re + z * a.u_zs, a.u_im + z * a.u_zs, a.up + z * 2 * a.rows * a.P, a.o, a.v, a.ij_pairs, s);
}

static void ladder_unpack(const LadderArgs& a, hipStream_t s) {
  for (int z = 0; z < a.nz; ++z)
    launch_unpack_pairs(a.rp + z * 2 * a.rows * a.P, a.out_re + z * a.out_zs, a.out_im + z * a.out_zs, a.o, a.v, 1.0,
                        a.ij_pairs, s);
}

static void run_ladder(const LadderArgs& a, hipStream_t s) {
  ladder_pack(a, s);
  if (a.S > 0) {
    launch_ozaki_split(a.su, false, s);
    launch_ozgemm(a.oz, s);
  } else {
    launch_dgemm(ladder_gemm(a), true, s);
  }
  const int64_t packed = 2 * a.rows * a.nz * a.P;
  launch_sum_splits(a.rp, a.splits, packed, packed, s);
  ladder_unpack(a, s);
}

void Graph::run(hipStream_t s0) const {
  // two-stream step: fork the second stream off the caller's, and join it
  // back before the first finalize (or at the end)
  bool joined = side_.empty() || !sb_;
  if (!joined) {
    TDCC_HIP(hipEventRecord(fork_ev_, s0));
    TDCC_HIP(hipStreamWaitEvent(sb_, fork_ev_, 0));
  }
  auto join = [&] {
    TDCC_HIP(hipEventRecord(join_ev_, sb_));
    TDCC_HIP(hipStreamWaitEvent(s0, join_ev_, 0));
    joined = true;
  };
  for (size_t j = 0; j < launches_.size(); ++j) {
    const auto& L = launches_[j];
    if (!joined && (int)j == first_post_) join();
    hipStream_t s = s0;
    if (!joined) {
      for (int e : lwait_[j]) TDCC_HIP(hipStreamWaitEvent(sb_, events_[e], 0));
      if (lstream_[j]) s = sb_;
    }
    switch (L->kind) {
      case Launch::GETT:
        if (L->g.M == 0 || L->g.N == 0) break;
        if (L->gp.dot) {
          launch_dot(L->dt, L->sc, s);
          break;
        }
        if (L->gp.on && L->gp.S > 0) {
          if (L->gp.a_cache < 0 && L->gp.a_split) launch_ozaki_split(L->sa, L->gp.a_cols, s);
          if (L->gp.b_cache < 0 && L->gp.b_split) launch_ozaki_split(L->sb, L->gp.b_cols, s);
          launch_ozgemm(L->oz, s);
          if (L->gp.m3) launch_ozgemm(L->oz2, s);
          else if (!L->gp.fused) launch_gemm_scatter(L->sc, s);
          break;
        }
        if (L->gp.on) {
          launch_gather2d(L->ga, s);
          launch_gather2d(L->gb, s);
          launch_dgemm(L->dg, L->gp.bt, s);
          launch_gemm_scatter(L->sc, s);
          break;
        }
        if (L->skinny) {
          launch_skinny(L->sk, s);
          break;
        }
        launch_gett(L->g, L->ac, L->bc, L->akc, L->bkc, s);
        if (L->g.splits > 1) launch_gett_reduce(L->g, L->c_cplx, s);
        break;
      case Launch::LIN: launch_lincomb(L->l, s); break;
      case Launch::LADDER: run_ladder(L->lad, s); break;
      case Launch::FIELD: launch_field(L->f, s); break;
      case Launch::FIN: launch_finalize(L->fin, s); break;
      case Launch::TRACE: launch_trace(L->tr, s); break;
      case Launch::FVJP: launch_field_vjp(L->fv, s); break;
      case Launch::SEED: launch_seed(L->sd, s); break;
      case Launch::TVJP: launch_trace_vjp(L->tv, s); break;
    }
    if (!joined && lrecord_[j] >= 0) TDCC_HIP(hipEventRecord(events_[lrecord_[j]], s0));
  }
  if (!joined) join();
  TDCC_HIP(hipGetLastError());
}

void Graph::run_profiled(hipStream_t s, std::vector<LaunchRecord>& out) const {
  std::vector<hipEvent_t> ev;  // one pair per timed launch, created as needed
  ev.reserve(8 * launches_.size() + 8);
  std::vector<LaunchRecord> recs;
  std::string detail;
  auto timed = [&](const std::string& name, double flops, double bytes, auto&& fn) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TDCC_HIP(hipEventCreate(&e0));
    ev.push_back(e0);
    TDCC_HIP(hipEventCreate(&e1));
    ev.push_back(e1);
    TDCC_HIP(hipEventRecord(e0, s));
    fn();
    TDCC_HIP(hipEventRecord(e1, s));
    recs.push_back({name, flops, bytes, 0.f, detail});
    detail.clear();
  };
  try {
    for (const auto& L : launches_) {
      switch (L->kind) {
        case Launch::GETT: {
          if (L->g.M == 0 || L->g.N == 0) break;
          const double mul = (L->ac && L->bc) ? 8 : ((L->ac || L->bc) ? 4 : 2);
          const double fl = mul * L->g.M * (double)L->g.N * L->g.K * L->g.batch;
          char nm[96];
          std::snprintf(nm, sizeof nm, "gett_kernel<%d,%d,%d,%d>%s", (int)L->ac, (int)L->bc, (int)L->akc, (int)L->bkc,
                        L->g.splits > 1 ? "+splitk" : "");
          char dt[160];
          if (L->gp.dot) {
            // algorithmic bytes: each operand read once, the partials written once
            const DotArgs& d = L->dt;
            const double pa = d.a_im ? 2 : 1, pb = d.b_im ? 2 : 1;
            const double dbytes = 8.0 * ((pa * d.M + pb * d.N) * (double)d.K + pa * d.M * pb * d.N * d.splits);
            std::snprintf(dt, sizeof dt, "%s rows=%d N=%d K=%lld splits=%d", nodes_[L->node].spec.c_str(),
                          (int)(pa * d.M), (int)(pb * d.N), (long long)d.K, d.splits);
            detail = dt;
            timed("dot_kernel", fl, dbytes, [&] { launch_dot(L->dt, L->sc, s); });
            break;
          }
          if (L->gp.on && L->gp.S > 0) {
            const GemmPath& gp = L->gp;
            const double abytes = (8.0 + gp.S) * L->sa.planes * (double)L->sa.Rb * L->sa.K;
            auto split_detail = [&](const SplitArgs& x, const char* side) {
              std::snprintf(dt, sizeof dt, "%s %s R=%lld K=%lld", nodes_[L->node].spec.c_str(), side,
                            (long long)(x.planes * x.Rb), (long long)x.K);
              detail = dt;
            };
            if (gp.a_cache < 0 && gp.a_split) {
              split_detail(L->sa, "A");
              timed(gp.a_cols ? "ozaki_split_cols_kernel" : "ozaki_split_rows_kernel", 0.0, abytes,
                    [&] { launch_ozaki_split(L->sa, gp.a_cols, s); });
            }
            const double bbytes = (8.0 + gp.S) * L->sb.planes * (double)L->sb.Rb * L->sb.K;
            if (gp.b_cache < 0 && gp.b_split) {
              split_detail(L->sb, "B");
              timed(gp.b_cols ? "ozaki_split_cols_kernel" : "ozaki_split_rows_kernel", 0.0, bbytes,
                    [&] { launch_ozaki_split(L->sb, gp.b_cols, s); });
            }
            if (gp.m3) {
              // 3M: two thirds of the FP64-equivalent flops in the P1, P2 launch
              std::snprintf(dt, sizeof dt, "%s rows=%d N=%d K=%d splits=%d digits=%d 3M P1+P2",
                            nodes_[L->node].spec.c_str(), L->oz.M, L->oz.N, L->g.K, gp.splits, gp.S);
              detail = dt;
              timed(ozgemm_kernel_name(L->oz.S), fl * 2.0 / 3.0, 0.0, [&] { launch_ozgemm(L->oz, s); });
              std::snprintf(dt, sizeof dt, "%s rows=%d N=%d K=%d splits=%d digits=%d 3M P3",
                            nodes_[L->node].spec.c_str(), L->oz2.M, L->oz2.N, L->g.K, gp.splits, gp.S);
              detail = dt;
              timed(ozgemm_kernel_name(L->oz2.S), fl / 3.0, 0.0, [&] { launch_ozgemm(L->oz2, s); });
              break;
            }
            std::snprintf(dt, sizeof dt, "%s rows=%d N=%d K=%d splits=%d digits=%d", nodes_[L->node].spec.c_str(),
                          L->oz.M, L->oz.N, L->g.K, gp.splits, gp.S);
            detail = dt;
            timed(ozgemm_kernel_name(L->oz.S), fl, 0.0, [&] { launch_ozgemm(L->oz, s); });
            const double sbytes = 8.0 * L->sc.M * (double)L->sc.N * (L->sc.kind == 2 ? 4 : 2) * L->sc.splits;
            if (!gp.fused) timed("gemm_scatter_kernel", 0.0, sbytes, [&] { launch_gemm_scatter(L->sc, s); });
            break;
          }
          if (L->gp.on) {
            const double gbytes = 8.0 * ((double)L->ga.planes * L->ga.Rb * L->ga.Cb) * 2;
            timed("gather2d_kernel", 0.0, gbytes, [&] { launch_gather2d(L->ga, s); });
            const double bbytes = 8.0 * ((double)L->gb.planes * L->gb.Rb * L->gb.Cb) * 2;
            timed("gather2d_kernel", 0.0, bbytes, [&] { launch_gather2d(L->gb, s); });
            std::snprintf(dt, sizeof dt, "%s rows=%d N=%d K=%d splits=%d", nodes_[L->node].spec.c_str(), L->dg.M,
                          L->dg.N, L->dg.K, L->gp.splits);
            detail = dt;
            timed(L->gp.bt ? "dgemm_kernel<BT>" : "dgemm_kernel<NN>", fl, 0.0,
                  [&] { launch_dgemm(L->dg, L->gp.bt, s); });
            const double sbytes = 8.0 * L->sc.M * (double)L->sc.N * (L->sc.kind == 2 ? 4 : 2) * L->sc.splits;
            timed("gemm_scatter_kernel", 0.0, sbytes, [&] { launch_gemm_scatter(L->sc, s); });
            break;
          }
          if (L->skinny) {
            const SkinnyArgs& k = L->sk;
            std::snprintf(dt, sizeof dt, "%s I=%d J=%lld K=%lld kmajor=%d S=%d", nodes_[L->node].spec.c_str(), k.I,
                          (long long)k.J, (long long)k.K, (int)k.g_kmajor, k.splits);
            detail = dt;
            const double pl = (L->ac || L->bc) ? 2.0 : 1.0;
            const double bytes = 8.0 * ((L->small_is_a ? (L->ac ? 2 : 1) : (L->bc ? 2 : 1)) * (double)k.I * k.K +
                                        (L->small_is_a ? (L->bc ? 2 : 1) : (L->ac ? 2 : 1)) * (double)k.J * k.K +
                                        pl * k.I * (double)k.J);
            timed(k.run > 0 ? "skinny_runs_kernel" : (k.g_kmajor ? "skinny_k_kernel" : "skinny_rows_kernel"), fl, bytes,
                  [&] { launch_skinny(k, s); });
            break;
          }
          std::snprintf(dt, sizeof dt, "%s M=%d N=%d K=%d B=%d S=%d", nodes_[L->node].spec.c_str(), L->g.M, L->g.N,
                        L->g.K, L->g.batch, L->g.splits);
          detail = dt;
          timed(nm, fl, 0.0, [&] {
            launch_gett(L->g, L->ac, L->bc, L->akc, L->bkc, s);
            if (L->g.splits > 1) launch_gett_reduce(L->g, L->c_cplx, s);
          });
          break;
        }
        case Launch::LIN: {
          // algorithmic bytes: the output (read too when accumulating) and
          // every term at the elements the kernel sums (the canonical ones
          // for an antisymmetric result)
          const char* kname = lincomb_kernel_name(L->l);
          double summed = (double)L->l.total;
          if (std::strcmp(kname, "lincomb_anti_kernel") == 0) {
            const double d0 = (double)L->l.dims[0], d1 = (double)L->l.dims[1], d3 = (double)L->l.dims[3];
            summed = ((L->l.anti & 1) ? d0 * (d0 - 1) / 2 : d0 * d1) *
                     ((L->l.anti & 2) ? d3 * (d3 + 1) / 2 : (double)L->l.dims[2] * d3);
          }
          double term_planes = 0.0;
          for (int k = 0; k < L->l.nterms; ++k)
            if (!L->l.t[k].re2) term_planes += L->l.t[k].im ? 2.0 : 1.0;  // product terms: two small factors
          const double bytes =
              8.0 * ((double)L->l.total * (L->l.out_im ? 2 : 1) * (1 + L->l.accumulate) + summed * term_planes);
          char dt[160];
          int perm = 0;  // terms read through a non-identity permutation
          for (int k = 0; k < L->l.nterms; ++k)
            for (int d = 0; d + 1 < L->l.rank; ++d)
              if (L->l.t[k].st[d] != 0 && L->l.t[k].st[d] < L->l.t[k].st[d + 1]) {
                ++perm;
                break;
              }
          std::snprintf(dt, sizeof dt, "node %d terms=%d permuted=%d n=%lld acc=%d", L->node, L->l.nterms, perm,
                        (long long)L->l.total, L->l.accumulate);
          detail = dt;
          timed(kname, 0.0, bytes, [&] { launch_lincomb(L->l, s); });
          break;
        }
        case Launch::LADDER: {
          const LadderArgs& a = L->lad;
          const double elems = 2.0 * a.o * a.o * a.v * a.v * a.nz, packed = 2.0 * a.rows * a.nz * a.P;
          timed("pack_pairs_kernel", 0.0, 8.0 * (2.0 * a.rows * a.nz * a.v * a.v + packed), [&] { ladder_pack(a, s); });
          const DgemmArgs g = ladder_gemm(a);
          char dt[96];
          if (a.S > 0) {
            timed("ozaki_split_rows_kernel", 0.0, (8.0 + a.S) * packed, [&] { launch_ozaki_split(a.su, false, s); });
            std::snprintf(dt, sizeof dt, "packed ladder M=%d N=%d K=%d digits=%d", g.M, g.N, g.K, a.S);
            detail = dt;
            timed(ozgemm_kernel_name(a.oz.S), 2.0 * g.M * (double)g.N * g.K, 0.0, [&] { launch_ozgemm(a.oz, s); });
          } else {
            std::snprintf(dt, sizeof dt, "packed ladder M=%d N=%d K=%d", g.M, g.N, g.K);
            detail = dt;
            timed("dgemm_kernel<BT>", 2.0 * g.M * (double)g.N * g.K, 0.0, [&] { launch_dgemm(g, true, s); });
          }
          if (a.splits > 1)
            timed("sum_splits_kernel", 0.0, 8.0 * packed * (a.splits + 1),
                  [&] { launch_sum_splits(a.rp, a.splits, (int64_t)packed, (int64_t)packed, s); });
          timed("unpack_pairs_kernel", 0.0, 8.0 * (elems + packed), [&] { ladder_unpack(a, s); });
          break;
        }
        case Launch::FIELD:
          timed("field_kernel", 0.0, 8.0 * 4 * L->f.n2, [&] { launch_field(L->f, s); });
          break;
        case Launch::FIN:
          timed("finalize_kernel", 0.0, 16.0 * 7 * L->fin.n, [&] { launch_finalize(L->fin, s); });
          break;
        case Launch::TRACE:
          timed("trace_kernel", 0.0, 0.0, [&] { launch_trace(L->tr, s); });
          break;
        case Launch::FVJP:
          timed("field_vjp_kernel", 0.0, 0.0, [&] { launch_field_vjp(L->fv, s); });
          break;
        case Launch::SEED:
          timed("seed_kernel", 0.0, 0.0, [&] { launch_seed(L->sd, s); });
          break;
        case Launch::TVJP:
          timed("trace_vjp_kernel", 0.0, 0.0, [&] { launch_trace_vjp(L->tv, s); });
          break;
      }
    }
    TDCC_HIP(hipGetLastError());
    TDCC_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < recs.size(); ++i) TDCC_HIP(hipEventElapsedTime(&recs[i].ms, ev[2 * i], ev[2 * i + 1]));
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  out.insert(out.end(), recs.begin(), recs.end());
}

int64_t Graph::contract_flops() const {
  int64_t f = 0;
  for (const auto& L : launches_) {
    if (L->kind != Launch::GETT) continue;
    const int64_t mul = (L->ac && L->bc) ? 8 : ((L->ac || L->bc) ? 4 : 2);
    f += mul * (int64_t)L->g.M * L->g.N * L->g.K * L->g.batch;
  }
  const int64_t P = (int64_t)v_ * (v_ - 1) / 2;
  for (size_t i = 0; i < nodes_.size(); ++i)
    if (nodes_[i].op == Op::Ladder && i < live_.size() && live_[i])
      f += 2 * ladder_rows2(o_, nodes_[i].ij_pairs) * (nodes_[i].zdep ? nz_ : 1) * P * P;
  return f;
}

}  // namespace tdcc
