// C ABI of the propagation library (declared in include/tdcc.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tdcc.h"
#include "cc.hpp"
#include "engine.hpp"

using namespace tdcc;

namespace {

struct Planes {  // complex tensor as two planes in one allocation
  int64_t n = 0;
  double* d = nullptr;
  double* re() const { return d; }
  double* im() const { return d + n; }
};

struct Program {
  std::unique_ptr<Graph> g;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool bound = false;
  bool step = false;
  // runtime parameters the launch records were bound with
  double negA = 0, dt6 = 0, scale = 0;
  int pol = -1;
  const void* table = nullptr;
  const void* mu_out = nullptr;
  bool batched = false;
  std::vector<Runtime::Traj> batch;  // trajectory parameters the records were bound with
  const void* wtab = nullptr;
  const void* ttab = nullptr;
  const void* grad = nullptr;
};

}  // namespace

struct tdcc_ctx {
  int o = 0, v = 0, n = 0, device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  double* tei = nullptr;
  double* F = nullptr;
  double* mu = nullptr;
  bool have_tei = false, have_system = false;
  Planes state[4], base[4], backup[4], io_in[4], io_out, io_F, int_out;
  double* staging = nullptr;
  int64_t staging_n = 0;   // complex elements the staging buffer holds
  int64_t* counter = nullptr;
  double* table = nullptr;
  int64_t table_cap = 0;
  double* mu_out = nullptr;
  int64_t mu_cap = 0;
  double* arena = nullptr;
  int64_t arena_bytes = 0;
  double* vpack[2] = {nullptr, nullptr};   // packed <ab||ef>, <ef||ab> (may alias)
  int8_t* vpack8[2] = {nullptr, nullptr};  // their digit planes in split precision (alias alike)
  double* vpack8_scale[2] = {nullptr, nullptr};
  unsigned long long* counts = nullptr;    // kCountBlocks + 1 partial counts
  double* amp_check = nullptr;             // 2 kCountBlocks + 2 partial maxima (device amplitude checks)
  Runtime rt;
  std::map<std::string, Program> progs;
  // t2 and lam2 of the state (of the checkpoint) are antisymmetric in
  // (i, j): the ladders then run on the i < j rows only
  bool pairs = true, backup_pairs = true;
  // the whole integral array is antisymmetric in each index pair
  bool tei_antisym = false;
  // batched trajectories of this system (tdcc_set_batch): per-slot state and
  // rounded base, per-slot field tables and dipole rows ([slot][cap][3][2])
  int nb = 0;
  std::vector<std::array<Planes, 4>> bstate, bbase;
  // one allocation per amplitude block for all slots (state, then base):
  // slot b of block k at bblk[.][k] + b * 2 * block_size(k), so that the
  // stacked program reaches slot b at a fixed element stride
  double* bblk[2][4] = {{nullptr, nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr, nullptr}};
  std::vector<char> slot_pairs;
  double* btable = nullptr;
  double* bmu = nullptr;
  int64_t bcap = 0;
  // adjoint buffers (tdcc_dipole_gradient): state before a step, adjoint in /
  // out, per-step weights and stage times, gradient accumulators
  std::array<Planes, 4> adj_y, adj_a, adj_b;
  double* wtab = nullptr;
  double* ttab = nullptr;
  int64_t vcap = 0;
  double* grad = nullptr;  // [4] dL/dA, dL/dw0 (complex), then (n,n) re and im planes of dL/dmu_pol
  // no-allocation dipole (tdcc_dipole_device_noalloc): its own runtime (no
  // counter; the trace writes row 0 of na_mu), a program planned and bound at
  // tdcc_create with its arena in the device's static scratch array, or the
  // reason it could not be
  Runtime rt_na;
  double* na_mu = nullptr;
  std::unique_ptr<Graph> na_prog;
  std::string na_err;
};

namespace {

int64_t block_size(const tdcc_ctx* c, int k) {
  const int64_t o = c->o, v = c->v;
  return (k == 0 || k == 2) ? o * v : o * o * v * v;
}

void alloc_planes(Planes& p, int64_t n) {
  p.n = n;
  TDCC_HIP(hipMalloc(&p.d, 2 * n * sizeof(double)));
  TDCC_HIP(hipMemset(p.d, 0, 2 * n * sizeof(double)));
}

void free_planes(Planes& p) {
  if (p.d) (void)hipFree(p.d);
  p.d = nullptr;
}

void ensure_staging(tdcc_ctx* c, int64_t n) {
  if (n <= c->staging_n) return;
  TDCC_HIP(hipStreamSynchronize(c->stream));
  if (c->staging) TDCC_HIP(hipFree(c->staging));
  c->staging = nullptr;
  c->staging_n = 0;
  TDCC_HIP(hipMalloc(&c->staging, 2 * n * sizeof(double)));
  c->staging_n = n;
}

// True when x (interleaved complex, (o,o,v,v)) is antisymmetric in its two
// index pairs to 1e-12 of its largest entry, as CC amplitudes are: the
// ladders may then compute the i < j rows only and mirror them, and sums
// over an antisymmetric index pair may run over p < q, which agrees with the
// full sums to that tolerance.  TDCC_IJ_PAIRS=0 turns this off.
bool antisym_ij(const double* x, int o, int v) {
  const char* e = std::getenv("TDCC_IJ_PAIRS");
  if (e && e[0] == '0') return false;
  if (!x) return true;
  const int64_t len = 2 * (int64_t)v * v;
  double mx = 0.0, dev = 0.0;
  for (int i = 0; i < o; ++i)
    for (int j = 0; j < o; ++j) {
      const double* a = x + ((int64_t)i * o + j) * len;
      const double* b = x + ((int64_t)j * o + i) * len;
      for (int p = 0; p < v; ++p)
        for (int q = 0; q < v; ++q)
          for (int c = 0; c < 2; ++c) {
            const double xa = a[2 * ((int64_t)p * v + q) + c];
            mx = std::max(mx, std::fabs(xa));
            dev = std::max(dev, std::fabs(xa + b[2 * ((int64_t)p * v + q) + c]));
            dev = std::max(dev, std::fabs(xa + a[2 * ((int64_t)q * v + p) + c]));
          }
    }
  return dev <= 1e-12 * mx;
}

// antisym_ij for an amplitude in device memory (same bound, on the GPU).
bool antisym_ij_device(tdcc_ctx* c, const double* x) {
  const char* e = std::getenv("TDCC_IJ_PAIRS");
  if (e && e[0] == '0') return false;
  if (!x) return true;
  if (!c->amp_check) TDCC_HIP(hipMalloc(&c->amp_check, (2 * kCountBlocks + 2) * sizeof(double)));
  launch_amp_antisym_check(x, c->o, c->v, c->amp_check, c->stream);
  TDCC_HIP(hipGetLastError());
  double r[2];
  TDCC_HIP(hipMemcpyAsync(r, c->amp_check + 2 * kCountBlocks, sizeof r, hipMemcpyDeviceToHost, c->stream));
  TDCC_HIP(hipStreamSynchronize(c->stream));
  return r[1] <= 1e-12 * r[0];
}

// A pointer handed to a *_device entry point must be device (or managed)
// memory of the context's GPU.
void require_device(const tdcc_ctx* c, const void* p, const char* what) {
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    throw Error(std::string(what) + " is not device memory");
  }
  if (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)
    throw Error(std::string(what) + " is not device memory");
  if (at.device != c->device)
    throw Error(std::string(what) + " is on device " + std::to_string(at.device) + ", the context on device " +
                std::to_string(c->device));
}

// Interleaved complex device array -> the context's planes (no host copy).
void upload_complex_device(tdcc_ctx* c, const double* dev, const Planes& p) {
  launch_split_complex(dev, p.re(), p.im(), p.n, c->stream);
  TDCC_HIP(hipGetLastError());
}

void upload_complex(tdcc_ctx* c, const double* host, const Planes& p) {
  ensure_staging(c, p.n);
  TDCC_HIP(hipMemcpyAsync(c->staging, host, 2 * p.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  launch_split_complex(c->staging, p.re(), p.im(), p.n, c->stream);
  TDCC_HIP(hipGetLastError());
  TDCC_HIP(hipStreamSynchronize(c->stream));
}

void download_complex(tdcc_ctx* c, const Planes& p, double* host) {
  ensure_staging(c, p.n);
  launch_join_complex(p.re(), p.im(), c->staging, p.n, c->stream);
  TDCC_HIP(hipGetLastError());
  TDCC_HIP(hipMemcpyAsync(host, c->staging, 2 * p.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TDCC_HIP(hipStreamSynchronize(c->stream));
}

Shape shape_of_block(int k) {
  Shape s;
  if (k == 0 || k == 2) {
    s.rank = 2;
    s.cls[0] = Cls::O;
    s.cls[1] = Cls::V;
  } else {
    s.rank = 4;
    s.cls[0] = s.cls[1] = Cls::O;
    s.cls[2] = s.cls[3] = Cls::V;
  }
  return s;
}

int ext_planes(Graph& g, const Planes& p, const Shape& s, const char* name) {
  Dev d;
  d.re = p.re();
  d.im = p.im();
  int64_t acc = 1;
  for (int k = s.rank - 1; k >= 0; --k) {
    d.st[k] = acc;
    acc *= g.extent(s.cls[k]);
  }
  return g.external(s, true, d, name);
}

int ext_tei(tdcc_ctx* c, Graph& g) {
  Shape s;
  s.rank = 4;
  s.cls[0] = s.cls[1] = s.cls[2] = s.cls[3] = Cls::N;
  Dev d;
  d.re = c->tei;
  const int64_t n = c->n;
  d.st[0] = n * n * n;
  d.st[1] = n * n;
  d.st[2] = n;
  d.st[3] = 1;
  const int id = g.external(s, false, d, "tei");
  if (c->tei_antisym) g.set_anti(id, anti_bit(0, 1) | anti_bit(2, 3));
  return id;
}

// t2 and lam2 (blocks 1 and 3) of an antisymmetric state are antisymmetric
// in both index pairs.
int mark_amplitude(Graph& g, int id, int k, bool pairs) {
  if (pairs && (k == 1 || k == 3)) g.set_anti(id, anti_bit(0, 1) | anti_bit(2, 3));
  return id;
}

void invalidate_exec(Program& p) {
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  p.exec = nullptr;
  p.graph = nullptr;
  p.bound = false;
}

void ensure_arena(tdcc_ctx* c, int64_t bytes) {
  if (bytes <= c->arena_bytes) return;
  TDCC_HIP(hipStreamSynchronize(c->stream));
  for (auto& kv : c->progs) invalidate_exec(kv.second);
  if (c->arena) TDCC_HIP(hipFree(c->arena));
  c->arena = nullptr;
  c->arena_bytes = 0;
  TDCC_HIP(hipMalloc(&c->arena, bytes));
  c->arena_bytes = bytes;
}

void bind(tdcc_ctx* c, Program& p) {
  const bool stale = !p.bound || p.negA != c->rt.negA || p.dt6 != c->rt.dt6 || p.scale != c->rt.round_scale ||
                     p.pol != c->rt.pol || p.table != c->table || p.mu_out != c->mu_out ||
                     (p.batched && p.batch != c->rt.batch) || p.wtab != c->rt.wtab || p.ttab != c->rt.ttab ||
                     p.grad != c->rt.g;
  if (!stale) return;
  invalidate_exec(p);
  p.g->bind(c->arena);
  p.bound = true;
  p.negA = c->rt.negA;
  p.dt6 = c->rt.dt6;
  p.scale = c->rt.round_scale;
  p.pol = c->rt.pol;
  p.table = c->table;
  p.mu_out = c->mu_out;
  if (p.batched) p.batch = c->rt.batch;
  p.wtab = c->rt.wtab;
  p.ttab = c->rt.ttab;
  p.grad = c->rt.g;
}

Program& finish_program(tdcc_ctx* c, const std::string& key, std::unique_ptr<Graph> g, bool step) {
  g->plan();
  Program& p = c->progs[key];
  p.g = std::move(g);
  p.step = step;
  ensure_arena(c, p.g->arena_bytes());
  return p;
}

// Program key suffix of the pair shortcuts a program was built with.
std::string pairs_suffix(const tdcc_ctx* c, bool pairs) {
  return pairs ? (c->tei_antisym ? ":pk" : ":p") : ":f";
}

int build_rhs_node(CCBuilder& b, int kind, int F, int t1, int t2, int l1, int l2) {
  switch (kind) {
    case TDCC_RHS_T1: return b.T1eq_rhs(t1, t2, F);
    case TDCC_RHS_T2: return b.T2eq_rhs(t1, t2, F);
    case TDCC_RHS_L1: return b.lam_1eq_rhs(t1, t2, l1, l2, F);
    case TDCC_RHS_L2: return b.lam2eq_rhs(t1, t2, l1, l2, F);
    case TDCC_RHS_T2_CC2: return b.T2eq_rhs_CC2(t1, t2, F);
    case TDCC_RHS_L1_CC2: return b.L1eq_rhs_CC2(t1, t2, l1, l2, F);
    case TDCC_RHS_L2_CC2: return b.L2eq_rhs_CC2(t1, t2, l1, l2, F);
    default: throw Error("unknown right-hand side kind " + std::to_string(kind));
  }
}

Program& rhs_program(tdcc_ctx* c, int kind, bool pairs) {
  const std::string key = "rhs:" + std::to_string(kind) + pairs_suffix(c, pairs);
  auto it = c->progs.find(key);
  if (it != c->progs.end()) return it->second;
  auto g = std::make_unique<Graph>(c->o, c->v, &c->rt);
  g->set_ij_pairs(pairs);
  g->set_k_pairs(pairs && c->tei_antisym);
  const int tei = ext_tei(c, *g);
  Shape sF;
  sF.rank = 2;
  sF.cls[0] = sF.cls[1] = Cls::N;
  const int F = ext_planes(*g, c->io_F, sF, "F");
  int y[4];
  for (int k = 0; k < 4; ++k) y[k] = mark_amplitude(*g, ext_planes(*g, c->io_in[k], shape_of_block(k), "in"), k, pairs);
  CCBuilder b(*g, tei);
  const int r = build_rhs_node(b, kind, F, y[0], y[1], y[2], y[3]);
  const int out_k = (kind == TDCC_RHS_T1 || kind == TDCC_RHS_L1 || kind == TDCC_RHS_L1_CC2) ? 0 : 1;
  Planes out = c->io_out;
  out.n = block_size(c, out_k);
  // io_out holds the largest block; view it with the residual's shape
  Dev d;
  d.re = c->io_out.re();
  d.im = c->io_out.re() + out.n;
  const Shape so = shape_of_block(out_k);
  int64_t acc = 1;
  for (int k = so.rank - 1; k >= 0; --k) {
    d.st[k] = acc;
    acc *= g->extent(so.cls[k]);
  }
  const int dst = g->external(so, true, d, "out");
  g->assign(dst, r);
  return finish_program(c, key, std::move(g), false);
}

Program& dipole_program(tdcc_ctx* c) {
  const std::string key = "dipole";
  auto it = c->progs.find(key);
  if (it != c->progs.end()) return it->second;
  auto g = std::make_unique<Graph>(c->o, c->v, &c->rt);
  const int tei = ext_tei(c, *g);
  int y[4];
  for (int k = 0; k < 4; ++k) y[k] = ext_planes(*g, c->io_in[k], shape_of_block(k), "in");
  CCBuilder b(*g, tei);
  g->trace(b.Dij(y[0], y[1], y[2], y[3]), y[2], b.Dai(y[0], y[1], y[2], y[3]), b.Dab(y[0], y[1], y[2], y[3]));
  return finish_program(c, key, std::move(g), false);
}

int build_intermediate(CCBuilder& b, const std::string& name, int F, int t1, int t2, int l1, int l2) {
  if (name == "Fae") return b.Fae(t1, t2, F);
  if (name == "Fmi") return b.Fmi(t1, t2, F);
  if (name == "Fme") return b.Fme(t1, t2, F);
  if (name == "Wmnij") return b.Wmnij(t1, t2, F);
  if (name == "Wmnij_2") return b.Wmnij_2(t1, t2, F);
  if (name == "Wabef") return b.Wabef(t1, t2, F);
  if (name == "Wabef_2") return b.Wabef_2(t1, t2, F);
  if (name == "Wmbej") return b.Wmbej(t1, t2, F);
  if (name == "LRFea") return b.LRFea(t1, t2, F);
  if (name == "LRFim") return b.LRFim(t1, t2, F);
  if (name == "LSWieam") return b.LSWieam(t1, t2, F);
  if (name == "LRWefam") return b.LRWefam(t1, t2, F);
  if (name == "LRWibjm") return b.LRWibjm(t1, t2, F);
  if (name == "LSWmnij") return b.LSWmnij(t1, t2, F);
  if (name == "LSWabef") return b.LSWabef(t1, t2, F);
  if (name == "Gfe") return b.Gfe(t2, l2);
  if (name == "Gmn") return b.Gmn(t2, l2);
  if (name == "LWfiea") return b.LWfiea(t1);
  if (name == "LWmina") return b.LWmina(t1);
  if (name == "Dij") return b.Dij(t1, t2, l1, l2);
  if (name == "Dab") return b.Dab(t1, t2, l1, l2);
  if (name == "Dai") return b.Dai(t1, t2, l1, l2);
  throw Error("unknown intermediate " + name);
}

Program& intermediate_program(tdcc_ctx* c, const std::string& name, int64_t* out_len) {
  const std::string key = "int:" + name;
  auto g = std::make_unique<Graph>(c->o, c->v, &c->rt);
  const int tei = ext_tei(c, *g);
  Shape sF;
  sF.rank = 2;
  sF.cls[0] = sF.cls[1] = Cls::N;
  const int F = ext_planes(*g, c->io_F, sF, "F");
  int y[4];
  for (int k = 0; k < 4; ++k) y[k] = ext_planes(*g, c->io_in[k], shape_of_block(k), "in");
  CCBuilder b(*g, tei);
  const int r = build_intermediate(b, name, F, y[0], y[1], y[2], y[3]);
  const Shape so = g->node(r).shape;
  const int64_t len = g->size(so);
  if (c->int_out.n < len) {
    TDCC_HIP(hipStreamSynchronize(c->stream));
    free_planes(c->int_out);
    for (auto& kv : c->progs)
      if (kv.first.rfind("int:", 0) == 0) invalidate_exec(kv.second);
    c->progs.erase(key);
    alloc_planes(c->int_out, len);
  }
  Planes out = c->int_out;
  out.n = len;
  const int dst = ext_planes(*g, out, so, "out");
  g->assign(dst, r);
  *out_len = len;
  auto it = c->progs.find(key);
  if (it != c->progs.end()) {
    // rebuilt graph replaces the cached one (the output buffer may have moved)
    invalidate_exec(it->second);
    c->progs.erase(it);
  }
  return finish_program(c, key, std::move(g), false);
}

std::string hexd(double x) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", x);
  return buf;
}

Program& step_program(tdcc_ctx* c, int method, double dt) {
  const std::string key = "step:" + std::to_string(method) + ":" + hexd(dt) + pairs_suffix(c, c->pairs);
  auto it = c->progs.find(key);
  if (it != c->progs.end()) return it->second;
  auto g = std::make_unique<Graph>(c->o, c->v, &c->rt);
  g->set_ij_pairs(c->pairs);
  g->set_k_pairs(c->pairs && c->tei_antisym);
  StepIO io;
  io.tei = ext_tei(c, *g);
  for (int k = 0; k < 4; ++k) {
    io.y[k] = mark_amplitude(*g, ext_planes(*g, c->state[k], shape_of_block(k), "y"), k, c->pairs);
    io.base[k] = mark_amplitude(*g, ext_planes(*g, c->base[k], shape_of_block(k), "base"), k, c->pairs);
  }
  for (int k = 0; k < 4; ++k) io.ynew[k] = ext_planes(*g, c->state[k], shape_of_block(k), "ynew");
  g->set_lambda_externals({io.y[2], io.y[3], io.base[2], io.base[3]});
  build_rk4_step(*g, io, method, dt);
  return finish_program(c, key, std::move(g), true);
}

// One program advancing every batch slot by one step.  Stacked (default): the
// step graph of one trajectory with a trajectory axis over the slots
// (Graph::set_stack), so that a contraction of a shared integral with the
// slots' amplitudes is one launch over all of them and a linear combination
// covers every slot in one launch.  TDCC_STACK=0 puts the slots' step graphs
// side by side instead (integral-only nodes shared), one launch per slot.
bool stack_enabled() {
  const char* e = std::getenv("TDCC_STACK");
  return !(e && e[0] == '0');
}

Program& batch_step_program(tdcc_ctx* c, int method, double dt) {
  bool pairs = true;
  for (char f : c->slot_pairs) pairs = pairs && f;
  const bool stacked = stack_enabled();
  const std::string key = std::string(stacked ? "zstep:" : "bstep:") + std::to_string(method) + ":" + hexd(dt) + ":" +
                          std::to_string(c->nb) + pairs_suffix(c, pairs);
  auto it = c->progs.find(key);
  if (it != c->progs.end()) return it->second;
  auto g = std::make_unique<Graph>(c->o, c->v, &c->rt);
  g->set_ij_pairs(pairs);
  g->set_k_pairs(pairs && c->tei_antisym);
  const int tei = ext_tei(c, *g);
  if (stacked) {
    g->set_stack(c->nb);
    StepIO io;
    io.tei = tei;
    auto ext = [&](const Planes& p, int k, const char* name) {
      const int id = ext_planes(*g, p, shape_of_block(k), name);
      g->set_zstride(id, 2 * block_size(c, k));
      return id;
    };
    for (int k = 0; k < 4; ++k) {
      io.y[k] = mark_amplitude(*g, ext(c->bstate[0][k], k, "y"), k, pairs);
      io.base[k] = mark_amplitude(*g, ext(c->bbase[0][k], k, "base"), k, pairs);
    }
    for (int k = 0; k < 4; ++k) io.ynew[k] = ext(c->bstate[0][k], k, "ynew");
    g->set_lambda_externals({io.y[2], io.y[3], io.base[2], io.base[3]});
    build_rk4_step(*g, io, method, dt);
  } else {
    std::vector<int> lam;
    for (int b = 0; b < c->nb; ++b) {
      g->set_traj(b);
      StepIO io;
      io.tei = tei;
      for (int k = 0; k < 4; ++k) {
        io.y[k] = mark_amplitude(*g, ext_planes(*g, c->bstate[b][k], shape_of_block(k), "y"), k, pairs);
        io.base[k] = mark_amplitude(*g, ext_planes(*g, c->bbase[b][k], shape_of_block(k), "base"), k, pairs);
      }
      for (int k = 0; k < 4; ++k) io.ynew[k] = ext_planes(*g, c->bstate[b][k], shape_of_block(k), "ynew");
      for (int k = 2; k < 4; ++k) lam.insert(lam.end(), {io.y[k], io.base[k]});
      build_rk4_step(*g, io, method, dt);
    }
    g->set_traj(-1);
    g->set_lambda_externals(lam);
  }
  Program& p = finish_program(c, key, std::move(g), true);
  p.batched = true;
  return p;
}

// Adjoint programs of the windowed-dipole gradient (tdcc_dipole_gradient):
// their externals are the context's adjoint buffers.
Program& vjp_step_program(tdcc_ctx* c, int method, double dt) {
  const std::string key = "vstep:" + std::to_string(method) + ":" + hexd(dt) + (c->pairs ? ":p" : ":f");
  auto it = c->progs.find(key);
  if (it != c->progs.end()) return it->second;
  auto g = std::make_unique<Graph>(c->o, c->v, &c->rt);
  g->set_ij_pairs(c->pairs);  // forward ladders only; their adjoints run on full rows
  StepVjpIO io;
  io.tei = ext_tei(c, *g);
  for (int k = 0; k < 4; ++k) {
    io.y[k] = ext_planes(*g, c->adj_y[k], shape_of_block(k), "y");
    io.ybar_in[k] = ext_planes(*g, c->adj_a[k], shape_of_block(k), "ybar_in");
    io.ybar_out[k] = ext_planes(*g, c->adj_b[k], shape_of_block(k), "ybar_out");
  }
  build_rk4_step_vjp(*g, io, method, dt);
  return finish_program(c, key, std::move(g), false);
}

Program& vjp_readout_program(tdcc_ctx* c) {
  const std::string key = "vread";
  auto it = c->progs.find(key);
  if (it != c->progs.end()) return it->second;
  auto g = std::make_unique<Graph>(c->o, c->v, &c->rt);
  ReadoutVjpIO io;
  io.tei = ext_tei(c, *g);
  for (int k = 0; k < 4; ++k) {
    io.y[k] = ext_planes(*g, c->adj_y[k], shape_of_block(k), "y");
    io.ybar[k] = ext_planes(*g, c->adj_a[k], shape_of_block(k), "ybar");
  }
  build_readout_vjp(*g, io);
  return finish_program(c, key, std::move(g), false);
}

void copy_state(tdcc_ctx* c, const std::array<Planes, 4>& dst, const Planes* src) {
  for (int k = 0; k < 4; ++k)
    TDCC_HIP(hipMemcpyAsync(dst[k].d, src[k].d, 2 * src[k].n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
}

void run_program(tdcc_ctx* c, Program& p, bool use_graph) {
  bind(c, p);
  if (!use_graph) {
    if (p.step) launch_counter_inc(c->counter, c->stream);
    p.g->run(c->stream);
    return;
  }
  if (!p.exec) {
    TDCC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
      if (p.step) launch_counter_inc(c->counter, c->stream);
      p.g->run(c->stream);
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(c->stream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    TDCC_HIP(hipStreamEndCapture(c->stream, &p.graph));
    TDCC_HIP(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
  }
  TDCC_HIP(hipGraphLaunch(p.exec, c->stream));
}

bool graphs_enabled() {
  const char* e = std::getenv("TDCC_NO_GRAPH");
  return !(e && e[0] == '1');
}

void ensure_rows(tdcc_ctx* c, int64_t rows) {
  if (rows <= c->table_cap && rows <= c->mu_cap) return;
  TDCC_HIP(hipStreamSynchronize(c->stream));
  const int64_t cap = std::max<int64_t>(rows, 1024);
  if (c->table) TDCC_HIP(hipFree(c->table));
  if (c->mu_out) TDCC_HIP(hipFree(c->mu_out));
  c->table = c->mu_out = nullptr;
  TDCC_HIP(hipMalloc(&c->table, cap * 6 * sizeof(double)));
  TDCC_HIP(hipMalloc(&c->mu_out, cap * 6 * sizeof(double)));
  c->table_cap = c->mu_cap = cap;
  c->rt.field_table = c->table;
  c->rt.mu_out = c->mu_out;
}

template <class F>
int guarded(tdcc_ctx* c, F&& f) {
  if (!c) return 1;
  try {
    TDCC_HIP(hipSetDevice(c->device));
    f();
    c->err.clear();
    return 0;
  } catch (const std::exception& e) {
    c->err = e.what();
    return 1;
  }
}

unsigned long long device_count(tdcc_ctx* c) {
  unsigned long long n = 0;
  TDCC_HIP(hipMemcpyAsync(&n, c->counts + kCountBlocks, sizeof n, hipMemcpyDeviceToHost, c->stream));
  TDCC_HIP(hipStreamSynchronize(c->stream));
  return n;
}

void free_vpack8(tdcc_ctx* c) {
  for (int k : {1, 0}) {
    if (k == 1 && c->vpack8[1] == c->vpack8[0]) {
      c->vpack8[1] = nullptr;
      c->vpack8_scale[1] = nullptr;
    }
    if (c->vpack8[k]) TDCC_HIP(hipFree(c->vpack8[k]));
    if (c->vpack8_scale[k]) TDCC_HIP(hipFree(c->vpack8_scale[k]));
    c->vpack8[k] = nullptr;
    c->vpack8_scale[k] = nullptr;
    c->rt.vpack8[k] = nullptr;
    c->rt.vpack8_scale[k] = nullptr;
  }
}

// Digit planes of the packed integrals for split precision: rows q(ab) of
// vpack, split along q(ef), with the runtime's digit count.
void build_vpack8(tdcc_ctx* c) {
  free_vpack8(c);
  const int S = c->rt.split_digits;
  const int64_t P = (int64_t)c->v * (c->v - 1) / 2;
  if (S <= 0 || P == 0 || !c->vpack[0]) return;
  const int64_t Kp = (P + 63) / 64 * 64;
  const int64_t plane = (P * Kp + 255) / 256 * 256;
  double* pmax = nullptr;
  TDCC_HIP(hipMalloc(&pmax, ozaki_split_workspace(P, Kp, false)));
  for (int k = 0; k < 2; ++k) {
    if (k == 1 && c->vpack[1] == c->vpack[0]) {
      c->vpack8[1] = c->vpack8[0];
      c->vpack8_scale[1] = c->vpack8_scale[0];
      break;
    }
    TDCC_HIP(hipMalloc(&c->vpack8[k], S * plane));
    TDCC_HIP(hipMalloc(&c->vpack8_scale[k], P * sizeof(double)));
    SplitArgs a{};
    a.src_re = c->vpack[k];
    a.ldx = P;
    a.Rb = P;
    a.planes = 1;
    a.K = P;
    a.Kp = Kp;
    a.S = S;
    a.D = c->vpack8[k];
    a.plane = plane;
    a.scale = c->vpack8_scale[k];
    a.pmax = pmax;
    launch_ozaki_split(a, false, c->stream);
    TDCC_HIP(hipGetLastError());
  }
  TDCC_HIP(hipStreamSynchronize(c->stream));
  TDCC_HIP(hipFree(pmax));
  for (int k = 0; k < 2; ++k) {
    c->rt.vpack8[k] = c->vpack8[k];
    c->rt.vpack8_scale[k] = c->vpack8_scale[k];
  }
  c->rt.vpack_Kp = Kp;
  c->rt.vpack_plane = plane;
}

// Packed antisymmetric <ab||ef> for the particle-particle ladders.  The
// packing is exact only for antisymmetrized integrals, which is checked here.
void build_vpack(tdcc_ctx* c) {
  TDCC_HIP(hipStreamSynchronize(c->stream));
  for (auto& kv : c->progs) invalidate_exec(kv.second);
  free_vpack8(c);
  if (c->vpack[1] && c->vpack[1] != c->vpack[0]) TDCC_HIP(hipFree(c->vpack[1]));
  if (c->vpack[0]) TDCC_HIP(hipFree(c->vpack[0]));
  c->vpack[0] = c->vpack[1] = nullptr;
  c->rt.vpack[0] = c->rt.vpack[1] = nullptr;
  if (!c->counts) TDCC_HIP(hipMalloc(&c->counts, (kCountBlocks + 1) * sizeof(unsigned long long)));
  launch_vvvv_antisym_check(c->tei, c->o, c->v, c->counts, c->stream);
  TDCC_HIP(hipGetLastError());
  const unsigned long long bad = device_count(c);
  if (bad) throw Error("the <ab||ef> block is not antisymmetric (" + std::to_string(bad) + " entries); "
                       "the integrals must be antisymmetrized spin-orbital <pq||rs>");
  launch_tei_antisym_check(c->tei, c->n, c->counts, c->stream);
  TDCC_HIP(hipGetLastError());
  c->tei_antisym = device_count(c) == 0;
  const int64_t P = (int64_t)c->v * (c->v - 1) / 2;
  if (P == 0) return;
  TDCC_HIP(hipMalloc(&c->vpack[0], P * P * sizeof(double)));
  TDCC_HIP(hipMalloc(&c->vpack[1], P * P * sizeof(double)));
  launch_build_vpack(c->tei, c->o, c->v, 0, c->vpack[0], c->stream);
  launch_build_vpack(c->tei, c->o, c->v, 1, c->vpack[1], c->stream);
  launch_count_mismatch(c->vpack[0], c->vpack[1], P * P, c->counts, c->stream);
  TDCC_HIP(hipGetLastError());
  if (device_count(c) == 0) {  // <ab||ef> == <ef||ab>: one copy serves both ladders
    TDCC_HIP(hipFree(c->vpack[1]));
    c->vpack[1] = c->vpack[0];
  }
  c->rt.vpack[0] = c->vpack[0];
  c->rt.vpack[1] = c->vpack[1];
  build_vpack8(c);
}

void require_system(tdcc_ctx* c) {
  if (!c->have_system || !c->have_tei) throw Error("system not set: call tdcc_set_system (and tdcc_synth_tei when tei was NULL)");
}

// Field table, evaluated as the reference does: t advances by dt before the
// step; stage times t, t + dt/2, t + dt; exp(i * ((w0*2)*pi) * t).
void field_table(int nsteps, double t0, double dt, double w0, std::vector<double>& tab, std::vector<double>& times) {
  tab.assign((size_t)nsteps * 6, 0.0);
  times.assign(nsteps, 0.0);
  double t = t0;
  const double wpi = (w0 * 2.0) * M_PI;
  for (int i = 0; i < nsteps; ++i) {
    t += dt;
    times[i] = t;
    const double ts[3] = {t, t + dt / 2.0, t + dt};
    for (int s = 0; s < 3; ++s) {
      const double th = wpi * ts[s];
      tab[(size_t)i * 6 + 2 * s] = std::cos(th);
      tab[(size_t)i * 6 + 2 * s + 1] = std::sin(th);
    }
  }
}

void check_run_args(int method, int nsteps, int pol, int precs, double dt) {
  if (method != TDCC_CCSD && method != TDCC_CC2) throw Error("method must be TDCC_CCSD or TDCC_CC2");
  if (nsteps < 0) throw Error("nsteps must be >= 0");
  if (pol < 0 || pol > 2) throw Error("pol must be 0, 1 or 2");
  if (precs < 0 || precs > 22) throw Error("precs must be in [0, 22]");
  if (!(dt > 0.0) || !std::isfinite(dt)) throw Error("dt must be positive");
}

void free_batch(tdcc_ctx* c) {
  for (auto& blk : c->bblk)
    for (double*& p : blk) {
      if (p) (void)hipFree(p);
      p = nullptr;
    }
  c->bstate.clear();
  c->bbase.clear();
  if (c->btable) (void)hipFree(c->btable);
  if (c->bmu) (void)hipFree(c->bmu);
  c->btable = c->bmu = nullptr;
  c->bcap = 0;
  c->nb = 0;
  c->slot_pairs.clear();
  c->rt.batch.clear();
}

Program* prepare_run(tdcc_ctx* c, int method, int nsteps, double t0, double dt, double w0, double A, int pol,
                     int precs, std::vector<double>& times) {
  require_system(c);
  check_run_args(method, nsteps, pol, precs, dt);
  if (nsteps == 0) return nullptr;
  ensure_rows(c, nsteps);
  std::vector<double> tab;
  field_table(nsteps, t0, dt, w0, tab, times);
  TDCC_HIP(hipMemcpyAsync(c->table, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // the pageable source must outlive the copy
  TDCC_HIP(hipStreamSynchronize(c->stream));
  c->rt.negA = -A;
  c->rt.pol = pol;
  c->rt.dt6 = dt / 6.0;
  double scale = 1.0;
  for (int k = 0; k < precs; ++k) scale *= 10.0;
  c->rt.round_scale = scale;
  Program& p = step_program(c, method, dt);
  launch_counter_set(c->counter, 0, c->stream);
  return &p;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation dipole entry point: the arena of its program
// (density blocks, contraction temporaries, reduction partials) lives in this
// file-scope device array instead of a hipMalloc'd buffer.  Each GPU holds its
// own copy for the life of the process; the host-side name is not a device
// address, so host code uses only the pointer that hipGetSymbolAddress returns
// for the current device (scratch_ptr).  Every call on one GPU shares it:
// calls that overlap on one GPU race.
constexpr int64_t kNoAllocScratchBytes = int64_t(64) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device -> address of its scratch copy

// Address of the scratch array on `device`, which must be the current device.
void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) return it->second;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw Error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return p;
}

// The dipole program of tdcc_dipole, planned once at tdcc_create and bound to
// the static scratch (when its arena fits there).
void build_noalloc_dipole(tdcc_ctx* c) {
  c->rt_na = c->rt;
  c->rt_na.counter = nullptr;
  c->rt_na.mu_out = c->na_mu;
  c->rt_na.batch.clear();
  auto g = std::make_unique<Graph>(c->o, c->v, &c->rt_na);
  const int tei = ext_tei(c, *g);
  int y[4];
  for (int k = 0; k < 4; ++k) y[k] = ext_planes(*g, c->io_in[k], shape_of_block(k), "in");
  CCBuilder b(*g, tei);
  g->trace(b.Dij(y[0], y[1], y[2], y[3]), y[2], b.Dai(y[0], y[1], y[2], y[3]), b.Dab(y[0], y[1], y[2], y[3]));
  g->plan();
  if (g->arena_bytes() > kNoAllocScratchBytes) {
    c->na_err = "the dipole program of o=" + std::to_string(c->o) + ", v=" + std::to_string(c->v) + " needs " +
                std::to_string(g->arena_bytes()) + " bytes of scratch; the static scratch holds " +
                std::to_string(kNoAllocScratchBytes);
    return;
  }
  g->bind(static_cast<double*>(scratch_ptr(c->device)));
  c->na_prog = std::move(g);
}

}  // namespace

extern "C" {

int tdcc_create(int nocc, int nvir, int device, tdcc_ctx** out) {
  if (!out) return 1;
  *out = nullptr;
  if (nocc <= 0 || nvir <= 0) return 1;
  auto* c = new tdcc_ctx();
  c->o = nocc;
  c->v = nvir;
  c->n = nocc + nvir;
  c->device = device;
  const int rc = guarded(c, [&] {
    int count = 0;
    TDCC_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw Error("no such HIP device");
    TDCC_HIP(hipStreamCreate(&c->stream));
    const int64_t n = c->n;
    TDCC_HIP(hipMalloc(&c->F, n * n * sizeof(double)));
    TDCC_HIP(hipMalloc(&c->mu, 3 * n * n * sizeof(double)));
    TDCC_HIP(hipMalloc(&c->tei, n * n * n * n * sizeof(double)));
    for (int k = 0; k < 4; ++k) {
      alloc_planes(c->state[k], block_size(c, k));
      alloc_planes(c->base[k], block_size(c, k));
      alloc_planes(c->io_in[k], block_size(c, k));
    }
    const int64_t big = std::max<int64_t>(block_size(c, 1), n * n);
    alloc_planes(c->io_out, big);
    alloc_planes(c->io_F, n * n);
    ensure_staging(c, big);
    TDCC_HIP(hipMalloc(&c->counter, sizeof(int64_t)));
    launch_counter_set(c->counter, 0, c->stream);
    ensure_rows(c, 1024);
    c->rt.o = c->o;
    c->rt.v = c->v;
    c->rt.stream = c->stream;
    c->rt.F = c->F;
    c->rt.mu = c->mu;
    c->rt.counter = c->counter;
    c->pairs = c->backup_pairs = antisym_ij(nullptr, c->o, c->v);  // zero amplitudes
    TDCC_HIP(hipMalloc(&c->na_mu, 6 * sizeof(double)));
    build_noalloc_dipole(c);
    TDCC_HIP(hipStreamSynchronize(c->stream));
  });
  if (rc) {
    tdcc_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

void tdcc_destroy(tdcc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->progs) invalidate_exec(kv.second);
  c->progs.clear();
  c->na_prog.reset();
  if (c->na_mu) (void)hipFree(c->na_mu);
  for (int k = 0; k < 4; ++k) {
    free_planes(c->state[k]);
    free_planes(c->base[k]);
    free_planes(c->backup[k]);
    free_planes(c->io_in[k]);
  }
  free_planes(c->io_out);
  free_planes(c->io_F);
  free_planes(c->int_out);
  free_batch(c);
  for (int k = 0; k < 4; ++k) {
    free_planes(c->adj_y[k]);
    free_planes(c->adj_a[k]);
    free_planes(c->adj_b[k]);
  }
  for (double* p : {c->wtab, c->ttab, c->grad})
    if (p) (void)hipFree(p);
  for (double* p : {c->tei, c->F, c->mu, c->staging, c->table, c->mu_out, c->arena})
    if (p) (void)hipFree(p);
  if (c->vpack[1] && c->vpack[1] != c->vpack[0]) (void)hipFree(c->vpack[1]);
  if (c->vpack[0]) (void)hipFree(c->vpack[0]);
  for (int k : {1, 0}) {
    if (k == 1 && c->vpack8[1] == c->vpack8[0]) continue;
    if (c->vpack8[k]) (void)hipFree(c->vpack8[k]);
    if (c->vpack8_scale[k]) (void)hipFree(c->vpack8_scale[k]);
  }
  if (c->counts) (void)hipFree(c->counts);
  if (c->amp_check) (void)hipFree(c->amp_check);
  if (c->counter) (void)hipFree(c->counter);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* tdcc_last_error(const tdcc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int tdcc_set_system(tdcc_ctx* c, const double* tei, const double* fock, const double* mu) {
  return guarded(c, [&] {
    if (!fock || !mu) throw Error("fock and mu are required");
    const int64_t n = c->n;
    if (tei) {
      TDCC_HIP(hipMemcpy(c->tei, tei, n * n * n * n * sizeof(double), hipMemcpyHostToDevice));
      c->have_tei = false;
      build_vpack(c);
      c->have_tei = true;
    }
    TDCC_HIP(hipMemcpy(c->F, fock, n * n * sizeof(double), hipMemcpyHostToDevice));
    TDCC_HIP(hipMemcpy(c->mu, mu, 3 * n * n * sizeof(double), hipMemcpyHostToDevice));
    c->have_system = true;
  });
}

int tdcc_set_precision(tdcc_ctx* c, int digits) {
  return guarded(c, [&] {
    if (digits != 0 && (digits < 2 || digits > kOzMaxSlices))
      throw Error("digits must be 0 (FP64) or in [2, " + std::to_string(kOzMaxSlices) + "]");
    if (digits == c->rt.split_digits) return;
    TDCC_HIP(hipStreamSynchronize(c->stream));
    for (auto& kv : c->progs) invalidate_exec(kv.second);
    c->progs.clear();  // programs are planned for one precision
    c->rt.split_digits = digits;
    build_vpack8(c);
  });
}

int tdcc_synth_tei(tdcc_ctx* c, uint32_t config_id, double sigma) {
  return guarded(c, [&] {
    launch_synth_tei(c->tei, c->n, config_id, sigma, c->stream);
    TDCC_HIP(hipGetLastError());
    TDCC_HIP(hipStreamSynchronize(c->stream));
    c->have_tei = false;
    build_vpack(c);
    c->have_tei = true;
  });
}

int tdcc_set_state(tdcc_ctx* c, const double* t1, const double* t2, const double* l1, const double* l2) {
  return guarded(c, [&] {
    const double* h[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k) {
      if (!h[k]) throw Error("all four amplitude arrays are required");
      upload_complex(c, h[k], c->state[k]);
    }
    c->pairs = antisym_ij(t2, c->o, c->v) && antisym_ij(l2, c->o, c->v);
  });
}

int tdcc_set_state_device(tdcc_ctx* c, const double* t1, const double* t2, const double* l1, const double* l2) {
  return guarded(c, [&] {
    const double* d[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k) {
      if (!d[k]) throw Error("all four amplitude arrays are required");
      require_device(c, d[k], "amplitude");
    }
    const bool pairs = antisym_ij_device(c, t2) && antisym_ij_device(c, l2);
    for (int k = 0; k < 4; ++k) upload_complex_device(c, d[k], c->state[k]);
    TDCC_HIP(hipStreamSynchronize(c->stream));
    c->pairs = pairs;
  });
}

int tdcc_get_state_device(tdcc_ctx* c, int which, double* t1, double* t2, double* l1, double* l2) {
  return guarded(c, [&] {
    if (which != 0 && which != 1) throw Error("which must be 0 (state) or 1 (rounded base)");
    double* d[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k)
      if (d[k]) require_device(c, d[k], "amplitude");
    for (int k = 0; k < 4; ++k) {
      if (!d[k]) continue;
      const Planes& p = which == 0 ? c->state[k] : c->base[k];
      launch_join_complex(p.re(), p.im(), d[k], p.n, c->stream);
    }
    TDCC_HIP(hipGetLastError());
    TDCC_HIP(hipStreamSynchronize(c->stream));
  });
}

int tdcc_get_state(tdcc_ctx* c, int which, double* t1, double* t2, double* l1, double* l2) {
  return guarded(c, [&] {
    if (which != 0 && which != 1) throw Error("which must be 0 (state) or 1 (rounded base)");
    double* h[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k)
      if (h[k]) download_complex(c, which == 0 ? c->state[k] : c->base[k], h[k]);
  });
}

int tdcc_set_batch(tdcc_ctx* c, int nb) {
  return guarded(c, [&] {
    if (nb < 0) throw Error("nb must be >= 0");
    TDCC_HIP(hipStreamSynchronize(c->stream));
    for (auto it = c->progs.begin(); it != c->progs.end();) {
      if (it->second.batched) {
        invalidate_exec(it->second);
        it = c->progs.erase(it);
      } else {
        ++it;
      }
    }
    free_batch(c);
    c->bstate.resize(nb);
    c->bbase.resize(nb);
    for (int w = 0; w < 2 && nb > 0; ++w)
      for (int k = 0; k < 4; ++k) {
        const int64_t slot = 2 * block_size(c, k);
        TDCC_HIP(hipMalloc(&c->bblk[w][k], (size_t)nb * slot * sizeof(double)));
        TDCC_HIP(hipMemset(c->bblk[w][k], 0, (size_t)nb * slot * sizeof(double)));
        for (int b = 0; b < nb; ++b) {
          Planes& p = (w == 0 ? c->bstate : c->bbase)[b][k];
          p.n = block_size(c, k);
          p.d = c->bblk[w][k] + (int64_t)b * slot;
        }
      }
    c->nb = nb;
    c->slot_pairs.assign(nb, antisym_ij(nullptr, c->o, c->v));  // zero amplitudes
    c->rt.batch.assign(nb, Runtime::Traj{});
  });
}

int tdcc_set_state_slot(tdcc_ctx* c, int slot, const double* t1, const double* t2, const double* l1,
                        const double* l2) {
  return guarded(c, [&] {
    if (slot < 0 || slot >= c->nb) throw Error("no such batch slot");
    const double* h[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k) {
      if (!h[k]) throw Error("all four amplitude arrays are required");
      upload_complex(c, h[k], c->bstate[slot][k]);
    }
    c->slot_pairs[slot] = antisym_ij(t2, c->o, c->v) && antisym_ij(l2, c->o, c->v);
  });
}

int tdcc_get_state_slot(tdcc_ctx* c, int slot, int which, double* t1, double* t2, double* l1, double* l2) {
  return guarded(c, [&] {
    if (slot < 0 || slot >= c->nb) throw Error("no such batch slot");
    if (which != 0 && which != 1) throw Error("which must be 0 (state) or 1 (rounded base)");
    double* h[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k)
      if (h[k]) download_complex(c, which == 0 ? c->bstate[slot][k] : c->bbase[slot][k], h[k]);
  });
}

int tdcc_rk4_run_batch(tdcc_ctx* c, int method, int nsteps, double t0, double dt, const int* pol,
                       const double* w0, const double* A, int precs, double* mu_out, double* t_out) {
  return guarded(c, [&] {
    require_system(c);
    if (c->nb == 0) throw Error("no batch slots: call tdcc_set_batch");
    if (!pol || !w0 || !A) throw Error("pol, w0 and A arrays are required");
    for (int b = 0; b < c->nb; ++b) check_run_args(method, nsteps, pol[b], precs, dt);
    if (nsteps == 0) return;
    if (nsteps > c->bcap) {
      TDCC_HIP(hipStreamSynchronize(c->stream));
      if (c->btable) TDCC_HIP(hipFree(c->btable));
      if (c->bmu) TDCC_HIP(hipFree(c->bmu));
      c->btable = c->bmu = nullptr;
      c->bcap = 0;
      const int64_t cap = std::max<int64_t>(nsteps, 1024);
      TDCC_HIP(hipMalloc(&c->btable, (size_t)c->nb * cap * 6 * sizeof(double)));
      TDCC_HIP(hipMalloc(&c->bmu, (size_t)c->nb * cap * 6 * sizeof(double)));
      c->bcap = cap;
    }
    // one host table per trajectory, all alive until the copies are done
    std::vector<std::vector<double>> tabs(c->nb);
    std::vector<double> times;
    for (int b = 0; b < c->nb; ++b) {
      field_table(nsteps, t0, dt, w0[b], tabs[b], times);
      double* dst = c->btable + (size_t)b * c->bcap * 6;
      TDCC_HIP(hipMemcpyAsync(dst, tabs[b].data(), tabs[b].size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
      Runtime::Traj& tr = c->rt.batch[b];
      tr.pol = pol[b];
      tr.negA = -A[b];
      tr.field_table = dst;
      tr.mu_out = c->bmu + (size_t)b * c->bcap * 6;
    }
    TDCC_HIP(hipStreamSynchronize(c->stream));
    c->rt.dt6 = dt / 6.0;
    double scale = 1.0;
    for (int k = 0; k < precs; ++k) scale *= 10.0;
    c->rt.round_scale = scale;
    Program& p = batch_step_program(c, method, dt);
    launch_counter_set(c->counter, 0, c->stream);
    const bool use_graph = graphs_enabled();
    for (int i = 0; i < nsteps; ++i) run_program(c, p, use_graph);
    TDCC_HIP(hipGetLastError());
    if (mu_out)
      for (int b = 0; b < c->nb; ++b)
        TDCC_HIP(hipMemcpyAsync(mu_out + (size_t)b * nsteps * 6, c->bmu + (size_t)b * c->bcap * 6,
                                (size_t)nsteps * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TDCC_HIP(hipStreamSynchronize(c->stream));
    if (t_out) std::memcpy(t_out, times.data(), nsteps * sizeof(double));
  });
}

int tdcc_dipole_gradient(tdcc_ctx* c, int method, int nsteps, double t0, double dt, double w0, double A, int pol,
                         int precs, const double* weights, int checkpoint_every, double* loss, double* dA,
                         double* dw0, double* dmu) {
  return guarded(c, [&] {
    require_system(c);
    check_run_args(method, nsteps, pol, precs, dt);
    if (nsteps < 1) throw Error("nsteps must be >= 1");
    if (!weights) throw Error("weights are required");
    const int K = checkpoint_every > 0 ? checkpoint_every : nsteps;
    const int64_t n = c->n, n2 = n * n;
    // tables: field (as the forward run), stage times, weights
    std::vector<double> times;
    Program* fwd = prepare_run(c, method, nsteps, t0, dt, w0, A, pol, precs, times);
    if (nsteps > c->vcap) {
      if (c->wtab) TDCC_HIP(hipFree(c->wtab));
      if (c->ttab) TDCC_HIP(hipFree(c->ttab));
      c->wtab = c->ttab = nullptr;
      c->vcap = 0;
      TDCC_HIP(hipMalloc(&c->wtab, (size_t)nsteps * sizeof(double)));
      TDCC_HIP(hipMalloc(&c->ttab, (size_t)nsteps * 3 * sizeof(double)));
      c->vcap = nsteps;
    }
    std::vector<double> tt((size_t)nsteps * 3);
    for (int i = 0; i < nsteps; ++i) {
      tt[3 * i] = times[i];
      tt[3 * i + 1] = times[i] + dt / 2.0;
      tt[3 * i + 2] = times[i] + dt;
    }
    TDCC_HIP(hipMemcpyAsync(c->ttab, tt.data(), tt.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TDCC_HIP(hipMemcpyAsync(c->wtab, weights, (size_t)nsteps * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (!c->grad) TDCC_HIP(hipMalloc(&c->grad, (4 + 2 * n2) * sizeof(double)));
    for (int k = 0; k < 4; ++k) {
      if (!c->adj_y[k].d) {
        alloc_planes(c->adj_y[k], block_size(c, k));
        alloc_planes(c->adj_a[k], block_size(c, k));
        alloc_planes(c->adj_b[k], block_size(c, k));
      }
    }
    c->rt.wtab = c->wtab;
    c->rt.ttab = c->ttab;
    c->rt.g = c->grad;
    c->rt.gmu_re = c->grad + 4;
    c->rt.gmu_im = c->grad + 4 + n2;
    // checkpoints (every K steps) and one segment of states
    const int nseg = (nsteps + K - 1) / K;
    std::vector<std::array<Planes, 4>> ck(nseg), stack(K);
    auto free_all = [&] {
      for (auto& s4 : ck)
        for (auto& p : s4) free_planes(p);
      for (auto& s4 : stack)
        for (auto& p : s4) free_planes(p);
    };
    try {
      for (auto* v : {&ck, &stack})
        for (auto& s4 : *v)
          for (int k = 0; k < 4; ++k) alloc_planes(s4[k], block_size(c, k));
      const bool use_graph = graphs_enabled();
      // forward, keeping the state at every segment start
      copy_state(c, ck[0], c->state);
      for (int i = 1; i <= nsteps; ++i) {
        run_program(c, *fwd, use_graph);
        if (i % K == 0 && i < nsteps) copy_state(c, ck[i / K], c->state);
      }
      std::vector<double> mu((size_t)nsteps * 6);
      TDCC_HIP(hipMemcpyAsync(mu.data(), c->mu_out, mu.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      std::array<Planes, 4> last;
      for (int k = 0; k < 4; ++k) alloc_planes(last[k], block_size(c, k));
      copy_state(c, last, c->state);
      TDCC_HIP(hipStreamSynchronize(c->stream));
      double lr = 0.0, li = 0.0;
      for (int i = 0; i < nsteps; ++i) {
        lr += weights[i] * mu[(size_t)i * 6 + 2 * pol];
        li += weights[i] * mu[(size_t)i * 6 + 2 * pol + 1];
      }
      // backward
      TDCC_HIP(hipMemsetAsync(c->grad, 0, (4 + 2 * n2) * sizeof(double), c->stream));
      for (int k = 0; k < 4; ++k) TDCC_HIP(hipMemsetAsync(c->adj_a[k].d, 0, 2 * c->adj_a[k].n * sizeof(double), c->stream));
      Program& vstep = vjp_step_program(c, method, dt);
      Program& vread = vjp_readout_program(c);
      copy_state(c, c->adj_y, c->state);
      launch_counter_set(c->counter, nsteps, c->stream);
      run_program(c, vread, use_graph);
      for (int j = nseg - 1; j >= 0; --j) {
        const int n0 = j * K, n1 = std::min(nsteps, n0 + K);
        copy_state(c, stack[0], ck[j].data());
        copy_state(c, std::array<Planes, 4>{c->state[0], c->state[1], c->state[2], c->state[3]}, ck[j].data());
        launch_counter_set(c->counter, n0, c->stream);
        for (int m = 1; m < n1 - n0; ++m) {
          run_program(c, *fwd, use_graph);
          copy_state(c, stack[m], c->state);
        }
        for (int s = n1; s > n0; --s) {
          copy_state(c, c->adj_y, stack[s - 1 - n0].data());
          launch_counter_set(c->counter, s, c->stream);
          run_program(c, vstep, use_graph);
          copy_state(c, c->adj_a, c->adj_b.data());
          if (s - 1 >= 1) {
            launch_counter_set(c->counter, s - 1, c->stream);
            run_program(c, vread, use_graph);
          }
        }
      }
      TDCC_HIP(hipGetLastError());
      std::vector<double> gr(4 + 2 * n2);
      TDCC_HIP(hipMemcpyAsync(gr.data(), c->grad, gr.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      // leave the state at the end of the propagation, as tdcc_rk4_run does
      copy_state(c, std::array<Planes, 4>{c->state[0], c->state[1], c->state[2], c->state[3]}, last.data());
      TDCC_HIP(hipStreamSynchronize(c->stream));
      for (auto& p : last) free_planes(p);
      if (loss) {
        loss[0] = lr;
        loss[1] = li;
      }
      if (dA) {
        dA[0] = gr[0];
        dA[1] = gr[1];
      }
      if (dw0) {
        dw0[0] = gr[2];
        dw0[1] = gr[3];
      }
      if (dmu)
        for (int64_t e = 0; e < n2; ++e) {
          dmu[2 * e] = gr[4 + e];
          dmu[2 * e + 1] = gr[4 + n2 + e];
        }
    } catch (...) {
      free_all();
      throw;
    }
    free_all();
  });
}

int tdcc_rhs(tdcc_ctx* c, int kind, const double* F, const double* t1, const double* t2, const double* l1,
             const double* l2, double* out) {
  return guarded(c, [&] {
    require_system(c);
    if (!F || !t1 || !t2 || !out) throw Error("F, t1, t2 and out are required");
    const bool lam = kind == TDCC_RHS_L1 || kind == TDCC_RHS_L2 || kind == TDCC_RHS_L1_CC2 || kind == TDCC_RHS_L2_CC2;
    if (lam && (!l1 || !l2)) throw Error("lam1 and lam2 are required for a Lambda right-hand side");
    const bool pairs = antisym_ij(t2, c->o, c->v) && (!lam || antisym_ij(l2, c->o, c->v));
    Program& p = rhs_program(c, kind, pairs);
    upload_complex(c, F, c->io_F);
    const double* h[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k)
      if (h[k]) upload_complex(c, h[k], c->io_in[k]);
    run_program(c, p, false);
    TDCC_HIP(hipStreamSynchronize(c->stream));
    const int out_k = (kind == TDCC_RHS_T1 || kind == TDCC_RHS_L1 || kind == TDCC_RHS_L1_CC2) ? 0 : 1;
    Planes o = c->io_out;
    o.n = block_size(c, out_k);
    download_complex(c, o, out);
  });
}

int tdcc_rhs_device(tdcc_ctx* c, int kind, const double* F, const double* t1, const double* t2, const double* l1,
                    const double* l2, double* out) {
  return guarded(c, [&] {
    require_system(c);
    if (!F || !t1 || !t2 || !out) throw Error("F, t1, t2 and out are required");
    const bool lam = kind == TDCC_RHS_L1 || kind == TDCC_RHS_L2 || kind == TDCC_RHS_L1_CC2 || kind == TDCC_RHS_L2_CC2;
    if (lam && (!l1 || !l2)) throw Error("lam1 and lam2 are required for a Lambda right-hand side");
    const double* d[4] = {t1, t2, lam ? l1 : nullptr, lam ? l2 : nullptr};
    require_device(c, F, "F");
    require_device(c, out, "out");
    for (int k = 0; k < 4; ++k)
      if (d[k]) require_device(c, d[k], "amplitude");
    const bool pairs = antisym_ij_device(c, t2) && (!lam || antisym_ij_device(c, l2));
    Program& p = rhs_program(c, kind, pairs);
    upload_complex_device(c, F, c->io_F);
    for (int k = 0; k < 4; ++k)
      if (d[k]) upload_complex_device(c, d[k], c->io_in[k]);
    run_program(c, p, false);
    const int out_k = (kind == TDCC_RHS_T1 || kind == TDCC_RHS_L1 || kind == TDCC_RHS_L1_CC2) ? 0 : 1;
    Planes o = c->io_out;
    o.n = block_size(c, out_k);
    launch_join_complex(o.re(), o.im(), out, o.n, c->stream);
    TDCC_HIP(hipGetLastError());
    TDCC_HIP(hipStreamSynchronize(c->stream));
  });
}

int tdcc_intermediate(tdcc_ctx* c, const char* name, const double* F, const double* t1, const double* t2,
                      const double* l1, const double* l2, double* out, int64_t out_len) {
  return guarded(c, [&] {
    require_system(c);
    if (!name || !F || !t1 || !t2 || !l1 || !l2 || !out) throw Error("all arguments are required");
    int64_t len = 0;
    Program& p = intermediate_program(c, name, &len);
    if (out_len != len) throw Error("out_len does not match the intermediate's size " + std::to_string(len));
    upload_complex(c, F, c->io_F);
    const double* h[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k) upload_complex(c, h[k], c->io_in[k]);
    run_program(c, p, false);
    TDCC_HIP(hipStreamSynchronize(c->stream));
    Planes o = c->int_out;
    o.n = len;
    download_complex(c, o, out);
  });
}

int tdcc_dipole(tdcc_ctx* c, const double* t1, const double* t2, const double* l1, const double* l2, double* mu_out) {
  return guarded(c, [&] {
    require_system(c);
    if (!t1 || !t2 || !l1 || !l2 || !mu_out) throw Error("all arguments are required");
    Program& p = dipole_program(c);
    const double* h[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k) upload_complex(c, h[k], c->io_in[k]);
    launch_counter_set(c->counter, 1, c->stream);
    run_program(c, p, false);
    TDCC_HIP(hipMemcpyAsync(mu_out, c->mu_out, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TDCC_HIP(hipStreamSynchronize(c->stream));
  });
}

int tdcc_dipole_device_noalloc(tdcc_ctx* c, const double* t1, const double* t2, const double* l1, const double* l2,
                               double* mu_out) {
  return guarded(c, [&] {
    if (!c->have_system) throw Error("system not set: call tdcc_set_system");
    if (!t1 || !t2 || !l1 || !l2 || !mu_out) throw Error("all arguments are required");
    if (!c->na_prog) throw Error(c->na_err.empty() ? "no-allocation dipole program missing" : c->na_err);
    const double* d[4] = {t1, t2, l1, l2};
    for (int k = 0; k < 4; ++k) launch_split_complex(d[k], c->io_in[k].re(), c->io_in[k].im(), c->io_in[k].n, c->stream);
    c->na_prog->run(c->stream);
    TDCC_HIP(hipMemcpyAsync(mu_out, c->na_mu, 6 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    TDCC_HIP(hipGetLastError());
    TDCC_HIP(hipStreamSynchronize(c->stream));
  });
}

int tdcc_rk4_run(tdcc_ctx* c, int method, int nsteps, double t0, double dt, double w0, double A, int pol,
                 int precs, double* mu_out, double* t_out) {
  return guarded(c, [&] {
    std::vector<double> times;
    Program* p = prepare_run(c, method, nsteps, t0, dt, w0, A, pol, precs, times);
    if (!p) return;
    const bool use_graph = graphs_enabled();
    for (int i = 0; i < nsteps; ++i) run_program(c, *p, use_graph);
    TDCC_HIP(hipGetLastError());
    if (mu_out)
      TDCC_HIP(hipMemcpyAsync(mu_out, c->mu_out, (size_t)nsteps * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TDCC_HIP(hipStreamSynchronize(c->stream));
    if (t_out) std::memcpy(t_out, times.data(), nsteps * sizeof(double));
  });
}

int tdcc_profile_step(tdcc_ctx* c, int method, double t0, double dt, double w0, double A, int pol, int precs,
                      char* report, int64_t report_len) {
  return guarded(c, [&] {
    if (!report || report_len <= 0) throw Error("report buffer required");
    std::vector<double> times;
    Program* p = prepare_run(c, method, 1, t0, dt, w0, A, pol, precs, times);
    bind(c, *p);
    launch_counter_inc(c->counter, c->stream);
    std::vector<Graph::LaunchRecord> recs;
    p->g->run_profiled(c->stream, recs);
    std::string txt;
    char line[256];
    for (const auto& r : recs) {
      std::snprintf(line, sizeof line, "%s\t%.17g\t%.17g\t%.9g\t%s\n", r.name.c_str(), r.flops, r.bytes, (double)r.ms,
                    r.detail.c_str());
      txt += line;
    }
    if ((int64_t)txt.size() + 1 > report_len) throw Error("report buffer too small: need " + std::to_string(txt.size() + 1));
    std::memcpy(report, txt.c_str(), txt.size() + 1);
  });
}

int tdcc_checkpoint(tdcc_ctx* c, int op) {
  return guarded(c, [&] {
    if (op != 0 && op != 1) throw Error("op must be 0 (save) or 1 (restore)");
    for (int k = 0; k < 4; ++k) {
      if (!c->backup[k].d) {
        if (op == 1) throw Error("no checkpoint saved");
        alloc_planes(c->backup[k], c->state[k].n);
      }
      const Planes& src = op == 0 ? c->state[k] : c->backup[k];
      const Planes& dst = op == 0 ? c->backup[k] : c->state[k];
      TDCC_HIP(hipMemcpyAsync(dst.d, src.d, 2 * src.n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    if (op == 0) c->backup_pairs = c->pairs;
    else c->pairs = c->backup_pairs;
    TDCC_HIP(hipStreamSynchronize(c->stream));
  });
}

int tdcc_step_info(tdcc_ctx* c, int method, double dt, int64_t* flops, int64_t* arena_bytes) {
  return guarded(c, [&] {
    Program& p = step_program(c, method, dt);
    if (flops) *flops = p.g->contract_flops();
    if (arena_bytes) *arena_bytes = p.g->arena_bytes();
  });
}

}  // extern "C"
