// Builders of the CCSD / CC2 right-hand sides, their intermediates, the
// correlated density and the RK4 step as expression graphs.
#pragma once

#include <initializer_list>
#include <vector>

#include "engine.hpp"

namespace tdcc {

// A growable list of lincomb terms with the antisymmetrizers the reference
// writes as `x - x.swapaxes(..)`.
struct Terms {
  std::vector<Term> v;
  Terms& add(int x, double c = 1.0) { v.push_back(tm(x, c)); return *this; }
  Terms& addp(int x, std::initializer_list<int> p, double c = 1.0) { v.push_back(tmp(x, p, c)); return *this; }
  // c * (x - x with positions 2,3 exchanged)
  Terms& anti23(int x, double c = 1.0) { add(x, c); return addp(x, {0, 1, 3, 2}, -c); }
  // c * (x - x with positions 0,1 exchanged)
  Terms& anti01(int x, double c = 1.0) { add(x, c); return addp(x, {1, 0, 2, 3}, -c); }
  // c * (x - x(2,3) - x(0,1) + x(0,1)(2,3))
  Terms& anti_both(int x, double c = 1.0) {
    add(x, c);
    addp(x, {0, 1, 3, 2}, -c);
    addp(x, {1, 0, 2, 3}, -c);
    return addp(x, {1, 0, 3, 2}, c);
  }
};

class CCBuilder {
 public:
  CCBuilder(Graph& g, int tei) : g_(g), tei_(tei) {}

  int blk(const char* c) { return g_.block(tei_, c); }
  int fb(int F, const char* c) { return g_.fblock(F, c); }
  int C(const char* spec, int a, int b, double alpha = 1.0) { return g_.contract(spec, a, b, alpha); }
  int C3(const char* spec, int a, int b, int c, double alpha = 1.0) { return g_.contract3(spec, a, b, c, alpha); }
  int L(const Terms& t) { return g_.lincomb(t.v); }
  int t1t1(int t1) { return C("ia,jb->ijab", t1, t1); }

  // CCSD T side (CCSD_Helper.py:136-309)
  int Fae(int t1, int t2, int F);
  int Fmi(int t1, int t2, int F);
  int Fme(int t1, int t2, int F);
  int T1eq_rhs(int t1, int t2, int F);
  int Wmnij(int t1, int t2, int F);
  int Wmnij_2(int t1, int t2, int F);
  int Wabef(int t1, int t2, int F);
  int Wabef_2(int t1, int t2, int F);
  int Wmbej(int t1, int t2, int F);
  int Wmbej_jbme(int t1, int t2);  // Wmbej laid out (j, b, m, e) for the T2 ring
  int T2eq_rhs(int t1, int t2, int F);
  // CCSD Lambda side (CCSD_Helper.py:447-648)
  int LRFea(int t1, int t2, int F);
  int LRFim(int t1, int t2, int F);
  int LSWieam(int t1, int t2, int F);
  int LRWefam(int t1, int t2, int F);
  int LRWibjm(int t1, int t2, int F);
  int Gfe(int t2, int l2);
  int Gmn(int t2, int l2);
  int LWfiea(int t1);
  int LWmina(int t1);
  int gef_weifa_term(int t1, int gef);  // Gfe . LWfiea, distributed
  int mn_antisym(int x);                 // antisymmetric part in the last index pair
  int weifa_l1_term(int t1, int l1);    // LWfiea . lam1, distributed
  int lam_1eq_rhs(int t1, int t2, int l1, int l2, int F);
  int LSWmnij(int t1, int t2, int F);
  int LSWabef(int t1, int t2, int F);
  int lam2eq_rhs(int t1, int t2, int l1, int l2, int F);
  // CC2 (CC2_Helper.py:12-175)
  int T2eq_rhs_CC2(int t1, int t2, int F);
  int L1eq_rhs_CC2(int t1, int t2, int l1, int l2, int F);
  int L2eq_rhs_CC2(int t1, int t2, int l1, int l2, int F);
  // density blocks (CCSD_Helper.py:984-1006)
  int Dij(int t1, int t2, int l1, int l2);
  int Dab(int t1, int t2, int l1, int l2);
  int Dai(int t1, int t2, int l1, int l2);

  // Reassociated forms of the terms that contract a v^4 intermediate
  // (mathematically identical sums; no complex v^4 tensor is formed).
  // 0.5 Wabef.t2 + 0.5 Wabef_2.tau of T2eq_rhs (CCSD_Helper.py:293, :295)
  int T2_vvvv_terms(int t1, int t2, int tau);
  // 0.5 Wabef_2.tau of T2eq_rhs_CC2 (CC2_Helper.py:34)
  int T2_vvvv_terms_CC2(int t1, int tau);
  // LSWabef.lam2 of lam2eq_rhs (CCSD_Helper.py:631)
  int L2_vvvv_term(int t1, int t2, int l2);
  // Wefam.lam2 of lam_1eq_rhs (CCSD_Helper.py:576), Wefam = LRWefam
  int L1_wefam_term(int t1, int t2, int l2, int F);

 private:
  Graph& g_;
  int tei_;
};

enum RhsKind { RHS_T1 = 0, RHS_T2 = 1, RHS_L1 = 2, RHS_L2 = 3, RHS_T2_CC2 = 4, RHS_L1_CC2 = 5, RHS_L2_CC2 = 6 };
enum Method { METHOD_CCSD = 0, METHOD_CC2 = 1 };

// Externals of a step program: state (updated in place), rounded base.
struct StepIO {
  int tei;
  int y[4];      // t1, t2, l1, l2 (step-start values)
  int base[4];   // rounded base written by the step
  int ynew[4];   // aliases of y used after the update (dipole)
};

// One RK4 step of the reference's propagator: four decoupled partial-RK4
// chains, fused finalize (rounding + update) per block, then the dipole trace.
void build_rk4_step(Graph& g, const StepIO& io, int method, double dt);

// Adjoint programs of the windowed-dipole gradient.
struct StepVjpIO {
  int tei;
  int y[4];          // state before the step
  int ybar_in[4];    // adjoint of the state after the step
  int ybar_out[4];   // written: adjoint of the state before the step
};
void build_rk4_step_vjp(Graph& g, const StepVjpIO& io, int method, double dt);

struct ReadoutVjpIO {
  int tei;
  int y[4];       // state whose dipole is read out
  int ybar[4];    // adjoint accumulated in place
};
void build_readout_vjp(Graph& g, const ReadoutVjpIO& io);

}  // namespace tdcc
