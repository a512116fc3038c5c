// CCSD / CC2 right-hand sides as expression graphs.  Each builder follows one
// reference function term by term (same einsum strings, same coefficients);
// scalar factors sit in the linear combinations so that identical
// contractions are shared between functions, stages and chains.
#include "cc.hpp"

#include <cstdlib>

namespace tdcc {

// ------------------------------------------------------------------ T side
int CCBuilder::Fae(int t1, int t2, int F) {  // CCSD_Helper.py:136-147
  const int tau = L(Terms().add(t2).add(t1t1(t1)));
  return L(Terms()
               .add(fb(F, "vv"))
               .add(C("me,ma->ae", fb(F, "ov"), t1), -0.5)
               .add(C("mafe,mf->ae", blk("ovvv"), t1))
               .add(C("mnef,mnaf->ae", blk("oovv"), tau), -0.5));
}

int CCBuilder::Fmi(int t1, int t2, int F) {  // CCSD_Helper.py:150-160
  const int tau = L(Terms().add(t2).add(t1t1(t1)));
  return L(Terms()
               .add(fb(F, "oo"))
               .add(C("me,ie->mi", fb(F, "ov"), t1), 0.5)
               .add(C("mnie,ne->mi", blk("ooov"), t1))
               .add(C("mnef,inef->mi", blk("oovv"), tau), 0.5));
}

int CCBuilder::Fme(int t1, int t2, int F) {  // CCSD_Helper.py:163-170
  (void)t2;
  return L(Terms().add(fb(F, "ov")).add(C("mnef,nf->me", blk("oovv"), t1)));
}

int CCBuilder::T1eq_rhs(int t1, int t2, int F) {  // CCSD_Helper.py:177-196
  const int fae = Fae(t1, t2, F), fmi = Fmi(t1, t2, F), fme = Fme(t1, t2, F);
  return L(Terms()
               .add(fb(F, "ov"))
               .add(C("ae,ie->ia", fae, t1))
               .add(C("mi,ma->ia", fmi, t1), -1.0)
               .add(C("me,imae->ia", fme, t2))
               .add(C("naif,nf->ia", blk("ovov"), t1), -1.0)
               .add(C("nmei,mnae->ia", blk("oovo"), t2), -0.5)
               .add(C("maef,imef->ia", blk("ovvv"), t2), -0.5));
}

int CCBuilder::Wmnij(int t1, int t2, int F) {  // CCSD_Helper.py:199-210
  (void)F;
  const int tt = t1t1(t1);
  // 'ib,ja->ijab' is t1t1 with positions 2,3 exchanged
  const int tau = L(Terms().add(t2, 0.25).add(tt, 0.5).addp(tt, {0, 1, 3, 2}, -0.5));
  const int x = C("mnie,je->mnij", blk("ooov"), t1);
  return L(Terms().add(blk("oooo")).anti23(x).add(C("mnef,ijef->mnij", blk("oovv"), tau)));
}

int CCBuilder::Wmnij_2(int t1, int t2, int F) {  // CCSD_Helper.py:213-224
  (void)t2;
  (void)F;
  const int x = C("mnie,je->mnij", blk("ooov"), t1);
  const int y = C("mnef,ijef->mnij", blk("oovv"), t1t1(t1));
  return L(Terms().add(blk("oooo")).anti23(x).anti23(y, 0.25));
}

int CCBuilder::Wabef(int t1, int t2, int F) {  // CCSD_Helper.py:227-239
  (void)F;
  const int x = C("amef,mb->abef", blk("vovv"), t1);
  const int y = C("mnef,mnab->abef", blk("oovv"), t2);
  const int z = C("mnef,mnab->abef", blk("oovv"), t1t1(t1));
  return L(Terms().add(blk("vvvv")).anti01(x, -1.0).add(y, 0.25).anti01(z, 0.5));
}

int CCBuilder::Wabef_2(int t1, int t2, int F) {  // CCSD_Helper.py:242-252
  (void)t2;
  (void)F;
  const int tt = t1t1(t1);
  const int tau = L(Terms().add(tt).addp(tt, {0, 1, 3, 2}, -1.0));
  const int x = C("amef,mb->abef", blk("vovv"), t1);
  return L(Terms().add(blk("vvvv")).anti01(x, -1.0).add(C("mnef,mnab->abef", blk("oovv"), tau), 0.25));
}

int CCBuilder::Wmbej(int t1, int t2, int F) {  // CCSD_Helper.py:255-265
  (void)F;
  const int t2t1 = L(Terms().add(t2, 0.5).add(C("jf,nb->jnfb", t1, t1)));
  return L(Terms()
               .add(blk("ovvo"))
               .add(C("mnej,nb->mbej", blk("oovo"), t1), -1.0)
               .add(C("mnef,jnfb->mbej", blk("oovv"), t2t1), -1.0)
               .add(C("mbef,jf->mbej", blk("ovvv"), t1)));
}

// Wmbej (CCSD_Helper.py:255-265) stored with its axes in the order (j, b, m,
// e): the T2 ring sums it over (m, e) against t_imae, so each (j, b) row is
// contiguous along the summed letters in the ring's order (a row-oriented
// digit split, and stores of the o^3 v^3 build along its contiguous axis).
// The integral term is one permuted copy of the ovvo block, shared by the
// stages.
int CCBuilder::Wmbej_jbme(int t1, int t2) {
  const int t2t1 = L(Terms().add(t2, 0.5).add(C("jf,nb->jnfb", t1, t1)));
  const int ovvo = g_.materialize(Terms().addp(blk("ovvo"), {3, 1, 0, 2}).v);
  return L(Terms()
               .add(ovvo)
               .add(C("mnej,nb->jbme", blk("oovo"), t1), -1.0)
               .add(C("mnef,jnfb->jbme", blk("oovv"), t2t1), -1.0)
               .add(C("mbef,jf->jbme", blk("ovvv"), t1)));
}

int CCBuilder::T2eq_rhs(int t1, int t2, int F) {  // CCSD_Helper.py:268-309
  const int fae = Fae(t1, t2, F), fmi = Fmi(t1, t2, F), fme = Fme(t1, t2, F);
  // TDCC_W_JBME=0: the ring reads Wmbej in its (m, b, e, j) layout
  const char* wl = std::getenv("TDCC_W_JBME");
  const bool jbme = !wl || std::atoi(wl) != 0;
  const int wmnij = Wmnij(t1, t2, F), wmbej = jbme ? Wmbej_jbme(t1, t2) : Wmbej(t1, t2, F),
            wmnij_2 = Wmnij_2(t1, t2, F);
  const int fvv = L(Terms().add(fae).add(C("me,mb->be", fme, t1), -0.5));
  const int foo = L(Terms().add(fmi).add(C("me,je->mj", fme, t1), 0.5));
  const int tt = t1t1(t1);
  // 'ma,nb->mnab' - 'na,mb->mnab'
  const int tau = L(Terms().add(tt).addp(tt, {1, 0, 2, 3}, -1.0));
  const int ring = L(Terms()
                         .add(C(jbme ? "jbme,imae->ijab" : "mbej,imae->ijab", wmbej, t2))
                         .add(C3("mbej,ie,ma->ijab", blk("ovvo"), t1, t1), -1.0));
  // terms sharing an antisymmetrizer are summed first
  const int p23 = L(Terms().add(C("be,ijae->ijab", fvv, t2)).add(C("mbij,ma->ijab", blk("ovoo"), t1), -1.0));
  const int p01 = L(Terms().add(C("mj,imab->ijab", foo, t2), -1.0).add(C("abej,ie->ijab", blk("vvvo"), t1)));
  const int ladders = L(Terms()
                            .add(C("mnij,mnab->ijab", wmnij, t2), 0.5)
                            .add(C("mnij,mnab->ijab", wmnij_2, tau), 0.5)
                            .add(T2_vvvv_terms(t1, t2, tau)));
  return L(Terms().add(blk("oovv")).add(ladders).anti23(p23).anti01(p01).anti_both(ring));
}

// -------------------------------------------------------------- Lambda side
int CCBuilder::LRFea(int t1, int t2, int F) {  // CCSD_Helper.py:447-459
  const int tau = L(Terms().add(t2, 0.5).add(t1t1(t1)));
  return L(Terms()
               .add(fb(F, "vv"))
               .add(C("ma,me->ea", fb(F, "ov"), t1), -1.0)
               .add(C("emaf,mf->ea", blk("vovv"), t1))
               .add(C("mnaf,mnef->ea", blk("oovv"), tau), -1.0));
}

int CCBuilder::LRFim(int t1, int t2, int F) {  // CCSD_Helper.py:462-472
  const int tau = L(Terms().add(t2, 0.5).add(t1t1(t1)));
  return L(Terms()
               .add(fb(F, "oo"))
               .add(C("ie,me->im", fb(F, "ov"), t1))
               .add(C("inmf,nf->im", blk("ooov"), t1))
               .add(C("inef,mnef->im", blk("oovv"), tau)));
}

int CCBuilder::LSWieam(int t1, int t2, int F) {  // CCSD_Helper.py:475-485
  (void)F;
  const int tau = L(Terms().add(t2).add(t1t1(t1)));
  return L(Terms()
               .add(blk("ovvo"))
               .add(C("eifa,mf->ieam", blk("vovv"), t1))
               .add(C("nima,ne->ieam", blk("ooov"), t1), -1.0)
               .add(C("ijab,mjbe->ieam", blk("oovv"), tau), -1.0));
}

int CCBuilder::LRWefam(int t1, int t2, int F) {  // CCSD_Helper.py:488-505
  const int fme = Fme(t1, t2, F);
  const int wabef = LSWabef(t1, t2, F);
  const int tau = L(Terms().add(t2, 0.25).add(t1t1(t1), 0.5));
  const int ring_a = L(Terms().add(blk("ovvo"), -1.0).add(C("jnab,nmfb->jfam", blk("oovv"), t2)));
  return L(Terms()
               .add(blk("vvvo"), 0.5)
               .add(C("na,mnef->efam", fme, t2), 0.5)
               .add(C("efab,mb->efam", wabef, t1))
               .add(C("jfam,je->efam", ring_a, t1))
               .add(C("jnam,jnef->efam", blk("oovo"), tau))
               .add(C("jfab,jmeb->efam", blk("ovvv"), t2), -1.0));
}

int CCBuilder::LRWibjm(int t1, int t2, int F) {  // CCSD_Helper.py:508-524
  const int fme = Fme(t1, t2, F);
  const int wmnij = LSWmnij(t1, t2, F);
  const int tau = L(Terms().add(t2, 0.25).add(t1t1(t1), 0.5));
  const int ring_a = L(Terms().add(blk("ovvo"), -1.0).add(C("inef,nmfb->ibem", blk("oovv"), t2), -1.0));
  return L(Terms()
               .add(blk("ovoo"), -0.5)
               .add(C("ie,jmbe->ibjm", fme, t2), 0.5)
               .add(C("injm,nb->ibjm", wmnij, t1))
               .add(C("ibem,je->ibjm", ring_a, t1))
               .add(C("ibef,jmef->ibjm", blk("ovvv"), tau), -1.0)
               .add(C("inem,jneb->ibjm", blk("oovo"), t2)));
}

int CCBuilder::Gfe(int t2, int l2) { return C("mnfb,mneb->fe", l2, t2, -0.5); }  // CCSD_Helper.py:526-527
int CCBuilder::Gmn(int t2, int l2) { return C("njed,mjed->nm", l2, t2, 0.5); }   // CCSD_Helper.py:529-530

int CCBuilder::LWfiea(int t1) {  // CCSD_Helper.py:533-540
  return L(Terms().add(blk("vovv")).add(C("jiea,jf->fiea", blk("oovv"), t1), -1.0));
}

// The (v,o,v,v) LWfiea = <fi||ea> - sum_j <ji||ea> t_jf is only ever
// contracted with an (o,v) or (v,v) tensor; distributing that contraction
// over its two terms replaces the o v^3 intermediate by o^2 v^3 and o^3 v^2
// contractions of the integral blocks themselves.
//   sum_fe G_fe W_fiea = sum_fe G_fe <fi||ea> - sum_je (sum_f G_fe t_jf) <ji||ea>
int CCBuilder::gef_weifa_term(int t1, int gef) {  // 'fe,fiea->ia' of CCSD_Helper.py:578
  const int y = C("fe,jf->je", gef, t1);
  return L(Terms().add(C("fe,fiea->ia", gef, blk("vovv"))).add(C("je,jiea->ia", y, blk("oovv")), -1.0));
}

//   sum_e W_ejab l_ie = sum_e <ej||ab> l_ie - sum_m (sum_e l_ie t_me) <mj||ab>
int CCBuilder::weifa_l1_term(int t1, int l1) {  // 'ejab,ie->ijab' of CCSD_Helper.py:632, CC2_Helper.py:162
  const int z = C("ie,me->im", l1, t1);
  return L(Terms().add(C("ejab,ie->ijab", blk("vovv"), l1)).add(C("mjab,im->ijab", blk("oovv"), z), -1.0));
}

// With antisymmetric amplitudes, the part of x antisymmetric in its last two
// indices, which is all that a sum against an amplitude over them sees (and
// makes that sum run over pairs); otherwise x itself.
int CCBuilder::mn_antisym(int x) {
  if (!g_.k_pairs()) return x;
  return L(Terms().add(x, 0.5).addp(x, {0, 1, 3, 2}, -0.5));
}

int CCBuilder::LWmina(int t1) {  // CCSD_Helper.py:543-550
  return L(Terms().add(blk("ooov")).add(C("mifa,nf->mina", blk("oovv"), t1)));
}

int CCBuilder::lam_1eq_rhs(int t1, int t2, int l1, int l2, int F) {  // CCSD_Helper.py:555-581
  const int fia = Fme(t1, t2, F), fea = LRFea(t1, t2, F), fim = LRFim(t1, t2, F);
  const int wieam = LSWieam(t1, t2, F), wibjm = LRWibjm(t1, t2, F);
  const int gef = Gfe(t2, l2), gmn = Gmn(t2, l2);
  const int wmina = LWmina(t1);
  return L(Terms()
               .add(fia)
               .add(C("ea,ie->ia", fea, l1))
               .add(C("im,ma->ia", fim, l1), -1.0)
               .add(C("ieam,me->ia", wieam, l1))
               .add(L1_wefam_term(t1, t2, l2, F))
               .add(C("ibjm,jmab->ia", wibjm, l2))
               .add(gef_weifa_term(t1, gef), -1.0)
               .add(C("nm,mina->ia", gmn, wmina), -1.0));
}

int CCBuilder::LSWmnij(int t1, int t2, int F) {  // CCSD_Helper.py:586-595
  (void)F;
  const int tau = L(Terms().add(t2, 0.25).add(t1t1(t1), 0.5));
  return L(Terms()
               .add(blk("oooo"), 0.5)
               .add(C("ijme,ne->ijmn", blk("ooov"), t1))
               .add(C("ijfe,mnfe->ijmn", blk("oovv"), tau)));
}

int CCBuilder::LSWabef(int t1, int t2, int F) {  // CCSD_Helper.py:598-607
  (void)F;
  const int tau = L(Terms().add(t2, 0.25).add(t1t1(t1), 0.5));
  return L(Terms()
               .add(blk("vvvv"), 0.5)
               .add(C("emab,mf->efab", blk("vovv"), t1), -1.0)
               .add(C("nmab,nmef->efab", blk("oovv"), tau)));
}

int CCBuilder::lam2eq_rhs(int t1, int t2, int l1, int l2, int F) {  // CCSD_Helper.py:610-648
  const int feb = LRFea(t1, t2, F), fjm = LRFim(t1, t2, F);
  const int wijmn = LSWmnij(t1, t2, F), wjebm = LSWieam(t1, t2, F);
  const int wijmb = LWmina(t1), fjb = Fme(t1, t2, F);
  const int gbe = Gfe(t2, l2), gmj = Gmn(t2, l2);
  const int oovv = blk("oovv");
  const int ring = L(Terms().add(C("jebm,imae->ijab", wjebm, l2)).add(C("jb,ia->ijab", fjb, l1)));
  const int p23 = L(Terms()
                        .add(C("eb,ijae->ijab", feb, l2))
                        .add(C("ijmb,ma->ijab", wijmb, l1), -1.0)
                        .add(C("ijfb,af->ijab", oovv, gbe)));
  const int p01 = L(Terms()
                        .add(C("jm,imab->ijab", fjm, l2), -1.0)
                        .add(weifa_l1_term(t1, l1))
                        .add(C("mjab,im->ijab", oovv, gmj), -1.0));
  const int direct = L(Terms()
                           .add(oovv)
                           .add(C("ijmn,mnab->ijab", mn_antisym(wijmn), l2))
                           .add(L2_vvvv_term(t1, t2, l2)));
  return L(Terms().add(direct).anti23(p23).anti01(p01).anti_both(ring));
}

// --------------------------------------------------------------------- CC2
int CCBuilder::T2eq_rhs_CC2(int t1, int t2, int F) {  // CC2_Helper.py:12-47
  const int wmnij_2 = Wmnij_2(t1, t2, F);
  const int tt = t1t1(t1);
  const int tau = L(Terms().add(tt).addp(tt, {1, 0, 2, 3}, -1.0));
  const int ring = C3("mbej,ie,ma->ijab", blk("ovvo"), t1, t1, -1.0);
  const int p23 = L(Terms().add(C("be,ijae->ijab", fb(F, "vv"), t2)).add(C("mbij,ma->ijab", blk("ovoo"), t1), -1.0));
  const int p01 = L(Terms().add(C("mj,imab->ijab", fb(F, "oo"), t2), -1.0).add(C("abej,ie->ijab", blk("vvvo"), t1)));
  const int ladders = L(Terms()
                            .add(C("mnij,mnab->ijab", wmnij_2, tau), 0.5)
                            .add(T2_vvvv_terms_CC2(t1, tau)));
  return L(Terms().add(blk("oovv")).add(ladders).anti23(p23).anti01(p01).anti_both(ring));
}

int CCBuilder::L1eq_rhs_CC2(int t1, int t2, int l1, int l2, int F) {  // CC2_Helper.py:50-125
  const int oovv = blk("oovv");
  const int tt = t1t1(t1);
  // Wmbej, CC2_Helper.py:71-75
  const int wmbej = L(Terms()
                          .add(blk("ovvo"))
                          .add(C("mbef,jf->mbej", blk("ovvv"), t1))
                          .add(C("mnej,nb->mbej", blk("oovo"), t1), -1.0)
                          .add(C("mnef,jnfb->mbej", oovv, tt), -1.0));
  const int gbj = C("me,mjbe->bj", l1, t2, -1.0);  // :77
  // Wmnij, :86-90 (the ijmn-labelled result is added positionally)
  const int tau_as = L(Terms().add(tt).addp(tt, {0, 1, 3, 2}, -1.0));
  const int wmnij = L(Terms()
                          .add(blk("oooo"))
                          .add(C("mnie,je->mnij", blk("ooov"), t1))
                          .add(C("mnje,ie->mnij", blk("ooov"), t1), -1.0)
                          .add(C("ijef,mnfe->ijmn", oovv, tau_as), -0.5));
  // Wmbij, :93-99
  const int zmejb_t1 = L(Terms()
                             .add(C("mbei,je->mbij", blk("ovvo"), t1))
                             .add(C("mbej,ie->mbij", blk("ovvo"), t1), -1.0));
  const int wmbij = L(Terms()
                          .add(blk("ovoo"))
                          .add(C("nb,mnij->mbij", t1, wmnij), -1.0)
                          .add(C("mbef,ijef->mbij", blk("ovvv"), tau_as), 0.5)
                          .add(zmejb_t1, -1.0));
  // Wabei, :105-120
  const int zmbej = L(Terms()
                          .add(blk("ovvo"))
                          .add(C("mbef,jf->mbej", blk("ovvv"), t1))
                          .add(C("mnej,nb->mbej", blk("oovo"), t1), -0.5)
                          .add(C("mnef,jnfb->mbej", oovv, tt), -0.5));
  const int wabei = L(Terms()
                          .add(blk("vvvo"))
                          .add(C("abef,if->abei", blk("vvvv"), t1))
                          .add(C("maei,mb->abei", zmbej, t1))
                          .add(C("mbei,ma->abei", zmbej, t1), -1.0));
  return L(Terms()
               .add(Fme(t1, t2, F))
               .add(C("ea,ie->ia", LRFea(t1, t2, F), l1))
               .add(C("im,ma->ia", LRFim(t1, t2, F), l1), -1.0)
               .add(C("ijab,bj->ia", oovv, gbj))
               .add(C("ieam,me->ia", wmbej, l1))
               .add(C("iemn,mnae->ia", wmbij, l2), -0.5)
               .add(C("efam,imef->ia", wabei, l2), 0.5));
}

int CCBuilder::L2eq_rhs_CC2(int t1, int t2, int l1, int l2, int F) {  // CC2_Helper.py:128-175
  const int fvv = fb(F, "vv"), foo = fb(F, "oo");
  const int wmina = LWmina(t1), fme = Fme(t1, t2, F);
  Terms t;
  t.add(blk("oovv"));
  t.add(C("ijab,eb->ijab", l2, fvv)).add(C("ijab,eb->ijba", l2, fvv), -1.0);
  t.add(C("imab,jm->ijab", l2, foo), -1.0).add(C("imab,jm->jiab", l2, foo));
  t.add(C("ijmb,ma->ijab", wmina, l1), -1.0).add(C("ijmb,ma->ijba", wmina, l1));
  const int wl = weifa_l1_term(t1, l1);
  t.add(wl).addp(wl, {1, 0, 2, 3}, -1.0);
  const int pp = C("ia,jb->ijab", l1, fme);
  t.anti_both(pp);  // ijab - ijba - jiab + jiba of the same outer product
  return L(t);
}

// ----------------------------------------------------------------- density
int CCBuilder::Dij(int t1, int t2, int l1, int l2) {  // CCSD_Helper.py:984-988
  Graph::Fp64Scope fp64(g_);  // the dipole stays FP64 in split precision
  return L(Terms().add(C("je,ie->ij", l1, t1), -1.0).add(C("jmea,imea->ij", l2, t2), -0.5));
}

int CCBuilder::Dab(int t1, int t2, int l1, int l2) {  // CCSD_Helper.py:991-995
  Graph::Fp64Scope fp64(g_);  // the dipole stays FP64 in split precision
  return L(Terms().add(C("nb,na->ab", l1, t1)).add(C("mneb,mnea->ab", l2, t2), 0.5));
}

int CCBuilder::Dai(int t1, int t2, int l1, int l2) {  // CCSD_Helper.py:998-1006
  Graph::Fp64Scope fp64(g_);  // the dipole stays FP64 in split precision
  return L(Terms()
               .addp(t1, {1, 0})
               .add(C("me,miea->ai", l1, t2))
               .add(C3("me,ma,ie->ai", l1, t1, t1), -1.0)
               .add(C3("mnef,mnaf,ie->ai", l2, t2, t1), -0.5)
               .add(C3("mnef,inef,ma->ai", l2, t2, t1), -0.5));
}

// ------------------------------------------- reassociated v^4 contractions
// Each function returns a sum that equals, term by term, the reference's
// contraction of a (v,v,v,v) intermediate, with every piece of that
// intermediate contracted into the amplitudes first.  Only the bare
// <ab||ef> block is ever contracted at O(o^2 v^4); the t1/t2-dressed pieces
// become O(o^3 v^3), O(o^4 v^2) or O(o v^4) contractions.

int CCBuilder::T2_vvvv_terms(int t1, int t2, int tau) {
  // 0.5 sum_ef Wabef(t1,t2)_abef t2_ijef + 0.5 sum_ef Wabef_2(t1)_abef tau_ijef
  const int V = blk("vvvv"), vovv = blk("vovv"), oovv = blk("oovv");
  const int tt = t1t1(t1);
  const int u = L(Terms().add(t2).add(tau));
  // <ab||ef> part of both, on u = t2 + tau ('abef,ijef->ijab', over pairs e<f, a<b)
  (void)V;
  const int ladder = g_.ladder(u, 0);
  // -P(ab) sum_m <am||ef> t_mb  ->  sum_m t_mb (sum_ef <am||ef> u_ijef)
  const int q = C("amef,ijef->ijam", vovv, u);
  const int wq = C("ijam,mb->ijab", q, t1);
  // 1/4 sum_mn <mn||ef> t_mnab and P(ab) 1/2 sum_mn <mn||ef> t_ma t_nb, on t2
  const int g = C("mnef,ijef->ijmn", oovv, t2);
  const int y = C("ijmn,mnab->ijab", g, t2);
  // g is antisymmetric in (m, n) (from <mn||ef>), so only the (m, n)
  // antisymmetric part of t_ma t_nb, half of tau, meets it
  const int z = g_.k_pairs() ? C("ijmn,mnab->ijab", g, tau, 0.5) : C("ijmn,mnab->ijab", g, tt);
  // 1/4 sum_mn <mn||ef> tau2_mnab of Wabef_2, on tau
  const int tau2 = L(Terms().add(tt).addp(tt, {0, 1, 3, 2}, -1.0));
  const int g2 = C("mnef,ijef->ijmn", oovv, tau);
  const int y2 = C("ijmn,mnab->ijab", g2, tau2);
  if (g_.k_pairs()) {
    // tau is antisymmetric in (a, b), so anti23(z, 1/4) = 1/4 g.tau, and
    // 1/8 g.t2 + 1/4 g.tau = g.(1/8 t2 + 1/4 tau): one contraction (the
    // summation order differs from the reference's two)
    (void)y;
    (void)z;
    const int yz = C("ijmn,mnab->ijab", g, L(Terms().add(t2, 0.125).add(tau, 0.25)));
    return L(Terms().add(ladder, 0.5).anti23(wq, -0.5).add(yz).add(y2, 0.125));
  }
  return L(Terms().add(ladder, 0.5).anti23(wq, -0.5).add(y, 0.125).anti23(z, 0.25).add(y2, 0.125));
}

int CCBuilder::T2_vvvv_terms_CC2(int t1, int tau) {
  // 0.5 sum_ef Wabef_2(t1)_abef tau_ijef with tau_ijef = t_ie t_jf - t_je t_if:
  // sum_ef <ab||ef> tau_ijef = X - X(i<->j), X_ijab = sum_e t_ie sum_f <ab||ef> t_jf
  const int V = blk("vvvv"), vovv = blk("vovv"), oovv = blk("oovv");
  const int tt = t1t1(t1);
  const int bj = C("abef,jf->jabe", V, t1);
  const int x = C("jabe,ie->ijab", bj, t1);
  const int q = C("amef,ijef->ijam", vovv, tau);
  const int wq = C("ijam,mb->ijab", q, t1);
  const int tau2 = L(Terms().add(tt).addp(tt, {0, 1, 3, 2}, -1.0));
  const int g2 = C("mnef,ijef->ijmn", oovv, tau);
  const int y2 = C("ijmn,mnab->ijab", g2, tau2);
  return L(Terms().anti01(x, 0.5).anti23(wq, -0.5).add(y2, 0.125));
}

int CCBuilder::L2_vvvv_term(int t1, int t2, int l2) {
  // sum_ef LSWabef_efab lam_ijef, LSWabef = 1/2 <ef||ab> - <em||ab> t_mf + <nm||ab> tau'_nmef
  const int vovv = blk("vovv"), oovv = blk("oovv");
  // Only the part of tau antisymmetric in (e, f) meets lam_ijef; with
  // antisymmetric amplitudes that part is used, which makes h antisymmetric
  // in (n, m) too and lets both sums over those pairs run over n < m, e < f.
  const int tt = t1t1(t1);
  const int tau = g_.k_pairs() ? L(Terms().add(t2, 0.25).add(tt, 0.25).addp(tt, {0, 1, 3, 2}, -0.25))
                               : L(Terms().add(t2, 0.25).add(tt, 0.5));
  // 'efab,ijef->ijab' over pairs e<f, a<b
  const int ladder = g_.ladder(l2, 1);
  const int p = C("mf,ijef->ijem", t1, l2);
  const int h = C("nmef,ijef->ijnm", tau, l2);
  return L(Terms()
               .add(ladder, 0.5)
               .add(C("emab,ijem->ijab", vovv, p), -1.0)
               .add(C("nmab,ijnm->ijab", oovv, h)));
}

int CCBuilder::L1_wefam_term(int t1, int t2, int l2, int F) {
  // sum_efm Wefam_efam lam_imef with Wefam = LRWefam (CCSD_Helper.py:488-505)
  const int oovv = blk("oovv");
  const int fme = Fme(t1, t2, F);
  const int tau = L(Terms().add(t2, 0.25).add(t1t1(t1), 0.5));
  const int ring_a = L(Terms().add(blk("ovvo"), -1.0).add(C("jnab,nmfb->jfam", oovv, t2)));
  // 0.5 <ef||am>
  const int c1 = C("efam,imef->ia", blk("vvvo"), l2);
  // 0.5 Fme_na t_mnef
  const int k = C("mnef,imef->in", t2, l2);
  const int c2 = C("na,in->ia", fme, k);
  // sum_efm sum_b LSWabef_efab t_mb lam_imef = sum_mb t_mb W_imab with
  // W_imab = sum_ef LSWabef_efab lam_imef, the <ef||ab> term of lam2eq_rhs at
  // this lam2 (the same node the L2 chain evaluates at its first stage)
  const int c3 = C("imab,mb->ia", L2_vvvv_term(t1, t2, l2), t1);
  // ring_a_jfam t_je
  const int u = C("je,imef->ijmf", t1, l2);
  const int c4 = C("jfam,ijmf->ia", ring_a, u);
  // <jn||am> tau_jnef
  const int vv = C("jnef,imef->jnim", tau, l2);
  const int c5 = C("jnam,jnim->ia", blk("oovo"), vv);
  // -<jf||ab> t_jmeb
  const int w = C("jmeb,imef->jbif", t2, l2);
  const int c6 = C("jfab,jbif->ia", blk("ovvv"), w);
  Terms t;
  t.add(c1, 0.5).add(c2, 0.5).add(c3).add(c4).add(c5).add(c6, -1.0);
  return L(t);
}

// ------------------------------------------------------------ the RK4 step
// The 16 stage derivatives k[chain][s] of one step of the four decoupled
// chains (CCSD_Helper.py:1158-1187, CC2_Helper.py:412-433): chain c evaluates
// its right-hand side at y_c + c_s k[c][s-1] with the other blocks at y.
static void stage_derivatives(Graph& g, CCBuilder& b, const int y[4], int method, double dt, int k[4][4]) {
  const int Fs[4] = {g.field(0), g.field(1), g.field(1), g.field(2)};
  const double cs[4] = {0.0, dt / 2.0, dt / 2.0, dt};
  const int t1 = y[0], t2 = y[1], l1 = y[2], l2 = y[3];
  for (int chain = 0; chain < 4; ++chain) {
    for (int s = 0; s < 4; ++s) {
      const int y0 = y[chain];
      const int ys = s == 0 ? y0 : g.lincomb({tm(y0), tm(k[chain][s - 1], cs[s])});
      const int F = Fs[s];
      int r = -1;
      if (method == METHOD_CCSD) {
        switch (chain) {
          case 0: r = b.T1eq_rhs(ys, t2, F); break;
          case 1: r = b.T2eq_rhs(t1, ys, F); break;
          case 2: r = b.lam_1eq_rhs(t1, t2, ys, l2, F); break;
          default: r = b.lam2eq_rhs(t1, t2, l1, ys, F); break;
        }
      } else {
        switch (chain) {
          case 0: r = b.T1eq_rhs(ys, t2, F); break;  // CC2 keeps the CCSD T1 equation
          case 1: r = b.T2eq_rhs_CC2(t1, ys, F); break;
          case 2: r = b.L1eq_rhs_CC2(t1, t2, ys, l2, F); break;
          default: r = b.L2eq_rhs_CC2(t1, t2, l1, ys, F); break;
        }
      }
      k[chain][s] = r;
    }
  }
}

void build_rk4_step(Graph& g, const StepIO& io, int method, double dt) {
  CCBuilder b(g, io.tei);
  int k[4][4];
  stage_derivatives(g, b, io.y, method, dt, k);
  for (int chain = 0; chain < 4; ++chain) g.finalize_rk(io.y[chain], io.base[chain], k[chain], chain >= 2);
  const int n1 = io.ynew[0], n2 = io.ynew[1], m1 = io.ynew[2], m2 = io.ynew[3];
  g.trace(b.Dij(n1, n2, m1, m2), m1, b.Dai(n1, n2, m1, m2), b.Dab(n1, n2, m1, m2));
}

// Adjoint of one step with the rounding taken as the identity: the update
// y_c' = y_c + sigma_c dt/6 (k1 + 2 k2 + 2 k3 + k4), sigma = -i (T), +i (Lambda),
// gives ybar_c = ybar'_c + (adjoint of y_c through the 16 stage evaluations
// seeded with kbar[c][s] = sigma_c dt/6 w_s ybar'_c), and the stage fields'
// adjoints go to the runtime's gradient accumulators.
void build_rk4_step_vjp(Graph& g, const StepVjpIO& io, int method, double dt) {
  CCBuilder b(g, io.tei);
  int k[4][4];
  stage_derivatives(g, b, io.y, method, dt, k);
  const double w[4] = {1.0, 2.0, 2.0, 1.0};
  std::vector<std::pair<int, Term>> seeds;
  for (int chain = 0; chain < 4; ++chain) {
    const double sig = chain < 2 ? -1.0 : 1.0;
    for (int s = 0; s < 4; ++s) seeds.push_back({k[chain][s], Term{io.ybar_in[chain], {0, 1, 2, 3}, 0.0, sig * dt / 6.0 * w[s]}});
  }
  const Graph::VjpResult r = g.vjp(seeds);
  for (int chain = 0; chain < 4; ++chain) {
    Terms t;
    t.add(io.ybar_in[chain]);
    auto it = r.ext.find(io.y[chain]);
    if (it != r.ext.end()) t.add(it->second);
    g.assign(io.ybar_out[chain], b.L(t));
  }
  for (const auto& fv : r.field_views) g.field_vjp(fv.second, fv.first);
}

// Adjoint of the weighted read-out w mu_pol . D(y) (CCSD_Helper.py:1066-1092)
// added into ybar, and its direct term w D into the dipole-matrix gradient.
void build_readout_vjp(Graph& g, const ReadoutVjpIO& io) {
  CCBuilder b(g, io.tei);
  const int t1 = io.y[0], t2 = io.y[1], l1 = io.y[2], l2 = io.y[3];
  const int dij = b.Dij(t1, t2, l1, l2), dai = b.Dai(t1, t2, l1, l2), dab = b.Dab(t1, t2, l1, l2);
  g.trace_vjp(dij, l1, dai, dab);
  const int sij = g.seed(Cls::O, Cls::O), sov = g.seed(Cls::O, Cls::V), sai = g.seed(Cls::V, Cls::O),
            sab = g.seed(Cls::V, Cls::V);
  const Graph::VjpResult r = g.vjp({{dij, tm(sij)}, {dai, tm(sai)}, {dab, tm(sab)}});
  for (int c = 0; c < 4; ++c) {
    Terms t;
    t.add(io.ybar[c]);
    auto it = r.ext.find(io.y[c]);
    if (it != r.ext.end()) t.add(it->second);
    if (c == 2) t.add(sov);  // D[o][v] = lam1
    g.assign(io.ybar[c], b.L(t));
  }
}

}  // namespace tdcc
