// Device kernels for the TDCC propagation library (gfx950 / CDNA4).
#include "kernels.hpp"

#include <cstdio>
#include <utility>
#include <cstdlib>

namespace tdcc {

typedef double d4 __attribute__((ext_vector_type(4)));

// ===========================================================================
// GETT: strided tensor contraction on v_mfma_f64_16x16x4_f64.
//
// Workgroup tile 64x64 (M x N), K step 16, 256 threads = 4 waves in a 2x2
// arrangement; each wave owns a 32x32 sub-tile = 2x2 MFMA blocks of 16x16.
// Complex operands are FP64 planes; a complex product uses 4 MFMAs
// (re*re - im*im, re*im + im*re), real x complex uses 2.
//
// MFMA 16x16x4 f64 operand layout (lane l of the wave):
//   A[i = l%16][k = l/16], B[k = l/16][j = l%16],
//   D[i = l/16 + 4*r][j = l%16] for the four accumulator elements r
//   (rows interleaved by 4; measured on gfx950, unlike the f32 variant).
// ===========================================================================
constexpr int BM = 64, BN = 64, BKK = 16, NT = 256;
constexpr int LDA = BM + 1, LDB = BN + 1;

template <bool AC, bool BC, bool AKC, bool BKC>
__global__ __launch_bounds__(NT) void gett_kernel(GettArgs g) {
  constexpr int PA = AC ? 2 : 1;
  constexpr int PB = BC ? 2 : 1;
  __shared__ double As[2][PA][BKK][LDA];
  __shared__ double Bs[2][PB][BKK][LDB];

  const int tiles = g.tiles_m * g.tiles_n;
  const int L = remap_tile(blockIdx.x, tiles);
  const int tm = L % g.tiles_m, tn = L / g.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int split = blockIdx.y, bz = blockIdx.z;
  const int kb = split * g.k_chunk;
  const int ke = min(g.K, kb + g.k_chunk);

  const int64_t aoff = g.ab ? g.ab[bz] : 0;
  const int64_t boff = g.bb ? g.bb[bz] : 0;
  const double* __restrict__ a_re = g.a_re + aoff;
  const double* __restrict__ a_im = AC ? g.a_im + aoff : nullptr;
  const double* __restrict__ b_re = g.b_re + boff;
  const double* __restrict__ b_im = BC ? g.b_im + boff : nullptr;

  const int tid = threadIdx.x;

  // ---- loader state -------------------------------------------------------
  // A tile element (mm, kk); fixed per-thread m offsets, per-step k offsets.
  int64_t a_mo[4];
  int a_lm[4], a_lk[4];
  if (AKC) {
    const int kk = tid & 15, mr = tid >> 4;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int m = m0 + mr + 16 * p;
      a_mo[p] = (m < g.M) ? g.am[m] : -1;
      a_lm[p] = mr + 16 * p;
      a_lk[p] = kk;
    }
  } else {
    // kr is the wave index, uniform across the wave: this operand's k
    // offsets are then read with scalar loads
    const int mm = tid & 63, kr = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int m = m0 + mm;
    const int64_t mo = (m < g.M) ? g.am[m] : -1;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      a_mo[p] = mo;
      a_lm[p] = mm;
      a_lk[p] = kr + 4 * p;
    }
  }
  int64_t b_no[4];
  int b_ln[4], b_lk[4];
  if (BKC) {
    const int kk = tid & 15, nr = tid >> 4;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int n = n0 + nr + 16 * p;
      b_no[p] = (n < g.N) ? g.bn[n] : -1;
      b_ln[p] = nr + 16 * p;
      b_lk[p] = kk;
    }
  } else {
    const int nn = tid & 63, kr = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n = n0 + nn;
    const int64_t no = (n < g.N) ? g.bn[n] : -1;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      b_no[p] = no;
      b_ln[p] = nn;
      b_lk[p] = kr + 4 * p;
    }
  }

  double ra[PA][4], rb[PB][4];

  auto load_global = [&](int k0) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int k = k0 + a_lk[p];
      const bool ok = (a_mo[p] >= 0) && (k < ke);
      int64_t off;
      if constexpr (AKC) {
        off = ok ? a_mo[p] + (g.ak_step >= 0 ? (int64_t)k * g.ak_step : g.ak[k]) : 0;
      } else {
        // uniform across the wave: a scalar load
        const int64_t kofs = k < ke ? (g.ak_step >= 0 ? (int64_t)k * g.ak_step : g.ak[k]) : 0;
        off = ok ? a_mo[p] + kofs : 0;
      }
      ra[0][p] = ok ? a_re[off] : 0.0;
      if (AC) ra[PA - 1][p] = ok ? a_im[off] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int k = k0 + b_lk[p];
      const bool ok = (b_no[p] >= 0) && (k < ke);
      int64_t off;
      if constexpr (BKC) {
        off = ok ? b_no[p] + (g.bk_step >= 0 ? (int64_t)k * g.bk_step : g.bk[k]) : 0;
      } else {
        const int64_t kofs = k < ke ? (g.bk_step >= 0 ? (int64_t)k * g.bk_step : g.bk[k]) : 0;
        off = ok ? b_no[p] + kofs : 0;
      }
      rb[0][p] = ok ? b_re[off] : 0.0;
      if (BC) rb[PB - 1][p] = ok ? b_im[off] : 0.0;
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
#pragma unroll
      for (int q = 0; q < PA; ++q) As[buf][q][a_lk[p]][a_lm[p]] = ra[q][p];
#pragma unroll
      for (int q = 0; q < PB; ++q) Bs[buf][q][b_lk[p]][b_ln[p]] = rb[q][p];
    }
  };

  // ---- accumulators -------------------------------------------------------
  d4 acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_re[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      acc_im[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }

  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;

  if (kb < ke) {
    load_global(kb);
    store_lds(0);
  }
  __syncthreads();
  int buf = 0;
  for (int k0 = kb; k0 < ke; k0 += BKK) {
    const bool more = (k0 + BKK) < ke;
    if (more) load_global(k0 + BKK);
#pragma unroll
    for (int k4 = 0; k4 < BKK / 4; ++k4) {
      const int kr = k4 * 4 + lk;
      double a[2][PA], b[2][PB];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < PA; ++q) a[i][q] = As[buf][q][kr][wm + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < PB; ++q) b[j][q] = Bs[buf][q][kr][wn + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], b[j][0], acc_re[i][j], 0, 0, 0);
          if (AC && BC) {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][1], -b[j][1], acc_re[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], b[j][1], acc_im[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][1], b[j][0], acc_im[i][j], 0, 0, 0);
          } else if (BC) {
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], b[j][PB - 1], acc_im[i][j], 0, 0, 0);
          } else if (AC) {
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][PA - 1], b[j][0], acc_im[i][j], 0, 0, 0);
          }
        }
    }
    if (more) store_lds(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // ---- epilogue -----------------------------------------------------------
  constexpr bool CC = AC || BC;
  if (g.splits > 1) {
    const int64_t plane = (int64_t)g.splits * g.batch * g.M * g.N;
    double* P = g.partial + ((int64_t)split * g.batch + bz) * g.M * (int64_t)g.N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = n0 + wn + j * 16 + lr;
        if (col >= g.N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + i * 16 + lk + 4 * r;
          if (row >= g.M) continue;
          const int64_t o = (int64_t)row * g.N + col;
          P[o] = acc_re[i][j][r];
          if (CC) P[plane + o] = acc_im[i][j][r];
        }
      }
    return;
  }
  const int64_t coff = g.cb ? g.cb[bz] : 0;
  double* __restrict__ c_re = g.c_re + coff;
  double* __restrict__ c_im = g.c_im ? g.c_im + coff : nullptr;
  const double ar = g.alpha_re, ai = g.alpha_im;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = n0 + wn + j * 16 + lr;
    if (col >= g.N) continue;
    const int64_t cno = g.cn[col];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + lk + 4 * r;
        if (row >= g.M) continue;
        const int64_t o = g.cm[row] + cno;
        const double xr = acc_re[i][j][r];
        const double xi = CC ? acc_im[i][j][r] : 0.0;
        double vr = ar * xr - ai * xi;
        double vi = ar * xi + ai * xr;
        if (g.accumulate) {
          vr += c_re[o];
          if (c_im) vi += c_im[o];
        }
        c_re[o] = vr;
        if (c_im) c_im[o] = vi;
      }
  }
}

__global__ void gett_reduce_kernel(GettArgs g, int c_cplx) {
  const int64_t mn = (int64_t)g.M * g.N;
  const int64_t total = mn * g.batch;
  const int64_t plane = (int64_t)g.splits * total;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t bz = idx / mn;
    const int64_t rem = idx - bz * mn;
    const int row = (int)(rem / g.N), col = (int)(rem - (int64_t)(rem / g.N) * g.N);
    double xr = 0.0, xi = 0.0;
    for (int s = 0; s < g.splits; ++s) {
      const int64_t o = (int64_t)s * total + idx;
      xr += g.partial[o];
      if (c_cplx) xi += g.partial[plane + o];
    }
    const int64_t o = (g.cb ? g.cb[bz] : 0) + g.cm[row] + g.cn[col];
    double vr = g.alpha_re * xr - g.alpha_im * xi;
    double vi = g.alpha_re * xi + g.alpha_im * xr;
    if (g.accumulate) {
      vr += g.c_re[o];
      if (g.c_im) vi += g.c_im[o];
    }
    g.c_re[o] = vr;
    if (g.c_im) g.c_im[o] = vi;
  }
}

template <bool AC, bool BC>
static void launch_gett_ab(const GettArgs& g, bool akc, bool bkc, hipStream_t s) {
  dim3 grid(g.tiles_m * g.tiles_n, g.splits, g.batch);
  if (akc && bkc) hipLaunchKernelGGL((gett_kernel<AC, BC, true, true>), grid, dim3(NT), 0, s, g);
  else if (akc) hipLaunchKernelGGL((gett_kernel<AC, BC, true, false>), grid, dim3(NT), 0, s, g);
  else if (bkc) hipLaunchKernelGGL((gett_kernel<AC, BC, false, true>), grid, dim3(NT), 0, s, g);
  else hipLaunchKernelGGL((gett_kernel<AC, BC, false, false>), grid, dim3(NT), 0, s, g);
}

void launch_gett(const GettArgs& g, bool a_cplx, bool b_cplx, bool akc, bool bkc, hipStream_t s) {
  if (a_cplx && b_cplx) launch_gett_ab<true, true>(g, akc, bkc, s);
  else if (a_cplx) launch_gett_ab<true, false>(g, akc, bkc, s);
  else if (b_cplx) launch_gett_ab<false, true>(g, akc, bkc, s);
  else launch_gett_ab<false, false>(g, akc, bkc, s);
}

void launch_gett_reduce(const GettArgs& g, bool c_cplx, hipStream_t s) {
  const int64_t total = (int64_t)g.M * g.N * g.batch;
  int blocks = (int)((total + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(gett_reduce_kernel, dim3(blocks), dim3(256), 0, s, g, c_cplx ? 1 : 0);
}

// ===========================================================================
// Linear combination of permuted tensors.  __dmul_rn/__dadd_rn keep every
// product and sum individually rounded (no contraction into FMA), which is
// how numpy evaluates `y + c * k`.
// ===========================================================================
// Value of term T at source offsets so (and so2 for a product term): x, or
// x * y for a product term, each product and sum individually rounded.
__device__ __forceinline__ void lincomb_term_value(const LincombTerm& T, int64_t so, int64_t so2, double& xr,
                                                   double& xi) {
  xr = T.re[so];
  xi = T.im ? T.im[so] : 0.0;
  if (T.re2) {
    const double yr = T.re2[so2], yi = T.im2 ? T.im2[so2] : 0.0;
    const double pr = __dsub_rn(__dmul_rn(xr, yr), __dmul_rn(xi, yi));
    const double pi = __dadd_rn(__dmul_rn(xr, yi), __dmul_rn(xi, yr));
    xr = pr;
    xi = pi;
  }
}

// Index decomposition in 32-bit arithmetic when the element count allows
// (64-bit division is a long instruction sequence on the GPU).
template <bool I32>
__global__ void lincomb_kernel(LincombArgs a) {
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < a.total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t ix[4] = {0, 0, 0, 0};
    if (I32) {
      uint32_t rem = (uint32_t)idx;
      for (int d = a.rank - 1; d >= 0; --d) {
        const uint32_t n = (uint32_t)a.dims[d], q = rem / n;
        ix[d] = rem - q * n;
        rem = q;
      }
    } else {
      int64_t rem = idx;
      for (int d = a.rank - 1; d >= 0; --d) {
        const int64_t q = rem / a.dims[d];
        ix[d] = rem - q * a.dims[d];
        rem = q;
      }
    }
    const int64_t zo = blockIdx.y;
    int64_t oo = zo * a.ozs;
    for (int d = 0; d < a.rank; ++d) oo += ix[d] * a.ost[d];
    double sr = 0.0, si = 0.0;
    if (a.accumulate) {
      sr = a.out_re[oo];
      if (a.out_im) si = a.out_im[oo];
    }
    for (int t = 0; t < a.nterms; ++t) {
      const LincombTerm& T = a.t[t];
      int64_t so = zo * T.zs, so2 = zo * T.zs2;
      for (int d = 0; d < a.rank; ++d) {
        so += ix[d] * T.st[d];
        so2 += ix[d] * T.st2[d];
      }
      double xr, xi;
      lincomb_term_value(T, so, so2, xr, xi);
      double pr, pi;
      if (T.ci == 0.0) {
        pr = __dmul_rn(T.cr, xr);
        pi = __dmul_rn(T.cr, xi);
      } else {
        pr = __dsub_rn(__dmul_rn(T.cr, xr), __dmul_rn(T.ci, xi));
        pi = __dadd_rn(__dmul_rn(T.cr, xi), __dmul_rn(T.ci, xr));
      }
      if (t == 0 && !a.accumulate) {
        sr = pr;
        si = pi;
      } else {
        sr = __dadd_rn(sr, pr);
        si = __dadd_rn(si, pi);
      }
    }
    a.out_re[oo] = sr;
    if (a.out_im) a.out_im[oo] = si;
  }
}

// Every plain term laid out like the contiguous output (no permutation):
// element e of each such operand sits at offset e, so a thread handles two
// consecutive elements with 16-byte accesses.  Product terms (x * y of two
// small permuted operands, such as t1 t1 formed in place) are read element by
// element at the offsets of e's index tuple (decomposed in 32-bit arithmetic;
// e and e + 1 share every index but the innermost, whose extent is even).
// Same arithmetic and term order as lincomb_kernel.
template <bool PROD>
__global__ __launch_bounds__(256) void lincomb_flat_kernel(LincombArgs a) {
  const int64_t n2 = a.total >> 1;
  const int64_t zo = blockIdx.y, oz = zo * a.ozs;
  const int R = a.rank;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q <= n2; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = 2 * q;
    if (e >= a.total) break;
    const bool pair = e + 1 < a.total;
    uint32_t ix[4] = {0, 0, 0, 0};
    if (PROD) {
      uint32_t rem = (uint32_t)e;
      for (int d = R - 1; d >= 0; --d) {
        const uint32_t n = (uint32_t)a.dims[d], qd = rem / n;
        ix[d] = rem - qd * n;
        rem = qd;
      }
    }
    double sr[2] = {0.0, 0.0}, si[2] = {0.0, 0.0};
    if (a.accumulate) {
      sr[0] = a.out_re[oz + e];
      if (pair) sr[1] = a.out_re[oz + e + 1];
      if (a.out_im) {
        si[0] = a.out_im[oz + e];
        if (pair) si[1] = a.out_im[oz + e + 1];
      }
    }
    for (int t = 0; t < a.nterms; ++t) {
      const LincombTerm& T = a.t[t];
      double xr[2], xi[2] = {0.0, 0.0};
      if (PROD && T.re2) {
        int64_t so = zo * T.zs, so2 = zo * T.zs2;
        for (int d = 0; d < R; ++d) {
          so += (int64_t)ix[d] * T.st[d];
          so2 += (int64_t)ix[d] * T.st2[d];
        }
        lincomb_term_value(T, so, so2, xr[0], xi[0]);
        if (pair) lincomb_term_value(T, so + T.st[R - 1], so2 + T.st2[R - 1], xr[1], xi[1]);
        else xr[1] = xi[1] = 0.0;
      } else {
        // a term laid out like the output, or (with the index tuple) one whose
        // rows are contiguous along the output's innermost axis at other strides
        int64_t te = zo * T.zs + e;
        if (PROD && T.rowwise) {
          te = zo * T.zs;
          for (int d = 0; d < R; ++d) te += (int64_t)ix[d] * T.st[d];
        }
        if (pair) {
          const double2 v = *reinterpret_cast<const double2*>(T.re + te);
          xr[0] = v.x;
          xr[1] = v.y;
          if (T.im) {
            const double2 w = *reinterpret_cast<const double2*>(T.im + te);
            xi[0] = w.x;
            xi[1] = w.y;
          }
        } else {
          xr[0] = T.re[te];
          xr[1] = 0.0;
          if (T.im) xi[0] = T.im[te];
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double pr, pi;
        if (T.ci == 0.0) {
          pr = __dmul_rn(T.cr, xr[h]);
          pi = __dmul_rn(T.cr, xi[h]);
        } else {
          pr = __dsub_rn(__dmul_rn(T.cr, xr[h]), __dmul_rn(T.ci, xi[h]));
          pi = __dadd_rn(__dmul_rn(T.cr, xi[h]), __dmul_rn(T.ci, xr[h]));
        }
        if (t == 0 && !a.accumulate) {
          sr[h] = pr;
          si[h] = pi;
        } else {
          sr[h] = __dadd_rn(sr[h], pr);
          si[h] = __dadd_rn(si[h], pi);
        }
      }
    }
    if (pair) {
      *reinterpret_cast<double2*>(a.out_re + oz + e) = make_double2(sr[0], sr[1]);
      if (a.out_im) *reinterpret_cast<double2*>(a.out_im + oz + e) = make_double2(si[0], si[1]);
    } else {
      a.out_re[oz + e] = sr[0];
      if (a.out_im) a.out_im[oz + e] = si[0];
    }
  }
}

// Whether every plain operand of `a` is laid out like its contiguous output,
// with 16-byte aligned planes (product terms, read element by element, only
// with an even innermost extent and fewer than 2^31 elements).
// TDCC_LINCOMB_FLAT_PROD=0: linear combinations with product terms on the
// element-wise kernel (a parity-test mode).
static bool lincomb_flat_prod_on() {
  const char* e = std::getenv("TDCC_LINCOMB_FLAT_PROD");
  return !e || std::atoi(e) != 0;
}

// A plain term not laid out like the output but contiguous along its innermost
// axis with every other stride even (16-byte aligned pairs), read at the
// offsets of the index tuple (e.g. an integral block viewed in the n^4 array).
static bool lincomb_term_rowwise(const LincombArgs& a, const LincombTerm& T) {
  if (T.re2 || a.rank < 1 || a.dims[a.rank - 1] % 2 || a.total >= ((int64_t)1 << 31) || !lincomb_flat_prod_on())
    return false;
  bool same = true;
  for (int d = 0; d < a.rank; ++d) same = same && (T.st[d] == a.ost[d] || a.dims[d] == 1);
  if (same) return false;
  if (T.st[a.rank - 1] != 1) return false;
  for (int d = 0; d < a.rank - 1; ++d)
    if (T.st[d] & 1) return false;
  return true;
}

// Product terms or row-wise terms: the flat kernel decomposes each element's
// index tuple.
static bool lincomb_has_prod(const LincombArgs& a) {
  for (int t = 0; t < a.nterms; ++t)
    if (a.t[t].re2 || lincomb_term_rowwise(a, a.t[t])) return true;
  return false;
}

static bool lincomb_is_flat(const LincombArgs& a) {
  auto al16 = [](const double* p) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  if (!al16(a.out_re) || !al16(a.out_im)) return false;
  int64_t st = 1;
  for (int d = a.rank - 1; d >= 0; --d) {
    if (a.ost[d] != st && a.dims[d] != 1) return false;
    st *= a.dims[d];
  }
  if (a.nz > 1 && (a.ozs & 1)) return false;
  if (lincomb_has_prod(a) && ((a.dims[a.rank - 1] & 1) || a.total >= ((int64_t)1 << 31) || !lincomb_flat_prod_on()))
    return false;
  for (int t = 0; t < a.nterms; ++t) {
    const LincombTerm& T = a.t[t];
    if (T.re2) continue;
    if (!al16(T.re) || !al16(T.im) || (a.nz > 1 && (T.zs & 1))) return false;
    if (!lincomb_term_rowwise(a, T)) {
      for (int d = 0; d < a.rank; ++d)
        if (T.st[d] != a.ost[d] && a.dims[d] != 1) return false;
    }
  }
  return true;
}

static int grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (int)b;
}

// Tiled variant for outputs whose innermost axis is contiguous while some
// term reads it with a stride and has its own contiguous axis on the output's
// second-innermost axis (a transposed pair of trailing indices, e.g. 'ijba'
// summed into 'ijab').  A block owns one 32 x 32 tile of the two trailing axes
// at one outer index; such a term is loaded along its contiguous axis into
// LDS and read back transposed, so every global access is coalesced.  The
// arithmetic per element is that of lincomb_kernel, term by term in order.
constexpr int kLinTile = 32;

__device__ __forceinline__ bool lincomb_transposed(const LincombTerm& T, int R) {
  return !T.re2 && T.st[R - 1] != 1 && T.st[R - 2] == 1;
}

// Adds the terms of `a` over the 32 x 32 tile (r0.., c0..) of the two
// trailing axes at the leading index whose source offsets are sbase[t] into
// sr/si (element k of a thread: row r0 + ly + 8k, column c0 + lx), term by
// term in order with lincomb_kernel's arithmetic; sr/si hold the output's
// old value when a.accumulate.  tr/ti are the block's staging tiles.
__device__ __forceinline__ void lincomb_tile_terms(const LincombArgs& a, const int64_t* lead, int nlead, int64_t r0,
                                                   int64_t c0, int64_t D2, int64_t D3,
                                                   double (*tr)[kLinTile + 1], double (*ti)[kLinTile + 1],
                                                   double* sr, double* si) {
  const int R = a.rank;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  const int64_t zo = blockIdx.y;
  for (int t = 0; t < a.nterms; ++t) {
    const LincombTerm& T = a.t[t];
    int64_t sbase = zo * T.zs, sbase2 = zo * T.zs2;
    for (int d = 0; d < nlead; ++d) {
      sbase += lead[d] * T.st[d];
      sbase2 += lead[d] * T.st2[d];
    }
    double xr[4], xi[4];
    if (lincomb_transposed(T, R)) {
      // stage: thread (lx, ly) loads trailing-axis element (r0 + lx, c0 + ly + 8k)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t r = r0 + lx, c = c0 + ly + 8 * k;
        const bool ok = r < D2 && c < D3;
        const int64_t so = sbase + r + c * T.st[R - 1];
        tr[ly + 8 * k][lx] = ok ? T.re[so] : 0.0;
        ti[ly + 8 * k][lx] = (ok && T.im) ? T.im[so] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        xr[k] = tr[lx][ly + 8 * k];
        xi[k] = ti[lx][ly + 8 * k];
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t r = r0 + ly + 8 * k, c = c0 + lx;
        const bool ok = r < D2 && c < D3;
        xr[k] = 0.0;
        xi[k] = 0.0;
        if (ok)
          lincomb_term_value(T, sbase + r * T.st[R - 2] + c * T.st[R - 1],
                             sbase2 + r * T.st2[R - 2] + c * T.st2[R - 1], xr[k], xi[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double pr, pi;
      if (T.ci == 0.0) {
        pr = __dmul_rn(T.cr, xr[k]);
        pi = __dmul_rn(T.cr, xi[k]);
      } else {
        pr = __dsub_rn(__dmul_rn(T.cr, xr[k]), __dmul_rn(T.ci, xi[k]));
        pi = __dadd_rn(__dmul_rn(T.cr, xi[k]), __dmul_rn(T.ci, xr[k]));
      }
      if (t == 0 && !a.accumulate) {
        sr[k] = pr;
        si[k] = pi;
      } else {
        sr[k] = __dadd_rn(sr[k], pr);
        si[k] = __dadd_rn(si[k], pi);
      }
    }
  }
}

__global__ __launch_bounds__(256) void lincomb_tiled_kernel(LincombArgs a) {
  __shared__ double tr[kLinTile][kLinTile + 1];
  __shared__ double ti[kLinTile][kLinTile + 1];
  const int R = a.rank;
  const int64_t D2 = a.dims[R - 2], D3 = a.dims[R - 1];
  const int64_t n2 = (D2 + kLinTile - 1) / kLinTile, n3 = (D3 + kLinTile - 1) / kLinTile;
  const int64_t b = blockIdx.x;
  const int64_t outer = b / (n2 * n3), tile = b - outer * (n2 * n3);
  const int64_t r0 = (tile / n3) * kLinTile, c0 = (tile % n3) * kLinTile;
  // outer index -> leading axes (rank 4: (i0, i1); rank 3: i0; rank 2: none)
  int64_t lead[2] = {0, 0};
  if (R == 4) {
    lead[0] = outer / a.dims[1];
    lead[1] = outer - lead[0] * a.dims[1];
  } else if (R == 3) {
    lead[0] = outer;
  }
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  int64_t obase = (int64_t)blockIdx.y * a.ozs;
  for (int d = 0; d < R - 2; ++d) obase += lead[d] * a.ost[d];
  double sr[4], si[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    sr[k] = 0.0;
    si[k] = 0.0;
    const int64_t r = r0 + ly + 8 * k, c = c0 + lx;
    if (a.accumulate && r < D2 && c < D3) {
      const int64_t oo = obase + r * a.ost[R - 2] + c;
      sr[k] = a.out_re[oo];
      if (a.out_im) si[k] = a.out_im[oo];
    }
  }
  lincomb_tile_terms(a, lead, R - 2, r0, c0, D2, D3, tr, ti, sr, si);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t r = r0 + ly + 8 * k, c = c0 + lx;
    if (r >= D2 || c >= D3) continue;
    const int64_t oo = obase + r * a.ost[R - 2] + c;
    a.out_re[oo] = sr[k];
    if (a.out_im) a.out_im[oo] = si[k];
  }
}

// p-th pair (r, c) of {0..n-1} with r < c (strict) or r <= c, row by row.
__device__ __forceinline__ void lincomb_tri_index(int64_t p, int64_t n, bool strict, int64_t& r, int64_t& c) {
  int64_t row = 0, len = strict ? n - 1 : n;
  while (p >= len) {
    p -= len;
    ++row;
    --len;
  }
  r = row;
  c = row + (strict ? 1 : 0) + p;
}

// Rank-4 results antisymmetric in axes (0, 1) and/or (2, 3), innermost axis
// contiguous.  A block owns one leading pair (i, j), i < j when axes 0, 1 are
// antisymmetric, and one 32 x 32 tile (A, B) of the trailing axes, A <= B when
// axes 2, 3 are; it sums the terms over the tile as lincomb_tiled_kernel does
// and stores the tile with its mirror images x[j][i] = -x[i][j] and
// x[..][b][a] = -x[..][a][b].  A diagonal tile takes its lower triangle from
// the upper one and stores zeros on the diagonal; the blocks past the last
// pair store the zero slices x[i][i].  Each source is read at the stored
// canonical elements only, so a term and its permuted partners (x, x with
// i, j exchanged, ...) read the source once between them.
__global__ __launch_bounds__(256) void lincomb_anti_kernel(LincombArgs a, int64_t npairs, int64_t ntiles) {
  __shared__ double tr[kLinTile][kLinTile + 1];
  __shared__ double ti[kLinTile][kLinTile + 1];
  const bool aij = (a.anti & 1) != 0, aab = (a.anti & 2) != 0;
  const int64_t D0 = a.dims[0], D1 = a.dims[1], D2 = a.dims[2], D3 = a.dims[3];
  const int64_t n3 = (D3 + kLinTile - 1) / kLinTile;
  const int64_t b = blockIdx.x;
  const int64_t pidx = b / ntiles, tidx = b - pidx * ntiles;
  int64_t tA, tB;
  if (aab) {
    lincomb_tri_index(tidx, n3, false, tA, tB);
  } else {
    tA = tidx / n3;
    tB = tidx - tA * n3;
  }
  const int64_t r0 = tA * kLinTile, c0 = tB * kLinTile;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  const bool diag_tile = aab && tA == tB;
  const int64_t os2 = a.ost[2];
  const int64_t oz = (int64_t)blockIdx.y * a.ozs;
  if (pidx >= npairs) {  // zero slice x[i][i]
    const int64_t i = pidx - npairs;
    const int64_t ob = oz + i * (a.ost[0] + a.ost[1]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t r = r0 + ly + 8 * k, c = c0 + lx;
      if (r < D2 && c < D3) {
        a.out_re[ob + r * os2 + c] = 0.0;
        if (a.out_im) a.out_im[ob + r * os2 + c] = 0.0;
      }
      const int64_t r2 = c0 + ly + 8 * k, c2 = r0 + lx;
      if (aab && !diag_tile && r2 < D2 && c2 < D3) {
        a.out_re[ob + r2 * os2 + c2] = 0.0;
        if (a.out_im) a.out_im[ob + r2 * os2 + c2] = 0.0;
      }
    }
    return;
  }
  int64_t lead[2];
  if (aij) {
    lincomb_tri_index(pidx, D0, true, lead[0], lead[1]);
  } else {
    lead[0] = pidx / D1;
    lead[1] = pidx - lead[0] * D1;
  }
  const int64_t ob = oz + lead[0] * a.ost[0] + lead[1] * a.ost[1];
  const int64_t obm = oz + lead[1] * a.ost[0] + lead[0] * a.ost[1];  // x[j][i] (aij)
  double sr[4], si[4];
  lincomb_tile_terms(a, lead, 2, r0, c0, D2, D3, tr, ti, sr, si);
  if (diag_tile) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      tr[ly + 8 * k][lx] = sr[k];
      ti[ly + 8 * k][lx] = si[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int rl = ly + 8 * k;
      if (rl > lx) {
        sr[k] = -tr[lx][rl];
        si[k] = -ti[lx][rl];
      } else if (rl == lx) {
        sr[k] = 0.0;
        si[k] = 0.0;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t r = r0 + ly + 8 * k, c = c0 + lx;
    if (r >= D2 || c >= D3) continue;
    a.out_re[ob + r * os2 + c] = sr[k];
    if (a.out_im) a.out_im[ob + r * os2 + c] = si[k];
    if (aij) {
      a.out_re[obm + r * os2 + c] = -sr[k];
      if (a.out_im) a.out_im[obm + r * os2 + c] = -si[k];
    }
  }
  if (aab && !diag_tile) {
    // x[..][c][r] = -x[..][r][c]: thread (lx, ly) stores row c0 + ly + 8k,
    // column r0 + lx, i.e. minus tile element (lx, ly + 8k)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      tr[ly + 8 * k][lx] = sr[k];
      ti[ly + 8 * k][lx] = si[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t r = c0 + ly + 8 * k, c = r0 + lx;
      if (r >= D2 || c >= D3) continue;
      const double vr = tr[lx][ly + 8 * k], vi = ti[lx][ly + 8 * k];
      a.out_re[ob + r * os2 + c] = -vr;
      if (a.out_im) a.out_im[ob + r * os2 + c] = -vi;
      if (aij) {
        a.out_re[obm + r * os2 + c] = vr;
        if (a.out_im) a.out_im[obm + r * os2 + c] = vi;
      }
    }
  }
}

// Whether the antisymmetric variant applies: rank 4, innermost output axis
// contiguous, the antisymmetric axes of equal extent, one launch.
static bool lincomb_use_anti(const LincombArgs& a) {
  if (a.anti == 0 || a.rank != 4 || a.accumulate || a.ost[3] != 1) return false;
  if ((a.anti & 1) && a.dims[0] != a.dims[1]) return false;
  if ((a.anti & 2) && a.dims[2] != a.dims[3]) return false;
  return true;
}

// Whether the tiled variant applies (and pays) for these arguments.
static bool lincomb_use_tiled(const LincombArgs& a) {
  if (a.rank < 2 || a.ost[a.rank - 1] != 1) return false;
  for (int t = 0; t < a.nterms; ++t)
    if (!a.t[t].re2 && a.t[t].st[a.rank - 1] != 1 && a.t[t].st[a.rank - 2] == 1) return true;
  return false;
}

// The arguments with the index axes reordered so that the tiled kernel
// applies: a term read along a leading axis d (unit stride there, not along
// the output's innermost axis) has axis d exchanged with axis rank - 2.  The
// sum is taken over the same index tuples, term by term in the same order, so
// every element gets the same value; only the order in which blocks visit the
// elements changes.  False when no term is read that way.
// TDCC_LINCOMB_SWAP=0: no reordering (such sums on the element-wise kernel, a
// parity-test mode).
static bool lincomb_swap_on() {
  const char* e = std::getenv("TDCC_LINCOMB_SWAP");
  return !e || std::atoi(e) != 0;
}
static bool lincomb_tiled_view(const LincombArgs& a, LincombArgs& b) {
  b = a;
  if (lincomb_use_tiled(a)) return true;
  const int R = a.rank;
  if (R < 3 || a.ost[R - 1] != 1 || !lincomb_swap_on()) return false;
  for (int t = 0; t < a.nterms; ++t) {
    const LincombTerm& T = a.t[t];
    if (T.re2 || T.st[R - 1] == 1) continue;
    for (int d = 0; d < R - 2; ++d) {
      if (T.st[d] != 1) continue;
      std::swap(b.dims[d], b.dims[R - 2]);
      std::swap(b.ost[d], b.ost[R - 2]);
      for (int u = 0; u < b.nterms; ++u) {
        std::swap(b.t[u].st[d], b.t[u].st[R - 2]);
        std::swap(b.t[u].st2[d], b.t[u].st2[R - 2]);
      }
      return true;
    }
  }
  return false;
}

void launch_lincomb(const LincombArgs& a, hipStream_t s) {
  if (a.total <= 0) return;
  const unsigned nz = (unsigned)(a.nz < 1 ? 1 : a.nz);
  if (lincomb_use_anti(a)) {
    const int64_t n2 = (a.dims[2] + kLinTile - 1) / kLinTile, n3 = (a.dims[3] + kLinTile - 1) / kLinTile;
    const int64_t npairs = (a.anti & 1) ? a.dims[0] * (a.dims[0] - 1) / 2 : a.dims[0] * a.dims[1];
    const int64_t ntiles = (a.anti & 2) ? n3 * (n3 + 1) / 2 : n2 * n3;
    const int64_t blocks = (npairs + ((a.anti & 1) ? a.dims[0] : 0)) * ntiles;
    if (blocks > 0 && blocks < (int64_t)1 << 31) {
      hipLaunchKernelGGL(lincomb_anti_kernel, dim3((unsigned)blocks, nz), dim3(256), 0, s, a, npairs, ntiles);
      return;
    }
  }
  if (lincomb_is_flat(a)) {
    if (lincomb_has_prod(a)) {
      LincombArgs b = a;
      for (int t = 0; t < b.nterms; ++t) b.t[t].rowwise = lincomb_term_rowwise(a, a.t[t]);
      hipLaunchKernelGGL(lincomb_flat_kernel<true>, dim3(grid_for((a.total + 1) / 2), nz), dim3(256), 0, s, b);
    } else {
      hipLaunchKernelGGL(lincomb_flat_kernel<false>, dim3(grid_for((a.total + 1) / 2), nz), dim3(256), 0, s, a);
    }
    return;
  }
  LincombArgs v;
  if (lincomb_tiled_view(a, v)) {
    const int R = v.rank;
    const int64_t n2 = (v.dims[R - 2] + kLinTile - 1) / kLinTile, n3 = (v.dims[R - 1] + kLinTile - 1) / kLinTile;
    int64_t outer = 1;
    for (int d = 0; d < R - 2; ++d) outer *= v.dims[d];
    const int64_t blocks = outer * n2 * n3;
    if (blocks < (int64_t)1 << 31) {
      hipLaunchKernelGGL(lincomb_tiled_kernel, dim3((unsigned)blocks, nz), dim3(256), 0, s, v);
      return;
    }
  }
  if (a.total < ((int64_t)1 << 31)) hipLaunchKernelGGL(lincomb_kernel<true>, dim3(grid_for(a.total), nz), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(lincomb_kernel<false>, dim3(grid_for(a.total), nz), dim3(256), 0, s, a);
}

const char* lincomb_kernel_name(const LincombArgs& a) {
  if (a.total > 0 && lincomb_use_anti(a)) return "lincomb_anti_kernel";
  if (a.total > 0 && lincomb_is_flat(a)) return lincomb_has_prod(a) ? "lincomb_flat_kernel<prod>" : "lincomb_flat_kernel";
  LincombArgs v;
  if (a.total > 0 && lincomb_tiled_view(a, v)) return "lincomb_tiled_kernel";
  return "lincomb_kernel";
}

// ===========================================================================
// Field + Fock build (FP64): bitwise the same operations as F + V(t) in the
// reference, V(t) = (-A * mu) * exp(i theta).
// ===========================================================================
__global__ void field_kernel(FieldArgs a) {
  const int64_t row = *a.counter - 1;
  const double c = a.table[(row * 3 + a.stage) * 2 + 0];
  const double s = a.table[(row * 3 + a.stage) * 2 + 1];
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < a.n2;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const double x = __dmul_rn(a.negA, a.mu_pol[idx]);
    a.out_re[idx] = __dadd_rn(a.F[idx], __dmul_rn(x, c));
    a.out_im[idx] = __dmul_rn(x, s);
  }
}

void launch_field(const FieldArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(field_kernel, dim3(grid_for(a.n2)), dim3(256), 0, s, a);
}

// ===========================================================================
// Fused RK4 finalize: rounded base + (-/+ i) dt/6 (k1 + 2k2 + 2k3 + k4).
// ===========================================================================
__global__ void finalize_kernel(FinalizeArgs a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double yr = a.y_re[i], yi = a.y_im[i];
    const double br = __ddiv_rn(rint(__dmul_rn(yr, a.scale)), a.scale);
    const double bi = __ddiv_rn(rint(__dmul_rn(yi, a.scale)), a.scale);
    double sr = __dadd_rn(a.k_re[0][i], __dmul_rn(2.0, a.k_re[1][i]));
    double si = __dadd_rn(a.k_im[0][i], __dmul_rn(2.0, a.k_im[1][i]));
    sr = __dadd_rn(sr, __dmul_rn(2.0, a.k_re[2][i]));
    si = __dadd_rn(si, __dmul_rn(2.0, a.k_im[2][i]));
    sr = __dadd_rn(sr, a.k_re[3][i]);
    si = __dadd_rn(si, a.k_im[3][i]);
    const double pr = __dmul_rn(a.dt6, sr);
    const double pi = __dmul_rn(a.dt6, si);
    // -1j * (pr + i pi) = pi - i pr ;  +1j * (pr + i pi) = -pi + i pr
    const double ur = a.lambda_sign ? -pi : pi;
    const double ui = a.lambda_sign ? pr : -pr;
    a.y_re[i] = __dadd_rn(br, ur);
    a.y_im[i] = __dadd_rn(bi, ui);
    if (a.base_re) {
      a.base_re[i] = br;
      a.base_im[i] = bi;
    }
  }
}

void launch_finalize(const FinalizeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(finalize_kernel, dim3(grid_for(a.n)), dim3(256), 0, s, a);
}

// ===========================================================================
// Dipole trace over the assembled correlated density, one workgroup,
// fixed-order tree reduction.
// ===========================================================================
__global__ __launch_bounds__(1024) void trace_kernel(TraceArgs a) {
  __shared__ double red[6][1024];
  const int n = a.o + a.v;
  const int64_t n2 = (int64_t)n * n;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t idx = threadIdx.x; idx < n2; idx += blockDim.x) {
    const int p = (int)(idx / n), q = (int)(idx - (int64_t)p * n);
    double dr, di;
    if (p < a.o && q < a.o) {
      const int64_t k = (int64_t)p * a.o + q;
      dr = a.dij_re[k]; di = a.dij_im[k];
    } else if (p < a.o) {
      const int64_t k = (int64_t)p * a.v + (q - a.o);
      dr = a.l1_re[k]; di = a.l1_im[k];
    } else if (q < a.o) {
      const int64_t k = (int64_t)(p - a.o) * a.o + q;
      dr = a.dai_re[k]; di = a.dai_im[k];
    } else {
      const int64_t k = (int64_t)(p - a.o) * a.v + (q - a.o);
      dr = a.dab_re[k]; di = a.dab_im[k];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double m = a.mu[c * n2 + idx];
      acc[2 * c] += m * dr;
      acc[2 * c + 1] += m * di;
    }
  }
  for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 6) {
    const int64_t row = a.counter ? (*a.counter - 1) : 0;
    a.out[row * 6 + threadIdx.x] = red[threadIdx.x][0];
  }
}

void launch_trace(const TraceArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(trace_kernel, dim3(1), dim3(1024), 0, s, a);
}

__global__ void counter_inc_kernel(int64_t* c) { c[0] += 1; }
__global__ void counter_set_kernel(int64_t* c, int64_t v) { c[0] = v; }

void launch_counter_inc(int64_t* counter, hipStream_t s) {
  hipLaunchKernelGGL(counter_inc_kernel, dim3(1), dim3(1), 0, s, counter);
}
void launch_counter_set(int64_t* counter, int64_t value, hipStream_t s) {
  hipLaunchKernelGGL(counter_set_kernel, dim3(1), dim3(1), 0, s, counter, value);
}

// ===========================================================================
// Synthetic TEI: Philox4x32-10 keyed by (0x20261203, config id), counter
// (component 3, canonical orbit index), Box-Muller cosine branch.
// ===========================================================================
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  const double x = (double)(a >> 5), y = (double)(b >> 6);
  return __dmul_rn(__dadd_rn(__dadd_rn(__dmul_rn(x, 67108864.0), y), 0.5), 1.0 / 9007199254740992.0);
}

__global__ void synth_tei_kernel(double* tei, int n, uint32_t cid, double sigma) {
  const int64_t n2 = (int64_t)n * n, total = n2 * n2;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t s = idx % n, r = (idx / n) % n, q = (idx / n2) % n, p = idx / (n2 * n);
    double sign = (p == q || r == s) ? 0.0 : 1.0;
    if (p > q) { const int64_t t = p; p = q; q = t; sign = -sign; }
    if (r > s) { const int64_t t = r; r = s; s = t; sign = -sign; }
    int64_t a = p * n + q, b = r * n + s;
    if (a > b) { const int64_t t = a; a = b; b = t; }
    const uint64_t cidx = (uint64_t)(a * n2 + b);
    uint32_t c[4] = {3u, (uint32_t)cidx, (uint32_t)(cidx >> 32), 0u};
    philox10(c, 0x20261203u, cid);
    const double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]);
    const double z = __dmul_rn(sqrt(__dmul_rn(-2.0, log(u1))), cos(__dmul_rn(6.283185307179586, u2)));
    tei[idx] = __dmul_rn(sign, __dmul_rn(sigma, z));
  }
}

void launch_synth_tei(double* tei, int n, uint32_t cid, double sigma, hipStream_t s) {
  const int64_t total = (int64_t)n * n * n * n;
  hipLaunchKernelGGL(synth_tei_kernel, dim3(grid_for(total)), dim3(256), 0, s, tei, n, cid, sigma);
}

}  // namespace tdcc

namespace tdcc {

__global__ void split_complex_kernel(const double* in, double* re, double* im, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    re[i] = in[2 * i];
    im[i] = in[2 * i + 1];
  }
}

__global__ void join_complex_kernel(const double* re, const double* im, double* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    out[2 * i] = re[i];
    out[2 * i + 1] = im ? im[i] : 0.0;
  }
}

void launch_split_complex(const double* inter, double* re, double* im, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(split_complex_kernel, dim3(grid_for(n)), dim3(256), 0, s, inter, re, im, n);
}

void launch_join_complex(const double* re, const double* im, double* inter, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(join_complex_kernel, dim3(grid_for(n)), dim3(256), 0, s, re, im, inter, n);
}

}  // namespace tdcc

namespace tdcc {

// ===========================================================================
// Dense real DGEMM, 128 x 128 x 16 tiles on v_mfma_f64_16x16x4_f64.
// 256 threads = 4 waves (2 x 2), each wave owns 64 x 64 = 4 x 4 MFMA blocks.
// Global -> registers -> LDS double buffering, one barrier per K step.
// ===========================================================================
constexpr int DG_BM = 128, DG_BN = 128, DG_BK = 16, DG_LD = 130;

template <bool BT, bool VEC>
__global__ __launch_bounds__(256) void dgemm_kernel(DgemmArgs g) {
  __shared__ double As[2][DG_BK][DG_LD];
  __shared__ double Bs[2][DG_BK][DG_LD];
  const int tiles_m = (g.M + DG_BM - 1) / DG_BM, tiles_n = (g.N + DG_BN - 1) / DG_BN;
  const int L = remap_tile(blockIdx.x, tiles_m * tiles_n);
  const int tm = L % tiles_m, tn = L / tiles_m;
  const int m0 = tm * DG_BM, n0 = tn * DG_BN;
  const int tid = threadIdx.x;
  // A: row ar, k offset ak (8 consecutive k); B likewise (BT) or k row bk, n offset bn (!BT)
  const int ar = tid >> 1, ak = (tid & 1) * 8;
  const int brow = BT ? (tid >> 1) : (tid >> 4);
  const int bcol = BT ? (tid & 1) * 8 : (tid & 15) * 8;
  double ra[8], rb[8];
  // split-K range of this block and its output slice
  const int kb = g.k_chunk > 0 ? (int)blockIdx.y * g.k_chunk : 0;
  const int ke = g.k_chunk > 0 ? min(g.K, kb + g.k_chunk) : g.K;
  double* const Cout = g.C + (int64_t)blockIdx.y * g.c_split_stride;
  auto load8 = [&](const double* p, bool row_ok, int c0, int cmax, double* r) {
    if (VEC && row_ok && c0 + 8 <= cmax) {
      const double2* q = reinterpret_cast<const double2*>(p + c0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 x = q[j];
        r[2 * j] = x.x;
        r[2 * j + 1] = x.y;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) r[j] = (row_ok && c0 + j < cmax) ? p[c0 + j] : 0.0;
    }
  };
  auto load_global = [&](int k0) {
    const int m = m0 + ar;
    load8(g.A + (int64_t)(m < g.M ? m : 0) * g.lda, m < g.M, k0 + ak, ke, ra);
    if (BT) {
      const int n = n0 + brow;
      load8(g.B + (int64_t)(n < g.N ? n : 0) * g.ldb, n < g.N, k0 + bcol, ke, rb);
    } else {
      const int k = k0 + brow;
      load8(g.B + (int64_t)(k < ke ? k : 0) * g.ldb, k < ke, n0 + bcol, g.N, rb);
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 8; ++j) As[buf][ak + j][ar] = ra[j];
    if (BT) {
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[buf][bcol + j][brow] = rb[j];
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[buf][brow][bcol + j] = rb[j];
    }
  };
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  load_global(kb);
  store_lds(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = kb; k0 < ke; k0 += DG_BK) {
    const bool more = k0 + DG_BK < ke;
    if (more) load_global(k0 + DG_BK);
#pragma unroll
    for (int k4 = 0; k4 < DG_BK / 4; ++k4) {
      const int kr = k4 * 4 + lk;
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[buf][kr][wm + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[buf][kr][wn + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) store_lds(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = n0 + wn + j * 16 + lr;
    if (col >= g.N) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + lk + 4 * r;
        if (row >= g.M) continue;
        double* c = Cout + (int64_t)row * g.ldc + col;
        const double x = g.alpha * acc[i][j][r];
        *c = g.beta != 0.0 ? x + g.beta * *c : x;
      }
  }
}

void launch_dgemm(const DgemmArgs& g, bool b_trans, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  const int tiles = ((g.M + DG_BM - 1) / DG_BM) * ((g.N + DG_BN - 1) / DG_BN);
  const int splits = g.k_chunk > 0 ? g.splits : 1;
  // 16-byte vector loads need even leading dimensions and 16-byte aligned bases
  const bool vec = (g.lda % 2 == 0) && (g.ldb % 2 == 0) && ((reinterpret_cast<uintptr_t>(g.A) & 15) == 0) &&
                   ((reinterpret_cast<uintptr_t>(g.B) & 15) == 0);
  const dim3 grid(tiles, splits);
  if (b_trans) {
    if (vec) hipLaunchKernelGGL((dgemm_kernel<true, true>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((dgemm_kernel<true, false>), grid, dim3(256), 0, s, g);
  } else {
    if (vec) hipLaunchKernelGGL((dgemm_kernel<false, true>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((dgemm_kernel<false, false>), grid, dim3(256), 0, s, g);
  }
}

// ===========================================================================
// Antisymmetric pair packing / unpacking of the trailing (v, v) index pair.
// ===========================================================================
__device__ __forceinline__ int64_t pair_index(int a, int b, int v) {
  return (int64_t)a * v - (int64_t)a * (a + 1) / 2 + (b - a - 1);
}

// Rows of the packed ladder operands: all (i, j) (r = i o + j), or with
// ij_pairs only i < j (r = pair index of (i, j) among o, the order in which
// lincomb_tri_index enumerates the strict pairs), the (i, j) antisymmetry of
// the amplitudes making the other half redundant.

// Both kernels work on 32 x 32 tiles (A, B), A <= B, of the (v, v) plane at
// one leading row, 256 threads, element k of a thread at (A 32 + ly + 8k,
// B 32 + lx); the transposed tile (B, A) goes through LDS, so every global
// access runs along a contiguous axis.  pack reads x[e][f] and x[f][e] of the
// tile pair and writes the e < f entries; unpack reads the a < b entries of
// its row, writes y[a][b] and y[b][a] (and, with ij_pairs, the mirrored row
// y[j][i] with the opposite sign; blocks past the last pair write the zero
// rows y[i][i]).  The values are those of the elementwise definition above.
__global__ __launch_bounds__(256) void pack_pairs_kernel(const double* xr, const double* xi, double* X, int o, int v,
                                                         int ij_pairs, int64_t ntiles) {
  __shared__ double tr[kLinTile][kLinTile + 1];
  const int64_t P = (int64_t)v * (v - 1) / 2, vv = (int64_t)v * v;
  const int64_t rows = ij_pairs ? (int64_t)o * (o - 1) / 2 : (int64_t)o * o;
  const int64_t nt = (v + kLinTile - 1) / kLinTile;
  const int64_t blk = blockIdx.x;
  const int64_t pr = blk / ntiles, t = blk - pr * ntiles;  // plane * rows + r, tile
  const int64_t pl = pr / rows, r = pr - pl * rows;
  int64_t i, j, tA, tB;
  if (ij_pairs) lincomb_tri_index(r, o, true, i, j);
  else {
    i = r / o;
    j = r - i * o;
  }
  lincomb_tri_index(t, nt, false, tA, tB);
  const double* x = (pl ? xi : xr) + (i * o + j) * vv;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  const int64_t r0 = tA * kLinTile, c0 = tB * kLinTile;
  // x[f][e] of this tile: row f = c0 + .., column e = r0 + ..
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t f = c0 + ly + 8 * k, e = r0 + lx;
    tr[ly + 8 * k][lx] = (f < v && e < v) ? x[f * v + e] : 0.0;
  }
  __syncthreads();
  double* Xrow = X + (pl * rows + r) * P;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t e = r0 + ly + 8 * k, f = c0 + lx;
    if (e >= f || f >= v) continue;
    const double xfe = tr[lx][ly + 8 * k];
    Xrow[pair_index((int)e, (int)f, v)] = __dsub_rn(x[e * v + f], xfe);
  }
}

__global__ __launch_bounds__(256) void unpack_pairs_kernel(const double* Y, double* yr, double* yi, int o, int v,
                                                           double alpha, int ij_pairs, int64_t npairs, int64_t ntiles) {
  __shared__ double tr[kLinTile][kLinTile + 1];
  const int64_t P = (int64_t)v * (v - 1) / 2, vv = (int64_t)v * v;
  const int64_t rows = ij_pairs ? (int64_t)o * (o - 1) / 2 : (int64_t)o * o;
  const int64_t nt = (v + kLinTile - 1) / kLinTile;
  const int64_t blk = blockIdx.x;
  const int64_t pp = blk / ntiles, t = blk - pp * ntiles;  // plane * npairs + pair, tile
  const int64_t pl = pp / npairs, p = pp - pl * npairs;
  int64_t tA, tB;
  lincomb_tri_index(t, nt, false, tA, tB);
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  const int64_t r0 = tA * kLinTile, c0 = tB * kLinTile;
  const bool diag_tile = tA == tB;
  double* yp = pl ? yi : yr;
  if (p >= rows) {  // zero row y[i][i] (ij_pairs)
    const int64_t i = p - rows;
    double* y = yp + (i * o + i) * vv;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t a = r0 + ly + 8 * k, b = c0 + lx;
      if (a < v && b < v) y[a * v + b] = 0.0;
      const int64_t a2 = c0 + ly + 8 * k, b2 = r0 + lx;
      if (!diag_tile && a2 < v && b2 < v) y[a2 * v + b2] = 0.0;
    }
    return;
  }
  int64_t i, j;
  if (ij_pairs) lincomb_tri_index(p, o, true, i, j);
  else {
    i = p / o;
    j = p - i * o;
  }
  const double* Yrow = Y + (pl * rows + p) * P;
  // w = alpha * Y[pair(a, b)] for a < b of the tile, 0 elsewhere
  double w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t a = r0 + ly + 8 * k, b = c0 + lx;
    w[k] = (a < b && b < v) ? __dmul_rn(alpha, Yrow[pair_index((int)a, (int)b, v)]) : 0.0;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) tr[ly + 8 * k][lx] = w[k];
  __syncthreads();
  double* y = yp + (i * o + j) * vv;
  double* ym = yp + (j * o + i) * vv;  // mirrored row (ij_pairs)
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t a = r0 + ly + 8 * k, b = c0 + lx;
    if (a >= v || b >= v) continue;
    // on a diagonal tile the lower triangle is minus the upper one
    const double val = (diag_tile && a > b) ? -tr[lx][ly + 8 * k] : w[k];
    y[a * v + b] = val;
    if (ij_pairs) ym[a * v + b] = -val;
  }
  if (!diag_tile) {
    // y[b][a] = -y[a][b]: thread stores row c0 + ly + 8k, column r0 + lx
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t a = c0 + ly + 8 * k, b = r0 + lx;
      if (a >= v || b >= v) continue;
      const double val = -tr[lx][ly + 8 * k];
      y[a * v + b] = val;
      if (ij_pairs) ym[a * v + b] = -val;
    }
  }
}

void launch_pack_pairs(const double* xr, const double* xi, double* X, int o, int v, bool ij_pairs, hipStream_t s) {
  const int64_t nt = (v + kLinTile - 1) / kLinTile, ntiles = nt * (nt + 1) / 2;
  const int64_t rows = ij_pairs ? (int64_t)o * (o - 1) / 2 : (int64_t)o * o;
  const int64_t blocks = 2 * rows * ntiles;
  if (blocks <= 0) return;
  hipLaunchKernelGGL(pack_pairs_kernel, dim3((unsigned)blocks), dim3(256), 0, s, xr, xi, X, o, v, (int)ij_pairs, ntiles);
}

void launch_unpack_pairs(const double* Y, double* yr, double* yi, int o, int v, double alpha, bool ij_pairs,
                         hipStream_t s) {
  const int64_t nt = (v + kLinTile - 1) / kLinTile, ntiles = nt * (nt + 1) / 2;
  const int64_t npairs = ij_pairs ? (int64_t)o * (o - 1) / 2 + o : (int64_t)o * o;
  const int64_t blocks = 2 * npairs * ntiles;
  if (blocks <= 0) return;
  hipLaunchKernelGGL(unpack_pairs_kernel, dim3((unsigned)blocks), dim3(256), 0, s, Y, yr, yi, o, v, alpha,
                     (int)ij_pairs, npairs, ntiles);
}

__global__ void sum_splits_kernel(double* y, int splits, int64_t stride, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = y[i];
    for (int c = 1; c < splits; ++c) acc = __dadd_rn(acc, y[c * stride + i]);
    y[i] = acc;
  }
}

void launch_sum_splits(double* y, int splits, int64_t stride, int64_t n, hipStream_t s) {
  if (splits <= 1 || n <= 0) return;
  hipLaunchKernelGGL(sum_splits_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, splits, stride, n);
}

__global__ void build_vpack_kernel(const double* tei, int o, int v, int ket, double* Vp) {
  const int64_t n = o + v, P = (int64_t)v * (v - 1) / 2, vv = (int64_t)v * v, total = vv * vv;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t abi = idx / vv, efi = idx - abi * vv;
    const int a = (int)(abi / v), b = (int)(abi - (int64_t)a * v);
    const int e = (int)(efi / v), f = (int)(efi - (int64_t)e * v);
    if (a >= b || e >= f) continue;
    const int64_t p = o + a, q = o + b, r = o + e, t = o + f;
    const int64_t src = ket ? ((r * n + t) * n + p) * n + q : ((p * n + q) * n + r) * n + t;
    Vp[pair_index(a, b, v) * P + pair_index(e, f, v)] = tei[src];
  }
}

void launch_build_vpack(const double* tei, int o, int v, int ket, double* Vp, hipStream_t s) {
  const int64_t total = (int64_t)v * v * v * v;
  hipLaunchKernelGGL(build_vpack_kernel, dim3(grid_for(total)), dim3(256), 0, s, tei, o, v, ket, Vp);
}

// Deterministic counting: per-block partial counts, then one block sums them.

__device__ void block_count_store(unsigned long long local, unsigned long long* partial) {
  __shared__ unsigned long long red[256];
  red[threadIdx.x] = local;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void antisym_check_kernel(const double* tei, int o, int v, unsigned long long* partial) {
  const int64_t n = o + v, vv = (int64_t)v * v, total = vv * vv;
  unsigned long long bad = 0;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t abi = idx / vv, efi = idx - abi * vv;
    const int64_t a = o + abi / v, b = o + abi % v, e = o + efi / v, f = o + efi % v;
    const double x = tei[((a * n + b) * n + e) * n + f];
    const double y = tei[((b * n + a) * n + e) * n + f];
    const double z = tei[((a * n + b) * n + f) * n + e];
    bad += (x != -y) || (x != -z);
  }
  block_count_store(bad, partial);
}

// Entries of the whole (n,n,n,n) block that break <pq||rs> = -<qp||rs> =
// -<pq||sr> (exact comparison).
__global__ void tei_antisym_check_kernel(const double* tei, int n, unsigned long long* partial) {
  const int64_t nn = (int64_t)n * n, total = nn * nn;
  unsigned long long bad = 0;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pq = idx / nn, rs = idx - pq * nn;
    const int64_t p = pq / n, q = pq - p * n, r = rs / n, t = rs - r * n;
    const double x = tei[idx];
    const double y = tei[(q * n + p) * nn + rs];
    const double z = tei[pq * nn + t * n + r];
    bad += (x != -y) || (x != -z);
  }
  block_count_store(bad, partial);
}

__global__ void mismatch_kernel(const double* x, const double* y, int64_t n, unsigned long long* partial) {
  unsigned long long bad = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    bad += x[i] != y[i];
  block_count_store(bad, partial);
}

__global__ void sum_counts_kernel(const unsigned long long* partial, int nb, unsigned long long* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    unsigned long long s = 0;
    for (int i = 0; i < nb; ++i) s += partial[i];
    out[0] = s;
  }
}

// `count` must hold kCountBlocks + 1 entries; the total lands in count[kCountBlocks].
void launch_vvvv_antisym_check(const double* tei, int o, int v, unsigned long long* count, hipStream_t s) {
  hipLaunchKernelGGL(antisym_check_kernel, dim3(kCountBlocks), dim3(256), 0, s, tei, o, v, count);
  hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(64), 0, s, count, kCountBlocks, count + kCountBlocks);
}

void launch_tei_antisym_check(const double* tei, int n, unsigned long long* count, hipStream_t s) {
  hipLaunchKernelGGL(tei_antisym_check_kernel, dim3(kCountBlocks), dim3(256), 0, s, tei, n, count);
  hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(64), 0, s, count, kCountBlocks, count + kCountBlocks);
}

// (i, j)- and (a, b)-antisymmetry of an interleaved complex (o, o, v, v)
// amplitude in device memory: per-block maxima of |x| and of the deviations
// |x_ijab + x_jiab|, |x_ijab + x_ijba| (NaN entries ignored, as std::max
// ignores them on the host), then one block takes the maxima of the partials.
__global__ void amp_antisym_kernel(const double* x, int o, int v, double* partial) {
  __shared__ double rm[256], rd[256];
  const int64_t vv = (int64_t)v * v, ovv = (int64_t)o * vv, total = 2 * o * ovv;
  double mx = 0.0, dev = 0.0;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = idx >> 1, c = idx & 1;
    const int64_t i = e / ovv, r1 = e - i * ovv, j = r1 / vv, ab = r1 - j * vv, a = ab / v, b = ab - a * v;
    const double xa = x[idx];
    mx = fmax(mx, fabs(xa));
    dev = fmax(dev, fabs(xa + x[2 * ((j * o + i) * vv + ab) + c]));
    dev = fmax(dev, fabs(xa + x[2 * ((i * o + j) * vv + b * v + a) + c]));
  }
  rm[threadIdx.x] = mx;
  rd[threadIdx.x] = dev;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rm[threadIdx.x] = fmax(rm[threadIdx.x], rm[threadIdx.x + w]);
      rd[threadIdx.x] = fmax(rd[threadIdx.x], rd[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = rm[0];
    partial[2 * blockIdx.x + 1] = rd[0];
  }
}

__global__ void max2_reduce_kernel(const double* partial, int nb, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double m = 0.0, d = 0.0;
    for (int i = 0; i < nb; ++i) {
      m = fmax(m, partial[2 * i]);
      d = fmax(d, partial[2 * i + 1]);
    }
    out[0] = m;
    out[1] = d;
  }
}

void launch_amp_antisym_check(const double* x, int o, int v, double* work, hipStream_t s) {
  hipLaunchKernelGGL(amp_antisym_kernel, dim3(kCountBlocks), dim3(256), 0, s, x, o, v, work);
  hipLaunchKernelGGL(max2_reduce_kernel, dim3(1), dim3(64), 0, s, work, kCountBlocks, work + 2 * kCountBlocks);
}

void launch_count_mismatch(const double* x, const double* y, int64_t n, unsigned long long* count, hipStream_t s) {
  hipLaunchKernelGGL(mismatch_kernel, dim3(kCountBlocks), dim3(256), 0, s, x, y, n, count);
  hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(64), 0, s, count, kCountBlocks, count + kCountBlocks);
}

}  // namespace tdcc
