// Skinny tensor contractions on gfx950: one operand has at most 32 rows.
//
//   C[cs[i] + cb[j]] = alpha * sum_k S[i, k] * G[j, k]
//   S[i, k] = s_plane[sr[i] + sk[k]]   (small operand, i < I <= 32)
//   G[j, k] = g_plane[gr[j] + gk[k]]   (big operand, j < J)
//
// These shapes (matrix-vector products with t1, outer products over a short
// index, and dot products over o^2 v^2 or o v^3) are HBM-bound: the big
// operand is streamed once and the small one sits in LDS.  FP64 VALU, fixed
// summation order, optional split over K with the partial sums reduced in
// split order.
#include "kernels.hpp"

namespace tdcc {

constexpr int SK_KC = 64;  // k chunk staged in LDS
constexpr int SK_KU = 8;   // big-operand loads in flight per thread (rows kernel)

// Small-operand chunk in LDS, [i][k] with a padded row so that lanes reading
// consecutive k hit distinct banks.
template <bool SC, int IT>
struct SmallChunk {
  double re[IT][SK_KC + 1];
  double im[SC ? IT : 1][SK_KC + 1];
};

template <bool SC, int IT>
__device__ __forceinline__ void load_small(SmallChunk<SC, IT>& sm, const SkinnyArgs& a, int64_t k0, int64_t k1) {
  for (int e = threadIdx.x; e < IT * SK_KC; e += blockDim.x) {
    const int i = e / SK_KC, kk = e - i * SK_KC;
    const int64_t k = k0 + kk;
    const bool ok = i < a.I && k < k1;
    const int64_t off = ok ? a.sr[i] + a.sk[k] : 0;
    sm.re[i][kk] = ok ? a.s_re[off] : 0.0;
    if (SC) sm.im[i][kk] = ok ? a.s_im[off] : 0.0;
  }
}

// Store (or write a partial of) the IT accumulators of big row j, which hold
// small rows i0 .. i0 + IT - 1.
template <bool SC, bool GC, int IT>
__device__ __forceinline__ void store_row(const SkinnyArgs& a, int64_t j, const double* acc_re, const double* acc_im,
                                          int i0 = 0) {
  constexpr bool CC = SC || GC;
  if (a.splits > 1) {
    double* p = a.partial + (int64_t)blockIdx.y * (CC ? 2 : 1) * a.I * a.J;
#pragma unroll
    for (int ii = 0; ii < IT; ++ii) {
      const int i = i0 + ii;
      if (i < a.I) {
        p[(int64_t)i * a.J + j] = acc_re[ii];
        if (CC) p[(int64_t)(a.I + i) * a.J + j] = acc_im[ii];
      }
    }
    return;
  }
#pragma unroll
  for (int ii = 0; ii < IT; ++ii) {
    const int i = i0 + ii;
    if (i < a.I) {
      const int64_t o = a.cs[i] + a.cb[j];
      const double yr = a.alpha_re * acc_re[ii] - (CC ? a.alpha_im * acc_im[ii] : 0.0);
      a.c_re[o] = yr;
      if (CC) a.c_im[o] = a.alpha_re * acc_im[ii] + a.alpha_im * acc_re[ii];
    }
  }
}

// The big operand's contiguous index is a row index: one lane per big row,
// loop over k (reads coalesced across lanes).  With IG > 1 the IT small rows
// are split over the block's IG waves, which then cover the same 64 big rows
// (the repeated reads hit the CU's cache): fewer accumulators per thread,
// more waves in flight.
template <bool SC, bool GC, int IT, int IG>
__global__ __launch_bounds__(256) void skinny_rows_kernel(SkinnyArgs a) {
  __shared__ SmallChunk<SC, IT> sm;
  constexpr int RT = IT / IG;  // small rows per thread
  const int ig = IG > 1 ? (int)(threadIdx.x >> 6) : 0;
  const int i0 = ig * RT;
  const int64_t j = (int64_t)blockIdx.x * (256 / IG) + (IG > 1 ? (threadIdx.x & 63) : threadIdx.x);
  const bool live = j < a.J;
  const int64_t kb = (int64_t)blockIdx.y * a.k_chunk, ke = min(a.K, kb + a.k_chunk);
  const double* gre = a.g_re + (live ? a.gr[j] : 0);
  const double* gim = GC ? a.g_im + (live ? a.gr[j] : 0) : nullptr;
  double acc_re[RT], acc_im[RT];
#pragma unroll
  for (int i = 0; i < RT; ++i) acc_re[i] = acc_im[i] = 0.0;
  for (int64_t k0 = kb; k0 < ke; k0 += SK_KC) {
    __syncthreads();
    load_small<SC, IT>(sm, a, k0, ke);
    __syncthreads();
    const int kn = (int)min((int64_t)SK_KC, ke - k0);
    if (!live || i0 >= a.I) continue;
    // the big-operand loads of SK_KU consecutive k are issued together, so a
    // thread has that many in flight; the sums still run in k order
    for (int kb8 = 0; kb8 < kn; kb8 += SK_KU) {
      double br[SK_KU], bi[SK_KU];
#pragma unroll
      for (int u = 0; u < SK_KU; ++u) {
        const bool ok = kb8 + u < kn;
        const int64_t go = ok ? a.gk[k0 + kb8 + u] : 0;
        br[u] = ok ? gre[go] : 0.0;
        bi[u] = (GC && ok) ? gim[go] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < SK_KU; ++u) {
        if (kb8 + u >= kn) break;
        const int kk = kb8 + u;
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          const double sr = sm.re[i0 + i][kk];
          const double si = SC ? sm.im[i0 + i][kk] : 0.0;
          acc_re[i] = fma(sr, br[u], acc_re[i]);
          if (SC && GC) acc_re[i] = fma(-si, bi[u], acc_re[i]);
          if (GC) acc_im[i] = fma(sr, bi[u], acc_im[i]);
          if (SC) acc_im[i] = fma(si, br[u], acc_im[i]);
        }
      }
    }
  }
  if (live) store_row<SC, GC, RT>(a, j, acc_re, acc_im, i0);
}

// The big operand's contiguous index lies in k: one wave per big row, lanes
// over k, then a lane reduction.  4 rows per block share the LDS chunk of the
// small operand, which holds KC k (32 KB); lane l takes k = k0 + l + 64u of the
// chunk, u < KC / 64, in order, with the offsets of all its k and then eight
// big-operand loads at a time in flight.
constexpr int sk_kchunk(bool sc, int it) {
  const int kc = 32768 / (it * 8 * (sc ? 2 : 1)) / 64 * 64;
  return kc < 64 ? 64 : (kc > 1024 ? 1024 : kc);
}

template <bool SC, int IT, int KC>
struct SmallChunkK {
  double re[IT][KC + 1];
  double im[SC ? IT : 1][KC + 1];
};

template <bool SC, bool GC, int IT>
__global__ __launch_bounds__(256) void skinny_k_kernel(SkinnyArgs a) {
  constexpr int KC = sk_kchunk(SC, IT), KU = KC / 64;
  __shared__ SmallChunkK<SC, IT, KC> sm;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j = (int64_t)blockIdx.x * 4 + w;
  const bool live = j < a.J;
  const int64_t kb = (int64_t)blockIdx.y * a.k_chunk, ke = min(a.K, kb + a.k_chunk);
  const double* gre = a.g_re + (live ? a.gr[j] : 0);
  const double* gim = GC ? a.g_im + (live ? a.gr[j] : 0) : nullptr;
  double acc_re[IT], acc_im[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) acc_re[i] = acc_im[i] = 0.0;
  for (int64_t k0 = kb; k0 < ke; k0 += KC) {
    __syncthreads();
    for (int e = threadIdx.x; e < IT * KC; e += blockDim.x) {
      const int i = e / KC, kk = e - i * KC;
      const int64_t k = k0 + kk;
      const bool ok = i < a.I && k < ke;
      const int64_t off = ok ? a.sr[i] + a.sk[k] : 0;
      sm.re[i][kk] = ok ? a.s_re[off] : 0.0;
      if (SC) sm.im[i][kk] = ok ? a.s_im[off] : 0.0;
    }
    __syncthreads();
    if (!live) continue;
    int64_t go[KU];
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int64_t k = k0 + 64 * u + lane;
      go[u] = k < ke ? a.gk[k] : -1;
    }
#pragma unroll
    for (int u0 = 0; u0 < KU; u0 += 8) {
      double br[8], bi[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const bool ok = u0 + u < KU && go[u0 + u < KU ? u0 + u : 0] >= 0;
        const int64_t o = ok ? go[u0 + u] : 0;
        br[u] = ok ? gre[o] : 0.0;
        bi[u] = (GC && ok) ? gim[o] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (u0 + u >= KU) break;
        const int kk = 64 * (u0 + u) + lane;
#pragma unroll
        for (int i = 0; i < IT; ++i) {
          const double sr = sm.re[i][kk];
          const double si = SC ? sm.im[i][kk] : 0.0;
          acc_re[i] = fma(sr, br[u], acc_re[i]);
          if (SC && GC) acc_re[i] = fma(-si, bi[u], acc_re[i]);
          if (GC) acc_im[i] = fma(sr, bi[u], acc_im[i]);
          if (SC) acc_im[i] = fma(si, br[u], acc_im[i]);
        }
      }
    }
  }
  // butterfly reduction: every lane ends with the same, order-fixed sum
#pragma unroll
  for (int i = 0; i < IT; ++i)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      acc_re[i] += __shfl_xor(acc_re[i], off);
      if (SC || GC) acc_im[i] += __shfl_xor(acc_im[i], off);
    }
  if (live && lane == 0) store_row<SC, GC, IT>(a, j, acc_re, acc_im);
}

__global__ void skinny_reduce_kernel(SkinnyArgs a, int cplx) {
  const int64_t total = (int64_t)a.I * a.J;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / a.J, j = e - i * a.J;
    double vr = 0.0, vi = 0.0;
    for (int s = 0; s < a.splits; ++s) {
      const double* p = a.partial + (int64_t)s * (cplx ? 2 : 1) * total;
      vr += p[e];
      if (cplx) vi += p[total + e];
    }
    const int64_t o = a.cs[i] + a.cb[j];
    a.c_re[o] = a.alpha_re * vr - a.alpha_im * vi;
    if (cplx) a.c_im[o] = a.alpha_re * vi + a.alpha_im * vr;
  }
}

// Short runs: the big operand's contiguous index is summed but runs only
// `run` elements (e.g. 20 consecutive m of ieam summed over (e, m)), so a wave
// over consecutive k touches a few separate short runs per load.  Instead one
// block takes one (jh, kh) block of run_rows x run contiguous elements: loads
// it coalesced into LDS (row jl at jl * (run + 1), the padding spreads the
// rows' banks), and thread jl sums its row against the small operand's run in
// r order.  Partial kh goes to partial[kh] in the layout skinny_reduce_kernel
// reads; skinny_runs_reduce_kernel sums the partials in kh order.
template <bool SC, bool GC>
__global__ __launch_bounds__(256) void skinny_runs_kernel(SkinnyArgs a) {
  constexpr bool CC = SC || GC;
  __shared__ double gre[kSkinnyRunsMax], gim[GC ? kSkinnyRunsMax : 1];
  __shared__ double sre[4][64], sim[SC ? 4 : 1][64];
  const int R = a.run, Jl = a.run_rows, n = R * Jl;
  const int64_t kh = blockIdx.x, jh = blockIdx.y;
  const int64_t base = a.gr[jh * Jl] + a.gk[kh * R];
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int jl = e / R, r = e - jl * R;
    gre[jl * (R + 1) + r] = a.g_re[base + e];
    if (GC) gim[jl * (R + 1) + r] = a.g_im[base + e];
  }
  for (int e = threadIdx.x; e < a.I * R; e += blockDim.x) {
    const int i = e / R, r = e - i * R;
    const int64_t off = a.sr[i] + a.sk[kh * R + r];
    sre[i][r] = a.s_re[off];
    if (SC) sim[i][r] = a.s_im[off];
  }
  __syncthreads();
  const int64_t total = (int64_t)a.I * a.J;
  double* const p = a.partial + kh * (CC ? 2 : 1) * total;
  for (int jl = threadIdx.x; jl < Jl; jl += blockDim.x) {
    double acc_re[4] = {0.0, 0.0, 0.0, 0.0}, acc_im[4] = {0.0, 0.0, 0.0, 0.0};
    for (int r = 0; r < R; ++r) {
      const double br = gre[jl * (R + 1) + r], bi = GC ? gim[jl * (R + 1) + r] : 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i >= a.I) break;
        const double s_r = sre[i][r], s_i = SC ? sim[i][r] : 0.0;
        acc_re[i] = fma(s_r, br, acc_re[i]);
        if (SC && GC) acc_re[i] = fma(-s_i, bi, acc_re[i]);
        if (GC) acc_im[i] = fma(s_r, bi, acc_im[i]);
        if (SC) acc_im[i] = fma(s_i, br, acc_im[i]);
      }
    }
    const int64_t j = jh * Jl + jl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= a.I) break;
      p[(int64_t)i * a.J + j] = acc_re[i];
      if (CC) p[total + (int64_t)i * a.J + j] = acc_im[i];
    }
  }
}

// Sum of the short-run partials: 64 outputs per block, wave w adds partials
// w, w + 4, ... in order, then the four wave sums are added in wave order.
__global__ __launch_bounds__(256) void skinny_runs_reduce_kernel(SkinnyArgs a, int cplx) {
  __shared__ double wr[4][64], wi[4][64];
  const int64_t total = (int64_t)a.I * a.J;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * 64 + lane;
  double vr = 0.0, vi = 0.0;
  if (e < total) {
#pragma unroll 8
    for (int s = w; s < a.splits; s += 4) {
      const double* p = a.partial + (int64_t)s * (cplx ? 2 : 1) * total;
      vr += p[e];
      if (cplx) vi += p[total + e];
    }
  }
  wr[w][lane] = vr;
  wi[w][lane] = vi;
  __syncthreads();
  if (w == 0 && e < total) {
    vr = ((wr[0][lane] + wr[1][lane]) + wr[2][lane]) + wr[3][lane];
    vi = ((wi[0][lane] + wi[1][lane]) + wi[2][lane]) + wi[3][lane];
    const int64_t i = e / a.J, j = e - i * a.J;
    const int64_t o = a.cs[i] + a.cb[j];
    a.c_re[o] = a.alpha_re * vr - a.alpha_im * vi;
    if (cplx) a.c_im[o] = a.alpha_re * vi + a.alpha_im * vr;
  }
}

template <bool SC, bool GC, int IT>
static void skinny_launch(const SkinnyArgs& a, hipStream_t s) {
  // small rows split over the block's 4 waves for 8 or more of them, except
  // for a short K (at most 32: one wave holds all rows and streams its k;
  // measured 3-11 % faster at I = 20, K = 20, and 40-60 % slower at K = 160)
  const int ig = (IT >= 8 && a.K > 32) ? 4 : 1;
  const int64_t per = a.g_kmajor ? 4 : (ig == 4 ? 64 : 256);
  const dim3 grid((unsigned)((a.J + per - 1) / per), (unsigned)a.splits), block(256);
  if (a.g_kmajor) hipLaunchKernelGGL((skinny_k_kernel<SC, GC, IT>), grid, block, 0, s, a);
  else if (ig == 4) hipLaunchKernelGGL((skinny_rows_kernel<SC, GC, IT, 4>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((skinny_rows_kernel<SC, GC, IT, 1>), grid, block, 0, s, a);
}

template <bool SC, bool GC>
static void skinny_dispatch(const SkinnyArgs& a, hipStream_t s) {
  if (a.I <= 4) skinny_launch<SC, GC, 4>(a, s);
  else if (a.I <= 8) skinny_launch<SC, GC, 8>(a, s);
  else if (a.I <= 16) skinny_launch<SC, GC, 16>(a, s);
  else if (a.I <= 20) skinny_launch<SC, GC, 20>(a, s);  // o = 20: no padded rows
  else skinny_launch<SC, GC, 32>(a, s);
}

void launch_skinny(const SkinnyArgs& a, hipStream_t s) {
  if (a.I <= 0 || a.J <= 0) return;
  const bool sc = a.s_im != nullptr, gc = a.g_im != nullptr;
  if (a.run > 0) {
    // K / run blocks along x, J / run_rows along y (checked when lowered)
    const dim3 grid((unsigned)(a.K / a.run), (unsigned)(a.J / a.run_rows)), block(256);
    if (sc && gc) hipLaunchKernelGGL((skinny_runs_kernel<true, true>), grid, block, 0, s, a);
    else if (sc) hipLaunchKernelGGL((skinny_runs_kernel<true, false>), grid, block, 0, s, a);
    else if (gc) hipLaunchKernelGGL((skinny_runs_kernel<false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((skinny_runs_kernel<false, false>), grid, block, 0, s, a);
    const int64_t total = (int64_t)a.I * a.J;
    hipLaunchKernelGGL(skinny_runs_reduce_kernel, dim3((unsigned)((total + 63) / 64)), dim3(256), 0, s, a,
                       (int)(sc || gc));
    return;
  }
  if (sc && gc) skinny_dispatch<true, true>(a, s);
  else if (sc) skinny_dispatch<true, false>(a, s);
  else if (gc) skinny_dispatch<false, true>(a, s);
  else skinny_dispatch<false, false>(a, s);
  if (a.splits > 1) {
    const int64_t total = (int64_t)a.I * a.J;
    int64_t b = (total + 255) / 256;
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(skinny_reduce_kernel, dim3((unsigned)b), dim3(256), 0, s, a, (int)(sc || gc));
  }
}

}  // namespace tdcc
