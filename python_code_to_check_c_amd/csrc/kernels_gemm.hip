// Gather / scatter kernels that put a tensor contraction into dense GEMM form
// (gfx950), and the dot-product contraction kernel.  The operands of a
// contraction are addressed through the same per-index-group offset tables as
// the strided contraction kernel.
#include "kernels.hpp"

namespace tdcc {

typedef double d4 __attribute__((ext_vector_type(4)));

static int grid_of(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (int)b;
}

// dst[r][c] (row-major, ld) = src_plane[rtab[r % Rb] + ctab[c % Cb]] with the
// plane chosen by r / Rb (planes_along_rows) or c / Cb.
__global__ void gather2d_kernel(Gather2dArgs g) {
  const int64_t R = g.planes_along_rows ? g.planes * g.Rb : g.Rb;
  const int64_t C = g.planes_along_rows ? g.Cb : g.planes * g.Cb;
  const int64_t total = R * C;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const bool i32 = total < ((int64_t)1 << 31);
    const int64_t r = i32 ? (int64_t)((uint32_t)idx / (uint32_t)C) : idx / C, c = idx - r * C;
    int64_t rb = r, cb = c, p = 0;
    if (g.planes_along_rows) {
      p = r >= g.Rb;
      rb = r - p * g.Rb;
    } else {
      p = c >= g.Cb;
      cb = c - p * g.Cb;
    }
    const double* src = p ? g.src_im : g.src_re;
    g.dst[r * g.ld + c] = src[g.rtab[rb] + g.ctab[cb]];
  }
}

void launch_gather2d(const Gather2dArgs& g, hipStream_t s) {
  const int64_t total = g.planes * g.Rb * g.Cb;
  hipLaunchKernelGGL(gather2d_kernel, dim3(grid_of(total)), dim3(256), 0, s, g);
}

// out[cm[m] + cn[n]] = alpha * value(m, n) with value recombined from the
// GEMM result Z (row-major, ldz):
//   kind 0 (real x real):       Z[m][n]
//   kind 1 (complex x real):    Z[m][n] + i Z[M+m][n]
//   kind 2 (complex x complex): (Z[m][n] - Z[M+m][N+n]) + i (Z[m][N+n] + Z[M+m][n])
__global__ void gemm_scatter_kernel(GemmScatterArgs g) {
  const int64_t total = (int64_t)g.M * g.N;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    // 32-bit division when the element count allows (64-bit division is a
    // long instruction sequence)
    const int64_t m = total < ((int64_t)1 << 31) ? (int64_t)((uint32_t)idx / (uint32_t)g.N) : idx / g.N;
    const int64_t n = idx - m * g.N;
    double vr = 0.0, vi = 0.0;
    for (int s = 0; s < g.splits; ++s) {
      const double* z = g.Z + s * g.z_stride;
      if (g.kind == 0) {
        vr += z[m * g.ldz + n];
      } else if (g.kind == 1) {
        vr += z[m * g.ldz + n];
        vi += z[(m + g.M) * g.ldz + n];
      } else {
        vr += z[m * g.ldz + n] - z[(m + g.M) * g.ldz + g.N + n];
        vi += z[m * g.ldz + g.N + n] + z[(m + g.M) * g.ldz + n];
      }
    }
    const int64_t o = g.cm[m] + g.cn[n];
    double yr = g.alpha_re * vr - g.alpha_im * vi;
    double yi = g.alpha_re * vi + g.alpha_im * vr;
    if (g.accumulate) {
      yr += g.out_re[o];
      if (g.out_im) yi += g.out_im[o];
    }
    store_mirrored(g.out_re, g.out_im, g.cm[m], g.cn[n], g.cm2 ? g.cm2[m] : -1, g.cn2 ? g.cn2[n] : -1, yr, yi);
  }
}

void launch_gemm_scatter(const GemmScatterArgs& g, hipStream_t s) {
  hipLaunchKernelGGL(gemm_scatter_kernel, dim3(grid_of((int64_t)g.M * g.N)), dim3(256), 0, s, g);
}

// ---------------------------------------------------------------------------
// Dot-product contraction: a block owns 64 rows of A' (grid z, zero padded)
// and 64 rows of B' (grid x) over one K range (grid y), in 32-wide K steps staged through
// LDS (two buffers, one barrier per step); wave w computes rows 16w .. 16w+15
// as four 16 x 16 blocks on v_mfma_f64_16x16x4_f64 (lane l: A[row l%16][k
// l/16], B[k l/16][col l%16], D[row l/16 + 4r][col l%16], as in the DGEMM
// kernel).  The loader thread owns k = tid % 32 of every step and the rows
// tid / 32 + 8p, p < 8, of both tiles (row offsets read once per block).  The
// shapes are HBM-bound, so the loads run two steps ahead of the MFMAs in two
// register sets (steps of even and odd parity), and the k offsets one step
// ahead of the loads.
// ---------------------------------------------------------------------------
constexpr int DK_BK = 32, DK_BN = 64, DK_LD = 65;

template <int PA, int PB, bool AKC, bool BKC>
__global__ __launch_bounds__(256) void dot_kernel(DotArgs g) {
  __shared__ double As[2][DK_BK][DK_LD];
  __shared__ double Bs[2][DK_BK][DK_LD];
  const int Mr = PA * g.M, Nc = PB * g.N;
  const int n0 = blockIdx.x * DK_BN, m0 = blockIdx.z * DK_BN;
  const int64_t kb = (int64_t)blockIdx.y * g.k_chunk;
  const int64_t ke = kb + g.k_chunk < g.K ? kb + g.k_chunk : g.K;
  const int64_t nsteps = (ke - kb + DK_BK - 1) / DK_BK;
  const int tid = threadIdx.x;
  // the wave index, known to the compiler as uniform across the wave, so that
  // the k offsets of an operand contiguous along its rows (one k per wave)
  // are read with scalar loads
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  // loader element p of an operand contiguous along k: row tid/32 + 8p, k
  // tid%32; of one contiguous along its rows: row tid%64, k tid/64 + 4p
  auto lrow = [&](bool kc, int p) { return kc ? (tid >> 5) + 8 * p : (tid & 63); };
  auto lk = [&](bool kc, int p) { return kc ? (tid & 31) : wv + 4 * p; };
  const double* arow[8];
  const double* brow[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const int r = m0 + lrow(AKC, p);
    arow[p] = nullptr;
    if (r < Mr) {
      const bool im = PA == 2 && r >= g.M;
      arow[p] = (im ? g.a_im : g.a_re) + g.am[im ? r - g.M : r];
    }
    const int c = n0 + lrow(BKC, p);
    brow[p] = nullptr;
    if (c < Nc) {
      const bool im = PB == 2 && c >= g.N;
      brow[p] = (im ? g.b_im : g.b_re) + g.bn[im ? c - g.N : c];
    }
  }
  // k offsets of step t for this thread's elements (-1: past the range)
  auto offsets = [&](int64_t t, int64_t* oa, int64_t* ob) {
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int64_t ka = kb + t * DK_BK + lk(AKC, p), kbb = kb + t * DK_BK + lk(BKC, p);
      oa[p] = (t < nsteps && ka < ke) ? g.ak[ka] : -1;
      ob[p] = (t < nsteps && kbb < ke) ? g.bk[kbb] : -1;
    }
  };
  auto load = [&](const int64_t* oa, const int64_t* ob, double* ra, double* rb) {
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      ra[p] = (oa[p] >= 0 && arow[p]) ? arow[p][oa[p]] : 0.0;
      rb[p] = (ob[p] >= 0 && brow[p]) ? brow[p][ob[p]] : 0.0;
    }
  };
  auto store = [&](int buf, const double* ra, const double* rb) {
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      As[buf][lk(AKC, p)][lrow(AKC, p)] = ra[p];
      Bs[buf][lk(BKC, p)][lrow(BKC, p)] = rb[p];
    }
  };
  d4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = d4{0.0, 0.0, 0.0, 0.0};
  const int lane = tid & 63, w = tid >> 6, lr = lane & 15, lq = lane >> 4;
  const bool active = m0 + 16 * w < Mr;
  auto compute = [&](int buf) {
    if (!active) return;
#pragma unroll
    for (int k4 = 0; k4 < DK_BK / 4; ++k4) {
      const double a = As[buf][4 * k4 + lq][16 * w + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Bs[buf][4 * k4 + lq][16 * j + lr], acc[j], 0, 0, 0);
    }
  };
  double xa[8], xb[8], ya[8], yb[8];
  int64_t oxa[8], oxb[8], oya[8], oyb[8];
  offsets(0, oxa, oxb);
  load(oxa, oxb, xa, xb);
  store(0, xa, xb);
  offsets(1, oya, oyb);
  load(oya, oyb, ya, yb);
  offsets(2, oxa, oxb);
  __syncthreads();
  // entering step t (even): LDS 0 holds step t, set y step t + 1 (in flight),
  // ox the offsets of step t + 2
  for (int64_t t = 0; t < nsteps; t += 2) {
    load(oxa, oxb, xa, xb);        // step t + 2
    offsets(t + 3, oya, oyb);
    compute(0);                    // step t
    if (t + 1 < nsteps) store(1, ya, yb);
    __syncthreads();
    if (t + 1 >= nsteps) break;
    load(oya, oyb, ya, yb);        // step t + 3
    offsets(t + 4, oxa, oxb);
    compute(1);                    // step t + 1
    if (t + 2 < nsteps) store(0, xa, xb);
    __syncthreads();
  }
  if (!active) return;
  double* Zs = g.Z + (int64_t)blockIdx.y * Mr * Nc;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = n0 + 16 * j + lr;
    if (col >= Nc) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = m0 + 16 * w + lq + 4 * r;
      if (row < Mr) Zs[(int64_t)row * Nc + col] = acc[j][r];
    }
  }
}

// Sum of the split partials and the scatter of gemm_scatter_kernel in one
// pass: one wave per output element (m, n); lane l sums the splits l, l + 64,
// ... in order for each of the (up to four) Z entries the element combines,
// then a butterfly over the lanes; lane 0 stores alpha * value.
__global__ __launch_bounds__(256) void dot_reduce_kernel(DotArgs d, GemmScatterArgs g) {
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= (int64_t)g.M * g.N) return;  // whole wave leaves together
  const int64_t m = e / g.N, n = e - m * g.N;
  const int pa = d.a_im ? 2 : 1, pb = d.b_im ? 2 : 1;
  const int64_t ldz = pb * (int64_t)g.N, zs = pa * (int64_t)g.M * ldz;
  double rr = 0.0, ii = 0.0, ri = 0.0, ir = 0.0;
  for (int s = lane; s < d.splits; s += 64) {
    const double* z = d.Z + (int64_t)s * zs;
    rr += z[m * ldz + n];
    if (g.kind >= 1) ir += z[(m + g.M) * ldz + n];
    if (g.kind == 2) {
      ri += z[m * ldz + g.N + n];
      ii += z[(m + g.M) * ldz + g.N + n];
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    rr += __shfl_xor(rr, off);
    ir += __shfl_xor(ir, off);
    ri += __shfl_xor(ri, off);
    ii += __shfl_xor(ii, off);
  }
  if (lane != 0) return;
  const double vr = g.kind == 2 ? rr - ii : rr;
  const double vi = g.kind == 2 ? ri + ir : ir;
  store_mirrored(g.out_re, g.out_im, g.cm[m], g.cn[n], -1, -1, g.alpha_re * vr - g.alpha_im * vi,
                 g.alpha_re * vi + g.alpha_im * vr);
}

void launch_dot(const DotArgs& d, const GemmScatterArgs& sc, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0) return;
  const int pa = d.a_im ? 2 : 1, pb = d.b_im ? 2 : 1;
  const int64_t Mr = pa * (int64_t)d.M, Nc = pb * (int64_t)d.N;
  const dim3 grid((unsigned)((Nc + DK_BN - 1) / DK_BN), (unsigned)d.splits, (unsigned)((Mr + DK_BN - 1) / DK_BN));
  const dim3 block(256);
  const int kc = (d.a_kc ? 2 : 0) + (d.b_kc ? 1 : 0);
#define TDCC_DOT_LAUNCH(PA, PB)                                                                    \
  switch (kc) {                                                                                    \
    case 3: hipLaunchKernelGGL((dot_kernel<PA, PB, true, true>), grid, block, 0, s, d); break;     \
    case 2: hipLaunchKernelGGL((dot_kernel<PA, PB, true, false>), grid, block, 0, s, d); break;    \
    case 1: hipLaunchKernelGGL((dot_kernel<PA, PB, false, true>), grid, block, 0, s, d); break;    \
    default: hipLaunchKernelGGL((dot_kernel<PA, PB, false, false>), grid, block, 0, s, d); break;  \
  }
  if (pa == 2 && pb == 2) { TDCC_DOT_LAUNCH(2, 2) }
  else if (pa == 2) { TDCC_DOT_LAUNCH(2, 1) }
  else if (pb == 2) { TDCC_DOT_LAUNCH(1, 2) }
  else { TDCC_DOT_LAUNCH(1, 1) }
#undef TDCC_DOT_LAUNCH
  const int64_t n = (int64_t)sc.M * sc.N;
  hipLaunchKernelGGL(dot_reduce_kernel, dim3((unsigned)((n + 3) / 4)), block, 0, s, d, sc);
}

}  // namespace tdcc
