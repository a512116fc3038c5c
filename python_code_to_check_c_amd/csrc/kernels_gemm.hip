// Gather / scatter kernels that put a tensor contraction into dense GEMM form
// (gfx950).  The operands of a contraction are addressed through the same
// per-index-group offset tables as the strided contraction kernel.
#include "kernels.hpp"

namespace tdcc {

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
    const int64_t r = idx / C, c = idx - r * C;
    int64_t rb = r, cb = c, p = 0;
    if (g.planes_along_rows) {
      p = r / g.Rb;
      rb = r - p * g.Rb;
    } else {
      p = c / g.Cb;
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
    const int64_t m = idx / g.N, n = idx - m * g.N;
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

}  // namespace tdcc
