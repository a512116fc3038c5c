// Adjoint kernels of the field build and the dipole read-out (gradient of a
// windowed dipole with respect to the field amplitude, frequency and dipole
// matrix).  Single-block reductions with a fixed order: deterministic.
#include "kernels.hpp"

#include <cmath>

namespace tdcc {

__global__ __launch_bounds__(256) void field_vjp_kernel(FieldVjpArgs a) {
  __shared__ double red[4][256];
  const int64_t row = *a.counter - 1;
  const double c = a.table[(row * 3 + a.stage) * 2 + 0];
  const double s = a.table[(row * 3 + a.stage) * 2 + 1];
  const double t = a.ttab[row * 3 + a.stage];
  const double tw = 2.0 * M_PI * t;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int64_t total = (int64_t)a.rows * a.cols;
  for (int64_t e = threadIdx.x; e < total; e += 256) {
    const int i = (int)(e / a.cols), j = (int)(e - (int64_t)i * a.cols);
    const int64_t pq = (int64_t)(a.p0 + i) * a.n + (a.q0 + j);
    const double br = a.ab_re[e], bi = a.ab_im[e];
    const double mu = a.mu_pol[pq];
    // Fbar * e^{i theta}
    const double er = br * c - bi * s, ei = br * s + bi * c;
    acc[0] -= mu * er;
    acc[1] -= mu * ei;
    // Fbar * (-A) mu e^{i theta} * i 2 pi t  =  (-A) mu 2 pi t * i (er + i ei)
    acc[2] += a.negA * mu * tw * (-ei);
    acc[3] += a.negA * mu * tw * er;
    a.gmu_re[pq] += a.negA * er;
    a.gmu_im[pq] += a.negA * ei;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) a.g[threadIdx.x] += red[threadIdx.x][0];
}

void launch_field_vjp(const FieldVjpArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(field_vjp_kernel, dim3(1), dim3(256), 0, s, a);
}

__global__ void seed_kernel(SeedArgs a) {
  const double w = a.wtab[*a.counter - 1];
  const int64_t total = (int64_t)a.rows * a.cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / a.cols), j = (int)(e - (int64_t)i * a.cols);
    a.out_re[e] = w * a.mu_pol[(int64_t)(a.p0 + i) * a.n + (a.q0 + j)];
    a.out_im[e] = 0.0;
  }
}

void launch_seed(const SeedArgs& a, hipStream_t s) {
  const int64_t total = (int64_t)a.rows * a.cols;
  int64_t b = (total + 255) / 256;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(seed_kernel, dim3((unsigned)b), dim3(256), 0, s, a);
}

__global__ void trace_vjp_kernel(TraceVjpArgs a) {
  const double w = a.wtab[*a.counter - 1];
  const int n = a.o + a.v;
  const int64_t n2 = (int64_t)n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n2; idx += (int64_t)gridDim.x * blockDim.x) {
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
    a.gmu_re[idx] += w * dr;
    a.gmu_im[idx] += w * di;
  }
}

void launch_trace_vjp(const TraceVjpArgs& a, hipStream_t s) {
  const int n = a.o + a.v;
  int64_t b = ((int64_t)n * n + 255) / 256;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(trace_vjp_kernel, dim3((unsigned)b), dim3(256), 0, s, a);
}

}  // namespace tdcc
