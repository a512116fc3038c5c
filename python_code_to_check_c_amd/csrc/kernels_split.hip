// Split-precision GEMM for gfx950: FP64 operands are cut into int8 digit
// planes (Ozaki scheme) and multiplied on v_mfma_i32_16x16x64_i8, whose int32
// accumulation is exact, so the only error is the truncation of the digit
// expansion.  The result is assembled in FP64.
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace tdcc {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// Row split in three launches:
//   1. partial maxima of |x| over k tiles:  pmax[kt][r]
//   2. row scales: e = ilogb(max_k |x_rk|) + 2, scale[r] = 2^e
//   3. digits: y = x 2^-e (|y| < 1/2), d_s = rint(254 y), y <- 254 y - d_s
// so every digit lies in [-127, 127] and every remainder in [-1/2, 1/2].  The
// scaling by 2^-e, rint and the subtraction of the integer part are exact in
// FP64; the product 254 y rounds at 2^-53 of |254 y| <= 127, which moves the
// represented value by about 2^-53 of the row scale, far below the truncation
// 0.5 * 254^-S.  A row with a non-finite entry gets a NaN scale and zero
// digits, so its products come out NaN.
// Tiles: along k (rows orientation) 16 rows x 256 k, a thread owning 4
// consecutive k of 4 rows; along r (columns orientation) 64 rows x 64 k, a
// thread owning 16 consecutive k of one row.  Offset tables are read once per
// thread and reused across its elements.
// ---------------------------------------------------------------------------
constexpr int SR_TR = 16, SR_TK = 256;  // rows orientation tile
constexpr int SC_TR = 64, SC_TK = 64;   // columns orientation tile

// Source of one row: element o is p[o], or p[o] + q[o] for the third plane
// of a three-plane operand (re, im, re + im).
struct RowSrc {
  const double* p;
  const double* q;
};

__device__ __forceinline__ RowSrc split_row_base(const SplitArgs& a, int64_t r) {
  const bool il = a.interleave && a.planes == 2;
  const int64_t p = il ? (r & 1) : r / a.Rb, rb = il ? (r >> 1) : r - p * a.Rb;
  const int64_t off = a.rtab ? a.rtab[rb] : rb * a.ldx;
  if (p == 2) return {a.src_re + off, a.src_im + off};
  return {(p ? a.src_im : a.src_re) + off, nullptr};
}

__device__ __forceinline__ double row_at(const RowSrc& b, int64_t o) { return b.q ? b.p[o] + b.q[o] : b.p[o]; }

__device__ __forceinline__ int64_t split_koff(const SplitArgs& a, int64_t k) { return a.ktab ? a.ktab[k] : k; }

// Address of digit s of element (r, k) (k a multiple of 4; a group of 4 or
// 16 digits starting at k stays inside one 32-wide K block).
template <int S>
__device__ __forceinline__ int8_t* digit_ptr(const SplitArgs& a, int64_t r, int64_t k, int s) {
  return a.dil ? a.D + r * a.Kp * S + (k >> 5) * (32 * S) + 32 * s + (k & 31) : a.D + s * a.plane + r * a.Kp + k;
}

__device__ __forceinline__ double abs_or_inf(double v) { return std::isfinite(v) ? fabs(v) : INFINITY; }

// The block's wave index, known to the compiler as uniform across the wave:
// in the column-oriented kernels a wave's lanes are rows and its k range
// depends on the wave only, so the k offset table is read with scalar loads
// (one per k for the wave) instead of a vector load per lane (a split of
// 51,200 rows x K = 190 through a pair table: 72 -> 36 us).
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// x[k .. k+3] of a row (0 past K): two 16-byte loads when the row's k are
// contiguous (no k table) and the address is 16-byte aligned, else one load
// per element through the k table.
__device__ __forceinline__ void load_row4(const SplitArgs& a, const RowSrc& base, int64_t k, double v[4],
                                          int64_t kend = -1) {
  if (kend < 0) kend = a.K;
  if (!a.ktab && k + 3 < kend && (reinterpret_cast<uintptr_t>(base.p + k) & 15) == 0 &&
      (!base.q || (reinterpret_cast<uintptr_t>(base.q + k) & 15) == 0)) {
    const double2 x = *reinterpret_cast<const double2*>(base.p + k);
    const double2 y = *reinterpret_cast<const double2*>(base.p + k + 2);
    v[0] = x.x;
    v[1] = x.y;
    v[2] = y.x;
    v[3] = y.y;
    if (base.q) {
      const double2 x2 = *reinterpret_cast<const double2*>(base.q + k);
      const double2 y2 = *reinterpret_cast<const double2*>(base.q + k + 2);
      v[0] += x2.x;
      v[1] += x2.y;
      v[2] += y2.x;
      v[3] += y2.y;
    }
  } else if (a.ktab && k + 3 < kend && (reinterpret_cast<uintptr_t>(a.ktab + k) & 15) == 0) {
    // the four offsets with two 16-byte loads
    const longlong2 o01 = *reinterpret_cast<const longlong2*>(a.ktab + k);
    const longlong2 o23 = *reinterpret_cast<const longlong2*>(a.ktab + k + 2);
    v[0] = row_at(base, o01.x);
    v[1] = row_at(base, o01.y);
    v[2] = row_at(base, o23.x);
    v[3] = row_at(base, o23.y);
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = k + q < kend ? row_at(base, split_koff(a, k + q)) : 0.0;
  }
}


// next digit of 4 values, packed little-endian into 4 bytes
__device__ __forceinline__ uint32_t next_digit4(double v[4]) {
  uint32_t packed = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double y = v[j] * kOzRadix;
    const double d = rint(y);
    v[j] = y - d;
    packed |= (uint32_t)(uint8_t)(int8_t)(int)d << (8 * j);
  }
  return packed;
}

__global__ __launch_bounds__(256) void split_max_rows_kernel(SplitArgs a) {
  const int64_t R = (int64_t)a.planes * a.Rb;
  const int q = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int64_t k0 = (int64_t)blockIdx.y * SR_TK + 4 * q;
  int64_t ko[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) ko[j] = k0 + j < a.K ? split_koff(a, k0 + j) : -1;
#pragma unroll
  for (int i = 0; i < SR_TR / 4; ++i) {
    const int64_t r = (int64_t)blockIdx.x * SR_TR + rg + 4 * i;
    double m = 0.0;
    if (r < R) {
      const RowSrc base = split_row_base(a, r);
      if (!a.ktab) {
        double v[4];
        load_row4(a, base, k0, v);
#pragma unroll
        for (int j = 0; j < 4; ++j) m = fmax(m, abs_or_inf(v[j]));
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (ko[j] >= 0) m = fmax(m, abs_or_inf(row_at(base, ko[j])));
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (q == 0 && r < R) a.pmax[(int64_t)blockIdx.y * R + r] = m;
  }
}

__global__ __launch_bounds__(256) void split_max_cols_kernel(SplitArgs a) {
  __shared__ double red[4][SC_TR];
  const int64_t R = (int64_t)a.planes * a.Rb;
  const int rl = threadIdx.x & 63, kg = wave_index();
  const int64_t r = (int64_t)blockIdx.x * SC_TR + rl;
  const int64_t k0 = (int64_t)blockIdx.y * SC_TK + 16 * kg;
  double m = 0.0;
  if (r < R) {
    const RowSrc base = split_row_base(a, r);
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (k0 + j < a.K) m = fmax(m, abs_or_inf(row_at(base, split_koff(a, k0 + j))));
  }
  red[kg][rl] = m;
  __syncthreads();
  if (kg == 0 && r < R)
    a.pmax[(int64_t)blockIdx.y * R + r] = fmax(fmax(red[0][rl], red[1][rl]), fmax(red[2][rl], red[3][rl]));
}

// Row scales from the partial maxima.  Few k tiles: one thread per row.
// Many k tiles (long rows): one 256-thread block per row.
__global__ __launch_bounds__(256) void split_scale_kernel(SplitArgs a, int64_t ktiles) {
  const int64_t R = (int64_t)a.planes * a.Rb;
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  double m = 0.0;
  for (int64_t t = 0; t < ktiles; ++t) m = fmax(m, a.pmax[t * R + r]);
  a.scale[r] = std::isfinite(m) ? ldexp(1.0, m > 0.0 ? ilogb(m) + 2 : 0) : NAN;
}

__global__ __launch_bounds__(256) void split_scale_long_kernel(SplitArgs a, int64_t ktiles) {
  __shared__ double red[4];
  const int64_t R = (int64_t)a.planes * a.Rb;
  const int64_t r = blockIdx.x;
  double m = 0.0;
  for (int64_t t = threadIdx.x; t < ktiles; t += 256) m = fmax(m, a.pmax[t * R + r]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    a.scale[r] = std::isfinite(m) ? ldexp(1.0, m > 0.0 ? ilogb(m) + 2 : 0) : NAN;
  }
}

// Rows of at most 256 digits: one wave per row, 4 rows per block, thread
// (lane l) holds k = 4 l + (0..3).
template <int S>
__global__ __launch_bounds__(256) void split_rows_wave_kernel(SplitArgs a) {
  const int64_t R = (int64_t)a.planes * a.Rb;
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;  // whole wave leaves together
  const RowSrc base = split_row_base(a, r);
  double v[4];
  double m = 0.0;
  load_row4(a, base, 4 * lane, v);
#pragma unroll
  for (int q = 0; q < 4; ++q) m = fmax(m, abs_or_inf(v[q]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  const bool finite = std::isfinite(m);
  const int e = finite ? (m > 0.0 ? ilogb(m) + 2 : 0) : 0;
  if (lane == 0) a.scale[r] = finite ? ldexp(1.0, e) : NAN;
  const int64_t k = 4 * lane;
  if (k >= a.Kp) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = finite ? ldexp(v[q], -e) : 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) *reinterpret_cast<uint32_t*>(digit_ptr<S>(a, r, k, s)) = next_digit4(v);
}

template <int S>
__global__ __launch_bounds__(256) void split_digits_rows_kernel(SplitArgs a) {
  const int64_t R = (int64_t)a.planes * a.Rb;
  const int q = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int64_t k0 = (int64_t)blockIdx.y * SR_TK + 4 * q;
  if (k0 >= a.Kp) return;
  int64_t ko[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) ko[j] = k0 + j < a.K ? split_koff(a, k0 + j) : -1;
#pragma unroll
  for (int i = 0; i < SR_TR / 4; ++i) {
    const int64_t r = (int64_t)blockIdx.x * SR_TR + rg + 4 * i;
    if (r >= R) break;
    const double sc = a.scale[r];
    const bool finite = std::isfinite(sc);
    const int e = finite ? ilogb(sc) : 0;
    const RowSrc base = split_row_base(a, r);
    double v[4];
    if (!a.ktab) {
      load_row4(a, base, k0, v);
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = finite ? ldexp(v[j], -e) : 0.0;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = (finite && ko[j] >= 0) ? ldexp(row_at(base, ko[j]), -e) : 0.0;
    }
#pragma unroll
    for (int s = 0; s < S; ++s) *reinterpret_cast<uint32_t*>(digit_ptr<S>(a, r, k0, s)) = next_digit4(v);
  }
}

template <int S>
__global__ __launch_bounds__(256) void split_digits_cols_kernel(SplitArgs a) {
  const int64_t R = (int64_t)a.planes * a.Rb;
  const int rl = threadIdx.x & 63, kg = wave_index();
  const int64_t r = (int64_t)blockIdx.x * SC_TR + rl;
  const int64_t k0 = (int64_t)blockIdx.y * SC_TK + 16 * kg;
  if (r >= R) return;
  const double sc = a.scale[r];
  const bool finite = std::isfinite(sc);
  const int e = finite ? ilogb(sc) : 0;
  const RowSrc base = split_row_base(a, r);
  uint32_t out[4][S];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t k = k0 + 4 * q + j;
      v[j] = (finite && k < a.K) ? ldexp(row_at(base, split_koff(a, k)), -e) : 0.0;
    }
#pragma unroll
    for (int s = 0; s < S; ++s) out[q][s] = next_digit4(v);
  }
#pragma unroll
  for (int s = 0; s < S; ++s)
    *reinterpret_cast<uint4*>(digit_ptr<S>(a, r, k0, s)) = make_uint4(out[0][s], out[1][s], out[2][s], out[3][s]);
}

// Single-pass splits for rows that fit in registers: the block reads its
// rows once, reduces the maximum, and writes the digits from registers.
// Rows orientation: one row per block, thread t holds k = 1024 j + 4 t + (0..3)
// for j < NJ (Kp <= 1024 NJ).  Columns orientation: 64 rows per block, thread
// (row r, phase g) holds k = 64 j + 16 g + (0..15) for j < NJ (Kp <= 64 NJ).
template <int S, int NJ>
__global__ __launch_bounds__(256) void split_rows_resident_kernel(SplitArgs a) {
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t r = blockIdx.x;
  // K chunk blockIdx.y of its own scale (sk_chunk > 0), else the whole row
  const int64_t klo = (int64_t)blockIdx.y * a.sk_chunk;
  const int64_t khi = a.sk_chunk > 0 ? min(a.K, klo + a.sk_chunk) : a.K;
  const int64_t kphi = a.sk_chunk > 0 ? min(a.Kp, klo + a.sk_chunk) : a.Kp;
  const RowSrc base = split_row_base(a, r);
  double v[NJ][4];
  double m = 0.0;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    load_row4(a, base, klo + 1024 * (int64_t)j + 4 * tid, v[j], khi);
#pragma unroll
    for (int q = 0; q < 4; ++q) m = fmax(m, abs_or_inf(v[j][q]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if (lane == 0) red[w] = m;
  __syncthreads();
  m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const bool finite = std::isfinite(m);
  const int e = finite ? (m > 0.0 ? ilogb(m) + 2 : 0) : 0;
  if (tid == 0) a.scale[(int64_t)blockIdx.y * a.planes * a.Rb + r] = finite ? ldexp(1.0, e) : NAN;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int64_t k = klo + 1024 * (int64_t)j + 4 * tid;
    if (k >= kphi) break;
#pragma unroll
    for (int q = 0; q < 4; ++q) v[j][q] = finite ? ldexp(v[j][q], -e) : 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) *reinterpret_cast<uint32_t*>(digit_ptr<S>(a, r, k, s)) = next_digit4(v[j]);
  }
}

template <int S, int NJ>
__global__ __launch_bounds__(256) void split_cols_resident_kernel(SplitArgs a) {
  __shared__ double red[4][SC_TR];
  const int64_t R = (int64_t)a.planes * a.Rb;
  const int rl = threadIdx.x & 63, g = wave_index();
  const int64_t r = (int64_t)blockIdx.x * SC_TR + rl;
  const bool live = r < R;
  const RowSrc base = live ? split_row_base(a, r) : RowSrc{nullptr, nullptr};
  double v[NJ][16];
  double m = 0.0;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int64_t k = 64 * (int64_t)j + 16 * g + q;
      v[j][q] = (live && k < a.K) ? row_at(base, split_koff(a, k)) : 0.0;
      m = fmax(m, abs_or_inf(v[j][q]));
    }
  red[g][rl] = m;
  __syncthreads();
  if (!live) return;
  m = fmax(fmax(red[0][rl], red[1][rl]), fmax(red[2][rl], red[3][rl]));
  const bool finite = std::isfinite(m);
  const int e = finite ? (m > 0.0 ? ilogb(m) + 2 : 0) : 0;
  if (g == 0) a.scale[r] = finite ? ldexp(1.0, e) : NAN;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int64_t k0 = 64 * (int64_t)j + 16 * g;
    if (k0 >= a.Kp) break;
    uint32_t out[4][S];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double x[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = finite ? ldexp(v[j][4 * q + i], -e) : 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s) out[q][s] = next_digit4(x);
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
      *reinterpret_cast<uint4*>(digit_ptr<S>(a, r, k0, s)) = make_uint4(out[0][s], out[1][s], out[2][s], out[3][s]);
  }
}

// Three-plane operands (re, im, re + im of a 3-product complex GEMM) split
// from one read of each base row: the kernels below take base row rb of both
// source planes once and write rows rb, Rb + rb and 2 Rb + rb (re, im and
// re + im, the sum formed as row_at forms it), with the digits, scales and
// maxima the per-plane kernels give.
template <int S, int NJ>
__global__ __launch_bounds__(256) void split_rows_resident3_kernel(SplitArgs a) {
  __shared__ double red[3][4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t rb = blockIdx.x;
  const RowSrc b0 = split_row_base(a, rb), b1 = split_row_base(a, a.Rb + rb);
  double x[NJ][4], y[NJ][4];
  double m[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int64_t k = 1024 * (int64_t)j + 4 * tid;
    load_row4(a, b0, k, x[j]);
    load_row4(a, b1, k, y[j]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      m[0] = fmax(m[0], abs_or_inf(x[j][q]));
      m[1] = fmax(m[1], abs_or_inf(y[j][q]));
      m[2] = fmax(m[2], abs_or_inf(x[j][q] + y[j][q]));
    }
  }
#pragma unroll
  for (int p = 0; p < 3; ++p) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m[p] = fmax(m[p], __shfl_xor(m[p], off));
    if (lane == 0) red[p][w] = m[p];
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const double mp = fmax(fmax(red[p][0], red[p][1]), fmax(red[p][2], red[p][3]));
    const bool finite = std::isfinite(mp);
    const int e = finite ? (mp > 0.0 ? ilogb(mp) + 2 : 0) : 0;
    const int64_t r = p * a.Rb + rb;
    if (tid == 0) a.scale[r] = finite ? ldexp(1.0, e) : NAN;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int64_t k = 1024 * (int64_t)j + 4 * tid;
      if (k >= a.Kp) break;
      double v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double z = p == 0 ? x[j][q] : (p == 1 ? y[j][q] : x[j][q] + y[j][q]);
        v[q] = finite ? ldexp(z, -e) : 0.0;
      }
#pragma unroll
      for (int s = 0; s < S; ++s) *reinterpret_cast<uint32_t*>(digit_ptr<S>(a, r, k, s)) = next_digit4(v);
    }
  }
}

__global__ __launch_bounds__(256) void split_max_cols3_kernel(SplitArgs a) {
  __shared__ double red[3][4][SC_TR];
  const int64_t R = 3 * a.Rb;
  const int rl = threadIdx.x & 63, kg = wave_index();
  const int64_t rb = (int64_t)blockIdx.x * SC_TR + rl;
  const int64_t k0 = (int64_t)blockIdx.y * SC_TK + 16 * kg;
  double m[3] = {0.0, 0.0, 0.0};
  if (rb < a.Rb) {
    const RowSrc b0 = split_row_base(a, rb), b1 = split_row_base(a, a.Rb + rb);
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (k0 + j < a.K) {
        const int64_t o = split_koff(a, k0 + j);
        const double xr = row_at(b0, o), xi = row_at(b1, o);
        m[0] = fmax(m[0], abs_or_inf(xr));
        m[1] = fmax(m[1], abs_or_inf(xi));
        m[2] = fmax(m[2], abs_or_inf(xr + xi));
      }
  }
#pragma unroll
  for (int p = 0; p < 3; ++p) red[p][kg][rl] = m[p];
  __syncthreads();
  if (kg == 0 && rb < a.Rb)
#pragma unroll
    for (int p = 0; p < 3; ++p)
      a.pmax[(int64_t)blockIdx.y * R + p * a.Rb + rb] =
          fmax(fmax(red[p][0][rl], red[p][1][rl]), fmax(red[p][2][rl], red[p][3][rl]));
}

template <int S>
__global__ __launch_bounds__(256) void split_digits_cols3_kernel(SplitArgs a) {
  const int rl = threadIdx.x & 63, kg = wave_index();
  const int64_t rb = (int64_t)blockIdx.x * SC_TR + rl;
  const int64_t k0 = (int64_t)blockIdx.y * SC_TK + 16 * kg;
  if (rb >= a.Rb) return;
  const RowSrc b0 = split_row_base(a, rb), b1 = split_row_base(a, a.Rb + rb);
  double x[16], y[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int64_t k = k0 + j;
    const int64_t o = k < a.K ? split_koff(a, k) : 0;
    x[j] = k < a.K ? row_at(b0, o) : 0.0;
    y[j] = k < a.K ? row_at(b1, o) : 0.0;
  }
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int64_t r = p * a.Rb + rb;
    const double sc = a.scale[r];
    const bool finite = std::isfinite(sc);
    const int e = finite ? ilogb(sc) : 0;
    uint32_t out[4][S];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int jj = 4 * q + j;
        const double z = p == 0 ? x[jj] : (p == 1 ? y[jj] : x[jj] + y[jj]);
        v[j] = (finite && k0 + jj < a.K) ? ldexp(z, -e) : 0.0;
      }
#pragma unroll
      for (int s = 0; s < S; ++s) out[q][s] = next_digit4(v);
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
      *reinterpret_cast<uint4*>(digit_ptr<S>(a, r, k0, s)) = make_uint4(out[0][s], out[1][s], out[2][s], out[3][s]);
  }
}

// TDCC_SPLIT3_FUSED=0: three-plane operands split plane by plane (the third
// plane reading both source planes again).
static bool split3_fused_on() {
  const char* e = std::getenv("TDCC_SPLIT3_FUSED");
  return !e || std::atoi(e) != 0;
}

int64_t ozaki_split_workspace(int64_t R, int64_t Kp, bool by_columns) {
  const int64_t tk = by_columns ? SC_TK : SR_TK;
  return ((Kp + tk - 1) / tk) * R * (int64_t)sizeof(double);
}

// Column-oriented short-K operands up to this many rows (planes counted)
// take the one-wave-per-row kernel; TDCC_WAVE_COLS_MAX_ROWS overrides 2048.
static int64_t wave_cols_max_rows() {
  const char* e = std::getenv("TDCC_WAVE_COLS_MAX_ROWS");
  return e ? std::atoll(e) : 2048;
}

template <int S>
static void split_launch(const SplitArgs& a, bool by_columns, hipStream_t s) {
  const int64_t R = (int64_t)a.planes * a.Rb;
  const dim3 block(256);
  // Short rows: one wave per row.  Column-oriented operands with few rows go
  // here too: 64-row column tiles would give them a handful of blocks whose
  // serial loads dominate, while the whole operand is small enough that
  // uncoalesced loads cost little.
  if (a.sk_chunk == 0 && !a.multipass && a.Kp <= 256 && (!by_columns || R <= wave_cols_max_rows())) {
    hipLaunchKernelGGL(split_rows_wave_kernel<S>, dim3((unsigned)((R + 3) / 4)), block, 0, s, a);
    return;
  }
  const bool three = a.planes == 3 && !a.interleave && a.sk_chunk == 0 && !a.multipass && split3_fused_on();
  if (three && !by_columns && a.Kp <= 16384) {
    const dim3 grid((unsigned)a.Rb);
    if (a.Kp <= 1024) hipLaunchKernelGGL((split_rows_resident3_kernel<S, 1>), grid, block, 0, s, a);
    else if (a.Kp <= 2048) hipLaunchKernelGGL((split_rows_resident3_kernel<S, 2>), grid, block, 0, s, a);
    else if (a.Kp <= 4096) hipLaunchKernelGGL((split_rows_resident3_kernel<S, 4>), grid, block, 0, s, a);
    else if (a.Kp <= 8192) hipLaunchKernelGGL((split_rows_resident3_kernel<S, 8>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((split_rows_resident3_kernel<S, 16>), grid, block, 0, s, a);
    return;
  }
  if (three && by_columns && a.Kp > 256) {
    const int64_t kt = (a.Kp + SC_TK - 1) / SC_TK;
    const dim3 grid((unsigned)((a.Rb + SC_TR - 1) / SC_TR), (unsigned)kt);
    hipLaunchKernelGGL(split_max_cols3_kernel, grid, block, 0, s, a);
    if (kt > 64) hipLaunchKernelGGL(split_scale_long_kernel, dim3((unsigned)R), block, 0, s, a, kt);
    else hipLaunchKernelGGL(split_scale_kernel, dim3((unsigned)((R + 255) / 256)), block, 0, s, a, kt);
    hipLaunchKernelGGL(split_digits_cols3_kernel<S>, grid, block, 0, s, a);
    return;
  }
  // rows in one pass: whole rows up to 16384, or K chunks with their own
  // scales (always single pass; the chunk is at most 16384)
  if (a.sk_chunk > 0 || (!a.multipass && !by_columns && a.Kp <= 16384)) {
    const int64_t len = a.sk_chunk > 0 ? a.sk_chunk : a.Kp;
    const dim3 grid((unsigned)R, (unsigned)(a.sk_chunk > 0 ? (a.Kp + a.sk_chunk - 1) / a.sk_chunk : 1));
    if (len <= 1024) hipLaunchKernelGGL((split_rows_resident_kernel<S, 1>), grid, block, 0, s, a);
    else if (len <= 2048) hipLaunchKernelGGL((split_rows_resident_kernel<S, 2>), grid, block, 0, s, a);
    else if (len <= 4096) hipLaunchKernelGGL((split_rows_resident_kernel<S, 4>), grid, block, 0, s, a);
    else if (len <= 8192) hipLaunchKernelGGL((split_rows_resident_kernel<S, 8>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((split_rows_resident_kernel<S, 16>), grid, block, 0, s, a);
    return;
  }
  if (!a.multipass && by_columns && a.Kp <= 256) {
    const dim3 grid((unsigned)((R + SC_TR - 1) / SC_TR));
    if (a.Kp <= 64) hipLaunchKernelGGL((split_cols_resident_kernel<S, 1>), grid, block, 0, s, a);
    else if (a.Kp <= 128) hipLaunchKernelGGL((split_cols_resident_kernel<S, 2>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((split_cols_resident_kernel<S, 4>), grid, block, 0, s, a);
    return;
  }
  const int64_t tr = by_columns ? SC_TR : SR_TR, tk = by_columns ? SC_TK : SR_TK;
  const int64_t kt = (a.Kp + tk - 1) / tk;
  const dim3 grid((unsigned)((R + tr - 1) / tr), (unsigned)kt);
  if (by_columns) hipLaunchKernelGGL(split_max_cols_kernel, grid, block, 0, s, a);
  else hipLaunchKernelGGL(split_max_rows_kernel, grid, block, 0, s, a);
  if (kt > 64) hipLaunchKernelGGL(split_scale_long_kernel, dim3((unsigned)R), block, 0, s, a, kt);
  else hipLaunchKernelGGL(split_scale_kernel, dim3((unsigned)((R + 255) / 256)), block, 0, s, a, kt);
  if (by_columns) hipLaunchKernelGGL(split_digits_cols_kernel<S>, grid, block, 0, s, a);
  else hipLaunchKernelGGL(split_digits_rows_kernel<S>, grid, block, 0, s, a);
}

void launch_ozaki_split(const SplitArgs& a, bool by_columns, hipStream_t s) {
  if (a.planes * a.Rb <= 0) return;
  switch (a.S) {
    case 2: split_launch<2>(a, by_columns, s); break;
    case 3: split_launch<3>(a, by_columns, s); break;
    case 4: split_launch<4>(a, by_columns, s); break;
    case 5: split_launch<5>(a, by_columns, s); break;
    case 6: split_launch<6>(a, by_columns, s); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// Digit-pair GEMM.  Tile 128 x (32 JB) x 64, 256 threads = 4 waves (2 x 2);
// a wave owns 64 x 16 JB = 4 x JB MFMA blocks of 16 x 16 per digit group.
// K step = 64 = one v_mfma_i32_16x16x64_i8.  Lane l of a wave supplies
// A[row l%16][k 16(l/16) .. +15] and B[col l%16][same k] as 16 bytes; the
// int32 result sits at D[4(l/16) + r][l%16] (measured on gfx950).
// LDS rows are 64 bytes with the 16-byte chunk index XOR-swizzled by
// (row >> 1) & 3, so the 16-byte reads of 8 consecutive lanes hit 8 distinct
// bank groups.  Global -> registers -> LDS double buffering, one barrier per
// K step.
// ---------------------------------------------------------------------------
constexpr int OZ_BM = 128, OZ_BK = 64;

__device__ __forceinline__ int oz_swz(int row, int kc) { return row * 4 + (kc ^ ((row >> 1) & 3)); }

template <int S, int JB>
__global__ __launch_bounds__(256) void ozgemm_kernel(OzGemmArgs g) {
  constexpr int BN = 32 * JB;
  constexpr int CA = OZ_BM * 4 / 256, CB = BN * 4 / 256;  // 16-byte chunks per thread and digit
  __shared__ v4i As[2][S][OZ_BM * 4];
  __shared__ v4i Bs[2][S][BN * 4];
  __shared__ double sA[OZ_BM], sB[BN];  // row and column scales of the tile
  const int tiles_m = (g.M + OZ_BM - 1) / OZ_BM, tiles_n = (g.N + BN - 1) / BN;
  const int L = remap_tile(blockIdx.x, tiles_m * tiles_n);
  const int tm = L % tiles_m, tn = L / tiles_m;
  const int m0 = tm * OZ_BM, n0 = tn * BN;
  const int tid = threadIdx.x;
  const int kb = g.k_chunk > 0 ? (int)blockIdx.y * g.k_chunk : 0;
  const int ke = g.k_chunk > 0 ? min(g.Kp, kb + g.k_chunk) : g.Kp;
  double* const Cout = g.C + (int64_t)blockIdx.y * g.c_split_stride;

  v4i ra[S][CA], rb[S][CB];
  auto load_global = [&](int k0) {
#pragma unroll
    for (int h = 0; h < CA; ++h) {
      const int c = tid + 256 * h, row = c >> 2, kc = c & 3;
      const int m = m0 + row;
#pragma unroll
      for (int s = 0; s < S; ++s)
        ra[s][h] = m < g.M ? *reinterpret_cast<const v4i*>(g.A + s * g.a_plane + (int64_t)m * g.Kp + k0 + kc * 16)
                           : v4i{0, 0, 0, 0};
    }
#pragma unroll
    for (int h = 0; h < CB; ++h) {
      const int c = tid + 256 * h, row = c >> 2, kc = c & 3;
      const int n = n0 + row;
#pragma unroll
      for (int s = 0; s < S; ++s)
        rb[s][h] = n < g.N ? *reinterpret_cast<const v4i*>(g.B + s * g.b_plane + (int64_t)n * g.Kp + k0 + kc * 16)
                           : v4i{0, 0, 0, 0};
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int h = 0; h < CA; ++h) {
      const int c = tid + 256 * h;
#pragma unroll
      for (int s = 0; s < S; ++s) As[buf][s][oz_swz(c >> 2, c & 3)] = ra[s][h];
    }
#pragma unroll
    for (int h = 0; h < CB; ++h) {
      const int c = tid + 256 * h;
#pragma unroll
      for (int s = 0; s < S; ++s) Bs[buf][s][oz_swz(c >> 2, c & 3)] = rb[s][h];
    }
  };

  v4i acc[S][4][JB];
#pragma unroll
  for (int q = 0; q < S; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < JB; ++j) acc[q][i][j] = v4i{0, 0, 0, 0};
  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * (16 * JB);
  const int lr = lane & 15, lk = lane >> 4;

  load_global(kb);
  if (tid < OZ_BM) sA[tid] = m0 + tid < g.M ? g.sa[(int64_t)blockIdx.y * g.sa_split_stride + m0 + tid] : 0.0;
  else if (tid - OZ_BM < BN) sB[tid - OZ_BM] = n0 + tid - OZ_BM < g.N ? g.sb[(int64_t)blockIdx.y * g.sb_split_stride + n0 + tid - OZ_BM] : 0.0;
  store_lds(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = kb; k0 < ke; k0 += OZ_BK) {
    const bool more = k0 + OZ_BK < ke;
    if (more) load_global(k0 + OZ_BK);
    v4i a[S][4];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int i = 0; i < 4; ++i) a[s][i] = As[buf][s][oz_swz(wm + i * 16 + lr, lk)];
#pragma unroll
    for (int t = 0; t < S; ++t) {
      v4i b[JB];
#pragma unroll
      for (int j = 0; j < JB; ++j) b[j] = Bs[buf][t][oz_swz(wn + j * 16 + lr, lk)];
#pragma unroll
      for (int s = 0; s + t < S; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < JB; ++j)
            acc[s + t][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[s][i], b[j], acc[s + t][i][j], 0, 0, 0);
    }
    if (more) store_lds(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // digit group q (pairs with s + t = q + 2, 1-based digits) weighs 254^-(q+2)
  if (g.epi < 0) {
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      const int col = n0 + wn + j * 16 + lr;
      if (col >= g.N) continue;
      const double sbc = sB[wn + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + i * 16 + 4 * lk + r;
          if (row >= g.M) continue;
          double v = 0.0;
#pragma unroll
          for (int q = S - 1; q >= 0; --q) v += (double)acc[q][i][j][r] * oz_group_weight(q);
          Cout[(int64_t)row * g.ldc + col] = v * (sA[row - m0] * sbc);
        }
    }
    return;
  }
  // Fused scatter.  A lane's four rows 4 lk + r start at an even row, so the
  // interleaved (re, im) rows 2m, 2m+1 of A sit in one lane (r = 0,1 and
  // 2,3); the interleaved columns 2n, 2n+1 of B sit in lanes lr and lr ^ 1.
  const double ar = g.alpha_re, ai = g.alpha_im;
#pragma unroll
  for (int j = 0; j < JB; ++j) {
    const int col = n0 + wn + j * 16 + lr;
    const bool colok = col < g.N;
    const double sbc = sB[wn + j * 16 + lr];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row0 = m0 + wm + i * 16 + 4 * lk;
      double z[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double v = 0.0;
#pragma unroll
        for (int q = S - 1; q >= 0; --q) v += (double)acc[q][i][j][r] * oz_group_weight(q);
        z[r] = (colok && row0 + r < g.M) ? v * (sA[row0 + r - m0] * sbc) : 0.0;
      }
      if (g.epi == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (colok && row0 + r < g.M)
            store_mirrored(g.out_re, nullptr, g.cm[row0 + r], g.cn[col], g.cm2 ? g.cm2[row0 + r] : -1,
                           g.cn2 ? g.cn2[col] : -1, ar * z[r], 0.0);
      } else if (g.epi == 1) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (!colok || row0 + 2 * h >= g.M) continue;
          const double re = z[2 * h], im = z[2 * h + 1];
          const int mr = (row0 >> 1) + h;
          store_mirrored(g.out_re, g.out_im, g.cm[mr], g.cn[col], g.cm2 ? g.cm2[mr] : -1, g.cn2 ? g.cn2[col] : -1,
                         ar * re - ai * im, ar * im + ai * re);
        }
      } else {
        double w[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) w[r] = __shfl_xor(z[r], 1);
        if ((lr & 1) == 0 && colok) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if (row0 + 2 * h >= g.M) continue;
            const double re = z[2 * h] - w[2 * h + 1];   // rr - ii
            const double im = w[2 * h] + z[2 * h + 1];   // ri + ir
            const int mr = (row0 >> 1) + h, nc = col >> 1;
            store_mirrored(g.out_re, g.out_im, g.cm[mr], g.cn[nc], g.cm2 ? g.cm2[mr] : -1, g.cn2 ? g.cn2[nc] : -1,
                           ar * re - ai * im, ar * im + ai * re);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Digit-pair GEMM for up to 4 digits on v_mfma_i32_32x32x32_i8: tile 128 x 128
// x 64, 512 threads = 8 waves (2 x 4), two per SIMD so that one wave's LDS
// and memory waits overlap the other's MFMAs; a wave owns 64 x 32 = 2 MFMA
// blocks of 32 x 32 per digit group, accumulators in VGPRs.  Lane l supplies
// A[row l%32][k 16(l/32) .. +15] of a 32-wide K slice (two slices per K step)
// and holds D[8(r/4) + 4(l/32) + r%4][l%32] in accumulator element r
// (checked against the 16x16x64 kernel's output).  Operands are read with
// buffer loads from rows clamped to the matrix (the extra rows only reach
// outputs that are not stored); LDS layout and swizzle as above.
// With few K steps per tile the blocks are persistent: block b takes tiles b,
// b + gridDim.x, ... (the same XCD for a grid that is a multiple of 8), and
// the first K step of its next tile is loaded while the epilogue of the
// current one stores (at 3 K steps per tile 8-17 % faster than one block per
// tile).
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(512, 1) void ozgemm8_kernel(OzGemmArgs g) {
  constexpr int BN = 128;
  __shared__ v4i As[2][S][OZ_BM * 4];
  __shared__ v4i Bs[2][S][BN * 4];
  __shared__ double sA[OZ_BM], sB[BN];
  const int tiles_m = (g.M + OZ_BM - 1) / OZ_BM, tiles_n = (g.N + BN - 1) / BN;
  const int tiles = tiles_m * tiles_n;
  const int tid = threadIdx.x;
  const int kb = g.k_chunk > 0 ? (int)blockIdx.y * g.k_chunk : 0;
  const int ke = g.k_chunk > 0 ? min(g.Kp, kb + g.k_chunk) : g.Kp;
  double* const Cout = g.C + (int64_t)blockIdx.y * g.c_split_stride;
  const __amdgpu_buffer_rsrc_t rA =
      __builtin_amdgcn_make_buffer_rsrc((void*)g.A, (short)0, (int)(S * g.a_plane), 0x00020000);
  const __amdgpu_buffer_rsrc_t rB =
      __builtin_amdgcn_make_buffer_rsrc((void*)g.B, (short)0, (int)(S * g.b_plane), 0x00020000);
  // tile origin and the thread's 16-byte chunk of A and of B per digit
  int m0 = 0, n0 = 0, va = 0, vb = 0;
  auto setup = [&](int t) {
    const int L = remap_tile(t, tiles);
    int tm, tn;
    grouped_tile(L, tiles_m, tiles_n, g.tile_group, tm, tn);
    m0 = tm * OZ_BM;
    n0 = tn * BN;
    va = min(m0 + (tid >> 2), g.M - 1) * g.Kp + (tid & 3) * 16;
    vb = min(n0 + (tid >> 2), g.N - 1) * g.Kp + (tid & 3) * 16;
  };
  v4i ra[S], rb[S];
  auto load_global = [&](int k0) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      ra[s] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rA, va, (int)(s * g.a_plane) + k0, 0));
      rb[s] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rB, vb, (int)(s * g.b_plane) + k0, 0));
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      As[buf][s][oz_swz(tid >> 2, tid & 3)] = ra[s];
      Bs[buf][s][oz_swz(tid >> 2, tid & 3)] = rb[s];
    }
  };
  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 32;
  const int lr = lane & 31, lh = lane >> 5;
  const double ar = g.alpha_re, ai = g.alpha_im;

  int t = blockIdx.x;
  setup(t);
  load_global(kb);
  for (;;) {
    v16i acc[S][2];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[q][i][e] = 0;
    if (tid < OZ_BM) sA[tid] = m0 + tid < g.M ? g.sa[(int64_t)blockIdx.y * g.sa_split_stride + m0 + tid] : 0.0;
    else if (tid - OZ_BM < BN) sB[tid - OZ_BM] = n0 + tid - OZ_BM < g.N ? g.sb[(int64_t)blockIdx.y * g.sb_split_stride + n0 + tid - OZ_BM] : 0.0;
    store_lds(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += OZ_BK) {
      const bool more = k0 + OZ_BK < ke;
      if (more) load_global(k0 + OZ_BK);
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        v4i a[S][2], b[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
          for (int i = 0; i < 2; ++i) a[s][i] = As[buf][s][oz_swz(wm + i * 32 + lr, 2 * kk + lh)];
          b[s] = Bs[buf][s][oz_swz(wn + lr, 2 * kk + lh)];
        }
#pragma unroll
        for (int tt = 0; tt < S; ++tt)
#pragma unroll
          for (int s = 0; s + tt < S; ++s)
#pragma unroll
            for (int i = 0; i < 2; ++i)
              acc[s + tt][i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[s][i], b[tt], acc[s + tt][i], 0, 0, 0);
      }
      if (more) store_lds(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
    // this tile's origin; the next tile's first K step goes in flight now
    const int em0 = m0, en0 = n0;
    const int tnext = t + (int)gridDim.x;
    if (tnext < tiles) {
      setup(tnext);
      load_global(kb);
    }
    // digit group q (pairs with s + t = q + 2, 1-based digits) weighs 254^-(q+2)
    const int col = en0 + wn + lr;
    const bool colok = col < g.N;
    const double sbc = sB[wn + lr];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int gq = 0; gq < 4; ++gq) {
        // rows row0 .. row0 + 3 of this lane; row0 is even
        const int row0 = em0 + wm + i * 32 + 8 * gq + 4 * lh;
        double z[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double v = 0.0;
#pragma unroll
          for (int q = S - 1; q >= 0; --q) v += (double)acc[q][i][4 * gq + r] * oz_group_weight(q);
          z[r] = (colok && row0 + r < g.M) ? v * (sA[row0 + r - em0] * sbc) : 0.0;
        }
        if (g.epi < 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (colok && row0 + r < g.M) Cout[(int64_t)(row0 + r) * g.ldc + col] = z[r];
        } else if (g.epi == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (colok && row0 + r < g.M)
              store_mirrored(g.out_re, nullptr, g.cm[row0 + r], g.cn[col], g.cm2 ? g.cm2[row0 + r] : -1,
                             g.cn2 ? g.cn2[col] : -1, ar * z[r], 0.0);
        } else if (g.epi == 1) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if (!colok || row0 + 2 * h >= g.M) continue;
            const double re = z[2 * h], im = z[2 * h + 1];
            const int mr = (row0 >> 1) + h;
            store_mirrored(g.out_re, g.out_im, g.cm[mr], g.cn[col], g.cm2 ? g.cm2[mr] : -1, g.cn2 ? g.cn2[col] : -1,
                           ar * re - ai * im, ar * im + ai * re);
          }
        } else {
          double x[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) x[r] = __shfl_xor(z[r], 1);
          if ((lr & 1) == 0 && colok) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              if (row0 + 2 * h >= g.M) continue;
              const double re = z[2 * h] - x[2 * h + 1];  // rr - ii
              const double im = x[2 * h] + z[2 * h + 1];  // ri + ir
              const int mr = (row0 >> 1) + h, nc = col >> 1;
              store_mirrored(g.out_re, g.out_im, g.cm[mr], g.cn[nc], g.cm2 ? g.cm2[mr] : -1, g.cn2 ? g.cn2[nc] : -1,
                             ar * re - ai * im, ar * im + ai * re);
            }
          }
        }
      }
    if (tnext >= tiles) break;
    // every wave has read this tile's scales before they are replaced
    __syncthreads();
    t = tnext;
  }
}

// ---------------------------------------------------------------------------
// Digit-pair GEMM for 5 and 6 digits on v_mfma_i32_32x32x32_i8: the 8-wave
// 128 x 128 tile of the kernel above with K steps of 32 (one 32-wide slice),
// so that the double-buffered digit tiles of A and B fit the 160 KB of LDS
// (2 x S x 8 KB per operand).  Threads 0..255 load the A tile, 256..511 the B
// tile (one 16-byte chunk per digit each), and a wave reads its A digits for
// one 32-row block at a time and one B digit at a time, which keeps the
// S * 32 accumulator registers, operands and prefetch within 256 registers
// (two waves per SIMD).  LDS rows are 32 bytes with the chunk index XORed by
// (row >> 3) & 1: 16 lanes reading rows r..r+15 at one chunk then cover the 64
// banks once.  Epilogue, tile order and persistence as in ozgemm8_kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int oz_swz2(int row, int kc) { return row * 2 + (kc ^ ((row >> 3) & 1)); }

template <int S>
__global__ __launch_bounds__(512, 1) void ozgemm8n_kernel(OzGemmArgs g) {
  constexpr int BN = 128, BK = 32;
  __shared__ v4i As[2][S][OZ_BM * 2];
  __shared__ v4i Bs[2][S][BN * 2];
  __shared__ double sA[OZ_BM], sB[BN];
  const int tiles_m = (g.M + OZ_BM - 1) / OZ_BM, tiles_n = (g.N + BN - 1) / BN;
  const int tiles = tiles_m * tiles_n;
  const int tid = threadIdx.x;
  const int kb = g.k_chunk > 0 ? (int)blockIdx.y * g.k_chunk : 0;
  const int ke = g.k_chunk > 0 ? min(g.Kp, kb + g.k_chunk) : g.Kp;
  double* const Cout = g.C + (int64_t)blockIdx.y * g.c_split_stride;
  const bool on_b = tid >= 256;
  const int lt = tid & 255, lrow = lt >> 1, lchunk = lt & 1;
  const __amdgpu_buffer_rsrc_t rX = on_b
      ? __builtin_amdgcn_make_buffer_rsrc((void*)g.B, (short)0, (int)(S * g.b_plane), 0x00020000)
      : __builtin_amdgcn_make_buffer_rsrc((void*)g.A, (short)0, (int)(S * g.a_plane), 0x00020000);
  const int xplane = on_b ? (int)g.b_plane : (int)g.a_plane;
  int m0 = 0, n0 = 0, vx = 0;
  auto setup = [&](int t) {
    const int L = remap_tile(t, tiles);
    int tm, tn;
    grouped_tile(L, tiles_m, tiles_n, g.tile_group, tm, tn);
    m0 = tm * OZ_BM;
    n0 = tn * BN;
    vx = on_b ? min(n0 + lrow, g.N - 1) * g.Kp + lchunk * 16 : min(m0 + lrow, g.M - 1) * g.Kp + lchunk * 16;
  };
  v4i rx[S];
  auto load_global = [&](int k0) {
#pragma unroll
    for (int s = 0; s < S; ++s)
      rx[s] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rX, vx, s * xplane + k0, 0));
  };
  auto store_lds = [&](int buf) {
    if (on_b) {
#pragma unroll
      for (int s = 0; s < S; ++s) Bs[buf][s][oz_swz2(lrow, lchunk)] = rx[s];
    } else {
#pragma unroll
      for (int s = 0; s < S; ++s) As[buf][s][oz_swz2(lrow, lchunk)] = rx[s];
    }
  };
  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 32;
  const int lr = lane & 31, lh = lane >> 5;
  const double ar = g.alpha_re, ai = g.alpha_im;

  int t = blockIdx.x;
  setup(t);
  load_global(kb);
  for (;;) {
    v16i acc[S][2];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[q][i][e] = 0;
    if (tid < OZ_BM) sA[tid] = m0 + tid < g.M ? g.sa[(int64_t)blockIdx.y * g.sa_split_stride + m0 + tid] : 0.0;
    else if (tid - OZ_BM < BN) sB[tid - OZ_BM] = n0 + tid - OZ_BM < g.N ? g.sb[(int64_t)blockIdx.y * g.sb_split_stride + n0 + tid - OZ_BM] : 0.0;
    store_lds(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += BK) {
      const bool more = k0 + BK < ke;
      if (more) load_global(k0 + BK);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        v4i a[S];
#pragma unroll
        for (int s = 0; s < S; ++s) a[s] = As[buf][s][oz_swz2(wm + i * 32 + lr, lh)];
#pragma unroll
        for (int tt = 0; tt < S; ++tt) {
          const v4i b = Bs[buf][tt][oz_swz2(wn + lr, lh)];
#pragma unroll
          for (int s = 0; s + tt < S; ++s)
            acc[s + tt][i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[s], b, acc[s + tt][i], 0, 0, 0);
        }
      }
      if (more) store_lds(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
    const int em0 = m0, en0 = n0;
    const int tnext = t + (int)gridDim.x;
    if (tnext < tiles) {
      setup(tnext);
      load_global(kb);
    }
    // digit group q (pairs with s + t = q + 2, 1-based digits) weighs 254^-(q+2)
    const int col = en0 + wn + lr;
    const bool colok = col < g.N;
    const double sbc = sB[wn + lr];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int gq = 0; gq < 4; ++gq) {
        const int row0 = em0 + wm + i * 32 + 8 * gq + 4 * lh;
        double z[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double v = 0.0;
#pragma unroll
          for (int q = S - 1; q >= 0; --q) v += (double)acc[q][i][4 * gq + r] * oz_group_weight(q);
          z[r] = (colok && row0 + r < g.M) ? v * (sA[row0 + r - em0] * sbc) : 0.0;
        }
        if (g.epi < 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (colok && row0 + r < g.M) Cout[(int64_t)(row0 + r) * g.ldc + col] = z[r];
        } else if (g.epi == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (colok && row0 + r < g.M)
              store_mirrored(g.out_re, nullptr, g.cm[row0 + r], g.cn[col], g.cm2 ? g.cm2[row0 + r] : -1,
                             g.cn2 ? g.cn2[col] : -1, ar * z[r], 0.0);
        } else if (g.epi == 1) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if (!colok || row0 + 2 * h >= g.M) continue;
            const double re = z[2 * h], im = z[2 * h + 1];
            const int mr = (row0 >> 1) + h;
            store_mirrored(g.out_re, g.out_im, g.cm[mr], g.cn[col], g.cm2 ? g.cm2[mr] : -1, g.cn2 ? g.cn2[col] : -1,
                           ar * re - ai * im, ar * im + ai * re);
          }
        } else {
          double x[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) x[r] = __shfl_xor(z[r], 1);
          if ((lr & 1) == 0 && colok) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              if (row0 + 2 * h >= g.M) continue;
              const double re = z[2 * h] - x[2 * h + 1];  // rr - ii
              const double im = x[2 * h] + z[2 * h + 1];  // ri + ir
              const int mr = (row0 >> 1) + h, nc = col >> 1;
              store_mirrored(g.out_re, g.out_im, g.cm[mr], g.cn[nc], g.cm2 ? g.cm2[mr] : -1, g.cn2 ? g.cn2[nc] : -1,
                             ar * re - ai * im, ar * im + ai * re);
            }
          }
        }
      }
    if (tnext >= tiles) break;
    __syncthreads();
    t = tnext;
  }
}

// ---------------------------------------------------------------------------
// The 8-wave digit-pair GEMM with asynchronous global-to-LDS copies
// (global_load_lds_dwordx4: lane l of a wave writes its 16 bytes at the wave's
// LDS base + 16 l) into three LDS buffers, loads issued two K steps ahead: a K
// step of 32 (S 15- or 10-MFMA groups per 32-row block) is shorter than a
// global round trip, which a one-step register prefetch left exposed.  No
// staging registers and no LDS store instructions; one barrier per K step.
// Tile 128 x 128, 8 waves (2 x 4), each wave 64 x 32 outputs; LDS rows of 32
// bytes as in ozgemm8n_kernel (chunk XORed with (row >> 3) & 1): lane l of
// wave w (A side w < 4, B side w >= 4) fills physical chunk c = 64 (w & 3) + l
// of a digit plane, i.e. row c >> 1, logical chunk (c & 1) ^ ((row >> 3) & 1).
// Epilogue, tile order and persistence as in ozgemm8_kernel.
// ---------------------------------------------------------------------------
// One 16-byte global -> LDS copy per lane, LDS address = dst (wave-uniform)
// + 16 * lane.  Written as one asm block so that the compiler does not track
// it: its waits are explicit s_waitcnt vmcnt(n) in the caller, which lets two
// K steps stay in flight across the step's barrier (the compiler's own
// tracking of these copies waits for all of them at every barrier).  The block
// has no register outputs; the compiler sets m0 (an input operand), and the
// s_nop covers the m0 write -> LDS copy hazard.
__device__ __forceinline__ void lds_copy16(const int8_t* src, v4i* dst) {
  const uint32_t lds = __builtin_amdgcn_readfirstlane(
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) v4i*)dst);
  asm volatile("s_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "{m0}"(lds) : "memory");
}

template <int S, int BM>
__global__ __launch_bounds__(512, 1) void ozgemm8d_kernel(OzGemmArgs g) {
  // BM = 128: waves 0-3 copy A, 4-7 copy B, wave tile 64 x 32 (two 32-row
  // blocks); BM = 64 (A with few rows): waves 0-1 copy A, 2-5 copy B, 6-7
  // none, wave tile 32 x 32
  constexpr int BN = 128, BK = 32, NBUF = 3, IB = BM / 64;
  __shared__ v4i As[NBUF][S][BM * 2];
  __shared__ v4i Bs[NBUF][S][BN * 2];
  __shared__ double sA[BM], sB[BN];
  const int tiles_m = (g.M + BM - 1) / BM, tiles_n = (g.N + BN - 1) / BN;
  const int tiles1 = tiles_m * tiles_n, tiles = tiles1 * g.nb;  // nb products one after another
  const int tid = threadIdx.x;
  const int kb = g.k_chunk > 0 ? (int)blockIdx.y * g.k_chunk : 0;
  const int ke = g.k_chunk > 0 ? min(g.Kp, kb + g.k_chunk) : g.Kp;
  const int nk = (ke - kb + BK - 1) / BK;
  double* Cout = g.C + (int64_t)blockIdx.y * g.c_split_stride;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int AW = BM / 32;  // waves copying A (2 chunks per row and plane)
  const bool on_b = wave >= AW;
  const bool copies = wave < AW + 4;
  const int lane = tid & 63;
  const int cw = on_b ? wave - AW : wave;                // wave index within its operand
  const int pc = 64 * cw + lane;                         // physical chunk this lane fills
  const int lrow = pc >> 1, lchunk = (pc & 1) ^ ((lrow >> 3) & 1);
  const int8_t* xbase = on_b ? g.B : g.A;
  int zb = 0;  // product of the current tile
  // planar digits: row stride Kp, K offset k, digit stride plane; digit-
  // interleaved rows: row stride Kp S, K offset k S (k a multiple of 32),
  // digit stride 32
  const int64_t xplane = g.dil ? 32 : (on_b ? g.b_plane : g.a_plane);
  const int sl = g.Sl > 0 ? g.Sl : S;
  const int64_t xrstride = g.dil ? (int64_t)g.Kp * sl : g.Kp;
  const int kmul = g.dil ? sl : 1;
  int m0 = 0, n0 = 0;
  int64_t xrow = 0;
  auto setup = [&](int t) {
    const int L0 = remap_tile(t, tiles);
    zb = L0 / tiles1;
    const int L = L0 - zb * tiles1;
    xbase = on_b ? g.B + zb * g.b_zstride : g.A + zb * g.a_zstride;
    Cout = g.C ? g.C + (int64_t)blockIdx.y * g.c_split_stride + zb * g.c_zstride : nullptr;
    int tm, tn;
    grouped_tile(L, tiles_m, tiles_n, g.tile_group, tm, tn);
    m0 = tm * BM;
    n0 = tn * BN;
    xrow = on_b ? (int64_t)min(n0 + min(lrow, BN - 1), g.N - 1) * xrstride
                : (int64_t)min(m0 + min(lrow, BM - 1), g.M - 1) * xrstride;
  };
  // copies K step j of the current tile into buffer j % 3: one 16-byte chunk
  // per lane per digit plane, written by the copy engine to LDS at the wave's
  // base + 16 * lane
  auto issue = [&](int j) {
    if (!copies) return;
    const int buf = j % NBUF, k0 = kb + j * BK;
    const int8_t* src = xbase + xrow + (int64_t)k0 * kmul + lchunk * 16;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      v4i* dst = on_b ? &Bs[buf][s][64 * cw] : &As[buf][s][64 * cw];
      lds_copy16(src + s * xplane, dst);
    }
  };
  const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * 32;
  const int lr = lane & 31, lh = lane >> 5;
  const double ar = g.alpha_re, ai = g.alpha_im;

  int t = blockIdx.x;
  setup(t);
  for (;;) {
    // the tile's scales first (their loads complete before the copies start),
    // then the copies of its first two K steps
    if (tid < BM)
      sA[tid] = m0 + tid < g.M ? g.sa[(int64_t)blockIdx.y * g.sa_split_stride + zb * g.sa_zstride + m0 + tid] : 0.0;
    else if (tid - BM < BN)
      sB[tid - BM] = n0 + tid - BM < g.N
          ? g.sb[(int64_t)blockIdx.y * g.sb_split_stride + zb * g.sb_zstride + n0 + tid - BM] : 0.0;
    if (nk > 0) issue(0);
    if (nk > 1) issue(1);
    v16i acc[S][IB];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[q][i][e] = 0;
    for (int j = 0; j < nk; ++j) {
      // step j's copies are the oldest outstanding: wait until only step
      // j + 1's S copies (if issued) remain in flight
      if (j + 1 < nk) {
        if constexpr (S == 5) __builtin_amdgcn_s_waitcnt(0xF75);  // vmcnt(5)
        else if constexpr (S == 4) __builtin_amdgcn_s_waitcnt(0xF74);  // vmcnt(4)
        else __builtin_amdgcn_s_waitcnt(0);
      } else {
        __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
      }
      __syncthreads();
      if (j + 2 < nk) issue(j + 2);
      const int buf = j % NBUF;
#pragma unroll
      for (int i = 0; i < IB; ++i) {
        v4i a[S];
#pragma unroll
        for (int s = 0; s < S; ++s) a[s] = As[buf][s][oz_swz2(wm + i * 32 + lr, lh)];
#pragma unroll
        for (int tt = 0; tt < S; ++tt) {
          const v4i b = Bs[buf][tt][oz_swz2(wn + lr, lh)];
#pragma unroll
          for (int s = 0; s + tt < S; ++s)
            acc[s + tt][i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[s], b, acc[s + tt][i], 0, 0, 0);
        }
      }
    }
    const int em0 = m0, en0 = n0;
    const int tnext = t + (int)gridDim.x;
    const double sbc = sB[wn + lr];
    const int col = en0 + wn + lr;
    const bool colok = col < g.N;
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
      for (int gq = 0; gq < 4; ++gq) {
        const int row0 = em0 + wm + i * 32 + 8 * gq + 4 * lh;
        double z[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double v = 0.0;
#pragma unroll
          for (int q = S - 1; q >= 0; --q) v += (double)acc[q][i][4 * gq + r] * oz_group_weight(q);
          z[r] = (colok && row0 + r < g.M) ? v * (sA[row0 + r - em0] * sbc) : 0.0;
        }
        if (g.epi < 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (colok && row0 + r < g.M) Cout[(int64_t)(row0 + r) * g.ldc + col] = z[r];
        } else if (g.epi == 3) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (colok && row0 + r < g.M) {
              const int64_t po = (int64_t)(row0 + r) * g.p12_ld + col;
              const double p1 = g.p12[po], p2 = g.p12[g.p12_stride + po];
              const double re = p1 - p2, im = (z[r] - p1) - p2;
              store_mirrored(g.out_re, g.out_im, g.cm[row0 + r], g.cn[col], g.cm2 ? g.cm2[row0 + r] : -1,
                             g.cn2 ? g.cn2[col] : -1, ar * re - ai * im, ar * im + ai * re);
            }
        } else if (g.epi == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (colok && row0 + r < g.M)
              store_mirrored(g.out_re, nullptr, g.cm[row0 + r], g.cn[col], g.cm2 ? g.cm2[row0 + r] : -1,
                             g.cn2 ? g.cn2[col] : -1, ar * z[r], 0.0);
        } else if (g.epi == 1) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if (!colok || row0 + 2 * h >= g.M) continue;
            const double re = z[2 * h], im = z[2 * h + 1];
            const int mr = (row0 >> 1) + h;
            store_mirrored(g.out_re, g.out_im, g.cm[mr], g.cn[col], g.cm2 ? g.cm2[mr] : -1, g.cn2 ? g.cn2[col] : -1,
                           ar * re - ai * im, ar * im + ai * re);
          }
        } else {
          double x[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) x[r] = __shfl_xor(z[r], 1);
          if ((lr & 1) == 0 && colok) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              if (row0 + 2 * h >= g.M) continue;
              const double re = z[2 * h] - x[2 * h + 1];  // rr - ii
              const double im = x[2 * h] + z[2 * h + 1];  // ri + ir
              const int mr = (row0 >> 1) + h, nc = col >> 1;
              store_mirrored(g.out_re, g.out_im, g.cm[mr], g.cn[nc], g.cm2 ? g.cm2[mr] : -1, g.cn2 ? g.cn2[nc] : -1,
                             ar * re - ai * im, ar * im + ai * re);
            }
          }
        }
      }
    if (tnext >= tiles) break;
    // every wave is done with the LDS buffers and the scales before the next
    // tile's scales and copies replace them
    __syncthreads();
    t = tnext;
    setup(t);
  }
}

int64_t ozgemm_max_k(int S) {
  // S pairs share the largest group; |d| <= 127 for every digit
  const int64_t lim = ((int64_t)1 << 31) - 1;
  return lim / ((int64_t)kOzDigitMax * kOzDigitMax * S) / OZ_BK * OZ_BK;
}

// Bands of tile rows for the 4-digit kernels' tile order: bands of 4 on grids
// at least 40 tiles wide (measured 3-7 % faster at 50 x 50 tiles), otherwise
// one band (tile row index fastest), which measured 8-20 % faster at 50 x 25.
// TDCC_OZ_GROUP overrides it.
static int oz_tile_group(int tiles_m, int tiles_n) {
  static const int env = [] {
    const char* e = std::getenv("TDCC_OZ_GROUP");
    return e ? std::max(1, std::atoi(e)) : 0;
  }();
  if (env > 0) return env;
  return tiles_n >= 40 ? 4 : tiles_m;
}

// Blocks of the persistent 4-digit kernel with one split: one per CU (the
// kernel's LDS holds one block per CU).  TDCC_OZ_PERSIST overrides it: 0 gives
// one block per tile, n > 0 about n blocks (a multiple of 8).
static int oz_persistent_blocks() {
  if (const char* e = std::getenv("TDCC_OZ_PERSIST")) {
    const int n = std::atoi(e);
    return n <= 0 ? 0 : std::max(8, n / 8 * 8);
  }
  static const int cus = [] {
    int dev = 0, c = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    return c / 8 * 8;
  }();
  return cus;
}

// The LDS-direct kernel (global -> LDS copies, two K steps in flight) for 4
// and 5 digits; TDCC_OZ_DIRECT=0 selects the register-staged kernels.
bool ozgemm_direct() {
  static const bool on = [] {
    const char* e = std::getenv("TDCC_OZ_DIRECT");
    return e ? std::atoi(e) != 0 : true;
  }();
  return on;
}

bool ozgemm_interleaved_ok(int S) { return (S == 4 || S == 5) && ozgemm_direct(); }

// The LDS-direct kernel's 64-row tile for an A operand of at most 64 rows
// (TDCC_OZ_TILE64=0: always 128 rows; 2: every launch, a test mode).
static bool oz_small_tile(const OzGemmArgs& g) {
  const char* e = std::getenv("TDCC_OZ_TILE64");
  const int m = e ? std::atoi(e) : 1;
  return m == 2 || (m == 1 && g.M <= 64);
}

const char* ozgemm_kernel_name(int S) {
  if ((S == 4 || S == 5) && ozgemm_direct()) return "ozgemm8d_kernel";
  return S <= 4 ? "ozgemm8_kernel" : (S == 5 ? "ozgemm8n_kernel" : "ozgemm_kernel");
}

void launch_ozgemm(const OzGemmArgs& args, hipStream_t s) {
  if (args.M <= 0 || args.N <= 0) return;
  OzGemmArgs g = args;
  g.tile_group = g.tune_group > 0 ? g.tune_group : oz_tile_group((g.M + OZ_BM - 1) / OZ_BM, (g.N + 127) / 128);
  const int bn = g.S <= 5 ? 128 : 64;
  // several products per launch (nb > 1) and the 3M epilogue exist only in the
  // LDS-direct kernel
  if ((g.nb > 1 || g.epi == 3) && !((g.S == 4 || g.S == 5) && ozgemm_direct())) return;
  const int tiles = ((g.M + OZ_BM - 1) / OZ_BM) * ((g.N + bn - 1) / bn) * g.nb;
  const int splits = g.k_chunk > 0 ? g.splits : 1;
  const int pb = oz_persistent_blocks();
  // persistent only with few K steps per tile (at 50 K steps one block per
  // tile measured 12 % faster: the dispatcher balances long tiles better)
  int gx8 = (g.S <= 5 && splits == 1 && pb > 0 && tiles > pb && g.Kp <= 8 * OZ_BK) ? pb : tiles;
  if (g.tune_pb >= 0 && g.S <= 5) gx8 = (g.tune_pb > 0 && splits == 1 && g.tune_pb < tiles) ? g.tune_pb : tiles;
  const dim3 grid(tiles, splits), grid8(gx8, splits), block(256), block8(512);
  if (ozgemm_direct() && (g.S == 4 || g.S == 5)) {
    if (oz_small_tile(g)) {
      // one 64-row tile of A rows: half the MFMA work of a 128-row tile
      const int t64 = ((g.M + 63) / 64) * ((g.N + 127) / 128) * g.nb;
      int gx = (splits == 1 && pb > 0 && t64 > pb && g.Kp <= 8 * OZ_BK) ? pb : t64;
      if (g.tune_pb >= 0) gx = (g.tune_pb > 0 && splits == 1 && g.tune_pb < t64) ? g.tune_pb : t64;
      const dim3 grid64(gx, splits);
      if (g.S == 4) hipLaunchKernelGGL((ozgemm8d_kernel<4, 64>), grid64, block8, 0, s, g);
      else hipLaunchKernelGGL((ozgemm8d_kernel<5, 64>), grid64, block8, 0, s, g);
      return;
    }
    if (g.S == 4) hipLaunchKernelGGL((ozgemm8d_kernel<4, 128>), grid8, block8, 0, s, g);
    else hipLaunchKernelGGL((ozgemm8d_kernel<5, 128>), grid8, block8, 0, s, g);
    return;
  }
  switch (g.S) {
    case 2: hipLaunchKernelGGL((ozgemm8_kernel<2>), grid8, block8, 0, s, g); break;
    case 3: hipLaunchKernelGGL((ozgemm8_kernel<3>), grid8, block8, 0, s, g); break;
    case 4: hipLaunchKernelGGL((ozgemm8_kernel<4>), grid8, block8, 0, s, g); break;
    case 5: hipLaunchKernelGGL((ozgemm8n_kernel<5>), grid8, block8, 0, s, g); break;
    case 6: hipLaunchKernelGGL((ozgemm_kernel<6, 2>), grid, block, 0, s, g); break;
    default: break;
  }
}

}  // namespace tdcc
