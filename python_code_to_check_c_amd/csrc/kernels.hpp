// Device kernels of the TDCC propagation library and their launch arguments.
//
// Complex tensors are stored as two FP64 planes (re, im) of the same shape;
// a real tensor has im == nullptr.  Every kernel here is deterministic: no
// floating-point atomics, fixed reduction orders.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tdcc {

__device__ __forceinline__ int remap_tile(int bid, int tiles) {
  // Workgroups are dispatched round-robin over the 8 XCDs (bid % 8).  Give
  // every XCD a contiguous range of tiles so that tiles sharing a B panel
  // (same tn, consecutive tm) run on one XCD and share its L2.
  const int q = tiles >> 3, r = tiles & 7;
  const int x = bid & 7, loc = bid >> 3;
  return x * q + (x < r ? x : r) + loc;
}

// Tile (tm, tn) of the L-th tile when the tile grid is walked in bands of G
// tile rows, each band column by column: consecutive tiles (those one XCD runs
// at the same time after remap_tile) then form G x (32 / G) patches that read
// G A panels and 32 / G B panels instead of 32 A panels and one or two B panels.
__device__ __forceinline__ void grouped_tile(int L, int tiles_m, int tiles_n, int G, int& tm, int& tn) {
  const int width = G * tiles_n, band = L / width, first = band * G;
  const int gs = tiles_m - first < G ? tiles_m - first : G, r = L - band * width;
  tm = first + r % gs;
  tn = r / gs;
}

// ---------------------------------------------------------------------------
// Strided tensor contraction on FP64 MFMA ("GEMM-like tensor-tensor", GETT).
// The M, N, K and batch index groups of an einsum are flattened on the host
// into offset tables, so C[cm[m] + cn[n] + cb[b]] = alpha * sum_k
// A[am[m] + ak[k] + ab[b]] * B[bn[n] + bk[k] + bb[b]] (+ C if accumulate).
struct GettArgs {
  const double* a_re; const double* a_im;
  const double* b_re; const double* b_im;
  double* c_re; double* c_im;
  const int64_t* am; const int64_t* ak;
  const int64_t* bn; const int64_t* bk;
  const int64_t* cm; const int64_t* cn;
  const int64_t* ab; const int64_t* bb; const int64_t* cb;  // batch (may be null when batch == 1)
  int M, N, K, batch;
  int tiles_m, tiles_n;
  int splits, k_chunk;          // split-K: splits > 1 writes partials
  double alpha_re, alpha_im;
  int accumulate;
  double* partial;              // [splits][batch][M][N] x planes, when splits > 1
  // >= 0: the K offsets are k * step (ak / bk unused; their first entry is
  // folded into the operand's pointers)
  int64_t ak_step = -1, bk_step = -1;
};

// Loader mapping: whether A's / B's memory-contiguous index lies in K.
void launch_gett(const GettArgs& g, bool a_cplx, bool b_cplx, bool a_kcontig,
                 bool b_kcontig, hipStream_t s);
void launch_gett_reduce(const GettArgs& g, bool c_cplx, hipStream_t s);

// ---------------------------------------------------------------------------
// Linear combination with per-term permuted strides:
//   out[idx] (+)= sum_t coef_t * src_t[sum_d idx_d * st_t[d]]
constexpr int kMaxTerms = 24;   // terms per launch
constexpr int kInlineTerms = 8;  // a source linear combination is inlined while the total stays within this
struct LincombTerm {
  const double* re; const double* im;
  int64_t st[4];
  double cr, ci;
  // product term: coef * x * y with y = (re2, im2) read at st2 (re2 null: none)
  const double* re2 = nullptr; const double* im2 = nullptr;
  int64_t st2[4] = {0, 0, 0, 0};
  // stacked trajectories: element offset of trajectory z is z * zs (zs2 for
  // the product factor); 0 for an operand shared by all trajectories
  int64_t zs = 0, zs2 = 0;
  // flat kernel: read at the offsets of the index tuple (rows contiguous along
  // the innermost axis, other strides unlike the output's); set at launch
  bool rowwise = false;
};
struct LincombArgs {
  double* out_re; double* out_im;
  int64_t ost[4];
  int64_t dims[4];
  int rank;
  int64_t total;
  int nterms;
  int accumulate;
  LincombTerm t[kMaxTerms];
  // rank 4: bit 0 set when the result changes sign under exchanging axes 0
  // and 1, bit 1 for axes 2 and 3 (every term is then read at the elements
  // with i < j / a <= b tiles only and the rest is stored as mirror images)
  int anti = 0;
  // stacked trajectories: nz trajectories over grid.y, the output of
  // trajectory z at element offset z * ozs
  int nz = 1;
  int64_t ozs = 0;
};
void launch_lincomb(const LincombArgs& a, hipStream_t s);
// Name of the kernel launch_lincomb runs for these arguments (profiles).
const char* lincomb_kernel_name(const LincombArgs& a);

// ---------------------------------------------------------------------------
// Field + Fock build for one stage time of the current step:
//   x = (-A) * mu_pol;  F_s = (F + x*cos) + i (x*sin)
// (cos, sin) come from table[(counter - 1) * 3 + stage].
struct FieldArgs {
  const double* F; const double* mu_pol;
  double* out_re; double* out_im;
  int64_t n2;
  double negA;
  const double* table;   // [steps][3][2]
  const int64_t* counter;
  int stage;
};
void launch_field(const FieldArgs& a, hipStream_t s);

// ---------------------------------------------------------------------------
// Fused RK4 finalize for one amplitude block:
//   base = rint(y * scale) / scale
//   inc  = dt6 * (((k1 + 2 k2) + 2 k3) + k4)
//   y    = base + sign_i * inc      (sign_i = -1j for T, +1j for Lambda)
struct FinalizeArgs {
  double* y_re; double* y_im;
  double* base_re; double* base_im;   // may be null
  const double* k_re[4]; const double* k_im[4];
  int64_t n;
  double dt6, scale;
  int lambda_sign;                    // 0: -1j (T), 1: +1j (Lambda)
};
void launch_finalize(const FinalizeArgs& a, hipStream_t s);

// ---------------------------------------------------------------------------
// Dipole trace: out[c] = sum_pq mu[c][p][q] * D[p][q] with the correlated
// density assembled from its blocks [[Dij, lam1], [Dai, Dab]].
struct TraceArgs {
  const double* dij_re; const double* dij_im;
  const double* l1_re;  const double* l1_im;
  const double* dai_re; const double* dai_im;
  const double* dab_re; const double* dab_im;
  const double* mu;     // [3][n][n]
  int o, v;
  double* out;          // [rows][3][2]
  const int64_t* counter;   // row = counter - 1 (row 0 when null)
};
void launch_trace(const TraceArgs& a, hipStream_t s);

void launch_counter_inc(int64_t* counter, hipStream_t s);
void launch_counter_set(int64_t* counter, int64_t value, hipStream_t s);

// ---------------------------------------------------------------------------
// Synthetic antisymmetrized TEI on the device (same generator as
// python_code_to_check_c_amd/synthetic.py).
void launch_synth_tei(double* tei, int n, uint32_t config_id, double sigma, hipStream_t s);

}  // namespace tdcc

namespace tdcc {
// ---------------------------------------------------------------------------
// Dense real GEMM on FP64 MFMA: C = alpha * A * op(B) + beta * C, all row-major.
// A is M x K (lda); op(B) = B^T with B stored N x K (ldb) when b_trans, else
// B stored K x N (ldb); C is M x N (ldc).
struct DgemmArgs {
  const double* A; const double* B; double* C;
  int M, N, K;
  int64_t lda, ldb, ldc;
  double alpha, beta;
  // split-K: grid.y = splits, split s covers k in [s*k_chunk, (s+1)*k_chunk)
  // and writes C + s * c_split_stride (k_chunk == 0: one split over all K)
  int k_chunk = 0;
  int splits = 1;
  int64_t c_split_stride = 0;
};
void launch_dgemm(const DgemmArgs& g, bool b_trans, hipStream_t s);

// Antisymmetric pair packing of the virtual-virtual index pair (e < f,
// lexicographic; P = v (v - 1) / 2) of a complex (o, o, v, v) tensor:
//   pack:   X[p * rows + r][q] = x[p][i][j][e_q][f_q] - x[p][i][j][f_q][e_q]
//           for planes p, written as a (2 rows) x P real matrix;
//   unpack: y[p][i][j][a][b] = alpha * Y[p * rows + r][q(a,b)] for a < b,
//           -alpha * Y[...][q(b,a)] for a > b, 0 for a == b.
// Rows r = i o + j (rows = o^2), or with ij_pairs only i < j, r = q(i, j)
// (rows = o (o - 1) / 2), and unpack writes y[j][i] = -y[i][j], y[i][i] = 0:
// for tensors antisymmetric in (i, j) the other half is redundant.
void launch_pack_pairs(const double* x_re, const double* x_im, double* X, int o, int v, bool ij_pairs,
                       hipStream_t s);
void launch_unpack_pairs(const double* Y, double* y_re, double* y_im, int o, int v, double alpha, bool ij_pairs,
                         hipStream_t s);
// y[i] = y[i] + y[stride + i] + ... over `splits` partial results, in order.
void launch_sum_splits(double* y, int splits, int64_t stride, int64_t n, hipStream_t s);
// Packed <ab||ef> of the virtual block of a (n,n,n,n) TEI with o occupied:
//   ladder_ket == 0: Vp[q(a,b)][q(e,f)] = <ab||ef>
//   ladder_ket == 1: Vp[q(a,b)][q(e,f)] = <ef||ab>
void launch_build_vpack(const double* tei, int o, int v, int ladder_ket, double* Vp, hipStream_t s);
// Counts entries of the virtual block violating <ab||ef> = -<ba||ef> = -<ab||fe>.
// `count` holds kCountBlocks + 1 entries; the total lands in count[kCountBlocks].
constexpr int kCountBlocks = 1024;
// Antisymmetry of a device (o, o, v, v) interleaved complex amplitude in (i, j)
// and (a, b): `work` holds 2 kCountBlocks + 2 doubles; the maximum of |x| and
// the largest deviation land in work[2 kCountBlocks] and work[2 kCountBlocks + 1].
void launch_amp_antisym_check(const double* x, int o, int v, double* work, hipStream_t s);
void launch_vvvv_antisym_check(const double* tei, int o, int v, unsigned long long* count, hipStream_t s);
// The same count over the whole (n,n,n,n) array.
void launch_tei_antisym_check(const double* tei, int n, unsigned long long* count, hipStream_t s);
// Counts mismatches between two equally shaped arrays.
void launch_count_mismatch(const double* x, const double* y, int64_t n, unsigned long long* count, hipStream_t s);

// Dense GEMM form of a contraction: operand gather into row-major panels
// and result scatter (see kernels_gemm.hip).
struct Gather2dArgs {
  const double* src_re; const double* src_im;
  const int64_t* rtab; const int64_t* ctab;
  int64_t Rb, Cb;        // base rows / columns (per plane)
  int planes;            // 1 or 2
  int planes_along_rows; // planes stacked along rows (1) or columns (0)
  double* dst;
  int64_t ld;
};
void launch_gather2d(const Gather2dArgs& g, hipStream_t s);
struct GemmScatterArgs {
  const double* Z;
  int64_t ldz;
  int splits;            // Z holds `splits` partial results z_stride apart, summed in order
  int64_t z_stride;
  int M, N;
  int kind;              // 0 real x real, 1 complex x real, 2 complex x complex
  const int64_t* cm; const int64_t* cn;
  double* out_re; double* out_im;
  double alpha_re, alpha_im;
  int accumulate;
  // output antisymmetric in a pair of row (column) letters and computed on the
  // ordered rows (columns) only: offsets of the mirrored row (column), -1 on
  // the diagonal; nullptr without such a pair (never with accumulate)
  const int64_t* cm2 = nullptr; const int64_t* cn2 = nullptr;
};
void launch_gemm_scatter(const GemmScatterArgs& g, hipStream_t s);

// ---------------------------------------------------------------------------
// Dot-product contractions (both free groups small, the summed group long) on
// FP64 MFMA, read through the offset tables without a gather:
//   Z[s][r][c] = sum_{k in split s} A'[r][k] * B'[c][k]
// with A' = [A_re; A_im] (row r = p M + m reads plane p at am[m] + ak[k]) and
// B' likewise (pb N rows); pa M, pb N <= kDotMaxRows.  launch_dot then sums the
// `splits` partial results in a fixed order and stores them as
// launch_gemm_scatter would (complex recombination by kind, alpha, output
// offset tables; no mirrors, no accumulation).
constexpr int kDotMaxRows = 512;
struct DotArgs {
  const double* a_re; const double* a_im;   // a_im / b_im null: real operand
  const double* b_re; const double* b_im;
  const int64_t* am; const int64_t* ak;
  const int64_t* bn; const int64_t* bk;
  int M, N;
  int64_t K;
  int64_t k_chunk;                          // multiple of 32
  int splits;
  int a_kc = 1, b_kc = 1;                   // operand contiguous along k (else along its rows)
  double* Z;                                // [splits][pa M][pb N]
};
void launch_dot(const DotArgs& a, const GemmScatterArgs& out, hipStream_t s);

// Store of one output element and its mirror images (see GemmScatterArgs):
// y at (om, on), -y at the mirrored row or column, y at both.
__device__ __forceinline__ void store_mirrored(double* re, double* im, int64_t om, int64_t on, int64_t om2,
                                               int64_t on2, double yr, double yi) {
  re[om + on] = yr;
  if (im) im[om + on] = yi;
  if (om2 >= 0) {
    re[om2 + on] = -yr;
    if (im) im[om2 + on] = -yi;
  }
  if (on2 >= 0) {
    re[om + on2] = -yr;
    if (im) im[om + on2] = -yi;
    if (om2 >= 0) {
      re[om2 + on2] = yr;
      if (im) im[om2 + on2] = yi;
    }
  }
}

// ---------------------------------------------------------------------------
// Split-precision GEMM on int8 MFMA (Ozaki scheme with exact accumulation).
//
// A row x of an FP64 matrix is written as x = 2^e * sum_{s=1..S} d_s R^-s
// (+ a remainder of at most 0.5 * R^-S * 2^e), radix R = kOzRadix = 254, with
// the row exponent e chosen so that |x| < 2^(e-1) and digits d_s =
// round-to-nearest in [-127, 127] (R = 254 is the largest radix whose rounded
// digits of remainders in [-1/2, 1/2] always fit int8).  A product of two split
// matrices keeps the digit pairs with s + t <= S + 1; pairs of equal s + t share
// one int32 accumulator, which sums exactly while pairs * 127^2 * k_chunk < 2^31.
// The FP64 result is the weighted sum of the accumulators (weights R^-(s+t))
// times the two row scales.
constexpr int kOzMaxSlices = 6;
constexpr double kOzRadix = 254.0;
constexpr int kOzDigitMax = 127;
// R^-(q + 2) for digit group q = s + t - 2 (1-based digits), q < kOzMaxSlices;
// 254^k is exact in FP64 for k <= 7, so each weight is one correctly rounded
// division.
__host__ __device__ constexpr double oz_group_weight(int q) {
  double p = kOzRadix * kOzRadix;
  for (int k = 0; k < q; ++k) p *= kOzRadix;
  return 1.0 / p;
}

struct SplitArgs {
  // Element (r, k) of the R x K source, R = planes * Rb, is
  //   plane(r / Rb)[rtab[r % Rb] + ktab[k]]   (offset tables), or
  //   plane(r / Rb)[(r % Rb) * ldx + k]       (rtab == ktab == nullptr);
  // plane 0 is src_re, plane 1 src_im, and with planes == 3 (not interleaved)
  // plane 2 is src_re + src_im (the third operand of a 3-product complex GEMM).
  const double* src_re; const double* src_im;
  const int64_t* rtab; const int64_t* ktab;
  int64_t ldx;
  int64_t Rb; int planes;
  int64_t K, Kp;        // Kp: padded row length of the digit planes (multiple of 64)
  int S;                // digits per element
  int8_t* D;            // S planes of R x Kp, `plane` bytes apart
  int64_t plane;
  double* scale;        // R row scales 2^e
  double* pmax;         // workspace: ozaki_split_workspace(R, Kp, by_columns) bytes
  bool multipass;       // always use the three-pass kernels (else single-pass when rows fit)
  bool interleave;      // with 2 planes: row r is plane r & 1 of base row r >> 1 (else plane r / Rb)
  // Digit-interleaved rows: digit s of (r, k) at D + r Kp S + (k / 32) 32 S +
  // 32 s + k % 32 (the S digits of a 32-wide K block of a row side by side),
  // else at D + s plane + r Kp + k.
  bool dil = false;
  // K chunks of their own scale (> 0, rows orientation, single pass): row r's
  // k in [c sk_chunk, (c+1) sk_chunk) gets scale[c R + r]; 0: one scale per row
  int64_t sk_chunk = 0;
};
// Three launches: partial row maxima over k tiles, row scales, digits.
// by_columns: threads run along r (for sources whose row index is the
// contiguous one) instead of along k.
void launch_ozaki_split(const SplitArgs& a, bool by_columns, hipStream_t s);
int64_t ozaki_split_workspace(int64_t R, int64_t Kp, bool by_columns);

struct OzGemmArgs {
  const int8_t* A; const int8_t* B;   // digit planes: A[s * a_plane + m * Kp + k], B[s * b_plane + n * Kp + k]
  int64_t a_plane, b_plane;
  const double* sa; const double* sb; // row scales of A (M) and B (N)
  double* C;                          // C[m * ldc + n] = sum_k A[m][k] * B[n][k]
  int64_t ldc;
  int M, N, Kp, S;                    // S: digits multiplied (at most the planes' Sl)
  int Sl = 0;                         // digits per element in the planes (0: S); the
                                      // digit-interleaved layout's strides follow it
  int k_chunk, splits;                // split-K over grid.y (k_chunk multiple of 64), partials
  int64_t c_split_stride;             // c_split_stride apart
  // Fused scatter epilogue (epi >= 0, one split only): instead of C, write
  //   out[cm[m] + cn[n]] = alpha * value(m, n)
  // with the planes of A (and B) interleaved by row:
  //   epi 0 (real x real):       value = Z[m][n]
  //   epi 1 (complex x real):    value = Z[2m][n] + i Z[2m+1][n]
  //   epi 2 (complex x complex): value = (Z[2m][2n] - Z[2m+1][2n+1]) + i (Z[2m][2n+1] + Z[2m+1][2n])
  int epi = -1;
  const int64_t* cm = nullptr; const int64_t* cn = nullptr;
  const int64_t* cm2 = nullptr; const int64_t* cn2 = nullptr;  // mirrored rows / columns, as in GemmScatterArgs
  double* out_re = nullptr; double* out_im = nullptr;
  double alpha_re = 1.0, alpha_im = 0.0;
  int tile_group = 1;                 // bands of tile rows (grouped_tile); 1: row index fastest
  // launch choices of an autotuner (bitwise neutral: every output is summed by
  // one block in the same order either way): 0 / -1 keep launch_ozgemm's
  // defaults; tune_group > 0 bands of tile rows; tune_pb 0 one block per
  // tile, > 0 that many persistent blocks
  int tune_group = 0;
  int tune_pb = -1;
  // both operands' digits in the digit-interleaved row layout (SplitArgs::dil);
  // only with ozgemm_interleaved_ok(S)
  bool dil = false;
  // scales per K split (SplitArgs::sk_chunk == k_chunk): split y reads
  // sa + y * sa_split_stride (sb likewise); 0: one scale per row for all splits
  int64_t sa_split_stride = 0, sb_split_stride = 0;
  // nb independent products of the same shape in one launch (LDS-direct
  // kernel, one split): product z reads A + z a_zstride, B + z b_zstride,
  // sa + z sa_zstride, sb + z sb_zstride and writes C + z c_zstride
  int nb = 1;
  int64_t a_zstride = 0, b_zstride = 0, sa_zstride = 0, sb_zstride = 0, c_zstride = 0;
  // epi 3 (complex x complex from three real products, 3M): this launch's
  // product is P3 = (A_re + A_im)(B_re + B_im)^T, P1 = A_re B_re^T and
  // P2 = A_im B_im^T are read from p12[m * p12_ld + n] and
  // p12[p12_stride + m * p12_ld + n], and
  //   value(m, n) = (P1 - P2) + i (P3 - P1 - P2)
  // goes through the output tables as in epi 0..2 (rows and columns complex)
  const double* p12 = nullptr;
  int64_t p12_ld = 0, p12_stride = 0;
};
void launch_ozgemm(const OzGemmArgs& a, hipStream_t s);
// Whether the digit GEMM for S digits reads digit-interleaved operands.
bool ozgemm_interleaved_ok(int S);

// Name of the kernel launch_ozgemm runs for S digits (profiles).
const char* ozgemm_kernel_name(int S);
// Largest K range one int32 accumulator may sum exactly for S digits.
int64_t ozgemm_max_k(int S);

// ---------------------------------------------------------------------------
// Skinny contraction (one operand with at most kSkinnyMax rows), FP64 VALU:
//   C[cs[i] + cb[j]] = alpha * sum_k S[i,k] G[j,k],
//   S[i,k] = s[sr[i] + sk[k]] (I <= kSkinnyMax), G[j,k] = g[gr[j] + gk[k]].
// g_kmajor: G's contiguous index lies in k.  s_im / g_im null for real
// operands.  splits > 1: split s covers k in [s k_chunk, (s+1) k_chunk) and
// writes partial[s][plane][i][j], reduced in split order.
constexpr int kSkinnyMax = 32;
struct SkinnyArgs {
  const double* s_re; const double* s_im;
  const double* g_re; const double* g_im;
  const int64_t* sr; const int64_t* sk;
  const int64_t* gr; const int64_t* gk;
  const int64_t* cs; const int64_t* cb;
  double* c_re; double* c_im;
  int I; int64_t J, K;
  int64_t k_chunk; int splits;
  double* partial;
  double alpha_re, alpha_im;
  bool g_kmajor;
  // Short runs (run > 0, big operand k-major): k = kh * run + r with
  // gk[k] = gk[kh * run] + r, and big rows j = jh * run_rows + jl with
  // gr[j] = gr[jh * run_rows] + jl * run, so each (jh, kh) block of the big
  // operand is run_rows x run contiguous elements.  One block per (jh, kh)
  // stages it through LDS and writes a partial per kh (splits = K / run,
  // k_chunk = run); a second kernel sums the partials in kh order.
  int run = 0, run_rows = 0;
};
// Largest run_rows * (run + 1) of the short-run kernel (its LDS block).
constexpr int kSkinnyRunsMax = 3584;
void launch_skinny(const SkinnyArgs& a, hipStream_t s);

// ---------------------------------------------------------------------------
// Adjoint kernels for d(sum_n w_n mu_pol(n))/d(A, w0, mu_pol).  The step
// counter selects the row of the per-step tables as in field_kernel.
//
// field_vjp: Fbar, the adjoint of one (rows x cols) block at (p0, q0) of the
// stage field F_s = F + (-A) mu_pol exp(i theta_s), theta_s = 2 pi w0 t_s:
//   g[0..1] += sum Fbar (-mu e^{i theta})                 (d/dA)
//   g[2..3] += sum Fbar (-A) mu e^{i theta} i 2 pi t_s     (d/dw0)
//   gmu[p][q] += Fbar (-A) e^{i theta}                     (d/dmu_pol)
struct FieldVjpArgs {
  const double* ab_re; const double* ab_im;  // contiguous (rows, cols) block adjoint
  int rows, cols, p0, q0, n;
  const double* mu_pol;
  double negA;
  const double* table;   // [steps][3][2] cos, sin
  const double* ttab;    // [steps][3] stage times
  const int64_t* counter;
  int stage;
  double* g;             // [4] (+=)
  double* gmu_re; double* gmu_im;  // (n, n) (+=)
};
void launch_field_vjp(const FieldVjpArgs& a, hipStream_t s);

// seed: out = w mu_pol[p0.., q0..] (a real block stored complex), w = wtab[counter - 1].
struct SeedArgs {
  double* out_re; double* out_im;
  int rows, cols, p0, q0, n;
  const double* mu_pol;
  const double* wtab;
  const int64_t* counter;
};
void launch_seed(const SeedArgs& a, hipStream_t s);

// trace_vjp: gmu[p][q] += w D[p][q] over the density blocks [[Dij, lam1], [Dai, Dab]].
struct TraceVjpArgs {
  const double* dij_re; const double* dij_im;
  const double* l1_re;  const double* l1_im;
  const double* dai_re; const double* dai_im;
  const double* dab_re; const double* dab_im;
  int o, v;
  const double* wtab;
  const int64_t* counter;
  double* gmu_re; double* gmu_im;
};
void launch_trace_vjp(const TraceVjpArgs& a, hipStream_t s);

// Interleaved complex <-> two planes.
void launch_split_complex(const double* inter, double* re, double* im, int64_t n, hipStream_t s);
void launch_join_complex(const double* re, const double* im, double* inter, int64_t n, hipStream_t s);
}  // namespace tdcc
