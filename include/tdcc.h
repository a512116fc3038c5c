/*
 * tdcc.h -- C ABI of the MI355X real-time TDCCSD / TDCC2 propagation library
 * (libtdcc_hip.so, built for gfx950).
 *
 * The reference (yramis/python_code_to_check_C) is pure Python with no FFI;
 * its boundary for this path is the Python surface of CCSD_Helper /
 * CC2_Helper / CC_Calculator.  Each entry point below replaces one piece of
 * that surface and cites it.  The Python mirror of the reference surface
 * (python_code_to_check_c_amd/CCSD_Helper.py, CC2_Helper.py, CC_Calculator.py)
 * binds these symbols through ctypes.
 *
 * Conventions
 *   - status: 0 = ok, nonzero = error; tdcc_last_error() gives the message.
 *   - complex arrays are interleaved (re, im) doubles in C order, i.e. the
 *     memory of a numpy complex128 array; real arrays are C-order doubles.
 *   - shapes: o = occupied, v = virtual spin orbitals, n = o + v;
 *     t1/lam1 (o,v), t2/lam2 (o,o,v,v), F (n,n), TEI (n,n,n,n) = <pq||rs>
 *     occupied first, mu (3,n,n).
 *   - host buffers are borrowed for the duration of a call only; the context
 *     owns all device memory; calls are synchronous for the caller.
 *   - a context is not thread-safe; separate contexts may be used concurrently.
 */
#ifndef TDCC_H
#define TDCC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdcc_ctx tdcc_ctx;

enum tdcc_method { TDCC_CCSD = 0, TDCC_CC2 = 1 };

enum tdcc_rhs_kind {
  TDCC_RHS_T1 = 0,      /* CCSD_Helper.T1eq_rhs(t1, t2, F)             CCSD_Helper.py:177 */
  TDCC_RHS_T2 = 1,      /* CCSD_Helper.T2eq_rhs(t1, t2, F)             CCSD_Helper.py:268 */
  TDCC_RHS_L1 = 2,      /* CCSD_Helper.lam_1eq_rhs(t1,t2,lam1,lam2,F)  CCSD_Helper.py:555 */
  TDCC_RHS_L2 = 3,      /* CCSD_Helper.lam2eq_rhs(t1,t2,lam1,lam2,F)   CCSD_Helper.py:610 */
  TDCC_RHS_T2_CC2 = 4,  /* CC2_Helper.T2eq_rhs_CC2(t1, t2, F)          CC2_Helper.py:12   */
  TDCC_RHS_L1_CC2 = 5,  /* CC2_Helper.L1eq_rhs_CC2(...)                CC2_Helper.py:50   */
  TDCC_RHS_L2_CC2 = 6   /* CC2_Helper.L2eq_rhs_CC2(...)                CC2_Helper.py:128  */
};

/* Create a context for one system of o occupied / v virtual spin orbitals on
 * HIP device `device`.  Replaces the psi4-backed CCSD_Helper.__init__
 * (CCSD_Helper.py:26-53) as far as this path needs it (sizes, slices). */
int tdcc_create(int nocc, int nvir, int device, tdcc_ctx** out);
void tdcc_destroy(tdcc_ctx* ctx);
const char* tdcc_last_error(const tdcc_ctx* ctx);

/* Integrals, Fock and dipoles (CCSD_Helper.py:53 self.TEI, :82-93 F_MO,
 * :970-980 Defd_dipole).  tei may be NULL when tdcc_synth_tei() provides it. */
int tdcc_set_system(tdcc_ctx* ctx, const double* tei, const double* fock, const double* mu);

/* Generate the antisymmetrized synthetic TEI on the device (seeded
 * Philox4x32-10 generator, identical to python_code_to_check_c_amd.synthetic). */
int tdcc_synth_tei(tdcc_ctx* ctx, uint32_t config_id, double sigma);

/* Arithmetic of the large contractions (the o^2 v^4 ladders and every
 * contraction above 2 GFLOP, i.e. the o^3 v^3 rings and o^2 v^3 terms).
 * digits = 0: FP64 (the default).  digits = 2..6: split precision -- both
 * operands are cut into `digits` int8 digit planes of radix 254 (digits in
 * [-127, 127]) with one power-of-two scale per row (Ozaki scheme), the digit
 * pairs are multiplied on int8 MFMA with exact int32 accumulation, and the
 * result is assembled in FP64; the truncation per operand element is at most
 * 0.5 * 254^-digits of its row scale (< 2 * 254^-digits of the row maximum).
 * The RK4 state, the field and Fock build, the small contractions and the
 * dipole stay FP64.  Replaces the reference's complex128 einsum arithmetic
 * (opt_einsum contract, CCSD_Helper.py:20) for those terms.
 *
 * Launch autotuning (environment): TDCC_AUTOTUNE=1 times candidate tile
 * orders and persistent grids for every digit-GEMM launch shape when a
 * program is bound, and keeps the fastest for the process (on by default for
 * batched programs); these choices leave every result bit for bit unchanged.
 * TDCC_AUTOTUNE=2 also tries fewer K splits (down to 2), which changes the
 * order of the FP64 partial sums.  TDCC_AUTOTUNE_FILE appends each choice to
 * that file. */
int tdcc_set_precision(tdcc_ctx* ctx, int digits);

/* Propagated state t1, t2, lam1, lam2 (complex, interleaved). */
int tdcc_set_state(tdcc_ctx* ctx, const double* t1, const double* t2, const double* l1, const double* l2);
/* which = 0: current state; which = 1: rounded base of the last step taken
 * (the t1min/t2min/L1min/L2min the reference saves on a stop,
 * CCSD_Helper.py:1207-1214, :1243). */
int tdcc_get_state(tdcc_ctx* ctx, int which, double* t1, double* t2, double* l1, double* l2);

/* Device-memory variants of tdcc_set_state / tdcc_get_state: t1, t2, lam1,
 * lam2 are DEVICE pointers (interleaved complex in C order, the memory of a
 * complex128 tensor) on the context's GPU; there is no host copy.  The
 * (i, j) antisymmetry test that picks the i < j shortcuts runs on the GPU with
 * the host test's bound.  Inputs must be complete when the call is made (the
 * caller synchronizes its own stream first); the call returns after its
 * results are written.  A host pointer is rejected with an error.  Same
 * state as the reference's Runge_Kutta_solver arguments
 * (CCSD_Helper.py:1191) and its saved t1min/t2min/L1min/L2min (:1207-1214). */
int tdcc_set_state_device(tdcc_ctx* ctx, const double* t1, const double* t2, const double* l1,
                          const double* l2);
int tdcc_get_state_device(tdcc_ctx* ctx, int which, double* t1, double* t2, double* l1, double* l2);

/* One right-hand side at the given complex F (n,n); out has the shape of the
 * residual (o,v) or (o,o,v,v).  Unused amplitude arguments may be NULL only
 * for the kinds that do not read them (T1/T2 kinds ignore lam1/lam2). */
int tdcc_rhs(tdcc_ctx* ctx, int kind, const double* F, const double* t1, const double* t2,
             const double* l1, const double* l2, double* out);
/* tdcc_rhs on DEVICE pointers (F, amplitudes and out as for
 * tdcc_set_state_device): the residual functions T1eq_rhs, T2eq_rhs,
 * lam_1eq_rhs, lam2eq_rhs (CCSD_Helper.py:177, :268, :555, :610) and the CC2
 * kinds T2eq_rhs_CC2, L1eq_rhs_CC2, L2eq_rhs_CC2 (CC2_Helper.py:12, :50,
 * :128) with their arguments and result in GPU memory; bit for bit the result
 * of tdcc_rhs. */
int tdcc_rhs_device(tdcc_ctx* ctx, int kind, const double* F, const double* t1, const double* t2,
                    const double* l1, const double* l2, double* out);

/* One named intermediate of the reference's residuals, evaluated with the
 * same graph builders the residuals use (parity checking): "Fae", "Fmi",
 * "Fme", "Wmnij", "Wmnij_2", "Wabef", "Wabef_2", "Wmbej" (CCSD_Helper.py:136-265),
 * "LRFea", "LRFim", "LSWieam", "LRWefam", "LRWibjm", "Gfe", "Gmn", "LWfiea",
 * "LWmina", "LSWmnij", "LSWabef" (:447-607), "Dij", "Dab", "Dai" (:984-1006).
 * out receives out_len complex values in the reference's index order. */
int tdcc_intermediate(tdcc_ctx* ctx, const char* name, const double* F, const double* t1,
                      const double* t2, const double* l1, const double* l2, double* out,
                      int64_t out_len);

/* Dipole expectation value of the correlated density for (t1,t2,lam1,lam2):
 * mu_out[3] complex (x, y, z).  CCSD_Helper.dipole_moment, CCSD_Helper.py:1066. */
int tdcc_dipole(tdcc_ctx* ctx, const double* t1, const double* t2, const double* l1,
                const double* l2, double* mu_out);

/* No-allocation variant of tdcc_dipole for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  It has no counterpart in the reference.
 * t1, t2, lam1, lam2 and mu_out are DEVICE pointers: the amplitudes
 * interleaved complex in C order (the memory of a complex128 tensor), mu_out
 * 3 complex (x, y, z; 6 doubles).  The result is that of tdcc_dipole for the
 * same amplitudes (the same program, bit for bit).  The call allocates
 * nothing: the program is planned at tdcc_create, and its scratch (density
 * blocks, contraction temporaries, reduction partials) lives in a static
 * device array of 64 MiB per GPU, shared by every call on that GPU and kept
 * until the process ends (calls that overlap on one GPU race).  A system
 * whose dipole program needs more scratch (o, v far beyond the workloads
 * here) gets an error naming the size.  The amplitudes pass through the
 * context's own input planes; the call returns after the result is written. */
int tdcc_dipole_device_noalloc(tdcc_ctx* ctx, const double* t1, const double* t2, const double* l1,
                               const double* l2, double* mu_out);

/* Run `nsteps` iterations of the reference's RK4 loop body on the current state
 * (CCSD_Helper.py:1206-1240 / CC2_Helper.py:453-485 without the stop rules and
 * the file output, which the caller applies): time starts at t0 and is
 * advanced by dt before each step; the field is -A mu[pol] exp(i 2 pi w0 t).
 * Amplitudes are rounded to `precs` decimals before each update.
 * mu_out: nsteps x 3 complex (dipole after each step); t_out: nsteps times. */
int tdcc_rk4_run(tdcc_ctx* ctx, int method, int nsteps, double t0, double dt, double w0,
                 double A, int pol, int precs, double* mu_out, double* t_out);

/* Batched trajectories of this system (many field polarisations, frequencies
 * and amplitudes; the reference runs one trajectory per Runge_Kutta_solver
 * call, CCSD_Helper.py:1191).  tdcc_set_batch allocates nb state slots
 * (0 frees them); slots are filled and read like the single state. */
int tdcc_set_batch(tdcc_ctx* ctx, int nb);
int tdcc_set_state_slot(tdcc_ctx* ctx, int slot, const double* t1, const double* t2, const double* l1,
                        const double* l2);
int tdcc_get_state_slot(tdcc_ctx* ctx, int slot, int which, double* t1, double* t2, double* l1, double* l2);

/* The loop body of tdcc_rk4_run for every slot at once: slot b uses the field
 * -A[b] mu[pol[b]] exp(i 2 pi w0[b] t); all slots share t0, dt and precs.  One
 * captured graph advances all slots by one step: one step program with a
 * trajectory axis over the slots, in which contractions of shared integral
 * blocks with the slots' amplitudes run as one GEMM over the slots' stacked
 * rows (TDCC_STACK=0: the slots' step programs side by side, one launch per
 * slot).  mu_out: nb x nsteps x 3 complex (slot-major); t_out: nsteps times. */
int tdcc_rk4_run_batch(tdcc_ctx* ctx, int method, int nsteps, double t0, double dt, const int* pol,
                       const double* w0, const double* A, int precs, double* mu_out, double* t_out);

/* Gradient of the windowed dipole L = sum_n weights[n] mu_pol(t_n) over nsteps
 * RK4 steps from the current state (the reference's loop, CCSD_Helper.py:
 * 1206-1240, without stop rules), with respect to the field amplitude A, the
 * frequency w0 and the dipole matrix mu_pol (which enters both the field
 * -A mu_pol exp(i 2 pi w0 t) and the read-out).  Hand-written adjoint of the
 * step (transposed contractions; the rounding to precs decimals is treated as
 * the identity), checkpoints every checkpoint_every steps with recompute of
 * each segment.  Outputs (complex, interleaved): loss[1], dA[1], dw0[1],
 * dmu[n*n] (n = o + v).  The state is left at the end of the propagation.
 * There is no reference counterpart (the reference has no derivatives). */
int tdcc_dipole_gradient(tdcc_ctx* ctx, int method, int nsteps, double t0, double dt, double w0, double A,
                         int pol, int precs, const double* weights, int checkpoint_every, double* loss,
                         double* dA, double* dw0, double* dmu);

/* Take one step of the same program eagerly with a HIP event pair around
 * every kernel launch; `report` receives one line per launch:
 * "<kernel name>\t<algorithmic flops>\t<algorithmic bytes>\t<milliseconds>\n". */
int tdcc_profile_step(tdcc_ctx* ctx, int method, double t0, double dt, double w0, double A, int pol,
                      int precs, char* report, int64_t report_len);

/* Device-side copy of the state: op 0 saves it, op 1 restores it. */
int tdcc_checkpoint(tdcc_ctx* ctx, int op);

/* Executed contraction flops (real) of one RK4 step program, and its device
 * arena size in bytes. */
int tdcc_step_info(tdcc_ctx* ctx, int method, double dt, int64_t* flops, int64_t* arena_bytes);

/* Restart block writer, replacing write_2data / write_4data
 * (CCSD_Helper.py:1107-1125): x (rank 2 or 4, C order, float64) as one text
 * row per element, "i\ta\tvalue\n" or "i\tj\ta\tb\tvalue\n", value =
 * numpy.around(x, precs) printed as Python's csv module prints a numpy
 * float64.  Byte-identical to the reference's files; no context, no device.
 * Status 0 = ok; tdcc_io_error() gives the message of the last failure on
 * the calling thread. */
int tdcc_write_restart_block(const double* x, int rank, const int64_t* dims, int precs, const char* path);

/* Restart block reader, replacing convert_2data / convert_4data
 * (CC_Calculator.py:85-113): out (dims, C order, float64) is zeroed, then each
 * row "i<TAB>a<TAB>value" or "i<TAB>j<TAB>a<TAB>b<TAB>value" assigns value
 * (parsed to the nearest double) at the indices (numbers rounded to the
 * nearest integer), in file order; blank lines are skipped, a malformed row or
 * an index outside dims is an error.  No context, no device. */
int tdcc_read_restart_block(const char* path, int rank, const int64_t* dims, double* out);

/* Exact binary snapshot of a stopped run, for resuming bit for bit (the text
 * blocks above quantize the state and the reference resumes one step later in
 * time, CC_Calculator.py:77-80; there is no reference counterpart).  Layout,
 * little-endian: 16-byte magic "TDCC-STATE-V1", int32 o, v, method, precs,
 * float64 t, dt, w0, A (params[0..3]), float64 F[(o+v)^2], then t1, t2, lam1,
 * lam2 as interleaved complex in C order, then a 64-bit FNV-1a checksum of all
 * preceding bytes.  Written to path + ".tmp" and renamed. */
int tdcc_write_state_file(const char* path, int o, int v, int method, int precs, const double* params,
                          const double* F, const double* t1, const double* t2, const double* l1, const double* l2);
int tdcc_read_state_file_header(const char* path, int* o, int* v, int* method, int* precs, double* params);
int tdcc_read_state_file(const char* path, double* F, double* t1, double* t2, double* l1, double* l2);
const char* tdcc_io_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TDCC_H */
