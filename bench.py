"""Benchmark: TDCCSD (T + Lambda) RK4 steps/s on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--method CCSD] [--o 20 --v 160]

Workload (defaults): real-time TDCCSD at o = 20, v = 160 spin orbitals, the
reference's partial-RK4 propagator (four decoupled chains, 16 right-hand sides
per step, rounding to 15 decimals, dipole after every step), dt = 0.001,
z field with w0 = 0.923, A = 0.002.  Inputs are synthetic (seeded Philox
generator; the TEI is generated on the device), the initial state is the MP2
guess for t2 with t1 = 0 and lam = t (the ground-state solve is not part of the
timed path).

Two legs run from the same initial state at t = 0: the FP64 path, then the
split-precision path (--digits int8 digits per operand for the large
contractions; FP64 state).  `value` is steps/s of the split leg (of the FP64 leg
with --digits 0); `config` carries the FP64 rate, the speedup and the max
|mu_z| difference between the legs over the timed steps (--error-steps N adds
an N-step comparison).  Each timed region is graph-replayed steps with the
inputs resident in HBM, synchronised on both sides.

`roofline` comes from one eager profiled step after the timed region (HIP
events around every launch): the dominant kernel is the kernel name with the
largest summed time.  `cpu_baseline` times the oracle (numpy complex128, the
reference's einsum strings) on a bounded sample and scales it to the workload
by F_ref.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the digit-GEMM launch autotuner (tile order, persistent grid, K splits; see
# include/tdcc.h) runs at program bind, which the warm-up steps cover
os.environ.setdefault("TDCC_AUTOTUNE", "1")

METRIC = "TDCCSD RK4 steps/s (T+Λ) o=20,v=160: FP32-split vs FP64 at dipole err≤1e-8"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
PEAK_INT8_TOPS = 5000.0


def initial_state(sysm, o, v):
    from python_code_to_check_c_amd import synthetic
    eps = np.linalg.eigh(sysm.F)[0]
    eo, ev = eps[:o], eps[o:]
    D = eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev
    oovv = synthetic.tei_block(sysm.config_id, o, v, "oovv", sysm.tei_sigma)
    t2 = (oovv / D).astype(complex)
    t1 = np.zeros((o, v), complex)
    return t1, t2, t1.copy(), t2.copy()


def _host_info():
    """Host BLAS, CPU model and thread count, for the CPU-baseline record."""
    info = {"cores": len(os.sched_getaffinity(0))}
    try:
        import threadpoolctl
        np.dot(np.ones((64, 64)), np.ones((64, 64)))
        pools = [p for p in threadpoolctl.threadpool_info() if p.get("user_api") == "blas"]
        if pools:
            p = pools[0]
            info["blas"] = f"{p.get('internal_api')} {p.get('version')} ({p.get('architecture')}), " \
                           f"{p.get('num_threads')} threads"
    except Exception as e:  # the record still carries the rest
        info["blas"] = f"unknown ({e})"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def cpu_baseline(method, o_full, v_full, sample=(20, 96)):
    """The oracle (numpy complex128, the reference's einsum strings with
    optimize='optimal', host BLAS) on one evaluation of each of the method's
    four right-hand sides at the sample shape: a quarter of an RK4 step's
    as-written work (each chain's first stage).  The step time is four times
    that plus the dipole; it is scaled to the workload by F_ref(full) /
    F_ref(sample), the flop ratio of the two shapes' as-written steps."""
    import resource

    from oracle import rk4
    from oracle.cc_equations import CCEquations
    from oracle.flops import rhs_flops, step_flops
    from python_code_to_check_c_amd import synthetic
    o, v = min(o_full, sample[0]), min(v_full, sample[1])
    s = synthetic.SyntheticSystem(90, o, v, tei_sigma=0.05)
    eq = CCEquations(s.TEI, o)
    t1, t2, l1, l2 = synthetic.perturbed_state(90, o, v, 0.05)
    F = s.F + rk4.field(s.mu[2], 0.002, 0.923, 0.001)
    names = ("T1eq_rhs", "T2eq_rhs", "lam_1eq_rhs", "lam2eq_rhs") if method == "CCSD" else \
        ("T1eq_rhs", "T2eq_rhs_CC2", "L1eq_rhs_CC2", "L2eq_rhs_CC2")
    t0 = time.perf_counter()
    for nm in names:
        f = getattr(eq, nm)
        f(t1, t2, F) if nm in ("T1eq_rhs", "T2eq_rhs", "T2eq_rhs_CC2") else f(t1, t2, l1, l2, F)
    el_rhs = time.perf_counter() - t0
    t0 = time.perf_counter()
    eq.dipole_moment(t1, t2, l1, l2, s.mu)
    el_dip = time.perf_counter() - t0
    step_s = 4.0 * el_rhs + el_dip
    f_small = step_flops(o, v, method)[0]
    f_full = step_flops(o_full, v_full, method)[0]
    f_quarter = sum(rhs_flops(o, v, nm) for nm in names)
    info = _host_info()
    rss_gb = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 2**20
    return {"value": (1.0 / step_s) * f_small / f_full, "unit": "steps/s", "cores": info["cores"], "kind": "port",
            "sample": f"oracle (numpy complex128, reference einsum strings, optimize='optimal') at o={o}, v={v}: "
                      f"one evaluation of each of {', '.join(names)} ({f_quarter:.3e} flop as written, a quarter "
                      f"of the step) in {el_rhs:.2f} s ({f_quarter / el_rhs / 1e9:.1f} GFLOP/s) and the dipole in "
                      f"{el_dip:.3f} s -> {step_s:.1f} s per RK4 step; scaled to o={o_full}, v={v_full} by F_ref "
                      f"{f_small:.3e} -> {f_full:.3e} flop/step (x{f_full / f_small:.2f})",
            "blas": info.get("blas"), "cpu_model": info.get("cpu_model"),
            "process_peak_rss_GiB": round(rss_gb, 2)}


def roofline(records, digits):
    """Dominant kernel (largest summed time) of one profiled step.  MFMA kernels
    are priced against the FP64 matrix peak, or, for the int8 digit GEMM,
    against the int8 dense peak divided by the digit products per FP64 product."""
    groups = {}
    for name, fl, by, ms, _ in records:
        g = groups.setdefault(name, [0.0, 0.0, 0.0, 0])
        g[0] += fl
        g[1] += by
        g[2] += ms
        g[3] += 1
    name, (fl, by, ms, calls) = max(groups.items(), key=lambda kv: kv[1][2])
    avg_ms = ms / calls
    if fl > 0:
        achieved = fl / (ms * 1e-3) / 1e12
        peak, note = PEAK_FP64_TFLOPS, "FP64 dense MFMA"
        if name.startswith("ozgemm"):
            pairs = digits * (digits + 1) // 2
            peak = PEAK_INT8_TOPS / pairs
            note = (f"FP64-equivalent flops; int8 dense MFMA {PEAK_INT8_TOPS:.0f} TOP/s / {pairs} digit products "
                    f"per FP64 product ({digits} digits)")
        return name, {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                      "frac": achieved / peak, "traffic": None, "kernel": name, "peak_note": note,
                      "calls_per_step": calls, "avg_ms": avg_ms, "flops_per_launch": fl / calls}, groups
    achieved = by / (ms * 1e-3) / 1e9
    return name, {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                  "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": name,
                  "calls_per_step": calls, "avg_ms": avg_ms}, groups


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--method", default="CCSD", choices=["CCSD", "CC2"])
    ap.add_argument("--o", type=int, default=20)
    ap.add_argument("--v", type=int, default=160)
    ap.add_argument("--config-id", type=int, default=3)
    ap.add_argument("--tei-sigma", type=float, default=0.05)
    ap.add_argument("--digits", type=int, default=5,
                    help="int8 digits per operand of the split-precision path (0: FP64 only)")
    ap.add_argument("--error-steps", type=int, default=0,
                    help="also propagate this many steps in both precisions and report the dipole difference")
    ap.add_argument("--error-digits", default=None,
                    help="comma-separated digit counts for the --error-steps comparison (default: --digits)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--dump-profile", default=None, help="write every profiled launch (tsv) to this file")
    ap.add_argument("--split-only", action="store_true",
                    help="time only the split-precision leg (profiling runs; no FP64 comparison)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy kernel over MIB MiB (HBM counter calibration) and exit")
    ap.add_argument("--pmc-summary", nargs=5, default=None,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV", "COPY_MIB"),
                    help="only turn rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE counter files (workload and copy "
                         "calibration) into the per-kernel HBM traffic table on stdout, and exit")
    ap.add_argument("--counter-summary", default=None, metavar="COUNTER_CSV",
                    help="only sum a rocprofv3 --pmc counter csv per kernel (with the MFMA-busy fraction "
                         "SQ_VALU_MFMA_BUSY_CYCLES / (4 SIMDs x SQ_BUSY_CU_CYCLES) and LDS bank-conflict rate "
                         "where those counters are present) and exit")
    ap.add_argument("--trace-summary", nargs="+", default=None, metavar="ARG",
                    help="KERNEL_TRACE_CSV STEPS [BLOCK]: only summarise a rocprofv3 --kernel-trace csv of a bench.py "
                         "leg per step and exit (BLOCK, default 1: the last block of dispatches between the runtime's "
                         "copy kernels, the timed run with --no-profile or the profiled step without it; 2: the one "
                         "before)")
    ap.add_argument("--batched", action="store_true",
                    help="the batched workload instead: 4 molecule shapes x 3 polarisations x 8 frequencies, "
                         "one context and one captured step per molecule (prints its own JSON line)")
    ap.add_argument("--batch-sigma", type=float, default=0.005,
                    help="TEI sigma of the batched workload's molecules")
    ap.add_argument("--gradient", action="store_true",
                    help="the differentiable workload instead: o=10, v=38, 500 steps, gradient of the Hann-windowed "
                         "mu_z with respect to A, w0 and the z dipole matrix (prints its own JSON line)")
    args = ap.parse_args()
    if args.gradient:
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        gradient(args)
        return
    if args.batched:
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        batched(args)
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.counter_summary:
        counter_summary(args.counter_summary)
        return
    if args.trace_summary:
        trace_summary(args.trace_summary[0], int(args.trace_summary[1]),
                      int(args.trace_summary[2]) if len(args.trace_summary) > 2 else 1)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of this "
              "path is a design, not a run.", file=sys.stderr)
        sys.exit(2)

    from python_code_to_check_c_amd import _lib, synthetic
    from oracle.flops import step_flops

    o, v = args.o, args.v
    dt, w0, A, pol, precs = 0.001, 0.923, 0.002, 2, 15
    t_setup = time.perf_counter()
    sysm = synthetic.SyntheticSystem(args.config_id, o, v, tei_sigma=args.tei_sigma, tei_on_device=True)
    ctx = _lib.Context(o, v)
    ctx.set_system(None, sysm.F, sysm.mu)
    ctx.synth_tei(args.config_id, args.tei_sigma)
    ctx.set_state(*initial_state(sysm, o, v))
    ctx.checkpoint(0)  # X0: both legs start here at t = 0
    flops_exec, arena = ctx.step_info(args.method, dt)
    log(f"setup {time.perf_counter() - t_setup:.1f} s; FP64 step arena {arena / 2**30:.2f} GiB; "
        f"executed contraction flops/step {flops_exec:.4e}")

    def leg(digits):
        ctx.set_precision(digits)
        if args.warmup:
            ctx.checkpoint(1)
            ctx.rk4_run(args.method, args.warmup, 0.0, dt, w0, A, pol, precs)
        ctx.checkpoint(1)
        t0 = time.perf_counter()
        mu, ts = ctx.rk4_run(args.method, args.steps, 0.0, dt, w0, A, pol, precs)
        el = time.perf_counter() - t0
        if not np.all(np.isfinite(mu)):
            raise RuntimeError("non-finite dipole in the timed region")
        log(f"{'FP64' if digits == 0 else f'split({digits} digits)'}: timed {args.steps} steps in {el:.3f} s")
        return args.steps / el, el, mu, float(ts[-1])

    split = args.digits > 0
    if args.split_only and split:
        sps, el, mu, t_end = leg(args.digits)
        sps64, el64, mu64 = float("nan"), float("nan"), None
    else:
        sps64, el64, mu64, t_end = leg(0)
        if split:
            sps, el, mu, t_end = leg(args.digits)
        else:
            sps, el, mu = sps64, el64, mu64

    f_ref, f_alg = step_flops(o, v, args.method)
    prec = (f"split: int8 digit planes ({args.digits} digits of radix 254 per operand, Ozaki scheme, exact int32 "
            "accumulation on int8 MFMA; not FP32 hi/lo) for the ladders and contractions above 2 GFLOP; FP64 state, "
            "field, small and dot-product contractions and dipole"
            if split else "fp64")
    out = {
        "metric": METRIC, "value": sps, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+int8-split" if split else "f64",
        "data": "synthetic",
        "config": {"workload": f"TD{args.method} (T+Lambda) partial-RK4 propagation, o={o}, v={v} spin orbitals, "
                               f"dt={dt}, z field w0={w0} A={A}, precs={precs}",
                   "method": args.method, "precision": prec, "digits": args.digits, "o": o, "v": v,
                   "tei_sigma": args.tei_sigma, "config_id": args.config_id,
                   "initial_state": "MP2 t2, t1=0, lam=t",
                   "fp64_steps_per_s": sps64, "fp64_ms_per_step": 1e3 * el64 / args.steps,
                   "speedup_vs_fp64": sps / sps64,
                   "F_ref_flop_per_step": f_ref, "F_alg_flop_per_step": f_alg,
                   "F_exec_contraction_flop_per_step": flops_exec,
                   "exec_TFLOPs": flops_exec * sps / 1e12},
    }
    if split and mu64 is not None:
        err = mu[:, 2] - mu64[:, 2]
        out["config"]["dipole_z_err_vs_fp64"] = {"steps": args.steps, "re": float(np.abs(err.real).max()),
                                                 "im": float(np.abs(err.imag).max())}
    if split and not args.split_only:
        out["config"]["residual_err_vs_fp64_at_initial_state"] = residual_errors(
            ctx, args.method, args.digits, sysm, o, v, rk4_field(sysm, A, w0, dt, pol))
    if split and args.error_steps > 0:
        out["config"]["dipole_z_err_vs_fp64_long"] = long_error(ctx, args, dt, w0, A, pol, precs)
    if not args.no_profile:
        ctx.set_precision(args.digits)
        recs = ctx.profile_step(args.method, t_end, dt, w0, A, pol, precs)
        if args.dump_profile:
            with open(args.dump_profile, "w") as f:
                f.write("kernel\tflops\tbytes\tms\tdetail\n")
                for r in sorted(recs, key=lambda r: -r[3]):
                    f.write("\t".join(str(x) for x in r) + "\n")
        name, roof, groups = roofline(recs, args.digits)
        roof["traffic"] = traffic_of(name)
        if roof["traffic"] is not None:
            roof["traffic_source"] = os.path.relpath(TRAFFIC_TABLE, ROOT)
        out["roofline"] = roof
        top = sorted(groups.items(), key=lambda kv: -kv[1][2])[:6]
        out["config"]["profile_top"] = {k: {"ms_per_step": g[2], "calls": g[3],
                                            "TFLOPs": (g[0] / (g[2] * 1e-3) / 1e12) if g[0] else None}
                                        for k, g in top}
        out["config"]["profiled_step_ms"] = sum(r[3] for r in recs)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.method, o, v)
    ctx.close()
    print(json.dumps(out), flush=True)


BATCH_SHAPES = [(8, 24), (10, 38), (14, 58), (20, 96)]


def batched(args):
    """Batched TDCCSD: for each molecule shape (config id 40 + index, TEI sigma
    --batch-sigma) 24 trajectories (x/y/z field x 8 frequencies w0 ~ U[0.3,
    1.2] from the seeded generator, A = 0.002) in one stacked program per
    molecule (one step graph with a trajectory axis: shared integral operands
    against the 24 trajectories' stacked rows); split precision as --digits.
    Each molecule is timed alone over --steps steps after --warmup; then the
    four molecules run at once, one context (stream) each from its own host
    thread, and that wall time gives `value` = trajectory-steps/s over all 96
    trajectories."""
    import threading

    from python_code_to_check_c_amd import _lib, synthetic
    from oracle.flops import step_flops
    dt, A, precs = 0.001, 0.002, 15
    sigma = args.batch_sigma
    mols = []
    for k, (o, v) in enumerate(BATCH_SHAPES):
        cid = 40 + k
        sysm = synthetic.SyntheticSystem(cid, o, v, tei_sigma=sigma, tei_on_device=True)
        ctx = _lib.Context(o, v)
        ctx.set_system(None, sysm.F, sysm.mu)
        ctx.synth_tei(cid, sigma)
        ctx.set_precision(args.digits)
        w0s = synthetic.w0_draws(cid, 8)
        pol = [p for p in (0, 1, 2) for _ in range(8)]
        w0 = [float(w) for _ in (0, 1, 2) for w in w0s]
        amp = [A] * len(pol)
        state = initial_state(sysm, o, v)
        ctx.set_batch(len(pol))
        mols.append(dict(o=o, v=v, ctx=ctx, pol=pol, w0=w0, amp=amp, state=state))

    def reset(m):
        for b in range(len(m["pol"])):
            m["ctx"].set_state_slot(b, *m["state"])

    def run(m, nsteps):
        mu, _ = m["ctx"].rk4_run_batch(args.method, nsteps, 0.0, dt, m["pol"], m["w0"], m["amp"], precs)
        m["finite"] = bool(np.all(np.isfinite(mu)))

    per_mol, flops = [], 0.0
    for m in mols:
        reset(m)
        if args.warmup:
            run(m, args.warmup)
            reset(m)
        t0 = time.perf_counter()
        run(m, args.steps)
        el = time.perf_counter() - t0
        n_ts = len(m["pol"]) * args.steps
        per_mol.append({"o": m["o"], "v": m["v"], "trajectories": len(m["pol"]), "seconds_alone": el,
                        "trajectory_steps_per_s_alone": n_ts / el, "finite": m["finite"]})
        flops += len(m["pol"]) * step_flops(m["o"], m["v"], args.method)[1]
        log(f"batch o={m['o']} v={m['v']}: {len(m['pol'])} trajectories x {args.steps} steps in {el:.3f} s")
    for m in mols:
        reset(m)
    threads = [threading.Thread(target=run, args=(m, args.steps)) for m in mols]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    total_ts = sum(len(m["pol"]) for m in mols) * args.steps
    log(f"all four molecules at once: {total_ts} trajectory-steps in {wall:.3f} s")
    for m in mols:
        m["ctx"].close()
    value = total_ts / wall
    print(json.dumps({
        "metric": "batched TDCCSD trajectory-steps/s (4 shapes x 24 trajectories)", "value": value,
        "unit": "trajectory-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+int8-split" if args.digits else "f64", "data": "synthetic",
        "config": {"workload": "batched TDCCSD, shapes (8,24),(10,38),(14,58),(20,96), 3 polarisations x 8 "
                               f"frequencies each, A=0.002, dt=0.001, TEI sigma {sigma}, MP2 start",
                   "program": "stacked" if os.environ.get("TDCC_STACK", "1") != "0" else "side by side",
                   "digits": args.digits, "per_molecule": per_mol,
                   "sum_of_alone_seconds": sum(p["seconds_alone"] for p in per_mol),
                   "concurrent_seconds": wall,
                   "F_alg_flop_per_batched_step": flops}}), flush=True)


def gradient(args, o=10, v=38, nsteps=500, ckpt=50):
    """Windowed-dipole gradient: L = sum_n hann_n dt mu_z(t_n) over 500 steps
    at o=10, v=38 (config id 2, TEI sigma 0.02, MP2 start), FP64, checkpoints
    every 50 steps.  Times the forward + adjoint call against one plain
    propagation, and checks dL/dA and dL/dw0 by central differences
    (h = 1e-5 relative) of the propagation itself."""
    from python_code_to_check_c_amd import _lib, synthetic
    dt, w0, A, precs = 0.001, 0.923, 0.002, 15
    sysm = synthetic.SyntheticSystem(2, o, v, tei_sigma=0.02)
    ctx = _lib.Context(o, v)
    ctx.set_system(sysm.TEI, sysm.F, sysm.mu)
    X0 = initial_state(sysm, o, v)
    w = np.hanning(nsteps + 2)[1:-1] * dt

    def loss(A_, w0_):
        ctx.set_state(*X0)
        mu, _ = ctx.rk4_run("CCSD", nsteps, 0.0, dt, w0_, A_, 2, precs)
        return np.sum(w * mu[:, 2])

    loss(A, w0)  # warm the programs
    ctx.set_state(*X0)
    ctx.dipole_gradient("CCSD", 2, 0.0, dt, w0, A, w[:2], pol=2, precs=precs, checkpoint_every=ckpt)
    t0 = time.perf_counter()
    L0 = loss(A, w0)
    t_fwd = time.perf_counter() - t0
    ctx.set_state(*X0)
    t0 = time.perf_counter()
    L, dA, dw, dmu = ctx.dipole_gradient("CCSD", nsteps, 0.0, dt, w0, A, w, pol=2, precs=precs,
                                         checkpoint_every=ckpt)
    t_grad = time.perf_counter() - t0
    hA, hw = 1e-5 * A, 1e-5 * w0
    fdA = (loss(A + hA, w0) - loss(A - hA, w0)) / (2 * hA)
    fdw = (loss(A, w0 + hw) - loss(A, w0 - hw)) / (2 * hw)
    ctx.close()
    c = lambda z: [float(np.real(z)), float(np.imag(z))]
    print(json.dumps({
        "metric": "windowed-dipole gradient wall time (forward + adjoint, 500 RK4 steps)", "value": t_grad,
        "unit": "s", "n_gpus": 1, "steps": nsteps, "warmup": 1, "ms_per_step": 1e3 * t_grad / nsteps,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"TDCCSD o={o}, v={v}, L = sum hann*dt*mu_z over {nsteps} steps, dL/d(A, w0, mu_z), "
                               f"checkpoints every {ckpt} steps",
                   "forward_only_s": t_fwd, "gradient_over_forward": t_grad / t_fwd,
                   "loss": c(L), "loss_forward": c(L0), "dL_dA": c(dA), "dL_dw0": c(dw),
                   "fd_dL_dA": c(fdA), "fd_dL_dw0": c(fdw),
                   "rel_err_dA": float(abs(dA - fdA) / abs(fdA)), "rel_err_dw0": float(abs(dw - fdw) / abs(fdw)),
                   "dL_dmu_max_abs": float(np.abs(dmu).max())}}), flush=True)


def calibrate_copy(mib):
    """A plain device-to-device copy kernel of known size (y = x * 1.0 over
    mib MiB of doubles: mib MiB read, mib MiB written), for scaling the HBM
    traffic counters (FETCH_SIZE / WRITE_SIZE) of the other kernels."""
    import torch
    n = mib * 2**20 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 launches of y = x * 1.0, {mib} MiB read + {mib} MiB written each")


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic_tdccsd_c3_o20_v160.txt")


def _counter_rows(path):
    import csv
    with open(path) as f:
        return [(r["Kernel_Name"].split("(")[0].replace("void ", ""), float(r["Counter_Value"]))
                for r in csv.DictReader(f)]


def counter_summary(path):
    """Per-kernel sums of every counter in a rocprofv3 --pmc csv, kernels by
    SQ_BUSY_CU_CYCLES (or by their first counter)."""
    import csv
    acc, launches = {}, {}
    with open(path) as f:
        for r in csv.DictReader(f):
            n = r["Kernel_Name"].split("(")[0].replace("void ", "")
            acc.setdefault(n, {}).setdefault(r["Counter_Name"], 0.0)
            acc[n][r["Counter_Name"]] += float(r["Counter_Value"])
            launches.setdefault(n, set()).add(r["Dispatch_Id"])
    names = sorted({c for v in acc.values() for c in v})
    key = "SQ_BUSY_CU_CYCLES" if "SQ_BUSY_CU_CYCLES" in names else names[0]
    extra = []
    if {"SQ_VALU_MFMA_BUSY_CYCLES", "SQ_BUSY_CU_CYCLES"} <= set(names):
        extra.append(("mfma_busy_frac", lambda v: v["SQ_VALU_MFMA_BUSY_CYCLES"] / (4 * v["SQ_BUSY_CU_CYCLES"])
                      if v["SQ_BUSY_CU_CYCLES"] else 0.0))
    if {"SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"} <= set(names):
        extra.append(("lds_conflict_rate", lambda v: v["SQ_LDS_BANK_CONFLICT"] / v["SQ_LDS_IDX_ACTIVE"]
                      if v["SQ_LDS_IDX_ACTIVE"] else 0.0))
    print("\t".join(["kernel", "launches"] + names + [e[0] for e in extra]))
    for n, v in sorted(acc.items(), key=lambda kv: -kv[1].get(key, 0.0)):
        print("\t".join([n, str(len(launches[n]))] + [f"{v.get(c, 0.0):.4e}" for c in names] +
                        [f"{e[1](v):.4f}" for e in extra]))


def trace_summary(path, steps, block=1):
    """Per-kernel statistics per RK4 step of a rocprofv3 kernel trace of a
    bench.py leg.  The dispatches fall into blocks separated by the runtime's
    copy kernels; `block` counts them from the end.  With --no-profile the
    last block is the timed run (`steps` graph replays after the state
    upload; set-up, autotuner trials and the warm-up come before it); with the
    profiled step it is that eager step (steps = 1, launches one after another
    on one stream, each between two HIP events: the durations behind the
    roofline figure) and the timed run is block 2.  The replays run a
    two-stream graph, so the order of dispatches differs between steps; the
    counts per step do not."""
    import csv
    with open(path) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Dispatch_Id"]))
    blocks, cur = [], []
    for r in rows:
        if r["Kernel_Name"].startswith("__amd_rocclr"):
            if cur:
                blocks.append(cur)
            cur = []
        else:
            cur.append(r)
    if cur:
        blocks.append(cur)
    sel = blocks[-int(block)]
    acc = {}
    for r in sel:
        acc.setdefault(r["Kernel_Name"], []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
    total = sum(sum(v) for v in acc.values())
    # timeline: wall clock, time with at least one and with two or more kernels running
    ev = sorted([(int(r["Start_Timestamp"]), 1) for r in sel] + [(int(r["End_Timestamp"]), -1) for r in sel])
    level, last, busy, multi = 0, ev[0][0], 0, 0
    for t, d in ev:
        busy += (t - last) if level > 0 else 0
        multi += (t - last) if level > 1 else 0
        level, last = level + d, t
    wall = ev[-1][0] - ev[0][0]
    print(f"# {len(sel) // steps} dispatches per step over {steps} steps ({total * 1e-3:.2f} ms of kernel time, "
          f"{total * 1e-3 / steps:.2f} ms per step)")
    print(f"# timeline per step: {wall * 1e-6 / steps:.2f} ms wall (first start to last end), {busy * 1e-6 / steps:.2f} ms "
          f"with a kernel running, {multi * 1e-6 / steps:.2f} ms with two or more (kernels that share the GPU "
          "run longer than alone, so their durations add up to more than the wall time)")
    print("| kernel | calls per step | ms per step | avg us | min us | max us | % |")
    print("|---|---|---|---|---|---|---|")
    for n, v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{n}` | {len(v) / steps:g} | {sum(v) * 1e-3 / steps:.3f} | {np.mean(v):.1f} | {min(v):.1f} | "
              f"{max(v):.1f} | {100 * sum(v) / total:.1f} |")


def pmc_summary(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, copy_mib):
    """Per-kernel HBM bytes per launch from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch), each scaled so that the copy kernel's counter equals the bytes
    it moves (copy_mib MiB each way)."""
    true_kib = float(copy_mib) * 1024.0
    cf = [v for n, v in _counter_rows(copy_fetch_csv) if "MulFunctor" in n]
    cw = [v for n, v in _counter_rows(copy_write_csv) if "MulFunctor" in n]
    f_scale, w_scale = true_kib / np.mean(cf), true_kib / np.mean(cw)
    acc, wacc = {}, {}
    for n, fv in _counter_rows(fetch_csv):
        a = acc.setdefault(n, [0, 0.0, 0.0])
        a[0] += 1
        a[1] += fv * f_scale * 1024.0
    for n, wv in _counter_rows(write_csv):
        w = wacc.setdefault(n, [0, 0.0])
        w[0] += 1
        w[1] += wv * w_scale * 1024.0
    if {n: a[0] for n, a in acc.items()} != {n: w[0] for n, w in wacc.items()}:
        raise RuntimeError("the FETCH_SIZE and WRITE_SIZE runs dispatched different kernels")
    for n, w in wacc.items():
        acc[n][2] = w[1]
    print(f"# FETCH_SIZE scale {f_scale:.4f}, WRITE_SIZE scale {w_scale:.4f} (copy kernel, {copy_mib} MiB each way)")
    print("# kernel\tlaunches\tfetch_bytes_per_launch\twrite_bytes_per_launch")
    for n, (c, fb, wb) in sorted(acc.items(), key=lambda kv: -(kv[1][1] + kv[1][2])):
        print(f"{n}\t{c}\t{fb / c:.6e}\t{wb / c:.6e}")


def traffic_of(kernel_name):
    """HBM bytes per launch of a kernel from the committed counter table (the
    base name, e.g. ozgemm8_kernel, matches tdcc::ozgemm8_kernel<4>), or None."""
    if not os.path.exists(TRAFFIC_TABLE):
        return None
    base = kernel_name.split("<")[0]
    best = None
    with open(TRAFFIC_TABLE) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            n, c, fb, wb = line.rstrip("\n").split("\t")
            if n.split("::")[-1].split("<")[0] == base and (best is None or int(c) > best[0]):
                best = (int(c), float(fb) + float(wb))
    return best[1] if best else None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rk4_field(sysm, A, w0, t, pol):
    """F + V(t) with V(t) = -A mu_pol exp(i 2 pi w0 t) (CCSD_Helper.py:1197-1200)."""
    import cmath
    return sysm.F + (-A * sysm.mu[pol] * cmath.exp(1j * w0 * 2 * np.pi * t))


def residual_errors(ctx, method, digits, sysm, o, v, F):
    """Split-precision residuals against the FP64 ones at the initial state and
    F + V(dt): max-abs difference and its ratio to the FP64 residual's largest
    element, per residual."""
    from python_code_to_check_c_amd import _lib
    kinds = ((("T1eq_rhs", _lib.RHS_T1), ("T2eq_rhs", _lib.RHS_T2), ("lam_1eq_rhs", _lib.RHS_L1),
              ("lam2eq_rhs", _lib.RHS_L2)) if method == "CCSD" else
             (("T1eq_rhs", _lib.RHS_T1), ("T2eq_rhs_CC2", _lib.RHS_T2_CC2), ("L1eq_rhs_CC2", _lib.RHS_L1_CC2),
              ("L2eq_rhs_CC2", _lib.RHS_L2_CC2)))
    X0 = initial_state(sysm, o, v)
    ref = {}
    ctx.set_precision(0)
    for name, kind in kinds:
        ref[name] = ctx.rhs(kind, F, *X0)
    ctx.set_precision(digits)
    out = {}
    for name, kind in kinds:
        d = np.abs(ctx.rhs(kind, F, *X0) - ref[name]).max()
        m = np.abs(ref[name]).max()
        out[name] = {"max_abs": float(d), "rel_to_max": float(d / m) if m > 0 else None, "max_abs_fp64": float(m)}
    return out


ERROR_WINDOWS = (1, 10, 20, 40, 60, 100, 250, 500, 1000, 1500, 2000)


def long_error(ctx, args, dt, w0, A, pol, precs, chunk=250):
    """Propagate error_steps steps from X0 in FP64 and in split precision (the
    bench's digit count, or each of --error-digits) and return the max |mu_z|
    difference (re and im), also over the first n steps for n in
    ERROR_WINDOWS.  Where the FP64 trajectory leaves the finite numbers the
    comparison ends at its last finite step (the split runs stop there too);
    `reference_stop_step` is the first step at which the reference's loop
    would stop on divergence (|Re mu_z| > 100, CCSD_Helper.py:1248), and the
    difference is also reported over the steps before it."""
    def run(digits, nmax):
        ctx.set_precision(digits)
        ctx.checkpoint(1)
        t, mus, done = 0.0, [], 0
        while done < nmax:
            n = min(chunk, nmax - done)
            mu, ts = ctx.rk4_run(args.method, n, t, dt, w0, A, pol, precs)
            mus.append(mu)
            t = float(ts[-1])
            done += n
            log(f"error run digits={digits}: {done}/{nmax} steps")
            if not np.all(np.isfinite(mu[:, 2])):
                break
        return np.concatenate(mus)[:, 2]

    ref = run(0, args.error_steps)
    fin = np.isfinite(ref)
    n_fin = int(np.argmin(fin)) if not fin.all() else len(ref)
    big = np.abs(ref.real) > 100
    n_stop = int(np.argmax(big)) + 1 if big.any() else None
    res = {"steps": args.error_steps, "compared_steps": n_fin,
           "mu_z_excursion_fp64": float(np.abs(ref[:n_fin] - ref[0]).max()) if n_fin else 0.0,
           "fp64_first_nonfinite_step": None if fin.all() else n_fin + 1,
           "reference_stop_step": n_stop, "by_digits": {}}
    digit_list = [int(x) for x in args.error_digits.split(",")] if args.error_digits else [args.digits]
    for digits in digit_list:
        m = run(digits, n_fin)
        n = min(n_fin, len(m))
        d = m[:n] - ref[:n]
        r = {"compared_steps": n, "re": float(np.abs(d.real).max()), "im": float(np.abs(d.imag).max()),
             "first_n_steps": {str(w): [float(np.abs(d[:w].real).max()), float(np.abs(d[:w].imag).max())]
                               for w in ERROR_WINDOWS if w <= n}}
        if n_stop:
            w = min(n_stop - 1, n)
            r["before_reference_stop"] = ({"steps": w, "re": float(np.abs(d[:w].real).max()),
                                           "im": float(np.abs(d[:w].imag).max())} if w > 0 else {"steps": 0})
        res["by_digits"][str(digits)] = r
        if digits == args.digits:
            res.update({k: r[k] for k in ("re", "im")})
            if "before_reference_stop" in r:
                res["before_reference_stop"] = r["before_reference_stop"]
    return res


if __name__ == "__main__":
    main()
