"""Oracle: CPU restatement of the reference's real-time TDCCSD/TDCC2 path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this package; only
`tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`
may use it, and only as the checker (or the timed CPU baseline), never as the
thing measured on the GPU.

Modules:
  cc_equations  numpy complex128 restatement of the CCSD T/Lambda residuals,
                their intermediates, the CC2 residuals and the 1-RDM / dipole
                (CCSD_Helper.py:135-321, :446-648, :984-1092; CC2_Helper.py:12-175)
  rk4           the reference's partial-RK4 propagator with its rounding and
                time-advance semantics (CCSD_Helper.py:1158-1256,
                CC2_Helper.py:412-500)
  ground_state  Jacobi + DIIS solver for the ground-state T and Lambda inputs
  fci           brute-force two-electron FCI used to pin the CCSD fixed point
  flops         symbolic evaluation of the residuals that counts FLOPs per RK4
                step (as written, and with stage/chain de-duplication)

Pins (each function's check is in tests/):
  * golden vectors written from the reference itself (tests/golden/) for every
    residual, intermediate, the dipole and a short RK4 trajectory;
  * zero-amplitude closed forms, CC2/MP2 fixed point, antisymmetry, affinity in
    Lambda, holomorphy;
  * o=2 CCSD == FCI (energy) and Lambda 1-RDM == FCI 1-RDM;
  * the RK4 driver against the closed form for a linear mock right-hand side.
"""
