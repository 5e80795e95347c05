"""Mixed-precision Poisson solve (fribo_set_precision): FP32 MFMA transforms
with FP32 accumulation, FP64 per-mode line solves, and n sweeps of
FP64-residual iterative refinement.

Bars:
* the FP32 transform kernel against numpy's FP64 product within the
  rounding bound of FP32 inputs and accumulation, (K + 2) 2^-24 |A||B|
  elementwise (a wrong fragment layout gives O(1) errors);
* per-field errors against the FP64 step and the oracle, and the largest
  cell divergence after the projection, for 0..2 refinement sweeps
  (measured on the 384x288 cylinder after 3 steps: without refinement the
  fields are off by 1e-6 and the divergence is 7e-5 against the FP64
  step's 1.4e-13; one sweep brings them to 3e-11 and 1.4e-9, two sweeps to
  1e-13 and 1.2e-14): the fields within 1e-4 / 1e-10 / 1e-12 of the
  oracle's, the divergence below 1e-3 / 1e-8 / 10x the FP64 step's;
* 1000 coupled steps of the spring-mounted 384x288 cylinder: the lift,
  drag, moment and displacement histories against the FP64 oracle's
  (tests/golden/cyl384_spring_1000steps.npz), relative to their largest
  value after step 20 (measured: lift 1.5e-4 without refinement, 2.5e-9
  with one sweep): within 1e-3 without refinement (FP32 transforms alone
  miss the 1e-6 lift bar), 1e-8 with one sweep.
"""
import os

import numpy as np
import pytest

from fribois2d_amd import cases
from helpers import GOLDEN, case, decomposition, divergence, gpu_for, oracle_for, relmax, run_steps

pytestmark = pytest.mark.gpu


def test_gemm32_matches_numpy():
    from fribois2d_amd.gpu import gemm32
    rng = np.random.default_rng(5)
    for (M, N, K) in ((200, 300, 150), (257, 129, 333), (64, 64, 16)):
        A = rng.standard_normal((M, K))
        B = rng.standard_normal((K, N))
        C = gemm32(A, B)
        ref = A @ B
        bound = (K + 2) * 2.0 ** -24 * (np.abs(A) @ np.abs(B))
        worst = float(np.max(np.abs(C - ref) / bound))
        print(f"{M}x{N}x{K}: max |C32 - C64| / bound = {worst:.3f}", flush=True)
        assert worst <= 1.0


@pytest.mark.parametrize("refine", [0, 1, 2])
def test_mixed_solve_against_fp64_and_the_oracle(refine):
    c = case("cyl384")
    dec = decomposition(c)
    g64 = gpu_for(c, dec, axis="y")
    gm = gpu_for(c, dec, axis="y")
    gm.set_precision(refine)
    assert gm.precision == refine
    o = oracle_for(c, dec)
    n = 3
    h64, hm, ho = [], [], []
    run_steps(g64, c, n, hist=h64)
    run_steps(gm, c, n, hist=hm)
    run_steps(o, c, n, hist=ho)
    s64, sm = g64.get_state(), gm.get_state()
    d64, dm = divergence(s64, c), divergence(sm, c)
    print(f"refine {refine}: max |div| FP64 {d64:.2e}, mixed {dm:.2e}", flush=True)
    for k in ("q2", "q3", "pr"):
        e64 = relmax(sm[k], s64[k])
        eo = relmax(sm[k], o.get(k))
        print(f"  {k}: vs FP64 step {e64:.2e}, vs oracle {eo:.2e}", flush=True)
        assert eo <= (1e-4, 1e-10, 1e-12)[refine], k
    assert dm <= (1e-3, 1e-8, 10 * d64)[refine]
    g64.close()
    gm.close()


@pytest.mark.parametrize("refine", [0, 1])
def test_mixed_coupled_lift_over_1000_steps(refine):
    import sys
    sys.path.insert(0, GOLDEN)
    from make_coupled_golden import COLS, NSTEPS, setup, trajectory
    ref = np.load(os.path.join(GOLDEN, f"cyl384_spring_{NSTEPS}steps.npz"))
    c, p, body = setup("cyl384_spring")
    g = gpu_for(c, decomposition(c), params=p, axis="y")
    g.set_precision(refine)
    hist = trajectory(g, body, c, NSTEPS)
    bar = 1e-3 if refine == 0 else 1e-8
    for col, name in enumerate(COLS):
        if name not in ("disp", "lift", "drag", "moment"):
            continue
        scale = np.max(np.abs(ref["hist"][20:, col]))
        err = float(np.max(np.abs(hist[:, col] - ref["hist"][:, col])) / scale)
        print(f"refine {refine}: {name} max |diff| / max after step 20 = {err:.2e}", flush=True)
        assert err <= bar, name
    g.close()
