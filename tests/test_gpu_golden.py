"""1000-step parity on the 384x288 cylinder workload against the committed
oracle trajectory (tests/golden/cyl384_1000steps.npz, make_golden.py).

Bars: velocity/pressure max difference <= 1e-10 relative to the
field's maximum after 1000 steps; per-step histories (dt, CFL, immersed-boundary
force sums) within 1e-9 relative to the largest value of each history.
"""
import os

import numpy as np
import pytest

from helpers import GOLDEN, case, decomposition, gpu_for

pytestmark = pytest.mark.gpu


def test_cyl384_1000_steps():
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import NSTEPS, summary, trajectory
    ref = np.load(os.path.join(GOLDEN, "cyl384_1000steps.npz"))
    c = case("cyl384")
    g = gpu_for(c, decomposition(c))
    print("running 1000 GPU steps of cyl384", flush=True)
    hist = trajectory(g, c, NSTEPS, g.get)
    st = g.get_state()
    sub, norms = summary({k: st[k] for k in ("q2", "q3", "pr")})
    href = ref["hist"]
    for col, name in ((1, "dt"), (2, "cflm"), (3, "sum fey"), (4, "sum fez")):
        err = np.max(np.abs(hist[:, col] - href[:, col])) / np.max(np.abs(href[:, col]))
        print(f"history {name}: max rel err {err:.2e}", flush=True)
        assert err <= 1e-9, name
    for k in ("q2", "q3", "pr"):
        a, b = sub[f"{k}_sub"], ref[f"{k}_sub"]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        nerr = abs(norms[f"{k}_l2"] - float(ref[f"{k}_l2"])) / float(ref[f"{k}_l2"])
        print(f"{k}: max rel err {err:.2e}, L2-norm rel err {nerr:.2e}", flush=True)
        assert err <= 1e-10, k
        assert nerr <= 1e-10, k
