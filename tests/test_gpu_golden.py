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


@pytest.mark.parametrize("axis", ["y", "z"])
@pytest.mark.parametrize("name", ["cyl384_spring", "osccyl_spring"])
def test_spring_mounted_body_1000_steps(name, axis):
    """A body on a spring, weakly coupled as main.F90:79-101 does it (rigid()
    with the previous loads, the fluid step, loads_transfer_c on the device),
    1000 steps against the oracle's run of the same inputs
    (tests/golden/make_coupled_golden.py): the reference's own OscCyl1DOF
    example and the 384x288 cylinder at m* = 10, Ur = 5.
    Bars: displacement, velocity, lift, drag and moment histories within
    1e-9 of each history's largest value after the start transient (steps
    20 on; the impulsive start makes the first loads 10-100 times the
    later ones), dt and cflm within 1e-9 relative, and the final fields
    within 1e-10 of their maximum."""
    import sys
    sys.path.insert(0, GOLDEN)
    from make_coupled_golden import COLS, NSTEPS, setup, summary, trajectory
    ref = np.load(os.path.join(GOLDEN, f"{name}_{NSTEPS}steps.npz"))
    c, p, body = setup(name)
    g = gpu_for(c, decomposition(c), params=p, axis=axis)
    print(f"running {NSTEPS} coupled GPU steps of {name}, transform along {axis}", flush=True)
    hist = trajectory(g, body, c, NSTEPS)
    href = ref["hist"]
    for col, cname in enumerate(COLS):
        if cname == "ns":
            continue
        scale = np.max(np.abs(href[20:, col]))
        err = np.max(np.abs(hist[:, col] - href[:, col])) / scale
        print(f"{cname}: max |diff| / max after step 20 = {err:.2e}", flush=True)
        assert err <= 1e-9, cname
    st = g.get_state()
    sub, norms = summary({k: st[k] for k in ("q2", "q3", "pr")})
    for k in ("q2", "q3", "pr"):
        a, b = sub[f"{k}_sub"], ref[f"{k}_sub"]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        nerr = abs(norms[f"{k}_l2"] - float(ref[f"{k}_l2"])) / float(ref[f"{k}_l2"])
        print(f"{k}: max rel err {err:.2e}, L2-norm rel err {nerr:.2e}", flush=True)
        assert err <= 1e-10 and nerr <= 1e-10, k
    g.close()


@pytest.mark.parametrize("axis", ["y", "z"])
def test_rect2048_against_the_oracle(axis):
    """The benchmarked workload (2048x1536 cells, 1000 markers) against the
    oracle's 3 steps of it (tests/golden/make_rect2048_golden.py): marker
    indices bit-exact, dt/cflm within 1e-12 relative, q2, q3 and pr on every
    16th node within 1e-10 of their maximum and their L2 norms over the
    whole arrays within 1e-10 relative (measured: velocities 1e-14, pressure
    1.3e-11 along y and 2e-13 along z, whose order the oracle shares)."""
    import sys
    sys.path.insert(0, GOLDEN)
    from make_rect2048_golden import FIELDS, NSTEPS, summary
    from helpers import run_steps
    ref = np.load(os.path.join(GOLDEN, f"rect2048_{NSTEPS}steps.npz"))
    c = case("rect2048")
    g = gpu_for(c, decomposition(c), axis=axis)
    hist = []
    run_steps(g, c, NSTEPS, hist=hist)
    hist = np.array(hist)
    np.testing.assert_allclose(hist[:, 1:], ref["hist"][:, 1:], rtol=1e-12, atol=0)
    out = summary(g.get)
    np.testing.assert_array_equal(g.get("lmind").astype(np.int16), ref["lmind"])
    np.testing.assert_array_equal(g.get("eulins").astype(np.int16), ref["eulins"])
    for k in FIELDS:
        a, b = out[f"{k}_sub"], ref[f"{k}_sub"]
        err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        nerr = abs(float(out[f"{k}_l2"]) - float(ref[f"{k}_l2"])) / float(ref[f"{k}_l2"])
        print(f"{k} ({axis}): max rel err {err:.2e}, L2 rel err {nerr:.2e}, "
              f"sum diff {abs(float(out[f'{k}_sum']) - float(ref[f'{k}_sum'])):.2e}", flush=True)
        assert err <= 1e-10 and nerr <= 1e-10, k
    g.close()
