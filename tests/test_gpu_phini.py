"""phini on the GPU (fribo_phini_device: the symmetrised z operator of
phini.F90:121-142 decomposed by rocSOLVER's divide-and-conquer STEDC) against
the operator itself and against the host decomposition (fribo_phini,
implicit QL).  Bars: eigen-residual and inverse pair to 1e-12 (relative to
the operator's and the identity's scale), sorted eigenvalues to 1e-12 of
the largest, and the same step as with the host decomposition (multi-step
bar 1e-10)."""
import numpy as np
import pytest

from fribois2d_amd import cases
from helpers import case, gpu_for, relmax, run_steps

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["cyl64", "cyl384", "rect2048"])
def test_device_phini_decomposes_the_operator(name):
    from fribois2d_amd import gpu
    c = case(name)
    zd, ztd, wd = gpu.phini(c.zz, c.params.alx3, device=True)
    zh, zth, wh = gpu.phini(c.zz, c.params.alx3)
    n = wd.size
    am, ac, ap = cases.tridiag_z(c.zz, c.params.alx3)
    A = np.diag(ac) + np.diag(am[1:], -1) + np.diag(ap[:-1], 1)
    anorm = np.max(np.abs(A))
    # each column is an eigenvector: A Y = Y diag(w), Y-scaled residual
    res = np.max(np.abs(A @ zd - zd * wd[None, :])) / (anorm * np.max(np.abs(zd)))
    print(f"{name}: n={n} eigen-residual {res:.2e}", flush=True)
    assert res <= 1e-12
    inv = np.max(np.abs(ztd @ zd - np.eye(n)))
    print(f"{name}: |Yinv Y - I| {inv:.2e}", flush=True)
    assert inv <= 1e-12
    ev = np.max(np.abs(np.sort(wd) - np.sort(wh))) / np.max(np.abs(wh))
    print(f"{name}: eigenvalues vs host QL {ev:.2e}", flush=True)
    assert ev <= 1e-12
    assert np.all(np.diff(wd) >= 0.0)          # STEDC: increasing order


def test_device_phini_gives_the_same_step():
    from fribois2d_amd import gpu
    c = case("cyl64")
    dh = gpu.phini(c.zz, c.params.alx3)
    dd = gpu.phini(c.zz, c.params.alx3, device=True)
    gh = gpu_for(c, dh, axis="z")
    gd = gpu_for(c, dd, axis="z")
    run_steps(gh, c, 10)
    run_steps(gd, c, 10)
    sh, sd = gh.get_state(), gd.get_state()
    for f in ("q2", "q3", "pr"):
        e = relmax(sd[f], sh[f])
        print(f"device vs host phini after 10 steps, {f}: {e:.2e}", flush=True)
        assert e <= 1e-10, f
