"""Grid shapes the generated workloads do not reach, on the GPU against the
oracle: the widest y extent this build accepts (n2-1 = 4096: 64 rows per
j-chunk, so the j-lines take the kernels whose factors stay in global
memory) and a width between the LDS-staged shapes (n2-1 = 1900), both with
few planes, so the Poisson transform runs along z (the reference's order)
through the small-tile GEMM.  One past the limit is refused.

Bars: the multi-step bar, 1e-10 of the field maximum, after three steps."""
import numpy as np
import pytest

from fribois2d_amd import cases
from helpers import decomposition, gpu_for, oracle_for, relmax, run_steps

pytestmark = pytest.mark.gpu


def _case(n2m, n3m=48):
    dy, dz = 0.01, 0.02
    rc = np.linspace(0.0, n2m * dy, n2m + 1)
    zz = np.linspace(0.0, n3m * dz, n3m + 1)
    p = cases.FluidParams(n2=n2m + 1, n3=n3m + 1, alx3=n3m * dz, rext2=n2m * dy, visc=0.01, dt0=1.0e-4)
    yc, zc = 0.5 * n2m * dy, 0.5 * n3m * dz
    verts = cases.cylinder_vertices(yc, zc, 0.06, 24)
    return cases.Case(f"wide{n2m}", p, rc, zz, verts, (yc, zc), None, 20261202)


@pytest.mark.parametrize("n2m", [1900, 4096])
def test_wide_grids_match_the_oracle(n2m):
    c = _case(n2m)
    dec = decomposition(c)
    o = oracle_for(c, dec)
    g = gpu_for(c, dec)
    assert g.transform == "z"
    ho, hg = [], []
    run_steps(o, c, 3, hist=ho)
    run_steps(g, c, 3, hist=hg)
    assert relmax([h[1] for h in hg], [h[1] for h in ho]) <= 1e-12
    so, sg = {k: o.get(k) for k in ("q2", "q3", "pr")}, g.get_state()
    for f in ("q2", "q3", "pr"):
        e = relmax(sg[f], so[f])
        print(f"n2m={n2m} {f}: {e:.2e}", flush=True)
        assert e <= 1e-10, f


def test_too_wide_grid_is_refused():
    from fribois2d_amd import gpu
    c = _case(4097, 16)
    with pytest.raises(gpu.FriboError) as ei:
        gpu_for(c, decomposition(c))
    assert "n2" in str(ei.value)
