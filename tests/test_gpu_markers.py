"""Marker placements the generated workloads do not reach, on the GPU against
the oracle: nearest-node ties (lgrdistribution.F90:38-67, MINLOC keeps the
first minimum), positions between the domain edge and the first cell centre
(clamped to cell 1 / n*m), supports shifted inward at the walls and at the
inflow and outflow planes (MLSint.F90:57-72), and markers on the planes
z = 0 and z = alx3, which forcing_MLS.F90:56 leaves out of the forcing.

Uniform 16 x 32 grid of spacing 1/4, so every tie is exact in binary.
Bars: indices bit-exact; weights 1e-14; fields after three steps 1e-10 of
the field maximum (the multi-step bar)."""
import numpy as np
import pytest

from fribois2d_amd import cases
from helpers import decomposition, gpu_for, oracle_for, relmax

pytestmark = pytest.mark.gpu

POINTS = [
    (1.0, 4.0), (1.125, 4.125),             # ties on both grids
    (0.1, 3.0), (3.9, 5.0),                 # first / last cell across y
    (0.05, 0.1), (3.95, 7.9),               # corners: clamped, shifted supports
    (2.0, 0.0), (2.3, 8.0),                 # on the inflow / outflow plane: not forced
    (0.0, 2.5),                             # on the wall y = 0
    (1.6, 4.25), (3.0, 7.95), (0.3, 3.0), (3.7, 5.0),
    (2.0, 0.3), (2.0, 7.7), (0.125, 2.0), (3.875, 6.0), (2.0, 0.125), (2.0, 7.875),
]


def _case():
    n2m, n3m = 16, 32
    rc = np.linspace(0.0, 4.0, n2m + 1)
    zz = np.linspace(0.0, 8.0, n3m + 1)
    p = cases.FluidParams(n2=n2m + 1, n3=n3m + 1, alx3=8.0, rext2=4.0, visc=0.02)
    return cases.Case("grid16", p, rc, zz, None, (0.0, 0.0), None, 20261202)


@pytest.mark.parametrize("axis", ["z", "y"])
def test_edge_markers_match_the_oracle(axis):
    c = _case()
    dec = decomposition(c)
    o = oracle_for(c, dec)
    g = gpu_for(c, dec, axis=axis)
    lmcc = np.array(POINTS, dtype=np.float64)
    rng = np.random.default_rng(11)
    lmvel = rng.uniform(-0.2, 0.2, lmcc.shape)
    lmA = np.full(len(POINTS), 0.05)
    dt = c.params.dt0
    for ns in range(3):
        o.set_markers(lmcc, lmvel, lmA)
        g.set_markers(lmcc, lmvel, lmA)
        rc, dto, cfo = o.fluid_tsch(ns, dt, dt)
        assert rc == 0
        dtg, cfg = g.fluid_tsch(ns, dt, dt)
        assert abs(dtg - dto) <= 1e-12 * dto
        assert abs(cfg - cfo) <= 1e-12 * abs(cfo)
        if ns == 0:
            np.testing.assert_array_equal(g.get("lmind"), o.get("lmind"))
            np.testing.assert_array_equal(g.get("eulins"), o.get("eulins"))
            assert relmax(g.get("phi"), o.get("phi")) <= 1e-14
            assert relmax(g.get("cl"), o.get("cl")) <= 1e-14
        dt = dto
    so, sg = {k: o.get(k) for k in ("q2", "q3", "pr")}, g.get_state()
    for f in ("q2", "q3", "pr"):
        assert np.all(np.isfinite(sg[f])), f
        e = relmax(sg[f], so[f])
        print(f"edge markers {axis} {f}: {e:.2e}", flush=True)
        assert e <= 1e-10, f


def _body_case(verts, center, motion):
    c = cases.make_case("cyl384")
    return cases.Case(c.name, c.params, c.rc, c.zz, verts, center, motion, c.seed)


@pytest.mark.parametrize("which", ["many", "large"])
def test_contribution_order_fallbacks_match_the_oracle(which):
    """The two fallbacks of the (cell, marker) ordering of the forcing
    contributions: more than 16384 keys (2400 markers, nine supports each:
    LDS bitonic blocks and global merge passes) and a body whose support box
    exceeds 16384 cells (a 7 x 2.5 rectangle, about 101 x 281 cells: LDS
    radix sort instead of the cell counting sort).  cyl384 grid, three steps.
    A lost or repeated contribution would show as an O(1) error in the
    forcing.  Bars: indices bit-exact, weights 1e-14, forcing of the last
    substep 1e-11, fields 1e-10."""
    if which == "many":
        c = _body_case(cases.cylinder_vertices(15.0, 10.0, 0.5, 2400), (15.0, 10.0),
                       cases.make_case("cyl384").motion)
    else:
        c = _body_case(cases.rectangle_vertices(15.0, 12.0, 7.0, 2.5, 600), (15.0, 12.0), None)
    dec = decomposition(c)
    o = oracle_for(c, dec)
    g = gpu_for(c, dec)
    t, dt = 0.0, c.params.dt0
    for ns in range(3):
        lm = c.markers(t)
        o.set_markers(*lm)
        g.set_markers(*lm)
        rc, dto, _ = o.fluid_tsch(ns, dt, dt)
        assert rc == 0
        dtg, _ = g.fluid_tsch(ns, dt, dt)
        assert abs(dtg - dto) <= 1e-12 * dto
        if ns == 0:
            np.testing.assert_array_equal(g.get("lmind"), o.get("lmind"))
            np.testing.assert_array_equal(g.get("eulins"), o.get("eulins"))
            assert relmax(g.get("phi"), o.get("phi")) <= 1e-14
        # the forcing of the last substep: its inputs carry the solvers' rounding
        assert relmax(g.get("fey"), o.get("fey")) <= 1e-11
        assert relmax(g.get("fez"), o.get("fez")) <= 1e-11
        t += dto
        dt = dto
    so, sg = {k: o.get(k) for k in ("q2", "q3", "pr")}, g.get_state()
    for f in ("q2", "q3", "pr"):
        e = relmax(sg[f], so[f])
        print(f"{which} {f}: {e:.2e}", flush=True)
        assert e <= 1e-10, f
