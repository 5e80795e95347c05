"""Fluid-structure coupling around the fluid step (CPU, on the oracle).

The rigid-body restatement is pinned by closed forms (free oscillation of the
RK4 spring, constrained degrees of freedom), the strongly coupled step by its
fixed point (the converged loads reproduce themselves) and by the weak
scheme it refines (same trajectory to O(dt) for a heavy body, m* = 10).
"""
import math

import numpy as np
import pytest

from fribois2d_amd import cases, coupling
from helpers import OracleFluid, case, oracle_for


def cylinder_body(c, rbmcon=(0, 1, 1)):
    # m* = 10, Ur = 5 (ky = 12.4025): the cyl384 cylinder of the 1000-step runs
    return coupling.RigidBody(vert=c.verts, cm=(c.center[0], c.center[1], 0.0), rbden=10.0,
                              rba=(0.7853982, 0.0981748), kk=(12.4025, 0.0, 0.0), rbmcon=rbmcon)


def test_rigid_body_free_oscillation():
    """No load: the RK4 spring oscillates at sqrt(k/m) (rigid.F90:165-228);
    locked degrees of freedom stay put."""
    c = case("cyl64")
    b = cylinder_body(c, rbmcon=(0, 1, 1))
    b.cm[0] += 0.1   # released from 0.1 off its rest position
    b.cmposi = b.cmposi.copy()
    dt = 1e-3
    m = b.rba[0] * b.rbden
    w = math.sqrt(b.kk[0] / m)
    zero = (np.zeros(2), 0.0)
    for _ in range(1000):
        b.advance(dt, *zero)
    y = b.cm[0] - b.cmposi[0]
    assert abs(y - 0.1 * math.cos(w * 1.0)) < 1e-9
    assert b.cm[1] == b.kinori[0][1] and b.cm[2] == b.kinori[0][2]
    # the markers moved with the body (released 0.1 from rest: moved y - 0.1)
    np.testing.assert_allclose(b.lmcc[:, 0] - c.markers(0.0)[0][:, 0], y - 0.1, atol=1e-12)


def test_strong_coupling_reaches_its_fixed_point():
    c = case("cyl64")
    p = cases.FluidParams(**{**c.params.__dict__, "idtv": True})
    o = OracleFluid(oracle_for(c, params=p))
    body = cylinder_body(c)
    o.set_markers(*body.markers())
    loads = (np.zeros(2), 0.0)
    dt = 1e-3
    its = []
    for ns in range(3):
        loads, body, it, dt, _ = coupling.strong_step(o, body, loads, ns, dt, tol=1e-8)
        its.append(it)
    # the converged step reproduces its loads: one more fluid step from the
    # rolled-back state with the converged body gives the same loads
    o.restore_state()
    o.set_markers(*body.markers())
    o.fluid_tsch(2, dt, dt)
    press, tau = o.loads_transfer(1.0, body.lmnor, body.lmacc)
    f, m = body.int_stress(press, tau)
    assert np.allclose(f, loads[0], rtol=1e-7, atol=1e-9)
    assert max(its) < 50
    print("sub-iterations per step:", its)


def test_strong_and_weak_coupling_agree_for_a_heavy_body():
    c = case("cyl64")
    p = cases.FluidParams(**{**c.params.__dict__, "idtv": True})
    ow, os_ = OracleFluid(oracle_for(c, params=p)), OracleFluid(oracle_for(c, params=p))
    bw, bs = cylinder_body(c), cylinder_body(c)
    lw = ls = (np.zeros(2), 0.0)
    dt = 1e-3
    for ns in range(4):
        lw, _, _ = coupling.weak_step(ow, bw, lw, ns, dt)
        ls, bs, _, _, _ = coupling.strong_step(os_, bs, ls, ns, dt)
    dy = abs(bw.cm[0] - bs.cm[0])
    print(f"weak vs strong displacement after 4 steps: {dy:.3e}")
    assert dy < 1e-5
