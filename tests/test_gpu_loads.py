"""GPU loads_transfer_c / checkbcs against the oracle and the closed form.

Bars: the same state on both sides gives the loads to 1e-9 of their maximum
after steps of the fluid (the fields agree to ~1e-12; the loads differentiate
them once and twice on cells of size ~1/40); on the linear-field closed form
(test_oracle_loads) to 1e-11 absolute.
"""
import numpy as np
import pytest

from fribois2d_amd import cases
from helpers import case, decomposition, gpu_for, oracle_for, relmax, run_steps
from test_oracle_loads import A, B, P, linear_fields, uniform_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,nsteps", [("cyl64", 5), ("cyl384", 2)])
def test_loads_match_the_oracle(name, nsteps):
    c = case(name)
    dec = decomposition(c)
    o = oracle_for(c, dec)
    g = gpu_for(c, dec)
    ho, hg = [], []
    run_steps(o, c, nsteps, hist=ho)
    run_steps(g, c, nsteps, hist=hg)
    t = sum(h[1] for h in ho)
    nor, acc = c.surface(t)
    rc, po, to = o.loads_transfer(1.0, nor, acc)
    assert rc == 0
    pg, tg = g.loads_transfer(1.0, nor, acc)
    ep, et = relmax(pg, po), relmax(tg, to)
    print(f"{name}: press {ep:.2e}, tau {et:.2e} (max |press| {np.abs(po).max():.3g})", flush=True)
    assert ep <= 1e-9 and et <= 1e-9
    bvo, bwo, nvo, nwo = o.checkbcs()
    bvg, bwg, nvg, nwg = g.checkbcs()
    assert np.max(np.abs(bvg - bvo)) <= 1e-10 * max(np.abs(bvo).max(), 1e-12)
    assert np.max(np.abs(bwg - bwo)) <= 1e-10 * max(np.abs(bwo).max(), 1e-12)
    assert abs(nvg - nvo) <= 1e-10 * nvo and abs(nwg - nwo) <= 1e-10 * nwo


def test_loads_closed_form_on_the_gpu():
    c = uniform_case()
    g = gpu_for(c)
    lm = c.markers(0.0)
    g.set_markers(*lm)
    g.fluid_tsch(0, 1e-3)
    q2, q3, pr = linear_fields(c)
    st = g.get_state()
    st.update(q2=q2, q3=q3, pr=pr)
    g.set_state(st)
    nor, acc = c.surface(0.0)
    kprobe = 1.3
    press, tau = g.loads_transfer(kprobe, nor, acc)
    h = min(np.diff(c.rc).min(), np.diff(c.zz).min())
    hpr = h * kprobe * kprobe
    prb = lm[0] + nor * hpr
    pp = P[0] + P[1] * prb[:, 0] + P[2] * prb[:, 1]
    dpdnw = -(acc[:, 0] * nor[:, 0] + acc[:, 1] * nor[:, 1])
    dpdne = P[1] * nor[:, 0] + P[2] * nor[:, 1]
    np.testing.assert_allclose(press, pp - 0.5 * (dpdnw + dpdne) * hpr, rtol=0, atol=1e-11)
    uy = A[0] + A[1] * prb[:, 0] + A[2] * prb[:, 1]
    uz = B[0] + B[1] * prb[:, 0] + B[2] * prb[:, 1]
    vn = uy * nor[:, 0] + uz * nor[:, 1]
    ty, tz = uy - vn * nor[:, 0], uz - vn * nor[:, 1]
    tn = np.hypot(ty, tz)
    ty, tz = ty / tn, tz / tn
    ducdn = A[1] * ty * nor[:, 0] + A[2] * ty * nor[:, 1] + B[1] * tz * nor[:, 0] + B[2] * tz * nor[:, 1]
    np.testing.assert_allclose(tau[:, 0], c.params.visc * ducdn * ty, rtol=0, atol=1e-12)
    np.testing.assert_allclose(tau[:, 1], c.params.visc * ducdn * tz, rtol=0, atol=1e-12)


def test_tangent_carries_over_at_rest():
    """Zero velocity: no probe has a tangential part, so every marker keeps
    the initial zero tangent and the friction stress vanishes
    (loads_transfer.F90:47,94-96)."""
    c = uniform_case()
    g = gpu_for(c)
    lm = c.markers(0.0)
    g.set_markers(*lm)
    g.fluid_tsch(0, 1e-3)
    st = g.get_state()
    for k in ("q2", "q3"):
        st[k] = np.zeros_like(st[k])
    g.set_state(st)
    nor, acc = c.surface(0.0)
    press, tau = g.loads_transfer(1.0, nor, acc)
    assert np.all(tau == 0.0)
    assert np.all(np.isfinite(press))


@pytest.mark.parametrize("name", ["cyl64", "cyl384"])
def test_diagnostics_match_the_oracle(name):
    """ttime's diagnostics: extrema exact, tree sums to 1e-13."""
    c = case(name)
    dec = decomposition(c)
    o = oracle_for(c, dec)
    g = gpu_for(c, dec)
    st = cases.initial_state(c.params, c.seed)
    # the same state on both sides: the GPU's own state after steps
    run_steps(g, c, 3)
    sg = g.get_state()
    o.set_state({k: sg[k] for k in ("q2", "q3", "pr")})
    do, dg = o.diagnostics(), g.diagnostics()
    for k in ("vmax2", "vmax3", "vmin2", "vmin3", "divmax", "cflm"):
        assert dg[k] == do[k], k
    for k in ("ektot", "q3avg", "vorti"):
        assert abs(dg[k] - do[k]) <= 1e-13 * max(abs(do[k]), 1e-30) + 1e-16, k
    if name == "cyl384":
        assert do["vorti"] != 0.0


def test_strong_coupling_on_the_gpu_matches_the_oracle():
    """The strongly coupled step with the GPU fluid (device rollback between
    sub-iterations) converges like the oracle's and to the same loads."""
    from fribois2d_amd import coupling
    from helpers import OracleFluid
    from test_coupling import cylinder_body
    c = case("cyl64")
    p = cases.FluidParams(**{**c.params.__dict__, "idtv": True})
    dec = decomposition(c)
    g = gpu_for(c, dec, params=p)
    o = OracleFluid(oracle_for(c, dec, params=p))
    bg, bo = cylinder_body(c), cylinder_body(c)
    lg = lo = (np.zeros(2), 0.0)
    dt = 1e-3
    for ns in range(3):
        lg, bg, itg, _, _ = coupling.strong_step(g, bg, lg, ns, dt)
        lo, bo, ito, _, _ = coupling.strong_step(o, bo, lo, ns, dt)
        print(f"step {ns}: sub-iterations gpu {itg} oracle {ito}; lift {lg[0][0]:.6e} / {lo[0][0]:.6e}", flush=True)
        assert abs(itg - ito) <= 1
        assert abs(lg[0][0] - lo[0][0]) <= 1e-8 * max(abs(lo[0][0]), 1e-3)
    assert abs(bg.cm[0] - bo.cm[0]) <= 1e-12
