"""Batched steps of uneven instances (fribo_batch_*) against the same
instances stepped one by one (fribo_fluid_tsch).

Bar: bit-identical fields and step scalars when the batch's transform tiles
are the instances' own (every kernel runs the same arithmetic on the same
data; only the launch grouping differs).  With other transform tiles the
transforms themselves are still bit-identical (each output element sums
over k in the same order) but the Poisson gauge constant is summed by tile
column, so it rounds differently; the pressure accumulates that constant's
rounding over the steps (1.6e-14 of its maximum after 3 steps), so the bar
is 1e-14 of the largest value for the velocities and dt, 1e-13 for the
pressure.  The captured (graph) batch equals the eager one bit for bit.
"""
import numpy as np
import pytest

from fribois2d_amd import cases

pytestmark = pytest.mark.gpu

SIZES = [(96, 72, 100.0), (128, 96, 150.0), (160, 120, 60.0), (224, 168, 200.0), (256, 192, 120.0)]
FIELDS = ("q2", "q3", "pr", "ru2", "ru3")


def _members():
    out = []
    for i, (n3m, n2m, re) in enumerate(SIZES):
        out.append(cases.ensemble_case(i, n3m, n2m, re, amp=0.05 + 0.02 * i, freq=0.15 + 0.05 * i))
    return out


def _step(c):
    from fribois2d_amd.gpu import FluidStep
    g = FluidStep(c.params, c.rc, c.zz)   # the library's own phini decomposition
    g.set_state(cases.initial_state(c.params, c.seed))
    g.set_markers(*c.markers(0.0))
    return g


def _run_single(cs, nsteps):
    out = []
    for c in cs:
        g = _step(c)
        t, dt, hist = 0.0, c.params.dt0, []
        for ns in range(nsteps):
            g.set_markers(*c.markers(t))
            dt, cf = g.fluid_tsch(ns, dt, dt)
            hist.append((dt, cf))
            t += dt
        out.append((g.get_state(), np.array(hist)))
        g.close()
    return out


def _run_batch(cs, nsteps, graph=False, shape=None, bad=None, tune=None):
    from fribois2d_amd.gpu import FluidBatch
    gs = [_step(c) for c in cs]
    b = FluidBatch(gs)
    if graph:
        b.set_graph(True)
    if shape is not None:
        b.set_gemm_shape(shape)
    if tune is not None:
        tune.append(b.autotune(3))
    n = len(cs)
    t = np.zeros(n)
    dt = np.array([c.params.dt0 for c in cs])
    hist = [[] for _ in cs]
    errs = []
    for ns in range(nsteps):
        mk = [c.markers(t[i]) for i, c in enumerate(cs)]
        if bad is not None and ns == nsteps - 1:
            lmcc = mk[bad][0].copy()
            lmcc[0, 1] = -1.0
            mk[bad] = (lmcc, mk[bad][1], mk[bad][2])
        b.set_markers(mk)
        dtn, cf, err = b.step(ns, dt, dt)
        errs.append(err)
        for i in range(n):
            if err[i] == 0:
                hist[i].append((dtn[i], cf[i]))
        t += dtn
        dt = dtn
    states = [g.get_state() for g in gs]
    shape_used = b.gemm_shape()
    b.close()
    for g in gs:
        g.close()
    return states, [np.array(h) for h in hist], errs, shape_used


def _compare(single, batch, bar):
    worst = 0.0
    for (st1, h1), st2, h2 in zip(single, batch[0], batch[1]):
        if bar == 0.0:
            np.testing.assert_array_equal(h1[:, 0], h2[:, 0])
        else:
            np.testing.assert_allclose(h2[:, 0], h1[:, 0], rtol=bar, atol=0)
        for k in FIELDS:
            a, b = st1[k], st2[k]
            d = float(np.max(np.abs(a - b)) / max(np.max(np.abs(a)), 1e-300))
            worst = max(worst, d)
            if bar == 0.0:
                assert d == 0.0, (k, d)
            else:
                assert d <= (10 * bar if k == "pr" else bar), (k, d)
    return worst


def test_batch_equals_single_instances():
    cs = _members()
    single = _run_single(cs, 4)
    batch = _run_batch(cs, 4)
    print(f"batch tile shape {batch[3]}", flush=True)
    # every member is small enough to take the 32x32 transform tiles on its own
    assert batch[3] == 0
    _compare(single, batch, 0.0)


def test_autotune_keeps_the_fastest_shape_and_no_state():
    """fribo_batch_autotune times every tile shape the members can run (the
    forward transform alone, and with the chunk pass of the per-mode z-lines
    where the shape does not fuse it) and keeps the fastest by the second
    figure; it changes no state, so the batch then steps like the single
    instances (other-tiles bar)."""
    cs = _members()
    single = _run_single(cs, 3)
    tune = []
    batch = _run_batch(cs, 3, tune=tune)
    best, ms, tot = tune[0]
    avail = tot > 0
    print(f"autotune: shape {best}, transform ms {ms}, with unfused chunk pass {tot}", flush=True)
    assert batch[3] == best and avail[best]
    assert np.array_equal(avail, ms > 0)
    assert tot[best] == tot[avail].min()
    _compare(single, batch, 1e-14)


def test_batch_with_other_transform_tiles():
    cs = _members()
    single = _run_single(cs, 3)
    for shape in (1, 2, 3, 4):
        batch = _run_batch(cs, 3, shape=shape)
        assert batch[3] == shape
        worst = _compare(single, batch, 1e-14)
        print(f"tile shape {shape}: largest relative difference {worst:.2e}", flush=True)


@pytest.mark.parametrize("n3m,n2ms", [(1024, (768, 768)), (256, (192, 96))])
def test_batch_fused_chunk_pass_on_both_direct_tiles(n3m, n2ms):
    """Both direct-to-LDS tile shapes (96x128 of 16 waves, the
    single-instance shape, and 96x64 of 8 waves with 16-deep K-slices) match
    the single-instance runs to the other-tiles bar.  1024 z-rows: chunks of
    16, one per wave group of the tile, so the forward transform's epilogue
    runs the first pass of the per-mode z-lines; 256 z-rows: chunks of 4,
    more than a tile has wave groups, so that pass keeps its own kernel (one
    member with a symmetric y grid, in the mirror form, one without)."""
    cs = [cases.ensemble_case(i, n3m, n2m, re, amp=0.1, freq=0.2) for i, (n2m, re) in enumerate(zip(n2ms, (90.0, 180.0)))]
    single = _run_single(cs, 3)
    for shape in (3, 4):
        batch = _run_batch(cs, 3, shape=shape)
        assert batch[3] == shape
        worst = _compare(single, batch, 1e-14)
        print(f"tile shape {shape}: largest relative difference {worst:.2e}", flush=True)


def test_batch_three_pass_k_lines():
    """Members set to the three-pass k-line solve (fribo_set_ksolve 0) take
    the batched three-pass kernels; the forms are bit-identical, so the batch
    equals the single-instance (one-pass) runs bit for bit."""
    from fribois2d_amd.gpu import FluidBatch
    cs = _members()[:3]
    single = _run_single(cs, 3)
    gs = [_step(c) for c in cs]
    for g in gs:
        g.set_ksolve("3pass")
    b = FluidBatch(gs)
    t = np.zeros(len(cs))
    dt = np.array([c.params.dt0 for c in cs])
    hist = [[] for _ in cs]
    for ns in range(3):
        b.set_markers([c.markers(t[i]) for i, c in enumerate(cs)])
        dtn, cf, err = b.step(ns, dt, dt)
        assert not err.any()
        for i in range(len(cs)):
            hist[i].append((dtn[i], cf[i]))
        t += dtn
        dt = dtn
    states = [g.get_state() for g in gs]
    b.close()
    for g in gs:
        g.close()
    _compare(single, (states, [np.array(h) for h in hist], None, None), 0.0)


def test_ksolve_form_is_fixed_while_in_a_batch():
    """fribo_set_ksolve refuses an instance that is in a batch (the batch
    holds a copy of its settings) and works again once the batch is gone."""
    from fribois2d_amd.gpu import FluidBatch, FriboError
    cs = _members()[:2]
    gs = [_step(c) for c in cs]
    b = FluidBatch(gs)
    with pytest.raises(FriboError) as e:
        gs[0].set_ksolve("3pass")
    assert "ARG" in str(e.value)
    assert gs[1].ymirror   # 128 x 96: symmetric y grid
    with pytest.raises(FriboError) as e:
        gs[1].set_ymirror(False)
    assert "ARG" in str(e.value)
    b.close()
    gs[0].set_ksolve("3pass")
    assert gs[0].ksolve == "3pass"
    for g in gs:
        g.close()


def test_batch_graph_replay_equals_eager():
    cs = _members()[:3]
    eager = _run_batch(cs, 6)
    graph = _run_batch(cs, 6, graph=True)
    for s1, s2 in zip(eager[0], graph[0]):
        for k in FIELDS:
            np.testing.assert_array_equal(s1[k], s2[k])
    for h1, h2 in zip(eager[1], graph[1]):
        np.testing.assert_array_equal(h1, h2)


def test_batch_instance_error_leaves_the_others():
    """One member's marker leaves the domain on the last step: it reports
    FRIBO_ERR_MARKER_OUTSIDE and keeps its state; the others step on."""
    cs = _members()[:3]
    ok = _run_batch(cs, 3)
    bad = _run_batch(cs, 3, bad=1)
    assert list(bad[2][-1]) == [0, 3, 0]
    ref2 = _run_batch(cs, 2)
    for k in FIELDS:
        np.testing.assert_array_equal(bad[0][0][k], ok[0][0][k])
        np.testing.assert_array_equal(bad[0][2][k], ok[0][2][k])
        np.testing.assert_array_equal(bad[0][1][k], ref2[0][1][k])


def test_instances_step_on_alone_after_the_batch():
    """Two batched steps, the batch destroyed, two more steps of each
    instance on its own: the same as four single-instance steps."""
    from fribois2d_amd.gpu import FluidBatch
    cs = _members()[:3]
    single = _run_single(cs, 4)
    gs = [_step(c) for c in cs]
    b = FluidBatch(gs)
    t = np.zeros(len(cs))
    dt = np.array([c.params.dt0 for c in cs])
    for ns in range(2):
        b.set_markers([c.markers(t[i]) for i, c in enumerate(cs)])
        dtn, _, err = b.step(ns, dt, dt)
        assert not err.any()
        t += dtn
        dt = dtn
    b.close()
    for i, (c, g) in enumerate(zip(cs, gs)):
        ti, di = t[i], dt[i]
        for ns in range(2, 4):
            g.set_markers(*c.markers(ti))
            di, _ = g.fluid_tsch(ns, di, di)
            ti += di
        st = g.get_state()
        for k in FIELDS:
            np.testing.assert_array_equal(st[k], single[i][0][k])
        g.close()
