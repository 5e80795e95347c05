"""The allocation-free loads entry point (fribo_loads_transfer_static)
against fribo_loads_transfer, both called through the C ABI on the same
state: bit-identical press and tau on two calls in a row (the second after
a step with fewer markers), and the refusal above the static capacity."""
import ctypes as C

import numpy as np
import pytest

from fribois2d_amd import cases
from helpers import case, decomposition, gpu_for

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _loads(L, fn, g, kprobe, nor, acc):
    nor = np.asfortranarray(nor, dtype=np.float64)
    acc = np.asfortranarray(acc, dtype=np.float64)
    nlm = nor.shape[0]
    press = np.full(nlm, np.nan)
    tau = np.full((nlm, 2), np.nan, order="F")
    rc = getattr(L, fn)(g.h, C.c_double(kprobe), _dp(nor), _dp(acc), _dp(press), _dp(tau))
    return rc, press, tau


def _subset(lm, sel):
    return tuple(np.asfortranarray(a[sel]) for a in lm)


def test_static_loads_equal_the_allocating_entry_point():
    from fribois2d_amd import gpu
    L = gpu.load()
    c = case("cyl64")
    g = gpu_for(c, decomposition(c))
    t, dt = 0.0, c.params.dt0
    lm = c.markers(t)
    g.set_markers(*lm)
    dt, _ = g.fluid_tsch(0, dt, dt)
    nor, acc = c.surface(dt)
    rc_a, pa, ta = _loads(L, "fribo_loads_transfer", g, 1.0, nor, acc)
    rc_s, ps, ts = _loads(L, "fribo_loads_transfer_static", g, 1.0, nor, acc)
    print(f"first call: {lm[2].size} markers, rc {rc_a} {rc_s}", flush=True)
    assert rc_a == 0 and rc_s == 0
    assert np.all(np.isfinite(pa)) and np.abs(pa).max() > 0.0
    np.testing.assert_array_equal(ps, pa)
    np.testing.assert_array_equal(ts, ta)
    # second call: a step with half the markers, then both entry points again
    sel = np.arange(0, lm[2].size, 2)
    lm2 = _subset(c.markers(dt), sel)
    g.set_markers(*lm2)
    dt2, _ = g.fluid_tsch(1, dt, dt)
    nor2, acc2 = (np.asfortranarray(a[sel]) for a in c.surface(dt + dt2))
    rc_s, ps, ts = _loads(L, "fribo_loads_transfer_static", g, 1.3, nor2, acc2)
    rc_a, pa, ta = _loads(L, "fribo_loads_transfer", g, 1.3, nor2, acc2)
    print(f"second call: {sel.size} markers, rc {rc_a} {rc_s}", flush=True)
    assert rc_a == 0 and rc_s == 0
    np.testing.assert_array_equal(ps, pa)
    np.testing.assert_array_equal(ts, ta)
    g.close()


def test_static_loads_refuse_more_markers_than_the_scratch_holds():
    from fribois2d_amd import gpu
    L = gpu.load()
    c = case("cyl64")
    g = gpu_for(c, decomposition(c))
    n = 8193
    th = 2 * np.pi * np.arange(n) / n
    lmcc = np.asfortranarray(np.stack([3.5 + 0.5 * np.cos(th), 3.0 + 0.5 * np.sin(th)], axis=1))
    g.set_markers(lmcc, np.zeros((n, 2), order="F"), np.full(n, 1e-4))
    nor = np.asfortranarray(np.stack([np.cos(th), np.sin(th)], axis=1))
    rc, _, _ = _loads(L, "fribo_loads_transfer_static", g, 1.0, nor, np.zeros((n, 2), order="F"))
    assert rc == 7
    assert b"at most 8192" in L.fribo_last_error()
    g.close()
