"""Shared set-up for the parity tests: one case, the oracle and the GPU step
fed with identical inputs."""
import os

import numpy as np

from fribois2d_amd import cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def case(name):
    if name in ("osccyl", "gallrect"):
        return cases.reference_example(os.path.join(GOLDEN, name), name)
    return cases.make_case(name)


def decomposition(c, method="sym"):
    return cases.phini(c.zz, c.params.alx3, method)


def oracle_for(c, dec=None, state=None, params=None):
    from oracle import Oracle
    p = params or c.params
    zmm, zmmt, wr = dec or decomposition(c)
    o = Oracle(p, c.rc, c.zz, zmm, zmmt, wr)
    o.set_state(state if state is not None else cases.initial_state(p, c.seed))
    return o


def gpu_for(c, dec=None, state=None, params=None, axis=None):
    """GPU step for case c; axis "z" / "y" selects the Poisson transform
    (default: the library's choice, y when it is the shorter axis)."""
    from fribois2d_amd.gpu import FluidStep
    p = params or c.params
    zmm, zmmt, wr = dec or decomposition(c)
    g = FluidStep(p, c.rc, c.zz, zmm, zmmt, wr)
    if axis is not None:
        g.set_transform(axis)
    g.set_state(state if state is not None else cases.initial_state(p, c.seed))
    return g


def relmax(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def poisson_kappa(c):
    """Condition-number estimate of the discrete pressure Poisson operator:
    largest eigenvalue 4/hy^2 + 4/hz^2 (smallest spacings) over the smallest
    non-zero one, pi^2/L^2 of the longer side."""
    hy, hz = np.diff(c.rc).min(), np.diff(c.zz).min()
    ly, lz = c.rc[-1] - c.rc[0], c.zz[-1] - c.zz[0]
    return (4.0 / hy**2 + 4.0 / hz**2) / (np.pi**2 / max(ly, lz) ** 2)


def divergence(st, c):
    """max |div u| over cells (divergence.F90:28-39 without the 1/(dt al))."""
    n2m, n3m = c.params.n2 - 1, c.params.n3 - 1
    rc, zz = c.rc, c.zz
    dx2 = n2m / c.params.rext2
    dx3 = n3m / c.params.alx3
    udx2m = dx2 / ((rc[1:] - rc[:-1]) * dx2)
    udx3m = dx3 / ((zz[1:] - zz[:-1]) * dx3)
    q2, q3 = st["q2"], st["q3"]
    d = (q2[1:n2m + 1, 3:n3m + 3] - q2[0:n2m, 3:n3m + 3]) * udx2m[:, None] + \
        (q3[0:n2m, 4:n3m + 4] - q3[0:n2m, 3:n3m + 3]) * udx3m[None, :]
    return float(np.max(np.abs(d)))


def run_steps(solver, c, nsteps, dt=None, t0=0.0, ns0=0, hist=None):
    """Advance with prescribed marker motion; returns (t, dt)."""
    t = t0
    dt = c.params.dt0 if dt is None else dt
    dto = dt
    for ns in range(ns0, ns0 + nsteps):
        if c.params.ibact and c.verts is not None:
            solver.set_markers(*c.markers(t))
        out = solver.fluid_tsch(ns, dt, dto)
        if isinstance(out, tuple) and len(out) == 3:
            rc, dtn, cf = out
            assert rc == 0, f"oracle error {rc} at step {ns}"
        else:
            dtn, cf = out
        if hist is not None:
            hist.append((ns, dtn, cf))
        t += dtn
        dto = dtn
        dt = dtn
    return t, dt


class OracleFluid:
    """The oracle behind the GPU step's method signatures (errors raise),
    so the coupling drivers run unchanged on either."""

    def __init__(self, o):
        self.o = o

    def set_markers(self, *a):
        self.o.set_markers(*a)

    def fluid_tsch(self, ns, dt, dt_o=None):
        rc, dtn, cflm = self.o.fluid_tsch(ns, dt, dt_o)
        if rc:
            raise RuntimeError(f"oracle step failed ({rc})")
        return dtn, cflm

    def loads_transfer(self, kprobe, lmnor, lmacc):
        rc, press, tau = self.o.loads_transfer(kprobe, lmnor, lmacc)
        if rc:
            raise RuntimeError(f"oracle loads failed ({rc})")
        return press, tau

    def save_state(self):
        self.o.save_state()

    def restore_state(self):
        self.o.restore_state()


def offset_grid_case():
    """tiny with every node but the first moved 2^30 up in y: the MLS moment
    matrices of the markers near y = 2^30 lose all their digits, so the
    partition of unity fails."""
    c = case("tiny")
    off = 2.0 ** 30
    c.rc = np.concatenate([[0.0], off + np.linspace(0.0, 6.0, c.params.n2 - 1)])
    p = cases.FluidParams(**{**c.params.__dict__, "rext2": float(c.rc[-1])})
    lmcc, lmvel, lmA = c.markers(0.0)
    lmcc = lmcc.copy()
    lmcc[:, 0] += off
    return c, p, (lmcc, lmvel, lmA)
