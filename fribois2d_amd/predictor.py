"""Learned first iterate of the strong fluid-structure coupling loop.

At low mass ratio (m* = 1) the loose coupling of the reference
(main.F90:79-101) is unstable, and the strongly coupled step
(``coupling.strong_step``: fixed point on the loads with Aitken relaxation,
the fluid rolled back on the device between sub-iterations) needs many
sub-iterations.  Its first iterate is free: any guess converges to the same
loads within the tolerance, because the plain iteration decides.  This
module proposes that guess with the library's MLP (``gpu.Mlp``: hand-written
HIP forward, backward and Adam kernels, FP32 training, BF16 inference) from
the last HIST steps of the body's displacement, velocity and lift, trained
on coupling histories the solver itself generates.

Features of step n: (disp, vel, lift, drag, moment) of steps n-HIST .. n-1
relative to step n-1's values (plus step n-1's values themselves), scaled
by the training set's spread; target: the change of the converged loads
(lift, drag, moment) from step n-1 to step n, scaled likewise.
"""
from __future__ import annotations

import math

import numpy as np

from . import cases, coupling

HIST = 8
NI = 5 * HIST + 5
NO = 3


def coupled_case(run: int, n3m: int = 512, n2m: int = 384, re: float = 150.0, dt: float = 0.005, seed: int = 7):
    """A cylinder on a transverse spring at m* = 1 (rbden = 1), Re = re,
    reduced velocity Ur uniform in [4, 8] and an initial transverse velocity
    of 0.05-0.2 drawn from (seed, run); fixed dt (the strongly coupled step
    repeats one dt).  Returns (case, params, body)."""
    rng = np.random.default_rng([seed, run])
    ur = float(rng.uniform(4.0, 8.0))
    v0 = float(rng.uniform(0.05, 0.2)) * (1.0 if rng.random() < 0.5 else -1.0)
    c = cases.ensemble_case(run, n3m, n2m, re)
    p = cases.FluidParams(**{**c.params.__dict__, "idtv": True, "dt0": dt, "cfllim": 2.0})
    rba = (0.7853982, 0.0981748)
    k = 1.0 * rba[0] * (2.0 * math.pi / ur) ** 2
    body = coupling.RigidBody(vert=c.verts, cm=(c.center[0], c.center[1], 0.0), rbden=1.0, rba=rba,
                              kk=(k, 0.0, 0.0), rbmcon=(0, 1, 1), cmv=np.array([v0, 0.0, 0.0]))
    return c, p, body


def features(hist):
    """Input vector of the step after the recorded ones: hist is a list of
    (disp, vel, lift, drag, moment) of the previous steps (at least HIST).
    Drag and moment enter too: the coupling tolerance applies to the whole
    load vector, whose size the drag dominates."""
    h = np.asarray(hist[-HIST:], dtype=np.float64)
    last = h[-1]
    return np.concatenate([(h - last).ravel(), last])


def samples(hist):
    """Training pairs (features, load change) of one run's history."""
    X, T = [], []
    for n in range(HIST, len(hist)):
        X.append(features(hist[:n]))
        T.append(np.asarray(hist[n][2:5]) - np.asarray(hist[n - 1][2:5]))
    return np.array(X).reshape(-1, NI), np.array(T).reshape(-1, NO)


def run_coupled(fluid, body, nsteps, dt, guess=None, tol=1e-8, maxit=40, stats=None):
    """Strongly coupled run; guess(hist, loads) proposes the first iterate of
    a step (None: the previous step's converged loads).  Each step's Aitken
    factor starts from the one the previous step converged with.  Returns
    the history rows (disp, vel, lift, drag, moment) and the sub-iterations
    per step.  A step whose iterates straddle a switch of the marker
    supports (see coupling.strong_step) ends after maxit sub-iterations;
    stats (a dict) gets the count of such steps under "failed"."""
    y0 = body.cmposi[0]
    loads = (np.zeros(2), 0.0)
    hist, its = [], []
    state = {}
    for ns in range(nsteps):
        g = guess(hist, loads) if guess is not None else loads
        loads, body, it, _, _ = coupling.strong_step(fluid, body, g, ns, dt, tol=tol, maxit=maxit, state=state,
                                                     strict=False)
        hist.append((body.cm[0] - y0, body.cmv[0], loads[0][0], loads[0][1], loads[1]))
        its.append(it)
    if stats is not None:
        stats["failed"] = stats.get("failed", 0) + state.get("failed", 0)
    return hist, np.array(its)


class Predictor:
    """The trained network with the scalings of its training set."""

    def __init__(self, net, xm, xs, ts):
        self.net, self.xm, self.xs, self.ts = net, xm, xs, ts

    @classmethod
    def train(cls, X, T, epochs=300, batch=256, lr=2e-3, seed=1):
        from .gpu import Mlp
        xm, xs = X.mean(0), X.std(0) + 1e-30
        ts = T.std(0) + 1e-30
        Xn = ((X - xm) / xs).astype(np.float32)
        Tn = (T / ts).astype(np.float32)
        net = Mlp(NI, NO, seed=seed)
        rng = np.random.default_rng(seed)
        loss = None
        for e in range(epochs):
            perm = rng.permutation(Xn.shape[0])
            loss = net.train(Xn[perm], Tn[perm], 1, batch, lr if e < epochs // 2 else lr / 4)
        p = cls(net, xm, xs, ts)
        p.loss = loss
        return p

    def __call__(self, hist, loads):
        """First iterate of the next step: the previous loads plus the
        predicted change (the previous loads alone until HIST steps exist)."""
        if len(hist) < HIST:
            return loads
        x = ((features(hist) - self.xm) / self.xs).astype(np.float32)
        d = self.net.predict(x)[0].astype(np.float64) * self.ts
        return (np.asarray(loads[0]) + d[:2], float(loads[1] + d[2]))


def extrapolate(hist, loads):
    """Linear extrapolation of the last two converged loads (a plain
    non-learned guess, for comparison)."""
    if len(hist) < 2:
        return loads
    a, b = np.asarray(hist[-1][2:5]), np.asarray(hist[-2][2:5])
    g = 2.0 * a - b
    return (g[:2], float(g[2]))
