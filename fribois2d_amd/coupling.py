"""Fluid-structure coupling around the fluid step.

The reference couples the rigid body to the flow loosely: every step
(main.F90:79-101) ``rigid()`` advances the body with the loads of the
previous step, then ``fluid_tsch`` advances the flow with the new marker
kinematics and ``loads_transfer_c`` computes the loads for the next step.
``RigidBody`` restates ``rigid.F90`` for one body:

* ``int_stress`` (rigid.F90:373-466): pressure and friction forces of every
  marker, their resultant and the moment about the centre of mass;
* ``kinem_RK4`` + ``modfn`` (rigid.F90:165-228, 341-369): explicit RK4 of
  M x'' + C x' + K x + K3 x^3 = F for the three rigid degrees of freedom,
  acceleration from the velocity difference;
* the ``rbmcon`` constraints (rigid.F90:60-76) and the marker update
  (rigid.F90:78-123): positions rotated about the centre, velocities and
  accelerations from the old lever arm (as the reference does), vertices
  moved and normals rebuilt by ``createnorm``.

``weak_step`` is the reference's staggered step.  ``strong_step`` iterates
the same three stages inside one step until the loads stop changing (the
strong coupling the reference prepares for with its unused _o arrays,
declarations.F90:155-159): the fluid state is rolled back before every
repeat, the load iterate is relaxed with Aitken's dynamic factor
(Kuettler and Wall, Comput. Mech. 43 (2008) 61-72), and the number of
sub-iterations is returned.  Both drivers work with the GPU step
(``fribois2d_amd.gpu.FluidStep``) and with any object offering the same
methods.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np

from .cases import createnorm, markers_from_vertices


@dataclass
class RigidBody:
    """One rigid body: closed polygon, centre of mass and the rbm inputs."""
    vert: np.ndarray                    # (nlm+1, 2) vertices (y, z)
    cm: np.ndarray                      # centre of mass (y, z, angle)
    rbden: float                        # density ratio
    rba: tuple                          # (area, polar moment) (rba(1), rba(2))
    kk: tuple = (0.0, 0.0, 0.0)         # linear stiffness (ky, kz, ka)
    cc: tuple = (0.0, 0.0, 0.0)         # damping
    kk3: tuple = (0.0, 0.0, 0.0)        # cubic stiffness
    rbmcon: tuple = (0, 0, 0)           # 1: degree of freedom locked
    rbgravF: float = 0.0
    cmv: np.ndarray = field(default_factory=lambda: np.zeros(3))
    cma: np.ndarray = field(default_factory=lambda: np.zeros(3))

    def __post_init__(self):
        self.vert = np.array(self.vert, dtype=np.float64)
        self.cm = np.array(self.cm, dtype=np.float64)
        self.cmposi = self.cm.copy()                 # spring rest position
        self.kinori = np.stack([self.cm.copy(), self.cmv.copy(), self.cma.copy()])
        self.lmcc, self.lmA = markers_from_vertices(self.vert)
        self.lmvel = np.zeros_like(self.lmcc)
        self.lmacc = np.zeros_like(self.lmcc)
        self.lmnor = createnorm(self.vert, -1.0)

    def copy(self):
        return copy.deepcopy(self)

    @property
    def nlm(self):
        return self.lmA.size

    def markers(self):
        return self.lmcc, self.lmvel, self.lmA

    # rigid.F90:373-466
    def int_stress(self, press, tau):
        fp = -press[:, None] * self.lmnor * self.lmA[:, None]
        fv = tau * self.lmA[:, None]
        ftot = fp + fv
        # the marker sums in marker order (cumulative sums run left to
        # right, like the reference's loops)
        fyp, fzp = np.cumsum(fp[:, 0])[-1], np.cumsum(fp[:, 1])[-1]
        fyv, fzv = np.cumsum(fv[:, 0])[-1], np.cumsum(fv[:, 1])[-1]
        forg = (1.0 - 1.0 / self.rbden) * self.rba[0] * self.rbgravF
        forcm = np.array([fyp + fyv, fzp + fzv]) + forg
        ry, rz = self.lmcc[:, 0] - self.cm[0], self.lmcc[:, 1] - self.cm[1]
        mom = float(np.cumsum(ry * ftot[:, 1] - rz * ftot[:, 0])[-1])
        return forcm, mom

    # rigid.F90:341-369
    def _modfn(self, x, ff):
        ky, kz, ka = self.kk
        cy, cz, ca = self.cc
        k3 = self.kk3
        mi = np.array([1.0, 1.0, 1.0, 1.0 / (self.rba[0] * self.rbden), 1.0 / (self.rba[0] * self.rbden),
                       1.0 / (self.rba[1] * self.rbden)])
        rhs = np.array([x[3], x[4], x[5],
                        -ky * x[0] - cy * x[3] - k3[0] * x[0] ** 3,
                        -kz * x[1] - cz * x[4] - k3[1] * x[1] ** 3,
                        -ka * x[2] - ca * x[5] - k3[2] * x[2] ** 3]) + ff
        return mi * rhs

    def advance(self, dt, forcm, momcm):
        """rigid(): RK4 with the given loads, constraints, marker update."""
        ff = np.array([0.0, 0.0, 0.0, forcm[0], forcm[1], momcm])
        xo = np.array([self.cm[0] - self.cmposi[0], self.cm[1] - self.cmposi[1], self.cm[2] - self.cmposi[2],
                       self.cmv[0], self.cmv[1], self.cmv[2]])
        k1 = self._modfn(xo, ff)
        k2 = self._modfn(xo + k1 * dt / 2.0, ff)
        k3 = self._modfn(xo + k2 * dt / 2.0, ff)
        k4 = self._modfn(xo + k3 * dt, ff)
        xn = xo + dt * (k1 / 6.0 + k2 / 3.0 + k3 / 3.0 + k4 / 6.0)
        cmp1 = self.cmposi + xn[:3]
        cmv1 = xn[3:].copy()
        cma1 = (cmv1 - self.cmv) / dt
        con = np.stack([cmp1, cmv1, cma1])
        for i in range(3):
            if self.rbmcon[i] == 1:
                con[:, i] = self.kinori[:, i]
        cmp1, cmv1, cma1 = con
        rot = cmp1[2] - self.cm[2]
        c, s = math.cos(rot), math.sin(rot)
        rho = self.lmcc - self.cm[:2]
        pos = cmp1[:2] + np.stack([c * rho[:, 0] - s * rho[:, 1], s * rho[:, 0] + c * rho[:, 1]], axis=1)
        c1 = np.stack([-cmv1[2] * rho[:, 1], cmv1[2] * rho[:, 0]], axis=1)
        vel = cmv1[:2] + c1
        c2 = np.stack([-cma1[2] * rho[:, 1], cma1[2] * rho[:, 0]], axis=1)
        c3 = np.stack([-cmv1[2] * c1[:, 1], cmv1[2] * c1[:, 0]], axis=1)
        acc = cma1[:2] + c2 + c3
        rv = self.vert - self.cm[:2]
        self.vert = cmp1[:2] + np.stack([c * rv[:, 0] - s * rv[:, 1], s * rv[:, 0] + c * rv[:, 1]], axis=1)
        self.lmcc, self.lmvel, self.lmacc = pos, vel, acc
        self.cm, self.cmv, self.cma = cmp1, cmv1, cma1
        self.lmnor = createnorm(self.vert, -1.0)


def weak_step(fluid, body, loads, ns, dt, kprobe=1.0):
    """One step of the reference's staggered scheme (main.F90:79-101):
    rigid() with the previous loads, the fluid step, the new loads.
    ``loads`` is (forcm, momcm) of the previous step; returns the new
    (forcm, momcm), dt and cflm."""
    body.advance(dt, *loads)
    fluid.set_markers(*body.markers())
    dtn, cflm = fluid.fluid_tsch(ns, dt, dt)
    press, tau = fluid.loads_transfer(kprobe, body.lmnor, body.lmacc)
    return body.int_stress(press, tau), dtn, cflm


OMEGA_MIN, OMEGA_MAX = 0.01, 2.0


def strong_step(fluid, body, guess, ns, dt, kprobe=1.0, tol=1e-8, maxit=50, omega0=0.5, state=None,
                strict=True):
    """One strongly coupled step: fixed-point iteration on the loads
    (forcm, momcm) with Aitken relaxation until the relative load change is
    below tol.  ``guess`` is the first load iterate (e.g. the previous step's
    loads, or a predictor's).  The fluid state is rolled back on the device
    before every repeat.  Returns (loads, body after the step, iterations,
    dt, cflm).  dt must not be changed by the step (fixed dt, idtv).
    The relaxation factor stays within [OMEGA_MIN, OMEGA_MAX] (an Aitken
    update outside it restarts from omega0); with a ``state`` dict the
    factor the previous step converged with starts the next one.
    The loads are a discontinuous function of the body position where a
    marker's support or probe changes cell; an iterate that straddles such a
    switch oscillates between two load levels and cannot converge.  With
    strict=False the step then ends after maxit sub-iterations with the last
    iterate (state["failed"] counts such steps) instead of raising."""
    fluid.save_state()
    body0 = body.copy()
    x = np.array([guess[0][0], guess[0][1], guess[1]], dtype=np.float64)
    r_prev = None
    omega = state.get("omega", omega0) if state is not None else omega0
    for it in range(1, maxit + 1):
        b = body0.copy()
        b.advance(dt, x[:2], x[2])
        if it > 1:
            fluid.restore_state()
        fluid.set_markers(*b.markers())
        dtn, cflm = fluid.fluid_tsch(ns, dt, dt)
        press, tau = fluid.loads_transfer(kprobe, b.lmnor, b.lmacc)
        f, m = b.int_stress(press, tau)
        y = np.array([f[0], f[1], m])
        r = y - x
        if np.linalg.norm(r) <= tol * max(np.linalg.norm(y), 1e-300):
            if state is not None:
                state["omega"] = omega
            return (y[:2], y[2]), b, it, dtn, cflm
        if r_prev is not None:
            d = r - r_prev
            dd = float(d @ d)
            if dd > 0.0:
                omega = -omega * float(r_prev @ d) / dd
            if not OMEGA_MIN <= omega <= OMEGA_MAX:
                omega = omega0
        x = x + omega * r
        r_prev = r
    if not strict:
        if state is not None:
            state["failed"] = state.get("failed", 0) + 1
            state["omega"] = omega0
        return (y[:2], y[2]), b, maxit, dtn, cflm
    raise RuntimeError(f"strong coupling did not converge in {maxit} iterations")
