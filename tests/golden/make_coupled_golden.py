"""Oracle trajectories of spring-mounted bodies, weakly coupled as the
reference's main loop does it (main.F90:79-101: rigid() with the previous
step's loads, fluid_tsch, loads_transfer_c), 1000 steps each:

* ``osccyl_spring_1000steps.npz`` -- the reference's own OscCyl1DOF example
  (tests/golden/osccyl: 811x673 grid, 377 markers, fixed dt = 1e-3, Re = 100;
  rigid_par.in: m* = 10, k = 12.25, y free, initial velocity 0.1);
* ``cyl384_spring_1000steps.npz`` -- the 384x288 cylinder workload on a
  transverse spring, m* = 10, Ur = 5 (k = 12.4025), CFL 0.3.

Stored per step: ns, dt, cflm, displacement y - y0, velocity, lift (y
force), drag (z force) and moment of int_stress (rigid.F90:373-466); at the
end q2, q3, pr on every 4th node and their L2 norms.  CPU oracle only:
about 20 minutes (osccyl) and 4 minutes (cyl384) on one core.

    python tests/golden/make_coupled_golden.py [osccyl_spring|cyl384_spring]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

NSTEPS = 1000
CASES = ("osccyl_spring", "cyl384_spring")
COLS = ("ns", "dt", "cflm", "disp", "vel", "lift", "drag", "moment")


def setup(name):
    """(case, params, body): the fluid case and its rigid body at rest."""
    from fribois2d_amd import coupling
    from helpers import case
    if name == "osccyl_spring":
        c = case("osccyl")
        # rigid_par.in of test_OscCyl1DOF
        body = coupling.RigidBody(vert=c.verts, cm=(30.0010, 15.0010, 0.0), rbden=10.0,
                                  rba=(0.78539816, 0.09817477), kk=(12.25, 0.0, 0.0), rbmcon=(0, 1, 1),
                                  cmv=np.array([0.1, 0.0, 0.0]))
        return c, c.params, body
    if name == "cyl384_spring":
        c = case("cyl384")
        body = coupling.RigidBody(vert=c.verts, cm=(c.center[0], c.center[1], 0.0), rbden=10.0,
                                  rba=(0.7853982, 0.0981748), kk=(12.4025, 0.0, 0.0), rbmcon=(0, 1, 1))
        return c, c.params, body
    raise KeyError(name)


def trajectory(fluid, body, c, nsteps, kprobe=1.0, progress=None):
    """Weakly coupled run; returns the (nsteps, 8) history."""
    from fribois2d_amd import coupling
    hist = np.zeros((nsteps, len(COLS)))
    loads = (np.zeros(2), 0.0)
    y0 = body.cmposi[0]
    dt = c.params.dt0
    for ns in range(nsteps):
        loads, dtn, cflm = coupling.weak_step(fluid, body, loads, ns, dt, kprobe)
        hist[ns] = (ns, dtn, cflm, body.cm[0] - y0, body.cmv[0], loads[0][0], loads[0][1], loads[1])
        dt = dtn
        if progress and (ns + 1) % progress == 0:
            print(f"step {ns + 1}", flush=True)
    return hist


def summary(fields):
    sub = {f"{k}_sub": v[::4, ::4].copy() for k, v in fields.items()}
    norms = {f"{k}_l2": np.sqrt(np.sum(v * v)) for k, v in fields.items()}
    return sub, norms


def main(names):
    from helpers import OracleFluid, decomposition, oracle_for
    for name in names:
        c, p, body = setup(name)
        o = oracle_for(c, decomposition(c), params=p)
        t0 = time.time()
        hist = trajectory(OracleFluid(o), body, c, NSTEPS, progress=100)
        sub, norms = summary({k: o.get(k) for k in ("q2", "q3", "pr")})
        out = os.path.join(HERE, f"{name}_{NSTEPS}steps.npz")
        np.savez_compressed(out, hist=hist, cols=np.array(COLS), **sub, **norms)
        print(f"wrote {out} in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or CASES)
