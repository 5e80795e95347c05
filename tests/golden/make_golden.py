"""Write cyl384_1000steps.npz: the oracle's trajectory of the 384x288 cylinder
workload (CFL 0.3, heaving body, seeded perturbation), 1000 time steps.

Stored: per-step history (ns, dt, cflm, sum of fey, sum of fez -- the total
immersed-boundary forcing, a lift/drag proxy available inside the fluid step),
the final q2, q3, pr on every 4th node, and the L2 norms of the full fields.
Runs the CPU oracle only (about 3 minutes on one core).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

NSTEPS = 1000


def trajectory(solver, c, nsteps, get):
    hist = np.zeros((nsteps, 5))
    t, dt = 0.0, c.params.dt0
    for ns in range(nsteps):
        solver.set_markers(*c.markers(t))
        out = solver.fluid_tsch(ns, dt, dt)
        dtn, cf = (out[1], out[2]) if len(out) == 3 else out
        if len(out) == 3:
            assert out[0] == 0, out
        hist[ns] = (ns, dtn, cf, float(np.sum(get("fey"))), float(np.sum(get("fez"))))
        t += dtn
        dt = dtn
    return hist


def summary(fields):
    sub = {f"{k}_sub": v[::4, ::4].copy() for k, v in fields.items()}
    norms = {f"{k}_l2": np.sqrt(np.sum(v * v)) for k, v in fields.items()}
    return sub, norms


def main():
    from helpers import case, decomposition, oracle_for
    c = case("cyl384")
    o = oracle_for(c, decomposition(c))
    t0 = time.time()
    hist = trajectory(o, c, NSTEPS, o.get)
    fields = {k: o.get(k) for k in ("q2", "q3", "pr")}
    sub, norms = summary(fields)
    out = os.path.join(HERE, "cyl384_1000steps.npz")
    np.savez_compressed(out, hist=hist, **sub, **norms)
    print(f"wrote {out} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
