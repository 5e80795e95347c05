"""Oracle run of the 2048x1536 rectangle workload (the benchmarked size): 3
CFL-controlled steps with the pitching body, from the seeded initial state.

Stored: per-step (ns, dt, cflm); after the last step q2, q3, pr and dph on
every 16th node, their L2 norms and sums over the full arrays, and the
marker index arrays lmind, eulins of the last step (bit-exact integers).
CPU oracle only (about 25 s on one core).
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

NSTEPS = 3
FIELDS = ("q2", "q3", "pr")


def summary(get):
    out = {}
    for k in FIELDS:
        v = get(k)
        out[f"{k}_sub"] = v[::16, ::16].copy()
        out[f"{k}_l2"] = np.sqrt(np.sum(v * v))
        out[f"{k}_sum"] = np.sum(v)
    return out


def main():
    from helpers import case, decomposition, oracle_for, run_steps
    c = case("rect2048")
    o = oracle_for(c, decomposition(c))
    t0 = time.time()
    hist = []
    run_steps(o, c, NSTEPS, hist=hist)
    out = summary(o.get)
    out["hist"] = np.array(hist, dtype=np.float64)
    out["lmind"] = o.get("lmind").astype(np.int16)
    out["eulins"] = o.get("eulins").astype(np.int16)
    path = os.path.join(HERE, f"rect2048_{NSTEPS}steps.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} in {time.time() - t0:.0f} s ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
