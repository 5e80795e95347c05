"""The Fortran binding (fribois2d_amd/fortran/fribo_gpu_mod.f90) drives the same
library as the Python one: the Fortran demo driver (main.F90-style loop,
20 CFL-controlled steps of a cylinder) and a Python run on the inputs it
wrote must agree bit for bit, and the CPU oracle to 1e-10; ttime's
diagnostics the driver read through the Fortran shims every step
(fribo_ttime_mod.f90) equal the library's bit for bit and the oracle's to
1e-12 of each value (the rounding-level divergence and vorticity integral
to 1e-11 and 1e-12 absolute)."""
import os
import subprocess

import numpy as np
import pytest

from fribois2d_amd import cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "fribois2d_amd", "fortran", "fribo_demo")


def _read(path):
    raw = open(path, "rb").read()
    n2, n3, nlm, nsteps = np.frombuffer(raw[:16], dtype=np.int32)
    f = np.frombuffer(raw[16:], dtype=np.float64)
    n2m, n3m = n2 - 1, n3 - 1
    out, o = {}, 0
    for name, shape in (("rc", (n2,)), ("zz", (n3,)), ("zmm", (n3m, n3m)), ("zmmt", (n3m, n3m)), ("wr", (n3m,)),
                        ("lmcc", (nlm, 2)), ("lmvel", (nlm, 2)), ("lmA", (nlm,)), ("dt", (nsteps,)),
                        ("q2", (n2, n3m + 6)), ("q3", (n2, n3m + 6)), ("pr", (n2, n3m + 6)),
                        ("diag", (9, nsteps))):
        size = int(np.prod(shape))
        out[name] = f[o:o + size].reshape(shape, order="F")
        o += size
    assert o == f.size
    return int(n2), int(n3), int(nsteps), out


def test_fortran_driver_matches_python_driver(tmp_path):
    if not os.path.exists(DEMO):
        pytest.skip("Fortran demo not built (amdflang missing)")
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.join(ROOT, "fribois2d_amd") + ":" + env.get("LD_LIBRARY_PATH", "")
    print("running the Fortran driver", flush=True)
    r = subprocess.run([DEMO], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    n2, n3, nsteps, d = _read(os.path.join(tmp_path, "fribo_demo_out.bin"))
    p = cases.FluidParams(n2=n2, n3=n3, alx3=8.0, rext2=6.0, visc=0.02)
    st = cases.initial_state(p, None)
    from fribois2d_amd.gpu import FluidStep
    g = FluidStep(p, d["rc"], d["zz"], d["zmm"], d["zmmt"], d["wr"])
    g.set_state(st)
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    o = Oracle(p, d["rc"], d["zz"], d["zmm"], d["zmmt"], d["wr"])
    o.set_state(st)
    dt = dto = 1e-3
    dts = []
    from oracle import DIAG_KEYS
    for ns in range(nsteps):
        g.set_markers(d["lmcc"], d["lmvel"], d["lmA"])
        o.set_markers(d["lmcc"], d["lmvel"], d["lmA"])
        dtg, _ = g.fluid_tsch(ns, dt, dto)
        rc, dtc, _ = o.fluid_tsch(ns, dt, dto)
        assert rc == 0
        dts.append(dtg)
        dt = dto = dtg
        # ttime's diagnostics through the Fortran shims: the library's
        # values bit for bit, the oracle's within 1e-12 of each value (the
        # largest divergence and the vorticity integral are rounding-level
        # quantities here: within 1e-11 and 1e-12 absolute)
        dg = g.diagnostics()
        do = o.diagnostics()
        atol = {"divmax": 1e-11, "vorti": 1e-12}
        for i, k in enumerate(DIAG_KEYS):
            assert d["diag"][i, ns] == dg[k], (ns, k)
            assert abs(d["diag"][i, ns] - do[k]) <= 1e-12 * abs(do[k]) + atol.get(k, 0.0), (ns, k)
    np.testing.assert_array_equal(np.array(dts), d["dt"])
    sg = g.get_state()
    for k in ("q2", "q3", "pr"):
        np.testing.assert_array_equal(sg[k], d[k], err_msg=k)
        ref = o.get(k)
        assert np.max(np.abs(d[k] - ref)) <= 1e-10 * np.max(np.abs(ref)), k
