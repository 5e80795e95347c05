"""Build libfribo_gpu.so (gfx950) in-tree with hipcc, and the Fortran binding.

``python -m fribois2d_amd.build`` compiles every HIP source of csrc/ into
fribois2d_amd/libfribo_gpu.so.  Flags:
  --offload-arch=gfx950                  MI355X only
  -ffp-contract=off                      keep the reference's unfused FP64
                                         arithmetic in the stencil/MLS kernels
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                         no packed-FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfribo_gpu.so")
SOURCES = ["fribo_api.hip", "fribo_stencil.hip", "fribo_tridiag.hip", "fribo_mls.hip", "fribo_gemm.hip",
           "fribo_eig.hip", "fribo_ysolve.hip", "fribo_mlp.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def build_lib(verbose=False):
    objs = []
    tmpdir = os.path.join(HERE, "_build")
    os.makedirs(tmpdir, exist_ok=True)
    procs = []
    for s in SOURCES:
        o = os.path.join(tmpdir, s.replace(".hip", ".o"))
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", os.path.join(CSRC, s), "-o", o]
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(o)
    failed = False
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0 or verbose:
            sys.stdout.write(out.decode(errors="replace"))
        if p.returncode != 0:
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    """Fortran interface module + the demo driver that calls the C ABI."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        return None
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    exe = os.path.join(fdir, "fribo_demo")
    cmd = [fc, "-O2", "-module-dir", bdir, os.path.join(fdir, "fribo_gpu_mod.f90"),
           os.path.join(fdir, "fribo_ttime_mod.f90"),
           os.path.join(fdir, "fribo_demo.f90"), "-o", exe + ".tmp", "-L", HERE, "-lfribo_gpu",
           "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stdout.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("amdflang failed")
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build_lib(verbose="-v" in sys.argv))
