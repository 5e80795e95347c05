"""The C ABI library loads and exports every entry point include/fribo_gpu.h
declares (CPU only: no kernel runs), and host-side argument checks reject bad
calls before touching the device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fribo_gpu.h")
FMOD = os.path.join(ROOT, "fribois2d_amd", "fortran", "fribo_gpu_mod.f90")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|const char\*)\s+(fribo_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_step_interface():
    names = declared()
    for n in ("fribo_init", "fribo_fluid_tsch", "fribo_set_state", "fribo_get_state", "fribo_set_markers",
              "fribo_finalize", "fribo_get_array"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from fribois2d_amd import gpu
    lib = ctypes.CDLL(gpu.LIB_PATH)
    for n in declared():
        assert hasattr(lib, n), n
    assert set(declared()) <= set(gpu.EXPORTS)


def test_fortran_module_binds_the_same_names():
    txt = open(FMOD).read()
    bound = set(re.findall(r'bind\(c,\s*name="(fribo_\w+)"\)', txt))
    for n in ("fribo_init", "fribo_finalize", "fribo_set_state", "fribo_get_state", "fribo_set_markers",
              "fribo_fluid_tsch", "fribo_cfl_dt", "fribo_substep", "fribo_get_array", "fribo_array_size",
              "fribo_sync", "fribo_last_error"):
        assert n in bound, n
    assert bound <= set(declared())


def test_argument_errors_are_reported_without_a_device():
    from fribois2d_amd import gpu
    L = gpu.load()
    h = ctypes.c_void_p()
    # null parameter block: rejected before any HIP call
    rc = L.fribo_init(None, None, None, None, None, 0, None, ctypes.byref(h))
    assert rc == 5
    assert b"null" in L.fribo_last_error()
    assert L.fribo_fluid_tsch(None, 0, None, 0.0, None) == 5
    assert L.fribo_set_markers(None, -1, None, None, None) == 5
    assert L.fribo_array_size(None, b"q2") == 0
    # phini on either side: too few nodes / missing outputs
    zz = (ctypes.c_double * 2)(0.0, 1.0)
    for fn in (L.fribo_phini, L.fribo_phini_device):
        assert fn(2, zz, 1.0, 1, None, None, None) == 5


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
