"""ctypes wrapper of the CPU oracle (libfribo_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Params(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("n2", "n3", "nsst", "idtv", "ibact", "forst", "wfun",
                                        "inslws", "inslwn")] + \
               [(n, C.c_double) for n in ("alx3", "rext2", "visc", "cou", "q3uw", "q3lw", "dtmax",
                                           "cflmax", "cfllim", "radiusSD")]


def lib_path():
    return os.path.join(_HERE, "libfribo_oracle.so")


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(lib_path()):
            raise RuntimeError("oracle library not built (run make -C oracle)")
        L = C.CDLL(lib_path())
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(_Params), dp, dp, dp, dp, dp]
        L.orc_destroy.argtypes = [vp]
        L.orc_size.restype = C.c_long
        L.orc_size.argtypes = [vp, C.c_char_p]
        L.orc_get.argtypes = [vp, C.c_char_p, vp]
        L.orc_set.argtypes = [vp, C.c_char_p, vp]
        L.orc_set_markers.argtypes = [vp, C.c_int, dp, dp, dp]
        L.orc_null_mode.argtypes = [vp]
        L.orc_fluid_tsch.argtypes = [vp, C.c_int, dp, C.c_double, dp]
        L.orc_cfl_dt.argtypes = [vp, C.c_int, dp, C.c_double, dp]
        L.orc_substep.argtypes = [vp, C.c_int, C.c_int]
        L.orc_test_dgtsv.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_test_gesolve.argtypes = [dp, dp, dp]
        L.orc_test_phcalc.argtypes = [vp, dp]
        L.orc_test_weight1d.argtypes = [C.c_double, C.c_int]
        L.orc_test_weight1d.restype = C.c_double
        L.orc_loads_transfer.argtypes = [vp, C.c_double, dp, dp, dp, dp]
        L.orc_checkbcs.argtypes = [vp, dp, dp, dp, dp]
        L.orc_diagnostics.argtypes = [vp, dp]
        L.orc_set_scalars.argtypes = [vp, dp]
        L.orc_test_stage.argtypes = [vp, C.c_char_p]
        _LIB = L
    return _LIB


DIAG_KEYS = ("vmax2", "vmax3", "vmin2", "vmin3", "divmax", "ektot", "vorti", "cflm", "q3avg")

_INT_ARRAYS = {"lmind", "eulins", "jmv", "jpv", "kmv", "kpv"}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One reference-layout fluid state advanced by the CPU restatement."""

    def __init__(self, params, rc, zz, zmm, zmmt, wr):
        L = _lib()
        p = _Params()
        for f, _ in _Params._fields_:
            setattr(p, f, int(getattr(params, f)) if _ in (C.c_int,) else float(getattr(params, f)))
        self.params = params
        self.n2, self.n3 = params.n2, params.n3
        self._keep = [np.ascontiguousarray(rc, dtype=np.float64),
                      np.ascontiguousarray(zz, dtype=np.float64),
                      np.asfortranarray(zmm, dtype=np.float64),
                      np.asfortranarray(zmmt, dtype=np.float64),
                      np.ascontiguousarray(wr, dtype=np.float64)]
        self.h = L.orc_create(C.byref(p), *[_dp(a) for a in self._keep])
        if not self.h:
            raise ValueError("orc_create rejected the parameters")

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_destroy(self.h)
            self.h = None

    def _shape(self, name, n):
        n2, n3m = self.n2, self.n3 - 1
        if n == n2 * (n3m + 6):
            return (n2, n3m + 6)
        if n == n2 * (n3m + 5):
            return (n2, n3m + 5)
        if n == n2 * n3m and name not in ("lmcc",):
            return (n2, n3m)
        return (n,)

    def get(self, name):
        L = _lib()
        n = L.orc_size(self.h, name.encode())
        if n <= 0:
            if name == "scalars":
                out = np.zeros(12)
                L.orc_get(self.h, b"scalars", out.ctypes.data_as(C.c_void_p))
                return out
            raise KeyError(name)
        dt = np.int32 if name in _INT_ARRAYS else np.float64
        out = np.zeros(n, dtype=dt)
        rc = L.orc_get(self.h, name.encode(), out.ctypes.data_as(C.c_void_p))
        if rc:
            raise KeyError(name)
        return out.reshape(self._shape(name, n), order="F")

    def set(self, name, arr):
        L = _lib()
        n = L.orc_size(self.h, name.encode())
        dt = np.int32 if name in _INT_ARRAYS else np.float64
        a = np.asfortranarray(arr, dtype=dt).ravel(order="F")
        if a.size != n:
            raise ValueError(f"{name}: {a.size} values, {n} expected")
        rc = L.orc_set(self.h, name.encode(), a.ctypes.data_as(C.c_void_p))
        if rc:
            raise KeyError(name)

    def set_state(self, st):
        for k, v in st.items():
            self.set(k, v)

    def set_markers(self, lmcc, lmvel, lmA):
        lmcc = np.asfortranarray(lmcc, dtype=np.float64)
        lmvel = np.asfortranarray(lmvel, dtype=np.float64)
        lmA = np.ascontiguousarray(lmA, dtype=np.float64)
        self._mk = (lmcc, lmvel, lmA)
        rc = _lib().orc_set_markers(self.h, lmA.size, _dp(lmcc), _dp(lmvel), _dp(lmA))
        if rc:
            raise ValueError("orc_set_markers")

    @property
    def null_mode(self):
        return _lib().orc_null_mode(self.h)

    def fluid_tsch(self, ns, dt, dt_o=None):
        d = C.c_double(dt)
        cf = C.c_double(0.0)
        rc = _lib().orc_fluid_tsch(self.h, ns, C.byref(d), float(dt if dt_o is None else dt_o), C.byref(cf))
        return rc, d.value, cf.value

    def cfl_dt(self, ns, dt, dt_o=None):
        d = C.c_double(dt)
        cf = C.c_double(0.0)
        rc = _lib().orc_cfl_dt(self.h, ns, C.byref(d), float(dt if dt_o is None else dt_o), C.byref(cf))
        return rc, d.value, cf.value

    def substep(self, n, ns):
        return _lib().orc_substep(self.h, n, ns)

    _SNAP = ("q2", "q3", "pr", "ru2", "ru3", "qb2s", "qb3s", "qb2n", "qb3n", "dq2x2o", "dq3x2o")

    def save_state(self):
        """Rollback copy of the prognostic state (iterated coupling)."""
        self._snap = {k: self.get(k).copy() for k in self._SNAP}

    def restore_state(self):
        for k, v in self._snap.items():
            self.set(k, v)

    def loads_transfer(self, kprobe, lmnor, lmacc):
        """loads_transfer_c after a step: (rc, press (nlm,), tau (nlm, 2))."""
        lmnor = np.asfortranarray(lmnor, dtype=np.float64)
        lmacc = np.asfortranarray(lmacc, dtype=np.float64)
        nlm = lmnor.shape[0]
        press = np.zeros(nlm)
        tau = np.zeros((nlm, 2), order="F")
        rc = _lib().orc_loads_transfer(self.h, float(kprobe), _dp(lmnor), _dp(lmacc), _dp(press), _dp(tau))
        return rc, press, tau

    def set_scalars(self, dt, beta, al, ga, ro):
        """The step scalars a stage reads (test hook)."""
        v = np.array([dt, beta, al, ga, ro], dtype=np.float64)
        _lib().orc_set_scalars(self.h, _dp(v))

    def stage(self, name):
        """Run one named stage of the substep on the current state (test hook)."""
        return _lib().orc_test_stage(self.h, name.encode())

    def diagnostics(self):
        """ttime's per-step diagnostics as a dict."""
        out = np.zeros(9)
        _lib().orc_diagnostics(self.h, _dp(out))
        return dict(zip(DIAG_KEYS, out))

    def checkbcs(self):
        """checkbcs after a step: (bcerrv, bcerrw, nbcv, nbcw)."""
        nlm = self._mk[2].size
        bv, bw = np.zeros(nlm), np.zeros(nlm)
        nv, nw = C.c_double(0.0), C.c_double(0.0)
        _lib().orc_checkbcs(self.h, _dp(bv), _dp(bw), C.byref(nv), C.byref(nw))
        return bv, bw, nv.value, nw.value


def dgtsv(dl, d, du, b):
    """The oracle's DGTTRF+DGTTRS restatement on one system."""
    dl, d, du = (np.ascontiguousarray(a, dtype=np.float64) for a in (dl, d, du))
    x = np.array(b, dtype=np.float64)
    _lib().orc_test_dgtsv(d.size, _dp(dl), _dp(d), _dp(du), _dp(x))
    return x


def gesolve(a, b):
    """The oracle's GEsolve restatement: (x, info)."""
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(3)
    info = _lib().orc_test_gesolve(_dp(a), _dp(b), _dp(x))
    return x, info


def phcalc(orc, rhs):
    """phcalc_outflow of the oracle applied to rhs (n2m, n3m)."""
    f = np.asfortranarray(rhs, dtype=np.float64).copy(order="F")
    _lib().orc_test_phcalc(orc.h, _dp(f))
    return f


def weight1d(rc, wfun):
    """One factor of the MLS weight function at scaled distance rc."""
    return _lib().orc_test_weight1d(float(rc), int(wfun))
