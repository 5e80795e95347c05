"""ctypes binding of libfribo_gpu.so -- the fluid step as the reference drives it.

``FluidStep`` mirrors the reference's calling pattern (main.F90:79-101):
upload the module state once, then per time step set the marker kinematics
produced by rigid() and call ``fluid_tsch(ns, dt, dt_o)``; errors that the
reference sends to stopcase (fluid_tsch.F90:43-54, lgrdistribution.F90:32-36,
MLSint.F90:157-163) raise ``FriboError`` with the same code the C ABI returns.

The library must be built (``python -m fribois2d_amd.build``); there is no
fallback path: importing this module without it raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfribo_gpu.so")

ERR_NAMES = {1: "DT_TOO_SMALL", 2: "CFL_TOO_LARGE", 3: "MARKER_OUTSIDE", 4: "PARTITION_UNITY",
             5: "ARG", 6: "HIP", 7: "LIMIT"}

EXPORTS = ["fribo_init", "fribo_finalize", "fribo_set_state", "fribo_get_state", "fribo_set_markers",
           "fribo_fluid_tsch", "fribo_cfl_dt", "fribo_substep", "fribo_substep_folded", "fribo_array_size",
           "fribo_get_array", "fribo_sync", "fribo_last_error", "fribo_test_gemm", "fribo_test_gemm_nn", "fribo_set_profiling", "fribo_get_profile",
           "fribo_reset_profile", "fribo_profile_name", "fribo_set_graph", "fribo_test_copy",
           "fribo_set_transform", "fribo_get_transform", "fribo_set_ksolve", "fribo_get_ksolve",
           "fribo_set_ymirror", "fribo_get_ymirror", "fribo_test_ytransform", "fribo_test_ymirror",
           "fribo_loads_transfer",
           "fribo_checkbcs", "fribo_diagnostics", "fribo_save_state", "fribo_restore_state",
           "fribo_phini", "fribo_phini_device", "fribo_loads_transfer_static",
           "fribo_batch_create", "fribo_batch_destroy", "fribo_batch_set_markers", "fribo_batch_step",
           "fribo_batch_set_graph", "fribo_batch_set_gemm_shape", "fribo_batch_get_gemm_shape",
           "fribo_batch_autotune", "fribo_set_precision", "fribo_get_precision", "fribo_test_gemm32",
           "fribo_mlp_create", "fribo_mlp_destroy", "fribo_mlp_size", "fribo_mlp_get_params", "fribo_mlp_grad",
           "fribo_mlp_train", "fribo_mlp_predict", "fribo_fetch_fields"]

NPROF = 12

DIAG_KEYS = ("vmax2", "vmax3", "vmin2", "vmin3", "divmax", "ektot", "vorti", "cflm", "q3avg")


class FriboError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fribo error {code} ({ERR_NAMES.get(code, '?')}): {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("n2", "n3", "nsst", "idtv", "ibact", "forst", "wfun",
                                          "inslws", "inslwn")] + \
               [(n, C.c_double) for n in ("alx3", "rext2", "visc", "cou", "q3uw", "q3lw", "dtmax",
                                           "cflmax", "cfllim", "radiusSD")]

    @classmethod
    def from_fluid(cls, fp):
        p = cls()
        for f, t in cls._fields_:
            v = getattr(fp, f)
            setattr(p, f, int(v) if t is C.c_int32 else float(v))
        return p


_lib = None


def load():
    """Load libfribo_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: build it with `python -m fribois2d_amd.build`")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.fribo_init.argtypes = [C.POINTER(Params), dp, dp, dp, dp, C.c_int, dp, C.POINTER(vp)]
    L.fribo_finalize.argtypes = [vp]
    L.fribo_set_state.argtypes = [vp] + [dp] * 11
    L.fribo_get_state.argtypes = [vp] + [dp] * 11
    L.fribo_set_markers.argtypes = [vp, C.c_int32, dp, dp, dp]
    L.fribo_fluid_tsch.argtypes = [vp, C.c_int32, dp, C.c_double, dp]
    L.fribo_cfl_dt.argtypes = [vp, C.c_int32, dp, C.c_double, dp]
    L.fribo_substep.argtypes = [vp, C.c_int32, C.c_int32]
    L.fribo_substep_folded.argtypes = [vp, C.c_int32, C.c_int32]
    L.fribo_array_size.argtypes = [vp, C.c_char_p]
    L.fribo_array_size.restype = C.c_long
    L.fribo_get_array.argtypes = [vp, C.c_char_p, vp, C.c_long]
    L.fribo_sync.argtypes = [vp]
    L.fribo_last_error.restype = C.c_char_p
    L.fribo_test_gemm.argtypes = [C.c_int32, C.c_int32, C.c_int32, dp, C.c_int32, dp, C.c_int32, dp, C.c_int32]
    L.fribo_test_gemm_nn.argtypes = [C.c_int32, C.c_int32, C.c_int32, dp, C.c_int32, dp, C.c_int32, dp, C.c_int32]
    L.fribo_set_profiling.argtypes = [vp, C.c_int32]
    L.fribo_get_profile.argtypes = [vp, dp, C.POINTER(C.c_int32), C.c_int32]
    L.fribo_reset_profile.argtypes = [vp]
    L.fribo_profile_name.argtypes = [C.c_int32]
    L.fribo_profile_name.restype = C.c_char_p
    L.fribo_set_graph.argtypes = [vp, C.c_int32]
    L.fribo_test_copy.argtypes = [C.c_int64, C.c_int32, dp]
    L.fribo_set_transform.argtypes = [vp, C.c_int32]
    L.fribo_get_transform.argtypes = [vp]
    L.fribo_set_ksolve.argtypes = [vp, C.c_int32]
    L.fribo_get_ksolve.argtypes = [vp]
    L.fribo_test_ytransform.argtypes = [C.c_int32, dp, dp, dp, dp, dp, dp]
    L.fribo_set_ymirror.argtypes = [vp, C.c_int32]
    L.fribo_get_ymirror.argtypes = [vp]
    L.fribo_test_ymirror.argtypes = [C.c_int32, dp, dp, dp, C.c_double, dp, dp, dp]
    L.fribo_loads_transfer.argtypes = [vp, C.c_double, dp, dp, dp, dp]
    L.fribo_loads_transfer_static.argtypes = [vp, C.c_double, dp, dp, dp, dp]
    L.fribo_checkbcs.argtypes = [vp, dp, dp, dp, dp]
    L.fribo_diagnostics.argtypes = [vp, dp]
    L.fribo_save_state.argtypes = [vp]
    L.fribo_phini.argtypes = [C.c_int32, dp, C.c_double, C.c_int32, dp, dp, dp]
    L.fribo_phini_device.argtypes = [C.c_int32, dp, C.c_double, C.c_int32, dp, dp, dp]
    L.fribo_restore_state.argtypes = [vp]
    L.fribo_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
    L.fribo_batch_destroy.argtypes = [vp]
    L.fribo_batch_set_markers.argtypes = [vp, dp]
    L.fribo_batch_step.argtypes = [vp, C.c_int32, dp, dp, dp, C.POINTER(C.c_int32)]
    L.fribo_batch_set_graph.argtypes = [vp, C.c_int32]
    L.fribo_batch_set_gemm_shape.argtypes = [vp, C.c_int32]
    L.fribo_batch_get_gemm_shape.argtypes = [vp]
    L.fribo_batch_autotune.argtypes = [vp, C.c_int32, dp, C.POINTER(C.c_int32)]
    L.fribo_fetch_fields.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, dp, dp, dp, dp, dp]
    L.fribo_set_precision.argtypes = [vp, C.c_int32]
    L.fribo_get_precision.argtypes = [vp]
    fp = C.POINTER(C.c_float)
    L.fribo_mlp_create.argtypes = [C.c_int32, C.c_int32, fp, C.c_uint64, C.POINTER(vp)]
    L.fribo_mlp_destroy.argtypes = [vp]
    L.fribo_mlp_size.argtypes = [vp]
    L.fribo_mlp_get_params.argtypes = [vp, fp]
    L.fribo_mlp_grad.argtypes = [vp, C.c_int32, fp, fp, fp, fp]
    L.fribo_mlp_train.argtypes = [vp, C.c_int32, fp, fp, C.c_int32, C.c_int32, C.c_float, fp]
    L.fribo_mlp_predict.argtypes = [vp, C.c_int32, fp, fp]
    L.fribo_test_gemm32.argtypes = [C.c_int32, C.c_int32, C.c_int32, dp, C.c_int32, dp, C.c_int32, dp, C.c_int32]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise FriboError(rc, load().fribo_last_error().decode())


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


_INT = {"lmind", "eulins"}


class FluidStep:
    """One fluid instance on the GPU (reference-layout host interface)."""

    def __init__(self, params, rc, zz, zmm=None, zmmt=None, wr=None):
        L = load()
        self.params = params
        self.n2, self.n3 = params.n2, params.n3
        p = Params.from_fluid(params)
        h = C.c_void_p()
        if zmm is None:
            # the library computes phini's decomposition itself
            zz = np.ascontiguousarray(zz, dtype=np.float64)
            rc = np.ascontiguousarray(rc, dtype=np.float64)
            _check(L.fribo_init(C.byref(p), _dp(rc), _dp(zz), None, None, 0, None, C.byref(h)))
            self.h = h
            self.nlm = 0
            return
        zmm = np.asfortranarray(zmm, dtype=np.float64)
        zmmt = np.asfortranarray(zmmt, dtype=np.float64)
        rc = np.ascontiguousarray(rc, dtype=np.float64)
        zz = np.ascontiguousarray(zz, dtype=np.float64)
        wr = np.ascontiguousarray(wr, dtype=np.float64)
        _check(L.fribo_init(C.byref(p), _dp(rc), _dp(zz), _dp(zmm), _dp(zmmt), zmm.shape[0], _dp(wr), C.byref(h)))
        self.h = h
        self.nlm = 0

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().fribo_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ------------------------------------------------------------
    def set_state(self, st):
        keys = ("q2", "q3", "pr", "ru2", "ru3", "qb2s", "qb3s", "qb2n", "qb3n", "dq2x2o", "dq3x2o")
        arrs = [None if st.get(k) is None else np.asfortranarray(st[k], dtype=np.float64) for k in keys]
        _check(load().fribo_set_state(self.h, *[_dp(a) for a in arrs]))

    def get_state(self):
        n2, n3m = self.n2, self.n3 - 1
        st = {k: np.zeros((n2, n3m + 6), order="F") for k in ("q2", "q3", "pr")}
        st.update({k: np.zeros((n2, n3m), order="F") for k in ("ru2", "ru3")})
        st.update({k: np.zeros(n2) for k in ("qb2s", "qb3s", "qb2n", "qb3n", "dq2x2o", "dq3x2o")})
        keys = ("q2", "q3", "pr", "ru2", "ru3", "qb2s", "qb3s", "qb2n", "qb3n", "dq2x2o", "dq3x2o")
        _check(load().fribo_get_state(self.h, *[_dp(st[k]) for k in keys]))
        return st

    def fetch_fields(self, kfirst, klast, st, fields=("q2", "q3", "pr", "ru2", "ru3")):
        """Planes kfirst..klast of the named fields into the arrays of st (the
        dict get_state returns; other planes untouched)."""
        bits = {"q2": 1, "q3": 2, "pr": 4, "ru2": 8, "ru3": 16}
        mask = sum(bits[f] for f in fields)
        arrs = [st[k] if k in fields else None for k in ("q2", "q3", "pr", "ru2", "ru3")]
        for a in arrs:
            assert a is None or (a.flags.f_contiguous and a.dtype == np.float64)
        _check(load().fribo_fetch_fields(self.h, mask, int(kfirst), int(klast), *[_dp(a) for a in arrs]))

    def set_markers(self, lmcc, lmvel, lmA):
        lmcc = np.asfortranarray(lmcc, dtype=np.float64)
        lmvel = np.asfortranarray(lmvel, dtype=np.float64)
        lmA = np.ascontiguousarray(lmA, dtype=np.float64)
        self._mk = (lmcc, lmvel, lmA)
        self.nlm = lmA.size
        _check(load().fribo_set_markers(self.h, lmA.size, _dp(lmcc), _dp(lmvel), _dp(lmA)))

    # -- stepping -----------------------------------------------------------
    def fluid_tsch(self, ns, dt, dt_o=None):
        d = C.c_double(dt)
        cf = C.c_double(0.0)
        _check(load().fribo_fluid_tsch(self.h, ns, C.byref(d), float(dt if dt_o is None else dt_o), C.byref(cf)))
        return d.value, cf.value

    def fluid_tsch_rc(self, ns, dt, dt_o=None):
        """Like fluid_tsch but returns (code, dt, cflm) instead of raising."""
        d = C.c_double(dt)
        cf = C.c_double(0.0)
        rc = load().fribo_fluid_tsch(self.h, ns, C.byref(d), float(dt if dt_o is None else dt_o), C.byref(cf))
        return rc, d.value, cf.value

    def cfl_dt(self, ns, dt, dt_o=None):
        d = C.c_double(dt)
        cf = C.c_double(0.0)
        _check(load().fribo_cfl_dt(self.h, ns, C.byref(d), float(dt if dt_o is None else dt_o), C.byref(cf)))
        return d.value, cf.value

    def substep(self, n, ns, folded=False):
        """Substep n; folded: the launch sequence of a whole step (the mean
        of substep n-1 runs inside substep n's boundary kernel)."""
        L = load()
        _check((L.fribo_substep_folded if folded else L.fribo_substep)(self.h, n, ns))

    def sync(self):
        _check(load().fribo_sync(self.h))

    # -- arrays -------------------------------------------------------------
    def get(self, name):
        L = load()
        n = L.fribo_array_size(self.h, name.encode())
        if n <= 0:
            raise KeyError(name)
        out = np.zeros(n, dtype=np.int32 if name in _INT else np.float64)
        _check(L.fribo_get_array(self.h, name.encode(), out.ctypes.data_as(C.c_void_p), n))
        n2, n3m = self.n2, self.n3 - 1
        if n == n2 * (n3m + 6) and name not in _INT:
            return out.reshape((n2, n3m + 6), order="F")
        if n == n2 * (n3m + 5) and name in ("fey", "fez"):
            return out.reshape((n2, n3m + 5), order="F")
        if n == n2 * n3m and name in ("ru2", "ru3", "rhs2", "rhs3"):
            return out.reshape((n2, n3m), order="F")
        return out

    def loads_transfer(self, kprobe, lmnor, lmacc):
        """loads_transfer_c after the last step: (press (nlm,), tau (nlm, 2))."""
        lmnor = np.asfortranarray(lmnor, dtype=np.float64)
        lmacc = np.asfortranarray(lmacc, dtype=np.float64)
        press = np.zeros(self.nlm)
        tau = np.zeros((self.nlm, 2), order="F")
        _check(load().fribo_loads_transfer(self.h, float(kprobe), _dp(lmnor), _dp(lmacc), _dp(press), _dp(tau)))
        return press, tau

    def save_state(self):
        """Device rollback copy of the prognostic state (iterated coupling)."""
        _check(load().fribo_save_state(self.h))

    def restore_state(self):
        _check(load().fribo_restore_state(self.h))

    def diagnostics(self):
        """ttime's per-step diagnostics (dict, same keys as the oracle's)."""
        out = np.zeros(9)
        _check(load().fribo_diagnostics(self.h, _dp(out)))
        return dict(zip(DIAG_KEYS, out))

    def checkbcs(self):
        """checkbcs after the last step: (bcerrv, bcerrw, nbcv, nbcw)."""
        bv, bw = np.zeros(self.nlm), np.zeros(self.nlm)
        nv, nw = C.c_double(0.0), C.c_double(0.0)
        _check(load().fribo_checkbcs(self.h, _dp(bv), _dp(bw), C.byref(nv), C.byref(nw)))
        return bv, bw, nv.value, nw.value

    def set_transform(self, axis):
        """Poisson eigen-transform axis: "z" (reference order), "y" or "auto"."""
        a = {"z": 0, "y": 1, "auto": -1}[axis]
        _check(load().fribo_set_transform(self.h, a))

    @property
    def transform(self):
        return "y" if load().fribo_get_transform(self.h) == 1 else "z"

    def set_ksolve(self, form):
        """k-line solve form: "fused" (one pass, lines of at most 2048 rows)
        or "3pass"."""
        _check(load().fribo_set_ksolve(self.h, {"3pass": 0, "fused": 1}[form]))

    @property
    def ksolve(self):
        return "fused" if load().fribo_get_ksolve(self.h) == 1 else "3pass"

    def set_ymirror(self, on):
        """Mirror form of the transform along y (grids symmetric about the
        channel centre): True halves the transforms' flops, False takes the
        dense form."""
        _check(load().fribo_set_ymirror(self.h, 1 if on else 0))

    @property
    def ymirror(self):
        return load().fribo_get_ymirror(self.h) == 1

    def set_precision(self, refine):
        """Poisson transforms: refine < 0 FP64; n >= 0 FP32 MFMA plus n
        FP64-residual refinement sweeps."""
        _check(load().fribo_set_precision(self.h, int(refine)))

    @property
    def precision(self):
        return load().fribo_get_precision(self.h)

    def set_graph(self, on=True):
        """Replay the step as one captured HIP graph (same results)."""
        _check(load().fribo_set_graph(self.h, 1 if on else 0))

    # -- profiling ----------------------------------------------------------
    def set_profiling(self, on=True, gemm_only=False):
        """Event timing of every kernel class (or of the transform GEMMs only)."""
        _check(load().fribo_set_profiling(self.h, (2 if gemm_only else 1) if on else 0))

    def reset_profile(self):
        _check(load().fribo_reset_profile(self.h))

    def profile(self):
        L = load()
        ms = np.zeros(NPROF)
        cnt = (C.c_int32 * NPROF)()
        _check(L.fribo_get_profile(self.h, _dp(ms), cnt, NPROF))
        return {L.fribo_profile_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(NPROF)}


def _padded(X, ld):
    """Column-major copy of X with leading dimension ld (zero padding)."""
    P = np.zeros((ld, X.shape[1]), order="F")
    P[: X.shape[0]] = X
    return P


def gemm_nt(A, B, pad=0):
    """C = A @ B.T on the GPU through the library's MFMA kernels (test hook).
    ``pad`` rounds the leading dimensions up to a multiple of it (the layout
    of the library's own operands, which admits the direct-to-LDS kernel)."""
    L = load()
    M, K = A.shape
    N = B.shape[0]
    la = -(-M // pad) * pad if pad else M
    lb = -(-N // pad) * pad if pad else N
    Ap = _padded(np.asarray(A, dtype=np.float64), la)
    Bp = _padded(np.asarray(B, dtype=np.float64), lb)
    Cm = np.zeros((la, N), order="F")
    _check(L.fribo_test_gemm(M, N, K, _dp(Ap), la, _dp(Bp), lb, _dp(Cm), la))
    return Cm[:M]


def gemm_nn(A, B, pad=0):
    """C = A @ B on the GPU (the form of the transforms along y; test hook)."""
    L = load()
    M, K = A.shape
    N = B.shape[1]
    la = -(-M // pad) * pad if pad else M
    lb = -(-K // pad) * pad if pad else K
    Ap = _padded(np.asarray(A, dtype=np.float64), la)
    Bp = _padded(np.asarray(B, dtype=np.float64), lb)
    Cm = np.zeros((la, N), order="F")
    _check(L.fribo_test_gemm_nn(M, N, K, _dp(Ap), la, _dp(Bp), lb, _dp(Cm), la))
    return Cm[:M]


def gemm32(A, B):
    """C = A @ B by the FP32 transform kernel (operands rounded to FP32,
    FP32 accumulation; test hook)."""
    L = load()
    M, K = A.shape
    N = B.shape[1]
    Ap = np.asfortranarray(A, dtype=np.float64)
    Bp = np.asfortranarray(B, dtype=np.float64)
    Cm = np.zeros((M, N), order="F")
    _check(L.fribo_test_gemm32(M, N, K, _dp(Ap), M, _dp(Bp), K, _dp(Cm), M))
    return Cm


def plain_copy(nbytes, reps):
    """Time `reps` device-to-device streaming copies of `nbytes`; ms per copy."""
    ms = C.c_double(0.0)
    _check(load().fribo_test_copy(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def ytransform(am, ac, ap):
    """Host eigen-decomposition of a tridiagonal operator (rows am, ac, ap):
    (Y, Yinv, wy) with operator @ Y = Y * wy."""
    n = len(ac)
    Y = np.zeros((n, n), order="F")
    Yi = np.zeros((n, n), order="F")
    w = np.zeros(n)
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (am, ac, ap)]
    _check(load().fribo_test_ytransform(n, *[_dp(x) for x in arr], _dp(Y), _dp(Yi), _dp(w)))
    return Y, Yi, w


def ymirror(am, ac, ap, tol=1e-10):
    """Mirror form of the decomposition of an operator symmetric about its
    centre: (Ys, Ysi, wys), block-diagonal backward and forward transforms
    on the pair sums and differences and the eigenvalues in their mode
    order; None when the operator is not symmetric to within tol."""
    n = len(ac)
    Ys = np.zeros((n, n), order="F")
    Ysi = np.zeros((n, n), order="F")
    w = np.zeros(n)
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (am, ac, ap)]
    rc = load().fribo_test_ymirror(n, *[_dp(x) for x in arr], float(tol), _dp(Ys), _dp(Ysi), _dp(w))
    if rc == 7:
        return None
    _check(rc)
    return Ys, Ysi, w


def phini(zz, alx3, device=False):
    """phini's (zmm, zmmt, wr) by the library's symmetrised decomposition:
    implicit QL on the host, or (device=True) rocSOLVER's STEDC on the GPU."""
    zz = np.ascontiguousarray(zz, dtype=np.float64)
    n = zz.size - 1
    zmm = np.zeros((n, n), order="F")
    zmmt = np.zeros((n, n), order="F")
    wr = np.zeros(n)
    fn = load().fribo_phini_device if device else load().fribo_phini
    _check(fn(zz.size, _dp(zz), float(alx3), n, _dp(zmm), _dp(zmmt), _dp(wr)))
    return zmm, zmmt, wr


class FluidBatch:
    """Independent FluidStep instances advanced together, one kernel launch
    per stage of the step for the whole batch (fribo_batch_*)."""

    def __init__(self, steps):
        self.steps = list(steps)
        L = load()
        hs = (C.c_void_p * len(self.steps))(*[s.h.value for s in self.steps])
        b = C.c_void_p()
        _check(L.fribo_batch_create(hs, len(self.steps), C.byref(b)))
        self.b = b

    def close(self):
        if getattr(self, "b", None) is not None and self.b.value:
            load().fribo_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_markers(self, markers):
        """markers[i] = (lmcc, lmvel, lmA) of instance i, with the counts of
        the instance's last set_markers: one staged copy for the batch."""
        parts = []
        for (lmcc, lmvel, lmA), s in zip(markers, self.steps):
            if np.size(lmA) != s.nlm:
                raise ValueError("marker count differs from the instance's last set_markers")
            parts += [np.asarray(lmcc, dtype=np.float64).ravel(order="F"),
                      np.asarray(lmvel, dtype=np.float64).ravel(order="F"),
                      np.asarray(lmA, dtype=np.float64).ravel()]
        buf = np.ascontiguousarray(np.concatenate(parts)) if parts else np.zeros(1)
        _check(load().fribo_batch_set_markers(self.b, _dp(buf)))

    def step(self, ns, dt, dt_o=None):
        """One step of every instance: (dt, cflm, err) arrays."""
        n = len(self.steps)
        d = np.array(dt, dtype=np.float64).reshape(n).copy()
        do = None if dt_o is None else np.ascontiguousarray(dt_o, dtype=np.float64)
        cf = np.zeros(n)
        err = (C.c_int32 * n)()
        _check(load().fribo_batch_step(self.b, ns, _dp(d), _dp(do), _dp(cf), err))
        return d, cf, np.array(err[:], dtype=np.int32)

    def set_graph(self, on):
        _check(load().fribo_batch_set_graph(self.b, 1 if on else 0))

    def set_gemm_shape(self, cfg):
        _check(load().fribo_batch_set_gemm_shape(self.b, int(cfg)))

    def gemm_shape(self):
        return load().fribo_batch_get_gemm_shape(self.b)

    def autotune(self, reps=5):
        """Pick the transform tile shape by timing; returns (best, ms per
        shape of the transform alone, ms per shape with the chunk pass of
        the per-mode z-lines that the shape does not fuse, which decides)."""
        ms = np.zeros(10)
        best = C.c_int32(-1)
        _check(load().fribo_batch_autotune(self.b, int(reps), _dp(ms), C.byref(best)))
        return best.value, ms[:5], ms[5:]

    def set_markers_flat(self, buf):
        """Marker kinematics of all instances, already laid out as
        fribo_batch_set_markers takes them."""
        buf = np.ascontiguousarray(buf, dtype=np.float64)
        _check(load().fribo_batch_set_markers(self.b, _dp(buf)))


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The library's coupling-predictor network (fribo_mlp_*): FP32 training
    on the GPU, BF16 inference."""

    def __init__(self, ni, no, params=None, seed=1):
        L = load()
        h = C.c_void_p()
        init = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        _check(L.fribo_mlp_create(int(ni), int(no), _fp(init), int(seed), C.byref(h)))
        self.h, self.ni, self.no = h, int(ni), int(no)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().fribo_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self):
        return load().fribo_mlp_size(self.h)

    def params(self):
        out = np.zeros(self.size, dtype=np.float32)
        _check(load().fribo_mlp_get_params(self.h, _fp(out)))
        return out

    def grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.size, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(load().fribo_mlp_grad(self.h, X.shape[0], _fp(X), _fp(T), _fp(g), _fp(loss)))
        return g, float(loss[0])

    def train(self, X, T, epochs, batch, lr):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(load().fribo_mlp_train(self.h, X.shape[0], _fp(X), _fp(T), int(epochs), int(batch), float(lr),
                                      _fp(loss)))
        return float(loss[0])

    def predict(self, X):
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.float32)
        Y = np.zeros((X.shape[0], self.no), dtype=np.float32)
        _check(load().fribo_mlp_predict(self.h, X.shape[0], _fp(X), _fp(Y)))
        return Y
