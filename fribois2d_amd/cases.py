"""Input generation for the fluid step: grids, markers, initial state, phini.

These mirror the reference's pre-processing so that runs can be built without
MATLAB or the Fortran set-up code:

* ``cubicstr`` restates ``preproc/cubicstr.m:2-38`` (uniform core plus cubic
  stretching on both sides); it reproduces the shipped ``ycoord.dat`` /
  ``zcoord.dat`` files of both example cases to their 14 printed decimals.
* ``read_lmdata`` follows ``read_rigid`` (``src/utils_proc.F90:519-578``),
  including the marker count ``wc -l - 1`` that drops the closing segment of
  files whose last line has no newline.
* ``phini`` produces ``zmm``, ``zmmt`` and ``wr`` as ``src/phini.F90:4-99``
  does (eigen-decomposition of the z pressure operator).  ``method="dgeev"``
  calls LAPACK's DGEEV/DGESV through scipy exactly like the reference;
  ``method="sym"`` uses the diagonal similarity that makes the operator
  symmetric and a tridiagonal symmetric eigen-solver (much faster, same
  eigen-pairs up to scaling and order, which the Poisson solve ignores).
* ``initial_state`` follows ``inqpr`` (``src/initial_sol.F90:104-152``) and
  adds the seeded 1e-3 perturbation (splitmix64) the reference lacks.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, Optional

import numpy as np

__all__ = [
    "FluidParams", "Case", "cubicstr", "read_grid_file", "read_lmdata", "markers_from_vertices",
    "cylinder_vertices", "rectangle_vertices", "createnorm", "tridiag_z", "phini", "initial_state",
    "splitmix64_uniform", "make_case", "CASES",
]


@dataclasses.dataclass
class FluidParams:
    """The scalar inputs ``fluid_tsch`` reads (fluid_par.in / global_par.in)."""
    n2: int
    n3: int
    alx3: float
    rext2: float
    visc: float
    nsst: int = 3
    idtv: bool = False       # True: fixed dt, CFL checked against cfllim
    ibact: bool = True
    forst: int = 1
    wfun: int = 1
    inslws: bool = False
    inslwn: bool = False
    cou: float = 0.8
    q3uw: float = 0.0
    q3lw: float = 0.0
    dtmax: float = 0.01
    cflmax: float = 0.3
    cfllim: float = 2.0
    radiusSD: float = 1.6
    dt0: float = 1.0e-3


# --------------------------------------------------------------------------
# grids

def cubicstr(interval, nint, li, le):
    """preproc/cubicstr.m:2-38 -- node coordinates, uniform core in ``interval``."""
    it = np.asarray(interval, dtype=np.float64) - li
    le = float(le) - li
    n1, n2, n3 = (int(v) for v in nint)
    xnm = np.linspace(it[0], it[1], n2 + 2)[1:-1]
    dx0 = (it[1] - it[0]) / (n2 + 2)
    B = dx0 * n3
    A = (le - it[1]) - B
    s = np.linspace(0.0, 1.0, n3)
    xn2 = A * s ** 3 + B * s + it[1]
    B = dx0 * n1
    A = it[0] - B
    s = np.linspace(0.0, 1.0, n1)
    xn1 = -(A * s ** 3 + B * s) + it[0]
    xn1 = xn1[::-1]
    return np.concatenate([xn1, xnm, xn2]) + li


def read_grid_file(path, n):
    """cordin.F90:243-278 (readgrid): ``n`` list-directed reals."""
    vals = np.array(open(path).read().split(), dtype=np.float64)
    if vals.size < n:
        raise ValueError(f"{path}: {vals.size} values, {n} needed")
    return vals[:n].copy()


def read_lmdata(path):
    """utils_proc.F90:519-578 (read_rigid): vertices with nlm = (newlines - 1)."""
    raw = open(path, "rb").read()
    nlm = raw.count(b"\n") - 1
    toks = raw.decode().split()
    verts = np.array(toks[: 2 * (nlm + 1)], dtype=np.float64).reshape(nlm + 1, 2)
    return verts


def markers_from_vertices(verts):
    """utils_proc.F90:560-565: marker = segment midpoint, lmA = segment length.

    Returns ``lmcc`` (nlm, 2) as (y, z) and ``lmA`` (nlm,).
    """
    v = np.asarray(verts, dtype=np.float64)
    lmcc = (v[:-1] + v[1:]) * 0.5
    d = v[1:] - v[:-1]
    lmA = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2)
    return lmcc, lmA


def cylinder_vertices(yc, zc, radius, nlm):
    """Closed polygon of nlm segments (preproc/MAIN_input_OscCyl1DOF.m:69-72)."""
    theta = np.linspace(0.0, 2.0 * math.pi, nlm + 1)
    return np.stack([radius * np.sin(theta) + yc, radius * np.cos(theta) + zc], axis=1)


def rectangle_vertices(yc, zc, chord, thick, nlm, angle=0.0):
    """Rectangle (chord along z) with nlm equal perimeter segments, rotated by angle."""
    per = 2.0 * (chord + thick)
    s = np.linspace(0.0, per, nlm + 1)
    s[-1] = 0.0
    hz, hy = chord / 2.0, thick / 2.0
    pts = []
    for t in s:
        if t < chord:
            z, y = -hz + t, -hy
        elif t < chord + thick:
            z, y = hz, -hy + (t - chord)
        elif t < 2 * chord + thick:
            z, y = hz - (t - chord - thick), hy
        else:
            z, y = -hz, hy - (t - 2 * chord - thick)
        pts.append((y, z))
    p = np.array(pts)
    c, s_ = math.cos(angle), math.sin(angle)
    y = c * p[:, 0] + s_ * p[:, 1]
    z = -s_ * p[:, 0] + c * p[:, 1]
    return np.stack([y + yc, z + zc], axis=1)


# --------------------------------------------------------------------------
# pressure operator eigen-decomposition

def tridiag_z(zz, alx3):
    """phini.F90:133-142: 1-D Neumann pressure operator along z (amphk, acphk, apphk)."""
    zz = np.asarray(zz, dtype=np.float64)
    n3 = zz.size
    n3m = n3 - 1
    dx3 = 1.0 / (alx3 / n3m)
    dx3q = dx3 * dx3
    g3rm = (zz[1:] - zz[:-1]) * dx3                       # k = 1..n3m
    g3rc = np.empty(n3)
    g3rc[1:n3m] = (zz[2:] - zz[:-2]) * dx3 * 0.5          # k = 2..n3m
    g3rc[0] = (zz[1] - zz[0]) * dx3
    g3rc[n3m] = (zz[n3m] - zz[n3m - 1]) * dx3
    am = np.zeros(n3m)
    ap = np.zeros(n3m)
    k = np.arange(n3m)
    kmc = np.where(k == 0, 0, 1)
    kpc = np.where(k == n3m - 1, 0, 1)
    kp = np.minimum(k + 1, n3m - 1)
    ug = 1.0 / g3rm
    am = (kmc * dx3q / g3rc[k]) * ug
    ap = (kpc * dx3q / g3rc[kp]) * ug
    ac = -(am + ap)
    return am, ac, ap


def phini(zz, alx3, method="sym"):
    """zmm (eigenvectors as columns), zmmt = zmm^-1, wr, each n3m-sized."""
    am, ac, ap = tridiag_z(zz, alx3)
    n = ac.size
    if method == "dgeev":
        import scipy.linalg as sla
        A = np.diag(ac) + np.diag(ap[:-1], 1) + np.diag(am[1:], -1)
        w, vr = sla.eig(A, left=False, right=True, check_finite=False)
        zmm = np.ascontiguousarray(vr.real)
        zmmt = np.linalg.solve(zmm, np.eye(n))
        return np.asfortranarray(zmm), np.asfortranarray(zmmt), w.real.copy()
    import scipy.linalg as sla
    # A = D S D^-1 with S symmetric: (d_{i+1}/d_i)^2 = a_{i+1}/c_i
    sub = am[1:]
    sup = ap[:-1]
    ratio = np.sqrt(sub / sup)
    d = np.concatenate([[1.0], np.cumprod(ratio)])
    # rescale to avoid overflow on long stretched grids
    d = d / np.sqrt(np.max(d) * np.min(d))
    off = np.sqrt(sub * sup)
    w, q = sla.eigh_tridiagonal(ac, off, lapack_driver="stemr")
    v = d[:, None] * q                      # eigenvectors of A
    norms = np.linalg.norm(v, axis=0)
    zmm = v / norms[None, :]
    zmmt = (q.T / d[None, :]) * norms[:, None]
    return np.asfortranarray(zmm), np.asfortranarray(zmmt), w.copy()


# --------------------------------------------------------------------------
# initial state

def splitmix64_uniform(seed, count):
    """``count`` values in [-1, 1) from splitmix64 (53-bit mantissa draws)."""
    mask = (1 << 64) - 1
    out = np.empty(count, dtype=np.float64)
    x = seed & mask
    for i in range(count):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / 9007199254740992.0)
    return 2.0 * out - 1.0


def _splitmix64_uniform_np(seed, count):
    """Vectorised splitmix64 (bit-identical to ``splitmix64_uniform``)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, count + 1, dtype=np.uint64)
        x = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def initial_state(p: FluidParams, seed: Optional[int] = None, amp: float = 1.0e-3):
    """inqpr (initial_sol.F90:104-152) plus the seeded velocity perturbation.

    Returns a dict of arrays in the reference layout (Fortran order):
    q2, q3, pr (n2, n3m+6 planes k=-2..n3m+3), ru2, ru3 (n2, n3m) and the
    outflow arrays qb2s, qb3s, qb2n, qb3n, dq2x2o, dq3x2o (n2).
    Perturbed entries: q2(2:n2m, 1:n3m) then q3(1:n2m, 2:n3m), j fastest.
    """
    n2, n3 = p.n2, p.n3
    n2m, n3m = n2 - 1, n3 - 1
    shape = (n2, n3m + 6)
    q2 = np.zeros(shape, order="F")
    q3 = np.zeros(shape, order="F")
    pr = np.zeros(shape, order="F")
    # initia zeroes planes 1..n3; inqpr sets planes kstart-1..kend+1 = 0..n3
    off = 2
    q3[:, off + 0: off + n3 + 1] = 1.0
    if seed is not None:
        cnt2 = (n2m - 1) * n3m
        cnt3 = n2m * (n3m - 1)
        u = _splitmix64_uniform_np(seed, cnt2 + cnt3)
        q2[1:n2m, off + 1: off + n3m + 1] += amp * u[:cnt2].reshape((n2m - 1, n3m), order="F")
        q3[0:n2m, off + 2: off + n3m + 1] += amp * u[cnt2:].reshape((n2m, n3m - 1), order="F")
    qb = {k: np.zeros(n2) for k in ("qb2s", "qb3s", "qb2n", "qb3n", "dq2x2o", "dq3x2o")}
    qb["qb3s"][:n2m] = 1.0
    qb["qb3n"][:n2m] = 1.0
    st = dict(q2=q2, q3=q3, pr=pr, ru2=np.zeros((n2, n3m), order="F"),
              ru3=np.zeros((n2, n3m), order="F"))
    st.update(qb)
    return st


# --------------------------------------------------------------------------
# cases

@dataclasses.dataclass
class Case:
    name: str
    params: FluidParams
    rc: np.ndarray
    zz: np.ndarray
    verts: Optional[np.ndarray]          # closed polygon (nlm+1, 2) as (y, z)
    center: tuple = (0.0, 0.0)           # (y, z) of the body reference point
    motion: Optional[Callable] = None    # t -> (dy, dz, dtheta, vy, vz, omega)
    seed: Optional[int] = None
    # markers of the body at rest (markers_from_vertices), computed once
    _rest: Optional[tuple] = dataclasses.field(default=None, repr=False, compare=False)

    @property
    def nlm(self):
        return 0 if self.verts is None else self.verts.shape[0] - 1

    def markers(self, t=0.0):
        """lmcc (nlm,2), lmvel (nlm,2), lmA (nlm,) for the body at time t
        (the 2-column arrays in Fortran order, as the library takes them)."""
        if self._rest is None:
            lmcc0, lmA = markers_from_vertices(self.verts)
            yc, zc = self.center
            self._rest = (lmcc0, lmA, lmcc0[:, 0] - yc, lmcc0[:, 1] - zc)
        lmcc0, lmA, ry, rz = self._rest
        if self.motion is None:
            return np.asfortranarray(lmcc0), np.zeros((lmA.size, 2), order="F"), lmA.copy()
        dy, dz, th, vy, vz, om = self.motion(t)
        yc, zc = self.center
        c, s = math.cos(th), math.sin(th)
        ryr, rzr = c * ry - s * rz, s * ry + c * rz
        lmcc = np.empty((lmA.size, 2), order="F")
        lmvel = np.empty((lmA.size, 2), order="F")
        np.add(yc + dy, ryr, out=lmcc[:, 0])
        np.add(zc + dz, rzr, out=lmcc[:, 1])
        np.subtract(vy, om * rzr, out=lmvel[:, 0])
        np.add(vz, om * ryr, out=lmvel[:, 1])
        return lmcc, lmvel, lmA.copy()

    def surface(self, t=0.0):
        """lmnor (nlm,2) and lmacc (nlm,2) as rigid() leaves them at time t:
        normals of the moved polygon by createnorm (rigid.F90:468-495, called
        with or=-1 at rigid.F90:123) and the rigid-body acceleration of every
        marker (rigid.F90:84-104)."""
        lmcc0, _ = markers_from_vertices(self.verts)
        yc, zc = self.center
        if self.motion is None:
            dy = dz = th = om = ay = az = al = 0.0
        else:
            dy, dz, th, _, _, om = self.motion(t)[:6]
            ay, az, al = self.motion.accel(t) if hasattr(self.motion, "accel") else (0.0, 0.0, 0.0)
        c, s_ = math.cos(th), math.sin(th)
        vr = self.verts - np.array([yc, zc])
        verts = np.stack([yc + dy + c * vr[:, 0] - s_ * vr[:, 1], zc + dz + s_ * vr[:, 0] + c * vr[:, 1]], axis=1)
        lmnor = createnorm(verts, -1.0)
        ry, rz = lmcc0[:, 0] - yc, lmcc0[:, 1] - zc
        ryr, rzr = c * ry - s_ * rz, s_ * ry + c * rz
        # accP = a_c + alpha x rho + omega x (omega x rho)
        c1y, c1z = -om * rzr, om * ryr
        lmacc = np.stack([ay - al * rzr - om * c1z, az + al * ryr + om * c1y], axis=1)
        return lmnor, lmacc


def createnorm(verts, orient=-1.0):
    """rigid.F90:468-495: unit normal of every segment from the cross product
    of the segment (x=0 plane) with the vector to its midpoint lifted to x=or."""
    v = np.asarray(verts, dtype=np.float64)
    nlm = v.shape[0] - 1
    out = np.zeros((nlm, 2))
    for i in range(nlm):
        t1 = np.array([0.0, v[i, 0], v[i, 1]])
        t2 = np.array([0.0, v[i + 1, 0], v[i + 1, 1]])
        t3 = np.array([orient, (v[i, 0] + v[i + 1, 0]) * 0.5, (v[i, 1] + v[i + 1, 1]) * 0.5])
        a, b = t2 - t1, t3 - t1
        cr = np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])
        cr = cr / math.sqrt(float(np.sum(cr ** 2)))
        out[i] = cr[1:]
    return out


def _heave(amp, freq):
    w = 2.0 * math.pi * freq

    def motion(t):
        return (amp * math.sin(w * t), 0.0, 0.0, amp * w * math.cos(w * t), 0.0, 0.0)
    motion.accel = lambda t: (-amp * w * w * math.sin(w * t), 0.0, 0.0)
    motion.amp, motion.w = amp, w
    return motion


def _pitch(amp, freq):
    w = 2.0 * math.pi * freq

    def motion(t):
        return (0.0, 0.0, amp * math.sin(w * t), 0.0, 0.0, amp * w * math.cos(w * t))
    motion.accel = lambda t: (0.0, 0.0, -amp * w * w * math.sin(w * t))
    return motion


def make_case(name: str, instance: int = 0) -> Case:
    """Named workloads.

    ``cyl384``   cylinder D=1, Re=100, 384x288 cells, core dx=D/40, 126 markers,
                 CFL 0.3, heaving (1-DOF transverse) marker motion.
    ``rect2048`` rectangle chord 1 / thickness 0.25, Re=200, 2048x1536 cells,
                 core dx=D/100, 1000 markers, pitching about its centre.
    ``cyl64``    small cylinder case (64x48 cells) for fast parity tests.
    ``tiny``     uniform 24x18 grid, 20 markers.
    ``osccyl``/``gallrect`` the reference's own example inputs (test_*/input_FSI),
                 read from a directory given by the ``FRIBO_REF_INPUTS`` mapping
                 in the tests (not shipped here).
    """
    seed = 20261202 + instance
    if name == "cyl384":
        zz = cubicstr([8.0, 16.0], [17, 319, 49], 0.0, 40.0)
        rc = cubicstr([13.0, 17.0], [65, 159, 65], 0.0, 30.0)
        p = FluidParams(n2=rc.size, n3=zz.size, alx3=40.0, rext2=30.0, visc=0.01)
        verts = cylinder_vertices(15.0, 10.0, 0.5, 126)
        return Case(name, p, rc, zz, verts, (15.0, 10.0), _heave(0.1, 0.2), seed)
    if name == "rect2048":
        zz = cubicstr([8.0, 20.0], [243, 1199, 607], 0.0, 40.0)
        rc = cubicstr([12.0, 18.0], [469, 599, 469], 0.0, 30.0)
        p = FluidParams(n2=rc.size, n3=zz.size, alx3=40.0, rext2=30.0, visc=0.005)
        verts = rectangle_vertices(15.0, 10.0, 1.0, 0.25, 1000)
        return Case(name, p, rc, zz, verts, (15.0, 10.0), _pitch(math.radians(5.0), 0.125), seed)
    if name == "cyl64":
        zz = cubicstr([2.0, 6.0], [9, 39, 17], 0.0, 10.0)
        rc = cubicstr([2.0, 5.0], [9, 31, 9], 0.0, 7.0)
        p = FluidParams(n2=rc.size, n3=zz.size, alx3=10.0, rext2=7.0, visc=0.01)
        verts = cylinder_vertices(3.5, 3.0, 0.5, 32)
        return Case(name, p, rc, zz, verts, (3.5, 3.0), _heave(0.1, 0.5), seed)
    if name == "tiny":
        n2m, n3m = 18, 24
        rc = np.linspace(0.0, 6.0, n2m + 1)
        zz = np.linspace(0.0, 8.0, n3m + 1)
        p = FluidParams(n2=n2m + 1, n3=n3m + 1, alx3=8.0, rext2=6.0, visc=0.02)
        verts = cylinder_vertices(3.0, 3.0, 1.0, 20)
        return Case(name, p, rc, zz, verts, (3.0, 3.0), _heave(0.1, 0.5), seed)
    raise KeyError(name)


def ensemble_case(instance: int, n3m: int, n2m: int, re: float, amp: float = 0.1, freq: float = 0.2) -> Case:
    """One member of a batched ensemble: cylinder D = 1 at (y, z) = (15, 10)
    in [0, 30] x [0, 40] on an n3m x n2m grid whose uniform core covers
    z in [8, 16], y in [13, 17] with about 83% / 55% of the nodes (the
    proportions of ``cyl384``), Reynolds number re, heaving with amplitude
    amp and frequency freq; markers about pi D / dx, clipped to 63..250."""
    n3, n2 = n3m + 1, n2m + 1
    zc = max(int(round(0.83 * n3)), 8)
    zb = max(int(round(0.045 * n3)), 3)
    yc = max(int(round(0.55 * n2)), 8)
    yb = max((n2 - yc) // 2, 3)
    zz = cubicstr([8.0, 16.0], [zb, zc, n3 - zb - zc], 0.0, 40.0)
    rc = cubicstr([13.0, 17.0], [yb, yc, n2 - yb - yc], 0.0, 30.0)
    dx = 8.0 / (zc - 1)
    nlm = int(min(250, max(63, round(math.pi / dx))))
    p = FluidParams(n2=rc.size, n3=zz.size, alx3=40.0, rext2=30.0, visc=1.0 / re)
    verts = cylinder_vertices(15.0, 10.0, 0.5, nlm)
    return Case(f"ens{instance}", p, rc, zz, verts, (15.0, 10.0), _heave(amp, freq), 20261202 + instance)


def reference_example(directory, which="osccyl"):
    """The reference's own example inputs (test_OscCyl1DOF / test_GallRect1DOF).

    Parameters follow their fluid_par.in / global_par.in: fixed dt = 1e-3,
    idtv = .true., cfllim = 2, Re = 100 (OscCyl) or 250 (GallRect).
    """
    import os
    if which == "osccyl":
        n3, n2, visc = 811, 673, 0.01
    elif which == "gallrect":
        n3, n2, visc = 731, 561, 0.004
    else:
        raise KeyError(which)
    rc = read_grid_file(os.path.join(directory, "ycoord.dat"), n2)
    zz = read_grid_file(os.path.join(directory, "zcoord.dat"), n3)
    verts = read_lmdata(os.path.join(directory, "lmdata.in"))
    p = FluidParams(n2=n2, n3=n3, alx3=40.0, rext2=60.0, visc=visc, idtv=True, dt0=1.0e-3,
                    dtmax=5.0e-3, cflmax=0.2, cfllim=2.0)
    center = (float(np.mean(verts[:-1, 0])), float(np.mean(verts[:-1, 1])))
    return Case(which, p, rc, zz, verts, center, _heave(0.05, 0.2), 20261202)


CASES = ("tiny", "cyl64", "cyl384", "rect2048")
