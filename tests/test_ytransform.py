"""Host eigen-decomposition behind the Poisson transform along y (CPU only:
the set-up code of libfribo_gpu, no kernel runs)."""
import numpy as np
import pytest

from fribois2d_amd import cases


def _ops(c):
    return cases.tridiag_z(c.rc, c.params.rext2)


def test_uniform_grid_matches_the_closed_form():
    """Uniform spacing: the Neumann operator dx2q*tridiag(1,-2,1) (ends -1)
    has eigenvalues -4*dx2q*sin^2(pi*m/(2n)), m = 0..n-1."""
    from fribois2d_amd import gpu
    n, L = 40, 3.0
    rc = np.linspace(0.0, L, n + 1)
    am, ac, ap = cases.tridiag_z(rc, L)
    Y, Yi, w = gpu.ytransform(am, ac, ap)
    dx2q = (n / L) ** 2
    ref = np.sort(-4.0 * dx2q * np.sin(np.pi * np.arange(n) / (2 * n)) ** 2)
    np.testing.assert_allclose(np.sort(w), ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    # eigenvectors: cos(pi*m*(j-1/2)/n) up to scale
    m = int(np.argsort(w)[-2])            # second smallest |w|: m = 1
    v = Y[:, m] / Y[0, m]
    j = np.arange(1, n + 1)
    np.testing.assert_allclose(v, np.cos(np.pi * (j - 0.5) / n) / np.cos(np.pi * 0.5 / n), atol=1e-12)


@pytest.mark.parametrize("name", ["cyl64", "cyl384"])
def test_stretched_grid_decomposition(name):
    from fribois2d_amd import gpu
    c = cases.make_case(name)
    am, ac, ap = _ops(c)
    n = ac.size
    A = np.diag(ac) + np.diag(ap[:-1], 1) + np.diag(am[1:], -1)
    Y, Yi, w = gpu.ytransform(am, ac, ap)
    scale = np.abs(A).max()
    assert np.abs(A @ Y - Y * w).max() <= 1e-13 * scale
    assert np.abs(Yi @ Y - np.eye(n)).max() <= 1e-13
    # exactly one Neumann null mode, the rest strictly negative
    ws = np.sort(w)
    assert abs(ws[-1]) <= 1e-9 * scale and ws[-2] < -1e-9 * scale
    # the same spectrum as a dense eigen-solver
    np.testing.assert_allclose(ws, np.sort(np.linalg.eigvals(A).real), rtol=0, atol=1e-11 * scale)


def test_library_phini_matches_the_operator():
    """fribo_phini: the z operator of tridiag_matrices decomposed without
    LAPACK; same spectrum as DGEEV (scipy) and a valid inverse pair."""
    from fribois2d_amd import gpu
    c = cases.make_case("cyl384")
    am, ac, ap = cases.tridiag_z(c.zz, c.params.alx3)
    A = np.diag(ac) + np.diag(ap[:-1], 1) + np.diag(am[1:], -1)
    zmm, zmmt, wr = gpu.phini(c.zz, c.params.alx3)
    scale = np.abs(A).max()
    assert np.abs(A @ zmm - zmm * wr).max() <= 1e-13 * scale
    assert np.abs(zmmt @ zmm - np.eye(ac.size)).max() <= 1e-13
    _, _, wref = cases.phini(c.zz, c.params.alx3, method="dgeev")
    np.testing.assert_allclose(np.sort(wr), np.sort(wref), rtol=0, atol=1e-10 * scale)


def _butterfly(n):
    """The pair butterfly (x_j, x_{n-1-j}) -> (x_j + x_{n-1-j}, x_j - x_{n-1-j})."""
    h = n // 2
    B = np.zeros((n, n))
    for j in range(h):
        B[j, j] = B[j, n - 1 - j] = 1.0
        B[n - 1 - j, j] = 1.0
        B[n - 1 - j, n - 1 - j] = -1.0
    return B


def test_mirror_form_uniform_grid_closed_form():
    """Uniform spacing, n = 40: the even block holds the eigenvalues
    -4 dx2q sin^2(pi m / (2n)) of even m (cosines symmetric about the
    centre), the odd block those of odd m, and the block transforms with the
    butterfly invert each other and diagonalise the operator."""
    from fribois2d_amd import gpu
    n, L = 40, 3.0
    rc = np.linspace(0.0, L, n + 1)
    am, ac, ap = cases.tridiag_z(rc, L)
    Ys, Ysi, w = gpu.ymirror(am, ac, ap)
    h = n // 2
    dx2q = (n / L) ** 2
    lam = -4.0 * dx2q * np.sin(np.pi * np.arange(n) / (2 * n)) ** 2
    tol = 1e-11 * np.abs(lam).max()
    np.testing.assert_allclose(np.sort(w[:h]), np.sort(lam[0::2]), rtol=0, atol=tol)
    np.testing.assert_allclose(np.sort(w[h:]), np.sort(lam[1::2]), rtol=0, atol=tol)
    B = _butterfly(n)
    Tf, Tb = Ysi @ B, B @ Ys
    A = np.diag(ac) + np.diag(ap[:-1], 1) + np.diag(am[1:], -1)
    assert np.abs(Tf @ Tb - np.eye(n)).max() <= 1e-13
    assert np.abs(Tb @ np.diag(w) @ Tf - A).max() <= 1e-12 * np.abs(A).max()


@pytest.mark.parametrize("name", ["cyl64", "cyl384"])
def test_mirror_form_stretched_grid(name):
    """Symmetric stretched grids: block-diagonal halves (the off-diagonal
    blocks exactly zero), the same spectrum as the dense decomposition, and
    with the butterfly an inverse pair that diagonalises the operator."""
    from fribois2d_amd import gpu
    c = cases.make_case(name)
    am, ac, ap = _ops(c)
    n = ac.size
    h = n // 2
    Ys, Ysi, w = gpu.ymirror(am, ac, ap)
    for M in (Ys, Ysi):
        assert not M[:h, h:].any() and not M[h:, :h].any()
    _, _, wd = gpu.ytransform(am, ac, ap)
    scale = np.abs(ac).max()
    np.testing.assert_allclose(np.sort(w), np.sort(wd), rtol=0, atol=1e-11 * scale)
    B = _butterfly(n)
    Tf, Tb = Ysi @ B, B @ Ys
    A = np.diag(ac) + np.diag(ap[:-1], 1) + np.diag(am[1:], -1)
    assert np.abs(Tf @ Tb - np.eye(n)).max() <= 1e-12
    assert np.abs(A @ Tb - Tb * w).max() <= 1e-12 * scale
    # exactly one null mode, in the even block
    k = int(np.argmin(np.abs(w)))
    assert k < h and abs(w[k]) <= 1e-9 * scale and np.sort(np.abs(w))[1] > 1e-9 * scale


def test_mirror_form_refuses_asymmetric_operators():
    """A grid that is not symmetric about the centre, or an odd size, has no
    mirror form."""
    from fribois2d_amd import gpu
    c = cases.make_case("cyl64")
    am, ac, ap = _ops(c)
    ac2 = ac.copy()
    ac2[3] *= 1.0 + 1e-8
    assert gpu.ymirror(am, ac2, ap) is None
    assert gpu.ymirror(am, ac2, ap, tol=1e-6) is not None
    rc = np.linspace(0.0, 3.0, 40) ** 1.2
    am3, ac3, ap3 = cases.tridiag_z(rc, float(rc[-1]))
    assert gpu.ymirror(am3, ac3, ap3) is None
    rc = np.linspace(0.0, 3.0, 42)   # 41 cells
    am4, ac4, ap4 = cases.tridiag_z(rc, 3.0)
    assert gpu.ymirror(am4, ac4, ap4) is None


@pytest.mark.parametrize("name", ["cyl384", "rect2048"])
def test_mirror_form_solves_shifted_systems(name):
    """The per-mode solve of the Poisson transform along y in the mirror
    form: butterfly, forward block transform, division by (w - sigma),
    backward block transform, butterfly, against a dense LU solve of
    (A - sigma I) y = x for a shift just past the spectrum's small end
    (condition number about 1e5) and one past its large end, on the
    benchmarked grid's operator too (n = 1536): within twice the error of
    the same solve with the dense eigen-decomposition (or 1e-13)."""
    from fribois2d_amd import gpu
    c = cases.make_case(name)
    am, ac, ap = _ops(c)
    n = ac.size
    A = np.diag(ac) + np.diag(ap[:-1], 1) + np.diag(am[1:], -1)
    Ys, Ysi, w = gpu.ymirror(am, ac, ap)
    Y, Yi, wd = gpu.ytransform(am, ac, ap)
    B = _butterfly(n)
    x = np.random.default_rng(7).standard_normal(n)
    for sigma in (0.37, 1.1 * np.abs(w).max()):
        ref = np.linalg.solve(A - sigma * np.eye(n), x)
        scale = np.abs(ref).max()
        y = B @ (Ys @ ((Ysi @ (B @ x)) / (w - sigma)))
        yd = Y @ ((Yi @ x) / (wd - sigma))
        e, ed = np.abs(y - ref).max() / scale, np.abs(yd - ref).max() / scale
        print(f"{name} sigma {sigma:.3g}: against LU, mirror form {e:.2e}, dense form {ed:.2e}", flush=True)
        assert e <= max(2.0 * ed, 1e-13)
