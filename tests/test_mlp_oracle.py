"""CPU pins of the coupling-predictor oracle (oracle/mlp_oracle.py): the
analytic gradient against central finite differences, the Adam step against
its closed form on a quadratic, and bfloat16 rounding against known values."""
import numpy as np

import mlp_oracle as M


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    ni, no, n = 7, 3, 11
    p = rng.standard_normal(M.layout(ni, no)["np"]) * 0.3
    X = rng.standard_normal((n, ni))
    T = rng.standard_normal((n, no))
    g = M.grad(p, X, T)
    h = 1e-6
    idx = rng.choice(p.size, 60, replace=False)
    for q in idx:
        e = np.zeros_like(p)
        e[q] = h
        fd = (M.loss(p + e, X, T) - M.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(g[q])), q


def test_adam_closed_form():
    """Adam's first step moves every parameter by lr sign(g) (bias-corrected
    moments are g and g^2), whatever the gradient's size."""
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -1e-3, 40.0])
    p1, m, v = M.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p1, p - 0.01 * np.sign(g) * np.abs(g) / (np.abs(g) + 1e-8), rtol=1e-15)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -10, 3.14159265, -0.1, 1.0 + 3 * 2.0 ** -8])
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # tie to even
    assert r[2] == 1.0 + 2.0 ** -7           # above the tie: up
    assert r[5] == 1.0 + 2.0 ** -6           # tie between odd and even: to even (up)
    assert abs(r[3] - 3.140625) < 1e-15
    assert abs(r[4] + 0.10009765625) < 1e-15
