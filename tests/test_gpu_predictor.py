"""The coupling predictor's network on the GPU (fribo_mlp_*) against the FP64
oracle (oracle/mlp_oracle.py), and what the trained network gains in a
strongly coupled run.

Bars: gradients of the FP32 kernels within 1e-4 of the largest oracle
gradient component, the loss within 1e-5 relative, five Adam steps within
1e-4 of the oracle's parameters, BF16 inference within 2e-3 of the oracle's
BF16 forward (the hidden activations are rounded to 8 mantissa bits, and a
rounding flip moves a value by 2^-8 relative); the strongly coupled run with
the learned first iterate converges to the same loads as the plain one
(within 1e-6 of their largest value, the coupling tolerance being 1e-8 on
each step's load change) in fewer sub-iterations.
"""
import numpy as np
import pytest

import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _data(ni, no, n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, ni)).astype(np.float32), rng.standard_normal((n, no)).astype(np.float32)


def test_mlp_gradients_match_the_oracle():
    from fribois2d_amd.gpu import Mlp
    ni, no = 27, 3
    net = Mlp(ni, no, seed=11)
    p = net.params().astype(np.float64)
    X, T = _data(ni, no, 300, 1)
    g, loss = net.grad(X, T)
    gr = M.grad(p, X, T)
    lr = M.loss(p, X, T)
    err = float(np.max(np.abs(g - gr)) / np.max(np.abs(gr)))
    print(f"gradient: max |gpu - oracle| / max |oracle| = {err:.2e}; loss {loss:.6e} vs {lr:.6e}", flush=True)
    assert err <= 1e-4
    assert abs(loss - lr) <= 1e-5 * lr
    net.close()


def test_mlp_adam_steps_match_the_oracle():
    from fribois2d_amd.gpu import Mlp
    ni, no = 27, 3
    net = Mlp(ni, no, seed=12)
    p = net.params().astype(np.float64)
    X, T = _data(ni, no, 128, 2)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in range(1, 6):
        net.train(X, T, 1, 128, 1e-3)
        p, m, v = M.adam(p, M.grad(p, X, T), m, v, t, 1e-3)
    err = float(np.max(np.abs(net.params() - p)))
    print(f"5 Adam steps: max |gpu - oracle| = {err:.2e}", flush=True)
    assert err <= 1e-4
    net.close()


def test_mlp_bf16_inference():
    from fribois2d_amd.gpu import Mlp
    ni, no = 27, 3
    net = Mlp(ni, no, seed=13)
    X, _ = _data(ni, no, 64, 3)
    Y = net.predict(X)
    Yr = M.forward_bf16(net.params(), X, no)
    err = float(np.max(np.abs(Y - Yr)) / np.max(np.abs(Yr)))
    print(f"BF16 inference: max |gpu - oracle| / max = {err:.2e}", flush=True)
    assert err <= 2e-3
    net.close()


def test_learned_first_iterate_saves_subiterations():
    from fribois2d_amd import cases, predictor as P
    from fribois2d_amd.gpu import FluidStep
    n3m, n2m, nsteps = 128, 96, 30
    c, p, _ = P.coupled_case(0, n3m, n2m)
    g = FluidStep(p, c.rc, c.zz)
    X, T = [], []
    for run in range(10):
        c, p, body = P.coupled_case(run, n3m, n2m)
        g.set_state(cases.initial_state(p, c.seed))
        g.set_markers(*body.markers())
        hist, its = P.run_coupled(g, body, nsteps, p.dt0)
        x, t = P.samples(hist)
        X.append(x)
        T.append(t)
    pred = P.Predictor.train(np.concatenate(X), np.concatenate(T), epochs=200)
    base_its, pred_its = [], []
    for run in (100, 101, 102):
        out = []
        for guess in (None, pred):
            c, p, body = P.coupled_case(run, n3m, n2m)
            g.set_state(cases.initial_state(p, c.seed))
            g.set_markers(*body.markers())
            out.append(P.run_coupled(g, body, nsteps, p.dt0, guess=guess))
        (hb, ib), (hp, ip) = out
        hb, hp = np.array(hb), np.array(hp)
        scale = np.max(np.abs(hb), axis=0)
        diff = float(np.max(np.abs(hp - hb) / scale))
        print(f"run {run}: sub-iterations plain {ib[P.HIST:].mean():.2f}, learned {ip[P.HIST:].mean():.2f}; "
              f"histories differ by {diff:.1e}", flush=True)
        assert diff <= 1e-6
        base_its.append(ib[P.HIST:])
        pred_its.append(ip[P.HIST:])
    print(f"training loss {pred.loss:.3e}", flush=True)
    assert np.mean(np.concatenate(pred_its)) < np.mean(np.concatenate(base_its))
    g.close()
