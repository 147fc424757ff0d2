"""CPU checks of the learned-guess network's checker and host logic.

The network has no counterpart in the reference; its CPU checker (oracle/mlp_oracle.cpp)
is pinned here by central finite differences of the loss (gradient), a one-path closed
form (forward pass) and the invariances the six features are built to have."""
import json
import math
import os

import numpy as np
import pytest

import oracle_bind as ob
from athena_amd import ml

NW = 4673
NP = 4687
NI, NH = 6, 64
OB1, OW2, OB2, OW3, OB3 = 384, 448, 4544, 4608, 4672


def _random_params(rng, scale=0.4):
    P = np.zeros(NP)
    P[:NW] = rng.normal(scale=scale, size=NW)
    P[NW + NI:NW + 2 * NI] = 1.0
    P[NW + 2 * NI + 1] = 1.0
    return P


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    P = _random_params(rng)
    x = rng.normal(size=(7, NI))
    t = rng.normal(size=7)
    g, loss = ob.mlp_grad(P, x, t)
    # every bias and output weight, and a random sample of the matrix weights
    idx = np.concatenate([np.arange(OB1, OW2), np.arange(OB2, NW),
                          rng.choice(OB1, 40, replace=False),
                          rng.choice(np.arange(OW2, OB2), 80, replace=False)])
    h = 1e-6
    fd = np.empty(idx.size)
    for n, q in enumerate(idx):
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += h
        Pm[q] -= h
        fd[n] = (ob.mlp_grad(Pp, x, t)[1] - ob.mlp_grad(Pm, x, t)[1]) / (2 * h)
    err = np.abs(fd - g[idx])
    assert np.all(err <= 1e-7 + 1e-6 * np.abs(fd)), err.max()
    # and the loss is the mean squared error of the forward pass
    y = ob.mlp_forward(P, x)
    assert loss == pytest.approx(np.mean((y - t) ** 2), rel=1e-13)


def test_oracle_forward_single_path_closed_form():
    # one active path x_k -> h1_a -> h2_c -> y
    P = np.zeros(NP)
    a, c, k = 5, 17, 3
    w1, b1, w2, b2, w3, b3 = 0.7, -0.2, 1.3, 0.1, -0.9, 0.25
    P[a * NI + k] = w1
    P[OB1 + a] = b1
    P[OW2 + c * NH + a] = w2
    P[OB2:OB2 + NH] = b2
    P[OW3 + c] = w3
    P[OB3] = b3
    ym, ys = 0.5, 2.0
    P[NW + 2 * NI] = ym
    P[NW + 2 * NI + 1] = ys
    x = np.zeros((3, NI))
    x[:, k] = [-1.5, 0.0, 2.0]
    y = ob.mlp_forward(P, x)
    want = (b3 + w3 * np.tanh(b2 + w2 * np.tanh(b1 + w1 * x[:, k]))) * ys + ym
    np.testing.assert_allclose(y, want, rtol=1e-14)


def _random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.diag(r))


def test_oracle_features_invariances():
    cons, bcc, _ = ml.conserved_batch(64, 11)
    x = ob.mlp_features(cons, bcc)
    rng = np.random.default_rng(2)
    R = _random_rotation(rng)
    c2 = cons.copy()
    c2[:, 1:4] = cons[:, 1:4] @ R.T
    np.testing.assert_allclose(ob.mlp_features(c2, bcc @ R.T), x, rtol=1e-12, atol=1e-12)
    lam = 37.5   # D, E, M scale with lambda, B with sqrt(lambda)
    np.testing.assert_allclose(ob.mlp_features(cons * lam, bcc * math.sqrt(lam)), x,
                               rtol=1e-12, atol=1e-12)


def test_oracle_features_special_cases():
    # D=1, M=(0.3,0,0), E=2, B along M -> cos^2 = 1; B across M -> 0; B=0 -> floors
    cons = np.array([[1.0, 0.3, 0.0, 0.0, 2.0]] * 3)
    bcc = np.array([[0.5, 0.0, 0.0], [0.0, 0.0, 0.5], [0.0, 0.0, 0.0]])
    x = ob.mlp_features(cons, bcc)
    np.testing.assert_allclose(x[:, 0], 0.0, atol=1e-15)                 # log10(E/D - 1)
    np.testing.assert_allclose(x[:, 1], 0.15, rtol=1e-15)                # |M|/E
    np.testing.assert_allclose(x[:, 2], math.log10(1 - 0.15 ** 2), rtol=1e-14)
    np.testing.assert_allclose(x[:2, 3], math.log10(0.25 / 2 + 1e-12), rtol=1e-14)
    np.testing.assert_allclose(x[:, 4], [1.0, 0.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(x[2, 3], -12.0, rtol=1e-14)
    np.testing.assert_allclose(x[:, 5], math.log10(math.sqrt(4 - 0.09) - 1), rtol=1e-14)
    # at rest, unmagnetised and cold: E = D, so both energy inputs hit their floor
    x0 = ob.mlp_features(np.array([[1.0, 0.0, 0.0, 0.0, 1.0]]), np.zeros((1, 3)))
    np.testing.assert_allclose(x0[0, [0, 1, 2, 5]], [-12.0, 0.0, 0.0, -12.0], atol=1e-14)


def test_sampled_states_follow_the_distribution():
    prim, b = ml.sample_states(50_000, 3)
    rho, u1, u2, u3, p = prim
    u0 = np.sqrt(1 + u1 ** 2 + u2 ** 2 + u3 ** 2)
    b0 = b[0] * u1 + b[1] * u2 + b[2] * u3
    bi = [(b[i] + b0 * u) / u0 for i, u in enumerate((u1, u2, u3))]
    sigma = (-(b0 ** 2) + sum(v ** 2 for v in bi)) / rho
    for v, (lo, hi) in ((u0, ml.DIST["lorentz"]), (rho, ml.DIST["rho"]),
                        (sigma, ml.DIST["sigma"]), (p / rho, ml.DIST["temperature"])):
        assert v.min() >= lo * (1 - 1e-9) and v.max() <= hi * (1 + 1e-9)
        # log-uniform: the median sits at the geometric mean
        assert abs(math.log(np.median(v)) - 0.5 * math.log(lo * hi)) < 0.05 * math.log(hi / lo)


def test_committed_weights_are_complete():
    assert os.path.exists(ml.WEIGHTS), "trained weights are missing"
    P = ml.load_params()
    assert P.shape == (NP,) and np.all(np.isfinite(P))
    with open(ml.WEIGHTS_META) as f:
        info = json.load(f)
    assert info["n_train"] >= 3_000_000
    assert info["val_rmse_bf16"] < 0.05
