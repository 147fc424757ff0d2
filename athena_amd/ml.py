"""Learned initial guess for the SR primitive recovery: data, training and loading.

The network (``athena_amd/csrc/mlp.hpp``) maps six rotation- and scale-invariant
inputs of a cell's conserved state to ``y = log10(p_gas / (E - D))``.  Its training
pairs come from the reference's own recovery: conserved states are built from sampled
primitive states and the reference Newman-Hamlin iteration
(``adiabatic_mhd_sr.cpp:404-522``, mode 0 of the batched GPU recovery, which matches the
CPU restatement bit for bit) recovers the gas pressure that is the target.  Training is
FP32 Adam on the GPU by the library's own forward/backward kernels
(``ath_ml_train``); inference is BF16 MFMA inside the recovery kernel.

The sampled distribution (per cell, all independent):

* Lorentz factor log-uniform in [1, 10], isotropic direction;
* rest-mass density log-uniform in [1e-3, 1];
* magnetisation sigma = b^2 / rho (comoving field strength) log-uniform in [0.01, 10],
  field direction isotropic;
* temperature p / rho log-uniform in [1e-2, 10];
* adiabatic index 4/3.

Nothing here runs on the CPU as a substitute for the GPU: every call goes through the
C ABI and raises if the library is missing.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import gpu
from .mesh import DEFAULT_FLOOR
from .pgen.sr import prim2cons

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "sr_c2p_mlp.npz")
WEIGHTS_META = os.path.join(HERE, "data", "sr_c2p_mlp.json")

NI, NH = 6, 64
NW = gpu.NW
NP = NW + 2 * NI + 2
GAMMA_AD = 4.0 / 3.0

DIST = dict(lorentz=(1.0, 10.0), rho=(1.0e-3, 1.0), sigma=(1.0e-2, 10.0),
            temperature=(1.0e-2, 10.0), gamma=GAMMA_AD)


def _log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n))


def _isotropic(rng, n):
    v = rng.normal(size=(3, n))
    return v / np.linalg.norm(v, axis=0)


def sample_states(n, seed, dist=DIST):
    """Primitive states (5, n) = rho, u1, u2, u3, p_gas (u: spatial 4-velocity) and
    cell-centred fields (3, n) drawn from `dist`."""
    rng = np.random.default_rng(seed)
    lor = _log_uniform(rng, *dist["lorentz"], n)
    rho = _log_uniform(rng, *dist["rho"], n)
    sig = _log_uniform(rng, *dist["sigma"], n)
    temp = _log_uniform(rng, *dist["temperature"], n)
    vhat = _isotropic(rng, n)
    bhat = _isotropic(rng, n)
    vmag = np.sqrt(1.0 - 1.0 / lor ** 2)
    u = vhat * (lor * vmag)
    # comoving b^2 = B^2 / Gamma^2 + (B.v)^2 for lab-frame B = |B| bhat
    bv = np.sum(bhat * vhat, axis=0) * vmag
    bmag = np.sqrt(sig * rho / (1.0 / lor ** 2 + bv ** 2))
    prim = np.array([rho, u[0], u[1], u[2], rho * temp])
    return prim, bhat * bmag


def conserved_batch(n, seed, dist=DIST):
    """(cons (n,5), bcc (n,3), prim (n,5)) of `n` sampled states."""
    prim, bcc = sample_states(n, seed, dist)
    cons = prim2cons(prim, bcc, dist["gamma"])
    return (np.ascontiguousarray(cons.T), np.ascontiguousarray(bcc.T),
            np.ascontiguousarray(prim.T))


def reference_pressure(cons, bcc, gamma=GAMMA_AD, floor=DEFAULT_FLOOR):
    """Gas pressure recovered by the reference iteration from a cold seed (the
    reference's seed is the output register; zero there means the floor), and whether
    the first solve converged."""
    seed = np.zeros_like(cons)
    _, prim, it, _, _ = gpu.sr_con2prim_modes(0, cons, seed, bcc, gamma, floor, floor)
    return prim[:, 4], it >= 0


def targets(cons, p):
    return np.log10(p / (cons[:, 4] - cons[:, 0]))


def make_dataset(n, seed, gamma=GAMMA_AD):
    """Raw features (m,6), targets (m) and conserved/field arrays of the states whose
    reference solve converged to the sampled pressure (relative 1e-6)."""
    cons, bcc, prim = conserved_batch(n, seed)
    p, ok = reference_pressure(cons, bcc, gamma)
    ok &= np.abs(p - prim[:, 4]) <= 1.0e-6 * prim[:, 4]
    cons, bcc, p = cons[ok], bcc[ok], p[ok]
    x = gpu.ml_features(cons, bcc)
    return x, targets(cons, p), cons, bcc


def init_params(x, y, seed):
    """Glorot-uniform weights, zero biases, and the input/output normalisation."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NP, dtype=np.float32)
    o = 0

    def glorot(fan_out, fan_in):
        a = math.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-a, a, fan_out * fan_in).astype(np.float32)

    P[o:o + NH * NI] = glorot(NH, NI); o += NH * NI
    o += NH                                                  # b1
    P[o:o + NH * NH] = glorot(NH, NH); o += NH * NH
    o += NH                                                  # b2
    P[o:o + NH] = glorot(1, NH); o += NH
    o += 1                                                   # b3
    assert o == NW
    xm, xs = x.mean(axis=0), x.std(axis=0)
    P[NW:NW + NI] = xm
    P[NW + NI:NW + 2 * NI] = 1.0 / np.maximum(xs, 1.0e-6)
    P[NW + 2 * NI] = y.mean()
    P[NW + 2 * NI + 1] = max(float(y.std()), 1.0e-6)
    return P


def normalise(P, x, y=None):
    xn = ((x - P[NW:NW + NI]) * P[NW + NI:NW + 2 * NI]).astype(np.float32)
    if y is None:
        return xn
    return xn, ((y - P[NW + 2 * NI]) / P[NW + 2 * NI + 1]).astype(np.float32)


def train(n_train=4_000_000, n_val=200_000, steps=40_000, batch=8192, lr0=3e-3,
          lr1=3e-5, chunk=1000, seed=20261202, log=print):
    """Trains a network on the GPU; returns (Mlp, params, info)."""
    x, y, _, _ = make_dataset(n_train, seed)
    xv, yv, _, _ = make_dataset(n_val, seed + 1)
    P = init_params(x, y, seed + 2)
    xn, yn = normalise(P, x, y)
    xvn = normalise(P, xv)
    net = gpu.Mlp(P)
    rng = np.random.default_rng(seed + 3)
    hist = []
    nchunks = (steps + chunk - 1) // chunk
    for c in range(nchunks):
        lr = lr0 * (lr1 / lr0) ** (c / max(1, nchunks - 1))
        s = min(chunk, steps - c * chunk)
        order = rng.integers(0, yn.size, size=s * batch, dtype=np.int32)
        h = net.train(xn, yn, order, batch, lr)
        hist.append(h)
        if log:
            yp = net.infer(xvn, bf16=False)
            log("chunk %d/%d lr %.2e loss %.3e val rmse(log10 p) %.4f"
                % (c + 1, nchunks, lr, float(h[-50:].mean()),
                   float(np.sqrt(np.mean((yp - yv) ** 2)))))
    P = net.params()
    yp32 = net.infer(xvn, bf16=False)
    yp16 = net.infer(xvn, bf16=True)
    info = dict(n_train=int(yn.size), n_val=int(yv.size), steps=steps, batch=batch,
                lr0=lr0, lr1=lr1, seed=seed, dist={k: v for k, v in DIST.items()},
                val_rmse_fp32=float(np.sqrt(np.mean((yp32 - yv) ** 2))),
                val_rmse_bf16=float(np.sqrt(np.mean((yp16 - yv) ** 2))),
                val_maxerr_bf16=float(np.max(np.abs(yp16 - yv))))
    return net, P, np.concatenate(hist), info


def save(P, loss_hist, info, path=WEIGHTS, meta=WEIGHTS_META):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=np.asarray(P, dtype=np.float32),
             loss=np.asarray(loss_hist, dtype=np.float32)[::100])
    with open(meta, "w") as f:
        json.dump(info, f, indent=1, sort_keys=True)


def load_params(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def load_mlp(path=WEIGHTS):
    """The committed trained network on the device."""
    return gpu.Mlp(load_params(path))


if __name__ == "__main__":
    import argparse
    import time

    ap = argparse.ArgumentParser(description="train the SR recovery guess network")
    ap.add_argument("--n-train", type=int, default=4_000_000)
    ap.add_argument("--steps", type=int, default=40_000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr0", type=float, default=3e-3)
    ap.add_argument("--lr1", type=float, default=3e-5)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    t0 = time.time()
    _, P, hist, info = train(n_train=a.n_train, steps=a.steps, batch=a.batch, lr0=a.lr0,
                             lr1=a.lr1, chunk=max(1000, a.steps // 60),
                             log=lambda s: print(s, flush=True))
    info["train_seconds"] = time.time() - t0
    save(P, hist, info, a.out, os.path.splitext(a.out)[0] + ".json")
    print(json.dumps(info), flush=True)
