"""GPU checks of the learned guess for the SR primitive recovery: the FP32 training
kernels against the CPU checker, the BF16 MFMA inference against the FP32 path, and the
learned-seed recovery against the reference iteration (CPU restatement) it must agree
with, including the evaluations it saves."""
import numpy as np
import pytest

import oracle_bind as ob
from athena_amd import gpu, ml
from athena_amd.mesh import DEFAULT_FLOOR

pytestmark = pytest.mark.gpu

NW, NP, NI = 4673, 4687, 6
GAM = ml.GAMMA_AD
FL = DEFAULT_FLOOR


def _random_params(rng, scale=0.3):
    P = np.zeros(NP, dtype=np.float32)
    P[:NW] = rng.normal(scale=scale, size=NW)
    P[NW + NI:NW + 2 * NI] = 1.0
    P[NW + 2 * NI + 1] = 1.0
    return P


def test_gpu_gradient_matches_oracle():
    rng = np.random.default_rng(7)
    P = _random_params(rng)
    x = rng.normal(size=(1000, NI)).astype(np.float32)
    t = rng.normal(size=1000).astype(np.float32)
    net = gpu.Mlp(P)
    g, loss = net.grad(x, t)
    go, lo = ob.mlp_grad(P.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    # FP32 accumulation over 1000 samples: relative error of the whole vector ~1e-6
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print("  gradient rel err %.3e, loss %.6e vs %.6e" % (rel, loss, lo))
    assert rel < 2e-5
    assert loss == pytest.approx(lo, rel=1e-5)
    for lo_, hi_ in ((0, 384), (384, 448), (448, 4544), (4544, 4608), (4608, 4672),
                     (4672, 4673)):   # every parameter group separately
        seg = np.linalg.norm(g[lo_:hi_] - go[lo_:hi_]) / np.linalg.norm(go[lo_:hi_])
        assert seg < 1e-4, (lo_, hi_, seg)


def test_gpu_adam_first_step():
    # Adam's bias-corrected first step is lr * g / (|g| + eps) (Kingma & Ba 2015, alg. 1)
    rng = np.random.default_rng(8)
    P = _random_params(rng)
    x = rng.normal(size=(512, NI)).astype(np.float32)
    t = rng.normal(size=512).astype(np.float32)
    net = gpu.Mlp(P)
    g, _ = net.grad(x, t)
    lr = 1e-3
    net.train(x, t, np.arange(512, dtype=np.int32), 512, lr)
    step = net.params()[:NW].astype(np.float64) - P[:NW]
    want = -lr * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(step, want, rtol=1e-3, atol=2e-9)
    assert np.array_equal(net.params()[NW:], P[NW:])   # normalisation untouched


def test_gpu_training_fits_a_smooth_function():
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, size=(20000, NI)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) * x[:, 1] + 0.5 * x[:, 2] ** 2 - x[:, 5]).astype(np.float32)
    P = _random_params(rng, 0.2)
    net = gpu.Mlp(P)
    hist = net.train(x, t, rng.integers(0, 20000, size=2000 * 256, dtype=np.int32), 256,
                     3e-3)
    print("  loss first %.3e last %.3e" % (hist[:20].mean(), hist[-20:].mean()))
    assert hist[-20:].mean() < 0.02 * hist[:20].mean()


def test_gpu_forward_fp32_matches_oracle():
    rng = np.random.default_rng(10)
    P = _random_params(rng)
    P[NW + 2 * NI] = 0.3
    P[NW + 2 * NI + 1] = 1.7
    x = rng.normal(size=(3000, NI)).astype(np.float32)
    y = gpu.Mlp(P).infer(x, bf16=False)
    yo = ob.mlp_forward(P.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(y, yo, rtol=1e-5, atol=1e-5)


def test_gpu_bf16_mfma_matches_fp32():
    # with the first layer in split bf16, the remaining error is the single-bf16 rounding
    # of W2 and of the hidden activations: ~5e-3 rms in log10 p (CPU emulation of the
    # rounding of W2 alone gives 3e-3); a wrong operand or accumulator layout in the MFMA
    # path gives O(1) errors
    net = ml.load_mlp()
    x, y, _, _ = ml.make_dataset(100_000, 777)
    xn = ml.normalise(net.params(), x)
    y32 = net.infer(xn, bf16=False)
    y16 = net.infer(xn, bf16=True)
    d = np.abs(y16 - y32)
    print("  |bf16 - fp32| rms %.2e max %.2e; fp32 rms err vs reference %.3e"
          % (np.sqrt(np.mean(d ** 2)), d.max(), np.sqrt(np.mean((y32 - y) ** 2))))
    assert np.sqrt(np.mean(d ** 2)) < 0.01 and d.max() < 0.15
    # the same on a ragged count (last wavefront partly empty) and a single cell
    for n in (1, 63, 65, 1000):
        np.testing.assert_array_equal(net.infer(xn[:n], bf16=True), y16[:n])


def test_gpu_features_match_oracle():
    cons, bcc, _ = ml.conserved_batch(20000, 12)
    x = gpu.ml_features(cons, bcc)
    # the device takes the logarithms in FP32 (v_log_f32, ~1 ulp)
    np.testing.assert_allclose(x, ob.mlp_features(cons, bcc), rtol=1e-6, atol=1e-6)


@pytest.fixture(scope="module")
def c2p_runs():
    net = ml.load_mlp()
    cons, bcc, prim = ml.conserved_batch(200_000, 4242)
    seed = np.zeros_like(cons)
    out = {m: gpu.sr_con2prim_modes(m, cons, seed, bcc, GAM, FL, FL, mlp=net)
           for m in (0, 1, 2)}
    return cons, bcc, prim, out


def test_learned_recovery_agrees_with_reference(c2p_runs):
    cons, bcc, prim, out = c2p_runs
    _, p0, it0, _, _ = out[0]
    # mode 0 is the reference iteration: the CPU restatement from the same cold seed
    # agrees to the last bits of the device's acos/cos/cbrt
    _, po, ito = ob.sr_con2prim(cons[:20000], np.zeros((20000, 5)), bcc[:20000], GAM, FL, FL)
    assert np.array_equal(it0[:20000] >= 0, ito >= 0)
    assert np.mean(it0[:20000] == ito) > 0.99
    assert np.all(np.abs(p0[:20000] - po) <= 1e-10 * np.abs(po))
    ok = it0 >= 0
    for mode in (1, 2):
        _, pm, itm, _, fb = out[mode]
        both = ok & (itm >= 0)
        assert both.sum() >= ok.sum() - 2
        # both solutions satisfy |p - g(p)| < 1e-12; g contracts, so they differ by a
        # few 1e-12 in p and relatively as little in the other primitives
        dp = np.abs(pm[both, 4] - p0[both, 4])
        assert dp.max() < 1e-10, dp.max()
        rel = np.abs(pm[both] - p0[both]) / (np.abs(p0[both]) + 1e-300)
        print("  mode %d: max |dp| %.2e, max rel diff %.2e" % (mode, dp.max(), rel.max()))
        assert np.all(rel[:, [0, 4]] < 1e-6)


def test_learned_answer_is_a_reference_fixed_point(c2p_runs):
    # the reference iteration seeded with the learned-mode answer finds it already
    # converged: its first update moves it by less than the 1e-12 tolerance
    cons, bcc, _, out = c2p_runs
    _, p2, it2, _, _ = out[2]
    sel = np.flatnonzero(it2 >= 0)[:20000]
    _, po, ito = ob.sr_con2prim(cons[sel], p2[sel], bcc[sel], GAM, FL, FL)
    assert np.all(ito == 0)
    assert np.max(np.abs(po[:, 4] - p2[sel, 4])) < 1e-12


def test_learned_guess_saves_evaluations(c2p_runs):
    _, _, _, out = c2p_runs
    ev = {m: out[m][3].astype(np.float64) for m in (0, 1, 2)}
    fb = out[2][4]
    print("  mean evaluations: reference %.3f, Newton cold %.3f, learned %.3f; "
          "fallback rate %.2e" % (ev[0].mean(), ev[1].mean(), ev[2].mean(), fb.mean()))
    assert ev[2].mean() < 0.75 * ev[0].mean()
    assert ev[2].mean() < 0.9 * ev[1].mean()
    assert fb.mean() < 1e-3
