"""CPU checks of the screen's oracle (features and MLP) and of the host-side
parameter layout.  The oracle is the build's own restatement, pinned here by
closed forms and a finite-difference gradient check."""
import numpy as np
import pytest

import oracle.screen as OS
from centre_amd import screen


def test_param_layout_matches():
    assert OS.N_PARAMS == screen.N_PARAMS == 4817
    assert screen.OFF_MU == OS.OFF_MU and screen.OFF_SD == OS.OFF_SD


def random_params(seed):
    rng = np.random.default_rng(seed)
    feat = rng.normal(size=(200, 8))
    p = screen.init_params(feat, seed).astype(np.float64)
    p[OS.OFF_B1:OS.OFF_W2] = rng.normal(0, 0.3, 64)    # non-zero biases exercise every term
    p[OS.OFF_B2:OS.OFF_W3] = rng.normal(0, 0.3, 64)
    p[OS.OFF_B3] = 0.2
    return p, feat


def test_oracle_gradient_matches_finite_differences():
    p, feat = random_params(3)
    t = np.random.default_rng(4).normal(size=len(feat))
    _, g = OS.loss_grad(p, feat, t)
    rng = np.random.default_rng(5)
    trainable = np.arange(OS.OFF_MU)
    for q in list(rng.choice(trainable, 60, replace=False)) + [OS.OFF_B3, OS.OFF_W1, OS.OFF_W2]:
        h = 1e-6
        e = np.zeros_like(p); e[q] = h
        lp, _ = OS.loss_grad(p + e, feat, t)
        lm, _ = OS.loss_grad(p - e, feat, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[q]) <= 1e-5 * max(1.0, abs(fd)), (q, fd, g[q])
    assert np.all(g[OS.OFF_MU:] == 0)


def test_oracle_forward_is_the_mlp():
    p, feat = random_params(7)
    W1, b1, W2, b2, w3, b3, mu, sd = OS.unpack(p)
    f = feat[3]
    x = (f - mu) / sd
    h1 = [max(0.0, sum(W1[a, k] * x[k] for k in range(8)) + b1[a]) for a in range(64)]
    h2 = [max(0.0, sum(W2[a, b] * h1[b] for b in range(64)) + b2[a]) for a in range(64)]
    y = sum(w3[a] * h2[a] for a in range(64)) + b3
    assert OS.forward(p, feat[3:4])[0] == pytest.approx(y, rel=1e-12)


def test_oracle_features_closed_forms():
    rng = np.random.default_rng(11)
    nb, Fr = 10, 4000
    a = rng.integers(0, nb, Fr).astype(np.uint8)
    codes = np.stack([a, a, (nb - 1 - a).astype(np.uint8), rng.integers(0, nb, Fr).astype(np.uint8)])
    types = np.array([3, 3, 1, 2], np.uint8)
    ext = np.tile([0.0, 1.0, 0.5, 0.0], (4, 1))
    f = OS.features(codes, types, ext, nb, 1000)
    # pair order: (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
    assert f[0, 3] == pytest.approx(1.0, abs=1e-12)          # identical -> linear r = 1
    assert f[1, 3] == pytest.approx(-1.0, abs=1e-12)         # mirrored -> r = -1
    assert f[0, 2] == pytest.approx(1.0, abs=1e-12)          # identical -> circular r = 1
    assert abs(f[2, 3]) < 0.1 and abs(f[2, 2]) < 0.1         # independent -> small
    # MI of a DOF with itself on the subsample = its (ML, no Jacobian) entropy
    # over the bins merged to 8 (code * 8 // 10: bins 0-1 and 5-6 merge)
    sub = a[::4][:1000]
    merged = sub.astype(int) * 8 // nb
    assert sorted(set((np.arange(nb) * 8 // nb).tolist())) == list(range(8))
    p = np.bincount(merged, minlength=8) / 1000.0
    h = -sum(v * np.log(v) for v in p if v > 0)
    assert f[0, 4] == pytest.approx(h, rel=1e-10)
    assert f[0, 7] == 1.0 and f[2, 7] == 0.0
    assert f[2, 5] == pytest.approx(1.0) and f[2, 6] == pytest.approx(2 / 3)


def test_target_and_keep_mask():
    t = screen.target(np.array([0.0, 1e-3, 0.5]))
    assert t[0] == pytest.approx(np.log(screen.MI_FLOOR))
    assert t[1] == pytest.approx(np.log(1e-3), rel=1e-6)
    m = screen.keep_mask(t, 1e-3)
    assert m.tolist() == [False, True, True]
