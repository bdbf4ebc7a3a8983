"""The learned pair screen on the GPU against its CPU oracle (oracle/screen.py):
features, MLP inference and the analytic gradient, then training and what
the trained network gains over screening with the subsample MI alone."""
import numpy as np
import pytest

import oracle.screen as OS
from centre_amd import hip, screen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def synth_codes(ctx, seed, nbd, nan, ndh, F, nb):
    bat = hip.synth_bat(seed, nbd, nan, ndh, F)
    D = nbd + nan + ndh
    d_bat = ctx.to_device(bat)
    d_codes = ctx.alloc((D, F), np.uint8)
    d_ext = ctx.alloc((D, 4), np.float64)
    types = np.array([1] * nbd + [2] * nan + [3] * ndh, np.uint8)
    col = 0
    for t, n in ((1, nbd), (2, nan), (3, ndh)):
        if n:
            ctx.dev_discretize_hist(d_bat.ptr + 8 * col, D, n, F, t, [nb], nb, True,
                                    d_codes.ptr + col * F, d_ext.ptr + 32 * col)
        col += n
    ctx.sync()
    d_bat.free()
    return d_codes, d_ext, types, ctx.to_device(types)


def test_features_match_oracle(ctx):
    nb, F, n_sub = 20, 8000, 1000
    d_codes, d_ext, types, d_types = synth_codes(ctx, 0xCE27E + 40, 8, 8, 8, F, nb)
    print("features D=24 F=8000 n_sub=1000", flush=True)
    d_feat = screen.features(ctx, d_codes, F, d_types, d_ext, nb, n_sub)
    got = d_feat.numpy().astype(np.float64)
    want = OS.features(d_codes.numpy(), types, d_ext.numpy(), nb, n_sub)
    np.testing.assert_allclose(got[:, :2], want[:, :2], rtol=1e-6)
    np.testing.assert_allclose(got[:, 2:4], want[:, 2:4], atol=2e-5)
    np.testing.assert_allclose(got[:, 4], want[:, 4], rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(got[:, 5:], want[:, 5:].astype(np.float32))


def test_forward_and_gradient_match_oracle(ctx):
    rng = np.random.default_rng(2)
    n = 1000                                   # not a multiple of the 64-sample block
    feat = rng.normal(size=(n, 8)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    p = screen.init_params(feat, 9)
    p[OS.OFF_B1:OS.OFF_W2] = rng.normal(0, 0.3, 64)
    p[OS.OFF_B2:OS.OFF_W3] = rng.normal(0, 0.3, 64)
    d_p, d_f, d_t = ctx.to_device(p), ctx.to_device(feat), ctx.to_device(t)
    d_y = ctx.alloc(n, np.float32)
    print("mlp forward / grad n=1000", flush=True)
    ctx.mlp_forward(d_p.ptr, d_f.ptr, n, d_y.ptr)
    want_y = OS.forward(p.astype(np.float64), feat.astype(np.float64))
    np.testing.assert_allclose(d_y.numpy(), want_y, rtol=1e-4, atol=1e-5)
    d_g = ctx.alloc(screen.N_PARAMS, np.float32)
    loss = ctx.mlp_loss_grad(d_p.ptr, d_f.ptr, d_t.ptr, n, d_g.ptr)
    want_l, want_g = OS.loss_grad(p.astype(np.float64), feat.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(want_l, rel=1e-4)
    g = d_g.numpy().astype(np.float64)
    scale = np.abs(want_g).max()
    np.testing.assert_allclose(g, want_g, rtol=1e-3, atol=1e-5 * scale)
    # the normalisation entries are not trained
    assert np.all(g[OS.OFF_MU:] == 0)


def plain_mi(ctx, d_codes, d_types, d_ext, D, F, nb, S=10):
    P = D * (D - 1) // 2
    plan = hip.PairPlan(ctx, D, nb)
    d_s1 = ctx.alloc((D, 1, S), np.float64)
    d_s2 = ctx.alloc((P, 1, S), np.float64)
    d_mi = ctx.alloc((P, S), np.float64)
    ctx.dev_pair_entropy(plan, d_codes.ptr, F, F, d_types.ptr, d_ext.ptr, S, [0], True, False,
                         d_s1.ptr, d_s2.ptr, d_mi.ptr)
    plan.close()
    return d_mi.numpy()[:, S - 1]


def test_training_and_gain_over_subsample_mi(ctx):
    """Train on two systems, evaluate on a third: the network's predictions
    of the full-pass MI beat the subsample MI used alone, both in squared
    error of log MI and in how many of the truly strongest pairs a screen
    keeping 20% of pairs retains."""
    nb, F, n_sub = 50, 40000, 2000
    feats, targets = [], []
    for k in range(3):
        d_codes, d_ext, types, d_types = synth_codes(ctx, 0xCE27E + 100 + k, 60, 60, 60, F, nb)
        D = len(types)
        print(f"system {k}: D={D} F={F}", flush=True)
        d_feat = screen.features(ctx, d_codes, F, d_types, d_ext, nb, n_sub)
        mi = plain_mi(ctx, d_codes, d_types, d_ext, D, F, nb)
        feats.append(d_feat.numpy())
        targets.append(screen.target(mi))
    ftr = np.concatenate(feats[:2])
    ttr = np.concatenate(targets[:2])
    net = screen.Screen(ctx, screen.init_params(ftr, 1))
    d_f, d_t = ctx.to_device(ftr), ctx.to_device(ttr)
    hist = net.train(d_f, d_t, len(ftr), epochs=30, batch=1024, lr=3e-3, seed=5)
    print("train loss by epoch", hist.round(4).tolist(), flush=True)
    assert hist[-1] < 0.5 * hist[0]
    # held-out system
    fte, tte = feats[2], targets[2].astype(np.float64)
    pred = net.predict(ctx.to_device(fte), len(fte)).numpy().astype(np.float64)
    mse_net = np.mean((pred - tte) ** 2)
    # the method without the network: subsample MI, best affine map to the target
    base = np.log(np.maximum(fte[:, 4].astype(np.float64), 1e-5))
    A = np.stack([base, np.ones_like(base)], 1)
    coef, *_ = np.linalg.lstsq(A, tte, rcond=None)
    mse_base = np.mean((A @ coef - tte) ** 2)
    k = len(tte) // 5
    top = set(np.argsort(-tte)[:k // 2])
    rec_net = len(top & set(np.argsort(-pred)[:k])) / len(top)
    rec_base = len(top & set(np.argsort(-base)[:k])) / len(top)
    print(f"held-out: mse net {mse_net:.4f} base {mse_base:.4f}; recall net {rec_net:.3f} "
          f"base {rec_base:.3f}", flush=True)
    assert mse_net < mse_base
    assert rec_net >= rec_base


def test_tile_plan_and_tile_selection(ctx):
    """A tile-subset plan writes exactly the pairs of its tiles, with the
    values of the all-pairs plan; tile selection keeps the tiles holding a
    pair at or above the cutoff plus each DOF's best predicted partners."""
    nb, F, S = 50, 6000, 3
    d_codes, d_ext, types, d_types = synth_codes(ctx, 0xCE27E + 60, 9, 8, 9, F, nb)
    D = len(types)
    P = D * (D - 1) // 2
    R = int(hip.lib().centre_pair_dofs_per_tile(nb))
    rng = np.random.default_rng(3)
    pred = rng.normal(-5, 1, P).astype(np.float32)
    d_pred = ctx.to_device(pred)
    cut = float(np.exp(np.quantile(pred, 0.9)))
    print("tile selection D=26", flush=True)
    tiles = ctx.screen_tiles(d_pred.ptr, D, nb, cut, 2)
    iu, ju = np.triu_indices(D, 1)
    nblk = (D + R - 1) // R
    want = np.zeros((nblk, nblk), bool)
    np.logical_or.at(want, (iu // R, ju // R), pred >= np.float32(np.log(cut)))
    M = np.full((D, D), -np.inf)
    M[iu, ju] = pred
    M[ju, iu] = pred
    for i in range(D):
        for j in np.argsort(-M[i])[:2]:
            a, b = min(i, j), max(i, j)
            want[a // R, b // R] = True
    got = np.zeros_like(want)
    got[tiles[:, 0], tiles[:, 1]] = True
    np.testing.assert_array_equal(got, want)
    assert np.all(tiles[:, 0] <= tiles[:, 1])
    # tile plan against the all-pairs plan
    outs = []
    for t in (None, tiles):
        plan = hip.PairPlan(ctx, D, nb, tiles=t) if t is not None else hip.PairPlan(ctx, D, nb)
        d_s1 = ctx.alloc((D, 1, S), np.float64)
        d_s2 = ctx.alloc((P, 1, S), np.float64)
        d_mi = ctx.alloc((P, S), np.float64)
        ctx.copy_d2d(d_s2.ptr, ctx.to_device(np.full((P, 1, S), np.nan)).ptr, 8 * P * S)
        ctx.dev_pair_entropy(plan, d_codes.ptr, F, F, d_types.ptr, d_ext.ptr, S, [0], True,
                             False, d_s1.ptr, d_s2.ptr, d_mi.ptr)
        plan.close()
        outs.append(d_s2.numpy()[:, 0, :])
    full, part = outs
    inside = got[iu // R, ju // R]
    np.testing.assert_array_equal(part[inside], full[inside])
    assert np.all(np.isnan(part[~inside]))
