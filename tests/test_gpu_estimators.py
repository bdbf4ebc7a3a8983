"""Chao-Shen and James-Stein (and the count output) through the grouped pass:
the pair kernel materialises the counts of a group of row blocks at a time,
then every estimator, MI and the record-order counts come from them on the
device.  Checked against the oracle's entropy2D (Estimators.cpp:391-651) and
binData2D on the same codes.  Bars as in test_gpu_parity: entropies 1e-10
relative, MI 1e-10 absolute, counts bit-exact."""
import numpy as np
import pytest

import oracle as O
from centre_amd import hip

pytestmark = pytest.mark.gpu

REL = 1e-10


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def coupled_codes(seed, D, F, nb):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
    codes[1::2] = (codes[0::2][: len(codes[1::2])] + rng.integers(0, 3, size=F)) % nb
    types = np.array([1, 2, 3] * (D // 3 + 1), np.uint8)[:D]
    ext = np.zeros((D, 4))
    ext[:, 0] = rng.uniform(0.5, 1.0, D)
    ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
    return codes, types, ext


def test_chao_shen_james_stein_at_headline_size():
    """600 DOF x 100,000 frames, 50 bins, 10 steps, all 179,700 pairs with
    ests (CS, JS, ML): 18 GB of counts, taken in groups of at most 2 GiB.
    Every S1 and an evenly strided sample of 60 pairs against the oracle; MI
    of CS (estimator 0) from the device's own S1."""
    nbd = nan = ndh = 200
    D, F, nb, S = 600, 100000, 50, 10
    P = D * (D - 1) // 2
    ests = (2, 3, 0)
    bat = hip.synth_bat(0xCE27E + 1, nbd, nan, ndh, F)
    types = np.array([1] * nbd + [2] * nan + [3] * ndh, np.uint8)
    print(f"CS/JS headline size D={D} F={F} nb={nb} S={S} ests={ests}", flush=True)
    with hip.Context(0) as ctx:
        d_bat = ctx.to_device(bat)
        d_codes = ctx.alloc((D, F), np.uint8)
        d_ext = ctx.alloc((D, 4), np.float64)
        d_types = ctx.to_device(types)
        col = 0
        for t, n in ((1, nbd), (2, nan), (3, ndh)):
            ctx.dev_discretize_hist(d_bat.ptr + 8 * col, D, n, F, t, [nb], nb, True,
                                    d_codes.ptr + col * F, d_ext.ptr + 32 * col)
            col += n
        ctx.sync()
        d_bat.free()
        plan = hip.PairPlan(ctx, D, nb)
        d_s1 = ctx.alloc((D, len(ests), S), np.float64)
        d_s2 = ctx.alloc((P, len(ests), S), np.float64)
        d_mi = ctx.alloc((P, S), np.float64)
        ctx.dev_pair_entropy(plan, d_codes.ptr, F, F, d_types.ptr, d_ext.ptr, S, ests, True,
                             False, d_s1.ptr, d_s2.ptr, d_mi.ptr)
        print(f"grouped pass {ctx.last_pair_kernel_ms():.1f} ms", flush=True)
        s1, s2, mi = d_s1.numpy(), d_s2.numpy(), d_mi.numpy()
        codes, ext = d_codes.numpy(), d_ext.numpy()
        plan.close()
    assert np.all(np.isfinite(s2)) and np.all(np.isfinite(mi))
    iu, ju = np.triu_indices(D, 1)
    sample = np.arange(0, P, P // 60)[:60]
    pairs = np.stack([iu[sample], ju[sample]], 1)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb], ests)
    assert rel_err(s2[sample], w2[:, :, 0, :]) < REL
    w1 = O.dof_entropies(codes, types, ext, S, [nb], ests)
    assert rel_err(s1, w1[:, :, 0, :]) < REL
    assert np.max(np.abs(mi - (s1[iu, 0] + s1[ju, 0] - s2[:, 0]))) < 1e-10


@pytest.mark.parametrize("nb,D,F,S", [(50, 40, 6000, 3), (7, 90, 5000, 2), (255, 6, 3000, 2)])
def test_explicit_records_all_estimators(nb, D, F, S):
    """Explicit lists (duplicates and reversed records) through the grouped
    pass, every estimator in one call; counts in each record's own order."""
    codes, types, ext = coupled_codes(nb + D, D, F, nb)
    rng = np.random.default_rng(D)
    pairs = rng.integers(0, D, size=(3 * D, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    pairs = np.concatenate([pairs, pairs[:7, ::-1], pairs[:3]]).astype(np.uint32)
    P = len(pairs)
    ests = (3, 1, 2, 0)
    with hip.Context(0) as ctx:
        plan = hip.PairPlan(ctx, D, nb, pairs)
        dc, dt, dx = ctx.to_device(codes), ctx.to_device(types), ctx.to_device(ext)
        d_s2 = ctx.alloc((P, 4, S), np.float64)
        d_mi = ctx.alloc((P, S), np.float64)
        d_s1 = ctx.alloc((D, 4, S), np.float64)
        cnt = ctx.alloc((P, S, nb * nb), np.uint32)
        print(f"explicit grouped nb={nb} D={D} P={P}", flush=True)
        ctx.dev_pair_entropy(plan, dc.ptr, F, F, dt.ptr, dx.ptr, S, ests, True, False,
                             d_s1.ptr, d_s2.ptr, d_mi.ptr, cnt.ptr)
        s1, s2, mi, c = d_s1.numpy(), d_s2.numpy(), d_mi.numpy(), cnt.numpy()
        plan.close()
    w1 = O.dof_entropies(codes, types, ext, S, [nb], ests)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb], ests)
    assert rel_err(s2, w2[:, :, 0, :]) < REL
    mi_w = w1[pairs[:, 0], 0, 0, :] + w1[pairs[:, 1], 0, 0, :] - w2[:, 0, 0, :]
    assert np.max(np.abs(mi - mi_w)) < 1e-10
    want = O.pair_counts(codes, nb, pairs, S)
    assert np.array_equal(c.reshape(P, S, nb, nb), want.astype(np.uint32))


def test_all_pairs_count_output():
    nb, D, F, S = 30, 25, 4000, 4
    codes, types, ext = coupled_codes(3, D, F, nb)
    P = D * (D - 1) // 2
    with hip.Context(0) as ctx:
        plan = hip.PairPlan(ctx, D, nb)
        dc, dt, dx = ctx.to_device(codes), ctx.to_device(types), ctx.to_device(ext)
        d_s2 = ctx.alloc((P, 1, S), np.float64)
        cnt = ctx.alloc((P, S, nb * nb), np.uint32)
        ctx.dev_pair_entropy(plan, dc.ptr, F, F, dt.ptr, dx.ptr, S, [0], True, False,
                             0, d_s2.ptr, 0, cnt.ptr)
        s2, c = d_s2.numpy(), cnt.numpy()
        plan.close()
    iu, ju = np.triu_indices(D, 1)
    pairs = np.stack([iu, ju], 1)
    assert np.array_equal(c.reshape(P, S, nb, nb), O.pair_counts(codes, nb, pairs, S).astype(np.uint32))
    assert rel_err(s2, O.pair_entropies(codes, types, ext, pairs, S, [nb])[:, :, 0, :]) < REL


@pytest.mark.parametrize("ests", [(2,), (0, 2), (3, 1)])
def test_tile_subset_plan_non_fused_writes_only_listed_pairs(ests):
    """A tile-subset plan with Chao-Shen / James-Stein writes S2 and MI of the
    pairs inside its tiles only; the rest of the caller's buffers stays as it
    was (NaN here), and the written pairs equal the oracle's."""
    nb, D, F, S = 20, 50, 3000, 3
    codes, types, ext = coupled_codes(17, D, F, nb)
    R = int(hip.lib().centre_pair_dofs_per_tile(nb))
    nblk = (D + R - 1) // R
    assert nblk >= 4
    tiles = np.array([[0, 0], [0, 3], [1, 2], [nblk - 1, nblk - 1], [1, 2]], np.int32)
    P = D * (D - 1) // 2
    nest = len(ests)
    with hip.Context(0) as ctx:
        plan = hip.PairPlan(ctx, D, nb, tiles=tiles)
        dc, dt, dx = ctx.to_device(codes), ctx.to_device(types), ctx.to_device(ext)
        d_s2 = ctx.to_device(np.full((P, nest, S), np.nan))
        d_mi = ctx.to_device(np.full((P, S), np.nan))
        print(f"tile subset ests={ests}", flush=True)
        ctx.dev_pair_entropy(plan, dc.ptr, F, F, dt.ptr, dx.ptr, S, ests, True, False,
                             0, d_s2.ptr, d_mi.ptr)
        s2, mi = d_s2.numpy(), d_mi.numpy()
        plan.close()
    iu, ju = np.triu_indices(D, 1)
    inside = np.zeros(P, bool)
    for I, J in tiles:
        inside |= (iu // R == I) & (ju // R == J)
    assert inside.sum() > 0 and (~inside).sum() > 0
    assert np.all(np.isnan(s2[~inside])) and np.all(np.isnan(mi[~inside]))
    pairs = np.stack([iu[inside], ju[inside]], 1)
    w1 = O.dof_entropies(codes, types, ext, S, [nb], ests)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb], ests)
    assert rel_err(s2[inside], w2[:, :, 0, :]) < REL
    mi_w = w1[pairs[:, 0], 0, 0, :] + w1[pairs[:, 1], 0, 0, :] - w2[:, 0, 0, :]
    assert np.max(np.abs(mi[inside] - mi_w)) < 1e-10


@pytest.mark.parametrize("explicit", [False, True])
def test_several_bin_schemes_in_one_call(explicit):
    """Bin schemes 30 and 50 of a 48-DOF trajectory discretised together
    (codes [D][2][F]); S1, S2 and MI in the [record][estimator][scheme][step]
    layout of contri_S, against the oracle's entropy2D over both schemes."""
    F, S, schemes, ests = 8000, 4, [30, 50], (0, 2)
    bat = hip.synth_bat(0xCE27E, 16, 16, 16, F)
    types = np.array([1] * 16 + [2] * 16 + [3] * 16, np.uint8)
    D = 48
    with hip.Context(0) as ctx:
        codes = np.zeros((D, 2, F), np.uint8)
        ext = np.zeros((D, 4))
        for t, sl in ((1, slice(0, 16)), (2, slice(16, 32)), (3, slice(32, 48))):
            c, e = ctx.discretize_hist(bat[:, sl], t, schemes, 30, True)
            codes[sl] = c
            ext[sl] = e
        iu, ju = np.triu_indices(D, 1)
        pairs = np.stack([iu, ju], 1)
        if explicit:
            rng = np.random.default_rng(4)
            pairs = pairs[rng.choice(len(pairs), 300, replace=False)]
            pairs[::3] = pairs[::3, ::-1]
        P = len(pairs)
        plans = [hip.PairPlan(ctx, D, nb, pairs.astype(np.uint32) if explicit else None)
                 for nb in schemes]
        dc, dt, dx = ctx.to_device(codes), ctx.to_device(types), ctx.to_device(ext)
        d_s1 = ctx.alloc((D, len(ests), 2, S), np.float64)
        d_s2 = ctx.alloc((P, len(ests), 2, S), np.float64)
        d_mi = ctx.alloc((P, 2, S), np.float64)
        print(f"two schemes explicit={explicit} P={P}", flush=True)
        ctx.dev_pair_entropy_schemes(plans, dc.ptr, F, F, dt.ptr, dx.ptr, S, ests, True, False,
                                     d_s1.ptr, d_s2.ptr, d_mi.ptr)
        s1, s2, mi = d_s1.numpy(), d_s2.numpy(), d_mi.numpy()
        for pl in plans:
            pl.close()
    w1 = O.dof_entropies(codes, types, ext, S, schemes, ests)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, schemes, ests)
    assert rel_err(s1, w1) < REL
    assert rel_err(s2, w2) < REL
    mi_w = w1[pairs[:, 0], 0] + w1[pairs[:, 1], 0] - w2[:, 0]
    assert np.max(np.abs(mi - mi_w)) < 1e-10
