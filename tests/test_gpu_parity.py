"""HIP path vs the CPU oracle and the reference's stored outputs.

Every test calls the C ABI (libcentre_hip.so) through ctypes; the oracle is
only the checker.  Bars: codes, counts and MST edge sets bit-exact;
entropies within 1e-10 relative (the pair kernel sums c*log(c) per pair in a
different order than entropy2D, and the entropies are O(1), so the
difference stays near 1e-14); MI, which can be close to zero, within 1e-10
absolute.
"""
import numpy as np
import pytest

import oracle as O
from centre_amd import hip

pytestmark = pytest.mark.gpu

REL = 1e-10


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


@pytest.fixture(params=[0, 1], ids=["fp4", "int8"])
def fctx(ctx, request):
    """The module context with the pair tiles' one-hot operands in FP4
    (the default wherever counts stay below 2^24) or int8."""
    ctx.set_mfma_format(request.param)
    yield ctx
    ctx.set_mfma_format(0)


# ---------------------------------------------------------------------------
# reference goldens: example-1 codes -> entcontri S1/S2 (30 bins, 4 steps)
TYPES = {"b": 1, "a": 2, "t": 3}


def ex1_all(golden):
    codes = np.concatenate([golden["ex1_codes_b"], golden["ex1_codes_a"],
                            golden["ex1_codes_t"]])
    ext = np.concatenate([golden["ex1_extreme_b"], golden["ex1_extreme_a"],
                          golden["ex1_extreme_t"]])
    nb_, na_ = len(golden["ex1_codes_b"]), len(golden["ex1_codes_a"])
    types = np.array([1] * nb_ + [2] * na_ + [3] * len(golden["ex1_codes_t"]), np.uint8)
    off = {"b": 0, "a": nb_, "t": nb_ + na_}
    return codes, ext, types, off


def test_example1_pair_entropies_vs_reference(ctx, golden):
    codes, ext, types, off = ex1_all(golden)
    kinds = {"bnd": "bb", "ang": "aa", "tor": "tt", "ba": "ba", "bd": "bt", "ad": "at"}
    pairs, want = [], []
    for k, (x, y) in kinds.items():
        ids = golden[f"ex1_id_{k}2d"].astype(np.int64)
        pairs.append(np.stack([ids[:, 0] + off[x], ids[:, 1] + off[y]], 1))
        want.append(golden[f"ex1_S_{k}2d"])
    pairs = np.concatenate(pairs)
    want = np.concatenate(want)
    print(f"example-1: {len(pairs)} pairs, {codes.shape} codes", flush=True)
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, 30, 4, pairs=pairs)
    assert rel_err(s2[:, 0, :], want) < REL
    for k, t in (("bnd", "b"), ("ang", "a"), ("tor", "t")):
        ids = golden[f"ex1_id_{k}1d"][:, 0].astype(np.int64)
        assert rel_err(s1[ids + off[t], 0, :], golden[f"ex1_S_{k}1d"]) < REL
    mi_want = s1[pairs[:, 0], 0, :] + s1[pairs[:, 1], 0, :] - want
    assert np.max(np.abs(mi - mi_want)) < 1e-10


def test_example1_counts_bit_exact(ctx, golden):
    codes, ext, types, off = ex1_all(golden)
    D, F = codes.shape
    plan = hip.PairPlan(ctx, D, 30)
    P = plan.records
    dc, dt, dx = ctx.to_device(codes), ctx.to_device(types), ctx.to_device(ext)
    s2 = ctx.alloc((P, 1, 4), np.float64)
    cnt = ctx.alloc((P, 4, 900), np.uint32)
    ctx.dev_pair_entropy(plan, dc.ptr, F, F, dt.ptr, dx.ptr, 4, [0], True, False,
                         0, s2.ptr, 0, cnt.ptr)
    got = cnt.numpy()
    iu, ju = np.triu_indices(D, 1)
    pick = np.random.default_rng(0).choice(P, 300, replace=False)
    want = O.pair_counts(codes, 30, np.stack([iu[pick], ju[pick]], 1), 4)
    assert np.array_equal(got[pick].reshape(-1, 4, 30, 30), want.astype(np.uint32))


# ---------------------------------------------------------------------------
# synthetic configuration, 48 DOF (16 B / 16 A / 16 T) x 10,000 frames
@pytest.fixture(scope="module")
def cfg1():
    bat = hip.synth_bat(0xCE27E + 0, 16, 16, 16, 10000)
    return bat


def oracle_discretize(bat, nb, nbins_ref, optimize=True):
    F, D = bat.shape
    codes = np.zeros((D, 1, F), np.uint8)
    ext = np.zeros((D, 4))
    types = np.array([1] * 16 + [2] * 16 + [3] * 16, np.uint8)
    for d in range(D):
        codes[d], ext[d] = O.real2int(bat[:, d].astype(np.float32), types[d], [nb],
                                      nbins_ref, optimize)
    return codes, ext, types


@pytest.mark.parametrize("optimize", [True, False])
def test_discretize_hist_vs_oracle(ctx, cfg1, optimize):
    nb = 30
    want_c, want_e, types = oracle_discretize(cfg1, nb, nb, optimize)
    for t, sl in ((1, slice(0, 16)), (2, slice(16, 32)), (3, slice(32, 48))):
        codes, ext = ctx.discretize_hist(cfg1[:, sl], t, [nb], nb, optimize)
        assert np.array_equal(codes, want_c[sl])
        assert np.array_equal(ext[:, [0, 1, 3]], want_e[sl][:, [0, 1, 3]])
        assert np.max(np.abs(ext[:, 2] - want_e[sl][:, 2])) < 1e-12


def test_discretize_multi_scheme(ctx, cfg1):
    for t, sl in ((1, slice(0, 16)), (3, slice(32, 48))):
        codes, ext = ctx.discretize_hist(cfg1[:, sl], t, [10, 30, 50], 30, True)
        for d in range(16):
            wc, we = O.real2int(cfg1[:, sl][:, d].astype(np.float32), t, [10, 30, 50], 30, True)
            assert np.array_equal(codes[d], wc)
            assert np.array_equal(ext[d, [0, 1, 3]], we[[0, 1, 3]])


@pytest.mark.parametrize("ests", [(0,), (1,), (0, 1), (2,), (3,), (0, 1, 2, 3)])
def test_cfg1_pair_and_marginal_entropies(ctx, cfg1, ests):
    nb, S = 30, 5
    codes, ext, types = oracle_discretize(cfg1, nb, nb)
    codes = codes[:, 0, :]
    D = codes.shape[0]
    iu, ju = np.triu_indices(D, 1)
    pairs = np.stack([iu, ju], 1)
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, ests=ests)
    w1 = O.dof_entropies(codes, types, ext, S, [nb], ests)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb], ests)
    assert rel_err(s1[:, :, :], w1[:, :, 0, :]) < REL
    assert rel_err(s2, w2[:, :, 0, :]) < REL
    mi_w = w1[iu, 0, 0, :] + w1[ju, 0, 0, :] - w2[:, 0, 0, :]
    assert np.max(np.abs(mi - mi_w)) < 1e-10


@pytest.mark.parametrize("nb,D,F,S", [(50, 7, 3001, 3), (5, 70, 2049, 2),
                                      (3, 130, 999, 1), (255, 3, 1000, 4),
                                      (64, 9, 50, 7), (31, 13, 4096, 16)])
def test_ragged_shapes_vs_oracle(ctx, nb, D, F, S):
    rng = np.random.default_rng(nb * 1000 + D)
    codes = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
    # correlate some rows so counts are not uniform
    codes[1::2] = (codes[0::2][: len(codes[1::2])] + rng.integers(0, 2, size=F)) % nb
    types = rng.integers(1, 4, size=D).astype(np.uint8)
    ext = np.zeros((D, 4))
    ext[:, 0] = rng.uniform(0.5, 1.0, D)
    ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
    iu, ju = np.triu_indices(D, 1)
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, ests=(0, 1))
    w2 = O.pair_entropies(codes, types, ext, np.stack([iu, ju], 1), S, [nb], (0, 1))
    assert rel_err(s2, w2[:, :, 0, :]) < REL


@pytest.mark.parametrize("ests", [(0, 1), (0,)])
@pytest.mark.parametrize("nb,D,F,S", [(5, 70, 3001, 3), (8, 64, 2048, 2), (2, 33, 999, 1),
                                      (3, 100, 5000, 5), (7, 40, 130, 2), (1, 35, 700, 2)])
def test_small_bin_tiles_vs_oracle(fctx, nb, D, F, S, ests):
    """nb <= 8 runs every tile, those of a partial last block included, on
    the eight-wave small-bin kernel (32 DOFs a side, 8 one-hot rows per DOF
    built by byte permutes); D past and below a block multiple, steps of
    partial chunks, ML alone and with MM, all pairs and explicit records."""
    rng = np.random.default_rng(nb * 11 + D)
    codes = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
    codes[1::3] = (codes[0::3][: len(codes[1::3])] + rng.integers(0, 2, size=F)) % nb
    codes[2::7] = 0                                                   # one bin
    types = rng.integers(1, 4, size=D).astype(np.uint8)
    ext = np.zeros((D, 4))
    ext[:, 0] = rng.uniform(0.5, 1.0, D)
    ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
    iu, ju = np.triu_indices(D, 1)
    ctx = fctx
    print(f"small-bin tiles nb={nb} D={D} F={F} S={S} ests={ests}", flush=True)
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, ests=ests)
    w2 = O.pair_entropies(codes, types, ext, np.stack([iu, ju], 1), S, [nb], ests)
    assert rel_err(s2, w2[:, :, 0, :]) < REL
    w1 = O.dof_entropies(codes, types, ext, S, [nb], ests)
    assert rel_err(mi, w1[iu, 0, 0, :] + w1[ju, 0, 0, :] - w2[:, 0, 0, :]) < 1e-9
    pairs = np.stack([iu, ju], 1)[rng.permutation(len(iu))[:80]]
    pairs[::2] = pairs[::2, ::-1]
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, pairs=pairs, ests=ests[::-1])
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb], ests[::-1])
    assert rel_err(s2, w2[:, :, 0, :]) < REL


@pytest.mark.parametrize("nb,D", [(50, 160), (20, 300), (6, 600)])
def test_pair_stage_is_deterministic(ctx, nb, D):
    """The same inputs give bit-identical S2 and MI on every call: the
    per-pair c log c sums are integers in LDS (fixed point), so the order in
    which lanes and waves add them does not matter.  Eight-wave tile
    (nb 50), four-wave tile (nb 20) and small-bin tile (nb 6), each with
    hundreds of concurrently running tiles; checked against the oracle too."""
    rng = np.random.default_rng(nb)
    F, S = 12000, 6
    codes = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
    codes[1::2] = (codes[0::2][: len(codes[1::2])] + rng.integers(0, 2, size=F)) % nb
    types = np.array([1, 2, 3] * D, np.uint8)[:D]
    ext = np.zeros((D, 4))
    ext[:, 0] = rng.uniform(0.5, 1.0, D)
    ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
    runs = [ctx.pair_entropy(codes, types, ext, nb, S, ests=(0, 1)) for _ in range(4)]
    for s1, s2, mi in runs[1:]:
        assert np.array_equal(s2.view(np.uint64), runs[0][1].view(np.uint64))
        assert np.array_equal(mi.view(np.uint64), runs[0][2].view(np.uint64))
    iu, ju = np.triu_indices(D, 1)
    pick = np.linspace(0, len(iu) - 1, 60).astype(np.int64)
    w2 = O.pair_entropies(codes, types, ext, np.stack([iu[pick], ju[pick]], 1), S, [nb], (0, 1))
    assert rel_err(runs[0][1][pick], w2[:, :, 0, :]) < REL


@pytest.mark.parametrize("ests", [(0, 1), (0,)])
@pytest.mark.parametrize("nb,D,F,S", [(50, 20, 4000, 4), (42, 18, 2500, 5), (36, 14, 3333, 3),
                                      (50, 23, 6002, 2), (29, 17, 1500, 2), (31, 16, 2000, 3),
                                      (33, 15, 1000, 2), (40, 13, 1200, 2), (45, 11, 1100, 2),
                                      (130, 4, 700, 2)])
def test_full_tiles_vs_oracle(fctx, nb, D, F, S, ests):
    """Tiles with R DOFs on both sides (diagonal ones included) run on the
    eight-wave kernel for nb >= 29; shapes with whole blocks only, a
    partial last block, and step lengths that are and are not a multiple of
    the 64-frame chunk.  With ML alone the kernel leaves out the nonzero-bin
    counts and splits each 32-row block at its DOF boundary group-wise, with
    MM it counts them and splits row by row: both forms, at bin counts that
    put the boundaries at every row offset mod 4."""
    rng = np.random.default_rng(nb * 7 + D)
    codes = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
    codes[1::3] = (codes[0::3][: len(codes[1::3])] + rng.integers(0, 3, size=F)) % nb
    codes[2::5] = rng.integers(0, 4, size=(len(codes[2::5]), F))      # peaked
    types = rng.integers(1, 4, size=D).astype(np.uint8)
    ext = np.zeros((D, 4))
    ext[:, 0] = rng.uniform(0.5, 1.0, D)
    ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
    iu, ju = np.triu_indices(D, 1)
    ctx = fctx
    print(f"full tiles nb={nb} D={D} F={F} S={S} ests={ests}", flush=True)
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, ests=ests)
    w2 = O.pair_entropies(codes, types, ext, np.stack([iu, ju], 1), S, [nb], ests)
    assert rel_err(s2, w2[:, :, 0, :]) < REL
    w1 = O.dof_entropies(codes, types, ext, S, [nb], ests)
    wmi = w1[iu, 0, 0, :] + w1[ju, 0, 0, :] - w2[:, 0, 0, :]
    assert rel_err(mi, wmi) < 1e-9
    pairs = np.stack([iu, ju], 1)[rng.permutation(len(iu))[:60]]
    pairs[::2] = pairs[::2, ::-1]
    rev = ests[::-1]
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, pairs=pairs, ests=rev)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb], rev)
    assert rel_err(s2, w2[:, :, 0, :]) < REL


@pytest.mark.parametrize("nb,D,F,S", [(50, 160, 12000, 6), (36, 70, 9001, 4), (6, 300, 10000, 5),
                                      (3, 96, 777, 3)])
def test_mfma_formats_bit_identical(ctx, nb, D, F, S):
    """FP4 and int8 one-hot operands give the same counts, so S2 and MI
    (fixed-point sums of the same counts) are bit-identical: eight-wave
    tile (nb 50, 36) and small-bin tile (nb 6, 3), partial chunks, ML + MM."""
    rng = np.random.default_rng(nb * 5 + D)
    codes = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
    codes[1::2] = (codes[0::2][: len(codes[1::2])] + rng.integers(0, 2, size=F)) % nb
    codes[3::7] = 0
    types = rng.integers(1, 4, size=D).astype(np.uint8)
    ext = np.zeros((D, 4))
    ext[:, 0] = rng.uniform(0.5, 1.0, D)
    ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
    print(f"formats nb={nb} D={D} F={F} S={S}", flush=True)
    out = []
    try:
        for fmt in (0, 1):
            ctx.set_mfma_format(fmt)
            out.append(ctx.pair_entropy(codes, types, ext, nb, S, ests=(0, 1)))
    finally:
        ctx.set_mfma_format(0)
    assert np.array_equal(out[0][1].view(np.uint64), out[1][1].view(np.uint64))
    assert np.array_equal(out[0][2].view(np.uint64), out[1][2].view(np.uint64))


def test_counts_past_the_fp4_limit(ctx):
    """One step of 18,500,003 frames with joint bin (0, 0) holding about 99%
    of them: the count passes 2^24, where f32 accumulation of FP4 products
    would round, so the launch takes int8 operands; S2 against the oracle."""
    nb, D, F = 50, 5, 18_500_003
    rng = np.random.default_rng(24)
    codes = np.zeros((D, F), np.uint8)
    hot = rng.random((D, F)) < 0.005
    codes[hot] = rng.integers(0, nb, size=int(hot.sum()), dtype=np.uint8)
    types = np.array([1, 2, 3, 1, 2], np.uint8)
    ext = np.zeros((D, 4))
    ext[:, 0] = 1.0
    ext[:, 1] = 2.0
    iu, ju = np.triu_indices(D, 1)
    print(f"past the fp4 limit nb={nb} D={D} F={F}", flush=True)
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, 1, ests=(0,))
    w2 = O.pair_entropies(codes, types, ext, np.stack([iu, ju], 1), 1, [nb], (0,))
    assert rel_err(s2, w2[:, :, 0, :]) < REL


def test_explicit_pairs_order_and_duplicates(ctx, cfg1):
    nb, S = 30, 4
    codes, ext, types = oracle_discretize(cfg1, nb, nb)
    codes = codes[:, 0, :]
    rng = np.random.default_rng(5)
    pairs = rng.integers(0, 48, size=(200, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    pairs = np.concatenate([pairs, pairs[:10, ::-1], pairs[:5]])
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, pairs=pairs)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb])
    assert rel_err(s2[:, 0, :], w2[:, 0, 0, :]) < REL


def test_explicit_reversed_pairs_all_estimators_and_counts(ctx, cfg1):
    """A record (i, j) with i > j keeps the reference's orientation: row =
    first DOF.  Chao-Shen evaluates the second DOF's Jacobian at the row bin
    (Estimators.cpp:558,562), so its S2 depends on the order; the counts of
    such a record are the transpose of (j, i)'s."""
    nb, S = 30, 4
    codes, ext, types = oracle_discretize(cfg1, nb, nb)
    codes = codes[:, 0, :]
    D, F = codes.shape
    rng = np.random.default_rng(11)
    pairs = rng.integers(0, D, size=(120, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    pairs = np.concatenate([pairs, pairs[:20, ::-1]]).astype(np.uint32)
    ests = (2, 0, 1, 3)
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, pairs=pairs, ests=ests)
    w1 = O.dof_entropies(codes, types, ext, S, [nb], ests)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb], ests)
    assert rel_err(s2, w2[:, :, 0, :]) < REL
    # the swap changes Chao-Shen where both DOFs carry a Jacobian
    mi_w = w1[pairs[:, 0], 0, 0, :] + w1[pairs[:, 1], 0, 0, :] - w2[:, 0, 0, :]
    assert np.max(np.abs(mi - mi_w)) < 1e-10
    plan = hip.PairPlan(ctx, D, nb, pairs)
    P = len(pairs)
    dc, dt, dx = ctx.to_device(codes), ctx.to_device(types), ctx.to_device(ext)
    ds2 = ctx.alloc((P, 1, S), np.float64)
    cnt = ctx.alloc((P, S, nb * nb), np.uint32)
    print(f"explicit plan counts: {P} records", flush=True)
    ctx.dev_pair_entropy(plan, dc.ptr, F, F, dt.ptr, dx.ptr, S, [0], True, False,
                         0, ds2.ptr, 0, cnt.ptr)
    plan.close()
    want = O.pair_counts(codes, nb, pairs, S)
    assert np.array_equal(cnt.numpy().reshape(P, S, nb, nb), want.astype(np.uint32))
    assert rel_err(ds2.numpy()[:, 0, :], w2[:, 1, 0, :]) < REL


@pytest.mark.parametrize("bad,F", [(30, 100), (255, 100), (255, 4096), (31, 4096)])
def test_code_out_of_range_is_an_error(ctx, bad, F):
    codes = np.zeros((4, F), np.uint8)
    codes[2, F // 2] = bad
    with pytest.raises(hip.CentreError) as e:
        ctx.pair_entropy(codes, np.ones(4, np.uint8), np.tile([1.0, 2.0, 0, 0], (4, 1)), 30, 2)
    assert e.value.code == -4


def test_mfma_format_argument(ctx):
    """Only 0 (FP4 where exact) and 1 (int8) are operand formats; another
    value is refused with -3 and leaves the context's format as it was."""
    for bad in (-1, 2, 7):
        with pytest.raises(hip.CentreError) as e:
            ctx.set_mfma_format(bad)
        assert e.value.code == -3
    ctx.set_mfma_format(0)


# ---------------------------------------------------------------------------
def kruskal_set(n, src, dst, w):
    w = np.asarray(w, np.float64) + 0.0          # -0.0 and +0.0 tie
    order = np.lexsort((np.arange(len(w)), -w)).tolist()
    par = list(range(n))

    def find(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x
    out = []
    for e in order:
        a, b = find(src[e]), find(dst[e])
        if a != b:
            par[a] = b
            out.append(e)
            if len(out) == n - 1:
                break
    return out


@pytest.mark.parametrize("seed", range(5))
def test_mst_vs_oracle(ctx, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 300))
    m = int(rng.integers(1, 3000))
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = rng.integers(0, 7, len(src)).astype(float) / 3 - 0.5
    got = ctx.mst(n, src, dst, w)
    want = O.boruvka(n, src, dst, w)
    assert sorted(got.tolist()) == sorted(want.tolist())
    assert got.tolist() == kruskal_set(n, src.tolist(), dst.tolist(), w.tolist())


def test_mst_signed_zero_weights_tie(ctx):
    """-0.0 and +0.0 compare equal in boruvkaMST (operator <, Centre.cpp:43-53),
    so among zero-weight edges the lowest index wins whatever the sign."""
    rng = np.random.default_rng(7)
    n, m = 60, 400
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = rng.choice([0.0, -0.0, 0.25], size=len(src))
    w[0::3] = -0.0
    assert np.signbit(w).any() and (~np.signbit(w) & (w == 0)).any()
    got = ctx.mst(n, src, dst, w)
    want = O.boruvka(n, src, dst, w)
    assert sorted(got.tolist()) == sorted(want.tolist())
    assert got.tolist() == kruskal_set(n, src.tolist(), dst.tolist(), w.tolist())


def test_mst_neighbour_graph_scale(ctx):
    """The neighbour-list graph of the vmKDE workload's size (3,000 nodes,
    pairs within 30 of each other, ~90k edges) with weights from a small set
    (heavy ties): the same forest as the oracle's Boruvka and Kruskal."""
    rng = np.random.default_rng(99)
    n = 3000
    i, j = np.triu_indices(n, 1)
    keep = (j - i) <= 30
    src, dst = i[keep], j[keep]
    w = rng.integers(0, 50, len(src)).astype(float) / 7.0
    print(f"mst n={n} m={len(src)}", flush=True)
    got = ctx.mst(n, src, dst, w)
    want = O.boruvka(n, src, dst, w)
    assert len(got) == n - 1
    assert sorted(got.tolist()) == sorted(want.tolist())
    assert got.tolist() == kruskal_set(n, src.tolist(), dst.tolist(), w.tolist())


def test_mst_complete_graph_in_pair_order(ctx):
    """All pairs of 2,500 DOFs in the pair stage's (i < j) order, weights from
    a small set (heavy ties): rows of the triangle give waves whose edges
    share a source, which the max pass reduces across the wave before its
    LDS table; the forest equals the oracle's Boruvka and Kruskal."""
    rng = np.random.default_rng(5)
    n = 2500
    src, dst = np.triu_indices(n, 1)
    w = rng.integers(0, 200, len(src)).astype(float) / 9.0
    print(f"mst complete n={n} m={len(src)}", flush=True)
    got = ctx.mst(n, src, dst, w)
    want = O.boruvka(n, src, dst, w)
    assert len(got) == n - 1
    assert sorted(got.tolist()) == sorted(want.tolist())
    assert got.tolist() == kruskal_set(n, src.tolist(), dst.tolist(), w.tolist())


def test_mst_more_nodes_than_the_lds_table(ctx):
    """40,000 nodes (beyond the 16,384-node LDS table, so keys go straight to
    global memory) in a few components, with runs of edges sharing an
    endpoint and tied weights: the same forest as the oracle's."""
    rng = np.random.default_rng(11)
    n, m = 40000, 300000
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    src[:64000] = np.repeat(rng.integers(0, n, 1000), 64)   # wave-uniform sources
    keep = (src != dst) & (src % 3 == dst % 3)                # three components at least
    src, dst = src[keep], dst[keep]
    w = rng.integers(0, 40, len(src)).astype(float) / 3.0
    print(f"mst n={n} m={len(src)}", flush=True)
    got = ctx.mst(n, src, dst, w)
    want = O.boruvka(n, src, dst, w)
    assert sorted(got.tolist()) == sorted(want.tolist())
    assert got.tolist() == kruskal_set(n, src.tolist(), dst.tolist(), w.tolist())


def test_mst_nan_weights_are_absent_edges(ctx):
    """Edges of NaN weight (pairs a screened pass skipped) are left out: the
    forest equals the oracle's Boruvka over the other edges, in the original
    edge indices, on a complete graph and on a sparse one that the NaN edges
    would otherwise connect."""
    rng = np.random.default_rng(23)
    for n, dense in ((700, True), (5000, False)):
        if dense:
            src, dst = np.triu_indices(n, 1)
        else:
            src = rng.integers(0, n, 40000)
            dst = rng.integers(0, n, 40000)
            keep = src != dst
            src, dst = src[keep], dst[keep]
        w = rng.integers(0, 60, len(src)).astype(float) / 7.0
        w[rng.random(len(src)) < 0.6] = np.nan
        got = ctx.mst(n, src, dst, w)
        ok = np.flatnonzero(~np.isnan(w))
        want = ok[O.boruvka(n, src[ok], dst[ok], w[ok])]
        assert not np.isnan(w[got]).any()
        assert sorted(got.tolist()) == sorted(want.tolist())
        assert got.tolist() == ok[kruskal_set(n, src[ok].tolist(), dst[ok].tolist(),
                                               w[ok].tolist())].tolist()


def test_mst_endpoint_out_of_range_is_an_error(ctx):
    """An endpoint >= n_node is refused (-3) before any round runs; the
    context stays usable."""
    src = np.array([0, 1, 2], np.uint32)
    dst = np.array([1, 2, 5], np.uint32)
    with pytest.raises(hip.CentreError) as e:
        ctx.mst(5, src, dst, np.ones(3))
    assert e.value.code == -3
    assert sorted(ctx.mst(6, src, dst, np.ones(3)).tolist()) == [0, 1, 2]


def test_mst_example2_trees(ctx, golden):
    from test_oracle_scorer import score
    for step in range(4):
        _, _, _, src, dst, w, nt, nid = score(golden, "ex2", step)
        got = ctx.mst(len(nt), src, dst, w)
        want = O.boruvka(len(nt), src, dst, w)
        assert sorted(got.tolist()) == sorted(want.tolist())
        assert len(got) == 1036


# ---------------------------------------------------------------------------
def test_vmkde_grid_and_states(ctx):
    bat = hip.synth_bat(0xCE27E + 2, 0, 0, 12, 3000)
    # even grids pair points half a turn apart; odd ones evaluate each point
    for G, kappa, nmax in ((360, 1.0, 3), (64, 10.0, 5), (45, 4.0, 4), (2, 1.0, 2)):
        print(f"vmkde G={G} kappa={kappa}", flush=True)
        codes, mins, nf, grid = ctx.discretize_vmkde(bat, nmax, kappa, 5, 1e-4, 1000, G,
                                                     want_grid=True)
        for d in range(bat.shape[1]):
            x = bat[:, d].astype(np.float32)
            wg = O.vmkde_grid(x, kappa, G)
            g = grid[d].reshape(-1)
            scale = np.max(np.abs(wg[: G + 1]))
            assert np.max(np.abs(g[: G + 1] - wg[: G + 1])) <= 1e-11 * scale
            # the state rule on the device grid is bit-exact with the oracle's
            wc, wm, wn = O.vmkde_states(x, g, nmax, 5, 1e-4, 1000, G)
            assert nf[d] == wn
            assert np.array_equal(mins[d, :wn], wm)
            assert np.array_equal(codes[d], wc)
            # end to end: the oracle's own grid gives the same states
            ec, em, en = O.vmkde(x, nmax, kappa, 5, 1e-4, 1000, G)
            assert nf[d] == en and np.array_equal(codes[d], ec)
            assert np.max(np.abs(mins[d, :en] - em), initial=0.0) <= 1e-12


def test_vmkde_eight_states():
    """An eight-mode torsion mixture: eight minima descend on eight lanes of
    the state kernel (nmaxconf 9), on even grids and an odd one."""
    rng = np.random.default_rng(77)
    F, D = 4000, 3
    centers = np.arange(8) * (np.pi / 4) - np.pi + 0.2
    bat = np.zeros((F, D))
    for d in range(D):
        k = rng.integers(0, 8, F)
        bat[:, d] = np.angle(np.exp(1j * (centers[k] + rng.normal(0, 0.08, F))))
    with hip.Context(0) as ctx:
        for G, kappa in ((360, 10.0), (120, 6.0), (91, 2.0)):
            print(f"vmkde eight states G={G} kappa={kappa}", flush=True)
            codes, mins, nf, grid = ctx.discretize_vmkde(bat, 9, kappa, 5, 1e-4, 1000, G,
                                                         want_grid=True)
            for d in range(D):
                x = bat[:, d].astype(np.float32)
                wg = O.vmkde_grid(x, kappa, G)
                g = grid[d].reshape(-1)
                scale = np.max(np.abs(wg[: G + 1]))
                assert np.max(np.abs(g[: G + 1] - wg[: G + 1])) <= 1e-11 * scale
                wc, wm, wn = O.vmkde_states(x, g, 9, 5, 1e-4, 1000, G)
                assert wn >= 6 and nf[d] == wn
                assert np.array_equal(mins[d, :wn], wm)
                assert np.array_equal(codes[d], wc)
                ec, em, en = O.vmkde(x, 9, kappa, 5, 1e-4, 1000, G)
                assert nf[d] == en and np.array_equal(codes[d], ec)
                assert np.max(np.abs(mins[d, :en] - em), initial=0.0) <= 1e-12


# ---------------------------------------------------------------------------
# edge cases: one bin, one frame per step, two DOFs, empty pair lists, the
# count-materialisation limit
def test_single_bin_and_single_frame_steps(ctx):
    rng = np.random.default_rng(21)
    for nb, D, F, S in ((1, 5, 1000, 3), (7, 2, 64, 64), (50, 2, 255, 255)):
        codes = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
        types = rng.integers(1, 4, size=D).astype(np.uint8)
        ext = np.zeros((D, 4))
        ext[:, 0] = rng.uniform(0.5, 1.0, D)
        ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
        iu, ju = np.triu_indices(D, 1)
        print(f"edge nb={nb} D={D} F={F} S={S}", flush=True)
        s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, ests=(0, 1, 2, 3))
        w1 = O.dof_entropies(codes, types, ext, S, [nb], (0, 1, 2, 3))
        w2 = O.pair_entropies(codes, types, ext, np.stack([iu, ju], 1), S, [nb], (0, 1, 2, 3))
        assert rel_err(s1, w1[:, :, 0, :]) < REL
        assert rel_err(s2, w2[:, :, 0, :]) < REL


def test_empty_pair_list(ctx):
    codes = np.zeros((4, 100), np.uint8)
    s1, s2, mi = ctx.pair_entropy(codes, np.ones(4, np.uint8), np.tile([1.0, 2.0, 0, 0], (4, 1)),
                                  10, 2, pairs=np.zeros((0, 2), np.uint32))
    assert s2.shape == (0, 1, 2) and mi.shape == (0, 2)
    assert np.all(np.isfinite(s1))


@pytest.mark.parametrize("nb,D,F,S,ests,explicit", [
    (5, 40, 20000, 3, (0,), False),      # one tile, eight parts, short last part
    (16, 50, 9000, 2, (1, 0), True),     # ten tiles, MI from MM
    (3, 20, 4096, 4, (0, 1), True),      # two parts of 512 frames
])
def test_frame_split_vs_oracle(ctx, nb, D, F, S, ests, explicit):
    """Plans with few tiles at small nb split every step's frames over several
    workgroups per tile and sum their count dumps before the entropies."""
    rng = np.random.default_rng(nb * 31 + D)
    codes = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
    codes[1::2] = (codes[0::2][: len(codes[1::2])] + rng.integers(0, 2, size=F)) % nb
    types = rng.integers(1, 4, size=D).astype(np.uint8)
    ext = np.zeros((D, 4))
    ext[:, 0] = rng.uniform(0.5, 1.0, D)
    ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
    iu, ju = np.triu_indices(D, 1)
    pairs = np.stack([iu, ju], 1)
    if explicit:
        pick = rng.choice(len(pairs), len(pairs) // 3, replace=False)
        pairs = pairs[pick]
        flip = rng.random(len(pairs)) < 0.4
        pairs[flip] = pairs[flip][:, ::-1]
    print(f"frame split nb={nb} D={D} F={F} S={S} ests={ests} explicit={explicit}", flush=True)
    s1, s2, mi = ctx.pair_entropy(codes, types, ext, nb, S, pairs=pairs if explicit else None,
                                  ests=ests)
    w1 = O.dof_entropies(codes, types, ext, S, [nb], ests)
    w2 = O.pair_entropies(codes, types, ext, pairs, S, [nb], ests)
    assert rel_err(s2, w2[:, :, 0, :]) < REL
    mi_w = w1[pairs[:, 0], 0, 0, :] + w1[pairs[:, 1], 0, 0, :] - w2[:, 0, 0, :]
    assert np.max(np.abs(mi - mi_w)) < 1e-10


# ---------------------------------------------------------------------------
# the headline workload at full size: 600 DOF (200 B / 200 A / 200 T) x
# 100,000 frames, 50 bins, 10 steps, all 179,700 pairs
def test_full_size_headline_properties():
    """At full size the checks are size-independent: the plug-in (ML) MI is a
    Kullback-Leibler divergence (the bin widths and Jacobians cancel between
    S1_i + S1_j and S2), so it is >= 0 for every pair and step up to
    rounding; S2 is symmetric in the pair's order; and an evenly strided
    sample of pairs matches the oracle's entropy2D within 1e-10."""
    nbd = nan = ndh = 200
    D, F, nb, S = 600, 100000, 50, 10
    P = D * (D - 1) // 2
    bat = hip.synth_bat(0xCE27E + 1, nbd, nan, ndh, F)
    types = np.array([1] * nbd + [2] * nan + [3] * ndh, np.uint8)
    print(f"full size D={D} F={F} nb={nb} S={S}", flush=True)
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
        d_s1 = ctx.alloc((D, 1, S), np.float64)
        d_s2 = ctx.alloc((P, 1, S), np.float64)
        d_mi = ctx.alloc((P, S), np.float64)
        ctx.dev_pair_entropy(plan, d_codes.ptr, F, F, d_types.ptr, d_ext.ptr, S, [0], True,
                             False, d_s1.ptr, d_s2.ptr, d_mi.ptr)
        s1, s2, mi = d_s1.numpy(), d_s2.numpy(), d_mi.numpy()
        codes, ext = d_codes.numpy(), d_ext.numpy()
        plan.close()
        for buf in (d_s1, d_s2, d_mi, d_codes, d_ext, d_types):
            buf.free()
        # reversed explicit records give the same S2 (ML is symmetric)
        rng = np.random.default_rng(5)
        iu, ju = np.triu_indices(D, 1)
        pick = np.sort(rng.choice(P, 2000, replace=False))
        rev = np.stack([ju[pick], iu[pick]], 1)
        _, s2r, _ = ctx.pair_entropy(codes, types, ext, nb, S, pairs=rev, ests=(0,))
    assert np.all(np.isfinite(s2)) and np.all(np.isfinite(mi))
    assert mi.min() > -1e-11
    assert rel_err(s2r, s2[pick]) < REL
    sample = np.arange(0, P, P // 48)[:48]
    w2 = O.pair_entropies(codes, types, ext, np.stack([iu[sample], ju[sample]], 1), S, [nb], (0,))
    assert rel_err(s2[sample], w2[:, :, 0, :]) < REL
    w1 = O.dof_entropies(codes, types, ext, S, [nb], (0,))
    assert rel_err(s1, w1[:, :, 0, :]) < REL


# ---------------------------------------------------------------------------
# the vmKDE workload at full size: 3,000 torsions x 50,000 frames, kappa 10,
# 64-interval grid, at most 5 states, pairs within 30 DOF, 10 steps
def test_full_size_vmkde_properties():
    """Size-independent checks at full size: every torsion's state count is
    in 0..5 and its codes below it (or all 0); ML MI >= 0 on every
    neighbour pair and step; a sample of torsions matches the oracle's
    states given the device grid, and a strided sample of pairs its
    entropy2D (this plan has 118 tiles, so it runs the frame split)."""
    D, F, G, kappa, nmax, S, cut = 3000, 50000, 64, 10.0, 5, 10, 30
    bat = hip.synth_bat(0xCE27E + 7, 0, 0, D, F)
    ii = np.concatenate([np.full(min(D, i + cut + 1) - i - 1, i) for i in range(D)])
    jj = np.concatenate([np.arange(i + 1, min(D, i + cut + 1)) for i in range(D)])
    pairs = np.stack([ii, jj], 1).astype(np.uint32)
    P = len(pairs)
    types = np.full(D, 3, np.uint8)
    print(f"full size vmkde D={D} F={F} P={P}", flush=True)
    with hip.Context(0) as ctx:
        d_bat = ctx.to_device(bat)
        d_codes = ctx.alloc((D, F), np.uint8)
        d_min = ctx.alloc((D, 9), np.float64)
        d_nf = ctx.alloc(D, np.uint8)
        d_grid = ctx.alloc((D, 2, G + 1), np.float64)
        ctx.dev_discretize_vmkde(d_bat.ptr, D, D, F, nmax, kappa, d_codes.ptr, d_min.ptr,
                                 d_nf.ptr, grid_intervals=G, d_grid=d_grid.ptr)
        ctx.sync()
        d_bat.free()
        d_types, d_ext = ctx.to_device(types), ctx.to_device(np.zeros((D, 4)))
        plan = hip.PairPlan(ctx, D, nmax, pairs)
        d_s1 = ctx.alloc((D, 1, S), np.float64)
        d_s2 = ctx.alloc((P, 1, S), np.float64)
        d_mi = ctx.alloc((P, S), np.float64)
        ctx.dev_pair_entropy(plan, d_codes.ptr, F, F, d_types.ptr, d_ext.ptr, S, [0], True, True,
                             d_s1.ptr, d_s2.ptr, d_mi.ptr)
        codes, mins, nf, grid = d_codes.numpy(), d_min.numpy(), d_nf.numpy(), d_grid.numpy()
        s2, mi = d_s2.numpy(), d_mi.numpy()
        plan.close()
        for buf in (d_codes, d_min, d_nf, d_grid, d_types, d_ext, d_s1, d_s2, d_mi):
            buf.free()
    assert nf.max() <= nmax
    assert np.all(codes.max(axis=1) < np.maximum(nf, 1))
    assert np.all(np.isfinite(s2)) and mi.min() > -1e-11
    for d in (0, 1, 777, 1500, 2998, 2999):
        x = bat[:, d].astype(np.float32)
        wc, wm, wn = O.vmkde_states(x, grid[d].reshape(-1), nmax, 5, 1e-4, 1000, G)
        assert nf[d] == wn and np.array_equal(mins[d, :wn], wm)
        assert np.array_equal(codes[d], wc)
    # end to end on 40 strided torsions: the oracle's own grid at full size
    # (h = 643 for 50k frames) gives the device's grid within 1e-11 of its
    # largest value, the same state count and codes, and minima within 1e-12
    for d in range(0, D, D // 40):
        x = bat[:, d].astype(np.float32)
        wg = O.vmkde_grid(x, kappa, G)
        g = grid[d].reshape(-1)
        assert np.max(np.abs(g[: G + 1] - wg[: G + 1])) <= 1e-11 * np.max(np.abs(wg[: G + 1]))
        ec, em, en = O.vmkde_states(x, wg, nmax, 5, 1e-4, 1000, G)
        assert nf[d] == en and np.array_equal(codes[d], ec), d
        assert np.max(np.abs(mins[d, :en] - em), initial=0.0) <= 1e-12
    sample = np.arange(0, P, P // 500)[:500]
    w2 = O.pair_entropies(codes, types, np.zeros((D, 4)), pairs[sample], S, [nmax], (0,),
                          jacobian=True, kde=True)
    assert rel_err(s2[sample], w2[:, :, 0, :]) < REL
