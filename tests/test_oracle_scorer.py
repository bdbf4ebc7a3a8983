"""Pin MI, the MIE sums and the MIST (Boruvka) oracle to the stored CSVs.

The stored entropy contributions (entcontri_*.nc) of example-1 and
example-2 are scored the way EntropyScorer::run does (EntropyScorer.cpp:
287-865): MI = S1_i + S1_j - S2_ij (Convergance.cpp:183-194), MIE S2 =
sum S1 - sum MI, and the MIST tree over nodes B, A, T with edges in the
order BB, BA, BD, AA, AD, DD.  The convergence CSVs (4 dp) and the tree CSVs
(6 dp, edge sets) are the expected outputs.
"""
import numpy as np
import pytest

import oracle as O

EDGE_ORDER = ["bnd", "ba", "bd", "ang", "ad", "tor"]
ENDS = {"bnd": (0, 0), "ba": (0, 1), "bd": (0, 2), "ang": (1, 1),
        "ad": (1, 2), "tor": (2, 2)}
ONE = {0: "bnd", 1: "ang", 2: "tor"}


def score(golden, ex, step):
    S1 = {t: golden[f"{ex}_S_{ONE[t]}1d"][:, step] for t in range(3)}
    ids1 = {t: golden[f"{ex}_id_{ONE[t]}1d"][:, 0] for t in range(3)}
    idx = {t: {int(d): k for k, d in enumerate(ids1[t])} for t in range(3)}
    off = {0: 0, 1: len(ids1[0]), 2: len(ids1[0]) + len(ids1[1])}
    src, dst, w, sums, mis = [], [], [], {}, {}
    for t in EDGE_ORDER:
        a, b = ENDS[t]
        ids = golden[f"{ex}_id_{t}2d"]
        S2 = golden[f"{ex}_S_{t}2d"][:, step]
        ia = np.array([idx[a][int(i)] for i in ids[:, 0]])
        ib = np.array([idx[b][int(i)] for i in ids[:, 1]])
        mi = S1[a][ia] + S1[b][ib] - S2
        src.append(ia + off[a])
        dst.append(ib + off[b])
        w.append(mi)
        sums[t] = S2.sum()
        mis[t] = mi.sum()
    nodes_t = np.concatenate([np.full(len(ids1[t]), t) for t in range(3)])
    nodes_id = np.concatenate([ids1[t] for t in range(3)])
    return (S1, sums, mis, np.concatenate(src), np.concatenate(dst),
            np.concatenate(w), nodes_t, nodes_id)


@pytest.mark.parametrize("ex", ["ex1", "ex2"])
def test_mie_and_mist_csv(golden, ex):
    mie = golden[f"{ex}h_MIE_csv"]
    mist = golden[f"{ex}h_MIST_csv"]
    for step in range(4):
        S1, sums, mis, src, dst, w, nt, nid = score(golden, ex, step)
        singlet = sum(S1[t].sum() for t in range(3))
        row = mie[step]
        assert abs(row[1] - S1[0].sum()) < 6e-5
        assert abs(row[2] - S1[1].sum()) < 6e-5
        assert abs(row[3] - S1[2].sum()) < 6e-5
        cols = {"bnd": 4, "ang": 6, "tor": 8, "ba": 10, "bd": 12, "ad": 14}
        for t, c in cols.items():
            assert abs(row[c] - sums[t]) < 6e-5
            assert abs(row[c + 1] - mis[t]) < 6e-5
        assert abs(row[16] - singlet) < 6e-5
        assert abs(row[17] - (singlet - sum(mis.values()))) < 6e-5

        tree = O.boruvka(len(nt), src, dst, w)
        want = golden[f"{ex}h_tree{step}"]
        assert len(tree) == len(want)
        got = set()
        for e in tree:
            a, b = int(src[e]), int(dst[e])
            got.add((int(nt[a]), int(nt[b]), int(nid[a]), int(nid[b])))
        exp = set((int(r[0]), int(r[1]), int(r[2]), int(r[3])) for r in want)
        assert got == exp
        assert np.allclose(np.sort(w[tree])[::-1], want[:, 4], atol=6e-7)
        # MIST row: I2 columns are per-type tree sums
        mrow = mist[step]
        tsum = w[tree].sum()
        assert abs(mrow[17] - (singlet - tsum)) < 6e-5


def kruskal(n, src, dst, w):
    order = sorted(range(len(w)), key=lambda e: (-w[e], e))
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
    return sorted(out)


@pytest.mark.parametrize("seed", range(6))
def test_boruvka_equals_kruskal_total_order(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 60))
    m = int(rng.integers(1, 400))
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    # heavy ties: weights drawn from a small set
    w = rng.integers(0, 5, len(src)).astype(float) / 4
    tree = O.boruvka(n, src, dst, w)
    assert sorted(tree.tolist()) == kruskal(n, src.tolist(), dst.tolist(), w.tolist())
