"""Pin the CPU oracle to the reference's stored example outputs.

example-1 (histogram run, 30 bins, 4 steps, ML estimator, Jacobian on) ships
the uint8 codes it computed (bin_*.nc) and the entropies it derived from them
(entcontri_*.nc), so the whole count -> entropy chain is checked here
record by record.  example-2 ships 1-D torsion counts (hist_tor-1d.nc) next
to their entropies.
"""
import numpy as np
import pytest

import oracle as O

TYPES = {"bnd": O.BOND, "ang": O.ANGLE, "tor": O.DIHEDRAL}
CODES = {"bnd": "b", "ang": "a", "tor": "t"}
PAIRS = {"bnd": ("bnd", "bnd"), "ang": ("ang", "ang"), "tor": ("tor", "tor"),
         "ba": ("bnd", "ang"), "bd": ("bnd", "tor"), "ad": ("ang", "tor")}
NSTEPS, NBINS = 4, 30


def rel_err(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize("t", ["bnd", "ang", "tor"])
def test_entropy_1d_example1(golden, t):
    codes = golden[f"ex1_codes_{CODES[t]}"]
    ext = golden[f"ex1_extreme_{CODES[t]}"]
    ids = golden[f"ex1_id_{t}1d"][:, 0]
    want = golden[f"ex1_S_{t}1d"]
    F = codes.shape[1]
    got = []
    for d in ids:
        f, _ = O.bin_data_1d(codes[d], NSTEPS, [NBINS], ext[d, 0], ext[d, 1])
        got.append(O.entropy_1d(TYPES[t], [O.ML], NSTEPS, F // NSTEPS,
                                ext[d, 0], ext[d, 1], True, False, [NBINS], f))
    assert rel_err(np.array(got), want) < 1e-12


@pytest.mark.parametrize("t", ["bnd", "ang", "tor", "ba", "bd", "ad"])
def test_entropy_2d_example1(golden, t):
    t1, t2 = PAIRS[t]
    c1, c2 = golden[f"ex1_codes_{CODES[t1]}"], golden[f"ex1_codes_{CODES[t2]}"]
    e1, e2 = golden[f"ex1_extreme_{CODES[t1]}"], golden[f"ex1_extreme_{CODES[t2]}"]
    ids = golden[f"ex1_id_{t}2d"]
    want = golden[f"ex1_S_{t}2d"]
    # through the whole-stage driver: one concatenated code matrix
    codes = np.concatenate([c1, c2])
    ext = np.concatenate([e1, e2])
    types = np.array([TYPES[t1]] * len(c1) + [TYPES[t2]] * len(c2), np.uint8)
    pairs = np.stack([ids[:, 0], ids[:, 1] + len(c1)], axis=1)
    got = O.pair_entropies(codes, types, ext, pairs, NSTEPS, [NBINS])
    assert rel_err(got[:, 0, 0, :], want) < 1e-12
    # and through the per-pair functions for the first records
    F = c1.shape[1]
    for r in range(5):
        i, j = ids[r]
        f, _ = O.bin_data_2d(c1[i], c2[j], NSTEPS, [NBINS], e1[i, 0], e1[i, 1],
                             e2[j, 0], e2[j, 1])
        s = O.entropy_2d(TYPES[t1], TYPES[t2], [O.ML], NSTEPS, F // NSTEPS,
                         e1[i, 0], e1[i, 1], e2[j, 0], e2[j, 1], True, False,
                         [NBINS], f)
        assert rel_err(s, want[r]) < 1e-12


def test_entropy_1d_from_stored_counts_example2(golden):
    """hist_tor-1d.nc holds the step-3 (all 10,000 frames) torsion counts."""
    freq = golden["ex2_hist_freq"].astype(np.uint64)
    ext = golden["ex2_hist_extrema"]
    ids = golden["ex2_hist_id"]
    step = int(golden["ex2_hist_firststep"][0])
    S1 = golden["ex2_S_tor1d"]
    id1d = golden["ex2_id_tor1d"][:, 0]
    row = {int(d): r for r, d in enumerate(id1d)}
    F = int(freq[0].sum())
    assert F == 10000
    for k, d in enumerate(ids):
        # entropy1D over one step whose prefix count is the stored one
        s = O.entropy_1d(O.DIHEDRAL, [O.ML], 1, F, ext[k, 0], ext[k, 1], True,
                         False, [30], freq[k])
        assert abs(s[0] - S1[row[int(d)], step]) <= 1e-12 * abs(S1[row[int(d)], step])
    # stored bin mids are binData1D's, bit for bit: the offset grows by one
    # bin width per bin (min + j*w + w/2 differs in the last bits)
    w = (ext[:, 1] - ext[:, 0]) / 30
    mids = np.zeros((len(ext), 30))
    o = ext[:, 0].copy()
    for j in range(30):
        mids[:, j] = o + w / 2.0
        o = o + w
    np.testing.assert_array_equal(mids, golden["ex2_hist_mid"])
    for k in range(0, len(ext), 53):
        f, m = O.bin_data_1d(np.zeros(4, np.uint8), 1, [30], ext[k, 0], ext[k, 1])
        np.testing.assert_array_equal(m, golden["ex2_hist_mid"][k])


def test_counts_match_bruteforce(golden):
    codes = golden["ex1_codes_t"]
    rng = np.random.default_rng(1)
    pairs = rng.integers(0, codes.shape[0], size=(20, 2))
    got = O.pair_counts(codes, 30, pairs, NSTEPS)
    ss = codes.shape[1] // NSTEPS
    for p, (i, j) in enumerate(pairs):
        for s in range(NSTEPS):
            n = (s + 1) * ss
            want = np.bincount(codes[i, :n].astype(int) * 30 + codes[j, :n],
                               minlength=900).reshape(30, 30)
            assert np.array_equal(got[p, s], want)
        f, _ = O.bin_data_2d(codes[i], codes[j], NSTEPS, [30], 0, 1, 0, 1)
        assert np.array_equal(f.reshape(NSTEPS, 30, 30), got[p])
