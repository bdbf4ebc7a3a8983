"""Pins for the oracle functions the stored example outputs cannot reach
(no BAT inputs ship with the reference): closed forms, published tables and
hand-built data whose answer is known."""
import numpy as np
import pytest
from scipy import special

import oracle as O

TWO_PI = 6.283185307179586


# ---- Bessel functions (Abramowitz & Stegun 9.8.1-9.8.4 accuracy) ---------------
# The reference's I0 polynomial carries 1.2067429 for the t^6 coefficient
# where A&S 9.8.1 prints 1.2067492 (Utils.cpp:121-122); the restatement keeps
# the reference's value, so below x = 3.75 it sits 6.3e-6 * t^6 under the
# true I0 -- checked here, which pins that coefficient as well.
@pytest.mark.parametrize("n", [0, 1, 2, 3])
def test_bessel_against_scipy(n):
    for x in [0.01, 0.5, 1.0, 2.0, 3.7, 3.8, 5.0, 10.0, 20.0, 50.0]:
        got, want = O.bessel_i(n, x), special.iv(n, x)
        i0 = special.iv(0, x)
        shift = -6.3e-6 * (x / 3.75) ** 6 if x < 3.75 else 0.0
        if n == 0:
            tol = 2e-7 if x < 3.75 else 5e-7 * want   # 9.8.1 absolute, 9.8.2 relative
            assert abs(got - (want + shift)) <= tol, (n, x, got, want)
        elif n == 1:
            assert abs(got - want) <= 5e-7 * want, (n, x, got, want)
        else:
            # Miller recurrence normalised by the reference's I0; its fixed
            # 2*(n + int(sqrt(40 n))) terms stop converging beyond x ~ 20
            tol = 5e-7 if x <= 20 else 3e-5
            assert abs(got - want * (1 + shift / i0)) <= tol * want, (n, x, got, want)


def test_bandwidth_closed_form():
    k, n = 1.0, 10000
    want = (3 * n * k * k * special.iv(2, 2 * k) /
            (4 * np.sqrt(np.pi) * special.iv(0, k) ** 2)) ** 0.4
    assert abs(O.vmkde_bandwidth(n, k) - want) < 1e-6 * want


# ---- bond / angle binning ------------------------------------------------------
def test_bond_binning_closed_form():
    x = np.array([1.0, 1.25, 1.5, 1.75, 2.0], np.float32)
    codes, ext = O.real2int(x, O.BOND, [4])
    mn, mx = 1.0 - 5e-9, 2.0 + 5e-9
    assert ext[0] == mn and ext[1] == mx and ext[2] == 1.5 and ext[3] == 0
    w = (mx - mn) / 4
    assert codes[0].tolist() == [int(np.floor((v - mn) / w)) for v in x.astype(float)]
    assert codes[0].tolist() == [0, 1, 2, 2, 3]


def test_bond_min_clamp_and_narrow_range():
    x = np.full(10, 1.0, np.float32)
    codes, ext = O.real2int(x, O.BOND, [10])
    assert ext[1] - ext[0] > 0.05            # max += 0.05 when range < 1e-4
    assert (codes == 0).all()
    x = np.array([0.0, 1e-10, 1.0], np.float32)
    _, ext = O.real2int(x, O.BOND, [10])
    assert ext[0] == 0.0                      # min clamped at 0


def test_angle_clamped_at_pi_and_top_bin_folds():
    x = np.array([1.0, np.float32(np.pi), 2.0], np.float32)
    codes, ext = O.real2int(x, O.ANGLE, [10])
    assert ext[1] == np.pi
    # float32(pi) > pi, so (x - min)/w reaches nbins and folds to nbins - 1
    assert codes[0, 1] == 9


def test_codes_monotone_and_in_range():
    rng = np.random.default_rng(0)
    x = rng.normal(1.5, 0.03, 5000).astype(np.float32)
    for nb in (7, 30, 50, 255):
        codes, _ = O.real2int(x, O.BOND, [nb])
        c = codes[0]
        assert c.max() < nb and c.min() == 0 and c.max() == nb - 1
        o = np.argsort(x, kind="stable")
        assert (np.diff(c[o].astype(int)) >= 0).all()


# ---- torsion gap shift -----------------------------------------------------------
def test_gap_shift_single_arc():
    x = np.linspace(1.0, 2.0, 1000).astype(np.float32)
    m = 90
    bs = (TWO_PI + 5e-9) / m
    hi = int(np.floor(2.0 / bs))
    want = TWO_PI - (hi + 1 + 0.5) * bs     # run starts right after the arc
    assert abs(O.dih_gap_shift(x, m) - want) < 1e-15
    codes, ext = O.real2int(x, O.DIHEDRAL, [30], 30, True)
    assert abs(ext[3] - want) < 1e-15
    shifted = np.fmod(x.astype(float) + want, TWO_PI)
    assert abs(ext[0] - (shifted.min() - 5e-9)) < 1e-15
    o = np.argsort(x)
    assert (np.diff(codes[0][o].astype(int)) >= 0).all()


def test_gap_shift_wrapping_arc_becomes_contiguous():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(5.8, TWO_PI - 1e-3, 500),
                        rng.uniform(0.0, 0.4, 500)]).astype(np.float32)
    codes, ext = O.real2int(x, O.DIHEDRAL, [30], 30, True)
    # after the shift the data occupy one short arc: range well below pi
    assert ext[1] - ext[0] < 1.2
    c = codes[0]
    assert c.min() == 0 and c.max() == 29
    # values just below 2pi map below values just above 0 (order preserved
    # along the arc through the cut)
    lo = c[x.astype(float) > 6.2].max()
    hi = c[x.astype(float) < 0.1].min()
    assert lo <= hi


def test_gap_shift_full_occupancy_uses_least_populated_bin():
    m = 60
    bs = (TWO_PI + 5e-9) / m
    x = (np.arange(m * 10) % m + 0.5) * bs
    x = np.concatenate([x, np.full(5, 10.5 * bs)])      # bin 10 heavier
    x = x[~((x > 25 * bs) & (x < 26 * bs))][:-0] if False else x
    x = np.concatenate([x[np.floor(x / bs) != 37], np.full(3, 37.5 * bs)])
    # every bin populated; bin 37 holds 3 (the minimum)
    assert abs(O.dih_gap_shift(x.astype(np.float32), m) - (TWO_PI - 37.5 * bs)) < 1e-12


def test_dihedral_without_optimisation_picks_narrower_range():
    x = np.concatenate([np.linspace(5.9, 6.28, 50), np.linspace(0.0, 0.3, 50)])
    codes, ext = O.real2int(x.astype(np.float32), O.DIHEDRAL, [10], 10, False)
    assert ext[3] == np.pi          # pi-shifted range chosen
    assert ext[1] - ext[0] < 1.0


# ---- vmKDE -------------------------------------------------------------------------
def test_vmkde_planted_two_states():
    rng = np.random.default_rng(2)
    a = np.concatenate([rng.vonmises(np.deg2rad(60), 20, 4000),
                        rng.vonmises(np.deg2rad(240) - TWO_PI, 20, 4000)])
    a = np.mod(a, TWO_PI).astype(np.float32)
    codes, mins, nf = O.vmkde(a, nmaxconf=3, kappa=1.0, sdo_step=5,
                              conv=1e-4, max_iter=1000)
    assert nf == 2
    deg = np.rad2deg(mins)
    assert abs(deg[0] - 150) < 8 and abs(deg[1] - 330) < 8
    # states: frames between the minima are state 1, the rest state 0
    inside = (a > mins[0]) & (a < mins[1])
    assert (codes[inside] == 1).all() and (codes[~inside] == 0).all()


def test_vmkde_single_state_all_zero():
    rng = np.random.default_rng(3)
    a = np.mod(rng.vonmises(1.0, 30, 3000), TWO_PI).astype(np.float32)
    codes, mins, nf = O.vmkde(a, nmaxconf=3, kappa=1.0)
    assert nf == 1 and (codes == 0).all()
    assert abs(np.rad2deg(mins[0]) - (np.rad2deg(1.0) + 180) % 360) < 3


def test_vmkde_reference_minima_files_are_plausible(golden):
    for key, nmax in (("ex1v_minima", 3), ("ex2v_minima", 3)):
        rec = golden[key]
        for r in rec:
            n = int(r[1])
            assert 1 <= n <= nmax
            m = r[2:2 + n]
            assert (np.diff(m) > 0).all() and m.min() >= 0 and m.max() < TWO_PI


# ---- estimators: closed forms -----------------------------------------------------
def test_entropy_uniform_and_delta_closed_forms():
    nb, S, n = 8, 1, 8000
    f = np.full(nb, n // nb, np.uint64)
    w = 1.0
    s = O.entropy_1d(O.DIHEDRAL, [O.ML, O.MM, O.JS], S, n, 0.0, nb * w, False,
                     False, [nb], f)
    assert abs(s[0] - np.log(nb * w)) < 1e-13
    assert abs(s[1] - (np.log(nb) + (nb - 1) / (2 * n))) < 1e-13
    # JS on exactly uniform counts: sum (t - p)^2 = 0 makes lambda infinite
    # and every shrunk frequency NaN, which the p > 0 test skips -> 0
    assert s[2] == 0.0
    # JS on two bins: Hausser & Strimmer (2009) shrinkage towards 1/2
    n2, k = 1000, 100
    q = k / n2
    lam = (1 - ((1 - q) ** 2 + q ** 2)) / ((n2 - 1) * 2 * (0.5 - q) ** 2)
    ps = lam * 0.5 + (1 - lam) * np.array([1 - q, q])
    js = O.entropy_1d(O.DIHEDRAL, [O.JS], 1, n2, 0.0, 2.0, False, False, [2],
                      np.array([n2 - k, k], np.uint64))[0]
    assert abs(js - (-np.sum(ps * np.log(ps)))) < 1e-14
    d = np.zeros(nb, np.uint64)
    d[3] = n
    s = O.entropy_1d(O.DIHEDRAL, [O.ML, O.MM, O.CS], S, n, 0.0, nb * w, False,
                     False, [nb], d)
    assert abs(s[0]) < 1e-15 and abs(s[1]) < 1e-15
    assert abs(s[2]) < 1e-15                        # CS: no singletons


def test_entropy_2d_independent_is_sum_of_marginals():
    rng = np.random.default_rng(4)
    nb, n = 6, 6000
    a = rng.integers(0, nb, n).astype(np.uint8)
    b = rng.integers(0, nb, n).astype(np.uint8)
    f2, _ = O.bin_data_2d(a, b, 1, [nb], 1.0, 2.0, 0.5, 1.5)
    ca = np.bincount(a, minlength=nb).astype(float)
    cb = np.bincount(b, minlength=nb).astype(float)
    joint = f2.reshape(nb, nb).astype(float)
    s2 = O.entropy_2d(O.BOND, O.ANGLE, [O.ML], 1, n, 1.0, 2.0, 0.5, 1.5, True,
                      False, [nb], f2)[0]
    # direct evaluation of -sum p log(p / (w1 w2 J1 J2)) with numpy
    w1, w2 = 1.0 / nb, 1.0 / nb
    m1 = 1.0 + np.arange(nb) * w1 + w1 / 2
    m2 = 0.5 + np.arange(nb) * w2 + w2 / 2
    J = np.outer(m1 ** 2, np.sin(m2))
    p = joint / n
    k = p > 0
    direct = -np.sum(p[k] * np.log(p[k] / (w1 * w2 * J[k])))
    assert abs(s2 - direct) < 1e-12
    # MI = S1a + S1b - S2 >= 0 for the plug-in estimator
    s1a = O.entropy_1d(O.BOND, [O.ML], 1, n, 1.0, 2.0, True, False, [nb],
                       ca.astype(np.uint64))[0]
    s1b = O.entropy_1d(O.ANGLE, [O.ML], 1, n, 0.5, 1.5, True, False, [nb],
                       cb.astype(np.uint64))[0]
    assert s1a + s1b - s2 >= -1e-12


def test_shuffle_swap_semantics():
    data = np.arange(6, dtype=np.uint8)
    idx = np.array([2, 2, 0, 5, 4, 3], np.uint64)
    got = O.shuffle(idx, data)
    ref = list(range(6))
    for i, j in enumerate(idx):
        ref[i], ref[int(j)] = ref[int(j)], ref[i]
    assert got.tolist() == ref


def test_step_count_outside_uint8_is_refused():
    """Step counts are unsigned char in the reference; 0 and 256 are refused
    rather than wrapped."""
    codes = np.zeros((2, 256), np.uint8)
    ext = np.tile([1.0, 2.0, 0.0, 0.0], (2, 1))
    for bad in (0, 256):
        with pytest.raises(ValueError):
            O.dof_entropies(codes, np.ones(2, np.uint8), ext, bad, [4])
        with pytest.raises(ValueError):
            O.pair_counts(codes, 4, [[0, 1]], bad)
