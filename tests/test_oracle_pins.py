"""Closed-form pins for the estimator branches and the vmKDE grid of the
oracle that the stored example outputs (ML with Jacobian only) cannot reach.

Every expected value is evaluated here from the published definition of the
estimator, in 40-digit decimal arithmetic over explicit cell lists, with the
reference's own deviations from the textbook stated where they apply:

* Chao-Shen (Chao & Shen 2003): coverage C = 1 - f1/n (f1 = cells holding
  exactly one frame), p_i = C c_i / n, and the Horvitz-Thompson weight
  1 / (1 - (1 - p_i)^n) on every term (Estimators.cpp:288-310, 477-484,
  534-584).  In 2-D the reference takes the second DOF's Jacobian at the ROW
  bin, mid2[j] (Estimators.cpp:558, 562).
* James-Stein shrinkage (Hausser & Strimmer 2009):
  lambda = (1 - sum p^2) / ((n - 1) sum (t - p)^2), p' = lambda t +
  (1 - lambda) p.  The reference's 2-D target is t = 1/B, not 1/B^2
  (Estimators.cpp:508-530).
* Miller-Madow: ML + (nonzero cells - 1) / (2n) (Estimators.cpp:275-283,
  637-641).
* von Mises KDE (Real2Int.cpp:659-698): the density of one frame a is
  exp(h cos(X - a)) / (2 pi I0(h)); the grid holds its first and second
  derivatives averaged over frames.

Each case is built so that the plausible slip (coverage left out, exponent
n - 1, the column-bin Jacobian, target 1/B^2, a missing 1/N or h) changes
the value by far more than the tolerance; the slip is evaluated too and
asserted to be distinguishable.
"""
from decimal import Decimal, getcontext

import numpy as np
import pytest
from scipy import special

import oracle as O

getcontext().prec = 40
TWO_PI = 6.283185307179586


def _d(x) -> Decimal:
    return Decimal(repr(float(x))) if not isinstance(x, Decimal) else x


def _ln(x: Decimal) -> Decimal:
    return x.ln()


# ---- Chao-Shen ----------------------------------------------------------------
def _cs_cells(counts, n, widths_jac):
    """-sum_i p_i log(p_i / (w_i)) / (1 - (1 - p_i)^n) over cells with c > 0;
    widths_jac[i] is the cell's area times its Jacobian."""
    counts = [int(c) for c in counts]
    f1 = sum(1 for c in counts if c == 1)
    cov = Decimal(1) - Decimal(f1) / Decimal(n)
    s = Decimal(0)
    for c, wj in zip(counts, widths_jac):
        if c == 0:
            continue
        p = cov * Decimal(c) / Decimal(n)
        s -= p * _ln(p / _d(wj)) / (Decimal(1) - (Decimal(1) - p) ** n)
    return float(s)


def test_chao_shen_1d_singletons_hand_value():
    # counts {2, 1, 1, 0}: f1 = 2, n = 4, coverage 1/2, p = 1/4, 1/8, 1/8
    f = np.array([2, 1, 1, 0], np.uint64)
    got = O.entropy_1d(O.DIHEDRAL, [O.CS], 1, 4, 0.0, 2.0, False, False, [4], f)[0]
    # 40-digit evaluation of -sum p log(2p) / (1 - (1 - p)^4)
    assert abs(got - 1.0909955521612381) < 1e-15
    assert abs(got - _cs_cells(f, 4, [0.5] * 4)) < 1e-15
    # slips: no coverage factor; HT exponent n - 1; no HT weight
    no_cov = -sum((c / 4) * np.log(c / 4 / 0.5) / (1 - (1 - c / 4) ** 4) for c in (2, 1, 1))
    exp_m1 = -sum(p * np.log(p / 0.5) / (1 - (1 - p) ** 3) for p in (0.25, 0.125, 0.125))
    no_ht = -sum(p * np.log(p / 0.5) for p in (0.25, 0.125, 0.125))
    for slip in (no_cov, exp_m1, no_ht):
        assert abs(slip - got) > 1e-2


@pytest.mark.parametrize("dtype", [O.BOND, O.ANGLE])
def test_chao_shen_1d_jacobian(dtype):
    rng = np.random.default_rng(11)
    nb, mn, mx = 6, 1.1, 1.7
    f = rng.integers(0, 4, nb).astype(np.uint64)
    f[1], f[4] = 1, 1          # singletons present
    n = int(f.sum())
    got = O.entropy_1d(dtype, [O.CS], 1, n, mn, mx, True, False, [nb], f)[0]
    w = (mx - mn) / nb
    mid = mn + np.arange(nb) * w + w / 2
    jac = mid ** 2 if dtype == O.BOND else np.sin(mid)
    assert abs(got - _cs_cells(f, n, w * jac)) < 1e-14
    # without the Jacobian flag the factor is absent
    plain = O.entropy_1d(dtype, [O.CS], 1, n, mn, mx, False, False, [nb], f)[0]
    assert abs(plain - _cs_cells(f, n, [w] * nb)) < 1e-14
    assert abs(plain - got) > 1e-3


def test_chao_shen_2d_row_bin_jacobian_quirk():
    # bond (row DOF) x angle (column DOF), 2 x 2 cells, counts [[3, 1], [2, 0]]
    # (the off-diagonal counts differ, so moving the column Jacobian between
    # them changes the sum)
    c = np.array([[3, 1], [2, 0]], np.uint64)
    n = 6
    mn1, mx1, mn2, mx2 = 1.4, 1.6, 1.8, 2.0
    got = O.entropy_2d(O.BOND, O.ANGLE, [O.CS], 1, n, mn1, mx1, mn2, mx2, True,
                       False, [2], c.ravel())[0]
    w1, w2 = (mx1 - mn1) / 2, (mx2 - mn2) / 2
    m1 = mn1 + np.arange(2) * w1 + w1 / 2
    m2 = mn2 + np.arange(2) * w2 + w2 / 2
    quirk = [w1 * w2 * m1[j] ** 2 * np.sin(m2[j]) for j in range(2) for k in range(2)]
    textbook = [w1 * w2 * m1[j] ** 2 * np.sin(m2[k]) for j in range(2) for k in range(2)]
    want = _cs_cells(c.ravel(), n, quirk)
    assert abs(got - want) < 1e-14
    assert abs(got - _cs_cells(c.ravel(), n, textbook)) > 1e-3
    # the quirk makes the estimate depend on which DOF is the row
    ct = np.ascontiguousarray(c.T)
    swapped = O.entropy_2d(O.ANGLE, O.BOND, [O.CS], 1, n, mn2, mx2, mn1, mx1,
                           True, False, [2], ct.ravel())[0]
    quirk_t = [w2 * w1 * np.sin(m2[j]) * m1[j] ** 2 for j in range(2) for k in range(2)]
    assert abs(swapped - _cs_cells(ct.ravel(), n, quirk_t)) < 1e-14
    assert abs(swapped - got) > 1e-3


def test_chao_shen_2d_prefix_steps_use_their_own_coverage():
    # two steps: step 0 sees n = 4 with f1 = 2; step 1 (prefix) n = 8, f1 = 1
    f = np.array([[2, 1, 1, 0], [4, 3, 1, 0]], np.uint64)
    got = O.entropy_2d(O.DIHEDRAL, O.DIHEDRAL, [O.CS], 2, 4, 0.0, 1.0, 0.0, 1.0,
                       True, False, [2], f.ravel())
    a = 0.25
    for s, n in ((0, 4), (1, 8)):
        assert abs(got[s] - _cs_cells(f[s], n, [a] * 4)) < 1e-14


# ---- James-Stein ------------------------------------------------------------------
def _js_cells(counts, n, target, widths_jac):
    p = [Decimal(int(c)) / Decimal(n) for c in counts]
    t = _d(target)
    s2 = sum(x * x for x in p)
    tk = sum((t - x) * (t - x) for x in p)
    lam = (Decimal(1) - s2) / ((Decimal(n) - 1) * tk)
    s = Decimal(0)
    for x, wj in zip(p, widths_jac):
        q = lam * t + (Decimal(1) - lam) * x
        if q > 0:
            s -= q * _ln(q / _d(wj))
    return float(s), float(lam)


def test_james_stein_2d_target_one_over_bins():
    c = np.array([[40, 5, 1], [0, 30, 9], [2, 0, 13]], np.uint64)
    n = int(c.sum())
    got = O.entropy_2d(O.DIHEDRAL, O.DIHEDRAL, [O.JS], 1, n, 0.0, 3.0, 0.0, 1.5,
                       False, False, [3], c.ravel())[0]
    area = (3.0 / 3) * (1.5 / 3)
    want, lam = _js_cells(c.ravel(), n, 1.0 / 3, [area] * 9)
    assert 0 < lam < 1
    assert abs(got - want) < 1e-14
    # shrinking toward 1/B^2 (the textbook target) gives another value
    textbook, _ = _js_cells(c.ravel(), n, 1.0 / 9, [area] * 9)
    assert abs(got - textbook) > 1e-4
    # empty cells are shrunk above zero and contribute
    ml = O.entropy_2d(O.DIHEDRAL, O.DIHEDRAL, [O.ML], 1, n, 0.0, 3.0, 0.0, 1.5,
                      False, False, [3], c.ravel())[0]
    assert got > ml


def test_james_stein_2d_with_jacobian():
    rng = np.random.default_rng(12)
    nb = 4
    c = rng.integers(0, 20, (nb, nb)).astype(np.uint64)
    n = int(c.sum())
    mn1, mx1, mn2, mx2 = 1.45, 1.55, 1.85, 1.95
    got = O.entropy_2d(O.BOND, O.ANGLE, [O.JS], 1, n, mn1, mx1, mn2, mx2, True,
                       False, [nb], c.ravel())[0]
    w1, w2 = (mx1 - mn1) / nb, (mx2 - mn2) / nb
    m1 = mn1 + np.arange(nb) * w1 + w1 / 2
    m2 = mn2 + np.arange(nb) * w2 + w2 / 2
    wj = [w1 * w2 * m1[j] ** 2 * np.sin(m2[k]) for j in range(nb) for k in range(nb)]
    want, _ = _js_cells(c.ravel(), n, 1.0 / nb, wj)
    assert abs(got - want) < 1e-13 * max(1, abs(want))


def test_james_stein_1d_target_one_over_bins():
    f = np.array([50, 20, 0, 5, 25], np.uint64)
    n = int(f.sum())
    got = O.entropy_1d(O.DIHEDRAL, [O.JS], 1, n, 0.0, 5.0, False, False, [5], f)[0]
    want, _ = _js_cells(f, n, 0.2, [1.0] * 5)
    assert abs(got - want) < 1e-14


# ---- Miller-Madow -----------------------------------------------------------------
def test_miller_madow_2d_uniform_table():
    nb, per = 5, 40
    n = nb * nb * per
    c = np.full(nb * nb, per, np.uint64)
    mn1, mx1, mn2, mx2 = 0.0, 2.5, 0.0, 0.5
    got = O.entropy_2d(O.DIHEDRAL, O.DIHEDRAL, [O.ML, O.MM], 1, n, mn1, mx1, mn2,
                       mx2, True, False, [nb], c)
    area = (mx1 - mn1) / nb * (mx2 - mn2) / nb
    assert abs(got[0] - np.log(nb * nb * area)) < 1e-13
    assert abs(got[1] - (np.log(nb * nb * area) + (nb * nb - 1) / (2.0 * n))) < 1e-13


def test_miller_madow_2d_counts_only_nonzero_cells():
    c = np.array([[6, 0, 2], [0, 0, 4], [1, 3, 0]], np.uint64)
    n = int(c.sum())
    got = O.entropy_2d(O.DIHEDRAL, O.DIHEDRAL, [O.ML, O.MM], 1, n, 0.0, 3.0, 0.0,
                       3.0, False, False, [3], c.ravel())
    p = c.ravel()[c.ravel() > 0] / n
    ml = -np.sum(p * np.log(p))
    assert abs(got[0] - ml) < 1e-14
    assert abs(got[1] - (ml + (5 - 1) / (2.0 * n))) < 1e-14
    # counting every cell instead of the nonzero ones is off by 4 / (2n)
    assert abs(got[1] - (ml + (9 - 1) / (2.0 * n))) > 1e-3


def test_miller_madow_2d_bond_angle_jacobian():
    rng = np.random.default_rng(13)
    nb = 3
    c = rng.integers(0, 9, (nb, nb)).astype(np.uint64)
    c[0, 1] = 0
    n = int(c.sum())
    mn1, mx1, mn2, mx2 = 1.3, 1.9, 1.6, 2.2
    got = O.entropy_2d(O.BOND, O.ANGLE, [O.MM], 1, n, mn1, mx1, mn2, mx2, True,
                       False, [nb], c.ravel())[0]
    w1, w2 = (mx1 - mn1) / nb, (mx2 - mn2) / nb
    m1 = mn1 + np.arange(nb) * w1 + w1 / 2
    m2 = mn2 + np.arange(nb) * w2 + w2 / 2
    J = np.outer(m1 ** 2, np.sin(m2))
    p = c / n
    k = p > 0
    want = -np.sum(p[k] * np.log(p[k] / (w1 * w2 * J[k]))) + (k.sum() - 1) / (2.0 * n)
    assert abs(got - want) < 1e-13


# ---- von Mises KDE grid -------------------------------------------------------------
def _grid(x, kappa, G):
    g = O.vmkde_grid(np.asarray(x, np.float32), kappa, G)
    return g[:G + 1], g[G + 1:]


def test_vmkde_grid_single_frame_is_the_von_mises_derivative():
    a = np.float32(1.3)
    kappa, G = 0.4, 360
    h = O.vmkde_bandwidth(1, kappa)
    gr, sd = _grid([a], kappa, G)
    X = np.arange(G + 1) * (360.0 / G) * np.pi / 180
    d = X - float(a)
    e = np.exp(h * np.cos(d)) / (TWO_PI * special.i0(h))
    want_g = -h * np.sin(d) * e
    want_s = h * (h * np.sin(d) ** 2 - np.cos(d)) * e
    # the reference's I0 (Abramowitz & Stegun 9.8.1/9.8.2, its own digits)
    # sits within 2e-7 relative of the exact I0 at this h
    scale = np.max(np.abs(want_g))
    assert np.max(np.abs(gr - want_g)) < 1e-6 * scale
    assert np.max(np.abs(sd - want_s)) < 1e-6 * np.max(np.abs(want_s))
    # with the reference's own I0 the agreement is to rounding
    e_ref = np.exp(h * np.cos(d)) / (TWO_PI * O.bessel_i(0, h))
    assert np.max(np.abs(gr - (-h * np.sin(d) * e_ref))) < 1e-13 * scale


def test_vmkde_grid_averages_frames():
    rng = np.random.default_rng(14)
    a = rng.uniform(0, TWO_PI, 7).astype(np.float32)
    kappa, G = 2.0, 72
    h = O.vmkde_bandwidth(len(a), kappa)
    gr, sd = _grid(a, kappa, G)
    X = np.arange(G + 1) * (360.0 / G) * np.pi / 180
    d = X[:, None] - a.astype(float)[None, :]
    e = np.exp(h * np.cos(d)) / (TWO_PI * O.bessel_i(0, h))
    want_g = np.mean(-h * np.sin(d) * e, axis=1)
    want_s = np.mean(h * (h * np.sin(d) ** 2 - np.cos(d)) * e, axis=1)
    assert np.max(np.abs(gr - want_g)) < 1e-13 * np.max(np.abs(want_g))
    assert np.max(np.abs(sd - want_s)) < 1e-13 * np.max(np.abs(want_s))


def test_vmkde_grid_invariants_on_a_mixture():
    rng = np.random.default_rng(15)
    a = np.mod(np.concatenate([rng.vonmises(1.0, 6, 600), rng.vonmises(4.0, 3, 400)]),
               TWO_PI).astype(np.float32)
    kappa, G = 1.0, 1440
    gr, sd = _grid(a, kappa, G)
    dx = TWO_PI / G
    # periodic: the last point repeats the first
    assert abs(gr[G] - gr[0]) < 1e-12 * np.max(np.abs(gr))
    # the density is periodic, so its derivatives integrate to zero over the
    # circle (trapezoid rule, spectrally accurate for periodic functions)
    assert abs(np.sum(gr[:G]) * dx) < 1e-12 * np.max(np.abs(gr))
    assert abs(np.sum(sd[:G]) * dx) < 1e-12 * np.max(np.abs(sd))
    # the second derivative is the derivative of the gradient (centred
    # differences, O(dx^2))
    fd = (np.roll(gr[:G], -1) - np.roll(gr[:G], 1)) / (2 * dx)
    assert np.max(np.abs(fd - sd[:G])) < 2e-3 * np.max(np.abs(sd))
    # and integrating the gradient recovers a density whose mean is 1 / 2pi
    # up to the constant: check the cumulative integral closes on itself
    cum = np.cumsum(gr[:G]) * dx
    assert abs(cum[-1]) < 1e-12 * np.max(np.abs(cum))


def test_vmkde_states_symmetric_mixture_exact_minima():
    # two identical clusters mirrored about 150.5 and 330.5 degrees: the
    # maxima straddle 60.5 and 240.5 on the 1-degree grid, the initial minima
    # are their midpoints, and the interpolated gradient there is zero by
    # symmetry, so the descent stops at once
    rng = np.random.default_rng(16)
    off = np.abs(rng.normal(0, np.deg2rad(12), 3000))
    c0, c1 = np.deg2rad(60.5), np.deg2rad(240.5)
    a = np.concatenate([c0 + off, c0 - off, c1 + off, c1 - off])
    a32 = np.mod(a, TWO_PI).astype(np.float32)
    codes, mins, nf = O.vmkde(a32, nmaxconf=3, kappa=1.0, sdo_step=5, conv=1e-4,
                              max_iter=1000)
    assert nf == 2
    # float32 rounding of the frames breaks the mirror symmetry at 1e-7
    assert abs(np.rad2deg(mins[0]) - 150.5) < 1e-3
    assert abs(np.rad2deg(mins[1]) - 330.5) < 1e-3
    inside = (a32 > mins[0]) & (a32 < mins[1])
    assert (codes[inside] == 1).all() and (codes[~inside] == 0).all()
