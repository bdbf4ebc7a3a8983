/*
 * centre_oracle.c -- CPU restatement of CENTRE's entropy hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path and
 * the CPU baseline that the benchmark times beside it.  Nothing in the
 * shipped library links or calls it; only tests/, the smoke check and the
 * benchmark's CPU-baseline leg load it.
 *
 * Every function restates one reference function (file:line given in its
 * comment) from its published behaviour, in plain C, loop for loop where
 * the loop order matters for floating-point results.  Reference paths are
 * relative to the CENTRE source tree.
 *
 * Pinning (what checks each function against something independent):
 *   or_bin_data_1d / or_bin_data_2d  : brute-force bincount + golden
 *                                      example-2 hist_tor-1d counts.
 *   or_entropy_1d / or_entropy_2d    : golden entcontri_*.nc (example-1,
 *                                      all 75 S1 and 2533 S2 records x 4
 *                                      steps, ML + Jacobian) and the
 *                                      example-2 stored counts; MM/CS/JS by
 *                                      closed forms: uniform and delta
 *                                      distributions, and 40-digit
 *                                      evaluations of Chao-Shen with
 *                                      singletons (1-D, 2-D with the row-bin
 *                                      Jacobian, prefix steps), 2-D
 *                                      James-Stein (target 1/B) and 2-D
 *                                      Miller-Madow, each also showing the
 *                                      plausible slip fails.
 *   or_bnd/ang_real2int              : closed form, invariants and the
 *                                      example-1 stored extrema ranges.
 *   or_dih_real2int                  : hand-built arcs with known gaps.
 *                                      On real data: parity unpinned (the
 *                                      reference ships no BAT inputs).
 *   or_bessel_i                      : Abramowitz & Stegun tables (scipy).
 *   or_vmkde                         : the grid against the analytic von
 *                                      Mises derivatives (one frame), their
 *                                      mean (several frames) and its
 *                                      invariants (periodicity, zero
 *                                      integrals, second derivative as the
 *                                      derivative of the first); the states
 *                                      against planted multi-modal data and
 *                                      a mirrored mixture with exact minima;
 *                                      stored minima only bound
 *                                      plausibility.  No stored trajectory
 *                                      reaches it with real data.
 *   or_boruvka                       : Kruskal under the same total order,
 *                                      golden tree CSVs (example-1/2).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_DX 5.0e-9                      /* configcentre.h:22 */
static const double OR_PI = 3.141592653589793;     /* Constants.h:16 */
static const double OR_TWOPI = 6.283185307179586;  /* Constants.h:18 */
static const double OR_SQRTPI = 1.7724538509055159;
static const double OR_DEGRAD = 0.017453292519943295;

enum { T_NONE = 0, T_BOND = 1, T_ANGLE = 2, T_DIH = 3 };   /* Centre.h:42 */
enum { E_ML = 0, E_MM = 1, E_CS = 2, E_JS = 3 };           /* Centre.h:89 */

/* ------------------------------------------------------------------------ */
/* Real2Int::shuffleArray (Real2Int.cpp:12-25): in-place sequential swaps.  */
int or_shuffle(const uint64_t *rand_idx, uint8_t *data, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        uint8_t t = data[i];
        data[i] = data[rand_idx[i]];
        data[rand_idx[i]] = t;
    }
    return 0;
}

/* Shared tail of bnd/angReal2Int (Real2Int.cpp:52-68, 160-182). */
static void or_bin_range(const float *x, uint64_t n, const uint8_t *schemes,
                         int nsch, double mn, double mx, uint8_t *codes) {
    for (int b = 0; b < nsch; ++b) {
        double w = (mx - mn) / (int)schemes[b];
        uint8_t *out = codes + (uint64_t)b * n;
        for (uint64_t i = 0; i < n; ++i) {
            uint8_t ti = (uint8_t)floor(((double)x[i] - mn) / w);
            out[i] = (ti == schemes[b]) ? (uint8_t)(ti - 1) : ti;
        }
    }
}

/* Real2Int::bndReal2Int (Real2Int.cpp:27-73).  ext = {min,max,avg,0}. */
int or_bnd_real2int(const float *x, uint64_t n, const uint8_t *schemes,
                    int nsch, uint8_t *codes, double *ext) {
    double mn = 1.0e100, mx = -1.0e100, sum = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        sum += x[i];
        if (mn > x[i]) mn = x[i];
        if (mx < x[i]) mx = x[i];
    }
    mn -= OR_DX;
    if (mn < 0.0) mn = 0.0;
    mx += OR_DX;
    if ((mx - mn) < 1.0e-4) mx += 0.05;
    or_bin_range(x, n, schemes, nsch, mn, mx, codes);
    ext[0] = mn; ext[1] = mx; ext[2] = sum / (double)n; ext[3] = 0.0;
    return 0;
}

/* Real2Int::angReal2Int (Real2Int.cpp:131-184): as bonds, max clamped to pi. */
int or_ang_real2int(const float *x, uint64_t n, const uint8_t *schemes,
                    int nsch, uint8_t *codes, double *ext) {
    double mn = 1.0e100, mx = -1.0e100, sum = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        sum += x[i];
        if (mn > x[i]) mn = x[i];
        if (mx < x[i]) mx = x[i];
    }
    mn -= OR_DX;
    if (mn < 0.0) mn = 0.0;
    mx += OR_DX;
    if (mx > OR_PI) mx = OR_PI;
    if ((mx - mn) < 1.0e-4) mx += 0.05;
    or_bin_range(x, n, schemes, nsch, mn, mx, codes);
    ext[0] = mn; ext[1] = mx; ext[2] = sum / (double)n; ext[3] = 0.0;
    return 0;
}

/* Gap search of Real2Int::dihReal2Int (Real2Int.cpp:280-350): returns the
 * phase shift 2*pi - centre[minidx] that moves the largest empty circular
 * run of the m-bin histogram (or the least-populated bin) to the cut. */
double or_dih_gap_shift(const float *x, uint64_t n, int m) {
    const double lo = 0.0, hi = OR_TWOPI + OR_DX;
    const double bs = (hi - lo) / m;
    uint64_t *prob = (uint64_t *)calloc((size_t)m, sizeof(uint64_t));
    unsigned char *zh = (unsigned char *)calloc((size_t)2 * m, 1);
    for (uint64_t i = 0; i < n; ++i) {
        int t1 = (int)fabs(floor(((double)x[i] - lo) / bs));
        uint8_t code = (uint8_t)t1;            /* stored through a u8 */
        if (code >= m) {                       /* reference: out of bounds */
            free(prob);
            free(zh);
            return NAN;
        }
        prob[code]++;
    }
    unsigned nzero = 0;
    for (int i = 0; i < m; ++i) {
        zh[i] = prob[i] == 0;
        nzero += zh[i];
        zh[m + i] = 0;
    }
    unsigned minidx = 0;
    if (nzero == 0) {
        double rec = (double)prob[0];
        for (int j = 1; j < m; ++j)
            if ((double)prob[j] < rec) { rec = (double)prob[j]; minidx = j; }
    } else {
        int j = 0;
        while (zh[j] == 0 && j < m) ++j;
        minidx = j;
        int first = j;
        for (int k = 0; k < m; ++k) zh[k + m] = zh[k];
        unsigned cnt = 0, rec = 0;
        for (j = first; j < 2 * m; ++j) {
            if (zh[j]) {
                cnt++;
            } else {
                if (cnt > rec) { minidx = j - cnt; rec = cnt; }
                cnt = 0;
            }
        }
    }
    double centre = lo + (minidx + 0.5) * bs;
    free(prob);
    free(zh);
    return OR_TWOPI - centre;
}

/* Real2Int::dihReal2Int (Real2Int.cpp:244-479).  ext = {min,max,avg,phase}. */
int or_dih_real2int(const float *x, uint64_t n, const uint8_t *schemes,
                    int nsch, uint8_t nbins_ref, int optimize, uint8_t *codes,
                    double *ext) {
    int m = (3 * (int)nbins_ref > 180) ? 180 : 3 * (int)nbins_ref;
    double displac = 0.0;
    if (optimize) {
        displac = or_dih_gap_shift(x, n, m);
        if (isnan(displac)) return -1;      /* torsion outside [-2pi, 2pi] */
    }
    /* initial values as the reference leaves them (Real2Int.cpp:269-276) */
    double mn1 = 6.5, mx1 = 0.0, mn2 = 6.5, mx2 = 0.0;
    double mn3 = 1.0e100, mx3 = 0.0, sum = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        double v = x[i];
        sum += x[i];
        if (mn1 > v) mn1 = v;
        if (mx1 < v) mx1 = v;
        double s = fmod(v + OR_PI, OR_TWOPI);
        if (s < mn2) mn2 = s;
        if (s > mx2) mx2 = s;
        if (optimize) {
            double o = fmod(v + displac, OR_TWOPI);
            if (o < mn3) mn3 = o;
            if (o > mx3) mx3 = o;
        }
    }
    double mn, mx;
    int dp;
    if (!optimize) {
        if (fabs((mx1 - mn1) - (mx2 - mn2)) <= 1.0e-1) { mn = mn1; mx = mx1; dp = 0; }
        else if ((mx1 - mn1) < (mx2 - mn2)) { mn = mn1; mx = mx1; dp = 0; }
        else { mn = mn2; mx = mx2; dp = 1; }
    } else {
        mn = mn3; mx = mx3; dp = 2;
    }
    mn -= OR_DX;
    if (mn < 0.0) mn = 0.0;
    mx += OR_DX;
    if (mx > OR_TWOPI) mx = OR_TWOPI;
    if ((mx - mn) < 1.0e-4) mx += 0.05;
    double avg = sum / (double)n;
    if (dp == 1) avg = fmod(avg + OR_PI, OR_TWOPI);
    else if (dp == 2) avg = fmod(avg + displac, OR_TWOPI);
    ext[0] = mn; ext[1] = mx; ext[2] = avg;
    ext[3] = optimize ? displac : dp * OR_PI;
    double shift = (dp == 1) ? OR_PI : displac;
    for (int b = 0; b < nsch; ++b) {
        double w = (mx - mn) / (int)schemes[b];
        uint8_t *out = codes + (uint64_t)b * n;
        for (uint64_t j = 0; j < n; ++j) {
            double v = (dp == 0) ? (double)x[j] : fmod((double)x[j] + shift, OR_TWOPI);
            uint8_t tt = (uint8_t)floor((v - mn) / w);
            out[j] = (tt == schemes[b]) ? (uint8_t)(tt - 1) : tt;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Modified Bessel functions I0, I1 (Utils.cpp:118-168): the polynomial
 * approximations of Abramowitz & Stegun 9.8.1-9.8.4; In for n >= 2 by
 * Miller's downward recurrence normalised with I0 (Utils.cpp:60-116). */
static double or_i0(double x) {
    if (fabs(x) < 3.75) {
        double y = x / 3.75;
        y = y * y;
        double t = 1.2067429 + y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2));
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * t));
    }
    double ax = fabs(x), y = 3.75 / ax;
    double bx = exp(ax) / sqrt(ax);
    double t = -0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2));
    double p = 0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * t))));
    return p * bx;
}

static double or_i1(double x) {
    if (fabs(x) < 3.75) {
        double y = x / 3.75;
        y = y * y;
        double t = 0.15084934 + y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3));
        return x * (0.5 + y * (0.87890594 + y * (0.51498869 + y * t)));
    }
    double ax = fabs(x), y = 3.75 / ax;
    double bx = exp(ax) / sqrt(ax);
    double t = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 + y * -0.420059e-2));
    double p = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y * (-0.1031555e-1 + y * t))));
    return p * bx;
}

double or_bessel_i(int n, double x) {
    if (n == 0) return or_i0(x);
    if (n == 1) return or_i1(x);
    if (fabs(x) < 1.e-10) return 0.0;
    double tox = 2.0 / x, bip = 0.0, bi = 1.0, res = 0.0;
    int iters = 2 * (n + (int)sqrtf((float)(40 * n)));
    for (int i = iters; i >= 1; --i) {
        double bim = bip + (double)(i * tox * bi);
        bip = bi;
        bi = bim;
        if (fabs(bi) > 1.e10) { bi *= 1.e-10; bip *= 1.e-10; res *= 1.e-10; }
        if (i == n) res = bip;
    }
    return res * or_i0(x) / bi;
}

/* Kernel bandwidth of getMinimaPositions (Real2Int.cpp:509-514). */
double or_vmkde_bandwidth(uint64_t nframes, double kappa) {
    double i0 = or_bessel_i(0, kappa);
    return pow((3 * (double)nframes * (kappa * kappa) * or_bessel_i(2, 2 * kappa)
                / (4 * OR_SQRTPI * pow(i0, 2.0))), (2.0 / 5.0));
}

/* Real2Int::vonMises (Real2Int.cpp:659-698): gradient and second derivative
 * of the von Mises KDE at X (radians), normalised by 2*pi*N*I0(h). */
void or_vonmises(double h, double X, const float *a, uint64_t n,
                 double *grad, double *sec) {
    double g = 0.0, s = 0.0;
    for (uint64_t j = 0; j < n; ++j) {
        double d = X - (double)a[j];
        double c = cos(d), sn = sin(d);
        double e = exp(h * c);
        g = g - sn * e;
        s = s + (h * sn * sn - c) * e;
    }
    double denom = OR_TWOPI * (double)n * or_bessel_i(0, h);
    *grad = g * h / denom;
    *sec = s * h / denom;
}

/* Real2Int::getMinimaPositions (Real2Int.cpp:481-657), histogram-free
 * conformer-state discretizer.  The reference evaluates the KDE on the
 * 361-point grid 0..360 degrees (grid_intervals = 360).  grid_intervals
 * = G evaluates it at G+1 points spaced 360/G degrees; every degree-valued
 * rule (maxima midpoints, 5-degree step clip, wrap at 360) is unchanged and
 * the interpolated gradient uses the grid spacing, so G = 360 is exactly
 * the reference.  The gradient/second-derivative grid is an input (2*(G+1)
 * doubles: gradients then second derivatives, from or_vmkde_grid or from
 * the device grid kernel) so that one state rule serves both.
 * codes[] must be zeroed by the caller (Discretizer.cpp:94-95). */
int or_vmkde_states(const float *x, uint64_t n, uint8_t nmaxconf,
                    uint32_t sdo_step, double conv, uint32_t max_iter,
                    uint32_t grid_intervals, const double *grid,
                    double *minima, uint8_t *nfound, uint8_t *codes) {
    const int G = (int)grid_intervals;
    const double hdeg = 360.0 / G;
    double *gr = (double *)malloc(sizeof(double) * (G + 1));
    double *sd = (double *)malloc(sizeof(double) * (G + 1));
    double cmax[9], smax[9] = {0};
    int nobs = 0;
    if (!grid || G < 2) {
        free(gr);
        free(sd);
        return -3;
    }
    memcpy(gr, grid, sizeof(double) * (G + 1));
    memcpy(sd, grid + (G + 1), sizeof(double) * (G + 1));
    for (int i = 0; i <= G; ++i) {
        if (i > 0 && (gr[i - 1] * gr[i]) <= 0.0 && sd[i] < -1.0e-4 && nobs < 9) {
            cmax[nobs] = (i * hdeg) / 2.0 + ((i - 1) * hdeg) / 2.0;
            smax[nobs] = sd[i] / 2.0 + sd[i - 1] / 2.0;
            nobs++;
        }
    }
    double t;
    for (int i = 0; i < nobs - 1; ++i)
        for (int j = i + 1; j < nobs; ++j)
            if (smax[i] > smax[j]) {
                t = smax[i]; smax[i] = smax[j]; smax[j] = t;
                t = cmax[i]; cmax[i] = cmax[j]; cmax[j] = t;
            }
    int nf = (nmaxconf < nobs) ? nmaxconf : nobs;
    for (int i = 0; i < nf - 1; ++i)
        for (int j = i + 1; j < nf; ++j)
            if (cmax[i] > cmax[j]) { t = cmax[i]; cmax[i] = cmax[j]; cmax[j] = t; }
    double mins[9] = {0};
    if (nf == 1) {
        mins[0] = cmax[0] - 180.0;
        if (mins[0] < 0.0) mins[0] += 360.0;
    } else {
        for (int i = 0; i < nf; ++i) {
            if (i == 0) {
                mins[i] = (cmax[i] + cmax[nf - 1] - 360.0) / 2.0;
                if (mins[i] < 0.0) mins[i] += 360;
            } else {
                mins[i] = (cmax[i] + cmax[i - 1]) / 2.0;
            }
        }
    }
    if (nf > 1) {
        for (int m = 0; m < nf; ++m) {
            int it = 0, conv_ok = 0;
            while (!conv_ok && it < (int)max_iter) {
                it++;
                double u = mins[m] / hdeg;
                int iu = (int)u;
                double g = fabs(u - 1 - iu) * gr[iu] + fabs(u - iu) * gr[1 + iu];
                if (fabs(g) < conv) {
                    conv_ok = 1;
                } else {
                    double st;
                    if (fabs(sdo_step * g) > 5.0)
                        st = ((sdo_step * g) / fabs(sdo_step * g)) * 5.0;
                    else
                        st = sdo_step * g;
                    mins[m] = mins[m] - st;
                    if (mins[m] < 0.0) mins[m] += 360.0;
                    if (mins[m] >= 360.0) mins[m] -= 360.0;
                }
            }
        }
    }
    for (int i = 0; i < nf - 1; ++i)
        for (int j = i + 1; j < nf; ++j)
            if (mins[i] > mins[j]) { t = mins[i]; mins[i] = mins[j]; mins[j] = t; }
    for (int i = 0; i < nf; ++i) mins[i] *= OR_DEGRAD;
    for (uint64_t k = 0; k < n; ++k) {
        if (nf == 1) {
            codes[k] = 0;
        } else {
            double v = x[k];
            for (int j = 0; j < nf; ++j) {
                if (j == 0) {
                    if (v < mins[0] || v > mins[nf - 1]) codes[k] = 0;
                } else if (v < mins[j] && v > mins[j - 1]) {
                    codes[k] = (uint8_t)j;
                }
            }
        }
    }
    for (int i = 0; i < nf; ++i) minima[i] = mins[i];
    *nfound = (uint8_t)nf;
    free(gr);
    free(sd);
    return 0;
}

/* Grid part of getMinimaPositions (Real2Int.cpp:518-532): gradient and
 * second derivative at the G+1 grid points.  out = 2*(G+1) doubles. */
int or_vmkde_grid(const float *x, uint64_t n, double kappa,
                  uint32_t grid_intervals, double *out) {
    const int G = (int)grid_intervals;
    double h = or_vmkde_bandwidth(n, kappa);
    for (int i = 0; i <= G; ++i) {
        double X = (i * (360.0 / G)) * OR_DEGRAD;
        or_vonmises(h, X, x, n, &out[i], &out[G + 1 + i]);
    }
    return 0;
}

/* Whole vmKDE discretizer for one torsion: grid + states. */
int or_vmkde(const float *x, uint64_t n, uint8_t nmaxconf, double kappa,
             uint32_t sdo_step, double conv, uint32_t max_iter,
             uint32_t grid_intervals, double *minima, uint8_t *nfound,
             uint8_t *codes) {
    double *grid = (double *)malloc(sizeof(double) * 2 * (grid_intervals + 1));
    or_vmkde_grid(x, n, kappa, grid_intervals, grid);
    memset(codes, 0, n);
    int r = or_vmkde_states(x, n, nmaxconf, sdo_step, conv, max_iter,
                            grid_intervals, grid, minima, nfound, codes);
    free(grid);
    return r;
}

/* ------------------------------------------------------------------------ */
/* binData1D (Estimators.cpp:56-118): per-step-chunk counts, then a prefix
 * sum over steps.  freqs (caller-zeroed) = nsteps * sum(schemes) uint64,
 * mids = sum(schemes) doubles. */
int or_bin_data_1d(uint8_t nsteps, const uint8_t *schemes, int nsch,
                   uint64_t nframes, double mn, double mx,
                   const uint8_t *data, uint64_t *freqs, double *mids) {
    size_t bsum = 0;
    for (int i = 0; i < nsch; ++i) bsum += schemes[i];
    int off = 0;
    for (int s = 0; s < nsch; ++s) {
        int nb = schemes[s];
        double w = (mx - mn) / nb, hw = w / 2.0, o = mn;
        for (int i = 0; i < nb; ++i) { mids[off + i] = o + hw; o += w; }
        off += nb;
    }
    uint64_t ss = nframes / nsteps, fshift = 0;
    unsigned bshift = 0;
    for (int s = 0; s < nsch; ++s) {
        int nb = schemes[s];
        for (int i = 0; i < nsteps; ++i) {
            uint64_t o = (uint64_t)i * ss;
            size_t of = bsum * i + bshift;
            for (uint64_t j = 0; j < ss; ++j) freqs[of + data[fshift + o + j]] += 1;
        }
        for (int i = 1; i < nsteps; ++i) {
            size_t of = bsum * i + bshift;
            for (int j = 0; j < nb; ++j) freqs[of + j] += freqs[of - bsum + j];
        }
        fshift += nframes;
        bshift += nb;
    }
    return 0;
}

/* binData2D (Estimators.cpp:130-205) with get2D21DIndex (:120-128):
 * joint index b1*nbins + b2.  freqs = nsteps * sum(b^2) uint64 (zeroed),
 * mids = 2*sum(b) doubles laid out per scheme as [mid1(b) | mid2(b)]. */
int or_bin_data_2d(uint8_t nsteps, const uint8_t *schemes, int nsch,
                   uint64_t nframes, double mn1, double mx1, double mn2,
                   double mx2, const uint8_t *d1, const uint8_t *d2,
                   uint64_t *freqs, double *mids) {
    size_t bsq = 0;
    for (int i = 0; i < nsch; ++i) bsq += (size_t)schemes[i] * schemes[i];
    int off = 0;
    for (int s = 0; s < nsch; ++s) {
        int nb = schemes[s];
        double w1 = (mx1 - mn1) / nb, h1 = w1 / 2, o1 = mn1;
        double w2 = (mx2 - mn2) / nb, h2 = w2 / 2, o2 = mn2;
        for (int i = 0; i < nb; ++i) {
            mids[off + i] = o1 + h1; o1 += w1;
            mids[off + nb + i] = o2 + h2; o2 += w2;
        }
        off += 2 * nb;
    }
    uint64_t ss = nframes / nsteps, fshift = 0;
    unsigned bshift = 0;
    for (int s = 0; s < nsch; ++s) {
        int nb = schemes[s];
        size_t nb2 = (size_t)nb * nb;
        for (int i = 0; i < nsteps; ++i) {
            uint64_t o = (uint64_t)i * ss;
            size_t of = bsq * i + bshift;
            for (uint64_t j = 0; j < ss; ++j) {
                size_t idx = (size_t)d1[fshift + o + j] * nb + d2[fshift + o + j];
                freqs[of + idx] += 1;
            }
        }
        for (int i = 1; i < nsteps; ++i) {
            size_t of = bsq * i + bshift;
            for (size_t j = 0; j < nb2; ++j) freqs[of + j] += freqs[of - bsq + j];
        }
        fshift += nframes;
        bshift += nb2;
    }
    return 0;
}

static double or_jac(int type, double mid) {
    if (type == T_BOND) return mid * mid;   /* pow(mid, 2) */
    if (type == T_ANGLE) return sin(mid);
    return 1.0;
}

/* entropy1D (Estimators.cpp:217-389).  entropy = nest*nsch*nsteps. */
int or_entropy_1d(int type, const int *ests, int nest, uint8_t nsteps,
                  uint64_t step_size, double mn, double mx, int jacobian,
                  int kde, const uint8_t *schemes, int nsch,
                  const uint64_t *freqs, double *entropy) {
    unsigned bsum = 0;
    for (int i = 0; i < nsch; ++i) bsum += schemes[i];
    double fr[256], mid[256];
    for (int e = 0; e < nest; ++e) {
        int est = ests[e];
        unsigned eoff = (unsigned)e * nsteps * nsch;
        unsigned foff = 0;
        for (int s = 0, entoff = 0; s < nsch; entoff += nsteps, ++s) {
            int nb = schemes[s];
            double w = (mx - mn) / nb, hw = w / 2;
            if (kde) w = 1.0;
            int hasj = (type == T_BOND || type == T_ANGLE);
            for (int j = 0; j < nb; ++j) mid[j] = hasj ? mn + j * w + hw : 0.0;
            for (int i = 0; i < nsteps; ++i) {
                const uint64_t *f = freqs + (uint64_t)i * bsum + foff;
                double n = (i + 1.0) * step_size;
                for (int j = 0; j < nb; ++j) fr[j] = f[j] / n;
                double sum = 0.0;
                uint8_t blank = 0, single = 0;
                if (est == E_MM) {
                    for (int j = 0; j < nb; ++j) if (f[j] == 0) ++blank;
                } else if (est == E_CS) {
                    for (int j = 0; j < nb; ++j) if (f[j] == 1) ++single;
                    for (int j = 0; j < nb; ++j) fr[j] = (1 - (single / n)) * (f[j] / n);
                } else if (est == E_JS) {
                    double tk = 1.0 / (double)nb, s2 = 0, tks = 0;
                    for (int j = 0; j < nb; ++j) {
                        s2 += fr[j] * fr[j];
                        tks += ((tk - fr[j]) * (tk - fr[j]));
                    }
                    double lam = (1 - s2) / ((n - 1) * tks);
                    for (int j = 0; j < nb; ++j) fr[j] = (lam * tk) + ((1 - lam) * fr[j]);
                }
                double jf = 1.0;
                for (int j = 0; j < nb; ++j) {
                    double pd = fr[j] / w;
                    if (pd > 0.0) {
                        double denom = (est == E_CS) ? 1 - pow((1 - fr[j]), n) : 1.0;
                        if (jacobian && hasj) {
                            jf = or_jac(type, mid[j]);
                            pd = pd / jf;
                            sum += (est == E_CS) ? (jf * pd * log(pd) / denom)
                                                 : (jf * pd * log(pd));
                        } else {
                            sum += (est == E_CS) ? (pd * log(pd) / denom)
                                                 : (pd * log(pd));
                        }
                    }
                }
                sum *= -w;
                if (est == E_MM) sum += ((nb - blank) - 1) / (2.0 * n);
                entropy[eoff + entoff + i] = sum;
            }
            foff += nb;
        }
    }
    return 0;
}

/* entropy2D (Estimators.cpp:391-651), including the Chao-Shen second-DOF
 * Jacobian that is evaluated at the row bin (mid2[j], Estimators.cpp:558,
 * 562) and the James-Stein target 1/nbins (Estimators.cpp:510).
 * entropy = nest*nsch*nsteps. */
int or_entropy_2d(int t1, int t2, const int *ests, int nest, uint8_t nsteps,
                  uint64_t step_size, double mn1, double mx1, double mn2,
                  double mx2, int jacobian, int kde, const uint8_t *schemes,
                  int nsch, const uint64_t *freqs, double *entropy) {
    size_t bsq = 0;
    for (int i = 0; i < nsch; ++i) bsq += (size_t)schemes[i] * schemes[i];
    double *fr = (double *)malloc(sizeof(double) * 256 * 256);
    double mid1[256], mid2[256];
    for (int e = 0; e < nest; ++e) {
        int est = ests[e];
        unsigned eoff = (unsigned)e * nsteps * nsch;
        unsigned foff = 0;
        for (int s = 0, entoff = 0; s < nsch; entoff += nsteps, ++s) {
            int nb = schemes[s];
            unsigned nb2 = (unsigned)nb * nb;
            double w1 = (mx1 - mn1) / nb, hw1 = w1 / 2;
            double w2 = (mx2 - mn2) / nb, hw2 = w2 / 2;
            if (kde) { w1 = 1.0; w2 = 1.0; }
            int hj1 = (t1 == T_BOND || t1 == T_ANGLE);
            int hj2 = (t2 == T_BOND || t2 == T_ANGLE);
            for (int j = 0; j < nb; ++j) {
                mid1[j] = hj1 ? mn1 + (j * w1) + hw1 : 0.0;
                mid2[j] = hj2 ? mn2 + (j * w2) + hw2 : 0.0;
            }
            double area = w1 * w2;
            for (int i = 0; i < nsteps; ++i) {
                const uint64_t *f = freqs + (uint64_t)i * bsq + foff;
                double n = (i + 1.0) * step_size;
                for (unsigned j = 0; j < nb2; ++j) fr[j] = f[j] / n;
                double sum = 0.0;
                unsigned blank = 0, single = 0;
                if (est == E_MM) {
                    for (unsigned j = 0; j < nb2; ++j) if (f[j] == 0) ++blank;
                } else if (est == E_CS) {
                    for (unsigned j = 0; j < nb2; ++j) if (f[j] == 1) ++single;
                    for (unsigned j = 0; j < nb2; ++j) fr[j] = (1 - (single / n)) * (f[j] / n);
                } else if (est == E_JS) {
                    double tk = 1.0 / (double)nb, s2 = 0, tks = 0;
                    for (unsigned j = 0; j < nb2; ++j) {
                        s2 += fr[j] * fr[j];
                        tks += ((tk - fr[j]) * (tk - fr[j]));
                    }
                    double lam = (1 - s2) / ((n - 1) * tks);
                    for (unsigned j = 0; j < nb2; ++j) fr[j] = (lam * tk) + ((1 - lam) * fr[j]);
                }
                double jf1 = 1.0, jf2 = 1.0;
                for (int j = 0; j < nb; ++j) {
                    if (hj1) jf1 = or_jac(t1, mid1[j]);
                    double row = 0.0;
                    for (int k = 0; k < nb; ++k) {
                        double q = fr[j * nb + k];
                        double pd = q / area;
                        if (pd > 0.0) {
                            double denom = 1.0;
                            if (est == E_CS) {
                                denom = 1 - pow((1 - q), n);
                                if (hj2) jf2 = or_jac(t2, mid2[j]);
                            } else if (hj2) {
                                jf2 = or_jac(t2, mid2[k]);
                            }
                            if (jacobian) {
                                double jf = jf1 * jf2;
                                pd = pd / jf;
                                row += (est == E_CS) ? (jf * pd * log(pd) / denom)
                                                     : (jf * pd * log(pd));
                            } else {
                                row += (est == E_CS) ? (pd * log(pd) / denom)
                                                     : (pd * log(pd));
                            }
                        }
                    }
                    row *= -area;
                    sum += row;
                }
                if (est == E_MM) sum += ((nb2 - blank) - 1) / (2.0 * n);
                entropy[eoff + entoff + i] = sum;
            }
            foff += nb2;
        }
    }
    free(fr);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Whole-stage drivers.  codes: [D][nsch][F] uint8 (row stride nsch*F),
 * types[D] (1/2/3), ext[D][4].  These are what EntropyCalculator::run1d /
 * run2d_xx / run2d_xy do per DOF / per pair (EntropyCalculator.cpp:156,
 * 530-564): binData + entropy, OpenMP dynamic over items. */
int or_dof_entropies(const uint8_t *codes, uint32_t D, uint64_t F,
                     const uint8_t *types, const double *ext, uint8_t nsteps,
                     const uint8_t *schemes, int nsch, const int *ests,
                     int nest, int jacobian, int kde, double *s1,
                     int nthreads) {
    unsigned bsum = 0;
    for (int i = 0; i < nsch; ++i) bsum += schemes[i];
    uint64_t ss = F / nsteps;
    size_t per = (size_t)nest * nsch * nsteps;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
    for (int64_t d = 0; d < (int64_t)D; ++d) {
        uint64_t *f = (uint64_t *)calloc((size_t)nsteps * bsum, sizeof(uint64_t));
        double *mids = (double *)malloc(sizeof(double) * bsum);
        const double *e = ext + 4 * d;
        or_bin_data_1d(nsteps, schemes, nsch, F, e[0], e[1],
                       codes + (uint64_t)d * nsch * F, f, mids);
        or_entropy_1d(types[d], ests, nest, nsteps, ss, e[0], e[1], jacobian,
                      kde, schemes, nsch, f, s1 + (size_t)d * per);
        free(f);
        free(mids);
    }
    return 0;
}

int or_pair_entropies(const uint8_t *codes, uint32_t D, uint64_t F,
                      const uint8_t *types, const double *ext,
                      const uint32_t *pairs, uint64_t npair, uint8_t nsteps,
                      const uint8_t *schemes, int nsch, const int *ests,
                      int nest, int jacobian, int kde, double *s2,
                      int nthreads) {
    size_t bsq = 0;
    for (int i = 0; i < nsch; ++i) bsq += (size_t)schemes[i] * schemes[i];
    unsigned bsum = 0;
    for (int i = 0; i < nsch; ++i) bsum += schemes[i];
    uint64_t ss = F / nsteps;
    size_t per = (size_t)nest * nsch * nsteps;
    for (uint64_t p = 0; p < npair; ++p)
        if (pairs[2 * p] >= D || pairs[2 * p + 1] >= D) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        uint64_t *f = (uint64_t *)malloc(sizeof(uint64_t) * nsteps * bsq);
        double *mids = (double *)malloc(sizeof(double) * 2 * bsum);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t p = 0; p < (int64_t)npair; ++p) {
            uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
            const double *e1 = ext + 4 * i, *e2 = ext + 4 * j;
            memset(f, 0, sizeof(uint64_t) * nsteps * bsq);
            or_bin_data_2d(nsteps, schemes, nsch, F, e1[0], e1[1], e2[0], e2[1],
                           codes + (uint64_t)i * nsch * F,
                           codes + (uint64_t)j * nsch * F, f, mids);
            or_entropy_2d(types[i], types[j], ests, nest, nsteps, ss, e1[0],
                          e1[1], e2[0], e2[1], jacobian, kde, schemes, nsch, f,
                          s2 + (size_t)p * per);
        }
        free(f);
        free(mids);
    }
    return 0;
}

/* Joint counts of every listed pair at every step, as the prefix sums that
 * binData2D leaves (one scheme, nb bins): out = npair * nsteps * nb*nb. */
int or_pair_counts(const uint8_t *codes, uint32_t D, uint64_t F, int nb,
                   const uint32_t *pairs, uint64_t npair, uint8_t nsteps,
                   uint64_t *out) {
    uint64_t ss = F / nsteps;
    size_t nb2 = (size_t)nb * nb;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t p = 0; p < (int64_t)npair; ++p) {
        const uint8_t *a = codes + (uint64_t)pairs[2 * p] * F;
        const uint8_t *b = codes + (uint64_t)pairs[2 * p + 1] * F;
        uint64_t *o = out + (size_t)p * nsteps * nb2;
        memset(o, 0, sizeof(uint64_t) * nsteps * nb2);
        for (int s = 0; s < nsteps; ++s) {
            uint64_t *os = o + (size_t)s * nb2;
            if (s) memcpy(os, os - nb2, sizeof(uint64_t) * nb2);
            for (uint64_t fidx = (uint64_t)s * ss; fidx < (uint64_t)(s + 1) * ss; ++fidx)
                os[(size_t)a[fidx] * nb + b[fidx]] += 1;
        }
    }
    (void)D;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* boruvkaMST (Centre.cpp:5-115) with findRoot / unionComponents
 * (Centre.cpp:117-143): maximum spanning forest; per component the heaviest
 * cross edge, ties to the lowest edge index (strict <).  tree_idx receives
 * edge indices in insertion order; returns the count via *ntree. */
static int or_find(uint32_t *par, int i) {
    while (par[i] != (uint32_t)i) {
        par[i] = par[par[i]];
        i = (int)par[i];
    }
    return i;
}

int or_boruvka(uint32_t V, const uint32_t *src, const uint32_t *dst,
               const double *w, uint64_t E, uint32_t *tree_idx,
               uint32_t *ntree) {
    uint32_t *par = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
    uint32_t *rank = (uint32_t *)calloc(V ? V : 1, sizeof(uint32_t));
    int64_t *cheap = (int64_t *)malloc(sizeof(int64_t) * (V ? V : 1));
    for (uint32_t v = 0; v < V; ++v) { par[v] = v; cheap[v] = -1; }
    uint32_t ntrees = V, nt = 0;
    for (uint64_t e = 0; e < E; ++e)
        if (src[e] >= V || dst[e] >= V) { free(par); free(rank); free(cheap); return -1; }
    while (ntrees > 1) {
        int found = 0;
        for (uint64_t i = 0; i < E; ++i) {
            int s1 = or_find(par, src[i]), s2 = or_find(par, dst[i]);
            if (s1 == s2) continue;
            if (cheap[s1] == -1 || w[cheap[s1]] < w[i]) { cheap[s1] = (int64_t)i; found++; }
            if (cheap[s2] == -1 || w[cheap[s2]] < w[i]) { cheap[s2] = (int64_t)i; found++; }
        }
        if (!found) break;
        for (uint32_t v = 0; v < V; ++v) {
            if (cheap[v] == -1) continue;
            int s1 = or_find(par, src[cheap[v]]), s2 = or_find(par, dst[cheap[v]]);
            if (s1 == s2) continue;
            tree_idx[nt++] = (uint32_t)cheap[v];
            if (rank[s1] < rank[s2]) par[s1] = s2;
            else if (rank[s1] > rank[s2]) par[s2] = s1;
            else { par[s2] = s1; rank[s1]++; }
            ntrees--;
        }
        for (uint32_t v = 0; v < V; ++v) cheap[v] = -1;
    }
    *ntree = nt;
    free(par);
    free(rank);
    free(cheap);
    return 0;
}
