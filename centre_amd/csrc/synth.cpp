// Seeded synthetic BAT trajectories for tests and benchmarks.
//
// No reference counterpart (the reference reads real trajectories through
// netcdf-c).  Values are a pure function of (seed, frame, DOF): a
// counter-based Philox4x32-10 generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC 2011) keyed by the seed, so any
// thread count gives the same bytes.
//
// Model (columns: n_bond bonds, then n_angle angles, then n_dih torsions):
//   latent   one shared 4-state Markov chain z_f, z_0 = 0; each frame keeps
//            the state with probability 0.98, else draws a state uniformly.
//   bonds    N(1.5 + s(z) * 0.03, 0.03),  angles N(1.91 + s(z) * 0.05, 0.05),
//            with s(z) = 0.5 * (z - 1.5) / 1.5.
//   torsions 3-component von Mises mixture at 300, 60 and 180 degrees,
//            kappa = 8, per-torsion Dirichlet(1,1,1) weights; in state z the
//            weight of component (t + z) mod 3 is multiplied by e and the
//            weights renormalised; angles wrapped to [0, 2 pi).
// Normals use Box-Muller; von Mises draws invert a 65537-point tabulated CDF
// (trapezoid rule on exp(kappa cos x)) with linear interpolation.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/centre_hip.h"

namespace {

struct Philox {
    uint32_t k0, k1;
    static inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
        const uint64_t p = (uint64_t)a * b;
        hi = (uint32_t)(p >> 32);
        lo = (uint32_t)p;
    }
    inline void operator()(uint32_t c[4]) const {
        uint32_t key0 = k0, key1 = k1;
        for (int r = 0; r < 10; ++r) {
            uint32_t hi0, lo0, hi1, lo1;
            mulhilo(0xD2511F53u, c[0], hi0, lo0);
            mulhilo(0xCD9E8D57u, c[2], hi1, lo1);
            const uint32_t n0 = hi1 ^ c[1] ^ key0, n1 = lo1, n2 = hi0 ^ c[3] ^ key1, n3 = lo0;
            c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
            key0 += 0x9E3779B9u;
            key1 += 0xBB67AE85u;
        }
    }
};

inline double u01(uint32_t x) { return (x + 0.5) * (1.0 / 4294967296.0); }

constexpr double kTwoPi = 6.283185307179586;
constexpr int kTab = 65536;

struct VonMisesTable {
    std::vector<double> cdf, x;
    explicit VonMisesTable(double kappa) : cdf(kTab + 1), x(kTab + 1) {
        const double pi = 3.141592653589793;
        double acc = 0.0;
        double prev = std::exp(kappa * std::cos(-pi));
        for (int i = 0; i <= kTab; ++i) {
            x[i] = -pi + kTwoPi * i / kTab;
            const double v = std::exp(kappa * std::cos(x[i]));
            if (i) acc += 0.5 * (v + prev) * (kTwoPi / kTab);
            cdf[i] = acc;
            prev = v;
        }
        for (auto& c : cdf) c /= acc;
    }
    double inv(double u) const {
        int lo = 0, hi = kTab;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cdf[mid] <= u) lo = mid; else hi = mid;
        }
        const double span = cdf[hi] - cdf[lo];
        const double t = span > 0 ? (u - cdf[lo]) / span : 0.0;
        return x[lo] + t * (x[hi] - x[lo]);
    }
};

}  // namespace

extern "C" int centre_synth_bat(uint64_t seed, uint32_t n_bond, uint32_t n_angle,
                                uint32_t n_dih, uint64_t n_frame, double* out,
                                int n_threads) {
    if (!out && n_frame) return -3;
    const uint32_t D = n_bond + n_angle + n_dih;
    const Philox ph{(uint32_t)seed, (uint32_t)(seed >> 32)};
    // latent chain (sequential, cheap)
    std::vector<uint8_t> z(n_frame);
    uint8_t state = 0;
    for (uint64_t f = 0; f < n_frame; ++f) {
        uint32_t c[4] = {(uint32_t)f, (uint32_t)(f >> 32), 0xFFFFFFFFu, 1u};
        ph(c);
        if (f > 0 && u01(c[0]) >= 0.98) state = (uint8_t)(c[1] >> 30);
        z[f] = state;
    }
    // per-torsion mixture weights in each latent state
    std::vector<double> wts((size_t)n_dih * 4 * 3);
    for (uint32_t t = 0; t < n_dih; ++t) {
        uint32_t c[4] = {t, 0u, 0xFFFFFFFEu, 2u};
        ph(c);
        double g[3], sg = 0;
        for (int k = 0; k < 3; ++k) {
            g[k] = -std::log(u01(c[k]));
            sg += g[k];
        }
        for (int s = 0; s < 4; ++s) {
            double w[3], sw = 0;
            for (int k = 0; k < 3; ++k) {
                w[k] = g[k] / sg;
                if (k == (int)((t + s) % 3)) w[k] *= 2.718281828459045;
                sw += w[k];
            }
            for (int k = 0; k < 3; ++k) wts[((size_t)t * 4 + s) * 3 + k] = w[k] / sw;
        }
    }
    const VonMisesTable vm(8.0);
    const double mu[3] = {300.0 * kTwoPi / 360.0, 60.0 * kTwoPi / 360.0, 180.0 * kTwoPi / 360.0};
    auto work = [&](uint64_t f0, uint64_t f1) {
        for (uint64_t f = f0; f < f1; ++f) {
            const int s = z[f];
            const double shift = 0.5 * (s - 1.5) / 1.5;
            double* row = out + f * D;
            for (uint32_t d = 0; d < D; ++d) {
                uint32_t c[4] = {(uint32_t)f, (uint32_t)(f >> 32), d, 0u};
                ph(c);
                if (d < n_bond + n_angle) {
                    const double r = std::sqrt(-2.0 * std::log(u01(c[0])));
                    const double nrm = r * std::cos(kTwoPi * u01(c[1]));
                    const bool bond = d < n_bond;
                    const double m = bond ? 1.5 : 1.91, sd = bond ? 0.03 : 0.05;
                    row[d] = m + shift * sd + sd * nrm;
                } else {
                    const uint32_t t = d - n_bond - n_angle;
                    const double* w = &wts[((size_t)t * 4 + s) * 3];
                    const double u = u01(c[0]);
                    const int k = u < w[0] ? 0 : (u < w[0] + w[1] ? 1 : 2);
                    double a = mu[k] + vm.inv(u01(c[1]));
                    a = std::fmod(a, kTwoPi);
                    if (a < 0) a += kTwoPi;
                    // keep the float32 image below 2 pi (the reader rounds)
                    if (a >= kTwoPi || (double)(float)a >= kTwoPi) a = 0.0;
                    row[d] = a;
                }
            }
        }
    };
    int nt = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if ((uint64_t)nt > n_frame) nt = (int)(n_frame ? n_frame : 1);
    std::vector<std::thread> th;
    const uint64_t per = (n_frame + nt - 1) / nt;
    for (int i = 0; i < nt; ++i) {
        const uint64_t a = i * per, b = std::min<uint64_t>(n_frame, a + per);
        if (a < b) th.emplace_back(work, a, b);
    }
    for (auto& t : th) t.join();
    return 0;
}
