// Marginal stage and the estimator epilogue on materialised counts.
//
// dof1d_kernel replaces binData1D + entropy1D (Estimators.cpp:56-118,
// 217-389) for all DOFs at once: one workgroup per DOF histograms each step
// chunk in LDS, keeps the prefix counts, and evaluates every requested
// estimator with bin b on thread b: the reference's per-bin terms, summed in
// a fixed tree order (so S1 is the same from run to run and within a few ulp
// of the sequential sum).  It also leaves, per DOF and step, the marginal
// Jacobian sum L = sum_b m_b log J_b and the bin width that the pair
// epilogue needs.
//
// count_entropy_kernel evaluates entropy2D (Estimators.cpp:391-651) for
// every estimator, and MI, from prefix counts written by the pair kernel's
// dump mode: one wave per (record, step), in the record's own row/column
// order, including the Chao-Shen second Jacobian evaluated at mid2[j]
// (Estimators.cpp:558,562) and the James-Stein target 1/nbins
// (Estimators.cpp:510).
#include "centre_internal.h"

namespace {

constexpr int kML = 0, kMM = 1, kCS = 2, kJS = 3;

__device__ inline double jac(int type, double mid) {
    if (type == 1) return mid * mid;        // pow(mid, 2)
    if (type == 2) return sin(mid);
    return 1.0;
}

// Sum over the 256 threads of a workgroup in a fixed order (wave shuffles,
// then the four wave totals in order), so the result does not vary from run
// to run; every thread gets the total.  red: 4 doubles of LDS.
__device__ inline double block_sum256(double v, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();   // red is free (previous use finished)
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// entropy1D for one step and one estimator with bin b on thread b (threads
// b >= nb contribute nothing).  The bin statistics the estimator needs
// (empty and singleton bins, the James-Stein sums) are block sums first;
// then the per-bin terms and their block sum.  Same terms as the
// reference's loops (Estimators.cpp:217-389), summed in tree order.
__device__ double ent1d_par(int type, int est, int nb, uint64_t fb, bool live, double n,
                            double w, double midb, int jacobian, double* red) {
    double frb = live ? fb / n : 0.0;
    int blank = 0;
    if (est == kMM) {
        blank = (int)block_sum256((live && fb == 0) ? 1.0 : 0.0, red);
    } else if (est == kCS) {
        const double single = block_sum256((live && fb == 1) ? 1.0 : 0.0, red);
        frb = live ? (1 - (single / n)) * (fb / n) : 0.0;
    } else if (est == kJS) {
        const double tk = 1.0 / (double)nb;
        const double s2 = block_sum256(frb * frb, red);
        const double tks = block_sum256(live ? (tk - frb) * (tk - frb) : 0.0, red);
        const double lam = (1 - s2) / ((n - 1) * tks);
        frb = live ? (lam * tk) + ((1 - lam) * frb) : 0.0;
    }
    const bool hasj = (type == 1 || type == 2);
    double term = 0.0;
    double pd = frb / w;
    if (live && pd > 0.0) {
        const double denom = (est == kCS) ? 1 - pow((1 - frb), n) : 1.0;
        if (jacobian && hasj) {
            const double jf = jac(type, midb);
            pd = pd / jf;
            term = (est == kCS) ? (jf * pd * log(pd) / denom) : (jf * pd * log(pd));
        } else {
            term = (est == kCS) ? (pd * log(pd) / denom) : (pd * log(pd));
        }
    }
    double sum = block_sum256(term, red);
    sum *= -w;
    if (est == kMM) sum += ((nb - blank) - 1) / (2.0 * n);
    return sum;
}

// Up to kPrivBins bins each thread counts into its own LDS row (an odd row
// length keeps the 64 lanes of a wave in different banks); more bins share
// one histogram through LDS atomics, which collide less the more bins there
// are.
constexpr int kPrivBins = 32;

__global__ __launch_bounds__(256) void dof1d_kernel(Dof1dArgs a) {
    __shared__ uint32_t h[256];
    __shared__ uint32_t hp[256 * (kPrivBins + 1)];
    __shared__ uint64_t cum[256];
    __shared__ double mid[256];
    __shared__ double red[4];
    const int d = blockIdx.x, tid = threadIdx.x;
    const int nb = a.nb, S = a.nsteps;
    const uint8_t* row = a.codes + (uint64_t)d * a.stride;
    const int type = a.types[d];
    const double mn = a.ext[4 * d + 0], mx = a.ext[4 * d + 1];
    const bool hasj = (type == 1 || type == 2);
    double w = (mx - mn) / nb;
    const double hw = w / 2;
    if (a.kde) w = 1.0;
    for (int b = tid; b < 256; b += 256) {
        cum[b] = 0;
        mid[b] = (hasj && b < nb) ? mn + b * w + hw : 0.0;
    }
    const bool priv = nb <= kPrivBins;
    const int nbp = nb | 1;                        // private row length, odd
    uint32_t* const mine = hp + tid * nbp;
    auto add = [&](uint32_t c) __attribute__((always_inline)) {
        if (c < (uint32_t)nb) {
            if (priv) atomicAdd(&mine[c], 1u);     // private: never contended
            else atomicAdd(&h[c], 1u);
        } else {
            atomicOr(a.flag, 1);
        }
    };
    for (int s = 0; s < S; ++s) {
        if (priv) {
            for (int b = 0; b < nbp; ++b) mine[b] = 0;
        } else {
            h[tid] = 0;
        }
        __syncthreads();
        // the step's codes: bytes up to a 4-byte boundary, dwords, then the
        // remaining bytes
        const uint8_t* const p = row + (uint64_t)s * a.ss;
        const uint64_t len = a.ss;
        const uint64_t head = min((uint64_t)((4 - (reinterpret_cast<uintptr_t>(p) & 3)) & 3), len);
        if ((uint64_t)tid < head) add(p[tid]);
        const uint32_t* const p4 = reinterpret_cast<const uint32_t*>(p + head);
        const uint64_t n4 = (len - head) >> 2;
        for (uint64_t i = tid; i < n4; i += 256) {
            const uint32_t v = p4[i];
            add(v & 0xFFu);
            add((v >> 8) & 0xFFu);
            add((v >> 16) & 0xFFu);
            add(v >> 24);
        }
        for (uint64_t f = head + 4 * n4 + tid; f < len; f += 256) add(p[f]);
        __syncthreads();
        if (tid < nb) {
            uint32_t tot = 0;
            if (priv) {
                for (int t = 0; t < 256; ++t) tot += hp[t * nbp + tid];
            } else {
                tot = h[tid];
            }
            cum[tid] += tot;
            if (a.counts) a.counts[((uint64_t)d * S + s) * nb + tid] = cum[tid];
        }
        __syncthreads();
        {   // every estimator and the Jacobian sum, bin b on thread b
            const double n = (s + 1.0) * a.ss;
            const bool live = tid < nb;
            const uint64_t fb = live ? cum[tid] : 0;
            const double midb = live ? mid[tid] : 0.0;
            for (int e = 0; e < a.nest; ++e) {
                const double v = ent1d_par(type, a.ests[e], nb, fb, live, n, w, midb, a.jacobian, red);
                if (tid == 0) a.s1[((uint64_t)d * a.nest + e) * S + s] = v;
            }
            double L = 0.0;
            if (a.jacobian && hasj)
                L = block_sum256((live && fb) ? (double)fb * log(jac(type, midb)) : 0.0, red);
            if (tid == 0) a.L[(uint64_t)d * S + s] = L;
        }
        __syncthreads();
    }
    if (tid == 0) a.w[d] = w;
}

// Sum over the 64 lanes of a wave in a fixed order; every lane returns lane
// 0's total, so the value does not depend on the lane or the run.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return __shfl(v, 0);
}

__device__ inline unsigned wave_sum_u(unsigned v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// (i, j), i < j, of canonical index p in the row-major upper triangle of D
// DOFs (row i starts at i (2D - i - 1) / 2)
__device__ inline void tri_pair(uint64_t p, uint32_t D, uint32_t* pi, uint32_t* pj) {
    auto off = [D](uint64_t x) { return x * (2ull * D - x - 1) / 2; };
    const double b = 2.0 * D - 1.0;
    long long i = (long long)((b - sqrt(b * b - 8.0 * (double)p)) / 2.0);
    if (i < 0) i = 0;
    if (i > (long long)D - 2) i = (long long)D - 2;
    while (i > 0 && off((uint64_t)i) > p) --i;
    while (i + 1 <= (long long)D - 2 && off((uint64_t)i + 1) <= p) ++i;
    *pi = (uint32_t)i;
    *pj = (uint32_t)(p - off((uint64_t)i) + (uint64_t)i + 1);
}

// One wave per (record, step): the lanes take the record's nb*nb cells in
// turn.  The statistics the estimators need (empty and singleton cells, the
// James-Stein sums) come first, then each estimator's terms, as
// entropy2D writes them (Estimators.cpp:455-648): the Chao-Shen second
// Jacobian at the row bin (mid2[j], :558, :562), the James-Stein target
// 1/nbins (:510).  (1 - q)^n is below 2^-54 once n q > 40, where
// 1 - (1 - q)^n rounds to 1 exactly, so the Horvitz-Thompson divisor skips
// pow there.
__global__ __launch_bounds__(64) void count_entropy_kernel(CountEntropyArgs a) {
    __shared__ double j1[256], j2[256];
    const int S = a.nsteps, nb = a.nb, nest = a.nest;
    const uint64_t item = blockIdx.x;
    const int s = (int)blockIdx.y;
    const int lane = (int)threadIdx.x;
    uint64_t rec, canon;
    uint32_t i, j;
    if (a.mode == 2) {
        rec = a.items[item];
        canon = a.rec2canon[rec];
        i = a.rec_ij[2 * rec];
        j = a.rec_ij[2 * rec + 1];
    } else {
        canon = a.mode == 1 ? a.items[item] : a.item_first + item;
        rec = canon;
        tri_pair(canon, a.n_dof, &i, &j);
    }
    // counts are stored with the lower DOF index as rows; the reference's
    // row is the record's first DOF
    const bool flip = i > j;
    const unsigned rs = flip ? 1u : (unsigned)nb, cs = flip ? (unsigned)nb : 1u;
    const int t1 = a.types[i], t2 = a.types[j];
    const double mn1 = a.ext[4 * i], mx1 = a.ext[4 * i + 1];
    const double mn2 = a.ext[4 * j], mx2 = a.ext[4 * j + 1];
    double w1 = (mx1 - mn1) / nb, w2 = (mx2 - mn2) / nb;
    const double hw1 = w1 / 2, hw2 = w2 / 2;
    if (a.kde) {
        w1 = 1.0;
        w2 = 1.0;
    }
    const bool hj1 = (t1 == 1 || t1 == 2), hj2 = (t2 == 1 || t2 == 2);
    for (int b = lane; b < nb; b += 64) {
        j1[b] = hj1 ? jac(t1, mn1 + (b * w1) + hw1) : 1.0;
        j2[b] = hj2 ? jac(t2, mn2 + (b * w2) + hw2) : 1.0;
    }
    __syncthreads();
    const double area = w1 * w2;
    const double n = (s + 1.0) * a.ss;
    const unsigned nb2 = (unsigned)nb * nb;
    const uint32_t* f = a.counts + ((canon - a.count_base) * S + s) * nb2;
    unsigned blank = 0, single = 0;
    double sq = 0.0, tks = 0.0;
    const double tk = 1.0 / (double)nb;
    for (unsigned q = lane; q < nb2; q += 64) {
        const uint32_t c = f[q];
        blank += c == 0;
        single += c == 1;
        const double fq = c / n;
        sq += fq * fq;
        tks += (tk - fq) * (tk - fq);
    }
    blank = wave_sum_u(blank);
    single = wave_sum_u(single);
    sq = wave_sum(sq);
    tks = wave_sum(tks);
    const double lam = (1 - sq) / ((n - 1) * tks);
    const double cov = 1 - (single / n);
    for (int e = 0; e < nest; ++e) {
        const int est = a.ests[e];
        double part = 0.0;
        for (unsigned q = lane; q < nb2; q += 64) {
            const unsigned r = q / nb, c = q - r * nb;
            const uint32_t fc = f[r * rs + c * cs];
            double p = fc / n;
            if (est == kCS) p = cov * (fc / n);
            else if (est == kJS) p = (lam * tk) + ((1 - lam) * p);
            double pd = p / area;
            if (pd > 0.0) {
                double denom = 1.0, jf2;
                if (est == kCS) {
                    if (n * p <= 40.0) denom = 1 - pow((1 - p), n);
                    jf2 = j2[r];
                } else {
                    jf2 = j2[c];
                }
                if (a.jacobian) {
                    const double jf = j1[r] * jf2;
                    pd = pd / jf;
                    part += (est == kCS) ? (jf * pd * log(pd) / denom) : (jf * pd * log(pd));
                } else {
                    part += (est == kCS) ? (pd * log(pd) / denom) : (pd * log(pd));
                }
            }
        }
        double sum = wave_sum(part) * -area;
        if (est == kMM) sum += ((nb2 - blank) - 1) / (2.0 * n);
        if (lane == 0) {
            a.s2[(rec * nest + e) * S + s] = sum;
            if (e == 0 && a.mi)
                a.mi[rec * S + s] = (a.s1[((uint64_t)i * nest) * S + s] +
                                     a.s1[((uint64_t)j * nest) * S + s]) - sum;
        }
    }
    if (a.counts_out) {
        uint32_t* o = a.counts_out + (rec * S + s) * nb2;
        for (unsigned q = lane; q < nb2; q += 64) {
            const unsigned r = q / nb, c = q - r * nb;
            o[q] = f[r * rs + c * cs];
        }
    }
}

__global__ void gather_rec_kernel(const double* src, const uint64_t* rec2canon,
                                  uint64_t n_rec, int width, double* dst) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_rec * width) return;
    const uint64_t r = t / width;
    const int k = (int)(t - r * width);
    dst[t] = src[rec2canon[r] * width + k];
}

// Pitched copy of height rows of width bytes, in 8-byte words when every
// row start is 8-byte aligned (narrow rows: the runtime's 2-D copy issues
// one copy per row).
__global__ void copy_2d_kernel(uint8_t* dst, size_t dpitch, const uint8_t* src, size_t spitch,
                               size_t width, size_t height, int words) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = words ? width / 8 : width;
    if (t >= per * height) return;
    const uint64_t r = t / per, c = t - r * per;
    if (words)
        reinterpret_cast<uint64_t*>(dst + r * dpitch)[c] =
            reinterpret_cast<const uint64_t*>(src + r * spitch)[c];
    else
        dst[r * dpitch + c] = src[r * spitch + c];
}

}  // namespace

hipError_t launch_copy_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                          size_t height, hipStream_t st) {
    const int words = ((width | dpitch | spitch | reinterpret_cast<uintptr_t>(dst) |
                        reinterpret_cast<uintptr_t>(src)) & 7) == 0;
    const uint64_t n = (words ? width / 8 : width) * (uint64_t)height;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(copy_2d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       static_cast<uint8_t*>(dst), dpitch, static_cast<const uint8_t*>(src), spitch,
                       width, height, words);
    return hipGetLastError();
}

hipError_t launch_dof1d(const Dof1dArgs& a, hipStream_t st) {
    if (a.n_dof == 0) return hipSuccess;
    hipLaunchKernelGGL(dof1d_kernel, dim3(a.n_dof), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_count_entropy(const CountEntropyArgs& a, hipStream_t st) {
    if (!a.n_item) return hipSuccess;
    if (a.n_item > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(count_entropy_kernel, dim3((unsigned)a.n_item, (unsigned)a.nsteps),
                       dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_gather_rec(const double* src, const uint64_t* rec2canon,
                             uint64_t n_rec, int width, double* dst,
                             hipStream_t st) {
    const uint64_t n = n_rec * width;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(gather_rec_kernel, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, st, src, rec2canon, n_rec, width, dst);
    return hipGetLastError();
}
