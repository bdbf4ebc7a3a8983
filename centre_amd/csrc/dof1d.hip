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
// count_entropy_kernel evaluates entropy2D (Estimators.cpp:391-651) for any
// estimator from prefix counts written by the pair kernel's dump mode: one
// thread per (pair, step, estimator), in the reference's row/column order,
// including the Chao-Shen second Jacobian evaluated at mid2[j]
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

__global__ __launch_bounds__(64) void count_entropy_kernel(CountEntropyArgs a) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int S = a.nsteps, nest = a.nest, nb = a.nb;
    if (t >= a.n_pair * S * nest) return;
    const int e = (int)(t % nest);
    const int s = (int)((t / nest) % S);
    const uint64_t p = t / ((uint64_t)nest * S);
    const int est = a.ests[e];
    const uint32_t i = a.canon_ij[2 * p], j = a.canon_ij[2 * p + 1];
    const uint64_t pc = a.rec2canon ? a.rec2canon[p] : p;
    // counts are stored with the lower DOF index as rows; the reference's
    // row is the record's first DOF, which matters to the Chao-Shen Jacobian
    const bool flip = i > j;
    const unsigned rs = flip ? 1u : (unsigned)nb, cs = flip ? (unsigned)nb : 1u;
    const int t1 = a.types[i], t2 = a.types[j];
    const double mn1 = a.ext[4 * i], mx1 = a.ext[4 * i + 1];
    const double mn2 = a.ext[4 * j], mx2 = a.ext[4 * j + 1];
    double w1 = (mx1 - mn1) / nb, w2 = (mx2 - mn2) / nb;
    const double hw1 = w1 / 2, hw2 = w2 / 2;
    if (a.kde) { w1 = 1.0; w2 = 1.0; }
    const bool hj1 = (t1 == 1 || t1 == 2), hj2 = (t2 == 1 || t2 == 2);
    const double area = w1 * w2;
    const double n = (s + 1.0) * a.ss;
    const uint32_t* f = a.counts + ((size_t)pc * S + s) * nb * nb;
    const unsigned nb2 = (unsigned)nb * nb;
    unsigned blank = 0, single = 0;
    double lam = 0.0, tk = 1.0 / (double)nb;
    if (est == kMM) {
        for (unsigned q = 0; q < nb2; ++q) if (f[q] == 0) ++blank;
    } else if (est == kCS) {
        for (unsigned q = 0; q < nb2; ++q) if (f[q] == 1) ++single;
    } else if (est == kJS) {
        double s2 = 0, tks = 0;
        for (unsigned q = 0; q < nb2; ++q) {
            const double fq = f[q] / n;
            s2 += fq * fq;
            tks += ((tk - fq) * (tk - fq));
        }
        lam = (1 - s2) / ((n - 1) * tks);
    }
    double sum = 0.0, jf1 = 1.0, jf2 = 1.0;
    for (int r = 0; r < nb; ++r) {
        if (hj1) jf1 = jac(t1, mn1 + (r * w1) + hw1);
        double row = 0.0;
        for (int c = 0; c < nb; ++c) {
            const uint32_t fc = f[r * rs + c * cs];
            double q = fc / n;
            if (est == kCS) q = (1 - (single / n)) * (fc / n);
            else if (est == kJS) q = (lam * tk) + ((1 - lam) * q);
            double pd = q / area;
            if (pd > 0.0) {
                double denom = 1.0;
                if (est == kCS) {
                    denom = 1 - pow((1 - q), n);
                    if (hj2) jf2 = jac(t2, mn2 + (r * w2) + hw2);
                } else if (hj2) {
                    jf2 = jac(t2, mn2 + (c * w2) + hw2);
                }
                if (a.jacobian) {
                    const double jf = jf1 * jf2;
                    pd = pd / jf;
                    row += (est == kCS) ? (jf * pd * log(pd) / denom) : (jf * pd * log(pd));
                } else {
                    row += (est == kCS) ? (pd * log(pd) / denom) : (pd * log(pd));
                }
            }
        }
        row *= -area;
        sum += row;
    }
    if (est == kMM) sum += ((nb2 - blank) - 1) / (2.0 * n);
    a.s2[((size_t)p * nest + e) * S + s] = sum;
}

__global__ void gather_rec_kernel(const double* src, const uint64_t* rec2canon,
                                  uint64_t n_rec, int width, double* dst) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_rec * width) return;
    const uint64_t r = t / width;
    const int k = (int)(t - r * width);
    dst[t] = src[rec2canon[r] * width + k];
}

__global__ void gather_counts_kernel(const uint32_t* src, const uint64_t* rec2canon,
                                     const uint32_t* rec_ij, uint64_t n_rec, int S, int nb,
                                     uint32_t* dst) {
    const uint64_t nb2 = (uint64_t)nb * nb;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_rec * S * nb2) return;
    const uint64_t q = t % nb2;
    const uint64_t rs = t / nb2;           // r * S + s
    const uint64_t r = rs / S;
    const uint64_t s = rs - r * S;
    const bool flip = rec_ij[2 * r] > rec_ij[2 * r + 1];
    const uint64_t x = q / nb, y = q - x * nb;
    const uint64_t qs = flip ? y * nb + x : q;
    dst[t] = src[(rec2canon[r] * S + s) * nb2 + qs];
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

hipError_t launch_gather_counts(const uint32_t* src, const uint64_t* rec2canon,
                                const uint32_t* rec_ij, uint64_t n_rec, int nsteps, int nb,
                                uint32_t* dst, hipStream_t st) {
    const uint64_t n = n_rec * nsteps * (uint64_t)nb * nb;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(gather_counts_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       src, rec2canon, rec_ij, n_rec, nsteps, nb, dst);
    return hipGetLastError();
}

hipError_t launch_dof1d(const Dof1dArgs& a, hipStream_t st) {
    if (a.n_dof == 0) return hipSuccess;
    hipLaunchKernelGGL(dof1d_kernel, dim3(a.n_dof), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_count_entropy(const CountEntropyArgs& a, hipStream_t st) {
    const uint64_t n = a.n_pair * a.nsteps * a.nest;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(count_entropy_kernel, dim3((unsigned)((n + 63) / 64)),
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
