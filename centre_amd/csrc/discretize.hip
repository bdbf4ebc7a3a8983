// Histogram discretizer: float64 BAT columns -> uint8 bin codes + extrema.
//
// Replaces Real2Int::bndReal2Int / angReal2Int / dihReal2Int
// (Real2Int.cpp:27-73, 131-184, 244-479).  Input is the [frame][dof]
// float64 hyperslab as stored in the BAT file; each value is first rounded
// to float32, as nc_get_vara_float does for the reference
// (Netcdf_BAT.cpp:247,279,311), and all range arithmetic is then FP64 with
// IEEE division, floor and fmod, in the order the source writes it.
//
// Passes (all HBM-bound; reads are coalesced across the DOFs of a frame row):
//   stats    per (64-DOF group, frame chunk): min, max, sum of the rounded
//            values; torsions with optimisation also fill the m-bin gap
//            histogram, m = min(3*nbins_ref, 180) over [0, 2pi + 5e-9).
//   finalize per DOF: ranges of bonds/angles; for torsions the largest empty
//            circular run of the gap histogram -> phase shift.
//   shifted  (torsions) min/max of fmod(x + pi, 2pi) and fmod(x + shift, 2pi).
//   finalize2 (torsions) the reference's range choice (Real2Int.cpp:376-410).
//   bin      codes = floor((x' - min)/w), top bin folded, written transposed
//            to [dof][scheme][frame] through an LDS tile.
// min and max are order-independent and bit-exact; the average is a sum
// over chunks (the reference's own build reassociates it under -Ofast).
#include "centre_internal.h"

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kTwoPi = 6.283185307179586;
constexpr double kDx = 5.0e-9;
constexpr int kMaxGap = 180;

// partial record per (chunk, dof): mn, mx, sum, mnPi, mxPi, mnSh, mxSh, -
__global__ __launch_bounds__(256) void stats_kernel(DiscArgs a, int pass,
                                                    uint64_t fchunk) {
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const uint32_t d = blockIdx.x * 64 + tx;
    const int chunk = blockIdx.y;
    const uint64_t f0 = (uint64_t)chunk * fchunk;
    const uint64_t f1 = min(f0 + fchunk, a.n_frame);
    const bool ok = d < a.n_dof;
    const int m = (3 * a.nbins_ref > kMaxGap) ? kMaxGap : 3 * a.nbins_ref;
    const bool gap = pass == 0 && a.dtype == 3 && a.optimize;
    __shared__ uint32_t gh[64 * kMaxGap];
    __shared__ double red[4][64][4];
    if (gap)
        for (int i = threadIdx.x; i < 64 * m; i += 256) gh[i] = 0;
    __syncthreads();
    double v0 = 1.0e100, v1 = -1.0e100, v2 = 0.0, v3 = 1.0e100, v4 = -1.0e100;
    const double bs = (kTwoPi + kDx) / m;
    const double sh = ok && pass == 1 ? a.shift[d] : 0.0;
    // four frames of the thread's stride-4 sequence are loaded before any
    // is used, so each lane keeps four HBM reads in flight
    auto visit = [&](double x) {
        if (pass == 0) {
            v2 += x;
            if (v0 > x) v0 = x;
            if (v1 < x) v1 = x;
            if (gap) {
                const int t1 = (int)fabs(floor(x / bs));
                const uint32_t code = (uint8_t)t1;
                if (code < (uint32_t)m) atomicAdd(&gh[tx * m + code], 1u);
                else atomicOr(a.flag, 2);
            }
        } else {
            const double p = fmod(x + kPi, kTwoPi);
            if (p < v0) v0 = p;
            if (p > v1) v1 = p;
            if (a.optimize) {
                const double o = fmod(x + sh, kTwoPi);
                if (o < v3) v3 = o;
                if (o > v4) v4 = o;
            }
        }
    };
    if (ok) {
        const double* col = a.bat + d;
        uint64_t f = f0 + ty;
        for (; f + 12 < f1; f += 16) {
            const float x0 = (float)col[f * a.ld], x1 = (float)col[(f + 4) * a.ld];
            const float x2 = (float)col[(f + 8) * a.ld], x3 = (float)col[(f + 12) * a.ld];
            visit(x0);
            visit(x1);
            visit(x2);
            visit(x3);
        }
        for (; f < f1; f += 4) visit((double)(float)col[f * a.ld]);
    }
    red[ty][tx][0] = v0;
    red[ty][tx][1] = v1;
    red[ty][tx][2] = v2;
    red[ty][tx][3] = v3;
    __syncthreads();
    // v4 reduced in a second round through the same buffer
    if (ty == 0 && ok) {
        double r0 = red[0][tx][0], r1 = red[0][tx][1], r2 = red[0][tx][2], r3 = red[0][tx][3];
        for (int k = 1; k < 4; ++k) {
            r0 = fmin(r0, red[k][tx][0]);
            r1 = fmax(r1, red[k][tx][1]);
            r2 += red[k][tx][2];
            r3 = fmin(r3, red[k][tx][3]);
        }
        double* out = a.part + ((size_t)chunk * a.n_dof + d) * 8;
        if (pass == 0) {
            out[0] = r0;
            out[1] = r1;
            out[2] = r2;
        } else {
            out[3] = r0;
            out[4] = r1;
            out[5] = r3;
        }
    }
    __syncthreads();
    red[ty][tx][0] = v4;
    __syncthreads();
    if (ty == 0 && ok && pass == 1) {
        double r4 = red[0][tx][0];
        for (int k = 1; k < 4; ++k) r4 = fmax(r4, red[k][tx][0]);
        a.part[((size_t)chunk * a.n_dof + d) * 8 + 6] = r4;
    }
    if (gap) {
        __syncthreads();
        for (int i = threadIdx.x; i < 64 * m; i += 256) {
            const uint32_t dd = blockIdx.x * 64 + i / m;
            if (dd < a.n_dof && gh[i]) atomicAdd(&a.gaphist[(size_t)dd * kMaxGap + i % m], gh[i]);
        }
    }
}

// Largest empty circular run of the gap histogram, else the least-populated
// bin; returns 2pi - centre[minidx] (Real2Int.cpp:280-350).
__device__ double gap_shift(const uint32_t* prob, int m) {
    const double bs = (kTwoPi + kDx) / m;
    unsigned nzero = 0;
    for (int i = 0; i < m; ++i) nzero += prob[i] == 0;
    unsigned minidx = 0;
    if (nzero == 0) {
        double rec = (double)prob[0];
        for (int j = 1; j < m; ++j)
            if ((double)prob[j] < rec) { rec = (double)prob[j]; minidx = j; }
    } else {
        int j = 0;
        while (j < m && prob[j] != 0) ++j;
        minidx = j;
        unsigned cnt = 0, rec = 0;
        for (int jj = j; jj < 2 * m; ++jj) {
            const bool z = prob[jj % m] == 0;
            if (z) {
                cnt++;
            } else {
                if (cnt > rec) { minidx = jj - cnt; rec = cnt; }
                cnt = 0;
            }
        }
    }
    return kTwoPi - (0.0 + (minidx + 0.5) * bs);
}

__global__ void finalize_kernel(DiscArgs a, int nchunk, int pass) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= a.n_dof) return;
    const double* p = a.part + (size_t)d * 8;
    const size_t cs = (size_t)a.n_dof * 8;
    double* e = a.ext + 4 * (size_t)d;
    if (pass == 0) {
        // four chunk records per iteration, in independent accumulators
        double mn4[4] = {1.0e100, 1.0e100, 1.0e100, 1.0e100};
        double mx4[4] = {-1.0e100, -1.0e100, -1.0e100, -1.0e100};
        double sm4[4] = {0.0, 0.0, 0.0, 0.0};
        int c = 0;
        for (; c + 4 <= nchunk; c += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                mn4[u] = fmin(mn4[u], p[(c + u) * cs + 0]);
                mx4[u] = fmax(mx4[u], p[(c + u) * cs + 1]);
                sm4[u] += p[(c + u) * cs + 2];
            }
        }
        for (; c < nchunk; ++c) {
            mn4[0] = fmin(mn4[0], p[c * cs + 0]);
            mx4[0] = fmax(mx4[0], p[c * cs + 1]);
            sm4[0] += p[c * cs + 2];
        }
        double mn = fmin(fmin(mn4[0], mn4[1]), fmin(mn4[2], mn4[3]));
        double mx = fmax(fmax(mx4[0], mx4[1]), fmax(mx4[2], mx4[3]));
        const double sum = (sm4[0] + sm4[1]) + (sm4[2] + sm4[3]);
        const double avg = sum / (double)a.n_frame;
        if (a.dtype == 1 || a.dtype == 2) {
            mn -= kDx;
            if (mn < 0.0) mn = 0.0;
            mx += kDx;
            if (a.dtype == 2 && mx > kPi) mx = kPi;
            if ((mx - mn) < 1.0e-4) mx += 0.05;
            e[0] = mn; e[1] = mx; e[2] = avg; e[3] = 0.0;
        } else {
            // raw range with the reference's starting values (6.5, 0.0)
            e[0] = fmin(6.5, mn);
            e[1] = fmax(0.0, mx);
            e[2] = avg;
            const int m = (3 * a.nbins_ref > kMaxGap) ? kMaxGap : 3 * a.nbins_ref;
            a.shift[d] = a.optimize ? gap_shift(a.gaphist + (size_t)d * kMaxGap, m) : 0.0;
        }
        return;
    }
    // pass 1 (torsions): choose the range
    double mn2 = 6.5, mx2 = 0.0, mn3 = 1.0e100, mx3 = 0.0;
#pragma unroll 4
    for (int c = 0; c < nchunk; ++c) {
        mn2 = fmin(mn2, p[c * cs + 3]);
        mx2 = fmax(mx2, p[c * cs + 4]);
        mn3 = fmin(mn3, p[c * cs + 5]);
        mx3 = fmax(mx3, p[c * cs + 6]);
    }
    const double mn1 = e[0], mx1 = e[1], avg0 = e[2];
    double mn, mx;
    int dp;
    if (!a.optimize) {
        if (fabs((mx1 - mn1) - (mx2 - mn2)) <= 1.0e-1) { mn = mn1; mx = mx1; dp = 0; }
        else if ((mx1 - mn1) < (mx2 - mn2)) { mn = mn1; mx = mx1; dp = 0; }
        else { mn = mn2; mx = mx2; dp = 1; }
    } else {
        mn = mn3; mx = mx3; dp = 2;
    }
    mn -= kDx;
    if (mn < 0.0) mn = 0.0;
    mx += kDx;
    if (mx > kTwoPi) mx = kTwoPi;
    if ((mx - mn) < 1.0e-4) mx += 0.05;
    double avg = avg0;
    if (dp == 1) avg = fmod(avg0 + kPi, kTwoPi);
    else if (dp == 2) avg = fmod(avg0 + a.shift[d], kTwoPi);
    e[0] = mn; e[1] = mx; e[2] = avg;
    e[3] = a.optimize ? a.shift[d] : dp * kPi;
    a.shift[d] = (dp == 0) ? -1.0 : (dp == 1 ? kPi : a.shift[d]);
}

// codes[d][sch][f]; shift < 0 means unshifted (bonds, angles, dp == 0).
__global__ __launch_bounds__(256) void bin_kernel(DiscArgs a) {
    __shared__ uint8_t tile[64][64 + 4];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const uint32_t d = blockIdx.x * 64 + tx;
    const uint64_t f0 = (uint64_t)blockIdx.y * 64;
    const bool ok = d < a.n_dof;
    double mn = 0, mx = 0, sh = -1.0;
    if (ok) {
        mn = a.ext[4 * (size_t)d];
        mx = a.ext[4 * (size_t)d + 1];
        if (a.dtype == 3) sh = a.shift[d];
    }
    double xs[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint64_t f = f0 + ty + 4 * k;
        double x = 0.0;
        if (ok && f < a.n_frame) {
            x = (double)(float)a.bat[f * a.ld + d];
            if (sh >= 0.0) x = fmod(x + sh, kTwoPi);
        }
        xs[k] = x;
    }
    for (int s = 0; s < a.nsch; ++s) {
        const int nb = a.schemes[s];
        const double w = (mx - mn) / nb;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            uint8_t t = (uint8_t)(int)floor((xs[k] - mn) / w);
            if (t == nb) t = (uint8_t)(t - 1);
            tile[tx][ty + 4 * k] = t;
        }
        __syncthreads();
        // write: 64 DOF rows x 64 frames, 4 threads per row, 16 frames each
        const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
        const uint32_t dd = blockIdx.x * 64 + row;
        if (dd < a.n_dof) {
            uint8_t* out = a.codes + ((size_t)dd * a.nsch + s) * a.n_frame + f0;
            for (int k = 0; k < 16; ++k) {
                const uint64_t f = f0 + part * 16 + k;
                if (f < a.n_frame) out[part * 16 + k] = tile[row][part * 16 + k];
            }
        }
        __syncthreads();
    }
}

}  // namespace

// Frame chunks of the stats passes: about 1,024 workgroups over the chip
// (four per CU), each chunk at least 512 frames, at most 128 chunks so the
// per-DOF finalize loop stays short.
int disc_chunks(uint64_t n_frame, uint32_t n_dof) {
    const uint64_t groups = (n_dof + 63) / 64;
    uint64_t c = (1024 + groups - 1) / groups;
    const uint64_t cap = (n_frame + 511) / 512;
    if (c > cap) c = cap;
    if (c > 128) c = 128;
    return (int)(c < 1 ? 1 : c);
}

hipError_t run_discretize_hist(const DiscArgs& a, int nchunk, hipStream_t st) {
    const uint64_t fchunk = (a.n_frame + nchunk - 1) / nchunk;
    const dim3 gs((a.n_dof + 63) / 64, nchunk);
    const dim3 gf((a.n_dof + 255) / 256);
    hipError_t e;
    if (a.dtype == 3 && a.optimize) {
        e = hipMemsetAsync(a.gaphist, 0, sizeof(uint32_t) * kMaxGap * a.n_dof, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(stats_kernel, gs, dim3(256), 0, st, a, 0, fchunk);
    hipLaunchKernelGGL(finalize_kernel, gf, dim3(256), 0, st, a, nchunk, 0);
    if (a.dtype == 3) {
        hipLaunchKernelGGL(stats_kernel, gs, dim3(256), 0, st, a, 1, fchunk);
        hipLaunchKernelGGL(finalize_kernel, gf, dim3(256), 0, st, a, nchunk, 1);
    }
    const dim3 gb((a.n_dof + 63) / 64, (unsigned)((a.n_frame + 63) / 64));
    hipLaunchKernelGGL(bin_kernel, gb, dim3(256), 0, st, a);
    return hipGetLastError();
}
