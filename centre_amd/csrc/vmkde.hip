// von Mises KDE conformer-state discretizer for torsions.
//
// Replaces Real2Int::getMinimaPositions / vonMises (Real2Int.cpp:481-698)
// and the Bessel functions they use (Utils.cpp:60-168).
//
//   prep     float32 rounding of the [frame][dof] float64 input (as
//            nc_get_vara_float), transposed to [dof][frame], plus cos/sin of
//            every frame angle.
//   grid     gradient and second derivative of the KDE at G+1 points
//            (0..360 degrees, G = 360 is the reference's 1-degree grid):
//            sum_f exp(h cos(X - a_f)) * {-sin(X - a_f), h sin^2 - cos},
//            with cos/sin(X - a) by the angle-difference identities from the
//            per-frame and per-point cos/sin.  FP64 throughout; frames are
//            staged in LDS and shared by the grid points of a workgroup.
//            For even G the points half a turn apart share one reduction
//            and one polynomial pair (exp(y) and exp(-y)).
//   states   one lane per torsion: maxima (gradient sign change with second
//            derivative < -1e-4, at most 9), the two selection sorts, the
//            initial minima, steepest descent on the linearly interpolated
//            grid gradient, final sort (Real2Int.cpp:518-634).
//   assign   state code of every frame from the sorted minima; frames exactly
//            on a minimum keep code 0 (Real2Int.cpp:636-653).
#include "centre_internal.h"

namespace {

constexpr double kDegRad = 0.017453292519943295;

// 64 frames x 64 torsions per workgroup, transposed through LDS so that both
// the [frame][dof] reads and the [dof][frame] writes coalesce
__global__ __launch_bounds__(256) void vm_prep_kernel(VmkdeArgs a) {
    __shared__ float tile[64][65];
    const uint64_t f0 = (uint64_t)blockIdx.x * 64;
    const uint32_t d0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const uint64_t f = f0 + r;
        const uint32_t d = d0 + tx;
        tile[r][tx] = (f < a.n_frame && d < a.n_dof) ? (float)a.bat[f * a.ld + d] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const uint32_t d = d0 + r;
        const uint64_t f = f0 + tx;
        if (d < a.n_dof && f < a.n_frame) {
            const float x = tile[tx][r];
            const uint64_t o = (uint64_t)d * a.n_frame + f;
            a.xt[o] = x;
            double sn, cs;
            sincos((double)x, &sn, &cs);
            reinterpret_cast<double2*>(a.trig)[o] = make_double2(cs, sn);
        }
    }
}

constexpr int kStage = 1024;  // frames staged per LDS pass

// A workgroup covers np grid points of one torsion with 256 / np frame lanes
// per point (threads past np * lanes only stage frames); see grid_split.
__global__ __launch_bounds__(256) void vm_grid_kernel(VmkdeArgs a, double h, double denom, int np) {
    __shared__ double2 fr[kStage];
    __shared__ double red[256][2];
    const int G = a.grid;
    const int nl = 256 / np;                       // frame lanes per point
    const int pl = threadIdx.x % np, fl = threadIdx.x / np;
    const int p = blockIdx.x * np + pl;
    const bool live = fl < nl && p <= G;
    double cx = 0, sx = 0;
    if (live) sincos((p * (360.0 / G)) * kDegRad, &sx, &cx);
    for (uint32_t d = blockIdx.y; d < a.n_dof; d += gridDim.y) {
        double g = 0.0, sd = 0.0;
        const double2* trig = reinterpret_cast<const double2*>(a.trig) + (uint64_t)d * a.n_frame;
        for (uint64_t f0 = 0; f0 < a.n_frame; f0 += kStage) {
            const int nf = (int)min((uint64_t)kStage, a.n_frame - f0);
            __syncthreads();
            for (int i = threadIdx.x; i < nf; i += 256) fr[i] = trig[f0 + i];
            __syncthreads();
            if (live) {
                for (int i = fl; i < nf; i += nl) {
                    const double ca = fr[i].x, sa = fr[i].y;
                    const double c = cx * ca + sx * sa;      // cos(X - a)
                    const double s = sx * ca - cx * sa;      // sin(X - a)
                    const double e = exp(h * c);
                    g = g - s * e;
                    sd = sd + (h * s * s - c) * e;
                }
            }
        }
        red[threadIdx.x][0] = g;
        red[threadIdx.x][1] = sd;
        __syncthreads();
        if (fl == 0 && p <= G) {
            double gg = 0.0, ss = 0.0;
            for (int k = 0; k < nl; ++k) {
                gg += red[k * np + pl][0];
                ss += red[k * np + pl][1];
            }
            double* out = a.gridbuf + (uint64_t)d * 2 * (G + 1);
            out[p] = gg * h / denom;
            out[G + 1 + p] = ss * h / denom;
        }
    }
}

// exp(y) and exp(-y) together, for the paired grid points below.  Cody-Waite
// reduction y = k ln2 + r, |r| <= ln2/2 (ln2 split into its double rounding
// and the remainder), then exp(+-r) = C(r) +- S(r) from the Taylor series of
// cosh (to r^12) and sinh (to r^13), whose first omitted terms are below
// 5e-18 on that interval; 2^(+-k) applied by ldexp.  Within 2 ulp of the
// correctly rounded values; overflow and underflow as ldexp gives them.
// a * b + c in one v_fma_f64 with c held in an SGPR pair: written as the
// compiler's own fma, a Horner step with a constant addend becomes a VGPR copy
// of the constant plus v_fmac_f64.  Arithmetic only (no memory operand).
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

__device__ __forceinline__ void exp_pm(double y, double& ep, double& em) {
    constexpr double kLog2e = 1.4426950408889634;
    constexpr double kLn2 = 0.6931471805599453;
    constexpr double kLn2Lo = 2.3190468138462996e-17;
    const double kf = rint(y * kLog2e);
    double r = fma(-kf, kLn2, y);
    r = fma(-kf, kLn2Lo, r);
    const int k = (int)kf;
    const double u = r * r;
    // 1/(2j)! and 1/(2j+1)!, j = 6..0
    double c = 2.08767569878681e-09;
    c = fma_sc(c, u, 2.755731922398589e-07);
    c = fma_sc(c, u, 2.48015873015873e-05);
    c = fma_sc(c, u, 0.001388888888888889);
    c = fma_sc(c, u, 0.041666666666666664);
    c = fma_sc(c, u, 0.5);
    c = fma_sc(c, u, 1.0);
    double o = 1.6059043836821613e-10;
    o = fma_sc(o, u, 2.505210838544172e-08);
    o = fma_sc(o, u, 2.7557319223985893e-06);
    o = fma_sc(o, u, 0.0001984126984126984);
    o = fma_sc(o, u, 0.008333333333333333);
    o = fma_sc(o, u, 0.16666666666666666);
    o = fma_sc(o, u, 1.0);
    o = o * r;
    ep = ldexp(c + o, k);
    em = ldexp(c - o, -k);
}

// Even G: point p and point p + G/2 lie half a turn apart, so their
// cos(X - a) and sin(X - a) differ only in sign and one exp_pm serves both.
// A thread owns base point q in 0..G/2 and its partner q + G/2 (the partner
// of q = 0 is base point G/2 and is not written); otherwise as vm_grid_kernel.
__global__ __launch_bounds__(256) void vm_grid_pair_kernel(VmkdeArgs a, double h, double denom,
                                                           int np) {
    __shared__ double2 fr[kStage];
    __shared__ double red[256][4];
    const int G = a.grid;
    const int Q = G / 2 + 1;
    const int nl = 256 / np;
    const int pl = threadIdx.x % np, fl = threadIdx.x / np;
    const int q = blockIdx.x * np + pl;
    const bool live = fl < nl && q < Q;
    double cx = 0, sx = 0;
    if (live) sincos((q * (360.0 / G)) * kDegRad, &sx, &cx);
    for (uint32_t d = blockIdx.y; d < a.n_dof; d += gridDim.y) {
        double g0 = 0.0, s0 = 0.0, g1 = 0.0, s1 = 0.0;
        const double2* trig = reinterpret_cast<const double2*>(a.trig) + (uint64_t)d * a.n_frame;
        for (uint64_t f0 = 0; f0 < a.n_frame; f0 += kStage) {
            const int nf = (int)min((uint64_t)kStage, a.n_frame - f0);
            __syncthreads();
            for (int i = threadIdx.x; i < nf; i += 256) fr[i] = trig[f0 + i];
            __syncthreads();
            if (live) {
                for (int i = fl; i < nf; i += nl) {
                    const double ca = fr[i].x, sa = fr[i].y;
                    const double c = cx * ca + sx * sa;      // cos(X - a)
                    const double s = sx * ca - cx * sa;      // sin(X - a)
                    double e, em;
                    exp_pm(h * c, e, em);
                    const double hss = h * s * s;
                    g0 = g0 - s * e;
                    s0 = s0 + (hss - c) * e;
                    g1 = g1 + s * em;                        // sin, cos negated
                    s1 = s1 + (hss + c) * em;
                }
            }
        }
        red[threadIdx.x][0] = g0;
        red[threadIdx.x][1] = s0;
        red[threadIdx.x][2] = g1;
        red[threadIdx.x][3] = s1;
        __syncthreads();
        if (fl == 0 && q < Q) {
            double v[4] = {0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < nl; ++k)
                for (int j = 0; j < 4; ++j) v[j] += red[k * np + pl][j];
            double* out = a.gridbuf + (uint64_t)d * 2 * (G + 1);
            out[q] = v[0] * h / denom;
            out[G + 1 + q] = v[1] * h / denom;
            if (q > 0) {
                out[q + G / 2] = v[2] * h / denom;
                out[G + 1 + q + G / 2] = v[3] * h / denom;
            }
        }
    }
}

// Points per workgroup for npts grid points: the fewest frame passes per
// torsion, ceil(npts/np) workgroups x F/floor(256/np) frames each, with at
// most 32 workgroups per torsion (each re-reads the torsion's frames).
int grid_split(int npts) {
    int best = 32;
    double cost = 1e300;
    for (int np = 8; np <= 128; ++np) {
        const int nwg = (npts + np - 1) / np;
        if (nwg > 32) continue;
        const double c = (double)nwg / (double)(256 / np);
        if (c < cost - 1e-12) {
            cost = c;
            best = np;
        }
    }
    return best;
}

// Seven torsions per 64-lane workgroup, nine lanes each.  The first lane of
// a group finds the maxima, sorts them and places the initial minima; each
// minimum then descends on its own lane; the first lane sorts the result.
constexpr int kStatesPerBlock = 7;

__global__ __launch_bounds__(64) void vm_states_kernel(VmkdeArgs a) {
    // every product and sum rounded on its own, as the reference's FP64
    // expressions are (no fused multiply-add in the descent step)
#pragma clang fp contract(off)
    __shared__ double s_cmax[kStatesPerBlock][9], s_smax[kStatesPerBlock][9];
    __shared__ double s_mins[kStatesPerBlock][9];
    __shared__ int s_nf[kStatesPerBlock];
    const int grp = threadIdx.x / 9, lane = threadIdx.x % 9;
    const uint32_t d = blockIdx.x * kStatesPerBlock + grp;
    const bool valid = grp < kStatesPerBlock && d < a.n_dof;
    const int G = a.grid;
    const double hdeg = 360.0 / G;
    const double* gr = a.gridbuf + (uint64_t)(valid ? d : 0) * 2 * (G + 1);
    const double* sd = gr + (G + 1);
    if (valid && lane == 0) {
        double* cmax = s_cmax[grp];
        double* smax = s_smax[grp];
        double* mins = s_mins[grp];
        int nobs = 0;
        for (int i = 0; i < 9; ++i) { cmax[i] = 0.0; smax[i] = 0.0; mins[i] = 0.0; }
        for (int i = 1; i <= G; ++i) {
            if ((gr[i - 1] * gr[i]) <= 0.0 && sd[i] < -1.0e-4 && nobs < 9) {
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
        const int nf = (a.nmaxconf < nobs) ? a.nmaxconf : nobs;
        for (int i = 0; i < nf - 1; ++i)
            for (int j = i + 1; j < nf; ++j)
                if (cmax[i] > cmax[j]) { t = cmax[i]; cmax[i] = cmax[j]; cmax[j] = t; }
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
        s_nf[grp] = nf;
    }
    __syncthreads();
    // steepest descent of one minimum on the linearly interpolated gradient
    if (valid && s_nf[grp] > 1 && lane < s_nf[grp]) {
        double x = s_mins[grp][lane];
        uint32_t it = 0;
        bool conv = false;
        while (!conv && it < a.sdo_iter) {
            it++;
            const double u = x / hdeg;
            int iu = (int)u;
            // x in [0, 360) gives 0 <= iu <= G (G only when x / hdeg rounds up
            // to G, where gr[1 + iu], the first second derivative, has weight
            // 0); the clamp only guards a NaN or infinite grid
            iu = iu < 0 ? 0 : (iu > G ? G : iu);
            const double gg = fabs(u - 1 - iu) * gr[iu] + fabs(u - iu) * gr[1 + iu];
            if (fabs(gg) < a.sdo_conv) {
                conv = true;
            } else {
                const double st0 = a.sdo_step * gg;
                const double st = (fabs(st0) > 5.0) ? (st0 / fabs(st0)) * 5.0 : st0;
                x = x - st;
                if (x < 0.0) x += 360.0;
                if (x >= 360.0) x -= 360.0;
            }
        }
        s_mins[grp][lane] = x;
    }
    __syncthreads();
    if (valid && lane == 0) {
        double* mins = s_mins[grp];
        const int nf = s_nf[grp];
        double t;
        for (int i = 0; i < nf - 1; ++i)
            for (int j = i + 1; j < nf; ++j)
                if (mins[i] > mins[j]) { t = mins[i]; mins[i] = mins[j]; mins[j] = t; }
        for (int i = 0; i < 9; ++i)
            a.minima[(uint64_t)d * 9 + i] = (i < nf) ? mins[i] * kDegRad : 0.0;
        a.nfound[d] = (uint8_t)nf;
    }
}

// kAssignSpan frames along x per workgroup (each thread every 256th of
// them, so loads and stores coalesce), torsions along y (strided past the
// grid's y extent)
constexpr int kAssignSpan = 4096;

__global__ __launch_bounds__(256) void vm_assign_kernel(VmkdeArgs a) {
    const uint64_t fb = (uint64_t)blockIdx.x * kAssignSpan + threadIdx.x;
    for (uint32_t d = blockIdx.y; d < a.n_dof; d += gridDim.y) {
        const int nf = a.nfound[d];
        double m[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) m[j] = a.minima[(uint64_t)d * 9 + j];
        const float* x = a.xt + (uint64_t)d * a.n_frame;
        uint8_t* out = a.codes + (uint64_t)d * a.n_frame;
        for (int k = 0; k < kAssignSpan / 256; ++k) {
            const uint64_t f = fb + (uint64_t)k * 256;
            if (f >= a.n_frame) break;
            const double v = x[f];
            // code j for m[j-1] < v < m[j]; the reference's j = 0 branch
            // (outside [m[0], m[nf-1]] -> 0) only restates the initial 0
            uint8_t code = 0;
#pragma unroll
            for (int j = 1; j < 9; ++j)
                if (j < nf && v < m[j] && v > m[j - 1]) code = (uint8_t)j;
            out[f] = code;
        }
    }
}

}  // namespace

hipError_t run_vmkde(const VmkdeArgs& a, hipStream_t st) {
    // bandwidth and normaliser are the same for every torsion (they depend
    // on the frame count and kappa only); computed on the host in FP64.
    const double h = centre_vm_bandwidth(a.n_frame, a.kappa);
    const double denom = 6.283185307179586 * (double)a.n_frame * centre_bessel_i(0, h);
    const unsigned ny = a.n_dof < 65535u ? a.n_dof : 65535u;  // torsions per y pass
    hipLaunchKernelGGL(vm_prep_kernel, dim3((unsigned)((a.n_frame + 63) / 64), (a.n_dof + 63) / 64),
                       dim3(256), 0, st, a);
    if (a.grid % 2 == 0) {
        const int Q = a.grid / 2 + 1;
        const int np = grid_split(Q);
        hipLaunchKernelGGL(vm_grid_pair_kernel, dim3((Q + np - 1) / np, ny), dim3(256), 0, st, a,
                           h, denom, np);
    } else {
        const int np = grid_split(a.grid + 1);
        hipLaunchKernelGGL(vm_grid_kernel, dim3((a.grid + 1 + np - 1) / np, ny), dim3(256), 0, st,
                           a, h, denom, np);
    }
    hipLaunchKernelGGL(vm_states_kernel, dim3((a.n_dof + kStatesPerBlock - 1) / kStatesPerBlock),
                       dim3(64), 0, st, a);
    hipLaunchKernelGGL(vm_assign_kernel, dim3((unsigned)((a.n_frame + kAssignSpan - 1) / kAssignSpan), ny),
                       dim3(256), 0, st, a);
    return hipGetLastError();
}
