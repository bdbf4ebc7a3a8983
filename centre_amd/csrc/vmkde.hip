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
    const uint32_t d = blockIdx.y;
    const int nl = 256 / np;                       // frame lanes per point
    const int pl = threadIdx.x % np, fl = threadIdx.x / np;
    const int p = blockIdx.x * np + pl;
    const bool live = fl < nl && p <= G;
    double cx = 0, sx = 0;
    if (live) sincos((p * (360.0 / G)) * kDegRad, &sx, &cx);
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

// Points per workgroup for G+1 grid points: the fewest frame passes per
// torsion, ceil((G+1)/np) workgroups x F/floor(256/np) frames each, with at
// most 32 workgroups per torsion (each re-reads the torsion's frames).
int grid_split(int G) {
    int best = 32;
    double cost = 1e300;
    for (int np = 8; np <= 128; ++np) {
        const int nwg = (G + 1 + np - 1) / np;
        if (nwg > 32) continue;
        const double c = (double)nwg / (double)(256 / np);
        if (c < cost - 1e-12) {
            cost = c;
            best = np;
        }
    }
    return best;
}

__global__ void vm_states_kernel(VmkdeArgs a) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= a.n_dof) return;
    const int G = a.grid;
    const double hdeg = 360.0 / G;
    const double* gr = a.gridbuf + (uint64_t)d * 2 * (G + 1);
    const double* sd = gr + (G + 1);
    double cmax[9], smax[9];
    int nobs = 0;
    for (int i = 0; i < 9; ++i) { cmax[i] = 0.0; smax[i] = 0.0; }
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
    double mins[9];
    for (int i = 0; i < 9; ++i) mins[i] = 0.0;
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
            uint32_t it = 0;
            bool conv = false;
            while (!conv && it < a.sdo_iter) {
                it++;
                const double u = mins[m] / hdeg;
                const int iu = (int)u;
                const double gg = fabs(u - 1 - iu) * gr[iu] + fabs(u - iu) * gr[1 + iu];
                if (fabs(gg) < a.sdo_conv) {
                    conv = true;
                } else {
                    const double st0 = a.sdo_step * gg;
                    const double st = (fabs(st0) > 5.0) ? (st0 / fabs(st0)) * 5.0 : st0;
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
    for (int i = 0; i < 9; ++i) a.minima[(uint64_t)d * 9 + i] = (i < nf) ? mins[i] * kDegRad : 0.0;
    a.nfound[d] = (uint8_t)nf;
}

__global__ void vm_assign_kernel(VmkdeArgs a) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (uint64_t)a.n_dof * a.n_frame;
    if (t >= n) return;
    const uint64_t d = t / a.n_frame;
    const int nf = a.nfound[d];
    const double* mins = a.minima + d * 9;
    const double v = a.xt[t];
    uint8_t code = 0;
    if (nf > 1) {
        for (int j = 0; j < nf; ++j) {
            if (j == 0) {
                if (v < mins[0] || v > mins[nf - 1]) code = 0;
            } else if (v < mins[j] && v > mins[j - 1]) {
                code = (uint8_t)j;
            }
        }
    }
    a.codes[t] = code;
}

}  // namespace

hipError_t run_vmkde(const VmkdeArgs& a, hipStream_t st) {
    // bandwidth and normaliser are the same for every torsion (they depend
    // on the frame count and kappa only); computed on the host in FP64.
    const double h = centre_vm_bandwidth(a.n_frame, a.kappa);
    const double denom = 6.283185307179586 * (double)a.n_frame * centre_bessel_i(0, h);
    const uint64_t n = (uint64_t)a.n_dof * a.n_frame;
    hipLaunchKernelGGL(vm_prep_kernel, dim3((unsigned)((a.n_frame + 63) / 64), (a.n_dof + 63) / 64),
                       dim3(256), 0, st, a);
    const int np = grid_split(a.grid);
    hipLaunchKernelGGL(vm_grid_kernel, dim3((a.grid + 1 + np - 1) / np, a.n_dof), dim3(256), 0, st,
                       a, h, denom, np);
    hipLaunchKernelGGL(vm_states_kernel, dim3((a.n_dof + 63) / 64), dim3(64), 0, st, a);
    hipLaunchKernelGGL(vm_assign_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}
