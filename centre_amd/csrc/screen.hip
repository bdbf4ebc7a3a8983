// Learned pair screen: cheap per-pair features and a small MLP that predicts
// the plain path's pair MI, so that pairs predicted below a cutoff can be left
// out of the full pass.  No reference counterpart: the reference's A-MIE /
// A-MIST restrict pairs by a neighbour list given by the user
// (Centre.cpp:459-593); this screen derives the list from the data instead.
//
// Features of a pair (i < j), all from device-resident codes:
//   0, 1  S1 of i and of j (ML, all frames)
//   2     circular correlation of the bin angles theta = 2 pi (code + 1/2) / nb
//         on a frame subsample (Jammalamadaka-SenGupta)
//   3     linear (Pearson) correlation of the codes on the subsample
//   4     MI on the subsample (ML plug-in) with the bins merged to
//         nbc = min(nb, 8): code * nbc / nb; computed by the pair kernel's
//         small-bin tile (plug-in bias (nbc - 1)^2 / (2 n_sub) instead of
//         (nb - 1)^2 / (2 n_sub))
//   5, 6  DOF types of i and j (1 bond, 2 angle, 3 torsion, scaled by 1/3)
//   7     1 when i and j have the same type
// Network: x = (f - mu) / sd, h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2),
// y = w3 . h2 + b3 (FP32); y predicts log(MI).  Loss 0.5 (y - t)^2 per sample.
#include "centre_internal.h"

namespace {

// Frames k * step of every DOF; with coarse != nullptr also the merged codes
// code * nbc / nb.
__global__ __launch_bounds__(256) void subsample_kernel(const uint8_t* __restrict__ codes,
                                                        uint64_t stride, uint64_t step,
                                                        uint32_t n_sub, uint8_t* __restrict__ out,
                                                        uint8_t* __restrict__ coarse, int nb,
                                                        int nbc) {
    const uint32_t d = blockIdx.y;
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n_sub) return;
    const uint8_t c = codes[(uint64_t)d * stride + (uint64_t)k * step];
    out[(uint64_t)d * n_sub + k] = c;
    if (coarse) coarse[(uint64_t)d * n_sub + k] = (uint8_t)((int)c * nbc / nb);
}

// Per DOF over the subsample: standardised linear values X and circular
// deviations Y, scaled so that the dot product of two rows is the
// correlation coefficient.
__global__ __launch_bounds__(256) void dof_moments_kernel(const uint8_t* __restrict__ sub,
                                                          uint32_t n_sub, int nb,
                                                          float* __restrict__ X,
                                                          float* __restrict__ Y) {
    __shared__ double red[4][256];
    const uint32_t d = blockIdx.x;
    const int tid = threadIdx.x;
    const uint8_t* row = sub + (uint64_t)d * n_sub;
    const double w = 6.283185307179586 / nb;
    double s1 = 0, s2 = 0, ss = 0, sc = 0;
    for (uint32_t k = tid; k < n_sub; k += 256) {
        const double c = row[k];
        s1 += c;
        s2 += c * c;
        double sn, cs;
        sincos((c + 0.5) * w, &sn, &cs);
        ss += sn;
        sc += cs;
    }
    red[0][tid] = s1;
    red[1][tid] = s2;
    red[2][tid] = ss;
    red[3][tid] = sc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o)
            for (int q = 0; q < 4; ++q) red[q][tid] += red[q][tid + o];
        __syncthreads();
    }
    const double n = n_sub;
    const double mean = red[0][0] / n;
    const double var = red[1][0] / n - mean * mean;
    const double mu = atan2(red[2][0], red[3][0]);
    __syncthreads();
    double sq = 0;
    for (uint32_t k = tid; k < n_sub; k += 256) {
        const double v = sin((row[k] + 0.5) * w - mu);
        sq += v * v;
    }
    red[0][tid] = sq;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[0][tid] += red[0][tid + o];
        __syncthreads();
    }
    const double lin_scale = var > 1e-12 ? 1.0 / sqrt(var * n) : 0.0;
    const double circ_scale = red[0][0] > 1e-12 ? 1.0 / sqrt(red[0][0]) : 0.0;
    for (uint32_t k = tid; k < n_sub; k += 256) {
        const double c = row[k];
        X[(uint64_t)d * n_sub + k] = (float)((c - mean) * lin_scale);
        Y[(uint64_t)d * n_sub + k] = (float)(sin((c + 0.5) * w - mu) * circ_scale);
    }
}

__device__ inline uint64_t tri(uint64_t i, uint64_t j, uint64_t D) {
    return i * D - i * (i + 1) / 2 + (j - i - 1);
}

// Row dot products of X and of Y for all i < j: 64 x 64 DOF blocks, 256
// threads with 4 x 4 outputs each, frames staged through LDS 32 at a time.
constexpr int kCB = 64, kCK = 32;
__global__ __launch_bounds__(256) void corr_kernel(const float* __restrict__ X,
                                                   const float* __restrict__ Y, uint32_t D,
                                                   uint32_t n, const int2* __restrict__ blocks,
                                                   float* __restrict__ out_lin,
                                                   float* __restrict__ out_circ) {
    __shared__ float xa[kCK][kCB + 1], xb[kCK][kCB + 1], ya[kCK][kCB + 1], yb[kCK][kCB + 1];
    const int2 blk = blocks[blockIdx.x];
    const uint32_t i0 = blk.x * kCB, j0 = blk.y * kCB;
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    float al[4][4] = {}, ac[4][4] = {};
    for (uint32_t k0 = 0; k0 < n; k0 += kCK) {
        // stage: 64 rows x 32 frames per array; 256 threads x 8 elements
        for (int e = tid; e < kCB * kCK; e += 256) {
            const int r = e / kCK, k = e % kCK;
            const uint32_t kk = k0 + k;
            const bool okk = kk < n;
            const uint32_t ri = i0 + r, rj = j0 + r;
            xa[k][r] = (okk && ri < D) ? X[(uint64_t)ri * n + kk] : 0.f;
            ya[k][r] = (okk && ri < D) ? Y[(uint64_t)ri * n + kk] : 0.f;
            xb[k][r] = (okk && rj < D) ? X[(uint64_t)rj * n + kk] : 0.f;
            yb[k][r] = (okk && rj < D) ? Y[(uint64_t)rj * n + kk] : 0.f;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < kCK; ++k) {
            float a[4], b[4], c[4], d[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = xa[k][ti * 4 + q];
                b[q] = xb[k][tj * 4 + q];
                c[q] = ya[k][ti * 4 + q];
                d[q] = yb[k][tj * 4 + q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    al[p][q] = fmaf(a[p], b[q], al[p][q]);
                    ac[p][q] = fmaf(c[p], d[q], ac[p][q]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t i = i0 + ti * 4 + p, j = j0 + tj * 4 + q;
            if (i < j && j < D) {
                const uint64_t c = tri(i, j, D);
                out_lin[c] = al[p][q];
                out_circ[c] = ac[p][q];
            }
        }
}

__global__ __launch_bounds__(256) void feature_kernel(uint32_t D, const double* __restrict__ s1,
                                                      const uint8_t* __restrict__ types,
                                                      const double* __restrict__ mi_sub,
                                                      const float* __restrict__ lin,
                                                      const float* __restrict__ circ,
                                                      float* __restrict__ feat) {
    const uint32_t i = blockIdx.y;
    const uint32_t j = i + 1 + blockIdx.x * 256 + threadIdx.x;
    if (j >= D) return;
    const uint64_t p = tri(i, j, D);
    float* f = feat + p * kScreenFeat;
    f[0] = (float)s1[i];
    f[1] = (float)s1[j];
    f[2] = circ[p];
    f[3] = lin[p];
    f[4] = (float)mi_sub[p];
    f[5] = types[i] * (1.f / 3.f);
    f[6] = types[j] * (1.f / 3.f);
    f[7] = types[i] == types[j] ? 1.f : 0.f;
}

// ---- MLP -------------------------------------------------------------------
constexpr int kF = kScreenFeat, kH = kScreenHidden;

// Parameters are read straight from global memory: every index below is
// uniform across the wave, so the reads become scalar loads shared by all
// lanes.
__device__ inline void normalise(const float* sp, const float* __restrict__ f, float (&x)[kF]) {
#pragma unroll
    for (int k = 0; k < kF; ++k) x[k] = (f[k] - sp[kOffMu + k]) / sp[kOffSd + k];
}

__device__ inline void layer1(const float* sp, const float (&x)[kF], float (&h1)[kH]) {
#pragma unroll
    for (int a = 0; a < kH; ++a) {
        float s = sp[kOffB1 + a];
#pragma unroll
        for (int k = 0; k < kF; ++k) s = fmaf(sp[kOffW1 + a * kF + k], x[k], s);
        h1[a] = fmaxf(s, 0.f);
    }
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* __restrict__ params,
                                                          const float* __restrict__ feat,
                                                          uint64_t n, float* __restrict__ out) {
    const float* __restrict__ sp = params;
    const uint64_t s = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    float x[kF], h1[kH];
    normalise(sp, feat + s * kF, x);
    layer1(sp, x, h1);
    float y = sp[kOffB3];
    for (int a = 0; a < kH; ++a) {
        float v = sp[kOffB2 + a];
#pragma unroll
        for (int b = 0; b < kH; ++b) v = fmaf(sp[kOffW2 + a * kH + b], h1[b], v);
        y = fmaf(sp[kOffW3 + a], fmaxf(v, 0.f), y);
    }
    out[s] = y;
}

// Gradient of sum_s 0.5 (y_s - t_s)^2 over the samples idx[0..n) (or 0..n
// when idx is null).  One wave of 64 samples per block: each lane runs its
// sample forward and back to the layer inputs, the activations go to LDS,
// and lane a then sums row a of every weight gradient over the 64 samples
// before one atomic add per weight.
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float* __restrict__ params,
                                                      const float* __restrict__ feat,
                                                      const float* __restrict__ target,
                                                      const uint32_t* __restrict__ idx,
                                                      uint64_t n, float* __restrict__ grad,
                                                      double* __restrict__ loss) {
    const float* __restrict__ sp = params;
    __shared__ float sx[64][kF + 1], sh1[64][kH + 1], sd1[64][kH + 1], sh2[64][kH + 1],
        sd2[64][kH + 1];
    __shared__ float sdy[64];
    const int lane = threadIdx.x;
    const uint64_t q = (uint64_t)blockIdx.x * 64 + lane;
    const bool live = q < n;
    const uint64_t s = live ? (idx ? idx[q] : q) : 0;
    float x[kF], h1[kH];
    normalise(sp, feat + s * kF, x);
    layer1(sp, x, h1);
    float y = sp[kOffB3];
    for (int a = 0; a < kH; ++a) {
        float v = sp[kOffB2 + a];
#pragma unroll
        for (int b = 0; b < kH; ++b) v = fmaf(sp[kOffW2 + a * kH + b], h1[b], v);
        v = fmaxf(v, 0.f);
        sh2[lane][a] = v;
        y = fmaf(sp[kOffW3 + a], v, y);
    }
    const float dy = live ? y - target[s] : 0.f;
    sdy[lane] = dy;
#pragma unroll
    for (int k = 0; k < kF; ++k) sx[lane][k] = x[k];
    // back through layer 2: d2[a] = dy w3[a] [h2[a] > 0]; d1[b] = sum_a d2[a] W2[a][b] [h1[b] > 0]
    float d1[kH];
#pragma unroll
    for (int b = 0; b < kH; ++b) d1[b] = 0.f;
    for (int a = 0; a < kH; ++a) {
        const float d2 = sh2[lane][a] > 0.f ? dy * sp[kOffW3 + a] : 0.f;
        sd2[lane][a] = d2;
#pragma unroll
        for (int b = 0; b < kH; ++b) d1[b] = fmaf(d2, sp[kOffW2 + a * kH + b], d1[b]);
    }
#pragma unroll
    for (int b = 0; b < kH; ++b) {
        sh1[lane][b] = h1[b];
        sd1[lane][b] = h1[b] > 0.f ? d1[b] : 0.f;
    }
    __syncthreads();
    // lane a: row a of dW2, dW1 and entries a of db1, db2, dw3
    const int a = lane;
    float gw2[kH];
#pragma unroll
    for (int b = 0; b < kH; ++b) gw2[b] = 0.f;
    float gw1[kF];
#pragma unroll
    for (int k = 0; k < kF; ++k) gw1[k] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f;
    for (int t = 0; t < 64; ++t) {
        const float d2 = sd2[t][a], d1a = sd1[t][a];
#pragma unroll
        for (int b = 0; b < kH; ++b) gw2[b] = fmaf(d2, sh1[t][b], gw2[b]);
#pragma unroll
        for (int k = 0; k < kF; ++k) gw1[k] = fmaf(d1a, sx[t][k], gw1[k]);
        gb1 += d1a;
        gb2 += d2;
        gw3 = fmaf(sdy[t], sh2[t][a], gw3);
    }
#pragma unroll
    for (int b = 0; b < kH; ++b) atomicAdd(&grad[kOffW2 + a * kH + b], gw2[b]);
#pragma unroll
    for (int k = 0; k < kF; ++k) atomicAdd(&grad[kOffW1 + a * kF + k], gw1[k]);
    atomicAdd(&grad[kOffB1 + a], gb1);
    atomicAdd(&grad[kOffB2 + a], gb2);
    atomicAdd(&grad[kOffW3 + a], gw3);
    if (lane == 0) {
        float gb3 = 0.f;
        double l = 0.0;
        for (int t = 0; t < 64; ++t) {
            gb3 += sdy[t];
            l += 0.5 * (double)sdy[t] * (double)sdy[t];
        }
        atomicAdd(&grad[kOffB3], gb3);
        if (loss) atomicAdd(loss, l);
    }
}

// Adam on the trainable parameters (the normalisation mu, sd stay fixed).
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ params,
                                                   const float* __restrict__ grad,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   float scale, float lr, float b1, float b2,
                                                   float eps, float c1, float c2) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= kOffMu) return;
    const float g = grad[q] * scale;
    const float mq = b1 * m[q] + (1.f - b1) * g;
    const float vq = b2 * v[q] + (1.f - b2) * g * g;
    m[q] = mq;
    v[q] = vq;
    params[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// ---- tile selection -----------------------------------------------------------
// Largest predicted value over the pairs i < j of DOF tile (I, J), R DOFs a side.
__global__ __launch_bounds__(64) void tile_max_kernel(const float* __restrict__ pred, uint32_t D,
                                                      int R, int nblk, float* __restrict__ tmax) {
    const int I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;   // whole block: uniform
    float m = -INFINITY;
    for (int q = threadIdx.x; q < R * R; q += 64) {
        const uint32_t i = I * R + q / R, j = J * R + q % R;
        if (i < j && j < D) m = fmaxf(m, pred[tri(i, j, D)]);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (threadIdx.x == 0) tmax[(size_t)I * nblk + J] = m;
}

// Marks the tiles holding the k best predicted partners of every DOF, so that
// the kept pair graph stays connected wherever the predictions allow.
constexpr int kTopK = 4;
__global__ __launch_bounds__(256) void topk_kernel(const float* __restrict__ pred, uint32_t D,
                                                   int R, int nblk, int k,
                                                   uint8_t* __restrict__ mark) {
    __shared__ float sv[256 * kTopK];
    __shared__ uint32_t sj[256 * kTopK];
    const uint32_t i = blockIdx.x;
    float v[kTopK];
    uint32_t jj[kTopK];
#pragma unroll
    for (int t = 0; t < kTopK; ++t) {
        v[t] = -INFINITY;
        jj[t] = 0xFFFFFFFFu;
    }
    for (uint32_t j = threadIdx.x; j < D; j += 256) {
        if (j == i) continue;
        const float x = i < j ? pred[tri(i, j, D)] : pred[tri(j, i, D)];
        if (x > v[kTopK - 1]) {   // insert into the sorted local list
            float cv = x;
            uint32_t cj = j;
#pragma unroll
            for (int t = 0; t < kTopK; ++t)
                if (cv > v[t]) {
                    const float tv = v[t];
                    const uint32_t tj = jj[t];
                    v[t] = cv;
                    jj[t] = cj;
                    cv = tv;
                    cj = tj;
                }
        }
    }
#pragma unroll
    for (int t = 0; t < kTopK; ++t) {
        sv[threadIdx.x * kTopK + t] = v[t];
        sj[threadIdx.x * kTopK + t] = jj[t];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int r = 0; r < k; ++r) {   // k rounds of arg-max over the candidates
            int best = -1;
            for (int q = 0; q < 256 * kTopK; ++q)
                if (sj[q] != 0xFFFFFFFFu && (best < 0 || sv[q] > sv[best])) best = q;
            if (best < 0) break;
            const uint32_t j = sj[best];
            sj[best] = 0xFFFFFFFFu;
            const uint32_t a = min(i, j), b = max(i, j);
            mark[(size_t)(a / R) * nblk + b / R] = 1;
        }
    }
}

__global__ void fill_f32(float* p, uint64_t n, float v) {
    const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) p[q] = v;
}

}  // namespace

hipError_t launch_subsample(const uint8_t* codes, uint64_t stride, uint32_t D, uint64_t step,
                            uint32_t n_sub, uint8_t* out, uint8_t* coarse, int nb, int nbc,
                            hipStream_t st) {
    hipLaunchKernelGGL(subsample_kernel, dim3((n_sub + 255) / 256, D), dim3(256), 0, st, codes,
                       stride, step, n_sub, out, coarse, nb, nbc);
    return hipGetLastError();
}

hipError_t launch_dof_moments(const uint8_t* sub, uint32_t D, uint32_t n_sub, int nb, float* X,
                              float* Y, hipStream_t st) {
    hipLaunchKernelGGL(dof_moments_kernel, dim3(D), dim3(256), 0, st, sub, n_sub, nb, X, Y);
    return hipGetLastError();
}

hipError_t launch_corr(const float* X, const float* Y, uint32_t D, uint32_t n, const int2* blocks,
                       int n_blocks, float* lin, float* circ, hipStream_t st) {
    if (n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(corr_kernel, dim3(n_blocks), dim3(256), 0, st, X, Y, D, n, blocks, lin, circ);
    return hipGetLastError();
}

int corr_block_size() { return kCB; }

hipError_t launch_features(uint32_t D, const double* s1, const uint8_t* types, const double* mi_sub,
                           const float* lin, const float* circ, float* feat, hipStream_t st) {
    if (D < 2) return hipSuccess;
    hipLaunchKernelGGL(feature_kernel, dim3((D + 255) / 256, D - 1), dim3(256), 0, st, D, s1,
                       types, mi_sub, lin, circ, feat);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const float* params, const float* feat, uint64_t n, float* out,
                              hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       params, feat, n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* params, const float* feat, const float* target,
                           const uint32_t* idx, uint64_t n, float* grad, double* loss,
                           hipStream_t st) {
    hipLaunchKernelGGL(fill_f32, dim3((kScreenParams + 255) / 256), dim3(256), 0, st, grad,
                       (uint64_t)kScreenParams, 0.f);
    if (!n) return hipGetLastError();
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, params,
                       feat, target, idx, n, grad, loss);
    return hipGetLastError();
}

hipError_t launch_tile_select(const float* pred, uint32_t D, int R, int k, float* tmax,
                              uint8_t* mark, hipStream_t st) {
    const int nblk = (int)((D + R - 1) / R);
    hipLaunchKernelGGL(tile_max_kernel, dim3(nblk, nblk), dim3(64), 0, st, pred, D, R, nblk, tmax);
    if (k > 0)
        hipLaunchKernelGGL(topk_kernel, dim3(D), dim3(256), 0, st, pred, D, R, nblk,
                           k < kTopK ? k : kTopK, mark);
    return hipGetLastError();
}

hipError_t launch_adam(float* params, const float* grad, float* m, float* v, float scale,
                       float lr, int t, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3((kScreenParams + 255) / 256), dim3(256), 0, st, params,
                       grad, m, v, scale, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}
