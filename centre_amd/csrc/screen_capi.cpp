// C ABI of the learned pair screen (features, MLP inference, gradients and
// training).  No reference counterpart; see screen.hip for the model.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "centre_internal.h"

namespace {

int sfail(centre_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define SCHK(x)                                                                       \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            return sfail(ctx, (int)e_, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Buf {
    void* p = nullptr;
    Buf() = default;
    Buf(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 16); }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace

extern "C" int centre_screen_n_params(void) { return kScreenParams; }

extern "C" int centre_screen_features(centre_ctx* ctx, const uint8_t* d_codes, uint64_t code_stride,
                                      uint32_t n_dof, uint64_t n_frame, const uint8_t* d_types,
                                      const double* d_extrema, int nbins, uint32_t n_sub,
                                      float* d_feat) {
    if (!ctx) return -3;
    if (!d_codes || !d_types || !d_extrema || !d_feat) return sfail(ctx, -3, "null pointer");
    if (n_dof < 2) return sfail(ctx, -3, "need at least two DOFs");
    if (nbins < 1 || nbins > 255) return sfail(ctx, -3, "nbins must be in 1..255");
    if (n_sub < 4 || (n_sub & 3) || n_sub > n_frame)
        return sfail(ctx, -3, "n_sub must be a multiple of 4 in [4, n_frame]");
    if (code_stride < n_frame) return sfail(ctx, -3, "code_stride < n_frame");
    SCHK(hipSetDevice(ctx->device));
    const uint64_t D = n_dof, P = D * (D - 1) / 2;
    const int ml = 0;
    // S1 of every DOF over all frames (one step, ML, with the Jacobian as in
    // the plain path; only its ordering across DOFs matters to the model)
    Buf s1;
    SCHK(s1.alloc(sizeof(double) * D));
    if (int r = centre_dev_dof_entropy(ctx, d_codes, code_stride, n_dof, n_frame, d_types,
                                       d_extrema, nbins, 1, &ml, 1, 1, 0, s1.as<double>(), nullptr))
        return r;
    // evenly strided frame subsample, and its codes merged to nbc bins
    const int nbc = std::min(nbins, kScreenMiBins);
    Buf sub, coarse;
    SCHK(sub.alloc(D * n_sub));
    SCHK(coarse.alloc(D * n_sub));
    SCHK(launch_subsample(d_codes, code_stride, n_dof, n_frame / n_sub, n_sub, sub.as<uint8_t>(),
                          coarse.as<uint8_t>(), nbins, nbc, ctx->stream));
    // MI of the merged codes through the pair kernel (one step, ML, no
    // Jacobian: widths and Jacobians cancel in MI, unit widths keep S2 finite)
    Buf s2, mi, unit;
    SCHK(s2.alloc(sizeof(double) * P));
    SCHK(mi.alloc(sizeof(double) * P));
    SCHK(unit.alloc(sizeof(double) * 4 * D));
    {
        std::vector<double> ext(4 * D, 0.0);
        for (uint64_t d = 0; d < D; ++d) ext[4 * d + 1] = (double)nbc;
        SCHK(hipMemcpyAsync(unit.p, ext.data(), sizeof(double) * ext.size(),
                            hipMemcpyHostToDevice, ctx->stream));
        centre_pair_plan* plan = nullptr;
        if (int r = centre_pair_plan_create(ctx, n_dof, nbc, nullptr, 0, &plan)) return r;
        const int r = centre_dev_pair_entropy(ctx, plan, coarse.as<uint8_t>(), n_sub, n_sub,
                                              d_types, unit.as<double>(), 1, &ml, 1, 0, 0,
                                              nullptr, s2.as<double>(), mi.as<double>(), nullptr);
        centre_pair_plan_destroy(plan);
        if (r) return r;
    }
    // linear and circular correlations on the subsample
    Buf X, Y, lin, circ, blocks;
    SCHK(X.alloc(sizeof(float) * D * n_sub));
    SCHK(Y.alloc(sizeof(float) * D * n_sub));
    SCHK(lin.alloc(sizeof(float) * P));
    SCHK(circ.alloc(sizeof(float) * P));
    SCHK(launch_dof_moments(sub.as<uint8_t>(), n_dof, n_sub, nbins, X.as<float>(), Y.as<float>(),
                            ctx->stream));
    const int cb = corr_block_size();
    const int nblk = (int)((D + cb - 1) / cb);
    std::vector<int2> bl;
    for (int I = 0; I < nblk; ++I)
        for (int J = I; J < nblk; ++J) bl.push_back(make_int2(I, J));
    SCHK(blocks.alloc(sizeof(int2) * bl.size()));
    SCHK(hipMemcpyAsync(blocks.p, bl.data(), sizeof(int2) * bl.size(), hipMemcpyHostToDevice,
                        ctx->stream));
    SCHK(launch_corr(X.as<float>(), Y.as<float>(), n_dof, n_sub, blocks.as<int2>(), (int)bl.size(),
                     lin.as<float>(), circ.as<float>(), ctx->stream));
    SCHK(launch_features(n_dof, s1.as<double>(), d_types, mi.as<double>(), lin.as<float>(),
                         circ.as<float>(), d_feat, ctx->stream));
    SCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int centre_screen_tiles(centre_ctx* ctx, const float* d_pred, uint32_t n_dof,
                                   int nbins, float cutoff, int topk, int32_t* tiles,
                                   uint64_t* n_tiles) {
    if (!ctx) return -3;
    if (!d_pred || !tiles || !n_tiles) return sfail(ctx, -3, "null pointer");
    if (n_dof < 2 || nbins < 1 || nbins > 255) return sfail(ctx, -3, "bad n_dof or nbins");
    if (topk < 0 || topk > 4) return sfail(ctx, -3, "topk must be in 0..4");
    if (!(cutoff > 0.f)) return sfail(ctx, -3, "cutoff must be > 0");
    SCHK(hipSetDevice(ctx->device));
    const int R = dofs_per_tile(nbins);
    const int nblk = (int)((n_dof + R - 1) / R);
    const size_t nt = (size_t)nblk * nblk;
    Buf tmax, mark;
    SCHK(tmax.alloc(sizeof(float) * nt));
    SCHK(mark.alloc(nt));
    SCHK(hipMemsetAsync(mark.p, 0, nt, ctx->stream));
    SCHK(launch_tile_select(d_pred, n_dof, R, topk, tmax.as<float>(), mark.as<uint8_t>(),
                            ctx->stream));
    std::vector<float> hm(nt);
    std::vector<uint8_t> hk(nt);
    SCHK(hipMemcpyAsync(hm.data(), tmax.p, sizeof(float) * nt, hipMemcpyDeviceToHost, ctx->stream));
    SCHK(hipMemcpyAsync(hk.data(), mark.p, nt, hipMemcpyDeviceToHost, ctx->stream));
    SCHK(hipStreamSynchronize(ctx->stream));
    const float lc = logf(cutoff);
    uint64_t n = 0;
    for (int I = 0; I < nblk; ++I)
        for (int J = I; J < nblk; ++J) {
            const size_t t = (size_t)I * nblk + J;
            if (hm[t] >= lc || hk[t]) {
                tiles[2 * n] = I;
                tiles[2 * n + 1] = J;
                ++n;
            }
        }
    *n_tiles = n;
    return 0;
}

extern "C" int centre_mlp_forward(centre_ctx* ctx, const float* d_params, const float* d_feat,
                                  uint64_t n, float* d_out) {
    if (!ctx) return -3;
    if (!n) return 0;
    if (!d_params || !d_feat || !d_out) return sfail(ctx, -3, "null pointer");
    SCHK(hipSetDevice(ctx->device));
    SCHK(launch_mlp_forward(d_params, d_feat, n, d_out, ctx->stream));
    SCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int centre_mlp_loss_grad(centre_ctx* ctx, const float* d_params, const float* d_feat,
                                    const float* d_target, uint64_t n, float* d_grad,
                                    double* loss) {
    if (!ctx) return -3;
    if (!d_params || !d_grad || (n && (!d_feat || !d_target))) return sfail(ctx, -3, "null pointer");
    SCHK(hipSetDevice(ctx->device));
    Buf dl;
    SCHK(dl.alloc(sizeof(double)));
    SCHK(hipMemsetAsync(dl.p, 0, sizeof(double), ctx->stream));
    SCHK(launch_mlp_grad(d_params, d_feat, d_target, nullptr, n, d_grad, dl.as<double>(),
                         ctx->stream));
    double l = 0.0;
    SCHK(hipMemcpyAsync(&l, dl.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SCHK(hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l;
    return 0;
}

extern "C" int centre_mlp_train(centre_ctx* ctx, float* d_params, const float* d_feat,
                                const float* d_target, uint64_t n, int epochs, int batch, float lr,
                                uint64_t seed, double* epoch_loss) {
    if (!ctx) return -3;
    if (!d_params || !d_feat || !d_target) return sfail(ctx, -3, "null pointer");
    if (n == 0 || n > 0xFFFFFFFFull) return sfail(ctx, -3, "n must be in 1..2^32-1");
    if (epochs < 1 || batch < 1 || !(lr > 0.f)) return sfail(ctx, -3, "bad epochs, batch or lr");
    SCHK(hipSetDevice(ctx->device));
    Buf m, v, g, idx, dl;
    SCHK(m.alloc(sizeof(float) * kScreenParams));
    SCHK(v.alloc(sizeof(float) * kScreenParams));
    SCHK(g.alloc(sizeof(float) * kScreenParams));
    SCHK(idx.alloc(sizeof(uint32_t) * n));
    SCHK(dl.alloc(sizeof(double)));
    SCHK(hipMemsetAsync(m.p, 0, sizeof(float) * kScreenParams, ctx->stream));
    SCHK(hipMemsetAsync(v.p, 0, sizeof(float) * kScreenParams, ctx->stream));
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937_64 gen(seed);
    int t = 0;
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), gen);
        SCHK(hipMemcpyAsync(idx.p, perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice,
                            ctx->stream));
        SCHK(hipMemsetAsync(dl.p, 0, sizeof(double), ctx->stream));
        for (uint64_t off = 0; off < n; off += (uint64_t)batch) {
            const uint64_t nb = std::min<uint64_t>((uint64_t)batch, n - off);
            SCHK(launch_mlp_grad(d_params, d_feat, d_target, idx.as<uint32_t>() + off, nb,
                                 g.as<float>(), dl.as<double>(), ctx->stream));
            SCHK(launch_adam(d_params, g.as<float>(), m.as<float>(), v.as<float>(), 1.f / nb, lr,
                             ++t, ctx->stream));
        }
        double l = 0.0;
        SCHK(hipMemcpyAsync(&l, dl.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SCHK(hipStreamSynchronize(ctx->stream));
        if (epoch_loss) epoch_loss[e] = l / (double)n;
    }
    return 0;
}
