// Maximum spanning forest (MIST) by Boruvka rounds on the device, and the
// segmented FP64 sums behind the MIE totals.
//
// Replaces boruvkaMST (Centre.cpp:5-143).  The reference picks, per
// component, the heaviest cross edge with ties to the lowest edge index
// (strict <, Centre.cpp:43-53), which makes the result the unique maximum
// spanning forest under the total order (weight desc, index asc).  Each
// device round computes the same per-component choice with two atomic
// passes -- atomicMax of an order-preserving 64-bit image of the FP64
// weight, then atomicMin of the edge index among edges that hit that
// maximum -- hooks every component to its choice (a 2-cycle through one
// edge is broken towards the lower label), and relabels by pointer jumping.
// Rounds stop when no cross edge remains, so disconnected graphs (A-MIST
// neighbour lists) give a forest, as in the reference.
#include "centre_internal.h"

namespace {

// -0.0 and +0.0 compare equal in the reference (operator <), so both map to
// the key of +0.0 and the edge index breaks the tie.
__device__ inline unsigned long long order_key(double w) {
    if (w == 0.0) w = 0.0;
    const unsigned long long b = (unsigned long long)__double_as_longlong(w);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ void mst_init_kernel(MstArgs a) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.n_node) a.comp[t] = (uint32_t)t;
    if (t < a.n_edge) a.in_tree[t] = 0;
}

__global__ void mst_reset_kernel(MstArgs a) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.n_node) {
        a.best_w[t] = 0ull;
        a.best_e[t] = 0xFFFFFFFFu;
    }
    if (t == 0) *a.changed = 0;
}

__global__ void mst_max_kernel(MstArgs a) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.n_edge) return;
    const uint32_t cu = a.comp[a.src[e]], cv = a.comp[a.dst[e]];
    if (cu == cv) return;
    const unsigned long long k = order_key(a.w[e]);
    atomicMax(&a.best_w[cu], k);
    atomicMax(&a.best_w[cv], k);
}

__global__ void mst_argmin_kernel(MstArgs a) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.n_edge) return;
    const uint32_t cu = a.comp[a.src[e]], cv = a.comp[a.dst[e]];
    if (cu == cv) return;
    const unsigned long long k = order_key(a.w[e]);
    if (k == a.best_w[cu]) atomicMin(&a.best_e[cu], (unsigned int)e);
    if (k == a.best_w[cv]) atomicMin(&a.best_e[cv], (unsigned int)e);
}

// parent[] is kept in best_w's low word slot reuse: use comp as labels and
// write the hook target into best_w (no longer needed after argmin).
__global__ void mst_hook_kernel(MstArgs a) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n_node) return;
    unsigned long long parent = c;
    if (a.comp[c] == c && a.best_e[c] != 0xFFFFFFFFu) {
        const uint32_t e = a.best_e[c];
        const uint32_t cu = a.comp[a.src[e]], cv = a.comp[a.dst[e]];
        const uint32_t other = (cu == c) ? cv : cu;
        a.in_tree[e] = 1;
        *a.changed = 1;
        // mutual choice through the same edge: the lower label stays root
        if (!(a.best_e[other] == e && c < other)) parent = other;
    }
    // stash the hook in best_w (a root's best_w is not read again this round)
    a.best_w[c] = parent;
}

__global__ void mst_jump_kernel(MstArgs a) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.n_node) return;
    // follow hooks from v's label to the new root
    uint32_t r = a.comp[v];
    for (uint32_t guard = 0; guard < a.n_node; ++guard) {
        const uint32_t p = (uint32_t)a.best_w[r];
        if (p == r) break;
        r = p;
    }
    a.comp[v] = r;
}

__global__ __launch_bounds__(256) void segsum_kernel(const double* x, const uint64_t* beg,
                                                     double* out) {
    __shared__ double red[256];
    const int s = blockIdx.x;
    const uint64_t b = beg[s], e = beg[s + 1];
    double acc = 0.0;
    for (uint64_t i = b + threadIdx.x; i < e; i += 256) acc += x[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = red[0];
}

}  // namespace

// *a.changed |= 1 when an edge endpoint is not below n_node
__global__ void mst_check_kernel(MstArgs a) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < a.n_edge && (a.src[e] >= a.n_node || a.dst[e] >= a.n_node)) atomicOr(a.changed, 1);
}

hipError_t mst_check(const MstArgs& a, hipStream_t st) {
    if (!a.n_edge) return hipSuccess;
    hipLaunchKernelGGL(mst_check_kernel, dim3((unsigned)((a.n_edge + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t mst_init(const MstArgs& a, hipStream_t st) {
    const uint64_t n = a.n_edge > a.n_node ? a.n_edge : a.n_node;
    hipLaunchKernelGGL(mst_init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t run_mst_round(const MstArgs& a, hipStream_t st) {
    const dim3 gv((a.n_node + 255) / 256), ge((unsigned)((a.n_edge + 255) / 256));
    hipLaunchKernelGGL(mst_reset_kernel, gv, dim3(256), 0, st, a);
    if (a.n_edge) {
        hipLaunchKernelGGL(mst_max_kernel, ge, dim3(256), 0, st, a);
        hipLaunchKernelGGL(mst_argmin_kernel, ge, dim3(256), 0, st, a);
    }
    hipLaunchKernelGGL(mst_hook_kernel, gv, dim3(256), 0, st, a);
    hipLaunchKernelGGL(mst_jump_kernel, gv, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_segment_sum(const double* x, const uint64_t* d_beg, int n_seg, double* d_out,
                              hipStream_t st) {
    if (n_seg <= 0) return hipSuccess;
    hipLaunchKernelGGL(segsum_kernel, dim3(n_seg), dim3(256), 0, st, x, d_beg, d_out);
    return hipGetLastError();
}
