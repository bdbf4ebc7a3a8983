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
// neighbour lists) give a forest, as in the reference.  Edges whose weight is
// NaN are left out (a screened pair pass leaves the pairs it skipped NaN).
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

// Largest key over the wave's lanes (inactive lanes pass 0).
__device__ inline unsigned long long wave_max(unsigned long long k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(k, off);
        k = o > k ? o : k;
    }
    return k;
}

// atomicMax of k into table[c] for every active lane; when every lane of the
// wave names the same c (the source side of a row of the pair triangle, or
// late rounds with few components) the wave reduces first and one lane
// writes, instead of 64 atomics on one word.
__device__ inline void max_into(unsigned long long* table, uint32_t c, bool act,
                                unsigned long long k) {
    const uint32_t c0 = __builtin_amdgcn_readfirstlane(c);
    if (__all(c == c0)) {
        const unsigned long long m = wave_max(act ? k : 0ull);
        if ((threadIdx.x & 63) == 0 && m) atomicMax(&table[c0], m);
    } else if (act) {
        atomicMax(&table[c], k);
    }
}

// Graphs whose node table does not fit in LDS: keys straight into best_w.
__global__ void mst_max_kernel(MstArgs a) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = e < a.n_edge;
    uint32_t cu = 0, cv = 0;
    unsigned long long k = 0;
    bool num = false;
    if (ok) {
        cu = a.comp[a.src[e]];
        cv = a.comp[a.dst[e]];
        const double w = a.w[e];
        num = !isnan(w);
        k = order_key(w);
    }
    const bool act = ok && num && cu != cv;
    max_into(a.best_w, cu, act, k);
    max_into(a.best_w, cv, act, k);
}

// Graphs of at most kMstLdsNodes nodes: each workgroup takes a grid-strided
// share of the edges, reduces their keys per component in an LDS table and
// writes the table whole to its row of `part` ([groups][n_node]); the
// column maxima then go to best_w without a single global atomic.
__global__ __launch_bounds__(1024) void mst_max_lds_kernel(MstArgs a, unsigned long long* part) {
    extern __shared__ unsigned long long tab[];
    for (uint32_t c = threadIdx.x; c < a.n_node; c += blockDim.x) tab[c] = 0ull;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // every lane runs the same number of iterations (max_into needs whole waves)
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < a.n_edge; base += stride) {
        const uint64_t e = base + threadIdx.x;
        const bool ok = e < a.n_edge;
        uint32_t cu = 0, cv = 0;
        unsigned long long k = 0;
        bool num = false;
        if (ok) {
            cu = a.comp[a.src[e]];
            cv = a.comp[a.dst[e]];
            const double w = a.w[e];
            num = !isnan(w);
            k = order_key(w);
        }
        const bool act = ok && num && cu != cv;
        max_into(tab, cu, act, k);
        max_into(tab, cv, act, k);
    }
    __syncthreads();
    unsigned long long* row = part + (size_t)blockIdx.x * a.n_node;
    for (uint32_t c = threadIdx.x; c < a.n_node; c += blockDim.x) row[c] = tab[c];
}

__global__ __launch_bounds__(256) void mst_colmax_kernel(MstArgs a, const unsigned long long* part,
                                                         int groups) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n_node) return;
    // four rows per iteration in independent maxima, so their loads overlap
    unsigned long long m[4] = {0ull, 0ull, 0ull, 0ull};
    int g = 0;
    for (; g + 4 <= groups; g += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned long long v = part[(size_t)(g + u) * a.n_node + c];
            m[u] = v > m[u] ? v : m[u];
        }
    }
    for (; g < groups; ++g) {
        const unsigned long long v = part[(size_t)g * a.n_node + c];
        m[0] = v > m[0] ? v : m[0];
    }
    const unsigned long long m01 = m[0] > m[1] ? m[0] : m[1];
    const unsigned long long m23 = m[2] > m[3] ? m[2] : m[3];
    a.best_w[c] = m01 > m23 ? m01 : m23;
}

__global__ void mst_argmin_kernel(MstArgs a) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.n_edge) return;
    const uint32_t cu = a.comp[a.src[e]], cv = a.comp[a.dst[e]];
    const double w = a.w[e];
    if (cu == cv || isnan(w)) return;
    const unsigned long long k = order_key(w);
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

// Tree edges appended to (idx, w) through a counter in *a.changed (their
// order is fixed afterwards by a sort on the host).
__global__ void mst_collect_kernel(MstArgs a, uint32_t* idx, double* w, uint32_t cap) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.n_edge || !a.in_tree[e]) return;
    const uint32_t slot = (uint32_t)atomicAdd(a.changed, 1);
    if (slot < cap) {
        idx[slot] = (uint32_t)e;
        w[slot] = a.w[e];
    }
}

hipError_t mst_collect(const MstArgs& a, uint32_t* idx, double* w, uint32_t cap, hipStream_t st) {
    if (!a.n_edge) return hipSuccess;
    hipLaunchKernelGGL(mst_collect_kernel, dim3((unsigned)((a.n_edge + 255) / 256)), dim3(256), 0,
                       st, a, idx, w, cap);
    return hipGetLastError();
}

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

size_t mst_part_bytes(uint32_t n_node) {
    return n_node <= kMstLdsNodes ? sizeof(unsigned long long) * kMstGroups * (size_t)n_node : 0;
}

hipError_t run_mst_round(const MstArgs& a, hipStream_t st) {
    const dim3 gv((a.n_node + 255) / 256), ge((unsigned)((a.n_edge + 255) / 256));
    hipLaunchKernelGGL(mst_reset_kernel, gv, dim3(256), 0, st, a);
    if (a.n_edge) {
        if (a.part && a.n_node <= kMstLdsNodes) {
            // at least 16 edges per thread, so that small graphs leave few
            // table rows for the column pass
            const uint64_t want = (a.n_edge + 16383) / 16384;
            const int groups = want < (uint64_t)kMstGroups ? (int)want : kMstGroups;
            const size_t lds = sizeof(unsigned long long) * a.n_node;
            if (lds > 65536) {
                const hipError_t e = hipFuncSetAttribute(
                    reinterpret_cast<const void*>(&mst_max_lds_kernel),
                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(mst_max_lds_kernel, dim3(groups), dim3(1024), lds, st, a, a.part);
            hipLaunchKernelGGL(mst_colmax_kernel, gv, dim3(256), 0, st, a,
                               (const unsigned long long*)a.part, groups);
        } else {
            hipLaunchKernelGGL(mst_max_kernel, ge, dim3(256), 0, st, a);
        }
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
