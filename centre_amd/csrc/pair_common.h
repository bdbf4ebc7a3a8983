// Pieces shared by the pair-stage kernels (pair_tiles.hip, pair_tile8.hip):
// the LDS layout of a tile, the log helpers of the entropy epilogue and the
// triangle index of a pair.
#pragma once
#include "centre_internal.h"

#include <cmath>
#include <type_traits>

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

namespace {

constexpr int kOpBytes = kTileRows * kLdsRow;   // one operand, one buffer
constexpr int kGroups = kChunk / 4;             // dword groups per DOF row
constexpr int kSuper = 4;                       // chunks per super-chunk

struct Smem {
    uint8_t* oh;        // [2 buf][2 operand][kTileRows * kLdsRow]
    uint32_t* cl;       // [2 buf][2R DOFs][kGroups] codes staged per slot
    uint8_t* row2dof;   // [256] local row DOF (0xFF = padding)
    uint8_t* col2dof;   // [256]
    unsigned long long* slotT;  // [R*R][nstripe] sum c log c per local pair (fixed point)
    int* slotN;         // [R*R][nstripe] nonzero bins per local pair
    double* logt;       // [256] log(1 + j/256)
    double* invt;       // [256] 1 / (1 + j/256)
    double* ctab;       // [ntab] c log c for small counts (four-wave tile)
    unsigned long long* ctabq;  // [ntab] the same in fixed point (eight-wave tile)
};

constexpr int kTab = 4096;                 // c log c table entries when LDS allows
constexpr size_t kLdsCap = 160 * 1024;

struct LdsLayout {
    size_t cl, map, T, N, log, park, tab, total;
    int nstripe;   // copies of each pair slot, picked by lane % nstripe, so
                   // that the lanes of one wave rarely add to the same word
    int ntab;      // 0 (no table) or kTab
};

__host__ __device__ inline LdsLayout lds_layout(int R, int nwave = 4) {
    LdsLayout L;
    size_t off = (size_t)2 * 2 * kOpBytes;
    L.cl = off;
    off += (size_t)2 * 2 * R * kGroups * 4;
    L.map = off;
    off += 512 + kLdsRow;   // row/col DOF maps, then a trash row
    off = (off + 15) & ~(size_t)15;
    L.nstripe = R <= 8 ? 16 : (R <= 16 ? 4 : 1);
    L.T = off;
    off += sizeof(unsigned long long) * R * R * L.nstripe;
    L.N = off;
    off += sizeof(int) * R * R * L.nstripe;
    off = (off + 15) & ~(size_t)15;
    L.log = off;   // log(1 + j/256), 1/(1 + j/256), then 8 series coefficients
    off += (2 * 256 + 8) * sizeof(double);
    // epilogue parking area of the four-wave tile's fast path (R <= 32):
    // 4 KB per wave; nwave = 0 for the eight-wave tile, which reads its
    // accumulators in place
    L.park = off;
    if (R <= 32) off += (size_t)nwave * 64 * 16 * sizeof(int);
    L.tab = off;
    L.ntab = off + sizeof(double) * kTab <= kLdsCap ? kTab : 0;
    off += sizeof(double) * L.ntab;
    L.total = off;
    return L;
}

// log(c) for an integer count c >= 1: c = 2^(e-1) * m, m in [1, 2),
// m = t_j (1 + r) with t_j = 1 + j/256 from a table, log(1 + r) by its Taylor
// series (|r| < 2^-8, truncation < 2e-18).  About 15 instructions instead of
// the ~100 of the general FP64 log; accurate to a few ulp.
__device__ inline double log_count(int c, const double* logt, const double* invt) {
    int e;
    const double m2 = 2.0 * frexp((double)c, &e);          // [1, 2), exact
    const double f = m2 - 1.0;                              // exact
    const int j = (int)(f * 256.0);
    const double r = (f - j * (1.0 / 256.0)) * invt[j];
    const double p = r * (1.0 + r * (-0.5 + r * (1.0 / 3.0 + r * (-0.25 + r * (0.2 + r * (-1.0 / 6.0))))));
    return (double)(e - 1) * 0.6931471805599453 + logt[j] + p;
}

// The same for any positive finite x, with the series coefficients and log 2
// read from LDS (coef = {1, -1/2, 1/3, -1/4, 1/5, -1/6, log 2}): as
// literals they would be materialised in registers for the whole kernel.
__device__ inline double log_pos_lds(double x, const double* logt, const double* invt,
                                     const double* coef) {
    int e;
    const double m2 = 2.0 * frexp(x, &e);
    const double f = m2 - 1.0;
    const int j = (int)(f * 256.0);
    const double r = (f - j * (1.0 / 256.0)) * invt[j];
    const double p =
        r * (coef[0] + r * (coef[1] + r * (coef[2] + r * (coef[3] + r * (coef[4] + r * coef[5])))));
    return (double)(e - 1) * coef[6] + logt[j] + p;
}

// The per-pair sums of c log c go into the LDS slots as 64-bit integers in
// units of 2^-K, with K from a frame count n so that T = sum c log c <=
// n log n stays below 2^60: the four-wave tile takes n of each step and
// rounds each addend, the eight-wave tile takes the launch's largest step and
// holds its c log c table in that fixed point (its sums are then integer adds
// throughout).  Integer addition is associative, so a slot's total, and S2,
// do not depend on the order in which lanes and waves add to it (with FP64
// atomics the last bits of S2 would vary from run to run).  Each table entry
// or addend is rounded once, by at most 2^-(K+1); at n = 1e5 frames K = 39,
// so the sum of a pair's 2,500 cells is within 2.3e-9 of the exact T in the
// worst case: 2.3e-13 of S2 (about 5e-14 relative) at a 10,000-frame step.
__host__ __device__ inline int fixed_k(double n, double logn) {
    int e;
    (void)frexp(n * logn + 1.0, &e);
    return 60 - e;
}

__device__ inline unsigned long long to_fixed(double x, int k) {
    return (unsigned long long)__double2ll_rn(ldexp(x, k));
}

__device__ inline double from_fixed(unsigned long long v, int k) {
    return ldexp((double)(long long)v, -k);
}

__device__ inline long long tri_index(long long i, long long j, long long D) {
    return i * D - i * (i + 1) / 2 + (j - i - 1);
}

}  // namespace
