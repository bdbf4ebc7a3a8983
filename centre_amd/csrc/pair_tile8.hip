// Eight-wave pair kernel for full tiles (pair_tile8_kernel) and its batch form.
//
// What it replaces: binData2D + entropy2D (Estimators.cpp:130-205, 391-651)
// for every pair of a tile, as pair_tiles.hip does (its header comment
// describes the one-hot formulation and the step-end epilogue).
#include "pair_common.h"

namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// FP4 form of the tile (kF4): one-hot operands as e2m1 nibbles (0x2 = 1.0,
// 0x0 = 0), two frames a byte (the even frame in the low nibble), multiplied
// by v_mfma_scale_f32_32x32x64_f8f6f4 with both block scales 2^0 (E8M0 127):
// a 64-frame chunk is one MFMA a block (the int8 form takes two, each in the
// same issue time), from LDS rows of 32 bytes (kLdsRowF4 apart).  Every
// product is 0 or 1 and the f32 accumulator holds integers, exact up to 2^24,
// so the counts equal the int8 path's while a pair's total frame count is at
// most 2^24 (the launch checks it).  A lane
// holds 32 consecutive k of its row or column (k = 32 * (lane / 32) + t, t in
// its 16 bytes low nibble first); the operands are laid out alike, so only the
// pairing of A's and B's k matters, and it is the same frame on both sides.
// LDS row stride of the FP4 one-hot operands: 32 bytes of nibbles (64
// frames) and 16 of padding, so that the 16 lanes of a 16-byte fragment read
// cover distinct banks; the smaller buffers leave room for a larger c log c
// table
constexpr int kLdsRowF4 = 48;

__device__ __forceinline__ v16f mfma_f4(const v4i& a, const v4i& b, const v16f& c) {
    const v8i a8 = {a[0], a[1], a[2], a[3], 0, 0, 0, 0};
    const v8i b8 = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, c, 4, 4, 0, 127, 0, 127);
}

__device__ __forceinline__ v16i counts_of(const v16i& c) { return c; }
__device__ __forceinline__ v16i counts_of(const v16f& c) {
    v16i r;
#pragma unroll
    for (int k = 0; k < 16; ++k) r[k] = (int)c[k];
    return r;
}

// ---- eight-wave kernel for full tiles --------------------------------------------
// Tiles with R DOFs on both sides (diagonal tiles included: both operands
// are then the same rows, and only pairs i < j are kept) and 4-byte aligned
// step starts, with 8 waves in a 2x4 layout: each wave owns 128 rows x 64
// columns (4x2 MFMA blocks, 128 accumulator registers), so two waves share a
// SIMD and one wave's fragment reads, staging and epilogue issue while the
// other one's MFMAs run.  Only the first ceil(live slots/64) waves stage
// codes (one slot per lane, lanes past the live slots repeat live ones of
// their own wave); the others only multiply.
template <bool kNz>
__device__ __forceinline__ void epi_block8(const v16i& blk, int brow, int gcol, int lane,
                                           const Smem& sm, int nb, int R, int nstr, int ntab,
                                           int kfix) {
    // the lane holds column gcol and rows row0 + 8*(r/4) + r%4, r = 0..15,
    // row0 = brow + 4*(lane/32) in the 32-row block starting at brow; with
    // two waves per SIMD the accumulators are ordinary VGPRs and are read in
    // place (the four-wave tile parks its AGPRs in LDS instead).  The table
    // holds 0 for counts 0 and 1, so every count is one lookup.  Rows and
    // columns past the tile's last DOF are never staged, so their counts are
    // 0 and add nothing.  Nonzero bins are counted only with kNz (the
    // Miller-Madow correction needs them).
    //
    // nb >= 28: a block's 32 rows touch at most two DOFs.  A block whose
    // first and last rows hold one DOF, or one DOF and padding rows, needs no
    // split (wave-uniform); otherwise the next DOF starts at local row bl
    // (wave-uniform) and the lane's sum is split into rows before and from
    // bl (tot - hi and hi).  With kNz the nonzero counts are split the same
    // way, at the lane's own next DOF (kk).
    const int cd = sm.col2dof[gcol];
    const int stripe = lane % nstr;
    const int row0 = brow + 4 * (lane >> 5);
    const int dfirst = sm.row2dof[brow], dlast = sm.row2dof[brow + 31];
    const bool one = dfirst == dlast || dlast == 0xFF;
    const int dlo = sm.row2dof[row0];
    const int kk = (dlo + 1) * nb - row0;                          // kNz split, lane-relative
    const int bl = one ? 32 : __builtin_amdgcn_readfirstlane((dfirst + 1) * nb - brow);
    const int tl = bl - 4 * (lane >> 5);                           // bl, lane-relative
    const int split = kNz ? kk : tl;
    int mx = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) mx = max(mx, blk[r]);
    // sums in the table's fixed point (units of 2^-kfix)
    unsigned long long tot = 0ull, hi = 0ull;
    int ntot = 0, nhi = 0;
    if (__builtin_amdgcn_ballot_w64(mx >= ntab) == 0) {
        // every count of the block is in the table (wave-uniform); the 16
        // lookups are issued together and summed as a tree, so that their
        // latencies overlap
        if (one) {
            unsigned long long v[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = sm.ctabq[blk[r]];
#pragma unroll
            for (int r = 0; r < 8; ++r) v[r] = v[2 * r] + v[2 * r + 1];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = v[2 * r] + v[2 * r + 1];
            tot = (v[0] + v[1]) + (v[2] + v[3]);
            if constexpr (kNz) {
#pragma unroll
                for (int r = 0; r < 16; ++r) ntot += min(blk[r], 1);
            }
        } else if constexpr (kNz) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const bool h = 8 * (r >> 2) + (r & 3) >= kk;
                const unsigned long long v = sm.ctabq[blk[r]];
                tot += v;
                hi += h ? v : 0ull;
                ntot += min(blk[r], 1);
                nhi += h ? min(blk[r], 1) : 0;
            }
        } else {
            // the lane's rows come in four groups of four (q): local rows
            // 8q..8q+3 in the first half-wave, 8q+4..8q+7 in the second, so
            // only the group holding bl is split, and per lane only there:
            // it adds the suffix of its four values from the first row >= bl
            unsigned long long v[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = sm.ctabq[blk[r]];
            unsigned long long g[4];
            hi = 0ull;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const unsigned long long s3 = v[4 * q + 3];
                const unsigned long long s2 = s3 + v[4 * q + 2];
                const unsigned long long s1 = s2 + v[4 * q + 1];
                g[q] = s1 + v[4 * q];
                if (bl <= 8 * q) {
                    hi += g[q];
                } else if (bl < 8 * q + 8) {
                    const int t = tl - 8 * q;   // rows of the group before bl
                    unsigned long long part = t == 3 ? s3 : 0ull;
                    part = t == 2 ? s2 : part;
                    part = t == 1 ? s1 : part;
                    part = t <= 0 ? g[q] : part;
                    hi += part;
                }
            }
            tot = (g[0] + g[1]) + (g[2] + g[3]);
        }
    } else {
        // a count past the table: c log c from the log tables
#pragma unroll 1
        for (int r = 0; r < 16; ++r) {
            const int c = blk[r];
            const bool h = 8 * (r >> 2) + (r & 3) >= split;
            const unsigned long long v =
                c < ntab ? sm.ctabq[c]
                         : to_fixed((double)c * log_pos_lds((double)c, sm.logt, sm.invt, sm.invt + 256), kfix);
            tot += v;
            hi += h ? v : 0ull;
            ntot += min(c, 1);
            nhi += h ? min(c, 1) : 0;
        }
    }
    if constexpr (kNz) {
        if (cd == 0xFF || dlo == 0xFF) return;
        if (ntot > nhi) {
            atomicAdd(&sm.slotT[(dlo * R + cd) * nstr + stripe], tot - hi);
            atomicAdd(&sm.slotN[(dlo * R + cd) * nstr + stripe], ntot - nhi);
        }
        const int dhi = sm.row2dof[row0 + 27];
        if (nhi && dhi != 0xFF && dhi != dlo) {
            atomicAdd(&sm.slotT[(dhi * R + cd) * nstr + stripe], hi);
            atomicAdd(&sm.slotN[(dhi * R + cd) * nstr + stripe], nhi);
        }
    } else {
        if (cd == 0xFF || dfirst == 0xFF) return;
        atomicAdd(&sm.slotT[(dfirst * R + cd) * nstr + stripe], tot - hi);
        if (!one) atomicAdd(&sm.slotT[((dfirst + 1) * R + cd) * nstr + stripe], hi);
    }
}

// Epilogue block of the small-bin tile (nb <= 8): every DOF has 8 one-hot
// rows (bins past nb stay zero), so a 32-row block holds four row DOFs and
// the lane's four row groups q (rows 8q..8q+3 in the first half-wave,
// 8q+4..8q+7 in the second) each lie in one of them; the lane's column is
// bin gcol % 8 of column DOF gcol / 8.  The eight lanes of a column DOF sum
// their group totals by DPP row shifts, and their last lane adds the sum to
// the pair's slot (one stripe per half-wave).
template <int kCtl>
__device__ __forceinline__ double dpp_d(double v) {
    // a DPP move of both halves of a double; lanes whose source lies outside
    // their 16-lane row read 0
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, kCtl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), kCtl, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// lane l ends with the sum over lanes l-7..l of its 16-lane row (row_shr 1,
// 2, 4), so lane 8k+7 holds the sum of lanes 8k..8k+7
__device__ __forceinline__ double dpp_sum8(double v) {
    v += dpp_d<0x111>(v);
    v += dpp_d<0x112>(v);
    v += dpp_d<0x114>(v);
    return v;
}

template <int kCtl>
__device__ __forceinline__ unsigned long long dpp_q(unsigned long long u) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, kCtl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), kCtl, 0xF, 0xF, true);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// dpp_sum8 on 64-bit integers
__device__ __forceinline__ unsigned long long dpp_sum8_q(unsigned long long v) {
    v += dpp_q<0x111>(v);
    v += dpp_q<0x112>(v);
    v += dpp_q<0x114>(v);
    return v;
}

__device__ __forceinline__ int dpp_sum8_i(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true);
    return v;
}

template <bool kNz>
__device__ __forceinline__ void epi_block_small(const v16i& blk, int brow, int gcol, int lane,
                                                const Smem& sm, int R, int nstr, int ntab,
                                                int kfix) {
    const double* const coef = sm.invt + 256;
    int mx = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) mx = max(mx, blk[r]);
    const bool big = __builtin_amdgcn_ballot_w64(mx >= ntab) != 0;   // wave-uniform
    const int cd = gcol >> 3;
    const int stripe = lane >> 5;
    const bool last = (lane & 7) == 7;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        unsigned long long v[4];
        if (!big) {
#pragma unroll
            for (int m = 0; m < 4; ++m) v[m] = sm.ctabq[blk[4 * q + m]];
        } else {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int c = blk[4 * q + m];
                v[m] = c < ntab ? sm.ctabq[c]
                                : to_fixed((double)c * log_pos_lds((double)c, sm.logt, sm.invt, coef), kfix);
            }
        }
        const unsigned long long g = dpp_sum8_q((v[0] + v[1]) + (v[2] + v[3]));
        const int rd = (brow >> 3) + q;
        if (last) atomicAdd(&sm.slotT[(rd * R + cd) * nstr + stripe], g);
        if constexpr (kNz) {
            int nz = 0;
#pragma unroll
            for (int m = 0; m < 4; ++m) nz += min(blk[4 * q + m], 1);
            nz = dpp_sum8_i(nz);
            if (last) atomicAdd(&sm.slotN[(rd * R + cd) * nstr + stripe], nz);
        }
    }
}

// LDS of the eight-wave tile: kBuf8 one-hot operand buffers (chunk g is
// multiplied from one, chunk g+1's fragments are read from the next while
// that happens, chunk g+2 is staged into the third), the DOF maps and a
// trash row, the per-pair slots, the log tables and as large a c log c table
// as the rest leaves room for.
constexpr int kBuf8 = 3;
constexpr int kTabSmall = 16384;   // c log c table cap of the small-bin tile

struct Lay8 {
    size_t map, T, N, log, dofc, tmap, tab, total;
    int nstripe, ntab;
};

// nz: nonzero-bin slots (Miller-Madow) are kept; without them the c log c
// table takes their room.  The small-bin tile's map area also holds the
// 64-bit mask of the 32x32 blocks that contain a pair to evaluate.
__host__ __device__ inline Lay8 lay8(int R, bool small, bool nz, bool f4) {
    Lay8 L;
    size_t off = (size_t)kBuf8 * 2 * kTileRows * (f4 ? kLdsRowF4 : kLdsRow);
    L.map = off;
    off += 512 + kLdsRow + 16;
    off = (off + 15) & ~(size_t)15;
    L.nstripe = small ? 2 : 16;
    L.T = off;
    off += sizeof(unsigned long long) * R * R * L.nstripe;
    L.N = off;
    if (nz) off += sizeof(int) * R * R * L.nstripe;
    off = (off + 15) & ~(size_t)15;
    L.log = off;   // log(1 + j/256), 1/(1 + j/256), then 8 series coefficients
    off += (2 * 256 + 8) * sizeof(double);
    // small-bin tile: its finalize inputs (the big-bin tile's finalize reads
    // them from global memory; LDS addresses held for it cost registers)
    L.dofc = off;   // the tile's 2R DOFs: bin width, then L and S1 of the step
    if (small) off += 3 * 64 * sizeof(double);
    L.tmap = off;   // the tile's canonical pair indices (explicit plans)
    if (small) off += sizeof(int) * R * R;
    off = (off + 15) & ~(size_t)15;
    L.tab = off;
    // 16 bytes stay free past the layout: the persistent batch kernel's
    // tile counter word
    const size_t room = off + 16 < kLdsCap ? (kLdsCap - off - 16) / sizeof(double) : 0;
    // the small-bin tile's counts are large (few bins), so its table takes
    // all the room there is
    const size_t cap = small ? (size_t)kTabSmall : (size_t)kTab;
    L.ntab = (int)((room < cap ? room : cap) / 256 * 256);
    off += sizeof(double) * L.ntab;
    L.total = off;
    return L;
}

// Tile-independent LDS contents of the eight-wave tile: the row/column DOF
// maps, log(1 + j/256) and 1/(1 + j/256), the log-series coefficients and
// the c log c table.  The persistent batch kernel writes them once before
// its tile loop (computed inside the loop they would be hoisted out of it
// and held in registers across every tile).
template <bool kNz, bool kSmall, bool kF4>
__device__ __forceinline__ void tile8_tables(int R, int nb, int kfix) {
    extern __shared__ __align__(16) uint8_t smem_raw[];
    const Lay8 lay = lay8(R, kSmall, kNz, kF4);
    uint8_t* const row2dof = smem_raw + lay.map;
    uint8_t* const col2dof = smem_raw + lay.map + 256;
    double* const logt = reinterpret_cast<double*>(smem_raw + lay.log);
    double* const invt = logt + 256;
    unsigned long long* const ctab = reinterpret_cast<unsigned long long*>(smem_raw + lay.tab);
    const int tid = threadIdx.x;
    if (tid < 256) {
        const int d = tid / (kSmall ? 8 : nb);
        row2dof[tid] = (uint8_t)(d < R ? d : 0xFF);
        col2dof[tid] = (uint8_t)(d < R ? d : 0xFF);
        const double t = 1.0 + tid * (1.0 / 256.0);
        logt[tid] = log(t);
        invt[tid] = 1.0 / t;
        if (tid < 8) {
            const double cf[8] = {1.0, -0.5, 1.0 / 3.0, -0.25, 0.2, -1.0 / 6.0,
                                  0.6931471805599453, 0.0};
            invt[256 + tid] = cf[tid];
        }
    }
    // c log c by the same log_pos_lds as the epilogue's counts past the
    // table, so that a count's term does not depend on the table's size
    __syncthreads();
    for (int c = tid; c < lay.ntab; c += 512)
        ctab[c] = c > 1 ? to_fixed((double)c * log_pos_lds((double)c, logt, invt, invt + 256), kfix) : 0ull;
}

// kTables: also write the tile-independent LDS contents (tile8_tables);
// false when the caller wrote them already.  kF4: the FP4 form (mfma_f4).
template <bool kBatch, bool kNz, bool kSmall, bool kF4, bool kTables = true>
__device__ __forceinline__ void tile8_body(const PairKernelArgs& a_in, int tix) {
    extern __shared__ __align__(16) uint8_t smem_raw[];
    const PairKernelArgs& a = a_in;
    // a batch entry takes codes, sizes and outputs from its system; the
    // output pointers are read at each step end rather than kept live
    const BatchSystem* const sdp = kBatch ? &a.systems[a.tile_sys[tix]] : nullptr;
    const uint8_t* const codes = kBatch ? sdp->codes : a.codes;
    const uint64_t cstride = kBatch ? sdp->stride : a.stride;
    const uint64_t ss = kBatch ? sdp->ss : a.ss;
    const uint64_t sstride = kBatch ? sdp->sstride : a.sstride;
    const int R = a.R, nb = a.nb, S = a.nsteps;
    const Lay8 lay = lay8(R, kSmall, kNz, kF4);
    const int nstr = lay.nstripe, ntab = lay.ntab;
    // one-hot row stride and operand size of this form
    constexpr int kRow = kF4 ? kLdsRowF4 : kLdsRow;
    constexpr int kOpB = kTileRows * kRow;
    Smem sm;
    sm.oh = smem_raw;
    sm.cl = nullptr;
    sm.row2dof = smem_raw + lay.map;
    sm.col2dof = smem_raw + lay.map + 256;
    sm.slotT = reinterpret_cast<unsigned long long*>(smem_raw + lay.T);
    sm.slotN = reinterpret_cast<int*>(smem_raw + lay.N);
    sm.logt = reinterpret_cast<double*>(smem_raw + lay.log);
    sm.invt = sm.logt + 256;
    sm.ctab = nullptr;
    sm.ctabq = reinterpret_cast<unsigned long long*>(smem_raw + lay.tab);
    uint8_t* const trash = smem_raw + lay.map + 512;

    int tid = threadIdx.x;
    // in the persistent loop, values derived from the thread index alone
    // would be hoisted out of the loop and held across tiles: hide it
    if constexpr (!kTables) asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave >> 2, wc = wave & 3;
    const int2 tile = a.tiles[tix];
    const int r0 = tile.x * R, c0 = tile.y * R;
    const bool diag = tile.x == tile.y;
    const size_t boff = diag ? 0 : kOpB;   // column operand
    const long long D = kBatch ? sdp->n_dof : a.n_dof;

    for (int i = tid; i < kBuf8 * 2 * kOpB / 16; i += 512)
        reinterpret_cast<uint4*>(sm.oh)[i] = make_uint4(0, 0, 0, 0);
    if constexpr (kTables) tile8_tables<kNz, kSmall, kF4>(R, nb, a.kfix);
    for (int i = tid; i < R * R * nstr; i += 512) {
        sm.slotT[i] = 0ull;
        if constexpr (kNz) sm.slotN[i] = 0;
    }
    // small-bin tile, per-tile finalize inputs: bin widths of the tile's
    // DOFs and, for an explicit plan, its canonical pair indices (rows past
    // the system's last DOF read the last one; their pairs are never written)
    double* const dcw = reinterpret_cast<double*>(smem_raw + lay.dofc);
    double* const dcL = dcw + 64;
    double* const dcS = dcw + 128;
    int* const tmap = reinterpret_cast<int*>(smem_raw + lay.tmap);
    if constexpr (kSmall) {
        const double* const dw = kBatch ? sdp->dof_w : a.dof_w;
        for (int t = tid; t < 2 * R; t += 512) {
            const long long g = t < R ? r0 + t : c0 + (t - R);
            dcw[t] = dw[g < D ? g : D - 1];
        }
        if (a.tilemap)
            for (int pq = tid; pq < R * R; pq += 512) tmap[pq] = a.tilemap[(size_t)tix * R * R + pq];
    }
    if (tid == 0) *reinterpret_cast<int*>(trash) = *reinterpret_cast<volatile int*>(a.flag);
    // small-bin tile: bit 8*rb + cb marks a 32x32 MFMA block (row DOFs
    // 4rb..4rb+3, column DOFs 4cb..4cb+3) holding a pair to evaluate; the
    // epilogue skips the others (a diagonal tile's lower half, the pairs
    // an explicit list leaves out)
    unsigned long long* const blkmask = reinterpret_cast<unsigned long long*>(trash + kLdsRow);
    if (kSmall && tid == 0) *blkmask = 0ull;
    __syncthreads();
    // codes were range-checked by the marginal pass; with the flag set the
    // results are discarded
    if (*reinterpret_cast<const int*>(trash) != 0) return;
    if constexpr (kSmall) {
        for (int pq = tid; pq < R * R; pq += 512) {
            const int rd = pq / R, cd = pq - rd * R;
            const long long gi = (long long)tile.x * R + rd, gj = (long long)tile.y * R + cd;
            const bool want = gi < gj && gj < D &&
                              (a.tilemap == nullptr || a.tilemap[(size_t)tix * R * R + pq] >= 0);
            if (want) atomicOr(blkmask, 1ull << (8 * (rd >> 2) + (cd >> 2)));
        }
    }

    constexpr int kCh = kChunk;   // frames a chunk
    const int nch = (int)((ss + kCh - 1) / kCh);
    const int G = S * nch;
    const int nlive = (diag ? 1 : 2) * R * kGroups;
    // this lane's slot: DOF ld, frames kfr..kfr+3 of each chunk.  Waves run on
    // SIMD wave % 4.  The stagers are waves 1, 2, 3 (and 0 when nlive > 192),
    // one slot per lane (lanes past nlive repeat live slots); SIMD 0 keeps
    // two multiply-only waves, and one of them, wave 0 (else wave 4),
    // finalizes each step's sums.  Against stagers 0, 1, 2 with wave 7
    // finalizing: 24.9 against 25.8 ms on the headline tiles; two stagers on
    // one SIMD (waves 0, 2, 4) took 32.1 ms, and spreading the slots over all
    // eight waves (nlive/8 lanes each) ran 20% slower.
    const int swave = kSmall ? wave : (wave < 4 ? (wave + 3) & 3 : wave);
    const bool stager = kSmall ? wave * 64 < (diag ? 1 : 2) * R * 4 : swave * 64 < nlive;
    const int q = (swave * 64 + lane) % nlive;
    const int fwave = nlive > 192 ? 4 : 0;   // finalizer, never a stager
    // kfr: the slot's first frame in the chunk; kb: its byte offset in its
    // LDS rows (the FP4 form holds the 4 frames in 2 bytes)
    const int ld = q / kGroups, kfr = (q - ld * kGroups) * 4;
    const int kb = kF4 ? kfr / 2 : kfr;
    const bool isA = ld < R;
    // rows past the system's last DOF (a partial last block) read the last
    // row; their pairs are never written
    const long long grow = isA ? r0 + ld : c0 + (ld - R);
    const uint8_t* const src = codes + (uint64_t)(grow < D ? grow : D - 1) * cstride + kfr;
    const uint32_t dst = (uint32_t)((isA ? 0 : kOpB) + (isA ? ld : ld - R) * nb * kRow + kb);
    const uint32_t trash_off = (uint32_t)(lay.map + 512 + 16 + kb);
    // what this slot set in the buffers staged one, two and three chunks
    // before the next (the buffer staged next is the one staged three ago)
    uint32_t set3 = 0u, set2 = 0u, set1 = 0u;
    // Small-bin tile: this lane's item is DOF sd of the two operands and
    // frames g16..g16+15 of each chunk, one row segment per bin (16 bytes;
    // 8 in the FP4 form, at byte gb);
    // the first waves stage (256 items, 128 on a diagonal tile).  Rows past
    // the system's last DOF read row 0 and stage no frames.
    const int nlive_s = (diag ? 1 : 2) * R * 4;
    const int qs = (wave * 64 + lane) % nlive_s;
    const int sd = qs >> 2, g16 = (qs & 3) * 16, gb = kF4 ? g16 / 2 : g16;
    const bool sA = sd < R;
    const long long srow = sA ? r0 + sd : c0 + (sd - R);
    const bool srow_ok = srow < D;
    const uint8_t* const src_s = codes + (uint64_t)(srow_ok ? srow : 0) * cstride + g16;
    const uint32_t dst_s = (uint32_t)((sA ? 0 : kOpB) + (sA ? sd : sd - R) * 8 * kRow + gb);
    const bool stager_s = wave * 64 < nlive_s;   // wave-uniform
    int ls = 0, lc = 0;   // next chunk to load
    auto advance = [](int& s, int& c, int n) __attribute__((always_inline)) {
        if (++c == n) { c = 0; ++s; }
    };
    const int nv_last = (int)(ss - (uint64_t)(nch - 1) * kCh);   // a step's last chunk
    auto nvalid_of = [&](int s, int c) __attribute__((always_inline)) -> int {
        return s < S ? (c + 1 < nch ? kCh : nv_last) : 0;
    };
    // this slot's code dword of the next chunk to load (past the last chunk
    // a dword every row has, unused)
    auto load1 = [&]() __attribute__((always_inline)) -> uint32_t {
        const uint64_t f0 = (uint64_t)ls * sstride + (uint64_t)lc * kCh;
        const int nv = nvalid_of(ls, lc);
        uint32_t v;
        if (nv >= kCh)   // wave-uniform: a full chunk, every lane's dword is in the step
            v = *reinterpret_cast<const uint32_t*>(src + f0);
        else
            v = *reinterpret_cast<const uint32_t*>(kfr < nv ? src + f0 : codes);
        advance(ls, lc, nch);
        return v;
    };
    // small-bin tile: the 16 code bytes of the next chunk to load (past the
    // last chunk or the step's valid frames, bytes of the first row, unused)
    auto load4 = [&]() __attribute__((always_inline)) -> uint4 {
        const uint64_t f0 = (uint64_t)ls * sstride + (uint64_t)lc * kCh;
        const int nv = nvalid_of(ls, lc);
        uint32_t v[4];
        if (nv >= kCh) {   // wave-uniform: every dword is in the step
#pragma unroll
            for (int d = 0; d < 4; ++d) v[d] = *reinterpret_cast<const uint32_t*>(src_s + f0 + 4 * d);
        } else {
#pragma unroll
            for (int d = 0; d < 4; ++d)
                v[d] = *reinterpret_cast<const uint32_t*>(g16 + 4 * d < nv ? src_s + f0 + 4 * d : codes);
        }
        advance(ls, lc, nch);
        return make_uint4(v[0], v[1], v[2], v[3]);
    };
    // small-bin tile: chunk with nv valid frames (codes w) into buffer b.
    // Each bin's 16 one-hot bytes are one byte permute per code dword: the
    // permute selects byte `code` of an 8-byte table holding 1 at the bin,
    // and selector 12 (an invalid frame) reads 0.  Rows of bins past nb stay
    // zero, every other row is overwritten whole, so nothing is cleared.
    //
    // FP4 form: the item's 16 frames in 8 bytes a bin, two a byte.  Its even
    // and its odd frames are first gathered into selector dwords (frames 8m,
    // 8m+2, 8m+4, 8m+6 and 8m+1, ..., 8m+7 of code dwords 2m, 2m+1); a bin's
    // byte j of dword m is then the permute of the even selector through a
    // table holding 0x02 at the bin, or-ed with that of the odd selector
    // through one holding 0x20: frames 8m+2j (low nibble) and 8m+2j+1.
    auto stage4 = [&](int b, int nv, uint4 w) __attribute__((always_inline)) {
        const int nvk = (srow_ok ? nv : 0) - g16;
        uint32_t x[4] = {w.x, w.y, w.z, w.w};
        if (nvk < 16) {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int n = nvk - 4 * d;
                const uint32_t keep = n >= 4 ? 0xFFFFFFFFu : (n <= 0 ? 0u : 0xFFFFFFFFu >> (8 * (4 - n)));
                x[d] = (x[d] & keep) | (0x0C0C0C0Cu & ~keep);
            }
        }
        const uint32_t base = (uint32_t)(b * 2 * kOpB) + dst_s;
        if constexpr (kF4) {
            uint32_t ev[2], od[2];
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                ev[m] = __builtin_amdgcn_perm(x[2 * m + 1], x[2 * m], 0x06040200u);
                od[m] = __builtin_amdgcn_perm(x[2 * m + 1], x[2 * m], 0x07050301u);
            }
#pragma unroll
            for (int bin = 0; bin < 8; ++bin) {
                if (bin >= nb) break;   // uniform
                const uint32_t lo = bin < 4 ? 2u << (8 * bin) : 0u;
                const uint32_t hi = bin < 4 ? 0u : 2u << (8 * (bin - 4));
                uint32_t o[2];
#pragma unroll
                for (int m = 0; m < 2; ++m)
                    o[m] = __builtin_amdgcn_perm(hi, lo, ev[m]) | __builtin_amdgcn_perm(hi << 4, lo << 4, od[m]);
                *reinterpret_cast<uint2*>(smem_raw + base + bin * kRow) = make_uint2(o[0], o[1]);
            }
        } else {
#pragma unroll
            for (int bin = 0; bin < 8; ++bin) {
                if (bin >= nb) break;   // uniform
                const uint32_t lo = bin < 4 ? 1u << (8 * bin) : 0u;
                const uint32_t hi = bin < 4 ? 0u : 1u << (8 * (bin - 4));
                uint4 o;
                o.x = __builtin_amdgcn_perm(hi, lo, x[0]);
                o.y = __builtin_amdgcn_perm(hi, lo, x[1]);
                o.z = __builtin_amdgcn_perm(hi, lo, x[2]);
                o.w = __builtin_amdgcn_perm(hi, lo, x[3]);
                *reinterpret_cast<uint4*>(smem_raw + base + bin * kRow) = o;
            }
        }
    };
    // chunk with nv valid frames (codes w) into buffer b: clear what this
    // slot set when buffer b was last staged, three chunks ago (old), and set
    // the new bytes; returns what it set.  Every chunk but a step's last is
    // full and takes the short path; in a partial chunk the frames past nv
    // go to the trash row.  Offsets stay integers so the stores compile to
    // ds_write.
    auto stage = [&](int b, int nv, uint32_t w, uint32_t old) __attribute__((always_inline)) -> uint32_t {
        const uint32_t rowoff = (uint32_t)(b * 2 * kOpB) + dst;
#pragma unroll
        for (int e = 0; e < 4; ++e) smem_raw[rowoff + ((old >> (8 * e)) & 0xFFu) * kRow + e] = 0;
        if (nv >= kChunk) {
#pragma unroll
            for (int e = 0; e < 4; ++e) smem_raw[rowoff + ((w >> (8 * e)) & 0xFFu) * kRow + e] = 1;
            return w;
        }
        const int nvk = nv - kfr;
        const uint32_t vmask = nvk >= 4 ? 0xFFFFFFFFu : (nvk <= 0 ? 0u : 0xFFFFFFFFu >> (8 * (4 - nvk)));
        const uint32_t wv = w & vmask;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t o = nvk > e ? rowoff : trash_off;
            smem_raw[o + ((wv >> (8 * e)) & 0xFFu) * kRow + e] = 1;
        }
        return wv;
    };
    // FP4 form of stage: the slot's 4 frames in 2 bytes of each row, frame
    // 2e in the low and 2e+1 in the high nibble of byte e.  Clearing writes 0
    // to byte e of both old rows; setting writes 0x02 to row c0 and 0x20 to
    // row c1, or 0x22 to both (one byte) when the two codes are equal.  The
    // slot owns these bytes of every row, so whole-byte writes lose nothing.
    auto stage_f4 = [&](int b, int nv, uint32_t w, uint32_t old) __attribute__((always_inline)) -> uint32_t {
        const uint32_t rowoff = (uint32_t)(b * 2 * kOpB) + dst;
#pragma unroll
        for (int i = 0; i < 4; ++i) smem_raw[rowoff + ((old >> (8 * i)) & 0xFFu) * kRow + (i >> 1)] = 0;
        if (nv >= kCh) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const uint32_t c0 = (w >> (16 * e)) & 0xFFu, c1 = (w >> (16 * e + 8)) & 0xFFu;
                const bool same = c0 == c1;
                smem_raw[rowoff + c0 * kRow + e] = same ? 0x22 : 0x02;
                smem_raw[rowoff + c1 * kRow + e] = same ? 0x22 : 0x20;
            }
            return w;
        }
        // a partial chunk: frames past nv go to the trash row
        const int nvk = nv - kfr;
        const uint32_t vmask = nvk >= 4 ? 0xFFFFFFFFu : (nvk <= 0 ? 0u : 0xFFFFFFFFu >> (8 * (4 - nvk)));
        const uint32_t wv = w & vmask;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const uint32_t c0 = (wv >> (16 * e)) & 0xFFu, c1 = (wv >> (16 * e + 8)) & 0xFFu;
            const bool v0 = nvk > 2 * e, v1 = nvk > 2 * e + 1;   // v1 implies v0
            const bool same = c0 == c1;
            smem_raw[(v0 ? rowoff : trash_off) + c0 * kRow + e] = v1 && same ? 0x22 : 0x02;
            smem_raw[(v1 ? rowoff : trash_off) + c1 * kRow + e] = v0 && same ? 0x22 : 0x20;
        }
        return wv;
    };
    using CodeT = std::conditional_t<kSmall, uint4, uint32_t>;
    auto loadc = [&]() __attribute__((always_inline)) -> CodeT {
        if constexpr (kSmall) return load4();
        else return load1();
    };
    // stages a chunk; returns what the byte scatter set (0 for the small-bin tile)
    auto stagec = [&](int b, int nv, CodeT w, uint32_t old) __attribute__((always_inline)) -> uint32_t {
        if constexpr (kSmall) {
            stage4(b, nv, w);
            return 0u;
        } else if constexpr (kF4) {
            return stage_f4(b, nv, w, old);
        } else {
            return stage(b, nv, w, old);
        }
    };
    using AccT = std::conditional_t<kF4, v16f, v16i>;
    AccT acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = AccT{};
    // fragment i of the row operand / j of the column operand, 32 frames kk.
    // The wave's 4 row blocks and 2 column blocks of 32: contiguous (rows
    // 128wr.., columns 64wc..), or on the small-bin tile interleaved (rows
    // 32(wr + 2i), columns 32(wc + 4j)), so that the blocks whose epilogue
    // is skipped (a triangle of the tile) are spread over the waves.
    const int lrow = (lane & 31) * kRow + (lane >> 5) * 16;
    auto rowblk = [&](int i) __attribute__((always_inline)) {
        return kSmall ? 32 * (wr + 2 * i) : 128 * wr + 32 * i;
    };
    auto colblk = [&](int j) __attribute__((always_inline)) {
        return kSmall ? 32 * (wc + 4 * j) : 64 * wc + 32 * j;
    };
    // buffer offsets stay integers and every address is formed from the
    // __shared__ array itself, so the reads compile to ds_read (a generic
    // pointer carried through the loop compiles to flat loads)
    auto rdA = [&](uint32_t bo, int kk, int i) __attribute__((always_inline)) -> v4i {
        return *reinterpret_cast<const v4i*>(smem_raw + bo + rowblk(i) * kRow + lrow +
                                             kk * 32);
    };
    auto rdB = [&](uint32_t bo, int kk, int j) __attribute__((always_inline)) -> v4i {
        return *reinterpret_cast<const v4i*>(smem_raw + bo + boff + colblk(j) * kRow +
                                             lrow + kk * 32);
    };
    v4i fa0[4], fb0[2], fa1[4], fb1[2];
    auto read_all = [&](uint32_t base) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < 4; ++i) fa0[i] = rdA(base, 0, i);
#pragma unroll
        for (int j = 0; j < 2; ++j) fb0[j] = rdB(base, 0, j);
        if constexpr (!kF4) {
#pragma unroll
            for (int i = 0; i < 4; ++i) fa1[i] = rdA(base, 1, i);
#pragma unroll
            for (int j = 0; j < 2; ++j) fb1[j] = rdB(base, 1, j);
        }
    };
    // The MFMAs of 32 frames (kk; the FP4 form's one kk is all 64); with kPref each fragment register is
    // reloaded with chunk g+1's fragment (buffer nbase) as soon as its last
    // MFMA of chunk g has issued, so the next chunk's LDS reads overlap this
    // chunk's MFMAs and need no second register set.
    auto mfma_kk = [&](auto pref, v4i (&af)[4], v4i (&bf)[2], uint32_t nbase,
                       int kk) __attribute__((always_inline)) {
        constexpr bool kPref = decltype(pref)::value;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if constexpr (kF4) {
                acc[i][0] = mfma_f4(af[i], bf[0], acc[i][0]);
                acc[i][1] = mfma_f4(af[i], bf[1], acc[i][1]);
            } else {
                acc[i][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[0], acc[i][0], 0, 0, 0);
                acc[i][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[1], acc[i][1], 0, 0, 0);
            }
            if constexpr (kPref) {
                __builtin_amdgcn_sched_barrier(0);
                af[i] = rdA(nbase, kk, i);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if constexpr (kPref) {
            bf[0] = rdB(nbase, kk, 0);
            bf[1] = rdB(nbase, kk, 1);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // Step-end epilogue, all waves: each block's c log c and nonzero bins
    // into the per-pair LDS slots.
    auto epilogue = [&](int s) __attribute__((always_inline)) {
        __syncthreads();   // the pending finalize has read and cleared the slots
        // small-bin tile: the step's L and S1 of the tile's DOFs for the
        // finalize, loaded by a multiply-only wave before its blocks and
        // stored to LDS after them, so that the blocks hide the load latency
        const bool ldw = kSmall && wave == 7 && lane < 2 * R;
        double eL = 0.0, eS = 0.0;
        if (ldw) {
            const double* const dL = kBatch ? sdp->dof_L : a.dof_L;
            const double* const dS = kBatch ? sdp->dof_S1 : a.dof_S1;
            const double* const mo = kBatch ? sdp->mi : a.mi;
            const long long g0 = lane < R ? r0 + lane : c0 + (lane - R);
            const long long g = g0 < D ? g0 : D - 1;
            eL = dL[g * S + s];
            if (mo) eS = dS[g * S + s];
        }
        const int kfix = a.kfix;
        unsigned long long need = ~0ull;
        if constexpr (kSmall) {
            const unsigned long long m = *blkmask;
            need = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(m >> 32)) << 32) |
                   (unsigned)__builtin_amdgcn_readfirstlane((unsigned)m);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if constexpr (kSmall) {
                    if ((need >> (8 * (rowblk(i) >> 5) + (colblk(j) >> 5))) & 1)
                        epi_block_small<kNz>(counts_of(acc[i][j]), rowblk(i), colblk(j) + (lane & 31), lane, sm,
                                             R, nstr, ntab, kfix);
                }
                else
                    epi_block8<kNz>(counts_of(acc[i][j]), rowblk(i), colblk(j) + (lane & 31), lane,
                                    sm, nb, R, nstr, ntab, kfix);
                __builtin_amdgcn_sched_barrier(0);
            }
        if (ldw) {
            dcL[lane] = eL;
            dcS[lane] = eS;
        }
    };
    // Slot sums of step fs -> S2 / MI, by wave fwave (never a stager; on the
    // small-bin tile by every wave) one iteration after the epilogue, so
    // that it overlaps the next chunk; the iteration barrier orders it after
    // the atomics, and the next epilogue's opening barrier after it.
    auto finalize = [&](int fs) __attribute__((always_inline)) {
        const double* const coef = sm.invt + 256;
        const double n = (double)(fs + 1) * (double)ss;
        const double logn = log_pos_lds(n, sm.logt, sm.invt, coef);
        const int kfix = a.kfix;
        asm volatile("" ::: "memory");   // read the output pointers here
        double* const s2 = kBatch ? sdp->s2 : a.s2;
        double* const mi = kBatch ? sdp->mi : a.mi;
        // small-bin tile: inputs from the LDS caches (bin widths and pair
        // indices of the tile, L and S1 of the step), no global load
        const double* const dof_w = kBatch ? sdp->dof_w : a.dof_w;
        const double* const dof_L = kBatch ? sdp->dof_L : a.dof_L;
        const double* const dof_S1 = kBatch ? sdp->dof_S1 : a.dof_S1;
        for (int pq = kSmall ? tid : lane; pq < R * R; pq += kSmall ? 512 : 64) {
            const int rd = pq / R, cd = pq - rd * R;
            const long long gi = r0 + rd, gj = c0 + cd;
            const long long p =
                gi >= gj || gj >= D
                    ? -1
                    : (a.tilemap ? (long long)(kSmall ? tmap[pq] : a.tilemap[(size_t)tix * R * R + pq])
                                 : tri_index(gi, gj, D));
            unsigned long long Tq = 0ull;
            int nz = 0;
            for (int k = 0; k < nstr; ++k) {
                Tq += sm.slotT[pq * nstr + k];
                sm.slotT[pq * nstr + k] = 0ull;
                if constexpr (kNz) {
                    nz += sm.slotN[pq * nstr + k];
                    sm.slotN[pq * nstr + k] = 0;
                }
            }
            if (p >= 0) {
                const double T = from_fixed(Tq, kfix);
                const double L = kSmall ? dcL[rd] + dcL[R + cd] : dof_L[gi * S + fs] + dof_L[gj * S + fs];
                const double area = kSmall ? dcw[rd] * dcw[R + cd] : dof_w[gi] * dof_w[gj];
                const double ml = (logn + log_pos_lds(area, sm.logt, sm.invt, coef)) - T / n + L / n;
                const double mm = ml + (double)(unsigned)(nz - 1) / (2.0 * n);
                if (a.slot_ml >= 0) s2[((size_t)p * a.nest + a.slot_ml) * S + fs] = ml;
                if (a.slot_mm >= 0) s2[((size_t)p * a.nest + a.slot_mm) * S + fs] = mm;
                if (mi) {
                    const double s1 =
                        kSmall ? dcS[rd] + dcS[R + cd] : dof_S1[gi * S + fs] + dof_S1[gj * S + fs];
                    mi[(size_t)p * S + fs] = s1 - (a.mi_from_mm ? mm : ml);
                }
            }
        }
    };

    // One barrier per chunk.  Chunk g multiplies from buffer g % 3 while its
    // fragment registers are reloaded with chunk g+1 (buffer (g+1) % 3,
    // staged during chunk g-1) and the stager waves stage chunk g+2 into
    // buffer (g+2) % 3, whose last reads (chunk g-1) all waves finished
    // before the previous barrier.  At a step's last chunk the epilogue needs
    // the registers, so that chunk reads nothing ahead and the next chunk's
    // fragments are read after the barrier.
    int ps = 0, pc = 0;   // chunk to stage next
    int fin = -1;         // step whose finalize is pending
    auto run = [&](auto stg) __attribute__((always_inline)) {
        constexpr bool kStg = decltype(stg)::value;
        // ring of code dwords: chunks g+2 and g+3 before chunk g's staging
        CodeT w0{}, w1{};
        if constexpr (kStg) {
            w0 = loadc();
            w1 = loadc();
            set2 = stagec(0, nvalid_of(ps, pc), w0, 0u);
        }
        advance(ps, pc, nch);
        if constexpr (kStg) {
            set1 = stagec(1, nvalid_of(ps, pc), w1, 0u);
            w0 = loadc();
            w1 = loadc();
        }
        advance(ps, pc, nch);
        __syncthreads();
        read_all(0);
        int b = 0;   // buffer of the chunk being multiplied
        // one chunk: staging of the chunk two ahead (stager waves), MFMAs
        // reading the next chunk's fragments ahead with kPref, or the pending
        // finalize
        auto chunk = [&](auto pref) __attribute__((always_inline)) {
            const int bn = b == kBuf8 - 1 ? 0 : b + 1;
            const int bs = bn == kBuf8 - 1 ? 0 : bn + 1;
            const uint32_t nbase = (uint32_t)(bn * 2 * kOpB);
            CodeT w2{};
            if constexpr (kStg) w2 = loadc();   // chunk g+4
            // the staging goes first: its scattered byte writes (bank
            // conflicts depend on the codes) then drain behind the MFMAs
            // instead of holding up the barrier (29.8 against 31.4 ms on the
            // headline tiles)
            if constexpr (kStg) {
                // past the last chunk nvalid is 0: everything goes to the trash row
                const uint32_t now = stagec(bs, nvalid_of(ps, pc), w0, set3);
                set3 = set2;
                set2 = set1;
                set1 = now;
                w0 = w1;
            }
            mfma_kk(pref, fa0, fb0, nbase, 0);
            if constexpr (!kF4) mfma_kk(pref, fa1, fb1, nbase, 1);
            if constexpr (kStg) w1 = w2;
            // the small-bin tile's 1,024 pairs: every wave, two pairs a lane
            if constexpr (kSmall) {
                if (fin >= 0) finalize(fin);
            } else if constexpr (!kStg) {
                if (fin >= 0 && wave == fwave) finalize(fin);
            }
            fin = -1;
            advance(ps, pc, nch);
            b = bn;
            return nbase;
        };
        for (int s = 0; s < S; ++s) {
            for (int c = 0; c + 1 < nch; ++c) {
                chunk(std::true_type{});
                __syncthreads();
            }
            // the step's last chunk: the epilogue needs the registers, so the
            // next chunk's fragments are read after the barrier
            const uint32_t nbase = chunk(std::false_type{});
            epilogue(s);
            fin = s;
            __syncthreads();
            if (s + 1 < S) read_all(nbase);
        }
    };
    if (stager) run(std::true_type{});
    else run(std::false_type{});
    if (fin >= 0 && (kSmall || wave == fwave)) finalize(fin);
}

template <bool kNz, bool kSmall, bool kF4>
__global__ __launch_bounds__(512) void pair_tile8_kernel(PairKernelArgs a) {
    tile8_body<false, kNz, kSmall, kF4>(a, blockIdx.x);
}

// Batch form of the eight-wave tile, persistent: one workgroup per CU takes
// entries of the ordered batch tile list (every system's tiles, those of a
// partial last block included) from a shared counter until the list is
// exhausted.  The counter value is handed to the workgroup through the word
// after the tile body's LDS.  The workgroup finishing a system's last tile
// stamps the system's completion (100 MHz wall clock).
// Persistent batch kernel: the tile-independent LDS tables are written once,
// then each tile of the loop writes only its operand buffers and pair slots.
template <bool kNz, bool kSmall, bool kF4>
__global__ __launch_bounds__(512) void pair_batch8_kernel(PairKernelArgs a, uint32_t lds_next) {
    extern __shared__ __align__(16) uint8_t smem_raw[];
    int* const s_t = reinterpret_cast<int*>(smem_raw + lds_next);
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.t0) *a.t0 = wall_clock64();
    tile8_tables<kNz, kSmall, kF4>(a.R, a.nb, a.kfix);
    for (;;) {
        __syncthreads();   // the previous tile is done with LDS and s_t
        if (threadIdx.x == 0) *s_t = atomicAdd(a.work, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(*s_t);
        if (t >= a.n_work) break;
        tile8_body<true, kNz, kSmall, kF4, false>(a, t);
        if (threadIdx.x == 0 && a.sys_count) {
            const int sys = a.tile_sys[t];
            __threadfence();
            if (atomicAdd(&a.sys_count[sys], 1) == a.sys_tiles[sys] - 1)
                a.sys_time[sys] = wall_clock64();
        }
    }
}

}  // namespace

// Small-bin tiles (nb <= 8) have 32 DOFs a side (8 rows each) and take the
// tiles of a partial last block too; otherwise R <= 8, i.e. nb >= 29, so that
// a lane's 16 rows of a block touch at most two DOFs.
bool pair_tile8_small(int nb) { return nb >= 1 && nb <= 8; }

bool pair_tile8_supported(int R, int nb) {
    const bool small = pair_tile8_small(nb);
    const Lay8 L = lay8(R, small, true, false);
    if (L.ntab < 256 || L.total > kLdsCap) return false;
    return small ? R == 32 : (R <= 8 && 2 * R * kGroups <= 512);
}

int pair_fixed_k(double n_max) { return fixed_k(n_max, std::log(n_max)); }

bool pair_f4_exact(double n_max) { return n_max <= 16777216.0; }

namespace {

// the eight-wave kernel of a launch: batch or not, nonzero-bin slots,
// small-bin tile, FP4 operands
template <bool kNz, bool kSmall, bool kF4>
const void* tile8_fn(bool batch) {
    return batch ? reinterpret_cast<const void*>(&pair_batch8_kernel<kNz, kSmall, kF4>)
                 : reinterpret_cast<const void*>(&pair_tile8_kernel<kNz, kSmall, kF4>);
}

const void* tile8_pick(bool batch, bool nz, bool small, bool f4) {
    static const void* (*const tab[8])(bool) = {
        tile8_fn<false, false, false>, tile8_fn<false, false, true>,
        tile8_fn<false, true, false>,  tile8_fn<false, true, true>,
        tile8_fn<true, false, false>,  tile8_fn<true, false, true>,
        tile8_fn<true, true, false>,   tile8_fn<true, true, true>};
    return tab[(nz ? 4 : 0) + (small ? 2 : 0) + (f4 ? 1 : 0)](batch);
}

}  // namespace

hipError_t launch_pair_tiles8(const PairKernelArgs& a_in, int n_tiles, hipStream_t st) {
    if (n_tiles <= 0) return hipSuccess;
    if (!pair_tile8_supported(a_in.R, a_in.nb)) return hipErrorInvalidValue;
    // the fixed point of the slot sums, from the last step's frame count
    PairKernelArgs a = a_in;
    const double n_max = (double)a.nsteps * (double)a.ss;
    a.kfix = pair_fixed_k(n_max);
    const bool f4 = !a.int8_only && pair_f4_exact(n_max);
    const bool small = pair_tile8_small(a.nb);
    // nonzero bins are counted only for the Miller-Madow estimator
    const bool nz = a.slot_mm >= 0 || a.mi_from_mm;
    const size_t lds = lay8(a.R, small, nz, f4).total;
    const void* fn = tile8_pick(false, nz, small, f4);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    void* args[] = {&a};
    e = hipLaunchKernel(fn, dim3(n_tiles), dim3(512), args, lds, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// the caller sets a.int8_only when the batch's largest frame count is past
// pair_f4_exact
hipError_t launch_pair_batch8(const PairKernelArgs& a_in, int grid, hipStream_t st) {
    if (a_in.n_work <= 0 || grid <= 0) return hipSuccess;
    if (!pair_tile8_supported(a_in.R, a_in.nb)) return hipErrorInvalidValue;
    PairKernelArgs a = a_in;
    const bool small = pair_tile8_small(a.nb);
    const bool nz = a.slot_mm >= 0 || a.mi_from_mm;
    const bool f4 = !a.int8_only;
    const size_t body = lay8(a.R, small, nz, f4).total;
    uint32_t next = (uint32_t)((body + 15) & ~(size_t)15);
    const size_t lds = next + 16;
    if (lds > kLdsCap) return hipErrorInvalidValue;
    const void* fn = tile8_pick(true, nz, small, f4);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    void* args[] = {&a, &next};
    e = hipLaunchKernel(fn, dim3(grid), dim3(512), args, lds, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}
