// Pair stage: joint counts as one-hot int8 MFMA contractions over frames,
// with the plogp entropy epilogue fused at every convergence-step boundary.
//
// What it replaces: binData2D + entropy2D (Estimators.cpp:130-205, 391-651),
// called once per DOF pair by EntropyCalculator::run2d_xx / run2d_xy
// (EntropyCalculator.cpp:530-564).
//
// Formulation.  For the DOFs of one tile, the one-hot matrix X has one row
// per (DOF, bin) and one column per frame: X[d*nb + b][f] = (code[d][f] == b).
// The joint counts of every DOF pair of the tile are the blocks of X_r X_c^T,
// an int8 GEMM with exact int32 accumulation (v_mfma_i32_32x32x32_i8).
// A tile is 256 one-hot rows x 256 one-hot columns: R = floor(256/nb) whole
// DOFs per side (5 DOFs = 250 rows for nb = 50), so every pair lies inside
// one tile and its counts are never written to memory.
//
// The frame axis is walked in chunks of kChunk frames aligned to the step
// boundaries (step s = frames [s*ss, (s+1)*ss), remainder frames dropped as
// Estimators.cpp:162 does).  Accumulators are never reset, so at the end of
// step s they hold the prefix counts binData2D produces after its prefix sum
// (Estimators.cpp:183-188).  The epilogue then adds c*log(c) over each pair's
// bins and writes, per pair and step,
//     S2 = log(n) + log(w_i w_j) - T/n + (L_i + L_j)/n,   T = sum c log c,
// which is entropy2D's ML sum -sum (c/n) log(c / (n w_i w_j J_i J_j)) with
// the Jacobian part folded into the marginal sums L_i = sum_b m_ib log J_ib
// (m = 1-D counts of the same frames).  MM adds (nonzero bins - 1)/(2n).
//
// One-hot staging.  Codes of the tile's DOFs arrive in 256-frame
// super-chunks with 16-byte loads, issued one super-chunk (four chunks of
// MFMA work) ahead of use so their latency is hidden.  The one-hot rows are
// written into LDS by setting one byte per (DOF, frame) and clearing the
// byte the same slot set two chunks earlier, so a chunk costs a few byte
// writes per thread instead of rebuilding 32 KB of zeros.  With nb <= 8 the
// four-wave general path writes each slot's nb rows as whole dwords instead
// (nothing to clear).  Two LDS buffers alternate: chunk g+1 is staged while
// chunk g feeds the MFMAs.
//
// Waves: 4 per workgroup in a 2x2 layout, each owning a 128x128 quadrant
// (4x4 MFMA 32x32 blocks = 256 int32 accumulators per lane).  Blocks with no
// wanted pair (padding rows/cols, the lower triangle of diagonal tiles) are
// skipped by a wave-uniform mask.
//
// Tiles with R DOFs on both sides and 29 <= nb take the eight-wave form
// below (pair_tile8_kernel): two waves per SIMD, 128 accumulators each, so
// that one wave's staging and epilogue overlap the other's MFMAs.  The
// four-wave form keeps the rest: tiles of a partial block, count dumps and
// smaller bin counts.
//
// Frame split.  A plan with few tiles at small nb can run K workgroups per
// tile, each counting a K-th of every step into its own count dump
// (pair_tile_kernel in dump mode with nsplit = K); split_entropy_kernel then
// sums the K dumps and evaluates S2 and MI with the epilogue's formulas.
#include "pair_common.h"

namespace {

// One tile: all DOF pairs of row block tile.x x column block tile.y of
// tile list entry tix, over every frame and step.  With a batch (a_in.systems)
// the entry belongs to system a_in.tile_sys[tix], whose codes, sizes and
// outputs replace those of a_in.
template <bool kDump, int kSlots, bool kBatch>
__device__ __forceinline__ void tile_body(const PairKernelArgs& a_in, int tix) {
    extern __shared__ __align__(16) uint8_t smem_raw[];
    PairKernelArgs a = a_in;
    if constexpr (kBatch) {
        const BatchSystem& sd = a_in.systems[a_in.tile_sys[tix]];
        a.codes = sd.codes;
        a.stride = sd.stride;
        a.n_dof = sd.n_dof;
        a.ss = sd.ss;
        a.sstride = sd.sstride;
        a.s2 = sd.s2;
        a.mi = sd.mi;
        a.dof_w = sd.dof_w;
        a.dof_L = sd.dof_L;
        a.dof_S1 = sd.dof_S1;
    }
    const int R = a.R, nb = a.nb, S = a.nsteps;
    const LdsLayout lay = lds_layout(R);
    const size_t omap = lay.map;
    const int nstr = lay.nstripe, ntab = lay.ntab;
    Smem sm;
    sm.oh = smem_raw;
    sm.cl = reinterpret_cast<uint32_t*>(smem_raw + lay.cl);
    sm.row2dof = smem_raw + omap;
    sm.col2dof = smem_raw + omap + 256;
    sm.slotT = reinterpret_cast<unsigned long long*>(smem_raw + lay.T);
    sm.slotN = reinterpret_cast<int*>(smem_raw + lay.N);
    sm.logt = reinterpret_cast<double*>(smem_raw + lay.log);
    sm.invt = sm.logt + 256;
    sm.ctab = reinterpret_cast<double*>(smem_raw + lay.tab);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
    const int wr = wave >> 1, wc = wave & 1;
    const int2 tile = a.tiles[tix];
    const int r0 = tile.x * R, c0 = tile.y * R;
    const int nr = min(R, a.n_dof - r0), nc = min(R, a.n_dof - c0);
    const bool diag = tile.x == tile.y;
    const int ngen = diag ? nr : nr + nc;   // diagonal tiles share operands
    const long long D = a.n_dof;

    // ---- LDS initialisation ------------------------------------------------
    {
        uint4* p = reinterpret_cast<uint4*>(sm.oh);
        const uint4 z = make_uint4(0, 0, 0, 0);
        for (int i = tid; i < 2 * 2 * kOpBytes / 16; i += 256) p[i] = z;
        for (int i = tid; i < 2 * 2 * R * kGroups; i += 256) sm.cl[i] = 0xFFFFFFFFu;
        for (int i = tid; i < 256; i += 256) {
            int d = i / nb;
            sm.row2dof[i] = (uint8_t)(d < nr ? d : 0xFF);
            sm.col2dof[i] = (uint8_t)(d < nc ? d : 0xFF);
        }
        for (int i = tid; i < R * R * nstr; i += 256) {
            sm.slotT[i] = 0ull;
            sm.slotN[i] = 0;
        }
        {
            const double t = 1.0 + tid * (1.0 / 256.0);
            sm.logt[tid] = log(t);
            sm.invt[tid] = 1.0 / t;
        }
        for (int c = tid; c < ntab; c += 256) sm.ctab[c] = c > 1 ? (double)c * log((double)c) : 0.0;
    }

    // ---- wave-uniform mask of MFMA blocks holding at least one wanted pair --
    uint32_t mask = 0;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) {
            const int rlo = 128 * wr + 32 * bi, clo = 128 * wc + 32 * bj;
            const int drlo = rlo / nb, dclo = clo / nb;
            const int dchi = min((clo + 31) / nb, nc - 1);
            bool need = drlo < nr && dclo < nc;
            if (diag) need = need && drlo < dchi;
            if (need) mask |= 1u << (bi * 4 + bj);
        }
    }

    const int nch = (int)((a.ss + kChunk - 1) / kChunk);
    const int G = S * nch;
    // dword staging needs every step to start on a 4-byte boundary; the
    // step length itself may be any number of frames
    const bool aligned = ((a.stride & 3) == 0) && ((a.sstride & 3) == 0) &&
                         ((reinterpret_cast<uintptr_t>(a.codes) & 3) == 0);

    // ---- code staging --------------------------------------------------------
    // A staging slot is 4 consecutive frames of one DOF within a chunk: slot
    // q -> DOF q / kGroups, frames 4 * (q % kGroups) .. +3.  Codes arrive in
    // super-chunks of kSuper chunks: one dword load per slot and chunk, all
    // issued kSuper chunks (of MFMA work) before they are needed.  0xFF marks
    // a frame outside its chunk or beyond the last chunk.
    auto load_super = [&](int j, uint32_t (&dst)[kSlots][kSuper]) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < kSuper; ++c) {
            const int g = j * kSuper + c;
            int nvalid = 0;
            uint64_t f0 = 0;
            if (g < G) {
                const int s = g / nch, cc = g - s * nch;
                f0 = (uint64_t)s * a.sstride + (uint64_t)cc * kChunk;
                nvalid = (int)min((uint64_t)kChunk, a.ss - (uint64_t)cc * kChunk);
            }
#pragma unroll
            for (int k = 0; k < kSlots; ++k) {
                const int q = tid + 256 * k;
                const int ld = q / kGroups, k4 = (q - ld * kGroups) * 4;
                uint32_t w = 0xFFFFFFFFu;
                if (ld < ngen && k4 < nvalid) {
                    const int dof = ld < nr ? r0 + ld : c0 + (ld - nr);
                    const uint8_t* src = a.codes + (uint64_t)dof * a.stride + f0 + k4;
                    if (aligned && k4 + 4 <= nvalid) {
                        w = *reinterpret_cast<const uint32_t*>(src);
                    } else {
                        w = 0;
                        for (int e = 0; e < 4; ++e)
                            w |= (k4 + e < nvalid ? (uint32_t)src[e] : 0xFFu) << (8 * e);
                    }
                }
                dst[k][c] = w;
            }
        }
    };
    // One-hot rows of chunk g into buffer g&1 from the dwords of its
    // super-chunk.  A code >= nb raises the error flag and is dropped.
    auto write_onehot = [&](int g, const uint32_t (&cw)[kSlots][kSuper]) __attribute__((always_inline)) {
        const int b = g & 1;
        const int cis = g - (g / kSuper) * kSuper;   // chunk within super-chunk
        const int sg = g / nch, cg = g - sg * nch;
        const int nvalid = (int)min((uint64_t)kChunk, a.ss - (uint64_t)cg * kChunk);
        (void)sg;
#pragma unroll
        for (int k = 0; k < kSlots; ++k) {
            const int q = tid + 256 * k;
            const int ld = q / kGroups, k4 = (q - ld * kGroups) * 4;
            if (ld < ngen) {
                // select dword cis with masks (a select chain on a runtime
                // index would be turned into an indexed scratch access)
                uint32_t wv = 0;
#pragma unroll
                for (int c = 0; c < kSuper; ++c) wv |= cw[k][c] & (0u - (uint32_t)(cis == c));
                const bool isA = ld < nr;
                const int rbase = (isA ? ld : ld - nr) * nb;
                uint8_t* op = sm.oh + (size_t)(b * 2 + (isA ? 0 : 1)) * kOpBytes + k4;
                if (nb <= 8) {   // workgroup-uniform
                    // Whole dwords: the slot's 4 frames in each of the DOF's
                    // nb rows are rewritten, so nothing from two chunks ago
                    // needs clearing.  Byte e of row r is 1 where code e == r
                    // (exact per-byte zero test of code ^ r, no carries
                    // between bytes); frames outside the chunk read as 0xFF,
                    // which matches no row.
                    const int nv = nvalid - k4;
                    const uint32_t vm = nv >= 4 ? 0xFFFFFFFFu : (nv <= 0 ? 0u : 0xFFFFFFFFu >> (8 * (4 - nv)));
                    const uint32_t wx = wv | ~vm;
                    bool bad = false;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        bad |= (e < nv) && ((wx >> (8 * e)) & 0xFFu) >= (uint32_t)nb;
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        if (r < nb) {
                            const uint32_t x = wx ^ (0x01010101u * (uint32_t)r);
                            const uint32_t z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
                            *reinterpret_cast<uint32_t*>(op + (rbase + r) * kLdsRow) = z >> 7;
                        }
                    }
                    if (bad) atomicOr(a.flag, 1);
                    continue;
                }
                uint32_t* slot = sm.cl + ((size_t)b * 2 * R + ld) * kGroups + (k4 >> 2);
                const uint32_t old = *slot;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t ob = (old >> (8 * e)) & 0xFFu;
                    if (ob != 0xFFu) op[(rbase + ob) * kLdsRow + e] = 0;
                }
                bool bad = false;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t cv = (wv >> (8 * e)) & 0xFFu;
                    if (k4 + e >= nvalid) continue;          // frame outside the chunk
                    if (cv >= (uint32_t)nb) {                // includes a stored 0xFF
                        bad = true;
                        wv |= 0xFFu << (8 * e);
                    } else {
                        op[(rbase + cv) * kLdsRow + e] = 1;
                    }
                }
                if (bad) atomicOr(a.flag, 1);
                *slot = wv;
            }
        }
    };

    v16i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v16i{};

    // MFMA half kk (32 frames) of chunk g from buffer g&1.
    auto mfma_half = [&](int g, int kk) __attribute__((always_inline)) {
        if (!mask) return;
        const uint8_t* A = sm.oh + (size_t)((g & 1) * 2) * kOpBytes;
        const uint8_t* B = diag ? A : A + kOpBytes;
        const int koff = kk * 32 + (lane >> 5) * 16;
        v4i af[4], bf[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            af[i] = *reinterpret_cast<const v4i*>(
                A + (128 * wr + 32 * i + (lane & 31)) * kLdsRow + koff);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            bf[j] = *reinterpret_cast<const v4i*>(
                B + (128 * wc + 32 * j + (lane & 31)) * kLdsRow + koff);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (mask & (1u << (i * 4 + j)))
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                        af[i], bf[j], acc[i][j], 0, 0, 0);
    };

    // Step-boundary epilogue: per-pair sum of c*log(c) and nonzero bins.
    // park: 4 KB of LDS per wave where one block's counts are parked so the
    // accumulators are read once, straight from AGPRs; with rezero_g >= 0 the
    // park area lies in buffer rezero_g & 1, which is cleared afterwards.
    // S2 / MI wanted from this launch (a count-only dump wants neither)
    const bool fused = a.slot_ml >= 0 || a.slot_mm >= 0 || a.mi != nullptr;
    auto epilogue = [&](int s, int* park, int rezero_g) __attribute__((always_inline)) {
        // One 32x32 block: the lane holds column (lane & 31) and rows
        // 8*(r/4) + 4*(lane/32) + r%4 for r = 0..15.
        const bool two_dof = nb >= 28 && !kDump;
        // frames of the step and the fixed-point scale of its slot sums
        const double n = (double)(s + 1) * (double)a.ss;
        const double logn = log(n);
        const int kfix = fixed_k(n, logn);
        auto epi = [&](const v16i& blk, const int i, const int j) __attribute__((always_inline)) {
            if (!(mask & (1u << (i * 4 + j)))) return;   // wave-uniform
            const int gcol = 128 * wc + 32 * j + (lane & 31);
            const int cd = sm.col2dof[gcol];
            const int row0 = 128 * wr + 32 * i + 4 * (lane >> 5);
            // park the block's counts (4 KB per wave) and read them back: the
            // accumulators leave the AGPRs in four 16-byte stores per lane
            v4i* pk = reinterpret_cast<v4i*>(park) + lane;
            pk[0] = v4i{blk[0], blk[1], blk[2], blk[3]};
            pk[64] = v4i{blk[4], blk[5], blk[6], blk[7]};
            pk[128] = v4i{blk[8], blk[9], blk[10], blk[11]};
            pk[192] = v4i{blk[12], blk[13], blk[14], blk[15]};
            asm volatile("" ::: "memory");
            if (two_dof) {
                // nb >= 28: the lane's 16 rows (row0 .. row0+27) touch at most
                // two DOFs; split them at the first row of the second one.
                const int dlo = sm.row2dof[row0];
                const bool ok = (cd != 0xFF) & (dlo != 0xFF);
                const int kk = (dlo + 1) * nb - row0;   // offsets >= kk: second DOF
                int c[16];
                int mx = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const v4i c4 = pk[64 * q];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        c[4 * q + t] = ok ? c4[t] : 0;
                        mx = max(mx, c[4 * q + t]);
                    }
                }
                double tot = 0.0, hi = 0.0;
                int ntot = 0, nhi = 0;
                auto add = [&](int r, double v) __attribute__((always_inline)) {
                    const bool h = 8 * (r >> 2) + (r & 3) >= kk;
                    const int nz = min(c[r], 1);
                    tot += v;
                    hi += h ? v : 0.0;
                    ntot += nz;
                    nhi += h ? nz : 0;
                };
                // counts below the table size (the usual case) are one LDS
                // read each, all issued before use; the terms of larger counts
                // are recomputed by the series, per count position and only
                // where some lane of the wave holds one
                if (ntab > 0) {
                    double tv[16];
#pragma unroll
                    for (int r = 0; r < 16; ++r) tv[r] = sm.ctab[min(c[r], ntab - 1)];
                    if (__builtin_amdgcn_ballot_w64(mx >= ntab) != 0) {
#pragma unroll
                        for (int r = 0; r < 16; ++r)
                            if (__builtin_amdgcn_ballot_w64(c[r] >= ntab) != 0 && c[r] >= ntab)
                                tv[r] = (double)c[r] * log_count(c[r], sm.logt, sm.invt);
                    }
#pragma unroll
                    for (int r = 0; r < 16; ++r) add(r, tv[r]);
                } else {
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        add(r, (double)c[r] * log_count(max(c[r], 1), sm.logt, sm.invt));
                }
                if (!ok) return;
                const int stripe = lane % nstr;
                if (ntot > nhi) {
                    atomicAdd(&sm.slotT[(dlo * R + cd) * nstr + stripe], to_fixed(tot - hi, kfix));
                    atomicAdd(&sm.slotN[(dlo * R + cd) * nstr + stripe], ntot - nhi);
                }
                const int dhi = sm.row2dof[row0 + 27];
                if (nhi && dhi != 0xFF && dhi != dlo) {
                    atomicAdd(&sm.slotT[(dhi * R + cd) * nstr + stripe], to_fixed(hi, kfix));
                    atomicAdd(&sm.slotN[(dhi * R + cd) * nstr + stripe], nhi);
                }
                return;
            }
            if (cd == 0xFF) return;
            // general: one LDS atomic per nonzero count (none when the launch
            // only dumps counts)
#pragma unroll 1
            for (int q = 0; q < 4; ++q) {
                const v4i c4 = pk[64 * q];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int grow = row0 + 8 * q + t;
                    const int rd = sm.row2dof[grow];
                    const int cnt = c4[t];
                    if (fused && rd != 0xFF && cnt > 0) {
                        const double v = cnt < ntab ? sm.ctab[cnt]
                                                    : (double)cnt * log_count(cnt, sm.logt, sm.invt);
                        atomicAdd(&sm.slotT[(rd * R + cd) * nstr + lane % nstr], to_fixed(v, kfix));
                        atomicAdd(&sm.slotN[(rd * R + cd) * nstr + lane % nstr], 1);
                    }
                    if constexpr (kDump) {
                        if (rd != 0xFF) {
                            const long long gi = r0 + rd, gj = c0 + cd;
                            if (gi < gj) {
                                const long long p = a.tilemap ? (long long)a.tilemap[(size_t)tix * R * R + rd * R + cd]
                                                              : tri_index(gi, gj, D);
                                if (p >= 0) {
                                    const int bi_ = grow - rd * nb, bj_ = gcol - cd * nb;
                                    const int rec = a.split_rec ? a.split_rec : nb * nb;
                                    a.counts[((size_t)(p - (long long)a.count_base) * S + s) * rec +
                                             (size_t)bi_ * nb + bj_] =
                                        (uint32_t)cnt;
                                }
                            }
                        }
                    }
                }
            }
        };
#define CENTRE_EPI(I, J) \
    epi(acc[I][J], I, J);  \
    __builtin_amdgcn_sched_barrier(0);
        CENTRE_EPI(0, 0) CENTRE_EPI(0, 1) CENTRE_EPI(0, 2) CENTRE_EPI(0, 3)
        CENTRE_EPI(1, 0) CENTRE_EPI(1, 1) CENTRE_EPI(1, 2) CENTRE_EPI(1, 3)
        CENTRE_EPI(2, 0) CENTRE_EPI(2, 1) CENTRE_EPI(2, 2) CENTRE_EPI(2, 3)
        CENTRE_EPI(3, 0) CENTRE_EPI(3, 1) CENTRE_EPI(3, 2) CENTRE_EPI(3, 3)
#undef CENTRE_EPI
        __syncthreads();
        for (int q = fused ? tid : nr * nc; q < nr * nc; q += 256) {
            const int rd = q / nc, cd = q - rd * nc;
            const long long gi = r0 + rd, gj = c0 + cd;
            if (gi < gj) {
                const long long p = a.tilemap ? (long long)a.tilemap[(size_t)tix * R * R + rd * R + cd]
                                              : tri_index(gi, gj, D);
                if (p >= 0) {
                    unsigned long long Tq = 0ull;
                    int nz = 0;
                    for (int k = 0; k < nstr; ++k) {
                        Tq += sm.slotT[(rd * R + cd) * nstr + k];
                        nz += sm.slotN[(rd * R + cd) * nstr + k];
                    }
                    const double T = from_fixed(Tq, kfix);
                    const double L = a.dof_L[gi * S + s] + a.dof_L[gj * S + s];
                    const double area = a.dof_w[gi] * a.dof_w[gj];
                    const double ml = (logn + log(area)) - T / n + L / n;
                    const double mm = ml + (double)(unsigned)(nz - 1) / (2.0 * n);
                    if (a.slot_ml >= 0) a.s2[((size_t)p * a.nest + a.slot_ml) * S + s] = ml;
                    if (a.slot_mm >= 0) a.s2[((size_t)p * a.nest + a.slot_mm) * S + s] = mm;
                    if (a.mi) {
                        const double s1 = a.dof_S1[gi * S + s] + a.dof_S1[gj * S + s];
                        a.mi[(size_t)p * S + s] = s1 - (a.mi_from_mm ? mm : ml);
                    }
                }
            }
            for (int k = 0; k < nstr; ++k) {
                sm.slotT[(rd * R + cd) * nstr + k] = 0ull;
                sm.slotN[(rd * R + cd) * nstr + k] = 0;
            }
        }
        // parked counts overwrote buffer g&1: clear it and its staged codes
        // so the next staging into it starts from zeros
        if (rezero_g >= 0) {
            const int g = rezero_g;
            uint4* p = reinterpret_cast<uint4*>(sm.oh + (size_t)((g & 1) * 2) * kOpBytes);
            const uint4 z = make_uint4(0, 0, 0, 0);
            for (int i = tid; i < 2 * kOpBytes / 16; i += 256) p[i] = z;
            for (int i = tid; i < 2 * R * kGroups; i += 256)
                sm.cl[(size_t)(g & 1) * 2 * R * kGroups + i] = 0xFFFFFFFFu;
        }
    };

    // ---- fast path -----------------------------------------------------------
    // Off-diagonal tiles with R DOFs on both sides and 4-byte aligned codes
    // (the bulk of the work) run a branch-free loop: all 16 MFMA blocks of a
    // wave unconditionally (blocks of padding rows multiply zeros and are
    // skipped by the epilogue), byte writes that have nothing to do go to a
    // trash row instead of being predicated, and the codes a slot staged
    // two chunks ago stay in registers.  The loop body is straight-line code,
    // so the scheduler can spread the staging over the issue slots between
    // MFMAs, and the fragment reads of the next chunk are software-pipelined
    // behind the MFMAs of the current one.
    bool done = false;
    if constexpr (!kDump && kSlots <= 4) {
        // Codes were range-checked by the marginal pass that runs first on the
        // same stream (any code >= nb sets the flag); with the flag set the
        // results are discarded, so skip the work.
        const bool fast = !diag && nr == R && nc == R && aligned;   // WG-uniform
        if (fast) {
            done = true;
            uint8_t* const trash = smem_raw + omap + 512;   // kLdsRow bytes
            // one read of the flag for the whole workgroup
            __syncthreads();
            if (tid == 0) *reinterpret_cast<int*>(trash) = *reinterpret_cast<volatile int*>(a.flag);
            __syncthreads();
            if (*reinterpret_cast<const int*>(trash) != 0) return;
            // slot k of this thread: DOF ld = tid/16 + 16k, frames 4*(tid%16)..+3.
            // Slots beyond the 2R*16 live ones repeat a live slot of another
            // wave (same bytes to the same addresses), so that every lane does
            // the same work without predication.
            const int k4 = (tid & 15) * 4;
            const int nlive = 2 * R * kGroups;
            const uint8_t* src[kSlots];
            uint32_t dst[kSlots];
#pragma unroll
            for (int k = 0; k < kSlots; ++k) {
                const int q = (tid + 256 * k) % nlive;
                const int ld = q / kGroups;
                const bool isA = ld < R;
                const int dof = isA ? r0 + ld : c0 + (ld - R);
                src[k] = a.codes + (uint64_t)dof * a.stride + k4;
                dst[k] = (uint32_t)((isA ? 0 : kOpBytes) + (isA ? ld : ld - R) * nb * kLdsRow + k4);
            }
            // codes each slot set in buffer 0 / 1 (code 0 where nothing was set:
            // clearing a zero byte of the slot's own frame column is harmless)
            uint32_t cur[kSlots][kSuper], nxt[kSlots][kSuper], last0[kSlots], last1[kSlots];
#pragma unroll
            for (int k = 0; k < kSlots; ++k) last0[k] = last1[k] = 0u;
            auto advance = [](int& s, int& c, int n) __attribute__((always_inline)) {
                if (++c == n) { c = 0; ++s; }
            };
            auto nvalid_of = [&](int s, int c) __attribute__((always_inline)) -> int {
                return s < S ? (int)min((uint64_t)kChunk, a.ss - (uint64_t)c * kChunk) : 0;
            };
            int ls = 0, lc = 0;   // next chunk to load
            auto load4 = [&](uint32_t (&d)[kSlots][kSuper]) __attribute__((always_inline)) {
#pragma unroll
                for (int c = 0; c < kSuper; ++c) {
                    const uint64_t f0 = (uint64_t)ls * a.sstride + (uint64_t)lc * kChunk;
                    const int nv = nvalid_of(ls, lc);
#pragma unroll
                    for (int k = 0; k < kSlots; ++k) {
                        // frames outside the chunk read a harmless in-bounds word;
                        // staging ignores them by position
                        d[k][c] = *reinterpret_cast<const uint32_t*>(k4 < nv ? src[k] + f0 : a.codes);
                    }
                    advance(ls, lc, nch);
                }
            };
            // chunk with nv valid frames from dword cis of cw into buffer b:
            // clear the bytes the slot set two chunks ago, set the new ones.
            // Frames beyond the chunk (the tail of a step) set code 0 in the
            // trash row and are remembered as code 0.
            // kFull: all kChunk frames valid (every chunk but a step's last).
            auto stage = [&](auto full, int b, int cis, int nv, const uint32_t (&cw)[kSlots][kSuper])
                __attribute__((always_inline)) {
                constexpr bool kFull = decltype(full)::value;
                uint8_t* const buf = sm.oh + (size_t)b * 2 * kOpBytes;
                const int nvk = kFull ? 4 : nv - k4;   // valid frames of this lane's 4
                const uint32_t vmask = nvk >= 4 ? 0xFFFFFFFFu
                                                : (nvk <= 0 ? 0u : 0xFFFFFFFFu >> (8 * (4 - nvk)));
#pragma unroll
                for (int k = 0; k < kSlots; ++k) {
                    // select dword cis with masks (a select chain on a runtime
                    // index becomes an indexed scratch access)
                    uint32_t wv = 0;
#pragma unroll
                    for (int c = 0; c < kSuper; ++c) wv |= cw[k][c] & (0u - (uint32_t)(cis == c));
                    if constexpr (!kFull) wv &= vmask;
                    const uint32_t old = b ? last1[k] : last0[k];
                    uint8_t* const row = buf + dst[k];
#pragma unroll
                    for (int e = 0; e < 4; ++e) row[((old >> (8 * e)) & 0xFFu) * kLdsRow + e] = 0;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        uint8_t* const srow = (kFull || nvk > e) ? row : trash + k4;
                        srow[((wv >> (8 * e)) & 0xFFu) * kLdsRow + e] = 1;
                    }
                    last1[k] = b ? wv : last1[k];
                    last0[k] = b ? last0[k] : wv;
                }
            };
            auto frags = [&](const uint8_t* base, int kk, v4i (&af)[4], v4i (&bf)[4])
                __attribute__((always_inline)) {
                const int koff = kk * 32 + (lane >> 5) * 16;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    af[i] = *reinterpret_cast<const v4i*>(
                        base + (128 * wr + 32 * i + (lane & 31)) * kLdsRow + koff);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    bf[j] = *reinterpret_cast<const v4i*>(
                        base + kOpBytes + (128 * wc + 32 * j + (lane & 31)) * kLdsRow + koff);
            };
            auto mfma16 = [&](const v4i (&af)[4], const v4i (&bf)[4]) __attribute__((always_inline)) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
            };

            int* const parkf = reinterpret_cast<int*>(smem_raw + lay.park) + wave * 64 * 16;
            int ps = 0, pc = 0;      // position of the next chunk to stage
            int ms = 0, mc = 0;      // position of chunk g
            __syncthreads();
            load4(cur);
            load4(nxt);
            stage(std::false_type{}, 0, 0, nvalid_of(ps, pc), cur);
            advance(ps, pc, nch);
            __syncthreads();
            v4i xa0[4], xb0[4], xa1[4], xb1[4];   // fragments of a chunk in buffer 0
            v4i ya0[4], yb0[4], ya1[4], yb1[4];   // fragments of a chunk in buffer 1
            frags(sm.oh, 0, xa0, xb0);
            frags(sm.oh, 1, xa1, xb1);
            stage(std::false_type{}, 1, 1, nvalid_of(ps, pc), cur);
            advance(ps, pc, nch);
            mfma16(xa0, xb0);
            __syncthreads();
            // Iteration g (chunk g in buffer b = g & 1, its second-half
            // fragments already in registers; chunk g+1 staged in buffer b^1):
            // read chunk g+1's fragments and multiply chunk g's second half
            // while they arrive; run the step epilogue if chunk g ends a step;
            // stage chunk g+2 into buffer b (all of its reads completed before
            // the last barrier) between the MFMAs of chunk g+1's first half.
            // No MFMA waits on an LDS read issued after the preceding barrier.
            auto iter = [&](int g, int b, const v4i (&ha)[4], const v4i (&hb)[4],
                            v4i (&na0)[4], v4i (&nb0)[4], v4i (&na1)[4], v4i (&nb1)[4])
                __attribute__((always_inline)) {
                if (((g + 2) & (kSuper - 1)) == 0) {   // chunk g+2 opens a super-chunk
#pragma unroll
                    for (int k = 0; k < kSlots; ++k)
#pragma unroll
                        for (int c = 0; c < kSuper; ++c) cur[k][c] = nxt[k][c];
                    load4(nxt);
                }
                const uint8_t* nbase = sm.oh + (size_t)(b ^ 1) * 2 * kOpBytes;
                frags(nbase, 0, na0, nb0);
                frags(nbase, 1, na1, nb1);
                mfma16(ha, hb);
                // a wave may have at most 15 LDS operations in flight: issue 6
                // reads, then one read per MFMA, so that no read waits for a
                // slot while the MFMA pipe idles
                __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
#pragma unroll
                for (int i = 0; i < 10; ++i) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                }
#pragma unroll
                for (int i = 0; i < 6; ++i) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                if (mc == nch - 1) {
                    __syncthreads();
                    epilogue(ms, parkf, -1);
                }
                advance(ms, mc, nch);
                stage(std::false_type{}, b, (g + 2) & (kSuper - 1), nvalid_of(ps, pc), cur);
                advance(ps, pc, nch);
                mfma16(na0, nb0);
                // one MFMA at a time with a share of the staging VALU work and
                // one LDS write in its shadow
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x002, 2 * kSlots, 0);
                    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
                }
                __syncthreads();
            };
            for (int g = 0; g < G; g += 2) {
                iter(g, 0, xa1, xb1, ya0, yb0, ya1, yb1);
                if (g + 1 >= G) break;
                iter(g + 1, 1, ya1, yb1, xa0, xb0, xa1, xb1);
            }
        }
    }
    if (done) return;

    // ---- general path ------------------------------------------------------
    // iteration g: multiply the first half of chunk g, stage chunk g+1 while
    // those MFMAs run, multiply the second half.  The super-chunk holding
    // chunk g+1 was loaded kSuper iterations earlier; the loads of the next
    // super-chunk are issued as soon as the current one is taken over.
    uint32_t cur[kSlots][kSuper], nxt[kSlots][kSuper];
    __syncthreads();
    load_super(0, cur);
    if (G > kSuper) load_super(1, nxt);
    write_onehot(0, cur);
    __syncthreads();
    for (int g = 0; g < G; ++g) {
        mfma_half(g, 0);
        __builtin_amdgcn_sched_barrier(0);
        if (g + 1 < G) {
            if ((g + 1) % kSuper == 0) {
#pragma unroll
                for (int k = 0; k < kSlots; ++k)
#pragma unroll
                    for (int c = 0; c < kSuper; ++c) cur[k][c] = nxt[k][c];
                const int jn = (g + 1) / kSuper + 1;
                if (jn * kSuper < G) load_super(jn, nxt);
            }
            write_onehot(g + 1, cur);
        }
        __builtin_amdgcn_sched_barrier(0);
        mfma_half(g, 1);
        if (g % nch == nch - 1) {
            __syncthreads();                 // every wave is done with buffer g&1
            epilogue(g / nch,
                     reinterpret_cast<int*>(sm.oh + (size_t)((g & 1) * 2) * kOpBytes) + wave * 64 * 16,
                     g);
        }
        __syncthreads();
    }
}

template <bool kDump, int kSlots>
__global__ __launch_bounds__(256) void pair_tile_kernel(PairKernelArgs a) {
    int tix = blockIdx.x;
    if constexpr (kDump) {
        if (a.nsplit > 1) {
            const int k = tix / a.split_tiles;
            tix -= k * a.split_tiles;
            a.codes += (uint64_t)k * a.split_len;
            a.ss = (k == a.nsplit - 1) ? a.ss - (uint64_t)k * a.split_len : a.split_len;
            a.counts += (size_t)k * a.split_cstride;
        }
    }
    tile_body<kDump, kSlots, false>(a, tix);
}

// Workgroups (tile, slice of 256 (pair, step) items), one thread per item: sum the partial
// counts, then S2 = log n + log(w_i w_j) - sum c log c / n + (L_i + L_j) / n,
// the MM correction (nonzero bins - 1) / 2n and MI = S1_i + S1_j - S2, as the
// step epilogue of the tile kernels computes them.
__global__ __launch_bounds__(256) void split_entropy_kernel(PairKernelArgs a) {
    const int tix = blockIdx.x;
    const int R = a.R, nb = a.nb, S = a.nsteps;
    const int2 tile = a.tiles[tix];
    const int r0 = tile.x * R, c0 = tile.y * R;
    const int nr = min(R, a.n_dof - r0), nc = min(R, a.n_dof - c0);
    const long long D = a.n_dof;
    const int nbb = nb * nb;
    const int item = blockIdx.y * 256 + threadIdx.x;
    if (item < nr * nc * S) {
        const int q = item / S, s = item - q * S;
        const int rd = q / nc, cd = q - rd * nc;
        const long long gi = r0 + rd, gj = c0 + cd;
        if (gi >= gj) return;
        const long long p = a.tilemap ? (long long)a.tilemap[(size_t)tix * R * R + rd * R + cd]
                                      : tri_index(gi, gj, D);
        if (p < 0) return;
        // records of split_rec words (nb^2 rounded up to a multiple of 4,
        // 16-byte aligned); the padding words are never written
        const uint4* c0p = reinterpret_cast<const uint4*>(a.counts + ((size_t)p * S + s) * a.split_rec);
        const size_t kstride = a.split_cstride / 4;
        double T = 0.0;
        int nz = 0;
        for (int b4 = 0; b4 < a.split_rec / 4; ++b4) {
            uint4 c = c0p[b4];
            for (int k = 1; k < a.nsplit; ++k) {
                const uint4 v = c0p[(size_t)k * kstride + b4];
                c.x += v.x; c.y += v.y; c.z += v.z; c.w += v.w;
            }
            const uint32_t cv[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (4 * b4 + e >= nbb) break;
                nz += cv[e] > 0;
                if (cv[e] > 1) T += (double)cv[e] * log((double)cv[e]);
            }
        }
        const double n = (double)(s + 1) * (double)a.ss;
        const double L = a.dof_L[gi * S + s] + a.dof_L[gj * S + s];
        const double area = a.dof_w[gi] * a.dof_w[gj];
        const double ml = (log(n) + log(area)) - T / n + L / n;
        const double mm = ml + (double)(unsigned)(nz - 1) / (2.0 * n);
        if (a.slot_ml >= 0) a.s2[((size_t)p * a.nest + a.slot_ml) * S + s] = ml;
        if (a.slot_mm >= 0) a.s2[((size_t)p * a.nest + a.slot_mm) * S + s] = mm;
        if (a.mi) {
            const double s1 = a.dof_S1[gi * S + s] + a.dof_S1[gj * S + s];
            a.mi[(size_t)p * S + s] = s1 - (a.mi_from_mm ? mm : ml);
        }
    }
}

// Persistent batch: every workgroup takes tile list entries from a shared
// counter until the list is exhausted; the workgroup finishing the last tile
// of a system stamps that system's completion time (100 MHz wall clock).
template <int kSlots>
__global__ __launch_bounds__(256) void pair_persist_kernel(PairKernelArgs a) {
    __shared__ int s_t;
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.t0) *a.t0 = wall_clock64();
    for (;;) {
        __syncthreads();   // the previous tile is done with LDS and s_t
        if (threadIdx.x == 0) s_t = atomicAdd(a.work, 1);
        __syncthreads();
        const int t = s_t;
        if (t >= a.n_work) break;
        tile_body<false, kSlots, true>(a, t);
        if (threadIdx.x == 0 && a.sys_count) {
            const int sys = a.tile_sys[t];
            __threadfence();
            if (atomicAdd(&a.sys_count[sys], 1) == a.sys_tiles[sys] - 1)
                a.sys_time[sys] = wall_clock64();
        }
    }
}


// Copies the ss frames of every step of every DOF to rows where each step
// starts on a 4-byte boundary: out[d][s * ss4 + f] = codes[d][s * ss + f]
// (padding bytes 0, never counted).
__global__ __launch_bounds__(256) void repack_kernel(const uint8_t* __restrict__ codes,
                                                     uint64_t stride, uint64_t ss, int S,
                                                     uint64_t ss4, uint8_t* __restrict__ out) {
    const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t d = blockIdx.y;
    const uint64_t row = (uint64_t)S * ss4;
    if (q >= row) return;
    const uint64_t s = q / ss4, f = q - s * ss4;
    out[d * row + q] = f < ss ? codes[d * stride + s * ss + f] : 0;
}

}  // namespace

hipError_t launch_repack(const uint8_t* codes, uint64_t stride, uint32_t D, uint64_t ss, int S,
                         uint64_t ss4, uint8_t* out, hipStream_t st) {
    const uint64_t row = (uint64_t)S * ss4;
    hipLaunchKernelGGL(repack_kernel, dim3((unsigned)((row + 255) / 256), D), dim3(256), 0, st,
                       codes, stride, ss, S, ss4, out);
    return hipGetLastError();
}

size_t pair_tile_lds_bytes(int R) { return lds_layout(R).total; }

template <bool kDump, int kSlots>
static hipError_t launch_t(const PairKernelArgs& a, int n_tiles, size_t lds, hipStream_t st) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&pair_tile_kernel<kDump, kSlots>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int grid = (kDump && a.nsplit > 1) ? n_tiles * a.nsplit : n_tiles;
    hipLaunchKernelGGL((pair_tile_kernel<kDump, kSlots>), dim3(grid), dim3(256), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_split_entropy(const PairKernelArgs& a, int n_tiles, hipStream_t st) {
    if (n_tiles <= 0) return hipSuccess;
    const int ny = (a.R * a.R * a.nsteps + 255) / 256;
    hipLaunchKernelGGL(split_entropy_kernel, dim3(n_tiles, ny), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <bool kDump>
static hipError_t launch_d(const PairKernelArgs& a, int n_tiles, size_t lds, hipStream_t st) {
    // staging slots per thread: 2R DOFs x kGroups slots over 256 threads
    const int need = (2 * a.R * kGroups + 255) / 256;
    if (need <= 1) return launch_t<kDump, 1>(a, n_tiles, lds, st);
    if (need <= 2) return launch_t<kDump, 2>(a, n_tiles, lds, st);
    if (need <= 4) return launch_t<kDump, 4>(a, n_tiles, lds, st);
    return launch_t<kDump, 8>(a, n_tiles, lds, st);
}

template <int kSlots>
static hipError_t launch_p(const PairKernelArgs& a, int grid, size_t lds, hipStream_t st) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&pair_persist_kernel<kSlots>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((pair_persist_kernel<kSlots>), dim3(grid), dim3(256), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_pair_persistent(const PairKernelArgs& a, int grid, hipStream_t st) {
    if (a.n_work <= 0 || grid <= 0) return hipSuccess;
    const size_t lds = pair_tile_lds_bytes(a.R);
    const int need = (2 * a.R * kGroups + 255) / 256;
    if (need <= 1) return launch_p<1>(a, grid, lds, st);
    if (need <= 2) return launch_p<2>(a, grid, lds, st);
    if (need <= 4) return launch_p<4>(a, grid, lds, st);
    return launch_p<8>(a, grid, lds, st);
}

hipError_t launch_pair_tiles(const PairKernelArgs& a, int n_tiles, bool dump,
                             hipStream_t st) {
    if (n_tiles <= 0) return hipSuccess;
    const size_t lds = pair_tile_lds_bytes(a.R);
    return dump ? launch_d<true>(a, n_tiles, lds, st) : launch_d<false>(a, n_tiles, lds, st);
}
