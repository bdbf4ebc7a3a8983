/*
 * centre_hip.h -- C ABI of the MI355X (gfx950) entropy hot path of CENTRE.
 *
 * Plain pointers and sizes only.  Functions named centre_dev_* take DEVICE
 * pointers (memory the caller allocated on the context's GPU) and enqueue
 * work on the context's stream; functions without _dev take HOST pointers,
 * copy, run and copy back synchronously.
 *
 * Return value: 0 on success; -1 / -2 are the reference's size-mismatch
 * codes (frequency / entropy buffer, Estimators.cpp:64,138,228-231,403-406);
 * -3 = invalid argument; -4 = input code >= nbins; > 0 = a hipError_t.
 * centre_last_error() returns a message for the last failure on a context.
 *
 * Layouts (one bin scheme per call; call once per scheme for several):
 *   BAT columns   double  [n_frame][ld], DOF j of this type at column j
 *                         (the [frame][dof] hyperslab Netcdf_BAT.cpp:247
 *                         reads; rounded to float32 as nc_get_vara_float does)
 *   codes         uint8   [n_dof][code_stride] (frames contiguous)
 *   extrema       double  [n_dof][4] = {min, max, avg, phase}
 *   S1            double  [n_dof][nest][nsteps]
 *   S2            double  [n_pair][nest][nsteps]
 *   MI            double  [n_pair][nsteps]   (estimator ests[0])
 * DOF types: 1 bond, 2 angle, 3 torsion (BAT_t, Centre.h:42-47).
 * Estimators: 0 ML, 1 MM, 2 CS, 3 JS (EntropyEstimators, Centre.h:89-94).
 *
 * Threading: one host thread per context; calls on one context serialise
 * on its stream.  Contexts on the same GPU are independent.
 */
#ifndef CENTRE_HIP_H
#define CENTRE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct centre_ctx centre_ctx;
typedef struct centre_pair_plan centre_pair_plan;

/* Context: owns a HIP stream and small device scratch on `device`.
 * No reference counterpart: the reference calls host functions directly. */
int centre_init(int device, centre_ctx **out);
void centre_destroy(centre_ctx *ctx);
const char *centre_last_error(const centre_ctx *ctx);
int centre_sync(centre_ctx *ctx);
/* The hipStream_t the context enqueues on (for the caller's event timing). */
void *centre_stream(centre_ctx *ctx);
/* Milliseconds of the last pair-stage kernel launch (HIP events). */
float centre_last_pair_kernel_ms(const centre_ctx *ctx);
/* Operand format of the pair stage's one-hot MFMA tiles.  No reference
 * counterpart (the reference counts on the CPU).  0 (the default): FP4 (e2m1)
 * one-hot operands, two frames a byte, whenever the largest joint count a
 * call can reach (steps x frames a step) is at most 2^24, where the f32
 * accumulation of 0/1 products is exact; int8 operands beyond.  1: int8
 * operands always.  The counts, and so every result, are the same either
 * way.  Returns 0, or -3 for another value. */
int centre_set_mfma_format(centre_ctx *ctx, int format);

/* ---- discretizer ----------------------------------------------------------
 * Replaces Real2Int::bndReal2Int / angReal2Int / dihReal2Int
 * (Real2Int.cpp:27-73, 131-184, 244-479) as called per DOF by
 * Discretizer::run_dtype (Discretizer.cpp:138-216): float -> uint8 codes +
 * extrema.  dof_type applies to all n_dof columns (one BAT type per call, as
 * Discretizer::run calls run_dtype per type, Discretizer.cpp:314-361).
 * codes: [n_dof][n_schemes][n_frame]. */
int centre_dev_discretize_hist(centre_ctx *ctx, const double *d_bat,
                               uint64_t ld, uint32_t n_dof, uint64_t n_frame,
                               int dof_type, const uint8_t *schemes,
                               int n_schemes, int nbins_ref, int optimize,
                               uint8_t *d_codes, double *d_extrema);
int centre_discretize_hist(centre_ctx *ctx, const double *bat, uint64_t ld,
                           uint32_t n_dof, uint64_t n_frame, int dof_type,
                           const uint8_t *schemes, int n_schemes,
                           int nbins_ref, int optimize, uint8_t *codes,
                           double *extrema);

/* Replaces Real2Int::getMinimaPositions / vonMises and
 * Utils::modifiedBesselI* (Real2Int.cpp:481-698, Utils.cpp:60-168) for
 * torsions: von Mises KDE on a (grid_intervals+1)-point grid over 0..360
 * degrees (360 = the reference's 1-degree grid), maxima, steepest-descent
 * minima, and conformer-state codes [n_dof][n_frame].  minima: [n_dof][9]
 * radians (first n_found[d] valid); grid_out (optional, may be NULL):
 * [n_dof][2][grid_intervals+1] gradient and second derivative. */
int centre_dev_discretize_vmkde(centre_ctx *ctx, const double *d_bat,
                                uint64_t ld, uint32_t n_dof, uint64_t n_frame,
                                int nmaxconf, double kappa, uint32_t sdo_step,
                                double sdo_conv, uint32_t sdo_iter,
                                uint32_t grid_intervals, uint8_t *d_codes,
                                double *d_minima, uint8_t *d_nfound,
                                double *d_grid);
int centre_discretize_vmkde(centre_ctx *ctx, const double *bat, uint64_t ld,
                            uint32_t n_dof, uint64_t n_frame, int nmaxconf,
                            double kappa, uint32_t sdo_step, double sdo_conv,
                            uint32_t sdo_iter, uint32_t grid_intervals,
                            uint8_t *codes, double *minima, uint8_t *n_found,
                            double *grid);

/* ---- marginals -------------------------------------------------------------
 * Replaces binData1D + entropy1D (Estimators.cpp:56-118, 217-389) for every
 * DOF at once (EntropyCalculator::run1d, EntropyCalculator.cpp:6-279).
 * Frames used: nsteps * (n_frame / nsteps) from the start of each row.
 * counts (optional): [n_dof][nsteps][nbins] prefix counts (hist_*-1d). */
int centre_dev_dof_entropy(centre_ctx *ctx, const uint8_t *d_codes,
                           uint64_t code_stride, uint32_t n_dof,
                           uint64_t n_frame, const uint8_t *d_types,
                           const double *d_extrema, int nbins, int nsteps,
                           const int *ests, int nest, int jacobian, int kde,
                           double *d_s1, uint64_t *d_counts);

/* ---- pair stage --------------------------------------------------------------
 * A plan fixes the DOF count, bin count and the pair set: pairs == NULL
 * means every i < j (row-major upper-triangle order, n_pair =
 * n_dof*(n_dof-1)/2); otherwise an explicit list [n_pair][2] of DOF indices
 * (any order, i != j; the reference's neighbour-list pairs,
 * Centre.cpp:382-593).  Records keep the list order. */
int centre_pair_plan_create(centre_ctx *ctx, uint32_t n_dof, int nbins,
                            const uint32_t *pairs, uint64_t n_pair,
                            centre_pair_plan **out);
/* A plan over every i < j pair of a subset of DOF tiles: tiles[n_tiles][2]
 * are (row block, column block) indices with row <= column, a block being
 * centre_pair_dofs_per_tile(nbins) consecutive DOFs.  Outputs keep the all-
 * pairs triangle layout; entries of pairs outside the listed tiles are not
 * written.  No reference counterpart (used by the learned pair screen). */
int centre_pair_plan_from_tiles(centre_ctx *ctx, uint32_t n_dof, int nbins,
                                const int32_t *tiles, uint64_t n_tiles,
                                centre_pair_plan **out);
int centre_pair_dofs_per_tile(int nbins);
void centre_pair_plan_destroy(centre_pair_plan *plan);
uint64_t centre_pair_plan_records(const centre_pair_plan *plan);
int centre_pair_plan_tiles(const centre_pair_plan *plan);

/* Replaces binData2D + entropy2D (Estimators.cpp:130-205, 391-651) over all
 * pairs of the plan (EntropyCalculator::run2d_xx / run2d_xy,
 * EntropyCalculator.cpp:281-1088), plus MI = S1_i + S1_j - S2_ij
 * (getMaxMI2D, Convergance.cpp:183-194).  Joint counts are exact MFMA
 * contractions of one-hot codes (FP4 operands with f32 counts while counts
 * stay <= 2^24, else int8 with int32 counts; see centre_set_mfma_format);
 * ML and MM are fused into the count pass.
 * With CS or JS among the estimators, or with d_counts, the counts of groups
 * of row blocks (at most 2 GiB at a time, or one row block) are materialised
 * and every estimator and MI is evaluated from them on the device, each
 * record in its own row/column order.  The marginals of the same
 * frames are computed inside the call; d_s1 (optional) receives them as
 * [n_dof][nest][nsteps].  d_counts (optional, NULL = never materialised):
 * [record][nsteps][nbins*nbins] uint32 prefix counts, cell x*nbins + y for
 * bin x of the record's first DOF and bin y of its second (the binData2D
 * layout behind hist_*-2d, Estimators.cpp:130-205). */
int centre_dev_pair_entropy(centre_ctx *ctx, const centre_pair_plan *plan,
                            const uint8_t *d_codes, uint64_t code_stride,
                            uint64_t n_frame, const uint8_t *d_types,
                            const double *d_extrema, int nsteps,
                            const int *ests, int nest, int jacobian, int kde,
                            double *d_s1, double *d_s2, double *d_mi,
                            uint32_t *d_counts);

/* Several bin schemes in one call, as entropy2D loops its bin_schemes
 * (Estimators.cpp:408-647) and binData2D bins each scheme's codes
 * (Estimators.cpp:130-205): plans[s] is the plan of scheme s (the same DOFs
 * and pair list for every scheme, nbins = that scheme's bin count); codes
 * [n_dof][n_schemes][code_stride] (the coord layout of bin_*.nc,
 * Netcdf_TrjInt.cpp:542-755).  Outputs in the reference's contri_S layout
 * (Netcdf_EntContri.cpp:55-73): S1 [n_dof][nest][n_schemes][nsteps] (optional),
 * S2 [record][nest][n_schemes][nsteps], MI [record][n_schemes][nsteps]
 * (optional, estimator ests[0]). */
int centre_dev_pair_entropy_schemes(centre_ctx *ctx,
                                    const centre_pair_plan *const *plans,
                                    int n_schemes, const uint8_t *d_codes,
                                    uint64_t code_stride, uint64_t n_frame,
                                    const uint8_t *d_types,
                                    const double *d_extrema, int nsteps,
                                    const int *ests, int nest, int jacobian,
                                    int kde, double *d_s1, double *d_s2,
                                    double *d_mi);

/* Bytes of the static device array behind centre_dev_pair_entropy_noalloc
 * (one per GPU the library runs on, held until the process ends). */
#define CENTRE_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): centre_dev_pair_entropy for the ML and MM estimators
 * without count output, calling no allocator.  Its temporaries (marginals,
 * per-DOF widths and Jacobian sums, a step-aligned copy of misaligned codes,
 * canonical-order outputs of explicit pair lists, frame-split count dumps)
 * live in one static device array of CENTRE_NOALLOC_SCRATCH_BYTES per GPU;
 * the plan is made beforehand (centre_pair_plan_create allocates).  Returns
 * -3 when this call's temporaries exceed the array (the frame split is then
 * left out first; it is used only where its dumps fit).  Every call on one
 * GPU shares the array and leaves its contents behind: calls that overlap on
 * one GPU (several contexts or host threads) race. */
int centre_dev_pair_entropy_noalloc(centre_ctx *ctx,
                                    const centre_pair_plan *plan,
                                    const uint8_t *d_codes,
                                    uint64_t code_stride, uint64_t n_frame,
                                    const uint8_t *d_types,
                                    const double *d_extrema, int nsteps,
                                    const int *ests, int nest, int jacobian,
                                    int kde, double *d_s1, double *d_s2,
                                    double *d_mi);

/* Host-buffer convenience: discretized codes in, S1/S2/MI out. */
int centre_pair_entropy(centre_ctx *ctx, const uint8_t *codes,
                        uint32_t n_dof, uint64_t n_frame,
                        const uint8_t *types, const double *extrema,
                        int nbins, int nsteps, const uint32_t *pairs,
                        uint64_t n_pair, const int *ests, int nest,
                        int jacobian, int kde, double *s1, double *s2,
                        double *mi);

/* ---- MIST --------------------------------------------------------------------
 * Replaces boruvkaMST (Centre.cpp:5-143) as called by EntropyScorer::run
 * (EntropyScorer.cpp:632-741): maximum spanning forest under the total order
 * (weight descending, edge index ascending).  tree_edge_idx receives the
 * forest's edge indices sorted by that order; *n_tree their count.  Edges
 * whose weight is NaN are left out (no counterpart: the reference has no
 * missing pairs; a screened pass leaves the pairs it skipped NaN, so its MI
 * array in triangle order is the edge list as it stands). */
int centre_dev_mst(centre_ctx *ctx, const uint32_t *d_src,
                   const uint32_t *d_dst, const double *d_w, uint64_t n_edge,
                   uint32_t n_node, uint32_t *tree_edge_idx,
                   uint32_t *n_tree);
int centre_mst(centre_ctx *ctx, const uint32_t *src, const uint32_t *dst,
               const double *w, uint64_t n_edge, uint32_t n_node,
               uint32_t *tree_edge_idx, uint32_t *n_tree);

/* Sum of each of n_seg contiguous segments of x (FP64), into out[n_seg]:
 * the MIE totals (sumContrib1D/2D, Convergance.cpp:12-43).  seg_begin is a
 * HOST array of n_seg + 1 ascending offsets into d_x. */
int centre_dev_segment_sum(centre_ctx *ctx, const double *d_x,
                           const uint64_t *seg_begin, int n_seg,
                           double *out);

/* ---- device memory plumbing --------------------------------------------------
 * No reference counterpart: allocation and copies for callers that have no
 * device allocator of their own (tests, the benchmark, the C++ host).
 * Copies are synchronous on the context's stream. */
int centre_dev_alloc(centre_ctx *ctx, size_t bytes, void **out);
int centre_dev_free(centre_ctx *ctx, void *d_ptr);
int centre_memcpy_h2d(centre_ctx *ctx, void *d_dst, const void *src,
                      size_t bytes);
int centre_memcpy_d2h(centre_ctx *ctx, void *dst, const void *d_src,
                      size_t bytes);
int centre_memcpy_d2d(centre_ctx *ctx, void *d_dst, const void *d_src,
                      size_t bytes);
/* Pitched device copy: height rows of width bytes (e.g. one step's MI
 * column of [pair][step] as MST edge weights). */
int centre_memcpy_d2d_2d(centre_ctx *ctx, void *d_dst, size_t dpitch,
                         const void *d_src, size_t spitch, size_t width,
                         size_t height);
/* hipDeviceSynchronize on the context's device. */
int centre_device_sync(centre_ctx *ctx);

/* ---- batched pair stage ------------------------------------------------------
 * No reference counterpart (centre_OMP runs one system per process): many
 * independent systems, each with its own DOF count and frames, share one
 * persistent launch.  Every system gets all its i < j pairs (triangle order)
 * and its marginals, exactly as centre_dev_pair_entropy with an all-pairs
 * plan would give them; ML and MM only.
 * One persistent launch runs every tile of every system from one work
 * queue (one workgroup per CU taking tiles from a shared counter).
 * order: tile order of the queue -- 1 longest tiles first, 2 system by
 * system, 3 systems by total work, largest first.
 * tile_dofs: DOFs a tile takes per side (its shape: tile_dofs * nbins of the
 * 256 one-hot rows), 1 .. centre_pair_dofs_per_tile(nbins) (exactly that at
 * nbins <= 8).  0 in either knob = autotune it: every candidate (the three
 * orders; the largest shape and one DOF fewer) runs on a calibration share
 * of the batch (2% of each system's tiles, steps cut to 1,024 frames) and
 * the fastest is run.  The frame chunk (64 frames per pipeline stage) is
 * fixed by the tile's LDS budget.
 * sys_ms[k] (optional): time from the launch to the completion of system k;
 * total_ms: the whole launch; order_used / tile_dofs_used: what was run. */
typedef struct centre_batch_system {
    const uint8_t *d_codes;    /* [n_dof][code_stride] */
    uint64_t code_stride;
    uint32_t n_dof;
    uint64_t n_frame;
    const uint8_t *d_types;    /* [n_dof] 1 bond, 2 angle, 3 torsion */
    const double *d_extrema;   /* [n_dof][4] */
    double *d_s1;              /* [n_dof][nest][nsteps] or NULL */
    double *d_s2;              /* [P][nest][nsteps] */
    double *d_mi;              /* [P][nsteps] or NULL */
} centre_batch_system;
int centre_batch_pair_entropy(centre_ctx *ctx, const centre_batch_system *sys,
                              int n_sys, int nbins, int nsteps,
                              const int *ests, int nest, int jacobian,
                              int kde, int order, int tile_dofs,
                              double *sys_ms, double *total_ms,
                              int *order_used, int *tile_dofs_used);

/* ---- learned pair screen ------------------------------------------------------
 * No reference counterpart (the reference restricts pairs only by a user
 * neighbour list, Centre.cpp:459-593).  An MLP 8 -> 64 -> 64 -> 1 (FP32, ReLU)
 * predicts log(MI) of a pair from 8 features; callers skip pairs whose
 * predicted MI is below a cutoff.  Parameter vector (centre_screen_n_params()
 * floats): W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3, mu[8], sd[8]
 * (inputs are standardised as (f - mu) / sd; training leaves mu, sd alone).
 *
 * centre_screen_features: d_feat[P][8] (device, FP32) for all P = D(D-1)/2
 * pairs i < j in triangle order: S1_i, S1_j (ML, all n_frame frames),
 * circular and linear correlation on n_sub evenly strided frames (n_sub a
 * multiple of 4), ML MI on the same frames with the bins merged to
 * min(nbins, 8) (code * nbc / nbins), type_i / 3, type_j / 3, same-type
 * flag.
 * centre_mlp_forward: d_out[n] = predicted log(MI).
 * centre_mlp_loss_grad: d_grad = gradient of sum 0.5 (y - t)^2 over the n
 * samples (FP32, accumulated with atomics), *loss = that sum.
 * centre_mlp_train: Adam (beta 0.9 / 0.999), minibatches of `batch` samples
 * in a seeded random order, `epochs` passes; epoch_loss[e] = mean loss. */
int centre_screen_n_params(void);
/* Tiles of the pair stage to run after screening: every tile (row block <=
 * column block, blocks of centre_pair_dofs_per_tile(nbins) DOFs) holding a
 * pair whose predicted log MI (d_pred, triangle order of all pairs) reaches
 * log(cutoff), plus the tiles holding each DOF's topk (<= 4) best predicted
 * partners.  tiles[] needs room for nblk*(nblk+1)/2 pairs of int32. */
int centre_screen_tiles(centre_ctx *ctx, const float *d_pred, uint32_t n_dof,
                        int nbins, float cutoff, int topk, int32_t *tiles,
                        uint64_t *n_tiles);
int centre_screen_features(centre_ctx *ctx, const uint8_t *d_codes,
                           uint64_t code_stride, uint32_t n_dof,
                           uint64_t n_frame, const uint8_t *d_types,
                           const double *d_extrema, int nbins, uint32_t n_sub,
                           float *d_feat);
int centre_mlp_forward(centre_ctx *ctx, const float *d_params,
                       const float *d_feat, uint64_t n, float *d_out);
int centre_mlp_loss_grad(centre_ctx *ctx, const float *d_params,
                         const float *d_feat, const float *d_target,
                         uint64_t n, float *d_grad, double *loss);
int centre_mlp_train(centre_ctx *ctx, float *d_params, const float *d_feat,
                     const float *d_target, uint64_t n, int epochs, int batch,
                     float lr, uint64_t seed, double *epoch_loss);

/* ---- synthetic input ---------------------------------------------------------
 * No reference counterpart (test and benchmark input).  Fills out[n_frame]
 * [n_bond + n_angle + n_dih] (host, float64, [frame][dof]) with a seeded
 * BAT trajectory: Gaussian bonds and angles and 3-component von Mises
 * mixture torsions, all coupled through one shared 4-state Markov latent.
 * The bytes depend only on the seed and sizes, not on n_threads. */
int centre_synth_bat(uint64_t seed, uint32_t n_bond, uint32_t n_angle,
                     uint32_t n_dih, uint64_t n_frame, double *out,
                     int n_threads);

#ifdef __cplusplus
}
#endif
#endif /* CENTRE_HIP_H */
