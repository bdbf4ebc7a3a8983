// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/centre_hip.h"

struct centre_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int* d_flag = nullptr;          // device error word (code validation etc.)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_pair_ms = 0.f;       // duration of the last pair-kernel launch
    // Grow-only scratch buffers kept across calls (0: vmKDE float32 columns,
    // 1: vmKDE cos/sin pairs, 2: frame-split count dumps, 3: MST round
    // state).  Calls on one context run on its stream in order and
    // synchronise before returning, so the next call may reuse them; freed
    // by centre_destroy.
    // 4: count dumps of the grouped estimator pass (CS / JS, count output).
    // 5: the fused pair stage's temporaries (marginals, step logs, weights,
    // repacked codes, canonical outputs of explicit plans).
    void* scratch[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t scratch_bytes[6] = {0, 0, 0, 0, 0, 0};
    // device address of the library's static scratch array on this device
    // (the allocation-free pair-stage entry point carves it)
    void* noalloc = nullptr;
    // eight-wave pair tile operands (centre_set_mfma_format): 0 = FP4 where
    // the counts stay exact, 1 = int8 always
    int int8_only = 0;
};

struct centre_pair_plan {
    uint32_t n_dof = 0;
    uint8_t nbins = 0;
    int dofs_per_tile = 0;          // R: DOFs per tile side
    int n_tiles = 0;
    int2* d_tiles = nullptr;        // (row block, col block)
    int* d_tilemap = nullptr;       // [tile][R*R] canonical index or -1; null = all i<j
    uint64_t n_canon = 0;           // canonical (i<j) pairs computed
    uint64_t n_rec = 0;             // records requested by the caller
    uint64_t* d_rec2canon = nullptr;  // [n_rec] (explicit lists only)
    uint32_t* d_rec_ij = nullptr;     // [n_rec][2] as given (explicit lists only)
    bool all_pairs = true;
    bool tile_subset = false;         // all-pairs layout, listed tiles only
    int n_full = 0;                   // leading tiles with R DOFs on both sides
    // host copies for the grouped estimator pass (launch order)
    std::vector<int2> h_tiles;
    std::vector<int> h_tilemap;       // explicit lists only
    std::vector<uint32_t> h_rec_ij;   // explicit lists only
    std::vector<uint64_t> h_rec2canon;
    // Row-block grouping, built on first use by the grouped pass.  The
    // canonical (i < j, row-major) indices of the pairs whose row DOF lies in
    // row block I form the range [canon_off[I], canon_off[I + 1]); the
    // plan's tiles of row block I are byrow_tiles[tile_off[I] ..
    // tile_off[I + 1]); the records (explicit lists: record ids sorted by
    // canonical index; tile subsets: canonical indices of the listed pairs)
    // are items[item_off[I] .. item_off[I + 1]).
    mutable bool grouped = false;
    mutable std::vector<uint64_t> canon_off, item_off;
    mutable std::vector<int> tile_off;
    mutable int2* d_byrow_tiles = nullptr;
    mutable int* d_byrow_tilemap = nullptr;
    mutable uint64_t* d_items = nullptr;
};

// Bins of one-hot rows per tile side (a multiple of 32, the MFMA block).
constexpr int kTileRows = 256;
// Frames per LDS chunk (the K extent staged per barrier).
constexpr int kChunk = 64;
// Padded LDS row stride (bytes) of the one-hot tiles: 16 extra bytes spread
// the 16-byte fragment reads of one row block over all banks.
constexpr int kLdsRow = kChunk + 16;

// DOFs per tile side: 32 for nb <= 8, whose tiles give each DOF 8 one-hot
// rows; otherwise as many as fit 256 rows, capped at 60 so that the per-pair
// LDS slots still fit next to the one-hot buffers (160 KB LDS).
inline int dofs_per_tile(int nb) {
    if (nb <= 8) return 32;
    int r = kTileRows / nb;
    return r > 60 ? 60 : r;
}

// One system of a batched pair stage (device pointers; same nbins and steps
// for every system of a batch).
struct BatchSystem {
    const uint8_t* codes;
    uint64_t stride;
    int n_dof;
    uint64_t ss;
    uint64_t sstride;
    double* s2;
    double* mi;
    const double* dof_w;
    const double* dof_L;
    const double* dof_S1;
};

struct PairKernelArgs {
    const uint8_t* codes;   // [D] rows of `stride` bytes; frame f at +f
    uint64_t stride;
    int n_dof;
    int nb;
    int R;
    int nsteps;
    uint64_t ss;            // frames per convergence step
    uint64_t sstride;       // code-row offset between step starts (>= ss)
    const int2* tiles;
    const int* tilemap;     // [tile][R*R] canonical pair index or -1 of local (row, col)
                            // DOFs; null => all i<j, canonical index by triangle
    const double* dof_w;    // [D] bin width (1.0 under kde)
    const double* dof_L;    // [D][S] sum_b m_b(s) log J_b (0 if no Jacobian)
    const double* dof_S1;   // [D][S] S1 of the MI estimator, or null
    int slot_ml;            // output slot of ML in s2 (-1 = not requested)
    int slot_mm;            // output slot of MM (-1 = not requested)
    int nest;               // estimators in the s2 layout
    double* s2;             // [P][nest][S]
    double* mi;             // [P][S] or null
    int mi_from_mm;         // MI uses the MM S2 (when MM is the first estimator)
    uint32_t* counts;       // dump mode: [P][S][nb*nb] uint32 prefix counts
    uint64_t count_base;    // dump mode: pair p is written at record p - count_base
    // frame split (dump mode only, nsplit > 1): workgroup w counts tile
    // w % split_tiles over frames [k*split_len, k*split_len + len_k) of every
    // step, k = w / split_tiles (the last part takes the rest of the step),
    // into counts + k*split_cstride
    int nsplit;
    int split_tiles;
    uint64_t split_len;
    uint64_t split_cstride;
    int split_rec;          // words per (pair, step) count record; 0 = nb*nb
    int* flag;              // error word: bit0 = code >= nbins
    // batch (persistent launch): per-entry system of the tile list
    const BatchSystem* systems;   // null = single system in the fields above
    const int* tile_sys;          // [n_work] system of each tile list entry
    int* work;                    // shared tile counter (starts at 0)
    int n_work;                   // tile list entries
    int* sys_count;               // [n_sys] finished tiles (start at 0), or null
    const int* sys_tiles;         // [n_sys] tiles per system
    unsigned long long* sys_time; // [n_sys] completion stamp (100 MHz)
    unsigned long long* t0;       // launch stamp
    // eight-wave tile: c log c sums in units of 2^-kfix (pair_fixed_k of the
    // largest frame count a step of the launch reaches)
    int kfix;
    // eight-wave tile: 1 = int8 one-hot operands; 0 = FP4 operands when
    // pair_f4_exact holds for the launch's largest frame count, else int8
    int int8_only;
};
int pair_fixed_k(double n_max);
// FP4 one-hot counts accumulate in f32: exact while no count exceeds 2^24
bool pair_f4_exact(double n_max);


// Launchers (return hipError_t of the launch).
// Eight-wave kernel for tiles with R DOFs on both sides and step-aligned codes
// (no count dump); supported for R <= 8 (nb >= 29) when its LDS layout keeps
// the c log c table.
bool pair_tile8_small(int nb);
bool pair_tile8_supported(int R, int nb);
hipError_t launch_pair_tiles8(const PairKernelArgs& a, int n_tiles, hipStream_t st);
// Batch form (full tiles only, step-aligned codes): one workgroup per entry
// of a.tiles[0 .. a.n_work), dispatched in list order.
hipError_t launch_pair_batch8(const PairKernelArgs& a, int grid, hipStream_t st);
hipError_t launch_pair_tiles(const PairKernelArgs& a, int n_tiles, bool dump,
                             hipStream_t st);
// S2 (ML / MM slots) and MI of every pair of n_tiles tiles from the
// a.nsplit partial count dumps at a.counts (stride a.split_cstride), with the
// same formulas as the fused step epilogue
hipError_t launch_split_entropy(const PairKernelArgs& a, int n_tiles, hipStream_t st);
size_t pair_tile_lds_bytes(int R);
hipError_t launch_pair_persistent(const PairKernelArgs& a, int grid, hipStream_t st);
hipError_t launch_repack(const uint8_t* codes, uint64_t stride, uint32_t D, uint64_t ss, int S,
                         uint64_t ss4, uint8_t* out, hipStream_t st);

struct Dof1dArgs {
    const uint8_t* codes;   // [D] rows of `stride` bytes
    uint64_t stride;
    int n_dof;
    int nb;
    int nsteps;
    uint64_t ss;
    const uint8_t* types;   // [D] 1 bond, 2 angle, 3 torsion
    const double* ext;      // [D][4] (min, max, ...)
    int jacobian, kde;
    int ests[4];            // [nest] estimator ids
    int nest;
    double* s1;             // [D][nest][S]
    double* L;              // [D][S] sum_b m_b log J_b (for the pair epilogue)
    double* w;              // [D] bin width used by the pair epilogue
    uint64_t* counts;       // optional [D][S][nb] prefix counts
    int* flag;
};
hipError_t launch_dof1d(const Dof1dArgs& a, hipStream_t st);

// Every estimator of entropy2D (Estimators.cpp:391-651) and MI from counts
// materialised by the pair kernel's dump mode, for the records of one group
// of row blocks.  Record t of the launch:
//   mode 0: canonical index = record = item_first + t (all-pairs plans)
//   mode 1: canonical index = record = items[t] (tile subsets)
//   mode 2: record = items[t], canonical rec2canon[record], DOFs
//           rec_ij[record] (explicit lists; i > j reads the counts transposed)
struct CountEntropyArgs {
    const uint32_t* counts;   // [canonical - count_base][S][nb*nb], row DOF < column DOF
    uint64_t count_base;
    uint64_t n_item;
    int mode;
    uint64_t item_first;
    const uint64_t* items;
    const uint64_t* rec2canon;
    const uint32_t* rec_ij;
    uint32_t n_dof;
    int nb, nsteps;
    uint64_t ss;
    const uint8_t* types;
    const double* ext;
    int jacobian, kde;
    int ests[4];
    int nest;
    double* s2;               // [record][nest][S]
    const double* s1;         // [D][nest][S] marginals, or null (no MI)
    double* mi;               // [record][S] of estimator ests[0], or null
    uint32_t* counts_out;     // [record][S][nb*nb] in the record's own order, or null
};
hipError_t launch_count_entropy(const CountEntropyArgs& a, hipStream_t st);

// ---- discretizer -------------------------------------------------------------
struct DiscArgs {
    const double* bat;      // [F][ld] float64, DOF column j at bat[f*ld + j]
    uint64_t ld;
    uint32_t n_dof;
    uint64_t n_frame;
    int dtype;              // 1 bond, 2 angle, 3 torsion
    const uint8_t* schemes; // device [nsch]
    int nsch;
    int nbins_ref;
    int optimize;
    uint8_t* codes;         // [D][nsch][F]
    double* ext;            // [D][4]
    // scratch
    double* part;           // [nchunk][D][8] partial min/max/sum
    uint32_t* gaphist;      // [D][180]
    double* shift;          // [D]
    int* flag;
};
hipError_t run_discretize_hist(const DiscArgs& a, int nchunk, hipStream_t st);
int disc_chunks(uint64_t n_frame, uint32_t n_dof);

// ---- vmKDE -------------------------------------------------------------------
struct VmkdeArgs {
    const double* bat;      // [F][ld]
    uint64_t ld;
    uint32_t n_dof;
    uint64_t n_frame;
    double kappa;
    int nmaxconf;
    uint32_t sdo_step;
    double sdo_conv;
    uint32_t sdo_iter;
    int grid;               // grid intervals G (360 = reference)
    double* gridbuf;        // [D][2][G+1] gradient, second derivative
    double* minima;         // [D][9] radians
    uint8_t* nfound;        // [D]
    uint8_t* codes;         // [D][F]
    float* xt;              // scratch [D][F] float32 transposed columns
    double* trig;           // scratch [D][F][2] cos/sin of frames
};
hipError_t run_vmkde(const VmkdeArgs& a, hipStream_t st);
// Host FP64 helpers shared with the vmKDE launcher (Utils.cpp:60-168).
double centre_bessel_i(int n, double x);
double centre_vm_bandwidth(uint64_t n, double kappa);

// ---- MST -----------------------------------------------------------------------
struct MstArgs {
    const uint32_t* src;
    const uint32_t* dst;
    const double* w;
    uint64_t n_edge;
    uint32_t n_node;
    uint32_t* comp;         // [V]
    unsigned long long* best_w;   // [V] ordered weight key
    unsigned int* best_e;   // [V]
    uint8_t* in_tree;       // [E]
    int* changed;           // device word
    unsigned long long* part;  // [kMstGroups][V] per-workgroup maxima (V <= kMstLdsNodes)
};
// Boruvka's max pass reduces per workgroup in LDS up to this many nodes
// (8 B each), over kMstGroups workgroups.
constexpr uint32_t kMstLdsNodes = 16384;
constexpr int kMstGroups = 256;
size_t mst_part_bytes(uint32_t n_node);
hipError_t run_mst_round(const MstArgs& a, hipStream_t st);
hipError_t mst_collect(const MstArgs& a, uint32_t* idx, double* w, uint32_t cap,
                       hipStream_t st);
hipError_t mst_init(const MstArgs& a, hipStream_t st);
hipError_t mst_check(const MstArgs& a, hipStream_t st);   // sets *changed on a bad endpoint

// ---- reductions ------------------------------------------------------------------
hipError_t launch_segment_sum(const double* x, const uint64_t* d_beg, int n_seg,
                              double* d_out, hipStream_t st);
hipError_t launch_gather_rec(const double* src, const uint64_t* rec2canon,
                             uint64_t n_rec, int width, double* dst,
                             hipStream_t st);
hipError_t launch_copy_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                          size_t height, hipStream_t st);
// ---- learned pair screen (screen.hip) ----------------------------------------
constexpr int kScreenFeat = 8;     // features per pair
constexpr int kScreenHidden = 64;  // width of both hidden layers
// parameter vector (floats): W1[H][F], b1[H], W2[H][H], b2[H], w3[H], b3, mu[F], sd[F]
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kScreenHidden * kScreenFeat;
constexpr int kOffW2 = kOffB1 + kScreenHidden;
constexpr int kOffB2 = kOffW2 + kScreenHidden * kScreenHidden;
constexpr int kOffW3 = kOffB2 + kScreenHidden;
constexpr int kOffB3 = kOffW3 + kScreenHidden;
constexpr int kOffMu = kOffB3 + 1;
constexpr int kOffSd = kOffMu + kScreenFeat;
constexpr int kScreenParams = kOffSd + kScreenFeat;
// Bins of the subsample MI feature: the codes merged to min(nb, 8) bins.
constexpr int kScreenMiBins = 8;
hipError_t launch_subsample(const uint8_t* codes, uint64_t stride, uint32_t D, uint64_t step,
                            uint32_t n_sub, uint8_t* out, uint8_t* coarse, int nb, int nbc,
                            hipStream_t st);
hipError_t launch_dof_moments(const uint8_t* sub, uint32_t D, uint32_t n_sub, int nb, float* X,
                              float* Y, hipStream_t st);
hipError_t launch_corr(const float* X, const float* Y, uint32_t D, uint32_t n, const int2* blocks,
                       int n_blocks, float* lin, float* circ, hipStream_t st);
int corr_block_size();
hipError_t launch_features(uint32_t D, const double* s1, const uint8_t* types, const double* mi_sub,
                           const float* lin, const float* circ, float* feat, hipStream_t st);
hipError_t launch_mlp_forward(const float* params, const float* feat, uint64_t n, float* out,
                              hipStream_t st);
hipError_t launch_mlp_grad(const float* params, const float* feat, const float* target,
                           const uint32_t* idx, uint64_t n, float* grad, double* loss,
                           hipStream_t st);
hipError_t launch_tile_select(const float* pred, uint32_t D, int R, int k, float* tmax,
                              uint8_t* mark, hipStream_t st);
hipError_t launch_adam(float* params, const float* grad, float* m, float* v, float scale,
                       float lr, int t, hipStream_t st);
