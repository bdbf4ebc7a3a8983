// extern "C" layer: argument checking, device scratch, plans, launches.
#include <algorithm>
#include <mutex>
#include <numeric>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "centre_internal.h"

namespace {

int fail(centre_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHK(x)                                                            \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess)                                                \
            return fail(ctx, (int)e_, std::string(#x) + ": " +               \
                                          hipGetErrorString(e_));            \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 16); }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Scratch slot `slot` of the context with at least `bytes` bytes (grown, never
// shrunk; the old buffer is freed first, which waits for the device).
hipError_t ctx_scratch(centre_ctx* ctx, int slot, size_t bytes, void** out) {
    if (ctx->scratch_bytes[slot] < bytes || !ctx->scratch[slot]) {
        if (ctx->scratch[slot]) {
            hipError_t e = hipFree(ctx->scratch[slot]);
            ctx->scratch[slot] = nullptr;
            ctx->scratch_bytes[slot] = 0;
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipMalloc(&ctx->scratch[slot], bytes ? bytes : 16);
        if (e != hipSuccess) {
            ctx->scratch[slot] = nullptr;
            return e;
        }
        ctx->scratch_bytes[slot] = bytes;
    }
    *out = ctx->scratch[slot];
    return hipSuccess;
}

// Scratch of the allocation-free pair stage (centre_dev_pair_entropy_noalloc):
// one static array per GPU, which the loader places with the code object.
// Host code never uses its name as an address: centre_init resolves the
// device address once per device (hipGetSymbolAddress, with that device
// current) and the entry point passes that address to the kernels.
constexpr size_t kNoAllocBytes = CENTRE_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current).
hipError_t noalloc_address(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_noalloc_ptr[device] = p;
    }
    *out = g_noalloc_ptr[device];
    return hipSuccess;
}

int check_flag(centre_ctx* ctx) {
    int f = 0;
    HIPCHK(hipMemcpyAsync(&f, ctx->d_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (f & 1) return fail(ctx, -4, "input code >= nbins");
    if (f & 2) return fail(ctx, -4, "torsion value outside the gap histogram range");
    return 0;
}

int check_ests(centre_ctx* ctx, const int* ests, int nest) {
    if (!ests || nest <= 0 || nest > 4) return fail(ctx, -3, "need 1..4 estimators");
    for (int e = 0; e < nest; ++e)
        if (ests[e] < 0 || ests[e] > 3) return fail(ctx, -3, "unknown estimator id");
    return 0;
}

}  // namespace

extern "C" {

int centre_init(int device, centre_ctx** out) {
    if (!out) return -3;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return (int)e;
    if (device < 0 || device >= n) return -3;
    e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    centre_ctx* ctx = new centre_ctx();
    ctx->device = device;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&ctx->d_flag, sizeof(int))) != hipSuccess ||
        (e = hipEventCreate(&ctx->ev0)) != hipSuccess ||
        (e = hipEventCreate(&ctx->ev1)) != hipSuccess ||
        (e = noalloc_address(device, &ctx->noalloc)) != hipSuccess) {
        centre_destroy(ctx);
        return (int)e;
    }
    *out = ctx;
    return 0;
}

void centre_destroy(centre_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : ctx->scratch)
        if (p) (void)hipFree(p);
    if (ctx->d_flag) (void)hipFree(ctx->d_flag);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* centre_last_error(const centre_ctx* ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

int centre_sync(centre_ctx* ctx) {
    if (!ctx) return -3;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

void* centre_stream(centre_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int centre_set_mfma_format(centre_ctx* ctx, int format) {
    if (!ctx) return -3;
    if (format != 0 && format != 1) return fail(ctx, -3, "mfma format must be 0 (FP4 where exact) or 1 (int8)");
    ctx->int8_only = format;
    return 0;
}

float centre_last_pair_kernel_ms(const centre_ctx* ctx) {
    return ctx ? ctx->last_pair_ms : 0.f;
}

// ---------------------------------------------------------------------------
int centre_dev_discretize_hist(centre_ctx* ctx, const double* d_bat, uint64_t ld,
                               uint32_t n_dof, uint64_t n_frame, int dof_type,
                               const uint8_t* schemes, int n_schemes,
                               int nbins_ref, int optimize, uint8_t* d_codes,
                               double* d_extrema) {
    if (!ctx) return -3;
    if (!d_bat || !d_codes || !d_extrema || !schemes || n_schemes <= 0)
        return fail(ctx, -3, "null buffer or no bin scheme");
    if (dof_type < 1 || dof_type > 3) return fail(ctx, -3, "dof_type must be 1, 2 or 3");
    if (ld < n_dof || n_frame == 0) return fail(ctx, -3, "ld < n_dof or no frames");
    for (int s = 0; s < n_schemes; ++s)
        if (schemes[s] == 0) return fail(ctx, -3, "zero bins in a scheme");
    if (dof_type == 3 && optimize && nbins_ref <= 0)
        return fail(ctx, -3, "nbins_ref required for torsion gap optimisation");
    if (n_dof == 0) return 0;
    HIPCHK(hipSetDevice(ctx->device));
    const int nchunk = disc_chunks(n_frame, n_dof);
    DevBuf sch, part, gh, shift;
    HIPCHK(sch.alloc(n_schemes));
    HIPCHK(part.alloc(sizeof(double) * 8 * (size_t)nchunk * n_dof));
    HIPCHK(gh.alloc(sizeof(uint32_t) * 180 * (size_t)n_dof));
    HIPCHK(shift.alloc(sizeof(double) * n_dof));
    HIPCHK(hipMemcpyAsync(sch.p, schemes, n_schemes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_flag, 0, sizeof(int), ctx->stream));
    DiscArgs a{};
    a.bat = d_bat;
    a.ld = ld;
    a.n_dof = n_dof;
    a.n_frame = n_frame;
    a.dtype = dof_type;
    a.schemes = sch.as<uint8_t>();
    a.nsch = n_schemes;
    a.nbins_ref = nbins_ref;
    a.optimize = optimize;
    a.codes = d_codes;
    a.ext = d_extrema;
    a.part = part.as<double>();
    a.gaphist = gh.as<uint32_t>();
    a.shift = shift.as<double>();
    a.flag = ctx->d_flag;
    HIPCHK(run_discretize_hist(a, nchunk, ctx->stream));
    return check_flag(ctx);
}

int centre_discretize_hist(centre_ctx* ctx, const double* bat, uint64_t ld,
                           uint32_t n_dof, uint64_t n_frame, int dof_type,
                           const uint8_t* schemes, int n_schemes, int nbins_ref,
                           int optimize, uint8_t* codes, double* extrema) {
    if (!ctx) return -3;
    if (!bat || !codes || !extrema) return fail(ctx, -3, "null buffer");
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf b, c, e;
    HIPCHK(b.alloc(sizeof(double) * ld * n_frame));
    HIPCHK(c.alloc((size_t)n_dof * n_schemes * n_frame));
    HIPCHK(e.alloc(sizeof(double) * 4 * n_dof));
    HIPCHK(hipMemcpyAsync(b.p, bat, sizeof(double) * ld * n_frame, hipMemcpyHostToDevice, ctx->stream));
    int r = centre_dev_discretize_hist(ctx, b.as<double>(), ld, n_dof, n_frame, dof_type,
                                       schemes, n_schemes, nbins_ref, optimize,
                                       c.as<uint8_t>(), e.as<double>());
    if (r) return r;
    HIPCHK(hipMemcpyAsync(codes, c.p, (size_t)n_dof * n_schemes * n_frame, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(extrema, e.p, sizeof(double) * 4 * n_dof, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---------------------------------------------------------------------------
namespace {
int run_dof1d(centre_ctx* ctx, const uint8_t* d_codes, uint64_t stride,
              uint32_t n_dof, uint64_t n_frame, const uint8_t* d_types,
              const double* d_ext, int nbins, int nsteps, const int* ests, int nest,
              int jacobian, int kde, double* d_s1, double* d_L, double* d_w,
              uint64_t* d_counts) {
    Dof1dArgs a{};
    a.codes = d_codes;
    a.stride = stride;
    a.n_dof = (int)n_dof;
    a.nb = nbins;
    a.nsteps = nsteps;
    a.ss = n_frame / nsteps;
    a.types = d_types;
    a.ext = d_ext;
    a.jacobian = jacobian;
    a.kde = kde;
    for (int e = 0; e < nest; ++e) a.ests[e] = ests[e];   // nest <= 4 (check_ests)
    a.nest = nest;
    a.s1 = d_s1;
    a.L = d_L;
    a.w = d_w;
    a.counts = d_counts;
    a.flag = ctx->d_flag;
    HIPCHK(launch_dof1d(a, ctx->stream));
    return 0;
}
}  // namespace

int centre_dev_dof_entropy(centre_ctx* ctx, const uint8_t* d_codes,
                           uint64_t code_stride, uint32_t n_dof, uint64_t n_frame,
                           const uint8_t* d_types, const double* d_extrema,
                           int nbins, int nsteps, const int* ests, int nest,
                           int jacobian, int kde, double* d_s1, uint64_t* d_counts) {
    if (!ctx) return -3;
    if (int r = check_ests(ctx, ests, nest)) return r;
    if (!d_codes || !d_types || !d_extrema || !d_s1) return fail(ctx, -3, "null buffer");
    if (nbins < 1 || nbins > 255 || nsteps < 1 || nsteps > 255)
        return fail(ctx, -3, "nbins and nsteps must be in 1..255");
    if (n_frame / nsteps == 0) return fail(ctx, -3, "fewer frames than steps");
    if (code_stride < n_frame) return fail(ctx, -3, "code_stride < n_frame");
    if (n_dof == 0) return 0;
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf L, w;
    HIPCHK(L.alloc(sizeof(double) * n_dof * nsteps));
    HIPCHK(w.alloc(sizeof(double) * n_dof));
    HIPCHK(hipMemsetAsync(ctx->d_flag, 0, sizeof(int), ctx->stream));
    if (int r = run_dof1d(ctx, d_codes, code_stride, n_dof, n_frame, d_types, d_extrema,
                          nbins, nsteps, ests, nest, jacobian, kde, d_s1,
                          L.as<double>(), w.as<double>(), d_counts))
        return r;
    return check_flag(ctx);
}

// ---------------------------------------------------------------------------
// Tiles with a full block of R DOFs on both sides go first: the eight-wave
// kernel takes them; tiles of the last, partial block follow.  Returns how
// many lead.
// Launch order of the first n tiles for the L2s of the 8 XCDs, which take
// workgroups round-robin (workgroup w runs on XCD w % 8).  The tiles are put
// in 4 x 8 blocks of (row block, column block), the blocked list is cut into
// 8 contiguous runs, and run x goes to positions x, x + 8, x + 16, ...: the
// ~32 tiles an XCD has in flight then share 4 row and 8 column code blocks
// instead of one row block and 32 column blocks.  Results do not depend on
// the order.
static void xcd_blocked_order(std::vector<int2>& t, size_t n) {
    constexpr size_t kXcd = 8;
    constexpr int kBI = 4, kBJ = 8;
    if (n < 2 * kXcd) return;
    std::stable_sort(t.begin(), t.begin() + n, [](const int2& a, const int2& b) {
        if (a.x / kBI != b.x / kBI) return a.x / kBI < b.x / kBI;
        if (a.y / kBJ != b.y / kBJ) return a.y / kBJ < b.y / kBJ;
        return a.x != b.x ? a.x < b.x : a.y < b.y;
    });
    std::vector<int2> out(n);
    const size_t per = n / kXcd, rem = n % kXcd;
    for (size_t x = 0; x < kXcd; ++x) {
        const size_t start = x * per + std::min(x, rem), cnt = per + (x < rem ? 1 : 0);
        for (size_t q = 0; q < cnt; ++q) out[q * kXcd + x] = t[start + q];
    }
    std::copy(out.begin(), out.end(), t.begin());
}

static int full_tiles_first(std::vector<int2>& tiles, uint32_t n_dof, int R) {
    const int nfull = (int)(n_dof / (uint32_t)R);   // blocks 0..nfull-1 are full
    auto full = [&](const int2& t) { return t.x < nfull && t.y < nfull; };
    auto mid = std::stable_partition(tiles.begin(), tiles.end(), full);
    const int n = (int)(mid - tiles.begin());
    xcd_blocked_order(tiles, (size_t)n);
    return n;
}

int centre_pair_dofs_per_tile(int nbins) {
    return (nbins < 1 || nbins > 255) ? -3 : dofs_per_tile(nbins);
}

int centre_pair_plan_from_tiles(centre_ctx* ctx, uint32_t n_dof, int nbins, const int32_t* tiles,
                                uint64_t n_tiles, centre_pair_plan** out) {
    if (!ctx || !out) return -3;
    *out = nullptr;
    if (nbins < 1 || nbins > 255) return fail(ctx, -3, "nbins must be in 1..255");
    if (n_dof < 2) return fail(ctx, -3, "need at least two DOFs");
    if (n_tiles && !tiles) return fail(ctx, -3, "null tile list");
    const int R = dofs_per_tile(nbins);
    const int64_t nblk = (n_dof + R - 1) / R;
    std::vector<int2> tl(n_tiles);
    for (uint64_t t = 0; t < n_tiles; ++t) {
        const int32_t I = tiles[2 * t], J = tiles[2 * t + 1];
        if (I < 0 || J < I || J >= nblk) return fail(ctx, -3, "tile index out of range or row > column");
        tl[t] = make_int2(I, J);
    }
    HIPCHK(hipSetDevice(ctx->device));
    auto* plan = new centre_pair_plan();
    plan->n_dof = n_dof;
    plan->nbins = (uint8_t)nbins;
    plan->dofs_per_tile = R;
    plan->all_pairs = true;
    plan->tile_subset = true;
    plan->n_canon = (uint64_t)n_dof * (n_dof - 1) / 2;
    plan->n_rec = plan->n_canon;
    plan->n_full = full_tiles_first(tl, n_dof, R);
    plan->n_tiles = (int)tl.size();
    plan->h_tiles = tl;
    hipError_t e = hipMalloc(&plan->d_tiles, sizeof(int2) * (tl.size() ? tl.size() : 1));
    if (e == hipSuccess && !tl.empty())
        e = hipMemcpy(plan->d_tiles, tl.data(), sizeof(int2) * tl.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        centre_pair_plan_destroy(plan);
        return fail(ctx, (int)e, std::string("plan tiles: ") + hipGetErrorString(e));
    }
    *out = plan;
    return 0;
}

int centre_pair_plan_create(centre_ctx* ctx, uint32_t n_dof, int nbins,
                            const uint32_t* pairs, uint64_t n_pair,
                            centre_pair_plan** out) {
    if (!ctx || !out) return -3;
    *out = nullptr;
    if (nbins < 1 || nbins > 255) return fail(ctx, -3, "nbins must be in 1..255");
    if (n_dof < 2) return fail(ctx, -3, "need at least two DOFs");
    HIPCHK(hipSetDevice(ctx->device));
    auto* plan = new centre_pair_plan();
    plan->n_dof = n_dof;
    plan->nbins = (uint8_t)nbins;
    const int R = dofs_per_tile(nbins);
    plan->dofs_per_tile = R;
    const int nblk = (int)((n_dof + R - 1) / R);
    std::vector<int2> tiles;
    if (!pairs) {
        plan->all_pairs = true;
        plan->n_canon = (uint64_t)n_dof * (n_dof - 1) / 2;
        plan->n_rec = plan->n_canon;
        for (int I = 0; I < nblk; ++I)
            for (int J = I; J < nblk; ++J) tiles.push_back(make_int2(I, J));
    } else {
        plan->all_pairs = false;
        plan->n_rec = n_pair;
        std::vector<uint64_t> key(n_pair);
        for (uint64_t p = 0; p < n_pair; ++p) {
            uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
            if (i >= n_dof || j >= n_dof || i == j) {
                delete plan;
                return fail(ctx, -3, "pair index out of range or i == j");
            }
            if (i > j) std::swap(i, j);
            key[p] = ((uint64_t)i << 32) | j;
        }
        std::vector<uint64_t> uniq(key);
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        plan->n_canon = uniq.size();
        // canonical index = position in the sorted unique list
        std::vector<uint64_t> r2c(n_pair);
        for (uint64_t p = 0; p < n_pair; ++p)
            r2c[p] = (uint64_t)(std::lower_bound(uniq.begin(), uniq.end(), key[p]) - uniq.begin());
        // tiles holding at least one pair (block pairs in row-major order)
        std::vector<uint64_t> tkey;
        tkey.reserve(uniq.size());
        for (uint64_t k : uniq)
            tkey.push_back(((uint64_t)((uint32_t)(k >> 32) / R) << 32) | ((uint32_t)k / R));
        std::sort(tkey.begin(), tkey.end());
        tkey.erase(std::unique(tkey.begin(), tkey.end()), tkey.end());
        for (uint64_t k : tkey) tiles.push_back(make_int2((int)(k >> 32), (int)(uint32_t)k));
        plan->n_full = full_tiles_first(tiles, n_dof, R);
        // per-tile map of local (row DOF, column DOF) to canonical index, in
        // the launch order of the tiles
        std::vector<uint64_t> tord(tiles.size());
        for (size_t q = 0; q < tiles.size(); ++q)
            tord[q] = ((uint64_t)tiles[q].x << 32) | (uint32_t)tiles[q].y;
        std::vector<size_t> tpos(tiles.size());
        std::iota(tpos.begin(), tpos.end(), (size_t)0);
        std::sort(tpos.begin(), tpos.end(), [&](size_t x, size_t y) { return tord[x] < tord[y]; });
        std::vector<int> tmap(tiles.size() * (size_t)R * R, -1);
        for (size_t c = 0; c < uniq.size(); ++c) {
            const uint32_t i = (uint32_t)(uniq[c] >> 32), j = (uint32_t)uniq[c];
            const uint64_t tk = ((uint64_t)(i / R) << 32) | (j / R);
            const size_t q = tpos[std::lower_bound(tpos.begin(), tpos.end(), tk,
                                                   [&](size_t x, uint64_t v) { return tord[x] < v; }) -
                                  tpos.begin()];
            tmap[q * R * R + (i % R) * R + (j % R)] = (int)c;
        }
        plan->h_rec_ij.assign(pairs, pairs + 2 * n_pair);
        plan->h_rec2canon = r2c;
        hipError_t e;
        if ((e = hipMalloc(&plan->d_tilemap, sizeof(int) * (tmap.size() ? tmap.size() : 1))) != hipSuccess ||
            (!tmap.empty() && (e = hipMemcpy(plan->d_tilemap, tmap.data(), sizeof(int) * tmap.size(),
                                             hipMemcpyHostToDevice)) != hipSuccess) ||
            (e = hipMalloc(&plan->d_rec2canon, sizeof(uint64_t) * (n_pair ? n_pair : 1))) != hipSuccess ||
            (n_pair && (e = hipMemcpy(plan->d_rec2canon, r2c.data(), sizeof(uint64_t) * n_pair,
                                      hipMemcpyHostToDevice)) != hipSuccess) ||
            (e = hipMalloc(&plan->d_rec_ij, sizeof(uint32_t) * 2 * (n_pair ? n_pair : 1))) != hipSuccess ||
            (n_pair && (e = hipMemcpy(plan->d_rec_ij, pairs, sizeof(uint32_t) * 2 * n_pair,
                                      hipMemcpyHostToDevice)) != hipSuccess)) {
            centre_pair_plan_destroy(plan);
            return fail(ctx, (int)e, std::string("plan upload: ") + hipGetErrorString(e));
        }
        plan->h_tilemap = std::move(tmap);
    }
    if (!pairs) plan->n_full = full_tiles_first(tiles, n_dof, R);
    plan->n_tiles = (int)tiles.size();
    plan->h_tiles = tiles;
    hipError_t e = hipMalloc(&plan->d_tiles, sizeof(int2) * (tiles.size() ? tiles.size() : 1));
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpy(plan->d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        centre_pair_plan_destroy(plan);
        return fail(ctx, (int)e, std::string("plan tiles: ") + hipGetErrorString(e));
    }
    *out = plan;
    return 0;
}

void centre_pair_plan_destroy(centre_pair_plan* plan) {
    if (!plan) return;
    if (plan->d_tiles) (void)hipFree(plan->d_tiles);
    if (plan->d_tilemap) (void)hipFree(plan->d_tilemap);
    if (plan->d_rec2canon) (void)hipFree(plan->d_rec2canon);
    if (plan->d_rec_ij) (void)hipFree(plan->d_rec_ij);
    if (plan->d_byrow_tiles) (void)hipFree(plan->d_byrow_tiles);
    if (plan->d_byrow_tilemap) (void)hipFree(plan->d_byrow_tilemap);
    if (plan->d_items) (void)hipFree(plan->d_items);
    delete plan;
}

uint64_t centre_pair_plan_records(const centre_pair_plan* plan) {
    return plan ? plan->n_rec : 0;
}

int centre_pair_plan_tiles(const centre_pair_plan* plan) {
    return plan ? plan->n_tiles : 0;
}

}  // extern "C"

namespace {

// The pair kernel stages codes in 4-byte words, so every step of every code
// row must start on a 4-byte boundary.  Codes laid out otherwise (a row
// stride or a step length that is not a multiple of 4) are copied once to
// rows of nsteps steps of round_up(ss, 4) frames.
bool codes_step_aligned(const uint8_t* codes, uint64_t stride, uint64_t ss) {
    return (stride & 3) == 0 && (ss & 3) == 0 && (reinterpret_cast<uintptr_t>(codes) & 3) == 0;
}

size_t repack_bytes(uint32_t D, uint64_t ss, int S) {
    return (size_t)D * S * ((ss + 3) & ~3ull);
}

// The codes as the pair kernel reads them: the caller's rows when they are
// step-aligned, else a copy in buf (repack_bytes bytes).
int stage_codes(centre_ctx* ctx, const uint8_t* codes, uint64_t stride, uint32_t D, uint64_t ss,
                int S, uint8_t* buf, const uint8_t** out, uint64_t* out_stride,
                uint64_t* out_sstride) {
    if (codes_step_aligned(codes, stride, ss)) {
        *out = codes;
        *out_stride = stride;
        *out_sstride = ss;
        return 0;
    }
    const uint64_t ss4 = (ss + 3) & ~3ull;
    HIPCHK(launch_repack(codes, stride, D, ss, S, ss4, buf, ctx->stream));
    *out = buf;
    *out_stride = (uint64_t)S * ss4;
    *out_sstride = ss4;
    return 0;
}

int step_aligned(centre_ctx* ctx, const uint8_t* codes, uint64_t stride, uint32_t D, uint64_t ss,
                 int S, DevBuf& buf, const uint8_t** out, uint64_t* out_stride,
                 uint64_t* out_sstride) {
    if (!codes_step_aligned(codes, stride, ss)) HIPCHK(buf.alloc(repack_bytes(D, ss, S)));
    return stage_codes(ctx, codes, stride, D, ss, S, buf.as<uint8_t>(), out, out_stride,
                       out_sstride);
}

// Tile list of a batch: entries (row block, column block, system).
struct BatchTile {
    int I, J, sys;
    double cost;   // frames x (1/2 on the diagonal)
};

// order: 1 = longest tiles first; 2 = system by system (row-major tiles);
// 3 = systems by total cost, largest first, each system's tiles together.
// Orders 1 and 3 are stable sorts of the row-major list; both reduce to
// ordering groups of equal-cost tiles (a system's diagonal tiles, its
// off-diagonal tiles) and emitting each group in row-major order, which is
// linear in the number of tiles.  sample < 1 keeps an evenly spread share of
// each system's tiles (calibration) without generating the rest.
std::vector<BatchTile> batch_tiles(const centre_batch_system* sys, int n_sys, int R, int nsteps,
                                   int order, uint64_t max_frames_per_step, double sample) {
    struct Group {
        int sys;
        bool diag;
        double cost;   // per tile
    };
    std::vector<int> nblk(n_sys);
    std::vector<double> frames(n_sys), sys_cost(n_sys, 0.0);
    std::vector<uint64_t> ntile(n_sys), nkeep(n_sys);
    size_t total = 0;
    for (int k = 0; k < n_sys; ++k) {
        nblk[k] = (int)((sys[k].n_dof + R - 1) / R);
        frames[k] = (double)std::min<uint64_t>(sys[k].n_frame / nsteps, max_frames_per_step) * nsteps;
        ntile[k] = (uint64_t)nblk[k] * (nblk[k] + 1) / 2;
        nkeep[k] = sample < 1.0 ? std::max<uint64_t>(1, (uint64_t)(ntile[k] * sample)) : ntile[k];
        total += nkeep[k];
    }
    // tiles of system k, in row-major order, that the sample keeps; with
    // only = 0 / 1 just the off-diagonal / diagonal ones
    auto emit = [&](int k, int only, std::vector<BatchTile>& out) {
        const uint64_t n = ntile[k], m = nkeep[k];
        uint64_t lin = 0, q = 0, next = 0;   // next kept linear index: q * n / m
        for (int I = 0; I < nblk[k] && q < m; ++I)
            for (int J = I; J < nblk[k] && q < m; ++J, ++lin) {
                if (lin != next) continue;
                ++q;
                next = q * n / m;
                const bool d = I == J;
                if (only < 0 || (only == 1) == d)
                    out.push_back({I, J, k, frames[k] * (d ? 0.5 : 1.0)});
            }
    };
    // system cost of the kept tiles
    {
        std::vector<BatchTile> tmp;
        for (int k = 0; k < n_sys; ++k) {
            if (sample < 1.0) {
                tmp.clear();
                emit(k, -1, tmp);
                for (const BatchTile& t : tmp) sys_cost[k] += t.cost;
            } else {
                sys_cost[k] = frames[k] * ((double)ntile[k] - 0.5 * nblk[k]);
            }
        }
    }
    std::vector<BatchTile> out;
    out.reserve(total);
    if (order == 1) {
        std::vector<Group> g;
        for (int k = 0; k < n_sys; ++k) {
            g.push_back({k, false, frames[k]});
            g.push_back({k, true, 0.5 * frames[k]});
        }
        // equal costs keep the row-major list's order: system, then position
        std::stable_sort(g.begin(), g.end(), [](const Group& x, const Group& y) {
            return x.cost > y.cost || (x.cost == y.cost && x.sys < y.sys);
        });
        for (const Group& q : g) emit(q.sys, q.diag ? 1 : 0, out);
    } else {
        std::vector<int> ord(n_sys);
        std::iota(ord.begin(), ord.end(), 0);
        if (order == 3)
            std::stable_sort(ord.begin(), ord.end(),
                             [&](int x, int y) { return sys_cost[x] > sys_cost[y]; });
        for (int k : ord) emit(k, -1, out);
    }
    return out;
}

struct BatchRun {
    double total_ms = 0.0;
    std::vector<double> sys_ms;
};

// One persistent launch over a prepared tile list, in the given order: the
// eight-wave tile when the bin count allows it (every tile, those of a
// system's partial last block included), else the four-wave tile.  Every
// system's completion stamp is taken by the workgroup finishing its last
// tile.
int batch_launch(centre_ctx* ctx, const PairKernelArgs& base, const std::vector<BatchSystem>& bs,
                 const std::vector<BatchTile>& tl, int grid, BatchRun* run) {
    const int n_sys = (int)bs.size();
    const bool w8 = pair_tile8_supported(base.R, base.nb);
    std::vector<int2> tiles;
    std::vector<int> tsys, per(n_sys, 0);
    tiles.reserve(tl.size());
    tsys.reserve(tl.size());
    for (const BatchTile& q : tl) {
        tiles.push_back(make_int2(q.I, q.J));
        tsys.push_back(q.sys);
        per[q.sys]++;
    }
    DevBuf dt, ds, dsys, dper, dcnt, dtime, dwork;
    HIPCHK(dt.alloc(sizeof(int2) * tiles.size()));
    HIPCHK(ds.alloc(sizeof(int) * tsys.size()));
    HIPCHK(dsys.alloc(sizeof(BatchSystem) * n_sys));
    HIPCHK(dper.alloc(sizeof(int) * n_sys));
    HIPCHK(dcnt.alloc(sizeof(int) * n_sys));
    HIPCHK(dtime.alloc(sizeof(unsigned long long) * (n_sys + 1)));
    HIPCHK(dwork.alloc(sizeof(int)));
    HIPCHK(hipMemcpyAsync(dt.p, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ds.p, tsys.data(), sizeof(int) * tsys.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dsys.p, bs.data(), sizeof(BatchSystem) * n_sys, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dper.p, per.data(), sizeof(int) * n_sys, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(dcnt.p, 0, sizeof(int) * n_sys, ctx->stream));
    HIPCHK(hipMemsetAsync(dtime.p, 0, sizeof(unsigned long long) * (n_sys + 1), ctx->stream));
    HIPCHK(hipMemsetAsync(dwork.p, 0, sizeof(int), ctx->stream));
    PairKernelArgs a = base;
    a.systems = dsys.as<BatchSystem>();
    a.tiles = dt.as<int2>();
    a.tile_sys = ds.as<int>();
    a.work = dwork.as<int>();
    a.n_work = (int)tiles.size();
    a.sys_count = dcnt.as<int>();
    a.sys_tiles = dper.as<int>();
    a.sys_time = dtime.as<unsigned long long>() + 1;
    a.t0 = dtime.as<unsigned long long>();
    uint64_t ss_max = 1;
    for (const BatchSystem& b : bs) ss_max = std::max<uint64_t>(ss_max, b.ss);
    a.kfix = pair_fixed_k((double)base.nsteps * (double)ss_max);
    a.int8_only = ctx->int8_only || !pair_f4_exact((double)base.nsteps * (double)ss_max);
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    if (w8) HIPCHK(launch_pair_batch8(a, grid, ctx->stream));
    else HIPCHK(launch_pair_persistent(a, grid, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<unsigned long long> stamps(n_sys + 1);
    HIPCHK(hipMemcpyAsync(stamps.data(), dtime.p, sizeof(unsigned long long) * (n_sys + 1),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    int khz = 100000;
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device);
    run->total_ms = ms;
    run->sys_ms.assign(n_sys, 0.0);
    for (int k = 0; k < n_sys; ++k)
        run->sys_ms[k] = stamps[k + 1] >= stamps[0] && per[k]
                             ? (double)(stamps[k + 1] - stamps[0]) / (double)khz
                             : 0.0;
    return 0;
}

}  // namespace

extern "C" int centre_batch_pair_entropy(centre_ctx* ctx, const centre_batch_system* sys, int n_sys,
                                         int nbins, int nsteps, const int* ests, int nest,
                                         int jacobian, int kde, int order, int tile_dofs,
                                         double* sys_ms, double* total_ms, int* order_used,
                                         int* tile_dofs_used) {
    if (!ctx) return -3;
    if (!sys || n_sys < 1) return fail(ctx, -3, "no systems");
    if (nbins < 1 || nbins > 255) return fail(ctx, -3, "nbins must be in 1..255");
    if (nsteps < 1 || nsteps > 255) return fail(ctx, -3, "nsteps must be in 1..255");
    if (!ests || nest < 1) return fail(ctx, -3, "no estimators");
    if (order < 0 || order > 3) return fail(ctx, -3, "order must be 0 (autotune) .. 3");
    const int r_max = dofs_per_tile(nbins);
    if (tile_dofs < 0 || tile_dofs > r_max || (pair_tile8_small(nbins) && tile_dofs != 0 && tile_dofs != r_max))
        return fail(ctx, -3, "tile_dofs must be 0 (autotune) or 1.." + std::to_string(r_max) +
                                 (pair_tile8_small(nbins) ? " (32 at nbins <= 8)" : ""));
    int slot_ml = -1, slot_mm = -1;
    for (int e = 0; e < nest; ++e) {
        if (ests[e] == 0 && slot_ml < 0) slot_ml = e;
        else if (ests[e] == 1 && slot_mm < 0) slot_mm = e;
        else return fail(ctx, -3, "batched pair stage supports the ML and MM estimators");
    }
    for (int k = 0; k < n_sys; ++k) {
        const centre_batch_system& b = sys[k];
        if (!b.d_codes || !b.d_types || !b.d_extrema || !b.d_s2)
            return fail(ctx, -3, "null pointer in system " + std::to_string(k));
        if (b.n_dof < 2) return fail(ctx, -3, "system " + std::to_string(k) + " has < 2 DOFs");
        if (b.n_frame < (uint64_t)nsteps || b.code_stride < b.n_frame)
            return fail(ctx, -3, "system " + std::to_string(k) + ": bad n_frame or code_stride");
    }
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemsetAsync(ctx->d_flag, 0, sizeof(int), ctx->stream));
    // marginals of every system
    std::vector<DevBuf> s1(n_sys), L(n_sys), w(n_sys), s1e0(n_sys), packed(n_sys);
    std::vector<BatchSystem> bs(n_sys);
    for (int k = 0; k < n_sys; ++k) {
        const centre_batch_system& b = sys[k];
        const uint64_t D = b.n_dof;
        HIPCHK(s1[k].alloc(sizeof(double) * D * nest * nsteps));
        HIPCHK(L[k].alloc(sizeof(double) * D * nsteps));
        HIPCHK(w[k].alloc(sizeof(double) * D));
        if (int r = run_dof1d(ctx, b.d_codes, b.code_stride, b.n_dof, b.n_frame, b.d_types,
                              b.d_extrema, nbins, nsteps, ests, nest, jacobian, kde,
                              s1[k].as<double>(), L[k].as<double>(), w[k].as<double>(), nullptr))
            return r;
        if (b.d_s1)
            HIPCHK(hipMemcpyAsync(b.d_s1, s1[k].p, sizeof(double) * D * nest * nsteps,
                                  hipMemcpyDeviceToDevice, ctx->stream));
        const double* s1_first = s1[k].as<double>();
        if (nest > 1) {   // the kernel reads [D][S] of the first estimator
            HIPCHK(s1e0[k].alloc(sizeof(double) * D * nsteps));
            HIPCHK(launch_copy_2d(s1e0[k].p, sizeof(double) * nsteps, s1[k].p,
                                  sizeof(double) * nsteps * nest, sizeof(double) * nsteps, D,
                                  ctx->stream));
            s1_first = s1e0[k].as<double>();
        }
        const uint8_t* pc = nullptr;
        uint64_t pstride = 0, psstride = 0;
        if (int r = step_aligned(ctx, b.d_codes, b.code_stride, b.n_dof, b.n_frame / nsteps, nsteps,
                                 packed[k], &pc, &pstride, &psstride))
            return r;
        bs[k] = BatchSystem{pc, pstride, (int)b.n_dof, b.n_frame / nsteps, psstride, b.d_s2,
                            b.d_mi, w[k].as<double>(), L[k].as<double>(), s1_first};
    }
    PairKernelArgs base{};
    base.nb = nbins;
    base.R = r_max;
    base.nsteps = nsteps;
    base.dof_w = nullptr;
    base.slot_ml = slot_ml;
    base.slot_mm = slot_mm;
    base.nest = nest;
    base.mi_from_mm = ests[0] == 1;
    base.flag = ctx->d_flag;
    int dev_cu = 0;
    HIPCHK(hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int grid = dev_cu;   // one workgroup per CU: the tile's LDS fills most of a CU
    // Autotuner over the unset knobs: the tile-list order (1..3) and the
    // tile shape, R DOFs a side (the largest that fits the 256-row tile, or
    // one fewer: more, smaller tiles spread a batch of small systems wider).
    // Each candidate runs on a calibration share of the batch (2% of each
    // system's tiles, at most 1,024 frames per step); about 0.2% of the full
    // run's work per candidate.
    std::vector<int> orders, shapes;
    if (order) orders.push_back(order);
    else orders = {1, 2, 3};
    if (tile_dofs) shapes.push_back(tile_dofs);
    else {
        shapes.push_back(r_max);
        if (!pair_tile8_small(nbins) && r_max > 1) shapes.push_back(r_max - 1);
    }
    int chosen = orders[0], chosen_R = shapes[0];
    if (orders.size() * shapes.size() > 1) {
        double best = 1e300;
        for (int R : shapes)
            for (int cand : orders) {
                PairKernelArgs cb = base;
                cb.R = R;
                auto tl = batch_tiles(sys, n_sys, R, nsteps, cand, 1024, 0.02);
                // calibration outputs land in the real output buffers and
                // are overwritten by the full run below
                BatchRun run;
                std::vector<BatchSystem> cbs = bs;
                for (auto& c : cbs) c.ss = std::min<uint64_t>(c.ss, 1024);
                if (int r = batch_launch(ctx, cb, cbs, tl, grid, &run)) return r;
                if (run.total_ms < best) {
                    best = run.total_ms;
                    chosen = cand;
                    chosen_R = R;
                }
            }
    }
    base.R = chosen_R;
    auto tl = batch_tiles(sys, n_sys, chosen_R, nsteps, chosen, ~0ull, 1.0);
    BatchRun run;
    if (int r = batch_launch(ctx, base, bs, tl, grid, &run)) return r;
    if (sys_ms)
        for (int k = 0; k < n_sys; ++k) sys_ms[k] = run.sys_ms[k];
    if (total_ms) *total_ms = run.total_ms;
    if (order_used) *order_used = chosen;
    if (tile_dofs_used) *tile_dofs_used = chosen_R;
    ctx->last_pair_ms = (float)run.total_ms;
    return check_flag(ctx);
}

namespace {

// ---- pair stage ---------------------------------------------------------------
// Device temporaries of one fused (ML / MM, no count output) pair-stage call.
struct PairWork {
    double* s1 = nullptr;      // [D][nest][S] marginals of the same frames
    double* L = nullptr;       // [D][S] sum_b m_b log J_b
    double* w = nullptr;       // [D] bin widths
    double* s1e0 = nullptr;    // [D][S] S1 of estimator 0 (nest > 1 with MI)
    uint8_t* packed = nullptr; // step-aligned copy of the codes, when needed
    double* cs2 = nullptr;     // [P][nest][S] canonical-order S2 (explicit lists)
    double* cmi = nullptr;     // [P][S] canonical-order MI (explicit lists)
    uint32_t* split = nullptr; // frame-split count dumps
};

// Sizes (bytes) of the parts of PairWork a call needs (0 = not needed) and
// the frame-split shape.
struct PairLayout {
    size_t s1 = 0, L = 0, w = 0, s1e0 = 0, packed = 0, cs2 = 0, cmi = 0, split = 0;
    int nsplit = 1;
    uint64_t split_len = 0;
    int rec = 0;
    // total of every part but the split dumps, each rounded up to 256 bytes
    size_t base_total() const {
        auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
        return up(s1) + up(L) + up(w) + up(s1e0) + up(packed) + up(cs2) + up(cmi);
    }
};

PairLayout pair_layout(const centre_pair_plan* plan, const uint8_t* d_codes, uint64_t code_stride,
                       uint64_t n_frame, int nsteps, int nest, bool want_mi) {
    const uint64_t D = plan->n_dof, P = plan->n_canon, S = (uint64_t)nsteps;
    const uint64_t ss = n_frame / nsteps;
    PairLayout l;
    l.s1 = sizeof(double) * D * nest * S;
    l.L = sizeof(double) * D * S;
    l.w = sizeof(double) * D;
    if (nest > 1 && want_mi) l.s1e0 = sizeof(double) * D * S;
    if (!codes_step_aligned(d_codes, code_stride, ss)) l.packed = repack_bytes(plan->n_dof, ss, nsteps);
    if (!plan->all_pairs) {
        l.cs2 = sizeof(double) * P * nest * S;
        if (want_mi) l.cmi = sizeof(double) * P * S;
    }
    return l;
}

// Frame split for plans with few tiles at small nb: the four-wave kernel is
// bound by one-hot staging there, so K workgroups per tile each count a K-th
// of every step into their own count dump and a second kernel sums the
// dumps and evaluates the entropies.  Chosen when the dumps fit max_bytes.
int choose_split(centre_ctx* ctx, const centre_pair_plan* plan, uint64_t ss, int nsteps,
                 size_t max_bytes, PairLayout* l) {
    const int nb = plan->nbins;
    l->rec = (nb * nb + 3) & ~3;
    // nb <= 8 runs on the small-bin eight-wave tile, which needs no split
    if (nb > 16 || plan->n_tiles <= 0 || pair_tile8_small(nb)) return 0;
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    int k = std::min<int>(8, ncu / plan->n_tiles);
    k = (int)std::min<uint64_t>((uint64_t)k, ss / 512);
    if (k < 2) return 0;
    const uint64_t len = ((ss + k - 1) / k + 63) / 64 * 64;   // whole chunks, 4-byte aligned
    const int parts = (int)((ss + len - 1) / len);
    const size_t bytes = sizeof(uint32_t) * plan->n_canon * nsteps * (size_t)l->rec * parts;
    if (parts < 2 || bytes > max_bytes) return 0;
    l->nsplit = parts;
    l->split_len = len;
    l->split = bytes;
    return 0;
}

// Fused pair stage (ML / MM into the count pass, no count output) with the
// temporaries in `ws` (sized by `l`).
int run_pair_fused(centre_ctx* ctx, const centre_pair_plan* plan, const uint8_t* d_codes,
                   uint64_t code_stride, uint64_t n_frame, const uint8_t* d_types,
                   const double* d_extrema, int nsteps, const int* ests, int nest, int jacobian,
                   int kde, double* d_s1, double* d_s2, double* d_mi, const PairLayout& l,
                   const PairWork& ws) {
    const uint32_t D = plan->n_dof;
    const int nb = plan->nbins;
    const uint64_t ss = n_frame / nsteps;
    HIPCHK(hipMemsetAsync(ctx->d_flag, 0, sizeof(int), ctx->stream));
    if (int r = run_dof1d(ctx, d_codes, code_stride, D, n_frame, d_types, d_extrema, nb, nsteps,
                          ests, nest, jacobian, kde, ws.s1, ws.L, ws.w, nullptr))
        return r;
    if (d_s1)
        HIPCHK(hipMemcpyAsync(d_s1, ws.s1, sizeof(double) * D * nest * nsteps,
                              hipMemcpyDeviceToDevice, ctx->stream));
    int slot_ml = -1, slot_mm = -1;
    for (int e = 0; e < nest; ++e) {
        if (ests[e] == 0 && slot_ml < 0) slot_ml = e;
        else if (ests[e] == 1 && slot_mm < 0) slot_mm = e;
    }
    const bool all = plan->all_pairs;
    PairKernelArgs a{};
    a.int8_only = ctx->int8_only;
    if (int r = stage_codes(ctx, d_codes, code_stride, D, ss, nsteps, ws.packed, &a.codes,
                            &a.stride, &a.sstride))
        return r;
    a.n_dof = (int)D;
    a.nb = nb;
    a.R = plan->dofs_per_tile;
    a.nsteps = nsteps;
    a.ss = ss;
    a.tiles = plan->d_tiles;
    a.tilemap = plan->d_tilemap;
    a.dof_w = ws.w;
    a.dof_L = ws.L;
    a.dof_S1 = ws.s1;
    a.slot_ml = slot_ml;
    a.slot_mm = slot_mm;
    a.nest = nest;
    a.s2 = all ? d_s2 : ws.cs2;
    a.mi = d_mi ? (all ? d_mi : ws.cmi) : nullptr;
    a.mi_from_mm = (ests[0] == 1);
    a.flag = ctx->d_flag;
    // dof_S1 is [D][nest][S]; the kernel reads [D][S] of estimator 0
    if (nest > 1 && d_mi) {
        HIPCHK(launch_copy_2d(ws.s1e0, sizeof(double) * nsteps, ws.s1,
                              sizeof(double) * nsteps * nest, sizeof(double) * nsteps, D,
                              ctx->stream));
        a.dof_S1 = ws.s1e0;
    }
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    if (l.nsplit > 1) {
        PairKernelArgs sp = a;
        sp.counts = ws.split;
        sp.nsplit = l.nsplit;
        sp.split_tiles = plan->n_tiles;
        sp.split_len = l.split_len;
        sp.split_rec = l.rec;
        sp.split_cstride = plan->n_canon * nsteps * (uint64_t)l.rec;
        PairKernelArgs dumpargs = sp;   // counts only
        dumpargs.slot_ml = dumpargs.slot_mm = -1;
        dumpargs.mi = nullptr;
        HIPCHK(launch_pair_tiles(dumpargs, plan->n_tiles, true, ctx->stream));
        HIPCHK(launch_split_entropy(sp, plan->n_tiles, ctx->stream));
    } else if (pair_tile8_supported(a.R, nb) && pair_tile8_small(nb) &&
               ((a.stride | a.sstride | reinterpret_cast<uintptr_t>(a.codes)) & 3) == 0) {
        // every tile, partial ones included, on the small-bin eight-wave kernel
        HIPCHK(launch_pair_tiles8(a, plan->n_tiles, ctx->stream));
    } else if (pair_tile8_supported(a.R, nb) && plan->n_full > 0 &&
               ((a.stride | a.sstride | reinterpret_cast<uintptr_t>(a.codes)) & 3) == 0) {
        // full tiles on the eight-wave kernel, those of a partial block after
        HIPCHK(launch_pair_tiles8(a, plan->n_full, ctx->stream));
        PairKernelArgs r = a;
        r.tiles = plan->d_tiles + plan->n_full;
        if (r.tilemap) r.tilemap += (size_t)plan->n_full * a.R * a.R;
        HIPCHK(launch_pair_tiles(r, plan->n_tiles - plan->n_full, false, ctx->stream));
    } else {
        HIPCHK(launch_pair_tiles(a, plan->n_tiles, false, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    if (!all) {
        HIPCHK(launch_gather_rec(ws.cs2, plan->d_rec2canon, plan->n_rec, nest * nsteps, d_s2,
                                 ctx->stream));
        if (d_mi)
            HIPCHK(launch_gather_rec(ws.cmi, plan->d_rec2canon, plan->n_rec, nsteps, d_mi,
                                     ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_pair_ms = ms;
    return check_flag(ctx);
}

// Row-block grouping of a plan (see centre_pair_plan), built once.
int ensure_grouping(centre_ctx* ctx, const centre_pair_plan* plan) {
    if (plan->grouped) return 0;
    const int R = plan->dofs_per_tile;
    const uint64_t D = plan->n_dof;
    const int nblk = (int)((D + R - 1) / R);
    const size_t nt = plan->h_tiles.size();
    // tiles by row block, launch order kept within a row block; a tile-subset
    // list may name a tile twice
    std::vector<size_t> ord(nt);
    std::iota(ord.begin(), ord.end(), (size_t)0);
    std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
        const int2 a = plan->h_tiles[x], b = plan->h_tiles[y];
        return a.x != b.x ? a.x < b.x : (plan->tile_subset && a.y < b.y);
    });
    std::vector<int2> bt;
    std::vector<int> btm;
    std::vector<int> tile_off(nblk + 1, 0);
    for (size_t q = 0; q < nt; ++q) {
        const int2 t = plan->h_tiles[ord[q]];
        if (plan->tile_subset && !bt.empty() && bt.back().x == t.x && bt.back().y == t.y) continue;
        bt.push_back(t);
        tile_off[t.x + 1]++;
        if (!plan->h_tilemap.empty())
            btm.insert(btm.end(), plan->h_tilemap.begin() + ord[q] * R * R,
                       plan->h_tilemap.begin() + (ord[q] + 1) * R * R);
    }
    for (int I = 0; I < nblk; ++I) tile_off[I + 1] += tile_off[I];
    // canonical index ranges of the row blocks
    std::vector<uint64_t> canon_off(nblk + 1, 0), item_off(nblk + 1, 0);
    std::vector<uint64_t> items;
    auto row_off = [&](uint64_t i) { return i * (2 * D - i - 1) / 2; };   // first pair of row i
    if (plan->all_pairs) {
        for (int I = 0; I <= nblk; ++I) canon_off[I] = row_off(std::min<uint64_t>((uint64_t)I * R, D));
        if (plan->tile_subset) {
            // canonical indices of the listed tiles' pairs, ascending per row block
            for (int I = 0; I < nblk; ++I) {
                const size_t first = items.size();
                for (int q = tile_off[I]; q < tile_off[I + 1]; ++q) {
                    const uint64_t J = (uint64_t)bt[q].y;
                    for (uint64_t i = (uint64_t)I * R; i < std::min<uint64_t>((uint64_t)(I + 1) * R, D); ++i)
                        for (uint64_t j = std::max(J * R, i + 1); j < std::min((J + 1) * R, D); ++j)
                            items.push_back(row_off(i) + (j - i - 1));
                }
                std::sort(items.begin() + first, items.end());
                items.erase(std::unique(items.begin() + first, items.end()), items.end());
                item_off[I + 1] = items.size();
            }
        }
    } else {
        // canonical indices are ordered by (lower DOF, higher DOF): count the
        // distinct ones per row block of the lower DOF
        std::vector<uint8_t> seen(plan->n_canon, 0);
        std::vector<uint64_t> per(nblk, 0), rper(nblk, 0);
        const uint64_t NR = plan->n_rec;
        for (uint64_t r = 0; r < NR; ++r) {
            const uint32_t lo = std::min(plan->h_rec_ij[2 * r], plan->h_rec_ij[2 * r + 1]);
            const uint64_t c = plan->h_rec2canon[r];
            rper[lo / R]++;
            if (!seen[c]) {
                seen[c] = 1;
                per[lo / R]++;
            }
        }
        for (int I = 0; I < nblk; ++I) {
            canon_off[I + 1] = canon_off[I] + per[I];
            item_off[I + 1] = item_off[I] + rper[I];
        }
        // record ids sorted by canonical index (record order among equals)
        items.resize(NR);
        std::iota(items.begin(), items.end(), (uint64_t)0);
        std::stable_sort(items.begin(), items.end(), [&](uint64_t x, uint64_t y) {
            return plan->h_rec2canon[x] < plan->h_rec2canon[y];
        });
    }
    hipError_t e = hipSuccess;
    if (!bt.empty() && (e = hipMalloc(&plan->d_byrow_tiles, sizeof(int2) * bt.size())) == hipSuccess)
        e = hipMemcpy(plan->d_byrow_tiles, bt.data(), sizeof(int2) * bt.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !btm.empty() &&
        (e = hipMalloc(&plan->d_byrow_tilemap, sizeof(int) * btm.size())) == hipSuccess)
        e = hipMemcpy(plan->d_byrow_tilemap, btm.data(), sizeof(int) * btm.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !items.empty() &&
        (e = hipMalloc(&plan->d_items, sizeof(uint64_t) * items.size())) == hipSuccess)
        e = hipMemcpy(plan->d_items, items.data(), sizeof(uint64_t) * items.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(ctx, (int)e, std::string("plan grouping: ") + hipGetErrorString(e));
    plan->tile_off = std::move(tile_off);
    plan->canon_off = std::move(canon_off);
    plan->item_off = std::move(item_off);
    plan->grouped = true;
    return 0;
}

// Count budget of one group of the grouped pass (a single row block may
// exceed it and then forms a group of its own).
constexpr size_t kGroupCountBytes = (size_t)2 << 30;

// Pair stage through materialised counts: any estimator (Chao-Shen and
// James-Stein need whole count tables), and the count output.  The plan's
// row blocks are taken in groups whose counts fit kGroupCountBytes: the
// pair kernel dumps a group's counts, then count_entropy_kernel evaluates
// every estimator, MI and the record-order count output of its records.
int run_pair_grouped(centre_ctx* ctx, const centre_pair_plan* plan, const uint8_t* d_codes,
                     uint64_t code_stride, uint64_t n_frame, const uint8_t* d_types,
                     const double* d_extrema, int nsteps, const int* ests, int nest,
                     int jacobian, int kde, double* d_s1, double* d_s2, double* d_mi,
                     uint32_t* d_counts) {
    const uint32_t D = plan->n_dof;
    const int nb = plan->nbins;
    const int R = plan->dofs_per_tile;
    const uint64_t ss = n_frame / nsteps;
    DevBuf s1, L, w;
    HIPCHK(s1.alloc(sizeof(double) * D * nest * nsteps));
    HIPCHK(L.alloc(sizeof(double) * D * nsteps));
    HIPCHK(w.alloc(sizeof(double) * D));
    HIPCHK(hipMemsetAsync(ctx->d_flag, 0, sizeof(int), ctx->stream));
    if (int r = run_dof1d(ctx, d_codes, code_stride, D, n_frame, d_types, d_extrema, nb, nsteps,
                          ests, nest, jacobian, kde, s1.as<double>(), L.as<double>(),
                          w.as<double>(), nullptr))
        return r;
    if (d_s1)
        HIPCHK(hipMemcpyAsync(d_s1, s1.p, sizeof(double) * D * nest * nsteps,
                              hipMemcpyDeviceToDevice, ctx->stream));
    if (int r = ensure_grouping(ctx, plan)) return r;
    const int nblk = (int)((D + R - 1) / R);
    const size_t per = sizeof(uint32_t) * (size_t)nsteps * nb * nb;   // counts of one pair
    std::vector<std::pair<int, int>> groups;
    size_t most = 0;
    for (int I0 = 0; I0 < nblk;) {
        int I1 = I0 + 1;
        while (I1 < nblk && (plan->canon_off[I1 + 1] - plan->canon_off[I0]) * per <= kGroupCountBytes)
            ++I1;
        groups.emplace_back(I0, I1);
        most = std::max(most, (size_t)(plan->canon_off[I1] - plan->canon_off[I0]) * per);
        I0 = I1;
    }
    void* cnt = nullptr;
    HIPCHK(ctx_scratch(ctx, 4, most, &cnt));
    PairKernelArgs a{};
    DevBuf packed;
    if (int r = step_aligned(ctx, d_codes, code_stride, D, ss, nsteps, packed, &a.codes, &a.stride,
                             &a.sstride))
        return r;
    a.n_dof = (int)D;
    a.nb = nb;
    a.R = R;
    a.nsteps = nsteps;
    a.ss = ss;
    a.dof_w = w.as<double>();
    a.dof_L = L.as<double>();
    a.slot_ml = a.slot_mm = -1;   // counts only
    a.nest = nest;
    a.counts = static_cast<uint32_t*>(cnt);
    a.flag = ctx->d_flag;
    CountEntropyArgs c{};
    c.counts = static_cast<const uint32_t*>(cnt);
    c.mode = plan->all_pairs ? (plan->tile_subset ? 1 : 0) : 2;
    c.rec2canon = plan->d_rec2canon;
    c.rec_ij = plan->d_rec_ij;
    c.n_dof = D;
    c.nb = nb;
    c.nsteps = nsteps;
    c.ss = ss;
    c.types = d_types;
    c.ext = d_extrema;
    c.jacobian = jacobian;
    c.kde = kde;
    for (int e = 0; e < nest; ++e) c.ests[e] = ests[e];
    c.nest = nest;
    c.s2 = d_s2;
    c.s1 = s1.as<double>();
    c.mi = d_mi;
    c.counts_out = d_counts;
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    for (const auto& g : groups) {
        const int nt = plan->tile_off[g.second] - plan->tile_off[g.first];
        if (nt == 0) continue;
        PairKernelArgs ag = a;
        ag.tiles = plan->d_byrow_tiles + plan->tile_off[g.first];
        ag.tilemap = plan->d_byrow_tilemap
                         ? plan->d_byrow_tilemap + (size_t)plan->tile_off[g.first] * R * R
                         : nullptr;
        ag.count_base = plan->canon_off[g.first];
        HIPCHK(launch_pair_tiles(ag, nt, true, ctx->stream));
        CountEntropyArgs cg = c;
        cg.count_base = plan->canon_off[g.first];
        if (c.mode == 0) {
            cg.item_first = plan->canon_off[g.first];
            cg.n_item = plan->canon_off[g.second] - plan->canon_off[g.first];
        } else {
            cg.items = plan->d_items + plan->item_off[g.first];
            cg.n_item = plan->item_off[g.second] - plan->item_off[g.first];
        }
        HIPCHK(launch_count_entropy(cg, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_pair_ms = ms;
    return check_flag(ctx);
}

int check_pair_args(centre_ctx* ctx, const centre_pair_plan* plan, const uint8_t* d_codes,
                    uint64_t code_stride, uint64_t n_frame, const uint8_t* d_types,
                    const double* d_extrema, int nsteps, const int* ests, int nest,
                    const double* d_s2) {
    if (!plan) return fail(ctx, -3, "null plan");
    if (int r = check_ests(ctx, ests, nest)) return r;
    if (!d_codes || !d_types || !d_extrema || !d_s2) return fail(ctx, -3, "null buffer");
    if (nsteps < 1 || nsteps > 255) return fail(ctx, -3, "nsteps must be in 1..255");
    if (n_frame / nsteps == 0) return fail(ctx, -3, "fewer frames than steps");
    if (code_stride < n_frame) return fail(ctx, -3, "code_stride < n_frame");
    return 0;
}

// The fused pair stage's temporaries laid out back to back from `base`, each
// part on a 256-byte boundary, in base_total()'s order; the split dumps follow
// them when `with_split` is set.
PairWork carve_pair_work(uint8_t* base, const PairLayout& l, bool with_split) {
    uint8_t* next = base;
    auto take = [&next](size_t bytes) -> void* {
        if (!bytes) return nullptr;
        void* p = next;
        next += (bytes + 255) & ~(size_t)255;
        return p;
    };
    PairWork ws;
    ws.s1 = static_cast<double*>(take(l.s1));
    ws.L = static_cast<double*>(take(l.L));
    ws.w = static_cast<double*>(take(l.w));
    ws.s1e0 = static_cast<double*>(take(l.s1e0));
    ws.packed = static_cast<uint8_t*>(take(l.packed));
    ws.cs2 = static_cast<double*>(take(l.cs2));
    ws.cmi = static_cast<double*>(take(l.cmi));
    if (with_split) ws.split = static_cast<uint32_t*>(take(l.split));
    return ws;
}

bool fused_estimators(const int* ests, int nest) {
    bool ml = false, mm = false;
    for (int e = 0; e < nest; ++e) {
        if (ests[e] == 0 && !ml) ml = true;
        else if (ests[e] == 1 && !mm) mm = true;
        else return false;
    }
    return true;
}

}  // namespace

extern "C" int centre_dev_pair_entropy(centre_ctx* ctx, const centre_pair_plan* plan,
                                       const uint8_t* d_codes, uint64_t code_stride,
                                       uint64_t n_frame, const uint8_t* d_types,
                                       const double* d_extrema, int nsteps,
                                       const int* ests, int nest, int jacobian, int kde,
                                       double* d_s1, double* d_s2, double* d_mi,
                                       uint32_t* d_counts) {
    if (!ctx) return -3;
    if (int r = check_pair_args(ctx, plan, d_codes, code_stride, n_frame, d_types, d_extrema,
                                nsteps, ests, nest, d_s2))
        return r;
    HIPCHK(hipSetDevice(ctx->device));
    if (d_counts || !fused_estimators(ests, nest))
        return run_pair_grouped(ctx, plan, d_codes, code_stride, n_frame, d_types, d_extrema,
                                nsteps, ests, nest, jacobian, kde, d_s1, d_s2, d_mi, d_counts);
    PairLayout l = pair_layout(plan, d_codes, code_stride, n_frame, nsteps, nest, d_mi != nullptr);
    if (int r = choose_split(ctx, plan, n_frame / nsteps, nsteps, (size_t)1 << 30, &l)) return r;
    // temporaries in the context's grow-only scratch (no allocation once a
    // call of this size has run)
    void* base = nullptr;
    HIPCHK(ctx_scratch(ctx, 5, l.base_total(), &base));
    PairWork ws = carve_pair_work(static_cast<uint8_t*>(base), l, false);
    if (l.split) {   // allocated before the timed region, kept by the context
        void* sp = nullptr;
        HIPCHK(ctx_scratch(ctx, 2, l.split, &sp));
        ws.split = static_cast<uint32_t*>(sp);
    }
    return run_pair_fused(ctx, plan, d_codes, code_stride, n_frame, d_types, d_extrema, nsteps,
                          ests, nest, jacobian, kde, d_s1, d_s2, d_mi, l, ws);
}

extern "C" int centre_dev_pair_entropy_schemes(centre_ctx* ctx,
                                               const centre_pair_plan* const* plans,
                                               int n_schemes, const uint8_t* d_codes,
                                               uint64_t code_stride, uint64_t n_frame,
                                               const uint8_t* d_types, const double* d_extrema,
                                               int nsteps, const int* ests, int nest,
                                               int jacobian, int kde, double* d_s1,
                                               double* d_s2, double* d_mi) {
    if (!ctx) return -3;
    if (!plans || n_schemes < 1 || n_schemes > 255) return fail(ctx, -3, "need 1..255 schemes");
    for (int sc = 0; sc < n_schemes; ++sc) {
        if (!plans[sc]) return fail(ctx, -3, "null plan");
        if (plans[sc]->n_dof != plans[0]->n_dof || plans[sc]->n_rec != plans[0]->n_rec ||
            plans[sc]->all_pairs != plans[0]->all_pairs ||
            plans[sc]->tile_subset != plans[0]->tile_subset ||
            plans[sc]->h_rec_ij != plans[0]->h_rec_ij)
            return fail(ctx, -3, "the schemes' plans must hold the same DOFs and pairs");
    }
    if (int r = check_pair_args(ctx, plans[0], d_codes, code_stride, n_frame, d_types, d_extrema,
                                nsteps, ests, nest, d_s2))
        return r;
    if (n_schemes == 1)
        return centre_dev_pair_entropy(ctx, plans[0], d_codes, code_stride, n_frame, d_types,
                                       d_extrema, nsteps, ests, nest, jacobian, kde, d_s1, d_s2,
                                       d_mi, nullptr);
    HIPCHK(hipSetDevice(ctx->device));
    const uint64_t D = plans[0]->n_dof, P = plans[0]->n_rec, S = (uint64_t)nsteps;
    DevBuf s1, s2, mi;
    HIPCHK(s1.alloc(sizeof(double) * D * nest * S));
    HIPCHK(s2.alloc(sizeof(double) * P * nest * S));
    if (d_mi) HIPCHK(mi.alloc(sizeof(double) * P * S));
    const size_t row = sizeof(double) * S;   // one step series
    float total_ms = 0.f;
    for (int sc = 0; sc < n_schemes; ++sc) {
        // scheme sc of DOF d is row d * n_schemes + sc of the code matrix
        if (int r = centre_dev_pair_entropy(ctx, plans[sc], d_codes + (size_t)sc * code_stride,
                                            code_stride * n_schemes, n_frame, d_types, d_extrema,
                                            nsteps, ests, nest, jacobian, kde, s1.as<double>(),
                                            s2.as<double>(), d_mi ? mi.as<double>() : nullptr,
                                            nullptr))
            return r;
        total_ms += ctx->last_pair_ms;
        // [x][S] series into the [x][n_schemes][S] layout
        HIPCHK(launch_copy_2d(reinterpret_cast<uint8_t*>(d_s2) + sc * row, row * n_schemes,
                              s2.p, row, row, P * nest, ctx->stream));
        if (d_s1)
            HIPCHK(launch_copy_2d(reinterpret_cast<uint8_t*>(d_s1) + sc * row, row * n_schemes,
                                  s1.p, row, row, D * nest, ctx->stream));
        if (d_mi)
            HIPCHK(launch_copy_2d(reinterpret_cast<uint8_t*>(d_mi) + sc * row, row * n_schemes,
                                  mi.p, row, row, P, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->last_pair_ms = total_ms;
    return 0;
}

extern "C" int centre_dev_pair_entropy_noalloc(centre_ctx* ctx, const centre_pair_plan* plan,
                                               const uint8_t* d_codes, uint64_t code_stride,
                                               uint64_t n_frame, const uint8_t* d_types,
                                               const double* d_extrema, int nsteps,
                                               const int* ests, int nest, int jacobian,
                                               int kde, double* d_s1, double* d_s2,
                                               double* d_mi) {
    if (!ctx) return -3;
    if (int r = check_pair_args(ctx, plan, d_codes, code_stride, n_frame, d_types, d_extrema,
                                nsteps, ests, nest, d_s2))
        return r;
    if (!fused_estimators(ests, nest))
        return fail(ctx, -3, "the allocation-free pair stage takes the ML and MM estimators only");
    if (!ctx->noalloc) return fail(ctx, -3, "no static scratch on this device");
    HIPCHK(hipSetDevice(ctx->device));
    PairLayout l = pair_layout(plan, d_codes, code_stride, n_frame, nsteps, nest, d_mi != nullptr);
    const size_t base = l.base_total();
    if (base > kNoAllocBytes)
        return fail(ctx, -3, "the temporaries of this call (" + std::to_string(base) +
                                 " bytes) exceed the static scratch (" +
                                 std::to_string(kNoAllocBytes) + " bytes)");
    // the frame split only where its count dumps fit what is left
    if (int r = choose_split(ctx, plan, n_frame / nsteps, nsteps, kNoAllocBytes - base, &l))
        return r;
    PairWork ws = carve_pair_work(static_cast<uint8_t*>(ctx->noalloc), l, true);
    return run_pair_fused(ctx, plan, d_codes, code_stride, n_frame, d_types, d_extrema, nsteps,
                          ests, nest, jacobian, kde, d_s1, d_s2, d_mi, l, ws);
}

extern "C" int centre_pair_entropy(centre_ctx* ctx, const uint8_t* codes, uint32_t n_dof,
                                   uint64_t n_frame, const uint8_t* types,
                                   const double* extrema, int nbins, int nsteps,
                                   const uint32_t* pairs, uint64_t n_pair, const int* ests,
                                   int nest, int jacobian, int kde, double* s1, double* s2,
                                   double* mi) {
    if (!ctx) return -3;
    if (!codes || !types || !extrema || !s2) return fail(ctx, -3, "null buffer");
    if (int r = check_ests(ctx, ests, nest)) return r;
    HIPCHK(hipSetDevice(ctx->device));
    centre_pair_plan* plan = nullptr;
    if (int r = centre_pair_plan_create(ctx, n_dof, nbins, pairs, n_pair, &plan)) return r;
    struct PlanGuard {
        centre_pair_plan* p;
        ~PlanGuard() { centre_pair_plan_destroy(p); }
    } guard{plan};
    const uint64_t P = plan->n_rec;
    DevBuf dc, dt, dx, ds1, ds2, dmi;
    HIPCHK(dc.alloc((size_t)n_dof * n_frame));
    HIPCHK(dt.alloc(n_dof));
    HIPCHK(dx.alloc(sizeof(double) * 4 * n_dof));
    HIPCHK(ds1.alloc(sizeof(double) * n_dof * nest * nsteps));
    HIPCHK(ds2.alloc(sizeof(double) * P * nest * nsteps));
    if (mi) HIPCHK(dmi.alloc(sizeof(double) * P * nsteps));
    HIPCHK(hipMemcpyAsync(dc.p, codes, (size_t)n_dof * n_frame, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dt.p, types, n_dof, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dx.p, extrema, sizeof(double) * 4 * n_dof, hipMemcpyHostToDevice, ctx->stream));
    if (int r = centre_dev_pair_entropy(ctx, plan, dc.as<uint8_t>(), n_frame, n_frame,
                                        dt.as<uint8_t>(), dx.as<double>(), nsteps, ests, nest,
                                        jacobian, kde, ds1.as<double>(), ds2.as<double>(),
                                        mi ? dmi.as<double>() : nullptr, nullptr))
        return r;
    if (s1)
        HIPCHK(hipMemcpyAsync(s1, ds1.p, sizeof(double) * n_dof * nest * nsteps,
                              hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(s2, ds2.p, sizeof(double) * P * nest * nsteps, hipMemcpyDeviceToHost,
                          ctx->stream));
    if (mi)
        HIPCHK(hipMemcpyAsync(mi, dmi.p, sizeof(double) * P * nsteps, hipMemcpyDeviceToHost,
                              ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Modified Bessel functions I0, I1 (Abramowitz & Stegun 9.8.1-9.8.4
// polynomial approximations) and In, n >= 2, by Miller's downward recurrence
// normalised with I0 -- the scheme Utils::modifiedBesselI uses
// (Utils.cpp:60-168), including its iteration count 2*(n + int(sqrt(40n))).
static double bessel_i0(double x) {
    if (fabs(x) < 3.75) {
        double y = x / 3.75;
        y = y * y;
        const double t = 1.2067429 + y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2));
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * t));
    }
    const double ax = fabs(x), y = 3.75 / ax;
    const double bx = exp(ax) / sqrt(ax);
    const double t = -0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2));
    const double p = 0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 +
                     y * (-0.157565e-2 + y * (0.916281e-2 + y * t))));
    return p * bx;
}

static double bessel_i1(double x) {
    if (fabs(x) < 3.75) {
        double y = x / 3.75;
        y = y * y;
        const double t = 0.15084934 + y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3));
        return x * (0.5 + y * (0.87890594 + y * (0.51498869 + y * t)));
    }
    const double ax = fabs(x), y = 3.75 / ax;
    const double bx = exp(ax) / sqrt(ax);
    const double t = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 + y * -0.420059e-2));
    const double p = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 +
                     y * (0.163801e-2 + y * (-0.1031555e-1 + y * t))));
    return p * bx;
}

double centre_bessel_i(int n, double x) {
    if (n == 0) return bessel_i0(x);
    if (n == 1) return bessel_i1(x);
    if (fabs(x) < 1.e-10) return 0.0;
    const double tox = 2.0 / x;
    double bip = 0.0, bi = 1.0, res = 0.0;
    const int iters = 2 * (n + (int)sqrtf((float)(40 * n)));
    for (int i = iters; i >= 1; --i) {
        const double bim = bip + (double)(i * tox * bi);
        bip = bi;
        bi = bim;
        if (fabs(bi) > 1.e10) {
            bi *= 1.e-10;
            bip *= 1.e-10;
            res *= 1.e-10;
        }
        if (i == n) res = bip;
    }
    return res * bessel_i0(x) / bi;
}

// Kernel bandwidth of getMinimaPositions (Real2Int.cpp:509-514).
double centre_vm_bandwidth(uint64_t n, double kappa) {
    const double i0 = centre_bessel_i(0, kappa);
    return pow((3 * (double)n * (kappa * kappa) * centre_bessel_i(2, 2 * kappa) /
                (4 * 1.7724538509055159 * pow(i0, 2.0))),
               (2.0 / 5.0));
}

extern "C" int centre_dev_discretize_vmkde(centre_ctx* ctx, const double* d_bat, uint64_t ld,
                                           uint32_t n_dof, uint64_t n_frame, int nmaxconf,
                                           double kappa, uint32_t sdo_step, double sdo_conv,
                                           uint32_t sdo_iter, uint32_t grid_intervals,
                                           uint8_t* d_codes, double* d_minima,
                                           uint8_t* d_nfound, double* d_grid) {
    if (!ctx) return -3;
    if (!d_bat || !d_codes || !d_minima || !d_nfound) return fail(ctx, -3, "null buffer");
    if (nmaxconf < 1 || nmaxconf > 9) return fail(ctx, -3, "nmaxconf must be in 1..9");
    if (grid_intervals < 2 || grid_intervals > 3600)
        return fail(ctx, -3, "grid_intervals must be in 2..3600");
    if (ld < n_dof || n_frame == 0) return fail(ctx, -3, "ld < n_dof or no frames");
    if (n_dof == 0) return 0;
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf grid;
    double* g = d_grid;
    if (!g) {
        HIPCHK(grid.alloc(sizeof(double) * 2 * (grid_intervals + 1) * (size_t)n_dof));
        g = grid.as<double>();
    }
    void *xt = nullptr, *trig = nullptr;
    HIPCHK(ctx_scratch(ctx, 0, sizeof(float) * n_frame * (size_t)n_dof, &xt));
    HIPCHK(ctx_scratch(ctx, 1, sizeof(double) * 2 * n_frame * (size_t)n_dof, &trig));
    VmkdeArgs a{};
    a.bat = d_bat;
    a.ld = ld;
    a.n_dof = n_dof;
    a.n_frame = n_frame;
    a.kappa = kappa;
    a.nmaxconf = nmaxconf;
    a.sdo_step = sdo_step;
    a.sdo_conv = sdo_conv;
    a.sdo_iter = sdo_iter;
    a.grid = (int)grid_intervals;
    a.gridbuf = g;
    a.minima = d_minima;
    a.nfound = d_nfound;
    a.codes = d_codes;
    a.xt = static_cast<float*>(xt);
    a.trig = static_cast<double*>(trig);
    HIPCHK(run_vmkde(a, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int centre_discretize_vmkde(centre_ctx* ctx, const double* bat, uint64_t ld,
                                       uint32_t n_dof, uint64_t n_frame, int nmaxconf,
                                       double kappa, uint32_t sdo_step, double sdo_conv,
                                       uint32_t sdo_iter, uint32_t grid_intervals,
                                       uint8_t* codes, double* minima, uint8_t* n_found,
                                       double* grid) {
    if (!ctx) return -3;
    if (!bat || !codes || !minima || !n_found) return fail(ctx, -3, "null buffer");
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf b, c, m, nf, gr;
    HIPCHK(b.alloc(sizeof(double) * ld * n_frame));
    HIPCHK(c.alloc((size_t)n_dof * n_frame));
    HIPCHK(m.alloc(sizeof(double) * 9 * n_dof));
    HIPCHK(nf.alloc(n_dof));
    const size_t gbytes = sizeof(double) * 2 * (grid_intervals + 1) * (size_t)n_dof;
    if (grid) HIPCHK(gr.alloc(gbytes));
    HIPCHK(hipMemcpyAsync(b.p, bat, sizeof(double) * ld * n_frame, hipMemcpyHostToDevice,
                          ctx->stream));
    int r = centre_dev_discretize_vmkde(ctx, b.as<double>(), ld, n_dof, n_frame, nmaxconf, kappa,
                                        sdo_step, sdo_conv, sdo_iter, grid_intervals,
                                        c.as<uint8_t>(), m.as<double>(), nf.as<uint8_t>(),
                                        grid ? gr.as<double>() : nullptr);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(codes, c.p, (size_t)n_dof * n_frame, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(minima, m.p, sizeof(double) * 9 * n_dof, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(n_found, nf.p, n_dof, hipMemcpyDeviceToHost, ctx->stream));
    if (grid) HIPCHK(hipMemcpyAsync(grid, gr.p, gbytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---------------------------------------------------------------------------
extern "C" int centre_dev_mst(centre_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst,
                              const double* d_w, uint64_t n_edge, uint32_t n_node,
                              uint32_t* tree_edge_idx, uint32_t* n_tree) {
    if (!ctx) return -3;
    if (!n_tree || (!tree_edge_idx && n_node > 1)) return fail(ctx, -3, "null output");
    if (n_edge && (!d_src || !d_dst || !d_w)) return fail(ctx, -3, "null edge buffer");
    if (n_edge >= 0xFFFFFFFFull) return fail(ctx, -3, "too many edges (max 2^32-2)");
    *n_tree = 0;
    if (n_node == 0) return 0;
    HIPCHK(hipSetDevice(ctx->device));
    // round state in one scratch region: best weight (8 B per node), then
    // component and best edge (4 B each), the change flag, edge flags
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_comp = up(sizeof(unsigned long long) * n_node);
    const size_t o_be = o_comp + up(sizeof(uint32_t) * n_node);
    const size_t o_ch = o_be + up(sizeof(unsigned int) * n_node);
    const size_t o_ti = o_ch + 256;                       // tree edge indices
    const size_t o_tw = o_ti + up(sizeof(uint32_t) * n_node);   // their weights
    const size_t o_pt = o_tw + up(sizeof(double) * n_node);     // per-workgroup maxima
    const size_t o_it = o_pt + up(mst_part_bytes(n_node));
    void* mem = nullptr;
    HIPCHK(ctx_scratch(ctx, 3, o_it + (n_edge ? n_edge : 1), &mem));
    uint8_t* const base = static_cast<uint8_t*>(mem);
    MstArgs a{};
    a.src = d_src;
    a.dst = d_dst;
    a.w = d_w;
    a.n_edge = n_edge;
    a.n_node = n_node;
    a.best_w = reinterpret_cast<unsigned long long*>(base);
    a.comp = reinterpret_cast<uint32_t*>(base + o_comp);
    a.best_e = reinterpret_cast<unsigned int*>(base + o_be);
    a.changed = reinterpret_cast<int*>(base + o_ch);
    a.in_tree = base + o_it;
    a.part = mst_part_bytes(n_node) ? reinterpret_cast<unsigned long long*>(base + o_pt) : nullptr;
    {
        // endpoint range check before any kernel indexes by an endpoint
        int bad = 0;
        HIPCHK(hipMemsetAsync(a.changed, 0, sizeof(int), ctx->stream));
        HIPCHK(mst_check(a, ctx->stream));
        HIPCHK(hipMemcpyAsync(&bad, a.changed, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (bad) return fail(ctx, -3, "edge endpoint out of range");
    }
    HIPCHK(mst_init(a, ctx->stream));
    for (uint32_t round = 0; round < 64; ++round) {
        HIPCHK(run_mst_round(a, ctx->stream));
        int changed = 0;
        HIPCHK(hipMemcpyAsync(&changed, a.changed, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (!changed) break;
    }
    // tree edges compacted on the device: a forest has at most n_node - 1
    int n_in = 0;
    HIPCHK(hipMemsetAsync(a.changed, 0, sizeof(int), ctx->stream));
    uint32_t* const d_ti = reinterpret_cast<uint32_t*>(base + o_ti);
    double* const d_tw = reinterpret_cast<double*>(base + o_tw);
    HIPCHK(mst_collect(a, d_ti, d_tw, n_node, ctx->stream));
    HIPCHK(hipMemcpyAsync(&n_in, a.changed, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (n_in < 0 || (uint64_t)n_in >= n_node)
        return fail(ctx, -3, "internal: forest larger than n_node - 1");
    std::vector<uint32_t> idx(n_in);
    std::vector<double> wt(n_in);
    if (n_in) {
        HIPCHK(hipMemcpyAsync(idx.data(), d_ti, sizeof(uint32_t) * n_in, hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipMemcpyAsync(wt.data(), d_tw, sizeof(double) * n_in, hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    std::vector<size_t> ord(idx.size());
    for (size_t k = 0; k < ord.size(); ++k) ord[k] = k;
    std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
        if (wt[x] != wt[y]) return wt[x] > wt[y];
        return idx[x] < idx[y];
    });
    for (size_t k = 0; k < ord.size(); ++k) tree_edge_idx[k] = idx[ord[k]];
    *n_tree = (uint32_t)idx.size();
    return 0;
}

extern "C" int centre_mst(centre_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                          const double* w, uint64_t n_edge, uint32_t n_node,
                          uint32_t* tree_edge_idx, uint32_t* n_tree) {
    if (!ctx) return -3;
    if (n_edge && (!src || !dst || !w)) return fail(ctx, -3, "null edge buffer");
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf s, d, ww;
    HIPCHK(s.alloc(sizeof(uint32_t) * n_edge));
    HIPCHK(d.alloc(sizeof(uint32_t) * n_edge));
    HIPCHK(ww.alloc(sizeof(double) * n_edge));
    if (n_edge) {
        HIPCHK(hipMemcpy(s.p, src, sizeof(uint32_t) * n_edge, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d.p, dst, sizeof(uint32_t) * n_edge, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ww.p, w, sizeof(double) * n_edge, hipMemcpyHostToDevice));
    }
    return centre_dev_mst(ctx, s.as<uint32_t>(), d.as<uint32_t>(), ww.as<double>(), n_edge,
                          n_node, tree_edge_idx, n_tree);
}

extern "C" int centre_dev_segment_sum(centre_ctx* ctx, const double* d_x,
                                      const uint64_t* seg_begin, int n_seg, double* out) {
    if (!ctx) return -3;
    if (n_seg <= 0) return 0;
    if (!d_x || !seg_begin || !out) return fail(ctx, -3, "null buffer");
    for (int s = 0; s < n_seg; ++s)
        if (seg_begin[s + 1] < seg_begin[s]) return fail(ctx, -3, "segments must be ordered");
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf b, o;
    HIPCHK(b.alloc(sizeof(uint64_t) * (n_seg + 1)));
    HIPCHK(o.alloc(sizeof(double) * n_seg));
    HIPCHK(hipMemcpyAsync(b.p, seg_begin, sizeof(uint64_t) * (n_seg + 1), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(launch_segment_sum(d_x, b.as<uint64_t>(), n_seg, o.as<double>(), ctx->stream));
    HIPCHK(hipMemcpyAsync(out, o.p, sizeof(double) * n_seg, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Device memory plumbing for callers without their own allocator.
extern "C" int centre_dev_alloc(centre_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return -3;
    *out = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMalloc(out, bytes ? bytes : 16));
    return 0;
}

extern "C" int centre_dev_free(centre_ctx* ctx, void* p) {
    if (!ctx) return -3;
    if (!p) return 0;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipFree(p));
    return 0;
}

extern "C" int centre_memcpy_h2d(centre_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
    if (!ctx) return -3;
    if (!bytes) return 0;
    if (!d_dst || !src) return fail(ctx, -3, "null pointer");
    HIPCHK(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int centre_memcpy_d2d(centre_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    if (!ctx) return -3;
    if (!bytes) return 0;
    if (!d_dst || !d_src) return fail(ctx, -3, "null pointer");
    HIPCHK(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int centre_memcpy_d2d_2d(centre_ctx* ctx, void* d_dst, size_t dpitch, const void* d_src,
                                    size_t spitch, size_t width, size_t height) {
    if (!ctx) return -3;
    if (!width || !height) return 0;
    if (!d_dst || !d_src) return fail(ctx, -3, "null pointer");
    if (dpitch < width || spitch < width) return fail(ctx, -3, "pitch < width");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(launch_copy_2d(d_dst, dpitch, d_src, spitch, width, height, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int centre_memcpy_d2h(centre_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
    if (!ctx) return -3;
    if (!bytes) return 0;
    if (!dst || !d_src) return fail(ctx, -3, "null pointer");
    HIPCHK(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int centre_device_sync(centre_ctx* ctx) {
    if (!ctx) return -3;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipDeviceSynchronize());
    return 0;
}
