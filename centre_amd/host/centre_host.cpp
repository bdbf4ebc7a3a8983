// centre_hip: the C++ host of the MI355X entropy path, a drop-in for the
// three stages of centre_OMP (main_smp.cpp:22-373): discretize, entropy
// contributions, scoring (MIE / MIST).  Reads the same input file, the same
// BAT trajectory and writes the same outputs:
//   bin_{bonds,angles,torsions}.nc   (Netcdf_TrjInt.cpp:542-755, CDF-5)
//   entcontri_{bnd,ang,tor}-1d.nc, entcontri_{bnd,ang,tor,ba,bd,ad}-2d.nc
//        (Netcdf_EntContri.cpp:577-760; written CDF-5 with the same
//         dimensions, variables and attributes -- netcdf-c reads both)
//   hist_*.nc when histogram.writefreq is on (Netcdf_HistUtils.cpp:808-1029)
//   MIE-/MIST-<est>_convergence_b-<B>.<ext>, Tree-MIST-<est>_b-<B>_s-<k>.<ext>
//        (EntropyScorer.cpp:571-865)
//   vmkde_torsions_minimas.out for the von Mises KDE discretizer.
// All numerics run through the C ABI of libcentre_hip.so on one GPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/centre_hip.h"
#include "config.hpp"
#include "nc_classic.hpp"
#include "ncfile.hpp"

using namespace centre;

namespace {

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void ck(centre_ctx* ctx, int rc, const char* what) {
    if (rc != 0) {
        std::ostringstream o;
        o << what << " failed (" << rc << "): " << centre_last_error(ctx);
        throw Fail(o.str());
    }
}

struct Dev {
    centre_ctx* ctx = nullptr;
    void* p = nullptr;
    Dev(centre_ctx* c, size_t bytes) : ctx(c) {
        if (bytes) ck(ctx, centre_dev_alloc(ctx, bytes, &p), "alloc");
    }
    ~Dev() { if (p) centre_dev_free(ctx, p); }
    Dev(const Dev&) = delete;
    template <class T> T* as() const { return static_cast<T*>(p); }
    void put(const void* src, size_t n) { ck(ctx, centre_memcpy_h2d(ctx, p, src, n), "h2d"); }
    void get(void* dst, size_t n) const { ck(ctx, centre_memcpy_d2h(ctx, dst, p, n), "d2h"); }
};

std::string join(const std::string& dir, const std::string& f) {
    if (dir.empty()) return f;
    return dir.back() == '/' ? dir + f : dir + "/" + f;
}

bool file_exists(const std::string& p) { return std::ifstream(p).good(); }

const char* TYPE_VAR[3] = {"bond", "angle", "dihedral"};
const char* BIN_FILE[3] = {"bin_bonds.nc", "bin_angles.nc", "bin_torsions.nc"};
const char* ONE_TAG[3] = {"bnd", "ang", "tor"};
const char* ONE_CONTENT[3] = {"B", "A", "D"};   // entcontri-1d content
const char* BIN_CONTENT[3] = {"B", "A", "T"};   // bin file content (Discretizer.cpp:40-60)
const uint32_t ONE_SET[3] = {B1D, A1D, D1D};
const char* EST_NAME[4] = {"ML", "MM", "CS", "JS"};

struct Kind {
    const char* tag;      // file tag
    const char* content;  // content attribute
    int a, b;             // row / column type
    uint32_t set;         // BATSet bit
};
const Kind KINDS[6] = {{"bnd", "B/B", 0, 0, BB2D}, {"ang", "A/A", 1, 1, AA2D},
                       {"tor", "D/D", 2, 2, DD2D}, {"ba", "B/A", 0, 1, BA2D},
                       {"bd", "B/D", 0, 2, BD2D}, {"ad", "A/D", 1, 2, AD2D}};
// MIST edge order: BB, BA, BD, AA, AD, DD (EntropyScorer.cpp:675-741)
const int EDGE_ORDER[6] = {0, 3, 4, 1, 5, 2};

// ---- subset / neighbour files (Centre.cpp:206-268, 459-593) -------------------
std::vector<std::vector<std::string>> records_of(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw Fail("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::vector<std::vector<std::string>> out;
    std::string rec;
    std::istringstream in(ss.str());
    while (std::getline(in, rec, ';')) {
        std::istringstream r(rec);
        std::vector<std::string> tok;
        std::string t;
        while (r >> t) tok.push_back(t);
        if (tok.empty() || tok[0][0] == '#') continue;
        out.push_back(tok);
    }
    return out;
}

using NeighMap = std::map<uint32_t, std::vector<uint32_t>>;   // sorted keys

// Record order of one pair kind.  Rows: blocks of `block` neighbour keys.
// Columns: blocks of block * procs subset members; each is split among the
// MPI ranks of the build being emulated, ranks 1.. first (ceil(bc / procs)
// columns each when the block is partial), rank 0 (the master) last, and a
// rank's records are its rows in key order, each row's neighbours in list
// order, kept when inside its column range.  procs = 1 is centre_OMP
// (EntropyCalculator.cpp:406-473, 827-890); procs > 1 is centre_OMPMPI
// (EntropyCalculator_mpi.cpp:815-1110, 1952-2100), whose order the stored
// example runs carry (two ranks).
std::vector<std::pair<uint32_t, uint32_t>> record_order(const NeighMap& nm,
                                                        const std::vector<uint32_t>& cols,
                                                        unsigned block, unsigned procs) {
    std::vector<uint32_t> keys;
    for (auto& kv : nm) keys.push_back(kv.first);
    std::vector<std::pair<uint32_t, uint32_t>> out;
    const size_t cached = (size_t)block * procs;
    for (size_t bi = 0; bi < keys.size(); bi += block) {
        const size_t br = std::min<size_t>(block, keys.size() - bi);
        for (size_t bj = 0; bj < cols.size(); bj += cached) {
            const size_t bc = std::min<size_t>(cached, cols.size() - bj);
            size_t per = block, ranks = procs;
            if (bc < cached) {
                per = (bc + procs - 1) / procs;
                ranks = (bc + per - 1) / per;
            }
            for (size_t p = 1; p <= ranks; ++p) {
                const size_t c0 = bj + (p - 1) * per, c1 = p < ranks ? bj + p * per : bj + bc;
                const uint32_t lo = cols[c0], hi = cols[c1 - 1];
                for (size_t r = bi; r < bi + br; ++r)
                    for (uint32_t cj : nm.at(keys[r]))
                        if (cj >= lo && cj <= hi) out.emplace_back(keys[r], cj);
            }
        }
    }
    return out;
}

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double s() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// ---- output writers ---------------------------------------------------------------
void write_bin_file(const std::string& path, int type, const std::vector<uint8_t>& schemes,
                    uint32_t n_dof, uint64_t F, const std::vector<uint8_t>& codes,
                    const std::vector<double>& ext) {
    nc::Writer w(5);
    const int dd = w.add_dim("dims", n_dof, true);
    const int de = w.add_dim("extremes", 4);
    const int ds = w.add_dim("bin_scheme_count", schemes.size());
    const int df = w.add_dim("frames", F);
    const std::string title = std::string(BIN_CONTENT[type]) + " Integer Vectors";
    w.add_attr("title", nc::Attr::of_text(title));
    w.add_attr("application", nc::Attr::of_text("CENTRE"));
    w.add_attr("program", nc::Attr::of_text("Real2Int"));
    w.add_attr("programVersion", nc::Attr::of_text("1.0.0"));
    w.add_attr("convention", nc::Attr::of_text("CENTRE_Real"));
    w.add_attr("conventionVersion", nc::Attr::of_text("1.0"));
    w.add_attr("content", nc::Attr::of_text(BIN_CONTENT[type]));
    w.add_attr("ncoordims", nc::Attr::of<uint32_t>(nc::UINT, {n_dof}));
    w.add_attr("nbinSchemes", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)schemes.size()}));
    w.add_attr("binSchemes", nc::Attr::of<uint8_t>(nc::UBYTE, schemes));
    std::vector<uint32_t> ids(n_dof);
    std::iota(ids.begin(), ids.end(), 0u);
    w.add_var("dimid", nc::UINT, {dd}, ids.data());
    w.add_var("extreme", nc::DOUBLE, {dd, de}, ext.data());
    w.add_var("coord", nc::UBYTE, {dd, ds, df}, codes.data());
    w.write(path);
}

// Entropy-contribution file (Netcdf_EntContri.cpp:577-760).  The reference
// creates it NC_NETCDF4 with contri_S chunked (16, estimators, schemes,
// steps) and id chunked (1, order) (netcdf-c's default for a record
// variable); cdf5 = true writes the same dims, vars and attrs as CDF-5.
void write_entcontri(const std::string& path, const std::string& content, int order,
                     bool upper, const std::vector<uint64_t>& dim_lens,
                     const std::vector<uint8_t>& schemes, int nsteps, int nest,
                     const std::vector<uint32_t>& ids, const std::vector<double>& contri,
                     bool netcdf4) {
    nc::Writer w(5);
    const uint64_t nrec = ids.size() / order;
    const int dr = w.add_dim("records", nrec, true);
    const int de = w.add_dim("estimators", nest);
    const int db = w.add_dim("nbin_schemes", schemes.size());
    const int dn = w.add_dim("nsteps", nsteps);
    const int di = w.add_dim("ids", order);
    // "title" is written twice by the reference (the second write, the
    // diagonal flag, wins: Netcdf_EntContri.h:32,38)
    w.add_attr("title", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)(upper ? 1 : 0)}));
    w.add_attr("application", nc::Attr::of_text("CENTRE"));
    w.add_attr("program", nc::Attr::of_text("Real2Int"));
    w.add_attr("programVersion", nc::Attr::of_text("1.0.0"));
    w.add_attr("convention", nc::Attr::of_text("EntropyCont"));
    w.add_attr("conventionVersion", nc::Attr::of_text("1.0"));
    w.add_attr("content", nc::Attr::of_text(content));
    w.add_attr("order", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)order}));
    w.add_attr("nsteps", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)nsteps}));
    w.add_attr("storagetype", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)(upper ? 3 : 0)}));
    w.add_attr("ndims", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)order}));
    w.add_attr("dimLengths", nc::Attr::of<uint64_t>(nc::UINT64, dim_lens));
    w.add_attr("nbinSchemes", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)schemes.size()}));
    w.add_attr("binSchemes", nc::Attr::of<uint8_t>(nc::UBYTE, schemes));
    w.add_var("id", nc::UINT, {dr, di}, ids.data());
    w.add_var("contri_S", nc::DOUBLE, {dr, de, db, dn}, contri.data());
    if (!netcdf4) {
        w.write(path);
        return;
    }
    w.set_chunks("id", {1, (uint64_t)order});
    w.set_chunks("contri_S", {16, (uint64_t)nest, schemes.size(), (uint64_t)nsteps});
    ncf::write_netcdf4(w, path);
}

// Bin mids as binData1D / binData2D accumulate them (Estimators.cpp:67-77,
// 140-160): the offset grows by one bin width per bin, which is not bit for
// bit min + i*w.
void append_mids(std::vector<double>& out, double mn, double mx, int nb) {
    const double w = (mx - mn) / nb, hw = w / 2.0;
    double o = mn;
    for (int i = 0; i < nb; ++i) {
        out.push_back(o + hw);
        o += w;
    }
}

// Count file of one kind (Netcdf_HistUtils.cpp:216-262 constructor clamps,
// 808-1029 NC_create; records as writeRecords, 714-772).  ndiag shares the
// attribute name "title" and overwrites it in place (Netcdf_HistUtils.h:26,32).
struct HistFile {
    std::string content;
    int ndims = 1;
    int storage = 0;                    // TensorType: FULL 0, UPPER 3
    std::vector<uint64_t> dim_lens;
    std::vector<uint32_t> ids;          // [rec][ndims]
    std::vector<double> extrema;        // [rec][ndims][4]
    std::vector<double> mids;           // [rec][ndims * sum(schemes)]
    std::vector<uint64_t> freq;         // [rec][steps_eff][sum(schemes^ndims)]
};

void write_hist(const std::string& path, const HistFile& h, const std::vector<uint8_t>& schemes,
                int first_step, int nsteps, int stride, int steps_eff) {
    uint64_t bsum = 0, bpow = 0;
    for (uint8_t b : schemes) {
        bsum += b;
        bpow += h.ndims == 1 ? (uint64_t)b : (uint64_t)b * b;
    }
    nc::Writer w(5);
    const uint64_t nrec = h.ids.size() / h.ndims;
    const int dr = w.add_dim("records", nrec, true);
    const int dm = w.add_dim("dimsXbin_sum", bsum * h.ndims);
    const int dp = w.add_dim("bin_pow_dims_sum", bpow);
    const int dn = w.add_dim("nsteps_eff", steps_eff);
    const int di = w.add_dim("ids", h.ndims);
    const int de = w.add_dim("extremes", 4);
    w.add_attr("title", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)(h.storage != 0 ? 1 : 0)}));
    w.add_attr("appication", nc::Attr::of_text("CENTRE"));
    w.add_attr("program", nc::Attr::of_text("Real2Int"));
    w.add_attr("programVersion", nc::Attr::of_text("1.0.0"));
    w.add_attr("convention", nc::Attr::of_text("CENTRE_HIST"));
    w.add_attr("conventionVersion", nc::Attr::of_text("1.0"));
    w.add_attr("content", nc::Attr::of_text(h.content));
    w.add_attr("order", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)h.ndims}));
    w.add_attr("firstStep", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)first_step}));
    w.add_attr("nsteps", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)nsteps}));
    w.add_attr("stepStride", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)stride}));
    w.add_attr("storagetype", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)h.storage}));
    w.add_attr("ndims", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)h.ndims}));
    w.add_attr("dimLengths", nc::Attr::of<uint64_t>(nc::UINT64, h.dim_lens));
    w.add_attr("nbinSchemes", nc::Attr::of<uint8_t>(nc::UBYTE, {(uint8_t)schemes.size()}));
    w.add_attr("binSchemes", nc::Attr::of<uint8_t>(nc::UBYTE, schemes));
    w.add_var("id", nc::UINT, {dr, di}, h.ids.data());
    w.add_var("extrema", nc::DOUBLE, {dr, di, de}, h.extrema.data());
    w.add_var("mid", nc::DOUBLE, {dr, dm}, h.mids.data());
    w.add_var("freq", nc::UINT64, {dr, dn, dp}, h.freq.data());
    w.write(path);
}

// ---- the three stages -------------------------------------------------------------
struct TypeData {
    uint32_t n_all = 0;                 // DOFs of this type in the BAT file
    std::vector<uint8_t> codes;         // [n_all][nsch][F]
    std::vector<double> ext;            // [n_all][4]
    std::vector<uint32_t> subset;       // member ids
};

struct Host {
    Inputs in;
    centre_ctx* ctx = nullptr;
    uint64_t F_file = 0;
    TypeData td[3];
    std::map<std::string, double> times;
    std::ofstream* log = nullptr;

    std::string out(const std::string& f) const { return join(in.outfilepath, f); }

    // Frame shuffle index (main_smp.cpp:194-201, 312-325): std::shuffle of
    // 0..F-1 with mt19937(randseed + 1000), randseed 0 = random_device.
    std::vector<uint64_t> shuffle_index(uint64_t F) {
        std::random_device rd;
        uint32_t seed = rd();
        if (in.randseed != 0) seed = (uint32_t)(in.randseed + 1000);
        std::mt19937 gen(seed);
        std::vector<uint64_t> idx(F);
        std::iota(idx.begin(), idx.end(), 0ull);
        std::shuffle(idx.begin(), idx.end(), gen);
        std::ofstream ofp(out("shuffle_indices.bin"), std::ios::binary);
        ofp.write(reinterpret_cast<const char*>(idx.data()), idx.size() * sizeof(idx[0]));
        return idx;
    }

    // Discretizer::run (Discretizer.cpp:314-361).  DOF counts come from the
    // input file (discretization.nbond/nangle/ndihed, first DOFs of the BAT
    // file), frames from discretization.nframe (all frames when absent).
    void discretize(const ncf::File& bat) {
        Timer t;
        const std::vector<uint8_t> schemes = in.nbins;
        std::ofstream minfile;
        if (in.pdfmethod == PdfMethod::VonMisesKDE) minfile.open(out("vmkde_torsions_minimas.out"));
        uint64_t F = F_file;
        if (!in.nframes.empty()) {
            F = 0;
            for (auto v : in.nframes) F += v;
            if (F > F_file) throw Fail("discretization.nframe exceeds the frames of the BAT file");
        }
        // Real2Int::shuffleArray swaps over codes of F frames; with several
        // bin schemes the code vector is longer than the index and the
        // reference leaves it unshuffled (Real2Int.cpp:12-25)
        std::vector<uint64_t> shuf;
        if (in.shuffleframes) shuf = shuffle_index(F);
        const unsigned ncfg[3] = {in.nbond, in.nangle, in.ndihed};
        for (int ty = 0; ty < 3; ++ty) {
            TypeData& T = td[ty];
            const uint32_t nfile = bat.has_var(TYPE_VAR[ty]) ? (uint32_t)bat.shape(TYPE_VAR[ty])[1] : 0;
            T.n_all = ncfg[ty] ? std::min<uint32_t>(ncfg[ty], nfile) : nfile;
            if (!T.n_all) continue;
            std::vector<double> x = bat.read_f64(TYPE_VAR[ty], {0, 0}, {F, nfile});
            if (T.n_all != nfile) {   // keep the first n_all DOFs of each frame
                for (uint64_t f = 0; f < F; ++f)
                    std::copy_n(&x[f * nfile], T.n_all, &x[f * T.n_all]);
                x.resize(F * T.n_all);
            }
            Dev dx(ctx, x.size() * 8);
            dx.put(x.data(), x.size() * 8);
            const bool vm = in.pdfmethod == PdfMethod::VonMisesKDE && ty == 2;
            const size_t nsch = vm ? 1 : schemes.size();
            Dev dc(ctx, (size_t)T.n_all * nsch * F), de(ctx, (size_t)T.n_all * 32);
            T.codes.resize((size_t)T.n_all * nsch * F);
            T.ext.assign((size_t)T.n_all * 4, 1000.0);   // DEFAULT_BAT_VALUE
            if (!vm) {
                ck(ctx, centre_dev_discretize_hist(ctx, dx.as<double>(), T.n_all, T.n_all, F, ty + 1,
                                                   schemes.data(), (int)nsch, in.referencenbins,
                                                   in.optimizedih ? 1 : 0, dc.as<uint8_t>(),
                                                   de.as<double>()),
                   "discretize_hist");
                de.get(T.ext.data(), T.ext.size() * 8);
            } else {
                Dev dm(ctx, (size_t)T.n_all * 9 * 8), dn(ctx, T.n_all);
                ck(ctx, centre_dev_discretize_vmkde(ctx, dx.as<double>(), T.n_all, T.n_all, F,
                                                    in.nmaxconf, in.kappa, in.sdosteps,
                                                    in.sdoconvlimit, in.sdoiterations, 360,
                                                    dc.as<uint8_t>(), dm.as<double>(),
                                                    dn.as<uint8_t>(), nullptr),
                   "discretize_vmkde");
                std::vector<double> mins((size_t)T.n_all * 9);
                std::vector<uint8_t> nf(T.n_all);
                dm.get(mins.data(), mins.size() * 8);
                dn.get(nf.data(), nf.size());
                for (uint32_t d = 0; d < T.n_all; ++d) {
                    minfile << "torsion " << std::setw(6) << d << " " << (int)nf[d];
                    for (int k = 0; k < nf[d]; ++k)
                        minfile << " " << std::fixed << std::setprecision(8) << mins[d * 9 + k];
                    minfile << ";\n";
                }
            }
            dc.get(T.codes.data(), T.codes.size());
            if (!shuf.empty() && nsch == 1)
                for (uint32_t d = 0; d < T.n_all; ++d) {
                    uint8_t* row = &T.codes[(size_t)d * F];
                    for (uint64_t i = 0; i < F; ++i) std::swap(row[i], row[shuf[i]]);
                }
            write_bin_file(out(BIN_FILE[ty]), ty,
                           vm ? std::vector<uint8_t>{(uint8_t)in.nmaxconf} : schemes, T.n_all, F,
                           T.codes, T.ext);
        }
        F_file = F;
        times["discretize"] = t.s();
    }

    void load_bins() {
        for (int ty = 0; ty < 3; ++ty) {
            TypeData& T = td[ty];
            if (!file_exists(out(BIN_FILE[ty]))) continue;
            auto r = ncf::open(out(BIN_FILE[ty]));
            auto shp = r->shape("coord");
            T.n_all = (uint32_t)shp[0];
            F_file = shp[2];
            T.codes = r->read_u8("coord");
            T.ext = r->read_f64("extreme");
        }
    }

    // codes of one type restricted to the entropy frames
    // (startframe / strideframe / nframe, Netcdf_TrjInt.cpp:353-359)
    std::vector<uint8_t> frames_of(int ty, uint32_t dof, int sch, uint64_t& Fe) const {
        const TypeData& T = td[ty];
        const size_t nsch = T.codes.size() / ((size_t)T.n_all * F_file);
        const uint8_t* row = T.codes.data() + ((size_t)dof * nsch + sch) * F_file;
        std::vector<uint8_t> v;
        const uint64_t last = std::min<uint64_t>(F_file, in.nframe);
        for (uint64_t f = in.startframe; f < last; f += in.strideframe) v.push_back(row[f]);
        Fe = v.size();
        return v;
    }

    void run_entropy();
    void run_scoring();
};

// EntropyCalculator::run (EntropyCalculator.cpp:1090-1140)
void Host::run_entropy() {
    Timer t;
    // subset and neighbours (Inputs.cpp:676-712)
    NeighMap nm[6];
    for (int ty = 0; ty < 3; ++ty) {
        td[ty].subset.clear();
        if (!(in.workset & ONE_SET[ty])) continue;
        for (uint32_t i = 0; i < td[ty].n_all; ++i) td[ty].subset.push_back(i);
    }
    if (in.usesubset) {
        for (int ty = 0; ty < 3; ++ty) td[ty].subset.clear();
        for (auto& r : records_of(join(in.infilepath, in.subsetfile))) {
            for (int ty = 0; ty < 3; ++ty) {
                static const char* w[3] = {"bond", "angle", "torsion"};
                if (r[0] != w[ty] || !(in.workset & ONE_SET[ty])) continue;
                const size_t n = std::stoul(r.at(2));
                for (size_t k = 0; k < n; ++k) td[ty].subset.push_back((uint32_t)std::stoul(r.at(3 + k)));
            }
        }
    }
    if (in.usesubset && in.useneighbor) {
        static const char* w[6] = {"bond", "angle", "torsion", "bacross", "bdcross", "adcross"};
        for (auto& r : records_of(join(in.infilepath, in.neighborfile))) {
            const size_t n = std::stoul(r.at(2));
            if (n == 0) continue;
            for (int k = 0; k < 6; ++k) {
                if (r[0] != w[k] || !(in.workset & KINDS[k].set)) continue;
                std::vector<uint32_t> v;
                for (size_t q = 0; q < n; ++q) v.push_back((uint32_t)std::stoul(r.at(3 + q)));
                nm[k].emplace((uint32_t)std::stoul(r[1]), v);
            }
        }
    } else {
        for (int k = 0; k < 6; ++k) {
            if (!(in.workset & KINDS[k].set)) continue;
            const auto& rows = td[KINDS[k].a].subset;
            const auto& cols = td[KINDS[k].b].subset;
            if (KINDS[k].a == KINDS[k].b) {
                for (size_t i = 0; i + 1 < rows.size(); ++i)
                    nm[k][rows[i]] = std::vector<uint32_t>(rows.begin() + i + 1, rows.end());
            } else {
                for (auto r : rows) nm[k][r] = cols;
            }
        }
    }
    const int S = in.nsteps, nest = (int)in.estimators.size();
    const std::vector<uint8_t> schemes = in.nbins;
    const int nsch = (int)schemes.size();
    const bool kde = in.pdfmethod == PdfMethod::VonMisesKDE;
    const unsigned block = std::max(1u, in.cacheentcdimspercpu * (unsigned)in.threads_for_order / 2);

    // concatenated device code matrix of every subset DOF, per scheme
    std::vector<uint32_t> base(3, 0);
    uint32_t Dall = 0;
    for (int ty = 0; ty < 3; ++ty) { base[ty] = Dall; Dall += (uint32_t)td[ty].subset.size(); }
    if (Dall == 0) throw Fail("empty workset");
    uint64_t Fe = 0;
    frames_of(td[0].subset.empty() ? (td[1].subset.empty() ? 2 : 1) : 0,
              td[0].subset.empty() ? (td[1].subset.empty() ? td[2].subset[0] : td[1].subset[0]) : td[0].subset[0],
              0, Fe);
    if (Fe / S == 0) throw Fail("fewer frames than steps");
    std::vector<uint8_t> types(Dall);
    std::vector<double> ext((size_t)Dall * 4);
    for (int ty = 0; ty < 3; ++ty)
        for (size_t k = 0; k < td[ty].subset.size(); ++k) {
            types[base[ty] + k] = (uint8_t)(ty + 1);
            std::copy_n(&td[ty].ext[(size_t)td[ty].subset[k] * 4], 4, &ext[(size_t)(base[ty] + k) * 4]);
        }
    Dev dt(ctx, Dall), dx(ctx, (size_t)Dall * 32), dcodes(ctx, (size_t)Dall * Fe);
    dt.put(types.data(), Dall);
    dx.put(ext.data(), ext.size() * 8);

    // count files (EntropyCalculator.h:75-101; the step range clamps as the
    // Netcdf_HistUtil constructor does, Netcdf_HistUtils.cpp:225-231)
    const bool wfreq = kde ? in.vm_writefreq : in.hist_writefreq;
    const uint32_t wset = kde ? in.vm_writeset : in.hist_writeset;
    auto hist_on = [&](uint32_t set) { return wfreq && (wset & set) != 0; };
    const int hfirst = std::min(std::max(in.hist_writestepstart, 0), S - 1);
    const int hstride = std::min(std::max(in.hist_writestepstride, 1), S);
    const int Se = (S - hfirst + hstride - 1) / hstride;
    uint64_t bsum = 0, bsq = 0;
    std::vector<uint64_t> off1(nsch), off2(nsch);
    for (int sc = 0; sc < nsch; ++sc) {
        off1[sc] = bsum;
        off2[sc] = bsq;
        bsum += schemes[sc];
        bsq += (uint64_t)schemes[sc] * schemes[sc];
    }
    bool any1 = false;
    std::vector<uint64_t> hist1[3];     // [rec][Se][bsum]
    for (int ty = 0; ty < 3; ++ty)
        if (hist_on(ONE_SET[ty]) && !td[ty].subset.empty()) {
            any1 = true;
            hist1[ty].assign(td[ty].subset.size() * Se * bsum, 0);
        }

    // outputs gathered over schemes: S1 [dof][est][sch][S], S2 [rec][est][sch][S]
    std::vector<double> s1((size_t)Dall * nest * nsch * S);
    std::vector<std::vector<uint32_t>> ids2(6);
    std::vector<std::vector<uint32_t>> pairs(6);
    for (int k = 0; k < 6; ++k) {
        if (!(in.workset & KINDS[k].set) || nm[k].empty()) continue;
        std::map<uint32_t, uint32_t> ra, rb;
        for (size_t i = 0; i < td[KINDS[k].a].subset.size(); ++i) ra[td[KINDS[k].a].subset[i]] = (uint32_t)i;
        for (size_t i = 0; i < td[KINDS[k].b].subset.size(); ++i) rb[td[KINDS[k].b].subset[i]] = (uint32_t)i;
        for (auto& pr : record_order(nm[k], td[KINDS[k].b].subset, block,
                                     (unsigned)std::max(1, in.procs_for_order))) {
            if (!ra.count(pr.first) || !rb.count(pr.second)) continue;
            ids2[k].push_back(pr.first);
            ids2[k].push_back(pr.second);
            pairs[k].push_back(base[KINDS[k].a] + ra[pr.first]);
            pairs[k].push_back(base[KINDS[k].b] + rb[pr.second]);
        }
    }
    std::vector<std::vector<double>> s2(6);
    std::vector<std::vector<uint64_t>> hist2(6);   // [rec][Se][bsq]
    for (int k = 0; k < 6; ++k)
        if (hist_on(KINDS[k].set)) hist2[k].assign(pairs[k].size() / 2 * Se * bsq, 0);
    std::vector<double> kernel_ms;
    // Every pair of the concatenated subset DOFs (all six kinds, no neighbour
    // list, no 2-D count files): one all-pairs launch, whose triangle order
    // then gives each kind's records.  Otherwise one explicit plan per kind.
    uint64_t n_req = 0;
    for (int k = 0; k < 6; ++k) n_req += pairs[k].size() / 2;
    bool allpairs = !(in.usesubset && in.useneighbor) &&
                    n_req == (uint64_t)Dall * (Dall - 1) / 2 && Dall >= 2;
    for (int k = 0; k < 6; ++k) allpairs = allpairs && hist2[k].empty();
    for (int sc = 0; sc < nsch; ++sc) {
        std::vector<uint8_t> all((size_t)Dall * Fe);
        for (int ty = 0; ty < 3; ++ty)
            for (size_t k = 0; k < td[ty].subset.size(); ++k) {
                uint64_t fe;
                auto v = frames_of(ty, td[ty].subset[k], sc, fe);
                std::copy(v.begin(), v.end(), all.begin() + (size_t)(base[ty] + k) * Fe);
            }
        dcodes.put(all.data(), all.size());
        const int nb = schemes[sc];
        Dev ds1(ctx, (size_t)Dall * nest * S * 8);
        Dev dc1(ctx, any1 ? (size_t)Dall * S * nb * 8 : 0);
        ck(ctx, centre_dev_dof_entropy(ctx, dcodes.as<uint8_t>(), Fe, Dall, Fe, dt.as<uint8_t>(),
                                       dx.as<double>(), nb, S, in.estimators.data(), nest,
                                       in.jacobian, kde, ds1.as<double>(),
                                       any1 ? dc1.as<uint64_t>() : nullptr),
           "dof_entropy");
        if (any1) {
            std::vector<uint64_t> c1((size_t)Dall * S * nb);
            dc1.get(c1.data(), c1.size() * 8);
            for (int ty = 0; ty < 3; ++ty) {
                if (hist1[ty].empty()) continue;
                for (size_t r = 0; r < td[ty].subset.size(); ++r)
                    for (int e = 0; e < Se; ++e)
                        std::copy_n(&c1[((size_t)(base[ty] + r) * S + hfirst + e * hstride) * nb], nb,
                                    &hist1[ty][(r * Se + e) * bsum + off1[sc]]);
            }
        }
        std::vector<double> h1((size_t)Dall * nest * S);
        ds1.get(h1.data(), h1.size() * 8);
        for (uint32_t d = 0; d < Dall; ++d)
            for (int e = 0; e < nest; ++e)
                for (int s = 0; s < S; ++s)
                    s1[(((size_t)d * nest + e) * nsch + sc) * S + s] = h1[((size_t)d * nest + e) * S + s];
        if (allpairs) {
            const uint64_t PA = (uint64_t)Dall * (Dall - 1) / 2;
            centre_pair_plan* plan = nullptr;
            ck(ctx, centre_pair_plan_create(ctx, Dall, nb, nullptr, 0, &plan), "plan");
            Dev ds2(ctx, PA * nest * S * 8);
            const int rc = centre_dev_pair_entropy(ctx, plan, dcodes.as<uint8_t>(), Fe, Fe,
                                                   dt.as<uint8_t>(), dx.as<double>(), S,
                                                   in.estimators.data(), nest, in.jacobian, kde,
                                                   nullptr, ds2.as<double>(), nullptr, nullptr);
            centre_pair_plan_destroy(plan);
            ck(ctx, rc, "pair_entropy");
            kernel_ms.push_back(centre_last_pair_kernel_ms(ctx));
            std::vector<double> h2(PA * nest * S);
            ds2.get(h2.data(), h2.size() * 8);
            for (int k = 0; k < 6; ++k) {
                const uint64_t P = pairs[k].size() / 2;
                s2[k].resize(P * nest * nsch * S);
                for (uint64_t p = 0; p < P; ++p) {
                    // i < j: rows come from the earlier type (or the earlier DOF)
                    const uint64_t i = pairs[k][2 * p], j = pairs[k][2 * p + 1];
                    const uint64_t t = i * Dall - i * (i + 1) / 2 + (j - i - 1);
                    for (int e = 0; e < nest; ++e)
                        for (int s = 0; s < S; ++s)
                            s2[k][((p * nest + e) * nsch + sc) * S + s] = h2[(t * nest + e) * S + s];
                }
            }
            continue;
        }
        for (int k = 0; k < 6; ++k) {
            const uint64_t P = pairs[k].size() / 2;
            if (!P) continue;
            s2[k].resize(P * nest * nsch * S);
            // with a count file, records go in chunks of at most 512 MiB of counts
            const bool hk = !hist2[k].empty();
            const uint64_t nb2 = (uint64_t)nb * nb;
            const uint64_t chunk = hk ? std::max<uint64_t>(1, (512ull << 20) / (S * nb2 * 4)) : P;
            for (uint64_t r0 = 0; r0 < P; r0 += chunk) {
                const uint64_t n = std::min(chunk, P - r0);
                centre_pair_plan* plan = nullptr;
                ck(ctx, centre_pair_plan_create(ctx, Dall, nb, pairs[k].data() + 2 * r0, n, &plan),
                   "plan");
                Dev ds2(ctx, n * nest * S * 8);
                Dev dcn(ctx, hk ? n * S * nb2 * 4 : 0);
                const int rc = centre_dev_pair_entropy(ctx, plan, dcodes.as<uint8_t>(), Fe, Fe,
                                                       dt.as<uint8_t>(), dx.as<double>(), S,
                                                       in.estimators.data(), nest, in.jacobian, kde,
                                                       nullptr, ds2.as<double>(), nullptr,
                                                       hk ? dcn.as<uint32_t>() : nullptr);
                centre_pair_plan_destroy(plan);
                ck(ctx, rc, "pair_entropy");
                kernel_ms.push_back(centre_last_pair_kernel_ms(ctx));
                std::vector<double> h2(n * nest * S);
                ds2.get(h2.data(), h2.size() * 8);
                for (uint64_t p = 0; p < n; ++p)
                    for (int e = 0; e < nest; ++e)
                        for (int s = 0; s < S; ++s)
                            s2[k][(((r0 + p) * nest + e) * nsch + sc) * S + s] =
                                h2[(p * nest + e) * S + s];
                if (hk) {
                    std::vector<uint32_t> c2(n * S * nb2);
                    dcn.get(c2.data(), c2.size() * 4);
                    for (uint64_t p = 0; p < n; ++p)
                        for (int e = 0; e < Se; ++e)
                            std::copy_n(&c2[(p * S + hfirst + e * hstride) * nb2], nb2,
                                        &hist2[k][((r0 + p) * Se + e) * bsq + off2[sc]]);
                }
            }
        }
    }
    // 1-D files
    for (int ty = 0; ty < 3; ++ty) {
        if (!(in.workset & ONE_SET[ty]) || td[ty].subset.empty()) continue;
        const size_t n = td[ty].subset.size();
        std::vector<double> c(s1.begin() + (size_t)base[ty] * nest * nsch * S,
                              s1.begin() + (size_t)(base[ty] + n) * nest * nsch * S);
        write_entcontri(out(std::string("entcontri_") + ONE_TAG[ty] + "-1d.nc"), ONE_CONTENT[ty], 1,
                        false, {n}, schemes, S, nest, td[ty].subset, c, in.entcontri_netcdf4);
    }
    for (int k = 0; k < 6; ++k) {
        if (!(in.workset & KINDS[k].set)) continue;
        const bool same = KINDS[k].a == KINDS[k].b;
        write_entcontri(out(std::string("entcontri_") + KINDS[k].tag + "-2d.nc"), KINDS[k].content, 2,
                        same, {td[KINDS[k].a].subset.size(), td[KINDS[k].b].subset.size()}, schemes,
                        S, nest, ids2[k], s2[k], in.entcontri_netcdf4);
    }
    // count files (EntropyCalculator.cpp:61-92, 173-214, 362-372, 575-617,
    // 786-795): 1-D FULL, same-type 2-D FULL, cross-type 2-D UPPER
    for (int ty = 0; ty < 3; ++ty) {
        if (hist1[ty].empty()) continue;
        HistFile h;
        h.content = ONE_CONTENT[ty];
        h.ndims = 1;
        h.dim_lens = {td[ty].subset.size()};
        h.ids = td[ty].subset;
        for (size_t r = 0; r < td[ty].subset.size(); ++r) {
            const double* e = &ext[(size_t)(base[ty] + r) * 4];
            h.extrema.insert(h.extrema.end(), e, e + 4);
            for (int sc = 0; sc < nsch; ++sc) append_mids(h.mids, e[0], e[1], schemes[sc]);
        }
        h.freq = std::move(hist1[ty]);
        write_hist(out(std::string("hist_") + ONE_TAG[ty] + "-1d.nc"), h, schemes, hfirst, S,
                   hstride, Se);
    }
    for (int k = 0; k < 6; ++k) {
        if (!hist_on(KINDS[k].set)) continue;
        HistFile h;
        h.content = KINDS[k].content;
        h.ndims = 2;
        h.storage = KINDS[k].a == KINDS[k].b ? 0 : 3;
        h.dim_lens = {td[KINDS[k].a].subset.size(), td[KINDS[k].b].subset.size()};
        h.ids = ids2[k];
        for (size_t p = 0; p < pairs[k].size() / 2; ++p) {
            const double* ea = &ext[(size_t)pairs[k][2 * p] * 4];
            const double* eb = &ext[(size_t)pairs[k][2 * p + 1] * 4];
            h.extrema.insert(h.extrema.end(), ea, ea + 4);
            h.extrema.insert(h.extrema.end(), eb, eb + 4);
            for (int sc = 0; sc < nsch; ++sc) {
                append_mids(h.mids, ea[0], ea[1], schemes[sc]);
                append_mids(h.mids, eb[0], eb[1], schemes[sc]);
            }
        }
        h.freq = std::move(hist2[k]);
        write_hist(out(std::string("hist_") + KINDS[k].tag + "-2d.nc"), h, schemes, hfirst, S,
                   hstride, Se);
    }
    times["entropy"] = t.s();
    double kms = 0;
    for (double v : kernel_ms) kms += v;
    times["pair_kernels"] = kms / 1000.0;
}

// One entropy-contribution file read whole: ids [records][order] and
// contri_S [records][estimators][schemes][steps].  The scorer takes every
// (estimator, scheme, step) slice from it, so each file is opened and read
// once (a NetCDF-4 file's `id` is stored one record per chunk, which makes a
// read per slice as slow as the number of records).
struct Contri {
    std::vector<uint32_t> ids;
    std::vector<double> S;
    std::vector<uint64_t> shape;
    int order = 0;
};

const Contri& contri_file(std::map<std::string, Contri>& cache, const std::string& path) {
    auto it = cache.find(path);
    if (it != cache.end()) return it->second;
    auto r = ncf::open(path);
    Contri c;
    c.shape = r->shape("contri_S");
    c.order = (int)r->shape("id")[1];
    c.ids = r->read_u32("id");
    if (c.shape[0]) c.S = r->read_f64("contri_S");
    return cache.emplace(path, std::move(c)).first->second;
}

std::vector<double> contri_of(std::map<std::string, Contri>& cache, const std::string& path,
                              int est, int sch, int step, std::vector<uint32_t>& ids, int& order) {
    const Contri& c = contri_file(cache, path);
    ids = c.ids;
    order = c.order;
    const uint64_t n = c.shape[0], ne = c.shape[1], ns = c.shape[2], S = c.shape[3];
    std::vector<double> v(n);
    for (uint64_t r = 0; r < n; ++r) v[r] = c.S[((r * ne + est) * ns + sch) * S + step];
    return v;
}

// EntropyScorer::run (EntropyScorer.cpp:7-875)
void Host::run_scoring() {
    Timer t;
    const int S = in.nsteps;
    const std::vector<uint8_t> schemes = in.nbins;
    const uint64_t F_used = [&] {
        uint64_t n = 0;
        const uint64_t last = std::min<uint64_t>(F_file, in.nframe);
        for (uint64_t f = in.startframe; f < last; f += in.strideframe) ++n;
        return n;
    }();
    const uint64_t step_size = F_used / S;
    std::map<std::string, Contri> cache;
    for (size_t e = 0; e < in.estimators.size(); ++e) {
        for (size_t sc = 0; sc < schemes.size(); ++sc) {
            const std::string suffix = std::string(EST_NAME[in.estimators[e]]) + "_convergence_b-" +
                                       std::to_string((int)schemes[sc]) + "." + in.reportfileext;
            std::ofstream mie(out("MIE-" + suffix));
            std::ofstream mist;
            const bool do_mist = in.method == Method::MIST;
            if (do_mist) mist.open(out("MIST-" + suffix));
            const int fw = 13;
            const char* hdr[18] = {"TIME", "BOND-S1", "ANGLE-S1", "DIHED-S1", "B/B-S2", "B/B-I2",
                                   "A/A-S2", "A/A-I2", "D/D-S2", "D/D-I2", "B/A-S2", "B/A-I2",
                                   "B/D-S2", "B/D-I2", "A/D-S2", "A/D-I2", "S1", "S2"};
            for (std::ofstream* f : {&mie, &mist}) {
                if (!f->is_open()) continue;
                for (int h = 0; h < 18; ++h)
                    *f << std::setw(fw) << std::left << hdr[h] << (h < 17 ? ", " : "\n");
            }
            for (int st = 0; st < S; ++st) {
                double sum1[3] = {0, 0, 0};
                std::vector<double> S1v[3];
                std::vector<uint32_t> id1[3];
                std::map<uint32_t, uint32_t> idx1[3];
                for (int ty = 0; ty < 3; ++ty) {
                    const std::string p = out(std::string("entcontri_") + ONE_TAG[ty] + "-1d.nc");
                    if (!file_exists(p)) continue;
                    int order;
                    S1v[ty] = contri_of(cache, p, (int)e, (int)sc, st, id1[ty], order);
                    for (size_t k = 0; k < id1[ty].size(); ++k) idx1[ty][id1[ty][k]] = (uint32_t)k;
                    for (double v : S1v[ty]) sum1[ty] += v;
                }
                double s2sum[6] = {0}, misum[6] = {0};
                std::vector<uint32_t> src, dst;
                std::vector<double> w;
                std::vector<int> ekind;
                uint32_t off[3] = {0, (uint32_t)id1[0].size(), (uint32_t)(id1[0].size() + id1[1].size())};
                std::vector<std::vector<double>> mik(6);
                std::vector<std::vector<uint32_t>> ik(6);
                for (int k = 0; k < 6; ++k) {
                    const std::string p = out(std::string("entcontri_") + KINDS[k].tag + "-2d.nc");
                    if (!file_exists(p)) continue;
                    int order;
                    auto S2 = contri_of(cache, p, (int)e, (int)sc, st, ik[k], order);
                    mik[k].resize(S2.size());
                    for (size_t r = 0; r < S2.size(); ++r) {
                        const uint32_t a = idx1[KINDS[k].a].at(ik[k][2 * r]);
                        const uint32_t b = idx1[KINDS[k].b].at(ik[k][2 * r + 1]);
                        mik[k][r] = S1v[KINDS[k].a][a] + S1v[KINDS[k].b][b] - S2[r];
                        s2sum[k] += S2[r];
                        misum[k] += mik[k][r];
                    }
                }
                for (int q = 0; q < 6; ++q) {
                    const int k = EDGE_ORDER[q];
                    for (size_t r = 0; r < mik[k].size(); ++r) {
                        src.push_back(off[KINDS[k].a] + idx1[KINDS[k].a].at(ik[k][2 * r]));
                        dst.push_back(off[KINDS[k].b] + idx1[KINDS[k].b].at(ik[k][2 * r + 1]));
                        w.push_back(mik[k][r]);
                        ekind.push_back(k);
                    }
                }
                const double singlet = sum1[0] + sum1[1] + sum1[2];
                const uint64_t time = step_size * (st + 1);
                // columns: BB, AA, DD, BA, BD, AD
                const int colk[6] = {0, 1, 2, 3, 4, 5};
                auto row = [&](std::ofstream& f, const double* i2, double s2tot) {
                    f << std::setw(fw) << std::left << time << ", ";
                    f << std::fixed << std::setprecision(4);
                    for (int ty = 0; ty < 3; ++ty) f << std::setw(fw) << std::left << sum1[ty] << ", ";
                    for (int c = 0; c < 6; ++c)
                        f << std::setw(fw) << std::left << s2sum[colk[c]] << ", " << std::setw(fw)
                          << std::left << i2[colk[c]] << ", ";
                    f << std::setw(fw) << std::left << singlet << ", " << std::setw(fw) << std::left
                      << s2tot << "\n";
                };
                double mie_tot = singlet;
                for (int k = 0; k < 6; ++k) mie_tot -= misum[k];
                row(mie, misum, mie_tot);
                if (do_mist) {
                    const uint32_t V = off[2] + (uint32_t)id1[2].size();
                    std::vector<uint32_t> tree(std::max<uint32_t>(V, 1));
                    uint32_t nt = 0;
                    ck(ctx, centre_mst(ctx, src.data(), dst.data(), w.data(), w.size(), V, tree.data(), &nt),
                       "mst");
                    double tk[6] = {0};
                    std::ofstream tf(out("Tree-MIST-" + std::string(EST_NAME[in.estimators[e]]) + "_b-" +
                                         std::to_string((int)schemes[sc]) + "_s-" + std::to_string(st) +
                                         "." + in.reportfileext));
                    tf << "TD1, TD2, IndexDim1, IndexDim2, MI\n";
                    const char tc[3] = {'B', 'A', 'T'};
                    auto type_of = [&](uint32_t v) { return v < off[1] ? 0 : (v < off[2] ? 1 : 2); };
                    auto id_of = [&](uint32_t v) { int ty = type_of(v); return id1[ty][v - off[ty]]; };
                    // edges by MI, largest first (cmpEdgeMIST, main_common.cpp:295)
                    std::vector<uint32_t> order(tree.begin(), tree.begin() + nt);
                    std::sort(order.begin(), order.end(),
                              [&](uint32_t x, uint32_t y) { return w[x] > w[y]; });
                    for (uint32_t ei : order) {
                        tk[ekind[ei]] += w[ei];
                        tf << std::setw(3) << std::left << tc[type_of(src[ei])] << ", " << std::setw(3)
                           << std::left << tc[type_of(dst[ei])] << ", " << std::setw(9) << std::left
                           << id_of(src[ei]) << ", " << std::setw(9) << std::left << id_of(dst[ei])
                           << ", " << std::setw(3) << std::left << std::fixed << std::setprecision(6)
                           << w[ei] << "\n";
                    }
                    double mist_tot = singlet;
                    for (int k = 0; k < 6; ++k) mist_tot -= tk[k];
                    row(mist, tk, mist_tot);
                }
            }
        }
    }
    times["scoring"] = t.s();
}

void usage() {
    std::cout << "usage: centre_hip -i <input.cfg> [-O]\n"
                 "  -i  CENTRE input file (same keys as centre_OMP)\n"
                 "  -O  overwrite existing output files\n"
                 "       centre_hip --inspect <file.nc> [var ...]\n"
                 "  print the format, global text attributes and, per variable,\n"
                 "  its shape and the sum of its values (no GPU is used)\n"
                 "       centre_hip --print-config <input.cfg> [frames in the trajectory]\n"
                 "  print the parsed input parameters (no GPU is used)\n"
                 "       centre_hip --convert <entcontri.nc> <out.nc> [netcdf4|cdf5]\n"
                 "  rewrite an entropy-contribution file of either format in the\n"
                 "  layout the entropy stage writes (default netcdf4; no GPU is used)\n";
}

int convert(const std::string& src, const std::string& dst, const std::string& fmt) {
    if (fmt != "netcdf4" && fmt != "cdf5") throw Fail("format must be netcdf4 or cdf5");
    auto f = ncf::open(src);
    if (f->text_attr("convention") != "EntropyCont" || !f->has_var("contri_S") || !f->has_var("id"))
        throw Fail(src + ": not an entropy-contribution file");
    const auto shp = f->shape("contri_S");   // records, estimators, schemes, steps
    const auto order = f->uint_attr("order"), st = f->uint_attr("storagetype");
    const auto dl = f->uint_attr("dimLengths"), bs = f->uint_attr("binSchemes");
    if (shp.size() != 4 || order.size() != 1 || st.size() != 1 || bs.size() != shp[2])
        throw Fail(src + ": attributes do not match contri_S");
    const std::vector<uint8_t> schemes(bs.begin(), bs.end());
    write_entcontri(dst, f->text_attr("content"), (int)order[0], st[0] == 3, dl, schemes,
                    (int)shp[3], (int)shp[1], f->read_u32("id"), f->read_f64("contri_S"),
                    fmt == "netcdf4");
    return 0;
}

int inspect(int argc, char** argv, int first) {
    auto f = ncf::open(argv[first]);
    std::cout << "format " << f->format() << "\n";
    for (const char* a : {"title", "convention", "content", "program", "application"}) {
        const std::string t = f->text_attr(a);
        if (!t.empty()) std::cout << "attr " << a << " " << t << "\n";
    }
    for (int i = first + 1; i < argc; ++i) {
        const std::string v = argv[i];
        if (!f->has_var(v)) {
            std::cout << "var " << v << " absent\n";
            continue;
        }
        auto shp = f->shape(v);
        std::cout << "var " << v << " shape";
        for (size_t k = 0; k < shp.size(); ++k) std::cout << (k ? "," : " ") << shp[k];
        const auto x = f->read_f64(v);
        double sum = 0.0;
        for (double d : x) sum += d;
        std::cout << " sum " << std::setprecision(17) << sum << "\n";
    }
    return 0;
}

// frames_in_file > 0: also apply the values derived from a trajectory of
// that many frames (frame range clamps, defaults) before printing.
int print_config(const char* path, uint64_t frames_in_file) {
    std::ifstream f(path);
    if (!f) throw Fail(std::string("cannot read input file ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    Inputs in = inputs_from_symbols(parse_ini(ss.str()));
    if (frames_in_file > 0 && in.ok()) derive(in, frames_in_file);
    for (auto& e : in.errors) std::cout << "error " << e << "\n";
    std::cout << "nsteps " << in.nsteps << "\ndiscretize " << in.discretize << "\ncalcentropy "
              << in.calcentropy << "\ngenconvgdata " << in.genconvgdata << "\ninfilepath "
              << in.infilepath << "\noutfilepath " << in.outfilepath << "\npdfmethod "
              << (in.pdfmethod == PdfMethod::Histogram ? "histogram" : "vonmiseskde")
              << "\nnbins";
    for (auto b : in.nbins) std::cout << " " << (int)b;
    std::cout << "\nreferencenbins " << in.referencenbins << "\nworkset " << in.workset
              << "\nusesubset " << in.usesubset << "\nuseneighbor " << in.useneighbor
              << "\njacobian " << in.jacobian << "\nestimators";
    for (auto e : in.estimators) std::cout << " " << EST_NAME[e];
    std::cout << "\nmethod " << (int)in.method << "\nkappa " << in.kappa << "\nnmaxconf "
              << in.nmaxconf << "\ncacheentcdimspercpu " << in.cacheentcdimspercpu
              << "\nthreadsfororder " << in.threads_for_order << "\nprocsfororder "
              << in.procs_for_order << "\nstartframe " << in.startframe << "\nstrideframe " << in.strideframe
              << "\nnframe " << in.nframe << "\n";
    return in.ok() ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    std::string cfg;
    bool overwrite = false;
    try {
        if (argc >= 3 && !std::strcmp(argv[1], "--inspect")) return inspect(argc, argv, 2);
        if (argc >= 3 && !std::strcmp(argv[1], "--print-config"))
            return print_config(argv[2], argc >= 4 ? std::strtoull(argv[3], nullptr, 10) : 0);
        if (argc >= 4 && !std::strcmp(argv[1], "--convert"))
            return convert(argv[2], argv[3], argc >= 5 ? argv[4] : "netcdf4");
    } catch (const std::exception& e) {
        std::cerr << "centre_hip: " << e.what() << "\n";
        return 2;
    }
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "-i") && i + 1 < argc) cfg = argv[++i];
        else if (!std::strcmp(argv[i], "-O")) overwrite = true;
        else if (!std::strcmp(argv[i], "-h")) { usage(); return 0; }
    }
    if (cfg.empty()) { usage(); return 1; }
    Host h;
    try {
        std::ifstream f(cfg);
        if (!f) throw Fail("cannot read input file " + cfg);
        std::stringstream ss;
        ss << f.rdbuf();
        h.in = inputs_from_symbols(parse_ini(ss.str()));
        if (!h.in.ok()) {
            for (auto& e : h.in.errors) std::cerr << "InputReadError:: " << e << "\n";
            return 1;
        }
        if (h.in.calcentropy && h.in.entcontri_netcdf4 && !ncf::netcdf4_available())
            throw Fail("NetCDF-4 output needs libhdf5 and libhdf5_hl (or set "
                       "mi355x.entcontriformat = cdf5)");
        if (h.in.discretize) {
            const std::string batp = join(h.in.infilepath, h.in.fnames[0]);
            auto bat = ncf::open(batp);
            // NetcdfFile::getNetcdfConventions (NetcdfFile.cpp:88-110)
            if (bat->text_attr("convention") != "CENTRE")
                throw Fail(batp + ": global attribute convention is not \"CENTRE\"");
            for (const char* v : TYPE_VAR)
                if (bat->has_var(v)) {
                    h.F_file = bat->shape(v)[0];
                    break;
                }
            derive(h.in, h.F_file);
            if (!overwrite && file_exists(join(h.in.outfilepath, BIN_FILE[0])))
                throw Fail("output files exist in " + h.in.outfilepath + " (use -O to overwrite)");
            ck(nullptr, centre_init(h.in.device, &h.ctx), "centre_init (no GPU?)");
            h.discretize(*bat);
        } else {
            // entropy from the bin files a previous run wrote (outfilepath)
            h.load_bins();
            ck(nullptr, centre_init(h.in.device, &h.ctx), "centre_init (no GPU?)");
        }
        derive(h.in, h.F_file);   // entropy frames within the discretised ones
        if (h.in.calcentropy) h.run_entropy();
        if (h.in.genconvgdata) h.run_scoring();
        std::ofstream info(join(h.in.outfilepath, h.in.infofile));
        info << std::fixed << std::setprecision(3);
        for (auto& kv : h.times) {
            info << std::setw(16) << kv.first << " " << kv.second << " s\n";
            std::cout << "TIME: " << kv.first << " " << kv.second << " s\n";
        }
        centre_destroy(h.ctx);
    } catch (const std::exception& e) {
        std::cerr << "centre_hip: " << e.what() << "\n";
        if (h.ctx) centre_destroy(h.ctx);
        return 2;
    }
    return 0;
}
