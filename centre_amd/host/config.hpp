// Input-file parameters of CENTRE (the drop-in configuration surface).
//
// Same INI dialect as Utils::parseConfiguration (Utils.cpp:198-270):
// "[section]" headers, "key = v1, v2" lines, '#' comments (whole line or
// inline token), a double-quoted single value.  Same keys and defaults as
// Inputs (Inputs.cpp:14-74 defaults, :92-553 keys, :556-751 derived values).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace centre {

using Symbols = std::map<std::string, std::vector<std::string>>;
Symbols parse_ini(const std::string& text);

// BATSet bits (Centre.h:96-109)
enum BATSet : uint32_t {
    B1D = 1, A1D = 2, D1D = 4, X1D = 7, BB2D = 8, AA2D = 16, DD2D = 32,
    XX2D = 63, BA2D = 64, BD2D = 128, AD2D = 256, XY2D = 511
};
uint32_t parse_workset(const std::vector<std::string>& tokens, uint32_t dflt);

enum class PdfMethod { Histogram, VonMisesKDE };
enum class Method { MIE, MIST, AMIE, AMIST };

struct Inputs {
    // [control]
    std::string infilepath = "./", outfilepath = "./", infofile = "progress.info";
    int nsteps = 20;
    bool discretize = true, calcentropy = true, genconvgdata = true;
    unsigned cachedscrtdimspercpu = 5, cacheentcdimspercpu = 10;
    unsigned dscrinfofreq = 25, entcinfofreq = 500;
    // [discretization]
    std::vector<std::string> fnames;
    std::vector<uint64_t> nframes;
    unsigned nbond = 0, nangle = 0, ndihed = 0;
    uint64_t randseed = 0;
    bool shuffleframes = false, shuffledofs = false, optimizedih = true;
    PdfMethod pdfmethod = PdfMethod::Histogram;
    // [histogram]
    std::vector<uint8_t> nbins;
    int referencenbins = 0;
    bool hist_writefreq = false;
    uint32_t hist_writeset = 0;
    int hist_writestepstart = 255, hist_writestepstride = 1;
    // [vonmiseskde]
    int nmaxconf = 5;
    double kappa = 1.0;
    unsigned sdosteps = 5, sdoiterations = 1000;
    double sdoconvlimit = 0.0001;
    bool vm_writefreq = false;
    uint32_t vm_writeset = 0;
    int vm_writestepstart = 255, vm_writestepstride = 1;
    // [entropy]
    uint32_t workset = XY2D;
    bool usesubset = false, useneighbor = false, jacobian = true;
    std::string subsetfile, neighborfile, reportfileext = "csv";
    uint64_t startframe = 0, strideframe = 1, nframe = 0;
    Method method = Method::MIST;
    std::vector<int> estimators;            // 0 ML, 1 MM, 2 CS, 3 JS
    // options of this build (no reference counterpart)
    int threads_for_order = 1;              // OpenMP threads the record order emulates
    int procs_for_order = 1;                // MPI ranks the record order emulates (1: centre_OMP)
    int device = 0;
    bool entcontri_netcdf4 = true;          // entcontri_* as NetCDF-4 (the reference's format) or CDF-5

    std::vector<std::string> errors;
    bool ok() const { return errors.empty(); }
};

// Fills an Inputs from parsed symbols (keys absent keep their defaults).
Inputs inputs_from_symbols(const Symbols& sym);
// Derived values after the BAT header is known (Inputs.cpp:556-751).
void derive(Inputs& in, uint64_t frames_in_file);

}  // namespace centre
