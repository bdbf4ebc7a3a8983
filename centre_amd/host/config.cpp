// INI parsing and the Inputs key table (see config.hpp).
#include "config.hpp"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <sstream>

namespace centre {

static std::string trim(const std::string& s, const char* ws = " \t\r\n") {
    const auto b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    const auto e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

Symbols parse_ini(const std::string& text) {
    Symbols out;
    std::istringstream in(text);
    std::string line, section;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        if (line.size() > 2 && line.front() == '[' && line.back() == ']') {
            section = line.substr(1, line.size() - 2);
            continue;
        }
        const auto eq = line.find('=');
        std::string key = trim(line.substr(0, eq));
        if (!section.empty()) key = section + "." + key;
        std::string value = eq == std::string::npos ? "" : trim(line.substr(eq + 1));
        std::vector<std::string> values;
        const auto q1 = value.find('"'), q2 = value.rfind('"');
        if (q1 != std::string::npos && q2 != std::string::npos && q2 > q1) {
            values.push_back(value.substr(q1 + 1, q2 - q1 - 1));
        } else {
            std::istringstream words(value);
            std::string tok;
            while (words >> tok) {
                if (!tok.empty() && tok[0] == '#') break;
                tok = trim(tok, ",");
                if (!tok.empty()) values.push_back(tok);
            }
        }
        out[key] = values;
    }
    return out;
}

static bool to_bool(const std::string& s) {
    std::string t = s;
    std::transform(t.begin(), t.end(), t.begin(), ::tolower);
    return t == "true" || t == "1" || t == "yes";
}

uint32_t parse_workset(const std::vector<std::string>& tokens, uint32_t dflt) {
    static const std::map<std::string, uint32_t> names = {
        {"B1D", B1D}, {"A1D", A1D}, {"D1D", D1D}, {"1D", X1D}, {"BB2D", BB2D},
        {"AA2D", AA2D}, {"DD2D", DD2D}, {"XX2D", XX2D}, {"BA2D", BA2D},
        {"BD2D", BD2D}, {"AD2D", AD2D}, {"2D", XY2D}};
    if (tokens.empty()) return dflt;
    uint32_t set = 0;
    for (auto& t : tokens) {
        if (t == "NONE") return 0;
        auto it = names.find(t);
        if (it != names.end()) set |= it->second;
    }
    return set;
}

Inputs inputs_from_symbols(const Symbols& sym) {
    Inputs in;
    auto get = [&](const char* k) -> const std::vector<std::string>* {
        auto it = sym.find(k);
        return (it == sym.end() || it->second.empty()) ? nullptr : &it->second;
    };
    auto s1 = [&](const char* k, std::string& dst) { if (auto v = get(k)) dst = (*v)[0]; };
    auto i1 = [&](const char* k, auto& dst) { if (auto v = get(k)) dst = (std::remove_reference_t<decltype(dst)>)std::atoll((*v)[0].c_str()); };
    auto d1 = [&](const char* k, double& dst) { if (auto v = get(k)) dst = std::atof((*v)[0].c_str()); };
    auto b1 = [&](const char* k, bool& dst) { if (auto v = get(k)) dst = to_bool((*v)[0]); };

    i1("control.nsteps", in.nsteps);
    b1("control.discretize", in.discretize);
    b1("control.calcentropy", in.calcentropy);
    b1("control.genconvgdata", in.genconvgdata);
    i1("control.dscrinfofreq", in.dscrinfofreq);
    i1("control.entcinfofreq", in.entcinfofreq);
    i1("control.cachedscrtdimspercpu", in.cachedscrtdimspercpu);
    i1("control.cacheentcdimspercpu", in.cacheentcdimspercpu);
    s1("control.outfilepath", in.outfilepath);
    s1("control.infilepath", in.infilepath);
    s1("control.infofile", in.infofile);
    i1("discretization.nbond", in.nbond);
    i1("discretization.nangle", in.nangle);
    i1("discretization.ndihed", in.ndihed);
    if (auto v = get("discretization.fname")) in.fnames.push_back((*v)[0]);
    if (auto v = get("discretization.nframe")) in.nframes.push_back(std::atoll((*v)[0].c_str()));
    i1("discretization.randseed", in.randseed);
    b1("discretization.shuffleframes", in.shuffleframes);
    b1("discretization.shuffledofs", in.shuffledofs);
    b1("discretization.optimizedih", in.optimizedih);
    if (auto v = get("discretization.pdfmethod")) {
        std::string m = (*v)[0];
        std::transform(m.begin(), m.end(), m.begin(), ::tolower);
        if (m == "histogram") in.pdfmethod = PdfMethod::Histogram;
        else if (m == "vonmiseskde") in.pdfmethod = PdfMethod::VonMisesKDE;
    }
    i1("vonmiseskde.nmaxconf", in.nmaxconf);
    b1("vonmiseskde.writefreq", in.vm_writefreq);
    if (auto v = get("vonmiseskde.writeset")) in.vm_writeset = parse_workset(*v, 0);
    i1("vonmiseskde.writestepstart", in.vm_writestepstart);
    i1("vonmiseskde.writestepstride", in.vm_writestepstride);
    d1("vonmiseskde.kappa", in.kappa);
    i1("vonmiseskde.sdosteps", in.sdosteps);
    i1("vonmiseskde.sdoiterations", in.sdoiterations);
    d1("vonmiseskde.sdoconvlimit", in.sdoconvlimit);
    i1("histogram.referencenbins", in.referencenbins);
    if (auto v = get("histogram.nbins"))
        for (auto& t : *v) in.nbins.push_back((uint8_t)std::atoi(t.c_str()));
    b1("histogram.writefreq", in.hist_writefreq);
    if (auto v = get("histogram.writeset")) in.hist_writeset = parse_workset(*v, 0);
    i1("histogram.writestepstart", in.hist_writestepstart);
    i1("histogram.writestepstride", in.hist_writestepstride);
    if (auto v = get("entropy.workset")) in.workset = parse_workset(*v, XY2D);
    b1("entropy.usesubset", in.usesubset);
    b1("entropy.useneighbor", in.useneighbor);
    b1("entropy.jacobian", in.jacobian);
    s1("entropy.subsetfile", in.subsetfile);
    s1("entropy.neighborfile", in.neighborfile);
    s1("entropy.reportfileext", in.reportfileext);
    i1("entropy.startframe", in.startframe);
    i1("entropy.strideframe", in.strideframe);
    i1("entropy.nframe", in.nframe);
    if (auto v = get("entropy.scoringmethod")) {
        std::string m = (*v)[0];
        std::transform(m.begin(), m.end(), m.begin(), ::toupper);
        in.method = m == "MIE" ? Method::MIE : m == "AMIE" ? Method::AMIE
                  : m == "AMIST" ? Method::AMIST : Method::MIST;
    }
    if (auto v = get("entropy.estimator")) {
        for (auto t : *v) {
            std::transform(t.begin(), t.end(), t.begin(), ::toupper);
            int e = t == "ML" ? 0 : t == "MM" ? 1 : t == "CS" ? 2 : t == "JS" ? 3 : -1;
            if (e >= 0 && std::find(in.estimators.begin(), in.estimators.end(), e) == in.estimators.end())
                in.estimators.push_back(e);
        }
    }
    // options of the MI355X host (no counterpart in the reference's cfg)
    i1("mi355x.threadsfororder", in.threads_for_order);
    i1("mi355x.procsfororder", in.procs_for_order);
    if (in.threads_for_order < 1) in.errors.push_back("mi355x.threadsfororder must be at least 1");
    if (in.procs_for_order < 1) in.errors.push_back("mi355x.procsfororder must be at least 1");
    i1("mi355x.device", in.device);
    if (auto v = get("mi355x.entcontriformat")) {
        std::string m = (*v)[0];
        std::transform(m.begin(), m.end(), m.begin(), ::tolower);
        if (m == "netcdf4") in.entcontri_netcdf4 = true;
        else if (m == "cdf5") in.entcontri_netcdf4 = false;
        else in.errors.push_back("mi355x.entcontriformat must be netcdf4 or cdf5");
    }

    if (in.nbins.empty() && in.pdfmethod == PdfMethod::Histogram)
        in.errors.push_back("histogram.nbins has no default and must be given");
    if (in.referencenbins <= 0 && in.pdfmethod == PdfMethod::Histogram)
        in.errors.push_back("histogram.referencenbins has no default and must be given");
    if (in.nsteps < 1 || in.nsteps > 255) in.errors.push_back("control.nsteps must be in 1..255");
    if (in.fnames.empty()) in.errors.push_back("discretization.fname is required");
    // a stride of 0 never advances (the reference divides by it,
    // Inputs.cpp:639, Netcdf_TrjInt.cpp:122)
    if (in.strideframe < 1) in.errors.push_back("entropy.strideframe must be at least 1");
    return in;
}

void derive(Inputs& in, uint64_t frames_in_file) {
    if (in.nframe == 0 || in.nframe > frames_in_file) in.nframe = frames_in_file;
    // the frame reader's clamps (Netcdf_TrjInt.cpp:120-121): a start at or
    // past the last frame reads from frame 0, a stride longer than the
    // trajectory reads every frame
    if (in.startframe >= in.nframe) in.startframe = 0;
    if (in.strideframe > in.nframe) in.strideframe = 1;
    if (in.pdfmethod == PdfMethod::VonMisesKDE) {
        in.referencenbins = in.nmaxconf;
        in.nbins.assign(1, (uint8_t)in.nmaxconf);
        in.hist_writefreq = in.vm_writefreq;
        in.hist_writeset = in.vm_writeset;
        in.hist_writestepstart = in.vm_writestepstart;
        in.hist_writestepstride = in.vm_writestepstride;
        uint32_t ws = in.workset, nws = 0;
        if (ws & D1D) nws = D1D;
        if (ws & DD2D) nws = D1D | DD2D;
        in.workset = nws;
    }
    if (in.hist_writefreq && in.hist_writestepstart == 255) in.hist_writestepstart = in.nsteps - 1;
    if (in.estimators.empty()) in.estimators.push_back(1);   // MM (Inputs.cpp:729-732)
}

}  // namespace centre
