// Format dispatch for NetCDF reads (see ncfile.hpp).
#include "ncfile.hpp"

#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "nc_classic.hpp"

namespace ncf {

namespace {

class Classic final : public File {
public:
    explicit Classic(const std::string& p) : r_(p) {}
    bool has_var(const std::string& n) const override { return r_.has_var(n); }
    std::vector<uint64_t> shape(const std::string& n) const override { return r_.shape(n); }
    std::vector<double> read_f64(const std::string& n, std::vector<uint64_t> s,
                                 std::vector<uint64_t> c) const override {
        return r_.read<double>(n, s, c);
    }
    std::vector<uint8_t> read_u8(const std::string& n, std::vector<uint64_t> s,
                                 std::vector<uint64_t> c) const override {
        return r_.read<uint8_t>(n, s, c);
    }
    std::vector<uint32_t> read_u32(const std::string& n, std::vector<uint64_t> s,
                                   std::vector<uint64_t> c) const override {
        return r_.read<uint32_t>(n, s, c);
    }
    std::string text_attr(const std::string& n) const override {
        auto it = r_.attrs().find(n);
        if (it == r_.attrs().end() || it->second.type != nc::CHAR) return "";
        std::string t = it->second.text();
        while (!t.empty() && t.back() == '\0') t.pop_back();
        return t;
    }
    const char* format() const override { return "classic"; }

private:
    nc::Reader r_;
};

// ---- NetCDF-4 through libhdf5, loaded at run time ---------------------------
// libhdf5 is opened with dlopen the first time a NetCDF-4 file is read, so the
// executable neither links it nor inherits the C++ runtime of the prefix it
// lives in.  Types follow the HDF5 1.10+ C API (hid_t is 64-bit).
using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kP_DEFAULT = 0, kS_ALL = 0;
constexpr unsigned kF_ACC_RDONLY = 0u;
constexpr int kS_SELECT_SET = 0, kT_STRING = 3, kT_STR_NULLPAD = 1;
constexpr size_t kT_VARIABLE = (size_t)-1;

struct Api {
    void* lib = nullptr;
    herr_t (*open_)();
    herr_t (*Eset_auto2)(hid_t, void*, void*);
    hid_t (*Fopen)(const char*, unsigned, hid_t);
    herr_t (*Fclose)(hid_t);
    htri_t (*Lexists)(hid_t, const char*, hid_t);
    hid_t (*Dopen2)(hid_t, const char*, hid_t);
    herr_t (*Dclose)(hid_t);
    hid_t (*Dget_space)(hid_t);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*Sclose)(hid_t);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
    herr_t (*Sselect_hyperslab)(hid_t, int, const hsize_t*, const hsize_t*, const hsize_t*, const hsize_t*);
    hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*);
    htri_t (*Aexists)(hid_t, const char*);
    hid_t (*Aopen)(hid_t, const char*, hid_t);
    herr_t (*Aclose)(hid_t);
    hid_t (*Aget_type)(hid_t);
    herr_t (*Aread)(hid_t, hid_t, void*);
    herr_t (*Tclose)(hid_t);
    int (*Tget_class)(hid_t);
    htri_t (*Tis_variable_str)(hid_t);
    hid_t (*Tcopy)(hid_t);
    herr_t (*Tset_size)(hid_t, size_t);
    size_t (*Tget_size)(hid_t);
    herr_t (*Tset_strpad)(hid_t, int);
    herr_t (*free_memory)(void*);
    hid_t native_double, native_uchar, native_uint, c_s1;

    template <class F> void sym(F& f, const char* name) {
        f = reinterpret_cast<F>(dlsym(lib, name));
        if (!f) throw nc::Error(std::string("libhdf5 lacks ") + name);
    }
    hid_t global(const char* name) {
        auto* p = reinterpret_cast<hid_t*>(dlsym(lib, name));
        if (!p) throw nc::Error(std::string("libhdf5 lacks ") + name);
        return *p;
    }
    Api() {
        const char* env = std::getenv("CENTRE_HDF5_LIB");
        const char* cands[] = {env, "libhdf5.so", "/opt/conda/lib/libhdf5.so",
                               "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
                               "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so", "/usr/local/lib/libhdf5.so"};
        for (const char* c : cands)
            if (c && (lib = dlopen(c, RTLD_NOW | RTLD_LOCAL))) break;
        if (!lib)
            throw nc::Error("NetCDF-4 (HDF5) input needs libhdf5 (set CENTRE_HDF5_LIB to its path)");
        sym(open_, "H5open");
        sym(Eset_auto2, "H5Eset_auto2");
        sym(Fopen, "H5Fopen");
        sym(Fclose, "H5Fclose");
        sym(Lexists, "H5Lexists");
        sym(Dopen2, "H5Dopen2");
        sym(Dclose, "H5Dclose");
        sym(Dget_space, "H5Dget_space");
        sym(Dread, "H5Dread");
        sym(Sclose, "H5Sclose");
        sym(Sget_simple_extent_ndims, "H5Sget_simple_extent_ndims");
        sym(Sget_simple_extent_dims, "H5Sget_simple_extent_dims");
        sym(Sselect_hyperslab, "H5Sselect_hyperslab");
        sym(Screate_simple, "H5Screate_simple");
        sym(Aexists, "H5Aexists");
        sym(Aopen, "H5Aopen");
        sym(Aclose, "H5Aclose");
        sym(Aget_type, "H5Aget_type");
        sym(Aread, "H5Aread");
        sym(Tclose, "H5Tclose");
        sym(Tget_class, "H5Tget_class");
        sym(Tis_variable_str, "H5Tis_variable_str");
        sym(Tcopy, "H5Tcopy");
        sym(Tset_size, "H5Tset_size");
        sym(Tget_size, "H5Tget_size");
        sym(Tset_strpad, "H5Tset_strpad");
        sym(free_memory, "H5free_memory");
        if (open_() < 0) throw nc::Error("H5open failed");
        Eset_auto2(0, nullptr, nullptr);   // errors are reported by this reader
        native_double = global("H5T_NATIVE_DOUBLE_g");
        native_uchar = global("H5T_NATIVE_UCHAR_g");
        native_uint = global("H5T_NATIVE_UINT_g");
        c_s1 = global("H5T_C_S1_g");
    }
};

Api& api() {
    static Api a;   // loaded once, on first use
    return a;
}

// closes an HDF5 identifier with the matching call
struct H {
    hid_t id;
    herr_t (*close)(hid_t);
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H() {
        if (id >= 0) close(id);
    }
    H(const H&) = delete;
    operator hid_t() const { return id; }
};

class Hdf5 final : public File {
public:
    explicit Hdf5(const std::string& p) : path_(p), A_(api()) {
        f_ = A_.Fopen(p.c_str(), kF_ACC_RDONLY, kP_DEFAULT);
        if (f_ < 0) throw nc::Error(p + ": cannot open as NetCDF-4/HDF5");
    }
    ~Hdf5() override {
        if (f_ >= 0) A_.Fclose(f_);
    }
    bool has_var(const std::string& n) const override {
        return A_.Lexists(f_, n.c_str(), kP_DEFAULT) > 0;
    }
    std::vector<uint64_t> shape(const std::string& n) const override {
        H d(open_ds(n), A_.Dclose);
        H sp(A_.Dget_space(d), A_.Sclose);
        return dims_of(sp);
    }
    std::vector<double> read_f64(const std::string& n, std::vector<uint64_t> s,
                                 std::vector<uint64_t> c) const override {
        return read<double>(n, s, c, A_.native_double);
    }
    std::vector<uint8_t> read_u8(const std::string& n, std::vector<uint64_t> s,
                                 std::vector<uint64_t> c) const override {
        return read<uint8_t>(n, s, c, A_.native_uchar);
    }
    std::vector<uint32_t> read_u32(const std::string& n, std::vector<uint64_t> s,
                                   std::vector<uint64_t> c) const override {
        return read<uint32_t>(n, s, c, A_.native_uint);
    }
    std::string text_attr(const std::string& n) const override {
        if (A_.Aexists(f_, n.c_str()) <= 0) return "";
        H a(A_.Aopen(f_, n.c_str(), kP_DEFAULT), A_.Aclose);
        H t(A_.Aget_type(a), A_.Tclose);
        if (A_.Tget_class(t) != kT_STRING) return "";
        std::string out;
        H mt(A_.Tcopy(A_.c_s1), A_.Tclose);
        if (A_.Tis_variable_str(t) > 0) {
            char* s = nullptr;
            A_.Tset_size(mt, kT_VARIABLE);
            if (A_.Aread(a, mt, &s) >= 0 && s) {
                out = s;
                A_.free_memory(s);
            }
        } else {
            const size_t len = A_.Tget_size(t);
            std::vector<char> buf(len + 1, 0);
            A_.Tset_size(mt, len);
            A_.Tset_strpad(mt, kT_STR_NULLPAD);
            if (A_.Aread(a, mt, buf.data()) >= 0) out.assign(buf.data(), strnlen(buf.data(), len));
        }
        return out;
    }
    const char* format() const override { return "netcdf4-hdf5"; }

private:
    std::string path_;
    Api& A_;
    hid_t f_ = -1;
    hid_t open_ds(const std::string& n) const {
        const hid_t d = A_.Dopen2(f_, n.c_str(), kP_DEFAULT);
        if (d < 0) throw nc::Error(path_ + ": no variable '" + n + "'");
        return d;
    }
    std::vector<uint64_t> dims_of(hid_t sp) const {
        const int nd = A_.Sget_simple_extent_ndims(sp);
        std::vector<hsize_t> dims(nd > 0 ? nd : 0);
        if (nd > 0) A_.Sget_simple_extent_dims(sp, dims.data(), nullptr);
        return std::vector<uint64_t>(dims.begin(), dims.end());
    }
    template <class T>
    std::vector<T> read(const std::string& n, std::vector<uint64_t> s, std::vector<uint64_t> c,
                        hid_t memtype) const {
        H d(open_ds(n), A_.Dclose);
        H sp(A_.Dget_space(d), A_.Sclose);
        const std::vector<uint64_t> dims = dims_of(sp);
        if (s.empty()) s.assign(dims.size(), 0);
        if (c.empty())
            for (size_t k = 0; k < dims.size(); ++k) c.push_back(dims[k] - s[k]);
        if (s.size() != dims.size() || c.size() != dims.size())
            throw nc::Error(path_ + ": rank mismatch reading '" + n + "'");
        uint64_t total = 1;
        for (size_t k = 0; k < dims.size(); ++k) {
            if (s[k] + c[k] > dims[k]) throw nc::Error(path_ + ": read past the end of '" + n + "'");
            total *= c[k];
        }
        std::vector<T> out(total);
        if (!total) return out;
        herr_t rc;
        if (!dims.empty()) {
            std::vector<hsize_t> hs(s.begin(), s.end()), hc(c.begin(), c.end());
            A_.Sselect_hyperslab(sp, kS_SELECT_SET, hs.data(), nullptr, hc.data(), nullptr);
            H ms(A_.Screate_simple((int)hc.size(), hc.data(), nullptr), A_.Sclose);
            rc = A_.Dread(d, memtype, ms, sp, kP_DEFAULT, out.data());
        } else {
            rc = A_.Dread(d, memtype, kS_ALL, kS_ALL, kP_DEFAULT, out.data());
        }
        if (rc < 0) throw nc::Error(path_ + ": HDF5 read of '" + n + "' failed");
        return out;
    }
};

}  // namespace

std::unique_ptr<File> open(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw nc::Error("cannot open " + path);
    char magic[8] = {0};
    f.read(magic, 8);
    if (magic[0] == 'C' && magic[1] == 'D' && magic[2] == 'F') return std::make_unique<Classic>(path);
    if (std::memcmp(magic, "\x89HDF\r\n\x1a\n", 8) == 0) return std::make_unique<Hdf5>(path);
    throw nc::Error(path + ": not a NetCDF file");
}

}  // namespace ncf
