// Format dispatch for NetCDF reads (see ncfile.hpp).
#include "ncfile.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
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
    std::vector<uint64_t> uint_attr(const std::string& n) const override {
        auto it = r_.attrs().find(n);
        if (it == r_.attrs().end()) return {};
        const nc::Attr& a = it->second;
        std::vector<uint64_t> out;
        switch (a.type) {
            case nc::BYTE: for (auto v : a.values<int8_t>()) out.push_back((uint64_t)v); break;
            case nc::UBYTE: for (auto v : a.values<uint8_t>()) out.push_back(v); break;
            case nc::SHORT: for (auto v : a.values<int16_t>()) out.push_back((uint64_t)v); break;
            case nc::USHORT: for (auto v : a.values<uint16_t>()) out.push_back(v); break;
            case nc::INT: for (auto v : a.values<int32_t>()) out.push_back((uint64_t)v); break;
            case nc::UINT: for (auto v : a.values<uint32_t>()) out.push_back(v); break;
            case nc::INT64: for (auto v : a.values<int64_t>()) out.push_back((uint64_t)v); break;
            case nc::UINT64: out = a.values<uint64_t>(); break;
            default: break;
        }
        return out;
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
constexpr int kS_SELECT_SET = 0, kT_INTEGER = 0, kT_STRING = 3, kT_STR_NULLPAD = 1;
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
    hid_t (*Aget_space)(hid_t);
    long long (*Sget_simple_extent_npoints)(hid_t);
    hid_t native_double, native_uchar, native_uint, native_ullong, c_s1;

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
        sym(Aget_space, "H5Aget_space");
        sym(Sget_simple_extent_npoints, "H5Sget_simple_extent_npoints");
        if (open_() < 0) throw nc::Error("H5open failed");
        Eset_auto2(0, nullptr, nullptr);   // errors are reported by this reader
        native_double = global("H5T_NATIVE_DOUBLE_g");
        native_uchar = global("H5T_NATIVE_UCHAR_g");
        native_uint = global("H5T_NATIVE_UINT_g");
        native_ullong = global("H5T_NATIVE_ULLONG_g");
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
    std::vector<uint64_t> uint_attr(const std::string& n) const override {
        if (A_.Aexists(f_, n.c_str()) <= 0) return {};
        H a(A_.Aopen(f_, n.c_str(), kP_DEFAULT), A_.Aclose);
        H t(A_.Aget_type(a), A_.Tclose);
        if (A_.Tget_class(t) != kT_INTEGER) return {};
        H sp(A_.Aget_space(a), A_.Sclose);
        const long long n_pts = A_.Sget_simple_extent_npoints(sp);
        if (n_pts <= 0) return {};
        std::vector<uint64_t> out((size_t)n_pts);
        if (A_.Aread(a, A_.native_ullong, out.data()) < 0) return {};
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

// ---- NetCDF-4 writing ---------------------------------------------------------
// The write side needs more of libhdf5 and the dimension-scale calls of
// libhdf5_hl (the library netcdf-c itself uses for dimensions); both are
// loaded on the first write, from the directory libhdf5 was found in.
constexpr unsigned kF_ACC_TRUNC = 0x0002u;
constexpr unsigned kCRT_ORDER = 0x0001u | 0x0002u;   // tracked | indexed
constexpr int kS_SCALAR = 0, kT_STR_NULLTERM = 0, kT_CSET_ASCII = 0;
constexpr int kLIBVER_EARLIEST = 0, kLIBVER_V18 = 1;
constexpr hsize_t kS_UNLIMITED = ~(hsize_t)0;

struct WApi {
    Api& A;
    void* hl = nullptr;
    hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t);
    hid_t (*Pcreate)(hid_t);
    herr_t (*Pclose)(hid_t);
    herr_t (*Pset_link_creation_order)(hid_t, unsigned);
    herr_t (*Pset_attr_creation_order)(hid_t, unsigned);
    herr_t (*Pset_libver_bounds)(hid_t, int, int);
    herr_t (*Pset_chunk)(hid_t, int, const hsize_t*);
    herr_t (*Pset_fill_value)(hid_t, hid_t, const void*);
    herr_t (*Pset_obj_track_times)(hid_t, int);
    hid_t (*Screate)(int);
    hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    hid_t (*Acreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Awrite)(hid_t, hid_t, const void*);
    herr_t (*Tset_cset)(hid_t, int);
    herr_t (*get_libversion)(unsigned*, unsigned*, unsigned*);
    herr_t (*DSset_scale)(hid_t, const char*);
    herr_t (*DSattach_scale)(hid_t, hid_t, unsigned);
    hid_t cls_fcreate, cls_faccess, cls_dcreate;
    hid_t t_i8, t_u8, t_i16, t_u16, t_i32, t_u32, t_i64, t_u64, t_f32, t_f64, t_f32be;

    template <class F> void hsym(F& f, const char* name) {
        f = reinterpret_cast<F>(dlsym(hl, name));
        if (!f) throw nc::Error(std::string("libhdf5_hl lacks ") + name);
    }
    explicit WApi(Api& a) : A(a) {
        // libhdf5_hl from the directory of the libhdf5 already loaded
        Dl_info info{};
        std::string dir;
        if (dladdr(reinterpret_cast<void*>(A.open_), &info) && info.dli_fname) {
            dir = info.dli_fname;
            dir = dir.substr(0, dir.find_last_of('/') + 1);
        }
        const char* env = std::getenv("CENTRE_HDF5_HL_LIB");
        const std::string cands[] = {env ? env : "", dir + "libhdf5_hl.so", "libhdf5_hl.so",
                                     "/opt/conda/lib/libhdf5_hl.so"};
        for (const std::string& c : cands)
            if (!c.empty() && (hl = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
        if (!hl)
            throw nc::Error("NetCDF-4 output needs libhdf5_hl (set CENTRE_HDF5_HL_LIB to its path)");
        A.sym(Fcreate, "H5Fcreate");
        A.sym(Pcreate, "H5Pcreate");
        A.sym(Pclose, "H5Pclose");
        A.sym(Pset_link_creation_order, "H5Pset_link_creation_order");
        A.sym(Pset_attr_creation_order, "H5Pset_attr_creation_order");
        A.sym(Pset_libver_bounds, "H5Pset_libver_bounds");
        A.sym(Pset_chunk, "H5Pset_chunk");
        A.sym(Pset_fill_value, "H5Pset_fill_value");
        A.sym(Pset_obj_track_times, "H5Pset_obj_track_times");
        A.sym(Screate, "H5Screate");
        A.sym(Dcreate2, "H5Dcreate2");
        A.sym(Dwrite, "H5Dwrite");
        A.sym(Acreate2, "H5Acreate2");
        A.sym(Awrite, "H5Awrite");
        A.sym(Tset_cset, "H5Tset_cset");
        A.sym(get_libversion, "H5get_libversion");
        hsym(DSset_scale, "H5DSset_scale");
        hsym(DSattach_scale, "H5DSattach_scale");
        cls_fcreate = A.global("H5P_CLS_FILE_CREATE_ID_g");
        cls_faccess = A.global("H5P_CLS_FILE_ACCESS_ID_g");
        cls_dcreate = A.global("H5P_CLS_DATASET_CREATE_ID_g");
        t_i8 = A.global("H5T_STD_I8LE_g");
        t_u8 = A.global("H5T_STD_U8LE_g");
        t_i16 = A.global("H5T_STD_I16LE_g");
        t_u16 = A.global("H5T_STD_U16LE_g");
        t_i32 = A.global("H5T_STD_I32LE_g");
        t_u32 = A.global("H5T_STD_U32LE_g");
        t_i64 = A.global("H5T_STD_I64LE_g");
        t_u64 = A.global("H5T_STD_U64LE_g");
        t_f32 = A.global("H5T_IEEE_F32LE_g");
        t_f64 = A.global("H5T_IEEE_F64LE_g");
        t_f32be = A.global("H5T_IEEE_F32BE_g");
    }
    // little-endian file type of a NetCDF type (also the memory type: the
    // host is little-endian, so no conversion happens)
    hid_t type_of(int t) const {
        switch (t) {
            case nc::BYTE: return t_i8;
            case nc::UBYTE: return t_u8;
            case nc::SHORT: return t_i16;
            case nc::USHORT: return t_u16;
            case nc::INT: return t_i32;
            case nc::UINT: return t_u32;
            case nc::INT64: return t_i64;
            case nc::UINT64: return t_u64;
            case nc::FLOAT: return t_f32;
            case nc::DOUBLE: return t_f64;
        }
        throw nc::Error("NetCDF-4 writer: unsupported variable type");
    }
};

WApi& wapi() {
    static WApi w(api());
    return w;
}

// NetCDF's default fill value of each type (netcdf.h NC_FILL_*), which
// netcdf-c records as the HDF5 fill value of every variable it creates.
std::vector<uint8_t> default_fill(int t) {
    auto bytes = [](auto v) {
        std::vector<uint8_t> b(sizeof(v));
        std::memcpy(b.data(), &v, sizeof(v));
        return b;
    };
    switch (t) {
        case nc::BYTE: return bytes((int8_t)-127);
        case nc::UBYTE: return bytes((uint8_t)255);
        case nc::SHORT: return bytes((int16_t)-32767);
        case nc::USHORT: return bytes((uint16_t)65535);
        case nc::INT: return bytes((int32_t)-2147483647);
        case nc::UINT: return bytes((uint32_t)4294967295u);
        case nc::INT64: return bytes((int64_t)-9223372036854775806LL);
        case nc::UINT64: return bytes((uint64_t)18446744073709551614ULL);
        case nc::FLOAT: return bytes(9.9692099683868690e+36f);
        case nc::DOUBLE: return bytes(9.9692099683868690e+36);
    }
    return {};
}

void ck_h5(long long rc, const std::string& what) {
    if (rc < 0) throw nc::Error("NetCDF-4 writer: " + what + " failed");
}

// scalar: a one-value attribute with a scalar dataspace (netcdf-c's form of
// _Netcdf4Dimid) instead of a one-element array
void put_attr(WApi& W, hid_t loc, const std::string& name, const nc::Attr& a, bool scalar = false) {
    Api& A = W.A;
    if (a.type == nc::CHAR) {
        // fixed-length, null-terminated ASCII of the text's own length, as
        // netcdf-c stores NC_CHAR attributes
        const std::string t = a.text();
        H ty(A.Tcopy(A.c_s1), A.Tclose);
        ck_h5(A.Tset_size(ty, t.empty() ? 1 : t.size()), "string type of " + name);
        ck_h5(A.Tset_strpad(ty, kT_STR_NULLTERM), "string padding of " + name);
        ck_h5(W.Tset_cset(ty, kT_CSET_ASCII), "string set of " + name);
        H sp(W.Screate(kS_SCALAR), A.Sclose);
        H at(W.Acreate2(loc, name.c_str(), ty, sp, kP_DEFAULT, kP_DEFAULT), A.Aclose);
        ck_h5(at, "attribute " + name);
        const std::string buf = t.empty() ? std::string(1, '\0') : t;
        ck_h5(W.Awrite(at, ty, buf.data()), "attribute " + name);
        return;
    }
    const hid_t ty = W.type_of(a.type);
    const hsize_t n = a.bytes.size() / nc::type_size(a.type);
    if (scalar && n != 1) throw nc::Error("scalar attribute " + name + " needs one value");
    H sp(scalar ? W.Screate(kS_SCALAR) : A.Screate_simple(1, &n, nullptr), A.Sclose);
    H at(W.Acreate2(loc, name.c_str(), ty, sp, kP_DEFAULT, kP_DEFAULT), A.Aclose);
    ck_h5(at, "attribute " + name);
    ck_h5(W.Awrite(at, ty, a.bytes.data()), "attribute " + name);
}

}  // namespace

void write_netcdf4(const nc::Writer& w, const std::string& path) {
    WApi& W = wapi();
    Api& A = W.A;
    H fcpl(W.Pcreate(W.cls_fcreate), W.Pclose);
    ck_h5(W.Pset_link_creation_order(fcpl, kCRT_ORDER), "link creation order");
    ck_h5(W.Pset_attr_creation_order(fcpl, kCRT_ORDER), "attribute creation order");
    H fapl(W.Pcreate(W.cls_faccess), W.Pclose);
    ck_h5(W.Pset_libver_bounds(fapl, kLIBVER_EARLIEST, kLIBVER_V18), "format bounds");
    H f(W.Fcreate(path.c_str(), kF_ACC_TRUNC, fcpl, fapl), A.Fclose);
    if (f < 0) throw nc::Error(path + ": cannot create as NetCDF-4/HDF5");
    for (auto& kv : w.attrs()) put_attr(W, f, kv.first, kv.second);

    // dimensions: one dimension scale per dimension, in definition order
    const auto& dims = w.dims();
    std::vector<std::unique_ptr<H>> scale;
    for (size_t d = 0; d < dims.size(); ++d) {
        const bool unl = (int)d == w.unlimited();
        H dcpl(W.Pcreate(W.cls_dcreate), W.Pclose);
        ck_h5(W.Pset_attr_creation_order(dcpl, kCRT_ORDER), "attribute creation order");
        ck_h5(W.Pset_obj_track_times(dcpl, 0), "time tracking");
        const hsize_t cur = unl ? 0 : dims[d].second, mx = unl ? kS_UNLIMITED : cur, one = 1;
        if (unl) ck_h5(W.Pset_chunk(dcpl, 1, &one), "chunks of " + dims[d].first);
        H sp(A.Screate_simple(1, &cur, &mx), A.Sclose);
        scale.push_back(std::make_unique<H>(
            W.Dcreate2(f, dims[d].first.c_str(), W.t_f32be, sp, kP_DEFAULT, dcpl, kP_DEFAULT),
            A.Dclose));
        const hid_t s = *scale.back();
        ck_h5(s, "dimension " + dims[d].first);
        char nm[96];
        std::snprintf(nm, sizeof nm, "This is a netCDF dimension but not a netCDF variable.%10llu",
                      (unsigned long long)cur);
        ck_h5(W.DSset_scale(s, nm), "dimension scale " + dims[d].first);
        put_attr(W, s, "_Netcdf4Dimid", nc::Attr::of<int32_t>(nc::INT, {(int32_t)d}), true);
    }
    // variables, each chunked when it has the unlimited dimension or chunks
    // were asked for
    for (const auto& v : w.vars()) {
        const size_t nd = v.dims.size();
        std::vector<hsize_t> cur(nd), mx(nd), ch(nd);
        bool has_unl = false;
        for (size_t k = 0; k < nd; ++k) {
            cur[k] = dims[v.dims[k]].second;
            has_unl |= v.dims[k] == w.unlimited();
            mx[k] = v.dims[k] == w.unlimited() ? kS_UNLIMITED : cur[k];
            ch[k] = !v.chunks.empty() ? v.chunks[k]
                    : (v.dims[k] == w.unlimited() ? 1 : std::max<hsize_t>(cur[k], 1));
        }
        H dcpl(W.Pcreate(W.cls_dcreate), W.Pclose);
        ck_h5(W.Pset_attr_creation_order(dcpl, kCRT_ORDER), "attribute creation order");
        ck_h5(W.Pset_obj_track_times(dcpl, 0), "time tracking");
        if (nd && (has_unl || !v.chunks.empty()))
            ck_h5(W.Pset_chunk(dcpl, (int)nd, ch.data()), "chunks of " + v.name);
        const hid_t ty = W.type_of(v.type);
        const std::vector<uint8_t> fill = default_fill(v.type);
        ck_h5(W.Pset_fill_value(dcpl, ty, fill.data()), "fill value of " + v.name);
        H sp(nd ? A.Screate_simple((int)nd, cur.data(), mx.data()) : W.Screate(kS_SCALAR), A.Sclose);
        H ds(W.Dcreate2(f, v.name.c_str(), ty, sp, kP_DEFAULT, dcpl, kP_DEFAULT), A.Dclose);
        ck_h5(ds, "variable " + v.name);
        if (nd > 1)
            put_attr(W, ds, "_Netcdf4Coordinates",
                     nc::Attr::of<int32_t>(nc::INT, std::vector<int32_t>(v.dims.begin(), v.dims.end())));
        for (size_t k = 0; k < nd; ++k)
            ck_h5(W.DSattach_scale(ds, *scale[v.dims[k]], (unsigned)k), "dimension list of " + v.name);
        for (auto& kv : v.attrs) put_attr(W, ds, kv.first, kv.second);
        if (!v.data.empty())
            ck_h5(W.Dwrite(ds, ty, kS_ALL, kS_ALL, kP_DEFAULT, v.data.data()), "data of " + v.name);
    }
    unsigned maj = 0, mnr = 0, rel = 0;
    W.get_libversion(&maj, &mnr, &rel);
    put_attr(W, f, "_NCProperties",
             nc::Attr::of_text("version=2,centre_hip=1.0.0,hdf5=" + std::to_string(maj) + "." +
                               std::to_string(mnr) + "." + std::to_string(rel)));
}

bool netcdf4_available() {
    try {
        wapi();
        return true;
    } catch (const nc::Error&) {
        return false;
    }
}

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
