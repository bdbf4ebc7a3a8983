// NetCDF classic-family files (CDF-1, CDF-2, CDF-5) without netcdf-c.
//
// Written from the published NetCDF classic / 64-bit offset / 64-bit data
// format specification.  Covers what CENTRE reads and writes: the BAT input
// (Netcdf_BAT.cpp:24-330), the bin-code files (Netcdf_TrjInt.cpp:542-755,
// CDF-5), the count files (Netcdf_HistUtils.cpp:808-1029, CDF-5) and the
// entropy-contribution layout (Netcdf_EntContri.cpp:577-760) written as CDF-5.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace nc {

enum Type : int {
    BYTE = 1, CHAR = 2, SHORT = 3, INT = 4, FLOAT = 5, DOUBLE = 6,
    UBYTE = 7, USHORT = 8, UINT = 9, INT64 = 10, UINT64 = 11
};
int type_size(int t);

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Attr {
    int type = CHAR;
    std::vector<uint8_t> bytes;   // native byte order
    std::string text() const { return std::string(bytes.begin(), bytes.end()); }
    template <class T> std::vector<T> values() const {
        std::vector<T> v(bytes.size() / sizeof(T));
        std::copy(bytes.begin(), bytes.begin() + v.size() * sizeof(T),
                  reinterpret_cast<uint8_t*>(v.data()));
        return v;
    }
    static Attr of_text(const std::string& s);
    template <class T> static Attr of(int type, const std::vector<T>& v) {
        Attr a;
        a.type = type;
        a.bytes.assign(reinterpret_cast<const uint8_t*>(v.data()),
                       reinterpret_cast<const uint8_t*>(v.data()) + v.size() * sizeof(T));
        return a;
    }
};

struct Var {
    std::string name;
    std::vector<int> dims;          // dimension indices
    std::map<std::string, Attr> attrs;
    int type = DOUBLE;
    uint64_t vsize = 0, begin = 0;
    bool record = false;
};

// Read-only view of a file (the whole file is loaded).
class Reader {
public:
    explicit Reader(const std::string& path);
    int version() const { return version_; }
    uint64_t numrecs() const { return numrecs_; }
    const std::vector<std::pair<std::string, uint64_t>>& dims() const { return dims_; }
    const std::map<std::string, Attr>& attrs() const { return gatts_; }
    bool has_var(const std::string& n) const { return vars_.count(n) != 0; }
    const Var& var(const std::string& n) const;
    std::vector<uint64_t> shape(const std::string& n) const;
    // Hyperslab read converted to T (T must match the stored type size class:
    // integers to integers of any width, FLOAT/DOUBLE to double/float).
    template <class T>
    std::vector<T> read(const std::string& n, std::vector<uint64_t> start = {},
                        std::vector<uint64_t> count = {}) const;

private:
    std::vector<uint8_t> buf_;
    int version_ = 1;
    uint64_t numrecs_ = 0, recsize_ = 0;
    int unlimited_ = -1;
    std::vector<std::pair<std::string, uint64_t>> dims_;
    std::map<std::string, Attr> gatts_;
    std::map<std::string, Var> vars_;
    double element(const Var& v, uint64_t byte_off) const;
    int64_t ielement(const Var& v, uint64_t byte_off) const;
};

// Whole-file writer: declare dims, attrs and variables with their full data,
// then write().  Record variables have the unlimited dimension first.
class Writer {
public:
    explicit Writer(int version = 5) : version_(version) {}
    int add_dim(const std::string& name, uint64_t len, bool unlimited = false);
    void add_attr(const std::string& name, const Attr& a) { gatts_.emplace_back(name, a); }
    // data: native-endian bytes of the full variable
    void add_var(const std::string& name, int type, const std::vector<int>& dims,
                 const void* data, const std::map<std::string, Attr>& attrs = {});
    // Chunk shape of a variable for the NetCDF-4 encoding (the classic
    // formats have no chunks and ignore it).
    void set_chunks(const std::string& var, const std::vector<uint64_t>& chunks);
    void write(const std::string& path) const;

    struct WVar {
        std::string name;
        int type;
        std::vector<int> dims;
        std::vector<uint8_t> data;
        std::vector<std::pair<std::string, Attr>> attrs;
        std::vector<uint64_t> chunks;
    };
    // The declared file, for other encoders of the same description.
    const std::vector<std::pair<std::string, uint64_t>>& dims() const { return dims_; }
    int unlimited() const { return unlimited_; }
    const std::vector<std::pair<std::string, Attr>>& attrs() const { return gatts_; }
    const std::vector<WVar>& vars() const { return vars_; }

private:
    int version_;
    int unlimited_ = -1;
    std::vector<std::pair<std::string, uint64_t>> dims_;
    std::vector<std::pair<std::string, Attr>> gatts_;
    std::vector<WVar> vars_;
};

}  // namespace nc
