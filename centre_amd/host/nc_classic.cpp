// NetCDF classic-family reader and writer (see nc_classic.hpp).
#include "nc_classic.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <numeric>

namespace nc {

int type_size(int t) {
    switch (t) {
        case BYTE: case CHAR: case UBYTE: return 1;
        case SHORT: case USHORT: return 2;
        case INT: case FLOAT: case UINT: return 4;
        case DOUBLE: case INT64: case UINT64: return 8;
    }
    throw Error("unknown nc_type " + std::to_string(t));
}

static uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

static void swap_to_native(uint8_t* p, uint64_t n, int ts) {
    if (ts == 1) return;
    for (uint64_t i = 0; i < n; ++i) std::reverse(p + i * ts, p + (i + 1) * ts);
}

Attr Attr::of_text(const std::string& s) {
    Attr a;
    a.type = CHAR;
    a.bytes.assign(s.begin(), s.end());
    return a;
}

// ---------------------------------------------------------------- reader --
namespace {
struct Cursor {
    const std::vector<uint8_t>& b;
    uint64_t p = 0;
    int version = 1;
    uint32_t u32() {
        if (p + 4 > b.size()) throw Error("truncated header");
        uint32_t v = (uint32_t(b[p]) << 24) | (uint32_t(b[p + 1]) << 16) |
                     (uint32_t(b[p + 2]) << 8) | b[p + 3];
        p += 4;
        return v;
    }
    uint64_t u64() {
        uint64_t hi = u32();
        return (hi << 32) | u32();
    }
    uint64_t nonneg() { return version == 5 ? u64() : u32(); }
    uint64_t offset() { return version == 1 ? u32() : u64(); }
    std::string name() {
        uint64_t n = nonneg();
        if (p + n > b.size()) throw Error("truncated name");
        std::string s(b.begin() + p, b.begin() + p + n);
        p += pad4(n);
        return s;
    }
    std::map<std::string, Attr> atts() {
        std::map<std::string, Attr> out;
        uint32_t tag = u32();
        uint64_t n = nonneg();
        if (tag == 0 && n == 0) return out;
        if (tag != 0x0C) throw Error("expected NC_ATTRIBUTE");
        for (uint64_t i = 0; i < n; ++i) {
            std::string nm = name();
            Attr a;
            a.type = (int)u32();
            uint64_t k = nonneg();
            const int ts = type_size(a.type);
            if (p + k * ts > b.size()) throw Error("truncated attribute");
            a.bytes.assign(b.begin() + p, b.begin() + p + k * ts);
            swap_to_native(a.bytes.data(), k, ts);
            p += pad4(k * ts);
            out[nm] = a;
        }
        return out;
    }
};
}  // namespace

Reader::Reader(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("cannot open " + path);
    buf_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    if (buf_.size() < 4 || buf_[0] != 'C' || buf_[1] != 'D' || buf_[2] != 'F')
        throw Error(path + ": not a NetCDF classic-format file");
    version_ = buf_[3];
    if (version_ != 1 && version_ != 2 && version_ != 5)
        throw Error(path + ": unsupported CDF version");
    Cursor c{buf_, 4, version_};
    uint64_t nr = c.nonneg();
    const bool streaming = (version_ != 5 && nr == 0xFFFFFFFFull);
    uint32_t tag = c.u32();
    uint64_t nd = c.nonneg();
    if (!(tag == 0 && nd == 0)) {
        if (tag != 0x0A) throw Error("expected NC_DIMENSION");
        for (uint64_t i = 0; i < nd; ++i) {
            std::string nm = c.name();
            uint64_t len = c.nonneg();
            if (len == 0) unlimited_ = (int)i;
            dims_.emplace_back(nm, len);
        }
    }
    gatts_ = c.atts();
    tag = c.u32();
    uint64_t nv = c.nonneg();
    std::vector<std::string> recvars;
    if (!(tag == 0 && nv == 0)) {
        if (tag != 0x0B) throw Error("expected NC_VARIABLE");
        for (uint64_t i = 0; i < nv; ++i) {
            Var v;
            v.name = c.name();
            uint64_t k = c.nonneg();
            for (uint64_t j = 0; j < k; ++j) v.dims.push_back((int)c.nonneg());
            v.attrs = c.atts();
            v.type = (int)c.u32();
            v.vsize = c.nonneg();
            v.begin = c.offset();
            v.record = !v.dims.empty() && v.dims[0] == unlimited_;
            if (v.record) recvars.push_back(v.name);
            vars_[v.name] = v;
        }
    }
    if (recvars.size() == 1) {
        const Var& v = vars_[recvars[0]];
        uint64_t n = type_size(v.type);
        for (size_t j = 1; j < v.dims.size(); ++j) n *= dims_[v.dims[j]].second;
        recsize_ = n;
    } else {
        for (auto& n : recvars) recsize_ += vars_[n].vsize;
    }
    numrecs_ = nr;
    if (streaming && !recvars.empty()) {
        uint64_t first = UINT64_MAX;
        for (auto& n : recvars) first = std::min(first, vars_[n].begin);
        numrecs_ = recsize_ ? (buf_.size() - first) / recsize_ : 0;
    }
    if (unlimited_ >= 0) dims_[unlimited_].second = numrecs_;
}

const Var& Reader::var(const std::string& n) const {
    auto it = vars_.find(n);
    if (it == vars_.end()) throw Error("no variable " + n);
    return it->second;
}

std::vector<uint64_t> Reader::shape(const std::string& n) const {
    std::vector<uint64_t> s;
    for (int d : var(n).dims) s.push_back(dims_[d].second);
    return s;
}

double Reader::element(const Var& v, uint64_t off) const {
    const uint8_t* p = buf_.data() + off;
    uint8_t tmp[8];
    const int ts = type_size(v.type);
    for (int i = 0; i < ts; ++i) tmp[i] = p[ts - 1 - i];
    switch (v.type) {
        case FLOAT: { float f; std::memcpy(&f, tmp, 4); return f; }
        case DOUBLE: { double d; std::memcpy(&d, tmp, 8); return d; }
        default: return (double)ielement(v, off);
    }
}

int64_t Reader::ielement(const Var& v, uint64_t off) const {
    const uint8_t* p = buf_.data() + off;
    uint64_t u = 0;
    const int ts = type_size(v.type);
    for (int i = 0; i < ts; ++i) u = (u << 8) | p[i];
    switch (v.type) {
        case BYTE: return (int8_t)u;
        case SHORT: return (int16_t)u;
        case INT: return (int32_t)u;
        case INT64: return (int64_t)u;
        case FLOAT: case DOUBLE: return (int64_t)element(v, off);
        default: return (int64_t)u;
    }
}

template <class T>
std::vector<T> Reader::read(const std::string& n, std::vector<uint64_t> start,
                            std::vector<uint64_t> count) const {
    const Var& v = var(n);
    const std::vector<uint64_t> shp = shape(n);
    const size_t nd = shp.size();
    if (start.empty()) start.assign(nd, 0);
    if (count.empty()) {
        count.resize(nd);
        for (size_t i = 0; i < nd; ++i) count[i] = shp[i] - start[i];
    }
    if (start.size() != nd || count.size() != nd) throw Error(n + ": rank mismatch");
    for (size_t i = 0; i < nd; ++i)
        if (start[i] + count[i] > shp[i]) throw Error(n + ": hyperslab out of range");
    const int ts = type_size(v.type);
    // strides in elements within one record (or the whole variable)
    std::vector<uint64_t> stride(nd, 1);
    for (int i = (int)nd - 2; i >= 0; --i) stride[i] = stride[i + 1] * shp[i + 1];
    uint64_t total = 1;
    for (auto c : count) total *= c;
    std::vector<T> out(total);
    std::vector<uint64_t> idx(nd, 0);
    const bool fp = (v.type == FLOAT || v.type == DOUBLE);
    for (uint64_t k = 0; k < total; ++k) {
        uint64_t off;
        if (v.record) {
            uint64_t inner = 0;
            for (size_t i = 1; i < nd; ++i) inner += (start[i] + idx[i]) * stride[i];
            off = v.begin + (start[0] + idx[0]) * recsize_ + inner * ts;
        } else {
            uint64_t e = 0;
            for (size_t i = 0; i < nd; ++i) e += (start[i] + idx[i]) * stride[i];
            off = v.begin + e * ts;
        }
        out[k] = fp ? (T)element(v, off) : (T)ielement(v, off);
        for (int i = (int)nd - 1; i >= 0; --i) {
            if (++idx[i] < count[i]) break;
            idx[i] = 0;
        }
    }
    return out;
}

template std::vector<double> Reader::read<double>(const std::string&, std::vector<uint64_t>, std::vector<uint64_t>) const;
template std::vector<float> Reader::read<float>(const std::string&, std::vector<uint64_t>, std::vector<uint64_t>) const;
template std::vector<uint8_t> Reader::read<uint8_t>(const std::string&, std::vector<uint64_t>, std::vector<uint64_t>) const;
template std::vector<uint32_t> Reader::read<uint32_t>(const std::string&, std::vector<uint64_t>, std::vector<uint64_t>) const;
template std::vector<uint64_t> Reader::read<uint64_t>(const std::string&, std::vector<uint64_t>, std::vector<uint64_t>) const;
template std::vector<int32_t> Reader::read<int32_t>(const std::string&, std::vector<uint64_t>, std::vector<uint64_t>) const;

// ---------------------------------------------------------------- writer --
int Writer::add_dim(const std::string& name, uint64_t len, bool unlimited) {
    if (unlimited) {
        if (unlimited_ >= 0) throw Error("one unlimited dimension only");
        unlimited_ = (int)dims_.size();
    }
    dims_.emplace_back(name, len);
    return (int)dims_.size() - 1;
}

void Writer::add_var(const std::string& name, int type, const std::vector<int>& dims,
                     const void* data, const std::map<std::string, Attr>& attrs) {
    WVar v;
    v.name = name;
    v.type = type;
    v.dims = dims;
    uint64_t n = type_size(type);
    for (int d : dims) n *= dims_.at(d).second;
    v.data.assign((const uint8_t*)data, (const uint8_t*)data + n);
    for (auto& kv : attrs) v.attrs.emplace_back(kv.first, kv.second);
    vars_.push_back(std::move(v));
}

namespace {
struct Out {
    std::vector<uint8_t> b;
    int version;
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void u64(uint64_t v) {
        u32((uint32_t)(v >> 32));
        u32((uint32_t)v);
    }
    void nonneg(uint64_t v) { version == 5 ? u64(v) : u32((uint32_t)v); }
    void offset(uint64_t v) { version == 1 ? u32((uint32_t)v) : u64(v); }
    void name(const std::string& s) {
        nonneg(s.size());
        b.insert(b.end(), s.begin(), s.end());
        b.resize(b.size() + (pad4(s.size()) - s.size()), 0);
    }
    void big(const uint8_t* p, uint64_t n, int ts) {
        const size_t at = b.size();
        b.insert(b.end(), p, p + n * ts);
        swap_to_native(b.data() + at, n, ts);   // same swap both ways
    }
    void atts(const std::vector<std::pair<std::string, Attr>>& a) {
        if (a.empty()) {
            u32(0);
            nonneg(0);
            return;
        }
        u32(0x0C);
        nonneg(a.size());
        for (auto& kv : a) {
            name(kv.first);
            u32((uint32_t)kv.second.type);
            const int ts = type_size(kv.second.type);
            const uint64_t k = kv.second.bytes.size() / ts;
            nonneg(k);
            big(kv.second.bytes.data(), k, ts);
            b.resize(b.size() + (pad4(k * ts) - k * ts), 0);
        }
    }
};
}  // namespace

void Writer::set_chunks(const std::string& var, const std::vector<uint64_t>& chunks) {
    for (auto& v : vars_)
        if (v.name == var) {
            if (chunks.size() != v.dims.size()) throw Error("chunk rank mismatch for " + var);
            v.chunks = chunks;
            return;
        }
    throw Error("no variable " + var);
}

void Writer::write(const std::string& path) const {
    const uint64_t numrecs = unlimited_ >= 0 ? dims_[unlimited_].second : 0;
    std::vector<const WVar*> fixed, rec;
    for (auto& v : vars_)
        (!v.dims.empty() && v.dims[0] == unlimited_ ? rec : fixed).push_back(&v);
    auto inner_bytes = [&](const WVar& v) {
        uint64_t n = type_size(v.type);
        for (size_t i = (!v.dims.empty() && v.dims[0] == unlimited_) ? 1 : 0; i < v.dims.size(); ++i)
            n *= dims_[v.dims[i]].second;
        return n;
    };
    auto header = [&](const std::vector<uint64_t>& begins) {
        Out o{{}, version_};
        o.b = {'C', 'D', 'F', (uint8_t)version_};
        o.nonneg(numrecs);
        if (dims_.empty()) {
            o.u32(0);
            o.nonneg(0);
        } else {
            o.u32(0x0A);
            o.nonneg(dims_.size());
            for (size_t i = 0; i < dims_.size(); ++i) {
                o.name(dims_[i].first);
                o.nonneg((int)i == unlimited_ ? 0 : dims_[i].second);
            }
        }
        o.atts(gatts_);
        if (vars_.empty()) {
            o.u32(0);
            o.nonneg(0);
        } else {
            o.u32(0x0B);
            o.nonneg(vars_.size());
            size_t k = 0;
            for (auto* list : {&fixed, &rec}) {
                for (const WVar* v : *list) {
                    o.name(v->name);
                    o.nonneg(v->dims.size());
                    for (int d : v->dims) o.nonneg(d);
                    o.atts(v->attrs);
                    o.u32((uint32_t)v->type);
                    o.nonneg(pad4(inner_bytes(*v)));
                    o.offset(begins.empty() ? 0 : begins[k]);
                    ++k;
                }
            }
        }
        return o.b;
    };
    const uint64_t hlen = header({}).size();
    std::vector<uint64_t> begins;
    uint64_t pos = hlen;
    for (const WVar* v : fixed) { begins.push_back(pos); pos += pad4(inner_bytes(*v)); }
    for (const WVar* v : rec) { begins.push_back(pos); pos += pad4(inner_bytes(*v)); }
    Out body{header(begins), version_};
    for (const WVar* v : fixed) {
        const int ts = type_size(v->type);
        body.big(v->data.data(), v->data.size() / ts, ts);
        body.b.resize(body.b.size() + (pad4(v->data.size()) - v->data.size()), 0);
    }
    const bool single = rec.size() == 1;
    for (uint64_t r = 0; r < numrecs; ++r) {
        for (const WVar* v : rec) {
            const int ts = type_size(v->type);
            const uint64_t ib = inner_bytes(*v);
            body.big(v->data.data() + r * ib, ib / ts, ts);
            if (!single) body.b.resize(body.b.size() + (pad4(ib) - ib), 0);
        }
    }
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw Error("cannot create " + path);
    f.write((const char*)body.b.data(), (std::streamsize)body.b.size());
    if (!f) throw Error("write failed: " + path);
}

}  // namespace nc
