// ncclassic.cpp -- NetCDF classic-format codec (see ncclassic.h).
#include "ncclassic.h"

#include <cstring>

namespace nc {

namespace {

const uint32_t TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C;

uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

void swap_bytes(uint8_t *p, size_t n, size_t esz) {
    if (esz == 1) return;
    for (size_t i = 0; i < n; i++) {
        uint8_t *q = p + i * esz;
        for (size_t a = 0, b = esz - 1; a < b; a++, b--) std::swap(q[a], q[b]);
    }
}

struct Buf {
    std::vector<uint8_t> b;
    void u32(uint32_t v) { for (int s = 24; s >= 0; s -= 8) b.push_back(uint8_t(v >> s)); }
    void u64(uint64_t v) { for (int s = 56; s >= 0; s -= 8) b.push_back(uint8_t(v >> s)); }
    void bytes(const void *p, size_t n) {
        const uint8_t *q = (const uint8_t *)p;
        b.insert(b.end(), q, q + n);
        while (b.size() % 4) b.push_back(0);
    }
};

struct In {
    const uint8_t *p, *e;
    bool ok = true;
    uint32_t u32() {
        if (e - p < 4) { ok = false; return 0; }
        uint32_t v = (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3];
        p += 4;
        return v;
    }
    uint64_t u64() {
        uint64_t hi = u32();
        return (hi << 32) | u32();
    }
    std::string str(uint64_t n) {
        if ((uint64_t)(e - p) < pad4(n)) { ok = false; return ""; }
        std::string s((const char *)p, n);
        p += pad4(n);
        return s;
    }
};

}  // namespace

size_t type_size(int t) {
    switch (t) {
    case NC_BYTE: case NC_CHAR: case NC_UBYTE: return 1;
    case NC_SHORT: case NC_USHORT: return 2;
    case NC_INT: case NC_FLOAT: case NC_UINT: return 4;
    case NC_DOUBLE: case NC_INT64: case NC_UINT64: return 8;
    default: return 0;
    }
}

Att Att::text(const std::string &n, const std::string &v) {
    Att a;
    a.name = n;
    a.type = NC_CHAR;
    a.data.assign(v.begin(), v.end());
    a.count = v.size();
    return a;
}
Att Att::f32(const std::string &n, float v) {
    Att a;
    a.name = n;
    a.type = NC_FLOAT;
    a.data.resize(4);
    memcpy(a.data.data(), &v, 4);
    a.count = 1;
    return a;
}
Att Att::f64(const std::string &n, double v) {
    Att a;
    a.name = n;
    a.type = NC_DOUBLE;
    a.data.resize(8);
    memcpy(a.data.data(), &v, 8);
    a.count = 1;
    return a;
}
Att Att::i32(const std::string &n, int v) {
    Att a;
    a.name = n;
    a.type = NC_INT;
    a.data.resize(4);
    memcpy(a.data.data(), &v, 4);
    a.count = 1;
    return a;
}
std::string Att::as_text() const { return std::string(data.begin(), data.end()); }
double Att::as_double(size_t i) const {
    const uint8_t *p = data.data() + i * type_size(type);
    switch (type) {
    case NC_BYTE: return (double)(int8_t)p[0];
    case NC_UBYTE: return (double)p[0];
    case NC_SHORT: { int16_t v; memcpy(&v, p, 2); return v; }
    case NC_USHORT: { uint16_t v; memcpy(&v, p, 2); return v; }
    case NC_INT: { int32_t v; memcpy(&v, p, 4); return v; }
    case NC_UINT: { uint32_t v; memcpy(&v, p, 4); return v; }
    case NC_FLOAT: { float v; memcpy(&v, p, 4); return v; }
    case NC_DOUBLE: { double v; memcpy(&v, p, 8); return v; }
    case NC_INT64: { int64_t v; memcpy(&v, p, 8); return (double)v; }
    case NC_UINT64: { uint64_t v; memcpy(&v, p, 8); return (double)v; }
    default: return 0.0;
    }
}

// ------------------------------------------------------------------------- reader

static bool read_atts(In &in, int version, std::vector<Att> &out) {
    uint32_t tag = in.u32();
    uint64_t n = version == 5 ? in.u64() : in.u32();
    if (tag == 0 && n == 0) return in.ok;
    if (tag != TAG_ATT) return false;
    for (uint64_t k = 0; k < n && in.ok; k++) {
        Att a;
        uint64_t ln = version == 5 ? in.u64() : in.u32();
        a.name = in.str(ln);
        a.type = (int)in.u32();
        a.count = version == 5 ? in.u64() : in.u32();
        size_t es = type_size(a.type);
        if (!es) return false;
        uint64_t nb = a.count * es;
        if ((uint64_t)(in.e - in.p) < pad4(nb)) return false;
        a.data.assign(in.p, in.p + nb);
        swap_bytes(a.data.data(), a.count, es);
        in.p += pad4(nb);
        out.push_back(a);
    }
    return in.ok;
}

bool Reader::open(const std::string &path, std::string *err) {
    f_ = fopen(path.c_str(), "rb");
    if (!f_) { if (err) *err = "cannot open " + path; return false; }
    // read a generous header prefix; grow if needed
    std::vector<uint8_t> h(1 << 16);
    for (;;) {
        fseek(f_, 0, SEEK_SET);
        size_t got = fread(h.data(), 1, h.size(), f_);
        if (got < 8) { if (err) *err = "not a NetCDF file: " + path; return false; }
        if (memcmp(h.data(), "CDF", 3) != 0) {
            if (err) *err = path + ": not a NetCDF classic file (NetCDF-4/HDF5 is not supported)";
            return false;
        }
        version_ = h[3];
        if (version_ != 1 && version_ != 2 && version_ != 5) { if (err) *err = "unknown CDF version"; return false; }
        In in{h.data() + 4, h.data() + got};
        dims_.clear(); gatts_.clear(); vars_.clear();
        numrecs_ = version_ == 5 ? in.u64() : in.u32();
        if (numrecs_ == 0xFFFFFFFFu) numrecs_ = 0;
        uint32_t tag = in.u32();
        uint64_t nd = version_ == 5 ? in.u64() : in.u32();
        bool ok = in.ok;
        if (ok && !(tag == 0 && nd == 0)) {
            ok = tag == TAG_DIM;
            for (uint64_t k = 0; ok && k < nd; k++) {
                Dim d;
                uint64_t ln = version_ == 5 ? in.u64() : in.u32();
                d.name = in.str(ln);
                d.len = version_ == 5 ? in.u64() : in.u32();
                if (d.len == 0) rec_dim_ = (int)k;
                dims_.push_back(d);
                ok = in.ok;
            }
        }
        if (ok) ok = read_atts(in, version_, gatts_);
        if (ok) {
            tag = in.u32();
            uint64_t nv = version_ == 5 ? in.u64() : in.u32();
            ok = in.ok;
            if (ok && !(tag == 0 && nv == 0)) {
                ok = tag == TAG_VAR;
                for (uint64_t k = 0; ok && k < nv; k++) {
                    Var v;
                    uint64_t ln = version_ == 5 ? in.u64() : in.u32();
                    v.name = in.str(ln);
                    uint64_t nvd = version_ == 5 ? in.u64() : in.u32();
                    for (uint64_t q = 0; q < nvd; q++) v.dims.push_back((int)(version_ == 5 ? in.u64() : in.u32()));
                    ok = read_atts(in, version_, v.atts);
                    v.type = (int)in.u32();
                    v.vsize = version_ == 5 ? in.u64() : in.u32();
                    v.begin = version_ == 1 ? in.u32() : in.u64();
                    v.is_record = !v.dims.empty() && v.dims[0] == rec_dim_;
                    if (v.is_record) recsize_ += v.vsize;
                    vars_.push_back(v);
                    ok = ok && in.ok;
                }
            }
        }
        if (ok) break;
        if (got < h.size()) { if (err) *err = "corrupt NetCDF header: " + path; return false; }
        h.resize(h.size() * 4);
        rec_dim_ = -1;
        recsize_ = 0;
    }
    int nrec = 0;
    for (auto &v : vars_) nrec += v.is_record;
    if (nrec == 1)
        for (auto &v : vars_)
            if (v.is_record) recsize_ = v.vsize;  // single record variable: no padding between records
    return true;
}

Reader::~Reader() {
    if (f_) fclose(f_);
}

int Reader::find_dim(const std::string &name) const {
    for (size_t i = 0; i < dims_.size(); i++)
        if (dims_[i].name == name) return (int)i;
    return -1;
}

int Reader::find_var(const std::string &name) const {
    for (size_t i = 0; i < vars_.size(); i++)
        if (vars_[i].name == name) return (int)i;
    return -1;
}

bool Reader::read_raw(int vi, std::vector<uint8_t> &raw, uint64_t &n, std::string *err) {
    const Var &v = vars_[vi];
    size_t es = type_size(v.type);
    n = 1;
    for (size_t k = v.is_record ? 1 : 0; k < v.dims.size(); k++) n *= dims_[v.dims[k]].len;
    uint64_t nrec = v.is_record ? numrecs_ : 1;
    raw.resize(n * es * nrec);
    for (uint64_t r = 0; r < nrec; r++) {
        uint64_t off = v.begin + (v.is_record ? r * recsize_ : 0);
        if (fseeko(f_, (off_t)off, SEEK_SET) != 0 ||
            fread(raw.data() + r * n * es, 1, n * es, f_) != n * es) {
            if (err) *err = "short read of variable " + v.name;
            return false;
        }
    }
    n *= nrec;
    swap_bytes(raw.data(), n, es);
    return true;
}

template <typename T>
static void convert(const std::vector<uint8_t> &raw, int type, uint64_t n, std::vector<T> &out) {
    out.resize(n);
    Att a;
    a.type = type;
    a.data = raw;
    for (uint64_t i = 0; i < n; i++) out[i] = (T)a.as_double(i);
}

bool Reader::read_double(int v, std::vector<double> &out, std::string *err) {
    std::vector<uint8_t> raw;
    uint64_t n;
    if (!read_raw(v, raw, n, err)) return false;
    if (vars_[v].type == NC_DOUBLE) {
        out.resize(n);
        memcpy(out.data(), raw.data(), n * 8);
    } else {
        convert(raw, vars_[v].type, n, out);
    }
    return true;
}

bool Reader::read_float(int v, std::vector<float> &out, std::string *err) {
    std::vector<uint8_t> raw;
    uint64_t n;
    if (!read_raw(v, raw, n, err)) return false;
    if (vars_[v].type == NC_FLOAT) {
        out.resize(n);
        memcpy(out.data(), raw.data(), n * 4);
    } else {
        convert(raw, vars_[v].type, n, out);
    }
    return true;
}

// ------------------------------------------------------------------------- writer

bool Writer::create(const std::string &path, int version, std::string *err) {
    version_ = version;
    f_ = fopen(path.c_str(), "wb+");
    if (!f_) { if (err) *err = "cannot create " + path; return false; }
    return true;
}

Writer::~Writer() {
    if (f_) {
        std::string e;
        close(&e);
    }
}

int Writer::def_dim(const std::string &name, uint64_t len) {
    dims_.push_back({name, len});
    if (len == 0) rec_dim_ = (int)dims_.size() - 1;
    return (int)dims_.size() - 1;
}

int Writer::def_var(const std::string &name, int type, const std::vector<int> &dims) {
    Var v;
    v.name = name;
    v.type = type;
    v.dims = dims;
    v.is_record = !dims.empty() && dims[0] == rec_dim_;
    vars_.push_back(v);
    return (int)vars_.size() - 1;
}

void Writer::put_att(int var, const Att &a) {
    if (var < 0) gatts_.push_back(a);
    else vars_[var].atts.push_back(a);
}

static void emit_atts(Buf &b, int version, const std::vector<Att> &atts) {
    if (atts.empty()) {
        b.u32(0);
        if (version == 5) b.u64(0); else b.u32(0);
        return;
    }
    b.u32(TAG_ATT);
    if (version == 5) b.u64(atts.size()); else b.u32((uint32_t)atts.size());
    for (const Att &a : atts) {
        if (version == 5) b.u64(a.name.size()); else b.u32((uint32_t)a.name.size());
        b.bytes(a.name.data(), a.name.size());
        b.u32((uint32_t)a.type);
        if (version == 5) b.u64(a.count); else b.u32((uint32_t)a.count);
        std::vector<uint8_t> d = a.data;
        swap_bytes(d.data(), a.count, type_size(a.type));
        b.bytes(d.data(), d.size());
    }
}

bool Writer::write_header(std::string *err) {
    Buf b;
    b.b = {'C', 'D', 'F', (uint8_t)version_};
    if (version_ == 5) b.u64(numrecs_); else b.u32((uint32_t)numrecs_);
    if (dims_.empty()) { b.u32(0); if (version_ == 5) b.u64(0); else b.u32(0); }
    else {
        b.u32(TAG_DIM);
        if (version_ == 5) b.u64(dims_.size()); else b.u32((uint32_t)dims_.size());
        for (const Dim &d : dims_) {
            if (version_ == 5) b.u64(d.name.size()); else b.u32((uint32_t)d.name.size());
            b.bytes(d.name.data(), d.name.size());
            if (version_ == 5) b.u64(d.len); else b.u32((uint32_t)d.len);
        }
    }
    emit_atts(b, version_, gatts_);
    if (vars_.empty()) { b.u32(0); if (version_ == 5) b.u64(0); else b.u32(0); }
    else {
        b.u32(TAG_VAR);
        if (version_ == 5) b.u64(vars_.size()); else b.u32((uint32_t)vars_.size());
        for (const Var &v : vars_) {
            if (version_ == 5) b.u64(v.name.size()); else b.u32((uint32_t)v.name.size());
            b.bytes(v.name.data(), v.name.size());
            if (version_ == 5) b.u64(v.dims.size()); else b.u32((uint32_t)v.dims.size());
            for (int d : v.dims) { if (version_ == 5) b.u64((uint64_t)d); else b.u32((uint32_t)d); }
            emit_atts(b, version_, v.atts);
            b.u32((uint32_t)v.type);
            if (version_ == 5) b.u64(v.vsize); else b.u32((uint32_t)(v.vsize > 0xFFFFFFFFull ? 0xFFFFFFFFu : v.vsize));
            if (version_ == 1) b.u32((uint32_t)v.begin); else b.u64(v.begin);
        }
    }
    if (fseeko(f_, 0, SEEK_SET) != 0 || fwrite(b.b.data(), 1, b.b.size(), f_) != b.b.size()) {
        if (err) *err = "header write failed";
        return false;
    }
    return true;
}

uint64_t Writer::header_size() const {
    Writer tmp;  // measure by emitting into a buffer through the same code path
    (void)tmp;
    uint64_t n = 4 + (version_ == 5 ? 8 : 4);
    auto nsz = [&](const std::string &s) { return (version_ == 5 ? 8 : 4) + pad4(s.size()); };
    auto cnt = [&]() { return (uint64_t)(version_ == 5 ? 8 : 4); };
    auto atts_sz = [&](const std::vector<Att> &atts) {
        uint64_t m = 4 + cnt();
        for (const Att &a : atts) m += nsz(a.name) + 4 + cnt() + pad4(a.data.size());
        return m;
    };
    n += 4 + cnt();
    for (const Dim &d : dims_) n += nsz(d.name) + cnt();
    n += atts_sz(gatts_);
    n += 4 + cnt();
    for (const Var &v : vars_)
        n += nsz(v.name) + cnt() + v.dims.size() * cnt() + atts_sz(v.atts) + 4 + cnt() + (version_ == 1 ? 4 : 8);
    return n;
}

bool Writer::enddef(std::string *err) {
    // sizes
    int nrec = 0;
    for (Var &v : vars_) {
        uint64_t n = type_size(v.type);
        for (size_t k = v.is_record ? 1 : 0; k < v.dims.size(); k++) n *= dims_[v.dims[k]].len;
        v.vsize = pad4(n);
        nrec += v.is_record;
        if (version_ != 5 && v.vsize > 0xFFFFFFFCull) {
            if (err) *err = "variable " + v.name + " needs CDF-5";
            return false;
        }
    }
    uint64_t off = pad4(header_size());
    for (Var &v : vars_)
        if (!v.is_record) { v.begin = off; off += v.vsize; }
    recsize_ = 0;
    for (Var &v : vars_)
        if (v.is_record) { v.begin = off + recsize_; recsize_ += v.vsize; }
    if (nrec == 1)
        for (Var &v : vars_)
            if (v.is_record) {
                uint64_t n = type_size(v.type);
                for (size_t k = 1; k < v.dims.size(); k++) n *= dims_[v.dims[k]].len;
                recsize_ = n;  // single record variable: records are not padded
            }
    defined_ = true;
    return write_header(err);
}

static bool write_at(FILE *f, uint64_t off, const void *data, size_t n, size_t es, std::string *err) {
    std::vector<uint8_t> tmp((const uint8_t *)data, (const uint8_t *)data + n * es);
    swap_bytes(tmp.data(), n, es);
    if (fseeko(f, (off_t)off, SEEK_SET) != 0 || fwrite(tmp.data(), 1, tmp.size(), f) != tmp.size()) {
        if (err) *err = "data write failed";
        return false;
    }
    return true;
}

bool Writer::put_var(int vi, const void *data, std::string *err) {
    const Var &v = vars_[vi];
    size_t es = type_size(v.type);
    uint64_t n = 1;
    for (int d : v.dims) n *= dims_[d].len;
    return write_at(f_, v.begin, data, n, es, err);
}

bool Writer::put_record(int vi, uint64_t rec, const void *data, std::string *err) {
    const Var &v = vars_[vi];
    size_t es = type_size(v.type);
    uint64_t n = 1;
    for (size_t k = 1; k < v.dims.size(); k++) n *= dims_[v.dims[k]].len;
    if (!write_at(f_, v.begin + rec * recsize_, data, n, es, err)) return false;
    if (rec + 1 > numrecs_) numrecs_ = rec + 1;
    return true;
}

bool Writer::sync(std::string *err) {
    if (!write_header(err)) return false;
    fflush(f_);
    return true;
}

bool Writer::close(std::string *err) {
    if (!f_) return true;
    bool ok = true;
    if (defined_) {
        // make the file length cover every fixed-size variable even if some were not written
        uint64_t end = 0;
        for (const Var &v : vars_)
            if (!v.is_record) end = std::max(end, v.begin + v.vsize);
        fseeko(f_, 0, SEEK_END);
        uint64_t cur = (uint64_t)ftello(f_);
        if (cur < end) {
            fseeko(f_, (off_t)(end - 1), SEEK_SET);
            fputc(0, f_);
        }
        ok = write_header(err);
    }
    fclose(f_);
    f_ = nullptr;
    return ok;
}

}  // namespace nc
