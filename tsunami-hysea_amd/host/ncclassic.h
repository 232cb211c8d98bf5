// ncclassic.h -- NetCDF classic-format (CDF-1, CDF-2, CDF-5) reader and writer.
//
// The reference reads its grids and writes its outputs through netcdf-c
// (src/GPU/netcdfTH.cu, src/GPU/netcdfSeries.cu); neither this machine nor the GPU
// box has netcdf-c, so the host carries its own codec of the published classic
// format (big-endian header: magic, numrecs, dim/att/var lists; fixed-size variables
// at their begin offsets, record variables interleaved per record).
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace nc {

enum Type { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6,
            NC_UBYTE = 7, NC_USHORT = 8, NC_UINT = 9, NC_INT64 = 10, NC_UINT64 = 11 };

size_t type_size(int t);

struct Att {
    std::string name;
    int type = NC_CHAR;
    std::vector<uint8_t> data;  // host byte order for numbers, raw for text
    size_t count = 0;
    static Att text(const std::string &n, const std::string &v);
    static Att f32(const std::string &n, float v);
    static Att f64(const std::string &n, double v);
    static Att i32(const std::string &n, int v);
    std::string as_text() const;
    double as_double(size_t i = 0) const;
};

struct Dim {
    std::string name;
    uint64_t len = 0;  // 0 = record (unlimited) dimension
};

struct Var {
    std::string name;
    int type = NC_FLOAT;
    std::vector<int> dims;
    std::vector<Att> atts;
    uint64_t vsize = 0, begin = 0;
    bool is_record = false;
};

class Reader {
  public:
    bool open(const std::string &path, std::string *err);
    ~Reader();
    int find_dim(const std::string &name) const;
    int find_var(const std::string &name) const;
    uint64_t dim_len(int d) const { return d == rec_dim_ ? numrecs_ : dims_[d].len; }
    const Var &var(int v) const { return vars_[v]; }
    const std::vector<Att> &global_atts() const { return gatts_; }
    // whole non-record (or single-record) variable converted to double / float
    bool read_double(int v, std::vector<double> &out, std::string *err);
    bool read_float(int v, std::vector<float> &out, std::string *err);
    uint64_t numrecs() const { return numrecs_; }

  private:
    FILE *f_ = nullptr;
    int version_ = 1;
    uint64_t numrecs_ = 0;
    int rec_dim_ = -1;
    uint64_t recsize_ = 0;
    std::vector<Dim> dims_;
    std::vector<Att> gatts_;
    std::vector<Var> vars_;
    bool read_raw(int v, std::vector<uint8_t> &raw, uint64_t &n, std::string *err);
};

class Writer {
  public:
    // version 2 (64-bit offsets) unless a variable needs CDF-5
    bool create(const std::string &path, int version, std::string *err);
    ~Writer();
    int def_dim(const std::string &name, uint64_t len);
    int def_var(const std::string &name, int type, const std::vector<int> &dims);
    void put_att(int var, const Att &a);  // var = -1: global
    bool enddef(std::string *err);
    // whole fixed-size variable (host-order values of the variable's type)
    bool put_var(int var, const void *data, std::string *err);
    // one record of a record variable
    bool put_record(int var, uint64_t rec, const void *data, std::string *err);
    bool sync(std::string *err);
    bool close(std::string *err);

  private:
    FILE *f_ = nullptr;
    int version_ = 2;
    bool defined_ = false;
    uint64_t numrecs_ = 0, recsize_ = 0;
    int rec_dim_ = -1;
    std::vector<Dim> dims_;
    std::vector<Att> gatts_;
    std::vector<Var> vars_;
    bool write_header(std::string *err);
    uint64_t header_size() const;
};

}  // namespace nc
