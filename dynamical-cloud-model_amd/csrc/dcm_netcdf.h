// dcm_netcdf.h -- a writer of the NetCDF classic format for the model's
// output file (the file `output` in driver_code.f90:725-1145 creates and
// appends to), host code only.
//
// The file holds, in definition order (driver_code.f90:794-958):
//   dims  times (unlimited), ip, jp, kp [, nq, nprec, l_q_names = 20]
//   vars  time(times) x(ip) y(jp) z(kp) rhoa(kp) thetan(kp) trefn(kp)
//         th u v w p div (times, ip, jp, kp)
//         [q_names(nq, l_q_names) q(times, ip, jp, kp, nq)
//          precip(times, ip, jp, kp, nprec)]
// every numeric variable NC_FLOAT with a "units" attribute (q_names has
// none).  The dimension lists are in C order, so k varies fastest in th..div
// and the field index fastest in q and precip (the reshapes of :1088-1100).
//
// The byte layout follows the published NetCDF classic format specification
// (CDF-1; CDF-2, with 64-bit offsets, when a variable would not fit the
// 32-bit one): big-endian header, fixed-size variables first in definition
// order, then the records, each record holding every record variable in
// definition order.  A record is therefore one contiguous block of bytes,
// which is what lets the GPU path produce it in one pass.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace dcm {

enum NcType { NC_CHAR_T = 2, NC_FLOAT_T = 5 };

struct NcVar {
    std::string name;
    std::vector<int> dims;  // dimension ids, C order
    int type = NC_FLOAT_T;
    std::string units;      // empty: no attribute
    bool record = false;
    uint64_t vsize = 0;     // bytes of one record (record vars) or of the var
    uint64_t begin = 0;
};

// the reference's output file layout for one grid
struct NcLayout {
    int ip = 0, jp = 0, kp = 0, nq = 0, nprec = 0;
    bool moisture = false;
    int version = 1;
    std::vector<std::pair<std::string, uint64_t>> dims;  // 0: unlimited
    std::vector<NcVar> vars;
    uint64_t header_bytes = 0, recsize = 0, rec_begin = 0;
    // byte offsets of the record variables inside one record
    uint64_t off_time = 0, off_th = 0, off_u = 0, off_v = 0, off_w = 0, off_p = 0,
             off_div = 0, off_q = 0, off_precip = 0;

    // returns false (and why) if a variable is too large for the format
    bool build(int ip, int jp, int kp, bool moisture, int nq, int nprec, std::string *why);
    std::vector<unsigned char> header(uint32_t numrecs) const;
    const NcVar *var(const std::string &name) const;
    uint64_t cells() const { return (uint64_t)ip * jp * kp; }
};

// NC_FILL_FLOAT, the value NetCDF gives data that was never written
constexpr uint32_t kNcFillFloatBits = 0x7cf00000u;

inline uint32_t to_be32(uint32_t x) { return __builtin_bswap32(x); }
uint32_t float_be(double v);  // (float)v, big-endian bits

// An open output file: header and fixed variables written at create.
class NcFile {
public:
    NcLayout L;
    int fd = -1;
    uint32_t numrecs = 0;
    std::string path;

    // create (clobbering) and write the header and the fixed variables; the
    // 1-D arrays are the reference's (0:n+1) arrays, q_names nq*20 chars
    bool create(const std::string &path, const NcLayout &layout, const double *x,
                const double *y, const double *z, const double *rhoa, const double *thetan,
                const double *trefn, const char *q_names, std::string *why);
    // write record `rec` (one block of L.recsize bytes) and raise numrecs
    bool put_record(uint32_t rec, const unsigned char *block, std::string *why);
    bool close(std::string *why);
    ~NcFile();
};

// Pack one record from host arrays in the reference's padded layout
// ((0:kp+1, 0:jp+1, 0:ip+1), q with nq fields after each other, precip
// (1:kp, 0:jp+1, 0:ip+1, nprec)); precip may be null (fill value).
void pack_record_host(const NcLayout &L, double time, const double *u, const double *v,
                      const double *w, const double *th, const double *p, const double *div,
                      const double *q, const double *precip, unsigned char *block);
// the precip section of a record only
void pack_precip_host(const NcLayout &L, const double *precip, unsigned char *block);

}  // namespace dcm
