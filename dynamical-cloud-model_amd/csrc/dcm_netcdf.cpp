// dcm_netcdf.cpp -- the NetCDF classic writer of the model's output file
// (driver_code.f90:725-1145) and its host C-ABI (dcm_nc_*).
#include "dcm_netcdf.h"

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cstring>

#include "../../include/dcm_hip.h"

namespace dcm {

void set_error(const std::string &msg);

namespace {

constexpr uint32_t kNcDimension = 0x0a, kNcVariable = 0x0b, kNcAttribute = 0x0c;

struct Out {
    std::vector<unsigned char> b;
    void u32(uint32_t x) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(x >> s));
    }
    void u64(uint64_t x) {
        for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(x >> s));
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void bytes(const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        b.insert(b.end(), c, c + n);
        pad();
    }
    void name(const std::string &s) {
        u32((uint32_t)s.size());
        bytes(s.data(), s.size());
    }
};

uint64_t round4(uint64_t n) { return (n + 3) & ~(uint64_t)3; }

bool write_all(int fd, const void *p, size_t n, uint64_t off, std::string *why) {
    const unsigned char *c = (const unsigned char *)p;
    while (n > 0) {
        const ssize_t w = pwrite(fd, c, n, (off_t)off);
        if (w < 0) {
            if (errno == EINTR) continue;
            if (why) *why = std::string("write failed: ") + std::strerror(errno);
            return false;
        }
        c += w;
        n -= (size_t)w;
        off += (uint64_t)w;
    }
    return true;
}

}  // namespace

uint32_t float_be(double v) {
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return to_be32(u);
}

bool NcLayout::build(int ip_, int jp_, int kp_, bool moisture_, int nq_, int nprec_,
                     std::string *why) {
    ip = ip_;
    jp = jp_;
    kp = kp_;
    moisture = moisture_;
    nq = moisture ? nq_ : 0;
    nprec = moisture ? nprec_ : 0;
    // a dimension of length 0 is the record dimension in the classic format
    // (nf90_def_dim with 0 is NF90_UNLIMITED), so nq and nprec must be >= 1
    if (ip < 1 || jp < 1 || kp < 1 || (moisture && (nq < 1 || nprec < 1))) {
        if (why) *why = "output layout: bad grid or field counts (moisture needs nq, nprec >= 1)";
        return false;
    }
    dims = {{"times", 0}, {"ip", (uint64_t)ip}, {"jp", (uint64_t)jp}, {"kp", (uint64_t)kp}};
    if (moisture) {
        dims.push_back({"nq", (uint64_t)nq});
        dims.push_back({"nprec", (uint64_t)nprec});
        dims.push_back({"l_q_names", 20});
    }
    vars.clear();
    auto add = [&](const char *n, std::vector<int> d, const char *units, int type = NC_FLOAT_T) {
        NcVar v;
        v.name = n;
        v.dims = d;
        v.units = units;
        v.type = type;
        v.record = !d.empty() && d[0] == 0;
        vars.push_back(v);
    };
    add("time", {0}, "seconds");
    add("x", {1}, "m");
    add("y", {2}, "m");
    add("z", {3}, "m");
    add("rhoa", {3}, "kg/m3");
    add("thetan", {3}, "K");
    add("trefn", {3}, "K");
    add("th", {0, 1, 2, 3}, "K");
    add("u", {0, 1, 2, 3}, "m/s");
    add("v", {0, 1, 2, 3}, "m/s");
    add("w", {0, 1, 2, 3}, "m/s");
    add("p", {0, 1, 2, 3}, "pa");
    add("div", {0, 1, 2, 3}, "m/s**2");
    if (moisture) {
        add("q_names", {4, 6}, "", NC_CHAR_T);
        add("q", {0, 1, 2, 3, 4}, "kg or number per kg");
        add("precip", {0, 1, 2, 3, 5}, "mm hr-1 or number per m-2 s-1");
    }
    uint64_t largest = 0;
    for (NcVar &v : vars) {
        uint64_t n = v.type == NC_FLOAT_T ? 4 : 1;
        for (size_t d = v.record ? 1 : 0; d < v.dims.size(); ++d) n *= dims[v.dims[d]].second;
        v.vsize = round4(n);
        if (v.vsize > largest) largest = v.vsize;
    }
    // CDF-1 keeps every variable (one record of it) below 2 GiB and every
    // begin offset 32-bit; CDF-2 allows 4 GiB and 64-bit offsets
    if (largest >= (1ull << 32) - 4) {
        if (why) *why = "output layout: a variable exceeds the classic format's 4 GiB";
        return false;
    }
    version = largest < (1ull << 31) - 4 ? 1 : 2;
    for (;;) {
        header_bytes = header(0).size();
        uint64_t off = header_bytes;
        for (NcVar &v : vars)
            if (!v.record) {
                v.begin = off;
                off += v.vsize;
            }
        rec_begin = off;
        uint64_t last_begin = 0;
        for (NcVar &v : vars)
            if (v.record) {
                v.begin = off;
                last_begin = off;
                off += v.vsize;
            }
        recsize = off - rec_begin;
        if (version == 2 || last_begin < (1ull << 31)) break;
        version = 2;  // wider begins: lay the header out again
    }
    auto at = [&](const char *n) { return var(n) ? var(n)->begin - rec_begin : 0; };
    off_time = at("time");
    off_th = at("th");
    off_u = at("u");
    off_v = at("v");
    off_w = at("w");
    off_p = at("p");
    off_div = at("div");
    off_q = moisture ? at("q") : recsize;
    off_precip = moisture ? at("precip") : recsize;
    return true;
}

const NcVar *NcLayout::var(const std::string &name) const {
    for (const NcVar &v : vars)
        if (v.name == name) return &v;
    return nullptr;
}

std::vector<unsigned char> NcLayout::header(uint32_t numrecs) const {
    Out o;
    const unsigned char magic[4] = {'C', 'D', 'F', (unsigned char)version};
    o.b.insert(o.b.end(), magic, magic + 4);
    o.u32(numrecs);
    o.u32(kNcDimension);
    o.u32((uint32_t)dims.size());
    for (const auto &d : dims) {
        o.name(d.first);
        o.u32((uint32_t)d.second);
    }
    o.u32(0);  // no global attributes
    o.u32(0);
    o.u32(kNcVariable);
    o.u32((uint32_t)vars.size());
    for (const NcVar &v : vars) {
        o.name(v.name);
        o.u32((uint32_t)v.dims.size());
        for (int d : v.dims) o.u32((uint32_t)d);
        if (v.units.empty()) {
            o.u32(0);
            o.u32(0);
        } else {
            o.u32(kNcAttribute);
            o.u32(1);
            o.name("units");
            o.u32(NC_CHAR_T);
            o.u32((uint32_t)v.units.size());
            o.bytes(v.units.data(), v.units.size());
        }
        o.u32((uint32_t)v.type);
        o.u32((uint32_t)v.vsize);
        if (version == 1)
            o.u32((uint32_t)v.begin);
        else
            o.u64(v.begin);
    }
    return o.b;
}

bool NcFile::create(const std::string &p, const NcLayout &layout, const double *x,
                    const double *y, const double *z, const double *rhoa, const double *thetan,
                    const double *trefn, const char *q_names, std::string *why) {
    L = layout;
    path = p;
    fd = ::open(p.c_str(), O_CREAT | O_TRUNC | O_WRONLY, 0644);
    if (fd < 0) {
        if (why) *why = "cannot create " + p + ": " + std::strerror(errno);
        return false;
    }
    numrecs = 0;
    const std::vector<unsigned char> h = L.header(0);
    if (!write_all(fd, h.data(), h.size(), 0, why)) return false;
    auto put1d = [&](const char *name, const double *a, int n) {
        const NcVar *v = L.var(name);
        std::vector<uint32_t> buf(n);
        for (int t = 0; t < n; ++t) buf[t] = float_be(a[t + 1]);  // a(1:n)
        return write_all(fd, buf.data(), 4 * (size_t)n, v->begin, why);
    };
    if (!put1d("x", x, L.ip) || !put1d("y", y, L.jp) || !put1d("z", z, L.kp) ||
        !put1d("rhoa", rhoa, L.kp) || !put1d("thetan", thetan, L.kp) ||
        !put1d("trefn", trefn, L.kp))
        return false;
    if (L.moisture) {
        const NcVar *v = L.var("q_names");
        std::vector<char> names(v->vsize, 0);
        if (q_names)
            std::memcpy(names.data(), q_names, 20 * (size_t)L.nq);
        else
            std::memset(names.data(), ' ', 20 * (size_t)L.nq);
        if (!write_all(fd, names.data(), names.size(), v->begin, why)) return false;
    }
    return true;
}

bool NcFile::put_record(uint32_t rec, const unsigned char *block, std::string *why) {
    if (fd < 0) {
        if (why) *why = "output file is not open";
        return false;
    }
    if (!write_all(fd, block, L.recsize, L.rec_begin + (uint64_t)rec * L.recsize, why))
        return false;
    if (rec + 1 > numrecs) {
        numrecs = rec + 1;
        const uint32_t be = to_be32(numrecs);
        if (!write_all(fd, &be, 4, 4, why)) return false;
    }
    return true;
}

bool NcFile::close(std::string *why) {
    if (fd < 0) return true;
    const int r = ::close(fd);
    fd = -1;
    if (r != 0) {
        if (why) *why = std::string("close failed: ") + std::strerror(errno);
        return false;
    }
    return true;
}

NcFile::~NcFile() {
    if (fd >= 0) ::close(fd);
}

void pack_precip_host(const NcLayout &L, const double *precip, unsigned char *block) {
    if (!L.moisture || L.nprec == 0) return;
    uint32_t *o = (uint32_t *)(block + L.off_precip);
    const int ip = L.ip, jp = L.jp, kp = L.kp, np = L.nprec;
    const size_t slab = (size_t)kp * (jp + 2) * (ip + 2);  // one precip field
    for (int i = 0; i < ip; ++i)
        for (int j = 0; j < jp; ++j)
            for (int k = 0; k < kp; ++k)
                for (int n = 0; n < np; ++n) {
                    const size_t e = (((size_t)i * jp + j) * kp + k) * np + n;
                    o[e] = precip ? float_be(precip[n * slab +
                                                    ((size_t)(i + 1) * (jp + 2) + (j + 1)) * kp + k])
                                  : to_be32(kNcFillFloatBits);
                }
}

void pack_record_host(const NcLayout &L, double time, const double *u, const double *v,
                      const double *w, const double *th, const double *p, const double *div,
                      const double *q, const double *precip, unsigned char *block) {
    const int ip = L.ip, jp = L.jp, kp = L.kp;
    const size_t nk = kp + 2, nj = jp + 2, ntot = nk * nj * (ip + 2);
    const uint32_t t = float_be(time);
    std::memcpy(block + L.off_time, &t, 4);
    const double *f[6] = {th, u, v, w, p, div};
    const uint64_t off[6] = {L.off_th, L.off_u, L.off_v, L.off_w, L.off_p, L.off_div};
    for (int a = 0; a < 6; ++a) {
        uint32_t *o = (uint32_t *)(block + off[a]);
        for (int i = 0; i < ip; ++i)
            for (int j = 0; j < jp; ++j)
                for (int k = 0; k < kp; ++k)
                    o[((size_t)i * jp + j) * kp + k] =
                        float_be(f[a][((size_t)(i + 1) * nj + (j + 1)) * nk + k + 1]);
    }
    if (L.moisture) {
        uint32_t *o = (uint32_t *)(block + L.off_q);
        for (int i = 0; i < ip; ++i)
            for (int j = 0; j < jp; ++j)
                for (int k = 0; k < kp; ++k)
                    for (int n = 0; n < L.nq; ++n)
                        o[(((size_t)i * jp + j) * kp + k) * L.nq + n] = float_be(
                            q[n * ntot + ((size_t)(i + 1) * nj + (j + 1)) * nk + k + 1]);
        pack_precip_host(L, precip, block);
    }
}

}  // namespace dcm

struct dcm_ncfile {
    dcm::NcFile f;
    std::vector<unsigned char> block;
};

extern "C" {

int dcm_nc_create(const char *path, int ip, int jp, int kp, int moisture, int nq, int nprec,
                  const char *q_names, const double *x, const double *y, const double *z,
                  const double *rhoa, const double *thetan, const double *trefn,
                  dcm_ncfile **out) {
    if (!path || !x || !y || !z || !rhoa || !thetan || !trefn || !out) return DCM_BAD_ARG;
    *out = nullptr;
    dcm::NcLayout L;
    std::string why;
    if (!L.build(ip, jp, kp, moisture != 0, nq, nprec, &why)) {
        dcm::set_error("dcm_nc_create: " + why);
        return DCM_BAD_ARG;
    }
    dcm_ncfile *h = new dcm_ncfile;
    if (!h->f.create(path, L, x, y, z, rhoa, thetan, trefn, q_names, &why)) {
        delete h;
        dcm::set_error("dcm_nc_create: " + why);
        return DCM_IO_ERROR;
    }
    *out = h;
    return DCM_OK;
}

int dcm_nc_put_record(dcm_ncfile *h, int n, double time, const double *u, const double *v,
                      const double *w, const double *th, const double *p, const double *div,
                      const double *q, const double *precip) {
    if (!h || n < 1 || !u || !v || !w || !th || !p || !div || (h->f.L.moisture && !q))
        return DCM_BAD_ARG;
    h->block.assign(h->f.L.recsize, 0);
    dcm::pack_record_host(h->f.L, time, u, v, w, th, p, div, q, precip, h->block.data());
    std::string why;
    if (!h->f.put_record((uint32_t)(n - 1), h->block.data(), &why)) {
        dcm::set_error("dcm_nc_put_record: " + why);
        return DCM_IO_ERROR;
    }
    return DCM_OK;
}

int dcm_nc_close(dcm_ncfile *h) {
    if (!h) return DCM_BAD_ARG;
    std::string why;
    const bool ok = h->f.close(&why);
    delete h;
    if (!ok) {
        dcm::set_error("dcm_nc_close: " + why);
        return DCM_IO_ERROR;
    }
    return DCM_OK;
}

}  // extern "C"
