// dcm_output.hip -- the model's output file written from the GPU
// (driver_code.f90:725-1145, `output`), asynchronously.
//
// A NetCDF classic record is one contiguous block: time, th, u, v, w, p,
// div, q, precip, each in the file's C order and big-endian.  The pack
// kernels build that block in HBM straight from the padded FP64 fields
// (FP64 -> float conversion, byte swap, halo removal and, for q, the
// interleaving of the fields, in one pass), so the snapshot costs one read
// of the fields on the model's stream.  The block then goes to pinned host
// memory on a copy stream of its own, overlapping the following steps, and
// a host thread writes it to the file.  Two record buffers are in flight.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

#include "dcm_internal.h"
#include "dcm_netcdf.h"

namespace dcm {

namespace {

__device__ __forceinline__ unsigned be_float(double x) {
    return __builtin_bswap32(__float_as_uint((float)x));
}

// the six single fields th, u, v, w, p, div: one thread per interior cell of
// an i-plane, blockIdx.z selects the field
struct PackSrc {
    const double *f[6];
    unsigned long long off[6];  // byte offset of the field inside the record
};

__global__ __launch_bounds__(256) void k_pack_fields(PackSrc s, unsigned char *rec, Geo g) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int plane = g.jp * g.kp;
    if (t >= plane) return;
    const int i = blockIdx.y, a = blockIdx.z;
    const int j = t / g.kp, k = t - j * g.kp;
    const double x = s.f[a][pidx(g, k + 1, j + 1, i + 1)];
    unsigned *out = (unsigned *)(rec + s.off[a]);
    out[(long long)i * plane + t] = be_float(x);
}

// nf fields interleaved with the field index fastest (q, :1088-1093)
__global__ __launch_bounds__(256) void k_pack_interleaved(const double *q, int nf,
                                                          unsigned char *rec,
                                                          unsigned long long off, Geo g) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = g.kp * nf, plane = g.jp * col;
    if (t >= plane) return;
    const int i = blockIdx.y;
    const int j = t / col, r = t - j * col;
    const int k = r / nf, n = r - k * nf;
    const double x = q[(long long)n * g.ntot + pidx(g, k + 1, j + 1, i + 1)];
    unsigned *out = (unsigned *)(rec + off);
    out[(long long)i * plane + t] = be_float(x);
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

}  // namespace dcm

using namespace dcm;

struct dcm_output {
    struct Slot {
        unsigned char *dev = nullptr, *host = nullptr;
        hipEvent_t packed = nullptr, copied = nullptr;
        bool busy = false;
        bool precip_fill = false;  // the precip section holds the fill value
    };
    struct Job {
        int slot;
        uint32_t rec;
        double time;
    };
    dcm_model *m = nullptr;
    dcm_ctx *c = nullptr;
    int device = 0;
    NcFile file;
    hipStream_t copy = nullptr;
    Slot slot[2];
    std::thread writer;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> queue;
    bool stop = false;
    int pending = 0;
    std::string err;  // the writer's first failure
    uint32_t next_rec = 0;
    long long written = 0;
    double stall_s = 0.0, write_s = 0.0;
};

namespace {

void writer_main(dcm_output *o) {
    (void)hipSetDevice(o->device);
    for (;;) {
        dcm_output::Job j;
        {
            std::unique_lock<std::mutex> lk(o->mu);
            o->cv.wait(lk, [&] { return o->stop || !o->queue.empty(); });
            if (o->queue.empty()) return;  // stopped and drained
            j = o->queue.front();
            o->queue.pop_front();
        }
        dcm_output::Slot &s = o->slot[j.slot];
        std::string why;
        const hipError_t e = hipEventSynchronize(s.copied);
        bool ok = e == hipSuccess;
        if (!ok) why = std::string("record copy: ") + hipGetErrorString(e);
        const auto t0 = std::chrono::steady_clock::now();
        if (ok) {
            const uint32_t t = float_be(j.time);
            std::memcpy(s.host + o->file.L.off_time, &t, 4);
            ok = o->file.put_record(j.rec, s.host, &why);
        }
        const double dt = seconds_since(t0);
        {
            std::lock_guard<std::mutex> lk(o->mu);
            if (!ok && o->err.empty()) o->err = why;
            if (ok) ++o->written;
            o->write_s += dt;
            s.busy = false;
            --o->pending;
        }
        o->cv.notify_all();
    }
}

void release(dcm_output *o) {
    if (o->writer.joinable()) {
        {
            std::lock_guard<std::mutex> lk(o->mu);
            o->stop = true;
        }
        o->cv.notify_all();
        o->writer.join();
    }
    if (o->copy) (void)hipStreamSynchronize(o->copy);
    for (dcm_output::Slot &s : o->slot) {
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        if (s.packed) (void)hipEventDestroy(s.packed);
        if (s.copied) (void)hipEventDestroy(s.copied);
    }
    if (o->copy) (void)hipStreamDestroy(o->copy);
    o->file.close(nullptr);
    delete o;
}

}  // namespace

extern "C" {

int dcm_output_create(dcm_model *m, const char *path, int moisture, int nprec,
                      const char *q_names, const double *x, const double *y, const double *z,
                      const double *rhoa, const double *thetan, const double *trefn,
                      dcm_output **out) {
    if (!m || !path || !x || !y || !z || !rhoa || !thetan || !trefn || !out) return DCM_BAD_ARG;
    *out = nullptr;
    ModelView v;
    DCM_CHECK(model_view(m, &v));
    const Geo &g = v.c->geo;
    if (g.ip > 65535) {
        set_error("dcm_output_create: ip above 65535 (one grid row of blocks per i-plane)");
        return DCM_BAD_ARG;
    }
    if (moisture && v.nq < 1) {
        set_error("dcm_output_create: moisture output needs q fields in the model");
        return DCM_BAD_ARG;
    }
    NcLayout L;
    std::string why;
    if (!L.build(g.ip, g.jp, g.kp, moisture != 0, v.nq, nprec, &why)) {
        set_error("dcm_output_create: " + why);
        return DCM_BAD_ARG;
    }
    if (L.off_time != 0 || L.off_th != 4) {
        set_error("dcm_output_create: unexpected record layout");
        return DCM_BAD_ARG;
    }
    DCM_CHECK(model_enable_div(m));
    dcm_output *o = new dcm_output;
    o->m = m;
    o->c = v.c;
    o->device = v.c->device;
    if (!o->file.create(path, L, x, y, z, rhoa, thetan, trefn, q_names, &why)) {
        release(o);
        set_error("dcm_output_create: " + why);
        return DCM_IO_ERROR;
    }
    auto fail = [&](const char *what, hipError_t e) {
        release(o);
        set_error(std::string("dcm_output_create: ") + what + ": " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? DCM_OUT_OF_MEMORY
                                                                          : DCM_HIP_ERROR;
    };
    hipError_t e = hipStreamCreateWithFlags(&o->copy, hipStreamNonBlocking);
    if (e != hipSuccess) return fail("copy stream", e);
    for (dcm_output::Slot &s : o->slot) {
        if ((e = hipMalloc(&s.dev, L.recsize)) != hipSuccess) return fail("record buffer", e);
        if ((e = hipHostMalloc(&s.host, L.recsize, hipHostMallocDefault)) != hipSuccess)
            return fail("pinned record buffer", e);
        if ((e = hipEventCreateWithFlags(&s.packed, hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&s.copied, hipEventDisableTiming)) != hipSuccess)
            return fail("events", e);
        if (L.moisture) {
            pack_precip_host(L, nullptr, s.host);
            s.precip_fill = true;
        }
    }
    o->writer = std::thread(writer_main, o);
    *out = o;
    return DCM_OK;
}

int dcm_output_write(dcm_output *o, double time, const double *precip) {
    if (!o) return DCM_BAD_ARG;
    int sidx;
    {
        std::unique_lock<std::mutex> lk(o->mu);
        const auto t0 = std::chrono::steady_clock::now();
        o->cv.wait(lk, [&] { return !o->err.empty() || !o->slot[0].busy || !o->slot[1].busy; });
        o->stall_s += seconds_since(t0);
        if (!o->err.empty()) {
            set_error("dcm_output_write: " + o->err);
            return DCM_IO_ERROR;
        }
        sidx = o->slot[0].busy ? 1 : 0;
        o->slot[sidx].busy = true;
    }
    dcm_output::Slot &s = o->slot[sidx];
    auto release_slot = [&](int st) {
        {
            std::lock_guard<std::mutex> lk(o->mu);
            s.busy = false;
        }
        o->cv.notify_all();
        return st;
    };
    ModelView v;
    int st = model_view(o->m, &v);
    if (st != DCM_OK) return release_slot(st);
    const Geo &g = o->c->geo;
    const NcLayout &L = o->file.L;
    PackSrc ps;
    const double *f[6] = {v.th, v.u, v.v, v.w, v.p, v.div};
    const uint64_t off[6] = {L.off_th, L.off_u, L.off_v, L.off_w, L.off_p, L.off_div};
    for (int a = 0; a < 6; ++a) {
        ps.f[a] = f[a];
        ps.off[a] = off[a];
    }
    const unsigned plane = (unsigned)g.jp * (unsigned)g.kp;
    hipLaunchKernelGGL(k_pack_fields, dim3(blocks_for(plane), g.ip, 6), dim3(kBlock), 0,
                       o->c->stream, ps, s.dev, g);
    if (L.moisture)
        hipLaunchKernelGGL(k_pack_interleaved, dim3(blocks_for((long long)plane * L.nq), g.ip),
                           dim3(kBlock), 0, o->c->stream, v.q, L.nq, s.dev,
                           (unsigned long long)L.off_q, g);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(s.packed, o->c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(o->copy, s.packed, 0);
    // the time (first 4 bytes) and the precip section are filled on the host
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.host + 4, s.dev + 4, L.off_precip - 4, hipMemcpyDeviceToHost,
                           o->copy);
    if (e == hipSuccess) e = hipEventRecord(s.copied, o->copy);
    if (e != hipSuccess) {
        set_error(std::string("dcm_output_write: ") + hipGetErrorString(e));
        return release_slot(DCM_HIP_ERROR);
    }
    // precip is the caller's host array, packed here; without one the
    // section keeps (or gets back) the fill value
    if (L.moisture && (precip || !s.precip_fill)) {
        pack_precip_host(L, precip, s.host);
        s.precip_fill = precip == nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(o->mu);
        o->queue.push_back({sidx, o->next_rec++, time});
        ++o->pending;
    }
    o->cv.notify_all();
    return DCM_OK;
}

int dcm_output_flush(dcm_output *o) {
    if (!o) return DCM_BAD_ARG;
    std::unique_lock<std::mutex> lk(o->mu);
    o->cv.wait(lk, [&] { return o->pending == 0; });
    if (!o->err.empty()) {
        set_error("dcm_output: " + o->err);
        return DCM_IO_ERROR;
    }
    return DCM_OK;
}

int dcm_output_stats(dcm_output *o, long long *records, double *stall_s, double *write_s) {
    if (!o) return DCM_BAD_ARG;
    std::lock_guard<std::mutex> lk(o->mu);
    if (records) *records = o->written;
    if (stall_s) *stall_s = o->stall_s;
    if (write_s) *write_s = o->write_s;
    return DCM_OK;
}

int dcm_output_destroy(dcm_output *o) {
    if (!o) return DCM_BAD_ARG;
    const int st = dcm_output_flush(o);
    release(o);
    return st;
}

}  // extern "C"
