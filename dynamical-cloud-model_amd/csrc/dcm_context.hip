// dcm_context.hip -- context, device memory, host mirrors and profiling.
//
// The context holds what the reference keeps as module or driver state: the
// grid metrics every routine receives (initialisation.f90:429-446), the
// Laplacian coefficients that set_mat_a stores in eight 3-D arrays
// (dynamics.f90:15-17, 34-175) -- kept here as six 1-D arrays, since on a
// product grid each coefficient depends on one index only -- and the
// bicgstab_first flag (dynamics.f90:16).
#include "dcm_internal.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <cmath>

namespace dcm {

// The reference preconditions with s0 = r/a_p and three Jacobi sweeps: the
// residual polynomial (1 - lambda)^4 on the spectrum of D^-1 A, which lies
// in [0, 2] (Gershgorin: a_p is minus the sum of the off-diagonal
// coefficients, Neumann rows included).  With d divisions by a_p, the
// relaxation factors 1/l_i, l_i the roots of the degree-d Chebyshev
// polynomial shifted to [a, 2], make the residual polynomial
// T_d((2 + a - 2 lambda)/(2 - a)) / T_d((2 + a)/(2 - a)) instead: it stays
// below 1 in magnitude on [a, 2], so the preconditioned operator's
// eigenvalues there are bounded away from 0, and near lambda = 0 it grows
// like d^2 lambda rather than d lambda.  The factors are applied in the
// order whose partial products grow least over [0, 2] (so FP32 rounding is
// not amplified).
constexpr double kChebyshevA = 0.002;
constexpr int kChebyshevDivisions = 12;

void set_chebyshev(dcm_ctx *c, double a, int d) {
    c->cheb_a = a;
    c->cheb_on = a > 0.0;
    c->cheb_deg = c->cheb_on ? d : 4;
    for (int i = 0; i < 16; ++i) c->cheb_w[i] = 1.0;
    if (!c->cheb_on) return;
    const double b = 2.0, pi = 3.14159265358979323846;
    double w[16];
    for (int i = 0; i < d; ++i)
        w[i] = 1.0 / (0.5 * (b + a) - 0.5 * (b - a) * std::cos((2 * i + 1) * pi / (2.0 * d)));
    // greedy order: each next factor keeps the running product smallest in
    // magnitude over [0, 2] (for these roots it never exceeds 1)
    constexpr int kGrid = 400;
    double prod[kGrid + 1];
    for (int n = 0; n <= kGrid; ++n) prod[n] = 1.0;
    bool used[16] = {};
    for (int pos = 0; pos < d; ++pos) {
        int best = -1;
        double best_growth = 1e300;
        for (int i = 0; i < d; ++i) {
            if (used[i]) continue;
            double g = 0.0;
            for (int n = 0; n <= kGrid; ++n)
                g = std::max(g, std::fabs(prod[n] * (1.0 - w[i] * (b * n / kGrid))));
            if (g < best_growth) {
                best_growth = g;
                best = i;
            }
        }
        used[best] = true;
        for (int n = 0; n <= kGrid; ++n) prod[n] *= 1.0 - w[best] * (b * n / kGrid);
        c->cheb_w[pos] = w[best];
    }
}

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

// Scratch is allocated on first use.  An allocation (and the zeroing
// memset that follows it) must never land inside a stream capture, where it
// would either fail the capture or record a memset node into the graph: the
// caller has to run the work once without capture first.  Returns an error
// when c->stream is capturing.
int no_alloc_in_capture(dcm_ctx *c, const char *what) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
        set_error(std::string(what) + ": scratch would be allocated inside a stream capture "
                                      "(run the step once without capture first)");
        return DCM_HIP_ERROR;
    }
    return DCM_OK;
}

int work_field64(dcm_ctx *c, int slot, double **out) {
    if ((int)c->work64.size() <= slot) c->work64.resize(slot + 1, nullptr);
    if (!c->work64[slot]) {
        DCM_CHECK(no_alloc_in_capture(c, "work_field64"));
        void *p = nullptr;
        if (hipMalloc(&p, c->geo.ntot * sizeof(double)) != hipSuccess) {
            set_error("hipMalloc of a work field failed");
            return DCM_OUT_OF_MEMORY;
        }
        DCM_HIP(hipMemsetAsync(p, 0, c->geo.ntot * sizeof(double), c->stream));
        c->work64[slot] = (double *)p;
    }
    *out = c->work64[slot];
    return DCM_OK;
}

int work_field32(dcm_ctx *c, int slot, float **out) {
    if ((int)c->work32.size() <= slot) c->work32.resize(slot + 1, nullptr);
    if (!c->work32[slot]) {
        DCM_CHECK(no_alloc_in_capture(c, "work_field32"));
        void *p = nullptr;
        if (hipMalloc(&p, c->geo.ntot * sizeof(float)) != hipSuccess) {
            set_error("hipMalloc of a work field failed");
            return DCM_OUT_OF_MEMORY;
        }
        DCM_HIP(hipMemsetAsync(p, 0, c->geo.ntot * sizeof(float), c->stream));
        c->work32[slot] = (float *)p;
    }
    *out = c->work32[slot];
    return DCM_OK;
}

int ensure_partials_adv(dcm_ctx *c, size_t n) {
    if (c->partials_adv_cap >= n) return DCM_OK;
    DCM_CHECK(no_alloc_in_capture(c, "ensure_partials_adv"));
    if (c->partials_adv) {
        DCM_HIP(hipStreamSynchronize(c->stream));
        DCM_HIP(hipFree(c->partials_adv));
        c->partials_adv = nullptr;
    }
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc of reduction partials failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->partials_adv = (double *)p;
    c->partials_adv_cap = n;
    return DCM_OK;
}

int ensure_partials(dcm_ctx *c, size_t n) {
    if (c->partials_cap >= n) return DCM_OK;
    DCM_CHECK(no_alloc_in_capture(c, "ensure_partials"));
    if (c->partials) {
        DCM_HIP(hipStreamSynchronize(c->stream));
        DCM_HIP(hipFree(c->partials));
        c->partials = nullptr;
    }
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc of reduction partials failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->partials = (double *)p;
    c->partials_cap = n;
    return DCM_OK;
}

int ensure_solver_vectors64(dcm_ctx *c, size_t count) {
    if (c->sv64_count >= count) return DCM_OK;
    DCM_CHECK(no_alloc_in_capture(c, "ensure_solver_vectors64"));
    if (c->sv64) {
        DCM_HIP(hipStreamSynchronize(c->stream));
        DCM_HIP(hipFree(c->sv64));
        c->sv64 = nullptr;
    }
    const size_t n = (size_t)c->geo.ip * c->geo.jp * c->geo.kp;
    void *p = nullptr;
    if (hipMalloc(&p, count * n * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc of solver vectors failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->sv64 = (double *)p;
    c->sv64_count = count;
    return DCM_OK;
}

int ensure_host_flags(dcm_ctx *c, size_t n, int **out) {
    if (c->h_flags_cap < n) {
        if (c->h_flags) {
            DCM_HIP(hipStreamSynchronize(c->stream));
            DCM_HIP(hipHostFree(c->h_flags));
            c->h_flags = nullptr;
            c->h_flags_cap = 0;
        }
        void *p = nullptr;
        DCM_HIP(hipHostMalloc(&p, n * sizeof(int)));
        c->h_flags = (int *)p;
        c->h_flags_cap = n;
    }
    *out = c->h_flags;
    return DCM_OK;
}

int ensure_solver_vectors32(dcm_ctx *c, size_t count) {
    if (c->sv32_count >= count) return DCM_OK;
    DCM_CHECK(no_alloc_in_capture(c, "ensure_solver_vectors32"));
    if (c->sv32) {
        DCM_HIP(hipStreamSynchronize(c->stream));
        DCM_HIP(hipFree(c->sv32));
        c->sv32 = nullptr;
    }
    const size_t n = (size_t)c->geo.ip * c->geo.jp * c->geo.kp;
    void *p = nullptr;
    if (hipMalloc(&p, count * n * sizeof(float)) != hipSuccess) {
        set_error("hipMalloc of solver vectors failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->sv32 = (float *)p;
    c->sv32_count = count;
    return DCM_OK;
}

// Upload a short host array into slot `slot` of the 1-D scratch area.
int upload_1d(dcm_ctx *c, const double *h, size_t n, int slot, const double **out) {
    const size_t stride = (size_t)c->geo.kp + c->geo.jp + c->geo.ip + 16;
    const size_t need = stride * (size_t)(slot + 1);
    if (n > stride) {
        set_error("1-D array longer than the grid extent");
        return DCM_BAD_ARG;
    }
    DCM_CHECK(no_alloc_in_capture(c, "upload_1d"));
    if (c->d1d_cap < need) {
        // every slot is allocated at the first upload, so earlier slots of
        // the same call stay valid
        const size_t cap = stride * 16;
        if (slot >= 16) {
            set_error("1-D scratch slot out of range");
            return DCM_BAD_ARG;
        }
        if (c->d1d) {
            DCM_HIP(hipStreamSynchronize(c->stream));
            DCM_HIP(hipFree(c->d1d));
        }
        void *p = nullptr;
        DCM_HIP(hipMalloc(&p, cap * sizeof(double)));
        c->d1d = (double *)p;
        c->d1d_cap = cap;
    }
    double *d = c->d1d + stride * slot;
    // synchronous: the host array may be reused by the caller right away
    DCM_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    *out = d;
    return DCM_OK;
}

// the profile filter is a comma-separated list of kernel names (empty: all)
static bool prof_wanted(const dcm_ctx *c, const char *name) {
    if (!c->profiling) return false;
    if (c->prof_only.empty()) return true;
    const size_t n = strlen(name);
    size_t pos = 0;
    while (pos <= c->prof_only.size()) {
        size_t end = c->prof_only.find(',', pos);
        if (end == std::string::npos) end = c->prof_only.size();
        if (end - pos == n && c->prof_only.compare(pos, n, name) == 0) return true;
        pos = end + 1;
    }
    return false;
}

int prof_begin(dcm_ctx *c, const char *name) {
    if (!prof_wanted(c, name)) return -1;
    hipEvent_t a, b;
    if (c->event_pool.size() >= 2) {
        a = c->event_pool.back();
        c->event_pool.pop_back();
        b = c->event_pool.back();
        c->event_pool.pop_back();
    } else {
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1;
    }
    (void)hipEventRecord(a, c->stream);
    c->pending.push_back({name, {a, b}});
    return (int)c->pending.size() - 1;
}

bool prof_events(dcm_ctx *c, const char *name, hipEvent_t *start, hipEvent_t *stop) {
    if (!prof_wanted(c, name)) return false;
    hipEvent_t a, b;
    if (c->event_pool.size() >= 2) {
        a = c->event_pool.back();
        c->event_pool.pop_back();
        b = c->event_pool.back();
        c->event_pool.pop_back();
    } else {
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return false;
    }
    c->pending.push_back({name, {a, b}});
    *start = a;
    *stop = b;
    return true;
}

void prof_end(dcm_ctx *c, int token) {
    if (token < 0) return;
    (void)hipEventRecord(c->pending[token].second.second, c->stream);
}

int prof_collect(dcm_ctx *c) {
    if (c->pending.empty()) return DCM_OK;
    DCM_HIP(hipStreamSynchronize(c->stream));
    for (auto &p : c->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.second.first, p.second.second);
        auto &t = c->timers[p.first];
        t.total_ms += ms;
        t.launches += 1;
        c->event_pool.push_back(p.second.first);
        c->event_pool.push_back(p.second.second);
    }
    c->pending.clear();
    return DCM_OK;
}

}  // namespace dcm

using namespace dcm;

extern "C" {

const char *dcm_last_error(void) { return g_last_error.c_str(); }

const char *dcm_version(void) { return "dcm_hip 1.0 gfx950"; }

int dcm_device_memory(size_t *free_bytes, size_t *total_bytes) {
    if (!free_bytes || !total_bytes) return DCM_BAD_ARG;
    DCM_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return DCM_OK;
}

int dcm_context_create(int ip, int jp, int kp, const double *dx, const double *dy,
                       const double *dz, const double *dxn, const double *dyn,
                       const double *dzn, const double *rhoa, const double *rhoan,
                       dcm_ctx **out) {
    if (!out || ip < 1 || jp < 1 || kp < 1 || !dx || !dy || !dz || !dxn || !dyn || !dzn ||
        !rhoa || !rhoan) {
        set_error("dcm_context_create: bad arguments");
        return DCM_BAD_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device");
        return DCM_NO_DEVICE;
    }
    dcm_ctx *c = new dcm_ctx();
    Geo &g = c->geo;
    g.ip = ip; g.jp = jp; g.kp = kp;
    g.nk = kp + 2; g.nj = jp + 2; g.ni = ip + 2;
    g.sj = g.nk;
    g.si = (long long)g.nk * g.nj;
    g.ntot = (long long)g.nk * g.nj * g.ni;
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed");
        return DCM_HIP_ERROR;
    }
    auto cp = [](const double *p, int n) { return std::vector<double>(p, p + n + 2); };
    c->h_dx = cp(dx, ip); c->h_dy = cp(dy, jp); c->h_dz = cp(dz, kp);
    c->h_dxn = cp(dxn, ip); c->h_dyn = cp(dyn, jp); c->h_dzn = cp(dzn, kp);
    c->h_rhoa = cp(rhoa, kp); c->h_rhoan = cp(rhoan, kp);
    // Laplacian coefficients: set_mat_a, dynamics.f90:84-96.  a_e/a_w depend
    // on i only, a_n/a_s on j only, a_u/a_d on k only (the reference stores
    // them as full 3-D arrays).  Same IEEE expressions, so the same bits.
    std::vector<double> ae(ip + 2, 0.0), aw(ip + 2, 0.0), an(jp + 2, 0.0), as(jp + 2, 0.0),
        au(kp + 2, 0.0), ad(kp + 2, 0.0);
    for (int i = 1; i <= ip; ++i) {
        ae[i] = 1.0 / (dx[i - 1] * dxn[i]);
        aw[i] = 1.0 / (dx[i - 1] * dxn[i - 1]);
    }
    for (int j = 1; j <= jp; ++j) {
        an[j] = 1.0 / (dy[j - 1] * dyn[j]);
        as[j] = 1.0 / (dy[j - 1] * dyn[j - 1]);
    }
    for (int k = 1; k <= kp; ++k) {
        au[k] = 1.0 / (dz[k - 1] * dzn[k]);
        ad[k] = 1.0 / (dz[k - 1] * dzn[k - 1]);
    }
    const std::vector<double> *arrs[14] = {&c->h_dx, &c->h_dy, &c->h_dz, &c->h_dxn, &c->h_dyn,
                                           &c->h_dzn, &c->h_rhoa, &c->h_rhoan, &ae, &aw,
                                           &an, &as, &au, &ad};
    size_t total = 0;
    size_t offs[14];
    for (int a = 0; a < 14; ++a) {
        offs[a] = total;
        total += arrs[a]->size();
    }
    std::vector<double> h(total);
    std::vector<float> h32(total);
    for (int a = 0; a < 14; ++a)
        for (size_t q = 0; q < arrs[a]->size(); ++q) {
            h[offs[a] + q] = (*arrs[a])[q];
            h32[offs[a] + q] = (float)(*arrs[a])[q];
        }
    void *p = nullptr, *p32 = nullptr;
    if (hipMalloc(&p, total * sizeof(double)) != hipSuccess ||
        hipMalloc(&p32, total * sizeof(float)) != hipSuccess) {
        set_error("hipMalloc of metrics failed");
        delete c;
        return DCM_OUT_OF_MEMORY;
    }
    c->dmet = (double *)p;
    c->dmet32 = (float *)p32;
    (void)hipMemcpy(c->dmet, h.data(), total * sizeof(double), hipMemcpyHostToDevice);
    (void)hipMemcpy(c->dmet32, h32.data(), total * sizeof(float), hipMemcpyHostToDevice);
    const double **m64[14] = {&c->met.dx, &c->met.dy, &c->met.dz, &c->met.dxn, &c->met.dyn,
                              &c->met.dzn, &c->met.rhoa, &c->met.rhoan, &c->met.ae,
                              &c->met.aw, &c->met.an, &c->met.as, &c->met.au, &c->met.ad};
    const float **m32[14] = {&c->met32.dx, &c->met32.dy, &c->met32.dz, &c->met32.dxn,
                             &c->met32.dyn, &c->met32.dzn, &c->met32.rhoa, &c->met32.rhoan,
                             &c->met32.ae, &c->met32.aw, &c->met32.an, &c->met32.as,
                             &c->met32.au, &c->met32.ad};
    for (int a = 0; a < 14; ++a) {
        *m64[a] = c->dmet + offs[a];
        *m32[a] = c->dmet32 + offs[a];
    }
    // FP32 tables of the mixed path's MPDATA: each entry one FP64 quotient
    // rounded to FP32
    {
        const int ni = ip + 2, nj = jp + 2, nk = kp + 2;
        std::vector<float> t(4 * (size_t)ni + 4 * (size_t)nj + 7 * (size_t)nk, 0.f);
        float *x = t.data(), *y = x + 4 * ni, *z = y + 4 * nj;
        for (int i = 0; i < ni; ++i) {
            x[4 * i] = (float)dx[i];
            x[4 * i + 1] = (float)(1.0 / dxn[i]);
            x[4 * i + 2] = i > 0 ? (float)(1.0 / dx[i - 1]) : 0.f;
            x[4 * i + 3] = i > 0 ? (float)(2.0 / (dxn[i - 1] + dxn[i])) : 0.f;
        }
        for (int j = 0; j < nj; ++j) {
            y[j] = (float)dy[j];
            y[nj + j] = (float)(1.0 / dy[j]);
            y[2 * nj + j] = (float)(1.0 / dyn[j]);
            y[3 * nj + j] = j > 0 ? (float)(2.0 / (dyn[j - 1] + dyn[j])) : 0.f;
        }
        for (int k = 0; k < nk; ++k) {
            z[k] = (float)dz[k];
            z[nk + k] = (float)(1.0 / dz[k]);
            z[2 * nk + k] = (float)(1.0 / dzn[k]);
            z[3 * nk + k] = k > 0 ? (float)(2.0 / (dzn[k - 1] + dzn[k])) : 0.f;
            z[4 * nk + k] = (float)rhoa[k];
            z[5 * nk + k] = (float)rhoan[k];
            z[6 * nk + k] = (float)(1.0 / rhoan[k]);
        }
        void *pt = nullptr;
        if (hipMalloc(&pt, t.size() * sizeof(float)) != hipSuccess) {
            set_error("hipMalloc of the FP32 tables failed");
            return DCM_OUT_OF_MEMORY;
        }
        c->dtab32 = (float *)pt;
        (void)hipMemcpy(c->dtab32, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice);
        const float *dy32 = c->dtab32 + 4 * ni, *dz32 = dy32 + 4 * nj;
        c->tab32 = Tab32{(const float4 *)c->dtab32,
                         dy32,          dy32 + nj,     dy32 + 2 * nj, dy32 + 3 * nj,
                         dz32,          dz32 + nk,     dz32 + 2 * nk, dz32 + 3 * nk,
                         dz32 + 4 * nk, dz32 + 5 * nk, dz32 + 6 * nk};
    }
    if (hipMalloc(&p, 64 * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->dscal = (double *)p;
    (void)hipMemset(c->dscal, 0, 64 * sizeof(double));
    if (hipHostMalloc(&p, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->hscal = (double *)p;
    if (hipMalloc(&p, 2 * sizeof(SolverScalars)) != hipSuccess) {
        set_error("hipMalloc failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->solver = (SolverScalars *)p;
    (void)hipMemset(c->solver, 0, 2 * sizeof(SolverScalars));
    if (hipHostMalloc(&p, 2 * sizeof(SolverScalars), hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->h_solver = (SolverScalars *)p;
    if (hipMalloc(&p, sizeof(OuterState)) != hipSuccess) {
        set_error("hipMalloc failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->outer_d = (OuterState *)p;
    if (hipHostMalloc(&p, sizeof(OuterState), hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->outer_h = (OuterState *)p;
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("device initialisation failed");
        return DCM_HIP_ERROR;
    }
    // DCM_SUM_ORDER=sequential selects the serial summation order for callers
    // that cannot reach dcm_set_sum_order (the Fortran drop-in modules)
    const char *order = std::getenv("DCM_SUM_ORDER");
    c->seq_sums = order && std::strcmp(order, "sequential") == 0;
    // the mixed inner solve's preconditioner: Chebyshev factors on
    // [0.002, 2] with 12 divisions by a_p (measured on the 256x256x128 moist
    // bubble: 36 inner iterations per step against 142 with the reference's
    // four plain ones, the step 26.5 ms against 43.4; 8..16 divisions and
    // a = 0.001..0.005 perform within 8 %)
    set_chebyshev(c, kChebyshevA, kChebyshevDivisions);
    *out = c;
    return DCM_OK;
}

int dcm_context_destroy(dcm_ctx *c) {
    if (!c) return DCM_BAD_ARG;
    (void)hipStreamSynchronize(c->stream);
    mg_release(c);
    for (auto &m : c->mirrors) (void)hipFree(m.second.dev);
    for (double *p : c->work64) (void)hipFree(p);
    for (float *p : c->work32) (void)hipFree(p);
    (void)hipFree(c->sv64);
    (void)hipFree(c->sv32);
    (void)hipFree(c->partials);
    (void)hipFree(c->hm_partials);
    (void)hipFree(c->d1d);
    (void)hipFree(c->dmet);
    (void)hipFree(c->dmet32);
    (void)hipFree(c->dtab32);
    (void)hipFree(c->mp32);
    (void)hipFree(c->partials_adv);
    (void)hipFree(c->dscal);
    (void)hipHostFree(c->hscal);
    (void)hipFree(c->solver);
    (void)hipHostFree(c->h_solver);
    (void)hipFree(c->outer_d);
    if (c->outer_h) (void)hipHostFree(c->outer_h);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->iter_graph) (void)hipGraphExecDestroy(c->iter_graph);
    if (c->chunk_graph) (void)hipGraphExecDestroy(c->chunk_graph);
    c->solve_tables.release();
    c->flags.release();
    if (c->h_flags) (void)hipHostFree(c->h_flags);
    if (!c->borrowed_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DCM_OK;
}

size_t dcm_field_elems(const dcm_ctx *c) { return c ? (size_t)c->geo.ntot : 0; }

int dcm_malloc(dcm_ctx *c, size_t bytes, void **dptr) {
    if (!c || !dptr) return DCM_BAD_ARG;
    if (hipMalloc(dptr, bytes ? bytes : 8) != hipSuccess) {
        set_error("hipMalloc failed");
        return DCM_OUT_OF_MEMORY;
    }
    return DCM_OK;
}

int dcm_free(dcm_ctx *c, void *dptr) {
    if (!c) return DCM_BAD_ARG;
    DCM_HIP(hipStreamSynchronize(c->stream));
    DCM_HIP(hipFree(dptr));
    return DCM_OK;
}

int dcm_memcpy_htod(dcm_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return DCM_BAD_ARG;
    DCM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    return DCM_OK;
}

int dcm_memcpy_dtoh(dcm_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return DCM_BAD_ARG;
    DCM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    return DCM_OK;
}

int dcm_synchronize(dcm_ctx *c) {
    if (!c) return DCM_BAD_ARG;
    DCM_HIP(hipStreamSynchronize(c->stream));
    return DCM_OK;
}

int dcm_mirror(dcm_ctx *c, const void *host, size_t bytes, void **dptr) {
    if (!c || !host || !dptr) return DCM_BAD_ARG;
    auto it = c->mirrors.find(host);
    if (it != c->mirrors.end()) {
        if (it->second.bytes != bytes) {
            set_error("dcm_mirror: host array re-registered with another size");
            return DCM_BAD_ARG;
        }
        *dptr = it->second.dev;
        return DCM_OK;
    }
    void *d = nullptr;
    if (hipMalloc(&d, bytes ? bytes : 8) != hipSuccess) {
        set_error("hipMalloc of a mirror failed");
        return DCM_OUT_OF_MEMORY;
    }
    c->mirrors[host] = Mirror{d, bytes};
    *dptr = d;
    return DCM_OK;
}

int dcm_mirror_upload(dcm_ctx *c, const void *host) {
    if (!c) return DCM_BAD_ARG;
    auto it = c->mirrors.find(host);
    if (it == c->mirrors.end()) {
        set_error("dcm_mirror_upload: unknown host array");
        return DCM_BAD_ARG;
    }
    DCM_HIP(hipMemcpyAsync(it->second.dev, host, it->second.bytes, hipMemcpyHostToDevice,
                           c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    return DCM_OK;
}

int dcm_mirror_download(dcm_ctx *c, void *host) {
    if (!c) return DCM_BAD_ARG;
    auto it = c->mirrors.find(host);
    if (it == c->mirrors.end()) {
        set_error("dcm_mirror_download: unknown host array");
        return DCM_BAD_ARG;
    }
    DCM_HIP(hipMemcpyAsync(host, it->second.dev, it->second.bytes, hipMemcpyDeviceToHost,
                           c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    return DCM_OK;
}

int dcm_mirror_release_all(dcm_ctx *c) {
    if (!c) return DCM_BAD_ARG;
    DCM_HIP(hipStreamSynchronize(c->stream));
    for (auto &m : c->mirrors) (void)hipFree(m.second.dev);
    c->mirrors.clear();
    return DCM_OK;
}

}  // extern "C"
