// dcm_init_dev.hip -- the 3-D initial state of allocate_and_set
// (initialisation.f90:640-870) generated in HBM from the 1-D profiles of
// dcm_init_profiles: no host copy of a 3-D field is made or uploaded.
#include <vector>

#include "dcm_internal.h"

namespace dcm {

namespace {

// SplitMix64 finaliser (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): a counter-based uniform
// deviate per cell, reproducible for any launch shape
__device__ __forceinline__ double cell_uniform(unsigned long long seed, unsigned long long n) {
    unsigned long long z = seed + (n + 1ull) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * 0x1.0p-53;
}

struct InitArgs {
    double *u, *v, *w, *th, *q;
    int nq, q_on;
    const double *uprof, *vprof, *qprof, *z, *zn, *xn, *yn;
    int param_theta, bubble;
    double param_z, param_delz;
    unsigned long long seed;
};

__global__ __launch_bounds__(256) void k_init_fields(InitArgs a, Geo g) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ntot;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c / g.si);
        const long long r = c - (long long)i * g.si;
        const int j = (int)(r / g.sj);
        const int k = (int)(r - (long long)j * g.sj);
        // winds: the same column everywhere, halos included (:753-800)
        a.u[c] = a.uprof[k];
        a.v[c] = a.vprof[k];
        a.w[c] = 0.0;
        // th: perturbation from theta (:461); the param_theta noise in the
        // two layers around param_z (:706-731), on the cells the reference's
        // loop reaches (k, j, i <= kp, jp, ip), or the warm bubble (:804-827)
        double th = 0.0;
        if (a.param_theta && k <= g.kp && j <= g.jp && i <= g.ip) {
            const double zk = a.z[k];
            const double rr = cell_uniform(a.seed, (unsigned long long)c);
            if (zk > a.param_z && zk <= a.param_z + a.param_delz) th = -0.001 + (rr - 0.5) / 20.0;
            if (zk > a.param_z - a.param_delz && zk <= a.param_z) th = 0.001 - (rr - 0.5) / 20.0;
        }
        if (a.bubble) {
            th = 0.0;
            double rad = (a.zn[k] - 350.0) * (a.zn[k] - 350.0);
            if (g.ip > 1) rad = rad + a.xn[i] * a.xn[i];
            if (g.jp > 1) rad = rad + a.yn[j] * a.yn[j];
            rad = sqrt(rad);
            if (rad <= 300.0) th = th + 1.0;
        }
        a.th[c] = th;
        // moisture: the same columns everywhere (:642-701)
        for (int n = 0; n < a.nq; ++n)
            a.q[(long long)n * g.ntot + c] = a.q_on ? a.qprof[(long long)n * g.nk + k] : 0.0;
    }
}

}  // namespace

}  // namespace dcm

using namespace dcm;

extern "C" {

int dcm_model_init_fields(dcm_model *m, const dcm_init_config *cfg,
                          const dcm_init_profiles *prof) {
    if (!m || !cfg || !prof || !prof->u || !prof->v || !prof->z || !prof->zn || !prof->xn ||
        !prof->yn)
        return DCM_BAD_ARG;
    ModelView v;
    DCM_CHECK(model_view(m, &v));
    const Geo &g = v.c->geo;
    if (cfg->ip != g.ip || cfg->jp != g.jp || cfg->kp != g.kp) {
        set_error("dcm_model_init_fields: the configuration's grid is not the model's");
        return DCM_BAD_ARG;
    }
    if (cfg->moisture && (cfg->nq != v.nq || !prof->q)) {
        set_error("dcm_model_init_fields: moisture needs the model's nq q fields and profiles");
        return DCM_BAD_ARG;
    }
    // one upload of the columns: u, v, z, zn (nk each), xn (ni), yn (nj), q
    const size_t nk = g.nk, ni = g.ni, nj = g.nj;
    const size_t nqk = cfg->moisture ? (size_t)v.nq * nk : 0;
    std::vector<double> h(4 * nk + ni + nj + nqk);
    double *p = h.data();
    auto put = [&](const double *src, size_t n) {
        std::copy(src, src + n, p);
        p += n;
    };
    put(prof->u, nk);
    put(prof->v, nk);
    put(prof->z, nk);
    put(prof->zn, nk);
    put(prof->xn, ni);
    put(prof->yn, nj);
    if (nqk) put(prof->q, nqk);
    double *d = nullptr;
    DCM_HIP(hipMalloc(&d, h.size() * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                                  v.c->stream);
    InitArgs a;
    a.u = const_cast<double *>(v.u);
    a.v = const_cast<double *>(v.v);
    a.w = const_cast<double *>(v.w);
    a.th = const_cast<double *>(v.th);
    a.q = const_cast<double *>(v.q);
    a.nq = v.nq;
    a.q_on = cfg->moisture ? 1 : 0;
    a.uprof = d;
    a.vprof = d + nk;
    a.z = d + 2 * nk;
    a.zn = d + 3 * nk;
    a.xn = d + 4 * nk;
    a.yn = d + 4 * nk + ni;
    a.qprof = nqk ? d + 4 * nk + ni + nj : nullptr;
    a.param_theta = cfg->param_theta;
    a.bubble = cfg->bubble;
    a.param_z = cfg->param_z;
    a.param_delz = cfg->param_delz;
    a.seed = cfg->seed;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_init_fields, dim3(2048), dim3(kBlock), 0, v.c->stream, a, g);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(v.c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) {
        set_error(std::string("dcm_model_init_fields: ") + hipGetErrorString(e));
        return DCM_HIP_ERROR;
    }
    // halos and the previous time levels (:861-870, driver_code.f90:287-295)
    return dcm_model_init(m);
}

}  // extern "C"
