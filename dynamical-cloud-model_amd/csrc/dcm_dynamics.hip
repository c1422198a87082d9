// dcm_dynamics.hip -- momentum sources, horizontal means, the momentum and
// scalar updates, and the model_driver time step on device-resident state.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <string>

#include "dcm_device.h"

namespace dcm {

static constexpr double kGrav = 9.81;  // dynamics.f90:18

// ---------------------------------------------------------------------------
// sources, dynamics.f90:743-797 (interior k = 1..kp)
// ---------------------------------------------------------------------------
// sources, dynamics.f90:743-797 (inviscid part).  A thread owns one (k,j)
// column and marches kSrcPlanes i planes: the values of plane i-1 and i+1
// that a cell needs are the previous and next planes' own values, kept in
// registers, so each plane costs 19 loads instead of the 26 a cell reads;
// offsets are 32-bit element offsets from the uniform bases.  The
// expressions and their operand order are the reference's.
constexpr int kSrcPlanes = 8;
struct SrcPlane {
    // u at (k,j), (k,j+1), (k,j-1), (k+1,j), (k-1,j)
    double u, un, us, uu, ud;
    // v at (k,j), (k,j-1), (k,j+1), (k+1,j), (k-1,j), (k+1,j-1)
    double v, vs, vn, vu, vd, vus;
    // w at (k,j), (k-1,j), (k+1,j), (k,j-1), (k,j+1), (k-1,j+1)
    double w, wd, wu, ws, wn, wdn;
    double th, thu;  // th at (k,j), (k+1,j)
};
__device__ __forceinline__ void src_load(SrcPlane &P, const double *__restrict__ u,
                                         const double *__restrict__ v,
                                         const double *__restrict__ w,
                                         const double *__restrict__ th, int c, int SJ) {
    P.u = u[c];
    P.un = u[c + SJ];
    P.us = u[c - SJ];
    P.uu = u[c + 1];
    P.ud = u[c - 1];
    P.v = v[c];
    P.vs = v[c - SJ];
    P.vn = v[c + SJ];
    P.vu = v[c + 1];
    P.vd = v[c - 1];
    P.vus = v[c + 1 - SJ];
    P.w = w[c];
    P.wd = w[c - 1];
    P.wu = w[c + 1];
    P.ws = w[c - SJ];
    P.wn = w[c + SJ];
    P.wdn = w[c - 1 + SJ];
    P.th = th[c];
    P.thu = th[c + 1];
}

__global__ __launch_bounds__(256) void k_sources(const double *__restrict__ u,
                                                 const double *__restrict__ v,
                                                 const double *__restrict__ w,
                                                 const double *__restrict__ th,
                                                 const double *__restrict__ theta,
                                                 const double *__restrict__ thbar,
                                                 double *__restrict__ su, double *__restrict__ sv,
                                                 double *__restrict__ sw, Geo g, Metrics M) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i0 = blockIdx.z * kSrcPlanes + 1;
    if (k > g.kp || j > g.jp) return;
    const int i1 = min(i0 + kSrcPlanes - 1, g.ip);
    const int SI = (int)g.si, SJ = (int)g.sj;
    const double *dx = M.dx, *dy = M.dy, *dz = M.dz, *dxn = M.dxn, *dyn = M.dyn, *dzn = M.dzn;
    // per-column constants: the j and k denominators and densities
    const double dY_u = 2.0 * (dyn[j] + dyn[j - 1]), dY_v = 2.0 * (dy[j] + dy[j - 1]);
    const double dZ_u = 2.0 * (dzn[k] + dzn[k - 1]), dZ_w = 2.0 * (dz[k] + dz[k - 1]);
    const double rn = M.rhoan[k], ra = M.rhoa[k];
    const double tden = ((2.0 * theta[k]) + thbar[k]) + thbar[k + 1];
    int c = i0 * SI + j * SJ + k;
    // plane i0-1: the values a cell takes from its west neighbour
    double u_w = u[c - SI], un_w = u[c - SI + SJ], uu_w = u[c - SI + 1];
    double v_w = v[c - SI], w_w = w[c - SI];
    SrcPlane P, N, NN;
    src_load(P, u, v, w, th, c, SJ);
    src_load(N, u, v, w, th, c + SI, SJ);  // plane i0+1 <= ip+1: inside the padded array
    for (int i = i0; i <= i1; ++i) {
        // the loads of plane i+2 in flight while plane i is computed
        if (i < i1) src_load(NN, u, v, w, th, c + 2 * SI, SJ);
        {  // su: :743-757
            const double X = (u_w * (P.u + u_w) - N.u * (P.u + N.u)) / (2.0 * (dx[i] + dx[i - 1]));
            const double Y = (P.us * (P.vs + N.vs) - P.un * (P.v + N.v)) / dY_u;
            const double Z = (P.ud * (P.wd + N.wd) - P.uu * (P.w + N.w)) / dZ_u;
            su[c] = rn * ((X + Y) + Z);
        }
        {  // sv: :761-775
            const double X = (v_w * (u_w + un_w) - N.v * (P.u + P.un)) / (2.0 * (dxn[i] + dxn[i - 1]));
            const double Y = (P.vs * (P.v + P.vs) - P.vn * (P.v + P.vn)) / dY_v;
            const double Z = (P.vd * (P.wd + P.wdn) - P.vu * (P.w + P.wn)) / dZ_u;
            sv[c] = rn * ((X + Y) + Z);
        }
        {  // sw with buoyancy: :779-797
            const double X = (w_w * (u_w + uu_w) - N.w * (P.u + P.uu)) / (2.0 * (dxn[i] + dxn[i - 1]));
            const double Y = (P.ws * (P.vs + P.vus) - P.wn * (P.v + P.vu)) / (2.0 * (dyn[j] + dyn[j - 1]));
            const double Z = (P.wd * (P.w + P.wd) - P.wu * (P.w + P.wu)) / dZ_w;
            const double B = (kGrav * (P.th + P.thu)) / tden;
            sw[c] = ra * (((X + Y) + Z) + B);
        }
        u_w = P.u;
        un_w = P.un;
        uu_w = P.uu;
        v_w = P.v;
        w_w = P.w;
        P = N;
        N = NN;
        c += SI;
    }
}

// floor and lid rows of the sources (dynamics.f90:922-973), whole planes
__global__ void k_sources_rows(double *su, double *sv, double *sw, Geo g) {
    const int P = g.nj * g.ni;  // (32-bit index arithmetic: a plane is far below 2^31)
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
        const int i = (int)((unsigned)t / (unsigned)g.nj), j = t - i * g.nj;
        const long long c0 = pidx(g, 0, j, i);
        su[c0] = -su[c0 + 1];
        sv[c0] = -sv[c0 + 1];
        sw[c0] = 0.0;
        const long long ck = pidx(g, g.kp, j, i);
        sw[ck] = 0.0;
        sw[ck + 1] = 0.0;
        su[ck] = 0.0;
        su[ck + 1] = 0.0;
        sv[ck] = 0.0;
        sv[ck + 1] = 0.0;
    }
}

// Poisson right-hand side (dynamics.f90:977-998), k = 0..kp+1
__global__ __launch_bounds__(256) void k_rhs(const double *su, const double *sv,
                                             const double *sw, double *rhs, Geo g, Metrics M) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp + 1 || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const double X = (su[c] - su[c - g.si]) / M.dx[i - 1];
    const double Y = (sv[c] - sv[c - g.sj]) / M.dy[j - 1];
    if (k >= 1)
        rhs[c] = (X + Y) + (sw[c] - sw[c - 1]) / M.dz[k - 1];
    else
        rhs[c] = (X + Y) + (sw[c] + sw[c + 1]) / M.dz[1];
}

// ---------------------------------------------------------------------------
// horizontal means (diagnostics.f90:31-107): per level k = 0..kp+1, sum over
// the interior columns in a fixed blocked order, divided by ip*jp.
// ---------------------------------------------------------------------------
constexpr int kHmCols = 256;  // columns per block
__global__ void k_level_sums(const double *a, Geo g, double *partials) {
    // block: 64 lanes along k x 4 column groups; grid: (k chunks, column chunks)
    __shared__ double red[4][64];
    const int k = blockIdx.x * 64 + threadIdx.x;
    const long long ncol = (long long)g.ip * g.jp;
    const long long c0 = (long long)blockIdx.y * kHmCols;
    double acc = 0.0;
    if (k < g.nk) {
        // column col = (i-1)*jp + (j-1): one division for the first, then
        // steps of 4 columns with the wrap into the next i
        const long long cfirst = c0 + threadIdx.y;
        const long long cend = c0 + kHmCols < ncol ? c0 + kHmCols : ncol;
        int i = (int)(cfirst / g.jp) + 1;
        int j = (int)(cfirst - (long long)(i - 1) * g.jp) + 1;
#pragma unroll 8
        for (long long col = cfirst; col < cend; col += 4) {
            acc += a[pidx(g, k, j, i)];
            j += 4;
            while (j > g.jp) {
                j -= g.jp;
                ++i;
            }
        }
    }
    red[threadIdx.y][threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.y == 0 && k < g.nk)
        partials[(long long)blockIdx.y * g.nk + k] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

__global__ void k_level_means(const double *partials, int nchunks, Geo g, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.nk) return;
    double s = 0.0;
    // in chunk order; unrolled so that the loads are in flight together
#pragma unroll 16
    for (int q = 0; q < nchunks; ++q) s += partials[(long long)q * g.nk + k];
    out[k] = s / (double)(g.ip * g.jp);
}

// per-level means in the reference's loop order (the SUM intrinsic over
// (k, 1:jpp, 1:ipp), diagnostics.f90:55-60: j inner, i outer), one thread
// per level
__global__ void k_level_means_seq(const double *a, Geo g, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.nk) return;
    double s = 0.0;
    for (int i = 1; i <= g.ip; ++i)
        for (int j = 1; j <= g.jp; ++j) s = s + a[pidx(g, k, j, i)];
    out[k] = s / (double)(g.ip * g.jp);
}

int level_means(dcm_ctx *c, const double *a, double *out_dev) {
    const Geo &g = c->geo;
    if (c->seq_sums) {
        int tok = prof_begin(c, "hmeans");
        hipLaunchKernelGGL(k_level_means_seq, dim3((g.nk + 63) / 64), dim3(64), 0, c->stream, a,
                           g, out_dev);
        DCM_LAUNCH_CHECK();
        prof_end(c, tok);
        return DCM_OK;
    }
    const long long ncol = (long long)g.ip * g.jp;
    const int nchunks = (int)((ncol + kHmCols - 1) / kHmCols);
    const size_t need = (size_t)nchunks * g.nk;
    if (c->hm_cap < need) {
        DCM_CHECK(no_alloc_in_capture(c, "level_means"));
        if (c->hm_partials) {
            DCM_HIP(hipStreamSynchronize(c->stream));
            DCM_HIP(hipFree(c->hm_partials));
        }
        void *p = nullptr;
        DCM_HIP(hipMalloc(&p, need * sizeof(double)));
        c->hm_partials = (double *)p;
        c->hm_cap = need;
    }
    int tok = prof_begin(c, "hmeans");
    hipLaunchKernelGGL(k_level_sums, dim3((g.nk + 63) / 64, nchunks), dim3(64, 4), 0, c->stream,
                       a, g, c->hm_partials);
    hipLaunchKernelGGL(k_level_means, dim3((g.nk + 255) / 256), dim3(256), 0, c->stream,
                       c->hm_partials, nchunks, g, out_dev);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    return DCM_OK;
}

// ---------------------------------------------------------------------------
// advance_momentum (dynamics.f90:1053-1094): k = 0..kp, interior (j,i)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_advance_momentum(
    double *u, double *v, double *w, const double *zu, const double *zv, const double *zw,
    const double *su, const double *sv, const double *sw, const double *p, Geo g, Metrics M,
    double dt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const double pc = p[c];
    u[c] = zu[c] + ((su[c] - (p[c + g.si] - pc) / M.dxn[i]) * dt) / M.rhoan[k];
    v[c] = zv[c] + ((sv[c] - (p[c + g.sj] - pc) / M.dyn[j]) * dt) / M.rhoan[k];
    w[c] = zw[c] + ((sw[c] - (p[c + 1] - pc) / M.dzn[k]) * dt) / M.rhoa[k];
}

// advance_momentum and the time filter (driver_code.f90:618-626) of the same
// cells in one pass: fut = prev + dt (s - grad p)/rho as k_advance_momentum,
// then prev = 0.99 cur + 0.01 prev from the previous level the momentum
// update has just read.  The filtered current level goes over the previous
// level's buffer, which the rotation keeps as the next step's previous level
// (the reference's zu => ut); the current level's buffer, left as it was,
// becomes the next future level (the reference's tu => zut: every cell of it
// is rewritten before it is read, so the reference's filtered zu is never
// needed).  The current level is only read, so this may run while the
// advection still reads it.  k_filter_shell filters the padded cells outside
// k = 0..kp, interior (j,i).
__global__ __launch_bounds__(256) void k_momentum_filter(
    double *u, double *v, double *w, double *zu, double *zv, double *zw, const double *cu,
    const double *cv, const double *cw, const double *su, const double *sv, const double *sw,
    const double *p, Geo g, Metrics M, double dt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const double pc = p[c];
    const double za = zu[c], zb = zv[c], zc = zw[c];
    u[c] = za + ((su[c] - (p[c + g.si] - pc) / M.dxn[i]) * dt) / M.rhoan[k];
    v[c] = zb + ((sv[c] - (p[c + g.sj] - pc) / M.dyn[j]) * dt) / M.rhoan[k];
    w[c] = zc + ((sw[c] - (p[c + 1] - pc) / M.dzn[k]) * dt) / M.rhoa[k];
    zu[c] = 0.99 * cu[c] + 0.01 * za;
    zv[c] = 0.99 * cv[c] + 0.01 * zb;
    zw[c] = 0.99 * cw[c] + 0.01 * zc;
}

// the time filter of k_momentum_filter on the padded cells it does not
// cover: the plane k = kp+1, the rows j = 0 and jp+1 and the planes i = 0
// and ip+1 (levels 0..kp)
__global__ void k_filter_shell(const double *u, double *zu, const double *v, double *zv,
                               const double *w, double *zw, Geo g) {
    // 32-bit index arithmetic (unsigned divisions): the shell has far fewer
    // than 2^31 cells
    const int K = g.kp + 1;
    const int A = g.nj * g.ni, B = 2 * K * g.ni, C = 2 * K * g.jp;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < A + B + C;
         t += gridDim.x * blockDim.x) {
        int k, j, i;
        if (t < A) {
            k = g.kp + 1;
            i = (int)((unsigned)t / (unsigned)g.nj);
            j = t - i * g.nj;
        } else if (t < A + B) {
            const int e = t - A, r = (int)((unsigned)e / (unsigned)K);
            k = e - r * K;
            j = (r & 1) ? g.jp + 1 : 0;
            i = r >> 1;
        } else {
            const int e = t - A - B, r = (int)((unsigned)e / (unsigned)K);
            k = e - r * K;
            j = (r >> 1) + 1;
            i = (r & 1) ? g.ip + 1 : 0;
        }
        const long long c = pidx(g, k, j, i);
        zu[c] = 0.99 * u[c] + 0.01 * zu[c];
        zv[c] = 0.99 * v[c] + 0.01 * zv[c];
        zw[c] = 0.99 * w[c] + 0.01 * zw[c];
    }
}

// divergence_calc (diagnostics.f90:140-148) on the interior
__global__ __launch_bounds__(256) void k_divergence(const double *u, const double *v,
                                                    const double *w, double *div, Geo g,
                                                    Metrics M) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    div[c] = M.rhoan[k] * (u[c] - u[c - g.si]) / M.dx[i - 1] +
             M.rhoan[k] * (v[c] - v[c - g.sj]) / M.dy[j - 1] +
             (M.rhoa[k] * w[c] - M.rhoa[k - 1] * w[c - 1]) / M.dz[k - 1];
}

// advance_scalar_fields_3d (sgm/subgrid_3d.f90:601-621), whole arrays
__global__ void k_zero_field(double *a, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        a[t] = 0.0;
}

__global__ void k_axpy_dt(double *a, const double *s, double dt, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        a[t] = a[t] + s[t] * dt;
}

// plain copy, two doubles per thread (measurement yardstick)
__global__ void k_stream_copy(double *dst, const double *src, long long n) {
    const long long t = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 2;
    if (t + 1 < n) {
        const double2 v = *(const double2 *)(src + t);
        *(double2 *)(dst + t) = v;
    } else if (t < n) {
        dst[t] = src[t];
    }
}

// floor / lid rows of the new current level (driver_code.f90:663-674)
__global__ void k_momentum_rows(double *u, double *v, double *w, Geo g) {
    const int P = g.nj * g.ni;  // (32-bit index arithmetic: a plane is far below 2^31)
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
        const int i = (int)((unsigned)t / (unsigned)g.nj), j = t - i * g.nj;
        const long long c0 = pidx(g, 0, j, i), ck = pidx(g, g.kp + 1, j, i);
        w[c0] = 0.0;
        u[c0] = -u[c0 + 1];
        v[c0] = -v[c0 + 1];
        w[ck] = 0.0;
        u[ck] = u[ck - 1];
        v[ck] = v[ck - 1];
    }
}

// initial previous-level rows (driver_code.f90:287-295)
__global__ void k_init_rows(double *zu, double *zv, double *zw, Geo g) {
    const int P = g.nj * g.ni;  // (32-bit index arithmetic: a plane is far below 2^31)
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
        const int i = (int)((unsigned)t / (unsigned)g.nj), j = t - i * g.nj;
        zw[pidx(g, 0, j, i)] = -zw[pidx(g, 1, j, i)];
        zu[pidx(g, 1, j, i)] = 0.0;
        zv[pidx(g, 1, j, i)] = 0.0;
    }
}
__global__ void k_init_rows_top(double *zu, double *zv, double *zw, Geo g) {
    const int P = g.nj * g.ni;  // (32-bit index arithmetic: a plane is far below 2^31)
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
        const int i = (int)((unsigned)t / (unsigned)g.nj), j = t - i * g.nj;
        zw[pidx(g, g.kp, j, i)] = -zw[pidx(g, g.kp - 1, j, i)];
        zw[pidx(g, g.kp + 1, j, i)] = 0.0;
        zu[pidx(g, g.kp, j, i)] = 0.0;
        zv[pidx(g, g.kp, j, i)] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// internal entry points used by the model and the C-ABI
// ---------------------------------------------------------------------------
// sources (dynamics.f90:639-1000).  ph == nullptr: the inviscid branch;
// otherwise the damping layer, forcing and subsidence terms and, when
// ph->viscous, the subgrid model on sf.
int sources_dev(dcm_ctx *c, const double *u, const double *v, const double *w, const double *th,
                const double *theta_d, const double *thbar_d, double *su, double *sv, double *sw,
                double *sth, double *rhs, const PhysicsDev *ph, const SubgridFields *sf) {
    const Geo &g = c->geo;
    Interior L = interior_launch(g);
    L.grid.z = (unsigned)((g.ip + kSrcPlanes - 1) / kSrcPlanes);
    int tok = prof_begin(c, "sources");
    hipLaunchKernelGGL(k_sources, L.grid, L.block, 0, c->stream, u, v, w, th, theta_d, thbar_d,
                       su, sv, sw, g, c->met);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    if (ph) DCM_CHECK(relax_dev(c, *ph, sf->zu, sf->zv, sf->zw, th, su, sv, sw, sth));
    if (ph && ph->viscous) {
        // the subgrid model fills su, sv, sw, sth itself (:878-906)
        DCM_CHECK(subgrid_dev(c, *ph, u, v, w, th, *sf, su, sv, sw, sth));
    } else {
        // inviscid branch: exchange_full(su,sv,sw,sth) (dynamics.f90:907-917),
        // the four fills in one launch
        FillSet<double> fs{};
        fs.a[0] = su;
        fs.a[1] = sv;
        fs.a[2] = sw;
        fs.a[3] = sth;
        DCM_CHECK(fill_set<double>(c, fs, sth ? 4 : 3, FILL_FULL, 0.0, 0.0));
    }
    hipLaunchKernelGGL(k_sources_rows, dim3(256), dim3(kBlock), 0, c->stream, su, sv, sw, g);
    DCM_LAUNCH_CHECK();
    Interior R = interior_launch(g, 2);
    tok = prof_begin(c, "rhs");
    hipLaunchKernelGGL(k_rhs, R.grid, R.block, 0, c->stream, su, sv, sw, rhs, g, c->met);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    return DCM_OK;
}

// the model's momentum update with the time filter of the same step
// (k_momentum_filter, k_filter_shell): fut from prev, then prev = the
// filtered current level
int momentum_filter_dev(dcm_ctx *c, double dt, double *const fut[3], double *const prev[3],
                        const double *const cur[3], const double *su, const double *sv,
                        const double *sw, const double *p) {
    const Geo &g = c->geo;
    Interior L = interior_launch(g, 1);
    int tok = prof_begin(c, "advance_momentum");
    hipLaunchKernelGGL(k_momentum_filter, L.grid, L.block, 0, c->stream, fut[0], fut[1], fut[2],
                       prev[0], prev[1], prev[2], cur[0], cur[1], cur[2], su, sv, sw, p, g,
                       c->met, dt);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    tok = prof_begin(c, "filter");
    const long long shell = (long long)g.nj * g.ni + 2LL * (g.kp + 1) * (g.ni + g.jp);
    hipLaunchKernelGGL(k_filter_shell, dim3(blocks_for(shell)), dim3(kBlock), 0, c->stream,
                       cur[0], prev[0], cur[1], prev[1], cur[2], prev[2], g);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    return DCM_OK;
}

int advance_momentum_dev(dcm_ctx *c, double dt, double *u, double *v, double *w,
                         const double *zu, const double *zv, const double *zw, const double *su,
                         const double *sv, const double *sw, const double *p) {
    const Geo &g = c->geo;
    Interior L = interior_launch(g, 1);
    int tok = prof_begin(c, "advance_momentum");
    hipLaunchKernelGGL(k_advance_momentum, L.grid, L.block, 0, c->stream, u, v, w, zu, zv, zw, su,
                       sv, sw, p, g, c->met, dt);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    return DCM_OK;
}

}  // namespace dcm

// sets *bad when the shift m does not round-trip the bcs: (bc - m) + m != bc
__global__ void k_bc_roundtrip(const double *m, double lbc, double ubc, int *bad) {
    if (threadIdx.x != 0) return;
    const double mm = *m;
    if ((lbc - mm) + mm != lbc || (ubc - mm) + mm != ubc) *bad = 1;
}

namespace dcm {
// a captured phase of the step (the kernels before or after the pressure
// solve), valid while the key it was captured for -- n mod 2 (the buffers
// of the time levels), precision, summation order, MPDATA variant and the
// model's work generation of every model it covers -- is unchanged
struct PhaseGraph {
    std::vector<int> key;
    hipGraphExec_t exec = nullptr;
};
}  // namespace dcm

// ---------------------------------------------------------------------------
// model state
// ---------------------------------------------------------------------------
struct dcm_model {
    dcm_ctx *c = nullptr;
    int nq = 0, ncat = 0;
    std::vector<int> c_s, c_e;
    double dt = 1.0, ptol = 1e-10;
    int kord = 2, monotone = 1, advection_scheme = 1;
    int precision = DCM_FP64;
    double *lev[3][3] = {};  // [component u,v,w][buffer ut,zut,tut]
    double *th = nullptr, *p = nullptr, *psrc = nullptr;
    double *su = nullptr, *sv = nullptr, *sw = nullptr, *sth = nullptr;
    double *q = nullptr;
    double *oned = nullptr;  // theta, thetan, thbar (kp+2 each)
    float *shadow = nullptr; // FP32 u,v,w for the mixed advection
    std::vector<double> xg, yg, zg;
    int n = 0;               // completed steps (pointer rotation)
    double t1 = 0.0, t2 = 0.0;
    // advection_scheme 2: the bcs of theta + thetan, thetan(0) and thetan(kp+1)
    // (find_base_top, mpi_module.f90:306-329), and a device flag set when a
    // step's (bc - m) + m differs from bc
    double thbase = 0.0, thtop = 0.0;
    int *bc_inexact = nullptr;
    double inner_rtol = 1e-4;  // fewest inner iterations at 256x256x128 (1e-3..1e-5 measured)
    bool check_residual = false;
    std::vector<void *> allocs;
    // physics options of sources (dcm_model_set_physics): device profiles
    // in phys1d (ubar vbar wbar dampfacn dampfac u_force v_force w_subs
    // lamsq thetan, kp+2 each) and the subgrid fields
    bool phys_on = false;
    dcm::PhysicsDev phys;
    double *phys1d = nullptr;
    double *strain = nullptr, *vism = nullptr, *vist = nullptr, *viss = nullptr, *sq = nullptr;
    // divergence_calc's diagnostic (driver_code.f90:443-444), kept once an
    // output file is attached
    double *div = nullptr;
    // bumped whenever the work of a step changes (physics options, div), so
    // that an ensemble does not replay graphs captured before the change
    int gen = 0;
    // dcm_model_step: the phases before and after the solve replayed from
    // HIP graphs, each captured once its key has run eagerly (so every
    // scratch buffer exists: nothing is allocated or cleared in a capture)
    bool use_graphs = true;
    std::vector<std::vector<int>> eager[4];
    dcm::PhaseGraph pre[3], post[3], adv[3], mom[3];
    // the mixed path advects theta and q (they do not depend on the
    // pressure) on a stream of its own while the pressure solve runs:
    // forked from the main stream after the sources, joined before the
    // momentum update
    hipStream_t adv_stream = nullptr;
    hipEvent_t adv_fork = nullptr, adv_join = nullptr;

    double *theta() { return oned; }
    double *thetan() { return oned + (c->geo.kp + 2); }
    double *thbar() { return oned + 2 * (c->geo.kp + 2); }
    // driver_code.f90:634-646: after n steps
    // the buffers of the three levels at step n: the previous level stays in
    // buffer 1 (the filter writes the filtered current level there), the
    // current and future levels swap buffers 0 and 2 every step
    void levels(double **cur, double **prev, double **fut) {
        static const int map[2][3] = {{0, 1, 2}, {2, 1, 0}};  // [n%2][cur,prev,fut]
        const int r = n % 2;
        for (int d = 0; d < 3; ++d) {
            cur[d] = lev[d][map[r][0]];
            prev[d] = lev[d][map[r][1]];
            fut[d] = lev[d][map[r][2]];
        }
    }
};

using namespace dcm;

static int model_alloc(dcm_model *m, size_t bytes, void **out) {
    DCM_CHECK(no_alloc_in_capture(m->c, "dcm_model"));
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        set_error("dcm_model: out of device memory");
        return DCM_OUT_OF_MEMORY;
    }
    DCM_HIP(hipMemsetAsync(p, 0, bytes, m->c->stream));
    m->allocs.push_back(p);
    *out = p;
    return DCM_OK;
}

namespace dcm {

int model_view(dcm_model *m, ModelView *v) {
    if (!m || !v) return DCM_BAD_ARG;
    double *cur[3], *prev[3], *fut[3];
    m->levels(cur, prev, fut);
    v->c = m->c;
    v->u = cur[0];
    v->v = cur[1];
    v->w = cur[2];
    v->th = m->th;
    v->p = m->p;
    v->q = m->q;
    v->nq = m->nq;
    v->div = m->div;
    return DCM_OK;
}

int model_enable_div(dcm_model *m) {
    if (!m) return DCM_BAD_ARG;
    if (m->div) return DCM_OK;
    void *p = nullptr;
    DCM_CHECK(model_alloc(m, (size_t)m->c->geo.ntot * sizeof(double), &p));  // div = 0 (:461)
    m->div = (double *)p;
    ++m->gen;
    return DCM_OK;
}

}  // namespace dcm

extern "C" {

int dcm_hmeans(dcm_ctx *c, const double *u, const double *v, const double *w, const double *th,
               const double *q, int nq, double *ubar, double *vbar, double *wbar, double *thbar,
               double *qbar) {
    if (!c || !th) return DCM_BAD_ARG;
    const Geo &g = c->geo;
    double *d;
    DCM_CHECK(work_field64(c, 0, &d));  // scratch for the 1-D results
    const double *src[4] = {u, v, w, th};
    double *dst[4] = {ubar, vbar, wbar, thbar};
    for (int a = 0; a < 4; ++a) {
        if (!src[a] || !dst[a]) continue;
        DCM_CHECK(level_means(c, src[a], d));
        DCM_HIP(hipMemcpyAsync(dst[a], d, g.nk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        DCM_HIP(hipStreamSynchronize(c->stream));
    }
    if (q && qbar)
        for (int n = 0; n < nq; ++n) {
            DCM_CHECK(level_means(c, q + (size_t)n * g.ntot, d));
            DCM_HIP(hipMemcpyAsync(qbar + (size_t)n * g.nk, d, g.nk * sizeof(double),
                                   hipMemcpyDeviceToHost, c->stream));
            DCM_HIP(hipStreamSynchronize(c->stream));
        }
    return DCM_OK;
}

int dcm_sources(dcm_ctx *c, const double *u, const double *v, const double *w, const double *th,
                const double *theta, const double *thbar, double *su, double *sv, double *sw,
                double *sth, double *rhs) {
    if (!c || !u || !v || !w || !th || !theta || !thbar || !su || !sv || !sw || !rhs)
        return DCM_BAD_ARG;
    const double *dth, *dthbar;
    DCM_CHECK(upload_1d(c, theta, (size_t)c->geo.kp + 2, 1, &dth));
    DCM_CHECK(upload_1d(c, thbar, (size_t)c->geo.kp + 2, 2, &dthbar));
    return sources_dev(c, u, v, w, th, dth, dthbar, su, sv, sw, sth, rhs, nullptr, nullptr);
}

// the device form of a dcm_physics: host profiles uploaded to 1-D scratch
// slots 3..12 (ubar..wbar may be NULL when damping and subsidence are off)
static int physics_upload(dcm_ctx *c, const dcm_physics *h, const double *ubar,
                          const double *vbar, const double *wbar, int nq, PhysicsDev *ph) {
    const size_t n = (size_t)c->geo.kp + 2;
    *ph = PhysicsDev();
    ph->viscous = h->viscous != 0;
    ph->damping_layer = h->damping_layer != 0;
    ph->forcing = h->forcing != 0;
    ph->divergence = h->divergence != 0;
    ph->moisture = nq > 0;
    ph->nq = nq;
    ph->forcing_tau = h->forcing_tau;
    ph->z0 = h->z0;
    ph->z0th = h->z0th;
    auto up = [&](const double *a, int slot, const double **out, const char *what) -> int {
        if (!a) {
            set_error(std::string("dcm_physics: ") + what + " is required by the options set");
            return DCM_BAD_ARG;
        }
        return upload_1d(c, a, n, slot, out);
    };
    if (ph->damping_layer) {
        DCM_CHECK(up(ubar, 3, &ph->ubar, "ubar"));
        DCM_CHECK(up(vbar, 4, &ph->vbar, "vbar"));
        DCM_CHECK(up(h->dampfacn, 6, &ph->dampfacn, "dampfacn"));
        DCM_CHECK(up(h->dampfac, 7, &ph->dampfac, "dampfac"));
    }
    if (ph->forcing) {
        if (!(h->forcing_tau != 0.0)) {
            set_error("dcm_physics: forcing_tau must be non-zero");
            return DCM_BAD_ARG;
        }
        DCM_CHECK(up(h->u_force, 8, &ph->u_force, "u_force"));
        DCM_CHECK(up(h->v_force, 9, &ph->v_force, "v_force"));
    }
    if (ph->divergence) {
        DCM_CHECK(up(wbar, 5, &ph->wbar, "wbar"));
        DCM_CHECK(up(h->w_subs, 10, &ph->w_subs, "w_subs"));
    }
    if (ph->viscous) {
        if (!h->z || !h->theta || !h->thetan || !(h->z0 > 0.0)) {
            set_error("dcm_physics: the subgrid model needs z, theta, thetan and z0 > 0");
            return DCM_BAD_ARG;
        }
        DCM_CHECK(up(h->lamsq, 11, &ph->lamsq, "lamsq"));
        DCM_CHECK(up(h->thetan, 12, &ph->thetan, "thetan"));
        ph->z1 = h->z[1];
        ph->theta0 = h->theta[0];
        ph->thetan1 = h->thetan[1];
        ph->log_z1_z0 = std::log(h->z[1] / h->z0);
    }
    return DCM_OK;
}

int dcm_subgrid(dcm_ctx *c, const dcm_physics *ph, const double *u, const double *v,
                const double *w, const double *zu, const double *zv, const double *zw,
                const double *th, const double *q, int nq, double *vism, double *vist,
                double *viss, double *strain, double *su, double *sv, double *sw, double *sth,
                double *sq) {
    if (!c || !ph || !u || !v || !w || !zu || !zv || !zw || !th || !vism || !vist || !strain ||
        !su || !sv || !sw || !sth || nq < 0 || (nq > 0 && (!q || !sq || !viss)))
        return DCM_BAD_ARG;
    dcm_physics h = *ph;
    h.viscous = 1;
    h.damping_layer = h.forcing = h.divergence = 0;
    PhysicsDev d;
    DCM_CHECK(physics_upload(c, &h, nullptr, nullptr, nullptr, nq, &d));
    SubgridFields sf;
    sf.zu = zu;
    sf.zv = zv;
    sf.zw = zw;
    sf.q = q;
    sf.sq = sq;
    sf.strain = strain;
    sf.vism = vism;
    sf.vist = vist;
    sf.viss = viss;
    return subgrid_dev(c, d, u, v, w, th, sf, su, sv, sw, sth);
}

int dcm_sources_phys(dcm_ctx *c, const dcm_physics *ph, const double *ubar, const double *vbar,
                     const double *wbar, const double *u, const double *v, const double *w,
                     const double *zu, const double *zv, const double *zw, const double *th,
                     const double *q, int nq, const double *theta, const double *thbar,
                     double *su, double *sv, double *sw, double *sth, double *sq,
                     double *strain, double *vism, double *vist, double *viss, double *rhs) {
    if (!c || !ph || !u || !v || !w || !zu || !zv || !zw || !th || !theta || !thbar || !su ||
        !sv || !sw || !sth || !rhs || nq < 0)
        return DCM_BAD_ARG;
    if (ph->viscous && (!strain || !vism || !vist || (nq > 0 && (!q || !sq || !viss))))
        return DCM_BAD_ARG;
    PhysicsDev d;
    DCM_CHECK(physics_upload(c, ph, ubar, vbar, wbar, nq, &d));
    const double *dth, *dthbar;
    DCM_CHECK(upload_1d(c, theta, (size_t)c->geo.kp + 2, 1, &dth));
    DCM_CHECK(upload_1d(c, thbar, (size_t)c->geo.kp + 2, 2, &dthbar));
    SubgridFields sf;
    sf.zu = zu;
    sf.zv = zv;
    sf.zw = zw;
    sf.q = q;
    sf.sq = sq;
    sf.strain = strain;
    sf.vism = vism;
    sf.vist = vist;
    sf.viss = viss;
    return sources_dev(c, u, v, w, th, dth, dthbar, su, sv, sw, sth, rhs, &d, &sf);
}

int dcm_advance_momentum(dcm_ctx *c, double dt, double *u, double *v, double *w,
                         const double *zu, const double *zv, const double *zw, const double *su,
                         const double *sv, const double *sw, const double *p) {
    if (!c || !u || !v || !w || !zu || !zv || !zw || !su || !sv || !sw || !p) return DCM_BAD_ARG;
    return advance_momentum_dev(c, dt, u, v, w, zu, zv, zw, su, sv, sw, p);
}

int dcm_divergence(dcm_ctx *c, const double *u, const double *v, const double *w, double *div) {
    if (!c || !u || !v || !w || !div) return DCM_BAD_ARG;
    Interior L = interior_launch(c->geo, 0);
    hipLaunchKernelGGL(k_divergence, L.grid, L.block, 0, c->stream, u, v, w, div, c->geo, c->met);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

int dcm_advance_scalars(dcm_ctx *c, double dt, double *q, const double *sq, double *th,
                        const double *sth, int nq, int moisture) {
    if (!c || !th || !sth) return DCM_BAD_ARG;
    const long long n = c->geo.ntot;
    hipLaunchKernelGGL(k_axpy_dt, dim3(2048), dim3(kBlock), 0, c->stream, th, sth, dt, n);
    DCM_LAUNCH_CHECK();
    if (moisture && nq > 0) {
        if (!q || !sq) return DCM_BAD_ARG;
        hipLaunchKernelGGL(k_axpy_dt, dim3(2048), dim3(kBlock), 0, c->stream, q, sq, dt,
                           n * nq);
        DCM_LAUNCH_CHECK();
    }
    return DCM_OK;
}

int dcm_model_create(dcm_ctx *c, int nq, int ncat, const int *c_s, const int *c_e, double dt,
                     int kord, int monotone, int advection_scheme, double ptol,
                     const double *theta, const double *thetan, const double *xg,
                     const double *yg, const double *zg, dcm_model **out) {
    if (!c || !out || nq < 0 || ncat < 0 || (ncat > 0 && (!c_s || !c_e)) || !theta || !thetan ||
        advection_scheme < 0 || advection_scheme > 2 || kord < 1) {
        set_error("dcm_model_create: bad arguments");
        return DCM_BAD_ARG;
    }
    for (int q = 0; q < ncat; ++q)
        if (c_s[q] < 1 || c_e[q] > nq || c_e[q] < c_s[q] || c_e[q] - c_s[q] + 1 > 32) {
            set_error("dcm_model_create: bad category bounds");
            return DCM_BAD_ARG;
        }
    dcm_model *m = new dcm_model();
    m->c = c;
    m->nq = nq;
    m->ncat = ncat;
    m->c_s.assign(c_s, c_s + ncat);
    m->c_e.assign(c_e, c_e + ncat);
    m->dt = dt;
    m->kord = kord;
    m->monotone = monotone;
    m->advection_scheme = advection_scheme;
    m->ptol = ptol;
    const Geo &g = c->geo;
    const size_t fb = (size_t)g.ntot * sizeof(double);
    void *p;
    int st = DCM_OK;
    for (int d = 0; d < 3 && st == DCM_OK; ++d)
        for (int l = 0; l < 3 && st == DCM_OK; ++l) {
            st = model_alloc(m, fb, &p);
            m->lev[d][l] = (double *)p;
        }
    double **singles[7] = {&m->th, &m->p, &m->psrc, &m->su, &m->sv, &m->sw, &m->sth};
    for (int a = 0; a < 7 && st == DCM_OK; ++a) {
        st = model_alloc(m, fb, &p);
        *singles[a] = (double *)p;
    }
    if (st == DCM_OK && nq > 0) {
        st = model_alloc(m, fb * nq, &p);
        m->q = (double *)p;
    }
    if (st == DCM_OK) {
        st = model_alloc(m, 3 * (g.kp + 2) * sizeof(double), &p);
        m->oned = (double *)p;
    }
    if (st == DCM_OK && advection_scheme == 2) {
        st = model_alloc(m, 256, &p);
        m->bc_inexact = (int *)p;
        m->thbase = thetan[0];
        m->thtop = thetan[g.kp + 1];
    }
    if (st != DCM_OK) {
        dcm_model_destroy(m);
        return st;
    }
    (void)hipMemcpyAsync(m->theta(), theta, (g.kp + 2) * sizeof(double), hipMemcpyHostToDevice,
                   c->stream);
    (void)hipMemcpyAsync(m->thetan(), thetan, (g.kp + 2) * sizeof(double), hipMemcpyHostToDevice,
                   c->stream);
    if (xg) m->xg.assign(xg, xg + g.ip + 2);
    if (yg) m->yg.assign(yg, yg + g.jp + 2);
    if (zg) m->zg.assign(zg, zg + g.kp + 2);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("dcm_model_create: upload failed");
        dcm_model_destroy(m);
        return DCM_HIP_ERROR;
    }
    if (hipStreamCreateWithFlags(&m->adv_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&m->adv_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->adv_join, hipEventDisableTiming) != hipSuccess) {
        set_error("dcm_model_create: stream or event creation failed");
        dcm_model_destroy(m);
        return DCM_HIP_ERROR;
    }
    *out = m;
    return DCM_OK;
}

int dcm_model_destroy(dcm_model *m) {
    if (!m) return DCM_BAD_ARG;
    (void)hipStreamSynchronize(m->c->stream);
    for (int r = 0; r < 3; ++r)
        for (dcm::PhaseGraph *g : {&m->pre[r], &m->post[r], &m->adv[r], &m->mom[r]})
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (m->adv_stream) (void)hipStreamDestroy(m->adv_stream);
    if (m->adv_fork) (void)hipEventDestroy(m->adv_fork);
    if (m->adv_join) (void)hipEventDestroy(m->adv_join);
    for (void *p : m->allocs) (void)hipFree(p);
    if (m->shadow) (void)hipFree(m->shadow);
    delete m;
    return DCM_OK;
}

int dcm_model_set_physics(dcm_model *m, const dcm_physics *ph) {
    if (!m) return DCM_BAD_ARG;
    DCM_HIP(hipStreamSynchronize(m->c->stream));
    ++m->gen;
    if (!ph) {
        m->phys_on = false;
        return DCM_OK;
    }
    dcm_ctx *c = m->c;
    const Geo &g = c->geo;
    const size_t nk = (size_t)g.kp + 2;
    if (ph->viscous && (!ph->z || !ph->theta || !ph->thetan || !ph->lamsq || !(ph->z0 > 0.0))) {
        set_error("dcm_model_set_physics: the subgrid model needs lamsq, z, theta, thetan, z0 > 0");
        return DCM_BAD_ARG;
    }
    if (ph->damping_layer && (!ph->dampfacn || !ph->dampfac)) {
        set_error("dcm_model_set_physics: the damping layer needs dampfacn and dampfac");
        return DCM_BAD_ARG;
    }
    if (ph->forcing && (!ph->u_force || !ph->v_force || !(ph->forcing_tau != 0.0))) {
        set_error("dcm_model_set_physics: forcing needs u_force, v_force and forcing_tau != 0");
        return DCM_BAD_ARG;
    }
    if (ph->divergence && !ph->w_subs) {
        set_error("dcm_model_set_physics: subsidence needs w_subs");
        return DCM_BAD_ARG;
    }
    void *p;
    if (!m->phys1d) {
        DCM_CHECK(model_alloc(m, 10 * nk * sizeof(double), &p));
        m->phys1d = (double *)p;
    }
    if (!m->sth) {
        DCM_CHECK(model_alloc(m, g.ntot * sizeof(double), &p));
        m->sth = (double *)p;
    }
    if (ph->viscous) {
        double **f[3] = {&m->strain, &m->vism, &m->vist};
        for (double **a : f)
            if (!*a) {
                DCM_CHECK(model_alloc(m, g.ntot * sizeof(double), &p));
                *a = (double *)p;
            }
        if (m->nq > 0 && !m->sq) {
            DCM_CHECK(model_alloc(m, (size_t)m->nq * g.ntot * sizeof(double), &p));
            m->sq = (double *)p;
            DCM_CHECK(model_alloc(m, (size_t)m->nq * g.ntot * sizeof(double), &p));
            m->viss = (double *)p;
        }
    }
    // profiles: 0 ubar 1 vbar 2 wbar (computed each step) 3 dampfacn
    // 4 dampfac 5 u_force 6 v_force 7 w_subs 8 lamsq 9 thetan
    std::vector<double> h(10 * nk, 0.0);
    const double *src[10] = {nullptr, nullptr, nullptr, ph->dampfacn, ph->dampfac, ph->u_force,
                             ph->v_force, ph->w_subs, ph->lamsq, ph->thetan};
    for (int a = 3; a < 10; ++a)
        if (src[a]) std::memcpy(h.data() + a * nk, src[a], nk * sizeof(double));
    DCM_HIP(hipMemcpyAsync(m->phys1d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    PhysicsDev d;
    d.viscous = ph->viscous != 0;
    d.damping_layer = ph->damping_layer != 0;
    d.forcing = ph->forcing != 0;
    d.divergence = ph->divergence != 0;
    d.moisture = m->nq > 0 && m->sq != nullptr;
    d.nq = m->nq;
    d.forcing_tau = ph->forcing_tau;
    d.z0 = ph->z0;
    d.z0th = ph->z0th;
    const double *P = m->phys1d;
    d.ubar = P;
    d.vbar = P + nk;
    d.wbar = P + 2 * nk;
    d.dampfacn = P + 3 * nk;
    d.dampfac = P + 4 * nk;
    d.u_force = P + 5 * nk;
    d.v_force = P + 6 * nk;
    d.w_subs = P + 7 * nk;
    d.lamsq = P + 8 * nk;
    d.thetan = P + 9 * nk;
    if (d.viscous) {
        d.z1 = ph->z[1];
        d.theta0 = ph->theta[0];
        d.thetan1 = ph->thetan[1];
        d.log_z1_z0 = std::log(ph->z[1] / ph->z0);
    }
    m->phys = d;
    m->phys_on = true;
    return DCM_OK;
}

int dcm_model_set_precision(dcm_model *m, int precision) {
    if (!m || (precision != DCM_FP64 && precision != DCM_MIXED)) return DCM_BAD_ARG;
    m->precision = precision;
    if (precision == DCM_MIXED && !m->shadow) {
        void *p = nullptr;
        if (hipMalloc(&p, 3 * (size_t)m->c->geo.ntot * sizeof(float)) != hipSuccess) {
            set_error("dcm_model_set_precision: out of device memory");
            return DCM_OUT_OF_MEMORY;
        }
        m->shadow = (float *)p;
    }
    return DCM_OK;
}

static int field_ptr(dcm_model *m, int field, int index, double **ptr, size_t *count) {
    const Geo &g = m->c->geo;
    double *cur[3], *prev[3], *fut[3];
    m->levels(cur, prev, fut);
    *count = (size_t)g.ntot;
    switch (field) {
        case DCM_F_U: *ptr = cur[0]; break;
        case DCM_F_V: *ptr = cur[1]; break;
        case DCM_F_W: *ptr = cur[2]; break;
        case DCM_F_ZU: *ptr = prev[0]; break;
        case DCM_F_ZV: *ptr = prev[1]; break;
        case DCM_F_ZW: *ptr = prev[2]; break;
        case DCM_F_TH: *ptr = m->th; break;
        case DCM_F_P: *ptr = m->p; break;
        case DCM_F_PSRC: *ptr = m->psrc; break;
        case DCM_F_SU: *ptr = m->su; break;
        case DCM_F_SV: *ptr = m->sv; break;
        case DCM_F_SW: *ptr = m->sw; break;
        case DCM_F_STH: *ptr = m->sth; break;
        case DCM_F_Q:
            if (index < 0 || index >= m->nq) return DCM_BAD_ARG;
            *ptr = m->q + (size_t)index * g.ntot;
            break;
        case DCM_F_THBAR:
            *ptr = m->thbar();
            *count = (size_t)g.kp + 2;
            break;
        default:
            return DCM_BAD_ARG;
    }
    return DCM_OK;
}

int dcm_model_set(dcm_model *m, int field, int index, const double *host) {
    if (!m || !host) return DCM_BAD_ARG;
    double *d;
    size_t n;
    DCM_CHECK(field_ptr(m, field, index, &d, &n));
    DCM_HIP(hipMemcpyAsync(d, host, n * sizeof(double), hipMemcpyHostToDevice, m->c->stream));
    DCM_HIP(hipStreamSynchronize(m->c->stream));
    return DCM_OK;
}

int dcm_model_get(dcm_model *m, int field, int index, double *host) {
    if (!m || !host) return DCM_BAD_ARG;
    double *d;
    size_t n;
    DCM_CHECK(field_ptr(m, field, index, &d, &n));
    DCM_HIP(hipMemcpyAsync(host, d, n * sizeof(double), hipMemcpyDeviceToHost, m->c->stream));
    DCM_HIP(hipStreamSynchronize(m->c->stream));
    return DCM_OK;
}

int dcm_model_field(dcm_model *m, int field, int index, double **dptr) {
    if (!m || !dptr) return DCM_BAD_ARG;
    size_t n;
    return field_ptr(m, field, index, dptr, &n);
}

int dcm_model_init(dcm_model *m) {
    if (!m) return DCM_BAD_ARG;
    dcm_ctx *c = m->c;
    const Geo &g = c->geo;
    double *cur[3], *prev[3], *fut[3];
    m->levels(cur, prev, fut);
    // initialisation.f90:861-870 (every q field here)
    for (int d = 0; d < 3; ++d) DCM_CHECK(fill<double>(c, cur[d], FILL_FULL, 0.0, 0.0));
    DCM_CHECK(fill<double>(c, m->th, FILL_FULL, 0.0, 0.0));
    for (int n = 0; n < m->nq; ++n)
        DCM_CHECK(fill<double>(c, m->q + (size_t)n * g.ntot, FILL_FULL, 0.0, 0.0));
    // previous and future levels start as copies of the current one
    for (int d = 0; d < 3; ++d) {
        DCM_HIP(hipMemcpyAsync(prev[d], cur[d], g.ntot * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
        DCM_HIP(hipMemcpyAsync(fut[d], cur[d], g.ntot * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
    }
    hipLaunchKernelGGL(k_init_rows, dim3(256), dim3(kBlock), 0, c->stream, prev[0], prev[1],
                       prev[2], g);
    hipLaunchKernelGGL(k_init_rows_top, dim3(256), dim3(kBlock), 0, c->stream, prev[0], prev[1],
                       prev[2], g);
    DCM_LAUNCH_CHECK();
    DCM_HIP(hipMemsetAsync(m->sth, 0, g.ntot * sizeof(double), c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    return DCM_OK;
}

// one pass of the model_driver loop, driver_code.f90:304-676 (hot path)
// model_driver up to the pressure solve (driver_code.f90:433-487)
static int model_pre_solve(dcm_model *m) {
    dcm_ctx *c = m->c;
    double *cur[3], *prev[3], *fut[3];
    m->levels(cur, prev, fut);
    // Without physics options sth = 0 (:340) stays zero (sources' inviscid
    // branch only fills it) and so does sq: the scalar update of :601-603
    // (q += sq*dt, th += sth*dt) is an exact no-op and is skipped.
    if (m->phys_on) {
        // :340-343: sth = 0 and sq = 0 before the sources add to them; the
        // means the damping layer and subsidence read (:433-435)
        const long long n = c->geo.ntot;
        hipLaunchKernelGGL(k_zero_field, dim3(2048), dim3(kBlock), 0, c->stream, m->sth, n);
        if (m->sq)
            hipLaunchKernelGGL(k_zero_field, dim3(2048), dim3(kBlock), 0, c->stream, m->sq,
                               n * m->nq);
        DCM_LAUNCH_CHECK();
        if (m->phys.damping_layer) {
            DCM_CHECK(level_means(c, cur[0], (double *)m->phys.ubar));
            DCM_CHECK(level_means(c, cur[1], (double *)m->phys.vbar));
        }
        if (m->phys.divergence) DCM_CHECK(level_means(c, cur[2], (double *)m->phys.wbar));
    }
    DCM_CHECK(level_means(c, m->th, m->thbar()));                  // :433-435
    if (m->div) {                                                   // :443-444
        const Interior L = interior_launch(c->geo);
        hipLaunchKernelGGL(k_divergence, L.grid, L.block, 0, c->stream, cur[0], cur[1], cur[2],
                           m->div, c->geo, c->met);
        DCM_LAUNCH_CHECK();
    }
    if (m->phys_on) {
        SubgridFields sf;
        sf.zu = prev[0];
        sf.zv = prev[1];
        sf.zw = prev[2];
        sf.q = m->q;
        sf.sq = m->sq;
        sf.strain = m->strain;
        sf.vism = m->vism;
        sf.vist = m->vist;
        sf.viss = m->viss;
        DCM_CHECK(sources_dev(c, cur[0], cur[1], cur[2], m->th, m->theta(), m->thbar(), m->su,
                              m->sv, m->sw, m->sth, m->psrc, &m->phys, &sf));  // :455-479
    } else {
        DCM_CHECK(sources_dev(c, cur[0], cur[1], cur[2], m->th, m->theta(), m->thbar(), m->su,
                              m->sv, m->sw, nullptr, m->psrc, nullptr, nullptr));  // :455-479
    }
    DCM_CHECK(fill<double>(c, m->psrc, FILL_ALONG_DIM_WO, 0.0, 0.0)); // :485-487
    return DCM_OK;
}

static void add_stats(dcm_step_stats *stats, const SolverOut &so, double true_resid) {
    if (!stats) return;
    stats->iters_total += so.iters;
    if (so.iters > stats->iters_max) stats->iters_max = so.iters;
    stats->outer_total += so.outer;
    stats->fallbacks += so.fallback;
    stats->last_err = so.err;
    if (true_resid > stats->max_true_resid) stats->max_true_resid = true_resid;
}

// model_driver after the pressure solve (driver_code.f90:506-674)
// The advection of theta and q after the solve (driver_code.f90:506-580):
// the reference-state advection and MPDATA of every field, from the current
// velocities; nothing here reads the pressure.
static int model_advect(dcm_model *m) {
    dcm_ctx *c = m->c;
    const Geo &g = c->geo;
    double *cur[3], *prev[3], *fut[3];
    m->levels(cur, prev, fut);
    int st;
    // :506-507 exchange_along_dim_wo(p): the solver already left the faces
    // periodic, so the fill is an exact no-op here.
    if (m->advection_scheme == 0 || m->advection_scheme == 1) {
        DCM_CHECK(adv_ref_state_dev(c, m->dt, cur[2], m->th, m->thetan()));  // :518-520
        DCM_CHECK(fill<double>(c, m->th, FILL_FULL, 0.0, 0.0));            // :526-527
    }
    const bool mixed = m->precision == DCM_MIXED;
    // the bcs of every q category: q1 = q2 = 0 (:571-580)
    std::vector<double> zeros(m->nq > 0 ? m->nq : 1, 0.0);
    if (m->advection_scheme == 0) {
        // :539-546 first_order_upstream_3d, then exchange_full(th, t1, t2);
        // FP64 on both paths (one pass, no intermediates to narrow)
        DCM_CHECK(upwind_dev(c, m->dt, cur[0], cur[1], cur[2], m->th, 0));
        DCM_CHECK(fill<double>(c, m->th, FILL_FULL, m->t1, m->t2));
    }
    if (mixed) {
        // theta (mpdata_3d, :548-556) and every q category (mpdata_vec_3d,
        // :571-580) in one batched FP32 MPDATA: the fields are independent
        const bool add = m->advection_scheme == 2;
        double lb = add ? m->thbase : m->t1, ub = add ? m->thtop : m->t2;
        std::vector<Mp32Cat> cats;
        if (m->advection_scheme != 0) cats.push_back(Mp32Cat{m->th, 1, &lb, &ub, 1});
        for (int cat = 0; cat < m->ncat; ++cat) {
            const int nf = m->c_e[cat] - m->c_s[cat] + 1;
            cats.push_back(Mp32Cat{m->q + (size_t)(m->c_s[cat] - 1) * g.ntot, nf,
                                   zeros.data() + (m->c_s[cat] - 1),
                                   zeros.data() + (m->c_s[cat] - 1), 0});
        }
        if (add) DCM_CHECK(add_profile_dev(c, m->th, m->thetan(), 1.0));
        if (!cats.empty())
            DCM_CHECK(mpdata32_batch(c, (float)m->dt, cur[0], cur[1], cur[2], m->shadow,
                                     cats.data(), (int)cats.size(), m->kord, m->monotone, 0));
        if (add) {
            hipLaunchKernelGGL(k_bc_roundtrip, dim3(1), dim3(64), 0, c->stream,
                               (const double *)(c->dscal + 16), lb, ub, m->bc_inexact);
            DCM_LAUNCH_CHECK();
            DCM_CHECK(add_profile_dev(c, m->th, m->thetan(), -1.0));
        }
    } else if (m->advection_scheme != 0) {
        // :548-551 mpdata_3d(th, t1, t2); :552-556 mpdata_3d_add: theta +
        // thetan advected with thbase, thtop as its bcs
        const bool add = m->advection_scheme == 2;
        double lb = add ? m->thbase : m->t1, ub = add ? m->thtop : m->t2;
        if (add) DCM_CHECK(add_profile_dev(c, m->th, m->thetan(), 1.0));
        st = mpdata_core<double>(c, m->dt, cur[0], cur[1], cur[2], m->th, 1, &lb, &ub, m->kord,
                                 m->monotone, 0, 0);
        if (st != DCM_OK) return st;
        if (add) {
            // t1 = t2 = 0 come back from (0 - m) + m exactly; thbase and
            // thtop only when the round trip is exact, which is checked
            hipLaunchKernelGGL(k_bc_roundtrip, dim3(1), dim3(64), 0, c->stream,
                               (const double *)(c->dscal + 16), lb, ub, m->bc_inexact);
            DCM_LAUNCH_CHECK();
            DCM_CHECK(add_profile_dev(c, m->th, m->thetan(), -1.0));
        }
    }
    if (!mixed && m->nq > 0) {  // :571-580 with q1 = q2 = 0
        for (int cat = 0; cat < m->ncat; ++cat) {
            const int nf = m->c_e[cat] - m->c_s[cat] + 1;
            double *f = m->q + (size_t)(m->c_s[cat] - 1) * g.ntot;
            const double *z = zeros.data() + (m->c_s[cat] - 1);
            st = mpdata_core<double>(c, m->dt, cur[0], cur[1], cur[2], f, nf, z, z, m->kord,
                                     m->monotone, 0, 1);
            if (st != DCM_OK) return st;
        }
    }
    return DCM_OK;
}

// The momentum update (driver_code.f90:592-594): the new level from the
// previous one, the sources and the pressure; it reads nothing the
// advection writes.
// advance_momentum (dynamics.f90:1053-1094) and the time filter
// (driver_code.f90:618-626) in one pass; the filter only reads the current
// level, which the advection reads too
static int model_momentum(dcm_model *m) {
    double *cur[3], *prev[3], *fut[3];
    m->levels(cur, prev, fut);
    return momentum_filter_dev(m->c, 2.0 * m->dt, fut, prev, cur, m->su, m->sv, m->sw, m->p);
}

// The scalar update of the sources options (advance_scalar_fields_3d,
// driver_code.f90:601-603), after the advection.
static int model_scalars(dcm_model *m) {
    if (!m->phys_on) return DCM_OK;
    dcm_ctx *c = m->c;
    const long long n = c->geo.ntot;
    hipLaunchKernelGGL(k_axpy_dt, dim3(2048), dim3(kBlock), 0, c->stream, m->th, m->sth, m->dt,
                       n);
    if (m->sq && m->nq > 0)
        hipLaunchKernelGGL(k_axpy_dt, dim3(2048), dim3(kBlock), 0, c->stream, m->q, m->sq,
                           m->dt, n * m->nq);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

// The velocities' end of the step after the momentum update and the filter:
// the rotation and the halos of the new current level (driver_code.f90:
// 634-674), which the advection does not read.
static int model_filter_rotate(dcm_model *m, dcm_step_stats *stats) {
    dcm_ctx *c = m->c;
    const Geo &g = c->geo;
    double *cur[3], *prev[3], *fut[3];
    m->n += 1;  // :634-646
    m->levels(cur, prev, fut);
    {  // exchange_full of u, v and w (:651-656), one launch
        FillSet<double> fs{};
        for (int d = 0; d < 3; ++d) fs.a[d] = cur[d];
        DCM_CHECK(fill_set<double>(c, fs, 3, FILL_FULL, 0.0, 0.0));
    }
    hipLaunchKernelGGL(k_momentum_rows, dim3(256), dim3(kBlock), 0, c->stream, cur[0], cur[1],
                       cur[2], g);  // :663-674
    DCM_LAUNCH_CHECK();
    if (stats) stats->steps += 1;
    return DCM_OK;
}

static int model_post_solve(dcm_model *m, dcm_step_stats *stats) {
    DCM_CHECK(model_advect(m));
    DCM_CHECK(model_momentum(m));
    DCM_CHECK(model_scalars(m));
    return model_filter_rotate(m, stats);
}

// advection_scheme 2: the reference leaves thbase and thtop as (bc - m) + m
// after every mpdata_3d call (INTENT(INOUT) scalars, advection_3d.F90:
// 278-280, 711-713); the model keeps them fixed, the same values whenever
// the round trip is exact.  For potential temperatures it always is: m <= bc
// <= 2m, so bc - m is exact (Sterbenz) and adding m back restores bc.  A
// step where it was not is reported here as an error, not left to drift.
static int model_check_bcs(dcm_model *m) {
    if (!m->bc_inexact) return DCM_OK;
    int bad = 0;
    DCM_HIP(hipMemcpy(&bad, m->bc_inexact, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) {
        set_error("dcm_model_step: advection_scheme 2 changed thbase/thtop by rounding "
                  "((bc - m) + m != bc); this model keeps them fixed");
        return DCM_BAD_ARG;
    }
    return DCM_OK;
}

extern "C++" {
namespace dcm {
// Run one phase of the step on `stream`: replay its graph when one was
// captured for `key`, else capture it (the host code of the phase runs
// once, at capture) and launch it.  `host_replay` repeats the phase's
// host-side bookkeeping on a replay.
template <typename Capture, typename Replay>
static int run_phase(hipStream_t stream, PhaseGraph *slots, const std::vector<int> &key,
                     Capture &&capture, Replay &&host_replay) {
    for (int r = 0; r < 3; ++r)
        if (slots[r].exec && slots[r].key == key) {
            host_replay();
            DCM_HIP(hipGraphLaunch(slots[r].exec, stream));
            return DCM_OK;
        }
    PhaseGraph *slot = &slots[0];
    for (int r = 0; r < 3; ++r)
        if (!slots[r].exec) {
            slot = &slots[r];
            break;
        }
    if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
    slot->exec = nullptr;
    slot->key.clear();
    DCM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
    const int st = capture();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(stream, &graph);
    if (st != DCM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (ec != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        return DCM_HIP_ERROR;
    }
    const hipError_t ei = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        slot->exec = nullptr;
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        return DCM_HIP_ERROR;
    }
    slot->key = key;
    DCM_HIP(hipGraphLaunch(slot->exec, stream));
    return DCM_OK;
}

}  // namespace dcm

// what the kernels of a captured phase and their arguments depend on
static std::vector<int> phase_key(const dcm_model *m) {
    return {m->n % 2 + 3 * (m->precision == DCM_MIXED) + 6 * (m->c->seq_sums ? 1 : 0) +
            12 * m->c->mpdata_variant + 24 * m->gen};
}

// Run phase 0 (before the solve) or 1 (after it) of a model step: eagerly
// the first time its key is seen, from a graph afterwards (not while kernels
// are being timed one by one).
template <typename Body, typename Replay>
static int model_phase(dcm_model *m, int phase, Body &&body, Replay &&host_replay,
                       int extra = 0) {
    std::vector<int> key = phase_key(m);
    key.push_back(extra);
    std::vector<std::vector<int>> &seen = m->eager[phase];
    const bool known = std::find(seen.begin(), seen.end(), key) != seen.end();
    PhaseGraph *slots = phase == 0 ? m->pre : (phase == 1 ? m->post : (phase == 2 ? m->adv : m->mom));
    if (m->use_graphs && !m->c->profiling && known)
        return run_phase(m->c->stream, slots, key, body, host_replay);
    DCM_CHECK(body());
    if (!known) {
        if (seen.size() >= 48) seen.erase(seen.begin());
        seen.push_back(key);
    }
    return DCM_OK;
}

}  // extern "C++"

// The advection runs beside the solve when it touches no scratch the solve
// uses: the mixed path's batched FP32 MPDATA (own scratch and partial sums)
// with advection_scheme 1 or 2, and not while kernels are timed one by one.
static bool advect_overlaps(const dcm_model *m) {
    const dcm_ctx *c = m->c;
    return m->adv_stream && m->precision == DCM_MIXED && m->advection_scheme != 0 &&
           !c->profiling && mpdata32_fast(c, m->kord);
}

static int model_step_once(dcm_model *m, dcm_step_stats *stats) {
    dcm_ctx *c = m->c;
    DCM_CHECK(model_phase(m, 0, [&] { return model_pre_solve(m); }, [] {}));
    const bool overlap = advect_overlaps(m);
    if (overlap) {
        DCM_HIP(hipEventRecord(m->adv_fork, c->stream));
        DCM_HIP(hipStreamWaitEvent(m->adv_stream, m->adv_fork, 0));
        hipStream_t main = c->stream;
        c->stream = m->adv_stream;
        const int sa = model_phase(m, 2, [&] { return model_advect(m); }, [] {});
        c->stream = main;
        if (sa != DCM_OK) return sa;
        DCM_HIP(hipEventRecord(m->adv_join, m->adv_stream));
    }
    SolverOut so;
    int st;
    if (m->precision == DCM_MIXED)
        st = bicgstab_mixed(c, m->sw, m->p, m->psrc, m->ptol, 9999, m->inner_rtol, &so);
    else {
        c->first_solve = false;
        st = bicgstab64(c, m->sw, m->p, m->psrc, m->ptol, 9999, &so);  // :496-500
    }
    if (st != DCM_OK) return st;
    double true_resid = 0.0;
    if (m->precision == DCM_MIXED)
        true_resid = so.err;  // the refinement loop's exit test is the true residual
    else if (m->check_residual)
        DCM_CHECK(dcm_true_residual(c, m->p, m->psrc, &true_resid));
    add_stats(stats, so, true_resid);
    // the host part of a phase with the rotation, on a replay: the level
    // rotation and the step count
    auto rotate = [&] {
        m->n += 1;
        if (stats) stats->steps += 1;
    };
    if (overlap) {
        // the momentum update, the filter and the rotation need only the
        // solve; the scalar update (and the next step) wait for the advection
        DCM_CHECK(model_phase(
            m, 3,
            [&] {
                DCM_CHECK(model_momentum(m));
                return model_filter_rotate(m, stats);
            },
            rotate));
        DCM_HIP(hipStreamWaitEvent(c->stream, m->adv_join, 0));
        if (!m->phys_on) return DCM_OK;
        return model_phase(m, 1, [&] { return model_scalars(m); }, [] {}, 1);
    }
    return model_phase(m, 1, [&] { return model_post_solve(m, stats); }, rotate);
}

int dcm_model_step(dcm_model *m, int nsteps, dcm_step_stats *stats) {
    if (!m || nsteps < 0) return DCM_BAD_ARG;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    for (int s = 0; s < nsteps; ++s) DCM_CHECK(model_step_once(m, stats));
    DCM_HIP(hipStreamSynchronize(m->c->stream));
    DCM_CHECK(prof_collect(m->c));
    DCM_CHECK(model_check_bcs(m));
    return DCM_OK;
}

int dcm_model_profile(dcm_model *m, int enable) {
    if (!m) return DCM_BAD_ARG;
    DCM_CHECK(prof_collect(m->c));
    m->c->profiling = enable != 0;
    m->c->timers.clear();
    return DCM_OK;
}

int dcm_model_profile_filter(dcm_model *m, const char *name) {
    if (!m) return DCM_BAD_ARG;
    m->c->prof_only = name ? name : "";
    return DCM_OK;
}

int dcm_model_set_inner_rtol(dcm_model *m, double rtol) {
    if (!m || !(rtol > 0.0 && rtol < 1.0)) return DCM_BAD_ARG;
    m->inner_rtol = rtol;
    return DCM_OK;
}

int dcm_model_set_graphs(dcm_model *m, int on) {
    if (!m) return DCM_BAD_ARG;
    m->use_graphs = on != 0;
    m->c->graphs = on != 0;
    return DCM_OK;
}

int dcm_model_set_check_residual(dcm_model *m, int on) {
    if (!m) return DCM_BAD_ARG;
    m->check_residual = on != 0;
    return DCM_OK;
}

// ---------------------------------------------------------------------------
// ensembles: independent models of uneven size stepped together.  Every
// member's kernels go to one stream; the pressure solves of all members run
// in one batched BiCGSTAB (each launch serves every member that has not
// converged), so the iteration loop costs one set of launches per iteration
// for the whole ensemble.  The stages before and after the solve (a few
// dozen small kernels per member) are captured into HIP graphs, one per
// buffer arrangement of the time levels (n mod 2), and replayed from then on.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace dcm {
// side streams of an ensemble (the box's hardware queues per process)
constexpr int kEnsembleLanes = 4;
}  // namespace dcm

struct dcm_ensemble {
    std::vector<dcm_model *> mem;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> saved;  // the members' own streams
    int precision = DCM_FP64;
    double ptol = 1e-10;
    int march_planes = 0;
    bool use_graphs = true;
    // the member keys each phase has already run with, without capture: a
    // phase is captured only for a key seen before, so every scratch buffer
    // its kernels use exists by then (no allocation or memset is recorded)
    std::vector<std::vector<int>> eager[2];
    dcm::PhaseGraph pre[3], post[3];
    // the stages around the solve run the members on side streams (members
    // balanced by cells), forked from and joined back into `stream`
    std::vector<hipStream_t> side;
    std::vector<int> lane;  // member -> side stream
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
};

namespace dcm {

static void drop_graphs(dcm_ensemble *e) {
    for (int r = 0; r < 3; ++r) {
        for (PhaseGraph *g : {&e->pre[r], &e->post[r]}) {
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
            g->exec = nullptr;
            g->key.clear();
        }
    }
}

// Run `body` on every member with the members spread over the side
// streams: the side streams wait for the work already queued on the
// ensemble stream, and the ensemble stream waits for all of them after.
template <typename Body> static int fan_out(dcm_ensemble *e, Body &&body) {
    const int n = (int)e->mem.size();
    DCM_HIP(hipEventRecord(e->fork, e->stream));
    for (hipStream_t s : e->side) DCM_HIP(hipStreamWaitEvent(s, e->fork, 0));
    int st = DCM_OK;
    for (int q = 0; q < n && st == DCM_OK; ++q) {
        dcm_model *m = e->mem[q];
        m->c->stream = e->side[e->lane[q]];
        st = body(m);
        m->c->stream = e->stream;
    }
    for (size_t k = 0; k < e->side.size(); ++k) {
        DCM_HIP(hipEventRecord(e->join[k], e->side[k]));
        DCM_HIP(hipStreamWaitEvent(e->stream, e->join[k], 0));
    }
    return st;
}

static int ensemble_solve(dcm_ensemble *e, std::vector<SolverOut> &outs) {
    const int n = (int)e->mem.size();
    std::vector<SolveTask> tasks(n);
    for (int q = 0; q < n; ++q) {
        dcm_model *m = e->mem[q];
        tasks[q] = SolveTask{m->c, m->sw, m->p, m->psrc};
        m->c->first_solve = false;
    }
    std::vector<int> st(n, DCM_OK);
    outs.assign(n, SolverOut());
    dcm_ctx *c0 = e->mem[0]->c;
    int r;
    int n_mg = 0;
    for (int q = 0; q < n; ++q) n_mg += (e->mem[q]->c->mg_on && e->mem[q]->c->mg) ? 1 : 0;
    if (e->precision == DCM_MIXED && n_mg > 0 && n_mg < n) {
        // the batched solve has no V-cycle: a member with the multigrid
        // setting would silently evolve differently from its solo run
        set_error("dcm_ensemble_step: either every member or none uses the multigrid "
                  "preconditioner");
        return DCM_BAD_ARG;
    }
    const bool all_mg = e->precision == DCM_MIXED && n_mg == n;
    if (all_mg) {
        // every member with the multigrid preconditioner: its own solve, as
        // when it is stepped alone (the batched iteration has no V-cycle).
        // With graphs on, the members' solves run concurrently, each on its
        // side stream (same kernels per member, so the same bits);
        // otherwise one after another on the ensemble's stream.
        bool conc = true;
        for (int q = 0; q < n; ++q)
            conc = conc && e->mem[q]->c->graphs && !e->mem[q]->c->profiling;
        if (conc && !e->side.empty()) {
            std::vector<hipStream_t> streams(n);
            std::vector<double> rtol(n);
            for (int q = 0; q < n; ++q) {
                streams[q] = e->side[e->lane[q]];
                rtol[q] = e->mem[q]->inner_rtol;
            }
            DCM_HIP(hipEventRecord(e->fork, e->stream));
            for (hipStream_t s : e->side) DCM_HIP(hipStreamWaitEvent(s, e->fork, 0));
            const int rq = bicgstab_mixed_concurrent(tasks.data(), n, streams.data(), e->ptol,
                                                     9999, rtol.data(), outs.data(), st.data());
            for (size_t k = 0; k < e->side.size(); ++k) {
                DCM_HIP(hipEventRecord(e->join[k], e->side[k]));
                DCM_HIP(hipStreamWaitEvent(e->stream, e->join[k], 0));
            }
            return rq;
        }
        int worst = DCM_OK;
        for (int q = 0; q < n; ++q) {
            const int rq = bicgstab_mixed_batch(e->mem[q]->c, &tasks[q], 1, e->ptol, 9999,
                                                e->mem[q]->inner_rtol, &outs[q], &st[q]);
            if (rq != DCM_OK && worst == DCM_OK) worst = rq;
        }
        return worst;
    }
    if (e->precision == DCM_MIXED)
        r = bicgstab_mixed_batch(c0, tasks.data(), n, e->ptol, 9999, e->mem[0]->inner_rtol,
                                 outs.data(), st.data());
    else
        r = bicgstab64_batch(c0, tasks.data(), n, e->ptol, 9999, outs.data(), st.data());
    return r;
}

// Time a few batched iterations of the current systems at each candidate
// plane count and keep the fastest.  Only solver scratch is touched: the
// iterate and residual vectors are rebuilt from p and psrc by every solve.
static int ensemble_autotune(dcm_ensemble *e, int *best_out) {
    static const int cand[] = {4, 8, 16};
    const int n = (int)e->mem.size();
    dcm_ctx *c0 = e->mem[0]->c;
    hipEvent_t a, b;
    DCM_HIP(hipEventCreate(&a));
    DCM_HIP(hipEventCreate(&b));
    float best_ms = 1e30f;
    int best = 8;
    std::vector<SolveTask> tasks(n);
    for (int q = 0; q < n; ++q)
        tasks[q] = SolveTask{e->mem[q]->c, e->mem[q]->sw, e->mem[q]->p, e->mem[q]->psrc};
    for (int ci = 0; ci < 3; ++ci) {
        for (int q = 0; q < n; ++q) e->mem[q]->c->march_planes = cand[ci];
        std::vector<SolverOut> o(n);
        std::vector<int> st(n);
        DCM_HIP(hipEventRecord(a, e->stream));
        DCM_CHECK(bicgstab_probe(c0, tasks.data(), n, 8, o.data(), st.data()));
        DCM_HIP(hipEventRecord(b, e->stream));
        DCM_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        DCM_HIP(hipEventElapsedTime(&ms, a, b));
        if (ms < best_ms) {
            best_ms = ms;
            best = cand[ci];
        }
    }
    for (int q = 0; q < n; ++q) e->mem[q]->c->march_planes = best;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->march_planes = best;
    if (best_out) *best_out = best;
    return DCM_OK;
}

}  // namespace dcm

extern "C" {

int dcm_ensemble_create(int n, dcm_model *const *models, dcm_ensemble **out) {
    if (!out || n < 1 || !models) return DCM_BAD_ARG;
    for (int q = 0; q < n; ++q) {
        if (!models[q]) return DCM_BAD_ARG;
        for (int u = 0; u < q; ++u)
            if (models[u] == models[q] || models[u]->c == models[q]->c) {
                set_error("dcm_ensemble_create: members need their own models and contexts");
                return DCM_BAD_ARG;
            }
        if (models[q]->precision != models[0]->precision || models[q]->ptol != models[0]->ptol) {
            set_error("dcm_ensemble_create: members must share precision and ptol");
            return DCM_BAD_ARG;
        }
    }
    dcm_ensemble *e = new dcm_ensemble();
    e->precision = models[0]->precision;
    e->ptol = models[0]->ptol;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        set_error("dcm_ensemble_create: hipStreamCreate failed");
        return DCM_HIP_ERROR;
    }
    for (int q = 0; q < n; ++q) {
        dcm_ctx *c = models[q]->c;
        (void)hipStreamSynchronize(c->stream);
        e->saved.push_back(c->stream);
        c->stream = e->stream;
        c->borrowed_stream = true;
        e->mem.push_back(models[q]);
    }
    // side streams for the stages around the solve; members go to the
    // least-loaded stream, largest first
    const int nside = n < kEnsembleLanes ? n : kEnsembleLanes;
    bool ok = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < nside && ok; ++k) {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
        if (s) e->side.push_back(s);
        if (ev) e->join.push_back(ev);
    }
    if (!ok) {
        (void)dcm_ensemble_destroy(e);
        set_error("dcm_ensemble_create: stream or event creation failed");
        return DCM_HIP_ERROR;
    }
    std::vector<int> order(n);
    for (int q = 0; q < n; ++q) order[q] = q;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        return models[a]->c->geo.ntot > models[b]->c->geo.ntot;
    });
    std::vector<long long> load(nside, 0);
    e->lane.assign(n, 0);
    for (int q : order) {
        int k = 0;
        for (int u = 1; u < nside; ++u)
            if (load[u] < load[k]) k = u;
        e->lane[q] = k;
        load[k] += models[q]->c->geo.ntot;
    }
    *out = e;
    return DCM_OK;
}

int dcm_ensemble_destroy(dcm_ensemble *e) {
    if (!e) return DCM_BAD_ARG;
    (void)hipStreamSynchronize(e->stream);
    for (hipStream_t s : e->side) (void)hipStreamSynchronize(s);
    drop_graphs(e);
    for (size_t q = 0; q < e->mem.size(); ++q) {
        dcm_ctx *c = e->mem[q]->c;
        c->stream = e->saved[q];
        c->borrowed_stream = false;
    }
    for (hipStream_t s : e->side) (void)hipStreamDestroy(s);
    for (hipEvent_t ev : e->join) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    (void)hipStreamDestroy(e->stream);
    delete e;
    return DCM_OK;
}

int dcm_ensemble_autotune(dcm_ensemble *e, int *march_planes) {
    if (!e) return DCM_BAD_ARG;
    return ensemble_autotune(e, march_planes);
}

int dcm_ensemble_set_graphs(dcm_ensemble *e, int enable) {
    if (!e) return DCM_BAD_ARG;
    DCM_HIP(hipStreamSynchronize(e->stream));
    e->use_graphs = enable != 0;
    if (!e->use_graphs) drop_graphs(e);
    return DCM_OK;
}

int dcm_ensemble_step(dcm_ensemble *e, int nsteps, dcm_step_stats *stats) {
    if (!e || nsteps < 0) return DCM_BAD_ARG;
    const int n = (int)e->mem.size();
    if (stats) std::memset(stats, 0, sizeof(dcm_step_stats) * n);
    std::vector<SolverOut> outs;
    std::vector<int> key(n);
    auto pre = [&]() -> int {
        return fan_out(e, [](dcm_model *m) { return model_pre_solve(m); });
    };
    auto post = [&]() -> int {
        return fan_out(e, [](dcm_model *m) { return model_post_solve(m, nullptr); });
    };
    auto advance_levels = [&]() {
        for (int q = 0; q < n; ++q) e->mem[q]->n += 1;  // model_post_solve's host part
    };
    auto seen = [&](int phase) {
        for (const std::vector<int> &k : e->eager[phase])
            if (k == key) return true;
        return false;
    };
    auto remember = [&](int phase) {
        std::vector<std::vector<int>> &v = e->eager[phase];
        if (seen(phase)) return;
        if (v.size() >= 48) v.erase(v.begin());
        v.push_back(key);
    };
    for (int s = 0; s < nsteps; ++s) {
        bool graphs = e->use_graphs;
        for (int q = 0; q < n; ++q) {
            // what the captured kernels and their arguments depend on
            const dcm_model *m = e->mem[q];
            key[q] = m->n % 2 + 3 * (m->precision == DCM_MIXED) + 6 * (m->c->seq_sums ? 1 : 0) +
                     12 * m->c->mpdata_variant + 24 * m->gen;
            graphs = graphs && !m->c->profiling;
        }
        if (graphs && seen(0)) {
            DCM_CHECK(run_phase(e->stream, e->pre, key, pre, [] {}));
        } else {
            DCM_CHECK(pre());
            remember(0);
        }
        DCM_CHECK(ensemble_solve(e, outs));
        for (int q = 0; q < n; ++q) {
            dcm_model *m = e->mem[q];
            double true_resid = 0.0;
            if (m->precision == DCM_MIXED)
                true_resid = outs[q].err;
            else if (m->check_residual)
                DCM_CHECK(dcm_true_residual(m->c, m->p, m->psrc, &true_resid));
            add_stats(stats ? stats + q : nullptr, outs[q], true_resid);
            if (stats) stats[q].steps += 1;
        }
        if (graphs && seen(1)) {
            DCM_CHECK(run_phase(e->stream, e->post, key, post, advance_levels));
        } else {
            DCM_CHECK(post());
            remember(1);
        }
    }
    DCM_HIP(hipStreamSynchronize(e->stream));
    for (dcm_model *m : e->mem) DCM_CHECK(model_check_bcs(m));
    return DCM_OK;
}

int dcm_stream_copy(dcm_ctx *c, double *dst, const double *src, size_t n, double *ms) {
    if (!c || !dst || !src) return DCM_BAD_ARG;
    hipEvent_t a = nullptr, b = nullptr;
    if (ms) {
        DCM_HIP(hipEventCreate(&a));
        DCM_HIP(hipEventCreate(&b));
        DCM_HIP(hipEventRecord(a, c->stream));
    }
    const unsigned nb = (unsigned)((n / 2 + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_stream_copy, dim3(nb > 0 ? nb : 1), dim3(kBlock), 0, c->stream, dst,
                       src, (long long)n);
    DCM_LAUNCH_CHECK();
    if (ms) {
        DCM_HIP(hipEventRecord(b, c->stream));
        DCM_HIP(hipEventSynchronize(b));
        float f = 0.f;
        DCM_HIP(hipEventElapsedTime(&f, a, b));
        *ms = f;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    return DCM_OK;
}

int dcm_model_kernel_time(dcm_model *m, const char *name, double *avg_ms, int *launches) {
    if (!m || !name) return DCM_BAD_ARG;
    DCM_CHECK(prof_collect(m->c));
    auto it = m->c->timers.find(name);
    if (it == m->c->timers.end()) return DCM_BAD_ARG;
    if (avg_ms) *avg_ms = it->second.launches ? it->second.total_ms / it->second.launches : 0.0;
    if (launches) *launches = it->second.launches;
    return DCM_OK;
}

}  // extern "C"
