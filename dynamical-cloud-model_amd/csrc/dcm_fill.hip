// dcm_fill.hip -- on-device boundary fills replacing the MPI halo swaps of
// mpi_module.f90 for a domain held by one GPU, and the generic reductions.
//
// With one rank the reference's exchanges reduce to these rules (halo 1):
//   * x,y faces (k in 1..kp): periodic copy of the opposite interior column
//     (mpi_module.f90:857-870, 1424-1437);
//   * z planes k=0 / k=kp+1: lbc / ubc over the whole plane, lateral halo
//     included (mpi_module.f90:874-883; exchange_along_z :651-660);
//   * exchange_full only: vertical edges periodic, and every edge or corner on
//     a z plane takes the z-wrapped interior value, because its receive comes
//     from MPI_PROC_NULL and leaves the locally pre-filled send buffer
//     (mpi_module.f90:965-1136, 1180-1196, 1244-1263).
// Every halo cell therefore has a closed-form source, so one kernel writes all
// of them at once, reading the interior only: no ordering between the
// reference's face, edge and corner stages is needed.
#include "dcm_device.h"

namespace dcm {

template <typename T>
__global__ void k_fill(FillSet<T> set, Geo g, int kind, double lbc, double ubc,
                       const double *shift, int mode, const int *skip_if) {
    if (skip_if && *skip_if) return;
    T *a = set.a[blockIdx.y];  // one array per grid row
    // 32-bit index arithmetic (unsigned divisions): the halo shell of any
    // grid whose padded planes have fewer than 2^30 cells
    const int P = g.nj * g.ni;
    const int R = 2 * g.ni + 2 * g.jp;
    const int nz = 2 * P;
    const int n = nz + g.kp * R;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int k, j, i;
        bool zplane;
        if (t < nz) {
            const int plane = t >= P ? 1 : 0;
            const int rem = t - plane * P;
            i = (int)((unsigned)rem / (unsigned)g.nj);
            j = rem - i * g.nj;
            k = plane ? g.kp + 1 : 0;
            zplane = true;
        } else {
            const int tt = t - nz;
            const int r = (int)((unsigned)tt / (unsigned)g.kp);
            k = tt - r * g.kp + 1;
            if (r < g.ni) {
                j = 0;
                i = r;
            } else if (r < 2 * g.ni) {
                j = g.jp + 1;
                i = r - g.ni;
            } else {
                const int r2 = r - 2 * g.ni;
                if (r2 < g.jp) {
                    i = 0;
                    j = r2 + 1;
                } else {
                    i = g.ip + 1;
                    j = r2 - g.jp + 1;
                }
            }
            zplane = false;
        }
        const bool jh = (j == 0 || j == g.jp + 1), ih = (i == 0 || i == g.ip + 1);
        const bool lateral_interior = !jh && !ih;
        bool use_bc = false, use_src = false;
        int sk = k;
        if (zplane) {
            switch (kind) {
                case FILL_FULL:
                    if (lateral_interior) use_bc = true;
                    else { use_src = true; sk = zwrap(k, g.kp); }
                    break;
                case FILL_FULL_WO:
                    if (!lateral_interior) { use_src = true; sk = zwrap(k, g.kp); }
                    break;
                case FILL_ALONG_DIM:
                case FILL_ALONG_Z:
                    use_bc = true;
                    break;
                default:
                    break;
            }
        } else {
            const bool face = jh != ih;  // exactly one lateral index in the halo
            switch (kind) {
                case FILL_FULL:
                case FILL_FULL_WO:
                    use_src = true;
                    break;
                case FILL_ALONG_DIM:
                case FILL_ALONG_DIM_WO:
                    use_src = face;
                    break;
                default:
                    break;
            }
        }
        if (use_bc) {
            double b = (k == 0) ? lbc : ubc;
            if (mode == 1) b = b - *shift;
            else if (mode == 2) b = (b - *shift) + *shift;
            a[pidx(g, k, j, i)] = (T)b;
        } else if (use_src) {
            a[pidx(g, k, j, i)] = a[pidx(g, sk, wrap(j, g.jp), wrap(i, g.ip))];
        }
    }
}

template <typename T>
int fill(dcm_ctx *c, T *a, FillKind kind, double lbc, double ubc, const double *bc_shift,
         int bc_mode, const int *skip_if) {
    const Geo &g = c->geo;
    const long long n = 2LL * g.nj * g.ni + (long long)g.kp * (2LL * g.ni + 2LL * g.jp);
    unsigned nb = blocks_for(n);
    if (nb > 4096) nb = 4096;
    int mode = bc_shift ? (bc_mode == 2 ? 2 : 1) : 0;
    int tok = prof_begin(c, "fill");
    FillSet<T> set{};
    set.a[0] = a;
    hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(kBlock), 0, c->stream, set, g, (int)kind, lbc,
                       ubc, bc_shift, mode, skip_if);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    return DCM_OK;
}

template <typename T>
int fill_set(dcm_ctx *c, const FillSet<T> &set, int count, FillKind kind, double lbc,
             double ubc, const int *skip_if) {
    const Geo &g = c->geo;
    if (count < 1 || count > kFillSetMax) {
        set_error("fill_set: 1..6 arrays");
        return DCM_BAD_ARG;
    }
    const long long n = 2LL * g.nj * g.ni + (long long)g.kp * (2LL * g.ni + 2LL * g.jp);
    unsigned nb = blocks_for(n);
    if (nb > 4096) nb = 4096;
    int tok = prof_begin(c, "fill");
    hipLaunchKernelGGL(k_fill<T>, dim3(nb, (unsigned)count), dim3(kBlock), 0, c->stream, set, g,
                       (int)kind, lbc, ubc, (const double *)nullptr, 0, skip_if);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    return DCM_OK;
}
template int fill_set<double>(dcm_ctx *, const FillSet<double> &, int, FillKind, double, double,
                              const int *);
template int fill_set<float>(dcm_ctx *, const FillSet<float> &, int, FillKind, double, double,
                             const int *);
template int fill<double>(dcm_ctx *, double *, FillKind, double, double, const double *, int,
                          const int *);
template int fill<float>(dcm_ctx *, float *, FillKind, double, double, const double *, int,
                         const int *);

// x(0,:,:)=x(1,:,:) and x(kp+1,:,:)=x(kp,:,:) over whole planes
// (mat_ax, dynamics.f90:206-211)
__global__ void k_neumann_z(double *x, Geo g) {
    const int P = g.nj * g.ni;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
        const int i = (int)((unsigned)t / (unsigned)g.nj), j = t - i * g.nj;
        x[pidx(g, 0, j, i)] = x[pidx(g, 1, j, i)];
        x[pidx(g, g.kp + 1, j, i)] = x[pidx(g, g.kp, j, i)];
    }
}

int fill_neumann_z(dcm_ctx *c, double *x) {
    const Geo &g = c->geo;
    unsigned nb = blocks_for((long long)g.nj * g.ni);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_neumann_z, dim3(nb), dim3(kBlock), 0, c->stream, x, g);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

// ---------------------------------------------------------------------------
// generic finalize kernels
// ---------------------------------------------------------------------------
__global__ void k_finalize_sums(const double *partials, int nblocks, int nsums, double *dst) {
    for (int s = 0; s < nsums; ++s) {
        double r = reduce_slot_1024(partials + (long long)s * nblocks, nblocks);
        if (threadIdx.x == 0) dst[s] = r;
    }
}

__global__ void k_finalize_min(const double *partials, int nblocks, double *dst) {
    double r = reduce_min_1024(partials, nblocks);
    if (threadIdx.x == 0) *dst = r;
}

int finalize_sums(dcm_ctx *c, const double *partials, int nblocks, int nsums, double *dst) {
    hipLaunchKernelGGL(k_finalize_sums, dim3(1), dim3(1024), 0, c->stream, partials, nblocks,
                       nsums, dst);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

int finalize_min(dcm_ctx *c, const double *partials, int nblocks, double *dst) {
    hipLaunchKernelGGL(k_finalize_min, dim3(1), dim3(1024), 0, c->stream, partials, nblocks,
                       dst);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

}  // namespace dcm

using namespace dcm;

extern "C" {

int dcm_fill_full(dcm_ctx *c, double *a, double lbc, double ubc) {
    if (!c || !a) return DCM_BAD_ARG;
    return fill<double>(c, a, FILL_FULL, lbc, ubc);
}
int dcm_fill_along_dim(dcm_ctx *c, double *a, double lbc, double ubc) {
    if (!c || !a) return DCM_BAD_ARG;
    return fill<double>(c, a, FILL_ALONG_DIM, lbc, ubc);
}
int dcm_fill_along_dim_wo(dcm_ctx *c, double *a) {
    if (!c || !a) return DCM_BAD_ARG;
    return fill<double>(c, a, FILL_ALONG_DIM_WO, 0.0, 0.0);
}
int dcm_fill_full_wo(dcm_ctx *c, double *a) {
    if (!c || !a) return DCM_BAD_ARG;
    return fill<double>(c, a, FILL_FULL_WO, 0.0, 0.0);
}
int dcm_fill_along_z(dcm_ctx *c, double *a, double lbc, double ubc) {
    if (!c || !a) return DCM_BAD_ARG;
    return fill<double>(c, a, FILL_ALONG_Z, lbc, ubc);
}

}  // extern "C"
