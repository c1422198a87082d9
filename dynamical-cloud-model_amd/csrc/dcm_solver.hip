// dcm_solver.hip -- the anelastic pressure solve (dynamics.f90:24-482) on one
// GPU: Jacobi(3)-preconditioned BiCGSTAB on the 7-point Laplacian.
//
// Operator.  set_mat_a (dynamics.f90:34-175) stores a_e..a_p as eight 3-D
// arrays; on the reference's product grid each coefficient depends on one
// index, so the matvec rebuilds them from six 1-D arrays with the same IEEE
// expressions (a_p by the same left-to-right sum).  mat_ax (:187-229) first
// fills the halo of its input: periodic x,y faces, then Neumann copies
// x(0)=x(1), x(kp+1)=x(kp).  Those fills only ever feed the stencil, so here
// the stencil reads the interior through wrapped / clamped indices and the
// solver's work vectors are compact (kp*jp*ip, no halo).
//
// Iteration (dynamics.f90:390-430), with every whole-array statement fused
// into the stencil pass that produces or consumes it:
//   jac_first  t = r - bet*(omg*v1); s1 = s0 + (t - A s0)/a_p, s0 = t/a_p
//   jac        s2 = s1 + (t - A s1)/a_p
//   jac_p      s3 = s2 + (t - A s2)/a_p;  p = s3 + bet*p
//   matvec_dot v1 = A p;                   <cr,v1>  -> alf = rho/nrm
//   jac_first  s = r - alf*v1;  cs1 ...
//   jac, jac   cs2, cs
//   matvec_dot t = A cs;                   <t,t>, <t,s> -> omg
//   update     x = x + alf*p + omg*cs; r = s - omg*t; <r,r>, <r,cr>
// Each value is computed by the reference's expression in its operand order,
// so only the dot products' summation order differs from the reference (a
// fixed two-level tree here, deterministic run to run).  The recurrence
// scalars live on the device; the host reads them once per iteration.
//
// Mixed precision (dcm_bicgstab_mixed): the same kernels with FP32 vectors
// and FP64 dot products solve A d = r for the FP64 residual r = b - A x; the
// outer loop adds d to x in FP64 and recomputes the true FP64 residual until
// ||b-Ax||/max(||b||,0.01) < tol (iterative refinement).  The operator is
// singular (periodic + Neumann), so each inner right-hand side has its mean
// removed.
#include <cfloat>
#include <cmath>

#include "dcm_device.h"

namespace dcm {

struct CGeo {
    int ip, jp, kp;
    long long n;
};

__device__ __forceinline__ long long cidx(const CGeo &g, int k, int j, int i) {
    return ((long long)(i - 1) * g.jp + (j - 1)) * g.kp + (k - 1);
}

template <typename R> struct Coef;
template <> struct Coef<double> {
    const double *ae, *aw, *an, *as, *au, *ad;
    __host__ __device__ Coef(const Metrics &m, const Metrics32 &)
        : ae(m.ae), aw(m.aw), an(m.an), as(m.as), au(m.au), ad(m.ad) {}
};
template <> struct Coef<float> {
    const float *ae, *aw, *an, *as, *au, *ad;
    __host__ __device__ Coef(const Metrics &, const Metrics32 &m)
        : ae(m.ae), aw(m.aw), an(m.an), as(m.as), au(m.au), ad(m.ad) {}
};

struct CoefPack {
    Metrics m64;
    Metrics32 m32;
};

// a_p of set_mat_a (dynamics.f90:90-92)
template <typename R>
__device__ __forceinline__ R coef_ap(const Coef<R> &C, int k, int j, int i) {
    return (((((-C.ae[i]) - C.aw[i]) - C.an[j]) - C.as[j]) - C.au[k]) - C.ad[k];
}

// neighbour indices of (k,j,i): periodic in i,j; Neumann copy in k
struct Nbr {
    int iw, ie, js, jn, kd, ku;
};
__device__ __forceinline__ Nbr nbrs(int k, int j, int i, int kp, int jp, int ip) {
    Nbr n;
    n.iw = i == 1 ? ip : i - 1;
    n.ie = i == ip ? 1 : i + 1;
    n.js = j == 1 ? jp : j - 1;
    n.jn = j == jp ? 1 : j + 1;
    n.kd = k == 1 ? 1 : k - 1;
    n.ku = k == kp ? kp : k + 1;
    return n;
}

// the stencil sum of mat_ax (dynamics.f90:216-224), coefficient at (k,j,i)
template <typename R>
__device__ __forceinline__ R stencil(const Coef<R> &C, int k, int j, int i, R ap, R xw, R xe,
                                     R xs, R xn, R xd, R xu, R xc) {
    return ((((((C.aw[i] * xw + C.ae[i] * xe) + C.as[j] * xs) + C.an[j] * xn) + C.ad[k] * xd) +
             C.au[k] * xu) +
            ap * xc) +
           R(0);
}

#define SOLVER_PROLOGUE                                              \
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;         \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;         \
    const int i = blockIdx.z + 1;                                    \
    const bool act = k <= g.kp && j <= g.jp;

// ---------------------------------------------------------------------------
// initial residual (dynamics.f90:370-381): r = b - A x, cr = r, p = v1 = 0;
// partial sums of <b,b> and <r,r>.  x and b are padded FP64 arrays.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_resid0(const double *x, const double *b, double *r,
                                                double *cr, double *p, double *v1, Geo pg,
                                                CGeo g, CoefPack cp, double *partials,
                                                long long stride) {
    SOLVER_PROLOGUE
    const Coef<double> C(cp.m64, cp.m32);
    double s[2] = {0.0, 0.0};
    if (act) {
        const Nbr n = nbrs(k, j, i, g.kp, g.jp, g.ip);
        const double ap = coef_ap(C, k, j, i);
        const double ax =
            stencil<double>(C, k, j, i, ap, x[pidx(pg, k, j, n.iw)], x[pidx(pg, k, j, n.ie)],
                            x[pidx(pg, k, n.js, i)], x[pidx(pg, k, n.jn, i)],
                            x[pidx(pg, n.kd, j, i)], x[pidx(pg, n.ku, j, i)], x[pidx(pg, k, j, i)]);
        const double bb = b[pidx(pg, k, j, i)];
        const double rr = bb - ax;
        const long long c = cidx(g, k, j, i);
        r[c] = rr;
        cr[c] = rr;
        p[c] = 0.0;
        v1[c] = 0.0;
        s[0] = bb * bb;
        s[1] = rr * rr;
    }
    block_sum_store<2>(s, partials, stride);
}

// FP64 true residual with an FP32 copy for the inner solve:
// r32 = float(b - A x); partial sums of <r,r>, sum(r) (and <b,b> if asked)
__global__ __launch_bounds__(256) void k_resid_true(const double *x, const double *b,
                                                    float *r32, Geo pg, CGeo g, CoefPack cp,
                                                    double *partials, long long stride) {
    SOLVER_PROLOGUE
    const Coef<double> C(cp.m64, cp.m32);
    double s[3] = {0.0, 0.0, 0.0};
    if (act) {
        const Nbr n = nbrs(k, j, i, g.kp, g.jp, g.ip);
        const double ap = coef_ap(C, k, j, i);
        const double ax =
            stencil<double>(C, k, j, i, ap, x[pidx(pg, k, j, n.iw)], x[pidx(pg, k, j, n.ie)],
                            x[pidx(pg, k, n.js, i)], x[pidx(pg, k, n.jn, i)],
                            x[pidx(pg, n.kd, j, i)], x[pidx(pg, n.ku, j, i)], x[pidx(pg, k, j, i)]);
        const double bb = b[pidx(pg, k, j, i)];
        const double rr = bb - ax;
        if (r32) r32[cidx(g, k, j, i)] = (float)rr;
        s[0] = rr * rr;
        s[1] = rr;
        s[2] = bb * bb;
    }
    block_sum_store<3>(s, partials, stride);
}

// ---------------------------------------------------------------------------
// preconditioner sweeps (dynamics.f90:241-270)
// ---------------------------------------------------------------------------
// MODE 0: t = r - bet*(omg*v1)     (dynamics.f90:394)
// MODE 1: t = r - alf*v1           (dynamics.f90:405, the "s" vector)
template <typename R, int MODE>
__device__ __forceinline__ R tval(const R *r, const R *v1, long long q, R a, R b) {
    if (MODE == 0) return r[q] - a * (b * v1[q]);
    return r[q] - a * v1[q];
}

template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_jac_first(const R *r, const R *v1, R *t_out, R *s_out,
                                                   CGeo g, CoefPack cp,
                                                   const SolverScalars *sc) {
    if (sc->done || sc->breakdown) return;
    SOLVER_PROLOGUE
    if (!act) return;
    const Coef<R> C(cp.m64, cp.m32);
    const R a = MODE == 0 ? (R)sc->bet : (R)sc->alf;
    const R bo = (R)sc->omg;
    const Nbr n = nbrs(k, j, i, g.kp, g.jp, g.ip);
    const long long c = cidx(g, k, j, i);
    // s0 = t/a_p at the centre and the six (wrapped/clamped) neighbours
    auto s0 = [&](int kk, int jj, int ii) -> R {
        const long long q = cidx(g, kk, jj, ii);
        return tval<R, MODE>(r, v1, q, a, bo) / coef_ap(C, kk, jj, ii);
    };
    const R tc = tval<R, MODE>(r, v1, c, a, bo);
    const R ap = coef_ap(C, k, j, i);
    const R sc0 = tc / ap;
    const R As = stencil<R>(C, k, j, i, ap, s0(k, j, n.iw), s0(k, j, n.ie), s0(k, n.js, i),
                            s0(k, n.jn, i), s0(n.kd, j, i), s0(n.ku, j, i), sc0);
    t_out[c] = tc;
    s_out[c] = sc0 + (tc - As) / ap;
}

// s_out = s + (t - A s)/a_p; with P_UPDATE: p = s_out + bet*p instead
template <typename R, bool P_UPDATE>
__global__ __launch_bounds__(256) void k_jac(const R *s, const R *t, R *s_out, CGeo g,
                                             CoefPack cp, const SolverScalars *sc) {
    if (sc->done || sc->breakdown) return;
    SOLVER_PROLOGUE
    if (!act) return;
    const Coef<R> C(cp.m64, cp.m32);
    const Nbr n = nbrs(k, j, i, g.kp, g.jp, g.ip);
    const long long c = cidx(g, k, j, i);
    const R ap = coef_ap(C, k, j, i);
    const R sc0 = s[c];
    const R As = stencil<R>(C, k, j, i, ap, s[cidx(g, k, j, n.iw)], s[cidx(g, k, j, n.ie)],
                            s[cidx(g, k, n.js, i)], s[cidx(g, k, n.jn, i)],
                            s[cidx(g, n.kd, j, i)], s[cidx(g, n.ku, j, i)], sc0);
    const R snew = sc0 + (t[c] - As) / ap;
    if (P_UPDATE) {
        const R bet = (R)sc->bet;
        s_out[c] = snew + bet * s_out[c];  // p = s + bet*p (dynamics.f90:399)
    } else {
        s_out[c] = snew;
    }
}

// y = A x; partial sums of <a,y> (and <y,y> when TWO)
template <typename R, bool TWO>
__global__ __launch_bounds__(256) void k_matvec_dot(const R *x, const R *a, R *y, CGeo g,
                                                    CoefPack cp, const SolverScalars *sc,
                                                    double *partials, long long stride) {
    if (sc->done || sc->breakdown) return;
    SOLVER_PROLOGUE
    const Coef<R> C(cp.m64, cp.m32);
    double s[2] = {0.0, 0.0};
    if (act) {
        const Nbr n = nbrs(k, j, i, g.kp, g.jp, g.ip);
        const long long c = cidx(g, k, j, i);
        const R ap = coef_ap(C, k, j, i);
        const R yv = stencil<R>(C, k, j, i, ap, x[cidx(g, k, j, n.iw)], x[cidx(g, k, j, n.ie)],
                                x[cidx(g, k, n.js, i)], x[cidx(g, k, n.jn, i)],
                                x[cidx(g, n.kd, j, i)], x[cidx(g, n.ku, j, i)], x[c]);
        y[c] = yv;
        if (TWO) {
            s[0] = (double)yv * (double)yv;   // tt = <t,t>
            s[1] = (double)yv * (double)a[c]; // ts = <t,s>
        } else {
            s[0] = (double)a[c] * (double)yv; // nrm = <cr,v1>
        }
    }
    if (TWO)
        block_sum_store<2>(s, partials, stride);
    else {
        double s1[1] = {s[0]};
        block_sum_store<1>(s1, partials, stride);
    }
}

// x = x + alf*p + omg*cs (padded x, FP64); r = s - omg*t; <r,r>, <r,cr>
template <typename R>
__global__ __launch_bounds__(256) void k_update(void *xv, const R *p, const R *cs, const R *s,
                                                const R *t, const R *cr, R *r, Geo pg, CGeo g,
                                                const SolverScalars *sc, double *partials,
                                                long long stride, int x_is_compact) {
    if (sc->done || sc->breakdown) return;
    SOLVER_PROLOGUE
    double acc[2] = {0.0, 0.0};
    if (act) {
        const long long c = cidx(g, k, j, i);
        const R alf = (R)sc->alf, omg = (R)sc->omg;
        if (x_is_compact) {
            // inner correction d, kept in the vector precision
            R *xr = (R *)xv;
            xr[c] = (xr[c] + alf * p[c]) + omg * cs[c];
        } else {
            double *x = (double *)xv;
            const long long xi = pidx(pg, k, j, i);
            x[xi] = (x[xi] + sc->alf * (double)p[c]) + sc->omg * (double)cs[c];
        }
        const R rv = s[c] - omg * t[c];
        r[c] = rv;
        acc[0] = (double)rv * (double)rv;
        acc[1] = (double)rv * (double)cr[c];
    }
    block_sum_store<2>(acc, partials, stride);
}

// ---------------------------------------------------------------------------
// scalar recurrences (one block; reduce the partial sums first)
// ---------------------------------------------------------------------------
// tiny = 1.e-16 is a default-real literal (dynamics.f90:321)
__device__ constexpr double kTinyOmega = (double)1.e-16f;

__global__ void k_fin_init(const double *partials, int nb, SolverScalars *sc, double tol,
                           int use_bb) {
    const double bb = reduce_slot_1024(partials, nb);
    const double rr = reduce_slot_1024(partials + nb, nb);
    if (threadIdx.x == 0) {
        double sc_err = use_bb ? sqrt(bb) : sc->sc_err;
        if (use_bb && sc_err < 0.01) sc_err = 0.01;  // :371
        sc->bb = bb;
        sc->sc_err = sc_err;
        sc->rr = rr;
        sc->err = sqrt(rr) / sc_err;
        sc->alf = 1.0;
        sc->omg = 1.0;
        sc->nrm = 1.0;
        sc->rho = rr;  // <r,cr> with cr = r (:391)
        sc->bet = (sc->rho / sc->nrm) * (sc->alf / sc->omg);
        sc->breakdown = 0;
        sc->nonfinite = !isfinite(sc->err);
        sc->done = (sc->err < tol) || sc->nonfinite;
    }
}

__global__ void k_fin_v(const double *partials, int nb, SolverScalars *sc) {
    if (sc->done || sc->breakdown) return;
    const double nrm = reduce_slot_1024(partials, nb);
    if (threadIdx.x == 0) {
        sc->nrm = nrm;
        sc->alf = sc->rho / nrm;  // :404
    }
}

__global__ void k_fin_t(const double *partials, int nb, SolverScalars *sc) {
    if (sc->done || sc->breakdown) return;
    const double tt = reduce_slot_1024(partials, nb);
    const double ts = reduce_slot_1024(partials + nb, nb);
    if (threadIdx.x == 0) {
        sc->tt = tt;
        sc->ts = ts;
        sc->omg = ts / tt;  // :412
        if (fabs(sc->omg) < kTinyOmega) sc->breakdown = 1;  // :414
    }
}

__global__ void k_fin_r(const double *partials, int nb, SolverScalars *sc, double tol) {
    if (sc->done || sc->breakdown) return;
    const double rr = reduce_slot_1024(partials, nb);
    const double rcr = reduce_slot_1024(partials + nb, nb);
    if (threadIdx.x == 0) {
        sc->rr = rr;
        sc->err = sqrt(rr) / sc->sc_err;  // :423
        sc->nrm = sc->rho;                // :425
        sc->rho = rcr;                    // next iteration's <r,cr> (:391)
        sc->bet = (sc->rho / sc->nrm) * (sc->alf / sc->omg);
        sc->nonfinite = !isfinite(sc->err);
        sc->done = (sc->err < tol) || sc->nonfinite;
    }
}

// final boundary rows of x (dynamics.f90:437-442), or the Neumann copies left
// by mat_ax on the early-return path (:383); whole z planes
__global__ void k_x_rows(double *x, const double *sw, Geo g, double dzn0, int final_rows) {
    const long long P = (long long)g.nj * g.ni;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < P;
         t += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(t % g.nj), i = (int)(t / g.nj);
        x[pidx(g, g.kp + 1, j, i)] = x[pidx(g, g.kp, j, i)];
        if (final_rows)
            x[pidx(g, 0, j, i)] = x[pidx(g, 1, j, i)] - sw[pidx(g, 0, j, i)] * dzn0;
        else
            x[pidx(g, 0, j, i)] = x[pidx(g, 1, j, i)];
    }
}

// ---------------------------------------------------------------------------
// whole-array helpers for the standalone precondition API
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ap_whole(const Geo &g, const Coef<double> &C, long long t) {
    const int k = (int)(t % g.nk);
    const long long r = t / g.nk;
    const int j = (int)(r % g.nj), i = (int)(r / g.nj);
    if (k < 1 || k > g.kp || j < 1 || j > g.jp || i < 1 || i > g.ip) return 1.0;  // :57
    return coef_ap(C, k, j, i);
}
__global__ void k_div_ap(double *s, const double *r, Geo g, CoefPack cp) {
    const Coef<double> C(cp.m64, cp.m32);
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < g.ntot;
         t += (long long)gridDim.x * blockDim.x)
        s[t] = r[t] / ap_whole(g, C, t);
}
__global__ void k_sweep_whole(double *s, const double *r, const double *t, Geo g, CoefPack cp) {
    const Coef<double> C(cp.m64, cp.m32);
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.ntot;
         q += (long long)gridDim.x * blockDim.x)
        s[q] = s[q] + (r[q] - t[q]) / ap_whole(g, C, q);
}
// ax = A x on the interior of padded arrays whose halo is already filled
__global__ void k_matvec_padded(const double *x, double *ax, Geo g, CoefPack cp) {
    const Coef<double> C(cp.m64, cp.m32);
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    ax[c] = stencil<double>(C, k, j, i, coef_ap(C, k, j, i), x[c - g.si], x[c + g.si],
                            x[c - g.sj], x[c + g.sj], x[c - 1], x[c + 1], x[c]);
}
__global__ void k_dot_padded(const double *x, const double *y, Geo g, double *partials) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    double s[1] = {0.0};
    if (k <= g.kp && j <= g.jp) {
        const long long c = pidx(g, k, j, i);
        s[0] = x[c] * y[c];
    }
    block_sum_store<1>(s, partials, 0);
}
// d (compact FP32) added to the padded FP64 x: x = x + d
__global__ void k_add_correction(double *x, const float *d, Geo pg, CGeo g) {
    SOLVER_PROLOGUE
    if (!act) return;
    x[pidx(pg, k, j, i)] += (double)d[cidx(g, k, j, i)];
}
// inner start: r = r32 - mean, cr = r, p = v1 = d = 0; <r,r>
__global__ void k_inner_init(const float *r32, float *r, float *cr, float *p, float *v1,
                             float *d, CGeo g, const double *mean_sum, double *partials,
                             long long stride) {
    SOLVER_PROLOGUE
    double s[2] = {0.0, 0.0};
    if (act) {
        const long long c = cidx(g, k, j, i);
        const float mean = (float)(*mean_sum / (double)g.n);
        const float rv = r32[c] - mean;
        r[c] = rv;
        cr[c] = rv;
        p[c] = 0.f;
        v1[c] = 0.f;
        d[c] = 0.f;
        s[0] = (double)rv * (double)rv;
    }
    block_sum_store<2>(s, partials, stride);
}

// three generic sums into dst[0..2]
__global__ void k_finalize_sums_3(const double *partials, int nb, double *dst) {
    const double a = reduce_slot_1024(partials, nb);
    const double b = reduce_slot_1024(partials + nb, nb);
    const double d = reduce_slot_1024(partials + 2 * (long long)nb, nb);
    if (threadIdx.x == 0) {
        dst[0] = a;
        dst[1] = b;
        dst[2] = d;
    }
}

// inner solve start: residuals measured relative to ||r_inner,0||
__global__ void k_fin_inner_init(const double *partials, int nb, SolverScalars *sc,
                                 double rtol) {
    const double rr = reduce_slot_1024(partials, nb);
    if (threadIdx.x == 0) {
        sc->sc_err = rr > 0.0 ? sqrt(rr) : 1.0;
        sc->rr = rr;
        sc->err = sqrt(rr) / sc->sc_err;
        sc->alf = 1.0;
        sc->omg = 1.0;
        sc->nrm = 1.0;
        sc->rho = rr;
        sc->bet = (sc->rho / sc->nrm) * (sc->alf / sc->omg);
        sc->breakdown = 0;
        sc->nonfinite = !isfinite(rr);
        sc->done = (rr == 0.0) || (sc->err < rtol) || sc->nonfinite;
    }
}

#undef SOLVER_PROLOGUE

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static CGeo cgeo(const Geo &g) {
    CGeo c;
    c.ip = g.ip;
    c.jp = g.jp;
    c.kp = g.kp;
    c.n = (long long)g.ip * g.jp * g.kp;
    return c;
}

// One BiCGSTAB run on compact vectors of precision R.  The vectors r, cr, p,
// v1 must already hold the initial state and sc the initial scalars.
// x: padded FP64 (FP64 path) or a compact R correction (mixed inner).
template <typename R>
static int bicgstab_iterate(dcm_ctx *c, R *const *V, void *x, int x_compact, double tol,
                            int itmax, int check_every, SolverOut *out) {
    const Geo &pg = c->geo;
    const CGeo g = cgeo(pg);
    const Interior L = interior_launch(pg);
    const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
    CoefPack cp{c->met, c->met32};
    R *r = V[0], *cr = V[1], *p = V[2], *v1 = V[3], *t = V[4], *s = V[5], *cs = V[6],
      *wa = V[7], *wb = V[8];
    SolverScalars *sc = c->solver;
    double *part = c->partials;
    int it = 0;
    if (itmax < 1) itmax = 1;
    for (it = 1; it <= itmax; ++it) {
        int tok = prof_begin(c, "jacobi");
        hipLaunchKernelGGL((k_jac_first<R, 0>), L.grid, L.block, 0, c->stream, r, v1, t, wa, g,
                           cp, sc);
        hipLaunchKernelGGL((k_jac<R, false>), L.grid, L.block, 0, c->stream, wa, t, wb, g, cp, sc);
        hipLaunchKernelGGL((k_jac<R, true>), L.grid, L.block, 0, c->stream, wb, t, p, g, cp, sc);
        prof_end(c, tok);
        tok = prof_begin(c, "matvec_dot");
        hipLaunchKernelGGL((k_matvec_dot<R, false>), L.grid, L.block, 0, c->stream, p, cr, v1, g,
                           cp, sc, part, (long long)nb);
        prof_end(c, tok);
        hipLaunchKernelGGL(k_fin_v, dim3(1), dim3(1024), 0, c->stream, part, nb, sc);
        tok = prof_begin(c, "jacobi");
        hipLaunchKernelGGL((k_jac_first<R, 1>), L.grid, L.block, 0, c->stream, r, v1, s, wa, g,
                           cp, sc);
        hipLaunchKernelGGL((k_jac<R, false>), L.grid, L.block, 0, c->stream, wa, s, wb, g, cp, sc);
        hipLaunchKernelGGL((k_jac<R, false>), L.grid, L.block, 0, c->stream, wb, s, cs, g, cp, sc);
        prof_end(c, tok);
        tok = prof_begin(c, "matvec_dot");
        hipLaunchKernelGGL((k_matvec_dot<R, true>), L.grid, L.block, 0, c->stream, cs, s, t, g,
                           cp, sc, part, (long long)nb);
        prof_end(c, tok);
        hipLaunchKernelGGL(k_fin_t, dim3(1), dim3(1024), 0, c->stream, part, nb, sc);
        tok = prof_begin(c, "update");
        hipLaunchKernelGGL(k_update<R>, L.grid, L.block, 0, c->stream, x, p, cs, s, t, cr, r, pg,
                           g, sc, part, (long long)nb, x_compact);
        prof_end(c, tok);
        hipLaunchKernelGGL(k_fin_r, dim3(1), dim3(1024), 0, c->stream, part, nb, sc, tol);
        DCM_LAUNCH_CHECK();
        if (it % check_every == 0 || it == itmax) {
            DCM_HIP(hipMemcpyAsync(c->h_solver, sc, sizeof(SolverScalars), hipMemcpyDeviceToHost,
                                   c->stream));
            DCM_HIP(hipStreamSynchronize(c->stream));
            if (c->h_solver->breakdown || c->h_solver->done) break;
        }
    }
    // the device stops at the first converged iteration; count from the host
    // copy of the scalars is exact only when check_every == 1
    out->iters = it > itmax ? itmax : it;
    out->err = c->h_solver->err;
    if (c->h_solver->breakdown) return DCM_BREAKDOWN;
    if (c->h_solver->nonfinite) {
        set_error("bicgstab: non-finite residual");
        return DCM_NONFINITE;
    }
    if (!(c->h_solver->err < tol)) return DCM_NOT_CONVERGED;
    return DCM_OK;
}

// FP64 BiCGSTAB of the reference on padded x, b (dynamics.f90:293-444)
int bicgstab64(dcm_ctx *c, const double *sw, double *x, double *b, double tol, int itmax,
               SolverOut *out) {
    const Geo &pg = c->geo;
    const CGeo g = cgeo(pg);
    const Interior L = interior_launch(pg);
    const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
    CoefPack cp{c->met, c->met32};
    DCM_CHECK(ensure_partials(c, 4 * (size_t)nb + 64));
    DCM_CHECK(ensure_solver_vectors64(c, 9));
    double *V[9];
    for (int q = 0; q < 9; ++q) V[q] = c->sv64 + (size_t)q * g.n;
    int tok = prof_begin(c, "resid0");
    hipLaunchKernelGGL(k_resid0, L.grid, L.block, 0, c->stream, x, b, V[0], V[1], V[2], V[3], pg,
                       g, cp, c->partials, (long long)nb);
    prof_end(c, tok);
    hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(1024), 0, c->stream, c->partials, nb, c->solver,
                       tol, 1);
    DCM_LAUNCH_CHECK();
    DCM_HIP(hipMemcpyAsync(c->h_solver, c->solver, sizeof(SolverScalars), hipMemcpyDeviceToHost,
                           c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    int st = DCM_OK;
    bool early = false;
    if (c->h_solver->nonfinite) {
        set_error("bicgstab: non-finite initial residual");
        st = DCM_NONFINITE;
    } else if (c->h_solver->err < tol) {
        early = true;  // :383, return before the loop
        out->iters = 0;
        out->err = c->h_solver->err;
    } else {
        st = bicgstab_iterate<double>(c, V, x, 0, tol, itmax, 1, out);
    }
    if (st == DCM_OK) {
        // the halo of x as the reference leaves it: periodic faces from
        // mat_ax, then the boundary rows of :437-442 (or mat_ax's Neumann
        // copies on the early return)
        DCM_CHECK(fill<double>(c, x, FILL_ALONG_DIM_WO, 0.0, 0.0));
        hipLaunchKernelGGL(k_x_rows, dim3(256), dim3(kBlock), 0, c->stream, x, sw, pg,
                           c->h_dzn[0], early ? 0 : 1);
        DCM_LAUNCH_CHECK();
    }
    return st;
}

// Mixed precision: FP32 inner BiCGSTAB inside FP64 iterative refinement.
int bicgstab_mixed(dcm_ctx *c, const double *sw, double *x, double *b, double tol, int itmax,
                   double inner_rtol, SolverOut *out) {
    const Geo &pg = c->geo;
    const CGeo g = cgeo(pg);
    const Interior L = interior_launch(pg);
    const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
    CoefPack cp{c->met, c->met32};
    DCM_CHECK(ensure_partials(c, 4 * (size_t)nb + 64));
    DCM_CHECK(ensure_solver_vectors32(c, 11));
    float *V[11];
    for (int q = 0; q < 11; ++q) V[q] = c->sv32 + (size_t)q * g.n;
    float *r32 = V[9], *d = V[10];
    SolverScalars *sc = c->solver;
    out->iters = 0;
    out->outer = 0;
    double sc_err = 0.0;
    const int max_outer = 50;
    for (int outer = 0; outer <= max_outer; ++outer) {
        // true FP64 residual, its FP32 copy and its mean
        int tok = prof_begin(c, "resid_true");
        hipLaunchKernelGGL(k_resid_true, L.grid, L.block, 0, c->stream, x, b, r32, pg, g, cp,
                           c->partials, (long long)nb);
        prof_end(c, tok);
        hipLaunchKernelGGL(k_finalize_sums_3, dim3(1), dim3(1024), 0, c->stream, c->partials, nb,
                           c->dscal + 40);
        DCM_LAUNCH_CHECK();
        DCM_HIP(hipMemcpyAsync(c->hscal + 40, c->dscal + 40, 3 * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
        DCM_HIP(hipStreamSynchronize(c->stream));
        const double rr = c->hscal[40], bb = c->hscal[42];
        if (outer == 0) {
            sc_err = sqrt(bb);
            if (sc_err < 0.01) sc_err = 0.01;
        }
        const double err = sqrt(rr) / sc_err;
        out->err = err;
        if (!std::isfinite(err)) {
            set_error("bicgstab_mixed: non-finite residual");
            return DCM_NONFINITE;
        }
        if (err < tol) break;
        if (outer == max_outer || out->iters >= itmax) return DCM_NOT_CONVERGED;
        out->outer += 1;
        // inner solve A d = r - mean(r), to a relative reduction inner_rtol
        hipLaunchKernelGGL(k_inner_init, L.grid, L.block, 0, c->stream, r32, V[0], V[1], V[2],
                           V[3], d, g, c->dscal + 41, c->partials, (long long)nb);
        DCM_LAUNCH_CHECK();
        // sc_err of the inner run = ||r_inner,0||: its recursive residual
        // err_inner = ||r_k|| / ||r_0||, stopped at inner_rtol, or once the
        // outer target is surely met.
        double rtol = inner_rtol;
        const double need = 0.5 * tol * sc_err / (sqrt(rr) > 0 ? sqrt(rr) : 1.0);
        if (need > rtol) rtol = need;
        hipLaunchKernelGGL(k_fin_inner_init, dim3(1), dim3(1024), 0, c->stream, c->partials, nb,
                           sc, rtol);
        DCM_LAUNCH_CHECK();
        SolverOut io;
        int st = bicgstab_iterate<float>(c, V, (void *)d, 1, rtol, itmax - out->iters, 1, &io);
        out->iters += io.iters;
        if (st == DCM_NONFINITE) return st;
        // a breakdown or stagnation of the inner solve still returns its
        // best correction; the outer loop judges the true residual
        hipLaunchKernelGGL(k_add_correction, L.grid, L.block, 0, c->stream, x, d, pg, g);
        DCM_LAUNCH_CHECK();
    }
    DCM_CHECK(fill<double>(c, x, FILL_ALONG_DIM_WO, 0.0, 0.0));
    hipLaunchKernelGGL(k_x_rows, dim3(256), dim3(kBlock), 0, c->stream, x, sw, pg, c->h_dzn[0],
                       1);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

}  // namespace dcm

using namespace dcm;

extern "C" {

int dcm_matvec(dcm_ctx *c, double *x, double *ax) {
    if (!c || !x || !ax) return DCM_BAD_ARG;
    const Geo &g = c->geo;
    // exchange_along_dim then the Neumann rows (dynamics.f90:203-211)
    DCM_CHECK(fill<double>(c, x, FILL_ALONG_DIM_WO, 0.0, 0.0));
    DCM_CHECK(fill_neumann_z(c, x));
    DCM_HIP(hipMemsetAsync(ax, 0, g.ntot * sizeof(double), c->stream));  // :213
    const Interior L = interior_launch(g);
    CoefPack cp{c->met, c->met32};
    hipLaunchKernelGGL(k_matvec_padded, L.grid, L.block, 0, c->stream, x, ax, g, cp);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

int dcm_precondition(dcm_ctx *c, double *s, const double *r, int prit) {
    if (!c || !s || !r) return DCM_BAD_ARG;
    double *t;
    DCM_CHECK(work_field64(c, 0, &t));
    CoefPack cp{c->met, c->met32};
    hipLaunchKernelGGL(k_div_ap, dim3(2048), dim3(kBlock), 0, c->stream, s, r, c->geo, cp);
    DCM_LAUNCH_CHECK();
    for (int it = 1; it <= prit; ++it) {
        DCM_CHECK(dcm_matvec(c, s, t));
        hipLaunchKernelGGL(k_sweep_whole, dim3(2048), dim3(kBlock), 0, c->stream, s, r, t,
                           c->geo, cp);
        DCM_LAUNCH_CHECK();
    }
    return DCM_OK;
}

int dcm_inn_prod(dcm_ctx *c, const double *x, const double *y, double *out) {
    if (!c || !x || !y || !out) return DCM_BAD_ARG;
    const Interior L = interior_launch(c->geo);
    const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
    DCM_CHECK(ensure_partials(c, (size_t)nb + 16));
    hipLaunchKernelGGL(k_dot_padded, L.grid, L.block, 0, c->stream, x, y, c->geo, c->partials);
    DCM_LAUNCH_CHECK();
    DCM_CHECK(finalize_sums(c, c->partials, nb, 1, c->dscal + 40));
    DCM_HIP(hipMemcpyAsync(c->hscal + 40, c->dscal + 40, sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    *out = c->hscal[40];
    return DCM_OK;
}

int dcm_true_residual(dcm_ctx *c, const double *x, const double *b, double *out) {
    if (!c || !x || !b || !out) return DCM_BAD_ARG;
    const Geo &pg = c->geo;
    const CGeo g = cgeo(pg);
    const Interior L = interior_launch(pg);
    const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
    CoefPack cp{c->met, c->met32};
    DCM_CHECK(ensure_partials(c, 4 * (size_t)nb + 64));
    hipLaunchKernelGGL(k_resid_true, L.grid, L.block, 0, c->stream, x, b, (float *)nullptr, pg,
                       g, cp, c->partials, (long long)nb);
    hipLaunchKernelGGL(k_finalize_sums_3, dim3(1), dim3(1024), 0, c->stream, c->partials, nb,
                       c->dscal + 40);
    DCM_LAUNCH_CHECK();
    DCM_HIP(hipMemcpyAsync(c->hscal + 40, c->dscal + 40, 3 * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    double sc_err = sqrt(c->hscal[42]);
    if (sc_err < 0.01) sc_err = 0.01;
    *out = sqrt(c->hscal[40]) / sc_err;
    return DCM_OK;
}

int dcm_bicgstab(dcm_ctx *c, const double *xg, const double *yg, const double *zg,
                 const double *sw, double *x, double *b, double tol, int test_solver, int itmax,
                 int *iters, double *err) {
    if (!c || !sw || !x || !b) return DCM_BAD_ARG;
    const Geo &g = c->geo;
    if (c->first_solve) {
        // first call (dynamics.f90:330-367): the coefficients already live
        // in the context; test_solver overwrites b with a Gaussian (:354-364),
        // evaluated on the host with the C library's exp
        if (test_solver) {
            if (!xg || !yg || !zg) return DCM_BAD_ARG;
            std::vector<double> hb((size_t)g.ntot);
            DCM_HIP(hipMemcpyAsync(hb.data(), b, g.ntot * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
            DCM_HIP(hipStreamSynchronize(c->stream));
            for (int i = 1; i <= g.ip; ++i)
                for (int j = 1; j <= g.jp; ++j)
                    for (int k = 1; k <= g.kp; ++k) {
                        const double zz = zg[k] - 4000.0;
                        hb[pidx(g, k, j, i)] =
                            exp(-(1.0 / 5.e6) * ((xg[i] * xg[i] + yg[j] * yg[j]) + zz * zz));
                    }
            DCM_HIP(hipMemcpyAsync(b, hb.data(), g.ntot * sizeof(double), hipMemcpyHostToDevice,
                                   c->stream));
            DCM_HIP(hipStreamSynchronize(c->stream));
        }
        c->first_solve = false;
    }
    SolverOut o;
    int st = bicgstab64(c, sw, x, b, tol, itmax > 0 ? itmax : 9999, &o);
    if (iters) *iters = o.iters;
    if (err) *err = o.err;
    return st;
}

int dcm_bicgstab_mixed(dcm_ctx *c, const double *sw, double *x, double *b, double tol, int itmax,
                       int *iters, int *outer, double *err) {
    if (!c || !sw || !x || !b) return DCM_BAD_ARG;
    SolverOut o;
    int st = bicgstab_mixed(c, sw, x, b, tol, itmax > 0 ? itmax : 9999, 1e-5, &o);
    if (iters) *iters = o.iters;
    if (outer) *outer = o.outer;
    if (err) *err = o.err;
    return st;
}

}  // extern "C"
