/*
 * dcm_oracle.c -- CPU restatement of the DCM dynamical-core hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by the
 * tests, by the smoke check and by the CPU-baseline leg of the benchmark.
 *
 * Each function below restates a routine of the reference Fortran, FP64 build
 * (VAR_TYPE=1), single MPI rank, keeping the reference's loop order, operand
 * order and parenthesisation.  Build with -O2 -ffp-contract=off so that no
 * multiply-add is contracted (the reference is built for x86-64 without FMA).
 *
 * Reference files (all citations relative to the reference tree):
 *   mpi_module.f90                boundary fills (exchange_*)
 *   pamm/sfvt/advection_3d.F90    upwind / MPDATA / reference-state advection
 *   diagnostics.f90               horizontal_means
 *   dynamics.f90                  sources, bicgstab, advance_momentum
 *   sgm/subgrid_3d.f90            advance_scalar_fields_3d
 *   driver_code.f90               model_driver time step
 */
#include "dcm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Host threads of the element-wise loops (or_set_threads; 1, the default:
 * serial).  Every such loop writes disjoint elements from inputs it does not
 * write, so its results do not depend on the thread count; the global sums
 * keep the reference's serial order unless or_set_sum_blocked(2) asks for a
 * parallel reduction (the CPU baseline of the benchmark only). */
static int g_threads = 1;
#define OMP_FOR _Pragma("omp parallel for schedule(static) if (g_threads > 1)")
int or_set_threads(int n) {
#ifdef _OPENMP
    g_threads = n > 1 ? n : 1;
    omp_set_num_threads(g_threads);
    return g_threads;
#else
    (void)n;
    g_threads = 1;
    return 1;
#endif
}

#define NK(g) ((size_t)(g)->kp + 2)
#define NJ(g) ((size_t)(g)->jp + 2)
#define NI(g) ((size_t)(g)->ip + 2)
#define NTOT(g) (NK(g) * NJ(g) * NI(g))
/* element (k,j,i) of a padded field, Fortran indices 0..n+1 */
#define A3(a, k, j, i) ((a)[((size_t)(i) * NJ(g) + (size_t)(j)) * NK(g) + (size_t)(k)])

/* small: pamm/sfvt/advection_3d.F90:19 */
static const double SMALL = 1e-60;
/* grav: dynamics.f90:18 */
static const double GRAV = 9.81;
/* tiny = 1.e-16 is a default-real (binary32) literal: dynamics.f90:321 */
static const double TINY_OMEGA = (double)1.e-16f;

static double *zalloc(size_t n) {
    double *p = (double *)calloc(n, sizeof(double));
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/* Fortran max(x,0._wp) / min(x,0._wp) with a fixed choice for signed zero */
static inline double pos0(double x) { return x > 0.0 ? x : 0.0; }
static inline double neg0(double x) { return x < 0.0 ? x : 0.0; }
static inline double min3(double a, double b, double c) {
    double m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    return m;
}

/* ------------------------------------------------------------------------ */
/* boundary fills                                                           */
/* ------------------------------------------------------------------------ */

static void fill_faces_xy(const or_grid *g, double *a) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    /* x faces, k 1..kp, j 1..jp: mpi_module.f90:857-862 (s_west == id) */
    for (int j = 1; j <= jp; ++j)
        for (int k = 1; k <= kp; ++k) {
            A3(a, k, j, 0) = A3(a, k, j, ip);
            A3(a, k, j, ip + 1) = A3(a, k, j, 1);
        }
    /* y faces, k 1..kp, i 1..ip: mpi_module.f90:864-870 */
    for (int i = 1; i <= ip; ++i)
        for (int k = 1; k <= kp; ++k) {
            A3(a, k, 0, i) = A3(a, k, jp, i);
            A3(a, k, jp + 1, i) = A3(a, k, 1, i);
        }
}

static void fill_zplanes(const or_grid *g, double *a, double lbc, double ubc) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    /* whole planes including the lateral halo: mpi_module.f90:874-883 */
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) A3(a, kp + 1, j, i) = ubc;
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) A3(a, 0, j, i) = lbc;
}

/* exchange_along_dim, mpi_module.f90:700-910 (one rank: lateral edges stale) */
void or_fill_along_dim(const or_grid *g, double *a, double lbc, double ubc) {
    fill_faces_xy(g, a);
    fill_zplanes(g, a, lbc, ubc);
}

/* exchange_along_dim_wo, mpi_module.f90:1285-1494 (no z planes) */
void or_fill_along_dim_wo(const or_grid *g, double *a) { fill_faces_xy(g, a); }

/* exchange_along_z, mpi_module.f90:584-685 (one rank: z planes only) */
void or_fill_along_z(const or_grid *g, double *a, double lbc, double ubc) {
    fill_zplanes(g, a, lbc, ubc);
}

/* The 12 edges and 8 corners of exchange_full / exchange_full_wo for one
 * rank.  Each receive buffer is pre-filled from the local send region
 * (mpi_module.f90:1180-1196 corners, 1244-1246 edges); x-y neighbours are the
 * rank itself, so nothing is received, and z neighbours are MPI_PROC_NULL, so
 * the receive leaves the local copy: every edge or corner touching a z plane
 * takes the z-wrapped interior value (mpi_module.f90:1250-1263). */
static void fill_edges_corners(const or_grid *g, double *a) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    /* vertical edges, k 1..kp: ws_bt, wn_bt, es_bt, en_bt */
    for (int k = 1; k <= kp; ++k) {
        double ws = A3(a, k, 1, 1), wn = A3(a, k, jp, 1);
        double es = A3(a, k, 1, ip), en = A3(a, k, jp, ip);
        A3(a, k, jp + 1, ip + 1) = ws;
        A3(a, k, 0, ip + 1) = wn;
        A3(a, k, jp + 1, 0) = es;
        A3(a, k, 0, 0) = en;
    }
    /* edges along i: bs_we, bn_we, ts_we, tn_we */
    for (int i = 1; i <= ip; ++i) {
        double bs = A3(a, 1, 1, i), bn = A3(a, 1, jp, i);
        double ts = A3(a, kp, 1, i), tn = A3(a, kp, jp, i);
        A3(a, kp + 1, jp + 1, i) = bs;
        A3(a, kp + 1, 0, i) = bn;
        A3(a, 0, jp + 1, i) = ts;
        A3(a, 0, 0, i) = tn;
    }
    /* edges along j: bw_sn, be_sn, tw_sn, te_sn */
    for (int j = 1; j <= jp; ++j) {
        double bw = A3(a, 1, j, 1), be = A3(a, 1, j, ip);
        double tw = A3(a, kp, j, 1), te = A3(a, kp, j, ip);
        A3(a, kp + 1, j, ip + 1) = bw;
        A3(a, kp + 1, j, 0) = be;
        A3(a, 0, j, ip + 1) = tw;
        A3(a, 0, j, 0) = te;
    }
    /* corners: wsb wnb wst wnt esb enb est ent */
    double wsb = A3(a, 1, 1, 1), wnb = A3(a, 1, jp, 1);
    double wst = A3(a, kp, 1, 1), wnt = A3(a, kp, jp, 1);
    double esb = A3(a, 1, 1, ip), enb = A3(a, 1, jp, ip);
    double est = A3(a, kp, 1, ip), ent = A3(a, kp, jp, ip);
    A3(a, kp + 1, jp + 1, ip + 1) = wsb;
    A3(a, kp + 1, 0, ip + 1) = wnb;
    A3(a, 0, jp + 1, ip + 1) = wst;
    A3(a, 0, 0, ip + 1) = wnt;
    A3(a, kp + 1, jp + 1, 0) = esb;
    A3(a, kp + 1, 0, 0) = enb;
    A3(a, 0, jp + 1, 0) = est;
    A3(a, 0, 0, 0) = ent;
}

/* exchange_full, mpi_module.f90:926-1271 */
void or_fill_full(const or_grid *g, double *a, double lbc, double ubc) {
    or_fill_along_dim(g, a, lbc, ubc);
    fill_edges_corners(g, a);
}

/* exchange_full_wo, mpi_module.f90:1511-1855 */
void or_fill_full_wo(const or_grid *g, double *a) {
    or_fill_along_dim_wo(g, a);
    fill_edges_corners(g, a);
}

/* ------------------------------------------------------------------------ */
/* advection                                                                */
/* ------------------------------------------------------------------------ */

/* first_order_upstream_3d, pamm/sfvt/advection_3d.F90:39-116 */
void or_upwind3d(const or_grid *g, double dt, const double *u, const double *v,
                 const double *w, double *psi, int neumann) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t n = (size_t)ip * jp * kp;
    double *fxr = zalloc(n), *fxl = zalloc(n), *fyr = zalloc(n);
    double *fyl = zalloc(n), *fzr = zalloc(n), *fzl = zalloc(n);
#define F3(f, k, j, i) f[((size_t)((i)-1) * jp + ((j)-1)) * kp + ((k)-1)]
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                double uc = A3(u, k, j, i), ul = A3(u, k, j, i - 1);
                double vc = A3(v, k, j, i), vl = A3(v, k, j - 1, i);
                double wc = A3(w, k, j, i), wl = A3(w, k - 1, j, i);
                F3(fxr, k, j, i) = (((uc + fabs(uc)) * A3(psi, k, j, i) +
                                     (uc - fabs(uc)) * A3(psi, k, j, i + 1)) * dt) /
                                   (2.0 * g->dxn[i]);
                F3(fxl, k, j, i) = (((ul + fabs(ul)) * A3(psi, k, j, i - 1) +
                                     (ul - fabs(ul)) * A3(psi, k, j, i)) * dt) /
                                   (2.0 * g->dxn[i]);
                F3(fyr, k, j, i) = (((vc + fabs(vc)) * A3(psi, k, j, i) +
                                     (vc - fabs(vc)) * A3(psi, k, j + 1, i)) * dt) /
                                   (2.0 * g->dyn[j]);
                F3(fyl, k, j, i) = (((vl + fabs(vl)) * A3(psi, k, j - 1, i) +
                                     (vl - fabs(vl)) * A3(psi, k, j, i)) * dt) /
                                   (2.0 * g->dyn[j]);
                F3(fzr, k, j, i) =
                    ((((wc + fabs(wc)) * g->rhoa[k]) * A3(psi, k, j, i) +
                      ((wc - fabs(wc)) * g->rhoa[k]) * A3(psi, k + 1, j, i)) * dt) /
                    ((2.0 * g->dzn[k]) * g->rhoan[k]);
                F3(fzl, k, j, i) =
                    ((((wl + fabs(wl)) * g->rhoa[k - 1]) * A3(psi, k - 1, j, i) +
                      ((wl - fabs(wl)) * g->rhoa[k - 1]) * A3(psi, k, j, i)) * dt) /
                    ((2.0 * g->dzn[k]) * g->rhoan[k]);
            }
    /* neumann flag: advection_3d.F90:95-111 */
    if (neumann == 1 || neumann == 2) {
        OMP_FOR
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j)
                F3(fzl, 1, j, i) = neumann == 1 ? F3(fzr, 1, j, i) : -F3(fzr, 1, j, i);
        OMP_FOR
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j)
                F3(fzr, kp, j, i) = neumann == 1 ? F3(fzl, kp, j, i) : -F3(fzl, kp, j, i);
    }
    /* advection_3d.F90:115 */
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k)
                A3(psi, k, j, i) = ((A3(psi, k, j, i) - (F3(fxr, k, j, i) - F3(fxl, k, j, i))) -
                                    (F3(fyr, k, j, i) - F3(fyl, k, j, i))) -
                                   (F3(fzr, k, j, i) - F3(fzl, k, j, i));
#undef F3
    free(fxr); free(fxl); free(fyr); free(fyl); free(fzr); free(fzl);
}

/* Which copy of advection_3d.F90 the MPDATA routines follow.  0: the one
 * linked into the model, pamm/sfvt/advection_3d.F90.  1: sfvt/advection_3d.F90,
 * whose limiter bounds each direction separately with unweighted
 * denominators (:555-682) and whose vector routines use unweighted
 * antidiffusive velocities (:1384-1507, :1874-1909). */
static int g_mpdata_variant = 0;
void or_set_mpdata_variant(int v) { g_mpdata_variant = v; }

/* Antidiffusive velocities, pamm/sfvt/advection_3d.F90:318-506 (identical
 * arithmetic at :1355-1545 for the vector variant). */
static void antidiffusive(const or_grid *g, double dt, const double *ut,
                          const double *vt, const double *wt,
                          const double *po, double *us, double *vs,
                          double *ws) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *dx = g->dx, *dy = g->dy, *dz = g->dz;
    const double *dxn = g->dxn, *dyn = g->dyn, *dzn = g->dzn;
    const double *ra = g->rhoa, *rn = g->rhoan;
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double g1 = rn[k], g2 = rn[k];
                const double g3 = 0.5 * (rn[k] + rn[k + 1]);
#define U(kk, jj, ii) A3(ut, kk, jj, ii)
#define V(kk, jj, ii) A3(vt, kk, jj, ii)
#define W(kk, jj, ii) A3(wt, kk, jj, ii)
#define P(kk, jj, ii) A3(po, kk, jj, ii)
                double u_div1 =
                    (ra[k] * (W(k, j, i) + W(k, j, i + 1)) -
                     ra[k - 1] * (W(k - 1, j, i) + W(k - 1, j, i + 1))) / (dz[k - 1] * g1) +
                    (rn[k] * (((V(k, j, i) + V(k, j, i + 1)) - V(k, j - 1, i)) - V(k, j - 1, i + 1))) /
                        (dy[j - 1] * g1);
                double u_div2 =
                    (ra[k] * (W(k, j, i) + W(k, j + 1, i)) -
                     ra[k - 1] * (W(k - 1, j, i) + W(k - 1, j + 1, i))) / (dz[k - 1] * g2) +
                    (rn[k] * (((U(k, j, i) + U(k, j + 1, i)) - U(k, j, i - 1)) - U(k, j + 1, i - 1))) /
                        (dx[i - 1] * g2);
                double u_div3 =
                    (rn[k] * (U(k, j, i) - U(k, j, i - 1)) +
                     rn[k + 1] * (U(k + 1, j, i) - U(k + 1, j, i - 1))) / (dx[i - 1] * g3) +
                    (rn[k] * (V(k, j, i) - V(k, j - 1, i)) +
                     rn[k + 1] * (V(k + 1, j, i) - V(k + 1, j - 1, i))) / (dy[j - 1] * g3);

                /* u: z and y cross terms (eqs 13/14 of Smolarkiewicz 1984) */
                double tz1 =
                    ((((0.25 * (((ra[k] * W(k, j, i + 1) + ra[k] * W(k, j, i)) +
                                 ra[k - 1] * W(k - 1, j, i + 1)) + ra[k - 1] * W(k - 1, j, i))) / g1) *
                      (((P(k + 1, j, i + 1) + P(k + 1, j, i)) - P(k - 1, j, i + 1)) - P(k - 1, j, i))) /
                     ((((P(k + 1, j, i + 1) + P(k + 1, j, i)) + P(k - 1, j, i + 1)) + P(k - 1, j, i)) + SMALL)) /
                    (0.5 * (dzn[k - 1] + dzn[k]));
                double ty1 =
                    ((((0.25 * (((rn[k] * V(k, j, i + 1) + rn[k] * V(k, j, i)) +
                                 rn[k] * V(k, j - 1, i + 1)) + rn[k] * V(k, j - 1, i))) / g1) *
                      (((P(k, j + 1, i + 1) + P(k, j + 1, i)) - P(k, j - 1, i + 1)) - P(k, j - 1, i))) /
                     ((((P(k, j + 1, i + 1) + P(k, j + 1, i)) + P(k, j - 1, i + 1)) + P(k, j - 1, i)) + SMALL)) /
                    (0.5 * (dyn[j - 1] + dyn[j]));
                double ujb1 = ((0.5 * dt) * U(k, j, i)) * (tz1 + ty1);

                /* v: z and x cross terms */
                double tz2 =
                    ((((0.25 * (((ra[k] * W(k, j + 1, i) + ra[k] * W(k, j, i)) +
                                 ra[k - 1] * W(k - 1, j + 1, i)) + ra[k - 1] * W(k - 1, j, i))) / g2) *
                      (((P(k + 1, j + 1, i) + P(k + 1, j, i)) - P(k - 1, j + 1, i)) - P(k - 1, j, i))) /
                     ((((P(k + 1, j + 1, i) + P(k + 1, j, i)) + P(k - 1, j + 1, i)) + P(k - 1, j, i)) + SMALL)) /
                    (0.5 * (dzn[k - 1] + dzn[k]));
                double tx2 =
                    ((((0.25 * (((rn[k] * U(k, j + 1, i) + rn[k] * U(k, j, i)) +
                                 rn[k] * U(k, j + 1, i - 1)) + rn[k] * U(k, j, i - 1))) / g2) *
                      (((P(k, j + 1, i + 1) + P(k, j, i + 1)) - P(k, j + 1, i - 1)) - P(k, j, i - 1))) /
                     ((((P(k, j + 1, i + 1) + P(k, j, i + 1)) + P(k, j + 1, i - 1)) + P(k, j, i - 1)) + SMALL)) /
                    (0.5 * (dxn[i - 1] + dxn[i]));
                double ujb2 = ((0.5 * dt) * V(k, j, i)) * (tz2 + tx2);

                /* w: y and x cross terms */
                double ty3 =
                    ((((0.25 * (((rn[k + 1] * V(k + 1, j, i) + rn[k] * V(k, j, i)) +
                                 rn[k + 1] * V(k + 1, j - 1, i)) + rn[k] * V(k, j - 1, i))) / g3) *
                      (((P(k + 1, j + 1, i) + P(k, j + 1, i)) - P(k + 1, j - 1, i)) - P(k, j - 1, i))) /
                     ((((P(k + 1, j + 1, i) + P(k, j + 1, i)) + P(k + 1, j - 1, i)) + P(k, j - 1, i)) + SMALL)) /
                    (0.5 * (dyn[j - 1] + dyn[j]));
                double tx3 =
                    ((((0.25 * (((rn[k + 1] * U(k + 1, j, i) + rn[k] * U(k, j, i)) +
                                 rn[k + 1] * U(k + 1, j, i - 1)) + rn[k] * U(k, j, i - 1))) / g3) *
                      (((P(k + 1, j, i + 1) + P(k, j, i + 1)) - P(k + 1, j, i - 1)) - P(k, j, i - 1))) /
                     ((((P(k + 1, j, i + 1) + P(k, j, i + 1)) + P(k + 1, j, i - 1)) + P(k, j, i - 1)) + SMALL)) /
                    (0.5 * (dxn[i - 1] + dxn[i]));
                double ujb3 = ((0.5 * dt) * W(k, j, i)) * (ty3 + tx3);

                const double uc = U(k, j, i), vc = V(k, j, i), wc = W(k, j, i);
                double usv = ((((fabs(uc) * dx[i]) - ((dt * uc) * uc)) * (P(k, j, i + 1) - P(k, j, i))) /
                              ((P(k, j, i + 1) + P(k, j, i)) + SMALL)) / dxn[i] - ujb1;
                double vsv = ((((fabs(vc) * dy[j]) - ((dt * vc) * vc)) * (P(k, j + 1, i) - P(k, j, i))) /
                              ((P(k, j + 1, i) + P(k, j, i)) + SMALL)) / dyn[j] - ujb2;
                double wsv = ((((fabs(wc) * dz[k]) - (((dt * (ra[k] / g3)) * wc) * wc)) *
                               (P(k + 1, j, i) - P(k, j, i))) /
                              ((P(k + 1, j, i) + P(k, j, i)) + SMALL)) / dzn[k] - ujb3;
                /* divergent-flow correction, eq 38 of Smolarkiewicz 1984 */
                usv = usv - ((0.25 * dt) * uc) * (((U(k, j, i + 1) - U(k, j, i - 1)) / dx[i - 1]) + u_div1);
                vsv = vsv - ((0.25 * dt) * vc) * (((V(k, j + 1, i) - V(k, j - 1, i)) / dy[j - 1]) + u_div2);
                wsv = wsv - ((0.25 * dt) * wc) *
                                (((ra[k + 1] * W(k + 1, j, i) - ra[k - 1] * W(k - 1, j, i)) / (dz[k - 1] * g3)) +
                                 u_div3);
                A3(us, k, j, i) = usv;
                A3(vs, k, j, i) = vsv;
                A3(ws, k, j, i) = wsv;
#undef U
#undef V
#undef W
#undef P
            }
}

/* Unweighted antidiffusive velocities of sfvt/advection_3d.F90's
 * mpdata_vec_3d (:1384-1507): eqs 13, 14 and 38 of Smolarkiewicz (1984)
 * without the density factors. */
static void antidiffusive_plain(const or_grid *g, double dt, const double *ut,
                                const double *vt, const double *wt,
                                const double *po, double *us, double *vs,
                                double *ws) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *dx = g->dx, *dy = g->dy, *dz = g->dz;
    const double *dxn = g->dxn, *dyn = g->dyn, *dzn = g->dzn;
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
#define U(kk, jj, ii) A3(ut, kk, jj, ii)
#define V(kk, jj, ii) A3(vt, kk, jj, ii)
#define W(kk, jj, ii) A3(wt, kk, jj, ii)
#define P(kk, jj, ii) A3(po, kk, jj, ii)
                /* :1386-1401 */
                double u_div1 = (((W(k, j, i) + W(k, j, i + 1)) - W(k - 1, j, i)) - W(k - 1, j, i + 1)) / dz[k - 1] +
                                (((V(k, j, i) + V(k, j, i + 1)) - V(k, j - 1, i)) - V(k, j - 1, i + 1)) / dy[j - 1];
                double u_div2 = (((W(k, j, i) + W(k, j + 1, i)) - W(k - 1, j, i)) - W(k - 1, j + 1, i)) / dz[k - 1] +
                                (((U(k, j, i) + U(k, j + 1, i)) - U(k, j, i - 1)) - U(k, j + 1, i - 1)) / dx[i - 1];
                double u_div3 = (((U(k, j, i) + U(k + 1, j, i)) - U(k, j, i - 1)) - U(k + 1, j, i - 1)) / dx[i - 1] +
                                (((V(k, j, i) + V(k + 1, j, i)) - V(k, j - 1, i)) - V(k + 1, j - 1, i)) / dy[j - 1];
                /* :1412-1473 */
                double tz1 = (((0.25 * (((W(k, j, i + 1) + W(k, j, i)) + W(k - 1, j, i + 1)) + W(k - 1, j, i))) *
                               (((P(k + 1, j, i + 1) + P(k + 1, j, i)) - P(k - 1, j, i + 1)) - P(k - 1, j, i))) /
                              ((((P(k + 1, j, i + 1) + P(k + 1, j, i)) + P(k - 1, j, i + 1)) + P(k - 1, j, i)) + SMALL)) /
                             (0.5 * (dzn[k - 1] + dzn[k]));
                double ty1 = (((0.25 * (((V(k, j, i + 1) + V(k, j, i)) + V(k, j - 1, i + 1)) + V(k, j - 1, i))) *
                               (((P(k, j + 1, i + 1) + P(k, j + 1, i)) - P(k, j - 1, i + 1)) - P(k, j - 1, i))) /
                              ((((P(k, j + 1, i + 1) + P(k, j + 1, i)) + P(k, j - 1, i + 1)) + P(k, j - 1, i)) + SMALL)) /
                             (0.5 * (dyn[j - 1] + dyn[j]));
                double ujb1 = ((0.5 * dt) * U(k, j, i)) * (tz1 + ty1);
                double tz2 = (((0.25 * (((W(k, j + 1, i) + W(k, j, i)) + W(k - 1, j + 1, i)) + W(k - 1, j, i))) *
                               (((P(k + 1, j + 1, i) + P(k + 1, j, i)) - P(k - 1, j + 1, i)) - P(k - 1, j, i))) /
                              ((((P(k + 1, j + 1, i) + P(k + 1, j, i)) + P(k - 1, j + 1, i)) + P(k - 1, j, i)) + SMALL)) /
                             (0.5 * (dzn[k - 1] + dzn[k]));
                double tx2 = (((0.25 * (((U(k, j + 1, i) + U(k, j, i)) + U(k, j + 1, i - 1)) + U(k, j, i - 1))) *
                               (((P(k, j + 1, i + 1) + P(k, j, i + 1)) - P(k, j + 1, i - 1)) - P(k, j, i - 1))) /
                              ((((P(k, j + 1, i + 1) + P(k, j, i + 1)) + P(k, j + 1, i - 1)) + P(k, j, i - 1)) + SMALL)) /
                             (0.5 * (dxn[i - 1] + dxn[i]));
                double ujb2 = ((0.5 * dt) * V(k, j, i)) * (tz2 + tx2);
                double ty3 = (((0.25 * (((V(k + 1, j, i) + V(k, j, i)) + V(k + 1, j - 1, i)) + V(k, j - 1, i))) *
                               (((P(k + 1, j + 1, i) + P(k, j + 1, i)) - P(k + 1, j - 1, i)) - P(k, j - 1, i))) /
                              ((((P(k + 1, j + 1, i) + P(k, j + 1, i)) + P(k + 1, j - 1, i)) + P(k, j - 1, i)) + SMALL)) /
                             (0.5 * (dyn[j - 1] + dyn[j]));
                double tx3 = (((0.25 * (((U(k + 1, j, i) + U(k, j, i)) + U(k + 1, j, i - 1)) + U(k, j, i - 1))) *
                               (((P(k + 1, j, i + 1) + P(k, j, i + 1)) - P(k + 1, j, i - 1)) - P(k, j, i - 1))) /
                              ((((P(k + 1, j, i + 1) + P(k, j, i + 1)) + P(k + 1, j, i - 1)) + P(k, j, i - 1)) + SMALL)) /
                             (0.5 * (dxn[i - 1] + dxn[i]));
                double ujb3 = ((0.5 * dt) * W(k, j, i)) * (ty3 + tx3);
                /* :1482-1493 */
                const double uc = U(k, j, i), vc = V(k, j, i), wc = W(k, j, i);
                double usv = ((((fabs(uc) * dx[i]) - ((dt * uc) * uc)) * (P(k, j, i + 1) - P(k, j, i))) /
                              ((P(k, j, i + 1) + P(k, j, i)) + SMALL)) / dxn[i] - ujb1;
                double vsv = ((((fabs(vc) * dy[j]) - ((dt * vc) * vc)) * (P(k, j + 1, i) - P(k, j, i))) /
                              ((P(k, j + 1, i) + P(k, j, i)) + SMALL)) / dyn[j] - ujb2;
                double wsv = ((((fabs(wc) * dz[k]) - ((dt * wc) * wc)) * (P(k + 1, j, i) - P(k, j, i))) /
                              ((P(k + 1, j, i) + P(k, j, i)) + SMALL)) / dzn[k] - ujb3;
                /* :1502-1507 */
                usv = usv - ((0.25 * dt) * uc) * (((U(k, j, i + 1) - U(k, j, i - 1)) / dx[i - 1]) + u_div1);
                vsv = vsv - ((0.25 * dt) * vc) * (((V(k, j + 1, i) - V(k, j - 1, i)) / dy[j - 1]) + u_div2);
                wsv = wsv - ((0.25 * dt) * wc) * (((W(k + 1, j, i) - W(k - 1, j, i)) / dz[k - 1]) + u_div3);
                A3(us, k, j, i) = usv;
                A3(vs, k, j, i) = vsv;
                A3(ws, k, j, i) = wsv;
#undef U
#undef V
#undef W
#undef P
            }
}

/* eq 18 of Smolarkiewicz & Grabowski (1990) for one face */
static inline double limited_face(double x, double bdn_c, double bup_c, double bup_n,
                                  double bdn_n) {
    return min3(1.0, bdn_c, bup_n) * pos0(x) + min3(1.0, bup_c, bdn_n) * neg0(x);
}

/* Per-direction limiter of sfvt/advection_3d.F90:555-682 (vector variant
 * :1561-1688): bounds from the three cells along each axis, one pair of
 * unweighted denominators (grid spacings dx(i-1), dy(j-1), dz(k-1), small
 * inside the dt product) shared by the three directions. */
static void limiter_directional(const or_grid *g, double dt, const double *ut,
                                const double *vt, const double *wt, const double *psi,
                                const double *po, double *us, double *vs, double *ws) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    const double *dx = g->dx, *dy = g->dy, *dz = g->dz;
    double *bu[3], *bd[3];
    for (int d = 0; d < 3; ++d) {
        bu[d] = zalloc(N);
        bd[d] = zalloc(N);
    }
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                /* eqs 20a,b per direction: :555-591 */
                const int off[3][3] = {{0, 0, 1}, {0, 1, 0}, {1, 0, 0}}; /* (k,j,i) steps */
                double mx[3], mn[3];
                for (int d = 0; d < 3; ++d) {
                    const int sk = off[d][0], sj = off[d][1], si = off[d][2];
                    const double vals[6] = {
                        A3(psi, k - sk, j - sj, i - si), A3(psi, k, j, i),
                        A3(psi, k + sk, j + sj, i + si), A3(po, k - sk, j - sj, i - si),
                        A3(po, k, j, i),                 A3(po, k + sk, j + sj, i + si)};
                    mx[d] = vals[0];
                    mn[d] = vals[0];
                    for (int q = 1; q < 6; ++q) {
                        if (vals[q] > mx[d]) mx[d] = vals[q];
                        if (vals[q] < mn[d]) mn[d] = vals[q];
                    }
                }
                /* eqs 19a,b: :608-636 */
                const double pc = A3(po, k, j, i);
                const double d1 =
                    dt * ((((pos0(A3(ut, k, j, i - 1)) * A3(po, k, j, i - 1) -
                             neg0(A3(ut, k, j, i)) * A3(po, k, j, i + 1)) / dx[i - 1] +
                            (pos0(A3(vt, k, j - 1, i)) * A3(po, k, j - 1, i) -
                             neg0(A3(vt, k, j, i)) * A3(po, k, j + 1, i)) / dy[j - 1]) +
                           (pos0(A3(wt, k - 1, j, i)) * A3(po, k - 1, j, i) -
                            neg0(A3(wt, k, j, i)) * A3(po, k + 1, j, i)) / dz[k - 1]) +
                          SMALL);
                const double d2 =
                    dt * ((((pos0(A3(ut, k, j, i)) * pc - neg0(A3(ut, k, j, i - 1)) * pc) / dx[i - 1] +
                            (pos0(A3(vt, k, j, i)) * pc - neg0(A3(vt, k, j - 1, i)) * pc) / dy[j - 1]) +
                           (pos0(A3(wt, k, j, i)) * pc - neg0(A3(wt, k - 1, j, i)) * pc) / dz[k - 1]) +
                          SMALL);
                for (int d = 0; d < 3; ++d) {
                    A3(bu[d], k, j, i) = (mx[d] - pc) / d1;
                    A3(bd[d], k, j, i) = (pc - mn[d]) / d2;
                }
            }
    /* :649-659 */
    for (int d = 0; d < 3; ++d) {
        or_fill_along_dim(g, bu[d], 0.0, 0.0);
        or_fill_along_dim(g, bd[d], 0.0, 0.0);
    }
    /* eq 18: :666-682 */
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                A3(us, k, j, i) = limited_face(A3(ut, k, j, i), A3(bd[0], k, j, i), A3(bu[0], k, j, i),
                                               A3(bu[0], k, j, i + 1), A3(bd[0], k, j, i + 1));
                A3(vs, k, j, i) = limited_face(A3(vt, k, j, i), A3(bd[1], k, j, i), A3(bu[1], k, j, i),
                                               A3(bu[1], k, j + 1, i), A3(bd[1], k, j + 1, i));
                A3(ws, k, j, i) = limited_face(A3(wt, k, j, i), A3(bd[2], k, j, i), A3(bu[2], k, j, i),
                                               A3(bu[2], k + 1, j, i), A3(bd[2], k + 1, j, i));
            }
    for (int d = 0; d < 3; ++d) {
        free(bu[d]);
        free(bd[d]);
    }
}

/* Non-oscillatory limiter, pamm/sfvt/advection_3d.F90:542-675 (vector variant
 * :1576-1708).  psi = psi_store (start-of-step field), po = psi_old. */
static void limiter(const or_grid *g, double dt, const double *ut,
                    const double *vt, const double *wt, const double *psi,
                    const double *po, double *us, double *vs, double *ws) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    double *pmax = zalloc(N), *pmin = zalloc(N), *bup = zalloc(N), *bdn = zalloc(N);
    const double *dxn = g->dxn, *dyn = g->dyn, *dzn = g->dzn;
    const double *ra = g->rhoa, *rn = g->rhoan;
    /* eqs 20a,b of Smolarkiewicz & Grabowski (1990): :549-575 */
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double vals[14] = {
                    A3(psi, k, j, i), A3(psi, k, j, i - 1), A3(psi, k, j, i + 1),
                    A3(psi, k, j - 1, i), A3(psi, k, j + 1, i), A3(psi, k - 1, j, i),
                    A3(psi, k + 1, j, i), A3(po, k, j, i), A3(po, k, j, i - 1),
                    A3(po, k, j, i + 1), A3(po, k, j - 1, i), A3(po, k, j + 1, i),
                    A3(po, k - 1, j, i), A3(po, k + 1, j, i)};
                double mx = vals[0], mn = vals[0];
                for (int q = 1; q < 14; ++q) {
                    if (vals[q] > mx) mx = vals[q];
                    if (vals[q] < mn) mn = vals[q];
                }
                A3(pmax, k, j, i) = mx;
                A3(pmin, k, j, i) = mn;
            }
    /* eqs 19a,b: :585-616 */
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double pc = A3(po, k, j, i);
                double d1 = (dt * ((((pos0(A3(ut, k, j, i - 1)) * rn[k]) * A3(po, k, j, i - 1) -
                                     (neg0(A3(ut, k, j, i)) * rn[k]) * A3(po, k, j, i + 1)) / dxn[i] +
                                    ((pos0(A3(vt, k, j - 1, i)) * rn[k]) * A3(po, k, j - 1, i) -
                                     (neg0(A3(vt, k, j, i)) * rn[k]) * A3(po, k, j + 1, i)) / dyn[j]) +
                                   ((pos0(A3(wt, k - 1, j, i)) * ra[k - 1]) * A3(po, k - 1, j, i) -
                                    (neg0(A3(wt, k, j, i)) * ra[k]) * A3(po, k + 1, j, i)) / dzn[k])) +
                            SMALL;
                double d2 = dt * (((((pos0(A3(ut, k, j, i)) * rn[k]) * pc -
                                     (neg0(A3(ut, k, j, i - 1)) * rn[k]) * pc) / dxn[i] +
                                    ((pos0(A3(vt, k, j, i)) * rn[k]) * pc -
                                     (neg0(A3(vt, k, j - 1, i)) * rn[k]) * pc) / dyn[j]) +
                                   ((pos0(A3(wt, k, j, i)) * ra[k]) * pc -
                                    (neg0(A3(wt, k - 1, j, i)) * ra[k - 1]) * pc) / dzn[k]) +
                                  SMALL);
                A3(bup, k, j, i) = (rn[k] * (A3(pmax, k, j, i) - pc)) / d1;
                A3(bdn, k, j, i) = (rn[k] * (pc - A3(pmin, k, j, i))) / d2;
            }
    /* :624-627 */
    or_fill_along_dim(g, bup, 0.0, 0.0);
    or_fill_along_dim(g, bdn, 0.0, 0.0);
    /* eq 18: :634-656 */
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double uc = A3(ut, k, j, i), vc = A3(vt, k, j, i), wc = A3(wt, k, j, i);
                A3(us, k, j, i) = min3(1.0, A3(bdn, k, j, i), A3(bup, k, j, i + 1)) * pos0(uc) +
                                  min3(1.0, A3(bup, k, j, i), A3(bdn, k, j, i + 1)) * neg0(uc);
                A3(vs, k, j, i) = min3(1.0, A3(bdn, k, j, i), A3(bup, k, j + 1, i)) * pos0(vc) +
                                  min3(1.0, A3(bup, k, j, i), A3(bdn, k, j + 1, i)) * neg0(vc);
                A3(ws, k, j, i) = min3(1.0, A3(bdn, k, j, i), A3(bup, k + 1, j, i)) * pos0(wc) +
                                  min3(1.0, A3(bup, k, j, i), A3(bdn, k + 1, j, i)) * neg0(wc);
            }
    free(pmax); free(pmin); free(bup); free(bdn);
}

/* Shared body of mpdata_3d (:219-714) and mpdata_vec_3d (:1250-1753).  The
 * shift/early-exit differ and are done by the callers; fields[0] supplies
 * psi_old and psi_store (vector variant :1343-1346). */
static void mpdata_core(const or_grid *g, double dt, const double *u,
                        const double *v, const double *w, double *fields,
                        int nf, const double *lbc, const double *ubc, int kord,
                        int monotone, int neumann, int vec) {
    const int plain = vec && g_mpdata_variant == 1;
    const size_t N = NTOT(g);
    double *us1 = zalloc(N), *vs1 = zalloc(N), *ws1 = zalloc(N);
    double *us2 = zalloc(N), *vs2 = zalloc(N), *ws2 = zalloc(N); /* zeroed :296-298 */
    double *psi_store = NULL;
    const double *ut = u, *vt = v, *wt = w;
    double *ut_sav = us1, *vt_sav = vs1, *wt_sav = ws1;
    double *psi_old = fields; /* field 1 */
    if (kord > 1) {
        psi_store = zalloc(N);
        memcpy(psi_store, fields, N * sizeof(double));
    }
    int it2 = 0;
    for (int it = 1; it <= kord; ++it) {
        if (it > 1) {
            it2 += 1;
            if (plain)
                antidiffusive_plain(g, dt, ut, vt, wt, psi_old, ut_sav, vt_sav, wt_sav);
            else
                antidiffusive(g, dt, ut, vt, wt, psi_old, ut_sav, vt_sav, wt_sav);
            if (it2 % 2 == 1) {
                ut = us1; vt = vs1; wt = ws1;
                ut_sav = us2; vt_sav = vs2; wt_sav = ws2;
            } else {
                ut = us2; vt = vs2; wt = ws2;
                ut_sav = us1; vt_sav = vs1; wt_sav = ws1;
            }
            or_fill_full(g, (double *)ut, 0.0, 0.0);
            or_fill_full(g, (double *)vt, 0.0, 0.0);
            or_fill_full(g, (double *)wt, 0.0, 0.0);
        }
        if (it > 1 && monotone) {
            it2 += 1;
            if (g_mpdata_variant == 1)
                limiter_directional(g, dt, ut, vt, wt, psi_store, psi_old, ut_sav, vt_sav, wt_sav);
            else
                limiter(g, dt, ut, vt, wt, psi_store, psi_old, ut_sav, vt_sav, wt_sav);
            ut = us2; vt = vs2; wt = ws2;
            ut_sav = us1; vt_sav = vs1; wt_sav = ws1;
            or_fill_full(g, (double *)ut, 0.0, 0.0);
            or_fill_full(g, (double *)vt, 0.0, 0.0);
            or_fill_full(g, (double *)wt, 0.0, 0.0);
        }
        for (int n = 0; n < nf; ++n) {
            double *f = fields + (size_t)n * N;
            or_upwind3d(g, dt, ut, vt, wt, f, neumann);
            if (it <= kord && kord >= 1) or_fill_full(g, f, lbc[n], ubc[n]);
        }
    }
    free(us1); free(vs1); free(ws1); free(us2); free(vs2); free(ws2);
    free(psi_store);
}

/* mpdata_3d, pamm/sfvt/advection_3d.F90:219-714 */
int or_mpdata3d(const or_grid *g, double dt, const double *u, const double *v,
                const double *w, double *psi, double *lbc, double *ubc,
                int kord, int monotone, int neumann) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    /* minimum over the whole array including halos, and lbc/ubc: :275-276 */
    double m = psi[0];
    for (size_t q = 1; q < N; ++q)
        if (psi[q] < m) m = psi[q];
    if (*lbc < m) m = *lbc;
    if (*ubc < m) m = *ubc;
    for (size_t q = 0; q < N; ++q) psi[q] = psi[q] - m;
    *lbc = *lbc - m;
    *ubc = *ubc - m;
    /* interior sum, column-major order: :282-283 */
    double s = 0.0;
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) s = s + A3(psi, k, j, i);
    if (s < SMALL) { /* :284-289 */
        for (size_t q = 0; q < N; ++q) psi[q] = psi[q] + m;
        *lbc = *lbc + m;
        *ubc = *ubc + m;
        return 1;
    }
    mpdata_core(g, dt, u, v, w, psi, 1, lbc, ubc, kord, monotone, neumann, 0);
    for (size_t q = 0; q < N; ++q) psi[q] = psi[q] + m; /* :711-713 */
    *lbc = *lbc + m;
    *ubc = *ubc + m;
    return 0;
}

/* mpdata_vec_3d, pamm/sfvt/advection_3d.F90:1250-1753 */
int or_mpdata_vec3d(const or_grid *g, double dt, const double *u,
                    const double *v, const double *w, double *psi, int nq,
                    double *lbc, double *ubc, int kord, int monotone,
                    int neumann) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    double *mg = zalloc((size_t)nq);
    /* per-field minimum over the interior only, and lbc/ubc: :1306-1312 */
    for (int n = 0; n < nq; ++n) {
        double *f = psi + (size_t)n * N;
        double m = A3(f, 1, 1, 1);
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j)
                for (int k = 1; k <= kp; ++k)
                    if (A3(f, k, j, i) < m) m = A3(f, k, j, i);
        if (lbc[n] < m) m = lbc[n];
        if (ubc[n] < m) m = ubc[n];
        mg[n] = m;
        for (size_t q = 0; q < N; ++q) f[q] = f[q] - m;
        lbc[n] = lbc[n] - m;
        ubc[n] = ubc[n] - m;
    }
    /* sum of field 1: :1314-1315 */
    double s = 0.0;
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) s = s + A3(psi, k, j, i);
    if (s < SMALL) {
        for (int n = 0; n < nq; ++n) {
            double *f = psi + (size_t)n * N;
            for (size_t q = 0; q < N; ++q) f[q] = f[q] + mg[n];
            lbc[n] = lbc[n] + mg[n];
            ubc[n] = ubc[n] + mg[n];
        }
        free(mg);
        return 1;
    }
    mpdata_core(g, dt, u, v, w, psi, nq, lbc, ubc, kord, monotone, neumann, 1);
    for (int n = 0; n < nq; ++n) {
        double *f = psi + (size_t)n * N;
        for (size_t q = 0; q < N; ++q) f[q] = f[q] + mg[n];
        lbc[n] = lbc[n] + mg[n];
        ubc[n] = ubc[n] + mg[n];
    }
    free(mg);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* vertical-only advection (used by PAMM sedimentation)                     */
/* ------------------------------------------------------------------------ */

/* first_order_upstream_vert_3d, pamm/sfvt/advection_3d.F90:737-791.
 * neumann 2: the bottom face of level 1 takes w(1) (:762-766); neumann 1:
 * zero-gradient fluxes at the floor and lid (:780-787). */
void or_upwind_vert3d(const or_grid *g, double dt, const double *w, double *psi,
                      int neumann) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *dzn = g->dzn, *ra = g->rhoa, *rn = g->rhoan;
    const size_t n = (size_t)ip * jp * kp;
    double *fr = zalloc(n), *fl = zalloc(n);
#define F3(f, k, j, i) f[((size_t)((i)-1) * jp + ((j)-1)) * kp + ((k)-1)]
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double wk = A3(w, k, j, i);
                const double wm = (k == 1 && neumann == 2) ? wk : A3(w, k - 1, j, i);
                F3(fr, k, j, i) = ((((wk + fabs(wk)) * ra[k]) * A3(psi, k, j, i) +
                                    ((wk - fabs(wk)) * ra[k]) * A3(psi, k + 1, j, i)) * dt) /
                                  ((2.0 * dzn[k]) * rn[k]);
                F3(fl, k, j, i) = ((((wm + fabs(wm)) * ra[k - 1]) * A3(psi, k - 1, j, i) +
                                    ((wm - fabs(wm)) * ra[k - 1]) * A3(psi, k, j, i)) * dt) /
                                  ((2.0 * dzn[k]) * rn[k]);
            }
    if (neumann == 1) {
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j) F3(fl, 1, j, i) = F3(fr, 1, j, i);
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j) F3(fr, kp, j, i) = F3(fl, kp, j, i);
    }
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k)
                A3(psi, k, j, i) = A3(psi, k, j, i) - (F3(fr, k, j, i) - F3(fl, k, j, i));
#undef F3
    free(fr);
    free(fl);
}

/* Vertical antidiffusive velocity, pamm/sfvt/advection_3d.F90:942-958
 * (vector variant :1905-1921): eq 13 of Smolarkiewicz (1984) with the
 * cross terms zero and the divergence term of eq 38. */
static void antidiffusive_vert(const or_grid *g, double dt, const double *wt,
                               const double *po, double *ws, int plain) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *dz = g->dz, *dzn = g->dzn, *ra = g->rhoa, *rn = g->rhoan;
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double gb3 = 0.5 * (rn[k] + rn[k + 1]);
                const double u_div3 = 0.0, u_j_bar3 = 0.0;
                const double wc = A3(wt, k, j, i);
                const double p0 = A3(po, k, j, i), p1 = A3(po, k + 1, j, i);
                double v;
                if (plain) { /* sfvt/advection_3d.F90:1897-1909 */
                    v = ((((fabs(wc) * dz[k] - (dt * wc) * wc) * (p1 - p0)) / ((p1 + p0) + SMALL)) /
                         dzn[k]) - u_j_bar3;
                    v = v - ((0.25 * dt) * wc) *
                                (((A3(wt, k + 1, j, i) - A3(wt, k - 1, j, i)) / dz[k - 1]) + u_div3);
                } else {
                    v = ((((fabs(wc) * dz[k] - ((dt * (ra[k] / gb3)) * wc) * wc) * (p1 - p0)) /
                          ((p1 + p0) + SMALL)) / dzn[k]) - u_j_bar3;
                    v = v - ((0.25 * dt) * wc) *
                                (((ra[k + 1] * A3(wt, k + 1, j, i) - ra[k - 1] * A3(wt, k - 1, j, i)) /
                                  (dz[k - 1] * gb3)) + u_div3);
                }
                A3(ws, k, j, i) = v;
            }
}

/* Vertical limiter, pamm/sfvt/advection_3d.F90:1000-1080 (vector variant
 * :1964-2036); psi = psi_store, po = psi_old. */
static void limiter_vert(const or_grid *g, double dt, const double *wt,
                         const double *psi, const double *po, double *ws) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    const double *dz = g->dz, *dzn = g->dzn, *ra = g->rhoa, *rn = g->rhoan;
    double *bup = zalloc(N), *bdn = zalloc(N);
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double vals[6] = {A3(psi, k - 1, j, i), A3(psi, k, j, i),
                                        A3(psi, k + 1, j, i), A3(po, k - 1, j, i),
                                        A3(po, k, j, i),      A3(po, k + 1, j, i)};
                double mx = vals[0], mn = vals[0];
                for (int q = 1; q < 6; ++q) {
                    if (vals[q] > mx) mx = vals[q];
                    if (vals[q] < mn) mn = vals[q];
                }
                const double pc = A3(po, k, j, i);
                const double wm = A3(wt, k - 1, j, i), wc = A3(wt, k, j, i);
                if (g_mpdata_variant == 1) { /* sfvt/advection_3d.F90:1051-1062 */
                    const double d1 = dt * ((pos0(wm) * A3(po, k - 1, j, i) -
                                             neg0(wc) * A3(po, k + 1, j, i)) / dz[k - 1] + SMALL);
                    const double d2 = dt * ((pos0(wc) * pc - neg0(wm) * pc) / dz[k - 1] + SMALL);
                    A3(bup, k, j, i) = (mx - pc) / d1;
                    A3(bdn, k, j, i) = (pc - mn) / d2;
                    continue;
                }
                const double d1 = (dt * ((pos0(wm) * ra[k - 1]) * A3(po, k - 1, j, i) -
                                         (neg0(wc) * ra[k]) * A3(po, k + 1, j, i))) / dzn[k] +
                                  SMALL;
                const double d2 =
                    (dt * ((pos0(wc) * ra[k]) * pc - (neg0(wm) * ra[k - 1]) * pc)) / dzn[k] + SMALL;
                A3(bup, k, j, i) = (rn[k] * (mx - pc)) / d1;
                A3(bdn, k, j, i) = (rn[k] * (pc - mn)) / d2;
            }
    or_fill_along_z(g, bup, 0.0, 0.0);
    or_fill_along_z(g, bdn, 0.0, 0.0);
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double wc = A3(wt, k, j, i);
                A3(ws, k, j, i) = min3(1.0, A3(bdn, k, j, i), A3(bup, k + 1, j, i)) * pos0(wc) +
                                  min3(1.0, A3(bup, k, j, i), A3(bdn, k + 1, j, i)) * neg0(wc);
            }
    free(bup);
    free(bdn);
}

/* Pass loop shared by mpdata_vert_3d (:896-1115) and mpdata_vec_vert_3d
 * (:1859-2071): fields[0] supplies psi_old and psi_store. */
static void mpdata_vert_core(const or_grid *g, double dt, const double *w, double *fields,
                             int nf, const double *lbc, const double *ubc, int kord,
                             int monotone, int neumann, int vec) {
    const int plain = vec && g_mpdata_variant == 1;
    const size_t N = NTOT(g);
    double *ws1 = zalloc(N), *ws2 = zalloc(N); /* w_store2 zeroed :876 */
    double *psi_store = NULL;
    const double *wt = w;
    double *wt_sav = ws1;
    if (kord > 1) {
        psi_store = zalloc(N); /* field 1 is the one read (:1851-1855) */
        memcpy(psi_store, fields, N * sizeof(double));
    }
    int it2 = 0;
    for (int it = 1; it <= kord; ++it) {
        if (it > 1) {
            it2 += 1;
            antidiffusive_vert(g, dt, wt, fields, wt_sav, plain);
            if (it2 % 2 == 1) {
                wt = ws1;
                wt_sav = ws2;
            } else {
                wt = ws2;
                wt_sav = ws1;
            }
            or_fill_along_z(g, (double *)wt, 0.0, 0.0);
        }
        if (it > 1 && monotone) {
            it2 += 1;
            limiter_vert(g, dt, wt, psi_store, fields, wt_sav);
            wt = ws2;
            wt_sav = ws1;
            or_fill_along_z(g, (double *)wt, 0.0, 0.0);
        }
        for (int n = 0; n < nf; ++n) {
            double *f = fields + (size_t)n * N;
            or_upwind_vert3d(g, dt, wt, f, neumann);
            if (it <= kord && kord >= 1) or_fill_along_z(g, f, lbc[n], ubc[n]);
        }
    }
    free(ws1);
    free(ws2);
    free(psi_store);
}

/* mpdata_vert_3d, pamm/sfvt/advection_3d.F90:820-1126 */
int or_mpdata_vert3d(const or_grid *g, double dt, const double *w, double *psi,
                     double *lbc, double *ubc, int kord, int monotone, int neumann) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    /* minimum over the whole array and lbc/ubc: :861-866 */
    double m = psi[0];
    for (size_t q = 1; q < N; ++q)
        if (psi[q] < m) m = psi[q];
    if (*lbc < m) m = *lbc;
    if (*ubc < m) m = *ubc;
    for (size_t q = 0; q < N; ++q) psi[q] = psi[q] - m;
    *lbc = *lbc - m;
    *ubc = *ubc - m;
    double s = 0.0; /* :868-869 */
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) s = s + A3(psi, k, j, i);
    if (s < SMALL) { /* :870-875 */
        for (size_t q = 0; q < N; ++q) psi[q] = psi[q] + m;
        *lbc = *lbc + m;
        *ubc = *ubc + m;
        return 1;
    }
    mpdata_vert_core(g, dt, w, psi, 1, lbc, ubc, kord, monotone, neumann, 0);
    for (size_t q = 0; q < N; ++q) psi[q] = psi[q] + m; /* :1123-1125 */
    *lbc = *lbc + m;
    *ubc = *ubc + m;
    return 0;
}

/* mpdata_vec_vert_3d, pamm/sfvt/advection_3d.F90:1779-2087 */
int or_mpdata_vec_vert3d(const or_grid *g, double dt, const double *w, double *psi, int nq,
                         double *lbc, double *ubc, int kord, int monotone, int neumann) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    double *mg = zalloc((size_t)nq);
    /* per-field minimum over the interior and lbc/ubc: :1821-1827 */
    for (int n = 0; n < nq; ++n) {
        double *f = psi + (size_t)n * N;
        double m = A3(f, 1, 1, 1);
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j)
                for (int k = 1; k <= kp; ++k)
                    if (A3(f, k, j, i) < m) m = A3(f, k, j, i);
        if (lbc[n] < m) m = lbc[n];
        if (ubc[n] < m) m = ubc[n];
        mg[n] = m;
        for (size_t q = 0; q < N; ++q) f[q] = f[q] - m;
        lbc[n] = lbc[n] - m;
        ubc[n] = ubc[n] - m;
    }
    double s = 0.0; /* field 1: :1829-1830 */
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) s = s + A3(psi, k, j, i);
    if (s < SMALL) {
        for (int n = 0; n < nq; ++n) {
            double *f = psi + (size_t)n * N;
            for (size_t q = 0; q < N; ++q) f[q] = f[q] + mg[n];
            lbc[n] = lbc[n] + mg[n];
            ubc[n] = ubc[n] + mg[n];
        }
        free(mg);
        return 1;
    }
    mpdata_vert_core(g, dt, w, psi, nq, lbc, ubc, kord, monotone, neumann, 1);
    for (int n = 0; n < nq; ++n) {
        double *f = psi + (size_t)n * N;
        for (size_t q = 0; q < N; ++q) f[q] = f[q] + mg[n];
        lbc[n] = lbc[n] + mg[n];
        ubc[n] = ubc[n] + mg[n];
    }
    free(mg);
    return 0;
}

/* mpdata_3d_add, pamm/sfvt/advection_3d.F90:139-190: a 1-D profile added
 * to every column (halo levels included), mpdata_3d, the profile removed. */
int or_mpdata3d_add(const or_grid *g, double dt, const double *u, const double *v,
                    const double *w, double *psi, const double *psi_1d, double *lbc,
                    double *ubc, int kord, int monotone, int neumann) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    for (int k = 0; k <= kp + 1; ++k)
        for (int i = 0; i <= ip + 1; ++i)
            for (int j = 0; j <= jp + 1; ++j) A3(psi, k, j, i) = A3(psi, k, j, i) + psi_1d[k];
    const int e = or_mpdata3d(g, dt, u, v, w, psi, lbc, ubc, kord, monotone, neumann);
    for (int k = 0; k <= kp + 1; ++k)
        for (int i = 0; i <= ip + 1; ++i)
            for (int j = 0; j <= jp + 1; ++j) A3(psi, k, j, i) = A3(psi, k, j, i) - psi_1d[k];
    return e;
}

/* adv_ref_state, pamm/sfvt/advection_3d.F90:1148-1218 (halo columns too) */
void or_adv_ref_state(const or_grid *g, double dt, const double *w,
                      double *psi, const double *ref) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *dz = g->dz, *ra = g->rhoa, *rn = g->rhoan;
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j)
            for (int k = 1; k <= kp; ++k)
                A3(psi, k, j, i) =
                    A3(psi, k, j, i) -
                    (dt * ((((0.5 / dz[k - 1]) * A3(w, k - 1, j, i)) * ra[k - 1]) * (ref[k] - ref[k - 1]) +
                           (((0.5 / dz[k]) * A3(w, k, j, i)) * ra[k]) * (ref[k + 1] - ref[k]))) /
                        rn[k];
}

/* ------------------------------------------------------------------------ */
/* diagnostics                                                              */
/* ------------------------------------------------------------------------ */

static double level_mean(const or_grid *g, const double *a, int k) {
    double s = 0.0;
    for (int i = 1; i <= g->ip; ++i)
        for (int j = 1; j <= g->jp; ++j) s = s + A3(a, k, j, i);
    return s / (double)(g->ip * g->jp);
}

/* horizontal_means, diagnostics.f90:31-107.  The reference's MPI_Allreduce
 * count (4+nq)*(kpp+2) is shorter than its buffer (diagnostics.f90:50,82), so
 * with nq=0 the top thbar entries are left undefined; here every level of
 * every mean is computed. */
void or_hmeans(const or_grid *g, const double *u, const double *v,
               const double *w, const double *th, const double *q, int nq,
               double *ubar, double *vbar, double *wbar, double *thbar,
               double *qbar) {
    const size_t N = NTOT(g);
    for (int k = 0; k <= g->kp + 1; ++k) {
        if (ubar) ubar[k] = level_mean(g, u, k);
        if (vbar) vbar[k] = level_mean(g, v, k);
        if (wbar) wbar[k] = level_mean(g, w, k);
        thbar[k] = level_mean(g, th, k);
        for (int n = 0; n < nq && q && qbar; ++n)
            qbar[(size_t)n * (g->kp + 2) + k] = level_mean(g, q + (size_t)n * N, k);
    }
}

/* ------------------------------------------------------------------------ */
/* momentum sources and Poisson right-hand side                            */
/* ------------------------------------------------------------------------ */

/* advective tendencies and buoyancy of sources, dynamics.f90:743-797 */
static void sources_advective(const or_grid *g, const double *u, const double *v,
                              const double *w, const double *th, const double *theta,
                              const double *thbar, double *su, double *sv, double *sw) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *dx = g->dx, *dy = g->dy, *dz = g->dz;
    const double *dxn = g->dxn, *dyn = g->dyn, *dzn = g->dzn;
    const double *ra = g->rhoa, *rn = g->rhoan;
    /* su: :743-757 */
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                double X = (A3(u, k, j, i - 1) * (A3(u, k, j, i) + A3(u, k, j, i - 1)) -
                            A3(u, k, j, i + 1) * (A3(u, k, j, i) + A3(u, k, j, i + 1))) /
                           (2.0 * (dx[i] + dx[i - 1]));
                double Y = (A3(u, k, j - 1, i) * (A3(v, k, j - 1, i) + A3(v, k, j - 1, i + 1)) -
                            A3(u, k, j + 1, i) * (A3(v, k, j, i) + A3(v, k, j, i + 1))) /
                           (2.0 * (dyn[j] + dyn[j - 1]));
                double Z = (A3(u, k - 1, j, i) * (A3(w, k - 1, j, i) + A3(w, k - 1, j, i + 1)) -
                            A3(u, k + 1, j, i) * (A3(w, k, j, i) + A3(w, k, j, i + 1))) /
                           (2.0 * (dzn[k] + dzn[k - 1]));
                A3(su, k, j, i) = rn[k] * ((X + Y) + Z);
            }
    /* sv: :761-775 */
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                double X = (A3(v, k, j, i - 1) * (A3(u, k, j, i - 1) + A3(u, k, j + 1, i - 1)) -
                            A3(v, k, j, i + 1) * (A3(u, k, j, i) + A3(u, k, j + 1, i))) /
                           (2.0 * (dxn[i] + dxn[i - 1]));
                double Y = (A3(v, k, j - 1, i) * (A3(v, k, j, i) + A3(v, k, j - 1, i)) -
                            A3(v, k, j + 1, i) * (A3(v, k, j, i) + A3(v, k, j + 1, i))) /
                           (2.0 * (dy[j] + dy[j - 1]));
                double Z = (A3(v, k - 1, j, i) * (A3(w, k - 1, j, i) + A3(w, k - 1, j + 1, i)) -
                            A3(v, k + 1, j, i) * (A3(w, k, j, i) + A3(w, k, j + 1, i))) /
                           (2.0 * (dzn[k] + dzn[k - 1]));
                A3(sv, k, j, i) = rn[k] * ((X + Y) + Z);
            }
    /* sw with buoyancy: :779-797 */
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                double X = (A3(w, k, j, i - 1) * (A3(u, k, j, i - 1) + A3(u, k + 1, j, i - 1)) -
                            A3(w, k, j, i + 1) * (A3(u, k, j, i) + A3(u, k + 1, j, i))) /
                           (2.0 * (dxn[i] + dxn[i - 1]));
                double Y = (A3(w, k, j - 1, i) * (A3(v, k, j - 1, i) + A3(v, k + 1, j - 1, i)) -
                            A3(w, k, j + 1, i) * (A3(v, k, j, i) + A3(v, k + 1, j, i))) /
                           (2.0 * (dyn[j] + dyn[j - 1]));
                double Z = (A3(w, k - 1, j, i) * (A3(w, k, j, i) + A3(w, k - 1, j, i)) -
                            A3(w, k + 1, j, i) * (A3(w, k, j, i) + A3(w, k + 1, j, i))) /
                           (2.0 * (dz[k] + dz[k - 1]));
                double B = (GRAV * (A3(th, k, j, i) + A3(th, k + 1, j, i))) /
                           (((2.0 * theta[k]) + thbar[k]) + thbar[k + 1]);
                A3(sw, k, j, i) = ra[k] * (((X + Y) + Z) + B);
            }
}

/* floor and lid rows and the Poisson right-hand side, dynamics.f90:922-998 */
static void sources_finish(const or_grid *g, double *su, double *sv, double *sw, double *rhs) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *dx = g->dx, *dy = g->dy, *dz = g->dz;
    /* floor and lid rows: :922-973 */
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) {
            A3(su, 0, j, i) = -A3(su, 1, j, i);
            A3(sv, 0, j, i) = -A3(sv, 1, j, i);
            A3(sw, 0, j, i) = 0.0;
        }
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) {
            A3(sw, kp, j, i) = 0.0; A3(sw, kp + 1, j, i) = 0.0;
            A3(su, kp, j, i) = 0.0; A3(su, kp + 1, j, i) = 0.0;
            A3(sv, kp, j, i) = 0.0; A3(sv, kp + 1, j, i) = 0.0;
        }
    /* Poisson right-hand side: :977-985, and the k=0 row :988-998 */
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j) {
            for (int k = 1; k <= kp + 1; ++k)
                A3(rhs, k, j, i) = ((A3(su, k, j, i) - A3(su, k, j, i - 1)) / dx[i - 1] +
                                    (A3(sv, k, j, i) - A3(sv, k, j - 1, i)) / dy[j - 1]) +
                                   (A3(sw, k, j, i) - A3(sw, k - 1, j, i)) / dz[k - 1];
        }
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            A3(rhs, 0, j, i) = ((A3(su, 0, j, i) - A3(su, 0, j, i - 1)) / dx[i - 1] +
                                (A3(sv, 0, j, i) - A3(sv, 0, j - 1, i)) / dy[j - 1]) +
                               (A3(sw, 0, j, i) + A3(sw, 1, j, i)) / dz[1];
}

/* sources, dynamics.f90:639-1000 (viscous, damping_layer, forcing and
 * divergence all off) */
void or_sources(const or_grid *g, const double *u, const double *v,
                const double *w, const double *th, const double *theta,
                const double *thbar, double *su, double *sv, double *sw,
                double *sth, double *rhs) {
    sources_advective(g, u, v, w, th, theta, thbar, su, sv, sw);
    /* inviscid branch: exchange_full on su, sv, sw, sth (:909-916) */
    or_fill_full(g, su, 0.0, 0.0);
    or_fill_full(g, sv, 0.0, 0.0);
    or_fill_full(g, sw, 0.0, 0.0);
    or_fill_full(g, sth, 0.0, 0.0);
    sources_finish(g, su, sv, sw, rhs);
}

/* ------------------------------------------------------------------------ */
/* subgrid model (sgm/subgrid_3d.f90) and the other terms of sources        */
/* ------------------------------------------------------------------------ */

/* parameters of subgrid_3d.f90:13-16 */
static const double SG_PRN = 0.95, SG_SUBB = 40.0, SG_SUBC = 16.0, SG_SUBG = 1.2,
                    SG_SUBH = 0.0, SG_RIC = 0.25, SG_SMALL = 1e-10, SG_KVON = 0.4;
#define SG_SUBA (1.0 / SG_PRN)
#define SG_SUBF SG_SUBA
static inline double sq2(double x) { return x * x; }
/* x**4._wp: evaluated as two squarings (the reference's libm pow is within
 * an ulp of it) */
static inline double pow4(double x) {
    const double t = x * x;
    return t * t;
}

/* 2 S_ij S_ij on the w points, subgrid_3d.f90:91-131 (the sixth term of
 * the 2 s12 s12 average repeats the second, as written there) */
static double strain_rate(const or_grid *g, const double *u, const double *v, const double *w,
                          int k, int j, int i) {
    const double *dx = g->dx, *dy = g->dy, *dz = g->dz;
    const double *dxn = g->dxn, *dyn = g->dyn, *dzn = g->dzn;
    double s = sq2((A3(u, k + 1, j, i) - A3(u, k + 1, j, i - 1)) / dx[i - 1]) +
               sq2((A3(u, k, j, i) - A3(u, k, j, i - 1)) / dx[i - 1]);
    s = (s + sq2((A3(v, k + 1, j, i) - A3(v, k + 1, j - 1, i)) / dy[j - 1])) +
        sq2((A3(v, k, j, i) - A3(v, k, j - 1, i)) / dy[j - 1]);
    s = (s + sq2((A3(w, k, j, i) - A3(w, k - 1, j, i)) / dz[k - 1])) +
        sq2((A3(w, k + 1, j, i) - A3(w, k, j, i)) / dz[k]);
    s = s + 0.5 * (sq2((A3(u, k + 1, j, i) - A3(u, k, j, i)) / dzn[k] +
                       (A3(w, k, j, i + 1) - A3(w, k, j, i)) / dxn[i]) +
                   sq2((A3(u, k + 1, j, i - 1) - A3(u, k, j, i - 1)) / dz[k] +
                       (A3(w, k, j, i) - A3(w, k, j, i - 1)) / dxn[i - 1]));
    s = s + 0.5 * (sq2((A3(w, k, j, i) - A3(w, k, j - 1, i)) / dyn[j - 1] +
                       (A3(v, k + 1, j - 1, i) - A3(v, k, j - 1, i)) / dzn[k]) +
                   sq2((A3(w, k, j + 1, i) - A3(w, k, j, i)) / dyn[j] +
                       (A3(v, k + 1, j, i) - A3(v, k, j, i)) / dzn[k]));
    const double p1 = (A3(u, k, j, i - 1) - A3(u, k, j - 1, i - 1)) / dyn[j - 1] +
                      (A3(v, k, j - 1, i) - A3(v, k, j - 1, i - 1)) / dxn[i - 1];
    const double p2 = (A3(u, k, j + 1, i - 1) - A3(u, k, j, i - 1)) / dyn[j] +
                      (A3(v, k, j, i) - A3(v, k, j, i - 1)) / dxn[i - 1];
    const double p3 = (A3(u, k, j, i) - A3(u, k, j - 1, i)) / dyn[j - 1] +
                      (A3(v, k, j - 1, i + 1) - A3(v, k, j - 1, i)) / dxn[i];
    const double p4 = (A3(u, k, j + 1, i) - A3(u, k, j, i)) / dyn[j] +
                      (A3(v, k, j, i + 1) - A3(v, k, j, i)) / dxn[i];
    const double p5 = (A3(u, k + 1, j, i - 1) - A3(u, k + 1, j - 1, i - 1)) / dyn[j - 1] +
                      (A3(v, k + 1, j - 1, i) - A3(v, k + 1, j - 1, i - 1)) / dxn[i - 1];
    const double p6 = (A3(u, k, j + 1, i - 1) - A3(u, k, j, i - 1)) / dyn[j] +
                      (A3(v, k, j, i) - A3(v, k, j, i - 1)) / dxn[i - 1];
    const double p7 = (A3(u, k + 1, j, i) - A3(u, k + 1, j - 1, i)) / dyn[j - 1] +
                      (A3(v, k + 1, j - 1, i + 1) - A3(v, k + 1, j - 1, i)) / dxn[i];
    const double p8 = (A3(u, k + 1, j + 1, i) - A3(u, k + 1, j, i)) / dyn[j] +
                      (A3(v, k + 1, j, i + 1) - A3(v, k + 1, j, i)) / dxn[i];
    s = s + 0.125 * (((((((sq2(p1) + sq2(p2)) + sq2(p3)) + sq2(p4)) + sq2(p5)) + sq2(p6)) +
                      sq2(p7)) + sq2(p8));
    return s;
}

/* richardson, subgrid_3d.f90:465-491: the gradient Richardson number and
 * the stability functions fm, fh.  1.e-15 is a default-real literal. */
static void richardson_fns(double strain, double th_k, double th_kp1, double thn_k,
                           double thn_kp1, double dzn_k, double *fm, double *fh) {
    const double rip = ((GRAV * (th_kp1 / thn_kp1 - th_k / thn_k)) / dzn_k) /
                       (strain + (double)1.e-15f);
    if (rip < 0.0) {
        *fm = sqrt(1.0 - SG_SUBC * rip); /* (.)**0.5_wp */
        *fh = SG_SUBA * sqrt(1.0 - SG_SUBB * rip);
    } else if (rip >= 0.0 && rip < SG_RIC) {
        const double x = pow4(1.0 - rip / SG_RIC);
        *fm = x * (1.0 - SG_SUBH * rip);
        *fh = (SG_SUBF * x) * (1.0 - SG_SUBG * rip);
    } else {
        *fm = 0.0;
        *fh = 0.0;
    }
}

/* monin_obukhov, subgrid_3d.f90:373-433: surface-layer viscosities at k=0
 * (Jacobson's non-iterative scheme) */
static void monin_obukhov(const or_grid *g, const or_physics *ph, const double *th,
                          const double *u, const double *v, double *vism, double *vist) {
    const int ip = g->ip, jp = g->jp;
    const double kv2 = SG_KVON * SG_KVON, L = ph->log_z1_z0, z1 = ph->z[1];
    const double z0 = ph->z0, z0th = ph->z0th, th0 = ph->theta[0], thn1 = ph->thetan[1];
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j) {
            const double vmag = 0.5 * sqrt(sq2(A3(u, 1, j, i) + A3(u, 1, j, i - 1)) +
                                           sq2(A3(v, 1, j, i) + A3(v, 1, j - 1, i))) +
                                SG_SMALL;
            const double dth = (A3(th, 1, j, i) + thn1) - th0;
            const double rib = ((GRAV * dth) * sq2(z1 - z0)) / ((th0 * sq2(vmag)) * (z1 - z0th));
            double gm, gh;
            if (rib <= 0.0) {
                const double q = sqrt(fabs((rib * z1) / z0));
                gm = 1.0 - (9.4 * rib) / (1.0 + ((70.0 * kv2) * q) / sq2(L));
                gh = 1.0 - (9.4 * rib) / (1.0 + ((50.0 * kv2) * q) / sq2(L));
            } else {
                gm = 1.0 / sq2(1.0 + 4.7 * rib);
                gh = 1.0 / sq2(1.0 + 4.7 * rib);
            }
            const double ustar = ((SG_KVON * vmag) / L) * sqrt(gm);
            const double thstar = (((kv2 * vmag) * dth) / ((ustar * SG_PRN) * sq2(L))) * gh;
            A3(vism, 0, j, i) = (sq2(ustar) * (z1 - z0)) / vmag;
            A3(vist, 0, j, i) = (ustar * thstar) / (dth / (z1 - z0th) + SG_SMALL);
        }
}

/* viscous flux divergence of a scalar (sth :288-312, sq :322-349); vt is
 * vist, already multiplied by rhoa */
static void diffuse_scalar(const or_grid *g, const double *vt, const double *T, double *sT) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *dx = g->dx, *dy = g->dy, *dz = g->dz, *dxn = g->dxn, *dyn = g->dyn;
    const double *rn = g->rhoan;
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double ax = ((A3(vt, k, j, i) + A3(vt, k, j, i - 1)) + A3(vt, k - 1, j, i)) +
                                  A3(vt, k - 1, j, i - 1);
                const double bx = ((A3(vt, k, j, i + 1) + A3(vt, k, j, i)) + A3(vt, k - 1, j, i + 1)) +
                                  A3(vt, k - 1, j, i);
                const double ay = ((A3(vt, k, j, i) + A3(vt, k, j - 1, i)) + A3(vt, k - 1, j, i)) +
                                  A3(vt, k - 1, j - 1, i);
                const double by = ((A3(vt, k, j + 1, i) + A3(vt, k, j, i)) + A3(vt, k - 1, j + 1, i)) +
                                  A3(vt, k - 1, j, i);
                const double tc = A3(T, k, j, i);
                const double X = (0.25 * ((ax * (A3(T, k, j, i - 1) - tc)) / dxn[i - 1] -
                                          (bx * (tc - A3(T, k, j, i + 1))) / dxn[i])) /
                                 dx[i - 1];
                const double Y = (0.25 * ((ay * (A3(T, k, j - 1, i) - tc)) / dyn[j - 1] -
                                          (by * (tc - A3(T, k, j + 1, i))) / dyn[j])) /
                                 dy[j - 1];
                const double Z = ((A3(vt, k - 1, j, i) * (A3(T, k - 1, j, i) - tc)) / dz[k - 1] -
                                  (A3(vt, k, j, i) * (tc - A3(T, k + 1, j, i))) / dz[k]) /
                                 dz[k - 1];
                A3(sT, k, j, i) = A3(sT, k, j, i) + ((X + Y) + Z) / rn[k];
            }
}

/* calculate_subgrid_3d, sgm/subgrid_3d.f90:38-357 */
void or_subgrid(const or_grid *g, const or_physics *ph, const double *u, const double *v,
                const double *w, const double *zu, const double *zv, const double *zw,
                const double *th, const double *q, double *vism, double *vist, double *viss,
                double *strain, double *su, double *sv, double *sw, double *sth, double *sq) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    const double *dx = g->dx, *dy = g->dy, *dz = g->dz;
    const double *dxn = g->dxn, *dyn = g->dyn, *dzn = g->dzn;
    const double *ra = g->rhoa, *rn = g->rhoan, *thn = ph->thetan;
    /* strain rate :91-131, Richardson functions :137, viscosities :143-159 */
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double s = strain_rate(g, u, v, w, k, j, i);
                double fm, fh;
                richardson_fns(s, A3(th, k, j, i), A3(th, k + 1, j, i), thn[k], thn[k + 1],
                               dzn[k], &fm, &fh);
                A3(strain, k, j, i) = s;
                A3(vism, k, j, i) = (ph->lamsq[k] * sqrt(s)) * fm;
                A3(vist, k, j, i) = ((ph->lamsq[k] * sqrt(s)) * fh) * ra[k];
            }
    /* no viscosity in the two levels below the lid :163-166 */
    for (int k = kp - 1; k <= kp; ++k)
        for (int i = 0; i <= ip + 1; ++i)
            for (int j = 0; j <= jp + 1; ++j) {
                A3(vism, k, j, i) = 0.0;
                A3(vist, k, j, i) = 0.0;
            }
    monin_obukhov(g, ph, th, u, v, vism, vist); /* :167-172 */
    or_fill_full_wo(g, vism);                    /* :176-179 */
    or_fill_full_wo(g, vist);
    if (ph->moisture) /* :182-187 */
        for (int n = 0; n < ph->nq; ++n) memcpy(viss + (size_t)n * N, vist, N * sizeof(double));
    /* stress tensor :189-223 */
    double *t11 = zalloc(N), *t22 = zalloc(N), *t33 = zalloc(N);
    double *t12 = zalloc(N), *t13 = zalloc(N), *t23 = zalloc(N);
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double r1 = rn[k] * 1.0;
                const double vv = A3(vism, k, j, i) + A3(vism, k - 1, j, i);
                A3(t11, k, j, i) = ((r1 * vv) * (A3(zu, k, j, i) - A3(zu, k, j, i - 1))) / dx[i - 1];
                A3(t22, k, j, i) = ((r1 * vv) * (A3(zv, k, j, i) - A3(zv, k, j - 1, i))) / dy[j - 1];
                A3(t33, k, j, i) = ((r1 * vv) * (A3(zw, k, j, i) - A3(zw, k - 1, j, i))) / dz[k - 1];
                const double s8 =
                    ((((((A3(vism, k, j, i) + A3(vism, k, j + 1, i)) + A3(vism, k - 1, j, i)) +
                        A3(vism, k - 1, j + 1, i)) + A3(vism, k, j, i + 1)) +
                      A3(vism, k, j + 1, i + 1)) + A3(vism, k - 1, j, i + 1)) +
                    A3(vism, k - 1, j + 1, i + 1);
                A3(t12, k, j, i) = ((rn[k] * 0.125) * s8) *
                                   ((A3(zu, k, j + 1, i) - A3(zu, k, j, i)) / dyn[j] +
                                    (A3(zv, k, j, i + 1) - A3(zv, k, j, i)) / dxn[i]);
            }
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                A3(t13, k, j, i) = ((ra[k] * 0.5) * (A3(vism, k, j, i) + A3(vism, k, j, i + 1))) *
                                   ((A3(zu, k + 1, j, i) - A3(zu, k, j, i)) / dzn[k] +
                                    (A3(zw, k, j, i + 1) - A3(zw, k, j, i)) / dxn[i]);
                A3(t23, k, j, i) = ((ra[k] * 0.5) * (A3(vism, k, j, i) + A3(vism, k, j + 1, i))) *
                                   ((A3(zw, k, j + 1, i) - A3(zw, k, j, i)) / dyn[j] +
                                    (A3(zv, k + 1, j, i) - A3(zv, k, j, i)) / dzn[k]);
            }
    double *taus[6] = {t11, t22, t33, t12, t13, t23}; /* :228-239 */
    for (int n = 0; n < 6; ++n) or_fill_along_dim(g, taus[n], 0.0, 0.0);
    /* momentum sources :241-278: sw on 1..kp-1, su and sv on 2..kp-1 */
    const int ktmp = kp - 1, ktmp1 = 2;
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j) {
            for (int k = 1; k <= ktmp; ++k)
                A3(sw, k, j, i) = A3(sw, k, j, i) +
                                  (((A3(t33, k + 1, j, i) - A3(t33, k, j, i)) / dzn[k] +
                                    (A3(t13, k, j, i) - A3(t13, k, j, i - 1)) / dx[i - 1]) +
                                   (A3(t23, k, j, i) - A3(t23, k, j - 1, i)) / dy[j - 1]);
            for (int k = ktmp1; k <= ktmp; ++k) {
                A3(su, k, j, i) = A3(su, k, j, i) +
                                  (((A3(t11, k, j, i + 1) - A3(t11, k, j, i)) / dxn[i] +
                                    (A3(t12, k, j, i) - A3(t12, k, j - 1, i)) / dy[j - 1]) +
                                   (A3(t13, k, j, i) - A3(t13, k - 1, j, i)) / dz[k - 1]);
                A3(sv, k, j, i) = A3(sv, k, j, i) +
                                  (((A3(t22, k, j + 1, i) - A3(t22, k, j, i)) / dyn[j] +
                                    (A3(t12, k, j, i) - A3(t12, k, j, i - 1)) / dx[i - 1]) +
                                   (A3(t23, k, j, i) - A3(t23, k - 1, j, i)) / dz[k - 1]);
            }
        }
    or_fill_full(g, su, 0.0, 0.0); /* :279-284 */
    or_fill_full(g, sv, 0.0, 0.0);
    or_fill_full(g, sw, 0.0, 0.0);
    diffuse_scalar(g, vist, th, sth); /* :288-315 */
    or_fill_full(g, sth, 0.0, 0.0);
    if (ph->moisture) /* :320-354 */
        for (int n = 0; n < ph->nq; ++n) {
            diffuse_scalar(g, vist, q + (size_t)n * N, sq + (size_t)n * N);
            or_fill_full(g, sq + (size_t)n * N, 0.0, 0.0);
        }
    for (int n = 0; n < 6; ++n) free(taus[n]);
}

/* sources, dynamics.f90:639-1000, every option but radiation */
void or_sources_phys(const or_grid *g, const or_physics *ph, const double *u, const double *v,
                     const double *w, const double *zu, const double *zv, const double *zw,
                     const double *th, const double *q, const double *theta,
                     const double *thbar, double *su, double *sv, double *sw, double *sth,
                     double *sq, double *strain, double *vism, double *vist, double *viss,
                     double *rhs) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const double *ra = g->rhoa, *rn = g->rhoan;
    sources_advective(g, u, v, w, th, theta, thbar, su, sv, sw);
    if (ph->damping_layer) /* :802-839 */
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j)
                for (int k = 1; k <= kp; ++k) {
                    A3(su, k, j, i) = A3(su, k, j, i) +
                                      (rn[k] * ph->dampfacn[k]) * (A3(zu, k, j, i) - ph->ubar[k]);
                    A3(sv, k, j, i) = A3(sv, k, j, i) +
                                      (rn[k] * ph->dampfacn[k]) * (A3(zv, k, j, i) - ph->vbar[k]);
                    A3(sw, k, j, i) = A3(sw, k, j, i) +
                                      (ra[k] * ph->dampfac[k]) * (A3(zw, k, j, i) - 0.0);
                    A3(sth, k, j, i) = A3(sth, k, j, i) + ph->dampfacn[k] * (A3(th, k, j, i) - 0.0);
                }
    if (ph->forcing) /* :841-860 */
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j)
                for (int k = 1; k <= kp; ++k) {
                    A3(su, k, j, i) = A3(su, k, j, i) -
                                      (rn[k] * (A3(zu, k, j, i) - ph->u_force[k])) / ph->forcing_tau;
                    A3(sv, k, j, i) = A3(sv, k, j, i) -
                                      (rn[k] * (A3(zv, k, j, i) - ph->v_force[k])) / ph->forcing_tau;
                }
    if (ph->divergence) /* :863-873 */
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j)
                for (int k = 1; k <= kp; ++k)
                    A3(sw, k, j, i) = A3(sw, k, j, i) - rn[k] * (ph->wbar[k] - ph->w_subs[k]);
    if (ph->viscous) { /* :878-906 */
        or_subgrid(g, ph, u, v, w, zu, zv, zw, th, q, vism, vist, viss, strain, su, sv, sw, sth,
                   sq);
    } else { /* :907-917 */
        or_fill_full(g, su, 0.0, 0.0);
        or_fill_full(g, sv, 0.0, 0.0);
        or_fill_full(g, sw, 0.0, 0.0);
        or_fill_full(g, sth, 0.0, 0.0);
    }
    sources_finish(g, su, sv, sw, rhs);
}

/* ------------------------------------------------------------------------ */
/* pressure solver                                                          */
/* ------------------------------------------------------------------------ */

static inline int is_interior(const or_grid *g, int k, int j, int i) {
    return k >= 1 && k <= g->kp && j >= 1 && j <= g->jp && i >= 1 && i <= g->ip;
}

/* The stencil's coefficients as set_mat_a stores them (dynamics.f90:84-96):
 * the reference keeps a_e..a_p as arrays computed once per solve, so its
 * stencil costs loads, not divisions.  The same values here come from
 * per-axis tables (a_e, a_w depend on i only, a_n, a_s on j, a_u, a_d on k)
 * and a_p is summed from them in the order of :91-93, which gives the same
 * bits as the inline expression of coef_ap. */
typedef struct {
    double *ae, *aw, *an, *as, *au, *ad; /* index 1..n */
} coef_tables;

static coef_tables coef_tables_make(const or_grid *g) {
    coef_tables t;
    t.ae = zalloc((size_t)g->ip + 2);
    t.aw = zalloc((size_t)g->ip + 2);
    t.an = zalloc((size_t)g->jp + 2);
    t.as = zalloc((size_t)g->jp + 2);
    t.au = zalloc((size_t)g->kp + 2);
    t.ad = zalloc((size_t)g->kp + 2);
    for (int i = 1; i <= g->ip; ++i) {
        t.ae[i] = 1.0 / (g->dx[i - 1] * g->dxn[i]);
        t.aw[i] = 1.0 / (g->dx[i - 1] * g->dxn[i - 1]);
    }
    for (int j = 1; j <= g->jp; ++j) {
        t.an[j] = 1.0 / (g->dy[j - 1] * g->dyn[j]);
        t.as[j] = 1.0 / (g->dy[j - 1] * g->dyn[j - 1]);
    }
    for (int k = 1; k <= g->kp; ++k) {
        t.au[k] = 1.0 / (g->dz[k - 1] * g->dzn[k]);
        t.ad[k] = 1.0 / (g->dz[k - 1] * g->dzn[k - 1]);
    }
    return t;
}

static void coef_tables_free(coef_tables *t) {
    free(t->ae);
    free(t->aw);
    free(t->an);
    free(t->as);
    free(t->au);
    free(t->ad);
}

static inline double tab_ap(const coef_tables *t, int k, int j, int i) {
    return ((((-t->ae[i] - t->aw[i]) - t->an[j]) - t->as[j]) - t->au[k]) - t->ad[k];
}

static void matvec_tab(const or_grid *g, const coef_tables *t, double *x, double *ax) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    or_fill_along_dim(g, x, 0.0, 0.0);
    OMP_FOR
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) A3(x, 0, j, i) = A3(x, 1, j, i);
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) A3(x, kp + 1, j, i) = A3(x, kp, j, i);
    const size_t slab = (size_t)(jp + 2) * (kp + 2);
    OMP_FOR
    for (int i = 0; i <= ip + 1; ++i) memset(ax + slab * i, 0, slab * sizeof(double));
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const double ae = t->ae[i], aw = t->aw[i];
                const double an = t->an[j], as = t->as[j];
                const double au = t->au[k], ad = t->ad[k];
                const double ap = tab_ap(t, k, j, i);
                A3(ax, k, j, i) = ((((((aw * A3(x, k, j, i - 1) + ae * A3(x, k, j, i + 1)) +
                                       as * A3(x, k, j - 1, i)) + an * A3(x, k, j + 1, i)) +
                                     ad * A3(x, k - 1, j, i)) + au * A3(x, k + 1, j, i)) +
                                   ap * A3(x, k, j, i)) + 0.0;
            }
}

/* mat_ax, dynamics.f90:187-229.  Mutates the halo of x. */
void or_matvec(const or_grid *g, double *x, double *ax) {
    coef_tables t = coef_tables_make(g);
    matvec_tab(g, &t, x, ax);
    coef_tables_free(&t);
}

/* precondition, dynamics.f90:241-270 (whole-array statements; a_p is 1 in
 * the halo, :57) */
static void precondition_tab(const or_grid *g, const coef_tables *ct, double *s, const double *r,
                             int prit, double *t) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    OMP_FOR
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j)
            for (int k = 0; k <= kp + 1; ++k)
                A3(s, k, j, i) = A3(r, k, j, i) / (is_interior(g, k, j, i) ? tab_ap(ct, k, j, i) : 1.0);
    for (int it = 1; it <= prit; ++it) {
        matvec_tab(g, ct, s, t);
        OMP_FOR
        for (int i = 0; i <= ip + 1; ++i)
            for (int j = 0; j <= jp + 1; ++j)
                for (int k = 0; k <= kp + 1; ++k)
                    A3(s, k, j, i) = A3(s, k, j, i) + (A3(r, k, j, i) - A3(t, k, j, i)) /
                                                          (is_interior(g, k, j, i) ? tab_ap(ct, k, j, i) : 1.0);
    }
}

void or_precondition(const or_grid *g, double *s, const double *r, int prit) {
    coef_tables t = coef_tables_make(g);
    double *w = zalloc(NTOT(g));
    precondition_tab(g, &t, s, r, prit, w);
    free(w);
    coef_tables_free(&t);
}

/* Summation order of or_inn_prod.  0: the reference's serial loop.  1: 64
 * interleaved partial sums combined pairwise, a different but equally valid
 * order, used only to measure how far results move when only the order of
 * the dot products changes.  2: a parallel reduction over the host threads
 * (an order that depends on the thread count; for timing only). */
static int g_sum_blocked = 0;
void or_set_sum_blocked(int on) { g_sum_blocked = on; }

/* inn_prod, dynamics.f90:458-482: sequential, i-j-k loop order */
double or_inn_prod(const or_grid *g, const double *x, const double *y) {
    if (g_sum_blocked == 2) {  /* a parallel reduction (timing only) */
        double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s) if (g_threads > 1)
        for (int i = 1; i <= g->ip; ++i)
            for (int j = 1; j <= g->jp; ++j)
                for (int k = 1; k <= g->kp; ++k) s += A3(x, k, j, i) * A3(y, k, j, i);
        return s;
    }
    if (g_sum_blocked) {
        double part[64] = {0};
        size_t c = 0;
        for (int i = 1; i <= g->ip; ++i)
            for (int j = 1; j <= g->jp; ++j)
                for (int k = 1; k <= g->kp; ++k) {
                    part[c & 63] = part[c & 63] + A3(x, k, j, i) * A3(y, k, j, i);
                    ++c;
                }
        for (int w = 32; w >= 1; w >>= 1)
            for (int q = 0; q < w; ++q) part[q] = part[q] + part[q + w];
        return part[0];
    }
    double s = 0.0;
    for (int i = 1; i <= g->ip; ++i)
        for (int j = 1; j <= g->jp; ++j)
            for (int k = 1; k <= g->kp; ++k) s = s + A3(x, k, j, i) * A3(y, k, j, i);
    return s;
}

/* bicgstab, dynamics.f90:293-444 */
int or_bicgstab(const or_grid *g, const double *xg, const double *yg,
                const double *zg, const double *sw, double *x, double *b,
                double tol, int test_solver, int first_call, int itmax,
                int *iters, double *errp) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    if (first_call && test_solver) { /* :354-364 */
        OMP_FOR
        for (int i = 1; i <= ip; ++i)
            for (int j = 1; j <= jp; ++j)
                for (int k = 1; k <= kp; ++k) {
                    const double zz = zg[k] - 4000.0;
                    A3(b, k, j, i) = exp(-(1.0 / 5.e6) * ((xg[i] * xg[i] + yg[j] * yg[j]) + zz * zz));
                }
    }
    double *ax = zalloc(N), *r = zalloc(N), *cr = zalloc(N), *cs = zalloc(N);
    double *p = zalloc(N), *v1 = zalloc(N), *t = zalloc(N), *s = zalloc(N);
    double *work = zalloc(N); /* precondition's local t */
    coef_tables ct = coef_tables_make(g);
    int status = OR_OK, it_done = 0;
    double sc_err = sqrt(or_inn_prod(g, b, b));
    if (sc_err < 0.01) sc_err = 0.01; /* :371 */
    matvec_tab(g, &ct, x, ax);
    OMP_FOR
    for (size_t q = 0; q < N; ++q) {
        r[q] = b[q] - ax[q];
        cr[q] = r[q];
        p[q] = 0.0;
        v1[q] = 0.0;
    }
    double err = sqrt(or_inn_prod(g, r, r)) / sc_err;
    if (err < tol) { /* :383 early return: no final boundary rows */
        *iters = 0;
        *errp = err;
        free(ax); free(r); free(cr); free(cs); free(p); free(v1); free(t); free(s);
        free(work);
        coef_tables_free(&ct);
        return OR_OK;
    }
    double alf = 1.0, omg = 1.0, nrm = 1.0, rho, bet, tt, ts;
    int it;
    for (it = 1; it <= itmax; ++it) {
        rho = or_inn_prod(g, r, cr);
        bet = (rho / nrm) * (alf / omg);
        OMP_FOR
        for (size_t q = 0; q < N; ++q) t[q] = r[q] - bet * (omg * v1[q]);
        precondition_tab(g, &ct, s, t, 3, work);
        OMP_FOR
        for (size_t q = 0; q < N; ++q) p[q] = s[q] + bet * p[q];
        matvec_tab(g, &ct, p, v1);
        nrm = or_inn_prod(g, cr, v1);
        alf = rho / nrm;
        OMP_FOR
        for (size_t q = 0; q < N; ++q) s[q] = r[q] - alf * v1[q];
        precondition_tab(g, &ct, cs, s, 3, work);
        matvec_tab(g, &ct, cs, t);
        tt = or_inn_prod(g, t, t);
        ts = or_inn_prod(g, t, s);
        omg = ts / tt;
        if (fabs(omg) < TINY_OMEGA) { /* :414-417 */
            status = OR_BREAKDOWN;
            it_done = it;
            break;
        }
        OMP_FOR
        for (size_t q = 0; q < N; ++q) x[q] = (x[q] + alf * p[q]) + omg * cs[q];
        OMP_FOR
        for (size_t q = 0; q < N; ++q) r[q] = s[q] - omg * t[q];
        err = sqrt(or_inn_prod(g, r, r)) / sc_err;
        nrm = rho;
        it_done = it;
        if (err < tol) break;
    }
    if (status == OR_OK && err > tol) status = OR_NOT_CONVERGED; /* :431-434 */
    if (status == OR_OK) { /* :437-442 */
        OMP_FOR
        for (int i = 0; i <= ip + 1; ++i)
            for (int j = 0; j <= jp + 1; ++j) A3(x, kp + 1, j, i) = A3(x, kp, j, i);
        for (int i = 0; i <= ip + 1; ++i)
            for (int j = 0; j <= jp + 1; ++j)
                A3(x, 0, j, i) = A3(x, 1, j, i) - A3(sw, 0, j, i) * g->dzn[0];
    }
    *iters = it_done;
    *errp = err;
    free(ax); free(r); free(cr); free(cs); free(p); free(v1); free(t); free(s);
    free(work);
    coef_tables_free(&ct);
    return status;
}

/* divergence_calc, diagnostics.f90:140-148 */
void or_divergence(const or_grid *g, const double *u, const double *v,
                   const double *w, double *div) {
    for (int i = 1; i <= g->ip; ++i)
        for (int j = 1; j <= g->jp; ++j)
            for (int k = 1; k <= g->kp; ++k)
                A3(div, k, j, i) =
                    g->rhoan[k] * (A3(u, k, j, i) - A3(u, k, j, i - 1)) / g->dx[i - 1] +
                    g->rhoan[k] * (A3(v, k, j, i) - A3(v, k, j - 1, i)) / g->dy[j - 1] +
                    (g->rhoa[k] * A3(w, k, j, i) - g->rhoa[k - 1] * A3(w, k - 1, j, i)) /
                        g->dz[k - 1];
}

/* advance_momentum, dynamics.f90:1019-1099 (dt is the leapfrog 2*dt) */
void or_advance_momentum(const or_grid *g, double dt, double *u, double *v,
                         double *w, const double *zu, const double *zv,
                         const double *zw, const double *su, const double *sv,
                         const double *sw, const double *p) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 0; k <= kp; ++k)
                A3(u, k, j, i) = A3(zu, k, j, i) +
                                 ((A3(su, k, j, i) - (A3(p, k, j, i + 1) - A3(p, k, j, i)) / g->dxn[i]) * dt) /
                                     g->rhoan[k];
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 0; k <= kp; ++k)
                A3(v, k, j, i) = A3(zv, k, j, i) +
                                 ((A3(sv, k, j, i) - (A3(p, k, j + 1, i) - A3(p, k, j, i)) / g->dyn[j]) * dt) /
                                     g->rhoan[k];
    OMP_FOR
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 0; k <= kp; ++k)
                A3(w, k, j, i) = A3(zw, k, j, i) +
                                 ((A3(sw, k, j, i) - (A3(p, k + 1, j, i) - A3(p, k, j, i)) / g->dzn[k]) * dt) /
                                     g->rhoa[k];
}

/* advance_scalar_fields_3d, sgm/subgrid_3d.f90:582-623 (whole arrays) */
void or_advance_scalars(const or_grid *g, double dt, double *q,
                        const double *sq, double *th, const double *sth,
                        int nq, int moisture) {
    const size_t N = NTOT(g);
    for (size_t x = 0; x < N; ++x) th[x] = th[x] + sth[x] * dt;
    if (moisture)
        for (size_t x = 0; x < N * (size_t)nq; ++x) q[x] = q[x] + sq[x] * dt;
}

/* ------------------------------------------------------------------------ */
/* the time step of model_driver                                            */
/* ------------------------------------------------------------------------ */

void or_model_levels(or_model *m, double **u, double **v, double **w,
                     double **zu, double **zv, double **zw, double **tu,
                     double **tv, double **tw) {
    /* pointer rotation, driver_code.f90:634-646 */
    switch (m->n % 3) {
    case 0:
        *u = m->ut; *v = m->vt; *w = m->wt;
        *zu = m->zut; *zv = m->zvt; *zw = m->zwt;
        *tu = m->tut; *tv = m->tvt; *tw = m->twt;
        break;
    case 1:
        *u = m->tut; *v = m->tvt; *w = m->twt;
        *zu = m->ut; *zv = m->vt; *zw = m->wt;
        *tu = m->zut; *tv = m->zvt; *tw = m->zwt;
        break;
    default:
        *u = m->zut; *v = m->zvt; *w = m->zwt;
        *zu = m->tut; *zv = m->tvt; *zw = m->twt;
        *tu = m->ut; *tv = m->vt; *tw = m->wt;
        break;
    }
}

/* driver_code.f90:287-295 */
void or_model_init_bcs(const or_grid *g, or_model *m) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    double *u, *v, *w, *zu, *zv, *zw, *tu, *tv, *tw;
    or_model_levels(m, &u, &v, &w, &zu, &zv, &zw, &tu, &tv, &tw);
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) {
            A3(zw, 0, j, i) = -A3(zw, 1, j, i);
            A3(zu, 1, j, i) = 0.0;
            A3(zv, 1, j, i) = 0.0;
        }
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) {
            A3(zw, kp, j, i) = -A3(zw, kp - 1, j, i);
            A3(zw, kp + 1, j, i) = 0.0;
            A3(zu, kp, j, i) = 0.0;
            A3(zv, kp, j, i) = 0.0;
        }
}

/* one pass of the time loop, driver_code.f90:304-676, hot path only:
 * output, microphysics, radiation, land surface, divergence_calc (a
 * diagnostic) and the viscous / damping / forcing branches are not taken. */
int or_model_step(const or_grid *g, or_model *m) {
    const int ip = g->ip, jp = g->jp, kp = g->kp;
    const size_t N = NTOT(g);
    double *u, *v, *w, *zu, *zv, *zw, *tu, *tv, *tw;
    or_model_levels(m, &u, &v, &w, &zu, &zv, &zw, &tu, &tv, &tw);

    memset(m->sth, 0, N * sizeof(double)); /* :340 */
    if (m->moisture) memset(m->sq, 0, N * (size_t)m->nq * sizeof(double));
    or_hmeans(g, u, v, w, m->th, m->moisture ? m->q : NULL,
              m->moisture ? m->nq : 0, m->ubar, m->vbar, m->wbar, m->thbar,
              m->qbar); /* :433-435 */
    if (m->phys) { /* :455-479 with the physics options */
        or_physics ph = *m->phys;
        ph.ubar = m->ubar;
        ph.vbar = m->vbar;
        ph.wbar = m->wbar;
        or_sources_phys(g, &ph, u, v, w, zu, zv, zw, m->th, m->q, m->theta, m->thbar, m->su,
                        m->sv, m->sw, m->sth, m->sq, m->strain, m->vism, m->vist, m->viss,
                        m->psrc);
    } else {
        or_sources(g, u, v, w, m->th, m->theta, m->thbar, m->su, m->sv, m->sw,
                   m->sth, m->psrc); /* :455-479 */
    }
    or_fill_along_dim_wo(g, m->psrc); /* :485-487 */
    int iters = 0;
    double err = 0.0;
    int st = or_bicgstab(g, m->xg, m->yg, m->zg, m->sw, m->p, m->psrc, m->ptol,
                         0, m->solver_first, 9999, &iters, &err); /* :496-500 */
    m->solver_first = 0;
    m->last_iters = iters;
    m->last_err = err;
    if (st != OR_OK) return st;
    or_fill_along_dim_wo(g, m->p); /* :506-507 */
    if (m->advection_scheme == 0 || m->advection_scheme == 1) {
        or_adv_ref_state(g, m->dt, w, m->th, m->thetan); /* :518-520 */
        or_fill_full(g, m->th, 0.0, 0.0);                /* :526-527 */
    }
    if (m->advection_scheme == 0) { /* :539-546 */
        or_upwind3d(g, m->dt, u, v, w, m->th, 0);
        or_fill_full(g, m->th, m->t1, m->t2);
    } else if (m->advection_scheme == 1) {
        or_mpdata3d(g, m->dt, u, v, w, m->th, &m->t1, &m->t2, m->kord,
                    m->monotone, 0); /* :548-551 */
    } else { /* :552-556: theta + thetan advected, thbase/thtop its bcs */
        or_mpdata3d_add(g, m->dt, u, v, w, m->th, m->thetan, &m->thbase, &m->thtop,
                        m->kord, m->monotone, 0);
    }
    if (m->moisture) { /* :571-580, q1 = q2 = 0 */
        for (int c = 0; c < m->ncat; ++c) {
            const int nqc = m->c_e[c] - m->c_s[c] + 1;
            double *q1 = zalloc((size_t)nqc), *q2 = zalloc((size_t)nqc);
            or_mpdata_vec3d(g, m->dt, u, v, w, m->q + (size_t)(m->c_s[c] - 1) * N,
                            nqc, q1, q2, m->kord, m->monotone, 0);
            free(q1); free(q2);
        }
    }
    or_advance_momentum(g, 2.0 * m->dt, tu, tv, tw, zu, zv, zw, m->su, m->sv,
                        m->sw, m->p); /* :592-594 */
    or_advance_scalars(g, m->dt, m->q, m->sq, m->th, m->sth, m->nq,
                       m->moisture); /* :601-603 */
    /* time filter, :618-626 */
    double *lev[3][2] = {{u, zu}, {v, zv}, {w, zw}};
    for (int c = 0; c < 3; ++c) {
        double *a = lev[c][0], *z = lev[c][1];
        OMP_FOR
        for (size_t x = 0; x < N; ++x) {
            const double tmp = 0.99 * a[x] + 0.01 * z[x];
            z[x] = 0.01 * a[x] + 0.99 * z[x];
            a[x] = tmp;
        }
    }
    m->n += 1; /* rotation, :634-646 */
    or_model_levels(m, &u, &v, &w, &zu, &zv, &zw, &tu, &tv, &tw);
    or_fill_full(g, u, 0.0, 0.0); /* :656-661 */
    or_fill_full(g, v, 0.0, 0.0);
    or_fill_full(g, w, 0.0, 0.0);
    OMP_FOR
    for (int i = 0; i <= ip + 1; ++i) /* :663-674 */
        for (int j = 0; j <= jp + 1; ++j) {
            A3(w, 0, j, i) = 0.0;
            A3(u, 0, j, i) = -A3(u, 1, j, i);
            A3(v, 0, j, i) = -A3(v, 1, j, i);
        }
    OMP_FOR
    for (int i = 0; i <= ip + 1; ++i)
        for (int j = 0; j <= jp + 1; ++j) {
            A3(w, kp + 1, j, i) = 0.0;
            A3(u, kp + 1, j, i) = A3(u, kp, j, i);
            A3(v, kp + 1, j, i) = A3(v, kp, j, i);
        }
    return OR_OK;
}
