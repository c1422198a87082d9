/*
 * abi_mock.c -- the C-ABI of include/dcm_hip.h implemented on the CPU by the
 * oracle, for the entry points the Fortran drop-in modules call.
 *
 * TEST INFRASTRUCTURE ONLY.  Preloaded (LD_PRELOAD) under the Fortran
 * drop-in check program on a host without a GPU, it lets the tests check the
 * Fortran layer itself -- argument order, array bounds, logical conversion,
 * the per-argument device buffers and their uploads and downloads -- against
 * the oracle, with "device" buffers in host memory.  The product never loads
 * it; on a GPU host the same program runs on libdcm_hip.so.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "dcm_oracle.h"

typedef struct {
    or_grid g;
    double *met;  /* 8 metric arrays, copied */
    int first_solve;
} mock_ctx;

static const char *g_err = "";

const char *dcm_last_error(void) { return g_err; }

int dcm_context_create(int ip, int jp, int kp, const double *dx, const double *dy,
                       const double *dz, const double *dxn, const double *dyn,
                       const double *dzn, const double *rhoa, const double *rhoan,
                       void **out) {
    mock_ctx *c = calloc(1, sizeof(mock_ctx));
    const size_t ni = (size_t)ip + 2, nj = (size_t)jp + 2, nk = (size_t)kp + 2;
    c->met = malloc(sizeof(double) * (2 * ni + 2 * nj + 4 * nk));
    double *p = c->met;
    const double *src[8] = {dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan};
    const size_t len[8] = {ni, nj, nk, ni, nj, nk, nk, nk};
    const double *dst[8];
    for (int a = 0; a < 8; ++a) {
        memcpy(p, src[a], len[a] * sizeof(double));
        dst[a] = p;
        p += len[a];
    }
    c->g.ip = ip;
    c->g.jp = jp;
    c->g.kp = kp;
    c->g.dx = dst[0];
    c->g.dy = dst[1];
    c->g.dz = dst[2];
    c->g.dxn = dst[3];
    c->g.dyn = dst[4];
    c->g.dzn = dst[5];
    c->g.rhoa = dst[6];
    c->g.rhoan = dst[7];
    c->first_solve = 1;
    *out = c;
    return 0;
}

int dcm_context_destroy(void *ctx) {
    mock_ctx *c = ctx;
    free(c->met);
    free(c);
    return 0;
}

int dcm_synchronize(void *ctx) { (void)ctx; return 0; }
int dcm_set_sum_order(void *ctx, int order) { (void)ctx; (void)order; return 0; }

int dcm_malloc(void *ctx, size_t bytes, void **d) {
    (void)ctx;
    /* poison new buffers, as fresh device memory holds arbitrary bits */
    *d = malloc(bytes ? bytes : 8);
    memset(*d, 0xff, bytes);
    return 0;
}
int dcm_free(void *ctx, void *d) { (void)ctx; free(d); return 0; }
int dcm_memcpy_htod(void *ctx, void *dst, const void *src, size_t n) {
    (void)ctx;
    memcpy(dst, src, n);
    return 0;
}
int dcm_memcpy_dtoh(void *ctx, void *dst, const void *src, size_t n) {
    (void)ctx;
    memcpy(dst, src, n);
    return 0;
}

int dcm_hmeans(void *ctx, const double *u, const double *v, const double *w, const double *th,
               const double *q, int nq, double *ubar, double *vbar, double *wbar,
               double *thbar, double *qbar) {
    mock_ctx *c = ctx;
    or_hmeans(&c->g, u, v, w, th, q, nq, ubar, vbar, wbar, thbar, qbar);
    return 0;
}

int dcm_divergence(void *ctx, const double *u, const double *v, const double *w, double *div) {
    mock_ctx *c = ctx;
    or_divergence(&c->g, u, v, w, div);
    return 0;
}

int dcm_sources(void *ctx, const double *u, const double *v, const double *w, const double *th,
                const double *theta, const double *thbar, double *su, double *sv, double *sw,
                double *sth, double *rhs) {
    mock_ctx *c = ctx;
    or_sources(&c->g, u, v, w, th, theta, thbar, su, sv, sw, sth, rhs);
    return 0;
}

int dcm_bicgstab(void *ctx, const double *xg, const double *yg, const double *zg,
                 const double *sw, double *x, double *b, double tol, int test_solver,
                 int itmax, int *iters, double *err) {
    mock_ctx *c = ctx;
    int it = 0;
    double e = 0.0;
    const int st = or_bicgstab(&c->g, xg, yg, zg, sw, x, b, tol, test_solver, c->first_solve,
                               itmax > 0 ? itmax : 9999, &it, &e);
    c->first_solve = 0;
    if (iters) *iters = it;
    if (err) *err = e;
    if (st == 0 && !isfinite(e)) return 3;
    return st;
}

int dcm_upwind3d(void *ctx, double dt, const double *u, const double *v, const double *w,
                 double *psi, int neumann) {
    mock_ctx *c = ctx;
    or_upwind3d(&c->g, dt, u, v, w, psi, neumann);
    return 0;
}

int dcm_mpdata3d(void *ctx, double dt, const double *u, const double *v, const double *w,
                 double *psi, double *lbc, double *ubc, int kord, int monotone, int neumann,
                 int *early_exit) {
    mock_ctx *c = ctx;
    const int e = or_mpdata3d(&c->g, dt, u, v, w, psi, lbc, ubc, kord, monotone, neumann);
    if (early_exit) *early_exit = e;
    return 0;
}

int dcm_mpdata_vec3d(void *ctx, double dt, const double *u, const double *v, const double *w,
                     double *psi, int nq, double *lbc, double *ubc, int kord, int monotone,
                     int neumann, int *early_exit) {
    mock_ctx *c = ctx;
    const int e = or_mpdata_vec3d(&c->g, dt, u, v, w, psi, nq, lbc, ubc, kord, monotone,
                                  neumann);
    if (early_exit) *early_exit = e;
    return 0;
}

int dcm_mpdata3d_add(void *ctx, double dt, const double *u, const double *v, const double *w,
                     double *psi, const double *psi_1d, double *lbc, double *ubc, int kord,
                     int monotone, int neumann, int *early_exit) {
    mock_ctx *c = ctx;
    const int e = or_mpdata3d_add(&c->g, dt, u, v, w, psi, psi_1d, lbc, ubc, kord, monotone,
                                  neumann);
    if (early_exit) *early_exit = e;
    return 0;
}

int dcm_upwind_vert3d(void *ctx, double dt, const double *w, double *psi, int neumann) {
    mock_ctx *c = ctx;
    or_upwind_vert3d(&c->g, dt, w, psi, neumann);
    return 0;
}

int dcm_mpdata_vert3d(void *ctx, double dt, const double *w, double *psi, double *lbc,
                      double *ubc, int kord, int monotone, int neumann, int *early_exit) {
    mock_ctx *c = ctx;
    const int e = or_mpdata_vert3d(&c->g, dt, w, psi, lbc, ubc, kord, monotone, neumann);
    if (early_exit) *early_exit = e;
    return 0;
}

int dcm_mpdata_vec_vert3d(void *ctx, double dt, const double *w, double *psi, int nq,
                          double *lbc, double *ubc, int kord, int monotone, int neumann,
                          int *early_exit) {
    mock_ctx *c = ctx;
    const int e = or_mpdata_vec_vert3d(&c->g, dt, w, psi, nq, lbc, ubc, kord, monotone, neumann);
    if (early_exit) *early_exit = e;
    return 0;
}

int dcm_adv_ref_state(void *ctx, double dt, const double *w, double *psi, const double *ref) {
    mock_ctx *c = ctx;
    or_adv_ref_state(&c->g, dt, w, psi, ref);
    return 0;
}

int dcm_advance_momentum(void *ctx, double dt, double *u, double *v, double *w,
                         const double *zu, const double *zv, const double *zw, const double *su,
                         const double *sv, const double *sw, const double *p) {
    mock_ctx *c = ctx;
    or_advance_momentum(&c->g, dt, u, v, w, zu, zv, zw, su, sv, sw, p);
    return 0;
}

/* dcm_physics of include/dcm_hip.h, field for field */
typedef struct {
    int viscous, damping_layer, forcing, divergence;
    double forcing_tau, z0, z0th;
    const double *lamsq, *dampfacn, *dampfac, *u_force, *v_force, *w_subs;
    const double *z, *theta, *thetan;
} mock_physics;

static or_physics to_or_physics(const mock_physics *ph, const double *ubar, const double *vbar,
                                const double *wbar, int nq) {
    or_physics o;
    memset(&o, 0, sizeof(o));
    o.viscous = ph->viscous;
    o.moisture = nq > 0;
    o.damping_layer = ph->damping_layer;
    o.forcing = ph->forcing;
    o.divergence = ph->divergence;
    o.nq = nq;
    o.forcing_tau = ph->forcing_tau;
    o.z0 = ph->z0;
    o.z0th = ph->z0th;
    o.log_z1_z0 = log(ph->z[1] / ph->z0);
    o.ubar = ubar;
    o.vbar = vbar;
    o.wbar = wbar;
    o.dampfacn = ph->dampfacn;
    o.dampfac = ph->dampfac;
    o.u_force = ph->u_force;
    o.v_force = ph->v_force;
    o.w_subs = ph->w_subs;
    o.lamsq = ph->lamsq;
    o.z = ph->z;
    o.theta = ph->theta;
    o.thetan = ph->thetan;
    return o;
}

int dcm_subgrid(void *ctx, const mock_physics *ph, const double *u, const double *v,
                const double *w, const double *zu, const double *zv, const double *zw,
                const double *th, const double *q, int nq, double *vism, double *vist,
                double *viss, double *strain, double *su, double *sv, double *sw, double *sth,
                double *sq) {
    mock_ctx *c = ctx;
    const or_physics o = to_or_physics(ph, NULL, NULL, NULL, nq);
    or_subgrid(&c->g, &o, u, v, w, zu, zv, zw, th, q, vism, vist, viss, strain, su, sv, sw,
               sth, sq);
    return 0;
}

int dcm_sources_phys(void *ctx, const mock_physics *ph, const double *ubar, const double *vbar,
                     const double *wbar, const double *u, const double *v, const double *w,
                     const double *zu, const double *zv, const double *zw, const double *th,
                     const double *q, int nq, const double *theta, const double *thbar,
                     double *su, double *sv, double *sw, double *sth, double *sq,
                     double *strain, double *vism, double *vist, double *viss, double *rhs) {
    mock_ctx *c = ctx;
    const or_physics o = to_or_physics(ph, ubar, vbar, wbar, nq);
    or_sources_phys(&c->g, &o, u, v, w, zu, zv, zw, th, q, theta, thbar, su, sv, sw, sth, sq,
                    strain, vism, vist, viss, rhs);
    return 0;
}
