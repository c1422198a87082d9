/*
 * dcm_oracle.h -- CPU restatement of the DCM dynamical-core hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  The product path never links or calls it.
 *
 * Every function restates one routine of the reference Fortran (FP64 build,
 * VAR_TYPE=1, one MPI rank) in plain C, in the reference's own loop and
 * operand order, so that with -ffp-contract=off the results are the ones the
 * reference computes.  Citations are to files of the reference tree.
 *
 * Array conventions (identical to the reference's explicit-shape dummies):
 *   3-D fields are (0:kp+1, 0:jp+1, 0:ip+1), column major, k fastest:
 *       a[((i)*(jp+2) + (j))*(kp+2) + (k)]
 *   1-D metric / base-state arrays have length n+2 and index 0..n+1, i.e. the
 *   C index equals the Fortran index (halo width 1 on both sides).
 */
#ifndef DCM_ORACLE_H
#define DCM_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int ip, jp, kp;
    const double *dx, *dy, *dz;     /* cell widths            (n+2) */
    const double *dxn, *dyn, *dzn;  /* staggered widths       (n+2) */
    const double *rhoa, *rhoan;     /* density on w / scalar levels (kp+2) */
} or_grid;

/* boundary fills: mpi_module.f90, single-rank semantics */
void or_fill_along_dim(const or_grid *g, double *a, double lbc, double ubc);
void or_fill_along_dim_wo(const or_grid *g, double *a);
void or_fill_full(const or_grid *g, double *a, double lbc, double ubc);
void or_fill_full_wo(const or_grid *g, double *a);
void or_fill_along_z(const or_grid *g, double *a, double lbc, double ubc);

/* advection: pamm/sfvt/advection_3d.F90 (the copy linked into main.exe) */
void or_upwind3d(const or_grid *g, double dt, const double *u, const double *v,
                 const double *w, double *psi, int neumann);
/* returns 1 when the zero-sum early exit was taken, else 0 */
int or_mpdata3d(const or_grid *g, double dt, const double *u, const double *v,
                const double *w, double *psi, double *lbc, double *ubc,
                int kord, int monotone, int neumann);
int or_mpdata_vec3d(const or_grid *g, double dt, const double *u,
                    const double *v, const double *w, double *psi, int nq,
                    double *lbc, double *ubc, int kord, int monotone,
                    int neumann);
void or_adv_ref_state(const or_grid *g, double dt, const double *w,
                      double *psi, const double *psi_ref);
/* vertical-only advection (PAMM sedimentation) and mpdata_3d_add */
void or_upwind_vert3d(const or_grid *g, double dt, const double *w, double *psi,
                      int neumann);
int or_mpdata_vert3d(const or_grid *g, double dt, const double *w, double *psi,
                     double *lbc, double *ubc, int kord, int monotone, int neumann);
int or_mpdata_vec_vert3d(const or_grid *g, double dt, const double *w, double *psi, int nq,
                         double *lbc, double *ubc, int kord, int monotone, int neumann);
int or_mpdata3d_add(const or_grid *g, double dt, const double *u, const double *v,
                    const double *w, double *psi, const double *psi_1d, double *lbc,
                    double *ubc, int kord, int monotone, int neumann);

/* diagnostics.f90 horizontal_means (computed over every level 0..kp+1) */
void or_hmeans(const or_grid *g, const double *u, const double *v,
               const double *w, const double *th, const double *q, int nq,
               double *ubar, double *vbar, double *wbar, double *thbar,
               double *qbar);

/* dynamics.f90 sources (inviscid branch, no damping/forcing/subsidence) */
void or_sources(const or_grid *g, const double *u, const double *v,
                const double *w, const double *th, const double *theta,
                const double *thbar, double *su, double *sv, double *sw,
                double *sth, double *rhs);
/* Physics options of sources (dynamics.f90:639-1000) beyond the inviscid
 * branch.  1-D arrays have kp+2 entries (index 0..kp+1). */
typedef struct {
    int viscous, moisture, damping_layer, forcing, divergence;
    int nq;                               /* advected q fields (moisture) */
    double forcing_tau, z0, z0th;
    /* log(z(1)/z0): one scalar per grid, supplied by the caller so that
     * every implementation uses the same rounded value */
    double log_z1_z0;
    const double *ubar, *vbar, *wbar;     /* horizontal means */
    const double *dampfacn, *dampfac;     /* damping rates, scalar / w levels */
    const double *u_force, *v_force, *w_subs;
    const double *lamsq;                  /* squared mixing length, w levels */
    const double *z;                      /* heights of the w levels */
    const double *theta, *thetan;         /* base state, w / scalar levels */
} or_physics;
/* calculate_subgrid_3d, sgm/subgrid_3d.f90:38-357 (with richardson
 * :448-493 and monin_obukhov :373-433): adds the viscous terms to su, sv, sw,
 * sth (and sq when moisture); fills vism, vist, strain (and viss) */
void or_subgrid(const or_grid *g, const or_physics *ph, const double *u, const double *v,
                const double *w, const double *zu, const double *zv, const double *zw,
                const double *th, const double *q, double *vism, double *vist, double *viss,
                double *strain, double *su, double *sv, double *sw, double *sth, double *sq);
/* sources with every option of the reference except radiation */
void or_sources_phys(const or_grid *g, const or_physics *ph, const double *u, const double *v,
                     const double *w, const double *zu, const double *zv, const double *zw,
                     const double *th, const double *q, const double *theta,
                     const double *thbar, double *su, double *sv, double *sw, double *sth,
                     double *sq, double *strain, double *vism, double *vist, double *viss,
                     double *rhs);

/* dynamics.f90 pressure solver pieces */
void or_matvec(const or_grid *g, double *x, double *ax);  /* mat_ax */
void or_precondition(const or_grid *g, double *s, const double *r, int prit);
double or_inn_prod(const or_grid *g, const double *x, const double *y);

/* status codes shared with the HIP library */
#define OR_OK 0
#define OR_NOT_CONVERGED 1
#define OR_BREAKDOWN 2

/* bicgstab.  first_call != 0 reproduces the first-call side effects
 * (test_solver overwrite of b).  Returns OR_* status; *iters and *err get the
 * iteration count and the final recursive relative residual. */
int or_bicgstab(const or_grid *g, const double *xg, const double *yg,
                const double *zg, const double *sw, double *x, double *b,
                double tol, int test_solver, int first_call, int itmax,
                int *iters, double *err);

/* divergence_calc, diagnostics.f90:120-150 (interior; halo of div kept) */
void or_divergence(const or_grid *g, const double *u, const double *v,
                   const double *w, double *div);

/* 1: blocked summation in or_inn_prod (sensitivity measurements only) */
void or_set_sum_blocked(int on);
/* host threads of the element-wise loops (1: serial, the default); returns
 * the count in effect (1 without OpenMP) */
int or_set_threads(int n);
/* 0: pamm/sfvt/advection_3d.F90 (linked into the model), 1: sfvt/advection_3d.F90 */
void or_set_mpdata_variant(int v);

void or_advance_momentum(const or_grid *g, double dt, double *u, double *v,
                         double *w, const double *zu, const double *zv,
                         const double *zw, const double *su, const double *sv,
                         const double *sw, const double *p);
void or_advance_scalars(const or_grid *g, double dt, double *q,
                        const double *sq, double *th, const double *sth,
                        int nq, int moisture);

/* ---- the model_driver time step (driver_code.f90:304-676), hot path ---- */
typedef struct {
    /* configuration */
    double dt, ptol;
    int kord, monotone, advection_scheme;   /* 0 upwind, 1 mpdata, 2 mpdata_3d_add */
    int moisture, nq, ncat;
    const int *c_s, *c_e;                   /* 1-based category bounds */
    /* 1-D base state (kp+2) */
    const double *theta, *thetan;
    const double *xg, *yg, *zg;             /* x, y, z coordinates (n+2) */
    /* prognostic state: three momentum levels (ut,zut,tut ...) */
    double *ut, *zut, *tut, *vt, *zvt, *tvt, *wt, *zwt, *twt;
    double *th, *p, *q;                      /* q has nq padded fields */
    /* work fields owned by the caller */
    double *su, *sv, *sw, *sth, *psrc, *sq;
    double *ubar, *vbar, *wbar, *thbar, *qbar; /* (kp+2), qbar (kp+2)*nq */
    /* driver state */
    int n;              /* completed steps (drives the pointer rotation) */
    int solver_first;   /* 1 until the first bicgstab call */
    double t1, t2;      /* saved lbc/ubc of the theta advection */
    int last_iters;     /* bicgstab iterations of the last step */
    double last_err;    /* its final recursive residual */
    /* physics options of sources (NULL: the inviscid branch) and the
     * subgrid work fields (strain, vism, vist; viss has nq fields) */
    const or_physics *phys;
    double *strain, *vism, *vist, *viss;
    /* lbc/ubc of theta + thetan in advection_scheme 2: thetan(0) and
     * thetan(kp+1) (find_base_top, mpi_module.f90:306-329), updated in
     * place by each mpdata_3d call like the reference's INOUT scalars */
    double thbase, thtop;
} or_model;

/* initial z-boundary rows of the previous time level (driver_code.f90:287-295) */
void or_model_init_bcs(const or_grid *g, or_model *m);
/* one step; returns the bicgstab status */
int or_model_step(const or_grid *g, or_model *m);
/* current / previous / future level pointers after m->n steps */
void or_model_levels(or_model *m, double **u, double **v, double **w,
                     double **zu, double **zv, double **zw, double **tu,
                     double **tv, double **tw);

#ifdef __cplusplus
}
#endif
#endif
