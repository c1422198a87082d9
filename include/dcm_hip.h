/*
 * dcm_hip.h -- C-ABI of the MI355X dynamical-core hot path (libdcm_hip.so).
 *
 * Every entry point replaces one routine of the reference Fortran model
 * (UoM-maul1609/dynamical-cloud-model) for a domain that lives on one GPU.
 * The Fortran host binds these with ISO_C_BINDING (module dcm_hip,
 * fortran/dcm_hip.f90); Python binds them with ctypes (dcm_amd/hip.py).
 *
 * Conventions
 *   - All 3-D field arguments are DEVICE pointers to FP64 arrays in the
 *     reference's layout: Fortran (0:kp+1, 0:jp+1, 0:ip+1), column major,
 *     k fastest, halo width 1.  The element (k,j,i) is at offset
 *     (i*(jp+2) + j)*(kp+2) + k.
 *   - 1-D arrays (metrics, base state) are HOST pointers of length n+2
 *     (Fortran index 0..n+1) unless stated otherwise; the context keeps
 *     device copies.
 *   - Scalars passed by pointer (lbc, ubc, iteration counts) are HOST
 *     pointers, updated like the Fortran INTENT(INOUT) arguments.
 *   - Every call is enqueued on the context's stream and returns a dcm_status;
 *     calls that return host-visible results synchronise that stream.
 *   - One context serves one domain on the device that was current when it
 *     was created; calls on one context are not re-entrant.
 */
#ifndef DCM_HIP_H
#define DCM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DCM_OK = 0,
    DCM_NOT_CONVERGED = 1, /* dynamics.f90:431-434 ("convergence failure")  */
    DCM_BREAKDOWN = 2,     /* dynamics.f90:414-417 (|omega| < 1e-16)         */
    DCM_NONFINITE = 3,     /* a NaN/Inf residual (the reference loops on)    */
    DCM_BAD_ARG = 4,
    DCM_HIP_ERROR = 5,
    DCM_NO_DEVICE = 6,
    DCM_OUT_OF_MEMORY = 7, /* the reference's "Not enough memory" stops      */
    DCM_IO_ERROR = 8       /* the output file could not be written (the
                            * reference's check() stops on nf90 errors)     */
} dcm_status;

typedef enum { DCM_FP64 = 0, DCM_MIXED = 1 } dcm_precision;

typedef struct dcm_ctx dcm_ctx;
typedef struct dcm_model dcm_model;

/* Text of the last error seen by this thread ("" if none). */
const char *dcm_last_error(void);
/* Version string of the library build. */
const char *dcm_version(void);
/* Free and total device memory of the current device in bytes (the HBM
 * footprint of a workload is the drop of *free_bytes across it).  No
 * counterpart in the reference. */
int dcm_device_memory(size_t *free_bytes, size_t *total_bytes);

/* ---------------------------------------------------------------------- */
/* Context: grid metrics and module-level state.                          */
/* Replaces the module state of d_solver (dynamics.f90:15-17, the saved   */
/* coefficient arrays and bicgstab_first) and the grid arrays that every  */
/* routine receives (initialisation.f90:429-446, 539, 573).               */
/* ---------------------------------------------------------------------- */
int dcm_context_create(int ip, int jp, int kp, const double *dx,
                       const double *dy, const double *dz, const double *dxn,
                       const double *dyn, const double *dzn,
                       const double *rhoa, const double *rhoan,
                       dcm_ctx **out);
int dcm_context_destroy(dcm_ctx *ctx);
/* Number of padded elements (kp+2)*(jp+2)*(ip+2) of one 3-D field. */
size_t dcm_field_elems(const dcm_ctx *ctx);

/* Device memory and transfers on the context's stream. */
int dcm_malloc(dcm_ctx *ctx, size_t bytes, void **dptr);
int dcm_free(dcm_ctx *ctx, void *dptr);
int dcm_memcpy_htod(dcm_ctx *ctx, void *dst, const void *src, size_t bytes);
int dcm_memcpy_dtoh(dcm_ctx *ctx, void *dst, const void *src, size_t bytes);
int dcm_synchronize(dcm_ctx *ctx);

/* Host-array mirrors for the Fortran drop-in: returns (creating on first
 * use) the device copy of a host array of `bytes` bytes. */
int dcm_mirror(dcm_ctx *ctx, const void *host, size_t bytes, void **dptr);
int dcm_mirror_upload(dcm_ctx *ctx, const void *host);
int dcm_mirror_download(dcm_ctx *ctx, void *host);
int dcm_mirror_release_all(dcm_ctx *ctx);

/* ---------------------------------------------------------------------- */
/* a1. Boundary fills (one rank, halo 1), replacing the MPI halo swaps.   */
/* ---------------------------------------------------------------------- */
/* exchange_full (mpi_module.f90:926-1271): faces periodic in x,y; z planes
 * = lbc/ubc; edges and corners touching a z plane take z-wrapped values. */
int dcm_fill_full(dcm_ctx *ctx, double *a, double lbc, double ubc);
/* exchange_along_dim (mpi_module.f90:700-910): faces + z planes only. */
int dcm_fill_along_dim(dcm_ctx *ctx, double *a, double lbc, double ubc);
/* exchange_along_dim_wo (mpi_module.f90:1285-1494): x,y faces only. */
int dcm_fill_along_dim_wo(dcm_ctx *ctx, double *a);
/* exchange_full_wo (mpi_module.f90:1511-1855): as full, no z planes. */
int dcm_fill_full_wo(dcm_ctx *ctx, double *a);
/* exchange_along_z (mpi_module.f90:584-685): z planes only. */
int dcm_fill_along_z(dcm_ctx *ctx, double *a, double lbc, double ubc);

/* ---------------------------------------------------------------------- */
/* a2-a5. Advection (pamm/sfvt/advection_3d.F90, module advection_s_3d).  */
/* ---------------------------------------------------------------------- */
/* Which copy of advection_3d.F90 the MPDATA routines (dcm_mpdata3d,
 * dcm_mpdata_vec3d, dcm_mpdata_vert3d, dcm_mpdata_vec_vert3d, dcm_mpdata3d_add
 * and the model's advection) follow on this context.  DCM_MPDATA_LINKED (the
 * default): pamm/sfvt/advection_3d.F90, the copy linked into the reference's
 * main.exe.  DCM_MPDATA_SFVT: sfvt/advection_3d.F90 (same interfaces,
 * replacing that module when it is the one linked), whose limiter bounds
 * each direction separately with unweighted denominators (:555-682,
 * :1028-1091) and whose vector routines use unweighted antidiffusive
 * velocities (:1384-1507, :1874-1909).  A setting, no counterpart call in
 * the reference, where the choice is made at link time (pamm/Makefile:1-3). */
/* The preconditioner of the pressure solves of a single model, of
 * dcm_bicgstab_mixed and of dcm_bicgstab: on != 0 applies one geometric
 * multigrid V-cycle (damped-Jacobi smoothing, trilinear transfers, every
 * even dimension halved per level and an odd one kept) wherever the
 * reference's iteration applies its
 * Jacobi sweeps (precondition, dynamics.f90:241-270, inside :390-430), so
 * that a solve needs a few iterations at any size.  The mixed path's FP32
 * inner solve runs an FP32 V-cycle (its outer FP64 refinement still ends
 * every solve at the requested true residual); the FP64 BiCGSTAB runs the
 * same cycle in FP64, which is no longer the reference's arithmetic but
 * measures the precision gain at an equal preconditioner.  DCM_BAD_ARG (and
 * no change) when the grid has at most 1024 cells or its dimensions are all
 * odd above 1024 cells.  Off by default.  In a DCM_MIXED ensemble either
 * every member has the setting (each member is then solved with its own
 * V-cycle, the members side by side on the ensemble's streams, and evolves
 * bit for bit as when stepped alone) or none has it (one batched
 * solve with the Jacobi/Chebyshev sweeps); dcm_ensemble_step returns
 * DCM_BAD_ARG for a mix.  on = 1 runs the levels that fit the LDS of one
 * workgroup (with every level below them) as one single-block kernel per
 * V-cycle, the levels of at most 2^17 cells as one kernel per leg and the
 * finest-level transfers as marching kernels; on = 2 launches each pass on
 * its own, one thread per cell (the same arithmetic and bits; kept for
 * comparison).  No counterpart in the
 * reference. */
int dcm_set_precond_multigrid(dcm_ctx *ctx, int on);
#define DCM_MPDATA_LINKED 0
#define DCM_MPDATA_SFVT 1
int dcm_set_mpdata_variant(dcm_ctx *ctx, int variant);
/* first_order_upstream_3d (:39-116): donor cell, psi updated in place. */
int dcm_upwind3d(dcm_ctx *ctx, double dt, const double *u, const double *v,
                 const double *w, double *psi, int neumann);
/* mpdata_3d (:219-714): positive-definite shift, kord passes, optional
 * non-oscillatory limiter.  lbc and ubc are shifted and restored as in the
 * reference.  *early_exit (may be NULL) reports the zero-sum exit. */
int dcm_mpdata3d(dcm_ctx *ctx, double dt, const double *u, const double *v,
                 const double *w, double *psi, double *lbc, double *ubc,
                 int kord, int monotone, int neumann, int *early_exit);
/* mpdata_vec_3d (:1250-1753): nq contiguous padded fields share field 1's
 * antidiffusive and limited velocities. lbc/ubc are arrays of nq. */
int dcm_mpdata_vec3d(dcm_ctx *ctx, double dt, const double *u,
                     const double *v, const double *w, double *psi, int nq,
                     double *lbc, double *ubc, int kord, int monotone,
                     int neumann, int *early_exit);
/* adv_ref_state (:1148-1218): psi_ref is a host array (kp+2). */
int dcm_adv_ref_state(dcm_ctx *ctx, double dt, const double *w, double *psi,
                      const double *psi_ref);
/* mpdata_3d_add (:139-190): psi_1d (host array, kp+2) added to every
 * column, halo levels included, then mpdata_3d, then psi_1d removed. */
int dcm_mpdata3d_add(dcm_ctx *ctx, double dt, const double *u, const double *v,
                     const double *w, double *psi, const double *psi_1d, double *lbc,
                     double *ubc, int kord, int monotone, int neumann, int *early_exit);
/* Vertical-only advection, used by PAMM sedimentation (pamm/microphysics.f90:
 * 1522-1560).  first_order_upstream_vert_3d (:737-791): donor cell along z,
 * psi updated in place; neumann 2 gives the floor face of level 1 the
 * velocity w(1) (sedimentation out of the domain), neumann 1 zero-gradient
 * fluxes at floor and lid. */
int dcm_upwind_vert3d(dcm_ctx *ctx, double dt, const double *w, double *psi, int neumann);
/* mpdata_vert_3d (:820-1126): as dcm_mpdata3d with the vertical velocity
 * only; shift minimum over the whole array. */
int dcm_mpdata_vert3d(dcm_ctx *ctx, double dt, const double *w, double *psi, double *lbc,
                      double *ubc, int kord, int monotone, int neumann, int *early_exit);
/* mpdata_vec_vert_3d (:1779-2087): nq contiguous padded fields share field
 * 1's velocities and limiter; shift minima over the interior. */
int dcm_mpdata_vec_vert3d(dcm_ctx *ctx, double dt, const double *w, double *psi, int nq,
                          double *lbc, double *ubc, int kord, int monotone, int neumann,
                          int *early_exit);

/* ---------------------------------------------------------------------- */
/* a6. horizontal_means (diagnostics.f90:31-107).  Outputs are HOST arrays */
/* of length kp+2 (qbar: nq*(kp+2), field-major); any may be NULL.        */
/* ---------------------------------------------------------------------- */
int dcm_hmeans(dcm_ctx *ctx, const double *u, const double *v,
               const double *w, const double *th, const double *q, int nq,
               double *ubar, double *vbar, double *wbar, double *thbar,
               double *qbar);

/* divergence_calc (diagnostics.f90:120-150): div = rhoan(dU/dx + dV/dy) +
 * d(rhoa w)/dz on the interior; the halo of div is left as passed. */
int dcm_divergence(dcm_ctx *ctx, const double *u, const double *v,
                   const double *w, double *div);

/* ---------------------------------------------------------------------- */
/* a7. sources (dynamics.f90:639-1000), inviscid branch with damping,     */
/* forcing and subsidence off.  theta/thbar are HOST arrays (kp+2).       */
/* ---------------------------------------------------------------------- */
int dcm_sources(dcm_ctx *ctx, const double *u, const double *v,
                const double *w, const double *th, const double *theta,
                const double *thbar, double *su, double *sv, double *sw,
                double *sth, double *rhs);

/* Options of sources beyond the inviscid branch (dynamics.f90:639-1000):
 * the subgrid model (viscous; sgm/subgrid_3d.f90:38-493), the damping layer
 * (:802-839), forcing (:841-860) and subsidence (divergence, :863-873).
 * Profiles are HOST arrays of kp+2 and may be NULL when no option set reads
 * them: lamsq (squared mixing length on w levels), z (heights of the w
 * levels), theta/thetan (base state) for viscous; dampfacn/dampfac for the
 * damping layer; u_force/v_force for forcing; w_subs for subsidence.
 * Radiation is not on this path. */
typedef struct dcm_physics {
    int viscous, damping_layer, forcing, divergence;
    double forcing_tau, z0, z0th;
    const double *lamsq, *dampfacn, *dampfac, *u_force, *v_force, *w_subs;
    const double *z, *theta, *thetan;
} dcm_physics;
/* calculate_subgrid_3d (sgm/subgrid_3d.f90:38-357): Smagorinsky-Lilly
 * viscosities with Richardson-number stability functions and the
 * Monin-Obukhov surface layer; adds the viscous terms to su, sv, sw, sth
 * (and to the nq fields of sq when nq > 0, copying vist into viss) and
 * fills them as the reference does.  Device fields; zu, zv, zw are the
 * previous time level. */
int dcm_subgrid(dcm_ctx *ctx, const dcm_physics *ph, const double *u, const double *v,
                const double *w, const double *zu, const double *zv, const double *zw,
                const double *th, const double *q, int nq, double *vism, double *vist,
                double *viss, double *strain, double *su, double *sv, double *sw, double *sth,
                double *sq);
/* sources with every option but radiation: the inviscid terms, then the
 * options of ph, then the floor/lid rows and the Poisson right-hand side.
 * ubar, vbar, wbar are the horizontal means (HOST, kp+2; needed by the
 * damping layer and subsidence), theta and thbar HOST arrays as in
 * dcm_sources; the subgrid fields are used when ph->viscous. */
int dcm_sources_phys(dcm_ctx *ctx, const dcm_physics *ph, const double *ubar,
                     const double *vbar, const double *wbar, const double *u,
                     const double *v, const double *w, const double *zu, const double *zv,
                     const double *zw, const double *th, const double *q, int nq,
                     const double *theta, const double *thbar, double *su, double *sv,
                     double *sw, double *sth, double *sq, double *strain, double *vism,
                     double *vist, double *viss, double *rhs);

/* ---------------------------------------------------------------------- */
/* a8-a12. Pressure solver (dynamics.f90:24-482).                         */
/* ---------------------------------------------------------------------- */
/* mat_ax (:187-229): fills the halo of x like the reference, ax = A x. */
int dcm_matvec(dcm_ctx *ctx, double *x, double *ax);
/* precondition (:241-270): prit Jacobi sweeps, s = M r. */
int dcm_precondition(dcm_ctx *ctx, double *s, const double *r, int prit);
/* Order of every global sum of the context (dot products, horizontal
 * means).  DCM_SUM_TREE (default): a fixed two-level tree, deterministic run
 * to run.  DCM_SUM_SEQUENTIAL: the reference's serial loop order (i outer, j,
 * k inner; inn_prod :470-476, diagnostics.f90:55-60), one GPU thread per sum;
 * bit-identical to a serial CPU run, slow, meant for verification.  A new
 * context starts sequential when the environment has DCM_SUM_ORDER=sequential. */
enum { DCM_SUM_TREE = 0, DCM_SUM_SEQUENTIAL = 1 };
int dcm_set_sum_order(dcm_ctx *ctx, int order);
/* The mixed path's FP32 inner solve (no counterpart in the reference, whose
 * solve is FP64 throughout) may relax the four sweeps of the Jacobi
 * preconditioner with Chebyshev factors for the eigenvalue interval
 * [a, 2] of D^-1 A, with `divisions` (3..16) applications of 1/a_p per
 * preconditioner call (the reference's s = r/a_p and three sweeps are 4);
 * the default is a = 0.002 with 12 divisions.  a = 0 restores the
 * reference's four plain sweeps.  The preconditioner stays a fixed linear map at the same cost;
 * the FP64 path and the FP64 refinement are not affected. */
int dcm_set_precond_chebyshev(dcm_ctx *ctx, double a, int divisions);
/* inn_prod (:458-482): interior dot product, result to *out (host). */
int dcm_inn_prod(dcm_ctx *ctx, const double *x, const double *y, double *out);
/* dcm_inn_prod that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels, same result bits; the partial sums live in a fixed 2 MiB device
 * array (each device the library runs on holds it until the process ends)
 * instead of the context's grow-only buffer.  Grids whose interior launch
 * has more than 2^18 blocks (ceil(kp/64)*ceil(jp/4)*ip for kp >= 64;
 * 512x512x256 is the largest such cube-like grid) get DCM_BAD_ARG.  Every
 * call on one device shares that array: calls that overlap on one device
 * (different streams or threads) race and must be serialised by the
 * caller. */
int dcm_inn_prod_noalloc(dcm_ctx *ctx, const double *x, const double *y, double *out);
/* bicgstab (:293-444).  xg,yg,zg are host coordinate arrays used only for
 * the test_solver right-hand side on the context's first call.  itmax <= 0
 * means the reference's 9999.  *iters / *err (may be NULL) receive the
 * iteration count and the final recursive residual. */
int dcm_bicgstab(dcm_ctx *ctx, const double *xg, const double *yg,
                 const double *zg, const double *sw, double *x, double *b,
                 double tol, int test_solver, int itmax, int *iters,
                 double *err);
/* Mixed-precision variant: FP32 BiCGSTAB inner solves inside FP64
 * iterative refinement on the true residual ||b-Ax||/max(||b||,0.01).
 * *iters receives the total inner iterations, *outer the refinements. */
int dcm_bicgstab_mixed(dcm_ctx *ctx, const double *sw, double *x, double *b,
                       double tol, int itmax, int *iters, int *outer,
                       double *err);
/* True residual ||b-Ax||_2 / max(||b||_2, 0.01) of a padded x (host out). */
int dcm_true_residual(dcm_ctx *ctx, const double *x, const double *b,
                      double *out);

/* a13. advance_momentum (dynamics.f90:1019-1099); dt is the leapfrog 2*dt. */
int dcm_advance_momentum(dcm_ctx *ctx, double dt, double *u, double *v,
                         double *w, const double *zu, const double *zv,
                         const double *zw, const double *su, const double *sv,
                         const double *sw, const double *p);
/* a14. advance_scalar_fields_3d (sgm/subgrid_3d.f90:582-623). */
int dcm_advance_scalars(dcm_ctx *ctx, double dt, double *q, const double *sq,
                        double *th, const double *sth, int nq, int moisture);

/* ---------------------------------------------------------------------- */
/* a15. The model_driver time step (driver_code.f90:287-295, 304-676) on  */
/* device-resident state.  The model owns every field; the host reads or  */
/* writes them with dcm_model_get / dcm_model_set (host arrays).          */
/* ---------------------------------------------------------------------- */
enum {
    DCM_F_U = 0, DCM_F_V, DCM_F_W,        /* current momentum level        */
    DCM_F_ZU, DCM_F_ZV, DCM_F_ZW,         /* previous level                */
    DCM_F_TH, DCM_F_P, DCM_F_PSRC,
    DCM_F_SU, DCM_F_SV, DCM_F_SW, DCM_F_STH,
    DCM_F_Q,                              /* q field index in `index`      */
    DCM_F_THBAR,                          /* 1-D (kp+2)                    */
    DCM_F_COUNT
};

typedef struct {
    int steps;            /* steps completed by the call                    */
    int iters_total;      /* BiCGSTAB iterations (inner, for mixed)         */
    int iters_max;        /* largest per-step iteration count               */
    int outer_total;      /* refinement sweeps (mixed only)                 */
    int fallbacks;        /* mixed solves finished by the FP64 solver       */
    int reserved;         /* zero                                           */
    double last_err;      /* last step's final residual                     */
    double max_true_resid;/* largest FP64 true residual seen (0 if unchecked)*/
} dcm_step_stats;

/* theta/thetan/xg/yg/zg: host 1-D arrays; c_s/c_e: 1-based category bounds
 * into q (driver_code.f90:571-580).  advection_scheme (driver_code.f90:
 * 511-561): 0 first_order_upstream_3d of theta then exchange_full(th, t1,
 * t2) (FP64 on both precisions); 1 mpdata_3d; 2 mpdata_3d_add of theta +
 * thetan with thetan(0), thetan(kp+1) as its bcs and no adv_ref_state.  In
 * scheme 2 dcm_model_step returns DCM_BAD_ARG if a step's shift would have
 * changed those bcs by rounding ((bc - m) + m != bc; never for potential
 * temperatures, where m <= bc <= 2m). */
int dcm_model_create(dcm_ctx *ctx, int nq, int ncat, const int *c_s,
                     const int *c_e, double dt, int kord, int monotone,
                     int advection_scheme, double ptol, const double *theta,
                     const double *thetan, const double *xg, const double *yg,
                     const double *zg, dcm_model **out);
int dcm_model_destroy(dcm_model *m);
/* Physics options of the model's sources (see dcm_physics); NULL restores
 * the inviscid branch.  The profiles are copied; the model then zeroes sth
 * and sq each step (driver_code.f90:340-343), computes the horizontal means
 * the damping layer and subsidence read, and applies the scalar update
 * q += sq*dt, th += sth*dt (:601-603). */
int dcm_model_set_physics(dcm_model *m, const dcm_physics *ph);
int dcm_model_set_precision(dcm_model *m, int precision);
/* host <-> device copies of one model field (index: q field, else 0) */
int dcm_model_set(dcm_model *m, int field, int index, const double *host);
int dcm_model_get(dcm_model *m, int field, int index, double *host);
/* device pointer of a model field (for zero-copy callers) */
int dcm_model_field(dcm_model *m, int field, int index, double **dptr);
/* Initial halos (initialisation.f90:861-870) and the previous-level rows of
 * driver_code.f90:287-295; call once after the fields are set. */
int dcm_model_init(dcm_model *m);
/* Advance nsteps time steps; stats may be NULL. */
int dcm_model_step(dcm_model *m, int nsteps, dcm_step_stats *stats);
/* Average kernel time (ms) of the named kernel over the last timed region,
 * and the number of launches it had; returns DCM_BAD_ARG if unknown. */
int dcm_model_profile(dcm_model *m, int enable);
/* Time only the kernel groups named in `name`, a comma-separated list
 * (NULL or "": every group). */
int dcm_model_profile_filter(dcm_model *m, const char *name);
/* Relative residual reduction of each FP32 inner solve (mixed; default 1e-4). */
int dcm_model_set_inner_rtol(dcm_model *m, double rtol);
/* When on, every pressure solve of the FP64 path is followed by the FP64 true
 * residual ||b-Ax||/max(||b||,0.01); the largest goes to max_true_resid
 * (the mixed path always measures it). */
int dcm_model_set_check_residual(dcm_model *m, int on);
/* HIP graphs in dcm_model_step (default on): the kernels before and after
 * the pressure solve, and each multigrid inner iteration of the mixed solve,
 * are captured once and replayed (per-kernel timing, dcm_model_profile,
 * launches them one by one).  Off: every kernel is launched from the host.
 * The results are the same bits either way.  No counterpart in the
 * reference, which has no device launches. */
int dcm_model_set_graphs(dcm_model *m, int on);

/* ---------------------------------------------------------------------- */
/* Ensembles (batching): independent models of uneven size stepped        */
/* together.  Every member's kernels run on one stream of the ensemble,   */
/* and the pressure solves of all members are one batched BiCGSTAB: each  */
/* launch serves every member that has not converged.  No counterpart in  */
/* the reference, which runs one domain per job.  Members keep their own  */
/* models and contexts (created as usual) and must share precision and    */
/* ptol; while in an ensemble a member must not be stepped on its own.    */
/* ---------------------------------------------------------------------- */
typedef struct dcm_ensemble dcm_ensemble;
int dcm_ensemble_create(int n, dcm_model *const *models, dcm_ensemble **out);
int dcm_ensemble_destroy(dcm_ensemble *e);
/* Time the batched iteration at each candidate i-plane count of the solver
 * tiles (4, 8, 16) on the members' current systems and keep the fastest;
 * the model state is not changed.  *march_planes (may be NULL) gets it. */
int dcm_ensemble_autotune(dcm_ensemble *e, int *march_planes);
/* Advance every member nsteps steps; stats (may be NULL) is an array of n.
 * From the second step on, the stages before and after the pressure solve
 * are replayed from HIP graphs captured on first use (one per buffer
 * arrangement of the time levels), unless graphs are switched off or a member
 * is being profiled. */
int dcm_ensemble_step(dcm_ensemble *e, int nsteps, dcm_step_stats *stats);
/* Graph capture of the step's stages on (default) or off; off releases the
 * captured graphs. */
int dcm_ensemble_set_graphs(dcm_ensemble *e, int enable);

/* ---------------------------------------------------------------------- */
/* f2. Namelist and initialisation (main.f90:74-140, initialisation.f90).  */
/* ---------------------------------------------------------------------- */
/* Fortran namelist files (`read(8,nml=run_vars)`, main.f90:77-78 and
 * :137-138): groups, derived-type components (nm1%dt), subscripted
 * sections (q_read(1,1:7)), repeat counts, logicals, character values.
 * Names are case-insensitive; the last assignment wins.  Getters return
 * DCM_BAD_ARG when the object is absent or not of the asked type. */
typedef struct dcm_namelist dcm_namelist;
int dcm_nml_read(const char *path, dcm_namelist **out);
int dcm_nml_free(dcm_namelist *nl);
/* 1 if the group (name NULL) or the object exists, else 0 */
int dcm_nml_has(const dcm_namelist *nl, const char *group, const char *name);
int dcm_nml_get_real(const dcm_namelist *nl, const char *group, const char *name, double *v);
int dcm_nml_get_int(const dcm_namelist *nl, const char *group, const char *name, int *v);
int dcm_nml_get_logical(const dcm_namelist *nl, const char *group, const char *name, int *v);
int dcm_nml_get_string(const dcm_namelist *nl, const char *group, const char *name, char *buf,
                       size_t n);
/* an array of `rank` dimensions `shape` (Fortran order, column-major out):
 * every assignment to the object is applied in order; elements no
 * assignment touches keep what out held */
int dcm_nml_get_array(const dcm_namelist *nl, const char *group, const char *name, int rank,
                      const int *shape, double *out);

/* The run configuration allocate_and_set reads (the nm1 components of
 * variables.f90:59-93 it uses, and the sounding). */
typedef struct dcm_init_config {
    int ip, jp, kp;
    double dx, dy, dz, dt, runtime;
    int moisture, theta_flag, nq;
    int iqv, iqc, inc;          /* 1-based q fields of vapour, cloud mass, cloud number; 0: none */
    double drop_num;
    double psurf, tsurf;
    int adiabatic_prof;
    double adiabatic_frac, t_cbase, t_ctop, rh_above, th_jump, th_grad;
    int n_levels;
    const double *z_read, *theta_read, *u_read, *v_read;  /* n_levels each */
    const double *q_read;       /* q_read(nq, n_levels), Fortran order */
    int damping_layer;
    double damping_thickness, damping_tau;
    int forcing;
    double forcing_tau;
    int divergence;
    double divergence_val, divergence_hgt;
    double cvis, z0, z0th;
    int bubble, param_theta, param_wind;
    double param_vmax, param_z, param_sigz, param_delz;
    unsigned long long seed;    /* of the param_theta noise */
} dcm_init_config;
/* Profiles of allocate_and_set, in caller-allocated arrays: x, xn, dx, dxn
 * of ip+2; y, yn, dy, dyn of jp+2; the others kp+2 (Fortran index 0..n+1);
 * q nq*(kp+2), field n's column at q[(n-1)*(kp+2)]. */
typedef struct dcm_init_profiles {
    double *x, *y, *z, *xn, *yn, *zn, *dx, *dy, *dz, *dxn, *dyn, *dzn;
    double *theta, *thetan, *rhoa, *rhoan, *pref, *prefn, *tref, *trefn;
    double *lamsq, *lamsqn, *dampfac, *dampfacn, *u_force, *v_force, *w_subs;
    double *u, *v;              /* the initial winds, the same in every column */
    double *q;                  /* the initial moisture fields, the same in every column */
    double thbase, thtop;       /* thetan(0), thetan(kp+1) (find_base_top) */
    int ntim;                   /* ceiling(runtime/dt) */
} dcm_init_profiles;
/* The 1-D part of allocate_and_set (initialisation.f90:117-888) on one
 * rank: grid, mixing length, damping layer, hydrostatic reference state
 * (the sounding's or setup_adiabatic_profile's, :899-1311), winds, forcing,
 * subsidence and moisture columns.  Host code; VODE and zeroin are
 * replaced by the exact Exner integral where theta(z) is known and by
 * RK4 plus Brent's method on the moist adiabat. */
int dcm_init_compute_profiles(const dcm_init_config *cfg, dcm_init_profiles *prof);
/* The 3-D initial state generated on the GPU from the profiles: u, v from
 * the wind columns, w = 0, th the param_theta noise (a counter-based hash
 * of the cell and cfg->seed replaces Fortran's random_number) or the warm
 * bubble, q from the moisture columns; then the halos as dcm_model_init
 * fills them.  The model must have been created on a context of the
 * profiles' grid, with cfg->nq q fields when moisture is set. */
int dcm_model_init_fields(dcm_model *m, const dcm_init_config *cfg,
                          const dcm_init_profiles *prof);

/* ---------------------------------------------------------------------- */
/* f3. Output file (driver_code.f90:725-1145, `output`).                   */
/* ---------------------------------------------------------------------- */
/* The reference's NetCDF file, written in the classic format by this
 * library: dims times (unlimited), ip, jp, kp [, nq, nprec, l_q_names];
 * float variables time, x, y, z, rhoa, thetan, trefn, th, u, v, w, p, div
 * [, q_names, q, precip] with the reference's names, dimension order and
 * units.  x, y, z, rhoa, thetan, trefn are the reference's (0:n+1) host
 * arrays (entries 1..n are written); q_names is nq blank-padded 20-character
 * names (the Fortran character(len=20) array), or NULL for blanks. */
typedef struct dcm_ncfile dcm_ncfile;
/* Host writer, for callers whose fields are on the host: creates the file
 * (NF90_CLOBBER, :788-964) and writes the fixed variables (:1018-1046). */
int dcm_nc_create(const char *path, int ip, int jp, int kp, int moisture, int nq, int nprec,
                  const char *q_names, const double *x, const double *y, const double *z,
                  const double *rhoa, const double *thetan, const double *trefn,
                  dcm_ncfile **out);
/* Record n (1-based, the reference's `n`) from host arrays in the padded
 * reference layout; q holds nq fields one after another, precip is
 * (1:kp, 0:jp+1, 0:ip+1, nprec) or NULL (NetCDF's fill value). */
int dcm_nc_put_record(dcm_ncfile *f, int n, double time, const double *u, const double *v,
                      const double *w, const double *th, const double *p, const double *div,
                      const double *q, const double *precip);
int dcm_nc_close(dcm_ncfile *f);

/* Asynchronous writer attached to a resident model.  dcm_output_write
 * snapshots the model's fields on the model's stream (a pack kernel turns
 * them into the record's big-endian floats in HBM, so the next step may
 * start at once), copies the record to pinned host memory on a copy stream
 * and returns; a host thread writes it to the file.  Two records may be in
 * flight; a third call waits for the oldest.  div is the divergence of the
 * momentum at the start of the previous step (:443-444), so the model
 * computes it every step once an output is attached (zero before the first
 * step).  q is written when moisture is set; precip is a host array as in
 * dcm_nc_put_record, or NULL.  Destroy the output before its model. */
typedef struct dcm_output dcm_output;
int dcm_output_create(dcm_model *m, const char *path, int moisture, int nprec,
                      const char *q_names, const double *x, const double *y, const double *z,
                      const double *rhoa, const double *thetan, const double *trefn,
                      dcm_output **out);
int dcm_output_write(dcm_output *o, double time, const double *precip);
/* wait until every record written so far is in the file */
int dcm_output_flush(dcm_output *o);
/* records written, the host time spent waiting for a free record buffer
 * and the writer thread's time in file writes (seconds) */
int dcm_output_stats(dcm_output *o, long long *records, double *stall_s, double *write_s);
int dcm_output_destroy(dcm_output *o);

/* ---------------------------------------------------------------------- */
/* Measurement utility (no counterpart in the reference): a plain device   */
/* copy of n doubles, the STREAM-copy yardstick and the byte calibration   */
/* for hardware counters.  Returns the kernel time in ms when ms != NULL.  */
/* ---------------------------------------------------------------------- */
int dcm_stream_copy(dcm_ctx *ctx, double *dst, const double *src, size_t n,
                    double *ms);
int dcm_model_kernel_time(dcm_model *m, const char *name, double *avg_ms,
                          int *launches);

#ifdef __cplusplus
}
#endif
#endif
