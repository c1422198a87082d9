// dcm_internal.h -- shared declarations of the HIP implementation.
//
// Device layout: every 3-D field is the reference's padded Fortran array
// (0:kp+1, 0:jp+1, 0:ip+1), k fastest.  A wavefront of 64 lanes walks 64
// consecutive k of one (j,i) column, so every stencil load along k is one
// coalesced 512-byte access and the j/i neighbours are the same addresses one
// column (nk*8 bytes) or one plane (nj*nk*8 bytes) away.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dcm_hip.h"

namespace dcm {

void set_error(const std::string &msg);

#define DCM_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            ::dcm::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return DCM_HIP_ERROR;                                            \
        }                                                                    \
    } while (0)

#define DCM_CHECK(expr)                                                      \
    do {                                                                     \
        int s_ = (expr);                                                     \
        if (s_ != DCM_OK) return s_;                                         \
    } while (0)

#define DCM_LAUNCH_CHECK()                                                   \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess) {                                              \
            ::dcm::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
            return DCM_HIP_ERROR;                                            \
        }                                                                    \
    } while (0)

// Grid geometry passed by value to kernels.
struct Geo {
    int ip, jp, kp;
    int nk, nj, ni;       // padded extents kp+2, jp+2, ip+2
    long long sj, si;     // strides of j and i in elements (nk, nk*nj)
    long long ntot;       // padded elements
};

__host__ __device__ inline long long pidx(const Geo &g, int k, int j, int i) {
    return (long long)i * g.si + (long long)j * g.sj + k;
}

// Device-side 1-D metrics (length n+2, Fortran index).
struct Metrics {
    const double *dx, *dy, *dz, *dxn, *dyn, *dzn, *rhoa, *rhoan;
    // 1/(dx(i-1)*dxn(i)) style Laplacian coefficients (set_mat_a,
    // dynamics.f90:84-96), FP64, indexed by the cell index.
    const double *ae, *aw, *an, *as, *au, *ad;
};

struct Metrics32 {
    const float *dx, *dy, *dz, *dxn, *dyn, *dzn, *rhoa, *rhoan;
    const float *ae, *aw, *an, *as, *au, *ad;
};

// FP32 coefficient tables of the mixed path's MPDATA (dcm_mpdata32.hip),
// each quotient of FP64 metrics rounded once: spacings, their reciprocals
// (idy = 1/dy, idyn = 1/dyn), the reciprocal half-sums of the cell widths
// (ihy(j) = 2/(dyn(j-1) + dyn(j)), 0 at j = 0), and along z the densities
// and 1/rhoan.  Along x, read once per i plane by a whole block, the four
// values of plane i are packed: xpk(i) = (dx(i), 1/dxn(i), 1/dx(i-1),
// 2/(dxn(i-1) + dxn(i))), zeros where i-1 < 0.  Fortran index, n+2 entries.
struct Tab32 {
    const float4 *xpk;
    const float *dy, *idy, *idyn, *ihy;
    const float *dz, *idz, *idzn, *ihz, *ra, *rn, *irn;
};

// Scalars of the BiCGSTAB recurrences, resident on the device.
struct SolverScalars {
    double rho, nrm, alf, omg, bet;
    double sc_err, err, bb, rr;
    double tt, ts;
    double mean;       // mean removed from an inner right-hand side
    int breakdown;     // |omega| < tiny
    int done;          // converged (device-side early exit for chunks)
    int nonfinite;
    int iters;         // iterations performed (counted on the device)
    int guard;         // inner solve: a non-finite or zero-denominator scalar
                       // ends the iteration as a breakdown, before the update
    int best_it;       // inner solve: iteration of the best residual so far
    double best_err;
    double tol;        // convergence test of this solve (err < tol)
    int dvalid;        // inner solve: its correction has been updated at least once
                       // (its start does not store the zero correction)
    int pad_;
};

// State of the mixed path's FP64 refinement kept on the device, so that a
// refinement step (correction, true residual, decision, next inner start)
// can follow an inner iteration without a host round trip.
struct OuterState {
    double sc_err, err_prev, err, rtol;
    int first;        // no residual evaluated yet
    int active;       // an inner solve is under way
    int go;           // this transition runs (the inner solve has finished)
    int outer;        // inner solves started
    int status;       // DCM_NONFINITE / DCM_NOT_CONVERGED, else 0
    int fallback;     // the last correction did not reduce the true residual
    int total_iters;  // inner iterations of the finished inner solves
    int pad;
};

// Per-kernel timing (profiling mode): accumulated over launches.
struct KernelTimer {
    double total_ms = 0.0;
    int launches = 0;
};

struct Mirror {
    void *dev = nullptr;
    size_t bytes = 0;
};

// Result of one pressure solve.
// A grow-only device buffer, freed with its owner.
struct DeviceScratch {
    void *ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes, void **out) {
        if (cap < bytes) {
            if (ptr) {
                (void)hipDeviceSynchronize();
                (void)hipFree(ptr);
                ptr = nullptr;
                cap = 0;
            }
            if (hipMalloc(&ptr, bytes) != hipSuccess) return DCM_OUT_OF_MEMORY;
            cap = bytes;
        }
        *out = ptr;
        return DCM_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

struct SolverOut {
    int iters = 0;   // BiCGSTAB iterations (inner iterations for mixed)
    int outer = 0;   // refinement sweeps (mixed)
    int fallback = 0; // mixed: the FP64 solver finished the solve
    double err = 0.0;
};

}  // namespace dcm

struct dcm_ctx {
    dcm::Geo geo;
    int device = 0;
    hipStream_t stream = nullptr;
    // device metrics (fp64 and fp32 copies), one allocation
    double *dmet = nullptr;
    float *dmet32 = nullptr;
    dcm::Metrics met;
    dcm::Metrics32 met32;
    float *dtab32 = nullptr;         // the FP32 MPDATA tables, one allocation
    dcm::Tab32 tab32;
    // scratch of the batched FP32 MPDATA: mp32_count padded FP32 fields
    float *mp32 = nullptr;
    size_t mp32_count = 0;
    std::vector<double> h_dx, h_dy, h_dz, h_dxn, h_dyn, h_dzn, h_rhoa, h_rhoan;
    // reduction scratch
    double *partials = nullptr;      // per-block partial sums (several slots)
    size_t partials_cap = 0;         // doubles
    double *partials_adv = nullptr;  // the FP32 MPDATA's, which may run beside the solve
    size_t partials_adv_cap = 0;
    double *dscal = nullptr;         // generic device scalars (64 doubles)
    double *hscal = nullptr;         // pinned host mirror (64 doubles)
    dcm::SolverScalars *solver = nullptr;   // device
    dcm::SolverScalars *h_solver = nullptr; // pinned host
    // lazily allocated work fields (FP64 padded) and compact solver vectors
    std::vector<double *> work64;
    std::vector<float *> work32;
    double *sv64 = nullptr;          // solver vectors, fp64
    size_t sv64_count = 0;
    float *sv32 = nullptr;           // solver vectors, fp32
    size_t sv32_count = 0;
    double *hm_partials = nullptr;   // horizontal-mean partials
    size_t hm_cap = 0;
    double *d1d = nullptr;           // scratch for 1-D host arrays (theta...)
    size_t d1d_cap = 0;
    bool first_solve = true;         // bicgstab_first (dynamics.f90:16)
    // global sums in the reference's sequential loop order (one thread per
    // sum) instead of the two-level tree: bit-identical to a serial CPU run
    bool seq_sums = false;
    // Chebyshev relaxation of the mixed inner solve's Jacobi preconditioner
    // (set_chebyshev): the factors of its four sweeps, in order
    bool cheb_on = false;
    // MPDATA arithmetic: 0 pamm/sfvt/advection_3d.F90 (the copy linked into
    // the model), 1 sfvt/advection_3d.F90 (dcm_set_mpdata_variant)
    int mpdata_variant = 0;
    // multigrid V-cycle preconditioner of the mixed inner solve
    // (dcm_set_precond_multigrid; dcm_mg.hip): the hierarchy and the switch
    void *mg = nullptr;
    bool mg_on = false;
    // levels of at most 16384 cells in one single-block kernel (default), or
    // each pass of them in its own launch (the same arithmetic)
    bool mg_tail = true;
    double cheb_a = 0.0;
    int cheb_deg = 4;  // a_p divisions per application (the reference's 4)
    double cheb_w[16] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0,
                         1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    // i planes per marching block of the solver (0: the built-in default)
    int march_planes = 0;
    // device tables of a batched solve and the members' flags (device and
    // pinned host)
    dcm::DeviceScratch solve_tables, flags;
    int *h_flags = nullptr;
    size_t h_flags_cap = 0;
    // the stream is borrowed (an ensemble's), not owned
    bool borrowed_stream = false;
    std::unordered_map<const void *, dcm::Mirror> mirrors;
    // profiling
    bool profiling = false;
    std::string prof_only;           // when non-empty, time only this name
    std::unordered_map<std::string, dcm::KernelTimer> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    // the multigrid inner iteration captured as a HIP graph, valid while the
    // launch arguments it was captured with (the key's bytes) are unchanged
    std::vector<char> iter_key;
    hipGraphExec_t iter_graph = nullptr;
    // the single-system multigrid mixed solve: one inner iteration and the
    // predicated refinement step after it, as one graph; the refinement
    // state on the device and its pinned host copy
    std::vector<char> chunk_key;
    hipGraphExec_t chunk_graph = nullptr;
    dcm::OuterState *outer_d = nullptr;
    dcm::OuterState *outer_h = nullptr;
    // graphs on (dcm_set_graphs): the single-system mixed solve replays its
    // inner iteration from a graph
    bool graphs = true;
};

namespace dcm {

// ---------------------------------------------------------------------------
// helpers implemented in dcm_context.hip
// ---------------------------------------------------------------------------
// DCM_HIP_ERROR when c->stream is being captured (scratch must exist first)
int no_alloc_in_capture(dcm_ctx *c, const char *what);
int work_field64(dcm_ctx *c, int slot, double **out);   // padded fp64 scratch
int work_field32(dcm_ctx *c, int slot, float **out);    // padded fp32 scratch
int ensure_partials(dcm_ctx *c, size_t n);
int ensure_partials_adv(dcm_ctx *c, size_t n);  // the FP32 MPDATA's own partial sums
int ensure_solver_vectors64(dcm_ctx *c, size_t count);
int ensure_solver_vectors32(dcm_ctx *c, size_t count);
int ensure_host_flags(dcm_ctx *c, size_t n, int **out);  // pinned ints
int upload_1d(dcm_ctx *c, const double *h, size_t n, int slot, const double **out);
// profiling bracket: returns an event pair index, or -1 when disabled
int prof_begin(dcm_ctx *c, const char *name);
void prof_end(dcm_ctx *c, int token);
// events for one timed launch of `name` (false when not profiled); the
// launch records them itself (hipExtLaunchKernelGGL), so the measured time
// is the kernel's own execution, as a kernel trace reports it
bool prof_events(dcm_ctx *c, const char *name, hipEvent_t *start, hipEvent_t *stop);

template <typename F, typename... Args>
inline void launch_timed(dcm_ctx *c, const char *name, F kernel, dim3 grid, dim3 block,
                         Args... args) {
    hipEvent_t a = nullptr, b = nullptr;
    if (prof_events(c, name, &a, &b))
        hipExtLaunchKernelGGL(kernel, grid, block, 0, c->stream, a, b, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, c->stream, args...);
}
int prof_collect(dcm_ctx *c);

// ---------------------------------------------------------------------------
// Chebyshev relaxation factors of the mixed inner solve's preconditioner on
// the eigenvalue interval [a, 2] of D^-1 A (a = 0: the reference's Jacobi)
// ---------------------------------------------------------------------------
void set_chebyshev(dcm_ctx *c, double a, int divisions = 4);
// multigrid preconditioner (dcm_mg.hip)
bool mg_possible(const Geo &g);
int mg_prepare(dcm_ctx *c);
template <typename R> int mg_correct(dcm_ctx *c, const R *b, R *x, const SolverScalars *sc);
constexpr float kMgOmega = 0.8f;  // damped-Jacobi weight of the V-cycle's smoother
void mg_release(dcm_ctx *c);

// ---------------------------------------------------------------------------
// resident model fields, for the output writer (dcm_dynamics.hip)
// ---------------------------------------------------------------------------
struct ModelView {
    dcm_ctx *c;
    const double *u, *v, *w, *th, *p, *q, *div;  // current momentum level
    int nq;
};
int model_view(dcm_model *m, ModelView *v);
// allocate the model's divergence diagnostic (zero) and compute it each step
int model_enable_div(dcm_model *m);

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;

struct Interior {
    dim3 grid, block;
};
inline Interior interior_launch(const Geo &g, int kfirst_extra = 0) {
    // lanes along k, the rest of the 256 threads along j; one i per z block
    int nkk = g.kp + kfirst_extra;
    int bx = nkk >= 64 ? 64 : (nkk >= 32 ? 32 : 16);
    int by = kBlock / bx;
    Interior L;
    L.block = dim3(bx, by, 1);
    L.grid = dim3((nkk + bx - 1) / bx, (g.jp + by - 1) / by, g.ip);
    return L;
}
inline unsigned blocks_for(long long n, int bs = kBlock) {
    return (unsigned)((n + bs - 1) / bs);
}

// ---------------------------------------------------------------------------
// fills (dcm_fill.hip)
// ---------------------------------------------------------------------------
enum FillKind { FILL_FULL = 0, FILL_ALONG_DIM, FILL_ALONG_DIM_WO, FILL_FULL_WO, FILL_ALONG_Z };
// lbc/ubc either by value or, when bc_shift is non-null, (bc - *bc_shift)
// bc_mode 1: z planes get bc - *bc_shift; 2: (bc - *bc_shift) + *bc_shift.
// skip_if: when non-null and *skip_if != 0 on the device, the fill is a no-op.
template <typename T>
int fill(dcm_ctx *c, T *a, FillKind kind, double lbc, double ubc,
         const double *bc_shift = nullptr, int bc_mode = 0, const int *skip_if = nullptr);
// up to kFillSetMax arrays filled alike (lbc/ubc by value) in one launch
constexpr int kFillSetMax = 6;
template <typename T> struct FillSet {
    T *a[kFillSetMax];
};
template <typename T>
int fill_set(dcm_ctx *c, const FillSet<T> &set, int count, FillKind kind, double lbc,
             double ubc, const int *skip_if = nullptr);
// Neumann z rows for the matvec: x(0,:,:)=x(1,:,:), x(kp+1,:,:)=x(kp,:,:)
int fill_neumann_z(dcm_ctx *c, double *x);

// ---------------------------------------------------------------------------
// reductions (dcm_reduce.hip)
// ---------------------------------------------------------------------------
// finalize a set of `nsums` partial-sum slots of `nblocks` blocks each into
// dscal[dst + s] (deterministic order)
int finalize_sums(dcm_ctx *c, const double *partials, int nblocks, int nsums,
                  double *dst);
int finalize_min(dcm_ctx *c, const double *partials, int nblocks, double *dst);

// ---------------------------------------------------------------------------
// cross-file entry points
// ---------------------------------------------------------------------------
template <typename R>
int mpdata_core(dcm_ctx *c, R dt, const R *u, const R *v, const R *w, double *psi, int nf,
                const double *lbc, const double *ubc, int kord, int monotone, int neumann,
                int interior_min);
// One call of mpdata_3d (whole_min 1: the minimum over the whole array) or
// mpdata_vec_3d (whole_min 0: per-field interior minima) on nf consecutive
// FP64 fields at psi, with host z-plane values lbc[nf], ubc[nf].
struct Mp32Cat {
    double *psi;
    int nf;
    const double *lbc, *ubc;
    int whole_min;
};
// The mixed path's MPDATA of several such calls at once with the FP64
// velocities u, v, w (padded, halos filled), FP32 intermediates
// (dcm_mpdata32.hip).  Categories the batched kernels do not serve (kord !=
// 2, the sfvt/advection_3d.F90 arithmetic) go through mpdata_core<float>
// with FP32 shadows of u, v, w written to shadow (3 padded fields).  The
// shift of field f of the batch is left at dscal[16 + f].
// true when mpdata32_batch takes its batched FP32 kernels (kord 2, the
// linked copy's arithmetic, 32-bit offsets), whose scratch is its own: the
// FP32 arrays and partials_adv
bool mpdata32_fast(const dcm_ctx *c, int kord);
int mpdata32_batch(dcm_ctx *c, float dt, const double *u, const double *v, const double *w,
                   float *shadow, const Mp32Cat *cats, int ncat, int kord, int monotone,
                   int neumann);
// vertical-only MPDATA of nf fields (FP64); interior_min: the vector form
int mpdata_vert_core(dcm_ctx *c, double dt, const double *w, double *psi, int nf,
                     const double *lbc, const double *ubc, int kord, int monotone, int neumann,
                     int interior_min);
int velocity_shadow(dcm_ctx *c, const double *u, const double *v, const double *w, float *uf,
                    float *vf, float *wf);
int adv_ref_state_dev(dcm_ctx *c, double dt, const double *w, double *psi, const double *ref_dev);
int upwind_dev(dcm_ctx *c, double dt, const double *u, const double *v, const double *w,
               double *psi, int neumann);
int add_profile_dev(dcm_ctx *c, double *psi, const double *prof_dev, double sign);

// Physics options of sources beyond the inviscid branch, device side: 1-D
// profiles are device arrays of kp+2.  log_z1_z0 = log(z(1)/z0) is computed
// on the host once per grid.
struct PhysicsDev {
    int viscous = 0, moisture = 0, damping_layer = 0, forcing = 0, divergence = 0, nq = 0;
    double forcing_tau = 1.0, z0 = 1e-4, z0th = 1e-4, log_z1_z0 = 0.0;
    double z1 = 0.0, theta0 = 0.0, thetan1 = 0.0;
    const double *ubar = nullptr, *vbar = nullptr, *wbar = nullptr;
    const double *dampfacn = nullptr, *dampfac = nullptr;
    const double *u_force = nullptr, *v_force = nullptr, *w_subs = nullptr;
    const double *lamsq = nullptr, *thetan = nullptr;
};
// the 3-D fields the subgrid model reads and writes besides u, v, w, th and
// the momentum / theta tendencies (q, sq and viss have nq fields)
struct SubgridFields {
    const double *zu = nullptr, *zv = nullptr, *zw = nullptr, *q = nullptr;
    double *sq = nullptr, *strain = nullptr, *vism = nullptr, *vist = nullptr, *viss = nullptr;
};
// work-field slots of the six stress components
constexpr int kSlotStress = 80;
int subgrid_dev(dcm_ctx *c, const PhysicsDev &ph, const double *u, const double *v,
                const double *w, const double *th, const SubgridFields &sf, double *su,
                double *sv, double *sw, double *sth);
int relax_dev(dcm_ctx *c, const PhysicsDev &ph, const double *zu, const double *zv,
              const double *zw, const double *th, double *su, double *sv, double *sw,
              double *sth);
// one pressure system of a batched solve: its context (geometry,
// coefficients, scalars, work vectors) and its padded sw, x, b
struct SolveTask {
    dcm_ctx *c;
    const double *sw;
    double *x, *b;
};
int bicgstab64_batch(dcm_ctx *c, const SolveTask *tasks, int n, double tol, int itmax,
                     SolverOut *outs, int *status);
int bicgstab_mixed_batch(dcm_ctx *c, const SolveTask *tasks, int n, double tol, int itmax,
                         double inner_rtol, SolverOut *outs, int *status);
int bicgstab_probe(dcm_ctx *c, const SolveTask *tasks, int n, int iters, SolverOut *outs,
                   int *status);
// single-system multigrid mixed solves of n independent systems, system q on
// streams[q] while it runs (each context's own graphs and scratch)
int bicgstab_mixed_concurrent(const SolveTask *tasks, int n, const hipStream_t *streams,
                              double tol, int itmax, const double *inner_rtol, SolverOut *outs,
                              int *status);
int bicgstab64(dcm_ctx *c, const double *sw, double *x, double *b, double tol, int itmax,
               SolverOut *out);
int bicgstab_mixed(dcm_ctx *c, const double *sw, double *x, double *b, double tol, int itmax,
                   double inner_rtol, SolverOut *out);

}  // namespace dcm
