// dcm_advect.hip -- MPDATA advection (pamm/sfvt/advection_3d.F90, the copy
// linked into the reference main.exe) on one GPU.  The context's
// mpdata_variant 1 selects the arithmetic of sfvt/advection_3d.F90 instead:
// its limiter bounds each direction separately with unweighted
// denominators (:555-682, vertical :1028-1091), and its vector routines use
// unweighted antidiffusive velocities (:1384-1507, :1874-1909).
//
// Data flow of one mpdata_3d / mpdata_vec_3d call (kord=2, limiter on):
//   K1 min        m = min(psi)  (whole array for mpdata_3d, interior and per
//                 field for the vector variant), combined with lbc/ubc
//   K2 upwind-1   psi1 = donor(psi-m, u,v,w) into scratch; the same kernel
//                 sums psi-m over the interior for the zero-sum exit test
//   K3 antidiff   U*,V*,W* from (u,v,w, psi1)           Smolarkiewicz 1984
//   K4 beta       beta_up/beta_down from psi-m, psi1, U* Smolarkiewicz &
//                 Grabowski 1990 eqs 19-20
//   K5 upwind-2   donor(psi1) with the limited velocities (eq 18) evaluated on
//                 the fly at the six faces, + m, written straight into psi
// The reference's psi_store copy, the shift/unshift sweeps and the limited
// velocity arrays are never materialised: psi itself is left untouched until
// K5, so psi-m is recomputed wherever the reference reads psi_store.  Each
// quantity is still evaluated by the reference's expression in the
// reference's order, so the FP64 path reproduces the reference bit for bit
// (built with -ffp-contract=off).  The zero-sum exit (:284-289) is decided on
// the device; later kernels test the flag, so no host synchronisation occurs.
//
// The same kernels instantiated with R=float give the mixed-precision path:
// psi1, U*, beta and the velocity shadows are FP32 (half the bytes), while
// psi, the shift m and the output stay FP64.
#include <cfloat>
#include <vector>

#include "dcm_device.h"

namespace dcm {

template <typename R> struct Small;
template <> struct Small<double> { static constexpr double v = 1e-60; };  // :19
// 1e-60 underflows in binary32; the FP32 path uses the smallest normal.
template <> struct Small<float> { static constexpr float v = FLT_MIN; };

struct MetPack {
    Metrics m64;
    Metrics32 m32;
};

template <typename R> struct M1;
template <> struct M1<double> {
    const double *dx, *dy, *dz, *dxn, *dyn, *dzn, *ra, *rn;
    __host__ __device__ M1(const MetPack &p)
        : dx(p.m64.dx), dy(p.m64.dy), dz(p.m64.dz), dxn(p.m64.dxn), dyn(p.m64.dyn),
          dzn(p.m64.dzn), ra(p.m64.rhoa), rn(p.m64.rhoan) {}
};
template <> struct M1<float> {
    const float *dx, *dy, *dz, *dxn, *dyn, *dzn, *ra, *rn;
    __host__ __device__ M1(const MetPack &p)
        : dx(p.m32.dx), dy(p.m32.dy), dz(p.m32.dz), dxn(p.m32.dxn), dyn(p.m32.dyn),
          dzn(p.m32.dzn), ra(p.m32.rhoa), rn(p.m32.rhoan) {}
};

// a / b in the arithmetic of R.  FP64: the correctly rounded quotient of
// the reference.  FP32 (the mixed path's MPDATA): a multiply by the
// hardware reciprocal, within about 1 ulp and a fifth of the instructions;
// every denominator here is at least FLT_MIN (metrics, or sums of the
// non-negative shifted field plus small) and bounds its numerator where it
// can be that small, so no quotient overflows where the IEEE one would not.
template <typename R> __device__ __forceinline__ R qdiv(R a, R b);
template <> __device__ __forceinline__ double qdiv<double>(double a, double b) { return a / b; }
template <> __device__ __forceinline__ float qdiv<float>(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
}

template <typename R> __device__ __forceinline__ R pos0(R x) { return x > R(0) ? x : R(0); }
template <typename R> __device__ __forceinline__ R neg0(R x) { return x < R(0) ? x : R(0); }
template <typename R> __device__ __forceinline__ R min3(R a, R b, R c) {
    R m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    return m;
}

// first_order_upstream_3d arithmetic for one cell (advection_3d.F90:74-115)
template <typename R>
__device__ __forceinline__ R donor(R dt, R uc, R ul, R vc, R vl, R wc, R wl, R pc, R pe, R pw,
                                   R pn, R ps, R pu, R pd, R dxn, R dyn, R dzn, R rak, R rakm,
                                   R rnk, int k, int kp, int neumann) {
    const R two = R(2);
    R fxr = qdiv<R>((((uc + fabs(uc)) * pc + (uc - fabs(uc)) * pe) * dt), (two * dxn));
    R fxl = qdiv<R>((((ul + fabs(ul)) * pw + (ul - fabs(ul)) * pc) * dt), (two * dxn));
    R fyr = qdiv<R>((((vc + fabs(vc)) * pc + (vc - fabs(vc)) * pn) * dt), (two * dyn));
    R fyl = qdiv<R>((((vl + fabs(vl)) * ps + (vl - fabs(vl)) * pc) * dt), (two * dyn));
    R fzr = qdiv<R>(((((wc + fabs(wc)) * rak) * pc + ((wc - fabs(wc)) * rak) * pu) * dt), ((two * dzn) * rnk));
    R fzl = qdiv<R>(((((wl + fabs(wl)) * rakm) * pd + ((wl - fabs(wl)) * rakm) * pc) * dt), ((two * dzn) * rnk));
    if (neumann == 1) {  // :95-101
        if (k == 1) fzl = fzr;
        if (k == kp) fzr = fzl;
    } else if (neumann == 2) {  // :102-110
        if (k == 1) fzl = -fzr;
        if (k == kp) fzr = -fzl;
    }
    return ((pc - (fxr - fxl)) - (fyr - fyl)) - (fzr - fzl);
}

// ---------------------------------------------------------------------------
// K1: minimum (mpdata_3d: whole array :275; vector variant: interior :1307)
// ---------------------------------------------------------------------------
__global__ void k_min(const double *a, Geo g, int interior_only, double *partials) {
    double v = INFINITY;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < g.ntot;
         t += (long long)gridDim.x * blockDim.x) {
        if (interior_only) {
            const int k = (int)(t % g.nk);
            const long long r = t / g.nk;
            const int j = (int)(r % g.nj), i = (int)(r / g.nj);
            if (k < 1 || k > g.kp || j < 1 || j > g.jp || i < 1 || i > g.ip) continue;
        }
        const double x = a[t];
        v = x < v ? x : v;
    }
    block_min_store(v, partials);
}

// m = min(min(a), lbc, ubc); flag cleared
__global__ void k_min_finish(const double *partials, int nb, double lbc, double ubc,
                             double *m, int *flag) {
    double r = reduce_min_1024(partials, nb);
    if (threadIdx.x == 0) {
        if (lbc < r) r = lbc;
        if (ubc < r) r = ubc;
        *m = r;
        *flag = 0;
    }
}

// zero-sum exit test (:282-289): flag = (sum(psi-m) < small)
__global__ void k_exit_flag(const double *partials, int nb, int *flag) {
    double s = reduce_slot_1024(partials, nb);
    if (threadIdx.x == 0) *flag = (s < 1e-60) ? 1 : 0;
}

// Early-exit result: psi = (psi - m) + m over the whole array (:285).
__global__ void k_exit_unshift(double *psi, long long n, const double *m, const int *flag) {
    if (!*flag) return;
    const double mm = *m;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        psi[t] = (psi[t] - mm) + mm;
}

// Velocity shadow for the FP32 path: whole padded arrays.
__global__ void k_to_float3(const double *u, const double *v, const double *w, float *uf,
                            float *vf, float *wf, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        uf[t] = (float)u[t];
        vf[t] = (float)v[t];
        wf[t] = (float)w[t];
    }
}

// ---------------------------------------------------------------------------
// K2: first donor pass on psi-m (also the plain upwind of the shifted field)
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_upwind_first(
    const double *psi, const double *mptr, const R *u, const R *v, const R *w, R *out, Geo g,
    MetPack mp, R dt, int neumann, int do_sum, double *partials) {
    const M1<R> M(mp);
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    const bool act = k <= g.kp && j <= g.jp;
    const double m = *mptr;
    double s[1] = {0.0};
    if (act) {
        const long long c = pidx(g, k, j, i);
        auto P = [&](long long q) { return (R)(psi[q] - m); };
        const R pc = P(c);
        if (do_sum) s[0] = psi[c] - m;
        R res = donor<R>(dt, u[c], u[c - g.si], v[c], v[c - g.sj], w[c], w[c - 1], pc,
                         P(c + g.si), P(c - g.si), P(c + g.sj), P(c - g.sj), P(c + 1), P(c - 1),
                         M.dxn[i], M.dyn[j], M.dzn[k], M.ra[k], M.ra[k - 1], M.rn[k], k, g.kp,
                         neumann);
        out[c] = res;
    }
    if (do_sum) block_sum_store<1>(s, partials, 0);
}

// ---------------------------------------------------------------------------
// plain donor pass on a filled scratch field; output to scratch or to psi+m
// ---------------------------------------------------------------------------
template <typename R, bool TO_PSI>
__global__ __launch_bounds__(256) void k_upwind(const R *src, const R *u, const R *v,
                                                const R *w, R *out, double *psi_out,
                                                const double *mptr, const int *flag, Geo g,
                                                MetPack mp, R dt, int neumann) {
    if (flag && *flag) return;
    const M1<R> M(mp);
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    R res = donor<R>(dt, u[c], u[c - g.si], v[c], v[c - g.sj], w[c], w[c - 1], src[c],
                     src[c + g.si], src[c - g.si], src[c + g.sj], src[c - g.sj], src[c + 1],
                     src[c - 1], M.dxn[i], M.dyn[j], M.dzn[k], M.ra[k], M.ra[k - 1], M.rn[k], k,
                     g.kp, neumann);
    if (TO_PSI)
        psi_out[c] = (double)res + *mptr;
    else
        out[c] = res;
}

// ---------------------------------------------------------------------------
// K3: antidiffusive velocities (advection_3d.F90:318-506).  G: the
// density-weighted form; !G: the unweighted one of sfvt/advection_3d.F90's
// mpdata_vec_3d (:1384-1507).
// ---------------------------------------------------------------------------
template <typename R, bool G>
__global__ __launch_bounds__(256) void k_antidiff(const R *ut, const R *vt, const R *wt,
                                                  const R *po, R *us, R *vs, R *ws,
                                                  const int *flag, Geo g, MetPack mp, R dt) {
    if (flag && *flag) return;
    const M1<R> M(mp);
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long SI = g.si, SJ = g.sj;
    const long long c = pidx(g, k, j, i);
#define U(dk, dj, di) ut[c + (dk) + (dj)*SJ + (di)*SI]
#define V(dk, dj, di) vt[c + (dk) + (dj)*SJ + (di)*SI]
#define W(dk, dj, di) wt[c + (dk) + (dj)*SJ + (di)*SI]
#define P(dk, dj, di) po[c + (dk) + (dj)*SJ + (di)*SI]
    const R small = Small<R>::v;
    const R h = R(0.5), q = R(0.25);
    const R *ra = M.ra, *rn = M.rn;
    const R g1 = rn[k], g2 = rn[k];
    const R g3 = h * (rn[k] + rn[k + 1]);
    const R dxim = M.dx[i - 1], dyjm = M.dy[j - 1], dzkm = M.dz[k - 1];
    if constexpr (!G) {
        const R ud1 = qdiv<R>((((W(0, 0, 0) + W(0, 0, 1)) - W(-1, 0, 0)) - W(-1, 0, 1)), dzkm) +
                      qdiv<R>((((V(0, 0, 0) + V(0, 0, 1)) - V(0, -1, 0)) - V(0, -1, 1)), dyjm);
        const R ud2 = qdiv<R>((((W(0, 0, 0) + W(0, 1, 0)) - W(-1, 0, 0)) - W(-1, 1, 0)), dzkm) +
                      qdiv<R>((((U(0, 0, 0) + U(0, 1, 0)) - U(0, 0, -1)) - U(0, 1, -1)), dxim);
        const R ud3 = qdiv<R>((((U(0, 0, 0) + U(1, 0, 0)) - U(0, 0, -1)) - U(1, 0, -1)), dxim) +
                      qdiv<R>((((V(0, 0, 0) + V(1, 0, 0)) - V(0, -1, 0)) - V(1, -1, 0)), dyjm);
        const R hz = h * (M.dzn[k - 1] + M.dzn[k]), hy = h * (M.dyn[j - 1] + M.dyn[j]),
                hx = h * (M.dxn[i - 1] + M.dxn[i]);
        const R a1 = qdiv<R>((qdiv<R>(((q * (((W(0, 0, 1) + W(0, 0, 0)) + W(-1, 0, 1)) + W(-1, 0, 0))) *
                       (((P(1, 0, 1) + P(1, 0, 0)) - P(-1, 0, 1)) - P(-1, 0, 0))), ((((P(1, 0, 1) + P(1, 0, 0)) + P(-1, 0, 1)) + P(-1, 0, 0)) + small))), hz);
        const R b1 = qdiv<R>((qdiv<R>(((q * (((V(0, 0, 1) + V(0, 0, 0)) + V(0, -1, 1)) + V(0, -1, 0))) *
                       (((P(0, 1, 1) + P(0, 1, 0)) - P(0, -1, 1)) - P(0, -1, 0))), ((((P(0, 1, 1) + P(0, 1, 0)) + P(0, -1, 1)) + P(0, -1, 0)) + small))), hy);
        const R a2 = qdiv<R>((qdiv<R>(((q * (((W(0, 1, 0) + W(0, 0, 0)) + W(-1, 1, 0)) + W(-1, 0, 0))) *
                       (((P(1, 1, 0) + P(1, 0, 0)) - P(-1, 1, 0)) - P(-1, 0, 0))), ((((P(1, 1, 0) + P(1, 0, 0)) + P(-1, 1, 0)) + P(-1, 0, 0)) + small))), hz);
        const R b2 = qdiv<R>((qdiv<R>(((q * (((U(0, 1, 0) + U(0, 0, 0)) + U(0, 1, -1)) + U(0, 0, -1))) *
                       (((P(0, 1, 1) + P(0, 0, 1)) - P(0, 1, -1)) - P(0, 0, -1))), ((((P(0, 1, 1) + P(0, 0, 1)) + P(0, 1, -1)) + P(0, 0, -1)) + small))), hx);
        const R a3 = qdiv<R>((qdiv<R>(((q * (((V(1, 0, 0) + V(0, 0, 0)) + V(1, -1, 0)) + V(0, -1, 0))) *
                       (((P(1, 1, 0) + P(0, 1, 0)) - P(1, -1, 0)) - P(0, -1, 0))), ((((P(1, 1, 0) + P(0, 1, 0)) + P(1, -1, 0)) + P(0, -1, 0)) + small))), hy);
        const R b3 = qdiv<R>((qdiv<R>(((q * (((U(1, 0, 0) + U(0, 0, 0)) + U(1, 0, -1)) + U(0, 0, -1))) *
                       (((P(1, 0, 1) + P(0, 0, 1)) - P(1, 0, -1)) - P(0, 0, -1))), ((((P(1, 0, 1) + P(0, 0, 1)) + P(1, 0, -1)) + P(0, 0, -1)) + small))), hx);
        const R uc = U(0, 0, 0), vc = V(0, 0, 0), wc = W(0, 0, 0);
        R usv = qdiv<R>((qdiv<R>((((fabs(uc) * M.dx[i]) - ((dt * uc) * uc)) * (P(0, 0, 1) - P(0, 0, 0))), ((P(0, 0, 1) + P(0, 0, 0)) + small))), M.dxn[i]) - ((h * dt) * uc) * (a1 + b1);
        R vsv = qdiv<R>((qdiv<R>((((fabs(vc) * M.dy[j]) - ((dt * vc) * vc)) * (P(0, 1, 0) - P(0, 0, 0))), ((P(0, 1, 0) + P(0, 0, 0)) + small))), M.dyn[j]) - ((h * dt) * vc) * (a2 + b2);
        R wsv = qdiv<R>((qdiv<R>((((fabs(wc) * M.dz[k]) - ((dt * wc) * wc)) * (P(1, 0, 0) - P(0, 0, 0))), ((P(1, 0, 0) + P(0, 0, 0)) + small))), M.dzn[k]) - ((h * dt) * wc) * (a3 + b3);
        us[c] = usv - ((q * dt) * uc) * ((qdiv<R>((U(0, 0, 1) - U(0, 0, -1)), dxim)) + ud1);
        vs[c] = vsv - ((q * dt) * vc) * ((qdiv<R>((V(0, 1, 0) - V(0, -1, 0)), dyjm)) + ud2);
        ws[c] = wsv - ((q * dt) * wc) * ((qdiv<R>((W(1, 0, 0) - W(-1, 0, 0)), dzkm)) + ud3);
        return;
    }
    R u_div1 = qdiv<R>((ra[k] * (W(0, 0, 0) + W(0, 0, 1)) - ra[k - 1] * (W(-1, 0, 0) + W(-1, 0, 1))), (dzkm * g1)) +
               qdiv<R>((rn[k] * (((V(0, 0, 0) + V(0, 0, 1)) - V(0, -1, 0)) - V(0, -1, 1))), (dyjm * g1));
    R u_div2 = qdiv<R>((ra[k] * (W(0, 0, 0) + W(0, 1, 0)) - ra[k - 1] * (W(-1, 0, 0) + W(-1, 1, 0))), (dzkm * g2)) +
               qdiv<R>((rn[k] * (((U(0, 0, 0) + U(0, 1, 0)) - U(0, 0, -1)) - U(0, 1, -1))), (dxim * g2));
    R u_div3 = qdiv<R>((rn[k] * (U(0, 0, 0) - U(0, 0, -1)) + rn[k + 1] * (U(1, 0, 0) - U(1, 0, -1))), (dxim * g3)) +
               qdiv<R>((rn[k] * (V(0, 0, 0) - V(0, -1, 0)) + rn[k + 1] * (V(1, 0, 0) - V(1, -1, 0))), (dyjm * g3));
    R tz1 = qdiv<R>((qdiv<R>(((qdiv<R>((q * (((ra[k] * W(0, 0, 1) + ra[k] * W(0, 0, 0)) + ra[k - 1] * W(-1, 0, 1)) +
                     ra[k - 1] * W(-1, 0, 0))), g1)) *
              (((P(1, 0, 1) + P(1, 0, 0)) - P(-1, 0, 1)) - P(-1, 0, 0))), ((((P(1, 0, 1) + P(1, 0, 0)) + P(-1, 0, 1)) + P(-1, 0, 0)) + small))), (h * (M.dzn[k - 1] + M.dzn[k])));
    R ty1 = qdiv<R>((qdiv<R>(((qdiv<R>((q * (((rn[k] * V(0, 0, 1) + rn[k] * V(0, 0, 0)) + rn[k] * V(0, -1, 1)) +
                     rn[k] * V(0, -1, 0))), g1)) *
              (((P(0, 1, 1) + P(0, 1, 0)) - P(0, -1, 1)) - P(0, -1, 0))), ((((P(0, 1, 1) + P(0, 1, 0)) + P(0, -1, 1)) + P(0, -1, 0)) + small))), (h * (M.dyn[j - 1] + M.dyn[j])));
    R ujb1 = ((h * dt) * U(0, 0, 0)) * (tz1 + ty1);
    R tz2 = qdiv<R>((qdiv<R>(((qdiv<R>((q * (((ra[k] * W(0, 1, 0) + ra[k] * W(0, 0, 0)) + ra[k - 1] * W(-1, 1, 0)) +
                     ra[k - 1] * W(-1, 0, 0))), g2)) *
              (((P(1, 1, 0) + P(1, 0, 0)) - P(-1, 1, 0)) - P(-1, 0, 0))), ((((P(1, 1, 0) + P(1, 0, 0)) + P(-1, 1, 0)) + P(-1, 0, 0)) + small))), (h * (M.dzn[k - 1] + M.dzn[k])));
    R tx2 = qdiv<R>((qdiv<R>(((qdiv<R>((q * (((rn[k] * U(0, 1, 0) + rn[k] * U(0, 0, 0)) + rn[k] * U(0, 1, -1)) +
                     rn[k] * U(0, 0, -1))), g2)) *
              (((P(0, 1, 1) + P(0, 0, 1)) - P(0, 1, -1)) - P(0, 0, -1))), ((((P(0, 1, 1) + P(0, 0, 1)) + P(0, 1, -1)) + P(0, 0, -1)) + small))), (h * (M.dxn[i - 1] + M.dxn[i])));
    R ujb2 = ((h * dt) * V(0, 0, 0)) * (tz2 + tx2);
    R ty3 = qdiv<R>((qdiv<R>(((qdiv<R>((q * (((rn[k + 1] * V(1, 0, 0) + rn[k] * V(0, 0, 0)) + rn[k + 1] * V(1, -1, 0)) +
                     rn[k] * V(0, -1, 0))), g3)) *
              (((P(1, 1, 0) + P(0, 1, 0)) - P(1, -1, 0)) - P(0, -1, 0))), ((((P(1, 1, 0) + P(0, 1, 0)) + P(1, -1, 0)) + P(0, -1, 0)) + small))), (h * (M.dyn[j - 1] + M.dyn[j])));
    R tx3 = qdiv<R>((qdiv<R>(((qdiv<R>((q * (((rn[k + 1] * U(1, 0, 0) + rn[k] * U(0, 0, 0)) + rn[k + 1] * U(1, 0, -1)) +
                     rn[k] * U(0, 0, -1))), g3)) *
              (((P(1, 0, 1) + P(0, 0, 1)) - P(1, 0, -1)) - P(0, 0, -1))), ((((P(1, 0, 1) + P(0, 0, 1)) + P(1, 0, -1)) + P(0, 0, -1)) + small))), (h * (M.dxn[i - 1] + M.dxn[i])));
    R ujb3 = ((h * dt) * W(0, 0, 0)) * (ty3 + tx3);
    const R uc = U(0, 0, 0), vc = V(0, 0, 0), wc = W(0, 0, 0);
    R usv = qdiv<R>((qdiv<R>((((fabs(uc) * M.dx[i]) - ((dt * uc) * uc)) * (P(0, 0, 1) - P(0, 0, 0))), ((P(0, 0, 1) + P(0, 0, 0)) + small))), M.dxn[i]) -
            ujb1;
    R vsv = qdiv<R>((qdiv<R>((((fabs(vc) * M.dy[j]) - ((dt * vc) * vc)) * (P(0, 1, 0) - P(0, 0, 0))), ((P(0, 1, 0) + P(0, 0, 0)) + small))), M.dyn[j]) -
            ujb2;
    R wsv = qdiv<R>((qdiv<R>((((fabs(wc) * M.dz[k]) - (((dt * (qdiv<R>(ra[k], g3))) * wc) * wc)) *
              (P(1, 0, 0) - P(0, 0, 0))), ((P(1, 0, 0) + P(0, 0, 0)) + small))), M.dzn[k]) -
            ujb3;
    usv = usv - ((q * dt) * uc) * ((qdiv<R>((U(0, 0, 1) - U(0, 0, -1)), dxim)) + u_div1);
    vsv = vsv - ((q * dt) * vc) * ((qdiv<R>((V(0, 1, 0) - V(0, -1, 0)), dyjm)) + u_div2);
    wsv = wsv - ((q * dt) * wc) *
                    ((qdiv<R>((ra[k + 1] * W(1, 0, 0) - ra[k - 1] * W(-1, 0, 0)), (dzkm * g3))) + u_div3);
    us[c] = usv;
    vs[c] = vsv;
    ws[c] = wsv;
#undef U
#undef V
#undef W
#undef P
}

// ---------------------------------------------------------------------------
// K4: beta_up / beta_down (advection_3d.F90:549-616).  psi_store is psi-m,
// recomputed from the untouched psi.  The limiter factors of the x, y and z
// faces are up[0..2] / dn[0..2]: one shared pair for the linked copy, three
// pairs for the per-direction limiter of sfvt/advection_3d.F90.
// ---------------------------------------------------------------------------
template <typename R> struct Betas {
    R *up[3], *dn[3];
};

// sfvt/advection_3d.F90:555-636: bounds from the three cells along each
// axis; one pair of unweighted denominators (grid spacings dx(i-1), dy(j-1),
// dz(k-1), small inside the dt product) shared by the three directions
template <typename R>
__device__ __forceinline__ void beta_directional(const double *psi, double m, const R *po,
                                                 const R *ut, const R *vt, const R *wt,
                                                 Betas<R> b, long long c, int k, int j, int i,
                                                 const Geo &g, const M1<R> &M, R dt) {
    const long long st[3] = {g.si, g.sj, 1};
    const R pc = po[c];
    R mx[3], mn[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long s = st[d];
        const R vals[6] = {(R)(psi[c - s] - m), (R)(psi[c] - m), (R)(psi[c + s] - m),
                           po[c - s],           pc,              po[c + s]};
        mx[d] = vals[0];
        mn[d] = vals[0];
#pragma unroll
        for (int q = 1; q < 6; ++q) {
            if (vals[q] > mx[d]) mx[d] = vals[q];
            if (vals[q] < mn[d]) mn[d] = vals[q];
        }
    }
    const long long SI = g.si, SJ = g.sj;
    const R small = Small<R>::v;
    const R d1 = dt * (((qdiv<R>((pos0(ut[c - SI]) * po[c - SI] - neg0(ut[c]) * po[c + SI]), M.dx[i - 1]) +
                         qdiv<R>((pos0(vt[c - SJ]) * po[c - SJ] - neg0(vt[c]) * po[c + SJ]), M.dy[j - 1])) +
                        qdiv<R>((pos0(wt[c - 1]) * po[c - 1] - neg0(wt[c]) * po[c + 1]), M.dz[k - 1])) +
                       small);
    const R d2 = dt * (((qdiv<R>((pos0(ut[c]) * pc - neg0(ut[c - SI]) * pc), M.dx[i - 1]) +
                         qdiv<R>((pos0(vt[c]) * pc - neg0(vt[c - SJ]) * pc), M.dy[j - 1])) +
                        qdiv<R>((pos0(wt[c]) * pc - neg0(wt[c - 1]) * pc), M.dz[k - 1])) +
                       small);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        b.up[d][c] = qdiv<R>((mx[d] - pc), d1);
        b.dn[d][c] = qdiv<R>((pc - mn[d]), d2);
    }
}

template <typename R, bool PD>
__global__ __launch_bounds__(256) void k_beta(const double *psi, const double *mptr,
                                              const R *po, const R *ut, const R *vt,
                                              const R *wt, Betas<R> b, const int *flag,
                                              Geo g, MetPack mp, R dt) {
    if (flag && *flag) return;
    const M1<R> M(mp);
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long SI = g.si, SJ = g.sj;
    const long long c = pidx(g, k, j, i);
    const double m = *mptr;
    if constexpr (PD) {
        beta_directional<R>(psi, m, po, ut, vt, wt, b, c, k, j, i, g, M, dt);
        return;
    }
    R *bup = b.up[0], *bdn = b.dn[0];
    const R small = Small<R>::v;
    const R vals[14] = {(R)(psi[c] - m),      (R)(psi[c - SI] - m), (R)(psi[c + SI] - m),
                        (R)(psi[c - SJ] - m), (R)(psi[c + SJ] - m), (R)(psi[c - 1] - m),
                        (R)(psi[c + 1] - m),  po[c],                po[c - SI],
                        po[c + SI],           po[c - SJ],           po[c + SJ],
                        po[c - 1],            po[c + 1]};
    R mx = vals[0], mn = vals[0];
#pragma unroll
    for (int q = 1; q < 14; ++q) {
        if (vals[q] > mx) mx = vals[q];
        if (vals[q] < mn) mn = vals[q];
    }
    const R *ra = M.ra, *rn = M.rn;
    const R pc = po[c];
    R d1 = (dt * ((qdiv<R>(((pos0(ut[c - SI]) * rn[k]) * po[c - SI] - (neg0(ut[c]) * rn[k]) * po[c + SI]), M.dxn[i]) +
                   qdiv<R>(((pos0(vt[c - SJ]) * rn[k]) * po[c - SJ] - (neg0(vt[c]) * rn[k]) * po[c + SJ]), M.dyn[j])) +
                  qdiv<R>(((pos0(wt[c - 1]) * ra[k - 1]) * po[c - 1] - (neg0(wt[c]) * ra[k]) * po[c + 1]), M.dzn[k]))) +
           small;
    R d2 = dt * (((qdiv<R>(((pos0(ut[c]) * rn[k]) * pc - (neg0(ut[c - SI]) * rn[k]) * pc), M.dxn[i]) +
                   qdiv<R>(((pos0(vt[c]) * rn[k]) * pc - (neg0(vt[c - SJ]) * rn[k]) * pc), M.dyn[j])) +
                  qdiv<R>(((pos0(wt[c]) * ra[k]) * pc - (neg0(wt[c - 1]) * ra[k - 1]) * pc), M.dzn[k])) +
                 small);
    bup[c] = qdiv<R>((rn[k] * (mx - pc)), d1);
    bdn[c] = qdiv<R>((rn[k] * (pc - mn)), d2);
}

// limited velocity of eq 18 at the face owned by cell c (stride s to the
// neighbour across that face)
template <typename R>
__device__ __forceinline__ R limited(const R *us, const R *bup, const R *bdn, long long c,
                                     long long s) {
    const R x = us[c];
    return min3<R>(R(1), bdn[c], bup[c + s]) * pos0(x) + min3<R>(R(1), bup[c], bdn[c + s]) * neg0(x);
}

// eq 18 materialised (only for kord > 2, where the next antidiffusive pass
// needs the limited velocities as its input)
template <typename R>
__global__ __launch_bounds__(256) void k_limit(const R *us, const R *vs, const R *ws,
                                               Betas<R> b, R *ul, R *vl, R *wl,
                                               const int *flag, Geo g) {
    if (flag && *flag) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    ul[c] = limited<R>(us, b.up[0], b.dn[0], c, g.si);
    vl[c] = limited<R>(vs, b.up[1], b.dn[1], c, g.sj);
    wl[c] = limited<R>(ws, b.up[2], b.dn[2], c, 1);
}

// ---------------------------------------------------------------------------
// K5: last donor pass with the limited velocities computed at the six faces.
// The lower faces of cells i=1 / j=1 take the periodic copy of the face at
// ip / jp (exchange_full of the limited velocities, :669-674); the face
// below k=1 is the lbc=0 plane.
// ---------------------------------------------------------------------------
template <typename R, bool TO_PSI>
__global__ __launch_bounds__(256) void k_upwind_limited(
    const R *src, const R *us, const R *vs, const R *ws, Betas<R> b, R *out,
    double *psi_out, const double *mptr, const int *flag, Geo g, MetPack mp, R dt,
    int neumann) {
    if (flag && *flag) return;
    const M1<R> M(mp);
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const long long cw = pidx(g, k, j, i == 1 ? g.ip : i - 1);
    const long long cs = pidx(g, k, j == 1 ? g.jp : j - 1, i);
    const R uc = limited<R>(us, b.up[0], b.dn[0], c, g.si);
    const R ul = limited<R>(us, b.up[0], b.dn[0], cw, g.si);
    const R vc = limited<R>(vs, b.up[1], b.dn[1], c, g.sj);
    const R vl = limited<R>(vs, b.up[1], b.dn[1], cs, g.sj);
    const R wc = limited<R>(ws, b.up[2], b.dn[2], c, 1);
    const R wl = k == 1 ? R(0) : limited<R>(ws, b.up[2], b.dn[2], c - 1, 1);
    R res = donor<R>(dt, uc, ul, vc, vl, wc, wl, src[c], src[c + g.si], src[c - g.si],
                     src[c + g.sj], src[c - g.sj], src[c + 1], src[c - 1], M.dxn[i], M.dyn[j],
                     M.dzn[k], M.ra[k], M.ra[k - 1], M.rn[k], k, g.kp, neumann);
    if (TO_PSI)
        psi_out[c] = (double)res + *mptr;
    else
        out[c] = res;
}

// copy the interior of a scratch field into psi (standalone upwind API)
__global__ void k_copy_interior(const double *src, double *dst, Geo g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    dst[c] = src[c];
}

// ---------------------------------------------------------------------------
// adv_ref_state (advection_3d.F90:1148-1218): every (j,i) column including
// the lateral halo, k = 1..kp
// ---------------------------------------------------------------------------
__global__ void k_adv_ref_state(const double *w, double *psi, const double *ref, Geo g,
                                Metrics M, double dt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k > g.kp || j > g.jp + 1) return;
    const long long c = pidx(g, k, j, i);
    psi[c] = psi[c] - (dt * ((((0.5 / M.dz[k - 1]) * w[c - 1]) * M.rhoa[k - 1]) *
                                 (ref[k] - ref[k - 1]) +
                             (((0.5 / M.dz[k]) * w[c]) * M.rhoa[k]) * (ref[k + 1] - ref[k]))) /
                          M.rhoan[k];
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
template <typename R>
static int get_buf(dcm_ctx *c, int slot, R **out);
template <>
int get_buf<double>(dcm_ctx *c, int slot, double **out) { return work_field64(c, slot, out); }
template <>
int get_buf<float>(dcm_ctx *c, int slot, float **out) { return work_field32(c, slot, out); }

// Work-slot map (separate pools for FP64 and FP32):
//   0-2 antidiffusive U*,V*,W*   3-5 limited (kord>2)   6-7 beta up/down
//   8.. psi1 per field, then psi2 per field (kord>2)
//   80-83 the y and z beta pairs of the per-direction limiter
static constexpr int kSlotPsi = 8;
static constexpr int kSlotBetaYZ = 80;

// psi(interior) = psi1 + m (kord = 1)
template <typename R>
__global__ void k_store_shifted(const R *src, double *psi, const double *mptr, const int *flag,
                                Geo g) {
    if (*flag) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    psi[c] = (double)src[c] + *mptr;
}

// Shared body of mpdata_3d and mpdata_vec_3d.  u,v,w are the advecting
// velocities in precision R (for R=float, FP32 shadows).  psi: nf fields of
// ntot FP64 each.  lbc/ubc are host values (unshifted); the shifts m[n] and
// the exit flag stay on the device.
template <typename R>
int mpdata_core(dcm_ctx *c, R dt, const R *u, const R *v, const R *w, double *psi, int nf,
                const double *lbc, const double *ubc, int kord, int monotone, int neumann,
                int interior_min) {
    const Geo &g = c->geo;
    const Interior L = interior_launch(g);
    MetPack mp{c->met, c->met32};
    double *dm = c->dscal + 16;         // shifts of up to 32 fields
    int *flag = (int *)(c->dscal + 8);  // zero-sum exit flag
    if (nf < 1 || nf > 32) {
        set_error("mpdata: a category must hold 1..32 fields");
        return DCM_BAD_ARG;
    }
    // sfvt/advection_3d.F90: per-direction limiter in both routines,
    // unweighted antidiffusive velocities in the vector one (interior_min)
    const bool pd = c->mpdata_variant == 1;
    const bool weighted = !(pd && interior_min);
    const unsigned nbmin = 1024;
    const unsigned nbint = L.grid.x * L.grid.y * L.grid.z;
    DCM_CHECK(ensure_partials(c, (size_t)nbint + nbmin + 16));
    // K1: shifts (mpdata_3d :275-280, mpdata_vec_3d :1306-1312)
    for (int n = 0; n < nf; ++n) {
        const double *f = psi + (size_t)n * g.ntot;
        int tok = prof_begin(c, "mpdata_min");
        hipLaunchKernelGGL(k_min, dim3(nbmin), dim3(kBlock), 0, c->stream, f, g, interior_min,
                           c->partials);
        DCM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_min_finish, dim3(1), dim3(1024), 0, c->stream, c->partials,
                           (int)nbmin, lbc[n], ubc[n], dm + n, flag);
        DCM_LAUNCH_CHECK();
        prof_end(c, tok);
    }
    // K2: first donor pass of every field with (u,v,w); zero-sum test on field 1
    std::vector<R *> s1(nf), s2(nf);
    for (int n = 0; n < nf; ++n) DCM_CHECK(get_buf<R>(c, kSlotPsi + n, &s1[n]));
    for (int n = 0; n < nf; ++n) {
        int tok = prof_begin(c, "upwind_first");
        hipLaunchKernelGGL(k_upwind_first<R>, L.grid, L.block, 0, c->stream,
                           psi + (size_t)n * g.ntot, dm + n, u, v, w, s1[n], g, mp, dt, neumann,
                           n == 0 ? 1 : 0, c->partials);
        DCM_LAUNCH_CHECK();
        prof_end(c, tok);
        if (n == 0) {
            hipLaunchKernelGGL(k_exit_flag, dim3(1), dim3(1024), 0, c->stream, c->partials,
                               (int)nbint, flag);
            DCM_LAUNCH_CHECK();
        }
    }
    if (kord <= 1) {
        for (int n = 0; n < nf; ++n) {
            hipLaunchKernelGGL(k_store_shifted<R>, L.grid, L.block, 0, c->stream, s1[n],
                               psi + (size_t)n * g.ntot, dm + n, flag, g);
            DCM_LAUNCH_CHECK();
        }
    } else {
        for (int n = 0; n < nf; ++n)
            DCM_CHECK(fill<R>(c, s1[n], FILL_FULL, lbc[n], ubc[n], dm + n, 1, flag));
        if (kord > 2)
            for (int n = 0; n < nf; ++n) DCM_CHECK(get_buf<R>(c, kSlotPsi + nf + n, &s2[n]));
        // velocity triples: X = (u,v,w); A, B scratch
        R *A[3], *B[3];
        Betas<R> bt{};
        for (int d = 0; d < 3; ++d) DCM_CHECK(get_buf<R>(c, d, &A[d]));
        if (kord > 2)
            for (int d = 0; d < 3; ++d) DCM_CHECK(get_buf<R>(c, 3 + d, &B[d]));
        if (monotone) {
            DCM_CHECK(get_buf<R>(c, 6, &bt.up[0]));
            DCM_CHECK(get_buf<R>(c, 7, &bt.dn[0]));
            for (int d = 1; d < 3; ++d) {
                if (pd) {
                    DCM_CHECK(get_buf<R>(c, kSlotBetaYZ + 2 * (d - 1), &bt.up[d]));
                    DCM_CHECK(get_buf<R>(c, kSlotBetaYZ + 2 * (d - 1) + 1, &bt.dn[d]));
                } else {
                    bt.up[d] = bt.up[0];
                    bt.dn[d] = bt.dn[0];
                }
            }
        }
        const R *in[3] = {u, v, w};
        R **anti = A, **other = B;
        std::vector<R *> cur = s1, nxt = s2;
        for (int it = 2; it <= kord; ++it) {
            const bool last = it == kord;
            // antidiffusive velocities from the current advecting velocity
            int tok = prof_begin(c, "antidiff");
            if (weighted)
                hipLaunchKernelGGL((k_antidiff<R, true>), L.grid, L.block, 0, c->stream, in[0],
                                   in[1], in[2], cur[0], anti[0], anti[1], anti[2], flag, g, mp,
                                   dt);
            else
                hipLaunchKernelGGL((k_antidiff<R, false>), L.grid, L.block, 0, c->stream, in[0],
                                   in[1], in[2], cur[0], anti[0], anti[1], anti[2], flag, g, mp,
                                   dt);
            DCM_LAUNCH_CHECK();
            prof_end(c, tok);
            DCM_CHECK(fill_set<R>(c, FillSet<R>{{anti[0], anti[1], anti[2]}}, 3, FILL_FULL, 0.0,
                                  0.0, flag));
            if (monotone) {
                tok = prof_begin(c, "beta");
                if (pd)
                    hipLaunchKernelGGL((k_beta<R, true>), L.grid, L.block, 0, c->stream, psi, dm,
                                       cur[0], anti[0], anti[1], anti[2], bt, flag, g, mp, dt);
                else
                    hipLaunchKernelGGL((k_beta<R, false>), L.grid, L.block, 0, c->stream, psi,
                                       dm, cur[0], anti[0], anti[1], anti[2], bt, flag, g, mp, dt);
                DCM_LAUNCH_CHECK();
                prof_end(c, tok);
                const FillSet<R> bs{{bt.up[0], bt.dn[0], bt.up[1], bt.dn[1], bt.up[2], bt.dn[2]}};
                DCM_CHECK(fill_set<R>(c, bs, pd ? 6 : 2, FILL_ALONG_DIM, 0.0, 0.0, flag));
            }
            if (last) {
                for (int n = 0; n < nf; ++n) {
                    tok = prof_begin(c, monotone ? "upwind_limited" : "upwind_last");
                    if (monotone)
                        hipLaunchKernelGGL((k_upwind_limited<R, true>), L.grid, L.block, 0,
                                           c->stream, cur[n], anti[0], anti[1], anti[2], bt,
                                           (R *)nullptr, psi + (size_t)n * g.ntot, dm + n,
                                           flag, g, mp, dt, neumann);
                    else
                        hipLaunchKernelGGL((k_upwind<R, true>), L.grid, L.block, 0, c->stream,
                                           cur[n], anti[0], anti[1], anti[2], (R *)nullptr,
                                           psi + (size_t)n * g.ntot, dm + n, flag, g, mp, dt,
                                           neumann);
                    DCM_LAUNCH_CHECK();
                    prof_end(c, tok);
                }
                break;
            }
            // an intermediate pass (kord > 2)
            R **adv = anti;  // advecting velocity of this pass
            if (monotone) {
                // eq 18 into the other triple, which the next pass reads
                hipLaunchKernelGGL(k_limit<R>, L.grid, L.block, 0, c->stream, anti[0], anti[1],
                                   anti[2], bt, other[0], other[1], other[2], flag, g);
                DCM_LAUNCH_CHECK();
                DCM_CHECK(fill_set<R>(c, FillSet<R>{{other[0], other[1], other[2]}}, 3, FILL_FULL,
                                      0.0, 0.0, flag));
                adv = other;
            }
            for (int n = 0; n < nf; ++n) {
                hipLaunchKernelGGL((k_upwind<R, false>), L.grid, L.block, 0, c->stream, cur[n],
                                   adv[0], adv[1], adv[2], nxt[n], (double *)nullptr, dm + n,
                                   flag, g, mp, dt, neumann);
                DCM_LAUNCH_CHECK();
                DCM_CHECK(fill<R>(c, nxt[n], FILL_FULL, lbc[n], ubc[n], dm + n, 1, flag));
            }
            for (int d = 0; d < 3; ++d) in[d] = adv[d];
            if (!monotone) std::swap(anti, other);  // never overwrite the input
            std::swap(cur, nxt);
        }
    }
    // halo of the result: exchange_full(psi, lbc-m, ubc-m) then +m, i.e. the
    // z planes hold (bc-m)+m and every other halo cell copies the interior
    for (int n = 0; n < nf; ++n) {
        double *f = psi + (size_t)n * g.ntot;
        DCM_CHECK(fill<double>(c, f, FILL_FULL, lbc[n], ubc[n], dm + n, 2, flag));
    }
    // zero-sum exit: psi = (psi-m)+m over the whole array, nothing else
    for (int n = 0; n < nf; ++n) {
        hipLaunchKernelGGL(k_exit_unshift, dim3(1024), dim3(kBlock), 0, c->stream,
                           psi + (size_t)n * g.ntot, g.ntot, dm + n, flag);
        DCM_LAUNCH_CHECK();
    }
    return DCM_OK;
}

template int mpdata_core<double>(dcm_ctx *, double, const double *, const double *,
                                 const double *, double *, int, const double *, const double *,
                                 int, int, int, int);
template int mpdata_core<float>(dcm_ctx *, float, const float *, const float *, const float *,
                                double *, int, const double *, const double *, int, int, int,
                                int);

// adv_ref_state over the columns j = 0..jp+1, i = 0..ip+1 (ref on the device)
int adv_ref_state_dev(dcm_ctx *c, double dt, const double *w, double *psi, const double *ref) {
    const Geo &g = c->geo;
    Interior L = interior_launch(g);
    L.grid.y = (g.nj + L.block.y - 1) / L.block.y;
    L.grid.z = g.ni;
    int tok = prof_begin(c, "adv_ref_state");
    hipLaunchKernelGGL(k_adv_ref_state, L.grid, L.block, 0, c->stream, w, psi, ref, g, c->met, dt);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    return DCM_OK;
}

// Convert u,v,w (whole padded arrays) to the FP32 shadows used by the mixed path.
int velocity_shadow(dcm_ctx *c, const double *u, const double *v, const double *w, float *uf,
                    float *vf, float *wf) {
    hipLaunchKernelGGL(k_to_float3, dim3(2048), dim3(kBlock), 0, c->stream, u, v, w, uf, vf, wf,
                       c->geo.ntot);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

// ---------------------------------------------------------------------------
// Vertical-only advection, FP64: first_order_upstream_vert_3d
// (pamm/sfvt/advection_3d.F90:737-791), mpdata_vert_3d (:820-1126) and
// mpdata_vec_vert_3d (:1779-2087), the sedimentation transport of PAMM.
// As in mpdata_core, the input field stays untouched until the last kernel:
// the first pass reads psi - m, later passes ping-pong between scratch fields
// whose z planes hold bc - m, and the result is written back as value + m.
// ---------------------------------------------------------------------------
// one donor pass along z (:754-789) from src (minus *shift when given) into
// dst; neumann 2: the floor face of level 1 takes w(1) (:762-766); neumann
// 1: zero-gradient fluxes at floor and lid (:780-787).  With do_sum, block
// partial sums of the shifted source (the zero-sum test, :868-870).
__global__ __launch_bounds__(256) void k_vupwind(const double *src, const double *shift,
                                                 const double *w, double *dst, Geo g,
                                                 Metrics mt, double dt, int neumann,
                                                 const int *flag, int do_sum,
                                                 double *partials) {
    if (flag && *flag) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    const bool act = k <= g.kp && j <= g.jp;
    const double m = shift ? *shift : 0.0;
    double s[1] = {0.0};
    if (act) {
        const long long c = pidx(g, k, j, i);
        auto P = [&](long long q) { return shift ? src[q] - m : src[q]; };
        const double pc = P(c), pu = P(c + 1), pd = P(c - 1);
        const double wk = w[c];
        const double wm = (k == 1 && neumann == 2) ? wk : w[c - 1];
        const double den = (2.0 * mt.dzn[k]) * mt.rhoan[k];
        double fr = ((((wk + fabs(wk)) * mt.rhoa[k]) * pc + ((wk - fabs(wk)) * mt.rhoa[k]) * pu) *
                     dt) / den;
        double fl = ((((wm + fabs(wm)) * mt.rhoa[k - 1]) * pd +
                      ((wm - fabs(wm)) * mt.rhoa[k - 1]) * pc) * dt) / den;
        if (neumann == 1) {
            if (k == 1) fl = fr;
            if (k == g.kp) fr = fl;
        }
        dst[c] = pc - (fr - fl);
        s[0] = pc;
    }
    if (do_sum) block_sum_store<1>(s, partials, 0);
}

// antidiffusive vertical velocity (:942-958; vector variant :1905-1921);
// plain: the unweighted form of sfvt/advection_3d.F90's mpdata_vec_vert_3d
// (:1897-1909)
__global__ __launch_bounds__(256) void k_vantidiff(const double *wt, const double *po,
                                                   double *ws, Geo g, Metrics mt, double dt,
                                                   const int *flag, int plain) {
    if (*flag) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const double *dz = mt.dz, *ra = mt.rhoa, *rn = mt.rhoan;
    const double gb3 = 0.5 * (rn[k] + rn[k + 1]);
    const double u_div3 = 0.0, u_j_bar3 = 0.0;
    const double wc = wt[c], p0 = po[c], p1 = po[c + 1];
    if (plain) {
        const double v = ((((fabs(wc) * dz[k] - (dt * wc) * wc) * (p1 - p0)) /
                           ((p1 + p0) + Small<double>::v)) / mt.dzn[k]) - u_j_bar3;
        ws[c] = v - ((0.25 * dt) * wc) * (((wt[c + 1] - wt[c - 1]) / dz[k - 1]) + u_div3);
        return;
    }
    double v = ((((fabs(wc) * dz[k] - ((dt * (ra[k] / gb3)) * wc) * wc) * (p1 - p0)) /
                 ((p1 + p0) + Small<double>::v)) / mt.dzn[k]) - u_j_bar3;
    v = v - ((0.25 * dt) * wc) *
                (((ra[k + 1] * wt[c + 1] - ra[k - 1] * wt[c - 1]) / (dz[k - 1] * gb3)) + u_div3);
    ws[c] = v;
}

// limiter bounds (:1000-1046; vector variant :1964-2010): psi_store is the
// input field minus its shift, psi_old the current pass.  pd: the unweighted
// denominators of sfvt/advection_3d.F90 (:1051-1062, :1975-1985)
__global__ __launch_bounds__(256) void k_vbeta(const double *psi, const double *shift,
                                               const double *po, const double *wt, double *bup,
                                               double *bdn, Geo g, Metrics mt, double dt,
                                               const int *flag, int pd) {
    if (*flag) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const double m = *shift;
    const double vals[6] = {psi[c - 1] - m, psi[c] - m, psi[c + 1] - m,
                            po[c - 1],      po[c],      po[c + 1]};
    double mx = vals[0], mn = vals[0];
#pragma unroll
    for (int q = 1; q < 6; ++q) {
        if (vals[q] > mx) mx = vals[q];
        if (vals[q] < mn) mn = vals[q];
    }
    const double *ra = mt.rhoa, *rn = mt.rhoan, *dzn = mt.dzn;
    const double pc = po[c], wm = wt[c - 1], wc = wt[c];
    if (pd) {
        const double *dz = mt.dz;
        const double d1 =
            dt * ((pos0(wm) * po[c - 1] - neg0(wc) * po[c + 1]) / dz[k - 1] + Small<double>::v);
        const double d2 = dt * ((pos0(wc) * pc - neg0(wm) * pc) / dz[k - 1] + Small<double>::v);
        bup[c] = (mx - pc) / d1;
        bdn[c] = (pc - mn) / d2;
        return;
    }
    const double d1 = (dt * ((pos0(wm) * ra[k - 1]) * po[c - 1] - (neg0(wc) * ra[k]) * po[c + 1])) /
                          dzn[k] + Small<double>::v;
    const double d2 = (dt * ((pos0(wc) * ra[k]) * pc - (neg0(wm) * ra[k - 1]) * pc)) / dzn[k] +
                      Small<double>::v;
    bup[c] = (rn[k] * (mx - pc)) / d1;
    bdn[c] = (rn[k] * (pc - mn)) / d2;
}

// limited vertical velocity, eq 18 (:1071-1076; vector variant :2028-2033)
__global__ __launch_bounds__(256) void k_vlimit(const double *wt, const double *bup,
                                                const double *bdn, double *ws, Geo g,
                                                const int *flag) {
    if (*flag) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const double wc = wt[c];
    ws[c] = min3(1.0, bdn[c], bup[c + 1]) * pos0(wc) + min3(1.0, bup[c], bdn[c + 1]) * neg0(wc);
}

// result of a vertical MPDATA call for one field: the interior from the last
// pass plus m, the z planes (bc - m) + m (exchange_along_z after every
// pass), every other cell (psi - m) + m; after the zero-sum exit
// (psi - m) + m everywhere (:870-875, :1123-1125)
__global__ void k_vfinish(double *psi, const double *res, const double *shift, const int *flag,
                          Geo g, double lbc, double ubc, int passes) {
    const double m = *shift;
    const bool exit = *flag != 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < g.ntot;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nk);
        const long long r = t / g.nk;
        const int j = (int)(r % g.nj), i = (int)(r / g.nj);
        double v;
        if (exit || passes == 0) {
            v = (psi[t] - m) + m;
        } else if (k == 0 || k == g.kp + 1) {
            v = ((k == 0 ? lbc : ubc) - m) + m;
        } else if (j >= 1 && j <= g.jp && i >= 1 && i <= g.ip) {
            v = res[t] + m;
        } else {
            v = (psi[t] - m) + m;
        }
        psi[t] = v;
    }
}

// psi(k,:,:) += sign * prof(k) over the whole padded array (mpdata_3d_add,
// :172-174 and :183-185)
__global__ void k_add_profile(double *psi, const double *prof, double sign, Geo g) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < g.ntot;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nk);
        psi[t] = sign > 0.0 ? psi[t] + prof[k] : psi[t] - prof[k];
    }
}

// Shared driver of mpdata_vert_3d and mpdata_vec_vert_3d: nf fields of psi
// (field 1 drives the velocities and the limiter), lbc/ubc of each field
// (host values before the shift).  The shifts and the exit flag are left in
// c->dscal like mpdata_core's.
int mpdata_vert_core(dcm_ctx *c, double dt, const double *w, double *psi, int nf,
                     const double *lbc, const double *ubc, int kord, int monotone, int neumann,
                     int interior_min) {
    const Geo &g = c->geo;
    const Interior L = interior_launch(g);
    const Metrics mt = c->met;
    double *dm = c->dscal + 16;
    int *flag = (int *)(c->dscal + 8);
    const bool pd = c->mpdata_variant == 1;  // sfvt/advection_3d.F90's arithmetic
    if (nf < 1 || nf > 32) {
        set_error("vertical mpdata: a category must hold 1..32 fields");
        return DCM_BAD_ARG;
    }
    if (kord < 0) {
        set_error("vertical mpdata: kord must be >= 0");
        return DCM_BAD_ARG;
    }
    const unsigned nbmin = 1024;
    const unsigned nbint = L.grid.x * L.grid.y * L.grid.z;
    DCM_CHECK(ensure_partials(c, (size_t)nbint + nbmin + 16));
    // shifts: minimum of the whole array (scalar form, :861) or of the
    // interior (vector form, :1822), with lbc and ubc
    for (int n = 0; n < nf; ++n) {
        hipLaunchKernelGGL(k_min, dim3(nbmin), dim3(kBlock), 0, c->stream,
                           psi + (size_t)n * g.ntot, g, interior_min, c->partials);
        hipLaunchKernelGGL(k_min_finish, dim3(1), dim3(1024), 0, c->stream, c->partials,
                           (int)nbmin, lbc[n], ubc[n], dm + n, flag);
        DCM_LAUNCH_CHECK();
    }
    // velocity and limiter scratch, two ping-pong fields per advected field
    double *W1, *W2, *bup, *bdn;
    DCM_CHECK(work_field64(c, 0, &W1));
    DCM_CHECK(work_field64(c, 1, &W2));
    DCM_CHECK(work_field64(c, 6, &bup));
    DCM_CHECK(work_field64(c, 7, &bdn));
    std::vector<double *> cur(nf), nxt(nf);
    for (int n = 0; n < nf; ++n) {
        DCM_CHECK(work_field64(c, kSlotPsi + n, &cur[n]));
        DCM_CHECK(work_field64(c, kSlotPsi + nf + n, &nxt[n]));
    }
    const double *wt = w;
    double *wt_sav = W1;
    int it2 = 0;
    for (int it = 1; it <= kord; ++it) {
        if (it > 1) {
            it2 += 1;
            hipLaunchKernelGGL(k_vantidiff, L.grid, L.block, 0, c->stream, wt, cur[0], wt_sav, g,
                               mt, dt, flag, (pd && interior_min) ? 1 : 0);
            DCM_LAUNCH_CHECK();
            if (it2 % 2 == 1) {
                wt = W1;
                wt_sav = W2;
            } else {
                wt = W2;
                wt_sav = W1;
            }
            DCM_CHECK(fill<double>(c, (double *)wt, FILL_ALONG_Z, 0.0, 0.0, nullptr, 0, flag));
        }
        if (it > 1 && monotone) {
            it2 += 1;
            hipLaunchKernelGGL(k_vbeta, L.grid, L.block, 0, c->stream, psi, dm, cur[0], wt, bup,
                               bdn, g, mt, dt, flag, pd ? 1 : 0);
            DCM_LAUNCH_CHECK();
            DCM_CHECK(fill<double>(c, bup, FILL_ALONG_Z, 0.0, 0.0, nullptr, 0, flag));
            DCM_CHECK(fill<double>(c, bdn, FILL_ALONG_Z, 0.0, 0.0, nullptr, 0, flag));
            hipLaunchKernelGGL(k_vlimit, L.grid, L.block, 0, c->stream, wt, bup, bdn, wt_sav, g,
                               flag);
            DCM_LAUNCH_CHECK();
            wt = W2;
            wt_sav = W1;
            DCM_CHECK(fill<double>(c, (double *)wt, FILL_ALONG_Z, 0.0, 0.0, nullptr, 0, flag));
        }
        for (int n = 0; n < nf; ++n) {
            const bool first = it == 1;
            // the first pass reads the input minus its shift and, for field
            // 1, sums it for the zero-sum test
            hipLaunchKernelGGL(k_vupwind, L.grid, L.block, 0, c->stream,
                               first ? (const double *)(psi + (size_t)n * g.ntot) : cur[n],
                               first ? (const double *)(dm + n) : nullptr, wt, nxt[n], g, mt, dt,
                               neumann, first ? nullptr : flag, (first && n == 0) ? 1 : 0,
                               c->partials);
            DCM_LAUNCH_CHECK();
            if (first && n == 0) {
                hipLaunchKernelGGL(k_exit_flag, dim3(1), dim3(1024), 0, c->stream, c->partials,
                                   (int)nbint, flag);
                DCM_LAUNCH_CHECK();
            }
            DCM_CHECK(fill<double>(c, nxt[n], FILL_ALONG_Z, lbc[n], ubc[n], dm + n, 1, flag));
            std::swap(cur[n], nxt[n]);
        }
    }
    if (kord < 1) {
        // no pass: only the zero-sum test
        hipLaunchKernelGGL(k_vupwind, L.grid, L.block, 0, c->stream, (const double *)psi,
                           (const double *)dm, w, nxt[0], g, mt, dt, neumann, nullptr, 1,
                           c->partials);
        hipLaunchKernelGGL(k_exit_flag, dim3(1), dim3(1024), 0, c->stream, c->partials,
                           (int)nbint, flag);
        DCM_LAUNCH_CHECK();
    }
    for (int n = 0; n < nf; ++n) {
        hipLaunchKernelGGL(k_vfinish, dim3(2048), dim3(kBlock), 0, c->stream,
                           psi + (size_t)n * g.ntot, (const double *)cur[n],
                           (const double *)(dm + n), (const int *)flag, g, lbc[n], ubc[n],
                           kord);
        DCM_LAUNCH_CHECK();
    }
    return DCM_OK;
}

}  // namespace dcm

using namespace dcm;

extern "C" {

int dcm_set_mpdata_variant(dcm_ctx *c, int variant) {
    if (!c || (variant != DCM_MPDATA_LINKED && variant != DCM_MPDATA_SFVT)) {
        set_error("dcm_set_mpdata_variant: the variant is DCM_MPDATA_LINKED or DCM_MPDATA_SFVT");
        return DCM_BAD_ARG;
    }
    c->mpdata_variant = variant;
    return DCM_OK;
}

int dcm_upwind3d(dcm_ctx *c, double dt, const double *u, const double *v, const double *w,
                 double *psi, int neumann) {
    if (!c || !u || !v || !w || !psi) return DCM_BAD_ARG;
    return upwind_dev(c, dt, u, v, w, psi, neumann);
}

}  // extern "C"

namespace dcm {

// first_order_upstream_3d in place on psi (FP64; the model's advection_scheme 0)
int upwind_dev(dcm_ctx *c, double dt, const double *u, const double *v, const double *w,
               double *psi, int neumann) {
    const Geo &g = c->geo;
    const Interior L = interior_launch(g);
    MetPack mp{c->met, c->met32};
    double *tmp;
    DCM_CHECK(work_field64(c, kSlotPsi, &tmp));
    // fluxes from the old field into scratch, then the in-place update
    // (advection_3d.F90:74-115)
    hipLaunchKernelGGL((k_upwind<double, false>), L.grid, L.block, 0, c->stream, psi, u, v, w,
                       tmp, (double *)nullptr, (const double *)nullptr, (const int *)nullptr, g,
                       mp, dt, neumann);
    DCM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_copy_interior, L.grid, L.block, 0, c->stream, tmp, psi, g);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

// psi += sign * prof(k) over the whole padded array (mpdata_3d_add's profile)
int add_profile_dev(dcm_ctx *c, double *psi, const double *prof_dev, double sign) {
    hipLaunchKernelGGL(k_add_profile, dim3(2048), dim3(kBlock), 0, c->stream, psi, prof_dev,
                       sign, c->geo);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

}  // namespace dcm

extern "C" {

static int mpdata_api(dcm_ctx *c, double dt, const double *u, const double *v, const double *w,
                      double *psi, int nq, double *lbc, double *ubc, int kord, int monotone,
                      int neumann, int *early_exit, int interior_min) {
    if (!c || !u || !v || !w || !psi || !lbc || !ubc || nq < 1 || nq > 32) return DCM_BAD_ARG;
    DCM_CHECK(mpdata_core<double>(c, dt, u, v, w, psi, nq, lbc, ubc, kord, monotone, neumann,
                                  interior_min));
    // lbc/ubc leave as (bc - m) + m, like the Fortran INTENT(INOUT) scalars
    DCM_HIP(hipMemcpyAsync(c->hscal, c->dscal, 48 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    for (int n = 0; n < nq; ++n) {
        const double m = c->hscal[16 + n];
        lbc[n] = (lbc[n] - m) + m;
        ubc[n] = (ubc[n] - m) + m;
    }
    if (early_exit) *early_exit = *(int *)(c->hscal + 8);
    return DCM_OK;
}

int dcm_mpdata3d(dcm_ctx *c, double dt, const double *u, const double *v, const double *w,
                 double *psi, double *lbc, double *ubc, int kord, int monotone, int neumann,
                 int *early_exit) {
    return mpdata_api(c, dt, u, v, w, psi, 1, lbc, ubc, kord, monotone, neumann, early_exit, 0);
}

int dcm_mpdata_vec3d(dcm_ctx *c, double dt, const double *u, const double *v, const double *w,
                     double *psi, int nq, double *lbc, double *ubc, int kord, int monotone,
                     int neumann, int *early_exit) {
    return mpdata_api(c, dt, u, v, w, psi, nq, lbc, ubc, kord, monotone, neumann, early_exit, 1);
}

int dcm_mpdata3d_add(dcm_ctx *c, double dt, const double *u, const double *v, const double *w,
                     double *psi, const double *psi_1d, double *lbc, double *ubc, int kord,
                     int monotone, int neumann, int *early_exit) {
    if (!c || !psi_1d) return DCM_BAD_ARG;
    const Geo &g = c->geo;
    const double *prof;
    DCM_CHECK(upload_1d(c, psi_1d, (size_t)g.kp + 2, 1, &prof));
    hipLaunchKernelGGL(k_add_profile, dim3(2048), dim3(kBlock), 0, c->stream, psi, prof, 1.0, g);
    DCM_LAUNCH_CHECK();
    DCM_CHECK(mpdata_api(c, dt, u, v, w, psi, 1, lbc, ubc, kord, monotone, neumann, early_exit,
                         0));
    hipLaunchKernelGGL(k_add_profile, dim3(2048), dim3(kBlock), 0, c->stream, psi, prof, -1.0,
                       g);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

int dcm_upwind_vert3d(dcm_ctx *c, double dt, const double *w, double *psi, int neumann) {
    if (!c || !w || !psi) return DCM_BAD_ARG;
    const Geo &g = c->geo;
    const Interior L = interior_launch(g);
    double *tmp;
    DCM_CHECK(work_field64(c, kSlotPsi, &tmp));
    // fluxes from the old field into scratch, then the in-place update
    // (advection_3d.F90:754-789)
    hipLaunchKernelGGL(k_vupwind, L.grid, L.block, 0, c->stream, (const double *)psi,
                       (const double *)nullptr, w, tmp, g, c->met, dt, neumann,
                       (const int *)nullptr, 0, (double *)nullptr);
    DCM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_copy_interior, L.grid, L.block, 0, c->stream, tmp, psi, g);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

static int mpdata_vert_api(dcm_ctx *c, double dt, const double *w, double *psi, int nq,
                           double *lbc, double *ubc, int kord, int monotone, int neumann,
                           int *early_exit, int interior_min) {
    if (!c || !w || !psi || !lbc || !ubc || nq < 1 || nq > 32) return DCM_BAD_ARG;
    DCM_CHECK(mpdata_vert_core(c, dt, w, psi, nq, lbc, ubc, kord, monotone, neumann,
                               interior_min));
    // lbc/ubc leave as (bc - m) + m, like the Fortran INTENT(INOUT) scalars
    DCM_HIP(hipMemcpyAsync(c->hscal, c->dscal, 48 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    for (int n = 0; n < nq; ++n) {
        const double m = c->hscal[16 + n];
        lbc[n] = (lbc[n] - m) + m;
        ubc[n] = (ubc[n] - m) + m;
    }
    if (early_exit) *early_exit = *(int *)(c->hscal + 8);
    return DCM_OK;
}

int dcm_mpdata_vert3d(dcm_ctx *c, double dt, const double *w, double *psi, double *lbc,
                      double *ubc, int kord, int monotone, int neumann, int *early_exit) {
    return mpdata_vert_api(c, dt, w, psi, 1, lbc, ubc, kord, monotone, neumann, early_exit, 0);
}

int dcm_mpdata_vec_vert3d(dcm_ctx *c, double dt, const double *w, double *psi, int nq,
                          double *lbc, double *ubc, int kord, int monotone, int neumann,
                          int *early_exit) {
    return mpdata_vert_api(c, dt, w, psi, nq, lbc, ubc, kord, monotone, neumann, early_exit, 1);
}

int dcm_adv_ref_state(dcm_ctx *c, double dt, const double *w, double *psi,
                      const double *psi_ref) {
    if (!c || !w || !psi || !psi_ref) return DCM_BAD_ARG;
    const double *dref;
    DCM_CHECK(upload_1d(c, psi_ref, (size_t)c->geo.kp + 2, 0, &dref));
    return adv_ref_state_dev(c, dt, w, psi, dref);
}

}  // extern "C"
