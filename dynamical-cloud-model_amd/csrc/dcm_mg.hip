// dcm_mg.hip -- a geometric multigrid V-cycle as the preconditioner of the
// mixed path's FP32 inner BiCGSTAB (dcm_set_precond_multigrid).
//
// The reference preconditions with three Jacobi sweeps (precondition,
// dynamics.f90:241-270); BiCGSTAB then needs O(n^(1/3)) iterations.  Here the
// same iteration (dynamics.f90:390-430) applies one V-cycle wherever the
// reference applies its sweeps: p = V(r - bet*omg*v) + bet*p and cs = V(s).
// A V-cycle with a fixed number of sweeps is a fixed linear map, as BiCGSTAB
// requires of its preconditioner, and it reduces every frequency of the
// error alike, so the inner solve needs a few iterations at any size.  The
// FP64 path keeps the reference's algorithm; the outer FP64 refinement
// decides convergence either way.
//
// Cell-centred hierarchy on the compact (k fastest) layout of the inner
// solve, every dimension halved per level down to at most 1024 cells:
//   smoother     damped Jacobi x <- x + w (b - A x)/a_p, two sweeps before
//                and two after the coarse correction (the first from zero);
//                on level 0 these are the solver's own sweep kernels, the
//                first one fused into the pass that forms t or s, the last
//                one of the p branch into the update p = x + bet*p;
//   restriction  the mean of the eight children's residuals;
//   prolongation trilinear (weights 3/4, 1/4 per axis), periodic in x and y,
//                the Neumann copy in z, added to the fine iterate;
//   operators    level 0 is the solver's own (the 1-D coefficients of
//                set_mat_a, dynamics.f90:84-96); a coarse coefficient is the
//                mean of its two children's divided by 4 (twice the spacing),
//                exact on the reference's uniform grids;
//   coarsest     one 1024-thread block iterating the smoother in LDS.
// The kernels and the hierarchy exist in FP32 (the mixed path's inner solve)
// and FP64 (the FP64 BiCGSTAB with the V-cycle in place of the reference's
// sweeps, which separates the precision gain from the preconditioner gain).
#include <vector>

#include "dcm_internal.h"

namespace dcm {

namespace {

constexpr int kCoarseSweeps = 96;  // smoother iterations on the coarsest level
constexpr int kCoarseMax = 1024;   // cells of the coarsest level (one block)

template <typename R> struct Lv {
    int ip, jp, kp, n;
    const R *ae, *aw, *an, *as, *au, *ad;
};

template <typename R> __device__ __forceinline__ R inv_of(R ap);
template <> __device__ __forceinline__ float inv_of<float>(float ap) {
    return __builtin_amdgcn_rcpf(ap);
}
template <> __device__ __forceinline__ double inv_of<double>(double ap) { return 1.0 / ap; }

__device__ __forceinline__ bool stopped(const SolverScalars *sc) {
    return sc->breakdown != 0 || sc->done != 0;
}
template <typename R> __device__ __forceinline__ int at(const Lv<R> &L, int k, int j, int i) {
    return ((i - 1) * L.jp + (j - 1)) * L.kp + (k - 1);
}
template <typename R>
__device__ __forceinline__ void cell_of(const Lv<R> &L, int c, int &k, int &j, int &i) {
    const int r = c / L.kp;
    k = c - r * L.kp + 1;
    i = r / L.jp + 1;
    j = r - (i - 1) * L.jp + 1;
}
template <typename R> __device__ __forceinline__ R ap_of(const Lv<R> &L, int k, int j, int i) {
    return (((((-L.ae[i]) - L.aw[i]) - L.an[j]) - L.as[j]) - L.au[k]) - L.ad[k];
}
// A x at (k,j,i): periodic in i and j, the Neumann copy in k
template <typename R, typename P>
__device__ __forceinline__ R ax_of(const Lv<R> &L, P x, int k, int j, int i, R ap) {
    const int iw = i == 1 ? L.ip : i - 1, ie = i == L.ip ? 1 : i + 1;
    const int js = j == 1 ? L.jp : j - 1, jn = j == L.jp ? 1 : j + 1;
    const int kd = k == 1 ? 1 : k - 1, ku = k == L.kp ? L.kp : k + 1;
    return L.aw[i] * x[at(L, k, j, iw)] + L.ae[i] * x[at(L, k, j, ie)] +
           L.as[j] * x[at(L, k, js, i)] + L.an[j] * x[at(L, k, jn, i)] +
           L.ad[k] * x[at(L, kd, j, i)] + L.au[k] * x[at(L, ku, j, i)] + ap * x[at(L, k, j, i)];
}

// xo = x + w (b - A x)/a_p; x == nullptr: a sweep from zero, xo = w b/a_p
template <typename R>
__global__ __launch_bounds__(256) void k_mg_smooth(const R *__restrict__ b,
                                                   const R *__restrict__ x,
                                                   R *__restrict__ xo, Lv<R> L, R w,
                                                   const SolverScalars *sc) {
    if (stopped(sc)) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L.n) return;
    int k, j, i;
    cell_of(L, c, k, j, i);
    const R ap = ap_of(L, k, j, i);
    const R inv = inv_of<R>(ap);
    if (!x) {
        xo[c] = w * b[c] * inv;
        return;
    }
    xo[c] = x[c] + w * (b[c] - ax_of(L, x, k, j, i, ap)) * inv;
}

// bc = mean over the eight children of (b - A x) on the fine level
template <typename R>
__global__ __launch_bounds__(256) void k_mg_restrict(const R *__restrict__ b,
                                                     const R *__restrict__ x,
                                                     R *__restrict__ bc, Lv<R> F, Lv<R> C,
                                                     const SolverScalars *sc) {
    if (stopped(sc)) return;
    const int cc = blockIdx.x * blockDim.x + threadIdx.x;
    if (cc >= C.n) return;
    int K, J, I;
    cell_of(C, cc, K, J, I);
    R acc = R(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int k = 2 * K - 1 + (q & 1), j = 2 * J - 1 + ((q >> 1) & 1),
                  i = 2 * I - 1 + (q >> 2);
        const R ap = ap_of(F, k, j, i);
        acc += b[at(F, k, j, i)] - ax_of(F, x, k, j, i, ap);
    }
    bc[cc] = R(0.125) * acc;
}

// x += trilinear interpolation of the coarse correction xc
template <typename R>
__global__ __launch_bounds__(256) void k_mg_prolong(const R *__restrict__ xc,
                                                    R *__restrict__ x, Lv<R> F, Lv<R> C,
                                                    const SolverScalars *sc) {
    if (stopped(sc)) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= F.n) return;
    int k, j, i;
    cell_of(F, c, k, j, i);
    const int I = (i + 1) >> 1, J = (j + 1) >> 1, K = (k + 1) >> 1;
    // the second coarse cell along each axis: the one on this child's side
    int I2 = (i & 1) ? I - 1 : I + 1;
    int J2 = (j & 1) ? J - 1 : J + 1;
    int K2 = (k & 1) ? K - 1 : K + 1;
    I2 = I2 < 1 ? C.ip : (I2 > C.ip ? 1 : I2);
    J2 = J2 < 1 ? C.jp : (J2 > C.jp ? 1 : J2);
    K2 = K2 < 1 ? 1 : (K2 > C.kp ? C.kp : K2);
    const R a = R(0.75), q = R(0.25);
    const R v =
        a * (a * (a * xc[at(C, K, J, I)] + q * xc[at(C, K2, J, I)]) +
             q * (a * xc[at(C, K, J2, I)] + q * xc[at(C, K2, J2, I)])) +
        q * (a * (a * xc[at(C, K, J, I2)] + q * xc[at(C, K2, J, I2)]) +
             q * (a * xc[at(C, K, J2, I2)] + q * xc[at(C, K2, J2, I2)]));
    x[c] += v;
}

// the coarsest level: `sweeps` smoother iterations from zero in LDS, one block
template <typename R>
__global__ __launch_bounds__(kCoarseMax) void k_mg_coarse(const R *__restrict__ b,
                                                          R *__restrict__ x, Lv<R> L,
                                                          int sweeps, R w,
                                                          const SolverScalars *sc) {
    __shared__ R xa[kCoarseMax], xb[kCoarseMax];
    if (stopped(sc)) return;  // the same value for every thread of the block
    const int c = threadIdx.x;
    const bool act = c < L.n;
    int k = 1, j = 1, i = 1;
    R ap = R(-1), bc = R(0), inv = R(0);
    if (act) {
        cell_of(L, c, k, j, i);
        ap = ap_of(L, k, j, i);
        inv = inv_of<R>(ap);
        bc = b[c];
        xa[c] = w * bc * inv;
    }
    __syncthreads();
    R *cur = xa, *nxt = xb;
    for (int s = 1; s < sweeps; ++s) {
        if (act) nxt[c] = cur[c] + w * (bc - ax_of(L, cur, k, j, i, ap)) * inv;
        __syncthreads();
        R *t = cur;
        cur = nxt;
        nxt = t;
    }
    if (act) x[c] = cur[c];
}

}  // namespace

// The hierarchy: level 0 is the inner solve's grid (its vectors are passed
// to mg_correct); levels 1.. own their right-hand side, iterate and scratch,
// in FP32 and in FP64.
template <typename R> struct MgLevels {
    Lv<R> lv[16];
    R *b[16] = {}, *x[16] = {}, *t[16] = {};
    R *coef = nullptr;  // the coarse levels' 1-D coefficients
};
struct MgHier {
    int nlev = 0;
    MgLevels<float> f32;
    MgLevels<double> f64;
};
template <typename R> static MgLevels<R> &levels_of(MgHier *H);
template <> MgLevels<float> &levels_of<float>(MgHier *H) { return H->f32; }
template <> MgLevels<double> &levels_of<double>(MgHier *H) { return H->f64; }

template <typename R> static void free_levels(MgLevels<R> &V, int nlev) {
    for (int l = 1; l < nlev; ++l) {
        (void)hipFree(V.b[l]);
        (void)hipFree(V.x[l]);
        (void)hipFree(V.t[l]);
    }
    (void)hipFree(V.coef);
}

void mg_release(dcm_ctx *c) {
    MgHier *H = (MgHier *)c->mg;
    if (!H) return;
    free_levels(H->f32, H->nlev);
    free_levels(H->f64, H->nlev);
    delete H;
    c->mg = nullptr;
}

// Levels of the grid: every dimension halved while all three are even, down
// to at most kCoarseMax cells; false when that size is never reached.
static bool mg_levels(int ip, int jp, int kp, std::vector<int> &dims) {
    dims = {ip, jp, kp};
    while ((long long)ip * jp * kp > kCoarseMax) {
        if (ip % 2 || jp % 2 || kp % 2) return false;
        ip /= 2;
        jp /= 2;
        kp /= 2;
        dims.insert(dims.end(), {ip, jp, kp});
    }
    return dims.size() >= 6;  // at least one coarse level
}

bool mg_possible(const Geo &g) {
    std::vector<int> d;
    return mg_levels(g.ip, g.jp, g.kp, d);
}

// the device levels of one precision: coarse coefficients from h (the
// doubles of every coarse level's six 1-D arrays at off), level 0 from the
// context's own coefficients
template <typename R>
static int build_levels(MgLevels<R> &V, MgHier *H, const std::vector<int> &d,
                        const std::vector<double> &h, const std::vector<size_t> &off,
                        const R *ae, const R *aw, const R *an, const R *as, const R *au,
                        const R *ad) {
    std::vector<R> hr(h.begin(), h.end());
    if (!hr.empty()) {
        if (hipMalloc((void **)&V.coef, hr.size() * sizeof(R)) != hipSuccess)
            return DCM_OUT_OF_MEMORY;
        if (hipMemcpy(V.coef, hr.data(), hr.size() * sizeof(R), hipMemcpyHostToDevice) !=
            hipSuccess)
            return DCM_HIP_ERROR;
    }
    V.lv[0] = Lv<R>{d[0], d[1], d[2], d[0] * d[1] * d[2], ae, aw, an, as, au, ad};
    for (int l = 1; l < H->nlev; ++l) {
        Lv<R> &L = V.lv[l];
        L.ip = d[3 * l];
        L.jp = d[3 * l + 1];
        L.kp = d[3 * l + 2];
        L.n = L.ip * L.jp * L.kp;
        const R **dst[6] = {&L.ae, &L.aw, &L.an, &L.as, &L.au, &L.ad};
        for (int a = 0; a < 6; ++a) *dst[a] = V.coef + off[6 * l + a];
        for (R **p : {&V.b[l], &V.x[l], &V.t[l]})
            if (hipMalloc((void **)p, (size_t)L.n * sizeof(R)) != hipSuccess)
                return DCM_OUT_OF_MEMORY;
    }
    return DCM_OK;
}

int mg_prepare(dcm_ctx *c) {
    if (c->mg) return DCM_OK;
    const Geo &g = c->geo;
    std::vector<int> d;
    if (!mg_levels(g.ip, g.jp, g.kp, d)) {
        set_error("multigrid preconditioner: the grid does not halve (all of ip, jp, kp "
                  "even at every level) down to 1024 cells or fewer");
        return DCM_BAD_ARG;
    }
    MgHier *H = new MgHier();
    H->nlev = (int)d.size() / 3;
    // level-0 coefficients as set_mat_a builds them (1-based, 0 at the ends)
    const int ip = g.ip, jp = g.jp, kp = g.kp;
    const std::vector<double> &dx = c->h_dx, &dy = c->h_dy, &dz = c->h_dz;
    const std::vector<double> &dxn = c->h_dxn, &dyn = c->h_dyn, &dzn = c->h_dzn;
    std::vector<double> f[6];
    f[0].assign(ip + 2, 0.0); f[1].assign(ip + 2, 0.0);
    f[2].assign(jp + 2, 0.0); f[3].assign(jp + 2, 0.0);
    f[4].assign(kp + 2, 0.0); f[5].assign(kp + 2, 0.0);
    for (int i = 1; i <= ip; ++i) {
        f[0][i] = 1.0 / (dx[i - 1] * dxn[i]);
        f[1][i] = 1.0 / (dx[i - 1] * dxn[i - 1]);
    }
    for (int j = 1; j <= jp; ++j) {
        f[2][j] = 1.0 / (dy[j - 1] * dyn[j]);
        f[3][j] = 1.0 / (dy[j - 1] * dyn[j - 1]);
    }
    for (int k = 1; k <= kp; ++k) {
        f[4][k] = 1.0 / (dz[k - 1] * dzn[k]);
        f[5][k] = 1.0 / (dz[k - 1] * dzn[k - 1]);
    }
    // coarse levels: the mean of the two children's coefficients over 4
    std::vector<double> h;
    std::vector<size_t> off(6 * H->nlev, 0);
    for (int l = 1; l < H->nlev; ++l) {
        const int n3[3] = {d[3 * l], d[3 * l + 1], d[3 * l + 2]};
        for (int a = 0; a < 6; ++a) {
            const int n = n3[a / 2];
            std::vector<double> nx(n + 2, 0.0);
            for (int m = 1; m <= n; ++m) nx[m] = 0.125 * (f[a][2 * m - 1] + f[a][2 * m]);
            off[6 * l + a] = h.size();
            for (double v : nx) h.push_back(v);
            f[a] = nx;
        }
    }
    int st = build_levels<float>(H->f32, H, d, h, off, c->met32.ae, c->met32.aw, c->met32.an,
                                 c->met32.as, c->met32.au, c->met32.ad);
    if (st == DCM_OK)
        st = build_levels<double>(H->f64, H, d, h, off, c->met.ae, c->met.aw, c->met.an,
                                  c->met.as, c->met.au, c->met.ad);
    c->mg = H;
    if (st != DCM_OK) {
        mg_release(c);
        set_error("multigrid preconditioner: allocating the coarse levels failed");
    }
    return st;
}

// The coarse-grid correction of a V-cycle on the inner solve's grid: x0 +=
// P V(R (b0 - A x0)), with V the cycle over levels 1.. (two smoother sweeps
// before and after each coarser correction, the coarsest level solved in one
// block).  The caller smooths level 0 with the solver's own sweep kernel.
// Every kernel skips its work once the solve has stopped.
template <typename R>
int mg_correct(dcm_ctx *c, const R *b0, R *x0, const SolverScalars *sc) {
    MgHier *H = (MgHier *)c->mg;
    if (!H) {
        set_error("multigrid preconditioner used before mg_prepare");
        return DCM_BAD_ARG;
    }
    MgLevels<R> &V = levels_of<R>(H);
    const int L = H->nlev;
    auto blocks = [](int n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); };
    const R w = (R)kMgOmega;
    hipLaunchKernelGGL(k_mg_restrict<R>, blocks(V.lv[1].n), dim3(kBlock), 0, c->stream, b0,
                       (const R *)x0, V.b[1], V.lv[0], V.lv[1], sc);
    for (int l = 1; l < L - 1; ++l) {
        const Lv<R> &F = V.lv[l], &C = V.lv[l + 1];
        hipLaunchKernelGGL(k_mg_smooth<R>, blocks(F.n), dim3(kBlock), 0, c->stream,
                           (const R *)V.b[l], (const R *)nullptr, V.t[l], F, w, sc);
        hipLaunchKernelGGL(k_mg_smooth<R>, blocks(F.n), dim3(kBlock), 0, c->stream,
                           (const R *)V.b[l], (const R *)V.t[l], V.x[l], F, w, sc);
        hipLaunchKernelGGL(k_mg_restrict<R>, blocks(C.n), dim3(kBlock), 0, c->stream,
                           (const R *)V.b[l], (const R *)V.x[l], V.b[l + 1], F, C, sc);
    }
    hipLaunchKernelGGL(k_mg_coarse<R>, dim3(1), dim3(kCoarseMax), 0, c->stream,
                       (const R *)V.b[L - 1], V.x[L - 1], V.lv[L - 1], kCoarseSweeps, w, sc);
    for (int l = L - 2; l >= 1; --l) {
        const Lv<R> &F = V.lv[l], &C = V.lv[l + 1];
        hipLaunchKernelGGL(k_mg_prolong<R>, blocks(F.n), dim3(kBlock), 0, c->stream,
                           (const R *)V.x[l + 1], V.x[l], F, C, sc);
        hipLaunchKernelGGL(k_mg_smooth<R>, blocks(F.n), dim3(kBlock), 0, c->stream,
                           (const R *)V.b[l], (const R *)V.x[l], V.t[l], F, w, sc);
        hipLaunchKernelGGL(k_mg_smooth<R>, blocks(F.n), dim3(kBlock), 0, c->stream,
                           (const R *)V.b[l], (const R *)V.t[l], V.x[l], F, w, sc);
    }
    hipLaunchKernelGGL(k_mg_prolong<R>, blocks(V.lv[0].n), dim3(kBlock), 0, c->stream,
                       (const R *)V.x[1], x0, V.lv[0], V.lv[1], sc);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}
template int mg_correct<float>(dcm_ctx *, const float *, float *, const SolverScalars *);
template int mg_correct<double>(dcm_ctx *, const double *, double *, const SolverScalars *);

}  // namespace dcm

using namespace dcm;

extern "C" {

int dcm_set_precond_multigrid(dcm_ctx *c, int on) {
    if (!c) return DCM_BAD_ARG;
    if (on) {
        DCM_CHECK(mg_prepare(c));
        c->mg_on = true;
    } else {
        c->mg_on = false;
    }
    return DCM_OK;
}

}  // extern "C"
