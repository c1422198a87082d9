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
// solve, every even dimension halved per level (an odd one, such as a
// single cell in y, keeps its size: semi-coarsening) down to at most 1024
// cells, and on while some dimension is even, down to 64 cells or fewer:
//   smoother     damped Jacobi x <- x + w (b - A x)/a_p, two sweeps before
//                and two after the coarse correction (the first from zero);
//                on level 0 these are the solver's own sweep kernels, the
//                first one fused into the pass that forms t or s, the last
//                one of the p branch into the update p = x + bet*p;
//   restriction  the mean of the eight children's residuals (two or four
//                where an axis keeps its size);
//   prolongation trilinear (weights 3/4, 1/4 per axis; an axis that keeps
//                its size is copied), periodic in x and y, the Neumann copy
//                in z, added to the fine iterate;
//   operators    level 0 is the solver's own (the 1-D coefficients of
//                set_mat_a, dynamics.f90:84-96); a coarse coefficient along
//                a halved axis is the mean of its two children's divided by
//                4 (twice the spacing), exact on the reference's uniform
//                grids, and along an axis that keeps its size the fine one;
//   coarsest     solved exactly: x = M b with M = (A_c + e e^T)^-1, e the
//                constant vector (the null space of the periodic/Neumann
//                operator; for b in the range of A_c this is the zero-mean
//                solution), M precomputed on the host in FP64 by
//                Gauss-Jordan elimination and applied as one dense
//                matrix-vector product by one block; above 512 cells, or
//                where a pivot vanishes, the coarsest level iterates the
//                smoother 96 times instead.
// Launch granularity: a level of a few thousand cells takes as long to
// launch as to sweep, so every level from the first that fits the LDS of
// one workgroup (with all the levels below it), and has at most two cells
// per thread, down to the coarsest runs in ONE kernel of one 256-thread
// block (k_mg_tail): each pass of the cycle over those levels is a loop
// over the level's cells followed by a workgroup barrier, iterates,
// right-hand sides and 1-D coefficients held in LDS.  At 256x256x128 that
// is the levels of 256 and 32 cells: their 7 launches per V-cycle become
// one, with the same arithmetic.  The block is small on purpose: beside
// the advection kernels of the model step, which fill the CUs, a 256-thread
// block finds a free slot sooner than a 1024-thread one (the step 5.90 ->
// 5.84 ms, with the 2048-cell level moved to the fused legs).  A hierarchy
// whose level of 513..2048 cells cannot run as fused legs (a
// semi-coarsened pair) takes a 1024-thread block with that level in it.
// The kernels and the hierarchy exist in FP32 (the mixed path's inner solve)
// and FP64 (the FP64 BiCGSTAB with the V-cycle in place of the reference's
// sweeps, which separates the precision gain from the preconditioner gain).
#include <algorithm>
#include <cmath>
#include <vector>

#include "dcm_internal.h"

namespace dcm {

namespace {

constexpr int kCoarseSweeps = 96;  // smoother iterations on the coarsest level
constexpr int kCoarseMax = 1024;   // cells of the coarsest level (one block)
constexpr int kDenseMax = 512;     // coarsest levels solved by the dense inverse
                                   // (O(n^3) host work at set-up)
constexpr int kDenseTarget = 64;   // the hierarchy halves on toward this size
constexpr int kFuseMax = 1 << 17;  // largest level whose legs run as one kernel each

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

// whether C halves every dimension of F (else some axis keeps its size)
template <typename R> __host__ __device__ __forceinline__ bool halves(const Lv<R> &F,
                                                                      const Lv<R> &C) {
    return F.ip == 2 * C.ip && F.jp == 2 * C.jp && F.kp == 2 * C.kp;
}
// The transfers of a semi-coarsened pair of levels: along an axis whose
// size C keeps there is one child (the same index) and no interpolation.
// The mean of the children's (b - A x), in the child order of the halving
// kernels (k fastest, then j, then i).
template <typename R, typename PB, typename PX>
__device__ __forceinline__ R restrict_semi(const Lv<R> &F, const Lv<R> &C, PB b, PX x, int K,
                                           int J, int I) {
    const int ni = F.ip != C.ip ? 2 : 1, nj = F.jp != C.jp ? 2 : 1, nk = F.kp != C.kp ? 2 : 1;
    R acc = R(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dk = q & 1, dj = (q >> 1) & 1, di = q >> 2;
        if (dk >= nk || dj >= nj || di >= ni) continue;
        const int k = nk * (K - 1) + 1 + dk, j = nj * (J - 1) + 1 + dj, i = ni * (I - 1) + 1 + di;
        const R ap = ap_of(F, k, j, i);
        acc += b[at(F, k, j, i)] - ax_of(F, x, k, j, i, ap);
    }
    const int halved = ni + nj + nk - 3;
    const R wt = halved == 3 ? R(0.125) : (halved == 2 ? R(0.25) : (halved == 1 ? R(0.5) : R(1)));
    return wt * acc;
}
// P xc at fine cell (k,j,i): the trilinear weights (3/4, 1/4) along halved
// axes, (1, 0) along the others
template <typename R, typename PC>
__device__ __forceinline__ R prolong_semi(const Lv<R> &F, const Lv<R> &C, PC xc, int k, int j,
                                          int i) {
    const bool hi = F.ip != C.ip, hj = F.jp != C.jp, hk = F.kp != C.kp;
    const int I = hi ? (i + 1) >> 1 : i, J = hj ? (j + 1) >> 1 : j, K = hk ? (k + 1) >> 1 : k;
    int I2 = hi ? ((i & 1) ? I - 1 : I + 1) : I;
    int J2 = hj ? ((j & 1) ? J - 1 : J + 1) : J;
    int K2 = hk ? ((k & 1) ? K - 1 : K + 1) : K;
    I2 = I2 < 1 ? C.ip : (I2 > C.ip ? 1 : I2);
    J2 = J2 < 1 ? C.jp : (J2 > C.jp ? 1 : J2);
    K2 = K2 < 1 ? 1 : (K2 > C.kp ? C.kp : K2);
    const R ai = hi ? R(0.75) : R(1), qi = hi ? R(0.25) : R(0);
    const R aj = hj ? R(0.75) : R(1), qj = hj ? R(0.25) : R(0);
    const R ak = hk ? R(0.75) : R(1), qk = hk ? R(0.25) : R(0);
    return ai * (aj * (ak * xc[at(C, K, J, I)] + qk * xc[at(C, K2, J, I)]) +
                 qj * (ak * xc[at(C, K, J2, I)] + qk * xc[at(C, K2, J2, I)])) +
           qi * (aj * (ak * xc[at(C, K, J, I2)] + qk * xc[at(C, K2, J, I2)]) +
                 qj * (ak * xc[at(C, K, J2, I2)] + qk * xc[at(C, K2, J2, I2)]));
}

// One thread per cell of a level on a 3-D grid: x-threads along k (the
// fastest index), y-threads along j, one i per grid z (no divisions).
struct Grid3 {
    dim3 grid, block;
};
template <typename R> static Grid3 grid3(const Lv<R> &L) {
    const int bx = L.kp >= 64 ? 64 : (L.kp >= 32 ? 32 : (L.kp >= 16 ? 16 : 8));
    const int by = 256 / bx;
    return {dim3((L.kp + bx - 1) / bx, (L.jp + by - 1) / by, L.ip), dim3(bx, by, 1)};
}
template <typename R>
__device__ __forceinline__ bool cell3(const Lv<R> &L, int &k, int &j, int &i) {
    k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    i = blockIdx.z + 1;
    return k <= L.kp && j <= L.jp;
}

// xo = x + w (b - A x)/a_p; x == nullptr: a sweep from zero, xo = w b/a_p
template <typename R>
__global__ __launch_bounds__(256) void k_mg_smooth(const R *__restrict__ b,
                                                   const R *__restrict__ x,
                                                   R *__restrict__ xo, Lv<R> L, R w,
                                                   const SolverScalars *sc) {
    if (stopped(sc)) return;
    int k, j, i;
    if (!cell3(L, k, j, i)) return;
    const int c = at(L, k, j, i);
    const R ap = ap_of(L, k, j, i);
    const R inv = inv_of<R>(ap);
    if (!x) {
        xo[c] = w * b[c] * inv;
        return;
    }
    xo[c] = x[c] + w * (b[c] - ax_of(L, x, k, j, i, ap)) * inv;
}

// bc = mean over the eight children of (b - A x) on the fine level (Semi:
// C keeps the size of some axis)
template <typename R, bool Semi>
__global__ __launch_bounds__(256) void k_mg_restrict(const R *__restrict__ b,
                                                     const R *__restrict__ x,
                                                     R *__restrict__ bc, Lv<R> F, Lv<R> C,
                                                     const SolverScalars *sc) {
    if (stopped(sc)) return;
    int K, J, I;
    if (!cell3(C, K, J, I)) return;
    const int cc = at(C, K, J, I);
    if (Semi) {
        bc[cc] = restrict_semi(F, C, b, x, K, J, I);
        return;
    }
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

// x += trilinear interpolation of the coarse correction xc (Semi: C keeps
// the size of some axis)
template <typename R, bool Semi>
__global__ __launch_bounds__(256) void k_mg_prolong(const R *__restrict__ xc,
                                                    R *__restrict__ x, Lv<R> F, Lv<R> C,
                                                    const SolverScalars *sc) {
    if (stopped(sc)) return;
    int k, j, i;
    if (!cell3(F, k, j, i)) return;
    const int c = at(F, k, j, i);
    if (Semi) {
        x[c] += prolong_semi(F, C, xc, k, j, i);
        return;
    }
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

// x += P xc on a level with an even ip, marching: a thread owns one (k,j)
// column and walks kProPlanes fine planes, two per coarse plane, keeping
// the four coarse values (K or K2) x (J or J2) of coarse planes I-1, I, I+1
// in registers, so each coarse value is read once per column pair instead of
// once per fine cell.  The same interpolation expression as k_mg_prolong.
constexpr int kProPlanes = 8;
template <typename R> struct Quad {
    R kj, k2j, kj2, k2j2;
};
template <typename R>
__device__ __forceinline__ Quad<R> quad_at(const R *__restrict__ xc, int plane, int o,
                                           int dk, int dj) {
    const R *p = xc + plane + o;
    return Quad<R>{p[0], p[dk], p[dj], p[dk + dj]};
}
template <typename R>
__device__ __forceinline__ R interp(const Quad<R> &A, const Quad<R> &B) {
    const R a = R(0.75), q = R(0.25);
    return a * (a * (a * A.kj + q * A.k2j) + q * (a * A.kj2 + q * A.k2j2)) +
           q * (a * (a * B.kj + q * B.k2j) + q * (a * B.kj2 + q * B.k2j2));
}
template <typename R>
__global__ __launch_bounds__(256) void k_mg_prolong_march(const R *__restrict__ xc,
                                                          R *__restrict__ x, Lv<R> F, Lv<R> C,
                                                          const SolverScalars *sc) {
    if (stopped(sc)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (k > F.kp || j > F.jp) return;
    const int i0 = blockIdx.z * kProPlanes + 1;  // odd
    const int i1 = min(i0 + kProPlanes - 1, F.ip);
    const int K = (k + 1) >> 1, J = (j + 1) >> 1;
    int K2 = (k & 1) ? K - 1 : K + 1;
    int J2 = (j & 1) ? J - 1 : J + 1;
    J2 = J2 < 1 ? C.jp : (J2 > C.jp ? 1 : J2);
    K2 = K2 < 1 ? 1 : (K2 > C.kp ? C.kp : K2);
    const int o = (J - 1) * C.kp + (K - 1);  // (K,J) within a coarse plane
    const int dk = K2 - K, dj = (J2 - J) * C.kp;
    const int CP = C.jp * C.kp;
    auto plane = [&](int I) { return ((I < 1 ? C.ip : (I > C.ip ? 1 : I)) - 1) * CP; };
    int I = (i0 + 1) >> 1;
    Quad<R> Qm = quad_at(xc, plane(I - 1), o, dk, dj);
    Quad<R> Q0 = quad_at(xc, plane(I), o, dk, dj);
    Quad<R> Qp = quad_at(xc, plane(I + 1), o, dk, dj);
    const int FP = F.jp * F.kp;
    int c = (i0 - 1) * FP + (j - 1) * F.kp + (k - 1);
    for (int i = i0; i <= i1; i += 2) {
        // the next coarse plane in flight while this pair is computed
        Quad<R> Qn = Qp;
        if (i + 2 <= i1) Qn = quad_at(xc, plane(I + 2), o, dk, dj);
        x[c] += interp(Q0, Qm);  // fine plane 2I-1: the second coarse plane is I-1
        if (i + 1 <= i1) x[c + FP] += interp(Q0, Qp);  // 2I: I+1
        Qm = Q0;
        Q0 = Qp;
        Qp = Qn;
        ++I;
        c += 2 * FP;
    }
}

// bc = mean over the eight children of (b - A x), marching: a thread owns
// one coarse (K,J) column, i.e. four fine child columns, and walks
// kProPlanes/2 coarse planes; x of the child columns at fine planes i-1 and
// i+1 comes from the previous and next coarse planes (registers), so a
// coarse cell costs 32 loads instead of 64.  The same arithmetic, in the
// same child order, as k_mg_restrict.
template <typename R> struct Kid {
    R ae, aw, an, as, au, ad;
};
template <typename R>
__device__ __forceinline__ R kid_res(const Kid<R> &q, R xw, R xe, R xs, R xn, R xd, R xu, R xc,
                                     R bv) {
    const R ap = (((((-q.ae) - q.aw) - q.an) - q.as) - q.au) - q.ad;
    return bv - (q.aw * xw + q.ae * xe + q.as * xs + q.an * xn + q.ad * xd + q.au * xu + ap * xc);
}
template <typename R> struct Four {
    R a, b, c, d;  // child columns (ka,ja), (kb,ja), (ka,jb), (kb,jb)
};
template <typename R>
__global__ __launch_bounds__(256) void k_mg_restrict_march(const R *__restrict__ b,
                                                           const R *__restrict__ x,
                                                           R *__restrict__ bc, Lv<R> F, Lv<R> C,
                                                           const SolverScalars *sc) {
    if (stopped(sc)) return;
    const int K = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int J = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (K > C.kp || J > C.jp) return;
    const int I0 = blockIdx.z * (kProPlanes / 2) + 1;
    const int I1 = min(I0 + kProPlanes / 2 - 1, C.ip);
    const int ka = 2 * K - 1, kb = 2 * K, ja = 2 * J - 1, jb = 2 * J;
    const int jaS = ja == 1 ? F.jp : ja - 1, jbN = jb == F.jp ? 1 : jb + 1;
    const int kaD = ka == 1 ? 1 : ka - 1, kbU = kb == F.kp ? F.kp : kb + 1;
    const int FP = F.jp * F.kp;
    auto col = [&](int k, int j) { return (j - 1) * F.kp + (k - 1); };
    const int cA = col(ka, ja), cB = col(kb, ja), cC = col(ka, jb), cD = col(kb, jb);
    const int cS1 = col(ka, jaS), cS2 = col(kb, jaS), cN1 = col(ka, jbN), cN2 = col(kb, jbN);
    const int cD1 = col(kaD, ja), cD2 = col(kaD, jb), cU1 = col(kbU, ja), cU2 = col(kbU, jb);
    const R asa = F.as[ja], ana = F.an[ja], asb = F.as[jb], anb = F.an[jb];
    const R ada = F.ad[ka], aua = F.au[ka], adb = F.ad[kb], aub = F.au[kb];
    auto four = [&](int i) {
        const R *p = x + (i - 1) * FP;
        return Four<R>{p[cA], p[cB], p[cC], p[cD]};
    };
    auto wrapi = [&](int i) { return i < 1 ? F.ip : (i > F.ip ? 1 : i); };
    // the child columns at fine plane 2*I0-2 (west of the first child plane)
    Four<R> W = four(wrapi(2 * I0 - 2));
    Four<R> A = four(2 * I0 - 1);
    for (int I = I0; I <= I1; ++I) {
        const int ia = 2 * I - 1, ib = 2 * I;
        const Four<R> B = four(ib);
        const Four<R> E = four(wrapi(ib + 1));
        const R *pa = x + (ia - 1) * FP, *pb = x + (ib - 1) * FP;
        const R *ba = b + (ia - 1) * FP, *bb = b + (ib - 1) * FP;
        R acc = R(0);
        {  // plane ia: west W, east B
            const R ae = F.ae[ia], aw = F.aw[ia];
            acc += kid_res(Kid<R>{ae, aw, ana, asa, aua, ada}, W.a, B.a, pa[cS1], A.c, pa[cD1],
                           A.b, A.a, ba[cA]);
            acc += kid_res(Kid<R>{ae, aw, ana, asa, aub, adb}, W.b, B.b, pa[cS2], A.d, A.a,
                           pa[cU1], A.b, ba[cB]);
            acc += kid_res(Kid<R>{ae, aw, anb, asb, aua, ada}, W.c, B.c, A.a, pa[cN1], pa[cD2],
                           A.d, A.c, ba[cC]);
            acc += kid_res(Kid<R>{ae, aw, anb, asb, aub, adb}, W.d, B.d, A.b, pa[cN2], A.c,
                           pa[cU2], A.d, ba[cD]);
        }
        {  // plane ib: west A, east E
            const R ae = F.ae[ib], aw = F.aw[ib];
            acc += kid_res(Kid<R>{ae, aw, ana, asa, aua, ada}, A.a, E.a, pb[cS1], B.c, pb[cD1],
                           B.b, B.a, bb[cA]);
            acc += kid_res(Kid<R>{ae, aw, ana, asa, aub, adb}, A.b, E.b, pb[cS2], B.d, B.a,
                           pb[cU1], B.b, bb[cB]);
            acc += kid_res(Kid<R>{ae, aw, anb, asb, aua, ada}, A.c, E.c, B.a, pb[cN1], pb[cD2],
                           B.d, B.c, bb[cC]);
            acc += kid_res(Kid<R>{ae, aw, anb, asb, aub, adb}, A.d, E.d, B.b, pb[cN2], B.c,
                           pb[cU2], B.d, bb[cD]);
        }
        bc[((I - 1) * C.jp + (J - 1)) * C.kp + (K - 1)] = R(0.125) * acc;
        W = B;
        A = E;
    }
}

// row c of the dense coarsest solve: sum over r of M[c][r] b[r], M stored
// transposed (Mt[r n + c]) so that the threads of a block read whole rows
template <typename R, typename PB>
__device__ __forceinline__ R dense_row(const R *Mt, int n, PB b, int c) {
    R acc = R(0);
#pragma unroll 8
    for (int r = 0; r < n; ++r) acc += Mt[(size_t)r * n + c] * b[r];
    return acc;
}

// the coarsest level, one block: x = M b when Mt is given, else `sweeps`
// smoother iterations from zero in LDS
template <typename R>
__global__ __launch_bounds__(kCoarseMax) void k_mg_coarse(const R *__restrict__ b,
                                                          R *__restrict__ x, Lv<R> L,
                                                          int sweeps, R w,
                                                          const SolverScalars *sc,
                                                          const R *__restrict__ Mt) {
    __shared__ R xa[kCoarseMax], xb[kCoarseMax];
    if (stopped(sc)) return;  // the same value for every thread of the block
    const int c = threadIdx.x;
    const bool act = c < L.n;
    if (Mt) {
        if (act) xb[c] = b[c];
        __syncthreads();
        if (act) x[c] = dense_row(Mt, L.n, (const R *)xb, c);
        return;
    }
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

// ---------------------------------------------------------------------------
// A level's down-leg and up-leg in one kernel each (temporal blocking).  A
// block owns a TI x TJ tile of whole k-columns and keeps the tile with a
// halo of two columns on each side in LDS (box coordinates ii, jj; global
// indices wrap periodically), so the passes that follow one another need no
// grid-wide barrier: the halo cells are computed redundantly.
//   down: t = w b/a_p on the box, x = t + w (b - A t)/a_p on the box less
//         one column, x written on the tile, and the restriction of
//         (b - A x) into the coarse cells of the tile;
//   up:   x' = x + P xc on the box, t = x' + w (b - A x')/a_p on the box
//         less one column, x'' = t + w (b - A t)/a_p on the tile, written to
//         a second array (other blocks still read x in their halos).
// Threads: x along k (blockDim.x = kp), y over the box's columns.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrap1(int v, int n) {
    // 1-based periodic index of v (v may lie a few cells outside 1..n)
    int r = (v - 1) % n;
    if (r < 0) r += n;
    return r + 1;
}

struct Tile2 {
    int TI, TJ;  // tile size in i and j (each a divisor of the level's ip, jp)
};

// The box's 1-D coefficients: ae, aw at the BI box columns' i (wrapped), an,
// as at the BJ box rows' j, in LDS after the three arrays; au, ad of the
// thread's k in registers.
template <typename R> struct BoxCoef {
    const R *ae, *aw, *an, *as;
    R au, ad;
};
template <typename R>
__device__ __forceinline__ BoxCoef<R> box_coef(const Lv<R> &F, R *tab, int ib, int jb, int BI,
                                               int BJ, int k) {
    const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    const int nt = (int)(blockDim.x * blockDim.y);
    for (int e = tid; e < 2 * (BI + BJ); e += nt) {
        R v;
        if (e < BI) v = F.ae[wrap1(ib + e, F.ip)];
        else if (e < 2 * BI) v = F.aw[wrap1(ib + e - BI, F.ip)];
        else if (e < 2 * BI + BJ) v = F.an[wrap1(jb + e - 2 * BI, F.jp)];
        else v = F.as[wrap1(jb + e - 2 * BI - BJ, F.jp)];
        tab[e] = v;
    }
    return BoxCoef<R>{tab, tab + BI, tab + 2 * BI, tab + 2 * BI + BJ, F.au[k], F.ad[k]};
}
// a_p at box column ii, row jj: the order of ap_of
template <typename R>
__device__ __forceinline__ R box_ap(const BoxCoef<R> &B, int ii, int jj) {
    return (((((-B.ae[ii]) - B.aw[ii]) - B.an[jj]) - B.as[jj]) - B.au) - B.ad;
}
// A v at box element e (plane = BJ*kp, row = kp): the operator of ax_of
template <typename R>
__device__ __forceinline__ R ax_box(const BoxCoef<R> &B, const R *v, int e, int plane, int kp,
                                    int k, int ii, int jj, R ap) {
    const int kd = k == 1 ? 0 : 1, ku = k == kp ? 0 : 1;
    return B.aw[ii] * v[e - plane] + B.ae[ii] * v[e + plane] + B.as[jj] * v[e - kp] +
           B.an[jj] * v[e + kp] + B.ad * v[e - kd] + B.au * v[e + ku] + ap * v[e];
}

template <typename R>
__global__ __launch_bounds__(256) void k_mg_down(const R *__restrict__ b, R *__restrict__ xg,
                                                 R *__restrict__ bc, Lv<R> F, Lv<R> C,
                                                 Tile2 T, R w, const SolverScalars *sc) {
    extern __shared__ __align__(16) unsigned char mg_lds[];
    if (stopped(sc)) return;
    const int kp = F.kp, BI = T.TI + 4, BJ = T.TJ + 4;
    const int plane = BJ * kp, box = BI * plane;
    R *lb = reinterpret_cast<R *>(mg_lds), *lt = lb + box, *lx = lt + box;
    const int ntj = F.jp / T.TJ;
    const int ti = (int)blockIdx.x / ntj, tj = (int)blockIdx.x - ti * ntj;
    const int ib = ti * T.TI - 1, jb = tj * T.TJ - 1;  // global i, j of box column (0, 0)
    const int k = (int)threadIdx.x + 1;
    const BoxCoef<R> B = box_coef(F, lx + box, ib, jb, BI, BJ, k);
    // b of the box
#pragma unroll 4
    for (int col = threadIdx.y; col < BI * BJ; col += blockDim.y) {
        const int ii = col / BJ, jj = col - ii * BJ;
        const int i = wrap1(ib + ii, F.ip), j = wrap1(jb + jj, F.jp);
        lb[col * kp + (k - 1)] = b[at(F, k, j, i)];
    }
    __syncthreads();
    // t = w b/a_p on the whole box
    for (int col = threadIdx.y; col < BI * BJ; col += blockDim.y) {
        const int ii = col / BJ, jj = col - ii * BJ;
        const int e = col * kp + (k - 1);
        lt[e] = w * lb[e] * inv_of<R>(box_ap(B, ii, jj));
    }
    __syncthreads();
    // x = t + w (b - A t)/a_p on the box less one column
    for (int col = threadIdx.y; col < (BI - 2) * (BJ - 2); col += blockDim.y) {
        const int ii = 1 + col / (BJ - 2), jj = 1 + col % (BJ - 2);
        const int e = (ii * BJ + jj) * kp + (k - 1);
        const R ap = box_ap(B, ii, jj);
        const R inv = inv_of<R>(ap);
        lx[e] = lt[e] + w * (lb[e] - ax_box(B, (const R *)lt, e, plane, kp, k, ii, jj, ap)) * inv;
    }
    __syncthreads();
    // x on the tile
    for (int col = threadIdx.y; col < T.TI * T.TJ; col += blockDim.y) {
        const int ii = 2 + col / T.TJ, jj = 2 + col % T.TJ;
        xg[at(F, k, jb + jj, ib + ii)] = lx[(ii * BJ + jj) * kp + (k - 1)];
    }
    // the coarse cells of the tile: mean over the eight children of (b - A x)
    // (the children's k differ from the thread's: their au, ad from F)
    if (k <= C.kp) {
        const int K = k;
        for (int col = threadIdx.y; col < (T.TI / 2) * (T.TJ / 2); col += blockDim.y) {
            const int ci = col / (T.TJ / 2), cj = col % (T.TJ / 2);
            const int I = ti * (T.TI / 2) + ci + 1, J = tj * (T.TJ / 2) + cj + 1;
            R acc = R(0);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int kk = 2 * K - 1 + (q & 1);
                const int ii = 2 * ci + 2 + ((q >> 2) & 1), jj = 2 * cj + 2 + ((q >> 1) & 1);
                const int e = (ii * BJ + jj) * kp + (kk - 1);
                BoxCoef<R> Bq = B;
                Bq.au = F.au[kk];
                Bq.ad = F.ad[kk];
                const R ap = box_ap(Bq, ii, jj);
                acc += lb[e] - ax_box(Bq, (const R *)lx, e, plane, kp, kk, ii, jj, ap);
            }
            bc[at(C, K, J, I)] = R(0.125) * acc;
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_mg_up(const R *__restrict__ xc,
                                               const R *__restrict__ xg,
                                               const R *__restrict__ b, R *__restrict__ xo,
                                               Lv<R> F, Lv<R> C, Tile2 T, R w,
                                               const SolverScalars *sc) {
    extern __shared__ __align__(16) unsigned char mg_lds[];
    if (stopped(sc)) return;
    const int kp = F.kp, BI = T.TI + 4, BJ = T.TJ + 4;
    const int plane = BJ * kp, box = BI * plane;
    R *lb = reinterpret_cast<R *>(mg_lds), *lt = lb + box, *lx = lt + box;
    const int ntj = F.jp / T.TJ;
    const int ti = (int)blockIdx.x / ntj, tj = (int)blockIdx.x - ti * ntj;
    const int ib = ti * T.TI - 1, jb = tj * T.TJ - 1;
    const int k = (int)threadIdx.x + 1;
    const BoxCoef<R> B = box_coef(F, lx + box, ib, jb, BI, BJ, k);
    // x' = x + P xc on the whole box (trilinear, as k_mg_prolong), b staged
    const int K = (k + 1) >> 1;
    int K2 = (k & 1) ? K - 1 : K + 1;
    K2 = K2 < 1 ? 1 : (K2 > C.kp ? C.kp : K2);
#pragma unroll 2
    for (int col = threadIdx.y; col < BI * BJ; col += blockDim.y) {
        const int ii = col / BJ, jj = col - ii * BJ;
        const int i = wrap1(ib + ii, F.ip), j = wrap1(jb + jj, F.jp);
        const int e = col * kp + (k - 1);
        const int I = (i + 1) >> 1, J = (j + 1) >> 1;
        int I2 = (i & 1) ? I - 1 : I + 1;
        int J2 = (j & 1) ? J - 1 : J + 1;
        I2 = I2 < 1 ? C.ip : (I2 > C.ip ? 1 : I2);
        J2 = J2 < 1 ? C.jp : (J2 > C.jp ? 1 : J2);
        const R a = R(0.75), q = R(0.25);
        const R v =
            a * (a * (a * xc[at(C, K, J, I)] + q * xc[at(C, K2, J, I)]) +
                 q * (a * xc[at(C, K, J2, I)] + q * xc[at(C, K2, J2, I)])) +
            q * (a * (a * xc[at(C, K, J, I2)] + q * xc[at(C, K2, J, I2)]) +
                 q * (a * xc[at(C, K, J2, I2)] + q * xc[at(C, K2, J2, I2)]));
        const int c = at(F, k, j, i);
        lx[e] = xg[c] + v;
        lb[e] = b[c];
    }
    __syncthreads();
    // t = x' + w (b - A x')/a_p on the box less one column
    for (int col = threadIdx.y; col < (BI - 2) * (BJ - 2); col += blockDim.y) {
        const int ii = 1 + col / (BJ - 2), jj = 1 + col % (BJ - 2);
        const int e = (ii * BJ + jj) * kp + (k - 1);
        const R ap = box_ap(B, ii, jj);
        const R inv = inv_of<R>(ap);
        lt[e] = lx[e] + w * (lb[e] - ax_box(B, (const R *)lx, e, plane, kp, k, ii, jj, ap)) * inv;
    }
    __syncthreads();
    // x'' = t + w (b - A t)/a_p on the tile
    for (int col = threadIdx.y; col < T.TI * T.TJ; col += blockDim.y) {
        const int ii = 2 + col / T.TJ, jj = 2 + col % T.TJ;
        const int e = (ii * BJ + jj) * kp + (k - 1);
        const R ap = box_ap(B, ii, jj);
        const R inv = inv_of<R>(ap);
        xo[at(F, k, jb + jj, ib + ii)] =
            lt[e] + w * (lb[e] - ax_box(B, (const R *)lt, e, plane, kp, k, ii, jj, ap)) * inv;
    }
}

// ---------------------------------------------------------------------------
// The tail of the V-cycle in one block: levels t0..t0+nl-1 of the hierarchy
// (the last the coarsest).  Level t0's right-hand side comes from global
// memory (the restriction kernel of the level above writes it) and its
// result goes back there (for the prolongation into the level above);
// everything else stays in LDS.
// ---------------------------------------------------------------------------
constexpr int kTailThreads = 256;    // the tail's block
constexpr int kTailThreadsMax = 1024;  // ... where a larger first level pays
constexpr int kTailMax = 8;  // levels in the tail

template <typename R> struct TailArgs {
    Lv<R> lv[kTailMax];  // the tail's levels (global coefficient pointers)
    int nl;              // levels in the tail; the last is the coarsest
    int coef[kTailMax];  // LDS offset (elements of R) of each level's six 1-D arrays
    int xo[kTailMax], to[kTailMax], bo[kTailMax];  // LDS offsets of x, t and b (b: levels >= 1)
    const R *b0;         // right-hand side of the first tail level
    R *x0;               // its result
    const R *Mt;         // the coarsest level's dense solve (null: the sweeps)
    int mo;              // its LDS offset
};

// level l of the tail with its coefficient arrays in LDS
template <typename R>
__device__ __forceinline__ Lv<R> tail_level(const TailArgs<R> &A, R *lds, int l) {
    Lv<R> L = A.lv[l];
    R *c = lds + A.coef[l];
    const int ni = L.ip + 2, nj = L.jp + 2, nk = L.kp + 2;
    L.ae = c;
    L.aw = c + ni;
    L.an = c + 2 * ni;
    L.as = c + 2 * ni + nj;
    L.au = c + 2 * ni + 2 * nj;
    L.ad = c + 2 * ni + 2 * nj + nk;
    return L;
}

// t = w b/a_p (a sweep from zero)
template <typename R, typename PB, typename PO>
__device__ __forceinline__ void tail_smooth0(const Lv<R> &L, PB b, PO t, R w) {
    for (int c = threadIdx.x; c < L.n; c += (int)blockDim.x) {
        int k, j, i;
        cell_of(L, c, k, j, i);
        const R inv = inv_of<R>(ap_of(L, k, j, i));
        t[c] = w * b[c] * inv;
    }
}
// xo = x + w (b - A x)/a_p
template <typename R, typename PB, typename PX, typename PO>
__device__ __forceinline__ void tail_smooth(const Lv<R> &L, PB b, PX x, PO xo, R w) {
    for (int c = threadIdx.x; c < L.n; c += (int)blockDim.x) {
        int k, j, i;
        cell_of(L, c, k, j, i);
        const R ap = ap_of(L, k, j, i);
        const R inv = inv_of<R>(ap);
        xo[c] = x[c] + w * (b[c] - ax_of(L, x, k, j, i, ap)) * inv;
    }
}
// bc = mean over the eight children of (b - A x)
template <typename R, typename PB, typename PX, typename PC>
__device__ __forceinline__ void tail_restrict(const Lv<R> &F, const Lv<R> &C, PB b, PX x,
                                              PC bc) {
    const bool semi = !halves(F, C);  // the same for every thread
    for (int cc = threadIdx.x; cc < C.n; cc += (int)blockDim.x) {
        int K, J, I;
        cell_of(C, cc, K, J, I);
        if (semi) {
            bc[cc] = restrict_semi(F, C, b, x, K, J, I);
            continue;
        }
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
}
// x += trilinear interpolation of xc
template <typename R, typename PC, typename PX>
__device__ __forceinline__ void tail_prolong(const Lv<R> &F, const Lv<R> &C, PC xc, PX x) {
    const bool semi = !halves(F, C);  // the same for every thread
    for (int c = threadIdx.x; c < F.n; c += (int)blockDim.x) {
        int k, j, i;
        cell_of(F, c, k, j, i);
        if (semi) {
            x[c] += prolong_semi(F, C, xc, k, j, i);
            continue;
        }
        const int I = (i + 1) >> 1, J = (j + 1) >> 1, K = (k + 1) >> 1;
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
}

template <typename R>
__global__ __launch_bounds__(kTailThreadsMax) void k_mg_tail(TailArgs<R> A, int sweeps, R w,
                                                          const SolverScalars *sc) {
    extern __shared__ __align__(16) unsigned char tail_lds[];
    R *lds = reinterpret_cast<R *>(tail_lds);
    if (stopped(sc)) return;  // the same value for every thread of the block
    const int nl = A.nl;
    // the 1-D coefficients of every tail level into LDS
    for (int l = 0; l < nl; ++l) {
        const Lv<R> &G = A.lv[l];
        const int ni = G.ip + 2, nj = G.jp + 2, nk = G.kp + 2;
        R *c = lds + A.coef[l];
        for (int e = threadIdx.x; e < 2 * (ni + nj + nk); e += (int)blockDim.x) {
            R v;
            if (e < ni) v = G.ae[e];
            else if (e < 2 * ni) v = G.aw[e - ni];
            else if (e < 2 * ni + nj) v = G.an[e - 2 * ni];
            else if (e < 2 * ni + 2 * nj) v = G.as[e - 2 * ni - nj];
            else if (e < 2 * ni + 2 * nj + nk) v = G.au[e - 2 * ni - 2 * nj];
            else v = G.ad[e - 2 * ni - 2 * nj - nk];
            c[e] = v;
        }
    }
    // the first level's right-hand side and the dense coarsest solve
    for (int e = threadIdx.x; e < A.lv[0].n; e += (int)blockDim.x)
        lds[A.bo[0] + e] = A.b0[e];
    if (A.Mt) {
        const int nc = A.lv[nl - 1].n;
        for (int e = threadIdx.x; e < nc * nc; e += (int)blockDim.x) lds[A.mo + e] = A.Mt[e];
    }
    __syncthreads();
    const int last = nl - 1;
    const R *b0 = lds + A.bo[0];
    // down: the first tail level (right-hand side in global memory) ...
    if (last > 0) {
        const Lv<R> F = tail_level(A, lds, 0), C = tail_level(A, lds, 1);
        R *t = lds + A.to[0], *x = lds + A.xo[0];
        tail_smooth0(F, b0, t, w);
        __syncthreads();
        tail_smooth(F, b0, (const R *)t, x, w);
        __syncthreads();
        tail_restrict(F, C, b0, (const R *)x, lds + A.bo[1]);
        __syncthreads();
    }
    // ... and the levels below it, all in LDS
    for (int l = 1; l < last; ++l) {
        const Lv<R> F = tail_level(A, lds, l), C = tail_level(A, lds, l + 1);
        const R *b = lds + A.bo[l];
        R *t = lds + A.to[l], *x = lds + A.xo[l];
        tail_smooth0(F, b, t, w);
        __syncthreads();
        tail_smooth(F, b, (const R *)t, x, w);
        __syncthreads();
        tail_restrict(F, C, b, (const R *)x, lds + A.bo[l + 1]);
        __syncthreads();
    }
    // the coarsest level: the smoother from zero, `sweeps` times, ping-pong
    // between its x and t (at most kCoarseMax cells: one per thread)
    R *res;
    {
        const Lv<R> L = tail_level(A, lds, last);
        const int c = threadIdx.x;
        const bool act = c < L.n;
        int k = 1, j = 1, i = 1;
        R ap = R(-1), bc = R(0), inv = R(0);
        R *xa = lds + A.xo[last], *xb = lds + A.to[last];
        if (A.Mt) {
            // x = M b, both in LDS
            if (act) xa[c] = dense_row((const R *)(lds + A.mo), L.n, (const R *)(lds + A.bo[last]), c);
            __syncthreads();
            res = xa;
        } else {
            if (act) {
                cell_of(L, c, k, j, i);
                ap = ap_of(L, k, j, i);
                inv = inv_of<R>(ap);
                bc = lds[A.bo[last] + c];
                xa[c] = w * bc * inv;
            }
            __syncthreads();
            R *cur = xa, *nxt = xb;
            for (int s = 1; s < sweeps; ++s) {
                if (act)
                    nxt[c] = cur[c] + w * (bc - ax_of(L, (const R *)cur, k, j, i, ap)) * inv;
                __syncthreads();
                R *tmp = cur;
                cur = nxt;
                nxt = tmp;
            }
            res = cur;
        }
    }
    // up: prolongation and two sweeps per level, in LDS ...
    const R *xc = res;
    for (int l = last - 1; l >= 1; --l) {
        const Lv<R> F = tail_level(A, lds, l), C = tail_level(A, lds, l + 1);
        const R *b = lds + A.bo[l];
        R *t = lds + A.to[l], *x = lds + A.xo[l];
        tail_prolong(F, C, xc, x);
        __syncthreads();
        tail_smooth(F, b, (const R *)x, t, w);
        __syncthreads();
        tail_smooth(F, b, (const R *)t, x, w);
        __syncthreads();
        xc = x;
    }
    // ... then the first tail level, whose result goes to global memory
    if (last > 0) {
        const Lv<R> F = tail_level(A, lds, 0), C = tail_level(A, lds, 1);
        R *t = lds + A.to[0], *x = lds + A.xo[0];
        tail_prolong(F, C, xc, x);
        __syncthreads();
        tail_smooth(F, b0, (const R *)x, t, w);
        __syncthreads();
        tail_smooth(F, b0, (const R *)t, A.x0, w);
    } else {
        const int c = threadIdx.x;
        if (c < A.lv[0].n) A.x0[c] = res[c];
    }
}

}  // namespace

// The hierarchy: level 0 is the inner solve's grid (its vectors are passed
// to mg_correct); levels 1.. own their right-hand side, iterate and scratch,
// in FP32 and in FP64.
template <typename R> struct MgLevels {
    Lv<R> lv[16];
    R *b[16] = {}, *x[16] = {}, *t[16] = {};
    R *coef = nullptr;  // the coarse levels' 1-D coefficients
    R *Mt = nullptr;    // the coarsest level's dense solve, transposed (or null)
    Tile2 tile[16] = {};  // temporal-blocking tile of each level (TI = 0: none)
    size_t tile_lds[16] = {};
    int tail = 0;       // first level of the one-block tail (0: none)
    int tail_threads = kTailThreads;  // its block
    size_t tail_lds = 0;
    TailArgs<R> targs{};
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
    (void)hipFree(V.Mt);
}

void mg_release(dcm_ctx *c) {
    MgHier *H = (MgHier *)c->mg;
    if (!H) return;
    free_levels(H->f32, H->nlev);
    free_levels(H->f64, H->nlev);
    delete H;
    c->mg = nullptr;
}

// Levels of the grid: every even dimension halved per level (an odd one
// keeps its size) down to at most kCoarseMax cells; false when that size is
// never reached (every dimension odd above it) or the grid has no coarser
// level.
static bool mg_levels(int ip, int jp, int kp, std::vector<int> &dims) {
    dims = {ip, jp, kp};
    auto coarsen = [&]() {
        if (ip % 2 && jp % 2 && kp % 2) return false;
        if (ip % 2 == 0) ip /= 2;
        if (jp % 2 == 0) jp /= 2;
        if (kp % 2 == 0) kp /= 2;
        dims.insert(dims.end(), {ip, jp, kp});
        return true;
    };
    while ((long long)ip * jp * kp > kCoarseMax)
        if (!coarsen()) return false;
    if (dims.size() < 6) return false;  // at least one coarse level
    // further down while it halves, so that the dense coarsest solve is small
    while ((long long)ip * jp * kp > kDenseTarget && coarsen()) {
    }
    return dims.size() / 3 <= 16;  // (the level arrays' size)
}

bool mg_possible(const Geo &g) {
    std::vector<int> d;
    return mg_levels(g.ip, g.jp, g.kp, d);
}

// The first level of the one-block tail: the lowest l >= 1 whose levels
// l..nlev-1 (x, t and b each, the six 1-D coefficient arrays, the dense
// coarsest solve, every array padded to 16 bytes) fit the LDS of one
// workgroup, with at most
// kTailMax levels and at most 2 cells per thread on the first (a level of
// 16 cells per thread was measured slower in one block than as six
// launches: each pass is a chain of dependent loads per cell), and a
// coarsest level of at most one cell per thread (its solve is one cell per
// thread); else no tail.  The block has kTailThreads threads, or
// kTailThreadsMax when that admits a level that fused legs cannot take.
template <typename R> static bool fusable(const MgLevels<R> &V, int l, int nlev) {
    if (l < 1 || l >= nlev - 1) return false;
    const Lv<R> &L = V.lv[l];
    if (L.kp > 256 || L.n > kFuseMax || !halves(L, V.lv[l + 1])) return false;
    for (int t : {4, 2}) {
        const size_t bytes =
            (3 * (size_t)(t + 4) * (t + 4) * L.kp + 4 * (size_t)(t + 4)) * sizeof(R);
        if (L.ip % t == 0 && L.jp % t == 0 && bytes <= 64 * 1024) return true;
    }
    return false;
}
template <typename R> static int plan_tail(MgLevels<R> &V, int nlev) {
    V.tail = 0;
    V.tail_threads = kTailThreads;
    for (int t0 = 1; t0 < nlev; ++t0)
        if (V.lv[t0].n > 2 * kTailThreads && V.lv[t0].n <= 2 * kTailThreadsMax &&
            !fusable(V, t0, nlev)) {
            V.tail_threads = kTailThreadsMax;
            break;
        }
    const int NT = V.tail_threads;
    if (V.lv[nlev - 1].n > NT) return DCM_OK;
    int dev = 0, lds_max = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess)
        return DCM_HIP_ERROR;
    const size_t pad = 16 / sizeof(R);
    auto up = [&](size_t n) { return (n + pad - 1) / pad * pad; };
    for (int t0 = 1; t0 < nlev; ++t0) {
        const int nl = nlev - t0;
        if (nl > kTailMax || V.lv[t0].n > 2 * NT) continue;
        TailArgs<R> A{};
        size_t off = 0;
        for (int l = 0; l < nl; ++l) {
            const Lv<R> &L = V.lv[t0 + l];
            A.lv[l] = L;
            A.coef[l] = (int)off;
            off += up(2 * (size_t)(L.ip + L.jp + L.kp + 6));
            A.xo[l] = (int)off;
            off += up((size_t)L.n);
            A.to[l] = (int)off;
            off += up((size_t)L.n);
            A.bo[l] = (int)off;
            off += up((size_t)L.n);
        }
        A.mo = (int)off;
        if (V.Mt) off += up((size_t)V.lv[nlev - 1].n * V.lv[nlev - 1].n);
        const size_t bytes = off * sizeof(R);
        if (bytes > (size_t)lds_max) continue;
        A.nl = nl;
        A.b0 = V.b[t0];
        A.x0 = V.x[t0];
        A.Mt = V.Mt;
        // the attribute is the function's, shared by every context: the
        // device's limit, so that no context lowers another's
        if (hipFuncSetAttribute((const void *)k_mg_tail<R>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                lds_max) != hipSuccess)
            return DCM_HIP_ERROR;
        V.tail = t0;
        V.tail_lds = bytes;
        V.targs = A;
        return DCM_OK;
    }
    return DCM_OK;  // no level fits: every pass in its own launch
}

// Tiles of the levels above the tail: 4x4 (else 2x2) columns whose box with
// its halo (3 arrays of (TI+4)(TJ+4)kp and the box's 1-D coefficients) fits
// 64 KiB of LDS, kp <= 256, for levels of at most kFuseMax cells (above
// that the halo work costs more than the launches it saves: 128x128x64 in
// FP32 took 28 us per fused leg against 19 us for its three kernels).
template <typename R> static int plan_tiles(MgLevels<R> &V, int nlev) {
    size_t most = 0;
    for (int l = 1; l < nlev - 1; ++l) {
        const Lv<R> &L = V.lv[l];
        V.tile[l] = Tile2{0, 0};
        if (!fusable(V, l, nlev)) continue;
        for (int t : {4, 2}) {
            const size_t bytes =
                (3 * (size_t)(t + 4) * (t + 4) * L.kp + 4 * (size_t)(t + 4)) * sizeof(R);
            if (L.ip % t == 0 && L.jp % t == 0 && bytes <= 64 * 1024) {
                V.tile[l] = Tile2{t, t};
                V.tile_lds[l] = bytes;
                most = std::max(most, bytes);
                break;
            }
        }
    }
    if (most > 0) {
        // the bound every tile obeys (the attribute is shared by every context)
        if (hipFuncSetAttribute((const void *)k_mg_down<R>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                64 * 1024) != hipSuccess ||
            hipFuncSetAttribute((const void *)k_mg_up<R>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                64 * 1024) != hipSuccess)
            return DCM_HIP_ERROR;
    }
    return DCM_OK;
}

// the device levels of one precision: coarse coefficients from h (the
// doubles of every coarse level's six 1-D arrays at off), level 0 from the
// context's own coefficients
template <typename R>
static int build_levels(MgLevels<R> &V, MgHier *H, const std::vector<int> &d,
                        const std::vector<double> &h, const std::vector<size_t> &off,
                        const std::vector<double> &minv, const R *ae, const R *aw,
                        const R *an, const R *as, const R *au, const R *ad) {
    if (!minv.empty()) {
        std::vector<R> mr(minv.begin(), minv.end());
        if (hipMalloc((void **)&V.Mt, mr.size() * sizeof(R)) != hipSuccess)
            return DCM_OUT_OF_MEMORY;
        if (hipMemcpy(V.Mt, mr.data(), mr.size() * sizeof(R), hipMemcpyHostToDevice) !=
            hipSuccess)
            return DCM_HIP_ERROR;
    }
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
    DCM_CHECK(plan_tail(V, H->nlev));
    return plan_tiles(V, H->nlev);
}

// M^T for M = (A + e e^T)^-1 of the coarsest level (n cells, coefficients
// f[0..5] = ae, aw, an, as, au, ad, 1-based), by Gauss-Jordan elimination
// with partial pivoting in FP64; empty when a pivot vanishes.  A is the
// operator of ax_of: periodic in i and j, the Neumann copy in k.
static std::vector<double> coarse_inverse_t(int ip, int jp, int kp,
                                            const std::vector<double> *f) {
    const int n = ip * jp * kp;
    if (n > kDenseMax) return {};
    auto at_ = [&](int k, int j, int i) { return ((i - 1) * jp + (j - 1)) * kp + (k - 1); };
    // [A + e e^T | I], row-major
    std::vector<double> a((size_t)n * 2 * n, 0.0);
    auto A = [&](int r, int col) -> double & { return a[(size_t)r * 2 * n + col]; };
    for (int i = 1; i <= ip; ++i)
        for (int j = 1; j <= jp; ++j)
            for (int k = 1; k <= kp; ++k) {
                const int r = at_(k, j, i);
                const int iw = i == 1 ? ip : i - 1, ie = i == ip ? 1 : i + 1;
                const int js = j == 1 ? jp : j - 1, jn = j == jp ? 1 : j + 1;
                const int kd = k == 1 ? 1 : k - 1, ku = k == kp ? kp : k + 1;
                const double ap = (((((-f[0][i]) - f[1][i]) - f[2][j]) - f[3][j]) - f[4][k]) -
                                  f[5][k];
                A(r, at_(k, j, iw)) += f[1][i];
                A(r, at_(k, j, ie)) += f[0][i];
                A(r, at_(k, js, i)) += f[3][j];
                A(r, at_(k, jn, i)) += f[2][j];
                A(r, at_(kd, j, i)) += f[5][k];
                A(r, at_(ku, j, i)) += f[4][k];
                A(r, r) += ap;
            }
    double scale = 0.0;
    for (int r = 0; r < n; ++r) scale = std::max(scale, std::fabs(A(r, r)));
    // the rank-one term at the operator's scale keeps the matrix well scaled
    for (int r = 0; r < n; ++r)
        for (int col = 0; col < n; ++col) A(r, col) += scale / n;
    for (int r = 0; r < n; ++r) A(r, n + r) = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A(r, col)) > std::fabs(A(piv, col))) piv = r;
        if (!(std::fabs(A(piv, col)) > 1e-12 * scale)) return {};
        if (piv != col)
            for (int q = 0; q < 2 * n; ++q) std::swap(A(piv, q), A(col, q));
        const double inv = 1.0 / A(col, col);
        for (int q = 0; q < 2 * n; ++q) A(col, q) *= inv;
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double m = A(r, col);
            if (m == 0.0) continue;
            for (int q = 0; q < 2 * n; ++q) A(r, q) -= m * A(col, q);
        }
    }
    std::vector<double> mt((size_t)n * n);
    for (int r = 0; r < n; ++r)
        for (int col = 0; col < n; ++col) mt[(size_t)col * n + r] = A(r, n + col);
    return mt;
}

int mg_prepare(dcm_ctx *c) {
    if (c->mg) return DCM_OK;
    const Geo &g = c->geo;
    std::vector<int> d;
    if (!mg_levels(g.ip, g.jp, g.kp, d)) {
        set_error("multigrid preconditioner: the grid does not coarsen (some of ip, jp, kp "
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
    // coarse levels: along a halved axis the mean of the two children's
    // coefficients over 4, along an axis that keeps its size the fine ones
    std::vector<double> h;
    std::vector<size_t> off(6 * H->nlev, 0);
    for (int l = 1; l < H->nlev; ++l) {
        const int n3[3] = {d[3 * l], d[3 * l + 1], d[3 * l + 2]};
        for (int a = 0; a < 6; ++a) {
            const int n = n3[a / 2];
            const bool halved = d[3 * (l - 1) + a / 2] != n;
            std::vector<double> nx(n + 2, 0.0);
            for (int m = 1; m <= n; ++m)
                nx[m] = halved ? 0.125 * (f[a][2 * m - 1] + f[a][2 * m]) : f[a][m];
            off[6 * l + a] = h.size();
            for (double v : nx) h.push_back(v);
            f[a] = nx;
        }
    }
    const int lc = H->nlev - 1;
    const std::vector<double> minv = coarse_inverse_t(d[3 * lc], d[3 * lc + 1], d[3 * lc + 2], f);
    int st = build_levels<float>(H->f32, H, d, h, off, minv, c->met32.ae, c->met32.aw,
                                 c->met32.an, c->met32.as, c->met32.au, c->met32.ad);
    if (st == DCM_OK)
        st = build_levels<double>(H->f64, H, d, h, off, minv, c->met.ae, c->met.aw, c->met.an,
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
    const R w = (R)kMgOmega;
    // levels from `top` down are the tail's (the coarsest alone without it)
    const bool tail = c->mg_tail && V.tail > 0;
    const int top = tail ? V.tail : L - 1;
    // the per-cell transfers between levels l and l+1
    auto restrict_cells = [&](int l, const R *b, const R *x) {
        const Grid3 GC = grid3(V.lv[l + 1]);
        if (halves(V.lv[l], V.lv[l + 1]))
            hipLaunchKernelGGL((k_mg_restrict<R, false>), GC.grid, GC.block, 0, c->stream, b, x,
                               V.b[l + 1], V.lv[l], V.lv[l + 1], sc);
        else
            hipLaunchKernelGGL((k_mg_restrict<R, true>), GC.grid, GC.block, 0, c->stream, b, x,
                               V.b[l + 1], V.lv[l], V.lv[l + 1], sc);
    };
    auto prolong_cells = [&](int l, const R *xc, R *x) {
        const Grid3 GF = grid3(V.lv[l]);
        if (halves(V.lv[l], V.lv[l + 1]))
            hipLaunchKernelGGL((k_mg_prolong<R, false>), GF.grid, GF.block, 0, c->stream, xc, x,
                               V.lv[l], V.lv[l + 1], sc);
        else
            hipLaunchKernelGGL((k_mg_prolong<R, true>), GF.grid, GF.block, 0, c->stream, xc, x,
                               V.lv[l], V.lv[l + 1], sc);
    };
    // the marching finest transfers (mode 1, a halving first level); per-pass
    // mode keeps the one-thread-per-cell kernels
    const bool march = c->mg_tail && halves(V.lv[0], V.lv[1]);
    if (march) {
        Grid3 GM = grid3(V.lv[1]);
        GM.grid.z = (unsigned)((V.lv[1].ip + kProPlanes / 2 - 1) / (kProPlanes / 2));
        hipLaunchKernelGGL(k_mg_restrict_march<R>, GM.grid, GM.block, 0, c->stream, b0,
                           (const R *)x0, V.b[1], V.lv[0], V.lv[1], sc);
    } else {
        restrict_cells(0, b0, (const R *)x0);
    }
    // a level's legs as one kernel each where it has a tile (mode 1); the
    // up-leg then leaves the level's result in t
    auto fused = [&](int l) { return c->mg_tail && V.tile[l].TI > 0; };
    for (int l = 1; l < top; ++l) {
        const Lv<R> &F = V.lv[l], &C = V.lv[l + 1];
        if (fused(l)) {
            const Tile2 T = V.tile[l];
            const dim3 blk((unsigned)F.kp, (unsigned)(256 / F.kp));
            const unsigned nb = (unsigned)((F.ip / T.TI) * (F.jp / T.TJ));
            hipLaunchKernelGGL(k_mg_down<R>, dim3(nb), blk, V.tile_lds[l], c->stream,
                               (const R *)V.b[l], V.x[l], V.b[l + 1], F, C, T, w, sc);
            continue;
        }
        const Grid3 GF = grid3(F);
        hipLaunchKernelGGL(k_mg_smooth<R>, GF.grid, GF.block, 0, c->stream,
                           (const R *)V.b[l], (const R *)nullptr, V.t[l], F, w, sc);
        hipLaunchKernelGGL(k_mg_smooth<R>, GF.grid, GF.block, 0, c->stream,
                           (const R *)V.b[l], (const R *)V.t[l], V.x[l], F, w, sc);
        restrict_cells(l, (const R *)V.b[l], (const R *)V.x[l]);
    }
    if (tail)
        hipLaunchKernelGGL(k_mg_tail<R>, dim3(1), dim3((unsigned)V.tail_threads), V.tail_lds,
                           c->stream, V.targs, kCoarseSweeps, w, sc);
    else
        hipLaunchKernelGGL(k_mg_coarse<R>, dim3(1), dim3(kCoarseMax), 0, c->stream,
                           (const R *)V.b[L - 1], V.x[L - 1], V.lv[L - 1], kCoarseSweeps, w, sc,
                           (const R *)V.Mt);
    const R *below = V.x[top];  // the result of the level below
    for (int l = top - 1; l >= 1; --l) {
        const Lv<R> &F = V.lv[l], &C = V.lv[l + 1];
        if (fused(l)) {
            const Tile2 T = V.tile[l];
            const dim3 blk((unsigned)F.kp, (unsigned)(256 / F.kp));
            const unsigned nb = (unsigned)((F.ip / T.TI) * (F.jp / T.TJ));
            hipLaunchKernelGGL(k_mg_up<R>, dim3(nb), blk, V.tile_lds[l], c->stream, below,
                               (const R *)V.x[l], (const R *)V.b[l], V.t[l], F, C, T, w, sc);
            below = V.t[l];
            continue;
        }
        const Grid3 GF = grid3(F);
        prolong_cells(l, below, V.x[l]);
        hipLaunchKernelGGL(k_mg_smooth<R>, GF.grid, GF.block, 0, c->stream,
                           (const R *)V.b[l], (const R *)V.x[l], V.t[l], F, w, sc);
        hipLaunchKernelGGL(k_mg_smooth<R>, GF.grid, GF.block, 0, c->stream,
                           (const R *)V.b[l], (const R *)V.t[l], V.x[l], F, w, sc);
        below = V.x[l];
    }
    if (march) {
        Grid3 G = grid3(V.lv[0]);
        G.grid.z = (unsigned)((V.lv[0].ip + kProPlanes - 1) / kProPlanes);
        hipLaunchKernelGGL(k_mg_prolong_march<R>, G.grid, G.block, 0, c->stream, below, x0,
                           V.lv[0], V.lv[1], sc);
    } else {
        prolong_cells(0, below, x0);
    }
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
        c->mg_tail = on != 2;
    } else {
        c->mg_on = false;
    }
    return DCM_OK;
}

}  // extern "C"
