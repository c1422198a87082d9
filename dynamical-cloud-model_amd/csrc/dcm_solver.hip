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
// solver's work vectors, the iterate included, are compact (kp*jp*ip).
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
// scalars live on the device and every kernel tests the device-side
// convergence flag, so the host only polls every few iterations.
//
// Stencil kernels march: a 256-thread block owns a kt x jt tile of (k,j)
// columns (kt = 64 lanes along the contiguous k, or 32 for shallow grids) and
// walks `ic` consecutive i planes, keeping the i-1, i, i+1 values of its own
// cells in registers; j and k neighbours come through L1/L2.  Offsets are
// 32-bit and blocks are renumbered so that each XCD (block id mod 8) walks a
// contiguous range of tiles, which keeps the j-neighbour columns in its own
// L2.
//
// Mixed precision (dcm_bicgstab_mixed): the same kernels with FP32 vectors
// and FP64 dot products solve A d = r for the FP64 residual r = b - A x; the
// outer loop adds d to x in FP64 and recomputes the true FP64 residual until
// ||b-Ax||/max(||b||,0.01) < tol (iterative refinement).  The operator is
// singular (periodic + Neumann), so each inner right-hand side has its mean
// removed.
#include <cfloat>
#include <cstddef>
#include <cmath>
#include <algorithm>
#include <functional>
#include <mutex>

#include <cstring>

#include "dcm_device.h"

namespace dcm {

struct CGeo {
    int ip, jp, kp;
    long long n;
};

__device__ __forceinline__ long long cidx(const CGeo &g, int k, int j, int i) {
    return ((long long)(i - 1) * g.jp + (j - 1)) * g.kp + (k - 1);
}

// Global-memory pointers.  A pointer read from device memory (a member
// table) is a generic pointer to the compiler, which then issues FLAT
// accesses; the kernel bodies take their buffers as global pointers, which
// gives GLOBAL loads and stores whichever way the pointer arrived.
template <typename T> using gp = __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ gp<T> glob(T *p) { return (gp<T>)p; }

// the six 1-D coefficient arrays of the operator at precision R
template <typename R> struct Coef {
    gp<const R> __restrict__ ae;
    gp<const R> __restrict__ aw;
    gp<const R> __restrict__ an;
    gp<const R> __restrict__ as;
    gp<const R> __restrict__ au;
    gp<const R> __restrict__ ad;
    template <typename M> __device__ explicit Coef(const M &m)
        : ae(glob(m.ae)), aw(glob(m.aw)), an(glob(m.an)), as(glob(m.as)), au(glob(m.au)),
          ad(glob(m.ad)) {}
    __device__ Coef(const Metrics &m64, const Metrics32 &m32);
};
template <> __device__ inline Coef<double>::Coef(const Metrics &m, const Metrics32 &)
    : Coef(m) {}
template <> __device__ inline Coef<float>::Coef(const Metrics &, const Metrics32 &m)
    : Coef(m) {}

struct CoefPack {
    Metrics m64;
    Metrics32 m32;
};

// the iteration has stopped for this system (converged or broken down):
// breakdown and done are adjacent ints, tested with one 8-byte load
static_assert(offsetof(SolverScalars, done) == offsetof(SolverScalars, breakdown) + 4 &&
                  offsetof(SolverScalars, breakdown) % 8 == 0,
              "SolverScalars: breakdown and done must share an aligned 8-byte word");
__device__ __forceinline__ bool halted(gp<const SolverScalars> sc) {
    return *(gp<const unsigned long long>)&sc->breakdown != 0ull;
}
__device__ __forceinline__ bool halted(const SolverScalars *sc) { return halted(glob(sc)); }

// a_p of set_mat_a (dynamics.f90:90-92)
template <typename R>
__device__ __forceinline__ R coef_ap(const Coef<R> &C, int k, int j, int i) {
    return (((((-C.ae[i]) - C.aw[i]) - C.an[j]) - C.as[j]) - C.au[k]) - C.ad[k];
}
// the same sum with its first two terms (the i part) precomputed
template <typename R>
__device__ __forceinline__ R coef_ap_i(const Coef<R> &C, R api, int k, int j) {
    return (((api - C.an[j]) - C.as[j]) - C.au[k]) - C.ad[k];
}

// x / a_p.  FP64: the correctly rounded quotient of the reference.  FP32
// (the inner solve of the mixed path): a multiply by the hardware reciprocal
// (within about 1 ulp), which keeps the Jacobi preconditioner a fixed linear
// map at a fraction of the cost of an IEEE division.
template <typename R> __device__ __forceinline__ R div_ap(R x, R ap);
template <> __device__ __forceinline__ double div_ap<double>(double x, double ap) {
    return x / ap;
}
template <> __device__ __forceinline__ float div_ap<float>(float x, float ap) {
    return x * __builtin_amdgcn_rcpf(ap);
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

// the stencil sum of mat_ax (dynamics.f90:216-224) with the coefficients of
// the centre cell
template <typename R>
__device__ __forceinline__ R stencil(R aw, R ae, R as, R an, R ad, R au, R ap, R xw, R xe, R xs,
                                     R xn, R xd, R xu, R xc) {
    return ((((((aw * xw + ae * xe) + as * xs) + an * xn) + ad * xd) + au * xu) + ap * xc) + R(0);
}
template <typename R>
__device__ __forceinline__ R stencil(const Coef<R> &C, int k, int j, int i, R ap, R xw, R xe,
                                     R xs, R xn, R xd, R xu, R xc) {
    return stencil<R>(C.aw[i], C.ae[i], C.as[j], C.an[j], C.ad[k], C.au[k], ap, xw, xe, xs, xn,
                      xd, xu, xc);
}

// ---------------------------------------------------------------------------
// marching geometry
// ---------------------------------------------------------------------------
struct MG {
    int ip, jp, kp;
    int sj, si;            // strides of j and i in a compact vector
    int kt, jt;            // tile lanes along k and rows along j (kt*jt = 256)
    int nkb, njb, nic, ic; // tiles along k, j; i chunks and their length
    int nblocks;
};

static MG march_geo(const Geo &g, int ic) {
    MG m;
    m.ip = g.ip;
    m.jp = g.jp;
    m.kp = g.kp;
    m.sj = g.kp;
    m.si = g.kp * g.jp;
    // 64 lanes along k unless 32 leaves fewer idle lanes (kp = 96: three
    // full tiles of 32 instead of a full and a half-empty tile of 64)
    const int w64 = (64 - g.kp % 64) % 64, w32 = (32 - g.kp % 32) % 32;
    m.kt = g.kp > 32 && w64 <= w32 ? 64 : 32;
    m.jt = kBlock / m.kt;
    m.nkb = (g.kp + m.kt - 1) / m.kt;
    m.njb = (g.jp + m.jt - 1) / m.jt;
    m.ic = ic;
    m.nic = (g.ip + ic - 1) / ic;
    m.nblocks = m.nkb * m.njb * m.nic;
    return m;
}
static unsigned march_grid(const MG &m) { return (unsigned)(((m.nblocks + 7) / 8) * 8); }

// Tile of block b of a one-system launch: blocks are renumbered so that XCD
// x = b mod 8 owns the contiguous tile range [x*per, (x+1)*per).  -1 for the
// padding blocks.
__device__ __forceinline__ int march_tile_id(const MG &m, int b) {
    const int per = (m.nblocks + 7) >> 3;
    const int t = (b & 7) * per + (b >> 3);
    return t < m.nblocks ? t : -1;
}
// the cells of tile `tile` (kb fastest, then jb, then the i chunk); false
// for a padding block (tile < 0)
struct Tile {
    int k, j, i0, i1, tile;
    bool act;
};
__device__ __forceinline__ bool march_tile(const MG &m, int tile, Tile &T) {
    if (tile < 0) return false;
    T.tile = tile;
    const int kb = T.tile % m.nkb;
    const int r = T.tile / m.nkb;
    const int jb = r % m.njb;
    const int icb = r / m.njb;
    T.k = kb * m.kt + (int)(threadIdx.x % m.kt) + 1;
    T.j = jb * m.jt + (int)(threadIdx.x / m.kt) + 1;
    T.i0 = icb * m.ic + 1;
    T.i1 = T.i0 + m.ic - 1;
    if (T.i1 > m.ip) T.i1 = m.ip;
    T.act = T.k <= m.kp && T.j <= m.jp;
    return true;
}

// ---------------------------------------------------------------------------
// Solve members.  The iteration kernels serve any number of independent
// systems (one domain, or the members of an ensemble) in one launch: a block
// finds its member in a table and works on that member's vectors with that
// member's geometry, coefficients and scalars.  The marching table gives
// each block a (member, tile) pair: the members' tiles are concatenated in
// member order and the list is split into 8 contiguous ranges, one per XCD,
// as a single domain's tiles are, so a small member stays within one XCD's
// L2 and a large one is split into i-chunk ranges.  Partial sums are stored
// by tile, so a member's sums are those of the same system solved alone.
// The update grid is split into per-member block ranges; the one-block
// reductions run one block per member.  With one member no table is needed.
// ---------------------------------------------------------------------------
struct IterMember {
    MG m;
    CoefPack cp;
    SolverScalars *sc;
    double *part;     // partial sums: 2 slots of max(nblocks, nbu)
    void *V[10];      // r, cr, p, v1, t, s, cs, wa, wb, x (compact, precision R)
    int n, nbu;       // cells; update-kernel blocks
    int ublock0;      // first update block
};
enum { VR = 0, VCR, VP, VV1, VT, VS, VCS, VWA, VWB, VX };

// marching table entry: member << kTileBits | tile; kNoTile for padding
constexpr int kTileBits = 20;
constexpr unsigned kNoTile = 0xffffffffu;
constexpr int kMaxMembers = 1 << (32 - kTileBits);

struct IterArgs {
    const IterMember *mem;
    const unsigned *mmap;  // marching block -> member and tile (batched launches)
    const int *umap;       // update block -> member (batched launches)
};

// the member of a marching block and its tile (-1: a padding block)
__device__ __forceinline__ const IterMember &member_m(const IterArgs &A, int &tile) {
    const unsigned e = A.mmap[blockIdx.x];
    tile = e == kNoTile ? -1 : (int)(e & ((1u << kTileBits) - 1));
    return A.mem[e == kNoTile ? 0 : (e >> kTileBits)];
}
__device__ __forceinline__ const IterMember &member_u(const IterArgs &A, int &lb) {
    const IterMember &M = A.mem[A.umap[blockIdx.x]];
    lb = (int)blockIdx.x - M.ublock0;
    return M;
}

// column offsets of a tile cell: base (i = 1) and the j/k neighbour deltas
struct Cols {
    int base, ds, dn, dd, du;
};
__device__ __forceinline__ Cols cols(const MG &m, int k, int j) {
    Cols c;
    c.base = (j - 1) * m.sj + (k - 1);
    c.ds = (j == 1 ? m.jp - 1 : -1) * m.sj;
    c.dn = (j == m.jp ? -(m.jp - 1) : 1) * m.sj;
    c.dd = k == 1 ? 0 : -1;
    c.du = k == m.kp ? 0 : 1;
    return c;
}

// ---------------------------------------------------------------------------
// initial residual (dynamics.f90:370-381): r = b - A x, cr = r, p = v1 = 0,
// x gathered into the compact iterate xc; partial sums of <b,b> and <r,r>.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_resid0(const double *__restrict__ x,
                                                const double *__restrict__ b, double *r,
                                                double *cr, double *p, double *v1, double *xc,
                                                Geo pg, CGeo g, CoefPack cp, double *partials,
                                                long long stride) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    const Coef<double> C(cp.m64, cp.m32);
    double s[2] = {0.0, 0.0};
    if (k <= g.kp && j <= g.jp) {
        const Nbr n = nbrs(k, j, i, g.kp, g.jp, g.ip);
        const double ap = coef_ap(C, k, j, i);
        const double xcv = x[pidx(pg, k, j, i)];
        const double ax =
            stencil<double>(C, k, j, i, ap, x[pidx(pg, k, j, n.iw)], x[pidx(pg, k, j, n.ie)],
                            x[pidx(pg, k, n.js, i)], x[pidx(pg, k, n.jn, i)],
                            x[pidx(pg, n.kd, j, i)], x[pidx(pg, n.ku, j, i)], xcv);
        const double bb = b[pidx(pg, k, j, i)];
        const double rr = bb - ax;
        const long long c = cidx(g, k, j, i);
        r[c] = rr;
        cr[c] = rr;
        p[c] = 0.0;
        v1[c] = 0.0;
        xc[c] = xcv;
        s[0] = bb * bb;
        s[1] = rr * rr;
    }
    block_sum_store<2>(s, partials, stride);
}

// FP64 true residual r = b - A x (kept, compact, when r64 is not null) and
// the partial sums of <r,r>, sum(r), <b,b>
// r = b - A x (FP64) with the partial sums of <r,r>, sum r and <b,b>.  A
// thread owns one (k,j) column and marches kResPlanes i planes with x of
// planes i-1, i, i+1 in registers (5 loads per cell instead of 8), 32-bit
// offsets; periodic in i and j, the Neumann copy in k.  Launch with
// resid_launch (the grid's z covers the i chunks).
constexpr int kResPlanes = 8;
__global__ __launch_bounds__(256) void k_resid_true(const double *__restrict__ x,
                                                    const double *__restrict__ b, double *r64,
                                                    Geo pg, CGeo g, CoefPack cp,
                                                    double *partials, long long stride,
                                                    const OuterState *gate = nullptr) {
    if (gate && !gate->go) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i0 = blockIdx.z * kResPlanes + 1;
    const Coef<double> C(cp.m64, cp.m32);
    double s[3] = {0.0, 0.0, 0.0};
    if (k <= g.kp && j <= g.jp) {
        const int i1 = min(i0 + kResPlanes - 1, g.ip);
        const int SI = (int)pg.si, SJ = (int)pg.sj;
        const int js = j == 1 ? g.jp : j - 1, jn = j == g.jp ? 1 : j + 1;
        const int dS = (js - j) * SJ, dN = (jn - j) * SJ;
        const int dD = k == 1 ? 0 : -1, dU = k == g.kp ? 0 : 1;
        const double as = C.as[j], an = C.an[j], ad = C.ad[k], au = C.au[k];
        const int col = j * SJ + k;
        const int CS = g.kp * g.jp, CJ = g.kp;  // compact strides
        const int iw = i0 == 1 ? g.ip : i0 - 1;
        int ie = i0 == g.ip ? 1 : i0 + 1;
        int c = i0 * SI + col;
        double xw = x[iw * SI + col], xc = x[c], xe = x[ie * SI + col];
        double xs = x[c + dS], xn = x[c + dN], xd = x[c + dD], xu = x[c + dU], bb = b[c];
        for (int i = i0; i <= i1; ++i) {
            // the loads of plane i+1 in flight while plane i is computed
            double nxe = 0.0, nxs = 0.0, nxn = 0.0, nxd = 0.0, nxu = 0.0, nbb = 0.0;
            const int ce = ie * SI + col;
            if (i < i1) {
                const int nie = ie == g.ip ? 1 : ie + 1;
                nxe = x[nie * SI + col];
                nxs = x[ce + dS];
                nxn = x[ce + dN];
                nxd = x[ce + dD];
                nxu = x[ce + dU];
                nbb = b[ce];
                ie = nie;
            }
            const double ap = coef_ap_i(C, (-C.ae[i]) - C.aw[i], k, j);
            const double ax =
                stencil<double>(C.aw[i], C.ae[i], as, an, ad, au, ap, xw, xe, xs, xn, xd, xu, xc);
            const double rr = bb - ax;
            if (r64) r64[(i - 1) * CS + (j - 1) * CJ + (k - 1)] = rr;
            s[0] += rr * rr;
            s[1] += rr;
            s[2] += bb * bb;
            xw = xc;
            xc = xe;
            xe = nxe;
            xs = nxs;
            xn = nxn;
            xd = nxd;
            xu = nxu;
            bb = nbb;
            c = ce;
        }
    }
    block_sum_store<3>(s, partials, stride);
}

// the launch of k_resid_true: interior columns, i in chunks of kResPlanes
inline Interior resid_launch(const Geo &g) {
    Interior L = interior_launch(g);
    L.grid.z = (unsigned)((g.ip + kResPlanes - 1) / kResPlanes);
    return L;
}

// ---------------------------------------------------------------------------
// sums in the reference's loop order (inn_prod, dynamics.f90:470-476: i
// outer, j, k inner, one running sum).  One thread; used when the context
// asks for sequential sums, so that a run is bit-identical to a serial CPU
// run of the same arithmetic.  out[0] = <a0,b0>, out[1] = <a1,b1> (a1 may be
// null).
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_seq_dots(IterArgs A, int a0i, int b0i, int a1i, int b1i) {
    const IterMember &M = A.mem[blockIdx.x];
    if (threadIdx.x != 0) return;
    const SolverScalars *sc = M.sc;
    if (halted(sc)) return;
    const R *a0 = (const R *)M.V[a0i], *b0 = (const R *)M.V[b0i];
    const R *a1 = a1i >= 0 ? (const R *)M.V[a1i] : nullptr;
    const R *b1 = b1i >= 0 ? (const R *)M.V[b1i] : nullptr;
    double s0 = 0.0, s1 = 0.0;
    for (long long q = 0; q < M.n; ++q) {
        s0 = s0 + (double)a0[q] * (double)b0[q];
        if (a1) s1 = s1 + (double)a1[q] * (double)b1[q];
    }
    M.part[0] = s0;
    M.part[1] = s1;
}

// <b,b> over the interior of a padded b and <r,r> of the compact r
__global__ void k_seq_resid0(const double *__restrict__ b, const double *__restrict__ r, Geo pg,
                             CGeo g, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double bb = 0.0, rr = 0.0;
    long long q = 0;
    for (int i = 1; i <= g.ip; ++i)
        for (int j = 1; j <= g.jp; ++j)
            for (int k = 1; k <= g.kp; ++k, ++q) {
                const double bv = b[pidx(pg, k, j, i)];
                bb = bb + bv * bv;
                rr = rr + r[q] * r[q];
            }
    out[0] = bb;
    out[1] = rr;
}

// <x,y> over the interior of padded arrays
__global__ void k_seq_dot_padded(const double *__restrict__ x, const double *__restrict__ y,
                                 Geo g, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 1; i <= g.ip; ++i)
        for (int j = 1; j <= g.jp; ++j)
            for (int k = 1; k <= g.kp; ++k) {
                const long long c = pidx(g, k, j, i);
                s = s + x[c] * y[c];
            }
    out[0] = s;
}

// ---------------------------------------------------------------------------
// scalar recurrences (one block; reduce the partial sums first)
// ---------------------------------------------------------------------------
// tiny = 1.e-16 is a default-real literal (dynamics.f90:321)
__device__ constexpr double kTinyOmega = (double)1.e-16f;
// inner (FP32) solves: iterations without a 10% gain that count as stagnation
constexpr int kStall = 40;

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
        sc->iters = 0;
        sc->guard = 0;
        sc->tol = tol;
        sc->nonfinite = !isfinite(sc->err);
        sc->done = (sc->err < tol) || sc->nonfinite;
    }
}

// scalar recurrences after each reduction (dynamics.f90:404, :412-414,
// :423-425); called by a one-block kernel or by the last block of the
// kernel that produced the partial sums
__device__ __forceinline__ void fin_v_apply(SolverScalars *sc, double nrm) {
    sc->nrm = nrm;
    sc->alf = sc->rho / nrm;  // :404
    if (sc->guard && !(nrm != 0.0 && isfinite(sc->alf))) sc->breakdown = 1;
}
__device__ __forceinline__ void fin_t_apply(SolverScalars *sc, double tt, double ts) {
    sc->iters += 1;
    sc->tt = tt;
    sc->ts = ts;
    sc->omg = ts / tt;  // :412
    if (fabs(sc->omg) < kTinyOmega) sc->breakdown = 1;  // :414
    if (sc->guard && !(tt != 0.0 && isfinite(sc->omg))) sc->breakdown = 1;
}
__device__ __forceinline__ void fin_r_apply(SolverScalars *sc, double rr, double rcr) {
    sc->dvalid = 1;  // the update of this iteration has run
    sc->rr = rr;
    sc->err = sqrt(rr) / sc->sc_err;  // :423
    sc->nrm = sc->rho;                // :425
    sc->rho = rcr;                    // next iteration's <r,cr> (:391)
    sc->bet = (sc->rho / sc->nrm) * (sc->alf / sc->omg);
    sc->nonfinite = !isfinite(sc->err);
    sc->done = (sc->err < sc->tol) || sc->nonfinite;
    if (sc->guard) {
        // an inner solve that has not improved its best residual by 10% for
        // kStall iterations has stagnated at its precision: stop and let the
        // FP64 refinement restart it from the true residual
        if (sc->err < 0.9 * sc->best_err) {
            sc->best_err = sc->err;
            sc->best_it = sc->iters;
        } else if (sc->iters - sc->best_it >= kStall) {
            sc->done = 1;
        }
    }
}

// nb: the member's partial count (its tiles, or 1 after sequential sums)
__device__ __forceinline__ int fin_count(const IterMember &M, int kind, int seq) {
    if (seq) return 1;
    return kind == 0 ? M.m.nblocks : M.nbu;
}
__global__ void k_fin_v1(const double *partials, int nb, SolverScalars *sc) {
    if (halted(sc)) return;
    const double nrm = reduce_slot_1024(partials, nb);
    if (threadIdx.x == 0) fin_v_apply(sc, nrm);
}
__global__ void k_fin_t1(const double *partials, int nb, SolverScalars *sc) {
    if (halted(sc)) return;
    const double tt = reduce_slot_1024(partials, nb);
    const double ts = reduce_slot_1024(partials + nb, nb);
    if (threadIdx.x == 0) fin_t_apply(sc, tt, ts);
}
__global__ void k_fin_r1(const double *partials, int nb, SolverScalars *sc) {
    if (halted(sc)) return;
    const double rr = reduce_slot_1024(partials, nb);
    const double rcr = reduce_slot_1024(partials + nb, nb);
    if (threadIdx.x == 0) fin_r_apply(sc, rr, rcr);
}
__global__ void k_fin_v(IterArgs A, int seq) {
    const IterMember &M = A.mem[blockIdx.x];
    SolverScalars *sc = M.sc;
    if (halted(sc)) return;
    const double nrm = reduce_slot_1024(M.part, fin_count(M, 0, seq));
    if (threadIdx.x == 0) fin_v_apply(sc, nrm);
}
__global__ void k_fin_t(IterArgs A, int seq) {
    const IterMember &M = A.mem[blockIdx.x];
    SolverScalars *sc = M.sc;
    if (halted(sc)) return;
    const int nb = fin_count(M, 0, seq);
    const double tt = reduce_slot_1024(M.part, nb);
    const double ts = reduce_slot_1024(M.part + nb, nb);
    if (threadIdx.x == 0) fin_t_apply(sc, tt, ts);
}
__global__ void k_fin_r(IterArgs A, int seq) {
    const IterMember &M = A.mem[blockIdx.x];
    SolverScalars *sc = M.sc;
    if (halted(sc)) return;
    const int nb = fin_count(M, 1, seq);
    const double rr = reduce_slot_1024(M.part, nb);
    const double rcr = reduce_slot_1024(M.part + nb, nb);
    if (threadIdx.x == 0) fin_r_apply(sc, rr, rcr);
}
// flags of every member for the host: done | breakdown << 1 | nonfinite << 2
__global__ void k_gather_flags(IterArgs A, int nmem, int *flags) {
    for (int q = threadIdx.x; q < nmem; q += blockDim.x) {
        const SolverScalars *sc = A.mem[q].sc;
        flags[q] = (sc->done ? 1 : 0) | (sc->breakdown ? 2 : 0) | (sc->nonfinite ? 4 : 0);
    }
}

// ---------------------------------------------------------------------------
// preconditioner sweeps (dynamics.f90:241-270), marching kernels
// ---------------------------------------------------------------------------
// MODE 0: t = r - bet*(omg*v1)     (dynamics.f90:394)
// MODE 1: t = r - alf*v1           (dynamics.f90:405, the "s" vector)
template <typename R, int MODE>
__device__ __forceinline__ R tval(gp<const R> __restrict__ r, gp<const R> __restrict__ v1, int q, R a,
                                  R b) {
    if (MODE == 0) return r[q] - a * (b * v1[q]);
    return r[q] - a * v1[q];
}

// the same from loaded values
template <typename R, int MODE> __device__ __forceinline__ R tv(R r, R v1, R a, R b) {
    if (MODE == 0) return r - a * (b * v1);
    return r - a * v1;
}

// First sweep, variant 0: fused in registers, s0 = t/a_p evaluated at the
// centre and the six neighbours (dynamics.f90:254-266).
// W: Chebyshev relaxation, s0 = w0 t/a_p and s1 = s0 + w1 (t - A s0)/a_p
// (otherwise the reference's w0 = w1 = 1, and the factors are not applied)
template <typename R, int MODE, bool W = false>
__device__ __forceinline__ void jac_first_reg_body(gp<const R> __restrict__ r,
                                                   gp<const R> __restrict__ v1,
                                                   gp<R> __restrict__ t_out, gp<R> __restrict__ s_out,
                                                   const MG &m, const CoefPack &cp,
                                                   gp<const SolverScalars> __restrict__ sc, int tile,
                                                   R w0 = R(1), R w1 = R(1)) {
    if (halted(sc)) return;
    Tile T;
    if (!march_tile(m, tile, T) || !T.act) return;
    const Coef<R> C(cp.m64, cp.m32);
    const R a = MODE == 0 ? (R)sc->bet : (R)sc->alf;
    const R bo = (R)sc->omg;
    // v1 = 0 in a solve's first iteration (dynamics.f90:380): not read then
    const bool vz = MODE == 0 && sc->iters == 0;
    auto V1 = [&](int idx) { return vz ? R(0) : (R)v1[idx]; };
    const int k = T.k, j = T.j;
    const Cols q = cols(m, k, j);
    const int js = j == 1 ? m.jp : j - 1, jn = j == m.jp ? 1 : j + 1;
    const int kd = k == 1 ? 1 : k - 1, ku = k == m.kp ? m.kp : k + 1;
    const R as = C.as[j], an = C.an[j], ad = C.ad[k], au = C.au[k];
    // On the floor (lid) row the k-1 (k+1) neighbour is the cell itself
    // (q.dd = 0, kd = k): its s0 is then s0c by the same operations, so the
    // neighbour is loaded through cc + q.dd without a branch.
    const int iw = T.i0 == 1 ? m.ip : T.i0 - 1;
    int cc = (T.i0 - 1) * m.si + q.base;
    int ie = T.i0 == m.ip ? 1 : T.i0 + 1;
    int ce = (ie - 1) * m.si + q.base;
    const int cw = (iw - 1) * m.si + q.base;
    auto rw = [&](R x) { return W ? w0 * x : x; };
    R s0w = rw(div_ap<R>(tv<R, MODE>(r[cw], V1(cw), a, bo), coef_ap(C, k, j, iw)));
    R tc = tv<R, MODE>(r[cc], V1(cc), a, bo);
    R s0c = rw(div_ap<R>(tc, coef_ap(C, k, j, T.i0)));
    // r, v1 at the j and k neighbours of plane i, and at the centre of i+1
    R rs = r[cc + q.ds], vs = V1(cc + q.ds), rn = r[cc + q.dn], vn = V1(cc + q.dn);
    R rd = r[cc + q.dd], vd = V1(cc + q.dd), ru = r[cc + q.du], vu = V1(cc + q.du);
    R re = r[ce], ve = V1(ce);
    for (int i = T.i0; i <= T.i1; ++i) {
        // loads for plane i+1 in flight while plane i is computed
        R nrs = R(0), nvs = R(0), nrn = R(0), nvn = R(0), nrd = R(0), nvd = R(0), nru = R(0),
          nvu = R(0), nre = R(0), nve = R(0);
        int nie = ie, nce = ce;
        if (i < T.i1) {
            nie = ie == m.ip ? 1 : ie + 1;
            nce = (nie - 1) * m.si + q.base;
            nrs = r[ce + q.ds];
            nvs = V1(ce + q.ds);
            nrn = r[ce + q.dn];
            nvn = V1(ce + q.dn);
            nrd = r[ce + q.dd];
            nvd = V1(ce + q.dd);
            nru = r[ce + q.du];
            nvu = V1(ce + q.du);
            nre = r[nce];
            nve = V1(nce);
        }
        const R te = tv<R, MODE>(re, ve, a, bo);
        const R s0e = rw(div_ap<R>(te, coef_ap(C, k, j, ie)));
        const R api = (-C.ae[i]) - C.aw[i];
        const R ap = coef_ap_i(C, api, k, j);
        const R s0s = rw(div_ap<R>(tv<R, MODE>(rs, vs, a, bo), coef_ap_i(C, api, k, js)));
        const R s0n = rw(div_ap<R>(tv<R, MODE>(rn, vn, a, bo), coef_ap_i(C, api, k, jn)));
        const R s0d = rw(div_ap<R>(tv<R, MODE>(rd, vd, a, bo), coef_ap_i(C, api, kd, j)));
        const R s0u = rw(div_ap<R>(tv<R, MODE>(ru, vu, a, bo), coef_ap_i(C, api, ku, j)));
        const R As = stencil<R>(C.aw[i], C.ae[i], as, an, ad, au, ap, s0w, s0e, s0s, s0n, s0d,
                                s0u, s0c);
        t_out[cc] = tc;
        const R d1 = div_ap<R>(tc - As, ap);
        s_out[cc] = s0c + (W ? w1 * d1 : d1);
        s0w = s0c;
        s0c = s0e;
        tc = te;
        rs = nrs;
        vs = nvs;
        rn = nrn;
        vn = nvn;
        rd = nrd;
        vd = nvd;
        ru = nru;
        vu = nvu;
        re = nre;
        ve = nve;
        cc = ce;
        ce = nce;
        ie = nie;
    }
}

template <typename R, int MODE, bool W = false>
__global__ __launch_bounds__(256) void k_jac_first_reg(const R *__restrict__ r,
                                                       const R *__restrict__ v1,
                                                       R *__restrict__ t_out,
                                                       R *__restrict__ s_out, MG m, CoefPack cp,
                                                       const SolverScalars *__restrict__ sc,
                                                       R w0 = R(1), R w1 = R(1)) {
    jac_first_reg_body<R, MODE, W>(glob(r), glob(v1), glob(t_out), glob(s_out), m, cp, glob(sc),
                                   march_tile_id(m, (int)blockIdx.x), w0, w1);
}
template <typename R, int MODE, bool W = false>
__global__ __launch_bounds__(256) void k_jac_first_reg_b(IterArgs A, int ri, int vi, int ti,
                                                         int oi, R w0 = R(1), R w1 = R(1)) {
    int tile;
    const IterMember &M = member_m(A, tile);
    jac_first_reg_body<R, MODE, W>(glob((const R *)M.V[ri]), glob((const R *)M.V[vi]),
                                   glob((R *)M.V[ti]), glob((R *)M.V[oi]), M.m, M.cp,
                                   glob((const SolverScalars *)M.sc), tile, w0, w1);
}

// First sweep, variant 2: a streaming pass t = r - bet*(omg*v1) (or
// r - alf*v1) with s0 = t/a_p, then the plain sweep k_jac on (s0, t).
template <typename R, int MODE>
__device__ __forceinline__ void jac_prep_body(gp<const R> __restrict__ r, gp<const R> __restrict__ v1,
                                              gp<R> __restrict__ t_out, gp<R> __restrict__ s0_out,
                                              const MG &m, const CoefPack &cp,
                                              gp<const SolverScalars> __restrict__ sc, int tile) {
    if (halted(sc)) return;
    Tile T;
    if (!march_tile(m, tile, T) || !T.act) return;
    const Coef<R> C(cp.m64, cp.m32);
    const R a = MODE == 0 ? (R)sc->bet : (R)sc->alf;
    const R bo = (R)sc->omg;
    // v1 = 0 in a solve's first iteration (dynamics.f90:380): not read then
    const bool vz = MODE == 0 && sc->iters == 0;
    const int k = T.k, j = T.j;
    const Cols q = cols(m, k, j);
    int cc = (T.i0 - 1) * m.si + q.base;
    for (int i = T.i0; i <= T.i1; ++i, cc += m.si) {
        const R tc = vz ? tv<R, MODE>(r[cc], R(0), a, bo) : tval<R, MODE>(r, v1, cc, a, bo);
        t_out[cc] = tc;
        s0_out[cc] = div_ap<R>(tc, coef_ap_i(C, (-C.ae[i]) - C.aw[i], k, j));
    }
}

template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_jac_prep(const R *__restrict__ r,
                                                  const R *__restrict__ v1,
                                                  R *__restrict__ t_out, R *__restrict__ s0_out,
                                                  MG m, CoefPack cp,
                                                  const SolverScalars *__restrict__ sc) {
    jac_prep_body<R, MODE>(glob(r), glob(v1), glob(t_out), glob(s0_out), m, cp, glob(sc),
                           march_tile_id(m, (int)blockIdx.x));
}
template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_jac_prep_b(IterArgs A, int ri, int vi, int ti, int oi) {
    int tile;
    const IterMember &M = member_m(A, tile);
    jac_prep_body<R, MODE>(glob((const R *)M.V[ri]), glob((const R *)M.V[vi]),
                           glob((R *)M.V[ti]), glob((R *)M.V[oi]), M.m, M.cp,
                           glob((const SolverScalars *)M.sc), tile);
}

// s_out = s + (t - A s)/a_p; with P_UPDATE: p = s_out + bet*p instead.  The
// loads of plane i+1 are issued while plane i is computed.
template <typename R, bool P_UPDATE, bool W = false>
__device__ __forceinline__ void jac_body(gp<const R> __restrict__ s, gp<const R> __restrict__ t,
                                         gp<R> __restrict__ s_out, const MG &m, const CoefPack &cp,
                                         gp<const SolverScalars> __restrict__ sc, int tile,
                                         R w = R(1)) {
    if (halted(sc)) return;
    Tile T;
    if (!march_tile(m, tile, T) || !T.act) return;
    const Coef<R> C(cp.m64, cp.m32);
    const int k = T.k, j = T.j;
    const Cols q = cols(m, k, j);
    const R as = C.as[j], an = C.an[j], ad = C.ad[k], au = C.au[k];
    const R bet = P_UPDATE ? (R)sc->bet : R(0);
    // p = 0 in a solve's first iteration (dynamics.f90:379): not read then
    const bool pread = P_UPDATE && sc->iters != 0;
    const int iw = T.i0 == 1 ? m.ip : T.i0 - 1;
    int cc = (T.i0 - 1) * m.si + q.base;
    int ie = T.i0 == m.ip ? 1 : T.i0 + 1;
    int ce = (ie - 1) * m.si + q.base;
    R xw = s[(iw - 1) * m.si + q.base];
    R xc = s[cc], xe = s[ce];
    R xs = s[cc + q.ds], xn = s[cc + q.dn], xd = s[cc + q.dd], xu = s[cc + q.du];
    R tc = t[cc];
    R po = pread ? s_out[cc] : R(0);
    for (int i = T.i0; i <= T.i1; ++i) {
        R nxe = R(0), nxs = R(0), nxn = R(0), nxd = R(0), nxu = R(0), ntc = R(0), npo = R(0);
        int nie = ie, nce = ce;
        if (i < T.i1) {
            nie = ie == m.ip ? 1 : ie + 1;
            nce = (nie - 1) * m.si + q.base;
            nxe = s[nce];
            nxs = s[ce + q.ds];
            nxn = s[ce + q.dn];
            nxd = s[ce + q.dd];
            nxu = s[ce + q.du];
            ntc = t[ce];
            if (pread) npo = s_out[ce];
        }
        const R ap = coef_ap_i(C, (-C.ae[i]) - C.aw[i], k, j);
        const R As = stencil<R>(C.aw[i], C.ae[i], as, an, ad, au, ap, xw, xe, xs, xn, xd, xu, xc);
        const R d = div_ap<R>(tc - As, ap);
        const R snew = xc + (W ? w * d : d);
        if (P_UPDATE)
            s_out[cc] = snew + bet * po;  // p = s + bet*p (dynamics.f90:399)
        else
            s_out[cc] = snew;
        xw = xc;
        xc = xe;
        xe = nxe;
        xs = nxs;
        xn = nxn;
        xd = nxd;
        xu = nxu;
        tc = ntc;
        po = npo;
        cc = ce;
        ce = nce;
        ie = nie;
    }
}

template <typename R, bool P_UPDATE, bool W = false>
__global__ __launch_bounds__(256) void k_jac(const R *__restrict__ s, const R *__restrict__ t,
                                             R *__restrict__ s_out, MG m, CoefPack cp,
                                             const SolverScalars *__restrict__ sc, R w = R(1)) {
    jac_body<R, P_UPDATE, W>(glob(s), glob(t), glob(s_out), m, cp, glob(sc),
                             march_tile_id(m, (int)blockIdx.x), w);
}
template <typename R, bool P_UPDATE, bool W = false>
__global__ __launch_bounds__(256) void k_jac_b(IterArgs A, int si_, int ti, int oi, R w = R(1)) {
    int tile;
    const IterMember &M = member_m(A, tile);
    jac_body<R, P_UPDATE, W>(glob((const R *)M.V[si_]), glob((const R *)M.V[ti]),
                             glob((R *)M.V[oi]), M.m, M.cp,
                             glob((const SolverScalars *)M.sc), tile, w);
}

// y = A x; partial sums of <a,y> (and <y,y> first when TWO).  Loads of plane
// i+1 in flight while plane i is computed.  (A last-arriving-block reduction
// that applies the scalar recurrence in place was measured 1.6x slower
// overall: every block's device-scope release fence writes back its XCD's
// L2.  One-block kernels reduce instead.)
template <typename R, bool TWO, bool STORE = true>
__device__ __forceinline__ void matvec_dot_body(gp<const R> __restrict__ x, gp<const R> __restrict__ a,
                                                gp<R> __restrict__ y, const MG &m,
                                                const CoefPack &cp,
                                                gp<const SolverScalars> __restrict__ sc,
                                                double *partials, int tile) {
    if (halted(sc)) return;
    const int stride = m.nblocks;
    Tile T;
    if (!march_tile(m, tile, T)) return;
    double acc[2] = {0.0, 0.0};
    if (T.act) {
        const Coef<R> C(cp.m64, cp.m32);
        const int k = T.k, j = T.j;
        const Cols q = cols(m, k, j);
        const R as = C.as[j], an = C.an[j], ad = C.ad[k], au = C.au[k];
        const int iw = T.i0 == 1 ? m.ip : T.i0 - 1;
        int cc = (T.i0 - 1) * m.si + q.base;
        int ie = T.i0 == m.ip ? 1 : T.i0 + 1;
        int ce = (ie - 1) * m.si + q.base;
        R xw = x[(iw - 1) * m.si + q.base];
        R xc = x[cc], xe = x[ce];
        R xs = x[cc + q.ds], xn = x[cc + q.dn], xd = x[cc + q.dd], xu = x[cc + q.du];
        R ac = a[cc];
        for (int i = T.i0; i <= T.i1; ++i) {
            R nxe = R(0), nxs = R(0), nxn = R(0), nxd = R(0), nxu = R(0), nac = R(0);
            int nie = ie, nce = ce;
            if (i < T.i1) {
                nie = ie == m.ip ? 1 : ie + 1;
                nce = (nie - 1) * m.si + q.base;
                nxe = x[nce];
                nxs = x[ce + q.ds];
                nxn = x[ce + q.dn];
                nxd = x[ce + q.dd];
                nxu = x[ce + q.du];
                nac = a[ce];
            }
            const R ap = coef_ap_i(C, (-C.ae[i]) - C.aw[i], k, j);
            const R yv = stencil<R>(C.aw[i], C.ae[i], as, an, ad, au, ap, xw, xe, xs, xn, xd, xu,
                                    xc);
            if (STORE) y[cc] = yv;
            if (TWO) {
                acc[0] += (double)yv * (double)yv;  // tt = <t,t>
                acc[1] += (double)yv * (double)ac;  // ts = <t,s>
            } else {
                acc[0] += (double)ac * (double)yv;  // nrm = <cr,v1>
            }
            xw = xc;
            xc = xe;
            xe = nxe;
            xs = nxs;
            xn = nxn;
            xd = nxd;
            xu = nxu;
            ac = nac;
            cc = ce;
            ce = nce;
            ie = nie;
        }
    }
    block_sum_store_at<2>(acc, partials, stride, T.tile);
}

template <typename R, bool TWO, bool STORE = true>
__global__ __launch_bounds__(256) void k_matvec_dot(const R *__restrict__ x,
                                                    const R *__restrict__ a, R *__restrict__ y,
                                                    MG m, CoefPack cp,
                                                    const SolverScalars *__restrict__ sc,
                                                    double *partials) {
    matvec_dot_body<R, TWO, STORE>(glob(x), glob(a), glob(y), m, cp, glob(sc), partials,
                                   march_tile_id(m, (int)blockIdx.x));
}
template <typename R, bool TWO>
__global__ __launch_bounds__(256) void k_matvec_dot_b(IterArgs A, int xi, int ai, int yi) {
    int tile;
    const IterMember &M = member_m(A, tile);
    matvec_dot_body<R, TWO>(glob((const R *)M.V[xi]), glob((const R *)M.V[ai]),
                            glob((R *)M.V[yi]), M.m, M.cp,
                            glob((const SolverScalars *)M.sc), M.part, tile);
}

// x = x + alf*p + omg*cs; r = s - omg*t; partial sums of <r,r>, <r,cr>
// (dynamics.f90:419-423), all compact vectors, grid-stride
template <typename R>
__device__ __forceinline__ void update_body(gp<R> __restrict__ x, gp<const R> __restrict__ p,
                                            gp<const R> __restrict__ cs, gp<const R> __restrict__ s,
                                            gp<const R> __restrict__ t, gp<const R> __restrict__ cr,
                                            gp<R> __restrict__ r, int n, int nbu,
                                            gp<const SolverScalars> __restrict__ sc,
                                            double *partials, int lb) {
    if (halted(sc)) return;
    const int stride = nbu;
    const R alf = (R)sc->alf, omg = (R)sc->omg;
    // the correction of a mixed inner solve (guard set) starts at zero: not
    // read in its first iteration (iters counts it from the <t,t> step on)
    const bool xz = sc->guard && sc->iters == 1;
    double acc[2] = {0.0, 0.0};
    for (int q = lb * kBlock + threadIdx.x; q < n; q += nbu * kBlock) {
        x[q] = ((xz ? R(0) : x[q]) + alf * p[q]) + omg * cs[q];
        const R rv = s[q] - omg * t[q];
        r[q] = rv;
        acc[0] += (double)rv * (double)rv;
        acc[1] += (double)rv * (double)cr[q];
    }
    block_sum_store_at<2>(acc, partials, stride, lb);
}
template <typename R>
__global__ __launch_bounds__(256) void k_update(R *__restrict__ x, const R *__restrict__ p,
                                                const R *__restrict__ cs,
                                                const R *__restrict__ s,
                                                const R *__restrict__ t,
                                                const R *__restrict__ cr, R *__restrict__ r,
                                                int n, int nbu,
                                                const SolverScalars *__restrict__ sc,
                                                double *partials) {
    update_body<R>(glob(x), glob(p), glob(cs), glob(s), glob(t), glob(cr), glob(r), n, nbu,
                   glob(sc), partials, (int)blockIdx.x);
}
// The update of a single system with t = A cs recomputed instead of read
// (k_matvec_dot<R, true, false> computed <t,t> and <t,s> without storing t):
// marching tiles like the matvec, the same stencil expression, so the same t;
// x = x + alf*p + omg*cs, r = s - omg*t, partial sums of <r,r> and <r,cr>
// per tile (dynamics.f90:409, 419-423)
template <typename R>
__global__ __launch_bounds__(256) void k_update_march(R *__restrict__ x, const R *__restrict__ p,
                                                      const R *__restrict__ cs,
                                                      const R *__restrict__ s,
                                                      const R *__restrict__ cr,
                                                      R *__restrict__ r, MG m, CoefPack cp,
                                                      const SolverScalars *__restrict__ sc,
                                                      double *partials) {
    if (halted(sc)) return;
    Tile T;
    if (!march_tile(m, march_tile_id(m, (int)blockIdx.x), T)) return;
    double acc[2] = {0.0, 0.0};
    if (T.act) {
        const R alf = (R)sc->alf, omg = (R)sc->omg;
        const bool xz = sc->guard && sc->iters == 1;  // as update_body
        const Coef<R> C(cp.m64, cp.m32);
        const int k = T.k, j = T.j;
        const Cols q = cols(m, k, j);
        const R as = C.as[j], an = C.an[j], ad = C.ad[k], au = C.au[k];
        const int iw = T.i0 == 1 ? m.ip : T.i0 - 1;
        int cc = (T.i0 - 1) * m.si + q.base;
        int ie = T.i0 == m.ip ? 1 : T.i0 + 1;
        int ce = (ie - 1) * m.si + q.base;
        R xw = cs[(iw - 1) * m.si + q.base];
        R xc = cs[cc], xe = cs[ce];
        R xs = cs[cc + q.ds], xn = cs[cc + q.dn], xd = cs[cc + q.dd], xu = cs[cc + q.du];
        for (int i = T.i0; i <= T.i1; ++i) {
            R nxe = R(0), nxs = R(0), nxn = R(0), nxd = R(0), nxu = R(0);
            int nie = ie, nce = ce;
            if (i < T.i1) {
                nie = ie == m.ip ? 1 : ie + 1;
                nce = (nie - 1) * m.si + q.base;
                nxe = cs[nce];
                nxs = cs[ce + q.ds];
                nxn = cs[ce + q.dn];
                nxd = cs[ce + q.dd];
                nxu = cs[ce + q.du];
            }
            const R ap = coef_ap_i(C, (-C.ae[i]) - C.aw[i], k, j);
            const R tv = stencil<R>(C.aw[i], C.ae[i], as, an, ad, au, ap, xw, xe, xs, xn, xd, xu,
                                    xc);
            x[cc] = ((xz ? R(0) : x[cc]) + alf * p[cc]) + omg * xc;
            const R rv = s[cc] - omg * tv;
            r[cc] = rv;
            acc[0] += (double)rv * (double)rv;
            acc[1] += (double)rv * (double)cr[cc];
            xw = xc;
            xc = xe;
            xe = nxe;
            xs = nxs;
            xn = nxn;
            xd = nxd;
            xu = nxu;
            cc = ce;
            ce = nce;
            ie = nie;
        }
    }
    block_sum_store_at<2>(acc, partials, m.nblocks, T.tile);
}

template <typename R>
__global__ __launch_bounds__(256) void k_update_b(IterArgs A) {
    int lb;
    const IterMember &M = member_u(A, lb);
    update_body<R>(glob((R *)M.V[VX]), glob((const R *)M.V[VP]),
                   glob((const R *)M.V[VCS]), glob((const R *)M.V[VS]),
                   glob((const R *)M.V[VT]), glob((const R *)M.V[VCR]),
                   glob((R *)M.V[VR]), M.n, M.nbu, glob((const SolverScalars *)M.sc),
                   M.part, lb);
}

// ---------------------------------------------------------------------------
// inner-solve start and generic sums
// ---------------------------------------------------------------------------
// inner solve start: residuals measured relative to ||r_inner,0||
// the scalars of a fresh inner solve from <r,r> (in the partial sums)
__device__ __forceinline__ void fin_inner_init_body(const double *partials, int nb,
                                                    SolverScalars *sc, double rtol) {
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
        sc->iters = 0;
        sc->guard = 1;
        sc->tol = rtol;
        sc->best_err = sc->err;
        sc->best_it = 0;
        sc->dvalid = 0;
        sc->nonfinite = !isfinite(rr);
        sc->done = (rr == 0.0) || (sc->err < rtol) || sc->nonfinite;
    }
}
__global__ void k_fin_inner_init(const double *partials, int nb, SolverScalars *sc,
                                 double rtol) {
    fin_inner_init_body(partials, nb, sc, rtol);
}

// three generic sums into dst[0..2]
__global__ void k_finalize_sums_3(const double *partials, int nb, double *dst,
                                  const OuterState *gate = nullptr) {
    if (gate && !gate->go) return;
    const double a = reduce_slot_1024(partials, nb);
    const double b = reduce_slot_1024(partials + nb, nb);
    const double d = reduce_slot_1024(partials + 2 * (long long)nb, nb);
    if (threadIdx.x == 0) {
        dst[0] = a;
        dst[1] = b;
        dst[2] = d;
    }
}

// ---------------------------------------------------------------------------
// boundary rows and copies
// ---------------------------------------------------------------------------
// compact iterate back into the interior of the padded x
__global__ void k_scatter_x(const double *__restrict__ xc, double *__restrict__ x, Geo pg,
                            CGeo g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    x[pidx(pg, k, j, i)] = xc[cidx(g, k, j, i)];
}

// final boundary rows of x (dynamics.f90:437-442), or the Neumann copies left
// by mat_ax on the early-return path (:383); whole z planes
__global__ void k_x_rows(double *x, const double *sw, Geo g, double dzn0, int final_rows) {
    const int P = g.nj * g.ni;  // (32-bit index arithmetic: a plane is far below 2^31)
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
        const int i = (int)((unsigned)t / (unsigned)g.nj), j = t - i * g.nj;
        x[pidx(g, g.kp + 1, j, i)] = x[pidx(g, g.kp, j, i)];
        if (final_rows)
            x[pidx(g, 0, j, i)] = x[pidx(g, 1, j, i)] - sw[pidx(g, 0, j, i)] * dzn0;
        else
            x[pidx(g, 0, j, i)] = x[pidx(g, 1, j, i)];
    }
}

// d (compact FP32) added to the padded FP64 x: x = x + d; d = 0 when the
// inner solve made no update (its start does not store d)
__global__ void k_add_correction(double *x, const float *d, Geo pg, CGeo g,
                                 const SolverScalars *sc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const bool none = !sc->dvalid;
    x[pidx(pg, k, j, i)] += none ? 0.0 : (double)d[cidx(g, k, j, i)];
}

// ---------------------------------------------------------------------------
// The refinement step on the device (the single-system multigrid mixed solve;
// the host loop of bicgstab_mixed_batch, one decision at a time)
// ---------------------------------------------------------------------------
__global__ void k_outer_start(OuterState *os) {
    if (threadIdx.x != 0) return;
    os->sc_err = 0.0;
    os->err_prev = 0.0;
    os->err = 0.0;
    os->rtol = 0.0;
    os->first = 1;
    os->active = 1;
    os->go = 1;
    os->outer = 0;
    os->status = 0;
    os->fallback = 0;
    os->total_iters = 0;
}
// the step runs when an inner solve is under way and has finished
// (converged, stagnated or broken down), or when the host forces it
__global__ void k_outer_gate(OuterState *os, const SolverScalars *sc, int force) {
    if (threadIdx.x != 0) return;
    os->go = os->active && (force || halted(sc));
}
// x += d on the interior (the halo is not touched until the solve ends);
// d = 0 when the inner solve made no update.  Gated on the refinement step.
__global__ void k_add_correction_gated(double *x, const float *d, Geo pg, CGeo g,
                                       const OuterState *gate, const SolverScalars *sc) {
    if (!gate->go) return;
    const bool none = !sc->dvalid;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i0 = blockIdx.z * kResPlanes + 1;
    if (k > g.kp || j > g.jp) return;
    const int i1 = min(i0 + kResPlanes - 1, g.ip);
    const int SI = (int)pg.si, CS = g.kp * g.jp;
    int c = i0 * SI + j * (int)pg.sj + k;
    int q = (i0 - 1) * CS + (j - 1) * g.kp + (k - 1);
    for (int i = i0; i <= i1; ++i, c += SI, q += CS) {
        x[c] = x[c] + (none ? 0.0 : (double)d[q]);
    }
}
// x -= d on the interior: the iterate before the last correction (to the
// rounding of x + d - d), when that correction did not reduce the residual
__global__ void k_undo_correction(double *x, const float *d, Geo pg, CGeo g,
                                  const SolverScalars *sc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp || !sc->dvalid) return;
    x[pidx(pg, k, j, i)] -= (double)d[cidx(g, k, j, i)];
}
// the decision of the host loop from the true residual's sums
// (rr = sums[0], bb = sums[2]): stop (converged, fallback, failure) or start
// the next inner solve with its tolerance
__global__ void k_outer_decide(const double *sums, OuterState *os, SolverScalars *sc, double tol,
                               double inner_rtol, int itmax, int max_outer) {
    if (threadIdx.x != 0 || !os->go) return;
    const double rr = sums[0], bb = sums[2];
    if (os->first) {
        double s = sqrt(bb);
        if (s < 0.01) s = 0.01;
        os->sc_err = s;
    } else {
        os->total_iters += sc->iters;  // the inner solve that just finished
    }
    const double err = sqrt(rr) / os->sc_err;
    os->err = err;
    int active = 0;
    if (os->first && !isfinite(err)) {
        os->status = DCM_NONFINITE;
    } else if (!os->first && !(err < os->err_prev)) {
        os->fallback = 1;
    } else if (err < tol) {
        // converged
    } else if (os->outer == max_outer || os->total_iters >= itmax) {
        os->status = DCM_NOT_CONVERGED;
    } else {
        active = 1;
        os->err_prev = err;
        os->outer += 1;
        double rtol = inner_rtol;
        const double need_rel = 0.5 * tol * os->sc_err / (sqrt(rr) > 0 ? sqrt(rr) : 1.0);
        if (need_rel > rtol) rtol = need_rel;
        os->rtol = rtol;
    }
    os->first = 0;
    os->active = active;
    if (!active) sc->done = 1;  // the inner kernels that follow skip
}
__global__ void k_fin_inner_init_dev(const double *partials, int nb, SolverScalars *sc,
                                     const OuterState *gate) {
    if (!(gate->go && gate->active)) return;
    fin_inner_init_body(partials, nb, sc, gate->rtol);
}

// inner start: r = float(r64 - mean) (the mean removed in FP64, so that a
// large mean does not swamp the rest in FP32), cr = r; <r,r>.  p = v1 = d = 0
// are not stored: the first iteration's kernels take them as zero (sc->iters)
__global__ void k_inner_init(const double *r64, float *r, float *cr, CGeo g,
                             const double *mean_sum, double *partials, long long stride,
                             const OuterState *gate = nullptr) {
    if (gate && !(gate->go && gate->active)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i0 = blockIdx.z * kResPlanes + 1;
    double s[2] = {0.0, 0.0};
    if (k <= g.kp && j <= g.jp) {
        const int i1 = min(i0 + kResPlanes - 1, g.ip);
        const double mean = *mean_sum / (double)g.n;
        const int CS = g.kp * g.jp;
        int c = (i0 - 1) * CS + (j - 1) * g.kp + (k - 1);
        for (int i = i0; i <= i1; ++i, c += CS) {
            const float rv = (float)(r64[c] - mean);
            r[c] = rv;
            cr[c] = rv;
            s[0] += (double)rv * (double)rv;
        }
    }
    block_sum_store<2>(s, partials, stride);
}

// ---------------------------------------------------------------------------
// whole-array helpers for the standalone precondition / matvec API
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

// tuning: i planes per marching block, and the first-sweep variant per
// precision (0 fused in registers, 2 streaming pass + plain sweep).
// Measured at 256x256x128 on MI355X: FP64 is fastest split (its IEEE
// divisions make the fused kernel issue-bound), FP32 fused in registers; 8
// and 16 planes are equal, 4 and 32 slower.  (An LDS-shared first sweep and
// a temporally blocked preconditioner holding all three sweeps in LDS were
// both bit-identical and slower: the stencil kernels here are bound by load
// latency, not by their HBM bytes.)
#ifndef DCM_MARCH_PLANES
#define DCM_MARCH_PLANES 8
#endif
#ifndef DCM_JAC_FIRST64
#define DCM_JAC_FIRST64 2
#endif
#ifndef DCM_JAC_FIRST32
#define DCM_JAC_FIRST32 0
#endif
constexpr int kMarchPlanes = DCM_MARCH_PLANES;
constexpr int kUpdateBlocks = 2048; // largest grid of the streaming update
constexpr int kPollEvery = 4;       // host polls the device flags this often

// update-kernel blocks of a member: a full grid for large systems, about four
// elements per thread for small ones
static int update_blocks(long long n) {
    long long b = (n + 4LL * kBlock - 1) / (4LL * kBlock);
    if (b < 8) b = 8;
    return (int)(b > kUpdateBlocks ? kUpdateBlocks : b);
}

// Host-side description of one system of a batched solve.
struct SolveSys {
    dcm_ctx *c;       // the member's context: geometry, coefficients, scalars
    void *V[10];      // its compact vectors (precision of the solve)
    int march_planes; // i planes per marching block
};

// The device tables of a batched solve (member table and block maps),
// uploaded for each solve; `tables` is a device scratch the caller owns.
struct SolvePlan {
    std::vector<IterMember> mem;
    std::vector<unsigned> mmap;
    std::vector<int> umap;
    unsigned mgrid = 0, ugrid = 0;
    IterArgs args{};
};

static int plan_solve(hipStream_t st, const SolveSys *sys, int nsys, SolvePlan &P,
                      DeviceScratch &tables) {
    if (nsys > kMaxMembers) {
        set_error("batched solve: more than 4096 systems");
        return DCM_BAD_ARG;
    }
    P.mem.resize(nsys);
    P.mmap.clear();
    P.umap.clear();
    unsigned ub = 0;
    long long tiles = 0;
    for (int q = 0; q < nsys; ++q) {
        dcm_ctx *c = sys[q].c;
        IterMember &M = P.mem[q];
        M.m = march_geo(c->geo, sys[q].march_planes);
        M.n = c->geo.ip * c->geo.jp * c->geo.kp;
        M.nbu = update_blocks(M.n);
        const int np = M.m.nblocks > M.nbu ? M.m.nblocks : M.nbu;
        DCM_CHECK(ensure_partials(c, 2 * (size_t)np + 64));
        M.cp = CoefPack{c->met, c->met32};
        M.sc = c->solver;
        M.part = c->partials;
        for (int v = 0; v < 10; ++v) M.V[v] = sys[q].V[v];
        M.ublock0 = (int)ub;
        if (M.m.nblocks >= (1 << kTileBits)) {
            set_error("batched solve: a system has more than 2^20 tiles");
            return DCM_BAD_ARG;
        }
        if (nsys > 1) P.umap.insert(P.umap.end(), (size_t)M.nbu, q);
        tiles += M.m.nblocks;
        ub += (unsigned)M.nbu;
    }
    if (nsys == 1) {
        P.mgrid = march_grid(P.mem[0].m);
    } else {
        // the members' tiles in order, split into 8 contiguous XCD ranges:
        // block b runs entry (b mod 8) * per + b / 8 of the list
        const long long per = (tiles + 7) / 8;
        P.mgrid = (unsigned)(8 * per);
        P.mmap.assign(P.mgrid, kNoTile);
        long long e = 0;
        for (int q = 0; q < nsys; ++q)
            for (int t = 0; t < P.mem[q].m.nblocks; ++t, ++e) {
                const long long b = (e % per) * 8 + e / per;
                P.mmap[b] = ((unsigned)q << kTileBits) | (unsigned)t;
            }
    }
    P.ugrid = ub;
    const size_t bm = sizeof(IterMember) * nsys;
    const size_t bmap = sizeof(unsigned) * P.mmap.size() + sizeof(int) * P.umap.size();
    char *d;
    DCM_CHECK(tables.ensure(bm + bmap + 256, (void **)&d));
    DCM_HIP(hipMemcpyAsync(d, P.mem.data(), bm, hipMemcpyHostToDevice, st));
    unsigned *dmap = (unsigned *)(d + ((bm + 255) / 256) * 256);
    int *dumap = (int *)(dmap + P.mmap.size());
    if (nsys > 1) {
        DCM_HIP(hipMemcpyAsync(dmap, P.mmap.data(), sizeof(unsigned) * P.mmap.size(),
                               hipMemcpyHostToDevice, st));
        DCM_HIP(hipMemcpyAsync(dumap, P.umap.data(), sizeof(int) * P.umap.size(),
                               hipMemcpyHostToDevice, st));
    }
    DCM_HIP(hipStreamSynchronize(st));  // the host vectors may change after return
    P.args.mem = (const IterMember *)d;
    P.args.mmap = nsys > 1 ? dmap : nullptr;
    P.args.umap = nsys > 1 ? dumap : nullptr;
    return DCM_OK;
}

// Launches of one BiCGSTAB iteration.  One system: the kernels take its
// vectors and geometry as arguments.  Several: the batched kernels find
// their member in the device table.
template <typename R>
struct IterLaunch {
    dcm_ctx *c;
    const SolvePlan &P;
    int nsys, seq;
    bool one() const { return nsys == 1; }
    const IterMember &M() const { return P.mem[0]; }
    R *v(int i) const { return (R *)P.mem[0].V[i]; }

    // Chebyshev relaxation of the sweeps: only the FP32 inner solve, and only
    // with the fused first sweep (the factors of sweeps 0..3 are c->cheb_w)
    bool cheb() const { return sizeof(R) == 4 && c->cheb_on && DCM_JAC_FIRST32 == 0; }
    template <bool PU, bool W> void jac_launch(const char *name, int si, int ti, int oi, R w) const {
        if (one())
            launch_timed(c, name, k_jac<R, PU, W>, dim3(P.mgrid), dim3(kBlock), (const R *)v(si),
                         (const R *)v(ti), v(oi), M().m, M().cp, (const SolverScalars *)M().sc, w);
        else
            launch_timed(c, name, k_jac_b<R, PU, W>, dim3(P.mgrid), dim3(kBlock), P.args, si, ti,
                         oi, w);
    }
    // a plain sweep; wi: which relaxation factor under Chebyshev
    void jac(const char *name, int si, int ti, int oi, int wi) const {
        if (cheb())
            jac_launch<false, true>(name, si, ti, oi, (R)c->cheb_w[wi]);
        else
            jac_launch<false, false>(name, si, ti, oi, R(1));
    }
    void jac_p(int si, int ti, int oi) const {
        if (cheb())
            jac_launch<true, true>("jac_p", si, ti, oi, (R)c->cheb_w[c->cheb_deg - 1]);
        else
            jac_launch<true, false>("jac_p", si, ti, oi, R(1));
    }
    template <int MODE, bool W> void first_reg(int ti, R w0, R w1) const {
        if (one())
            hipLaunchKernelGGL((k_jac_first_reg<R, MODE, W>), dim3(P.mgrid), dim3(kBlock), 0,
                               c->stream, (const R *)v(VR), (const R *)v(VV1), v(ti), v(VWA),
                               M().m, M().cp, (const SolverScalars *)M().sc, w0, w1);
        else
            hipLaunchKernelGGL((k_jac_first_reg_b<R, MODE, W>), dim3(P.mgrid), dim3(kBlock), 0,
                               c->stream, P.args, (int)VR, (int)VV1, ti, (int)VWA, w0, w1);
    }
    template <int MODE> void first(int ti) const {
        const int variant = sizeof(R) == 8 ? DCM_JAC_FIRST64 : DCM_JAC_FIRST32;
        const int tok = prof_begin(c, "jac_first");
        if (variant == 0) {
            if (cheb())
                first_reg<MODE, true>(ti, (R)c->cheb_w[0], (R)c->cheb_w[1]);
            else
                first_reg<MODE, false>(ti, R(1), R(1));
        } else {
            if (one())
                hipLaunchKernelGGL((k_jac_prep<R, MODE>), dim3(P.mgrid), dim3(kBlock), 0,
                                   c->stream, (const R *)v(VR), (const R *)v(VV1), v(ti), v(VWB),
                                   M().m, M().cp, (const SolverScalars *)M().sc);
            else
                hipLaunchKernelGGL((k_jac_prep_b<R, MODE>), dim3(P.mgrid), dim3(kBlock), 0,
                                   c->stream, P.args, (int)VR, (int)VV1, ti, (int)VWB);
            jac("jac_first", VWB, ti, VWA, 1);
        }
        prof_end(c, tok);
    }
    template <bool TWO> void matvec(int xi, int ai, int yi) const {
        if (one())
            launch_timed(c, "matvec_dot", k_matvec_dot<R, TWO>, dim3(P.mgrid), dim3(kBlock),
                         (const R *)v(xi), (const R *)v(ai), v(yi), M().m, M().cp,
                         (const SolverScalars *)M().sc, M().part);
        else
            launch_timed(c, "matvec_dot", k_matvec_dot_b<R, TWO>, dim3(P.mgrid), dim3(kBlock),
                         P.args, xi, ai, yi);
    }
    void update() const {
        if (one())
            launch_timed(c, "update", k_update<R>, dim3(P.ugrid), dim3(kBlock), v(VX),
                         (const R *)v(VP), (const R *)v(VCS), (const R *)v(VS), (const R *)v(VT),
                         (const R *)v(VCR), v(VR), M().n, M().nbu, (const SolverScalars *)M().sc,
                         M().part);
        else
            launch_timed(c, "update", k_update_b<R>, dim3(P.ugrid), dim3(kBlock), P.args);
    }
    // the multigrid iteration's t = A cs, <t,t>, <t,s> without storing t, and
    // its update recomputing t (per-tile partial sums); with sequential sums,
    // which read t back, the stored t and the streaming update
    void matvec_t_mg() const {
        if (seq) {
            matvec<true>(VCS, VS, VT);
            return;
        }
        launch_timed(c, "matvec_dot", k_matvec_dot<R, true, false>, dim3(P.mgrid), dim3(kBlock),
                     (const R *)v(VCS), (const R *)v(VS), v(VT), M().m, M().cp,
                     (const SolverScalars *)M().sc, M().part);
    }
    void update_mg() const {
        if (seq) {
            update();
            return;
        }
        launch_timed(c, "update", k_update_march<R>, dim3(P.mgrid), dim3(kBlock), v(VX),
                     (const R *)v(VP), (const R *)v(VCS), (const R *)v(VS), (const R *)v(VCR),
                     v(VR), M().m, M().cp, (const SolverScalars *)M().sc, M().part);
    }
    // 0: alf (<cr,v1>), 1: omg (<t,t>, <t,s>), 2: the residual (<r,r>, <r,cr>);
    // tiles: the residual's partial sums are per marching tile (update_mg)
    void fin(int which, bool tiles = false) const {
        if (seq) {
            const int a0 = which == 0 ? VCR : (which == 1 ? VT : VR);
            const int b0 = which == 0 ? VV1 : (which == 1 ? VT : VR);
            const int a1 = which == 0 ? -1 : (which == 1 ? VT : VR);
            const int b1 = which == 0 ? -1 : (which == 1 ? VS : VCR);
            hipLaunchKernelGGL(k_seq_dots<R>, dim3(nsys), dim3(64), 0, c->stream, P.args, a0, b0,
                               a1, b1);
        }
        if (one()) {
            const int nb = seq ? 1 : (which == 2 && !tiles ? M().nbu : M().m.nblocks);
            if (which == 0)
                hipLaunchKernelGGL(k_fin_v1, dim3(1), dim3(1024), 0, c->stream, M().part, nb,
                                   M().sc);
            else if (which == 1)
                hipLaunchKernelGGL(k_fin_t1, dim3(1), dim3(1024), 0, c->stream, M().part, nb,
                                   M().sc);
            else
                hipLaunchKernelGGL(k_fin_r1, dim3(1), dim3(1024), 0, c->stream, M().part, nb,
                                   M().sc);
        } else {
            if (which == 0)
                hipLaunchKernelGGL(k_fin_v, dim3(nsys), dim3(1024), 0, c->stream, P.args, seq);
            else if (which == 1)
                hipLaunchKernelGGL(k_fin_t, dim3(nsys), dim3(1024), 0, c->stream, P.args, seq);
            else
                hipLaunchKernelGGL(k_fin_r, dim3(nsys), dim3(1024), 0, c->stream, P.args, seq);
        }
    }
    // the multigrid preconditioner of a single system: the mixed path's FP32
    // inner solve, or the FP64 BiCGSTAB with the V-cycle in place of the
    // reference's sweeps
    bool mg() const { return one() && c->mg_on && c->mg; }
    // the right-hand side of a V-cycle, ti = r - bet*(omg*v1) (MODE 0) or
    // r - alf*v1 (MODE 1), and the V-cycle's first two level-0 sweeps into
    // VWA: the first from zero, ti/a_p, the second damped.  FP32 in the
    // fused mode: one marching kernel that evaluates ti/a_p at the centre
    // and the six neighbours (the arithmetic of the two kernels below, which
    // write ti/a_p to VWB and sweep from it, so the same bits).
    template <int MODE> void vcycle_start(int ti) const {
        const R w = (R)kMgOmega;
        if (sizeof(R) == 4 && c->mg_tail) {
            first_reg<MODE, true>(ti, R(1), w);
            return;
        }
        hipLaunchKernelGGL((k_jac_prep<R, MODE>), dim3(P.mgrid), dim3(kBlock), 0, c->stream,
                           (const R *)v(VR), (const R *)v(VV1), v(ti), v(VWB), M().m, M().cp,
                           (const SolverScalars *)M().sc);
        jac_launch<false, true>("jac", VWB, ti, VWA, w);
    }
    // the rest of the V-cycle on the right-hand side in ti, from the second
    // sweep's result in VWA: the coarse-grid correction and two sweeps (the
    // last into oi; P: p = x + bet*p)
    template <bool P> void vcycle_rest(int ti, int oi) const {
        const R w = (R)kMgOmega;
        const int tok = prof_begin(c, "mg_correct");
        (void)mg_correct<R>(c, (const R *)v(ti), v(VWA), (const SolverScalars *)M().sc);
        prof_end(c, tok);
        jac_launch<false, true>("jac", VWA, ti, VWB, w);
        jac_launch<P, true>(P ? "jac_p" : "jac", VWB, ti, oi, w);
    }
    // dynamics.f90:391-425 with one V-cycle in place of each application of
    // the Jacobi sweeps: p = V(t) + bet*p, cs = V(s)
    void iteration_mg() const {
        {
            vcycle_start<0>(VT);             // t = r - bet*(omg*v1), two sweeps
            vcycle_rest<true>(VT, VP);       // p = V(t) + bet*p
            matvec<false>(VP, VCR, VV1);     // v1 = A p, <cr,v1>
            fin(0);
            vcycle_start<1>(VS);             // s = r - alf*v1, two sweeps
            vcycle_rest<false>(VS, VCS);     // cs = V(s)
            matvec_t_mg();                   // t = A cs, <t,t>, <t,s>
            fin(1);
            update_mg();                     // x, r (t recomputed); <r,r>, <r,cr>
            fin(2, true);
        }
    }
    // dynamics.f90:391-425
    void iteration() const {
        if (mg()) {
            iteration_mg();
            return;
        }
        // the sweeps between the fused first one and the last: d - 3 of
        // them, d the preconditioner's number of a_p divisions (4, as the
        // reference's s = r/a_p and three sweeps, unless Chebyshev sets it)
        const int d = cheb() ? c->cheb_deg : 4;
        first<0>(VT);               // t = r - bet*(omg*v1), sweep 1 -> wa
        int cur = VWA;
        for (int k = 2; k < d - 1; ++k) {
            const int nxt = cur == VWA ? VWB : VWA;
            jac("jac", cur, VT, nxt, k);  // sweep k
            cur = nxt;
        }
        jac_p(cur, VT, VP);         // the last sweep, p = s + bet*p
        matvec<false>(VP, VCR, VV1);  // v1 = A p, <cr,v1>
        fin(0);
        first<1>(VS);               // s = r - alf*v1, sweep 1 -> wa
        cur = VWA;
        for (int k = 2; k < d - 1; ++k) {
            const int nxt = cur == VWA ? VWB : VWA;
            jac("jac", cur, VS, nxt, k);
            cur = nxt;
        }
        jac("jac", cur, VS, VCS, d - 1);  // cs = M s
        matvec<true>(VCS, VS, VT);  // t = A cs, <t,t>, <t,s>
        fin(1);
        update();                   // x, r; <r,r>, <r,cr>
        fin(2);
    }
};

// The graph of one multigrid inner iteration of a single system (the
// iteration's kernels and the gather of its flags), captured on first use
// and whenever the launch arguments differ from the capture's: the key holds
// the system's device table entry, the grids, the flags buffer and the
// hierarchy.
template <typename R>
static int iteration_graph(dcm_ctx *c, const SolvePlan &P, const IterLaunch<R> &L,
                           int *dflags) {
    std::vector<char> key(sizeof(IterMember) + 4 * sizeof(void *) + 2 * sizeof(unsigned));
    char *k = key.data();
    std::memcpy(k, &P.mem[0], sizeof(IterMember));
    k += sizeof(IterMember);
    const void *ptrs[4] = {P.args.mem, dflags, c->mg, c->partials};
    std::memcpy(k, ptrs, sizeof(ptrs));
    k += sizeof(ptrs);
    std::memcpy(k, &P.mgrid, sizeof(unsigned));
    std::memcpy(k + sizeof(unsigned), &P.ugrid, sizeof(unsigned));
    if (c->iter_graph && key == c->iter_key) return DCM_OK;
    if (c->iter_graph) (void)hipGraphExecDestroy(c->iter_graph);
    c->iter_graph = nullptr;
    c->iter_key.clear();
    DCM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    L.iteration();
    hipLaunchKernelGGL(k_gather_flags, dim3(1), dim3(64), 0, c->stream, P.args, 1, dflags);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (ec != hipSuccess) {
        set_error(std::string("capturing the inner iteration: ") + hipGetErrorString(ec));
        return DCM_HIP_ERROR;
    }
    const hipError_t ei = hipGraphInstantiate(&c->iter_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c->iter_graph = nullptr;
        set_error(std::string("instantiating the inner iteration: ") + hipGetErrorString(ei));
        return DCM_HIP_ERROR;
    }
    c->iter_key = std::move(key);
    return DCM_OK;
}

// BiCGSTAB iterations (dynamics.f90:390-430) of a batch of independent
// systems in lockstep, every launch serving all of them; a member that has
// converged or broken down skips its blocks.  Each member's initial state and
// scalars (its tolerance included) must be set.  `c` supplies the stream, the
// profiling and the flags buffers; outs[q] receives member q's iterations and
// residual, status[q] its status.
template <typename R>
static int bicgstab_iterate_batch(dcm_ctx *c, const SolveSys *sys, int nsys, int itmax,
                                  SolverOut *outs, int *status) {
    SolvePlan P;
    DCM_CHECK(plan_solve(c->stream, sys, nsys, P, c->solve_tables));
    const IterArgs &A = P.args;
    const IterLaunch<R> L{c, P, nsys, c->seq_sums ? 1 : 0};
    int *dflags, *hflags;
    DCM_CHECK(c->flags.ensure(sizeof(int) * nsys, (void **)&dflags));
    DCM_CHECK(ensure_host_flags(c, nsys, &hflags));
    if (itmax < 1) itmax = 1;
    int launched = 0;
    // A multigrid iteration is about sixty launches, most of them small
    // coarse-level kernels that finish faster than the host can queue them:
    // it is captured once as a graph and replayed (not while kernels are
    // being timed one by one).
    const bool graph = L.mg() && c->graphs && !c->profiling;
    if (graph) DCM_CHECK(iteration_graph(c, P, L, dflags));
    while (launched < itmax) {
        // a V-cycle-preconditioned solve converges in one to three
        // iterations: poll after every one, so that no iteration is
        // launched after convergence
        const int every = L.mg() ? 1 : kPollEvery;
        const int chunk = (itmax - launched) < every ? (itmax - launched) : every;
        if (graph) {
            DCM_HIP(hipGraphLaunch(c->iter_graph, c->stream));
        } else {
            for (int q = 0; q < chunk; ++q) L.iteration();
            DCM_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_gather_flags, dim3(1), dim3(64), 0, c->stream, A, nsys,
                               dflags);
        }
        launched += chunk;
        DCM_HIP(hipMemcpyAsync(hflags, dflags, sizeof(int) * nsys, hipMemcpyDeviceToHost,
                               c->stream));
        DCM_HIP(hipStreamSynchronize(c->stream));
        bool all = true;
        for (int q = 0; q < nsys; ++q) all = all && (hflags[q] & 3) != 0;
        if (all) break;
    }
    // the kernels after convergence were no-ops; the device counted the
    // iterations that ran
    for (int q = 0; q < nsys; ++q) {
        dcm_ctx *mc = sys[q].c;
        DCM_HIP(hipMemcpyAsync(mc->h_solver, mc->solver, sizeof(SolverScalars),
                               hipMemcpyDeviceToHost, c->stream));
    }
    DCM_HIP(hipStreamSynchronize(c->stream));
    int worst = DCM_OK;
    for (int q = 0; q < nsys; ++q) {
        const SolverScalars *h = sys[q].c->h_solver;
        outs[q].iters = h->iters;
        outs[q].err = h->err;
        int st = DCM_OK;
        if (h->breakdown)
            st = DCM_BREAKDOWN;
        else if (h->nonfinite)
            st = DCM_NONFINITE;
        else if (!(h->err < h->tol))
            st = DCM_NOT_CONVERGED;
        status[q] = st;
        if (st != DCM_OK && worst == DCM_OK) worst = st;
    }
    if (worst == DCM_NONFINITE) set_error("bicgstab: non-finite residual");
    return worst;
}

// one system: the batch engine with one member


// The reference's FP64 BiCGSTAB (dynamics.f90:293-444) on padded x, b for a
// batch of independent systems (one per context) that share c's stream:
// every member's setup, one batched iteration loop, every member's finish.
int bicgstab64_batch(dcm_ctx *c, const SolveTask *tasks, int n, double tol, int itmax,
                     SolverOut *outs, int *status) {
    std::vector<SolveSys> sys;
    std::vector<int> idx;
    std::vector<char> early(n, 0), iterated(n, 0);
    for (int q = 0; q < n; ++q) {
        dcm_ctx *mc = tasks[q].c;
        const Geo &pg = mc->geo;
        const CGeo g = cgeo(pg);
        const Interior L = interior_launch(pg);
        const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
        DCM_CHECK(ensure_partials(mc, 4 * (size_t)nb + 64));
        DCM_CHECK(ensure_solver_vectors64(mc, 10));
        double *V = mc->sv64;
        CoefPack cp{mc->met, mc->met32};
        int tok = prof_begin(c, "resid0");
        hipLaunchKernelGGL(k_resid0, L.grid, L.block, 0, mc->stream, tasks[q].x, tasks[q].b, V,
                           V + g.n, V + 2 * g.n, V + 3 * g.n, V + 9 * g.n, pg, g, cp,
                           mc->partials, (long long)nb);
        prof_end(c, tok);
        if (c->seq_sums)
            hipLaunchKernelGGL(k_seq_resid0, dim3(1), dim3(64), 0, mc->stream, tasks[q].b, V, pg,
                               g, mc->partials);
        hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(1024), 0, mc->stream, mc->partials,
                           c->seq_sums ? 1 : nb, mc->solver, tol, 1);
        DCM_LAUNCH_CHECK();
        DCM_HIP(hipMemcpyAsync(mc->h_solver, mc->solver, sizeof(SolverScalars),
                               hipMemcpyDeviceToHost, mc->stream));
    }
    for (int q = 0; q < n; ++q) DCM_HIP(hipStreamSynchronize(tasks[q].c->stream));
    for (int q = 0; q < n; ++q) {
        dcm_ctx *mc = tasks[q].c;
        status[q] = DCM_OK;
        outs[q] = SolverOut();
        if (mc->h_solver->nonfinite) {
            set_error("bicgstab: non-finite initial residual");
            status[q] = DCM_NONFINITE;
        } else if (mc->h_solver->err < tol) {
            early[q] = 1;  // :383, return before the loop
            outs[q].err = mc->h_solver->err;
        } else {
            SolveSys sy;
            sy.c = mc;
            for (int v = 0; v < 10; ++v) sy.V[v] = mc->sv64 + (size_t)v * mc->geo.ip * mc->geo.jp *
                                                                 mc->geo.kp;
            sy.march_planes = mc->march_planes > 0 ? mc->march_planes : kMarchPlanes;
            sys.push_back(sy);
            idx.push_back(q);
            iterated[q] = 1;
        }
    }
    if (!sys.empty()) {
        std::vector<SolverOut> o(sys.size());
        std::vector<int> st(sys.size());
        (void)bicgstab_iterate_batch<double>(c, sys.data(), (int)sys.size(), itmax, o.data(),
                                             st.data());
        for (size_t u = 0; u < sys.size(); ++u) {
            outs[idx[u]] = o[u];
            status[idx[u]] = st[u];
        }
    }
    int worst = DCM_OK;
    for (int q = 0; q < n; ++q) {
        dcm_ctx *mc = tasks[q].c;
        const Geo &pg = mc->geo;
        const CGeo g = cgeo(pg);
        const Interior L = interior_launch(pg);
        if (iterated[q])
            hipLaunchKernelGGL(k_scatter_x, L.grid, L.block, 0, mc->stream,
                               mc->sv64 + 9 * (size_t)g.n, tasks[q].x, pg, g);
        if (status[q] == DCM_OK) {
            // the halo of x as the reference leaves it: periodic faces from
            // mat_ax, then the boundary rows of :437-442 (or mat_ax's Neumann
            // copies on the early return)
            DCM_CHECK(fill<double>(mc, tasks[q].x, FILL_ALONG_DIM_WO, 0.0, 0.0));
            hipLaunchKernelGGL(k_x_rows, dim3(256), dim3(kBlock), 0, mc->stream, tasks[q].x,
                               tasks[q].sw, pg, mc->h_dzn[0], early[q] ? 0 : 1);
        }
        DCM_LAUNCH_CHECK();
        if (status[q] != DCM_OK && worst == DCM_OK) worst = status[q];
    }
    return worst;
}

// A fixed number of FP64 iterations of a batch from its current x and b
// (tolerance 0, so no member stops early), for timing: only solver scratch
// (residual and iterate vectors, partials, scalars) is written.
int bicgstab_probe(dcm_ctx *c, const SolveTask *tasks, int n, int iters, SolverOut *outs,
                   int *status) {
    std::vector<SolveSys> sys(n);
    for (int q = 0; q < n; ++q) {
        dcm_ctx *mc = tasks[q].c;
        const Geo &pg = mc->geo;
        const CGeo g = cgeo(pg);
        const Interior L = interior_launch(pg);
        const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
        DCM_CHECK(ensure_partials(mc, 4 * (size_t)nb + 64));
        DCM_CHECK(ensure_solver_vectors64(mc, 10));
        double *V = mc->sv64;
        CoefPack cp{mc->met, mc->met32};
        hipLaunchKernelGGL(k_resid0, L.grid, L.block, 0, mc->stream, tasks[q].x, tasks[q].b, V,
                           V + g.n, V + 2 * g.n, V + 3 * g.n, V + 9 * g.n, pg, g, cp,
                           mc->partials, (long long)nb);
        hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(1024), 0, mc->stream, mc->partials, nb,
                           mc->solver, 0.0, 1);
        DCM_LAUNCH_CHECK();
        sys[q].c = mc;
        for (int v = 0; v < 10; ++v) sys[q].V[v] = V + (size_t)v * g.n;
        sys[q].march_planes = mc->march_planes > 0 ? mc->march_planes : kMarchPlanes;
    }
    (void)bicgstab_iterate_batch<double>(c, sys.data(), n, iters, outs, status);
    return DCM_OK;
}

int bicgstab64(dcm_ctx *c, const double *sw, double *x, double *b, double tol, int itmax,
               SolverOut *out) {
    const SolveTask t{c, sw, x, b};
    int st = DCM_OK;
    DCM_CHECK(bicgstab64_batch(c, &t, 1, tol, itmax, out, &st));
    return st;
}

// The single-system multigrid mixed solve with the refinement step on the
// device.  After the start (true residual, decision, first inner start),
// each chunk is one inner iteration followed by the refinement step
// (correction, true residual, decision, next inner start), every kernel of
// the step gated on the inner solve having finished; the chunk is replayed
// from one graph and the host polls the refinement state once per chunk
// instead of after the residual, after every iteration and after every
// inner solve.  The kernels and their order are those of the host loop of
// bicgstab_mixed_batch, so the bits are the same.
static int chunk_graph(dcm_ctx *c, const SolvePlan &P, const IterLaunch<float> &L,
                       const std::vector<const void *> &ptrs, const double *scal, int nscal,
                       const std::function<void()> &step) {
    std::vector<char> key(sizeof(IterMember) + ptrs.size() * sizeof(void *) +
                          (size_t)nscal * sizeof(double) + 2 * sizeof(unsigned));
    char *k = key.data();
    std::memcpy(k, &P.mem[0], sizeof(IterMember));
    k += sizeof(IterMember);
    std::memcpy(k, ptrs.data(), ptrs.size() * sizeof(void *));
    k += ptrs.size() * sizeof(void *);
    std::memcpy(k, scal, (size_t)nscal * sizeof(double));
    k += (size_t)nscal * sizeof(double);
    std::memcpy(k, &P.mgrid, sizeof(unsigned));
    std::memcpy(k + sizeof(unsigned), &P.ugrid, sizeof(unsigned));
    if (c->chunk_graph && key == c->chunk_key) return DCM_OK;
    if (c->chunk_graph) (void)hipGraphExecDestroy(c->chunk_graph);
    c->chunk_graph = nullptr;
    c->chunk_key.clear();
    DCM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    L.iteration();
    step();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (ec != hipSuccess) {
        set_error(std::string("capturing the inner iteration: ") + hipGetErrorString(ec));
        return DCM_HIP_ERROR;
    }
    const hipError_t ei = hipGraphInstantiate(&c->chunk_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c->chunk_graph = nullptr;
        set_error(std::string("instantiating the inner iteration: ") + hipGetErrorString(ei));
        return DCM_HIP_ERROR;
    }
    c->chunk_key = std::move(key);
    return DCM_OK;
}

// One system's chunked solve as a resumable state: start() queues the start
// and captures the chunk graph, launch() queues one chunk and the copy of
// the refinement state to the host, poll() reads it once the stream has
// drained, finish() handles the outcome.  Everything runs on c->stream.
struct ChunkedSolve {
    dcm_ctx *c = nullptr;
    SolveTask t{};
    double tol = 0.0, inner_rtol = 0.0;
    int itmax = 0;
    static constexpr int kMaxOuter = 50;
    static constexpr int kInnerCap = 2000;  // FP32 iterations per refinement sweep
    Geo pg{};
    CGeo g{};
    Interior IL{}, RL{};
    int nbr = 0;
    float *V[10] = {};
    double *r64 = nullptr;
    SolvePlan P;
    int seen_outer = 1, launched = 0, cap = 0;
    OuterState h{};

    // true residual, decision, inner start (each gated on os->go)
    void decide_and_start() {
        OuterState *os = c->outer_d;
        const OuterState *gos = os;
        SolverScalars *sc = c->solver;
        const CoefPack cp{c->met, c->met32};
        double *sums = c->dscal + 40;
        hipLaunchKernelGGL(k_resid_true, RL.grid, RL.block, 0, c->stream, (const double *)t.x,
                           (const double *)t.b, r64, pg, g, cp, c->partials, (long long)nbr, gos);
        hipLaunchKernelGGL(k_finalize_sums_3, dim3(1), dim3(1024), 0, c->stream,
                           (const double *)c->partials, nbr, sums, gos);
        hipLaunchKernelGGL(k_outer_decide, dim3(1), dim3(64), 0, c->stream, (const double *)sums,
                           os, sc, tol, inner_rtol, itmax, kMaxOuter);
        hipLaunchKernelGGL(k_inner_init, RL.grid, RL.block, 0, c->stream, (const double *)r64,
                           V[0], V[1], g, (const double *)(sums + 1), c->partials,
                           (long long)nbr, gos);
        hipLaunchKernelGGL(k_fin_inner_init_dev, dim3(1), dim3(1024), 0, c->stream,
                           (const double *)c->partials, nbr, sc, gos);
    }
    // the refinement step after an inner iteration
    void step(int force) {
        hipLaunchKernelGGL(k_outer_gate, dim3(1), dim3(64), 0, c->stream, c->outer_d,
                           (const SolverScalars *)c->solver, force);
        hipLaunchKernelGGL(k_add_correction_gated, RL.grid, RL.block, 0, c->stream, t.x,
                           (const float *)V[9], pg, g, (const OuterState *)c->outer_d,
                           (const SolverScalars *)c->solver);
        decide_and_start();
    }

    int start(dcm_ctx *c_, const SolveTask &t_, double tol_, int itmax_, double inner_rtol_) {
        c = c_;
        t = t_;
        tol = tol_;
        itmax = itmax_;
        inner_rtol = inner_rtol_;
        pg = c->geo;
        g = cgeo(pg);
        IL = interior_launch(pg);
        RL = resid_launch(pg);
        nbr = (int)(RL.grid.x * RL.grid.y * RL.grid.z);
        const int nb = (int)(IL.grid.x * IL.grid.y * IL.grid.z);
        DCM_CHECK(ensure_partials(c, 4 * (size_t)nb + 64));
        DCM_CHECK(ensure_solver_vectors32(c, 10));
        const size_t cells = (size_t)pg.ip * pg.jp * pg.kp;
        for (int v = 0; v < 10; ++v) V[v] = c->sv32 + (size_t)v * cells;
        DCM_CHECK(work_field64(c, 2, &r64));
        SolveSys sy;
        sy.c = c;
        for (int v = 0; v < 10; ++v) sy.V[v] = V[v];
        sy.march_planes = c->march_planes > 0 ? c->march_planes : kMarchPlanes;
        DCM_CHECK(plan_solve(c->stream, &sy, 1, P, c->solve_tables));
        const IterLaunch<float> L{c, P, 1, c->seq_sums ? 1 : 0};
        hipLaunchKernelGGL(k_outer_start, dim3(1), dim3(64), 0, c->stream, c->outer_d);
        decide_and_start();
        DCM_LAUNCH_CHECK();
        const std::vector<const void *> ptrs = {P.args.mem, c->mg, c->partials, t.x, t.b,
                                                r64, c->outer_d, c->solver, c->dscal,
                                                c->sv32};
        const double scal[3] = {tol, inner_rtol, (double)itmax};
        DCM_CHECK(chunk_graph(c, P, L, ptrs, scal, 3, [&]() { step(0); }));
        // the start leaves inner solve 1 under way (or nothing to do)
        seen_outer = 1;
        launched = 0;
        cap = std::min(kInnerCap, itmax);
        return DCM_OK;
    }
    int launch() {
        if (launched >= cap) {
            step(1);  // the inner solve has used its iterations: correct anyway
            DCM_LAUNCH_CHECK();
        } else {
            DCM_HIP(hipGraphLaunch(c->chunk_graph, c->stream));
            ++launched;
        }
        DCM_HIP(hipMemcpyAsync(c->outer_h, c->outer_d, sizeof(OuterState),
                               hipMemcpyDeviceToHost, c->stream));
        return DCM_OK;
    }
    // after the stream has drained: true while the solve goes on
    bool poll() {
        h = *c->outer_h;
        if (!h.active) return false;
        if (h.outer != seen_outer) {  // a new inner solve has started
            seen_outer = h.outer;
            launched = 0;
            cap = std::max(1, std::min(kInnerCap, itmax - h.total_iters));
        }
        return true;
    }
    int finish(SolverOut *out, int *status) {
        *out = SolverOut();
        out->iters = h.total_iters;
        out->outer = h.outer;
        out->err = h.err;
        if (h.status == DCM_NONFINITE) set_error("bicgstab_mixed: non-finite initial residual");
        *status = h.status;
        if (h.fallback) {
            // the last FP32 correction did not reduce the true residual:
            // take it back and let the reference's FP64 solver finish
            hipLaunchKernelGGL(k_undo_correction, IL.grid, IL.block, 0, c->stream, t.x,
                               (const float *)V[9], pg, g, (const SolverScalars *)c->solver);
            DCM_LAUNCH_CHECK();
            SolverOut o;
            int st = DCM_OK;
            (void)bicgstab64_batch(c, &t, 1, tol, itmax, &o, &st);
            out->iters += o.iters;
            out->err = o.err;
            out->fallback = 1;
            *status = st;
            return st;
        }
        if (*status == DCM_OK) {
            DCM_CHECK(fill<double>(c, t.x, FILL_ALONG_DIM_WO, 0.0, 0.0));
            hipLaunchKernelGGL(k_x_rows, dim3(256), dim3(kBlock), 0, c->stream, t.x, t.sw,
                               c->geo, c->h_dzn[0], 1);
            DCM_LAUNCH_CHECK();
        }
        return *status;
    }
};

static int mixed_one_chunked(dcm_ctx *c, const SolveTask &t, double tol, int itmax,
                             double inner_rtol, SolverOut *out, int *status) {
    ChunkedSolve cs;
    DCM_CHECK(cs.start(c, t, tol, itmax, inner_rtol));
    for (;;) {
        DCM_CHECK(cs.launch());
        DCM_HIP(hipStreamSynchronize(c->stream));
        if (!cs.poll()) break;
    }
    return cs.finish(out, status);
}

// Several single-system multigrid mixed solves at once, system q on
// streams[q] (its context's stream while it runs): every round queues one
// chunk of each unfinished system, then drains the streams and polls them,
// so the systems' small kernels overlap on the device.  Each system runs
// exactly the kernels of its solo solve, so it ends with the same bits.
int bicgstab_mixed_concurrent(const SolveTask *tasks, int n, const hipStream_t *streams,
                              double tol, int itmax, const double *inner_rtol, SolverOut *outs,
                              int *status) {
    std::vector<ChunkedSolve> cs(n);
    std::vector<hipStream_t> saved(n);
    std::vector<char> live(n, 0);
    int worst = DCM_OK;
    for (int q = 0; q < n; ++q) {
        dcm_ctx *mc = tasks[q].c;
        saved[q] = mc->stream;
        mc->stream = streams[q];
    }
    auto restore = [&]() {
        for (int q = 0; q < n; ++q) tasks[q].c->stream = saved[q];
    };
    for (int q = 0; q < n; ++q) {
        const int st = cs[q].start(tasks[q].c, tasks[q], tol, itmax, inner_rtol[q]);
        if (st != DCM_OK) {
            restore();
            return st;
        }
        live[q] = 1;
    }
    std::vector<hipStream_t> uniq(streams, streams + n);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    for (int left = n; left > 0;) {
        for (int q = 0; q < n; ++q)
            if (live[q]) {
                const int st = cs[q].launch();
                if (st != DCM_OK) {
                    restore();
                    return st;
                }
            }
        for (hipStream_t s : uniq) {
            const hipError_t e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                restore();
                set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
                return DCM_HIP_ERROR;
            }
        }
        for (int q = 0; q < n; ++q)
            if (live[q] && !cs[q].poll()) {
                live[q] = 0;
                --left;
                const int st = cs[q].finish(&outs[q], &status[q]);
                if (st != DCM_OK && worst == DCM_OK) worst = st;
            }
    }
    restore();
    return worst;
}

// Mixed precision for a batch: FP32 inner BiCGSTAB (all members in one
// batched loop) inside each member's FP64 iterative refinement.
int bicgstab_mixed_batch(dcm_ctx *c, const SolveTask *tasks, int n, double tol, int itmax,
                         double inner_rtol, SolverOut *outs, int *status) {
    if (n == 1 && tasks[0].c == c && c->mg_on && c->mg && c->graphs && !c->profiling)
        return mixed_one_chunked(c, tasks[0], tol, itmax, inner_rtol, outs, status);
    struct Mem {
        double sc_err = 0.0, err_prev = 0.0;
        double *r64 = nullptr, *xsave = nullptr;
        float *V[10] = {};
        bool active = true, fallback = false;
    };
    std::vector<Mem> M(n);
    const int max_outer = 50;
    const int inner_cap = 2000;  // FP32 iterations per refinement sweep
    for (int q = 0; q < n; ++q) {
        dcm_ctx *mc = tasks[q].c;
        const Geo &pg = mc->geo;
        const Interior L = interior_launch(pg);
        const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
        DCM_CHECK(ensure_partials(mc, 4 * (size_t)nb + 64));
        DCM_CHECK(ensure_solver_vectors32(mc, 10));
        const size_t cells = (size_t)pg.ip * pg.jp * pg.kp;
        for (int v = 0; v < 10; ++v) M[q].V[v] = mc->sv32 + (size_t)v * cells;
        DCM_CHECK(work_field64(mc, 2, &M[q].r64));
        outs[q] = SolverOut();
        status[q] = DCM_OK;
    }
    for (int outer = 0; outer <= max_outer; ++outer) {
        // true FP64 residual of every active member, its sums to the host
        for (int q = 0; q < n; ++q) {
            if (!M[q].active) continue;
            dcm_ctx *mc = tasks[q].c;
            const Geo &pg = mc->geo;
            const CGeo g = cgeo(pg);
            const Interior L = resid_launch(pg);
            const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
            CoefPack cp{mc->met, mc->met32};
            int tok = prof_begin(c, "resid_true");
            hipLaunchKernelGGL(k_resid_true, L.grid, L.block, 0, mc->stream, tasks[q].x,
                               tasks[q].b, M[q].r64, pg, g, cp, mc->partials, (long long)nb);
            prof_end(c, tok);
            hipLaunchKernelGGL(k_finalize_sums_3, dim3(1), dim3(1024), 0, mc->stream,
                               mc->partials, nb, mc->dscal + 40);
            DCM_LAUNCH_CHECK();
            DCM_HIP(hipMemcpyAsync(mc->hscal + 40, mc->dscal + 40, 3 * sizeof(double),
                                   hipMemcpyDeviceToHost, mc->stream));
        }
        for (int q = 0; q < n; ++q)
            if (M[q].active) DCM_HIP(hipStreamSynchronize(tasks[q].c->stream));
        std::vector<SolveSys> sys;
        std::vector<int> idx;
        for (int q = 0; q < n; ++q) {
            if (!M[q].active) continue;
            dcm_ctx *mc = tasks[q].c;
            const Geo &pg = mc->geo;
            const CGeo g = cgeo(pg);
            const Interior L = interior_launch(pg);
            const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
            const double rr = mc->hscal[40], bb = mc->hscal[42];
            if (outer == 0) {
                M[q].sc_err = sqrt(bb);
                if (M[q].sc_err < 0.01) M[q].sc_err = 0.01;
            }
            const double err = sqrt(rr) / M[q].sc_err;
            outs[q].err = err;
            if (outer == 0 && !std::isfinite(err)) {
                set_error("bicgstab_mixed: non-finite initial residual");
                status[q] = DCM_NONFINITE;
                M[q].active = false;
                continue;
            }
            if (outer > 0 && !(err < M[q].err_prev)) {
                // the last FP32 correction did not reduce the true residual
                // (an inner breakdown, stagnation at FP32 resolution, or
                // overflow): restore x and let the reference's FP64 solver
                // finish this member
                DCM_HIP(hipMemcpyAsync(tasks[q].x, M[q].xsave, pg.ntot * sizeof(double),
                                       hipMemcpyDeviceToDevice, mc->stream));
                M[q].active = false;
                M[q].fallback = true;
                continue;
            }
            if (err < tol) {
                M[q].active = false;
                continue;
            }
            if (outer == max_outer || outs[q].iters >= itmax) {
                status[q] = DCM_NOT_CONVERGED;
                M[q].active = false;
                continue;
            }
            M[q].err_prev = err;
            outs[q].outer += 1;
            // inner solve A d = r - mean(r), to a relative reduction
            // inner_rtol, or as soon as the outer target is surely met
            const Interior PL = resid_launch(pg);
            const int npl = (int)(PL.grid.x * PL.grid.y * PL.grid.z);
            hipLaunchKernelGGL(k_inner_init, PL.grid, PL.block, 0, mc->stream, M[q].r64, M[q].V[0],
                               M[q].V[1], g, mc->dscal + 41, mc->partials, (long long)npl);
            double rtol = inner_rtol;
            const double need_rel = 0.5 * tol * M[q].sc_err / (sqrt(rr) > 0 ? sqrt(rr) : 1.0);
            if (need_rel > rtol) rtol = need_rel;
            hipLaunchKernelGGL(k_fin_inner_init, dim3(1), dim3(1024), 0, mc->stream,
                               mc->partials, npl, mc->solver, rtol);
            DCM_LAUNCH_CHECK();
            SolveSys sy;
            sy.c = mc;
            for (int v = 0; v < 10; ++v) sy.V[v] = M[q].V[v];
            sy.march_planes = mc->march_planes > 0 ? mc->march_planes : kMarchPlanes;
            sys.push_back(sy);
            idx.push_back(q);
        }
        if (sys.empty()) break;
        int cap = inner_cap;
        for (int u : idx)
            if (itmax - outs[u].iters < cap) cap = itmax - outs[u].iters;
        if (cap < 1) cap = 1;
        // (each member's inner tolerance is in its scalars)
        std::vector<SolverOut> o(sys.size());
        std::vector<int> st(sys.size());
        (void)bicgstab_iterate_batch<float>(c, sys.data(), (int)sys.size(), cap, o.data(),
                                            st.data());
        for (size_t u = 0; u < sys.size(); ++u) {
            const int q = idx[u];
            dcm_ctx *mc = tasks[q].c;
            const Geo &pg = mc->geo;
            const CGeo g = cgeo(pg);
            const Interior L = interior_launch(pg);
            outs[q].iters += o[u].iters;
            // whatever the inner status, its correction is judged by the true
            // residual at the top of the loop; a guarded breakdown stops
            // before the update that would use a non-finite scalar
            if (!M[q].xsave) DCM_CHECK(work_field64(mc, 1, &M[q].xsave));
            DCM_HIP(hipMemcpyAsync(M[q].xsave, tasks[q].x, pg.ntot * sizeof(double),
                                   hipMemcpyDeviceToDevice, mc->stream));
            hipLaunchKernelGGL(k_add_correction, L.grid, L.block, 0, mc->stream, tasks[q].x,
                               M[q].V[9], pg, g, (const SolverScalars *)mc->solver);
            DCM_LAUNCH_CHECK();
        }
    }
    // members whose FP32 corrections stalled: the FP64 solver, batched
    std::vector<SolveTask> fb;
    std::vector<int> fidx;
    for (int q = 0; q < n; ++q)
        if (M[q].fallback) {
            fb.push_back(tasks[q]);
            fidx.push_back(q);
        }
    if (!fb.empty()) {
        std::vector<SolverOut> o(fb.size());
        std::vector<int> st(fb.size());
        (void)bicgstab64_batch(c, fb.data(), (int)fb.size(), tol, itmax, o.data(), st.data());
        for (size_t u = 0; u < fb.size(); ++u) {
            const int q = fidx[u];
            outs[q].iters += o[u].iters;
            outs[q].err = o[u].err;
            outs[q].fallback = 1;
            status[q] = st[u];
        }
    }
    int worst = DCM_OK;
    for (int q = 0; q < n; ++q) {
        if (status[q] == DCM_OK && !M[q].fallback) {
            dcm_ctx *mc = tasks[q].c;
            DCM_CHECK(fill<double>(mc, tasks[q].x, FILL_ALONG_DIM_WO, 0.0, 0.0));
            hipLaunchKernelGGL(k_x_rows, dim3(256), dim3(kBlock), 0, mc->stream, tasks[q].x,
                               tasks[q].sw, mc->geo, mc->h_dzn[0], 1);
            DCM_LAUNCH_CHECK();
        }
        if (status[q] != DCM_OK && worst == DCM_OK) worst = status[q];
    }
    return worst;
}

int bicgstab_mixed(dcm_ctx *c, const double *sw, double *x, double *b, double tol, int itmax,
                   double inner_rtol, SolverOut *out) {
    const SolveTask t{c, sw, x, b};
    int st = DCM_OK;
    DCM_CHECK(bicgstab_mixed_batch(c, &t, 1, tol, itmax, inner_rtol, out, &st));
    return st;
}

// ---------------------------------------------------------------------------
// Fixed device scratch of dcm_inn_prod_noalloc: the partial sums of one dot
// product and its result, in a file-scope array that each device holds from
// module load to process end.  2 MiB of partials cover any grid whose
// interior launch has at most kNoAllocPartials blocks (ceil(kp/64) x
// ceil(jp/4) x ip for kp >= 64; 512x512x256 needs 262144).  Host code never
// passes `scratch` itself to a kernel: its device address comes from
// hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocPartials = size_t(1) << 18;
constexpr size_t kNoAllocDoubles = kNoAllocPartials + 32;  // + the result slot, padded
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kNoAllocMaxDevices = 64;
static double *g_scratch_dev[kNoAllocMaxDevices];  // device address of `scratch` per device
static std::mutex g_scratch_mu;

// the device address of `scratch` on `device` (looked up on the first call
// for that device, with that device current)
static int noalloc_scratch(int device, double **out) {
    if (device < 0 || device >= kNoAllocMaxDevices) {
        set_error("dcm_inn_prod_noalloc: device ordinal " + std::to_string(device) +
                  " beyond the " + std::to_string(kNoAllocMaxDevices) + " supported");
        return DCM_BAD_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        int cur = -1;
        DCM_HIP(hipGetDevice(&cur));
        if (cur != device) DCM_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return DCM_HIP_ERROR;
        }
        g_scratch_dev[device] = (double *)p;
    }
    *out = g_scratch_dev[device];
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
    if (c->seq_sums) {
        hipLaunchKernelGGL(k_seq_dot_padded, dim3(1), dim3(64), 0, c->stream, x, y, c->geo,
                           c->dscal + 40);
        DCM_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(k_dot_padded, L.grid, L.block, 0, c->stream, x, y, c->geo,
                           c->partials);
        DCM_LAUNCH_CHECK();
        DCM_CHECK(finalize_sums(c, c->partials, nb, 1, c->dscal + 40));
    }
    DCM_HIP(hipMemcpyAsync(c->hscal + 40, c->dscal + 40, sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    *out = c->hscal[40];
    return DCM_OK;
}

int dcm_inn_prod_noalloc(dcm_ctx *c, const double *x, const double *y, double *out) {
    if (!c || !x || !y || !out) return DCM_BAD_ARG;
    const Interior L = interior_launch(c->geo);
    const size_t nb = (size_t)L.grid.x * L.grid.y * L.grid.z;
    if (nb > kNoAllocPartials) {
        set_error("dcm_inn_prod_noalloc: the grid's interior launch has " + std::to_string(nb) +
                  " blocks; the fixed scratch holds the partial sums of at most " +
                  std::to_string(kNoAllocPartials));
        return DCM_BAD_ARG;
    }
    double *s = nullptr;
    DCM_CHECK(noalloc_scratch(c->device, &s));
    double *res = s + kNoAllocPartials;  // every slot written before it is read
    if (c->seq_sums) {
        hipLaunchKernelGGL(k_seq_dot_padded, dim3(1), dim3(64), 0, c->stream, x, y, c->geo, res);
        DCM_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(k_dot_padded, L.grid, L.block, 0, c->stream, x, y, c->geo, s);
        DCM_LAUNCH_CHECK();
        DCM_CHECK(finalize_sums(c, s, (int)nb, 1, res));
    }
    // into the context's pinned host scalars (allocated with the context)
    DCM_HIP(hipMemcpyAsync(c->hscal + 40, res, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DCM_HIP(hipStreamSynchronize(c->stream));
    *out = c->hscal[40];
    return DCM_OK;
}

int dcm_true_residual(dcm_ctx *c, const double *x, const double *b, double *out) {
    if (!c || !x || !b || !out) return DCM_BAD_ARG;
    const Geo &pg = c->geo;
    const CGeo g = cgeo(pg);
    const Interior L = resid_launch(pg);
    const int nb = (int)(L.grid.x * L.grid.y * L.grid.z);
    CoefPack cp{c->met, c->met32};
    DCM_CHECK(ensure_partials(c, 4 * (size_t)nb + 64));
    hipLaunchKernelGGL(k_resid_true, L.grid, L.block, 0, c->stream, x, b, (double *)nullptr, pg,
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

int dcm_set_sum_order(dcm_ctx *c, int order) {
    if (!c || (order != DCM_SUM_TREE && order != DCM_SUM_SEQUENTIAL)) return DCM_BAD_ARG;
    c->seq_sums = order == DCM_SUM_SEQUENTIAL;
    return DCM_OK;
}

int dcm_set_precond_chebyshev(dcm_ctx *c, double a, int divisions) {
    if (!c || !(a >= 0.0) || a >= 2.0 || divisions < 3 || divisions > 16) return DCM_BAD_ARG;
    dcm::set_chebyshev(c, a, divisions);
    return DCM_OK;
}

int dcm_bicgstab_mixed(dcm_ctx *c, const double *sw, double *x, double *b, double tol, int itmax,
                       int *iters, int *outer, double *err) {
    if (!c || !sw || !x || !b) return DCM_BAD_ARG;
    SolverOut o;
    int st = bicgstab_mixed(c, sw, x, b, tol, itmax > 0 ? itmax : 9999, 1e-4, &o);
    if (iters) *iters = o.iters;
    if (outer) *outer = o.outer;
    if (err) *err = o.err;
    return st;
}

}  // extern "C"
