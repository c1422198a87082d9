// dcm_mpdata32.hip -- MPDATA of the mixed-precision model step: every field
// a step advects (theta, each q category) in one batched launch per pass,
// FP32 intermediates, FP64 fields.
//
// The algorithm is the linked copy's mpdata_3d / mpdata_vec_3d with kord 2
// (pamm/sfvt/advection_3d.F90:219-714, 1250-1753): shift by the minimum,
// zero-sum exit, a donor pass (:39-116), the antidiffusive velocities of
// Smolarkiewicz (1984) with the density weighting of :318-506, the
// Smolarkiewicz-Grabowski (1990) limiter of :542-675 and a donor pass with
// the limited velocities (:684-697); the vector routine takes velocities and
// limiter from the category's first field.  These FP32 kernels carry no
// bit-parity duty (the FP64 path's kernels in dcm_advect.hip do), so:
//   * a quotient of metrics becomes a product with a table of reciprocals
//     (Tab32), computed once per thread or per i plane; the nine quotients
//     per cell that depend on the field are hardware reciprocals, as in the
//     FP32 instantiation of dcm_advect.hip;
//   * the FP32 arrays (the velocity shadows and every intermediate) use an
//     aligned copy of the padded layout: each (j,i) column has a stride that
//     is a multiple of 32 elements and level k = 1 starts a 128-byte line,
//     so a wavefront's 64 consecutive levels are two whole lines (the
//     reference's kp+2 stride would straddle three);
//   * a thread owns one (k,j) column of a tile and marches a chunk of i
//     planes, keeping the column sums of planes i-1, i, i+1 in registers
//     (the flux through the lower x face of a cell is the upper x face's
//     flux of the cell before it);
//   * in the donor and antidiffusive-velocity passes a lane takes the k+1
//     and k-1 values of the arrays it has loaded from the adjacent lanes of
//     the wavefront (DPP shifts), so only the first and last lane of a
//     column segment load them;
//   * each pass writes, besides a cell, the halo cells that the reference's
//     exchange would copy from it: no fill kernel runs;
//   * the limiter factors never reach HBM: the last kernel computes them one
//     i plane ahead into a ring of four planes in LDS (the tile's rows plus
//     one halo row on each side) and applies the limited donor pass to the
//     plane behind, so beta and the last donor pass cost one read of their
//     inputs and one write of psi;
//   * each pass serves every field (or every category's first field) in one
//     launch: per step a minimum, three passes and two small reductions.
// Offsets are 32-bit: a padded field is below 2^31 elements (checked).
//
//   K0 k32_shadow                 FP32 aligned copies of u, v, w
//   K1 k32_min / k32_min_finish   m(f) = min(psi_f) (+ lbc, ubc)
//   K2 k32_donor1                 psi1_f = donor(psi_f - m), halo images;
//                                 psi_store = psi - m of each category's
//                                 first field; block sums of psi - m
//      k32_exit_flags             flag(L) = (sum < small), per category
//   K3 k32_antidiff               U*, V*, W* of each category (first field)
//   K4 k32_betalim                beta of the category (LDS), then
//                                 psi_f = donor(psi1_f, limited U*) + m
//      k32_exit_unshift           psi = (psi - m) + m where the exit was taken
#include <cfloat>
#include <algorithm>
#include <vector>

#include "dcm_device.h"

// The library is built without multiply-add contraction, which the FP64
// path's bit parity needs; these FP32 kernels contract freely.
#pragma clang fp contract(fast)

namespace dcm {

constexpr int kMp32Max = 16;     // fields per batched launch
constexpr int kMp32Planes = 16;  // i planes per marching tile
constexpr int kMinBlocks = 1024;
constexpr int kK4Threads = 1024; // largest block of the beta + limited kernel
constexpr int kK4MaxKt = 256;    // its column of levels: kp <= 256

// geometry of one field: the FP64 reference layout (sj, si) and the aligned
// FP32 layout (sj4, si4, level k at column offset k + kK0), and the
// 256-thread marching tiles of K2 and K3
struct PM {
    int ip, jp, kp;
    int sj, si;                        // FP64 strides
    int sj4, si4;                      // FP32 strides
    int kt, jt, nkb, njb, ic, nic;     // tile lanes (k), rows (j); tiles; i chunks
    int ntiles;                        // tiles per field
};
constexpr int kK0 = 31;  // column offset of level 0 in the aligned layout

static PM pm_geo(const Geo &g, int ic) {
    PM m;
    m.ip = g.ip;
    m.jp = g.jp;
    m.kp = g.kp;
    m.sj = (int)g.sj;
    m.si = (int)g.si;
    m.sj4 = ((g.kp + 2 + 31) / 32) * 32;
    m.si4 = m.sj4 * g.nj;
    const int w64 = (64 - g.kp % 64) % 64, w32 = (32 - g.kp % 32) % 32;
    m.kt = g.kp > 32 && w64 <= w32 ? 64 : 32;
    m.jt = kBlock / m.kt;
    m.nkb = (g.kp + m.kt - 1) / m.kt;
    m.njb = (g.jp + m.jt - 1) / m.jt;
    m.ic = ic;
    m.nic = (g.ip + ic - 1) / ic;
    m.ntiles = m.nkb * m.njb * m.nic;
    return m;
}
// elements of one aligned FP32 array (a multiple of 64)
static long long ntot4_of(const PM &m) {
    const long long n = (long long)m.si4 * (m.ip + 2) + 32;
    return ((n + 63) / 64) * 64;
}

// Block -> item of a launch.  The item list is split into 8 contiguous
// ranges, one per XCD (block b runs on XCD b mod 8), so tiles that share
// faces share an L2.
__device__ __forceinline__ bool xcd_item(int total, int &t) {
    const int per = (total + 7) >> 3;
    const int b = (int)blockIdx.x;
    t = (b & 7) * per + (b >> 3);
    return t < total;
}
static unsigned xcd_grid(long long total) { return (unsigned)(((total + 7) / 8) * 8); }

struct Col {
    int f, tile, k, j, i0, i1;
    bool act;
};
__device__ __forceinline__ bool pm_col(const PM &m, int nf, Col &C) {
    int t;
    if (!xcd_item(nf * m.ntiles, t)) return false;
    C.f = t / m.ntiles;
    C.tile = t - C.f * m.ntiles;
    const int kb = C.tile % m.nkb;
    const int r = C.tile / m.nkb;
    const int jb = r % m.njb;
    const int icb = r / m.njb;
    C.k = kb * m.kt + (int)(threadIdx.x % m.kt) + 1;
    C.j = jb * m.jt + (int)(threadIdx.x / m.kt) + 1;
    C.i0 = icb * m.ic + 1;
    C.i1 = C.i0 + m.ic - 1;
    if (C.i1 > m.ip) C.i1 = m.ip;
    C.act = C.k <= m.kp && C.j <= m.jp;
    return true;
}

// The batch: fields f = 0..nf-1; category leaders (first fields) L =
// 0..nlead-1.  FP32 scratch (aligned arrays of ntot4 elements): the velocity
// shadows u, v, w; psi1 of field f at 3+f; U*, V*, W* of leader L at
// 3+nf+3L; psi_store of leader L at 3+nf+3*nlead+L; and, for the separate
// limiter kernels (kp > 256), beta_up and beta_down of leader L at
// 3+nf+4*nlead+2L.
struct Mp32Args {
    double *psi[kMp32Max];
    double lbc[kMp32Max], ubc[kMp32Max];
    signed char lslot[kMp32Max];   // leader slot of field f
    signed char lfield[kMp32Max];  // field of leader slot L
    signed char whole[kMp32Max];   // minimum over the whole array (mpdata_3d)
    float *s32;
    long long ntot4;
    double *m;                     // shifts, m[f]
    int *flag;                     // zero-sum exit, flag[L]
    double *partials;              // min partials (K1), exit sums (K2)
    int nf, nlead;
    float dt;
    int neumann;
};
__device__ __forceinline__ float *arr(const Mp32Args &A, int slot) {
    return A.s32 + (size_t)slot * A.ntot4;
}
__device__ __forceinline__ float *psi1_of(const Mp32Args &A, int f) { return arr(A, 3 + f); }
__device__ __forceinline__ float *ustar_of(const Mp32Args &A, int L, int d) {
    return arr(A, 3 + A.nf + 3 * L + d);
}
__device__ __forceinline__ float *store_of(const Mp32Args &A, int L) {
    return arr(A, 3 + A.nf + 3 * A.nlead + L);
}
__device__ __forceinline__ float *beta_of(const Mp32Args &A, int L, int d) {
    return arr(A, 3 + A.nf + 4 * A.nlead + 2 * L + d);
}

template <typename T> __device__ __forceinline__ T pos0(T x) { return x > T(0) ? x : T(0); }
template <typename T> __device__ __forceinline__ T neg0(T x) { return x < T(0) ? x : T(0); }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float min3f(float a, float b, float c) {
    return fminf(fminf(a, b), c);
}
__device__ __forceinline__ int wrapi(int x, int n) { return x < 1 ? x + n : (x > n ? x - n : x); }

// The lanes of a marching tile hold consecutive levels of a column (lane l:
// level k0 + l), so the value at k+1 of an array a lane has loaded at k is
// the next lane's and the value at k-1 the previous lane's: DPP wave shifts
// (wave_shl:1 gives lane l the value of lane l+1, wave_shr:1 that of lane
// l-1) replace those loads.  A lane whose neighbour holds another column or
// no level (the first and last lane of a column segment, the level kp) gets
// 0 here and loads the value itself.  Every lane of the tile must execute
// the shift, outside any branch that depends on the lane.
__device__ __forceinline__ int dpp_next(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_prev(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(dpp_next(__float_as_int(v)));
}
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(dpp_prev(__float_as_int(v)));
}
__device__ __forceinline__ double lane_next(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_next((int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_next((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double lane_prev(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_prev((int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_prev((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// A uniform base pointer read at a 32-bit byte offset: the load takes the
// base from scalar registers and the offset from one vector register, so a
// stencil of loads costs no address arithmetic (every neighbour has a base
// of its own, shifted by a constant or a stride, and the k neighbours are
// immediates).
template <typename T> struct Ld {
    const char *b;
    __device__ __forceinline__ explicit Ld(const T *p) : b((const char *)p) {}
    // element dk along k from the byte offset off
    __device__ __forceinline__ T at(unsigned off, int dk = 0) const {
        return *(const T *)(b + off + dk * (int)sizeof(T));
    }
};
template <typename T> struct St {
    char *b;
    __device__ __forceinline__ explicit St(T *p) : b((char *)p) {}
    __device__ __forceinline__ void put(unsigned off, T v, int dk = 0) const {
        *(T *)(b + off + dk * (int)sizeof(T)) = v;
    }
};

// The halo cells that exchange_full copies from interior cell (k,j,i) at
// element c of an array with strides sj, si: the periodic images on the
// lateral faces and edges (j = jp+1 copies j = 1, j = 0 copies j = jp,
// likewise along i), and on the z planes the z-wrapped edges (the receive
// from MPI_PROC_NULL keeps the local copy, mpi_module.f90:1244-1263): a
// z-plane cell outside the lateral interior holds level kp at k = 0 and
// level 1 at kp+1.  The lateral-interior z-plane cells are stored by the
// caller (k = 1 and k = kp lanes).
template <typename T>
__device__ __forceinline__ void put_images(T *a, int c, T v, int k, int j, int i, const PM &m,
                                           int sj, int si) {
    int dj[3], di[3];
    int nj = 1, ni = 1;
    dj[0] = 0;
    di[0] = 0;
    if (j == 1) dj[nj++] = m.jp * sj;
    if (j == m.jp) dj[nj++] = -m.jp * sj;
    if (i == 1) di[ni++] = m.ip * si;
    if (i == m.ip) di[ni++] = -m.ip * si;
#pragma unroll
    for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
        for (int b1 = 0; b1 < 3; ++b1) {
            if (a1 >= nj || b1 >= ni || (a1 == 0 && b1 == 0)) continue;
            const int o = c + dj[a1] + di[b1];
            a[o] = v;
            if (k == m.kp) a[o - m.kp] = v;  // level 0 of a lateral halo column
            if (k == 1) a[o + m.kp] = v;     // level kp+1
        }
}
// psi_store = psi - m at the same halo cells (FP32, aligned c4) from the FP64
// field's own halo cells (FP64 layout c8), as the reference's psi_store copy
// of the whole array holds them
__device__ __forceinline__ void put_store_images(float *q, const double *psi, double mm, int c4,
                                                 int c8, int k, int j, int i, const PM &m) {
    int dj4[3], di4[3], dj8[3], di8[3];
    int nj = 1, ni = 1;
    dj4[0] = di4[0] = dj8[0] = di8[0] = 0;
    if (j == 1) {
        dj4[nj] = m.jp * m.sj4;
        dj8[nj++] = m.jp * m.sj;
    }
    if (j == m.jp) {
        dj4[nj] = -m.jp * m.sj4;
        dj8[nj++] = -m.jp * m.sj;
    }
    if (i == 1) {
        di4[ni] = m.ip * m.si4;
        di8[ni++] = m.ip * m.si;
    }
    if (i == m.ip) {
        di4[ni] = -m.ip * m.si4;
        di8[ni++] = -m.ip * m.si;
    }
#pragma unroll
    for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
        for (int b1 = 0; b1 < 3; ++b1) {
            if (a1 >= nj || b1 >= ni || (a1 == 0 && b1 == 0)) continue;
            const int o4 = c4 + dj4[a1] + di4[b1], o8 = c8 + dj8[a1] + di8[b1];
            q[o4] = (float)(psi[o8] - mm);
            if (k == m.kp) q[o4 - m.kp] = (float)(psi[o8 - m.kp] - mm);
            if (k == 1) q[o4 + m.kp] = (float)(psi[o8 + m.kp] - mm);
        }
}
// exchange_along_dim with lbc = ubc = 0 (the limiter factors, :624-627):
// periodic faces (the edges are left alone: nothing reads them) and zero z
// planes over the whole plane, lateral halo included.  Aligned FP32 layout.
__device__ __forceinline__ void put_faces(float *a, int c, float v, int k, int j, int i,
                                          const PM &m) {
    int dj[3], di[3];
    int nj = 1, ni = 1;
    dj[0] = 0;
    di[0] = 0;
    if (j == 1) dj[nj++] = m.jp * m.sj4;
    if (j == m.jp) dj[nj++] = -m.jp * m.sj4;
    if (i == 1) di[ni++] = m.ip * m.si4;
    if (i == m.ip) di[ni++] = -m.ip * m.si4;
#pragma unroll
    for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
        for (int b1 = 0; b1 < 3; ++b1) {
            if (a1 >= nj || b1 >= ni || (a1 == 0 && b1 == 0)) continue;
            const int o = c + dj[a1] + di[b1];
            if (a1 == 0 || b1 == 0) a[o] = v;  // a face, not an edge
            if (k == 1) a[o - 1] = 0.f;
            if (k == m.kp) a[o + 1] = 0.f;
        }
}

// ---------------------------------------------------------------------------
// K0: the aligned FP32 copies of u, v, w (every padded cell)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k32_shadow(const double *__restrict__ u,
                                                  const double *__restrict__ v,
                                                  const double *__restrict__ w, float *us,
                                                  float *vs, float *ws, PM m, int nk) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= nk || j > m.jp + 1) return;
    const int c8 = i * m.si + j * m.sj + k, c4 = i * m.si4 + j * m.sj4 + k + kK0;
    us[c4] = (float)u[c8];
    vs[c4] = (float)v[c8];
    ws[c4] = (float)w[c8];
}

// ---------------------------------------------------------------------------
// K1: minima (mpdata_3d: whole array :275; mpdata_vec_3d: interior :1307)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k32_min(Mp32Args A, Geo g) {
    const int f = blockIdx.y;
    const double *__restrict__ a = A.psi[f];
    double v = INFINITY;
    if (A.whole[f]) {
        for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < g.ntot;
             t += (long long)gridDim.x * blockDim.x) {
            const double x = a[t];
            v = x < v ? x : v;
        }
    } else {
        // one wave per (j,i) column, lanes along k
        const int lane = threadIdx.x & 63;
        const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
        const int nwaves = (int)(gridDim.x * (blockDim.x >> 6));
        const int ncol = g.ip * g.jp;
        for (int col = wave; col < ncol; col += nwaves) {
            const int i = col / g.jp + 1, j = col - (i - 1) * g.jp + 1;
            const long long base = (long long)i * g.si + (long long)j * g.sj;
            for (int k = lane + 1; k <= g.kp; k += 64) {
                const double x = a[base + k];
                v = x < v ? x : v;
            }
        }
    }
    block_min_store(v, A.partials);  // partials[f * gridDim.x + blockIdx.x]
}

__global__ void k32_min_finish(Mp32Args A, int nb) {
    const int f = blockIdx.x;
    double r = reduce_min_1024(A.partials + (size_t)f * nb, nb);
    if (threadIdx.x == 0) {
        if (A.lbc[f] < r) r = A.lbc[f];
        if (A.ubc[f] < r) r = A.ubc[f];
        A.m[f] = r;
    }
}

// ---------------------------------------------------------------------------
// K2: donor pass on psi - m (first_order_upstream_3d, :74-115), FP32 flux
// arithmetic from the FP64 difference psi - m; psi1 with the halo of
// exchange_full(psi1, lbc - m, ubc - m).  Leaders also store psi_store =
// psi - m (whole array) for the limiter and block sums of psi - m over the
// interior (the exit test, :282-289 / :1314-1315).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k32_donor1(Mp32Args A, PM m, Tab32 T) {
    Col C;
    if (!pm_col(m, A.nf, C)) return;
    const int f = C.f;
    const int L = A.lslot[f];
    const bool lead = A.lfield[L] == f;
    const double *psi = A.psi[f];
    float *out = psi1_of(A, f);
    float *sto = store_of(A, L);
    const float *u = arr(A, 0), *v = arr(A, 1), *w = arr(A, 2);
    const double mm = A.m[f];
    double s[1] = {0.0};
    if (C.act) {
        const int k = C.k, j = C.j, SI = m.si, SJ = m.sj, SI4 = m.si4, SJ4 = m.sj4;
        const Ld<double> Pe(psi + SI), Pn(psi + SJ), Ps(psi - SJ), Pc(psi);
        const Ld<float> Uc(u), Vc(v), Vs(v - SJ4), Wc(w);
        const St<float> O(out), Q(sto);
        const float dt = A.dt;
        const float cy = dt * T.idyn[j];
        const float czr = ((dt * T.ra[k]) * T.idzn[k]) * T.irn[k];
        const float czl = ((dt * T.ra[k - 1]) * T.idzn[k]) * T.irn[k];
        const float zb = (float)(A.lbc[f] - mm), zt = (float)(A.ubc[f] - mm);
        const int nm = A.neumann;
        const bool jedge = j == 1 || j == m.jp;
        // lanes that load their k-1 / k+1 values (see lane_next)
        const int lk = (int)threadIdx.x % m.kt;
        const bool fix_lo = lk == 0, fix_hi = lk == m.kt - 1 || k == m.kp;
        int c = C.i0 * SI + j * SJ + k;
        int c4 = C.i0 * SI4 + j * SJ4 + k + kK0;
        unsigned o4 = (unsigned)c4 * 4u, o8 = (unsigned)c * 8u;
        float pw = (float)(psi[c - SI] - mm);
        double dc = psi[c] - mm;
        float pc = (float)dc;
        const float uw = u[c4 - SI4];
        float fxl = pos0(uw) * pw + neg0(uw) * pc;
        for (int i = C.i0; i <= C.i1; ++i) {
            double xu = 0.0, xd = 0.0;
            float xw = 0.f;
            if (fix_lo) {
                xd = Pc.at(o8, -1);
                xw = Wc.at(o4, -1);
            }
            if (fix_hi) xu = Pc.at(o8, 1);
            const double de = Pe.at(o8) - mm;
            const float pe = (float)de;
            const float pn = (float)(Pn.at(o8) - mm), ps = (float)(Ps.at(o8) - mm);
            const float uc = Uc.at(o4), vc = Vc.at(o4), vl = Vs.at(o4), wc = Wc.at(o4);
            // psi - m at k+1 and k-1 of plane i, w at k-1: the centre values
            // of the adjacent lanes
            const double su = lane_next(dc), sd = lane_prev(dc);
            const float sw = lane_prev(wc);
            const float pu = (float)(fix_hi ? xu - mm : su);
            const float pd = (float)(fix_lo ? xd - mm : sd);
            const float wl = fix_lo ? xw : sw;
            const float fxr = pos0(uc) * pc + neg0(uc) * pe;
            const float fyr = pos0(vc) * pc + neg0(vc) * pn;
            const float fyl = pos0(vl) * ps + neg0(vl) * pc;
            float fzr = (pos0(wc) * pc + neg0(wc) * pu) * czr;
            float fzl = (pos0(wl) * pd + neg0(wl) * pc) * czl;
            if (nm == 1) {  // :95-101
                if (k == 1) fzl = fzr;
                if (k == m.kp) fzr = fzl;
            } else if (nm == 2) {  // :102-110
                if (k == 1) fzl = -fzr;
                if (k == m.kp) fzr = -fzl;
            }
            const float cx = dt * T.xpk[i].y;
            const float res = ((pc - (fxr - fxl) * cx) - (fyr - fyl) * cy) - (fzr - fzl);
            O.put(o4, res);
            if (k == 1) O.put(o4, zb, -1);
            if (k == m.kp) O.put(o4, zt, 1);
            const bool edge = jedge || i == 1 || i == m.ip;
            if (edge) put_images<float>(out, c4, res, k, j, i, m, SJ4, SI4);
            if (lead) {
                s[0] += dc;
                Q.put(o4, pc);
                if (k == 1) Q.put(o4, pd, -1);
                if (k == m.kp) Q.put(o4, pu, 1);
                if (edge) put_store_images(sto, psi, mm, c4, c, k, j, i, m);
            }
            fxl = fxr;
            pc = pe;
            dc = de;
            c += SI;
            c4 += SI4;
            o4 += (unsigned)SI4 * 4u;
            o8 += (unsigned)SI * 8u;
        }
    }
    if (lead) block_sum_store_at<1>(s, A.partials + (size_t)L * m.ntiles, 0, C.tile);
}

// flag(L) = sum(psi - m) < small over the category's first field
__global__ void k32_exit_flags(Mp32Args A, int ntiles) {
    const int L = blockIdx.x;
    const double s = reduce_slot_1024(A.partials + (size_t)L * ntiles, ntiles);
    if (threadIdx.x == 0) A.flag[L] = (s < 1e-60) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// K3: antidiffusive velocities of each category's first field, the
// density-weighted form of :318-506 (g1 = g2 = rhoan(k), g3 the mean of
// rhoan(k), rhoan(k+1)), from the FP32 velocity shadows and psi1.  Column
// sums of one i plane, with P = psi1:
//   D = P(k+1)-P(k-1), S = P(k+1)+P(k-1), Dy = P(j+1)-P(j-1),
//   Sy = P(j+1)+P(j-1), E = P(j+1)+P, F = P(k+1)+P,
//   Uy = U(j+1)+U, Uz = rhoan(k+1)U(k+1)+rhoan(k)U, Vs = V+V(j-1),
//   Vd = V-V(j-1), Z = rhoa(k)W+rhoa(k-1)W(k-1), Wd = rhoa(k)W-rhoa(k-1)W(k-1);
// each four-point sum of the reference is the sum or difference of two of
// them at planes i and i+1 (x-face terms) or i-1 and i+1 (centred terms).
// ---------------------------------------------------------------------------
struct Plane {
    float A, B, Cn, Dm, O;     // psi1 at k+1, k-1, j+1, j-1, centre
    float U0, Uy, Uz;
    float V0, Vm;
    float W0, Wm;
};

__device__ __forceinline__ void plane_at(Plane &q, const float *p1, const float *u,
                                         const float *v, const float *w, int c, int SJ,
                                         float rnk, float rnk1) {
    q.A = p1[c + 1];
    q.B = p1[c - 1];
    q.Cn = p1[c + SJ];
    q.Dm = p1[c - SJ];
    q.O = p1[c];
    q.U0 = u[c];
    q.Uy = u[c + SJ] + q.U0;
    q.Uz = rnk1 * u[c + 1] + rnk * q.U0;
    q.V0 = v[c];
    q.Vm = v[c - SJ];
    q.W0 = w[c];
    q.Wm = w[c - 1];
}

__global__ __launch_bounds__(256) void k32_antidiff(Mp32Args A, PM m, Tab32 T) {
    Col C;
    if (!pm_col(m, A.nlead, C)) return;
    const int L = C.f;
    if (A.flag[L]) return;
    if (!C.act) return;
    const int fld = A.lfield[L];
    const float *p1 = psi1_of(A, fld);
    const float *u = arr(A, 0), *v = arr(A, 1), *w = arr(A, 2);
    float *us = ustar_of(A, L, 0), *vs = ustar_of(A, L, 1), *ws = ustar_of(A, L, 2);
    const int k = C.k, j = C.j, SI = m.si4, SJ = m.sj4;
    // plane i+1 (the march's next plane) and plane i's own values
    const Ld<float> nP(p1 + SI), nPn(p1 + SI + SJ), nPs(p1 + SI - SJ), nU(u + SI),
        nUj(u + SI + SJ), nV(v + SI), nVm(v + SI - SJ), nW(w + SI);
    const Ld<float> Pj(p1 + SJ), Pm(p1 - SJ), Vc(v), Vm(v - SJ), Vj(v + SJ), Wc(w), Wj(w + SJ);
    const St<float> SU(us), SV(vs), SW(ws);
    const float dt = A.dt, hdt = 0.5f * dt, qdt = 0.25f * dt, q = 0.25f;
    const float small = FLT_MIN;
    const float rnk = T.rn[k], rnk1 = T.rn[k + 1];
    const float rak = T.ra[k], rakm = T.ra[k - 1], rak1 = T.ra[k + 1];
    const float ig1 = T.irn[k];                       // 1/g1 = 1/g2
    const float ig3 = 1.0f / (0.5f * (rnk + rnk1));   // 1/g3
    const float idzkm = T.idz[k - 1], idyjm = T.idy[j - 1];
    const float c_zg1 = idzkm * ig1, c_zg3 = idzkm * ig3, c_yg3 = idyjm * ig3;
    const float ihz = T.ihz[k], ihy = T.ihy[j];
    const float dzk = T.dz[k], dyj = T.dy[j], idznk = T.idzn[k], idynj = T.idyn[j];
    const float qg1 = q * ig1, qg3 = q * ig3;
    const float wfac = dt * (rak * ig3);
    const bool jedge = j == 1 || j == m.jp;
    // lanes that load their k-1 / k+1 values (see lane_next)
    const int lk = (int)threadIdx.x % m.kt;
    const bool fix_lo = lk == 0, fix_hi = lk == m.kt - 1 || k == m.kp;
    int c = C.i0 * SI + j * SJ + k + kK0;
    unsigned o4 = (unsigned)c * 4u;
    Plane pw, pc, pn;
    plane_at(pw, p1, u, v, w, c - SI, SJ, rnk, rnk1);
    plane_at(pc, p1, u, v, w, c, SJ, rnk, rnk1);
    for (int i = C.i0; i <= C.i1; ++i) {
        // k+1 values (x*) and k-1 values (y*) of the lanes that load them
        float xa = 0.f, xu = 0.f, xkj = 0.f, xkm = 0.f, xvk = 0.f, xvkm = 0.f, xwk = 0.f;
        float yb = 0.f, ywm = 0.f, ymj = 0.f, ywmj = 0.f;
        if (fix_hi) {
            xa = nP.at(o4, 1);
            xu = nU.at(o4, 1);
            xkj = Pj.at(o4, 1);
            xkm = Pm.at(o4, 1);
            xvk = Vc.at(o4, 1);
            xvkm = Vm.at(o4, 1);
            xwk = Wc.at(o4, 1);
        }
        if (fix_lo) {
            yb = nP.at(o4, -1);
            ywm = nW.at(o4, -1);
            ymj = Pj.at(o4, -1);
            ywmj = Wj.at(o4, -1);
        }
        pn.Cn = nPn.at(o4);
        pn.Dm = nPs.at(o4);
        pn.O = nP.at(o4);
        pn.U0 = nU.at(o4);
        const float uj = nUj.at(o4);
        pn.V0 = nV.at(o4);
        pn.Vm = nVm.at(o4);
        pn.W0 = nW.at(o4);
        const float vj = Vj.at(o4), wj = Wj.at(o4);
        // the values at k+1 and k-1: plane i+1's from this step's centre
        // loads, plane i's from the previous step's (pc)
        const float sa = lane_next(pn.O), sb = lane_prev(pn.O), su = lane_next(pn.U0),
                    swm = lane_prev(pn.W0);
        const float skj = lane_next(pc.Cn), smj = lane_prev(pc.Cn), skm = lane_next(pc.Dm);
        const float svk = lane_next(pc.V0), svkm = lane_next(pc.Vm), swk = lane_next(pc.W0),
                    swmj = lane_prev(wj);
        pn.A = fix_hi ? xa : sa;
        pn.B = fix_lo ? yb : sb;
        pn.Uy = uj + pn.U0;
        pn.Uz = rnk1 * (fix_hi ? xu : su) + rnk * pn.U0;
        pn.Wm = fix_lo ? ywm : swm;
        const float pkj = fix_hi ? xkj : skj, pmj = fix_lo ? ymj : smj, pkm = fix_hi ? xkm : skm;
        const float vk = fix_hi ? xvk : svk, vkm = fix_hi ? xvkm : svkm;
        const float wmj = fix_lo ? ywmj : swmj, wk = fix_hi ? xwk : swk;
        const float4 xt = T.xpk[i];
        const float dxi = xt.x, idxni = xt.y, idxim = xt.z, ihx = xt.w;
        // column sums
        const float Dc = pc.A - pc.B, Dn = pn.A - pn.B;
        const float Sc = pc.A + pc.B, Sn = pn.A + pn.B;
        const float Dyc = pc.Cn - pc.Dm, Dyn = pn.Cn - pn.Dm;
        const float Syc = pc.Cn + pc.Dm, Syn = pn.Cn + pn.Dm;
        const float En = pn.Cn + pn.O, Ew = pw.Cn + pw.O;
        const float Fn = pn.A + pn.O, Fw = pw.A + pw.O;
        const float Vsc = pc.V0 + pc.Vm, Vsn = pn.V0 + pn.Vm;
        const float Vdc = pc.V0 - pc.Vm, Vdn = pn.V0 - pn.Vm;
        const float Zc = rak * pc.W0 + rakm * pc.Wm, Zn = rak * pn.W0 + rakm * pn.Wm;
        const float Wdc = rak * pc.W0 - rakm * pc.Wm, Wdn = rak * pn.W0 - rakm * pn.Wm;
        // divergence terms (u_div1..3)
        const float ud1 = (Wdc + Wdn) * c_zg1 + (Vdc + Vdn) * idyjm;
        const float ud2 = (rak * (pc.W0 + wj) - rakm * (pc.Wm + wmj)) * c_zg1 +
                          (pc.Uy - pw.Uy) * idxim;
        const float ud3 = (pc.Uz - pw.Uz) * (idxim * ig3) +
                          (rnk * Vdc + rnk1 * (vk - vkm)) * c_yg3;
        // cross terms of each component
        const float tz1 = ((qg1 * (Zn + Zc)) * (Dn + Dc)) * rcp((Sn + Sc) + small) * ihz;
        const float ty1 = ((q * (Vsn + Vsc)) * (Dyn + Dyc)) * rcp((Syn + Syc) + small) * ihy;
        const float pa = pkj + pc.A, pb = pmj + pc.B;
        const float tz2 = ((qg1 * (rak * (wj + pc.W0) + rakm * (wmj + pc.Wm))) * (pa - pb)) *
                          rcp((pa + pb) + small) * ihz;
        const float tx2 = ((q * (pc.Uy + pw.Uy)) * (En - Ew)) * rcp((En + Ew) + small) * ihx;
        const float pj = pkj + pc.Cn, pm = pkm + pc.Dm;
        const float ty3 = ((qg3 * (rnk1 * (vk + vkm) + rnk * Vsc)) * (pj - pm)) *
                          rcp((pj + pm) + small) * ihy;
        const float tx3 = ((qg3 * (pc.Uz + pw.Uz)) * (Fn - Fw)) * rcp((Fn + Fw) + small) * ihx;
        const float uc = pc.U0, vc = pc.V0, wc = pc.W0;
        float usv = ((fabsf(uc) * dxi - (dt * uc) * uc) * (pn.O - pc.O)) *
                        rcp((pn.O + pc.O) + small) * idxni -
                    (hdt * uc) * (tz1 + ty1);
        float vsv = ((fabsf(vc) * dyj - (dt * vc) * vc) * (pc.Cn - pc.O)) *
                        rcp((pc.Cn + pc.O) + small) * idynj -
                    (hdt * vc) * (tz2 + tx2);
        float wsv = ((fabsf(wc) * dzk - (wfac * wc) * wc) * (pc.A - pc.O)) *
                        rcp((pc.A + pc.O) + small) * idznk -
                    (hdt * wc) * (ty3 + tx3);
        usv -= (qdt * uc) * ((pn.U0 - pw.U0) * idxim + ud1);
        vsv -= (qdt * vc) * ((vj - pc.Vm) * idyjm + ud2);
        wsv -= (qdt * wc) * ((rak1 * wk - rakm * pc.Wm) * c_zg3 + ud3);
        SU.put(o4, usv);
        SV.put(o4, vsv);
        SW.put(o4, wsv);
        // exchange_full(., 0, 0): zero z planes in the lateral interior
        if (k == 1) {
            SU.put(o4, 0.f, -1);
            SV.put(o4, 0.f, -1);
            SW.put(o4, 0.f, -1);
        }
        if (k == m.kp) {
            SU.put(o4, 0.f, 1);
            SV.put(o4, 0.f, 1);
            SW.put(o4, 0.f, 1);
        }
        if (jedge || i == 1 || i == m.ip) {
            put_images<float>(us, c, usv, k, j, i, m, SJ, SI);
            put_images<float>(vs, c, vsv, k, j, i, m, SJ, SI);
            put_images<float>(ws, c, wsv, k, j, i, m, SJ, SI);
        }
        pw = pc;
        pc = pn;
        c += SI;
        o4 += (unsigned)SI * 4u;
    }
}

// ---------------------------------------------------------------------------
// Limiter factors (eqs 19-20 of Smolarkiewicz & Grabowski 1990, :549-616) of
// one cell: bounds over psi_store = psi - m and psi1 at the cell and its six
// neighbours, density-weighted in/outflow denominators.
// ---------------------------------------------------------------------------
struct BetaIn {
    float qc, qw, qe, qn, qs, qu, qd;  // psi_store
    float oc, ow, oe, on, os, ou, od;  // psi1
    float uc, uw, vc, vl, wc, wl;      // U*, V*, W* at the six faces
};
__device__ __forceinline__ void beta_cell(const BetaIn &x, float rnk, float cx, float cy,
                                          float czp, float czm, float dt, float &bu,
                                          float &bd) {
    const float small = FLT_MIN;
    float mx = fmaxf(fmaxf(x.qc, x.qw), fmaxf(x.qe, x.qn));
    mx = fmaxf(fmaxf(mx, x.qs), fmaxf(x.qu, x.qd));
    mx = fmaxf(fmaxf(mx, x.oc), fmaxf(x.ow, x.oe));
    mx = fmaxf(fmaxf(mx, x.on), fmaxf(x.os, fmaxf(x.ou, x.od)));
    float mn = fminf(fminf(x.qc, x.qw), fminf(x.qe, x.qn));
    mn = fminf(fminf(mn, x.qs), fminf(x.qu, x.qd));
    mn = fminf(fminf(mn, x.oc), fminf(x.ow, x.oe));
    mn = fminf(fminf(mn, x.on), fminf(x.os, fminf(x.ou, x.od)));
    const float d1 = dt * ((cx * (pos0(x.uw) * x.ow - neg0(x.uc) * x.oe) +
                            cy * (pos0(x.vl) * x.os - neg0(x.vc) * x.on)) +
                           (czm * pos0(x.wl) * x.od - czp * neg0(x.wc) * x.ou)) +
                     small;
    const float d2 = dt * (x.oc * ((cx * (pos0(x.uc) - neg0(x.uw)) + cy * (pos0(x.vc) - neg0(x.vl))) +
                                   (czp * pos0(x.wc) - czm * neg0(x.wl))) +
                           small);
    bu = (rnk * (mx - x.oc)) * rcp(d1);
    bd = (rnk * (x.oc - mn)) * rcp(d2);
}

// The inputs of beta from the aligned FP32 arrays at byte offset o (all
// bases uniform; k neighbours are immediates).  With `next`, the west
// values are the previous plane's centre values, passed in.
struct BetaSrc {
    Ld<float> q, qe, qw, qn, qs, o, oe, ow, on, os, u, uw, v, vs, w;
    __device__ BetaSrc(const float *sq, const float *p1, const float *us, const float *vs_,
                       const float *ws, int SI4, int SJ4)
        : q(sq), qe(sq + SI4), qw(sq - SI4), qn(sq + SJ4), qs(sq - SJ4), o(p1), oe(p1 + SI4),
          ow(p1 - SI4), on(p1 + SJ4), os(p1 - SJ4), u(us), uw(us - SI4), v(vs_), vs(vs_ - SJ4),
          w(ws) {}
};
__device__ __forceinline__ void beta_load(BetaIn &x, const BetaSrc &S, unsigned o) {
    x.qc = S.q.at(o);
    x.qe = S.qe.at(o);
    x.qn = S.qn.at(o);
    x.qs = S.qs.at(o);
    x.qu = S.q.at(o, 1);
    x.qd = S.q.at(o, -1);
    x.oc = S.o.at(o);
    x.oe = S.oe.at(o);
    x.on = S.on.at(o);
    x.os = S.os.at(o);
    x.ou = S.o.at(o, 1);
    x.od = S.o.at(o, -1);
    x.uc = S.u.at(o);
    x.vc = S.v.at(o);
    x.vl = S.vs.at(o);
    x.wc = S.w.at(o);
    x.wl = S.w.at(o, -1);
}
__device__ __forceinline__ void beta_load_full(BetaIn &x, const BetaSrc &S, unsigned o) {
    beta_load(x, S, o);
    x.qw = S.qw.at(o);
    x.ow = S.ow.at(o);
    x.uw = S.uw.at(o);
}

// eq 18: the limited velocity at a face with antidiffusive velocity x, the
// factors of the cell below the face (_c) and above it (_o)
__device__ __forceinline__ float limited(float x, float bd_c, float bu_c, float bu_o,
                                         float bd_o) {
    return min3f(1.f, bd_c, bu_o) * pos0(x) + min3f(1.f, bu_c, bd_o) * neg0(x);
}

// ---------------------------------------------------------------------------
// K4: beta of the category's first field and the last donor pass of field f
// (:634-697), + m; psi with the halo of exchange_full(psi, lbc, ubc) after
// the unshift, i.e. z planes (bc - m) + m.
//
// A block owns whole columns of kp <= 256 levels (thread t: level
// t mod KT + 1, row t / KT of R rows) and marches a chunk of i planes.  In the
// step for plane i it computes beta of plane i+1 -- its own rows, and the row
// below and above the tile (the first and last rows' threads compute those
// too; a lateral halo row takes the periodic copy's values, as
// exchange_along_dim leaves them) -- into an LDS ring of four planes, waits
// for the block, then applies the limited donor pass to plane i, reading the
// factors of planes i-1, i, i+1 from the ring.  The inputs of beta for plane
// i+1 stay in registers for the donor pass of the next step.  Four planes in
// the ring make one barrier per step enough: the plane written in a step was
// last read two steps earlier.  The kernel reads psi_store, never psi, which
// it writes.
// ---------------------------------------------------------------------------
struct K4 {
    int KT, R, njb, ntiles;  // levels per column slot, rows, tiles
};
static K4 k4_geo(const PM &m, int rows) {
    K4 q;
    q.KT = ((m.kp + 31) / 32) * 32;
    q.R = std::max(1, std::min(rows, kK4Threads / q.KT));
    q.njb = (m.jp + q.R - 1) / q.R;
    q.ntiles = q.njb * m.nic;
    return q;
}

__global__ __launch_bounds__(kK4Threads) void k32_betalim(Mp32Args A, PM m, K4 G, Tab32 T) {
    // LDS ring: [plane & 3][bu, bd][row slot 0..R+1][level 0..KT-1]
    extern __shared__ float ring[];
    int t;
    if (!xcd_item(A.nf * G.ntiles, t)) return;
    const int f = t / G.ntiles;
    const int tile = t - f * G.ntiles;
    const int L = A.lslot[f];
    if (A.flag[L]) return;
    const int lf = A.lfield[L];
    const int jb = tile % G.njb, icb = tile / G.njb;
    const int r = (int)threadIdx.x / G.KT;
    const int k = (int)threadIdx.x % G.KT + 1;
    const int j0 = jb * G.R + 1;
    const int j = j0 + r;
    const int jlast = min(j0 + G.R - 1, m.jp);
    const int i0 = icb * m.ic + 1;
    const int i1 = min(i0 + m.ic - 1, m.ip);
    const bool act = r < G.R && k <= m.kp && j <= jlast;
    const int KT = G.KT, RS = G.R + 2, PL = 2 * RS * KT;
    auto slot = [&](int p, int a, int rs, int kk) -> float & {
        return ring[(p & 3) * PL + a * RS * KT + rs * KT + kk];
    };
    const float *sto = store_of(A, L);
    const float *lp1 = psi1_of(A, lf);
    const float *us = ustar_of(A, L, 0), *vs = ustar_of(A, L, 1), *ws = ustar_of(A, L, 2);
    const float *p1 = psi1_of(A, f);
    double *psi = A.psi[f];
    const double mm = A.m[f];
    const float dt = A.dt;
    const int SI = m.si, SJ = m.sj, SI4 = m.si4, SJ4 = m.sj4;
    // per-thread constants (inactive threads read valid table entries)
    const int kk = act ? k : 1;
    const int jj = act ? j : 1;
    const float rnk = T.rn[kk];
    const float czp = T.ra[kk] * T.idzn[kk], czm = T.ra[kk - 1] * T.idzn[kk];
    const float cyd = dt * T.idyn[jj];
    const float czr = ((dt * T.ra[kk]) * T.idzn[kk]) * T.irn[kk];
    const float czl = ((dt * T.ra[kk - 1]) * T.idzn[kk]) * T.irn[kk];
    const double zb = (A.lbc[f] - mm) + mm, zt = (A.ubc[f] - mm) + mm;
    const int nm = A.neumann;
    // the halo rows this thread also computes: the row below the tile (first
    // row) and above it (last row), as periodic copies
    const bool below = act && j == j0;
    const bool above = act && j == jlast;
    const int jbelow = wrapi(j0 - 1, m.jp), jabove = wrapi(jlast + 1, m.jp);
    const int rs_above = jlast - j0 + 2;
    const BetaSrc S(sto, lp1, us, vs, ws, SI4, SJ4);
    // beta at plane p, row jr, every input loaded
    auto beta_at = [&](int p, int jr, BetaIn &x, float &bu, float &bd) {
        const int pp = wrapi(p, m.ip);
        const unsigned o = (unsigned)(pp * SI4 + jr * SJ4 + k + kK0) * 4u;
        beta_load_full(x, S, o);
        beta_cell(x, rnk, rnk * T.xpk[pp].y, rnk * T.idyn[jr], czp, czm, dt, bu, bd);
    };
    auto halo_rows = [&](int p) {
        BetaIn x;
        float bu, bd;
        if (below) {
            beta_at(p, jbelow, x, bu, bd);
            slot(p, 0, 0, k - 1) = bu;
            slot(p, 1, 0, k - 1) = bd;
        }
        if (above) {
            beta_at(p, jabove, x, bu, bd);
            slot(p, 0, rs_above, k - 1) = bu;
            slot(p, 1, rs_above, k - 1) = bd;
        }
    };
    BetaIn cur, nxt;
    // prologue: beta of planes i0-1 (the centre of plane i0-1 is all the
    // donor pass reads of it) and i0
    if (act) {
        float bu, bd;
        beta_at(i0 - 1, j, cur, bu, bd);
        slot(i0 - 1, 0, r + 1, k - 1) = bu;
        slot(i0 - 1, 1, r + 1, k - 1) = bd;
        beta_at(i0, j, cur, bu, bd);
        slot(i0, 0, r + 1, k - 1) = bu;
        slot(i0, 1, r + 1, k - 1) = bd;
        halo_rows(i0);
    }
    const bool jedge = j == 1 || j == m.jp;
    int c8 = i0 * SI + j * SJ + k;
    int c4 = i0 * SI4 + j * SJ4 + k + kK0;
    unsigned o4n = (unsigned)(c4 + SI4) * 4u;  // plane i+1 of the own column
    float fxl = 0.f;
    for (int i = i0; i <= i1; ++i) {
        // beta of plane i+1
        if (act) {
            float bu, bd;
            if (i + 1 <= m.ip) {
                // the plane before i+1 is cur's: its centre values are the
                // west values of plane i+1
                beta_load(nxt, S, o4n);
                nxt.qw = cur.qc;
                nxt.ow = cur.oc;
                nxt.uw = cur.uc;
                beta_cell(nxt, rnk, rnk * T.xpk[i + 1].y, rnk * T.idyn[j], czp, czm, dt, bu, bd);
            } else {
                beta_at(i + 1, j, nxt, bu, bd);
            }
            slot(i + 1, 0, r + 1, k - 1) = bu;
            slot(i + 1, 1, r + 1, k - 1) = bd;
            if (i < i1) halo_rows(i + 1);
        }
        __syncthreads();
        if (act) {
            // the donor pass of field f at plane i with the limited velocities
            const float buc = slot(i, 0, r + 1, k - 1), bdc = slot(i, 1, r + 1, k - 1);
            const float bue = slot(i + 1, 0, r + 1, k - 1), bde = slot(i + 1, 1, r + 1, k - 1);
            const float bun = slot(i, 0, r + 2, k - 1), bdn = slot(i, 1, r + 2, k - 1);
            const float bus = slot(i, 0, r, k - 1), bds = slot(i, 1, r, k - 1);
            // the factors vanish on the z planes (lbc = ubc = 0)
            const float buu = k < m.kp ? slot(i, 0, r + 1, k) : 0.f;
            const float bdu = k < m.kp ? slot(i, 1, r + 1, k) : 0.f;
            const float uc = limited(cur.uc, bdc, buc, bue, bde);
            const float vc = limited(cur.vc, bdc, buc, bun, bdn);
            const float vl = limited(cur.vl, bds, bus, buc, bdc);
            const float wc = limited(cur.wc, bdc, buc, buu, bdu);
            float wl = 0.f;  // the face below k = 1 is the lbc = 0 plane
            if (k > 1)
                wl = limited(cur.wl, slot(i, 1, r + 1, k - 2), slot(i, 0, r + 1, k - 2), buc,
                             bdc);
            // psi1 of field f (for the category's first field: the values
            // beta read)
            float pc = cur.oc, pw = cur.ow, pe = cur.oe, pn = cur.on, ps = cur.os, pu = cur.ou,
                  pd = cur.od;
            if (f != lf) {
                pc = p1[c4];
                pw = p1[c4 - SI4];
                pe = p1[c4 + SI4];
                pn = p1[c4 + SJ4];
                ps = p1[c4 - SJ4];
                pu = p1[c4 + 1];
                pd = p1[c4 - 1];
            }
            if (i == i0) {
                const float ul = limited(cur.uw, slot(i - 1, 1, r + 1, k - 1),
                                         slot(i - 1, 0, r + 1, k - 1), buc, bdc);
                fxl = pos0(ul) * pw + neg0(ul) * pc;
            }
            const float fxr = pos0(uc) * pc + neg0(uc) * pe;
            const float fyr = pos0(vc) * pc + neg0(vc) * pn;
            const float fyl = pos0(vl) * ps + neg0(vl) * pc;
            float fzr = (pos0(wc) * pc + neg0(wc) * pu) * czr;
            float fzl = (pos0(wl) * pd + neg0(wl) * pc) * czl;
            if (nm == 1) {
                if (k == 1) fzl = fzr;
                if (k == m.kp) fzr = fzl;
            } else if (nm == 2) {
                if (k == 1) fzl = -fzr;
                if (k == m.kp) fzr = -fzl;
            }
            const float cx = dt * T.xpk[i].y;
            const float res = ((pc - (fxr - fxl) * cx) - (fyr - fyl) * cyd) - (fzr - fzl);
            const double rv = (double)res + mm;
            psi[c8] = rv;
            if (k == 1) psi[c8 - 1] = zb;
            if (k == m.kp) psi[c8 + 1] = zt;
            if (jedge || i == 1 || i == m.ip) put_images<double>(psi, c8, rv, k, j, i, m, SJ, SI);
            fxl = fxr;
            cur = nxt;
        }
        c8 += SI;
        c4 += SI4;
        o4n += (unsigned)SI4 * 4u;
    }
}

// ---------------------------------------------------------------------------
// Separate limiter kernels for columns of more than 256 levels: beta of each
// category's first field into two aligned arrays, then the last donor pass.
// Without the limiter the last donor pass runs with U*, V*, W* themselves.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k32_beta(Mp32Args A, PM m, Tab32 T) {
    Col C;
    if (!pm_col(m, A.nlead, C)) return;
    const int L = C.f;
    if (A.flag[L]) return;
    if (!C.act) return;
    const int fld = A.lfield[L];
    const float *sto = store_of(A, L);
    const float *p1 = psi1_of(A, fld);
    const float *us = ustar_of(A, L, 0), *vs = ustar_of(A, L, 1), *ws = ustar_of(A, L, 2);
    float *bu = beta_of(A, L, 0), *bd = beta_of(A, L, 1);
    const int k = C.k, j = C.j;
    const BetaSrc S(sto, p1, us, vs, ws, m.si4, m.sj4);
    const float dt = A.dt;
    const float rnk = T.rn[k];
    const float cy = rnk * T.idyn[j];
    const float czp = T.ra[k] * T.idzn[k], czm = T.ra[k - 1] * T.idzn[k];
    const bool jedge = j == 1 || j == m.jp;
    for (int i = C.i0; i <= C.i1; ++i) {
        const int c4 = i * m.si4 + j * m.sj4 + k + kK0;
        BetaIn x;
        beta_load_full(x, S, (unsigned)c4 * 4u);
        float vu, vd;
        beta_cell(x, rnk, rnk * T.xpk[i].y, cy, czp, czm, dt, vu, vd);
        bu[c4] = vu;
        bd[c4] = vd;
        if (k == 1) {
            bu[c4 - 1] = 0.f;
            bd[c4 - 1] = 0.f;
        }
        if (k == m.kp) {
            bu[c4 + 1] = 0.f;
            bd[c4 + 1] = 0.f;
        }
        if (jedge || i == 1 || i == m.ip) {
            put_faces(bu, c4, vu, k, j, i, m);
            put_faces(bd, c4, vd, k, j, i, m);
        }
    }
}

template <bool LIM>
__global__ __launch_bounds__(256) void k32_limited(Mp32Args A, PM m, Tab32 T) {
    Col C;
    if (!pm_col(m, A.nf, C)) return;
    const int f = C.f;
    const int L = A.lslot[f];
    if (A.flag[L]) return;
    if (!C.act) return;
    double *psi = A.psi[f];
    const double mm = A.m[f];
    const float *p1 = psi1_of(A, f);
    const float *us = ustar_of(A, L, 0), *vs = ustar_of(A, L, 1), *ws = ustar_of(A, L, 2);
    const float *bu = LIM ? beta_of(A, L, 0) : us, *bd = LIM ? beta_of(A, L, 1) : us;
    const int k = C.k, j = C.j, SI = m.si, SJ = m.sj, SI4 = m.si4, SJ4 = m.sj4;
    const Ld<float> Pe(p1 + SI4), Pn(p1 + SJ4), Ps(p1 - SJ4), Pc(p1);
    const Ld<float> Uc(us), Vc(vs), Vl(vs - SJ4), Wc(ws);
    const Ld<float> Bue(bu + SI4), Bde(bd + SI4), Bun(bu + SJ4), Bdn(bd + SJ4), Bus(bu - SJ4),
        Bds(bd - SJ4), Buc(bu), Bdc(bd);
    const St<double> S(psi);
    const float dt = A.dt;
    const float cy = dt * T.idyn[j];
    const float czr = ((dt * T.ra[k]) * T.idzn[k]) * T.irn[k];
    const float czl = ((dt * T.ra[k - 1]) * T.idzn[k]) * T.irn[k];
    const double zb = (A.lbc[f] - mm) + mm, zt = (A.ubc[f] - mm) + mm;
    const int nm = A.neumann;
    const bool jedge = j == 1 || j == m.jp;
    int c = C.i0 * SI + j * SJ + k;
    int c4 = C.i0 * SI4 + j * SJ4 + k + kK0;
    unsigned o4 = (unsigned)c4 * 4u, o8 = (unsigned)c * 8u;
    float pw = p1[c4 - SI4], pc = p1[c4];
    float buc = 0.f, bdc = 0.f;
    float ul = us[c4 - SI4];
    if (LIM) {
        buc = bu[c4];
        bdc = bd[c4];
        ul = limited(ul, bd[c4 - SI4], bu[c4 - SI4], buc, bdc);
    }
    float fxl = pos0(ul) * pw + neg0(ul) * pc;
    for (int i = C.i0; i <= C.i1; ++i) {
        const float pe = Pe.at(o4), pn = Pn.at(o4), ps = Ps.at(o4), pu = Pc.at(o4, 1),
                    pd = Pc.at(o4, -1);
        float uc = Uc.at(o4), vc = Vc.at(o4), vl = Vl.at(o4), wc = Wc.at(o4), wl = Wc.at(o4, -1);
        float bue = 0.f, bde = 0.f;
        if (LIM) {
            bue = Bue.at(o4);
            bde = Bde.at(o4);
            const float bun = Bun.at(o4), bdn = Bdn.at(o4), bus = Bus.at(o4), bds = Bds.at(o4);
            const float buu = Buc.at(o4, 1), bdu = Bdc.at(o4, 1), bud = Buc.at(o4, -1),
                        bdd = Bdc.at(o4, -1);
            uc = limited(uc, bdc, buc, bue, bde);
            vc = limited(vc, bdc, buc, bun, bdn);
            vl = limited(vl, bds, bus, buc, bdc);
            wc = limited(wc, bdc, buc, buu, bdu);
            wl = k == 1 ? 0.f : limited(wl, bdd, bud, buc, bdc);
        } else if (k == 1) {
            wl = 0.f;  // the lbc = 0 plane of the antidiffusive w
        }
        const float fxr = pos0(uc) * pc + neg0(uc) * pe;
        const float fyr = pos0(vc) * pc + neg0(vc) * pn;
        const float fyl = pos0(vl) * ps + neg0(vl) * pc;
        float fzr = (pos0(wc) * pc + neg0(wc) * pu) * czr;
        float fzl = (pos0(wl) * pd + neg0(wl) * pc) * czl;
        if (nm == 1) {
            if (k == 1) fzl = fzr;
            if (k == m.kp) fzr = fzl;
        } else if (nm == 2) {
            if (k == 1) fzl = -fzr;
            if (k == m.kp) fzr = -fzl;
        }
        const float cx = dt * T.xpk[i].y;
        const float res = ((pc - (fxr - fxl) * cx) - (fyr - fyl) * cy) - (fzr - fzl);
        const double rv = (double)res + mm;
        S.put(o8, rv);
        if (k == 1) S.put(o8, zb, -1);
        if (k == m.kp) S.put(o8, zt, 1);
        if (jedge || i == 1 || i == m.ip) put_images<double>(psi, c, rv, k, j, i, m, SJ, SI);
        fxl = fxr;
        pw = pc;
        pc = pe;
        buc = bue;
        bdc = bde;
        c += SI;
        o4 += (unsigned)SI4 * 4u;
        o8 += (unsigned)SI * 8u;
    }
}

// zero-sum exit (:284-289): psi = (psi - m) + m over the whole array
__global__ void k32_exit_unshift(Mp32Args A, long long n) {
    const int f = blockIdx.y;
    if (!A.flag[A.lslot[f]]) return;
    double *psi = A.psi[f];
    const double mm = A.m[f];
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        psi[t] = (psi[t] - mm) + mm;
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
static int ensure_mp32(dcm_ctx *c, size_t count, size_t ntot4) {
    if (c->mp32_count >= count * ntot4) return DCM_OK;
    DCM_CHECK(no_alloc_in_capture(c, "mpdata32 scratch"));
    if (c->mp32) {
        DCM_HIP(hipStreamSynchronize(c->stream));
        DCM_HIP(hipFree(c->mp32));
        c->mp32 = nullptr;
        c->mp32_count = 0;
    }
    void *p = nullptr;
    const size_t bytes = count * ntot4 * sizeof(float);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        set_error("hipMalloc of the FP32 MPDATA scratch failed");
        return DCM_OUT_OF_MEMORY;
    }
    // zero once: halo cells no pass writes (the lateral edges of the limiter
    // factors, the padding between aligned columns) then hold a defined value
    DCM_HIP(hipMemsetAsync(p, 0, bytes, c->stream));
    c->mp32 = (float *)p;
    c->mp32_count = count * ntot4;
    return DCM_OK;
}

// rows of a K4 tile (the last kernel's block holds whole columns)
constexpr int kK4Rows = 8;

// one batched launch sequence over at most kMp32Max fields
static int mpdata32_run(dcm_ctx *c, float dt, const double *u, const double *v,
                        const double *w, const Mp32Cat *cats, int ncat, int monotone,
                        int neumann) {
    const Geo &g = c->geo;
    Mp32Args A{};
    int nf = 0, nlead = 0;
    for (int q = 0; q < ncat; ++q) {
        A.lfield[nlead] = (signed char)nf;
        for (int n = 0; n < cats[q].nf; ++n) {
            A.psi[nf] = cats[q].psi + (size_t)n * g.ntot;
            A.lbc[nf] = cats[q].lbc[n];
            A.ubc[nf] = cats[q].ubc[n];
            A.lslot[nf] = (signed char)nlead;
            A.whole[nf] = (signed char)(cats[q].whole_min ? 1 : 0);
            ++nf;
        }
        ++nlead;
    }
    const PM m = pm_geo(g, kMp32Planes);
    const K4 G = k4_geo(m, kK4Rows);
    const bool fused = monotone && m.kp <= kK4MaxKt;
    const size_t ntot4 = (size_t)ntot4_of(m);
    DCM_CHECK(ensure_mp32(c, 3 + (size_t)nf + 6 * (size_t)nlead, ntot4));
    const size_t npart = std::max((size_t)kMinBlocks * nf, (size_t)m.ntiles * nlead);
    DCM_CHECK(ensure_partials_adv(c, npart + 16));
    A.s32 = c->mp32;
    A.ntot4 = (long long)ntot4;
    A.m = c->dscal + 16;
    A.flag = (int *)(c->dscal + 48);
    A.partials = c->partials_adv;
    A.nf = nf;
    A.nlead = nlead;
    A.dt = dt;
    A.neumann = neumann;
    const Tab32 &T = c->tab32;
    float *s = c->mp32;
    hipLaunchKernelGGL(k32_shadow, dim3((g.nk + 63) / 64, (g.nj + 3) / 4, g.ni), dim3(256), 0,
                       c->stream, u, v, w, s, s + ntot4, s + 2 * ntot4, m, g.nk);
    int tok = prof_begin(c, "mpdata32_min");
    hipLaunchKernelGGL(k32_min, dim3(kMinBlocks, nf), dim3(kBlock), 0, c->stream, A, g);
    hipLaunchKernelGGL(k32_min_finish, dim3(nf), dim3(1024), 0, c->stream, A, kMinBlocks);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    launch_timed(c, "k32_donor1", k32_donor1, dim3(xcd_grid((long long)nf * m.ntiles)),
                 dim3(kBlock), A, m, T);
    hipLaunchKernelGGL(k32_exit_flags, dim3(nlead), dim3(1024), 0, c->stream, A, m.ntiles);
    DCM_LAUNCH_CHECK();
    launch_timed(c, "k32_antidiff", k32_antidiff, dim3(xcd_grid((long long)nlead * m.ntiles)),
                 dim3(kBlock), A, m, T);
    DCM_LAUNCH_CHECK();
    if (fused) {
        const unsigned threads = (unsigned)(((G.KT * G.R + 63) / 64) * 64);
        const size_t lds = 4 * 2 * (size_t)(G.R + 2) * G.KT * sizeof(float);
        const dim3 grid(xcd_grid((long long)nf * G.ntiles));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (prof_events(c, "k32_betalim", &e0, &e1))
            hipExtLaunchKernelGGL(k32_betalim, grid, dim3(threads), lds, c->stream, e0, e1, 0, A,
                                  m, G, T);
        else
            hipLaunchKernelGGL(k32_betalim, grid, dim3(threads), lds, c->stream, A, m, G, T);
    } else if (monotone) {
        launch_timed(c, "k32_beta", k32_beta, dim3(xcd_grid((long long)nlead * m.ntiles)),
                     dim3(kBlock), A, m, T);
        launch_timed(c, "k32_limited", k32_limited<true>,
                     dim3(xcd_grid((long long)nf * m.ntiles)), dim3(kBlock), A, m, T);
    } else {
        launch_timed(c, "k32_limited", k32_limited<false>,
                     dim3(xcd_grid((long long)nf * m.ntiles)), dim3(kBlock), A, m, T);
    }
    DCM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k32_exit_unshift, dim3(256, nf), dim3(kBlock), 0, c->stream, A, g.ntot);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

bool mpdata32_fast(const dcm_ctx *c, int kord) {
    const Geo &g = c->geo;
    const PM m = pm_geo(g, kMp32Planes);
    // byte offsets of the aligned FP32 arrays and of the FP64 fields fit 32 bits
    const bool fits = (long long)ntot4_of(m) * 4 < (1LL << 32) && g.ntot * 8 < (1LL << 32);
    return kord == 2 && c->mpdata_variant == 0 && fits;
}

int mpdata32_batch(dcm_ctx *c, float dt, const double *u, const double *v, const double *w,
                   float *shadow, const Mp32Cat *cats, int ncat, int kord, int monotone,
                   int neumann) {
    const Geo &g = c->geo;
    const bool fast = mpdata32_fast(c, kord);
    float *uf = shadow, *vf = shadow + g.ntot, *wf = shadow + 2 * g.ntot;
    bool shadowed = false;
    auto generic = [&](const Mp32Cat &q) -> int {
        if (!shadowed) {
            DCM_CHECK(velocity_shadow(c, u, v, w, uf, vf, wf));
            shadowed = true;
        }
        return mpdata_core<float>(c, dt, uf, vf, wf, q.psi, q.nf, q.lbc, q.ubc, kord, monotone,
                                  neumann, q.whole_min ? 0 : 1);
    };
    if (!fast) {
        for (int q = 0; q < ncat; ++q) DCM_CHECK(generic(cats[q]));
        return DCM_OK;
    }
    // batches of whole categories, at most kMp32Max fields each; the shift
    // of field f of a batch is dscal[16 + f]
    int q = 0;
    while (q < ncat) {
        if (cats[q].nf > kMp32Max) {
            DCM_CHECK(generic(cats[q]));
            ++q;
            continue;
        }
        int q1 = q, nf = 0;
        while (q1 < ncat && cats[q1].nf <= kMp32Max && nf + cats[q1].nf <= kMp32Max)
            nf += cats[q1++].nf;
        DCM_CHECK(mpdata32_run(c, dt, u, v, w, cats + q, q1 - q, monotone, neumann));
        q = q1;
    }
    return DCM_OK;
}

}  // namespace dcm
