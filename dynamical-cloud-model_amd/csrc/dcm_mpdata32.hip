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
//   * a thread owns one (k,j) column of a 64x4 or 32x8 tile and marches a
//     chunk of i planes, keeping the column sums of planes i-1, i, i+1 in
//     registers (the flux through the lower x face of a cell is the upper x
//     face's flux of the cell before it);
//   * each pass writes, besides a cell, the halo cells that the reference's
//     exchange would copy from it (put_full / put_faces): no fill kernel runs;
//   * each pass serves every field (or every category's first field) in one
//     launch: four passes, a minimum and two small reductions per step
//     instead of about a dozen launches per field.
// Offsets are 32-bit: a padded field is below 2^31 elements (checked).
//
//   K1 k32_min / k32_min_finish   m(f) = min(psi_f) (+ lbc, ubc)
//   K2 k32_donor1                 psi1_f = donor(psi_f - m), halo images;
//                                 block sums of psi - m for the exit test
//      k32_exit_flags             flag(L) = (sum < small), per category
//   K3 k32_antidiff               U*, V*, W* of each category (first field)
//   K4 k32_beta                   beta_up, beta_down of each category
//   K5 k32_limited                psi_f = donor(psi1_f, limited U*) + m
//      k32_exit_unshift           psi = (psi - m) + m where the exit was taken
#include <cfloat>
#include <vector>

#include "dcm_device.h"

// The library is built without multiply-add contraction, which the FP64
// path's bit parity needs; these FP32 kernels contract freely.
#pragma clang fp contract(fast)

namespace dcm {

constexpr int kMp32Max = 16;   // fields per batched launch
constexpr int kMp32Planes = 16; // i planes per marching tile
constexpr int kMinBlocks = 1024;

// marching over the padded layout of one field
struct PM {
    int ip, jp, kp;
    int sj, si;                        // padded strides
    int kt, jt, nkb, njb, ic, nic;     // tile lanes (k), rows (j); tiles; i chunks
    int ntiles;                        // tiles per field
};

static PM pm_geo(const Geo &g, int ic) {
    PM m;
    m.ip = g.ip;
    m.jp = g.jp;
    m.kp = g.kp;
    m.sj = (int)g.sj;
    m.si = (int)g.si;
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

// Block -> (field of the launch, tile).  The (field, tile) list is split
// into 8 contiguous ranges, one per XCD (block b runs on XCD b mod 8), so
// tiles that share faces share an L2.
struct Col {
    int f, tile, k, j, i0, i1;
    bool act;
};
__device__ __forceinline__ bool pm_col(const PM &m, int nf, Col &C) {
    const int total = nf * m.ntiles;
    const int per = (total + 7) >> 3;
    const int b = (int)blockIdx.x;
    const int t = (b & 7) * per + (b >> 3);
    if (t >= total) return false;
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
static unsigned pm_grid(const PM &m, int nf) {
    const long long total = (long long)nf * m.ntiles;
    return (unsigned)(((total + 7) / 8) * 8);
}

// The batch: fields f = 0..nf-1; category leaders (first fields) L =
// 0..nlead-1.  Scratch: psi1 of field f at s32 + f*ntot; U*, V*, W*,
// beta_up, beta_down of leader L at s32 + (nf + 5L + 0..4)*ntot.
struct Mp32Args {
    double *psi[kMp32Max];
    double lbc[kMp32Max], ubc[kMp32Max];
    signed char lslot[kMp32Max];   // leader slot of field f
    signed char lfield[kMp32Max];  // field of leader slot L
    signed char whole[kMp32Max];   // minimum over the whole array (mpdata_3d)
    float *s32;
    long long ntot;
    double *m;                     // shifts, m[f]
    int *flag;                     // zero-sum exit, flag[L]
    double *partials;              // min partials (K1), exit sums (K2)
    int nf, nlead;
    float dt;
    int neumann;
};

template <typename T> __device__ __forceinline__ T pos0(T x) { return x > T(0) ? x : T(0); }
template <typename T> __device__ __forceinline__ T neg0(T x) { return x < T(0) ? x : T(0); }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float min3f(float a, float b, float c) {
    return fminf(fminf(a, b), c);
}

// A uniform base pointer read at a 32-bit byte offset: the load takes the
// base from scalar registers and the offset from one vector register, so a
// stencil of loads costs no address arithmetic (every neighbour has a base
// of its own, shifted by a constant or a stride).
template <typename T> struct Ld {
    const char *b;
    __device__ __forceinline__ explicit Ld(const T *p) : b((const char *)p) {}
    // element dk along k (an immediate) from the byte offset off
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

// The halo cells that exchange_full copies from interior cell (k,j,i)
// (offset c): the periodic images on the lateral faces and edges, and on
// the z planes the z-wrapped edges (the receive from MPI_PROC_NULL keeps the
// local copy, mpi_module.f90:1244-1263): a z-plane cell outside the lateral
// interior holds level kp at k = 0 and level 1 at kp+1.  The lateral-interior
// z-plane cells are stored by the caller (k = 1 and k = kp lanes).
template <typename T>
__device__ __forceinline__ void put_images(T *a, int c, T v, int k, int j, int i, const PM &m) {
    int dj[3], di[3];
    int nj = 1, ni = 1;
    dj[0] = 0;
    di[0] = 0;
    if (j == 1) dj[nj++] = m.jp * m.sj;    // j = jp+1 copies j = 1
    if (j == m.jp) dj[nj++] = -m.jp * m.sj; // j = 0 copies j = jp
    if (i == 1) di[ni++] = m.ip * m.si;
    if (i == m.ip) di[ni++] = -m.ip * m.si;
    for (int a1 = 0; a1 < nj; ++a1)
        for (int b1 = 0; b1 < ni; ++b1) {
            if (a1 == 0 && b1 == 0) continue;
            const int o = c + dj[a1] + di[b1];
            a[o] = v;
            if (k == m.kp) a[o - m.kp] = v;  // level 0 of a lateral halo column
            if (k == 1) a[o + m.kp] = v;     // level kp+1
        }
}
// exchange_along_dim with lbc = ubc = 0 (the limiter factors, :624-627):
// periodic faces (the edges are left alone: nothing reads them) and zero z
// planes over the whole plane, lateral halo included.
__device__ __forceinline__ void put_faces(float *a, int c, float v, int k, int j, int i,
                                       const PM &m) {
    int dj[3], di[3];
    int nj = 1, ni = 1;
    dj[0] = 0;
    di[0] = 0;
    if (j == 1) dj[nj++] = m.jp * m.sj;
    if (j == m.jp) dj[nj++] = -m.jp * m.sj;
    if (i == 1) di[ni++] = m.ip * m.si;
    if (i == m.ip) di[ni++] = -m.ip * m.si;
    for (int a1 = 0; a1 < nj; ++a1)
        for (int b1 = 0; b1 < ni; ++b1) {
            if (a1 == 0 && b1 == 0) continue;
            const int o = c + dj[a1] + di[b1];
            if (a1 == 0 || b1 == 0) a[o] = v;  // a face
            if (k == 1) a[o - 1] = 0.f;
            if (k == m.kp) a[o + 1] = 0.f;
        }
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
// exchange_full(psi1, lbc - m, ubc - m).  Leaders also store block sums of
// psi - m over the interior (the exit test, :282-289 / :1314-1315).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k32_donor1(Mp32Args A, PM m, Tab32 T,
                                                  const float *__restrict__ u,
                                                  const float *__restrict__ v,
                                                  const float *__restrict__ w) {
    Col C;
    if (!pm_col(m, A.nf, C)) return;
    const int f = C.f;
    const int L = A.lslot[f];
    const bool lead = A.lfield[L] == f;
    const double *psi = A.psi[f];
    float *out = A.s32 + (size_t)f * A.ntot;
    const double mm = A.m[f];
    double s[1] = {0.0};
    if (C.act) {
        const int k = C.k, j = C.j, SI = m.si, SJ = m.sj;
        const Ld<double> Pe(psi + SI), Pn(psi + SJ), Ps(psi - SJ), Pc(psi);
        const Ld<float> Uc(u), Vc(v), Vs(v - SJ), Wc(w);
        const St<float> O(out);
        const float dt = A.dt;
        const float cy = dt * T.idyn[j];
        const float czr = ((dt * T.ra[k]) * T.idzn[k]) * T.irn[k];
        const float czl = ((dt * T.ra[k - 1]) * T.idzn[k]) * T.irn[k];
        const float zb = (float)(A.lbc[f] - mm), zt = (float)(A.ubc[f] - mm);
        const int nm = A.neumann;
        const bool jedge = j == 1 || j == m.jp;
        int c = C.i0 * SI + j * SJ + k;
        unsigned o4 = (unsigned)c * 4u, o8 = (unsigned)c * 8u;
        float pw = (float)(psi[c - SI] - mm);
        double dc = psi[c] - mm;
        float pc = (float)dc;
        const float uw = u[c - SI];
        float fxl = pos0(uw) * pw + neg0(uw) * pc;
        for (int i = C.i0; i <= C.i1; ++i) {
            const double de = Pe.at(o8) - mm;
            const float pe = (float)de;
            const float pn = (float)(Pn.at(o8) - mm), ps = (float)(Ps.at(o8) - mm);
            const float pu = (float)(Pc.at(o8, 1) - mm), pd = (float)(Pc.at(o8, -1) - mm);
            const float uc = Uc.at(o4), vc = Vc.at(o4), vl = Vs.at(o4), wc = Wc.at(o4),
                        wl = Wc.at(o4, -1);
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
            if (jedge || i == 1 || i == m.ip) put_images<float>(out, c, res, k, j, i, m);
            if (lead) s[0] += dc;
            fxl = fxr;
            pc = pe;
            dc = de;
            c += SI;
            o4 += (unsigned)SI * 4u;
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

__global__ __launch_bounds__(256) void k32_antidiff(Mp32Args A, PM m, Tab32 T,
                                                    const float *__restrict__ u,
                                                    const float *__restrict__ v,
                                                    const float *__restrict__ w) {
    Col C;
    if (!pm_col(m, A.nlead, C)) return;
    const int L = C.f;
    if (A.flag[L]) return;
    if (!C.act) return;
    const int fld = A.lfield[L];
    const float *p1 = A.s32 + (size_t)fld * A.ntot;
    float *us = A.s32 + (size_t)(A.nf + 5 * L) * A.ntot;
    float *vs = us + A.ntot;
    float *ws = vs + A.ntot;
    const int k = C.k, j = C.j, SI = m.si, SJ = m.sj;
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
    int c = C.i0 * SI + j * SJ + k;
    unsigned o4 = (unsigned)c * 4u;
    Plane pw, pc, pn;
    plane_at(pw, p1, u, v, w, c - SI, SJ, rnk, rnk1);
    plane_at(pc, p1, u, v, w, c, SJ, rnk, rnk1);
    for (int i = C.i0; i <= C.i1; ++i) {
        pn.A = nP.at(o4, 1);
        pn.B = nP.at(o4, -1);
        pn.Cn = nPn.at(o4);
        pn.Dm = nPs.at(o4);
        pn.O = nP.at(o4);
        pn.U0 = nU.at(o4);
        pn.Uy = nUj.at(o4) + pn.U0;
        pn.Uz = rnk1 * nU.at(o4, 1) + rnk * pn.U0;
        pn.V0 = nV.at(o4);
        pn.Vm = nVm.at(o4);
        pn.W0 = nW.at(o4);
        pn.Wm = nW.at(o4, -1);
        const float pkj = Pj.at(o4, 1), pmj = Pj.at(o4, -1), pkm = Pm.at(o4, 1);
        const float vk = Vc.at(o4, 1), vkm = Vm.at(o4, 1), vj = Vj.at(o4);
        const float wj = Wj.at(o4), wmj = Wj.at(o4, -1), wk = Wc.at(o4, 1);
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
            put_images<float>(us, c, usv, k, j, i, m);
            put_images<float>(vs, c, vsv, k, j, i, m);
            put_images<float>(ws, c, wsv, k, j, i, m);
        }
        pw = pc;
        pc = pn;
        c += SI;
        o4 += (unsigned)SI * 4u;
    }
}

// ---------------------------------------------------------------------------
// K4: limiter factors (eqs 19-20 of Smolarkiewicz & Grabowski 1990,
// :549-616) of each category's first field: bounds over psi_store = psi - m
// (recomputed from the untouched FP64 field) and psi1 at the cell and its six
// neighbours, density-weighted in/outflow denominators.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k32_beta(Mp32Args A, PM m, Tab32 T) {
    Col C;
    if (!pm_col(m, A.nlead, C)) return;
    const int L = C.f;
    if (A.flag[L]) return;
    if (!C.act) return;
    const int fld = A.lfield[L];
    const double *psi = A.psi[fld];
    const double mm = A.m[fld];
    const float *p1 = A.s32 + (size_t)fld * A.ntot;
    float *const base = A.s32 + (size_t)(A.nf + 5 * L) * A.ntot;
    const float *us = base;
    const float *vs = base + A.ntot;
    const float *ws = base + 2 * A.ntot;
    float *bu = base + 3 * A.ntot;
    float *bd = base + 4 * A.ntot;
    const int k = C.k, j = C.j, SI = m.si, SJ = m.sj;
    const Ld<double> Qe(psi + SI), Qn(psi + SJ), Qs(psi - SJ), Qc(psi);
    const Ld<float> Oe(p1 + SI), On(p1 + SJ), Os(p1 - SJ), Oc(p1);
    const Ld<float> Uc(us), Vc(vs), Vl(vs - SJ), Wc(ws);
    const St<float> BU(bu), BD(bd);
    const float dt = A.dt, small = FLT_MIN;
    const float rnk = T.rn[k];
    const float cy = rnk * T.idyn[j];
    const float czp = T.ra[k] * T.idzn[k], czm = T.ra[k - 1] * T.idzn[k];
    const bool jedge = j == 1 || j == m.jp;
    int c = C.i0 * SI + j * SJ + k;
    unsigned o4 = (unsigned)c * 4u, o8 = (unsigned)c * 8u;
    float qw = (float)(psi[c - SI] - mm), qc = (float)(psi[c] - mm);
    float ow = p1[c - SI], oc = p1[c];
    float uw = us[c - SI];
    for (int i = C.i0; i <= C.i1; ++i) {
        const float qe = (float)(Qe.at(o8) - mm);
        const float qn = (float)(Qn.at(o8) - mm), qs = (float)(Qs.at(o8) - mm);
        const float qu = (float)(Qc.at(o8, 1) - mm), qd = (float)(Qc.at(o8, -1) - mm);
        const float oe = Oe.at(o4), on = On.at(o4), os = Os.at(o4), ou = Oc.at(o4, 1),
                    od = Oc.at(o4, -1);
        const float uc = Uc.at(o4), vc = Vc.at(o4), vl = Vl.at(o4), wc = Wc.at(o4),
                    wl = Wc.at(o4, -1);
        float mx = fmaxf(fmaxf(qc, qw), fmaxf(qe, qn));
        mx = fmaxf(fmaxf(mx, qs), fmaxf(qu, qd));
        mx = fmaxf(fmaxf(mx, oc), fmaxf(ow, oe));
        mx = fmaxf(fmaxf(mx, on), fmaxf(os, fmaxf(ou, od)));
        float mn = fminf(fminf(qc, qw), fminf(qe, qn));
        mn = fminf(fminf(mn, qs), fminf(qu, qd));
        mn = fminf(fminf(mn, oc), fminf(ow, oe));
        mn = fminf(fminf(mn, on), fminf(os, fminf(ou, od)));
        const float cx = rnk * T.xpk[i].y;
        const float d1 = dt * ((cx * (pos0(uw) * ow - neg0(uc) * oe) +
                                cy * (pos0(vl) * os - neg0(vc) * on)) +
                               (czm * pos0(wl) * od - czp * neg0(wc) * ou)) +
                         small;
        const float d2 = dt * (oc * ((cx * (pos0(uc) - neg0(uw)) + cy * (pos0(vc) - neg0(vl))) +
                                     (czp * pos0(wc) - czm * neg0(wl))) +
                               small);
        const float vu = (rnk * (mx - oc)) * rcp(d1), vd = (rnk * (oc - mn)) * rcp(d2);
        BU.put(o4, vu);
        BD.put(o4, vd);
        if (k == 1) {
            BU.put(o4, 0.f, -1);
            BD.put(o4, 0.f, -1);
        }
        if (k == m.kp) {
            BU.put(o4, 0.f, 1);
            BD.put(o4, 0.f, 1);
        }
        if (jedge || i == 1 || i == m.ip) {
            put_faces(bu, c, vu, k, j, i, m);
            put_faces(bd, c, vd, k, j, i, m);
        }
        qw = qc;
        qc = qe;
        ow = oc;
        oc = oe;
        uw = uc;
        c += SI;
        o4 += (unsigned)SI * 4u;
        o8 += (unsigned)SI * 8u;
    }
}

// ---------------------------------------------------------------------------
// K5: donor pass of each field's psi1 with the limited velocities of eq 18
// (:634-697) evaluated at the six faces (or with U*, V*, W* themselves when
// the limiter is off), + m; psi with the halo of exchange_full(psi, lbc, ubc)
// after the unshift, i.e. z planes (bc - m) + m.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float limited(float x, float bd_c, float bu_c, float bu_o,
                                         float bd_o) {
    return min3f(1.f, bd_c, bu_o) * pos0(x) + min3f(1.f, bu_c, bd_o) * neg0(x);
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
    const float *p1 = A.s32 + (size_t)f * A.ntot;
    const float *us = A.s32 + (size_t)(A.nf + 5 * L) * A.ntot;
    const float *vs = us + A.ntot;
    const float *ws = vs + A.ntot;
    const float *bu = ws + A.ntot;
    const float *bd = bu + A.ntot;
    const int k = C.k, j = C.j, SI = m.si, SJ = m.sj;
    const Ld<float> Pe(p1 + SI), Pn(p1 + SJ), Ps(p1 - SJ), Pc(p1);
    const Ld<float> Uc(us), Vc(vs), Vl(vs - SJ), Wc(ws);
    const Ld<float> Bue(bu + SI), Bde(bd + SI), Bun(bu + SJ), Bdn(bd + SJ), Bus(bu - SJ),
        Bds(bd - SJ), Buc(bu), Bdc(bd);
    const St<double> S(psi);
    const float dt = A.dt;
    const float cy = dt * T.idyn[j];
    const float czr = ((dt * T.ra[k]) * T.idzn[k]) * T.irn[k];
    const float czl = ((dt * T.ra[k - 1]) * T.idzn[k]) * T.irn[k];
    const double zb = (A.lbc[f] - mm) + mm, zt = (A.ubc[f] - mm) + mm;
    const int nm = A.neumann;
    const bool jedge = j == 1 || j == m.jp;
    int c = C.i0 * SI + j * SJ + k;
    unsigned o4 = (unsigned)c * 4u, o8 = (unsigned)c * 8u;
    float pw = p1[c - SI], pc = p1[c];
    float buc = 0.f, bdc = 0.f;
    float ul = us[c - SI];
    if (LIM) {
        buc = bu[c];
        bdc = bd[c];
        ul = limited(ul, bd[c - SI], bu[c - SI], buc, bdc);
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
        const double r = (double)res + mm;
        S.put(o8, r);
        if (k == 1) S.put(o8, zb, -1);
        if (k == m.kp) S.put(o8, zt, 1);
        if (jedge || i == 1 || i == m.ip) put_images<double>(psi, c, r, k, j, i, m);
        fxl = fxr;
        pw = pc;
        pc = pe;
        buc = bue;
        bdc = bde;
        c += SI;
        o4 += (unsigned)SI * 4u;
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
static int ensure_mp32(dcm_ctx *c, size_t count) {
    if (c->mp32_count >= count) return DCM_OK;
    DCM_CHECK(no_alloc_in_capture(c, "mpdata32 scratch"));
    if (c->mp32) {
        DCM_HIP(hipStreamSynchronize(c->stream));
        DCM_HIP(hipFree(c->mp32));
        c->mp32 = nullptr;
        c->mp32_count = 0;
    }
    void *p = nullptr;
    const size_t bytes = count * (size_t)c->geo.ntot * sizeof(float);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        set_error("hipMalloc of the FP32 MPDATA scratch failed");
        return DCM_OUT_OF_MEMORY;
    }
    // zero once: halo cells no pass writes (the lateral edges of the limiter
    // factors) then hold a defined value
    DCM_HIP(hipMemsetAsync(p, 0, bytes, c->stream));
    c->mp32 = (float *)p;
    c->mp32_count = count;
    return DCM_OK;
}

// one batched launch sequence over at most kMp32Max fields
static int mpdata32_run(dcm_ctx *c, float dt, const float *u, const float *v, const float *w,
                        const Mp32Cat *cats, int ncat, int monotone, int neumann) {
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
    DCM_CHECK(ensure_mp32(c, (size_t)nf + 5 * (size_t)nlead));
    const size_t npart = std::max((size_t)kMinBlocks * nf, (size_t)m.ntiles * nlead);
    DCM_CHECK(ensure_partials(c, npart + 16));
    A.s32 = c->mp32;
    A.ntot = g.ntot;
    A.m = c->dscal + 16;
    A.flag = (int *)(c->dscal + 48);
    A.partials = c->partials;
    A.nf = nf;
    A.nlead = nlead;
    A.dt = dt;
    A.neumann = neumann;
    const Tab32 &T = c->tab32;
    int tok = prof_begin(c, "mpdata32_min");
    hipLaunchKernelGGL(k32_min, dim3(kMinBlocks, nf), dim3(kBlock), 0, c->stream, A, g);
    hipLaunchKernelGGL(k32_min_finish, dim3(nf), dim3(1024), 0, c->stream, A, kMinBlocks);
    DCM_LAUNCH_CHECK();
    prof_end(c, tok);
    launch_timed(c, "k32_donor1", k32_donor1, dim3(pm_grid(m, nf)), dim3(kBlock), A, m, T, u, v,
                 w);
    hipLaunchKernelGGL(k32_exit_flags, dim3(nlead), dim3(1024), 0, c->stream, A, m.ntiles);
    DCM_LAUNCH_CHECK();
    launch_timed(c, "k32_antidiff", k32_antidiff, dim3(pm_grid(m, nlead)), dim3(kBlock), A, m, T,
                 u, v, w);
    DCM_LAUNCH_CHECK();
    if (monotone) {
        launch_timed(c, "k32_beta", k32_beta, dim3(pm_grid(m, nlead)), dim3(kBlock), A, m, T);
        launch_timed(c, "k32_limited", k32_limited<true>, dim3(pm_grid(m, nf)), dim3(kBlock), A,
                     m, T);
    } else {
        launch_timed(c, "k32_limited", k32_limited<false>, dim3(pm_grid(m, nf)), dim3(kBlock), A,
                     m, T);
    }
    DCM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k32_exit_unshift, dim3(256, nf), dim3(kBlock), 0, c->stream, A, g.ntot);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

int mpdata32_batch(dcm_ctx *c, float dt, const float *u, const float *v, const float *w,
                   const Mp32Cat *cats, int ncat, int kord, int monotone, int neumann) {
    const Geo &g = c->geo;
    const bool fast = kord == 2 && c->mpdata_variant == 0 && g.ntot < (1LL << 31);
    if (!fast) {
        for (int q = 0; q < ncat; ++q)
            DCM_CHECK(mpdata_core<float>(c, dt, u, v, w, cats[q].psi, cats[q].nf, cats[q].lbc,
                                         cats[q].ubc, kord, monotone, neumann,
                                         cats[q].whole_min ? 0 : 1));
        return DCM_OK;
    }
    // batches of whole categories, at most kMp32Max fields each; the shift
    // of field f of a batch is dscal[16 + f]
    int q = 0;
    while (q < ncat) {
        if (cats[q].nf > kMp32Max) {
            DCM_CHECK(mpdata_core<float>(c, dt, u, v, w, cats[q].psi, cats[q].nf, cats[q].lbc,
                                         cats[q].ubc, kord, monotone, neumann,
                                         cats[q].whole_min ? 0 : 1));
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
