// dcm_subgrid.hip -- the options of the reference's sources beyond the
// inviscid branch, on the device: the Smagorinsky-Lilly subgrid model of
// calculate_subgrid_3d (sgm/subgrid_3d.f90:38-357, with richardson :448-493
// and monin_obukhov :373-433) and the damping-layer, forcing and subsidence
// terms (dynamics.f90:802-873).
//
// Each kernel evaluates the reference's expressions in its operand order
// (no FMA contraction), so FP64 results are bit-identical to a serial CPU
// run of the same arithmetic.  Two library calls are replaced by exactly
// rounded forms: x**0.5 by sqrt and x**4 by two squarings (the reference's
// libm pow is within an ulp of both), and log(z(1)/z0), one scalar per grid,
// is computed on the host.
#include <cmath>

#include "dcm_device.h"

namespace dcm {

namespace {
// parameters of subgrid_3d.f90:13-16
constexpr double kPrn = 0.95, kSubb = 40.0, kSubc = 16.0, kSubg = 1.2, kSubh = 0.0,
                 kRic = 0.25, kSgSmall = 1e-10, kKvon = 0.4, kGrav = 9.81;
constexpr double kSuba = 1.0 / kPrn, kSubf = kSuba;

__device__ __forceinline__ double sq2(double x) { return x * x; }
__device__ __forceinline__ double pow4(double x) {
    const double t = x * x;
    return t * t;
}
}  // namespace

// 2 S_ij S_ij on the w points (subgrid_3d.f90:91-131; the sixth term of the
// 2 s12 s12 average repeats the second, as written there)
__device__ __forceinline__ double strain_rate(const double *__restrict__ u,
                                              const double *__restrict__ v,
                                              const double *__restrict__ w, const Geo &g,
                                              const Metrics &mt, int k, int j, int i) {
    auto U = [&](int kk, int jj, int ii) { return u[pidx(g, kk, jj, ii)]; };
    auto V = [&](int kk, int jj, int ii) { return v[pidx(g, kk, jj, ii)]; };
    auto W = [&](int kk, int jj, int ii) { return w[pidx(g, kk, jj, ii)]; };
    const double *dx = mt.dx, *dy = mt.dy, *dz = mt.dz, *dxn = mt.dxn, *dyn = mt.dyn,
                 *dzn = mt.dzn;
    double s = sq2((U(k + 1, j, i) - U(k + 1, j, i - 1)) / dx[i - 1]) +
               sq2((U(k, j, i) - U(k, j, i - 1)) / dx[i - 1]);
    s = (s + sq2((V(k + 1, j, i) - V(k + 1, j - 1, i)) / dy[j - 1])) +
        sq2((V(k, j, i) - V(k, j - 1, i)) / dy[j - 1]);
    s = (s + sq2((W(k, j, i) - W(k - 1, j, i)) / dz[k - 1])) +
        sq2((W(k + 1, j, i) - W(k, j, i)) / dz[k]);
    s = s + 0.5 * (sq2((U(k + 1, j, i) - U(k, j, i)) / dzn[k] +
                       (W(k, j, i + 1) - W(k, j, i)) / dxn[i]) +
                   sq2((U(k + 1, j, i - 1) - U(k, j, i - 1)) / dz[k] +
                       (W(k, j, i) - W(k, j, i - 1)) / dxn[i - 1]));
    s = s + 0.5 * (sq2((W(k, j, i) - W(k, j - 1, i)) / dyn[j - 1] +
                       (V(k + 1, j - 1, i) - V(k, j - 1, i)) / dzn[k]) +
                   sq2((W(k, j + 1, i) - W(k, j, i)) / dyn[j] +
                       (V(k + 1, j, i) - V(k, j, i)) / dzn[k]));
    const double p1 = (U(k, j, i - 1) - U(k, j - 1, i - 1)) / dyn[j - 1] +
                      (V(k, j - 1, i) - V(k, j - 1, i - 1)) / dxn[i - 1];
    const double p2 = (U(k, j + 1, i - 1) - U(k, j, i - 1)) / dyn[j] +
                      (V(k, j, i) - V(k, j, i - 1)) / dxn[i - 1];
    const double p3 = (U(k, j, i) - U(k, j - 1, i)) / dyn[j - 1] +
                      (V(k, j - 1, i + 1) - V(k, j - 1, i)) / dxn[i];
    const double p4 = (U(k, j + 1, i) - U(k, j, i)) / dyn[j] + (V(k, j, i + 1) - V(k, j, i)) / dxn[i];
    const double p5 = (U(k + 1, j, i - 1) - U(k + 1, j - 1, i - 1)) / dyn[j - 1] +
                      (V(k + 1, j - 1, i) - V(k + 1, j - 1, i - 1)) / dxn[i - 1];
    const double p6 = p2;
    const double p7 = (U(k + 1, j, i) - U(k + 1, j - 1, i)) / dyn[j - 1] +
                      (V(k + 1, j - 1, i + 1) - V(k + 1, j - 1, i)) / dxn[i];
    const double p8 = (U(k + 1, j + 1, i) - U(k + 1, j, i)) / dyn[j] +
                      (V(k + 1, j, i + 1) - V(k + 1, j, i)) / dxn[i];
    s = s + 0.125 * (((((((sq2(p1) + sq2(p2)) + sq2(p3)) + sq2(p4)) + sq2(p5)) + sq2(p6)) +
                      sq2(p7)) + sq2(p8));
    return s;
}

// strain rate, Richardson functions and the viscosities (subgrid_3d.f90:
// 91-159, richardson :465-491); zero in the two levels below the lid
// (:163-166).  1.e-15 in the Richardson number is a default-real literal.
__global__ __launch_bounds__(256) void k_sg_viscosity(
    const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
    const double *__restrict__ th, const double *__restrict__ lamsq,
    const double *__restrict__ thetan, double *__restrict__ strain, double *__restrict__ vism,
    double *__restrict__ vist, Geo g, Metrics mt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const double s = strain_rate(u, v, w, g, mt, k, j, i);
    const double rip = ((kGrav * (th[c + 1] / thetan[k + 1] - th[c] / thetan[k])) / mt.dzn[k]) /
                       (s + (double)1.e-15f);
    double fm, fh;
    if (rip < 0.0) {
        fm = sqrt(1.0 - kSubc * rip);
        fh = kSuba * sqrt(1.0 - kSubb * rip);
    } else if (rip >= 0.0 && rip < kRic) {
        const double x = pow4(1.0 - rip / kRic);
        fm = x * (1.0 - kSubh * rip);
        fh = (kSubf * x) * (1.0 - kSubg * rip);
    } else {
        fm = 0.0;
        fh = 0.0;
    }
    strain[c] = s;
    const bool lid = k >= g.kp - 1;
    vism[c] = lid ? 0.0 : (lamsq[k] * sqrt(s)) * fm;
    vist[c] = lid ? 0.0 : ((lamsq[k] * sqrt(s)) * fh) * mt.rhoa[k];
}

// surface-layer viscosities at k = 0 (monin_obukhov, subgrid_3d.f90:390-428)
__global__ void k_sg_surface(const double *__restrict__ u, const double *__restrict__ v,
                             const double *__restrict__ th, double *__restrict__ vism,
                             double *__restrict__ vist, Geo g, double z1, double z0, double z0th,
                             double theta0, double thetan1, double logz) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y + 1;
    if (j > g.jp) return;
    const double kv2 = kKvon * kKvon;
    const double vmag = 0.5 * sqrt(sq2(u[pidx(g, 1, j, i)] + u[pidx(g, 1, j, i - 1)]) +
                                   sq2(v[pidx(g, 1, j, i)] + v[pidx(g, 1, j - 1, i)])) +
                        kSgSmall;
    const double dth = (th[pidx(g, 1, j, i)] + thetan1) - theta0;
    const double rib = ((kGrav * dth) * sq2(z1 - z0)) / ((theta0 * sq2(vmag)) * (z1 - z0th));
    double gm, gh;
    if (rib <= 0.0) {
        const double q = sqrt(fabs((rib * z1) / z0));
        gm = 1.0 - (9.4 * rib) / (1.0 + ((70.0 * kv2) * q) / sq2(logz));
        gh = 1.0 - (9.4 * rib) / (1.0 + ((50.0 * kv2) * q) / sq2(logz));
    } else {
        gm = 1.0 / sq2(1.0 + 4.7 * rib);
        gh = 1.0 / sq2(1.0 + 4.7 * rib);
    }
    const double ustar = ((kKvon * vmag) / logz) * sqrt(gm);
    const double thstar = (((kv2 * vmag) * dth) / ((ustar * kPrn) * sq2(logz))) * gh;
    vism[pidx(g, 0, j, i)] = (sq2(ustar) * (z1 - z0)) / vmag;
    vist[pidx(g, 0, j, i)] = (ustar * thstar) / (dth / (z1 - z0th) + kSgSmall);
}

// the stress tensor (subgrid_3d.f90:189-223) from the previous-level winds
__global__ __launch_bounds__(256) void k_sg_stress(
    const double *__restrict__ zu, const double *__restrict__ zv, const double *__restrict__ zw,
    const double *__restrict__ vism, double *__restrict__ t11, double *__restrict__ t22,
    double *__restrict__ t33, double *__restrict__ t12, double *__restrict__ t13,
    double *__restrict__ t23, Geo g, Metrics mt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const long long si = g.si, sj = g.sj;
    const double rn = mt.rhoan[k], ra = mt.rhoa[k];
    const double r1 = rn * 1.0;
    const double vv = vism[c] + vism[c - 1];
    t11[c] = ((r1 * vv) * (zu[c] - zu[c - si])) / mt.dx[i - 1];
    t22[c] = ((r1 * vv) * (zv[c] - zv[c - sj])) / mt.dy[j - 1];
    t33[c] = ((r1 * vv) * (zw[c] - zw[c - 1])) / mt.dz[k - 1];
    const double s8 = ((((((vism[c] + vism[c + sj]) + vism[c - 1]) + vism[c - 1 + sj]) +
                         vism[c + si]) + vism[c + sj + si]) + vism[c - 1 + si]) +
                      vism[c - 1 + sj + si];
    t12[c] = ((rn * 0.125) * s8) *
             ((zu[c + sj] - zu[c]) / mt.dyn[j] + (zv[c + si] - zv[c]) / mt.dxn[i]);
    t13[c] = ((ra * 0.5) * (vism[c] + vism[c + si])) *
             ((zu[c + 1] - zu[c]) / mt.dzn[k] + (zw[c + si] - zw[c]) / mt.dxn[i]);
    t23[c] = ((ra * 0.5) * (vism[c] + vism[c + sj])) *
             ((zw[c + sj] - zw[c]) / mt.dyn[j] + (zv[c + 1] - zv[c]) / mt.dzn[k]);
}

// viscous momentum tendencies (subgrid_3d.f90:241-278): sw on 1..kp-1, su
// and sv on 2..kp-1
__global__ __launch_bounds__(256) void k_sg_momentum(
    const double *__restrict__ t11, const double *__restrict__ t22,
    const double *__restrict__ t33, const double *__restrict__ t12,
    const double *__restrict__ t13, const double *__restrict__ t23, double *__restrict__ su,
    double *__restrict__ sv, double *__restrict__ sw, Geo g, Metrics mt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp - 1 || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const long long si = g.si, sj = g.sj;
    sw[c] = sw[c] + (((t33[c + 1] - t33[c]) / mt.dzn[k] + (t13[c] - t13[c - si]) / mt.dx[i - 1]) +
                     (t23[c] - t23[c - sj]) / mt.dy[j - 1]);
    if (k >= 2) {
        su[c] = su[c] + (((t11[c + si] - t11[c]) / mt.dxn[i] + (t12[c] - t12[c - sj]) / mt.dy[j - 1]) +
                         (t13[c] - t13[c - 1]) / mt.dz[k - 1]);
        sv[c] = sv[c] + (((t22[c + sj] - t22[c]) / mt.dyn[j] + (t12[c] - t12[c - si]) / mt.dx[i - 1]) +
                         (t23[c] - t23[c - 1]) / mt.dz[k - 1]);
    }
}

// viscous flux divergence of the scalars (sth :288-312, sq :322-349): th
// and then the nq fields of q, in one pass that reads the neighbourhood of
// vt (vist, already multiplied by rhoa) once per cell; each field's
// arithmetic is the reference's.  With viss, viss(:,:,:,n) = vist is written
// on the interior in the same pass (:185; the halo shell by k_viss_shell).
__global__ __launch_bounds__(256) void k_sg_diffuse(const double *__restrict__ vt,
                                                    const double *__restrict__ th,
                                                    double *__restrict__ sth,
                                                    const double *__restrict__ q,
                                                    double *__restrict__ sq, int nq,
                                                    double *__restrict__ viss, Geo g,
                                                    Metrics mt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const long long si = g.si, sj = g.sj;
    const double v0 = vt[c], vd = vt[c - 1];
    const double ax = ((v0 + vt[c - si]) + vd) + vt[c - 1 - si];
    const double bx = ((vt[c + si] + v0) + vt[c - 1 + si]) + vd;
    const double ay = ((v0 + vt[c - sj]) + vd) + vt[c - 1 - sj];
    const double by = ((vt[c + sj] + v0) + vt[c - 1 + sj]) + vd;
    const double dxm = mt.dx[i - 1], dxnm = mt.dxn[i - 1], dxn = mt.dxn[i];
    const double dym = mt.dy[j - 1], dynm = mt.dyn[j - 1], dyn = mt.dyn[j];
    const double dzm = mt.dz[k - 1], dzk = mt.dz[k], rn = mt.rhoan[k];
    for (int f = 0; f <= nq; ++f) {
        const double *T = f == 0 ? th : q + (long long)(f - 1) * g.ntot;
        double *sT = f == 0 ? sth : sq + (long long)(f - 1) * g.ntot;
        const double tc = T[c];
        const double X = (0.25 * ((ax * (T[c - si] - tc)) / dxnm - (bx * (tc - T[c + si])) / dxn)) /
                         dxm;
        const double Y = (0.25 * ((ay * (T[c - sj] - tc)) / dynm - (by * (tc - T[c + sj])) / dyn)) /
                         dym;
        const double Z = ((vd * (T[c - 1] - tc)) / dzm - (v0 * (tc - T[c + 1])) / dzk) / dzm;
        sT[c] = sT[c] + ((X + Y) + Z) / rn;
        if (viss && f > 0) viss[(long long)(f - 1) * g.ntot + c] = v0;
    }
}

// viss(:,:,:,n) = vist on the cells k_sg_diffuse does not write: the halo
// shell (k or j or i outside the interior)
__global__ void k_viss_shell(const double *__restrict__ vist, double *__restrict__ viss, int nq,
                             Geo g) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ntot;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c / g.si);
        const long long r = c - (long long)i * g.si;
        const int j = (int)(r / g.sj);
        const int k = (int)(r - (long long)j * g.sj);
        if (i >= 1 && i <= g.ip && j >= 1 && j <= g.jp && k >= 1 && k <= g.kp) continue;
        const double v = vist[c];
        for (int n = 0; n < nq; ++n) viss[(long long)n * g.ntot + c] = v;
    }
}

// damping layer (dynamics.f90:802-839), forcing (:841-860) and subsidence
// (:863-873), in that order, on the interior
__global__ __launch_bounds__(256) void k_relax(
    const double *__restrict__ zu, const double *__restrict__ zv, const double *__restrict__ zw,
    const double *__restrict__ th, double *__restrict__ su, double *__restrict__ sv,
    double *__restrict__ sw, double *__restrict__ sth, PhysicsDev ph, Geo g, Metrics mt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.kp || j > g.jp) return;
    const long long c = pidx(g, k, j, i);
    const double rn = mt.rhoan[k], ra = mt.rhoa[k];
    double a = su[c], b = sv[c], d = sw[c];
    if (ph.damping_layer) {
        a = a + (rn * ph.dampfacn[k]) * (zu[c] - ph.ubar[k]);
        b = b + (rn * ph.dampfacn[k]) * (zv[c] - ph.vbar[k]);
        d = d + (ra * ph.dampfac[k]) * (zw[c] - 0.0);
        sth[c] = sth[c] + ph.dampfacn[k] * (th[c] - 0.0);
    }
    if (ph.forcing) {
        a = a - (rn * (zu[c] - ph.u_force[k])) / ph.forcing_tau;
        b = b - (rn * (zv[c] - ph.v_force[k])) / ph.forcing_tau;
    }
    if (ph.divergence) d = d - rn * (ph.wbar[k] - ph.w_subs[k]);
    su[c] = a;
    sv[c] = b;
    sw[c] = d;
}

// calculate_subgrid_3d (subgrid_3d.f90:38-357)
int subgrid_dev(dcm_ctx *c, const PhysicsDev &ph, const double *u, const double *v,
                const double *w, const double *th, const SubgridFields &sf, double *su,
                double *sv, double *sw, double *sth) {
    const Geo &g = c->geo;
    const Interior L = interior_launch(g);
    const Metrics mt = c->met;
    int tok = prof_begin(c, "subgrid");
    hipLaunchKernelGGL(k_sg_viscosity, L.grid, L.block, 0, c->stream, u, v, w, th, ph.lamsq,
                       ph.thetan, sf.strain, sf.vism, sf.vist, g, mt);
    DCM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sg_surface, dim3((g.jp + 63) / 64, g.ip), dim3(64), 0, c->stream, u, v,
                       th, sf.vism, sf.vist, g, ph.z1, ph.z0, ph.z0th, ph.theta0, ph.thetan1,
                       ph.log_z1_z0);
    DCM_LAUNCH_CHECK();
    DCM_CHECK(fill<double>(c, sf.vism, FILL_FULL_WO, 0.0, 0.0));  // :176-179
    DCM_CHECK(fill<double>(c, sf.vist, FILL_FULL_WO, 0.0, 0.0));
    const bool moist = ph.moisture && sf.q && sf.sq && ph.nq > 0;
    if (moist && sf.viss) {  // :185, the shell here, the interior with the diffusion
        hipLaunchKernelGGL(k_viss_shell, dim3(2048), dim3(kBlock), 0, c->stream, sf.vist,
                           sf.viss, ph.nq, g);
        DCM_LAUNCH_CHECK();
    }
    double *t[6];
    for (int n = 0; n < 6; ++n) DCM_CHECK(work_field64(c, kSlotStress + n, &t[n]));
    hipLaunchKernelGGL(k_sg_stress, L.grid, L.block, 0, c->stream, sf.zu, sf.zv, sf.zw, sf.vism,
                       t[0], t[1], t[2], t[3], t[4], t[5], g, mt);
    DCM_LAUNCH_CHECK();
    for (int n = 0; n < 6; ++n) DCM_CHECK(fill<double>(c, t[n], FILL_ALONG_DIM, 0.0, 0.0));
    hipLaunchKernelGGL(k_sg_momentum, L.grid, L.block, 0, c->stream, t[0], t[1], t[2], t[3],
                       t[4], t[5], su, sv, sw, g, mt);
    DCM_LAUNCH_CHECK();
    DCM_CHECK(fill<double>(c, su, FILL_FULL, 0.0, 0.0));  // :279-284
    DCM_CHECK(fill<double>(c, sv, FILL_FULL, 0.0, 0.0));
    DCM_CHECK(fill<double>(c, sw, FILL_FULL, 0.0, 0.0));
    hipLaunchKernelGGL(k_sg_diffuse, L.grid, L.block, 0, c->stream, sf.vist, th, sth,
                       moist ? sf.q : nullptr, moist ? sf.sq : nullptr, moist ? ph.nq : 0,
                       moist ? sf.viss : nullptr, g, mt);
    DCM_LAUNCH_CHECK();
    DCM_CHECK(fill<double>(c, sth, FILL_FULL, 0.0, 0.0));  // :314-315
    if (moist)
        for (int n = 0; n < ph.nq; ++n)
            DCM_CHECK(fill<double>(c, sf.sq + (size_t)n * g.ntot, FILL_FULL, 0.0, 0.0));  // :351-352
    prof_end(c, tok);
    return DCM_OK;
}

int relax_dev(dcm_ctx *c, const PhysicsDev &ph, const double *zu, const double *zv,
              const double *zw, const double *th, double *su, double *sv, double *sw,
              double *sth) {
    if (!ph.damping_layer && !ph.forcing && !ph.divergence) return DCM_OK;
    const Interior L = interior_launch(c->geo);
    hipLaunchKernelGGL(k_relax, L.grid, L.block, 0, c->stream, zu, zv, zw, th, su, sv, sw, sth,
                       ph, c->geo, c->met);
    DCM_LAUNCH_CHECK();
    return DCM_OK;
}

}  // namespace dcm
