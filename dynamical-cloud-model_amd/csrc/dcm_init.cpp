// dcm_init.cpp -- the 1-D part of the model's initialisation
// (allocate_and_set, initialisation.f90:117-888, and
// setup_adiabatic_profile, :899-1311), host code: grid, mixing length,
// damping layer, hydrostatic reference state, sounding winds, forcing and
// subsidence profiles, and the columns of the moisture fields.  The 3-D
// fields are generated from these profiles on the GPU
// (dcm_model_init_fields).
//
// The reference integrates the hydrostatic equation with VODE (relative
// tolerance 1e-5) and finds temperatures with zeroin (tolerance 1e-5).
// Here, wherever potential temperature is a known function of height (the
// sounding, the dry adiabat, the layer above cloud top) the integral is
// exact: with the Exner function pi = (p/1e5)^(ra/cp) the hydrostatic
// equation dp/dz = -g p / (ra T) becomes dpi/dz = -g / (cp theta(z)), and
// theta is piecewise linear in z.  Only the moist adiabat (temperature
// implicit in p) is integrated numerically: classical fourth-order
// Runge-Kutta with 0.5 m steps and a Brent root find of the temperature to
// 1e-12 K at every stage.  Both are far below the reference's own
// tolerances, so the results agree with it to those tolerances.
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/dcm_hip.h"

namespace dcm {
void set_error(const std::string &msg);
}

namespace {

// initialisation.f90:9-10
constexpr double kRa = 287.0, kRv = 461.5, kEps1 = kRa / kRv, kGrav = 9.81, kCp = 1005.0,
                 kCw = 4190.0, kTtr = 273.15, kLv = 2.5e6;
constexpr double kKappa = kRa / kCp;

// svp_liq (:1592-1600), Buck's fit over liquid water
double svp_liq(double t) {
    return 100.0 * 6.1121 * std::exp((18.678 - (t - kTtr) / 234.5) * (t - kTtr) / (257.14 + (t - kTtr)));
}

// find_pos (numerics find_pos.f90): bisection for the interval holding x
int find_pos(const double *xa, int n, double x) {
    const bool ascnd = xa[n - 1] >= xa[0];
    int ilow = 0, iupp = n + 1;
    while (iupp - ilow > 1) {
        const int imid = (iupp + ilow) / 2;
        if (ascnd == (x >= xa[imid - 1]))
            ilow = imid;
        else
            iupp = imid;
    }
    if (x == xa[0]) return 1;
    if (x == xa[n - 1]) return n - 1;
    return ilow;
}

// poly_int on two points (Neville's scheme as numerics poly_int.f90 runs it)
double poly_int2(double x1, double x2, double y1, double y2, double x) {
    const double d1 = x1 - x, d2 = x2 - x;
    const int ns = std::fabs(d2) < std::fabs(d1) ? 1 : 0;
    double y = ns == 0 ? y1 : y2;
    const double f = (y2 - y1) / (d1 - d2);
    y += ns == 0 ? d1 * f : d2 * f;
    return y;
}

// the sounding interpolation of allocate_and_set (:539-546): locate the
// interval, clamp it to 1..n-1, interpolate at min(z, z_read(n))
double sounding(const double *zr, const double *vr, int n, double z) {
    int iloc = find_pos(zr, n, z);
    iloc = std::min(n - 1, iloc);
    iloc = std::max(1, iloc);
    return poly_int2(zr[iloc - 1], zr[iloc], vr[iloc - 1], vr[iloc], std::min(z, zr[n - 1]));
}

double exner(double p) { return std::pow(p / 1e5, kKappa); }
double pressure(double pi) { return 1e5 * std::pow(pi, 1.0 / kKappa); }

// integral of dz / theta(z) for theta linear on [z1, z2] (theta1 at z1,
// theta2 at z2): (z2 - z1) log(theta2 / theta1) / (theta2 - theta1)
double inv_theta_integral(double z1, double z2, double th1, double th2) {
    const double u = (th2 - th1) / th1;
    if (std::fabs(u) < 1e-12) return (z2 - z1) / th1 * (1.0 - 0.5 * u);
    return (z2 - z1) * std::log1p(u) / (th2 - th1);
}

// the Exner integral int_0^z dz / theta_sounding(z), theta_sounding the
// reference's interpolation (linear between levels, extrapolated below the
// first, constant above the last)
double sounding_inv_theta(const double *zr, const double *thr, int n, double z) {
    std::vector<double> brk;  // breakpoints between 0 and z
    const double lo = std::min(0.0, z), hi = std::max(0.0, z);
    brk.push_back(lo);
    for (int l = 1; l < n - 1; ++l)  // interior levels switch segments
        if (zr[l] > lo && zr[l] < hi) brk.push_back(zr[l]);
    if (zr[n - 1] > lo && zr[n - 1] < hi) brk.push_back(zr[n - 1]);
    brk.push_back(hi);
    double s = 0.0;
    for (size_t b = 0; b + 1 < brk.size(); ++b) {
        const double a = brk[b], c = brk[b + 1];
        if (c <= a) continue;
        s += inv_theta_integral(a, c, sounding(zr, thr, n, a), sounding(zr, thr, n, c));
    }
    return z >= 0.0 ? s : -s;
}

// Brent's root finder (Brent 1973, "Algorithms for Minimization without
// Derivatives", ch. 4), on [a, b] with f(a) f(b) <= 0
double brent(double a, double b, const std::function<double(double)> &f, double tol,
             bool *ok) {
    double fa = f(a), fb = f(b);
    *ok = true;
    if (fa == 0.0) return a;
    if (fb == 0.0) return b;
    if ((fa > 0) == (fb > 0)) {
        *ok = false;
        return b;
    }
    double c = a, fc = fa, d = b - a, e = d;
    for (int it = 0; it < 200; ++it) {
        if ((fb > 0) == (fc > 0)) {
            c = a;
            fc = fa;
            d = e = b - a;
        }
        if (std::fabs(fc) < std::fabs(fb)) {
            a = b;
            b = c;
            c = a;
            fa = fb;
            fb = fc;
            fc = fa;
        }
        const double t = 2.0 * 2.2e-16 * std::fabs(b) + 0.5 * tol;
        const double m = 0.5 * (c - b);
        if (std::fabs(m) <= t || fb == 0.0) return b;
        if (std::fabs(e) >= t && std::fabs(fa) > std::fabs(fb)) {
            double s = fb / fa, p, q;
            if (a == c) {
                p = 2.0 * m * s;
                q = 1.0 - s;
            } else {
                const double qq = fa / fc, r = fb / fc;
                p = s * (2.0 * m * qq * (qq - r) - (b - a) * (r - 1.0));
                q = (qq - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0)
                q = -q;
            else
                p = -p;
            if (2.0 * p < std::min(3.0 * m * q - std::fabs(t * q), std::fabs(e * q))) {
                e = d;
                d = p / q;
            } else {
                d = m;
                e = d;
            }
        } else {
            d = m;
            e = d;
        }
        a = b;
        fa = fb;
        b += std::fabs(d) > t ? d : (m > 0 ? t : -t);
        fb = f(b);
    }
    return b;
}

struct Adiabat {
    // the state of setup_adiabatic_profile's module globals
    double rv = 0.0, frac2 = 0.0, theta_q_sat = 0.0;
    // calc_theta_q (:1567-1579): moist potential temperature at (t, p)
    double theta_q(double t, double p) const {
        const double ws = kEps1 * svp_liq(t) / (p - svp_liq(t)) * frac2;
        return t * std::pow(1e5 / p, kRa / (kCp + kCw * rv)) *
               std::exp(kLv * ws / (t * (kCp + kCw * rv)));
    }
    // the temperature on the moist adiabat at pressure p (the zeroin
    // brackets of hydrostatic2b/2c: 20 K .. 1.1 t_cbase)
    double temperature(double p, double t_cbase, bool *ok) const {
        return brent(20.0, 1.1 * t_cbase, [&](double t) { return theta_q(t, p) - theta_q_sat; },
                     1e-12, ok);
    }
};

template <typename F> double rk4(double y, double x0, double x1, double hmax, F &&f, bool *ok) {
    const double span = x1 - x0;
    if (span == 0.0) return y;
    const int n = std::max(1, (int)std::ceil(std::fabs(span) / hmax));
    const double h = span / n;
    double x = x0;
    for (int s = 0; s < n && *ok; ++s) {
        const double k1 = f(x, y, ok);
        const double k2 = f(x + 0.5 * h, y + 0.5 * h * k1, ok);
        const double k3 = f(x + 0.5 * h, y + 0.5 * h * k2, ok);
        const double k4 = f(x + h, y + h * k3, ok);
        y += h / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
        x = x0 + (s + 1) * h;
    }
    return y;
}

struct Column {
    // one set of levels (w levels z or scalar levels zn) of the reference state
    const double *z;
    double *p, *t, *theta, *rho, *qv, *qc;
};

// setup_adiabatic_profile (:899-1311) on one set of levels; `cloud_water`
// false for the scalar levels, where the reference sets qc1d to zero
bool adiabatic_column(const dcm_init_config &c, int nk, Column col, bool cloud_water,
                      std::string *why) {
    const double psurf = c.psurf, tsurf = c.tsurf, t_cbase = c.t_cbase, t_ctop = c.t_ctop;
    const double pis = exner(psurf);
    const double zcb = (tsurf - t_cbase) * kCp / kGrav;
    for (int k = 0; k < nk; ++k) col.qv[k] = col.qc[k] = 0.0;
    // 1. the dry adiabat (theta = tsurf) up to cloud base
    for (int k = 0; k < nk; ++k) {
        const double p = pressure(pis - kGrav * col.z[k] / (kCp * tsurf));
        const double t = tsurf * std::pow(p / 1e5, kKappa);
        if (t < t_cbase) break;
        col.p[k] = p;
        col.t[k] = t;
        col.theta[k] = tsurf;
        col.rho[k] = p / (kRa * t);
    }
    // 2. cloud-base pressure and the saturation mixing ratio there
    const double pcb = pressure(pis - kGrav * zcb / (kCp * tsurf));
    const double rv_sub = kEps1 * svp_liq(t_cbase) / (pcb - svp_liq(t_cbase));
    // 3. the moist adiabat from cloud base
    Adiabat A;
    A.rv = c.theta_flag ? rv_sub : 0.0;
    A.frac2 = c.theta_flag ? c.adiabatic_frac : 0.0;
    A.theta_q_sat = A.theta_q(t_cbase, pcb);
    bool ok = true;
    auto dpdz = [&](double, double p, bool *okp) {
        bool o;
        const double t = A.temperature(p, t_cbase, &o);
        if (!o) *okp = false;
        return -(kGrav * p) / (kRa * t);
    };
    double zprev = zcb, pprev = pcb;
    for (int k = 0; k < nk; ++k) {
        if (zcb >= col.z[k]) {
            col.qv[k] = rv_sub;
            continue;
        }
        const double p = rk4(pprev, zprev, col.z[k], 0.5, dpdz, &ok);
        zprev = col.z[k];
        pprev = p;
        bool o;
        const double t = A.temperature(p, t_cbase, &o);
        if (!ok || !o) {
            if (why) *why = "adiabatic profile: no moist-adiabat temperature";
            return false;
        }
        if (t < t_ctop) break;
        col.p[k] = p;
        col.t[k] = t;
        col.theta[k] = t * std::pow(psurf / p, kKappa);
        col.rho[k] = p / (kRa * t);
        col.qv[k] = kEps1 * svp_liq(t) / (p - svp_liq(t)) * 0.95;
        col.qc[k] = cloud_water ? std::max((rv_sub - col.qv[k]) * c.adiabatic_frac, 0.0) : 0.0;
    }
    // 4. cloud-top pressure: the moist adiabat reaches t_ctop (calc_theta_q2)
    const double pctop_dry = pcb * std::exp(kCp / kRa * std::log(t_ctop / t_cbase));
    bool o;
    const double pct = brent(pctop_dry / 2.0, pcb,
                             [&](double p) { return A.theta_q(t_ctop, p) - A.theta_q_sat; },
                             1e-9, &o);
    if (!o) {
        if (why) *why = "adiabatic profile: cloud-top pressure not bracketed";
        return false;
    }
    // 5. cloud-top height: dz/dp along the moist adiabat (hydrostatic2c)
    auto dzdp = [&](double p, double, bool *okp) {
        bool oo;
        const double t = A.temperature(p, t_cbase, &oo);
        if (!oo) *okp = false;
        return -(kRa * t) / (kGrav * p);
    };
    const double zct = rk4(zcb, pcb, pct, 5.0, dzdp, &ok);
    if (!ok) {
        if (why) *why = "adiabatic profile: no temperature on the way to cloud top";
        return false;
    }
    // 6. above cloud top: theta = theta(cloud top) + th_jump + th_grad (z - zct)
    const double th_ct = (t_ctop + c.th_jump) * std::pow(1e5 / pct, kKappa);
    const double pict = exner(pct);
    for (int k = 0; k < nk; ++k) {
        if (zct >= col.z[k]) continue;
        const double th = th_ct + c.th_grad * (col.z[k] - zct);
        const double p = pressure(pict - kGrav / kCp * inv_theta_integral(zct, col.z[k], th_ct, th));
        const double t = th * std::pow(p / 1e5, kKappa);
        col.p[k] = p;
        col.t[k] = t;
        col.theta[k] = t * std::pow(1e5 / p, kKappa);
        col.rho[k] = p / (kRa * t);
        col.qv[k] = c.rh_above * kEps1 * svp_liq(t) / (p - svp_liq(t));
    }
    return true;
}

}  // namespace

extern "C" {

int dcm_init_compute_profiles(const dcm_init_config *c, dcm_init_profiles *o) {
    if (!c || !o) return DCM_BAD_ARG;
    const int ip = c->ip, jp = c->jp, kp = c->kp, nk = kp + 2;
    if (ip < 1 || jp < 1 || kp < 2 || !(c->dx > 0) || !(c->dy > 0) || !(c->dz > 0) ||
        c->n_levels < 2 || !c->z_read || !c->theta_read || !(c->dt > 0) ||
        (c->moisture && (c->nq < 1 || (!c->adiabatic_prof && !c->q_read))) ||
        !o->x || !o->y || !o->z || !o->xn || !o->yn || !o->zn || !o->dx || !o->dy || !o->dz ||
        !o->dxn || !o->dyn || !o->dzn || !o->theta || !o->thetan || !o->rhoa || !o->rhoan ||
        !o->pref || !o->prefn || !o->tref || !o->trefn || !o->lamsq || !o->lamsqn ||
        !o->dampfac || !o->dampfacn || !o->u_force || !o->v_force || !o->w_subs || !o->u ||
        !o->v || (c->moisture && !o->q)) {
        dcm::set_error("dcm_init_compute_profiles: bad configuration or missing output array");
        return DCM_BAD_ARG;
    }
    const int nl = c->n_levels;
    const double *zr = c->z_read, *thr = c->theta_read;
    o->ntim = (int)std::ceil(c->runtime / c->dt);  // :228
    // grid (:429-447), one rank: ipstart = jpstart = kpstart = 1
    for (int f = 0; f < ip + 2; ++f) {
        o->dx[f] = o->dxn[f] = c->dx;
        o->x[f] = c->dx * f - (double)(ip - 1) / 2.0 * c->dx;
        o->xn[f] = o->x[f] - 0.5 * c->dx;
    }
    for (int f = 0; f < jp + 2; ++f) {
        o->dy[f] = o->dyn[f] = c->dy;
        o->y[f] = c->dy * f - (double)(jp - 1) / 2.0 * c->dy;
        o->yn[f] = o->y[f] - 0.5 * c->dy;
    }
    for (int f = 0; f < nk; ++f) {
        o->dz[f] = o->dzn[f] = c->dz;
        o->z[f] = c->dz * (f - 1) + 1.0 * c->dz;
        o->zn[f] = o->z[f] - 0.5 * c->dz;
    }
    // mixing length (:449-452)
    for (int f = 0; f < nk; ++f) {
        const double a = c->cvis * ((c->dx + c->dy) + o->dz[f]) / 3.0;
        const double b = 0.4 * (o->z[f] + c->z0);
        o->lamsq[f] = 1.0 / (1.0 / (a * a) + 1.0 / (b * b));
        const double an = c->cvis * ((c->dx + c->dy) + o->dzn[f]) / 3.0;
        const double bn = 0.4 * (o->zn[f] + c->z0);
        o->lamsqn[f] = 1.0 / (1.0 / (an * an) + 1.0 / (bn * bn));
    }
    // damping layer above zn(kp) - damping_thickness (:466-491)
    const double ztop = o->zn[kp];
    for (int f = 0; f < nk; ++f) {
        o->dampfacn[f] = o->dampfac[f] = 0.0;
        if (c->damping_layer) {
            if (o->zn[f] > ztop - c->damping_thickness) o->dampfacn[f] = -1.0 / c->damping_tau;
            if (o->z[f] > ztop - c->damping_thickness) o->dampfac[f] = -1.0 / c->damping_tau;
        }
    }
    // reference state (:500-640 and setup_adiabatic_profile)
    std::vector<double> qv(nk, 0.0), qc(nk, 0.0), qvn(nk, 0.0), qcn(nk, 0.0);
    if (!c->adiabatic_prof) {
        const double pis = exner(c->psurf);
        for (int f = 0; f < nk; ++f) {
            // hydrostatic1a from the surface; z <= 0 keeps psurf (:511-516)
            const double zf = o->z[f];
            const double p = zf <= 0.0
                                 ? c->psurf
                                 : pressure(pis - kGrav / kCp * sounding_inv_theta(zr, thr, nl, zf));
            o->theta[f] = sounding(zr, thr, nl, zf);
            o->rhoa[f] = p / (kRa * o->theta[f] * std::pow(p / 100000.0, kKappa));
            o->pref[f] = p;
            o->tref[f] = o->theta[f] * std::pow(p / 100000.0, kKappa);
            const double zg = o->zn[f];
            const double pn = pressure(pis - kGrav / kCp * sounding_inv_theta(zr, thr, nl, zg));
            o->thetan[f] = sounding(zr, thr, nl, zg);
            o->rhoan[f] = pn / (kRa * o->thetan[f] * std::pow(pn / 100000.0, kKappa));
            o->prefn[f] = pn;
            o->trefn[f] = o->thetan[f] * std::pow(pn / 100000.0, kKappa);
        }
    } else {
        std::string why;
        Column w{o->z, o->pref, o->tref, o->theta, o->rhoa, qv.data(), qc.data()};
        Column s{o->zn, o->prefn, o->trefn, o->thetan, o->rhoan, qvn.data(), qcn.data()};
        if (!adiabatic_column(*c, nk, w, true, &why) || !adiabatic_column(*c, nk, s, false, &why)) {
            dcm::set_error("dcm_init_compute_profiles: " + why);
            return DCM_BAD_ARG;
        }
    }
    o->thbase = o->thetan[0];  // find_base_top (mpi_module.f90:310-333)
    o->thtop = o->thetan[kp + 1];
    // winds on the scalar levels, zero in the two lowest rows and row kp
    // (:753-800), or the param_wind erf profile (:737-751)
    for (int f = 0; f < nk; ++f) {
        double u = 0.0, v = 0.0;
        if (c->param_wind) {
            v = (0.5 * std::erf((o->zn[f] - c->param_z) / c->param_sigz) + 0.5) * c->param_vmax;
        } else {
            u = c->u_read ? sounding(zr, c->u_read, nl, o->zn[f]) : 0.0;
            v = c->v_read ? sounding(zr, c->v_read, nl, o->zn[f]) : 0.0;
        }
        if (f == kp || f <= 1) u = v = 0.0;
        o->u[f] = u;
        o->v[f] = v;
        // forcing toward column (1,1) (:838-843), subsidence (:846-853)
        o->u_force[f] = c->forcing ? u : 0.0;
        o->v_force[f] = c->forcing ? v : 0.0;
        o->w_subs[f] = !c->divergence ? 0.0
                       : (o->z[f] <= c->divergence_hgt ? -c->divergence_val * o->z[f]
                                                       : -c->divergence_val * c->divergence_hgt);
    }
    // columns of the moisture fields (:642-690 and :693-701)
    if (c->moisture) {
        const int nq = c->nq;
        double *q = o->q;
        // q_read is the Fortran array q_read(nq, n_levels): row 1 is strided
        std::vector<double> q1(nl, 0.0);
        if (c->q_read)
            for (int l = 0; l < nl; ++l) q1[l] = c->q_read[(size_t)l * nq];
        for (int t = 0; t < nq * nk; ++t) q[t] = 0.0;
        auto Q = [&](int field, int f) -> double * {  // field 1-based, 0: absent
            return field >= 1 && field <= nq ? &q[(size_t)(field - 1) * nk + f] : nullptr;
        };
        auto put = [&](int field, int f, double v) {
            if (double *d = Q(field, f)) *d = v;
        };
        for (int f = 0; f < nk; ++f) {
            if (!c->adiabatic_prof) {
                // q_read(1,:) on the scalar levels, split at saturation
                const double var = sounding(zr, q1.data(), nl, o->zn[f]);  // q_read(1, :)
                put(1, f, var);
                const double t = o->thetan[f] * std::pow(o->prefn[f] / c->psurf, kKappa);
                const double qsat = kEps1 * svp_liq(t) / (o->prefn[f] - svp_liq(t));
                if (var >= qsat) {
                    put(c->iqc, f, var - qsat);
                    put(c->iqv, f, qsat);
                    put(c->inc, f, c->drop_num);
                } else {
                    put(c->inc, f, 0.0);
                }
            } else {
                // the scalar-level pass of setup_adiabatic_profile decides
                put(c->inc, f, 0.0);
                put(c->iqc, f, qcn[f]);
                put(c->iqv, f, qvn[f]);
                if (qcn[f] > 0.0) put(c->inc, f, c->drop_num);
            }
        }
    }
    return DCM_OK;
}

}  // extern "C"
