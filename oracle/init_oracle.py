"""TEST INFRASTRUCTURE ONLY: a restatement of the 1-D part of the model's
initialisation (allocate_and_set, initialisation.f90:117-888, and
setup_adiabatic_profile, :899-1311) with general-purpose numerical tools,
to check the library's dcm_init_compute_profiles.

Where the library uses the exact Exner integral and RK4 with Brent's
method, this integrates the reference's own ODEs (hydrostatic1a/1b/2b/2c/2e)
with scipy's adaptive Dormand-Prince method (rtol 1e-12) and finds the
temperatures with scipy.optimize.brentq, in the reference's order of steps:
the two agree to about 1e-10 where the library is exact and 1e-8 on the
moist adiabat, far inside the reference's own VODE/zeroin tolerance (1e-5).
The sounding interpolation is numpy's, with the reference's clamping
(linear extrapolation below the first level, constant above the last)."""
import math

import numpy as np
from scipy.integrate import solve_ivp
from scipy.optimize import brentq

# initialisation.f90:9-10
RA, RV, GRAV, CP, CW, TTR, LV = 287.0, 461.5, 9.81, 1005.0, 4190.0, 273.15, 2.5e6
EPS1 = RA / RV
KAPPA = RA / CP


def svp_liq(t):
    """svp_liq, initialisation.f90:1592-1600 (Buck)."""
    return 100.0 * 6.1121 * np.exp((18.678 - (t - TTR) / 234.5) * (t - TTR) / (257.14 + (t - TTR)))


def sounding(zr, vr, z):
    """Linear in z between levels, the first segment extended below the
    first level, constant above the last (:539-546)."""
    z = np.minimum(np.asarray(z, dtype=np.float64), zr[-1])
    out = np.interp(z, zr, vr)
    below = z < zr[0]
    slope = (vr[1] - vr[0]) / (zr[1] - zr[0])
    return np.where(below, vr[0] + slope * (z - zr[0]), out)


def _integrate(f, y0, x0, x1):
    if x1 == x0:
        return y0
    r = solve_ivp(f, (x0, x1), [y0], method="DOP853", rtol=1e-12, atol=1e-9)
    assert r.success
    return float(r.y[0, -1])


def hydrostatic_sounding(zr, thr, psurf, z):
    """hydrostatic1a from the surface (z <= 0 at w levels keeps psurf is
    the caller's business), integrated level to level of the sounding so
    that no step crosses a kink of theta(z)."""
    f = lambda zz, p: [-(GRAV * p[0]) / (RA * float(sounding(zr, thr, zz)) *
                                          (p[0] / 1e5) ** KAPPA)]
    lo, hi = min(0.0, z), max(0.0, z)
    cuts = [c for c in zr[1:] if lo < c < hi]
    pts = [0.0] + (cuts if z > 0 else cuts[::-1]) + [z]
    p = psurf
    for a, b in zip(pts[:-1], pts[1:]):
        p = _integrate(f, p, a, b)
    return p


def grid(ip, jp, kp, dx, dy, dz):
    """:429-447 on one rank."""
    f = np.arange(ip + 2.0)
    x = dx * f - (ip - 1) / 2.0 * dx
    f = np.arange(jp + 2.0)
    y = dy * f - (jp - 1) / 2.0 * dy
    f = np.arange(kp + 2.0)
    z = dz * (f - 1) + dz
    return x, y, z, x - 0.5 * dx, y - 0.5 * dy, z - 0.5 * dz


def reference_state_sounding(nm, zr, thr, z, zn):
    out = {k: np.zeros(len(z)) for k in ("pref", "prefn", "theta", "thetan", "rhoa", "rhoan",
                                          "tref", "trefn")}
    for f in range(len(z)):
        p = nm["psurf"] if z[f] <= 0 else hydrostatic_sounding(zr, thr, nm["psurf"], z[f])
        th = float(sounding(zr, thr, z[f]))
        out["pref"][f], out["theta"][f] = p, th
        out["tref"][f] = th * (p / 1e5) ** KAPPA
        out["rhoa"][f] = p / (RA * out["tref"][f])
        pn = hydrostatic_sounding(zr, thr, nm["psurf"], zn[f])
        thn = float(sounding(zr, thr, zn[f]))
        out["prefn"][f], out["thetan"][f] = pn, thn
        out["trefn"][f] = thn * (pn / 1e5) ** KAPPA
        out["rhoan"][f] = pn / (RA * out["trefn"][f])
    return out


def adiabatic_column(nm, z, cloud_water):
    """setup_adiabatic_profile on one set of levels, step by step."""
    n = len(z)
    p_, t_, th_, rho_ = (np.zeros(n) for _ in range(4))
    qv, qc = np.zeros(n), np.zeros(n)
    psurf, tsurf, tcb, tct = nm["psurf"], nm["tsurf"], nm["t_cbase"], nm["t_ctop"]
    zcb = (tsurf - tcb) * CP / GRAV
    dry = lambda zz, p: [-(GRAV * p[0]) / (RA * tsurf * (p[0] / 1e5) ** KAPPA)]
    for k in range(n):  # 1. the dry adiabat
        p = _integrate(dry, psurf, 0.0, z[k])
        t = tsurf * (p / 1e5) ** KAPPA
        if t < tcb:
            break
        p_[k], t_[k], th_[k], rho_[k] = p, t, tsurf, p / (RA * t)
    pcb = _integrate(dry, psurf, 0.0, zcb)  # 2.
    rv_sub = EPS1 * svp_liq(tcb) / (pcb - svp_liq(tcb))
    rv = rv_sub if nm["theta_flag"] else 0.0
    frac2 = nm["adiabatic_frac"] if nm["theta_flag"] else 0.0

    def theta_q(t, p):
        ws = EPS1 * svp_liq(t) / (p - svp_liq(t)) * frac2
        return t * (1e5 / p) ** (RA / (CP + CW * rv)) * math.exp(LV * ws / (t * (CP + CW * rv)))

    tq_sat = theta_q(tcb, pcb)
    temp = lambda p: brentq(lambda t: theta_q(t, p) - tq_sat, 20.0, 1.1 * tcb, xtol=1e-13)
    moist = lambda zz, p: [-(GRAV * p[0]) / (RA * temp(p[0]))]
    for k in range(n):  # 3. the moist adiabat
        if zcb >= z[k]:
            qv[k] = rv_sub
            continue
        p = _integrate(moist, pcb, zcb, z[k])
        t = temp(p)
        if t < tct:
            break
        p_[k], t_[k] = p, t
        th_[k] = t * (psurf / p) ** KAPPA
        rho_[k] = p / (RA * t)
        qv[k] = EPS1 * svp_liq(t) / (p - svp_liq(t)) * 0.95
        qc[k] = max((rv_sub - qv[k]) * nm["adiabatic_frac"], 0.0) if cloud_water else 0.0
    pctop_dry = pcb * math.exp(CP / RA * math.log(tct / tcb))  # 4.
    pct = brentq(lambda p: theta_q(tct, p) - tq_sat, pctop_dry / 2.0, pcb, xtol=1e-10)
    zct = _integrate(lambda p, zz: [-(RA * temp(p)) / (GRAV * p)], zcb, pcb, pct)  # 5.
    th_ct = (tct + nm["th_jump"]) * (1e5 / pct) ** KAPPA  # 6.
    above = lambda zz, p: [-(GRAV * p[0]) / (RA * (th_ct + nm["th_grad"] * (zz - zct)) *
                                              (p[0] / 1e5) ** KAPPA)]
    for k in range(n):
        if zct >= z[k]:
            continue
        p = _integrate(above, pct, zct, z[k])
        t = (th_ct + nm["th_grad"] * (z[k] - zct)) * (p / 1e5) ** KAPPA
        p_[k], t_[k] = p, t
        th_[k] = t * (1e5 / p) ** KAPPA
        rho_[k] = p / (RA * t)
        qv[k] = nm["rh_above"] * EPS1 * svp_liq(t) / (p - svp_liq(t))
    return dict(p=p_, t=t_, theta=th_, rho=rho_, qv=qv, qc=qc, zcb=zcb, zct=zct, pcb=pcb,
                pct=pct, theta_q=theta_q, tq_sat=tq_sat)
