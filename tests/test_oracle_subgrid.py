"""Pins of the oracle's subgrid model (calculate_subgrid_3d with richardson
and monin_obukhov, sgm/subgrid_3d.f90) and of the other options of sources
(damping layer, forcing, subsidence; dynamics.f90:802-917) against checks that
do not restate its code: closed forms of 2 S_ij S_ij for shear flows,
Smagorinsky's neutral viscosity nu = lambda^2 |S| with the turbulent Prandtl
number, continuity of the stability functions and no mixing beyond the
critical Richardson number, the neutral log law of the surface layer, the
discrete-Laplacian eigenfunction of the scalar diffusion, zero tendencies
where the flow meets the relaxation targets, and equality with the inviscid
sources when every option is off."""
import math

import numpy as np
import pytest

import oracle as O
from dcm_amd import workloads as W
from helpers import field, interior, small_domain

G = 9.81
KAPPA = 0.4
PRN = 0.95


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def _uniform(ip=8, jp=6, kp=10, d=40.0):
    return W.make_domain(ip, jp, kp, d, d, d)     # constant theta: thetan = 300


def _run_subgrid(dom, ph, u, v, w, th, nq=0, q=None):
    g = O.Grid(dom)
    P = O.Physics(dom, ph, nq)
    z = lambda: np.zeros(dom.shape)
    out = dict(vism=z(), vist=z(), strain=z(), su=z(), sv=z(), sw=z(), sth=z())
    viss = np.zeros((max(nq, 1),) + dom.shape)
    sq = np.zeros((max(nq, 1),) + dom.shape)
    qq = np.zeros((1,) + dom.shape) if q is None else q
    O.subgrid(g, P, u, v, w, u, v, w, th, qq, out["vism"], out["vist"], viss, out["strain"],
              out["su"], out["sv"], out["sw"], out["sth"], sq)
    out["sq"], out["viss"] = sq, viss
    return out


def _fields(dom):
    return (np.zeros(dom.shape) for _ in range(4))


@pytest.mark.parametrize("which", ["du/dz", "dw/dx", "dv/dx", "du/dx"])
def test_strain_of_simple_shears(which):
    """2 S_ij S_ij: du/dz = a or dw/dx = a gives a^2 (the s13 pair), dv/dx =
    a gives a^2 (the eight-point s12 average), du/dx = a gives 2 a^2."""
    dom = _uniform()
    ph = W.physics_profiles(dom)
    u, v, w, th = _fields(dom)
    a = 0.015
    I, J, K = np.meshgrid(np.arange(dom.ip + 2), np.arange(dom.jp + 2), np.arange(dom.kp + 2),
                          indexing="ij")
    d = dom.dx[1]
    if which == "du/dz":
        u[...] = a * dom.zn[K]
        want = a * a
    elif which == "dw/dx":
        w[...] = a * d * I
        want = a * a
    elif which == "dv/dx":
        v[...] = a * d * I
        want = a * a
    else:
        u[...] = a * d * I
        want = 2 * a * a
    out = _run_subgrid(dom, ph, u, v, w, th)
    np.testing.assert_allclose(interior(out["strain"])[1:-1, 1:-1, :], want, rtol=1e-12)


def test_neutral_smagorinsky_viscosity_and_prandtl_number():
    """Neutral shear (Ri = 0): nu_m = lambda^2 |S|, nu_h = nu_m / Pr (x rhoa),
    both zero in the two levels below the lid."""
    dom = _uniform()
    ph = W.physics_profiles(dom)
    u, v, w, th = _fields(dom)
    a = 0.02
    u[...] = a * dom.zn[None, None, :]
    out = _run_subgrid(dom, ph, u, v, w, th)
    kp = dom.kp
    lam = ph.lamsq[None, None, 1:kp - 1]
    vism = interior(out["vism"])[:, :, :kp - 2]
    vist = interior(out["vist"])[:, :, :kp - 2]
    np.testing.assert_allclose(vism, lam * a * np.ones_like(vism), rtol=1e-12)
    np.testing.assert_allclose(vist, vism / PRN * dom.rhoa[None, None, 1:kp - 1], rtol=1e-12)
    assert not interior(out["vism"])[:, :, kp - 2:].any()
    assert not interior(out["vist"])[:, :, kp - 2:].any()


def _ri_case(ri):
    """Linear shear a with a linear theta perturbation giving Richardson
    number ri everywhere: Ri = g (dth/dz) / (theta0 a^2)."""
    dom = _uniform()
    ph = W.physics_profiles(dom)
    u, v, w, th = _fields(dom)
    a = 0.02
    u[...] = a * dom.zn[None, None, :]
    th[...] = (ri * 300.0 * a * a / G) * dom.zn[None, None, :]
    out = _run_subgrid(dom, ph, u, v, w, th)
    k = 3
    lam = ph.lamsq[k]
    fm = out["vism"][4, 3, k] / (lam * a)
    fh = out["vist"][4, 3, k] / (lam * a * dom.rhoa[k])
    return fm, fh


def test_stability_functions_are_continuous_and_vanish_beyond_critical():
    fm0m, fh0m = _ri_case(-1e-9)
    fm0p, fh0p = _ri_case(1e-9)
    assert abs(fm0m - 1.0) < 1e-6 and abs(fm0p - 1.0) < 1e-6
    assert abs(fh0m - 1 / PRN) < 1e-6 and abs(fh0p - 1 / PRN) < 1e-6
    fmc, fhc = _ri_case(0.25 * (1 - 1e-4))          # just below Ri_c = 0.25
    assert 0.0 < fmc < 1e-12 and 0.0 < fhc < 1e-12
    assert _ri_case(0.3) == (0.0, 0.0)                # supercritical: no mixing
    vals = [_ri_case(r)[0] for r in (-0.5, -0.1, 0.0, 0.05, 0.1, 0.2)]
    assert all(x > y for x, y in zip(vals, vals[1:]))  # mixing decreases with stability
    assert _ri_case(-0.5)[0] > 1.0                     # convective enhancement


def test_surface_layer_neutral_log_law():
    """Neutral surface layer (theta(1) + th = theta(0)): u* = kappa U / ln(z1/z0)
    and the level-0 viscosity u*^2 (z1 - z0) / U; no heat flux."""
    dom = _uniform()
    ph = W.physics_profiles(dom, z0=0.05)
    u, v, w, th = _fields(dom)
    U = 3.5
    u[...] = U
    th[...] = dom.theta[0] - dom.thetan[1]          # th(1) + thetan(1) - theta(0) = 0
    out = _run_subgrid(dom, ph, u, v, w, th)
    z1 = dom.z[1]
    vmag = U + 1e-10
    ustar = KAPPA * vmag / math.log(z1 / 0.05)
    np.testing.assert_allclose(out["vism"][1:-1, 1:-1, 0], ustar ** 2 * (z1 - 0.05) / vmag,
                               rtol=1e-13)
    assert not out["vist"][1:-1, 1:-1, 0].any()
    # a warm surface (unstable) mixes more than a cold one (stable)
    th2 = th + 2.0
    warm = _run_subgrid(dom, ph, u, v, w, th2)["vism"][1:-1, 1:-1, 0]
    th3 = th - 2.0
    cold = _run_subgrid(dom, ph, u, v, w, th3)["vism"][1:-1, 1:-1, 0]
    # air warmer than the surface (dth > 0) is stable: less momentum mixing
    # than neutral; colder air is unstable: more
    assert (warm < out["vism"][1:-1, 1:-1, 0]).all()
    assert (cold > out["vism"][1:-1, 1:-1, 0]).all()


def test_scalar_diffusion_is_the_discrete_laplacian():
    """theta varying only in x, neutral shear in u: the x diffusion of a
    cosine is an eigenfunction of the second difference, weighted by the
    four surrounding viscosities: sth = 0.25 P (2cos(k dx) - 2) th / (dx^2 rhoan),
    P = 2 vist(k) + 2 vist(k-1) (columns uniform in x)."""
    dom = _uniform(16, 4, 8)
    ph = W.physics_profiles(dom)
    u, v, w, th = _fields(dom)
    u[...] = 0.02 * dom.zn[None, None, :]
    n = dom.ip
    kx = 2 * math.pi / n
    ii = np.arange(dom.ip + 2)
    th[...] = 0.5 * np.cos(kx * (ii - 1))[:, None, None]   # periodic, halos included
    out = _run_subgrid(dom, ph, u, v, w, th)
    vt = out["vist"]
    d = dom.dx[1]
    P = 2 * vt[1:-1, 1:-1, 1:-1] + 2 * vt[1:-1, 1:-1, :-2]
    want = 0.25 * P * (2 * math.cos(kx) - 2) * th[1:-1, 1:-1, 1:-1] / (d * d) \
        / dom.rhoan[None, None, 1:-1]
    # level 1 sits on the surface layer, whose viscosity follows th(1) and so
    # varies along x: the columns are uniform from level 2 up
    np.testing.assert_allclose(interior(out["sth"])[:, :, 1:], want[:, :, 1:], rtol=1e-10,
                               atol=1e-18)


def test_uniform_flow_has_no_viscous_momentum_tendency():
    dom = _uniform()
    ph = W.physics_profiles(dom)
    u, v, w, th = _fields(dom)
    u[...], v[...] = 2.0, -1.0
    out = _run_subgrid(dom, ph, u, v, w, th)
    assert not interior(out["strain"]).any()
    for name in ("su", "sv", "sw"):
        assert not interior(out[name])[:, :, 1:].any(), name


def _sources(dom, ph, u, v, w, th, ubar=None, vbar=None, wbar=None):
    g = O.Grid(dom)
    P = O.Physics(dom, ph, 0)
    for arr, val in ((P.ubar, ubar), (P.vbar, vbar), (P.wbar, wbar)):
        if val is not None:
            arr[:] = val
    z = lambda: np.zeros(dom.shape)
    out = dict(su=z(), sv=z(), sw=z(), sth=z(), rhs=z())
    thbar = np.zeros(dom.kp + 2)
    q = np.zeros((1,) + dom.shape)
    O.sources_phys(g, P, u, v, w, u, v, w, th, q, dom.theta, thbar, out["su"], out["sv"],
                   out["sw"], out["sth"], np.zeros((1,) + dom.shape), z(), z(), z(),
                   np.zeros((1,) + dom.shape), out["rhs"])
    return out


def test_sources_with_every_option_off_is_the_inviscid_branch():
    dom = small_domain(8, 7, 9, varying_theta=True)
    g = O.Grid(dom)
    u, v, w, th = (field(dom, 40 + n) for n in range(4))
    for a in (u, v, w, th):
        O.fill_full(g, a)
    ph = W.physics_profiles(dom, viscous=False)
    got = _sources(dom, ph, u, v, w, th)
    z = lambda: np.zeros(dom.shape)
    su, sv, sw, sth, rhs = z(), z(), z(), z(), z()
    O.sources(g, u, v, w, th, dom.theta, np.zeros(dom.kp + 2), su, sv, sw, sth, rhs)
    for name, ref in (("su", su), ("sv", sv), ("sw", sw), ("sth", sth), ("rhs", rhs)):
        np.testing.assert_array_equal(got[name], ref, err_msg=name)


def test_relaxation_terms_vanish_on_their_targets():
    """Damping toward the horizontal mean (and w, th toward 0), forcing toward
    the forcing profile and subsidence of the mean w: no tendency when the
    flow meets each target; a non-zero tendency of the right sign otherwise."""
    dom = _uniform(6, 5, 12)
    u, v, w, th = _fields(dom)
    prof_u = 1.0 + 0.01 * dom.zn
    prof_v = -0.5 + 0.02 * dom.zn
    u[...] = prof_u[None, None, :]
    v[...] = prof_v[None, None, :]
    base = _sources(dom, W.physics_profiles(dom, viscous=False), u, v, w, th)
    ph = W.physics_profiles(dom, viscous=False, damping_layer=True, forcing=True, u=u, v=v,
                            divergence=True, divergence_val=1e-5)
    on = _sources(dom, ph, u, v, w, th, ubar=prof_u, vbar=prof_v, wbar=ph.w_subs)
    for name in ("su", "sv", "sw", "sth"):
        np.testing.assert_array_equal(on[name], base[name], err_msg=name)
    # a warm anomaly in the damping layer is damped (dampfacn < 0)
    th2 = th.copy()
    th2[:, :, dom.kp - 1] = 1.0
    d = _sources(dom, ph, u, v, w, th2, ubar=prof_u, vbar=prof_v, wbar=ph.w_subs)
    assert (interior(d["sth"])[:, :, dom.kp - 2] < 0).all()


def test_scalar_diffusion_along_y_is_the_discrete_laplacian():
    dom = _uniform(4, 16, 8)
    ph = W.physics_profiles(dom)
    u, v, w, th = _fields(dom)
    u[...] = 0.02 * dom.zn[None, None, :]
    kx = 2 * math.pi / dom.jp
    jj = np.arange(dom.jp + 2)
    th[...] = 0.5 * np.cos(kx * (jj - 1))[None, :, None]
    out = _run_subgrid(dom, ph, u, v, w, th)
    vt = out["vist"]
    d = dom.dy[1]
    P = 2 * vt[1:-1, 1:-1, 1:-1] + 2 * vt[1:-1, 1:-1, :-2]
    want = 0.25 * P * (2 * math.cos(kx) - 2) * th[1:-1, 1:-1, 1:-1] / (d * d) \
        / dom.rhoan[None, None, 1:-1]
    np.testing.assert_allclose(interior(out["sth"])[:, :, 1:], want[:, :, 1:], rtol=1e-10,
                               atol=1e-18)


def test_vertical_scalar_diffusion_is_in_flux_form():
    """theta varying only in z: the rhoan dz weighted column sum of the
    tendency telescopes to the surface flux vist(0) (th(0) - th(1)) / dz(0)
    (the viscosity is zero below the lid)."""
    dom = small_domain(5, 4, 12, varying_theta=True)
    ph = W.physics_profiles(dom)
    u, v, w, th = _fields(dom)
    u[...] = 0.015 * dom.zn[None, None, :]
    th[...] = (0.3 * np.sin(0.7 * np.arange(dom.kp + 2)))[None, None, :]
    out = _run_subgrid(dom, ph, u, v, w, th)
    wgt = (dom.rhoan[1:-1] * dom.dz[:-2])[None, None, :]
    col = (interior(out["sth"]) * wgt).sum(axis=2)
    vt0 = out["vist"][1:-1, 1:-1, 0]
    flux = vt0 * (th[1:-1, 1:-1, 0] - th[1:-1, 1:-1, 1]) / dom.dz[0]
    np.testing.assert_allclose(col, flux, rtol=1e-9, atol=1e-15)
