"""The run set-up from the reference's own namelists: the namelist reader
against the values written in the files (tests/golden/namelists, copies of
the reference's namelist.in and two run_scripts namelists), and the 1-D
profiles of dcm_init_compute_profiles against closed forms and against an
independent restatement of allocate_and_set / setup_adiabatic_profile with
scipy's ODE integrator and root finder (oracle/init_oracle.py)."""
import math
import os

import numpy as np
import pytest

import init_oracle as IO
from dcm_amd import hip as H
from dcm_amd import init as I

pytestmark = pytest.mark.skipif(not os.path.exists(H.LIB_PATH), reason="libdcm_hip.so not built")

GOLD = os.path.join(os.path.dirname(__file__), "golden", "namelists")
NML = os.path.join(GOLD, "namelist.in")
WI = os.path.join(GOLD, "westbrook_illingworth_nml.in")
KH = os.path.join(GOLD, "namelist_kelvin_helmholtz.in")


def test_reference_namelist_values():
    rc = I.RunConfig(NML)
    nm = rc.nm
    assert (nm["ip"], nm["jp"], nm["kp"], nm["nq"], nm["nprec"], nm["n_levels"]) == \
        (1, 101, 120, 3, 1, 7)
    assert (nm["dt"], nm["runtime"], nm["dx"], nm["dz"], nm["ptol"]) == (2.5, 28800.0, 200.0,
                                                                        100.0, 1e-8)
    assert (nm["kord"], nm["advection_scheme"], nm["subgrid_model"]) == (3, 1, 2)
    assert nm["monotone"] and nm["moisture"] and nm["adiabatic_prof"] and nm["param_theta"]
    assert not nm["divergence"] and not nm["bubble"] and not nm["param_wind"]
    assert (nm["t_cbase"], nm["t_ctop"], nm["tsurf"], nm["psurf"]) == (281.0, 273.0, 290.9, 1e5)
    assert nm["outputfile"] == "/tmp/output.nc"
    np.testing.assert_array_equal(rc.z_read, [0, 740, 1600, 1601, 6260, 6261, 12000])
    np.testing.assert_array_equal(rc.theta_read, [290.9] * 5 + [300.9, 400.0])
    np.testing.assert_array_equal(rc.v_read, [0, 0, 0, 1, 1, 0, 0])
    q = rc.q_read.reshape(7, 3).T  # q_read(nq, n_levels), Fortran order
    np.testing.assert_array_equal(q[0], [0.0099, 0.0099, 0.0099, 1.4e-4, 1.4e-4, 1.4e-11, 1.4e-11])
    assert not q[1:].any()


def test_long_sounding_namelist():
    rc = I.RunConfig(WI)
    assert rc.nm["n_levels"] == 69 and not rc.nm["adiabatic_prof"]
    assert rc.z_read.shape == (69,) and np.all(np.diff(rc.z_read) > 0)
    assert rc.z_read[0] == 0.0 and rc.theta_read[0] == 280.0
    assert rc.nm["divergence"] and rc.nm["divergence_val"] > 0


def test_namelist_syntax(tmp_path):
    p = tmp_path / "t.nml"
    p.write_text("""junk before the group
&grp
  A = 1.5d0, ! a comment
  b(2:4) = 3*7,
  C(1,2) = 1, 2,   ! fills on from element (1,2)
  s = 'it''s, here',
  t = "q/uote"
  l1 = .true., L2 = F, l3=.f.
  arr = 1., , 3.,
  e = 2.5e-3 /
&other x = 4 /
""")
    with I.Namelist(p) as nl:
        assert nl.has("grp") and nl.has("GRP", "a") and not nl.has("grp", "zz")
        assert nl.real("grp", "a") == 1.5
        np.testing.assert_array_equal(nl.array("grp", "b", (5,)), [0, 7, 7, 7, 0])
        c = nl.array("grp", "c", (2, 3), fill=-1.0)
        assert c[0, 1] == 1.0 and c[1, 1] == 2.0 and c[0, 0] == -1.0
        assert nl.string("grp", "s") == "it's, here"
        assert nl.string("grp", "t") == "q/uote"
        assert nl.logical("grp", "l1") is True
        assert nl.logical("grp", "l2") is False and nl.logical("grp", "l3") is False
        np.testing.assert_array_equal(nl.array("grp", "arr", (3,), fill=9.0), [1, 9, 3])
        assert nl.real("grp", "e") == 2.5e-3
        assert nl.int("other", "x") == 4
        assert nl.real("grp", "missing", "dflt") == "dflt"
        with pytest.raises(H.DcmError):
            nl.array("grp", "b", (3,))  # b(2:4) past a length-3 array


def _small(path, kp=None, jp=None):
    rc = I.RunConfig(path)
    if kp:
        rc.cfg.kp = kp
    if jp:
        rc.cfg.jp = jp
    return rc


def test_grid_mixing_length_damping_and_subsidence():
    rc = _small(WI)
    c, nm = rc.cfg, rc.nm
    pr = I.compute_profiles(rc)
    x, y, z, xn, yn, zn = IO.grid(c.ip, c.jp, c.kp, c.dx, c.dy, c.dz)
    for name, ref in (("x", x), ("y", y), ("z", z), ("xn", xn), ("yn", yn), ("zn", zn)):
        np.testing.assert_array_equal(pr[name], ref, err_msg=name)
    # y = dy f - (jp-1)/2 dy counts f from the halo: y = 0 at f = (jp-1)/2
    assert pr["y"][(c.jp - 1) // 2] == 0.0
    lam = 1.0 / (1.0 / (nm["cvis"] * (c.dx + c.dy + c.dz) / 3.0) ** 2 +
                 1.0 / (0.4 * (z + nm["z0"])) ** 2)
    np.testing.assert_allclose(pr["lamsq"], lam, rtol=1e-15)
    ztop = zn[c.kp]
    np.testing.assert_array_equal(pr["dampfacn"] != 0, zn > ztop - nm["damping_thickness"])
    np.testing.assert_array_equal(pr["dampfac"] != 0, z > ztop - nm["damping_thickness"])
    assert set(np.unique(pr["dampfac"])) <= {0.0, -1.0 / nm["damping_tau"]}
    ws = np.where(z <= nm["divergence_hgt"], -nm["divergence_val"] * z,
                  -nm["divergence_val"] * nm["divergence_hgt"])
    np.testing.assert_array_equal(pr["w_subs"], ws)
    assert pr["ntim"] == math.ceil(nm["runtime"] / nm["dt"])


def test_constant_theta_sounding_is_the_dry_adiabat():
    """theta = 300 K everywhere: pi(z) = pi_s - g z / (cp theta) exactly."""
    rc = _small(WI)
    rc.theta_read[:] = 300.0  # the configuration points at this array
    pr = I.compute_profiles(rc)
    pis = (rc.nm["psurf"] / 1e5) ** IO.KAPPA
    for zz, pp, name in ((pr["z"][1:], pr["pref"][1:], "w"), (pr["zn"], pr["prefn"], "scalar")):
        exact = 1e5 * (pis - IO.GRAV * zz / (IO.CP * 300.0)) ** (1.0 / IO.KAPPA)
        np.testing.assert_allclose(pp, exact, rtol=2e-14, err_msg=name)
    assert pr["pref"][0] == rc.nm["psurf"]
    np.testing.assert_array_equal(pr["theta"], 300.0)


def test_sounding_reference_state_matches_the_ode():
    rc = _small(WI, kp=60)
    pr = I.compute_profiles(rc)
    ref = IO.reference_state_sounding(rc.nm, rc.z_read, rc.theta_read, pr["z"], pr["zn"])
    for name in ("theta", "thetan"):
        np.testing.assert_allclose(pr[name], ref[name], rtol=1e-14, err_msg=name)
    for name in ("pref", "prefn", "tref", "trefn", "rhoa", "rhoan"):
        np.testing.assert_allclose(pr[name], ref[name], rtol=1e-10, err_msg=name)


def test_moisture_split_at_saturation():
    """Non-adiabatic start: q_read(1,:) on the scalar levels, the excess over
    saturation (at thetan, prefn) becomes cloud with drop_num drops."""
    rc = _small(WI, kp=60)
    pr = I.compute_profiles(rc)
    q1 = IO.sounding(rc.z_read, rc.q_read.reshape(-1, 3)[:, 0], pr["zn"])
    t = pr["thetan"] * (pr["prefn"] / rc.nm["psurf"]) ** IO.KAPPA
    qsat = IO.EPS1 * IO.svp_liq(t) / (pr["prefn"] - IO.svp_liq(t))
    cloudy = q1 >= qsat
    np.testing.assert_allclose(pr["q"][0], np.where(cloudy, qsat, q1), rtol=1e-12)
    np.testing.assert_allclose(pr["q"][1], np.where(cloudy, q1 - qsat, 0.0), rtol=1e-9,
                               atol=1e-15)
    np.testing.assert_array_equal(pr["q"][2], np.where(cloudy, rc.nm["drop_num"], 0.0))


def test_adiabatic_profile_matches_the_ode_and_conserves_theta_q():
    rc = _small(NML, kp=80)
    nm = dict(rc.nm)
    pr = I.compute_profiles(rc)
    w = IO.adiabatic_column(nm, pr["z"], True)
    s = IO.adiabatic_column(nm, pr["zn"], False)
    for mine, ref in (("pref", w["p"]), ("tref", w["t"]), ("theta", w["theta"]),
                      ("rhoa", w["rho"]), ("prefn", s["p"]), ("trefn", s["t"]),
                      ("thetan", s["theta"]), ("rhoan", s["rho"])):
        np.testing.assert_allclose(pr[mine], ref, rtol=2e-9, err_msg=mine)
    np.testing.assert_allclose(pr["q"][0], s["qv"], rtol=2e-9)  # vapour from the scalar pass
    assert not pr["q"][1].any() and not pr["q"][2].any()  # that pass sets no cloud water
    # physics: below cloud base the dry adiabat (theta = tsurf); inside the
    # cloud the moist potential temperature of cloud base; above cloud top
    # vapour at rh_above of saturation
    zn, p, t = pr["zn"], pr["prefn"], pr["trefn"]
    below = zn <= s["zcb"]
    np.testing.assert_allclose(pr["thetan"][below], nm["tsurf"], rtol=1e-14)
    inside = (zn > s["zcb"]) & (zn < s["zct"])
    assert inside.sum() > 3
    tq = np.array([s["theta_q"](tt, pp) for tt, pp in zip(t[inside], p[inside])])
    np.testing.assert_allclose(tq, s["tq_sat"], rtol=1e-11)
    above = zn > s["zct"]
    qsat = IO.EPS1 * IO.svp_liq(t[above]) / (p[above] - IO.svp_liq(t[above]))
    np.testing.assert_allclose(pr["q"][0][above], nm["rh_above"] * qsat, rtol=1e-13)


def test_param_wind_profile():
    rc = _small(KH, kp=100)
    nm = rc.nm
    assert nm["param_wind"]
    pr = I.compute_profiles(rc)
    zn = pr["zn"]
    from scipy.special import erf
    v = (0.5 * erf((zn - nm["param_z"]) / nm["param_sigz"]) + 0.5) * nm["param_vmax"]
    v[[0, 1, rc.cfg.kp]] = 0.0
    np.testing.assert_allclose(pr["v"], v, rtol=1e-14, atol=1e-300)
    assert not pr["u"].any()
    np.testing.assert_array_equal(pr["v_force"], pr["v"] if nm["forcing"] else 0.0)


def test_bad_configuration_is_an_error():
    rc = _small(WI)
    rc.cfg.dz = 0.0
    with pytest.raises(H.DcmError):
        I.compute_profiles(rc)
