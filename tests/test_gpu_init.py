"""Runs set up from the reference's own namelists on the GPU: the 3-D
initial state generated in HBM by dcm_model_init_fields from the profiles
(winds and moisture columns, the param_theta noise, the warm bubble, the
halos of initialisation.f90:861-870), and a few steps of such a run
bit-identical to the oracle's model started from the same state."""
import dataclasses
import os

import numpy as np
import pytest

import oracle as O
from dcm_amd import hip as H
from dcm_amd import init as I
from test_gpu_parity import _compare_models

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "namelists")
NML = os.path.join(GOLD, "namelist.in")
WI = os.path.join(GOLD, "westbrook_illingworth_nml.in")

M64 = (1 << 64) - 1


def _noise(seed, c):
    """The documented per-cell deviate (SplitMix64 of seed + (c+1) golden
    gamma), in uint64 arithmetic."""
    z = np.uint64((seed + (c + 1) * 0x9E3779B97F4A7C15) & M64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def _expected(case):
    """The initial fields built on the host from the same profiles."""
    rc, pr = case.init
    dom, cfg = case.dom, rc.cfg
    shape = dom.shape
    g = O.Grid(dom)
    u = np.broadcast_to(pr["u"], shape).copy()
    v = np.broadcast_to(pr["v"], shape).copy()
    w = np.zeros(shape)
    th = np.zeros(shape)
    if cfg.param_theta:
        c = np.arange(int(np.prod(shape)), dtype=np.uint64).reshape(shape)
        r = _noise(cfg.seed, c)
        zk = np.broadcast_to(pr["z"], shape)
        inside = np.zeros(shape, bool)
        inside[: dom.ip + 1, : dom.jp + 1, : dom.kp + 1] = True
        up = inside & (zk > cfg.param_z) & (zk <= cfg.param_z + cfg.param_delz)
        dn = inside & (zk > cfg.param_z - cfg.param_delz) & (zk <= cfg.param_z)
        th = np.where(up, -0.001 + (r - 0.5) / 20.0, th)
        th = np.where(dn, 0.001 - (r - 0.5) / 20.0, th)
    q = [np.broadcast_to(pr["q"][n], shape).copy() for n in range(case.nq)]
    for a in [u, v, w, th] + q:
        O.fill_full(g, a, 0.0, 0.0)
    return {"u": u, "v": v, "w": w, "th": th, "q": q}


@pytest.mark.parametrize("path,overrides", [
    (NML, {"jp": 24, "kp": 60}),                   # adiabatic profile, noise near 2500 m
    (WI, {"jp": 20, "kp": 40, "radiation": False}),  # a 69-level sounding with winds
])
def test_device_initial_state(path, overrides):
    case = I.case_from_namelist(path, steps=1, overrides=overrides)
    exp = _expected(case)
    with H.Context(case.dom) as c:
        m = H.Model(c, case)
        try:
            got = {"u": m.get(H.F_U), "v": m.get(H.F_V), "w": m.get(H.F_W), "th": m.get(H.F_TH),
                   "q": [m.get(H.F_Q, n) for n in range(case.nq)]}
        finally:
            m.close()
    for name in ("u", "v", "w", "th"):
        np.testing.assert_array_equal(got[name], exp[name], err_msg=name)
    for n in range(case.nq):
        np.testing.assert_array_equal(got["q"][n], exp["q"][n], err_msg=f"q{n}")
    rc = case.init[0]
    if rc.cfg.param_theta:
        layer = np.abs(got["th"][1:-1, 1:-1, 1:-1]) > 0
        assert layer.any()
        assert np.abs(got["th"]).max() <= 0.001 + 0.025 + 1e-15


def test_warm_bubble_option():
    case = I.case_from_namelist(WI, steps=1, overrides={"jp": 20, "kp": 16, "ip": 12,
                                                        "bubble": True, "radiation": False})
    with H.Context(case.dom) as c:
        m = H.Model(c, case)
        try:
            th = m.get(H.F_TH)
        finally:
            m.close()
    dom = case.dom
    zn, xn, yn = dom.zn[None, None, :], dom.xn[:, None, None], dom.yn[None, :, None]
    rad = np.sqrt((zn - 350.0) ** 2 + xn ** 2 + yn ** 2)
    inner = th[1:-1, 1:-1, 1:-1]
    np.testing.assert_array_equal(inner, np.where(rad <= 300.0, 1.0, 0.0)[1:-1, 1:-1, 1:-1])


def test_namelist_run_matches_oracle():
    """Four steps of the westbrook_illingworth run (subgrid model, damping
    layer, subsidence, 3 moisture fields; its radiation switched off) on a
    reduced grid, from the device-generated state, bit-identical to the
    oracle's model."""
    case = I.case_from_namelist(WI, steps=4, overrides={"jp": 24, "kp": 40,
                                                        "radiation": False})
    with H.Context(case.dom) as c:
        m = H.Model(c, case)
        try:
            fields = {"u": m.get(H.F_U), "v": m.get(H.F_V), "w": m.get(H.F_W),
                      "th": m.get(H.F_TH),
                      "q": np.stack([m.get(H.F_Q, n) for n in range(case.nq)])}
        finally:
            m.close()
    host = dataclasses.replace(case, **fields)
    host.init = None
    _compare_models(host, 4, seq=True)


def test_run_from_namelist_writes_the_output_file(tmp_path):
    """The model_driver loop from namelist.in on a reduced grid: five steps
    of 2.5 s with a record every 5 s (times 0, 5, 10, at the start of steps
    1, 3, 5), the first record the initial state."""
    from scipy.io import netcdf_file

    from dcm_amd import run as R
    over = {"jp": 24, "kp": 40, "output_interval": 5.0}
    path = tmp_path / "run.nc"
    res = R.run_namelist(NML, steps=5, output=str(path), overrides=over)
    assert res["steps"] == 5 and res["records"] == 3
    assert res["max_true_resid"] < 1e-8
    case = I.case_from_namelist(NML, steps=1, overrides=over)
    with H.Context(case.dom) as c:
        m = H.Model(c, case, H.MIXED)
        try:
            th0, q0 = m.get(H.F_TH), [m.get(H.F_Q, n) for n in range(case.nq)]
        finally:
            m.close()
    with netcdf_file(path, "r", mmap=False) as nc:
        np.testing.assert_array_equal(nc.variables["time"][:], np.array([0, 5, 10], np.float32))
        assert nc.variables["q"].shape == (3, 1, 24, 40, case.nq)
        np.testing.assert_array_equal(nc.variables["th"][0],
                                      th0[1:-1, 1:-1, 1:-1].astype(np.float32))
        for n in range(case.nq):
            np.testing.assert_array_equal(nc.variables["q"][0][..., n],
                                          q0[n][1:-1, 1:-1, 1:-1].astype(np.float32))
        for name in ("u", "v", "w", "th", "p", "div"):
            assert np.isfinite(nc.variables[name][:]).all(), name
        np.testing.assert_array_equal(nc.variables["trefn"][:],
                                      case.trefn[1:-1].astype(np.float32))
