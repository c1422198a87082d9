"""The output file of driver_code.f90:725-1145 written by the library's
NetCDF classic writer from host arrays (dcm_nc_*), read back with scipy's
independent reader of the classic format: the reference's dimensions,
variables, dimension order, units and values, and the indexing its plotting
script (python/example_plot.py) uses."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from dcm_amd import hip as H
from helpers import field, interior, small_domain

pytestmark = pytest.mark.skipif(not os.path.exists(H.LIB_PATH), reason="libdcm_hip.so not built")

VARS = ["time", "x", "y", "z", "rhoa", "thetan", "trefn", "th", "u", "v", "w", "p", "div"]
MOIST = ["q_names", "q", "precip"]
UNITS = {"time": b"seconds", "x": b"m", "y": b"m", "z": b"m", "rhoa": b"kg/m3",
         "thetan": b"K", "trefn": b"K", "th": b"K", "u": b"m/s", "v": b"m/s", "w": b"m/s",
         "p": b"pa", "div": b"m/s**2", "q": b"kg or number per kg",
         "precip": b"mm hr-1 or number per m-2 s-1"}
F32 = np.float32


def _records(dom, nq, nprec, nrec, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrec):
        fl = [field(dom, 1000 * seed + 10 * r + a, -5.0, 5.0) for a in range(6)]
        q = np.stack([field(dom, 500 + 10 * r + n, 0.0, 1e-2) for n in range(nq)]) if nq else None
        pr = rng.random((nprec, dom.ip + 2, dom.jp + 2, dom.kp)) if nprec else None
        out.append((0.5 * r + 0.25, fl, q, pr))
    return out


def _write(path, dom, recs, moisture, nq, nprec, names=None, trefn=None, precip=True):
    with H.NcFile(path, dom, moisture=moisture, nq=nq, nprec=nprec, q_names=names,
                  trefn=trefn) as f:
        for r, (t, fl, q, pr) in enumerate(recs):
            f.put_record(r + 1, t, *fl, q=q, precip=pr if precip else None)


def test_moist_file_layout_and_values(tmp_path):
    dom = small_domain(7, 5, 6, stretched=True, varying_theta=True)
    nq, nprec = 3, 2
    trefn = 280.0 + np.arange(dom.kp + 2)
    recs = _records(dom, nq, nprec, 3)
    path = tmp_path / "out.nc"
    _write(path, dom, recs, True, nq, nprec, names=["vapour", "cloud", "rain"], trefn=trefn)
    assert open(path, "rb").read(4) == b"CDF\x01"
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.dimensions == {"times": None, "ip": 7, "jp": 5, "kp": 6, "nq": 3, "nprec": 2,
                                 "l_q_names": 20}
        assert list(nc.variables) == VARS + MOIST
        for name, var in nc.variables.items():
            assert getattr(var, "units", None) == UNITS.get(name), name
        for name in ("th", "u", "v", "w", "p", "div"):
            assert nc.variables[name].dimensions == ("times", "ip", "jp", "kp")
        assert nc.variables["q"].dimensions == ("times", "ip", "jp", "kp", "nq")
        assert nc.variables["precip"].dimensions == ("times", "ip", "jp", "kp", "nprec")
        assert nc.variables["q_names"].dimensions == ("nq", "l_q_names")
        names = [b"".join(r).decode() for r in nc.variables["q_names"][:]]
        assert names == ["vapour".ljust(20), "cloud".ljust(20), "rain".ljust(20)]
        for name, prof in (("x", dom.x), ("y", dom.y), ("z", dom.z), ("rhoa", dom.rhoa),
                           ("thetan", dom.thetan), ("trefn", trefn)):
            np.testing.assert_array_equal(nc.variables[name][:], prof[1:-1].astype(F32))
        np.testing.assert_array_equal(nc.variables["time"][:],
                                      np.array([r[0] for r in recs], F32))
        for r, (t, fl, q, pr) in enumerate(recs):
            for name, a in zip(("u", "v", "w", "th", "p", "div"), fl):
                np.testing.assert_array_equal(nc.variables[name][r], interior(a).astype(F32),
                                              err_msg=f"{name} record {r}")
            for n in range(nq):
                np.testing.assert_array_equal(nc.variables["q"][r][..., n],
                                              interior(q[n]).astype(F32))
            for n in range(nprec):
                np.testing.assert_array_equal(nc.variables["precip"][r][..., n],
                                              pr[n][1:-1, 1:-1, :].astype(F32))


def test_the_plot_scripts_slice(tmp_path):
    """python/example_plot.py draws nc['q'][ii, 0, :, :, n] against y and z:
    the y-z section at the first x point of field n, (jp, kp)."""
    dom = small_domain(6, 8, 5)
    recs = _records(dom, 2, 1, 2, seed=3)
    path = tmp_path / "out.nc"
    _write(path, dom, recs, True, 2, 1)
    with netcdf_file(path, "r", mmap=False) as nc:
        sl = np.squeeze(nc.variables["q"][-1, 0, :, :, 1])
        assert sl.shape == (dom.jp, dom.kp)
        np.testing.assert_array_equal(sl, recs[-1][2][1][1, 1:-1, 1:-1].astype(F32))
        np.testing.assert_array_equal(nc.variables["y"][:], dom.y[1:-1].astype(F32))
        np.testing.assert_array_equal(nc.variables["z"][:], dom.z[1:-1].astype(F32))


def test_dry_file_has_no_moisture_variables(tmp_path):
    dom = small_domain(4, 3, 5)
    recs = _records(dom, 0, 0, 2, seed=5)
    path = tmp_path / "dry.nc"
    _write(path, dom, recs, False, 0, 0)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.dimensions == {"times": None, "ip": 4, "jp": 3, "kp": 5}
        assert list(nc.variables) == VARS
        assert nc.variables["th"].shape == (2, 4, 3, 5)
        np.testing.assert_array_equal(nc.variables["th"][1], interior(recs[1][1][3]).astype(F32))


def test_missing_precip_is_the_netcdf_fill_value(tmp_path):
    dom = small_domain(4, 4, 4)
    recs = _records(dom, 1, 2, 1, seed=7)
    path = tmp_path / "fill.nc"
    _write(path, dom, recs, True, 1, 2, precip=False)
    with netcdf_file(path, "r", mmap=False) as nc:
        pr = nc.variables["precip"][0]
        assert pr.shape == (4, 4, 4, 2)
        assert np.all(pr == F32(9.9692099683868690e36))


def test_records_out_of_order_and_rewritten(tmp_path):
    """A record is addressed by its index n (start = (/.., n/)): writing 3
    before 2 and rewriting 1 leaves three records with the last values."""
    dom = small_domain(3, 3, 3)
    recs = _records(dom, 0, 0, 4, seed=9)
    path = tmp_path / "order.nc"
    with H.NcFile(path, dom) as f:
        for n, r in ((1, 0), (3, 2), (2, 1), (1, 3)):
            t, fl, _, _ = recs[r]
            f.put_record(n, t, *fl)
    with netcdf_file(path, "r", mmap=False) as nc:
        th = nc.variables["th"]
        assert th.shape[0] == 3
        for n, r in ((0, 3), (1, 1), (2, 2)):
            np.testing.assert_array_equal(th[n], interior(recs[r][1][3]).astype(F32))


def test_bad_arguments_are_errors(tmp_path):
    dom = small_domain(3, 3, 3)
    with pytest.raises(H.DcmError):
        H.NcFile(tmp_path / "x.nc", dom, moisture=True, nq=0, nprec=1)
    with pytest.raises(H.DcmError):  # a zero-length dimension would be a second record one
        H.NcFile(tmp_path / "x.nc", dom, moisture=True, nq=2, nprec=0)
    with pytest.raises(H.DcmError):
        H.NcFile(tmp_path / "no_such_dir" / "x.nc", dom)
    with H.NcFile(tmp_path / "ok.nc", dom) as f:
        fl = [field(dom, a) for a in range(6)]
        with pytest.raises(H.DcmError):
            f.put_record(0, 0.0, *fl)  # n is 1-based
