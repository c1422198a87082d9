"""The output file written from the resident model (dcm_output_*): records
packed on the GPU, copied on a side stream and written by a host thread.
Each record must hold the fields as they were when dcm_output_write was
called (float32 of the FP64 state), and div the divergence of the momentum
at the start of the previous step, as model_driver writes it
(driver_code.f90:310-327 and :443-444); checked with scipy's reader."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import oracle as O
from dcm_amd import hip as H
from dcm_amd import workloads as W
from helpers import interior

pytestmark = pytest.mark.gpu
F32 = np.float32


def _snap(m, nq):
    s = {name: m.get(fid) for fid, name in ((H.F_U, "u"), (H.F_V, "v"), (H.F_W, "w"),
                                            (H.F_TH, "th"), (H.F_P, "p"))}
    s["q"] = [m.get(H.F_Q, n) for n in range(nq)]
    return s


def _div(dom, u, v, w):
    d = np.zeros(dom.shape)
    O.divergence(O.Grid(dom), u, v, w, d)
    return d


def _check_record(nc, r, dom, snap, div, t):
    assert nc.variables["time"][r] == F32(t)
    for name in ("u", "v", "w", "th", "p"):
        np.testing.assert_array_equal(nc.variables[name][r], interior(snap[name]).astype(F32),
                                      err_msg=f"{name} record {r}")
    np.testing.assert_array_equal(nc.variables["div"][r], interior(div).astype(F32),
                                  err_msg=f"div record {r}")
    for n, q in enumerate(snap["q"]):
        np.testing.assert_array_equal(nc.variables["q"][r][..., n], interior(q).astype(F32),
                                      err_msg=f"q{n} record {r}")


@pytest.mark.parametrize("precision", [H.FP64, H.MIXED])
def test_model_records_hold_the_state_at_each_write(tmp_path, precision):
    case = W.moist_bubble(20, 12, 16, delta=50.0, dt=0.5, steps=4, ntracer=1, seed=31)
    dom, nq = case.dom, case.nq
    path = tmp_path / "model.nc"
    expect = []
    with H.Context(dom) as ctx:
        m = H.Model(ctx, case, precision)
        try:
            with H.Output(m, path, q_names=["qv", "tracer"]) as out:
                out.write(0.0)  # before any step: div = 0 (initialisation.f90:461)
                expect.append((0.0, _snap(m, nq), np.zeros(dom.shape)))
                for r in range(1, 4):
                    prev = _snap(m, nq)
                    m.step(1)
                    t = r * case.dt
                    out.write(t)
                    expect.append((t, _snap(m, nq), _div(dom, prev["u"], prev["v"], prev["w"])))
                out.flush()
                n, stall, wr = out.stats()
                assert n == 4 and stall >= 0.0 and wr >= 0.0
        finally:
            m.close()
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.variables["th"].shape == (4, dom.ip, dom.jp, dom.kp)
        assert nc.variables["q"].shape == (4, dom.ip, dom.jp, dom.kp, nq)
        assert np.all(nc.variables["precip"][:] == F32(9.9692099683868690e36))
        for r, (t, snap, div) in enumerate(expect):
            _check_record(nc, r, dom, snap, div, t)


def test_back_to_back_writes_wait_for_a_free_buffer(tmp_path):
    """Six writes without a step between them (two record buffers): every
    record complete and equal, the host precip array packed per record."""
    case = W.warm_bubble(16, 16, 12, steps=1, seed=33)
    dom = case.dom
    path = tmp_path / "burst.nc"
    with H.Context(dom) as ctx:
        m = H.Model(ctx, case)
        try:
            m.step(2)
            snap = _snap(m, 0)
            with H.Output(m, path) as out:
                for r in range(6):
                    out.write(float(r))
        finally:
            m.close()
    with netcdf_file(path, "r", mmap=False) as nc:
        assert "q" not in nc.variables
        assert nc.variables["th"].shape[0] == 6
        for r in range(6):
            assert nc.variables["time"][r] == F32(r)
            for name in ("u", "w", "th", "p"):
                np.testing.assert_array_equal(nc.variables[name][r],
                                              interior(snap[name]).astype(F32))


def test_precip_from_the_host(tmp_path):
    case = W.moist_bubble(12, 10, 8, delta=50.0, dt=0.5, steps=1, ntracer=0, seed=35)
    dom = case.dom
    nprec = 2
    rng = np.random.default_rng(1)
    prs = [rng.random((nprec, dom.ip + 2, dom.jp + 2, dom.kp)) for _ in range(2)]
    path = tmp_path / "precip.nc"
    with H.Context(dom) as ctx:
        m = H.Model(ctx, case)
        try:
            with H.Output(m, path, nprec=nprec) as out:
                for pr in prs:
                    m.step(1)
                    out.write(1.0, precip=pr)
                # without precip the record buffers hold the fill value again
                out.write(2.0)
                out.write(3.0)
        finally:
            m.close()
    with netcdf_file(path, "r", mmap=False) as nc:
        for r, pr in enumerate(prs):
            for n in range(nprec):
                np.testing.assert_array_equal(nc.variables["precip"][r][..., n],
                                              pr[n][1:-1, 1:-1, :].astype(F32))
        assert np.all(nc.variables["precip"][2:] == F32(9.9692099683868690e36))


def test_output_attached_inside_an_ensemble(tmp_path):
    """An output attached after the ensemble captured its graphs: the member's
    steps must compute div from then on (new graphs), and the record equals
    the member's state; the other member is unaffected (bit-identical to a
    run without the output)."""
    def members():
        return [W.warm_bubble(16, 12, 10, steps=1, amp=2.0, seed=41, name="a"),
                W.moist_bubble(12, 16, 8, delta=50.0, dt=0.5, steps=1, ntracer=1, seed=42)]

    def run(with_output):
        cases = members()
        ctxs = [H.Context(c.dom) for c in cases]
        models = [H.Model(x, c, H.MIXED) for x, c in zip(ctxs, cases)]
        got = {}
        try:
            with H.Ensemble(models) as e:
                e.step(4)  # captures the graphs of all three phases
                out = H.Output(models[1], tmp_path / "member.nc") if with_output else None
                prev = _snap(models[1], cases[1].nq)
                e.step(1)
                if out:
                    out.write(2.5)
                    out.close()
                got["snap"] = _snap(models[1], cases[1].nq)
                got["prev"] = prev
                got["other"] = _snap(models[0], 0)
        finally:
            for m in models:
                m.close()
            for x in ctxs:
                x.close()
        return cases, got

    cases, a = run(True)
    _, b = run(False)
    dom = cases[1].dom
    with netcdf_file(tmp_path / "member.nc", "r", mmap=False) as nc:
        _check_record(nc, 0, dom, a["snap"], _div(dom, a["prev"]["u"], a["prev"]["v"],
                                                  a["prev"]["w"]), 2.5)
    for name in ("u", "v", "w", "th", "p"):
        np.testing.assert_array_equal(a["other"][name], b["other"][name])
        np.testing.assert_array_equal(a["snap"][name], b["snap"][name])
