"""The MPDATA arithmetic of sfvt/advection_3d.F90 (dcm_set_mpdata_variant
with DCM_MPDATA_SFVT) against the oracle under the same setting.

That copy of the module bounds each direction of the limiter separately
with unweighted denominators (:555-682, vertical :1028-1091) and uses
unweighted antidiffusive velocities in its vector routines (:1384-1507,
:1874-1909).  Every routine must match the oracle bit for bit on stretched
grids with varying density; the model must match the oracle's model bit for
bit under sequential sums; the mixed path must stay within the FP32 inner
tolerance of FP64; an ensemble must evolve its members as they evolve alone,
also when the setting changes between captured graphs."""
import numpy as np
import pytest

import oracle as O
from dcm_amd import hip as H
from dcm_amd import workloads as W
from helpers import field, lateral_edge_mask, rel_l2, small_domain

pytestmark = pytest.mark.gpu

SHAPES = [(8, 6, 5), (12, 10, 8), (5, 4, 16), (16, 16, 16)]


@pytest.fixture(scope="module", autouse=True)
def _oracle_sfvt():
    O.build()
    O.set_mpdata_variant(1)
    yield
    O.set_mpdata_variant(0)


def _ctx(dom):
    c = H.Context(dom)
    c.set_mpdata_variant(H.MPDATA_SFVT)
    return c


def _vel(dom, seed, courant=0.3, dt=1.0):
    g = O.Grid(dom)
    u = field(dom, seed, -1, 1) * courant * dom.dx[1] / dt
    v = field(dom, seed + 1, -1, 1) * courant * dom.dy[1] / dt
    w = field(dom, seed + 2, -1, 1) * courant * dom.dz[1] / dt
    for a in (u, v, w):
        O.fill_full(g, a, 0.0, 0.0)
    return u, v, w


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kord,monotone", [(2, True), (2, False), (3, True), (4, True)])
def test_sfvt_mpdata_bitwise(shape, kord, monotone):
    dom = small_domain(*shape, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    u, v, w = _vel(dom, 131)
    psi = field(dom, 134, -1.0, 3.0)
    O.fill_full(g, psi, 0.0, 0.0)
    with _ctx(dom) as c:
        du, dv, dw, dpsi = c.array(u), c.array(v), c.array(w), c.array(psi)
        ex, lb, ub = c.mpdata3d(1.0, du, dv, dw, dpsi, 0.0, 0.0, kord, monotone, 0)
        r, olb, oub = O.mpdata3d(g, 1.0, u, v, w, psi, 0.0, 0.0, kord, monotone, 0)
        assert ex == r == 0 and (lb, ub) == (olb, oub)
        np.testing.assert_array_equal(dpsi.get(), psi)


def test_sfvt_mpdata_bc_and_neumann():
    dom = small_domain(12, 10, 8, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    u, v, w = _vel(dom, 141)
    psi = field(dom, 144, 280.0, 300.0)
    O.fill_full(g, psi, 290.0, 295.0)
    with _ctx(dom) as c:
        du, dv, dw, dpsi = c.array(u), c.array(v), c.array(w), c.array(psi)
        ex, lb, ub = c.mpdata3d(0.7, du, dv, dw, dpsi, 290.0, 295.0, 2, True, 1)
        r, olb, oub = O.mpdata3d(g, 0.7, u, v, w, psi, 290.0, 295.0, 2, True, 1)
        assert (lb, ub) == (olb, oub)
        np.testing.assert_array_equal(dpsi.get(), psi)


@pytest.mark.parametrize("nq", [1, 3])
@pytest.mark.parametrize("kord,monotone", [(2, True), (2, False), (3, True)])
def test_sfvt_mpdata_vec_bitwise(nq, kord, monotone):
    dom = small_domain(12, 10, 8, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    u, v, w = _vel(dom, 161)
    q = np.stack([field(dom, 170 + n, 0.0, 1.0 + n) for n in range(nq)])
    for n in range(nq):
        O.fill_full(g, q[n], 0.0, 0.0)
    with _ctx(dom) as c:
        du, dv, dw, dq = c.array(u), c.array(v), c.array(w), c.array(q)
        c.mpdata_vec3d(1.0, du, dv, dw, dq, nq, kord=kord, monotone=monotone)
        O.mpdata_vec3d(g, 1.0, u, v, w, q, kord=kord, monotone=monotone)
        np.testing.assert_array_equal(dq.get(), q)


@pytest.mark.parametrize("kord,monotone", [(2, True), (3, True), (2, False)])
def test_sfvt_vertical_bitwise(kord, monotone):
    dom = small_domain(9, 7, 14, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    w = field(dom, 181, -1, 1) * 0.4 * dom.dz[1]
    O.fill_along_z(g, w, 0.0, 0.0)
    psi = field(dom, 182, 0.0, 2.0)
    O.fill_along_z(g, psi, 0.0, 0.0)
    q = np.stack([field(dom, 183 + n, 0.0, 1.0) for n in range(3)])
    with _ctx(dom) as c:
        dw, dpsi, dq = c.array(w), c.array(psi), c.array(q)
        c.mpdata_vert3d(1.0, dw, dpsi, 0.0, 0.0, kord, monotone, 0)
        O.mpdata_vert3d(g, 1.0, w, psi, 0.0, 0.0, kord, monotone, 0)
        np.testing.assert_array_equal(dpsi.get(), psi)
        c.mpdata_vec_vert3d(1.0, dw, dq, 3, kord=kord, monotone=monotone)
        O.mpdata_vec_vert3d(g, 1.0, w, q, kord=kord, monotone=monotone)
        np.testing.assert_array_equal(dq.get(), q)


def test_sfvt_mpdata3d_add_bitwise():
    dom = small_domain(10, 8, 9, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    u, v, w = _vel(dom, 191)
    psi = field(dom, 194, -0.5, 0.5)
    prof = dom.thetan.copy()
    with _ctx(dom) as c:
        du, dv, dw, dpsi = c.array(u), c.array(v), c.array(w), c.array(psi)
        c.mpdata3d_add(1.0, du, dv, dw, dpsi, prof, 0.0, 0.0, 2, True, 0)
        O.mpdata3d_add(g, 1.0, u, v, w, psi, prof, 0.0, 0.0, 2, True, 0)
        np.testing.assert_array_equal(dpsi.get(), psi)


def test_sfvt_setting_is_checked():
    dom = small_domain(4, 4, 4)
    with H.Context(dom) as c:
        with pytest.raises(H.DcmError):
            c.set_mpdata_variant(2)


def _moist(steps=4):
    return W.moist_bubble(12, 16, 10, delta=50.0, dt=0.5, steps=steps, ntracer=1, seed=12)


def test_sfvt_model_matches_oracle():
    """Moist bubble (theta by mpdata_3d, the q categories by mpdata_vec_3d)
    with the reference's sequential sums: bit-identical to the oracle."""
    case = _moist()
    om = O.Model(case)
    with _ctx(case.dom) as c:
        c.set_sum_order(H.SUM_SEQUENTIAL)
        gm = H.Model(c, case)
        try:
            for s in range(4):
                st = gm.step(1)
                om.step(1)
                assert st.iters_total == om.iters[-1], (s, st.iters_total, om.iters[-1])
            (u, v, w), _ = om.levels()
            mask = ~lateral_edge_mask(case.dom)
            for name, fid, n, ref in (("u", H.F_U, 0, u), ("v", H.F_V, 0, v), ("w", H.F_W, 0, w),
                                      ("th", H.F_TH, 0, om.th)):
                np.testing.assert_array_equal(gm.get(fid, n), ref, err_msg=name)
            for n in range(case.nq):
                np.testing.assert_array_equal(gm.get(H.F_Q, n), om.q[n], err_msg=f"q{n}")
            np.testing.assert_array_equal(gm.get(H.F_P)[mask], om.p[mask])
        finally:
            gm.close()


def test_sfvt_mixed_model_tracks_fp64():
    case = _moist(6)
    out = {}
    for prec in (H.FP64, H.MIXED):
        with _ctx(case.dom) as c:
            m = H.Model(c, case, prec)
            try:
                st = m.step(6)
                assert st.max_true_resid < 1e-9
                out[prec] = {n: m.get(f, 0) for n, f in (("u", H.F_U), ("w", H.F_W),
                                                        ("th", H.F_TH))}
                out[prec]["q0"] = m.get(H.F_Q, 0)
            finally:
                m.close()
    for name in out[H.FP64]:
        assert rel_l2(out[H.FP64][name], out[H.MIXED][name]) < 1e-5, name


def _snapshot(case, m):
    out = {name: m.get(fid) for fid, name in ((H.F_U, "u"), (H.F_W, "w"), (H.F_TH, "th"))}
    for n in range(case.nq):
        out[f"q{n}"] = m.get(H.F_Q, n)
    return out


def test_sfvt_ensemble_recaptures_when_the_setting_changes():
    """Two copies of a two-member ensemble, one replaying graphs and one
    eager; after four steps (graphs captured under the linked copy's
    arithmetic) every context switches to sfvt's, and the replaying copy must
    capture again rather than replay the old kernels."""
    cases = [_moist(8), W.warm_bubble(16, 12, 10, steps=8, amp=2.0, seed=11, name="a")]
    sets = []
    for _ in range(2):
        ctxs = [H.Context(cs.dom) for cs in cases]
        sets.append((ctxs, [H.Model(c, cs, H.MIXED) for c, cs in zip(ctxs, cases)]))
    try:
        (cg, mg), (ce, me) = sets
        with H.Ensemble(mg) as eg, H.Ensemble(me) as ee:
            ee.set_graphs(False)
            for s in range(8):
                if s == 4:
                    for c in cg + ce:
                        c.set_mpdata_variant(H.MPDATA_SFVT)
                sg, se = eg.step(1), ee.step(1)
                for q, case in enumerate(cases):
                    assert sg[q].iters_total == se[q].iters_total, (s, q)
                    a, b = _snapshot(case, mg[q]), _snapshot(case, me[q])
                    for name in a:
                        np.testing.assert_array_equal(a[name], b[name],
                                                      err_msg=f"step {s} member {q} {name}")
    finally:
        for ctxs, models in sets:
            for m in models:
                m.close()
            for c in ctxs:
                c.close()
