"""The subgrid model (calculate_subgrid_3d) and the full sources (damping
layer, forcing, subsidence, viscous terms) through the C-ABI, bit for bit
against the oracle: every output field, halos included."""
import numpy as np
import pytest

import oracle as O
from dcm_amd import hip as H
from dcm_amd import workloads as W
from helpers import field, small_domain

pytestmark = pytest.mark.gpu

SHAPES = [(8, 6, 7), (12, 10, 16), (5, 4, 33)]


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    O.build()


def _state(dom, seed, nq):
    g = O.Grid(dom)
    u, v, w = (field(dom, seed + n, -3.0, 3.0) for n in range(3))
    zu, zv, zw = (a + 0.1 * field(dom, seed + 10 + n) for n, a in enumerate((u, v, w)))
    th = field(dom, seed + 20, -0.5, 0.5) + 0.002 * dom.zn[None, None, :] * 50.0
    for a in (u, v, w, zu, zv, zw, th):
        O.fill_full(g, a, 0.0, 0.0)
    q = np.stack([field(dom, seed + 30 + n, 0.0, 1e-3) for n in range(max(nq, 1))])
    return u, v, w, zu, zv, zw, th, q


def _zeros(dom, nq=0):
    z = lambda: np.zeros(dom.shape)
    return dict(vism=z(), vist=z(), strain=z(), su=z(), sv=z(), sw=z(), sth=z(),
                viss=np.zeros((max(nq, 1),) + dom.shape), sq=np.zeros((max(nq, 1),) + dom.shape))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("nq", [0, 2])
def test_subgrid_bitwise(shape, nq):
    dom = small_domain(*shape, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    ph = W.physics_profiles(dom, z0=0.02, z0th=0.01)
    u, v, w, zu, zv, zw, th, q = _state(dom, 100, nq)
    # tendencies already hold the inviscid terms: start from random values
    base = {k: field(dom, 200 + n) for n, k in enumerate(("su", "sv", "sw", "sth"))}
    exp = _zeros(dom, nq)
    for k, a in base.items():
        exp[k] = a.copy()
    exp["sq"] = np.stack([field(dom, 210 + n) for n in range(max(nq, 1))])
    got0 = {k: a.copy() for k, a in exp.items()}
    O.subgrid(g, O.Physics(dom, ph, nq), u, v, w, zu, zv, zw, th, q, exp["vism"], exp["vist"],
              exp["viss"], exp["strain"], exp["su"], exp["sv"], exp["sw"], exp["sth"], exp["sq"])
    with H.Context(dom) as c:
        d = {k: c.array(a) for k, a in got0.items()}
        du, dv, dw, dzu, dzv, dzw, dth, dq = (c.array(a) for a in (u, v, w, zu, zv, zw, th, q))
        c.subgrid(ph, du, dv, dw, dzu, dzv, dzw, dth, dq if nq else None, nq, d["vism"],
                  d["vist"], d["viss"] if nq else None, d["strain"], d["su"], d["sv"], d["sw"],
                  d["sth"], d["sq"] if nq else None)
        for k in ("vism", "vist", "strain", "su", "sv", "sw", "sth") + (("sq", "viss") if nq else ()):
            np.testing.assert_array_equal(d[k].get(), exp[k], err_msg=k)


@pytest.mark.parametrize("opts", ["viscous", "relax", "all", "none"])
@pytest.mark.parametrize("nq", [0, 1])
def test_sources_phys_bitwise(opts, nq):
    dom = small_domain(10, 8, 14, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    u, v, w, zu, zv, zw, th, q = _state(dom, 300, nq)
    ph = W.physics_profiles(dom, u=u, v=v, viscous=opts in ("viscous", "all"),
                            damping_layer=opts in ("relax", "all"),
                            forcing=opts in ("relax", "all"), forcing_tau=900.0,
                            divergence=opts in ("relax", "all"), divergence_val=2e-5,
                            divergence_hgt=300.0)
    means = O.hmeans(g, u, v, w, th, q if nq else None)
    ubar, vbar, wbar, thbar = means[:4]
    P = O.Physics(dom, ph, nq)
    P.ubar[:], P.vbar[:], P.wbar[:] = ubar, vbar, wbar
    exp = _zeros(dom, nq)
    exp["rhs"] = np.zeros(dom.shape)
    O.sources_phys(g, P, u, v, w, zu, zv, zw, th, q, dom.theta, thbar, exp["su"], exp["sv"],
                   exp["sw"], exp["sth"], exp["sq"], exp["strain"], exp["vism"], exp["vist"],
                   exp["viss"], exp["rhs"])
    with H.Context(dom) as c:
        z = _zeros(dom, nq)
        z["rhs"] = np.zeros(dom.shape)
        d = {k: c.array(a) for k, a in z.items()}
        du, dv, dw, dzu, dzv, dzw, dth, dq = (c.array(a) for a in (u, v, w, zu, zv, zw, th, q))
        c.sources_phys(ph, ubar, vbar, wbar, du, dv, dw, dzu, dzv, dzw, dth,
                       dq if nq else None, nq, dom.theta, thbar, d["su"], d["sv"], d["sw"],
                       d["sth"], d["sq"], d["strain"], d["vism"], d["vist"], d["viss"], d["rhs"])
        names = ["su", "sv", "sw", "sth", "rhs"]
        if opts in ("viscous", "all"):
            names += ["vism", "vist", "strain"] + (["sq", "viss"] if nq else [])
        for k in names:
            np.testing.assert_array_equal(d[k].get(), exp[k], err_msg=k)


def test_full_size_viscous_sources_bitwise():
    """256x256x128, every option on, one tracer: the GPU sources equal the
    oracle's bit for bit on all 8.4 M cells."""
    dom = W.make_domain(256, 256, 128, 50.0, 50.0, 50.0)
    g = O.Grid(dom)
    print("full-size viscous sources 256x256x128", flush=True)
    u, v, w, zu, zv, zw, th, q = _state(dom, 500, 1)
    ph = W.physics_profiles(dom, u=u, v=v, damping_layer=True, forcing=True, divergence=True,
                            divergence_val=1e-5)
    means = O.hmeans(g, u, v, w, th, q)
    ubar, vbar, wbar, thbar = means[:4]
    P = O.Physics(dom, ph, 1)
    P.ubar[:], P.vbar[:], P.wbar[:] = ubar, vbar, wbar
    exp = _zeros(dom, 1)
    exp["rhs"] = np.zeros(dom.shape)
    O.sources_phys(g, P, u, v, w, zu, zv, zw, th, q, dom.theta, thbar, exp["su"], exp["sv"],
                   exp["sw"], exp["sth"], exp["sq"], exp["strain"], exp["vism"], exp["vist"],
                   exp["viss"], exp["rhs"])
    with H.Context(dom) as c:
        z = _zeros(dom, 1)
        z["rhs"] = np.zeros(dom.shape)
        d = {k: c.array(a) for k, a in z.items()}
        du, dv, dw, dzu, dzv, dzw, dth, dq = (c.array(a) for a in (u, v, w, zu, zv, zw, th, q))
        c.sources_phys(ph, ubar, vbar, wbar, du, dv, dw, dzu, dzv, dzw, dth, dq, 1, dom.theta,
                       thbar, d["su"], d["sv"], d["sw"], d["sth"], d["sq"], d["strain"],
                       d["vism"], d["vist"], d["viss"], d["rhs"])
        for k in ("su", "sv", "sw", "sth", "sq", "vism", "vist", "rhs"):
            np.testing.assert_array_equal(d[k].get(), exp[k], err_msg=k)
