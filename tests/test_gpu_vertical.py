"""Vertical-only advection (first_order_upstream_vert_3d, mpdata_vert_3d,
mpdata_vec_vert_3d) and mpdata_3d_add through the C-ABI, bit for bit
against the oracle; PAMM's own call (kord 1, no limiter, neumann 2, three
fields, sub-stepped) included, and a full-size sedimentation step checked
against the closed-form loss through the floor."""
import numpy as np
import pytest

import oracle as O
from dcm_amd import hip as H
from dcm_amd import workloads as W
from helpers import field, interior, small_domain

pytestmark = pytest.mark.gpu

SHAPES = [(8, 6, 5), (12, 10, 8), (5, 4, 16), (16, 16, 33)]


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    O.build()


def _w(dom, seed, courant=0.4, dt=1.0, falling=False):
    w = field(dom, seed, -1, 1) * courant * dom.dz[1] / dt
    if falling:
        w = -np.abs(w)
    return w


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("neumann", [0, 1, 2])
def test_upwind_vert_bitwise(shape, neumann):
    dom = small_domain(*shape, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    w = _w(dom, 3)
    psi = field(dom, 4, 0.0, 2.0)
    with H.Context(dom) as c:
        dw, dpsi = c.array(w), c.array(psi)
        c.upwind_vert3d(0.8, dw, dpsi, neumann)
        O.upwind_vert3d(g, 0.8, w, psi, neumann)
        np.testing.assert_array_equal(dpsi.get(), psi)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kord,monotone", [(0, False), (1, False), (2, False), (2, True),
                                           (3, True), (3, False)])
@pytest.mark.parametrize("neumann", [0, 2])
def test_mpdata_vert_bitwise(shape, kord, monotone, neumann):
    dom = small_domain(*shape, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    w = _w(dom, 5)
    psi = field(dom, 6, -1.0, 3.0)
    O.fill_along_z(g, psi, 0.4, 0.2)
    with H.Context(dom) as c:
        dw, dpsi = c.array(w), c.array(psi)
        ex, lb, ub = c.mpdata_vert3d(1.0, dw, dpsi, 0.4, 0.2, kord, monotone, neumann)
        r, olb, oub = O.mpdata_vert3d(g, 1.0, w, psi, 0.4, 0.2, kord, monotone, neumann)
        assert ex == r == 0
        assert (lb, ub) == (olb, oub)
        np.testing.assert_array_equal(dpsi.get(), psi)


@pytest.mark.parametrize("nq", [1, 3])
@pytest.mark.parametrize("kord,monotone", [(1, False), (2, True), (3, False)])
def test_mpdata_vec_vert_bitwise(nq, kord, monotone):
    dom = small_domain(10, 7, 12, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    w = _w(dom, 7)
    psi = np.stack([field(dom, 8 + n, 0.0, 1.0 + n) for n in range(nq)])
    lbc = [0.1 * n for n in range(nq)]
    ubc = [0.05 * n for n in range(nq)]
    with H.Context(dom) as c:
        dw, dpsi = c.array(w), c.array(psi)
        ex, lb, ub = c.mpdata_vec_vert3d(0.9, dw, dpsi, nq, lbc, ubc, kord, monotone, 1)
        r, olb, oub = O.mpdata_vec_vert3d(g, 0.9, w, psi, lbc, ubc, kord, monotone, 1)
        assert ex == r == 0
        np.testing.assert_array_equal(lb, olb)
        np.testing.assert_array_equal(ub, oub)
        np.testing.assert_array_equal(dpsi.get(), psi)


def test_mpdata_vert_zero_sum_exit():
    dom = small_domain(8, 6, 5)
    g = O.Grid(dom)
    w = _w(dom, 9)
    psi = np.full(dom.shape, 0.75)
    with H.Context(dom) as c:
        dw, dpsi = c.array(w), c.array(psi)
        ex, lb, ub = c.mpdata_vert3d(1.0, dw, dpsi, 0.75, 0.75)
        r, olb, oub = O.mpdata_vert3d(g, 1.0, w, psi, 0.75, 0.75)
        assert ex == r == 1 and (lb, ub) == (olb, oub)
        np.testing.assert_array_equal(dpsi.get(), psi)


def test_pamm_sedimentation_substeps_bitwise():
    """The call of pamm/microphysics.f90:1522-1530: a category of three
    hydrometeor fields falling with one speed field, kord 1, no limiter,
    neumann 2, sub-stepped n times with dt/n."""
    dom = small_domain(20, 16, 40, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    vq = _w(dom, 11, courant=0.8, dt=2.0, falling=True)
    O.fill_along_z(g, vq, 0.0, 0.0)
    psi = np.stack([field(dom, 12 + n, 0.0, 1e-3 * (n + 1)) for n in range(3)])
    lbc, ubc = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
    nsub, dt = 4, 2.0
    with H.Context(dom) as c:
        dw, dpsi = c.array(vq), c.array(psi)
        for _ in range(nsub):
            ex, lb, ub = c.mpdata_vec_vert3d(dt / nsub, dw, dpsi, 3, lbc, ubc, 1, False, 2)
            r, olb, oub = O.mpdata_vec_vert3d(g, dt / nsub, vq, psi, lbc, ubc, 1, False, 2)
            assert ex == r
            np.testing.assert_array_equal(lb, olb)
            lbc, ubc = list(lb), list(ub)
        np.testing.assert_array_equal(dpsi.get(), psi)


@pytest.mark.parametrize("kord,monotone", [(2, True), (3, False)])
def test_mpdata3d_add_bitwise(kord, monotone):
    dom = small_domain(9, 8, 10, stretched=True, varying_theta=True)
    g = O.Grid(dom)
    u, v, w = (field(dom, 20 + n, -1, 1) * 0.3 * dom.dx[1] for n in range(3))
    for a in (u, v, w):
        O.fill_full(g, a)
    prof = 280.0 + 0.003 * np.arange(dom.kp + 2) * dom.dz[1]
    psi = field(dom, 24, -0.5, 0.5)
    O.fill_full(g, psi, 0.0, 0.0)
    with H.Context(dom) as c:
        du, dv, dw, dpsi = c.array(u), c.array(v), c.array(w), c.array(psi)
        ex, lb, ub = c.mpdata3d_add(1.0, du, dv, dw, dpsi, prof, 280.0, 281.0, kord, monotone)
        r, olb, oub = O.mpdata3d_add(g, 1.0, u, v, w, psi, prof, 280.0, 281.0, kord, monotone)
        assert ex == r and (lb, ub) == (olb, oub)
        np.testing.assert_array_equal(dpsi.get(), psi)


def test_full_size_sedimentation_loses_the_floor_flux():
    """256x256x128, a rain field falling at up to 8 m/s, one PAMM-style
    step (kord 1, neumann 2): the density-weighted column mass changes by
    exactly the flux through the floor face, dt * w(1) * rhoa(0) * (psi(1)-m),
    m the positivity shift (here 0), to rounding."""
    dom = W.make_domain(256, 256, 128, 50.0, 50.0, 50.0)
    dt = 2.0
    rng = np.random.default_rng(77)
    psi = np.zeros(dom.shape)
    psi[1:-1, 1:-1, 1:-1] = 1e-3 * rng.random((256, 256, 128))
    vq = np.zeros(dom.shape)
    vq[1:-1, 1:-1, 0:-1] = -8.0 * rng.random((256, 256, 129))
    vq[:, :, -2] = 0.0                                  # nothing through the lid
    wgt = (dom.rhoan[1:-1] * dom.dzn[1:-1])[None, None, :]
    m0 = float((interior(psi) * wgt).sum())
    loss = dt * float((vq[1:-1, 1:-1, 1] * dom.rhoa[0] * psi[1:-1, 1:-1, 1]).sum())
    with H.Context(dom) as c:
        dw, dpsi = c.array(vq), c.array(psi[None])
        ex, lb, ub = c.mpdata_vec_vert3d(dt, dw, dpsi, 1, [0.0], [0.0], 1, False, 2)
        got = dpsi.get()[0]
    assert ex == 0
    m1 = float((interior(got) * wgt).sum())
    assert abs((m1 - m0) - loss) <= 1e-11 * m0, (m1 - m0, loss)
    assert interior(got).min() >= 0.0
