"""The multigrid V-cycle preconditioner of the mixed path's FP32 inner solve
(dcm_set_precond_multigrid).

It replaces the Jacobi sweeps of the reference's iteration
(dynamics.f90:390-430) by one V-cycle per application, so it is judged by
what the mixed path promises: every solve ends at the requested FP64 true
residual, the model stays within the FP32 inner tolerance of the FP64 path
and of the mixed path with the reference's preconditioner, and the inner
solve needs a few iterations where the Jacobi sweeps need tens."""
import numpy as np
import pytest

from dcm_amd import hip as H
from dcm_amd import workloads as W
from helpers import rel_l2, small_domain

pytestmark = pytest.mark.gpu


def _rhs(dom, seed):
    """A compatible right-hand side (zero interior mean) in the padded layout."""
    r = np.random.default_rng(seed)
    b = np.zeros(dom.shape)
    b[1:-1, 1:-1, 1:-1] = r.standard_normal((dom.ip, dom.jp, dom.kp))
    b[1:-1, 1:-1, 1:-1] -= b[1:-1, 1:-1, 1:-1].mean()
    return b


@pytest.mark.parametrize("shape", [(32, 32, 16), (64, 64, 32), (16, 16, 64)])
def test_mixed_solve_converges_in_few_iterations(shape):
    dom = small_domain(*shape)
    b = _rhs(dom, 5)
    sw = np.zeros(dom.shape)
    iters = {}
    for mg in (False, True):
        with H.Context(dom) as c:
            if mg:
                c.set_precond_multigrid(True)
            x = c.array(np.zeros(dom.shape))
            it, outer, err = c.bicgstab_mixed(c.array(sw), x, c.array(b), tol=1e-10)
            assert err < 1e-10, (mg, it, outer, err)
            iters[mg] = it
    assert iters[True] * 3 <= iters[False], iters


def test_multigrid_model_tracks_fp64_and_the_jacobi_path():
    case = W.moist_bubble(32, 32, 16, delta=50.0, dt=0.5, steps=8, ntracer=1, seed=12)
    out, iters = {}, {}
    for name, prec, mg in (("fp64", H.FP64, False), ("jacobi", H.MIXED, False),
                           ("mg", H.MIXED, True)):
        with H.Context(case.dom) as c:
            if mg:
                c.set_precond_multigrid(True)
            m = H.Model(c, case, prec)
            try:
                m.set_check_residual(True)
                st = m.step(8)
                assert st.max_true_resid < 1e-10, (name, st.max_true_resid)
                assert st.fallbacks == 0
                iters[name] = st.iters_total
                out[name] = {f: m.get(f) for f in (H.F_U, H.F_V, H.F_W, H.F_TH)}
                out[name]["q0"] = m.get(H.F_Q, 0)
                out[name]["p"] = m.get(H.F_P)
            finally:
                m.close()
    assert iters["mg"] * 2 <= iters["jacobi"], iters
    for f in out["fp64"]:
        center = f == "p"
        a, b, g = out["fp64"][f], out["jacobi"][f], out["mg"][f]
        if center:
            a, b, g = a - a.mean(), b - b.mean(), g - g.mean()
        assert rel_l2(a, g) < 1e-5, f
        assert rel_l2(b, g) < 1e-5, f


def test_multigrid_is_refused_where_the_grid_does_not_coarsen():
    dom = small_domain(12, 10, 8)   # 960 cells: no coarser level
    with H.Context(dom) as c:
        with pytest.raises(H.DcmError):
            c.set_precond_multigrid(True)
    dom = small_domain(15, 9, 11)   # 1485 cells, every dimension odd
    with H.Context(dom) as c:
        with pytest.raises(H.DcmError):
            c.set_precond_multigrid(True)


@pytest.mark.parametrize("shape", [(40, 6, 40), (64, 2, 32), (128, 4, 64), (96, 8, 16)])
def test_semi_coarsened_grids_converge_in_few_iterations(shape):
    """A dimension that turns odd keeps its size on the coarser levels
    (40x6x40 -> 20x3x20 -> 10x3x10 -> 5x3x5, odd everywhere; 64x2x32 ->
    32x1x16 -> 16x1x8 -> ...; 96x8x16 -> 48x4x8 -> 24x2x4 -> 12x1x2) and
    the transfers copy along it: the V-cycle still cuts the inner
    iterations at least threefold against the Jacobi sweeps, and the solve
    ends at the FP64 true residual asked for."""
    dom = small_domain(*shape)
    b = _rhs(dom, 7)
    sw = np.zeros(dom.shape)
    iters, sols = {}, {}
    for mg in (False, True):
        with H.Context(dom) as c:
            if mg:
                c.set_precond_multigrid(True)
            x = c.array(np.zeros(dom.shape))
            it, outer, err = c.bicgstab_mixed(c.array(sw), x, c.array(b), tol=1e-10)
            assert err < 1e-10, (mg, it, outer, err)
            iters[mg] = it
            sols[mg] = x.get()[1:-1, 1:-1, 1:-1]
    assert iters[True] * 3 <= iters[False], iters
    a, g = sols[False], sols[True]
    assert rel_l2(a - a.mean(), g - g.mean()) < 1e-8


def test_multigrid_switches_off_and_on():
    dom = small_domain(32, 32, 16)
    b = _rhs(dom, 9)
    sw = np.zeros(dom.shape)
    with H.Context(dom) as c:
        c.set_precond_multigrid(True)
        c.set_precond_multigrid(False)
        x = c.array(np.zeros(dom.shape))
        off = c.bicgstab_mixed(c.array(sw), x, c.array(b), tol=1e-10)
        c.set_precond_multigrid(True)
        x2 = c.array(np.zeros(dom.shape))
        on = c.bicgstab_mixed(c.array(sw), x2, c.array(b), tol=1e-10)
        assert off[2] < 1e-10 and on[2] < 1e-10
        assert on[0] < off[0]
        a, g = x.get()[1:-1, 1:-1, 1:-1], x2.get()[1:-1, 1:-1, 1:-1]
        assert rel_l2(a - a.mean(), g - g.mean()) < 1e-8


def _q_mass(dom, a):
    w = (dom.rhoan[1:-1] * dom.dz[1:-1])[None, None, :]
    return float(np.sum(a[1:-1, 1:-1, 1:-1] * w))


@pytest.mark.parametrize("name", ["warm_bubble_64", "moist_bubble_256x256x128",
                                  "kelvin_helmholtz_512x4x256", "large_domain_512x512x256"])
def test_multigrid_at_full_size(name):
    """The workloads' full sizes (the 512x4x256 Kelvin-Helmholtz grid stops
    halving in y at one cell and is semi-coarsened below that): every solve
    at FP64 true residual below 1e-10, every q conserved to FP32 flux
    rounding, within 1e-5 of the FP64 path, and a few inner iterations per
    solve."""
    case = W.CONFIGS[name](steps=2)
    dom = case.dom
    print(f"multigrid at full size {name}: {dom.ip}x{dom.jp}x{dom.kp}", flush=True)
    out = {}
    for prec in (H.FP64, H.MIXED):
        with H.Context(dom) as c:
            if prec == H.MIXED:
                c.set_precond_multigrid(True)
            m = H.Model(c, case, prec)
            try:
                m.set_check_residual(True)
                st = m.step(2)
                print(f"  precision {prec}: iterations {st.iters_total}, outer {st.outer_total}, "
                      f"true residual {st.max_true_resid:.3g}", flush=True)
                assert st.max_true_resid < 1e-10 and st.fallbacks == 0
                if prec == H.MIXED:
                    assert st.iters_total <= 6 * max(st.outer_total, 1), st.iters_total
                out[prec] = {"th": m.get(H.F_TH), "u": m.get(H.F_U), "w": m.get(H.F_W)}
                for n in range(case.nq):
                    q = m.get(H.F_Q, n)
                    out[prec][f"q{n}"] = q
                    ref = _q_mass(dom, case.q[n])
                    scale = max(abs(ref), _q_mass(dom, np.abs(case.q[n])))
                    assert abs(_q_mass(dom, q) - ref) <= 1e-6 * scale, (prec, n)
            finally:
                m.close()
    for key in out[H.FP64]:
        a, b = out[H.FP64][key][1:-1, 1:-1, 1:-1], out[H.MIXED][key][1:-1, 1:-1, 1:-1]
        assert rel_l2(a, b) < 1e-5, key


def test_ensemble_with_multigrid_members_evolves_them_as_alone():
    """Every member with the multigrid preconditioner: the ensemble solves
    each with its own V-cycle, so each evolves bit for bit as when stepped
    alone (graphs capture the stages around the solve)."""
    cases = [W.warm_bubble(16, 12, 10, steps=7, amp=2.0, seed=11, name="a"),
             W.moist_bubble(12, 16, 8, delta=50.0, dt=0.5, steps=7, ntracer=1, seed=12),
             W.warm_bubble(20, 8, 40, steps=7, amp=1.0, radius=300.0, seed=13, name="c"),
             W.warm_bubble(16, 16, 16, steps=7, amp=3.0, seed=14, name="d")]

    def snap(case, m):
        out = {f: m.get(f) for f in (H.F_U, H.F_V, H.F_W, H.F_TH, H.F_P)}
        for n in range(case.nq):
            out[f"q{n}"] = m.get(H.F_Q, n)
        return out

    alone = []
    for case in cases:
        with H.Context(case.dom) as c:
            c.set_precond_multigrid(True)
            m = H.Model(c, case, H.MIXED)
            try:
                st = m.step(7)
                alone.append((snap(case, m), st.iters_total))
            finally:
                m.close()
    ctxs = [H.Context(cs.dom) for cs in cases]
    models = []
    try:
        for c, cs in zip(ctxs, cases):
            c.set_precond_multigrid(True)
            models.append(H.Model(c, cs, H.MIXED))
        with H.Ensemble(models) as e:
            stats = e.step(7)
        for q, case in enumerate(cases):
            exp, iters = alone[q]
            assert stats[q].iters_total == iters, (q, stats[q].iters_total, iters)
            got = snap(case, models[q])
            for name in exp:
                np.testing.assert_array_equal(got[name], exp[name], err_msg=f"member {q} {name}")
    finally:
        for m in models:
            m.close()
        for c in ctxs:
            c.close()


def test_multigrid_iteration_graph_replays_match_eager_launches():
    """The single-system multigrid inner iteration is replayed from a HIP
    graph and the advection runs on its own stream beside the solve; with
    per-kernel timing on (profile) the same kernels are launched one by one
    on one stream.  Both give the same bits after every step."""
    case = W.moist_bubble(32, 32, 16, delta=50.0, dt=0.5, steps=6, ntracer=1, seed=21)
    ctxs = [H.Context(case.dom), H.Context(case.dom)]
    try:
        models = []
        for n, c in enumerate(ctxs):
            c.set_precond_multigrid(True)
            m = H.Model(c, case, H.MIXED)
            if n == 1:
                m.profile(True)
            models.append(m)
        for step in range(6):
            stats = [m.step(1) for m in models]
            assert stats[0].iters_total == stats[1].iters_total, step
            for f in (H.F_U, H.F_W, H.F_TH, H.F_P):
                np.testing.assert_array_equal(models[0].get(f), models[1].get(f),
                                              err_msg=f"field {f} after step {step}")
        for m in models:
            m.close()
    finally:
        for c in ctxs:
            c.close()


def test_fp64_solve_with_the_v_cycle_matches_the_reference_sweeps():
    """The FP64 BiCGSTAB with the V-cycle in place of the reference's Jacobi
    sweeps: every solve reaches the FP64 true residual bar in a fraction of
    the iterations, and the model stays within the solver tolerance of the
    reference's FP64 path."""
    case = W.moist_bubble(32, 32, 16, delta=50.0, dt=0.5, steps=6, ntracer=1, seed=31)
    out, iters = {}, {}
    for mg in (False, True):
        with H.Context(case.dom) as c:
            if mg:
                c.set_precond_multigrid(True)
            m = H.Model(c, case, H.FP64)
            try:
                m.set_check_residual(True)
                st = m.step(6)
                assert st.max_true_resid < 1e-10, (mg, st.max_true_resid)
                iters[mg] = st.iters_total
                out[mg] = {f: m.get(f) for f in (H.F_U, H.F_W, H.F_TH, H.F_P)}
            finally:
                m.close()
    assert iters[True] * 3 <= iters[False], iters
    for f in out[False]:
        a, b = out[False][f], out[True][f]
        if f == H.F_P:
            a, b = a - a.mean(), b - b.mean()
        assert rel_l2(a, b) < 1e-7, f


@pytest.mark.parametrize("precision", [H.MIXED, H.FP64])
@pytest.mark.parametrize("shape,steps", [((64, 64, 32), 4), ((256, 128, 64), 2),
                                         ((128, 4, 64), 3), ((256, 4, 64), 2)])
def test_one_block_tail_matches_per_pass_launches(precision, shape, steps):
    """The levels that fit one workgroup's LDS run as one single-block kernel,
    the levels of at most 2^17 cells as one kernel per leg, and the
    finest-level transfers as marching kernels; with
    dcm_set_precond_multigrid(ctx, 2) every pass is its own one-thread-per-
    cell launch.  64x64x32 has every coarse level fused or in the tail;
    256x128x64 has a first coarse level of 2^18 cells that keeps its
    per-pass kernels; 128x4x64 reaches one cell in y at level 2, and its
    tail transfers are semi-coarsened; 256x4x64 has a semi-coarsened level
    of 1024 cells that the fused legs cannot take, so its tail is a
    1024-thread block.  Same arithmetic, so the same bits after every
    step."""
    case = W.moist_bubble(*shape, delta=50.0, dt=0.5, steps=steps, ntracer=1, seed=41)
    snaps = {}
    for mode in (1, 2):
        with H.Context(case.dom) as c:
            c.set_precond_multigrid(mode)
            m = H.Model(c, case, precision)
            try:
                out = []
                for step in range(steps):
                    st = m.step(1)
                    out.append((st.iters_total, {f: m.get(f) for f in
                                                 (H.F_U, H.F_W, H.F_TH, H.F_P)}))
                snaps[mode] = out
            finally:
                m.close()
    for step, ((ia, a), (ib, b)) in enumerate(zip(snaps[1], snaps[2])):
        assert ia == ib, (step, ia, ib)
        for f in a:
            np.testing.assert_array_equal(a[f], b[f], err_msg=f"field {f} after step {step}")
