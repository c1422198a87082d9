"""Ensembles: independent members of uneven size stepped together, their
pressure solves batched into one BiCGSTAB loop and the stages around the
solve replayed from captured HIP graphs.  A member must evolve exactly as it
does on its own (the batched kernels run each member's arithmetic, partial
sums included, unchanged), and with the reference's sequential sums exactly
as the oracle."""
import numpy as np
import pytest

import oracle as O
from dcm_amd import hip as H
from dcm_amd import workloads as W
from helpers import lateral_edge_mask

pytestmark = pytest.mark.gpu

FIELDS = ((H.F_U, "u"), (H.F_V, "v"), (H.F_W, "w"), (H.F_TH, "th"), (H.F_P, "p"))


def _members(steps=7):
    """Uneven members: different shapes, depths, bubbles and tracer counts."""
    return [W.warm_bubble(16, 12, 10, steps=steps, amp=2.0, seed=11, name="a"),
            W.moist_bubble(12, 16, 8, delta=50.0, dt=0.5, steps=steps, ntracer=1, seed=12),
            W.warm_bubble(20, 8, 40, steps=steps, amp=1.0, radius=300.0, seed=13, name="c"),
            W.warm_bubble(8, 8, 16, steps=steps, amp=3.0, seed=14, name="d")]


def _open(cases, precision, seq=False):
    ctxs, models = [], []
    for case in cases:
        c = H.Context(case.dom)
        if seq:
            c.set_sum_order(H.SUM_SEQUENTIAL)
        ctxs.append(c)
        models.append(H.Model(c, case, precision))
    return ctxs, models


def _close(ctxs, models):
    for m in models:
        m.close()
    for c in ctxs:
        c.close()


def _snapshot(case, m):
    out = {name: m.get(fid) for fid, name in FIELDS}
    for n in range(case.nq):
        out[f"q{n}"] = m.get(H.F_Q, n)
    return out


@pytest.mark.parametrize("precision", [H.FP64, H.MIXED])
def test_ensemble_members_evolve_as_alone(precision):
    """Nine steps: the first three run each time-level phase eagerly (which
    allocates every scratch buffer), the next three capture the graphs of
    the three phases, the last three replay them."""
    cases = _members(9)
    steps = 9
    # alone
    alone = []
    for case in cases:
        ctxs, models = _open([case], precision)
        try:
            st = models[0].step(steps)
            alone.append((_snapshot(case, models[0]), st.iters_total))
        finally:
            _close(ctxs, models)
    # together
    ctxs, models = _open(cases, precision)
    try:
        with H.Ensemble(models) as e:
            stats = e.step(steps)
        for q, case in enumerate(cases):
            got = _snapshot(case, models[q])
            exp, iters = alone[q]
            assert stats[q].iters_total == iters, (q, stats[q].iters_total, iters)
            for name in exp:
                np.testing.assert_array_equal(got[name], exp[name], err_msg=f"member {q} {name}")
    finally:
        _close(ctxs, models)


def test_graph_replays_match_eager_steps():
    """Two copies of an ensemble, one replaying captured graphs and one
    launching every kernel, compared after every step (eager solves and host
    synchronisations run between the replays)."""
    cases = _members()
    ctx_g, mod_g = _open(cases, H.MIXED)
    ctx_e, mod_e = _open(cases, H.MIXED)
    try:
        with H.Ensemble(mod_g) as eg, H.Ensemble(mod_e) as ee:
            ee.set_graphs(False)
            for s in range(7):
                sg = eg.step(1)
                se = ee.step(1)
                for q, case in enumerate(cases):
                    assert sg[q].iters_total == se[q].iters_total, (s, q)
                    a, b = _snapshot(case, mod_g[q]), _snapshot(case, mod_e[q])
                    for name in a:
                        np.testing.assert_array_equal(a[name], b[name],
                                                      err_msg=f"step {s} member {q} {name}")
    finally:
        _close(ctx_g, mod_g)
        _close(ctx_e, mod_e)


def test_ensemble_with_sequential_sums_matches_oracle():
    cases = _members(4)
    ctxs, models = _open(cases, H.FP64, seq=True)
    oracles = [O.Model(case) for case in cases]
    try:
        with H.Ensemble(models) as e:
            planes = e.autotune()
            assert planes in (4, 8, 16)
            for s in range(4):
                stats = e.step(1)
                for q, om in enumerate(oracles):
                    om.step(1)
                    assert stats[q].iters_total == om.iters[-1], (s, q)
        for q, (case, om) in enumerate(zip(cases, oracles)):
            (u, v, w), _ = om.levels()
            got = _snapshot(case, models[q])
            for name, ref in (("u", u), ("v", v), ("w", w), ("th", om.th)):
                np.testing.assert_array_equal(got[name], ref, err_msg=f"member {q} {name}")
            mask = ~lateral_edge_mask(case.dom)  # p's edges are stale in the reference
            np.testing.assert_array_equal(got["p"][mask], om.p[mask])
    finally:
        _close(ctxs, models)


def test_autotune_leaves_the_state_unchanged():
    cases = _members(1)[:2]
    ctxs, models = _open(cases, H.FP64)
    try:
        before = [_snapshot(c, m) for c, m in zip(cases, models)]
        with H.Ensemble(models) as e:
            e.autotune()
        after = [_snapshot(c, m) for c, m in zip(cases, models)]
        for b, a in zip(before, after):
            for name in b:
                np.testing.assert_array_equal(a[name], b[name])
    finally:
        _close(ctxs, models)


def test_uneven_64_member_ensemble_converges():
    """The 64-member uneven ensemble (nx, ny in 32..128, nz in 32..96): three
    mixed-precision steps (the last two through captured graphs), every
    member's FP64 true residual below 1e-10."""
    cases = W.ensemble_members(64, seed=404, steps=3)
    print(f"ensemble of {len(cases)}: {sum(c.dom.cells for c in cases)} cells", flush=True)
    ctxs, models = _open(cases, H.MIXED)
    try:
        with H.Ensemble(models) as e:
            stats = e.step(3)
        assert all(s.max_true_resid < 1e-10 for s in stats), [s.max_true_resid for s in stats]
    finally:
        _close(ctxs, models)


def test_ensemble_physics_switched_on_after_capture():
    """Two copies of a two-member ensemble, one replaying graphs and one
    eager.  After six steps (graphs of every phase captured) member 0 gets
    the subgrid model, whose scratch fields do not exist yet: the replaying
    copy runs the new phases eagerly once before capturing them (no
    allocation or memset inside a capture), and both copies stay bit for
    bit equal."""
    cases = [W.warm_bubble(16, 12, 10, steps=12, amp=2.0, seed=11, name="a"),
             W.moist_bubble(12, 8, 16, delta=50.0, dt=0.5, steps=12, ntracer=1, seed=12)]
    ph = W.physics_profiles(cases[0].dom)
    sets = [_open(cases, H.MIXED) for _ in range(2)]
    try:
        (cg, mg), (ce, me) = sets
        with H.Ensemble(mg) as eg, H.Ensemble(me) as ee:
            ee.set_graphs(False)
            for s in range(12):
                if s == 6:
                    mg[0].set_physics(ph)
                    me[0].set_physics(ph)
                sg, se = eg.step(1), ee.step(1)
                for q, case in enumerate(cases):
                    assert sg[q].iters_total == se[q].iters_total, (s, q)
                    a, b = _snapshot(case, mg[q]), _snapshot(case, me[q])
                    for name in a:
                        np.testing.assert_array_equal(a[name], b[name],
                                                      err_msg=f"step {s} member {q} {name}")
    finally:
        for ctxs, models in sets:
            _close(ctxs, models)


def test_ensemble_refuses_a_multigrid_mix():
    """Mixed-precision members either all use the multigrid preconditioner
    or none: the batched solve has no V-cycle."""
    cases = [W.warm_bubble(16, 16, 16, steps=2, seed=11, name="a"),
             W.warm_bubble(16, 16, 16, steps=2, seed=12, name="b")]
    ctxs, models = _open(cases, H.MIXED)
    try:
        ctxs[0].set_precond_multigrid(True)
        with H.Ensemble(models) as e:
            with pytest.raises(H.DcmError):
                e.step(1)
    finally:
        _close(ctxs, models)
