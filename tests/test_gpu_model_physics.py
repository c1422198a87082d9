"""The resident model with the physics options of sources on (subgrid
viscosity with the surface layer, damping layer, forcing, subsidence):
bit for bit against the oracle's model_driver with the reference's
sequential sums, a mixed-precision run near the FP64 one, and an ensemble
of such members evolving as when stepped alone."""
import numpy as np
import pytest

import oracle as O
from dcm_amd import hip as H
from dcm_amd import workloads as W
from helpers import interior, rel_l2
from test_gpu_parity import _compare_models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    O.build()


def sheared_bubble(ip, jp, kp, steps=3, ntracer=2, seed=71, damping=True, forcing=True,
                   subsidence=True):
    """A moist bubble in a sheared wind with every option of sources on (the
    reference's namelist.in sets viscous and damping)."""
    return W.sheared_moist_bubble(ip, jp, kp, steps=steps, ntracer=ntracer, seed=seed,
                                  damping=damping, forcing=forcing, subsidence=subsidence)


@pytest.mark.parametrize("shape", [(16, 12, 20), (12, 16, 9)])
def test_model_with_physics_matches_oracle(shape):
    case = sheared_bubble(*shape)
    _compare_models(case, 3, seq=True)


def test_model_viscous_only_matches_oracle():
    case = sheared_bubble(10, 10, 14, damping=False, forcing=False, subsidence=False)
    _compare_models(case, 3, seq=True)


def test_mixed_precision_with_physics_tracks_fp64():
    case = sheared_bubble(32, 32, 24, steps=4)
    out = {}
    for prec in (H.FP64, H.MIXED):
        with H.Context(case.dom) as c:
            m = H.Model(c, case, prec)
            try:
                m.set_check_residual(True)
                st = m.step(4)
                assert st.max_true_resid < 1e-10, (prec, st.max_true_resid)
                out[prec] = {"th": m.get(H.F_TH), "u": m.get(H.F_U), "w": m.get(H.F_W)}
            finally:
                m.close()
    for k in out[H.FP64]:
        err = rel_l2(interior(out[H.FP64][k]), interior(out[H.MIXED][k]))
        assert err < 1e-5, (k, err)


def test_ensemble_members_with_physics_evolve_as_alone():
    cases = [sheared_bubble(12, 10, 14, steps=5, seed=80),
             sheared_bubble(10, 12, 9, steps=5, seed=81, forcing=False)]
    alone = []
    for case in cases:
        with H.Context(case.dom) as c:
            m = H.Model(c, case, H.MIXED)
            try:
                m.step(5)
                alone.append(m.get(H.F_TH))
            finally:
                m.close()
    ctxs = [H.Context(cs.dom) for cs in cases]
    models = [H.Model(x, cs, H.MIXED) for x, cs in zip(ctxs, cases)]
    try:
        with H.Ensemble(models) as e:
            e.step(5)
        for q, m in enumerate(models):
            np.testing.assert_array_equal(m.get(H.F_TH), alone[q], err_msg=f"member {q}")
    finally:
        for m in models:
            m.close()
        for x in ctxs:
            x.close()
