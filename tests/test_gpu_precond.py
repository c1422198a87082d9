"""The Chebyshev relaxation of the mixed inner solve's Jacobi preconditioner
(dcm_set_precond_chebyshev): the FP64 path, the reference's algorithm, is
bit for bit the same whatever the setting; the mixed path reaches the same
FP64 true residual with the reference's plain sweeps and with the Chebyshev
factors, agrees within the solver tolerance, and needs fewer iterations
with them."""
import numpy as np
import pytest

from dcm_amd import hip as H
from dcm_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ((H.F_U, "u"), (H.F_V, "v"), (H.F_W, "w"), (H.F_TH, "th"), (H.F_P, "p"))


def _run(case, precision, a, divisions, steps):
    with H.Context(case.dom) as c:
        c.set_precond_chebyshev(a, divisions)
        m = H.Model(c, case, precision)
        try:
            m.set_check_residual(True)
            st = m.step(steps)
            return {n: m.get(f) for f, n in FIELDS}, st
        finally:
            m.close()


def test_fp64_path_ignores_the_setting():
    case = W.moist_bubble(24, 20, 32, delta=50.0, dt=0.5, steps=3, ntracer=1, seed=51)
    a, sa = _run(case, H.FP64, 0.0, 4, 3)
    b, sb = _run(case, H.FP64, 0.002, 12, 3)
    assert sa.iters_total == sb.iters_total
    for name in a:
        np.testing.assert_array_equal(a[name], b[name], err_msg=name)


@pytest.mark.parametrize("a,divisions", [(0.005, 4), (0.002, 12), (0.001, 16)])
def test_mixed_chebyshev_reaches_the_tolerance_in_fewer_iterations(a, divisions):
    case = W.moist_bubble(64, 64, 48, delta=50.0, dt=0.5, steps=3, ntracer=1, seed=52)
    plain, sp = _run(case, H.MIXED, 0.0, 4, 3)
    cheb, sc = _run(case, H.MIXED, a, divisions, 3)
    assert sp.max_true_resid < 1e-10 and sc.max_true_resid < 1e-10
    assert sc.iters_total < sp.iters_total, (sc.iters_total, sp.iters_total)
    for name in plain:
        ref, got = plain[name], cheb[name]
        if name == "p":
            ref, got = ref - ref.mean(), got - got.mean()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 1e-7, (name, err)


def test_bad_settings_are_errors():
    case = W.warm_bubble(8, 8, 8, steps=1)
    with H.Context(case.dom) as c:
        for a, d in ((-1.0, 4), (2.0, 4), (0.01, 2), (0.01, 17)):
            with pytest.raises(H.DcmError):
                c.set_precond_chebyshev(a, d)
