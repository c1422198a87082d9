"""Host logic: the seeded synthetic workloads are deterministic and have the
stated shapes, base states and initial fields."""
import numpy as np
import pytest

from dcm_amd import workloads as W


def test_rng_is_counter_based_and_uniform():
    a = W.uniform(7, 3, np.arange(10000, dtype=np.uint64))
    b = W.uniform(7, 3, np.arange(10000, dtype=np.uint64))
    np.testing.assert_array_equal(a, b)
    assert 0.0 <= a.min() and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.02
    c = W.uniform(8, 3, np.arange(10000, dtype=np.uint64))
    assert not np.array_equal(a, c)


def test_grid_and_hydrostatic_base_state():
    d = W.make_domain(8, 6, 10, 50.0, 50.0, 50.0)
    assert d.shape == (10, 8, 12)
    np.testing.assert_allclose(d.z, 50.0 * np.arange(12))
    np.testing.assert_allclose(d.zn, d.z - 25.0)
    np.testing.assert_allclose(d.x[1] - d.x[0], 50.0)
    # isentropic: pi = 1 - g z / (cp theta); rho = p / (R theta pi)
    pi = 1.0 - W.GRAV * d.z / (W.CP * 300.0)
    rho = 1e5 * pi ** (W.CP / W.RA) / (W.RA * 300.0 * pi)
    np.testing.assert_allclose(d.rhoa, rho, rtol=1e-10)
    assert np.all(np.diff(d.rhoan) < 0)


@pytest.mark.parametrize("make", [
    lambda: W.warm_bubble(16, 16, 16, steps=1),
    lambda: W.moist_bubble(16, 12, 10, steps=1, ntracer=2),
    lambda: W.kelvin_helmholtz(24, 4, 16, steps=1),
    lambda: W.large_domain(16, 16, 8, steps=1, ntracer=1),
])
def test_cases_are_deterministic(make):
    a, b = make(), make()
    for name in ("u", "v", "w", "th"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name))
    if a.q is not None:
        np.testing.assert_array_equal(a.q, b.q)
        assert a.q.shape[0] == a.nq == len(a.c_s)


def test_warm_bubble_amplitude_and_noise():
    c = W.warm_bubble(32, 32, 32, amp=2.0, radius=500.0, zc=500.0, noise=0.01)
    th = c.th[1:-1, 1:-1, 1:-1]
    assert 1.9 < th.max() <= 2.01
    assert th.min() >= -0.01
    assert np.all(c.th[0] == 0.0)  # halos left for the fills


def test_ensemble_members_are_uneven_and_seeded():
    ms = W.ensemble_members(16, seed=404, steps=1)
    shapes = {(m.dom.ip, m.dom.jp, m.dom.kp) for m in ms}
    assert len(shapes) > 4
    for m in ms:
        assert m.dom.ip in (32, 48, 64, 96, 128) and m.dom.kp in (32, 64, 96)
