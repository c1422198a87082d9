"""The oracle's host threads change its speed, not its results.

The CPU baseline of the benchmark runs the oracle's element-wise loops on
several OpenMP threads; every such loop writes each cell from values that
no other iteration writes, so the model state after a step must be
bit-identical to the serial run.  The dot products stay serial (the
reference's order) unless a parallel reduction is asked for explicitly.
"""
import numpy as np
import pytest

import oracle as O
from dcm_amd import workloads as W


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def _run(case, threads, steps=2):
    O.set_threads(threads)
    try:
        m = O.Model(case)
        m.step(steps)
        (u, v, w), _ = m.levels()
        return m.iters, {"u": u.copy(), "v": v.copy(), "w": w.copy(), "th": m.th.copy(),
                         "p": m.p.copy(), "q": m.q.copy()}
    finally:
        O.set_threads(1)


def test_model_step_is_bit_identical_on_four_threads():
    case = W.moist_bubble(24, 20, 16, steps=2)
    it1, a = _run(case, 1)
    it4, b = _run(case, 4)
    assert it1 == it4
    for name in a:
        np.testing.assert_array_equal(a[name], b[name], err_msg=name)


def test_parallel_dot_products_change_only_the_last_bits():
    """With the parallel reduction (timing only) the sums' order changes:
    the state stays within 1e-9 of the serial run after a step."""
    case = W.moist_bubble(24, 20, 16, steps=1)
    _, a = _run(case, 1, steps=1)
    O.set_sum_blocked(2)
    try:
        _, b = _run(case, 4, steps=1)
    finally:
        O.set_sum_blocked(0)
    for name in ("u", "w", "th"):
        d = np.abs(a[name] - b[name]).max()
        s = np.abs(a[name]).max()
        assert d <= 1e-9 * max(s, 1.0), (name, d, s)
