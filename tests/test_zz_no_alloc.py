"""dcm_inn_prod_noalloc, the allocation-free variant of dcm_inn_prod: the
same kernels over a fixed device array instead of the context's partial-sum
buffer, so the result is the same bits.  Two calls in a row, the second on
a smaller grid (stale partials of the first must not leak into it), both
summation orders, and the size limit."""
import numpy as np
import pytest

from dcm_amd import hip as H
from dcm_amd import workloads as W

pytestmark = pytest.mark.gpu


def _pair(ctx, dom, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(dom.shape)
    y = rng.standard_normal(dom.shape)
    dx = H.DeviceArray(ctx, dom.shape).set(x)
    dy = H.DeviceArray(ctx, dom.shape).set(y)
    return x, y, dx, dy


def _check(ctx, dom, seed):
    x, y, dx, dy = _pair(ctx, dom, seed)
    try:
        got = ctx.inn_prod_noalloc(dx, dy)
        ref = ctx.inn_prod(dx, dy)
        print(f"  {dom.ip}x{dom.jp}x{dom.kp}: noalloc {got!r}, inn_prod {ref!r}", flush=True)
        assert got == ref  # same kernels, same order: the same bits
        exact = float(np.sum(x[1:-1, 1:-1, 1:-1] * y[1:-1, 1:-1, 1:-1]))
        assert abs(got - exact) <= 1e-12 * np.sum(np.abs(x * y))
        # a second call on the same inputs gives the same bits again
        assert ctx.inn_prod_noalloc(dx, dy) == got
    finally:
        dx.free()
        dy.free()


@pytest.mark.parametrize("order", [H.SUM_TREE, H.SUM_SEQUENTIAL])
def test_matches_inn_prod_on_two_calls_the_second_smaller(order):
    big = W.make_domain(96, 80, 72, 50.0, 50.0, 40.0)
    small = W.make_domain(24, 20, 16, 50.0, 50.0, 40.0)
    for dom, seed in ((big, 11), (small, 12)):
        with H.Context(dom) as c:
            c.set_sum_order(order)
            _check(c, dom, seed)


def test_rejects_grids_beyond_the_fixed_scratch():
    """4096 x 260 x 64 needs 1 x 65 x 4096 = 266240 partial sums, more than
    the 2^18 the array holds: refused before any launch."""
    dom = W.make_domain(4096, 260, 64, 50.0, 50.0, 40.0)
    with H.Context(dom) as c:
        dx = H.DeviceArray(c, dom.shape)
        try:
            with pytest.raises(H.DcmError, match="fixed scratch"):
                c.inn_prod_noalloc(dx, dx)
        finally:
            dx.free()
