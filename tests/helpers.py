"""Test helpers: seeded random fields, masks of defined halo cells, norms."""
import numpy as np

from dcm_amd import workloads as W


def rng(seed):
    return np.random.default_rng(seed)


def small_domain(ip, jp, kp, delta=50.0, stretched=False, varying_theta=False):
    """A small domain; optionally non-uniform metrics to exercise every index."""
    theta_fn = None
    if varying_theta:
        def theta_fn(z):
            z = np.asarray(z, dtype=np.float64)
            return 300.0 + 0.004 * z + 0.5 * np.tanh((z - kp * delta / 2) / 60.0)
    d = W.make_domain(ip, jp, kp, delta, delta * 0.9, delta * 1.1, theta_fn)
    if stretched:
        r = rng(7)
        for name, n in (("dx", ip), ("dy", jp), ("dz", kp)):
            a = delta * (0.8 + 0.4 * r.random(n + 2))
            setattr(d, name, a)
            setattr(d, name + "n", delta * (0.8 + 0.4 * r.random(n + 2)))
    return d


def field(dom, seed, lo=-1.0, hi=1.0):
    return lo + (hi - lo) * rng(seed).random(dom.shape)


def interior(a):
    return a[..., 1:-1, 1:-1, 1:-1]


def rel_l2(a, b):
    na = np.linalg.norm(a)
    return np.linalg.norm(a - b) / (na if na > 0 else 1.0)


def lateral_edge_mask(dom):
    """Cells whose value the reference leaves undefined in the pressure:
    lateral edges (both j and i in the halo) at every level."""
    m = np.zeros(dom.shape, dtype=bool)
    m[0, 0, :] = m[0, -1, :] = m[-1, 0, :] = m[-1, -1, :] = True
    return m
