"""Seeded synthetic workloads for the dynamical-core hot path.

The reference builds its grid, hydrostatic base state and initial fields in
``initialisation.f90`` (grid: ``:429-446``; density: ``:539,573``; bubble
and noise: ``:677-680,720-748,810-833``) from namelists, with VODE for the
hydrostatic integral and the compiler's RNG for noise.  Those inputs are
replaced here by closed forms and a counter-based RNG so that the CPU oracle
and the GPU path receive bit-identical inputs on any machine:

* grid (``initialisation.f90:429-446``): uniform spacing, ``x(i)=i*dx-(ip-1)dx/2``,
  ``xn=x-dx/2`` (same for y), ``z(k)=k*dz``, ``zn=z-dz/2``, Fortran indices
  ``0..n+1`` stored at numpy index ``0..n+1``;
* base state: Exner function ``pi(z)=1-(g/cp)*int_0^z dz'/theta(z')``,
  ``p=1e5*pi^(cp/R)``, ``rho=p/(R*theta*pi)`` (the reference's
  ``rhoa=p/(ra*theta*(p/1e5)^(ra/cp))``), on w levels (``rhoa``) and scalar
  levels (``rhoan``);
* noise: SplitMix64 keyed by (seed, field, linear cell index), so values do
  not depend on launch shape or decomposition.

Fields are numpy arrays of shape ``(ip+2, jp+2, kp+2)`` in C order, which is
the reference's Fortran ``(0:kp+1, 0:jp+1, 0:ip+1)`` layout with k fastest.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

GRAV = 9.81
CP = 1005.0
RA = 287.0

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, field: int, counter: np.ndarray) -> np.ndarray:
    """Uniform [0,1) doubles keyed by (seed, field, counter)."""
    key = _splitmix64(np.uint64(seed))
    ctr = (np.uint64(field) << np.uint64(48)) | np.asarray(counter, dtype=np.uint64)
    r = _splitmix64(key ^ ctr)
    return (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass
class Domain:
    """Grid metrics and 1-D base state (arrays of length n+2)."""

    ip: int
    jp: int
    kp: int
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray
    dxn: np.ndarray
    dyn: np.ndarray
    dzn: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    xn: np.ndarray
    yn: np.ndarray
    zn: np.ndarray
    rhoa: np.ndarray
    rhoan: np.ndarray
    theta: np.ndarray
    thetan: np.ndarray

    @property
    def shape(self):
        return (self.ip + 2, self.jp + 2, self.kp + 2)

    @property
    def cells(self) -> int:
        return self.ip * self.jp * self.kp


def _exner(zs: np.ndarray, theta_fn, psurf_pi: float = 1.0) -> np.ndarray:
    """pi(z) = 1 - g/cp * int_0^z dz/theta, by composite Simpson on a fine grid."""
    out = np.empty_like(zs)
    for n, zz in enumerate(zs):
        m = 2 * max(8, int(abs(zz) / 5.0) + 1)
        t = np.linspace(0.0, zz, m + 1)
        f = 1.0 / theta_fn(t)
        h = (zz - 0.0) / m
        integral = h / 3.0 * (f[0] + f[-1] + 4.0 * f[1:-1:2].sum() + 2.0 * f[2:-1:2].sum())
        out[n] = psurf_pi - GRAV / CP * integral
    return out


def make_domain(ip: int, jp: int, kp: int, dx: float, dy: float, dz: float,
                theta_fn=None) -> Domain:
    """Uniform grid with a hydrostatic base state theta_fn(z) (default 300 K)."""
    if theta_fn is None:
        def theta_fn(z):
            return np.full_like(np.asarray(z, dtype=np.float64), 300.0)
    i = np.arange(ip + 2, dtype=np.float64)
    j = np.arange(jp + 2, dtype=np.float64)
    k = np.arange(kp + 2, dtype=np.float64)
    x = dx * i - (ip - 1) / 2.0 * dx
    y = dy * j - (jp - 1) / 2.0 * dy
    z = dz * k
    xn, yn, zn = x - 0.5 * dx, y - 0.5 * dy, z - 0.5 * dz
    theta = np.asarray(theta_fn(z), dtype=np.float64)
    thetan = np.asarray(theta_fn(zn), dtype=np.float64)
    pi_w = _exner(z, theta_fn)
    pi_s = _exner(zn, theta_fn)
    rhoa = 1e5 * pi_w ** (CP / RA) / (RA * theta * pi_w)
    rhoan = 1e5 * pi_s ** (CP / RA) / (RA * thetan * pi_s)
    full = lambda n, v: np.full(n + 2, float(v))
    return Domain(ip, jp, kp, full(ip, dx), full(jp, dy), full(kp, dz),
                  full(ip, dx), full(jp, dy), full(kp, dz),
                  x, y, z, xn, yn, zn, rhoa, rhoan, theta, thetan)


@dataclasses.dataclass
class Physics:
    """The options of the reference's sources beyond the inviscid branch
    (dynamics.f90:639-1000): Smagorinsky-Lilly subgrid viscosity with the
    Monin-Obukhov surface layer, a damping layer, large-scale forcing and
    subsidence.  1-D profiles have kp+2 entries."""

    viscous: bool = True
    damping_layer: bool = False
    forcing: bool = False
    divergence: bool = False
    forcing_tau: float = 3600.0
    z0: float = 1e-4
    z0th: float = 1e-4
    lamsq: np.ndarray = None
    dampfacn: np.ndarray = None
    dampfac: np.ndarray = None
    u_force: np.ndarray = None
    v_force: np.ndarray = None
    w_subs: np.ndarray = None


def physics_profiles(dom: Domain, u=None, v=None, viscous=True, cvis=0.23, z0=1e-4, z0th=1e-4,
                     damping_layer=False, damping_thickness=None, damping_tau=1000.0,
                     forcing=False, forcing_tau=3600.0, divergence=False, divergence_val=0.0,
                     divergence_hgt=1000.0) -> Physics:
    """Profiles as the reference's allocate_and_set builds them
    (initialisation.f90:449-452 mixing length, :466-491 damping layer above
    zn(kp) - damping_thickness, :838-853 forcing profile from column (1,1)
    and the subsidence profile).  dx, dy are the domain's uniform spacings."""
    dx, dy = float(dom.dx[1]), float(dom.dy[1])
    lamsq = 1.0 / (1.0 / (cvis * (dx + dy + dom.dz) / 3.0) ** 2 + 1.0 / (0.4 * (dom.z + z0)) ** 2)
    n = dom.kp + 2
    dampfacn, dampfac = np.zeros(n), np.zeros(n)
    if damping_layer:
        thick = damping_thickness if damping_thickness is not None else 0.25 * dom.zn[dom.kp]
        ztop = dom.zn[dom.kp]
        dampfacn[dom.zn > ztop - thick] = -1.0 / damping_tau
        dampfac[dom.z > ztop - thick] = -1.0 / damping_tau
    u_force = np.zeros(n) if u is None else np.array(u[1, 1, :], dtype=np.float64)
    v_force = np.zeros(n) if v is None else np.array(v[1, 1, :], dtype=np.float64)
    w_subs = np.where(dom.z <= divergence_hgt, -divergence_val * dom.z,
                      -divergence_val * divergence_hgt)
    return Physics(viscous, damping_layer, forcing, divergence, forcing_tau, z0, z0th, lamsq,
                   dampfacn, dampfac, u_force, v_force, w_subs)


@dataclasses.dataclass
class Case:
    """A runnable configuration: domain, run settings and initial fields."""

    name: str
    dom: Domain
    dt: float
    steps: int
    kord: int = 2
    monotone: bool = True
    ptol: float = 1e-10
    advection_scheme: int = 1
    nq: int = 0
    # 1-based inclusive category bounds into q (the reference's c_s/c_e)
    c_s: tuple = ()
    c_e: tuple = ()
    seed: int = 0
    u: np.ndarray = None
    v: np.ndarray = None
    w: np.ndarray = None
    th: np.ndarray = None
    q: np.ndarray = None  # (nq, ip+2, jp+2, kp+2)
    physics: Physics = None  # None: the inviscid branch of sources

    @property
    def moisture(self) -> bool:
        return self.nq > 0

    @property
    def ncat(self) -> int:
        return len(self.c_s)


def _cell_counter(dom: Domain) -> np.ndarray:
    ii, jj, kk = np.meshgrid(np.arange(dom.ip + 2), np.arange(dom.jp + 2),
                             np.arange(dom.kp + 2), indexing="ij")
    return ((ii * (dom.jp + 2) + jj) * (dom.kp + 2) + kk).astype(np.uint64)


def _interior(a: np.ndarray) -> np.ndarray:
    return a[1:-1, 1:-1, 1:-1]


def _box(centres: np.ndarray, c: float, radius: float) -> slice:
    """Index range of the cells whose centre is within radius of c along one
    axis (with a relative margin, so every cell the exact test accepts is
    inside); cells outside the range are outside the sphere."""
    idx = np.nonzero(np.abs(centres - c) <= radius * (1.0 + 1e-9))[0]
    if idx.size == 0:
        return slice(0, 0)
    return slice(int(idx[0]), int(idx[-1]) + 1)


def _bubble(dom: Domain, amp: float, radius: float, x0: float, y0: float,
            z0: float) -> np.ndarray:
    """amp cos^2(pi r / 2R) for r <= R, else 0, on every cell (halos too)."""
    out = np.zeros(dom.shape)
    bx, by, bz = _box(dom.xn, x0, radius), _box(dom.yn, y0, radius), _box(dom.zn, z0, radius)
    X, Y, Z = np.meshgrid(dom.xn[bx], dom.yn[by], dom.zn[bz], indexing="ij")
    r = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2 + (Z - z0) ** 2)
    out[bx, by, bz] = np.where(r <= radius, amp * np.cos(math.pi * r / (2.0 * radius)) ** 2, 0.0)
    return out


def _noise(dom: Domain, seed: int, field: int, amp: float) -> np.ndarray:
    out = np.zeros(dom.shape)
    u = uniform(seed, field, _cell_counter(dom))
    _interior(out)[...] = _interior(-amp + 2.0 * amp * u)
    return out


def _blobs(dom: Domain, seed: int, field: int, nblob: int, rmin: float,
           rmax: float) -> np.ndarray:
    """nblob seeded spheres of a seeded value (later blobs overwrite earlier
    ones), zero on the halo."""
    out = np.zeros(dom.shape)
    ctr = np.arange(5 * nblob, dtype=np.uint64) + np.uint64(1 << 40)
    r = uniform(seed, field, ctr).reshape(nblob, 5)
    xlo, xhi = dom.x[0], dom.x[dom.ip]
    ylo, yhi = dom.y[0], dom.y[dom.jp]
    zhi = dom.z[dom.kp]
    for b in range(nblob):
        cx = xlo + (xhi - xlo) * r[b, 0]
        cy = ylo + (yhi - ylo) * r[b, 1]
        cz = zhi * r[b, 2]
        rad = rmin + (rmax - rmin) * r[b, 3]
        val = r[b, 4]
        bx, by, bz = _box(dom.xn, cx, rad), _box(dom.yn, cy, rad), _box(dom.zn, cz, rad)
        X, Y, Z = np.meshgrid(dom.xn[bx], dom.yn[by], dom.zn[bz], indexing="ij")
        inside = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= rad * rad
        sub = out[bx, by, bz]
        out[bx, by, bz] = np.where(inside, val, sub)
    out[0, :, :] = 0.0
    out[-1, :, :] = 0.0
    out[:, 0, :] = 0.0
    out[:, -1, :] = 0.0
    out[:, :, 0] = 0.0
    out[:, :, -1] = 0.0
    return out


def warm_bubble(ip=64, jp=64, kp=64, delta=50.0, dt=1.0, steps=100, amp=2.0,
                radius=500.0, zc=500.0, noise=0.01, seed=101, name=None,
                ptol=1e-10) -> Case:
    """Dry warm bubble (the reference's CPU-runnable case at 64^3)."""
    dom = make_domain(ip, jp, kp, delta, delta, delta)
    th = _bubble(dom, amp, radius, 0.0, 0.0, zc) + _noise(dom, seed, 1, noise)
    zero = np.zeros(dom.shape)
    return Case(name or f"warm_bubble_{ip}x{jp}x{kp}", dom, dt, steps,
                seed=seed, ptol=ptol, u=zero.copy(), v=zero.copy(),
                w=zero.copy(), th=th)


def moist_bubble(ip=256, jp=256, kp=128, delta=50.0, dt=0.5, steps=200,
                 amp=1.5, radius=500.0, zc=500.0, noise=0.01, ntracer=8,
                 seed=202, name=None, ptol=1e-10) -> Case:
    """Moist bubble with qv and passive tracers, each its own SFVT category."""
    dom = make_domain(ip, jp, kp, delta, delta, delta)
    th = _bubble(dom, amp, radius, 0.0, 0.0, zc) + _noise(dom, seed, 1, noise)
    nq = 1 + ntracer
    q = np.zeros((nq,) + dom.shape)
    qv = 0.014 * np.exp(-dom.zn / 2000.0)
    q[0][1:-1, 1:-1, 1:-1] = qv[None, None, 1:-1]
    for t in range(ntracer):
        q[1 + t] = _blobs(dom, seed, 10 + t, 4, 200.0, 600.0)
    zero = np.zeros(dom.shape)
    cats = tuple(range(1, nq + 1))
    return Case(name or f"moist_bubble_{ip}x{jp}x{kp}", dom, dt, steps,
                seed=seed, ptol=ptol, nq=nq, c_s=cats, c_e=cats,
                u=zero.copy(), v=zero.copy(), w=zero.copy(), th=th, q=q)


def sheared_moist_bubble(ip=256, jp=256, kp=128, delta=50.0, dt=0.5, steps=200, ntracer=8,
                         seed=71, damping=True, forcing=True, subsidence=True, name=None,
                         ptol=1e-10) -> Case:
    """The moist bubble in a linearly sheared wind (u = 2 + 0.004 z, v = -u/2)
    with the subgrid model on, a 300 s damping layer over the top quarter,
    600 s relaxation towards the initial column (1,1) and subsidence of
    2e-4 1/s below mid-depth, as the options of the reference's sources
    (dynamics.f90:639-1000); the surface roughness lengths are 0.05 m
    (momentum) and 0.01 m (heat)."""
    case = moist_bubble(ip, jp, kp, delta=delta, dt=dt, steps=steps, ntracer=ntracer,
                        seed=seed, ptol=ptol,
                        name=name or f"sheared_moist_bubble_{ip}x{jp}x{kp}")
    dom = case.dom
    shear = 2.0 + 0.004 * dom.zn
    case.u = case.u + shear[None, None, :]
    case.v = case.v - 0.5 * shear[None, None, :]
    case.physics = physics_profiles(dom, u=case.u, v=case.v, z0=0.05, z0th=0.01,
                                    damping_layer=damping, damping_tau=300.0,
                                    forcing=forcing, forcing_tau=600.0,
                                    divergence=subsidence, divergence_val=2e-4,
                                    divergence_hgt=0.5 * dom.z[kp])
    return case


def kelvin_helmholtz(ip=512, jp=4, kp=256, delta=10.0, dt=0.2, steps=1000,
                     seed=303, name=None, ptol=1e-10) -> Case:
    """Shear layer u=10 tanh((z-1280)/50), stable theta plus a 0.5 K jump."""
    zmid = kp * delta / 2.0

    def theta_fn(z):
        z = np.asarray(z, dtype=np.float64)
        return 300.0 + 0.003 * z + 0.25 * (1.0 + np.tanh((z - zmid) / 50.0))

    dom = make_domain(ip, jp, kp, delta, delta, delta, theta_fn)
    u = np.zeros(dom.shape)
    _interior(u)[...] = (10.0 * np.tanh((dom.zn - zmid) / 50.0))[None, None, 1:-1]
    w = _noise(dom, seed, 3, 1e-3)
    return Case(name or f"kelvin_helmholtz_{ip}x{jp}x{kp}", dom, dt, steps,
                seed=seed, ptol=ptol, u=u, v=np.zeros(dom.shape), w=w,
                th=np.zeros(dom.shape))


def large_domain(ip=512, jp=512, kp=256, delta=25.0, dt=0.25, steps=100,
                 ntracer=4, seed=505, name=None, ptol=1e-10) -> Case:
    """3x3 array of +2 K bubbles (R=400 m) with theta, qv and tracers."""
    dom = make_domain(ip, jp, kp, delta, delta, delta)
    Lx, Ly = ip * delta, jp * delta
    th = _noise(dom, seed, 1, 0.02)
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            th = th + _bubble(dom, 2.0, 400.0, a * Lx / 4.0, b * Ly / 4.0, 500.0)
    nq = 1 + ntracer
    q = np.zeros((nq,) + dom.shape)
    q[0][1:-1, 1:-1, 1:-1] = (0.014 * np.exp(-dom.zn / 2000.0))[None, None, 1:-1]
    for t in range(ntracer):
        q[1 + t] = _blobs(dom, seed, 10 + t, 4, 200.0, 600.0)
    zero = np.zeros(dom.shape)
    cats = tuple(range(1, nq + 1))
    return Case(name or f"large_domain_{ip}x{jp}x{kp}", dom, dt, steps,
                seed=seed, ptol=ptol, nq=nq, c_s=cats, c_e=cats,
                u=zero.copy(), v=zero.copy(), w=zero.copy(), th=th, q=q)


def ensemble_members(n=64, seed=404, dt=1.0, steps=200):
    """Uneven warm-bubble members: nx,ny in {32..128}, nz in {32,64,96}."""
    sizes_xy = (32, 48, 64, 96, 128)
    sizes_z = (32, 64, 96)
    r = uniform(seed, 0, np.arange(5 * n, dtype=np.uint64)).reshape(n, 5)
    out = []
    for m in range(n):
        ip = sizes_xy[int(r[m, 0] * 5)]
        jp = sizes_xy[int(r[m, 1] * 5)]
        kp = sizes_z[int(r[m, 2] * 3)]
        amp = 0.5 + 2.5 * r[m, 3]
        rad = 200.0 + 600.0 * r[m, 4]
        out.append(warm_bubble(ip, jp, kp, 50.0, dt, steps, amp, rad,
                               zc=min(500.0, 0.4 * kp * 50.0), noise=0.01,
                               seed=seed * 1000 + m, name=f"member{m}"))
    return out


CONFIGS = {
    "warm_bubble_64": lambda **kw: warm_bubble(**kw),
    "moist_bubble_256x256x128": lambda **kw: moist_bubble(**kw),
    "kelvin_helmholtz_512x4x256": lambda **kw: kelvin_helmholtz(**kw),
    "large_domain_512x512x256": lambda **kw: large_domain(**kw),
}
