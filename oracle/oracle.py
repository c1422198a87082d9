"""ctypes binding of the CPU oracle (``libdcm_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
CPU-baseline leg of the benchmark as the checker.  The product path never
imports this module.

Arrays are numpy float64, C-contiguous, shape ``(ip+2, jp+2, kp+2)`` (the
reference's Fortran ``(0:kp+1,0:jp+1,0:ip+1)`` layout); every call works in
place exactly like the Fortran routine it restates.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DCM_ORACLE_LIB: another build of dcm_oracle.c (the mutation tests load
# copies with a planted mistake this way)
LIB_PATH = os.environ.get("DCM_ORACLE_LIB") or os.path.join(HERE, "libdcm_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Grid(C.Structure):
    _fields_ = [("ip", C.c_int), ("jp", C.c_int), ("kp", C.c_int),
                ("dx", _dp), ("dy", _dp), ("dz", _dp),
                ("dxn", _dp), ("dyn", _dp), ("dzn", _dp),
                ("rhoa", _dp), ("rhoan", _dp)]


class _Physics(C.Structure):
    _fields_ = [("viscous", C.c_int), ("moisture", C.c_int), ("damping_layer", C.c_int),
                ("forcing", C.c_int), ("divergence", C.c_int), ("nq", C.c_int),
                ("forcing_tau", C.c_double), ("z0", C.c_double), ("z0th", C.c_double),
                ("log_z1_z0", C.c_double),
                ("ubar", _dp), ("vbar", _dp), ("wbar", _dp),
                ("dampfacn", _dp), ("dampfac", _dp),
                ("u_force", _dp), ("v_force", _dp), ("w_subs", _dp),
                ("lamsq", _dp), ("z", _dp), ("theta", _dp), ("thetan", _dp)]


class _Model(C.Structure):
    _fields_ = [("dt", C.c_double), ("ptol", C.c_double),
                ("kord", C.c_int), ("monotone", C.c_int),
                ("advection_scheme", C.c_int),
                ("moisture", C.c_int), ("nq", C.c_int), ("ncat", C.c_int),
                ("c_s", _ip), ("c_e", _ip),
                ("theta", _dp), ("thetan", _dp),
                ("xg", _dp), ("yg", _dp), ("zg", _dp),
                ("ut", _dp), ("zut", _dp), ("tut", _dp),
                ("vt", _dp), ("zvt", _dp), ("tvt", _dp),
                ("wt", _dp), ("zwt", _dp), ("twt", _dp),
                ("th", _dp), ("p", _dp), ("q", _dp),
                ("su", _dp), ("sv", _dp), ("sw", _dp), ("sth", _dp),
                ("psrc", _dp), ("sq", _dp),
                ("ubar", _dp), ("vbar", _dp), ("wbar", _dp), ("thbar", _dp),
                ("qbar", _dp),
                ("n", C.c_int), ("solver_first", C.c_int),
                ("t1", C.c_double), ("t2", C.c_double),
                ("last_iters", C.c_int), ("last_err", C.c_double),
                ("phys", C.POINTER(_Physics)),
                ("strain", _dp), ("vism", _dp), ("vist", _dp), ("viss", _dp),
                ("thbase", C.c_double), ("thtop", C.c_double)]


class Physics:
    """An or_physics struct built from a workloads.Physics and its domain;
    the horizontal means are arrays of this object (set them before use)."""

    def __init__(self, dom, ph, nq=0):
        n = dom.kp + 2
        self.ubar, self.vbar, self.wbar = np.zeros(n), np.zeros(n), np.zeros(n)
        arr = lambda a: np.ascontiguousarray(np.zeros(n) if a is None else a, dtype=np.float64)
        self._keep = [arr(a) for a in (ph.dampfacn, ph.dampfac, ph.u_force, ph.v_force,
                                        ph.w_subs, ph.lamsq, dom.z, dom.theta, dom.thetan)]
        dfn, df, uf, vf, ws, lam, z, th, thn = self._keep
        self.s = _Physics(int(ph.viscous), int(nq > 0), int(ph.damping_layer), int(ph.forcing),
                          int(ph.divergence), nq, ph.forcing_tau, ph.z0, ph.z0th,
                          math.log(dom.z[1] / ph.z0), ptr(self.ubar), ptr(self.vbar),
                          ptr(self.wbar), ptr(dfn), ptr(df), ptr(uf), ptr(vf), ptr(ws),
                          ptr(lam), ptr(z), ptr(th), ptr(thn))
        self.ref = C.byref(self.s)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        g = C.POINTER(_Grid)
        d = C.c_double
        L.or_fill_along_dim.argtypes = [g, _dp, d, d]
        L.or_fill_along_dim_wo.argtypes = [g, _dp]
        L.or_fill_full.argtypes = [g, _dp, d, d]
        L.or_fill_full_wo.argtypes = [g, _dp]
        L.or_fill_along_z.argtypes = [g, _dp, d, d]
        L.or_upwind3d.argtypes = [g, d, _dp, _dp, _dp, _dp, C.c_int]
        L.or_mpdata3d.argtypes = [g, d, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
        L.or_mpdata3d.restype = C.c_int
        L.or_mpdata_vec3d.argtypes = [g, d, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp,
                                      C.c_int, C.c_int, C.c_int]
        L.or_mpdata_vec3d.restype = C.c_int
        L.or_adv_ref_state.argtypes = [g, d, _dp, _dp, _dp]
        ph = C.POINTER(_Physics)
        L.or_subgrid.argtypes = [g, ph] + [_dp] * 17
        L.or_sources_phys.argtypes = [g, ph] + [_dp] * 20
        L.or_upwind_vert3d.argtypes = [g, d, _dp, _dp, C.c_int]
        L.or_mpdata_vert3d.argtypes = [g, d, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
        L.or_mpdata_vert3d.restype = C.c_int
        L.or_mpdata_vec_vert3d.argtypes = [g, d, _dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                           C.c_int]
        L.or_mpdata_vec_vert3d.restype = C.c_int
        L.or_mpdata3d_add.argtypes = [g, d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                      C.c_int, C.c_int]
        L.or_mpdata3d_add.restype = C.c_int
        L.or_hmeans.argtypes = [g, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.or_sources.argtypes = [g, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_matvec.argtypes = [g, _dp, _dp]
        L.or_precondition.argtypes = [g, _dp, _dp, C.c_int]
        L.or_inn_prod.argtypes = [g, _dp, _dp]
        L.or_inn_prod.restype = d
        L.or_bicgstab.argtypes = [g, _dp, _dp, _dp, _dp, _dp, _dp, d, C.c_int, C.c_int,
                                  C.c_int, _ip, _dp]
        L.or_bicgstab.restype = C.c_int
        L.or_advance_momentum.argtypes = [g, d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_divergence.argtypes = [g, _dp, _dp, _dp, _dp]
        L.or_set_sum_blocked.argtypes = [C.c_int]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = C.c_int
        L.or_set_mpdata_variant.argtypes = [C.c_int]
        L.or_advance_scalars.argtypes = [g, d, _dp, _dp, _dp, _dp, C.c_int, C.c_int]
        m = C.POINTER(_Model)
        L.or_model_init_bcs.argtypes = [g, m]
        L.or_model_step.argtypes = [g, m]
        L.or_model_step.restype = C.c_int
        _lib = L
    return _lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "need C-contiguous float64"
    return a.ctypes.data_as(_dp)


class Grid:
    """Grid metrics of a workloads.Domain, in the oracle's struct."""

    def __init__(self, dom):
        self.dom = dom
        self._keep = [np.ascontiguousarray(getattr(dom, n), dtype=np.float64)
                      for n in ("dx", "dy", "dz", "dxn", "dyn", "dzn", "rhoa", "rhoan")]
        self.s = _Grid(dom.ip, dom.jp, dom.kp, *[ptr(a) for a in self._keep])
        self.ref = C.byref(self.s)

    @property
    def shape(self):
        return self.dom.shape


# ---- thin wrappers (in place, like the Fortran routines) -------------------

def fill_full(g: Grid, a, lbc=0.0, ubc=0.0):
    lib().or_fill_full(g.ref, ptr(a), lbc, ubc)


def fill_full_wo(g: Grid, a):
    lib().or_fill_full_wo(g.ref, ptr(a))


def fill_along_dim(g: Grid, a, lbc=0.0, ubc=0.0):
    lib().or_fill_along_dim(g.ref, ptr(a), lbc, ubc)


def fill_along_dim_wo(g: Grid, a):
    lib().or_fill_along_dim_wo(g.ref, ptr(a))


def fill_along_z(g: Grid, a, lbc=0.0, ubc=0.0):
    lib().or_fill_along_z(g.ref, ptr(a), lbc, ubc)


def upwind3d(g: Grid, dt, u, v, w, psi, neumann=0):
    lib().or_upwind3d(g.ref, dt, ptr(u), ptr(v), ptr(w), ptr(psi), neumann)


def mpdata3d(g: Grid, dt, u, v, w, psi, lbc=0.0, ubc=0.0, kord=2, monotone=True,
             neumann=0):
    lb, ub = C.c_double(lbc), C.c_double(ubc)
    r = lib().or_mpdata3d(g.ref, dt, ptr(u), ptr(v), ptr(w), ptr(psi),
                          C.byref(lb), C.byref(ub), kord, int(monotone), neumann)
    return r, lb.value, ub.value


def mpdata_vec3d(g: Grid, dt, u, v, w, psi, lbc=None, ubc=None, kord=2,
                 monotone=True, neumann=0):
    nq = psi.shape[0]
    lb = np.zeros(nq) if lbc is None else np.array(lbc, dtype=np.float64)
    ub = np.zeros(nq) if ubc is None else np.array(ubc, dtype=np.float64)
    r = lib().or_mpdata_vec3d(g.ref, dt, ptr(u), ptr(v), ptr(w), ptr(psi), nq,
                              ptr(lb), ptr(ub), kord, int(monotone), neumann)
    return r, lb, ub


def upwind_vert3d(g: Grid, dt, w, psi, neumann=0):
    lib().or_upwind_vert3d(g.ref, dt, ptr(w), ptr(psi), neumann)


def mpdata_vert3d(g: Grid, dt, w, psi, lbc=0.0, ubc=0.0, kord=2, monotone=True, neumann=0):
    lb, ub = C.c_double(lbc), C.c_double(ubc)
    r = lib().or_mpdata_vert3d(g.ref, dt, ptr(w), ptr(psi), C.byref(lb), C.byref(ub), kord,
                               int(monotone), neumann)
    return r, lb.value, ub.value


def mpdata_vec_vert3d(g: Grid, dt, w, psi, lbc=None, ubc=None, kord=2, monotone=True,
                      neumann=0):
    nq = psi.shape[0]
    lb = np.zeros(nq) if lbc is None else np.array(lbc, dtype=np.float64)
    ub = np.zeros(nq) if ubc is None else np.array(ubc, dtype=np.float64)
    r = lib().or_mpdata_vec_vert3d(g.ref, dt, ptr(w), ptr(psi), nq, ptr(lb), ptr(ub), kord,
                                   int(monotone), neumann)
    return r, lb, ub


def mpdata3d_add(g: Grid, dt, u, v, w, psi, psi_1d, lbc=0.0, ubc=0.0, kord=2, monotone=True,
                 neumann=0):
    prof = np.ascontiguousarray(psi_1d, dtype=np.float64)
    lb, ub = C.c_double(lbc), C.c_double(ubc)
    r = lib().or_mpdata3d_add(g.ref, dt, ptr(u), ptr(v), ptr(w), ptr(psi), ptr(prof),
                              C.byref(lb), C.byref(ub), kord, int(monotone), neumann)
    return r, lb.value, ub.value


def subgrid(g: Grid, ph: Physics, u, v, w, zu, zv, zw, th, q, vism, vist, viss, strain, su, sv,
            sw, sth, sq):
    lib().or_subgrid(g.ref, ph.ref, *[ptr(a) for a in (u, v, w, zu, zv, zw, th, q, vism, vist,
                                                        viss, strain, su, sv, sw, sth, sq)])


def sources_phys(g: Grid, ph: Physics, u, v, w, zu, zv, zw, th, q, theta, thbar, su, sv, sw,
                 sth, sq, strain, vism, vist, viss, rhs):
    lib().or_sources_phys(g.ref, ph.ref, *[ptr(np.ascontiguousarray(a)) if a is not None else None
                                           for a in (u, v, w, zu, zv, zw, th, q, theta, thbar)],
                          *[ptr(a) for a in (su, sv, sw, sth, sq, strain, vism, vist, viss,
                                             rhs)])


def adv_ref_state(g: Grid, dt, w, psi, ref):
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    lib().or_adv_ref_state(g.ref, dt, ptr(w), ptr(psi), ptr(ref))


def hmeans(g: Grid, u, v, w, th, q=None):
    n = g.dom.kp + 2
    nq = 0 if q is None else q.shape[0]
    out = [np.zeros(n) for _ in range(4)]
    qbar = np.zeros((max(nq, 1), n))
    lib().or_hmeans(g.ref, ptr(u), ptr(v), ptr(w), ptr(th),
                    ptr(q) if q is not None else None, nq,
                    *[ptr(a) for a in out], ptr(qbar))
    return out + [qbar[:nq]]


def sources(g: Grid, u, v, w, th, theta, thbar, su, sv, sw, sth, rhs):
    lib().or_sources(g.ref, ptr(u), ptr(v), ptr(w), ptr(th),
                     ptr(np.ascontiguousarray(theta)), ptr(np.ascontiguousarray(thbar)),
                     ptr(su), ptr(sv), ptr(sw), ptr(sth), ptr(rhs))


def matvec(g: Grid, x, ax):
    lib().or_matvec(g.ref, ptr(x), ptr(ax))


def precondition(g: Grid, s, r, prit=3):
    lib().or_precondition(g.ref, ptr(s), ptr(r), prit)


def inn_prod(g: Grid, x, y) -> float:
    return lib().or_inn_prod(g.ref, ptr(x), ptr(y))


def bicgstab(g: Grid, sw, x, b, tol=1e-10, test_solver=False, first_call=True,
             itmax=9999):
    d = g.dom
    it, err = C.c_int(0), C.c_double(0.0)
    st = lib().or_bicgstab(g.ref, ptr(d.x), ptr(d.y), ptr(d.z), ptr(sw), ptr(x),
                           ptr(b), tol, int(test_solver), int(first_call), itmax,
                           C.byref(it), C.byref(err))
    return st, it.value, err.value


def set_mpdata_variant(v: int):
    """0: pamm/sfvt/advection_3d.F90 (the copy linked into the model);
    1: sfvt/advection_3d.F90 (per-direction limiter, unweighted vector
    antidiffusive velocities)."""
    lib().or_set_mpdata_variant(int(v))


def set_sum_blocked(on):
    """Blocked (64-way) summation in the dot products instead of the serial
    loop; 2: a parallel reduction over the host threads (timing only)."""
    lib().or_set_sum_blocked(int(on))


def set_threads(n: int) -> int:
    """Host threads of the element-wise loops (1: serial, the default); the
    results of those loops do not depend on it.  Returns the count in effect."""
    return lib().or_set_threads(int(n))


def divergence(g: Grid, u, v, w, div):
    lib().or_divergence(g.ref, ptr(u), ptr(v), ptr(w), ptr(div))


def advance_momentum(g: Grid, dt, u, v, w, zu, zv, zw, su, sv, sw, p):
    lib().or_advance_momentum(g.ref, dt, ptr(u), ptr(v), ptr(w), ptr(zu), ptr(zv),
                              ptr(zw), ptr(su), ptr(sv), ptr(sw), ptr(p))


def advance_scalars(g: Grid, dt, q, sq, th, sth, moisture):
    nq = 0 if q is None else q.shape[0]
    lib().or_advance_scalars(g.ref, dt, ptr(q) if q is not None else None,
                             ptr(sq) if sq is not None else None, ptr(th), ptr(sth),
                             nq, int(moisture))


class Model:
    """The oracle's model_driver time loop over a workloads.Case."""

    def __init__(self, case):
        self.case = case
        d = case.dom
        self.g = Grid(d)
        sh = d.shape
        z = lambda: np.zeros(sh)
        self.ut, self.vt, self.wt = case.u.copy(), case.v.copy(), case.w.copy()
        self.th = case.th.copy()
        for a, lbc in ((self.ut, 0.0), (self.vt, 0.0), (self.wt, 0.0), (self.th, 0.0)):
            fill_full(self.g, a, lbc, 0.0)  # initialisation.f90:861-868
        self.zut, self.zvt, self.zwt = self.ut.copy(), self.vt.copy(), self.wt.copy()
        self.tut, self.tvt, self.twt = self.ut.copy(), self.vt.copy(), self.wt.copy()
        self.p, self.psrc = z(), z()
        self.su, self.sv, self.sw, self.sth = z(), z(), z(), z()
        nq = max(case.nq, 1)
        self.q = np.zeros((nq,) + sh) if case.q is None else case.q.copy()
        if case.q is not None:
            for n in range(case.nq):
                fill_full(self.g, self.q[n], 0.0, 0.0)
        self.sq = np.zeros((nq,) + sh)
        kk = d.kp + 2
        self.ubar, self.vbar, self.wbar, self.thbar = (np.zeros(kk) for _ in range(4))
        self.qbar = np.zeros((nq, kk))
        self.cs = (C.c_int * max(case.ncat, 1))(*(case.c_s or (1,)))
        self.ce = (C.c_int * max(case.ncat, 1))(*(case.c_e or (1,)))
        self._keep = [np.ascontiguousarray(a) for a in (d.theta, d.thetan, d.x, d.y, d.z)]
        th_, thn_, xg, yg, zg = self._keep
        self.m = _Model(case.dt, case.ptol, case.kord, int(case.monotone),
                        case.advection_scheme, int(case.moisture), case.nq,
                        case.ncat, self.cs, self.ce, ptr(th_), ptr(thn_),
                        ptr(xg), ptr(yg), ptr(zg),
                        ptr(self.ut), ptr(self.zut), ptr(self.tut),
                        ptr(self.vt), ptr(self.zvt), ptr(self.tvt),
                        ptr(self.wt), ptr(self.zwt), ptr(self.twt),
                        ptr(self.th), ptr(self.p), ptr(self.q),
                        ptr(self.su), ptr(self.sv), ptr(self.sw), ptr(self.sth),
                        ptr(self.psrc), ptr(self.sq),
                        ptr(self.ubar), ptr(self.vbar), ptr(self.wbar),
                        ptr(self.thbar), ptr(self.qbar), 0, 1, 0.0, 0.0, 0, 0.0)
        # find_base_top (initialisation.f90:877): the bcs of scheme 2
        self.m.thbase, self.m.thtop = float(d.thetan[0]), float(d.thetan[d.kp + 1])
        self.phys = None
        if case.physics is not None:
            self.phys = Physics(d, case.physics, case.nq)
            self.strain, self.vism, self.vist = z(), z(), z()
            self.viss = np.zeros((nq,) + sh)
            self.m.phys = C.pointer(self.phys.s)
            self.m.strain, self.m.vism = ptr(self.strain), ptr(self.vism)
            self.m.vist, self.m.viss = ptr(self.vist), ptr(self.viss)
        lib().or_model_init_bcs(self.g.ref, C.byref(self.m))
        self.iters = []

    def step(self, n=1):
        for _ in range(n):
            st = lib().or_model_step(self.g.ref, C.byref(self.m))
            self.iters.append(self.m.last_iters)
            if st != 0:
                raise RuntimeError(f"oracle step failed with status {st}")

    def levels(self):
        """(current, previous, future) momentum levels as in driver_code.f90:634-646."""
        r = self.m.n % 3
        U = (self.ut, self.tut, self.zut)
        V = (self.vt, self.tvt, self.zvt)
        W = (self.wt, self.twt, self.zwt)
        Z = (self.zut, self.ut, self.tut)
        ZV = (self.zvt, self.vt, self.tvt)
        ZW = (self.zwt, self.wt, self.twt)
        return (U[r], V[r], W[r]), (Z[r], ZV[r], ZW[r])
