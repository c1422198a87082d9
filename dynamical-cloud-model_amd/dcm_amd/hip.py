"""ctypes binding of the C-ABI library ``libdcm_hip.so`` (include/dcm_hip.h).

This is the Python host of the HIP path: the tests and the benchmark drive
the kernels through exactly the entry points a Fortran host binds.  Device
memory is owned by the library (``dcm_malloc``), so no other GPU runtime is
loaded into the process.  There is no CPU fallback: if the library cannot be
loaded or no GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libdcm_hip.so")

STATUS = {0: "DCM_OK", 1: "DCM_NOT_CONVERGED", 2: "DCM_BREAKDOWN", 3: "DCM_NONFINITE",
          4: "DCM_BAD_ARG", 5: "DCM_HIP_ERROR", 6: "DCM_NO_DEVICE", 7: "DCM_OUT_OF_MEMORY"}

FP64, MIXED = 0, 1
# the copy of advection_3d.F90 MPDATA follows (dcm_set_mpdata_variant)
MPDATA_LINKED, MPDATA_SFVT = 0, 1
SUM_TREE, SUM_SEQUENTIAL = 0, 1
(F_U, F_V, F_W, F_ZU, F_ZV, F_ZW, F_TH, F_P, F_PSRC, F_SU, F_SV, F_SW, F_STH, F_Q,
 F_THBAR) = range(15)

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_d = C.c_double
_i = C.c_int


class DcmError(RuntimeError):
    def __init__(self, fn, status, msg=""):
        self.status = status
        super().__init__(f"{fn} failed: {STATUS.get(status, status)} {msg}")


class StepStats(C.Structure):
    _fields_ = [("steps", _i), ("iters_total", _i), ("iters_max", _i), ("outer_total", _i),
                ("fallbacks", _i), ("reserved", _i), ("last_err", _d), ("max_true_resid", _d)]


class PhysicsStruct(C.Structure):
    """dcm_physics: options of sources beyond the inviscid branch."""
    _fields_ = [("viscous", _i), ("damping_layer", _i), ("forcing", _i), ("divergence", _i),
                ("forcing_tau", _d), ("z0", _d), ("z0th", _d),
                ("lamsq", _dp), ("dampfacn", _dp), ("dampfac", _dp), ("u_force", _dp),
                ("v_force", _dp), ("w_subs", _dp), ("z", _dp), ("theta", _dp), ("thetan", _dp)]


def physics_struct(dom, ph):
    """A dcm_physics for a workloads.Physics on its domain; returns the
    struct and the arrays it points into (keep both alive)."""
    keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
            for a in (ph.lamsq, ph.dampfacn, ph.dampfac, ph.u_force, ph.v_force, ph.w_subs,
                      dom.z, dom.theta, dom.thetan)]
    st = PhysicsStruct(int(ph.viscous), int(ph.damping_layer), int(ph.forcing),
                       int(ph.divergence), ph.forcing_tau, ph.z0, ph.z0th,
                       *[_dptr(a) if a is not None else None for a in keep])
    return st, keep


class InitConfig(C.Structure):
    """dcm_init_config of include/dcm_hip.h, field for field."""

    _fields_ = [("ip", _i), ("jp", _i), ("kp", _i),
                ("dx", _d), ("dy", _d), ("dz", _d), ("dt", _d), ("runtime", _d),
                ("moisture", _i), ("theta_flag", _i), ("nq", _i),
                ("iqv", _i), ("iqc", _i), ("inc", _i),
                ("drop_num", _d), ("psurf", _d), ("tsurf", _d),
                ("adiabatic_prof", _i),
                ("adiabatic_frac", _d), ("t_cbase", _d), ("t_ctop", _d), ("rh_above", _d),
                ("th_jump", _d), ("th_grad", _d),
                ("n_levels", _i),
                ("z_read", _dp), ("theta_read", _dp), ("u_read", _dp), ("v_read", _dp),
                ("q_read", _dp),
                ("damping_layer", _i), ("damping_thickness", _d), ("damping_tau", _d),
                ("forcing", _i), ("forcing_tau", _d),
                ("divergence", _i), ("divergence_val", _d), ("divergence_hgt", _d),
                ("cvis", _d), ("z0", _d), ("z0th", _d),
                ("bubble", _i), ("param_theta", _i), ("param_wind", _i),
                ("param_vmax", _d), ("param_z", _d), ("param_sigz", _d), ("param_delz", _d),
                ("seed", C.c_ulonglong)]


PROFILE_ARRAYS = ("x", "y", "z", "xn", "yn", "zn", "dx", "dy", "dz", "dxn", "dyn", "dzn",
                   "theta", "thetan", "rhoa", "rhoan", "pref", "prefn", "tref", "trefn",
                   "lamsq", "lamsqn", "dampfac", "dampfacn", "u_force", "v_force", "w_subs",
                   "u", "v", "q")


class InitProfiles(C.Structure):
    """dcm_init_profiles of include/dcm_hip.h."""

    _fields_ = [(n, _dp) for n in PROFILE_ARRAYS] + [("thbase", _d), ("thtop", _d),
                                                      ("ntim", _i)]


_SIGS = {
    "dcm_last_error": ([], C.c_char_p),
    "dcm_version": ([], C.c_char_p),
    "dcm_device_memory": ([C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], _i),
    "dcm_context_create": ([_i, _i, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                            C.POINTER(_vp)], _i),
    "dcm_context_destroy": ([_vp], _i),
    "dcm_field_elems": ([_vp], C.c_size_t),
    "dcm_malloc": ([_vp, C.c_size_t, C.POINTER(_vp)], _i),
    "dcm_free": ([_vp, _vp], _i),
    "dcm_memcpy_htod": ([_vp, _vp, _vp, C.c_size_t], _i),
    "dcm_memcpy_dtoh": ([_vp, _vp, _vp, C.c_size_t], _i),
    "dcm_synchronize": ([_vp], _i),
    "dcm_mirror": ([_vp, _vp, C.c_size_t, C.POINTER(_vp)], _i),
    "dcm_mirror_upload": ([_vp, _vp], _i),
    "dcm_mirror_download": ([_vp, _vp], _i),
    "dcm_mirror_release_all": ([_vp], _i),
    "dcm_fill_full": ([_vp, _vp, _d, _d], _i),
    "dcm_fill_along_dim": ([_vp, _vp, _d, _d], _i),
    "dcm_fill_along_dim_wo": ([_vp, _vp], _i),
    "dcm_fill_full_wo": ([_vp, _vp], _i),
    "dcm_fill_along_z": ([_vp, _vp, _d, _d], _i),
    "dcm_upwind3d": ([_vp, _d, _vp, _vp, _vp, _vp, _i], _i),
    "dcm_mpdata3d": ([_vp, _d, _vp, _vp, _vp, _vp, _dp, _dp, _i, _i, _i, _ip], _i),
    "dcm_mpdata_vec3d": ([_vp, _d, _vp, _vp, _vp, _vp, _i, _dp, _dp, _i, _i, _i, _ip], _i),
    "dcm_adv_ref_state": ([_vp, _d, _vp, _vp, _dp], _i),
    "dcm_subgrid": ([_vp, C.POINTER(PhysicsStruct)] + [_vp] * 8 + [_i] + [_vp] * 9, _i),
    "dcm_sources_phys": ([_vp, C.POINTER(PhysicsStruct), _dp, _dp, _dp] + [_vp] * 8 + [_i] +
                         [_dp, _dp] + [_vp] * 10, _i),
    "dcm_mpdata3d_add": ([_vp, _d, _vp, _vp, _vp, _vp, _dp, _dp, _dp, _i, _i, _i, _ip], _i),
    "dcm_upwind_vert3d": ([_vp, _d, _vp, _vp, _i], _i),
    "dcm_mpdata_vert3d": ([_vp, _d, _vp, _vp, _dp, _dp, _i, _i, _i, _ip], _i),
    "dcm_mpdata_vec_vert3d": ([_vp, _d, _vp, _vp, _i, _dp, _dp, _i, _i, _i, _ip], _i),
    "dcm_hmeans": ([_vp, _vp, _vp, _vp, _vp, _vp, _i, _dp, _dp, _dp, _dp, _dp], _i),
    "dcm_sources": ([_vp, _vp, _vp, _vp, _vp, _dp, _dp, _vp, _vp, _vp, _vp, _vp], _i),
    "dcm_matvec": ([_vp, _vp, _vp], _i),
    "dcm_precondition": ([_vp, _vp, _vp, _i], _i),
    "dcm_inn_prod": ([_vp, _vp, _vp, _dp], _i),
    "dcm_inn_prod_noalloc": ([_vp, _vp, _vp, _dp], _i),
    "dcm_set_sum_order": ([_vp, _i], _i),
    "dcm_set_precond_chebyshev": ([_vp, _d, _i], _i),
    "dcm_set_mpdata_variant": ([_vp, _i], _i),
    "dcm_set_precond_multigrid": ([_vp, _i], _i),
    "dcm_bicgstab": ([_vp, _dp, _dp, _dp, _vp, _vp, _vp, _d, _i, _i, _ip, _dp], _i),
    "dcm_bicgstab_mixed": ([_vp, _vp, _vp, _vp, _d, _i, _ip, _ip, _dp], _i),
    "dcm_true_residual": ([_vp, _vp, _vp, _dp], _i),
    "dcm_advance_momentum": ([_vp, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "dcm_divergence": ([_vp, _vp, _vp, _vp, _vp], _i),
    "dcm_advance_scalars": ([_vp, _d, _vp, _vp, _vp, _vp, _i, _i], _i),
    "dcm_model_create": ([_vp, _i, _i, _ip, _ip, _d, _i, _i, _i, _d, _dp, _dp, _dp, _dp, _dp,
                          C.POINTER(_vp)], _i),
    "dcm_model_destroy": ([_vp], _i),
    "dcm_model_set_physics": ([_vp, C.POINTER(PhysicsStruct)], _i),
    "dcm_model_set_precision": ([_vp, _i], _i),
    "dcm_model_set": ([_vp, _i, _i, _dp], _i),
    "dcm_model_get": ([_vp, _i, _i, _dp], _i),
    "dcm_model_field": ([_vp, _i, _i, C.POINTER(_vp)], _i),
    "dcm_model_init": ([_vp], _i),
    "dcm_model_step": ([_vp, _i, C.POINTER(StepStats)], _i),
    "dcm_model_profile": ([_vp, _i], _i),
    "dcm_model_profile_filter": ([_vp, C.c_char_p], _i),
    "dcm_model_set_inner_rtol": ([_vp, _d], _i),
    "dcm_model_set_check_residual": ([_vp, _i], _i),
    "dcm_model_set_graphs": ([_vp, _i], _i),
    "dcm_stream_copy": ([_vp, _vp, _vp, C.c_size_t, _dp], _i),
    "dcm_model_kernel_time": ([_vp, C.c_char_p, _dp, _ip], _i),
    "dcm_ensemble_create": ([_i, C.POINTER(_vp), C.POINTER(_vp)], _i),
    "dcm_ensemble_destroy": ([_vp], _i),
    "dcm_ensemble_autotune": ([_vp, _ip], _i),
    "dcm_ensemble_step": ([_vp, _i, C.POINTER(StepStats)], _i),
    "dcm_ensemble_set_graphs": ([_vp, _i], _i),
    "dcm_nc_create": ([C.c_char_p, _i, _i, _i, _i, _i, _i, C.c_char_p, _dp, _dp, _dp, _dp, _dp,
                       _dp, C.POINTER(_vp)], _i),
    "dcm_nc_put_record": ([_vp, _i, _d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp], _i),
    "dcm_nc_close": ([_vp], _i),
    "dcm_output_create": ([_vp, C.c_char_p, _i, _i, C.c_char_p, _dp, _dp, _dp, _dp, _dp, _dp,
                           C.POINTER(_vp)], _i),
    "dcm_output_write": ([_vp, _d, _dp], _i),
    "dcm_output_flush": ([_vp], _i),
    "dcm_output_stats": ([_vp, C.POINTER(C.c_longlong), _dp, _dp], _i),
    "dcm_output_destroy": ([_vp], _i),
    "dcm_nml_read": ([C.c_char_p, C.POINTER(_vp)], _i),
    "dcm_nml_free": ([_vp], _i),
    "dcm_nml_has": ([_vp, C.c_char_p, C.c_char_p], _i),
    "dcm_nml_get_real": ([_vp, C.c_char_p, C.c_char_p, _dp], _i),
    "dcm_nml_get_int": ([_vp, C.c_char_p, C.c_char_p, C.POINTER(_i)], _i),
    "dcm_nml_get_logical": ([_vp, C.c_char_p, C.c_char_p, C.POINTER(_i)], _i),
    "dcm_nml_get_string": ([_vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t], _i),
    "dcm_nml_get_array": ([_vp, C.c_char_p, C.c_char_p, _i, C.POINTER(_i), _dp], _i),
    "dcm_init_compute_profiles": ([C.POINTER(InitConfig), C.POINTER(InitProfiles)], _i),
    "dcm_model_init_fields": ([_vp, C.POINTER(InitConfig), C.POINTER(InitProfiles)], _i),
}

_lib = None


def declared_symbols():
    return sorted(_SIGS)


def load():
    """Load libdcm_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _call(name, *args):
    st = getattr(load(), name)(*args)
    if st != 0:
        raise DcmError(name, st, load().dcm_last_error().decode())
    return st


def device_memory():
    """(free, total) bytes of device memory of the current device."""
    free, total = C.c_size_t(0), C.c_size_t(0)
    _call("dcm_device_memory", C.byref(free), C.byref(total))
    return free.value, total.value


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class DeviceArray:
    """A device buffer owned by the library, with a numpy shape."""

    def __init__(self, ctx: "Context", shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        _call("dcm_malloc", ctx.h, max(self.nbytes, 8), C.byref(p))
        self.ptr = p

    def set(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _call("dcm_memcpy_htod", self.ctx.h, self.ptr, a.ctypes.data_as(_vp), self.nbytes)
        return self

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _call("dcm_memcpy_dtoh", self.ctx.h, out.ctypes.data_as(_vp), self.ptr, self.nbytes)
        return out

    def free(self):
        if self.ptr:
            _call("dcm_free", self.ctx.h, self.ptr)
            self.ptr = None


class Context:
    """A dcm_ctx for one workloads.Domain."""

    def __init__(self, dom):
        self.dom = dom
        self._keep = [np.ascontiguousarray(getattr(dom, n), dtype=np.float64)
                      for n in ("dx", "dy", "dz", "dxn", "dyn", "dzn", "rhoa", "rhoan")]
        h = _vp()
        _call("dcm_context_create", dom.ip, dom.jp, dom.kp, *[_dptr(a) for a in self._keep],
              C.byref(h))
        self.h = h

    def close(self):
        if self.h:
            _call("dcm_context_destroy", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- memory ---------------------------------------------------------------
    def array(self, a: np.ndarray) -> DeviceArray:
        return DeviceArray(self, a.shape, a.dtype).set(a)

    def zeros(self, shape=None) -> DeviceArray:
        shape = self.dom.shape if shape is None else shape
        return DeviceArray(self, shape).set(np.zeros(shape))

    def sync(self):
        _call("dcm_synchronize", self.h)

    # -- a1 fills -------------------------------------------------------------
    def fill_full(self, a, lbc=0.0, ubc=0.0):
        _call("dcm_fill_full", self.h, a.ptr, lbc, ubc)

    def fill_along_dim(self, a, lbc=0.0, ubc=0.0):
        _call("dcm_fill_along_dim", self.h, a.ptr, lbc, ubc)

    def fill_along_dim_wo(self, a):
        _call("dcm_fill_along_dim_wo", self.h, a.ptr)

    def fill_full_wo(self, a):
        _call("dcm_fill_full_wo", self.h, a.ptr)

    def fill_along_z(self, a, lbc=0.0, ubc=0.0):
        _call("dcm_fill_along_z", self.h, a.ptr, lbc, ubc)

    # -- a2-a5 advection ------------------------------------------------------
    def upwind3d(self, dt, u, v, w, psi, neumann=0):
        _call("dcm_upwind3d", self.h, dt, u.ptr, v.ptr, w.ptr, psi.ptr, neumann)

    def mpdata3d(self, dt, u, v, w, psi, lbc=0.0, ubc=0.0, kord=2, monotone=True, neumann=0):
        lb, ub, ex = _d(lbc), _d(ubc), _i(0)
        _call("dcm_mpdata3d", self.h, dt, u.ptr, v.ptr, w.ptr, psi.ptr, C.byref(lb),
              C.byref(ub), kord, int(monotone), neumann, C.byref(ex))
        return ex.value, lb.value, ub.value

    def mpdata_vec3d(self, dt, u, v, w, psi, nq, lbc=None, ubc=None, kord=2, monotone=True,
                     neumann=0):
        lb = np.zeros(nq) if lbc is None else np.array(lbc, dtype=np.float64)
        ub = np.zeros(nq) if ubc is None else np.array(ubc, dtype=np.float64)
        ex = _i(0)
        _call("dcm_mpdata_vec3d", self.h, dt, u.ptr, v.ptr, w.ptr, psi.ptr, nq, _dptr(lb),
              _dptr(ub), kord, int(monotone), neumann, C.byref(ex))
        return ex.value, lb, ub

    def adv_ref_state(self, dt, w, psi, ref):
        ref = np.ascontiguousarray(ref, dtype=np.float64)
        _call("dcm_adv_ref_state", self.h, dt, w.ptr, psi.ptr, _dptr(ref))

    def mpdata3d_add(self, dt, u, v, w, psi, psi_1d, lbc=0.0, ubc=0.0, kord=2, monotone=True,
                     neumann=0):
        prof = np.ascontiguousarray(psi_1d, dtype=np.float64)
        lb, ub, ex = _d(lbc), _d(ubc), _i(0)
        _call("dcm_mpdata3d_add", self.h, dt, u.ptr, v.ptr, w.ptr, psi.ptr, _dptr(prof),
              C.byref(lb), C.byref(ub), kord, int(monotone), neumann, C.byref(ex))
        return ex.value, lb.value, ub.value

    def upwind_vert3d(self, dt, w, psi, neumann=0):
        _call("dcm_upwind_vert3d", self.h, dt, w.ptr, psi.ptr, neumann)

    def mpdata_vert3d(self, dt, w, psi, lbc=0.0, ubc=0.0, kord=2, monotone=True, neumann=0):
        lb, ub, ex = _d(lbc), _d(ubc), _i(0)
        _call("dcm_mpdata_vert3d", self.h, dt, w.ptr, psi.ptr, C.byref(lb), C.byref(ub), kord,
              int(monotone), neumann, C.byref(ex))
        return ex.value, lb.value, ub.value

    def mpdata_vec_vert3d(self, dt, w, psi, nq, lbc=None, ubc=None, kord=2, monotone=True,
                          neumann=0):
        lb = np.zeros(nq) if lbc is None else np.array(lbc, dtype=np.float64)
        ub = np.zeros(nq) if ubc is None else np.array(ubc, dtype=np.float64)
        ex = _i(0)
        _call("dcm_mpdata_vec_vert3d", self.h, dt, w.ptr, psi.ptr, nq, _dptr(lb), _dptr(ub),
              kord, int(monotone), neumann, C.byref(ex))
        return ex.value, lb, ub

    # -- a6, a7 ---------------------------------------------------------------
    def hmeans(self, u, v, w, th, q=None, nq=0):
        n = self.dom.kp + 2
        out = [np.zeros(n) for _ in range(4)]
        qbar = np.zeros(max(nq, 1) * n)
        _call("dcm_hmeans", self.h, u.ptr, v.ptr, w.ptr, th.ptr, q.ptr if q else None, nq,
              *[_dptr(a) for a in out], _dptr(qbar))
        return out + [qbar[: nq * n].reshape(nq, n)]

    def sources(self, u, v, w, th, theta, thbar, su, sv, sw, sth, rhs):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        thbar = np.ascontiguousarray(thbar, dtype=np.float64)
        _call("dcm_sources", self.h, u.ptr, v.ptr, w.ptr, th.ptr, _dptr(theta), _dptr(thbar),
              su.ptr, sv.ptr, sw.ptr, sth.ptr if sth else None, rhs.ptr)

    def subgrid(self, ph, u, v, w, zu, zv, zw, th, q, nq, vism, vist, viss, strain, su, sv, sw,
                sth, sq):
        st, keep = physics_struct(self.dom, ph)
        _call("dcm_subgrid", self.h, C.byref(st), u.ptr, v.ptr, w.ptr, zu.ptr, zv.ptr, zw.ptr,
              th.ptr, q.ptr if q is not None else None, nq, vism.ptr, vist.ptr,
              viss.ptr if viss is not None else None, strain.ptr, su.ptr, sv.ptr, sw.ptr,
              sth.ptr, sq.ptr if sq is not None else None)

    def sources_phys(self, ph, ubar, vbar, wbar, u, v, w, zu, zv, zw, th, q, nq, theta, thbar,
                     su, sv, sw, sth, sq, strain, vism, vist, viss, rhs):
        st, keep = physics_struct(self.dom, ph)
        host = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                for a in (ubar, vbar, wbar, theta, thbar)]
        ub, vb, wb, tht, thb = host
        opt = lambda a: a.ptr if a is not None else None
        _call("dcm_sources_phys", self.h, C.byref(st), *[_dptr(a) if a is not None else None
                                                         for a in (ub, vb, wb)],
              u.ptr, v.ptr, w.ptr, zu.ptr, zv.ptr, zw.ptr, th.ptr, opt(q), nq, _dptr(tht),
              _dptr(thb), su.ptr, sv.ptr, sw.ptr, sth.ptr, opt(sq), opt(strain), opt(vism),
              opt(vist), opt(viss), rhs.ptr)

    # -- a8-a12 solver --------------------------------------------------------
    def matvec(self, x, ax):
        _call("dcm_matvec", self.h, x.ptr, ax.ptr)

    def precondition(self, s, r, prit=3):
        _call("dcm_precondition", self.h, s.ptr, r.ptr, prit)

    def set_sum_order(self, order):
        """SUM_TREE (default) or SUM_SEQUENTIAL (the reference's serial order)."""
        _call("dcm_set_sum_order", self.h, order)

    def set_precond_multigrid(self, on=True):
        """One multigrid V-cycle in place of the Jacobi sweeps (single models,
        dcm_bicgstab_mixed, and the FP64 BiCGSTAB).  on=2: the same cycle
        with every pass of the coarse levels in its own launch (the default,
        True or 1, runs the levels that fit one workgroup's LDS as one
        single-block kernel; the bits are the same)."""
        _call("dcm_set_precond_multigrid", self.h, 2 if on == 2 else int(bool(on)))

    def set_mpdata_variant(self, variant):
        """MPDATA_LINKED (pamm/sfvt/advection_3d.F90, the default) or
        MPDATA_SFVT (sfvt/advection_3d.F90's limiter and vector routines)."""
        _call("dcm_set_mpdata_variant", self.h, int(variant))

    def set_precond_chebyshev(self, a, divisions=4):
        """Chebyshev factors on [a, 2] for the mixed inner solve's sweeps, with
        `divisions` applications of 1/a_p (0: the reference's Jacobi sweeps)."""
        _call("dcm_set_precond_chebyshev", self.h, float(a), int(divisions))

    def inn_prod(self, x, y) -> float:
        out = _d()
        _call("dcm_inn_prod", self.h, x.ptr, y.ptr, C.byref(out))
        return out.value

    def inn_prod_noalloc(self, x, y) -> float:
        """inn_prod with its partial sums in the library's fixed device
        scratch (no allocation; calls overlapping on one device race)."""
        out = _d()
        _call("dcm_inn_prod_noalloc", self.h, x.ptr, y.ptr, C.byref(out))
        return out.value

    def bicgstab(self, sw, x, b, tol=1e-10, test_solver=False, itmax=0, check=True):
        d = self.dom
        it, err = _i(0), _d(0.0)
        keep = [np.ascontiguousarray(a) for a in (d.x, d.y, d.z)]
        st = load().dcm_bicgstab(self.h, *[_dptr(a) for a in keep], sw.ptr, x.ptr, b.ptr, tol,
                                 int(test_solver), itmax, C.byref(it), C.byref(err))
        if check and st != 0:
            raise DcmError("dcm_bicgstab", st, load().dcm_last_error().decode())
        return st, it.value, err.value

    def bicgstab_mixed(self, sw, x, b, tol=1e-10, itmax=0):
        it, outer, err = _i(0), _i(0), _d(0.0)
        _call("dcm_bicgstab_mixed", self.h, sw.ptr, x.ptr, b.ptr, tol, itmax, C.byref(it),
              C.byref(outer), C.byref(err))
        return it.value, outer.value, err.value

    def true_residual(self, x, b) -> float:
        out = _d()
        _call("dcm_true_residual", self.h, x.ptr, b.ptr, C.byref(out))
        return out.value

    def stream_copy(self, dst, src, n=None) -> float:
        """Plain device copy of n doubles; returns its kernel time in ms."""
        n = dst.nbytes // 8 if n is None else n
        ms = _d()
        _call("dcm_stream_copy", self.h, dst.ptr, src.ptr, n, C.byref(ms))
        return ms.value

    def divergence(self, u, v, w, div):
        _call("dcm_divergence", self.h, u.ptr, v.ptr, w.ptr, div.ptr)

    # -- a13, a14 -------------------------------------------------------------
    def advance_momentum(self, dt, u, v, w, zu, zv, zw, su, sv, sw, p):
        _call("dcm_advance_momentum", self.h, dt, u.ptr, v.ptr, w.ptr, zu.ptr, zv.ptr, zw.ptr,
              su.ptr, sv.ptr, sw.ptr, p.ptr)

    def advance_scalars(self, dt, q, sq, th, sth, nq, moisture):
        _call("dcm_advance_scalars", self.h, dt, q.ptr if q else None, sq.ptr if sq else None,
              th.ptr, sth.ptr, nq, int(moisture))


class Model:
    """dcm_model: the model_driver time step on device-resident state."""

    def __init__(self, ctx: Context, case, precision=FP64):
        self.ctx = ctx
        self.case = case
        d = case.dom
        ncat = case.ncat
        cs = (_i * max(ncat, 1))(*(case.c_s or (1,)))
        ce = (_i * max(ncat, 1))(*(case.c_e or (1,)))
        self._keep = [np.ascontiguousarray(a, dtype=np.float64)
                      for a in (d.theta, d.thetan, d.x, d.y, d.z)]
        h = _vp()
        _call("dcm_model_create", ctx.h, case.nq, ncat, cs, ce, case.dt, case.kord,
              int(case.monotone), case.advection_scheme, case.ptol,
              *[_dptr(a) for a in self._keep], C.byref(h))
        self.h = h
        self.set_precision(precision)
        if getattr(case, "physics", None) is not None:
            self.set_physics(case.physics)
        init = getattr(case, "init", None)
        if init is not None:
            # the initial state generated on the GPU (dcm_amd.init)
            rc, pr = init
            _call("dcm_model_init_fields", h, C.byref(rc.cfg), C.byref(pr["_struct"]))
            return
        self.set(F_U, case.u)
        self.set(F_V, case.v)
        self.set(F_W, case.w)
        self.set(F_TH, case.th)
        for n in range(case.nq):
            self.set(F_Q, case.q[n], n)
        _call("dcm_model_init", h)

    def set_physics(self, ph):
        """Physics options of the sources (a workloads.Physics; None: inviscid)."""
        if ph is None:
            _call("dcm_model_set_physics", self.h, None)
            return
        st, keep = physics_struct(self.case.dom, ph)
        _call("dcm_model_set_physics", self.h, C.byref(st))

    def set_precision(self, precision):
        _call("dcm_model_set_precision", self.h, precision)

    def set_inner_rtol(self, rtol):
        _call("dcm_model_set_inner_rtol", self.h, rtol)

    def set_check_residual(self, on=True):
        _call("dcm_model_set_check_residual", self.h, int(on))

    def set_graphs(self, on=True):
        """Replay the step's kernels from HIP graphs (default) or launch them
        one by one (same bits)."""
        _call("dcm_model_set_graphs", self.h, int(on))

    def set(self, field, a, index=0):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _call("dcm_model_set", self.h, field, index, _dptr(a))

    def get(self, field, index=0) -> np.ndarray:
        if field == F_THBAR:
            out = np.empty(self.case.dom.kp + 2)
        else:
            out = np.empty(self.case.dom.shape)
        _call("dcm_model_get", self.h, field, index, _dptr(out))
        return out

    def step(self, n=1) -> StepStats:
        st = StepStats()
        _call("dcm_model_step", self.h, n, C.byref(st))
        return st

    def profile(self, enable=True, only=None):
        _call("dcm_model_profile", self.h, int(enable))
        _call("dcm_model_profile_filter", self.h, only.encode() if only else None)

    def kernel_time(self, name):
        avg, n = _d(), _i()
        st = load().dcm_model_kernel_time(self.h, name.encode(), C.byref(avg), C.byref(n))
        if st != 0:
            return None, 0
        return avg.value, n.value

    def close(self):
        if self.h:
            _call("dcm_model_destroy", self.h)
            self.h = None


class Ensemble:
    """dcm_ensemble: independent models stepped together; their pressure
    solves run as one batched BiCGSTAB.  The models stay owned by the
    caller and must not be stepped on their own while in the ensemble."""

    def __init__(self, models):
        self.models = list(models)
        arr = (_vp * len(self.models))(*[m.h for m in self.models])
        h = _vp()
        _call("dcm_ensemble_create", len(self.models), arr, C.byref(h))
        self.h = h

    def autotune(self) -> int:
        planes = _i()
        _call("dcm_ensemble_autotune", self.h, C.byref(planes))
        return planes.value

    def set_graphs(self, on=True):
        """Replay the stages around the solve from captured HIP graphs (on by
        default) or launch them one by one."""
        _call("dcm_ensemble_set_graphs", self.h, 1 if on else 0)

    def step(self, n=1):
        """Advance every member n steps; one StepStats per member."""
        st = (StepStats * len(self.models))()
        _call("dcm_ensemble_step", self.h, n, st)
        return list(st)

    def close(self):
        if self.h:
            _call("dcm_ensemble_destroy", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def q_names_bytes(names, nq):
    """nq names as the reference's character(len=20) array (blank padded)."""
    names = list(names or [])
    names += [""] * (nq - len(names))
    return b"".join(n.encode()[:20].ljust(20) for n in names[:nq])


def _profiles(dom, trefn):
    n = dom.kp + 2
    t = np.zeros(n) if trefn is None else trefn
    return [np.ascontiguousarray(a, dtype=np.float64) for a in (dom.x, dom.y, dom.z, dom.rhoa,
                                                                dom.thetan, t)]


class NcFile:
    """The reference's output file (driver_code.f90:725-1145) written from
    host arrays in the padded layout: dcm_nc_create / dcm_nc_put_record.
    trefn is the reference temperature profile (kp+2; zeros if not given)."""

    def __init__(self, path, dom, moisture=False, nq=0, nprec=0, q_names=None, trefn=None):
        self.dom = dom
        self.moisture = bool(moisture)
        self._keep = _profiles(dom, trefn)
        names = q_names_bytes(q_names, nq) if moisture else None
        h = _vp()
        _call("dcm_nc_create", os.fsencode(path), dom.ip, dom.jp, dom.kp, int(self.moisture), nq,
              nprec, names, *[_dptr(a) for a in self._keep], C.byref(h))
        self.h = h

    def put_record(self, n, time, u, v, w, th, p, div, q=None, precip=None):
        """Record n (1-based); q is (nq, *shape), precip (nprec, ip+2, jp+2, kp)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                for a in (u, v, w, th, p, div, q, precip)]
        _call("dcm_nc_put_record", self.h, n, time,
              *[_dptr(a) if a is not None else None for a in arrs])

    def close(self):
        if self.h:
            _call("dcm_nc_close", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Output:
    """dcm_output: the output file written asynchronously from a resident
    model (pack on the GPU, copy on a side stream, file write on a host
    thread).  Close it before its model.  With moisture, nprec defaults to
    one precipitation field (written as NetCDF's fill value unless write()
    is given precip)."""

    def __init__(self, model, path, moisture=None, nprec=None, q_names=None, trefn=None):
        dom = model.case.dom
        nq = model.case.nq
        self.moisture = (nq > 0) if moisture is None else bool(moisture)
        if nprec is None:
            nprec = 1 if self.moisture else 0
        self._keep = _profiles(dom, trefn)
        names = q_names_bytes(q_names, nq) if self.moisture else None
        h = _vp()
        _call("dcm_output_create", model.h, os.fsencode(path), int(self.moisture), nprec, names,
              *[_dptr(a) for a in self._keep], C.byref(h))
        self.h = h

    def write(self, time, precip=None):
        pr = None if precip is None else np.ascontiguousarray(precip, dtype=np.float64)
        self._precip = pr
        _call("dcm_output_write", self.h, time, _dptr(pr) if pr is not None else None)

    def flush(self):
        _call("dcm_output_flush", self.h)

    def stats(self):
        """(records written, seconds waiting for a buffer, seconds in file writes)"""
        n, stall, wr = C.c_longlong(), _d(), _d()
        _call("dcm_output_stats", self.h, C.byref(n), C.byref(stall), C.byref(wr))
        return n.value, stall.value, wr.value

    def close(self):
        if self.h:
            _call("dcm_output_destroy", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
