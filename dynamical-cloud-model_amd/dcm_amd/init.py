"""Runs from the reference's own namelists: the namelist reader
(dcm_nml_*), the run configuration allocate_and_set reads
(initialisation.f90:117-888), its 1-D profiles (dcm_init_compute_profiles)
and a Case whose 3-D fields are generated on the GPU
(dcm_model_init_fields) instead of being uploaded."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import hip as H
from . import workloads as W

_d, _i, _dp, _vp = C.c_double, C.c_int, C.POINTER(C.c_double), C.c_void_p


class Namelist:
    """A Fortran namelist file, read by the library's reader."""

    def __init__(self, path):
        h = _vp()
        H._call("dcm_nml_read", os.fsencode(path), C.byref(h))
        self.h = h

    def _get(self, fn, group, name, box):
        return getattr(H.load(), fn)(self.h, group.encode(), name.encode(), C.byref(box)) == 0

    def has(self, group, name=None) -> bool:
        return bool(H.load().dcm_nml_has(self.h, group.encode(),
                                       name.encode() if name else None))

    def real(self, group, name, default=None):
        x = _d()
        return x.value if self._get("dcm_nml_get_real", group, name, x) else default

    def int(self, group, name, default=None):
        x = _i()
        return x.value if self._get("dcm_nml_get_int", group, name, x) else default

    def logical(self, group, name, default=None):
        x = _i()
        return bool(x.value) if self._get("dcm_nml_get_logical", group, name, x) else default

    def string(self, group, name, default=None):
        buf = C.create_string_buffer(1024)
        if H.load().dcm_nml_get_string(self.h, group.encode(), name.encode(), buf, 1024) != 0:
            return default
        return buf.value.decode()

    def array(self, group, name, shape, fill=0.0):
        """The object as a Fortran array of `shape` (returned in Fortran
        order, so a[i-1, j-1] is name(i, j)); untouched elements = fill."""
        shape = tuple(int(s) for s in shape)
        out = np.full(int(np.prod(shape)), float(fill))
        sh = (_i * len(shape))(*shape)
        H._call("dcm_nml_get_array", self.h, group.encode(), name.encode(), len(shape), sh,
                out.ctypes.data_as(_dp))
        return out.reshape(shape, order="F")

    def close(self):
        if self.h:
            H.load().dcm_nml_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class RunConfig:
    """The run configuration of a namelist file (&run_vars, &sounding_vars):
    `nm` holds every nm1 component the dynamical core and allocate_and_set
    use, `cfg` the dcm_init_config pointing at this object's arrays.
    iqv/iqc/inc are the q fields of vapour, cloud mass and cloud number (the
    reference takes them from its microphysics set-up)."""

    # the defaults of variables.f90:59-93 (components without one read as 0)
    DEFAULTS = {"ptol": 1e-8, "adiabatic_prof": False, "param_wind": False, "param_theta": True,
                "bubble": False}
    REALS = ("dt", "runtime", "dx", "dy", "dz", "psurf", "tsurf", "adiabatic_frac", "t_cbase",
             "t_ctop", "rh_above", "th_jump", "th_grad", "damping_thickness", "damping_tau",
             "forcing_tau", "divergence_val", "divergence_hgt", "cvis", "z0", "z0th",
             "param_vmax", "param_z", "param_sigz", "param_delz", "ptol", "drop_num",
             "output_interval")
    INTS = ("ip", "jp", "kp", "nq", "nprec", "n_levels", "kord", "advection_scheme",
            "subgrid_model", "microphysics_flag")
    LOGICALS = ("moisture", "theta_flag", "adiabatic_prof", "damping_layer", "forcing",
                "divergence", "viscous_dissipation", "monotone", "bubble", "param_theta",
                "param_wind", "radiation")

    def __init__(self, path, iqv=1, iqc=2, inc=3, seed=2):
        nm = {}
        with Namelist(path) as nl:
            for n in self.REALS:
                nm[n] = nl.real("run_vars", f"nm1%{n}", self.DEFAULTS.get(n, 0.0))
            for n in self.INTS:
                nm[n] = nl.int("run_vars", f"nm1%{n}", 0)
            for n in self.LOGICALS:
                nm[n] = nl.logical("run_vars", f"nm1%{n}", self.DEFAULTS.get(n, False))
            nm["outputfile"] = nl.string("run_vars", "nm1%outputfile", "output")
            nl_lev = nm["n_levels"]
            nq = max(nm["nq"], 1)
            self.z_read = np.ascontiguousarray(nl.array("sounding_vars", "z_read", (nl_lev,)))
            self.theta_read = np.ascontiguousarray(
                nl.array("sounding_vars", "theta_read", (nl_lev,)))
            self.u_read = np.ascontiguousarray(nl.array("sounding_vars", "u_read", (nl_lev,)))
            self.v_read = np.ascontiguousarray(nl.array("sounding_vars", "v_read", (nl_lev,)))
            qr = nl.array("sounding_vars", "q_read", (nq, nl_lev))
            self.q_read = np.ascontiguousarray(qr.ravel(order="F"))  # Fortran order
        self.nm = nm
        self.path = os.fspath(path)
        p = lambda a: a.ctypes.data_as(_dp)
        self.cfg = H.InitConfig(
            nm["ip"], nm["jp"], nm["kp"], nm["dx"], nm["dy"], nm["dz"], nm["dt"], nm["runtime"],
            int(nm["moisture"]), int(nm["theta_flag"]), nm["nq"] if nm["moisture"] else 0,
            iqv, iqc, inc, nm["drop_num"], nm["psurf"], nm["tsurf"], int(nm["adiabatic_prof"]),
            nm["adiabatic_frac"], nm["t_cbase"], nm["t_ctop"], nm["rh_above"], nm["th_jump"],
            nm["th_grad"], nm["n_levels"], p(self.z_read), p(self.theta_read), p(self.u_read),
            p(self.v_read), p(self.q_read), int(nm["damping_layer"]), nm["damping_thickness"],
            nm["damping_tau"], int(nm["forcing"]), nm["forcing_tau"], int(nm["divergence"]),
            nm["divergence_val"], nm["divergence_hgt"], nm["cvis"], nm["z0"], nm["z0th"],
            int(nm["bubble"]), int(nm["param_theta"]), int(nm["param_wind"]), nm["param_vmax"],
            nm["param_z"], nm["param_sigz"], nm["param_delz"], seed)


def compute_profiles(rc: RunConfig) -> dict:
    """dcm_init_compute_profiles: the 1-D arrays of allocate_and_set."""
    c = rc.cfg
    nk = c.kp + 2
    sizes = {"x": c.ip + 2, "xn": c.ip + 2, "dx": c.ip + 2, "dxn": c.ip + 2,
             "y": c.jp + 2, "yn": c.jp + 2, "dy": c.jp + 2, "dyn": c.jp + 2,
             "q": max(c.nq, 1) * nk}
    out = {n: np.zeros(sizes.get(n, nk)) for n in H.PROFILE_ARRAYS}
    prof = H.InitProfiles(*[out[n].ctypes.data_as(_dp) for n in H.PROFILE_ARRAYS], 0.0, 0.0, 0)
    H._call("dcm_init_compute_profiles", C.byref(c), C.byref(prof))
    out["q"] = out["q"].reshape(max(c.nq, 1), nk)
    out["thbase"], out["thtop"], out["ntim"] = prof.thbase, prof.thtop, prof.ntim
    out["_struct"] = prof
    return out


def case_from_namelist(path, steps=None, overrides=None, **kw) -> W.Case:
    """A Case for the resident model from a namelist file: grid and base
    state from the profiles, the options of sources from &run_vars, and
    the initial fields left to the GPU (Case.init).  `overrides` replaces
    scalar settings of the file (e.g. {"jp": 32, "kp": 40})."""
    rc = RunConfig(path, **kw)
    nm = rc.nm
    for key, val in (overrides or {}).items():
        nm[key] = val
        if any(key == f[0] for f in H.InitConfig._fields_):
            setattr(rc.cfg, key, int(val) if isinstance(val, bool) else val)
    if nm["advection_scheme"] not in (0, 1, 2):
        raise NotImplementedError(f"advection_scheme {nm['advection_scheme']} is not coded "
                                  "(the reference stops with 'not coded' too)")
    if nm["radiation"]:
        raise NotImplementedError("radiation is not on the GPU path")
    pr = compute_profiles(rc)
    dom = W.Domain(nm["ip"], nm["jp"], nm["kp"], pr["dx"], pr["dy"], pr["dz"], pr["dxn"],
                   pr["dyn"], pr["dzn"], pr["x"], pr["y"], pr["z"], pr["xn"], pr["yn"],
                   pr["zn"], pr["rhoa"], pr["rhoan"], pr["theta"], pr["thetan"])
    nq = rc.cfg.nq
    cats = tuple(range(1, nq + 1))
    physics = None
    if nm["viscous_dissipation"] or nm["damping_layer"] or nm["forcing"] or nm["divergence"]:
        physics = W.Physics(bool(nm["viscous_dissipation"]), bool(nm["damping_layer"]),
                            bool(nm["forcing"]), bool(nm["divergence"]), nm["forcing_tau"],
                            nm["z0"], nm["z0th"], pr["lamsq"], pr["dampfacn"], pr["dampfac"],
                            pr["u_force"], pr["v_force"], pr["w_subs"])
    case = W.Case(os.path.basename(rc.path), dom, nm["dt"],
                  steps if steps is not None else pr["ntim"], kord=nm["kord"],
                  monotone=bool(nm["monotone"]), ptol=nm["ptol"],
                  advection_scheme=nm["advection_scheme"], nq=nq, c_s=cats, c_e=cats,
                  physics=physics)
    case.init = (rc, pr)
    case.trefn = pr["trefn"]
    return case


