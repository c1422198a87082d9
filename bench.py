#!/usr/bin/env python3
"""Benchmark of the dynamical-core time step on one MI355X.

Workload: the moist bubble with tracers, 256x256x128 cells at 50 m, dt=0.5 s,
theta + qv + 8 passive tracers (F=10 advected fields, each its own SFVT
category), MPDATA kord=2 with the limiter, BiCGSTAB to a 1e-10 relative
residual.  Inputs are the seeded synthetic fields of
``dcm_amd.workloads.moist_bubble``; they are resident in HBM before the timed
region starts.

A step is one pass of the model_driver time loop (driver_code.f90:304-676):
horizontal means, sources, pressure solve, reference-state advection, MPDATA
of every field, momentum update, time filter and halo fills.  Both paths run
W warm-up steps, then K timed steps from the same state:

* FP64  -- the reference arithmetic (parity path);
* mixed -- FP32 MPDATA intermediates and FP32 inner BiCGSTAB inside FP64
  iterative refinement, every solve ending at FP64 true residual < 1e-10;
  the inner solve is preconditioned by one multigrid V-cycle per
  application (dcm_set_precond_multigrid).  The same mixed path with the
  Chebyshev-relaxed and with the reference's plain Jacobi sweeps is
  reported beside it (paths.mixed_chebyshev, paths.mixed_plain_jacobi).

``value`` is the mixed path's cell-steps/s (interior cells x timed steps /
wall seconds).  The ``roofline`` object describes the kernel with the
largest time per step of that path, chosen among its candidates (the three
FP32 MPDATA passes, each serving all ten fields in one launch, and the
multigrid's finest-level sweep), each timed with HIP events on the model's
stream in two further steps after the timed ones.  Its algorithmic bytes
are the canonical stage bytes of the stages the kernel performs (the last
MPDATA kernel performs two stages, the limiter and the last donor pass, so
fusion can put its canonical figure above what it moves; ``traffic`` is what
it moves, from the committed rocprofv3 counters).

Further legs: the FP64 path with the same V-cycle (precision gain at an
equal preconditioner), the Kelvin-Helmholtz shear layer (512x4x256, mixed
against FP64 error), the large domain (512x512x256, HBM footprint), the
64-member ensemble, and the options of sources, vertical MPDATA, the
sfvt/advection_3d.F90 arithmetic, initialisation and output.  The CPU
baseline is the oracle (a C restatement of the reference) on the host's
cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dynamical-cloud-model_amd"))

METRIC = "cell-steps/s mixed vs FP64 at 1e-10 pressure resid; % HBM roofline"
HBM_PEAK_GBPS = 8000.0   # MI355X HBM3E
DOMINANT = "jac"         # Jacobi sweep kernel k_jac<R,false>: s' = s + (t - A s)/a_p
JAC_WORDS = 3            # reads s and t, writes s'
# candidates for the dominant kernel of the mixed path: (timer name, kernel,
# canonical FP32-path bytes per cell and field of the stages it performs)
MPDATA32 = (("k32_donor1", "k32_donor1 (first donor pass, every field)", 24),
            ("k32_antidiff", "k32_antidiff (antidiffusive velocities, every field)", 28),
            ("k32_betalim", "k32_betalim (limiter factors + limited donor pass, every field)",
             64))
# the headline mixed path preconditions its FP32 inner solve with a multigrid
# V-cycle whose finest-level sweeps are the same k_jac<float,false,true>
# (weight 0.8), five of them per inner iteration ("jac"; the sixth updates p)
MG_DOMINANT = DOMINANT
MG_SWEEPS_PER_ITER = 5
MG_WORDS_PER_ITER = 51.7 # FP32 words per cell of one inner iteration (canonical)
PROFILE_STEPS = 2        # steps whose dominant-kernel launches are timed
# the mixed inner solve's preconditioner: Chebyshev factors on [CHEB_A, 2]
# with CHEB_DIVISIONS divisions by a_p (the library's default, set explicitly)
CHEB_A, CHEB_DIVISIONS = 0.002, 12


def jac_per_iter(divisions):
    """k_jac<R,false> launches per BiCGSTAB iteration: the middle sweeps of
    both preconditioner applications and the last sweep of the second."""
    return 2 * (divisions - 3) + 1


def canonical_bytes_fp64(F, iters, scalar_update=False):
    """Canonical stage model, FP64 words per cell-step: 39 + 30F + 37*it
    (the scalar update's 3 words per field only when the model runs it:
    without the options of sources it is an exact no-op and skipped)."""
    return 8.0 * (39 + (30 if scalar_update else 27) * F + 37 * iters)


def canonical_bytes_mixed(F, inner_iters, outer, divisions=4, multigrid=False,
                          scalar_update=False):
    """Same stages with each word at its storage precision (bytes/cell-step).

    FP64: state (u,v,w,theta,q,p,b,x), sources, RHS, the true residual and
    the correction of each refinement sweep.  FP32: MPDATA intermediates, the
    velocity shadows and the inner solver vectors.
    """
    b = 8 + 56 + 32                      # thbar, sources, Poisson RHS
    b += outer * (20 + 24 + 20)          # true residual, inner start, x += d
    # 37 FP32 words per inner iteration with the reference's four divisions,
    # each further division one more sweep (3 words) in both applications
    if multigrid:
        # two preparations (4 words each, the first sweep included), two
        # V-cycles (three sweeps of 3 words, the restriction's 2 reads and the
        # prolongation's read and write on the finest level: 13 words, 8/7
        # of that over all levels), two matvec-dots (6), the update (8)
        b += 4 * MG_WORDS_PER_ITER * inner_iters
    else:
        b += 4 * (37 + 6 * (divisions - 4)) * inner_iters
    b += 24 + 36                         # adv_ref_state, FP32 velocity shadows
    b += F * (8 + 24 + 28 + 32 + 32)     # MPDATA per field (min, 4 passes)
    if scalar_update:
        b += F * 24                      # scalar update (skipped without sources options)
    b += 80 + 72                         # advance_momentum, time filter
    return float(b)


def run_path(H, case, precision, warmup, steps, inner_rtol=None, precond="chebyshev",
             candidates=()):
    """precond: "multigrid" (one V-cycle per application), "chebyshev"
    (Chebyshev-relaxed Jacobi sweeps) or "plain" (the reference's four
    Jacobi sweeps).  The profiled steps time the finest-level sweep and
    every kernel named in ``candidates``."""
    ctx = H.Context(case.dom)
    profile_only = DOMINANT
    if precond == "plain":
        ctx.set_precond_chebyshev(0.0)  # the reference's four plain Jacobi sweeps
    else:
        ctx.set_precond_chebyshev(CHEB_A, CHEB_DIVISIONS)
    if precond == "multigrid":
        ctx.set_precond_multigrid(True)
        profile_only = MG_DOMINANT
    model = H.Model(ctx, case, precision)
    model.set_check_residual(True)
    if inner_rtol:
        model.set_inner_rtol(inner_rtol)
    if warmup:
        model.step(warmup)
    t0 = time.perf_counter()
    st = model.step(steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    # the dominant kernel's launches timed with HIP events in passes of their
    # own after the timed steps (events on every launch would slow them)
    names = [profile_only] + [n for n in candidates if n != profile_only]
    model.profile(True, only=",".join(names))
    sp = model.step(PROFILE_STEPS)
    timed = {}
    for nm in names:
        a, n = model.kernel_time(nm)
        timed[nm] = {"avg_ms": a, "launches": n,
                     "ms_per_step": a * n / PROFILE_STEPS if a else None}
    avg_ms, launches = timed[profile_only]["avg_ms"], timed[profile_only]["launches"]
    model.profile(False)
    cells = case.dom.cells
    snap = {"th": model.get(H.F_TH), "u": model.get(H.F_U), "w": model.get(H.F_W),
            "p": model.get(H.F_P), "q0": model.get(H.F_Q, 0), "q1": model.get(H.F_Q, 1)}
    res = {
        "value": cells * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "iters_per_step": st.iters_total / steps,
        "iters_total": st.iters_total,
        "fallbacks": st.fallbacks,
        "iters_max": st.iters_max,
        "outer_per_step": st.outer_total / steps,
        "max_true_resid": st.max_true_resid,
        "kernel": {"name": profile_only, "avg_ms": avg_ms, "launches": launches,
                   "profiled_steps": PROFILE_STEPS, "profiled_iters": sp.iters_total},
        "kernels": timed,
    }
    return res, snap, (ctx, model)


def run_ensemble(H, W, warmup, steps, members=64, seed=404, profile=False):
    """The uneven warm-bubble ensemble three ways: every member's pressure
    solve batched into one mixed-precision BiCGSTAB per step with the
    Chebyshev-relaxed sweeps (after the autotuner picks the solver tile; the
    stages around the solve replayed from HIP graphs); every member with its
    own multigrid-preconditioned mixed solve, the members' solves running
    concurrently on the ensemble's side streams; and the members one after
    another on the plain FP64 path.  Cell-steps/s over the whole ensemble.

    With ``profile`` the batched Jacobi sweeps are timed with HIP events
    (which switches the graphs off) for the roofline: algorithmic bytes are
    12 B x the plain sweeps per inner iteration (jac_per_iter) x the cells of
    each member that iterated, summed over members, against the summed
    launch durations."""
    cases = W.ensemble_members(members, seed=seed, steps=warmup + steps)
    cells = sum(c.dom.cells for c in cases)
    out = {"members": len(cases), "cells": cells,
           "shapes": sorted({f"{c.dom.ip}x{c.dom.jp}x{c.dom.kp}" for c in cases})}
    ctxs = [H.Context(c.dom) for c in cases]
    for x in ctxs:
        x.set_precond_chebyshev(CHEB_A, CHEB_DIVISIONS)
    models = [H.Model(x, c, H.MIXED) for x, c in zip(ctxs, cases)]
    try:
        with H.Ensemble(models) as e:
            t0 = time.perf_counter()
            planes = e.autotune()
            tune_s = time.perf_counter() - t0
            if warmup:
                e.step(warmup)
            if profile:
                models[0].profile(True, only=DOMINANT)
            t0 = time.perf_counter()
            stats = e.step(steps)
            wall = time.perf_counter() - t0
            kernel = None
            if profile:
                avg_ms, launches = models[0].kernel_time(DOMINANT)
                models[0].profile(False)
                swept = sum(c.dom.cells * s.iters_total for c, s in zip(cases, stats))
                kernel = {"name": "k_jac_b<float,false,true>", "avg_ms": avg_ms,
                          "launches": launches,
                          "algorithmic_bytes": JAC_WORDS * 4 * jac_per_iter(CHEB_DIVISIONS) *
                          swept}
        out["batched_mixed"] = {
            "value": cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "autotuned_march_planes": planes, "autotune_s": tune_s,
            "graphs": not profile,
            "iters_total": sum(s.iters_total for s in stats),
            "iters_max_member": max(s.iters_max for s in stats),
            "fallbacks": sum(s.fallbacks for s in stats),
            "max_true_resid": max(s.max_true_resid for s in stats),
            "kernel": kernel}
    finally:
        for m in models:
            m.close()
        for x in ctxs:
            x.close()
    if not profile:
        # the same ensemble with every member on the multigrid preconditioner:
        # graphs around the solve, each member's V-cycle-preconditioned solve
        ctxs = [H.Context(c.dom) for c in cases]
        models = []
        try:
            for x in ctxs:
                x.set_precond_multigrid(True)
            models = [H.Model(x, c, H.MIXED) for x, c in zip(ctxs, cases)]
            with H.Ensemble(models) as e:
                if warmup:
                    e.step(warmup)
                t0 = time.perf_counter()
                stats = e.step(steps)
                wall = time.perf_counter() - t0
            out["multigrid_mixed"] = {
                "value": cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
                "iters_total": sum(s.iters_total for s in stats),
                "fallbacks": sum(s.fallbacks for s in stats),
                "max_true_resid": max(s.max_true_resid for s in stats)}
        finally:
            for m in models:
                m.close()
            for x in ctxs:
                x.close()
    wall = 0.0
    iters = 0
    resid = 0.0
    for c in cases:
        with H.Context(c.dom) as x:
            m = H.Model(x, c, H.FP64)
            try:
                m.set_check_residual(True)
                if warmup:
                    m.step(warmup)
                t0 = time.perf_counter()
                st = m.step(steps)
                wall += time.perf_counter() - t0
                iters += st.iters_total
                resid = max(resid, st.max_true_resid)
            finally:
                m.close()
    out["sequential_fp64"] = {"value": cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
                              "iters_total": iters, "max_true_resid": resid}
    out["speedup_batched_mixed_vs_sequential_fp64"] = (out["batched_mixed"]["value"] /
                                                       out["sequential_fp64"]["value"])
    if "multigrid_mixed" in out:
        out["speedup_multigrid_mixed_vs_sequential_fp64"] = (out["multigrid_mixed"]["value"] /
                                                             out["sequential_fp64"]["value"])
        out["best_mixed"] = max(("batched_mixed", "multigrid_mixed"),
                                key=lambda k: out[k]["value"])
    return out


def sedimentation(H, W, calls=20, cpu=True):
    """PAMM's sedimentation call (pamm/microphysics.f90:1522-1530):
    mpdata_vec_vert_3d of a three-field category falling at up to 8 m/s,
    kord 1, no limiter, neumann 2, on the 256x256x128 grid.  Wall time per
    call through the C-ABI (each call returns lbc/ubc to the host, as the
    Fortran interface does).  Canonical bytes per cell: read the three
    fields and w, write the three fields (7 words); the positivity shift and
    the zero-sum test each read the fields once more in this implementation
    (minimum per field, sum of field 1), and the pass writes to scratch
    before the result is copied back: 6 words per field plus w."""
    dom = W.make_domain(256, 256, 128, 50.0, 50.0, 50.0)
    nq = 3
    rng = np.random.default_rng(606)
    psi = np.zeros((nq,) + dom.shape)
    psi[:, 1:-1, 1:-1, 1:-1] = 1e-3 * rng.random((nq, 256, 256, 128))
    vq = np.zeros(dom.shape)
    vq[1:-1, 1:-1, 0:-1] = -8.0 * rng.random((256, 256, 129))
    vq[:, :, -2] = 0.0
    cells = dom.cells
    with H.Context(dom) as c:
        dw, dpsi = c.array(vq), c.array(psi)
        c.mpdata_vec_vert3d(0.5, dw, dpsi, nq, [0.0] * nq, [0.0] * nq, 1, False, 2)
        t0 = time.perf_counter()
        for _ in range(calls):
            c.mpdata_vec_vert3d(0.5, dw, dpsi, nq, [0.0] * nq, [0.0] * nq, 1, False, 2)
        wall = (time.perf_counter() - t0) / calls
    canonical = 8.0 * (2 * nq + 1) * cells
    moved = 8.0 * (6 * nq + nq) * cells
    out = {"call": "mpdata_vec_vert_3d, 3 fields, kord 1, no limiter, neumann 2",
           "grid": "256x256x128", "ms_per_call": 1e3 * wall,
           "cell_steps_per_s": cells / wall,
           "canonical_GBps": canonical / wall / 1e9, "moved_GBps": moved / wall / 1e9}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        O.build()
        sd = W.make_domain(64, 64, 128, 50.0, 50.0, 50.0)
        g = O.Grid(sd)
        sp = np.zeros((nq,) + sd.shape)
        sp[:, 1:-1, 1:-1, 1:-1] = 1e-3 * rng.random((nq, 64, 64, 128))
        sw = np.zeros(sd.shape)
        sw[1:-1, 1:-1, 0:-1] = -8.0 * rng.random((64, 64, 129))
        t0 = time.perf_counter()
        n = 5
        for _ in range(n):
            O.mpdata_vec_vert3d(g, 0.5, sw, sp, np.zeros(nq), np.zeros(nq), 1, False, 2)
        cw = (time.perf_counter() - t0) / n
        out["cpu_baseline"] = {"value": sd.cells / cw, "unit": "cell-steps/s", "cores": 1,
                               "kind": "port", "sample": f"oracle, same call on 64x64x128, "
                                                         f"{n} calls, {n * cw:.2f} s"}
    return out


def mpdata_variants(H, W, calls=10, model_steps=6, cpu=True):
    """The two copies of advection_3d.F90 side by side on the 256x256x128
    moist-bubble grid: mpdata_3d of theta (kord 2, limiter) and
    mpdata_vec_3d of a three-field category, FP64 through the C-ABI, with
    the linked copy's arithmetic and with sfvt's (per-direction limiter:
    six limiter fields instead of two).  Canonical bytes per cell: read the
    fields and u, v, w, write the fields.  Also the mixed model step under
    sfvt's arithmetic."""
    case = W.moist_bubble(256, 256, 128, steps=model_steps)
    dom = case.dom
    rng = np.random.default_rng(707)
    th = np.zeros(dom.shape)
    th[1:-1, 1:-1, 1:-1] = rng.random((256, 256, 128))
    q = np.zeros((3,) + dom.shape)
    q[:, 1:-1, 1:-1, 1:-1] = 1e-3 * rng.random((3, 256, 256, 128))
    vel = [np.zeros(dom.shape) for _ in range(3)]
    for a, d in zip(vel, (dom.dx[1], dom.dy[1], dom.dz[1])):
        a[1:-1, 1:-1, 1:-1] = 0.6 * d * (rng.random((256, 256, 128)) - 0.5)
    cells = dom.cells
    out = {"grid": "256x256x128", "kord": 2, "monotone": True}
    for name, variant in (("linked", H.MPDATA_LINKED), ("sfvt", H.MPDATA_SFVT)):
        with H.Context(dom) as c:
            c.set_mpdata_variant(variant)
            du, dv, dw = (c.array(a) for a in vel)
            dth, dq = c.array(th), c.array(q)
            res = {}
            for call, fn, nf in (("mpdata_3d", lambda: c.mpdata3d(1.0, du, dv, dw, dth), 1),
                                 ("mpdata_vec_3d, 3 fields",
                                  lambda: c.mpdata_vec3d(1.0, du, dv, dw, dq, 3), 3)):
                fn()
                c.sync()
                t0 = time.perf_counter()
                for _ in range(calls):
                    fn()
                c.sync()
                wall = (time.perf_counter() - t0) / calls
                res[call] = {"ms_per_call": 1e3 * wall, "cell_steps_per_s": cells / wall,
                             "canonical_GBps": 8.0 * (2 * nf + 3) * cells / wall / 1e9}
            out[name] = res
    for call in out["linked"]:
        out["sfvt"][call]["time_vs_linked"] = (out["sfvt"][call]["ms_per_call"] /
                                               out["linked"][call]["ms_per_call"])
    with H.Context(dom) as c:
        c.set_mpdata_variant(H.MPDATA_SFVT)
        c.set_precond_chebyshev(CHEB_A, CHEB_DIVISIONS)
        c.set_precond_multigrid(True)
        m = H.Model(c, case, H.MIXED)
        try:
            m.step(2)
            c.sync()
            t0 = time.perf_counter()
            st = m.step(model_steps)
            c.sync()
            wall = (time.perf_counter() - t0) / model_steps
        finally:
            m.close()
    out["sfvt_mixed_model"] = {"ms_per_step": 1e3 * wall, "value": cells / wall,
                               "unit": "cell-steps/s", "max_true_resid": st.max_true_resid}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        O.build()
        O.set_mpdata_variant(1)
        try:
            sd = W.make_domain(64, 64, 64, 50.0, 50.0, 50.0)
            g = O.Grid(sd)
            sp = np.zeros(sd.shape)
            sp[1:-1, 1:-1, 1:-1] = rng.random((64, 64, 64))
            sv = [np.zeros(sd.shape) for _ in range(3)]
            for a in sv:
                a[1:-1, 1:-1, 1:-1] = 0.6 * 50.0 * (rng.random((64, 64, 64)) - 0.5)
            n = 5
            t0 = time.perf_counter()
            for _ in range(n):
                O.mpdata3d(g, 1.0, sv[0], sv[1], sv[2], sp)
            cw = (time.perf_counter() - t0) / n
        finally:
            O.set_mpdata_variant(0)
        out["cpu_baseline"] = {"value": sd.cells / cw, "unit": "cell-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle, sfvt mpdata_3d on 64x64x64, {n} calls, "
                                         f"{n * cw:.2f} s"}
    return out


def physics(H, W, steps, warmup=2, base=None, cpu=True):
    """The sheared moist bubble with every option of sources but radiation
    (subgrid viscosity and surface layer, damping layer, forcing,
    subsidence) on the 256x256x128 grid: model steps in both precisions, and
    the subgrid stage alone (calculate_subgrid_3d, sgm/subgrid_3d.f90:38-357)
    timed through dcm_subgrid.  Canonical bytes of the subgrid stage per
    cell: read u, v, w, zu, zv, zw, th and the nq fields of q; read and
    write su, sv, sw, sth and the nq fields of sq; write strain, vism, vist
    and the nq fields of viss: 18 + 4 nq words."""
    case = W.sheared_moist_bubble(256, 256, 128)
    cells = case.dom.cells
    out = {"workload": case.name, "options": "viscous, damping_layer, forcing, divergence",
           "fields": 1 + case.nq}
    for pth, prec in (("mixed", H.MIXED), ("fp64", H.FP64)):
        res, _, handles = run_path(H, case, prec, warmup, steps, precond="multigrid")
        handles[1].close()
        handles[0].close()
        out[pth] = {k: res[k] for k in ("value", "ms_per_step", "iters_per_step",
                                        "max_true_resid")}
        if base and pth in base:
            out[pth]["vs_inviscid_step"] = res["ms_per_step"] / base[pth]["ms_per_step"]
        print(f"[bench] physics {pth}: {res['value']:.4g} cell-steps/s, "
              f"{res['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    nq = case.nq
    dom = case.dom
    with H.Context(dom) as c:
        flds = {n: c.array(a) for n, a in (("u", case.u), ("v", case.v), ("w", case.w),
                                           ("th", case.th))}
        for n in ("zu", "zv", "zw"):
            flds[n] = c.array(flds[n[1]].get())
        q = c.array(case.q)
        one = [c.zeros() for _ in range(7)]
        many = [c.zeros((nq,) + dom.shape) for _ in range(2)]
        vism, vist, strain, su, sv, sw, sth = one
        viss, sq = many
        args = (case.physics, flds["u"], flds["v"], flds["w"], flds["zu"], flds["zv"],
                flds["zw"], flds["th"], q, nq, vism, vist, viss, strain, su, sv, sw, sth, sq)
        c.subgrid(*args)
        c.sync()
        calls = 10
        t0 = time.perf_counter()
        for _ in range(calls):
            c.subgrid(*args)
        c.sync()
        wall = (time.perf_counter() - t0) / calls
    canonical = 8.0 * (18 + 4 * nq) * cells
    out["subgrid"] = {"call": f"dcm_subgrid, nq {nq}", "ms_per_call": 1e3 * wall,
                      "cells_per_s": cells / wall,
                      "canonical_GBps": canonical / wall / 1e9,
                      "frac_of_8TBps": canonical / wall / 8e12}
    if cpu:
        out["subgrid"]["cpu_baseline"] = subgrid_cpu(W, nq)
    return out


def subgrid_cpu(W, nq):
    """The oracle's calculate_subgrid_3d (single thread) on a 64x64x128 sample
    of the same case, one call."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    case = W.sheared_moist_bubble(64, 64, 128, ntracer=nq - 1)
    dom = case.dom
    g = O.Grid(dom)
    ph = O.Physics(dom, case.physics, nq=nq)
    z = lambda *sh: np.zeros(sh if sh else dom.shape)
    vism, vist, strain, su, sv, sw, sth = (z() for _ in range(7))
    viss, sq = z(nq, *dom.shape), z(nq, *dom.shape)
    t0 = time.perf_counter()
    O.subgrid(g, ph, case.u, case.v, case.w, case.u, case.v, case.w, case.th, case.q, vism, vist,
              viss, strain, su, sv, sw, sth, sq)
    wall = time.perf_counter() - t0
    return {"value": dom.cells / wall, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"oracle subgrid model, sheared moist bubble 64x64x128, nq {nq}, one "
                      f"call, {wall:.2f} s"}


def output(H, W, steps=20, every=5):
    """The output file (driver_code.f90:725-1145) from the resident mixed
    model on the 256x256x128 moist bubble: two identical models step the
    same `steps` steps (so the solver does the same work), one writing a
    record every `every` steps; plus the pack kernel's time (a snapshot of
    15 fields into one record in HBM: read 8 and write 4 bytes per cell and
    field) and the writer thread's file rate.  The file goes to a temporary
    directory that is removed afterwards."""
    import shutil
    import tempfile
    case = W.moist_bubble(256, 256, 128)
    cells, nf = case.dom.cells, 6 + case.nq
    tmp = tempfile.mkdtemp(prefix="dcm_out_")

    def timed(with_output):
        with H.Context(case.dom) as ctx:
            m = H.Model(ctx, case, H.MIXED)
            res = {}
            try:
                m.step(2)
                if not with_output:
                    t0 = time.perf_counter()
                    st = m.step(steps)
                    res["loop"] = time.perf_counter() - t0
                    res["iters"] = st.iters_total
                    return res
                with H.Output(m, os.path.join(tmp, "out.nc")) as out:
                    out.write(0.0)  # first record: buffers and file warm
                    out.flush()
                    ctx.sync()
                    t0 = time.perf_counter()
                    out.write(0.0)
                    ctx.sync()  # the pack kernel has run (the copy and write go on)
                    res["pack"] = time.perf_counter() - t0
                    out.flush()
                    n0, stall0, wr0 = out.stats()
                    iters = 0
                    t0 = time.perf_counter()
                    for s in range(steps):
                        iters += m.step(1).iters_total
                        if (s + 1) % every == 0:
                            out.write(float(s + 1))
                    res["loop"] = time.perf_counter() - t0
                    out.flush()
                    res["total"] = time.perf_counter() - t0
                    n1, stall1, wr1 = out.stats()
                    res.update(iters=iters, recs=n1 - n0, stall=stall1 - stall0, wr=wr1 - wr0)
                return res
            finally:
                m.close()

    try:
        plain = timed(False)
        wout = timed(True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    rec_bytes = 4.0 * nf * cells
    recs = wout["recs"]
    # the host alternative: the library's host writer packing one record
    # from FP64 host arrays of the same state and writing it (dcm_nc_*)
    tmp2 = tempfile.mkdtemp(prefix="dcm_out_")
    try:
        f = [case.u, case.v, case.w, case.th, np.zeros(case.dom.shape), np.zeros(case.dom.shape)]
        with H.NcFile(os.path.join(tmp2, "host.nc"), case.dom, moisture=True, nq=case.nq,
                      nprec=1) as nc:
            t0 = time.perf_counter()
            nc.put_record(1, 0.0, *f, q=case.q)
            t_host = time.perf_counter() - t0
    finally:
        shutil.rmtree(tmp2, ignore_errors=True)
    return {"workload": "moist_bubble_256x256x128, mixed", "record_MB": rec_bytes / 1e6,
            "records": recs, "every_steps": every, "steps": steps,
            "solver_iterations_equal": plain["iters"] == wout["iters"],
            "ms_per_step_without_output": 1e3 * plain["loop"] / steps,
            "ms_per_step_with_output": 1e3 * wout["loop"] / steps,
            "ms_per_step_including_final_flush": 1e3 * wout["total"] / steps,
            "host_stall_s": wout["stall"],
            "pack_ms": 1e3 * wout["pack"], "pack_GBps": 12.0 * nf * cells / wout["pack"] / 1e9,
            "writer_s_per_record": wout["wr"] / max(recs, 1),
            "writer_GBps": rec_bytes * recs / max(wout["wr"], 1e-9) / 1e9,
            "cpu_baseline": {"value": 1e3 * t_host, "unit": "ms per record", "cores": 1,
                             "kind": "port",
                             "sample": "one record packed from FP64 host arrays and written "
                                       "by the library's host writer (dcm_nc_put_record)"}}


def initialisation(H, W):
    """A run set up from the reference's namelist.in on the 256x256x128 grid
    (dcm_amd.init): parsing and the 1-D profiles on the host, then the 3-D
    initial state generated in HBM (dcm_model_init_fields), against the same
    model created from host arrays of the same state (uploads)."""
    import dataclasses
    from dcm_amd import init as I
    path = os.path.join(ROOT, "tests", "golden", "namelists", "namelist.in")
    over = {"ip": 256, "jp": 256, "kp": 128}
    t0 = time.perf_counter()
    case = I.case_from_namelist(path, steps=1, overrides=over)
    t_prof = time.perf_counter() - t0
    with H.Context(case.dom) as ctx:
        ctx.sync()
        t0 = time.perf_counter()
        m = H.Model(ctx, case, H.MIXED)
        ctx.sync()
        t_dev = time.perf_counter() - t0
        fields = {"u": m.get(H.F_U), "v": m.get(H.F_V), "w": m.get(H.F_W),
                  "th": m.get(H.F_TH),
                  "q": np.stack([m.get(H.F_Q, n) for n in range(case.nq)])}
        m.close()
    host = dataclasses.replace(case, **fields)
    host.init = None
    with H.Context(case.dom) as ctx:
        ctx.sync()
        t0 = time.perf_counter()
        m = H.Model(ctx, host, H.MIXED)
        ctx.sync()
        t_host = time.perf_counter() - t0
        m.close()
    nf = 4 + case.nq
    # the host alternative: the same 3-D state built with numpy from the
    # profiles (columns broadcast, the noise layers), before any upload
    rc, pr = case.init
    t0 = time.perf_counter()
    shape = case.dom.shape
    hu = np.broadcast_to(pr["u"], shape).copy()
    hv = np.broadcast_to(pr["v"], shape).copy()
    hw = np.zeros(shape)
    hq = np.stack([np.broadcast_to(pr["q"][n], shape).copy() for n in range(case.nq)])
    zk = np.broadcast_to(pr["z"], shape)
    r = np.random.default_rng(2).random(shape)
    layer = (zk > rc.cfg.param_z) & (zk <= rc.cfg.param_z + rc.cfg.param_delz)
    hth = np.where(layer, -0.001 + (r - 0.5) / 20.0, 0.0)
    t_cpu = time.perf_counter() - t0
    del hu, hv, hw, hq, hth, r
    return {"namelist": "namelist.in (adiabatic profile, param_theta noise), grid 256x256x128",
            "profiles_ms": 1e3 * t_prof, "model_create_device_init_ms": 1e3 * t_dev,
            "model_create_from_host_arrays_ms": 1e3 * t_host,
            "fields": nf, "field_MB": 8.0 * nf * case.dom.cells / 1e6,
            "cpu_baseline": {"value": 1e3 * t_cpu, "unit": "ms", "cores": 1, "kind": "port",
                             "sample": "the same 3-D initial state built with numpy on the "
                                       "host from the profiles (no upload), full grid"}}


def ensemble_line(H, W, args):
    """One JSON line for the 64-member ensemble alone: value = the faster
    mixed path's cell-steps/s (graphs on: the multigrid members' concurrent
    solves or the batched Chebyshev solve), the roofline of the batched
    sweep from a second, event-timed run of the same steps."""
    print(f"[bench] ensemble: 64 uneven members, warmup {args.warmup}, timed {args.steps} steps",
          file=sys.stderr, flush=True)
    ens = run_ensemble(H, W, args.warmup, args.steps)
    prof = run_ensemble(H, W, args.warmup, args.steps, profile=True)["batched_mixed"]["kernel"]
    best = ens.get("best_mixed", "batched_mixed")
    b = ens[best]
    print(f"[bench] ensemble: batched mixed {ens['batched_mixed']['value']:.4g}, multigrid "
          f"mixed {ens.get('multigrid_mixed', {}).get('value', float('nan')):.4g}, sequential "
          f"FP64 {ens['sequential_fp64']['value']:.4g} cell-steps/s", file=sys.stderr, flush=True)
    achieved = None
    if prof and prof["avg_ms"]:
        achieved = prof["algorithmic_bytes"] / (prof["avg_ms"] * prof["launches"] * 1e-3) / 1e9
    line = {
        "metric": "cell-steps/s, 64-member ensemble: mixed vs sequential FP64",
        "value": b["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": b["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "ensemble64_uneven_warm_bubbles", "members": ens["members"],
                   "cells": ens["cells"], "delta_m": 50.0, "dt_s": 1.0, "fields": 1, "kord": 2,
                   "monotone": True, "ptol": 1e-10, "seed": 404, "parallelism": "single GPU",
                   "path": best},
        "ensemble": ens,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS if achieved else None, "traffic": None,
                     "kernel": "k_jac_b<float,false>", "timing": prof},
    }
    print(json.dumps(line))
    return 0


def rel_l2(a, b, center=False):
    a = a[1:-1, 1:-1, 1:-1]
    b = b[1:-1, 1:-1, 1:-1]
    if center:
        a = a - a.mean()
        b = b - b.mean()
    n = np.linalg.norm(a)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def kelvin_helmholtz(H, W, steps, warmup=2):
    """The shear layer (512x4x256 cells at 10 m, dt=0.2 s) on both paths from
    the same state: after ``steps`` timed steps, the mixed path's relative L2
    error against FP64 for p (mean removed), u, w and theta, with both paths'
    iterations, FP64 fallbacks and worst true residual.  Both paths use the
    V-cycle, semi-coarsened below 128x1x64 (y stops halving at one cell)."""
    case = W.kelvin_helmholtz(steps=warmup + steps)
    out, snaps = {}, {}
    for name, prec in (("fp64", H.FP64), ("mixed", H.MIXED)):
        with H.Context(case.dom) as ctx:
            ctx.set_precond_chebyshev(CHEB_A, CHEB_DIVISIONS)
            ctx.set_precond_multigrid(True)
            model = H.Model(ctx, case, prec)
            try:
                model.set_check_residual(True)
                if warmup:
                    model.step(warmup)
                t0 = time.perf_counter()
                st = model.step(steps)
                wall = time.perf_counter() - t0
                snaps[name] = {"p": model.get(H.F_P), "u": model.get(H.F_U),
                               "w": model.get(H.F_W), "th": model.get(H.F_TH)}
            finally:
                model.close()
        out[name] = {"value": case.dom.cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
                     "iters_per_step": st.iters_total / steps, "iters_max": st.iters_max,
                     "outer_per_step": st.outer_total / steps, "fallbacks": st.fallbacks,
                     "max_true_resid": st.max_true_resid}
    out["grid"] = f"{case.dom.ip}x{case.dom.jp}x{case.dom.kp}"
    out["steps"] = warmup + steps
    out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["mixed_error_vs_fp64"] = {k: rel_l2(snaps["fp64"][k], snaps["mixed"][k],
                                            center=(k == "p")) for k in snaps["fp64"]}
    return out


def large_domain(H, W, steps, warmup=1):
    """The 512x512x256 domain (3x3 bubbles, theta + qv + 4 tracers, 25 m,
    dt=0.25 s) on both paths, each with the V-cycle: cell-steps/s, the
    device memory the model holds (free memory before creation minus after
    its first step), and the canonical stage bytes per cell-step times the
    rate."""
    case = W.large_domain(steps=warmup + steps)
    F = 1 + case.nq
    out = {"grid": f"{case.dom.ip}x{case.dom.jp}x{case.dom.kp}", "fields": F}
    for name, prec in (("fp64", H.FP64), ("mixed", H.MIXED)):
        free0, total = H.device_memory()
        with H.Context(case.dom) as ctx:
            ctx.set_precond_chebyshev(CHEB_A, CHEB_DIVISIONS)
            ctx.set_precond_multigrid(True)
            model = H.Model(ctx, case, prec)
            try:
                model.set_check_residual(True)
                model.step(warmup)
                free1, _ = H.device_memory()
                t0 = time.perf_counter()
                st = model.step(steps)
                wall = time.perf_counter() - t0
            finally:
                model.close()
        value = case.dom.cells * steps / wall
        it = st.iters_total / steps
        bpc = (canonical_bytes_mixed(F, it, st.outer_total / steps, multigrid=True)
               if prec == H.MIXED else canonical_bytes_fp64(F, it))
        out[name] = {"value": value, "ms_per_step": 1e3 * wall / steps, "iters_per_step": it,
                     "outer_per_step": st.outer_total / steps, "fallbacks": st.fallbacks,
                     "max_true_resid": st.max_true_resid,
                     "hbm_footprint_GiB": (free0 - free1) / 2**30,
                     "hbm_total_GiB": total / 2**30,
                     "canonical_bytes_per_cell_step": bpc,
                     "canonical_GBps": bpc * value / 1e9}
    out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(W, ip, jp, steps, case=None, threads=None):
    """The CPU oracle (a C restatement of the reference) on the host's
    cores: the model step of the same workload on an ip x jp x 128 domain
    (the full grid by default), its element-wise loops over x-slabs on
    ``threads`` OpenMP threads (at most 16, the host share of one GPU) and
    its dot products as parallel reductions (an order that changes the last
    bits of the iteration, not its cost)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    if threads is None:
        threads = max(1, min(16, os.cpu_count() or 1))
    O.set_threads(threads)
    O.set_sum_blocked(2)
    if case is None or (case.dom.ip, case.dom.jp) != (ip, jp):
        case = W.moist_bubble(ip, jp, 128, steps=steps)
    try:
        m = O.Model(case)
        t0 = time.perf_counter()
        m.step(steps)
        wall = time.perf_counter() - t0
    finally:
        O.set_threads(1)
        O.set_sum_blocked(0)
    return {"value": case.dom.cells * steps / wall, "unit": "cell-steps/s", "cores": threads,
            "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"oracle model step (C restatement of the reference, OpenMP over "
                      f"x-slabs), moist bubble {ip}x{jp}x128, {steps} step(s) from the "
                      f"initial state, {wall:.1f} s; BiCGSTAB iterations {m.iters}"}


def stream_copy_gbps(H, dom):
    n = 1 << 27  # 1 GiB per array
    ctx = H.Context(dom)
    a = H.DeviceArray(ctx, (n,))
    b = H.DeviceArray(ctx, (n,))
    ctx.stream_copy(b, a)
    best = min(ctx.stream_copy(b, a) for _ in range(5))
    a.free()
    b.free()
    ctx.close()
    return 2.0 * n * 8 / (best * 1e-3) / 1e9


def load_busy(workload):
    """Kernel time per step of a path from the committed rocprofv3 kernel
    trace summary (sum of every kernel's duration over the traced steps)."""
    path = os.path.join(ROOT, "profiles", f"busy_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return {"kernel_ms_per_step": d["kernel_ms_per_step"], "source": f"profiles/busy_{workload}.json"}


def load_traffic(workload, kernel=DOMINANT):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", default="256x256x128")
    ap.add_argument("--paths", default="fp64,mixed")
    ap.add_argument("--inner-rtol", type=float, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chebyshev-path", action="store_true",
                    help="skip the mixed path with Chebyshev-relaxed sweeps")
    ap.add_argument("--no-plain-sweeps", action="store_true",
                    help="skip the mixed path with the reference's plain Jacobi sweeps")
    ap.add_argument("--cpu-sample", default="256x256:2",
                    help="IPxJP:STEPS of the CPU baseline (the oracle on the host's cores)")
    ap.add_argument("--no-fp64-multigrid", action="store_true",
                    help="skip the FP64 path with the V-cycle")
    ap.add_argument("--kh-steps", type=int, default=998,
                    help="timed steps per path of the Kelvin-Helmholtz leg after 2 warm-up "
                         "steps (default: the 1000 steps of the shear-layer run; 0: skip)")
    ap.add_argument("--large-steps", type=int, default=3,
                    help="timed steps per path of the 512x512x256 leg (0: skip)")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--mpdata-variant-calls", type=int, default=10,
                    help="calls per MPDATA variant and routine (0: skip that leg)")
    ap.add_argument("--ensemble-steps", type=int, default=10,
                    help="timed steps of the 64-member ensemble comparison (0: skip)")
    ap.add_argument("--physics-steps", type=int, default=10,
                    help="timed steps of the sheared moist bubble with the options of "
                         "sources on (0: skip)")
    ap.add_argument("--output-steps", type=int, default=20,
                    help="steps of the output-file leg (a record every 5; 0: skip)")
    ap.add_argument("--workload", choices=("moist_bubble", "ensemble"), default="moist_bubble",
                    help="ensemble: only the 64-member uneven ensemble, batched mixed vs "
                         "sequential FP64, --steps timed steps")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: the library runs a domain on exactly one GPU (asked for "
              f"{args.gpus}); the whole domain lives on one MI355X.", file=sys.stderr)
        return 2

    from dcm_amd import hip as H
    from dcm_amd import workloads as W

    if args.workload == "ensemble":
        return ensemble_line(H, W, args)

    ip, jp, kp = (int(x) for x in args.grid.split("x"))
    case = W.moist_bubble(ip, jp, kp)
    workload = f"moist_bubble_{ip}x{jp}x{kp}"
    F = 1 + case.nq
    paths = args.paths.split(",")
    out = {}
    snaps = {}
    keep = []
    for pth in paths:
        prec = H.MIXED if pth == "mixed" else H.FP64
        print(f"[bench] {pth}: warmup {args.warmup}, timed {args.steps} steps", file=sys.stderr,
              flush=True)
        res, snap, handles = run_path(H, case, prec, args.warmup, args.steps,
                                      inner_rtol=args.inner_rtol if pth == "mixed" else None,
                                      precond="multigrid" if pth == "mixed" else "chebyshev",
                                      candidates=[n for n, _, _ in MPDATA32]
                                      if pth == "mixed" else ())
        out[pth] = res
        snaps[pth] = snap
        handles[1].close()
        handles[0].close()
        print(f"[bench] {pth}: {res['value']:.4g} cell-steps/s, {res['ms_per_step']:.2f} "
              f"ms/step, {res['iters_per_step']:.1f} it/step", file=sys.stderr, flush=True)

    if "mixed" in out and "fp64" in out and not args.no_fp64_multigrid:
        # the FP64 path with the same V-cycle: the gain of the precision alone
        print(f"[bench] FP64 with the V-cycle: warmup {args.warmup}, timed {args.steps} steps",
              file=sys.stderr, flush=True)
        res, snap, handles = run_path(H, case, H.FP64, args.warmup, args.steps,
                                      precond="multigrid")
        handles[1].close()
        handles[0].close()
        out["fp64_multigrid"] = res
        snaps["fp64_multigrid"] = snap
        print(f"[bench] FP64, V-cycle: {res['value']:.4g} cell-steps/s, "
              f"{res['ms_per_step']:.2f} ms/step, {res['iters_per_step']:.1f} it/step",
              file=sys.stderr, flush=True)
    if "mixed" in out and not args.no_chebyshev_path:
        # the mixed path with Chebyshev-relaxed Jacobi sweeps instead of the
        # V-cycle, to separate the multigrid gain
        print(f"[bench] mixed with Chebyshev sweeps: warmup {args.warmup}, timed "
              f"{args.steps} steps", file=sys.stderr, flush=True)
        res, _, handles = run_path(H, case, H.MIXED, args.warmup, args.steps,
                                   inner_rtol=args.inner_rtol, precond="chebyshev")
        handles[1].close()
        handles[0].close()
        out["mixed_chebyshev"] = res
        print(f"[bench] mixed, Chebyshev sweeps: {res['value']:.4g} cell-steps/s, "
              f"{res['ms_per_step']:.2f} ms/step, {res['iters_per_step']:.1f} it/step",
              file=sys.stderr, flush=True)
    if "mixed" in out and not args.no_plain_sweeps:
        # the mixed path with the reference's preconditioner (four plain
        # Jacobi sweeps), to separate the FP32 gain from the Chebyshev one
        print(f"[bench] mixed with plain Jacobi sweeps: warmup {args.warmup}, timed "
              f"{args.steps} steps", file=sys.stderr, flush=True)
        res, _, handles = run_path(H, case, H.MIXED, args.warmup, args.steps,
                                   inner_rtol=args.inner_rtol, precond="plain")
        handles[1].close()
        handles[0].close()
        out["mixed_plain_jacobi"] = res
        print(f"[bench] mixed, plain sweeps: {res['value']:.4g} cell-steps/s, "
              f"{res['ms_per_step']:.2f} ms/step, {res['iters_per_step']:.1f} it/step",
              file=sys.stderr, flush=True)
    main_path = "mixed" if "mixed" in out else paths[0]
    r = out[main_path]
    cells = case.dom.cells
    word = 4 if main_path == "mixed" else 8
    k = r["kernel"]
    cands = {}
    if k["avg_ms"]:
        # Launches after convergence (up to a polling interval) return at
        # once; the sweep's launch time is averaged over the launches that
        # swept: the iterations times the launches per iteration.
        # avg_launch_ms (every launch) is the figure that rocprofv3's kernel
        # statistics report.
        total_ms = k["avg_ms"] * k["launches"]
        per_iter = MG_SWEEPS_PER_ITER if main_path == "mixed" else jac_per_iter(4)
        active = min(k["launches"], per_iter * k["profiled_iters"])
        cands[DOMINANT] = {
            "kernel": ("k_jac<float,false,true> (a damped-Jacobi sweep of the V-cycle's "
                       "finest level)" if word == 4 else "k_jac<double,false,false>"),
            "ms_per_step": total_ms / PROFILE_STEPS,
            "algorithmic_bytes_per_launch": JAC_WORDS * word * cells,
            "avg_active_launch_ms": total_ms / active,
            "avg_launch_ms": k["avg_ms"], "launches": k["launches"]}
    if main_path == "mixed":
        for name, label, bpcf in MPDATA32:
            t = r["kernels"].get(name) or {}
            if not t.get("avg_ms"):
                continue
            # one launch serves every field of the step (F <= 16)
            per_launch = bpcf * cells * F
            cands[name] = {"kernel": label, "ms_per_step": t["ms_per_step"],
                           "algorithmic_bytes_per_launch": per_launch,
                           "avg_active_launch_ms": t["avg_ms"],
                           "avg_launch_ms": t["avg_ms"], "launches": t["launches"]}
    for c in cands.values():
        c["achieved_GBps"] = c["algorithmic_bytes_per_launch"] / (c["avg_active_launch_ms"] *
                                                                   1e-3) / 1e9
        c["frac"] = c["achieved_GBps"] / HBM_PEAK_GBPS
    dom_name = max(cands, key=lambda n: cands[n]["ms_per_step"]) if cands else DOMINANT
    d = cands.get(dom_name, {})
    achieved = d.get("achieved_GBps")
    traffic = load_traffic(workload + ("_mixed" if main_path == "mixed" else "_fp64"), dom_name)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBPS) if achieved else None,
                "traffic": traffic,
                "kernel": d.get("kernel", dom_name),
                "algorithmic_bytes_per_launch": d.get("algorithmic_bytes_per_launch"),
                "avg_active_launch_ms": d.get("avg_active_launch_ms"),
                "avg_launch_ms": d.get("avg_launch_ms"), "launches": d.get("launches"),
                "ms_per_step": d.get("ms_per_step"),
                "candidates": {n: {kk: c[kk] for kk in ("ms_per_step", "achieved_GBps", "frac")}
                               for n, c in cands.items()}}
    busy = load_busy(workload + "_mixed") if main_path == "mixed" else None
    if busy:
        busy = dict(busy)
        busy["gpu_busy_fraction"] = busy["kernel_ms_per_step"] / r["ms_per_step"]
    step_roof = {}
    for pth, res in out.items():
        if pth.startswith("mixed"):
            bpc = canonical_bytes_mixed(F, res["iters_per_step"], res["outer_per_step"],
                                        4 if pth == "mixed_plain_jacobi" else CHEB_DIVISIONS,
                                        multigrid=pth == "mixed")
        elif pth == "fp64_multigrid":
            # the V-cycle's canonical words in FP64 (twice the FP32 bytes)
            bpc = 8.0 * (39 + 27 * F) + 2 * 4 * MG_WORDS_PER_ITER * res["iters_per_step"]
        else:
            bpc = canonical_bytes_fp64(F, res["iters_per_step"])
        step_roof[pth] = {"bytes_per_cell_step": bpc,
                          "GBps": bpc * res["value"] / 1e9,
                          "frac_of_8TBps": bpc * res["value"] / 8e12}
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_path == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "grid": args.grid, "delta_m": 50.0, "dt_s": 0.5,
                   "fields": F, "kord": 2, "monotone": True, "ptol": 1e-10,
                   "parallelism": "single GPU"},
        "paths": out,
        "roofline": roofline,
        "step_roofline": step_roof,
    }
    if busy:
        line["gpu_busy"] = busy
    if "fp64" in out and "mixed" in out:
        line["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    if "fp64_multigrid" in out:
        line["speedup_mixed_vs_fp64_same_precond"] = (out["mixed"]["value"] /
                                                      out["fp64_multigrid"]["value"])
        a, b = snaps["fp64"], snaps["mixed"]
        line["mixed_error_vs_fp64"] = {
            name: rel_l2(a[name], b[name], center=(name == "p")) for name in a}
    print("[bench] sedimentation (vertical MPDATA) 256x256x128", file=sys.stderr, flush=True)
    line["sedimentation"] = sedimentation(H, W, cpu=not args.no_cpu)
    if args.mpdata_variant_calls > 0:
        print("[bench] MPDATA of sfvt/advection_3d.F90 beside the linked copy, 256x256x128",
              file=sys.stderr, flush=True)
        line["mpdata_variants"] = mpdata_variants(H, W, calls=args.mpdata_variant_calls,
                                                  cpu=not args.no_cpu)
    if args.physics_steps > 0:
        print(f"[bench] physics options: sheared moist bubble 256x256x128, "
              f"{args.physics_steps} steps", file=sys.stderr, flush=True)
        line["physics"] = physics(H, W, args.physics_steps, base=out, cpu=not args.no_cpu)
    print("[bench] initialisation from namelist.in", file=sys.stderr, flush=True)
    line["initialisation"] = initialisation(H, W)
    if args.output_steps > 0:
        print(f"[bench] output file: {args.output_steps} steps, a record every 5",
              file=sys.stderr, flush=True)
        line["output"] = output(H, W, args.output_steps)
    if args.ensemble_steps > 0:
        print(f"[bench] ensemble: 64 uneven members, {args.ensemble_steps} steps",
              file=sys.stderr, flush=True)
        line["ensemble"] = run_ensemble(H, W, 1, args.ensemble_steps)
        ens = line["ensemble"]
        print(f"[bench] ensemble: batched mixed {ens['batched_mixed']['value']:.4g}, "
              f"multigrid mixed {ens['multigrid_mixed']['value']:.4g}, "
              f"sequential FP64 {ens['sequential_fp64']['value']:.4g} cell-steps/s",
              file=sys.stderr, flush=True)
    if args.kh_steps > 0:
        print(f"[bench] Kelvin-Helmholtz 512x4x256: {args.kh_steps} steps per path",
              file=sys.stderr, flush=True)
        line["kelvin_helmholtz"] = kelvin_helmholtz(H, W, args.kh_steps)
        kh = line["kelvin_helmholtz"]
        print(f"[bench] Kelvin-Helmholtz: mixed {kh['mixed']['value']:.4g}, FP64 "
              f"{kh['fp64']['value']:.4g} cell-steps/s, errors {kh['mixed_error_vs_fp64']}",
              file=sys.stderr, flush=True)
    if args.large_steps > 0:
        print(f"[bench] large domain 512x512x256: {args.large_steps} steps per path",
              file=sys.stderr, flush=True)
        line["large_domain"] = large_domain(H, W, args.large_steps)
        ld = line["large_domain"]
        print(f"[bench] large domain: mixed {ld['mixed']['value']:.4g}, FP64 "
              f"{ld['fp64']['value']:.4g} cell-steps/s", file=sys.stderr, flush=True)
    if not args.no_stream:
        line["stream_copy_GBps"] = stream_copy_gbps(H, case.dom)
    if not args.no_cpu:
        cip, steps = args.cpu_sample.split(":")
        cx, cy = (int(x) for x in cip.split("x"))
        print(f"[bench] CPU baseline: oracle, moist bubble {cx}x{cy}x128, {steps} step(s)",
              file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(W, cx, cy, int(steps), case=case)
        print(f"[bench] CPU baseline: {line['cpu_baseline']['value']:.4g} cell-steps/s on "
              f"{line['cpu_baseline']['cores']} threads", file=sys.stderr, flush=True)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
