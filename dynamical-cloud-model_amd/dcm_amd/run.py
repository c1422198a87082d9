"""A run of the resident model from one of the reference's namelist files,
in the order of main.f90 and model_driver (driver_code.f90:304-676): the
initial state generated on the GPU, the time loop, and the output file
written every output_interval at the start of a step (:310-333), from the
GPU asynchronously.

    python -m dcm_amd.run namelist.in [--steps N] [--precision mixed|fp64]
                                      [--output PATH] [--set kp=64 ...]

Radiation, land surface and microphysics are not on the GPU path: a
namelist that switches radiation on needs --set radiation=false; the
moisture fields are advected (and written) without microphysics."""
from __future__ import annotations

import argparse
import sys
import time as _time

from . import hip as H
from . import init as I


def run_namelist(path, steps=None, precision=H.MIXED, output=None, overrides=None,
                 q_names=None, progress=None):
    """Run `steps` time steps (default: the namelist's runtime / dt).
    output: the file (default nm1%outputfile; False: no output).  Returns
    a dict with the step statistics, the records written and wall times."""
    case = I.case_from_namelist(path, overrides=overrides)
    rc, pr = case.init
    nm = rc.nm
    ntim = pr["ntim"] if steps is None else int(steps)
    out_path = nm["outputfile"] if output is None else output
    dt, interval = nm["dt"], nm["output_interval"]
    res = {"steps": 0, "records": 0, "iters": 0, "max_true_resid": 0.0}
    t0 = _time.perf_counter()
    with H.Context(case.dom) as ctx:
        model = H.Model(ctx, case, precision)
        out = None
        try:
            if out_path:
                names = q_names or [f"q{n + 1}" for n in range(case.nq)]
                out = H.Output(model, out_path, nprec=max(nm["nprec"], 1) if case.nq else 0,
                               q_names=names, trefn=pr["trefn"])
            last = -interval
            t_init = _time.perf_counter() - t0
            for n in range(1, ntim + 1):
                time = (n - 1) * dt  # :309
                if out is not None and time - last >= interval:
                    out.write(time)
                    last = time
                    res["records"] += 1
                st = model.step(1)
                res["steps"] += 1
                res["iters"] += st.iters_total
                res["max_true_resid"] = max(res["max_true_resid"], st.max_true_resid)
                if progress and n % progress == 0:
                    print(f"step {n} of {ntim}, time {n * dt:.1f} s, {st.iters_total} "
                          f"iterations", file=sys.stderr, flush=True)
            if out is not None:
                out.flush()
                res["writer"] = out.stats()
        finally:
            if out is not None:
                out.close()
            model.close()
    res["init_s"] = t_init
    res["wall_s"] = _time.perf_counter() - t0
    res["cells"] = case.dom.cells
    return res


def _value(text):
    low = text.lower()
    if low in ("true", ".true.", "t"):
        return True
    if low in ("false", ".false.", "f"):
        return False
    try:
        return int(text)
    except ValueError:
        return float(text)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("namelist")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed")
    ap.add_argument("--output", default=None, help="output file (default nm1%%outputfile)")
    ap.add_argument("--no-output", action="store_true")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="override a run_vars setting, e.g. kp=64 or radiation=false")
    a = ap.parse_args(argv)
    overrides = {}
    for kv in a.set:
        k, v = kv.split("=", 1)
        overrides[k.strip().lower()] = _value(v.strip())
    res = run_namelist(a.namelist, a.steps, H.MIXED if a.precision == "mixed" else H.FP64,
                       False if a.no_output else a.output, overrides, progress=100)
    rate = res["cells"] * res["steps"] / max(res["wall_s"] - res["init_s"], 1e-9)
    print(f"{res['steps']} steps, {res['records']} records, {res['iters']} solver iterations, "
          f"max residual {res['max_true_resid']:.2e}, {rate:.3g} cell-steps/s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
