"""Error and speed of the mixed-precision contraction modes against the FP64
GPU path on the Taylor-Green vortex.

For each mode (fp32, bf16x3) and each input state it reports, per conserved
field, the relative L2 error ||x - x64|| / ||x64|| and the relative L-inf
error max|x - x64| / max|x64|:

  * of one RHS evaluation, on the TGV field with a seeded 1e-3 relative
    perturbation (the throughput input) and on the clean TGV field;
  * of the state after ``steps`` RK steps from the clean TGV field;

and the time per RK step of every mode.

    python -m dsem_amd.mixed_study --n 22 --p 8 --steps 1000 --out study.json
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import cases
from . import mesh as M
from .solver import DSEM

FIELDS = ("rho", "rho_u", "rho_v", "rho_w", "rho_e")


def field_errors(x: np.ndarray, ref: np.ndarray) -> dict:
    """x, ref [K,5,np] -> {field: {"rel_l2", "rel_linf"}}."""
    out = {}
    for v, name in enumerate(FIELDS):
        d, r = x[:, v] - ref[:, v], ref[:, v]
        out[name] = {"rel_l2": float(np.linalg.norm(d) / max(np.linalg.norm(r), 1e-300)),
                     "rel_linf": float(np.abs(d).max() / max(np.abs(r).max(), 1e-300))}
    return out


def run(n: int = 22, p: int = 8, steps: int = 1000, modes=("fp32", "bf16x3"), re: float = 1600.0,
        cfl: float = 0.5, device: int = 0, timing_steps: int = 20) -> dict:
    mesh = M.periodic_box(n, p)
    res = {"config": dict(n=n, p=p, steps=steps, re=re, cfl=cfl, dof=mesh.dof), "modes": {}}
    solvers = {m: DSEM(mesh, re=re, pr=0.71, precision=m, metrics="per_point", device=device)
               for m in ("fp64",) + tuple(modes)}
    try:
        s64 = solvers["fp64"]
        xg, jac = s64.gauss_points()
        clean = cases.tgv_state(xg, jac, "isothermal")
        pert = cases.tgv_state(xg, jac, "reference", perturb=1e-3, seed=1234)
        s64.set_state(clean)
        dt = s64.dtcalc(cfl)
        rhs64 = {}
        for name, st in (("perturbed", pert), ("clean", clean)):
            s64.set_state(st)
            rhs64[name] = s64.rhs()
        final = {}
        tstep = {}
        for mode, s in solvers.items():
            s.set_state(clean)
            s.steps(steps, dt, check=True)
            final[mode] = s.get_state()
            s.set_state(pert)
            s.steps(2, dt, check=False)
            s.synchronize()
            t0 = time.perf_counter()
            s.steps(timing_steps, dt, check=False)
            s.synchronize()
            tstep[mode] = (time.perf_counter() - t0) * 1e3 / timing_steps
        res["ms_per_step"] = {"fp64": tstep["fp64"]}
        for mode in modes:
            s = solvers[mode]
            r = {}
            for name, st in (("perturbed", pert), ("clean", clean)):
                s.set_state(st)
                r[f"rhs_{name}"] = field_errors(s.rhs(), rhs64[name])
            r[f"state_after_{steps}_steps"] = field_errors(final[mode], final["fp64"])
            r["ms_per_step"] = tstep[mode]
            r["speedup_vs_fp64"] = tstep["fp64"] / tstep[mode]
            res["modes"][mode] = r
    finally:
        for s in solvers.values():
            s.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=22)
    ap.add_argument("--p", type=int, default=8)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    r = run(a.n, a.p, a.steps)
    txt = json.dumps(r, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(txt + "\n")
    print(txt)


if __name__ == "__main__":
    main()
