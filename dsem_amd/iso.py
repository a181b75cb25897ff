"""Decaying compressible isotropic turbulence with entropy viscosity.

Initial field from ``cases.isotropic_turbulence`` (seeded solenoidal Fourier
synthesis, E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, Ma_t = 0.3, uniform rho and T);
Re of the non-dimensional equations chosen so that Re_lambda = 100.  Entropy
viscosity with c_E = Cmu = 1 and the c_max = 0.5 cap (DSEM_EV_CMAX_CAP);
fixed dt = CFL * dtcalc of the initial state.

    python -m dsem_amd.iso --n 16 --p 6 --steps 500 --precision fp32
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import cases
from . import mesh as M
from .solver import DSEM


def run(n: int = 16, p: int = 6, steps: int = 500, precision: str = "fp32", cfl: float = 0.5,
        seed: int = 1234, device: int = 0, record_every: int = 10) -> dict:
    mesh = M.periodic_box(n, p)
    with DSEM(mesh, re=1.0, pr=0.71, metrics="per_point", device=device) as probe:
        xg, jac = probe.gauss_points()
    JQ0, info = cases.isotropic_turbulence(xg, jac, seed=seed)
    re = info["re_for_re_lambda_100"]
    with DSEM(mesh, re=re, pr=0.71, precision=precision, metrics="per_point", shock=True, cmu=1.0, mumax=1e30,
              cmax=0.5, cmax_cap=True, device=device) as s:
        s.set_state(JQ0)
        dt = s.dtcalc(cfl)
        hist = [(0, 0.0, s.kinetic_energy())]
        t0 = time.perf_counter()
        for k in range(1, steps + 1):
            s.tstep(k * dt, dt)
            if k % record_every == 0 or k == steps:
                hist.append((k, k * dt, s.kinetic_energy()))
        s.synchronize()
        wall = time.perf_counter() - t0
        _, muh = s.get_viscosity()
        JQ = s.get_state()
    return {"history": np.array(hist), "JQ": JQ, "JQ0": JQ0, "dt": dt, "re": re, "wall_s": wall,
            "mu_h_max": float(np.abs(muh).max()),
            "config": dict(n=n, p=p, steps=steps, precision=precision, cfl=cfl, seed=seed, dof=mesh.dof,
                           u_rms=info["u_rms"], taylor_microscale=info["taylor_microscale"])}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=16)
    ap.add_argument("--p", type=int, default=6)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "bf16x3"])
    a = ap.parse_args(argv)
    r = run(a.n, a.p, a.steps, a.precision)
    h = r["history"]
    print(json.dumps({"config": r["config"], "re": r["re"], "dt": r["dt"], "wall_s": r["wall_s"],
                      "ms_per_step": r["wall_s"] * 1e3 / a.steps, "Ek0": h[0, 2], "Ek_end": h[-1, 2],
                      "mu_h_max": r["mu_h_max"]}))


if __name__ == "__main__":
    main()
