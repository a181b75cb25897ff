"""Taylor-Green vortex run on the GPU path.

The reference driver's time loop (src/3DDriver_parfmdf.f90:430-503) around
the GPU step, with the statistics that the TGV case produces every step:

  * ``Ek``    -- turbk, (time, turbk) to fort.22
                 (caseFiles/tgv/BackstepProblem-stoch.f90:679,695);
  * ``dissp`` -- -dEk/dt as the first difference (Ek_prev - Ek)/dt, fort.23
                 (BackstepProblem-stoch.f90:692,696);
  * ``eps``   -- the enstrophy dissipation of energydissipation
                 (src/Energydissipation/energydissipation.f90:86-88), which the
                 driver evaluates right after tstep and before advancing the
                 clock (3DDriver_parfmdf.f90:466-471,502).

dt is CFL times the inviscid bound of the initial state, fixed for the run
(global time stepping, 3DDriver_parfmdf.f90:436-457).  Row k of the history
is the state after step k; ``time`` is the clock after the step.  The
reference starts its first difference from prestep = 0
(BackstepProblem-stoch.f90:459-461), which makes its first fort.23 value
-Ek/dt; here the first difference starts from the Ek of the initial state.

    python -m dsem_amd.tgv --n 4 --p 4 --steps 100 --out tgv_4x4x4_p4
writes <out>.npz (final state JQ [K,5,np], Gauss points, history) and
<out>.txt (step, time, Ek, -dEk/dt, eps).
"""
from __future__ import annotations

import argparse

import numpy as np

from . import cases
from . import mesh as M
from .solver import DSEM


def run(n: int = 4, p: int = 4, steps: int = 100, re: float = 1600.0, pr: float = 0.71, cfl: float = 0.5,
        variant: str = "isothermal", perturb: float = 0.0, seed: int = 1234, device: int = 0,
        metrics: str = "per_point", every: int = 1, progress: int = 0) -> dict:
    """History rows after every ``every``-th step (and the last); -dEk/dt is
    then the first difference over those steps.  ``progress`` > 0 prints a
    line every that many steps."""
    mesh = M.periodic_box(n, p)
    every = max(1, int(every))
    with DSEM(mesh, re=re, pr=pr, metrics=metrics, device=device) as s:
        xg, jac = s.gauss_points()
        s.set_state(cases.tgv_state(xg, jac, variant, perturb=perturb, seed=seed))
        s.keep_lagged_states(True)
        dt = s.dtcalc(cfl)
        prev, kprev = s.kinetic_energy(), 0
        rows = [(0, 0.0, prev, np.nan, np.nan)]
        t = 0.0
        for k in range(1, steps + 1):
            s.tstep(t, dt)
            t += dt
            if k % every == 0 or k == steps:
                ek, eps = s.diagnostics()
                rows.append((k, t, ek, (prev - ek) / ((k - kprev) * dt), eps))
                prev, kprev = ek, k
            if progress > 0 and k % progress == 0:
                print(f"step {k} t={t:.4f} Ek={rows[-1][2]:.10f} -dEk/dt={rows[-1][3]:.6e}", flush=True)
        JQ = s.get_state()
    return {"history": np.array(rows), "JQ": JQ, "xg": xg, "jacob": jac, "dt": dt,
            "config": dict(n=n, p=p, steps=steps, re=re, pr=pr, cfl=cfl, variant=variant, perturb=perturb,
                           every=every)}


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=4)
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--re", type=float, default=1600.0)
    ap.add_argument("--pr", type=float, default=0.71)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--variant", default="isothermal", choices=["isothermal", "reference"])
    ap.add_argument("--every", type=int, default=1, help="history row every k steps")
    ap.add_argument("--progress", type=int, default=0, help="print a line every k steps")
    ap.add_argument("--out", default="tgv")
    a = ap.parse_args(argv)
    r = run(a.n, a.p, a.steps, a.re, a.pr, a.cfl, a.variant, every=a.every, progress=a.progress)
    np.savez(a.out + ".npz", JQ=r["JQ"], xg=r["xg"], jacob=r["jacob"], history=r["history"], dt=r["dt"])
    np.savetxt(a.out + ".txt", r["history"], fmt=["%d", "%.10e", "%.16e", "%.16e", "%.16e"],
               header="step time Ek -dEk/dt eps")
    h = r["history"]
    print(f"dt={r['dt']:.6e}  Ek(0)={h[0, 2]:.12f}  Ek(end)={h[-1, 2]:.12f}  -dEk/dt(end)={h[-1, 3]:.6e}")


if __name__ == "__main__":
    main()
