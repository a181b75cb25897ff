"""Adjoint of the Taylor-Green solve: gradient check and cost.

Objective J = int 1/2 rho |u|^2 dV of the state after ``steps`` RK steps
(Gauss-point quadrature with the bmatu1b weights).  Gradients, by the
hand-written adjoint kernels behind torch.autograd with a checkpoint every
``every`` steps:
  * dJ/du0 -- with respect to the initial velocity field (u, v, w at every
    Gauss point; density and pressure of the initial state held fixed);
  * dJ/dnu -- with respect to nu = 1/Re.
Checks: central finite differences (step h) of J along ``ndir`` seeded random
directions of u0 (numpy default_rng(seed)), each direction normalised to unit
RMS per entry, and along nu (relative step 1e-4).  Cost: wall time of the backward pass
(recomputation included) over the forward pass.

    python -m dsem_amd.adjoint_study --n 8 --p 6 --steps 50 --every 10 --ndir 32
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import autograd as A
from . import cases
from . import mesh as M
from .solver import DSEM

GAMMA = 1.4


def _fixed_fields(xg, jac, device):
    """rho, p (the isothermal TGV field, held fixed) and jacob as device
    tensors, evaluated once outside any timed region."""
    rho, _, _, _, p = cases.tgv_primitive(xg[..., 0], xg[..., 1], xg[..., 2], "isothermal")
    return (torch.as_tensor(rho, device=device), torch.as_tensor(p, device=device),
            torch.as_tensor(jac, device=device))


def _ic_torch(fixed, vel):
    """JQ0 [K,5,np] from the velocity vel [3,K,np] and the fixed fields."""
    rho, p, jac = fixed
    u, v, w = vel[0], vel[1], vel[2]
    rhoe = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    Q = torch.stack([rho, rho * u, rho * v, rho * w, rhoe], dim=1)
    return (Q / jac[:, None, :]).contiguous()


def run(n=8, p=6, steps=50, every=10, ndir=32, h=1e-6, seed=2024, re=1600.0, cfl=0.5, device=0) -> dict:
    mesh = M.periodic_box(n, p)
    dev = torch.device("cuda", device)
    out = {"config": dict(n=n, p=p, steps=steps, every=every, ndir=ndir, h=h, seed=seed, re=re, cfl=cfl,
                          dof=mesh.dof, objective="J = int 1/2 rho |u|^2 dV at the final step")}
    with DSEM(mesh, re=re, pr=0.71, metrics="per_point", device=device) as s:
        xg, jac = s.gauss_points()
        rho, u, v, w, _ = cases.tgv_primitive(xg[..., 0], xg[..., 1], xg[..., 2], "isothermal")
        vel0 = torch.as_tensor(np.stack([u, v, w]), device=dev)
        fixed = _fixed_fields(xg, jac, dev)
        s.set_state(_ic_torch(fixed, vel0).cpu().numpy())
        dt = s.dtcalc(cfl)
        bmat = A.bmat_tensor(s, dev)
        nu = torch.tensor(1.0 / re, dtype=torch.float64, device=dev, requires_grad=True)

        def objective(vel, nu_t, grad=True):
            JQ0 = _ic_torch(fixed, vel)
            return A.kinetic_energy_integral(A.solve(s, JQ0, steps, dt, every, nu_t), bmat)

        # warm-up, then timed forward and backward
        vel = vel0.clone().requires_grad_(True)
        objective(vel, nu).backward()
        vel = vel0.clone().requires_grad_(True)
        nu.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = objective(vel, nu)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        gvel = vel.grad.detach().cpu().numpy()
        gnu = float(nu.grad)
        out["J"] = float(J.detach())
        out["forward_ms"] = (t1 - t0) * 1e3
        out["backward_ms"] = (t2 - t1) * 1e3
        out["adjoint_over_forward"] = (t2 - t1) / (t1 - t0)

        def J_of(vel_np, re_s=None):
            if re_s is None:
                ctx = s
            else:
                ctx = DSEM(mesh, re=re_s, pr=0.71, metrics="per_point", device=device)
            try:
                ctx.set_state(_ic_torch(fixed, torch.as_tensor(vel_np, device=dev)).cpu().numpy())
                ctx.steps(steps, dt, check=False)
                q = ctx.get_state()
            finally:
                if ctx is not s:
                    ctx.close()
            bm = s.quadrature_weights()
            return 0.5 * float(np.sum(bm * (q[:, 1] ** 2 + q[:, 2] ** 2 + q[:, 3] ** 2) / q[:, 0]))

        if ndir <= 0:
            return out                   # cost only, no finite-difference checks
        rng = np.random.default_rng(seed)
        v0 = vel0.cpu().numpy()
        errs, rows = [], []
        for _ in range(ndir):
            d = rng.standard_normal(v0.shape)
            d /= np.sqrt(np.mean(d * d))
            fd = (J_of(v0 + h * d) - J_of(v0 - h * d)) / (2 * h)
            ad = float(np.sum(gvel * d))
            errs.append(abs(ad - fd) / max(abs(fd), 1e-300))
            rows.append((ad, fd))
        nu0 = 1.0 / re
        dnu = 1e-4 * nu0   # J depends weakly on nu: a larger relative step keeps rounding out
        fdn = (J_of(v0, 1.0 / (nu0 + dnu)) - J_of(v0, 1.0 / (nu0 - dnu))) / (2 * dnu)
        out["velocity_directions"] = {"max_rel_err": max(errs), "median_rel_err": float(np.median(errs)),
                                      "first": [{"adjoint": a, "fd": f} for a, f in rows[:4]]}
        out["nu"] = {"adjoint": gnu, "fd": fdn, "rel_err": abs(gnu - fdn) / abs(fdn)}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--p", type=int, default=6)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--ndir", type=int, default=32)
    ap.add_argument("--h", type=float, default=1e-6)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    r = run(a.n, a.p, a.steps, a.every, a.ndir, a.h)
    txt = json.dumps(r, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(txt + "\n")
    print(txt)


if __name__ == "__main__":
    main()
