"""Adjoint of the entropy-viscosity path (no counterpart in the reference,
which has no derivative code) against central differences:

* dsem_ev_vjp, the reverse of k_ev (Entropy_Viscosity, EViscosity_GL.f90:19-101;
  RHS_ent_trans, RHS.f90:21-190; smoothing, viscosity_smoothing.f90:4-195;
  EV_Prolong, Prolong_gtol.f90:7-51): <muhbar, mu_h(JQ + h dJQ, Slag + h dSlag)
  - mu_h(JQ - h dJQ, Slag - h dSlag)> / 2h against <JQbar, dJQ> + <Slagbar, dSlag>,
  with the Ducros sensor, the smoothing (p = 1, 2 and higher: its three
  schedules) and the c_max cap each switched on, and with the mumax cap
  binding at part of the points;
* a shock run through torch.autograd (RKSolve with entropy viscosity: the
  state, the lagged states and their adjoints carried across steps, mu_h's
  adjoint through k_volume's and k_ev's reverses): dJ/dJQ0 of
  J = int 1/2 rho |u|^2 after 4 steps along seeded directions, and dJ/dnu.
"""
import numpy as np
import pytest

from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM
from helpers import distorted_box

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _lagged_setup(mesh, kw, seed):
    """A state and the lagged states one step of the same run leaves."""
    s = DSEM(mesh, re=200.0, shock=True, **kw)
    xg, jac = s.gauss_points()
    JQ = cases.tgv_state(xg, jac, "isothermal", perturb=3e-2, seed=seed)
    s.set_state(JQ)
    dt = 0.3 * s.dtcalc()
    s.tstep(0.0, dt)
    return s, s.get_state(), s.get_lagged_states()


def _muh(s, JQ, stag):
    s.set_state(JQ)
    s.set_lagged_states(stag)
    s.ev_update()
    return s.get_viscosity()[1]


@pytest.mark.parametrize("p,opts", [(3, {}), (3, {"ducros": True}), (4, {"smoothing": True}), (1, {"smoothing": True}),
                                    (2, {"smoothing": True, "ducros": True}), (3, {"cmax_cap": True}),
                                    (5, {"ducros": True, "smoothing": True, "cmax_cap": True})])
def test_ev_vjp_matches_central_differences(p, opts):
    print(f"\n[case] ev vjp p={p} {opts}", flush=True)
    mesh = distorted_box(3, p, amp=0.04, seed=p)
    kw = dict(cmu=2.0, mumax=0.02, cmax=0.5, **opts)
    s, JQ, stag = _lagged_setup(mesh, kw, seed=10 + p)
    with s:
        mu0 = _muh(s, JQ, stag)
        if not opts.get("cmax_cap"):
            capped = np.mean(np.isclose(mu0, kw["mumax"], rtol=0, atol=0)) if mu0.size else 0
            print(f"mu_h max {mu0.max():.3e}, share at the mumax cap on the staggered points (after prolongation) "
                  f"{capped:.3f}", flush=True)
        rng = np.random.default_rng(p)
        for trial in range(2):
            lam = rng.standard_normal(mu0.shape)
            dJQ = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
            dS = rng.standard_normal(stag.shape) * np.abs(stag).mean(axis=(0, 1, 3), keepdims=True)
            s.set_state(JQ)
            s.set_lagged_states(stag)
            jb, sb = s.ev_vjp(lam)
            h = 1e-6
            fp = float(np.sum(lam * _muh(s, JQ + h * dJQ, stag + h * dS)))
            fm = float(np.sum(lam * _muh(s, JQ - h * dJQ, stag - h * dS)))
            fd = (fp - fm) / (2 * h)
            ad = float(np.sum(jb * dJQ) + np.sum(sb * dS))
            print(f"trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
            assert _rel(ad, fd) < 1e-6


def test_shock_run_gradient_through_autograd():
    import torch
    from dsem_amd import autograd as A
    print("\n[case] shock run adjoint 3^3 p=3, 4 steps", flush=True)
    mesh = distorted_box(3, 3, amp=0.04, seed=2)
    kw = dict(re=200.0, shock=True, cmu=2.0, mumax=0.05, ducros=True, smoothing=True)
    nsteps = 4
    with DSEM(mesh, **kw) as s:
        xg, jac = s.gauss_points()
        JQ = cases.tgv_state(xg, jac, "isothermal", perturb=3e-2, seed=4)
        s.set_state(JQ)
        dt = 0.3 * s.dtcalc()
        bmat = A.bmat_tensor(s, "cuda")
        grads = {}
        for snaps, tape in ((None, None), (1, 1)):
            x = torch.tensor(JQ, device="cuda", requires_grad=True)
            nu = torch.tensor(1.0 / 200.0, dtype=torch.float64, device="cuda", requires_grad=True)
            J = A.kinetic_energy_integral(A.solve(s, x, nsteps, dt, every=2, nu=nu, snaps=snaps, tape=tape), bmat)
            J.backward()
            grads[(snaps, tape)] = (x.grad.cpu().numpy(), float(nu.grad))
        g, gnu = grads[(None, None)]
        g2, gnu2 = grads[(1, 1)]
        assert np.linalg.norm(g2 - g) <= 1e-13 * np.linalg.norm(g) and abs(gnu2 - gnu) <= 1e-12 * abs(gnu)
        bm = s.quadrature_weights()

        def Jf(q0, ctx=s):
            ctx.set_state(q0)
            ctx.steps(nsteps, dt, check=False)
            q = ctx.get_state()
            return 0.5 * float(np.sum(bm * (q[:, 1] ** 2 + q[:, 2] ** 2 + q[:, 3] ** 2) / q[:, 0]))

        rng = np.random.default_rng(2024)
        h = 1e-6
        for trial in range(3):
            d = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
            fd = (Jf(JQ + h * d) - Jf(JQ - h * d)) / (2 * h)
            ad = float(np.sum(g * d))
            print(f"trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
            assert _rel(ad, fd) < 1e-6
    nu0, dnu = 1.0 / 200.0, 1e-4 / 200.0
    vals = []
    for nu_s in (nu0 + dnu, nu0 - dnu):
        kw2 = dict(kw)
        kw2["re"] = 1.0 / nu_s
        with DSEM(mesh, **kw2) as s2:
            vals.append(Jf(JQ, s2))
    fdn = (vals[0] - vals[1]) / (2 * dnu)
    print(f"nu: adjoint {gnu:.12e} fd {fdn:.12e} rel {_rel(gnu, fdn):.2e}", flush=True)
    assert _rel(gnu, fdn) < 1e-5
