"""Adjoint of the static Smagorinsky path (no counterpart in the reference,
which has no derivative code) against central differences.

tstep computes nu_t once per step from the step's start state and the
lagged staggered states the previous step's last stage left (simpleSmag,
staticSmag.f90:32-103; filteredStrain :126-179; nut_prolong :206-261), and
the viscous fluxes read amu = 1 + nu_t rho (+ mu_h) (Vis_Fluxes_new.f90:123-125).
A run through torch.autograd (RKSolve carrying the lagged states and their
adjoint across steps; nu_t's adjoint through k_volume's reverse, k_smag's
reverse and its gather) is checked by dJ/dJQ0 of J = int 1/2 rho |u|^2 after
4 steps along seeded directions and by dJ/dnu (nu_t = Re (C_s Delta)^2 S is
itself a function of nu = 1/Re), with the numax cap off and binding at part of
the points, on distorted and rotated meshes (p = 3 on the scalar volume
reverse, p >= 4 on its matrix-core form), alone and with entropy viscosity.
Two checkpoint schedules must give the same gradient.  The gradients of
Smagorinsky and entropy-viscosity runs are also checked against central
differences through the CPU oracle's steps.  Where the cap binds,
nu_t has kinks, and a central difference whose step straddles one at some
point is off there: those cases take steps of 1e-7 (on the rotated mesh one
direction measured 1.0e-6 relative at 1e-6, 3.4e-8 at 1e-7; without the cap
both give 1e-9).
"""
import numpy as np
import pytest

from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM
from helpers import distorted_box

pytestmark = pytest.mark.gpu

CS = 0.12


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


CASES = [
    ("distorted p=3 no cap", lambda: distorted_box(3, 3, amp=0.04, seed=2), 1e9, {}),
    ("distorted 2^3 p=8 no cap", lambda: distorted_box(2, 8, amp=0.04, seed=8), 1e9, {}),
    ("distorted p=5", lambda: distorted_box(3, 5, amp=0.04, seed=5), 0.3, {}),
    ("rotated p=4", lambda: M.rotated_box(4, 4, seed=4), 0.3, {}),
    ("distorted p=4 with entropy viscosity", lambda: distorted_box(3, 4, amp=0.04, seed=3), 0.3,
     dict(shock=True, cmu=2.0, mumax=0.05, ducros=True)),
]


@pytest.mark.parametrize("name,make_mesh,numax,extra", CASES, ids=[c[0] for c in CASES])
def test_les_run_gradient_through_autograd(name, make_mesh, numax, extra):
    import torch
    from dsem_amd import autograd as A
    print(f"\n[case] Smagorinsky run adjoint, {name}, 4 steps", flush=True)
    mesh = make_mesh()
    kw = dict(re=200.0, **extra)
    nsteps = 4
    with DSEM(mesh, **kw) as s:
        s.set_les(True, CS, numax)
        xg, jac = s.gauss_points()
        JQ = cases.tgv_state(xg, jac, "isothermal", perturb=3e-2, seed=4)
        s.set_state(JQ)
        dt = 0.3 * s.dtcalc()
        s.tstep(0.0, dt)
        ng, _ = s.les_update()
        print(f"nu_t after one step: max {ng.max():.3e}, share at the numax cap {np.mean(ng >= numax):.3f}, "
              f"share at zero {np.mean(ng <= 0.0):.3f}", flush=True)
        assert np.mean(ng >= numax) < 1.0
        s.set_state(JQ)
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

        rng = np.random.default_rng(2025)
        h = 1e-6 if numax > 1e6 else 1e-7
        for trial in range(3):
            d = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
            fd = (Jf(JQ + h * d) - Jf(JQ - h * d)) / (2 * h)
            ad = float(np.sum(g * d))
            print(f"trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
            assert _rel(ad, fd) < 1e-6
    nu0 = 1.0 / 200.0
    dnu = nu0 * (1e-4 if numax > 1e6 else 1e-6)
    vals = []
    for nu_s in (nu0 + dnu, nu0 - dnu):
        kw2 = dict(kw)
        kw2["re"] = 1.0 / nu_s
        with DSEM(mesh, **kw2) as s2:
            s2.set_les(True, CS, numax)
            vals.append(Jf(JQ, s2))
    fdn = (vals[0] - vals[1]) / (2 * dnu)
    print(f"nu: adjoint {gnu:.12e} fd {fdn:.12e} rel {_rel(gnu, fdn):.2e}", flush=True)
    assert _rel(gnu, fdn) < 1e-5


def test_les_adjoint_refusals():
    """The plain entry points refuse the model; a tape reserved before
    switching the model on is laid out again."""
    from dsem_amd import _lib as L
    print("\n[case] Smagorinsky adjoint refusals 2^3 p=3", flush=True)
    mesh = distorted_box(2, 3, amp=0.04, seed=1)
    with DSEM(mesh, re=200.0) as s:
        xg, jac = s.gauss_points()
        s.set_state(cases.tgv_state(xg, jac, "isothermal", perturb=3e-2, seed=1))
        s.tape_reserve(2)
        s.set_les(True, CS, 0.3)
        with pytest.raises(L.DsemError):
            s.rk_step_record(1e-3, 0)     # the tape holds no nu_t yet
        s.tape_reserve(2)
        s.rk_step_record(1e-3, 0)
        with pytest.raises(L.DsemError):
            s.rhs_vjp(np.zeros((s.K, 5, s.np)))


@pytest.mark.parametrize("model", ["smagorinsky", "entropy viscosity"])
def test_lagged_model_gradient_against_oracle_finite_differences(model):
    """dJ/dJQ0 of a 4-step run with the Smagorinsky model (cap off) or with
    entropy viscosity (Ducros sensor and smoothing on, mumax cap off), from
    the GPU adjoint, against central differences through the CPU oracle's
    steps (or_rk_step_les / or_rk_step_ev: the restated reference, lagged
    states carried from step to step, the first step a fresh start)."""
    import torch
    import oracle as O
    from dsem_amd import autograd as A
    print(f"\n[case] {model} adjoint vs oracle FD 2^3 p=4, 4 steps", flush=True)
    mesh = M.periodic_box(2, 4)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=3e-2, seed=7)
    ph = O.phys(re=400.0)
    dt = 0.4 * O.dtcalc(g, JQ, ph)
    nsteps = 4
    les = model == "smagorinsky"
    kw = dict(re=400.0, metrics="per_point")
    if not les:
        kw.update(shock=True, cmu=2.0, mumax=1e9, ducros=True, smoothing=True)
    with DSEM(mesh, **kw) as s:
        if les:
            s.set_les(True, CS, 1e9)
        bmat = A.bmat_tensor(s, "cuda")
        x = torch.tensor(JQ, device="cuda", requires_grad=True)
        J = A.kinetic_energy_integral(A.solve(s, x, nsteps, dt, every=2), bmat)
        J.backward()
        grad = x.grad.cpu().numpy()
    ev = O.OrEv(2.0, 1e9, 1, 1)
    bm = O.ke_weights(mesh.p)

    def Jf(q0):
        q = q0.copy()
        stag = np.zeros((mesh.K, 3, 5, g.ns))
        for k in range(nsteps):
            if les:
                O.rk_step_les(g, q, dt, stag, k > 0, CS, 1e9, ph)
            else:
                O.rk_step_ev(g, q, dt, stag, k > 0, ev, ph)
        return 0.5 * float(np.sum(bm * (q[:, 1] ** 2 + q[:, 2] ** 2 + q[:, 3] ** 2) / q[:, 0]))

    rng = np.random.default_rng(99)
    h = 1e-6
    for trial in range(2):
        d = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
        fd = (Jf(JQ + h * d) - Jf(JQ - h * d)) / (2 * h)
        ad = float(np.sum(grad * d))
        print(f"trial {trial}: adjoint {ad:.12e} oracle fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
        assert _rel(ad, fd) < 1e-6


SENSOR_CASES = [
    ("shock sensor, entropy viscosity", dict(shock_sensor=True), True),
    ("density sensor, entropy viscosity", dict(rho_sensor=True), True),
    ("both sensors, entropy viscosity", dict(shock_sensor=True, rho_sensor=True), True),
    ("both sensors, no entropy viscosity", dict(shock_sensor=True, rho_sensor=True), False),
]


@pytest.mark.parametrize("name,sensors,ev", SENSOR_CASES, ids=[c[0] for c in SENSOR_CASES])
def test_les_sensor_run_gradient_through_autograd(name, sensors, ev):
    """The sensor switches (staticSmag.f90:68-69): nu_t times 1 - ducros
    and/or the density sensor, both k_ev's functions of the step's start
    state and lagged states.  4 steps on a distorted 3^3 p = 4 mesh, cap off:
    with entropy viscosity, dJ/dJQ0 and dJ/dnu against central differences;
    without it the sensors are the zero arrays nothing writes, the density
    switch makes nu_t 0, and the gradient must equal the one without the
    model (to 1e-13: the volume kernels then run other instances).  (That run's own central differences are off by up to
    2.6e-6 at a step of 1e-6 and 7.7e-7 at 1e-7, the same with the model
    off: kinks of the forward without entropy viscosity, not the adjoint.)"""
    import torch
    from dsem_amd import autograd as A
    print(f"\n[case] Smagorinsky sensors adjoint, {name}, 3^3 p=4, 4 steps", flush=True)
    mesh = distorted_box(3, 4, amp=0.04, seed=11)
    kw = dict(re=200.0)
    if ev:
        kw.update(shock=True, cmu=2.0, mumax=1e9)
    nsteps = 4

    def make(re):
        s = DSEM(mesh, **dict(kw, re=re))
        s.set_les(True, CS, 1e9)
        s.set_les_sensors(**sensors)
        return s

    def grad(s, JQ, dt, bmat):
        x = torch.tensor(JQ, device="cuda", requires_grad=True)
        nu = torch.tensor(1.0 / 200.0, dtype=torch.float64, device="cuda", requires_grad=True)
        J = A.kinetic_energy_integral(A.solve(s, x, nsteps, dt, every=2, nu=nu), bmat)
        J.backward()
        return x.grad.cpu().numpy(), float(nu.grad)

    with make(200.0) as s:
        xg, jac = s.gauss_points()
        JQ = cases.tgv_state(xg, jac, "isothermal", perturb=3e-2, seed=12)
        s.set_state(JQ)
        dt = 0.3 * s.dtcalc()
        bmat = A.bmat_tensor(s, "cuda")
        g, gnu = grad(s, JQ, dt, bmat)
        # the sensors' share of the gradient, so that the check below has power
        s.set_les_sensors()
        g0, _ = grad(s, JQ, dt, bmat)
        s.set_les_sensors(**sensors)
        share = np.linalg.norm(g - g0) / np.linalg.norm(g)
        print(f"|g - g(sensors off)| / |g| = {share:.3e}", flush=True)
        if not ev:
            s.set_les(False)
            g1, gnu1 = grad(s, JQ, dt, bmat)
            print(f"|g - g(model off)| = {np.abs(g - g1).max():.3e}, nu {abs(gnu - gnu1):.3e}", flush=True)
            assert np.abs(g - g1).max() <= 1e-13 * np.abs(g).max() and abs(gnu - gnu1) <= 1e-13 * abs(gnu)
            return
        assert share > 1e-4
        bm = s.quadrature_weights()

        def Jf(q0, ctx=s):
            ctx.set_state(q0)
            ctx.steps(nsteps, dt, check=False)
            q = ctx.get_state()
            return 0.5 * float(np.sum(bm * (q[:, 1] ** 2 + q[:, 2] ** 2 + q[:, 3] ** 2) / q[:, 0]))

        rng = np.random.default_rng(77)
        h = 1e-6
        for trial in range(3):
            d = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
            fd = (Jf(JQ + h * d) - Jf(JQ - h * d)) / (2 * h)
            ad = float(np.sum(g * d))
            print(f"trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
            assert _rel(ad, fd) < 1e-6
    nu0 = 1.0 / 200.0
    dnu = nu0 * 1e-4
    vals = []
    for nu_s in (nu0 + dnu, nu0 - dnu):
        with make(1.0 / nu_s) as s2:
            vals.append(Jf(JQ, s2))
    fdn = (vals[0] - vals[1]) / (2 * dnu)
    print(f"nu: adjoint {gnu:.12e} fd {fdn:.12e} rel {_rel(gnu, fdn):.2e}", flush=True)
    assert _rel(gnu, fdn) < 1e-5
