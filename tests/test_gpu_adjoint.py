"""Adjoint (reverse-mode) kernels against central finite differences of the
forward GPU path and of the oracle.

Dot-product form: for random cotangents lam and directions v,
<(dR/dJQ)^T lam, v> = <lam, (R(JQ + h v) - R(JQ - h v)) / 2h>, and likewise
for nu = 1/Re, for one RK step, and for J = int 1/2 rho |u|^2 after a run of
steps (two-level checkpointed adjoint through torch.autograd).
"""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM
from helpers import distorted_box

pytestmark = pytest.mark.gpu


def _state(mesh, seed, perturb=2e-2):
    g = O.Geometry(mesh)
    return g, cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=perturb, seed=seed)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("metrics", ["per_point", "per_element"])
def test_rhs_vjp_matches_central_differences(metrics):
    print(f"\n[case] rhs vjp distorted 3^3 p=4 {metrics}", flush=True)
    mesh = distorted_box(3, 4, amp=0.06, seed=5) if metrics == "per_point" else M.periodic_box(3, 4)
    _, JQ = _state(mesh, 11)
    rng = np.random.default_rng(7)
    with DSEM(mesh, re=400.0, metrics=metrics) as s:
        s.set_state(JQ)
        for trial in range(3):
            lam = rng.standard_normal(JQ.shape)
            v = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
            jb, _ = s.rhs_vjp(lam)
            h = 1e-6
            s.set_state(JQ + h * v)
            rp = s.rhs()
            s.set_state(JQ - h * v)
            rm = s.rhs()
            s.set_state(JQ)
            fd = float(np.sum(lam * (rp - rm)) / (2 * h))
            ad = float(np.sum(jb * v))
            print(f"trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
            assert _rel(ad, fd) < 1e-6


@pytest.mark.parametrize("p", [5, 6, 7, 8])
def test_rhs_vjp_high_orders(p):
    """Orders 7 and 8 run the forward through the persistent volume kernel
    (1024 work-items, next element copied into LDS); the adjoint must agree
    with central differences there too, and at 5, 6 beside them."""
    print(f"\n[case] rhs vjp distorted 2^3 p={p}", flush=True)
    mesh = distorted_box(2, p, amp=0.05, seed=p)
    _, JQ = _state(mesh, 20 + p)
    rng = np.random.default_rng(p)
    with DSEM(mesh, re=400.0) as s:
        s.set_state(JQ)
        for trial in range(2):
            lam = rng.standard_normal(JQ.shape)
            v = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
            jb, _ = s.rhs_vjp(lam)
            h = 1e-6
            s.set_state(JQ + h * v)
            rp = s.rhs()
            s.set_state(JQ - h * v)
            rm = s.rhs()
            s.set_state(JQ)
            fd = float(np.sum(lam * (rp - rm)) / (2 * h))
            ad = float(np.sum(jb * v))
            print(f"trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
            assert _rel(ad, fd) < 1e-6


def test_rhs_vjp_field_by_field():
    """Directions confined to one conserved field at a time, so that no
    field's adjoint can hide behind another's."""
    print("\n[case] rhs vjp per field 2^3 p=3", flush=True)
    mesh = distorted_box(2, 3, amp=0.05, seed=3)
    _, JQ = _state(mesh, 4)
    rng = np.random.default_rng(9)
    lam = rng.standard_normal(JQ.shape)
    with DSEM(mesh, re=250.0, metrics="per_point") as s:
        s.set_state(JQ)
        jb, _ = s.rhs_vjp(lam)
        for f in range(5):
            v = np.zeros_like(JQ)
            v[:, f] = rng.standard_normal(v[:, f].shape) * np.abs(JQ[:, f]).mean()
            h = 1e-6
            s.set_state(JQ + h * v)
            rp = s.rhs()
            s.set_state(JQ - h * v)
            rm = s.rhs()
            s.set_state(JQ)
            fd = float(np.sum(lam * (rp - rm)) / (2 * h))
            ad = float(np.sum(jb * v))
            assert _rel(ad, fd) < 1e-6, (f, ad, fd)


def test_rhs_vjp_transonic_states():
    """Sonic flow: u close to +c and v close to -c, so that on x faces the
    Roe flux takes the qtd >= 0 branch and on y faces the qtd < 0 branch,
    each with its eigenvalue (qtd -+ atd) near zero, where the entropy fix
    switches on.  TGV states at Mach 0.1 never reach that fix."""
    print("\n[case] rhs vjp transonic 3^3 p=3", flush=True)
    mesh = distorted_box(3, 3, amp=0.04, seed=6)
    g = O.Geometry(mesh)
    gamma = 1.4
    rng = np.random.default_rng(21)
    shp = g.jacob.shape
    rho = 1.0 + 0.1 * rng.standard_normal(shp)
    p = 1.0 / gamma * (1.0 + 0.1 * rng.standard_normal(shp))
    c = np.sqrt(gamma * p / rho)
    u = 1.0 + 0.15 * rng.standard_normal(shp)
    v = -1.0 + 0.15 * rng.standard_normal(shp)
    w = 0.15 * rng.standard_normal(shp)
    # a sizeable share of points lies within 5% of sonic along each direction
    assert np.mean(np.abs(u / c - 1.0) < 0.05) > 0.1
    assert np.mean(np.abs(v / c + 1.0) < 0.05) > 0.1
    Q = np.ascontiguousarray(np.transpose(cases.conservative(rho, u, v, w, p, gamma), (1, 0, 2)))
    JQ = Q / g.jacob[:, None, :]
    with DSEM(mesh, re=200.0, metrics="per_point") as s:
        s.set_state(JQ)
        for trial in range(3):
            lam = rng.standard_normal(JQ.shape)
            dv = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
            jb, _ = s.rhs_vjp(lam)
            h = 1e-7
            s.set_state(JQ + h * dv)
            rp = s.rhs()
            s.set_state(JQ - h * dv)
            rm = s.rhs()
            s.set_state(JQ)
            fd = float(np.sum(lam * (rp - rm)) / (2 * h))
            ad = float(np.sum(jb * dv))
            print(f"trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
            assert _rel(ad, fd) < 1e-6


def test_nu_adjoint_matches_difference_in_reynolds_number():
    print("\n[case] nu adjoint 3^3 p=4", flush=True)
    mesh = M.periodic_box(3, 4)
    _, JQ = _state(mesh, 2)
    lam = np.random.default_rng(1).standard_normal(JQ.shape)
    nu, dnu = 1.0 / 300.0, 1e-6 / 300.0
    with DSEM(mesh, re=300.0) as s:
        s.set_state(JQ)
        _, nb = s.rhs_vjp(lam)
    out = []
    for n in (nu + dnu, nu - dnu):
        with DSEM(mesh, re=1.0 / n) as s:
            s.set_state(JQ)
            out.append(s.rhs())
    fd = float(np.sum(lam * (out[0] - out[1])) / (2 * dnu))
    print(f"nubar {nb:.12e} fd {fd:.12e}", flush=True)
    assert _rel(nb, fd) < 1e-6


def test_rk_step_vjp_matches_central_differences():
    print("\n[case] rk step vjp distorted 3^3 p=4", flush=True)
    mesh = distorted_box(3, 4, amp=0.05, seed=8)
    g, JQ = _state(mesh, 3)
    dt = 0.4 * O.dtcalc(g, JQ, O.phys(re=500.0))
    rng = np.random.default_rng(5)
    qb = rng.standard_normal(JQ.shape)
    v = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
    with DSEM(mesh, re=500.0, metrics="per_point") as s:
        s.set_state(JQ)
        q0b, nb = s.rk_step_vjp(dt, qb)
        assert np.array_equal(s.get_state(), JQ)      # the start state is left in place
        h = 1e-6
        s.set_state(JQ + h * v)
        s.tstep(0.0, dt)
        qp = s.get_state()
        s.set_state(JQ - h * v)
        s.tstep(0.0, dt)
        qm = s.get_state()
    fd = float(np.sum(qb * (qp - qm)) / (2 * h))
    ad = float(np.sum(q0b * v))
    print(f"adjoint {ad:.12e} fd {fd:.12e}", flush=True)
    assert _rel(ad, fd) < 1e-6
    # nu: one step at Re shifted both ways
    out = []
    nu, dnu = 1.0 / 500.0, 1e-6 / 500.0
    for n in (nu + dnu, nu - dnu):
        with DSEM(mesh, re=1.0 / n, metrics="per_point") as s:
            s.set_state(JQ)
            s.tstep(0.0, dt)
            out.append(s.get_state())
    fdn = float(np.sum(qb * (out[0] - out[1])) / (2 * dnu))
    assert _rel(nb, fdn) < 1e-6


def test_adjoint_against_oracle_finite_differences():
    """The gradient of J after a few steps against central differences
    through the CPU oracle (the restated reference algorithm)."""
    import torch
    from dsem_amd import autograd as A
    print("\n[case] adjoint vs oracle FD 2^3 p=4, 4 steps", flush=True)
    mesh = M.periodic_box(2, 4)
    g, JQ = _state(mesh, 6, perturb=1e-2)
    ph = O.phys(re=800.0)
    dt = 0.4 * O.dtcalc(g, JQ, ph)
    nsteps = 4
    with DSEM(mesh, re=800.0, metrics="per_point") as s:
        bmat = A.bmat_tensor(s, "cuda")
        x = torch.tensor(JQ, device="cuda", requires_grad=True)
        J = A.kinetic_energy_integral(A.solve(s, x, nsteps, dt, every=2), bmat)
        J.backward()
        grad = x.grad.cpu().numpy()
    bm = O.ke_weights(mesh.p)
    Jf = lambda q: 0.5 * float(np.sum(bm * (q[:, 1] ** 2 + q[:, 2] ** 2 + q[:, 3] ** 2) / q[:, 0]))
    rng = np.random.default_rng(2024)
    for _ in range(2):
        d = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
        h = 1e-6
        vals = []
        for sgn in (1.0, -1.0):
            q = JQ + sgn * h * d
            for _ in range(nsteps):
                O.rk_step(g, q, dt, ph)
            vals.append(Jf(q))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = float(np.sum(grad * d))
        print(f"adjoint {ad:.12e} oracle fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
        assert _rel(ad, fd) < 1e-6


def test_taylor_green_adjoint_gradient_8x8x8_p6():
    """J = int 1/2 rho |u|^2 after 50 steps (8^3 hex8, p = 6, Re = 1600),
    checkpoint every 10 steps: dJ/du0 along 32 seeded directions and dJ/dnu
    against central differences."""
    from dsem_amd import adjoint_study
    print("\n[case] adjoint study 8^3 p=6 50 steps", flush=True)
    r = adjoint_study.run(n=8, p=6, steps=50, every=10, ndir=32)
    print({k: r[k] for k in ("forward_ms", "backward_ms", "adjoint_over_forward")}, r["velocity_directions"]["max_rel_err"],
          r["nu"], flush=True)
    assert r["velocity_directions"]["max_rel_err"] < 1e-6
    assert r["nu"]["rel_err"] < 1e-6


def test_binomial_schedules_give_the_uniform_gradient():
    """Memory-constrained binomial schedules (few snapshots, short tape)
    recompute states by plain steps; their gradients equal those of the
    uniform schedule (snapshot and tape of `every` steps)."""
    import torch
    from dsem_amd import autograd as A
    from dsem_amd.checkpoint import Plan
    print("\n[case] binomial checkpoint schedules 2^3 p=4, 12 steps", flush=True)
    mesh = M.periodic_box(2, 4)
    g, JQ = _state(mesh, 7, perturb=1e-2)
    dt = 0.4 * O.dtcalc(g, JQ, O.phys(re=800.0))
    n = 12
    grads = {}
    with DSEM(mesh, re=800.0, metrics="per_point") as s:
        bmat = A.bmat_tensor(s, "cuda")
        for key in [(None, None), (2, 1), (1, 3), (0, 2), (3, 2)]:
            snaps, tape = key
            x = torch.tensor(JQ, device="cuda", requires_grad=True)
            nu = torch.tensor(1.0 / 800.0, dtype=torch.float64, device="cuda", requires_grad=True)
            J = A.kinetic_energy_integral(A.solve(s, x, n, dt, every=4, nu=nu, snaps=snaps, tape=tape), bmat)
            J.backward()
            grads[key] = (x.grad.cpu().numpy(), float(nu.grad))
            if snaps is not None:
                print(f"snaps {snaps} tape {tape}: {Plan(n, snaps, tape).recomputed_steps()} recomputed steps",
                      flush=True)
    ref_x, ref_nu = grads[(None, None)]
    for key, (gx, gnu) in grads.items():
        assert np.linalg.norm(gx - ref_x) <= 1e-13 * np.linalg.norm(ref_x), key
        assert abs(gnu - ref_nu) <= 1e-12 * abs(ref_nu), key


def test_second_solve_before_backward_keeps_the_gradient():
    """The forward pass records the solve's last chunk on the context's tape.
    A second solve on the same context before the first one's backward
    replaces that recording; the first backward then records it again and
    its gradient is unchanged."""
    import torch
    from dsem_amd import autograd as A
    print("\n[case] interleaved solves 2^3 p=4, 8 steps", flush=True)
    mesh = M.periodic_box(2, 4)
    g, JQ = _state(mesh, 3, perturb=1e-2)
    _, JQ2 = _state(mesh, 4, perturb=1e-2)
    dt = 0.4 * O.dtcalc(g, JQ, O.phys(re=800.0))
    with DSEM(mesh, re=800.0, metrics="per_point") as s:
        bmat = A.bmat_tensor(s, "cuda")

        def grad(interleave):
            x = torch.tensor(JQ, device="cuda", requires_grad=True)
            J = A.kinetic_energy_integral(A.solve(s, x, 8, dt, every=4), bmat)
            if interleave:
                y = torch.tensor(JQ2, device="cuda", requires_grad=True)
                A.kinetic_energy_integral(A.solve(s, y, 8, dt, every=4), bmat)
            J.backward()
            return x.grad.cpu().numpy()

        alone, inter = grad(False), grad(True)
    assert np.linalg.norm(inter - alone) <= 1e-13 * np.linalg.norm(alone)
