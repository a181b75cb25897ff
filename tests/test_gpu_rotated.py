"""GPU RHS and RK steps on meshes with rotated conforming faces (all eight
orientations, rotate_to_mortar / rotate_flux_to_face / rotate_fv_to_mortar,
mortar_operations.f90:461-995) against the oracle, whose maps are pinned by
the geometry and by frame invariance (tests/test_oracle_rotated.py); the
adjoint on such meshes against central differences."""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM
from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p,seed,metrics", [(3, 0, "per_point"), (5, 1, "per_point"), (8, 2, "per_point"),
                                            (7, 3, "per_point"), (4, 2, "per_element")])
def test_rotated_faces_rhs_matches_oracle(p, seed, metrics):
    print(f"\n[case] rotated faces rhs p={p} seed={seed} {metrics}", flush=True)
    mesh = M.rotated_box(4, p, seed=seed)
    assert len(set(mesh.mortar_orient.tolist())) == 8
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-2, seed=seed)
    want = O.rhs(g, JQ, O.phys(re=400.0))
    with DSEM(mesh, re=400.0, metrics=metrics) as s:
        s.set_state(JQ)
        got = s.rhs()
    err = rel_l2(got, want)
    print("rel L2 per field", err, flush=True)
    assert np.all(err <= (1e-12 if metrics == "per_point" else 5e-12))


def test_rotated_faces_rk_steps_match_oracle():
    print("\n[case] rotated faces 3 RK steps p=6", flush=True)
    mesh = M.rotated_box(4, 6, seed=5)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-2, seed=2)
    ph = O.phys(re=400.0)
    dt = 0.5 * O.dtcalc(g, JQ, ph)
    ref = JQ.copy()
    for _ in range(3):
        _, _, r_ref = O.rk_step(g, ref, dt, ph)
    with DSEM(mesh, re=400.0) as s:
        s.set_state(JQ)
        for _ in range(3):
            r = s.tstep(0.0, dt)
        got = s.get_state()
    err = rel_l2(got - JQ, ref - JQ)
    print("state change rel L2 per field", err, "resid", r, r_ref, flush=True)
    assert np.all(err <= 1e-12)
    assert abs(r - r_ref) <= 1e-10 * abs(r_ref)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("p,seed", [(4, 1), (7, 3)])
def test_rhs_vjp_on_rotated_faces_matches_central_differences(p, seed):
    """The adjoint's face kernels follow the forward's face point maps: on a
    mesh with all eight orientations, <(dR/dJQ)^T lam, v> against
    <lam, (R(JQ + h v) - R(JQ - h v)) / 2h>, and the nu adjoint against R
    at Re shifted both ways."""
    print(f"\n[case] rhs vjp rotated faces p={p}", flush=True)
    mesh = M.rotated_box(4, p, seed=seed)
    assert len(set(mesh.mortar_orient.tolist())) == 8
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=seed)
    rng = np.random.default_rng(17)
    re = 400.0
    with DSEM(mesh, re=re) as s:
        s.set_state(JQ)
        for trial in range(2):
            lam = rng.standard_normal(JQ.shape)
            v = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
            jb, nb = s.rhs_vjp(lam)
            h = 1e-6
            s.set_state(JQ + h * v)
            rp = s.rhs()
            s.set_state(JQ - h * v)
            rm = s.rhs()
            s.set_state(JQ)
            fd = float(np.sum(lam * (rp - rm)) / (2 * h))
            ad = float(np.sum(jb * v))
            # error against the size of the terms of the dot product: a random
            # draw can make the sum itself small (cancellation), which would
            # turn the differences' own truncation error into a large ratio
            scale = float(np.sum(np.abs(lam * (rp - rm))) / (2 * h))
            print(f"trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e} "
                  f"vs terms {abs(ad - fd) / scale:.2e}", flush=True)
            assert abs(ad - fd) < 1e-7 * scale
    nu, dnu = 1.0 / re, 1e-6 / re
    out = []
    for n in (nu + dnu, nu - dnu):
        with DSEM(mesh, re=1.0 / n) as s:
            s.set_state(JQ)
            out.append(s.rhs())
    fdn = float(np.sum(lam * (out[0] - out[1])) / (2 * dnu))
    print(f"nu: adjoint {nb:.12e} fd {fdn:.12e}", flush=True)
    assert _rel(nb, fdn) < 1e-6


def test_rk_step_vjp_on_rotated_faces_matches_central_differences():
    print("\n[case] rk step vjp rotated faces p=5", flush=True)
    mesh = M.rotated_box(4, 5, seed=2)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=4)
    dt = 0.4 * O.dtcalc(g, JQ, O.phys(re=500.0))
    rng = np.random.default_rng(9)
    qb = rng.standard_normal(JQ.shape)
    v = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
    with DSEM(mesh, re=500.0) as s:
        s.set_state(JQ)
        q0b, _ = s.rk_step_vjp(dt, qb)
        h = 1e-6
        s.set_state(JQ + h * v)
        s.tstep(0.0, dt)
        qp = s.get_state()
        s.set_state(JQ - h * v)
        s.tstep(0.0, dt)
        qm = s.get_state()
    fd = float(np.sum(qb * (qp - qm)) / (2 * h))
    ad = float(np.sum(q0b * v))
    print(f"adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
    assert _rel(ad, fd) < 1e-6


def test_shock_run_gradient_on_rotated_faces():
    """Entropy viscosity on: a 3-step solve through torch.autograd (step
    tape, lagged states' adjoint) on a mesh with all eight orientations,
    dJ/dJQ0 along a seeded direction against central differences."""
    import torch
    from dsem_amd import autograd as A
    print("\n[case] shock run gradient rotated faces p=4", flush=True)
    mesh = M.rotated_box(4, 4, seed=6)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=8)
    dt = 0.4 * O.dtcalc(g, JQ, O.phys(re=300.0))
    rng = np.random.default_rng(21)
    v = rng.standard_normal(JQ.shape) * np.abs(JQ).mean(axis=(0, 2), keepdims=True)
    with DSEM(mesh, re=300.0, shock=True, cmu=1.0, mumax=0.02, ducros=True, smoothing=True) as s:
        bmat = A.bmat_tensor(s, "cuda")
        x = torch.tensor(JQ, device="cuda", requires_grad=True)
        J = A.kinetic_energy_integral(A.solve(s, x, 3, dt, every=3), bmat)
        J.backward()
        ad = float((x.grad.cpu().numpy() * v).sum())
        h = 1e-6
        vals = []
        for sgn in (1.0, -1.0):
            y = torch.tensor(JQ + sgn * h * v, device="cuda")
            vals.append(float(A.kinetic_energy_integral(A.solve(s, y, 3, dt, every=3), bmat)))
    fd = (vals[0] - vals[1]) / (2 * h)
    print(f"adjoint {ad:.12e} fd {fd:.12e} rel {_rel(ad, fd):.2e}", flush=True)
    assert _rel(ad, fd) < 1e-6


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_rotated_faces_mixed_precision_rhs(precision):
    print(f"\n[case] rotated faces rhs p=6 {precision}", flush=True)
    mesh = M.rotated_box(4, 6, seed=4)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-2, seed=9)
    want = O.rhs(g, JQ, O.phys(re=400.0))
    with DSEM(mesh, re=400.0, precision=precision) as s:
        s.set_state(JQ)
        got = s.rhs()
    err = rel_l2(got, want)
    print("rel L2 per field", err, flush=True)
    assert np.all(err <= (1e-5 if precision == "fp32" else 1e-3))


@pytest.mark.parametrize("p", [4, 7])
def test_rotated_faces_entropy_viscosity_steps_match_oracle(p):
    """Entropy-viscosity steps (Ducros sensor, smoothing, mumax cap) on a mesh
    whose conforming faces take all eight orientations: the lagged staggered
    states reach k_ev through the rotated face maps of k_volume."""
    print(f"\n[case] rotated faces 3 EV steps p={p}", flush=True)
    mesh = M.rotated_box(4, p, seed=4)
    assert len(set(mesh.mortar_orient.tolist())) == 8
    g = O.Geometry(mesh)
    JQ0 = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=6)
    ph = O.phys(re=300.0)
    dt = 0.4 * O.dtcalc(g, JQ0, ph)
    ev = O.OrEv(1.0, 0.02, 1, 1)
    ref = JQ0.copy()
    stag = np.zeros((mesh.K, 3, 5, g.ns))
    muh_r = None
    for k in range(3):
        ref, muh_r, _ = O.rk_step_ev(g, ref, dt, stag, k > 0, ev, ph)
    with DSEM(mesh, re=300.0, shock=True, cmu=1.0, mumax=0.02, ducros=True, smoothing=True,
              metrics="per_point") as s:
        s.set_state(JQ0)
        for _ in range(3):
            s.tstep(0.0, dt)
        got = s.get_state()
        _, muh = s.get_viscosity()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    emu = np.abs(muh - muh_r).max() / np.abs(muh_r).max()
    print(f"state rel max {err:.2e}, mu_h rel max {emu:.2e}", flush=True)
    assert np.abs(muh_r).max() > 0
    assert emu < 1e-11
    assert err < 1e-13
