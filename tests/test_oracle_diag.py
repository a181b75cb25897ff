"""Pins of the oracle's flow diagnostics: the kinetic energy of the TGV
statistics (caseFiles/tgv/BackstepProblem-stoch.f90:345-361,629-692) and the
enstrophy dissipation (src/Energydissipation/energydissipation.f90:2-89).

* the Gauss-Legendre rule against numpy and against the reference's gaussq;
* bmatu1b against exact integrals of polynomials (interpolatory rule on the
  Chebyshev-Gauss points: exact up to degree p per direction) and against
  its tensor-product structure;
* Ek against closed forms for velocity fields that are polynomial in each
  element, including a non-uniform density (u = (JQ)_2/(JQ)_1);
* eps against closed forms for linear shear and rigid rotation, with the
  lagged staggered states built from the exact fields.
"""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M

from test_oracle_ev import _setup


@pytest.mark.parametrize("n", [1, 2, 5, 9])
def test_gauss_legendre_rule(n):
    t, w = O.gauss_legendre(n)
    tn, wn = np.polynomial.legendre.leggauss(n)
    assert np.allclose(t, tn, atol=4e-16, rtol=0) and np.allclose(w, wn, atol=1e-15, rtol=0)
    if O.ref_available():
        tr, wr = np.zeros(n), np.zeros(n)
        O.ref().ref_gaussq(n, O.ptr(tr), O.ptr(wr))
        order = np.argsort(tr)
        assert np.allclose(t, tr[order], atol=6e-16, rtol=0)
        assert np.allclose(w, wr[order], atol=2e-15, rtol=0)


@pytest.mark.parametrize("p", [1, 3, 4, 8])
def test_ke_weights_integrate_polynomials_exactly(p):
    ncg = p + 1
    b = O.ke_weights(p).reshape(ncg, ncg, ncg)          # [l][j][i]
    _, cxg, _, _ = O.ops(p)
    x = cxg
    for a in range(ncg):
        for c in range(0, ncg, 2):
            for e in (0, p):
                f = (x[None, None, :] ** a) * (x[None, :, None] ** c) * (x[:, None, None] ** e)
                exact = 1.0 / ((a + 1) * (c + 1) * (e + 1))
                assert abs((b * f).sum() - exact) < 5e-15
    # tensor-product structure: bmat(i,j,l) = W_i W_j W_l with W_i = int_0^1 h_i,
    # the weights of Fejer's first rule on the Chebyshev-Gauss points
    # (closed form, e.g. Waldvogel, BIT 46 (2006) 195-202), mapped to [0,1]
    th = (2 * np.arange(ncg) + 1) * np.pi / (2 * ncg)
    wf = np.array([1.0 - 2.0 * sum(np.cos(2 * j * t) / (4 * j * j - 1) for j in range(1, ncg // 2 + 1))
                   for t in th]) / ncg
    xf = 0.5 * (1.0 + np.cos(th))
    W = np.zeros(ncg)
    for i in range(ncg):
        k = np.argmin(np.abs(xf - x[i]))
        assert abs(xf[k] - x[i]) < 1e-15
        W[i] = wf[k]
    assert np.allclose(b, W[None, None, :] * W[None, :, None] * W[:, None, None], rtol=1e-12, atol=1e-16)
    if O.ref_available():
        # the reference's polyn and gaussq in the reference's triple sum
        R = O.ref()
        tr, wr = np.zeros(ncg), np.zeros(ncg)
        R.ref_gaussq(ncg, O.ptr(tr), O.ptr(wr))
        z = np.ascontiguousarray(cxg)
        h = np.array([[R.ref_polyn(i + 1, 0.5 * (tr[k] + 1.0), ncg, O.ptr(z)) for k in range(ncg)]
                      for i in range(ncg)])
        W1 = 0.5 * (h * wr[None, :]).sum(axis=1)
        assert np.allclose(b, W1[None, None, :] * W1[None, :, None] * W1[:, None, None], rtol=1e-13, atol=1e-17)


def _state(g, rho, u, v, w, p):
    Q = cases.conservative(rho, u, v, w, p)                 # [5, K, np]
    return np.ascontiguousarray(np.moveaxis(Q, 0, 1) / g.jacob[:, None, :])


def test_kinetic_energy_closed_form():
    """Velocity linear in x and y (periodicity is irrelevant to the
    quadrature), density varying: Ek = 1/2 <u^2+v^2+w^2> exactly."""
    mesh = M.periodic_box(3, 3)
    g = O.Geometry(mesh)
    X, Y, Z = g.xg[..., 0], g.xg[..., 1], g.xg[..., 2]
    a0, a1, b0, b1, c0 = 0.3, 0.2, -0.1, 0.15, 0.05
    u, v, w = a0 + a1 * X, b0 + b1 * Y, c0 + 0 * X
    rho = 1.0 + 0.2 * np.sin(X) * np.cos(Z)
    JQ = _state(g, rho, u, v, w, 71.43 + 0 * X)
    L = 2 * np.pi
    mean_sq = lambda c0_, c1_: c0_ ** 2 + c0_ * c1_ * L + c1_ ** 2 * L ** 2 / 3
    exact = 0.5 * (mean_sq(a0, a1) + mean_sq(b0, b1) + c0 ** 2)
    assert abs(O.kinetic_energy(g, JQ) - exact) < 1e-13 * exact


def test_kinetic_energy_tgv_converges_to_one_eighth():
    errs = []
    for n, p in [(2, 4), (4, 6)]:
        mesh = M.periodic_box(n, p)
        g = O.Geometry(mesh)
        JQ = cases.tgv_state(g.xg, g.jacob, "reference")
        errs.append(abs(O.kinetic_energy(g, JQ) - 0.125))
    assert errs[1] < 1e-6 and errs[1] < 1e-2 * errs[0]


@pytest.mark.parametrize("alpha", [0.4, 1.3])
def test_enstrophy_shear_and_rotation(alpha):
    re = 1600.0
    ph = O.phys(re=re)
    shear = lambda x, y, z: (np.ones_like(x), alpha * y, 0 * x, 0 * x, 71.43 + 0 * x)
    _, g, JQ, stag = _setup(shear)
    assert abs(O.enstrophy(g, JQ, stag, ph) - 100.0 * alpha ** 2 / re) < 1e-12 * 100.0 * alpha ** 2 / re
    rot = lambda x, y, z: (1.0 + 0.1 * x, -alpha * y, alpha * x, 0.5 * alpha * x, 71.43 + 0 * x)
    _, g, JQ, stag = _setup(rot)
    # omega = (0, -alpha/2, 2 alpha) with a linear density: gradients of rho u are quadratic, still exact
    exact = 100.0 * (4 * alpha ** 2 + 0.25 * alpha ** 2) / re
    assert abs(O.enstrophy(g, JQ, stag, ph) - exact) < 1e-11 * exact


def test_enstrophy_of_tgv_initial_field():
    """<|omega|^2> of the TGV velocity is 3/4; the unweighted point mean over
    Chebyshev-Gauss points converges to it."""
    re = 1600.0
    errs = []
    for n, p in [(2, 4), (4, 6)]:
        mesh = M.periodic_box(n, p)
        g = O.Geometry(mesh)
        JQ = cases.tgv_state(g.xg, g.jacob, "reference")
        _, dbg = O.rhs(g, JQ, O.phys(re=re), debug=True)
        eps = O.enstrophy(g, JQ, np.ascontiguousarray(dbg["stag"]), O.phys(re=re))
        errs.append(abs(eps - 100.0 * 0.75 / re) / (100.0 * 0.75 / re))
    assert errs[1] < 2e-2 and errs[1] < errs[0]


def test_rk_step_keeps_last_stage_states():
    mesh = M.periodic_box(2, 3)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-2, seed=2)
    ph = O.phys(re=500.0)
    dt = 0.3 * O.dtcalc(g, JQ, ph)
    stag = np.zeros((mesh.K, 3, 5, g.ns))
    a = JQ.copy()
    O.rk_step(g, a, dt, ph, stag_out=stag)
    # the staggered states of the last stage equal those of an RHS evaluated
    # on the state entering stage 5: rebuild that state stage by stage
    ark, _, crk = O.rkcoefs()
    q, G = JQ.copy(), np.zeros_like(JQ)
    for k in range(4):
        Qt = O.rhs(g, q, ph)
        G = ark[k] * G + Qt
        q = q + crk[k] * dt * G
    _, dbg = O.rhs(g, q, ph, debug=True)
    assert np.abs(dbg["stag"] - stag).max() <= 1e-13 * np.abs(stag).max()
