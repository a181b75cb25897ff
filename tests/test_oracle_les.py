"""Pins of the oracle's static Smagorinsky model (src/turbulence/staticSmago/
staticSmag.f90; the file needs the `mpi` module and does not compile here).

Lagged staggered states from exact linear velocity fields (for which the
scheme's Gauss-point gradients are exact) give nu_t in closed form:
S = (2 S_ij S_ij)^(1/2), nu_t = Re (C_s Delta)^2 S, Delta = h/ncg on a box of
element size h; the cap numax and the prolongation of a constant are exact.
The eddy viscosity enters the RHS as amu = 1 + nu_t rho, which is checked
against an RHS with the equivalent entropy-viscosity field mu_h = nu_t rho.
"""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M

from test_oracle_ev import _setup, P0


@pytest.mark.parametrize("alpha", [0.4, 1.7])
def test_shear_closed_form(alpha):
    fn = lambda x, y, z: (np.ones_like(x), alpha * y, 0 * x, 0 * x, P0 + 0 * x)
    mesh, g, JQ, stag = _setup(fn)
    re, cs = 900.0, 0.17
    nut_g, nut_s = O.smagorinsky(g, JQ, stag, cs, 1e9, O.phys(re=re))
    h = 2 * np.pi / mesh.K ** (1 / 3)
    delta = h / (mesh.p + 1)
    expect = re * (cs * delta) ** 2 * abs(alpha)
    assert np.allclose(nut_g, expect, rtol=1e-10)
    assert np.allclose(nut_s, expect, rtol=1e-10)


def test_strain_of_mixed_linear_field():
    """u = (a y + b z, c x, d x + e y): S_ij from the symmetric part."""
    a, b, c, d, e = 0.3, -0.2, 0.5, 0.1, 0.7
    fn = lambda x, y, z: (np.ones_like(x), a * y + b * z, c * x, d * x + e * y, P0 + 0 * x)
    mesh, g, JQ, stag = _setup(fn)
    gradu = np.array([[0, a, b], [c, 0, 0], [d, e, 0]], dtype=float)
    Sij = 0.5 * (gradu + gradu.T)
    Sbar = np.sqrt(2 * np.sum(Sij * Sij))
    re, cs = 500.0, 0.2
    nut_g, _ = O.smagorinsky(g, JQ, stag, cs, 1e9, O.phys(re=re))
    delta = (2 * np.pi / mesh.K ** (1 / 3)) / (mesh.p + 1)
    assert np.allclose(nut_g, re * (cs * delta) ** 2 * Sbar, rtol=1e-10)


def test_cap_and_rhs_coupling():
    fn = lambda x, y, z: (np.ones_like(x), 1.1 * y, 0 * x, 0 * x, P0 + 0 * x)
    mesh, g, JQ, stag = _setup(fn)
    nut_g, nut_s = O.smagorinsky(g, JQ, stag, 10.0, 0.05, O.phys(re=1600.0))
    assert np.all(nut_g == 0.05) and np.allclose(nut_s, 0.05, rtol=1e-13)
    # amu = 1 + nu_t rho: the same RHS as an entropy viscosity mu_h = nu_t * rho_stag
    rng = np.random.default_rng(4)
    mesh = M.periodic_box(2, 4)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-2, seed=3)
    ph = O.phys(re=400.0)
    nut = np.abs(rng.standard_normal((mesh.K, 3, g.ns))) * 0.05
    _, dbg = O.rhs(g, JQ, ph, debug=True)
    rho_stag = dbg["stag"][:, :, 0, :]
    a = O.rhs_les(g, JQ, nut, ph)
    b = O.rhs(g, JQ, ph, muh=nut * rho_stag)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    assert np.abs(a - O.rhs(g, JQ, ph)).max() > 1e-6 * np.abs(b).max()   # the model acts


def test_sensor_switches():
    """shockSensor multiplies nu_t by 1 - ducros and rhoSensor by the density
    sensor, before the clamp (staticSmag.f90:67-73).  Closed forms: a pure
    shear has ducros = 0 (unchanged nu_t); an isotropic dilatation has
    ducros = 9a^2/(9a^2 + 1e-9) (nu_t almost removed)."""
    re, cs = 700.0, 0.17
    shear = lambda x, y, z: (np.ones_like(x), 0.8 * y, 0 * x, 0 * x, P0 + 0 * x)
    mesh, g, JQ, stag = _setup(shear)
    _, _, _, duc, rs = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0), rho_sensor=True)
    base, _ = O.smagorinsky(g, JQ, stag, cs, 1e9, O.phys(re=re))
    sh, _ = O.smagorinsky(g, JQ, stag, cs, 1e9, O.phys(re=re), O.LES_SHOCK_SENSOR, duc, rs)
    assert np.array_equal(sh, base)                      # ducros = 0 exactly for a shear
    alpha = 0.3
    dil = lambda x, y, z: (np.ones_like(x), alpha * x, alpha * y, alpha * z, P0 + 0 * x)
    mesh, g, JQ, stag = _setup(dil)
    _, _, _, duc, rs = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0), rho_sensor=True)
    base, _ = O.smagorinsky(g, JQ, stag, cs, 1e9, O.phys(re=re))
    sh, _ = O.smagorinsky(g, JQ, stag, cs, 1e9, O.phys(re=re), O.LES_SHOCK_SENSOR, duc, rs)
    d9 = (3 * alpha) ** 2
    assert np.allclose(sh, base * (1.0 - d9 / (d9 + 1e-9)), rtol=1e-6)
    # both switches on a general field: the product of the two factors, then the clamp
    fn = lambda x, y, z: (1.0 + 0.1 * np.sin(x), 1.3 * np.sin(y) * np.cos(z), 0.2 * np.cos(x), 0.1 * np.sin(z),
                          P0 + np.cos(z))
    mesh, g, JQ, stag = _setup(fn, n=2, p=5)
    _, _, _, duc, rs = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0), rho_sensor=True)
    base, _ = O.smagorinsky(g, JQ, stag, cs, 1e9, O.phys(re=re))
    both, both_s = O.smagorinsky(g, JQ, stag, cs, 1e9, O.phys(re=re), O.LES_SHOCK_SENSOR | O.LES_RHO_SENSOR, duc, rs)
    assert np.allclose(both, np.clip(base * (1.0 - duc) * rs, 0, 1e9), rtol=1e-13, atol=0)
    assert np.ptp(rs) > 0 and np.ptp(duc) > 0
    capped, _ = O.smagorinsky(g, JQ, stag, cs, 1e-7, O.phys(re=re), O.LES_RHO_SENSOR, duc, rs)
    assert np.allclose(capped, np.clip(base * rs, 0, 1e-7), rtol=1e-13, atol=0)
