"""GPU entropy viscosity against the oracle: the Entropy step alone (given the
same lagged staggered states) and whole EV time steps (mu_h from the previous
step's last stage, zero on a fresh start)."""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM
from helpers import distorted_box, rel_l2

pytestmark = pytest.mark.gpu


def _lagged(mesh, g, JQ, ph):
    """Staggered states after face averaging, as the last stage leaves them."""
    _, dbg = O.rhs(g, JQ, ph, debug=True)
    return np.ascontiguousarray(dbg["stag"])


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("ducros,smoothing", [(False, False), (True, False), (False, True), (True, True)])
def test_entropy_step_matches_oracle(ducros, smoothing):
    print(f"\n[case] ev update ducros={ducros} smoothing={smoothing}", flush=True)
    mesh = M.periodic_box(3, 5)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=5)
    ph = O.phys(re=400.0)
    stag = _lagged(mesh, g, JQ, ph)
    ev = O.OrEv(0.8, 0.05, int(ducros), int(smoothing))
    mu_r, muh_r, res_r, duc_r = O.ev_viscosity(g, JQ, stag, ev, ph)
    with DSEM(mesh, re=400.0, shock=True, cmu=0.8, mumax=0.05, ducros=ducros, smoothing=smoothing,
              metrics="per_point") as s:
        s.set_state(JQ)
        s.set_lagged_states(stag)
        mu, res, duc = s.ev_update()
        _, muh = s.get_viscosity()
    print("rel errors", _rel(res, res_r), _rel(duc, duc_r), _rel(mu, mu_r), _rel(muh, muh_r), flush=True)
    assert _rel(res, res_r) < 1e-12 and _rel(duc, duc_r) < 1e-12
    assert _rel(mu, mu_r) < 1e-12 and _rel(muh, muh_r) < 1e-12
    assert 0 < np.count_nonzero(mu_r == 0.05) < mu_r.size   # the cap is active somewhere, not everywhere


def test_ev_time_steps_match_oracle():
    print("\n[case] 3 EV steps n=3 p=4", flush=True)
    mesh = M.periodic_box(3, 4)
    g = O.Geometry(mesh)
    JQ0 = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=9)
    ph = O.phys(re=300.0)
    dt = 0.4 * O.dtcalc(g, JQ0, ph)
    ev = O.OrEv(1.0, 0.02, 0, 1)
    ref = JQ0.copy()
    stag = np.zeros((mesh.K, 3, 5, g.ns))
    muh_r = None
    for k in range(3):
        ref, muh_r, _ = O.rk_step_ev(g, ref, dt, stag, k > 0, ev, ph)
    with DSEM(mesh, re=300.0, shock=True, cmu=1.0, mumax=0.02, smoothing=True, metrics="per_point") as s:
        s.set_state(JQ0)
        for _ in range(3):
            s.tstep(0.0, dt)
        got = s.get_state()
        _, muh = s.get_viscosity()
        lag = s.get_lagged_states()
    assert np.abs(muh_r).max() > 0
    assert _rel(muh, muh_r) < 1e-11
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13
    assert np.abs(lag - stag).max() / np.abs(stag).max() < 1e-13


def test_entropy_step_p1_smoothing_in_reference_order():
    """p = 1: every Gauss point is a corner and each neighbours three others,
    so the in-place order of the corner updates decides the result."""
    print("\n[case] ev update p=1 smoothing", flush=True)
    mesh = distorted_box(3, 1, amp=0.05, seed=4)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=5)
    ph = O.phys(re=400.0)
    stag = _lagged(mesh, g, JQ, ph)
    ev = O.OrEv(0.8, 1e9, 0, 1)
    mu_r, muh_r, _, _ = O.ev_viscosity(g, JQ, stag, ev, ph)
    with DSEM(mesh, re=400.0, shock=True, cmu=0.8, mumax=1e9, smoothing=True) as s:
        s.set_state(JQ)
        s.set_lagged_states(stag)
        mu, _, _ = s.ev_update()
        _, muh = s.get_viscosity()
    print("rel errors", _rel(mu, mu_r), _rel(muh, muh_r), flush=True)
    assert _rel(mu, mu_r) < 1e-12 and _rel(muh, muh_r) < 1e-12


@pytest.mark.parametrize("p", [4, 6])
def test_cmax_cap_matches_oracle(p):
    """The c_max cap (EViscosity_GL.f90:82, used by the isotropic-turbulence
    case) against the oracle, on a mesh where it binds at some points only."""
    print(f"\n[case] ev cmax cap p={p}", flush=True)
    mesh = distorted_box(3, p, amp=0.05, seed=8)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=3e-2, seed=7)
    ph = O.phys(re=400.0)
    stag = _lagged(mesh, g, JQ, ph)
    ev = O.OrEv(30.0, 1e9, 1, 1)
    ev.cmax_cap, ev.cmax = 1, 0.5
    mu_r, muh_r, _, _ = O.ev_viscosity(g, JQ, stag, ev, ph)
    ev0 = O.OrEv(30.0, 1e9, 1, 0)
    ev0.cmax_cap, ev0.cmax = 1, 0.5
    capped0, _, _, _ = O.ev_viscosity(g, JQ, stag, ev0, ph)
    uncapped0, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(30.0, 1e9, 1, 0), ph)
    frac = np.mean(capped0 < uncapped0)
    print(f"cap binds at {frac:.3f} of the points", flush=True)
    assert 0.01 < frac < 0.5
    with DSEM(mesh, re=400.0, shock=True, cmu=30.0, mumax=1e9, cmax=0.5, ducros=True, smoothing=True,
              cmax_cap=True) as s:
        s.set_state(JQ)
        s.set_lagged_states(stag)
        mu, _, _ = s.ev_update()
        _, muh = s.get_viscosity()
    assert _rel(mu, mu_r) < 1e-12 and _rel(muh, muh_r) < 1e-12
