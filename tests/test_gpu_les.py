"""GPU static Smagorinsky model against the oracle: the eddy viscosity from
given lagged staggered states, the RHS with a prescribed eddy viscosity, and
whole steps (nu_t from the previous step's last stage, zero on a fresh
start)."""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM
from helpers import distorted_box

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n,p", [(3, 1), (3, 5), (2, 8)])
def test_eddy_viscosity_matches_oracle(n, p):
    print(f"\n[case] smagorinsky nu_t distorted {n}^3 p={p}", flush=True)
    mesh = distorted_box(n, p, amp=0.06, seed=6)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=5)
    ph = O.phys(re=800.0)
    _, dbg = O.rhs(g, JQ, ph, debug=True)
    stag = np.ascontiguousarray(dbg["stag"])
    numax = float(np.median(O.smagorinsky(g, JQ, stag, 0.17, 1e9, ph)[0]))
    ng_r, ns_r = O.smagorinsky(g, JQ, stag, 0.17, numax, ph)
    with DSEM(mesh, re=800.0, metrics="per_point") as s:
        s.set_state(JQ)
        s.set_les(True, 0.17, numax)
        s.set_lagged_states(stag)
        ng, ns = s.les_update()
    assert 0 < np.count_nonzero(ng_r == numax) < ng_r.size      # the cap acts at about half the points
    assert _rel(ng, ng_r) < 1e-12 and _rel(ns, ns_r) < 1e-12


@pytest.mark.parametrize("n,p", [(3, 4), (2, 8)])
def test_rhs_with_eddy_viscosity_matches_oracle(n, p):
    """p = 4 and p = 8 (1024 work-items per element, 151 KB of LDS)."""
    print(f"\n[case] rhs with nu_t {n}^3 p={p}", flush=True)
    mesh = M.periodic_box(n, p)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-2, seed=2)
    ph = O.phys(re=600.0)
    nut = np.abs(np.random.default_rng(1).standard_normal((mesh.K, 3, g.ns))) * 0.03
    ref = O.rhs_les(g, JQ, nut, ph)
    with DSEM(mesh, re=600.0, metrics="per_point") as s:
        s.set_state(JQ)
        s.set_les(True, 0.17, 1.0)
        s.set_les_viscosity(nut)
        got = s.rhs()
    for v in range(5):
        assert _rel(got[:, v], ref[:, v]) < 1e-12


def test_les_steps_match_oracle():
    print("\n[case] 3 LES steps 3^3 p=4", flush=True)
    mesh = M.periodic_box(3, 4)
    g = O.Geometry(mesh)
    JQ0 = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=9)
    ph = O.phys(re=1600.0)
    dt = 0.4 * O.dtcalc(g, JQ0, ph)
    ref = JQ0.copy()
    stag = np.zeros((mesh.K, 3, 5, g.ns))
    nut_r = None
    for k in range(3):
        ref, nut_r, _ = O.rk_step_les(g, ref, dt, stag, k > 0, 0.2, 0.05, ph)
    with DSEM(mesh, re=1600.0, metrics="per_point") as s:
        s.set_state(JQ0)
        s.set_les(True, 0.2, 0.05)
        for _ in range(3):
            s.tstep(0.0, dt)
        got = s.get_state()
    assert np.abs(nut_r).max() > 0
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize("sensors", [O.LES_SHOCK_SENSOR, O.LES_RHO_SENSOR, O.LES_SHOCK_SENSOR | O.LES_RHO_SENSOR])
def test_sensor_switches_match_oracle(sensors):
    """shockSensor / rhoSensor (staticSmag.f90:68-69): the sensors come from the
    entropy-viscosity update of the same lagged states (tstep runs Entropy
    before simpleSmag)."""
    print(f"\n[case] smagorinsky sensors={sensors} distorted 3^3 p=5", flush=True)
    mesh = distorted_box(3, 5, amp=0.06, seed=6)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=5)
    ph = O.phys(re=800.0)
    _, dbg = O.rhs(g, JQ, ph, debug=True)
    stag = np.ascontiguousarray(dbg["stag"])
    _, _, _, duc_r, rs_r = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0), ph, rho_sensor=True)
    ng_r, ns_r = O.smagorinsky(g, JQ, stag, 0.17, 1e9, ph, sensors, duc_r, rs_r)
    with DSEM(mesh, re=800.0, shock=True, cmu=1.0, mumax=1.0) as s:
        s.set_state(JQ)
        s.set_les(True, 0.17, 1e9)
        s.set_les_sensors(bool(sensors & O.LES_SHOCK_SENSOR), bool(sensors & O.LES_RHO_SENSOR))
        s.set_lagged_states(stag)
        s.ev_update()
        duc, rs = s.ev_sensors()
        ng, ns = s.les_update()
    assert _rel(duc, duc_r) < 1e-12 and _rel(rs, rs_r) < 1e-12
    assert _rel(ng, ng_r) < 1e-12 and _rel(ns, ns_r) < 1e-12
    base, _ = O.smagorinsky(g, JQ, stag, 0.17, 1e9, ph)
    assert _rel(ng, base) > 1e-3   # the switches act
