"""GPU flow diagnostics against the oracle: kinetic energy (turbk), the
enstrophy dissipation, and the Taylor-Green dissipation-rate history
(-dEk/dt per step) on the 4^3 p=4 case and, through committed oracle
fixtures, on the 22^3 p=8 production case (200 steps) and on 8^3 p=4 to
t = 10, through the rise, peak and decay of the dissipation rate."""
import os

import numpy as np
import pytest

import oracle as O
from dsem_amd import _lib as L
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd import tgv
from dsem_amd.solver import DSEM
from helpers import distorted_box

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "tgv22p8_history.txt")
GOLDEN8 = os.path.join(os.path.dirname(__file__), "golden", "tgv8p4_history.txt")


def test_kinetic_energy_matches_oracle_on_distorted_mesh():
    print("\n[case] Ek distorted 3^3 p=5", flush=True)
    mesh = distorted_box(3, 5, amp=0.08, seed=4)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=3e-2, seed=8)
    ref = O.kinetic_energy(g, JQ)
    with DSEM(mesh, re=800.0, metrics="per_point") as s:
        s.set_state(JQ)
        got = s.kinetic_energy()
    assert abs(got - ref) <= 1e-13 * abs(ref)


def test_enstrophy_needs_lagged_states():
    mesh = M.periodic_box(2, 3)
    with DSEM(mesh) as s:
        xg, jac = s.gauss_points()
        s.set_state(cases.tgv_state(xg, jac))
        with pytest.raises(L.DsemError):
            s.diagnostics()
        s.keep_lagged_states(True)
        with pytest.raises(L.DsemError):   # no step taken yet
            s.diagnostics()


def test_diagnostics_after_steps_match_oracle():
    print("\n[case] Ek/eps after 2 steps, distorted 3^3 p=4", flush=True)
    mesh = distorted_box(3, 4, amp=0.06, seed=2)
    g = O.Geometry(mesh)
    ph = O.phys(re=700.0)
    JQ0 = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=2e-2, seed=3)
    dt = 0.4 * O.dtcalc(g, JQ0, ph)
    ref = JQ0.copy()
    stag = np.zeros((mesh.K, 3, 5, g.ns))
    for _ in range(2):
        O.rk_step(g, ref, dt, ph, stag_out=stag)
    ek_r, eps_r = O.kinetic_energy(g, ref), O.enstrophy(g, ref, stag, ph)
    with DSEM(mesh, re=700.0, metrics="per_point") as s:
        s.set_state(JQ0)
        s.keep_lagged_states(True)
        s.tstep(0.0, dt)
        s.tstep(dt, dt)
        ek, eps = s.diagnostics()
        lag = s.get_lagged_states()
    assert np.abs(lag - stag).max() <= 1e-13 * np.abs(stag).max()
    assert abs(ek - ek_r) <= 1e-13 * ek_r
    assert abs(eps - eps_r) <= 1e-12 * eps_r


def _oracle_history(n, p, steps, re=1600.0, pr=0.71, cfl=0.5):
    ph = O.phys(re=re, pr=pr)
    mesh = M.periodic_box(n, p)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal")
    dt = cfl * O.dtcalc(g, JQ, ph)
    stag = np.zeros((mesh.K, 3, 5, g.ns))
    prev = O.kinetic_energy(g, JQ)
    rows = [(0, 0.0, prev, np.nan, np.nan)]
    t = 0.0
    for k in range(1, steps + 1):
        O.rk_step(g, JQ, dt, ph, stag_out=stag)
        ek = O.kinetic_energy(g, JQ)
        t += dt
        rows.append((k, t, ek, (prev - ek) / dt, O.enstrophy(g, JQ, stag, ph)))
        prev = ek
    return np.array(rows), JQ


def _compare_history(h, r):
    """Rounding-level agreement: the clean low-Mach TGV RHS carries ~1e-11
    relative FP64 rounding noise, which the two implementations accumulate
    differently over the run, and Ek sums ~8e6 points in a different order
    (tree vs sequential, ~sqrt(N) eps = 3e-13 relative); the first difference
    divided by dt turns that into ~1e-6 of -dEk/dt at 22^3."""
    assert h.shape == r.shape
    ek = np.abs(h[:, 2] - r[:, 2]).max() / np.abs(r[:, 2]).max()
    d, dr = h[1:, 3], r[1:, 3]
    de = np.abs(d - dr).max() / np.abs(dr).max()
    ee = np.abs(h[1:, 4] - r[1:, 4]).max() / np.abs(r[1:, 4]).max()
    print(f"history: Ek {ek:.2e}  -dEk/dt {de:.2e}  eps {ee:.2e}", flush=True)
    assert np.allclose(h[:, 1], r[:, 1], rtol=1e-14, atol=0)
    assert ek <= 1e-11
    assert de <= 1e-5
    assert ee <= 1e-10


def test_tgv_history_4x4x4_p4():
    print("\n[case] TGV 4^3 p=4 100 steps history", flush=True)
    r, JQ_r = _oracle_history(4, 4, 100)
    out = tgv.run(n=4, p=4, steps=100)
    _compare_history(out["history"], r)
    assert np.abs(out["JQ"] - JQ_r).max() <= 1e-12 * np.abs(JQ_r).max()
    # initial dissipation rate: nu <|omega|^2> = 0.75/Re for the TGV field
    assert abs(out["history"][1, 3] - 0.75 / 1600.0) < 0.03 * 0.75 / 1600.0


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="oracle history fixture not generated")
def test_tgv_history_22x22x22_p8():
    print("\n[case] TGV 22^3 p=8 200 steps history vs oracle fixture", flush=True)
    r = np.loadtxt(GOLDEN)
    steps = int(r[-1, 0])
    out = tgv.run(n=22, p=8, steps=steps)
    h = out["history"]
    _compare_history(h, r)
    # -dEk/dt against the enstrophy dissipation nu <|omega|^2> (eps carries
    # the reference's factor 100 and is an unweighted point mean): they agree
    # to within half a percent over the run
    assert np.abs(h[1:, 3] - h[1:, 4] / 100.0).max() < 1e-2 * np.abs(h[1:, 3]).max()


@pytest.mark.skipif(not os.path.exists(GOLDEN8), reason="oracle history fixture not generated")
def test_tgv_history_8x8x8_p4_through_the_dissipation_peak():
    """8^3 hex8, p = 4, Re = 1600, 8266 steps to t = 10: the GPU history
    against the oracle's (every 20th step), over the window where a wrong
    viscous term would show -- the dissipation rate grows 30-fold from
    0.75/Re, peaks and decays."""
    print("\n[case] TGV 8^3 p=4 to t=10 history vs oracle fixture", flush=True)
    r = np.loadtxt(GOLDEN8)
    steps = int(r[-1, 0])
    out = tgv.run(n=8, p=4, steps=steps)
    h = out["history"][r[:, 0].astype(int)]
    assert np.allclose(h[:, 1], r[:, 1], rtol=1e-13, atol=0)
    ek = np.abs(h[:, 2] - r[:, 2]).max() / np.abs(r[:, 2]).max()
    de = np.abs(h[1:, 3] - r[1:, 3]).max() / np.abs(r[1:, 3]).max()
    ee = np.abs(h[1:, 4] - r[1:, 4]).max() / np.abs(r[1:, 4]).max()
    ipk = int(np.argmax(r[1:, 3]))
    print(f"history to t={r[-1, 1]:.3f}: Ek {ek:.2e}  -dEk/dt {de:.2e}  eps {ee:.2e}; "
          f"peak -dEk/dt {h[1 + ipk, 3]:.6e} (oracle {r[1 + ipk, 3]:.6e}) at t={r[1 + ipk, 1]:.3f}", flush=True)
    # Measured: Ek 1.9e-9, -dEk/dt 3.5e-8, eps 3.2e-8 of their maxima at t = 10.
    # The two implementations round differently (~1e-13 per step); the
    # transitional flow amplifies that over the 8266 steps.  A 1 % change of
    # nu moves Ek by up to 9.6e-4 and -dEk/dt by up to 7.5e-3 of their maxima
    # over this window.
    assert ek <= 1e-8
    assert de <= 1e-6
    assert ee <= 1e-6
    early = r[:, 1] <= 1.0              # before the transition: rounding level
    assert np.abs(h[early, 2] - r[early, 2]).max() <= 1e-12 * r[0, 2]
    d = r[1:, 3]
    assert d[ipk] > 20 * d[0] and 0 < ipk < len(d) - 1 and d[-1] < 0.7 * d[ipk]
