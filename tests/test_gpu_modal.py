"""GPU modalfilter (energydissipation.f90:720-959) against the oracle, and its
round-off drift at the production size (the reference applies it every 50
steps with every mode kept, where it is the identity up to round-off)."""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM
from helpers import distorted_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,p", [(3, 1), (2, 5), (2, 8)])
def test_modal_filter_matches_oracle(n, p):
    print(f"\n[case] modal filter n={n} p={p}", flush=True)
    mesh = distorted_box(n, p, amp=0.05, seed=3) if p < 8 else M.periodic_box(n, p)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-2, seed=4)
    with DSEM(mesh) as s:
        for cutoff in sorted({p + 1, max(p - 1, 1)}):
            s.set_state(JQ)
            s.modal_filter(cutoff)
            got = s.get_state()
            want = O.modal_filter(g, JQ, cutoff)
            err = np.abs(got - want).max() / np.abs(JQ[:, 1:4]).max()
            print(f"cutoff {cutoff}: max abs diff / max |rho u| = {err:.2e}", flush=True)
            assert err <= 1e-14
            assert np.array_equal(got[:, [0, 4]], JQ[:, [0, 4]])


def test_modal_filter_drift_at_22cubed_p8():
    print("\n[case] modal filter drift 22^3 p=8", flush=True)
    mesh = M.periodic_box(22, 8)
    with DSEM(mesh) as s:
        xg, jac = s.gauss_points()
        JQ = cases.tgv_state(xg, jac, "reference", perturb=1e-3, seed=1234)
        s.set_state(JQ)
        s.modal_filter()
        got = s.get_state()
    scale = np.abs(JQ[:, 1:4]).max()
    drift = np.abs(got - JQ).max() / scale
    l2 = np.linalg.norm(got[:, 1:4] - JQ[:, 1:4]) / np.linalg.norm(JQ[:, 1:4])
    # Stated bound: the round trip is six passes of length-N direct sums
    # (N = p+1 = 9) between scalings.  With independent roundings each pass
    # adds about sqrt(N) u relative (u = 2^-53) in the L2 norm; three standard
    # deviations over the six passes: 3 sqrt(6 N) u = 2.4e-15.  The worst-case
    # bound for single entries, 6 N u = 6e-15 of max|rho u|, is held at 1e-14.
    u = 2.0 ** -53
    N = 9
    print(f"drift max {drift:.2e}, relative L2 {l2:.2e} (bound {3 * np.sqrt(6 * N) * u:.2e})", flush=True)
    assert drift <= 1e-14
    assert l2 <= 3 * np.sqrt(6 * N) * u
