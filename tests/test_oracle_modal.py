"""Pins of the oracle's modalfilter (energydissipation.f90:720-959).

The reference transforms with FFTW (REDFT10 forward, REDFT01 back), which is
absent here; scipy.fft's unnormalised DCT types 2 and 3 are the same
transforms (y_k = 2 sum x_j cos(pi (2j+1) k / 2N), and
y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (2k+1) / 2N)).  Pins:

* against a scipy DCT-II / DCT-III pair with the reference's rescaling, with
  every mode kept (the reference as shipped) and with modes cut;
* every mode kept: the identity, with the round-off drift bounded (the bound
  stated for the reference's every-50-steps application);
* modes cut: a projection (idempotent) that keeps tensor polynomials of
  degree < cutoff in the Chebyshev-Gauss reference coordinates exactly -- the
  DCT-II basis at those points is the Chebyshev polynomials T_k.
"""
import numpy as np
import pytest
import scipy.fft as F

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M


def scipy_filter(geom, JQ, cutoff):
    N = geom.p + 1
    out = JQ.copy()
    for e in range(geom.K):
        j0 = geom.jacob[e, 0]
        for v in (1, 2, 3):
            a = (JQ[e, v] * j0).reshape(N, N, N)          # [l][m][n]
            mid = F.dctn(a, type=2)
            idx = np.arange(N)
            sc = np.where(idx == 0, 2.0 * N, float(N))
            mid = mid / (sc[:, None, None] * sc[None, :, None] * sc[None, None, :])
            keep = (idx < cutoff)
            mid = mid * (keep[:, None, None] & keep[None, :, None] & keep[None, None, :])
            h = np.where(idx == 0, 1.0, 2.0)
            mid = mid / (h[:, None, None] * h[None, :, None] * h[None, None, :])
            out[e, v] = (F.dctn(mid, type=3) / j0).reshape(-1)
    return out


def _tgv(n, p, seed=3):
    mesh = M.periodic_box(n, p)
    g = O.Geometry(mesh)
    return mesh, g, cases.tgv_state(g.xg, g.jacob, "reference", perturb=1e-3, seed=seed)


@pytest.mark.parametrize("p", [1, 4, 8])
def test_filter_matches_scipy_dct_pair(p):
    _, g, JQ = _tgv(2, p)
    for cutoff in sorted({p + 1, max(p - 1, 1)}):
        got = O.modal_filter(g, JQ, cutoff)
        want = scipy_filter(g, JQ, cutoff)
        scale = np.abs(JQ[:, 1:4]).max()
        assert np.abs(got - want).max() <= 1e-14 * scale, cutoff
        assert np.array_equal(got[:, [0, 4]], JQ[:, [0, 4]])   # density and energy untouched


def test_all_modes_kept_is_identity_within_stated_drift():
    _, g, JQ = _tgv(4, 8)
    got = O.modal_filter(g, JQ)
    drift = np.abs(got - JQ).max() / np.abs(JQ[:, 1:4]).max()
    print(f"modalfilter drift (max abs / max |rho u|) at p=8: {drift:.2e}")
    assert drift <= 1e-14   # measured 3.7e-15 (direct sums; about 16 ulps of |rho u|)


def test_cut_filter_is_a_projection_keeping_low_polynomials():
    p, cutoff = 6, 4
    mesh, g, JQ = _tgv(2, p)
    once = O.modal_filter(g, JQ, cutoff)
    twice = O.modal_filter(g, once, cutoff)
    assert np.abs(twice - once).max() <= 1e-13 * np.abs(once).max()
    assert np.abs(once - JQ).max() > 1e-6 * np.abs(JQ).max()     # it does remove something
    # tensor polynomial of degree < cutoff in the element's reference coordinates
    N = p + 1
    xi = (g.xg - g.xg.min(axis=1, keepdims=True)) / (2 * np.pi / 2)   # element-local [0,1] coordinates
    poly = (1 + xi[..., 0] - 2 * xi[..., 0] ** 3) * (2 - xi[..., 1] ** 2) * (0.5 + xi[..., 2] ** 3)
    JQp = JQ.copy()
    for v in (1, 2, 3):
        JQp[:, v] = v * poly / g.jacob[:, :1]
    kept = O.modal_filter(g, JQp, cutoff)
    assert np.abs(kept - JQp).max() <= 1e-13 * np.abs(JQp).max()
    # one degree more is not kept
    JQq = JQp.copy()
    JQq[:, 1] = xi[..., 0] ** cutoff / g.jacob[:, :1]
    assert np.abs(O.modal_filter(g, JQq, cutoff) - JQq).max() > 1e-6
    assert N > cutoff
