"""Tecplot writer (dsem_amd.tecplot) against closed forms.

The plot array of an element holds the Gauss-point values inside and values
interpolated to the faces, edges and corners.  Interpolation on the ncg
Gauss points is exact for polynomials of degree <= p, so a polynomial field
and the coordinates of an affine element must come out exact at every one of
the (ncg+2)^3 plotted points, corners included; the file holds them to the 8
digits of E20.8.
"""
import math

import numpy as np
import pytest

from dsem_amd import tecplot as TP


def _box_element(p, x0=(0.2, -0.5, 1.0), h=(0.7, 0.4, 1.3)):
    """Gauss points of one axis-aligned box element, [1, ncg^3, 3], and the
    1D Gauss points on [0, 1]."""
    ncg = p + 1
    t = np.array([0.5 * (1.0 - math.cos((2.0 * n - 1.0) * math.pi / (2.0 * ncg))) for n in range(1, ncg + 1)])
    l, m, n = np.meshgrid(t, t, t, indexing="ij")     # [l, m, n]
    xg = np.stack([x0[0] + h[0] * n, x0[1] + h[1] * m, x0[2] + h[2] * l], axis=-1).reshape(1, -1, 3)
    return xg, t


@pytest.mark.parametrize("p", [1, 3, 6, 8])
def test_end_rows_reproduce_polynomials(p):
    bl, br = TP.end_rows(p)
    _, t = _box_element(p)
    for deg in range(p + 1):
        f = t ** deg
        assert abs(bl @ f - 0.0 ** deg) < 1e-13
        assert abs(br @ f - 1.0) < 1e-12
    # mirror symmetry of the Chebyshev-Gauss points: the right row is the left one reversed
    assert np.allclose(br, bl[::-1], rtol=0, atol=1e-13)


@pytest.mark.parametrize("p", [2, 5, 8])
def test_plotted_points_are_exact_for_polynomials(p):
    ncg = p + 1
    xg, _ = _box_element(p)
    x, y, z = xg[0, :, 0], xg[0, :, 1], xg[0, :, 2]
    # density, velocity and energy as polynomials of degree <= p in the conserved variables
    rho = 1.0 + 0.1 * x + 0.05 * y * z if p >= 2 else 1.0 + 0.1 * x
    rhou = 0.3 * x ** p - 0.2 * y
    rhov = 0.1 * z ** min(p, 2)
    rhow = -0.4 + 0.1 * x * y if p >= 2 else -0.4
    rhoe = 2.5 + 0.01 * (x + y + z) ** min(p, 3)
    Q = np.stack([rho, rhou, rhov, rhow, rhoe])[None]
    jac = np.full((1, ncg ** 3), 0.5)
    F = TP.plot_fields(xg, Q / jac[:, None, :], jac, p, gamma=1.4, mach=0.1, scale=True)
    N = ncg + 2
    assert F.shape == (1, 11, N, N, N)
    # plotted coordinates: the element's Gauss points inside, its faces/edges/corners outside
    t_ext = np.concatenate([[0.0], _box_element(p)[1], [1.0]])
    L, M, Nn = np.meshgrid(t_ext, t_ext, t_ext, indexing="ij")
    X = 0.2 + 0.7 * Nn
    Y = -0.5 + 0.4 * M
    Z = 1.0 + 1.3 * L
    assert np.abs(F[0, 0] - X).max() < 1e-13
    assert np.abs(F[0, 1] - Y).max() < 1e-13
    assert np.abs(F[0, 2] - Z).max() < 1e-13
    # conserved fields exact at every plotted point, so the primitives are too
    rhoX = 1.0 + 0.1 * X + 0.05 * Y * Z if p >= 2 else 1.0 + 0.1 * X
    uX = (0.3 * X ** p - 0.2 * Y) / rhoX
    assert np.abs(F[0, 6] - rhoX).max() < 1e-12
    assert np.abs(F[0, 3] - uX).max() < 1e-11
    wX = (-0.4 + 0.1 * X * Y if p >= 2 else -0.4 + 0 * X) / rhoX
    assert np.abs(F[0, 5] - wX).max() < 1e-11


def test_scale_flag_multiplies_out_the_jacobian():
    p, ncg = 3, 4
    xg, _ = _box_element(p)
    rng = np.random.default_rng(1)
    Q = np.stack([1.0 + 0.1 * rng.random(ncg ** 3), rng.random(ncg ** 3), rng.random(ncg ** 3),
                  rng.random(ncg ** 3), 3.0 + rng.random(ncg ** 3)])[None]
    jac = np.full((1, ncg ** 3), 0.25)
    JQ = Q / jac[:, None, :]
    Fs = TP.plot_fields(xg, JQ, jac, p, 1.4, 0.5, scale=True)
    Fu = TP.plot_fields(xg, JQ, jac, p, 1.4, 0.5, scale=False)
    # unscaled: rho carries J = 1/jacob = 4; velocities and temperature do not
    assert np.allclose(Fu[0, 6], 4.0 * Fs[0, 6], rtol=1e-14, atol=0)
    for i in (3, 4, 5, 7):
        assert np.allclose(Fu[0, i], Fs[0, i], rtol=1e-12, atol=1e-14)
    # temperature = p gamma M^2 / rho at an interior point
    q = Q[0, :, 0]
    pr = 0.4 * (q[4] - 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0])
    assert abs(Fs[0, 7, 1, 1, 1] - pr * 1.4 * 0.25 / q[0]) < 1e-13


@pytest.mark.parametrize("v,s", [(1.0, "0.10000000E+01"), (-123.456789, "-0.12345679E+03"), (0.0, "0.00000000E+00"),
                                  (9.999999999e-5, "0.10000000E-03"), (1.5e-120, "0.15000000-119"),
                                  (2.0 ** -3, "0.12500000E+00")])
def test_fortran_e_format(v, s):
    out = TP.fortran_e(v)
    assert len(out) == 20 and out.strip() == s


def test_write_and_read_back(tmp_path):
    p, ncg = 2, 3
    xg1, _ = _box_element(p)
    xg = np.concatenate([xg1, xg1 + np.array([0.7, 0.0, 0.0])])
    rng = np.random.default_rng(5)
    JQ = np.concatenate([np.stack([1.0 + 0.1 * rng.random((2, ncg ** 3))[e] for _ in range(1)] +
                                  [rng.random(ncg ** 3) for _ in range(3)] + [np.full(ncg ** 3, 3.0)])[None]
                         for e in range(2)])
    jac = np.full((2, ncg ** 3), 1.0)
    ev = rng.random((2, ncg ** 3))
    path = tmp_path / "sol.plt"
    TP.write_plt(str(path), 0.125, xg, JQ, jac, p, 1.4, 0.1, ev=ev)
    text = path.read_text().splitlines()
    assert text[0].strip() == 'TITLE = "test"'
    assert text[2].strip() == 'VARIABLES = "x","y","z","u","v","w","rho","T","react","EV","NUT"'
    t, zones = TP.read_plt(str(path))
    assert t == 0.125
    assert len(zones) == 2 and all(z.shape == ((ncg + 2) ** 3, 11) for z in zones)
    F = TP.plot_fields(xg, JQ, jac, p, 1.4, 0.1, ev=ev)
    for e in range(2):
        ref = F[e].reshape(11, -1).T
        # 8 significant digits of 0.dddddddd: half a unit in the last digit is <= 5e-8 relative
        assert np.allclose(zones[e], ref, rtol=5.1e-8, atol=1e-300)
    # the EV column holds ev inside the element (point (1,1,1) is Gauss point 0)
    assert abs(zones[0][(ncg + 2) ** 2 + (ncg + 2) + 1, 9] - ev[0, 0]) <= 5.1e-8 * abs(ev[0, 0])
