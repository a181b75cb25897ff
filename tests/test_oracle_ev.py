"""Mathematical pins of the oracle's entropy-viscosity restatement
(EV_routines/Main/RHS.f90, EViscosity_GL.f90, Filter/viscosity_smoothing.f90,
Prolong_gtol.f90; these files need the `mpi` module and do not compile here).

The lagged staggered states are built from exact linear velocity fields, for
which the Gauss-point gradients of the scheme are exact, so the residual
D = (Phi + Gamma)/(T Re), the Ducros sensor and mu_h have closed forms.
"""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M

GAMMA = 1.4


def _stag_points(g, mesh):
    """Physical coordinates of the three staggered grids [K,3,ns,3] (box mesh)."""
    p = mesh.p
    ncg, nc = p + 1, p + 2
    cx, cxg, _, _ = O.ops(p)
    K = mesh.K
    out = np.zeros((K, 3, nc * ncg * ncg, 3))
    for e in range(K):
        c0 = mesh.corners[e, 0]
        hvec = mesh.corners[e, 6] - mesh.corners[e, 0]
        for d in range(3):
            for b in range(ncg):
                for a in range(ncg):
                    for j in range(nc):
                        if d == 0:
                            X = (cx[j], cxg[a], cxg[b]); s = j + nc * (a + ncg * b)
                        elif d == 1:
                            X = (cxg[a], cx[j], cxg[b]); s = a + ncg * (j + nc * b)
                        else:
                            X = (cxg[a], cxg[b], cx[j]); s = a + ncg * (b + ncg * j)
                        out[e, d, s] = c0 + hvec * np.array(X)
    return out


def _field(fn, pts):
    rho, u, v, w, p = fn(pts[..., 0], pts[..., 1], pts[..., 2])
    return np.moveaxis(cases.conservative(rho, u, v, w, p), 0, -2)   # [..., 5, n]


def _setup(fn, n=2, p=4):
    mesh = M.periodic_box(n, p)
    g = O.Geometry(mesh)
    sp = _stag_points(g, mesh)
    stag = np.ascontiguousarray(np.moveaxis(cases.conservative(*fn(sp[..., 0], sp[..., 1], sp[..., 2])), 0, 2))
    Q = np.moveaxis(cases.conservative(*fn(g.xg[..., 0], g.xg[..., 1], g.xg[..., 2])), 0, 1)
    JQ = np.ascontiguousarray(Q / g.jacob[:, None, :])
    return mesh, g, JQ, stag


P0 = 71.43


def test_uniform_state_has_no_entropy_viscosity():
    fn = lambda x, y, z: (np.ones_like(x), 0.3 + 0 * x, -0.1 + 0 * x, 0.2 + 0 * x, P0 + 0 * x)
    mesh, g, JQ, stag = _setup(fn)
    mu, muh, res, duc = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 10.0, 0, 0))
    assert np.abs(res).max() < 1e-12 and np.abs(mu).max() < 1e-12 and np.abs(muh).max() < 1e-12


@pytest.mark.parametrize("alpha", [0.3, 1.7])
def test_linear_shear_closed_form(alpha):
    """u = alpha*y: Phi = alpha^2, grad T = 0, D = alpha^2/(T Re),
    mu = Cmu Re rho h^2 |D| with h = (element volume)^(1/3)/ncg."""
    fn = lambda x, y, z: (np.ones_like(x), alpha * y, 0 * x, 0 * x, P0 + 0 * x)
    mesh, g, JQ, stag = _setup(fn)
    re = 800.0
    ph = O.phys(re=re)
    mu, muh, res, duc = O.ev_viscosity(g, JQ, stag, O.OrEv(0.7, 1e6, 0, 0), ph)
    T = P0 * GAMMA / 1.0
    assert np.allclose(res, alpha ** 2 / (T * re), rtol=1e-10, atol=0)
    H = 2 * np.pi / mesh.K ** (1 / 3)
    h = H / (mesh.p + 1)
    assert np.allclose(mu, 0.7 * re * 1.0 * h * h * alpha ** 2 / (T * re), rtol=1e-10)
    # divergence-free shear: Ducros sensor = 0/(0 + |w|^2 + 1e-9) = 0
    assert np.abs(duc).max() < 1e-12
    # mu is constant, so its prolongation to the staggered grids is the same constant
    assert np.allclose(muh, mu[0, 0], rtol=1e-12)


def test_dilatation_and_rotation_ducros():
    alpha = 0.5
    dil = lambda x, y, z: (np.ones_like(x), alpha * x, alpha * y, alpha * z, P0 + 0 * x)
    _, g, JQ, stag = _setup(dil)
    _, _, res, duc = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0))
    assert np.allclose(duc, (3 * alpha) ** 2 / ((3 * alpha) ** 2 + 1e-9), rtol=1e-12)
    # Phi = 2*3a^2 - 2*(3a)^2/3 = 0 for isotropic dilatation
    assert np.abs(res).max() < 1e-12
    rot = lambda x, y, z: (np.ones_like(x), -alpha * y, alpha * x, 0 * x, P0 + 0 * x)
    _, g, JQ, stag = _setup(rot)
    _, _, res, duc = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0))
    assert np.abs(duc).max() < 1e-12 and np.abs(res).max() < 1e-12


def test_temperature_gradient_term():
    """Gamma = |grad T|^2/(T (gamma-1) Pr) for a linear pressure at rho = 1."""
    beta = 0.8
    fn = lambda x, y, z: (np.ones_like(x), 0 * x, 0 * x, 0 * x, P0 + beta * z)
    mesh, g, JQ, stag = _setup(fn)
    re, pr = 100.0, 0.71
    _, _, res, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0), O.phys(re=re, pr=pr))
    p = P0 + beta * g.xg[..., 2]
    T = p * GAMMA
    gam = (GAMMA * beta) ** 2 / (T * (GAMMA - 1) * pr)
    assert np.allclose(res, gam / T / re, rtol=1e-9)


def test_cap_and_ducros_switch():
    alpha = 2.0
    fn = lambda x, y, z: (np.ones_like(x), alpha * y, 0 * x, 0 * x, P0 + 0 * x)
    _, g, JQ, stag = _setup(fn)
    mu, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1e6, 0.25, 0, 0))
    assert np.all(mu == 0.25)
    mu, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 1, 0))
    assert np.abs(mu).max() < 1e-14   # Ducros sensor is zero for a shear


def test_smoothing_preserves_constants_and_bounds():
    rng = np.random.default_rng(3)
    # a field whose mu varies: perturbed shear
    fn = lambda x, y, z: (1.0 + 0.1 * np.sin(x), 1.3 * np.sin(y) * np.cos(z), 0.2 * np.cos(x), 0 * x,
                          P0 + np.cos(z))
    _, g, JQ, stag = _setup(fn, n=2, p=5)
    mu0, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1e9, 0, 0))
    mu1, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1e9, 0, 1))
    assert not np.allclose(mu0, mu1)
    for e in range(mu0.shape[0]):
        assert mu1[e].min() >= mu0[e].min() - 1e-15 and mu1[e].max() <= mu0[e].max() + 1e-15
    # a constant mu is kept by the smoothing: uniform shear
    fn2 = lambda x, y, z: (np.ones_like(x), 0.9 * y, 0 * x, 0 * x, P0 + 0 * x)
    _, g, JQ, stag = _setup(fn2)
    a, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1e9, 0, 0))
    b, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1e9, 0, 1))
    assert np.allclose(a, b, rtol=1e-13)


def test_cmax_cap_closed_form():
    """With the c_max cap (EViscosity_GL.f90:82, commented out in the
    reference) and a saturating Cmu, mu = c_max * max(rho) * h * max(|u| + c)
    over the element: for u = alpha*y at uniform rho and p this is
    c_max rho0 h (alpha max|y| + c0)."""
    alpha, cmax = 0.6, 0.5
    fn = lambda x, y, z: (np.ones_like(x), alpha * y, 0 * x, 0 * x, P0 + 0 * x)
    mesh, g, JQ, stag = _setup(fn)
    ev = O.OrEv(1e6, 1e-30, 0, 0)   # mumax tiny: the c_max cap must replace it
    ev.cmax_cap, ev.cmax = 1, cmax
    mu, _, _, _ = O.ev_viscosity(g, JQ, stag, ev)
    h = (2 * np.pi / mesh.K ** (1 / 3)) / (mesh.p + 1)
    c0 = np.sqrt(GAMMA * P0 / 1.0)
    want = cmax * 1.0 * h * (alpha * np.abs(g.xg[..., 1]).max(axis=1, keepdims=True) + c0)
    assert np.allclose(mu, np.broadcast_to(want, mu.shape), rtol=1e-12)
    # the cap varies from element to element with the element's own maxima
    assert np.ptp(want) > 0.01 * want.mean()
    # and without it the fixed mumax caps
    mu2, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1e6, 0.3, 0, 0))
    assert np.all(mu2 == 0.3)


def _neighbours(n, m, l, N):
    out = []
    for dn, dm, dl in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        a, b, c = n + dn, m + dm, l + dl
        if 1 <= a <= N and 1 <= b <= N and 1 <= c <= N:
            out.append((a, b, c))
    return out


def _reference_order(N):
    """viscosity_smoothing.f90:31-190: the eight corners in the listed order,
    then the interiors of the faces n=1, m=1, l=1, n=N, m=N, l=N (outer loop
    over the first free index), then the element interior (n, m, l nested)."""
    order = [(1, 1, 1), (N, 1, 1), (1, N, 1), (1, 1, N), (1, N, N), (N, 1, N), (N, N, 1), (N, N, N)]
    r = range(2, N)
    order += [(1, m, l) for m in r for l in r]
    order += [(n, 1, l) for n in r for l in r]
    order += [(n, m, 1) for n in r for m in r]
    order += [(N, m, l) for m in r for l in r]
    order += [(n, N, l) for n in r for l in r]
    order += [(n, m, N) for n in r for m in r]
    order += [(n, m, l) for n in r for m in r for l in r]
    return order


@pytest.mark.parametrize("p", [1, 2, 5])
def test_smoothing_is_in_place_in_the_reference_order(p):
    """Each smoothed point is 1/2 itself + 1/2 the mean of its in-element axis
    neighbours, taking the new value of every neighbour updated before it
    (in place, Gauss-Seidel); edge points are not smoothed.  A simultaneous
    (Jacobi) update fails this."""
    fn = lambda x, y, z: (1.0 + 0.1 * np.sin(x), 1.3 * np.sin(y) * np.cos(z), 0.2 * np.cos(x), 0 * x,
                          P0 + np.cos(z))
    _, g, JQ, stag = _setup(fn, n=2, p=p)
    mu0, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1e9, 0, 0))
    mu1, _, _, _ = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1e9, 0, 1))
    N = p + 1
    idx = lambda n, m, l: (n - 1) + N * ((m - 1) + N * (l - 1))
    order = _reference_order(N)
    rank = {pt: i for i, pt in enumerate(order)}
    assert len(rank) == len(order)
    for e in range(mu0.shape[0]):
        jac = mu0[e].copy()
        for pt in order:
            nb = _neighbours(*pt, N)
            seen = [mu1[e, idx(*q)] if rank.get(q, 1 << 30) < rank[pt] else mu0[e, idx(*q)] for q in nb]
            want = 0.5 * mu0[e, idx(*pt)] + 0.5 * sum(seen) / len(nb)
            assert abs(mu1[e, idx(*pt)] - want) <= 1e-14 * abs(mu0[e]).max(), (e, pt)
            jac[idx(*pt)] = 0.5 * mu0[e, idx(*pt)] + 0.5 * sum(mu0[e, idx(*q)] for q in nb) / len(nb)
        for n in range(1, N + 1):
            for m in range(1, N + 1):
                for l in range(1, N + 1):
                    if (n, m, l) not in rank:
                        assert mu1[e, idx(n, m, l)] == mu0[e, idx(n, m, l)]
        assert np.abs(jac - mu1[e]).max() > 1e-6 * np.abs(mu0[e]).max()   # control: not a Jacobi sweep


def test_density_sensor_closed_form():
    """rhoSensor (RHS.f90:263-275) for a linear density: the gradient (a, b, c)
    is exact, so the sensor is sqrt(sum over pairs 1/(1/(ga^2+eps) + 1/(gb^2+eps) + 1))."""
    a, b, c = 0.02, -0.05, 0.01
    fn = lambda x, y, z: (1.0 + a * x + b * y + c * z, 0.1 + 0 * x, 0 * x, 0 * x, P0 + 0 * x)
    _, g, JQ, stag = _setup(fn)
    *_, rs = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0), rho_sensor=True)
    eps = 1e-9
    t = lambda u, v: 1.0 / (1.0 / (u * u + eps) + 1.0 / (v * v + eps) + 1.0)
    want = np.sqrt(t(a, b) + t(a, c) + t(b, c))
    assert np.allclose(rs, want, rtol=1e-8)
    # uniform density: the sensor is sqrt(3 eps/(2 + eps)), about 3.9e-5
    fn0 = lambda x, y, z: (np.ones_like(x), 0.1 + 0 * x, 0 * x, 0 * x, P0 + 0 * x)
    _, g, JQ, stag = _setup(fn0)
    *_, rs0 = O.ev_viscosity(g, JQ, stag, O.OrEv(1.0, 1.0, 0, 0), rho_sensor=True)
    assert np.allclose(rs0, np.sqrt(3 * eps / (2 + eps)), rtol=1e-6)
