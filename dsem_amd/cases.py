"""Initial states for the Taylor-Green vortex and synthetic throughput inputs.

Case set-up (host side, once).  The TGV field restates ``exactsol``
(``caseFiles/tgv/BackstepProblem-stoch.f90:1533-1570``) with the reference's
non-dimensionalisation: ``mach`` = 1 in the equations and wall temperature
T0 = 100, so that p0 = T0/gamma = 71.43, c0 = 10 and the effective Mach number
is U0/c0 = 0.1 (``p8-22x22x22/bump.in:53-56``).

Two density choices:
  * ``"reference"`` -- rho = 1 everywhere, as ``exactsol`` sets it;
  * ``"isothermal"`` -- rho = rho0 * p / p0 ("rho set from p at constant
    temperature"), used where the density equation must not vanish.
"""
from __future__ import annotations

import numpy as np

GAMMA = 1.4


def tgv_primitive(x, y, z, variant: str = "reference", T0: float = 100.0,
                  gamma: float = GAMMA, U0: float = 1.0):
    u = U0 * np.sin(x) * np.cos(y) * np.cos(z)
    v = -U0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    p0 = T0 * 1.0 / gamma
    p = p0 + (np.cos(2.0 * x) + np.cos(2.0 * y)) * (np.cos(2.0 * z) + 2.0) / 16.0
    if variant == "reference":
        rho = np.ones_like(u)
    elif variant == "isothermal":
        rho = p / p0
    else:
        raise ValueError(f"unknown TGV variant {variant!r}")
    return rho, u, v, w, p


def conservative(rho, u, v, w, p, gamma: float = GAMMA):
    """[5, ...] conserved variables (rho, rho u, rho v, rho w, rho e)."""
    rhoe = p / (gamma - 1.0) + 0.5 * rho * (u ** 2 + v ** 2 + w ** 2)
    return np.stack([rho, rho * u, rho * v, rho * w, rhoe])


def tgv_state(xg: np.ndarray, jacob: np.ndarray, variant: str = "reference",
              perturb: float = 0.0, seed: int = 1234, T0: float = 100.0) -> np.ndarray:
    """J-weighted conserved state JQ [K, 5, np] at the Gauss points.

    xg [K, np, 3] physical Gauss points, jacob [K, np] (=1/J).  The driver
    converts Q to J*Q by dividing by jacob (3DDriver_parfmdf.f90:286-300).
    ``perturb`` > 0 multiplies every field by (1 + N(0, perturb)) with a
    seeded generator, so that no RHS component vanishes identically.
    """
    rho, u, v, w, p = tgv_primitive(xg[..., 0], xg[..., 1], xg[..., 2], variant, T0)
    Q = conservative(rho, u, v, w, p)                    # [5, K, np]
    Q = np.ascontiguousarray(np.transpose(Q, (1, 0, 2)))  # [K, 5, np]
    if perturb > 0:
        rng = np.random.default_rng(seed)
        Q = Q * (1.0 + perturb * rng.standard_normal(Q.shape))
    return Q / jacob[:, None, :]


def isotropic_turbulence(xg: np.ndarray, jacob: np.ndarray, seed: int = 1234, k0: float = 4.0,
                         mach_t: float = 0.3, kmax: int = 12, T0: float = 100.0, gamma: float = GAMMA):
    """Decaying compressible isotropic turbulence on the periodic [0, 2 pi]^3 box.

    Velocity: a random solenoidal Fourier field, u(x) = sum_k uh(k) exp(i k.x)
    over integer wavevectors |k| <= kmax, with uh(k) = P(k) xi(k), xi complex
    Gaussian (numpy default_rng(seed)), P the projector normal to k, Hermitian
    symmetry uh(-k) = conj(uh(k)), and every shell |k| in [n - 1/2, n + 1/2)
    rescaled to the energy E(n) proportional to n^4 exp(-2 (n/k0)^2).  The
    amplitude gives the turbulent Mach number Ma_t = <|u|^2>^(1/2) / c with
    c = sqrt(gamma p / rho) = sqrt(T0) in the reference's scaling (mach = 1).
    Density and temperature are uniform: rho = 1, p = T0 / gamma.

    The Gauss points of a box mesh lie on a tensor grid, so the synthesis is
    three 1-D contractions over the distinct coordinates.

    Returns JQ [K, 5, np] and a dict with u_rms (per component), the Taylor
    microscale lambda = (5 int E / int k^2 E)^(1/2) of the discrete spectrum,
    and re_for_re_lambda_100 = 100 / (u_rms lambda), the Reynolds number of
    the non-dimensional equations that gives Re_lambda = 100.
    """
    rng = np.random.default_rng(seed)
    ks = np.arange(-kmax, kmax + 1)
    KX, KY, KZ = np.meshgrid(ks, ks, ks, indexing="ij")
    kk = np.stack([KX, KY, KZ])                              # [3, n, n, n]
    k2 = (kk ** 2).sum(0).astype(float)
    xi = rng.standard_normal((3,) + k2.shape) + 1j * rng.standard_normal((3,) + k2.shape)
    xi = 0.5 * (xi + np.conj(xi[:, ::-1, ::-1, ::-1]))       # Hermitian: xi(-k) = conj(xi(k))
    with np.errstate(invalid="ignore", divide="ignore"):
        kd = (kk * xi).sum(0) / np.where(k2 > 0, k2, 1.0)
    uh = xi - kk * kd                                        # solenoidal projection
    kn = np.sqrt(k2)
    shell = np.rint(kn).astype(int)
    uh[:, (k2 == 0) | (kn > kmax)] = 0.0
    spec = np.zeros(kmax + 1)
    for n in range(1, kmax + 1):
        sel = shell == n
        e_now = 0.5 * np.sum(np.abs(uh[:, sel]) ** 2)
        target = n ** 4 * np.exp(-2.0 * (n / k0) ** 2)
        if e_now > 0:
            uh[:, sel] *= np.sqrt(target / e_now)
            spec[n] = target
    # physical-space values on the tensor grid of Gauss-point coordinates
    coords = [np.unique(np.round(xg[..., d], 12)) for d in range(3)]
    E = [np.exp(1j * np.outer(c, ks)) for c in coords]       # [nx, 2kmax+1]
    vel = np.einsum("ia,jb,lc,vabc->vijl", E[0], E[1], E[2], uh, optimize=True).real
    c0 = np.sqrt(T0)
    scale = mach_t * c0 / np.sqrt(np.mean((vel ** 2).sum(0)))
    vel *= scale
    idx = [np.searchsorted(coords[d], np.round(xg[..., d], 12)) for d in range(3)]
    u, v, w = (vel[c][idx[0], idx[1], idx[2]] for c in range(3))
    rho = np.ones_like(u)
    p = T0 / gamma * np.ones_like(u)
    Q = np.ascontiguousarray(np.transpose(conservative(rho, u, v, w, p, gamma), (1, 0, 2)))
    n = np.arange(kmax + 1)
    u_rms = mach_t * c0 / np.sqrt(3.0)
    lam = np.sqrt(5.0 * spec.sum() / (n ** 2 * spec).sum())
    info = {"u_rms": u_rms, "taylor_microscale": lam, "re_for_re_lambda_100": 100.0 / (u_rms * lam),
            "spectrum": spec * scale ** 2}
    return Q / jacob[:, None, :], info
