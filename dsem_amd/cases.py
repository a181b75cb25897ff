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
