"""Independent Fourier pseudo-spectral evaluation of the compressible
Navier-Stokes right-hand side on the periodic cube [0, 2pi]^3.

Used as a mathematical check of the oracle's whole RHS: for a smooth periodic
state the DSEM residual must converge spectrally to this one.  Conventions
match the reference's non-dimensionalisation: T = gamma Ma^2 p / rho,
mu = 1, kappa = 1, heat flux grad T / (Pr Ma^2 (gamma - 1)), viscous terms
divided by Re.
"""
import numpy as np


def _deriv(f, k, axis):
    F = np.fft.fftn(f)
    shape = [1, 1, 1]
    shape[axis] = -1
    return np.real(np.fft.ifftn(1j * k.reshape(shape) * F))


def ns_rhs(prim, N, re, pr, mach, gamma=1.4, inviscid_scale=1.0, viscous_scale=1.0):
    """dQ/dt on an N^3 grid.  prim(x, y, z) -> (rho, u, v, w, p)."""
    x = 2 * np.pi * np.arange(N) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    rho, u, v, w, p = prim(X, Y, Z)
    k = np.fft.fftfreq(N, 1.0 / N)
    k[N // 2] = 0.0
    d = lambda f, a: _deriv(f, k, a)
    vel = [u, v, w]
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)
    Q = [rho, rho * u, rho * v, rho * w, E]
    # inviscid fluxes F[j][eq]
    F = []
    for j in range(3):
        Fj = [rho * vel[j]]
        for i in range(3):
            Fj.append(rho * vel[i] * vel[j] + (p if i == j else 0.0))
        Fj.append((E + p) * vel[j])
        F.append(Fj)
    L = [[d(vel[i], j) for j in range(3)] for i in range(3)]    # du_i/dx_j
    div = L[0][0] + L[1][1] + L[2][2]
    tau = [[(L[i][j] + L[j][i] - (2.0 / 3.0) * div * (i == j)) for j in range(3)] for i in range(3)]
    T = gamma * mach * mach * p / rho
    qh = [d(T, j) / (pr * mach * mach * (gamma - 1)) for j in range(3)]
    Fv = []
    for j in range(3):
        Fvj = [np.zeros_like(rho)] + [tau[i][j] for i in range(3)]
        Fvj.append(sum(vel[i] * tau[i][j] for i in range(3)) + qh[j])
        Fv.append(Fvj)
    out = []
    for q in range(5):
        s = np.zeros_like(rho)
        for j in range(3):
            s = s - inviscid_scale * d(F[j][q], j) + viscous_scale * d(Fv[j][q], j) / re
        out.append(s)
    return np.stack(out)


def evaluate(fields, pts):
    """Trigonometric interpolant of fields [nf, N, N, N] at points [P, 3].

    Points must lie on a tensor grid (as the Gauss points of a structured
    box do); the evaluation is separable along x, y, z.
    """
    N = fields.shape[-1]
    k = np.fft.fftfreq(N, 1.0 / N)
    k[N // 2] = 0.0
    C = np.fft.fftn(fields, axes=(1, 2, 3)) / N ** 3
    ux, ix = np.unique(np.round(pts[:, 0], 11), return_inverse=True)
    uy, iy = np.unique(np.round(pts[:, 1], 11), return_inverse=True)
    uz, iz = np.unique(np.round(pts[:, 2], 11), return_inverse=True)
    Ex = np.exp(1j * np.outer(ux, k))
    Ey = np.exp(1j * np.outer(uy, k))
    Ez = np.exp(1j * np.outer(uz, k))
    T = np.einsum("ai,fijk->fajk", Ex, C)
    T = np.einsum("bj,fajk->fabk", Ey, T)
    T = np.einsum("ck,fabk->fabc", Ez, T)
    return np.real(T[:, ix.ravel(), iy.ravel(), iz.ravel()])
