"""Shared helpers for the GPU parity tests."""
import math

import numpy as np

from dsem_amd import mesh as M


def distorted_box(n: int, p: int, amp: float = 0.08, seed: int = 11) -> M.Mesh:
    """Periodic cube whose interior lattice nodes are moved randomly (boundary
    nodes stay on the lattice so the periodic pairing still matches), giving
    non-affine trilinear elements."""
    L = 2 * math.pi
    h = L / n
    rng = np.random.default_rng(seed)
    N1 = n + 1
    X = np.zeros((N1, N1, N1, 3))
    for i in range(N1):
        for j in range(N1):
            for k in range(N1):
                X[i, j, k] = (i * h, j * h, k * h)
                if 0 < i < n and 0 < j < n and 0 < k < n:
                    X[i, j, k] += amp * h * rng.standard_normal(3)
    corners = np.zeros((n ** 3, 8, 3))
    nodes = np.zeros((n ** 3, 8), dtype=np.int64)
    e = 0
    for iz in range(n):
        for ix in range(n):
            for iy in range(n):
                cn = [(ix, iy, iz), (ix + 1, iy, iz), (ix + 1, iy + 1, iz), (ix, iy + 1, iz),
                      (ix, iy, iz + 1), (ix + 1, iy, iz + 1), (ix + 1, iy + 1, iz + 1), (ix, iy + 1, iz + 1)]
                for a, (i, j, k) in enumerate(cn):
                    corners[e, a] = X[i, j, k]
                    nodes[e, a] = j + N1 * (i + N1 * k)
                e += 1
    return M.build_topology(p, corners, nodes, periodic=True)


def graded_box(nx: int, ny: int, nz: int, p: int, lengths=(2 * math.pi, 3.0, 5.0), ratio: float = 1.6) -> M.Mesh:
    """Periodic box of nx x ny x nz elements with edge lengths ``lengths`` and
    geometrically graded spacing along each axis (largest / smallest element
    = ratio): non-cubic, anisotropic, different counts per direction.
    Element order as the reference mesh files (y fastest, then x, then z)."""
    def axis(n, L):
        w = ratio ** (np.arange(n) / max(n - 1, 1))
        return np.concatenate([[0.0], np.cumsum(w / w.sum() * L)])
    xs, ys, zs = axis(nx, lengths[0]), axis(ny, lengths[1]), axis(nz, lengths[2])
    K = nx * ny * nz
    corners = np.zeros((K, 8, 3))
    nodes = np.zeros((K, 8), dtype=np.int64)
    e = 0
    for iz in range(nz):
        for ix in range(nx):
            for iy in range(ny):
                cn = [(ix, iy, iz), (ix + 1, iy, iz), (ix + 1, iy + 1, iz), (ix, iy + 1, iz),
                      (ix, iy, iz + 1), (ix + 1, iy, iz + 1), (ix + 1, iy + 1, iz + 1), (ix, iy + 1, iz + 1)]
                for a, (i, j, k) in enumerate(cn):
                    corners[e, a] = (xs[i], ys[j], zs[k])
                    nodes[e, a] = j + (ny + 1) * (i + (nx + 1) * k)
                e += 1
    return M.build_topology(p, corners, nodes, periodic=True)


def periodic_state(xg, jacob, lengths, seed=0, amp=0.05):
    """A smooth state periodic on the box: rho, u, v, w, p as low Fourier
    modes of the box periods with seeded phases; returned as J*Q [K,5,np]
    (xg [K,np,3], jacob = 1/J [K,np])."""
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, (5, 3))
    kx, ky, kz = (2 * np.pi / L for L in lengths)
    x, y, z = xg[..., 0], xg[..., 1], xg[..., 2]
    f = [np.sin(kx * x + ph[i, 0]) * np.cos(ky * y + ph[i, 1]) + np.sin(kz * z + ph[i, 2]) for i in range(5)]
    rho = 1.0 + amp * f[0]
    u, v, w = 0.3 * f[1], 0.3 * f[2], 0.3 * f[3]
    pr = 71.4 + amp * 10 * f[4]
    E = pr / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    Q = np.stack([rho, rho * u, rho * v, rho * w, E], axis=1)
    return Q / jacob[:, None, :]


def rel_l2(a, b):
    """Per-field relative L2 error of a against b, arrays [K, 5, np]."""
    num = np.sqrt(((a - b) ** 2).sum(axis=(0, 2)))
    den = np.sqrt((b ** 2).sum(axis=(0, 2)))
    return num / np.maximum(den, 1e-300)


def per_element_metrics(g):
    """Replace the oracle geometry's per-point metrics by their value at the
    first point of each element / mortar, as the library's per-element
    (affine) metric mode stores them."""
    g.jacob[:] = g.jacob[:, :1]
    g.gmetg[:] = g.gmetg[:, :1]
    for d in range(3):
        g.gmet[d][:] = g.gmet[d][:, :1]
    g.nhat[:] = g.nhat[:, :1]
    g.norm[:] = g.norm[:, :1]
    return g
