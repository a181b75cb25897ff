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
