"""Tecplot ASCII output of the solution, as the reference's `wrtdat`
(src/paraFMDF/Plotting/WrtPlotFiles_scalar_turb_3d.f90:13-286) writes it for
a run without scalars or FMDF.

File layout (one zone per element, in element order):

     TITLE = "test"
    TEXT X=50, Y=50, T="<time>"
     VARIABLES = "x","y","z","u","v","w","rho","T","react","EV","NUT"
     ZONE I=ncg+2 J=ncg+2 K=ncg+2 , F = POINT SOLUTIONTIME=<time>
    <(ncg+2)^3 lines, l outermost, n innermost, 11 values in Fortran E20.8>

Each zone holds the Gauss-point values in its interior and values carried
to the element's faces, then edges, then corners by 1D interpolation to the
two ends of a line (`fill_plot_arrays` :290-375, `interp_int_to_faces`,
`interp_face_to_edge`, `interp_edge_to_corner`, `interp_to_ends`
:792-1041).  The interpolation rows are those of the Gauss->Lobatto matrix
at the two Lobatto end points (`intrpmat`, src/Spectral/stac3m/matrices.f90;
points from src/Grid_Routines/stac3m_geomLib.f90:7-47, Lagrange basis from
src/Independentroutines/baselib.f90:18-80).

The conserved fields are interpolated first and the primitives formed at
every plotted point (`wrtdat` :236-243): u = Q2/Q1, rho = Q1,
p = (gamma-1)(Q5 - rho |u|^2 / 2), T = p gamma M^2 / rho.  With
`scale=False` (the end-of-run file) Q is the stored J*Q, so rho carries the
Jacobian while u, v, w and T do not; `scale=True` (the movie files,
3DDriver_parfmdf.f90:547) multiplies by 1/J first.  "react" is zero for this
physics, "EV" is the entropy viscosity and "NUT" the Smagorinsky viscosity
at the Gauss points, both extended like the state.
"""
from __future__ import annotations

import math

import numpy as np

VARIABLES = ("x", "y", "z", "u", "v", "w", "rho", "T", "react", "EV", "NUT")


def _lagrange(k: int, x: float, z: np.ndarray) -> float:
    """Lagrange basis polynomial k (0-based) of the nodes z at x, with the
    reference's factor order (baselib.f90:18-40)."""
    n = len(z)
    r = 1.0
    first = True
    for j in range(n):
        if j == k:
            continue
        f = (x - z[j]) / (z[k] - z[j])
        r = f if first else r * f
        first = False
    return r


def end_rows(p: int) -> tuple[np.ndarray, np.ndarray]:
    """Rows 1 and nc of the Gauss->Lobatto interpolation matrix of order p:
    the Lagrange basis of the ncg Chebyshev-Gauss points evaluated at the
    Lobatto end points 0 and 1."""
    ncg, nc = p + 1, p + 2
    cxg = np.array([0.5 * (1.0 - math.cos((2.0 * n - 1.0) * math.pi / (2.0 * (ncg - 1.0) + 2.0)))
                    for n in range(1, ncg + 1)])
    x0 = 0.5 * (1.0 - math.cos(0.0))
    x1 = 0.5 * (1.0 - math.cos((nc - 1) * math.pi / (nc - 1)))
    bl = np.array([_lagrange(k, x0, cxg) for k in range(ncg)])
    br = np.array([_lagrange(k, x1, cxg) for k in range(ncg)])
    return bl, br


def _ends(line: list[np.ndarray], bl: np.ndarray, br: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """interp_to_ends: left = sum_k bl[k] u_k, right = sum_k br[k] u_k, summed
    in k order; `line` holds the ncg values along the line (arrays over the
    other indices)."""
    left = np.zeros_like(line[0])
    right = np.zeros_like(line[0])
    for k in range(len(line)):
        left = left + bl[k] * line[k]
        right = right + br[k] * line[k]
    return left, right


def extend(u: np.ndarray, bl: np.ndarray, br: np.ndarray) -> np.ndarray:
    """Gauss-point values u[..., l, m, n] (ncg^3) -> plot array [..., l, m, n]
    of (ncg+2)^3 with faces, edges and corners filled in the reference's order."""
    ncg = u.shape[-1]
    N = ncg + 2
    U = np.zeros(u.shape[:-3] + (N, N, N))
    U[..., 1:-1, 1:-1, 1:-1] = u
    I = slice(1, N - 1)   # interior range
    # faces 6 and 4 (n ends), over interior (m, l)
    U[..., I, I, 0], U[..., I, I, N - 1] = _ends([U[..., I, I, 1 + k] for k in range(ncg)], bl, br)
    # faces 1 and 2 (m ends), over interior (n, l)
    U[..., I, 0, I], U[..., I, N - 1, I] = _ends([U[..., I, 1 + k, I] for k in range(ncg)], bl, br)
    # faces 3 and 5 (l ends), over interior (n, m)
    U[..., 0, I, I], U[..., N - 1, I, I] = _ends([U[..., 1 + k, I, I] for k in range(ncg)], bl, br)
    # edges: along faces 1 and 2 in n (left/right), then in l (top/bottom)
    for m in (0, N - 1):
        U[..., I, m, 0], U[..., I, m, N - 1] = _ends([U[..., I, m, 1 + k] for k in range(ncg)], bl, br)
    for m in (0, N - 1):
        U[..., 0, m, I], U[..., N - 1, m, I] = _ends([U[..., 1 + k, m, I] for k in range(ncg)], bl, br)
    # edges along faces 6 and 4 in l
    for n in (0, N - 1):
        U[..., 0, I, n], U[..., N - 1, I, n] = _ends([U[..., 1 + k, I, n] for k in range(ncg)], bl, br)
    # corners: along the four n-edges at (m, l) = (1,1), (N,1), (1,N), (N,N)
    for l, m in ((0, 0), (0, N - 1), (N - 1, 0), (N - 1, N - 1)):
        U[..., l, m, 0], U[..., l, m, N - 1] = _ends([U[..., l, m, 1 + k] for k in range(ncg)], bl, br)
    return U


def fortran_e(v: float, w: int = 20, d: int = 8) -> str:
    """Fortran Ew.d editing: 0.dddddddd E+ee, right-justified in w columns."""
    if v == 0.0 or not math.isfinite(v):
        if not math.isfinite(v):
            return f"{'NaN' if math.isnan(v) else ('-Infinity' if v < 0 else 'Infinity'):>{w}}"
        return f"{'0.' + '0' * d + 'E+00':>{w}}"
    s = f"{abs(v):.{d - 1}E}"             # d significant digits, correctly rounded
    mant, ex = s.split("E")
    digits = mant.replace(".", "")
    e = int(ex) + 1
    exp = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) < 100 else f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    return f"{('-' if v < 0 else '') + '0.' + digits + exp:>{w}}"


def plot_fields(xg: np.ndarray, JQ: np.ndarray, jacob: np.ndarray, p: int, gamma: float, mach: float,
                scale: bool = False, ev: np.ndarray | None = None, nut: np.ndarray | None = None) -> np.ndarray:
    """The 11 plotted variables at the (ncg+2)^3 points of every element:
    [K, 11, ncg+2, ncg+2, ncg+2] (l, m, n).

    xg [K, ncg^3, 3], JQ [K, 5, ncg^3], jacob [K, ncg^3] (=1/J), ev and nut
    [K, ncg^3] or None (zero)."""
    ncg = p + 1
    K = JQ.shape[0]
    shp = (K, ncg, ncg, ncg)
    bl, br = end_rows(p)
    Q = JQ * jacob[:, None, :] if scale else JQ
    Qp = extend(Q.reshape(K, 5, ncg, ncg, ncg), bl, br)
    X = extend(np.moveaxis(xg, 2, 1).reshape(K, 3, ncg, ncg, ncg), bl, br)
    zero = np.zeros(shp)
    S = extend(np.stack([zero if ev is None else ev.reshape(shp), zero if nut is None else nut.reshape(shp)], axis=1),
               bl, br)
    rho = Qp[:, 0]
    u, v, w = Qp[:, 1] / rho, Qp[:, 2] / rho, Qp[:, 3] / rho
    pr = (gamma - 1.0) * (Qp[:, 4] - 0.5 * rho * (u ** 2 + v ** 2 + w ** 2))
    T = pr * gamma * mach * mach / rho
    react = np.zeros_like(rho)
    return np.stack([X[:, 0], X[:, 1], X[:, 2], u, v, w, rho, T, react, S[:, 0], S[:, 1]], axis=1)


def write_plt(path: str, time: float, xg: np.ndarray, JQ: np.ndarray, jacob: np.ndarray, p: int,
              gamma: float = 1.4, mach: float = 1.0, scale: bool = False, ev: np.ndarray | None = None,
              nut: np.ndarray | None = None) -> None:
    """Write the `.plt` file of `wrtdat` for the state JQ [K, 5, ncg^3]."""
    F = plot_fields(xg, JQ, jacob, p, gamma, mach, scale, ev, nut)
    N = p + 3
    tstr = fortran_e(time, 25, 16)
    with open(path, "w") as fh:
        fh.write('  TITLE = "test"\n')
        fh.write(f' TEXT X=50, Y=50, T="{tstr}"\n')
        fh.write(" VARIABLES = " + ",".join(f'"{v}"' for v in VARIABLES) + "\n")
        for e in range(F.shape[0]):
            fh.write(f"  ZONE I={N:12d}  J={N:12d}  K={N:12d}  , F = POINT SOLUTIONTIME={tstr}\n")
            vals = F[e].reshape(len(VARIABLES), -1).T   # points in l, m, n order
            fh.write("".join("".join(fortran_e(x) for x in row) + "\n" for row in vals))


def read_plt(path: str) -> tuple[float, list[np.ndarray]]:
    """Parse a file written by write_plt: (time, [zone arrays [npoints, 11]])."""
    zones: list[np.ndarray] = []
    time = float("nan")
    cur: list[list[float]] = []
    npts = 0
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if s.startswith("TEXT"):
                time = float(s.split('T="')[1].rstrip('"'))
            elif s.startswith("ZONE"):
                parts = s.replace("=", " ").split()
                i, j, k = int(parts[parts.index("I") + 1]), int(parts[parts.index("J") + 1]), int(parts[parts.index("K") + 1])
                npts = i * j * k
                cur = []
                zones.append(None)   # placeholder until filled
            elif s and s[0] in "0123456789-." and zones:
                cur.append([float(x) for x in s.split()])
                if len(cur) == npts:
                    zones[-1] = np.array(cur)
    return time, zones
