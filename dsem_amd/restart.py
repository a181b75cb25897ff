"""The reference's binary restart file (write_restart / read_restart).

File layout, as write_restart writes it on one rank with the order matrices
computed (``defined_resolution`` false; src/Parabolic3D/stac3m/restart.f90:19-312):
a Fortran sequential unformatted file (every record framed by its byte
length as a 4-byte integer before and after) holding

* ``ngrid, nmort``;
* the order matrices (Order_Matrices, src/paraFMDF/initializationsc2.f90:143-170):
  ``num_orders[xyz]``, ``order_map[xyz]``, ``bx, by, bz, dmx, dmy, dmz``,
  ``num_mrtr_matrices``, ``mrtr_mat_map``, ``prolong``, ``restrict``;
* per element, 24 records of TYPE(domain) components
  (initializationsc2.f90:50-137): ``type``, ``node(26)``,
  ``orientation(6)``, ``mortar(2,6)``, ``which_surface(6)``, ``nc(3),ncg(3)``,
  ``Q, Qlgg, Qglg, Qggl, f, g, h, g_Q`` (each (nx,ny,nz,neq)),
  ``material_id``, ``cx(nmax,3),cxg(nmax,3)``, ``xg(3,nx,ny,nz)``,
  ``jacob(nx,ny,nz)``, ``gmet(3,3,nx,ny,nz)``, ``gmetg(3,3,nx,ny,nz)``,
  ``corner(3,8)``, ``ibtype(6)``, ``bcond(6)`` (CHARACTER(9)),
  ``domain_type`` (CHARACTER(8));
* per mortar, 13 records of TYPE(mortar) components
  (initializationsc2.f90:193-215): ``lenmortar(2)``, ``len(2,2)``,
  ``id(2)``, ``iface(2)``, ``orient``, ``sign``, ``nsign``,
  ``n_hat(3,nmax,nmax)``, ``norm(nmax,nmax)``, ``Q(nmax,nmax,neq,2)``,
  ``fv(nmax,nmax,neq,2)``, ``xg(3,nmax,nmax)``, ``conforming(2,2)``;
* ``num_prop`` and one record of material properties per material;
* the case record of the TGV set-up, ``time``
  (write_restart_case, caseFiles/tgv/BackstepProblem-stoch.f90:1876-1895).
Statistics and droplet records follow only when those subsystems are on;
this path writes neither and the reader stops before them.

Array sizes come from the case's size module (caseFiles/tgv/size.f90: nx =
ny = nz = nmax = 10, neq = 5, max_orders = 3, max_materials = 1), the
defaults of ``Sizes``.  Fortran arrays are stored column-major, so
``Q(nx,ny,nz,neq)`` of one element is the block [neq][nz][ny][nx] in C order
-- the state layout 1 of the C ABI.

What this path fills: the state J*Q, the lagged face-averaged staggered
states Qlgg/Qglg/Qggl when the caller has them (entropy viscosity reads
them at the next step), the geometry (corners, Gauss points, jacob, gmet,
gmetg, mortar normals), the order matrices and the topology.  The
per-stage arrays f, g, h, g_Q and the mortar traces Q, fv are written as
zeros: tstep rebuilds them before use (update_interior's first stage sets
g_Q = Qt, rkcoefs' A_1 = 0).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

from . import mesh as M


@dataclass(frozen=True)
class Sizes:
    nx: int = 10
    neq: int = 5
    max_orders: int = 3
    max_materials: int = 1

    @property
    def nmax(self) -> int:
        return self.nx

    @property
    def max_mrtr_mat(self) -> int:
        return self.max_orders * self.max_orders // 2 + 1


# ---------------------------------------------------------------- records
def write_records(path: str, records: list[bytes]) -> None:
    with open(path, "wb") as f:
        for r in records:
            n = struct.pack("<i", len(r))
            f.write(n)
            f.write(r)
            f.write(n)


def read_records(path: str) -> list[bytes]:
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        if n < 0:
            raise ValueError("records split into subrecords (> 2 GiB) are not expected in a restart file")
        rec = data[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        if n2 != n or len(rec) != n:
            raise ValueError(f"corrupt record framing at byte {pos}")
        out.append(rec)
        pos += 8 + n
    return out


def _i4(*a) -> bytes:
    """The items one after another, each in Fortran (column-major) order."""
    return b"".join(np.asarray(x, dtype="<i4").ravel(order="F").tobytes() for x in a)


def _f8(a) -> bytes:
    return np.asarray(a, dtype="<f8").ravel(order="F").tobytes()


def _chars(strings, width) -> bytes:
    return b"".join(s.encode("ascii").ljust(width)[:width] for s in strings)


# ---------------------------------------------------------------- layout helpers
def _pad_state(block: np.ndarray, p: int, sz: Sizes) -> np.ndarray:
    """Compact [5][ncg^3] (point n + ncg(m + ncg l)) -> Fortran (nx,ny,nz,neq)."""
    ncg = p + 1
    out = np.zeros((sz.nx, sz.nx, sz.nx, sz.neq), order="F")
    b = block.reshape(sz.neq, ncg, ncg, ncg)            # [v][l][m][n]
    out[:ncg, :ncg, :ncg, :] = np.transpose(b, (3, 2, 1, 0))
    return out


def _unpad_state(arr: np.ndarray, p: int) -> np.ndarray:
    ncg = p + 1
    return np.ascontiguousarray(np.transpose(arr[:ncg, :ncg, :ncg, :], (3, 2, 1, 0))).reshape(arr.shape[3], -1)


def _stag_index(d: int, p: int):
    """(i, j, k) Fortran positions (0-based) of the staggered points s of grid d."""
    ncg, nc = p + 1, p + 2
    idx = np.zeros((nc * ncg * ncg, 3), dtype=np.int64)
    for b in range(ncg):
        for a in range(ncg):
            for j in range(nc):
                if d == 0:
                    s, pos = j + nc * (a + ncg * b), (j, a, b)
                elif d == 1:
                    s, pos = a + ncg * (j + nc * b), (a, j, b)
                else:
                    s, pos = a + ncg * (b + ncg * j), (a, b, j)
                idx[s] = pos
    return idx


def _pad_stag(stag_e: np.ndarray, d: int, p: int, sz: Sizes) -> np.ndarray:
    """Grid-d staggered state [5][ns] -> Fortran (nx,ny,nz,neq) (Qlgg / Qglg / Qggl)."""
    out = np.zeros((sz.nx, sz.nx, sz.nx, sz.neq), order="F")
    ix = _stag_index(d, p)
    for v in range(sz.neq):
        out[ix[:, 0], ix[:, 1], ix[:, 2], v] = stag_e[v]
    return out


def _unpad_stag(arr: np.ndarray, d: int, p: int) -> np.ndarray:
    ix = _stag_index(d, p)
    return np.stack([arr[ix[:, 0], ix[:, 1], ix[:, 2], v] for v in range(arr.shape[3])])


# ---------------------------------------------------------------- writer
def write_restart(path: str, mesh: M.Mesh, JQ: np.ndarray, *, cx, cxg, dmx, bx, xg, jacob, gmetg, gmet, nhat,
                  norm, time: float = 0.0, stag: np.ndarray | None = None, sizes: Sizes = Sizes(),
                  num_prop: int = 0, material_property: np.ndarray | None = None) -> None:
    """Write the restart file of a state on ``mesh``.

    JQ [K][5][ncg^3] (compact, J*Q); stag [K][3][5][nc ncg^2] lagged staggered
    states or None; operators dmx [ncg][nc] (dmx(k,j)) and bx [nc][ncg]
    (bx(j,k)) and the geometry as the C ABI returns them (dsem_points,
    dsem_operators, dsem_gauss_points, dsem_get_metrics)."""
    sz = sizes
    K, p, Mn = mesh.K, mesh.p, mesh.nmortar
    ncg, nc = p + 1, p + 2
    if nc > sz.nx:
        raise ValueError(f"p = {p} needs nx >= {nc}")
    recs = [_i4(K, Mn)]
    # order matrices: one order in use, slot 1 of order_map holding nc, the
    # operators in slot 1 as set_matrices stores them (matrices.f90:20-90)
    recs.append(_i4(1, 1, 1))
    omap = np.zeros(sz.max_orders, dtype=np.int64)
    omap[0] = nc
    recs.append(_i4(omap, omap, omap))
    bxp = np.zeros((sz.nx, sz.nx, sz.max_orders), order="F")
    dmp = np.zeros((sz.nx, sz.nx, sz.max_orders), order="F")
    bxp[:nc, :ncg, 0] = np.asarray(bx).reshape(nc, ncg)        # bx(j,k): Gauss basis k at Lobatto point j
    dmp[:ncg, :nc, 0] = np.asarray(dmx).reshape(ncg, nc)       # dmx(k,j): d/dx of Lobatto basis j at Gauss k
    for a in (bxp, bxp, bxp, dmp, dmp, dmp):
        recs.append(_f8(a))
    recs.append(_i4(0))
    recs.append(_i4(np.zeros((2, sz.max_mrtr_mat), dtype=np.int64)))
    zpr = np.zeros((sz.nmax, sz.nmax, sz.max_mrtr_mat), order="F")
    recs.append(_f8(zpr))
    recs.append(_f8(zpr))
    # element records
    node_ids = _node_ids(mesh)
    e2m = {}
    for k in range(Mn):
        for s in range(2):
            e2m[(int(mesh.mortar_elem[k, s]), int(mesh.mortar_face[k, s]))] = (k, s)
    cxp = np.zeros((sz.nmax, 3), order="F")
    cgp = np.zeros((sz.nmax, 3), order="F")
    cxp[:nc, :] = np.asarray(cx)[:, None]
    cgp[:ncg, :] = np.asarray(cxg)[:, None]
    zero_state = np.zeros((sz.nx, sz.nx, sz.nx, sz.neq), order="F")
    stag_ix = [_stag_index(d, p) for d in range(3)]
    for e in range(K):
        recs.append(_i4(8))
        nodes = np.zeros(26, dtype=np.int64)
        nodes[:8] = node_ids[e] + 1
        recs.append(_i4(nodes))
        orient = np.zeros(6, dtype=np.int64)
        mort = np.zeros((2, 6), dtype=np.int64)
        for f in range(1, 7):
            k, s = e2m[(e, f)]
            mort[:, f - 1] = (k + 1, s + 1)
            if s == 1:
                orient[f - 1] = int(mesh.mortar_orient[k]) if mesh.mortar_orient is not None else 0
        recs.append(_i4(orient))
        recs.append(_i4(mort))
        recs.append(_i4(np.zeros(6, dtype=np.int64)))
        recs.append(_i4(nc, nc, nc, ncg, ncg, ncg))
        recs.append(_f8(_pad_state(JQ[e], p, sz)))
        for d in range(3):
            recs.append(_f8(_pad_stag(stag[e, d], d, p, sz) if stag is not None else zero_state))
        for _ in range(4):                                  # f, g, h, g_Q
            recs.append(_f8(zero_state))
        recs.append(_i4(1))
        recs.append(_f8(cxp) + _f8(cgp))
        xgp = np.zeros((3, sz.nx, sz.nx, sz.nx), order="F")
        jp = np.zeros((sz.nx, sz.nx, sz.nx), order="F")
        gg = np.zeros((3, 3, sz.nx, sz.nx, sz.nx), order="F")
        gm = np.zeros((3, 3, sz.nx, sz.nx, sz.nx), order="F")
        xe = xg[e].reshape(ncg, ncg, ncg, 3)                 # [l][m][n][c]
        xgp[:, :ncg, :ncg, :ncg] = np.transpose(xe, (3, 2, 1, 0))
        jp[:ncg, :ncg, :ncg] = np.transpose(jacob[e].reshape(ncg, ncg, ncg), (2, 1, 0))
        ge = gmetg[e].reshape(ncg, ncg, ncg, 3, 3)           # [l][m][n][i][j]
        gg[:, :, :ncg, :ncg, :ncg] = np.transpose(ge, (3, 4, 2, 1, 0))
        for d in range(3):
            ix = stag_ix[d]
            for c in range(3):
                gm[d, c, ix[:, 0], ix[:, 1], ix[:, 2]] = gmet[e, d, :, c]
        recs.append(_f8(xgp))
        recs.append(_f8(jp))
        recs.append(_f8(gm))
        recs.append(_f8(gg))
        recs.append(_f8(mesh.corners[e].T))                   # corner(3,8)
        recs.append(_i4(np.zeros(6, dtype=np.int64)))
        recs.append(_chars(["interface"] * 6, 9))
        recs.append(_chars(["fluid"], 8))
    # mortar records
    for k in range(Mn):
        em, es = int(mesh.mortar_elem[k, 0]), int(mesh.mortar_elem[k, 1])
        fm, fs = int(mesh.mortar_face[k, 0]), int(mesh.mortar_face[k, 1])
        recs.append(_i4(ncg, ncg))
        recs.append(_i4(np.full((2, 2), ncg)))
        recs.append(_i4(em + 1, es + 1))
        recs.append(_i4(fm, fs))
        recs.append(_i4(int(mesh.mortar_orient[k]) if mesh.mortar_orient is not None else 0))
        recs.append(_i4(int(mesh.mortar_sign[k])))
        recs.append(_i4(int(mesh.mortar_nsign[k])))
        nh = np.zeros((3, sz.nmax, sz.nmax), order="F")
        nr = np.zeros((sz.nmax, sz.nmax), order="F")
        nh[:, :ncg, :ncg] = np.transpose(nhat[k].reshape(ncg, ncg, 3), (2, 1, 0))
        nr[:ncg, :ncg] = norm[k].reshape(ncg, ncg).T
        recs.append(_f8(nh))
        recs.append(_f8(nr))
        zq = np.zeros((sz.nmax, sz.nmax, sz.neq, 2), order="F")
        recs.append(_f8(zq))
        recs.append(_f8(zq))
        mx = np.zeros((3, sz.nmax, sz.nmax), order="F")
        mx[:, :ncg, :ncg] = np.transpose(_face_points(mesh, em, fm, cxg).reshape(ncg, ncg, 3), (2, 1, 0))
        recs.append(_f8(mx))
        recs.append(_i4(np.ones((2, 2), dtype=np.int64)))   # conforming: .TRUE. (logical, 4 bytes)
    recs.append(_i4(num_prop))
    mp = np.zeros((num_prop, sz.max_materials)) if material_property is None else np.asarray(material_property)
    for mid in range(sz.max_materials):
        recs.append(_f8(mp[:num_prop, mid]))
    recs.append(_f8([time]))
    write_records(path, recs)


def _face_points(mesh: M.Mesh, e: int, f: int, cxg) -> np.ndarray:
    """Physical positions of face f's Gauss points, [j][i][3] (mortar xg)."""
    fa, fv, ai, aj = M.FACE_PARAM[f]
    n = len(cxg)
    out = np.zeros((n, n, 3))
    for j in range(n):
        for i in range(n):
            X = [0.0, 0.0, 0.0]
            X[fa], X[ai], X[aj] = fv, cxg[i], cxg[j]
            out[j, i] = M.hex8_point(mesh.corners[e], X)
    return out


def _node_ids(mesh: M.Mesh) -> np.ndarray:
    keys: dict = {}
    ids = np.zeros((mesh.K, 8), dtype=np.int64)
    for e in range(mesh.K):
        for a in range(8):
            ids[e, a] = keys.setdefault(tuple(np.round(mesh.corners[e, a], 12)), len(keys))
    return ids


# ---------------------------------------------------------------- reader
@dataclass
class Restart:
    mesh: M.Mesh
    JQ: np.ndarray            # [K][5][ncg^3]
    stag: np.ndarray          # [K][3][5][nc ncg^2]
    time: float
    corners: np.ndarray
    xg: np.ndarray            # [K][ncg^3][3]
    jacob: np.ndarray         # [K][ncg^3]
    records: int


def read_restart(path: str, sizes: Sizes = Sizes()) -> Restart:
    """Read a restart file of this layout (written by write_restart here or
    by the reference's write_restart on one rank)."""
    sz = sizes
    R = read_records(path)
    pos = 0

    def take(nbytes=None):
        nonlocal pos
        r = R[pos]
        if nbytes is not None and len(r) != nbytes:
            raise ValueError(f"record {pos}: {len(r)} bytes, expected {nbytes}")
        pos += 1
        return r

    i4 = lambda r: np.frombuffer(r, dtype="<i4")
    f8 = lambda r: np.frombuffer(r, dtype="<f8")
    K, Mn = i4(take(8))
    K, Mn = int(K), int(Mn)
    nmat = sz.nx * sz.nx * sz.max_orders * 8
    take(12)
    take(3 * sz.max_orders * 4)
    for _ in range(6):
        take(nmat)
    take(4)
    take(2 * sz.max_mrtr_mat * 4)
    take(sz.nmax * sz.nmax * sz.max_mrtr_mat * 8)
    take(sz.nmax * sz.nmax * sz.max_mrtr_mat * 8)
    st = sz.nx ** 3 * sz.neq * 8
    corners = np.zeros((K, 8, 3))
    p = None
    JQ = stag = xg = jac = None
    for e in range(K):
        take(4)
        take(26 * 4)
        take(6 * 4)
        take(12 * 4)
        take(6 * 4)
        ncs = i4(take(24))
        nc, ncg = int(ncs[0]), int(ncs[3])
        if p is None:
            p = ncg - 1
            npt, ns = ncg ** 3, nc * ncg * ncg
            JQ = np.zeros((K, sz.neq, npt))
            stag = np.zeros((K, 3, sz.neq, ns))
            xg = np.zeros((K, npt, 3))
            jac = np.zeros((K, npt))
        elif ncg - 1 != p:
            raise ValueError("elements of different orders are outside this path")
        shp = (sz.nx, sz.nx, sz.nx, sz.neq)
        JQ[e] = _unpad_state(f8(take(st)).reshape(shp, order="F"), p)
        for d in range(3):
            stag[e, d] = _unpad_stag(f8(take(st)).reshape(shp, order="F"), d, p)
        for _ in range(4):
            take(st)
        take(4)
        take(2 * sz.nmax * 3 * 8)
        xgp = f8(take(3 * sz.nx ** 3 * 8)).reshape((3, sz.nx, sz.nx, sz.nx), order="F")
        xg[e] = np.transpose(xgp[:, :ncg, :ncg, :ncg], (3, 2, 1, 0)).reshape(-1, 3)
        jp = f8(take(sz.nx ** 3 * 8)).reshape((sz.nx,) * 3, order="F")
        jac[e] = np.transpose(jp[:ncg, :ncg, :ncg], (2, 1, 0)).reshape(-1)
        take(9 * sz.nx ** 3 * 8)
        take(9 * sz.nx ** 3 * 8)
        corners[e] = f8(take(24 * 8)).reshape((3, 8), order="F").T
        take(24)
        take(54)
        take(8)
    me = np.zeros((Mn, 2), dtype=np.int32)
    mf = np.zeros((Mn, 2), dtype=np.int32)
    mo = np.zeros(Mn, dtype=np.int32)
    ms = np.zeros(Mn, dtype=np.int32)
    mn = np.zeros(Mn, dtype=np.int32)
    for k in range(Mn):
        take(8)
        take(16)
        me[k] = i4(take(8)) - 1
        mf[k] = i4(take(8))
        mo[k] = i4(take(4))[0]
        ms[k] = i4(take(4))[0]
        mn[k] = i4(take(4))[0]
        take(3 * sz.nmax ** 2 * 8)
        take(sz.nmax ** 2 * 8)
        take(sz.nmax ** 2 * sz.neq * 2 * 8)
        take(sz.nmax ** 2 * sz.neq * 2 * 8)
        take(3 * sz.nmax ** 2 * 8)
        take(16)
    nprop = int(i4(take(4))[0])
    for _ in range(sz.max_materials):
        take(nprop * 8)
    time = float(f8(take(8))[0])
    mesh = M.Mesh(p=p, corners=corners, mortar_elem=me, mortar_face=mf, mortar_sign=ms, mortar_nsign=mn,
                  mortar_orient=mo)
    mesh.check()
    return Restart(mesh=mesh, JQ=JQ, stag=stag, time=time, corners=corners, xg=xg, jacob=jac, records=pos)
