"""Conforming hex8 meshes and mortar topology for the DSEM hot path.

Host-side set-up, run once per mesh.  It produces what the C-ABI
``dsem_create`` consumes: per-element hex8 corners and the list of mortars
(element/face pairs, master side first, with the ``sign``/``nsign``
conventions of the reference).

Reference behaviour restated here:

* simple mesh format 1.1 reader -- ``src/Grid_Routines/read_mesh.f90:428-543``
  (``nodes`` / ``elements`` / ``boundary`` sections, one element header line
  ``id hex 8 material px py pz`` followed by its 8 node ids);
* face matching and master choice -- ``src/Grid_Routines/face_operations.f90:56-130``
  (elements visited in order, faces 1..6; the first element that presents a
  face becomes the master, side 1) and ``set_mortars.f90:80-220``
  (``sign_map``);
* periodic pairing of the Taylor-Green box -- ``caseFiles/tgv/BackstepProblem-stoch.f90:163-310``
  (z: face 3 master / face 5 slave; y: face 1 master / face 2 slave;
  x: face 4 of the x-max element master / face 6 slave -- the reference keeps
  the face-4 mortar there);
* ``nsign`` -- ``set_mortars.f90:227-380`` (sign of the dot product of the two
  faces' outward-coordinate normals, +1 on every aligned mesh);
* orientation of the slave face against the mortar --
  ``face_operations.f90:404-460`` (``faces_match``: the slave index and
  direction along which the master face's corner 1->2 and 1->4 edges run,
  mapped by ``orient_map``/``side_map`` of ``element_library.f90:99-145``
  to DFLT, B1F2, B1B2, F1B2, F2F1, B2F1, B2B1, F2B1 = 0..7), and the
  point map of ``rotate_to_mortar`` (``mortar_operations.f90:461-605``):
  mortar point (I, J) takes the slave's face point ``slave_point(o, I, J)``.

Face numbering (reference): 1 eta=0, 2 eta=1, 3 zeta=0, 4 xi=1, 5 zeta=1, 6 xi=0.
Faces are passed to the library 1-based, elements 0-based.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# element-local corner numbers of each face (element_library.f90: Face_Map_hex8)
FACE_NODES = {
    1: (0, 1, 5, 4),
    2: (3, 2, 6, 7),
    3: (0, 1, 2, 3),
    4: (1, 2, 6, 5),
    5: (4, 5, 6, 7),
    6: (0, 3, 7, 4),
}

# sign_map(ifl, ifr) of set_mortars.f90 (column-major RESHAPE of the table)
_SIGN_VALS = [-1, 1, -1, 1, 1, -1,
              1, -1, 1, -1, -1, 1,
              -1, 1, -1, 1, 1, -1,
              1, -1, 1, -1, -1, 1,
              1, -1, 1, -1, -1, 1,
              -1, 1, -1, 1, 1, -1]


def sign_map(ifl: int, ifr: int) -> int:
    return _SIGN_VALS[(ifr - 1) * 6 + (ifl - 1)]


# side_map(a, b) of element_library.f90 for face corners a, b = 1..4: the
# face index (1 or 2) that runs from corner a to corner b, negative when it
# runs backward; 0 where the corners are not neighbours.
_SIDE_MAP = {(1, 2): 1, (2, 1): -1, (1, 4): 2, (4, 1): -2, (2, 3): 2, (3, 2): -2, (4, 3): 1, (3, 4): -1}
# orient_map(is1, is2) of element_library.f90: DFLT, B1F2, B1B2, F1B2, F2F1,
# B2F1, B2B1, F2B1 = 0..7
_ORIENT_MAP = {(1, 2): 0, (-1, 2): 1, (-1, -2): 2, (1, -2): 3, (2, 1): 4, (-2, 1): 5, (-2, -1): 6, (2, -1): 7}


def faces_match_orientation(target_nodes, test_nodes) -> int:
    """Orientation of the test (slave) face against the target (master)
    face from their corner node ids, as faces_match computes it
    (face_operations.f90:404-460)."""
    c = [list(test_nodes).index(t) + 1 for t in target_nodes]
    is1 = _SIDE_MAP.get((c[0], c[1]), 0)
    is2 = _SIDE_MAP.get((c[0], c[3]), 0)
    o = _ORIENT_MAP.get((is1, is2))
    if o is None:
        raise ValueError("illegal orientation of element faces")
    return o


def slave_point(o: int, I, J, n: int):
    """Slave face point (i, j) that mortar point (I, J) takes under
    orientation o (rotate_to_mortar, mortar_operations.f90:461-605);
    0-based indices on an n x n face."""
    m = n - 1
    return {0: (I, J), 1: (m - I, J), 2: (m - I, m - J), 3: (I, m - J),
            4: (J, I), 5: (J, m - I), 6: (m - J, m - I), 7: (m - J, I)}[o]


# (fixed axis, fixed value, in-face axis of mortar index i, of index j)
FACE_PARAM = {
    1: (1, 0.0, 0, 2),
    2: (1, 1.0, 0, 2),
    3: (2, 0.0, 0, 1),
    5: (2, 1.0, 0, 1),
    4: (0, 1.0, 1, 2),
    6: (0, 0.0, 1, 2),
}


def hex8_point(c: np.ndarray, X) -> np.ndarray:
    """Trilinear map of reference point X in [0,1]^3 (GeomLib3D.f90:60-84)."""
    x0, x1, x2 = X
    w = np.array([(1 - x0) * (1 - x1) * (1 - x2), x0 * (1 - x1) * (1 - x2), x0 * x1 * (1 - x2),
                  (1 - x0) * x1 * (1 - x2), (1 - x0) * (1 - x1) * x2, x0 * (1 - x1) * x2,
                  x0 * x1 * x2, (1 - x0) * x1 * x2])
    return w @ c


def _face_param_corners(c: np.ndarray, face: int) -> np.ndarray:
    """Physical position of the face's 4 parameter corners (i,j) in {0,1}^2."""
    fa, fv, ai, aj = FACE_PARAM[face]
    out = []
    for j in (0.0, 1.0):
        for i in (0.0, 1.0):
            X = [0.0, 0.0, 0.0]
            X[fa] = fv
            X[ai] = i
            X[aj] = j
            out.append(hex8_point(c, X))
    return np.array(out)


@dataclass
class Mesh:
    """Conforming hex8 mesh with its mortar list (side 0 = master)."""

    p: int
    corners: np.ndarray                 # [K, 8, 3] float64
    mortar_elem: np.ndarray             # [M, 2] int32 (0-based)
    mortar_face: np.ndarray             # [M, 2] int32 (1..6)
    mortar_sign: np.ndarray             # [M] int32
    mortar_nsign: np.ndarray            # [M] int32
    mortar_orient: np.ndarray = field(default=None)  # [M] int32, 0..7 (faces_match orientation)

    @property
    def K(self) -> int:
        return int(self.corners.shape[0])

    @property
    def nmortar(self) -> int:
        return int(self.mortar_elem.shape[0])

    @property
    def ncg(self) -> int:
        return self.p + 1

    @property
    def nc(self) -> int:
        return self.p + 2

    @property
    def dof(self) -> int:
        return 5 * self.K * self.ncg ** 3

    def check(self) -> None:
        """Every element face appears in exactly one mortar (closed mesh)."""
        seen = np.zeros((self.K, 6), dtype=np.int32)
        for s in range(2):
            np.add.at(seen, (self.mortar_elem[:, s], self.mortar_face[:, s] - 1), 1)
        if not np.all(seen == 1):
            raise ValueError("mesh is not closed: every face must belong to exactly one mortar")


def _nsign(c_m: np.ndarray, f_m: int, c_s: np.ndarray, f_s: int) -> int:
    """Sign of the dot product of the two faces' grad-xi_d directions
    (set_mortars.f90:227-380)."""
    def direction(c, f):
        fa, fv, _, _ = FACE_PARAM[f]
        X = [0.5, 0.5, 0.5]
        X[fa] = fv
        h = 1e-3
        # grad of the map: columns dx/dX_k
        G = np.zeros((3, 3))
        for k in range(3):
            Xp = list(X)
            Xm = list(X)
            Xp[k] += h
            Xm[k] -= h
            G[:, k] = (hex8_point(c, Xp) - hex8_point(c, Xm)) / (2 * h)
        Gi = np.linalg.inv(G)
        v = Gi[fa, :]
        return v / np.linalg.norm(v)
    sj = float(direction(c_m, f_m) @ direction(c_s, f_s))
    return 1 if sj >= 0 else -1


def build_topology(p: int, corners: np.ndarray, node_ids: np.ndarray | None = None,
                   periodic: bool = True) -> Mesh:
    """Match faces of a conforming hex8 mesh and pair periodic boundary faces.

    ``node_ids`` [K,8] identifies shared nodes; when omitted nodes are matched
    by exact coordinates.  Boundary faces are paired as in the Taylor-Green
    case set-up (exact coordinate comparisons of the corners named there).
    """
    corners = np.ascontiguousarray(corners, dtype=np.float64)
    K = corners.shape[0]
    if node_ids is None:
        keys = {}
        node_ids = np.zeros((K, 8), dtype=np.int64)
        for e in range(K):
            for a in range(8):
                t = tuple(corners[e, a])
                node_ids[e, a] = keys.setdefault(t, len(keys))
    face_owner: dict = {}
    mortars = []  # [elem_m, face_m, elem_s, face_s, sign, orientation by node ids]
    for e in range(K):
        for f in range(1, 7):
            key = frozenset(int(node_ids[e, a]) for a in FACE_NODES[f])
            if key in face_owner:
                mi = face_owner.pop(key)
                em, fm = mortars[mi][0], mortars[mi][1]
                mortars[mi][2] = e
                mortars[mi][3] = f
                mortars[mi][4] = sign_map(fm, f)
                mortars[mi][5] = faces_match_orientation([int(node_ids[em, a]) for a in FACE_NODES[fm]],
                                                         [int(node_ids[e, a]) for a in FACE_NODES[f]])
            else:
                face_owner[key] = len(mortars)
                mortars.append([e, f, -1, 0, 1, None])
    boundary = [m for m in mortars if m[2] < 0]
    mortars = [m for m in mortars if m[2] >= 0]
    if boundary:
        if not periodic:
            raise ValueError("open boundaries are outside this path: only closed or periodic meshes")
        bset = {}
        for m in boundary:
            bset.setdefault(m[1], []).append(m[0])
        used = set()

        def pair(face_a, face_b, match, master_is_b):
            lst_b = bset.get(face_b, [])
            index = {}
            for eb in lst_b:
                index.setdefault(match(eb, True), []).append(eb)
            for ea in bset.get(face_a, []):
                cand = index.get(match(ea, False), [])
                if len(cand) != 1:
                    raise ValueError(f"periodic pairing failed for element {ea} face {face_a}")
                eb = cand[0]
                if master_is_b:
                    mortars.append([eb, face_b, ea, face_a, 1, None])
                else:
                    mortars.append([ea, face_a, eb, face_b, 1, None])
                used.add((ea, face_a))
                used.add((eb, face_b))

        C = corners
        # z: face 3 of idm (corner 1) <-> face 5 of ids (corner 5): x and y equal
        pair(3, 5, lambda e, isb: (C[e, 4, 0], C[e, 4, 1]) if isb else (C[e, 0, 0], C[e, 0, 1]), False)
        # y: face 1 (corner 1) <-> face 2 (corner 4): x and z equal
        pair(1, 2, lambda e, isb: (C[e, 3, 0], C[e, 3, 2]) if isb else (C[e, 0, 0], C[e, 0, 2]), False)
        # x: face 6 (corner 1) <-> face 4 (corner 2): z and y equal; face-4 mortar kept
        pair(6, 4, lambda e, isb: (C[e, 1, 2], C[e, 1, 1]) if isb else (C[e, 0, 2], C[e, 0, 1]), True)
        if len(used) != len(boundary):
            raise ValueError("unpaired boundary faces remain; only periodic boxes are supported")
    M = len(mortars)
    me = np.array([[m[0], m[2]] for m in mortars], dtype=np.int32).reshape(M, 2)
    mf = np.array([[m[1], m[3]] for m in mortars], dtype=np.int32).reshape(M, 2)
    ms = np.array([m[4] for m in mortars], dtype=np.int32)
    mn = np.ones(M, dtype=np.int32)
    mo = np.zeros(M, dtype=np.int32)
    # Orientation from the geometry: the o under which the slave face's
    # parameter corners land on the master's (periodic pairs after a shift
    # along one axis).  Faces matched by node ids must agree with
    # faces_match's orientation.
    for k in range(M):
        cm = _face_param_corners(corners[me[k, 0]], int(mf[k, 0])).reshape(2, 2, 3)   # [J][I]
        cs = _face_param_corners(corners[me[k, 1]], int(mf[k, 1])).reshape(2, 2, 3)
        tol = 1e-9 * (1 + np.abs(cm).max())
        found = []
        for o in range(8):
            mapped = np.array([[cs[slave_point(o, I, J, 2)[1], slave_point(o, I, J, 2)[0]] for I in (0, 1)]
                               for J in (0, 1)])
            dlt = (cm - mapped).reshape(4, 3)
            if np.allclose(dlt, dlt[0:1], rtol=0, atol=tol):
                found.append(o)
        if len(found) != 1:
            raise ValueError(f"mortar {k}: the two faces are not conforming")
        mo[k] = found[0]
        if mortars[k][5] is not None and mortars[k][5] != mo[k]:
            raise ValueError(f"mortar {k}: node-id orientation {mortars[k][5]} != geometric orientation {mo[k]}")
        mn[k] = _nsign(corners[me[k, 0]], int(mf[k, 0]), corners[me[k, 1]], int(mf[k, 1]))
    mesh = Mesh(p=p, corners=corners, mortar_elem=me, mortar_face=mf, mortar_sign=ms,
                mortar_nsign=mn, mortar_orient=mo)
    mesh.check()
    return mesh


def box_nodes(n: int, length: float = 2.0 * math.pi, rounded: bool = True):
    """Node lattice of the periodic cube as written in the reference mesh
    files: node (ix, iy, iz) has id iy + (n+1)*(ix + (n+1)*iz) (y fastest),
    coordinates i*length/n printed with 14 decimals."""
    h = length / n
    coords = np.arange(n + 1) * h
    if rounded:
        coords = np.array([float(f"{c:.14f}") for c in coords])
    return coords


def periodic_box(n: int, p: int, length: float = 2.0 * math.pi, rounded: bool = True) -> Mesh:
    """n^3 uniform hex8 elements on [0,length]^3, periodic in x, y, z.

    Element order follows the reference's tgv mesh files (y fastest, then x,
    then z), so mortar masters coincide with the reference set-up.
    """
    if n < 2:
        raise ValueError("n >= 2 elements per direction are needed for periodic pairing")
    x = box_nodes(n, length, rounded)
    N1 = n + 1

    def nid(ix, iy, iz):
        return iy + N1 * (ix + N1 * iz)

    corners = np.zeros((n ** 3, 8, 3))
    nodes = np.zeros((n ** 3, 8), dtype=np.int64)
    e = 0
    for iz in range(n):
        for ix in range(n):
            for iy in range(n):
                cn = [(ix, iy, iz), (ix + 1, iy, iz), (ix + 1, iy + 1, iz), (ix, iy + 1, iz),
                      (ix, iy, iz + 1), (ix + 1, iy, iz + 1), (ix + 1, iy + 1, iz + 1), (ix, iy + 1, iz + 1)]
                for a, (i, j, k) in enumerate(cn):
                    corners[e, a] = (x[i], x[j], x[k])
                    nodes[e, a] = nid(i, j, k)
                e += 1
    return build_topology(p, corners, nodes, periodic=True)


_REF_VERTS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])


def cube_rotations() -> list[np.ndarray]:
    """The 24 orientation-preserving symmetries of the unit cube, each as the
    corner relabelling perm with new corner a = old corner perm[a]."""
    import itertools
    out = []
    for axes in itertools.permutations(range(3)):
        for flips in itertools.product((0, 1), repeat=3):
            A = np.zeros((3, 3))
            for r, (a, f) in enumerate(zip(axes, flips)):
                A[r, a] = -1.0 if f else 1.0
            if np.linalg.det(A) < 0:
                continue
            # old reference point of new corner a: X_old = A X_new + shift
            shift = np.array([1.0 if f else 0.0 for f in flips])
            perm = []
            for v in _REF_VERTS:
                xo = A @ v + shift
                perm.append(int(np.nonzero(np.all(_REF_VERTS == xo, axis=1))[0][0]))
            out.append(np.array(perm))
    return out


def rotated_box(n: int, p: int, seed: int = 0, length: float = 2.0 * math.pi) -> Mesh:
    """Periodic n^3 box whose interior elements (those not on the periodic
    boundary, n >= 3) have their local frames turned by seeded random cube
    symmetries: the same physical mesh as periodic_box with conforming faces
    of every orientation 0..7 between neighbours."""
    if n < 3:
        raise ValueError("n >= 3: the boundary elements stay aligned for the periodic pairing")
    base = periodic_box(n, p, length)
    rots = cube_rotations()
    rng = np.random.default_rng(seed)
    x = box_nodes(n, length)
    N1 = n + 1
    corners = base.corners.copy()
    nodes = np.zeros((n ** 3, 8), dtype=np.int64)
    e = 0
    for iz in range(n):
        for ix in range(n):
            for iy in range(n):
                ids = [iy + N1 * (ix + N1 * iz), iy + N1 * (ix + 1 + N1 * iz), iy + 1 + N1 * (ix + 1 + N1 * iz),
                       iy + 1 + N1 * (ix + N1 * iz), iy + N1 * (ix + N1 * (iz + 1)), iy + N1 * (ix + 1 + N1 * (iz + 1)),
                       iy + 1 + N1 * (ix + 1 + N1 * (iz + 1)), iy + 1 + N1 * (ix + N1 * (iz + 1))]
                nodes[e] = ids
                if 0 < ix < n - 1 and 0 < iy < n - 1 and 0 < iz < n - 1:
                    perm = rots[int(rng.integers(len(rots)))]
                    corners[e] = base.corners[e][perm]
                    nodes[e] = np.array(ids)[perm]
                e += 1
    return build_topology(p, corners, nodes, periodic=True)


def read_mesh(path: str, periodic: bool = True) -> Mesh:
    """Read the reference's simple mesh format 1.1 (read_mesh.f90:428-543).

    Only hex8 elements with equal orders in the three directions are
    accepted; the staggered (non-'dg') method sets ncg = order+1, nc = ncg+1.
    """
    nodes = {}
    elems = []
    orders = set()
    with open(path) as fh:
        lines = iter(fh.read().splitlines())
        for line in lines:
            s = line.strip()
            if s == "nodes":
                for line2 in lines:
                    if line2.strip().startswith("end"):
                        break
                    t = line2.split()
                    nodes[int(t[0])] = (float(t[1]), float(t[2]), float(t[3]))
            elif s == "elements":
                for line2 in lines:
                    if line2.strip().startswith("end"):
                        break
                    t = line2.split()
                    etype = int(t[2])
                    if t[1] != "hex" or etype != 8:
                        raise ValueError("only hex8 elements are supported")
                    orders.add((int(t[4]), int(t[5]), int(t[6])))
                    ids = [int(v) for v in next(lines).split()[:8]]
                    elems.append((int(t[0]), ids))
    if len(orders) != 1:
        raise ValueError("all elements must have the same order")
    (px, py, pz), = orders
    if not (px == py == pz):
        raise ValueError("anisotropic orders are outside this path")
    elems.sort()
    K = len(elems)
    corners = np.zeros((K, 8, 3))
    node_ids = np.zeros((K, 8), dtype=np.int64)
    for e, (_, ids) in enumerate(elems):
        node_ids[e] = ids
        for a, i in enumerate(ids):
            corners[e, a] = nodes[i]
    return build_topology(px, corners, node_ids, periodic=periodic)
