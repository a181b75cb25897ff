"""The reference restart file (write_restart, src/Parabolic3D/stac3m/restart.f90:19-312)
written and read by dsem_amd.restart.

* Every record's byte length equals the size of the component it holds, as
  the reference's compiled declarations give it (TYPE(domain),
  TYPE(mortar), Order_Matrices, Material_Properties; ref_restart_record_sizes
  in oracle/_ref), over the header, every element and every mortar record.
* The Q record of an element, read as the C-order block [neq][nz][ny][nx],
  is the state's layout 1 (d(id)%Q(nx,ny,nz,neq)): checked entry by entry
  against the compact state.
* Write -> read returns the mesh (corners, mortar topology, orientations,
  signs), the state, the lagged staggered states, Gauss points, jacob and
  time exactly, on an aligned and on a rotated-face mesh.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd import restart as R


def _write(tmp_path, mesh, seed=0, num_prop=0):
    g = O.Geometry(mesh)
    cx, cxg, dmx, bx = O.ops(mesh.p)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-2, seed=seed)
    rng = np.random.default_rng(seed)
    stag = rng.standard_normal((mesh.K, 3, 5, (mesh.p + 2) * (mesh.p + 1) ** 2))
    path = str(tmp_path / "restart.rst")
    R.write_restart(path, mesh, JQ, cx=cx, cxg=cxg, dmx=dmx, bx=bx, xg=g.xg, jacob=g.jacob,
                    gmetg=g.gmetg.reshape(mesh.K, -1, 9), gmet=np.stack(g.gmet, axis=1), nhat=g.nhat,
                    norm=g.norm, time=1.25, stag=stag, num_prop=num_prop,
                    material_property=np.zeros((num_prop, 1)))
    return path, g, JQ, stag


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (needs /root/reference)")
def test_record_sizes_match_the_reference_declarations(tmp_path):
    mesh = M.periodic_box(2, 3)
    path, *_ = _write(tmp_path, mesh, num_prop=1)
    rec = [len(r) for r in R.read_records(path)]
    lib = O.ref()
    lib.ref_restart_record_sizes.argtypes = [C.POINTER(C.c_int), C.c_int]
    lib.ref_restart_record_sizes.restype = C.c_int
    buf = (C.c_int * 64)()
    n = lib.ref_restart_record_sizes(buf, 64)
    ref = list(buf[:n])
    head, dom, mor, tail = ref[:13], ref[13:37], ref[37:50], ref[50:]
    assert len(dom) == 24 and len(mor) == 13 and len(tail) == 3
    K, Mn = mesh.K, mesh.nmortar
    assert len(rec) == 13 + 24 * K + 13 * Mn + 1 + 1 + 1
    assert rec[:13] == head
    for e in range(K):
        assert rec[13 + 24 * e:13 + 24 * (e + 1)] == dom, e
    o = 13 + 24 * K
    for k in range(Mn):
        assert rec[o + 13 * k:o + 13 * (k + 1)] == mor, k
    assert rec[-3:] == tail


def test_q_record_is_the_layout_1_block(tmp_path):
    mesh = M.periodic_box(2, 4)
    path, g, JQ, _ = _write(tmp_path, mesh, seed=3)
    rec = R.read_records(path)
    ncg = mesh.p + 1
    for e in (0, 5):
        q = np.frombuffer(rec[13 + 24 * e + 6], dtype="<f8").reshape(5, 10, 10, 10)   # [v][l][m][n]
        for v in range(5):
            for l in range(ncg):
                for m in range(ncg):
                    for n in range(ncg):
                        assert q[v, l, m, n] == JQ[e, v, n + ncg * (m + ncg * l)]
        assert np.all(q[:, ncg:, :, :] == 0) and np.all(q[:, :, ncg:, :] == 0) and np.all(q[:, :, :, ncg:] == 0)


@pytest.mark.parametrize("rotated", [False, True])
def test_write_read_round_trip(tmp_path, rotated):
    mesh = M.rotated_box(3, 4, seed=2) if rotated else M.periodic_box(3, 5)
    path, g, JQ, stag = _write(tmp_path, mesh, seed=7)
    r = R.read_restart(path)
    assert r.time == 1.25
    assert np.array_equal(r.mesh.corners, mesh.corners)
    for a in ("mortar_elem", "mortar_face", "mortar_sign", "mortar_nsign", "mortar_orient"):
        assert np.array_equal(getattr(r.mesh, a), getattr(mesh, a)), a
    assert np.array_equal(r.JQ, JQ)
    assert np.array_equal(r.stag, stag)
    assert np.array_equal(r.xg, g.xg)
    assert np.array_equal(r.jacob, g.jacob)


def test_framing_errors_are_reported(tmp_path):
    path = str(tmp_path / "bad.rst")
    with open(path, "wb") as f:
        f.write(np.array([8], dtype="<i4").tobytes() + b"\0" * 8 + np.array([4], dtype="<i4").tobytes())
    with pytest.raises(ValueError):
        R.read_records(path)
