"""The C-ABI library loads, exports every symbol include/dsem_hip.h declares,
and rejects invalid meshes and parameters before touching a device.
No compute calls: these run without a GPU."""
import ctypes as C

import numpy as np
import pytest

from dsem_amd import _lib as L
from dsem_amd import mesh as M


def test_library_exports_every_declared_symbol():
    lib = L.load()
    syms = L.header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_timing_table_is_static():
    lib = L.load()
    assert lib.dsem_timing_count() == 11


def test_metric_default_is_the_reference_per_point_mode():
    """A zero-initialised dsem_params selects per-point metrics (the parity
    mode); per-element metrics are opt-in."""
    import re
    text = open(L.HEADER).read()
    vals = dict(re.findall(r"#define (DSEM_METRICS_\w+)\s+(\d+)", text))
    assert vals["DSEM_METRICS_PER_POINT"] == "0"
    assert L.METRICS_PER_POINT == 0
    import inspect
    from dsem_amd.solver import DSEM
    assert inspect.signature(DSEM.__init__).parameters["metrics"].default == "per_point"


def _create(mesh, **kw):
    lib = L.load()
    me = np.ascontiguousarray(mesh.mortar_elem, dtype=np.int32)
    mf = np.ascontiguousarray(mesh.mortar_face, dtype=np.int32)
    ms = np.ascontiguousarray(mesh.mortar_sign, dtype=np.int32)
    mn = np.ascontiguousarray(mesh.mortar_nsign, dtype=np.int32)
    cr = np.ascontiguousarray(mesh.corners, dtype=np.float64)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    desc = L.MeshDesc(kw.get("K", mesh.K), kw.get("p", mesh.p), kw.get("M", mesh.nmortar), 0,
                      ip(me), ip(mf), ip(ms), ip(mn), None, cr.ctypes.data_as(C.POINTER(C.c_double)))
    prm = L.Params(1.4, kw.get("re", 1600.0), 0.71, 1.0, 0, kw.get("precision", 0), 0, 0, 1.0, 1.0, 0.5)
    h = C.c_void_p()
    rc = lib.dsem_create(C.byref(desc), C.byref(prm), 0, C.byref(h))
    msg = lib.dsem_last_error(h).decode()
    lib.dsem_destroy(h)
    return rc, msg


def test_rejects_bad_order():
    m = M.periodic_box(2, 3)
    rc, msg = _create(m, p=9)
    assert rc == L.DSEM_ERR_ARG and "p must be" in msg


def test_rejects_open_mesh():
    m = M.periodic_box(2, 3)
    rc, msg = _create(m, M=m.nmortar - 1)
    assert rc == L.DSEM_ERR_UNSUPPORTED and "open element face" in msg


def test_rejects_nonpositive_reynolds():
    m = M.periodic_box(2, 3)
    rc, _ = _create(m, re=0.0)
    assert rc == L.DSEM_ERR_ARG


def test_null_arguments():
    lib = L.load()
    assert lib.dsem_create(None, None, 0, None) == L.DSEM_ERR_ARG
    assert lib.dsem_destroy(None) == 0


def test_mesh_topology_is_closed_and_aligned():
    m = M.periodic_box(3, 4)
    assert m.nmortar == 3 * m.K
    m.check()
    assert np.all(m.mortar_sign == 1) and np.all(m.mortar_nsign == 1)


def test_nonperiodic_open_mesh_is_rejected_by_the_builder():
    m = M.periodic_box(2, 2)
    with pytest.raises(ValueError):
        M.build_topology(2, m.corners[:4], periodic=False)
