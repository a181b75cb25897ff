"""The Fortran iso_c_binding module drives the GPU library with the
reference's calling convention (state handed over in the reference's padded
d(id)%Q layout, tstep-style stepping) and gets the same state as the Python
binding and the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "tgv_drive")


def test_fortran_driver_matches_python_binding_and_oracle(tmp_path):
    print("\n[case] fortran driver n=3 p=4, 4 steps", flush=True)
    if not os.path.exists(EXE):
        pytest.fail("tests/fortran/tgv_drive is not built (python -m dsem_amd.build)")
    mesh = M.periodic_box(3, 4)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal", perturb=1e-3)
    ph = O.phys(re=400.0)
    dt = 0.5 * O.dtcalc(g, JQ, ph)
    nsteps = 4
    ncg = mesh.p + 1
    pad = np.zeros((mesh.K, 5, 10, 10, 10))
    pad[:, :, :ncg, :ncg, :ncg] = JQ.reshape(mesh.K, 5, ncg, ncg, ncg)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        np.array([mesh.K, mesh.p, mesh.nmortar, nsteps], dtype=np.int32).tofile(fh)
        for a in (mesh.mortar_elem, mesh.mortar_face, mesh.mortar_sign, mesh.mortar_nsign):
            np.ascontiguousarray(a, dtype=np.int32).tofile(fh)
        np.ascontiguousarray(mesh.corners, dtype=np.float64).tofile(fh)
        pad.tofile(fh)
        np.array([1.4, 400.0, 0.71, 1.0, dt]).tofile(fh)
    out = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    raw = np.fromfile(fout, dtype=np.float64)
    resid_f, ek_f, eps_f, jq_f = raw[0], raw[1], raw[2], raw[3:].reshape(mesh.K, 5, -1)
    with DSEM(mesh, re=400.0, metrics="per_point") as s:
        s.set_state(JQ)
        s.keep_lagged_states(True)
        for _ in range(nsteps):
            r = s.tstep(0.0, dt)
        jq_p = s.get_state()
        ek_p, eps_p = s.diagnostics()
    assert np.array_equal(jq_f, jq_p) and resid_f == r
    assert ek_f == ek_p and eps_f == eps_p
    ref = JQ.copy()
    for _ in range(nsteps):
        ref, _, _ = O.rk_step(g, ref, dt, ph)
    assert np.abs(jq_f - ref).max() / np.abs(ref).max() < 1e-13
