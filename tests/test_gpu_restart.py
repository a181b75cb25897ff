"""Restart files from the GPU path (write_restart / read_restart,
src/Parabolic3D/stac3m/restart.f90:19-312): a run stopped after n steps,
written, read into a new context and continued equals the uninterrupted run
bit for bit -- with entropy viscosity on, whose first step after the restart
reads the lagged staggered states from the file -- and the geometry the file
carries (dsem_get_metrics, dsem_points, dsem_operators) equals the oracle's."""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd import restart as R
from dsem_amd.solver import DSEM

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shock", [False, True])
def test_restarted_run_equals_uninterrupted_run(tmp_path, shock):
    print(f"\n[case] restart continuation 3^3 p=5 shock={shock}", flush=True)
    mesh = M.periodic_box(3, 5)
    kw = dict(re=400.0, shock=shock, cmu=1.0, mumax=0.5)
    with DSEM(mesh, **kw) as s:
        xg, jac = s.gauss_points()
        JQ = cases.tgv_state(xg, jac, "isothermal", perturb=1e-2, seed=11)
        s.set_state(JQ)
        dt = 0.5 * s.dtcalc()
        s.steps(3, dt)
        path = str(tmp_path / "run.rst")
        s.write_restart(path, time=3 * dt)
        s.steps(2, dt)
        want = s.get_state()
    s2, t = DSEM.from_restart(path, **kw)
    with s2:
        assert t == 3 * dt
        s2.steps(2, dt)
        got = s2.get_state()
    assert np.array_equal(got, want)


def test_restart_geometry_equals_oracle(tmp_path):
    print("\n[case] restart geometry 3^3 p=4 rotated", flush=True)
    mesh = M.rotated_box(3, 4, seed=1)
    g = O.Geometry(mesh)
    cx, cxg, dmx, bx = O.ops(mesh.p)
    with DSEM(mesh) as s:
        path = str(tmp_path / "g.rst")
        s.write_restart(path)
    rec = R.read_records(path)
    assert np.array_equal(np.frombuffer(rec[3], dtype="<f8").reshape(3, 10, 10)[0, :mesh.p + 1, :mesh.p + 2], bx.T)
    assert np.array_equal(np.frombuffer(rec[6], dtype="<f8").reshape(3, 10, 10)[0, :mesh.p + 2, :mesh.p + 1], dmx.T)
    r = R.read_restart(path)
    assert np.array_equal(r.xg, g.xg) and np.array_equal(r.jacob, g.jacob)
    ncg = mesh.p + 1
    for e in (0, 13):
        gg = np.frombuffer(rec[13 + 24 * e + 19], dtype="<f8").reshape(10, 10, 10, 3, 3)   # gmetg: [l][m][n][j][i]
        for pt in (0, 7, ncg ** 3 - 1):
            n, m, l = pt % ncg, (pt // ncg) % ncg, pt // ncg ** 2
            assert np.array_equal(gg[l, m, n].T, g.gmetg[e, pt])
