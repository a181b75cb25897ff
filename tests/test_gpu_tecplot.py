"""Tecplot output of the GPU state: the interior points of every zone hold
the solver's Gauss-point values (to the 8 digits of the file)."""
import numpy as np
import pytest

from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd import tecplot as TP
from dsem_amd.solver import DSEM

pytestmark = pytest.mark.gpu


def test_write_tecplot_of_solver_state(tmp_path):
    print("\n[case] tecplot of a 2^3 p=3 TGV state", flush=True)
    mesh = M.periodic_box(2, 3)
    with DSEM(mesh, mach=0.1) as s:
        xg, jac = s.gauss_points()
        s.set_state(cases.tgv_state(xg, jac, "reference"))
        s.steps(2, 0.3 * s.dtcalc())
        JQ = s.get_state()
        path = tmp_path / "tgv.plt"
        s.write_tecplot(str(path), time=0.5, scale=True)
    t, zones = TP.read_plt(str(path))
    assert t == 0.5 and len(zones) == mesh.K
    ncg, N = 4, 6
    Q = JQ * jac[:, None, :]
    for e in range(mesh.K):
        Z = zones[e].reshape(N, N, N, 11)[1:-1, 1:-1, 1:-1].reshape(-1, 11)   # interior, l m n order
        assert np.allclose(Z[:, 0:3], xg[e], rtol=5.1e-8, atol=1e-12)
        assert np.allclose(Z[:, 6], Q[e, 0], rtol=5.1e-8, atol=0)
        assert np.allclose(Z[:, 3], Q[e, 1] / Q[e, 0], rtol=5.1e-8, atol=1e-12)
