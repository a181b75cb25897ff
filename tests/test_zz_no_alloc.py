"""dsem_rhs_device_noalloc (RHS intermediates in the library's static device
array, no allocation) against dsem_rhs_device (the context's buffers), called
through the C ABI on device buffers: bit-for-bit equal, on two calls in a row
whose second has the smaller mesh (the array keeps the first call's
contents), and a mesh beyond the array's size is refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from dsem_amd import _lib as L
from dsem_amd import cases
from dsem_amd import mesh as M
from dsem_amd.solver import DSEM
from helpers import distorted_box

pytestmark = pytest.mark.gpu


def _both(mesh, seed):
    """(Qt of dsem_rhs_device, Qt of dsem_rhs_device_noalloc) on one state."""
    lib = L.load()
    with DSEM(mesh, re=1600.0) as s:
        xg, jac = s.gauss_points()
        JQ = cases.tgv_state(xg, jac, "isothermal", perturb=1e-2, seed=seed)
        dJQ = torch.from_numpy(JQ).to("cuda")
        ref = torch.full_like(dJQ, np.nan)
        got = torch.full_like(dJQ, np.nan)
        torch.cuda.synchronize()
        print(f"[case] rhs_device K={mesh.K} p={mesh.p}", flush=True)
        assert lib.dsem_rhs_device(s._h, C.c_void_p(dJQ.data_ptr()), C.c_void_p(ref.data_ptr())) == 0
        assert lib.dsem_synchronize(s._h) == 0
        print(f"[case] rhs_device_noalloc K={mesh.K} p={mesh.p}", flush=True)
        rc = lib.dsem_rhs_device_noalloc(s._h, C.c_void_p(dJQ.data_ptr()), C.c_void_p(got.data_ptr()))
        assert rc == 0, lib.dsem_last_error(s._h).decode()
        assert lib.dsem_synchronize(s._h) == 0
        return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_rhs_equals_rhs_device_on_two_calls():
    big = distorted_box(4, 6, amp=0.05, seed=11)     # 64 elements, p = 6
    small = distorted_box(2, 6, amp=0.05, seed=12)   # 8 elements: reuses a dirty array
    for mesh, seed in ((big, 1), (small, 2)):
        ref, got = _both(mesh, seed)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(ref, got), np.abs(ref - got).max()


def test_noalloc_rhs_refuses_a_mesh_beyond_the_array():
    lib = L.load()
    mesh = M.periodic_box(9, 8)   # 729 elements at p = 8: about 71 MB of intermediates
    with DSEM(mesh) as s:
        d = torch.zeros(s.dof, dtype=torch.float64, device="cuda")
        o = torch.zeros_like(d)
        torch.cuda.synchronize()
        rc = lib.dsem_rhs_device_noalloc(s._h, C.c_void_p(d.data_ptr()), C.c_void_p(o.data_ptr()))
        assert rc == L.DSEM_ERR_ARG
        assert "bytes of scratch" in lib.dsem_last_error(s._h).decode()
