"""Oracle vs the reference's own leaf routines, compiled from the reference
sources into oracle/_ref/libdsemref.so (see oracle/Makefile).

Every comparison is bit-for-bit: the oracle restates each routine in the
reference's floating-point order.  Skipped where the reference sources (and
hence the harness library) are absent.
"""
import numpy as np
import pytest

import oracle as O
from dsem_amd import cases
from dsem_amd import mesh as M

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (no reference sources)")

NX = 10


def _pad_gauss(a, ncg):
    U = np.zeros((NX, NX, NX))
    U[:ncg, :ncg, :ncg] = a.reshape(ncg, ncg, ncg)
    return U


def _pad_stag(a, d, ncg, nc):
    U = np.zeros((NX, NX, NX))
    if d == 0:
        U[:ncg, :ncg, :nc] = a.reshape(ncg, ncg, nc)
    elif d == 1:
        U[:ncg, :nc, :ncg] = a.reshape(ncg, nc, ncg)
    else:
        U[:nc, :ncg, :ncg] = a.reshape(nc, ncg, ncg)
    return U


def _unpad_stag(U, d, ncg, nc):
    if d == 0:
        return U[:ncg, :ncg, :nc].reshape(-1)
    if d == 1:
        return U[:ncg, :nc, :ncg].reshape(-1)
    return U[:nc, :ncg, :ncg].reshape(-1)


@pytest.mark.parametrize("p", range(1, 9))
def test_points_and_operators_bitwise(oracle, p):
    nc, ncg = p + 2, p + 1
    cx, cxg, dmx, bx = O.ops(p)
    rcx, rcxg = np.zeros(nc), np.zeros(ncg)
    rd, rb = np.zeros((ncg, nc)), np.zeros((nc, ncg))
    O.ref().ref_ops(p, O.ptr(rcx), O.ptr(rcxg), O.ptr(rd), O.ptr(rb))
    assert np.array_equal(cx, rcx) and np.array_equal(cxg, rcxg)
    assert np.array_equal(dmx, rd), np.abs(dmx - rd).max()
    assert np.array_equal(bx, rb), np.abs(bx - rb).max()
    mg = np.zeros(ncg)
    O.ref().ref_mortar_gauss_points(ncg, O.ptr(mg))
    assert np.array_equal(mg, cxg)


def test_rk_coefficients_bitwise(oracle):
    a, b, c = O.rkcoefs()
    ra, rb, rc = np.zeros(5), np.zeros(5), np.zeros(5)
    O.ref().ref_rkcoefs(O.ptr(ra), O.ptr(rb), O.ptr(rc))
    assert np.array_equal(a, ra) and np.array_equal(b, rb) and np.array_equal(c, rc)


def _distorted_corners(seed=3):
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
    return (base * 0.7 + 0.12 * rng.standard_normal((8, 3)))


@pytest.mark.parametrize("p", [2, 4, 8])
@pytest.mark.parametrize("kind", ["box", "distorted"])
def test_geometry_bitwise(oracle, p, kind):
    if kind == "box":
        c = M.periodic_box(3, p).corners[5]
    else:
        c = _distorted_corners()
    ncg, nc = p + 1, p + 2
    corners = np.ascontiguousarray(c[None], dtype=np.float64)
    xg = np.zeros((1, ncg ** 3, 3)); jac = np.zeros((1, ncg ** 3)); gg = np.zeros((1, ncg ** 3, 3, 3))
    gm = [np.zeros((1, nc * ncg * ncg, 3)) for _ in range(3)]
    O.lib().or_geometry(1, p, O.ptr(corners), O.ptr(xg), O.ptr(jac), O.ptr(gg), O.ptr(gm[0]), O.ptr(gm[1]), O.ptr(gm[2]))
    rxg = np.zeros((NX, NX, NX, 3)); rjac = np.zeros((NX, NX, NX))
    rgg = np.zeros((NX, NX, NX, 3, 3)); rgm = np.zeros((NX, NX, NX, 3, 3))
    rnh = np.zeros((6, NX, NX, 3)); rno = np.zeros((6, NX, NX))
    cf = np.ascontiguousarray(c, dtype=np.float64)
    O.ref().ref_geometry(p, O.ptr(cf), O.ptr(rxg), O.ptr(rjac), O.ptr(rgg), O.ptr(rgm), O.ptr(rnh), O.ptr(rno))
    assert np.array_equal(xg[0].reshape(ncg, ncg, ncg, 3), rxg[:ncg, :ncg, :ncg])
    assert np.array_equal(jac[0].reshape(ncg, ncg, ncg), rjac[:ncg, :ncg, :ncg])
    # gmetg[pt, i, j] == Fortran gmetg(i+1, j+1, ...) stored [l,m,n,j,i]
    assert np.array_equal(gg[0].reshape(ncg, ncg, ncg, 3, 3), np.swapaxes(rgg[:ncg, :ncg, :ncg], -1, -2))
    assert np.array_equal(gm[0][0].reshape(ncg, ncg, nc, 3), rgm[:ncg, :ncg, :nc, :, 0])
    assert np.array_equal(gm[1][0].reshape(ncg, nc, ncg, 3), rgm[:ncg, :nc, :ncg, :, 1])
    assert np.array_equal(gm[2][0].reshape(nc, ncg, ncg, 3), rgm[:nc, :ncg, :ncg, :, 2])
    # mortar normals: each face as master of its own mortar
    me = np.zeros((6, 2), dtype=np.int32)
    mf = np.ascontiguousarray(np.repeat(np.arange(1, 7, dtype=np.int32)[:, None], 2, axis=1))
    nh = np.zeros((6, ncg * ncg, 3)); no = np.zeros((6, ncg * ncg))
    O.lib().or_mortar_geometry(6, p, O.ptr(corners), O.ptr(me), O.ptr(mf), O.ptr(nh), O.ptr(no))
    for f in range(6):
        assert np.array_equal(nh[f].reshape(ncg, ncg, 3), rnh[f, :ncg, :ncg]), f + 1
        assert np.array_equal(no[f].reshape(ncg, ncg), rno[f, :ncg, :ncg]), f + 1


def _tiny_case(p=3, n=2, variant="isothermal", perturb=1e-3):
    mesh = M.periodic_box(n, p)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, variant, perturb=perturb)
    return mesh, g, JQ


@pytest.mark.parametrize("p", [3, 8])
def test_interpolation_gradient_divergence_bitwise(oracle, p):
    """Int_fluxes interpolation (mxm), Gauss_Deriv and the update divergence
    inside or_rhs equal the reference routines applied to the same data."""
    mesh, g, JQ = _tiny_case(p=p)
    ph = O.phys(re=50.0)
    Qt, dbg = O.rhs(g, JQ, ph, debug=True)
    ncg, nc, npt = p + 1, p + 2, (p + 1) ** 3
    _, _, dmx, bx = O.ops(p)
    R = O.ref()
    bpad = np.zeros((NX, NX)); bpad[:ncg, :nc] = bx.T     # Fortran b(i,k) = bx[i][k]
    bt = np.zeros((NX, NX)); bt[:nc, :ncg] = bx            # Fortran dmatt(m,j) = bx[j][m]
    for e in range(mesh.K):
        Qp = JQ[e] * g.jacob[e][None]
        for v in range(5):
            U = _pad_gauss(Qp[v], ncg)
            outs = []
            o = np.zeros((NX, NX, NX)); R.ref_mxm(O.ptr(bpad), NX, O.ptr(U), NX, O.ptr(o), NX * NX); outs.append(o)
            o = np.zeros((NX, NX, NX))
            for k in range(NX):
                Uk = np.ascontiguousarray(U[k]); ok = np.zeros((NX, NX))
                R.ref_mxm(O.ptr(Uk), NX, O.ptr(bt), NX, O.ptr(ok), NX)
                o[k] = ok
            outs.append(o)
            o = np.zeros((NX, NX, NX)); R.ref_mxm(O.ptr(U), NX * NX, O.ptr(bt), NX, O.ptr(o), NX); outs.append(o)
            for d in range(3):
                ref_s = _unpad_stag(outs[d], d, ncg, nc)
                got = dbg["stag"][e, d, v]
                # interior Lobatto positions (faces hold the mortar average)
                jj = np.arange(nc * ncg * ncg)
                if d == 0:
                    jpos = jj % nc
                elif d == 1:
                    jpos = (jj // ncg) % nc
                else:
                    jpos = jj // (ncg * ncg)
                inner = (jpos > 0) & (jpos < nc - 1)
                assert np.array_equal(got[inner], ref_s[inner])
        # Gauss_Deriv on the face-averaged staggered states
        gp = np.swapaxes(g.gmetg[e].reshape(ncg, ncg, ncg, 3, 3), -1, -2)
        G = np.zeros((NX, NX, NX, 3, 3)); G[:ncg, :ncg, :ncg] = gp
        for v in range(5):
            W = [_pad_stag(dbg["stag"][e, d, v], d, ncg, nc) for d in range(3)]
            wx, wy, wz = np.zeros((NX, NX, NX)), np.zeros((NX, NX, NX)), np.zeros((NX, NX, NX))
            R.ref_gauss_deriv(p, O.ptr(G), O.ptr(W[0]), O.ptr(W[1]), O.ptr(W[2]), O.ptr(wx), O.ptr(wy), O.ptr(wz))
            for c, w in enumerate((wx, wy, wz)):
                assert np.array_equal(dbg["grad"][e, c, v], w[:ncg, :ncg, :ncg].reshape(-1))
        # update_interior divergence: Qt = -(dx3dga f + dy3dga g + dz3dga h)
        for v in range(5):
            outs = []
            for d in range(3):
                U = _pad_stag(dbg["flux"][e, d, v], d, ncg, nc)
                o = np.zeros((NX, NX, NX))
                R.ref_d3dga(p, d + 1, O.ptr(U), O.ptr(o))
                outs.append(o[:ncg, :ncg, :ncg].reshape(-1))
            ref_qt = -(outs[0] + outs[1] + outs[2])
            assert np.array_equal(Qt[e, v], ref_qt)


def test_prolongation_and_viscous_assembly_bitwise(oracle):
    """Vis_Fluxes: prolongation with the reference interp/mxv, then the
    oracle's pointwise viscous flux, reproduce r,s,t exactly."""
    p = 3
    mesh, g, JQ = _tiny_case(p=p)
    ph = O.phys(re=50.0)
    _, dbg = O.rhs(g, JQ, ph, debug=True)
    ncg, nc = p + 1, p + 2
    R = O.ref()
    for e in range(mesh.K):
        Ge = dbg["grad"][e].reshape(3, 5, ncg, ncg, ncg)   # [c, v, l, m, n]
        for d in range(3):
            S = dbg["stag"][e, d]
            gm = g.gmet[d][e]
            for b in range(ncg):
                for a in range(ncg):
                    lines = np.zeros((3, 5, nc))
                    for c in range(3):
                        for v in range(5):
                            if d == 0:
                                u = Ge[c, v, b, a, :]
                            elif d == 1:
                                u = Ge[c, v, b, :, a]
                            else:
                                u = Ge[c, v, :, b, a]
                            u = np.ascontiguousarray(u); ugl = np.zeros(nc)
                            R.ref_interp(p, O.ptr(u), O.ptr(ugl))
                            lines[c, v] = ugl
                    for j in range(nc):
                        if d == 0:
                            si = j + nc * (a + ncg * b)
                        elif d == 1:
                            si = a + ncg * (j + nc * b)
                        else:
                            si = a + ncg * (b + ncg * j)
                        f, gg, h = O.vis_flux(S[:, si], lines[0, :, j], lines[1, :, j], lines[2, :, j], 1.0, ph)
                        r = gm[si, 0] * f + gm[si, 1] * gg + gm[si, 2] * h
                        assert np.array_equal(dbg["visc"][e, d, :, si], r)


@pytest.mark.parametrize("p", [4, 8])
def test_dtcalc_bitwise(oracle, p):
    mesh, g, JQ = _tiny_case(p=p, n=2)
    ph = O.phys()
    got = O.dtcalc(g, JQ, ph)
    ncg, nc = p + 1, p + 2
    dt = np.finfo(float).max
    for e in range(mesh.K):
        q = np.zeros((5, NX, NX, NX))
        for v in range(5):
            q[v] = _pad_gauss(JQ[e, v], ncg)
        jac = _pad_gauss(g.jacob[e], ncg)
        gm = np.zeros((NX, NX, NX, 3, 3))
        gm[:ncg, :ncg, :nc, :, 0] = g.gmet[0][e].reshape(ncg, ncg, nc, 3)
        gm[:ncg, :nc, :ncg, :, 1] = g.gmet[1][e].reshape(ncg, nc, ncg, 3)
        gm[:nc, :ncg, :ncg, :, 2] = g.gmet[2][e].reshape(nc, ncg, ncg, 3)
        import ctypes
        d = ctypes.c_double(dt)
        O.ref().ref_dtcalc(p, 1.4, O.ptr(q), O.ptr(jac), O.ptr(gm), ctypes.byref(d))
        dt = d.value
    assert got == dt


def test_box_matches_reference_mesh_file():
    """The generated 22^3 box equals the reference's tgv-22x22x22-p8.mesh."""
    import os
    path = "/root/reference/p8-22x22x22/tgv-22x22x22-p8.mesh"
    if not os.path.exists(path):
        pytest.skip("reference mesh file absent")
    a = M.read_mesh(path)
    b = M.periodic_box(22, 8)
    assert a.p == 8 and a.K == 10648 and a.nmortar == 3 * 10648
    assert np.array_equal(a.corners, b.corners)
    ka = set(map(tuple, np.concatenate([a.mortar_elem, a.mortar_face], 1).tolist()))
    kb = set(map(tuple, np.concatenate([b.mortar_elem, b.mortar_face], 1).tolist()))
    assert ka == kb
    assert np.all(a.mortar_sign == 1) and np.all(a.mortar_nsign == 1)
