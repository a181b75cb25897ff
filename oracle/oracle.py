"""ctypes access to the CPU oracle (liboracle.so) and to the reference leaf
routines compiled from the reference sources (_ref/libdsemref.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg, and only there, as the checker.  The
product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdsemref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class OrMesh(C.Structure):
    _fields_ = [("K", C.c_int32), ("p", C.c_int32), ("nmort", C.c_int32), ("pad_", C.c_int32),
                ("jacob", _dp), ("gmetg", _dp), ("gmetx", _dp), ("gmety", _dp), ("gmetz", _dp),
                ("m_elem", _ip), ("m_face", _ip), ("m_sign", _ip), ("m_nsign", _ip),
                ("nhat", _dp), ("norm", _dp), ("m_orient", _ip)]


class OrPhys(C.Structure):
    _fields_ = [("gamma", C.c_double), ("re", C.c_double), ("pr", C.c_double), ("mach", C.c_double)]


class OrEv(C.Structure):
    _fields_ = [("cmu", C.c_double), ("mumax", C.c_double), ("ducros", C.c_int32), ("smoothing", C.c_int32),
                ("cmax_cap", C.c_int32), ("pad_", C.c_int32), ("cmax", C.c_double)]


class OrDebug(C.Structure):
    _fields_ = [("stag", _dp), ("flux", _dp), ("grad", _dp), ("visc", _dp), ("eta", _dp)]


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref/ when the reference sources are present)."""
    out = subprocess.run(["make", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if out.returncode:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if os.path.isdir("/root/reference"):
        out = subprocess.run(["make", "-C", HERE, "ref"], capture_output=True, text=True)
        if out.returncode:
            raise RuntimeError("reference harness build failed:\n" + out.stdout + out.stderr)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_ops.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.or_rkcoefs.argtypes = [_dp, _dp, _dp]
        L.or_geometry.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_mortar_geometry.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp]
        L.or_euler_flux.argtypes = [C.POINTER(OrPhys), _dp, _dp, _dp, _dp]
        L.or_riemann.argtypes = [C.POINTER(OrPhys), _dp, C.c_double, _dp, _dp, _dp]
        L.or_vis_flux.argtypes = [C.POINTER(OrPhys), _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]
        L.or_rhs.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), _dp, _dp, _dp, C.POINTER(OrDebug), C.c_int]
        L.or_rk_step.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), _dp, _dp, C.c_double, _dp, _dp, _dp,
                                 C.c_int]
        L.or_dtcalc.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), _dp]
        L.or_ev_viscosity.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), C.POINTER(OrEv), _dp, _dp, _dp, _dp,
                                      _dp, _dp, _dp]
        L.or_rk_step_ev.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), C.POINTER(OrEv), _dp, _dp, C.c_double,
                                    _dp, C.c_int, _dp, _dp, C.c_int]
        L.or_dtcalc.restype = C.c_double
        L.or_set_mutation.argtypes = [C.c_int]
        L.or_modal_filter.argtypes = [C.POINTER(OrMesh), _dp, C.c_int]
        L.or_smagorinsky_sensors.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), C.c_double, C.c_double, _dp,
                                             _dp, C.c_int, _dp, _dp, _dp, _dp]
        L.or_rhs_les.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), _dp, _dp, _dp, _dp, C.POINTER(OrDebug),
                                 C.c_int]
        L.or_smagorinsky.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), C.c_double, C.c_double, _dp, _dp, _dp,
                                     _dp]
        L.or_rk_step_les.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), C.c_double, C.c_double, _dp, _dp,
                                     C.c_double, _dp, C.c_int, _dp, _dp, C.c_int]
        L.or_gauss_legendre.argtypes = [C.c_int, _dp, _dp]
        L.or_ke_weights.argtypes = [C.c_int, _dp]
        L.or_kinetic_energy.argtypes = [C.POINTER(OrMesh), _dp, _dp]
        L.or_enstrophy.argtypes = [C.POINTER(OrMesh), C.POINTER(OrPhys), _dp, _dp, _dp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_ops.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        R.ref_mortar_gauss_points.argtypes = [C.c_int, _dp]
        R.ref_rkcoefs.argtypes = [_dp, _dp, _dp]
        R.ref_mxm.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int]
        R.ref_interp.argtypes = [C.c_int, _dp, _dp]
        R.ref_d3dga.argtypes = [C.c_int, C.c_int, _dp, _dp]
        R.ref_gauss_deriv.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        R.ref_geometry.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        R.ref_dtcalc.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        R.ref_gaussq.argtypes = [C.c_int, _dp, _dp]
        R.ref_polyn.argtypes = [C.c_int, C.c_double, C.c_int, _dp]
        R.ref_polyn.restype = C.c_double
        _ref = R
    return _ref


def ptr(a: np.ndarray):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def phys(gamma=1.4, re=1600.0, pr=0.71, mach=1.0) -> OrPhys:
    return OrPhys(gamma, re, pr, mach)


# ---------------------------------------------------------------- operators
def ops(p: int):
    nc, ncg = p + 2, p + 1
    cx, cxg = np.zeros(nc), np.zeros(ncg)
    dmx, bx = np.zeros((ncg, nc)), np.zeros((nc, ncg))
    rc = lib().or_ops(p, ptr(cx), ptr(cxg), ptr(dmx), ptr(bx))
    assert rc == 0
    return cx, cxg, dmx, bx


def rkcoefs():
    a, b, c = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().or_rkcoefs(ptr(a), ptr(b), ptr(c))
    return a, b, c


class Geometry:
    """Per-point metrics of a mesh as computed by the oracle."""

    def __init__(self, mesh):
        K, p = mesh.K, mesh.p
        ncg, nc = p + 1, p + 2
        np_, ns, nf = ncg ** 3, nc * ncg * ncg, ncg * ncg
        self.corners = np.ascontiguousarray(mesh.corners, dtype=np.float64)
        self.xg = np.zeros((K, np_, 3))
        self.jacob = np.zeros((K, np_))
        self.gmetg = np.zeros((K, np_, 3, 3))
        self.gmet = [np.zeros((K, ns, 3)) for _ in range(3)]
        rc = lib().or_geometry(K, p, ptr(self.corners), ptr(self.xg), ptr(self.jacob), ptr(self.gmetg),
                               ptr(self.gmet[0]), ptr(self.gmet[1]), ptr(self.gmet[2]))
        assert rc == 0
        M = mesh.nmortar
        self.m_elem = np.ascontiguousarray(mesh.mortar_elem, dtype=np.int32)
        self.m_face = np.ascontiguousarray(mesh.mortar_face, dtype=np.int32)
        self.m_sign = np.ascontiguousarray(mesh.mortar_sign, dtype=np.int32)
        self.m_nsign = np.ascontiguousarray(mesh.mortar_nsign, dtype=np.int32)
        orient = getattr(mesh, "mortar_orient", None)
        self.m_orient = np.ascontiguousarray(np.zeros(M, dtype=np.int32) if orient is None else orient, dtype=np.int32)
        self.nhat = np.zeros((M, nf, 3))
        self.norm = np.zeros((M, nf))
        rc = lib().or_mortar_geometry(M, p, ptr(self.corners), ptr(self.m_elem), ptr(self.m_face),
                                      ptr(self.nhat), ptr(self.norm))
        assert rc == 0
        self.K, self.p, self.M = K, p, M
        self.np, self.ns, self.nf = np_, ns, nf

    def cmesh(self) -> OrMesh:
        return OrMesh(self.K, self.p, self.M, 0, ptr(self.jacob), ptr(self.gmetg), ptr(self.gmet[0]),
                      ptr(self.gmet[1]), ptr(self.gmet[2]), ptr(self.m_elem), ptr(self.m_face),
                      ptr(self.m_sign), ptr(self.m_nsign), ptr(self.nhat), ptr(self.norm), ptr(self.m_orient))


def rhs(geom: Geometry, JQ: np.ndarray, ph: OrPhys | None = None, muh=None, debug=False, nthreads=0):
    ph = ph or phys()
    JQ = np.ascontiguousarray(JQ, dtype=np.float64)
    Qt = np.zeros_like(JQ)
    m = geom.cmesh()
    dbg = None
    out = {}
    if debug:
        K = geom.K
        out = dict(stag=np.zeros((K, 3, 5, geom.ns)), flux=np.zeros((K, 3, 5, geom.ns)),
                   grad=np.zeros((K, 3, 5, geom.np)), visc=np.zeros((K, 3, 5, geom.ns)),
                   eta=np.zeros((geom.M, geom.nf, 2)))
        dbg = OrDebug(ptr(out["stag"]), ptr(out["flux"]), ptr(out["grad"]), ptr(out["visc"]), ptr(out["eta"]))
    mu = None if muh is None else np.ascontiguousarray(muh, dtype=np.float64)
    rc = lib().or_rhs(C.byref(m), C.byref(ph), ptr(JQ), ptr(mu), ptr(Qt),
                      C.byref(dbg) if dbg is not None else None, nthreads)
    assert rc == 0, rc
    return (Qt, out) if debug else Qt


def rk_step(geom: Geometry, JQ: np.ndarray, dt: float, ph: OrPhys | None = None, muh=None, nthreads=0,
            stag_out=None):
    """Advance JQ in place by one RK4(5) step; returns (JQ, G, resid).
    stag_out [K,3,5,ns] (optional) receives the last stage's face-averaged
    staggered states."""
    ph = ph or phys()
    G = np.zeros_like(JQ)
    res = C.c_double(0.0)
    m = geom.cmesh()
    mu = None if muh is None else np.ascontiguousarray(muh, dtype=np.float64)
    if stag_out is not None:
        assert stag_out.flags.c_contiguous and stag_out.dtype == np.float64
    rc = lib().or_rk_step(C.byref(m), C.byref(ph), ptr(JQ), ptr(G), dt, ptr(mu), ptr(stag_out), C.byref(res),
                          nthreads)
    assert rc == 0, rc
    return JQ, G, res.value


def dtcalc(geom: Geometry, JQ: np.ndarray, ph: OrPhys | None = None) -> float:
    ph = ph or phys()
    m = geom.cmesh()
    return lib().or_dtcalc(C.byref(m), C.byref(ph), ptr(np.ascontiguousarray(JQ)))


def euler_flux(Q, ph=None):
    ph = ph or phys()
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    f, g, h = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().or_euler_flux(C.byref(ph), ptr(Q), ptr(f), ptr(g), ptr(h))
    return f, g, h


def modal_filter(geom: Geometry, JQ, cutoff=None):
    """modalfilter on a copy of JQ [K,5,np]; cutoff None keeps every mode."""
    out = np.array(JQ, dtype=np.float64, order="C", copy=True)
    m = geom.cmesh()
    rc = lib().or_modal_filter(C.byref(m), ptr(out), geom.p + 1 if cutoff is None else int(cutoff))
    assert rc == 0
    return out


MUT_ENTROPY_FIX_SIGN, MUT_SWAP_SIDES, MUT_ORIENT_INVERSE = 1, 2, 4


class mutation:
    """Context manager planting a mistake in the oracle (tests' controls only):
    MUT_ENTROPY_FIX_SIGN flips the sign of the Roe flux's entropy-fix term,
    MUT_SWAP_SIDES swaps the left/right states of the interface flux,
    MUT_ORIENT_INVERSE reads B2F1 and F2B1 slave faces with each other's
    point map (the inverse map)."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        lib().or_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().or_set_mutation(0)
        return False


def riemann(nhat, ds, Ql, Qr, ph=None):
    ph = ph or phys()
    out = np.zeros(5)
    lib().or_riemann(C.byref(ph), ptr(np.ascontiguousarray(nhat, dtype=np.float64)), float(ds),
                     ptr(np.ascontiguousarray(Ql, dtype=np.float64)),
                     ptr(np.ascontiguousarray(Qr, dtype=np.float64)), ptr(out))
    return out


def vis_flux(Q, Qx, Qy, Qz, amu=1.0, ph=None):
    ph = ph or phys()
    f, g, h = np.zeros(5), np.zeros(5), np.zeros(5)
    a = [np.ascontiguousarray(t, dtype=np.float64) for t in (Q, Qx, Qy, Qz)]
    lib().or_vis_flux(C.byref(ph), *[ptr(t) for t in a], float(amu), ptr(f), ptr(g), ptr(h))
    return f, g, h


# ---------------------------------------------------------------- entropy viscosity
def ev_viscosity(geom: Geometry, JQ, stag_lag, ev: OrEv, ph: OrPhys | None = None, rho_sensor=False):
    """(muhg [K,np], muh [K,3,ns], resid [K,np], ducros [K,np]) and, with
    rho_sensor, the density sensor [K,np] as a fifth entry."""
    ph = ph or phys()
    K = geom.K
    muhg, resid, duc = np.zeros((K, geom.np)), np.zeros((K, geom.np)), np.zeros((K, geom.np))
    rs = np.zeros((K, geom.np))
    muh = np.zeros((K, 3, geom.ns))
    m = geom.cmesh()
    rc = lib().or_ev_viscosity(C.byref(m), C.byref(ph), C.byref(ev), ptr(np.ascontiguousarray(JQ)),
                               ptr(np.ascontiguousarray(stag_lag)), ptr(muhg), ptr(muh), ptr(resid), ptr(duc),
                               ptr(rs))
    assert rc == 0
    return (muhg, muh, resid, duc, rs) if rho_sensor else (muhg, muh, resid, duc)


def rk_step_ev(geom: Geometry, JQ, dt, stag_lag, have_lag, ev: OrEv, ph: OrPhys | None = None, nthreads=0):
    """One EV step in place; stag_lag [K,3,5,ns] updated in place.  Returns (JQ, muh, resid)."""
    ph = ph or phys()
    G = np.zeros_like(JQ)
    muh = np.zeros((geom.K, 3, geom.ns))
    res = C.c_double(0.0)
    m = geom.cmesh()
    rc = lib().or_rk_step_ev(C.byref(m), C.byref(ph), C.byref(ev), ptr(JQ), ptr(G), dt, ptr(stag_lag),
                             int(have_lag), ptr(muh), C.byref(res), nthreads)
    assert rc == 0
    return JQ, muh, res.value


# ---------------------------------------------------------------- diagnostics
def gauss_legendre(n: int):
    t, w = np.zeros(n), np.zeros(n)
    lib().or_gauss_legendre(n, ptr(t), ptr(w))
    return t, w


def ke_weights(p: int) -> np.ndarray:
    """bmatu1b [ncg^3] (index i + ncg(j + ncg l))."""
    b = np.zeros((p + 1) ** 3)
    assert lib().or_ke_weights(p, ptr(b)) == 0
    return b


def kinetic_energy(geom: Geometry, JQ) -> float:
    out = C.c_double(0.0)
    m = geom.cmesh()
    assert lib().or_kinetic_energy(C.byref(m), ptr(np.ascontiguousarray(JQ)), C.byref(out)) == 0
    return out.value


def enstrophy(geom: Geometry, JQ, stag_lag, ph: OrPhys | None = None) -> float:
    ph = ph or phys()
    out = C.c_double(0.0)
    m = geom.cmesh()
    assert lib().or_enstrophy(C.byref(m), C.byref(ph), ptr(np.ascontiguousarray(JQ)),
                              ptr(np.ascontiguousarray(stag_lag)), C.byref(out)) == 0
    return out.value


# ---------------------------------------------------------------- static Smagorinsky
LES_SHOCK_SENSOR, LES_RHO_SENSOR = 1, 2


def smagorinsky(geom: Geometry, JQ, stag_lag, cs: float, numax: float, ph: OrPhys | None = None, sensors: int = 0,
                ducros=None, rho_sensor=None):
    """(nu_t at the Gauss points [K,np], nu_t on the staggered grids [K,3,ns]);
    sensors: LES_SHOCK_SENSOR | LES_RHO_SENSOR with their Gauss-point fields."""
    ph = ph or phys()
    g = np.zeros((geom.K, geom.np))
    st = np.zeros((geom.K, 3, geom.ns))
    m = geom.cmesh()
    d = None if ducros is None else np.ascontiguousarray(ducros, dtype=np.float64)
    r = None if rho_sensor is None else np.ascontiguousarray(rho_sensor, dtype=np.float64)
    assert lib().or_smagorinsky_sensors(C.byref(m), C.byref(ph), cs, numax, ptr(np.ascontiguousarray(JQ)),
                                        ptr(np.ascontiguousarray(stag_lag)), int(sensors), ptr(d), ptr(r),
                                        ptr(g), ptr(st)) == 0
    return g, st


def rhs_les(geom: Geometry, JQ, nut, ph: OrPhys | None = None, nthreads=0):
    ph = ph or phys()
    Qt = np.zeros_like(JQ)
    m = geom.cmesh()
    assert lib().or_rhs_les(C.byref(m), C.byref(ph), ptr(np.ascontiguousarray(JQ)), None,
                            ptr(np.ascontiguousarray(nut, dtype=np.float64)), ptr(Qt), None, nthreads) == 0
    return Qt


def rk_step_les(geom: Geometry, JQ, dt, stag_lag, have_lag, cs, numax, ph: OrPhys | None = None, nthreads=0):
    """One static-Smagorinsky step in place; stag_lag updated in place.  Returns (JQ, nut_s, resid)."""
    ph = ph or phys()
    G = np.zeros_like(JQ)
    nut = np.zeros((geom.K, 3, geom.ns))
    res = C.c_double(0.0)
    m = geom.cmesh()
    assert lib().or_rk_step_les(C.byref(m), C.byref(ph), cs, numax, ptr(JQ), ptr(G), dt, ptr(stag_lag),
                                int(have_lag), ptr(nut), C.byref(res), nthreads) == 0
    return JQ, nut, res.value
