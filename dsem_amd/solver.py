"""Host-side mirror of the reference's time-stepping interface on the
MI355X library.

``DSEM.tstep(time, dt)`` stands where the driver calls
``tstep(time,dt,d,ngrid,mrtr,nmort,resid)`` (src/tstep.f90:12,
src/3DDriver_parfmdf.f90:466); ``DSEM.rhs`` is ``Fluxes`` followed by the
divergence of ``update`` (fluxesH_newa.f90:17, updatea.f90:100-113);
``DSEM.dtcalc`` is the driver's time-step computation
(3DDriver_parfmdf.f90:362-370 with dtcalc.f90:9-58).  A NaN/Inf residual
raises, where the reference writes NAN.dat and aborts
(3DDriver_parfmdf.f90:489-499).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import Mesh

_PREC = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32, "bf16x3": L.PREC_BF16X3}
_METRICS = {"auto": L.METRICS_AUTO, "per_point": L.METRICS_PER_POINT, "per_element": L.METRICS_PER_ELEMENT}


class NonFiniteResidual(L.DsemError):
    pass


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class DSEM:
    def __init__(self, mesh: Mesh, gamma: float = 1.4, re: float = 1600.0, pr: float = 0.71,
                 mach: float = 1.0, shock: bool = False, precision: str = "fp64",
                 metrics: str = "per_point", device: int = 0, cmu: float = 1.0, mumax: float = 1.0,
                 cmax: float = 0.5, ducros: bool = False, smoothing: bool = False, cmax_cap: bool = False):
        self._L = L.load()
        self.mesh = mesh
        self._keep = [np.ascontiguousarray(mesh.mortar_elem, dtype=np.int32),
                      np.ascontiguousarray(mesh.mortar_face, dtype=np.int32),
                      np.ascontiguousarray(mesh.mortar_sign, dtype=np.int32),
                      np.ascontiguousarray(mesh.mortar_nsign, dtype=np.int32),
                      np.ascontiguousarray(mesh.corners, dtype=np.float64),
                      np.ascontiguousarray(np.zeros(mesh.nmortar) if mesh.mortar_orient is None else mesh.mortar_orient,
                                           dtype=np.int32)]
        me, mf, ms, mn, cr, mo = self._keep
        desc = L.MeshDesc(mesh.K, mesh.p, mesh.nmortar, 0, _ip(me), _ip(mf), _ip(ms), _ip(mn), _ip(mo), _dp(cr))
        flags = (L.EV_DUCROS if ducros else 0) | (L.EV_SMOOTHING if smoothing else 0) | \
            (L.EV_CMAX_CAP if cmax_cap else 0)
        prm = L.Params(gamma, re, pr, mach, int(bool(shock)), _PREC[precision], _METRICS[metrics], flags,
                       cmu, mumax, cmax)
        h = C.c_void_p()
        rc = self._L.dsem_create(C.byref(desc), C.byref(prm), device, C.byref(h))
        self._h = h
        if rc != L.DSEM_OK:
            msg = self._L.dsem_last_error(h).decode() if h.value else "create failed"
            self._L.dsem_destroy(h)
            self._h = C.c_void_p()
            raise L.DsemError(rc, msg)
        self.K, self.p = mesh.K, mesh.p
        self.ncg, self.nc = mesh.p + 1, mesh.p + 2
        self.np, self.ns = self.ncg ** 3, self.nc * self.ncg ** 2
        self.dof = 5 * self.K * self.np
        self.params = dict(gamma=gamma, re=re, pr=pr, mach=mach, shock=shock, precision=precision)
        self.les = False            # static Smagorinsky model (set_les)

    # ------------------------------------------------------------------ util
    def _check(self, rc: int):
        if rc == L.DSEM_OK:
            return
        msg = self._L.dsem_last_error(self._h).decode()
        if rc == L.DSEM_ERR_NONFINITE:
            raise NonFiniteResidual(rc, msg)
        raise L.DsemError(rc, msg)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.dsem_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------ data
    def write_tecplot(self, path: str, time: float = 0.0, scale: bool = False, ev: np.ndarray | None = None,
                      nut: np.ndarray | None = None) -> None:
        """Tecplot `.plt` of the current state as the reference's wrtdat writes
        it (dsem_amd.tecplot); scale=True divides out J (the movie files)."""
        from . import tecplot
        xg, jac = self.gauss_points()
        tecplot.write_plt(path, time, xg, self.get_state(), jac, self.p, self.params["gamma"], self.params["mach"],
                          scale, ev, nut)

    # ------------------------------------------------------------- restart files
    def write_restart(self, path: str, time: float = 0.0) -> None:
        """Write the current state as the reference's restart file
        (write_restart, src/Parabolic3D/stac3m/restart.f90:19-312; layout in
        dsem_amd.restart): J*Q, the lagged staggered states when the context
        keeps them (shock = 1 or keep_lagged_states), geometry, operators,
        topology and ``time``."""
        from . import restart as R
        M_, nf = self.mesh.nmortar, self.ncg * self.ncg
        cx, cxg = np.zeros(self.nc), np.zeros(self.ncg)
        self._check(self._L.dsem_points(self._h, _dp(cx), _dp(cxg)))
        dmx, bx = np.zeros((self.ncg, self.nc)), np.zeros((self.nc, self.ncg))
        self._check(self._L.dsem_operators(self._h, _dp(dmx), _dp(bx)))
        gmetg, gmet = np.zeros((self.K, self.np, 9)), np.zeros((self.K, 3, self.ns, 3))
        nhat, norm = np.zeros((M_, nf, 3)), np.zeros((M_, nf))
        self._check(self._L.dsem_get_metrics(self._h, _dp(gmetg), _dp(gmet), _dp(nhat), _dp(norm)))
        xg, jac = self.gauss_points()
        try:
            stag = self.get_lagged_states()
        except L.DsemError:
            stag = None
        R.write_restart(path, self.mesh, self.get_state(), cx=cx, cxg=cxg, dmx=dmx, bx=bx, xg=xg, jacob=jac,
                        gmetg=gmetg, gmet=gmet, nhat=nhat, norm=norm, time=time, stag=stag)

    def load_restart(self, rst) -> float:
        """Take the state of a restart file (dsem_amd.restart.read_restart's
        Restart, same mesh) and, when the context keeps them and the file has
        them, the lagged staggered states; returns the file's time."""
        if rst.JQ.shape != (self.K, 5, self.np):
            raise ValueError("the restart file holds a different mesh")
        self.set_state(rst.JQ)
        if np.any(rst.stag != 0.0):
            try:
                self.set_lagged_states(rst.stag)
            except L.DsemError:
                pass
        return rst.time

    @classmethod
    def from_restart(cls, path: str, **kw):
        """A context on the mesh of a restart file, holding its state; returns
        (solver, time)."""
        from . import restart as R
        rst = R.read_restart(path)
        s = cls(rst.mesh, **kw)
        return s, s.load_restart(rst)

    def gauss_points(self):
        xg = np.zeros((self.K, self.np, 3))
        jac = np.zeros((self.K, self.np))
        self._check(self._L.dsem_gauss_points(self._h, _dp(xg), _dp(jac)))
        return xg, jac

    def set_state(self, JQ: np.ndarray, layout: int = 0):
        shape = (self.K, 5, self.np) if layout == 0 else (self.K, 5, 10, 10, 10)
        JQ = np.ascontiguousarray(JQ, dtype=np.float64)
        if JQ.size != int(np.prod(shape)):
            raise ValueError(f"state must have {int(np.prod(shape))} values for layout {layout}")
        self._check(self._L.dsem_set_state(self._h, _dp(JQ), layout))

    def get_state(self, layout: int = 0) -> np.ndarray:
        shape = (self.K, 5, self.np) if layout == 0 else (self.K, 5, 10, 10, 10)
        out = np.zeros(shape)
        self._check(self._L.dsem_get_state(self._h, _dp(out), layout))
        return out

    def set_viscosity(self, muh: np.ndarray):
        muh = np.ascontiguousarray(muh, dtype=np.float64)
        if muh.size != self.K * 3 * self.ns:
            raise ValueError("muh must be [K][3][nc*ncg^2]")
        self._check(self._L.dsem_set_viscosity(self._h, _dp(muh)))

    def get_viscosity(self):
        """(mu_h at the Gauss points [K, np], mu_h on the staggered grids [K, 3, ns])."""
        g = np.zeros((self.K, self.np))
        st = np.zeros((self.K, 3, self.ns))
        self._check(self._L.dsem_get_viscosity(self._h, _dp(g), _dp(st)))
        return g, st

    def ev_update(self):
        """Entropy (EV_routines/Main/Driver.f90:21) on the current state; returns
        (mu_h, residual D, Ducros sensor) at the Gauss points."""
        mu, r, dc = (np.zeros((self.K, self.np)) for _ in range(3))
        self._check(self._L.dsem_ev_update(self._h, _dp(mu), _dp(r), _dp(dc)))
        return mu, r, dc

    def set_lagged_states(self, stag: np.ndarray):
        stag = np.ascontiguousarray(stag, dtype=np.float64)
        if stag.size != self.K * 3 * 5 * self.ns:
            raise ValueError("lagged states must be [K][3][5][nc*ncg^2]")
        self._check(self._L.dsem_set_lagged_states(self._h, _dp(stag)))

    def get_lagged_states(self) -> np.ndarray:
        out = np.zeros((self.K, 3, 5, self.ns))
        self._check(self._L.dsem_get_lagged_states(self._h, _dp(out)))
        return out

    # ------------------------------------------------------------- LES model
    def set_les(self, enable: bool = True, cs: float = 0.17, numax: float = 1.0) -> None:
        """Static Smagorinsky model (staticSmag.f90): nu_t = Re (C_s Delta)^2 |S|."""
        self._check(self._L.dsem_set_les(self._h, int(enable), cs, numax))
        self.les = bool(enable)

    def set_les_sensors(self, shock_sensor: bool = False, rho_sensor: bool = False) -> None:
        """Smagorinsky sensor switches (shockSensor / rhoSensor, staticSmag.f90:68-69)."""
        flags = (L.LES_SHOCK_SENSOR if shock_sensor else 0) | (L.LES_RHO_SENSOR if rho_sensor else 0)
        self._check(self._L.dsem_set_les_sensors(self._h, flags))

    def ev_sensors(self):
        """(Ducros sensor, density sensor) at the Gauss points [K,np] from the
        last entropy-viscosity update."""
        d = np.zeros((self.K, self.np))
        r = np.zeros((self.K, self.np))
        self._check(self._L.dsem_get_ev_sensors(self._h, _dp(d), _dp(r)))
        return d, r

    def modal_filter(self, cutoff: int | None = None) -> None:
        """modalfilter (energydissipation.f90:720-959) on the state; cutoff None
        keeps every mode, as the reference does (the identity up to round-off)."""
        self._check(self._L.dsem_modal_filter(self._h, self.ncg if cutoff is None else int(cutoff)))

    def les_update(self):
        """(nu_t at the Gauss points [K, np], nu_t on the staggered grids [K, 3, ns])
        from the current lagged staggered states."""
        g = np.zeros((self.K, self.np))
        st = np.zeros((self.K, 3, self.ns))
        self._check(self._L.dsem_les_update(self._h, _dp(g), _dp(st)))
        return g, st

    def set_les_viscosity(self, nut: np.ndarray) -> None:
        nut = np.ascontiguousarray(nut, dtype=np.float64)
        if nut.size != self.K * 3 * self.ns:
            raise ValueError("nu_t must be [K][3][nc*ncg^2]")
        self._check(self._L.dsem_set_les_viscosity(self._h, _dp(nut)))

    # ------------------------------------------------------------ diagnostics
    def keep_lagged_states(self, enable: bool = True) -> None:
        """Keep the last stage's staggered states after every step (needed by
        the enstrophy diagnostic when entropy viscosity is off)."""
        self._check(self._L.dsem_keep_lagged_states(self._h, int(enable)))

    def kinetic_energy(self) -> float:
        """turbk of the TGV statistics: 1/2 <u^2+v^2+w^2> over the (2 pi)^3 box."""
        ek = C.c_double(0.0)
        self._check(self._L.dsem_reduce_ke(self._h, C.byref(ek), None))
        return ek.value

    def diagnostics(self) -> tuple[float, float]:
        """(Ek, eps): kinetic energy and the enstrophy dissipation of
        energydissipation.f90 (needs lagged staggered states)."""
        ek, eps = C.c_double(0.0), C.c_double(0.0)
        self._check(self._L.dsem_reduce_ke(self._h, C.byref(ek), C.byref(eps)))
        return ek.value, eps.value

    # ------------------------------------------------------------------ path
    def rhs(self, t: float = 0.0) -> np.ndarray:
        out = np.zeros((self.K, 5, self.np))
        self._check(self._L.dsem_rhs(self._h, t, _dp(out)))
        return out

    # ---------------------------------------------------------------- adjoint
    def rhs_vjp(self, lam: np.ndarray) -> tuple[np.ndarray, float]:
        """((dR/dJQ)^T lam, <lam, dR/dnu>) at the current state, nu = 1/Re."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        assert lam.shape == (self.K, 5, self.np)
        out = np.zeros_like(lam)
        nb = C.c_double(0.0)
        self._check(self._L.dsem_rhs_vjp(self._h, _dp(lam), _dp(out), C.byref(nb)))
        return out, nb.value

    def rk_step_vjp(self, dt: float, qbar: np.ndarray) -> tuple[np.ndarray, float]:
        """Adjoint of one step from the current state (left unchanged):
        returns (adjoint of the start state, nu contribution)."""
        q = np.array(qbar, dtype=np.float64, copy=True, order="C")
        assert q.shape == (self.K, 5, self.np)
        nb = C.c_double(0.0)
        self._check(self._L.dsem_rk_step_vjp(self._h, dt, _dp(q), C.byref(nb)))
        return q, nb.value

    def rhs_vjp_device(self, jq_ptr: int, lam_ptr: int, out_ptr: int) -> float:
        nb = C.c_double(0.0)
        self._check(self._L.dsem_rhs_vjp_device(self._h, jq_ptr, lam_ptr, out_ptr, C.byref(nb)))
        return nb.value

    def rk_step_vjp_device(self, dt: float, qbar_ptr: int) -> float:
        nb = C.c_double(0.0)
        self._check(self._L.dsem_rk_step_vjp_device(self._h, dt, qbar_ptr, C.byref(nb)))
        return nb.value

    def tape_reserve(self, nslots: int) -> None:
        if nslots > getattr(self, "_tape_slots", 0) or self.les != getattr(self, "_tape_les", False):
            self.tape_owner = None  # a larger tape, or one laid out for nu_t, is a new allocation
            self._tape_slots = max(nslots, getattr(self, "_tape_slots", 0))
            self._tape_les = self.les
        self._check(self._L.dsem_tape_reserve(self._h, nslots))

    def rk_step_record(self, dt: float, slot: int) -> None:
        self.tape_owner = None      # whoever recorded the tape's contents before no longer owns them
        self._check(self._L.dsem_rk_step_record(self._h, dt, slot))

    def rk_step_vjp_tape(self, dt: float, slot: int, qbar_ptr: int) -> float:
        nb = C.c_double(0.0)
        self._check(self._L.dsem_rk_step_vjp_tape(self._h, dt, slot, qbar_ptr, C.byref(nb)))
        return nb.value

    # ---------------------------------------------- adjoint with entropy viscosity
    def ev_vjp(self, muhbar: np.ndarray):
        """((d mu_h/d JQ)^T muhbar, (d mu_h/d Slag)^T muhbar) at the current
        state and lagged states (shock = 1); muhbar [K][3][nc ncg^2]."""
        muhbar = np.ascontiguousarray(muhbar, dtype=np.float64)
        assert muhbar.size == self.K * 3 * self.ns
        jb = np.zeros((self.K, 5, self.np))
        sb = np.zeros((self.K, 3, 5, self.ns))
        self._check(self._L.dsem_ev_vjp(self._h, _dp(muhbar), _dp(jb), _dp(sb)))
        return jb, sb

    def rk_step_vjp_tape_lagged(self, dt: float, slot: int, qbar_ptr: int, slagbar_ptr: int) -> float:
        nb = C.c_double(0.0)
        self._check(self._L.dsem_rk_step_vjp_tape_lagged(self._h, dt, slot, qbar_ptr, slagbar_ptr, C.byref(nb)))
        return nb.value

    def get_lagged_states_device(self, ptr: int) -> bool:
        """Copy the lagged states to device memory at ptr; False (nothing
        copied) when the context holds none (a fresh start)."""
        valid = C.c_int32(0)
        self._check(self._L.dsem_get_lagged_states_device(self._h, ptr, C.byref(valid)))
        return bool(valid.value)

    def set_lagged_states_device(self, ptr: int | None) -> None:
        """Lagged states from device memory; None makes the next step a fresh
        start (mu_h = 0)."""
        self._check(self._L.dsem_set_lagged_states_device(self._h, ptr))

    def set_state_device(self, ptr: int) -> None:
        self._check(self._L.dsem_set_state_device(self._h, ptr))

    def get_state_device(self, ptr: int) -> None:
        self._check(self._L.dsem_get_state_device(self._h, ptr))

    def quadrature_weights(self) -> np.ndarray:
        w = np.zeros(self.np)
        self._check(self._L.dsem_quadrature_weights(self._h, _dp(w)))
        return w

    def tstep(self, time: float, dt: float) -> float:
        r = C.c_double(0.0)
        self._check(self._L.dsem_rk_step(self._h, time, dt, C.byref(r)))
        return r.value

    def steps(self, n: int, dt: float, time: float = 0.0, check: bool = True) -> float | None:
        if check:
            r = C.c_double(0.0)
            self._check(self._L.dsem_rk_steps(self._h, time, dt, n, C.byref(r)))
            return r.value
        self._check(self._L.dsem_rk_steps(self._h, time, dt, n, None))
        return None

    def dtcalc(self, cfl: float = 1.0) -> float:
        d = C.c_double(0.0)
        self._check(self._L.dsem_dtcalc(self._h, C.byref(d)))
        return d.value * cfl

    def synchronize(self):
        self._check(self._L.dsem_synchronize(self._h))

    def stream_handle(self) -> int:
        s = C.c_void_p()
        self._check(self._L.dsem_stream(self._h, C.byref(s)))
        return s.value or 0

    # ------------------------------------------------------------------ timing
    def timing(self, enable: bool):
        self._check(self._L.dsem_timing_enable(self._h, int(enable)))
        if enable:
            self._check(self._L.dsem_timing_reset(self._h))

    def timing_report(self) -> dict:
        out = {}
        for k in range(self._L.dsem_timing_count()):
            name = C.c_char_p()
            ms = C.c_double()
            n = C.c_int64()
            self._check(self._L.dsem_timing_query(self._h, k, C.byref(name), C.byref(ms), C.byref(n)))
            out[name.value.decode()] = (ms.value, n.value)
        return out
