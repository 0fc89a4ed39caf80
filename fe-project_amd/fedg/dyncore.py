"""Host-side mirror of the reference dynamical-core interfaces on the device.

* ``heve_cal_tend``      - atm_dyn_dgm_nonhydro3d_rhot_heve_cal_tend
                           (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:292-489)
* ``trcadvect_cal_tend`` - atm_dyn_dgm_trcadvect3d_heve_cal_tend
                           (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231)
* ``trcadvect_calc_fct_coef``, ``trcadvect_tmar``, ``trcadvect_save_massflux``
                         - the limiter and mass-flux routines of the same file (:232-404)
* ``calc_pressure``      - atm_dyn_dgm_nonhydro3d_common_DRHOT2PRES (..._common.F90:428-479)
* ``halo_exchange``      - MeshFieldComm_Exchange on one tile (DAT/scale_meshfieldcomm_*.F90)
* ``apply_bc``           - ATMOS_dyn_bnd_applyBC_prgvars_lc (FDS/scale_atm_dyn_dgm_bnd.F90:270-367)
* ``calc_phyd_hgrad``    - atm_dyn_dgm_nonhydro3d_common_calc_phyd_hgrad_lc (..._common.F90:624-777)
* ``modalfilter_apply``  - atm_dyn_dgm_modalfilter_apply (FDS/scale_atm_dyn_dgm_modalfilter.F90:48-130)
* ``TimeIntRK``          - timeint_rk (CMN/scale_timeint_rk.F90) with device buffers
* ``HeveStepper``        - AtmDynDGMDriver_nonhydro3d_update, HEVE branch
                           (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:614-963), fused:
                           one kernel per RK stage, state resident on the GPU.

Fields are device buffers (``DevField``) of the reference's (Np,NeA) layout.
Every call raises ``FedgError`` on a nonzero status, as the reference aborts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LIB, check
from .mesh import DeviceMesh

VARS = ("DDENS", "DRHOT", "MOMZ", "MOMX", "MOMY")  # PRGVAR ids 1..5


class DevField:
    """A device array of float64 (owned)."""

    def __init__(self, n: int, host: np.ndarray | None = None):
        self.n = int(n)
        p = C.c_void_p()
        check(LIB.fedg_dev_alloc(C.byref(p), max(1, self.n) * 8), "fedg_dev_alloc")
        self.ptr = p
        if host is not None:
            self.upload(host)
        else:
            check(LIB.fedg_memset_dev(self.ptr, 0, self.n * 8), "fedg_memset_dev")

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, dtype=np.float64)
        if h.size != self.n:
            raise ValueError(f"size mismatch {h.size} != {self.n}")
        check(LIB.fedg_memcpy_h2d(self.ptr, h.ctypes.data_as(C.c_void_p), self.n * 8), "fedg_memcpy_h2d")

    def download(self) -> np.ndarray:
        h = np.empty(self.n, dtype=np.float64)
        check(LIB.fedg_memcpy_d2h(h.ctypes.data_as(C.c_void_p), self.ptr, self.n * 8), "fedg_memcpy_d2h")
        return h

    def free(self):
        if self.ptr:
            LIB.fedg_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(f):
    return f.ptr if isinstance(f, (DevField, DevField32)) else f


def _massflux_fields(m, fields):
    """Check the five tracer-coupling fields of set_massflux (all or none) and return them.

    MFLX_x/_y/_z need Np*NeA values, alph_dens_M/_P NfpTot*Ne; the stepper writes them at every
    stage, so the caller's DevFields must live until the next set_massflux or close: the steppers
    keep a reference to them."""
    given = [f is not None for f in fields]
    if not any(given):
        return None
    if not all(given):
        raise ValueError("set_massflux takes all five fields or none")
    need = (m.Np * m.NeA,) * 3 + (m.NfpTot * m.Ne,) * 2
    for f, k, name in zip(fields, need, ("MFLX_x", "MFLX_y", "MFLX_z", "alph_dens_M", "alph_dens_P")):
        if not isinstance(f, DevField):
            raise TypeError(f"{name} must be a DevField")
        if f.n < k:
            raise ValueError(f"{name} holds {f.n} values, needs {k}")
    return tuple(fields)


def heve_cal_tend(dm: DeviceMesh, out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DPRES, DENS_hyd, PRES_hyd, PRES_hyd_ref,
                  THERM_hyd, CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy):
    """out = (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt): device fields."""
    args = [_p(a) for a in (*out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DPRES, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd,
                            CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy)]
    check(LIB.fedg_heve_cal_tend(dm.handle, *args), "fedg_heve_cal_tend")


def heve_cal_tend_cco(dm: DeviceMesh, out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DPRES, DENS_hyd, PRES_hyd, PRES_hyd_ref,
                      THERM_hyd, CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy):
    """Interior (volume) part of the HEVE tendency, cal_tend_cco; out as heve_cal_tend."""
    args = [_p(a) for a in (*out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DPRES, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd,
                            CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy)]
    check(LIB.fedg_heve_cal_tend_cco(dm.handle, *args), "fedg_heve_cal_tend_cco")


def heve_add_bnd_contrib(dm: DeviceMesh, dt, DDENS, MOMX, MOMY, MOMZ, DRHOT, DPRES, DENS_hyd, PRES_hyd, THERM_hyd,
                         Rtot=None, CVtot=None, CPtot=None):
    """Face contribution added in place to dt = (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt)
    (add_bnd_contrib_generalvc); DPRES's halo tail is recomputed first."""
    args = [_p(a) for a in (*dt, DDENS, MOMX, MOMY, MOMZ, DRHOT, DPRES, DENS_hyd, PRES_hyd, THERM_hyd, Rtot, CVtot,
                            CPtot)]
    check(LIB.fedg_heve_add_bnd_contrib(dm.handle, *args), "fedg_heve_add_bnd_contrib")


def trcadvect_cal_tend(dm: DeviceMesh, QTRC_dt, QTRC, MOMX, MOMY, MOMZ, alphDENS_M, alphDENS_P, fct_coef, RHOQ_tp,
                       J):
    """Advection tendency of one tracer, atm_dyn_dgm_trcadvect3d_heve_cal_tend
    (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231); alphDENS_M/_P are
    (NfpTot,Ne), J is (Np,Ne), the rest (Np,NeA) device fields."""
    args = [_p(a) for a in (QTRC_dt, QTRC, MOMX, MOMY, MOMZ, alphDENS_M, alphDENS_P, fct_coef, RHOQ_tp, J)]
    check(LIB.fedg_trcadvect_cal_tend(dm.handle, *args), "fedg_trcadvect_cal_tend")


def trcadvect_calc_fct_coef(dm: DeviceMesh, fct_coef, QTRC, MOMX, MOMY, MOMZ, RHOQ_tp, alphDENS_M, alphDENS_P,
                            DENS_hyd, DDENS, DDENS0, rk_c_ssm1, dt, J, disable_limiter=False):
    """FCT coefficient of each element, atm_dyn_dgm_trcadvect3d_heve_calc_fct_coef
    (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:232-305); written on the interior elements."""
    args = [_p(a) for a in (fct_coef, QTRC, MOMX, MOMY, MOMZ, RHOQ_tp, alphDENS_M, alphDENS_P, DENS_hyd, DDENS,
                            DDENS0)]
    check(LIB.fedg_trcadvect_calc_fct_coef(dm.handle, *args, float(rk_c_ssm1), float(dt), _p(J),
                                           int(bool(disable_limiter))), "fedg_trcadvect_calc_fct_coef")


def trcadvect_tmar(dm: DeviceMesh, QTRC, DENS_hyd, DDENS, J):
    """Truncation and mass-aware rescaling in place, atm_dyn_dgm_trcadvect3d_TMAR (..._trcadvect3d_heve.F90:308-337)."""
    check(LIB.fedg_trcadvect_tmar(dm.handle, _p(QTRC), _p(DENS_hyd), _p(DDENS), _p(J)), "fedg_trcadvect_tmar")


def _thermo(thermo, nq):
    th = [np.ascontiguousarray(t, dtype=np.float64) for t in thermo]
    if len(th) != 4 or any(t.size != nq for t in th):
        raise ValueError("thermo = (TRACER_MASS, TRACER_R, TRACER_CV, TRACER_CP), one value per tracer each")
    return th


def tracer_specific_heat(dm: DeviceMesh, Q, thermo, Rtot, CVtot, CPtot):
    """Rtot, CVtot, CPtot of the mixture of the tracers Q (device fields), as update_pressure_specific_heat
    (FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:653-716); thermo = (TRACER_MASS, _R, _CV, _CP)."""
    th = _thermo(thermo, len(Q))
    arr = (C.c_void_p * len(Q))(*[_p(a) for a in Q])
    check(LIB.fedg_tracer_specific_heat(dm.handle, len(Q), arr, *[t.ctypes.data_as(C.c_void_p) for t in th],
                                        _p(Rtot), _p(CVtot), _p(CPtot)), "fedg_tracer_specific_heat")


def trc_negative_fixer(dm: DeviceMesh, Q, DDENS, PRES, CVtot, CPtot, Rtot, DENS_hyd, PRES_hyd, J, zlev, thermo,
                       QLA=0, QIA=0, DRHOT=None, dt=0.0):
    """fix_negative_val of the coupled tracer step (atm_phy_mp_dgm_common_negative_fixer), in place."""
    th = _thermo(thermo, len(Q))
    arr = (C.c_void_p * len(Q))(*[_p(a) for a in Q])
    check(LIB.fedg_trc_negative_fixer(dm.handle, len(Q), arr, _p(DDENS), _p(PRES), _p(CVtot), _p(CPtot), _p(Rtot),
                                      _p(DENS_hyd), _p(PRES_hyd), _p(J), _p(zlev), float(dt), int(QLA), int(QIA),
                                      *[t.ctypes.data_as(C.c_void_p) for t in th],
                                      None if DRHOT is None else _p(DRHOT)), "fedg_trc_negative_fixer")


def atmos_calc_diagvar(dm: DeviceMesh, name: str, out, DDENS=None, MOMX=None, MOMY=None, MOMZ=None, PRES=None,
                       DENS_hyd=None, PRES_hyd=None, Rtot=None, CVtot=None, CPtot=None, Q=(), mass=None, cv=None,
                       engi0=None, zlev=None, pos_z=None):
    """A diagnostic variable by name into the device field out (Np*Ne values), as
    AtmosVars_CalcDiagvar (model/atm_nonhydro3d/src/atmos/mod_atmos_vars.F90:926-934 ->
    mod_atmos_vars_container.F90:378-452, 1100-1366): DENS, U, V, W, Umet, Vmet, PRES_diff, T,
    T_diff, PT, PT_diff, QDRY, ENGK, ENGP (zlev), ENGI, ENGT (pos_z, the computational z).
    Q: tracer device fields; mass / cv / engi0: TRACER_MASS, TRACER_CV, TRACER_ENGI0 per tracer."""
    nq = len(Q)
    arr = (C.c_void_p * max(nq, 1))(*[_p(a) for a in Q])
    tr = []
    for t in (mass, cv, engi0):
        if t is None:
            tr.append(None)
        else:
            t = np.ascontiguousarray(t, dtype=np.float64)
            if t.size != nq:
                raise ValueError("TRACER_MASS / _CV / _ENGI0 need one value per tracer")
            tr.append(t)
    ptrs = [None if t is None else t.ctypes.data_as(C.c_void_p) for t in tr]
    check(LIB.fedg_atmos_calc_diagvar(dm.handle, name.encode(), _p(out), *[_p(a) for a in (DDENS, MOMX, MOMY, MOMZ,
                                                                                          PRES)],
                                      nq, arr, *ptrs, *[_p(a) for a in (DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, zlev,
                                                                       pos_z)]), "fedg_atmos_calc_diagvar")


def trcadvect_save_massflux(dm: DeviceMesh, MFLX_x, MFLX_y, MFLX_z, alph_dens_M, alph_dens_P, DDENS, MOMX, MOMY, MOMZ,
                            DPRES, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, rkstage, tavg_weight_h, tavg_weight_v,
                            is_hevi=False):
    """Time-averaged mass fluxes and face damping coefficients of one RK stage,
    atm_dyn_dgm_trcadvect3d_save_massflux (..._trcadvect3d_heve.F90:340-404)."""
    args = [_p(a) for a in (MFLX_x, MFLX_y, MFLX_z, alph_dens_M, alph_dens_P, DDENS, MOMX, MOMY, MOMZ, DPRES,
                            DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot)]
    check(LIB.fedg_trcadvect_save_massflux(dm.handle, *args, int(rkstage), float(tavg_weight_h),
                                           float(tavg_weight_v), int(bool(is_hevi))),
          "fedg_trcadvect_save_massflux")


def hevi_cal_tend(dm: DeviceMesh, out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DPRES, DENS_hyd, PRES_hyd, PRES_hyd_ref,
                  THERM_hyd, CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy):
    """Explicit HEVI tendencies; out = (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt): device fields."""
    args = [_p(a) for a in (*out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DPRES, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd,
                            CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy)]
    check(LIB.fedg_hevi_cal_tend(dm.handle, *args), "fedg_hevi_cal_tend")


def hevi_cal_vi(dm: DeviceMesh, out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DENS_hyd, PRES_hyd, DDENS0, MOMX0, MOMY0,
                MOMZ0, DRHOT0, Rtot, CVtot, CPtot, impl_fac, dt, max_it=None, rtol=0.0, noalloc=False):
    """Implicit HEVI tendencies into out = (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt) (device fields).

    max_it None: the reference's single Newton iteration (fedg_hevi_cal_vi).
    Otherwise Newton to rtol (fedg_hevi_cal_vi_newton, or with noalloc the
    variant on the library's fixed workspace, fedg_hevi_cal_vi_noalloc);
    returns (iterations, residual norms)."""
    args = [_p(a) for a in (*out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DENS_hyd, PRES_hyd, DDENS0, MOMX0, MOMY0, MOMZ0,
                            DRHOT0, Rtot, CVtot, CPtot)]
    if max_it is None:
        check(LIB.fedg_hevi_cal_vi(dm.handle, *args, float(impl_fac), float(dt)), "fedg_hevi_cal_vi")
        return 1, None
    resid = (C.c_double * max(int(max_it), 1))()
    it = C.c_int(0)
    name = "fedg_hevi_cal_vi_noalloc" if noalloc else "fedg_hevi_cal_vi_newton"
    check(getattr(LIB, name)(dm.handle, *args, float(impl_fac), float(dt), int(max_it), float(rtol), resid,
                             C.byref(it)), name)
    n = min(it.value + 1, int(max_it))
    return it.value, np.array(resid[:n])


def hevi_cal_vi_guess(dm: DeviceMesh, out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DENS_hyd, PRES_hyd, DDENS0, MOMX0, MOMY0,
                      MOMZ0, DRHOT0, Rtot, CVtot, CPtot, impl_fac, dt, max_it, rtol, mlp, scale, Tex=None):
    """hevi_cal_vi to rtol started from mlp's prediction (fedg_hevi_cal_vi_guess).

    Tex: the previous stage's explicit tendency (5 device fields, VID order) or None.
    Returns (iterations, residual norms, lines that fell back to q^n)."""
    args = [_p(a) for a in (*out, DDENS, MOMX, MOMY, MOMZ, DRHOT, DENS_hyd, PRES_hyd, DDENS0, MOMX0, MOMY0, MOMZ0,
                            DRHOT0, Rtot, CVtot, CPtot)]
    if len(scale) != 15:
        raise ValueError("scale takes 15 values: state, tendency-feature and increment scales")
    sc = (C.c_double * 15)(*[float(x) for x in scale])
    tex = None if Tex is None else (C.c_void_p * 5)(*[_p(a) for a in Tex])
    resid = (C.c_double * max(int(max_it), 1))()
    it, fb = C.c_int(0), C.c_long(0)
    check(LIB.fedg_hevi_cal_vi_guess(dm.handle, *args, float(impl_fac), float(dt), int(max_it), float(rtol),
                                     mlp.handle, sc, tex, resid, C.byref(it), C.byref(fb)), "fedg_hevi_cal_vi_guess")
    n = min(it.value + 1, int(max_it))
    return it.value, np.array(resid[:n]), fb.value


def add_phy_tend(dm: DeviceMesh, out, DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p, PRES, Rtot, CPtot):
    """add_phy_tend in place on out = (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt): device fields."""
    args = [_p(a) for a in (*out, DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p, PRES, Rtot, CPtot)]
    check(LIB.fedg_add_phy_tend(dm.handle, *args), "fedg_add_phy_tend")


def _phys_fields(m, fields):
    """Check the six physics-tendency fields of set_phys_tend (all or none, DevFields of Np*NeA
    values) and return them; the steppers read them at every stage, so they keep a reference."""
    given = [f is not None for f in fields]
    if not any(given):
        return None
    if not all(given):
        raise ValueError("set_phys_tend takes all six fields or none")
    for f, name in zip(fields, ("DENS_tp", "MOMX_tp", "MOMY_tp", "MOMZ_tp", "RHOT_tp", "RHOH_p")):
        if not isinstance(f, DevField):
            raise TypeError(f"{name} must be a DevField")
        if f.n < m.Np * m.NeA:
            raise ValueError(f"{name} holds {f.n} values, needs {m.Np * m.NeA}")
    return tuple(fields)


def calc_pressure(dm: DeviceMesh, PRES, DPRES, DRHOT, PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot):
    check(LIB.fedg_drhot2pres(dm.handle, *[_p(a) for a in (PRES, DPRES, DRHOT, PRES_hyd, THERM_hyd, Rtot, CVtot,
                                                           CPtot)]), "fedg_drhot2pres")


def halo_exchange(dm: DeviceMesh, fields):
    arr = (C.c_void_p * len(fields))(*[_p(f) for f in fields])
    check(LIB.fedg_halo_exchange(dm.handle, len(fields), arr), "fedg_halo_exchange")


def apply_bc(dm: DeviceMesh, DDENS, MOMX, MOMY, MOMZ, THERM, DENS_hyd, PRES_hyd):
    check(LIB.fedg_apply_bc_prgvars(dm.handle, *[_p(a) for a in (DDENS, MOMX, MOMY, MOMZ, THERM, DENS_hyd,
                                                                 PRES_hyd)]), "fedg_apply_bc_prgvars")


def modalfilter_matrix(porder: int, etac: float = 2.0 / 3.0, alpha: float = 36.0, order: int = 16) -> np.ndarray:
    """ModalFilter%FilterMat of a 1D element (ELM/scale_element_modalfilter.F90:60-90), F[i, j] = F(i+1, j+1)."""
    n1 = porder + 1
    F = np.zeros(n1 * n1)
    check(LIB.fedg_modalfilter_matrix(int(porder), float(etac), float(alpha), int(order),
                                      F.ctypes.data_as(C.c_void_p)), "fedg_modalfilter_matrix")
    return F.reshape(n1, n1, order="F")


def modalfilter_apply(dm: DeviceMesh, DDENS, MOMX, MOMY, MOMZ, DRHOT, do_weight_Gsqrt=True):
    """atm_dyn_dgm_modalfilter_apply (FDS/scale_atm_dyn_dgm_modalfilter.F90:48-130) on device fields."""
    check(LIB.fedg_modalfilter_apply(dm.handle, *[_p(a) for a in (DDENS, MOMX, MOMY, MOMZ, DRHOT)],
                                     int(bool(do_weight_Gsqrt))), "fedg_modalfilter_apply")


def calc_phyd_hgrad(dm: DeviceMesh, PRES_hyd: np.ndarray, PRES_hyd_ref: np.ndarray):
    """Hydrostatic pressure gradient (host arrays in/out, halos of the inputs filled)."""
    n = dm.mesh.Np * dm.mesh.NeA
    P, R = DevField(n, PRES_hyd), DevField(n, PRES_hyd_ref)
    Dx, Dy = DevField(n), DevField(n)
    check(LIB.fedg_calc_phyd_hgrad(dm.handle, Dx.ptr, Dy.ptr, P.ptr, R.ptr), "fedg_calc_phyd_hgrad")
    out = Dx.download(), Dy.download()
    for f in (P, R, Dx, Dy):
        f.free()
    return out


class TimeIntRK:
    """timeint_rk with device buffers (var0, varTmp, tend_buf2D_ex/_im)."""

    def __init__(self, dm: DeviceMesh, scheme: str, dt: float):
        self.dm = dm
        h = C.c_void_p()
        check(LIB.fedg_tint_create(dm.handle, scheme.encode(), float(dt), C.byref(h)), "fedg_tint_create")
        self.handle = h
        ns, nb, im = C.c_int(), C.c_int(), C.c_int()
        check(LIB.fedg_tint_info(h, C.byref(ns), C.byref(nb), C.byref(im)), "fedg_tint_info")
        self.nstage, self.tend_buf_size, self.imex = ns.value, nb.value, bool(im.value)

    def tend_ex(self, stage: int, varID: int):
        return LIB.fedg_tint_tend_ex(self.handle, stage, varID)

    def tend_im(self, stage: int, varID: int):
        return LIB.fedg_tint_tend_im(self.handle, stage, varID)

    def store_var0(self, q, varID):
        check(LIB.fedg_tint_store_var0(self.handle, _p(q), varID), "fedg_tint_store_var0")

    def advance(self, stage, qs, varIDs):
        arr = (C.c_void_p * len(qs))(*[_p(q) for q in qs])
        ids = (C.c_int * len(varIDs))(*varIDs)
        check(LIB.fedg_tint_advance(self.handle, stage, len(qs), arr, ids), "fedg_tint_advance")

    def store_implicit(self, stage, q, varID):
        check(LIB.fedg_tint_store_implicit(self.handle, stage, _p(q), varID), "fedg_tint_store_implicit")

    def close(self):
        if self.handle:
            LIB.fedg_tint_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HeveStepper:
    """Fused HEVE RK stepper with the state resident on the GPU."""

    def __init__(self, dm: DeviceMesh, bs, DPhydDx: np.ndarray, DPhydDy: np.ndarray, dt: float,
                 scheme: str = "ERK_4s4o", dry: bool = True):
        self.dm = dm
        m = dm.mesh
        n = m.Np * m.Ne
        self._keep = [np.ascontiguousarray(a[:n]) for a in (bs.DENS_hyd, bs.PRES_hyd, bs.THERM_hyd, DPhydDx, DPhydDy)]
        cor = np.ascontiguousarray(bs.CORIOLIS)
        self._keep.append(cor)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        if dry:
            rt = cv = cp = None
        else:
            rt, cv, cp = (np.ascontiguousarray(a[:n]) for a in (bs.Rtot, bs.CVtot, bs.CPtot))
            self._keep += [rt, cv, cp]
        h = C.c_void_p()
        check(LIB.fedg_heve_create(dm.handle, scheme.encode(), float(dt), *[ptr(a) for a in self._keep[:3]],
                                   ptr(cor), ptr(self._keep[3]), ptr(self._keep[4]),
                                   None if rt is None else ptr(rt), None if cv is None else ptr(cv),
                                   None if cp is None else ptr(cp), C.byref(h)), "fedg_heve_create")
        self.handle = h
        self.n = n

    @property
    def flat(self) -> bool:
        return bool(LIB.fedg_heve_is_flat(self.handle))

    def set_massflux(self, MFLX_x=None, MFLX_y=None, MFLX_z=None, alph_dens_M=None, alph_dens_P=None):
        """Save the time-averaged mass fluxes and face damping coefficients of every stage into these
        device fields (atm_dyn_dgm_trcadvect3d_save_massflux, for the tracer step); no arguments: stop."""
        fl = _massflux_fields(self.dm.mesh, (MFLX_x, MFLX_y, MFLX_z, alph_dens_M, alph_dens_P))
        check(LIB.fedg_heve_set_massflux(self.handle, *([_p(a) for a in fl] if fl else [None] * 5)),
              "fedg_heve_set_massflux")
        self._mflx = fl

    def set_phys_tend(self, DENS_tp=None, MOMX_tp=None, MOMY_tp=None, MOMZ_tp=None, RHOT_tp=None, RHOH_p=None):
        """Physics tendencies added to every stage's tendency (add_phy_tend); no arguments: stop."""
        fl = _phys_fields(self.dm.mesh, (DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p))
        check(LIB.fedg_heve_set_phys_tend(self.handle, *([_p(a) for a in fl] if fl else [None] * 6)),
              "fedg_heve_set_phys_tend")
        self._tp = fl

    def set_precision(self, bits: int):
        """64: the reference's FP64 stage; 32: the FP32 perturbation-form tendency (FP64 state and RK)."""
        check(LIB.fedg_heve_set_precision(self.handle, int(bits)), "fedg_heve_set_precision")

    def set_modalfilter(self, on: bool = True):
        """MODALFILTER_FLAG: filter after every step (DeviceMesh.setup_modalfilter first)."""
        check(LIB.fedg_heve_set_modalfilter(self.handle, int(bool(on))), "fedg_heve_set_modalfilter")

    def set_state(self, q):
        hs = [np.ascontiguousarray(a[: self.n], dtype=np.float64) for a in q]
        arr = (C.c_void_p * 5)(*[h.ctypes.data_as(C.c_void_p) for h in hs])
        check(LIB.fedg_heve_set_state(self.handle, arr), "fedg_heve_set_state")

    def get_state(self):
        hs = [np.empty(self.n) for _ in range(5)]
        arr = (C.c_void_p * 5)(*[h.ctypes.data_as(C.c_void_p) for h in hs])
        check(LIB.fedg_heve_get_state(self.handle, arr), "fedg_heve_get_state")
        return hs

    def get_pressure(self):
        P, D = np.empty(self.n), np.empty(self.n)
        check(LIB.fedg_heve_get_pressure(self.handle, P.ctypes.data_as(C.c_void_p), D.ctypes.data_as(C.c_void_p)),
              "fedg_heve_get_pressure")
        return P, D

    def run(self, nsteps: int):
        check(LIB.fedg_heve_run(self.handle, int(nsteps)), "fedg_heve_run")

    def run_timed(self, nsteps: int):
        """(total ms, mean ms per stage-kernel launch) from HIP events on the stepper's stream."""
        tot, per = C.c_double(), C.c_double()
        check(LIB.fedg_heve_run_timed(self.handle, int(nsteps), C.byref(tot), C.byref(per)), "fedg_heve_run_timed")
        return tot.value, per.value

    def close(self):
        if self.handle:
            LIB.fedg_heve_destroy(self.handle)
            self.handle = None
        self._mflx = None  # the stepper no longer writes the mass-flux fields
        self._tp = None  # nor reads the physics tendencies

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HeviStepper:
    """HEVI IMEX step driver with the state resident on the GPU (fedg_hevi_*).

    max_it / rtol: Newton iterations of the implicit solve per stage (the
    reference runs one, rtol 0)."""

    def __init__(self, dm: DeviceMesh, bs, DPhydDx: np.ndarray, DPhydDy: np.ndarray, dt: float,
                 scheme: str = "IMEX_ARK324", max_it: int = 1, rtol: float = 0.0):
        self.dm = dm
        m = dm.mesh
        self.n = m.Np * m.Ne
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (bs.DENS_hyd, bs.PRES_hyd, bs.THERM_hyd, bs.CORIOLIS, DPhydDx, DPhydDy, bs.Rtot, bs.CVtot,
                       bs.CPtot)]
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        h = C.c_void_p()
        check(LIB.fedg_hevi_create(dm.handle, scheme.encode(), float(dt), *[ptr(a) for a in self._keep],
                                   int(max_it), float(rtol), C.byref(h)), "fedg_hevi_create")
        self.handle = h

    def set_state(self, q):
        hs = [np.ascontiguousarray(a[: self.n], dtype=np.float64) for a in q]
        arr = (C.c_void_p * 5)(*[h.ctypes.data_as(C.c_void_p) for h in hs])
        check(LIB.fedg_hevi_set_state(self.handle, arr), "fedg_hevi_set_state")

    def get_state(self):
        hs = [np.empty(self.n) for _ in range(5)]
        arr = (C.c_void_p * 5)(*[h.ctypes.data_as(C.c_void_p) for h in hs])
        check(LIB.fedg_hevi_get_state(self.handle, arr), "fedg_hevi_get_state")
        return hs

    def run(self, nsteps: int):
        check(LIB.fedg_hevi_run(self.handle, int(nsteps)), "fedg_hevi_run")

    def run_timed(self, nsteps: int):
        """(total ms, mean ms per stage of the implicit solve, of the explicit tendency)."""
        tot, im, ex = C.c_double(), C.c_double(), C.c_double()
        check(LIB.fedg_hevi_run_timed(self.handle, int(nsteps), C.byref(tot), C.byref(im), C.byref(ex)),
              "fedg_hevi_run_timed")
        return tot.value, im.value, ex.value

    def set_modalfilter(self, on: bool = True):
        """MODALFILTER_FLAG: filter after every step (DeviceMesh.setup_modalfilter first)."""
        check(LIB.fedg_hevi_set_modalfilter(self.handle, int(bool(on))), "fedg_hevi_set_modalfilter")

    def set_precision(self, bits: int):
        """64: the reference's FP64 explicit tendency; 32: the explicit tendency in FP32
        perturbation form (flat box, p = 7, dry air; state, IMEX and Newton solve FP64)."""
        check(LIB.fedg_hevi_set_precision(self.handle, int(bits)), "fedg_hevi_set_precision")

    def set_guess(self, mlp, scale):
        """Initial Newton iterate from mlp (None: q^n); scale: 5 state, 5 tendency-feature, 5 increment scales."""
        if len(scale) != 15:
            raise ValueError("scale takes 15 values: state, tendency-feature and increment scales")
        sc = (C.c_double * 15)(*[float(x) for x in scale])
        check(LIB.fedg_hevi_set_guess(self.handle, None if mlp is None else mlp.handle, sc), "fedg_hevi_set_guess")

    def guess_stats(self):
        lines, fb = C.c_long(), C.c_long()
        check(LIB.fedg_hevi_guess_stats(self.handle, C.byref(lines), C.byref(fb)), "fedg_hevi_guess_stats")
        return lines.value, fb.value

    def guess_residuals(self):
        """(||G(q^n)||, ||G(start)||), root of the sums of squares over the solves so far."""
        a, b = C.c_double(), C.c_double()
        check(LIB.fedg_hevi_guess_residuals(self.handle, C.byref(a), C.byref(b)), "fedg_hevi_guess_residuals")
        return float(np.sqrt(a.value)), float(np.sqrt(b.value))

    def harvest(self, X: DevField, Y: DevField, capacity: int, per_stage: int, seed: int = 1):
        """Collect (features, converged targets) rows into the fp32 device buffers X, Y during later steps."""
        check(LIB.fedg_hevi_harvest(self.handle, _p(X), _p(Y), int(capacity), int(per_stage), int(seed)),
              "fedg_hevi_harvest")

    def harvest_count(self):
        c = C.c_int()
        check(LIB.fedg_hevi_harvest_count(self.handle, C.byref(c)), "fedg_hevi_harvest_count")
        return c.value

    def newton_stats(self):
        it, calls = C.c_long(), C.c_long()
        check(LIB.fedg_hevi_newton_stats(self.handle, C.byref(it), C.byref(calls)), "fedg_hevi_newton_stats")
        return it.value, calls.value

    def set_phys_tend(self, DENS_tp=None, MOMX_tp=None, MOMY_tp=None, MOMZ_tp=None, RHOT_tp=None, RHOH_p=None):
        """Physics tendencies added to every stage's explicit tendency (add_phy_tend); no arguments: stop."""
        fl = _phys_fields(self.dm.mesh, (DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p))
        check(LIB.fedg_hevi_set_phys_tend(self.handle, *([_p(a) for a in fl] if fl else [None] * 6)),
              "fedg_hevi_set_phys_tend")
        self._tp = fl

    def set_massflux(self, MFLX_x=None, MFLX_y=None, MFLX_z=None, alph_dens_M=None, alph_dens_P=None):
        """Save the tracer mass fluxes of every stage (save_massflux, HEVI weights); no arguments: stop."""
        fl = _massflux_fields(self.dm.mesh, (MFLX_x, MFLX_y, MFLX_z, alph_dens_M, alph_dens_P))
        check(LIB.fedg_hevi_set_massflux(self.handle, *([_p(a) for a in fl] if fl else [None] * 5)),
              "fedg_hevi_set_massflux")
        self._mflx = fl

    def close(self):
        if self.handle:
            LIB.fedg_hevi_destroy(self.handle)
            self.handle = None
        self._mflx = None  # the stepper no longer writes the mass-flux fields
        self._tp = None  # nor reads the physics tendencies

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """Per-column MLP on the GPU (fedg_mlp_*): BF16 matrix-core products, FP32 accumulation and masters."""

    def __init__(self, n_in, n_hidden, n_hidden_layers, n_out, seed=1):
        h = C.c_void_p()
        check(LIB.fedg_mlp_create(int(n_in), int(n_hidden), int(n_hidden_layers), int(n_out), int(seed), C.byref(h)),
              "fedg_mlp_create")
        self.handle = h
        self.n_in, self.n_out = n_in, n_out
        n = C.c_long()
        check(LIB.fedg_mlp_param_count(h, C.byref(n)), "fedg_mlp_param_count")
        self.nparam = n.value

    def params(self):
        p = np.empty(self.nparam, dtype=np.float32)
        check(LIB.fedg_mlp_get_params(self.handle, p.ctypes.data_as(C.c_void_p), None), "fedg_mlp_get_params")
        return p

    def grads(self):
        g = np.empty(self.nparam, dtype=np.float32)
        check(LIB.fedg_mlp_get_params(self.handle, None, g.ctypes.data_as(C.c_void_p)), "fedg_mlp_get_params")
        return g

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(LIB.fedg_mlp_set_params(self.handle, p.ctypes.data_as(C.c_void_p)), "fedg_mlp_set_params")

    def forward(self, X: "DevField32", Y: "DevField32", batch: int):
        check(LIB.fedg_mlp_forward(self.handle, _p(X), _p(Y), int(batch)), "fedg_mlp_forward")

    def loss_grad(self, X, T, batch: int) -> float:
        loss = C.c_float()
        check(LIB.fedg_mlp_loss_grad(self.handle, _p(X), _p(T), int(batch), C.byref(loss)), "fedg_mlp_loss_grad")
        return loss.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(LIB.fedg_mlp_adam_step(self.handle, float(lr), float(beta1), float(beta2), float(eps)),
              "fedg_mlp_adam_step")

    def close(self):
        if self.handle:
            LIB.fedg_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DevField32:
    """A device float32 array (the MLP's inputs and outputs)."""

    def __init__(self, n: int, host: np.ndarray | None = None):
        self.n = int(n)
        self.ptr = C.c_void_p()
        check(LIB.fedg_dev_alloc(C.byref(self.ptr), 4 * max(self.n, 1)), "fedg_dev_alloc")
        if host is not None:
            self.upload(host)

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, dtype=np.float32).ravel()
        assert h.size <= self.n
        check(LIB.fedg_memcpy_h2d(self.ptr, h.ctypes.data_as(C.c_void_p), 4 * h.size), "fedg_memcpy_h2d")

    def download(self, n=None) -> np.ndarray:
        n = self.n if n is None else int(n)
        h = np.empty(n, dtype=np.float32)
        check(LIB.fedg_memcpy_d2h(h.ctypes.data_as(C.c_void_p), self.ptr, 4 * n), "fedg_memcpy_d2h")
        return h

    @property
    def _as_parameter_(self):
        return self.ptr

    def free(self):
        if self.ptr:
            LIB.fedg_dev_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(LIB.fedg_device_name(buf, 256), "fedg_device_name")
    return buf.value.decode()


def stage_probe(mode: int, NeX: int, NeY: int, NeZ: int, reps: int = 10) -> float:
    """Mean ms of a kernel with the loads and stores of one p=7 RK4 stage and no
    arithmetic (fedg_stage_probe): mode 0 / 1 / 2 = first / middle / last stage."""
    ms = C.c_double()
    check(LIB.fedg_stage_probe(int(mode), int(NeX), int(NeY), int(NeZ), int(reps), C.byref(ms)), "fedg_stage_probe")
    return ms.value
