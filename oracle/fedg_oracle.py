"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  The oracle takes
the same host arrays (mesh, operators, base state) as the product and
restates the reference Fortran on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)


class OracleMeshC(C.Structure):
    _fields_ = [
        ("n1", C.c_int), ("Np", C.c_int), ("Nfp_h", C.c_int), ("Nfp_v", C.c_int), ("NfpTot", C.c_int),
        ("Ne", C.c_int), ("NeA", C.c_int), ("Ne2D", C.c_int), ("nB", C.c_int),
        ("Escale", dp), ("Gsqrt", dp), ("GsqrtH", dp), ("GI3", dp), ("G11", dp), ("G12", dp), ("G22", dp),
        ("normal_fn", dp), ("Fscale", dp),
        ("vmapM", ip), ("vmapP", ip), ("vmapB", ip), ("halo_src", ip), ("velbc", ip),
        ("EMap3Dto2D", ip), ("IndexH2Dto3D", ip), ("IndexH2Dto3D_bnd", ip),
        ("D1D", dp), ("Lift_mat", dp), ("IntrpMat_tr", dp),
        ("GRAV", C.c_double), ("Rdry", C.c_double), ("CPdry", C.c_double), ("CVdry", C.c_double),
        ("PRES00", C.c_double), ("MF_h", dp), ("MF_v", dp),
    ]


class RkTabC(C.Structure):
    _fields_ = [("nstage", C.c_int), ("tend_buf_size", C.c_int), ("imex", C.c_int), ("dt", C.c_double),
                ("a_ex", C.c_double * 16), ("b_ex", C.c_double * 4), ("a_im", C.c_double * 16),
                ("b_im", C.c_double * 4), ("tend_buf_indmap", C.c_int * 4)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built: run `make -C {HERE}`")
    lib = C.CDLL(LIB_PATH)
    M = C.POINTER(OracleMeshC)
    lib.oracle_drhot2pres.argtypes = [M] + [dp] * 8
    lib.oracle_halo_exchange.argtypes = [M, dp]
    lib.oracle_apply_bc.argtypes = [M] + [dp] * 5
    lib.oracle_heve_numflux.argtypes = [M] + [dp] * 10
    lib.oracle_heve_cal_tend.argtypes = [M] + [dp] * 21
    lib.oracle_heve_cal_tend_range.argtypes = [M, C.c_int, C.c_int] + [dp] * 21
    lib.oracle_heve_cal_tend_cco.argtypes = [M] + [dp] * 16
    lib.oracle_heve_add_bnd_contrib.argtypes = [M] + [dp] * 17
    lib.oracle_calc_phyd_hgrad.argtypes = [M] + [dp] * 4
    lib.oracle_trcadvect_cal_tend.argtypes = [M] + [dp] * 11
    lib.oracle_trcadvect_calc_fct_coef.argtypes = [M] + [dp] * 12 + [C.c_double, C.c_double, dp, C.c_int]
    lib.oracle_trcadvect_tmar.argtypes = [M] + [dp] * 5
    lib.oracle_tracer_specific_heat.argtypes = [M, C.c_int, C.POINTER(dp)] + [dp] * 7
    lib.oracle_trc_negative_fixer.argtypes = [M, dp, C.c_int, C.POINTER(dp)] + [dp] * 14
    lib.oracle_atmos_calc_diagvar.argtypes = [M, C.c_char_p] + [dp] * 6 + [C.c_int, C.POINTER(dp)] + [dp] * 9
    lib.oracle_atmos_calc_diagvar.restype = C.c_int
    lib.oracle_trcadvect_cal_alphdens_advtest.argtypes = [M] + [dp] * 7
    lib.oracle_tracer_modalfilter_apply.argtypes = [M] + [dp] * 5
    lib.oracle_heve_set_massflux.argtypes = [dp] * 5
    lib.oracle_add_phy_tend.argtypes = [M, dp * 5, dp * 6, dp, dp, dp]
    lib.oracle_set_phys_tend.argtypes = [dp] * 6
    lib.oracle_trcvar_rk_advance.argtypes = [C.c_long, C.c_int, C.c_int, dp, dp, dp, C.c_double] + [dp] * 7
    lib.oracle_trcadvect_save_massflux.argtypes = [M] + [dp] * 12 + [C.c_int, C.c_double, C.c_double, C.c_int]
    lib.oracle_rk_tab_get.argtypes = [C.c_char_p, C.c_double, C.POINTER(RkTabC)]
    lib.oracle_rk_advance.argtypes = [C.POINTER(RkTabC), C.c_int, C.c_long] + [dp] * 5 + [C.c_long]
    lib.oracle_rk_store_impl.argtypes = [C.POINTER(RkTabC), C.c_int, C.c_long] + [dp] * 4 + [C.c_long]
    lib.oracle_heve_update.argtypes = [M, C.POINTER(RkTabC), C.c_int, dp * 5] + [dp] * 2 + [dp] * 10 + [dp]
    lib.oracle_num_threads.restype = C.c_int
    lib.oracle_set_pow_mode.argtypes = [C.c_int]
    lib.oracle_get_pow_mode.restype = C.c_int
    lib.oracle_pow_array.argtypes = [dp, dp, dp, C.c_long, C.c_int]
    V5 = dp * 5
    lib.oracle_hevi_numflux.argtypes = [M] + [dp] * 10
    lib.oracle_hevi_cal_tend.argtypes = [M] + [dp] * 21
    lib.oracle_hevi_newton.argtypes = [M, V5, V5, V5] + [dp] * 5 + [C.c_double, C.c_double, C.c_int,
                                                                     C.c_double, dp]
    lib.oracle_hevi_newton.restype = C.c_int
    lib.oracle_hevi_newton_r0.argtypes = [M, V5, V5, V5] + [dp] * 5 + [C.c_double, C.c_double, C.c_int,
                                                                        C.c_double, C.c_double, dp]
    lib.oracle_hevi_newton_r0.restype = C.c_int
    lib.oracle_modalfilter_apply.argtypes = [M] + [dp] * 5 + [C.c_int]
    lib.oracle_hevi_cal_vi.argtypes = [M, V5, V5, V5] + [dp] * 5 + [C.c_double, C.c_double, C.c_int,
                                                                     C.c_double, dp]
    lib.oracle_hevi_cal_vi.restype = C.c_int
    lib.oracle_hevi_linear_system.argtypes = [M, V5, V5, V5] + [dp] * 5 + [C.c_double, C.c_double] + [dp] * 8
    lib.oracle_hevi_residual.argtypes = [M, V5, V5, V5] + [dp] * 5 + [C.c_double, C.c_double, V5]
    lib.oracle_hevi_update.argtypes = [M, C.POINTER(RkTabC), C.c_int, V5] + [dp] * 2 + [dp] * 10 + [
        C.c_int, C.c_double, dp]
    return lib


LIB = _load()


def _d(a):
    return None if a is None else a.ctypes.data_as(dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(ip)


class OracleMesh:
    """Oracle view of a fedg BoxMesh (same host arrays)."""

    def __init__(self, mesh, velbc=None, consts=None, Gsqrt=None, GI3=None):
        from fedg import constants  # the oracle takes the constants as inputs too
        c = consts or constants.SCALE
        e = mesh.elem
        self.mesh = mesh
        self.velbc = mesh.velbc() if velbc is None else np.asarray(velbc, dtype=np.int32)
        self.Gsqrt = mesh.Gsqrt if Gsqrt is None else Gsqrt
        self.GI3 = mesh.GI3 if GI3 is None else GI3
        n2d = mesh.Nfp_v * mesh.Ne2D
        self.G11, self.G12, self.G22 = np.ones(n2d), np.zeros(n2d), np.ones(n2d)
        s = OracleMeshC()
        s.n1, s.Np, s.Nfp_h, s.Nfp_v, s.NfpTot = e.n1, e.Np, e.Nfp, e.Nfp, e.NfpTot
        s.Ne, s.NeA, s.Ne2D, s.nB = mesh.Ne, mesh.NeA, mesh.Ne2D, mesh.nB
        s.Escale, s.Gsqrt, s.GsqrtH, s.GI3 = _d(mesh.Escale), _d(self.Gsqrt), _d(mesh.GsqrtH), _d(self.GI3)
        s.G11, s.G12, s.G22 = _d(self.G11), _d(self.G12), _d(self.G22)
        s.normal_fn, s.Fscale = _d(mesh.normal_fn), _d(mesh.Fscale)
        s.vmapM, s.vmapP, s.vmapB, s.halo_src = _i(mesh.vmapM), _i(mesh.vmapP), _i(mesh.vmapB), _i(mesh.halo_src)
        s.velbc, s.EMap3Dto2D = _i(self.velbc), _i(mesh.EMap3Dto2D)
        s.IndexH2Dto3D, s.IndexH2Dto3D_bnd = _i(e.IndexH2Dto3D), _i(e.IndexH2Dto3D_bnd)
        s.D1D, s.Lift_mat, s.IntrpMat_tr = _d(e.D1D), _d(e.Lift_mat), _d(e.IntrpMat_VPOrdM1_tr)
        s.GRAV, s.Rdry, s.CPdry, s.CVdry, s.PRES00 = c.GRAV, c.Rdry, c.CPdry, c.CVdry, c.PRES00
        self.s = s

    @property
    def ref(self):
        return C.byref(self.s)

    def set_modalfilter(self, MF_h=None, MF_v=None):
        """Filter matrices ((n1,n1) Fortran order) applied after every step by the update loops; None: off."""
        if MF_h is None:
            self._mf = None
            self.s.MF_h = self.s.MF_v = None
            return
        self._mf = (np.asfortranarray(MF_h, dtype=np.float64).ravel(order="F"),
                    np.asfortranarray(MF_v, dtype=np.float64).ravel(order="F"))
        self.s.MF_h, self.s.MF_v = _d(self._mf[0]), _d(self._mf[1])

    def modalfilter_apply(self, q, do_weight_Gsqrt=True):
        """Filter q (VID order) in place with the matrices of set_modalfilter."""
        if getattr(self, "_mf", None) is None:
            raise ValueError("set_modalfilter first")
        LIB.oracle_modalfilter_apply(self.ref, _d(q[0]), _d(q[3]), _d(q[4]), _d(q[2]), _d(q[1]),
                                     int(bool(do_weight_Gsqrt)))

    # --- wrappers -----------------------------------------------------------
    def drhot2pres(self, DRHOT, bs):
        m = self.mesh
        P, D = m.field(), m.field()
        LIB.oracle_drhot2pres(self.ref, _d(P), _d(D), _d(DRHOT), _d(bs.PRES_hyd), _d(bs.THERM_hyd), _d(bs.Rtot),
                              _d(bs.CVtot), _d(bs.CPtot))
        return P, D

    def exchange(self, f):
        LIB.oracle_halo_exchange(self.ref, _d(f))

    def apply_bc(self, q):
        LIB.oracle_apply_bc(self.ref, _d(q[0]), _d(q[3]), _d(q[4]), _d(q[2]), _d(q[1]))

    def numflux(self, q, DPRES, bs):
        m = self.mesh
        out = np.zeros(m.NfpTot * 5 * m.Ne)
        LIB.oracle_heve_numflux(self.ref, _d(out), _d(q[0]), _d(q[3]), _d(q[4]), _d(q[2]), _d(q[1]), _d(DPRES),
                                _d(bs.DENS_hyd), _d(bs.PRES_hyd), _d(bs.THERM_hyd))
        return out

    def cal_tend(self, q, DPRES, bs, DPhydDx, DPhydDy, ke_range=None):
        """Tendencies (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt), each (Np,NeA)."""
        m = self.mesh
        out = [m.field() for _ in range(5)]
        args = [_d(a) for a in out] + [_d(q[0]), _d(q[3]), _d(q[4]), _d(q[2]), _d(q[1]), _d(DPRES),
                                       _d(bs.DENS_hyd), _d(bs.PRES_hyd), _d(bs.PRES_hyd_ref), _d(bs.THERM_hyd),
                                       _d(bs.CORIOLIS), _d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot), _d(DPhydDx),
                                       _d(DPhydDy)]
        if ke_range is None:
            LIB.oracle_heve_cal_tend(self.ref, *args)
        else:
            LIB.oracle_heve_cal_tend_range(self.ref, int(ke_range[0]), int(ke_range[1]), *args)
        return out

    def cal_tend_cco(self, q, DPRES, bs, DPhydDx, DPhydDy):
        """Interior tendencies of cal_tend_cco (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt)."""
        m = self.mesh
        out = [m.field() for _ in range(5)]
        LIB.oracle_heve_cal_tend_cco(self.ref, *[_d(a) for a in out], _d(q[0]), _d(q[3]), _d(q[4]), _d(q[2]),
                                     _d(q[1]), _d(DPRES), _d(bs.DENS_hyd), _d(bs.THERM_hyd), _d(bs.CORIOLIS),
                                     _d(DPhydDx), _d(DPhydDy))
        return out

    def add_bnd_contrib(self, dt, q, DPRES, bs, moist=False):
        """add_bnd_contrib_generalvc in place on dt = (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt); writes
        DPRES's halo tail.  moist: Rtot/CVtot/CPtot from bs, else the dry constants."""
        ext = [_d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot)] if moist else [None, None, None]
        LIB.oracle_heve_add_bnd_contrib(self.ref, *[_d(a) for a in dt], _d(q[0]), _d(q[3]), _d(q[4]), _d(q[2]),
                                        _d(q[1]), _d(DPRES), _d(bs.DENS_hyd), _d(bs.PRES_hyd), _d(bs.THERM_hyd),
                                        *ext)

    def trcadvect_cal_tend(self, QTRC, MOMX, MOMY, MOMZ, alphDENS_M, alphDENS_P, fct_coef, RHOQ_tp, J=None):
        """Tracer advection tendency QTRC_dt (Np,NeA); J defaults to the mesh's (Np,Ne) Jacobian."""
        m = self.mesh
        out = m.field()
        J = m.J if J is None else J
        LIB.oracle_trcadvect_cal_tend(self.ref, _d(m.elem.w1d), _d(out), _d(QTRC), _d(MOMX), _d(MOMY), _d(MOMZ),
                                      _d(alphDENS_M), _d(alphDENS_P), _d(fct_coef), _d(RHOQ_tp), _d(J))
        return out

    def trcadvect_calc_fct_coef(self, QTRC, MOMX, MOMY, MOMZ, RHOQ_tp, alphDENS_M, alphDENS_P, DENS_hyd, DDENS,
                                DDENS0, rk_c_ssm1, dt, J=None, disable_limiter=False):
        """FCT coefficient (Np,NeA), interior elements written (the halo left zero)."""
        m = self.mesh
        out = m.field()
        J = m.J if J is None else J
        LIB.oracle_trcadvect_calc_fct_coef(self.ref, _d(m.elem.w1d), _d(out), _d(QTRC), _d(MOMX), _d(MOMY), _d(MOMZ),
                                           _d(RHOQ_tp), _d(alphDENS_M), _d(alphDENS_P), _d(DENS_hyd), _d(DDENS),
                                           _d(DDENS0), float(rk_c_ssm1), float(dt), _d(J), int(bool(disable_limiter)))
        return out

    def trcadvect_tmar(self, QTRC, DENS_hyd, DDENS, J=None):
        """TMAR in place on QTRC."""
        m = self.mesh
        J = m.J if J is None else J
        LIB.oracle_trcadvect_tmar(self.ref, _d(m.elem.w1d), _d(QTRC), _d(DENS_hyd), _d(DDENS), _d(J))

    def tracer_specific_heat(self, Q, thermo, Rtot, CVtot, CPtot):
        """Rtot, CVtot, CPtot of the mixture of the tracers Q; thermo = (TRACER_MASS, _R, _CV, _CP)."""
        Qa = (dp * len(Q))(*[_d(a) for a in Q])
        th = [np.ascontiguousarray(t, dtype=np.float64) for t in thermo]
        LIB.oracle_tracer_specific_heat(self.ref, len(Q), Qa, *[_d(t) for t in th], _d(Rtot), _d(CVtot), _d(CPtot))

    def atmos_calc_diagvar(self, name, DDENS, MOMX, MOMY, MOMZ, PRES, DENS_hyd, PRES_hyd, Rtot, CPtot, Q=(),
                           mass=None, cv=None, engi0=None, zlev=None, pos_z=None):
        """A diagnostic variable by name on the interior nodes (vars_calc_diagnoseVar_lc)."""
        m = self.mesh
        out = np.zeros(m.Np * m.Ne)
        Qa = (dp * max(len(Q), 1))(*[_d(a) for a in Q])
        tr = [None if t is None else np.ascontiguousarray(t, dtype=np.float64) for t in (mass, cv, engi0)]
        rc = LIB.oracle_atmos_calc_diagvar(self.ref, name.encode(), _d(out), _d(DDENS), _d(MOMX), _d(MOMY), _d(MOMZ),
                                           _d(PRES), len(Q), Qa, *[None if t is None else _d(t) for t in tr],
                                           _d(DENS_hyd), _d(PRES_hyd), _d(Rtot), _d(CPtot),
                                           None if zlev is None else _d(zlev), None if pos_z is None else _d(pos_z))
        if rc != 0:
            raise ValueError(f"diagnostic variable {name} is not supported")
        return out

    def trc_negative_fixer(self, Q, DDENS, PRES, CVtot, CPtot, Rtot, DENS_hyd, PRES_hyd, J, zlev, thermo,
                           DRHOT=None):
        """The water species' negative fixer in place (Q, DDENS, PRES, CVtot, CPtot, Rtot, DRHOT)."""
        m = self.mesh
        Qa = (dp * len(Q))(*[_d(a) for a in Q])
        th = [np.ascontiguousarray(t, dtype=np.float64) for t in thermo]
        LIB.oracle_trc_negative_fixer(self.ref, _d(m.elem.w1d), len(Q), Qa, _d(DDENS), _d(PRES), _d(CVtot),
                                      _d(CPtot), _d(Rtot), _d(DENS_hyd), _d(PRES_hyd), _d(J), _d(zlev),
                                      *[_d(t) for t in th], _d(DRHOT) if DRHOT is not None else None)

    def trcadvect_save_massflux(self, MFLX, alph, DDENS, MOMX, MOMY, MOMZ, DPRES, DENS_hyd, PRES_hyd, rkstage,
                                twh, twv, is_hevi=False):
        """save_massflux in place on MFLX = (x, y, z) and alph = (M, P)."""
        LIB.oracle_trcadvect_save_massflux(self.ref, *[_d(a) for a in MFLX], *[_d(a) for a in alph], _d(DDENS),
                                           _d(MOMX), _d(MOMY), _d(MOMZ), _d(DPRES), _d(DENS_hyd), _d(PRES_hyd),
                                           int(rkstage), float(twh), float(twv), int(bool(is_hevi)))

    def trcadvect_cal_alphdens_advtest(self, DDENS, MOMX, MOMY, MOMZ, DENS_hyd):
        m = self.mesh
        aM, aP = np.zeros(m.NfpTot * m.Ne), np.zeros(m.NfpTot * m.Ne)
        LIB.oracle_trcadvect_cal_alphdens_advtest(self.ref, _d(aM), _d(aP), _d(DDENS), _d(MOMX), _d(MOMY), _d(MOMZ),
                                                  _d(DENS_hyd))
        return aM, aP

    def trcvar_rk_advance(self, nowstage, tab, dt, q, var0, varTmp, tend, DDENS, DDENS0, DENS_hyd):
        """Low-storage tracer RK update over the interior nodes; tab: fedg.tracer.RkTables."""
        n = self.mesh.Np * self.mesh.Ne
        LIB.oracle_trcvar_rk_advance(n, int(nowstage), tab.nstage, _d(tab.sig_f), _d(tab.gam_f), _d(tab.c), float(dt),
                                     _d(q), _d(var0), _d(varTmp), _d(tend), _d(DDENS), _d(DDENS0), _d(DENS_hyd))

    def heve_set_massflux(self, MFLX=None, alph=None):
        """heve_update saves the tracer mass fluxes of every stage into MFLX = (x, y, z) and
        alph = (M, P) (the arrays must stay alive); None: stop."""
        if MFLX is None:
            LIB.oracle_heve_set_massflux(None, None, None, None, None)
        else:
            LIB.oracle_heve_set_massflux(*[_d(a) for a in MFLX], *[_d(a) for a in alph])

    def add_phy_tend(self, T, tp, PRES, Rtot, CPtot):
        """add_phy_tend in place on T (5 fields, VID order); tp = (DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp,
        MOMY_tp, RHOH_p)."""
        Ta = (dp * 5)(*[_d(a) for a in T])
        ta = (dp * 6)(*[_d(a) for a in tp])
        LIB.oracle_add_phy_tend(self.ref, Ta, ta, _d(PRES), _d(Rtot), _d(CPtot))

    def set_phys_tend(self, tp=None):
        """heve_update / hevi_update add these physics tendencies (DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp,
        MOMY_tp, RHOH_p) at every stage (the arrays must stay alive); None: stop."""
        if tp is None:
            LIB.oracle_set_phys_tend(None, None, None, None, None, None)
        else:
            LIB.oracle_set_phys_tend(*[_d(a) for a in tp])

    def tracer_modalfilter_apply(self, QTRC, DENS_hyd, DDENS, MF_h, MF_v):
        """Density-weighted tracer filter in place; MF_h, MF_v (n1,n1) numpy matrices."""
        h = np.asarray(MF_h, dtype=np.float64).ravel(order="F").copy()
        v = np.asarray(MF_v, dtype=np.float64).ravel(order="F").copy()
        LIB.oracle_tracer_modalfilter_apply(self.ref, _d(QTRC), _d(DENS_hyd), _d(DDENS), _d(h), _d(v))

    def tracer_advection(self, Q, MOMX, MOMY, MOMZ, DDENS, DENS_hyd, dt, nsteps, scheme="ERK_SSP_3s3o",
                         disable_limiter=False, modal_filter=None):
        """AtmDynDGMDriver_trcadv3d_update with ONLY_TRACERADV_FLAG
        (FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:312-560), nsteps steps of
        the tracer Q (in place) by fixed mass fluxes; the flow fields' halos
        are exchanged here (closed boundaries: the caller's BC)."""
        from fedg.tracer import rk_tables
        m = self.mesh
        t = rk_tables(scheme)
        MX, MY, MZ, DD = (a.copy() for a in (MOMX, MOMY, MOMZ, DDENS))
        Dh = DENS_hyd.copy()
        for a in (MX, MY, MZ, DD, Dh):
            self.exchange(a)
        aM, aP = self.trcadvect_cal_alphdens_advtest(DD, MX, MY, MZ, Dh)
        tp = m.field()
        var0, varTmp = m.field(), m.field()
        for _ in range(nsteps):
            for s in range(1, t.nstage + 1):
                self.exchange(Q)
                dttmp = dt * t.gam[s, s - 1] / t.sig[s, s - 1]
                R = self.trcadvect_calc_fct_coef(Q, MX, MY, MZ, tp, aM, aP, Dh, DD, DD, t.c[s - 1], dttmp,
                                                 disable_limiter=disable_limiter)
                self.exchange(R)
                tend = self.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R, tp)
                self.trcvar_rk_advance(s, t, dt, Q, var0, varTmp, tend, DD, DD, Dh)
                if s == t.nstage and modal_filter is not None:
                    self.tracer_modalfilter_apply(Q, Dh, DD, *modal_filter)
                if s == t.nstage and not disable_limiter:
                    self.trcadvect_tmar(Q, Dh, DD)
        return Q

    # --- HEVI ---------------------------------------------------------------
    def hevi_numflux(self, q, DPRES, bs):
        m = self.mesh
        out = np.zeros(m.NfpTot * 5 * m.Ne)
        LIB.oracle_hevi_numflux(self.ref, _d(out), _d(q[0]), _d(q[3]), _d(q[4]), _d(q[2]), _d(q[1]), _d(DPRES),
                                _d(bs.DENS_hyd), _d(bs.PRES_hyd), _d(bs.THERM_hyd))
        return out

    def hevi_cal_tend(self, q, DPRES, bs, DPhydDx, DPhydDy):
        """Explicit HEVI tendencies (DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt), each (Np,NeA)."""
        m = self.mesh
        out = [m.field() for _ in range(5)]
        LIB.oracle_hevi_cal_tend(self.ref, *[_d(a) for a in out], _d(q[0]), _d(q[3]), _d(q[4]), _d(q[2]),
                                 _d(q[1]), _d(DPRES), _d(bs.DENS_hyd), _d(bs.PRES_hyd), _d(bs.PRES_hyd_ref),
                                 _d(bs.THERM_hyd), _d(bs.CORIOLIS), _d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot),
                                 _d(DPhydDx), _d(DPhydDy))
        return out

    def _v5(self, fields):
        return (dp * 5)(*[_d(a) for a in fields])

    def hevi_cal_vi(self, q, var0, bs, impl_fac, dt, max_it=1, rtol=0.0):
        """Implicit tendencies in VID order (each (Np,NeA)), iterations done, residual history."""
        m = self.mesh
        tend = [m.field() for _ in range(5)]
        resid = np.zeros(max(max_it, 1))
        it = LIB.oracle_hevi_cal_vi(self.ref, self._v5(tend), self._v5(q), self._v5(var0), _d(bs.DENS_hyd),
                                    _d(bs.PRES_hyd), _d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot), float(impl_fac),
                                    float(dt), int(max_it), float(rtol), _d(resid))
        return tend, it, resid[:min(it + 1, max(max_it, 1))]

    def hevi_newton(self, pv, qs, q0, bs, impl_fac, dt, max_it=1, rtol=0.0, r0_ref=0.0):
        """Newton iterations on pv (VID order, updated in place); returns (iterations, residual history).

        r0_ref > 0: the residual norm the tolerance is relative to (else the first iteration's)."""
        resid = np.zeros(max(max_it, 1))
        it = LIB.oracle_hevi_newton_r0(self.ref, self._v5(pv), self._v5(qs), self._v5(q0), _d(bs.DENS_hyd),
                                       _d(bs.PRES_hyd), _d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot), float(impl_fac),
                                       float(dt), int(max_it), float(rtol), float(r0_ref), _d(resid))
        return it, resid[:min(it + 1, max(max_it, 1))]

    def hevi_linear_system(self, pv, qs, q0, bs, impl_fac, dt):
        m = self.mesh
        N = m.elem.n1
        NeZ = m.Ne // m.Ne2D
        nc = NeZ * m.Nfp_v * m.Ne2D
        n3 = 3 * N
        D, L, U, b = np.zeros(nc * n3 * n3), np.zeros(nc * n3 * 3), np.zeros(nc * n3 * 3), np.zeros(nc * n3)
        Duv, Luv, Uuv, buv = np.zeros(nc * N * N), np.zeros(nc * N), np.zeros(nc * N), np.zeros(nc * N * 2)
        LIB.oracle_hevi_linear_system(self.ref, self._v5(pv), self._v5(qs), self._v5(q0), _d(bs.DENS_hyd),
                                      _d(bs.PRES_hyd), _d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot), float(impl_fac),
                                      float(dt), _d(D), _d(L), _d(U), _d(b), _d(Duv), _d(Luv), _d(Uuv), _d(buv))
        sh = (m.Ne2D, m.Nfp_v, NeZ)
        return {"D": D.reshape(sh + (n3, n3)).transpose(0, 1, 2, 4, 3),
                "L": L.reshape(sh + (3, n3)).transpose(0, 1, 2, 4, 3),
                "U": U.reshape(sh + (3, n3)).transpose(0, 1, 2, 4, 3),
                "b": b.reshape(sh + (n3,)),
                "Duv": Duv.reshape(sh + (N, N)).transpose(0, 1, 2, 4, 3),
                "Luv": Luv.reshape(sh + (N,)), "Uuv": Uuv.reshape(sh + (N,)),
                "buv": buv.reshape(sh + (2, N)).transpose(0, 1, 2, 4, 3)}

    def hevi_residual(self, pv, qs, q0, bs, impl_fac, dt):
        m = self.mesh
        G = [np.zeros(m.Np * m.Ne) for _ in range(5)]
        LIB.oracle_hevi_residual(self.ref, self._v5(pv), self._v5(qs), self._v5(q0), _d(bs.DENS_hyd),
                                 _d(bs.PRES_hyd), _d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot), float(impl_fac),
                                 float(dt), self._v5(G))
        return G

    def hevi_update(self, q, bs, DPhydDx, DPhydDy, dt, nsteps, scheme="IMEX_ARK324", max_it=1, rtol=0.0):
        """Advance q (VID order, (Np,NeA) host fields) in place by HEVI IMEX steps; returns (PRES, DPRES)."""
        m = self.mesh
        tab = rk_tab(scheme, dt)
        nA = m.Np * m.NeA
        work = np.zeros(nA * (10 + 10 * tab.tend_buf_size))
        P, D = m.field(), m.field()
        LIB.oracle_hevi_update(self.ref, C.byref(tab), int(nsteps), self._v5(q), _d(P), _d(D), _d(bs.DENS_hyd),
                               _d(bs.PRES_hyd), _d(bs.PRES_hyd_ref), _d(bs.THERM_hyd), _d(bs.CORIOLIS),
                               _d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot), _d(DPhydDx), _d(DPhydDy), int(max_it),
                               float(rtol), _d(work))
        return P, D

    def calc_phyd_hgrad(self, bs):
        m = self.mesh
        Dx, Dy = m.field(), m.field()
        LIB.oracle_calc_phyd_hgrad(self.ref, _d(Dx), _d(Dy), _d(bs.PRES_hyd), _d(bs.PRES_hyd_ref))
        return Dx, Dy

    def heve_update(self, q, bs, DPhydDx, DPhydDy, dt, nsteps, scheme="ERK_4s4o"):
        """Advance q (list of 5 (Np,NeA) host fields, VID order) in place; returns (PRES, DPRES)."""
        m = self.mesh
        tab = rk_tab(scheme, dt)
        nA = m.Np * m.NeA
        work = np.zeros(nA * (5 * (2 + tab.tend_buf_size) + 2))
        P, D = m.field(), m.field()
        qa = (dp * 5)(*[_d(a) for a in q])
        LIB.oracle_heve_update(self.ref, C.byref(tab), int(nsteps), qa, _d(P), _d(D), _d(bs.DENS_hyd),
                               _d(bs.PRES_hyd), _d(bs.PRES_hyd_ref), _d(bs.THERM_hyd), _d(bs.CORIOLIS),
                               _d(bs.Rtot), _d(bs.CVtot), _d(bs.CPtot), _d(DPhydDx), _d(DPhydDy), _d(work))
        return P, D


def rk_tab(name, dt):
    t = RkTabC()
    if LIB.oracle_rk_tab_get(name.encode(), float(dt), C.byref(t)) != 0:
        raise ValueError(f"unknown RK scheme {name}")
    return t


def rk_advance(tab, stage, q, var0, varTmp, tend_ex, tend_im=None, bs=None):
    n = q.size
    LIB.oracle_rk_advance(C.byref(tab), stage, n, _d(q), _d(var0), _d(varTmp), _d(tend_ex), _d(tend_im),
                          n if bs is None else bs)


def rk_store_impl(tab, stage, q, var0, varTmp, tend_im, bs=None):
    n = q.size
    LIB.oracle_rk_store_impl(C.byref(tab), stage, n, _d(q), _d(var0), _d(varTmp), _d(tend_im),
                             n if bs is None else bs)


def num_threads():
    return LIB.oracle_num_threads()


POW_LIBM, POW_PORTABLE = 0, 1


def set_pow_mode(mode: int):
    """0: the C library's pow (reference build semantics); 1: the portable pow the kernels use."""
    LIB.oracle_set_pow_mode(int(mode))


def get_pow_mode() -> int:
    return LIB.oracle_get_pow_mode()


class pow_mode:
    """Context manager selecting the oracle's pow."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.prev = get_pow_mode()
        set_pow_mode(self.mode)

    def __exit__(self, *a):
        set_pow_mode(self.prev)


def pow_array(x, y, mode):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
    out = np.empty_like(x)
    LIB.oracle_pow_array(_d(x), _d(y), _d(out), x.size, int(mode))
    return out
