"""Tracer advection on the device: the low-storage SSP Runge-Kutta tables, the
density-weighted tracer RK update, and a stepper that runs the reference's
tracer update loop for the tracer-only advection test.

* ``rk_tables``          - timeint_rk_butcher_tab_get for the low-storage SSP
                           schemes (CMN/scale_timeint_rk_butcher_tab.F90:126-324)
                           with ShuOsher2Butcher (:331-363)
* ``trcvar_rk_advance``  - timeint_rk%Advance_trcvar, low-storage form
                           (CMN/scale_timeint_rk.F90:1390-1498)
* ``tracer_modalfilter_apply`` - atm_dyn_dgm_tracer_modalfilter_apply
                           (FDS/scale_atm_dyn_dgm_modalfilter.F90:232-270)
* ``cal_alphdens_advtest`` - atm_dyn_dgm_trcadvect3d_heve_cal_alphdens_advtest
                           (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:407-445)
* ``TracerAdvection``    - AtmDynDGMDriver_trcadv3d_update with
                           ONLY_TRACERADV_FLAG (FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:312-560):
                           fixed mass fluxes and density, per RK stage the FCT
                           coefficient, the tendency and the tracer update, TMAR
                           after the last stage.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LIB, check
from .dyncore import (DevField, _p, halo_exchange, trcadvect_cal_tend, trcadvect_calc_fct_coef,
                      trcadvect_tmar)
from .mesh import DeviceMesh

# Shu-Osher coefficients sigma(i, j), gamma(i, j) (1-based, i = 2..nstage+1)
# of the reference's low-storage SSP schemes.
_SSP = {
    "ERK_SSP_2s2o": (2, {(2, 1): 1.0, (3, 1): 0.5, (3, 2): 0.5},
                     {(2, 1): 1.0, (3, 2): 0.5}),
    "ERK_SSP_3s3o": (3, {(2, 1): 1.0, (3, 1): 3.0 / 4.0, (3, 2): 1.0 / 4.0, (4, 1): 1.0 / 3.0, (4, 2): 0.0,
                         (4, 3): 2.0 / 3.0},
                     {(2, 1): 1.0, (3, 2): 1.0 / 4.0, (4, 3): 2.0 / 3.0}),
    "ERK_SSP_4s3o": (4, {(2, 1): 1.0, (3, 1): 0.0, (3, 2): 1.0, (4, 1): 2.0 / 3.0, (4, 2): 0.0, (4, 3): 1.0 / 3.0,
                         (5, 4): 1.0},
                     {(2, 1): 0.5, (3, 2): 0.5, (4, 3): 1.0 / 6.0, (5, 4): 0.5}),
    "ERK_SSP_5s3o_2N2*": (5, {(2, 1): 1.0, (3, 2): 1.0, (4, 1): 0.682342861037239, (4, 3): 0.317657138962761,
                              (5, 4): 1.0, (6, 1): 0.045230974482400, (6, 5): 0.954769025517600},
                          {(2, 1): 0.465388589249323, (3, 2): 0.465388589249323, (4, 3): 0.124745797313998,
                           (5, 4): 0.465388589249323, (6, 5): 0.154263303748666}),
}


def _ssp_10s4o():
    sig, gam = {}, {}
    for n in range(1, 5):
        sig[(n + 1, n)], gam[(n + 1, n)] = 1.0, 1.0 / 6.0
    sig[(6, 1)], sig[(6, 5)] = 3.0 / 5.0, 2.0 / 5.0
    gam[(6, 5)] = 1.0 / 15.0
    for n in range(6, 10):
        sig[(n + 1, n)], gam[(n + 1, n)] = 1.0, 1.0 / 6.0
    sig[(11, 1)], sig[(11, 5)], sig[(11, 10)] = 0.2 * 0.2, 1.8 * 0.2, 3.0 * 0.2
    gam[(11, 5)], gam[(11, 10)] = 1.8 / 30.0, 3.0 / 30.0
    return 10, sig, gam


_SSP["ERK_SSP_10s4o_2N"] = _ssp_10s4o()


class RkTables:
    """nstage, sig/gam (nstage+1, nstage) and the Butcher a, b, c of a low-storage scheme."""

    def __init__(self, name: str):
        if name not in _SSP:
            raise ValueError(f"not a low-storage SSP scheme: {name} (have {sorted(_SSP)})")
        N, s, g = _SSP[name]
        self.name, self.nstage = name, N
        self.sig = np.zeros((N + 1, N))
        self.gam = np.zeros((N + 1, N))
        for (i, j), v in s.items():
            self.sig[i - 1, j - 1] = v
        for (i, j), v in g.items():
            self.gam[i - 1, j - 1] = v
        # A = (I - SIG)^-1 GAM, SIG strictly lower triangular: forward substitution
        a = np.zeros((N, N))
        for i in range(N):
            for j in range(N):
                acc = self.gam[i, j]
                for k in range(i):
                    acc = acc + self.sig[i, k] * a[k, j]
                a[i, j] = acc
        self.a = a
        self.b = np.array([self.gam[N, j] + sum(self.sig[N, k] * a[k, j] for k in range(N)) for j in range(N)])
        self.c = np.array([sum(a[i, : i + 1]) for i in range(N)])
        self.sig_f = np.asfortranarray(self.sig).ravel(order="F").copy()
        self.gam_f = np.asfortranarray(self.gam).ravel(order="F").copy()


def rk_tables(name: str) -> RkTables:
    return RkTables(name)


def _dp(a):
    return a.ctypes.data_as(C.c_void_p)


def trcvar_rk_advance(dm: DeviceMesh, nowstage: int, tab: RkTables, dt: float, q, var0, varTmp, tend, DDENS,
                      DDENS0, DENS_hyd):
    check(LIB.fedg_trcvar_rk_advance(dm.handle, int(nowstage), tab.nstage, _dp(tab.sig_f), _dp(tab.gam_f),
                                     _dp(tab.c), float(dt), *[_p(a) for a in (q, var0, varTmp, tend, DDENS, DDENS0,
                                                                              DENS_hyd)]),
          "fedg_trcvar_rk_advance")


def cal_alphdens_advtest(dm: DeviceMesh, alph_dens_M, alph_dens_P, DDENS, MOMX, MOMY, MOMZ, DENS_hyd):
    check(LIB.fedg_trcadvect_cal_alphdens_advtest(dm.handle, *[_p(a) for a in (alph_dens_M, alph_dens_P, DDENS, MOMX,
                                                                              MOMY, MOMZ, DENS_hyd)]),
          "fedg_trcadvect_cal_alphdens_advtest")


def tracer_modalfilter_apply(dm: DeviceMesh, QTRC, DENS_hyd, DDENS, MF_h, MF_v):
    """Density-weighted modal filter of a tracer in place; MF_h, MF_v: (n1,n1) host
    matrices in the reference's column-major order (fedg.modalfilter_matrix)."""
    h = np.asarray(MF_h, dtype=np.float64).ravel(order="F").copy()
    v = np.asarray(MF_v, dtype=np.float64).ravel(order="F").copy()
    check(LIB.fedg_tracer_modalfilter_apply(dm.handle, _p(QTRC), _p(DENS_hyd), _p(DDENS), _dp(h), _dp(v)),
          "fedg_tracer_modalfilter_apply")


class TracerAdvection:
    """Tracer-only advection by fixed mass fluxes (ONLY_TRACERADV_FLAG).

    MOMX/MOMY/MOMZ, DDENS, DENS_hyd: host (Np,NeA) arrays of the flow, copied
    to the device once, with their halo exchanged (on closed boundaries the
    caller fills the halo with the boundary condition, as the reference's
    ApplyBC_PROGVARS_lc does).  J: lcmesh%J (Np,Ne).  step(Q) advances the
    device tracer field Q by one step of dt."""

    def __init__(self, dm: DeviceMesh, MOMX, MOMY, MOMZ, DDENS, DENS_hyd, J, dt, scheme="ERK_SSP_3s3o",
                 disable_limiter=False, modal_filter=None):
        mesh = dm.mesh
        self.dm, self.dt, self.disable = dm, float(dt), bool(disable_limiter)
        self.mf = modal_filter  # (MF_h, MF_v) of the tracer filter (MODALFILTER_FLAG), or None
        self.tab = rk_tables(scheme)
        nA, nf = mesh.Np * mesh.NeA, mesh.NfpTot * mesh.Ne
        f = lambda a: DevField(a.size, a)
        self.MX, self.MY, self.MZ = f(MOMX), f(MOMY), f(MOMZ)
        self.DD, self.Dh, self.J = f(DDENS), f(DENS_hyd), f(J)
        halo_exchange(dm, [self.MX, self.MY, self.MZ, self.DD, self.Dh])
        self.aM, self.aP = DevField(nf), DevField(nf)
        cal_alphdens_advtest(dm, self.aM, self.aP, self.DD, self.MX, self.MY, self.MZ, self.Dh)
        self.tp = DevField(nA)
        self.R, self.tend, self.var0, self.varTmp = DevField(nA), DevField(nA), DevField(nA), DevField(nA)

    def step(self, Q, nsteps=1):
        dm, t, dt = self.dm, self.tab, self.dt
        for _ in range(nsteps):
            for s in range(1, t.nstage + 1):
                halo_exchange(dm, [Q])
                dttmp = dt * t.gam[s, s - 1] / t.sig[s, s - 1]
                trcadvect_calc_fct_coef(dm, self.R, Q, self.MX, self.MY, self.MZ, self.tp, self.aM, self.aP, self.Dh,
                                        self.DD, self.DD, t.c[s - 1], dttmp, self.J, self.disable)
                halo_exchange(dm, [self.R])
                trcadvect_cal_tend(dm, self.tend, Q, self.MX, self.MY, self.MZ, self.aM, self.aP, self.R, self.tp,
                                   self.J)
                trcvar_rk_advance(dm, s, t, dt, Q, self.var0, self.varTmp, self.tend, self.DD, self.DD, self.Dh)
                if s == t.nstage and self.mf is not None:
                    tracer_modalfilter_apply(dm, Q, self.Dh, self.DD, *self.mf)
                if s == t.nstage and not self.disable:
                    trcadvect_tmar(dm, Q, self.Dh, self.DD, self.J)
            # QTRC = (DENS_hyd + DDENS_TRC) / (DENS_hyd + DDENS) * QTRC_tmp: DDENS_TRC = DDENS
            # here, so the factor is exactly one

    def close(self):
        for a in (self.MX, self.MY, self.MZ, self.DD, self.Dh, self.J, self.aM, self.aP, self.tp, self.R, self.tend,
                  self.var0, self.varTmp):
            a.free()
