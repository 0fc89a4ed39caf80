"""fedg: FElib regional nonhydrostatic DG dynamical-core step on AMD MI355X.

Python host mirror of the reference's dynamical-core interfaces over the C ABI
of libfedg.so (include/fedg.h).  Import fails when the native library is not
built: there is no CPU fallback.
"""
from . import cases, constants, history, restart, tracer  # noqa: F401
from ._lib import EXPORTED, LIB, LIB_PATH, FedgError, check  # noqa: F401
from .dyncore import (VARS, DevField, DevField32, HeveStepper, HeviStepper, Mlp, TimeIntRK, add_phy_tend,  # noqa: F401
                      apply_bc, atmos_calc_diagvar,
                      calc_phyd_hgrad, calc_pressure, device_name, halo_exchange, heve_add_bnd_contrib, heve_cal_tend,
                      heve_cal_tend_cco, hevi_cal_tend, trcadvect_cal_tend,
                      trcadvect_calc_fct_coef, trcadvect_save_massflux, trcadvect_tmar, tracer_specific_heat,
                      trc_negative_fixer,
                      hevi_cal_vi, hevi_cal_vi_guess, modalfilter_apply, modalfilter_matrix, stage_probe)
from .mesh import BND_NONE, BND_NOSLIP, BND_SLIP, BoxMesh, DeviceMesh, RefElem  # noqa: F401
