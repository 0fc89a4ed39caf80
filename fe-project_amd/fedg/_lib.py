"""ctypes binding of libfedg.so (the C ABI declared in include/fedg.h).

The library is built in-tree by ``make -C fe-project_amd`` (``__graft_entry__.build``).
Loading fails loudly when it is missing: there is no CPU fallback for any entry
point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfedg.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
vp = C.c_void_p


class BoxDesc(C.Structure):
    _fields_ = [
        ("porder", C.c_int32),
        ("NeX", C.c_int32), ("NeY", C.c_int32), ("NeZ", C.c_int32),
        ("xmin", C.c_double), ("xmax", C.c_double),
        ("ymin", C.c_double), ("ymax", C.c_double),
        ("zmin", C.c_double), ("zmax", C.c_double),
        ("periodic_x", C.c_int32), ("periodic_y", C.c_int32), ("periodic_z", C.c_int32),
    ]


class BoxSizes(C.Structure):
    _fields_ = [(n, C.c_int32) for n in
                ("n1", "Np", "Nfp_h", "Nfp_v", "NfpTot", "Ne", "NeA", "Ne2D", "Ne2DA", "nB")]


class BoxArrays(C.Structure):
    _fields_ = [
        ("pos_en", dp), ("J", dp), ("Escale", dp), ("sJ", dp), ("Fscale", dp), ("normal_fn", dp),
        ("vmapM", ip), ("vmapP", ip), ("vmapB", ip), ("halo_src", ip), ("halo_face", ip),
        ("EMap3Dto2D", ip), ("Gsqrt", dp), ("GI3", dp), ("GsqrtH", dp),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("porder", C.c_int32),
        ("Ne", C.c_int32), ("NeA", C.c_int32), ("Ne2D", C.c_int32), ("Ne2DA", C.c_int32), ("nB", C.c_int32),
        ("Escale", dp), ("Gsqrt", dp), ("GI3", dp), ("GsqrtH", dp),
        ("G11", dp), ("G12", dp), ("G22", dp),
        ("normal_fn", dp), ("Fscale", dp),
        ("vmapM", ip), ("vmapP", ip), ("EMap3Dto2D", ip), ("halo_src", ip), ("velbc", ip),
        ("D1D", dp), ("Lift_mat", dp), ("IntrpMat_VPOrdM1_tr", dp),
        ("NeX", C.c_int32), ("NeY", C.c_int32), ("NeZ", C.c_int32),
        ("GRAV", C.c_double), ("Rdry", C.c_double), ("CPdry", C.c_double), ("CVdry", C.c_double),
        ("PRES00", C.c_double),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"fedg: native library not built ({LIB_PATH}); run `make -C fe-project_amd` "
            "(or __graft_entry__.build()). There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "fedg_last_error": (C.c_char_p, []),
        "fedg_version": (C.c_char_p, []),
        "fedg_dev_alloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
        "fedg_dev_free": (C.c_int, [vp]),
        "fedg_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
        "fedg_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
        "fedg_memset_dev": (C.c_int, [vp, C.c_int, C.c_size_t]),
        "fedg_device_sync": (C.c_int, []),
        "fedg_device_name": (C.c_int, [C.c_char_p, C.c_int]),
        "fedg_stage_probe": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "fedg_refelem_tables": (C.c_int, [C.c_int, dp, dp, dp, dp, dp, dp, dp, dp, ip, ip]),
        "fedg_box_get_sizes": (C.c_int, [C.POINTER(BoxDesc), C.POINTER(BoxSizes)]),
        "fedg_box_build": (C.c_int, [C.POINTER(BoxDesc), C.POINTER(BoxArrays)]),
        "fedg_ctx_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(vp)]),
        "fedg_ctx_destroy": (C.c_int, [vp]),
        "fedg_ctx_sizes": (C.c_int, [vp, C.POINTER(BoxSizes)]),
        "fedg_ctx_is_flat": (C.c_int, [vp]),
        "fedg_drhot2pres": (C.c_int, [vp] + [vp] * 8),
        "fedg_halo_exchange": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "fedg_apply_bc_prgvars": (C.c_int, [vp] + [vp] * 7),
        "fedg_heve_cal_tend": (C.c_int, [vp] + [vp] * 21),
        "fedg_heve_cal_tend_cco": (C.c_int, [vp] + [vp] * 21),
        "fedg_heve_add_bnd_contrib": (C.c_int, [vp] + [vp] * 17),
        "fedg_trcadvect_cal_tend": (C.c_int, [vp] + [vp] * 10),
        "fedg_trcadvect_calc_fct_coef": (C.c_int, [vp] + [vp] * 11 + [C.c_double, C.c_double, vp, C.c_int]),
        "fedg_trcadvect_tmar": (C.c_int, [vp] * 5),
        "fedg_trcadvect_cal_alphdens_advtest": (C.c_int, [vp] * 8),
        "fedg_tracer_modalfilter_apply": (C.c_int, [vp] * 6),
        "fedg_trcvar_rk_advance": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, C.c_double] + [vp] * 7),
        "fedg_trcadvect_save_massflux": (C.c_int, [vp] + [vp] * 15 + [C.c_int, C.c_double, C.c_double, C.c_int]),
        "fedg_calc_phyd_hgrad": (C.c_int, [vp] + [vp] * 4),
        "fedg_hevi_cal_tend": (C.c_int, [vp] + [vp] * 21),
        "fedg_heve_set_precision": (C.c_int, [vp, C.c_int]),
        "fedg_hevi_set_precision": (C.c_int, [vp, C.c_int]),
        "fedg_modalfilter_matrix": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, vp]),
        "fedg_ctx_setup_modalfilter": (C.c_int, [vp, vp, vp]),
        "fedg_modalfilter_apply": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_int]),
        "fedg_heve_set_modalfilter": (C.c_int, [vp, C.c_int]),
        "fedg_hevi_set_modalfilter": (C.c_int, [vp, C.c_int]),
        "fedg_hevi_cal_vi_guess": (C.c_int, [vp] + [vp] * 20 + [C.c_double, C.c_double, C.c_int, C.c_double, vp,
                                                               C.POINTER(C.c_double), C.POINTER(vp), vp,
                                                               C.POINTER(C.c_int), C.POINTER(C.c_long)]),
        "fedg_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, C.POINTER(vp)]),
        "fedg_mlp_destroy": (C.c_int, [vp]),
        "fedg_mlp_param_count": (C.c_int, [vp, C.POINTER(C.c_long)]),
        "fedg_mlp_get_params": (C.c_int, [vp, vp, vp]),
        "fedg_mlp_set_params": (C.c_int, [vp, vp]),
        "fedg_mlp_forward": (C.c_int, [vp, vp, vp, C.c_int]),
        "fedg_mlp_loss_grad": (C.c_int, [vp, vp, vp, C.c_int, C.POINTER(C.c_float)]),
        "fedg_mlp_adam_step": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, C.c_float]),
        "fedg_hevi_set_guess": (C.c_int, [vp, vp, C.POINTER(C.c_double)]),
        "fedg_hevi_guess_stats": (C.c_int, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "fedg_hevi_guess_residuals": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "fedg_hevi_harvest": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_uint]),
        "fedg_hevi_harvest_count": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "fedg_hevi_create": (C.c_int, [vp, C.c_char_p, C.c_double] + [vp] * 9 + [C.c_int, C.c_double,
                                                                                 C.POINTER(vp)]),
        "fedg_hevi_destroy": (C.c_int, [vp]),
        "fedg_hevi_set_state": (C.c_int, [vp, vp]),
        "fedg_hevi_get_state": (C.c_int, [vp, vp]),
        "fedg_hevi_run": (C.c_int, [vp, C.c_int]),
        "fedg_hevi_run_timed": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]),
        "fedg_hevi_newton_stats": (C.c_int, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "fedg_hevi_cal_vi": (C.c_int, [vp] + [vp] * 20 + [C.c_double, C.c_double]),
        "fedg_hevi_cal_vi_newton": (C.c_int, [vp] + [vp] * 20 + [C.c_double, C.c_double, C.c_int, C.c_double,
                                                                 C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "fedg_hevi_cal_vi_noalloc": (C.c_int, [vp] + [vp] * 20 + [C.c_double, C.c_double, C.c_int, C.c_double,
                                                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "fedg_tint_create": (C.c_int, [vp, C.c_char_p, C.c_double, C.POINTER(vp)]),
        "fedg_tint_destroy": (C.c_int, [vp]),
        "fedg_tint_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "fedg_tint_tend_ex": (vp, [vp, C.c_int, C.c_int]),
        "fedg_tint_tend_im": (vp, [vp, C.c_int, C.c_int]),
        "fedg_tint_store_var0": (C.c_int, [vp, vp, C.c_int]),
        "fedg_tint_advance": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(C.c_int)]),
        "fedg_tint_store_implicit": (C.c_int, [vp, C.c_int, vp, C.c_int]),
        "fedg_heve_create": (C.c_int, [vp, C.c_char_p, C.c_double] + [vp] * 9 + [C.POINTER(vp)]),
        "fedg_heve_destroy": (C.c_int, [vp]),
        "fedg_heve_is_flat": (C.c_int, [vp]),
        "fedg_heve_set_massflux": (C.c_int, [vp] * 6),
        "fedg_hevi_set_massflux": (C.c_int, [vp] * 6),
        "fedg_add_phy_tend": (C.c_int, [vp] * 15),
        "fedg_box_topography_metrics": (C.c_int, [vp, vp, C.c_double, vp, vp, vp]),
        "fedg_tracer_specific_heat": (C.c_int, [vp, C.c_int, C.POINTER(vp)] + [vp] * 7),
        "fedg_trc_negative_fixer": (C.c_int, [vp, C.c_int, C.POINTER(vp)] + [vp] * 9 +
                                    [C.c_double, C.c_int, C.c_int] + [vp] * 5),
        "fedg_atmos_calc_diagvar": (C.c_int, [vp, C.c_char_p] + [vp] * 6 + [C.c_int, C.POINTER(vp)] + [vp] * 10),
        "fedg_heve_set_phys_tend": (C.c_int, [vp] * 7),
        "fedg_hevi_set_phys_tend": (C.c_int, [vp] * 7),
        "fedg_heve_set_state": (C.c_int, [vp, C.POINTER(vp)]),
        "fedg_heve_get_state": (C.c_int, [vp, C.POINTER(vp)]),
        "fedg_heve_state_dev": (C.c_int, [vp, C.POINTER(vp)]),
        "fedg_heve_run": (C.c_int, [vp, C.c_int]),
        "fedg_heve_run_timed": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "fedg_heve_get_pressure": (C.c_int, [vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


LIB = _load()

EXPORTED = (
    "fedg_last_error fedg_version fedg_dev_alloc fedg_dev_free fedg_memcpy_h2d fedg_memcpy_d2h "
    "fedg_memset_dev fedg_device_sync fedg_device_name fedg_stage_probe fedg_refelem_tables fedg_box_get_sizes "
    "fedg_box_build fedg_ctx_create fedg_ctx_destroy fedg_ctx_sizes fedg_ctx_is_flat fedg_drhot2pres fedg_halo_exchange "
    "fedg_apply_bc_prgvars fedg_heve_cal_tend fedg_heve_cal_tend_cco fedg_heve_add_bnd_contrib fedg_trcadvect_cal_tend fedg_trcadvect_calc_fct_coef fedg_trcadvect_tmar fedg_trcadvect_save_massflux fedg_trcadvect_cal_alphdens_advtest fedg_trcvar_rk_advance fedg_tracer_modalfilter_apply fedg_hevi_cal_tend fedg_hevi_cal_vi fedg_hevi_cal_vi_newton fedg_hevi_cal_vi_noalloc "
    "fedg_modalfilter_matrix fedg_ctx_setup_modalfilter fedg_modalfilter_apply fedg_heve_set_modalfilter "
    "fedg_hevi_set_modalfilter "
    "fedg_hevi_cal_vi_guess fedg_mlp_create fedg_mlp_destroy fedg_mlp_param_count fedg_mlp_get_params fedg_mlp_set_params "
    "fedg_mlp_forward fedg_mlp_loss_grad fedg_mlp_adam_step fedg_hevi_set_guess fedg_hevi_guess_stats "
    "fedg_hevi_harvest fedg_hevi_harvest_count fedg_hevi_guess_residuals "
    "fedg_heve_set_precision fedg_hevi_set_precision fedg_hevi_create fedg_hevi_destroy fedg_hevi_set_state fedg_hevi_get_state fedg_hevi_run "
    "fedg_hevi_run_timed fedg_hevi_newton_stats fedg_calc_phyd_hgrad fedg_tint_create fedg_tint_destroy "
    "fedg_tint_info fedg_tint_tend_ex fedg_tint_tend_im fedg_tint_store_var0 fedg_tint_advance "
    "fedg_tint_store_implicit fedg_heve_create fedg_heve_destroy fedg_heve_is_flat fedg_heve_set_massflux fedg_hevi_set_massflux fedg_heve_set_state "
    "fedg_heve_get_state fedg_heve_state_dev fedg_heve_run fedg_heve_run_timed fedg_heve_get_pressure "
    "fedg_add_phy_tend fedg_heve_set_phys_tend fedg_hevi_set_phys_tend fedg_box_topography_metrics "
    "fedg_tracer_specific_heat fedg_trc_negative_fixer fedg_atmos_calc_diagvar"
).split()


class FedgError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = LIB.fedg_last_error().decode(errors="replace")
        raise FedgError(f"{what or 'fedg call'} failed (status {rc}): {msg}")
