!-------------------------------------------------------------------------------
!> ISO_C_BINDING interfaces of libfedg.so (include/fedg.h) for a Fortran host:
!! every entry point of the C ABI, in the order of the header.
!!
!! Required arrays are assumed-size dummies: inside an OpenACC
!! `host_data use_device(...)` region (or an OpenMP `target data
!! use_device_addr(...)` region) their addresses are the device copies the
!! FElib host already keeps, so no data moves.  Arguments the C side accepts
!! as NULL are `type(c_ptr), value` (pass c_loc(x) or c_null_ptr).  Opaque
!! handles (context, integrator, steppers, network) are `type(c_ptr)`.
!! Every integer function returns 0 (FEDG_OK) or an error code;
!! fedg_last_error() gives the message (fedg_error_message converts it).
!-------------------------------------------------------------------------------
module fedg_c_binding
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: FEDG_OK = 0, FEDG_ERR_ARG = 1, FEDG_ERR_HIP = 2, &
                                       FEDG_ERR_NOMEM = 3, FEDG_ERR_UNSUPPORTED = 4
  integer(c_int), parameter, public :: FEDG_DDENS_ID = 1, FEDG_DRHOT_ID = 2, FEDG_MOMZ_ID = 3, &
                                       FEDG_MOMX_ID = 4, FEDG_MOMY_ID = 5
  integer(c_int), parameter, public :: FEDG_BND_NONE = 0, FEDG_BND_SLIP = 1, FEDG_BND_NOSLIP = 2

  !> structured box (fedg_box_desc)
  type, bind(C), public :: fedg_box_desc
    integer(c_int32_t) :: porder, NeX, NeY, NeZ
    real(c_double) :: xmin, xmax, ymin, ymax, zmin, zmax
    integer(c_int32_t) :: periodic_x, periodic_y, periodic_z
  end type fedg_box_desc

  !> sizes of a box or a context (fedg_box_sizes)
  type, bind(C), public :: fedg_box_sizes
    integer(c_int32_t) :: n1, Np, Nfp_h, Nfp_v, NfpTot
    integer(c_int32_t) :: Ne, NeA, Ne2D, Ne2DA, nB
  end type fedg_box_sizes

  !> host output arrays of fedg_box_build (fedg_box_arrays); any may be c_null_ptr
  type, bind(C), public :: fedg_box_arrays
    type(c_ptr) :: pos_en, J, Escale, sJ, Fscale, normal_fn, vmapM, vmapP, vmapB
    type(c_ptr) :: halo_src, halo_face, EMap3Dto2D, Gsqrt, GI3, GsqrtH
  end type fedg_box_arrays

  !> mesh description (fedg_mesh_desc); pointers are host arrays
  type, bind(C), public :: fedg_mesh_desc
    integer(c_int32_t) :: porder, Ne, NeA, Ne2D, Ne2DA, nB
    type(c_ptr) :: Escale, Gsqrt, GI3, GsqrtH, G11, G12, G22
    type(c_ptr) :: normal_fn, Fscale, vmapM, vmapP, EMap3Dto2D, halo_src, velbc
    type(c_ptr) :: D1D, Lift_mat, IntrpMat_VPOrdM1_tr
    integer(c_int32_t) :: NeX, NeY, NeZ
    real(c_double) :: GRAV, Rdry, CPdry, CVdry, PRES00
  end type fedg_mesh_desc

  public :: fedg_error_message
  public :: fedg_last_error, fedg_version
  public :: fedg_dev_alloc, fedg_dev_free, fedg_memcpy_h2d, fedg_memcpy_d2h, fedg_memset_dev
  public :: fedg_device_sync, fedg_device_name, fedg_stage_probe
  public :: fedg_refelem_tables, fedg_modalfilter_matrix, fedg_box_get_sizes, fedg_box_build
  public :: fedg_box_topography_metrics
  public :: fedg_ctx_create, fedg_ctx_destroy, fedg_ctx_setup_modalfilter, fedg_modalfilter_apply
  public :: fedg_ctx_sizes, fedg_ctx_is_flat
  public :: fedg_drhot2pres, fedg_halo_exchange, fedg_apply_bc_prgvars
  public :: fedg_heve_cal_tend, fedg_heve_cal_tend_cco, fedg_heve_add_bnd_contrib
  public :: fedg_trcadvect_cal_tend, fedg_trcadvect_calc_fct_coef, fedg_trcadvect_tmar
  public :: fedg_trcadvect_save_massflux, fedg_trcadvect_cal_alphdens_advtest
  public :: fedg_tracer_modalfilter_apply, fedg_trcvar_rk_advance
  public :: fedg_tracer_specific_heat, fedg_trc_negative_fixer, fedg_atmos_calc_diagvar
  public :: fedg_hevi_cal_tend, fedg_hevi_cal_vi, fedg_hevi_cal_vi_newton, fedg_hevi_cal_vi_noalloc
  public :: fedg_calc_phyd_hgrad, fedg_add_phy_tend, fedg_heve_set_phys_tend, fedg_hevi_set_phys_tend
  public :: fedg_tint_create, fedg_tint_destroy, fedg_tint_info, fedg_tint_tend_ex, fedg_tint_tend_im
  public :: fedg_tint_store_var0, fedg_tint_advance, fedg_tint_store_implicit
  public :: fedg_heve_create, fedg_heve_destroy, fedg_heve_is_flat, fedg_heve_set_massflux
  public :: fedg_heve_set_precision, fedg_heve_set_modalfilter, fedg_heve_set_state, fedg_heve_get_state
  public :: fedg_heve_state_dev, fedg_heve_run, fedg_heve_run_timed, fedg_heve_get_pressure
  public :: fedg_hevi_create, fedg_hevi_destroy, fedg_hevi_set_state, fedg_hevi_get_state
  public :: fedg_hevi_set_massflux, fedg_hevi_run, fedg_hevi_run_timed, fedg_hevi_newton_stats
  public :: fedg_hevi_set_modalfilter, fedg_hevi_set_precision
  public :: fedg_mlp_create, fedg_mlp_destroy, fedg_mlp_param_count, fedg_mlp_get_params
  public :: fedg_mlp_set_params, fedg_mlp_forward, fedg_mlp_loss_grad, fedg_mlp_adam_step
  public :: fedg_hevi_set_guess, fedg_hevi_cal_vi_guess, fedg_hevi_guess_stats, fedg_hevi_guess_residuals
  public :: fedg_hevi_harvest, fedg_hevi_harvest_count

  interface
    ! ---- errors, version, device memory -------------------------------------
    type(c_ptr) function fedg_last_error() bind(C, name="fedg_last_error")
      import :: c_ptr
    end function fedg_last_error

    type(c_ptr) function fedg_version() bind(C, name="fedg_version")
      import :: c_ptr
    end function fedg_version

    integer(c_int) function fedg_dev_alloc(ptr, bytes) bind(C, name="fedg_dev_alloc")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), intent(out) :: ptr
      integer(c_size_t), value :: bytes
    end function fedg_dev_alloc

    integer(c_int) function fedg_dev_free(ptr) bind(C, name="fedg_dev_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
    end function fedg_dev_free

    integer(c_int) function fedg_memcpy_h2d(dst_dev, src_host, bytes) bind(C, name="fedg_memcpy_h2d")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst_dev, src_host
      integer(c_size_t), value :: bytes
    end function fedg_memcpy_h2d

    integer(c_int) function fedg_memcpy_d2h(dst_host, src_dev, bytes) bind(C, name="fedg_memcpy_d2h")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst_host, src_dev
      integer(c_size_t), value :: bytes
    end function fedg_memcpy_d2h

    integer(c_int) function fedg_memset_dev(dst_dev, val, bytes) bind(C, name="fedg_memset_dev")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst_dev
      integer(c_int), value :: val
      integer(c_size_t), value :: bytes
    end function fedg_memset_dev

    integer(c_int) function fedg_device_sync() bind(C, name="fedg_device_sync")
      import :: c_int
    end function fedg_device_sync

    integer(c_int) function fedg_device_name(buf, buflen) bind(C, name="fedg_device_name")
      import :: c_int, c_char
      character(kind=c_char) :: buf(*)
      integer(c_int), value :: buflen
    end function fedg_device_name

    !> access-pattern probe of the p = 7 stage (measurement only)
    integer(c_int) function fedg_stage_probe(mode, NeX, NeY, NeZ, reps, ms) bind(C, name="fedg_stage_probe")
      import :: c_int, c_double
      integer(c_int), value :: mode, NeX, NeY, NeZ, reps
      real(c_double), intent(out) :: ms
    end function fedg_stage_probe

    ! ---- reference element, modal filter matrix, box mesh (host) ------------
    !> ElementOperationTensorProd3D / HexahedralElement set-up
    !! (ELM/scale_element_operation_tensorprod3D.F90:5512-5569); outputs may be c_null_ptr
    integer(c_int) function fedg_refelem_tables(porder, x1d, w1d, D1D, V1D, invV1D, invM1D, Lift_mat, &
        IntrpMat_VPOrdM1_tr, IndexH2Dto3D, IndexH2Dto3D_bnd) bind(C, name="fedg_refelem_tables")
      import :: c_int, c_ptr
      integer(c_int), value :: porder
      type(c_ptr), value :: x1d, w1d, D1D, V1D, invV1D, invM1D, Lift_mat, IntrpMat_VPOrdM1_tr
      type(c_ptr), value :: IndexH2Dto3D, IndexH2Dto3D_bnd
    end function fedg_refelem_tables

    !> ModalFilter_Init_line with get_exp_filter (ELM/scale_element_modalfilter.F90:60-90,196-236)
    integer(c_int) function fedg_modalfilter_matrix(porder, etac, alpha, ord, F) bind(C, name="fedg_modalfilter_matrix")
      import :: c_int, c_double
      integer(c_int), value :: porder, ord
      real(c_double), value :: etac, alpha
      real(c_double) :: F(*)
    end function fedg_modalfilter_matrix

    !> MeshCubeDom3D_setupLocalDom for one tile (MSH/scale_mesh_cubedom3d.F90:339-469)
    integer(c_int) function fedg_box_get_sizes(desc, sizes) bind(C, name="fedg_box_get_sizes")
      import :: c_int, fedg_box_desc, fedg_box_sizes
      type(fedg_box_desc), intent(in) :: desc
      type(fedg_box_sizes), intent(out) :: sizes
    end function fedg_box_get_sizes

    integer(c_int) function fedg_box_build(desc, out) bind(C, name="fedg_box_build")
      import :: c_int, fedg_box_desc, fedg_box_arrays
      type(fedg_box_desc), intent(in) :: desc
      type(fedg_box_arrays), intent(in) :: out
    end function fedg_box_build

    !> terrain-following metrics of the box (MSH/scale_mesh_topography.F90:159-258,
    !! MSH/scale_meshutil_vcoord.F90:81-209); outputs may be c_null_ptr
    integer(c_int) function fedg_box_topography_metrics(desc, topo, zTop, Gsqrt, GI3, zlev) &
        bind(C, name="fedg_box_topography_metrics")
      import :: c_int, c_ptr, c_double, fedg_box_desc
      type(fedg_box_desc), intent(in) :: desc
      real(c_double), intent(in) :: topo(*)
      real(c_double), value :: zTop
      type(c_ptr), value :: Gsqrt, GI3, zlev
    end function fedg_box_topography_metrics

    ! ---- context ---------------------------------------------------------------
    integer(c_int) function fedg_ctx_create(desc, ctx) bind(C, name="fedg_ctx_create")
      import :: c_int, c_ptr, fedg_mesh_desc
      type(fedg_mesh_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: ctx
    end function fedg_ctx_create

    integer(c_int) function fedg_ctx_destroy(ctx) bind(C, name="fedg_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function fedg_ctx_destroy

    !> Setup_ModalFilter (ELM/scale_element_operation_tensorprod3D.F90:5464-5509); host matrices
    integer(c_int) function fedg_ctx_setup_modalfilter(ctx, MFilter_h1D, MFilter_v1D) &
        bind(C, name="fedg_ctx_setup_modalfilter")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: MFilter_h1D(*), MFilter_v1D(*)
    end function fedg_ctx_setup_modalfilter

    !> atm_dyn_dgm_modalfilter_apply (FDS/scale_atm_dyn_dgm_modalfilter.F90:48-130)
    integer(c_int) function fedg_modalfilter_apply(ctx, DDENS, MOMX, MOMY, MOMZ, DRHOT, do_weight_Gsqrt) &
        bind(C, name="fedg_modalfilter_apply")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DDENS(*), MOMX(*), MOMY(*), MOMZ(*), DRHOT(*)
      integer(c_int), value :: do_weight_Gsqrt
    end function fedg_modalfilter_apply

    integer(c_int) function fedg_ctx_sizes(ctx, sizes) bind(C, name="fedg_ctx_sizes")
      import :: c_int, c_ptr, fedg_box_sizes
      type(c_ptr), value :: ctx
      type(fedg_box_sizes), intent(out) :: sizes
    end function fedg_ctx_sizes

    integer(c_int) function fedg_ctx_is_flat(ctx) bind(C, name="fedg_ctx_is_flat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function fedg_ctx_is_flat

    ! ---- drop-in entry points (device fields) ----------------------------------
    !> atm_dyn_dgm_nonhydro3d_common_DRHOT2PRES (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:428-479)
    integer(c_int) function fedg_drhot2pres(ctx, PRES, DPRES, DRHOT, PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot) &
        bind(C, name="fedg_drhot2pres")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: PRES(*), DPRES(*)
      real(c_double), intent(in) :: DRHOT(*), PRES_hyd(*), THERM_hyd(*), Rtot(*), CVtot(*), CPtot(*)
    end function fedg_drhot2pres

    !> MeshFieldComm_Exchange on one tile (DAT/scale_meshfieldcomm_base.F90:571-805); fields: device
    !! addresses of nfields (Np,NeA) fields
    integer(c_int) function fedg_halo_exchange(ctx, nfields, fields) bind(C, name="fedg_halo_exchange")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nfields
      type(c_ptr), intent(in) :: fields(*)
    end function fedg_halo_exchange

    !> ATMOS_dyn_bnd_applyBC_prgvars_lc (FDS/scale_atm_dyn_dgm_bnd.F90:270-367)
    integer(c_int) function fedg_apply_bc_prgvars(ctx, DDENS, MOMX, MOMY, MOMZ, THERM, DENS_hyd, PRES_hyd) &
        bind(C, name="fedg_apply_bc_prgvars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DDENS(*), MOMX(*), MOMY(*), MOMZ(*), THERM(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*)
    end function fedg_apply_bc_prgvars

    !> atm_dyn_dgm_nonhydro3d_rhot_heve_cal_tend (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:292-489);
    !! the argument list of atm_dyn_nonhydro3d_cal_tend_ex (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:152-199)
    integer(c_int) function fedg_heve_cal_tend(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES_, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd, &
        CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy) bind(C, name="fedg_heve_cal_tend")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*), DPRES_(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*), PRES_hyd_ref(*), THERM_hyd(*), CORIOLIS(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*), DPhydDx(*), DPhydDy(*)
    end function fedg_heve_cal_tend

    !> atm_dyn_dgm_nonhydro3d_rhot_heve_cal_tend_cco (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:493-678)
    integer(c_int) function fedg_heve_cal_tend_cco(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES_, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd, &
        CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy) bind(C, name="fedg_heve_cal_tend_cco")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*), DPRES_(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*), PRES_hyd_ref(*), THERM_hyd(*), CORIOLIS(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*), DPhydDx(*), DPhydDy(*)
    end function fedg_heve_cal_tend_cco

    !> atm_dyn_dgm_nonhydro3d_rhot_heve_add_bnd_contrib_generalvc
    !! (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:242-496); Rtot, CVtot, CPtot all
    !! given or all c_null_ptr (dry air)
    integer(c_int) function fedg_heve_add_bnd_contrib(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES, DENS_hyd, PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot) &
        bind(C, name="fedg_heve_add_bnd_contrib")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*)
      real(c_double) :: DPRES(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*), THERM_hyd(*)
      type(c_ptr), value :: Rtot, CVtot, CPtot
    end function fedg_heve_add_bnd_contrib

    !> atm_dyn_dgm_trcadvect3d_heve_cal_tend (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231)
    integer(c_int) function fedg_trcadvect_cal_tend(ctx, QTRC_dt, QTRC_, MOMX_, MOMY_, MOMZ_, &
        alphDENS_M, alphDENS_P, fct_coef, RHOQ_tp, J) bind(C, name="fedg_trcadvect_cal_tend")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: QTRC_dt(*)
      real(c_double), intent(in) :: QTRC_(*), MOMX_(*), MOMY_(*), MOMZ_(*), alphDENS_M(*), alphDENS_P(*)
      real(c_double), intent(in) :: fct_coef(*), RHOQ_tp(*), J(*)
    end function fedg_trcadvect_cal_tend

    !> atm_dyn_dgm_trcadvect3d_heve_calc_fct_coef (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:232-305);
    !! with disable_limiter /= 0 every argument but fct_coef may be c_null_ptr
    integer(c_int) function fedg_trcadvect_calc_fct_coef(ctx, fct_coef, QTRC_, MOMX_, MOMY_, MOMZ_, RHOQ_tp_, &
        alphDENS_M, alphDENS_P, DENS_hyd, DDENS_, DDENS0_, rk_c_ssm1, dt, J, disable_limiter) &
        bind(C, name="fedg_trcadvect_calc_fct_coef")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: fct_coef(*)
      type(c_ptr), value :: QTRC_, MOMX_, MOMY_, MOMZ_, RHOQ_tp_, alphDENS_M, alphDENS_P
      type(c_ptr), value :: DENS_hyd, DDENS_, DDENS0_
      real(c_double), value :: rk_c_ssm1, dt
      type(c_ptr), value :: J
      integer(c_int), value :: disable_limiter
    end function fedg_trcadvect_calc_fct_coef

    !> atm_dyn_dgm_trcadvect3d_TMAR (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:308-337)
    integer(c_int) function fedg_trcadvect_tmar(ctx, QTRC_, DENS_hyd, DDENS_, J) bind(C, name="fedg_trcadvect_tmar")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: QTRC_(*)
      real(c_double), intent(in) :: DENS_hyd(*), DDENS_(*), J(*)
    end function fedg_trcadvect_tmar

    !> atm_dyn_dgm_trcadvect3d_save_massflux (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:340-404);
    !! Rtot, CVtot, CPtot are unused (may be c_null_ptr)
    integer(c_int) function fedg_trcadvect_save_massflux(ctx, MFLX_x_tavg, MFLX_y_tavg, MFLX_z_tavg, &
        alph_dens_M, alph_dens_P, DDENS, MOMX, MOMY, MOMZ, DPRES, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, &
        rkstage, tavg_weight_h, tavg_weight_v, is_hevi) bind(C, name="fedg_trcadvect_save_massflux")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: MFLX_x_tavg(*), MFLX_y_tavg(*), MFLX_z_tavg(*), alph_dens_M(*), alph_dens_P(*)
      real(c_double), intent(in) :: DDENS(*), MOMX(*), MOMY(*), MOMZ(*), DPRES(*), DENS_hyd(*), PRES_hyd(*)
      type(c_ptr), value :: Rtot, CVtot, CPtot
      integer(c_int), value :: rkstage
      real(c_double), value :: tavg_weight_h, tavg_weight_v
      integer(c_int), value :: is_hevi
    end function fedg_trcadvect_save_massflux

    !> specific heats of the mixture, update_pressure_specific_heat
    !! (FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:653-716); QTRC: device addresses of nq fields
    integer(c_int) function fedg_tracer_specific_heat(ctx, nq, QTRC, TRACER_MASS, TRACER_R, TRACER_CV, TRACER_CP, &
        Rtot, CVtot, CPtot) bind(C, name="fedg_tracer_specific_heat")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nq
      type(c_ptr), intent(in) :: QTRC(*)
      real(c_double), intent(in) :: TRACER_MASS(*), TRACER_R(*), TRACER_CV(*), TRACER_CP(*)
      real(c_double) :: Rtot(*), CVtot(*), CPtot(*)
    end function fedg_tracer_specific_heat

    !> fix_negative_val (FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:596-651) with
    !! atm_phy_mp_dgm_common_negative_fixer (FElib/src/microphysics/scale_atm_phy_mp_dgm_common.F90:429-556);
    !! DRHOT may be c_null_ptr
    integer(c_int) function fedg_trc_negative_fixer(ctx, nq, QTRC, DDENS, PRES, CVtot, CPtot, Rtot, DENS_hyd, &
        PRES_hyd, J, zlev, dt, QLA, QIA, TRACER_MASS, TRACER_R, TRACER_CV, TRACER_CP, DRHOT) &
        bind(C, name="fedg_trc_negative_fixer")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nq
      type(c_ptr), intent(in) :: QTRC(*)
      real(c_double) :: DDENS(*), PRES(*), CVtot(*), CPtot(*), Rtot(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*), J(*), zlev(*)
      real(c_double), value :: dt
      integer(c_int), value :: QLA, QIA
      real(c_double), intent(in) :: TRACER_MASS(*), TRACER_R(*), TRACER_CV(*), TRACER_CP(*)
      type(c_ptr), value :: DRHOT
    end function fedg_trc_negative_fixer

    !> vars_calc_diagnoseVar_lc (model/atm_nonhydro3d/src/atmos/mod_atmos_vars_container.F90:1100-1366)
    !! by name: field_name is trim(name)//c_null_char; a field the named formula does not read
    !! may be c_null_ptr, so every field is passed by c_loc
    integer(c_int) function fedg_atmos_calc_diagvar(ctx, field_name, var_out, DDENS_, MOMX_, MOMY_, MOMZ_, &
        PRES_, nq, QTRC, TRACER_MASS, TRACER_CV, TRACER_ENGI0, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, &
        zlev, pos_z) bind(C, name="fedg_atmos_calc_diagvar")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: field_name(*)
      type(c_ptr), value :: var_out, DDENS_, MOMX_, MOMY_, MOMZ_, PRES_
      integer(c_int), value :: nq
      type(c_ptr), value :: QTRC, TRACER_MASS, TRACER_CV, TRACER_ENGI0
      type(c_ptr), value :: DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, zlev, pos_z
    end function fedg_atmos_calc_diagvar

    !> atm_dyn_dgm_trcadvect3d_heve_cal_alphdens_advtest (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:407-445)
    integer(c_int) function fedg_trcadvect_cal_alphdens_advtest(ctx, alph_dens_M, alph_dens_P, DDENS_, MOMX_, &
        MOMY_, MOMZ_, DENS_hyd) bind(C, name="fedg_trcadvect_cal_alphdens_advtest")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: alph_dens_M(*), alph_dens_P(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DENS_hyd(*)
    end function fedg_trcadvect_cal_alphdens_advtest

    !> atm_dyn_dgm_tracer_modalfilter_apply (FDS/scale_atm_dyn_dgm_modalfilter.F90:232-270);
    !! the filter matrices are host arrays
    integer(c_int) function fedg_tracer_modalfilter_apply(ctx, QTRC_, DENS_hyd, DDENS_, MFilter_tracer_h1D, &
        MFilter_tracer_v1D) bind(C, name="fedg_tracer_modalfilter_apply")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: QTRC_(*)
      real(c_double), intent(in) :: DENS_hyd(*), DDENS_(*), MFilter_tracer_h1D(*), MFilter_tracer_v1D(*)
    end function fedg_tracer_modalfilter_apply

    !> timeint_rk%Advance_trcvar, low-storage schemes (CMN/scale_timeint_rk.F90:1390-1498); the
    !! tables are the timeint_rk object's coef_sig_ex, coef_gam_ex (host) and coef_c_ex
    integer(c_int) function fedg_trcvar_rk_advance(ctx, nowstage, nstage, coef_sig_ex, coef_gam_ex, coef_c_ex, &
        dt, q, var0, varTmp, tend, DDENS, DDENS0, DENS_hyd) bind(C, name="fedg_trcvar_rk_advance")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nowstage, nstage
      real(c_double), intent(in) :: coef_sig_ex(*), coef_gam_ex(*), coef_c_ex(*)
      real(c_double), value :: dt
      real(c_double) :: q(*), var0(*), varTmp(*)
      real(c_double), intent(in) :: tend(*), DDENS(*), DDENS0(*), DENS_hyd(*)
    end function fedg_trcvar_rk_advance

    !> atm_dyn_dgm_nonhydro3d_rhot_hevi_cal_tend (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:289-482)
    integer(c_int) function fedg_hevi_cal_tend(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES_, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd, &
        CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy) bind(C, name="fedg_hevi_cal_tend")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*), DPRES_(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*), PRES_hyd_ref(*), THERM_hyd(*), CORIOLIS(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*), DPhydDx(*), DPhydDy(*)
    end function fedg_hevi_cal_tend

    !> atm_dyn_dgm_nonhydro3d_rhot_hevi_cal_vi (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:772-965);
    !! the argument list of atm_dyn_nonhydro3d_cal_vi (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:201-250)
    integer(c_int) function fedg_hevi_cal_vi(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DENS_hyd, PRES_hyd, &
        DDENS0_, MOMX0_, MOMY0_, MOMZ0_, DRHOT0_, Rtot, CVtot, CPtot, impl_fac, dt) &
        bind(C, name="fedg_hevi_cal_vi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*)
      real(c_double), intent(in) :: DDENS0_(*), MOMX0_(*), MOMY0_(*), MOMZ0_(*), DRHOT0_(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*)
      real(c_double), value :: impl_fac, dt
    end function fedg_hevi_cal_vi

    !> fedg_hevi_cal_vi iterated to a tolerance (no counterpart in the reference); resid (host,
    !! max_it) and iterations may be c_null_ptr
    integer(c_int) function fedg_hevi_cal_vi_newton(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DENS_hyd, PRES_hyd, &
        DDENS0_, MOMX0_, MOMY0_, MOMZ0_, DRHOT0_, Rtot, CVtot, CPtot, impl_fac, dt, max_it, rtol, &
        resid, iterations) bind(C, name="fedg_hevi_cal_vi_newton")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*)
      real(c_double), intent(in) :: DDENS0_(*), MOMX0_(*), MOMY0_(*), MOMZ0_(*), DRHOT0_(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*)
      real(c_double), value :: impl_fac, dt
      integer(c_int), value :: max_it
      real(c_double), value :: rtol
      type(c_ptr), value :: resid, iterations
    end function fedg_hevi_cal_vi_newton

    !> fedg_hevi_cal_vi_newton without any device allocation, for callers that cannot allocate
    !! (no counterpart in the reference): a fixed 64 MiB device workspace per device
    integer(c_int) function fedg_hevi_cal_vi_noalloc(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DENS_hyd, PRES_hyd, &
        DDENS0_, MOMX0_, MOMY0_, MOMZ0_, DRHOT0_, Rtot, CVtot, CPtot, impl_fac, dt, max_it, rtol, &
        resid, iterations) bind(C, name="fedg_hevi_cal_vi_noalloc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*)
      real(c_double), intent(in) :: DDENS0_(*), MOMX0_(*), MOMY0_(*), MOMZ0_(*), DRHOT0_(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*)
      real(c_double), value :: impl_fac, dt
      integer(c_int), value :: max_it
      real(c_double), value :: rtol
      type(c_ptr), value :: resid, iterations
    end function fedg_hevi_cal_vi_noalloc

    !> atm_dyn_dgm_nonhydro3d_common_calc_phyd_hgrad_lc (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:624-777)
    integer(c_int) function fedg_calc_phyd_hgrad(ctx, DPhydDx, DPhydDy, PRES_hyd, PRES_hyd_ref) &
        bind(C, name="fedg_calc_phyd_hgrad")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DPhydDx(*), DPhydDy(*)
      real(c_double), intent(in) :: PRES_hyd(*), PRES_hyd_ref(*)
    end function fedg_calc_phyd_hgrad

    !> add_phy_tend (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:1098-1178, ENTOT_CONSERVE_SCHEME_FLAG off)
    integer(c_int) function fedg_add_phy_tend(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p, PRES, Rtot, CPtot) bind(C, name="fedg_add_phy_tend")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DENS_tp(*), MOMX_tp(*), MOMY_tp(*), MOMZ_tp(*), RHOT_tp(*), RHOH_p(*)
      real(c_double), intent(in) :: PRES(*), Rtot(*), CPtot(*)
    end function fedg_add_phy_tend

    ! ---- time integration (timeint_rk, CMN/scale_timeint_rk.F90) --------------
    !> timeint_rk%Init (CMN/scale_timeint_rk.F90:135-229); scheme: a NUL-terminated name
    integer(c_int) function fedg_tint_create(ctx, scheme, dt, tint) bind(C, name="fedg_tint_create")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: scheme(*)
      real(c_double), value :: dt
      type(c_ptr), intent(out) :: tint
    end function fedg_tint_create

    integer(c_int) function fedg_tint_destroy(tint) bind(C, name="fedg_tint_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: tint
    end function fedg_tint_destroy

    integer(c_int) function fedg_tint_info(tint, nstage, tend_buf_size, imex) bind(C, name="fedg_tint_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: tint
      integer(c_int), intent(out) :: nstage, tend_buf_size, imex
    end function fedg_tint_info

    !> device address of tend_buf2D_ex(:,:,varID,tend_buf_indmap(stage))
    type(c_ptr) function fedg_tint_tend_ex(tint, stage, varID) bind(C, name="fedg_tint_tend_ex")
      import :: c_int, c_ptr
      type(c_ptr), value :: tint
      integer(c_int), value :: stage, varID
    end function fedg_tint_tend_ex

    type(c_ptr) function fedg_tint_tend_im(tint, stage, varID) bind(C, name="fedg_tint_tend_im")
      import :: c_int, c_ptr
      type(c_ptr), value :: tint
      integer(c_int), value :: stage, varID
    end function fedg_tint_tend_im

    !> timeint_rk_store_var0_2D (CMN/scale_timeint_rk.F90:624-647)
    integer(c_int) function fedg_tint_store_var0(tint, q, varID) bind(C, name="fedg_tint_store_var0")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: tint
      real(c_double), intent(in) :: q(*)
      integer(c_int), value :: varID
    end function fedg_tint_store_var0

    !> Advance_varlist, rk_advance_general2D (CMN/scale_timeint_rk.F90:536-573, 2201-2355): q holds
    !! the device addresses of nvar fields, varIDs their variable ids
    integer(c_int) function fedg_tint_advance(tint, stage, nvar, q, varIDs) bind(C, name="fedg_tint_advance")
      import :: c_int, c_ptr
      type(c_ptr), value :: tint
      integer(c_int), value :: stage, nvar
      type(c_ptr), intent(in) :: q(*)
      integer(c_int), intent(in) :: varIDs(*)
    end function fedg_tint_advance

    !> StoreImplicit, rk_storeimpl_general2D (CMN/scale_timeint_rk.F90:2510-2560)
    integer(c_int) function fedg_tint_store_implicit(tint, stage, q, varID) bind(C, name="fedg_tint_store_implicit")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: tint
      integer(c_int), value :: stage
      real(c_double) :: q(*)
      integer(c_int), value :: varID
    end function fedg_tint_store_implicit

    ! ---- fused HEVE stepper (AtmDynDGMDriver_nonhydro3d_update, HEVE) ---------
    !> base state: host (Np,NeA) arrays; Rtot, CVtot, CPtot: c_loc of host arrays or c_null_ptr (dry air)
    integer(c_int) function fedg_heve_create(ctx, scheme, dt, DENS_hyd, PRES_hyd, THERM_hyd, CORIOLIS, &
        DPhydDx, DPhydDy, Rtot, CVtot, CPtot, st) bind(C, name="fedg_heve_create")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: scheme(*)
      real(c_double), value :: dt
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*), THERM_hyd(*), CORIOLIS(*), DPhydDx(*), DPhydDy(*)
      type(c_ptr), value :: Rtot, CVtot, CPtot
      type(c_ptr), intent(out) :: st
    end function fedg_heve_create

    integer(c_int) function fedg_heve_destroy(st) bind(C, name="fedg_heve_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
    end function fedg_heve_destroy

    integer(c_int) function fedg_heve_is_flat(st) bind(C, name="fedg_heve_is_flat")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
    end function fedg_heve_is_flat

    !> per-stage save_massflux for the tracers (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:899-915):
    !! five device addresses, or five c_null_ptr to stop
    integer(c_int) function fedg_heve_set_massflux(st, MFLX_x_tavg, MFLX_y_tavg, MFLX_z_tavg, alph_dens_M, &
        alph_dens_P) bind(C, name="fedg_heve_set_massflux")
      import :: c_int, c_ptr
      type(c_ptr), value :: st, MFLX_x_tavg, MFLX_y_tavg, MFLX_z_tavg, alph_dens_M, alph_dens_P
    end function fedg_heve_set_massflux

    !> physics tendencies added at every stage (add_phy_tend): six device addresses or six c_null_ptr
    integer(c_int) function fedg_heve_set_phys_tend(st, DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p) &
        bind(C, name="fedg_heve_set_phys_tend")
      import :: c_int, c_ptr
      type(c_ptr), value :: st, DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p
    end function fedg_heve_set_phys_tend

    integer(c_int) function fedg_heve_set_precision(st, bits) bind(C, name="fedg_heve_set_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      integer(c_int), value :: bits
    end function fedg_heve_set_precision

    integer(c_int) function fedg_heve_set_modalfilter(st, on) bind(C, name="fedg_heve_set_modalfilter")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      integer(c_int), value :: on
    end function fedg_heve_set_modalfilter

    !> q_host: host addresses of the five (Np,Ne) arrays in variable-id order
    integer(c_int) function fedg_heve_set_state(st, q_host) bind(C, name="fedg_heve_set_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      type(c_ptr), intent(in) :: q_host(5)
    end function fedg_heve_set_state

    integer(c_int) function fedg_heve_get_state(st, q_host) bind(C, name="fedg_heve_get_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      type(c_ptr), intent(in) :: q_host(5)
    end function fedg_heve_get_state

    integer(c_int) function fedg_heve_state_dev(st, q_dev) bind(C, name="fedg_heve_state_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      type(c_ptr), intent(out) :: q_dev(5)
    end function fedg_heve_state_dev

    integer(c_int) function fedg_heve_run(st, nsteps) bind(C, name="fedg_heve_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      integer(c_int), value :: nsteps
    end function fedg_heve_run

    integer(c_int) function fedg_heve_run_timed(st, nsteps, ms_total, ms_per_stage_kernel) &
        bind(C, name="fedg_heve_run_timed")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: st
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: ms_total, ms_per_stage_kernel
    end function fedg_heve_run_timed

    integer(c_int) function fedg_heve_get_pressure(st, PRES_host, DPRES_host) bind(C, name="fedg_heve_get_pressure")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: st
      real(c_double) :: PRES_host(*), DPRES_host(*)
    end function fedg_heve_get_pressure

    ! ---- HEVI stepper (AtmDynDGMDriver_nonhydro3d_update, hevi_flag) ---------
    integer(c_int) function fedg_hevi_create(ctx, scheme, dt, DENS_hyd, PRES_hyd, THERM_hyd, CORIOLIS, &
        DPhydDx, DPhydDy, Rtot, CVtot, CPtot, max_it, rtol, st) bind(C, name="fedg_hevi_create")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: scheme(*)
      real(c_double), value :: dt
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*), THERM_hyd(*), CORIOLIS(*), DPhydDx(*), DPhydDy(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*)
      integer(c_int), value :: max_it
      real(c_double), value :: rtol
      type(c_ptr), intent(out) :: st
    end function fedg_hevi_create

    integer(c_int) function fedg_hevi_destroy(st) bind(C, name="fedg_hevi_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
    end function fedg_hevi_destroy

    integer(c_int) function fedg_hevi_set_state(st, q) bind(C, name="fedg_hevi_set_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      type(c_ptr), intent(in) :: q(5)
    end function fedg_hevi_set_state

    integer(c_int) function fedg_hevi_get_state(st, q) bind(C, name="fedg_hevi_get_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      type(c_ptr), intent(in) :: q(5)
    end function fedg_hevi_get_state

    integer(c_int) function fedg_hevi_set_massflux(st, MFLX_x_tavg, MFLX_y_tavg, MFLX_z_tavg, alph_dens_M, &
        alph_dens_P) bind(C, name="fedg_hevi_set_massflux")
      import :: c_int, c_ptr
      type(c_ptr), value :: st, MFLX_x_tavg, MFLX_y_tavg, MFLX_z_tavg, alph_dens_M, alph_dens_P
    end function fedg_hevi_set_massflux

    integer(c_int) function fedg_hevi_set_phys_tend(st, DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p) &
        bind(C, name="fedg_hevi_set_phys_tend")
      import :: c_int, c_ptr
      type(c_ptr), value :: st, DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p
    end function fedg_hevi_set_phys_tend

    integer(c_int) function fedg_hevi_run(st, nsteps) bind(C, name="fedg_hevi_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      integer(c_int), value :: nsteps
    end function fedg_hevi_run

    integer(c_int) function fedg_hevi_run_timed(st, nsteps, ms_total, ms_implicit_per_stage, &
        ms_explicit_per_stage) bind(C, name="fedg_hevi_run_timed")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: st
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: ms_total, ms_implicit_per_stage, ms_explicit_per_stage
    end function fedg_hevi_run_timed

    integer(c_int) function fedg_hevi_newton_stats(st, iterations, calls) bind(C, name="fedg_hevi_newton_stats")
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: st
      integer(c_long), intent(out) :: iterations, calls
    end function fedg_hevi_newton_stats

    integer(c_int) function fedg_hevi_set_modalfilter(st, on) bind(C, name="fedg_hevi_set_modalfilter")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      integer(c_int), value :: on
    end function fedg_hevi_set_modalfilter

    integer(c_int) function fedg_hevi_set_precision(st, bits) bind(C, name="fedg_hevi_set_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      integer(c_int), value :: bits
    end function fedg_hevi_set_precision

    ! ---- learned initial iterate of the HEVI Newton solve (no counterpart) ----
    integer(c_int) function fedg_mlp_create(n_in, n_hidden, n_hidden_layers, n_out, seed, mlp) &
        bind(C, name="fedg_mlp_create")
      import :: c_int, c_ptr
      integer(c_int), value :: n_in, n_hidden, n_hidden_layers, n_out, seed
      type(c_ptr), intent(out) :: mlp
    end function fedg_mlp_create

    integer(c_int) function fedg_mlp_destroy(mlp) bind(C, name="fedg_mlp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: mlp
    end function fedg_mlp_destroy

    integer(c_int) function fedg_mlp_param_count(mlp, n) bind(C, name="fedg_mlp_param_count")
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: mlp
      integer(c_long), intent(out) :: n
    end function fedg_mlp_param_count

    !> host copies of the parameters and gradients (either may be c_null_ptr)
    integer(c_int) function fedg_mlp_get_params(mlp, params_host, grads_host) bind(C, name="fedg_mlp_get_params")
      import :: c_int, c_ptr
      type(c_ptr), value :: mlp, params_host, grads_host
    end function fedg_mlp_get_params

    integer(c_int) function fedg_mlp_set_params(mlp, params_host) bind(C, name="fedg_mlp_set_params")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: mlp
      real(c_float), intent(in) :: params_host(*)
    end function fedg_mlp_set_params

    integer(c_int) function fedg_mlp_forward(mlp, X_dev, Y_dev, batch) bind(C, name="fedg_mlp_forward")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: mlp
      real(c_float), intent(in) :: X_dev(*)
      real(c_float) :: Y_dev(*)
      integer(c_int), value :: batch
    end function fedg_mlp_forward

    integer(c_int) function fedg_mlp_loss_grad(mlp, X_dev, T_dev, batch, loss) bind(C, name="fedg_mlp_loss_grad")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: mlp
      real(c_float), intent(in) :: X_dev(*), T_dev(*)
      integer(c_int), value :: batch
      real(c_float), intent(out) :: loss
    end function fedg_mlp_loss_grad

    integer(c_int) function fedg_mlp_adam_step(mlp, lr, beta1, beta2, eps) bind(C, name="fedg_mlp_adam_step")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: mlp
      real(c_float), value :: lr, beta1, beta2, eps
    end function fedg_mlp_adam_step

    !> mlp and scale may be c_null_ptr (no network)
    integer(c_int) function fedg_hevi_set_guess(st, mlp, scale) bind(C, name="fedg_hevi_set_guess")
      import :: c_int, c_ptr
      type(c_ptr), value :: st, mlp, scale
    end function fedg_hevi_set_guess

    !> Tex: c_loc of five device addresses (variable-id order) or c_null_ptr; resid, iterations,
    !! fallbacks may be c_null_ptr
    integer(c_int) function fedg_hevi_cal_vi_guess(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DENS_hyd, PRES_hyd, &
        DDENS0_, MOMX0_, MOMY0_, MOMZ0_, DRHOT0_, Rtot, CVtot, CPtot, impl_fac, dt, max_it, rtol, &
        mlp, scale, Tex, resid, iterations, fallbacks) bind(C, name="fedg_hevi_cal_vi_guess")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*)
      real(c_double), intent(in) :: DDENS0_(*), MOMX0_(*), MOMY0_(*), MOMZ0_(*), DRHOT0_(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*)
      real(c_double), value :: impl_fac, dt
      integer(c_int), value :: max_it
      real(c_double), value :: rtol
      type(c_ptr), value :: mlp
      real(c_double), intent(in) :: scale(*)
      type(c_ptr), value :: Tex, resid, iterations, fallbacks
    end function fedg_hevi_cal_vi_guess

    integer(c_int) function fedg_hevi_guess_stats(st, lines, fallbacks) bind(C, name="fedg_hevi_guess_stats")
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: st
      integer(c_long), intent(out) :: lines, fallbacks
    end function fedg_hevi_guess_stats

    integer(c_int) function fedg_hevi_guess_residuals(st, r2_plain, r2_start) &
        bind(C, name="fedg_hevi_guess_residuals")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: st
      real(c_double), intent(out) :: r2_plain, r2_start
    end function fedg_hevi_guess_residuals

    !> X_dev, Y_dev: device fp32 buffers (capacity rows)
    integer(c_int) function fedg_hevi_harvest(st, X_dev, Y_dev, capacity, per_stage, seed) &
        bind(C, name="fedg_hevi_harvest")
      import :: c_int, c_ptr
      type(c_ptr), value :: st, X_dev, Y_dev
      integer(c_int), value :: capacity, per_stage, seed
    end function fedg_hevi_harvest

    integer(c_int) function fedg_hevi_harvest_count(st, count) bind(C, name="fedg_hevi_harvest_count")
      import :: c_int, c_ptr
      type(c_ptr), value :: st
      integer(c_int), intent(out) :: count
    end function fedg_hevi_harvest_count
  end interface

contains

  !> fedg_last_error() as a Fortran string
  function fedg_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = fedg_last_error()
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function fedg_error_message
end module fedg_c_binding
