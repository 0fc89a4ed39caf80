!-------------------------------------------------------------------------------
!> ISO_C_BINDING interfaces of libfedg.so (include/fedg.h) for a Fortran host.
!!
!! Arrays are passed as assumed-size dummies: inside an OpenACC
!! `host_data use_device(...)` region (or an OpenMP `target data
!! use_device_addr(...)` region) their addresses are the device copies the
!! FElib host already keeps, so no data moves.  Every function returns 0
!! (FEDG_OK) or an error code; fedg_last_error() gives the message.
!-------------------------------------------------------------------------------
module fedg_c_binding
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: FEDG_OK = 0

  !> mesh description (include/fedg.h fedg_mesh_desc); pointers are host arrays
  type, bind(C), public :: fedg_mesh_desc
    integer(c_int) :: porder, Ne, NeA, Ne2D, Ne2DA, nB
    type(c_ptr) :: Escale, Gsqrt, GI3, GsqrtH, G11, G12, G22
    type(c_ptr) :: normal_fn, Fscale, vmapM, vmapP, EMap3Dto2D, halo_src, velbc
    type(c_ptr) :: D1D, Lift_mat, IntrpMat_VPOrdM1_tr
    integer(c_int) :: NeX, NeY, NeZ
    real(c_double) :: GRAV, Rdry, CPdry, CVdry, PRES00
  end type fedg_mesh_desc

  public :: fedg_ctx_create, fedg_ctx_destroy, fedg_last_error
  public :: fedg_heve_cal_tend, fedg_hevi_cal_tend, fedg_hevi_cal_vi
  public :: fedg_drhot2pres, fedg_calc_phyd_hgrad
  public :: fedg_heve_cal_tend_cco, fedg_heve_add_bnd_contrib
  public :: fedg_trcadvect_cal_tend, fedg_trcadvect_calc_fct_coef, fedg_trcadvect_tmar
  public :: fedg_trcadvect_save_massflux, fedg_trcadvect_cal_alphdens_advtest, fedg_trcvar_rk_advance
  public :: fedg_tracer_modalfilter_apply

  interface
    integer(c_int) function fedg_ctx_create(desc, ctx) bind(C, name="fedg_ctx_create")
      import :: c_int, c_ptr, fedg_mesh_desc
      type(fedg_mesh_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: ctx
    end function fedg_ctx_create

    integer(c_int) function fedg_ctx_destroy(ctx) bind(C, name="fedg_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function fedg_ctx_destroy

    type(c_ptr) function fedg_last_error() bind(C, name="fedg_last_error")
      import :: c_ptr
    end function fedg_last_error

    !> atm_dyn_dgm_nonhydro3d_rhot_heve_cal_tend (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:292-489)
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

    !> atm_dyn_dgm_nonhydro3d_rhot_hevi_cal_vi (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:772-965)
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

    !> atm_dyn_dgm_nonhydro3d_common_DRHOT2PRES (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:428-479)
    integer(c_int) function fedg_drhot2pres(ctx, PRES, DPRES, DRHOT, PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot) &
        bind(C, name="fedg_drhot2pres")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: PRES(*), DPRES(*)
      real(c_double), intent(in) :: DRHOT(*), PRES_hyd(*), THERM_hyd(*), Rtot(*), CVtot(*), CPtot(*)
    end function fedg_drhot2pres

    !> atm_dyn_dgm_nonhydro3d_common_calc_phyd_hgrad_lc (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:624-777)
    integer(c_int) function fedg_calc_phyd_hgrad(ctx, DPhydDx, DPhydDy, PRES_hyd, PRES_hyd_ref) &
        bind(C, name="fedg_calc_phyd_hgrad")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DPhydDx(*), DPhydDy(*)
      real(c_double), intent(in) :: PRES_hyd(*), PRES_hyd_ref(*)
    end function fedg_calc_phyd_hgrad

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
    !! (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:242-496); Rtot, CVtot, CPtot
    !! are all given or all c_null_ptr (dry-air constants)
    integer(c_int) function fedg_heve_add_bnd_contrib(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, &
        DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES, DENS_hyd, PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot) &
        bind(C, name="fedg_heve_add_bnd_contrib")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: DENS_dt(*), MOMX_dt(*), MOMY_dt(*), MOMZ_dt(*), RHOT_dt(*), DPRES(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DRHOT_(*)
      real(c_double), intent(in) :: DENS_hyd(*), PRES_hyd(*), THERM_hyd(*)
      type(c_ptr), value :: Rtot, CVtot, CPtot
    end function fedg_heve_add_bnd_contrib

    !> atm_dyn_dgm_trcadvect3d_heve_cal_tend (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231);
    !! FaceIntMat is built by the context, J is lcmesh%J
    integer(c_int) function fedg_trcadvect_cal_tend(ctx, QTRC_dt, QTRC_, MOMX_, MOMY_, MOMZ_, &
        alphDENS_M, alphDENS_P, fct_coef, RHOQ_tp, J) bind(C, name="fedg_trcadvect_cal_tend")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: QTRC_dt(*)
      real(c_double), intent(in) :: QTRC_(*), MOMX_(*), MOMY_(*), MOMZ_(*), alphDENS_M(*), alphDENS_P(*)
      real(c_double), intent(in) :: fct_coef(*), RHOQ_tp(*), J(*)
    end function fedg_trcadvect_cal_tend

    !> atm_dyn_dgm_trcadvect3d_heve_calc_fct_coef (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:232-305)
    integer(c_int) function fedg_trcadvect_calc_fct_coef(ctx, fct_coef, QTRC_, MOMX_, MOMY_, MOMZ_, RHOQ_tp_, &
        alphDENS_M, alphDENS_P, DENS_hyd, DDENS_, DDENS0_, rk_c_ssm1, dt, J, disable_limiter) &
        bind(C, name="fedg_trcadvect_calc_fct_coef")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: fct_coef(*)
      real(c_double), intent(in) :: QTRC_(*), MOMX_(*), MOMY_(*), MOMZ_(*), RHOQ_tp_(*)
      real(c_double), intent(in) :: alphDENS_M(*), alphDENS_P(*), DENS_hyd(*), DDENS_(*), DDENS0_(*), J(*)
      real(c_double), value :: rk_c_ssm1, dt
      integer(c_int), value :: disable_limiter
    end function fedg_trcadvect_calc_fct_coef

    !> atm_dyn_dgm_trcadvect3d_TMAR (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:308-337)
    integer(c_int) function fedg_trcadvect_tmar(ctx, QTRC_, DENS_hyd, DDENS_, J) bind(C, name="fedg_trcadvect_tmar")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: QTRC_(*)
      real(c_double), intent(in) :: DENS_hyd(*), DDENS_(*), J(*)
    end function fedg_trcadvect_tmar

    !> atm_dyn_dgm_trcadvect3d_save_massflux (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:340-404)
    integer(c_int) function fedg_trcadvect_save_massflux(ctx, MFLX_x_tavg, MFLX_y_tavg, MFLX_z_tavg, &
        alph_dens_M, alph_dens_P, DDENS, MOMX, MOMY, MOMZ, DPRES, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, &
        rkstage, tavg_weight_h, tavg_weight_v, is_hevi) bind(C, name="fedg_trcadvect_save_massflux")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: MFLX_x_tavg(*), MFLX_y_tavg(*), MFLX_z_tavg(*), alph_dens_M(*), alph_dens_P(*)
      real(c_double), intent(in) :: DDENS(*), MOMX(*), MOMY(*), MOMZ(*), DPRES(*), DENS_hyd(*), PRES_hyd(*)
      real(c_double), intent(in) :: Rtot(*), CVtot(*), CPtot(*)
      integer(c_int), value :: rkstage, is_hevi
      real(c_double), value :: tavg_weight_h, tavg_weight_v
    end function fedg_trcadvect_save_massflux

    !> atm_dyn_dgm_trcadvect3d_heve_cal_alphdens_advtest (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:407-445)
    integer(c_int) function fedg_trcadvect_cal_alphdens_advtest(ctx, alph_dens_M, alph_dens_P, DDENS_, MOMX_, &
        MOMY_, MOMZ_, DENS_hyd) bind(C, name="fedg_trcadvect_cal_alphdens_advtest")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: alph_dens_M(*), alph_dens_P(*)
      real(c_double), intent(in) :: DDENS_(*), MOMX_(*), MOMY_(*), MOMZ_(*), DENS_hyd(*)
    end function fedg_trcadvect_cal_alphdens_advtest

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

    !> atm_dyn_dgm_tracer_modalfilter_apply (FDS/scale_atm_dyn_dgm_modalfilter.F90:232-270);
    !! the filter matrices are host arrays
    integer(c_int) function fedg_tracer_modalfilter_apply(ctx, QTRC_, DENS_hyd, DDENS_, MFilter_tracer_h1D, &
        MFilter_tracer_v1D) bind(C, name="fedg_tracer_modalfilter_apply")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: QTRC_(*)
      real(c_double), intent(in) :: DENS_hyd(*), DDENS_(*), MFilter_tracer_h1D(*), MFilter_tracer_v1D(*)
    end function fedg_tracer_modalfilter_apply
  end interface
end module fedg_c_binding
