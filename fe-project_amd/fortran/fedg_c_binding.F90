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
  end interface
end module fedg_c_binding
