!-------------------------------------------------------------------------------
!> Host-array form of the drop-in tendency for a Fortran host without device
!! copies of its fields: the argument list of atm_dyn_nonhydro3d_cal_tend_ex
!! (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:152-199), its derived-type
!! arguments (element3D_operation, Dx..Lift, lmesh, elem, lmesh2D, elem2D)
!! replaced by the fedg context and the array extents they carry.  The fields
!! go to device buffers, fedg_heve_cal_tend runs on them, and the tendencies
!! come back.  A host whose fields already live on the device calls
!! fedg_heve_cal_tend itself inside `host_data use_device(...)`.
!-------------------------------------------------------------------------------
module fedg_dropin
  use iso_c_binding
  use fedg_c_binding
  implicit none
  private
  public :: fedg_dropin_cal_tend_ex

  type :: devbuf
    type(c_ptr) :: p = c_null_ptr
    real(c_double), pointer :: a(:) => null()
  end type devbuf

contains

  subroutine fedg_dropin_cal_tend_ex( &
      DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt,           & ! (out)
      DDENS_, MOMX_, MOMY_, MOMZ_, THERM_, DPRES_,           & ! (in)
      DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd, CORIOLIS, & ! (in)
      Rtot, CVtot, CPtot,                                    & ! (in)
      DPhydDx, DPhydDy,                                      & ! (in)
      ctx, Np, NeA, Np2D, NeA2D, stat)
    type(c_ptr), intent(in) :: ctx          ! element3D_operation, lmesh, elem, lmesh2D, elem2D
    integer, intent(in) :: Np, NeA          ! elem%Np, lmesh%NeA
    integer, intent(in) :: Np2D, NeA2D      ! elem2D%Np, lmesh2D%NeA
    real(c_double), intent(out), target :: DENS_dt(Np,NeA)
    real(c_double), intent(out), target :: MOMX_dt(Np,NeA)
    real(c_double), intent(out), target :: MOMY_dt(Np,NeA)
    real(c_double), intent(out), target :: MOMZ_dt(Np,NeA)
    real(c_double), intent(out), target :: RHOT_dt(Np,NeA)
    real(c_double), intent(in), target :: DDENS_(Np,NeA)
    real(c_double), intent(in), target :: MOMX_(Np,NeA)
    real(c_double), intent(in), target :: MOMY_(Np,NeA)
    real(c_double), intent(in), target :: MOMZ_(Np,NeA)
    real(c_double), intent(in), target :: THERM_(Np,NeA)
    real(c_double), intent(in), target :: DPRES_(Np,NeA)
    real(c_double), intent(in), target :: DENS_hyd(Np,NeA)
    real(c_double), intent(in), target :: PRES_hyd(Np,NeA)
    real(c_double), intent(in), target :: PRES_hyd_ref(Np,NeA)
    real(c_double), intent(in), target :: THERM_hyd(Np,NeA)
    real(c_double), intent(in), target :: CORIOLIS(Np2D,NeA2D)
    real(c_double), intent(in), target :: Rtot(Np,NeA)
    real(c_double), intent(in), target :: CVtot(Np,NeA)
    real(c_double), intent(in), target :: CPtot(Np,NeA)
    real(c_double), intent(in), target :: DPhydDx(Np,NeA)
    real(c_double), intent(in), target :: DPhydDy(Np,NeA)
    integer, intent(out) :: stat

    integer, parameter :: nbuf = 21
    type(devbuf) :: b(nbuf)
    integer(c_size_t) :: n3, n2, nb
    integer :: k

    n3 = int(Np, c_size_t) * int(NeA, c_size_t)
    n2 = int(Np2D, c_size_t) * int(NeA2D, c_size_t)
    stat = FEDG_OK
    do k = 1, nbuf
      nb = n3
      if (k == 16) nb = n2  ! CORIOLIS
      stat = fedg_dev_alloc(b(k)%p, nb * 8_c_size_t)
      if (stat /= FEDG_OK) exit
      call c_f_pointer(b(k)%p, b(k)%a, [nb])
    end do
    ! inputs, in the order of the argument list (buffers 6..21)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(6)%p, c_loc(DDENS_), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(7)%p, c_loc(MOMX_), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(8)%p, c_loc(MOMY_), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(9)%p, c_loc(MOMZ_), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(10)%p, c_loc(THERM_), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(11)%p, c_loc(DPRES_), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(12)%p, c_loc(DENS_hyd), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(13)%p, c_loc(PRES_hyd), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(14)%p, c_loc(PRES_hyd_ref), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(15)%p, c_loc(THERM_hyd), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(16)%p, c_loc(CORIOLIS), n2 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(17)%p, c_loc(Rtot), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(18)%p, c_loc(CVtot), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(19)%p, c_loc(CPtot), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(20)%p, c_loc(DPhydDx), n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_h2d(b(21)%p, c_loc(DPhydDy), n3 * 8)
    if (stat == FEDG_OK) &
      stat = fedg_heve_cal_tend(ctx, b(1)%a, b(2)%a, b(3)%a, b(4)%a, b(5)%a, &
                                b(6)%a, b(7)%a, b(8)%a, b(9)%a, b(10)%a, b(11)%a, &
                                b(12)%a, b(13)%a, b(14)%a, b(15)%a, b(16)%a, &
                                b(17)%a, b(18)%a, b(19)%a, b(20)%a, b(21)%a)
    if (stat == FEDG_OK) stat = fedg_memcpy_d2h(c_loc(DENS_dt), b(1)%p, n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_d2h(c_loc(MOMX_dt), b(2)%p, n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_d2h(c_loc(MOMY_dt), b(3)%p, n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_d2h(c_loc(MOMZ_dt), b(4)%p, n3 * 8)
    if (stat == FEDG_OK) stat = fedg_memcpy_d2h(c_loc(RHOT_dt), b(5)%p, n3 * 8)
    do k = 1, nbuf
      if (c_associated(b(k)%p)) then
        if (fedg_dev_free(b(k)%p) /= FEDG_OK .and. stat == FEDG_OK) stat = FEDG_ERR_HIP
      end if
    end do
  end subroutine fedg_dropin_cal_tend_ex
end module fedg_dropin
