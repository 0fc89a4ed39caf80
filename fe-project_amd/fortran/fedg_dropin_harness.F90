!-------------------------------------------------------------------------------
!> A Fortran host of the drop-in tendency: builds the box mesh and the
!! reference-element operators through the C ABI, creates the context, and
!! calls fedg_dropin_cal_tend_ex (the cal_tend_ex argument list) on the fields
!! of an input file.
!!
!! usage: fedg_dropin_harness <input> <output>
!!   input (stream, native byte order): int32 porder, NeX, NeY, NeZ,
!!     periodic_x, periodic_y, periodic_z; float64 xmin, xmax, ymin, ymax,
!!     zmin, zmax; float64 GRAV, Rdry, CPdry, CVdry, PRES00; then the
!!     (Np,NeA) fields DDENS_, MOMX_, MOMY_, MOMZ_, THERM_, DPRES_, DENS_hyd,
!!     PRES_hyd, PRES_hyd_ref, THERM_hyd, Rtot, CVtot, CPtot, DPhydDx,
!!     DPhydDy and the (Nfp_v,Ne2DA) CORIOLIS;
!!   output: DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, (Np,Ne) each.
!! Velocity boundary condition: SLIP on every closed boundary.
!-------------------------------------------------------------------------------
program fedg_dropin_harness
  use iso_c_binding
  use fedg_c_binding
  use fedg_dropin
  implicit none

  type(fedg_box_desc) :: bd
  type(fedg_box_sizes) :: sz
  type(fedg_box_arrays) :: ba
  type(fedg_mesh_desc) :: md
  type(c_ptr) :: ctx
  character(len=4096) :: fin, fout
  integer(c_int32_t) :: hdr(7)
  real(c_double) :: ext(6), cst(5)
  integer :: u, stat, Np, Ne, NeA, Nfp_v, Ne2DA, NfpTot, n1, nB, k
  real(c_double), allocatable, target :: pos_en(:), Jac(:), Escale(:), sJ(:), Fscale(:), normal_fn(:)
  real(c_double), allocatable, target :: Gsqrt(:), GI3(:), GsqrtH(:)
  integer(c_int32_t), allocatable, target :: vmapM(:), vmapP(:), vmapB(:), halo_src(:), halo_face(:), EMap(:)
  integer(c_int32_t), allocatable, target :: velbc(:)
  real(c_double), allocatable, target :: D1D(:), Lift(:), Itr(:)
  real(c_double), allocatable, target :: q(:,:,:), cor(:,:), dt(:,:,:)
  integer :: bnd_open(6)

  if (command_argument_count() /= 2) then
    write(*, '(a)') 'usage: fedg_dropin_harness <input> <output>'
    stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) hdr, ext, cst
  bd%porder = hdr(1); bd%NeX = hdr(2); bd%NeY = hdr(3); bd%NeZ = hdr(4)
  bd%periodic_x = hdr(5); bd%periodic_y = hdr(6); bd%periodic_z = hdr(7)
  bd%xmin = ext(1); bd%xmax = ext(2); bd%ymin = ext(3); bd%ymax = ext(4); bd%zmin = ext(5); bd%zmax = ext(6)
  call ok(fedg_box_get_sizes(bd, sz), 'fedg_box_get_sizes')
  n1 = sz%n1; Np = sz%Np; Ne = sz%Ne; NeA = sz%NeA; Nfp_v = sz%Nfp_v; Ne2DA = sz%Ne2DA
  NfpTot = sz%NfpTot; nB = sz%nB

  ! MeshCubeDom3D for one tile
  allocate(pos_en(Np * Ne * 3), Jac(Np * Ne), Escale(Np * Ne * 9), sJ(NfpTot * Ne), Fscale(NfpTot * Ne))
  allocate(normal_fn(NfpTot * Ne * 3), Gsqrt(Np * NeA), GI3(Np * NeA * 2), GsqrtH(Nfp_v * sz%Ne2D))
  allocate(vmapM(NfpTot * Ne), vmapP(NfpTot * Ne), vmapB(max(nB, 1)), halo_src(max(nB, 1)))
  allocate(halo_face(max(nB, 1)), EMap(Ne), velbc(max(nB, 1)))
  ba%pos_en = c_loc(pos_en); ba%J = c_loc(Jac); ba%Escale = c_loc(Escale); ba%sJ = c_loc(sJ)
  ba%Fscale = c_loc(Fscale); ba%normal_fn = c_loc(normal_fn); ba%vmapM = c_loc(vmapM)
  ba%vmapP = c_loc(vmapP); ba%vmapB = c_loc(vmapB); ba%halo_src = c_loc(halo_src)
  ba%halo_face = c_loc(halo_face); ba%EMap3Dto2D = c_loc(EMap); ba%Gsqrt = c_loc(Gsqrt)
  ba%GI3 = c_loc(GI3); ba%GsqrtH = c_loc(GsqrtH)
  call ok(fedg_box_build(bd, ba), 'fedg_box_build')

  ! ElementOperationTensorProd3D: D1D, Lift_mat, IntrpMat_VPOrdM1_tr
  allocate(D1D(n1 * n1), Lift(Np * 6), Itr(n1 * n1))
  call ok(fedg_refelem_tables(int(bd%porder, c_int), c_null_ptr, c_null_ptr, c_loc(D1D), c_null_ptr, &
                              c_null_ptr, c_null_ptr, c_loc(Lift), c_loc(Itr), c_null_ptr, c_null_ptr), &
          'fedg_refelem_tables')

  ! SLIP walls on the closed boundaries (y-, x+, y+, x-, z-, z+)
  bnd_open = [bd%periodic_y, bd%periodic_x, bd%periodic_y, bd%periodic_x, bd%periodic_z, bd%periodic_z]
  do k = 1, nB
    if (bnd_open(halo_face(k)) /= 0) then
      velbc(k) = FEDG_BND_NONE
    else
      velbc(k) = FEDG_BND_SLIP
    end if
  end do

  md%porder = bd%porder; md%Ne = Ne; md%NeA = NeA; md%Ne2D = sz%Ne2D; md%Ne2DA = Ne2DA; md%nB = nB
  md%Escale = c_loc(Escale); md%Gsqrt = c_loc(Gsqrt); md%GI3 = c_loc(GI3); md%GsqrtH = c_loc(GsqrtH)
  md%G11 = c_null_ptr; md%G12 = c_null_ptr; md%G22 = c_null_ptr
  md%normal_fn = c_loc(normal_fn); md%Fscale = c_loc(Fscale); md%vmapM = c_loc(vmapM)
  md%vmapP = c_loc(vmapP); md%EMap3Dto2D = c_loc(EMap); md%halo_src = c_loc(halo_src)
  md%velbc = c_loc(velbc); md%D1D = c_loc(D1D); md%Lift_mat = c_loc(Lift)
  md%IntrpMat_VPOrdM1_tr = c_loc(Itr)
  md%NeX = bd%NeX; md%NeY = bd%NeY; md%NeZ = bd%NeZ
  md%GRAV = cst(1); md%Rdry = cst(2); md%CPdry = cst(3); md%CVdry = cst(4); md%PRES00 = cst(5)
  call ok(fedg_ctx_create(md, ctx), 'fedg_ctx_create')

  ! fields: 1 DDENS_, 2 MOMX_, 3 MOMY_, 4 MOMZ_, 5 THERM_, 6 DPRES_, 7 DENS_hyd, 8 PRES_hyd,
  ! 9 PRES_hyd_ref, 10 THERM_hyd, 11 Rtot, 12 CVtot, 13 CPtot, 14 DPhydDx, 15 DPhydDy
  allocate(q(Np, NeA, 15), cor(Nfp_v, Ne2DA), dt(Np, NeA, 5))
  read(u) q, cor
  close(u)

  call fedg_dropin_cal_tend_ex( &
      dt(:,:,1), dt(:,:,2), dt(:,:,3), dt(:,:,4), dt(:,:,5),                 &
      q(:,:,1), q(:,:,2), q(:,:,3), q(:,:,4), q(:,:,5), q(:,:,6),            &
      q(:,:,7), q(:,:,8), q(:,:,9), q(:,:,10), cor,                          &
      q(:,:,11), q(:,:,12), q(:,:,13),                                       &
      q(:,:,14), q(:,:,15),                                                  &
      ctx, Np, NeA, Nfp_v, Ne2DA, stat)
  call ok(stat, 'fedg_dropin_cal_tend_ex')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) dt(:, 1:Ne, 1:5)
  close(u)
  call ok(fedg_ctx_destroy(ctx), 'fedg_ctx_destroy')
  write(*, '(a,i0,a,i0)') 'fedg_dropin_harness: cal_tend_ex on ', Ne, ' elements of p = ', bd%porder

contains

  subroutine ok(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= FEDG_OK) then
      write(*, '(a,a,i0,a,a)') what, ' failed: ', rc, ': ', fedg_error_message()
      stop 1
    end if
  end subroutine ok
end program fedg_dropin_harness
