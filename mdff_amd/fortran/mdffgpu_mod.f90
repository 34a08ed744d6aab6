! mdffgpu_mod.f90 -- ISO_C_BINDING interface of the MI355X force/field
! library (include/mdffgpu.h) for MDFF's Fortran host.
!
! The derived types mirror struct mdffgpu_params / mdffgpu_energies byte
! for byte (bind(C)); the tables have the shapes of the reference's module
! arrays (ntypemax = 16), so the host copies them as they are:
!   prm%epslj = epslj ; prm%poldip = poldip ; prm%dip_pol_damp_b = ...
! engforce_driver_gpu replaces the body of engforce_driver
! (src/field.f90:1111-1166) and keeps its error convention: a failure prints
! the library's message and STOPs (e.g. src/field.f90:303-305).
module mdffgpu_mod
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: MDFFGPU_NTYPEMAX = 16
  integer(c_int32_t), parameter, public :: MDFFGPU_PREC_FP64 = 0, MDFFGPU_PREC_MIXED = 1
  integer(c_int32_t), parameter, public :: MDFFGPU_EXT_POLY = 0, MDFFGPU_EXT_ASPC = 1
  integer(c_int32_t), parameter, public :: MDFFGPU_SCF_JACOBI = 0, MDFFGPU_SCF_PCG = 1

  type, bind(C), public :: mdffgpu_params
    integer(c_int32_t) :: natm, ntype
    integer(c_int32_t) :: lnmlj, lmorse, lbmhft, lbmhftd, lcoulomb, lsurf
    integer(c_int32_t) :: calc_mode, trunc
    real(c_double) :: cutshortrange, cutlongrange, skindiff, coul_unit, press_unit
    real(c_double) :: plj(16, 16), qlj(16, 16), epslj(16, 16), sigmalj(16, 16)
    real(c_double) :: epsmor(16, 16), sigmamor(16, 16), rhomor(16, 16)
    real(c_double) :: Abmhftd(16, 16), Bbmhftd(16, 16), Cbmhftd(16, 16), Dbmhftd(16, 16)
    real(c_double) :: BDbmhftd(16, 16)
    real(c_double) :: qch(16)
    real(c_double) :: dip(16, 3)
    real(c_double) :: poldip(16, 3, 3)
    integer(c_int32_t) :: ldip_polar(16)
    integer(c_int32_t) :: ldip_damping(16, 16, 16)
    real(c_double) :: dip_pol_damp_b(16, 16, 16), dip_pol_damp_c(16, 16, 16)
    integer(c_int32_t) :: dip_pol_damp_k(16, 16, 16)
    real(c_double) :: alphaES
    integer(c_int32_t) :: kES(3), task_coul(3)
    real(c_double) :: conv_tol_ind
    integer(c_int32_t) :: min_scf_pol_iter, max_scf_pol_iter, extrapolate_order
    integer(c_int32_t) :: algo_ext_dipole, doefield, doefg
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: mixed_fp64_parts
    integer(c_int32_t) :: scf_solver
    integer(c_int32_t) :: task_coul_quad(3), reserved(2)
    real(c_double) :: quad(16, 3, 3)
    real(c_double) :: polquad(16, 3, 3, 3)
    integer(c_int32_t) :: lquad_polar(16)
    integer(c_int32_t) :: lwfc(16)
    integer(c_int32_t) :: ldip_wfc, reserved2
    real(c_double) :: rcut_wfc
  end type

  type, bind(C), public :: mdffgpu_energies
    real(c_double) :: u_lj, vir_lj, u_morse, vir_morse, u_bmhft, vir_bmhft
    real(c_double) :: u_coul, u_pol, pvirial_coul
    real(c_double) :: tau_nonb(3, 3), tau_coul(3, 3)
    integer(c_int32_t) :: n_scf, n_es_calls
    real(c_double) :: rmsd, scf_tol
  end type

  interface
    integer(c_int) function mdffgpu_create(params, itype, A, ctx) bind(C, name="mdffgpu_create")
      import :: c_int, c_int32_t, c_double, c_ptr, mdffgpu_params
      type(mdffgpu_params), intent(in) :: params
      integer(c_int32_t), intent(in) :: itype(*)
      real(c_double), intent(in) :: A(3, 3)
      type(c_ptr), intent(out) :: ctx
    end function

    subroutine mdffgpu_destroy(ctx) bind(C, name="mdffgpu_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_int) function mdffgpu_set_cell(ctx, A) bind(C, name="mdffgpu_set_cell")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: A(3, 3)
    end function

    integer(c_int) function mdffgpu_engforce(ctx, rx, ry, rz, itime, fx, fy, fz, mu_t, theta_t, &
                                             ef_t, efg_t, en) bind(C, name="mdffgpu_engforce")
      import :: c_int, c_int32_t, c_double, c_ptr, mdffgpu_energies
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rx(*), ry(*), rz(*)
      integer(c_int32_t), value :: itime
      real(c_double), intent(inout) :: fx(*), fy(*), fz(*)
      real(c_double), intent(out) :: mu_t(*), theta_t(*), ef_t(*), efg_t(*)
      type(mdffgpu_energies), intent(out) :: en
    end function

    integer(c_int) function mdffgpu_get_dipole_history(ctx, hist) &
        bind(C, name="mdffgpu_get_dipole_history")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: hist(*)
    end function

    integer(c_int) function mdffgpu_set_dipole_history(ctx, hist) &
        bind(C, name="mdffgpu_set_dipole_history")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: hist(*)
    end function

    integer(c_int32_t) function mdffgpu_num_kvectors(ctx) bind(C, name="mdffgpu_num_kvectors")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function mdffgpu_last_error(ctx) bind(C, name="mdffgpu_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface

  public :: mdffgpu_create, mdffgpu_destroy, mdffgpu_set_cell, mdffgpu_engforce
  public :: mdffgpu_get_dipole_history, mdffgpu_set_dipole_history, mdffgpu_num_kvectors
  public :: mdffgpu_last_error, mdffgpu_check, mdffgpu_message, engforce_driver_gpu

contains

  ! Message of the last failure on ctx (c_null_ptr: last create).
  function mdffgpu_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    call c_f_pointer(mdffgpu_last_error(ctx), s, [512])
    do i = 1, 512
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function

  ! The reference's error convention: print and STOP.
  subroutine mdffgpu_check(status, ctx, where)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: where
    if (status /= 0) then
      write (*, '(a,a,a,i0,a,a)') 'ERROR ', where, ': mdffgpu status ', status, ' : ', &
        trim(mdffgpu_message(ctx))
      stop
    end if
  end subroutine

  ! Body of engforce_driver: forces (in/out, overwritten by the short-range
  ! terms and incremented by the Coulomb ones), moments, fields, EFG and the
  ! thermodynamic scalars of one evaluation.
  subroutine engforce_driver_gpu(ctx, natm, rx, ry, rz, itime, fx, fy, fz, mu_t, theta_t, &
                                 ef_t, efg_t, en)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: natm, itime
    real(c_double), intent(in) :: rx(natm), ry(natm), rz(natm)
    real(c_double), intent(inout) :: fx(natm), fy(natm), fz(natm)
    real(c_double), intent(out) :: mu_t(natm, 3), theta_t(natm, 3, 3)
    real(c_double), intent(out) :: ef_t(natm, 3), efg_t(natm, 3, 3)
    type(mdffgpu_energies), intent(out) :: en
    integer(c_int) :: st
    st = mdffgpu_engforce(ctx, rx, ry, rz, int(itime, c_int32_t), fx, fy, fz, mu_t, theta_t, &
                          ef_t, efg_t, en)
    call mdffgpu_check(st, ctx, 'engforce_driver')
  end subroutine

end module mdffgpu_mod
