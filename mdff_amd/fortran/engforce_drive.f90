! engforce_drive.f90 -- Fortran host program over the C ABI: reads one
! configuration, calls engforce_driver_gpu for nsteps frames, and writes
! the results.  It stands in for MDFF's md_run loop (src/runmd.f90:231)
! calling engforce_driver, and is what the GPU tests run to check the
! Fortran -> ISO_C_BINDING -> HIP path.
!
! input  (stream, native endianness): the mdffgpu_params struct, itype(natm),
!        A(3,3), nsteps, then nsteps frames of rx, ry, rz (natm each)
! output (stream): per frame fx, fy, fz, mu_t(natm,3), ef_t(natm,3),
!        efg_t(natm,3,3) and the mdffgpu_energies struct
program engforce_drive
  use iso_c_binding
  use mdffgpu_mod
  implicit none
  type(mdffgpu_params) :: prm
  type(mdffgpu_energies) :: en
  type(c_ptr) :: ctx
  integer(c_int32_t), allocatable :: itype(:)
  real(c_double) :: A(3, 3)
  real(c_double), allocatable :: rx(:), ry(:), rz(:), fx(:), fy(:), fz(:)
  real(c_double), allocatable :: mu_t(:, :), theta_t(:, :, :), ef_t(:, :), efg_t(:, :, :)
  integer(c_int32_t) :: nsteps
  integer :: natm, istep
  character(len=512) :: fin, fout
  integer :: uin, uout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  read (uin) prm
  natm = prm%natm
  allocate (itype(natm), rx(natm), ry(natm), rz(natm), fx(natm), fy(natm), fz(natm))
  allocate (mu_t(natm, 3), theta_t(natm, 3, 3), ef_t(natm, 3), efg_t(natm, 3, 3))
  read (uin) itype
  read (uin) A
  read (uin) nsteps

  call mdffgpu_check(mdffgpu_create(prm, itype, A, ctx), c_null_ptr, 'field_init')

  open (newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  do istep = 1, nsteps
    read (uin) rx, ry, rz
    fx = 0.0_c_double
    fy = 0.0_c_double
    fz = 0.0_c_double
    call engforce_driver_gpu(ctx, natm, rx, ry, rz, istep, fx, fy, fz, mu_t, theta_t, ef_t, &
                             efg_t, en)
    write (uout) fx, fy, fz, mu_t, ef_t, efg_t, en
  end do
  close (uout)
  close (uin)
  call mdffgpu_destroy(ctx)
end program engforce_drive
