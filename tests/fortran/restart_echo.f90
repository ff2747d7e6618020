! Reads a restart file with the statements of initial(isign=1) (PF:1913-1917),
! doubles every value and writes it back with those of output (PF:2476-2479).
! Used by tests/test_restart.py to check the Python record layout against the
! Fortran runtime's own unformatted I/O.
!   restart_echo nnx nny nnz infile outfile
program restart_echo
  implicit none
  integer :: nnx, nny, nnz, nx, ny, nz, u
  character(len=512) :: a
  real*8, allocatable, dimension(:,:,:) :: vnp1, omegaynp1, rhsvn, rhsomegayn, unp1, wnp1
  real*8, allocatable, dimension(:,:) :: clamb1n, clamb1nm1, clamb3n, clamb3nm1, au0, w0
  call get_command_argument(1, a); read(a, *) nnx
  call get_command_argument(2, a); read(a, *) nny
  call get_command_argument(3, a); read(a, *) nnz
  nx = nnx - 1; ny = nny - 1; nz = nnz
  allocate(vnp1(0:nx,0:ny,0:nz), omegaynp1(0:nx,0:ny,0:nz), rhsvn(0:nx,0:ny,0:nz))
  allocate(rhsomegayn(0:nx,0:ny,0:nz), unp1(0:nx,0:ny,0:nz), wnp1(0:nx,0:ny,0:nz))
  allocate(clamb1n(2,0:nz), clamb1nm1(2,0:nz), clamb3n(2,0:nz), clamb3nm1(2,0:nz))
  allocate(au0(2,0:nz), w0(2,0:nz))
  call get_command_argument(4, a)
  open(newunit=u, file=trim(a), status='old', form='unformatted')
  read(u) clamb1n, clamb1nm1, clamb3n, clamb3nm1, au0, w0
  read(u) vnp1, omegaynp1, rhsvn, rhsomegayn, unp1, wnp1
  close(u)
  clamb1n = 2 * clamb1n; clamb1nm1 = 2 * clamb1nm1; clamb3n = 2 * clamb3n
  clamb3nm1 = 2 * clamb3nm1; au0 = 2 * au0; w0 = 2 * w0
  vnp1 = 2 * vnp1; omegaynp1 = 2 * omegaynp1; rhsvn = 2 * rhsvn
  rhsomegayn = 2 * rhsomegayn; unp1 = 2 * unp1; wnp1 = 2 * wnp1
  ! a marker value at a known place: Re of streamwise mode 1, span mode 2, degree 3
  vnp1(2, 2, 3) = 12345.0d0
  call get_command_argument(5, a)
  open(newunit=u, file=trim(a), status='replace', form='unformatted')
  write(u) clamb1n, clamb1nm1, clamb3n, clamb3nm1, au0, w0
  write(u) vnp1, omegaynp1, rhsvn, rhsomegayn, unp1, wnp1
  close(u)
end program restart_echo
