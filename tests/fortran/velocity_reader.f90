! Reads a 3Ddata/velocity dump with the implied-do read statement of the
! reference's post-processing (post-processing/post_poiseuilleflow.f90:64-68)
! and prints, per component m = 1..3 (u + u0, w, v), the sum, the sum of
! squares and the value at (i, j, k) = (1, 2, 3).  Used by
! tests/test_fortran_driver.py to show that the driver's dump reads back
! through that statement.
!   velocity_reader nnx nny nnz file
program velocity_reader
  implicit none
  integer :: nnx, nny, nnz, nx, ny, nz, i, j, k, m
  character(len=512) :: a
  real*8, allocatable :: ftmp(:,:,:,:)
  call get_command_argument(1, a); read(a, *) nnx
  call get_command_argument(2, a); read(a, *) nny
  call get_command_argument(3, a); read(a, *) nnz
  nx = nnx - 1; ny = nny - 1; nz = nnz
  allocate(ftmp(0:nx, 0:ny, 0:nz, 3))
  call get_command_argument(4, a)
  open(12, file=trim(a), status='old', form='unformatted')
  read(12) ((((ftmp(i,j,k,m),i=0,nx),j=0,ny),k=0,nz),m=1,3)
  close(12)
  do m = 1, 3
    write(*, '(i2,3es26.17)') m, sum(ftmp(:,:,:,m)), sum(ftmp(:,:,:,m)**2), ftmp(1,2,3,m)
  end do
end program velocity_reader
