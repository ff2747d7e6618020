! Run parameters and state arrays of the GPU driver, with the names and
! shapes of paraandcons (PC:4-38) at one rank (nproc = 1).  The reference fixes
! them at compile time; here the same names are defaults that a namelist file
! (&channel ... /) can override at run time.
module paraandcons
  implicit none

  integer, parameter :: nproc = 1          ! PC:4 (whole domain on one GPU)
  integer :: nnx = 32, nny = 32, nnz = 32  ! PC:5-7
  integer :: nx, ny, nz                    ! PC:8-10, derived in setup_dims
  integer :: istepbegin = 0                ! PC:15
  integer :: nstep = 100                   ! PC:16
  integer :: interupt = 100                ! PC:17, restart output period
  integer :: outputstep = 1000             ! PC:20, statistics print period
  integer :: isign = 1                     ! PF:30: 1 = read restart, -1 = laminar cold start
  real*8 :: aphi = 1.0d0, beta = 0.5d0     ! PC:22
  real*8 :: dt = 1.0d-3                    ! PC:23
  real*8 :: re = 3000.0d0                  ! PC:24
  real*8 :: re_p = 2.0d0 / 3000.0d0        ! PC:25
  integer :: precision = 0                 ! 0 = FP64, 1 = mixed (no PC counterpart)
  character(len=256) :: restart_in = 'data00000002_instant.000.dat'
  character(len=256) :: output_dir = '.'

  ! state (PC:31-38), reference layout: Re/Im interleaved along dim 1
  real*8, allocatable, target, dimension(:,:,:) :: rhsvn, rhsomegayn, unp1, vnp1, wnp1, omegaynp1
  real*8, allocatable, target, dimension(:,:,:) :: omegaxnp1, omegaznp1
  real*8, allocatable, target, dimension(:,:) :: clamb1nm1, clamb3nm1, clamb1n, clamb3n, au0, w0

  namelist /channel/ nnx, nny, nnz, aphi, beta, dt, re, re_p, istepbegin, nstep, interupt, &
                     outputstep, isign, precision, restart_in, output_dir

contains

  subroutine read_parameters(fname)
    character(len=*), intent(in) :: fname
    integer :: u, ios
    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*, '(a,a)') 'cannot open parameter file ', trim(fname)
      error stop 1
    end if
    read(u, nml=channel, iostat=ios)
    if (ios /= 0) then
      write(*, '(a,a)') 'bad &channel namelist in ', trim(fname)
      error stop 1
    end if
    close(u)
  end subroutine read_parameters

  subroutine setup_dims()
    nx = nnx / nproc - 1   ! PC:8 (Re/Im of nnx/2 streamwise modes)
    ny = nny - 1           ! PC:9
    nz = nnz               ! PC:10
  end subroutine setup_dims

  ! PF:89-97
  subroutine allocate_state()
    call setup_dims()
    allocate(rhsvn(0:nx,0:ny,0:nz), rhsomegayn(0:nx,0:ny,0:nz))
    allocate(unp1(0:nx,0:ny,0:nz), vnp1(0:nx,0:ny,0:nz), wnp1(0:nx,0:ny,0:nz))
    allocate(omegaxnp1(0:nx,0:ny,0:nz), omegaynp1(0:nx,0:ny,0:nz), omegaznp1(0:nx,0:ny,0:nz))
    allocate(clamb1nm1(2,0:nz), clamb3nm1(2,0:nz), clamb1n(2,0:nz), clamb3n(2,0:nz))
    allocate(au0(2,0:nz), w0(2,0:nz))
    rhsvn = 0; rhsomegayn = 0; unp1 = 0; vnp1 = 0; wnp1 = 0
    omegaxnp1 = 0; omegaynp1 = 0; omegaznp1 = 0
    clamb1nm1 = 0; clamb3nm1 = 0; clamb1n = 0; clamb3n = 0; au0 = 0; w0 = 0
  end subroutine allocate_state

end module paraandcons
