! Time loop of poiseuilleflow (PF:220-532) on one MI355X through libchf.so.
!
!   poiseuilleflow_gpu [parameters.nml]
!
! The loop keeps the reference's order: the first iteration only solves (the
! state read by `initial` already holds h^n, PF:222); every later iteration
! runs getnonlinear, writes the restart record every `interupt` steps
! (output, PF:2460-2504) and then the per-mode and mean-mode solves
! (PF:242-474).  The restart files use the reference's own record layout and
! name, data00000002_instant.000.dat (PF:2466-2480), so either program can
! continue the other's run.  Every `outputstep` steps, `statistics` writes the
! reference's output files under output_dir (PF:2666-3198): the physical
! velocity dump 3Ddata/velocity<istep>.000.dat and energy.plt, and at steps
! that are also multiples of `interupt` the meanu/, stress/, rms/ and uplus/
! profiles and diagnostics.plt.  The directories are created when missing.
program poiseuilleflow_gpu
  use, intrinsic :: iso_c_binding
  use paraandcons
  use chf_binding
  implicit none

  type(chf_params) :: prm
  type(chf_state) :: st
  type(c_ptr) :: ctx
  character(len=512) :: arg
  integer :: istep, nsolve
  integer(8) :: c0, c1, crate
  real*8 :: secs

  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    call read_parameters(trim(arg))
  end if
  call allocate_state()
  call make_output_dirs()

  prm%nnx = nnx; prm%nny = nny; prm%nnz = nnz
  prm%aphi = aphi; prm%beta = beta; prm%dt = dt; prm%re = re; prm%re_p = re_p
  prm%precision = precision
  ctx = c_null_ptr
  call chf_check(chf_create(prm, ctx), ctx, 'chf_create')

  st%clamb1n = c_loc(clamb1n); st%clamb1nm1 = c_loc(clamb1nm1)
  st%clamb3n = c_loc(clamb3n); st%clamb3nm1 = c_loc(clamb3nm1)
  st%au0 = c_loc(au0); st%w0 = c_loc(w0)
  st%vnp1 = c_loc(vnp1); st%omegaynp1 = c_loc(omegaynp1)
  st%rhsvn = c_loc(rhsvn); st%rhsomegayn = c_loc(rhsomegayn)
  st%unp1 = c_loc(unp1); st%wnp1 = c_loc(wnp1)

  if (isign == 1) then
    call read_restart(trim(restart_in))
    call chf_check(chf_set_state(ctx, st), ctx, 'chf_set_state')
  else
    call laminar_start()
  end if
  write(*, '(a,i0,a,i0,a,i0,a,f8.3,a)') 'channel ', nnx, ' x ', nnz + 1, ' x ', nny, &
      ', device ', dble(chf_device_bytes(ctx)) / 1.0d9, ' GB'

  nsolve = 0
  call system_clock(c0, crate)
  do istep = istepbegin, nstep                  ! PF:220
    if (istep /= istepbegin) then
      if (mod(istep, outputstep) == 0) call statistics(istep)   ! PF:223
      call chf_check(chf_nonlinear(ctx), ctx, 'chf_nonlinear')   ! PF:224
      if (mod(istep, interupt) == 0) then        ! output, PF:225 / 2465
        call chf_check(chf_get_state(ctx, st), ctx, 'chf_get_state')
        call write_restart(istep)
      end if
    end if
    call chf_check(chf_solve(ctx), ctx, 'chf_solve')            ! PF:242-474
    nsolve = nsolve + 1
  end do
  call chf_check(chf_sync(ctx), ctx, 'chf_sync')
  call system_clock(c1)
  secs = dble(c1 - c0) / dble(crate)
  write(*, '(a,i0,a,f10.4,a,f10.3,a)') 'steps ', nsolve, ', wall ', secs, ' s, ', &
      dble(nsolve) / max(secs, 1.0d-12), ' steps/s'
  call chf_check(chf_destroy(ctx), c_null_ptr, 'chf_destroy')

contains

  ! initial(isign=1), PF:1861-1924: two unformatted records
  subroutine read_restart(fname)
    character(len=*), intent(in) :: fname
    integer :: u, ios
    open(newunit=u, file=fname, status='old', form='unformatted', action='read', iostat=ios)
    if (ios /= 0) then
      write(*, '(a,a)') 'cannot open restart file ', fname
      error stop 1
    end if
    read(u) clamb1n, clamb1nm1, clamb3n, clamb3nm1, au0, w0
    read(u) vnp1, omegaynp1, rhsvn, rhsomegayn, unp1, wnp1
    close(u)
    write(*, '(a,a)') 'read initial data from ', fname
  end subroutine read_restart

  ! output, PF:2460-2504
  subroutine write_restart(istep)
    integer, intent(in) :: istep
    character(len=600) :: fname
    integer :: u
    write(fname, '(a,a,i8.8,a,i3.3,a)') trim(output_dir), '/data', 2, '_instant.', 0, '.dat'
    open(newunit=u, file=trim(fname), status='replace', form='unformatted')
    write(u) clamb1n, clamb1nm1, clamb3n, clamb3nm1, au0, w0
    write(u) vnp1, omegaynp1, rhsvn, rhsomegayn, unp1, wnp1
    close(u)
    write(*, '(a,i0,a,a)') 'step ', istep, ': restart record written to ', trim(fname)
  end subroutine write_restart

  ! Laminar start U = (re re_p / 2)(1 - y^2), v = omega_y = 0: the exact
  ! steady state of the scheme (Chebyshev a0 = U_c/2, a2 = -U_c/2).
  subroutine laminar_start()
    real*8 :: uc
    uc = 0.5d0 * re * re_p
    au0 = 0.0d0
    w0 = 0.0d0
    au0(1, 0) = 0.5d0 * uc
    au0(1, 2) = -0.5d0 * uc
    vnp1 = 0.0d0
    omegaynp1 = 0.0d0
    call chf_check(chf_init_fields(ctx, vnp1, omegaynp1, au0, w0), ctx, 'chf_init_fields')
    call chf_check(chf_get_state(ctx, st), ctx, 'chf_get_state')
  end subroutine laminar_start

  ! output_dir and its 3Ddata, meanu, stress, rms and uplus folders, which the
  ! reference expects to exist (mkdir(2) through the C library; an existing
  ! folder is left as it is)
  subroutine make_output_dirs()
    interface
      integer(c_int) function c_mkdir(path, mode) bind(C, name="mkdir")
        import :: c_int, c_char
        character(kind=c_char), intent(in) :: path(*)
        integer(c_int), value :: mode
      end function c_mkdir
    end interface
    character(len=16), parameter :: sub(5) = [character(len=16) :: '3Ddata', 'meanu', 'stress', 'rms', 'uplus']
    integer :: i, rc
    rc = c_mkdir(trim(output_dir) // c_null_char, int(o'755', c_int))
    do i = 1, 5
      rc = c_mkdir(trim(output_dir) // '/' // trim(sub(i)) // c_null_char, int(o'755', c_int))
    end do
  end subroutine make_output_dirs

  ! statistics (PF:2666-3198) from the GPU's physical fields (chf_physical,
  ! transform(1, .) of PF:2723-2725) and plane reductions (chf_statistics,
  ! PF:2816-3020).  Every output step: the physical velocity (u + u0, w, v)
  ! as one unformatted record of real*8 (0:nx,0:ny,0:nz,3) in
  ! 3Ddata/velocity<istep>.000.dat (PF:2749-2764), and istep with the summed
  ! kinetic energy appended to energy.plt (PF:2766-2787).  At steps that are
  ! also multiples of interupt: the CFL numbers (PF:2816-2934), the profiles
  ! meanu/average_velocity, stress/{Reynolds_shear,viscous,total_shear}_stress,
  ! rms/{u,v,w}rms and uplus/uplus (PF:3073-3159), the printed wall shear
  ! velocities and Reynolds numbers, and diagnostics.plt (PF:3161-3189).  The
  ! reference's base flow u0, du0 is zero (SCHEB, PF:3319-3321), so u0 adds
  ! nothing.
  subroutine statistics(istep)
    integer, intent(in) :: istep
    real*8, allocatable :: ftmp(:,:,:,:)
    real*8 :: ps(0:nz, 12), dudy(0:nz), tau(0:nz), zz(0:nz), dy(0:nz), yplus(0:nz), sumn(0:nz)
    real*8 :: pi, dx, ddz, ddy, cfl_x, cfl_y, cfl_z, ubulk, ut_b, ut_t, sume
    character(len=600) :: fname
    integer :: k, u

    allocate(ftmp(0:nx, 0:ny, 0:nz, 3))
    call chf_check(chf_physical(ctx, 0, ftmp(:, :, :, 1)), ctx, 'chf_physical(u)')
    call chf_check(chf_physical(ctx, 2, ftmp(:, :, :, 2)), ctx, 'chf_physical(w)')
    call chf_check(chf_physical(ctx, 1, ftmp(:, :, :, 3)), ctx, 'chf_physical(v)')
    write(fname, '(a,a,i8.8,a,i3.3,a)') trim(output_dir), '/3Ddata/velocity', istep, '.', 0, '.dat'
    open(newunit=u, file=trim(fname), status='replace', form='unformatted')
    write(u) ftmp
    close(u)
    sume = 0.5d0 * sum(ftmp * ftmp)
    deallocate(ftmp)
    if (interupt >= 10) then
      if (mod(istep, interupt / 10) == 0) then
        write(*, '(a,i8,3x,a,f15.3,3x,a,es16.8)') 'istep=', istep, 'time=', dble(istep) * dt, 'engery=', sume
      end if
    end if
    open(newunit=u, file=trim(output_dir) // '/energy.plt', form='formatted', position='append')
    write(u, *) istep, sume
    close(u)

    if (mod(istep, interupt) /= 0) return
    call chf_check(chf_statistics(ctx, ps), ctx, 'chf_statistics')
    pi = 4.0d0 * atan(1.0d0)
    dx = aphi * 2.0d0 * pi / dble(nnx)
    ddz = beta * 2.0d0 * pi / dble(nny)
    ddy = pi / dble(nz)
    dy = 0.0d0
    do k = 1, nz / 2
      dy(k) = -(cos(dble(k) * ddy) - cos(dble(k - 1) * ddy))
      dy(nz - k) = dy(k)
    end do
    do k = 0, nz
      zz(k) = cos(dble(k) * ddy)
    end do
    cfl_x = maxval(ps(:, 10)) * dt / dx
    cfl_z = maxval(ps(:, 12)) * dt / ddz
    cfl_y = 0.0d0
    do k = 1, nz - 1
      cfl_y = max(cfl_y, ps(k, 11) * dt / dy(k))
    end do
    write(*, '(a,es14.6,a,es14.6,a,es14.6)') 'CFL_X= ', cfl_x, ' CFL_Y= ', cfl_y, ' CFL_Z= ', cfl_z

    sumn = ps(:, 1)                               ! ua + u0
    call derivative_at_points(ps(:, 1), dudy)
    tau = -ps(:, 7) + dudy / re                   ! total shear stress (PF:3086-3088)

    write(fname, '(a,a,i8.8,a)') trim(output_dir), '/meanu/average_velocity', istep, '.plt'
    call write_profile(fname, zz, sumn, nz + 1)
    write(fname, '(a,a,i8.8,a)') trim(output_dir), '/stress/Reynolds_shear_stress', istep, '.plt'
    call write_profile(fname, zz, -ps(:, 7) / tau(nz), nz + 1)
    write(fname, '(a,a,i8.8,a)') trim(output_dir), '/stress/viscous_stress', istep, '.plt'
    call write_profile(fname, zz, (dudy / re) / tau(nz), nz + 1)
    write(fname, '(a,a,i8.8,a)') trim(output_dir), '/stress/total_shear_stress', istep, '.plt'
    call write_profile(fname, zz, tau / tau(nz), nz + 1)
    write(fname, '(a,a,i8.8,a)') trim(output_dir), '/rms/urms', istep, '.plt'
    call write_profile(fname, zz, sqrt(ps(:, 4)) / sqrt(tau(nz)), nz + 1)
    write(fname, '(a,a,i8.8,a)') trim(output_dir), '/rms/vrms', istep, '.plt'
    call write_profile(fname, zz, sqrt(ps(:, 5)) / sqrt(tau(nz)), nz + 1)
    write(fname, '(a,a,i8.8,a)') trim(output_dir), '/rms/wrms', istep, '.plt'
    call write_profile(fname, zz, sqrt(ps(:, 6)) / sqrt(tau(nz)), nz + 1)
    do k = 0, nz / 2
      yplus(k) = re * ((1.0d0 - zz(k)) + (1.0d0 + zz(nz - k))) / 2.0d0 * sqrt(tau(nz))
    end do
    write(fname, '(a,a,i8.8,a)') trim(output_dir), '/uplus/uplus', istep, '.plt'
    call write_profile(fname, yplus, [((sumn(k) + sumn(nz - k)) / 2.0d0 / sqrt(tau(nz)), k = 0, nz / 2)], &
                       nz / 2 + 1)

    ut_b = sqrt(abs(tau(nz)))
    ut_t = sqrt(abs(tau(0)))
    ubulk = 0.0d0
    do k = nz / 2, nz - 1
      ubulk = ubulk + sumn(k) * abs(zz(k + 1) - zz(k))
    end do
    write(*, '(a,i8)') 'istep=', istep
    write(*, '(a,2es14.6)') 'wall shear velocity: ', ut_b, ut_t
    write(*, '(a,es14.6,a,2es14.6)') 'Mean pressure gradient[d<p>/dx]: ', -re_p, &
        ' & The shear-stress gradient: ', -tau(nz), tau(0)
    write(*, '(a,2es14.6)') 'Reynolds number R_t: ', ut_b * re, ut_t * re
    write(*, '(a,es14.6)') 'Reynolds number R_c: ', sumn(nz / 2) * re
    write(*, '(a,es14.6)') 'Reynolds number R_b: ', 2.0d0 * ubulk * re
    open(newunit=u, file=trim(output_dir) // '/diagnostics.plt', form='formatted', position='append')
    write(u, *) istep, ut_b, ut_t, -re_p, -tau(nz), tau(0), ut_b * re, ut_t * re, sumn(nz / 2) * re, &
        2.0d0 * ubulk * re
    close(u)
  end subroutine statistics

  ! one profile file: a line "x(k) f(k)" per point, list-directed as the
  ! reference writes them (PF:3073-3159)
  subroutine write_profile(fname, x, f, n)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: n
    real*8, intent(in) :: x(n), f(n)
    integer :: k, u
    open(newunit=u, file=trim(fname), status='replace')
    do k = 1, n
      write(u, *) x(k), f(k)
    end do
    close(u)
  end subroutine write_profile

  ! d/dy of the degree-nz interpolant of f(y_k) at y_k = cos(pi k/nz): the
  ! chebyshev(-1) -> dz -> chebyshev(1) sequence of PF:3083-3085.
  subroutine derivative_at_points(f, df)
    real*8, intent(in) :: f(0:nz)
    real*8, intent(out) :: df(0:nz)
    real*8 :: a(0:nz), b(0:nz + 2), pi, w, ck
    integer :: j, l
    pi = 4.0d0 * atan(1.0d0)
    do l = 0, nz
      a(l) = 0.0d0
      do j = 0, nz
        w = 1.0d0
        if (j == 0 .or. j == nz) w = 0.5d0
        a(l) = a(l) + w * f(j) * cos(pi * dble(j) * dble(l) / dble(nz))
      end do
      ck = 1.0d0
      if (l == 0 .or. l == nz) ck = 2.0d0
      a(l) = a(l) * 2.0d0 / (dble(nz) * ck)
    end do
    b = 0.0d0
    do j = nz - 1, 0, -1
      b(j) = b(j + 2) + 2.0d0 * dble(j + 1) * a(j + 1)
    end do
    b(0) = 0.5d0 * b(0)
    do j = 0, nz
      df(j) = 0.0d0
      do l = 0, nz
        df(j) = df(j) + b(l) * cos(pi * dble(j) * dble(l) / dble(nz))
      end do
    end do
  end subroutine derivative_at_points

end program poiseuilleflow_gpu
