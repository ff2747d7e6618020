! ISO_C_BINDING interface to libchf.so, one interface per entry point of
! include/chf.h.  The types mirror chf_params and chf_state field by field;
! every array argument is the caller's own reference-layout array
! (real*8 (0:nx,0:ny,0:nz) or (2,0:nz)) passed by address.
module chf_binding
  use, intrinsic :: iso_c_binding
  implicit none

  type, bind(C) :: chf_params
    integer(c_int) :: nnx, nny, nnz
    real(c_double) :: aphi, beta, dt, re, re_p
    integer(c_int) :: precision
  end type chf_params

  type, bind(C) :: chf_state
    type(c_ptr) :: clamb1n, clamb1nm1, clamb3n, clamb3nm1, au0, w0
    type(c_ptr) :: vnp1, omegaynp1, rhsvn, rhsomegayn, unp1, wnp1
  end type chf_state

  interface
    ! operator set-up (SCHEB PF:156, get_coef_v / get_coef_omegay + zgetri PF:181-214)
    integer(c_int) function chf_create(prm, ctx) bind(C, name="chf_create")
      import :: c_int, c_ptr, chf_params
      type(chf_params), intent(in) :: prm
      type(c_ptr), intent(out) :: ctx
    end function chf_create

    integer(c_int) function chf_destroy(ctx) bind(C, name="chf_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function chf_destroy

    type(c_ptr) function chf_last_error(ctx) bind(C, name="chf_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function chf_last_error

    ! initial(isign=1) restart read PF:1861-1924 / output PF:2460-2504
    integer(c_int) function chf_set_state(ctx, st) bind(C, name="chf_set_state")
      import :: c_int, c_ptr, chf_state
      type(c_ptr), value :: ctx
      type(chf_state), intent(in) :: st
    end function chf_set_state

    integer(c_int) function chf_get_state(ctx, st) bind(C, name="chf_get_state")
      import :: c_int, c_ptr, chf_state
      type(c_ptr), value :: ctx
      type(chf_state), intent(in) :: st
    end function chf_get_state

    ! cold start: kinematics, h^{-1} = 0, one getnonlinear (PF:1932-1946)
    integer(c_int) function chf_init_fields(ctx, vnp1, omegaynp1, au0, w0) bind(C, name="chf_init_fields")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vnp1(*), omegaynp1(*), au0(*), w0(*)
    end function chf_init_fields

    ! getnonlinear PF:1380-1612
    integer(c_int) function chf_nonlinear(ctx) bind(C, name="chf_nonlinear")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function chf_nonlinear

    ! per-mode solves, kinematics and mean mode PF:242-474
    integer(c_int) function chf_solve(ctx) bind(C, name="chf_solve")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function chf_solve

    ! n iterations of the PF:220-532 loop body
    integer(c_int) function chf_step(ctx, n) bind(C, name="chf_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
    end function chf_step

    integer(c_int) function chf_sync(ctx) bind(C, name="chf_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function chf_sync

    ! post-solve u, w, omega_x, omega_z (PF:387-472); c_null_ptr skips an array
    integer(c_int) function chf_get_fields(ctx, unp1, wnp1, omegaxnp1, omegaznp1) bind(C, name="chf_get_fields")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, unp1, wnp1, omegaxnp1, omegaznp1
    end function chf_get_fields

    ! transform(1, .) of field 0..5 (u, v, w, omega_x, omega_y, omega_z), PF:948-961
    integer(c_int) function chf_physical(ctx, field, out) bind(C, name="chf_physical")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(out) :: out(*)
    end function chf_physical

    ! plane statistics of statistics (PF:2816-3020): out(0:nz, 12)
    integer(c_int) function chf_statistics(ctx, out) bind(C, name="chf_statistics")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
    end function chf_statistics

    integer(c_int) function chf_dims(ctx, mx, mz, ny) bind(C, name="chf_dims")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: mx, mz, ny
    end function chf_dims

    integer(c_long_long) function chf_device_bytes(ctx) bind(C, name="chf_device_bytes")
      import :: c_long_long, c_ptr
      type(c_ptr), value :: ctx
    end function chf_device_bytes

    integer(c_int) function chf_set_timing(ctx, on) bind(C, name="chf_set_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function chf_set_timing

    integer(c_int) function chf_kernel_count() bind(C, name="chf_kernel_count")
      import :: c_int
    end function chf_kernel_count

    type(c_ptr) function chf_kernel_name(id) bind(C, name="chf_kernel_name")
      import :: c_ptr, c_int
      integer(c_int), value :: id
    end function chf_kernel_name

    integer(c_int) function chf_timing(ctx, id, total_ms, launches) bind(C, name="chf_timing")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(out) :: total_ms
      integer(c_int), intent(out) :: launches
    end function chf_timing

    ! autotuner report: chosen column blocks per wave of kernel class id
    integer(c_int) function chf_tile_choice(ctx, id, ms_by_choice) bind(C, name="chf_tile_choice")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(out) :: ms_by_choice(3)
    end function chf_tile_choice

    ! override of the tuned choice (1, 2 or 4); results do not change
    integer(c_int) function chf_set_tile_choice(ctx, id, cb) bind(C, name="chf_set_tile_choice")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: id, cb
    end function chf_set_tile_choice
  end interface

contains

  ! Fortran copy of a NUL-terminated C string
  function c_to_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
      integer(c_size_t) function strlen(q) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: q
      end function strlen
    end interface
    if (.not. c_associated(p)) then
      s = ""
      return
    end if
    n = int(strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    s = transfer(chars(1:n), s)
  end function c_to_string

  ! The reference ignores every status code; this driver stops on the first
  ! failure with the library's message.
  subroutine chf_check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      write(*, '(a,a,i0,a,a)') what, ' failed, status ', rc, ': ', c_to_string(chf_last_error(ctx))
      error stop 1
    end if
  end subroutine chf_check

end module chf_binding
