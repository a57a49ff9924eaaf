! Fortran 2003 interface to the tgrav C ABI (include/tgrav.h): the block a FLASH maintainer adds
! to call the library from interface.F90 (get_accel_gas_on_particles, interface.F90:1130-1216).
module tgrav_c
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: TGRAV_PLAIN_FP64 = 0, TGRAV_MIXED = 1, TGRAV_HOST = 0
  interface
    integer(c_int) function tgrav_create(ctx, device, tune_cache) bind(C, name="tgrav_create")
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
      integer(c_int), value :: device
      type(c_ptr), value :: tune_cache     ! c_null_ptr, or c_loc of a NUL-terminated path
    end function
    integer(c_int) function tgrav_destroy(ctx) bind(C, name="tgrav_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tgrav_set_constant(ctx, g) bind(C, name="tgrav_set_constant")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: g
    end function
    integer(c_int) function tgrav_set_sources(ctx, n, x, y, z, m, eps, kind) &
        bind(C, name="tgrav_set_sources")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: x(*), y(*), z(*), m(*), eps(*)
      integer(c_int), value :: kind
    end function
    integer(c_int) function tgrav_eval(ctx, n, eps, x, y, z, ax, ay, az, phi, mode, kind) &
        bind(C, name="tgrav_eval")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: eps(*), x(*), y(*), z(*)
      real(c_double), intent(out) :: ax(*), ay(*), az(*), phi(*)
      integer(c_int), value :: mode, kind
    end function
    integer(c_int) function tgrav_eval_kick(ctx, n, eps, x, y, z, dt, vx, vy, vz, ax, ay, az, phi, &
        mode, kind) bind(C, name="tgrav_eval_kick")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: eps(*), x(*), y(*), z(*)
      real(c_double), value :: dt
      real(c_double), intent(inout) :: vx(*), vy(*), vz(*)
      real(c_double), intent(out) :: ax(*), ay(*), az(*), phi(*)
      integer(c_int), value :: mode, kind
    end function
    type(c_ptr) function tgrav_last_error() bind(C, name="tgrav_last_error")
      import :: c_ptr
    end function
  end interface
contains
  ! tgrav_last_error() as a Fortran string.
  function tgrav_error_text() result(s)
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: p(:)
    integer :: n
    call c_f_pointer(tgrav_last_error(), p, [512])
    n = 0
    do while (n < 512)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: s)
    s = transfer(p(1:n), s)
  end function
end module tgrav_c
