! The reverse direction the way FLASH's get_accel_gas_on_particles would call it: gas cells as
! point masses (m = rho dV, eps = dx) onto particles (eps = 0), one tgrav_eval for all
! particles, then the fused half kick.  Argument: 0 = TGRAV_PLAIN_FP64, 1 = TGRAV_MIXED.
! Prints the inputs and results (S: cells, T: particles and fields, K: kicked velocities).
program reverse_drive
  use iso_c_binding
  use tgrav_c
  implicit none
  integer, parameter :: nc = 600, np = 257
  real(c_double), parameter :: gconst = 4.300917270d-3
  type(c_ptr) :: ctx
  real(c_double) :: xc(nc), yc(nc), zc(nc), mc(nc), ec(nc)
  real(c_double) :: xp(np), yp(np), zp(np), ep(np), gax(np), gay(np), gaz(np), pot(np)
  real(c_double) :: vx(np), vy(np), vz(np), kx(np), ky(np), kz(np), kp(np)
  integer(c_int) :: rc, mode
  integer :: i
  character(len=8) :: arg
  call get_command_argument(1, arg)
  read(arg, *) mode
  do i = 1, nc                      ! cells on a 1/64 pc lattice, masses and widths from a table
    xc(i) = dble(mod(37 * i + 11, 257) - 128) / 64d0
    yc(i) = dble(mod(91 * i + 5, 263) - 131) / 64d0
    zc(i) = dble(mod(53 * i + 29, 251) - 125) / 64d0
    mc(i) = 0.5d0 + dble(mod(i, 7)) * 0.25d0
    ec(i) = dble(1 + mod(i, 3)) / 128d0
  end do
  do i = 1, np
    xp(i) = dble(mod(61 * i + 3, 269) - 134) / 64d0
    yp(i) = dble(mod(17 * i + 41, 241) - 120) / 64d0
    zp(i) = dble(mod(73 * i + 7, 277) - 138) / 64d0
  end do
  ep = 0d0
  rc = tgrav_create(ctx, 0_c_int, c_null_ptr)
  call check(rc, 'tgrav_create')
  rc = tgrav_set_constant(ctx, gconst)
  call check(rc, 'tgrav_set_constant')
  rc = tgrav_eval(ctx, int(np, c_int64_t), ep, xp, yp, zp, gax, gay, gaz, pot, mode, TGRAV_HOST)
  if (rc /= -2) then                ! TGRAV_ENOSOURCES: evaluation before any source set
    write(0, *) 'expected TGRAV_ENOSOURCES, got', rc
    stop 1
  end if
  rc = tgrav_set_sources(ctx, int(nc, c_int64_t), xc, yc, zc, mc, ec, TGRAV_HOST)
  call check(rc, 'tgrav_set_sources')
  rc = tgrav_eval(ctx, -1_c_int64_t, ep, xp, yp, zp, gax, gay, gaz, pot, mode, TGRAV_HOST)
  if (rc /= -1) then                ! TGRAV_EINVAL: negative length
    write(0, *) 'expected TGRAV_EINVAL, got', rc
    stop 1
  end if
  rc = tgrav_eval(ctx, int(np, c_int64_t), ep, xp, yp, zp, gax, gay, gaz, pot, mode, TGRAV_HOST)
  call check(rc, 'tgrav_eval')
  vx = 0d0
  vy = 0d0
  vz = 0d0
  rc = tgrav_eval_kick(ctx, int(np, c_int64_t), ep, xp, yp, zp, 0.25d0, vx, vy, vz, kx, ky, kz, kp, &
                       mode, TGRAV_HOST)
  call check(rc, 'tgrav_eval_kick')
  rc = tgrav_destroy(ctx)
  call check(rc, 'tgrav_destroy')
  do i = 1, nc
    write(*, '(A,5(1X,ES25.17E3))') 'S ', xc(i), yc(i), zc(i), mc(i), ec(i)
  end do
  do i = 1, np
    write(*, '(A,7(1X,ES25.17E3))') 'T ', xp(i), yp(i), zp(i), gax(i), gay(i), gaz(i), pot(i)
  end do
  do i = 1, np
    write(*, '(A,3(1X,ES25.17E3))') 'K ', vx(i), vy(i), vz(i)
  end do
contains
  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write(0, *) what, ' -> ', code, ': ', tgrav_error_text()
      stop 1
    end if
  end subroutine
end program reverse_drive
