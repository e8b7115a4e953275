! Fortran host check of libcruna_hip.so, in the form of the reference's own
! unit test usr/unit_test/diff/computation_direct.f90 (D(0) = 0, D(1) = 0,
! D sin(x_d) = cos(x_d), mean-absolute error per direction), followed by a
! device-resident run driven from Fortran:
!   cruna_hip_check N NSTEPS DT MU0 T0 IN.bin OUT.bin [CLOSURE.bin]
! reads q(N,N,N,5) (raw little-endian float64, Fortran order) from IN.bin,
! runs NSTEPS of the cruna_hpc step-loop body (RK4 -> filters) and writes the
! state to OUT.bin.  Without file arguments only the derivative test runs.
! With CLOSURE.bin (16 float64: in_std(5), out_std(5), p_ref, gain(5); then
! the float32 parameters) the learned closure on the N^3 grid is attached to
! the direct RHS first.
program cruna_hip_check
  use iso_c_binding
  use cruna_hip
  implicit none
  type(cruna_hip_cfg) :: cfg
  type(c_ptr) :: ctx
  integer :: n, nsteps, i, j, k, d, u
  real(c_double) :: dt, mu0, t0, h, pi, err(3)
  real(c_double), allocatable :: y(:,:,:), y_x(:,:,:), x(:), q(:,:,:,:)
  character(len=256) :: arg, fin, fout, fcl
  type(c_ptr) :: cnn
  real(c_double) :: hdr(16)
  real(c_float), allocatable :: params(:)
  integer :: fsize

  pi = 4.0_c_double * atan(1.0_c_double)
  n = 32
  nsteps = 0
  dt = 1.0e-3_c_double
  mu0 = 1.0_c_double / 1600.0_c_double
  t0 = 1.0_c_double
  fin = ''
  fout = ''
  fcl = ''
  cnn = c_null_ptr
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg); read(arg, *) n
  end if
  if (command_argument_count() >= 7) then
    call get_command_argument(2, arg); read(arg, *) nsteps
    call get_command_argument(3, arg); read(arg, *) dt
    call get_command_argument(4, arg); read(arg, *) mu0
    call get_command_argument(5, arg); read(arg, *) t0
    call get_command_argument(6, fin)
    call get_command_argument(7, fout)
  end if
  if (command_argument_count() >= 8) call get_command_argument(8, fcl)

  h = 2.0_c_double * pi / real(n, c_double)
  call cruna_hip_default_cfg(cfg)
  cfg%n = n
  cfg%dx = h
  cfg%scheme = CRUNA_HIP_PERIODIC
  cfg%R = 1.0_c_double
  cfg%W = 1.0_c_double
  cfg%mu0 = mu0
  cfg%T0 = t0
  cfg%dt = dt
  cfg%filter_freq = 1
  call hip_create(cfg, ctx)

  allocate(y(n, n, n), y_x(n, n, n), x(n))
  do i = 1, n
    x(i) = real(i - 1, c_double) * h
  end do

  ! component 1: D(0) = 0
  y = 0.0_c_double
  do d = 1, 3
    call hip_diff(ctx, d, y_x, y, h)
    err(d) = sum(abs(y_x)) / real(n, c_double)**3
  end do
  write(*, '(a, 3es12.4)') " c1, tot_err:", err
  ! component 2: D(1) = 0
  y = 1.0_c_double
  do d = 1, 3
    call hip_diff(ctx, d, y_x, y, h)
    err(d) = sum(abs(y_x)) / real(n, c_double)**3
  end do
  write(*, '(a, 3es12.4)') " c2, tot_err:", err
  ! components 3..5: D_d sin(x_d) = cos(x_d)
  do d = 1, 3
    do k = 1, n
      do j = 1, n
        do i = 1, n
          if (d == 1) y(i, j, k) = sin(x(i))
          if (d == 2) y(i, j, k) = sin(x(j))
          if (d == 3) y(i, j, k) = sin(x(k))
        end do
      end do
    end do
    call hip_diff(ctx, d, y_x, y, h)
    do k = 1, n
      do j = 1, n
        do i = 1, n
          if (d == 1) y(i, j, k) = cos(x(i))
          if (d == 2) y(i, j, k) = cos(x(j))
          if (d == 3) y(i, j, k) = cos(x(k))
        end do
      end do
    end do
    err(d) = sum(abs(y_x - y)) / real(n, c_double)**3
  end do
  write(*, '(a, 3es12.4)') " c3-5, Dx1 Dx2 Dx3 sin -> cos, tot_err:", err

  if (len_trim(fin) > 0) then
    allocate(q(n, n, n, 5))
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) q
    close(u)
    if (len_trim(fcl) > 0) then
      open(newunit=u, file=trim(fcl), access='stream', form='unformatted', status='old')
      inquire(unit=u, size=fsize)
      allocate(params((fsize - 16 * 8) / 4))
      read(u) hdr, params
      close(u)
      call hip_closure_create(n, params, hdr(1:5), hdr(6:10), hdr(11), hdr(12:16), cnn)
      call hip_set_closure(ctx, cnn)
      write(*, '(a, i0, a)') " closure: ", size(params), " parameters"
    end if
    call hip_set_state(ctx, q)
    call hip_run_direct(ctx, 0, nsteps)
    call hip_synchronize(ctx)
    call hip_get_state(ctx, q)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) q
    close(u)
    write(*, '(a, i0, a)') " run_direct: ", nsteps, " steps written"
  end if
  call hip_destroy(ctx)
  if (c_associated(cnn)) call hip_closure_destroy(cnn)
end program cruna_hip_check
