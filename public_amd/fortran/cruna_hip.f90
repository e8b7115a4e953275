! Fortran binding of libcruna_hip.so (include/cruna_hip.h) for the CRUNA host.
!
! iso_c_binding interfaces to every C entry point, plus checked wrappers that
! follow the reference's error convention: a nonzero status prints
! "error in cruna_hip:<routine>: <message>" and stops, as the reference's
! routines do (e.g. src/boundaries/boundary_conditions.f90:129-131).
! Arrays are the reference's: real(c_double) q(n1,n2,n3,5), contiguous.
module cruna_hip
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: CRUNA_HIP_PERIODIC = 0, CRUNA_HIP_OPEN = 1
  integer, parameter, public :: CRUNA_HIP_VISC_SUTHERLAND = 0, CRUNA_HIP_VISC_POWER = 1
  integer, parameter, public :: CRUNA_HIP_FP64 = 0, CRUNA_HIP_FP32_MIXED = 1
  integer, parameter, public :: CRUNA_HIP_MAX_BC = 32

  ! cruna_hip_cfg (C layout; bc(1:10, r) is row r of boundary_conditions.dat)
  type, bind(C), public :: cruna_hip_cfg
    integer(c_int) :: n(3)
    real(c_double) :: dx(3)
    integer(c_int) :: scheme(3)
    real(c_double) :: filter_alpha
    real(c_double) :: gamma
    real(c_double) :: R
    real(c_double) :: W
    real(c_double) :: Pr
    real(c_double) :: mu0
    real(c_double) :: T0
    integer(c_int) :: visc_law
    real(c_double) :: visc_exp
    real(c_double) :: dt
    integer(c_int) :: filter_freq(3)
    integer(c_int) :: afilter_freq(3)
    integer(c_int) :: heat_unit_spacing
    integer(c_int) :: precision
    integer(c_int) :: tile_lines(3)
    integer(c_int) :: nbc
    real(c_double) :: bc(10, CRUNA_HIP_MAX_BC)
  end type cruna_hip_cfg

  interface
    subroutine cruna_hip_default_cfg(cfg) bind(C, name="cruna_hip_default_cfg")
      import :: cruna_hip_cfg
      type(cruna_hip_cfg), intent(out) :: cfg
    end subroutine
    integer(c_int) function c_create(cfg, ctx) bind(C, name="cruna_hip_create")
      import :: c_int, c_ptr, cruna_hip_cfg
      type(cruna_hip_cfg), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function c_destroy(ctx) bind(C, name="cruna_hip_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function c_last_error() bind(C, name="cruna_hip_last_error")
      import :: c_ptr
    end function
    integer(c_int) function c_set_state(ctx, q) bind(C, name="cruna_hip_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q(*)
    end function
    integer(c_int) function c_get_state(ctx, q) bind(C, name="cruna_hip_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: q(*)
    end function
    integer(c_int) function c_set_qref(ctx, q) bind(C, name="cruna_hip_set_qref")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q(*)
    end function
    integer(c_int) function c_set_sponge_shape(ctx, s) bind(C, name="cruna_hip_set_sponge_shape")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: s(*)
    end function
    integer(c_int) function c_diff(ctx, dir, q_x, q, h) bind(C, name="cruna_hip_diff")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: dir
      real(c_double), intent(out) :: q_x(*)
      real(c_double), intent(in) :: q(*)
      real(c_double), value :: h
    end function
    integer(c_int) function c_filter(ctx, dir, q, nv) bind(C, name="cruna_hip_filter")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: dir, nv
      real(c_double), intent(inout) :: q(*)
    end function
    integer(c_int) function c_rhs_direct(ctx, rhs, q, t) bind(C, name="cruna_hip_rhs_direct")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rhs(*)
      real(c_double), intent(in) :: q(*)
      real(c_double), value :: t
    end function
    integer(c_int) function c_time_stepper_direct(ctx, q1, q0, t) &
        bind(C, name="cruna_hip_time_stepper_direct")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: q1(*)
      real(c_double), intent(in) :: q0(*)
      real(c_double), value :: t
    end function
    integer(c_int) function c_run_direct(ctx, nt0, nsteps) bind(C, name="cruna_hip_run_direct")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt0, nsteps
    end function
    integer(c_int) function c_synchronize(ctx) bind(C, name="cruna_hip_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_rhs_adjoint(ctx, rhs, qs, q, g, t) bind(C, name="cruna_hip_rhs_adjoint")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rhs(*)
      real(c_double), intent(in) :: qs(*), q(*)
      type(c_ptr), value :: g
      real(c_double), value :: t
    end function
    integer(c_int) function c_time_stepper_adjoint(ctx, qs1, qs0, q0, g, t) &
        bind(C, name="cruna_hip_time_stepper_adjoint")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: qs1(*)
      real(c_double), intent(in) :: qs0(*), q0(*)
      type(c_ptr), value :: g
      real(c_double), value :: t
    end function
    integer(c_int) function c_set_adjoint_state(ctx, qs) bind(C, name="cruna_hip_set_adjoint_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: qs(*)
    end function
    integer(c_int) function c_get_adjoint_state(ctx, qs) bind(C, name="cruna_hip_get_adjoint_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: qs(*)
    end function
    integer(c_int) function c_trajectory_alloc(ctx, nslots) bind(C, name="cruna_hip_trajectory_alloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nslots
    end function
    integer(c_int) function c_trajectory_set(ctx, nt, q) bind(C, name="cruna_hip_trajectory_set")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
      real(c_double), intent(in) :: q(*)
    end function
    integer(c_int) function c_trajectory_get(ctx, nt, q) bind(C, name="cruna_hip_trajectory_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
      real(c_double), intent(out) :: q(*)
    end function
    integer(c_int) function c_trajectory_record(ctx, on) bind(C, name="cruna_hip_trajectory_record")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function c_run_adjoint(ctx, nt_end, nsteps) bind(C, name="cruna_hip_run_adjoint")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt_end, nsteps
    end function
  end interface

  public :: cruna_hip_default_cfg
  public :: hip_create, hip_destroy, hip_set_state, hip_get_state, hip_set_qref
  public :: hip_set_sponge_shape, hip_diff, hip_filter, hip_rhs_direct
  public :: hip_time_stepper_direct, hip_run_direct, hip_synchronize
  public :: hip_rhs_adjoint, hip_time_stepper_adjoint, hip_set_adjoint_state
  public :: hip_get_adjoint_state, hip_trajectory_alloc, hip_trajectory_set
  public :: hip_trajectory_get, hip_trajectory_record, hip_run_adjoint

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    character(len=512) :: text
    if (rc == 0) return
    call c_f_pointer(c_last_error(), msg, [512])
    text = ''
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*,*) "error in cruna_hip:"//where//": "//trim(text)
    stop 1
  end subroutine check

  subroutine hip_create(cfg, ctx)
    type(cruna_hip_cfg), intent(in) :: cfg
    type(c_ptr), intent(out) :: ctx
    call check(c_create(cfg, ctx), "create")
  end subroutine

  subroutine hip_destroy(ctx)
    type(c_ptr), intent(inout) :: ctx
    call check(c_destroy(ctx), "destroy")
    ctx = c_null_ptr
  end subroutine

  subroutine hip_set_state(ctx, q)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), contiguous :: q(:,:,:,:)
    call check(c_set_state(ctx, q), "set_state")
  end subroutine

  subroutine hip_get_state(ctx, q)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), contiguous :: q(:,:,:,:)
    call check(c_get_state(ctx, q), "get_state")
  end subroutine

  subroutine hip_set_qref(ctx, q)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), contiguous :: q(:,:,:,:)
    call check(c_set_qref(ctx, q), "set_qref")
  end subroutine

  subroutine hip_set_sponge_shape(ctx, s)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), contiguous :: s(:,:,:)
    call check(c_set_sponge_shape(ctx, s), "set_sponge_shape")
  end subroutine

  ! diff_3d(q_x, q, h) along direction dir = 1, 2, 3 (h optional, default 1)
  subroutine hip_diff(ctx, dir, q_x, q, h)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: dir
    real(c_double), intent(out), contiguous :: q_x(:,:,:)
    real(c_double), intent(in), contiguous :: q(:,:,:)
    real(c_double), intent(in), optional :: h
    real(c_double) :: hh
    hh = 1.0_c_double
    if (present(h)) hh = h
    call check(c_diff(ctx, int(dir - 1, c_int), q_x, q, hh), "diff")
  end subroutine

  ! filter_4d(q) along direction dir = 1, 2, 3, in place
  subroutine hip_filter(ctx, dir, q)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: dir
    real(c_double), intent(inout), contiguous :: q(:,:,:,:)
    call check(c_filter(ctx, int(dir - 1, c_int), q, int(size(q, 4), c_int)), "filter")
  end subroutine

  subroutine hip_rhs_direct(ctx, rhs, q, t)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), contiguous :: rhs(:,:,:,:)
    real(c_double), intent(in), contiguous :: q(:,:,:,:)
    real(c_double), intent(in) :: t
    call check(c_rhs_direct(ctx, rhs, q, t), "rhs_direct")
  end subroutine

  subroutine hip_time_stepper_direct(ctx, q1, q0, t)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), contiguous :: q1(:,:,:,:)
    real(c_double), intent(in), contiguous :: q0(:,:,:,:)
    real(c_double), intent(in) :: t
    call check(c_time_stepper_direct(ctx, q1, q0, t), "time_stepper_direct")
  end subroutine

  subroutine hip_run_direct(ctx, nt0, nsteps)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nt0, nsteps
    call check(c_run_direct(ctx, int(nt0, c_int), int(nsteps, c_int)), "run_direct")
  end subroutine

  subroutine hip_synchronize(ctx)
    type(c_ptr), intent(in) :: ctx
    call check(c_synchronize(ctx), "synchronize")
  end subroutine

  ! right_hand_side_adjoint_3d(rhs, qs_in, q, g, t); g optional
  subroutine hip_rhs_adjoint(ctx, rhs, qs, q, g, t)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), contiguous :: rhs(:,:,:,:)
    real(c_double), intent(in), contiguous :: qs(:,:,:,:), q(:,:,:,:)
    real(c_double), intent(in), contiguous, target, optional :: g(:,:,:,:)
    real(c_double), intent(in) :: t
    type(c_ptr) :: gp
    gp = c_null_ptr
    if (present(g)) gp = c_loc(g)
    call check(c_rhs_adjoint(ctx, rhs, qs, q, gp, t), "rhs_adjoint")
  end subroutine

  ! time_stepper_adjoint(qs1, qs0, q0, g, t); g optional
  subroutine hip_time_stepper_adjoint(ctx, qs1, qs0, q0, g, t)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), contiguous :: qs1(:,:,:,:)
    real(c_double), intent(in), contiguous :: qs0(:,:,:,:), q0(:,:,:,:)
    real(c_double), intent(in), contiguous, target, optional :: g(:,:,:,:)
    real(c_double), intent(in) :: t
    type(c_ptr) :: gp
    gp = c_null_ptr
    if (present(g)) gp = c_loc(g)
    call check(c_time_stepper_adjoint(ctx, qs1, qs0, q0, gp, t), "time_stepper_adjoint")
  end subroutine

  subroutine hip_set_adjoint_state(ctx, qs)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), contiguous :: qs(:,:,:,:)
    call check(c_set_adjoint_state(ctx, qs), "set_adjoint_state")
  end subroutine

  subroutine hip_get_adjoint_state(ctx, qs)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), contiguous :: qs(:,:,:,:)
    call check(c_get_adjoint_state(ctx, qs), "get_adjoint_state")
  end subroutine

  subroutine hip_trajectory_alloc(ctx, nslots)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nslots
    call check(c_trajectory_alloc(ctx, int(nslots, c_int)), "trajectory_alloc")
  end subroutine

  ! Q(:,:,:,:,nt) <- q
  subroutine hip_trajectory_set(ctx, nt, q)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nt
    real(c_double), intent(in), contiguous :: q(:,:,:,:)
    call check(c_trajectory_set(ctx, int(nt, c_int), q), "trajectory_set")
  end subroutine

  subroutine hip_trajectory_get(ctx, nt, q)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nt
    real(c_double), intent(out), contiguous :: q(:,:,:,:)
    call check(c_trajectory_get(ctx, int(nt, c_int), q), "trajectory_get")
  end subroutine

  subroutine hip_trajectory_record(ctx, on)
    type(c_ptr), intent(in) :: ctx
    logical, intent(in) :: on
    call check(c_trajectory_record(ctx, merge(1_c_int, 0_c_int, on)), "trajectory_record")
  end subroutine

  ! calc_adjoint's loop body for nt = nt_end, nt_end-1, ..., nt_end-nsteps+1
  subroutine hip_run_adjoint(ctx, nt_end, nsteps)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nt_end, nsteps
    call check(c_run_adjoint(ctx, int(nt_end, c_int), int(nsteps, c_int)), "run_adjoint")
  end subroutine

end module cruna_hip
