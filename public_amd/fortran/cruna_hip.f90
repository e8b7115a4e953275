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
  integer, parameter, public :: CRUNA_HIP_DIFF_O6_COMPACT = 0, CRUNA_HIP_DIFF_O4_EXPLICIT = 1, &
                                CRUNA_HIP_DIFF_O4_PADE = 2, CRUNA_HIP_DIFF_O2_EXPLICIT = 3
  integer, parameter, public :: CRUNA_HIP_EQ_NS = 0, CRUNA_HIP_EQ_EULER = 1

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
    integer(c_int) :: diff_kind(3)
    integer(c_int) :: wfilter_freq(3)
    integer(c_int) :: equations
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
    integer(c_int) function c_set_kvector(ctx, k) bind(C, name="cruna_hip_set_kvector")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: k(*)
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
    integer(c_int) function c_filter_weighted(ctx, dir, q, nv, weight, strength) &
        bind(C, name="cruna_hip_filter_weighted")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: dir, nv
      real(c_double), intent(inout) :: q(*)
      real(c_double), intent(in) :: weight(*), strength(*)
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
    integer(c_int) function c_set_convergence(ctx, cfreq, ctol) bind(C, name="cruna_hip_set_convergence")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: cfreq
      real(c_double), value :: ctol
    end function
    integer(c_int) function c_get_residual(ctx, res, nt_checked, nt_conv) bind(C, name="cruna_hip_get_residual")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: res
      integer(c_int), intent(out) :: nt_checked, nt_conv
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
    integer(c_int) function c_subsets_alloc(ctx, subsets, steps) bind(C, name="cruna_hip_subsets_alloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: subsets, steps
    end function
    integer(c_int) function c_run_direct_subsets(ctx, subsets, steps) bind(C, name="cruna_hip_run_direct_subsets")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: subsets, steps
    end function
    integer(c_int) function c_run_adjoint_subsets(ctx, subsets, steps) bind(C, name="cruna_hip_run_adjoint_subsets")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: subsets, steps
    end function
    integer(c_int) function c_set_metric(ctx, xm, xj) bind(C, name="cruna_hip_set_metric")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: xm(*), xj(*)
    end function
    integer(c_int) function c_set_penalization(ctx, phi, inv_phi) bind(C, name="cruna_hip_set_penalization")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: phi(*), inv_phi(*)
    end function
    integer(c_int) function c_autotune(ctx, steps, tiles) bind(C, name="cruna_hip_autotune")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: steps
      integer(c_int), intent(out) :: tiles(3)
    end function
    integer(c_int) function c_box_filter_state(ctx, f, qb, rb) bind(C, name="cruna_hip_box_filter_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, qb, rb
      integer(c_int), value :: f
    end function
    integer(c_int) function c_ensemble_create(cfgs, n, ens) bind(C, name="cruna_hip_ensemble_create")
      import :: c_int, c_ptr, cruna_hip_cfg
      type(cruna_hip_cfg), intent(in) :: cfgs(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: ens
    end function
    integer(c_int) function c_ensemble_destroy(ens) bind(C, name="cruna_hip_ensemble_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
    end function
    integer(c_int) function c_ensemble_member(ens, i, ctx) bind(C, name="cruna_hip_ensemble_member")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: i
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function c_ensemble_set_role(ens, i, role, nt) bind(C, name="cruna_hip_ensemble_set_role")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: i, role, nt
    end function
    integer(c_int) function c_ensemble_run(ens, nsteps) bind(C, name="cruna_hip_ensemble_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function c_ensemble_autotune(ens, steps, tiles) bind(C, name="cruna_hip_ensemble_autotune")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens, tiles
      integer(c_int), value :: steps
    end function
    integer(c_int) function c_ensemble_set_batching(ens, bins) bind(C, name="cruna_hip_ensemble_set_batching")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: bins
    end function
    integer(c_int) function c_ensemble_batched_steps(ens, steps) bind(C, name="cruna_hip_ensemble_batched_steps")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), intent(out) :: steps
    end function
    integer(c_int) function c_cnn_create(s, cnn) bind(C, name="cruna_hip_cnn_create")
      import :: c_int, c_ptr
      integer(c_int), value :: s
      type(c_ptr), intent(out) :: cnn
    end function
    integer(c_int) function c_cnn_destroy(cnn) bind(C, name="cruna_hip_cnn_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: cnn
    end function
    integer(c_long) function c_cnn_num_params(cnn) bind(C, name="cruna_hip_cnn_num_params")
      import :: c_long, c_ptr
      type(c_ptr), value :: cnn
    end function
    integer(c_int) function c_cnn_set_params(cnn, params) bind(C, name="cruna_hip_cnn_set_params")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: cnn
      real(c_float), intent(in) :: params(*)
    end function
    integer(c_int) function c_cnn_set_norm(cnn, in_std, out_std, p_ref) bind(C, name="cruna_hip_cnn_set_norm")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: cnn
      real(c_double), intent(in) :: in_std(5), out_std(5)
      real(c_double), value :: p_ref
    end function
    integer(c_int) function c_cnn_set_gain(cnn, gain) bind(C, name="cruna_hip_cnn_set_gain")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: cnn
      real(c_double), intent(in) :: gain(5)
    end function
    integer(c_int) function c_set_closure(ctx, cnn) bind(C, name="cruna_hip_set_closure")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, cnn
    end function
  end interface

  integer, parameter, public :: CRUNA_HIP_FORWARD = 0, CRUNA_HIP_ADJOINT = 1

  public :: cruna_hip_default_cfg
  public :: hip_create, hip_destroy, hip_set_state, hip_get_state, hip_set_qref
  public :: hip_set_sponge_shape, hip_set_kvector, hip_diff, hip_filter, hip_filter_weighted, hip_rhs_direct
  public :: hip_time_stepper_direct, hip_run_direct, hip_synchronize
  public :: hip_set_convergence, hip_get_residual
  public :: hip_rhs_adjoint, hip_time_stepper_adjoint, hip_set_adjoint_state
  public :: hip_get_adjoint_state, hip_trajectory_alloc, hip_trajectory_set
  public :: hip_trajectory_get, hip_trajectory_record, hip_run_adjoint
  public :: hip_set_metric, hip_set_penalization, hip_autotune, hip_box_filter_state
  public :: hip_subsets_alloc, hip_run_direct_subsets, hip_run_adjoint_subsets
  public :: hip_ensemble_create, hip_ensemble_destroy, hip_ensemble_member
  public :: hip_ensemble_set_role, hip_ensemble_run, hip_ensemble_autotune
  public :: hip_ensemble_set_batching, hip_ensemble_batched_steps
  public :: hip_closure_create, hip_closure_destroy, hip_set_closure

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

  ! kvector(n1,n2,n3,3): boundary normals per point for rows 202 / 203 / 205 / 206
  subroutine hip_set_kvector(ctx, k)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), contiguous :: k(:,:,:,:)
    call check(c_set_kvector(ctx, k), "set_kvector")
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

  ! filter_weighted_iN_mpi_cartesian::filter_4d(q, weight, strength) with the
  ! o2_e_c_p core (src/filter/filter_weighted_i1_mpi_cartesian.f90:15-54)
  subroutine hip_filter_weighted(ctx, dir, q, weight, strength)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: dir
    real(c_double), intent(inout), contiguous :: q(:,:,:,:)
    real(c_double), intent(in), contiguous :: weight(:,:,:), strength(:,:,:)
    call check(c_filter_weighted(ctx, int(dir - 1, c_int), q, int(size(q, 4), c_int), weight, strength), &
               "filter_weighted")
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

  ! run.cfreq / run.ctol of the cruna_hpc loop (cruna_hpc.f90:81-82, 185-208)
  subroutine hip_set_convergence(ctx, cfreq, ctol)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: cfreq
    real(c_double), intent(in) :: ctol
    call check(c_set_convergence(ctx, int(cfreq, c_int), ctol), "set_convergence")
  end subroutine

  ! last residual, the step it was computed at, and the step at which the
  ! last hip_run_direct reached run.ctol (0: not reached)
  subroutine hip_get_residual(ctx, res, nt_checked, nt_conv)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: res
    integer, intent(out) :: nt_checked, nt_conv
    integer(c_int) :: a, b
    call check(c_get_residual(ctx, res, a, b), "get_residual")
    nt_checked = a
    nt_conv = b
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

  ! time.subsets with device checkpoints (computation_direct.f90:55-133,
  ! computation_adjoint.f90:45-135): checkpoints instead of the subset cache
  subroutine hip_subsets_alloc(ctx, subsets, steps)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: subsets, steps
    call check(c_subsets_alloc(ctx, int(subsets, c_int), int(steps, c_int)), "subsets_alloc")
  end subroutine

  subroutine hip_run_direct_subsets(ctx, subsets, steps)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: subsets, steps
    call check(c_run_direct_subsets(ctx, int(subsets, c_int), int(steps, c_int)), "run_direct_subsets")
  end subroutine

  subroutine hip_run_adjoint_subsets(ctx, subsets, steps)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: subsets, steps
    call check(c_run_adjoint_subsets(ctx, int(subsets, c_int), int(steps, c_int)), "run_adjoint_subsets")
  end subroutine

  ! deformed grid: X_metric(n1,n2,n3,3,3), X_jacobian(n1,n2,n3) of data_geom
  ! (grid_metric::calc_metric, src/geometry/grid_metric.f90:16-188)
  subroutine hip_set_metric(ctx, X_metric, X_jacobian)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), contiguous :: X_metric(:,:,:,:,:), X_jacobian(:,:,:)
    call check(c_set_metric(ctx, X_metric, X_jacobian), "set_metric")
  end subroutine

  ! volume penalisation: penalization(:,:,:,1) (phi) and (:,:,:,2) (1/phi) of
  ! the case's volume_penalization (src/immersed/volume_penalization.f90)
  subroutine hip_set_penalization(ctx, phi, inv_phi)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), contiguous :: phi(:,:,:), inv_phi(:,:,:)
    call check(c_set_penalization(ctx, phi, inv_phi), "set_penalization")
  end subroutine

  ! lines per work-group per direction, tuned on a copy of the state
  subroutine hip_autotune(ctx, steps, tiles)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: steps
    integer(c_int), intent(out) :: tiles(3)
    call check(c_autotune(ctx, int(steps, c_int), tiles), "autotune")
  end subroutine

  ! f^3 box averages of the device state (and of its direct RHS)
  subroutine hip_box_filter_state(ctx, f, q_box, rhs_box)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: f
    real(c_double), intent(out), contiguous, target :: q_box(:,:,:,:)
    real(c_double), intent(out), contiguous, target, optional :: rhs_box(:,:,:,:)
    type(c_ptr) :: rb
    rb = c_null_ptr
    if (present(rhs_box)) rb = c_loc(rhs_box)
    call check(c_box_filter_state(ctx, int(f, c_int), c_loc(q_box), rb), "box_filter_state")
  end subroutine

  subroutine hip_ensemble_create(cfgs, ens)
    type(cruna_hip_cfg), intent(in) :: cfgs(:)
    type(c_ptr), intent(out) :: ens
    call check(c_ensemble_create(cfgs, int(size(cfgs), c_int), ens), "ensemble_create")
  end subroutine

  subroutine hip_ensemble_destroy(ens)
    type(c_ptr), intent(inout) :: ens
    call check(c_ensemble_destroy(ens), "ensemble_destroy")
    ens = c_null_ptr
  end subroutine

  ! member i (0-based) as a context for hip_set_state, trajectories, ...
  subroutine hip_ensemble_member(ens, i, ctx)
    type(c_ptr), intent(in) :: ens
    integer, intent(in) :: i
    type(c_ptr), intent(out) :: ctx
    call check(c_ensemble_member(ens, int(i, c_int), ctx), "ensemble_member")
  end subroutine

  subroutine hip_ensemble_set_role(ens, i, role, nt)
    type(c_ptr), intent(in) :: ens
    integer, intent(in) :: i, role, nt
    call check(c_ensemble_set_role(ens, int(i, c_int), int(role, c_int), int(nt, c_int)), "ensemble_set_role")
  end subroutine

  subroutine hip_ensemble_run(ens, nsteps)
    type(c_ptr), intent(in) :: ens
    integer, intent(in) :: nsteps
    call check(c_ensemble_run(ens, int(nsteps, c_int)), "ensemble_run")
  end subroutine

  ! batched launches: bins -1 automatic, 1 one launch per sweep kind over all
  ! members, 0 the per-member graphs
  subroutine hip_ensemble_set_batching(ens, bins)
    type(c_ptr), intent(in) :: ens
    integer, intent(in) :: bins
    call check(c_ensemble_set_batching(ens, int(bins, c_int)), "ensemble_set_batching")
  end subroutine

  ! steps of the last run that went through batched launches
  subroutine hip_ensemble_batched_steps(ens, steps)
    type(c_ptr), intent(in) :: ens
    integer, intent(out) :: steps
    integer(c_int) :: s
    call check(c_ensemble_batched_steps(ens, s), "ensemble_batched_steps")
    steps = int(s)
  end subroutine

  ! tiles(3, members) receives every member's lines per work-group
  subroutine hip_ensemble_autotune(ens, steps, tiles)
    type(c_ptr), intent(in) :: ens
    integer, intent(in) :: steps
    integer(c_int), intent(out), contiguous, target :: tiles(:,:)
    call check(c_ensemble_autotune(ens, int(steps, c_int), c_loc(tiles)), "ensemble_autotune")
  end subroutine

  ! learned closure on the periodic S^3 grid from trained parameters
  subroutine hip_closure_create(s, params, in_std, out_std, p_ref, gain, cnn)
    integer, intent(in) :: s
    real(c_float), intent(in) :: params(:)
    real(c_double), intent(in) :: in_std(5), out_std(5), p_ref, gain(5)
    type(c_ptr), intent(out) :: cnn
    call check(c_cnn_create(int(s, c_int), cnn), "cnn_create")
    if (size(params, kind=c_long) /= c_cnn_num_params(cnn)) then
      write(*,*) "error in cruna_hip:closure_create: wrong number of parameters"
      stop 1
    end if
    call check(c_cnn_set_norm(cnn, in_std, out_std, p_ref), "cnn_set_norm")
    call check(c_cnn_set_params(cnn, params), "cnn_set_params")
    call check(c_cnn_set_gain(cnn, gain), "cnn_set_gain")
  end subroutine

  subroutine hip_closure_destroy(cnn)
    type(c_ptr), intent(inout) :: cnn
    call check(c_cnn_destroy(cnn), "cnn_destroy")
    cnn = c_null_ptr
  end subroutine

  ! the closure becomes the force term of the context's direct RHS
  ! (apply_force, src/equations/ns_rho_rhou_p.f90:212); c_null_ptr removes it
  subroutine hip_set_closure(ctx, cnn)
    type(c_ptr), intent(in) :: ctx, cnn
    call check(c_set_closure(ctx, cnn), "set_closure")
  end subroutine

end module cruna_hip
