! TEST INFRASTRUCTURE: bind(C) entry points into the reference's own
! Fortran modules, compiled from the reference sources in place (see
! oracle/Makefile).  Used only to pin the C oracle in tests; nothing in the
! product loads it.  Each wrapper calls exactly one reference procedure.
!
! The reference cores cache their factorisations in `save` variables on first
! use (src/diff/diff_pen_core_o6_i_c_p.f90:27-32), so every core binds to the
! first line length it sees; callers run one size per process.
module ref_harness
  use iso_c_binding
  use parameter
  use diff_pen_core_o6_i_c_p, only: dp_i1 => diff_i1
  use diff_pen_core_o6_i_c_o, only: do_i1 => diff_i1
  use diff_pen_core_o4_e_c_p, only: d4p_i1 => diff_i1
  use diff_pen_core_o4_e_c_o, only: d4o_i1 => diff_i1
  use diff_pen_core_o2_e_c_o, only: d2o_i1 => diff_i1
  use diff_pen_core_o4_i_c_p_pade, only: dpp_i1 => diff_i1
  use diff_pen_core_o4_i_c_o_pade, only: dpo_i1 => diff_i1
  use diff_pen_core_o6_i_c_p_omp, only: dpomp_i1 => diff_i1
  use filter_pen_core_o6_i_c_p, only: fp_i1 => filter_i1
  use filter_pen_core_o6_i_c_o, only: fo_i1 => filter_i1
  use ns_rho_rhou_p_trafos, only: give_T, give_mu, give_lambda
  use nonreflecting, only: characteristic_q_k1d, characteristic_qs_k1d
  use sponge_layer_core, only: sponge_rhs, sponge_q, sponge_qs, sponge_rhss
  implicit none
contains

  subroutine ref_init(alpha, gamma, R, W, Pr, mu0, T0, npts) bind(C, name="ref_init")
    real(c_double), value :: alpha, gamma, R, W, Pr, mu0, T0
    integer(c_int), value :: npts
    parameter_array = ''
    call set_parameter(alpha, 'filter.alpha')
    call set_parameter(W, 'material.W')
    call set_parameter(Pr, 'material.Pr')
    call set_parameter(mu0, 'material.mu0')
    call set_parameter(T0, 'material.T0')
    params%material%gamma = gamma
    params%material%R = R
    params%equation%nbr_vars = 5
    params%geom%n1b = npts
    params%geom%n2b = 1
    params%geom%n3b = 1
  end subroutine ref_init

  subroutine ref_diff(scheme, q, n, nlines, dx) bind(C, name="ref_diff")
    integer(c_int), value :: scheme, n, nlines
    real(c_double), intent(inout) :: q(n, nlines)
    real(c_double), value :: dx
    if (scheme == 0) then
      call dp_i1(q, dx)
    else
      call do_i1(q, dx)
    end if
  end subroutine ref_diff

  ! the reference's OpenMP variant of the periodic o6 core, with
  ! parallelism.tnr_omp threads (read once, on its first call)
  subroutine ref_diff_omp(scheme, q, n, nlines, dx) bind(C, name="ref_diff_omp")
    integer(c_int), value :: scheme, n, nlines  ! periodic lines only (scheme 0)
    real(c_double), intent(inout) :: q(n, nlines)
    real(c_double), value :: dx
    call dpomp_i1(q, dx)
  end subroutine ref_diff_omp

  subroutine ref_set_threads(nthreads) bind(C, name="ref_set_threads")
    integer(c_int), value :: nthreads
    integer :: tnr
    tnr = nthreads
    call set_parameter(tnr, 'parallelism.tnr_omp')
  end subroutine ref_set_threads

  ! the 4th-order explicit cores (no saved state)
  subroutine ref_diff_o4(scheme, q, n, nlines, dx) bind(C, name="ref_diff_o4")
    integer(c_int), value :: scheme, n, nlines
    real(c_double), intent(inout) :: q(n, nlines)
    real(c_double), value :: dx
    if (scheme == 0) then
      call d4p_i1(q, dx)
    else
      call d4o_i1(q, dx)
    end if
  end subroutine ref_diff_o4

  ! the 2nd-order explicit open core
  subroutine ref_diff_o2(q, n, nlines, dx) bind(C, name="ref_diff_o2")
    integer(c_int), value :: n, nlines
    real(c_double), intent(inout) :: q(n, nlines)
    real(c_double), value :: dx
    call d2o_i1(q, dx)
  end subroutine ref_diff_o2

  ! the 4th-order Pade cores (factorisation saved on first use, like o6)
  subroutine ref_diff_pade(scheme, q, n, nlines, dx) bind(C, name="ref_diff_pade")
    integer(c_int), value :: scheme, n, nlines
    real(c_double), intent(inout) :: q(n, nlines)
    real(c_double), value :: dx
    if (scheme == 0) then
      call dpp_i1(q, dx)
    else
      call dpo_i1(q, dx)
    end if
  end subroutine ref_diff_pade

  subroutine ref_filter(scheme, q, n, nlines) bind(C, name="ref_filter")
    integer(c_int), value :: scheme, n, nlines
    real(c_double), intent(inout) :: q(n, nlines)
    if (scheme == 0) then
      call fp_i1(q)
    else
      call fo_i1(q)
    end if
  end subroutine ref_filter

  subroutine ref_trafos(q, npts, T, mu, lam) bind(C, name="ref_trafos")
    integer(c_int), value :: npts
    real(c_double), intent(in) :: q(npts, 1, 1, 5)
    real(c_double), intent(out) :: T(npts, 1, 1), mu(npts, 1, 1), lam(npts, 1, 1)
    T = give_T(q)
    mu = give_mu(T)
    lam = give_lambda(mu)
  end subroutine ref_trafos

  subroutine ref_characteristic_q(q, aim, npts, k) bind(C, name="ref_characteristic_q")
    integer(c_int), value :: npts
    real(c_double), intent(inout) :: q(npts, 1, 1, 5)
    real(c_double), intent(in) :: aim(npts, 1, 1, 5)
    real(c_double), intent(in) :: k(3)
    call characteristic_q_k1d(q, aim, k)
  end subroutine ref_characteristic_q

  subroutine ref_characteristic_qs(qs, q, npts, k) bind(C, name="ref_characteristic_qs")
    integer(c_int), value :: npts
    real(c_double), intent(inout) :: qs(npts, 1, 1, 5)
    real(c_double), intent(in) :: q(npts, 1, 1, 5)
    real(c_double), intent(in) :: k(3)
    call characteristic_qs_k1d(qs, q, k)
  end subroutine ref_characteristic_qs

  subroutine ref_sponge(kind, y, shape, q, qref, npts) bind(C, name="ref_sponge")
    ! kind 0: sponge_rhs(y=rhs,shape,q,qref)   1: sponge_q(y=q,shape,qref)
    !      2: sponge_qs(y=qs,shape)            3: sponge_rhss(y=rhss,shape,qs=qref,q)
    integer(c_int), value :: kind, npts
    real(c_double), intent(inout) :: y(npts, 1, 1, 5)
    real(c_double), intent(in) :: shape(npts, 1, 1)
    real(c_double), intent(in) :: q(npts, 1, 1, 5), qref(npts, 1, 1, 5)
    select case (kind)
    case (0)
      call sponge_rhs(y, shape, q, qref)
    case (1)
      call sponge_q(y, shape, qref)
    case (2)
      call sponge_qs(y, shape)
    case (3)
      call sponge_rhss(y, shape, qref, q)
    end select
  end subroutine ref_sponge

end module ref_harness
