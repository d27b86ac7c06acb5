! iso_c_binding interfaces of the MI355X MCTDHX library (include/cody_mctdhx.h), in the
! bind(C) pattern the reference uses for its native libraries
! (src/Utils/M_Utils_CombinationGslLib.f90:20-48).  Every function returns 0 on success;
! CodyHip_Check turns a non-zero status into `error stop`, the reference's error mechanism.
module M_CodyHip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CODY_OK = 0
  integer(c_int), parameter :: CODY_PREC_FP64 = 0, CODY_PREC_SPLIT_FP32 = 1, CODY_PREC_BF16X3 = 2
  integer(c_int), parameter :: CODY_PREC_INT8_SLICED = 3, CODY_PREC_BF16_SLICED = 4

  interface
    function cody_last_error() bind(C, name="cody_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function

    function cody_fedvr_npoints(nE, nLoc, xminExcl, xmaxExcl) bind(C, name="cody_fedvr_npoints") result(r)
      import :: c_int
      integer(c_int), value :: nE, nLoc, xminExcl, xmaxExcl
      integer(c_int) :: r
    end function

    function cody_fedvr_grid(xmin, xmax, nE, nLoc, xminExcl, xmaxExcl, x, w) &
        bind(C, name="cody_fedvr_grid") result(r)
      import :: c_int, c_double
      real(c_double), value :: xmin, xmax
      integer(c_int), value :: nE, nLoc, xminExcl, xmaxExcl
      real(c_double) :: x(*), w(*)
      integer(c_int) :: r
    end function

    function cody_create(ctx, device, nG, nBodies, nOrbs, reg) bind(C, name="cody_create") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: ctx
      integer(c_int), value :: device, nG, nBodies, nOrbs
      real(c_double), value :: reg
      integer(c_int) :: r
    end function

    function cody_destroy(ctx) bind(C, name="cody_destroy") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: r
    end function

    function cody_sizes(ctx, nC, nState) bind(C, name="cody_sizes") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: nC, nState
      integer(c_int) :: r
    end function

    function cody_set_grid(ctx, x, w) bind(C, name="cody_set_grid") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: x(*), w(*)
      integer(c_int) :: r
    end function

    function cody_set_kinetic_fedvr(ctx, xmin, xmax, nE, nLoc, xminExcl, xmaxExcl, mass) &
        bind(C, name="cody_set_kinetic_fedvr") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: xmin, xmax, mass
      integer(c_int), value :: nE, nLoc, xminExcl, xmaxExcl
      integer(c_int) :: r
    end function

    function cody_set_potential_softyukawa(ctx, n, pos, q, a1, a2, d) &
        bind(C, name="cody_set_potential_softyukawa") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_double) :: pos(*), q(*), a1(*), a2(*), d(*)
      integer(c_int) :: r
    end function

    function cody_set_interaction_softyukawa(ctx, s, a1, a2, d) &
        bind(C, name="cody_set_interaction_softyukawa") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: s, a1, a2, d
      integer(c_int) :: r
    end function

    function cody_set_laser(ctx, nK, c, T) bind(C, name="cody_set_laser") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nK
      real(c_double) :: c(*)
      real(c_double), value :: T
      integer(c_int) :: r
    end function

    function cody_upload_state(ctx, state) bind(C, name="cody_upload_state") result(r)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: state(*)
      integer(c_int) :: r
    end function

    function cody_download_state(ctx, state) bind(C, name="cody_download_state") result(r)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: state(*)
      integer(c_int) :: r
    end function

    function cody_time_derivative(ctx, state, dstate, t) bind(C, name="cody_time_derivative") result(r)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: state(*), dstate(*)
      real(c_double), value :: t
      integer(c_int) :: r
    end function

    !> The split-step vector fields TimeDerivativeOrbsLin and TimeDerivativeCoeffsPlusOrbsNonLin.
    function cody_time_derivative_orbs_lin(ctx, state, dstate, t) &
        bind(C, name="cody_time_derivative_orbs_lin") result(r)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: state(*), dstate(*)
      real(c_double), value :: t
      integer(c_int) :: r
    end function

    function cody_time_derivative_nonlin(ctx, state, dstate, t) &
        bind(C, name="cody_time_derivative_nonlin") result(r)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: state(*), dstate(*)
      real(c_double), value :: t
      integer(c_int) :: r
    end function

    function cody_energy(ctx, state, t, e) bind(C, name="cody_energy") result(r)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: state(*)
      real(c_double), value :: t
      real(c_double) :: e
      integer(c_int) :: r
    end function

    function cody_dipole(ctx, state, d) bind(C, name="cody_dipole") result(r)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: state(*)
      real(c_double) :: d
      integer(c_int) :: r
    end function

    function cody_set_precision(ctx, mode) bind(C, name="cody_set_precision") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      integer(c_int) :: r
    end function

    function cody_set_split(ctx, mode) bind(C, name="cody_set_split") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      integer(c_int) :: r
    end function

    function cody_set_graphs(ctx, enable) bind(C, name="cody_set_graphs") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      integer(c_int) :: r
    end function

    function cody_autotune(ctx, wsplit) bind(C, name="cody_autotune") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: wsplit
      integer(c_int) :: r
    end function

    !> Independent instances evaluated as one captured graph: ctxs(1:n) are context handles,
    !> dStates/states device pointers, one per instance.
    function cody_batch_create(batch, n, ctxs) bind(C, name="cody_batch_create") result(r)
      import :: c_ptr, c_int
      type(c_ptr) :: batch
      integer(c_int), value :: n
      type(c_ptr) :: ctxs(*)
      integer(c_int) :: r
    end function

    function cody_batch_destroy(batch) bind(C, name="cody_batch_destroy") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: batch
      integer(c_int) :: r
    end function

    function cody_batch_time_derivative_dev(batch, states, dstates, t, stream) &
        bind(C, name="cody_batch_time_derivative_dev") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: batch, stream
      type(c_ptr) :: states(*), dstates(*)
      real(c_double), value :: t
      integer(c_int) :: r
    end function

    function cody_strang(ctx, t0, t1, krylovDim, tol, maxRestarts, useMetric, converged) &
        bind(C, name="cody_strang") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t0, t1, tol
      integer(c_int), value :: krylovDim, maxRestarts, useMetric
      integer(c_int) :: converged
      integer(c_int) :: r
    end function

    !> Short iterative Lanczos on the linear split part and RK4 on the nonlinear part (full /= 0:
    !> on the full right-hand side), both on the context's device state.
    function cody_sil(ctx, t0, t1, krylovDim, tol, maxRestarts, useMetric, converged) &
        bind(C, name="cody_sil") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t0, t1, tol
      integer(c_int), value :: krylovDim, maxRestarts, useMetric
      integer(c_int) :: converged
      integer(c_int) :: r
    end function

    function cody_rk4(ctx, t0, t1, full) bind(C, name="cody_rk4") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t0, t1
      integer(c_int), value :: full
      integer(c_int) :: r
    end function

    !> Operator pieces (the reference's procedure-pointer semantics) and the CI densities.
    function cody_densities(ctx, coeffs, rdm1, rdm2, reginv) bind(C, name="cody_densities") result(r)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: coeffs(*), rdm1(*), rdm2(*), reginv(*)
      integer(c_int) :: r
    end function

    function cody_interaction_potential(ctx, nSrc, src, pot) bind(C, name="cody_interaction_potential") result(r)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: nSrc
      complex(c_double_complex) :: src(*), pot(*)
      integer(c_int) :: r
    end function

    function cody_apply_correlation(ctx, orbs, rdm1, rdm2, dorbs, h2) bind(C, name="cody_apply_correlation") result(r)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: orbs(*), rdm1(*), rdm2(*), dorbs(*), h2(*)
      integer(c_int) :: r
    end function

    function cody_project_orbs(ctx, orbs, dorbs) bind(C, name="cody_project_orbs") result(r)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: orbs(*), dorbs(*)
      integer(c_int) :: r
    end function

    function cody_apply_single_body_acc(ctx, orbs, t, dorbs, h1) bind(C, name="cody_apply_single_body_acc") result(r)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: orbs(*), dorbs(*), h1(*)
      real(c_double), value :: t
      integer(c_int) :: r
    end function

    !> i2, j2 are 0-based orbital indices
    function cody_apply_interaction(ctx, orbs, i2, j2, t, dorbs) bind(C, name="cody_apply_interaction") result(r)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: orbs(*), dorbs(*)
      integer(c_int), value :: i2, j2
      real(c_double), value :: t
      integer(c_int) :: r
    end function

    function cody_apply_kinetic(ctx, orbs, dorbs) bind(C, name="cody_apply_kinetic") result(r)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: orbs(*), dorbs(*)
      integer(c_int) :: r
    end function

    !> Adjoint: Chebyshev half-steps for steps of dt and fields |E| <= emax, one Chebyshev-Strang
    !> step of the device state, and <x> after nSteps such steps with its gradient in the laser
    !> coefficients (checkpoints every `every` steps).
    function cody_set_chebyshev(ctx, dt, emax, lo, hi, degree) bind(C, name="cody_set_chebyshev") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt, emax
      real(c_double) :: lo, hi
      integer(c_int) :: degree
      integer(c_int) :: r
    end function

    function cody_strang_cheb(ctx, t0, t1) bind(C, name="cody_strang_cheb") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t0, t1
      integer(c_int) :: r
    end function

    function cody_dipole_gradient(ctx, t0, dt, nSteps, every, J, gradc) bind(C, name="cody_dipole_gradient") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t0, dt
      integer(c_int), value :: nSteps, every
      real(c_double) :: J
      real(c_double) :: gradc(*)
      integer(c_int) :: r
    end function

    !> Restart files coeffs.in / orbBB_II.in / rdm2BtXX_YY.in in directory dir (a
    !> c_null_char-terminated string; "" is the working directory), the files the reference's
    !> Coeffs_SaveCoeffs, Orbs_SaveOrbs and Coeffs_SaveTwoRdm write.
    function cody_save_state_files(dir, nC, nG, nBodyTypes, nOrbs, state) &
        bind(C, name="cody_save_state_files") result(r)
      import :: c_char, c_int, c_double_complex
      character(kind=c_char) :: dir(*)
      integer(c_int), value :: nC, nG, nBodyTypes
      integer(c_int) :: nOrbs(*)
      complex(c_double_complex) :: state(*)
      integer(c_int) :: r
    end function

    function cody_load_state_files(dir, nC, nG, nBodyTypes, nOrbs, state) &
        bind(C, name="cody_load_state_files") result(r)
      import :: c_char, c_int, c_double_complex
      character(kind=c_char) :: dir(*)
      integer(c_int), value :: nC, nG, nBodyTypes
      integer(c_int) :: nOrbs(*)
      complex(c_double_complex) :: state(*)
      integer(c_int) :: r
    end function

    function cody_save_two_rdm(dir, nBodyTypes, nOrbs, rdm2) bind(C, name="cody_save_two_rdm") result(r)
      import :: c_char, c_int, c_double_complex
      character(kind=c_char) :: dir(*)
      integer(c_int), value :: nBodyTypes
      integer(c_int) :: nOrbs(*)
      complex(c_double_complex) :: rdm2(*)
      integer(c_int) :: r
    end function
  end interface

contains

  !> Turn a non-zero status into `error stop` with the library's message.
  subroutine CodyHip_Check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (status == CODY_OK) return
    text = ""
    call c_f_pointer(cody_last_error(), msg, [512])
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write (*, '(A,I0,A)') "cody error ", status, " in "//where//": "//trim(text)
    error stop "MI355X MCTDHX library call failed"
  end subroutine

end module
