! GPU leaf of the MCTDHX method: fabricating it binds the reference's Method and Propagator
! procedure pointers to the MI355X implementation behind the C ABI (include/cody_mctdhx.h).
! The input mirrors the JSON keys the reference reads for this system
! (grid.linear.fedvr, sysKinetic, sysPotential.linear.softYukawa,
! sysInteraction.linear.softYukawa, method.mb.orbBased, integratorList.sil).
module M_Method_Mb_OrbBased_MctdhxHip
  use iso_c_binding
  use M_Utils_Types
  implicit none

  type :: T_MctdhxHip_Input
    real(R64) :: xmin = -10.0_R64, xmax = 10.0_R64
    integer(I32) :: nElements = 10, nLocals = 5
    logical :: xminExcludedQ = .false., xmaxExcludedQ = .false.
    real(R64) :: bodyMass = 1.0_R64
    real(R64) :: position(1) = [0.0_R64], charge(1) = [1.0_R64]
    real(R64) :: softening1(1) = [1.0_R64], softening2(1) = [0.0_R64], dampening(1) = [0.0_R64]
    real(R64) :: strength = 1.0_R64, intSoftening1 = 1.0_R64, intSoftening2 = 0.0_R64
    real(R64) :: intDampening = 0.0_R64
    integer(I32) :: nBodies = 2, nOrbs = 4
    real(R64) :: regularizationParameter = 1.0e-10_R64
    integer(I32) :: krylovDim = 32, maxRestarts = 8
    real(R64) :: tolerance = 1.0e-10_R64
    logical :: metricQ = .true.
    !> .true.: the split step's nonlinear part is the full right-hand side minus the linear part
    !> (second order in dt); .false. (default): the reference's TimeDerivativeCoeffsPlusOrbsNonLin
    logical :: consistentSplitQ = .false.
    !> precision of the mean-field contraction (CODY_PREC_* of M_CodyHip): FP64 by default; the
    !> int8-sliced and BF16-sliced forms give W within 1e-10 of FP64 (the adjoint needs FP64)
    integer(I32) :: contractionPrecision = 0
    integer(I32) :: device = 0
  end type

  type(T_MctdhxHip_Input), target :: MctdhxHip_input
  type(c_ptr) :: MctdhxHip_ctx = c_null_ptr
  integer(I32) :: MctdhxHip_nCoeffs = 0
  integer(I32) :: MctdhxHip_nState = 0
  real(R64) :: MctdhxHip_laserBound = 0.0_R64  !< sum |c_k| of the laser coefficients: a bound of |E(t)|

  interface
    !> Binds Method_Setup, Method_TimeDerivative and Method_GetEnergy, and the orbital-based
    !> method's pointers (the split-step vector fields Method_Mb_OrbBased_TimeDerivativeOrbsLin and
    !> Method_Mb_OrbBased_TimeDerivativeCoeffsPlusOrbsNonLin, ApplySingleBodyOp, ApplyInteractionOp,
    !> ApplyCorrelationOp) to the GPU leaf: the pointers S_Method_Mb_OrbBased.f90:85-109 binds.
    module subroutine Method_Mb_OrbBased_MctdhxHip_Fabricate(input)
      type(T_MctdhxHip_Input), intent(in) :: input
    end subroutine
    !> Method_TimeDerivative (S_Method_Mb_OrbBased_Mctdhx.f90:101-149) on the GPU, host arrays.
    module subroutine MctdhxHip_TimeDerivative(dState, state, time)
      complex(R64), intent(out), contiguous, target :: dState(:)
      complex(R64), intent(in), contiguous, target :: state(:)
      real(R64), intent(in) :: time
    end subroutine
    !> The split parts of the GPU leaf: TimeDerivativeOrbsLin (S_Method_Mb_OrbBased_Mctdhx.f90:152-177)
    !> and TimeDerivativeCoeffsPlusOrbsNonLin (:180-226; with consistentSplitQ, the full right-hand
    !> side minus the linear part).
    module subroutine MctdhxHip_TimeDerivativeOrbsLin(dState, state, time)
      complex(R64), intent(out), contiguous, target :: dState(:)
      complex(R64), intent(in), contiguous, target :: state(:)
      real(R64), intent(in) :: time
    end subroutine
    module subroutine MctdhxHip_TimeDerivativeCoeffsPlusOrbsNonLin(dState, state, time)
      complex(R64), intent(out), contiguous, target :: dState(:)
      complex(R64), intent(in), contiguous, target :: state(:)
      real(R64), intent(in) :: time
    end subroutine
    !> ApplySingleBodyOp (S_Method_Mb_OrbBased.f90:976-1034): dOrbs += (T + V + E(t) x) orbs and
    !> h1_(i, j) = <phi_i|dOrbs_j> of the accumulated dOrbs.
    module subroutine MctdhxHip_ApplySingleBodyOp(dOrbs, orbs, time, h1_)
      complex(R64), intent(inout), contiguous, target :: dOrbs(:, :)
      complex(R64), intent(in), contiguous, target :: orbs(:, :)
      real(R64), intent(in) :: time
      complex(R64), intent(out), allocatable, optional :: h1_(:, :)
    end subroutine
    !> ApplyInteractionOp (S_Method_Mb_OrbBased.f90:491-542): dOrbs(:, i) += W orbs(:, i) for every
    !> orbital i, W the interaction potential of conj(phi_i2) phi_j2 w.
    module subroutine MctdhxHip_ApplyInteractionOp(dOrbs, orbs, i2, j2, time)
      complex(R64), intent(inout), contiguous :: dOrbs(:, :)
      complex(R64), intent(in), contiguous :: orbs(:, :)
      integer(I32), intent(in) :: i2
      integer(I32), intent(in) :: j2
      real(R64), intent(in) :: time
    end subroutine
    !> Binds Propagator_Setup / Propagator_Propagate to the device split step order 2
    !> (SIL on the linear part, RK4 on the nonlinear part).
    module subroutine Propagator_SplitStepHip_Fabricate
    end subroutine
    !> Releases the device context.
    module subroutine Method_Mb_OrbBased_MctdhxHip_Finalize
    end subroutine
    !> Dipole <x> = Re sum_ij rho1(i,j) <phi_j|x phi_i> of Method_state on the GPU: the value of
    !> PrinterObservableLinear_DumpDipole (src/Utils/M_Utils_PrinterObservableLinear.f90:25-70).
    module function Method_Mb_OrbBased_MctdhxHip_Dipole() result(res)
      real(R64) :: res
    end function
    !> Pieces with the abstract interfaces of the reference's procedure pointers, so that
    !> SysInteraction_FillInteractionPotential (src/SysInteraction/M_SysInteraction.f90:94-110),
    !> SysKinetic_MultiplyWithKineticOp (src/SysKinetic/M_SysKinetic.f90:141-160),
    !> Orbs_ProjectOnSubspace (src/Orbs/M_Orbs.f90:132-153) and
    !> Method_Mb_OrbBased_ApplyCorrelationOp (src/Method/Mb/OrbBased/M_Method_Mb_OrbBased.f90:211-231)
    !> can point at the GPU (one body type: the optional body-type arguments must be 1 or absent).
    module subroutine MctdhxHip_FillInteractionPotential(interactionPotential, src, time, bt1_, bt2_)
      complex(R64), intent(out), allocatable :: interactionPotential(:)
      complex(R64), intent(in), contiguous :: src(:)
      real(R64), intent(in) :: time
      integer(I32), intent(in), optional :: bt1_
      integer(I32), intent(in), optional :: bt2_
    end subroutine
    module subroutine MctdhxHip_MultiplyWithKineticOp(dOrb, orb, time, bt_)
      complex(R64), intent(out), contiguous :: dOrb(:)
      complex(R64), intent(in), contiguous :: orb(:)
      real(R64), intent(in) :: time
      integer(I32), intent(in), optional :: bt_
    end subroutine
    module subroutine MctdhxHip_ProjectOnSubspace(dOrbs, orbs)
      complex(R64), intent(inout), contiguous :: dOrbs(:, :)
      complex(R64), intent(in), contiguous :: orbs(:, :)
    end subroutine
    module subroutine MctdhxHip_ApplyCorrelationOp(dOrbs, orbs, rdm1, rdm2, time, h2_)
      complex(R64), intent(inout), contiguous, target :: dOrbs(:, :)
      complex(R64), intent(in), contiguous, target :: orbs(:, :)
      complex(R64), intent(in), contiguous :: rdm1(:, :)
      complex(R64), intent(in), contiguous :: rdm2(:, :, :, :)
      real(R64), intent(in) :: time
      complex(R64), intent(out), allocatable, optional :: h2_(:, :, :, :)
    end subroutine
    !> rho1, rho2 of the CI part of Method_state (the densities ApplyCorrelationOp takes).
    module subroutine MctdhxHip_Densities(rdm1, rdm2)
      complex(R64), intent(out) :: rdm1(:, :)
      complex(R64), intent(out) :: rdm2(:, :, :, :)
    end subroutine
    !> Length-gauge dipole coupling E(t) x with E(t) = sum_k c_k sin(k pi t / T) sin^2(pi t / T)
    !> for 0 <= t <= T (zero outside): the 1D counterpart of the time-dependent field of
    !> src/SysPotential/Ylm/CoulombLaser (SysPotential_timeIndependentQ = .false.).
    module subroutine Method_Mb_OrbBased_MctdhxHip_SetLaser(coefficients, period)
      real(R64), intent(in) :: coefficients(:)
      real(R64), intent(in) :: period
    end subroutine
    !> <x> after nSteps Strang steps of dt from Method_state (left at t0 + nSteps dt), the linear
    !> half-steps a fixed-degree Chebyshev expansion of exp(-i dt/2 h(E)), and its gradient in the
    !> laser coefficients by the discrete adjoint with a checkpoint every checkpointEvery steps.
    !> No counterpart in the reference, which has no adjoint.
    module function Method_Mb_OrbBased_MctdhxHip_DipoleGradient(t0, dt, nSteps, checkpointEvery, gradient, &
                                                                chebLo, chebHi, chebDegree) result(res)
      real(R64), intent(in) :: t0, dt
      integer(I32), intent(in) :: nSteps, checkpointEvery
      real(R64), intent(out) :: gradient(:)
      real(R64), intent(out), optional :: chebLo, chebHi
      integer(I32), intent(out), optional :: chebDegree
      real(R64) :: res
    end function
  end interface

end module
