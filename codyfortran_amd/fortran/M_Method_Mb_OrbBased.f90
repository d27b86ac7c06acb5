! Orbital-based method interface: the procedure pointers through which the reference's split-step
! drivers and operator code reach the method (src/Method/Mb/OrbBased/M_Method_Mb_OrbBased.f90:195-300).
! Same names, abstract interfaces and semantics, so that a program written against the reference
! wires the GPU leaf by fabricating it (test/simulationTests/He1d/T_He1d_06_TimePropagationMctdhxSil.f90:59-61).
module M_Method_Mb_OrbBased
  use M_Utils_Types
  implicit none

  !> dOrbs(:, i) += W[conj(phi_i2) phi_j2] orbs(:, i) for every orbital i (accumulates).
  procedure(I_Method_Mb_OrbBased_ApplyInteractionOp), pointer :: Method_Mb_OrbBased_ApplyInteractionOp => null()
  abstract interface
    subroutine I_Method_Mb_OrbBased_ApplyInteractionOp(dOrbs, orbs, i2, j2, time)
      import :: I32, R64
      complex(R64), intent(inout), contiguous :: dOrbs(:, :)
      complex(R64), intent(in), contiguous :: orbs(:, :)
      integer(I32), intent(in) :: i2
      integer(I32), intent(in) :: j2
      real(R64), intent(in) :: time
    end subroutine
  end interface

  !> dOrbs += the correlation term of the densities rdm1, rdm2; h2_ its two-body elements.
  procedure(I_Method_Mb_OrbBased_ApplyCorrelationOp), pointer :: Method_Mb_OrbBased_ApplyCorrelationOp => null()
  abstract interface
    subroutine I_Method_Mb_OrbBased_ApplyCorrelationOp(dOrbs, orbs, rdm1, rdm2, time, h2_)
      import :: R64
      complex(R64), intent(inout), contiguous, target :: dOrbs(:, :)
      complex(R64), intent(in), contiguous, target :: orbs(:, :)
      complex(R64), intent(in), contiguous :: rdm1(:, :)
      complex(R64), intent(in), contiguous :: rdm2(:, :, :, :)
      real(R64), intent(in) :: time
      complex(R64), intent(out), allocatable, optional :: h2_(:, :, :, :)
    end subroutine
  end interface

  !> dOrbs += (T + V + field) orbs; h1_(i, j) = <phi_i|dOrbs_j> of the accumulated dOrbs.
  procedure(I_Method_Mb_OrbBased_ApplySingleBodyOp), pointer :: Method_Mb_OrbBased_ApplySingleBodyOp => null()
  abstract interface
    subroutine I_Method_Mb_OrbBased_ApplySingleBodyOp(dOrbs, orbs, time, h1_)
      import :: R64
      complex(R64), intent(inout), contiguous, target :: dOrbs(:, :)
      complex(R64), intent(in), contiguous, target :: orbs(:, :)
      real(R64), intent(in) :: time
      complex(R64), intent(out), allocatable, optional :: h1_(:, :)
    end subroutine
  end interface

  !> The linear split part (orbitals under -i h, coefficients fixed); overwrites dState.
  procedure(I_Method_Mb_OrbBased_TimeDerivativeOrbsLin), pointer :: Method_Mb_OrbBased_TimeDerivativeOrbsLin => null()
  abstract interface
    subroutine I_Method_Mb_OrbBased_TimeDerivativeOrbsLin(dState, state, time)
      import :: R64
      complex(R64), intent(out), contiguous, target :: dState(:)
      complex(R64), intent(in), contiguous, target :: state(:)
      real(R64), intent(in) :: time
    end subroutine
  end interface

  !> The nonlinear split part (coefficients and the orbitals' mean-field terms); overwrites dState.
  procedure(I_Method_Mb_OrbBased_TimeDerivativeOrbsLin), pointer :: &
    Method_Mb_OrbBased_TimeDerivativeCoeffsPlusOrbsNonLin => null()

end module
