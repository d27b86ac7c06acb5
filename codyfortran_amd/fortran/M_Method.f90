! Method interface: the packed state and the procedure pointers a program calls.
! Same names, abstract interfaces and semantics as the reference's src/Method/M_Method.f90:47-109,
! so a driver written against the reference binds to the GPU leaf by fabricating it.
module M_Method
  use M_Utils_Types
  implicit none

  !> Packed state [C(nC) ; orbs(nG, nO)] (src/Method/M_Method.f90:47).
  complex(R64), allocatable, target :: Method_state(:)

  procedure(I_Method_Setup), pointer :: Method_Setup => null()
  abstract interface
    subroutine I_Method_Setup
    end subroutine
  end interface

  !> dState = -i H(state, t) state; overwrites dState (src/Method/M_Method.f90:82-90).
  procedure(I_Method_TimeDerivative), pointer :: Method_TimeDerivative => null()
  abstract interface
    subroutine I_Method_TimeDerivative(dState, state, time)
      import :: R64
      complex(R64), intent(out), contiguous, target :: dState(:)
      complex(R64), intent(in), contiguous, target :: state(:)
      real(R64), intent(in) :: time
    end subroutine
  end interface

  !> Energy of Method_state at `time` (src/Method/M_Method.f90:94-109).
  procedure(I_Method_GetEnergy), pointer :: Method_GetEnergy => null()
  abstract interface
    function I_Method_GetEnergy(time) result(res)
      import :: R64
      real(R64) :: res
      real(R64), intent(in) :: time
    end function
  end interface

end module
