! Propagator interface (the reference's src/Propagator/M_Propagator.f90:64-99).
module M_Propagator
  use M_Utils_Types
  implicit none

  procedure(I_Propagator_Setup), pointer :: Propagator_Setup => null()
  abstract interface
    subroutine I_Propagator_Setup
    end subroutine
  end interface

  !> In-place propagation state(t0) -> state(t1).
  procedure(I_Propagator_Propagate), pointer :: Propagator_Propagate => null()
  abstract interface
    subroutine I_Propagator_Propagate(state, t0, t1)
      import :: R64
      complex(R64), intent(inout), contiguous :: state(:)
      real(R64), intent(in) :: t0
      real(R64), intent(in) :: t1
    end subroutine
  end interface

end module
