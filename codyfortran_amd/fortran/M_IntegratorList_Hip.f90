! GPU integrator leaves of the integrator list, extending T_IntegratorList_E
! (src/IntegratorList/M_IntegratorList.f90:98-133):
!   silHip: short iterative Lanczos (the role of src/IntegratorList/Sil, S_IntegratorList_Sil.f90:114-312)
!           on the device, for the linear split part;
!   rk4Hip: classical RK4 (the role of rk.o4Expl, S_IntegratorList_Rk_O4Expl.f90:90-141) on the
!           device, for the nonlinear split part or the full right-hand side.
! A device integrator cannot call a host callback: each leaf integrates the vector field of the
! GPU MCTDHX leaf that its TimeDerivative points at, and stops with an error for any other callback.
! Each Integrate copies the state to the device and back.
module M_IntegratorList_Hip
  use M_Utils_Types
  use M_IntegratorList, only: T_IntegratorList_E
  implicit none

  type, extends(T_IntegratorList_E) :: T_IntegratorList_E_SilHip
    integer(I32) :: krylovDim = 32
    real(R64) :: tolerance = 1.0e-10_R64
    integer(I32) :: maxRestarts = 8
  contains
    procedure :: Fabricate => SilHip_Fabricate
    procedure :: Setup => SilHip_Setup
    procedure :: Integrate => SilHip_Integrate
  end type

  type, extends(T_IntegratorList_E) :: T_IntegratorList_E_Rk4Hip
  contains
    procedure :: Fabricate => Rk4Hip_Fabricate
    procedure :: Setup => Rk4Hip_Setup
    procedure :: Integrate => Rk4Hip_Integrate
  end type

  interface
    !> Allocates a silHip element (the counterpart of IntegratorList_Sil_Allocate).
    module subroutine IntegratorList_SilHip_Allocate(e, path)
      class(T_IntegratorList_E), allocatable, intent(out) :: e
      character(len=*), intent(in) :: path
    end subroutine
    !> Allocates an rk4Hip element.
    module subroutine IntegratorList_Rk4Hip_Allocate(e, path)
      class(T_IntegratorList_E), allocatable, intent(out) :: e
      character(len=*), intent(in) :: path
    end subroutine
    !> Reads krylovDim, tolerance and maxRestarts (the keys of integratorList.sil) from the leaf's input.
    module subroutine SilHip_Fabricate(this)
      class(T_IntegratorList_E_SilHip), intent(inout) :: this
    end subroutine
    module subroutine SilHip_Setup(this)
      class(T_IntegratorList_E_SilHip), intent(inout) :: this
    end subroutine
    !> Lanczos propagation of the linear part from t0 to t1 with step halving; metricWeights, when
    !> allocated, must be the state metric of the FEDVR grid (1 on the coefficients, the grid
    !> weights on every orbital block), the one metric the device integrator holds.
    module subroutine SilHip_Integrate(this, state, t0, t1)
      class(T_IntegratorList_E_SilHip), intent(inout) :: this
      complex(R64), intent(inout), contiguous :: state(:)
      real(R64), intent(in) :: t0
      real(R64), intent(in) :: t1
    end subroutine
    module subroutine Rk4Hip_Fabricate(this)
      class(T_IntegratorList_E_Rk4Hip), intent(inout) :: this
    end subroutine
    module subroutine Rk4Hip_Setup(this)
      class(T_IntegratorList_E_Rk4Hip), intent(inout) :: this
    end subroutine
    !> One RK4 step of t1 - t0 on the nonlinear part or the full right-hand side.
    module subroutine Rk4Hip_Integrate(this, state, t0, t1)
      class(T_IntegratorList_E_Rk4Hip), intent(inout) :: this
      complex(R64), intent(inout), contiguous :: state(:)
      real(R64), intent(in) :: t0
      real(R64), intent(in) :: t1
    end subroutine
  end interface

end module
