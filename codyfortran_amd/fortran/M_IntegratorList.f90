! Integrator-list interface (the reference's src/IntegratorList/M_IntegratorList.f90:57-230): the
! abstract integrator element with its Fabricate / Setup / Integrate lifecycle, the fabrication
! input carrying each element's time-derivative callback, and the global registry the split-step
! propagator indexes.  Same names, components and abstract interfaces, so that the GPU integrator
! leaves (M_IntegratorList_Hip) extend the reference's own type inside CodyFortran.
!
! The reference's IntegratorList_Fabricate names the elements from the JSON "integratorList"
! children; this mirror has no JSON reader and takes the same names from IntegratorList_childNames.
module M_IntegratorList
  use M_Utils_Types
  implicit none

  !> One element's construction input: the callback that computes d(state)/dt.
  type :: T_IntegratorList_FabricateInput
    procedure(I_TimeDerivative), pointer, nopass :: TimeDerivative
  end type

  interface
    !> Allocates integratorList(:) from the element names, binds each element's callback and
    !> fabricates it; binds IntegratorList_Setup to the setup of every element.
    module subroutine IntegratorList_Fabricate(input)
      type(T_IntegratorList_FabricateInput), intent(in) :: input(:)
    end subroutine
  end interface

  !> An integrator backend: advances a state from t0 to t1 in place.
  type, abstract :: T_IntegratorList_E
    procedure(I_TimeDerivative), pointer, nopass :: TimeDerivative
    !> configuration path of the element, e.g. "integratorList.silHip"
    character(len=:), allocatable :: path
    !> inner-product weights of metric-aware integrators (size of the state); unallocated: Euclidean
    real(R64), allocatable :: metricWeights(:)
  contains
    procedure(I_Fabricate), deferred :: Fabricate
    procedure(I_Setup), deferred :: Setup
    procedure(I_Integrate), deferred :: Integrate
  end type

  abstract interface
    subroutine I_Fabricate(this)
      import :: T_IntegratorList_E
      class(T_IntegratorList_E), intent(inout) :: this
    end subroutine
    subroutine I_Setup(this)
      import :: T_IntegratorList_E
      class(T_IntegratorList_E), intent(inout) :: this
    end subroutine
    subroutine I_Integrate(this, state, t0, t1)
      import :: R64, T_IntegratorList_E
      class(T_IntegratorList_E), intent(inout) :: this
      complex(R64), intent(inout), contiguous :: state(:)
      real(R64), intent(in) :: t0
      real(R64), intent(in) :: t1
    end subroutine
    subroutine I_TimeDerivative(dState, state, time)
      import :: R64
      complex(R64), intent(out), contiguous, target :: dState(:)
      complex(R64), intent(in), contiguous, target :: state(:)
      real(R64), intent(in) :: time
    end subroutine
  end interface

  !> Wrapper that lets integratorList(:) hold elements of different concrete types.
  type :: T_IntegratorList_Container
    class(T_IntegratorList_E), allocatable :: e
  end type

  type(T_IntegratorList_Container), allocatable :: integratorList(:)

  !> Element names, in order (the children of the reference's JSON "integratorList" node).
  character(len=32), allocatable :: IntegratorList_childNames(:)

  !> Setup of every element; a no-op until IntegratorList_Fabricate has run.
  procedure(I_IntegratorList_Setup), pointer :: IntegratorList_Setup => IntegratorList_NoSetup
  abstract interface
    subroutine I_IntegratorList_Setup
    end subroutine
  end interface

contains

  subroutine IntegratorList_NoSetup
  end subroutine

end module
