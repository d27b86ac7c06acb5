! Fabrication of the integrator list (the role of src/IntegratorList/S_IntegratorList.f90:84-110,
! which dispatches on the JSON child names): this mirror holds the GPU leaves only.  Inside
! CodyFortran the two branches go at the top of the reference's dispatch, matched on the whole
! name, since "silHip" contains "sil" and "rk4Hip" contains "rk".
submodule(M_IntegratorList) S_IntegratorList
  use M_IntegratorList_Hip
  implicit none

contains

  module subroutine IntegratorList_Fabricate(input)
    type(T_IntegratorList_FabricateInput), intent(in) :: input(:)
    integer :: i
    character(len=:), allocatable :: name
    if (.not. allocated(IntegratorList_childNames)) error stop "IntegratorList_Fabricate: no element names"
    if (size(input) /= size(IntegratorList_childNames)) &
      error stop "Input size does not match number of integratorList elements"
    if (allocated(integratorList)) deallocate (integratorList)
    allocate (integratorList(size(input)))
    do i = 1, size(input)
      name = trim(IntegratorList_childNames(i))
      select case (name)
      case ("silHip")
        call IntegratorList_SilHip_Allocate(integratorList(i)%e, "integratorList."//name)
      case ("rk4Hip")
        call IntegratorList_Rk4Hip_Allocate(integratorList(i)%e, "integratorList."//name)
      case default
        error stop "integratorList element must be one of: silHip, rk4Hip"
      end select
      integratorList(i)%e%TimeDerivative => input(i)%TimeDerivative
      call integratorList(i)%e%Fabricate()
    end do
    IntegratorList_Setup => SetupElements
  end subroutine

  subroutine SetupElements
    integer :: i
    do i = 1, size(integratorList)
      call integratorList(i)%e%Setup()
    end do
  end subroutine

end submodule
