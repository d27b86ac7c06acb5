! Grid data of the linear FEDVR grid (the reference's Grid_nPoints, src/Grid/M_Grid.f90:70,
! and Grid_Linear_xCoord / Grid_Linear_weights of src/Grid/Linear/M_Grid_Linear.f90), filled by
! the GPU leaf's Setup so a program can build its initial state on the same nodes.
module M_Grid
  use M_Utils_Types
  implicit none
  integer(I32) :: Grid_nPoints = 0
  real(R64), allocatable :: Grid_Linear_xCoord(:)
  real(R64), allocatable :: Grid_Linear_weights(:)

contains

  !> <fConjg|f> = sum conj(fConjg) f w (src/Grid/Linear/S_Grid_Linear.f90:53-61)
  pure function Grid_InnerProduct(fConjg, f) result(res)
    complex(R64), intent(in), contiguous :: fConjg(:)
    complex(R64), intent(in), contiguous :: f(:)
    complex(R64) :: res
    res = sum(conjg(fConjg(:)) * f(:) * Grid_Linear_weights(:))
  end function

end module
