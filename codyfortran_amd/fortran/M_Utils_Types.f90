! Kind parameters of the host API (the reference's src/Utils/M_Utils_Types.f90:17-24).
module M_Utils_Types
  use iso_fortran_env, only: int32, int64, real64
  implicit none
  integer, parameter :: I32 = int32
  integer, parameter :: I64 = int64
  integer, parameter :: R64 = real64
end module
