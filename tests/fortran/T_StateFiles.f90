! Byte compatibility of the library's restart files with Fortran stream I/O, the way the
! reference writes and reads them (unformatted, access='stream': src/Utils/M_Utils_DataStorage.f90).
!
! usage: t_state_files DIR    (DIR/f and DIR/g must exist)
!   1. writes coeffs.in and orbBB_II.in into DIR/f with Fortran stream writes, for two body
!      types with 3 and 2 orbitals, and loads them with cody_load_state_files;
!   2. saves the same state into DIR/g with cody_save_state_files and reads every file back
!      with Fortran stream reads.
! Prints "T_StateFiles PASS" when every value matches bit for bit.
program T_StateFiles
  use iso_c_binding
  use M_CodyHip
  implicit none

  integer, parameter :: nC = 6, nG = 7, nBt = 2
  integer(c_int) :: nOrbs(nBt) = [3_c_int, 2_c_int]
  complex(c_double_complex) :: ref(nC + nG * 5), got(nC + nG * 5), buf(nG), cbuf(nC)
  character(len=512) :: base
  character(len=32) :: name
  integer :: k, bt, i, o, io, nBad

  call get_command_argument(1, base)
  do k = 1, size(ref)
    ref(k) = cmplx(real(k, c_double) / 3.0_c_double, -sqrt(real(k, c_double)), kind=c_double)
  end do

  ! 1. Fortran writes, the library reads
  open (newunit=io, file=trim(base)//"/f/coeffs.in", form="unformatted", access="stream", &
        status="replace", action="write")
  write (io) ref(1:nC)
  close (io)
  o = nC
  do bt = 1, nBt
    do i = 1, nOrbs(bt)
      write (name, '("orb",I2.2,"_",I2.2,".in")') bt, i
      open (newunit=io, file=trim(base)//"/f/"//trim(name), form="unformatted", access="stream", &
            status="replace", action="write")
      write (io) ref(o + 1:o + nG)
      close (io)
      o = o + nG
    end do
  end do
  got = (0.0_c_double, 0.0_c_double)
  call CodyHip_Check(cody_load_state_files(trim(base)//"/f"//c_null_char, nC, nG, nBt, nOrbs, got), &
                     "cody_load_state_files")
  nBad = count(got /= ref)

  ! 2. the library writes, Fortran reads
  call CodyHip_Check(cody_save_state_files(trim(base)//"/g"//c_null_char, nC, nG, nBt, nOrbs, ref), &
                     "cody_save_state_files")
  open (newunit=io, file=trim(base)//"/g/coeffs.in", form="unformatted", access="stream", &
        status="old", action="read")
  read (io) cbuf
  close (io)
  nBad = nBad + count(cbuf /= ref(1:nC))
  o = nC
  do bt = 1, nBt
    do i = 1, nOrbs(bt)
      write (name, '("orb",I2.2,"_",I2.2,".in")') bt, i
      open (newunit=io, file=trim(base)//"/g/"//trim(name), form="unformatted", access="stream", &
            status="old", action="read")
      read (io) buf
      close (io)
      nBad = nBad + count(buf /= ref(o + 1:o + nG))
      o = o + nG
    end do
  end do

  if (nBad /= 0) then
    write (*, '(A,I0)') "T_StateFiles mismatching values: ", nBad
    error stop 1
  end if
  write (*, '(A)') "T_StateFiles PASS"
end program
