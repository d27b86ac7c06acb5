! Fortran driver of the GPU MCTDHX leaf, written like the reference's simulation tests
! (test/simulationTests/He1d/T_He1d_06_TimePropagationMctdhxSil.f90:49-107): fabricate, set up,
! evaluate Method_TimeDerivative and Method_GetEnergy, propagate with Propagator_Propagate.
! Usage: t_mctdhx_hip <initial state file> <output directory>
! State files are raw complex(R64) streams (the format of src/Utils/M_Utils_DataStorage.f90:20-65).
program T_MctdhxHip_Rhs
  use M_Utils_Types
  use M_Method
  use M_Propagator
  use M_Grid
  use M_Method_Mb_OrbBased_MctdhxHip
  implicit none

  type(T_MctdhxHip_Input) :: input
  character(len=1024) :: inFile, outDir, precArg
  complex(R64), allocatable :: dState(:)
  real(R64) :: t, dt, e0, e1, gJ, grad(3), chebLo, chebHi
  integer :: io, iStep
  integer(I32) :: chebK

  call get_command_argument(1, inFile)
  call get_command_argument(2, outDir)
  call get_command_argument(3, precArg)
  if (len_trim(inFile) == 0 .or. len_trim(outDir) == 0) error stop "usage: t_mctdhx_hip <state> <outdir> [precision]"

  ! 1D FEDVR MCTDHF: 100 elements x 5 GLL points on [-100, 100], N=2 fermions in M=4 orbitals
  input%xmin = -100.0_R64
  input%xmax = 100.0_R64
  input%nElements = 100
  input%nLocals = 5
  input%charge = [2.0_R64]
  input%softening1 = [1.0_R64]
  input%softening2 = [0.0_R64]
  input%strength = 1.0_R64
  input%intSoftening1 = 1.0_R64
  input%intSoftening2 = 0.0_R64
  input%nBodies = 2
  input%nOrbs = 4
  input%regularizationParameter = 1.0e-8_R64
  input%krylovDim = 16
  input%tolerance = 1.0e-12_R64
  if (len_trim(precArg) > 0) read (precArg, *) input%contractionPrecision

  call Method_Mb_OrbBased_MctdhxHip_Fabricate(input)
  call Propagator_SplitStepHip_Fabricate
  call Method_Setup
  call Propagator_Setup

  open (newunit=io, file=trim(inFile), access="stream", form="unformatted", status="old", action="read")
  read (io) Method_state
  close (io)

  allocate (dState(size(Method_state)))
  if (input%contractionPrecision == 0) call WritePieces
  call Method_TimeDerivative(dState, Method_state, 0.0_R64)
  open (newunit=io, file=trim(outDir)//"/dstate.bin", access="stream", form="unformatted", status="replace")
  write (io) dState
  close (io)

  e0 = Method_GetEnergy(0.0_R64)
  write (*, '(A,ES24.16)') "energy0 ", e0
  if (input%contractionPrecision /= 0) then  ! (the reduced-precision run: the right-hand side and energy)
    call Method_Mb_OrbBased_MctdhxHip_Finalize
    stop
  end if

  t = 0.0_R64
  dt = 0.01_R64
  do iStep = 1, 5
    call Propagator_Propagate(Method_state, t, t + dt)
    t = t + dt
  end do
  e1 = Method_GetEnergy(t)
  write (*, '(A,ES24.16)') "energy5 ", e1
  write (*, '(A,ES24.16)') "dipole5 ", Method_Mb_OrbBased_MctdhxHip_Dipole()
  write (*, '(A,ES24.16)') "gridsum ", sum(Grid_Linear_weights)
  open (newunit=io, file=trim(outDir)//"/state5.bin", access="stream", form="unformatted", status="replace")
  write (io) Method_state
  close (io)

  ! gradient of <x> after 10 Chebyshev-Strang steps in three laser coefficients, from the initial state
  open (newunit=io, file=trim(inFile), access="stream", form="unformatted", status="old", action="read")
  read (io) Method_state
  close (io)
  call Method_Mb_OrbBased_MctdhxHip_SetLaser([0.4_R64, -0.3_R64, 0.2_R64], 0.5_R64)
  gJ = Method_Mb_OrbBased_MctdhxHip_DipoleGradient(0.0_R64, 0.05_R64, 10, 3, grad, chebLo, chebHi, chebK)
  write (*, '(A,ES24.16)') "gradJ ", gJ
  write (*, '(A,ES24.16)') "grad1 ", grad(1)
  write (*, '(A,ES24.16)') "grad2 ", grad(2)
  write (*, '(A,ES24.16)') "grad3 ", grad(3)
  write (*, '(A,ES24.16)') "chebLo ", chebLo
  write (*, '(A,ES24.16)') "chebHi ", chebHi
  write (*, '(A,I0)') "chebK ", chebK

  call Method_Mb_OrbBased_MctdhxHip_Finalize
  write (*, '(A)') "done"

contains

  !> The operator pieces on the initial state, through the leaf's procedures with the reference's
  !> abstract interfaces: pieces.bin = [T phi_1 ; V src ; (1 - P)(x phi) ; corr(0) ; h2].
  subroutine WritePieces
    integer :: nC, nG, nO, j, u
    complex(R64), allocatable :: orbs(:, :), d(:, :), kin(:), pot(:), rdm1(:, :), rdm2(:, :, :, :), h2(:, :, :, :)
    nC = MctdhxHip_nCoeffs
    nG = Grid_nPoints
    nO = input%nOrbs
    orbs = reshape(Method_state(nC + 1:), [nG, nO])
    allocate (kin(nG), d(nG, nO), rdm1(nO, nO), rdm2(nO, nO, nO, nO))
    call MctdhxHip_MultiplyWithKineticOp(kin, orbs(:, 1), 0.0_R64)
    call MctdhxHip_FillInteractionPotential(pot, conjg(orbs(:, 1)) * orbs(:, 2) * Grid_Linear_weights, 0.0_R64)
    do j = 1, nO
      d(:, j) = Grid_Linear_xCoord * orbs(:, j)
    end do
    call MctdhxHip_ProjectOnSubspace(d, orbs)
    open (newunit=u, file=trim(outDir)//"/pieces.bin", access="stream", form="unformatted", status="replace")
    write (u) kin, pot, d
    call MctdhxHip_Densities(rdm1, rdm2)
    d = (0.0_R64, 0.0_R64)
    call MctdhxHip_ApplyCorrelationOp(d, orbs, rdm1, rdm2, 0.0_R64, h2)
    write (u) d, h2
    close (u)
  end subroutine
end program
