! Fortran driver of the split-step chain of the GPU leaf, shaped like the reference's
! test/simulationTests/He1d/T_He1d_06_TimePropagationMctdhxSil.f90:49-107: the integrator list is
! fabricated with the method's split-part callbacks (as at :59-61), a SIL element for the linear
! part and an RK4 element for the nonlinear part, and propagated by the order-2 split
! A(dt/2) B(dt) A(dt/2) over integratorList(1:2) (the sequence of S_Propagator_SplitStep_Order2.f90:52-77).
! Usage: t_mctdhx_split <initial state file> <output directory>
!        t_mctdhx_split --foreign-callback   (the SIL element refuses a callback it cannot integrate)
! State files are raw complex(R64) streams (the format of src/Utils/M_Utils_DataStorage.f90:20-65).
program T_MctdhxHip_Split
  use M_Utils_Types
  use M_Method
  use M_Method_Mb_OrbBased
  use M_IntegratorList
  use M_Grid
  use M_Method_Mb_OrbBased_MctdhxHip
  implicit none

  type(T_MctdhxHip_Input) :: input
  type(T_IntegratorList_FabricateInput) :: IntegratorListInput(2)
  character(len=1024) :: inFile, outDir
  complex(R64), allocatable :: dState(:), orbs(:, :), d(:, :), h1(:, :), probe(:)
  real(R64), allocatable :: metricWeights(:)
  real(R64) :: t, dt, startEnergy, energy, startNorm, finalNorm
  integer :: io, iStep, nC, nG, nO, i

  call get_command_argument(1, inFile)
  call get_command_argument(2, outDir)

  ! 1D FEDVR MCTDHF: 100 elements x 5 GLL points on [-100, 100], N=2 fermions in M=4 orbitals
  input%xmin = -100.0_R64
  input%xmax = 100.0_R64
  input%nElements = 100
  input%nLocals = 5
  input%charge = [2.0_R64]
  input%strength = 1.0_R64
  input%nBodies = 2
  input%nOrbs = 4
  input%regularizationParameter = 1.0e-8_R64
  input%krylovDim = 16
  input%tolerance = 1.0e-12_R64

  call Method_Mb_OrbBased_MctdhxHip_Fabricate(input)
  IntegratorList_childNames = [character(len=32) :: "silHip", "rk4Hip"]
  if (trim(inFile) == "--foreign-callback") then
    ! the SIL element is handed the full right-hand side, which the device Lanczos cannot integrate
    IntegratorListInput(1)%TimeDerivative => Method_TimeDerivative
    IntegratorListInput(2)%TimeDerivative => Method_Mb_OrbBased_TimeDerivativeCoeffsPlusOrbsNonLin
    call IntegratorList_Fabricate(IntegratorListInput)
    allocate (probe(8))
    probe = (0.0_R64, 0.0_R64)
    call integratorList(1)%e%Integrate(probe, 0.0_R64, 0.1_R64)
    error stop "the SIL element accepted a foreign callback"
  end if
  if (len_trim(inFile) == 0 .or. len_trim(outDir) == 0) error stop "usage: t_mctdhx_split <state> <outdir>"
  IntegratorListInput(1)%TimeDerivative => Method_Mb_OrbBased_TimeDerivativeOrbsLin
  IntegratorListInput(2)%TimeDerivative => Method_Mb_OrbBased_TimeDerivativeCoeffsPlusOrbsNonLin
  call IntegratorList_Fabricate(IntegratorListInput)

  call Method_Setup
  call IntegratorList_Setup

  open (newunit=io, file=trim(inFile), access="stream", form="unformatted", status="old", action="read")
  read (io) Method_state
  close (io)
  nC = MctdhxHip_nCoeffs
  nG = Grid_nPoints
  nO = input%nOrbs

  ! the split-step vector fields through the method's pointers
  allocate (dState(size(Method_state)))
  open (newunit=io, file=trim(outDir)//"/split_parts.bin", access="stream", form="unformatted", status="replace")
  call Method_Mb_OrbBased_TimeDerivativeOrbsLin(dState, Method_state, 0.0_R64)
  write (io) dState
  call Method_Mb_OrbBased_TimeDerivativeCoeffsPlusOrbsNonLin(dState, Method_state, 0.0_R64)
  write (io) dState

  ! ApplySingleBodyOp and ApplyInteractionOp, accumulating onto x phi
  orbs = reshape(Method_state(nC + 1:), [nG, nO])
  allocate (d(nG, nO))
  do i = 1, nO
    d(:, i) = Grid_Linear_xCoord * orbs(:, i)
  end do
  call Method_Mb_OrbBased_ApplySingleBodyOp(d, orbs, 0.0_R64, h1)
  write (io) d, h1
  do i = 1, nO
    d(:, i) = Grid_Linear_xCoord * orbs(:, i)
  end do
  call Method_Mb_OrbBased_ApplyInteractionOp(d, orbs, 2, 3, 0.0_R64)
  call Method_Mb_OrbBased_ApplyInteractionOp(d, orbs, 4, 1, 0.0_R64)
  write (io) d
  close (io)

  ! the state metric of the FEDVR grid on the SIL element, as the reference's laser drivers set it
  allocate (metricWeights(nC + nG * nO))
  metricWeights(1:nC) = 1.0_R64
  do i = 1, nO
    metricWeights(nC + (i - 1) * nG + 1:nC + i * nG) = Grid_Linear_weights
  end do
  integratorList(1)%e%metricWeights = metricWeights

  startEnergy = Method_GetEnergy(0.0_R64)
  startNorm = sqrt(sum(metricWeights * abs(Method_state)**2))
  t = 0.0_R64
  dt = 0.01_R64
  do iStep = 1, 5
    call PropagateOrder2(Method_state, t, t + dt)
    t = t + dt
  end do
  energy = Method_GetEnergy(t)
  finalNorm = sqrt(sum(metricWeights * abs(Method_state)**2))
  write (*, '(A,ES24.16)') "startEnergy ", startEnergy
  write (*, '(A,ES24.16)') "energy5 ", energy
  write (*, '(A,ES24.16)') "normDrift ", abs(finalNorm - startNorm)
  open (newunit=io, file=trim(outDir)//"/chain5.bin", access="stream", form="unformatted", status="replace")
  write (io) Method_state
  close (io)

  call Method_Mb_OrbBased_MctdhxHip_Finalize
  write (*, '(A)') "done"

contains

  subroutine PropagateOrder2(state, t0, t1)
    complex(R64), intent(inout), contiguous :: state(:)
    real(R64), intent(in) :: t0, t1
    real(R64) :: h
    h = t1 - t0
    call integratorList(1)%e%Integrate(state, t0, t0 + 0.5_R64 * h)
    call integratorList(2)%e%Integrate(state, t0, t0 + h)
    call integratorList(1)%e%Integrate(state, t0 + 0.5_R64 * h, t0 + h)
  end subroutine

end program
