! Implementation of the GPU MCTDHX leaf.
submodule(M_Method_Mb_OrbBased_MctdhxHip) S_Method_Mb_OrbBased_MctdhxHip
  use M_CodyHip
  implicit none

contains

  module subroutine Method_Mb_OrbBased_MctdhxHip_Fabricate(input)
    use M_Method
    use M_Method_Mb_OrbBased
    type(T_MctdhxHip_Input), intent(in) :: input
    MctdhxHip_input = input
    Method_Setup => Setup
    Method_TimeDerivative => MctdhxHip_TimeDerivative
    Method_GetEnergy => GetEnergy
    Method_Mb_OrbBased_TimeDerivativeOrbsLin => MctdhxHip_TimeDerivativeOrbsLin
    Method_Mb_OrbBased_TimeDerivativeCoeffsPlusOrbsNonLin => MctdhxHip_TimeDerivativeCoeffsPlusOrbsNonLin
    Method_Mb_OrbBased_ApplySingleBodyOp => MctdhxHip_ApplySingleBodyOp
    Method_Mb_OrbBased_ApplyInteractionOp => MctdhxHip_ApplyInteractionOp
    Method_Mb_OrbBased_ApplyCorrelationOp => MctdhxHip_ApplyCorrelationOp
  end subroutine

  module subroutine Propagator_SplitStepHip_Fabricate
    use M_Propagator
    Propagator_Setup => PropagatorSetup
    Propagator_Propagate => Propagate
  end subroutine

  module subroutine Method_Mb_OrbBased_MctdhxHip_Finalize
    integer(c_int) :: r
    if (c_associated(MctdhxHip_ctx)) r = cody_destroy(MctdhxHip_ctx)
    MctdhxHip_ctx = c_null_ptr
  end subroutine

  !> Creates the device context, the FEDVR grid, kinetic energy, potentials and the packed
  !> state (the allocations of Method_Setup, S_Method_Mb_OrbBased_Mctdhx.f90:40-68).
  subroutine Setup
    use M_Method
    use M_Grid
    type(T_MctdhxHip_Input), pointer :: in
    integer(c_int) :: lo, hi, nG, nC, nS
    in => MctdhxHip_input
    lo = merge(1, 0, in%xminExcludedQ)
    hi = merge(1, 0, in%xmaxExcludedQ)
    nG = cody_fedvr_npoints(in%nElements, in%nLocals, lo, hi)
    if (nG < 2) error stop "MctdhxHip Setup: invalid FEDVR layout"
    Grid_nPoints = nG
    if (allocated(Grid_Linear_xCoord)) deallocate (Grid_Linear_xCoord, Grid_Linear_weights)
    allocate (Grid_Linear_xCoord(nG), Grid_Linear_weights(nG))
    call CodyHip_Check(cody_fedvr_grid(in%xmin, in%xmax, in%nElements, in%nLocals, lo, hi, &
                                       Grid_Linear_xCoord, Grid_Linear_weights), "fedvr_grid")
    call CodyHip_Check(cody_create(MctdhxHip_ctx, in%device, nG, in%nBodies, in%nOrbs, &
                                   in%regularizationParameter), "create")
    call CodyHip_Check(cody_sizes(MctdhxHip_ctx, nC, nS), "sizes")
    MctdhxHip_nCoeffs = nC
    MctdhxHip_nState = nS
    call CodyHip_Check(cody_set_grid(MctdhxHip_ctx, Grid_Linear_xCoord, Grid_Linear_weights), "set_grid")
    call CodyHip_Check(cody_set_kinetic_fedvr(MctdhxHip_ctx, in%xmin, in%xmax, in%nElements, in%nLocals, &
                                              lo, hi, in%bodyMass), "set_kinetic_fedvr")
    call CodyHip_Check(cody_set_potential_softyukawa(MctdhxHip_ctx, 1, in%position, in%charge, &
                                                     in%softening1, in%softening2, in%dampening), &
                       "set_potential_softyukawa")
    call CodyHip_Check(cody_set_interaction_softyukawa(MctdhxHip_ctx, in%strength, in%intSoftening1, &
                                                       in%intSoftening2, in%intDampening), &
                       "set_interaction_softyukawa")
    if (in%consistentSplitQ) call CodyHip_Check(cody_set_split(MctdhxHip_ctx, 1_c_int), "set_split")
    if (in%contractionPrecision /= CODY_PREC_FP64) &
      call CodyHip_Check(cody_set_precision(MctdhxHip_ctx, int(in%contractionPrecision, c_int)), "set_precision")
    if (allocated(Method_state)) deallocate (Method_state)
    allocate (Method_state(nS))
    Method_state(:) = (0.0_R64, 0.0_R64)
  end subroutine

  subroutine CheckStateSizes(dState, state, where)
    complex(R64), intent(in) :: dState(:), state(:)
    character(len=*), intent(in) :: where
    if (.not. c_associated(MctdhxHip_ctx)) error stop "MctdhxHip: Method_Setup has not run"
    if (size(state) /= MctdhxHip_nState .or. size(dState) /= MctdhxHip_nState) then
      write (*, '(A)') "MctdhxHip "//where//": state size mismatch"
      error stop "MI355X MCTDHX: state size mismatch"
    end if
  end subroutine

  module subroutine MctdhxHip_TimeDerivative(dState, state, time)
    complex(R64), intent(out), contiguous, target :: dState(:)
    complex(R64), intent(in), contiguous, target :: state(:)
    real(R64), intent(in) :: time
    call CheckStateSizes(dState, state, "TimeDerivative")
    call CodyHip_Check(cody_time_derivative(MctdhxHip_ctx, state, dState, time), "time_derivative")
  end subroutine

  module subroutine MctdhxHip_TimeDerivativeOrbsLin(dState, state, time)
    complex(R64), intent(out), contiguous, target :: dState(:)
    complex(R64), intent(in), contiguous, target :: state(:)
    real(R64), intent(in) :: time
    call CheckStateSizes(dState, state, "TimeDerivativeOrbsLin")
    call CodyHip_Check(cody_time_derivative_orbs_lin(MctdhxHip_ctx, state, dState, time), "time_derivative_orbs_lin")
  end subroutine

  module subroutine MctdhxHip_TimeDerivativeCoeffsPlusOrbsNonLin(dState, state, time)
    complex(R64), intent(out), contiguous, target :: dState(:)
    complex(R64), intent(in), contiguous, target :: state(:)
    real(R64), intent(in) :: time
    call CheckStateSizes(dState, state, "TimeDerivativeCoeffsPlusOrbsNonLin")
    call CodyHip_Check(cody_time_derivative_nonlin(MctdhxHip_ctx, state, dState, time), "time_derivative_nonlin")
  end subroutine

  subroutine CheckOrbShapes(dOrbs, orbs, where)
    complex(R64), intent(in) :: dOrbs(:, :), orbs(:, :)
    character(len=*), intent(in) :: where
    if (size(orbs, 1) /= Grid_nPoints_of_ctx() .or. size(orbs, 2) /= MctdhxHip_input%nOrbs .or. &
        any(shape(dOrbs) /= shape(orbs))) then
      write (*, '(A)') "MctdhxHip "//where//": shapes must be nG x nOrbs"
      error stop "MI355X MCTDHX: orbital shape mismatch"
    end if
  end subroutine

  module subroutine MctdhxHip_ApplySingleBodyOp(dOrbs, orbs, time, h1_)
    complex(R64), intent(inout), contiguous, target :: dOrbs(:, :)
    complex(R64), intent(in), contiguous, target :: orbs(:, :)
    real(R64), intent(in) :: time
    complex(R64), intent(out), allocatable, optional :: h1_(:, :)
    complex(c_double_complex) :: h1(MctdhxHip_input%nOrbs, MctdhxHip_input%nOrbs)
    call CheckOrbShapes(dOrbs, orbs, "ApplySingleBodyOp")
    call CodyHip_Check(cody_apply_single_body_acc(MctdhxHip_ctx, orbs, time, dOrbs, h1), "apply_single_body_acc")
    if (present(h1_)) then
      allocate (h1_(size(h1, 1), size(h1, 2)))
      h1_ = h1
    end if
  end subroutine

  module subroutine MctdhxHip_ApplyInteractionOp(dOrbs, orbs, i2, j2, time)
    complex(R64), intent(inout), contiguous :: dOrbs(:, :)
    complex(R64), intent(in), contiguous :: orbs(:, :)
    integer(I32), intent(in) :: i2
    integer(I32), intent(in) :: j2
    real(R64), intent(in) :: time
    call CheckOrbShapes(dOrbs, orbs, "ApplyInteractionOp")
    if (i2 < 1 .or. i2 > size(orbs, 2) .or. j2 < 1 .or. j2 > size(orbs, 2)) &
      error stop "MctdhxHip ApplyInteractionOp: orbital index out of range"
    call CodyHip_Check(cody_apply_interaction(MctdhxHip_ctx, orbs, int(i2 - 1, c_int), int(j2 - 1, c_int), time, &
                                              dOrbs), "apply_interaction")
  end subroutine

  !> Method_GetEnergy (S_Method_Mb_OrbBased.f90:132-161) on the GPU.
  function GetEnergy(time) result(res)
    use M_Method
    real(R64) :: res
    real(R64), intent(in) :: time
    real(c_double) :: e
    call CodyHip_Check(cody_energy(MctdhxHip_ctx, Method_state, time, e), "energy")
    res = e
  end function

  module function Method_Mb_OrbBased_MctdhxHip_Dipole() result(res)
    use M_Method
    real(R64) :: res
    real(c_double) :: d
    call CodyHip_Check(cody_dipole(MctdhxHip_ctx, Method_state, d), "dipole")
    res = d
  end function

  subroutine CheckBodyType(bt, where)
    integer(I32), intent(in), optional :: bt
    character(len=*), intent(in) :: where
    if (present(bt)) then
      if (bt /= 1) then
        write (*, '(A)') "MctdhxHip "//where//": the GPU leaf holds one body type"
        error stop "MI355X MCTDHX: unsupported body type"
      end if
    end if
  end subroutine

  module subroutine MctdhxHip_FillInteractionPotential(interactionPotential, src, time, bt1_, bt2_)
    complex(R64), intent(out), allocatable :: interactionPotential(:)
    complex(R64), intent(in), contiguous :: src(:)
    real(R64), intent(in) :: time
    integer(I32), intent(in), optional :: bt1_
    integer(I32), intent(in), optional :: bt2_
    complex(c_double_complex), allocatable :: s(:), p(:)
    call CheckBodyType(bt1_, "FillInteractionPotential")
    call CheckBodyType(bt2_, "FillInteractionPotential")
    if (size(src) /= Grid_nPoints_of_ctx()) error stop "MctdhxHip FillInteractionPotential: source size"
    allocate (s(size(src)), p(size(src)))
    s = src
    call CodyHip_Check(cody_interaction_potential(MctdhxHip_ctx, 1_c_int, s, p), "interaction_potential")
    allocate (interactionPotential(size(src)))
    interactionPotential = p
  end subroutine

  module subroutine MctdhxHip_MultiplyWithKineticOp(dOrb, orb, time, bt_)
    complex(R64), intent(out), contiguous :: dOrb(:)
    complex(R64), intent(in), contiguous :: orb(:)
    real(R64), intent(in) :: time
    integer(I32), intent(in), optional :: bt_
    complex(c_double_complex), allocatable :: o(:, :), d(:, :)
    integer :: nG
    call CheckBodyType(bt_, "MultiplyWithKineticOp")
    nG = Grid_nPoints_of_ctx()
    if (size(orb) /= nG .or. size(dOrb) /= nG) error stop "MctdhxHip MultiplyWithKineticOp: orbital size"
    ! the library applies T to the context's nO columns: the orbital in column 1, zeros elsewhere
    allocate (o(nG, MctdhxHip_input%nOrbs), d(nG, MctdhxHip_input%nOrbs))
    o = (0.0_R64, 0.0_R64)
    o(:, 1) = orb
    call CodyHip_Check(cody_apply_kinetic(MctdhxHip_ctx, o, d), "apply_kinetic")
    dOrb = d(:, 1)
  end subroutine

  module subroutine MctdhxHip_ProjectOnSubspace(dOrbs, orbs)
    complex(R64), intent(inout), contiguous :: dOrbs(:, :)
    complex(R64), intent(in), contiguous :: orbs(:, :)
    if (size(orbs, 2) /= MctdhxHip_input%nOrbs .or. any(shape(dOrbs) /= shape(orbs))) &
      error stop "MctdhxHip ProjectOnSubspace: shapes must be nG x nOrbs"
    call CodyHip_Check(cody_project_orbs(MctdhxHip_ctx, orbs, dOrbs), "project_orbs")
  end subroutine

  module subroutine MctdhxHip_ApplyCorrelationOp(dOrbs, orbs, rdm1, rdm2, time, h2_)
    complex(R64), intent(inout), contiguous, target :: dOrbs(:, :)
    complex(R64), intent(in), contiguous, target :: orbs(:, :)
    complex(R64), intent(in), contiguous :: rdm1(:, :)
    complex(R64), intent(in), contiguous :: rdm2(:, :, :, :)
    real(R64), intent(in) :: time
    complex(R64), intent(out), allocatable, optional :: h2_(:, :, :, :)
    complex(c_double_complex), allocatable :: h2(:, :, :, :)
    integer :: n
    n = MctdhxHip_input%nOrbs
    if (size(orbs, 2) /= n .or. any(shape(dOrbs) /= shape(orbs))) &
      error stop "MctdhxHip ApplyCorrelationOp: shapes must be nG x nOrbs"
    allocate (h2(n, n, n, n))
    call CodyHip_Check(cody_apply_correlation(MctdhxHip_ctx, orbs, rdm1, rdm2, dOrbs, h2), "apply_correlation")
    if (present(h2_)) then
      allocate (h2_(n, n, n, n))
      h2_ = h2
    end if
  end subroutine

  module subroutine MctdhxHip_Densities(rdm1, rdm2)
    use M_Method
    complex(R64), intent(out) :: rdm1(:, :)
    complex(R64), intent(out) :: rdm2(:, :, :, :)
    complex(c_double_complex), allocatable :: ri(:, :)
    allocate (ri(size(rdm1, 1), size(rdm1, 2)))
    call CodyHip_Check(cody_densities(MctdhxHip_ctx, Method_state, rdm1, rdm2, ri), "densities")
  end subroutine

  integer function Grid_nPoints_of_ctx()
    use M_Grid
    Grid_nPoints_of_ctx = Grid_nPoints
  end function

  module subroutine Method_Mb_OrbBased_MctdhxHip_SetLaser(coefficients, period)
    real(R64), intent(in) :: coefficients(:)
    real(R64), intent(in) :: period
    real(c_double) :: c(max(1, size(coefficients)))
    c = 0.0_R64
    c(1:size(coefficients)) = coefficients
    call CodyHip_Check(cody_set_laser(MctdhxHip_ctx, int(size(coefficients), c_int), c, period), "set_laser")
    MctdhxHip_laserBound = sum(abs(coefficients))
  end subroutine

  module function Method_Mb_OrbBased_MctdhxHip_DipoleGradient(t0, dt, nSteps, checkpointEvery, gradient, &
                                                              chebLo, chebHi, chebDegree) result(res)
    use M_Method
    real(R64), intent(in) :: t0, dt
    integer(I32), intent(in) :: nSteps, checkpointEvery
    real(R64), intent(out) :: gradient(:)
    real(R64), intent(out), optional :: chebLo, chebHi
    integer(I32), intent(out), optional :: chebDegree
    real(R64) :: res
    real(c_double) :: lo, hi, J
    real(c_double) :: g(max(1, size(gradient)))
    integer(c_int) :: K
    call CodyHip_Check(cody_set_chebyshev(MctdhxHip_ctx, dt, MctdhxHip_laserBound, lo, hi, K), "set_chebyshev")
    call CodyHip_Check(cody_upload_state(MctdhxHip_ctx, Method_state), "upload_state")
    call CodyHip_Check(cody_dipole_gradient(MctdhxHip_ctx, t0, dt, int(nSteps, c_int), int(checkpointEvery, c_int), &
                                            J, g), "dipole_gradient")
    call CodyHip_Check(cody_download_state(MctdhxHip_ctx, Method_state), "download_state")
    gradient = g(1:size(gradient))
    if (present(chebLo)) chebLo = lo
    if (present(chebHi)) chebHi = hi
    if (present(chebDegree)) chebDegree = K
    res = J
  end function

  subroutine PropagatorSetup
  end subroutine

  !> Propagator_Propagate for splitStep.order2 (S_Propagator_SplitStep_Order2.f90:52-77) with
  !> SIL on the linear part and RK4 on the nonlinear part, device resident between the copies.
  subroutine Propagate(state, t0, t1)
    complex(R64), intent(inout), contiguous :: state(:)
    real(R64), intent(in) :: t0
    real(R64), intent(in) :: t1
    integer(c_int) :: conv
    type(T_MctdhxHip_Input), pointer :: in
    in => MctdhxHip_input
    call CodyHip_Check(cody_upload_state(MctdhxHip_ctx, state), "upload_state")
    call CodyHip_Check(cody_strang(MctdhxHip_ctx, t0, t1, in%krylovDim, in%tolerance, in%maxRestarts, &
                                   merge(1, 0, in%metricQ), conv), "strang")
    call CodyHip_Check(cody_download_state(MctdhxHip_ctx, state), "download_state")
    if (conv == 0) write (*, '(A,ES13.6)') "Warning: Short iterative Lanczos step exceeded tolerance at t=", t0
  end subroutine

end submodule
