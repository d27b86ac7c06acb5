! Implementation of the GPU integrator leaves.
submodule(M_IntegratorList_Hip) S_IntegratorList_Hip
  use iso_c_binding
  use M_CodyHip
  use M_Method_Mb_OrbBased_MctdhxHip
  implicit none

contains

  module subroutine IntegratorList_SilHip_Allocate(e, path)
    class(T_IntegratorList_E), allocatable, intent(out) :: e
    character(len=*), intent(in) :: path
    allocate (T_IntegratorList_E_SilHip :: e)
    e%path = path
  end subroutine

  module subroutine IntegratorList_Rk4Hip_Allocate(e, path)
    class(T_IntegratorList_E), allocatable, intent(out) :: e
    character(len=*), intent(in) :: path
    allocate (T_IntegratorList_E_Rk4Hip :: e)
    e%path = path
  end subroutine

  module subroutine SilHip_Fabricate(this)
    class(T_IntegratorList_E_SilHip), intent(inout) :: this
    this%krylovDim = MctdhxHip_input%krylovDim
    this%tolerance = MctdhxHip_input%tolerance
    this%maxRestarts = MctdhxHip_input%maxRestarts
    if (this%krylovDim < 1 .or. this%tolerance <= 0.0_R64 .or. this%maxRestarts < 0) &
      error stop "IntegratorList silHip: krylovDim >= 1, tolerance > 0 and maxRestarts >= 0 are required"
  end subroutine

  module subroutine SilHip_Setup(this)
    class(T_IntegratorList_E_SilHip), intent(inout) :: this
  end subroutine

  !> The metric the device Lanczos holds: 1 on the nC coefficients, the grid weights on each
  !> orbital block (relative agreement 1e-12).
  subroutine CheckStateMetric(weights)
    use M_Grid
    real(R64), intent(in) :: weights(:)
    integer :: nC, nG, i
    real(R64) :: dev
    nC = MctdhxHip_nCoeffs
    nG = Grid_nPoints
    dev = maxval(abs(weights(1:nC) - 1.0_R64))
    do i = 1, MctdhxHip_input%nOrbs
      dev = max(dev, maxval(abs(weights(nC + (i - 1) * nG + 1:nC + i * nG) - Grid_Linear_weights)) &
                / maxval(abs(Grid_Linear_weights)))
    end do
    if (dev > 1.0e-12_R64) &
      error stop "IntegratorList silHip: metricWeights must be 1 on the coefficients and the grid weights on the orbitals"
  end subroutine

  module subroutine SilHip_Integrate(this, state, t0, t1)
    class(T_IntegratorList_E_SilHip), intent(inout) :: this
    complex(R64), intent(inout), contiguous :: state(:)
    real(R64), intent(in) :: t0
    real(R64), intent(in) :: t1
    integer(c_int) :: conv, useMetric
    if (.not. associated(this%TimeDerivative, MctdhxHip_TimeDerivativeOrbsLin)) &
      error stop "IntegratorList silHip: TimeDerivative must be the GPU leaf's TimeDerivativeOrbsLin"
    if (size(state) /= MctdhxHip_nState) error stop "IntegratorList silHip: state size mismatch"
    useMetric = 0
    if (allocated(this%metricWeights)) then
      if (size(this%metricWeights) /= size(state)) &
        error stop "IntegratorList silHip: metricWeights size does not match state size"
      call CheckStateMetric(this%metricWeights)
      useMetric = 1
    end if
    call CodyHip_Check(cody_upload_state(MctdhxHip_ctx, state), "upload_state")
    call CodyHip_Check(cody_sil(MctdhxHip_ctx, t0, t1, int(this%krylovDim, c_int), this%tolerance, &
                                int(this%maxRestarts, c_int), useMetric, conv), "sil")
    call CodyHip_Check(cody_download_state(MctdhxHip_ctx, state), "download_state")
    if (conv == 0) write (*, '(A,ES13.6)') "Warning: Short iterative Lanczos step exceeded tolerance in the step from t=", t0
  end subroutine

  module subroutine Rk4Hip_Fabricate(this)
    class(T_IntegratorList_E_Rk4Hip), intent(inout) :: this
  end subroutine

  module subroutine Rk4Hip_Setup(this)
    class(T_IntegratorList_E_Rk4Hip), intent(inout) :: this
  end subroutine

  module subroutine Rk4Hip_Integrate(this, state, t0, t1)
    class(T_IntegratorList_E_Rk4Hip), intent(inout) :: this
    complex(R64), intent(inout), contiguous :: state(:)
    real(R64), intent(in) :: t0
    real(R64), intent(in) :: t1
    integer(c_int) :: full
    if (associated(this%TimeDerivative, MctdhxHip_TimeDerivativeCoeffsPlusOrbsNonLin)) then
      full = 0
    else if (associated(this%TimeDerivative, MctdhxHip_TimeDerivative)) then
      full = 1
    else
      error stop "IntegratorList rk4Hip: TimeDerivative must be the GPU leaf's CoeffsPlusOrbsNonLin or TimeDerivative"
    end if
    if (size(state) /= MctdhxHip_nState) error stop "IntegratorList rk4Hip: state size mismatch"
    call CodyHip_Check(cody_upload_state(MctdhxHip_ctx, state), "upload_state")
    call CodyHip_Check(cody_rk4(MctdhxHip_ctx, t0, t1, full), "rk4")
    call CodyHip_Check(cody_download_state(MctdhxHip_ctx, state), "download_state")
  end subroutine

end submodule
