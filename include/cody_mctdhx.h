/*
 * cody_mctdhx.h — C ABI of the MI355X (gfx950) MCTDHX right-hand side and propagation.
 *
 * Drop-in boundary for CodyFortran's Fabricate/Setup + procedure-pointer API.  A Fortran
 * program binds these entry points through iso_c_binding (the pattern of
 * src/Utils/M_Utils_CombinationGslLib.f90:20-48) and assigns them to the procedure
 * pointers it already uses (Method_TimeDerivative, Propagator_Propagate, ...).
 *
 * Conventions (identical to the reference):
 *   - complex(R64) arrays are interleaved (re, im) doubles, column-major;
 *   - the packed state is [C(nC) ; orbs(nG, nO)] (S_Method_Mb_OrbBased_Mctdhx.f90:229-249);
 *   - CI configurations are the lexicographic N-of-M combinations of one fermionic body type;
 *   - every entry point returns 0 on success and a CODY_ERR_* code otherwise; the message of
 *     the last failure is available from cody_last_error() (the Fortran wrapper turns a
 *     non-zero status into `error stop`, the reference's only error mechanism).
 * Pointers named d_* are device (HBM) pointers; all other pointers are host pointers.
 * A `stream` argument is a hipStream_t (NULL selects the context's own stream).
 * Contexts are not thread-safe; one context drives one GPU.
 */
#ifndef CODY_MCTDHX_H
#define CODY_MCTDHX_H

#ifdef __cplusplus
extern "C" {
#endif

#define CODY_OK 0
#define CODY_ERR_ARG 1         /* invalid argument or size */
#define CODY_ERR_HIP 2         /* HIP runtime failure */
#define CODY_ERR_STATE 3       /* call out of order (e.g. RHS before grid/potentials are set) */
#define CODY_ERR_UNSUPPORTED 4 /* configuration outside what the kernels implement */
#define CODY_ERR_IO 5          /* a restart file cannot be opened, or holds fewer bytes than needed */

/* precision of the mean-field contraction W = V . S (the dominant kernel) */
#define CODY_PREC_FP64 0       /* FP64 MFMA, FP64 accumulation */
#define CODY_PREC_SPLIT_FP32 1 /* V and S split into FP32 hi+lo, 3 FP32-MFMA products, FP64 chunk sums */
#define CODY_PREC_BF16X3 2     /* V and S split into 3 BF16 terms, 6 BF16-MFMA products, FP32 accumulate */
#define CODY_PREC_INT8_SLICED 3 /* V and S scaled and split into five balanced base-254 digits;
                                   exact int8-MFMA products with int32 accumulation, FP64
                                   recombination */
#define CODY_PREC_BF16_SLICED 4 /* the same digits as BF16 operands of BF16 MFMA with FP32
                                   accumulation (exact per 64-point block), FP64 recombination:
                                   W equal to CODY_PREC_INT8_SLICED's bit for bit; at most 80 pair
                                   columns (8 orbitals) and 16384 padded grid points */

typedef struct cody_ctx cody_ctx;
typedef struct cody_batch cody_batch;

const char* cody_last_error(void);

/* ---------------------------------------------------------------- grid helpers (host) */
/* Number of FEDVR grid points: src/Utils/M_Utils_Fedvr.f90:108-111. */
int cody_fedvr_npoints(int nE, int nLoc, int xminExcl, int xmaxExcl);
/* FEDVR nodes and summed weights: replaces Fedvr_CreateCtx, src/Utils/M_Utils_Fedvr.f90:76-149. */
int cody_fedvr_grid(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                    double* x, double* w);
/* Reference-element stiffness K(i,j) = -sum_k w_k D(k,i) D(k,j) on [-1/2,1/2]:
 * replaces DerivativeFedvr_CreateCtx, src/Utils/M_Utils_DerivativeFedvr.f90:29-76. */
int cody_fedvr_kref(int nLoc, double* kref);

/* ---------------------------------------------------------------- context (Setup) */
/* One spin-polarised fermionic body type: nBodies particles in nOrbs orbitals on nG points.
 * reg is method.mb.orbBased.regularizationParameter (S_Method_Mb_OrbBased.f90:79).
 * Replaces the allocations of Method_Setup (S_Method_Mb_OrbBased_Mctdhx.f90:40-68),
 * ConfigList_Setup (S_ConfigList_AllActive.f90:105-117) and Coeffs_Fabricate
 * (S_Coeffs_Generic.f90:55-82). */
int cody_create(cody_ctx** ctx, int device, int nG, int nBodies, int nOrbs, double reg);
int cody_destroy(cody_ctx* ctx);
int cody_sizes(const cody_ctx* ctx, int* nC, int* nState);
int cody_get_stream(const cody_ctx* ctx, void** stream);

/* Grid coordinates and quadrature weights: Grid_Linear_xCoord / Grid_Linear_weights
 * (src/Grid/Linear/S_Grid_Linear.f90:53-61 uses them in Grid_InnerProduct). */
int cody_set_grid(cody_ctx* ctx, const double* x, const double* w);
/* Linear FEDVR kinetic energy -(1/2m) d2/dx2 in weak form on nE equal elements of nLoc GLL
 * points (the FEDVR operator of src/Utils/M_Utils_DerivativeFedvr.f90:99-152 behind the
 * SysKinetic_MultiplyWithKineticOp pointer, src/SysKinetic/M_SysKinetic.f90:141-160). */
int cody_set_kinetic_fedvr(cody_ctx* ctx, double xmin, double xmax, int nE, int nLoc,
                           int xminExcl, int xmaxExcl, double mass);
/* External potential values V_ext(x_i) (SysPotential_FillExternalPotential,
 * src/SysPotential/M_SysPotential.f90:71-101). */
int cody_set_potential(cody_ctx* ctx, const double* vext);
/* Soft-Yukawa centres -sum_c q_c e^{-d_c r}/(sqrt(r^2+a1_c^2)+a2_c):
 * S_SysPotential_Linear_SoftYukawa_StdImpl.f90:45-84. */
int cody_set_potential_softyukawa(cody_ctx* ctx, int nCenters, const double* position,
                                  const double* charge, const double* soft1, const double* soft2,
                                  const double* damp);
/* Two-body kernel K(r) = s e^{-d r}/(sqrt(r^2+a1^2)+a2) at true distances r = |x_i - x_j|,
 * tabulated on the device as an nG x nG matrix: replaces SysInteraction_Setup +
 * FillInteractionPotential (src/SysInteraction/Linear/SoftYukawa/S_SysInteraction_Linear_SoftYukawa.f90:67-73,
 * src/Utils/M_Utils_ConvolutionIntegral.f90:71-107). */
int cody_set_interaction_softyukawa(cody_ctx* ctx, double strength, double soft1, double soft2,
                                    double damp);
/* Length-gauge dipole field V(x,t) = E(t) x, E(t) = sum_k c_k sin(k pi t/T) sin^2(pi t/T)
 * (a 1D leaf in the pattern of src/SysPotential/Ylm/CoulombLaser; nK = 0 switches it off). */
int cody_set_laser(cody_ctx* ctx, int nK, const double* c, double T);
/* Precision of the mean-field contraction (CODY_PREC_*). */
int cody_set_precision(cody_ctx* ctx, int mode);
/* Replay of captured launch sequences (no counterpart in the reference, whose right-hand side is
 * a chain of host calls).  By default the propagators (cody_sil, cody_strang, cody_strang_cheb)
 * capture a sub-step or step into a HIP graph on first use and replay it, and a single right-hand
 * side (the device-pointer entry points, the stages of cody_rk4) is launched kernel by kernel:
 * its few launches cost less than a graph launch.  enable != 0 also replays single right-hand
 * sides as graphs, one per (state, dState, mode); 0 launches everything individually. */
int cody_set_graphs(cody_ctx* ctx, int enable);
/* Zero sources of the mean-field contraction: with enable != 0 (the default) the contraction skips
 * every 64-point (int8 form) or 16-point (FP64 form) block of the contraction index where all pair
 * sources conj(phi_l) phi_k w are exactly zero, as the reference leaves zero sources out of its
 * convolution sums (ConvolutionIntegral_Apply, src/Utils/M_Utils_ConvolutionIntegral.f90:84-93).
 * The skipped products are exact zeros: W changes only by the summation order (~1e-16).  0
 * contracts every block (the dense work, for measurement).  Grids above 16384 points are always
 * contracted densely. */
int cody_set_source_skip(cody_ctx* ctx, int enable);
/* Project autotuner (no counterpart in the reference, which has no tiled kernels): times the FP64
 * contraction of this context's shape for each k split (1, 2, 4, 8 contiguous k ranges per row
 * group) and keeps the fastest; *wsplit (optional) receives the choice.  The timing uses the
 * context's own state.  With *wsplit > 0 on entry that split is set without timing. */
int cody_autotune(cody_ctx* ctx, int* wsplit);

/* ---------------------------------------------------------------- state transfer */
/* The context's device-resident state (Method_state, src/Method/M_Method.f90:47). */
int cody_upload_state(cody_ctx* ctx, const double* state);
int cody_download_state(cody_ctx* ctx, double* state);
int cody_state_device_ptr(cody_ctx* ctx, void** d_state);

/* ---------------------------------------------------------------- restart files (host, no context) */
/* CodyFortran's restart files in directory `dir` (NULL or "": the current directory), each a
 * headerless stream of native-endian complex128 values, as M_Utils_DataStorage writes them
 * (src/Utils/M_Utils_DataStorage.f90:20-105):
 *   coeffs.in       the nC CI coefficients: Coeffs_SaveCoeffs (src/Coeffs/S_Coeffs.f90:167-174)
 *                   and CoeffsInit load (src/CoeffsInit/Load/S_CoeffsInit_Load.f90:70-79);
 *   orbBB_II.in     orbital II (1-based) of body type BB, nG values: Orbs_SaveOrbs
 *                   (src/Orbs/S_Orbs.f90:251-275) and OrbsInit load
 *                   (src/OrbsInit/Load/S_OrbsInit_Load.f90:33-50).
 * `state` is the packed host state [C(nC) ; orbs(nG, sum nOrbs)], orbitals ordered by body type.
 * Saving replaces existing files; loading fails with CODY_ERR_IO when a file is missing or
 * shorter than needed.  A device-resident state goes through cody_download_state first (or
 * cody_upload_state after loading). */
int cody_save_state_files(const char* dir, int nC, int nG, int nBodyTypes, const int* nOrbs,
                          const double* state);
int cody_load_state_files(const char* dir, int nC, int nG, int nBodyTypes, const int* nOrbs,
                          double* state);
/* rdm2BtXX_YY.in: the (XX, YY) body-type block rho2(i1, i2, j1, j2) of the two-body density
 * (i1, j1 of body type XX; i2, j2 of YY), column-major, for every pair of body types:
 * Coeffs_SaveTwoRdm (src/Coeffs/S_Coeffs.f90:185-214).  rdm2 is the full nO^4 array over the
 * combined orbital index, as cody_densities returns it. */
int cody_save_two_rdm(const char* dir, int nBodyTypes, const int* nOrbs, const double* rdm2);

/* ---------------------------------------------------------------- Method */
/* Full MCTDHX right-hand side  dState = -i [ (1-P)(h phi + corr) ; (H - <H>) C ]:
 * replaces Method_TimeDerivative (I_Method_TimeDerivative, src/Method/M_Method.f90:82-90),
 * MCTDHX implementation S_Method_Mb_OrbBased_Mctdhx.f90:101-149.  Overwrites d_dstate. */
int cody_time_derivative_dev(cody_ctx* ctx, const void* d_state, void* d_dstate, double t,
                             void* stream);
/* Same, host arrays in and out (copies through the context's buffers; synchronous). */
int cody_time_derivative(cody_ctx* ctx, const double* state, double* dstate, double t);
/* Split-step vector fields: Method_Mb_OrbBased_TimeDerivativeOrbsLin (:152-177) and
 * Method_Mb_OrbBased_TimeDerivativeCoeffsPlusOrbsNonLin (:180-226). */
int cody_time_derivative_orbs_lin_dev(cody_ctx* ctx, const void* d_state, void* d_dstate, double t,
                                      void* stream);
int cody_time_derivative_nonlin_dev(cody_ctx* ctx, const void* d_state, void* d_dstate, double t,
                                    void* stream);
/* The same two vector fields with host arrays (copies through the context's buffers;
 * synchronous): the procedures the reference's Method_Mb_OrbBased_TimeDerivativeOrbsLin and
 * Method_Mb_OrbBased_TimeDerivativeCoeffsPlusOrbsNonLin pointers
 * (src/Method/Mb/OrbBased/M_Method_Mb_OrbBased.f90:272-300) bind to. */
int cody_time_derivative_orbs_lin(cody_ctx* ctx, const double* state, double* dstate, double t);
int cody_time_derivative_nonlin(cody_ctx* ctx, const double* state, double* dstate, double t);
/* Energy E = Tr[rho1 h1] + Tr[rho2 h2] of a host state: replaces Method_GetEnergy
 * (src/Method/Mb/OrbBased/S_Method_Mb_OrbBased.f90:132-161). */
int cody_energy(cody_ctx* ctx, const double* state, double t, double* energy);
/* Dipole <x> = Re sum_ij rho1(i,j) <phi_j|x phi_i>_w of a host state: replaces the value of
 * PrinterObservableLinear_DumpDipole (src/Utils/M_Utils_PrinterObservableLinear.f90:25-70). */
int cody_dipole(cody_ctx* ctx, const double* state, double* dipole);
/* Same for a device state; the value (one double) is written to d_dipole, enqueued on `stream`. */
int cody_dipole_dev(cody_ctx* ctx, const void* d_state, void* d_dipole, void* stream);

/* ---------------------------------------------------------------- operator pieces (host in/out) */
/* Mean-field potentials of the unordered pairs l <= k (pair p = l*nO - l(l-1)/2 + k-l):
 * W(:,p) = FillInteractionPotential(FillInteractionSrc(phi_l, phi_k)) (S_SysInteraction_Linear.f90:57-72).
 * orbs is nG x nO, W is nG x P (complex, column-major). */
int cody_meanfield_w(cody_ctx* ctx, const double* orbs, double* W);
/* dOrbs = (T + V) orbs and h1 = <phi_i|(T+V) phi_j> (ApplySingleBodyOp, S_Method_Mb_OrbBased.f90:976-1034). */
int cody_apply_single_body(cody_ctx* ctx, const double* orbs, double t, double* dorbs, double* h1);
/* Method_Mb_OrbBased_ApplySingleBodyOp with its accumulate semantics (I_Method_Mb_OrbBased_ApplySingleBodyOp,
 * M_Method_Mb_OrbBased.f90:253-269; S_Method_Mb_OrbBased.f90:976-1034): dorbs += (T + V + E(t) x) orbs,
 * then h1 (nO x nO, may be NULL) = <phi_i|dorbs_j>_w of the accumulated dorbs. */
int cody_apply_single_body_acc(cody_ctx* ctx, const double* orbs, double t, double* dorbs, double* h1);
/* Method_Mb_OrbBased_ApplyInteractionOp (I_Method_Mb_OrbBased_ApplyInteractionOp,
 * M_Method_Mb_OrbBased.f90:195-208; S_Method_Mb_OrbBased.f90:491-542): for the source pair (i2, j2)
 * (0-based orbital indices), dorbs(:, i) += W * orbs(:, i) for every orbital i, with W the
 * interaction potential of the source conj(phi_i2) phi_j2 w, on the FP64 MFMA contraction.
 * Accumulate semantics; nG x nO complex128 column-major host arrays. */
int cody_apply_interaction(cody_ctx* ctx, const double* orbs, int i2, int j2, double t, double* dorbs);
/* rho1, rho2 of a CI vector (Coeffs_ApplyH1FillRdm1 / Coeffs_ApplyH2FillRdm2 in density mode,
 * S_Coeffs_Generic.f90:103-319 with the unpack of :325-434) and the regularised inverse
 * rho1 (rho1^2 + eps)^-1 (S_Method_Mb_OrbBased.f90:733-744). */
int cody_densities(cody_ctx* ctx, const double* coeffs, double* rdm1, double* rdm2, double* reginv);

/* Operator pieces with the reference's procedure-pointer semantics (host arrays, column-major
 * complex128, nG x nO orbitals):
 *   cody_interaction_potential: pot(:, s) = sum_j K(x_i - x_j) src(x_j, s) for nSrc sources that
 *     already hold the weights (src = conj(phi_a) phi_b w): SysInteraction_FillInteractionPotential
 *     (src/SysInteraction/M_SysInteraction.f90:94-110, source convention
 *     src/SysInteraction/Linear/S_SysInteraction_Linear.f90:57-72), on the FP64 MFMA contraction;
 *   cody_apply_correlation: dorbs += the correlation term, h2 (nO^4, may be NULL) =
 *     1/2 <phi_i1|W_(b a) phi_j1> at (i1, a, j1, b): Method_Mb_OrbBased_ApplyCorrelationOp
 *     (src/Method/Mb/OrbBased/M_Method_Mb_OrbBased.f90:211-231, S_Method_Mb_OrbBased.f90:647-872),
 *     accumulate semantics;
 *   cody_project_orbs: dorbs <- dorbs - sum_i phi_i <phi_i|dorbs>: Orbs_ProjectOnSubspace
 *     (src/Orbs/M_Orbs.f90:132-153);
 *   cody_apply_kinetic: dorbs = T orbs: SysKinetic_MultiplyWithKineticOp
 *     (src/SysKinetic/M_SysKinetic.f90:141-160), overwrite semantics. */
int cody_interaction_potential(cody_ctx* ctx, int nSrc, const double* src, double* pot);
int cody_apply_correlation(cody_ctx* ctx, const double* orbs, const double* rdm1, const double* rdm2, double* dorbs,
                           double* h2);
int cody_project_orbs(cody_ctx* ctx, const double* orbs, double* dorbs);
int cody_apply_kinetic(cody_ctx* ctx, const double* orbs, double* dorbs);

/* No counterpart in the reference: the orbital projector of cody_project_orbs for callers that
 * cannot allocate device memory (code under HIP graph capture, tight loops, harnesses that forbid
 * hipMalloc).  Context-free, device pointers only, enqueued on `stream` (NULL: the null stream):
 *   d_dorbs <- d_dorbs - sum_i phi_i <phi_i|d_dorbs>_w  (in place; Orbs_ProjectOnSubspace,
 *   src/Orbs/M_Orbs.f90:132-153), d_orbs nG x nO complex128 column-major, d_w the nG weights.
 * It allocates nothing: its scratch (the overlap partial sums) is a static device array of
 * 576 KiB per GPU, so nG <= CODY_NOALLOC_MAX_POINTS and nO <= 12, else CODY_ERR_ARG.  All calls
 * on one GPU share that array: calls that overlap in time on one GPU race. */
#define CODY_NOALLOC_MAX_POINTS 65536
int cody_project_orbs_dev_noalloc(int nG, int nO, const double* d_w, const void* d_orbs, void* d_dorbs,
                                  void* stream);

/* ---------------------------------------------------------------- propagation (device state) */
/* The nonlinear part B of the split step, used by cody_time_derivative_nonlin_dev, cody_rk4(full=0)
 * and cody_strang:
 *   CODY_SPLIT_REFERENCE (default): TimeDerivativeCoeffsPlusOrbsNonLin
 *     (S_Method_Mb_OrbBased_Mctdhx.f90:180-226), which evolves C with the full h1 but leaves out
 *     the -P h phi term of the orbitals, so A + B differs from the full right-hand side;
 *   CODY_SPLIT_CONSISTENT: the full right-hand side minus the linear part, so that A + B is the
 *     full right-hand side and the Strang step is second order.  No counterpart in the reference;
 *     the option for runs that need the step-size convergence of the energy. */
#define CODY_SPLIT_REFERENCE 0
#define CODY_SPLIT_CONSISTENT 1
int cody_set_split(cody_ctx* ctx, int mode);
/* Short iterative Lanczos on the linear split part with full re-orthogonalisation and step
 * halving (S_IntegratorList_Sil.f90:114-312), metric = unit CI weights + grid weights per
 * orbital block when useMetric != 0.  Propagates the context's state from t0 to t1.
 * *converged receives 1 if every sub-step met the tolerance. */
int cody_sil(cody_ctx* ctx, double t0, double t1, int krylovDim, double tol, int maxRestarts,
             int useMetric, int* converged);
/* Classical RK4 on the nonlinear split part (S_IntegratorList_Rk_O4Expl.f90:90-141) or, with
 * full != 0, on the full right-hand side. */
int cody_rk4(cody_ctx* ctx, double t0, double t1, int full);
/* Strang step A(dt/2) B(dt) A(dt/2), A = SIL on the linear part, B = RK4 on the nonlinear part:
 * replaces Propagator_Propagate (src/Propagator/M_Propagator.f90:75-99) for splitStep.order2
 * (S_Propagator_SplitStep_Order2.f90:52-77). */
int cody_strang(cody_ctx* ctx, double t0, double t1, int krylovDim, double tol, int maxRestarts,
                int useMetric, int* converged);

/* ---------------------------------------------------------------- adjoint (differentiation) */
/* No counterpart in the reference, which has no adjoint: the backward pass of the right-hand
 * side and of a Strang propagation, for gradients of the final dipole <x>(T) with respect to the
 * laser coefficients of cody_set_laser.  Gradients are taken with respect to the real inner
 * product <a, b> = Re(sum_C conj(a) b + sum_orbitals sum_x w(x) conj(a(x)) b(x)) (the SIL
 * metric), in which h = T + V + E(t) x is self-adjoint.  They need the FP64 contraction. */
#define CODY_PART_FULL 0   /* Method_TimeDerivative (S_Method_Mb_OrbBased_Mctdhx.f90:101-149) */
#define CODY_PART_LIN 1    /* TimeDerivativeOrbsLin (:152-177) */
#define CODY_PART_NONLIN 2 /* the nonlinear split part of cody_set_split */
/* d_grad = J^T d_lam for the right-hand side `part` at (d_state, t) and *d_ebar = <d_lam, df/dE>
 * (device pointers; d_ebar one double), enqueued on `stream` (NULL: the context's stream). */
int cody_time_derivative_vjp_dev(cody_ctx* ctx, int part, const void* d_state, const void* d_lam, void* d_grad,
                                 double t, void* d_ebar, void* stream);
/* d_grad = gradient of <x> (cody_dipole_dev) at d_state. */
int cody_dipole_grad_dev(cody_ctx* ctx, const void* d_state, void* d_grad, void* stream);
/* Fixes the linear half-steps of cody_strang_cheb / cody_dipole_gradient: a Chebyshev expansion of
 * exp(-i dt/2 h(E)) on Gershgorin bounds [lo, hi] of h(E) for |E| <= emax, of the degree that
 * brings the dropped Bessel terms below 1e-17.  A fixed polynomial makes the step a smooth
 * function of the field (a Krylov step's dimension can change with it). */
int cody_set_chebyshev(cody_ctx* ctx, double dt, double emax, double* lo, double* hi, int* degree);
/* One Strang step of the context's state, A(dt/2) B(dt) A(dt/2) as in cody_strang
 * (S_Propagator_SplitStep_Order2.f90:52-77) with A the Chebyshev half-step (E at the half-step's
 * midpoint) and B classical RK4 on the nonlinear part; t1 - t0 must be the dt of
 * cody_set_chebyshev. */
int cody_strang_cheb(cody_ctx* ctx, double t0, double t1);
/* Propagates the context's state by nSteps Chebyshev-Strang steps of dt from t0 (the state is left
 * at t0 + nSteps dt), *J = <x> at the end and gradc[k] = dJ/dc_k for the nLaser coefficients, by
 * the discrete adjoint of the steps: the forward pass keeps a checkpoint every `checkpointEvery`
 * steps, the backward pass recomputes each segment from its checkpoint.  With checkpointEvery >=
 * nSteps the forward pass keeps every step's tape instead (about 37 MB per step at 4096 points,
 * M = 8) and nothing is recomputed. */
int cody_dipole_gradient(cody_ctx* ctx, double t0, double dt, int nSteps, int checkpointEvery, double* J,
                         double* gradc);

/* ---------------------------------------------------------------- batches of instances */
/* Independent instances (each its own context: grid, orbitals, interaction) evaluated together.
 * No counterpart in the reference, which runs one Method per process; this is the GPU form of
 * calling Method_TimeDerivative (S_Method_Mb_OrbBased_Mctdhx.f90:101-149) for every instance:
 * all right-hand sides are captured into one HIP graph (four parallel chains) on first use for
 * a set of state pointers and replayed afterwards.  The contexts must outlive the batch. */
int cody_batch_create(cody_batch** batch, int n, cody_ctx* const* ctxs);
int cody_batch_destroy(cody_batch* batch);
/* dStates[i] = TimeDerivative(states[i], t) of instance i, device pointers, enqueued on `stream`
 * (0: the batch's own stream). */
int cody_batch_time_derivative_dev(cody_batch* batch, void* const* d_states, void* const* d_dstates,
                                   double t, void* stream);
/* Schedule autotuner of a batch (no counterpart in the reference): times each instance's own
 * right-hand side once, then builds and times the batched graph for each candidate schedule and
 * keeps the fastest for these state pointers.  A schedule is (parallel chains, split threshold):
 * groups of one orbital count are split while a group's measured cost exceeds the threshold times
 * a chain's share, and groups and the other instances go to the chains longest first.  Candidates,
 * in order: (4, 1.0), (4, 0.7), (4, 0.5), (2, 1.0), (2, 0.5), (1, 1.0).  cand_ms (optional, max_cand
 * entries) receives each candidate's time per batched right-hand side, *pick (optional) the index
 * of the chosen one.  d_dstates are overwritten (one right-hand side at t = 0).  Results are
 * bit-identical under every schedule (the grouped launches run the instances' own bodies). */
int cody_batch_autotune(cody_batch* batch, void* const* d_states, void* const* d_dstates, double* cand_ms,
                        int max_cand, int* pick);
/* A fixed schedule instead of the autotuned one: parallel chains (1..4) and split threshold (> 0),
 * with the measured costs of the last cody_batch_autotune (else estimates).  Takes effect on the
 * next batched call. */
int cody_batch_set_schedule(cody_batch* batch, int chains, double split_threshold);

/* ---------------------------------------------------------------- measurement */
/* Runs `iters` right-hand sides of the context's state and returns the average duration of each
 * launch of the sequence, timed with HIP events on the context's stream (names in `names`,
 * ';'-separated, at most nMax entries).  Returns the number of launches per RHS in *nKernels. */
int cody_profile_rhs(cody_ctx* ctx, int iters, int nMax, double* avg_ms, char* names, int namesLen,
                     int* nKernels);

/* Average duration of one launch of the mean-field contraction kernel of the current precision
 * (the dominant kernel; for the int8-sliced form the one with the orbital correlation in its
 * epilogue, which reads R, rho2 and h phi of the last right-hand side), from `iters` back-to-back
 * launches after one warm-up on the context's stream bracketed by two HIP events; the pair
 * sources and the contraction operands are formed once from the context's state beforehand. */
int cody_profile_contraction(cody_ctx* ctx, int iters, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif
