/*
 * FP64 CPU oracle for the MCTDHX orbital + CI equations of motion.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the
 * smoke entry point and the CPU-baseline leg of the benchmark may load it.
 * The GPU product never links or calls it.
 *
 * It restates the reference algorithm (CodyFortran, Fortran 2008) in plain
 * C99 + OpenMP.  Every function cites the reference file:line it follows.
 * Array conventions follow the reference's Fortran storage: column-major,
 * complex(R64) interleaved (re, im), state = [C(nC) ; orbs(nG, nO)].
 */
#ifndef CODY_ORACLE_H
#define CODY_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

/* ---- grid helpers (stdlib gauss_legendre_lobatto, M_Utils_Fedvr) ---- */
int orc_gll(int n, double a, double b, double* x, double* w);
int orc_fedvr_npoints(int nE, int nLoc, int xminExcl, int xmaxExcl);
int orc_fedvr_grid(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                   double* x, double* w);
int orc_fedvr_kref(int nLoc, double* kref /* nLoc*nLoc col-major */, double* dref);
/* second / first derivative on an FEDVR grid (M_Utils_DerivativeFedvr.f90:121-152) */
int orc_fedvr_d2(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                 const double complex* f, double complex* d2f);
int orc_fedvr_d1(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                 const double complex* f, double complex* d1f);

/* ---- system ---- */
orc_sys* orc_create(int nG, const double* x, const double* w, int nBT, const int* nBodies,
                    const int* nOrbs, double reg);
void orc_destroy(orc_sys* s);
int orc_ncoeffs(const orc_sys* s);
int orc_norbs(const orc_sys* s);
int orc_set_kinetic_fedvr(orc_sys* s, double xmin, double xmax, int nE, int nLoc, int xminExcl,
                          int xmaxExcl, const double* mass);
int orc_set_kinetic_fd(orc_sys* s, double dx, const double* mass);
int orc_set_potential(orc_sys* s, const double* vext);
/* interaction K(|x_i - x_j|) as a dense nG x nG real matrix (copied) */
int orc_set_interaction(orc_sys* s, const double* vint);
/* dipole-field coupling E(t) x with E(t) = sum_k c_k sin(k pi t/T) sin^2(pi t/T) */
int orc_set_laser(orc_sys* s, int nC, const double* c, double T);

/* ---- CI (Coeffs_Generic) ---- */
int orc_rdm1(const orc_sys* s, const double complex* coeffs, double complex* rdm1);
int orc_rdm2(const orc_sys* s, const double complex* coeffs, double complex* rdm2);
int orc_apply_h1(const orc_sys* s, const double complex* coeffs, const double complex* h1,
                 double complex* dcoeffs);
int orc_apply_h2(const orc_sys* s, const double complex* coeffs, const double complex* h2,
                 double complex* dcoeffs);
int orc_config_codes(const orc_sys* s, int bt, unsigned* codes);

/* ---- operators (Method_Mb_OrbBased) ---- */
int orc_apply_single_body(const orc_sys* s, const double complex* orbs, double t,
                          double complex* dorbs /* += */, double complex* h1 /* may be NULL */);
int orc_meanfield_w(const orc_sys* s, const double complex* orbs, double complex* w /* nG x P */);
/* zero-source compaction of the mean-field sweep (on by default; off: every source point) and the
 * number of source points the sweep visits for given orbitals */
void orc_set_compaction(int on);
int orc_active_sources(const orc_sys* s, const double complex* orbs);
int orc_apply_correlation(const orc_sys* s, const double complex* orbs, const double complex* rdm1,
                          const double complex* rdm2, double complex* dorbs /* += */,
                          double complex* h2 /* may be NULL */);
int orc_fill_h1(const orc_sys* s, const double complex* orbs, double t, double complex* h1);
int orc_fill_h2(const orc_sys* s, const double complex* orbs, double complex* h2);
int orc_reg_inverse(const orc_sys* s, const double complex* rdm1, double complex* reginv);
int orc_project_orbs(const orc_sys* s, const double complex* orbs, double complex* dorbs);

/* ---- method (S_Method_Mb_OrbBased_Mctdhx) ---- */
int orc_time_derivative(const orc_sys* s, const double complex* state, double t,
                        double complex* dstate);
int orc_time_derivative_orbs_lin(const orc_sys* s, const double complex* state, double t,
                                 double complex* dstate);
int orc_time_derivative_nonlin(const orc_sys* s, const double complex* state, double t,
                               double complex* dstate);
double orc_energy(const orc_sys* s, const double complex* state, double t);
double orc_dipole(const orc_sys* s, const double complex* state);

/* ---- integrators / propagators ---- */
/* ORC_RHS_NONLIN is the reference's nonlinear split part (TimeDerivativeCoeffsPlusOrbsNonLin);
 * ORC_RHS_NONLIN_CONSISTENT is the full right-hand side minus the linear part, the nonlinear part
 * for which A + B is the full vector field (not a reference routine). */
enum { ORC_RHS_FULL = 0, ORC_RHS_LIN = 1, ORC_RHS_NONLIN = 2, ORC_RHS_IMAG = 3, ORC_RHS_NONLIN_CONSISTENT = 4 };
int orc_rk4(const orc_sys* s, int rhs, double complex* state, double t0, double t1);
int orc_rk1(const orc_sys* s, int rhs, double complex* state, double t0, double t1);
/* returns 1 if every sub-step converged, 0 otherwise; metric may be NULL */
int orc_sil(const orc_sys* s, int rhs, double complex* state, double t0, double t1, int krylovDim,
            double tol, int maxRestarts, const double* metric);
int orc_strang(const orc_sys* s, double complex* state, double t0, double t1, int krylovDim,
               double tol, int maxRestarts, const double* metric);
/* Strang step with B integrated on `rhsB` (ORC_RHS_NONLIN: the reference's step) */
int orc_strang_split(const orc_sys* s, int rhsB, double complex* state, double t0, double t1, int krylovDim,
                     double tol, int maxRestarts, const double* metric);
/* Adjoint (one body type).  Gradients are with respect to Re(sum_C conj(a) b + sum_orbs sum_x w
 * conj(a) b).  orc_rhs_vjp: grad = J^T lam of the right-hand side rhs (FULL, LIN, NONLIN,
 * NONLIN_CONSISTENT) at (state, t), *ebar = <lam, df/dE>.  orc_strang_cheb: one Strang step with
 * the linear half-steps as a degree-K Chebyshev expansion of exp(-i dt/2 h) on [lo, hi].
 * orc_dipole_gradient: <x> after nSteps such steps and its gradient in the laser coefficients. */
int orc_rhs_vjp(const orc_sys* s, int rhs, const double complex* state, double t, const double complex* lam,
                double complex* grad, double* ebar);
int orc_dipole_grad(const orc_sys* s, const double complex* state, double complex* grad);
int orc_cheb_degree(double tau, double lo, double hi);
int orc_strang_cheb(const orc_sys* s, int rhsB, double complex* state, double t0, double t1, double lo, double hi,
                    int K);
int orc_dipole_gradient(const orc_sys* s, int rhsB, double complex* state, double t0, double dt, int nSteps,
                        double lo, double hi, int K, double* J, double* gradc);
int orc_orthonormalize(const orc_sys* s, double complex* orbs);
int orc_normalize(const orc_sys* s, double complex* coeffs);

/* real symmetric eigen-decomposition (cyclic Jacobi), n <= 64 */
int orc_symeig(int n, const double* a, double* evals, double* evecs);

int orc_num_threads(void);
/* OpenMP threads of later calls (the single-thread CPU baseline sets 1) */
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
