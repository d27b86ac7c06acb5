/*
 * FP64 CPU oracle for the MCTDHX right-hand side and its propagation.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by the smoke
 * entry point and by the CPU-baseline leg of the benchmark; never by the
 * product path.
 *
 * The code restates, in C, the algorithm of the reference Fortran sources
 * cited per function (paths relative to the reference repository root).
 * Deviations from the reference, all deliberate:
 *   - the interaction uses true distances K(|x_i - x_j|) through a dense
 *     matrix, as the FEDVR grid is non-uniform (the reference's Toeplitz
 *     kernel K(|i-j| dx), src/Utils/M_Utils_ConvolutionIntegral.f90:52-55,
 *     is only valid on a uniform grid, where both agree);
 *   - the mean-field potentials of all unordered pairs are formed in one
 *     sweep over the interaction matrix instead of one sweep per pair;
 *   - LAPACK ZHEEV / DSYEV are replaced by a cyclic Jacobi eigensolver; the
 *     consumers (U f(n) U^H and U exp(-i L dt) U^T e1) are gauge invariant.
 * Only fermionic body types are supported (the hot path's configuration).
 */
#include "cody_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define MAXBT 4

typedef struct {
  int nOrbs, nBodies, shift, nConf;
  unsigned* code; /* occupation bit pattern per configuration (lexicographic order) */
  int* rank;      /* bit pattern -> configuration index, -1 if not a configuration */
  int sMax, dMax;
  int *sN, *sExc, *sCode;
  double* sFac;
  int *dN, *dExc, *dCode;
  double* dFac;
} cfg_t;

struct orc_sys {
  int nG, nBT, nO, nC;
  int nBodies[MAXBT], nOrbs[MAXBT], start[MAXBT], stride[MAXBT];
  int* btOf; /* body type of orbital */
  double *x, *w;
  double reg;
  /* kinetic */
  int kin; /* 0 none, 1 FEDVR weak form, 2 9-point FD with zero ends */
  int nE, nLoc, xminExcl, xmaxExcl;
  double elemSize;
  double* kref;
  int *eStart, *eEnd;
  double mass[MAXBT];
  double dx;
  /* potentials */
  double* vext;
  double* vint;
  int nLaser;
  double* laserC;
  double laserT;
  cfg_t cfg[MAXBT];
};

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ======================================================================== */
/* Grid helpers                                                              */
/* ======================================================================== */

/* Legendre P_n and P_n' at x by the three-term recurrence. */
static void legendre(int n, double x, double* p, double* dp) {
  double p0 = 1.0, p1 = x;
  if (n == 0) {
    *p = 1.0;
    *dp = 0.0;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  *p = p1;
  *dp = n * (x * p1 - p0) / (x * x - 1.0);
}

/* Gauss-Legendre-Lobatto nodes/weights on [a,b]: the ends plus the roots of
 * P'_{n-1}, weights 2/(n(n-1) P_{n-1}(x)^2).  Replaces fortran-stdlib's
 * gauss_legendre_lobatto used at src/Utils/M_Utils_Fedvr.f90:127 and
 * src/Utils/M_Utils_DerivativeFedvr.f90:50. */
int orc_gll(int n, double a, double b, double* x, double* w) {
  if (n < 2) return -1;
  int N = n - 1;
  double* t = (double*)malloc(sizeof(double) * n);
  double* wt = (double*)malloc(sizeof(double) * n);
  t[0] = -1.0;
  t[N] = 1.0;
  for (int i = 1; i < N; ++i) {
    /* Chebyshev-Gauss-Lobatto initial guess, Newton on (1-x^2) P_N'(x) */
    double xi = -cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(N, xi, &p, &dp);
      /* f = P_N'(x); f' = (2x P_N' - N(N+1) P_N) / (1 - x^2) */
      double f = dp;
      double fp = (2.0 * xi * dp - N * (N + 1.0) * p) / (1.0 - xi * xi);
      double dx = f / fp;
      xi -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    t[i] = xi;
  }
  for (int i = 0; i <= N; ++i) {
    double p, dp;
    if (i == 0 || i == N) {
      p = (i == 0 && (N % 2)) ? -1.0 : 1.0;
    } else {
      legendre(N, t[i], &p, &dp);
    }
    wt[i] = 2.0 / (N * (N + 1.0) * p * p);
  }
  for (int i = 0; i < n; ++i) {
    x[i] = 0.5 * (b - a) * t[i] + 0.5 * (b + a);
    w[i] = 0.5 * (b - a) * wt[i];
  }
  free(t);
  free(wt);
  return 0;
}

/* src/Utils/M_Utils_Fedvr.f90:108-111 */
int orc_fedvr_npoints(int nE, int nLoc, int xminExcl, int xmaxExcl) {
  int n = nE * (nLoc - 1) + 1;
  if (xminExcl) n -= 1;
  if (xmaxExcl) n -= 1;
  return n;
}

/* Element index ranges and the global grid: src/Utils/M_Utils_Fedvr.f90:76-149
 * (bridge points are shared and their weights summed, :139). */
static int fedvr_layout(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                        double* x, double* w, int* eStart, int* eEnd) {
  double h = (xmax - xmin) / nE;
  int nP = orc_fedvr_npoints(nE, nLoc, xminExcl, xmaxExcl);
  double* ex = (double*)malloc(sizeof(double) * nLoc);
  double* ew = (double*)malloc(sizeof(double) * nLoc);
  if (w)
    for (int i = 0; i < nP; ++i) w[i] = 0.0;
  int ig = -1;
  for (int e = 0; e < nE; ++e) {
    double lb = xmin + e * h, rb = xmin + (e + 1) * h;
    orc_gll(nLoc, lb, rb, ex, ew);
    int started = -1;
    for (int l = 0; l < nLoc; ++l) {
      if (e == 0 && l == 0 && xminExcl) continue;
      if (e == nE - 1 && l == nLoc - 1 && xmaxExcl) continue;
      if (e == 0 || l != 0) ig += 1;
      if (started < 0) started = ig;
      if (x) x[ig] = ex[l];
      if (w) w[ig] += ew[l];
    }
    if (eStart) eStart[e] = started;
    if (eEnd) eEnd[e] = ig;
  }
  free(ex);
  free(ew);
  return nP;
}

int orc_fedvr_grid(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                   double* x, double* w) {
  fedvr_layout(xmin, xmax, nE, nLoc, xminExcl, xmaxExcl, x, w, NULL, NULL);
  return 0;
}

/* dL_j/dxi at node i: src/Utils/M_Utils_DerivativeFedvr.f90:158-181 */
static double lagrange_deriv(const double* x, int j, int i, int n) {
  double r;
  if (j == i) {
    r = 0.0;
    for (int k = 0; k < n; ++k)
      if (k != j) r += 1.0 / (x[i] - x[k]);
  } else {
    r = 1.0 / (x[j] - x[i]);
    for (int k = 0; k < n; ++k)
      if (k != j && k != i) r *= (x[i] - x[k]) / (x[j] - x[k]);
  }
  return r;
}

/* Reference-element operators on [-1/2, 1/2]:
 * D(i,j) = L_j'(x_i); K(i,j) = -sum_k w_k D(k,i) D(k,j)
 * (src/Utils/M_Utils_DerivativeFedvr.f90:29-76). */
int orc_fedvr_kref(int nLoc, double* kref, double* dref) {
  double* xr = (double*)malloc(sizeof(double) * nLoc);
  double* wr = (double*)malloc(sizeof(double) * nLoc);
  double* D = (double*)malloc(sizeof(double) * nLoc * nLoc);
  orc_gll(nLoc, -0.5, 0.5, xr, wr);
  for (int i = 0; i < nLoc; ++i)
    for (int j = 0; j < nLoc; ++j) D[i + j * nLoc] = lagrange_deriv(xr, j, i, nLoc);
  for (int i = 0; i < nLoc; ++i)
    for (int j = 0; j < nLoc; ++j) {
      double s = 0.0;
      for (int k = 0; k < nLoc; ++k) s -= wr[k] * D[k + i * nLoc] * D[k + j * nLoc];
      kref[i + j * nLoc] = s;
    }
  if (dref) {
    /* weak first-derivative form F(i,j) = w_i D(i,j) (:70-74) */
    for (int i = 0; i < nLoc; ++i)
      for (int j = 0; j < nLoc; ++j) dref[i + j * nLoc] = wr[i] * D[i + j * nLoc];
  }
  free(xr);
  free(wr);
  free(D);
  return 0;
}

/* Element-wise weak-form operator, divided by the global weights:
 * src/Utils/M_Utils_DerivativeFedvr.f90:121-152. */
static void fedvr_apply(const double* op, int nLoc, int nE, int xminExcl, int xmaxExcl,
                        const int* eStart, const int* eEnd, double dxPow, const double* gw, int nP,
                        const cplx* f, cplx* res) {
  for (int i = 0; i < nP; ++i) res[i] = 0.0;
  for (int e = 0; e < nE; ++e) {
    int ls = 0, le = nLoc - 1;
    if (xminExcl && e == 0) ls = 1;
    if (xmaxExcl && e == nE - 1) le = nLoc - 2;
    int gs = eStart[e];
    for (int a = ls; a <= le; ++a) {
      cplx acc = 0.0;
      for (int b = ls; b <= le; ++b) acc += op[a + b * nLoc] * f[gs + (b - ls)];
      res[gs + (a - ls)] += acc / dxPow;
    }
    (void)eEnd;
  }
  for (int i = 0; i < nP; ++i) res[i] /= gw[i];
}

static int fedvr_deriv(int order, double xmin, double xmax, int nE, int nLoc, int xminExcl,
                       int xmaxExcl, const cplx* f, cplx* out) {
  int nP = orc_fedvr_npoints(nE, nLoc, xminExcl, xmaxExcl);
  double* x = (double*)malloc(sizeof(double) * nP);
  double* w = (double*)malloc(sizeof(double) * nP);
  int* es = (int*)malloc(sizeof(int) * nE);
  int* ee = (int*)malloc(sizeof(int) * nE);
  double* K = (double*)malloc(sizeof(double) * nLoc * nLoc);
  double* F = (double*)malloc(sizeof(double) * nLoc * nLoc);
  fedvr_layout(xmin, xmax, nE, nLoc, xminExcl, xmaxExcl, x, w, es, ee);
  orc_fedvr_kref(nLoc, K, F);
  double h = (xmax - xmin) / nE;
  if (order == 2)
    fedvr_apply(K, nLoc, nE, xminExcl, xmaxExcl, es, ee, h, w, nP, f, out);
  else
    fedvr_apply(F, nLoc, nE, xminExcl, xmaxExcl, es, ee, 1.0, w, nP, f, out);
  free(x);
  free(w);
  free(es);
  free(ee);
  free(K);
  free(F);
  return 0;
}

int orc_fedvr_d2(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                 const cplx* f, cplx* d2f) {
  return fedvr_deriv(2, xmin, xmax, nE, nLoc, xminExcl, xmaxExcl, f, d2f);
}
int orc_fedvr_d1(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                 const cplx* f, cplx* d1f) {
  return fedvr_deriv(1, xmin, xmax, nE, nLoc, xminExcl, xmaxExcl, f, d1f);
}

/* ======================================================================== */
/* Configuration lists (src/ConfigList/AllActive/...)                        */
/* ======================================================================== */

static int popcount_below(unsigned occ, int j) { return __builtin_popcount(occ & ((1u << j) - 1u)); }

/* a^dag_{cr[0]} ... a^dag_{cr[n-1]} a_{de[n-1]} ... a_{de[0]} |code>:
 * annihilations left to right, creations right to left, sign (-1)^(occupied
 * orbitals below): src/ConfigList/AllActive/Fermions/S_ConfigList_AllActive_Fermions.f90:232-330 */
static int excite(const cfg_t* c, const int* cr, const int* de, int n, int iC, double* fac) {
  unsigned occ = c->code[iC];
  double f = 1.0;
  for (int i = 0; i < n; ++i) {
    int d = de[i];
    if (!(occ & (1u << d))) return -1;
    if (popcount_below(occ, d) & 1) f = -f;
    occ &= ~(1u << d);
  }
  for (int i = n - 1; i >= 0; --i) {
    int a = cr[i];
    if (occ & (1u << a)) return -1;
    if (popcount_below(occ, a) & 1) f = -f;
    occ |= (1u << a);
  }
  *fac = f;
  return c->rank[occ];
}

static int binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return (int)r;
}

/* Lexicographic k-combinations (GSL order, src/Utils/M_Utils_CombinationGslLib.f90:53-86)
 * and the singles/doubles connectivity (src/ConfigList/AllActive/S_ConfigList_AllActive.f90:141-349). */
static int cfg_build(cfg_t* c, int nOrbs, int nBodies, int shift, int nO) {
  c->nOrbs = nOrbs;
  c->nBodies = nBodies;
  c->shift = shift;
  c->nConf = binom(nOrbs, nBodies);
  c->code = (unsigned*)malloc(sizeof(unsigned) * c->nConf);
  c->rank = (int*)malloc(sizeof(int) * (1u << nOrbs));
  for (unsigned i = 0; i < (1u << nOrbs); ++i) c->rank[i] = -1;
  int comb[32];
  for (int i = 0; i < nBodies; ++i) comb[i] = i;
  for (int iC = 0; iC < c->nConf; ++iC) {
    unsigned code = 0;
    for (int i = 0; i < nBodies; ++i) code |= 1u << comb[i];
    c->code[iC] = code;
    c->rank[code] = iC;
    int i = nBodies - 1;
    while (i >= 0 && comb[i] == nOrbs - nBodies + i) --i;
    if (i < 0) break;
    comb[i] += 1;
    for (int j = i + 1; j < nBodies; ++j) comb[j] = comb[j - 1] + 1;
  }
  /* singles: loop j1 outer, i1 inner (:165-175) */
  c->sN = (int*)calloc(c->nConf, sizeof(int));
  c->sMax = 0;
  for (int iC = 0; iC < c->nConf; ++iC) {
    int k = 0;
    for (int j1 = 0; j1 < nOrbs; ++j1)
      for (int i1 = 0; i1 < nOrbs; ++i1) {
        double f;
        if (excite(c, &i1, &j1, 1, iC, &f) >= 0) ++k;
      }
    c->sN[iC] = k;
    if (k > c->sMax) c->sMax = k;
  }
  c->sExc = (int*)malloc(sizeof(int) * c->sMax * c->nConf);
  c->sCode = (int*)malloc(sizeof(int) * c->sMax * c->nConf);
  c->sFac = (double*)malloc(sizeof(double) * c->sMax * c->nConf);
  for (int iC = 0; iC < c->nConf; ++iC) {
    int k = 0;
    for (int j1 = 0; j1 < nOrbs; ++j1)
      for (int i1 = 0; i1 < nOrbs; ++i1) {
        double f;
        int jn = excite(c, &i1, &j1, 1, iC, &f);
        if (jn < 0) continue;
        c->sExc[k + iC * c->sMax] = jn;
        c->sCode[k + iC * c->sMax] = (i1 + shift) * nO + (j1 + shift);
        c->sFac[k + iC * c->sMax] = f;
        ++k;
      }
  }
  /* doubles: j2, j1<=j2, i2, i1<=i2 (:288-329) */
  c->dN = (int*)calloc(c->nConf, sizeof(int));
  c->dMax = 0;
  for (int iC = 0; iC < c->nConf; ++iC) {
    int k = 0;
    for (int j2 = 0; j2 < nOrbs; ++j2)
      for (int j1 = 0; j1 <= j2; ++j1)
        for (int i2 = 0; i2 < nOrbs; ++i2)
          for (int i1 = 0; i1 <= i2; ++i1) {
            int cr[2] = {i1, i2}, de[2] = {j1, j2};
            double f;
            if (excite(c, cr, de, 2, iC, &f) >= 0) ++k;
          }
    c->dN[iC] = k;
    if (k > c->dMax) c->dMax = k;
  }
  int dm = c->dMax > 0 ? c->dMax : 1;
  c->dExc = (int*)malloc(sizeof(int) * dm * c->nConf);
  c->dCode = (int*)malloc(sizeof(int) * dm * c->nConf);
  c->dFac = (double*)malloc(sizeof(double) * dm * c->nConf);
  for (int iC = 0; iC < c->nConf; ++iC) {
    int k = 0;
    for (int j2 = 0; j2 < nOrbs; ++j2)
      for (int j1 = 0; j1 <= j2; ++j1)
        for (int i2 = 0; i2 < nOrbs; ++i2)
          for (int i1 = 0; i1 <= i2; ++i1) {
            int cr[2] = {i1, i2}, de[2] = {j1, j2};
            double f;
            int jn = excite(c, cr, de, 2, iC, &f);
            if (jn < 0) continue;
            c->dExc[k + iC * dm] = jn;
            c->dCode[k + iC * dm] =
                (((i1 + shift) * nO + (j1 + shift)) * nO + (i2 + shift)) * nO + (j2 + shift);
            c->dFac[k + iC * dm] = f;
            ++k;
          }
  }
  c->dMax = dm;
  return 0;
}

static void cfg_free(cfg_t* c) {
  free(c->code);
  free(c->rank);
  free(c->sN);
  free(c->sExc);
  free(c->sCode);
  free(c->sFac);
  free(c->dN);
  free(c->dExc);
  free(c->dCode);
  free(c->dFac);
}

int orc_config_codes(const orc_sys* s, int bt, unsigned* codes) {
  if (bt < 0 || bt >= s->nBT) return -1;
  memcpy(codes, s->cfg[bt].code, sizeof(unsigned) * s->cfg[bt].nConf);
  return s->cfg[bt].nConf;
}

/* ======================================================================== */
/* System                                                                    */
/* ======================================================================== */

orc_sys* orc_create(int nG, const double* x, const double* w, int nBT, const int* nBodies,
                    const int* nOrbs, double reg) {
  if (nBT < 1 || nBT > MAXBT) return NULL;
  orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
  s->nG = nG;
  s->nBT = nBT;
  s->reg = reg;
  s->x = (double*)malloc(sizeof(double) * nG);
  s->w = (double*)malloc(sizeof(double) * nG);
  memcpy(s->x, x, sizeof(double) * nG);
  memcpy(s->w, w, sizeof(double) * nG);
  int nO = 0;
  for (int b = 0; b < nBT; ++b) {
    if (nOrbs[b] > 22 || nBodies[b] > nOrbs[b]) {
      free(s);
      return NULL;
    }
    s->nBodies[b] = nBodies[b];
    s->nOrbs[b] = nOrbs[b];
    s->start[b] = nO;
    nO += nOrbs[b];
    s->mass[b] = 1.0;
  }
  s->nO = nO;
  s->btOf = (int*)malloc(sizeof(int) * nO);
  for (int b = 0; b < nBT; ++b)
    for (int i = 0; i < nOrbs[b]; ++i) s->btOf[s->start[b] + i] = b;
  /* CI dimension = product over body types (src/Coeffs/Generic/S_Coeffs_Generic.f90:62-66);
   * body type 0 varies fastest (:131-151) */
  int nC = 1;
  for (int b = 0; b < nBT; ++b) {
    cfg_build(&s->cfg[b], nOrbs[b], nBodies[b], s->start[b], nO);
    s->stride[b] = nC;
    nC *= s->cfg[b].nConf;
  }
  s->nC = nC;
  return s;
}

void orc_destroy(orc_sys* s) {
  if (!s) return;
  for (int b = 0; b < s->nBT; ++b) cfg_free(&s->cfg[b]);
  free(s->btOf);
  free(s->x);
  free(s->w);
  free(s->kref);
  free(s->eStart);
  free(s->eEnd);
  free(s->vext);
  free(s->vint);
  free(s->laserC);
  free(s);
}

int orc_ncoeffs(const orc_sys* s) { return s->nC; }
int orc_norbs(const orc_sys* s) { return s->nO; }

int orc_set_kinetic_fedvr(orc_sys* s, double xmin, double xmax, int nE, int nLoc, int xminExcl,
                          int xmaxExcl, const double* mass) {
  if (orc_fedvr_npoints(nE, nLoc, xminExcl, xmaxExcl) != s->nG) return -1;
  s->kin = 1;
  s->nE = nE;
  s->nLoc = nLoc;
  s->xminExcl = xminExcl;
  s->xmaxExcl = xmaxExcl;
  s->elemSize = (xmax - xmin) / nE;
  free(s->kref);
  free(s->eStart);
  free(s->eEnd);
  s->kref = (double*)malloc(sizeof(double) * nLoc * nLoc);
  s->eStart = (int*)malloc(sizeof(int) * nE);
  s->eEnd = (int*)malloc(sizeof(int) * nE);
  orc_fedvr_kref(nLoc, s->kref, NULL);
  fedvr_layout(xmin, xmax, nE, nLoc, xminExcl, xmaxExcl, NULL, NULL, s->eStart, s->eEnd);
  for (int b = 0; b < s->nBT; ++b) s->mass[b] = mass ? mass[b] : 1.0;
  return 0;
}

int orc_set_kinetic_fd(orc_sys* s, double dx, const double* mass) {
  s->kin = 2;
  s->dx = dx;
  for (int b = 0; b < s->nBT; ++b) s->mass[b] = mass ? mass[b] : 1.0;
  return 0;
}

int orc_set_potential(orc_sys* s, const double* vext) {
  free(s->vext);
  s->vext = (double*)malloc(sizeof(double) * s->nG);
  memcpy(s->vext, vext, sizeof(double) * s->nG);
  return 0;
}

int orc_set_interaction(orc_sys* s, const double* vint) {
  free(s->vint);
  size_t n = (size_t)s->nG * s->nG;
  s->vint = (double*)malloc(sizeof(double) * n);
  memcpy(s->vint, vint, sizeof(double) * n);
  return 0;
}

int orc_set_laser(orc_sys* s, int nC, const double* c, double T) {
  free(s->laserC);
  s->laserC = NULL;
  s->nLaser = nC;
  s->laserT = T;
  if (nC > 0) {
    s->laserC = (double*)malloc(sizeof(double) * nC);
    memcpy(s->laserC, c, sizeof(double) * nC);
  }
  return 0;
}

static double laser_field(const orc_sys* s, double t) {
  if (s->nLaser <= 0 || t < 0.0 || t > s->laserT) return 0.0;
  double env = sin(M_PI * t / s->laserT);
  env *= env;
  double e = 0.0;
  for (int k = 0; k < s->nLaser; ++k) e += s->laserC[k] * sin((k + 1) * M_PI * t / s->laserT);
  return e * env;
}

/* ======================================================================== */
/* CI: H-apply with simultaneous RDM fill (src/Coeffs/Generic/S_Coeffs_Generic.f90) */
/* ======================================================================== */

static void local_configs(const orc_sys* s, int iC, int* loc) {
  for (int b = 0; b < s->nBT; ++b) loc[b] = (iC / s->stride[b]) % s->cfg[b].nConf;
}

/* :103-178.  rdm1Lin(code(i,j)) += f conj(C_j) C_i ; dC_i += f h1Lin C_j */
static void apply_h1_fill_rdm1(const orc_sys* s, const cplx* C, cplx* rdm1Lin, cplx* dC,
                               const cplx* h1Lin) {
  int nO2 = s->nO * s->nO;
  if (rdm1Lin) memset(rdm1Lin, 0, sizeof(cplx) * nO2);
  for (int iC = 0; iC < s->nC; ++iC) {
    int loc[MAXBT];
    local_configs(s, iC, loc);
    for (int b = 0; b < s->nBT; ++b) {
      const cfg_t* c = &s->cfg[b];
      int il = loc[b];
      for (int k = 0; k < c->sN[il]; ++k) {
        int code = c->sCode[k + il * c->sMax];
        double f = c->sFac[k + il * c->sMax];
        int jC = iC + (c->sExc[k + il * c->sMax] - il) * s->stride[b];
        if (dC && h1Lin) dC[iC] += f * h1Lin[code] * C[jC];
        if (rdm1Lin) rdm1Lin[code] += f * conj(C[jC]) * C[iC];
      }
    }
  }
}

/* :197-319.  intra-body doubles and inter-body singles x singles */
static void apply_h2_fill_rdm2(const orc_sys* s, const cplx* C, cplx* rdm2Lin, cplx* dC,
                               const cplx* h2Lin) {
  int nO = s->nO, nO2 = nO * nO;
  if (rdm2Lin) memset(rdm2Lin, 0, sizeof(cplx) * nO2 * nO2);
  for (int iC = 0; iC < s->nC; ++iC) {
    int loc[MAXBT];
    local_configs(s, iC, loc);
    for (int b = 0; b < s->nBT; ++b) {
      const cfg_t* c = &s->cfg[b];
      int il = loc[b];
      for (int k = 0; k < c->dN[il]; ++k) {
        int code = c->dCode[k + il * c->dMax];
        double f = c->dFac[k + il * c->dMax];
        int jC = iC + (c->dExc[k + il * c->dMax] - il) * s->stride[b];
        if (dC && h2Lin) dC[iC] += f * h2Lin[code] * C[jC];
        if (rdm2Lin) rdm2Lin[code] += f * conj(C[jC]) * C[iC];
      }
      if (b == s->nBT - 1) continue;
      for (int k1 = 0; k1 < c->sN[il]; ++k1) {
        int code1 = c->sCode[k1 + il * c->sMax] * nO2;
        double f1 = c->sFac[k1 + il * c->sMax];
        int jCt = iC + (c->sExc[k1 + il * c->sMax] - il) * s->stride[b];
        for (int b2 = b + 1; b2 < s->nBT; ++b2) {
          const cfg_t* c2 = &s->cfg[b2];
          int il2 = loc[b2];
          for (int k2 = 0; k2 < c2->sN[il2]; ++k2) {
            int code = code1 + c2->sCode[k2 + il2 * c2->sMax];
            double f = f1 * c2->sFac[k2 + il2 * c2->sMax];
            int jC = jCt + (c2->sExc[k2 + il2 * c2->sMax] - il2) * s->stride[b2];
            if (dC && h2Lin) dC[iC] += f * h2Lin[code] * C[jC];
            if (rdm2Lin) rdm2Lin[code] += f * conj(C[jC]) * C[iC];
          }
        }
      }
    }
  }
}

#define IDX4(a, b, c, d, n) ((a) + (n) * ((b) + (n) * ((c) + (n) * (d))))

/* :325-349 */
static void rdm1_unpack(const orc_sys* s, const cplx* lin, cplx* rdm1) {
  int nO = s->nO;
  for (int j = 0; j < nO; ++j)
    for (int i = 0; i < nO; ++i) rdm1[i + j * nO] = lin[i * nO + j];
}

/* :362-434 (fermions, and the antisymmetric inter-body convention) */
static void rdm2_unpack(const orc_sys* s, const cplx* lin, cplx* rdm2) {
  int nO = s->nO;
  memset(rdm2, 0, sizeof(cplx) * nO * nO * nO * nO);
  for (int j2 = 0; j2 < nO; ++j2)
    for (int j1 = 0; j1 <= j2; ++j1)
      for (int i2 = 0; i2 < nO; ++i2)
        for (int i1 = 0; i1 <= i2; ++i1) {
          int bi1 = s->btOf[i1], bi2 = s->btOf[i2], bj1 = s->btOf[j1], bj2 = s->btOf[j2];
          if (bi1 == bi2 && bj1 != bj2) continue;
          if (bi1 != bi2 && bj1 == bj2) continue;
          cplx v = lin[((i1 * nO + j1) * nO + i2) * nO + j2];
          rdm2[IDX4(i1, i2, j1, j2, nO)] += v;
          rdm2[IDX4(i2, i1, j1, j2, nO)] -= v;
          rdm2[IDX4(i1, i2, j2, j1, nO)] -= v;
          rdm2[IDX4(i2, i1, j2, j1, nO)] += v;
        }
}

/* :440-464 */
static void h1_pack(const orc_sys* s, const cplx* h1, cplx* lin) {
  int nO = s->nO;
  for (int j = 0; j < nO; ++j)
    for (int i = 0; i < nO; ++i) lin[i * nO + j] = h1[i + j * nO];
}

/* :478-542 */
static void h2_pack(const orc_sys* s, const cplx* h2, cplx* lin) {
  int nO = s->nO;
  memset(lin, 0, sizeof(cplx) * nO * nO * nO * nO);
  for (int j2 = 0; j2 < nO; ++j2)
    for (int j1 = 0; j1 <= j2; ++j1)
      for (int i2 = 0; i2 < nO; ++i2)
        for (int i1 = 0; i1 <= i2; ++i1) {
          int bi1 = s->btOf[i1], bi2 = s->btOf[i2], bj1 = s->btOf[j1], bj2 = s->btOf[j2];
          if (bi1 == bi2 && bj1 != bj2) continue;
          if (bi1 != bi2 && bj1 == bj2) continue;
          int code = ((i1 * nO + j1) * nO + i2) * nO + j2;
          cplx v = h2[IDX4(i1, i2, j1, j2, nO)] - h2[IDX4(i1, i2, j2, j1, nO)] -
                   h2[IDX4(i2, i1, j1, j2, nO)] + h2[IDX4(i2, i1, j2, j1, nO)];
          if (i1 == i2 || j1 == j2) v = 0.0;
          lin[code] = v;
        }
}

int orc_rdm1(const orc_sys* s, const cplx* C, cplx* rdm1) {
  int nO2 = s->nO * s->nO;
  cplx* lin = (cplx*)malloc(sizeof(cplx) * nO2);
  apply_h1_fill_rdm1(s, C, lin, NULL, NULL);
  rdm1_unpack(s, lin, rdm1);
  free(lin);
  return 0;
}

int orc_rdm2(const orc_sys* s, const cplx* C, cplx* rdm2) {
  int n4 = s->nO * s->nO * s->nO * s->nO;
  cplx* lin = (cplx*)malloc(sizeof(cplx) * n4);
  apply_h2_fill_rdm2(s, C, lin, NULL, NULL);
  rdm2_unpack(s, lin, rdm2);
  free(lin);
  return 0;
}

int orc_apply_h1(const orc_sys* s, const cplx* C, const cplx* h1, cplx* dC) {
  cplx* lin = (cplx*)malloc(sizeof(cplx) * s->nO * s->nO);
  h1_pack(s, h1, lin);
  apply_h1_fill_rdm1(s, C, NULL, dC, lin);
  free(lin);
  return 0;
}

int orc_apply_h2(const orc_sys* s, const cplx* C, const cplx* h2, cplx* dC) {
  int n4 = s->nO * s->nO * s->nO * s->nO;
  cplx* lin = (cplx*)malloc(sizeof(cplx) * n4);
  h2_pack(s, h2, lin);
  apply_h2_fill_rdm2(s, C, NULL, dC, lin);
  free(lin);
  return 0;
}

/* ======================================================================== */
/* One-body operators                                                        */
/* ======================================================================== */

static cplx inner(const orc_sys* s, const cplx* a, const cplx* b) {
  /* src/Grid/Linear/S_Grid_Linear.f90:53-61 */
  cplx acc = 0.0;
  for (int i = 0; i < s->nG; ++i) acc += conj(a[i]) * b[i] * s->w[i];
  return acc;
}

/* 9-point central second difference with zero ends:
 * src/Utils/M_Utils_DerivativeFinDiff.f90:67-71,179-198 */
static void fd_d2_zero_ends(const orc_sys* s, const cplx* f, cplx* out) {
  static const double c[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
  int n = s->nG;
  for (int i = 0; i < n; ++i) {
    cplx acc = c[0] * f[i];
    for (int k = 1; k <= 4; ++k) {
      cplx l = (i - k >= 0) ? f[i - k] : 0.0;
      cplx r = (i + k < n) ? f[i + k] : 0.0;
      acc += c[k] * (l + r);
    }
    out[i] = acc / (s->dx * s->dx);
  }
}

/* T phi = -(1/2m) d2 phi: src/SysKinetic/Linear/Laplacian/FinDiff/S_SysKinetic_Linear_Laplacian_FinDiff.f90:62-91
 * with the FEDVR weak-form second derivative (src/Utils/M_Utils_DerivativeFedvr.f90:110-152)
 * in place of the finite difference on FEDVR grids. */
static void kinetic(const orc_sys* s, const cplx* f, int bt, cplx* out) {
  int n = s->nG;
  if (s->kin == 1) {
    fedvr_apply(s->kref, s->nLoc, s->nE, s->xminExcl, s->xmaxExcl, s->eStart, s->eEnd,
                s->elemSize, s->w, n, f, out);
  } else if (s->kin == 2) {
    fd_d2_zero_ends(s, f, out);
  } else {
    for (int i = 0; i < n; ++i) out[i] = 0.0;
    return;
  }
  double fac = -0.5 / s->mass[bt];
  for (int i = 0; i < n; ++i) out[i] *= fac;
}

/* dOrbs += (T + V_ext(t)) orbs ; h1(i,j) = <phi_i|dOrbs_j> (same body type):
 * src/Method/Mb/OrbBased/S_Method_Mb_OrbBased.f90:976-1034, :332-375, :425-488 */
int orc_apply_single_body(const orc_sys* s, const cplx* orbs, double t, cplx* dorbs, cplx* h1) {
  int nG = s->nG, nO = s->nO;
  double E = laser_field(s, t);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nO; ++i) {
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * nG);
    kinetic(s, orbs + (size_t)i * nG, s->btOf[i], tmp);
    for (int g = 0; g < nG; ++g) {
      double v = (s->vext ? s->vext[g] : 0.0) + E * s->x[g];
      dorbs[(size_t)i * nG + g] += tmp[g] + v * orbs[(size_t)i * nG + g];
    }
    free(tmp);
  }
  if (h1) {
    for (int j = 0; j < nO; ++j)
      for (int i = 0; i < nO; ++i) {
        h1[i + j * nO] = 0.0;
        if (s->btOf[i] != s->btOf[j]) continue;
        h1[i + j * nO] = inner(s, orbs + (size_t)i * nG, dorbs + (size_t)j * nG);
      }
  }
  return 0;
}

/* :164-224 */
int orc_fill_h1(const orc_sys* s, const cplx* orbs, double t, cplx* h1) {
  cplx* tmp = (cplx*)calloc((size_t)s->nG * s->nO, sizeof(cplx));
  orc_apply_single_body(s, orbs, t, tmp, h1);
  free(tmp);
  return 0;
}

/* ======================================================================== */
/* Interaction: mean-field potentials of the unordered pairs                 */
/* ======================================================================== */

/* W_{lk}(x_i) = sum_j K(x_i,x_j) conj(phi_l(x_j)) phi_k(x_j) w_j for l <= k, pair index
 * p enumerating (l=0..nO-1, k=l..nO-1).  Source: src/SysInteraction/Linear/S_SysInteraction_Linear.f90:57-72;
 * potential: src/Utils/M_Utils_ConvolutionIntegral.f90:71-107 (true distances, see header).
 * Output w is nG x P, column-major.  Pairs of different body types are computed too
 * (callers skip them). */
/* Zero sources: ConvolutionIntegral_Apply (src/Utils/M_Utils_ConvolutionIntegral.f90:84-93) first
 * collects the source points with |src(j)| > srcThreshold (default 0) and sums over those only.
 * With compaction on (the default) the sweep below leaves out every point j where the sources of
 * all pairs are exactly zero; the reference compacts each pair's source separately, which leaves
 * out the same points when the orbitals share their support (as in every workload here).  The
 * skipped terms are exact zeros, so W is the same to the last bit either way; off restores the
 * dense sweep (for timing). */
static int g_compact = 1;
void orc_set_compaction(int on) { g_compact = on != 0; }

/* Number of source points the sweep of orc_meanfield_w visits for these orbitals. */
int orc_active_sources(const orc_sys* s, const cplx* orbs) {
  int n = 0;
  for (int g = 0; g < s->nG; ++g) {
    int any = !g_compact;
    for (int l = 0; l < s->nO && !any; ++l)
      for (int k = l; k < s->nO && !any; ++k)
        any = conj(orbs[(size_t)l * s->nG + g]) * orbs[(size_t)k * s->nG + g] * s->w[g] != 0.0;
    n += any;
  }
  return n;
}

int orc_meanfield_w(const orc_sys* s, const cplx* orbs, cplx* W) {
  int nG = s->nG, nO = s->nO, P = nO * (nO + 1) / 2;
  cplx* src = (cplx*)malloc(sizeof(cplx) * (size_t)nG * P); /* [g][p] */
  int* active = (int*)malloc(sizeof(int) * (size_t)nG);
  int nAct = 0;
  for (int g = 0; g < nG; ++g) {
    int p = 0, any = !g_compact;
    for (int l = 0; l < nO; ++l)
      for (int k = l; k < nO; ++k, ++p) {
        src[(size_t)g * P + p] = conj(orbs[(size_t)l * nG + g]) * orbs[(size_t)k * nG + g] * s->w[g];
        any |= src[(size_t)g * P + p] != 0.0;
      }
    if (any) active[nAct++] = g;
  }
#pragma omp parallel
  {
    cplx* acc = (cplx*)malloc(sizeof(cplx) * P);
#pragma omp for schedule(static)
    for (int i = 0; i < nG; ++i) {
      for (int p = 0; p < P; ++p) acc[p] = 0.0;
      const double* row = s->vint + (size_t)i * nG;
      for (int jj = 0; jj < nAct; ++jj) {
        const int j = active[jj];
        double v = row[j];
        const cplx* sj = src + (size_t)j * P;
        for (int p = 0; p < P; ++p) acc[p] += v * sj[p];
      }
      for (int p = 0; p < P; ++p) W[(size_t)p * nG + i] = acc[p];
    }
    free(acc);
  }
  free(active);
  free(src);
  return 0;
}

static int pair_index(int nO, int l, int k) {
  /* l <= k */
  return l * nO - l * (l - 1) / 2 + (k - l);
}

/* ======================================================================== */
/* Small dense linear algebra                                                */
/* ======================================================================== */

/* cyclic Jacobi for a real symmetric n x n matrix (column-major) */
int orc_symeig(int n, const double* a, double* ev, double* V) {
  double* A = (double*)malloc(sizeof(double) * n * n);
  memcpy(A, a, sizeof(double) * n * n);
  for (int i = 0; i < n * n; ++i) V[i] = 0.0;
  for (int i = 0; i < n; ++i) V[i + i * n] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += A[i + j * n] * A[i + j * n];
        if (i != j) off += A[i + j * n] * A[i + j * n];
      }
    if (off <= 1e-34 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p + q * n];
        if (apq == 0.0) continue;
        double app = A[p + p * n], aqq = A[q + q * n];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = c * akp - sn * akq;
          A[k + q * n] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = c * apk - sn * aqk;
          A[q + k * n] = sn * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = c * vkp - sn * vkq;
          V[k + q * n] = sn * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) ev[i] = A[i + i * n];
  free(A);
  return 0;
}

/* regInvMat = U diag(n/(n^2+eps)) U^H for rho1 = U diag(n) U^H:
 * src/Method/Mb/OrbBased/S_Method_Mb_OrbBased.f90:733-744 with the eigen-decomposition of
 * src/Utils/M_Utils_RdmDiagonalize.f90:17-36.  The Hermitian nO x nO problem is solved through
 * its real 2nO x 2nO embedding [[Re, -Im], [Im, Re]]; f(embedding) = embedding of f(rho1). */
int orc_reg_inverse(const orc_sys* s, const cplx* rdm1, cplx* R) {
  int n = s->nO, m = 2 * n;
  double* E = (double*)calloc((size_t)m * m, sizeof(double));
  double* V = (double*)malloc(sizeof(double) * m * m);
  double* ev = (double*)malloc(sizeof(double) * m);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double re = creal(rdm1[i + j * n]), im = cimag(rdm1[i + j * n]);
      E[i + j * m] = re;
      E[(i + n) + (j + n) * m] = re;
      E[(i + n) + j * m] = im;
      E[i + (j + n) * m] = -im;
    }
  orc_symeig(m, E, ev, V);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double re = 0.0, im = 0.0;
      for (int k = 0; k < m; ++k) {
        double f = ev[k] / (ev[k] * ev[k] + s->reg);
        re += V[i + k * m] * f * V[j + k * m];
        im += V[(i + n) + k * m] * f * V[j + k * m];
      }
      R[i + j * n] = re + I * im;
    }
  free(E);
  free(V);
  free(ev);
  return 0;
}

/* ======================================================================== */
/* Correlation operator                                                      */
/* ======================================================================== */

/* src/Method/Mb/OrbBased/S_Method_Mb_OrbBased.f90:647-805 with AccumulateCorrelation :808-838
 * and FillH2CorrelationBlock :841-872 (unrestricted: rdm*spatial = rdm*). */
int orc_apply_correlation(const orc_sys* s, const cplx* orbs, const cplx* rdm1, const cplx* rdm2,
                          cplx* dorbs, cplx* h2) {
  int nG = s->nG, nO = s->nO, P = nO * (nO + 1) / 2;
  size_t n4 = (size_t)nO * nO * nO * nO;
  cplx* R = (cplx*)malloc(sizeof(cplx) * nO * nO);
  cplx* rho2p = (cplx*)calloc(n4, sizeof(cplx));
  orc_reg_inverse(s, rdm1, R);
  /* rho2'(i1,i2,j1,l2) = sum_k1 rho2(i1,i2,k1,l2) R(k1,j1)  (:746-753) */
  for (int l2 = 0; l2 < nO; ++l2)
    for (int j1 = 0; j1 < nO; ++j1)
      for (int k1 = 0; k1 < nO; ++k1) {
        cplx r = R[k1 + j1 * nO];
        for (int i2 = 0; i2 < nO; ++i2)
          for (int i1 = 0; i1 < nO; ++i1)
            rho2p[IDX4(i1, i2, j1, l2, nO)] += rdm2[IDX4(i1, i2, k1, l2, nO)] * r;
      }
  if (h2) memset(h2, 0, sizeof(cplx) * n4);
  cplx* W = (cplx*)malloc(sizeof(cplx) * (size_t)nG * P);
  orc_meanfield_w(s, orbs, W);
  /* pair loop l2 <= k2 (:762-789); orbTmps(:,k1) = W_{l2k2} phi_k1, swapped pair uses conj(W) */
  for (int l2 = 0; l2 < nO; ++l2)
    for (int k2 = l2; k2 < nO; ++k2) {
      if (s->btOf[l2] != s->btOf[k2]) continue;
      const cplx* Wp = W + (size_t)pair_index(nO, l2, k2) * nG;
      for (int swap = 0; swap < (k2 == l2 ? 1 : 2); ++swap) {
        int a = swap ? l2 : k2; /* AccumulateCorrelation(..., a, b) */
        int b = swap ? k2 : l2;
#pragma omp parallel for schedule(static)
        for (int j1 = 0; j1 < nO; ++j1) {
          for (int k1 = 0; k1 < nO; ++k1) {
            if (s->btOf[j1] != s->btOf[k1]) continue;
            cplx r = rho2p[IDX4(k1, a, j1, b, nO)];
            const cplx* pk = orbs + (size_t)k1 * nG;
            cplx* dj = dorbs + (size_t)j1 * nG;
            for (int g = 0; g < nG; ++g) {
              cplx wv = swap ? conj(Wp[g]) : Wp[g];
              dj[g] += wv * pk[g] * r;
            }
          }
        }
        if (h2) {
          /* h2(i1,b,j1,a) = 1/2 <phi_i1 | W_{b a} phi_j1>  (:841-872) */
          for (int j1 = 0; j1 < nO; ++j1)
            for (int i1 = 0; i1 < nO; ++i1) {
              if (s->btOf[i1] != s->btOf[j1]) continue;
              const cplx* pi = orbs + (size_t)i1 * nG;
              const cplx* pj = orbs + (size_t)j1 * nG;
              cplx acc = 0.0;
              for (int g = 0; g < nG; ++g) {
                cplx wv = swap ? conj(Wp[g]) : Wp[g];
                acc += conj(pi[g]) * wv * pj[g] * s->w[g];
              }
              h2[IDX4(i1, b, j1, a, nO)] = 0.5 * acc;
            }
        }
      }
    }
  free(W);
  free(R);
  free(rho2p);
  return 0;
}

/* FillH2: src/Method/Mb/OrbBased/S_Method_Mb_OrbBased.f90:227-329 */
int orc_fill_h2(const orc_sys* s, const cplx* orbs, cplx* h2) {
  int nG = s->nG, nO = s->nO, P = nO * (nO + 1) / 2;
  memset(h2, 0, sizeof(cplx) * nO * nO * nO * nO);
  cplx* W = (cplx*)malloc(sizeof(cplx) * (size_t)nG * P);
  orc_meanfield_w(s, orbs, W);
  for (int j2 = 0; j2 < nO; ++j2)
    for (int i2 = j2; i2 < nO; ++i2) {
      if (s->btOf[i2] != s->btOf[j2]) continue;
      /* source conj(phi_j2) phi_i2 for the stored pair (j2 <= i2); the block (i2,j2) uses its conjugate */
      const cplx* Wp = W + (size_t)pair_index(nO, j2, i2) * nG;
      for (int swap = 0; swap < (i2 == j2 ? 1 : 2); ++swap) {
        int a = swap ? j2 : i2, b = swap ? i2 : j2; /* block (a,b): potential from conj(phi_a) phi_b */
        int conjW = !swap;                          /* (i2,j2) = conj of stored (j2,i2) */
        if (i2 == j2) conjW = 0;
        for (int j1 = 0; j1 < nO; ++j1)
          for (int i1 = 0; i1 < nO; ++i1) {
            if (s->btOf[i1] != s->btOf[j1]) continue;
            cplx acc = 0.0;
            for (int g = 0; g < nG; ++g) {
              cplx wv = conjW ? conj(Wp[g]) : Wp[g];
              acc += conj(orbs[(size_t)i1 * nG + g]) * wv * orbs[(size_t)j1 * nG + g] * s->w[g];
            }
            h2[IDX4(i1, a, j1, b, nO)] = 0.5 * acc;
          }
      }
    }
  free(W);
  return 0;
}

/* ======================================================================== */
/* Projectors                                                                */
/* ======================================================================== */

/* (1-P) per body type: src/Orbs/S_Orbs.f90:198-225 -> src/Grid/S_Grid.f90:170-203 */
int orc_project_orbs(const orc_sys* s, const cplx* orbs, cplx* dorbs) {
  int nG = s->nG;
  for (int b = 0; b < s->nBT; ++b) {
    int st = s->start[b], n = s->nOrbs[b];
    cplx* inp = (cplx*)malloc(sizeof(cplx) * n * n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        inp[i + j * n] = inner(s, orbs + (size_t)(st + i) * nG, dorbs + (size_t)(st + j) * nG);
    for (int j = 0; j < n; ++j)
      for (int g = 0; g < nG; ++g) {
        cplx acc = 0.0;
        for (int i = 0; i < n; ++i) acc += orbs[(size_t)(st + i) * nG + g] * inp[i + j * n];
        dorbs[(size_t)(st + j) * nG + g] -= acc;
      }
    free(inp);
  }
  return 0;
}

/* dC -= <C|dC> C : src/Coeffs/S_Coeffs.f90:105-111 */
static void project_coeffs(const orc_sys* s, const cplx* C, cplx* dC) {
  cplx ov = 0.0;
  for (int i = 0; i < s->nC; ++i) ov += conj(C[i]) * dC[i];
  for (int i = 0; i < s->nC; ++i) dC[i] -= ov * C[i];
}

/* ======================================================================== */
/* Method: time derivatives and energy                                       */
/* ======================================================================== */

/* Full MCTDHX RHS: src/Method/Mb/OrbBased/Mctdhx/S_Method_Mb_OrbBased_Mctdhx.f90:101-149 */
int orc_time_derivative(const orc_sys* s, const cplx* state, double t, cplx* dstate) {
  int nC = s->nC, nO = s->nO, nG = s->nG;
  size_t n4 = (size_t)nO * nO * nO * nO;
  size_t nS = (size_t)nC + (size_t)nG * nO;
  memset(dstate, 0, sizeof(cplx) * nS);
  const cplx* C = state;
  const cplx* orbs = state + nC;
  cplx* dC = dstate;
  cplx* dorbs = dstate + nC;
  cplx* rdm1 = (cplx*)malloc(sizeof(cplx) * nO * nO);
  cplx* rdm2 = (cplx*)malloc(sizeof(cplx) * n4);
  cplx* h1 = (cplx*)malloc(sizeof(cplx) * nO * nO);
  cplx* h2 = (cplx*)malloc(sizeof(cplx) * n4);
  orc_rdm1(s, C, rdm1);
  orc_rdm2(s, C, rdm2);
  orc_apply_single_body(s, orbs, t, dorbs, h1);
  orc_apply_correlation(s, orbs, rdm1, rdm2, dorbs, h2);
  orc_apply_h1(s, C, h1, dC);
  orc_apply_h2(s, C, h2, dC);
  project_coeffs(s, C, dC);
  orc_project_orbs(s, orbs, dorbs);
  for (size_t i = 0; i < nS; ++i) dstate[i] *= -I;
  free(rdm1);
  free(rdm2);
  free(h1);
  free(h2);
  return 0;
}

/* Linear split part: :152-177 */
int orc_time_derivative_orbs_lin(const orc_sys* s, const cplx* state, double t, cplx* dstate) {
  int nC = s->nC;
  size_t nS = (size_t)nC + (size_t)s->nG * s->nO;
  memset(dstate, 0, sizeof(cplx) * nS);
  orc_apply_single_body(s, state + nC, t, dstate + nC, NULL);
  for (size_t i = 0; i < nS; ++i) dstate[i] *= -I;
  return 0;
}

/* Nonlinear split part: :180-226 */
int orc_time_derivative_nonlin(const orc_sys* s, const cplx* state, double t, cplx* dstate) {
  int nC = s->nC, nO = s->nO, nG = s->nG;
  size_t n4 = (size_t)nO * nO * nO * nO;
  size_t nS = (size_t)nC + (size_t)nG * nO;
  memset(dstate, 0, sizeof(cplx) * nS);
  const cplx* C = state;
  const cplx* orbs = state + nC;
  cplx* rdm1 = (cplx*)malloc(sizeof(cplx) * nO * nO);
  cplx* rdm2 = (cplx*)malloc(sizeof(cplx) * n4);
  cplx* h1 = (cplx*)malloc(sizeof(cplx) * nO * nO);
  cplx* h2 = (cplx*)malloc(sizeof(cplx) * n4);
  orc_rdm1(s, C, rdm1);
  orc_rdm2(s, C, rdm2);
  orc_fill_h1(s, orbs, t, h1);
  orc_apply_correlation(s, orbs, rdm1, rdm2, dstate + nC, h2);
  orc_apply_h1(s, C, h1, dstate);
  orc_apply_h2(s, C, h2, dstate);
  project_coeffs(s, C, dstate);
  orc_project_orbs(s, orbs, dstate + nC);
  for (size_t i = 0; i < nS; ++i) dstate[i] *= -I;
  free(rdm1);
  free(rdm2);
  free(h1);
  free(h2);
  return 0;
}

/* E = sum h1(k,l) rdm1(l,k) + sum h2(k1,k2,l1,l2) rdm2(l1,l2,k1,k2):
 * src/Method/Mb/OrbBased/S_Method_Mb_OrbBased.f90:132-161, src/Utils/M_Utils_RdmObservables.f90:17-52 */
double orc_energy(const orc_sys* s, const cplx* state, double t) {
  int nC = s->nC, nO = s->nO;
  size_t n4 = (size_t)nO * nO * nO * nO;
  cplx* rdm1 = (cplx*)malloc(sizeof(cplx) * nO * nO);
  cplx* rdm2 = (cplx*)malloc(sizeof(cplx) * n4);
  cplx* h1 = (cplx*)malloc(sizeof(cplx) * nO * nO);
  cplx* h2 = (cplx*)malloc(sizeof(cplx) * n4);
  orc_rdm1(s, state, rdm1);
  orc_rdm2(s, state, rdm2);
  orc_fill_h1(s, state + nC, t, h1);
  orc_fill_h2(s, state + nC, h2);
  cplx e = 0.0;
  for (int k = 0; k < nO; ++k)
    for (int l = 0; l < nO; ++l) e += h1[k + l * nO] * rdm1[l + k * nO];
  for (int k1 = 0; k1 < nO; ++k1)
    for (int k2 = 0; k2 < nO; ++k2)
      for (int l1 = 0; l1 < nO; ++l1)
        for (int l2 = 0; l2 < nO; ++l2)
          e += h2[IDX4(k1, k2, l1, l2, nO)] * rdm2[IDX4(l1, l2, k1, k2, nO)];
  free(rdm1);
  free(rdm2);
  free(h1);
  free(h2);
  return creal(e);
}

/* <x> = sum_kl rdm1(l,k) <phi_k|x|phi_l>: the density-based dipole of
 * src/Utils/M_Utils_PrinterObservableLinear.f90:25-103 in the orbital basis. */
double orc_dipole(const orc_sys* s, const cplx* state) {
  int nC = s->nC, nO = s->nO, nG = s->nG;
  cplx* rdm1 = (cplx*)malloc(sizeof(cplx) * nO * nO);
  orc_rdm1(s, state, rdm1);
  const cplx* orbs = state + nC;
  cplx d = 0.0;
  for (int k = 0; k < nO; ++k)
    for (int l = 0; l < nO; ++l) {
      if (s->btOf[k] != s->btOf[l]) continue;
      cplx acc = 0.0;
      for (int g = 0; g < nG; ++g)
        acc += conj(orbs[(size_t)k * nG + g]) * s->x[g] * orbs[(size_t)l * nG + g] * s->w[g];
      d += acc * rdm1[l + k * nO];
    }
  free(rdm1);
  return creal(d);
}

/* ======================================================================== */
/* Integrators                                                               */
/* ======================================================================== */

static int eval_rhs(const orc_sys* s, int rhs, const cplx* state, double t, cplx* d) {
  switch (rhs) {
    case ORC_RHS_FULL:
      return orc_time_derivative(s, state, t, d);
    case ORC_RHS_LIN:
      return orc_time_derivative_orbs_lin(s, state, t, d);
    case ORC_RHS_NONLIN:
      return orc_time_derivative_nonlin(s, state, t, d);
    case ORC_RHS_NONLIN_CONSISTENT: {
      size_t nS = (size_t)s->nC + (size_t)s->nG * s->nO;
      cplx* lin = (cplx*)malloc(sizeof(cplx) * nS);
      if (!lin) return -1;
      int r = orc_time_derivative(s, state, t, d);
      r |= orc_time_derivative_orbs_lin(s, state, t, lin);
      for (size_t i = 0; i < nS; ++i) d[i] -= lin[i];
      free(lin);
      return r;
    }
    case ORC_RHS_IMAG: {
      /* src/Utils/M_Utils_Actions.f90:66-78 */
      int r = orc_time_derivative(s, state, t, d);
      size_t nS = (size_t)s->nC + (size_t)s->nG * s->nO;
      for (size_t i = 0; i < nS; ++i) d[i] *= -I;
      return r;
    }
  }
  return -1;
}

static size_t state_size(const orc_sys* s) { return (size_t)s->nC + (size_t)s->nG * s->nO; }

/* classical RK4: src/IntegratorList/Rk/O4Expl/S_IntegratorList_Rk_O4Expl.f90:90-141 */
int orc_rk4(const orc_sys* s, int rhs, cplx* state, double t0, double t1) {
  size_t n = state_size(s);
  cplx* d = (cplx*)malloc(sizeof(cplx) * n);
  cplx* inc = (cplx*)malloc(sizeof(cplx) * n);
  cplx* tmp = (cplx*)malloc(sizeof(cplx) * n);
  double dt = t1 - t0, dt2 = dt / 2.0, dt3 = dt / 3.0, dt6 = dt / 6.0;
  eval_rhs(s, rhs, state, t0, d);
  for (size_t i = 0; i < n; ++i) {
    inc[i] = d[i] * dt6;
    tmp[i] = state[i] + d[i] * dt2;
  }
  eval_rhs(s, rhs, tmp, t0 + dt2, d);
  for (size_t i = 0; i < n; ++i) {
    inc[i] += d[i] * dt3;
    tmp[i] = state[i] + d[i] * dt2;
  }
  eval_rhs(s, rhs, tmp, t0 + dt2, d);
  for (size_t i = 0; i < n; ++i) {
    inc[i] += d[i] * dt3;
    tmp[i] = state[i] + d[i] * dt;
  }
  eval_rhs(s, rhs, tmp, t0 + dt, d);
  for (size_t i = 0; i < n; ++i) state[i] += inc[i] + d[i] * dt6;
  free(d);
  free(inc);
  free(tmp);
  return 0;
}

/* explicit Euler (propagator.single + rk o1Expl in the imaginary-time driver) */
int orc_rk1(const orc_sys* s, int rhs, cplx* state, double t0, double t1) {
  size_t n = state_size(s);
  cplx* d = (cplx*)malloc(sizeof(cplx) * n);
  eval_rhs(s, rhs, state, t0, d);
  for (size_t i = 0; i < n; ++i) state[i] += d[i] * (t1 - t0);
  free(d);
  return 0;
}

static cplx mdot(size_t n, const double* m, const cplx* a, const cplx* b) {
  cplx acc = 0.0;
  if (m)
    for (size_t i = 0; i < n; ++i) acc += conj(a[i]) * m[i] * b[i];
  else
    for (size_t i = 0; i < n; ++i) acc += conj(a[i]) * b[i];
  return acc;
}

/* phi = U exp(-i dt L) U^T e1, err = beta_m |phi_m|: src/IntegratorList/Sil/S_IntegratorList_Sil.f90:277-310 */
static double sil_subspace(int m, const double* alpha, const double* beta, double dt, cplx* phi) {
  double* T = (double*)calloc(m * m, sizeof(double));
  double* ev = (double*)malloc(sizeof(double) * m);
  double* U = (double*)malloc(sizeof(double) * m * m);
  for (int j = 0; j < m; ++j) T[j + j * m] = alpha[j];
  for (int j = 0; j < m - 1; ++j) {
    T[j + (j + 1) * m] = beta[j];
    T[(j + 1) + j * m] = beta[j];
  }
  orc_symeig(m, T, ev, U);
  for (int i = 0; i < m; ++i) {
    cplx acc = 0.0;
    for (int j = 0; j < m; ++j) acc += U[i + j * m] * U[0 + j * m] * cexp(-I * dt * ev[j]);
    phi[i] = acc;
  }
  double err = beta[m - 1] * cabs(phi[m - 1]);
  free(T);
  free(ev);
  free(U);
  return err;
}

/* Short iterative Lanczos with full re-orthogonalisation and step halving:
 * src/IntegratorList/Sil/S_IntegratorList_Sil.f90:114-268 */
int orc_sil(const orc_sys* s, int rhs, cplx* state, double t0, double t1, int krylovDim,
            double tol, int maxRestarts, const double* metric) {
  const double zt = 1.0e-14;
  if (fabs(t1 - t0) <= zt) return 1;
  size_t n = state_size(s);
  int maxDim = krylovDim < (int)n ? krylovDim : (int)n;
  if (maxDim < 1) maxDim = 1;
  cplx* basis = (cplx*)malloc(sizeof(cplx) * n * maxDim);
  cplx* work = (cplx*)malloc(sizeof(cplx) * n);
  cplx* trial = (cplx*)malloc(sizeof(cplx) * n);
  cplx* d = (cplx*)malloc(sizeof(cplx) * n);
  cplx* wv = (cplx*)malloc(sizeof(cplx) * n);
  cplx* phi = (cplx*)malloc(sizeof(cplx) * maxDim);
  double* alpha = (double*)malloc(sizeof(double) * maxDim);
  double* beta = (double*)malloc(sizeof(double) * maxDim);
  memcpy(work, state, sizeof(cplx) * n);
  double tc = t0, rem = t1 - t0;
  int allConv = 1;
  while (fabs(rem) > zt) {
    double dts = rem;
    int restarts = 0, conv = 0;
    for (;;) {
      /* PerformLanczosStep (:207-268) */
      double nrm = sqrt(creal(mdot(n, metric, work, work)));
      double err = 0.0;
      int mDim = maxDim;
      if (nrm <= zt) {
        memcpy(trial, work, sizeof(cplx) * n);
        conv = 1;
      } else {
        for (size_t i = 0; i < n; ++i) basis[i] = work[i] / nrm;
        double tev = tc + 0.5 * dts;
        for (int j = 0; j < maxDim; ++j) {
          cplx* v = basis + (size_t)j * n;
          eval_rhs(s, rhs, v, tev, d);
          for (size_t i = 0; i < n; ++i) wv[i] = I * d[i];
          if (j > 0)
            for (size_t i = 0; i < n; ++i) wv[i] -= beta[j - 1] * basis[(size_t)(j - 1) * n + i];
          alpha[j] = creal(mdot(n, metric, v, wv));
          for (size_t i = 0; i < n; ++i) wv[i] -= alpha[j] * v[i];
          for (int k = 0; k <= j; ++k) {
            cplx* bk = basis + (size_t)k * n;
            cplx ov = mdot(n, metric, bk, wv);
            for (size_t i = 0; i < n; ++i) wv[i] -= ov * bk[i];
          }
          beta[j] = sqrt(creal(mdot(n, metric, wv, wv)));
          int breakdown = beta[j] <= zt;
          err = sil_subspace(j + 1, alpha, beta, dts, phi);
          if (breakdown) err = 0.0;
          conv = err <= tol;
          if (conv || breakdown || j == maxDim - 1) {
            mDim = j + 1;
            break;
          }
          for (size_t i = 0; i < n; ++i) basis[(size_t)(j + 1) * n + i] = wv[i] / beta[j];
        }
        for (size_t i = 0; i < n; ++i) {
          cplx acc = 0.0;
          for (int j = 0; j < mDim; ++j) acc += basis[(size_t)j * n + i] * phi[j];
          trial[i] = nrm * acc;
        }
      }
      if (conv) break;
      if (fabs(dts) <= zt || restarts >= maxRestarts) break;
      dts *= 0.5;
      restarts++;
    }
    if (!conv) allConv = 0;
    memcpy(work, trial, sizeof(cplx) * n);
    tc += dts;
    rem -= dts;
  }
  memcpy(state, work, sizeof(cplx) * n);
  free(basis);
  free(work);
  free(trial);
  free(d);
  free(wv);
  free(phi);
  free(alpha);
  free(beta);
  return allConv;
}

/* Strang A(1/2) B(1) A(1/2) with A = SIL on the linear part and B = RK4 on the nonlinear part:
 * src/Propagator/SplitStep/Order2/S_Propagator_SplitStep_Order2.f90:52-77 */
int orc_strang_split(const orc_sys* s, int rhsB, cplx* state, double t0, double t1, int krylovDim, double tol,
                     int maxRestarts, const double* metric) {
  double dt = t1 - t0, t2 = t0 + 0.5 * dt;
  int ok = orc_sil(s, ORC_RHS_LIN, state, t0, t0 + 0.5 * dt, krylovDim, tol, maxRestarts, metric);
  orc_rk4(s, rhsB, state, t0, t0 + dt);
  ok &= orc_sil(s, ORC_RHS_LIN, state, t2, t2 + 0.5 * dt, krylovDim, tol, maxRestarts, metric);
  return ok;
}

int orc_strang(const orc_sys* s, cplx* state, double t0, double t1, int krylovDim, double tol,
               int maxRestarts, const double* metric) {
  return orc_strang_split(s, ORC_RHS_NONLIN, state, t0, t1, krylovDim, tol, maxRestarts, metric);
}

/* MGS twice per body type: src/Grid/S_Grid.f90:120-151 */
int orc_orthonormalize(const orc_sys* s, cplx* orbs) {
  int nG = s->nG;
  for (int b = 0; b < s->nBT; ++b) {
    int st = s->start[b];
    for (int i = 0; i < s->nOrbs[b]; ++i) {
      cplx* fi = orbs + (size_t)(st + i) * nG;
      for (int pass = 0; pass < 2; ++pass)
        for (int j = 0; j < i; ++j) {
          cplx* fj = orbs + (size_t)(st + j) * nG;
          cplx ov = inner(s, fj, fi);
          for (int g = 0; g < nG; ++g) fi[g] -= fj[g] * ov;
        }
      double nrm = creal(inner(s, fi, fi));
      if (!(nrm > 0.0)) return -1;
      double sc = 1.0 / sqrt(nrm);
      for (int g = 0; g < nG; ++g) fi[g] *= sc;
    }
  }
  return 0;
}

/* src/Coeffs/S_Coeffs.f90:84-93 */
int orc_normalize(const orc_sys* s, cplx* C) {
  double n2 = 0.0;
  for (int i = 0; i < s->nC; ++i) n2 += creal(conj(C[i]) * C[i]);
  double nrm = sqrt(n2);
  for (int i = 0; i < s->nC; ++i) C[i] /= nrm;
  return 0;
}

/* ======================================================================== */
/* Adjoint: vector-Jacobian products and the gradient of <x>(T)               */
/* ======================================================================== */
/* The reference has no adjoint.  What follows is the transpose, for one body type, of the maps
 * above: the right-hand sides of orc_time_derivative / _nonlin / _orbs_lin (and their
 * difference), the dipole, a Strang step whose linear half-steps are a fixed-degree Chebyshev
 * expansion of exp(-i tau h(E)), and its checkpoint-free discrete adjoint.  Gradients are taken
 * with respect to the real inner product
 *   <a, b> = Re( sum_C conj(a) b + sum_orbitals sum_x w(x) conj(a(x)) b(x) ),
 * in which h = T + V_ext + E x is self-adjoint. */

/* (T + V_ext + E x) f for one orbital (one body type) */
static void onebody(const orc_sys* s, const cplx* f, double E, cplx* out) {
  kinetic(s, f, 0, out);
  for (int g = 0; g < s->nG; ++g) out[g] += ((s->vext ? s->vext[g] : 0.0) + E * s->x[g]) * f[g];
}

/* out[jC] += f conj(hLin[code]) D[iC] over the singles and doubles tables: the conjugate
 * transpose of the H-apply of apply_h1_fill_rdm1 / apply_h2_fill_rdm2 (one body type) */
static void ci_apply_t(const orc_sys* s, const cplx* D, const cplx* h1Lin, const cplx* h2Lin, cplx* out) {
  const cfg_t* c = &s->cfg[0];
  for (int iC = 0; iC < s->nC; ++iC) {
    for (int k = 0; k < c->sN[iC]; ++k) {
      int q = k + iC * c->sMax;
      out[c->sExc[q]] += c->sFac[q] * conj(h1Lin[c->sCode[q]]) * D[iC];
    }
    if (!h2Lin) continue;
    for (int k = 0; k < c->dN[iC]; ++k) {
      int q = k + iC * c->dMax;
      out[c->dExc[q]] += c->dFac[q] * conj(h2Lin[c->dCode[q]]) * D[iC];
    }
  }
}

/* transition densities lin[code] = sum f conj(bra[jC]) ket[iC] (the fill of apply_h*_fill_rdm*
 * with two vectors; bra = ket gives rdm1Lin / rdm2Lin) */
static void ci_transition(const orc_sys* s, const cplx* bra, const cplx* ket, cplx* lin1, cplx* lin2) {
  const cfg_t* c = &s->cfg[0];
  int nO2 = s->nO * s->nO;
  memset(lin1, 0, sizeof(cplx) * nO2);
  memset(lin2, 0, sizeof(cplx) * nO2 * nO2);
  for (int iC = 0; iC < s->nC; ++iC) {
    for (int k = 0; k < c->sN[iC]; ++k) {
      int q = k + iC * c->sMax;
      lin1[c->sCode[q]] += c->sFac[q] * conj(bra[c->sExc[q]]) * ket[iC];
    }
    for (int k = 0; k < c->dN[iC]; ++k) {
      int q = k + iC * c->dMax;
      lin2[c->dCode[q]] += c->dFac[q] * conj(bra[c->dExc[q]]) * ket[iC];
    }
  }
}

/* gC += d/dC of Re sum conj(rho1bar) rho1 + Re sum conj(rho2bar) rho2 (either may be NULL) */
static void rdm_grad(const orc_sys* s, const cplx* C, const cplx* rho1bar, const cplx* rho2bar, cplx* gC) {
  int nO = s->nO;
  size_t n4 = (size_t)nO * nO * nO * nO;
  cplx* l1 = (cplx*)calloc(nO * nO, sizeof(cplx));
  cplx* l2 = (cplx*)calloc(n4, sizeof(cplx));
  if (rho1bar) h1_pack(s, rho1bar, l1); /* the adjoint of the unpack permutation */
  if (rho2bar) h2_pack(s, rho2bar, l2); /* the antisymmetrising pack is the unpack's transpose */
  apply_h1_fill_rdm1(s, C, NULL, gC, l1);
  apply_h2_fill_rdm2(s, C, NULL, gC, l2);
  ci_apply_t(s, C, l1, l2, gC);
  free(l1);
  free(l2);
}

/* X = A^{-1} B for a small complex n x n A (column-major), Gauss-Jordan with partial pivoting */
static int csolve(int n, const cplx* A, const cplx* B, cplx* X) {
  cplx* M = (cplx*)malloc(sizeof(cplx) * n * n);
  memcpy(M, A, sizeof(cplx) * n * n);
  memcpy(X, B, sizeof(cplx) * n * n);
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (cabs(M[i + k * n]) > cabs(M[piv + k * n])) piv = i;
    if (cabs(M[piv + k * n]) == 0.0) {
      free(M);
      return -1;
    }
    if (piv != k)
      for (int j = 0; j < n; ++j) {
        cplx tm = M[k + j * n];
        M[k + j * n] = M[piv + j * n];
        M[piv + j * n] = tm;
        tm = X[k + j * n];
        X[k + j * n] = X[piv + j * n];
        X[piv + j * n] = tm;
      }
    cplx inv = 1.0 / M[k + k * n];
    for (int j = 0; j < n; ++j) {
      M[k + j * n] *= inv;
      X[k + j * n] *= inv;
    }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      cplx f = M[i + k * n];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        M[i + j * n] -= f * M[k + j * n];
        X[i + j * n] -= f * X[k + j * n];
      }
    }
  }
  free(M);
  return 0;
}

/* W_ab(x) = V[conj(phi_b) phi_a w](x) for ordered (a, b) from the unordered pairs of orc_meanfield_w */
static cplx w_ab(const cplx* W, int nO, int nG, int a, int b, int g) {
  return a >= b ? W[(size_t)pair_index(nO, b, a) * nG + g] : conj(W[(size_t)pair_index(nO, a, b) * nG + g]);
}

/* grad = J^T lam for the right-hand side `rhs` (ORC_RHS_FULL, _LIN, _NONLIN, _NONLIN_CONSISTENT)
 * at (state, t), and *ebar = <lam, df/dE>. */
int orc_rhs_vjp(const orc_sys* s, int rhs, const cplx* state, double t, const cplx* lam, cplx* grad,
                double* ebar) {
  if (s->nBT != 1) return -1;
  if (rhs != ORC_RHS_FULL && rhs != ORC_RHS_LIN && rhs != ORC_RHS_NONLIN && rhs != ORC_RHS_NONLIN_CONSISTENT)
    return -1;
  const int nC = s->nC, nO = s->nO, nG = s->nG, P = nO * (nO + 1) / 2;
  const size_t n2 = (size_t)nO * nO, n4 = n2 * n2, nOG = (size_t)nO * nG;
  const cplx* C = state;
  const cplx* orbs = state + nC;
  const double E = laser_field(s, t);
  cplx* gC = grad;
  cplx* gO = grad + nC;
  memset(grad, 0, sizeof(cplx) * (nC + nOG));
  cplx* LO = (cplx*)malloc(sizeof(cplx) * nOG);
  cplx* hphi = (cplx*)malloc(sizeof(cplx) * nOG);
  cplx* hbar = (cplx*)calloc(nOG, sizeof(cplx));
  cplx* tmp = (cplx*)malloc(sizeof(cplx) * nG);
  for (size_t i = 0; i < nOG; ++i) LO[i] = I * lam[nC + i]; /* the adjoint of the factor -i */
  for (int j = 0; j < nO; ++j) onebody(s, orbs + (size_t)j * nG, E, hphi + (size_t)j * nG);
  if (rhs == ORC_RHS_LIN) {
    memcpy(hbar, LO, sizeof(cplx) * nOG);
  } else {
    const double ah = rhs == ORC_RHS_NONLIN ? 0.0 : 1.0, bh = rhs == ORC_RHS_NONLIN_CONSISTENT ? -1.0 : 0.0;
    cplx* rdm1 = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* rdm2 = (cplx*)malloc(sizeof(cplx) * n4);
    cplx* R = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* rho2p = (cplx*)calloc(n4, sizeof(cplx));
    cplx* h1 = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* h2 = (cplx*)malloc(sizeof(cplx) * n4);
    cplx* W = (cplx*)malloc(sizeof(cplx) * (size_t)nG * P);
    cplx* draw = (cplx*)calloc(nOG, sizeof(cplx));
    cplx* inp = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* h1Lin = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* h2Lin = (cplx*)malloc(sizeof(cplx) * n4);
    cplx* dCraw = (cplx*)calloc(nC, sizeof(cplx));
    cplx* D = (cplx*)malloc(sizeof(cplx) * nC);
    cplx* l1 = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* l2 = (cplx*)malloc(sizeof(cplx) * n4);
    cplx* h1b = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* h2b = (cplx*)malloc(sizeof(cplx) * n4);
    cplx* B = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* dbar = (cplx*)malloc(sizeof(cplx) * nOG);
    cplx* Wbar = (cplx*)calloc(n2 * nG, sizeof(cplx));
    cplx* r2pb = (cplx*)calloc(n4, sizeof(cplx));
    cplx* r2b = (cplx*)calloc(n4, sizeof(cplx));
    cplx* Rb = (cplx*)calloc(n2, sizeof(cplx));
    cplx* A = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* Q = (cplx*)malloc(sizeof(cplx) * n2);
    cplx* r1b = (cplx*)malloc(sizeof(cplx) * n2);
    /* forward quantities */
    orc_rdm1(s, C, rdm1);
    orc_rdm2(s, C, rdm2);
    orc_reg_inverse(s, rdm1, R);
    for (int l2 = 0; l2 < nO; ++l2)
      for (int j1 = 0; j1 < nO; ++j1)
        for (int k1 = 0; k1 < nO; ++k1)
          for (int i2 = 0; i2 < nO; ++i2)
            for (int i1 = 0; i1 < nO; ++i1)
              rho2p[IDX4(i1, i2, j1, l2, nO)] += rdm2[IDX4(i1, i2, k1, l2, nO)] * R[k1 + j1 * nO];
    for (int j = 0; j < nO; ++j)
      for (int i = 0; i < nO; ++i) h1[i + j * nO] = inner(s, orbs + (size_t)i * nG, hphi + (size_t)j * nG);
    orc_meanfield_w(s, orbs, W);
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < nO; ++a)
      for (int b = 0; b < nO; ++b)
        for (int j1 = 0; j1 < nO; ++j1)
          for (int i1 = 0; i1 < nO; ++i1) {
            cplx acc = 0.0;
            for (int g = 0; g < nG; ++g)
              acc += conj(orbs[(size_t)i1 * nG + g]) * w_ab(W, nO, nG, a, b, g) * orbs[(size_t)j1 * nG + g] * s->w[g];
            h2[IDX4(i1, b, j1, a, nO)] = 0.5 * acc;
          }
#pragma omp parallel for schedule(static)
    for (int g = 0; g < nG; ++g)
      for (int a = 0; a < nO; ++a)
        for (int b = 0; b < nO; ++b) {
          cplx wv = w_ab(W, nO, nG, a, b, g);
          for (int j = 0; j < nO; ++j) {
            cplx acc = 0.0;
            for (int k1 = 0; k1 < nO; ++k1) acc += orbs[(size_t)k1 * nG + g] * rho2p[IDX4(k1, a, j, b, nO)];
            draw[(size_t)j * nG + g] += wv * acc;
          }
        }
    for (size_t i = 0; i < nOG; ++i) draw[i] += ah * hphi[i];
    for (int j = 0; j < nO; ++j)
      for (int i = 0; i < nO; ++i) inp[i + j * nO] = inner(s, orbs + (size_t)i * nG, draw + (size_t)j * nG);
    h1_pack(s, h1, h1Lin);
    h2_pack(s, h2, h2Lin);
    apply_h1_fill_rdm1(s, C, NULL, dCraw, h1Lin);
    apply_h2_fill_rdm2(s, C, NULL, dCraw, h2Lin);
    cplx ov = 0.0, alpha = 0.0;
    for (int i = 0; i < nC; ++i) ov += conj(C[i]) * dCraw[i];
    /* CI projection and H-apply (dC = -i [(1 - C C^H) H(h1, h2) C]) */
    for (int i = 0; i < nC; ++i) alpha += conj(C[i]) * (I * lam[i]);
    for (int i = 0; i < nC; ++i) {
      cplx LC = I * lam[i];
      D[i] = LC - alpha * C[i];
      gC[i] = -conj(alpha) * dCraw[i] - conj(ov) * LC;
    }
    ci_apply_t(s, D, h1Lin, h2Lin, gC);
    ci_transition(s, C, D, l1, l2);
    rdm1_unpack(s, l1, h1b);
    rdm2_unpack(s, l2, h2b);
    /* orbital projection dphi = -i [(1 - P) draw + bh hphi] */
    for (int j = 0; j < nO; ++j)
      for (int i = 0; i < nO; ++i) B[i + j * nO] = inner(s, orbs + (size_t)i * nG, LO + (size_t)j * nG);
    for (int j = 0; j < nO; ++j)
      for (int g = 0; g < nG; ++g) {
        cplx acc = LO[(size_t)j * nG + g];
        for (int i = 0; i < nO; ++i) acc -= orbs[(size_t)i * nG + g] * B[i + j * nO];
        dbar[(size_t)j * nG + g] = acc;
      }
    for (int i = 0; i < nO; ++i)
      for (int g = 0; g < nG; ++g) {
        cplx acc = 0.0;
        for (int j = 0; j < nO; ++j)
          acc += conj(inp[i + j * nO]) * LO[(size_t)j * nG + g] + conj(B[i + j * nO]) * draw[(size_t)j * nG + g];
        gO[(size_t)i * nG + g] -= acc;
      }
    /* h1(i,j) = <phi_i|hphi_j> and the draw's h phi term */
    for (int j = 0; j < nO; ++j)
      for (int g = 0; g < nG; ++g) {
        cplx acc = ah * dbar[(size_t)j * nG + g] + bh * LO[(size_t)j * nG + g];
        for (int i = 0; i < nO; ++i) acc += h1b[i + j * nO] * orbs[(size_t)i * nG + g];
        hbar[(size_t)j * nG + g] = acc;
      }
    for (int i = 0; i < nO; ++i)
      for (int g = 0; g < nG; ++g) {
        cplx acc = 0.0;
        for (int j = 0; j < nO; ++j) acc += conj(h1b[i + j * nO]) * hphi[(size_t)j * nG + g];
        gO[(size_t)i * nG + g] += acc;
      }
    /* correlation sum_{a,b,k1} W_ab phi_k1 rho2'(k1,a,j1,b) and h2(i1,b,j1,a) = 1/2 <phi_i1|W_ab phi_j1> */
#pragma omp parallel for schedule(static)
    for (int g = 0; g < nG; ++g) {
      for (int a = 0; a < nO; ++a)
        for (int b = 0; b < nO; ++b) {
          cplx wv = w_ab(W, nO, nG, a, b, g), acc = 0.0;
          for (int k1 = 0; k1 < nO; ++k1) {
            cplx pk = orbs[(size_t)k1 * nG + g];
            for (int j1 = 0; j1 < nO; ++j1) {
              cplx dj = dbar[(size_t)j1 * nG + g], pj = orbs[(size_t)j1 * nG + g];
              cplx r = conj(rho2p[IDX4(k1, a, j1, b, nO)]);
              acc += r * conj(pk) * dj + 0.5 * h2b[IDX4(k1, b, j1, a, nO)] * pk * conj(pj);
              gO[(size_t)k1 * nG + g] += conj(wv) * r * dj + 0.5 * conj(h2b[IDX4(k1, b, j1, a, nO)]) * wv * pj;
              gO[(size_t)j1 * nG + g] += 0.5 * h2b[IDX4(k1, b, j1, a, nO)] * conj(wv) * pk;
            }
          }
          Wbar[(a + b * nO) * (size_t)nG + g] = acc;
        }
    }
    /* rho2'bar(k1,a,j1,b) = sum_x w conj(W_ab phi_k1) dbar_j1 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < nO; ++a)
      for (int b = 0; b < nO; ++b)
        for (int k1 = 0; k1 < nO; ++k1)
          for (int j1 = 0; j1 < nO; ++j1) {
            cplx acc = 0.0;
            for (int g = 0; g < nG; ++g)
              acc += s->w[g] * conj(w_ab(W, nO, nG, a, b, g) * orbs[(size_t)k1 * nG + g]) * dbar[(size_t)j1 * nG + g];
            r2pb[IDX4(k1, a, j1, b, nO)] = acc;
          }
    /* rho2' = rho2 x_3 R */
    for (int l2 = 0; l2 < nO; ++l2)
      for (int j1 = 0; j1 < nO; ++j1)
        for (int k = 0; k < nO; ++k)
          for (int i2 = 0; i2 < nO; ++i2)
            for (int i1 = 0; i1 < nO; ++i1) {
              cplx pb = r2pb[IDX4(i1, i2, j1, l2, nO)];
              r2b[IDX4(i1, i2, k, l2, nO)] += pb * conj(R[k + j1 * nO]);
              Rb[k + j1 * nO] += pb * conj(rdm2[IDX4(i1, i2, k, l2, nO)]);
            }
    /* R = (rho1^2 + eps)^{-1} rho1: rho1bar = Q - Q R^H rho1 - rho1 Q R^H, Q = (rho1^2 + eps)^{-1} Rbar */
    for (int j = 0; j < nO; ++j)
      for (int i = 0; i < nO; ++i) {
        cplx acc = i == j ? s->reg : 0.0;
        for (int k = 0; k < nO; ++k) acc += rdm1[i + k * nO] * rdm1[k + j * nO];
        A[i + j * nO] = acc;
      }
    csolve(nO, A, Rb, Q);
    for (int j = 0; j < nO; ++j)
      for (int i = 0; i < nO; ++i) {
        cplx acc = Q[i + j * nO];
        for (int k = 0; k < nO; ++k)
          for (int m = 0; m < nO; ++m) {
            acc -= Q[i + k * nO] * conj(R[m + k * nO]) * rdm1[m + j * nO];
            acc -= rdm1[i + k * nO] * Q[k + m * nO] * conj(R[j + m * nO]);
          }
        r1b[i + j * nO] = acc;
      }
    rdm_grad(s, C, r1b, r2b, gC);
    /* W_kl = V [conj(phi_l) phi_k w] for the stored pairs (l <= k): sbar = V (w Wbar) */
    for (int l = 0; l < nO; ++l)
      for (int k = l; k < nO; ++k) {
        for (int g = 0; g < nG; ++g) {
          cplx v = Wbar[(k + l * nO) * (size_t)nG + g];
          if (k != l) v += conj(Wbar[(l + k * nO) * (size_t)nG + g]);
          tmp[g] = s->w[g] * v;
        }
#pragma omp parallel for schedule(static)
        for (int g = 0; g < nG; ++g) {
          cplx sb = 0.0;
          const double* row = s->vint + (size_t)g * nG;
          for (int i = 0; i < nG; ++i) sb += row[i] * tmp[i];
          gO[(size_t)k * nG + g] += sb * orbs[(size_t)l * nG + g];
          gO[(size_t)l * nG + g] += conj(sb) * orbs[(size_t)k * nG + g];
        }
      }
    free(rdm1), free(rdm2), free(R), free(rho2p), free(h1), free(h2), free(W), free(draw), free(inp);
    free(h1Lin), free(h2Lin), free(dCraw), free(D), free(l1), free(l2), free(h1b), free(h2b), free(B);
    free(dbar), free(Wbar), free(r2pb), free(r2b), free(Rb), free(A), free(Q), free(r1b);
  }
  /* hphi = (T + V + E x) phi: self-adjoint in the weighted product */
  double eb = 0.0;
  for (int j = 0; j < nO; ++j) {
    onebody(s, hbar + (size_t)j * nG, E, tmp);
    for (int g = 0; g < nG; ++g) {
      gO[(size_t)j * nG + g] += tmp[g];
      eb += s->w[g] * s->x[g] * creal(conj(hbar[(size_t)j * nG + g]) * orbs[(size_t)j * nG + g]);
    }
  }
  *ebar = eb;
  free(LO), free(hphi), free(hbar), free(tmp);
  return 0;
}

/* gradient of <x> = sum_kl rdm1(l,k) <phi_k|x|phi_l> (orc_dipole) */
int orc_dipole_grad(const orc_sys* s, const cplx* state, cplx* grad) {
  if (s->nBT != 1) return -1;
  const int nC = s->nC, nO = s->nO, nG = s->nG;
  const cplx* orbs = state + nC;
  cplx* rdm1 = (cplx*)malloc(sizeof(cplx) * nO * nO);
  cplx* X = (cplx*)malloc(sizeof(cplx) * nO * nO);
  orc_rdm1(s, state, rdm1);
  for (int l = 0; l < nO; ++l)
    for (int k = 0; k < nO; ++k) {
      cplx acc = 0.0;
      for (int g = 0; g < nG; ++g)
        acc += conj(orbs[(size_t)k * nG + g]) * s->x[g] * orbs[(size_t)l * nG + g] * s->w[g];
      X[k + l * nO] = acc;
    }
  memset(grad, 0, sizeof(cplx) * (nC + (size_t)nO * nG));
  for (int k = 0; k < nO; ++k)
    for (int g = 0; g < nG; ++g) {
      cplx acc = 0.0;
      for (int l = 0; l < nO; ++l) acc += rdm1[l + k * nO] * orbs[(size_t)l * nG + g];
      grad[nC + (size_t)k * nG + g] = 2.0 * s->x[g] * acc;
    }
  rdm_grad(s, state, X, NULL, grad); /* rho1bar(l,k) = conj(X(k,l)) = X(l,k) */
  free(rdm1);
  free(X);
  return 0;
}

/* Bessel J_0..J_K(z), z > 0: Miller's backward recurrence normalised by J_0 + 2 sum J_2k = 1 */
static void bessel_j(int K, double z, double* J) {
  int M = K > (int)z ? K : (int)z;
  M = 2 * ((M + 20 + (int)sqrt(40.0 * (M + 1))) / 2);
  double jp = 0.0, jc = 1e-300, norm = 0.0;
  double* v = (double*)malloc(sizeof(double) * (M + 1));
  v[M] = jc;
  for (int k = M; k >= 1; --k) {
    double jm = (2.0 * k / z) * jc - jp;
    jp = jc;
    jc = jm;
    v[k - 1] = jc;
    if (fabs(jc) > 1e250) { /* rescale */
      for (int q = k - 1; q <= M; ++q) v[q] *= 1e-250;
      jp *= 1e-250;
      jc *= 1e-250;
    }
  }
  norm = v[0];
  for (int k = 2; k <= M; k += 2) norm += 2.0 * v[k];
  for (int k = 0; k <= K; ++k) J[k] = k <= M ? v[k] / norm : 0.0;
  free(v);
}

/* smallest degree K with |J_k(tau r)| < 1e-17 for k > K, r = (hi - lo)/2 */
int orc_cheb_degree(double tau, double lo, double hi) {
  double z = fabs(tau) * 0.5 * (hi - lo);
  int Kmax = (int)(z + 60 + 10 * cbrt(z + 1));
  double* J = (double*)malloc(sizeof(double) * (Kmax + 1));
  bessel_j(Kmax, z, J);
  int K = Kmax;
  while (K > 1 && fabs(J[K]) < 1e-17 && fabs(J[K - 1]) < 1e-17) --K;
  free(J);
  return K;
}

/* a_k of exp(-i tau (c + r y)) = sum_k a_k T_k(y): e^{-i tau c} (2 - delta_k0) (-i)^k J_k(tau r) */
static void cheb_coeffs(double tau, double lo, double hi, int K, cplx* a) {
  double c = 0.5 * (hi + lo), r = 0.5 * (hi - lo);
  double* J = (double*)malloc(sizeof(double) * (K + 1));
  if (tau >= 0) {
    bessel_j(K, tau * r, J);
  } else { /* J_k(-z) = (-1)^k J_k(z) */
    bessel_j(K, -tau * r, J);
    for (int k = 1; k <= K; k += 2) J[k] = -J[k];
  }
  cplx ph = cexp(-I * tau * c), mi = 1.0;
  for (int k = 0; k <= K; ++k) {
    a[k] = (k == 0 ? 1.0 : 2.0) * mi * J[k] * ph;
    mi *= -I;
  }
  free(J);
}

/* orbs <- exp(-i tau h(E)) orbs by the degree-K expansion on [lo, hi]; with U non-null the terms
 * u_k = T_k(h~) phi_j (h~ = (h - c)/r) are kept at U[(k nO + j) nG] */
static void cheb_apply(const orc_sys* s, cplx* orbs, double tau, double E, double lo, double hi, int K, cplx* U) {
  const int nO = s->nO, nG = s->nG;
  const double c = 0.5 * (hi + lo), r = 0.5 * (hi - lo);
  cplx* a = (cplx*)malloc(sizeof(cplx) * (K + 1));
  cheb_coeffs(tau, lo, hi, K, a);
  cplx* um = (cplx*)malloc(sizeof(cplx) * nG);
  cplx* uc = (cplx*)malloc(sizeof(cplx) * nG);
  cplx* un = (cplx*)malloc(sizeof(cplx) * nG);
  cplx* out = (cplx*)malloc(sizeof(cplx) * nG);
  for (int j = 0; j < nO; ++j) {
    cplx* f = orbs + (size_t)j * nG;
    memcpy(uc, f, sizeof(cplx) * nG);
    for (int g = 0; g < nG; ++g) out[g] = a[0] * uc[g];
    if (U) memcpy(U + (size_t)j * nG, uc, sizeof(cplx) * nG);
    for (int k = 1; k <= K; ++k) {
      onebody(s, uc, E, un);
      for (int g = 0; g < nG; ++g) {
        cplx ht = (un[g] - c * uc[g]) / r;
        un[g] = k == 1 ? ht : 2.0 * ht - um[g];
        out[g] += a[k] * un[g];
      }
      if (U) memcpy(U + ((size_t)k * nO + j) * nG, un, sizeof(cplx) * nG);
      cplx* tp = um;
      um = uc;
      uc = un;
      un = tp;
    }
    memcpy(f, out, sizeof(cplx) * nG);
  }
  free(a), free(um), free(uc), free(un), free(out);
}

/* lam <- A^T lam for the map above (orbital part; C passes through); returns <lam, dA/dE phi> */
static double cheb_adjoint(const orc_sys* s, cplx* lamOrbs, double tau, double E, double lo, double hi, int K,
                           const cplx* U) {
  const int nO = s->nO, nG = s->nG;
  const double c = 0.5 * (hi + lo), r = 0.5 * (hi - lo);
  cplx* a = (cplx*)malloc(sizeof(cplx) * (K + 1));
  cheb_coeffs(tau, lo, hi, K, a);
  cplx* ub = (cplx*)malloc(sizeof(cplx) * (K + 1) * (size_t)nG);
  cplx* hv = (cplx*)malloc(sizeof(cplx) * nG);
  double eb = 0.0;
  for (int j = 0; j < nO; ++j) {
    cplx* lj = lamOrbs + (size_t)j * nG;
    for (int k = 0; k <= K; ++k)
      for (int g = 0; g < nG; ++g) ub[(size_t)k * nG + g] = conj(a[k]) * lj[g];
    for (int k = K; k >= 1; --k) {
      cplx* bk = ub + (size_t)k * nG;
      const cplx* uprev = U + ((size_t)(k - 1) * nO + j) * nG;
      double fac = k == 1 ? 1.0 : 2.0;
      onebody(s, bk, E, hv);
      for (int g = 0; g < nG; ++g) {
        ub[(size_t)(k - 1) * nG + g] += fac * (hv[g] - c * bk[g]) / r;
        if (k >= 2) ub[(size_t)(k - 2) * nG + g] -= bk[g];
        eb += fac * s->w[g] * s->x[g] / r * creal(conj(bk[g]) * uprev[g]);
      }
    }
    memcpy(lj, ub, sizeof(cplx) * nG);
  }
  free(a), free(ub), free(hv);
  return eb;
}

static double laser_basis(const orc_sys* s, int k, double t) {
  if (s->nLaser <= 0 || t < 0.0 || t > s->laserT) return 0.0;
  double env = sin(M_PI * t / s->laserT);
  return sin((k + 1) * M_PI * t / s->laserT) * env * env;
}

typedef struct {
  cplx *U1, *U2, *y1, *y2, *y3, *y4; /* Chebyshev terms of both half-steps, RK4 stage inputs */
} strang_tape;

/* classical RK4 with the arithmetic of orc_rk4, recording the stage inputs */
static void rk4_taped(const orc_sys* s, int rhs, cplx* state, double t0, double dt, strang_tape* tp) {
  size_t n = state_size(s);
  cplx* d = (cplx*)malloc(sizeof(cplx) * n);
  cplx* inc = (cplx*)malloc(sizeof(cplx) * n);
  cplx* tmp = (cplx*)malloc(sizeof(cplx) * n);
  double dt2 = dt / 2.0, dt3 = dt / 3.0, dt6 = dt / 6.0;
  if (tp) memcpy(tp->y1, state, sizeof(cplx) * n);
  eval_rhs(s, rhs, state, t0, d);
  for (size_t i = 0; i < n; ++i) {
    inc[i] = d[i] * dt6;
    tmp[i] = state[i] + d[i] * dt2;
  }
  if (tp) memcpy(tp->y2, tmp, sizeof(cplx) * n);
  eval_rhs(s, rhs, tmp, t0 + dt2, d);
  for (size_t i = 0; i < n; ++i) {
    inc[i] += d[i] * dt3;
    tmp[i] = state[i] + d[i] * dt2;
  }
  if (tp) memcpy(tp->y3, tmp, sizeof(cplx) * n);
  eval_rhs(s, rhs, tmp, t0 + dt2, d);
  for (size_t i = 0; i < n; ++i) {
    inc[i] += d[i] * dt3;
    tmp[i] = state[i] + d[i] * dt;
  }
  if (tp) memcpy(tp->y4, tmp, sizeof(cplx) * n);
  eval_rhs(s, rhs, tmp, t0 + dt, d);
  for (size_t i = 0; i < n; ++i) state[i] += inc[i] + d[i] * dt6;
  free(d), free(inc), free(tmp);
}

/* A(dt/2) B(dt) A(dt/2): A = degree-K Chebyshev exp(-i dt/2 h(E)) with E at the half-step's
 * midpoint (the time at which the Lanczos step evaluates the linear part), B = RK4 on rhsB */
static void strang_cheb_step(const orc_sys* s, int rhsB, cplx* state, double t0, double dt, double lo, double hi,
                             int K, strang_tape* tp) {
  cheb_apply(s, state + s->nC, 0.5 * dt, laser_field(s, t0 + 0.25 * dt), lo, hi, K, tp ? tp->U1 : NULL);
  rk4_taped(s, rhsB, state, t0, dt, tp);
  cheb_apply(s, state + s->nC, 0.5 * dt, laser_field(s, t0 + 0.75 * dt), lo, hi, K, tp ? tp->U2 : NULL);
}

int orc_strang_cheb(const orc_sys* s, int rhsB, cplx* state, double t0, double t1, double lo, double hi, int K) {
  if (s->nBT != 1 || K < 1) return -1;
  strang_cheb_step(s, rhsB, state, t0, t1 - t0, lo, hi, K, NULL);
  return 0;
}

/* J = <x>(t0 + nSteps dt) after nSteps Strang-Chebyshev steps from `state` (overwritten by the
 * final state) and gradc[k] = dJ/dc_k for the laser coefficients of orc_set_laser, by the
 * discrete adjoint over all stored steps. */
int orc_dipole_gradient(const orc_sys* s, int rhsB, cplx* state, double t0, double dt, int nSteps, double lo,
                        double hi, int K, double* J, double* gradc) {
  if (s->nBT != 1 || K < 1 || nSteps < 1) return -1;
  const size_t n = state_size(s), nOG = (size_t)s->nO * s->nG;
  cplx* traj = (cplx*)malloc(sizeof(cplx) * n * (nSteps + 1));
  strang_tape tp;
  tp.U1 = (cplx*)malloc(sizeof(cplx) * nOG * (K + 1));
  tp.U2 = (cplx*)malloc(sizeof(cplx) * nOG * (K + 1));
  tp.y1 = (cplx*)malloc(sizeof(cplx) * n);
  tp.y2 = (cplx*)malloc(sizeof(cplx) * n);
  tp.y3 = (cplx*)malloc(sizeof(cplx) * n);
  tp.y4 = (cplx*)malloc(sizeof(cplx) * n);
  cplx* lam = (cplx*)malloc(sizeof(cplx) * n);
  cplx* yb = (cplx*)malloc(sizeof(cplx) * n);
  cplx* d1 = (cplx*)malloc(sizeof(cplx) * n);
  cplx* d2 = (cplx*)malloc(sizeof(cplx) * n);
  cplx* d3 = (cplx*)malloc(sizeof(cplx) * n);
  cplx* d4 = (cplx*)malloc(sizeof(cplx) * n);
  cplx* g = (cplx*)malloc(sizeof(cplx) * n);
  memcpy(traj, state, sizeof(cplx) * n);
  for (int m = 0; m < nSteps; ++m) {
    memcpy(traj + n * (m + 1), traj + n * m, sizeof(cplx) * n);
    strang_cheb_step(s, rhsB, traj + n * (m + 1), t0 + m * dt, dt, lo, hi, K, NULL);
  }
  memcpy(state, traj + n * nSteps, sizeof(cplx) * n);
  *J = orc_dipole(s, state);
  for (int k = 0; k < s->nLaser; ++k) gradc[k] = 0.0;
  orc_dipole_grad(s, state, lam);
  for (int m = nSteps - 1; m >= 0; --m) {
    const double tm = t0 + m * dt, dt2 = dt / 2.0, dt3 = dt / 3.0, dt6 = dt / 6.0;
    memcpy(yb, traj + n * m, sizeof(cplx) * n);
    strang_cheb_step(s, rhsB, yb, tm, dt, lo, hi, K, &tp); /* recompute with the tape */
    double e, ts[6] = {tm + 0.75 * dt, tm + dt, tm + dt2, tm + dt2, tm, tm + 0.25 * dt}, eb[6];
    eb[0] = cheb_adjoint(s, lam + s->nC, 0.5 * dt, laser_field(s, ts[0]), lo, hi, K, tp.U2);
    for (size_t i = 0; i < n; ++i) {
      d4[i] = dt6 * lam[i];
      d3[i] = dt3 * lam[i];
      d2[i] = dt3 * lam[i];
      d1[i] = dt6 * lam[i];
    }
    orc_rhs_vjp(s, rhsB, tp.y4, tm + dt, d4, g, &e);
    eb[1] = e;
    for (size_t i = 0; i < n; ++i) {
      lam[i] += g[i];
      d3[i] += dt * g[i];
    }
    orc_rhs_vjp(s, rhsB, tp.y3, tm + dt2, d3, g, &e);
    eb[2] = e;
    for (size_t i = 0; i < n; ++i) {
      lam[i] += g[i];
      d2[i] += dt2 * g[i];
    }
    orc_rhs_vjp(s, rhsB, tp.y2, tm + dt2, d2, g, &e);
    eb[3] = e;
    for (size_t i = 0; i < n; ++i) {
      lam[i] += g[i];
      d1[i] += dt2 * g[i];
    }
    orc_rhs_vjp(s, rhsB, tp.y1, tm, d1, g, &e);
    eb[4] = e;
    for (size_t i = 0; i < n; ++i) lam[i] += g[i];
    eb[5] = cheb_adjoint(s, lam + s->nC, 0.5 * dt, laser_field(s, ts[5]), lo, hi, K, tp.U1);
    for (int q = 0; q < 6; ++q)
      for (int k = 0; k < s->nLaser; ++k) gradc[k] += eb[q] * laser_basis(s, k, ts[q]);
  }
  free(traj), free(tp.U1), free(tp.U2), free(tp.y1), free(tp.y2), free(tp.y3), free(tp.y4);
  free(lam), free(yb), free(d1), free(d2), free(d3), free(d4), free(g);
  return 0;
}
