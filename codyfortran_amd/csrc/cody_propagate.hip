// Device-resident propagation of the context's state: short iterative Lanczos (SIL) on the
// linear split part, classical RK4, and the Strang split step.
//
// The Lanczos loop of one SIL sub-step (src/IntegratorList/Sil/S_IntegratorList_Sil.f90:207-268)
// runs without the host: every iteration is two launches (k_lz_pass1, then k_lz_next, which also
// starts the next iteration) whose first act is to test a device flag `done`.  The end of an
// iteration (lz_eval: the subspace exponential and the error estimate) sets it, together with the
// Krylov dimension, when the tolerance is met, on breakdown or at the maximal dimension; the
// launches after it return at once.  A sub-step is therefore a fixed launch sequence, captured once as a HIP graph and
// replayed, and the host reads the convergence flags once per sub-step to apply the reference's
// step halving (:155-170).  A Strang step (S_Propagator_SplitStep_Order2.f90:52-77) is one graph:
// SIL(dt/2), RK4 on the nonlinear part, SIL(dt/2), each SIL as a single sub-step, after a backup
// of the state.  If either SIL misses the tolerance in that single sub-step, the state is restored
// and the step is redone with host-driven halving, so the result is the reference's in every case.
//
// All vector algebra, Gram-Schmidt re-orthogonalisation and the tridiagonal exponential run in
// FP64.  Dot products are per-block partial sums that every block of the next launch adds in the
// same fixed order, so all blocks use bitwise identical coefficients.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <initializer_list>
#include <string>
#include <vector>

#include "../../include/cody_mctdhx.h"
#include "cody_ctx.hpp"
#include "fedvr_device.hpp"
#include "mctdhx_kernels.hpp"

namespace cody {
int enqueue_rhs(cody_ctx* c, const double2* state, double2* dstate, double t, int mode,
                hipStream_t st, double* energy, RhsTrace* tr);
int run_rhs(cody_ctx* c, const double2* state, double2* dstate, double t, int mode, hipStream_t st);
int check_ready(const cody_ctx* c);
double laser_field(const cody_ctx* c, double t);
}  // namespace cody

using namespace cody;

namespace {

constexpr int kLzMaxBlocks = 512;  // blocks of a Lanczos vector launch: one 256-entry tile each, at most this many
                                   // (partial-sum rows); larger states take several tiles per block
constexpr int kMaxKrylov = 32;     // maximal Krylov dimension
constexpr int kPS = 64;            // partial-sum row stride (up to kMaxKrylov dots and one norm)
constexpr double kZero = 1.0e-14;  // zeroThreshold of the reference
constexpr int kChebMax = 200;      // Chebyshev terms of the subspace exponential (else Taylor)
constexpr int kDctMax = 128;       // Chebyshev nodes of the coefficient transform (longer series: Bessel values)
static_assert((kChebMax + 1) + (kDctMax + 256) + 2 * kDctMax <= 3 * kMaxKrylov * kMaxKrylov,
              "the coefficient arrays fit the exponential workspace");

// parameter slots (device doubles) written by the host before a replay
enum {
  kParSubDt = 0, kParSubE = 1,               // one SIL sub-step: dt, E(midpoint)
  kParSil1Dt = 2, kParSil1E = 3,             // Strang: first SIL
  kParSil2Dt = 4, kParSil2E = 5,             //         second SIL
  kParRk0 = 6,                               //         E at t0, t0 + dt/2, t0 + dt (RK4 stages)
  kParCount = 16
};

__device__ __forceinline__ double metric_at(int i, int nC, int nG, const double* __restrict__ w) {
  return (w == nullptr || i < nC) ? 1.0 : w[(i - nC) % nG];
}

// sOut[k] = sum_b partial[b][k] for k < nv <= kPS over the nRows block rows, fixed order: G row
// groups (8 groups of 32 lanes for nv <= 32, else 4 of 64), group q taking rows q, q + G, ... in
// order, then the groups in order.  A group's rows are loaded 17 at a time before they are added,
// so up to 136 rows (8 groups) cost one memory round trip.  sTmp holds 8 x kPS values.
__device__ void lz_reduce(const double2* __restrict__ partial, int nRows, int nv, double2* sOut,
                          double2 (*sTmp)[kPS]) {
  constexpr int kBatch = 17;
  const int t = threadIdx.x;
  const int G = nv <= 32 ? 8 : 4, lanes = 256 / G;
  const int k = t % lanes, q = t / lanes;
  if (k < nv) {
    double2 acc = make_double2(0.0, 0.0);
    for (int r0 = q; r0 < nRows; r0 += G * kBatch) {
      double2 p[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int r = r0 + G * u;
        p[u] = r < nRows ? partial[(size_t)r * kPS + k] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        acc.x += p[u].x;
        acc.y += p[u].y;
      }
    }
    sTmp[q][k] = acc;
  }
  __syncthreads();
  if (t < nv) {
    double2 acc = sTmp[0][t];
    for (int g = 1; g < G; ++g) {
      acc.x += sTmp[g][t].x;
      acc.y += sTmp[g][t].y;
    }
    sOut[t] = acc;
  }
  __syncthreads();
}

// partial[block][k] = sum over this block's entries of m_i conj(V_k[i]) y[i] for k < nv, and with
// withNorm also partial[block][nv] = sum m_i |y_i|^2.  The vectors are taken eight at a time so
// that their loads are in flight together; each batch's 16 block sums go through LDS
// (block_sums256: sBuf 8 x 256 doubles, sRes 16 doubles), fixed order.
__device__ void lz_dots(const double2* __restrict__ V, size_t ld, int nv, int withNorm, const double2* y, int n,
                        int nb, int nC, int nG, const double* __restrict__ metric, double2* __restrict__ partial,
                        double* sBuf, double* sRes) {
  const int nt = nv + (withNorm ? 1 : 0);
  for (int k0 = 0; k0 < nt; k0 += 8) {
    double acc[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += nb * 256) {
      const double m = metric_at(i, nC, nG, metric);
      const double2 b = y[i];
      double2 a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u;
        a[u] = k < nv ? V[k * ld + i] : (k == nv ? b : make_double2(0.0, 0.0));
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        acc[u] = fma(m, a[u].x * b.x + a[u].y * b.y, acc[u]);
        acc[8 + u] = fma(m, a[u].x * b.y - a[u].y * b.x, acc[8 + u]);
      }
    }
    block_sums256<16>(acc, sBuf, sRes);
    for (int u = threadIdx.x; u < 8 && k0 + u < nt; u += 256)
      partial[(size_t)blockIdx.x * kPS + k0 + u] = make_double2(sRes[u], sRes[8 + u]);
    __syncthreads();  // sRes is reused by the next batch
  }
}

__global__ void k_set_par(double* __restrict__ par, int off, int n, double v0, double v1, double v2, double v3,
                          double v4, double v5, double v6) {
  const double v[7] = {v0, v1, v2, v3, v4, v5, v6};
  for (int k = 0; k < n; ++k) par[off + k] = v[k];
}

__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ a, double2* __restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Start of a sub-step: clear the loop control and form the partial sums of <in|in>_m.
__global__ __launch_bounds__(256) void k_lz_norm(const double2* __restrict__ in, int n, int nb, int nC, int nG,
                                                 const double* __restrict__ metric, double2* __restrict__ partial,
                                                 int* __restrict__ ctl) {
  __shared__ double sBuf[8 * 256], sRes[16];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl[0] = 0;
    ctl[1] = 0;
    ctl[2] = 0;
  }
  lz_dots(in, 0, 0, 1, in, n, nb, nC, nG, metric, partial, sBuf, sRes);
}

// v_0 = in / |in|_m; a zero input ends the loop at once (the sub-step returns its input).
__global__ __launch_bounds__(256) void k_lz_start(const double2* __restrict__ partial, const double2* __restrict__ in,
                                                  double2* __restrict__ v0, int n, int nb, double* __restrict__ nrm2,
                                                  int* __restrict__ ctl, double* __restrict__ flags) {
  __shared__ double2 sC[kPS], sTmp[8][kPS];
  lz_reduce(partial, nb, 1, sC, sTmp);
  const double s2 = sC[0].x, nr = sqrt(s2);
  if (nr <= kZero) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl[0] = 1;
      ctl[1] = 0;
      ctl[2] = 1;
      flags[0] = 0.0;
      flags[1] = 1.0;
      flags[2] = 0.0;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *nrm2 = s2;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += nb * 256)
    v0[i] = make_double2(in[i].x / nr, in[i].y / nr);
}

// First step of Lanczos iteration j on the linear split part (:228-236 with TimeDerivativeOrbsLin,
// S_Method_Mb_OrbBased_Mctdhx.f90:152-177, whose derivative -i H v the SIL multiplies by i):
//   w = H v_j - beta_{j-1} v_{j-1},  H v = (T + V_ext + E x) v on each orbital block, 0 on C,
// and the partial sums of c_k = <v_k|w>_m for k <= j.  c_j is alpha_j + i (rounding), so the
// reference's w -= alpha_j v_j followed by its projections on v_0..v_j (:237-246) is the single
// subtraction of sum_k c_k v_k that k_lz_pass1 performs.  Used for iteration 0; the later
// iterations form w inside k_lz_next.
__global__ __launch_bounds__(256) void k_lz_apply(const double2* __restrict__ basis, size_t ld, int j,
                                                  const double* __restrict__ beta, const int* __restrict__ ctl,
                                                  const double* __restrict__ w, const double* __restrict__ xg,
                                                  const double* __restrict__ vext, KinParams kp,
                                                  const double* __restrict__ efield, int nC, int nG, int n, int nb,
                                                  const double* __restrict__ metric, double2* __restrict__ wout,
                                                  double2* __restrict__ pout) {
  __shared__ double sBuf[8 * 256], sRes[16];
  if (ctl[0]) return;
  const double E = *efield;
  const double bprev = j > 0 ? beta[j - 1] : 0.0;
  const double2* v = basis + (size_t)j * ld;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += nb * 256) {
    const double2 vi = v[i];
    double2 h = cmk(0, 0);
    if (i >= nC) {
      const int o = (i - nC) / nG, g = (i - nC) - o * nG;
      h = kinetic_at(v + nC + (size_t)o * nG, g, kp, 1.0 / w[g]);
      const double pot = vext[g] + E * xg[g];
      h.x = fma(pot, vi.x, h.x);
      h.y = fma(pot, vi.y, h.y);
    }
    if (j > 0) {
      const double2 p = v[i - ld];
      h.x -= bprev * p.x;
      h.y -= bprev * p.y;
    }
    wout[i] = h;
  }
  lz_dots(basis, ld, j + 1, 0, wout, n, nb, nC, nG, metric, pout, sBuf, sRes);
}

// The Lanczos vector launches keep a thread's basis entries v_k[i], k < nv <= KMAX, in registers:
// they are loaded once, all together, and serve both the subtraction and the dot products.
constexpr int lz_kmax(int nv) { return nv <= 8 ? 8 : nv <= 16 ? 16 : kMaxKrylov; }

// Writes this block's nv + 1 partial sums from acc (slot 0: the extra sum, slots 2 + 2k: k < nv)
// through the LDS block sums.
template <int NS>
__device__ __forceinline__ void lz_store_partials(const double (&acc)[NS], int nv, double* sBuf, double* sRes,
                                                  double2* __restrict__ partial, int extraAt) {
  block_sums256<NS>(acc, sBuf, sRes, 2 * (nv + 1));
  for (int k = threadIdx.x; k <= nv; k += 256)
    partial[(size_t)blockIdx.x * kPS + k] =
        k == extraAt ? make_double2(sRes[0], sRes[1]) : make_double2(sRes[2 + 2 * k], sRes[3 + 2 * k]);
}

// First re-orthogonalisation pass: w -= sum_{k<=j} c_k v_k (c from pin), alpha_j = Re c_j; then
// the partial sums of the second pass, <v_k|w> for k <= j, and of <w|w>.
template <int KMAX>
__global__ __launch_bounds__(256) void k_lz_pass1(const int* __restrict__ ctl, const double2* __restrict__ pin,
                                                  int j, const double2* __restrict__ basis, size_t ld,
                                                  double2* __restrict__ w, int n, int nb, int nC, int nG,
                                                  const double* __restrict__ metric, double* __restrict__ alpha,
                                                  double2* __restrict__ pout) {
  constexpr int NS = 2 * (KMAX + 1);
  __shared__ double2 sC[kPS], sTmp[8][kPS];
  __shared__ double sBuf[8 * 256], sRes[NS];
  const int nv = j + 1;
  // the loop-control flag is loaded with the partial sums (one round trip), tested after them
  const int done = ctl[0];
  lz_reduce(pin, nb, nv, sC, sTmp);
  if (done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) alpha[j] = sC[j].x;
  double acc[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) acc[u] = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += nb * 256) {
    double2 v[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) v[k] = k < nv ? basis[k * ld + i] : make_double2(0.0, 0.0);
    double2 r = w[i];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (k < nv) {
        const double2 cf = sC[k];
        r.x -= cf.x * v[k].x - cf.y * v[k].y;
        r.y -= cf.x * v[k].y + cf.y * v[k].x;
      }
    }
    w[i] = r;
    const double m = metric_at(i, nC, nG, metric);
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      acc[2 + 2 * k] = fma(m, v[k].x * r.x + v[k].y * r.y, acc[2 + 2 * k]);
      acc[3 + 2 * k] = fma(m, v[k].x * r.y - v[k].y * r.x, acc[3 + 2 * k]);
    }
    acc[0] = fma(m, r.x * r.x + r.y * r.y, acc[0]);
  }
  lz_store_partials(acc, nv, sBuf, sRes, pout, nv);
}

// J_k(a), k = 0..K (0 < a), by Miller's backward recurrence J_{k-1} = (2k/a) J_k - J_{k+1} from
// far above K, normalised with J_0 + 2 sum_k J_2k = 1 (Abramowitz & Stegun 9.1.27, 9.1.46);
// rescaled against overflow.  One thread.
__device__ void bessel_j_miller(double a, int K, double* J) {
  int N = K + 2 * (int)sqrt(40.0 * (K + 1)) + 10;
  N += N & 1;
  const double ta = 2.0 / a;
  double jn1 = 0.0, jn = 1.0, sum = 0.0;
  for (int k = N; k >= 1; --k) {
    if (k <= K) J[k] = jn;
    if ((k & 1) == 0) sum += 2.0 * jn;
    const double jm = (k * ta) * jn - jn1;
    jn1 = jn;
    jn = jm;
    if (fabs(jn) > 1e250) {
      jn *= 1e-250;
      jn1 *= 1e-250;
      sum *= 1e-250;
      for (int q = k; q <= K; ++q) J[q] *= 1e-250;
    }
  }
  J[0] = jn;
  sum += jn;
  const double is = 1.0 / sum;
  for (int q = 0; q <= K; ++q) J[q] *= is;
}

template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// v of lane l - 1 (0 in lane 0) and of lane l + 1 (0 in lane 63): DPP wave_shr:1 / wave_shl:1
__device__ __forceinline__ double from_lane_below(double v) { return dpp_move<0x138>(v); }
__device__ __forceinline__ double from_lane_above(double v) { return dpp_move<0x130>(v); }

__device__ __forceinline__ double readlane_double(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Number of Chebyshev terms for exp(-i a X), |X| <= 1: the first k > a with (a/2)^k / k! < 1e-18
// (which bounds |J_k(a)| and the rest of the series); -1 when more than kChebMax would be needed.
// Whole wave 0: lane l tests k = k0 + l + 1, the products (a/2)^k / k! by a prefix product across
// the lanes.  Every lane returns the same value.
__device__ int cheb_terms_wave(double a) {
  const int l = threadIdx.x & 63;
  double base = 1.0;
  for (int k0 = 0; k0 < kChebMax; k0 += 64) {
    const int k = k0 + l + 1;
    double p = 0.5 * a / k;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double o = __shfl_up(p, off, 64);
      if (l >= off) p *= o;
    }
    p *= base;
    const unsigned long long hit = __ballot(k < kChebMax && k > a && p < 1e-18);
    if (hit) return k0 + __ffsll((long long)hit);
    base = readlane_double(p, 63);
  }
  return -1;
}

// phi = exp(-i dt T) e1 for the m x m Lanczos tridiagonal T(sA, sB) by scaling and squaring of a
// degree-12 Taylor polynomial; one workgroup.  Used when the Chebyshev series would be too long.
__device__ void tridiag_expm_taylor(const double* sA, const double* sB, int m, double dt, double2* sPhi,
                                    double2* E, double2* Tm, double2* Pw, int* sSq) {
  const int t = threadIdx.x, nt = blockDim.x, mm = m * m;
  if (t == 0) {
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) {
      double r = fabs(sA[i]);
      if (i > 0) r += fabs(sB[i - 1]);
      if (i < m - 1) r += fabs(sB[i]);
      nrm = fmax(nrm, r);
    }
    nrm *= fabs(dt);
    int s = 0;
    while (nrm > 0.25 && s < 60) {
      nrm *= 0.5;
      ++s;
    }
    *sSq = s;
  }
  __syncthreads();
  const int nsq = *sSq;
  const double sc = ldexp(1.0, -nsq);
  for (int e = t; e < mm; e += nt) {
    const int i = e % m, j = e / m;
    double tv = 0.0;
    if (i == j)
      tv = sA[i];
    else if (i == j + 1)
      tv = sB[j];
    else if (j == i + 1)
      tv = sB[i];
    Tm[e] = make_double2(0.0, -dt * sc * tv);
    E[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    Pw[e] = E[e];
  }
  __syncthreads();
  for (int k = 1; k <= 12; ++k) {  // |A| <= 1/4: the remainder is below 4e-18
    double2 r[4];
    int cnt = 0;
    for (int e = t; e < mm; e += nt, ++cnt) {
      const int i = e % m, j = e / m;
      double2 acc = make_double2(0.0, 0.0);
      for (int q = 0; q < m; ++q) {
        const double2 a = Pw[i + q * m], b = Tm[q + j * m];
        acc.x += a.x * b.x - a.y * b.y;
        acc.y += a.x * b.y + a.y * b.x;
      }
      r[cnt] = make_double2(acc.x / k, acc.y / k);
    }
    __syncthreads();
    cnt = 0;
    for (int e = t; e < mm; e += nt, ++cnt) {
      Pw[e] = r[cnt];
      E[e].x += r[cnt].x;
      E[e].y += r[cnt].y;
    }
    __syncthreads();
  }
  for (int sq = 0; sq < nsq; ++sq) {
    double2 r[4];
    int cnt = 0;
    for (int e = t; e < mm; e += nt, ++cnt) {
      const int i = e % m, j = e / m;
      double2 acc = make_double2(0.0, 0.0);
      for (int q = 0; q < m; ++q) {
        const double2 a = E[i + q * m], b = E[q + j * m];
        acc.x += a.x * b.x - a.y * b.y;
        acc.y += a.x * b.y + a.y * b.x;
      }
      r[cnt] = acc;
    }
    __syncthreads();
    cnt = 0;
    for (int e = t; e < mm; e += nt, ++cnt) E[e] = r[cnt];
    __syncthreads();
  }
  for (int i = t; i < m; i += nt) sPhi[i] = E[i];  // first column
  __syncthreads();
}

// phi = exp(-i dt T) e1 into sPhi (LDS) for the Lanczos tridiagonal T: alpha sA[0..m-1], beta
// sB[0..m-2].  The reference diagonalises T (EvaluateExpInSubspace, :277-310); here the
// Chebyshev expansion exp(-i dt T) = e^{-i dt c} sum_{k<=K} c_k T_k(X), X = (T - c)/r with c, r
// the centre and half-width of the Gershgorin interval (Tal-Ezer and Kosloff, J. Chem. Phys. 81
// (1984) 3967).  The coefficients of f(x) = exp(-i dt r x) are c_k = (2 - delta_k0) (-i)^k
// J_k(dt r) (Jacobi-Anger).  For K < kDctMax they come from f at N > K Chebyshev nodes
// x_j = cos theta_j, theta_j = pi (j + 1/2) / N, by the discrete cosine transform
//   c_k = (2 - delta_k0) / N sum_j f(x_j) cos(k theta_j),
// exact up to the aliased terms c_{2N-k}, c_{2N+k}, ... which the choice of K bounds by 1e-18
// (all threads: nodes, sums over node ranges); for longer series from Bessel values by Miller's
// recurrence.  The recurrence T_{k+1} = 2 X T_k - T_{k-1} on e1 runs on wave 0, vector component l
// in lane l, neighbours by DPP lane moves.  Whole workgroup; ws holds 3 kMaxKrylov^2 double2.
__device__ void subspace_exp(const double* sA, const double* sB, int m, double dt, double2* sPhi, double* sJ,
                             double* sPar, double2* ws, int* sSq) {
  const int t = threadIdx.x, l = t & 63;
  if (t < 64) {
    const double al = l < m ? sA[l] : 0.0;
    const double bu = l < m - 1 ? sB[l] : 0.0;           // coupling to l + 1
    const double bd = (l > 0 && l < m) ? sB[l - 1] : 0.0;  // coupling to l - 1
    double lo = l < m ? al - fabs(bu) - fabs(bd) : 1e300;
    double hi = l < m ? al + fabs(bu) + fabs(bd) : -1e300;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      lo = fmin(lo, __shfl_xor(lo, off, 64));
      hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    const double c = 0.5 * (lo + hi), r = 0.5 * (hi - lo), a = fabs(dt) * r;
    const int K = (r > 0.0 && dt != 0.0) ? cheb_terms_wave(a) : 0;
    if (l == 0) {
      sPar[0] = c;
      sPar[1] = r;
      sPar[2] = a;
      sPar[3] = (double)K;
    }
  }
  __syncthreads();
  const int K = (int)sPar[3];
  constexpr int K2 = kMaxKrylov * kMaxKrylov;
  if (K < 0) {
    tridiag_expm_taylor(sA, sB, m, dt, sPhi, ws, ws + K2, ws + 2 * K2, sSq);
    return;
  }
  const double c = sPar[0], r = sPar[1], a = sPar[2];
  // LDS layout in ws: coefficients, node values (and per-range partial sums), cosine table
  double2* sCoef = ws;                      // kChebMax + 1
  double2* sF = ws + (kChebMax + 1);        // kDctMax + blockDim.x
  double* sCos = reinterpret_cast<double*>(sF + kDctMax + 256);  // 4 kDctMax
  if (K > 0 && K < kDctMax) {
    const int N = K < 64 ? 64 : kDctMax, msk = 4 * N - 1;
    const double dr = dt * r;
    for (int q = t; q < 4 * N; q += blockDim.x) sCos[q] = cos(M_PI * q / (2 * N));  // cos(pi q / 2N)
    __syncthreads();
    if (t < N) {
      double sn, cs;
      sincos(dr * sCos[2 * t + 1], &sn, &cs);  // f(x_t) = exp(-i dr x_t)
      sF[t] = make_double2(cs, -sn);
    }
    __syncthreads();
    // thread t: coefficient k = t % N over node range t / N of blockDim.x / N
    const int k = t % N, part = t / N, nParts = blockDim.x / N, per = N / nParts;
    double re = 0.0, im = 0.0;
    if (k <= K) {
      for (int j = part * per; j < (part + 1) * per; ++j) {
        const double cw = sCos[(k * (2 * j + 1)) & msk];
        const double2 f = sF[j];
        re = fma(f.x, cw, re);
        im = fma(f.y, cw, im);
      }
    }
    sF[kDctMax + t] = make_double2(re, im);
    __syncthreads();
    if (t <= K) {
      double2 acc = sF[kDctMax + t];
      for (int q = 1; q < nParts; ++q) {
        const double2 v = sF[kDctMax + t + q * N];
        acc.x += v.x;
        acc.y += v.y;
      }
      const double sc = (t == 0 ? 1.0 : 2.0) / N;
      sCoef[t] = make_double2(acc.x * sc, acc.y * sc);
    }
    __syncthreads();
  } else if (K > 0) {
    if (t == 0) bessel_j_miller(a, K, sJ);
    __syncthreads();
    // (2 - delta_k0) (-i)^k J_k(a), conjugated for dt < 0 (the expansion is in exp(-i |dt| r X))
    const double sg = dt < 0.0 ? -1.0 : 1.0;
    for (int k = t; k <= K; k += blockDim.x) {
      const double v = (k == 0 ? 1.0 : 2.0) * sJ[k];
      const int q = k & 3;
      sCoef[k] = q == 0 ? make_double2(v, 0.0) : q == 1 ? make_double2(0.0, -sg * v)
               : q == 2 ? make_double2(-v, 0.0) : make_double2(0.0, sg * v);
    }
    __syncthreads();
  }
  if (t < 64) {
    double sre = 0.0, sim = 0.0;
    if (K == 0) {
      sre = l == 0 ? 1.0 : 0.0;  // T = c I (or dt = 0)
    } else {
      const double ir = 1.0 / r;
      const double al = l < m ? (sA[l] - c) * ir : 0.0;
      const double bu = l < m - 1 ? sB[l] * ir : 0.0;
      const double bd = (l > 0 && l < m) ? sB[l - 1] * ir : 0.0;
      double u0 = l == 0 ? 1.0 : 0.0;
      double u1 = al * u0 + bd * from_lane_below(u0) + bu * from_lane_above(u0);
      const double2 c0 = sCoef[0], c1 = sCoef[1];
      sre = c0.x * u0 + c1.x * u1;
      sim = c0.y * u0 + c1.y * u1;
      const double al2 = 2.0 * al, bu2 = 2.0 * bu, bd2 = 2.0 * bd;
      double2 cn = K >= 2 ? sCoef[2] : make_double2(0.0, 0.0);
      for (int k = 2; k <= K; ++k) {
        const double2 ck = cn;
        if (k < K) cn = sCoef[k + 1];  // the next coefficient's load overlaps this step
        double u2 = fma(al2, u1, -u0);
        u2 = fma(bd2, from_lane_below(u1), u2);
        u2 = fma(bu2, from_lane_above(u1), u2);
        sre = fma(ck.x, u2, sre);
        sim = fma(ck.y, u2, sim);
        u0 = u1;
        u1 = u2;
      }
    }
    // times e^{-i dt c}
    double sn, cs;
    sincos(dt * c, &sn, &cs);
    if (l < m) sPhi[l] = make_double2(cs * sre + sn * sim, cs * sim - sn * sre);
  }
  __syncthreads();
}

// The end of a Lanczos iteration with Krylov dimension m (one workgroup): phi = exp(-i dt T_m) e1
// for the tridiagonal alpha[0..m-1], beta[0..m-2], the error estimate bres |phi[m-1]| with the
// residual coupling bres = beta_{m-1} (0 on breakdown, bres <= zeroThreshold), flags = (error,
// converged), and the end of the loop (ctl) when it is within tol, on breakdown, or at m = maxDim
// (:238-254).  ws: 3 kMaxKrylov^2 values of LDS for the Taylor fallback of the exponential.
__device__ void lz_eval(int* ctl, const double* alpha, const double* beta, int m, double bres, int maxDim,
                        const double* dtp, double tol, double2* phi, double* flags, double2* ws) {
  __shared__ double2 sPhi[kMaxKrylov];
  __shared__ double sA[kMaxKrylov], sB[kMaxKrylov], sJ[kChebMax + 1], sPar[4];
  __shared__ int sSq;
  for (int k = threadIdx.x; k < m; k += blockDim.x) {
    sA[k] = alpha[k];
    sB[k] = k < m - 1 ? beta[k] : bres;
  }
  __syncthreads();
  const bool breakdown = bres <= kZero;
  subspace_exp(sA, sB, m, *dtp, sPhi, sJ, sPar, ws, &sSq);
  for (int k = threadIdx.x; k < m; k += blockDim.x) phi[k] = sPhi[k];
  if (threadIdx.x == 0) {
    const double2 last = sPhi[m - 1];
    const double err = breakdown ? 0.0 : bres * sqrt(last.x * last.x + last.y * last.y);
    const bool conv = err <= tol;
    flags[0] = err;
    flags[1] = conv ? 1.0 : 0.0;
    flags[2] = m;
    if (conv || breakdown || m == maxDim) {
      // The other workgroups of the same k_lz_next launch read ctl[0] at their entry: relaxed
      // atomic accesses, so a reader sees either value.  Both are correct: when this workgroup
      // ends the loop, that launch's vector results (v_{j+1}, w', their sums) are never used.
      __hip_atomic_store(&ctl[1], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// beta_j = |w2|_m = sqrt(|w|^2 - sum |c2_k|^2) from the second pass's sums (the v_k are
// orthonormal in the metric and the second pass removes rounding-size components only); every
// workgroup forms it from the same sums in the same order.
__device__ double lz_beta(const double2* __restrict__ pin, int nRows, int nv, double2* sC, double2 (*sTmp)[kPS],
                          double* sBeta) {
  lz_reduce(pin, nRows, nv + 1, sC, sTmp);
  if (threadIdx.x == 0) {
    double b2 = sC[nv].x;
    for (int k = 0; k < nv; ++k) b2 -= sC[k].x * sC[k].x + sC[k].y * sC[k].y;
    *sBeta = sqrt(fmax(b2, 0.0));
  }
  __syncthreads();
  return *sBeta;
}

// The end of iteration j and the first step of iteration j + 1 in one launch (j + 1 < maxDim):
//   v_{j+1} = (w - sum c2_k v_k) / beta_j (second pass, c2 from pin),
//   w' = H v_{j+1} - beta_j v_j (as k_lz_apply) and the partial sums of <v_k|w'>_m, k <= j + 1.
// v_{j+1} is formed on each 256-entry tile and on `halo` = nLoc - 1 entries either side (the
// reach of the FEDVR stencil) in LDS, so H needs no other workgroup's values; w' goes to the
// other w buffer, so no workgroup overwrites the w that another one's halo reads.  Each thread
// forms its tile entry and, for threads t < 2 halo, one halo entry, with the loads of both in
// flight together; the tile entry's basis values stay in registers for the dot products.
// Workgroup nb (the last) ends iteration j (lz_eval on T_{j+1}) while the others work; when it
// ends the loop the launch's vector results are never used.  Block 0 records beta_j.
template <int KMAX>
__global__ __launch_bounds__(256) void k_lz_next(int* ctl, const double2* __restrict__ pin,
                                                 const double* __restrict__ alpha, double* __restrict__ beta, int j,
                                                 int maxDim, const double* __restrict__ dtp, double tol,
                                                 double2* __restrict__ phi, double* __restrict__ flags,
                                                 double2* __restrict__ basis, size_t ld,
                                                 const double2* __restrict__ wcur, double2* __restrict__ wnext,
                                                 const double* __restrict__ wg, const double* __restrict__ xg,
                                                 const double* __restrict__ vext, KinParams kp,
                                                 const double* __restrict__ efield, int nC, int nG, int n, int nb,
                                                 const double* __restrict__ metric, double2* __restrict__ pout,
                                                 int halo) {
  constexpr int NS = 2 * (KMAX + 1);
  __shared__ double2 sWork[3 * kMaxKrylov * kMaxKrylov];  // exponential workspace, or the v tile
  __shared__ double2 sC[kPS], sTmp[8][kPS];
  __shared__ double sBuf[8 * 256], sRes[NS];
  __shared__ double sBeta;
  __shared__ double sK[(kMaxStencilReach + 1) * (kMaxStencilReach + 1)];  // K(a, b) of the reference element
  // (workgroup nb of this launch may set ctl[0]: relaxed atomic load, see lz_eval; loaded with the
  // partial sums of the beta reduction, tested after them)
  const int done = __hip_atomic_load(&ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int nv = j + 1;
  const int t = threadIdx.x;
  // the element matrix goes to LDS (its loads in flight during the reduction): a global K(a, b)
  // load per stencil neighbour, each waited for before the next, would dominate the apply
  const int nK2 = kp.nLoc * kp.nLoc;
  double kv[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) kv[u] = t + 256 * u < nK2 ? kp.kref[t + 256 * u] : 0.0;
  const double E = *efield;
  const double bj = lz_beta(pin, nb, nv, sC, sTmp, &sBeta);
  if (done) return;
  if (blockIdx.x == nb) {
    lz_eval(ctl, alpha, beta, nv, bj, maxDim, dtp, tol, phi, flags, sWork);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) beta[j] = bj;
  if (bj <= kZero) return;  // breakdown: workgroup nb ends the loop
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (t + 256 * u < nK2) sK[t + 256 * u] = kv[u];  // read after the tile's first barrier
  double2* vn = basis + (size_t)nv * ld;
  double2* sV = sWork;
  const bool hasHalo = t < 2 * halo;
  double acc[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) acc[u] = 0.0;
  for (int tile = blockIdx.x; tile * 256 < n; tile += nb) {
    const int i0 = tile * 256, lo = i0 - halo;
    const int i = i0 + t;
    // this thread's halo entry and its place in the tile (LDS slot p holds entry lo + p)
    const int ih = t < halo ? lo + t : i0 + 256 + (t - halo);
    const int ph = t < halo ? t : 256 + t;
    const bool inMain = i < n, inHalo = hasHalo && ih >= 0 && ih < n;
    // orbital block and grid point of the tile entry, with their per-point data
    const bool orbital = inMain && i >= nC;
    const int o = orbital ? (i - nC) / nG : 0, g = orbital ? (i - nC) - o * nG : 0;
    const double wgi = orbital ? wg[g] : 1.0, vxi = orbital ? vext[g] : 0.0, xgi = orbital ? xg[g] : 0.0;
    const double m = metric_at(i, nC, nG, metric);
    double2 v[KMAX], vh[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      v[k] = (k < nv && inMain) ? basis[k * ld + i] : make_double2(0.0, 0.0);
      vh[k] = (k < nv && inHalo) ? basis[k * ld + ih] : make_double2(0.0, 0.0);
    }
    double2 r = inMain ? wcur[i] : make_double2(0.0, 0.0);
    double2 rh = inHalo ? wcur[ih] : make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (k < nv) {
        const double2 cf = sC[k];
        r.x -= cf.x * v[k].x - cf.y * v[k].y;
        r.y -= cf.x * v[k].y + cf.y * v[k].x;
        rh.x -= cf.x * vh[k].x - cf.y * vh[k].y;
        rh.y -= cf.x * vh[k].y + cf.y * vh[k].x;
      }
    }
    const double2 vi = inMain ? make_double2(r.x / bj, r.y / bj) : make_double2(0.0, 0.0);
    __syncthreads();  // the previous tile's reads of sV are done
    sV[t + halo] = vi;
    if (hasHalo) sV[ph] = inHalo ? make_double2(rh.x / bj, rh.y / bj) : make_double2(0.0, 0.0);
    __syncthreads();
    if (inMain) {
      vn[i] = vi;
      double2 h = cmk(0, 0);
      if (orbital) {
        h = kinetic_at_lds_k(sV, lo - (nC + o * nG), g, kp, sK, 1.0 / wgi);
        const double pot = vxi + E * xgi;
        h.x = fma(pot, vi.x, h.x);
        h.y = fma(pot, vi.y, h.y);
      }
      double2 p = make_double2(0.0, 0.0);  // v_j[i], selected without a dynamic register index
#pragma unroll
      for (int k = 0; k < KMAX; ++k)
        if (k == j) p = v[k];
      h.x -= bj * p.x;
      h.y -= bj * p.y;
      wnext[i] = h;
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        acc[2 + 2 * k] = fma(m, v[k].x * h.x + v[k].y * h.y, acc[2 + 2 * k]);
        acc[3 + 2 * k] = fma(m, v[k].x * h.y - v[k].y * h.x, acc[3 + 2 * k]);
      }
      acc[0] = fma(m, vi.x * h.x + vi.y * h.y, acc[0]);  // <v_{j+1}|w'>
      acc[1] = fma(m, vi.x * h.y - vi.y * h.x, acc[1]);
    }
  }
  lz_store_partials(acc, nv, sBuf, sRes, pout, nv);
}

// The end of the last iteration j = maxDim - 1 (one workgroup): beta_j and lz_eval on T_maxDim,
// which ends the loop.
__global__ __launch_bounds__(256) void k_lz_last(int* ctl, const double2* __restrict__ pin,
                                                 const double* __restrict__ alpha, double* __restrict__ beta, int j,
                                                 int maxDim, const double* __restrict__ dtp, double tol,
                                                 double2* __restrict__ phi, double* __restrict__ flags, int nb) {
  __shared__ double2 sWork[3 * kMaxKrylov * kMaxKrylov];
  __shared__ double2 sC[kPS], sTmp[8][kPS];
  __shared__ double sBeta;
  if (ctl[0]) return;
  const double bj = lz_beta(pin, nb, j + 1, sC, sTmp, &sBeta);
  if (threadIdx.x == 0) beta[j] = bj;
  lz_eval(ctl, alpha, beta, j + 1, bj, maxDim, dtp, tol, phi, flags, sWork);
}

// out = |in|_m sum_{k < mDim} phi_k v_k, or out = in for a zero input (:258).  out may be in.
__global__ __launch_bounds__(256) void k_lz_combine(const int* __restrict__ ctl, const double2* __restrict__ basis,
                                                    size_t ld, const double2* __restrict__ phi,
                                                    const double* __restrict__ nrm2, const double2* in,
                                                    double2* out, int n) {
  const int m = ctl[1];
  const bool zero = ctl[2] != 0;
  const double nr = zero ? 0.0 : sqrt(*nrm2);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    if (zero) {
      out[i] = in[i];
      continue;
    }
    // eight basis vectors' loads in flight at a time, accumulated in order
    double2 acc = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < m; k0 += 8) {
      double2 a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = k0 + u < m ? basis[(k0 + u) * ld + i] : make_double2(0.0, 0.0);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (k0 + u < m) {
          const double2 p = phi[k0 + u];
          acc.x += a[u].x * p.x - a[u].y * p.y;
          acc.y += a[u].x * p.y + a[u].y * p.x;
        }
      }
    }
    out[i] = make_double2(nr * acc.x, nr * acc.y);
  }
}

// RK4 stage updates (S_IntegratorList_Rk_O4Expl.f90:107-137)
__global__ __launch_bounds__(256) void k_rk_stage(const double2* __restrict__ state,
                                                  const double2* __restrict__ d, double winc,
                                                  double wtmp, int first, double2* __restrict__ inc,
                                                  double2* __restrict__ tmp, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 di = d[i];
  double2 a = first ? make_double2(0.0, 0.0) : inc[i];
  a.x += winc * di.x;
  a.y += winc * di.y;
  inc[i] = a;
  tmp[i] = make_double2(state[i].x + wtmp * di.x, state[i].y + wtmp * di.y);
}
__global__ __launch_bounds__(256) void k_rk_final(double2* __restrict__ state,
                                                  const double2* __restrict__ inc,
                                                  const double2* __restrict__ d, double w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  state[i].x += inc[i].x + w * d[i].x;
  state[i].y += inc[i].y + w * d[i].y;
}

int perr(int code, const char* msg) {
  return set_error(code, msg);
}

#define HIPCHK(expr)                                             \
  do {                                                           \
    hipError_t e_ = (expr);                                      \
    if (e_ != hipSuccess) return perr(CODY_ERR_HIP, hipGetErrorString(e_)); \
  } while (0)

PropBuffers* buffers(cody_ctx* c, int kdim) {
  if (c->prop.kdim >= kdim) return &c->prop;
  c->prop.release();
  PropBuffers& b = c->prop;
  b.kdim = kdim;
  const size_t n = c->nS;
  if (ctx_alloc((void**)&b.basis, sizeof(double2) * n * kdim) || ctx_alloc((void**)&b.w, sizeof(double2) * n) ||
      ctx_alloc((void**)&b.w2, sizeof(double2) * n) ||
      ctx_alloc((void**)&b.partial, sizeof(double2) * 2 * kLzMaxBlocks * kPS) ||
      ctx_alloc((void**)&b.phi, sizeof(double2) * kMaxKrylov) ||
      ctx_alloc((void**)&b.trial, sizeof(double2) * n) || ctx_alloc((void**)&b.work, sizeof(double2) * n) ||
      ctx_alloc((void**)&b.backup, sizeof(double2) * n) ||
      ctx_alloc((void**)&b.alpha, sizeof(double) * kMaxKrylov) ||
      ctx_alloc((void**)&b.beta, sizeof(double) * kMaxKrylov) || ctx_alloc((void**)&b.nrm2, sizeof(double)) ||
      ctx_alloc((void**)&b.par, sizeof(double) * kParCount) || ctx_alloc((void**)&b.flags, sizeof(double) * 8) ||
      ctx_alloc((void**)&b.ctl, sizeof(int) * 4) ||
      ctx_alloc((void**)&b.rk_inc, sizeof(double2) * n) || ctx_alloc((void**)&b.rk_tmp, sizeof(double2) * n) ||
      ctx_alloc((void**)&b.rk_d, sizeof(double2) * n))
    return nullptr;
  if (hipHostMalloc((void**)&b.h_flags, sizeof(double) * 8, hipHostMallocDefault) != hipSuccess) return nullptr;
  return &b;
}

inline dim3 blocks_for(int n) { return dim3((n + 255) / 256); }

void set_par(PropBuffers* b, int off, std::initializer_list<double> v, hipStream_t st) {
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  int k = 0;
  for (double x : v)
    if (k < 7) a[k++] = x;
  hipLaunchKernelGGL(k_set_par, dim3(1), dim3(1), 0, st, b->par, off, k, a[0], a[1], a[2], a[3], a[4], a[5], a[6]);
}

// One SIL sub-step from `in` into `out` (PerformLanczosStep, :207-268): dt from *dtp, E of the
// linear part from *efp, flags[0..2] = (error, converged, Krylov dimension).  Launches only.
void enqueue_sil_substep(cody_ctx* c, PropBuffers* b, int maxDim, double tol, int useMetric, const double2* in,
                         double2* out, const double* dtp, const double* efp, double* flags, hipStream_t st) {
  const int n = c->nS;
  const size_t ld = (size_t)n;
  const double* metric = useMetric ? c->d_w : nullptr;
  double2* p0 = b->partial;
  double2* p1 = b->partial + kLzMaxBlocks * kPS;
  const int nb = std::min(kLzMaxBlocks, (n + 255) / 256);  // one 256-entry tile per block
  const dim3 g(nb), t(256);
  hipLaunchKernelGGL(k_lz_norm, g, t, 0, st, in, n, nb, c->nC, c->nG, metric, p0, b->ctl);
  hipLaunchKernelGGL(k_lz_start, g, t, 0, st, p0, in, b->basis, n, nb, b->nrm2, b->ctl, flags);
  const KinParams kp{c->kin.kref, c->kin.nE, c->kin.nLoc, c->kin.xminExcl, c->nG, c->kin.invh, c->kin.fac};
  // iteration j: k_lz_pass1 on w[j % 2], then k_lz_next (the end of j and w[(j + 1) % 2] of
  // j + 1, one more workgroup for the end of j) or, after the last iteration, k_lz_last
  double2* wb[2] = {b->w, b->w2};
  const int halo = kp.nLoc - 1;
  hipLaunchKernelGGL(k_lz_apply, g, t, 0, st, b->basis, ld, 0, b->beta, b->ctl, c->d_w, c->d_x, c->d_vext, kp, efp,
                     c->nC, c->nG, n, nb, metric, wb[0], p0);
  for (int j = 0; j < maxDim; ++j) {
    const int km = lz_kmax(j + 1);
    auto pass1 = km == 8 ? k_lz_pass1<8> : km == 16 ? k_lz_pass1<16> : k_lz_pass1<kMaxKrylov>;
    hipLaunchKernelGGL(pass1, g, t, 0, st, b->ctl, p0, j, b->basis, ld, wb[j % 2], n, nb, c->nC, c->nG, metric,
                       b->alpha, p1);
    auto next = km == 8 ? k_lz_next<8> : km == 16 ? k_lz_next<16> : k_lz_next<kMaxKrylov>;
    if (j + 1 < maxDim)
      hipLaunchKernelGGL(next, dim3(nb + 1), t, 0, st, b->ctl, p1, b->alpha, b->beta, j, maxDim, dtp, tol,
                         b->phi, flags, b->basis, ld, wb[j % 2], wb[(j + 1) % 2], c->d_w, c->d_x, c->d_vext, kp, efp,
                         c->nC, c->nG, n, nb, metric, p0, halo);
    else
      hipLaunchKernelGGL(k_lz_last, dim3(1), t, 0, st, b->ctl, p1, b->alpha, b->beta, j, maxDim, dtp, tol, b->phi,
                         flags, nb);
  }
  hipLaunchKernelGGL(k_lz_combine, blocks_for(n), t, 0, st, b->ctl, b->basis, ld, b->phi, b->nrm2, in, out, n);
}

// RK4 stages on `mode` (S_IntegratorList_Rk_O4Expl.f90:90-141).  With replay_rhs each right-hand
// side is the cached graph of run_rhs (which sets E(t) itself); otherwise the launches are
// enqueued (or captured) on st and read E from efp[0..2] = E(t0), E(t0 + dt/2), E(t0 + dt).
int enqueue_rk4(cody_ctx* c, PropBuffers* b, double t0, double t1, int mode, hipStream_t st, bool replay_rhs,
                const double* efp) {
  const int n = c->nS;
  const dim3 g = blocks_for(n);
  const double dt = t1 - t0, dt2 = dt / 2.0, dt3 = dt / 3.0, dt6 = dt / 6.0;
  auto rhs = [&](const double2* s, double t, int slot) {
    if (replay_rhs) return run_rhs(c, s, b->rk_d, t, mode, st);
    c->efield_src = efp ? efp + slot : nullptr;
    const int r = enqueue_rhs(c, s, b->rk_d, t, mode, st, nullptr, nullptr);
    c->efield_src = nullptr;
    return r;
  };
  int r;
  if ((r = rhs(c->d_state, t0, 0))) return r;
  hipLaunchKernelGGL(k_rk_stage, g, dim3(256), 0, st, c->d_state, b->rk_d, dt6, dt2, 1, b->rk_inc, b->rk_tmp, n);
  if ((r = rhs(b->rk_tmp, t0 + dt2, 1))) return r;
  hipLaunchKernelGGL(k_rk_stage, g, dim3(256), 0, st, c->d_state, b->rk_d, dt3, dt2, 0, b->rk_inc, b->rk_tmp, n);
  if ((r = rhs(b->rk_tmp, t0 + dt2, 1))) return r;
  hipLaunchKernelGGL(k_rk_stage, g, dim3(256), 0, st, c->d_state, b->rk_d, dt3, dt, 0, b->rk_inc, b->rk_tmp, n);
  if ((r = rhs(b->rk_tmp, t0 + dt, 2))) return r;
  hipLaunchKernelGGL(k_rk_final, g, dim3(256), 0, st, c->d_state, b->rk_inc, b->rk_d, dt6, n);
  return CODY_OK;
}

// Replays the cached graph of `key` (capturing `body` on first use) or, with graphs off or when
// capture fails, launches `body` directly on the context's stream.
int run_prop_graph(cody_ctx* c, PropBuffers* b, const PropGraph& key, const std::function<int(hipStream_t)>& body) {
  hipStream_t st = c->stream;
  if (c->graphs_on) {
    hipGraphExec_t ex = nullptr;
    for (const auto& g : b->graphs)
      if (g.kind == key.kind && g.in == key.in && g.out == key.out && g.maxDim == key.maxDim &&
          g.useMetric == key.useMetric && g.tol == key.tol && g.dt == key.dt)
        ex = g.exec;
    if (!ex) {
      if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        c->capturing = true;
        const int r = body(st);
        c->capturing = false;
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &gr);
        if (r == CODY_OK && e == hipSuccess && gr && hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) == hipSuccess) {
          if (b->graphs.size() >= 8) {
            hipGraphExecDestroy(b->graphs.front().exec);
            b->graphs.erase(b->graphs.begin());
          }
          PropGraph pg = key;
          pg.exec = ex;
          b->graphs.push_back(pg);
        } else {
          ex = nullptr;
          c->graphs_on = false;  // capture unsupported here: plain launches from now on
        }
        if (gr) hipGraphDestroy(gr);
        (void)hipGetLastError();
      } else {
        c->graphs_on = false;
      }
    }
    if (ex) {
      HIPCHK(hipGraphLaunch(ex, st));
      return CODY_OK;
    }
  }
  return body(st);
}

// One SIL sub-step of b->work over [tbase, tbase + dt] into b->trial; one host read of the flags.
int lanczos_substep(cody_ctx* c, PropBuffers* b, int maxDim, double tol, int useMetric, double tbase, double dt,
                    int* conv, double* err) {
  hipStream_t st = c->stream;
  set_par(b, kParSubDt, {dt, laser_field(c, tbase + 0.5 * dt)}, st);
  const PropGraph key{0, b->work, b->trial, maxDim, useMetric, tol, 0.0, nullptr};
  int r = run_prop_graph(c, b, key, [&](hipStream_t s) {
    enqueue_sil_substep(c, b, maxDim, tol, useMetric, b->work, b->trial, b->par + kParSubDt, b->par + kParSubE,
                        b->flags, s);
    return CODY_OK;
  });
  if (r) return r;
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->h_flags, b->flags, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *err = b->h_flags[0];
  *conv = b->h_flags[1] != 0.0;
  return CODY_OK;
}

int sil_impl(cody_ctx* c, double t0, double t1, int krylovDim, double tol, int maxRestarts,
             int useMetric, int* converged) {
  if (converged) *converged = 1;
  if (std::fabs(t1 - t0) <= kZero) return CODY_OK;
  const int maxDim = std::max(1, std::min(krylovDim, c->nS));
  PropBuffers* b = buffers(c, maxDim);
  if (!b) return perr(CODY_ERR_HIP, "SIL buffer allocation failed");
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(b->work, c->d_state, sizeof(double2) * c->nS, hipMemcpyDeviceToDevice, st));
  double tc = t0, rem = t1 - t0;
  int all = 1;
  // step halving on a missed tolerance (S_IntegratorList_Sil.f90:155-175)
  while (std::fabs(rem) > kZero) {
    double dts = rem;
    int restarts = 0, conv = 0;
    double err = 0.0;
    for (;;) {
      int r = lanczos_substep(c, b, maxDim, tol, useMetric, tc, dts, &conv, &err);
      if (r) return r;
      if (conv) break;
      if (std::fabs(dts) <= kZero || restarts >= maxRestarts) break;
      dts *= 0.5;
      ++restarts;
    }
    if (!conv) all = 0;
    HIPCHK(hipMemcpyAsync(b->work, b->trial, sizeof(double2) * c->nS, hipMemcpyDeviceToDevice, st));
    tc += dts;
    rem -= dts;
  }
  HIPCHK(hipMemcpyAsync(c->d_state, b->work, sizeof(double2) * c->nS, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  if (converged) *converged = all;
  return CODY_OK;
}

int rk4_impl(cody_ctx* c, double t0, double t1, int mode) {
  PropBuffers* b = buffers(c, 1);
  if (!b) return perr(CODY_ERR_HIP, "RK4 buffer allocation failed");
  int r = enqueue_rk4(c, b, t0, t1, mode, c->stream, true, nullptr);
  if (r) return r;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return CODY_OK;
}

// One Strang step A(dt/2) B(dt) A(dt/2) as one graph, each SIL a single sub-step with at most
// `dim` Lanczos iterations, after a backup of the state.  *ok: both SILs met the tolerance;
// *mDim: the larger of their Krylov dimensions.  The parameter slots are set by the caller.
int strang_graph_step(cody_ctx* c, PropBuffers* b, double t0, double t1, int dim, double tol, int useMetric,
                      bool* ok, int* mDim) {
  hipStream_t st = c->stream;
  const int n = c->nS;
  const PropGraph key{1, c->d_state, c->d_state, dim, useMetric, tol, t1 - t0, nullptr};
  int r = run_prop_graph(c, b, key, [&](hipStream_t s) {
    hipLaunchKernelGGL(k_copy, blocks_for(n), dim3(256), 0, s, c->d_state, b->backup, n);
    enqueue_sil_substep(c, b, dim, tol, useMetric, c->d_state, c->d_state, b->par + kParSil1Dt, b->par + kParSil1E,
                        b->flags, s);
    int rr = enqueue_rk4(c, b, t0, t1, kModeNonLin, s, false, b->par + kParRk0);
    if (rr) return rr;
    enqueue_sil_substep(c, b, dim, tol, useMetric, c->d_state, c->d_state, b->par + kParSil2Dt, b->par + kParSil2E,
                        b->flags + 3, s);
    return CODY_OK;
  });
  if (r) return r;
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->h_flags, b->flags, sizeof(double) * 6, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *ok = b->h_flags[1] != 0.0 && b->h_flags[4] != 0.0;
  *mDim = (int)std::max(b->h_flags[2], b->h_flags[5]);
  return CODY_OK;
}

// A(dt/2) B(dt) A(dt/2): S_Propagator_SplitStep_Order2.f90:61-75.  Fast path: one graph with each
// SIL as a single sub-step, its Lanczos loop capped one iteration above the largest Krylov
// dimension of the last eight steps (the launches after convergence return at once but still
// cost their dispatch).  A SIL that stops at the cap without meeting the tolerance has not run
// the reference's loop to its end: the state is restored and the step rerun with the full
// dimension, so an accepted step is the uncapped one.  If a SIL misses the tolerance at the full
// dimension, restore and use host-driven halving.
int strang_impl(cody_ctx* c, double t0, double t1, int krylovDim, double tol, int maxRestarts, int useMetric,
                int* converged) {
  const double dt = t1 - t0, t2 = t0 + 0.5 * dt;
  const int maxDim = std::max(1, std::min(krylovDim, c->nS));
  PropBuffers* b = buffers(c, maxDim);
  if (!b) return perr(CODY_ERR_HIP, "propagator buffer allocation failed");
  hipStream_t st = c->stream;
  const int n = c->nS;
  if (std::fabs(0.5 * dt) > kZero) {
    set_par(b, kParSil1Dt,
            {0.5 * dt, laser_field(c, t0 + 0.25 * dt), 0.5 * dt, laser_field(c, t2 + 0.25 * dt), laser_field(c, t0),
             laser_field(c, t0 + 0.5 * dt), laser_field(c, t0 + dt)},
            st);
    const int cap = b->krylov_cap(maxDim);
    bool ok = false;
    int m = 0, r;
    if ((r = strang_graph_step(c, b, t0, t1, cap, tol, useMetric, &ok, &m))) return r;
    if (!ok && cap < maxDim) {
      HIPCHK(hipMemcpyAsync(c->d_state, b->backup, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
      if ((r = strang_graph_step(c, b, t0, t1, maxDim, tol, useMetric, &ok, &m))) return r;
    }
    b->record_dim(m);
    if (ok) {
      if (converged) *converged = 1;
      return CODY_OK;
    }
    HIPCHK(hipMemcpyAsync(c->d_state, b->backup, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
  }
  int ok1 = 1, ok2 = 1, r;
  if ((r = sil_impl(c, t0, t0 + 0.5 * dt, krylovDim, tol, maxRestarts, useMetric, &ok1))) return r;
  if ((r = rk4_impl(c, t0, t0 + dt, kModeNonLin))) return r;
  if ((r = sil_impl(c, t2, t2 + 0.5 * dt, krylovDim, tol, maxRestarts, useMetric, &ok2))) return r;
  if (converged) *converged = ok1 && ok2;
  return CODY_OK;
}

}  // namespace

extern "C" {

int cody_sil(cody_ctx* c, double t0, double t1, int krylovDim, double tol, int maxRestarts,
             int useMetric, int* converged) {
  if (!c || krylovDim < 1 || krylovDim > kMaxKrylov || !(tol > 0.0) || maxRestarts < 0)
    return perr(CODY_ERR_ARG, "cody_sil: bad arguments (1 <= krylovDim <= 32, tol > 0)");
  int r = check_ready(c);
  if (r) return r;
  HIPCHK(hipSetDevice(c->device));
  return sil_impl(c, t0, t1, krylovDim, tol, maxRestarts, useMetric, converged);
}

int cody_rk4(cody_ctx* c, double t0, double t1, int full) {
  if (!c) return perr(CODY_ERR_ARG, "cody_rk4: null context");
  int r = check_ready(c);
  if (r) return r;
  HIPCHK(hipSetDevice(c->device));
  return rk4_impl(c, t0, t1, full ? kModeFull : kModeNonLin);
}

int cody_strang(cody_ctx* c, double t0, double t1, int krylovDim, double tol, int maxRestarts,
                int useMetric, int* converged) {
  if (!c || krylovDim < 1 || krylovDim > kMaxKrylov || !(tol > 0.0) || maxRestarts < 0)
    return perr(CODY_ERR_ARG, "cody_strang: bad arguments");
  int r = check_ready(c);
  if (r) return r;
  HIPCHK(hipSetDevice(c->device));
  return strang_impl(c, t0, t1, krylovDim, tol, maxRestarts, useMetric, converged);
}

}  // extern "C"
