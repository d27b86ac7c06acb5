// Device-resident propagation of the context's state: short iterative Lanczos (SIL) on the
// linear split part, classical RK4, and the Strang split step.  All vector algebra, Gram-Schmidt
// re-orthogonalisation and the tridiagonal exponential run in FP64 on the GPU; the host only
// reads back the convergence flags that drive the Krylov-dimension and step-halving control of
// the reference (src/IntegratorList/Sil/S_IntegratorList_Sil.f90:114-312).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/cody_mctdhx.h"
#include "cody_ctx.hpp"
#include "mctdhx_kernels.hpp"

namespace cody {
int enqueue_rhs(cody_ctx* c, const double2* state, double2* dstate, double t, int mode,
                hipStream_t st, double* energy, RhsTrace* tr);
int run_rhs(cody_ctx* c, const double2* state, double2* dstate, double t, int mode, hipStream_t st);
int check_ready(const cody_ctx* c);
}  // namespace cody

using namespace cody;

namespace {

constexpr int kDotBlocks = 120;
constexpr int kMaxKrylov = 32;

__device__ __forceinline__ double metric_at(int i, int nC, int nG, const double* __restrict__ w) {
  return (w == nullptr || i < nC) ? 1.0 : w[(i - nC) % nG];
}

// partial[b][k] = sum_{i in block b} conj(V_k[i]) m_i y[i], k = 0..nv-1 (V_k = V + k*ld)
__global__ __launch_bounds__(256) void k_multidot(const double2* __restrict__ V, size_t ld, int nv,
                                                  const double2* __restrict__ y, int n, int nC, int nG,
                                                  const double* __restrict__ w,
                                                  double2* __restrict__ partial) {
  __shared__ double sr[4][kMaxKrylov], si[4][kMaxKrylov];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < nv; ++k) {
    double ar = 0.0, ai = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const double m = metric_at(i, nC, nG, w);
      const double2 a = V[k * ld + i], b = y[i];
      ar = fma(m, a.x * b.x + a.y * b.y, ar);
      ai = fma(m, a.x * b.y - a.y * b.x, ai);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      ar += __shfl_xor(ar, off, 64);
      ai += __shfl_xor(ai, off, 64);
    }
    if (lane == 0) {
      sr[wv][k] = ar;
      si[wv][k] = ai;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nv; k += blockDim.x) {
    double r = 0.0, im = 0.0;
    for (int q = 0; q < 4; ++q) {
      r += sr[q][k];
      im += si[q][k];
    }
    partial[(size_t)blockIdx.x * kMaxKrylov + k] = make_double2(r, im);
  }
}

// out[k] = sum_b partial[b][k]
__global__ void k_multidot_final(const double2* __restrict__ partial, int nb, int nv,
                                 double2* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= nv) return;
  double r = 0.0, im = 0.0;
  for (int b = 0; b < nb; ++b) {
    r += partial[(size_t)b * kMaxKrylov + k].x;
    im += partial[(size_t)b * kMaxKrylov + k].y;
  }
  out[k] = make_double2(r, im);
}

// y -= sum_k coef[k] V_k   (coef on the device)
__global__ __launch_bounds__(256) void k_sub_combination(const double2* __restrict__ V, size_t ld,
                                                         int nv, const double2* __restrict__ coef,
                                                         double2* __restrict__ y, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 acc = y[i];
  for (int k = 0; k < nv; ++k) {
    const double2 a = V[k * ld + i], cf = coef[k];
    acc.x -= cf.x * a.x - cf.y * a.y;
    acc.y -= cf.x * a.y + cf.y * a.x;
  }
  y[i] = acc;
}

// w = i*deriv - beta_prev * vprev  (beta_prev from the device, 0 for j = 0)
__global__ __launch_bounds__(256) void k_lanczos_w(const double2* __restrict__ deriv,
                                                   const double2* __restrict__ vprev,
                                                   const double* __restrict__ beta, int j,
                                                   double2* __restrict__ w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 d = deriv[i];
  double2 r = make_double2(-d.y, d.x);
  if (j > 0) {
    const double b = beta[j - 1];
    const double2 v = vprev[i];
    r.x -= b * v.x;
    r.y -= b * v.y;
  }
  w[i] = r;
}

// y = s * x with s = scale / sqrt(Re(*nrm2))   (normalisation by a device-side norm)
__global__ __launch_bounds__(256) void k_scale_by_norm(const double2* __restrict__ x,
                                                       const double2* __restrict__ nrm2, double scale,
                                                       int invert, double2* __restrict__ y, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double nr = sqrt(nrm2->x);
  const double s = invert ? scale / nr : scale * nr;
  y[i] = make_double2(s * x[i].x, s * x[i].y);
}

// Lanczos scalars of iteration j (one thread): alpha_j from <v_j|w>, beta_j from <w|w>.
__global__ void k_store_alpha(const double2* __restrict__ dotv, double* __restrict__ alpha, int j) {
  alpha[j] = dotv[0].x;
}
__global__ __launch_bounds__(256) void k_axpy_real(const double* __restrict__ alpha, int j,
                                                   const double2* __restrict__ v,
                                                   double2* __restrict__ w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = alpha[j];
  w[i].x -= a * v[i].x;
  w[i].y -= a * v[i].y;
}
__global__ void k_store_beta(const double2* __restrict__ dotw, double* __restrict__ beta, int j) {
  beta[j] = sqrt(fmax(dotw[0].x, 0.0));
}

// phi = exp(-i dt T) e1 for the (m x m) Lanczos tridiagonal T(alpha, beta) by scaling and
// squaring of a degree-20 Taylor polynomial (one workgroup, FP64), and the residual estimate
// err = beta_{m-1} |phi_{m-1}| (:277-310).  flags[0] = err, flags[1] = beta_{m-1}.
__global__ __launch_bounds__(256) void k_tridiag_expm(const double* __restrict__ alpha,
                                                      const double* __restrict__ beta, int m,
                                                      double dt, double2* __restrict__ phi,
                                                      double* __restrict__ flags) {
  __shared__ double2 E[kMaxKrylov * kMaxKrylov], Tm[kMaxKrylov * kMaxKrylov], Pw[kMaxKrylov * kMaxKrylov];
  __shared__ int s_sq;
  const int t = threadIdx.x, nt = blockDim.x, mm = m * m;
  if (t == 0) {
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) {
      double r = fabs(alpha[i]);
      if (i > 0) r += fabs(beta[i - 1]);
      if (i < m - 1) r += fabs(beta[i]);
      nrm = fmax(nrm, r);
    }
    nrm *= fabs(dt);
    int s = 0;
    while (nrm > 0.25 && s < 60) {
      nrm *= 0.5;
      ++s;
    }
    s_sq = s;
  }
  __syncthreads();
  const double sc = ldexp(1.0, -s_sq);
  // A = -i dt T / 2^s ; E = I ; P = I
  for (int e = t; e < mm; e += nt) {
    const int i = e % m, j = e / m;
    double tv = 0.0;
    if (i == j) tv = alpha[i];
    else if (i == j + 1) tv = beta[j];
    else if (j == i + 1) tv = beta[i];
    Tm[e] = make_double2(0.0, -dt * sc * tv);
    E[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    Pw[e] = E[e];
  }
  __syncthreads();
  for (int k = 1; k <= 20; ++k) {
    // Pw = Pw * A / k
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
  for (int sq = 0; sq < s_sq; ++sq) {
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
  for (int i = t; i < m; i += nt) phi[i] = E[i];  // first column
  if (t == 0) {
    const double2 last = E[m - 1];
    flags[0] = beta[m - 1] * sqrt(last.x * last.x + last.y * last.y);
    flags[1] = beta[m - 1];
  }
}

// out = nrm * sum_j phi_j V_j
__global__ __launch_bounds__(256) void k_combine(const double2* __restrict__ V, size_t ld, int m,
                                                 const double2* __restrict__ phi,
                                                 const double2* __restrict__ nrm2,
                                                 double2* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double nr = sqrt(nrm2->x);
  double2 acc = make_double2(0.0, 0.0);
  for (int j = 0; j < m; ++j) {
    const double2 a = V[j * ld + i], p = phi[j];
    acc.x += a.x * p.x - a.y * p.y;
    acc.y += a.x * p.y + a.y * p.x;
  }
  out[i] = make_double2(nr * acc.x, nr * acc.y);
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
      ctx_alloc((void**)&b.deriv, sizeof(double2) * n) ||
      ctx_alloc((void**)&b.partial, sizeof(double2) * kDotBlocks * kMaxKrylov) ||
      ctx_alloc((void**)&b.dots, sizeof(double2) * kMaxKrylov) ||
      ctx_alloc((void**)&b.phi, sizeof(double2) * kMaxKrylov) || ctx_alloc((void**)&b.nrm2, sizeof(double2) * 2) ||
      ctx_alloc((void**)&b.trial, sizeof(double2) * n) || ctx_alloc((void**)&b.work, sizeof(double2) * n) ||
      ctx_alloc((void**)&b.alpha, sizeof(double) * kMaxKrylov) ||
      ctx_alloc((void**)&b.beta, sizeof(double) * kMaxKrylov) || ctx_alloc((void**)&b.flags, sizeof(double) * 4) ||
      ctx_alloc((void**)&b.rk_inc, sizeof(double2) * n) || ctx_alloc((void**)&b.rk_tmp, sizeof(double2) * n) ||
      ctx_alloc((void**)&b.rk_d, sizeof(double2) * n))
    return nullptr;
  if (hipHostMalloc((void**)&b.h_flags, sizeof(double) * 4, hipHostMallocDefault) != hipSuccess) return nullptr;
  return &b;
}

inline dim3 blocks_for(int n) { return dim3((n + 255) / 256); }

void multidot(cody_ctx* c, PropBuffers* b, const double2* V, int nv, const double2* y,
              const double* metric, double2* out) {
  const int n = c->nS;
  hipLaunchKernelGGL(k_multidot, dim3(kDotBlocks), dim3(256), 0, c->stream, V, (size_t)n, nv, y, n,
                     c->nC, c->nG, metric, b->partial);
  hipLaunchKernelGGL(k_multidot_final, dim3(1), dim3(kMaxKrylov), 0, c->stream, b->partial, kDotBlocks,
                     nv, out);
}

// One Lanczos sub-step from b->work over dt (PerformLanczosStep, :207-268).
// Returns 0 on success and sets *conv / *err.
int lanczos_step(cody_ctx* c, PropBuffers* b, int maxDim, double tol, const double* metric,
                 double tbase, double dt, int* conv, double* err) {
  const int n = c->nS;
  const dim3 g = blocks_for(n);
  hipStream_t st = c->stream;
  // norm of the input
  multidot(c, b, b->work, 1, b->work, metric, b->nrm2);
  HIPCHK(hipMemcpyAsync(b->h_flags, b->nrm2, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double zt = 1.0e-14;
  if (std::sqrt(b->h_flags[0]) <= zt) {
    HIPCHK(hipMemcpyAsync(b->trial, b->work, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
    *conv = 1;
    *err = 0.0;
    return 0;
  }
  hipLaunchKernelGGL(k_scale_by_norm, g, dim3(256), 0, st, b->work, b->nrm2, 1.0, 1, b->basis, n);
  const double tev = tbase + 0.5 * dt;
  int mDim = maxDim;
  for (int j = 0; j < maxDim; ++j) {
    double2* v = b->basis + (size_t)j * n;
    int r = enqueue_rhs(c, v, b->deriv, tev, kModeLin, st, nullptr, nullptr);
    if (r) return r;
    hipLaunchKernelGGL(k_lanczos_w, g, dim3(256), 0, st, b->deriv,
                       j > 0 ? b->basis + (size_t)(j - 1) * n : v, b->beta, j, b->w, n);
    multidot(c, b, v, 1, b->w, metric, b->dots);
    hipLaunchKernelGGL(k_store_alpha, dim3(1), dim3(1), 0, st, b->dots, b->alpha, j);
    hipLaunchKernelGGL(k_axpy_real, g, dim3(256), 0, st, b->alpha, j, v, b->w, n);
    // full re-orthogonalisation against v_0..v_j, done twice (classical Gram-Schmidt)
    for (int pass = 0; pass < 2; ++pass) {
      multidot(c, b, b->basis, j + 1, b->w, metric, b->dots);
      hipLaunchKernelGGL(k_sub_combination, g, dim3(256), 0, st, b->basis, (size_t)n, j + 1, b->dots, b->w, n);
    }
    multidot(c, b, b->w, 1, b->w, metric, b->dots);
    hipLaunchKernelGGL(k_store_beta, dim3(1), dim3(1), 0, st, b->dots, b->beta, j);
    hipLaunchKernelGGL(k_tridiag_expm, dim3(1), dim3(256), 0, st, b->alpha, b->beta, j + 1, dt, b->phi,
                       b->flags);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->h_flags, b->flags, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double e = b->h_flags[0];
    const bool breakdown = b->h_flags[1] <= zt;
    if (breakdown) e = 0.0;
    *err = e;
    *conv = e <= tol;
    if (*conv || breakdown || j == maxDim - 1) {
      mDim = j + 1;
      break;
    }
    // v_{j+1} = w / beta_j
    hipLaunchKernelGGL(k_scale_by_norm, g, dim3(256), 0, st, b->w, b->dots, 1.0, 1,
                       b->basis + (size_t)(j + 1) * n, n);
  }
  hipLaunchKernelGGL(k_combine, g, dim3(256), 0, st, b->basis, (size_t)n, mDim, b->phi, b->nrm2, b->trial, n);
  HIPCHK(hipGetLastError());
  return 0;
}

int sil_impl(cody_ctx* c, double t0, double t1, int krylovDim, double tol, int maxRestarts,
             int useMetric, int* converged) {
  const double zt = 1.0e-14;
  if (converged) *converged = 1;
  if (std::fabs(t1 - t0) <= zt) return CODY_OK;
  const int maxDim = std::max(1, std::min(krylovDim, c->nS));
  PropBuffers* b = buffers(c, maxDim);
  if (!b) return perr(CODY_ERR_HIP, "SIL buffer allocation failed");
  const double* metric = useMetric ? c->d_w : nullptr;
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(b->work, c->d_state, sizeof(double2) * c->nS, hipMemcpyDeviceToDevice, st));
  double tc = t0, rem = t1 - t0;
  int all = 1;
  while (std::fabs(rem) > zt) {
    double dts = rem;
    int restarts = 0, conv = 0;
    double err = 0.0;
    for (;;) {
      int r = lanczos_step(c, b, maxDim, tol, metric, tc, dts, &conv, &err);
      if (r) return r;
      if (conv) break;
      if (std::fabs(dts) <= zt || restarts >= maxRestarts) break;
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
  const int n = c->nS;
  const dim3 g = blocks_for(n);
  hipStream_t st = c->stream;
  const double dt = t1 - t0, dt2 = dt / 2.0, dt3 = dt / 3.0, dt6 = dt / 6.0;
  int r;
  if ((r = run_rhs(c, c->d_state, b->rk_d, t0, mode, st))) return r;
  hipLaunchKernelGGL(k_rk_stage, g, dim3(256), 0, st, c->d_state, b->rk_d, dt6, dt2, 1, b->rk_inc, b->rk_tmp, n);
  if ((r = run_rhs(c, b->rk_tmp, b->rk_d, t0 + dt2, mode, st))) return r;
  hipLaunchKernelGGL(k_rk_stage, g, dim3(256), 0, st, c->d_state, b->rk_d, dt3, dt2, 0, b->rk_inc, b->rk_tmp, n);
  if ((r = run_rhs(c, b->rk_tmp, b->rk_d, t0 + dt2, mode, st))) return r;
  hipLaunchKernelGGL(k_rk_stage, g, dim3(256), 0, st, c->d_state, b->rk_d, dt3, dt, 0, b->rk_inc, b->rk_tmp, n);
  if ((r = run_rhs(c, b->rk_tmp, b->rk_d, t0 + dt, mode, st))) return r;
  hipLaunchKernelGGL(k_rk_final, g, dim3(256), 0, st, c->d_state, b->rk_inc, b->rk_d, dt6, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
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
  // A(1/2) B(1) A(1/2): S_Propagator_SplitStep_Order2.f90:61-75
  const double dt = t1 - t0, t2 = t0 + 0.5 * dt;
  int ok1 = 1, ok2 = 1;
  if ((r = sil_impl(c, t0, t0 + 0.5 * dt, krylovDim, tol, maxRestarts, useMetric, &ok1))) return r;
  if ((r = rk4_impl(c, t0, t0 + dt, kModeNonLin))) return r;
  if ((r = sil_impl(c, t2, t2 + 0.5 * dt, krylovDim, tol, maxRestarts, useMetric, &ok2))) return r;
  if (converged) *converged = ok1 && ok2;
  return CODY_OK;
}

}  // extern "C"
