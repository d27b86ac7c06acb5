// CDNA4 (gfx950) kernels of the MCTDHX right-hand side.
//
// Data layout in HBM (one context, one fermionic body type, nO orbitals, nG grid points):
//   state      complex128 [C(nC) ; phi(nG, nO) column-major]   (the reference's Method_state)
//   V          FP64 nGp x nGp row-major, V[i][j] = K(|x_i - x_j|), zero padded to nGp = 64*ceil(nG/64)
//   ST         FP64 Pc x nGp: row 2p / 2p+1 = Re / Im of the pair source conj(phi_l) phi_k w,
//              p enumerating l <= k; Pc = 16*ceil(2P/16), unused rows zero
//   Wpart      FP64 nsplit x nGp x Pc: split-K partial products V . S
//   WT         FP64 Pc x nGp: mean-field potentials W_p(x) (Re/Im rows), summed over splits
// Complex scalars are double2 (x = re, y = im).
#include <hip/hip_runtime.h>

#include "mctdhx_kernels.hpp"

namespace cody {

// ------------------------------------------------------------------ complex helpers
__device__ __forceinline__ double2 cmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return cmk(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return cmk(a.x * s, a.y * s); }
// acc += a * b
__device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
  acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
// acc += conj(a) * b
__device__ __forceinline__ void cfmac(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
  acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
}
__device__ __forceinline__ double2 mul_minus_i(double2 a) { return cmk(a.y, -a.x); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int pair_index(int nO, int l, int k) {  // l <= k
  return l * nO - (l * (l - 1)) / 2 + (k - l);
}

__device__ __forceinline__ int idx4(int a, int b, int c, int d, int n) {
  return a + n * (b + n * (c + n * d));
}

// ------------------------------------------------------------------ interaction matrix
// V[i][j] = s e^{-d r}/(sqrt(r^2 + a1^2) + a2), r = |x_i - x_j|
// (src/SysInteraction/Linear/SoftYukawa/S_SysInteraction_Linear_SoftYukawa.f90:67-73), padded with 0.
__global__ void k_fill_softyukawa(const double* __restrict__ x, int nG, int nGp, double s,
                                  double a1, double a2, double damp, double* __restrict__ V) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nGp) return;
  double v = 0.0;
  if (i < nG && j < nG) {
    const double r = fabs(x[i] - x[j]);
    v = s * exp(-damp * r) / (sqrt(r * r + a1 * a1) + a2);
  }
  V[(size_t)i * nGp + j] = v;
}

// ------------------------------------------------------------------ pair sources
// ST[2p][x] + i ST[2p+1][x] = conj(phi_l(x)) phi_k(x) w(x), p = (l <= k)
// (FillInteractionSrc, src/SysInteraction/Linear/S_SysInteraction_Linear.f90:57-72).
template <int NO>
__global__ __launch_bounds__(256) void k_pair_src(const double2* __restrict__ orbs,
                                                  const double* __restrict__ w, int nG, int nGp,
                                                  int Pc, double* __restrict__ ST) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nGp) return;
  double2 phi[NO];
  const bool in = x < nG;
  const double wx = in ? w[x] : 0.0;
#pragma unroll
  for (int k = 0; k < NO; ++k) phi[k] = in ? orbs[(size_t)k * nG + x] : cmk(0, 0);
  int c = 0;
#pragma unroll
  for (int l = 0; l < NO; ++l)
#pragma unroll
    for (int k = l; k < NO; ++k) {
      const double2 s = cscale(cmul(cconj(phi[l]), phi[k]), wx);
      ST[(size_t)c * nGp + x] = s.x;
      ST[(size_t)(c + 1) * nGp + x] = s.y;
      c += 2;
    }
  for (; c < Pc; ++c) ST[(size_t)c * nGp + x] = 0.0;
}

// ------------------------------------------------------------------ mean-field contraction (FP64 MFMA)
// Wpart[s][r][c] = sum_{k in split s} V[r][k] S[k][c], S[k][c] = ST[c][k].
// v_mfma_f64_16x16x4f64: lane l holds A(row l%16, k l/16), B(k l/16, col l%16) and
// D(row 4i + l/16, col l%16), i = 0..3.  The contraction index is permuted inside each
// 16-wide k block (MFMA step s of lane group q uses k = kb + 4q + s) so that every lane
// fetches its four A and four B values with one 32-byte load each.
template <int NT>
__global__ __launch_bounds__(256) void k_meanfield_f64(const double* __restrict__ V,
                                                       const double* __restrict__ ST,
                                                       double* __restrict__ Wpart, int nGp, int Pc,
                                                       int kchunk) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane >> 4, rr = lane & 15;
  const int row = blockIdx.x * 64 + wave * 16 + rr;
  const int split = blockIdx.y;
  const int k0 = split * kchunk;
  typedef double double4_ __attribute__((ext_vector_type(4)));
  double4_ acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = (double4_){0.0, 0.0, 0.0, 0.0};
  const double* aptr = V + (size_t)row * nGp + k0 + 4 * q;
  const double* bptr[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) bptr[t] = ST + (size_t)(t * 16 + rr) * nGp + k0 + 4 * q;
  for (int kb = 0; kb < kchunk; kb += 16) {
    const double4_ a = *reinterpret_cast<const double4_*>(aptr + kb);
    double4_ b[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) b[t] = *reinterpret_cast<const double4_*>(bptr[t] + kb);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[t][s], acc[t], 0, 0, 0);
    }
  }
  // FP64 16x16 result layout: accumulator element i of lane l is row 4i + l/16, column l%16
  double* out = Wpart + (size_t)split * nGp * Pc;
  const int orow = blockIdx.x * 64 + wave * 16 + q;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(size_t)(orow + 4 * i) * Pc + t * 16 + rr] = acc[t][i];
}

// Sum the split-K partials in fixed order and transpose to WT[c][x].
__global__ __launch_bounds__(256) void k_wsum(const double* __restrict__ Wpart, int nsplit, int nGp,
                                              int Pc, double* __restrict__ WT) {
  __shared__ double tile[32][kMaxPc + 1];
  const int r0 = blockIdx.x * 32;
  for (int e = threadIdx.x; e < 32 * Pc; e += blockDim.x) {
    const int r = e / Pc, c = e % Pc;
    double s = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) s += Wpart[((size_t)sp * nGp + r0 + r) * Pc + c];
    tile[r][c] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * Pc; e += blockDim.x) {
    const int c = e / 32, r = e % 32;
    WT[(size_t)c * nGp + r0 + r] = tile[r][c];
  }
}

// ------------------------------------------------------------------ one-body operator at a point
// FEDVR weak-form kinetic energy -(1/2m)(1/w_x) sum_e sum_b K(a,b)/h f_b
// (src/Utils/M_Utils_DerivativeFedvr.f90:121-152): the point with global index g (numbered with
// the excluded xmin node counted) is node a = g mod (nLoc-1) of element g div (nLoc-1) and, for
// a = 0, also node nLoc-1 of the previous element.  Excluded end nodes contribute f = 0.
struct KinParams {
  const double* kref;
  int nE, nLoc, xminExcl, nG;
  double invh, fac;  // 1/h, -1/(2m)
};

__device__ __forceinline__ double2 kinetic_at(const double2* __restrict__ f, int x,
                                              const KinParams& kp, double rw) {
  const int gi = x + kp.xminExcl;
  const int per = kp.nLoc - 1;
  double2 acc = cmk(0, 0);
  int e = gi / per, a = gi % per;
  // element e (node a), if it exists
  for (int pass = 0; pass < 2; ++pass) {
    int ee = e, aa = a;
    if (pass == 1) {
      if (a != 0 || e == 0) break;
      ee = e - 1;
      aa = per;
    }
    if (ee >= kp.nE) continue;
    const int base = ee * per - kp.xminExcl;
    for (int b = 0; b < kp.nLoc; ++b) {
      const int j = base + b;
      if (j < 0 || j >= kp.nG) continue;
      const double kab = kp.kref[aa + b * kp.nLoc];
      const double2 fj = f[j];
      acc.x = fma(kab, fj.x, acc.x);
      acc.y = fma(kab, fj.y, acc.y);
    }
  }
  return cscale(acc, kp.invh * rw * kp.fac);
}

// ------------------------------------------------------------------ orbital right-hand side
// One workgroup = 64 grid points (one per lane) x 4 orbitals j (one per wave).
//   hphi_j   = (T + V_ext + E(t) x) phi_j                        (ApplySingleBodyOp :976-1034)
//   corr_j   = sum_{k1,(l2,k2)} W_{l2k2} phi_k1 rho2'(k1,k2,j,l2) (ApplyCorrelationOp :762-838)
//   draw_j   = [hphi_j] + [corr_j]  (mode)
// and per-workgroup partial sums of h1(i,j) = <phi_i|hphi_j>_w and inp(i,j) = <phi_i|draw_j>_w.
// The rho2' coefficients depend only on the wave's orbital j and the loop indices, so their
// loads are uniform across the wave.
template <int NO>
__global__ __launch_bounds__(256) void k_orbital(
    const double2* __restrict__ orbs, const double* __restrict__ w,
    const double* __restrict__ xg, const double* __restrict__ vext, KinParams kp, double efield,
    const double* __restrict__ WT, const double2* __restrict__ rho2p, int nG, int nGp, int mode,
    double2* __restrict__ draw, double2* __restrict__ part) {
  constexpr int n2 = NO * NO;
  const int lane = threadIdx.x & 63;
  const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + (threadIdx.x >> 6));
  const int x = blockIdx.x * 64 + lane;
  if (j >= NO) return;
  const bool in = x < nG;
  const int xc = in ? x : nG - 1;
  double2 phi[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) phi[k] = in ? orbs[(size_t)k * nG + xc] : cmk(0, 0);
  const double wx = in ? w[xc] : 0.0;
  double2 hphi = cmk(0, 0), corr = cmk(0, 0);
  if (in) {
    hphi = kinetic_at(orbs + (size_t)j * nG, xc, kp, 1.0 / wx);
    const double v = vext[xc] + efield * xg[xc];
    double2 pj = orbs[(size_t)j * nG + xc];
    hphi.x = fma(v, pj.x, hphi.x);
    hphi.y = fma(v, pj.y, hphi.y);
  }
  if (mode != kModeLin) {
    const double2* rj = rho2p + n2 * j;
    int p = 0;
#pragma unroll
    for (int l = 0; l < NO; ++l) {
#pragma unroll
      for (int k = l; k < NO; ++k, ++p) {
        const double2 W = in ? cmk(WT[(size_t)(2 * p) * nGp + x], WT[(size_t)(2 * p + 1) * nGp + x])
                             : cmk(0, 0);
        // A = sum_k1 phi_k1 rho2'(k1,k,j,l); B = sum_k1 phi_k1 rho2'(k1,l,j,k)
        const double2* ra = rj + NO * k + n2 * NO * l;
        double2 A = cmk(0, 0);
#pragma unroll
        for (int k1 = 0; k1 < NO; ++k1) cfma(A, phi[k1], ra[k1]);
        cfma(corr, W, A);
        if (k != l) {
          const double2* rb = rj + NO * l + n2 * NO * k;
          double2 B = cmk(0, 0);
#pragma unroll
          for (int k1 = 0; k1 < NO; ++k1) cfma(B, phi[k1], rb[k1]);
          cfma(corr, cconj(W), B);
        }
      }
    }
  }
  double2 d = cmk(0, 0);
  if (mode == kModeFull) d = cadd(hphi, corr);
  else if (mode == kModeNonLin) d = corr;
  else d = hphi;
  if (in) draw[(size_t)j * nG + x] = d;
  // partials: part[block][0][i + j nO] = h1, part[block][1][i + j nO] = inp
  double2* pb = part + (size_t)blockIdx.x * 2 * n2;
  const double2 hw = cscale(hphi, wx), dw = cscale(d, wx);
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    double2 h = cmk(0, 0), q = cmk(0, 0);
    cfmac(h, phi[i], hw);
    cfmac(q, phi[i], dw);
    h.x = wave_sum(h.x);
    h.y = wave_sum(h.y);
    q.x = wave_sum(q.x);
    q.y = wave_sum(q.y);
    if (lane == 0) {
      pb[i + j * NO] = h;
      pb[n2 + i + j * NO] = q;
    }
  }
}

// ------------------------------------------------------------------ h2 Gram matrix
// Gpart[s][c1][c2] = sum_{x in chunk s} ST[c1][x] WT[c2][x]  (16 x 16 tile per workgroup).
__global__ __launch_bounds__(256) void k_h2_gram(const double* __restrict__ ST,
                                                 const double* __restrict__ WT, int nGp, int Pc,
                                                 int xchunk, double* __restrict__ Gpart) {
  __shared__ double sa[16][65], sb[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int c1 = blockIdx.x * 16 + ty, c2 = blockIdx.y * 16 + tx;
  const int x0 = blockIdx.z * xchunk;
  double acc = 0.0;
  for (int xb = 0; xb < xchunk; xb += 64) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int r = e / 64, cx = e % 64;
      sa[r][cx] = ST[(size_t)(blockIdx.x * 16 + r) * nGp + x0 + xb + cx];
      sb[r][cx] = WT[(size_t)(blockIdx.y * 16 + r) * nGp + x0 + xb + cx];
    }
    __syncthreads();
#pragma unroll 8
    for (int cx = 0; cx < 64; ++cx) acc = fma(sa[ty][cx], sb[tx][cx], acc);
    __syncthreads();
  }
  Gpart[((size_t)blockIdx.z * Pc + c1) * Pc + c2] = acc;
}

// ------------------------------------------------------------------ projection + scaling
// dphi_j = -i (draw_j - sum_i phi_i inp(i,j))  (Grid_ProjectOnSubspace, src/Grid/S_Grid.f90:170-203)
template <int NO>
__global__ __launch_bounds__(256) void k_project(const double2* __restrict__ orbs,
                                                 const double2* __restrict__ draw,
                                                 const double2* __restrict__ inp, int nG,
                                                 int project, double2* __restrict__ dorbs) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nG) return;
  double2 phi[NO];
  if (project) {
#pragma unroll
    for (int i = 0; i < NO; ++i) phi[i] = orbs[(size_t)i * nG + x];
  }
#pragma unroll
  for (int j = 0; j < NO; ++j) {
    double2 d = draw[(size_t)j * nG + x];
    if (project) {
#pragma unroll
      for (int i = 0; i < NO; ++i) d = csub(d, cmul(phi[i], inp[i + j * NO]));
    }
    dorbs[(size_t)j * nG + x] = mul_minus_i(d);
  }
}

// ------------------------------------------------------------------ zero fill (graph-capturable)
__global__ void k_zero(double2* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_double2(0.0, 0.0);
}

// ------------------------------------------------------------------ launchers
void launch_zero(double2* p, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}

void launch_fill_softyukawa(const double* x, int nG, int nGp, double s, double a1, double a2,
                            double damp, double* V, hipStream_t st) {
  dim3 grid((nGp + 255) / 256, nGp);
  hipLaunchKernelGGL(k_fill_softyukawa, grid, dim3(256), 0, st, x, nG, nGp, s, a1, a2, damp, V);
}

#define CODY_NO_SWITCH(NOV, CALL) \
  switch (NOV) {                    \
    case 1: CALL(1); break;         \
    case 2: CALL(2); break;         \
    case 3: CALL(3); break;         \
    case 4: CALL(4); break;         \
    case 5: CALL(5); break;         \
    case 6: CALL(6); break;         \
    case 7: CALL(7); break;         \
    case 8: CALL(8); break;         \
    case 9: CALL(9); break;         \
    case 10: CALL(10); break;       \
    case 11: CALL(11); break;       \
    case 12: CALL(12); break;       \
    default: break;                 \
  }

void launch_pair_src(const double2* orbs, const double* w, int nG, int nGp, int nO, int Pc,
                     double* ST, hipStream_t st) {
#define CALL_PS(NO)                                                                            \
  hipLaunchKernelGGL(k_pair_src<NO>, dim3((nGp + 255) / 256), dim3(256), 0, st, orbs, w, nG, nGp, \
                     Pc, ST)
  CODY_NO_SWITCH(nO, CALL_PS)
#undef CALL_PS
}

template <int NT>
static void launch_mf(const double* V, const double* ST, double* Wpart, int nGp, int Pc, int nsplit,
                      hipStream_t st) {
  dim3 grid(nGp / 64, nsplit);
  hipLaunchKernelGGL(k_meanfield_f64<NT>, grid, dim3(256), 0, st, V, ST, Wpart, nGp, Pc,
                     nGp / nsplit);
}

void launch_meanfield_f64(const double* V, const double* ST, double* Wpart, int nGp, int Pc,
                          int nsplit, hipStream_t st) {
  switch (Pc / 16) {
    case 1: launch_mf<1>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 2: launch_mf<2>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 3: launch_mf<3>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 4: launch_mf<4>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 5: launch_mf<5>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 6: launch_mf<6>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 7: launch_mf<7>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 8: launch_mf<8>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 9: launch_mf<9>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    case 10: launch_mf<10>(V, ST, Wpart, nGp, Pc, nsplit, st); break;
    default: break;
  }
}

void launch_wsum(const double* Wpart, int nsplit, int nGp, int Pc, double* WT, hipStream_t st) {
  hipLaunchKernelGGL(k_wsum, dim3(nGp / 32), dim3(256), 0, st, Wpart, nsplit, nGp, Pc, WT);
}

void launch_orbital(const double2* orbs, const double* w, const double* x, const double* vext,
                    const KinLaunch& kl, double efield, const double* WT, const double2* rho2p,
                    int nG, int nGp, int nO, int mode, double2* draw, double2* part,
                    hipStream_t st) {
  KinParams kp{kl.kref, kl.nE, kl.nLoc, kl.xminExcl, nG, kl.invh, kl.fac};
  dim3 grid(nGp / 64, (nO + 3) / 4);
#define CALL_OR(NO)                                                                          \
  hipLaunchKernelGGL(k_orbital<NO>, grid, dim3(256), 0, st, orbs, w, x, vext, kp, efield, WT, \
                     rho2p, nG, nGp, mode, draw, part)
  CODY_NO_SWITCH(nO, CALL_OR)
#undef CALL_OR
}

void launch_h2_gram(const double* ST, const double* WT, int nGp, int Pc, int nxs, double* Gpart,
                    hipStream_t st) {
  dim3 grid(Pc / 16, Pc / 16, nxs);
  hipLaunchKernelGGL(k_h2_gram, grid, dim3(256), 0, st, ST, WT, nGp, Pc, nGp / nxs, Gpart);
}

void launch_project(const double2* orbs, const double2* draw, const double2* inp, int nG, int nO,
                    int project, double2* dorbs, hipStream_t st) {
#define CALL_PR(NO)                                                                                 \
  hipLaunchKernelGGL(k_project<NO>, dim3((nG + 255) / 256), dim3(256), 0, st, orbs, draw, inp, nG, \
                     project, dorbs)
  CODY_NO_SWITCH(nO, CALL_PR)
#undef CALL_PR
}

}  // namespace cody
