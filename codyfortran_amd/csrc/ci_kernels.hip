// CI-side kernels of the MCTDHX right-hand side (the CI space is small, n_C <= 1024):
//   k_ci_rdm      : rho1, rho2 from C, one wave per density element
//   k_ci_reginv   : R = rho1 (rho1^2 + eps)^{-1} (one wave)
//   k_gram_reduce : G = S^T W from the split-K partials of the Gram kernel; h1 from its partials
//   k_ci_apply    : CI H-apply with h2 formed from G, one wave per configuration
//   k_ci_energy   : E = Tr[rho1 h1] + Tr[rho2 h2]
//   k_sum_parts   : h1 / inp from per-workgroup partials
#include <hip/hip_runtime.h>

#include "mctdhx_kernels.hpp"

namespace cody {

namespace {

__device__ __forceinline__ double2 cmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return cmk(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return cmk(a.x * s, a.y * s); }
__device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
  acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return cmk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ int idx4(int a, int b, int c, int d, int n) {
  return a + n * (b + n * (c + n * d));
}
// index of the strict pair (a < b) among the nO(nO-1)/2 strict pairs
__device__ __forceinline__ int strict_pair(int nO, int a, int b) {
  return a * nO - (a * (a + 1)) / 2 + (b - a - 1);
}
__device__ __forceinline__ int pair_index(int nO, int l, int k) {  // l <= k
  return l * nO - (l * (l - 1)) / 2 + (k - l);
}

}  // namespace

// ------------------------------------------------------------------------------------ densities
__device__ __forceinline__ void strict_pair_decode(int nO, int p, int& a, int& b) {
  a = 0;
  while (p >= nO - 1 - a) {
    p -= nO - 1 - a;
    ++a;
  }
  b = a + 1 + p;
}

// One wave per density element: items 0..n2-1 are rho1(i,j) (e = i + j nO, code i nO + j),
// items n2.. the canonical rho2Lin (i1 < i2, j1 < j2; e = strict(i1,i2) + np strict(j1,j2), code
// ((i1 nO + j1) nO + i2) nO + j2).  Lanes stride over the element's CSR entries
// sum_q f_q conj(C_jc) C_ic (S_Coeffs_Generic.f90:362-434 builds the same sums), so the gathers
// of one element are issued together.  Each canonical rho2 value is also written to its four
// antisymmetric positions of the full rho2; entries with i1 = i2 or j1 = j2 are never written
// (they are zero from allocation).
__global__ __launch_bounds__(256) void k_ci_rdm(const double2* __restrict__ C, int nC, int nO,
                                                const int* __restrict__ r1_ptr,
                                                const CiEntryDev* __restrict__ r1,
                                                const int* __restrict__ r2_ptr,
                                                const CiEntryDev* __restrict__ r2,
                                                double2* __restrict__ rho1, double2* __restrict__ r2can,
                                                double2* __restrict__ rho2) {
  const int n2 = nO * nO, np = nO * (nO - 1) / 2, nCan = np * np;
  const int item = blockIdx.x * 4 + (threadIdx.x >> 6), ln = threadIdx.x & 63;
  extern __shared__ double2 sC[];  // C staged once per workgroup
  for (int i = threadIdx.x; i < nC; i += blockDim.x) sC[i] = C[i];
  __syncthreads();
  if (item >= n2 + nCan) return;
  int code, i1 = 0, i2 = 0, j1 = 0, j2 = 0;
  const int* ptr;
  const CiEntryDev* tab;
  if (item < n2) {
    code = (item % nO) * nO + item / nO;
    ptr = r1_ptr;
    tab = r1;
  } else {
    const int e = item - n2;
    strict_pair_decode(nO, e % np, i1, i2);
    strict_pair_decode(nO, e / np, j1, j2);
    code = ((i1 * nO + j1) * nO + i2) * nO + j2;
    ptr = r2_ptr;
    tab = r2;
  }
  double2 acc = cmk(0, 0);
  for (int q = ptr[code] + ln; q < ptr[code + 1]; q += 64) {
    const CiEntryDev en = tab[q];
    cfma(acc, cscale(cconj(sC[en.jc]), en.f), sC[en.ic]);
  }
  acc.x = wave_sum(acc.x);
  acc.y = wave_sum(acc.y);
  if (ln != 0) return;
  if (item < n2) {
    rho1[item] = acc;
  } else {
    r2can[item - n2] = acc;
    const double2 m = cmk(-acc.x, -acc.y);
    rho2[idx4(i1, i2, j1, j2, nO)] = acc;
    rho2[idx4(i2, i1, j1, j2, nO)] = m;
    rho2[idx4(i1, i2, j2, j1, nO)] = m;
    rho2[idx4(i2, i1, j2, j1, nO)] = acc;
  }
}

// R = rho1 (rho1^2 + eps)^{-1} (regInvMat, S_Method_Mb_OrbBased.f90:737-744) by Gauss-Jordan on
// rho1^2 + eps I (Hermitian positive definite: no pivoting) and one refinement step, in one wave
// (the barriers of a one-wave workgroup cost nothing).  rho2' = rho2 x_3 R is formed slice by
// slice inside the orbital kernel.
template <int NO>
__global__ __launch_bounds__(64) void k_ci_reginv(const double2* __restrict__ rho1g, double reg,
                                                  double2* __restrict__ reginv) {
  constexpr int n2 = NO * NO;
  __shared__ double2 s1[n2], sA0[n2], sM[2 * n2], sX[n2], sR[n2], sF[NO];
  const int t = threadIdx.x;
  for (int i = t; i < n2; i += 64) s1[i] = rho1g[i];
  __syncthreads();
  for (int e = t; e < n2; e += 64) {
    const int i = e % NO, j = e / NO;
    double2 acc = cmk(i == j ? reg : 0.0, 0.0);
#pragma unroll
    for (int k = 0; k < NO; ++k) cfma(acc, s1[i + k * NO], s1[k + j * NO]);
    sA0[e] = acc;
    sM[e] = acc;
    sM[n2 + e] = cmk(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int k = 0; k < NO; ++k) {
    const double2 piv = sM[k + k * NO];
    __syncthreads();
    if (t < 2 * NO) sM[k + t * NO] = cdiv(sM[k + t * NO], piv);
    if (t < NO) sF[t] = sM[t + k * NO];
    __syncthreads();
    for (int e = t; e < 2 * n2; e += 64) {
      const int r = e % NO, c = e / NO;
      if (r != k) sM[r + c * NO] = csub(sM[r + c * NO], cmul(sF[r], sM[k + c * NO]));
    }
    __syncthreads();
  }
  // X = A^-1 rho1, then one refinement X += A^-1 (rho1 - A X)
  const double2* sInv = sM + n2;
  for (int e = t; e < n2; e += 64) {
    const int i = e % NO, j = e / NO;
    double2 acc = cmk(0, 0);
#pragma unroll
    for (int k = 0; k < NO; ++k) cfma(acc, sInv[i + k * NO], s1[k + j * NO]);
    sX[e] = acc;
  }
  __syncthreads();
  for (int e = t; e < n2; e += 64) {
    const int i = e % NO, j = e / NO;
    double2 r = s1[e];
#pragma unroll
    for (int k = 0; k < NO; ++k) r = csub(r, cmul(sA0[i + k * NO], sX[k + j * NO]));
    sR[e] = r;  // residual
  }
  __syncthreads();
  for (int e = t; e < n2; e += 64) {
    const int i = e % NO, j = e / NO;
    double2 acc = sX[e];
#pragma unroll
    for (int k = 0; k < NO; ++k) cfma(acc, sInv[i + k * NO], sR[k + j * NO]);
    reginv[e] = acc;
  }
}

// ------------------------------------------------------------------------------------ partials
// out[e] = sum_b part[b stride + e], fixed order (h1 and inp from per-workgroup partials).
__global__ __launch_bounds__(256) void k_sum_parts(const double2* __restrict__ part, int nPart, int stride,
                                                   int n, double2* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double2 acc = cmk(0, 0);
#pragma unroll 8
  for (int b = 0; b < nPart; ++b) acc = cadd(acc, part[(size_t)b * stride + e]);
  out[e] = acc;
}

// G[c1][c2] = sum_s Gpart[s][c2][c1] = sum_x S_c1(x) W_c2(x), fixed order; threads past Pc^2 form
// h1 = sum_b part1[b][e] from the one-body kernel's per-workgroup partials.
__global__ __launch_bounds__(256) void k_gram_reduce(const double* __restrict__ Gpart, int nxs, int Pc,
                                                     double* __restrict__ G, const double2* __restrict__ part1,
                                                     int nPart1, int n2, double2* __restrict__ h1) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = Pc * Pc;
  if (g >= n) {
    const int e = g - n;
    if (e >= n2) return;
    double2 acc = cmk(0, 0);
#pragma unroll 8
    for (int b = 0; b < nPart1; ++b) acc = cadd(acc, part1[(size_t)b * n2 + e]);
    h1[e] = acc;
    return;
  }
  double acc = 0.0;
#pragma unroll 8
  for (int s = 0; s < nxs; ++s) acc += Gpart[(size_t)s * n + g];
  const int c2 = g / Pc, c1 = g % Pc;
  G[c1 * Pc + c2] = acc;
}

// h2(i1,a,j1,b) = 1/2 sum_x src_(i1,j1)(x) W_(a,b)(x) from G = S^T W (FillH2CorrelationBlock,
// S_Method_Mb_OrbBased.f90:841-872); src_(i,j) = conj(src_(j,i)), W_(a,b) = conj(W_(b,a)).
__device__ __forceinline__ double2 h2_from_G(const double* __restrict__ G, int Pc, int nO, int i1,
                                             int a, int j1, int b) {
  const int p1 = i1 <= j1 ? pair_index(nO, i1, j1) : pair_index(nO, j1, i1);
  const int p2 = a <= b ? pair_index(nO, a, b) : pair_index(nO, b, a);
  const double s1 = i1 <= j1 ? 1.0 : -1.0, s2 = a <= b ? 1.0 : -1.0;
  const double rr = G[(2 * p1) * Pc + 2 * p2], ii = G[(2 * p1 + 1) * Pc + 2 * p2 + 1];
  const double ri = G[(2 * p1) * Pc + 2 * p2 + 1], ir = G[(2 * p1 + 1) * Pc + 2 * p2];
  return cmk(0.5 * (rr - s1 * s2 * ii), 0.5 * (s2 * ri + s1 * ir));
}

// antisymmetrised h2Lin(i1 < i2, j1 < j2) (FillH2Lin, S_Coeffs_Generic.f90:478-542)
__device__ __forceinline__ double2 h2lin_from_G(const double* __restrict__ G, int Pc, int nO, int i1,
                                                int i2, int j1, int j2) {
  double2 v = h2_from_G(G, Pc, nO, i1, i2, j1, j2);
  v = csub(v, h2_from_G(G, Pc, nO, i1, i2, j2, j1));
  v = csub(v, h2_from_G(G, Pc, nO, i2, i1, j1, j2));
  return cadd(v, h2_from_G(G, Pc, nO, i2, i1, j2, j1));
}

// CI H-apply, one wave per configuration ic, lanes over its connections (ApplyH of
// S_Coeffs_Generic.f90 through the singles / doubles tables):
//   dCraw(ic) = sum_singles f h1(i,j) C(jc) + sum_doubles f h2Lin(i1,i2,j1,j2) C(jc)
// (the projection and the factor -i are applied by k_project).
__global__ __launch_bounds__(256) void k_ci_apply(const double2* __restrict__ C, int nC, int nO, int Pc,
                                                  const double* __restrict__ G, const double2* __restrict__ h1,
                                                  const int* __restrict__ s_ptr,
                                                  const CiEntryDev* __restrict__ sing,
                                                  const int* __restrict__ d_ptr,
                                                  const CiEntryDev* __restrict__ dbl,
                                                  double2* __restrict__ dCraw) {
  const int ic = blockIdx.x * 4 + (threadIdx.x >> 6), ln = threadIdx.x & 63;
  if (ic >= nC) return;
  double2 acc = cmk(0, 0);
  for (int q = s_ptr[ic] + ln; q < s_ptr[ic + 1]; q += 64) {
    const CiEntryDev en = sing[q];
    const int i = en.code / nO, j = en.code % nO;
    cfma(acc, cscale(h1[i + j * nO], en.f), C[en.jc]);
  }
  for (int q = d_ptr[ic] + ln; q < d_ptr[ic + 1]; q += 64) {
    const CiEntryDev en = dbl[q];
    int code = en.code;
    const int j2 = code % nO;
    code /= nO;
    const int i2 = code % nO;
    code /= nO;
    const int j1 = code % nO;
    const int i1 = code / nO;
    if (i1 == i2 || j1 == j2) continue;
    cfma(acc, cscale(h2lin_from_G(G, Pc, nO, i1, i2, j1, j2), en.f), C[en.jc]);
  }
  acc.x = wave_sum(acc.x);
  acc.y = wave_sum(acc.y);
  if (ln == 0) dCraw[ic] = acc;
}

// the full h2(i1,a,j1,b) tensor from G (the h2_ output of ApplyCorrelationOp,
// S_Method_Mb_OrbBased.f90:841-872), one element per thread
__global__ __launch_bounds__(256) void k_h2_full(const double* __restrict__ G, int Pc, int nO,
                                                 double2* __restrict__ h2) {
  const int n2 = nO * nO, e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n2 * n2) return;
  const int i1 = e % nO, a = (e / nO) % nO, j1 = (e / n2) % nO, b = e / (n2 * nO);
  h2[e] = h2_from_G(G, Pc, nO, i1, a, j1, b);
}

// E = sum h1(k,l) rho1(l,k) + sum h2(k1,k2,l1,l2) rho2(l1,l2,k1,k2)  (Method_GetEnergy), one
// workgroup of 256 threads.
__global__ __launch_bounds__(256) void k_ci_energy(int nO, int Pc, const double* __restrict__ G,
                                                   const double2* __restrict__ h1,
                                                   const double2* __restrict__ rho1,
                                                   const double2* __restrict__ rho2,
                                                   double* __restrict__ energy) {
  __shared__ double red[8];
  const int t = threadIdx.x, n2 = nO * nO, n4 = n2 * n2;
  double2 e = cmk(0, 0);
  for (int q = t; q < n2; q += blockDim.x) {
    const int k = q % nO, l = q / nO;
    cfma(e, h1[k + l * nO], rho1[l + k * nO]);
  }
  for (int q = t; q < n4; q += blockDim.x) {
    const int k1 = q % nO, k2 = (q / nO) % nO, l1 = (q / n2) % nO, l2 = q / (n2 * nO);
    cfma(e, h2_from_G(G, Pc, nO, k1, k2, l1, l2), rho2[idx4(l1, l2, k1, k2, nO)]);
  }
  double v = wave_sum(e.x);
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
    *energy = s;
  }
}

// <x> = sum_ij rho1(i,j) <phi_j|x phi_i>_w (PrinterObservableLinear_DumpDipole,
// src/Utils/M_Utils_PrinterObservableLinear.f90:57-68, occupation 1 for unrestricted orbitals)
// from the per-block partials of k_xmat (part1[b][i + j nO] = block b's share of <phi_i|x phi_j>),
// summed in fixed order; one workgroup of 256 threads writes the real part.
__global__ __launch_bounds__(256) void k_dipole(int nO, const double2* __restrict__ part1, int nPart,
                                                const double2* __restrict__ rho1, double* __restrict__ dipole) {
  __shared__ double red[4];
  const int t = threadIdx.x, n2 = nO * nO;
  double2 e = cmk(0, 0);
  for (int q = t; q < n2; q += blockDim.x) {
    const int i = q % nO, j = q / nO;
    double2 xji = cmk(0, 0);
    for (int b = 0; b < nPart; ++b) xji = cadd(xji, part1[(size_t)b * n2 + j + i * nO]);
    cfma(e, rho1[i + j * nO], xji);
  }
  const double v = wave_sum(e.x);
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  if (t == 0) *dipole = (red[0] + red[1]) + (red[2] + red[3]);
}

// ------------------------------------------------------------------------------------ launchers
// Dynamic LDS of the density kernels: C (k_ci_rdm); the others use static LDS.
size_t ci_density_lds_bytes(int nC, int nO) {
  (void)nO;
  return sizeof(double2) * (size_t)nC;
}

int ci_kernels_allow_lds(size_t bytes) {
  (void)bytes;  // static LDS only
  return 0;
}

void launch_ci_density(const double2* C, int nC, int nO, double reg, const DevCiTables& t,
                       double2* rho1, double2* rho2, double2* r2can, double2* reginv, hipStream_t st) {
  const int n2 = nO * nO, np = nO * (nO - 1) / 2, items = n2 + np * np;
  hipLaunchKernelGGL(k_ci_rdm, dim3((items + 3) / 4), dim3(256), sizeof(double2) * nC, st, C, nC, nO, t.r1_ptr,
                     t.r1, t.r2_ptr, t.r2, rho1, r2can, rho2);
#define CALL_RI(NO) hipLaunchKernelGGL(k_ci_reginv<NO>, dim3(1), dim3(64), 0, st, rho1, reg, reginv)
  switch (nO) {
    case 1: CALL_RI(1); break;
    case 2: CALL_RI(2); break;
    case 3: CALL_RI(3); break;
    case 4: CALL_RI(4); break;
    case 5: CALL_RI(5); break;
    case 6: CALL_RI(6); break;
    case 7: CALL_RI(7); break;
    case 8: CALL_RI(8); break;
    case 9: CALL_RI(9); break;
    case 10: CALL_RI(10); break;
    case 11: CALL_RI(11); break;
    case 12: CALL_RI(12); break;
    default: break;
  }
#undef CALL_RI
}

void launch_h2_full(const double* G, int Pc, int nO, double2* h2, hipStream_t st) {
  const int n4 = nO * nO * nO * nO;
  hipLaunchKernelGGL(k_h2_full, dim3((n4 + 255) / 256), dim3(256), 0, st, G, Pc, nO, h2);
}

void launch_reginv(const double2* rho1, int nO, double reg, double2* reginv, hipStream_t st) {
#define CALL_RG(NO) hipLaunchKernelGGL(k_ci_reginv<NO>, dim3(1), dim3(64), 0, st, rho1, reg, reginv)
  switch (nO) {
    case 1: CALL_RG(1); break;
    case 2: CALL_RG(2); break;
    case 3: CALL_RG(3); break;
    case 4: CALL_RG(4); break;
    case 5: CALL_RG(5); break;
    case 6: CALL_RG(6); break;
    case 7: CALL_RG(7); break;
    case 8: CALL_RG(8); break;
    case 9: CALL_RG(9); break;
    case 10: CALL_RG(10); break;
    case 11: CALL_RG(11); break;
    case 12: CALL_RG(12); break;
    default: break;
  }
#undef CALL_RG
}

void launch_sum_parts(const double2* part, int nPart, int stride, int n, double2* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_parts, dim3((n + 255) / 256), dim3(256), 0, st, part, nPart, stride, n, out);
}

void launch_gram_reduce(const double* Gpart, int nxs, int Pc, double* G, const double2* part1, int nPart1,
                        int n2, double2* h1, hipStream_t st) {
  hipLaunchKernelGGL(k_gram_reduce, dim3((Pc * Pc + n2 + 255) / 256), dim3(256), 0, st, Gpart, nxs, Pc, G, part1,
                     nPart1, n2, h1);
}

void launch_ci_apply(const double2* C, int nC, int nO, int Pc, const double* G, const double2* h1,
                     const DevCiTables& t, double2* dCraw, hipStream_t st) {
  hipLaunchKernelGGL(k_ci_apply, dim3((nC + 3) / 4), dim3(256), 0, st, C, nC, nO, Pc, G, h1, t.s_ptr, t.singles,
                     t.d_ptr, t.doubles, dCraw);
}

void launch_dipole(int nO, const double2* part1, int nPart, const double2* rho1, double* dipole, hipStream_t st) {
  hipLaunchKernelGGL(k_dipole, dim3(1), dim3(256), 0, st, nO, part1, nPart, rho1, dipole);
}

void launch_ci_energy(int nO, int Pc, const double* G, const double2* h1, const double2* rho1,
                      const double2* rho2, double* energy, hipStream_t st) {
  hipLaunchKernelGGL(k_ci_energy, dim3(1), dim3(256), 0, st, nO, Pc, G, h1, rho1, rho2, energy);
}

}  // namespace cody
