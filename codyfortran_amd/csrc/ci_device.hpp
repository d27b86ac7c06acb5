// Device-side bodies of the CI kernels (the CI space is small, n_C <= 1024), shared by the
// kernels of ci_kernels.hip and the fused right-hand-side launches of mctdhx_kernels.hip.  A body
// takes its workgroup index as an argument; every thread of the workgroup calls it.
#pragma once

#include <hip/hip_runtime.h>

#include "fedvr_device.hpp"
#include "mctdhx_kernels.hpp"

namespace cody {

__device__ __forceinline__ void ci_rdm_body(int bx, const double2* __restrict__ C, int nC, int nO,
                                                const int* __restrict__ r1_ptr,
                                                const CiEntryDev* __restrict__ r1,
                                                const int* __restrict__ r2_ptr,
                                                const CiEntryDev* __restrict__ r2,
                                                double2* __restrict__ rho1, double2* __restrict__ r2can,
                                                double2* __restrict__ rho2) {
  const int n2 = nO * nO, np = nO * (nO - 1) / 2, nCan = np * np, ipw = ci_rdm_items_per_wave(nC);
  const int ln = threadIdx.x & 63;
  extern __shared__ double2 sC[];  // C staged once per workgroup, four loads per thread in flight
  for (int i0 = threadIdx.x; i0 < nC; i0 += 4 * blockDim.x) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * blockDim.x;
      v[u] = i < nC ? C[i] : cmk(0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * blockDim.x;
      if (i < nC) sC[i] = v[u];
    }
  }
  __syncthreads();
  for (int it = 0; it < ipw; ++it) {
    const int item = (bx * 4 + (threadIdx.x >> 6)) * ipw + it;
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
    const int q1 = ptr[code + 1];
    for (int q0 = ptr[code] + ln; q0 < q1; q0 += 4 * 64) {  // four table entries' loads in flight
      CiEntryDev en[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = q0 + 64 * u;
        en[u] = q < q1 ? tab[q] : CiEntryDev{0, 0, 0, 0.0f};
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (q0 + 64 * u < q1) cfma(acc, cscale(cconj(sC[en[u].jc]), en[u].f), sC[en[u].ic]);
    }
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (ln == 0) {
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
  }
}

// R = rho1 (rho1^2 + eps)^{-1}: the body of k_ci_reginv.  The first nth >= 2 NO threads of the
// workgroup do the work; the others only meet its barriers.
template <int NO>
__device__ __forceinline__ void reginv_body(const double2* __restrict__ rho1g, double reg,
                                            double2* __restrict__ reginv, int nth) {
  constexpr int n2 = NO * NO;
  __shared__ double2 s1[n2], sA0[n2], sM[2 * n2], sX[n2], sR[n2], sF[NO];
  const int t = threadIdx.x < nth ? threadIdx.x : 1 << 20;  // (threads past nth only meet the barriers)
  for (int i = t; i < n2; i += nth) s1[i] = rho1g[i];
  __syncthreads();
  for (int e = t; e < n2; e += nth) {
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
    for (int e = t; e < 2 * n2; e += nth) {
      const int r = e % NO, c = e / NO;
      if (r != k) sM[r + c * NO] = csub(sM[r + c * NO], cmul(sF[r], sM[k + c * NO]));
    }
    __syncthreads();
  }
  // X = A^-1 rho1, then one refinement X += A^-1 (rho1 - A X)
  const double2* sInv = sM + n2;
  for (int e = t; e < n2; e += nth) {
    const int i = e % NO, j = e / NO;
    double2 acc = cmk(0, 0);
#pragma unroll
    for (int k = 0; k < NO; ++k) cfma(acc, sInv[i + k * NO], s1[k + j * NO]);
    sX[e] = acc;
  }
  __syncthreads();
  for (int e = t; e < n2; e += nth) {
    const int i = e % NO, j = e / NO;
    double2 r = s1[e];
#pragma unroll
    for (int k = 0; k < NO; ++k) r = csub(r, cmul(sA0[i + k * NO], sX[k + j * NO]));
    sR[e] = r;  // residual
  }
  __syncthreads();
  for (int e = t; e < n2; e += nth) {
    const int i = e % NO, j = e / NO;
    double2 acc = sX[e];
#pragma unroll
    for (int k = 0; k < NO; ++k) cfma(acc, sInv[i + k * NO], sR[k + j * NO]);
    reginv[e] = acc;
  }
}

// G from the Gram kernel's split-K partials and h1 from the one-body partials: the body of
// k_gram_reduce.
__device__ __forceinline__ void gram_reduce_body(int bx, const double* __restrict__ Gpart, int nxs, int Pc,
                                                     double* __restrict__ G, const double2* __restrict__ part1,
                                                     int nPart1, int n2, double2* __restrict__ h1) {
  const int g = bx * blockDim.x + threadIdx.x;
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

// dCraw(ic) = sum_singles f h1(i,j) C(jc) + sum_doubles f h2Lin(i1,i2,j1,j2) C(jc) for one
// configuration, the lanes of one wave over its connections; the wave sum is valid in every lane.
__device__ __forceinline__ double2 ci_config_sum(int ic, int ln, const double2* __restrict__ C, int nO, int Pc,
                                                 const double* __restrict__ G, const double2* __restrict__ h1,
                                                 const int* __restrict__ s_ptr, const CiEntryDev* __restrict__ sing,
                                                 const int* __restrict__ d_ptr, const CiEntryDev* __restrict__ dbl) {
  double2 acc = cmk(0, 0);
  const int s0 = s_ptr[ic], s1 = s_ptr[ic + 1], d0 = d_ptr[ic], d1 = d_ptr[ic + 1];
  for (int q = s0 + ln; q < s1; q += 64) {
    const CiEntryDev en = sing[q];
    const int i = en.code / nO, j = en.code % nO;
    cfma(acc, cscale(h1[i + j * nO], en.f), C[en.jc]);
  }
  // the doubles in rounds of R entries per lane: a round's table entries are loaded together, then
  // its G and C gathers, so a round costs two memory round trips (each lane still adds its entries
  // in the order q = ln, ln + 64, ...)
  constexpr int R = 4;
  for (int qb = d0 + ln; qb < d1; qb += R * 64) {
    CiEntryDev en[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int q = qb + 64 * u;
      en[u] = q < d1 ? dbl[q] : CiEntryDev{0, 0, 0, 0.0f};
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
      int code = en[u].code;
      const int j2 = code % nO;
      code /= nO;
      const int i2 = code % nO;
      code /= nO;
      const int j1 = code % nO;
      const int i1 = code / nO;
      if (qb + 64 * u >= d1 || i1 == i2 || j1 == j2) continue;
      cfma(acc, cscale(h2lin_from_G(G, Pc, nO, i1, i2, j1, j2), en[u].f), C[en[u].jc]);
    }
  }
  acc.x = wave_sum(acc.x);
  acc.y = wave_sum(acc.y);
  return acc;
}

// Coefficient of G(c1, c2) in the two-body energy Re sum h2(i1,a,j1,b) rho2(j1,b,i1,a) of
// k_ci_energy (Method_GetEnergy, S_Method_Mb_OrbBased.f90:132-161), h2 from G as h2_from_G forms
// it: the pair p1 = c1/2 is (i1, j1) in either order (sign s1), p2 = c2/2 is (a, b) (sign s2), and
// row / column parity picks the real or imaginary part of G.  Zero on padding columns.
__device__ __forceinline__ double energy_g_coef(int c1, int c2, int nO, const double2* __restrict__ rho2) {
  const int P = nO * (nO + 1) / 2;
  if (c1 >= 2 * P || c2 >= 2 * P) return 0.0;
  const int r1 = c1 & 1, r2 = c2 & 1;
  int l1 = 0, q1 = c1 >> 1, l2 = 0, q2 = c2 >> 1;
  while (q1 >= nO - l1) {
    q1 -= nO - l1;
    ++l1;
  }
  while (q2 >= nO - l2) {
    q2 -= nO - l2;
    ++l2;
  }
  const int k1 = l1 + q1, k2 = l2 + q2;
  double q = 0.0;
  for (int t1 = 0; t1 < (l1 == k1 ? 1 : 2); ++t1) {
    const int i1 = t1 ? k1 : l1, j1 = t1 ? l1 : k1;
    const double s1 = t1 ? -1.0 : 1.0;
    for (int t2 = 0; t2 < (l2 == k2 ? 1 : 2); ++t2) {
      const int a = t2 ? k2 : l2, b = t2 ? l2 : k2;
      const double s2 = t2 ? -1.0 : 1.0;
      const double2 r = rho2[idx4(j1, b, i1, a, nO)];
      if (r1 == 0 && r2 == 0) q += 0.5 * r.x;
      else if (r1 == 1 && r2 == 1) q -= 0.5 * s1 * s2 * r.x;
      else if (r1 == 0) q -= 0.5 * s2 * r.y;
      else q -= 0.5 * s1 * r.y;
    }
  }
  return q;
}

// gram_reduce_body, and in the same pass this workgroup's share of the energy
// E = sum h1(k,l) rho1(l,k) + sum G(c1,c2) Q(c1,c2) (Q = energy_g_coef), summed over its threads
// in fixed order into epart[bx].  E = <C|H C>, the overlap that the CI projection removes
// (Coeffs_ProjectOnSubspace, src/Coeffs/S_Coeffs.f90:105-111).
__device__ __forceinline__ void gram_reduce_energy_body(int bx, const double* __restrict__ Gpart, int nxs, int Pc,
                                                        double* __restrict__ G, const double2* __restrict__ part1,
                                                        int nPart1, int nO, double2* __restrict__ h1,
                                                        const double2* __restrict__ rho1,
                                                        const double2* __restrict__ rho2, double* __restrict__ epart) {
  __shared__ double sRed[16];
  const int g = bx * blockDim.x + threadIdx.x;
  const int n = Pc * Pc, n2 = nO * nO;
  double e = 0.0;
  if (g < n) {
    double acc = 0.0;
#pragma unroll 8
    for (int s = 0; s < nxs; ++s) acc += Gpart[(size_t)s * n + g];
    const int c2 = g / Pc, c1 = g % Pc;
    G[c1 * Pc + c2] = acc;
    e = acc * energy_g_coef(c1, c2, nO, rho2);
  } else if (g - n < n2) {
    const int q = g - n;
    double2 acc = cmk(0, 0);
#pragma unroll 8
    for (int b = 0; b < nPart1; ++b) acc = cadd(acc, part1[(size_t)b * n2 + q]);
    h1[q] = acc;
    const int k = q % nO, l = q / nO;
    const double2 r = rho1[l + k * nO];
    e = acc.x * r.x - acc.y * r.y;
  }
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) sRed[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sRed[w];
    epart[bx] = t;
  }
}


// CI H-apply and CI projection in one launch, one wave per configuration: dCraw = H C (as
// k_ci_apply) and dC = -i (dCraw - E C) (Coeffs_ProjectOnSubspace, src/Coeffs/S_Coeffs.f90:105-111),
// with <C|dCraw> = <C|H C> = E taken from the nE energy partials of the G reduction (summed in a
// fixed order by every wave) instead of a reduction over dCraw.
__device__ __forceinline__ void ci_apply_project_body(int bx, const double2* __restrict__ C, int nC, int nO, int Pc,
                                                      const double* __restrict__ G, const double2* __restrict__ h1,
                                                      const DevCiTables& t, const double* __restrict__ epart, int nE,
                                                      double2* __restrict__ dCraw, double2* __restrict__ dC) {
  const int ic = bx * 4 + (threadIdx.x >> 6), ln = threadIdx.x & 63;
  if (ic >= nC) return;
  // the energy partials are loaded before the configuration's sums and added after them
  double e = 0.0;
  for (int b = ln; b < nE; b += 64) e += epart[b];
  const double2 acc = ci_config_sum(ic, ln, C, nO, Pc, G, h1, t.s_ptr, t.singles, t.d_ptr, t.doubles);
  e = wave_sum(e);
  if (ln == 0) {
    dCraw[ic] = acc;
    dC[ic] = mul_minus_i(csub(acc, cscale(C[ic], e)));
  }
}


}  // namespace cody
