// Adjoint of the MCTDHX right-hand side and of a Strang propagation (gfx950).
//
// The reference has no adjoint.  This file transposes, kernel by kernel, the right-hand side of
// enqueue_rhs (S_Method_Mb_OrbBased_Mctdhx.f90:101-226 and the pieces it calls) and a Strang step
// A(dt/2) B(dt) A(dt/2) (S_Propagator_SplitStep_Order2.f90:52-77) whose linear half-steps A are a
// fixed-degree Chebyshev expansion of exp(-i dt/2 h(E)).  Gradients are taken with respect to
//   <a, b> = Re( sum_C conj(a) b + sum_orbitals sum_x w(x) conj(a(x)) b(x) ),
// in which h = T + V_ext + E x is self-adjoint, so the adjoint of every one-body application is
// the same application.
//
// Vector-Jacobian product of the right-hand side f(C, phi) = -i [ (1 - C C^H) H(h1, h2) C ;
// (1 - P) draw + bh h phi ], draw = ah h phi + corr, for a cotangent lam (Lambda = i lam):
//   CI part       D = Lambda_C - C (C^H Lambda_C); gC = -conj(C^H Lambda_C) dCraw - conj(C^H dCraw) Lambda_C
//                 + H^H D; h1bar, h2bar = transition densities <C| . |D>            (k_adj_ci_start,
//                 k_ci_apply on D, k_adj_transition)
//   projection    B = <phi_i|Lambda_j>; dbar_j = Lambda_j - sum_i phi_i B_ij;
//                 g_i = -sum_j [conj(inp_ij) Lambda_j + conj(B_ij) draw_j] + sum_j conj(h1bar_ij) hphi_j;
//                 hbar_j = ah dbar_j + bh Lambda_j + sum_i h1bar_ij phi_i           (k_adj_bdots, k_adj_orb1)
//   correlation   Wbar_ab = sum conj(rho2'(k1,a,j1,b)) conj(phi_k1) dbar_j1 + 1/2 sum h2bar(k1,b,j1,a) phi_k1 conj(phi_j1)
//                 and the matching orbital terms; rho2'bar = <W_ab phi_k1 | dbar_j1> (k_adj_local, k_adj_r2pb)
//   densities     rho2bar = rho2'bar x_3 R^H, Rbar; rho1bar from R = (rho1^2 + eps)^{-1} rho1;
//                 gC += H(rho1bar + rho1bar^H, rho2bar + rho2bar^H) C             (k_adj_rdm, k_adj_ci_apply_lin)
//   mean fields   sbar = V (w Wbar) (the FP64 MFMA contraction of the forward pass, V symmetric);
//                 g_k += sbar_p phi_l, g_l += conj(sbar_p) phi_k; g += h hbar; Ebar = <hbar, x phi>   (k_adj_final)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/cody_mctdhx.h"
#include "cody_ctx.hpp"
#include "fedvr_device.hpp"
#include "mctdhx_kernels.hpp"

namespace cody {
int enqueue_rhs(cody_ctx* c, const double2* state, double2* dstate, double t, int mode, hipStream_t st,
                double* energy, RhsTrace* tr);
int run_rhs(cody_ctx* c, const double2* state, double2* dstate, double t, int mode, hipStream_t st);
int check_ready(const cody_ctx* c);
double laser_field(const cody_ctx* c, double t);
}  // namespace cody

using namespace cody;

#define ADJ_NO_SWITCH(NOV, CALL) \
  switch (NOV) {                   \
    case 1: CALL(1); break;        \
    case 2: CALL(2); break;        \
    case 3: CALL(3); break;        \
    case 4: CALL(4); break;        \
    case 5: CALL(5); break;        \
    case 6: CALL(6); break;        \
    case 7: CALL(7); break;        \
    case 8: CALL(8); break;        \
    case 9: CALL(9); break;        \
    case 10: CALL(10); break;      \
    case 11: CALL(11); break;      \
    case 12: CALL(12); break;      \
    default: break;                \
  }

namespace {

__device__ __forceinline__ int idx4(int a, int b, int c, int d, int n) { return a + n * (b + n * (c + n * d)); }
__device__ __forceinline__ int pair_of(int nO, int l, int k) { return l * nO - (l * (l - 1)) / 2 + (k - l); }
__device__ __forceinline__ double2 mul_i(double2 a) { return cmk(-a.y, a.x); }

// fixed-order sum of one double over a 256-thread workgroup; valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// W_ab(x) = V[conj(phi_b) phi_a w](x); the stored pair p = (l <= k) holds W_{k l}
template <int NO>
__device__ __forceinline__ double2 w_ab(const double* __restrict__ WT, int nGp, int a, int b, int x) {
  if (a >= b) {
    const int p = pair_of(NO, b, a);
    return cmk(WT[(size_t)(2 * p) * nGp + x], WT[(size_t)(2 * p + 1) * nGp + x]);
  }
  const int p = pair_of(NO, a, b);
  return cmk(WT[(size_t)(2 * p) * nGp + x], -WT[(size_t)(2 * p + 1) * nGp + x]);
}

// offset of S^T[c][x] in the fragment-tiled layout of the FP64 contraction (k_pair_src_t)
__device__ __forceinline__ size_t tile_off(int c, int x, int nK) {
  const int l = (c & 15) + 16 * ((x & 15) >> 2), h = (x >> 1) & 1, e = x & 1;
  return ((size_t)((c >> 4) * nK + (x >> 4)) << 8) + 128 * h + 2 * l + e;
}

// ------------------------------------------------------------------------------- CI side
// D = Lambda_C - alpha C, gC = -conj(alpha) dCraw - conj(ov) Lambda_C, alpha = C^H Lambda_C,
// ov = C^H dCraw, Lambda_C = i lam_C (one workgroup)
__global__ __launch_bounds__(256) void k_adj_ci_start(const double2* __restrict__ C, const double2* __restrict__ lam,
                                                      const double2* __restrict__ dCraw, int nC,
                                                      double2* __restrict__ D, double2* __restrict__ gC) {
  __shared__ double red[4];
  double2 al = cmk(0, 0), ov = cmk(0, 0);
  for (int i = threadIdx.x; i < nC; i += 256) {
    cfmac(al, C[i], mul_i(lam[i]));
    cfmac(ov, C[i], dCraw[i]);
  }
  double v[4] = {al.x, al.y, ov.x, ov.y};
  for (int q = 0; q < 4; ++q) {
    v[q] = block_sum(v[q], red);
    __syncthreads();
  }
  al = cmk(v[0], v[1]);
  ov = cmk(v[2], v[3]);
  for (int i = threadIdx.x; i < nC; i += 256) {
    const double2 L = mul_i(lam[i]);
    D[i] = csub(L, cmul(al, C[i]));
    gC[i] = csub(cmk(0, 0), cadd(cmul(cconj(al), dCraw[i]), cmul(cconj(ov), L)));
  }
}

// transition densities lin[code] = sum f conj(bra_jc) ket_ic through the density tables (the
// layout of k_ci_rdm): h1b(i,j) at i + j nO, h2b unpacked to its four antisymmetric positions
__global__ __launch_bounds__(256) void k_adj_transition(const double2* __restrict__ bra, const double2* __restrict__ ket,
                                                        int nC, int nO, const int* __restrict__ r1_ptr,
                                                        const CiEntryDev* __restrict__ r1,
                                                        const int* __restrict__ r2_ptr,
                                                        const CiEntryDev* __restrict__ r2, double2* __restrict__ h1b,
                                                        double2* __restrict__ h2b) {
  const int n2 = nO * nO, np = nO * (nO - 1) / 2, nCan = np * np;
  const int item = blockIdx.x * 4 + (threadIdx.x >> 6), ln = threadIdx.x & 63;
  extern __shared__ double2 sv[];  // bra then ket
  for (int i = threadIdx.x; i < nC; i += blockDim.x) {
    sv[i] = bra[i];
    sv[nC + i] = ket[i];
  }
  __syncthreads();
  if (item >= n2 + nCan) return;
  int code, i1 = 0, i2 = 0, j1 = 0, j2 = 0;
  const int* ptr;
  const CiEntryDev* tab;
  auto strict = [&](int p, int& a, int& b) {
    a = 0;
    while (p >= nO - 1 - a) {
      p -= nO - 1 - a;
      ++a;
    }
    b = a + 1 + p;
  };
  if (item < n2) {
    code = (item % nO) * nO + item / nO;
    ptr = r1_ptr;
    tab = r1;
  } else {
    const int e = item - n2;
    strict(e % np, i1, i2);
    strict(e / np, j1, j2);
    code = ((i1 * nO + j1) * nO + i2) * nO + j2;
    ptr = r2_ptr;
    tab = r2;
  }
  double2 acc = cmk(0, 0);
  for (int q = ptr[code] + ln; q < ptr[code + 1]; q += 64) {
    const CiEntryDev en = tab[q];
    cfma(acc, cscale(cconj(sv[en.jc]), en.f), sv[nC + en.ic]);
  }
  acc.x = wave_sum(acc.x);
  acc.y = wave_sum(acc.y);
  if (ln != 0) return;
  if (item < n2) {
    h1b[item] = acc;
  } else {
    const double2 m = cmk(-acc.x, -acc.y);
    h2b[idx4(i1, i2, j1, j2, nO)] = acc;
    h2b[idx4(i2, i1, j1, j2, nO)] = m;
    h2b[idx4(i1, i2, j2, j1, nO)] = m;
    h2b[idx4(i2, i1, j2, j1, nO)] = acc;
  }
}

// gC[ic] += sum_singles f lin1[code] C[jc] + sum_doubles f lin2[code] C[jc] (the H-apply of
// k_ci_apply with explicit packed coefficients; lin2 may be null), one wave per configuration
__global__ __launch_bounds__(256) void k_adj_ci_apply_lin(const double2* __restrict__ C, int nC, int nO,
                                                          const int* __restrict__ s_ptr,
                                                          const CiEntryDev* __restrict__ sing,
                                                          const int* __restrict__ d_ptr,
                                                          const CiEntryDev* __restrict__ dbl,
                                                          const double2* __restrict__ lin1,
                                                          const double2* __restrict__ lin2, double2* __restrict__ gC) {
  const int ic = blockIdx.x * 4 + (threadIdx.x >> 6), ln = threadIdx.x & 63;
  if (ic >= nC) return;
  double2 acc = cmk(0, 0);
  for (int q = s_ptr[ic] + ln; q < s_ptr[ic + 1]; q += 64) {
    const CiEntryDev en = sing[q];
    cfma(acc, cscale(lin1[en.code], en.f), C[en.jc]);
  }
  if (lin2) {
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
      cfma(acc, cscale(lin2[en.code], en.f), C[en.jc]);
    }
  }
  acc.x = wave_sum(acc.x);
  acc.y = wave_sum(acc.y);
  if (ln == 0) gC[ic] = cadd(gC[ic], acc);
}

// ------------------------------------------------------------------------------- orbital side
// partial sums of B(i,j) = <phi_i|Lambda_j>_w over 256-point blocks (Lambda = i lam)
template <int NO>
__global__ __launch_bounds__(256) void k_adj_bdots(const double2* __restrict__ orbs, const double* __restrict__ w,
                                                   const double2* __restrict__ lamO, int nG,
                                                   double2* __restrict__ part) {
  __shared__ double2 sRed[4][NO];
  const int j = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = blockIdx.x * 256 + threadIdx.x;
  const bool in = x < nG;
  const int xc = in ? x : nG - 1;
  const double2 lw = in ? cscale(mul_i(lamO[(size_t)j * nG + xc]), w[xc]) : cmk(0, 0);
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    double2 acc = cmk(0, 0);
    if (in) cfmac(acc, orbs[(size_t)i * nG + xc], lw);
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (lane == 0) sRed[wave][i] = acc;
  }
  __syncthreads();
  if (threadIdx.x < NO) {
    const int i = threadIdx.x;
    part[(size_t)blockIdx.x * NO * NO + i + j * NO] =
        cadd(cadd(sRed[0][i], sRed[1][i]), cadd(sRed[2][i], sRed[3][i]));
  }
}

// projection and h1 terms, one grid point per thread
template <int NO>
__global__ __launch_bounds__(256) void k_adj_orb1(const double2* __restrict__ orbs, const double2* __restrict__ lamO,
                                                  const double2* __restrict__ draw, const double2* __restrict__ hphi,
                                                  const double2* __restrict__ Bm, const double2* __restrict__ inp,
                                                  const double2* __restrict__ h1b, int nG, double ah, double bh,
                                                  double2* __restrict__ dbar, double2* __restrict__ gO,
                                                  double2* __restrict__ hbar) {
  constexpr int n2 = NO * NO;
  __shared__ double2 sB[n2], sI[n2], sH[n2];
  for (int e = threadIdx.x; e < n2; e += 256) {
    sB[e] = Bm[e];
    sI[e] = inp[e];
    sH[e] = h1b[e];
  }
  __syncthreads();
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= nG) return;
  double2 phi[NO], L[NO], dr[NO], hp[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    phi[k] = orbs[(size_t)k * nG + x];
    L[k] = mul_i(lamO[(size_t)k * nG + x]);
    dr[k] = draw[(size_t)k * nG + x];
    hp[k] = hphi[(size_t)k * nG + x];
  }
#pragma unroll
  for (int j = 0; j < NO; ++j) {
    double2 db = L[j], hb = cscale(L[j], bh);
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      db = csub(db, cmul(phi[i], sB[i + j * NO]));
      cfma(hb, sH[i + j * NO], phi[i]);
    }
    dbar[(size_t)j * nG + x] = db;
    hb.x = fma(ah, db.x, hb.x);
    hb.y = fma(ah, db.y, hb.y);
    hbar[(size_t)j * nG + x] = hb;
  }
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    double2 g = cmk(0, 0);
#pragma unroll
    for (int j = 0; j < NO; ++j) {
      g = csub(g, cadd(cmul(cconj(sI[i + j * NO]), L[j]), cmul(cconj(sB[i + j * NO]), dr[j])));
      cfma(g, cconj(sH[i + j * NO]), hp[j]);
    }
    gO[(size_t)i * nG + x] = g;
  }
}

// rho2'(i1,i2,j1,l2) = sum_k rho2(i1,i2,k,l2) R(k,j1)
template <int NO>
__global__ __launch_bounds__(256) void k_adj_rho2p(const double2* __restrict__ rho2, const double2* __restrict__ R,
                                                   double2* __restrict__ rho2p) {
  constexpr int n2 = NO * NO, n4 = n2 * n2;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n4) return;
  const int i12 = e % n2, j1 = (e / n2) % NO, l2 = e / (n2 * NO);
  double2 acc = cmk(0, 0);
#pragma unroll
  for (int k = 0; k < NO; ++k) cfma(acc, rho2[i12 + n2 * (k + NO * l2)], R[k + j1 * NO]);
  rho2p[e] = acc;
}

// Wbar of the stored pairs (folded: Wbar_p = Wbar_kl + conj(Wbar_lk)), written times w into the
// tiled source of the backward contraction, and the orbital gradient terms of the correlation
// apply and of h2.  Workgroup = 64 grid points (one per lane) x 4 pairs (one per wave,
// p = 4 blockIdx.y + wave); the four waves' orbital terms are added in LDS in fixed order and
// written to gpart[blockIdx.y] (summed by k_adj_final).  Points past nG write zero sources.
template <int NO>
__global__ __launch_bounds__(256) void k_adj_local(const double2* __restrict__ orbs, const double* __restrict__ w,
                                                   const double* __restrict__ WT, int nG, int nGp,
                                                   const double2* __restrict__ rho2p, const double2* __restrict__ h2b,
                                                   const double2* __restrict__ dbar, double2* __restrict__ gpart,
                                                   double* __restrict__ STadj) {
  constexpr int NP = NO * (NO + 1) / 2;
  __shared__ double2 sG[4][NO][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p = __builtin_amdgcn_readfirstlane(4 * (int)blockIdx.y + wave);
  const int x = blockIdx.x * 64 + lane;
  const int nK = nGp >> 4;
  const bool in = x < nG;
  const int xc = in ? x : nG - 1;
  double2 g[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) g[k] = cmk(0, 0);
  if (p < NP) {
    int l = 0, r = p;
    while (r >= NO - l) {
      r -= NO - l;
      ++l;
    }
    const int k = l + r;
    double2 phi[NO], db[NO];
#pragma unroll
    for (int q = 0; q < NO; ++q) {
      phi[q] = orbs[(size_t)q * nG + xc];
      db[q] = dbar[(size_t)q * nG + xc];
    }
    const double2 Wp = cmk(WT[(size_t)(2 * p) * nGp + x], WT[(size_t)(2 * p + 1) * nGp + x]);  // W_{k l}
    double2 fold = cmk(0, 0);
    for (int s = 0; s < (k == l ? 1 : 2); ++s) {
      const int a = s ? l : k, b = s ? k : l;
      const double2 Wab = s ? cconj(Wp) : Wp;
      const double2 cW = cconj(Wab), hW = cscale(Wab, 0.5);
      // with rr = conj(rho2'(k1,a,j1,b)) and hb = h2bar(k1,b,j1,a):
      //   X     += sum_k1 conj(phi_k1) T1_k1 + 1/2 phi_k1 T2_k1,
      //   g_k1  += cW T1_k1 + hW conj(T2_k1),   g_j1 += conj(hW) sum_k1 hb phi_k1,
      // T1_k1 = sum_j1 rr dbar_j1, T2_k1 = sum_j1 hb conj(phi_j1)
      double2 X = cmk(0, 0);
      double2 U[NO];
#pragma unroll
      for (int j1 = 0; j1 < NO; ++j1) U[j1] = cmk(0, 0);
#pragma unroll
      for (int k1 = 0; k1 < NO; ++k1) {
        double2 T1 = cmk(0, 0), T2 = cmk(0, 0);
#pragma unroll
        for (int j1 = 0; j1 < NO; ++j1) {
          const double2 rr = cconj(rho2p[idx4(k1, a, j1, b, NO)]);
          const double2 hb = h2b[idx4(k1, b, j1, a, NO)];
          cfma(T1, rr, db[j1]);
          cfma(T2, hb, cconj(phi[j1]));
          cfma(U[j1], hb, phi[k1]);
        }
        cfma(X, cconj(phi[k1]), T1);
        cfma(X, cscale(phi[k1], 0.5), T2);
        cfma(g[k1], cW, T1);
        cfma(g[k1], hW, cconj(T2));
      }
      const double2 chW = cconj(hW);
#pragma unroll
      for (int j1 = 0; j1 < NO; ++j1) cfma(g[j1], chW, U[j1]);
      fold = cadd(fold, s ? cconj(X) : X);
    }
    const double wx = in ? w[xc] : 0.0;
    STadj[tile_off(2 * p, x, nK)] = wx * fold.x;
    STadj[tile_off(2 * p + 1, x, nK)] = wx * fold.y;
  }
#pragma unroll
  for (int q = 0; q < NO; ++q) sG[wave][q][lane] = g[q];
  __syncthreads();
  for (int e = threadIdx.x; e < NO * 64; e += 256) {
    const int q = e >> 6, ln = e & 63, xo = blockIdx.x * 64 + ln;
    if (xo >= nG) continue;
    gpart[((size_t)blockIdx.y * NO + q) * nG + xo] =
        cadd(cadd(sG[0][q][ln], sG[1][q][ln]), cadd(sG[2][q][ln], sG[3][q][ln]));
  }
}

// rho2'bar(k1,a,j1,b) = sum_x w conj(W_ab phi_k1) dbar_j1, one workgroup per (k1, a, b)
template <int NO>
__global__ __launch_bounds__(256) void k_adj_r2pb(const double2* __restrict__ orbs, const double* __restrict__ w,
                                                  const double* __restrict__ WT, int nG, int nGp,
                                                  const double2* __restrict__ dbar, double2* __restrict__ r2pb) {
  __shared__ double2 sRed[4][NO];
  const int k1 = blockIdx.x % NO, a = (blockIdx.x / NO) % NO, b = blockIdx.x / (NO * NO);
  double2 acc[NO];
#pragma unroll
  for (int j = 0; j < NO; ++j) acc[j] = cmk(0, 0);
#pragma unroll 4
  for (int x = threadIdx.x; x < nG; x += 256) {  // four points' loads in flight, added in order
    const double2 v = cscale(cconj(cmul(w_ab<NO>(WT, nGp, a, b, x), orbs[(size_t)k1 * nG + x])), w[x]);
#pragma unroll
    for (int j = 0; j < NO; ++j) cfma(acc[j], v, dbar[(size_t)j * nG + x]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NO; ++j) {
    const double2 s = cmk(wave_sum(acc[j].x), wave_sum(acc[j].y));
    if (lane == 0) sRed[wave][j] = s;
  }
  __syncthreads();
  if (threadIdx.x < NO) {
    const int j = threadIdx.x;
    r2pb[idx4(k1, a, j, b, NO)] = cadd(cadd(sRed[0][j], sRed[1][j]), cadd(sRed[2][j], sRed[3][j]));
  }
}

// rho2bar = rho2'bar x_3 R^H, Rbar, rho1bar = Q - Q R^H rho1 - rho1 Q R^H with
// Q = (rho1^2 + eps)^{-1} Rbar (Gauss-Jordan, A Hermitian positive definite), and the packed
// coefficients lin + lin^H of the density gradient (one workgroup)
template <int NO>
__global__ __launch_bounds__(256) void k_adj_rdm(const double2* __restrict__ rho1, const double2* __restrict__ R,
                                                 const double2* __restrict__ rho2, const double2* __restrict__ r2pb,
                                                 double reg, double2* __restrict__ r2b, double2* __restrict__ lin1,
                                                 double2* __restrict__ lin2) {
  constexpr int n2 = NO * NO, n3 = n2 * NO, n4 = n2 * n2;
  __shared__ double2 s1[n2], sR[n2], sM[2 * n2], sT[n2], sQ[n2], sF[NO], sRb[256];
  static_assert(n2 <= 256, "one Rbar output per thread group member");
  const int t = threadIdx.x;
  for (int e = t; e < n2; e += 256) {
    s1[e] = rho1[e];
    sR[e] = R[e];
  }
  __syncthreads();
  for (int e = t; e < n4; e += 256) {
    const int i12 = e % n2, k = (e / n2) % NO, l2 = e / n3;
    double2 acc = cmk(0, 0);
#pragma unroll
    for (int j1 = 0; j1 < NO; ++j1) cfma(acc, r2pb[i12 + n2 * (j1 + NO * l2)], cconj(sR[k + j1 * NO]));
    r2b[e] = acc;
  }
  // Rbar(k, j1) = sum_{l2, i12} conj(rho2(i12, k, l2)) r2pb(i12, j1, l2): output o = k + j1 NO
  // by G thread groups, group g taking l2 = g, g + G, ...; the group sums are added in order
  {
    constexpr int G = n2 <= 64 ? 4 : (n2 <= 128 ? 2 : 1), S = 256 / G;
    const int o = t % S, g = t / S;
    if (o < n2) {
      const int k = o % NO, j1 = o / NO;
      double2 acc = cmk(0, 0);
      for (int l2 = g; l2 < NO; l2 += G) {
        const double2* ra = rho2 + n2 * (k + NO * l2);
        const double2* rb = r2pb + n2 * (j1 + NO * l2);
#pragma unroll 8
        for (int i12 = 0; i12 < n2; ++i12) cfmac(acc, ra[i12], rb[i12]);
      }
      sRb[g * S + o] = acc;
    }
  }
  for (int e = t; e < n2; e += 256) {
    // [A | Rbar], A = rho1^2 + eps
    const int i = e % NO, j = e / NO;
    double2 a = cmk(i == j ? reg : 0.0, 0.0);
#pragma unroll
    for (int q = 0; q < NO; ++q) cfma(a, s1[i + q * NO], s1[q + j * NO]);
    sM[e] = a;
  }
  __syncthreads();
  {
    constexpr int G = n2 <= 64 ? 4 : (n2 <= 128 ? 2 : 1), S = 256 / G;
    for (int e = t; e < n2; e += 256) {
      double2 acc = sRb[e];
      for (int g = 1; g < G; ++g) acc = cadd(acc, sRb[g * S + e]);
      sM[n2 + e] = acc;  // Rbar(k, j1) at k + j1 NO
    }
  }
  __syncthreads();
  for (int k = 0; k < NO; ++k) {
    const double2 piv = sM[k + k * NO];
    __syncthreads();
    if (t < 2 * NO) {
      const double d = piv.x * piv.x + piv.y * piv.y;
      const double2 v = sM[k + t * NO];
      sM[k + t * NO] = cmk((v.x * piv.x + v.y * piv.y) / d, (v.y * piv.x - v.x * piv.y) / d);
    }
    if (t < NO) sF[t] = sM[t + k * NO];
    __syncthreads();
    for (int e = t; e < 2 * n2; e += 256) {
      const int r = e % NO, c = e / NO;
      if (r != k) sM[r + c * NO] = csub(sM[r + c * NO], cmul(sF[r], sM[k + c * NO]));
    }
    __syncthreads();
  }
  for (int e = t; e < n2; e += 256) sQ[e] = sM[n2 + e];
  __syncthreads();
  // T = Q R^H
  for (int e = t; e < n2; e += 256) {
    const int i = e % NO, j = e / NO;
    double2 acc = cmk(0, 0);
#pragma unroll
    for (int q = 0; q < NO; ++q) cfma(acc, sQ[i + q * NO], cconj(sR[j + q * NO]));
    sT[e] = acc;
  }
  __syncthreads();
  // rho1bar = Q - T rho1 - rho1 T, kept in sM[0, n2)
  for (int e = t; e < n2; e += 256) {
    const int i = e % NO, j = e / NO;
    double2 acc = sQ[e];
#pragma unroll
    for (int q = 0; q < NO; ++q) {
      acc = csub(acc, cmul(sT[i + q * NO], s1[q + j * NO]));
      acc = csub(acc, cmul(s1[i + q * NO], sT[q + j * NO]));
    }
    sM[e] = acc;
  }
  __syncthreads();
  // lin1[i nO + j] = rho1bar(i,j) + conj(rho1bar(j,i))
  for (int e = t; e < n2; e += 256) {
    const int i = e / NO, j = e % NO;
    lin1[e] = cadd(sM[i + j * NO], cconj(sM[j + i * NO]));
  }
  // lin2 at the canonical codes (i1 < i2, j1 < j2): pack(rho2bar) + conj(pack(rho2bar) at the
  // transposed code)
  constexpr int np = NO * (NO - 1) / 2;
  for (int e = t; e < np * np; e += 256) {
    int i1 = 0, i2 = 0, j1 = 0, j2 = 0;
    int p = e % np;
    while (p >= NO - 1 - i1) {
      p -= NO - 1 - i1;
      ++i1;
    }
    i2 = i1 + 1 + p;
    p = e / np;
    while (p >= NO - 1 - j1) {
      p -= NO - 1 - j1;
      ++j1;
    }
    j2 = j1 + 1 + p;
    const double2 v = cadd(csub(csub(r2b[idx4(i1, i2, j1, j2, NO)], r2b[idx4(i1, i2, j2, j1, NO)]),
                                r2b[idx4(i2, i1, j1, j2, NO)]),
                           r2b[idx4(i2, i1, j2, j1, NO)]);
    const double2 vt = cadd(csub(csub(r2b[idx4(j1, j2, i1, i2, NO)], r2b[idx4(j1, j2, i2, i1, NO)]),
                                 r2b[idx4(j2, j1, i1, i2, NO)]),
                            r2b[idx4(j2, j1, i2, i1, NO)]);
    lin2[((i1 * NO + j1) * NO + i2) * NO + j2] = cadd(v, cconj(vt));
  }
}

// mean-field and one-body terms, one grid point per thread: g += the pair groups' terms of
// k_adj_local (gpart, ngroups of them), g_k += sbar_p phi_l and
// g_l += conj(sbar_p) phi_k for the stored pairs (WTadj = V (w Wbar), may be null), g_j += h hbar_j,
// and per-workgroup partials of Ebar = sum_j <hbar_j, x phi_j>_w
// One thread per grid point and orbital j = blockIdx.y: the pair terms that land on g_j are taken in
// the stored pair order, so every g_j receives its terms in a fixed order; the Ebar partials are
// per (orbital, workgroup), epart[blockIdx.y gridDim.x + blockIdx.x].
template <int NO>
__global__ __launch_bounds__(256) void k_adj_final(const double2* __restrict__ orbs, const double* __restrict__ w,
                                                   const double* __restrict__ xg, const double* __restrict__ vext,
                                                   KinParams kp, double E, const double* __restrict__ WTadj, int nG,
                                                   int nGp, const double2* __restrict__ gpart, int ngroups,
                                                   const double2* __restrict__ hbar, double2* __restrict__ gO,
                                                   double* __restrict__ epart) {
  __shared__ double red[4];
  const int x = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  double e = 0.0;
  if (x < nG) {
    double2 g = gO[(size_t)j * nG + x];
#pragma unroll 4
    for (int q = 0; q < ngroups; ++q) g = cadd(g, gpart[((size_t)q * NO + j) * nG + x]);
    if (WTadj) {
      double2 phi[NO];
#pragma unroll
      for (int k = 0; k < NO; ++k) phi[k] = orbs[(size_t)k * nG + x];
      int p = 0;
#pragma unroll
      for (int l = 0; l < NO; ++l)
#pragma unroll
        for (int k = l; k < NO; ++k, ++p) {
          if (k != j && l != j) continue;  // uniform over the workgroup
          const double2 sb = cmk(WTadj[(size_t)(2 * p) * nGp + x], WTadj[(size_t)(2 * p + 1) * nGp + x]);
          if (k == j) cfma(g, sb, phi[l]);
          if (l == j) cfma(g, cconj(sb), phi[k]);
        }
    }
    const double wx = w[x], v = vext[x] + E * xg[x];
    const double2 hb = hbar[(size_t)j * nG + x];
    double2 hh = kinetic_at(hbar + (size_t)j * nG, x, kp, 1.0 / wx);
    hh.x = fma(v, hb.x, hh.x);
    hh.y = fma(v, hb.y, hh.y);
    gO[(size_t)j * nG + x] = cadd(g, hh);
    const double2 pj = orbs[(size_t)j * nG + x];
    e = (hb.x * pj.x + hb.y * pj.y) * (wx * xg[x]);
  }
  const double s = block_sum(e, red);
  if (threadIdx.x == 0) epart[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// out = sum of v[0..n) in a fixed order (one workgroup)
__global__ __launch_bounds__(256) void k_adj_sum(const double* __restrict__ v, int n, double* __restrict__ out) {
  __shared__ double red[4];
  double acc = 0.0;
#pragma unroll 16
  for (int i = threadIdx.x; i < n; i += 256) acc += v[i];  // 16 loads in flight, added in order
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_adj_hbar_lin(const double2* __restrict__ lamO, int n, double2* __restrict__ hbar) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) hbar[i] = mul_i(lamO[i]);
}

// gradient of <x> = sum_kl rho1(l,k) <phi_k|x phi_l>: g_k = 2 x sum_l rho1(l,k) phi_l
template <int NO>
__global__ __launch_bounds__(256) void k_adj_dipole_orb(const double2* __restrict__ orbs, const double* __restrict__ xg,
                                                        const double2* __restrict__ rho1, int nG,
                                                        double2* __restrict__ gO) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= nG) return;
  double2 phi[NO];
#pragma unroll
  for (int l = 0; l < NO; ++l) phi[l] = orbs[(size_t)l * nG + x];
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    double2 acc = cmk(0, 0);
#pragma unroll
    for (int l = 0; l < NO; ++l) cfma(acc, rho1[l + k * NO], phi[l]);
    gO[(size_t)k * nG + x] = cscale(acc, 2.0 * xg[x]);
  }
}

// lin1[i nO + j] = X(i,j) + conj(X(j,i)) for X at i + j nO (packed rho1bar + its transpose)
__global__ void k_adj_lin1_from(const double2* __restrict__ X, int nO, double2* __restrict__ lin1) {
  const int e = threadIdx.x;
  if (e >= nO * nO) return;
  const int i = e / nO, j = e % nO;
  lin1[e] = cadd(X[i + j * nO], cconj(X[j + i * nO]));
}

// ------------------------------------------------------------------------------- Chebyshev
// One term of the expansion, all orbitals: u_k = f ((h - c)/r) u_{k-1} - u_{k-2} (f = 1 and no
// u_{k-2} for k = 1), s = (k == 1 ? a_0 u_0 : acc) + a_k u_k, written to `out` (the last term)
// or to acc.
__global__ __launch_bounds__(256) void k_cheb_term(const double2* __restrict__ uprev, const double2* __restrict__ ucur,
                                                   double2* __restrict__ unext, double2* __restrict__ acc,
                                                   double2* __restrict__ out, KinParams kp,
                                                   const double* __restrict__ w, const double* __restrict__ xg,
                                                   const double* __restrict__ vext, double E, double cc, double rinv,
                                                   double2 a0, double2 ak, int first, int nG) {
  const int j = blockIdx.y, x = blockIdx.x * 256 + threadIdx.x;
  if (x >= nG) return;
  const size_t off = (size_t)j * nG + x;
  const double2 u = ucur[off];
  double2 hu = kinetic_at(ucur + (size_t)j * nG, x, kp, 1.0 / w[x]);
  const double v = vext[x] + E * xg[x] - cc;
  hu.x = fma(v, u.x, hu.x);
  hu.y = fma(v, u.y, hu.y);
  double2 un;
  if (first) {
    un = cscale(hu, rinv);
  } else {
    const double2 up = uprev[off];
    un = cmk(2.0 * rinv * hu.x - up.x, 2.0 * rinv * hu.y - up.y);
  }
  unext[off] = un;
  double2 s = first ? cmul(a0, u) : acc[off];
  cfma(s, ak, un);
  if (out)
    out[off] = s;
  else
    acc[off] = s;
}

// Two terms per launch: u_k on the workgroup's 256-point tile and a halo of one element
// (nLoc - 1 points) on either side, kept in LDS, then u_{k+1} = 2 ((h - c)/r) u_k - u_{k-1} on the
// tile from it; s += a_k u_k + a_{k+1} u_{k+1} (s = a_0 u_0 first when k = 1), written to `out`
// (the pair ends the series) or to acc.  Every value is formed by the operations, in the order,
// of two k_cheb_term launches.
constexpr int kChebHaloMax = 32;
__global__ __launch_bounds__(256) void k_cheb_term2(const double2* __restrict__ uprev, const double2* __restrict__ ucur,
                                                    double2* __restrict__ uk_out, double2* __restrict__ uk1_out,
                                                    double2* __restrict__ acc, double2* __restrict__ out, KinParams kp,
                                                    const double* __restrict__ w, const double* __restrict__ xg,
                                                    const double* __restrict__ vext, double E, double cc, double rinv,
                                                    double2 a0, double2 ak, double2 ak1, int first, int halo, int nG) {
  __shared__ double2 sU[256 + 2 * kChebHaloMax];
  const int j = blockIdx.y, x0 = blockIdx.x * 256, xlo = x0 - halo;
  const double2* uc = ucur + (size_t)j * nG;
  for (int t = threadIdx.x; t < 256 + 2 * halo; t += 256) {
    const int x = xlo + t;
    double2 un = cmk(0, 0);
    if (x >= 0 && x < nG) {
      const double2 u = uc[x];
      double2 hu = kinetic_at(uc, x, kp, 1.0 / w[x]);
      const double v = vext[x] + E * xg[x] - cc;
      hu.x = fma(v, u.x, hu.x);
      hu.y = fma(v, u.y, hu.y);
      if (first) {
        un = cscale(hu, rinv);
      } else {
        const double2 up = uprev[(size_t)j * nG + x];
        un = cmk(2.0 * rinv * hu.x - up.x, 2.0 * rinv * hu.y - up.y);
      }
    }
    sU[t] = un;
  }
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x >= nG) return;
  const size_t off = (size_t)j * nG + x;
  const double2 uk = sU[threadIdx.x + halo];
  uk_out[off] = uk;
  double2 s = first ? cmul(a0, uc[x]) : acc[off];
  cfma(s, ak, uk);
  double2 hu = kinetic_at_lds(sU, xlo, x, kp, 1.0 / w[x]);
  const double v = vext[x] + E * xg[x] - cc;
  hu.x = fma(v, uk.x, hu.x);
  hu.y = fma(v, uk.y, hu.y);
  const double2 up = uc[x];
  const double2 un = cmk(2.0 * rinv * hu.x - up.x, 2.0 * rinv * hu.y - up.y);
  uk1_out[off] = un;
  cfma(s, ak1, un);
  if (out)
    out[off] = s;
  else
    acc[off] = s;
}

// One Clenshaw step of the adjoint: b_k = conj(a_k) lam + f ((h - c)/r) b_{k+1} - b_{k+2}
// (terms absent when the pointers are null) and eacc += f <b_{k+1}, (x/r) u_k> per point
__global__ __launch_bounds__(256) void k_cheb_adj_term(const double2* __restrict__ lam,
                                                       const double2* __restrict__ bnext,
                                                       const double2* __restrict__ bnext2, double2* __restrict__ bout,
                                                       const double2* __restrict__ uk, double* __restrict__ eacc,
                                                       KinParams kp, const double* __restrict__ w,
                                                       const double* __restrict__ xg, const double* __restrict__ vext,
                                                       double E, double cc, double rinv, double2 cak, double f,
                                                       int nG) {
  const int j = blockIdx.y, x = blockIdx.x * 256 + threadIdx.x;
  if (x >= nG) return;
  const size_t off = (size_t)j * nG + x;
  double2 b = cmul(cak, lam[off]);
  if (bnext) {
    const double2 bn = bnext[off];
    double2 hb = kinetic_at(bnext + (size_t)j * nG, x, kp, 1.0 / w[x]);
    const double v = vext[x] + E * xg[x] - cc;
    hb.x = fma(v, bn.x, hb.x);
    hb.y = fma(v, bn.y, hb.y);
    b.x = fma(f * rinv, hb.x, b.x);
    b.y = fma(f * rinv, hb.y, b.y);
    const double2 u = uk[off];
    eacc[off] += f * rinv * w[x] * xg[x] * (bn.x * u.x + bn.y * u.y);
  }
  if (bnext2) b = csub(b, bnext2[off]);
  bout[off] = b;
}

// Two Clenshaw steps per launch, k and k - 1 (k >= 2, so both have b_{k+1} and f = 2): b_k on the
// tile and a one-element halo into LDS, then b_{k-1} on the tile from it; the eacc updates of both
// steps in order.  Reads b_{k+1}, b_{k+2} and writes b_k, b_{k-1}: four distinct buffers, so no
// workgroup overwrites what another one's halo still reads.  Same operations, in the same order,
// as two k_cheb_adj_term launches.
__global__ __launch_bounds__(256) void k_cheb_adj_term2(const double2* __restrict__ lam,
                                                        const double2* __restrict__ bnext,
                                                        const double2* __restrict__ bnext2, double2* __restrict__ bk_out,
                                                        double2* __restrict__ bk1_out, const double2* __restrict__ uk,
                                                        const double2* __restrict__ uk1, double* __restrict__ eacc,
                                                        KinParams kp, const double* __restrict__ w,
                                                        const double* __restrict__ xg, const double* __restrict__ vext,
                                                        double E, double cc, double rinv, double2 cak, double2 cak1,
                                                        int halo, int nG) {
  __shared__ double2 sB[256 + 2 * kChebHaloMax];
  const int j = blockIdx.y, x0 = blockIdx.x * 256, xlo = x0 - halo;
  const double2* bn_o = bnext + (size_t)j * nG;
  const double f = 2.0;
  for (int t = threadIdx.x; t < 256 + 2 * halo; t += 256) {
    const int x = xlo + t;
    double2 b = cmk(0, 0);
    if (x >= 0 && x < nG) {
      const size_t off = (size_t)j * nG + x;
      b = cmul(cak, lam[off]);
      const double2 bn = bn_o[x];
      double2 hb = kinetic_at(bn_o, x, kp, 1.0 / w[x]);
      const double v = vext[x] + E * xg[x] - cc;
      hb.x = fma(v, bn.x, hb.x);
      hb.y = fma(v, bn.y, hb.y);
      b.x = fma(f * rinv, hb.x, b.x);
      b.y = fma(f * rinv, hb.y, b.y);
      if (bnext2) b = csub(b, bnext2[off]);
    }
    sB[t] = b;
  }
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x >= nG) return;
  const size_t off = (size_t)j * nG + x;
  const double2 bn = bn_o[x];
  {
    const double2 u = uk[off];
    eacc[off] += f * rinv * w[x] * xg[x] * (bn.x * u.x + bn.y * u.y);
  }
  const double2 bk = sB[threadIdx.x + halo];
  bk_out[off] = bk;
  double2 b = cmul(cak1, lam[off]);
  double2 hb = kinetic_at_lds(sB, xlo, x, kp, 1.0 / w[x]);
  const double v = vext[x] + E * xg[x] - cc;
  hb.x = fma(v, bk.x, hb.x);
  hb.y = fma(v, bk.y, hb.y);
  b.x = fma(f * rinv, hb.x, b.x);
  b.y = fma(f * rinv, hb.y, b.y);
  const double2 u = uk1[off];
  eacc[off] += f * rinv * w[x] * xg[x] * (bk.x * u.x + bk.y * u.y);
  b = csub(b, bn);
  bk1_out[off] = b;
}

__global__ void k_zero_d(double* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Gershgorin interval of h = T + V + E x (|E| <= emax) at one row of its symmetric form
// W^{1/2} h W^{-1/2}: T_xj = S_xj / w_x with S symmetric, so the row has S_xx / w_x on the
// diagonal and |S_xj| / sqrt(w_x w_j) off it (the loops of kinetic_at)
__global__ void k_gersh(KinParams kp, const double* __restrict__ w, const double* __restrict__ xg,
                        const double* __restrict__ vext, int nG, double emax, double* __restrict__ lo,
                        double* __restrict__ hi) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nG) return;
  const int gi = x + kp.xminExcl, per = kp.nLoc - 1;
  const int e = gi / per, a = gi % per;
  double diag = 0.0, rad = 0.0;
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
      if (j < 0 || j >= nG) continue;
      const double s = kp.kref[aa + b * kp.nLoc] * kp.invh * kp.fac;
      if (j == x)
        diag += s;
      else
        rad += fabs(s) / sqrt(w[x] * w[j]);
    }
  }
  const double c = diag / w[x] + vext[x], ex = emax * fabs(xg[x]);
  lo[x] = c - rad - ex;
  hi[x] = c + rad + ex;
}

// ------------------------------------------------------------------------------- RK4 pieces
// inc = (first ? 0 : inc) + ai d; ynext = y1 + an d
__global__ void k_rk_stage(const double2* __restrict__ y1, const double2* __restrict__ d, double2* __restrict__ inc,
                           double2* __restrict__ ynext, double ai, double an, int first, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 dv = d[i], yv = y1[i];
  double2 c = first ? cmk(0, 0) : inc[i];
  c.x += dv.x * ai;
  c.y += dv.y * ai;
  inc[i] = c;
  ynext[i] = cmk(yv.x + dv.x * an, yv.y + dv.y * an);
}

// y = y1 + (inc + d dt6): the arithmetic of the forward RK4 (state += inc + d dt6)
__global__ void k_rk_last(double2* __restrict__ y, const double2* __restrict__ y1, const double2* __restrict__ inc,
                          const double2* __restrict__ d, double dt6, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 a = inc[i], dv = d[i], yv = y1[i];
  y[i] = cmk(yv.x + (a.x + dv.x * dt6), yv.y + (a.y + dv.y * dt6));
}

// adjoint of the RK4 combination: the stage cotangents dt6, dt3, dt3, dt6 times lam
__global__ void k_rk_adj_init(const double2* __restrict__ lam, double2* __restrict__ d1, double2* __restrict__ d2,
                              double2* __restrict__ d3, double2* __restrict__ d4, double dt6, double dt3, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 l = lam[i];
  d1[i] = cscale(l, dt6);
  d2[i] = cscale(l, dt3);
  d3[i] = cscale(l, dt3);
  d4[i] = cscale(l, dt6);
}

// lam += g; dd += a g (dd may be null)
__global__ void k_rk_adj_acc(double2* __restrict__ lam, double2* __restrict__ dd, const double2* __restrict__ g,
                             double a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 gv = g[i];
  lam[i] = cadd(lam[i], gv);
  if (dd) {
    double2 v = dd[i];
    v.x += a * gv.x;
    v.y += a * gv.y;
    dd[i] = v;
  }
}

// ------------------------------------------------------------------------------- host side
int aerr(int code, const char* msg) { return set_error(code, msg); }

#define HIPCHK(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) return aerr(CODY_ERR_HIP, hipGetErrorString(e_));   \
  } while (0)
#define CHK(expr)      \
  do {                 \
    int r_ = (expr);   \
    if (r_) return r_; \
  } while (0)

KinParams kin_params(const cody_ctx* c) {
  return KinParams{c->kin.kref, c->kin.nE, c->kin.nLoc, c->kin.xminExcl, c->nG, c->kin.invh, c->kin.fac};
}

unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

int alloc_vjp(cody_ctx* c) {
  AdjBuffers& a = c->adj;
  if (a.ready) return CODY_OK;
  const size_t n2 = (size_t)c->nO * c->nO, n4 = n2 * n2, nOG = (size_t)c->nO * c->nG, nS = c->nS;
  const size_t nb = (c->nG + 255) / 256, ns = (size_t)c->Pc * c->nGp;
  int r = 0;
#define AL(p, bytes) \
  if ((r = ctx_alloc((void**)&(p), (bytes))) != 0) return r;
  AL(a.f, sizeof(double2) * nS);
  AL(a.D, sizeof(double2) * c->nC);
  AL(a.tmpC, sizeof(double2) * c->nC);
  AL(a.Bpart, sizeof(double2) * nb * n2);
  AL(a.B, sizeof(double2) * n2);
  AL(a.inp, sizeof(double2) * n2);
  AL(a.h1b, sizeof(double2) * n2);
  AL(a.h2b, sizeof(double2) * n4);  // entries with a repeated index pair stay zero
  AL(a.rho2p, sizeof(double2) * n4);
  AL(a.r2pb, sizeof(double2) * n4);
  AL(a.r2b, sizeof(double2) * n4);
  AL(a.lin1, sizeof(double2) * n2);
  AL(a.lin2, sizeof(double2) * n4);  // only the canonical codes are written
  AL(a.dbar, sizeof(double2) * nOG);
  AL(a.hbar, sizeof(double2) * nOG);
  AL(a.X, sizeof(double2) * n2);
  AL(a.STadj, sizeof(double) * ns);  // pad rows (c >= 2P) stay zero
  AL(a.WTadj, sizeof(double) * ns);
  AL(a.epart, sizeof(double) * nb * kMaxOrbs);  // one row per orbital
  AL(a.eacc, sizeof(double) * nOG);
  AL(a.gpart, sizeof(double2) * nOG * ((c->P + 3) / 4));
  AL(a.J, sizeof(double) * 2);
  for (int q = 0; q < 4; ++q) {
    AL(a.y[q], sizeof(double2) * nS);
    AL(a.dl[q], sizeof(double2) * nS);
  }
  AL(a.rk_d, sizeof(double2) * nS);
  AL(a.rk_inc, sizeof(double2) * nS);
  AL(a.lam, sizeof(double2) * nS);
  AL(a.g, sizeof(double2) * nS);
  AL(a.ys, sizeof(double2) * nS);
#undef AL
  a.ready = true;
  return CODY_OK;
}

int check_adjoint(const cody_ctx* c) {
  CHK(check_ready(c));
  if (c->precision != CODY_PREC_FP64)
    return aerr(CODY_ERR_UNSUPPORTED, "the adjoint needs the FP64 contraction (cody_set_precision)");
  if (!c->d_Vt) return aerr(CODY_ERR_STATE, "FP64 contraction operand not prepared");
  return CODY_OK;
}

// grad = J^T lam of the right-hand side `part` at (state, t); *ebar = <lam, df/dE>
int enqueue_vjp(cody_ctx* c, int part, const double2* state, const double2* lam, double2* grad, double t,
                double* ebar, hipStream_t st) {
  AdjBuffers& a = c->adj;
  const int nC = c->nC, nO = c->nO, nG = c->nG, nGp = c->nGp, n2 = nO * nO;
  const double2* C = state;
  const double2* orbs = state + nC;
  double2* gO = grad + nC;
  const double E = laser_field(c, t);
  const KinParams kp = kin_params(c);
  const unsigned nb = blocks(nG);
  if (part == CODY_PART_LIN) {
    launch_zero(grad, nC, st);
    hipLaunchKernelGGL(k_adj_hbar_lin, dim3(blocks((size_t)nO * nG)), dim3(256), 0, st, lam + nC, nO * nG, a.hbar);
    launch_zero(gO, nO * nG, st);
#define CALL_F0(NO)                                                                                              \
  hipLaunchKernelGGL(k_adj_final<NO>, dim3(nb, NO), dim3(256), 0, st, orbs, c->d_w, c->d_x, c->d_vext, kp, E,    \
                     nullptr, nG, nGp, nullptr, 0, a.hbar, gO, a.epart)
    ADJ_NO_SWITCH(nO, CALL_F0)
#undef CALL_F0
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, st, a.epart, (int)(nb * nO), ebar);
    return CODY_OK;
  }
  int mode = part == CODY_PART_FULL ? kModeFull : kModeNonLin;
  if (mode == kModeNonLin && c->split_consistent) mode = kModeNonLinConsistent;
  const double ah = mode == kModeNonLin ? 0.0 : 1.0, bh = mode == kModeNonLinConsistent ? -1.0 : 0.0;
  // the forward pass leaves rho1, rho2, R, h phi, h1, G, W, draw, the inp partials and dCraw
  CHK(enqueue_rhs(c, state, a.f, t, mode, st, nullptr, nullptr));
  // CI projection and H-apply
  hipLaunchKernelGGL(k_adj_ci_start, dim3(1), dim3(256), 0, st, C, lam, c->d_dCraw, nC, a.D, grad);
  launch_ci_apply(a.D, nC, nO, c->Pc, c->d_G, c->d_h1, c->dci, a.tmpC, st);  // H^H D (H Hermitian)
  launch_accumulate(grad, a.tmpC, nC, st);
  const int np = nO * (nO - 1) / 2, items = n2 + np * np;
  hipLaunchKernelGGL(k_adj_transition, dim3((items + 3) / 4), dim3(256), sizeof(double2) * 2 * nC, st, C, a.D, nC,
                     nO, c->dci.r1_ptr, c->dci.r1, c->dci.r2_ptr, c->dci.r2, a.h1b, a.h2b);
  // orbital projection and h1
  launch_sum_parts(c->d_part, c->nPart, n2, n2, a.inp, st);
  const int nbB = (int)nb;
#define CALL_B(NO)                                                                                              \
  hipLaunchKernelGGL(k_adj_bdots<NO>, dim3(nb, NO), dim3(256), 0, st, orbs, c->d_w, lam + nC, nG, a.Bpart);      \
  launch_sum_parts(a.Bpart, nbB, n2, n2, a.B, st);                                                             \
  hipLaunchKernelGGL(k_adj_orb1<NO>, dim3(nb), dim3(256), 0, st, orbs, lam + nC, c->d_draw, c->d_hphi, a.B, a.inp, \
                     a.h1b, nG, ah, bh, a.dbar, gO, a.hbar);                                                     \
  hipLaunchKernelGGL(k_adj_rho2p<NO>, dim3((NO * NO * NO * NO + 255) / 256), dim3(256), 0, st, c->d_rho2,        \
                     c->d_reginv, a.rho2p);                                                                      \
  hipLaunchKernelGGL(k_adj_local<NO>, dim3(nGp / 64, (NO * (NO + 1) / 2 + 3) / 4), dim3(256), 0, st, orbs, c->d_w, \
                     c->d_WT, nG, nGp, a.rho2p, a.h2b, a.dbar, a.gpart, a.STadj);                                \
  hipLaunchKernelGGL(k_adj_r2pb<NO>, dim3(NO * NO * NO), dim3(256), 0, st, orbs, c->d_w, c->d_WT, nG, nGp, a.dbar, \
                     a.r2pb);                                                                                    \
  hipLaunchKernelGGL(k_adj_rdm<NO>, dim3(1), dim3(256), 0, st, c->d_rho1, c->d_reginv, c->d_rho2, a.r2pb, c->reg,  \
                     a.r2b, a.lin1, a.lin2)
  ADJ_NO_SWITCH(nO, CALL_B)
#undef CALL_B
  hipLaunchKernelGGL(k_adj_ci_apply_lin, dim3((nC + 3) / 4), dim3(256), 0, st, C, nC, nO, c->dci.s_ptr,
                     c->dci.singles, c->dci.d_ptr, c->dci.doubles, a.lin1, a.lin2, grad);
  // sbar = V (w Wbar) on the FP64 MFMA contraction of the forward pass
  if (int r = launch_mf_f64_w(c->d_Vt, a.STadj, c->d_Wpart, a.WTadj, nGp, c->Pc, c->wsplit, st)) return r;
#define CALL_F(NO)                                                                                               \
  hipLaunchKernelGGL(k_adj_final<NO>, dim3(nb, NO), dim3(256), 0, st, orbs, c->d_w, c->d_x, c->d_vext, kp, E,    \
                     a.WTadj, nG, nGp, a.gpart, (NO * (NO + 1) / 2 + 3) / 4, a.hbar, gO, a.epart)
  ADJ_NO_SWITCH(nO, CALL_F)
#undef CALL_F
  hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, st, a.epart, (int)(nb * nO), ebar);
  return CODY_OK;
}

int enqueue_dipole_grad(cody_ctx* c, const double2* state, double2* grad, hipStream_t st) {
  AdjBuffers& a = c->adj;
  const int nC = c->nC, nO = c->nO, nG = c->nG, n2 = nO * nO;
  launch_ci_density(state, nC, nO, c->reg, c->dci, c->d_rho1, c->d_rho2, c->d_r2can, c->d_reginv, st);
  launch_xmat(state + nC, c->d_w, c->d_x, nG, nO, c->d_part1, st);
  launch_sum_parts(c->d_part1, (nG + 255) / 256, n2, n2, a.X, st);
#define CALL_D(NO)                                                                                          \
  hipLaunchKernelGGL(k_adj_dipole_orb<NO>, dim3(blocks(nG)), dim3(256), 0, st, state + nC, c->d_x, c->d_rho1, \
                     nG, grad + nC)
  ADJ_NO_SWITCH(nO, CALL_D)
#undef CALL_D
  launch_zero(grad, nC, st);
  hipLaunchKernelGGL(k_adj_lin1_from, dim3(1), dim3(256), 0, st, a.X, nO, a.lin1);
  hipLaunchKernelGGL(k_adj_ci_apply_lin, dim3((nC + 3) / 4), dim3(256), 0, st, state, nC, nO, c->dci.s_ptr,
                     c->dci.singles, c->dci.d_ptr, c->dci.doubles, a.lin1, nullptr, grad);
  return CODY_OK;
}

// Bessel J_0..J_K(z), z >= 0, by Miller's backward recurrence normalised by J_0 + 2 sum J_2k = 1
std::vector<double> bessel(int K, double z) {
  std::vector<double> J(K + 1, 0.0);
  if (z == 0.0) {
    J[0] = 1.0;
    return J;
  }
  int M = std::max(K, (int)z);
  M = 2 * ((M + 20 + (int)std::sqrt(40.0 * (M + 1))) / 2);
  std::vector<double> v(M + 2, 0.0);
  v[M] = 1e-300;
  for (int k = M; k >= 1; --k) {
    v[k - 1] = (2.0 * k / z) * v[k] - v[k + 1];
    if (std::fabs(v[k - 1]) > 1e250)
      for (int q = k - 1; q <= M; ++q) v[q] *= 1e-250;
  }
  double norm = v[0];
  for (int k = 2; k <= M; k += 2) norm += 2.0 * v[k];
  for (int k = 0; k <= K && k <= M; ++k) J[k] = v[k] / norm;
  return J;
}

// orbitals <- sum_k a_k T_k((h - c)/r) orbitals, terms kept in U[k] (U[0] a copy of the input)
int cheb_half(cody_ctx* c, double2* orbs, double E, double2* U, hipStream_t st) {
  AdjBuffers& a = c->adj;
  const size_t nOG = (size_t)c->nO * c->nG;
  const double cc = 0.5 * (a.hi + a.lo), rinv = 2.0 / (a.hi - a.lo);
  const KinParams kp = kin_params(c);
  HIPCHK(hipMemcpyAsync(U, orbs, sizeof(double2) * nOG, hipMemcpyDeviceToDevice, st));
  const dim3 grid(blocks(c->nG), c->nO);
  const int halo = kp.nLoc - 1;
  const bool pairs = halo >= 1 && halo <= kChebHaloMax;
  int k = 1;
  for (; pairs && k + 1 <= a.K; k += 2)
    hipLaunchKernelGGL(k_cheb_term2, grid, dim3(256), 0, st, k >= 2 ? U + (size_t)(k - 2) * nOG : nullptr,
                       U + (size_t)(k - 1) * nOG, U + (size_t)k * nOG, U + (size_t)(k + 1) * nOG, a.cacc,
                       k + 1 == a.K ? orbs : nullptr, kp, c->d_w, c->d_x, c->d_vext, E, cc, rinv, a.a[0], a.a[k],
                       a.a[k + 1], k == 1 ? 1 : 0, halo, c->nG);
  for (; k <= a.K; ++k)
    hipLaunchKernelGGL(k_cheb_term, grid, dim3(256), 0, st, k >= 2 ? U + (size_t)(k - 2) * nOG : nullptr,
                       U + (size_t)(k - 1) * nOG, U + (size_t)k * nOG, a.cacc, k == a.K ? orbs : nullptr, kp, c->d_w,
                       c->d_x, c->d_vext, E, cc, rinv, a.a[0], a.a[k], k == 1 ? 1 : 0, c->nG);
  return CODY_OK;
}

// lamOrbs <- A^T lamOrbs for the half-step with terms U; *ebar = <lam, dA/dE phi>
int cheb_half_adj(cody_ctx* c, double2* lamOrbs, double E, const double2* U, double* ebar, hipStream_t st) {
  AdjBuffers& a = c->adj;
  const size_t nOG = (size_t)c->nO * c->nG;
  const double cc = 0.5 * (a.hi + a.lo), rinv = 2.0 / (a.hi - a.lo);
  const KinParams kp = kin_params(c);
  const dim3 grid(blocks(c->nG), c->nO);
  hipLaunchKernelGGL(k_zero_d, dim3(blocks(nOG)), dim3(256), 0, st, a.eacc, (int)nOG);
  const int K = a.K;
  auto conj_a = [&](int k) { return make_double2(a.a[k].x, -a.a[k].y); };
  // b_k lives in cb[k % 4] (b_0 in lamOrbs); steps K - 1 .. 2 run two per launch
  hipLaunchKernelGGL(k_cheb_adj_term, grid, dim3(256), 0, st, lamOrbs, nullptr, nullptr, a.cb[K % 4], nullptr,
                     a.eacc, kp, c->d_w, c->d_x, c->d_vext, E, cc, rinv, conj_a(K), 0.0, c->nG);
  const int halo = kp.nLoc - 1;
  const bool pairs = halo >= 1 && halo <= kChebHaloMax;
  int k = K - 1;
  for (; pairs && k - 1 >= 1; k -= 2)
    hipLaunchKernelGGL(k_cheb_adj_term2, grid, dim3(256), 0, st, lamOrbs, a.cb[(k + 1) % 4],
                       k + 2 <= K ? a.cb[(k + 2) % 4] : nullptr, a.cb[k % 4], a.cb[(k - 1) % 4], U + (size_t)k * nOG,
                       U + (size_t)(k - 1) * nOG, a.eacc, kp, c->d_w, c->d_x, c->d_vext, E, cc, rinv, conj_a(k),
                       conj_a(k - 1), halo, c->nG);
  for (; k >= 0; --k)
    hipLaunchKernelGGL(k_cheb_adj_term, grid, dim3(256), 0, st, lamOrbs, a.cb[(k + 1) % 4],
                       k + 2 <= K ? a.cb[(k + 2) % 4] : nullptr, k == 0 ? lamOrbs : a.cb[k % 4],
                       U + (size_t)k * nOG, a.eacc, kp, c->d_w, c->d_x, c->d_vext, E, cc, rinv, conj_a(k),
                       k >= 1 ? 2.0 : 1.0, c->nG);
  hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, st, a.eacc, (int)nOG, ebar);
  return CODY_OK;
}

// What the adjoint of one Strang step reads: the Chebyshev terms of both half-steps and the
// inputs of the four RK4 stages
struct StepTape {
  double2 *U1, *U2, *y[4];
};

StepTape default_tape(cody_ctx* c) {
  AdjBuffers& a = c->adj;
  return StepTape{a.U1, a.U2, {a.y[0], a.y[1], a.y[2], a.y[3]}};
}

size_t tape_elems(const cody_ctx* c) {
  return 2 * (size_t)(c->adj.K + 1) * c->nO * c->nG + 4 * (size_t)c->nS;
}

StepTape tape_slot(cody_ctx* c, size_t i) {
  double2* b = c->adj.tape + i * tape_elems(c);
  const size_t nu = (size_t)(c->adj.K + 1) * c->nO * c->nG;
  return StepTape{b, b + nu, {b + 2 * nu, b + 2 * nu + c->nS, b + 2 * nu + 2 * c->nS, b + 2 * nu + 3 * c->nS}};
}

// One Strang step of y at tm: A(dt/2) with E(tm + dt/4), RK4 on the nonlinear part, A(dt/2)
// with E(tm + 3dt/4); the Chebyshev terms go to tp.U1, tp.U2 and the stage inputs to tp.y[0..3].
// The stages are formed in the fixed buffers a.y[0..3] (the keys of the captured right-hand
// sides) and copied to the tape when it is elsewhere.
int strang_cheb_step(cody_ctx* c, double2* y, double tm, const StepTape& tp, hipStream_t st) {
  AdjBuffers& a = c->adj;
  const double dt = a.dt, dt2 = dt / 2.0, dt3 = dt / 3.0, dt6 = dt / 6.0;
  const int n = c->nS;
  const unsigned nb = blocks(n);
  auto keep = [&](int q) -> int {
    if (tp.y[q] != a.y[q])
      HIPCHK(hipMemcpyAsync(tp.y[q], a.y[q], sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
    return CODY_OK;
  };
  CHK(cheb_half(c, y + c->nC, laser_field(c, tm + 0.25 * dt), tp.U1, st));
  HIPCHK(hipMemcpyAsync(a.y[0], y, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
  CHK(keep(0));
  CHK(run_rhs(c, a.y[0], a.rk_d, tm, kModeNonLin, st));
  hipLaunchKernelGGL(k_rk_stage, dim3(nb), dim3(256), 0, st, a.y[0], a.rk_d, a.rk_inc, a.y[1], dt6, dt2, 1, n);
  CHK(keep(1));
  CHK(run_rhs(c, a.y[1], a.rk_d, tm + dt2, kModeNonLin, st));
  hipLaunchKernelGGL(k_rk_stage, dim3(nb), dim3(256), 0, st, a.y[0], a.rk_d, a.rk_inc, a.y[2], dt3, dt2, 0, n);
  CHK(keep(2));
  CHK(run_rhs(c, a.y[2], a.rk_d, tm + dt2, kModeNonLin, st));
  hipLaunchKernelGGL(k_rk_stage, dim3(nb), dim3(256), 0, st, a.y[0], a.rk_d, a.rk_inc, a.y[3], dt3, dt, 0, n);
  CHK(keep(3));
  CHK(run_rhs(c, a.y[3], a.rk_d, tm + dt, kModeNonLin, st));
  hipLaunchKernelGGL(k_rk_last, dim3(nb), dim3(256), 0, st, y, a.y[0], a.rk_inc, a.rk_d, dt6, n);
  CHK(cheb_half(c, y + c->nC, laser_field(c, tm + 0.75 * dt), tp.U2, st));
  return CODY_OK;
}

// a.lam <- (step at tm)^T a.lam from the step's tape; the six dE cotangents go to eb[0..5] at
// tm + 3dt/4, tm + dt, tm + dt/2, tm + dt/2, tm, tm + dt/4
int strang_cheb_adj(cody_ctx* c, double tm, const StepTape& tp, double* eb, hipStream_t st) {
  AdjBuffers& a = c->adj;
  const double dt = a.dt, dt2 = dt / 2.0, dt3 = dt / 3.0, dt6 = dt / 6.0;
  const int n = c->nS;
  const unsigned nb = blocks(n);
  CHK(cheb_half_adj(c, a.lam + c->nC, laser_field(c, tm + 0.75 * dt), tp.U2, eb + 0, st));
  // d1..d4 = dt6, dt3, dt3, dt6 times lam
  hipLaunchKernelGGL(k_rk_adj_init, dim3(nb), dim3(256), 0, st, a.lam, a.dl[0], a.dl[1], a.dl[2], a.dl[3], dt6, dt3,
                     n);
  CHK(enqueue_vjp(c, CODY_PART_NONLIN, tp.y[3], a.dl[3], a.g, tm + dt, eb + 1, st));
  hipLaunchKernelGGL(k_rk_adj_acc, dim3(nb), dim3(256), 0, st, a.lam, a.dl[2], a.g, dt, n);
  CHK(enqueue_vjp(c, CODY_PART_NONLIN, tp.y[2], a.dl[2], a.g, tm + dt2, eb + 2, st));
  hipLaunchKernelGGL(k_rk_adj_acc, dim3(nb), dim3(256), 0, st, a.lam, a.dl[1], a.g, dt2, n);
  CHK(enqueue_vjp(c, CODY_PART_NONLIN, tp.y[1], a.dl[1], a.g, tm + dt2, eb + 3, st));
  hipLaunchKernelGGL(k_rk_adj_acc, dim3(nb), dim3(256), 0, st, a.lam, a.dl[0], a.g, dt2, n);
  CHK(enqueue_vjp(c, CODY_PART_NONLIN, tp.y[0], a.dl[0], a.g, tm, eb + 4, st));
  hipLaunchKernelGGL(k_rk_adj_acc, dim3(nb), dim3(256), 0, st, a.lam, nullptr, a.g, 0.0, n);
  CHK(cheb_half_adj(c, a.lam + c->nC, laser_field(c, tm + 0.25 * dt), tp.U1, eb + 5, st));
  return CODY_OK;
}

double laser_basis(const cody_ctx* c, int k, double t) {
  if (c->nLaser <= 0 || t < 0.0 || t > c->laserT) return 0.0;
  const double env = std::sin(M_PI * t / c->laserT);
  return std::sin((k + 1) * M_PI * t / c->laserT) * env * env;
}

}  // namespace

extern "C" {

int cody_time_derivative_vjp_dev(cody_ctx* c, int part, const void* d_state, const void* d_lam, void* d_grad,
                                 double t, void* d_ebar, void* stream) {
  if (!c || !d_state || !d_lam || !d_grad || !d_ebar) return aerr(CODY_ERR_ARG, "vjp: null argument");
  if (part != CODY_PART_FULL && part != CODY_PART_LIN && part != CODY_PART_NONLIN)
    return aerr(CODY_ERR_ARG, "vjp: unknown right-hand-side part");
  if (d_grad == d_state || d_grad == d_lam) return aerr(CODY_ERR_ARG, "vjp: the gradient aliases an input");
  CHK(check_adjoint(c));
  HIPCHK(hipSetDevice(c->device));
  CHK(alloc_vjp(c));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  CHK(enqueue_vjp(c, part, static_cast<const double2*>(d_state), static_cast<const double2*>(d_lam),
                  static_cast<double2*>(d_grad), t, static_cast<double*>(d_ebar), st));
  HIPCHK(hipGetLastError());
  return CODY_OK;
}

int cody_dipole_grad_dev(cody_ctx* c, const void* d_state, void* d_grad, void* stream) {
  if (!c || !d_state || !d_grad) return aerr(CODY_ERR_ARG, "dipole_grad: null argument");
  if (d_grad == d_state) return aerr(CODY_ERR_ARG, "dipole_grad: the gradient aliases the state");
  if (!c->grid_set) return aerr(CODY_ERR_STATE, "dipole_grad: grid not set");
  HIPCHK(hipSetDevice(c->device));
  CHK(alloc_vjp(c));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  CHK(enqueue_dipole_grad(c, static_cast<const double2*>(d_state), static_cast<double2*>(d_grad), st));
  HIPCHK(hipGetLastError());
  return CODY_OK;
}

int cody_set_chebyshev(cody_ctx* c, double dt, double emax, double* lo, double* hi, int* degree) {
  if (!c || !(dt > 0.0) || !(emax >= 0.0)) return aerr(CODY_ERR_ARG, "cody_set_chebyshev: bad arguments");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  CHK(alloc_vjp(c));
  AdjBuffers& a = c->adj;
  const int nG = c->nG;
  double *dlo = nullptr, *dhi = nullptr;
  CHK(ctx_alloc((void**)&dlo, sizeof(double) * nG));
  if (ctx_alloc((void**)&dhi, sizeof(double) * nG)) {
    hipFree(dlo);
    return aerr(CODY_ERR_HIP, "cody_set_chebyshev: allocation failed");
  }
  hipLaunchKernelGGL(k_gersh, dim3(blocks(nG)), dim3(256), 0, c->stream, kin_params(c), c->d_w, c->d_x, c->d_vext, nG,
                     emax, dlo, dhi);
  std::vector<double> hl(nG), hh(nG);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(hl.data(), dlo, sizeof(double) * nG, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hh.data(), dhi, sizeof(double) * nG, hipMemcpyDeviceToHost);
  hipFree(dlo);
  hipFree(dhi);
  if (e != hipSuccess) return aerr(CODY_ERR_HIP, hipGetErrorString(e));
  double l = *std::min_element(hl.begin(), hl.end()), h = *std::max_element(hh.begin(), hh.end());
  const double pad = 1e-9 * std::max(1.0, h - l);  // rounding of the bound itself
  l -= pad;
  h += pad;
  const double tau = 0.5 * dt, z = tau * 0.5 * (h - l);
  const int Kmax = (int)(z + 60 + 10 * std::cbrt(z + 1));
  std::vector<double> J = bessel(Kmax, z);
  int K = Kmax;
  while (K > 1 && std::fabs(J[K]) < 1e-17 && std::fabs(J[K - 1]) < 1e-17) --K;
  // reallocate the term buffers when the degree grows
  const size_t nOG = (size_t)c->nO * nG;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (K > a.K || !a.U1) {
    a.free_cheb();
    int r = 0;
    if ((r = ctx_alloc((void**)&a.U1, sizeof(double2) * nOG * (K + 1))) ||
        (r = ctx_alloc((void**)&a.U2, sizeof(double2) * nOG * (K + 1))) ||
        (r = ctx_alloc((void**)&a.cacc, sizeof(double2) * nOG)) || (r = ctx_alloc((void**)&a.cb[0], sizeof(double2) * nOG)) ||
        (r = ctx_alloc((void**)&a.cb[1], sizeof(double2) * nOG)) || (r = ctx_alloc((void**)&a.cb[2], sizeof(double2) * nOG)) ||
        (r = ctx_alloc((void**)&a.cb[3], sizeof(double2) * nOG)))
      return r;
  }
  a.K = K;
  a.lo = l;
  a.hi = h;
  a.dt = dt;
  a.emax = emax;
  // a_k = (2 - delta_k0) (-i)^k J_k(tau r) e^{-i tau c}
  const double cc = 0.5 * (h + l), ph = -tau * cc;
  const double2 eph = make_double2(std::cos(ph), std::sin(ph));
  a.a.assign(K + 1, make_double2(0, 0));
  double2 mi = make_double2(1, 0);
  for (int k = 0; k <= K; ++k) {
    const double s = (k == 0 ? 1.0 : 2.0) * J[k];
    const double2 m = make_double2(mi.x * s, mi.y * s);
    a.a[k] = make_double2(m.x * eph.x - m.y * eph.y, m.x * eph.y + m.y * eph.x);
    mi = make_double2(mi.y, -mi.x);  // times -i
  }
  if (lo) *lo = l;
  if (hi) *hi = h;
  if (degree) *degree = K;
  return CODY_OK;
}

static int check_cheb_dt(const cody_ctx* c, double dt) {
  if (c->adj.K < 1)
    return aerr(CODY_ERR_STATE, "cody_set_chebyshev has not been called since the one-body operator was set");
  if (std::fabs(dt - c->adj.dt) > 1e-12 * std::fabs(c->adj.dt))
    return aerr(CODY_ERR_ARG, "the step differs from the dt of cody_set_chebyshev");
  // |E(t)| <= sum |c_k|: the expansion's interval must hold the spectrum for every field value
  double bound = 0.0;
  for (int k = 0; k < c->nLaser; ++k) bound += std::fabs(c->laserC[k]);
  if (bound > c->adj.emax * (1.0 + 1e-12))
    return aerr(CODY_ERR_ARG, "the laser field can exceed the emax of cody_set_chebyshev");
  return CODY_OK;
}

int cody_strang_cheb(cody_ctx* c, double t0, double t1) {
  if (!c) return aerr(CODY_ERR_ARG, "cody_strang_cheb: null context");
  CHK(check_adjoint(c));
  CHK(check_cheb_dt(c, t1 - t0));
  HIPCHK(hipSetDevice(c->device));
  CHK(strang_cheb_step(c, c->d_state, t0, default_tape(c), c->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return CODY_OK;
}

int cody_dipole_gradient(cody_ctx* c, double t0, double dt, int nSteps, int every, double* J, double* gradc) {
  if (!c || nSteps < 1 || every < 1 || !J || (c->nLaser > 0 && !gradc))
    return aerr(CODY_ERR_ARG, "cody_dipole_gradient: bad arguments");
  CHK(check_adjoint(c));
  CHK(check_cheb_dt(c, dt));
  HIPCHK(hipSetDevice(c->device));
  AdjBuffers& a = c->adj;
  const size_t nS = c->nS, nCk = (size_t)(nSteps + every - 1) / every;
  hipStream_t st = c->stream;
  HIPCHK(hipStreamSynchronize(st));
  if (a.nck < nCk) {
    if (a.ck) hipFree(a.ck);
    a.ck = nullptr;
    a.nck = 0;
    CHK(ctx_alloc((void**)&a.ck, sizeof(double2) * nS * nCk));
    a.nck = nCk;
  }
  // one tape per step of a segment (about 37 MB per step at 4096 points, M = 8, degree 32);
  // without room for them the backward pass re-runs each step into the single work tape
  const int seg = std::min(every, nSteps);
  const size_t tape_bytes = sizeof(double2) * tape_elems(c) * seg;
  if (a.ntape < (size_t)seg || a.tape_K != a.K) {
    if (a.tape) hipFree(a.tape);
    a.tape = nullptr;
    a.ntape = 0;
    if (hipMalloc((void**)&a.tape, tape_bytes) == hipSuccess) {
      a.ntape = seg;
      a.tape_K = a.K;
    } else {
      (void)hipGetLastError();
      a.tape = nullptr;
    }
  }
  if (!a.tape && a.nseg < (size_t)every) {
    if (a.seg) hipFree(a.seg);
    a.seg = nullptr;
    a.nseg = 0;
    CHK(ctx_alloc((void**)&a.seg, sizeof(double2) * nS * every));
    a.nseg = every;
  }
  if (a.nebar < 6 * (size_t)nSteps) {
    if (a.ebar) hipFree(a.ebar);
    a.ebar = nullptr;
    a.nebar = 0;
    CHK(ctx_alloc((void**)&a.ebar, sizeof(double) * 6 * nSteps));
    a.nebar = 6 * (size_t)nSteps;
  }
  // forward pass with checkpoints; with one segment over the whole run and room for its tape,
  // the forward pass records every step's tape itself and nothing is recomputed
  const StepTape work = default_tape(c);
  const bool taped_forward = a.tape && every >= nSteps;
  for (int m = 0; m < nSteps; ++m) {
    if (m % every == 0)
      HIPCHK(hipMemcpyAsync(a.ck + nS * (m / every), c->d_state, sizeof(double2) * nS, hipMemcpyDeviceToDevice, st));
    CHK(strang_cheb_step(c, c->d_state, t0 + m * dt, taped_forward ? tape_slot(c, m) : work, st));
  }
  // the objective and its gradient at the final state
  launch_ci_density(c->d_state, c->nC, c->nO, c->reg, c->dci, c->d_rho1, c->d_rho2, c->d_r2can, c->d_reginv, st);
  launch_xmat(c->d_state + c->nC, c->d_w, c->d_x, c->nG, c->nO, c->d_part1, st);
  launch_dipole(c->nO, c->d_part1, (c->nG + 255) / 256, c->d_rho1, a.J, st);
  CHK(enqueue_dipole_grad(c, c->d_state, a.lam, st));
  // backward pass, segment by segment from the last checkpoint
  for (long s = (long)nCk - 1; s >= 0; --s) {
    const int m0 = (int)s * every, m1 = std::min(nSteps, m0 + every);
    HIPCHK(hipMemcpyAsync(a.ys, a.ck + nS * s, sizeof(double2) * nS, hipMemcpyDeviceToDevice, st));
    if (a.tape) {
      // recompute the segment once (unless the forward pass kept its tape), keeping every
      // step's tape, then run its adjoint backwards
      if (!taped_forward)
        for (int m = m0; m < m1; ++m) CHK(strang_cheb_step(c, a.ys, t0 + m * dt, tape_slot(c, m - m0), st));
      for (int m = m1 - 1; m >= m0; --m)
        CHK(strang_cheb_adj(c, t0 + m * dt, tape_slot(c, m - m0), a.ebar + 6 * (size_t)m, st));
    } else {
      // the segment's states, then each step re-run into the work tape before its adjoint
      for (int m = m0; m < m1; ++m) {
        HIPCHK(hipMemcpyAsync(a.seg + nS * (m - m0), a.ys, sizeof(double2) * nS, hipMemcpyDeviceToDevice, st));
        if (m + 1 < m1) CHK(strang_cheb_step(c, a.ys, t0 + m * dt, work, st));
      }
      for (int m = m1 - 1; m >= m0; --m) {
        HIPCHK(hipMemcpyAsync(a.ys, a.seg + nS * (m - m0), sizeof(double2) * nS, hipMemcpyDeviceToDevice, st));
        CHK(strang_cheb_step(c, a.ys, t0 + m * dt, work, st));
        CHK(strang_cheb_adj(c, t0 + m * dt, work, a.ebar + 6 * (size_t)m, st));
      }
    }
    HIPCHK(hipGetLastError());
  }
  std::vector<double> eb(6 * (size_t)nSteps);
  HIPCHK(hipMemcpyAsync(eb.data(), a.ebar, sizeof(double) * eb.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(J, a.J, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int k = 0; k < c->nLaser; ++k) gradc[k] = 0.0;
  for (int m = 0; m < nSteps; ++m) {
    const double tm = t0 + m * dt;
    const double ts[6] = {tm + 0.75 * dt, tm + dt, tm + 0.5 * dt, tm + 0.5 * dt, tm, tm + 0.25 * dt};
    for (int q = 0; q < 6; ++q)
      for (int k = 0; k < c->nLaser; ++k) gradc[k] += eb[6 * (size_t)m + q] * laser_basis(c, k, ts[q]);
  }
  return CODY_OK;
}

}  // extern "C"
