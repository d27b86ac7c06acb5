// CI-side kernels of the MCTDHX right-hand side (the CI space is small, n_C <= 1024):
//   k_ci_rdm      : rho1, rho2 from C, one wave per density element
//   k_ci_reginv   : R = rho1 (rho1^2 + eps)^{-1} (one wave)
//   k_gram_reduce : G = S^T W from the split-K partials of the Gram kernel; h1 from its partials
//   k_ci_apply    : CI H-apply with h2 formed from G, one wave per configuration
//   k_ci_energy   : E = Tr[rho1 h1] + Tr[rho2 h2]
//   k_sum_parts   : h1 / inp from per-workgroup partials
#include <hip/hip_runtime.h>

#include "ci_device.hpp"

namespace cody {


// ------------------------------------------------------------------------------------ densities

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
  ci_rdm_body(blockIdx.x, C, nC, nO, r1_ptr, r1, r2_ptr, r2, rho1, r2can, rho2);
}

// R = rho1 (rho1^2 + eps)^{-1} (regInvMat, S_Method_Mb_OrbBased.f90:737-744) by Gauss-Jordan on
// rho1^2 + eps I (Hermitian positive definite: no pivoting) and one refinement step, in one wave
// (the barriers of a one-wave workgroup cost nothing).  rho2' = rho2 x_3 R is formed slice by
// slice inside the orbital kernel.
template <int NO>
__global__ __launch_bounds__(64) void k_ci_reginv(const double2* __restrict__ rho1g, double reg,
                                                  double2* __restrict__ reginv) {
  reginv_body<NO>(rho1g, reg, reginv, blockDim.x);
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
  gram_reduce_body(blockIdx.x, Gpart, nxs, Pc, G, part1, nPart1, n2, h1);
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
  const double2 acc = ci_config_sum(ic, ln, C, nO, Pc, G, h1, s_ptr, sing, d_ptr, dbl);
  if (ln == 0) dCraw[ic] = acc;
}

__global__ __launch_bounds__(256) void k_ci_apply_project(const double2* __restrict__ C, int nC, int nO, int Pc,
                                                          const double* __restrict__ G,
                                                          const double2* __restrict__ h1, DevCiTables t,
                                                          const double* __restrict__ epart, int nE,
                                                          double2* __restrict__ dCraw, double2* __restrict__ dC) {
  ci_apply_project_body(blockIdx.x, C, nC, nO, Pc, G, h1, t, epart, nE, dCraw, dC);
}

// the CI H-apply and projection of the instances of a grouped right-hand side
__global__ __launch_bounds__(256) void k_grp_apply(const GroupInstance* __restrict__ gi, const int2* __restrict__ map,
                                                   int nO) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  ci_apply_project_body(m.y, a.C, a.nC, nO, a.Pc, a.G, a.h1, a.t, a.epart, a.nRed, a.dCraw, a.dC);
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
  const int per = 4 * ci_rdm_items_per_wave(nC);
  hipLaunchKernelGGL(k_ci_rdm, dim3((items + per - 1) / per), dim3(256), sizeof(double2) * nC, st, C, nC, nO, t.r1_ptr,
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

void launch_ci_apply_project(const double2* C, int nC, int nO, int Pc, const double* G, const double2* h1,
                             const DevCiTables& t, const double* epart, int nE, double2* dCraw, double2* dC,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_ci_apply_project, dim3((nC + 3) / 4), dim3(256), 0, st, C, nC, nO, Pc, G, h1, t, epart, nE,
                     dCraw, dC);
}

void launch_grp_apply(const GroupInstance* gi, const int2* map, int nBlocks, int nO, hipStream_t st) {
  hipLaunchKernelGGL(k_grp_apply, dim3(nBlocks), dim3(256), 0, st, gi, map, nO);
}

void launch_dipole(int nO, const double2* part1, int nPart, const double2* rho1, double* dipole, hipStream_t st) {
  hipLaunchKernelGGL(k_dipole, dim3(1), dim3(256), 0, st, nO, part1, nPart, rho1, dipole);
}

void launch_ci_energy(int nO, int Pc, const double* G, const double2* h1, const double2* rho1,
                      const double2* rho2, double* energy, hipStream_t st) {
  hipLaunchKernelGGL(k_ci_energy, dim3(1), dim3(256), 0, st, nO, Pc, G, h1, rho1, rho2, energy);
}

}  // namespace cody
