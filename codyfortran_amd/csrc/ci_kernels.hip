// CI-side kernels of the MCTDHX right-hand side (one workgroup of 1024 threads each; the CI
// space is small, n_C <= 1024), written so that every phase is parallel over its outputs and
// reads its operands from LDS:
//   k_ci_density  : rho1, rho2 from C, R = rho1 (rho1^2 + eps)^{-1}, rho2' = rho2 x_3 R
//   k_ci_finalize : h2 from the Gram matrix G = S^T W, CI H-apply, projection, energy
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

// Block-wide complex sum (1024 threads = 16 waves), result broadcast to all threads.
__device__ double2 block_csum(double2 v, double* red) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) {
    red[2 * wv] = v.x;
    red[2 * wv + 1] = v.y;
  }
  __syncthreads();
  double2 s = cmk(0, 0);
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s = cadd(s, cmk(red[2 * k], red[2 * k + 1]));
  return s;
}

}  // namespace

// ------------------------------------------------------------------------------------ densities
// Dynamic LDS: C[nC] | r1lin[n2] | r2lin[nCan] | A[n2] | Aorig[n2] | Minv[2 n2] | X[n2] | f[nO]
__global__ __launch_bounds__(1024) void k_ci_density(
    const double2* __restrict__ C, int nC, int nO, double reg, const int* __restrict__ r1_ptr,
    const CiEntryDev* __restrict__ r1, const int* __restrict__ r2_ptr,
    const CiEntryDev* __restrict__ r2, double2* __restrict__ rho1, double2* __restrict__ rho2,
    double2* __restrict__ rho2p, double2* __restrict__ reginv) {
  extern __shared__ double2 lds[];
  const int t = threadIdx.x, nt = blockDim.x;
  const int n2 = nO * nO, n4 = n2 * n2, np = nO * (nO - 1) / 2, nCan = np * np;
  double2* sC = lds;
  double2* s1 = sC + nC;
  double2* s2 = s1 + n2;
  double2* sA = s2 + nCan;
  double2* sA0 = sA + n2;
  double2* sM = sA0 + n2;  // augmented [A | I], nO x 2nO, column-major
  double2* sX = sM + 2 * n2;
  double2* sF = sX + n2;
  for (int i = t; i < nC; i += nt) sC[i] = C[i];
  __syncthreads();
  // rho1Lin[code = i nO + j] = sum f conj(C_jc) C_ic  ->  rho1(i,j)
  for (int e = t; e < n2; e += nt) {
    const int i = e % nO, j = e / nO;
    const int code = i * nO + j;
    double2 acc = cmk(0, 0);
    for (int q = r1_ptr[code]; q < r1_ptr[code + 1]; ++q) {
      const CiEntryDev en = r1[q];
      cfma(acc, cscale(cconj(sC[en.jc]), en.f), sC[en.ic]);
    }
    s1[e] = acc;
    rho1[e] = acc;
  }
  // canonical rho2Lin for i1 < i2, j1 < j2 (code ((i1 nO + j1) nO + i2) nO + j2)
  for (int e = t; e < nCan; e += nt) {
    const int pi = e % np, pj = e / np;
    int i1 = 0, rem = pi;
    while (rem >= nO - 1 - i1) {
      rem -= nO - 1 - i1;
      ++i1;
    }
    const int i2 = i1 + 1 + rem;
    int j1 = 0;
    rem = pj;
    while (rem >= nO - 1 - j1) {
      rem -= nO - 1 - j1;
      ++j1;
    }
    const int j2 = j1 + 1 + rem;
    const int code = ((i1 * nO + j1) * nO + i2) * nO + j2;
    double2 acc = cmk(0, 0);
    for (int q = r2_ptr[code]; q < r2_ptr[code + 1]; ++q) {
      const CiEntryDev en = r2[q];
      cfma(acc, cscale(cconj(sC[en.jc]), en.f), sC[en.ic]);
    }
    s2[e] = acc;
  }
  __syncthreads();
  // full antisymmetric rho2 (unpack of S_Coeffs_Generic.f90:362-434)
  for (int e = t; e < n4; e += nt) {
    int a = e % nO, b = (e / nO) % nO, c = (e / n2) % nO, d = e / (n2 * nO);
    double2 v = cmk(0, 0);
    if (a != b && c != d) {
      double sg = 1.0;
      if (a > b) { int tmp = a; a = b; b = tmp; sg = -sg; }
      if (c > d) { int tmp = c; c = d; d = tmp; sg = -sg; }
      v = cscale(s2[strict_pair(nO, a, b) + np * strict_pair(nO, c, d)], sg);
    }
    rho2[e] = v;
  }
  // A = rho1^2 + eps I ; augmented [A | I]
  for (int e = t; e < n2; e += nt) {
    const int i = e % nO, j = e / nO;
    double2 acc = cmk(i == j ? reg : 0.0, 0.0);
    for (int k = 0; k < nO; ++k) cfma(acc, s1[i + k * nO], s1[k + j * nO]);
    sA[e] = acc;
    sA0[e] = acc;
    sM[e] = acc;
    sM[n2 + e] = cmk(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  // Gauss-Jordan inverse (A is Hermitian positive definite: no pivoting needed)
  for (int k = 0; k < nO; ++k) {
    const double2 piv = sM[k + k * nO];
    __syncthreads();
    for (int c = t; c < 2 * nO; c += nt) sM[k + c * nO] = cdiv(sM[k + c * nO], piv);
    for (int r = t; r < nO; r += nt) sF[r] = sM[r + k * nO];
    __syncthreads();
    for (int e = t; e < 2 * n2; e += nt) {
      const int r = e % nO, c = e / nO;
      if (r != k) sM[r + c * nO] = csub(sM[r + c * nO], cmul(sF[r], sM[k + c * nO]));
    }
    __syncthreads();
  }
  // X = A^-1 rho1, then one refinement X += A^-1 (rho1 - A X)
  const double2* sInv = sM + n2;
  for (int e = t; e < n2; e += nt) {
    const int i = e % nO, j = e / nO;
    double2 acc = cmk(0, 0);
    for (int k = 0; k < nO; ++k) cfma(acc, sInv[i + k * nO], s1[k + j * nO]);
    sX[e] = acc;
  }
  __syncthreads();
  for (int e = t; e < n2; e += nt) {
    const int i = e % nO, j = e / nO;
    double2 r = s1[e];
    for (int k = 0; k < nO; ++k) r = csub(r, cmul(sA0[i + k * nO], sX[k + j * nO]));
    sA[e] = r;  // residual
  }
  __syncthreads();
  for (int e = t; e < n2; e += nt) {
    const int i = e % nO, j = e / nO;
    double2 acc = sX[e];
    for (int k = 0; k < nO; ++k) cfma(acc, sInv[i + k * nO], sA[k + j * nO]);
    reginv[e] = acc;
    sA0[e] = acc;  // R
  }
  __syncthreads();
  // rho2'(i1,i2,j1,l2) = sum_k rho2(i1,i2,k,l2) R(k,j1)   (S_Method_Mb_OrbBased.f90:746-753)
  for (int e = t; e < n4; e += nt) {
    int a = e % nO, b = (e / nO) % nO;
    const int j1 = (e / n2) % nO, l2 = e / (n2 * nO);
    double2 acc = cmk(0, 0);
    if (a != b) {
      double sa = 1.0;
      if (a > b) { int tmp = a; a = b; b = tmp; sa = -1.0; }
      const int pa = strict_pair(nO, a, b);
      for (int k = 0; k < nO; ++k) {
        if (k == l2) continue;
        int c = k, d = l2;
        double sg = sa;
        if (c > d) { int tmp = c; c = d; d = tmp; sg = -sg; }
        cfma(acc, cscale(s2[pa + np * strict_pair(nO, c, d)], sg), sA0[k + j1 * nO]);
      }
    }
    rho2p[e] = acc;
  }
}

// ------------------------------------------------------------------------------------ partials
// h1/inp: sum over the per-tile partials; G: sum over the x-split partials.
__global__ __launch_bounds__(256) void k_reduce_partials(const double2* __restrict__ part, int nPart,
                                                         int n2, const double* __restrict__ Gpart,
                                                         int nxs, int nG2e, double2* __restrict__ h1,
                                                         double2* __restrict__ inp,
                                                         double* __restrict__ G) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 2 * n2) {
    double2 acc = cmk(0, 0);
    for (int b = 0; b < nPart; ++b) acc = cadd(acc, part[(size_t)b * 2 * n2 + e]);
    if (e < n2) h1[e] = acc;
    else inp[e - n2] = acc;
  } else if (e < 2 * n2 + nG2e) {
    const int g = e - 2 * n2;
    double acc = 0.0;
    for (int s = 0; s < nxs; ++s) acc += Gpart[(size_t)s * nG2e + g];
    G[g] = acc;
  }
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

// Dynamic LDS: C[nC] | h1[n2] | h2lin[nCan] | dC[nC] | red[64 doubles]
__global__ __launch_bounds__(1024) void k_ci_finalize(
    const double2* __restrict__ C, int nC, int nO, int Pc, const double* __restrict__ G,
    const double2* __restrict__ h1g, const int* __restrict__ s_ptr,
    const CiEntryDev* __restrict__ sing, const int* __restrict__ d_ptr,
    const CiEntryDev* __restrict__ dbl, const double2* __restrict__ rho1,
    const double2* __restrict__ rho2, double2* __restrict__ h2, double2* __restrict__ dC_out,
    double* __restrict__ energy) {
  extern __shared__ double2 lds[];
  const int t = threadIdx.x, nt = blockDim.x;
  const int n2 = nO * nO, n4 = n2 * n2, np = nO * (nO - 1) / 2, nCan = np * np;
  double2* sC = lds;
  double2* sh1 = sC + nC;
  double2* sh2 = sh1 + n2;
  double2* sdC = sh2 + nCan;
  double* red = reinterpret_cast<double*>(sdC + nC);
  for (int i = t; i < nC; i += nt) sC[i] = C[i];
  for (int e = t; e < n2; e += nt) sh1[e] = h1g[e];
  // antisymmetrised h2Lin for i1 < i2, j1 < j2 (FillH2Lin, S_Coeffs_Generic.f90:478-542)
  for (int e = t; e < nCan; e += nt) {
    const int pi = e % np, pj = e / np;
    int i1 = 0, rem = pi;
    while (rem >= nO - 1 - i1) {
      rem -= nO - 1 - i1;
      ++i1;
    }
    const int i2 = i1 + 1 + rem;
    int j1 = 0;
    rem = pj;
    while (rem >= nO - 1 - j1) {
      rem -= nO - 1 - j1;
      ++j1;
    }
    const int j2 = j1 + 1 + rem;
    double2 v = h2_from_G(G, Pc, nO, i1, i2, j1, j2);
    v = csub(v, h2_from_G(G, Pc, nO, i1, i2, j2, j1));
    v = csub(v, h2_from_G(G, Pc, nO, i2, i1, j1, j2));
    v = cadd(v, h2_from_G(G, Pc, nO, i2, i1, j2, j1));
    sh2[e] = v;
  }
  // full h2 for the energy and for callers
  for (int e = t; e < n4; e += nt) {
    const int i1 = e % nO, a = (e / nO) % nO, j1 = (e / n2) % nO, b = e / (n2 * nO);
    h2[e] = h2_from_G(G, Pc, nO, i1, a, j1, b);
  }
  __syncthreads();
  // CI H-apply: one wave per configuration, lanes over its connections
  const int wv = t >> 6, ln = t & 63, nw = nt >> 6;
  for (int ic = wv; ic < nC; ic += nw) {
    double2 acc = cmk(0, 0);
    for (int q = s_ptr[ic] + ln; q < s_ptr[ic + 1]; q += 64) {
      const CiEntryDev en = sing[q];
      const int i = en.code / nO, j = en.code % nO;
      cfma(acc, cscale(sh1[i + j * nO], en.f), sC[en.jc]);
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
      cfma(acc, cscale(sh2[strict_pair(nO, i1, i2) + np * strict_pair(nO, j1, j2)], en.f), sC[en.jc]);
    }
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (ln == 0) sdC[ic] = acc;
  }
  __syncthreads();
  // dC -= <C|dC> C ; dC_out = -i dC   (src/Coeffs/S_Coeffs.f90:105-111)
  double2 ov = cmk(0, 0);
  for (int i = t; i < nC; i += nt) {
    const double2 c = sC[i], d = sdC[i];
    ov = cadd(ov, cmk(c.x * d.x + c.y * d.y, c.x * d.y - c.y * d.x));
  }
  ov = block_csum(ov, red);
  for (int i = t; i < nC; i += nt) {
    const double2 r = csub(sdC[i], cmul(ov, sC[i]));
    dC_out[i] = cmk(r.y, -r.x);
  }
  if (energy) {
    // E = sum h1(k,l) rho1(l,k) + sum h2(k1,k2,l1,l2) rho2(l1,l2,k1,k2)
    double2 e = cmk(0, 0);
    for (int q = t; q < n2; q += nt) {
      const int k = q % nO, l = q / nO;
      cfma(e, sh1[k + l * nO], rho1[l + k * nO]);
    }
    for (int q = t; q < n4; q += nt) {
      const int k1 = q % nO, k2 = (q / nO) % nO, l1 = (q / n2) % nO, l2 = q / (n2 * nO);
      cfma(e, h2_from_G(G, Pc, nO, k1, k2, l1, l2), rho2[idx4(l1, l2, k1, k2, nO)]);
    }
    e = block_csum(e, red);
    if (t == 0) *energy = e.x;
  }
}

// ------------------------------------------------------------------------------------ launchers
size_t ci_density_lds_bytes(int nC, int nO) {
  const int n2 = nO * nO, np = nO * (nO - 1) / 2;
  return sizeof(double2) * (size_t)(nC + n2 + np * np + 6 * n2 + nO);
}
size_t ci_finalize_lds_bytes(int nC, int nO) {
  const int n2 = nO * nO, np = nO * (nO - 1) / 2;
  return sizeof(double2) * (size_t)(2 * nC + n2 + np * np) + 64 * sizeof(double);
}

int ci_kernels_allow_lds(size_t bytes) {
  hipError_t e1 = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ci_density),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  hipError_t e2 = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ci_finalize),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  return (e1 == hipSuccess && e2 == hipSuccess) ? 0 : 1;
}

void launch_ci_density(const double2* C, int nC, int nO, double reg, const DevCiTables& t,
                       double2* rho1, double2* rho2, double2* rho2p, double2* reginv, hipStream_t st) {
  hipLaunchKernelGGL(k_ci_density, dim3(1), dim3(1024), ci_density_lds_bytes(nC, nO), st, C, nC, nO,
                     reg, t.r1_ptr, t.r1, t.r2_ptr, t.r2, rho1, rho2, rho2p, reginv);
}

void launch_reduce_partials(const double2* part, int nPart, int n2, const double* Gpart, int nxs,
                            int nG2e, double2* h1, double2* inp, double* G, hipStream_t st) {
  const int n = 2 * n2 + nG2e;
  hipLaunchKernelGGL(k_reduce_partials, dim3((n + 255) / 256), dim3(256), 0, st, part, nPart, n2, Gpart,
                     nxs, nG2e, h1, inp, G);
}

void launch_ci_finalize(const double2* C, int nC, int nO, int Pc, const double* G, const double2* h1,
                        const DevCiTables& t, const double2* rho1, const double2* rho2, double2* h2,
                        double2* dC, double* energy, hipStream_t st) {
  hipLaunchKernelGGL(k_ci_finalize, dim3(1), dim3(1024), ci_finalize_lds_bytes(nC, nO), st, C, nC, nO, Pc,
                     G, h1, t.s_ptr, t.singles, t.d_ptr, t.doubles, rho1, rho2, h2, dC, energy);
}

}  // namespace cody
