// CDNA4 (gfx950) kernels of the MCTDHX right-hand side.
//
// Data layout in HBM (one context, one fermionic body type, nO orbitals, nG grid points):
//   state      complex128 [C(nC) ; phi(nG, nO) column-major]   (the reference's Method_state)
//   V          FP64 nGp x nGp row-major, V[i][j] = K(|x_i - x_j|), zero padded to nGp = 256*ceil(nG/256)
//   ST         FP64 Pc x nGp: row 2p / 2p+1 = Re / Im of the pair source conj(phi_l) phi_k w,
//              p enumerating l <= k; Pc = 16*ceil(2P/16), unused rows zero
//   WT         FP64 Pc x nGp: mean-field potentials W_p(x) (Re/Im rows), from meanfield_gemm.hip
// Complex scalars are double2 (x = re, y = im).
#include <hip/hip_runtime.h>

#include "ci_device.hpp"
#include "fedvr_device.hpp"
#include "gemm_device.hpp"
#include "mctdhx_kernels.hpp"

namespace cody {

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

// The same S^T fragment-tiled for the FP64 kernels, one output element per thread (coalesced):
// STt[(t nK + K) 256 + 128 h + 2 l + e] = S^T[c = 16 t + l % 16][x = 16 K + 4 (l / 16) + 2 h + e].
// A workgroup covers one 16 x 16 fragment.  With act non-null it writes act[fragment]
// = 1 if any of its values is non-zero, else 0 (a plain store every call: nothing to clear); the
// contraction kernels skip the k blocks whose fragments are all zero, the sources the reference
// leaves out of its convolution sums (src/Utils/M_Utils_ConvolutionIntegral.f90:84-93).
__device__ __forceinline__ void pair_src_t_body(int bx, const double2* __restrict__ orbs,
                                                    const double* __restrict__ w, int nG, int nGp, int nO,
                                                    int Pc, double* __restrict__ STt,
                                                    int* __restrict__ act) {
  const size_t o = (size_t)bx * blockDim.x + threadIdx.x;  // grid = Pc nGp / 256 exactly
  const int nK = nGp >> 4;
  const int e = o & 1, h = (o >> 7) & 1, l = (o >> 1) & 63;
  const size_t blk = o >> 8;
  const int t = (int)(blk / nK), K = (int)(blk % nK);
  const int c = 16 * t + (l & 15), x = 16 * K + 4 * (l >> 4) + 2 * h + e;
  double v = 0.0;
  const int p = c >> 1;
  if (x < nG && p < nO * (nO + 1) / 2) {
    int a = 0, r = p;
    while (r >= nO - a) {
      r -= nO - a;
      ++a;
    }
    const double2 s = cscale(cmul(cconj(orbs[(size_t)a * nG + x]), orbs[(size_t)(a + r) * nG + x]), w[x]);
    v = (c & 1) ? s.y : s.x;
  }
  STt[o] = v;
  if (act) {
    const int any = __syncthreads_or(v != 0.0);
    if (threadIdx.x == 0) act[blk] = any ? 1 : 0;
  }
}

// The pair sources of one 16-line x 64-point block (t, K64) for the int8 form: S^T in the FP64
// fragment order of pair_src_t_body, and its five balanced base-254 digits in the int8 fragment
// order (digit p of block (t, K64) is 1 KiB at ((t nK64 + K64) NS + p) 1024, lane 16 (K16 % 4) +
// line % 16 holding the 16 points of K16), under a scale 2^e per line and block: e = the exponent
// of the block's largest |S| on that line plus one (|S| 2^-e <= 1/2), at eSb[K64 Pc + line].  So
// the digits need no grid-wide maximum and are formed where S is.  act[t nK64 + K64] = 1 when a
// value of the block is non-zero.  Wave w takes K16 = 4 K64 + w, each thread four points of one line.
__device__ __forceinline__ void pair_src_i8_body(int bx, const double2* __restrict__ orbs,
                                                 const double* __restrict__ w, int nG, int nGp, int nO, int Pc,
                                                 double* __restrict__ STt, int8_t* __restrict__ Si8,
                                                 int* __restrict__ eSb, int* __restrict__ act) {
  constexpr int NS = kInt8Slices;
  __shared__ double sM[4][64];
  __shared__ int sE[16];
  const int nK16 = nGp >> 4, nK64 = nGp >> 6;
  const int t = bx / nK64, K64 = bx % nK64;
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, K16 = 4 * K64 + wv;
  const int c = 16 * t + (l & 15), p = c >> 1;
  const bool pin = p < nO * (nO + 1) / 2;
  int a = 0, r = p;
  if (pin)
    while (r >= nO - a) {
      r -= nO - a;
      ++a;
    }
  double v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {  // u = 2 h + e of the FP64 fragment
    const int x = 16 * K16 + 4 * (l >> 4) + u;
    double s = 0.0;
    if (pin && x < nG) {
      const double2 z = cscale(cmul(cconj(orbs[(size_t)a * nG + x]), orbs[(size_t)(a + r) * nG + x]), w[x]);
      s = (c & 1) ? z.y : z.x;
    }
    v[u] = s;
  }
  double* dst = STt + ((size_t)t * nK16 + K16) * 256 + 2 * l;
  *reinterpret_cast<double2*>(dst) = make_double2(v[0], v[1]);
  *reinterpret_cast<double2*>(dst + 128) = make_double2(v[2], v[3]);
  sM[wv][l] = fmax(fmax(fabs(v[0]), fabs(v[1])), fmax(fabs(v[2]), fabs(v[3])));
  const int any = __syncthreads_or(v[0] != 0.0 || v[1] != 0.0 || v[2] != 0.0 || v[3] != 0.0);
  if (threadIdx.x < 16) {
    double m = 0.0;
#pragma unroll
    for (int w2 = 0; w2 < 4; ++w2)
#pragma unroll
      for (int q = 0; q < 4; ++q) m = fmax(m, sM[w2][16 * q + threadIdx.x]);
    int e = 0;
    if (m > 0.0) {
      int ex;
      const double fr = frexp(m, &ex);
      e = (fr == 0.5 ? ex - 1 : ex) + 1;
    }
    sE[threadIdx.x] = e;
    eSb[(size_t)K64 * Pc + 16 * t + threadIdx.x] = e;
  }
  if (threadIdx.x == 0) act[(size_t)t * nK64 + K64] = any ? 1 : 0;
  __syncthreads();
  const int e = sE[l & 15];
  double rr[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) rr[u] = ldexp(v[u], -e);
  int8_t* o = Si8 + ((size_t)t * nK64 + K64) * NS * 1024 + 16 * (16 * wv + (l & 15)) + 4 * (l >> 4);
#pragma unroll
  for (int pp = 0; pp < NS; ++pp) {
    unsigned word = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double tt = rr[u] * kSliceBase;
      const double q = rint(tt);
      rr[u] = tt - q;
      word |= (unsigned)(uint8_t)static_cast<int8_t>(q) << (8 * u);
    }
    *reinterpret_cast<unsigned*>(o + (size_t)pp * 1024) = word;
  }
}

__global__ __launch_bounds__(256) void k_pair_src_i8(const double2* __restrict__ orbs, const double* __restrict__ w,
                                                     int nG, int nGp, int nO, int Pc, double* __restrict__ STt,
                                                     int8_t* __restrict__ Si8, int* __restrict__ eSb,
                                                     int* __restrict__ act) {
  pair_src_i8_body(blockIdx.x, orbs, w, nG, nGp, nO, Pc, STt, Si8, eSb, act);
}

__global__ __launch_bounds__(256) void k_pair_src_t(const double2* __restrict__ orbs,
                                                    const double* __restrict__ w, int nG, int nGp, int nO,
                                                    int Pc, double* __restrict__ STt,
                                                    int* __restrict__ act) {
  pair_src_t_body(blockIdx.x, orbs, w, nG, nGp, nO, Pc, STt, act);
}

// Arbitrary complex sources (ns columns of nG points, column-major) in the fragment-tiled S^T
// layout of k_pair_src_t: rows 2s / 2s+1 hold Re / Im of source s; other rows and points past nG
// are zero.  One output element per thread (grid = Pc nGp / 256).
__global__ __launch_bounds__(256) void k_tile_src(const double2* __restrict__ src, int ns, int nG, int nGp,
                                                  double* __restrict__ STt) {
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nK = nGp >> 4;
  const int e = o & 1, h = (o >> 7) & 1, l = (o >> 1) & 63;
  const size_t blk = o >> 8;
  const int t = (int)(blk / nK), K = (int)(blk % nK);
  const int c = 16 * t + (l & 15), x = 16 * K + 4 * (l >> 4) + 2 * h + e;
  double v = 0.0;
  if (x < nG && (c >> 1) < ns) {
    const double2 z = src[(size_t)(c >> 1) * nG + x];
    v = (c & 1) ? z.y : z.x;
  }
  STt[o] = v;
}

// Partial sums of the overlaps <phi_i|f_j>_w over 256-point blocks, part[block][i + j NO]
// (the overlaps of Grid_ProjectOnSubspace, src/Grid/S_Grid.f90:170-203).
template <int NO>
__global__ __launch_bounds__(256) void k_overlap(const double2* __restrict__ orbs, const double* __restrict__ w,
                                                 const double2* __restrict__ f, int nG,
                                                 double2* __restrict__ part) {
  __shared__ double2 sRed[4][NO];
  const int j = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = blockIdx.x * 256 + threadIdx.x;
  const bool in = x < nG;
  const int xc = in ? x : nG - 1;
  const double2 fw = in ? cscale(f[(size_t)j * nG + xc], w[xc]) : cmk(0, 0);
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    double2 acc = cmk(0, 0);
    if (in) cfmac(acc, orbs[(size_t)i * nG + xc], fw);
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

// ------------------------------------------------------------------ one-body part
// One workgroup = 256 grid points x one orbital j (blockIdx.y):
//   hphi_j = (T + V_ext + E(t) x) phi_j                   (ApplySingleBodyOp :976-1034)
// and the workgroup's partial sums of h1(i,j) = <phi_i|hphi_j>_w, part1[block][i + j NO].
template <int NO>
__device__ __forceinline__ void onebody_body(int bx, int j, const double2* __restrict__ orbs,
                                             const double* __restrict__ w, const double* __restrict__ xg,
                                             const double* __restrict__ vext, KinParams kp,
                                             const double* __restrict__ efield, int nG, double2* __restrict__ hphi,
                                             double2* __restrict__ part1) {
  __shared__ double sBuf[8 * 256];
  __shared__ double sRes[2 * NO];
  __shared__ double2 sF[256 + 2 * kMaxStencilReach];
  __shared__ double sK[(kMaxStencilReach + 1) * (kMaxStencilReach + 1)];
  const int x = bx * 256 + threadIdx.x;
  const bool in = x < nG;
  const int xc = in ? x : nG - 1;
  // orbital j on this tile and on the stencil's reach (nLoc - 1 points) either side, in LDS
  const double2* fj = orbs + (size_t)j * nG;
  const int halo = kp.nLoc - 1, lo = bx * 256 - halo;
  for (int t = threadIdx.x; t < 256 + 2 * halo; t += 256) {
    const int i = lo + t;
    sF[t] = (i >= 0 && i < nG) ? fj[i] : cmk(0, 0);
  }
  for (int t = threadIdx.x; t < kp.nLoc * kp.nLoc; t += 256) sK[t] = kp.kref[t];
  double2 phi[NO];
#pragma unroll
  for (int i = 0; i < NO; ++i) phi[i] = in ? orbs[(size_t)i * nG + xc] : cmk(0, 0);
  const double wx = in ? w[xc] : 0.0;
  const double v = in ? vext[xc] + *efield * xg[xc] : 0.0;
  __syncthreads();
  double2 h = cmk(0, 0);
  if (in) {
    h = kinetic_at_lds_k(sF, lo, xc, kp, sK, 1.0 / wx);
    const double2 pj = sF[xc - lo];
    h.x = fma(v, pj.x, h.x);
    h.y = fma(v, pj.y, h.y);
    hphi[(size_t)j * nG + x] = h;
  }
  const double2 hw = cscale(h, wx);
  double v2[2 * NO];
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    double2 acc = cmk(0, 0);
    if (in) cfmac(acc, phi[i], hw);
    v2[2 * i] = acc.x;
    v2[2 * i + 1] = acc.y;
  }
  block_sums256<2 * NO>(v2, sBuf, sRes);
  if (threadIdx.x < NO) {
    const int i = threadIdx.x;
    part1[(size_t)bx * NO * NO + i + j * NO] = cmk(sRes[2 * i], sRes[2 * i + 1]);
  }
}

template <int NO>
__global__ __launch_bounds__(256) void k_onebody(const double2* __restrict__ orbs, const double* __restrict__ w,
                                                 const double* __restrict__ xg, const double* __restrict__ vext,
                                                 KinParams kp, const double* __restrict__ efield, int nG,
                                                 double2* __restrict__ hphi, double2* __restrict__ part1) {
  onebody_body<NO>(blockIdx.x, blockIdx.y, orbs, w, xg, vext, kp, efield, nG, hphi, part1);
}

// Partial sums of the position matrix x(i,j) = <phi_i|x phi_j>_w over 256-point blocks,
// part1[block][i + j NO]: the orbital matrix of PrinterObservableLinear_DumpDipole
// (src/Utils/M_Utils_PrinterObservableLinear.f90:57-68).  Same reduction shape as k_onebody.
template <int NO>
__global__ __launch_bounds__(256) void k_xmat(const double2* __restrict__ orbs, const double* __restrict__ w,
                                              const double* __restrict__ xg, int nG,
                                              double2* __restrict__ part1) {
  __shared__ double2 sRed[4][NO];
  const int j = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = blockIdx.x * 256 + threadIdx.x;
  const bool in = x < nG;
  const int xc = in ? x : nG - 1;
  const double2 xw = in ? cscale(orbs[(size_t)j * nG + xc], xg[xc] * w[xc]) : cmk(0, 0);
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    double2 acc = cmk(0, 0);
    if (in) cfmac(acc, orbs[(size_t)i * nG + xc], xw);
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (lane == 0) sRed[wave][i] = acc;
  }
  __syncthreads();
  if (threadIdx.x < NO) {
    const int i = threadIdx.x;
    const double2 acc = cadd(cadd(sRed[0][i], sRed[1][i]), cadd(sRed[2][i], sRed[3][i]));
    part1[(size_t)blockIdx.x * NO * NO + i + j * NO] = acc;
  }
}

// ------------------------------------------------------------------ orbital right-hand side
// One workgroup = 64 grid points (one per lane) x one orbital j (blockIdx.y); its 4 waves split
// the unordered pairs (l <= k, pair p taken by wave p % 4):
//   corr_j   = sum_{k1,(l2,k2)} W_{l2k2} phi_k1 rho2'(k1,k2,j,l2) (ApplyCorrelationOp :762-838)
//   draw_j   = hphi_j + corr_j (full) or corr_j (nonlinear part)
// and per-workgroup partial sums of inp(i,j) = <phi_i|draw_j>_w, part[block][i + j NO] (the NO
// sums are split over the waves).  The slice rho2'(:,:,j,:) (NO^3 values) is staged in LDS and
// read as broadcasts; each wave loads all of its W values before using them, and the four
// correlation shares are added in fixed order.
// The slice rho2'(:, :, j, :) of rho2'(k1, k, j, l) = sum_m rho2(k1, k, m, l) R(m, j) at
// e = k1 + NO (k + NO l) (S_Method_Mb_OrbBased.f90:746-753), from the canonical rho2 s2 (np x np,
// strict pairs) and column j of R: the same operations in the same order wherever it is formed.
template <int NO>
__device__ __forceinline__ double2 rho2p_entry(int e, const double2* s2, const double2* Rj) {
  constexpr int n2 = NO * NO, np = NO * (NO - 1) / 2, kNone = -(n2 + 1);
  auto pairc = [](int a, int b) {  // strict-pair index of {a, b}; -1 - index when a > b; kNone when a = b
    return a == b ? kNone : (a < b ? a * NO - (a * (a + 1)) / 2 + (b - a - 1) : -1 - (b * NO - (b * (b + 1)) / 2 + (a - b - 1)));
  };
  const int ab = e % n2, l = e / n2;
  const int pab = pairc(ab % NO, ab / NO);
  double2 acc = cmk(0, 0);
  if (pab != kNone) {
    const int pa = pab >= 0 ? pab : -1 - pab;
    const double sa = pab >= 0 ? 1.0 : -1.0;
#pragma unroll
    for (int m = 0; m < NO; ++m) {
      const int pcd = pairc(m, l);
      if (pcd == kNone) continue;
      const int pc = pcd >= 0 ? pcd : -1 - pcd;
      cfma(acc, cscale(s2[pa + np * pc], pcd >= 0 ? sa : -sa), Rj[m]);
    }
  }
  return acc;
}

template <int NO>
__device__ __forceinline__ void orbital_body(int bx, int j, const double2* __restrict__ orbs, const double* __restrict__ w,
                                                 const double2* __restrict__ hphi, const double* __restrict__ WT,
                                                 const double2* __restrict__ r2can, const double2* __restrict__ reginv,
                                                 int nG, int nGp, int mode, double2* __restrict__ draw,
                                                 double2* __restrict__ part) {
  constexpr int n2 = NO * NO, n3 = n2 * NO, NP = NO * (NO + 1) / 2, PW = (NP + 3) / 4;
  constexpr int np = NO * (NO - 1) / 2, nCan = np * np;
  __shared__ double2 s2[nCan > 0 ? nCan : 1];
  __shared__ double2 sR[n3];
  __shared__ double2 sRj[NO];
  __shared__ double2 sCorr[4][64];
  __shared__ double2 sRed[NO];
  __shared__ int sLK[NP];  // l + 256 k of pair p = l NO - l (l - 1) / 2 + k - l, l <= k
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = bx * 64 + lane;
  if (threadIdx.x < NP) {
    int l = 0, r = threadIdx.x;
    while (r >= NO - l) {
      r -= NO - l;
      ++l;
    }
    sLK[threadIdx.x] = l + 256 * (l + r);
  }
  {
    // canonical rho2 (i1 < i2, j1 < j2) and column j of R (each thread's loads of the canonical
    // rho2 in flight together), then the slice
    constexpr int kStage = nCan > 0 ? (nCan + 255) / 256 : 1;
    double2 v[kStage];
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int e = threadIdx.x + 256 * u;
      v[u] = e < nCan ? r2can[e] : cmk(0, 0);
    }
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int e = threadIdx.x + 256 * u;
      if (e < nCan) s2[e] = v[u];
    }
    if (threadIdx.x < NO) sRj[threadIdx.x] = reginv[threadIdx.x + j * NO];
    __syncthreads();
    for (int e = threadIdx.x; e < n3; e += 256) sR[e] = rho2p_entry<NO>(e, s2, sRj);
  }
  const bool in = x < nG;
  const int xc = in ? x : nG - 1;
  double2 phi[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) phi[k] = in ? orbs[(size_t)k * nG + xc] : cmk(0, 0);
  const double wx = in ? w[xc] : 0.0;
  double2 Wv[PW];
#pragma unroll
  for (int m = 0; m < PW; ++m) {
    const int p = wave + 4 * m;
    Wv[m] = (in && p < NP) ? cmk(WT[(size_t)(2 * p) * nGp + x], WT[(size_t)(2 * p + 1) * nGp + x]) : cmk(0, 0);
  }
  const double2 hj = (in && (mode == kModeFull || mode == kModeNonLinConsistent)) ? hphi[(size_t)j * nG + x] : cmk(0, 0);
  __syncthreads();  // sR staged
  double2 corr = cmk(0, 0);
#pragma unroll
  for (int m = 0; m < PW; ++m) {
    const int p = wave + 4 * m;  // uniform over the wave
    if (p >= NP) break;
    const int lk = sLK[p], l = lk & 255, k = lk >> 8;
    // A = sum_k1 phi_k1 rho2'(k1,k,j,l); B = sum_k1 phi_k1 rho2'(k1,l,j,k), each over even and
    // odd k1 in two chains (shorter dependency chains), then added
    const double2* ra = sR + NO * (k + NO * l);
    const double2* rb = sR + NO * (l + NO * k);
    double2 A0 = cmk(0, 0), A1 = cmk(0, 0), B0 = cmk(0, 0), B1 = cmk(0, 0);
#pragma unroll
    for (int k1 = 0; k1 < NO; k1 += 2) {
      cfma(A0, phi[k1], ra[k1]);
      cfma(B0, phi[k1], rb[k1]);
      if (k1 + 1 < NO) {
        cfma(A1, phi[k1 + 1], ra[k1 + 1]);
        cfma(B1, phi[k1 + 1], rb[k1 + 1]);
      }
    }
    cfma(corr, Wv[m], cadd(A0, A1));
    if (k != l) cfma(corr, cconj(Wv[m]), cadd(B0, B1));
  }
  sCorr[wave][lane] = corr;
  __syncthreads();
  corr = cadd(cadd(sCorr[0][lane], sCorr[1][lane]), cadd(sCorr[2][lane], sCorr[3][lane]));
  const double2 d = cadd(hj, corr);
  if (in && wave == 0) draw[(size_t)j * nG + x] = d;
  // inp(i, j) partial sums, sum i formed by wave i % 4
  const double2 dw = cscale(d, wx);
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    if ((i & 3) != wave) continue;
    double2 acc = cmk(0, 0);
    cfmac(acc, phi[i], dw);
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (lane == 0) sRed[i] = acc;
  }
  __syncthreads();
  if (threadIdx.x < NO) part[(size_t)bx * n2 + threadIdx.x + j * NO] = sRed[threadIdx.x];
}

template <int NO>
__global__ __launch_bounds__(256) void k_orbital(const double2* __restrict__ orbs, const double* __restrict__ w,
                                                 const double2* __restrict__ hphi, const double* __restrict__ WT,
                                                 const double2* __restrict__ r2can, const double2* __restrict__ reginv,
                                                 int nG, int nGp, int mode, double2* __restrict__ draw,
                                                 double2* __restrict__ part) {
  orbital_body<NO>(blockIdx.x, blockIdx.y, orbs, w, hphi, WT, r2can, reginv, nG, nGp, mode, draw, part);
}

// ------------------------------------------------------------------ orbital correlation on FP64 MFMA
// From kCorrQMinOrbs orbitals the correlation of orbital_body (nO^4 products per point, each reading
// an LDS broadcast) runs as a product on FP64 MFMA followed by a short contraction with phi:
//   corr_j(x)   = sum_k1 phi_k1(x) O_{k1 j}(x),
//   O_{k1 j}(x) = sum_p a_p(x) (X + Y) + i b_p(x) (X - Y),
// where W_p = a_p + i b_p for the pair p = {l <= k}, X = rho2'(k1, k, j, l) and Y = rho2'(k1, l, j, k)
// (zero for k = l), since W_p X + conj(W_p) Y = a_p (X + Y) + i b_p (X - Y) (ApplyCorrelationOp,
// S_Method_Mb_OrbBased.f90:762-838).  O is the real product of the 16-point block of W^T (rows
// 2p + re/im, as the contraction writes them) with Q (2P x 2 nO^2 real): row 2p + 0 holds the re/im
// parts of X + Y, row 2p + 1 those of i (X - Y).  Q's columns come in 16-column tiles, tile
// t = j TPJ + th holding re/im of O_{k1 j} for k1 = 8 th .. 8 th + 7; Q is stored in the order the
// MFMA's B operand is read (lane l of step s of tile t: row 4 s + l / 16, column l % 16, at
// ((t KS + s) 64 + l)), formed once per right-hand side from the canonical rho2 and R.
// One thread per (pair slot p < 2 KS, k1 < 8 TPJ, j): X and Y once, the four entries of rows 2p, 2p + 1
// and columns re/im of (k1, j) written as two pairs of neighbours (zeros for the padding slots).
template <int NO>
__device__ __forceinline__ void qform_body(int blk, const double2* __restrict__ r2can,
                                           const double2* __restrict__ reginv, double* __restrict__ qf) {
  constexpr int KS = corr_q_steps(NO), TPJ = corr_q_tiles_per_orb(NO), NP = NO * (NO + 1) / 2, K1 = 8 * TPJ;
  const int e = blk * 256 + threadIdx.x;
  if (e >= 2 * KS * K1 * NO) return;
  const int p = e % (2 * KS), k1 = (e / (2 * KS)) % K1, j = e / (2 * KS * K1);
  double2 X = cmk(0, 0), Y = cmk(0, 0);
  if (p < NP && k1 < NO) {
    int l = 0, r = p;  // pair p -> (l <= k)
    while (r >= NO - l) {
      r -= NO - l;
      ++l;
    }
    const int k = l + r;
    double2 Rj[NO];
#pragma unroll
    for (int m = 0; m < NO; ++m) Rj[m] = reginv[m + j * NO];
    X = rho2p_entry<NO>(k1 + NO * (k + NO * l), r2can, Rj);
    if (k != l) Y = rho2p_entry<NO>(k1 + NO * (l + NO * k), r2can, Rj);
  }
  // row 2p (a_p): re/im of X + Y; row 2p + 1 (b_p): re/im of i (X - Y)
  const int t = j * TPJ + (k1 >> 3), s = p >> 1, n = 2 * (k1 & 7);
  double* base = qf + ((size_t)t * KS + s) * 64 + 16 * (2 * (p & 1)) + n;
  *reinterpret_cast<double2*>(base) = cmk(X.x + Y.x, X.y + Y.y);
  *reinterpret_cast<double2*>(base + 16) = cmk(-(X.y - Y.y), X.x - Y.x);
}

// One workgroup per 64 points (bx) and group g of the orbitals j (j = g, g + G, ..., G =
// kCorrQGroups), one wave per 16 of the points: the wave holds its W^T block as the MFMA A operand
// in registers (KS values per lane) and walks its orbitals' Q tiles with Q's fragments loaded one
// tile ahead (loads complete in issue order, so every wait covers one tile of MFMAs).  The result
// of tile t (element i of lane l: point 4 i + l / 16, column l % 16) is contracted with phi_k1 of
// its column in registers and the 16 columns of a lane group are added by a fixed butterfly.  Then
// draw_j = h phi_j + corr_j (h phi in the modes with the one-body part) and the per-workgroup
// partial sums part[bx][i + j nO] = sum_x conj(phi_i) w draw_j of its orbitals j, as orbital_body
// leaves them.
template <int NO>
__device__ __forceinline__ void corr_q_body(int bx, int g, const double2* __restrict__ orbs,
                                            const double* __restrict__ w, const double2* __restrict__ hphi,
                                            const double* __restrict__ WT, const double* __restrict__ qf, int nG,
                                            int nGp, int mode, double2* __restrict__ draw,
                                            double2* __restrict__ part) {
  constexpr int n2 = NO * NO, NP = NO * (NO + 1) / 2, G = kCorrQGroups;
  constexpr int KS = corr_q_steps(NO), TPJ = corr_q_tiles_per_orb(NO), JPG = (NO + G - 1) / G;
  __shared__ double2 sCorr[JPG][64];
  __shared__ double2 sPhi[NO][64];
  __shared__ double sw[64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, q = lane >> 4, n = lane & 15;
  const int x0 = bx * 64, xw = x0 + wave * 16;
  const int nj = (NO - g + G - 1) / G;  // orbitals j = g + G u of this group, u < nj
  double a[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int kk = 4 * s + q;
    a[s] = kk < 2 * NP ? WT[(size_t)kk * nGp + xw + n] : 0.0;
  }
  const double* qp = qf + lane;
  double qb[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) qb[s] = qp[((size_t)g * TPJ * KS + s) * 64];
  double pr[TPJ][4], pim[TPJ][4];
#pragma unroll
  for (int th = 0; th < TPJ; ++th)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k1 = 8 * th + (n >> 1), gx = xw + 4 * i + q;
      const double2 v = (k1 < NO && gx < nG) ? orbs[(size_t)k1 * nG + gx] : cmk(0, 0);
      pr[th][i] = v.x;
      pim[th][i] = v.y;
    }
  int any = 0;
  for (int e = tid; e < NO * 64; e += 256) {
    const int i = e >> 6, xx = e & 63;
    const double2 v = x0 + xx < nG ? orbs[(size_t)i * nG + x0 + xx] : cmk(0, 0);
    sPhi[i][xx] = v;
    any |= (v.x != 0.0 || v.y != 0.0);
  }
  if (tid < 64) sw[tid] = x0 + tid < nG ? w[x0 + tid] : 0.0;
  // 64 points with no orbital value: the correlation (sum_k1 phi_k1 O_k1j) and the overlap partials
  // (products with conj(phi_i)) are exact zeros, so draw_j = h phi_j without the MFMA work
  if (!__syncthreads_or(any)) {
    const int xx = tid & 63, gx = x0 + xx;
    const bool in = gx < nG, hOn = mode == kModeFull || mode == kModeNonLinConsistent;
    for (int u = tid >> 6; u < nj; u += 4) {
      const int j = g + G * u;
      const double2 hj = (in && hOn) ? hphi[(size_t)j * nG + gx] : cmk(0, 0);
      if (in) draw[(size_t)j * nG + gx] = cadd(hj, cmk(0, 0));
    }
    for (int f = tid; f < NO * nj; f += 256) part[(size_t)bx * n2 + f % NO + NO * (g + G * (f / NO))] = cmk(0, 0);
    return;
  }
  const bool reCol = (n & 1) == 0;
  for (int u = 0; u < nj; ++u) {
    const int j = g + G * u;
    double re[4] = {0.0, 0.0, 0.0, 0.0}, im[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int th = 0; th < TPJ; ++th) {
      // the tile after this one in the wave's walk (the last one reloads itself, unused)
      const int tn = th + 1 < TPJ ? j * TPJ + th + 1 : (u + 1 < nj ? (j + G) * TPJ : j * TPJ + th);
      d4_t acc = (d4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const double qv = qb[s];
        qb[s] = qp[((size_t)tn * KS + s) * 64];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], qv, acc, 0, 0, 0);
      }
      // phi_k1 (re + i im) at this lane's column: the re column adds phi re, the im column i phi im
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        re[i] = fma(reCol ? pr[th][i] : -pim[th][i], acc[i], re[i]);
        im[i] = fma(reCol ? pim[th][i] : pr[th][i], acc[i], im[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        re[i] += __shfl_xor(re[i], off, 16);
        im[i] += __shfl_xor(im[i], off, 16);
      }
    if (n < 4) {
      const double rv = n == 0 ? re[0] : n == 1 ? re[1] : n == 2 ? re[2] : re[3];
      const double iv = n == 0 ? im[0] : n == 1 ? im[1] : n == 2 ? im[2] : im[3];
      sCorr[u][wave * 16 + 4 * n + q] = cmk(rv, iv);
    }
  }
  __syncthreads();
  const int xx = tid & 63, gx = x0 + xx;
  const bool in = gx < nG, hOn = mode == kModeFull || mode == kModeNonLinConsistent;
  for (int u = tid >> 6; u < nj; u += 4) {
    const int j = g + G * u;
    const double2 hj = (in && hOn) ? hphi[(size_t)j * nG + gx] : cmk(0, 0);
    const double2 d = cadd(hj, sCorr[u][xx]);
    if (in) draw[(size_t)j * nG + gx] = d;
    sCorr[u][xx] = cscale(d, sw[xx]);
  }
  __syncthreads();
  for (int f = tid; f < NO * nj; f += 256) {
    const int i = f % NO, u = f / NO;
    double2 acc = cmk(0, 0);
#pragma unroll 8
    for (int x = 0; x < 64; ++x) cfmac(acc, sPhi[i][x], sCorr[u][x]);
    part[(size_t)bx * n2 + i + NO * (g + G * u)] = acc;
  }
}

// Q (the first corr_q_blocks workgroups) beside the Gram matrix's split-K partials (G = S^T W, the
// others), which the correlation launch then leaves out.
template <int NO>
__global__ __launch_bounds__(256) void k_rhs_qform(const double2* __restrict__ r2can, const double2* __restrict__ reginv,
                                                   double* __restrict__ qf, const double* __restrict__ STt,
                                                   const double* __restrict__ WT, int nGp, int Pc,
                                                   double* __restrict__ Gpart, int nxs, const int* __restrict__ act,
                                                   int actShift) {
  if constexpr (NO >= kCorrQMinOrbs) {
    constexpr int NT = (NO * (NO + 1) + 15) / 16, nQ = corr_q_blocks(NO);
    const int b = blockIdx.x;
    if (b < nQ) {
      qform_body<NO>(b, r2can, reginv, qf);
      return;
    }
    gram_f64_body<NT>((b - nQ) % NT, (b - nQ) / NT, STt, nGp >> 4, WT, nGp, Gpart, Pc, nGp / (4 * nxs),
                      (size_t)Pc * Pc, act, actShift);
  }
}


// ------------------------------------------------------------------ overlap sums
// inp(i, j) = sum_b part[b][i + j NO] over nPart per-workgroup partials, into sInp (LDS): G groups
// of threads sum interleaved partial blocks, then the groups are added in order, so every
// workgroup (and every kernel that uses this) forms bit-identical sums.
template <int NO>
constexpr int kOverlapGroups = NO * NO >= 256 ? 1 : 256 / (NO * NO);

template <int NO, int G = kOverlapGroups<NO>>
__device__ __forceinline__ void sum_overlap_parts(const double2* __restrict__ part, int nPart,
                                                  double2 (*sG)[NO * NO], double2* sInp) {
  constexpr int n2 = NO * NO;
  for (int e = threadIdx.x; e < G * n2; e += blockDim.x) {
    const int g = e / n2, i = e % n2;
    double2 acc = cmk(0, 0);
#pragma unroll 8
    for (int b = g; b < nPart; b += G) acc = cadd(acc, part[(size_t)b * n2 + i]);
    sG[g][i] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    double2 acc = sG[0][i];
    for (int g = 1; g < G; ++g) acc = cadd(acc, sG[g][i]);
    sInp[i] = acc;
  }
  __syncthreads();
}

// In-place projection f_j <- f_j - sum_i phi_i inp(i, j), inp from the k_overlap partials
// (Grid_ProjectOnSubspace, src/Grid/S_Grid.f90:170-203, without the right-hand side's -i).
template <int NO>
__global__ __launch_bounds__(256) void k_project_inplace(const double2* __restrict__ orbs,
                                                         const double2* __restrict__ part, int nPart, int nG,
                                                         double2* __restrict__ f) {
  constexpr int n2 = NO * NO, G = kOverlapGroups<NO>;
  __shared__ double2 sG[G][n2];
  __shared__ double2 sInp[n2];
  sum_overlap_parts<NO>(part, nPart, sG, sInp);
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nG) return;
  double2 phi[NO];
#pragma unroll
  for (int i = 0; i < NO; ++i) phi[i] = orbs[(size_t)i * nG + x];
#pragma unroll
  for (int j = 0; j < NO; ++j) {
    double2 d = f[(size_t)j * nG + x];
#pragma unroll
    for (int i = 0; i < NO; ++i) d = csub(d, cmul(phi[i], sInp[i + j * NO]));
    f[(size_t)j * nG + x] = d;
  }
}

// ------------------------------------------------------------------ projection + scaling
// dphi_j = -i (draw_j - sum_i phi_i inp(i,j))  (Grid_ProjectOnSubspace, src/Grid/S_Grid.f90:170-203),
// inp = sum_b part[b] over the orbital kernel's per-workgroup partials (summed by every workgroup
// into LDS, in fixed order).  With dC non-null the last workgroup applies the CI projection and
// scaling dC = -i (dCraw - <C|dCraw> C)  (src/Coeffs/S_Coeffs.f90:105-111).
template <int NO>
__device__ __forceinline__ void project_body(int bx, bool ciBlock, const double2* __restrict__ orbs,
                                                 const double2* __restrict__ draw,
                                                 const double2* __restrict__ part, int nPart, int nG,
                                                 int project, double2* __restrict__ dorbs,
                                                 const double2* __restrict__ C,
                                                 const double2* __restrict__ dCraw, int nC,
                                                 double2* __restrict__ dC, const double2* __restrict__ hsub) {
  constexpr int n2 = NO * NO, G = kOverlapGroups<NO>;
  if (ciBlock) {
    __shared__ double red[2][4];
    double2 ov = cmk(0, 0);
    for (int i = threadIdx.x; i < nC; i += blockDim.x) cfmac(ov, C[i], dCraw[i]);
    ov.x = wave_sum(ov.x);
    ov.y = wave_sum(ov.y);
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = ov.x;
      red[1][threadIdx.x >> 6] = ov.y;
    }
    __syncthreads();
    ov = cmk((red[0][0] + red[0][1]) + (red[0][2] + red[0][3]), (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]));
    for (int i = threadIdx.x; i < nC; i += blockDim.x) dC[i] = mul_minus_i(csub(dCraw[i], cmul(ov, C[i])));
    return;
  }
  __shared__ double2 sG[G][n2];
  __shared__ double2 sInp[n2];
  if (project) sum_overlap_parts<NO>(part, nPart, sG, sInp);
  const int x = bx * blockDim.x + threadIdx.x;
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
      for (int i = 0; i < NO; ++i) d = csub(d, cmul(phi[i], sInp[i + j * NO]));
    }
    if (hsub) d = csub(d, hsub[(size_t)j * nG + x]);
    dorbs[(size_t)j * nG + x] = mul_minus_i(d);
  }
}

// The projection of the int8 path, whose contraction left c_m(x) (the correlation before R) and
// the overlap sums A(i, j) = <phi_i|h_j>_w, B(i, m) = <phi_i|c_m>_w (partr: A, then B; h = h phi,
// or null where the right-hand side has no one-body part):
//   d_j = h_j + sum_m R(m, j) c_m,   <phi_i|d_j>_w = A(i, j) + sum_m B(i, m) R(m, j),
//   dphi_j = -i [d_j - sum_i phi_i <phi_i|d_j>_w - hsub_j]  (hsub optional).
template <int NO>
__device__ __forceinline__ void project_corr_body(int bx, const double2* __restrict__ orbs,
                                                  const double2* __restrict__ h, const double2* __restrict__ corrc,
                                                  const double2* __restrict__ reginv,
                                                  const double2* __restrict__ partr, int nG,
                                                  double2* __restrict__ dorbs, const double2* __restrict__ hsub) {
  constexpr int n2 = NO * NO;
  static_assert(n2 <= 256, "one overlap entry per thread");
  __shared__ double2 sR[n2], sB[n2], sInp[n2];
  // every global load first (one latency): the point's orbitals, c and h, then R and the sums
  const int x = bx * blockDim.x + threadIdx.x;
  const bool in = x < nG;
  double2 phi[NO], cm[NO], hv[NO];
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    phi[i] = in ? orbs[(size_t)i * nG + x] : cmk(0, 0);
    cm[i] = in ? corrc[(size_t)i * nG + x] : cmk(0, 0);
    hv[i] = (in && h) ? h[(size_t)i * nG + x] : cmk(0, 0);
  }
  const int e = threadIdx.x;
  double2 av = cmk(0, 0);
  if (e < n2) {
    av = partr[e];
    sR[e] = reginv[e];
    sB[e] = partr[n2 + e];
  }
  __syncthreads();
  if (e < n2) {
    const int i = e % NO, j = e / NO;
#pragma unroll
    for (int m = 0; m < NO; ++m) cfma(av, sB[i + NO * m], sR[m + NO * j]);
    sInp[e] = av;
  }
  __syncthreads();
  if (!in) return;
#pragma unroll
  for (int j = 0; j < NO; ++j) {
    double2 d = hv[j];
#pragma unroll
    for (int m = 0; m < NO; ++m) cfma(d, sR[m + NO * j], cm[m]);
#pragma unroll
    for (int i = 0; i < NO; ++i) d = csub(d, cmul(phi[i], sInp[i + j * NO]));
    if (hsub) d = csub(d, hsub[(size_t)j * nG + x]);
    dorbs[(size_t)j * nG + x] = mul_minus_i(d);
  }
}

template <int NO>
__global__ __launch_bounds__(256) void k_project(const double2* __restrict__ orbs,
                                                 const double2* __restrict__ draw,
                                                 const double2* __restrict__ part, int nPart, int nG,
                                                 int project, double2* __restrict__ dorbs,
                                                 const double2* __restrict__ C,
                                                 const double2* __restrict__ dCraw, int nC,
                                                 double2* __restrict__ dC, const double2* __restrict__ hsub) {
  project_body<NO>(blockIdx.x, dC && blockIdx.x == gridDim.x - 1, orbs, draw, part, nPart, nG, project, dorbs, C,
                   dCraw, nC, dC, hsub);
}

// ------------------------------------------------------------------ scalar store (graph-capturable)
__global__ void k_set_scalar(double* __restrict__ p, double v) { *p = v; }

void launch_set_scalar(double* p, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, st, p, v);
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

// ------------------------------------------------------------------ operator pieces
// y += x elementwise
__global__ void k_accumulate(double2* __restrict__ y, const double2* __restrict__ x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = cadd(y[i], x[i]);
}

void launch_accumulate(double2* y, const double2* x, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, st, y, x, n);
}

// One pair source src = conj(phi_a) phi_b w (FillInteractionSrc, S_SysInteraction_Linear.f90:57-72).
__global__ void k_pair_src_one(const double2* __restrict__ orbs, const double* __restrict__ w, int nG, int a, int b,
                               double2* __restrict__ src) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < nG) src[x] = cscale(cmul(cconj(orbs[(size_t)a * nG + x]), orbs[(size_t)b * nG + x]), w[x]);
}

void launch_pair_src_one(const double2* orbs, const double* w, int nG, int a, int b, double2* src, hipStream_t st) {
  hipLaunchKernelGGL(k_pair_src_one, dim3((nG + 255) / 256), dim3(256), 0, st, orbs, w, nG, a, b, src);
}

// f(:, i) += pot * orbs(:, i) for every orbital i, pot = WT(0, :) + i WT(1, :), the potential of
// the first source of a contraction (MultiplyPotentialOnOrbRange with MultiplyWithInteractionPotential,
// S_Method_Mb_OrbBased.f90:545-572, S_SysInteraction_Linear.f90:79-88).
__global__ void k_apply_potential(const double* __restrict__ WT, int nGp, const double2* __restrict__ orbs, int nG,
                                  double2* __restrict__ f) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nG) return;
  const size_t o = (size_t)blockIdx.y * nG + x;
  f[o] = cadd(f[o], cmul(cmk(WT[x], WT[(size_t)nGp + x]), orbs[o]));
}

void launch_apply_potential(const double* WT, int nGp, const double2* orbs, int nG, int nO, double2* f,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_apply_potential, dim3((nG + 255) / 256, nO), dim3(256), 0, st, WT, nGp, orbs, nG, f);
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

// STt (fragment-tiled, always; with the int8 digits when Si8 is non-null) and, when ST is
// non-null, S^T row-major.
void launch_pair_src(const double2* orbs, const double* w, int nG, int nGp, int nO, int Pc,
                     double* ST, double* STt, int* act, hipStream_t st, int8_t* Si8, int* eSb, int* actK) {
  const size_t n = (size_t)Pc * nGp;
  if (Si8)  // the int8 form: S^T with its digits (per line and 64-point block scales)
    hipLaunchKernelGGL(k_pair_src_i8, dim3((unsigned)(n / 1024)), dim3(256), 0, st, orbs, w, nG, nGp, nO, Pc, STt,
                       Si8, eSb, actK);
  else
    hipLaunchKernelGGL(k_pair_src_t, dim3((unsigned)(n / 256)), dim3(256), 0, st, orbs, w, nG, nGp, nO, Pc, STt,
                       act);
  if (!ST) return;
#define CALL_PS(NO)                                                                            \
  hipLaunchKernelGGL(k_pair_src<NO>, dim3((nGp + 255) / 256), dim3(256), 0, st, orbs, w, nG, nGp, \
                     Pc, ST)
  CODY_NO_SWITCH(nO, CALL_PS)
#undef CALL_PS
}

void launch_pair_src_rows(const double2* orbs, const double* w, int nG, int nGp, int nO, int Pc, double* ST,
                          hipStream_t st) {
#define CALL_PR(NO)                                                                                            \
  hipLaunchKernelGGL(k_pair_src<NO>, dim3((nGp + 255) / 256), dim3(256), 0, st, orbs, w, nG, nGp, \
                     Pc, ST)
  CODY_NO_SWITCH(nO, CALL_PR)
#undef CALL_PR
}

void launch_tile_src(const double2* src, int ns, int nG, int nGp, int Pc, double* STt, hipStream_t st) {
  hipLaunchKernelGGL(k_tile_src, dim3((unsigned)((size_t)Pc * nGp / 256)), dim3(256), 0, st, src, ns, nG, nGp, STt);
}

void launch_overlap_parts(const double2* orbs, const double* w, const double2* f, int nG, int nO, double2* part,
                          hipStream_t st) {
  dim3 grid((nG + 255) / 256, nO);
#define CALL_OV(NO) hipLaunchKernelGGL(k_overlap<NO>, grid, dim3(256), 0, st, orbs, w, f, nG, part)
  CODY_NO_SWITCH(nO, CALL_OV)
#undef CALL_OV
}

void launch_project_inplace(const double2* orbs, const double2* part, int nG, int nO, double2* f, hipStream_t st) {
  const int nb = (nG + 255) / 256;
#define CALL_PI(NO) hipLaunchKernelGGL(k_project_inplace<NO>, dim3(nb), dim3(256), 0, st, orbs, part, nb, nG, f)
  CODY_NO_SWITCH(nO, CALL_PI)
#undef CALL_PI
}

void launch_onebody(const double2* orbs, const double* w, const double* x, const double* vext,
                    const KinLaunch& kl, const double* efield, int nG, int nO, double2* hphi, double2* part1,
                    hipStream_t st) {
  KinParams kp{kl.kref, kl.nE, kl.nLoc, kl.xminExcl, nG, kl.invh, kl.fac};
  dim3 grid((nG + 255) / 256, nO);
#define CALL_OB(NO) \
  hipLaunchKernelGGL(k_onebody<NO>, grid, dim3(256), 0, st, orbs, w, x, vext, kp, efield, nG, hphi, part1)
  CODY_NO_SWITCH(nO, CALL_OB)
#undef CALL_OB
}

void launch_xmat(const double2* orbs, const double* w, const double* x, int nG, int nO, double2* part1,
                 hipStream_t st) {
  dim3 grid((nG + 255) / 256, nO);
#define CALL_XM(NO) hipLaunchKernelGGL(k_xmat<NO>, grid, dim3(256), 0, st, orbs, w, x, nG, part1)
  CODY_NO_SWITCH(nO, CALL_XM)
#undef CALL_XM
}

void launch_rhs_qform(int nO, const double2* r2can, const double2* reginv, double* qf, const double* STt,
                      const double* WT, int nGp, int Pc, double* Gpart, int nxs, hipStream_t st, const int* act,
                      int actShift) {
  const unsigned grid = (unsigned)(corr_q_blocks(nO) + Pc / 16 * nxs);
#define CALL_QF(NO)                                                                                          \
  hipLaunchKernelGGL(k_rhs_qform<NO>, dim3(grid), dim3(256), 0, st, r2can, reginv, qf, STt, WT, nGp, Pc, Gpart, \
                     nxs, act, actShift)
  CODY_NO_SWITCH(nO, CALL_QF)
#undef CALL_QF
}

void launch_orbital(const double2* orbs, const double* w, const double2* hphi, const double* WT,
                    const double2* r2can, const double2* reginv, int nG, int nGp, int nO, int mode,
                    double2* draw, double2* part, hipStream_t st) {
  dim3 grid(nGp / 64, nO);
#define CALL_OR(NO)                                                                                  \
  hipLaunchKernelGGL(k_orbital<NO>, grid, dim3(256), 0, st, orbs, w, hphi, WT, r2can, reginv, nG, nGp, mode, \
                     draw, part)
  CODY_NO_SWITCH(nO, CALL_OR)
#undef CALL_OR
}

void launch_project(const double2* orbs, const double2* draw, const double2* part, int nPart, int nG, int nO,
                    int project, double2* dorbs, const double2* C, const double2* dCraw, int nC, double2* dC,
                    const double2* hsub, hipStream_t st) {
  const int nb = (nG + 255) / 256 + (dC ? 1 : 0);
#define CALL_PR(NO)                                                                                           \
  hipLaunchKernelGGL(k_project<NO>, dim3(nb), dim3(256), 0, st, orbs, draw, part, nPart, nG, project, dorbs, C, \
                     dCraw, nC, dC, hsub)
  CODY_NO_SWITCH(nO, CALL_PR)
#undef CALL_PR
}


// ------------------------------------------------------------------ fused right-hand-side launches
// The right-hand side runs on one queue as a handful of launches.  Pieces that do not depend on
// each other share a launch as workgroup roles (each role a contiguous range of workgroup
// indices), so a latency-bound piece costs no launch of its own and no cross-queue dependency.

// Front: from the state alone, the CI densities (rho1, rho2, canonical rho2), the one-body part
// (h phi, h1 partials) and the pair sources (S^T tiled with activity flags, or for the int8 form
// with its digits, block exponents and per-block flags when Si8 is set).
template <int NO>
__global__ __launch_bounds__(256) void k_rhs_front(const double2* __restrict__ C, int nC, DevCiTables t,
                                                   double2* __restrict__ rho1, double2* __restrict__ r2can,
                                                   double2* __restrict__ rho2, int nRdm,
                                                   const double2* __restrict__ orbs, const double* __restrict__ w,
                                                   const double* __restrict__ xg, const double* __restrict__ vext,
                                                   KinParams kp, const double* __restrict__ efield, int nG,
                                                   double2* __restrict__ hphi, double2* __restrict__ part1, int nObx,
                                                   int nGp, int Pc, double* __restrict__ STt,
                                                   int* __restrict__ act, int8_t* __restrict__ Si8,
                                                   int* __restrict__ eSb, int* __restrict__ actK) {
  int b = blockIdx.x;
  if (b < nRdm) {
    ci_rdm_body(b, C, nC, NO, t.r1_ptr, t.r1, t.r2_ptr, t.r2, rho1, r2can, rho2);
    return;
  }
  b -= nRdm;
  if (b < nObx * NO) {
    onebody_body<NO>(b % nObx, b / nObx, orbs, w, xg, vext, kp, efield, nG, hphi, part1);
    return;
  }
  b -= nObx * NO;
  if (Si8)
    pair_src_i8_body(b, orbs, w, nG, nGp, NO, Pc, STt, Si8, eSb, actK);
  else
    pair_src_t_body(b, orbs, w, nG, nGp, NO, Pc, STt, act);
}

// After the contraction: R = rho1 (rho1^2 + eps)^{-1} (workgroup 0 when `first` is 1, eliminated
// by its first wave) and the k-split sum of W^T (the others; none when the contraction wrote W^T
// itself).
template <int NO>
__global__ __launch_bounds__(256) void k_rhs_wsum(const double2* __restrict__ rho1, double reg,
                                                  double2* __restrict__ reginv, const double* __restrict__ Wpart,
                                                  int nsplit, size_t plane, double* __restrict__ WT, int first) {
  if (blockIdx.x < first) {
    // every thread of the workgroup: the elimination's row updates (2 nO^2 entries per pivot) take
    // one round instead of up to five on one wave (the same arithmetic per entry either way)
    reginv_body<NO>(rho1, reg, reginv, blockDim.x);
    return;
  }
  wsum_body(blockIdx.x - first, Wpart, nsplit, plane, WT);
}

// After W: the Gram matrix's split-K partials (G = S^T W) and the orbital correlation with its
// projector-overlap partials, side by side (both only read W^T).
template <int NO>
__global__ __launch_bounds__(256) void k_rhs_mid(const double* __restrict__ STt, const double* __restrict__ WT,
                                                 int nGp, int Pc, double* __restrict__ Gpart, int nxs,
                                                 const double2* __restrict__ orbs, const double* __restrict__ w,
                                                 const double2* __restrict__ hphi, const double2* __restrict__ r2can,
                                                 const double2* __restrict__ reginv, int nG, int mode,
                                                 double2* __restrict__ draw, double2* __restrict__ part,
                                                 const double* __restrict__ qf, const int* __restrict__ act,
                                                 int actShift) {
  constexpr int NT = (NO * (NO + 1) + 15) / 16;  // = Pc / 16
  int b = blockIdx.x;
  if constexpr (NO >= kCorrQMinOrbs) {  // (the Gram partials come from the Q launch)
    const int nOrbX = nGp / 64;
    corr_q_body<NO>(b % nOrbX, b / nOrbX, orbs, w, hphi, WT, qf, nG, nGp, mode, draw, part);
  } else {
    const int nGram = NT * nxs;
    if (b < nGram) {
      gram_f64_body<NT>(b % NT, b / NT, STt, nGp >> 4, WT, nGp, Gpart, Pc, nGp / (4 * nxs), (size_t)Pc * Pc, act,
                        actShift);
      return;
    }
    b -= nGram;
    const int nOrbX = nGp / 64;
    orbital_body<NO>(b % nOrbX, b / nOrbX, orbs, w, hphi, WT, r2can, reginv, nG, nGp, mode, draw, part);
  }
}

// After the fused int8 contraction (which formed c_m(x) and its overlap partials), with 1024
// threads per workgroup:
//   - G = S^T W directly, one 16 x 16 tile (R, T) per workgroup (the first NT^2): 16 waves take
//     every 16th 16-point block of the active 64-point blocks of S's row tile R (actK, or every
//     block), four blocks' loads issued before their MFMAs; the wave tiles are added in LDS in
//     fixed order, G[c1 Pc + c2] written, and the tile's energy share sum G(c1, c2) Q(c1, c2) into
//     epart[tile] (Q = energy_g_coef);
//   - h1 from its partials and its energy share sum h1(k, l) rho1(l, k) into epart[NT^2];
//   - the fixed-order sums of the contraction's nPart overlap partials A(i, j) and B(i, m) into
//     partr[0, n2) and partr[n2, 2 n2) (one workgroup each);
//   - R = rho1 (rho1^2 + eps)^{-1} (one wave).
// E = sum of the NT^2 + 1 energy shares (gram_direct_energy_parts); the CI apply reads them.
template <int NO>
__device__ __forceinline__ void gram_tile_body(int tile, const double* __restrict__ STt, const double* __restrict__ WT,
                                               int nGp, int Pc, const int* __restrict__ actK,
                                               const double2* __restrict__ rho2, double* __restrict__ G,
                                               double* __restrict__ epart) {
  constexpr int NT = (NO * (NO + 1) + 15) / 16, NW = 16, U = 4;
  __shared__ int sList[kR16MaxBlocks64];
  __shared__ int sCount;
  __shared__ double sAcc[NW][16][17];
  __shared__ double sRed[4];
  const int R = tile / NT, T = tile % NT;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, rr = lane & 15;
  const int nK64 = nGp >> 6, nK = nGp >> 4;
  if (actK) {  // the active 64-point blocks of row tile R of S^T, in increasing order (wave 0)
    if (wave == 0) {
      const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
      int n = 0;
      for (int base = 0; base < nK64; base += 64) {
        const int b = base + lane;
        const int any = b < nK64 ? actK[(size_t)R * nK64 + b] : 0;
        const unsigned long long m = __ballot(any != 0);
        if (any) sList[n + __popcll(m & below)] = b;
        n += __popcll(m);
      }
      if (lane == 0) sCount = n;
    }
    __syncthreads();
  }
  const int nSub = 4 * (actK ? sCount : nK64);  // 16-point blocks to contract
  const double* ap = STt + (size_t)R * nK * 256 + 2 * lane;
  const double* bp = WT + (size_t)(16 * T + rr) * nGp + 4 * q;
  d4_t acc = (d4_t){0.0, 0.0, 0.0, 0.0};
  for (int j0 = wave; j0 < nSub; j0 += NW * U) {
    d4_t a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + NW * u;
      if (j < nSub) {
        const int k16 = 4 * (actK ? sList[j >> 2] : (j >> 2)) + (j & 3);
        const d2_t lo = *reinterpret_cast<const d2_t*>(ap + (size_t)k16 * 256);
        const d2_t hi = *reinterpret_cast<const d2_t*>(ap + (size_t)k16 * 256 + 128);
        a[u] = (d4_t){lo[0], lo[1], hi[0], hi[1]};
        b[u] = *reinterpret_cast<const d4_t*>(bp + 16 * k16);
      } else {
        a[u] = (d4_t){0.0, 0.0, 0.0, 0.0};
        b[u] = a[u];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][s], b[u][s], acc, 0, 0, 0);
  }
  // FP64 result layout: element i of lane l is row 4 i + l / 16 (c1), column l % 16 (c2)
#pragma unroll
  for (int i = 0; i < 4; ++i) sAcc[wave][4 * i + q][rr] = acc[i];
  __syncthreads();
  double e = 0.0;
  if (threadIdx.x < 256) {
    const int row = threadIdx.x >> 4, col = threadIdx.x & 15;
    double v = sAcc[0][row][col];
#pragma unroll
    for (int w = 1; w < NW; ++w) v += sAcc[w][row][col];
    const int c1 = 16 * R + row, c2 = 16 * T + col;
    G[c1 * Pc + c2] = v;
    e = v * energy_g_coef(c1, c2, NO, rho2);
    e = wave_sum(e);
    if (lane == 0) sRed[wave] = e;
  }
  __syncthreads();
  if (threadIdx.x == 0) epart[tile] = (sRed[0] + sRed[1]) + (sRed[2] + sRed[3]);
}

// h1 from its nPart1 partials and its energy share sum h1(k, l) rho1(l, k), in fixed order.
template <int NO>
__device__ __forceinline__ void h1_energy_body(const double2* __restrict__ part1, int nPart1, double2* __restrict__ h1,
                                               const double2* __restrict__ rho1, double* __restrict__ eslot) {
  constexpr int n2 = NO * NO;
  static_assert(n2 <= 256, "one entry per thread of the first four waves");
  __shared__ double sRed[4];
  double e = 0.0;
  if (threadIdx.x < 256) {
    if (threadIdx.x < n2) {
      const int q = threadIdx.x;
      double2 acc = cmk(0, 0);
#pragma unroll 8
      for (int b = 0; b < nPart1; ++b) acc = cadd(acc, part1[(size_t)b * n2 + q]);
      h1[q] = acc;
      const int k = q % NO, l = q / NO;
      const double2 r = rho1[l + k * NO];
      e = acc.x * r.x - acc.y * r.y;
    }
    e = wave_sum(e);
    if ((threadIdx.x & 63) == 0) sRed[threadIdx.x >> 6] = e;
  }
  __syncthreads();
  if (threadIdx.x == 0) *eslot = (sRed[0] + sRed[1]) + (sRed[2] + sRed[3]);
}

template <int NO>
__device__ __forceinline__ void gram_direct_body(int b, const double* __restrict__ STt, const double* __restrict__ WT,
                                                 int nGp, int Pc, const int* __restrict__ actK,
                                                 const double2* __restrict__ rho2, double* __restrict__ G,
                                                 double* __restrict__ epart, const double2* __restrict__ part1,
                                                 int nPart1, double2* __restrict__ h1,
                                                 const double2* __restrict__ part, int nPart,
                                                 double2* __restrict__ partr, const double2* __restrict__ rho1,
                                                 double reg, double2* __restrict__ reginv) {
  // (the overlap sums take all 1024 threads: 1024 / nO^2 groups of threads)
  constexpr int NT = (NO * (NO + 1) + 15) / 16, n2 = NO * NO, G2 = n2 >= 1024 ? 1 : 1024 / n2;
  if (b < NT * NT) {
    gram_tile_body<NO>(b, STt, WT, nGp, Pc, actK, rho2, G, epart);
    return;
  }
  b -= NT * NT;
  if (b == 0) {
    h1_energy_body<NO>(part1, nPart1, h1, rho1, epart + NT * NT);
    return;
  }
  if (b == 3) {
    // one wave: a wave that has ended no longer counts toward the workgroup barrier on CDNA, so
    // the remaining wave's barriers wait for nobody else
    if (threadIdx.x >= 64) return;
    reginv_body<NO>(rho1, reg, reginv, 64);
    return;
  }
  __shared__ double2 sG[G2][n2];
  __shared__ double2 sInp[n2];
  sum_overlap_parts<NO, G2>(part + (size_t)(b - 1) * nPart * n2, nPart, sG, sInp);
  for (int i = threadIdx.x; i < n2; i += blockDim.x) partr[(b - 1) * n2 + i] = sInp[i];
}

template <int NO>
__global__ __launch_bounds__(1024) void k_rhs_gram(const double* __restrict__ STt, const double* __restrict__ WT,
                                                   int nGp, int Pc, const int* __restrict__ actK,
                                                   const double2* __restrict__ rho2, double* __restrict__ G,
                                                   double* __restrict__ epart, const double2* __restrict__ part1,
                                                   int nPart1, double2* __restrict__ h1,
                                                   const double2* __restrict__ part, int nPart,
                                                   double2* __restrict__ partr, const double2* __restrict__ rho1,
                                                   double reg, double2* __restrict__ reginv) {
  gram_direct_body<NO>(blockIdx.x, STt, WT, nGp, Pc, actK, rho2, G, epart, part1, nPart1, h1, part, nPart, partr,
                       rho1, reg, reginv);
}

// The fused int8 path's last launch: the CI apply and projection (the first nCi workgroups, with
// E from the nE energy shares) beside the orbital projection with the R mixing
// (project_corr_body), which does not depend on G.
template <int NO>
__global__ __launch_bounds__(256) void k_rhs_final(const double2* __restrict__ C, int nC, int Pc,
                                                   const double* __restrict__ G, const double2* __restrict__ h1,
                                                   DevCiTables t, const double* __restrict__ epart, int nE,
                                                   double2* __restrict__ dCraw, double2* __restrict__ dC, int nCi,
                                                   const double2* __restrict__ orbs, const double2* __restrict__ h,
                                                   const double2* __restrict__ corrc,
                                                   const double2* __restrict__ reginv,
                                                   const double2* __restrict__ partr, int nG,
                                                   double2* __restrict__ dorbs, const double2* __restrict__ hsub) {
  const int b = blockIdx.x;
  if (b < nCi) {
    ci_apply_project_body(b, C, nC, NO, Pc, G, h1, t, epart, nE, dCraw, dC);
    return;
  }
  project_corr_body<NO>(b - nCi, orbs, h, corrc, reginv, partr, nG, dorbs, hsub);
}

// Then: G and h1 from their partials with the energy partials (the first nRed workgroups), and the
// orbital projection with the right-hand side's factor -i (the others).
template <int NO>
__global__ __launch_bounds__(256) void k_rhs_post(const double* __restrict__ Gpart, int nxs, int Pc,
                                                  double* __restrict__ G, const double2* __restrict__ part1,
                                                  int nPart1, double2* __restrict__ h1,
                                                  const double2* __restrict__ rho1, const double2* __restrict__ rho2,
                                                  double* __restrict__ epart, int nRed,
                                                  const double2* __restrict__ orbs, const double2* __restrict__ draw,
                                                  const double2* __restrict__ part, int nPart, int nG,
                                                  double2* __restrict__ dorbs, const double2* __restrict__ hsub) {
  int b = blockIdx.x;
  if (b < nRed) {
    gram_reduce_energy_body(b, Gpart, nxs, Pc, G, part1, nPart1, NO, h1, rho1, rho2, epart);
    return;
  }
  project_body<NO>(b - nRed, false, orbs, draw, part, nPart, nG, 1, dorbs, nullptr, nullptr, 0, nullptr, hsub);
}

// ------------------------------------------------------------------ grouped launches
// The fused launches above over several instances of one orbital count: workgroup b runs block
// map[b].y of instance map[b].x with that instance's arguments, exactly as that instance's own
// launch would (same bodies, same block decomposition, so the results are bit-identical).
template <int NO>
__global__ __launch_bounds__(256) void k_grp_front(const GroupInstance* __restrict__ gi, const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  int b = m.y;
  if (b < a.nRdm) {
    ci_rdm_body(b, a.C, a.nC, NO, a.t.r1_ptr, a.t.r1, a.t.r2_ptr, a.t.r2, a.rho1, a.r2can, a.rho2);
    return;
  }
  b -= a.nRdm;
  const double2* orbs = a.C + a.nC;
  if (b < a.nObx * NO) {
    const KinParams kp{a.kin.kref, a.kin.nE, a.kin.nLoc, a.kin.xminExcl, a.nG, a.kin.invh, a.kin.fac};
    onebody_body<NO>(b % a.nObx, b / a.nObx, orbs, a.w, a.x, a.vext, kp, a.efield, a.nG, a.hphi, a.part1);
    return;
  }
  b -= a.nObx * NO;
  if (a.Si8)
    pair_src_i8_body(b, orbs, a.w, a.nG, a.nGp, NO, a.Pc, a.STt, a.Si8, a.eSb, a.actK);
  else
    pair_src_t_body(b, orbs, a.w, a.nG, a.nGp, NO, a.Pc, a.STt, a.act);
}

template <int NO>
__global__ __launch_bounds__(256) void k_grp_wsum(const GroupInstance* __restrict__ gi, const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  if (m.y < a.wsumFirst) {  // (every thread, as in k_rhs_wsum)
    reginv_body<NO>(a.rho1, a.reg, a.reginv, blockDim.x);
    return;
  }
  wsum_body(m.y - a.wsumFirst, a.Wpart, a.nsplit, (size_t)a.Pc * a.nGp, a.WT);
}

// Q of the orbital correlation (nO >= kCorrQMinOrbs), after R: as k_rhs_qform.
template <int NO>
__global__ __launch_bounds__(256) void k_grp_qform(const GroupInstance* __restrict__ gi, const int2* __restrict__ map) {
  if constexpr (NO >= kCorrQMinOrbs) {
    constexpr int NT = (NO * (NO + 1) + 15) / 16, nQ = corr_q_blocks(NO);
    const int2 m = map[blockIdx.x];
    const GroupInstance& a = gi[m.x];
    if (m.y < nQ) {
      qform_body<NO>(m.y, a.r2can, a.reginv, a.qf);
      return;
    }
    const int b = m.y - nQ;
    gram_f64_body<NT>(b % NT, b / NT, a.STt, a.nGp >> 4, a.WT, a.nGp, a.Gpart, a.Pc, a.nGp / (4 * a.nxs),
                      (size_t)a.Pc * a.Pc, a.actC, a.Si8 ? 6 : 4);
  }
}

template <int NO>
__global__ __launch_bounds__(256) void k_grp_mid(const GroupInstance* __restrict__ gi, const int2* __restrict__ map) {
  constexpr int NT = (NO * (NO + 1) + 15) / 16;
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  int b = m.y;
  if constexpr (NO >= kCorrQMinOrbs) {  // (the Gram partials come from the Q launch)
    const int nOrbX = a.nGp / 64;
    corr_q_body<NO>(b % nOrbX, b / nOrbX, a.C + a.nC, a.w, a.hphi, a.WT, a.qf, a.nG, a.nGp, kModeFull, a.draw, a.part);
  } else {
    const int nGram = NT * a.nxs;
    if (b < nGram) {
      gram_f64_body<NT>(b % NT, b / NT, a.STt, a.nGp >> 4, a.WT, a.nGp, a.Gpart, a.Pc, a.nGp / (4 * a.nxs),
                        (size_t)a.Pc * a.Pc, a.actC, a.Si8 ? 6 : 4);
      return;
    }
    b -= nGram;
    const int nOrbX = a.nGp / 64;
    orbital_body<NO>(b % nOrbX, b / nOrbX, a.C + a.nC, a.w, a.hphi, a.WT, a.r2can, a.reginv, a.nG, a.nGp, kModeFull,
                     a.draw, a.part);
  }
}

template <int NO>
__global__ __launch_bounds__(1024) void k_grp_gram(const GroupInstance* __restrict__ gi, const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  gram_direct_body<NO>(m.y, a.STt, a.WT, a.nGp, a.Pc, a.actC, a.rho2, a.G, a.epart, a.part1, a.nPart1, a.h1, a.part,
                       a.nGp / 16, a.partr, a.rho1, a.reg, a.reginv);
}

// the fused int8 groups' last launch (as k_rhs_final)
template <int NO>
__global__ __launch_bounds__(256) void k_grp_final(const GroupInstance* __restrict__ gi, const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  const int nCi = (a.nC + 3) / 4;
  if (m.y < nCi) {
    ci_apply_project_body(m.y, a.C, a.nC, NO, a.Pc, a.G, a.h1, a.t, a.epart, a.nRed, a.dCraw, a.dC);
    return;
  }
  project_corr_body<NO>(m.y - nCi, a.C + a.nC, a.hphi, a.draw, a.reginv, a.partr, a.nG, a.dC + a.nC, nullptr);
}

template <int NO>
__global__ __launch_bounds__(256) void k_grp_post(const GroupInstance* __restrict__ gi, const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  if (m.y < a.nRed) {
    gram_reduce_energy_body(m.y, a.Gpart, a.nxs, a.Pc, a.G, a.part1, a.nPart1, NO, a.h1, a.rho1, a.rho2, a.epart);
    return;
  }
  project_body<NO>(m.y - a.nRed, false, a.C + a.nC, a.draw, a.part, a.nPart, a.nG, 1, a.dC + a.nC, nullptr, nullptr, 0,
                   nullptr, nullptr);
}

void launch_grp_front(int nO, const GroupInstance* gi, const int2* map, int nBlocks, int maxC, hipStream_t st) {
#define CALL_GF(NO) \
  hipLaunchKernelGGL(k_grp_front<NO>, dim3(nBlocks), dim3(256), sizeof(double2) * maxC, st, gi, map)
  CODY_NO_SWITCH(nO, CALL_GF)
#undef CALL_GF
}
void launch_grp_wsum(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st) {
  if (nBlocks == 0) return;
#define CALL_GW(NO) hipLaunchKernelGGL(k_grp_wsum<NO>, dim3(nBlocks), dim3(256), 0, st, gi, map)
  CODY_NO_SWITCH(nO, CALL_GW)
#undef CALL_GW
}
void launch_grp_qform(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st) {
  if (nBlocks == 0) return;
#define CALL_GR(NO) hipLaunchKernelGGL(k_grp_qform<NO>, dim3(nBlocks), dim3(256), 0, st, gi, map)
  CODY_NO_SWITCH(nO, CALL_GR)
#undef CALL_GR
}
void launch_grp_mid(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st) {
#define CALL_GM(NO) hipLaunchKernelGGL(k_grp_mid<NO>, dim3(nBlocks), dim3(256), 0, st, gi, map)
  CODY_NO_SWITCH(nO, CALL_GM)
#undef CALL_GM
}
void launch_grp_final(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st) {
#define CALL_GFI(NO) hipLaunchKernelGGL(k_grp_final<NO>, dim3(nBlocks), dim3(256), 0, st, gi, map)
  CODY_NO_SWITCH(nO, CALL_GFI)
#undef CALL_GFI
}
void launch_grp_gram(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st) {
#define CALL_GG(NO) hipLaunchKernelGGL(k_grp_gram<NO>, dim3(nBlocks), dim3(1024), 0, st, gi, map)
  CODY_NO_SWITCH(nO, CALL_GG)
#undef CALL_GG
}
void launch_grp_post(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st) {
#define CALL_GP(NO) hipLaunchKernelGGL(k_grp_post<NO>, dim3(nBlocks), dim3(256), 0, st, gi, map)
  CODY_NO_SWITCH(nO, CALL_GP)
#undef CALL_GP
}

void launch_rhs_front(const double2* C, int nC, int nO, const DevCiTables& t, double2* rho1, double2* r2can,
                      double2* rho2, const double2* orbs, const double* w, const double* x, const double* vext,
                      const KinLaunch& kl, const double* efield, int nG, double2* hphi, double2* part1, int nGp,
                      int Pc, double* STt, int* act, int8_t* Si8, int* eSb, int* actK, hipStream_t st) {
  const int n2 = nO * nO, np = nO * (nO - 1) / 2, per = 4 * ci_rdm_items_per_wave(nC);
  const int nRdm = (n2 + np * np + per - 1) / per;
  const int nObx = (nG + 255) / 256, nPair = (int)((size_t)Pc * nGp / (Si8 ? 1024 : 256));
  KinParams kp{kl.kref, kl.nE, kl.nLoc, kl.xminExcl, nG, kl.invh, kl.fac};
  const unsigned grid = (unsigned)(nRdm + nObx * nO + nPair);
#define CALL_FR(NO)                                                                                             \
  hipLaunchKernelGGL(k_rhs_front<NO>, dim3(grid), dim3(256), sizeof(double2) * nC, st, C, nC, t, rho1, r2can, rho2, \
                     nRdm, orbs, w, x, vext, kp, efield, nG, hphi, part1, nObx, nGp, Pc, STt, act, Si8, eSb, actK)
  CODY_NO_SWITCH(nO, CALL_FR)
#undef CALL_FR
}

void launch_rhs_wsum(const double2* rho1, int nO, double reg, double2* reginv, const double* Wpart, int nsplit,
                     size_t plane, double* WT, hipStream_t st, bool withReginv) {
  const unsigned nSum = nsplit > 1 ? (unsigned)((plane / 2 + 255) / 256) : 0u;
  const int first = withReginv ? 1 : 0;
  if (nSum + first == 0) return;
#define CALL_WS(NO)                                                                                        \
  hipLaunchKernelGGL(k_rhs_wsum<NO>, dim3(first + nSum), dim3(256), 0, st, rho1, reg, reginv, Wpart, nsplit, \
                     plane, WT, first)
  CODY_NO_SWITCH(nO, CALL_WS)
#undef CALL_WS
}

void launch_rhs_mid(const double* STt, const double* WT, int nGp, int Pc, double* Gpart, int nxs,
                    const double2* orbs, const double* w, const double2* hphi, const double2* r2can,
                    const double2* reginv, int nG, int nO, int mode, double2* draw, double2* part, hipStream_t st,
                    const double* qf, const int* act, int actShift) {
  const unsigned grid = (unsigned)(nO >= kCorrQMinOrbs ? nGp / 64 * kCorrQGroups : Pc / 16 * nxs + nGp / 64 * nO);
#define CALL_MI(NO)                                                                                             \
  hipLaunchKernelGGL(k_rhs_mid<NO>, dim3(grid), dim3(256), 0, st, STt, WT, nGp, Pc, Gpart, nxs, orbs, w, hphi, r2can, \
                     reginv, nG, mode, draw, part, qf, act, actShift)
  CODY_NO_SWITCH(nO, CALL_MI)
#undef CALL_MI
}

int gram_direct_energy_parts(int Pc) { return (Pc / 16) * (Pc / 16) + 1; }

void launch_rhs_gram(const double* STt, const double* WT, int nGp, int Pc, int nO, const int* actK,
                     const double2* rho2, double* G, double* epart, const double2* part1, int nPart1, double2* h1,
                     const double2* part, int nPart, double2* partr, const double2* rho1, double reg,
                     double2* reginv, hipStream_t st) {
  const unsigned grid = (unsigned)((Pc / 16) * (Pc / 16) + 4);
#define CALL_RG(NO)                                                                                          \
  hipLaunchKernelGGL(k_rhs_gram<NO>, dim3(grid), dim3(1024), 0, st, STt, WT, nGp, Pc, actK, rho2, G, epart, part1, \
                     nPart1, h1, part, nPart, partr, rho1, reg, reginv)
  CODY_NO_SWITCH(nO, CALL_RG)
#undef CALL_RG
}

void launch_rhs_final(const double2* C, int nC, int nO, int Pc, const double* G, const double2* h1,
                      const DevCiTables& t, const double* epart, int nE, double2* dCraw, double2* dC,
                      const double2* orbs, const double2* h, const double2* corrc, const double2* reginv,
                      const double2* partr, int nG, double2* dorbs, const double2* hsub, hipStream_t st) {
  const int nCi = (nC + 3) / 4;
  const unsigned grid = (unsigned)(nCi + (nG + 255) / 256);
#define CALL_FI(NO)                                                                                          \
  hipLaunchKernelGGL(k_rhs_final<NO>, dim3(grid), dim3(256), 0, st, C, nC, Pc, G, h1, t, epart, nE, dCraw, dC, nCi, \
                     orbs, h, corrc, reginv, partr, nG, dorbs, hsub)
  CODY_NO_SWITCH(nO, CALL_FI)
#undef CALL_FI
}

int rhs_post_energy_parts(int Pc, int nO) { return (Pc * Pc + nO * nO + 255) / 256; }

void launch_rhs_post(const double* Gpart, int nxs, int Pc, double* G, const double2* part1, int nPart1,
                     double2* h1, const double2* rho1, const double2* rho2, double* epart, const double2* orbs,
                     const double2* draw, const double2* part, int nPart, int nG, int nO, double2* dorbs,
                     const double2* hsub, hipStream_t st) {
  const int nRed = rhs_post_energy_parts(Pc, nO);
  const unsigned grid = (unsigned)(nRed + (nG + 255) / 256);
#define CALL_PO(NO)                                                                                                \
  hipLaunchKernelGGL(k_rhs_post<NO>, dim3(grid), dim3(256), 0, st, Gpart, nxs, Pc, G, part1, nPart1, h1, rho1, rho2, \
                     epart, nRed, orbs, draw, part, nPart, nG, dorbs, hsub)
  CODY_NO_SWITCH(nO, CALL_PO)
#undef CALL_PO
}


}  // namespace cody
