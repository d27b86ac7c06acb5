// Device helpers shared by the right-hand-side and propagator kernels: complex arithmetic on
// double2 (x = re, y = im), wavefront sums, and the FEDVR kinetic operator at one grid point.
#pragma once

#include <hip/hip_runtime.h>

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
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return cmk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// ------------------------------------------------------------------ orbital index helpers
// column-major index of a 4-index tensor of side n
__device__ __forceinline__ int idx4(int a, int b, int c, int d, int n) { return a + n * (b + n * (c + n * d)); }
// unordered pair l <= k among the nO (nO + 1) / 2 pairs
__device__ __forceinline__ int pair_index(int nO, int l, int k) { return l * nO - (l * (l - 1)) / 2 + (k - l); }
// strict pair a < b among the nO (nO - 1) / 2 strict pairs, and its inverse
__device__ __forceinline__ int strict_pair(int nO, int a, int b) { return a * nO - (a * (a + 1)) / 2 + (b - a - 1); }
__device__ __forceinline__ void strict_pair_decode(int nO, int p, int& a, int& b) {
  a = 0;
  while (p >= nO - 1 - a) {
    p -= nO - 1 - a;
    ++a;
  }
  b = a + 1 + p;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// NS sums over the 256 threads of a workgroup, each thread contributing v[s] to sum s, in rounds
// of at most 8 sums: the round's values go to LDS (sBuf, 8 x 256 doubles) once, each group of 16
// lanes adds 16 of one sum's values in order, and the 16 group partials meet in a 4-step
// butterfly.  The result lands in res[s] (LDS).  About ten LDS latencies per round, against six
// dependent cross-lane steps per sum for NS wave sums.  Fixed order (bit-reproducible).  Every
// thread calls it.  Only the first nUsed (uniform, <= NS) sums are formed.
template <int NS>
__device__ __forceinline__ void block_sums256(const double (&v)[NS], double* sBuf, double* res, int nUsed = NS) {
  const int ns = nUsed < NS ? nUsed : NS;
#pragma unroll
  for (int s0 = 0; s0 < NS; s0 += 8) {
    if (s0 >= ns) break;
    constexpr int kRound = 8;
#pragma unroll
    for (int s = 0; s < kRound; ++s)
      if (s0 + s < NS) sBuf[s * 256 + threadIdx.x] = v[s0 + s];
    __syncthreads();
    const int n = ns - s0 < kRound ? ns - s0 : kRound;
    if (threadIdx.x < n * 16) {
      const int s = threadIdx.x >> 4, q = threadIdx.x & 15;
      const double* b = sBuf + s * 256 + 16 * q;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) acc += b[k];
#pragma unroll
      for (int off = 8; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if (q == 0) res[s0 + s] = acc;
    }
    __syncthreads();
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

// kinetic_at with f staged in LDS: sF[j - xlo] holds f at grid index j, for every j that the
// stencil of x reaches ([x - (nLoc - 1), x + nLoc - 1] inside [0, nG)).  Same operations in the
// same order as kinetic_at.
__device__ __forceinline__ double2 kinetic_at_lds(const double2* sF, int xlo, int x, const KinParams& kp,
                                                  double rw) {
  const int gi = x + kp.xminExcl;
  const int per = kp.nLoc - 1;
  double2 acc = cmk(0, 0);
  int e = gi / per, a = gi % per;
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
      const double2 fj = sF[j - xlo];
      acc.x = fma(kab, fj.x, acc.x);
      acc.y = fma(kab, fj.y, acc.y);
    }
  }
  return cscale(acc, kp.invh * rw * kp.fac);
}

// kinetic_at_lds with the reference-element matrix staged in LDS too: sK[a + b nLoc] = K(a, b).
// Every operand of the stencil then comes from LDS (a global K(a, b) load per neighbour, each
// waited for before the next, would dominate the one-body kernel's time).  Same operations in the
// same order as kinetic_at.
__device__ __forceinline__ double2 kinetic_at_lds_k(const double2* sF, int xlo, int x, const KinParams& kp,
                                                    const double* sK, double rw) {
  const int gi = x + kp.xminExcl;
  const int per = kp.nLoc - 1;
  double2 acc = cmk(0, 0);
  int e = gi / per, a = gi % per;
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
      const double kab = sK[aa + b * kp.nLoc];
      const double2 fj = sF[j - xlo];
      acc.x = fma(kab, fj.x, acc.x);
      acc.y = fma(kab, fj.y, acc.y);
    }
  }
  return cscale(acc, kp.invh * rw * kp.fac);
}

}  // namespace cody
