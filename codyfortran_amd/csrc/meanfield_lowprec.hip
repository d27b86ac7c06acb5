// Reduced-precision forms of the mean-field contraction W = V S (gfx950 MFMA).
//
//  * int8-sliced (exact products): every row of V and every column of S is scaled by a power of
//    two into [-1/2, 1/2] and split into NS signed 7-bit slices, x = 2^e sum_p X_p 2^{-7p}
//    (round-to-nearest residuals, |X_p| <= 64).  Slice products run on v_mfma_i32_16x16x64_i8 with
//    exact int32 accumulation over the whole contraction (|sum| < 2^27); all products with
//    p + q <= NS + 1 that share the scale 2^{-7(p+q)} are summed in one int32 accumulator, and
//    the diagonals are recombined in FP64.  The only error is the slice truncation.
//  * split-FP32: x = hi + lo with hi = fl32(x), lo = fl32(x - hi); hi*hi + hi*lo + lo*hi on
//    v_mfma_f32_16x16x4f32, FP32 accumulation over 32-wide k chunks flushed into FP64.
//  * 3xBF16: x = b1 + b2 + b3 (successive BF16 roundings); the six products with p + q <= 4 on
//    v_mfma_f32_16x16x32_bf16, FP32 accumulation over 64-wide k chunks flushed into FP64.
// Output layout as the FP64 kernel: Wpart[split][row][col], FP64.
#include <hip/hip_runtime.h>

#include "mctdhx_kernels.hpp"

namespace cody {

typedef int i4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ slicing
// One workgroup per row (V) or per column of S (ST row): exponent e = ceil(log2 max|x|) + 1,
// slices written to out[p][line][k], p = 0..ns-1 (int8).
__global__ __launch_bounds__(256) void k_slice_rows(const double* __restrict__ X, int len, int nlines,
                                                    int ns, int8_t* __restrict__ out,
                                                    int* __restrict__ expo) {
  __shared__ double red[4];
  const int line = blockIdx.x;
  const double* row = X + (size_t)line * len;
  double m = 0.0;
  for (int k = threadIdx.x; k < len; k += blockDim.x) m = fmax(m, fabs(row[k]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  int e = 0;
  if (m > 0.0) {
    int ex;
    const double fr = frexp(m, &ex);  // m = fr 2^ex, fr in [0.5, 1)
    e = (fr == 0.5 ? ex - 1 : ex) + 1;  // ceil(log2 m) + 1
  }
  if (threadIdx.x == 0) expo[line] = e;
  const size_t plane = (size_t)nlines * len;
  for (int k = threadIdx.x; k < len; k += blockDim.x) {
    double r = ldexp(row[k], -e);
    for (int p = 1; p <= ns; ++p) {
      const double q = rint(ldexp(r, 7 * p));
      out[(size_t)(p - 1) * plane + (size_t)line * len + k] = static_cast<int8_t>(q);
      r -= ldexp(q, -7 * p);
    }
  }
}

// ------------------------------------------------------------------ int8-sliced contraction
// Lane l of a wave holds row (l%16) of a 16-row tile and the 16 consecutive k of group l/16 of
// each 64-wide k block, for A = V slices and B = S slices alike, so each i8 MFMA contracts 64 k.
template <int NT, int NS>
__global__ __launch_bounds__(256) void k_meanfield_i8(const int8_t* __restrict__ As,
                                                      const int* __restrict__ eA,
                                                      const int8_t* __restrict__ Bs,
                                                      const int* __restrict__ fB,
                                                      double* __restrict__ Wpart, int nGp, int Pc,
                                                      int kchunk) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane >> 4, rr = lane & 15;
  const int row = blockIdx.x * 64 + wave * 16 + rr;
  const int split = blockIdx.y;
  const int k0 = split * kchunk;
  const size_t planeA = (size_t)nGp * nGp, planeB = (size_t)Pc * nGp;
  i4_t acc[NT][NS];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int d = 0; d < NS; ++d) acc[t][d] = (i4_t){0, 0, 0, 0};
  const int8_t* abase = As + (size_t)row * nGp + k0 + 16 * q;
  for (int kb = 0; kb < kchunk; kb += 64) {
    i4_t a[NS];
#pragma unroll
    for (int p = 0; p < NS; ++p) a[p] = *reinterpret_cast<const i4_t*>(abase + p * planeA + kb);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int8_t* bbase = Bs + (size_t)(t * 16 + rr) * nGp + k0 + 16 * q + kb;
      i4_t b[NS];
#pragma unroll
      for (int p = 0; p < NS; ++p) b[p] = *reinterpret_cast<const i4_t*>(bbase + p * planeB);
      // diagonal d = p + q - 2 (0-based slice indices p, q with p + q <= NS - 1)
#pragma unroll
      for (int d = 0; d < NS; ++d)
#pragma unroll
        for (int p = 0; p <= d; ++p)
          acc[t][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[p], b[d - p], acc[t][d], 0, 0, 0);
    }
  }
  double* out = Wpart + (size_t)split * nGp * Pc;
  const int orow = blockIdx.x * 64 + wave * 16 + 4 * q;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = t * 16 + rr;
    const int fc = fB[col];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double s = 0.0;
#pragma unroll
      for (int d = NS - 1; d >= 0; --d) s += ldexp(static_cast<double>(acc[t][d][i]), -7 * (d + 2));
      out[(size_t)(orow + i) * Pc + col] = ldexp(s, eA[orow + i] + fc);
    }
  }
}

// ------------------------------------------------------------------ split-FP32 contraction
__global__ __launch_bounds__(256) void k_split_f32(const double* __restrict__ X, size_t n,
                                                   float* __restrict__ hi, float* __restrict__ lo) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = X[i];
  const float h = static_cast<float>(x);
  hi[i] = h;
  lo[i] = static_cast<float>(x - static_cast<double>(h));
}

template <int NT>
__global__ __launch_bounds__(256) void k_meanfield_f32x3(const float* __restrict__ Ah,
                                                         const float* __restrict__ Al,
                                                         const float* __restrict__ Bh,
                                                         const float* __restrict__ Bl,
                                                         double* __restrict__ Wpart, int nGp, int Pc,
                                                         int kchunk) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane >> 4, rr = lane & 15;
  const int row = blockIdx.x * 64 + wave * 16 + rr;
  const int split = blockIdx.y;
  const int k0 = split * kchunk;
  double acc64[NT][4];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc64[t][i] = 0.0;
  const size_t aoff = (size_t)row * nGp + k0 + 4 * q;
  for (int kc = 0; kc < kchunk; kc += 32) {
    f4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 32; kb += 16) {
      const f4_t ah = *reinterpret_cast<const f4_t*>(Ah + aoff + kc + kb);
      const f4_t al = *reinterpret_cast<const f4_t*>(Al + aoff + kc + kb);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const size_t boff = (size_t)(t * 16 + rr) * nGp + k0 + 4 * q + kc + kb;
        const f4_t bh = *reinterpret_cast<const f4_t*>(Bh + boff);
        const f4_t bl = *reinterpret_cast<const f4_t*>(Bl + boff);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[s], bh[s], acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[s], bl[s], acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[s], bh[s], acc[t], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc64[t][i] += static_cast<double>(acc[t][i]);
  }
  double* out = Wpart + (size_t)split * nGp * Pc;
  const int orow = blockIdx.x * 64 + wave * 16 + 4 * q;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(size_t)(orow + i) * Pc + t * 16 + rr] = acc64[t][i];
}

// ------------------------------------------------------------------ 3xBF16 contraction
__global__ __launch_bounds__(256) void k_split_bf16x3(const double* __restrict__ X, size_t n,
                                                      __bf16* __restrict__ b1, __bf16* __restrict__ b2,
                                                      __bf16* __restrict__ b3) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = X[i];
  const __bf16 v1 = static_cast<__bf16>(static_cast<float>(r));
  r -= static_cast<double>(static_cast<float>(v1));
  const __bf16 v2 = static_cast<__bf16>(static_cast<float>(r));
  r -= static_cast<double>(static_cast<float>(v2));
  const __bf16 v3 = static_cast<__bf16>(static_cast<float>(r));
  b1[i] = v1;
  b2[i] = v2;
  b3[i] = v3;
}

template <int NT>
__global__ __launch_bounds__(256) void k_meanfield_bf16x3(const __bf16* __restrict__ A, const __bf16* __restrict__ B,
                                                          double* __restrict__ Wpart, int nGp, int Pc,
                                                          int kchunk) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane >> 4, rr = lane & 15;
  const int row = blockIdx.x * 64 + wave * 16 + rr;
  const int split = blockIdx.y;
  const int k0 = split * kchunk;
  const size_t planeA = (size_t)nGp * nGp, planeB = (size_t)Pc * nGp;
  double acc64[NT][4];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc64[t][i] = 0.0;
  const __bf16* abase = A + (size_t)row * nGp + k0 + 8 * q;
  for (int kc = 0; kc < kchunk; kc += 64) {
    f4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 64; kb += 32) {
      bf8_t a[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) a[p] = *reinterpret_cast<const bf8_t*>(abase + p * planeA + kc + kb);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const __bf16* bbase = B + (size_t)(t * 16 + rr) * nGp + k0 + 8 * q + kc + kb;
        bf8_t b[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) b[p] = *reinterpret_cast<const bf8_t*>(bbase + p * planeB);
        // small terms first
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc64[t][i] += static_cast<double>(acc[t][i]);
  }
  double* out = Wpart + (size_t)split * nGp * Pc;
  const int orow = blockIdx.x * 64 + wave * 16 + 4 * q;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(size_t)(orow + i) * Pc + t * 16 + rr] = acc64[t][i];
}

// ------------------------------------------------------------------ launchers
void launch_slice_rows(const double* X, int len, int nlines, int ns, int8_t* out, int* expo,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_slice_rows, dim3(nlines), dim3(256), 0, st, X, len, nlines, ns, out, expo);
}

void launch_split_f32(const double* X, size_t n, float* hi, float* lo, hipStream_t st) {
  hipLaunchKernelGGL(k_split_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X, n, hi, lo);
}

void launch_split_bf16x3(const double* X, size_t n, __bf16* out, hipStream_t st) {
  hipLaunchKernelGGL(k_split_bf16x3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X, n, out,
                     out + n, out + 2 * n);
}

#define CODY_NT_SWITCH(PC, CALL) \
  switch ((PC) / 16) {           \
    case 1: CALL(1); break;      \
    case 2: CALL(2); break;      \
    case 3: CALL(3); break;      \
    case 4: CALL(4); break;      \
    case 5: CALL(5); break;      \
    case 6: CALL(6); break;      \
    case 7: CALL(7); break;      \
    case 8: CALL(8); break;      \
    case 9: CALL(9); break;      \
    case 10: CALL(10); break;    \
    default: break;              \
  }

void launch_meanfield_i8(const int8_t* As, const int* eA, const int8_t* Bs, const int* fB,
                         double* Wpart, int nGp, int Pc, int nsplit, hipStream_t st) {
  dim3 grid(nGp / 64, nsplit);
  const int kchunk = nGp / nsplit;
#define CALL_I8(NT)                                                                                   \
  hipLaunchKernelGGL((k_meanfield_i8<NT, kInt8Slices>), grid, dim3(256), 0, st, As, eA, Bs, fB, Wpart, \
                     nGp, Pc, kchunk)
  CODY_NT_SWITCH(Pc, CALL_I8)
#undef CALL_I8
}

void launch_meanfield_f32x3(const float* Ah, const float* Al, const float* Bh, const float* Bl,
                            double* Wpart, int nGp, int Pc, int nsplit, hipStream_t st) {
  dim3 grid(nGp / 64, nsplit);
  const int kchunk = nGp / nsplit;
#define CALL_F32(NT)                                                                              \
  hipLaunchKernelGGL(k_meanfield_f32x3<NT>, grid, dim3(256), 0, st, Ah, Al, Bh, Bl, Wpart, nGp, Pc, \
                     kchunk)
  CODY_NT_SWITCH(Pc, CALL_F32)
#undef CALL_F32
}

void launch_meanfield_bf16x3(const __bf16* A, const __bf16* B, double* Wpart, int nGp, int Pc,
                             int nsplit, hipStream_t st) {
  dim3 grid(nGp / 64, nsplit);
  const int kchunk = nGp / nsplit;
#define CALL_BF(NT) \
  hipLaunchKernelGGL(k_meanfield_bf16x3<NT>, grid, dim3(256), 0, st, A, B, Wpart, nGp, Pc, kchunk)
  CODY_NT_SWITCH(Pc, CALL_BF)
#undef CALL_BF
}

}  // namespace cody
