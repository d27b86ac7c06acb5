// Operand preparation of the reduced-precision forms of the mean-field contraction W = V S
// (the contraction kernels themselves are in meanfield_gemm.hip).
//
//  * int8-sliced (exact products): every row of V and every column of S is scaled by a power of
//    two into [-1/2, 1/2] and split into NS balanced base-254 digits, x = 2^e sum_p X_p 254^{-p}
//    (t = 254 r, X_p = rint(t), r = t - X_p: |r| <= 1/2 and |X_p| <= 127 at every step).  Digit
//    products run on v_mfma_i32_16x16x64_i8 with exact int32 accumulation over the contraction
//    (|sum| <= NS 127^2 N_g < 2^31); all products with p + q <= NS + 1 that share the scale
//    254^{-(p+q)} are summed in one int32 accumulator, and the diagonals are recombined in FP64.
//    The only error is the digit truncation (relative ~NS 254^{-NS}).
//  * split-FP32: x = hi + lo with hi = fl32(x), lo = fl32(x - hi); hi*hi + hi*lo + lo*hi on
//    v_mfma_f32_16x16x4f32, FP32 accumulation over 32-wide k chunks flushed into FP64.
//  * 3xBF16: x = b1 + b2 + b3 (successive BF16 roundings); the six products with p + q <= 4 on
//    v_mfma_f32_16x16x32_bf16, FP32 accumulation over 64-wide k chunks flushed into FP64.
#include <hip/hip_runtime.h>

#include "gemm_device.hpp"
#include "mctdhx_kernels.hpp"

namespace cody {

typedef int i4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ slicing
// One workgroup per row (V) or per column of S (ST row): exponent e = ceil(log2 max|x|) + 1,
// slices written (int8) either line-major, out[p][line][k], or fragment-tiled for the MFMA A
// operand: slice p of the 16-line x 64-k block (R, K) is 1 KB contiguous at
// ((R nK64 + K) ns + p) 1024, lane l = 16 (k%64 / 16) + line%16 holding bytes k%16.
__global__ __launch_bounds__(1024) void k_slice_rows(const double* __restrict__ X, int len, int nlines,
                                                     int ns, int tiled, int8_t* __restrict__ out,
                                                     int* __restrict__ expo) {
  __shared__ double red[16];
  const int line = blockIdx.x, nw = blockDim.x >> 6;
  const double* row = X + (size_t)line * len;
  double m = 0.0;
  for (int k = threadIdx.x; k < len; k += blockDim.x) m = fmax(m, fabs(row[k]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  m = 0.0;
  for (int w = 0; w < nw; ++w) m = fmax(m, red[w]);
  int e = 0;
  if (m > 0.0) {
    int ex;
    const double fr = frexp(m, &ex);  // m = fr 2^ex, fr in [0.5, 1)
    e = (fr == 0.5 ? ex - 1 : ex) + 1;  // ceil(log2 m) + 1
  }
  if (threadIdx.x == 0) expo[line] = e;
  const size_t plane = (size_t)nlines * len;
  const int nK64 = len >> 6;
  // four consecutive k per thread: one 32-bit store per slice (len is a multiple of 64)
  for (int k = 4 * threadIdx.x; k < len; k += 4 * blockDim.x) {
    double r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) r[u] = ldexp(row[k + u], -e);
    size_t o;
    if (tiled) {
      const int lane = 16 * ((k & 63) >> 4) + (line & 15);
      o = ((size_t)(line >> 4) * nK64 + (k >> 6)) * ns * 1024 + 16 * lane + (k & 15);
    } else {
      o = (size_t)line * len + k;
    }
    for (int p = 1; p <= ns; ++p) {
      unsigned word = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double t = r[u] * kSliceBase;
        const double q = rint(t);
        r[u] = t - q;
        word |= (unsigned)(uint8_t)static_cast<int8_t>(q) << (8 * u);
      }
      *reinterpret_cast<unsigned*>(out + o + (tiled ? (size_t)(p - 1) * 1024 : (size_t)(p - 1) * plane)) = word;
    }
  }
}

// ------------------------------------------------------------------ FP64 tiling
// Vt[(R nK + K) 256 + 128 h + 2 l + e] = V[16 R + l % 16][16 K + 4 (l / 16) + 2 h + e]: every 16 x 16
// block of V becomes 2 KB of contiguous MFMA fragments (two 1 KB loads per wave).
__global__ __launch_bounds__(256) void k_tile_f64(const double* __restrict__ V, int nGp,
                                                  double* __restrict__ Vt) {
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= (size_t)nGp * nGp) return;
  const int nK = nGp >> 4;
  const int e = o & 1, l = (o >> 1) & 63, h = (o >> 7) & 1;
  const size_t blk = o >> 8;
  const int R = (int)(blk / nK), K = (int)(blk % nK);
  const int row = 16 * R + (l & 15), k = 16 * K + 4 * (l >> 4) + 2 * h + e;
  Vt[o] = V[(size_t)row * nGp + k];
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

// ------------------------------------------------------------------ launchers
void launch_slice_rows(const double* X, int len, int nlines, int ns, int tiled, int8_t* out, int* expo,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_slice_rows, dim3(nlines), dim3(1024), 0, st, X, len, nlines, ns, tiled, out, expo);
}

void launch_tile_f64(const double* V, int nGp, double* Vt, hipStream_t st) {
  const size_t n = (size_t)nGp * nGp;
  hipLaunchKernelGGL(k_tile_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, V, nGp, Vt);
}

void launch_split_f32(const double* X, size_t n, float* hi, float* lo, hipStream_t st) {
  hipLaunchKernelGGL(k_split_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X, n, hi, lo);
}

void launch_split_bf16x3(const double* X, size_t n, __bf16* out, hipStream_t st) {
  hipLaunchKernelGGL(k_split_bf16x3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X, n, out,
                     out + n, out + 2 * n);
}

}  // namespace cody
