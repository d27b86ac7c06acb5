// Device-side pieces of the MFMA contractions shared by meanfield_gemm.hip and the fused
// right-hand-side launches of mctdhx_kernels.hip: vector types, the fixed-order LDS reduction of
// per-wave tiles, and the bodies of the Gram kernel (G = S^T W split-K partials), of the k-split
// sum of W^T and of the int8 slicing of the tiled S^T.  A body takes its workgroup indices as arguments; every thread of the
// workgroup calls it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mctdhx_kernels.hpp"

namespace cody {
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef int i4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// Sum the WAVES waves' 16 x (16 NT) FP64 tiles in LDS (pairwise halving, fixed order, so the
// result is bit-reproducible) and store W^T[c][row0 + r].
// vals[t][i] holds row rowOf(i) of column t*16 + l%16 of this wave's share of the contraction.
template <int NT, bool F64LAYOUT, int WAVES = 4>
__device__ __forceinline__ void reduce_store(double (&vals)[NT][4], double* __restrict__ WT, int ldc,
                                             int row0, double* red /* [WAVES/2][16][NT*16+1] */) {
  constexpr int LD = NT * 16 + 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, rr = lane & 15;
  auto rowOf = [&](int i) { return F64LAYOUT ? 4 * i + q : 4 * q + i; };
#pragma unroll
  for (int half = WAVES / 2; half >= 1; half >>= 1) {
    if (wave >= half && wave < 2 * half) {
      double* slot = red + (wave - half) * 16 * LD;
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) slot[rowOf(i) * LD + t * 16 + rr] = vals[t][i];
    }
    __syncthreads();
    if (wave < half) {
      const double* slot = red + wave * 16 * LD;
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) vals[t][i] += slot[rowOf(i) * LD + t * 16 + rr];
    }
    __syncthreads();
  }
  if (wave == 0) {
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) red[rowOf(i) * LD + t * 16 + rr] = vals[t][i];
  }
  __syncthreads();
  // coalesced transposed store: 16 consecutive rows of one column per 16 threads
  for (int e = threadIdx.x; e < 16 * NT * 16; e += blockDim.x) {
    const int c = e >> 4, r = e & 15;
    WT[(size_t)c * ldc + row0 + r] = red[r * LD + c];
  }
}

// Active k blocks of the pair sources.  Block b (of nBlk blocks, each g 16-point fragments wide)
// is active when any of its nT x g fragment flags act[t nK16 + g b + q] (k_pair_src_t) is set,
// i.e. when some pair source is non-zero there; an inactive block's products are all zero, as in
// the reference's convolution sums, which leave zero sources out
// (src/Utils/M_Utils_ConvolutionIntegral.f90:84-93).  The active blocks go to sList in increasing
// order, padded with one inactive block up to a multiple of `mult` (its sources are zero, so the
// padding adds nothing; with no inactive block nBlk is already a multiple of mult), so that every
// k range of a workgroup holds the same number of blocks.  Returns the padded count.  Called by
// every thread of the workgroup; sWave holds one int per wave plus one.
__device__ int active_blocks(const int* __restrict__ act, int nT, int nK16, int g, int nBlk, int mult,
                             int* __restrict__ sList, int* __restrict__ sWave) {
  const int nth = blockDim.x, nw = nth >> 6, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int per = (nBlk + nth - 1) / nth, b0 = min(tid * per, nBlk), b1 = min(b0 + per, nBlk);
  unsigned mine = 0;  // bit i: block b0 + i is active (per <= 32)
  int zero = nBlk;    // first inactive block of this thread's range
  for (int b = b0; b < b1; ++b) {
    int any = 0;
    for (int t = 0; t < nT; ++t)
      for (int q = 0; q < g; ++q) any |= act[(size_t)t * nK16 + g * b + q];
    if (any) mine |= 1u << (b - b0);
    else if (zero == nBlk) zero = b;
  }
  // exclusive prefix sum of the per-thread counts: within the wave by shuffles, then over waves
  const int cnt = __popc(mine);
  int incl = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (tid == 0) sWave[nw] = nBlk;
  if (lane == 63) sWave[wave] = incl;
  __syncthreads();
  if (zero < nBlk) atomicMin(&sWave[nw], zero);
  int off = incl - cnt;
  for (int v = 0; v < wave; ++v) off += sWave[v];
  int total = 0;
  for (int v = 0; v < nw; ++v) total += sWave[v];
  for (int i = 0; mine; ++i, mine >>= 1)
    if (mine & 1u) sList[off++] = b0 + i;
  __syncthreads();
  const int padded = (total + mult - 1) / mult * mult;
  const int z = sWave[nw];
  for (int k = total + tid; k < padded; k += nth) sList[k] = z;
  __syncthreads();
  return padded;
}


// Gram split-K partials Gpart[y][c2][c1] = sum_{x in range y} S^T[c1][x] W^T[c2][x] (G = S^T W,
// used by h2): A = S^T fragment-tiled (as k_pair_src_t writes it), B = W^T row-major; one
// workgroup per 16 rows of S^T (bx) and k range by, one quarter of it per wave.
// act (optional): the pair sources' activity flags, act[t (ldb >> actShift) + block] for column
// tile t of S^T and block of 2^actShift points; a wave whose k range holds no active block of its
// row tile bx skips its loads and MFMAs (those products are exact zeros).
template <int NT>
__device__ __forceinline__ void gram_f64_body(int bx, int by, const double* __restrict__ At, int nK,
                                                  const double* __restrict__ B, int ldb,
                                                  double* __restrict__ Cout, int ldc, int kq,
                                                  size_t cSplitStride, const int* __restrict__ act = nullptr,
                                                  int actShift = 4) {
  __shared__ double red[2 * 16 * (NT * 16 + 1)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, rr = lane & 15;
  const int row0 = bx * 16;
  const int kw = (by * 4 + wave) * kq;
  const double* ap = At + ((size_t)bx * nK + (kw >> 4)) * 256 + 2 * lane;
  const double* bp = B + (size_t)rr * ldb + kw + 4 * q;
  const size_t bstep = (size_t)16 * ldb;
  d4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = (d4_t){0.0, 0.0, 0.0, 0.0};
  // NBUF register stages rotate through the k blocks: the load issued at step j fills the stage
  // consumed NBUF - 1 steps later, so each MFMA batch only waits for loads that are that old.
  // Loads past the last block are clamped onto it (redundant, never consumed), so every load is
  // unconditional and the compiler's wait counts stay exact.
  constexpr int NBUF = NT <= 6 ? 3 : 2;
  const int nb = kq >> 4;
  d4_t a[NBUF];
  d4_t b[NBUF][NT];
  auto load = [&](int u, int blk) {
    const int kb = min(blk, nb - 1) << 4;
    const d2_t lo = *reinterpret_cast<const d2_t*>(ap + 16 * kb);
    const d2_t hi = *reinterpret_cast<const d2_t*>(ap + 16 * kb + 128);
    a[u] = (d4_t){lo[0], lo[1], hi[0], hi[1]};
#pragma unroll
    for (int t = 0; t < NT; ++t) b[u][t] = *reinterpret_cast<const d4_t*>(bp + t * bstep + kb);
  };
  auto mma = [&](int u) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][s], b[u][t][s], acc[t], 0, 0, 0);
  };
  bool live = true;
  if (act) {
    const int b0 = kw >> actShift, nbw = ((kw + kq - 1) >> actShift) - b0 + 1;
    const int* fl = act + (size_t)bx * (ldb >> actShift) + b0;
    int any = 0;
    for (int i = lane; i < nbw; i += 64) any |= fl[i];
    live = __ballot(any != 0) != 0ull;
  }
  if (live) {
#pragma unroll
    for (int j = 0; j < NBUF - 1; ++j) load(j, j);
    int i = 0;
    for (; i + NBUF <= nb; i += NBUF) {
#pragma unroll
      for (int j = 0; j < NBUF; ++j) {
        load((j + NBUF - 1) % NBUF, i + j + NBUF - 1);
        mma(j);
      }
    }
    // the remaining nb - i < NBUF blocks are already in stages 0 .. nb - i - 1
#pragma unroll
    for (int j = 0; j < NBUF - 1; ++j)
      if (j < nb - i) mma(j);
  }
  double vals[NT][4];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) vals[t][i] = acc[t][i];
  reduce_store<NT, true>(vals, Cout + by * cSplitStride, ldc, row0, red);
}

// W^T = sum of the nsplit partials, in fixed order (two doubles per thread).
__device__ __forceinline__ void wsum_body(int bx, const double* __restrict__ Wpart, int nsplit, size_t plane,
                                              double* __restrict__ WT) {
  const size_t i = 2 * ((size_t)bx * blockDim.x + threadIdx.x);
  if (i >= plane) return;
  d2_t v = *reinterpret_cast<const d2_t*>(Wpart + i);
  for (int ks = 1; ks < nsplit; ++ks) v += *reinterpret_cast<const d2_t*>(Wpart + ks * plane + i);
  *reinterpret_cast<d2_t*>(WT + i) = v;
}

}  // namespace cody
