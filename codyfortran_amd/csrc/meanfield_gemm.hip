// Mean-field contraction W^T = (V S)^T on gfx950 MFMA, all precisions (A4, the dominant kernel).
//
// Work decomposition: one workgroup (4 waves) owns 16 rows x all P_c columns of W; its four waves
// take one quarter of the contraction index each, so a 4096-point grid runs 256 workgroups /
// 1024 waves with every V element read exactly once.  Each wave keeps two or three k blocks of
// operands in registers (later blocks are in flight while the MFMAs of the current one issue).
// The four quarter sums are added in LDS in fixed order (bit-reproducible) and written
// transposed, W^T[c][x], which is the layout the per-grid-point kernels read coalesced.
//
// Operand layouts (k contiguous in both): V[row][k] (V is symmetric, row-major), S^T[col][k].
// Inside each k-block the contraction index is permuted so that every lane fetches the values
// of its MFMA steps with one wide load: lane group q = l/16 takes k = kb + w*q + s at step s,
// w = values per lane per MFMA (4 for FP64, 64/4 = 16 for int8, ...).
//
// Result layouts of the 16x16 MFMAs: FP64 (v_mfma_f64_16x16x4f64): element i of lane l is row
// 4i + l/16; 32-bit results (f32/bf16/i8): row 4(l/16) + i; column l%16 in both.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <type_traits>

#include "../../include/cody_mctdhx.h"
#include "fedvr_device.hpp"
#include "gemm_device.hpp"

namespace cody {


// ---------------------------------------------------------------------------------------- FP64
// Gram split-K partials Gpart[y][c2][c1] = sum_{x in range y} S^T[c1][x] W^T[c2][x] (G = S^T W,
// used by h2): A = S^T fragment-tiled (as k_pair_src_t writes it), B = W^T row-major; one
// workgroup per 16 rows of S^T and k range blockIdx.y, one quarter of it per wave.
template <int NT>
__global__ __launch_bounds__(256) void k_gram_f64(const double* __restrict__ At, int nK,
                                                  const double* __restrict__ B, int ldb,
                                                  double* __restrict__ Cout, int ldc, int kq,
                                                  size_t cSplitStride) {
  gram_f64_body<NT>(blockIdx.x, blockIdx.y, At, nK, B, ldb, Cout, ldc, kq, cSplitStride);
}

// W^T = (V S)^T with both operands fragment-tiled: the 16 x 16 block (R, K) of V is 2 KB
// contiguous at Vt + (R nK + K) 256 (k_tile_f64), as two 1 KB halves [h][lane][2] holding
// V[16R + l%16][16K + 4(l/16) + 2h + e]; S^T is tiled the same way with R the column tile
// (k_pair_src).
//
// Decomposition: a workgroup owns 64 rows (4 row tiles) and one of gridDim.y contiguous k ranges
// (the k split, 1..kMaxWSplit, chosen per shape by cody_autotune);
// its waves are (row tile r, k sub-range h).  Per 16-wide k block the S^T fragments of both k
// sub-ranges (KH x NT x 2 KB) are fetched once by the whole workgroup, 1 KB per wave and chunk,
// and shared through LDS, so S^T crosses L2 once per 64 rows (it is the S^T traffic, not V, that
// bounds a 16-row decomposition).  Each wave streams its V fragments straight into registers.
//
// Pipeline (D register stages, two LDS buffers, one barrier per block): at step j a wave issues
// the global loads of block j + D - 1 (V fragment and its S^T chunks), writes the S^T chunks of
// block j + 1 into LDS buffer (j + 1) % 2, reads block j's fragments from buffer j % 2 and issues
// its MFMAs.  Every global load is issued D - 1 steps ahead of its use: loads complete in issue
// order, so operands fetched at different distances would make every wait cover the farthest.
//
// The k sub-ranges are summed in LDS and each workgroup stores its k-split partial; k_wsum adds
// the partials in fixed order (bit-reproducible).  With a single k range the kernel writes W^T.  The partials are combined by a second
// kernel rather than by the last workgroup of a row group: the device-scope fences that a
// last-arriver scheme needs write back and invalidate whole L2 caches on this GPU, which cost
// more than the extra launch.
// The body takes the workgroup's row group bx, k range by of nsplit (the grid of k_mf_f64_w).
template <int NT>
__device__ __forceinline__ void mf_f64_w_body(int bx, int by, int nsplit, const double* __restrict__ Vt,
                                              const double* __restrict__ STt, int nGp,
                                              double* __restrict__ Wpart, const int* __restrict__ act) {
  constexpr int KH = NT <= 5 ? 2 : 1, W = 4 * KH, D = 4, LD = NT * 16 + 1;
  constexpr int NCH = KH * NT * 2, CPW = (NCH + W - 1) / W;  // 1 KB S^T chunks per block, per wave
  constexpr int BUF = W * CPW * 128;  // doubles per LDS buffer (chunks NCH.. are never read)
  constexpr int LDS = 2 * BUF > 4 * 16 * LD ? 2 * BUF : 4 * 16 * LD;
  __shared__ double lds[LDS];
  __shared__ int sList[kSkipMaxBlocks16];
  __shared__ int sWave[4 * KH + 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, rr = lane & 15;
  const int r = wave & 3, h = wave >> 2;
  const int nK = nGp >> 4;
  // k blocks of each (k range, sub-range): consecutive positions of the active-block list, or of
  // all blocks when act is null
  const int nbw = (act ? active_blocks(act, NT, nK, 1, nK, nsplit * KH, sList, sWave) : nK) / (nsplit * KH);
  auto kblock = [&](int hh, int blk) {
    const int j = (by * KH + hh) * nbw + min(blk, nbw - 1);
    return act ? sList[j] : j;
  };
  const double* ab = Vt + (size_t)(bx * 4 + r) * nK * 256 + 2 * lane;
  // this wave's S^T chunks: chunk c = wave + W m = (hh NT + t) 2 + half of every block (chunks
  // past NCH are clamped onto the last one for the load and not written)
  auto chunk_src = [&](int m, int blk) {
    const int c = min(wave + W * m, NCH - 1);
    const int half = c & 1, ht = c >> 1, hh = ht / NT, t = ht % NT;
    return STt + ((size_t)t * nK + kblock(hh, blk)) * 256 + 128 * half + 2 * lane;
  };
  d4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = (d4_t){0.0, 0.0, 0.0, 0.0};
  d4_t a[D];
  d2_t stg[D][CPW];
  auto gload = [&](int u, int blk) {
    const size_t o = (size_t)kblock(h, blk) * 256;
    const d2_t lo = *reinterpret_cast<const d2_t*>(ab + o);
    const d2_t hi = *reinterpret_cast<const d2_t*>(ab + o + 128);
    a[u] = (d4_t){lo[0], lo[1], hi[0], hi[1]};
#pragma unroll
    for (int m = 0; m < CPW; ++m) stg[u][m] = *reinterpret_cast<const d2_t*>(chunk_src(m, blk));
  };
  auto lwrite = [&](int u, int buf) {
#pragma unroll
    for (int m = 0; m < CPW; ++m) *reinterpret_cast<d2_t*>(lds + buf * BUF + (wave + W * m) * 128 + 2 * lane) = stg[u][m];
  };
  auto step = [&](int u, int buf) {
    d4_t bf[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const double* p = lds + buf * BUF + (h * NT + t) * 256 + 2 * lane;
      const d2_t lo = *reinterpret_cast<const d2_t*>(p);
      const d2_t hi = *reinterpret_cast<const d2_t*>(p + 128);
      bf[t] = (d4_t){lo[0], lo[1], hi[0], hi[1]};
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][s], bf[t][s], acc[t], 0, 0, 0);
  };
  // stage invariant at the top of an iteration (i a multiple of D, D even): register stages
  // 0..D-2 hold blocks i..i+D-2 (stage 0 already written to LDS buffer 0).  Loads past the last
  // block are clamped onto it and never consumed.
  static_assert(D == 4, "the steps below are written out for four stages");
  // (loops holding a barrier are not unrolled by the compiler, and the stage index must be a
  // compile-time constant for the stages to live in registers)
#define CODY_W_STEP(J)                                 \
  {                                                    \
    gload(((J) + D - 1) % D, i + (J) + D - 1);         \
    lwrite(((J) + 1) % D, ((J) + 1) & 1);              \
    step((J), (J) & 1);                                \
    __syncthreads();                                   \
  }
  if (nbw > 0) {  // (no active block: the partial stays zero)
    gload(0, 0);
    gload(1, 1);
    gload(2, 2);
    lwrite(0, 0);
    __syncthreads();
    int i = 0;
    for (; i + D <= nbw; i += D) {
      CODY_W_STEP(0)
      CODY_W_STEP(1)
      CODY_W_STEP(2)
      CODY_W_STEP(3)
    }
    // remaining nbw - i < D blocks (the conditions are uniform over the workgroup)
    if (i < nbw) CODY_W_STEP(0)
    if (i + 1 < nbw) CODY_W_STEP(1)
    if (i + 2 < nbw) CODY_W_STEP(2)
  }
#undef CODY_W_STEP
  // sum the k sub-ranges of each row tile into red[r] (element u of lane l: row 4u + l/16)
  double* slot = lds + r * 16 * LD;
  if (h == KH - 1) {
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int u = 0; u < 4; ++u) slot[(4 * u + q) * LD + t * 16 + rr] = acc[t][u];
  }
  __syncthreads();
  if (KH == 2 && h == 0) {
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int u = 0; u < 4; ++u) slot[(4 * u + q) * LD + t * 16 + rr] += acc[t][u];
  }
  __syncthreads();
  // store this k split's 64 x (16 NT) partial transposed: 64 consecutive rows per column
  const int row0 = bx * 64;
  const size_t plane = (size_t)NT * 16 * nGp;
  double* Wp = Wpart + by * plane;  // (the plane of a single k range is W^T itself)
  for (int e = threadIdx.x; e < 64 * NT * 16; e += blockDim.x) {
    const int c = e >> 6, rl = e & 63;
    Wp[(size_t)c * nGp + row0 + rl] = lds[(rl >> 4) * 16 * LD + (rl & 15) * LD + c];
  }
}

template <int NT>
__global__ __launch_bounds__(NT <= 5 ? 512 : 256) void k_mf_f64_w(
    const double* __restrict__ Vt, const double* __restrict__ STt, int nGp, double* __restrict__ Wpart,
    const int* __restrict__ act) {
  mf_f64_w_body<NT>(blockIdx.x, blockIdx.y, gridDim.y, Vt, STt, nGp, Wpart, act);
}

// W^T = sum of the nsplit partials, in fixed order (two doubles per thread).
__global__ __launch_bounds__(256) void k_wsum(const double* __restrict__ Wpart, int nsplit, size_t plane,
                                              double* __restrict__ WT) {
  wsum_body(blockIdx.x, Wpart, nsplit, plane, WT);
}

// Eight signed digit bytes (two words, lowest byte first) as eight BF16 values, exactly.
__device__ __forceinline__ bf8_t digits_bf16(int w0, int w1) {
  bf8_t r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[j] = static_cast<__bf16>(static_cast<float>((w0 << (24 - 8 * j)) >> 24));
    r[4 + j] = static_cast<__bf16>(static_cast<float>((w1 << (24 - 8 * j)) >> 24));
  }
  return r;
}

// ---------------------------------------------------------------------------------------- int8, 16-row
// The int8-sliced contraction for the right-hand side: one workgroup per 16-row tile of W (nGp / 16
// workgroups) and one wave per 16-column tile t, all waves walking the same k blocks of 64 points:
// the active ones (a block is active when some line's pair source in it is non-zero, per-block
// flags from pair_src_i8_body), or all blocks when the flags are null.  Both operands are
// fragment-tiled: digit p of the 16-line x 64-k block (R, K) is 1 KiB at ((R nK64 + K) NS + p) 1024,
// lane l holding bytes k = 64K + 16(l/16) .. +15 of line 16R + l%16.  V's digits carry one scale per
// row (eA), S's one per line and block (eSb[K Pc + line]).  Per block a wave loads the five V digit
// fragments of its rows (the same for every wave of the workgroup) and the five S digit fragments
// of its columns, one block ahead (loads complete in issue order; loads past the last block are
// clamped onto it).  Products with p + q <= NS - 1 that share the scale 254^{-(p+q+2)} accumulate
// exactly in one int32 tile per diagonal over the block (|sum| <= NS 127^2 64); the block's
// diagonals are recombined in FP64 with the block's scale while the next block's MFMAs run (two
// tile sets), and the blocks are added in list order.  (Measured at 4096 points with 18 of 64
// blocks active: two blocks in flight beat three or four; sharing the V fragments through LDS with
// a barrier per block is slower.)
//
// mf_i8_r16_acc: that loop for the workgroup's row tile rt; v[u] is then W at row 4(l/16) + u,
// column l%16 of the wave's tile.
//
// With BF the same digits run on BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation: a
// digit (|d| <= 127) is exact in BF16, a product of two exact in FP32, and a diagonal's sum over
// one block (|sum| <= 5 127^2 64 < 2^24) exact too, so the block sums, and W, equal the int8
// path's bit for bit.  The lane's 16 digit bytes (k = 16 (l/16) .. +15) feed two MFMAs of eight.
//
// pred >= 0: every thread's vote whether the tile is needed, taken at the loop's first barrier; when
// no thread votes 1 the function returns false at once (v unset).  pred < 0: no vote (true).
template <int NT, int PF, bool BF = false>
__device__ __forceinline__ bool mf_i8_r16_acc(int rt, const int8_t* __restrict__ As, const int* __restrict__ eA,
                                              const int8_t* __restrict__ Bs, const int* __restrict__ eSb,
                                              const int* __restrict__ actK, int nGp, int Pc, double (&v)[4],
                                              int pred = -1) {
  static_assert(PF == 2, "the loop below keeps one block in flight");
  using acc_t = typename std::conditional<BF, f4_t, i4_t>::type;
  constexpr int NS = kInt8Slices;
  __shared__ int sList[kR16MaxBlocks64];
  __shared__ int sCount;
  const int t = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, q = lane >> 4, rr = lane & 15;
  const int nK64 = nGp >> 6;
  if (actK) {  // the active blocks, in increasing order (wave 0)
    if (threadIdx.x < 64) {
      const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
      int n = 0;
      for (int base = 0; base < nK64; base += 64) {
        const int b = base + lane;
        int any = 0;
        if (b < nK64)
          for (int tt = 0; tt < NT; ++tt) any |= actK[(size_t)tt * nK64 + b];
        const unsigned long long m = __ballot(any != 0);
        if (any) sList[n + __popcll(m & below)] = b;
        n += __popcll(m);
      }
      if (lane == 0) sCount = n;
    }
    if (pred >= 0) {
      if (!__syncthreads_or(pred)) return false;
    } else {
      __syncthreads();
    }
  } else if (pred >= 0 && !__syncthreads_or(pred)) {
    return false;
  }
  const int total = actK ? __builtin_amdgcn_readfirstlane(sCount) : nK64;
  const int8_t* ab = As + (size_t)rt * nK64 * NS * 1024 + 16 * lane;
  const int8_t* bb = Bs + (size_t)t * nK64 * NS * 1024 + 16 * lane;
  const int line = t * 16 + rr;
#pragma unroll
  for (int u = 0; u < 4; ++u) v[u] = 0.0;
  i4_t a[2][NS], b[2][NS];
  acc_t acc[2][NS];
  int es[2];
  double sc[2];
  auto load = [&](int u, int j) {
    j = min(j, total - 1);
    const int K = actK ? __builtin_amdgcn_readfirstlane(sList[j]) : j;
    const size_t o = (size_t)K * NS * 1024;
#pragma unroll
    for (int p = 0; p < NS; ++p) a[u][p] = *reinterpret_cast<const i4_t*>(ab + o + p * 1024);
#pragma unroll
    for (int p = 0; p < NS; ++p) b[u][p] = *reinterpret_cast<const i4_t*>(bb + o + p * 1024);
    es[u] = eSb[(size_t)K * Pc + line];
  };
  auto mfma = [&](int u) {
#pragma unroll
    for (int d = 0; d < NS; ++d) acc[u][d] = (acc_t){0, 0, 0, 0};
    if constexpr (BF) {
      bf8_t ah[NS][2], bh[NS][2];
#pragma unroll
      for (int p = 0; p < NS; ++p) {
        ah[p][0] = digits_bf16(a[u][p][0], a[u][p][1]);
        ah[p][1] = digits_bf16(a[u][p][2], a[u][p][3]);
        bh[p][0] = digits_bf16(b[u][p][0], b[u][p][1]);
        bh[p][1] = digits_bf16(b[u][p][2], b[u][p][3]);
      }
#pragma unroll
      for (int p = 0; p < NS; ++p)
#pragma unroll
        for (int d = p; d < NS; ++d)
#pragma unroll
          for (int h = 0; h < 2; ++h)
            acc[u][d] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[p][h], bh[d - p][h], acc[u][d], 0, 0, 0);
    } else {
#pragma unroll
      for (int p = 0; p < NS; ++p)
#pragma unroll
        for (int d = p; d < NS; ++d)
          acc[u][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[u][p], b[u][d - p], acc[u][d], 0, 0, 0);
    }
    sc[u] = ldexp(1.0, es[u]);
  };
  auto flush = [&](int u) {  // the block in tile set u, scaled by its line-and-block factor
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double bs = 0.0;
#pragma unroll
      for (int d = NS - 1; d >= 0; --d) bs += static_cast<double>(acc[u][d][k]) * kSliceScale.v[d + 2];
      v[k] = fma(bs, sc[u], v[k]);
    }
  };
  if (total > 0) {
    load(0, 0);
    for (int j = 0; j < total; j += 2) {
      load(1, j + 1);
      mfma(0);              // block j
      if (j > 0) flush(1);  // block j - 1
      if (j + 1 < total) {
        load(0, j + 2);
        mfma(1);   // block j + 1
        flush(0);  // block j
      }
    }
    if ((total - 1) & 1)  // the last block (tile sets indexed by constants only: no scratch copy)
      flush(1);
    else
      flush(0);
  }
  // 32-bit MFMA results: element u of lane l is row 4(l/16) + u, column l%16
  const int row0 = rt * 16 + 4 * q;
#pragma unroll
  for (int u = 0; u < 4; ++u) v[u] = ldexp(v[u], eA[row0 + u]);
  return true;
}

// The body takes the workgroup's row tile rt.
// With orbs set (a right-hand side, whose W is only multiplied by phi and by the pair sources), a
// tile whose 16 points carry no orbital value writes zero rows of W^T instead (the vote rides on
// the contraction's first barrier); without (the mean fields themselves) every row is formed.
template <int NT, int PF, bool BF = false>
__device__ __forceinline__ void mf_i8_r16_body(int rt, const int8_t* __restrict__ As, const int* __restrict__ eA,
                                               const int8_t* __restrict__ Bs, const int* __restrict__ eSb,
                                               const int* __restrict__ actK, int nGp, int Pc, double* __restrict__ WT,
                                               const double2* __restrict__ orbs = nullptr, int nG = 0, int nO = 0) {
  int pred = -1;
  if (orbs) {  // (NT 64 >= 16 nO threads for every orbital count)
    const int px = threadIdx.x & 15, pj = threadIdx.x >> 4, gx = rt * 16 + px;
    const double2 ph = (pj < nO && gx < nG) ? orbs[(size_t)pj * nG + gx] : make_double2(0.0, 0.0);
    pred = (ph.x != 0.0 || ph.y != 0.0) ? 1 : 0;
  }
  double v[4];
  const bool live = mf_i8_r16_acc<NT, PF, BF>(rt, As, eA, Bs, eSb, actK, nGp, Pc, v, pred);
  const int t = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double* dst = WT + (size_t)(t * 16 + (lane & 15)) * nGp + rt * 16 + 4 * (lane >> 4);
#pragma unroll
  for (int u = 0; u < 4; ++u) dst[u] = live ? v[u] : 0.0;
}

// ------------------------------------------------------------------ int8 contraction + orbital correlation
// The 16-row int8 contraction with the orbital correlation of its 16 points in the epilogue
// (ApplyCorrelationOp, S_Method_Mb_OrbBased.f90:647-872): the workgroup holds W(x) for all pairs of
// its points once the k loop ends, so the correlation needs no pass of its own over W^T.  With
// rho2' = rho2 x_3 R, rho2 antisymmetric in its first index pair (a, b),
//   corr_j(x) = sum_{k1,k,l} phi_k1(x) M_kl(x) rho2'(k1, k, j, l) = sum_m R(m, j) c_m(x),
//   c_m(x)    = sum_{a<b, l} rho2(a, b, m, l) B_{ab,l}(x),   B_{ab,l} = phi_a M_bl - phi_b M_al,
// where M_kl = W_p (pair p = {l <= k}) and M_kl = conj(W_p) for k < l.  The B terms (half the
// products of the plain sum) are formed per (point, l) from registers; c = B^T rho2 is a 16-point x
// (2 NO real) x (2 np NO real) product, run on FP64 MFMA 16x16x4 with the complex values split into
// real and imaginary parts along k and along the columns, its k steps shared out over the waves and
// the wave tiles added in LDS in fixed order.  rho2(a, b, m, l) = sign(m, l) rho2can[pair(a, b),
// pair(m, l)] is a signed copy of the canonical rho2.  R is applied later (by the projection,
// project_corr_body), so this kernel does not wait for it: it writes c_m(x) and this tile's two
// overlap partials, A(i, j) = sum_x conj(phi_i) w h_j and B(i, m) = sum_x conj(phi_i) w c_m, at
// part[rt NO^2 + i + NO j] and part[(nGp / 16 + rt) NO^2 + i + NO m] (h phi only in the modes whose
// right-hand side has the one-body part).
template <int NO, int PF, bool BF = false>
__device__ __forceinline__ void mf_i8_orb_body(int rt, const int8_t* __restrict__ As, const int* __restrict__ eA,
                                               const int8_t* __restrict__ Bs, const int* __restrict__ eSb,
                                               const int* __restrict__ actK, int nGp, double* __restrict__ WT,
                                               const OrbCorrArgs& o) {
  constexpr int NT = (NO * (NO + 1) + 15) / 16, NTH = NT * 64, n2 = NO * NO, np = NO * (NO - 1) / 2;
  constexpr int NE = np * NO;                    // B terms per point (a < b, l)
  constexpr int NEP = NE > 0 ? NE : 1;
  constexpr int NR = NE * NO;                    // rho2(a < b, m, l) entries
  constexpr int KS = (2 * NE + 3) / 4;           // FP64 MFMA k steps (4 real k each)
  constexpr int RPT = NR > 0 ? (NR + NTH - 1) / NTH : 1;  // rho2 entries staged per thread
  __shared__ double sW[NT * 16][17];
  __shared__ double sBre[NEP][16], sBim[NEP][16];  // B_{e}(x), e = pair(a, b) + np l
  __shared__ double2 sRho[NR > 0 ? NR : 1];         // rho2(a < b, m, l) at e NO + m
  __shared__ double sC[NT][16][17];                 // the waves' c tiles: row x, column 2m + re/im
  __shared__ double2 sPhi[NO][16], sH[NO][16], sCm[NO][16];
  __shared__ double sw[16];
  __shared__ int sAB[np > 0 ? np : 1];              // a + 16 b of strict pair index pab
  const int tid = threadIdx.x, x0 = rt * 16;
  // epilogue operands, loaded ahead of the k loop (their latency hides behind it)
  double2 rR[RPT];
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int e = tid + NTH * u;  // e = (pab + np l) NO + m
    double2 r = cmk(0, 0);
    if (e < NR) {
      const int m = e % NO, ee = e / NO, pab = ee % np, l = ee / np;
      if (m != l) {
        r = o.r2can[pab + np * (m < l ? strict_pair(NO, m, l) : strict_pair(NO, l, m))];
        if (m > l) r = cmk(-r.x, -r.y);
      }
    }
    rR[u] = r;
  }
  const bool own = tid < NO * 16;  // (point tid % 16, orbital tid / 16)
  const int px = tid & 15, pj = tid >> 4, gx = x0 + px;
  const bool in = own && gx < o.nG;
  const double2 phiL = in ? o.orbs[(size_t)pj * o.nG + gx] : cmk(0, 0);
  const double2 hL = (in && (o.mode == kModeFull || o.mode == kModeNonLinConsistent))
                         ? o.hphi[(size_t)pj * o.nG + gx] : cmk(0, 0);
  const double wL = (tid < 16 && x0 + tid < o.nG) ? o.w[x0 + tid] : 0.0;
  if (tid < np) {
    int a = 0, r = tid;
    while (r >= NO - 1 - a) {
      r -= NO - 1 - a;
      ++a;
    }
    sAB[tid] = a + 16 * (a + 1 + r);
  }
  // A tile whose 16 points carry no orbital value at all (every phi exactly zero there) needs no
  // W: its correlation terms are products with phi and its Gram terms products with the pair
  // sources, both exact zeros.  It writes zeros for its rows of W^T, its c and its partials (the
  // test rides on the contraction's first barrier).
  double v[4];
  const bool live = mf_i8_r16_acc<NT, PF, BF>(rt, As, eA, Bs, eSb, actK, nGp, NT * 16, v,
                                              in && (phiL.x != 0.0 || phiL.y != 0.0) ? 1 : 0);
  const int t = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, q = lane >> 4, rr = lane & 15;
  double* dst = WT + (size_t)(t * 16 + rr) * nGp + x0 + 4 * q;
  if (!live) {
#pragma unroll
    for (int u = 0; u < 4; ++u) dst[u] = 0.0;
    if (in) o.corrc[(size_t)pj * o.nG + gx] = cmk(0, 0);
    for (int e = tid; e < 2 * n2; e += NTH) {
      const int half = e / n2, f = e % n2;
      o.part[((size_t)half * (nGp >> 4) + rt) * n2 + f] = cmk(0, 0);
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    dst[u] = v[u];
    sW[t * 16 + rr][4 * q + u] = v[u];
  }
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int e = tid + NTH * u;
    if (e < NR) sRho[e] = rR[u];
  }
  if (own) {
    sPhi[pj][px] = phiL;
    sH[pj][px] = hL;
  }
  if (tid < 16) sw[tid] = wL;
  __syncthreads();
  // B terms: thread (x, l) from phi(x) and the column l of M(x), both in registers
  if constexpr (np > 0) {
    if (own) {
      const int x = px, l = pj;
      double2 ph[NO], ml[NO];
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        ph[k] = sPhi[k][x];
        const int pp = l <= k ? pair_index(NO, l, k) : pair_index(NO, k, l);
        const double2 wv = cmk(sW[2 * pp][x], sW[2 * pp + 1][x]);
        ml[k] = l <= k ? wv : cconj(wv);
      }
#pragma unroll
      for (int a = 0; a < NO; ++a)
#pragma unroll
        for (int b = a + 1; b < NO; ++b) {
          const double2 bv = csub(cmul(ph[a], ml[b]), cmul(ph[b], ml[a]));
          const int e = strict_pair(NO, a, b) + np * l;
          sBre[e][x] = bv.x;
          sBim[e][x] = bv.y;
        }
    }
  }
  __syncthreads();
  // c = B^T rho2 on FP64 MFMA: step s takes real k = 4 s + lane / 16 (B term k / 2, part k % 2);
  // lane l supplies A[x = l % 16][k] and rho2's real embedding at [k][n = l % 16], n = 2 m + part
  {
    d4_t cacc = (d4_t){0.0, 0.0, 0.0, 0.0};
    const int n = rr, m = n >> 1, cn = n & 1;
    for (int s = t; s < KS; s += NT) {
      const int k = 4 * s + q, e = k >> 1, ck = k & 1;
      double av = 0.0, bv = 0.0;
      if (e < NE) {
        av = ck ? sBim[e][rr] : sBre[e][rr];
        if (m < NO) {
          const double2 r = sRho[e * NO + m];
          // real embedding: [re b, im b] x [[re r, im r], [-im r, re r]]
          bv = ck ? (cn ? r.x : -r.y) : (cn ? r.y : r.x);
        }
      }
      cacc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, cacc, 0, 0, 0);
    }
    // FP64 result layout: element i of lane l is row 4 i + l / 16, column l % 16
#pragma unroll
    for (int i = 0; i < 4; ++i) sC[t][4 * i + q][rr] = cacc[i];
  }
  __syncthreads();
  if (own) {  // c_m(x), m = pj, summed over the waves in order
    const int x = px, mm = pj;
    double cr = sC[0][x][2 * mm], ci = sC[0][x][2 * mm + 1];
    for (int w2 = 1; w2 < NT; ++w2) {
      cr += sC[w2][x][2 * mm];
      ci += sC[w2][x][2 * mm + 1];
    }
    const double2 cv = cmk(cr, ci);
    sCm[mm][x] = cv;
    if (in) o.corrc[(size_t)mm * o.nG + gx] = cv;
  }
  __syncthreads();
  for (int e = tid; e < 2 * n2; e += NTH) {
    const int half = e / n2, f = e % n2, i = f % NO, j = f / NO;
    double2 sacc = cmk(0, 0);
#pragma unroll
    for (int x = 0; x < 16; ++x) cfmac(sacc, sPhi[i][x], cscale(half ? sCm[j][x] : sH[j][x], sw[x]));
    o.part[((size_t)half * (nGp >> 4) + rt) * n2 + f] = sacc;
  }
}

template <int NO, bool BF = false>
__global__ __launch_bounds__(((NO * (NO + 1) + 15) / 16) * 64) void k_mf_i8_orb(
    const int8_t* __restrict__ As, const int* __restrict__ eA, const int8_t* __restrict__ Bs,
    const int* __restrict__ eSb, const int* __restrict__ actK, int nGp, double* __restrict__ WT, OrbCorrArgs o) {
  mf_i8_orb_body<NO, kR16Prefetch, BF>(blockIdx.x, As, eA, Bs, eSb, actK, nGp, WT, o);
}

template <int NT, bool BF = false>
__global__ __launch_bounds__(NT * 64) void k_mf_i8_r16(const int8_t* __restrict__ As, const int* __restrict__ eA,
                                                        const int8_t* __restrict__ Bs, const int* __restrict__ eSb,
                                                        const int* __restrict__ actK, int nGp, int Pc,
                                                        double* __restrict__ WT, const double2* __restrict__ orbs,
                                                        int nG, int nO) {
  mf_i8_r16_body<NT, kR16Prefetch, BF>(blockIdx.x, As, eA, Bs, eSb, actK, nGp, Pc, WT, orbs, nG, nO);
}

// ---------------------------------------------------------------------------------------- int8 slices
// (the untiled form for more than 80 pair columns)
template <int NT, int NS>
__global__ __launch_bounds__(256) void k_mf_i8(const int8_t* __restrict__ As, const int* __restrict__ eA,
                                               const int8_t* __restrict__ Bs, const int* __restrict__ fB,
                                               double* __restrict__ WT, int nGp, int Pc) {
  __shared__ double red[2 * 16 * (NT * 16 + 1)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, rr = lane & 15;
  const int row0 = blockIdx.x * 16;
  const int kq = nGp >> 2;
  const int k0 = wave * kq + 16 * q;
  const size_t planeB = (size_t)Pc * nGp, bstep = (size_t)16 * nGp;
  // A (the V slices) is fragment-tiled (k_slice_rows with tiled output): the NS slices of the
  // 16 x 64 block (R, K) are NS contiguous 1 KB fragments [lane][16 bytes]; this wave's blocks
  // are K = wave nb ... wave nb + nb - 1.  B addresses are a uniform base plus one lane offset.
  const int nK64 = nGp >> 6;
  const int8_t* ap = As + ((size_t)blockIdx.x * nK64 + wave * (kq >> 6)) * NS * 1024 + 16 * lane;
  const int8_t* bp = Bs + (size_t)rr * nGp + k0;
  const unsigned offB = (unsigned)(rr * nGp + k0);
  i4_t acc[NT][NS];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int d = 0; d < NS; ++d) acc[t][d] = (i4_t){0, 0, 0, 0};
  const int nb = kq >> 6;
  if constexpr (NT <= 6) {
    // one register set per column tile: as soon as tile t's MFMAs of block i have issued, tile t
    // of block i+1 is loaded into the same registers, so each load has NT - 1 tiles of MFMA work
    // to land; A of block i+1 is loaded at the start of block i.  Loads past the last block are
    // clamped onto it (unconsumed), so every load is unconditional.
    i4_t a[NS], b[NT][NS];
    auto loadA = [&](i4_t (&dst)[NS], int blk) {
      const int kb = min(blk, nb - 1) << 6;
#pragma unroll
      for (int p = 0; p < NS; ++p) dst[p] = *reinterpret_cast<const i4_t*>(ap + (size_t)(kb >> 6) * NS * 1024 + p * 1024);
    };
    auto loadB = [&](int t, int blk) {
      const int kb = min(blk, nb - 1) << 6;
#pragma unroll
      for (int p = 0; p < NS; ++p) b[t][p] = *reinterpret_cast<const i4_t*>(Bs + (t * bstep + p * planeB) + (offB + kb));
    };
    loadA(a, 0);
#pragma unroll
    for (int t = 0; t < NT; ++t) loadB(t, 0);
    for (int i = 0; i < nb; ++i) {
      i4_t an[NS];
      loadA(an, i + 1);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int d = 0; d < NS; ++d)
#pragma unroll
          for (int p = 0; p <= d; ++p)
            acc[t][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[p], b[t][d - p], acc[t][d], 0, 0, 0);
        loadB(t, i + 1);
      }
#pragma unroll
      for (int p = 0; p < NS; ++p) a[p] = an[p];
    }
  } else {
    // wide column sets: A one block ahead, B one column tile ahead
    i4_t a[NS];
#pragma unroll
    for (int p = 0; p < NS; ++p) a[p] = *reinterpret_cast<const i4_t*>(ap + p * 1024);
    auto block = [&](int kb) {
      i4_t b[NS];
#pragma unroll
      for (int p = 0; p < NS; ++p) b[p] = *reinterpret_cast<const i4_t*>(bp + p * planeB + kb);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        i4_t bn[NS];
        const int tn = min(t + 1, NT - 1);
#pragma unroll
        for (int p = 0; p < NS; ++p) bn[p] = *reinterpret_cast<const i4_t*>(bp + tn * bstep + p * planeB + kb);
#pragma unroll
        for (int d = 0; d < NS; ++d)
#pragma unroll
          for (int p = 0; p <= d; ++p)
            acc[t][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[p], b[d - p], acc[t][d], 0, 0, 0);
#pragma unroll
        for (int p = 0; p < NS; ++p) b[p] = bn[p];
      }
    };
    for (int i = 0; i + 1 < nb; ++i) {
      i4_t an[NS];
#pragma unroll
      for (int p = 0; p < NS; ++p) an[p] = *reinterpret_cast<const i4_t*>(ap + (size_t)(i + 1) * NS * 1024 + p * 1024);
      block(i << 6);
#pragma unroll
      for (int p = 0; p < NS; ++p) a[p] = an[p];
    }
    block((nb - 1) << 6);
  }
  double vals[NT][4];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int fc = fB[t * 16 + rr];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double s = 0.0;
#pragma unroll
      for (int d = NS - 1; d >= 0; --d) s += static_cast<double>(acc[t][d][i]) * kSliceScale.v[d + 2];
      vals[t][i] = ldexp(s, eA[row0 + 4 * q + i] + fc);
    }
  }
  reduce_store<NT, false>(vals, WT, nGp, row0, red);
}

// ---------------------------------------------------------------------------------------- split FP32
template <int NT>
__global__ __launch_bounds__(256) void k_mf_f32x3(const float* __restrict__ Ah, const float* __restrict__ Al,
                                                  const float* __restrict__ Bh, const float* __restrict__ Bl,
                                                  double* __restrict__ WT, int nGp) {
  __shared__ double red[2 * 16 * (NT * 16 + 1)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, rr = lane & 15;
  const int row0 = blockIdx.x * 16;
  const int kq = nGp >> 2;
  const int k0 = wave * kq + 4 * q;
  const size_t aoff = (size_t)(row0 + rr) * nGp + k0, bstep = (size_t)16 * nGp;
  const size_t boff = (size_t)rr * nGp + k0;
  double vals[NT][4];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) vals[t][i] = 0.0;
  for (int kc = 0; kc < kq; kc += 32) {
    f4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 32; kb += 16) {
      const f4_t ah = *reinterpret_cast<const f4_t*>(Ah + aoff + kc + kb);
      const f4_t al = *reinterpret_cast<const f4_t*>(Al + aoff + kc + kb);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const f4_t bh = *reinterpret_cast<const f4_t*>(Bh + boff + t * bstep + kc + kb);
        const f4_t bl = *reinterpret_cast<const f4_t*>(Bl + boff + t * bstep + kc + kb);
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
      for (int i = 0; i < 4; ++i) vals[t][i] += static_cast<double>(acc[t][i]);
  }
  reduce_store<NT, false>(vals, WT, nGp, row0, red);
}

// ---------------------------------------------------------------------------------------- 3xBF16
template <int NT>
__global__ __launch_bounds__(256) void k_mf_bf16x3(const __bf16* __restrict__ A, const __bf16* __restrict__ B,
                                                   double* __restrict__ WT, int nGp, int Pc) {
  __shared__ double red[2 * 16 * (NT * 16 + 1)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, rr = lane & 15;
  const int row0 = blockIdx.x * 16;
  const int kq = nGp >> 2;
  const int k0 = wave * kq + 8 * q;
  const size_t planeA = (size_t)nGp * nGp, planeB = (size_t)Pc * nGp, bstep = (size_t)16 * nGp;
  const __bf16* ap = A + (size_t)(row0 + rr) * nGp + k0;
  const __bf16* bp = B + (size_t)rr * nGp + k0;
  double vals[NT][4];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) vals[t][i] = 0.0;
  for (int kc = 0; kc < kq; kc += 64) {
    f4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 64; kb += 32) {
      bf8_t a[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) a[p] = *reinterpret_cast<const bf8_t*>(ap + p * planeA + kc + kb);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        bf8_t b[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) b[p] = *reinterpret_cast<const bf8_t*>(bp + t * bstep + p * planeB + kc + kb);
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
      for (int i = 0; i < 4; ++i) vals[t][i] += static_cast<double>(acc[t][i]);
  }
  reduce_store<NT, false>(vals, WT, nGp, row0, red);
}

// ---------------------------------------------------------------------------------------- launchers
#define CODY_NT_SWITCH2(PC, CALL) \
  switch ((PC) / 16) {            \
    case 1: CALL(1); break;       \
    case 2: CALL(2); break;       \
    case 3: CALL(3); break;       \
    case 4: CALL(4); break;       \
    case 5: CALL(5); break;       \
    case 6: CALL(6); break;       \
    case 7: CALL(7); break;       \
    case 8: CALL(8); break;       \
    case 9: CALL(9); break;       \
    case 10: CALL(10); break;     \
    default: break;               \
  }

// Shape check shared by the k-split contraction launchers, before anything is launched: nGp a
// positive multiple of 256, Pc a multiple of 16 with Pc / 16 in 1..maxTiles (the kernels'
// instantiations), nsplit one of 1, 2, 4, 8 (the partials buffer holds kMaxWSplit = 8 planes) and a
// divisor of the kernel's k blocks.  Returns 0 or CODY_ERR_ARG (1) with the reason in the library's
// last error.
int check_contraction_shape(const char* who, int nGp, int Pc, int nsplit, int maxTiles, int kBlocks) {
  char msg[192];
  if (nGp <= 0 || nGp % 256 != 0) {
    std::snprintf(msg, sizeof msg, "%s: padded grid %d is not a positive multiple of 256", who, nGp);
    return set_error(1, msg);
  }
  if (Pc <= 0 || Pc % 16 != 0 || Pc / 16 > maxTiles) {
    std::snprintf(msg, sizeof msg, "%s: %d pair columns are outside 16..%d in steps of 16", who, Pc, 16 * maxTiles);
    return set_error(1, msg);
  }
  if (nsplit != 1 && nsplit != 2 && nsplit != 4 && nsplit != 8) {
    std::snprintf(msg, sizeof msg, "%s: k split %d is not one of 1, 2, 4, 8", who, nsplit);
    return set_error(1, msg);
  }
  static_assert(kMaxWSplit == 8, "the split check above allows at most kMaxWSplit planes");
  if (kBlocks % nsplit != 0) {
    std::snprintf(msg, sizeof msg, "%s: k split %d does not divide the %d k blocks", who, nsplit, kBlocks);
    return set_error(1, msg);
  }
  return 0;
}

// Wpart holds nsplit x Pc x nGp doubles (unused for nsplit = 1, when the kernel writes WT).  With
// WT null only the partials are formed (k_wsum is not launched).  Returns 0, or CODY_ERR_ARG
// without launching anything when the shape is outside check_contraction_shape.
int launch_mf_f64_w(const double* Vt, const double* STt, double* Wpart, double* WT, int nGp, int Pc,
                    int nsplit, hipStream_t st, const int* act) {
  // per row group the kernel splits the nGp / 16 k blocks into nsplit ranges of KH sub-ranges
  const int kh = Pc / 16 <= 5 ? 2 : 1;
  if (int r = check_contraction_shape("k_mf_f64_w", nGp, Pc, nsplit, 10, nGp / 16 / kh)) return r;
  if (nGp / 16 > kSkipMaxBlocks16) act = nullptr;  // (the block list lives in LDS)
  double* out = (nsplit == 1 && WT) ? WT : Wpart;
#define CALL_W(NT)                                                                                    \
  hipLaunchKernelGGL(k_mf_f64_w<NT>, dim3(nGp / 64, nsplit), dim3((NT) <= 5 ? 512 : 256), 0, st, Vt, STt, \
                     nGp, out, act)
  CODY_NT_SWITCH2(Pc, CALL_W)
#undef CALL_W
  if (!WT || nsplit == 1) return 0;
  const size_t plane = (size_t)Pc * nGp;
  hipLaunchKernelGGL(k_wsum, dim3((unsigned)((plane / 2 + 255) / 256)), dim3(256), 0, st, Wpart, nsplit, plane,
                     WT);
  return 0;
}

// ------------------------------------------------------------------ grouped contractions
// Workgroup b runs block (bx, by) = (map[b].y mod nGp / 64, map[b].y div nGp / 64) of instance
// map[b].x (see GroupInstance).
template <int NT>
__global__ __launch_bounds__(NT <= 5 ? 512 : 256) void k_grp_mf_f64_w(const GroupInstance* __restrict__ gi,
                                                                      const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  const int nbx = a.nGp / 64;
  mf_f64_w_body<NT>(m.y % nbx, m.y / nbx, a.nsplit, a.Vt, a.STt, a.nGp, a.nsplit == 1 ? a.WT : a.Wpart, a.actC);
}

template <int NT>
__global__ __launch_bounds__(NT * 64) void k_grp_mf_i8_r16(const GroupInstance* __restrict__ gi,
                                                            const int2* __restrict__ map, int nO) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  mf_i8_r16_body<NT, kR16Prefetch>(m.y, a.Vi8, a.eV, a.Si8, a.eSb, a.actC, a.nGp, a.Pc, a.WT, a.C + a.nC, a.nG, nO);
}

int launch_grp_mf_f64_w(int Pc, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st) {
#define CALL_GW(NT) \
  hipLaunchKernelGGL(k_grp_mf_f64_w<NT>, dim3(nBlocks), dim3((NT) <= 5 ? 512 : 256), 0, st, gi, map)
  switch (Pc / 16) {
    case 1: CALL_GW(1); break;
    case 2: CALL_GW(2); break;
    case 3: CALL_GW(3); break;
    case 4: CALL_GW(4); break;
    case 5: CALL_GW(5); break;
    case 6: CALL_GW(6); break;
    case 7: CALL_GW(7); break;
    case 8: CALL_GW(8); break;
    case 9: CALL_GW(9); break;
    case 10: CALL_GW(10); break;
    default: return set_error(CODY_ERR_ARG, "k_grp_mf_f64_w: Pc / 16 outside 1..10");
  }
#undef CALL_GW
  return 0;
}

template <int NO>
__global__ __launch_bounds__(((NO * (NO + 1) + 15) / 16) * 64) void k_grp_mf_i8_orb(const GroupInstance* __restrict__ gi,
                                                                                    const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const GroupInstance& a = gi[m.x];
  const OrbCorrArgs o{a.C + a.nC, a.w, a.hphi, a.r2can, a.nG, kModeFull, a.draw, a.part};
  mf_i8_orb_body<NO, kR16Prefetch>(m.y, a.Vi8, a.eV, a.Si8, a.eSb, a.actC, a.nGp, a.WT, o);
}

int launch_grp_mf_i8_orb(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st) {
#define CALL_GO(NO) \
  hipLaunchKernelGGL(k_grp_mf_i8_orb<NO>, dim3(nBlocks), dim3(((NO * (NO + 1) + 15) / 16) * 64), 0, st, gi, map)
  switch (nO) {
    case 1: CALL_GO(1); break;
    case 2: CALL_GO(2); break;
    case 3: CALL_GO(3); break;
    case 4: CALL_GO(4); break;
    case 5: CALL_GO(5); break;
    case 6: CALL_GO(6); break;
    case 7: CALL_GO(7); break;
    case 8: CALL_GO(8); break;
    default: return set_error(CODY_ERR_ARG, "k_grp_mf_i8_orb: nO outside 1..8");
  }
#undef CALL_GO
  return 0;
}

int launch_grp_mf_i8_r16(int Pc, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st, int nO) {
#define CALL_GI(NT) hipLaunchKernelGGL(k_grp_mf_i8_r16<NT>, dim3(nBlocks), dim3((NT) * 64), 0, st, gi, map, nO)
  switch (Pc / 16) {
    case 1: CALL_GI(1); break;
    case 2: CALL_GI(2); break;
    case 3: CALL_GI(3); break;
    case 4: CALL_GI(4); break;
    case 5: CALL_GI(5); break;
    case 6: CALL_GI(6); break;
    case 7: CALL_GI(7); break;
    case 8: CALL_GI(8); break;
    case 9: CALL_GI(9); break;
    case 10: CALL_GI(10); break;
    default: return set_error(CODY_ERR_ARG, "k_grp_mf_i8_r16: Pc / 16 outside 1..10");
  }
#undef CALL_GI
  return 0;
}

// Gpart[s][c2][c1] = sum_{x in split s} ST[c1][x] WT[c2][x]; nsplit * 64 must divide nGp.
void launch_gram_f64(const double* STt, const double* WT, double* Gpart, int nGp, int Pc, int nsplit,
                     hipStream_t st) {
  const int kq = nGp / (4 * nsplit);
#define CALL_G(NT)                                                                                  \
  hipLaunchKernelGGL(k_gram_f64<NT>, dim3(Pc / 16, nsplit), dim3(256), 0, st, STt, nGp / 16, WT, nGp, Gpart, \
                     Pc, kq, (size_t)Pc * Pc)
  CODY_NT_SWITCH2(Pc, CALL_G)
#undef CALL_G
}

// The 16-row int8 contraction with the orbital correlation of each tile (nO <= 8, nGp / 64 <= 256):
// W^T, c_m(x) and two overlap partials per 16-point tile (part: 2 nGp / 16 blocks of nO^2).  Returns
// 0, or CODY_ERR_ARG without launching anything outside these limits.
int launch_mf_i8_orb(const int8_t* As, const int* eA, const int8_t* Bs, const int* eSb, const int* actK, double* WT,
                     int nGp, int nO, const OrbCorrArgs& o, hipStream_t st, bool bf16) {
  if (nGp < 64 || nGp % 64 || nGp / 64 > kR16MaxBlocks64 || nO < 1 || nO > 8 || o.nG > nGp)
    return set_error(CODY_ERR_ARG, "k_mf_i8_orb: nGp must be a positive multiple of 64, at most 64 x 256, above "
                                   "nG, and nO in 1..8");
#define CALL_OC(NO)                                                                                               \
  do {                                                                                                            \
    if (bf16)                                                                                                     \
      hipLaunchKernelGGL((k_mf_i8_orb<NO, true>), dim3(nGp / 16), dim3(((NO * (NO + 1) + 15) / 16) * 64), 0, st, \
                         As, eA, Bs, eSb, actK, nGp, WT, o);                                                      \
    else                                                                                                          \
      hipLaunchKernelGGL((k_mf_i8_orb<NO, false>), dim3(nGp / 16), dim3(((NO * (NO + 1) + 15) / 16) * 64), 0,    \
                         st, As, eA, Bs, eSb, actK, nGp, WT, o);                                                  \
  } while (0)
  switch (nO) {
    case 1: CALL_OC(1); break;
    case 2: CALL_OC(2); break;
    case 3: CALL_OC(3); break;
    case 4: CALL_OC(4); break;
    case 5: CALL_OC(5); break;
    case 6: CALL_OC(6); break;
    case 7: CALL_OC(7); break;
    case 8: CALL_OC(8); break;
  }
#undef CALL_OC
  return 0;
}

// 16-row int8 contraction (Pc <= 80, nGp / 64 <= 256): W^T written directly.  actK: the per-block
// activity flags of S (pair_src_i8_body), or null to contract every k block.  Returns 0, or
// CODY_ERR_ARG without launching anything when the shape is outside these limits.
int launch_mf_i8_r16(const int8_t* As, const int* eA, const int8_t* Bs, const int* eSb, const int* actK, double* WT,
                     int nGp, int Pc, hipStream_t st, bool bf16, const double2* orbs, int nG, int nO) {
  if (nGp < 64 || nGp % 64 || Pc < 16 || Pc > 16 * kR16MaxTiles || Pc % 16 || nGp / 64 > kR16MaxBlocks64)
    return set_error(CODY_ERR_ARG, "k_mf_i8_r16: nGp must be a positive multiple of 64, at most 64 x 256, "
                                   "and Pc a multiple of 16 up to 160");
#define CALL_R16(NT)                                                                                            \
  do {                                                                                                          \
    if (bf16)                                                                                                   \
      hipLaunchKernelGGL((k_mf_i8_r16<NT, true>), dim3(nGp / 16), dim3((NT) * 64), 0, st, As, eA, Bs, eSb, actK, \
                         nGp, Pc, WT, orbs, nG, nO);                                                            \
    else                                                                                                        \
      hipLaunchKernelGGL((k_mf_i8_r16<NT, false>), dim3(nGp / 16), dim3((NT) * 64), 0, st, As, eA, Bs, eSb,     \
                         actK, nGp, Pc, WT, orbs, nG, nO);                                                      \
  } while (0)
  switch (Pc / 16) {
    case 1: CALL_R16(1); break;
    case 2: CALL_R16(2); break;
    case 3: CALL_R16(3); break;
    case 4: CALL_R16(4); break;
    case 5: CALL_R16(5); break;
    case 6: CALL_R16(6); break;
    case 7: CALL_R16(7); break;
    case 8: CALL_R16(8); break;
    case 9: CALL_R16(9); break;
    case 10: CALL_R16(10); break;
  }
#undef CALL_R16
  return 0;
}

void launch_mf_i8(const int8_t* As, const int* eA, const int8_t* Bs, const int* fB, double* WT, int nGp,
                  int Pc, hipStream_t st) {
#define CALL_B(NT)                                                                                     \
  hipLaunchKernelGGL((k_mf_i8<NT, kInt8Slices>), dim3(nGp / 16), dim3(256), 0, st, As, eA, Bs, fB, WT, nGp, \
                     Pc)
  CODY_NT_SWITCH2(Pc, CALL_B)
#undef CALL_B
}

void launch_mf_f32x3(const float* Ah, const float* Al, const float* Bh, const float* Bl, double* WT, int nGp,
                     int Pc, hipStream_t st) {
#define CALL_C(NT) \
  hipLaunchKernelGGL(k_mf_f32x3<NT>, dim3(nGp / 16), dim3(256), 0, st, Ah, Al, Bh, Bl, WT, nGp)
  CODY_NT_SWITCH2(Pc, CALL_C)
#undef CALL_C
}

void launch_mf_bf16x3(const __bf16* A, const __bf16* B, double* WT, int nGp, int Pc, hipStream_t st) {
#define CALL_D(NT) hipLaunchKernelGGL(k_mf_bf16x3<NT>, dim3(nGp / 16), dim3(256), 0, st, A, B, WT, nGp, Pc)
  CODY_NT_SWITCH2(Pc, CALL_D)
#undef CALL_D
}

}  // namespace cody
