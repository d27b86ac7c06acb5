// Kernel launchers of the MCTDHX right-hand side (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cody {

int set_error(int code, const char* msg);  // records the library's last error, returns code

constexpr int kMaxOrbs = 12;   // orbitals per body type handled by the kernels
constexpr int kMaxStencilReach = 31;  // nLoc - 1 for the largest FEDVR element (nLoc <= 32)
constexpr int kMaxPc = 160;    // 16*ceil(M(M+1)/16) for M <= 12
// CI density role: each wave sums ci_rdm_items_per_wave(nC) entries of rho1 / canonical rho2, so
// that a workgroup stages C into LDS once for 4 ipw entries (a large CI space takes 4 per wave)
constexpr int kRdmWideConf = 256, kRdmWideItems = 4;
__host__ __device__ constexpr int ci_rdm_items_per_wave(int nC) { return nC > kRdmWideConf ? kRdmWideItems : 1; }
constexpr int kInt8Slices = 5; // balanced base-254 digits per operand of the int8-sliced contraction
// Digit base of the int8 slices.  Values scaled into [-1/2, 1/2] have balanced base-254 digits
// in [-127, 127] (254/2 = 127 fits int8 without a carry case); each digit carries log2 254 = 7.99 bits.
constexpr double kSliceBase = 254.0;
// kSliceScale[k] = 254^-k (correctly rounded at compile time)
struct SliceScales {
  double v[16];
  constexpr SliceScales() : v() {
    double s = 1.0;
    for (int k = 0; k < 16; ++k) {
      v[k] = s;
      s /= kSliceBase;
    }
  }
};
constexpr SliceScales kSliceScale{};
// right-hand-side modes: the full one, the reference's nonlinear split part, the linear part, and
// the consistent nonlinear part (full minus linear)
enum { kModeFull = 0, kModeNonLin = 1, kModeLin = 2, kModeNonLinConsistent = 3 };

struct CiEntryDev {
  int32_t ic, jc, code;
  float f;
};

struct DevCiTables {
  const int* s_ptr;
  const CiEntryDev* singles;
  const int* d_ptr;
  const CiEntryDev* doubles;
  const int* r1_ptr;
  const CiEntryDev* r1;
  const int* r2_ptr;
  const CiEntryDev* r2;
};

struct KinLaunch {
  const double* kref;
  int nE, nLoc, xminExcl;
  double invh, fac;
};

// One instance of a grouped right-hand side: independent systems with one orbital count are
// evaluated by one launch per kernel type, each workgroup taking its instance and its block index
// within that instance's own launch from a block map (int2: instance, local block).  The fields
// are the arguments of the fused launches of that instance.
struct GroupInstance {
  const double2* C;  // state: C, then the orbitals
  double2* dC;       // dState, same layout
  int nC, nG, nGp, Pc;
  int nRdm, nObx;         // front roles: CI density blocks, one-body blocks per orbital
  int nsplit;             // contraction k ranges (1: the contraction writes W^T itself)
  int wsumFirst;          // 1: the k-split-sum launch also forms R (workgroup 0)
  int nxs, nPart, nPart1, nRed;
  DevCiTables t;
  double2 *rho1, *r2can, *rho2, *hphi, *part1, *reginv, *draw, *part, *h1, *dCraw;
  const double *w, *x, *vext, *efield;
  KinLaunch kin;
  double reg;
  double *STt, *Wpart, *WT, *Gpart, *G, *epart;
  int* act;                  // pair-source activity flags (written by the front launch; FP64 path)
  int* actK;                 // int8 path: per-block activity flags (written by the front launch)
  const int* actC;           // the flags the contraction reads (null: no skipping)
  double2* partr;            // int8 path: the summed overlap partials (the projection reads these)
  double* qf;                // Q of the orbital correlation (nO >= kCorrQMinOrbs, FP64 path), or null
  const double* Vt;                          // FP64 contraction operand
  const int8_t* Vi8;                         // int8 contraction operands
  const int* eV;
  int8_t* Si8;  // int8 path: digits of S (written by the front launch), null on the FP64 path
  int* eSb;     // int8 path: their per-line, per-64-point-block exponents
};

// Grouped launches over the instances of one orbital count nO (the map holds nBlocks entries)
void launch_grp_front(int nO, const GroupInstance* gi, const int2* map, int nBlocks, int maxC, hipStream_t st);
void launch_grp_wsum(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st);
void launch_grp_qform(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st);
void launch_grp_mid(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st);
void launch_grp_post(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st);
void launch_grp_apply(const GroupInstance* gi, const int2* map, int nBlocks, int nO, hipStream_t st);
// the contractions, Pc = 16 NT for every instance of the group (CODY_ERR_ARG otherwise)
int launch_grp_mf_f64_w(int Pc, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st);
int launch_grp_mf_i8_r16(int Pc, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st, int nO);

void launch_zero(double2* p, int n, hipStream_t st);
// y += x (n complex values)
void launch_accumulate(double2* y, const double2* x, int n, hipStream_t st);
// src = conj(phi_a) phi_b w: one pair source (nG values)
void launch_pair_src_one(const double2* orbs, const double* w, int nG, int a, int b, double2* src, hipStream_t st);
// f(:, i) += (WT(0, :) + i WT(1, :)) orbs(:, i) for the nO orbitals: the first source's potential applied
void launch_apply_potential(const double* WT, int nGp, const double2* orbs, int nG, int nO, double2* f,
                            hipStream_t st);
void launch_fill_softyukawa(const double* x, int nG, int nGp, double s, double a1, double a2,
                            double damp, double* V, hipStream_t st);
void launch_ci_density(const double2* C, int nC, int nO, double reg, const DevCiTables& t,
                       double2* rho1, double2* rho2, double2* r2can, double2* reginv, hipStream_t st);
void launch_sum_parts(const double2* part, int nPart, int stride, int n, double2* out, hipStream_t st);
// operator pieces: sources in the tiled S^T layout, overlap partials <phi_i|f_j>, the full h2
// tensor from G, R from a given rho1
void launch_tile_src(const double2* src, int ns, int nG, int nGp, int Pc, double* STt, hipStream_t st);
void launch_overlap_parts(const double2* orbs, const double* w, const double2* f, int nG, int nO, double2* part,
                          hipStream_t st);
void launch_h2_full(const double* G, int Pc, int nO, double2* h2, hipStream_t st);
// f <- f - sum_i phi_i <phi_i|f>_w in place, from the ceil(nG / 256) x nO^2 partials of launch_overlap_parts
void launch_project_inplace(const double2* orbs, const double2* part, int nG, int nO, double2* f, hipStream_t st);
void launch_reginv(const double2* rho1, int nO, double reg, double2* reginv, hipStream_t st);
// position matrix partials (k_xmat) and <x> from them and rho1 (k_dipole)
void launch_xmat(const double2* orbs, const double* w, const double* x, int nG, int nO, double2* part1,
                 hipStream_t st);
void launch_dipole(int nO, const double2* part1, int nPart, const double2* rho1, double* dipole, hipStream_t st);
void launch_gram_reduce(const double* Gpart, int nxs, int Pc, double* G, const double2* part1, int nPart1,
                        int n2, double2* h1, hipStream_t st);
void launch_ci_apply(const double2* C, int nC, int nO, int Pc, const double* G, const double2* h1,
                     const DevCiTables& t, double2* dCraw, hipStream_t st);
void launch_ci_energy(int nO, int Pc, const double* G, const double2* h1, const double2* rho1,
                      const double2* rho2, double* energy, hipStream_t st);
// CI H-apply and CI projection in one launch: dCraw = H C, dC = -i (dCraw - E C), E the sum of the
// nE energy partials of launch_rhs_post (E = <C|H C>)
void launch_ci_apply_project(const double2* C, int nC, int nO, int Pc, const double* G, const double2* h1,
                             const DevCiTables& t, const double* epart, int nE, double2* dCraw, double2* dC,
                             hipStream_t st);
// Fused launches of the right-hand side (one queue):
//   front: CI densities, one-body part (h phi, h1 partials), pair sources (+ flags; with Si8 set
//          the int8 digits of S, their block exponents eSb and per-block flags actK instead of act);
//   wsum:  R = rho1 (rho1^2 + eps)^-1 and the k-split sum of W^T (nsplit <= 1: none);
//   mid:   Gram partials (nxs x-ranges) and the orbital correlation with its overlap partials;
//   post:  G and h1 from their partials with the energy partials epart (one per reduction
//          workgroup, rhs_post_energy_parts of them), and the orbital projection (times -i).
void launch_rhs_front(const double2* C, int nC, int nO, const DevCiTables& t, double2* rho1, double2* r2can,
                      double2* rho2, const double2* orbs, const double* w, const double* x, const double* vext,
                      const KinLaunch& kl, const double* efield, int nG, double2* hphi, double2* part1, int nGp,
                      int Pc, double* STt, int* act, int8_t* Si8, int* eSb, int* actK, hipStream_t st);
void launch_rhs_wsum(const double2* rho1, int nO, double reg, double2* reginv, const double* Wpart, int nsplit,
                     size_t plane, double* WT, hipStream_t st, bool withReginv = true);
void launch_rhs_mid(const double* STt, const double* WT, int nGp, int Pc, double* Gpart, int nxs,
                    const double2* orbs, const double* w, const double2* hphi, const double2* r2can,
                    const double2* reginv, int nG, int nO, int mode, double2* draw, double2* part, hipStream_t st,
                    const double* qf = nullptr, const int* act = nullptr, int actShift = 4);
int rhs_post_energy_parts(int Pc, int nO);
void launch_rhs_post(const double* Gpart, int nxs, int Pc, double* G, const double2* part1, int nPart1,
                     double2* h1, const double2* rho1, const double2* rho2, double* epart, const double2* orbs,
                     const double2* draw, const double2* part, int nPart, int nG, int nO, double2* dorbs,
                     const double2* hsub, hipStream_t st);
size_t ci_density_lds_bytes(int nC, int nO);
int ci_kernels_allow_lds(size_t bytes);
// pair sources S^T (fragment-tiled STt, optionally row-major ST) and per-fragment activity flags
// (act, optional: Pc/16 x nGp/16 ints, 1 = a non-zero value); with
// Si8 set, the int8 form's instead: S^T, its digits Si8, the per-line and 64-point-block exponents
// eSb (nGp/64 x Pc) and the per-block activity flags actK (Pc/16 x nGp/64)
void launch_pair_src(const double2* orbs, const double* w, int nG, int nGp, int nO, int Pc,
                     double* ST, double* STt, int* act, hipStream_t st,
                     int8_t* Si8 = nullptr, int* eSb = nullptr, int* actK = nullptr);
// row-major S^T alone (the operand of the split-FP32, BF16 and untiled int8 forms)
void launch_pair_src_rows(const double2* orbs, const double* w, int nG, int nGp, int nO, int Pc, double* ST,
                          hipStream_t st);
void launch_onebody(const double2* orbs, const double* w, const double* x, const double* vext,
                    const KinLaunch& kl, const double* efield, int nG, int nO, double2* hphi, double2* part1,
                    hipStream_t st);
void launch_set_scalar(double* p, double v, hipStream_t st);
void launch_orbital(const double2* orbs, const double* w, const double2* hphi, const double* WT,
                    const double2* r2can, const double2* reginv, int nG, int nGp, int nO, int mode,
                    double2* draw, double2* part, hipStream_t st);
void launch_slice_rows(const double* X, int len, int nlines, int ns, int tiled, int8_t* out, int* expo,
                       hipStream_t st);
void launch_split_f32(const double* X, size_t n, float* hi, float* lo, hipStream_t st);
void launch_split_bf16x3(const double* X, size_t n, __bf16* out, hipStream_t st);
constexpr int kMaxWSplit = 8;  // largest k split of the FP64 contraction (partials buffer depth)
// From this many orbitals the FP64 right-hand side runs the orbital correlation as an FP64 MFMA
// product of W (16 points x 2P real) with the rho2'-weighted matrix Q (2P x 2 nO^2 real), Q formed
// once per right-hand side in MFMA fragment order (corr_q_body, qform_body).
constexpr int kCorrQMinOrbs = 8;
constexpr int kCorrQGroups = 4;  // orbital groups of the correlation: workgroups per 64 points
// Q's k steps of 4 real rows (2P = nO (nO + 1) rows), its 16-column tiles per orbital j (columns
// re/im of k1 = 0..7, then 8..15) and its size in doubles: tile t = j tiles-per-j + th, step s,
// lane l at ((t KS + s) 64 + l)
constexpr int corr_q_steps(int nO) { return (nO * (nO + 1) + 3) / 4; }
constexpr int corr_q_tiles_per_orb(int nO) { return (2 * nO + 15) / 16; }
constexpr int corr_q_doubles(int nO) { return nO * corr_q_tiles_per_orb(nO) * corr_q_steps(nO) * 64; }
constexpr int corr_q_blocks(int nO) { return (corr_q_doubles(nO) / 4 + 255) / 256; }  // four doubles per thread
// Q beside the Gram partials (which the correlation launch then leaves out): the launch between the
// k-split sum (R) and launch_rhs_mid from kCorrQMinOrbs orbitals
void launch_rhs_qform(int nO, const double2* r2can, const double2* reginv, double* qf, const double* STt,
                      const double* WT, int nGp, int Pc, double* Gpart, int nxs, hipStream_t st,
                      const int* act = nullptr, int actShift = 4);
constexpr int kSkipMaxBlocks16 = 1024;  // longest grid (in 16-point blocks) whose zero k blocks are skipped
// shape check of the k-split contraction launchers (0 or CODY_ERR_ARG, nothing launched)
int check_contraction_shape(const char* who, int nGp, int Pc, int nsplit, int maxTiles, int kBlocks);
// act (optional): the pair-source activity flags of launch_pair_src; the k blocks whose sources
// are all zero are skipped (grids up to kSkipMaxBlocks16 16-point blocks; dense beyond)
int launch_mf_f64_w(const double* Vt, const double* STt, double* Wpart, double* WT, int nGp, int Pc,
                    int nsplit, hipStream_t st, const int* act = nullptr);
void launch_tile_f64(const double* V, int nGp, double* Vt, hipStream_t st);
void launch_gram_f64(const double* STt, const double* WT, double* Gpart, int nGp, int Pc, int nsplit,
                     hipStream_t st);
constexpr int kR16MaxBlocks64 = 256;  // longest grid (in 64-point blocks) of the 16-row int8 contraction
constexpr int kR16MaxTiles = 10;      // 16-column tiles of the 16-row int8 contraction (with the orbital epilogue: 5)
constexpr int kR16Prefetch = 2;       // k blocks of operands in flight per wave of the 16-row int8 contraction
// Orbital-correlation operands of the fused int8 contraction (launch_mf_i8_orb)
struct OrbCorrArgs {
  const double2* orbs;  // phi, nO x nG
  const double* w;      // FEDVR weights
  const double2* hphi;  // (T + V) phi
  const double2* r2can;  // canonical rho2 (a < b, m < l), np x np
  int nG, mode;
  double2* corrc;  // output: c_m(x) (the correlation before R), nO x nG
  double2* part;   // output: overlap partials, two nO^2 blocks per 16-point tile (h phi, then c)
};
int launch_mf_i8_orb(const int8_t* As, const int* eA, const int8_t* Bs, const int* eSb, const int* actK, double* WT,
                     int nGp, int nO, const OrbCorrArgs& o, hipStream_t st,
                     bool bf16 = false);
// after launch_mf_i8_orb: G = S^T W directly (a 1024-thread workgroup per 16 x 16 tile, over the
// active blocks of actK or every block) with its energy shares, h1 and its share, the fixed-order
// sums of the contraction's two sets of nPart overlap partials into partr (2 nO^2; one workgroup
// each) and R = rho1 (rho1^2 + eps)^-1 (one wave)
void launch_rhs_gram(const double* STt, const double* WT, int nGp, int Pc, int nO, const int* actK,
                     const double2* rho2, double* G, double* epart, const double2* part1, int nPart1, double2* h1,
                     const double2* part, int nPart, double2* partr, const double2* rho1, double reg,
                     double2* reginv, hipStream_t st);
// the energy shares the direct-G launch leaves (one per G tile, one for h1)
int gram_direct_energy_parts(int Pc);
// the fused int8 path's last launch: CI apply and projection | orbital projection with the R mixing
void launch_rhs_final(const double2* C, int nC, int nO, int Pc, const double* G, const double2* h1,
                      const DevCiTables& t, const double* epart, int nE, double2* dCraw, double2* dC,
                      const double2* orbs, const double2* h, const double2* corrc, const double2* reginv,
                      const double2* partr, int nG, double2* dorbs, const double2* hsub, hipStream_t st);
void launch_grp_final(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st);
int launch_grp_mf_i8_orb(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st);
void launch_grp_gram(int nO, const GroupInstance* gi, const int2* map, int nBlocks, hipStream_t st);
int launch_mf_i8_r16(const int8_t* As, const int* eA, const int8_t* Bs, const int* eSb, const int* actK, double* WT,
                     int nGp, int Pc, hipStream_t st,
                     bool bf16 = false,
                     const double2* orbs = nullptr, int nG = 0, int nO = 0);
void launch_mf_i8(const int8_t* As, const int* eA, const int8_t* Bs, const int* fB, double* WT, int nGp,
                  int Pc, hipStream_t st);
void launch_mf_f32x3(const float* Ah, const float* Al, const float* Bh, const float* Bl, double* WT, int nGp,
                     int Pc, hipStream_t st);
void launch_mf_bf16x3(const __bf16* A, const __bf16* B, double* WT, int nGp, int Pc, hipStream_t st);
// dorbs = -i [(1 - P) draw - hsub] (hsub optional), and the CI projection of dCraw into dC
void launch_project(const double2* orbs, const double2* draw, const double2* part, int nPart, int nG, int nO,
                    int project, double2* dorbs, const double2* C, const double2* dCraw, int nC, double2* dC,
                    const double2* hsub, hipStream_t st);

}  // namespace cody
