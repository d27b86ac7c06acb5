// C ABI of the MI355X MCTDHX path: context, setup, right-hand side orchestration,
// operator pieces and measurement.  Declared in include/cody_mctdhx.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <functional>
#include <string>
#include <vector>

#include "../../include/cody_mctdhx.h"
#include "batch_plan.hpp"
#include "host_setup.hpp"
#include "mctdhx_kernels.hpp"
#include "cody_ctx.hpp"

using namespace cody;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(CODY_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

#define CHK(expr)           \
  do {                      \
    int r_ = (expr);        \
    if (r_) return r_;      \
  } while (0)

namespace cody {

int ctx_alloc(void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(CODY_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemset(*p, 0, bytes);
  // hipMemset may return before the fill has run, and the context's streams are non-blocking
  // (not ordered after it): wait here, or a kernel launched on them right after the allocation
  // (the V tiling, the propagator's first step) could have its output zeroed by the late fill
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail(CODY_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
  return CODY_OK;
}

int set_error(int code, const char* msg) { return fail(code, msg); }

// Allocates the operands of the selected contraction precision and splits V (once per V).
int prepare_precision(cody_ctx* c) {
  const size_t nv = (size_t)c->nGp * c->nGp, ns = (size_t)c->Pc * c->nGp;
  int r = 0;
  switch (c->precision) {
    case CODY_PREC_INT8_SLICED:
    case CODY_PREC_BF16_SLICED:
      if (!c->d_Vi8) {
        // (eS: per-line exponents of the untiled form; eSb, actK: per-line-and-block exponents and
        // per-block flags of the tiled form)
        const size_t nb = (size_t)(c->nGp / 64) * c->Pc;
        if ((r = ctx_alloc((void**)&c->d_Vi8, nv * kInt8Slices)) || (r = ctx_alloc((void**)&c->d_Si8, ns * kInt8Slices)) ||
            (r = ctx_alloc((void**)&c->d_eV, sizeof(int) * c->nGp)) || (r = ctx_alloc((void**)&c->d_eS, sizeof(int) * c->Pc)) ||
            (r = ctx_alloc((void**)&c->d_eSb, sizeof(int) * nb)) || (r = ctx_alloc((void**)&c->d_actK, sizeof(int) * nb / 16)))
          return r;
      }
      if (c->int_set && !c->vprec_ready)
        launch_slice_rows(c->d_V, c->nGp, c->nGp, kInt8Slices, 1, c->d_Vi8, c->d_eV, c->stream);
      break;
    case CODY_PREC_SPLIT_FP32:
      if (!c->d_Vhi) {
        if ((r = ctx_alloc((void**)&c->d_Vhi, nv * sizeof(float))) || (r = ctx_alloc((void**)&c->d_Vlo, nv * sizeof(float))) ||
            (r = ctx_alloc((void**)&c->d_Shi, ns * sizeof(float))) || (r = ctx_alloc((void**)&c->d_Slo, ns * sizeof(float))))
          return r;
      }
      if (c->int_set && !c->vprec_ready) launch_split_f32(c->d_V, nv, c->d_Vhi, c->d_Vlo, c->stream);
      break;
    case CODY_PREC_BF16X3:
      if (!c->d_Vbf) {
        if ((r = ctx_alloc((void**)&c->d_Vbf, 3 * nv * sizeof(__bf16))) || (r = ctx_alloc((void**)&c->d_Sbf, 3 * ns * sizeof(__bf16))))
          return r;
      }
      if (c->int_set && !c->vprec_ready) launch_split_bf16x3(c->d_V, nv, c->d_Vbf, c->stream);
      break;
    default:
      if (!c->d_Vt && (r = ctx_alloc((void**)&c->d_Vt, nv * sizeof(double)))) return r;
      if (c->int_set && !c->vprec_ready) launch_tile_f64(c->d_V, c->nGp, c->d_Vt, c->stream);
      break;
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(CODY_ERR_HIP, std::string("prepare_precision: ") + hipGetErrorString(e));
  c->vprec_ready = c->int_set;
  return CODY_OK;
}

// The int8 form on the 16-row contraction over the active k blocks, S's digits fragment-tiled and
// formed by the pair-source kernel (Pc <= 160, grids up to kR16MaxBlocks64 64-point blocks); else
// the untiled int8 kernel, which like the other reduced-precision forms reads row-major S^T.
bool int8_tiled(const cody_ctx* c) {
  return (c->precision == CODY_PREC_INT8_SLICED || c->precision == CODY_PREC_BF16_SLICED) &&
         c->Pc <= 16 * kR16MaxTiles && c->nGp / 64 <= kR16MaxBlocks64;
}
// ... with the orbital correlation in the contraction's epilogue (up to 8 orbitals, 80 pair columns);
// above that the tiled contraction writes W^T and the FP64 form's correlation follows.
bool int8_fused(const cody_ctx* c) { return int8_tiled(c) && c->nO <= 8; }
double* plain_st(cody_ctx* c) { return (c->precision == CODY_PREC_FP64 || int8_tiled(c)) ? nullptr : c->d_ST; }
// activity flags of the pair sources for the contraction (null: every k block is contracted):
// per 16 x 16 fragment for the FP64 form, per 16 x 64 block for the tiled int8 form
const int* act_flags(const cody_ctx* c) {
  if (!c->skip_zero) return nullptr;
  return int8_tiled(c) ? c->d_actK : c->d_act;
}
// the flags the Gram partials skip by (as the grouped launches take them: the FP64 contraction's,
// which it leaves out above kSkipMaxBlocks16 16-point blocks) and their block size (2^shift points)
const int* gram_flags(const cody_ctx* c) {
  return (!int8_tiled(c) && c->nGp / 16 > kSkipMaxBlocks16) ? nullptr : act_flags(c);
}
int gram_flag_shift(const cody_ctx* c) { return int8_tiled(c) ? 6 : 4; }

// The pair sources of orbs for the context's contraction: S^T fragment-tiled with its activity
// flags, with the int8 digits where the tiled int8 contraction reads them, and row-major S^T for
// the forms that read that.
void enqueue_pair_src(cody_ctx* c, const double2* orbs, hipStream_t st) {
  if (int8_tiled(c))
    launch_pair_src(orbs, c->d_w, c->nG, c->nGp, c->nO, c->Pc, nullptr, c->d_STt, nullptr, st, c->d_Si8, c->d_eSb,
                    c->d_actK);
  else
    launch_pair_src(orbs, c->d_w, c->nG, c->nGp, c->nO, c->Pc, plain_st(c), c->d_STt, c->d_act, st);
}

// Enqueue W = V S for the context's precision (ST already formed).  With partialSplit null, W^T
// is complete at the end; otherwise a k-split contraction leaves its partial planes in d_Wpart
// and *partialSplit receives their count (0 when W^T itself was written), for a later sum.
// Returns 0, or the launcher's error (nothing of the contraction launched) when the shape is
// outside what the kernels take.
int enqueue_contraction(cody_ctx* c, hipStream_t st, const std::function<void(const char*)>& mark,
                        int* partialSplit = nullptr, const OrbCorrArgs* orb = nullptr,
                        const double2* rhsOrbs = nullptr) {
  const size_t ns = (size_t)c->Pc * c->nGp;
  if (partialSplit) *partialSplit = 0;
  const bool bf = c->precision == CODY_PREC_BF16_SLICED;
  switch (c->precision) {
    case CODY_PREC_INT8_SLICED:
    case CODY_PREC_BF16_SLICED:
      if (int8_tiled(c)) {  // 16-row kernel over the active k blocks (S's digits from the pair sources)
        if (orb) {
          CHK(launch_mf_i8_orb(c->d_Vi8, c->d_eV, c->d_Si8, c->d_eSb, act_flags(c), c->d_WT, c->nGp, c->nO, *orb,
                               st, bf));
          mark(bf ? "k_mf_i8_orb<bf16>" : "k_mf_i8_orb");
        } else {
          // (in a right-hand side, tiles with no orbital value skip W)
          CHK(launch_mf_i8_r16(c->d_Vi8, c->d_eV, c->d_Si8, c->d_eSb, act_flags(c), c->d_WT, c->nGp, c->Pc, st, bf,
                               rhsOrbs, c->nG, c->nO));
          mark(bf ? "k_mf_i8_r16<bf16>" : "k_mf_i8_r16");
        }
        return CODY_OK;
      }
      if (bf) return set_error(CODY_ERR_UNSUPPORTED, "the BF16-sliced contraction takes at most 80 pair columns");
      launch_slice_rows(c->d_ST, c->nGp, c->Pc, kInt8Slices, 0, c->d_Si8, c->d_eS, st);
      mark("k_slice_rows");
      launch_mf_i8(c->d_Vi8, c->d_eV, c->d_Si8, c->d_eS, c->d_WT, c->nGp, c->Pc, st);
      mark("k_mf_i8");
      return CODY_OK;
    case CODY_PREC_SPLIT_FP32:
      launch_split_f32(c->d_ST, ns, c->d_Shi, c->d_Slo, st);
      mark("k_split_f32");
      launch_mf_f32x3(c->d_Vhi, c->d_Vlo, c->d_Shi, c->d_Slo, c->d_WT, c->nGp, c->Pc, st);
      mark("k_mf_f32x3");
      return CODY_OK;
    case CODY_PREC_BF16X3:
      launch_split_bf16x3(c->d_ST, ns, c->d_Sbf, st);
      mark("k_split_bf16x3");
      launch_mf_bf16x3(c->d_Vbf, c->d_Sbf, c->d_WT, c->nGp, c->Pc, st);
      mark("k_mf_bf16x3");
      return CODY_OK;
    default: {
      const bool keep = partialSplit && c->wsplit > 1;  // (a single k range writes W^T itself)
      CHK(launch_mf_f64_w(c->d_Vt, c->d_STt, c->d_Wpart, keep ? nullptr : c->d_WT, c->nGp, c->Pc, c->wsplit, st,
                          act_flags(c)));
      if (keep) *partialSplit = c->wsplit;
      mark(keep ? "k_mf_f64_w" : "k_mf_f64_w + k_wsum");
      return CODY_OK;
    }
  }
}

}  // namespace cody

extern "C" {

const char* cody_last_error(void) { return g_err.c_str(); }

int cody_fedvr_npoints(int nE, int nLoc, int xminExcl, int xmaxExcl) {
  if (nE < 1 || nLoc < 2) return -1;
  return fedvr_npoints(nE, nLoc, xminExcl != 0, xmaxExcl != 0);
}

int cody_fedvr_grid(double xmin, double xmax, int nE, int nLoc, int xminExcl, int xmaxExcl,
                    double* x, double* w) {
  if (nE < 1 || nLoc < 2 || !(xmax > xmin) || !x || !w) return fail(CODY_ERR_ARG, "cody_fedvr_grid: bad arguments");
  fedvr_grid(xmin, xmax, nE, nLoc, xminExcl != 0, xmaxExcl != 0, x, w);
  return CODY_OK;
}

int cody_fedvr_kref(int nLoc, double* kref) {
  if (nLoc < 2 || !kref) return fail(CODY_ERR_ARG, "cody_fedvr_kref: bad arguments");
  fedvr_kref(nLoc, kref);
  return CODY_OK;
}

int cody_create(cody_ctx** out, int device, int nG, int nBodies, int nOrbs, double reg) {
  if (!out) return fail(CODY_ERR_ARG, "cody_create: null output");
  *out = nullptr;
  if (nG < 2 || nOrbs < 1 || nOrbs > kMaxOrbs || nBodies < 1 || nBodies > nOrbs)
    return fail(CODY_ERR_UNSUPPORTED, "cody_create: need nG >= 2 and 1 <= nBodies <= nOrbs <= 12");
  if (!(reg >= 0.0)) return fail(CODY_ERR_ARG, "cody_create: regularization must be >= 0");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(CODY_ERR_ARG, "cody_create: no such device");
  HIPCHK(hipSetDevice(device));
  auto* c = new cody_ctx();
  c->device = device;
  c->nG = nG;
  c->nGp = 256 * ((nG + 255) / 256);
  c->N = nBodies;
  c->nO = nOrbs;
  c->reg = reg;
  c->P = nOrbs * (nOrbs + 1) / 2;
  c->Pc = 16 * ((2 * c->P + 15) / 16);
  std::string err;
  if (!build_ci_tables(nOrbs, nBodies, c->ci, err)) {
    delete c;
    return fail(CODY_ERR_UNSUPPORTED, err);
  }
  c->nC = c->ci.nConf;
  c->nS = c->nC + nG * nOrbs;
  if (c->nC > 1024) {
    delete c;
    return fail(CODY_ERR_UNSUPPORTED, "cody_create: more than 1024 configurations");
  }
  // split-K factor of the mean-field contraction and x-split of the h2 Gram kernel
  const int k16 = c->nGp / 16;
  c->nsplit = 1;
  for (int s : {8, 7, 6, 5, 4, 3, 2, 1})
    if (k16 % s == 0) {
      c->nsplit = s;
      break;
    }
  // split-K factor of the Gram matrix G = S^T W (nGp / (64 nxs) must be an integer); at most 8
  // above 80 pair columns, where each workgroup's 16 x Pc tile has enough MFMA work per k range
  const int k64 = c->nGp / 64, maxXs = c->Pc > 80 ? 8 : 32;
  c->nxs = 1;
  for (int s : {32, 16, 8, 4, 2, 1})
    if (s <= maxXs && k64 % s == 0) {
      c->nxs = s;
      break;
    }
  c->nPart = c->nGp / 64;  // one partial per 64-point workgroup of k_orbital
  {
    const size_t lds = ci_density_lds_bytes(c->nC, nOrbs);
    if (lds > 160 * 1024) {
      delete c;
      return fail(CODY_ERR_UNSUPPORTED, "cody_create: CI workspace exceeds the 160 KiB of LDS");
    }
    if (lds > 64 * 1024 && ci_kernels_allow_lds(lds)) {
      delete c;
      return fail(CODY_ERR_HIP, "cody_create: could not raise the LDS limit of the CI kernels");
    }
  }
  int r = c->allocate();
  if (r) {
    delete c;
    return r;
  }
  *out = c;
  return CODY_OK;
}

int cody_destroy(cody_ctx* c) {
  if (!c) return CODY_OK;
  hipSetDevice(c->device);
  c->release();
  delete c;
  return CODY_OK;
}

int cody_sizes(const cody_ctx* c, int* nC, int* nState) {
  if (!c) return fail(CODY_ERR_ARG, "cody_sizes: null context");
  if (nC) *nC = c->nC;
  if (nState) *nState = c->nS;
  return CODY_OK;
}

int cody_get_stream(const cody_ctx* c, void** stream) {
  if (!c || !stream) return fail(CODY_ERR_ARG, "cody_get_stream: null argument");
  *stream = c->stream;
  return CODY_OK;
}

int cody_set_grid(cody_ctx* c, const double* x, const double* w) {
  if (!c || !x || !w) return fail(CODY_ERR_ARG, "cody_set_grid: null argument");
  for (int i = 0; i < c->nG; ++i)
    if (!(w[i] > 0.0)) return fail(CODY_ERR_ARG, "cody_set_grid: weights must be positive");
  HIPCHK(hipSetDevice(c->device));
  c->hx.assign(x, x + c->nG);
  c->hw.assign(w, w + c->nG);
  HIPCHK(hipMemcpy(c->d_x, x, sizeof(double) * c->nG, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_w, w, sizeof(double) * c->nG, hipMemcpyHostToDevice));
  c->grid_set = true;
  c->adj.K = 0;  // the Chebyshev bounds of the one-body operator are stale
  c->drop_graphs();
  c->int_set = false;  // the interaction matrix depends on the grid
  return CODY_OK;
}

int cody_set_kinetic_fedvr(cody_ctx* c, double xmin, double xmax, int nE, int nLoc, int xminExcl,
                           int xmaxExcl, double mass) {
  if (!c) return fail(CODY_ERR_ARG, "cody_set_kinetic_fedvr: null context");
  if (nE < 1 || nLoc < 2 || nLoc > 32 || !(xmax > xmin) || !(mass > 0.0))
    return fail(CODY_ERR_ARG, "cody_set_kinetic_fedvr: bad arguments");
  if (fedvr_npoints(nE, nLoc, xminExcl != 0, xmaxExcl != 0) != c->nG)
    return fail(CODY_ERR_ARG, "cody_set_kinetic_fedvr: element layout does not match nG");
  HIPCHK(hipSetDevice(c->device));
  std::vector<double> kref(nLoc * nLoc);
  fedvr_kref(nLoc, kref.data());
  HIPCHK(hipMemcpy(c->d_kref, kref.data(), sizeof(double) * nLoc * nLoc, hipMemcpyHostToDevice));
  c->kin = KinLaunch{c->d_kref, nE, nLoc, xminExcl != 0 ? 1 : 0, nE / (xmax - xmin), -0.5 / mass};
  c->kin_set = true;
  c->adj.K = 0;  // the Chebyshev bounds of the one-body operator are stale
  c->drop_graphs();
  return CODY_OK;
}

int cody_set_potential(cody_ctx* c, const double* vext) {
  if (!c || !vext) return fail(CODY_ERR_ARG, "cody_set_potential: null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpy(c->d_vext, vext, sizeof(double) * c->nG, hipMemcpyHostToDevice));
  c->pot_set = true;
  c->adj.K = 0;  // the Chebyshev bounds of the one-body operator are stale
  c->drop_graphs();
  return CODY_OK;
}

int cody_set_potential_softyukawa(cody_ctx* c, int n, const double* pos, const double* q,
                                  const double* a1, const double* a2, const double* d) {
  if (!c || n < 0 || (n > 0 && (!pos || !q || !a1 || !a2 || !d)))
    return fail(CODY_ERR_ARG, "cody_set_potential_softyukawa: bad arguments");
  if (!c->grid_set) return fail(CODY_ERR_STATE, "cody_set_potential_softyukawa: set the grid first");
  std::vector<double> v(c->nG, 0.0);
  for (int i = 0; i < c->nG; ++i)
    for (int k = 0; k < n; ++k) {
      const double r = std::fabs(c->hx[i] - pos[k]);
      v[i] -= q[k] * std::exp(-r * d[k]) / (std::sqrt(r * r + a1[k] * a1[k]) + a2[k]);
    }
  return cody_set_potential(c, v.data());
}

int cody_set_interaction_softyukawa(cody_ctx* c, double s, double a1, double a2, double damp) {
  if (!c) return fail(CODY_ERR_ARG, "cody_set_interaction_softyukawa: null context");
  if (!c->grid_set) return fail(CODY_ERR_STATE, "cody_set_interaction_softyukawa: set the grid first");
  if (a1 == 0.0 && a2 <= 0.0)
    return fail(CODY_ERR_ARG, "cody_set_interaction_softyukawa: kernel is singular at r = 0");
  HIPCHK(hipSetDevice(c->device));
  launch_fill_softyukawa(c->d_x, c->nG, c->nGp, s, a1, a2, damp, c->d_V, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->int_set = true;
  c->drop_graphs();
  c->vprec_ready = false;
  return prepare_precision(c);
}

int cody_set_laser(cody_ctx* c, int nK, const double* coef, double T) {
  if (!c || nK < 0 || nK > 64 || (nK > 0 && (!coef || !(T > 0.0))))
    return fail(CODY_ERR_ARG, "cody_set_laser: bad arguments");
  c->nLaser = nK;
  c->laserT = T;
  for (int k = 0; k < nK; ++k) c->laserC[k] = coef[k];
  if (nK == 0) {  // the field stays zero: no per-call update of the device scalar
    HIPCHK(hipSetDevice(c->device));
    const double zero = 0.0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->d_efield, &zero, sizeof(double), hipMemcpyHostToDevice));
  }
  return CODY_OK;
}

int cody_set_precision(cody_ctx* c, int mode) {
  if (!c) return fail(CODY_ERR_ARG, "cody_set_precision: null context");
  if (mode < CODY_PREC_FP64 || mode > CODY_PREC_BF16_SLICED)
    return fail(CODY_ERR_ARG, "cody_set_precision: unknown mode");
  if (mode == CODY_PREC_BF16_SLICED && (c->Pc > 80 || c->nGp / 64 > kR16MaxBlocks64))
    return fail(CODY_ERR_UNSUPPORTED, "cody_set_precision: the BF16-sliced contraction takes at most 80 pair "
                                      "columns (8 orbitals) and 16384 padded grid points");
  HIPCHK(hipSetDevice(c->device));
  c->precision = mode;
  c->vprec_ready = false;
  c->drop_graphs();
  return prepare_precision(c);
}

int cody_upload_state(cody_ctx* c, const double* state) {
  if (!c || !state) return fail(CODY_ERR_ARG, "cody_upload_state: null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpy(c->d_state, state, sizeof(double2) * c->nS, hipMemcpyHostToDevice));
  return CODY_OK;
}

int cody_download_state(cody_ctx* c, double* state) {
  if (!c || !state) return fail(CODY_ERR_ARG, "cody_download_state: null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(state, c->d_state, sizeof(double2) * c->nS, hipMemcpyDeviceToHost));
  return CODY_OK;
}

int cody_state_device_ptr(cody_ctx* c, void** d_state) {
  if (!c || !d_state) return fail(CODY_ERR_ARG, "cody_state_device_ptr: null argument");
  *d_state = c->d_state;
  return CODY_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// RHS orchestration
// ------------------------------------------------------------------------------------------
namespace cody {

double laser_field(const cody_ctx* c, double t) {
  if (c->nLaser <= 0 || t < 0.0 || t > c->laserT) return 0.0;
  double env = std::sin(M_PI * t / c->laserT);
  env *= env;
  double e = 0.0;
  for (int k = 0; k < c->nLaser; ++k) e += c->laserC[k] * std::sin((k + 1) * M_PI * t / c->laserT);
  return e * env;
}

int check_ready(const cody_ctx* c) {
  if (!c->grid_set) return set_error(CODY_ERR_STATE, "grid not set");
  if (!c->kin_set) return set_error(CODY_ERR_STATE, "kinetic operator not set");
  if (!c->pot_set) return set_error(CODY_ERR_STATE, "external potential not set");
  if (!c->int_set) return set_error(CODY_ERR_STATE, "interaction not set");
  return CODY_OK;
}

const double* efield_src(const cody_ctx* c) { return c->efield_src ? c->efield_src : c->d_efield; }

// Enqueue one right-hand side on `st`.  mode: kModeFull / kModeNonLin / kModeLin.
// When `ev` is non-null, an event is recorded after each launch (profiling).
int enqueue_rhs(cody_ctx* c, const double2* state, double2* dstate, double t, int mode,
                hipStream_t st, double* energy, RhsTrace* tr) {
  const double2* C = state;
  const double2* orbs = state + c->nC;
  double2* dC = dstate;
  double2* dorbs = dstate + c->nC;
  auto mark = [&](const char* name) {
    if (tr) {
      hipEvent_t e;
      hipEventCreate(&e);
      hipEventRecord(e, st);
      tr->ev.push_back(e);
      if (name) tr->names.emplace_back(name);
    }
  };
  mark(nullptr);
  // the nonlinear split part of the context's split: the reference's or full minus linear
  if (mode == kModeNonLin && c->split_consistent) mode = kModeNonLinConsistent;
  // E(t) is read from device memory, so that a captured right-hand side stays valid for any t
  if (c->nLaser > 0 && !c->capturing) launch_set_scalar(c->d_efield, laser_field(c, t), st);
  const int n2 = c->nO * c->nO, nPart1 = (c->nG + 255) / 256;
  if (mode == kModeLin) {
    // TimeDerivativeOrbsLin: dphi = -i (T + V) phi, CI frozen (S_Method_Mb_OrbBased_Mctdhx.f90:152-177)
    launch_onebody(orbs, c->d_w, c->d_x, c->d_vext, c->kin, efield_src(c), c->nG, c->nO, c->d_draw, c->d_part1,
                   st);
    mark("k_onebody");
    launch_project(orbs, c->d_draw, nullptr, 0, c->nG, c->nO, 0, dorbs, nullptr, nullptr, 0, nullptr, nullptr, st);
    mark("k_project");
    launch_zero(dC, c->nC, st);
    mark("k_zero");
    return CODY_OK;
  }
  // One queue, few launches: independent pieces share a launch as workgroup roles.
  //   front  CI densities | one-body part and h1 partials | pair sources and activity flags
  //          (the tiled int8 form: with S's digits)
  //   W = V S, leaving its k-split partials
  //   wsum   R = rho1 (rho1^2 + eps)^-1 | k-split sum of W^T
  //   mid    Gram partials of G = S^T W | orbital correlation and projector-overlap partials
  //   post   G and h1 from their partials | orbital projection
  //   post   also the energy E = <C|H C> in partial sums
  //   final  CI H-apply and CI projection with E
  // The tiled int8 form's contraction also forms the orbital correlation of its points before R
  // (c_m(x) and its overlap partials); then one launch forms the Gram partials, the partial sums
  // and R, and the projection applies R.
  const bool fused = int8_fused(c);
  launch_rhs_front(C, c->nC, c->nO, c->dci, c->d_rho1, c->d_r2can, c->d_rho2, orbs, c->d_w, c->d_x, c->d_vext, c->kin,
                   efield_src(c), c->nG, c->d_hphi, c->d_part1, c->nGp, c->Pc, c->d_STt, c->d_act,
                   int8_tiled(c) ? c->d_Si8 : nullptr, c->d_eSb, c->d_actK, st);
  mark("k_rhs_front");
  if (double* sr = plain_st(c)) launch_pair_src_rows(orbs, c->d_w, c->nG, c->nGp, c->nO, c->Pc, sr, st);
  int ps = 0;
  const OrbCorrArgs oa{orbs, c->d_w, c->d_hphi, c->d_r2can, c->nG, mode, c->d_draw, c->d_part};
  CHK(enqueue_contraction(c, st, mark, &ps, fused ? &oa : nullptr, orbs));
  if (fused) {
    // G directly with its energy shares | h1 | the overlap-partial sums | R; then the CI apply beside
    // the orbital projection (which does not depend on G)
    launch_rhs_gram(c->d_STt, c->d_WT, c->nGp, c->Pc, c->nO, gram_flags(c), c->d_rho2, c->d_G, c->d_epart, c->d_part1,
                    nPart1, c->d_h1, c->d_part, c->nGp / 16, c->d_partr, c->d_rho1, c->reg, c->d_reginv, st);
    mark("k_rhs_gram");
    const bool hOn = mode == kModeFull || mode == kModeNonLinConsistent;
    launch_rhs_final(C, c->nC, c->nO, c->Pc, c->d_G, c->d_h1, c->dci, c->d_epart, gram_direct_energy_parts(c->Pc),
                     c->d_dCraw, dC, orbs, hOn ? c->d_hphi : nullptr, c->d_draw, c->d_reginv, c->d_partr, c->nG,
                     dorbs, mode == kModeNonLinConsistent ? c->d_hphi : nullptr, st);
    mark("k_rhs_final");
    if (energy) launch_ci_energy(c->nO, c->Pc, c->d_G, c->d_h1, c->d_rho1, c->d_rho2, energy, st);
    return CODY_OK;
  } else {
    launch_rhs_wsum(c->d_rho1, c->nO, c->reg, c->d_reginv, c->d_Wpart, ps, (size_t)c->Pc * c->nGp, c->d_WT, st);
    mark("k_rhs_wsum");
    if (c->nO >= kCorrQMinOrbs) {
      launch_rhs_qform(c->nO, c->d_r2can, c->d_reginv, c->d_qf, c->d_STt, c->d_WT, c->nGp, c->Pc, c->d_Gpart, c->nxs,
                       st, gram_flags(c), gram_flag_shift(c));
      mark("k_rhs_qform");
    }
    launch_rhs_mid(c->d_STt, c->d_WT, c->nGp, c->Pc, c->d_Gpart, c->nxs, orbs, c->d_w, c->d_hphi, c->d_r2can,
                   c->d_reginv, c->nG, c->nO, mode, c->d_draw, c->d_part, st, c->d_qf, gram_flags(c),
                   gram_flag_shift(c));
    mark("k_rhs_mid");
  }
  launch_rhs_post(c->d_Gpart, c->nxs, c->Pc, c->d_G, c->d_part1, nPart1, c->d_h1, c->d_rho1, c->d_rho2, c->d_epart,
                  orbs, c->d_draw, c->d_part, c->nPart, c->nG, c->nO, dorbs,
                  mode == kModeNonLinConsistent ? c->d_hphi : nullptr, st);
  mark("k_rhs_post");
  launch_ci_apply_project(C, c->nC, c->nO, c->Pc, c->d_G, c->d_h1, c->dci, c->d_epart,
                          rhs_post_energy_parts(c->Pc, c->nO), c->d_dCraw, dC, st);
  mark("k_ci_apply_project");
  if (energy) launch_ci_energy(c->nO, c->Pc, c->d_G, c->d_h1, c->d_rho1, c->d_rho2, energy, st);
  return CODY_OK;
}

// Right-hand side through a captured graph of enqueue_rhs (captured on the context's own stream
// on first use for these buffers and mode, replayed on `st`); plain launches when graphs are off
// or capture fails.  The graph holds no memory-set nodes.
int run_rhs(cody_ctx* c, const double2* state, double2* dstate, double t, int mode, hipStream_t st) {
  if (c->graphs_on && c->rhs_graphs) {
    hipGraphExec_t ex = nullptr;
    for (const auto& g : c->graphs)
      if (g.state == state && g.dstate == dstate && g.mode == mode && g.energy == nullptr) ex = g.exec;
    if (!ex) {
      hipStream_t cs = c->stream;
      if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        c->capturing = true;
        const int r = enqueue_rhs(c, state, dstate, t, mode, cs, nullptr, nullptr);
        c->capturing = false;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (r == CODY_OK && e == hipSuccess && g &&
            hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess) {
          if (c->graphs.size() >= 16) {
            hipGraphExecDestroy(c->graphs.front().exec);
            c->graphs.erase(c->graphs.begin());
          }
          c->graphs.push_back(RhsGraph{state, dstate, mode, nullptr, ex});
        } else {
          ex = nullptr;
          c->graphs_on = false;  // capture unsupported here: plain launches from now on
        }
        if (g) hipGraphDestroy(g);
        (void)hipGetLastError();
      } else {
        c->graphs_on = false;
      }
    }
    if (ex) {
      if (c->nLaser > 0) launch_set_scalar(c->d_efield, laser_field(c, t), st);
      if (hipGraphLaunch(ex, st) != hipSuccess) return set_error(CODY_ERR_HIP, "hipGraphLaunch failed");
      return CODY_OK;
    }
  }
  return enqueue_rhs(c, state, dstate, t, mode, st, nullptr, nullptr);
}

}  // namespace cody

extern "C" {

int cody_autotune(cody_ctx* c, int* wsplit) {
  if (!c) return fail(CODY_ERR_ARG, "cody_autotune: null context");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  HIPCHK(hipStreamSynchronize(st));
  const int kBlocksF64 = c->nGp / 16 / (c->Pc / 16 <= 5 ? 2 : 1);  // k blocks the FP64 kernel splits
  if (wsplit && *wsplit > 0) {  // a given split instead of the timed choice
    if (check_contraction_shape("cody_autotune", c->nGp, c->Pc, *wsplit, 10, kBlocksF64)) return CODY_ERR_ARG;
    c->wsplit = *wsplit;
    c->drop_graphs();
    return CODY_OK;
  }
  if (c->precision == CODY_PREC_FP64) {
    enqueue_pair_src(c, c->d_state + c->nC, st);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    int pick = c->wsplit;
    for (int ks : {1, 2, 4, 8}) {
      // (splits the shape does not take are skipped; the check launches nothing)
      if (check_contraction_shape("k_mf_f64_w", c->nGp, c->Pc, ks, 10, kBlocksF64)) continue;
      CHK(launch_mf_f64_w(c->d_Vt, c->d_STt, c->d_Wpart, c->d_WT, c->nGp, c->Pc, ks, st, act_flags(c)));  // warm-up
      HIPCHK(hipEventRecord(e0, st));
      for (int i = 0; i < 8; ++i)
        CHK(launch_mf_f64_w(c->d_Vt, c->d_STt, c->d_Wpart, c->d_WT, c->nGp, c->Pc, ks, st, act_flags(c)));
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        pick = ks;
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c->wsplit = pick;
    c->drop_graphs();
  }
  if (wsplit) *wsplit = c->wsplit;
  return CODY_OK;
}

int cody_set_graphs(cody_ctx* c, int enable) {
  if (!c) return fail(CODY_ERR_ARG, "cody_set_graphs: null context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->drop_graphs();
  c->graphs_on = c->rhs_graphs = enable != 0;
  return CODY_OK;
}

int cody_set_source_skip(cody_ctx* c, int enable) {
  if (!c) return fail(CODY_ERR_ARG, "cody_set_source_skip: null context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->skip_zero = enable != 0;
  c->drop_graphs();  // captured contractions bake the flag pointer in
  return CODY_OK;
}

int cody_set_split(cody_ctx* c, int mode) {
  if (!c || (mode != CODY_SPLIT_REFERENCE && mode != CODY_SPLIT_CONSISTENT))
    return fail(CODY_ERR_ARG, "cody_set_split: unknown split");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->split_consistent = mode == CODY_SPLIT_CONSISTENT;
  c->drop_graphs();  // captured nonlinear parts and Strang steps bake the split in
  return CODY_OK;
}

static int rhs_dev(cody_ctx* c, const void* d_state, void* d_dstate, double t, void* stream,
                   int mode) {
  if (!c || !d_state || !d_dstate) return fail(CODY_ERR_ARG, "time derivative: null argument");
  if (d_state == d_dstate) return fail(CODY_ERR_ARG, "time derivative: state and dState alias");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  CHK(run_rhs(c, static_cast<const double2*>(d_state), static_cast<double2*>(d_dstate), t, mode, st));
  HIPCHK(hipGetLastError());
  return CODY_OK;
}

int cody_time_derivative_dev(cody_ctx* c, const void* s, void* d, double t, void* stream) {
  return rhs_dev(c, s, d, t, stream, kModeFull);
}
int cody_time_derivative_orbs_lin_dev(cody_ctx* c, const void* s, void* d, double t, void* stream) {
  return rhs_dev(c, s, d, t, stream, kModeLin);
}
int cody_time_derivative_nonlin_dev(cody_ctx* c, const void* s, void* d, double t, void* stream) {
  return rhs_dev(c, s, d, t, stream, kModeNonLin);
}

// One right-hand side of host arrays through the context's buffers (synchronous).
static int rhs_host(cody_ctx* c, const double* state, double* dstate, double t, int mode, const char* who) {
  if (!c || !state || !dstate) return fail(CODY_ERR_ARG, (std::string(who) + ": null argument").c_str());
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->d_tmp, state, sizeof(double2) * c->nS, hipMemcpyHostToDevice, c->stream));
  CHK(enqueue_rhs(c, c->d_tmp, c->d_tmp2, t, mode, c->stream, nullptr, nullptr));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dstate, c->d_tmp2, sizeof(double2) * c->nS, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CODY_OK;
}

int cody_time_derivative(cody_ctx* c, const double* state, double* dstate, double t) {
  return rhs_host(c, state, dstate, t, kModeFull, "cody_time_derivative");
}
int cody_time_derivative_orbs_lin(cody_ctx* c, const double* state, double* dstate, double t) {
  return rhs_host(c, state, dstate, t, kModeLin, "cody_time_derivative_orbs_lin");
}
int cody_time_derivative_nonlin(cody_ctx* c, const double* state, double* dstate, double t) {
  return rhs_host(c, state, dstate, t, kModeNonLin, "cody_time_derivative_nonlin");
}

int cody_energy(cody_ctx* c, const double* state, double t, double* energy) {
  if (!c || !state || !energy) return fail(CODY_ERR_ARG, "cody_energy: null argument");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->d_tmp, state, sizeof(double2) * c->nS, hipMemcpyHostToDevice, c->stream));
  CHK(enqueue_rhs(c, c->d_tmp, c->d_tmp2, t, kModeFull, c->stream, c->d_energy, nullptr));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(energy, c->d_energy, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CODY_OK;
}

// <x> of a device state into one double on `st`: rho1 from the CI part, the position matrix of
// the orbitals, then their contraction (PrinterObservableLinear_DumpDipole).
static int enqueue_dipole(cody_ctx* c, const double2* state, double* d_out, hipStream_t st) {
  launch_ci_density(state, c->nC, c->nO, c->reg, c->dci, c->d_rho1, c->d_rho2, c->d_r2can, c->d_reginv, st);
  launch_xmat(state + c->nC, c->d_w, c->d_x, c->nG, c->nO, c->d_part1, st);
  launch_dipole(c->nO, c->d_part1, (c->nG + 255) / 256, c->d_rho1, d_out, st);
  HIPCHK(hipGetLastError());
  return CODY_OK;
}

int cody_dipole(cody_ctx* c, const double* state, double* dipole) {
  if (!c || !state || !dipole) return fail(CODY_ERR_ARG, "cody_dipole: null argument");
  if (!c->grid_set) return fail(CODY_ERR_STATE, "cody_dipole: grid not set");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->d_tmp, state, sizeof(double2) * c->nS, hipMemcpyHostToDevice, c->stream));
  CHK(enqueue_dipole(c, c->d_tmp, c->d_energy + 1, c->stream));
  HIPCHK(hipMemcpyAsync(dipole, c->d_energy + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CODY_OK;
}

int cody_dipole_dev(cody_ctx* c, const void* d_state, void* d_dipole, void* stream) {
  if (!c || !d_state || !d_dipole) return fail(CODY_ERR_ARG, "cody_dipole_dev: null argument");
  if (!c->grid_set) return fail(CODY_ERR_STATE, "cody_dipole_dev: grid not set");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  return enqueue_dipole(c, static_cast<const double2*>(d_state), static_cast<double*>(d_dipole), st);
}

int cody_meanfield_w(cody_ctx* c, const double* orbs, double* W) {
  if (!c || !orbs || !W) return fail(CODY_ERR_ARG, "cody_meanfield_w: null argument");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  double2* d_orbs = c->d_tmp + c->nC;
  HIPCHK(hipMemcpyAsync(d_orbs, orbs, sizeof(double2) * c->nG * c->nO, hipMemcpyHostToDevice, c->stream));
  enqueue_pair_src(c, d_orbs, c->stream);
  CHK(enqueue_contraction(c, c->stream, [](const char*) {}));
  HIPCHK(hipGetLastError());
  std::vector<double> wt((size_t)c->Pc * c->nGp);
  HIPCHK(hipMemcpyAsync(wt.data(), c->d_WT, sizeof(double) * wt.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int p = 0; p < c->P; ++p)
    for (int x = 0; x < c->nG; ++x) {
      W[2 * ((size_t)p * c->nG + x)] = wt[(size_t)(2 * p) * c->nGp + x];
      W[2 * ((size_t)p * c->nG + x) + 1] = wt[(size_t)(2 * p + 1) * c->nGp + x];
    }
  return CODY_OK;
}

int cody_apply_single_body(cody_ctx* c, const double* orbs, double t, double* dorbs, double* h1) {
  if (!c || !orbs || !dorbs) return fail(CODY_ERR_ARG, "cody_apply_single_body: null argument");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  double2* d_orbs = c->d_tmp + c->nC;
  HIPCHK(hipMemcpyAsync(d_orbs, orbs, sizeof(double2) * c->nG * c->nO, hipMemcpyHostToDevice, c->stream));
  if (c->nLaser > 0) launch_set_scalar(c->d_efield, laser_field(c, t), c->stream);
  launch_onebody(d_orbs, c->d_w, c->d_x, c->d_vext, c->kin, efield_src(c), c->nG, c->nO, c->d_draw, c->d_part1,
                 c->stream);
  const int n2 = c->nO * c->nO;
  launch_sum_parts(c->d_part1, (c->nG + 255) / 256, n2, n2, c->d_h1, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dorbs, c->d_draw, sizeof(double2) * c->nG * c->nO, hipMemcpyDeviceToHost, c->stream));
  if (h1) HIPCHK(hipMemcpyAsync(h1, c->d_h1, sizeof(double2) * n2, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CODY_OK;
}

int cody_apply_single_body_acc(cody_ctx* c, const double* orbs, double t, double* dorbs, double* h1) {
  if (!c || !orbs || !dorbs) return fail(CODY_ERR_ARG, "cody_apply_single_body_acc: null argument");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const int n2 = c->nO * c->nO, nOG = c->nO * c->nG;
  double2* d_orbs = c->d_tmp + c->nC;
  double2* d_in = c->d_tmp2 + c->nC;
  HIPCHK(hipMemcpyAsync(d_orbs, orbs, sizeof(double2) * nOG, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_in, dorbs, sizeof(double2) * nOG, hipMemcpyHostToDevice, st));
  if (c->nLaser > 0) launch_set_scalar(c->d_efield, laser_field(c, t), st);
  launch_onebody(d_orbs, c->d_w, c->d_x, c->d_vext, c->kin, efield_src(c), c->nG, c->nO, c->d_draw, c->d_part1, st);
  launch_accumulate(c->d_draw, d_in, nOG, st);
  if (h1) {
    // h1 from the accumulated result, as the reference forms it
    launch_overlap_parts(d_orbs, c->d_w, c->d_draw, c->nG, c->nO, c->d_part1, st);
    launch_sum_parts(c->d_part1, (c->nG + 255) / 256, n2, n2, c->d_h1, st);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dorbs, c->d_draw, sizeof(double2) * nOG, hipMemcpyDeviceToHost, st));
  if (h1) HIPCHK(hipMemcpyAsync(h1, c->d_h1, sizeof(double2) * n2, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CODY_OK;
}

int cody_densities(cody_ctx* c, const double* coeffs, double* rdm1, double* rdm2, double* reginv) {
  if (!c || !coeffs) return fail(CODY_ERR_ARG, "cody_densities: null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->d_tmp, coeffs, sizeof(double2) * c->nC, hipMemcpyHostToDevice, c->stream));
  launch_ci_density(c->d_tmp, c->nC, c->nO, c->reg, c->dci, c->d_rho1, c->d_rho2, c->d_r2can, c->d_reginv,
                    c->stream);
  HIPCHK(hipGetLastError());
  const size_t n2 = (size_t)c->nO * c->nO, n4 = n2 * n2;
  if (rdm1) HIPCHK(hipMemcpyAsync(rdm1, c->d_rho1, sizeof(double2) * n2, hipMemcpyDeviceToHost, c->stream));
  if (rdm2) HIPCHK(hipMemcpyAsync(rdm2, c->d_rho2, sizeof(double2) * n4, hipMemcpyDeviceToHost, c->stream));
  if (reginv) HIPCHK(hipMemcpyAsync(reginv, c->d_reginv, sizeof(double2) * n2, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CODY_OK;
}

// The buffers of the interaction-potential pieces: c->d_src (Pc / 2 sources) and the FP64
// operand of V (kept current by the FP64 form; tiled here for the others).
static int prepare_potential(cody_ctx* c, hipStream_t st) {
  if (!c->d_src) CHK(ctx_alloc((void**)&c->d_src, sizeof(double2) * c->nG * (c->Pc / 2)));
  if (c->precision != CODY_PREC_FP64 || !c->d_Vt) {
    if (!c->d_Vt) CHK(ctx_alloc((void**)&c->d_Vt, sizeof(double) * c->nGp * c->nGp));
    launch_tile_f64(c->d_V, c->nGp, c->d_Vt, st);
  }
  return CODY_OK;
}

int cody_apply_interaction(cody_ctx* c, const double* orbs, int i2, int j2, double t, double* dorbs) {
  (void)t;  // the soft-Yukawa interaction does not depend on time
  if (!c || !orbs || !dorbs) return fail(CODY_ERR_ARG, "cody_apply_interaction: null argument");
  if (i2 < 0 || i2 >= c->nO || j2 < 0 || j2 >= c->nO)
    return fail(CODY_ERR_ARG, "cody_apply_interaction: orbital index out of range");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  CHK(prepare_potential(c, st));
  const int nOG = c->nO * c->nG;
  double2* d_orbs = c->d_tmp + c->nC;
  HIPCHK(hipMemcpyAsync(d_orbs, orbs, sizeof(double2) * nOG, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->d_draw, dorbs, sizeof(double2) * nOG, hipMemcpyHostToDevice, st));
  launch_pair_src_one(d_orbs, c->d_w, c->nG, i2, j2, c->d_src, st);
  launch_tile_src(c->d_src, 1, c->nG, c->nGp, c->Pc, c->d_STt, st);
  CHK(launch_mf_f64_w(c->d_Vt, c->d_STt, c->d_Wpart, c->d_WT, c->nGp, c->Pc, c->wsplit, st));
  launch_apply_potential(c->d_WT, c->nGp, d_orbs, c->nG, c->nO, c->d_draw, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dorbs, c->d_draw, sizeof(double2) * nOG, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CODY_OK;
}

int cody_interaction_potential(cody_ctx* c, int nSrc, const double* src, double* pot) {
  if (!c || nSrc < 0 || (nSrc > 0 && (!src || !pot)))
    return fail(CODY_ERR_ARG, "cody_interaction_potential: bad arguments");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const int chunk = c->Pc / 2;  // sources per contraction
  CHK(prepare_potential(c, st));
  std::vector<double> wt((size_t)c->Pc * c->nGp);
  for (int s0 = 0; s0 < nSrc; s0 += chunk) {
    const int ns = std::min(chunk, nSrc - s0);
    HIPCHK(hipMemcpyAsync(c->d_src, src + 2 * (size_t)s0 * c->nG, sizeof(double2) * c->nG * ns,
                          hipMemcpyHostToDevice, st));
    launch_tile_src(c->d_src, ns, c->nG, c->nGp, c->Pc, c->d_STt, st);
    CHK(launch_mf_f64_w(c->d_Vt, c->d_STt, c->d_Wpart, c->d_WT, c->nGp, c->Pc, c->wsplit, st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(wt.data(), c->d_WT, sizeof(double) * wt.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int q = 0; q < ns; ++q)
      for (int x = 0; x < c->nG; ++x) {
        pot[2 * ((size_t)(s0 + q) * c->nG + x)] = wt[(size_t)(2 * q) * c->nGp + x];
        pot[2 * ((size_t)(s0 + q) * c->nG + x) + 1] = wt[(size_t)(2 * q + 1) * c->nGp + x];
      }
  }
  return CODY_OK;
}

int cody_apply_correlation(cody_ctx* c, const double* orbs, const double* rdm1, const double* rdm2, double* dorbs,
                           double* h2) {
  if (!c || !orbs || !rdm1 || !rdm2 || !dorbs) return fail(CODY_ERR_ARG, "cody_apply_correlation: null argument");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const int nO = c->nO, nG = c->nG, np = nO * (nO - 1) / 2;
  const size_t n2 = (size_t)nO * nO, n4 = n2 * n2, nOG = (size_t)nO * nG;
  if (h2 && !c->d_h2) CHK(ctx_alloc((void**)&c->d_h2, sizeof(double2) * n4));
  // canonical rho2Lin (i1 < i2, j1 < j2), the operand of the orbital kernel
  std::vector<double> can(2 * std::max(1, np * np));
  auto strict = [&](int a, int b) { return a * nO - (a * (a + 1)) / 2 + (b - a - 1); };
  for (int j2 = 0; j2 < nO; ++j2)
    for (int j1 = 0; j1 < j2; ++j1)
      for (int i2 = 0; i2 < nO; ++i2)
        for (int i1 = 0; i1 < i2; ++i1) {
          const size_t e = strict(i1, i2) + (size_t)np * strict(j1, j2);
          const size_t f = i1 + nO * (i2 + nO * (j1 + (size_t)nO * j2));
          can[2 * e] = rdm2[2 * f];
          can[2 * e + 1] = rdm2[2 * f + 1];
        }
  double2* d_orbs = c->d_tmp + c->nC;
  HIPCHK(hipMemcpyAsync(d_orbs, orbs, sizeof(double2) * nOG, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->d_rho1, rdm1, sizeof(double2) * n2, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->d_r2can, can.data(), sizeof(double) * can.size(), hipMemcpyHostToDevice, st));
  launch_reginv(c->d_rho1, nO, c->reg, c->d_reginv, st);
  enqueue_pair_src(c, d_orbs, st);
  CHK(enqueue_contraction(c, st, [](const char*) {}));
  launch_orbital(d_orbs, c->d_w, c->d_hphi, c->d_WT, c->d_r2can, c->d_reginv, nG, c->nGp, nO, kModeNonLin,
                 c->d_draw, c->d_part, st);
  if (h2) {
    launch_gram_f64(c->d_STt, c->d_WT, c->d_Gpart, c->nGp, c->Pc, c->nxs, st);
    launch_gram_reduce(c->d_Gpart, c->nxs, c->Pc, c->d_G, c->d_part1, 0, 0, nullptr, st);
    launch_h2_full(c->d_G, c->Pc, nO, c->d_h2, st);
  }
  HIPCHK(hipGetLastError());
  std::vector<double> corr(2 * nOG);
  HIPCHK(hipMemcpyAsync(corr.data(), c->d_draw, sizeof(double2) * nOG, hipMemcpyDeviceToHost, st));
  if (h2) HIPCHK(hipMemcpyAsync(h2, c->d_h2, sizeof(double2) * n4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (size_t i = 0; i < 2 * nOG; ++i) dorbs[i] += corr[i];  // accumulate semantics
  return CODY_OK;
}

int cody_project_orbs(cody_ctx* c, const double* orbs, double* dorbs) {
  if (!c || !orbs || !dorbs) return fail(CODY_ERR_ARG, "cody_project_orbs: null argument");
  if (!c->grid_set) return fail(CODY_ERR_STATE, "cody_project_orbs: grid not set");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const size_t nOG = (size_t)c->nO * c->nG;
  double2* d_orbs = c->d_tmp + c->nC;
  HIPCHK(hipMemcpyAsync(d_orbs, orbs, sizeof(double2) * nOG, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->d_draw, dorbs, sizeof(double2) * nOG, hipMemcpyHostToDevice, st));
  launch_overlap_parts(d_orbs, c->d_w, c->d_draw, c->nG, c->nO, c->d_part1, st);
  launch_project(d_orbs, c->d_draw, c->d_part1, (c->nG + 255) / 256, c->nG, c->nO, 1, c->d_hphi, nullptr, nullptr, 0,
                 nullptr, nullptr, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dorbs, c->d_hphi, sizeof(double2) * nOG, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // the projection kernel scales by -i (the right-hand side's factor): undo it exactly
  for (size_t i = 0; i < nOG; ++i) {
    const double re = dorbs[2 * i], im = dorbs[2 * i + 1];
    dorbs[2 * i] = -im;
    dorbs[2 * i + 1] = re;
  }
  return CODY_OK;
}

int cody_apply_kinetic(cody_ctx* c, const double* orbs, double* dorbs) {
  if (!c || !orbs || !dorbs) return fail(CODY_ERR_ARG, "cody_apply_kinetic: null argument");
  if (!c->grid_set || !c->kin_set) return fail(CODY_ERR_STATE, "cody_apply_kinetic: grid or kinetic operator not set");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const size_t nOG = (size_t)c->nO * c->nG;
  double2* d_orbs = c->d_tmp + c->nC;
  HIPCHK(hipMemcpyAsync(d_orbs, orbs, sizeof(double2) * nOG, hipMemcpyHostToDevice, st));
  // the one-body kernel with a zero potential and a zero field
  launch_onebody(d_orbs, c->d_w, c->d_x, c->d_zeros, c->kin, c->d_zeros + c->nG, c->nG, c->nO, c->d_draw,
                 c->d_part1, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dorbs, c->d_draw, sizeof(double2) * nOG, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CODY_OK;
}

int cody_profile_rhs(cody_ctx* c, int iters, int nMax, double* avg_ms, char* names, int namesLen,
                     int* nKernels) {
  if (!c || iters < 1 || !avg_ms || nMax < 1) return fail(CODY_ERR_ARG, "cody_profile_rhs: bad arguments");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  std::vector<double> acc;
  std::string labels;
  int nk = 0;
  for (int it = 0; it < iters + 1; ++it) {
    RhsTrace tr;
    CHK(enqueue_rhs(c, c->d_state, c->d_tmp2, 0.0, kModeFull, c->stream, nullptr, &tr));
    HIPCHK(hipStreamSynchronize(c->stream));
    nk = static_cast<int>(tr.ev.size()) - 1;
    if (acc.empty()) {
      acc.assign(nk, 0.0);
      for (int k = 0; k < nk; ++k) labels += (k ? ";" : "") + tr.names[k];
    }
    for (int k = 0; k < nk; ++k) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, tr.ev[k], tr.ev[k + 1]);
      if (it > 0) acc[k] += ms;  // first pass is a warm-up
    }
    for (auto e : tr.ev) hipEventDestroy(e);
  }
  for (int k = 0; k < nk && k < nMax; ++k) avg_ms[k] = acc[k] / iters;
  if (nKernels) *nKernels = nk;
  if (names && namesLen > 0) {
    std::strncpy(names, labels.c_str(), namesLen - 1);
    names[namesLen - 1] = '\0';
  }
  return CODY_OK;
}

// ---------------------------------------------------------------- batches of instances
static_assert(kPlanRdmWideConf == kRdmWideConf && kPlanRdmWideItems == kRdmWideItems, "CI density blocks agree");
static_assert(kPlanMaxOrbs == kMaxOrbs && kPlanInt8MaxBlocks64 == kR16MaxBlocks64 && kPlanCorrQMinOrbs == kCorrQMinOrbs &&
                  kPlanCorrQGroups == kCorrQGroups,
              "the host planner and the device kernels agree on their limits");
static_assert(sizeof(BlockRef) == sizeof(int2), "block references are uploaded as int2");
static_assert(plan_corr_q_doubles(8) == corr_q_doubles(8) && plan_corr_q_doubles(9) == corr_q_doubles(9) &&
                  plan_corr_q_doubles(10) == corr_q_doubles(10) && plan_corr_q_doubles(11) == corr_q_doubles(11) &&
                  plan_corr_q_doubles(12) == corr_q_doubles(12),
              "the host planner and the device kernels agree on the size of Q");
namespace {

// Device time of one right-hand side, for the assignment of work to chains: the contraction at
// ~50 TFLOP/s plus ~50 us of latency-bound launches.
double rhs_cost_estimate(const cody_ctx* c) { return 2.0 * c->nGp * (double)c->nGp * c->Pc / 50e6 + 50.0; }

// Cost of instance i for the schedule: its measured right-hand side (cody_batch_autotune) or the
// estimate.
double batch_cost(const cody_batch* b, int i) {
  return i < (int)b->cost.size() ? b->cost[i] : rhs_cost_estimate(b->ctx[i]);
}

// Partitions the batch: instances on the FP64 or the 16-row int8 contraction are grouped by
// (orbital count, contraction), each group getting its argument table and block maps in device
// memory; the others (the split-FP32 / BF16 forms, the other int8 kernels) keep their own
// launches.  The maps repeat each instance's own launch grids (launch_rhs_front, launch_mf_f64_w /
// launch_mf_i8_orb, launch_rhs_wsum, launch_rhs_mid / launch_rhs_gram, launch_rhs_post,
// launch_ci_apply_project), so a grouped workgroup does what that instance's own would.
int batch_build_groups(cody_batch* b, void* const* d_states, void* const* d_dstates) {
  b->release_groups();
  const int n = static_cast<int>(b->ctx.size());
  for (int i = 0; i < n; ++i) {
    cody_ctx* c = b->ctx[i];
    // (the BF16-sliced form keeps its own launches)
    const bool fp64 = c->precision == CODY_PREC_FP64, i8 = c->precision == CODY_PREC_INT8_SLICED && int8_tiled(c);
    if (!fp64 && !i8) {
      b->singles.push_back(i);
      continue;
    }
    BatchGroup* grp = nullptr;
    for (auto& g : b->groups)
      if (g.nO == c->nO && g.int8 == i8) grp = &g;
    if (!grp) {
      b->groups.emplace_back();
      grp = &b->groups.back();
      grp->nO = c->nO;
      grp->Pc = c->Pc;
      grp->int8 = i8;
      grp->fused = i8 && int8_fused(c);
    }
    grp->members.push_back(i);
  }
  // a group whose cost exceeds split_share times a chain's share of the batch is split into two
  // groups of about equal cost (members longest first, each to the lighter half), so that its
  // launches run on two chains side by side instead of setting the batch's critical path alone
  {
    double total = 0.0;
    for (const auto& g : b->groups)
      for (int i : g.members) total += batch_cost(b, i);
    for (int i : b->singles) total += batch_cost(b, i);
    const double share = b->split_share * total / b->nchains;
    for (size_t gi0 = 0; gi0 < b->groups.size() && b->groups.size() < 3 * cody_batch::kChains; ++gi0) {
      BatchGroup& g = b->groups[gi0];
      double cost = 0.0;
      for (int i : g.members) cost += batch_cost(b, i);
      if (cost <= share || g.members.size() < 2) continue;
      std::vector<int> mem = g.members;
      std::stable_sort(mem.begin(), mem.end(), [&](int a, int c) { return batch_cost(b, a) > batch_cost(b, c); });
      BatchGroup h;
      h.nO = g.nO;
      h.Pc = g.Pc;
      h.int8 = g.int8;
      h.fused = g.fused;
      std::vector<int> keep;
      double ck = 0.0, ch = 0.0;
      for (int i : mem) {
        const double ci = batch_cost(b, i);
        if (ck <= ch) {
          keep.push_back(i);
          ck += ci;
        } else {
          h.members.push_back(i);
          ch += ci;
        }
      }
      std::sort(keep.begin(), keep.end());
      std::sort(h.members.begin(), h.members.end());
      g.members = keep;
      b->groups.push_back(h);
      --gi0;  // (look at the kept half again)
    }
  }
  for (auto& g : b->groups) {
    std::vector<GroupInstance> gi;
    std::vector<PlanShape> shapes;
    const int nO = g.nO, n2 = nO * nO, np = nO * (nO - 1) / 2;
    for (int mi = 0; mi < (int)g.members.size(); ++mi) {
      const int i = g.members[mi];
      cody_ctx* c = b->ctx[i];
      GroupInstance a{};
      a.C = static_cast<const double2*>(d_states[i]);
      a.dC = static_cast<double2*>(d_dstates[i]);
      a.nC = c->nC;
      a.nG = c->nG;
      a.nGp = c->nGp;
      a.Pc = c->Pc;
      a.nRdm = (n2 + np * np + 4 * ci_rdm_items_per_wave(c->nC) - 1) / (4 * ci_rdm_items_per_wave(c->nC));
      a.nObx = (c->nG + 255) / 256;
      a.nsplit = g.int8 ? 1 : c->wsplit;  // (the int8 contraction writes W^T itself)
      a.wsumFirst = g.fused ? 0 : 1;  // (the fused int8 path forms R in its Gram launch)
      a.nxs = c->nxs;
      a.nPart = c->nPart;
      a.nPart1 = (c->nG + 255) / 256;
      // energy shares the CI apply sums: the G reduction's, or on the fused path the direct G's
      a.nRed = g.fused ? gram_direct_energy_parts(c->Pc) : rhs_post_energy_parts(c->Pc, nO);
      a.t = c->dci;
      a.rho1 = c->d_rho1;
      a.r2can = c->d_r2can;
      a.rho2 = c->d_rho2;
      a.hphi = c->d_hphi;
      a.part1 = c->d_part1;
      a.reginv = c->d_reginv;
      a.draw = c->d_draw;
      a.part = c->d_part;
      a.h1 = c->d_h1;
      a.dCraw = c->d_dCraw;
      a.w = c->d_w;
      a.x = c->d_x;
      a.vext = c->d_vext;
      a.efield = efield_src(c);
      a.kin = c->kin;
      a.reg = c->reg;
      a.STt = c->d_STt;
      a.Wpart = c->d_Wpart;
      a.WT = c->d_WT;
      a.Gpart = c->d_Gpart;
      a.G = c->d_G;
      a.epart = c->d_epart;
      a.act = c->d_act;
      a.actK = c->d_actK;
      // (the FP64 kernel keeps its block list in LDS, up to kSkipMaxBlocks16 fragments)
      a.actC = gram_flags(c);
      a.partr = g.fused ? c->d_partr : nullptr;
      a.Vt = c->d_Vt;
      a.Vi8 = c->d_Vi8;
      a.eV = c->d_eV;
      a.Si8 = g.int8 ? c->d_Si8 : nullptr;  // (selects the front's pair-source role)
      a.eSb = c->d_eSb;
      gi.push_back(a);
      PlanShape sh;
      sh.nO = nO;
      sh.nC = c->nC;
      sh.nG = c->nG;
      sh.nGp = c->nGp;
      sh.Pc = c->Pc;
      sh.nsplit = a.nsplit;
      sh.nxs = c->nxs;
      sh.int8 = g.int8;
      sh.fused = g.fused;
      shapes.push_back(sh);
      g.maxC = std::max(g.maxC, c->nC);
      g.cost += batch_cost(b, i);
    }
    // the block maps, built and checked entry by entry on the host before any launch
    GroupPlan plan;
    const std::string why = build_group_plan(shapes, plan);
    if (!why.empty()) return fail(CODY_ERR_ARG, ("cody_batch_time_derivative_dev: " + why).c_str());
    for (int k = 0; k < kPlanKinds; ++k) {
      g.off[k] = static_cast<int>(plan.off[k]);
      g.cnt[k] = static_cast<int>(plan.cnt[k]);
    }
    if (nO >= kCorrQMinOrbs && !g.fused) {  // Q of every member's orbital correlation
      const size_t nq = (size_t)corr_q_doubles(nO);
      HIPCHK(hipMalloc((void**)&g.d_qf, sizeof(double) * nq * gi.size()));
      for (size_t m = 0; m < gi.size(); ++m) gi[m].qf = g.d_qf + m * nq;
    }
    HIPCHK(hipMalloc((void**)&g.d_inst, sizeof(GroupInstance) * gi.size()));
    HIPCHK(hipMalloc((void**)&g.d_map, sizeof(BlockRef) * plan.map.size()));
    HIPCHK(hipMemcpy(g.d_inst, gi.data(), sizeof(GroupInstance) * gi.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.d_map, plan.map.data(), sizeof(BlockRef) * plan.map.size(), hipMemcpyHostToDevice));
  }
  return CODY_OK;
}

// The right-hand sides of a group's instances: one launch per kernel type.
int enqueue_group_rhs(const BatchGroup& g, hipStream_t st) {
  auto map = [&](int k) { return g.d_map + g.off[k]; };
  launch_grp_front(g.nO, g.d_inst, map(kPlanFront), g.cnt[kPlanFront], g.maxC, st);
  if (g.fused) {
    CHK(launch_grp_mf_i8_orb(g.nO, g.d_inst, map(kPlanContract), g.cnt[kPlanContract], st));
  } else if (g.int8) {
    CHK(launch_grp_mf_i8_r16(g.Pc, g.d_inst, map(kPlanContract), g.cnt[kPlanContract], st, g.nO));
  } else {
    CHK(launch_grp_mf_f64_w(g.Pc, g.d_inst, map(kPlanContract), g.cnt[kPlanContract], st));
  }
  launch_grp_wsum(g.nO, g.d_inst, map(kPlanWsum), g.cnt[kPlanWsum], st);
  launch_grp_qform(g.nO, g.d_inst, map(kPlanQform), g.cnt[kPlanQform], st);
  if (g.fused) {  // G directly, h1, the overlap sums, R; then the CI apply beside the projection
    launch_grp_gram(g.nO, g.d_inst, map(kPlanMid), g.cnt[kPlanMid], st);
    launch_grp_final(g.nO, g.d_inst, map(kPlanApply), g.cnt[kPlanApply], st);
    HIPCHK(hipGetLastError());
    return CODY_OK;
  }
  launch_grp_mid(g.nO, g.d_inst, map(kPlanMid), g.cnt[kPlanMid], st);
  launch_grp_post(g.nO, g.d_inst, map(kPlanPost), g.cnt[kPlanPost], st);
  launch_grp_apply(g.d_inst, map(kPlanApply), g.cnt[kPlanApply], g.nO, st);
  HIPCHK(hipGetLastError());
  return CODY_OK;
}

}  // namespace

int cody_batch_create(cody_batch** out, int n, cody_ctx* const* ctxs) {
  if (!out || n < 1 || !ctxs) return fail(CODY_ERR_ARG, "cody_batch_create: bad arguments");
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return fail(CODY_ERR_ARG, "cody_batch_create: null context");
    if (ctxs[i]->device != ctxs[0]->device) return fail(CODY_ERR_ARG, "cody_batch_create: contexts on different devices");
    for (int k = 0; k < i; ++k)
      if (ctxs[k] == ctxs[i]) return fail(CODY_ERR_ARG, "cody_batch_create: a context appears twice");
  }
  auto* b = new cody_batch();
  b->ctx.assign(ctxs, ctxs + n);
  if (hipSetDevice(ctxs[0]->device) != hipSuccess || hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
    delete b;
    return fail(CODY_ERR_HIP, "cody_batch_create: stream / event creation failed");
  }
  for (int k = 0; k < cody_batch::kChains; ++k)
    if (hipStreamCreateWithFlags(&b->chain[k], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) != hipSuccess) {
      b->release();
      delete b;
      return fail(CODY_ERR_HIP, "cody_batch_create: stream / event creation failed");
    }
  *out = b;
  return CODY_OK;
}

int cody_batch_destroy(cody_batch* b) {
  if (!b) return CODY_OK;
  hipSetDevice(b->ctx[0]->device);
  b->release();
  delete b;
  return CODY_OK;
}

int cody_batch_time_derivative_dev(cody_batch* b, void* const* d_states, void* const* d_dstates, double t,
                                   void* stream) {
  if (!b || !d_states || !d_dstates) return fail(CODY_ERR_ARG, "cody_batch_time_derivative_dev: null argument");
  const int n = static_cast<int>(b->ctx.size());
  for (int i = 0; i < n; ++i) {
    if (!d_states[i] || !d_dstates[i] || d_states[i] == d_dstates[i])
      return fail(CODY_ERR_ARG, "cody_batch_time_derivative_dev: bad state pointers");
    CHK(check_ready(b->ctx[i]));
  }
  HIPCHK(hipSetDevice(b->ctx[0]->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : b->cs;
  std::vector<const void*> key;
  for (int i = 0; i < n; ++i) {
    key.push_back(d_states[i]);
    key.push_back(d_dstates[i]);
  }
  bool stale = !b->exec || key != b->key;
  for (int i = 0; i < n && !stale; ++i) stale = b->ctx[i]->generation != b->gen[i];
  if (stale) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    CHK(batch_build_groups(b, d_states, d_dstates));
    // one graph: each group's grouped launches on one chain, the other instances' own launches on
    // the chains too; groups and instances by longest-first greedy assignment of an estimated cost
    std::vector<std::pair<double, int>> items;  // (cost, group index g >= 0, or -1 - instance)
    for (int g = 0; g < (int)b->groups.size(); ++g) items.emplace_back(b->groups[g].cost, g);
    for (int i : b->singles) items.emplace_back(batch_cost(b, i), -1 - i);
    std::stable_sort(items.begin(), items.end(), [](const auto& a, const auto& c) { return a.first > c.first; });
    std::vector<double> load(b->nchains, 0.0);
    HIPCHK(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
    hipEventRecord(b->fork, b->cs);
    for (int k = 0; k < cody_batch::kChains; ++k) hipStreamWaitEvent(b->chain[k], b->fork, 0);
    int r = CODY_OK;
    for (const auto& it : items) {
      if (r != CODY_OK) break;
      const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      load[k] += it.first;
      if (it.second >= 0) {
        r = enqueue_group_rhs(b->groups[it.second], b->chain[k]);
      } else {
        cody_ctx* c = b->ctx[-1 - it.second];
        c->capturing = true;
        r = enqueue_rhs(c, static_cast<const double2*>(d_states[-1 - it.second]),
                        static_cast<double2*>(d_dstates[-1 - it.second]), t, kModeFull, b->chain[k], nullptr, nullptr);
        c->capturing = false;
      }
    }
    for (int k = 0; k < cody_batch::kChains; ++k) {
      hipEventRecord(b->join[k], b->chain[k]);
      hipStreamWaitEvent(b->cs, b->join[k], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cs, &g);
    if (r != CODY_OK || e != hipSuccess || !g || hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
      if (g) hipGraphDestroy(g);
      b->exec = nullptr;
      return r != CODY_OK ? r : fail(CODY_ERR_HIP, "cody_batch_time_derivative_dev: graph capture failed");
    }
    hipGraphDestroy(g);
    b->key = key;
    b->gen.resize(n);
    for (int i = 0; i < n; ++i) b->gen[i] = b->ctx[i]->generation;
  }
  for (int i = 0; i < n; ++i)
    if (b->ctx[i]->nLaser > 0) launch_set_scalar(b->ctx[i]->d_efield, laser_field(b->ctx[i], t), st);
  HIPCHK(hipGraphLaunch(b->exec, st));
  return CODY_OK;
}

int cody_batch_autotune(cody_batch* b, void* const* d_states, void* const* d_dstates, double* cand_ms,
                        int max_cand, int* pick_out) {
  if (!b) return fail(CODY_ERR_ARG, "cody_batch_autotune: null batch");
  // validates the arguments and evaluates once with the current schedule
  CHK(cody_batch_time_derivative_dev(b, d_states, d_dstates, 0.0, nullptr));
  const int n = static_cast<int>(b->ctx.size());
  hipStream_t st = b->cs;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) hipEventDestroy(e0);
    return fail(CODY_ERR_HIP, "cody_batch_autotune: event creation failed");
  }
  auto timed = [&](auto&& body, int reps, double* ms_out) -> int {
    CHK(body());  // warm-up
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) CHK(body());
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *ms_out = ms / reps;
    return CODY_OK;
  };
  int r = CODY_OK;
  // each instance's own right-hand side, in microseconds
  std::vector<double> cost(n, 0.0);
  for (int i = 0; i < n && r == CODY_OK; ++i) {
    cody_ctx* c = b->ctx[i];
    double ms = 0.0;
    r = timed([&] { return run_rhs(c, static_cast<const double2*>(d_states[i]), static_cast<double2*>(d_dstates[i]), 0.0,
                                   kModeFull, st); }, 3, &ms);
    cost[i] = 1000.0 * ms;
  }
  struct Cand {
    int chains;
    double share;
  };
  const Cand cand[] = {{4, 1.0}, {4, 0.7}, {4, 0.5}, {2, 1.0}, {2, 0.5}, {1, 1.0}};
  static_assert(cody_batch::kChains >= 4, "the candidates use up to four chains");
  constexpr int nCand = sizeof(cand) / sizeof(cand[0]);
  double best = 1e300;
  int pick = 0;
  if (r == CODY_OK) b->cost = cost;
  for (int k = 0; k < nCand && r == CODY_OK; ++k) {
    b->nchains = cand[k].chains;
    b->split_share = cand[k].share;
    b->key.clear();  // (rebuilds the groups and the graph on the next call)
    double ms = 0.0;
    r = timed([&] { return cody_batch_time_derivative_dev(b, d_states, d_dstates, 0.0, st); }, 5, &ms);
    if (r == CODY_OK && cand_ms && k < max_cand) cand_ms[k] = ms;
    if (r == CODY_OK && ms < best) {
      best = ms;
      pick = k;
    }
  }
  if (r == CODY_OK) {
    b->nchains = cand[pick].chains;
    b->split_share = cand[pick].share;
    b->key.clear();
    r = cody_batch_time_derivative_dev(b, d_states, d_dstates, 0.0, st);
    if (r == CODY_OK && hipStreamSynchronize(st) != hipSuccess) r = fail(CODY_ERR_HIP, "cody_batch_autotune: sync failed");
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (r != CODY_OK) return r;
  if (pick_out) *pick_out = pick;
  return CODY_OK;
}

int cody_batch_set_schedule(cody_batch* b, int chains, double split_threshold) {
  if (!b || chains < 1 || chains > cody_batch::kChains || !(split_threshold > 0.0))
    return fail(CODY_ERR_ARG, "cody_batch_set_schedule: bad arguments");
  b->nchains = chains;
  b->split_share = split_threshold;
  b->key.clear();
  return CODY_OK;
}

int cody_profile_contraction(cody_ctx* c, int iters, double* avg_ms) {
  if (!c || iters < 1 || !avg_ms) return fail(CODY_ERR_ARG, "cody_profile_contraction: bad arguments");
  CHK(check_ready(c));
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  enqueue_pair_src(c, c->d_state + c->nC, st);
  auto none = [](const char*) {};
  CHK(enqueue_contraction(c, st, none));  // operands (slices, block list) of the state
  // the int8 form's right-hand side runs the contraction with the orbital correlation in its
  // epilogue: that kernel is timed, with R, rho2 and h phi of the last right-hand side
  const OrbCorrArgs oa{c->d_state + c->nC, c->d_w, c->d_hphi, c->d_r2can, c->nG, kModeFull, c->d_draw, c->d_part};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (int i = -1; i < iters; ++i) {  // (launch -1 is a warm-up)
    if (i == 0) HIPCHK(hipEventRecord(e0, st));
    switch (c->precision) {
      case CODY_PREC_INT8_SLICED:  // the contraction kernel alone
      case CODY_PREC_BF16_SLICED:
        if (int8_fused(c))
          CHK(launch_mf_i8_orb(c->d_Vi8, c->d_eV, c->d_Si8, c->d_eSb, act_flags(c), c->d_WT, c->nGp, c->nO, oa, st,
                               c->precision == CODY_PREC_BF16_SLICED));
        else if (int8_tiled(c))
          CHK(launch_mf_i8_r16(c->d_Vi8, c->d_eV, c->d_Si8, c->d_eSb, act_flags(c), c->d_WT, c->nGp, c->Pc, st,
                               c->precision == CODY_PREC_BF16_SLICED));
        else
          launch_mf_i8(c->d_Vi8, c->d_eV, c->d_Si8, c->d_eS, c->d_WT, c->nGp, c->Pc, st);
        break;
      case CODY_PREC_SPLIT_FP32:
        launch_mf_f32x3(c->d_Vhi, c->d_Vlo, c->d_Shi, c->d_Slo, c->d_WT, c->nGp, c->Pc, st);
        break;
      case CODY_PREC_BF16X3:
        launch_mf_bf16x3(c->d_Vbf, c->d_Sbf, c->d_WT, c->nGp, c->Pc, st);
        break;
      default:  // the contraction kernel alone (its partial sum k_wsum is not timed)
        CHK(launch_mf_f64_w(c->d_Vt, c->d_STt, c->d_Wpart, nullptr, c->nGp, c->Pc, c->wsplit, st, act_flags(c)));
    }
  }
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *avg_ms = ms / iters;
  return CODY_OK;
}

}  // extern "C"
