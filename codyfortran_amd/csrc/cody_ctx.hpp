// Device context of one MCTDHX system (one GPU, one fermionic body type).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "batch_plan.hpp"
#include "host_setup.hpp"
#include "mctdhx_kernels.hpp"

namespace cody {
int ctx_alloc(void** p, size_t bytes);
int set_error(int code, const char* msg);
}  // namespace cody

// Events recorded after each launch of a right-hand side (profiling).
struct RhsTrace {
  std::vector<hipEvent_t> ev;
  std::vector<std::string> names;
};

// One captured right-hand side (both streams, all launches) for fixed buffers and mode.
struct RhsGraph {
  const void* state;
  const void* dstate;
  int mode;
  const double* energy;
  hipGraphExec_t exec;
};

// A captured propagation launch sequence: one SIL sub-step (kind 0) from `in` into `out`, or one
// Strang step of the context's state (kind 1), for the parameters it bakes in.
struct PropGraph {
  int kind;
  const void* in;
  const void* out;
  int maxDim, useMetric;
  double tol, dt;
  hipGraphExec_t exec;
};

// Work buffers of the device-side propagators (allocated on first use).
struct PropBuffers {
  int kdim = 0;
  double2 *basis = nullptr, *w = nullptr, *w2 = nullptr, *partial = nullptr, *phi = nullptr;
  double2 *trial = nullptr, *work = nullptr, *backup = nullptr;
  double *alpha = nullptr, *beta = nullptr, *nrm2 = nullptr, *par = nullptr, *flags = nullptr;
  double* h_flags = nullptr;  // pinned host copy of flags
  int* ctl = nullptr;         // Lanczos loop control: done, Krylov dimension, zero input
  double2 *rk_inc = nullptr, *rk_tmp = nullptr, *rk_d = nullptr;
  std::vector<PropGraph> graphs;
  // Krylov dimensions of the last Strang steps (the larger of the two SILs), to cap the next
  // step's Lanczos loop
  int dimHist[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int dimHistN = 0;
  int krylov_cap(int maxDim) const {
    if (dimHistN == 0) return maxDim;
    int m = 0;
    for (int i = 0; i < 8 && i < dimHistN; ++i) m = dimHist[i] > m ? dimHist[i] : m;
    return m + 1 < maxDim ? m + 1 : maxDim;
  }
  void record_dim(int m) { dimHist[dimHistN++ % 8] = m; }
  void drop_graphs() {
    for (auto& g : graphs) hipGraphExecDestroy(g.exec);
    graphs.clear();
  }

  void release() {
    drop_graphs();
    void* ptrs[] = {basis, w, w2, partial, phi, trial, work, backup, alpha, beta, nrm2, par, flags, ctl,
                    rk_inc, rk_tmp, rk_d};
    for (void* p : ptrs)
      if (p) hipFree(p);
    if (h_flags) hipHostFree(h_flags);
    *this = PropBuffers();
  }
};

// Work buffers of the adjoint (vector-Jacobian products, Chebyshev half-steps, checkpointed
// gradient), allocated on first use.
struct AdjBuffers {
  bool ready = false;
  double2 *f = nullptr, *D = nullptr, *tmpC = nullptr, *Bpart = nullptr, *B = nullptr, *inp = nullptr;
  double2 *h1b = nullptr, *h2b = nullptr, *rho2p = nullptr, *r2pb = nullptr, *r2b = nullptr;
  double2 *lin1 = nullptr, *lin2 = nullptr, *dbar = nullptr, *hbar = nullptr, *X = nullptr;
  double2* gpart = nullptr;  // orbital terms of the pair groups of the correlation adjoint
  double *STadj = nullptr, *WTadj = nullptr, *epart = nullptr, *eacc = nullptr, *J = nullptr;
  // Chebyshev half-steps exp(-i dt/2 h(E)) on [lo, hi] of degree K (cody_set_chebyshev)
  int K = 0;  // 0: not set up (or invalidated by a change of the grid, kinetic or potential)
  double lo = 0.0, hi = 0.0, dt = 0.0, emax = 0.0;
  std::vector<double2> a;  // expansion coefficients a_0..a_K
  double2 *U1 = nullptr, *U2 = nullptr, *cacc = nullptr;
  double2* cb[4] = {};  // Clenshaw ring of the Chebyshev adjoint
  // Strang-step tape, adjoint state and checkpoints
  double2* y[4] = {};
  double2* dl[4] = {};
  double2 *rk_d = nullptr, *rk_inc = nullptr, *lam = nullptr, *g = nullptr, *ys = nullptr;
  double2 *ck = nullptr, *seg = nullptr;
  size_t nck = 0, nseg = 0;
  double2* tape = nullptr;  // per-step tapes of one checkpoint segment
  size_t ntape = 0;
  int tape_K = 0;
  double* ebar = nullptr;
  size_t nebar = 0;
  void free_cheb() {
    void* p[] = {U1, U2, cacc, cb[0], cb[1], cb[2], cb[3]};
    for (void* q : p)
      if (q) hipFree(q);
    U1 = U2 = cacc = nullptr;
    cb[0] = cb[1] = cb[2] = cb[3] = nullptr;
    K = 0;
  }
  void release() {
    free_cheb();
    void* p[] = {f,     D,     tmpC,  Bpart, B,     inp,   h1b,   h2b,   rho2p, r2pb,  r2b,   lin1,  lin2, dbar,
                 hbar,  X,     gpart, STadj, WTadj, epart, eacc,  J,     y[0],  y[1],  y[2],  y[3],  dl[0], dl[1], dl[2],
                 dl[3], rk_d,  rk_inc, lam,  g,     ys,    ck,    seg,   ebar,  tape};
    for (void* q : p)
      if (q) hipFree(q);
    *this = AdjBuffers();
  }
};

struct cody_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool split_consistent = false;  // nonlinear split part = full minus linear (cody_set_split)
  // sizes
  int nG = 0, nGp = 0, N = 0, nO = 0, nC = 0, nS = 0, P = 0, Pc = 0;
  int nsplit = 1, nxs = 1, nPart = 1;
  int wsplit = 4;  // k split of the FP64 contraction (cody_autotune may change it)
  double reg = 0.0;
  int precision = 0;
  // setup flags
  bool grid_set = false, kin_set = false, pot_set = false, int_set = false;
  std::vector<double> hx, hw;
  cody::KinLaunch kin{};
  int nLaser = 0;
  double laserT = 0.0;
  double laserC[64] = {};
  // CI tables
  cody::CiTables ci;
  cody::DevCiTables dci{};
  void* ci_blob = nullptr;
  // device buffers
  double *d_x = nullptr, *d_w = nullptr, *d_vext = nullptr, *d_kref = nullptr;
  double *d_V = nullptr, *d_ST = nullptr, *d_WT = nullptr;
  double *d_STt = nullptr, *d_Wpart = nullptr;  // fragment-tiled S^T, k-split partials of W^T
  int* d_act = nullptr;        // per-fragment activity of S^T (k_pair_src_t), Pc/16 x nGp/16
  bool skip_zero = true;       // contraction skips the k blocks whose pair sources are all zero
  double *d_Gpart = nullptr, *d_G = nullptr, *d_energy = nullptr;
  double* d_epart = nullptr;  // energy partials of the G reduction (the CI projection's overlap)
  double* d_efield = nullptr;  // E(t) of the dipole coupling, read by the one-body kernel
  double2 *d_state = nullptr, *d_tmp = nullptr, *d_tmp2 = nullptr;
  double2 *d_rho1 = nullptr, *d_rho2 = nullptr, *d_reginv = nullptr;
  double2* d_r2can = nullptr;  // canonical rho2Lin (i1 < i2, j1 < j2)
  double2 *d_h1 = nullptr, *d_hphi = nullptr, *d_dCraw = nullptr;
  double2 *d_draw = nullptr, *d_part = nullptr, *d_part1 = nullptr;
  double* d_qf = nullptr;  // Q of the orbital correlation (nO >= kCorrQMinOrbs)
  double2* d_partr = nullptr;  // the int8 path's two sets of overlap partials summed (2 nO^2)
  double* d_zeros = nullptr;  // nG + 1 zeros: a zero potential and a zero field (the kinetic piece)
  double2* d_src = nullptr;   // interaction sources of cody_interaction_potential (nG x P, on use)
  double2* d_h2 = nullptr;    // full h2 tensor of cody_apply_correlation (nO^4, on use)
  PropBuffers prop;
  AdjBuffers adj;
  // contraction operands derived from V (allocated by cody_set_precision): FP64 fragment-tiled
  // copy, int8 slices, FP32 and BF16 splits
  double* d_Vt = nullptr;
  int8_t *d_Vi8 = nullptr, *d_Si8 = nullptr;
  int *d_eV = nullptr, *d_eS = nullptr;
  int *d_eSb = nullptr, *d_actK = nullptr;  // tiled int8 form: S's block exponents, block flags
  float *d_Vhi = nullptr, *d_Vlo = nullptr, *d_Shi = nullptr, *d_Slo = nullptr;
  __bf16 *d_Vbf = nullptr, *d_Sbf = nullptr;
  bool vprec_ready = false;  // V operands of the current precision are up to date

  int allocate() {
    using cody::ctx_alloc;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
      return cody::set_error(2, "hipStreamCreate failed");
    const size_t n2 = (size_t)nO * nO, n4 = n2 * n2;
    int r = 0;
#define ALLOC(ptr, bytes)                                   \
  if ((r = ctx_alloc((void**)&(ptr), (bytes))) != 0) return r;
    ALLOC(d_x, sizeof(double) * nG);
    ALLOC(d_w, sizeof(double) * nG);
    ALLOC(d_vext, sizeof(double) * nG);
    ALLOC(d_kref, sizeof(double) * 32 * 32);
    ALLOC(d_V, sizeof(double) * nGp * nGp);
    ALLOC(d_ST, sizeof(double) * Pc * nGp);
    ALLOC(d_WT, sizeof(double) * Pc * nGp);
    ALLOC(d_STt, sizeof(double) * Pc * nGp);
    ALLOC(d_act, sizeof(int) * (Pc / 16) * (nGp / 16));
    ALLOC(d_Wpart, sizeof(double) * cody::kMaxWSplit * Pc * nGp);
    ALLOC(d_Gpart, sizeof(double) * nxs * Pc * Pc);
    ALLOC(d_G, sizeof(double) * Pc * Pc);
    ALLOC(d_epart, sizeof(double) * ((Pc * Pc + nO * nO + 255) / 256));
    ALLOC(d_energy, sizeof(double) * 8);
    ALLOC(d_efield, sizeof(double));  // (every allocation starts zero-filled: E = 0)
    ALLOC(d_state, sizeof(double2) * nS);
    ALLOC(d_tmp, sizeof(double2) * nS);
    ALLOC(d_tmp2, sizeof(double2) * nS);
    ALLOC(d_rho1, sizeof(double2) * n2);
    // entries of rho2 with a repeated index pair stay zero (from the allocation's fill): the
    // kernels never write them
    ALLOC(d_rho2, sizeof(double2) * n4);
    ALLOC(d_r2can, sizeof(double2) * std::max<size_t>(1, (size_t)(nO * (nO - 1) / 2) * (nO * (nO - 1) / 2)));
    ALLOC(d_reginv, sizeof(double2) * n2);
    ALLOC(d_h1, sizeof(double2) * n2);
    ALLOC(d_hphi, sizeof(double2) * nG * nO);
    ALLOC(d_dCraw, sizeof(double2) * nC);
    ALLOC(d_draw, sizeof(double2) * nG * nO);
    // overlap partials: one per 64-point workgroup of the orbital kernel, or two per 16-point tile of
    // the fused int8 contraction
    ALLOC(d_part, sizeof(double2) * std::max(nPart, 2 * (nGp / 16)) * n2);
    ALLOC(d_partr, sizeof(double2) * 2 * n2);
    if (nO >= cody::kCorrQMinOrbs) ALLOC(d_qf, sizeof(double) * cody::corr_q_doubles(nO));
    ALLOC(d_part1, sizeof(double2) * ((nG + 255) / 256) * n2);
    ALLOC(d_zeros, sizeof(double) * (nG + 1));
    // CI tables in one blob: ptr arrays then entry arrays
    const size_t ns = ci.s_ptr.size(), nd = ci.d_ptr.size(), n1 = ci.r1_ptr.size(), nr2 = ci.r2_ptr.size();
    const size_t es = ci.singles.size(), ed = ci.doubles.size(), e1 = ci.r1.size(), e2 = ci.r2.size();
    const size_t ib = sizeof(int) * (ns + nd + n1 + nr2);
    const size_t ibp = 256 * ((ib + 255) / 256);
    const size_t eb = sizeof(cody::CiEntryDev) * (es + ed + e1 + e2);
    ALLOC(ci_blob, ibp + eb);
#undef ALLOC
    std::vector<char> host(ibp + eb, 0);
    int* ip = reinterpret_cast<int*>(host.data());
    auto cpyi = [&](const std::vector<int32_t>& v) {
      std::copy(v.begin(), v.end(), ip);
      ip += v.size();
    };
    cpyi(ci.s_ptr);
    cpyi(ci.d_ptr);
    cpyi(ci.r1_ptr);
    cpyi(ci.r2_ptr);
    auto* ep = reinterpret_cast<cody::CiEntryDev*>(host.data() + ibp);
    auto cpye = [&](const std::vector<cody::CiEntry>& v) {
      for (const auto& e : v) *ep++ = cody::CiEntryDev{e.ic, e.jc, e.code, e.f};
    };
    cpye(ci.singles);
    cpye(ci.doubles);
    cpye(ci.r1);
    cpye(ci.r2);
    if (hipMemcpy(ci_blob, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
      return cody::set_error(2, "hipMemcpy of CI tables failed");
    const int* ib0 = reinterpret_cast<const int*>(ci_blob);
    const auto* eb0 = reinterpret_cast<const cody::CiEntryDev*>(static_cast<char*>(ci_blob) + ibp);
    dci.s_ptr = ib0;
    dci.d_ptr = ib0 + ns;
    dci.r1_ptr = ib0 + ns + nd;
    dci.r2_ptr = ib0 + ns + nd + n1;
    dci.singles = eb0;
    dci.doubles = eb0 + es;
    dci.r1 = eb0 + es + ed;
    dci.r2 = eb0 + es + ed + e1;
    return 0;
  }

  // captured right-hand sides, replayed by the device-pointer entry points and the propagators;
  // dropped whenever a setter changes what the launches bake in (kernels, sizes, parameters)
  std::vector<RhsGraph> graphs;
  bool graphs_on = true;    // propagators replay captured steps (cody_set_graphs)
  bool rhs_graphs = false;  // single right-hand sides replay captured graphs (cody_set_graphs)
  bool capturing = false;  // enqueue_rhs is being captured (E(t) is then set per replay)
  unsigned generation = 0;  // bumped whenever captured launches become stale (batches check it)
  // E(t) as the one-body kernels read it; a captured Strang step points it at per-stage slots
  const double* efield_src = nullptr;
  void drop_graphs() {
    for (auto& g : graphs) hipGraphExecDestroy(g.exec);
    graphs.clear();
    prop.drop_graphs();
    ++generation;
  }

  void release() {
    drop_graphs();
    void* ptrs[] = {d_x,   d_w,    d_vext,  d_kref, d_V,     d_ST,    d_WT, d_STt, d_Wpart, d_act,
                    d_Gpart, d_G,  d_energy, d_epart, d_efield, d_state, d_tmp, d_tmp2, d_rho1, d_rho2,
                    d_reginv, d_r2can, d_h1, d_hphi, d_dCraw, d_draw, d_part, d_part1, d_zeros, d_src, d_h2, ci_blob,
                    d_partr, d_qf};
    for (void* p : ptrs)
      if (p) hipFree(p);
    prop.release();
    adj.release();
    void* lp[] = {d_Vt, d_Vi8, d_Si8, d_eV, d_eS, d_eSb, d_actK, d_Vhi, d_Vlo, d_Shi, d_Slo, d_Vbf, d_Sbf};
    for (void* p : lp)
      if (p) hipFree(p);
    if (stream) hipStreamDestroy(stream);
    stream = nullptr;
  }
};

// Several contexts whose right-hand sides are replayed as one captured graph.
// Instances of a batch with one orbital count and one contraction path, evaluated by one launch
// per kernel type: their argument table and block maps in device memory.
struct BatchGroup {
  int nO = 0, Pc = 0, maxC = 0;
  bool int8 = false;
  bool fused = false;  // int8 with the orbital correlation in the contraction (up to 8 orbitals)
  double cost = 0.0;  // estimated device time (for the assignment to chains)
  std::vector<int> members;
  cody::GroupInstance* d_inst = nullptr;
  double* d_qf = nullptr;  // Q of the members' orbital correlation (orbital counts from kCorrQMinOrbs)
  int2* d_map = nullptr;  // the kinds' maps back to back
  int off[cody::kPlanKinds] = {}, cnt[cody::kPlanKinds] = {};  // each kernel type's map
};

struct cody_batch {
  static constexpr int kChains = 4;  // parallel chains (the device exposes four hardware queues)
  std::vector<cody_ctx*> ctx;
  std::vector<unsigned> gen;
  std::vector<const void*> key;
  hipStream_t cs = nullptr;
  hipStream_t chain[kChains] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kChains] = {};
  hipGraphExec_t exec = nullptr;
  std::vector<BatchGroup> groups;  // grouped instances
  std::vector<int> singles;        // instances evaluated by their own launches
  // schedule: measured device time of each instance's own right-hand side (empty: the estimate
  // is used) and the split threshold, in chain shares, above which a group is split further
  std::vector<double> cost;
  double split_share = 1.0;
  int nchains = kChains;  // chains the schedule uses (1..kChains)
  void release_groups() {
    for (auto& g : groups) {
      if (g.d_inst) hipFree(g.d_inst);
      if (g.d_map) hipFree(g.d_map);
      if (g.d_qf) hipFree(g.d_qf);
    }
    groups.clear();
    singles.clear();
  }
  void release() {
    release_groups();
    if (exec) hipGraphExecDestroy(exec);
    exec = nullptr;
    for (int k = 0; k < kChains; ++k) {
      if (join[k]) hipEventDestroy(join[k]);
      if (chain[k]) hipStreamDestroy(chain[k]);
      join[k] = nullptr;
      chain[k] = nullptr;
    }
    if (fork) hipEventDestroy(fork);
    if (cs) hipStreamDestroy(cs);
    fork = nullptr;
    cs = nullptr;
  }
};
