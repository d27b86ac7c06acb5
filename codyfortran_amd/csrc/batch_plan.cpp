// Host-only planning of grouped right-hand sides (see batch_plan.hpp).
#include "batch_plan.hpp"

#include <climits>
#include <cstdio>

namespace cody {

namespace {

std::string fmt(const char* f, int a, int b = 0, int c = 0) {
  char buf[192];
  std::snprintf(buf, sizeof buf, f, a, b, c);
  return buf;
}

bool is_pow2_upto(int v, int hi) { return v >= 1 && v <= hi && (v & (v - 1)) == 0; }

}  // namespace

std::string check_plan_shape(const PlanShape& s) {
  if (s.nO < 1 || s.nO > kPlanMaxOrbs) return fmt("%d orbitals are outside 1..%d", s.nO, kPlanMaxOrbs);
  if (s.nC < 1 || s.nC > kPlanMaxConf) return fmt("%d configurations are outside 1..%d", s.nC, kPlanMaxConf);
  if (s.nG < 1 || s.nGp != 256 * ((s.nG + 255) / 256))
    return fmt("padded grid %d is not the multiple of 256 above %d points", s.nGp, s.nG);
  const int pc = 16 * ((s.nO * (s.nO + 1) + 15) / 16);
  if (s.Pc != pc) return fmt("%d pair columns for %d orbitals (expected %d)", s.Pc, s.nO, pc);
  const int k64 = s.nGp / 64;
  if (!is_pow2_upto(s.nxs, 32) || k64 % s.nxs != 0)
    return fmt("Gram x ranges %d do not divide the %d 64-point blocks", s.nxs, k64);
  if (s.fused != (s.int8 && s.nO <= kPlanFusedMaxOrbs))
    return fmt("the fused int8 path is for int8 instances of at most %d orbitals (%d orbitals)", kPlanFusedMaxOrbs, s.nO);
  if (s.int8) {
    if (s.Pc / 16 > kPlanInt8MaxTiles) return fmt("%d pair columns exceed the int8 contraction's %d", s.Pc, 16 * kPlanInt8MaxTiles);
    if (s.nsplit != 1) return fmt("int8 k split %d (the int8 contraction takes one range)", s.nsplit);
    if (k64 > kPlanInt8MaxBlocks64) return fmt("%d k blocks exceed the int8 contraction's %d", k64, kPlanInt8MaxBlocks64);
  } else {
    if (s.Pc / 16 > kPlanF64MaxTiles) return fmt("%d pair columns exceed the FP64 contraction's %d", s.Pc, 16 * kPlanF64MaxTiles);
    if (!is_pow2_upto(s.nsplit, 8)) return fmt("k split %d is not one of 1, 2, 4, 8", s.nsplit);
    const int kh = s.Pc / 16 <= 5 ? 2 : 1;
    if ((s.nGp / 16 / kh) % s.nsplit != 0)
      return fmt("k split %d does not divide the %d k blocks", s.nsplit, s.nGp / 16 / kh);
  }
  return {};
}

void plan_block_counts(const PlanShape& s, int64_t cnt[kPlanKinds]) {
  const int64_t n2 = (int64_t)s.nO * s.nO, np = (int64_t)s.nO * (s.nO - 1) / 2;
  const int64_t plane = (int64_t)s.Pc * s.nGp;
  const int64_t nObx = (s.nG + 255) / 256;
  const int64_t per = 4 * (s.nC > kPlanRdmWideConf ? kPlanRdmWideItems : 1);
  const int64_t nRdm = (n2 + np * np + per - 1) / per;
  const int64_t nRed = ((int64_t)s.Pc * s.Pc + n2 + 255) / 256;
  const int64_t wsumFirst = s.fused ? 0 : 1;  // R in the fused int8 path's Gram launch, else in the k-split sum's
  // pair sources: one 16 x 16 fragment per workgroup, or for the int8 path (digits with a scale per
  // line and 64-point block) one 16 x 64 block
  cnt[kPlanFront] = nRdm + nObx * s.nO + plane / (s.int8 ? 1024 : 256);
  // the int8 contraction: one workgroup per 16-row tile; the FP64 one: 64-row groups x k ranges
  cnt[kPlanContract] = s.int8 ? (int64_t)(s.nGp / 16) : (int64_t)(s.nGp / 64) * s.nsplit;
  cnt[kPlanWsum] = wsumFirst + (s.nsplit > 1 ? (plane / 2 + 255) / 256 : 0);
  const bool corrQ = !s.fused && s.nO >= kPlanCorrQMinOrbs;
  const int64_t nGram = (int64_t)(s.Pc / 16) * s.nxs;
  // Q (four doubles per thread) beside the Gram partials
  cnt[kPlanQform] = corrQ ? (plan_corr_q_doubles(s.nO) / 4 + 255) / 256 + nGram : 0;
  // Gram + orbital correlation per 64 points and orbital, or the correlation on FP64 MFMA per 64
  // points and orbital group; the int8 path's contraction forms the correlation, and beside the Gram
  // two workgroups sum its two sets of overlap partials and one forms R
  // (the fused int8 path: G directly, one 1024-thread workgroup per 16 x 16 tile, then h1, the two
  // overlap sums and R; no separate G reduction and projection launch: the projection runs beside
  // the CI apply)
  const int64_t nt = s.Pc / 16;
  cnt[kPlanMid] = s.fused ? nt * nt + 4 : (corrQ ? (int64_t)(s.nGp / 64) * kPlanCorrQGroups : nGram + (int64_t)(s.nGp / 64) * s.nO);
  cnt[kPlanPost] = s.fused ? 0 : nRed + nObx;
  cnt[kPlanApply] = (s.nC + 3) / 4 + (s.fused ? nObx : 0);
}

std::string build_group_plan(const std::vector<PlanShape>& members, GroupPlan& out) {
  out = GroupPlan();
  if (members.empty()) return "an empty group";
  const int nM = static_cast<int>(members.size());
  std::vector<int64_t> counts((size_t)nM * kPlanKinds);
  for (int m = 0; m < nM; ++m) {
    const PlanShape& s = members[m];
    const std::string why = check_plan_shape(s);
    if (!why.empty()) return fmt("instance %d: ", m) + why;
    if (s.nO != members[0].nO || s.int8 != members[0].int8 || s.fused != members[0].fused)
      return fmt("instance %d: another orbital count or path", m);
    plan_block_counts(s, &counts[(size_t)m * kPlanKinds]);
  }
  int64_t total = 0;
  for (int k = 0; k < kPlanKinds; ++k) {
    out.off[k] = total;
    for (int m = 0; m < nM; ++m) out.cnt[k] += counts[(size_t)m * kPlanKinds + k];
    if (out.cnt[k] > INT_MAX) return fmt("a grouped grid of kind %d exceeds 2^31 - 1 workgroups", k);
    total += out.cnt[k];
  }
  out.map.resize((size_t)total);
  for (int k = 0; k < kPlanKinds; ++k) {
    int64_t o = out.off[k];
    for (int m = 0; m < nM; ++m)
      for (int64_t b = 0; b < counts[(size_t)m * kPlanKinds + k]; ++b) out.map[(size_t)o++] = BlockRef{m, (int32_t)b};
  }
  return check_group_plan(members, out);
}

std::string check_group_plan(const std::vector<PlanShape>& members, const GroupPlan& plan) {
  const int nM = static_cast<int>(members.size());
  int64_t expect = 0;
  for (int k = 0; k < kPlanKinds; ++k) {
    if (plan.off[k] != expect) return fmt("kind %d starts at a gap or overlap", k);
    if (plan.cnt[k] < 0 || plan.off[k] + plan.cnt[k] > (int64_t)plan.map.size()) return fmt("kind %d overruns the map", k);
    expect += plan.cnt[k];
    // every member's blocks of this kind, once and in order
    int cur = 0;
    int64_t next = 0, own[kPlanKinds];
    if (nM > 0) plan_block_counts(members[0], own);
    for (int64_t i = plan.off[k]; i < plan.off[k] + plan.cnt[k]; ++i) {
      const BlockRef r = plan.map[(size_t)i];
      if (r.inst < 0 || r.inst >= nM) return fmt("kind %d, entry %d: instance %d out of range", k, (int)(i - plan.off[k]), r.inst);
      while (cur < nM && next == own[k]) {  // member `cur` done: move on
        ++cur;
        next = 0;
        if (cur < nM) plan_block_counts(members[cur], own);
      }
      if (r.inst != cur || r.block != next) return fmt("kind %d: instance %d block %d out of order", k, r.inst, r.block);
      if (r.block < 0 || r.block >= own[k]) return fmt("kind %d: block %d outside instance %d's grid", k, r.block, r.inst);
      ++next;
    }
    while (cur < nM && next == own[k]) {
      ++cur;
      next = 0;
      if (cur < nM) plan_block_counts(members[cur], own);
    }
    if (cur != nM) return fmt("kind %d: instance %d's blocks are missing", k, cur);
  }
  if (expect != (int64_t)plan.map.size()) return "the kinds do not tile the map";
  return {};
}

}  // namespace cody
