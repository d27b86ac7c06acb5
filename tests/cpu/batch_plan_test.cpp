// CPU unit test of the grouped-batch planner (codyfortran_amd/csrc/batch_plan.cpp), built with
// AddressSanitizer and UBSan by tests/test_batch_plan.py.  Exits non-zero on the first failure.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "batch_plan.hpp"

using namespace cody;

static int g_fail = 0;
#define EXPECT(cond, ...)                         \
  do {                                            \
    if (!(cond)) {                                \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                   \
      std::printf("\n");                          \
      ++g_fail;                                   \
    }                                             \
  } while (0)

static PlanShape shape(int nO, int nC, int nG, int nsplit, bool int8) {
  PlanShape s;
  s.nO = nO;
  s.nC = nC;
  s.nG = nG;
  s.nGp = 256 * ((nG + 255) / 256);
  s.Pc = 16 * ((nO * (nO + 1) + 15) / 16);
  s.nsplit = nsplit;
  const int k64 = s.nGp / 64;
  s.nxs = 1;
  for (int v : {32, 16, 8, 4, 2, 1})
    if (k64 % v == 0) {
      s.nxs = v;
      break;
    }
  s.int8 = int8;
  s.fused = int8 && nO <= kPlanFusedMaxOrbs;
  return s;
}

static long long binom(int n, int k) {
  long long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

int main() {
  // 1. known answer: the 4096-point system with 8 orbitals, 4 fermions (70 configurations), FP64
  //    contraction split in 4, as its own launches size it (launch_rhs_front: 212 CI-density
  //    blocks + 16 x 8 one-body blocks + 80 x 4096 / 256 pair-source blocks; k_mf_f64_w: 64 row
  //    groups x 4; k_rhs_wsum: 1 + 80 x 4096 / 512; k_rhs_qform: Q (8 x 18 x 64 / 4 doubles / 256) +
  //    5 x 32 Gram; k_rhs_mid: 64 x 4 correlation;
  //    k_rhs_post: 26 reduction + 16 projection; k_ci_apply_project: 18)
  {
    const PlanShape s = shape(8, 70, 4096, 4, false);
    EXPECT(check_plan_shape(s).empty(), "headline shape refused: %s", check_plan_shape(s).c_str());
    int64_t c[kPlanKinds];
    plan_block_counts(s, c);
    const int64_t want[kPlanKinds] = {212 + 128 + 1280, 256, 641, 9 + 160, 64 * 4, 26 + 16, 18};
    for (int k = 0; k < kPlanKinds; ++k) EXPECT(c[k] == want[k], "kind %d: %lld blocks, expected %lld", k, (long long)c[k], (long long)want[k]);
    // the int8 path: pair sources with their digits per 16 x 64 block (80 x 4096 / 1024), one
    // contraction workgroup per 16-row tile writing W^T (no k-split sum, R in the Gram launch)
    const PlanShape s8 = shape(8, 70, 4096, 1, true);
    plan_block_counts(s8, c);
    EXPECT(c[kPlanFront] == 212 + 128 + 320, "int8 front blocks %lld", (long long)c[kPlanFront]);
    EXPECT(c[kPlanContract] == 256, "int8 contraction blocks %lld", (long long)c[kPlanContract]);
    EXPECT(c[kPlanWsum] == 0, "int8 sum blocks %lld", (long long)c[kPlanWsum]);
    // the correlation is formed by the contraction: 5 x 5 direct G tiles + h1 + two partial-sum
    // workgroups + R; no G-reduction launch; the projection (16 workgroups) beside the CI apply
    EXPECT(c[kPlanMid] == 25 + 4, "int8 Gram blocks %lld", (long long)c[kPlanMid]);
    EXPECT(c[kPlanPost] == 0, "int8 post blocks %lld", (long long)c[kPlanPost]);
    EXPECT(c[kPlanApply] == 18 + 16, "int8 final blocks %lld", (long long)c[kPlanApply]);
    // 12 orbitals: Q (24 tiles x 39 k steps x 64 doubles, four per thread) is formed once per
    // right-hand side beside the Gram partials; the correlation takes one workgroup per 64 points and
    // group of 3 orbitals
    plan_block_counts(shape(12, 924, 4096, 4, false), c);
    EXPECT(c[kPlanQform] == 59 + 10 * 32, "Q + Gram blocks %lld", (long long)c[kPlanQform]);
    EXPECT(c[kPlanMid] == 64 * 4, "12-orbital correlation blocks %lld", (long long)c[kPlanMid]);
    // 12 orbitals on the int8 path: the 16-row contraction (no correlation epilogue) writes W^T;
    // R in the k-split-sum launch, Q + Gram, the MFMA correlation, as the FP64 path
    {
      int64_t d[kPlanKinds];
      const PlanShape u = shape(12, 924, 4096, 1, true);
      EXPECT(check_plan_shape(u).empty(), "unfused int8 shape refused: %s", check_plan_shape(u).c_str());
      plan_block_counts(u, d);
      EXPECT(d[kPlanContract] == 256 && d[kPlanWsum] == 1 && d[kPlanQform] == 59 + 10 * 32 && d[kPlanMid] == 64 * 4,
             "unfused int8 blocks %lld %lld %lld %lld", (long long)d[kPlanContract], (long long)d[kPlanWsum],
             (long long)d[kPlanQform], (long long)d[kPlanMid]);
      EXPECT(d[kPlanFront] == 282 + 192 + 640, "unfused int8 front blocks %lld", (long long)d[kPlanFront]);
    }
    // 924 configurations: four CI density entries per wave, 16 per workgroup: (144 + 66^2) / 16
    // = 281.25 -> 282 blocks, then 16 x 12 one-body and 160 x 4096 / 256 pair-source blocks
    EXPECT(c[kPlanFront] == 282 + 192 + 2560, "12-orbital front blocks %lld", (long long)c[kPlanFront]);
  }
  // 2. the batched workload's shapes (N_g in 511/1026/2046/4096, M in 4/6/8/12, N = M/2), every
  //    FP64 split and the int8 path: built plans pass the independent check, cover each
  //    instance's blocks exactly once and in order
  for (bool int8 : {false, true})
    for (int nO : {4, 6, 8, 12}) {
      std::vector<PlanShape> members;
      int k = 0;
      for (int nG : {511, 1026, 2046, 4096})
        for (int split : {1, 2, 4, 8}) {
          if (int8 && split != 1) continue;
          members.push_back(shape(nO, (int)binom(nO, nO / 2), nG, split, int8));
          ++k;
        }
      GroupPlan plan;
      const std::string why = build_group_plan(members, plan);
      EXPECT(why.empty(), "nO %d int8 %d refused: %s", nO, (int)int8, why.c_str());
      if (!why.empty()) continue;
      EXPECT(check_group_plan(members, plan).empty(), "check failed");
      // brute force: per kind, count each (instance, block) once and compare with the counts
      for (int kind = 0; kind < kPlanKinds; ++kind) {
        std::vector<std::vector<int>> seen(members.size());
        for (size_t m = 0; m < members.size(); ++m) {
          int64_t c[kPlanKinds];
          plan_block_counts(members[m], c);
          seen[m].assign((size_t)c[kind], 0);
        }
        for (int64_t i = plan.off[kind]; i < plan.off[kind] + plan.cnt[kind]; ++i) {
          const BlockRef r = plan.map[(size_t)i];
          if (r.inst < 0 || r.inst >= (int)members.size() || r.block < 0 || r.block >= (int)seen[r.inst].size()) {
            EXPECT(false, "entry out of range");
            break;
          }
          ++seen[r.inst][r.block];
        }
        for (auto& v : seen)
          for (int n : v) EXPECT(n == 1, "a block covered %d times", n);
      }
    }
  // 3. refused shapes
  {
    auto refused = [](PlanShape s) { return !check_plan_shape(s).empty(); };
    EXPECT(refused(shape(8, 70, 4096, 16, false)), "16-way split accepted");  // (the one that overran d_Wpart)
    EXPECT(refused(shape(8, 70, 4096, 3, false)), "3-way split accepted");
    EXPECT(refused(shape(12, 924, 4096, 4, true)), "int8 with a 4-way split accepted (12 orbitals)");
    {
      PlanShape f = shape(12, 924, 4096, 1, true);
      f.fused = true;
      EXPECT(refused(f), "the fused int8 path accepted for 12 orbitals");
    }
    EXPECT(refused(shape(8, 70, 4096, 4, true)), "int8 with a 4-way split accepted");
    EXPECT(refused(shape(8, 70, 16640, 1, true)), "int8 with 260 k blocks accepted");
    EXPECT(refused(shape(13, 70, 4096, 1, false)), "13 orbitals accepted");
    EXPECT(refused(shape(12, 1025, 4096, 1, false)), "1025 configurations accepted");
    PlanShape s = shape(8, 70, 4096, 4, false);
    s.Pc = 64;
    EXPECT(refused(s), "wrong pair-column count accepted");
    s = shape(8, 70, 4096, 4, false);
    s.nGp = 4000;
    EXPECT(refused(s), "unpadded grid accepted");
    s = shape(8, 70, 4096, 4, false);
    s.nxs = 3;
    EXPECT(refused(s), "3 Gram ranges accepted");
    // the smallest padded grid (256 points, 8 orbitals: 8 k blocks of 2 x 16 points) still takes
    // the 8-way split
    EXPECT(!refused(shape(8, 70, 256, 8, false)), "8-way split of 8 k blocks refused");
    GroupPlan plan;
    std::vector<PlanShape> mixed = {shape(8, 70, 4096, 4, false), shape(6, 20, 4096, 4, false)};
    EXPECT(!build_group_plan(mixed, plan).empty(), "mixed orbital counts accepted");
    std::vector<PlanShape> paths = {shape(8, 70, 4096, 4, false), shape(8, 70, 4096, 4, true)};
    EXPECT(!build_group_plan(paths, plan).empty(), "mixed contraction paths accepted");
    EXPECT(!build_group_plan({}, plan).empty(), "empty group accepted");
  }
  // 4. tampered maps are caught by the check
  {
    std::vector<PlanShape> members = {shape(6, 20, 1026, 2, false), shape(6, 20, 511, 1, false)};
    GroupPlan plan;
    EXPECT(build_group_plan(members, plan).empty(), "build failed");
    GroupPlan t = plan;
    std::swap(t.map[(size_t)t.off[kPlanMid]], t.map[(size_t)t.off[kPlanMid] + 1]);
    EXPECT(!check_group_plan(members, t).empty(), "swapped entries accepted");
    t = plan;
    t.map[(size_t)t.off[kPlanFront]].inst = 2;
    EXPECT(!check_group_plan(members, t).empty(), "instance out of range accepted");
    t = plan;
    t.map[(size_t)(t.off[kPlanPost] + t.cnt[kPlanPost] - 1)].block += 1;
    EXPECT(!check_group_plan(members, t).empty(), "block beyond the grid accepted");
    t = plan;
    t.map.pop_back();
    EXPECT(!check_group_plan(members, t).empty(), "truncated map accepted");
    t = plan;
    t.off[kPlanWsum] += 1;
    EXPECT(!check_group_plan(members, t).empty(), "shifted kind offset accepted");
  }
  // 5. grids beyond 2^31 - 1 workgroups are refused before anything is allocated
  {
    std::vector<PlanShape> huge(4, shape(12, 924, 1 << 30, 1, false));
    GroupPlan plan;
    const std::string why = build_group_plan(huge, plan);
    EXPECT(!why.empty() && plan.map.empty(), "2^31 overflow not refused");
  }
  std::printf(g_fail ? "batch_plan_test: %d failures\n" : "batch_plan_test: all passed\n", g_fail);
  return g_fail ? 1 : 0;
}
