// Host-only planning of grouped right-hand sides: the block maps that let one launch per kernel
// type run the right-hand sides of several independent instances of one orbital count.  Pure C++
// (no HIP), so that the maps and their checks are unit-tested on the CPU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cody {

// One workgroup of a grouped launch: block `block` of instance `inst` (the layout of HIP's int2).
struct BlockRef {
  int32_t inst, block;
};
static_assert(sizeof(BlockRef) == 8, "a block reference is uploaded as int2");

// Limits of the grouped kernels (the device code's kMaxOrbs, kR16MaxBlocks64 and contraction
// tiles; cody_api.hip asserts that they agree).
constexpr int kPlanMaxOrbs = 12;           // orbitals per instance
constexpr int kPlanMaxConf = 1024;         // CI configurations per instance
constexpr int kPlanInt8MaxTiles = 10;      // 16-column tiles of the int8 contraction
constexpr int kPlanFusedMaxOrbs = 8;       // orbitals of the int8 contraction with the correlation epilogue
constexpr int kPlanInt8MaxBlocks64 = 256;  // 64-point k blocks of the int8 contraction (one range)
constexpr int kPlanF64MaxTiles = 10;       // 16-column tiles of the FP64 contraction
// CI density role: entries per wave (the device's ci_rdm_items_per_wave): 1, or 4 above 256
// configurations; a workgroup holds four waves
constexpr int kPlanRdmWideConf = 256, kPlanRdmWideItems = 4;

// The kernel types of a right-hand side, in launch order.
enum PlanKind { kPlanFront, kPlanContract, kPlanWsum, kPlanQform, kPlanMid, kPlanPost, kPlanApply, kPlanKinds };

// Orbital counts from which the FP64 path runs the orbital correlation on FP64 MFMA: one workgroup
// per 64 points and each of kPlanCorrQGroups orbital groups, after a launch (kPlanQform) that forms
// the rho2'-weighted matrix Q of plan_corr_q_doubles(nO) doubles, four per thread, beside the Gram
// partials.
constexpr int kPlanCorrQMinOrbs = 8, kPlanCorrQGroups = 4;
constexpr int plan_corr_q_doubles(int nO) { return nO * ((2 * nO + 15) / 16) * ((nO * (nO + 1) + 3) / 4) * 64; }

// The sizes an instance's launches depend on.
struct PlanShape {
  int nO = 0;      // orbitals
  int nC = 0;      // configurations
  int nG = 0;      // grid points
  int nGp = 0;     // padded grid (a multiple of 256)
  int Pc = 0;      // pair columns, 16 ceil(nO (nO + 1) / 16)
  int nsplit = 1;  // contraction k ranges (the int8 path: 1, the 16-row kernel writes W^T)
  int nxs = 1;     // x ranges of the Gram partials
  bool int8 = false;
  bool fused = false;  // int8 with the orbital correlation in the contraction (exactly: int8, nO <= 8)
};

// Workgroups of each kernel type of one instance's right-hand side: the grids of its own launches
// (front: CI densities + one-body part + pair sources, on the int8 path with S's digits; the
// contraction, with the orbital correlation on the int8 path; the k-split sum with R (FP64 path);
// Q for nO >= kPlanCorrQMinOrbs (FP64 path); Gram + orbital correlation, or on the int8 path Gram + the sums
// of the overlap partials + R; G reduction + orbital projection; CI apply).  The shape must have
// passed check_plan_shape.
void plan_block_counts(const PlanShape& s, int64_t cnt[kPlanKinds]);

// Returns an empty string for a shape the grouped kernels accept, else the reason.
std::string check_plan_shape(const PlanShape& s);

// A group's maps, the kinds' maps back to back: map[off[k] + i], i < cnt[k].
struct GroupPlan {
  std::vector<BlockRef> map;
  int64_t off[kPlanKinds] = {};
  int64_t cnt[kPlanKinds] = {};
};

// Builds the maps of the instances `members` (all of one orbital count and contraction path):
// for every kind, each member's blocks 0..count-1 in member order.  Returns an empty string, or
// the reason it refuses (a shape check, mixed orbital counts or paths, a grid longer than 2^31 - 1
// workgroups).
std::string build_group_plan(const std::vector<PlanShape>& members, GroupPlan& out);

// Independent check of a plan against the shapes: every entry names a member and a block inside
// that member's grid for its kind, each member's blocks appear exactly once and in order, and the
// kinds' ranges tile the map.  Returns an empty string or the first violation.
std::string check_group_plan(const std::vector<PlanShape>& members, const GroupPlan& plan);

}  // namespace cody
