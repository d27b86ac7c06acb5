// Orbital projector for callers that cannot allocate device memory (code under HIP graph
// capture, tight loops, harnesses that forbid hipMalloc): cody_project_orbs_dev_noalloc.
//
// Its one scratch buffer, the per-workgroup partial sums of the overlaps <phi_i|f_j>_w, is the
// file-scope array `scratch` below instead of library-allocated memory.  Every GPU the library
// runs on holds its own copy of the array until the process ends.  In host code `scratch` is
// only a stand-in for the symbol; its device address on each GPU comes from hipGetSymbolAddress,
// taken once per device (with that device current) on the first call there and kept.  Only that
// address is passed to kernels.
//
// All calls on one GPU share that GPU's array and it keeps the last call's partial sums: two
// calls that overlap in time on one GPU (different streams, no ordering between them) race on
// it.  The overlap kernel writes every partial that the projection kernel then reads, so nothing
// has to be cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/cody_mctdhx.h"
#include "mctdhx_kernels.hpp"

constexpr int kMaxPoints = CODY_NOALLOC_MAX_POINTS;
constexpr int kMaxParts = (kMaxPoints + 255) / 256;  // one partial block per 256 grid points
// 256 x 12^2 double2 = 576 KiB per GPU
static __device__ __align__(256) double2 scratch[kMaxParts * cody::kMaxOrbs * cody::kMaxOrbs];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first call there.
int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return cody::set_error(CODY_ERR_HIP, (std::string("hipGetDevice: ") + hipGetErrorString(e)).c_str());
  if (dev < 0 || dev >= kMaxDevices) return cody::set_error(CODY_ERR_UNSUPPORTED, "cody_project_orbs_dev_noalloc: device index above 63");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return cody::set_error(CODY_ERR_HIP, (std::string("hipGetSymbolAddress(scratch): ") +
                                            (e != hipSuccess ? hipGetErrorString(e) : "null address")).c_str());
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return CODY_OK;
}

}  // namespace

extern "C" int cody_project_orbs_dev_noalloc(int nG, int nO, const double* d_w, const void* d_orbs, void* d_dorbs,
                                             void* stream) {
  if (nG < 1 || nO < 1 || !d_w || !d_orbs || !d_dorbs)
    return cody::set_error(CODY_ERR_ARG, "cody_project_orbs_dev_noalloc: bad arguments");
  if (nG > kMaxPoints || nO > cody::kMaxOrbs)
    return cody::set_error(CODY_ERR_ARG, "cody_project_orbs_dev_noalloc: more than 65536 grid points or 12 orbitals "
                                         "(the size of its static scratch)");
  if (d_orbs == d_dorbs) return cody::set_error(CODY_ERR_ARG, "cody_project_orbs_dev_noalloc: orbs and dorbs alias");
  void* part = nullptr;
  if (int r = scratch_address(&part)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const auto* orbs = static_cast<const double2*>(d_orbs);
  auto* f = static_cast<double2*>(d_dorbs);
  cody::launch_overlap_parts(orbs, d_w, f, nG, nO, static_cast<double2*>(part), st);
  cody::launch_project_inplace(orbs, static_cast<const double2*>(part), nG, nO, f, st);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return cody::set_error(CODY_ERR_HIP, (std::string("cody_project_orbs_dev_noalloc: ") + hipGetErrorString(e)).c_str());
  return CODY_OK;
}
