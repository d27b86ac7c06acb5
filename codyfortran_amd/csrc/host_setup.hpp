// Host-side setup for the MI355X MCTDHX path: FEDVR grid pieces and the CI connectivity
// tables that the device kernels consume (one fermionic body type).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cody {

// ---- FEDVR (replaces fortran-stdlib GLL + src/Utils/M_Utils_Fedvr.f90 + M_Utils_DerivativeFedvr.f90)
int fedvr_npoints(int nE, int nLoc, bool xminExcl, bool xmaxExcl);
void gll_nodes(int n, double a, double b, double* x, double* w);
void fedvr_grid(double xmin, double xmax, int nE, int nLoc, bool xminExcl, bool xmaxExcl,
                double* x, double* w);
void fedvr_kref(int nLoc, double* kref);

// ---- CI tables for one fermionic body type, lexicographic configuration order.
// Entry of a gather list: contribution  f * conj(C[jc]) * C[ic]  (densities) or
// f * h[code] * C[jc] accumulated into dC[ic] (H-apply).
struct CiEntry {
  int32_t ic, jc, code;
  float f;  // +-1
};

struct CiTables {
  int nOrbs = 0, nBodies = 0, nConf = 0;
  std::vector<uint32_t> occ;  // occupation bit pattern per configuration
  // H-apply lists grouped by target configuration ic (CSR over ic)
  std::vector<int32_t> s_ptr, d_ptr;
  std::vector<CiEntry> singles, doubles;
  // density gather lists grouped by code (CSR over code):
  // rho1Lin[code1], code1 = i*nO + j (i = created, j = destroyed)
  // rho2Lin[code2], code2 = ((i1*nO + j1)*nO + i2)*nO + j2, i1 < i2, j1 < j2
  std::vector<int32_t> r1_ptr, r2_ptr;
  std::vector<CiEntry> r1, r2;
};

bool build_ci_tables(int nOrbs, int nBodies, CiTables& t, std::string& err);

}  // namespace cody
