// Host-side setup: FEDVR grid pieces and CI connectivity tables.
#include "host_setup.hpp"

#include <algorithm>
#include <cmath>
#include <unordered_map>

namespace cody {

int fedvr_npoints(int nE, int nLoc, bool xminExcl, bool xmaxExcl) {
  // src/Utils/M_Utils_Fedvr.f90:108-111
  return nE * (nLoc - 1) + 1 - (xminExcl ? 1 : 0) - (xmaxExcl ? 1 : 0);
}

// Legendre P_m(t) and P_m'(t) (three-term recurrence; t strictly inside (-1, 1) for P').
static void legendre_pm(int m, double t, double& p, double& dp) {
  double pm1 = 1.0, pcur = t;
  if (m == 0) {
    p = 1.0;
    dp = 0.0;
    return;
  }
  for (int k = 1; k < m; ++k) {
    double pn = ((2 * k + 1) * t * pcur - k * pm1) / (k + 1);
    pm1 = pcur;
    pcur = pn;
  }
  p = pcur;
  dp = m * (pm1 - t * pcur) / (1.0 - t * t);
}

// Gauss-Legendre-Lobatto rule with n points on [a, b]: ends plus the zeros of P'_{n-1},
// located by Newton iteration from Chebyshev-Gauss-Lobatto guesses; weights 2/(n(n-1)P_{n-1}^2).
void gll_nodes(int n, double a, double b, double* x, double* w) {
  const int m = n - 1;
  std::vector<double> t(n), wt(n);
  t[0] = -1.0;
  t[m] = 1.0;
  for (int i = 1; i < m; ++i) {
    double ti = -std::cos(M_PI * i / m);
    for (int it = 0; it < 60; ++it) {
      double p, dp;
      legendre_pm(m, ti, p, dp);
      double d2p = (2.0 * ti * dp - m * (m + 1.0) * p) / (1.0 - ti * ti);
      double step = dp / d2p;
      ti -= step;
      if (std::fabs(step) < 1e-16) break;
    }
    t[i] = ti;
  }
  for (int i = 0; i < n; ++i) {
    double p = 1.0, dp;
    if (i == 0)
      p = (m % 2) ? -1.0 : 1.0;
    else if (i < m)
      legendre_pm(m, t[i], p, dp);
    wt[i] = 2.0 / (m * (m + 1.0) * p * p);
  }
  for (int i = 0; i < n; ++i) {
    x[i] = a + 0.5 * (b - a) * (t[i] + 1.0);
    w[i] = 0.5 * (b - a) * wt[i];
  }
}

void fedvr_grid(double xmin, double xmax, int nE, int nLoc, bool xminExcl, bool xmaxExcl,
                double* x, double* w) {
  // Shared bridge points carry the sum of both elements' weights (src/Utils/M_Utils_Fedvr.f90:128-145).
  const int nP = fedvr_npoints(nE, nLoc, xminExcl, xmaxExcl);
  std::fill(w, w + nP, 0.0);
  std::vector<double> ex(nLoc), ew(nLoc);
  const double h = (xmax - xmin) / nE;
  for (int e = 0; e < nE; ++e) {
    gll_nodes(nLoc, xmin + e * h, xmin + (e + 1) * h, ex.data(), ew.data());
    for (int l = 0; l < nLoc; ++l) {
      int g = e * (nLoc - 1) + l - (xminExcl ? 1 : 0);  // global index of this node
      if (g < 0 || g >= nP) continue;
      x[g] = ex[l];
      w[g] += ew[l];
    }
  }
}

void fedvr_kref(int nLoc, double* kref) {
  // K(i,j) = -sum_k w_k L_i'(x_k) L_j'(x_k) on the reference element [-1/2, 1/2]
  // (src/Utils/M_Utils_DerivativeFedvr.f90:52-67), with the Lagrange derivative in barycentric form.
  std::vector<double> xr(nLoc), wr(nLoc), bw(nLoc), D(nLoc * nLoc);
  gll_nodes(nLoc, -0.5, 0.5, xr.data(), wr.data());
  for (int j = 0; j < nLoc; ++j) {
    double p = 1.0;
    for (int k = 0; k < nLoc; ++k)
      if (k != j) p *= (xr[j] - xr[k]);
    bw[j] = 1.0 / p;
  }
  for (int i = 0; i < nLoc; ++i) {
    double diag = 0.0;
    for (int j = 0; j < nLoc; ++j) {
      if (j == i) continue;
      double dij = (bw[j] / bw[i]) / (xr[i] - xr[j]);
      D[i + j * nLoc] = dij;  // L_j'(x_i)
      diag -= dij;
    }
    D[i + i * nLoc] = diag;
  }
  for (int i = 0; i < nLoc; ++i)
    for (int j = 0; j < nLoc; ++j) {
      double s = 0.0;
      for (int k = 0; k < nLoc; ++k) s -= wr[k] * D[k + i * nLoc] * D[k + j * nLoc];
      kref[i + j * nLoc] = s;
    }
}

// ------------------------------------------------------------------------------------------
// CI tables
// ------------------------------------------------------------------------------------------

namespace {

// a+_{cr[0]} .. a+_{cr[n-1]} a_{de[n-1]} .. a_{de[0]} acting on an occupation pattern; returns
// false if the result vanishes.  Sign rule: one factor -1 per occupied orbital with a lower
// index at the moment each operator acts (annihilators in list order, creators in reverse),
// src/ConfigList/AllActive/Fermions/S_ConfigList_AllActive_Fermions.f90:232-330.
bool apply_string(uint32_t occ, const int* cr, const int* de, int n, uint32_t& out, int& sign) {
  int s = 1;
  for (int i = 0; i < n; ++i) {
    uint32_t bit = 1u << de[i];
    if (!(occ & bit)) return false;
    if (__builtin_popcount(occ & (bit - 1u)) & 1) s = -s;
    occ ^= bit;
  }
  for (int i = n - 1; i >= 0; --i) {
    uint32_t bit = 1u << cr[i];
    if (occ & bit) return false;
    if (__builtin_popcount(occ & (bit - 1u)) & 1) s = -s;
    occ |= bit;
  }
  out = occ;
  sign = s;
  return true;
}

template <class KeyFn>
void csr_by(const std::vector<CiEntry>& in, int nKeys, KeyFn key, std::vector<int32_t>& ptr,
            std::vector<CiEntry>& out) {
  ptr.assign(nKeys + 1, 0);
  for (const auto& e : in) ptr[key(e) + 1]++;
  for (int k = 0; k < nKeys; ++k) ptr[k + 1] += ptr[k];
  out.resize(in.size());
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (const auto& e : in) out[fill[key(e)]++] = e;
}

}  // namespace

bool build_ci_tables(int nOrbs, int nBodies, CiTables& t, std::string& err) {
  if (nOrbs < 1 || nOrbs > 16 || nBodies < 1 || nBodies > nOrbs) {
    err = "CI tables: need 1 <= nBodies <= nOrbs <= 16";
    return false;
  }
  t.nOrbs = nOrbs;
  t.nBodies = nBodies;
  t.occ.clear();
  // lexicographic N-of-M combinations (the GSL iterator order of
  // src/Utils/M_Utils_CombinationGslLib.f90:53-86)
  std::vector<int> comb(nBodies);
  for (int i = 0; i < nBodies; ++i) comb[i] = i;
  for (;;) {
    uint32_t o = 0;
    for (int c : comb) o |= 1u << c;
    t.occ.push_back(o);
    int i = nBodies - 1;
    while (i >= 0 && comb[i] == nOrbs - nBodies + i) --i;
    if (i < 0) break;
    ++comb[i];
    for (int j = i + 1; j < nBodies; ++j) comb[j] = comb[j - 1] + 1;
  }
  t.nConf = static_cast<int>(t.occ.size());
  std::unordered_map<uint32_t, int> rank;
  for (int i = 0; i < t.nConf; ++i) rank[t.occ[i]] = i;
  const int nO = nOrbs;
  std::vector<CiEntry> s_all, d_all;
  t.s_ptr.assign(t.nConf + 1, 0);
  t.d_ptr.assign(t.nConf + 1, 0);
  for (int ic = 0; ic < t.nConf; ++ic) {
    // singles a+_i a_j (src/ConfigList/AllActive/S_ConfigList_AllActive.f90:141-229)
    for (int j = 0; j < nO; ++j)
      for (int i = 0; i < nO; ++i) {
        uint32_t o;
        int sg;
        if (!apply_string(t.occ[ic], &i, &j, 1, o, sg)) continue;
        s_all.push_back({ic, rank[o], i * nO + j, static_cast<float>(sg)});
      }
    t.s_ptr[ic + 1] = static_cast<int32_t>(s_all.size());
    // doubles a+_i1 a+_i2 a_j2 a_j1, i1 <= i2, j1 <= j2 (:250-349)
    for (int j2 = 0; j2 < nO; ++j2)
      for (int j1 = 0; j1 <= j2; ++j1)
        for (int i2 = 0; i2 < nO; ++i2)
          for (int i1 = 0; i1 <= i2; ++i1) {
            int cr[2] = {i1, i2}, de[2] = {j1, j2};
            uint32_t o;
            int sg;
            if (!apply_string(t.occ[ic], cr, de, 2, o, sg)) continue;
            d_all.push_back({ic, rank[o], ((i1 * nO + j1) * nO + i2) * nO + j2, static_cast<float>(sg)});
          }
    t.d_ptr[ic + 1] = static_cast<int32_t>(d_all.size());
  }
  t.singles = s_all;
  t.doubles = d_all;
  csr_by(s_all, nO * nO, [](const CiEntry& e) { return e.code; }, t.r1_ptr, t.r1);
  csr_by(d_all, nO * nO * nO * nO, [](const CiEntry& e) { return e.code; }, t.r2_ptr, t.r2);
  return true;
}

}  // namespace cody
