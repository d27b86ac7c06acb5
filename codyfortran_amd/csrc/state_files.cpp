// CodyFortran's restart files: headerless streams of native-endian complex128 values (real part,
// then imaginary part), the bytes a Fortran `access='stream'` unit holds for a complex(R64) array
// (src/Utils/M_Utils_DataStorage.f90:20-65 reads, :67-105 writes).
//
//   coeffs.in          the n_C CI coefficients              (src/Coeffs/S_Coeffs.f90:167-174)
//   orbBB_II.in        orbital II of body type BB, N_g values (src/Orbs/S_Orbs.f90:251-275)
//   rdm2BtXX_YY.in     the (XX, YY) body-type block of rho2  (src/Coeffs/S_Coeffs.f90:185-214)
//
// BB, II, XX and YY are 1-based and written with two digits (Fortran I2.2).  Writing replaces
// an existing file; reading needs the file to exist and to hold at least the requested bytes
// (longer files are accepted, as by the reference's LoadData).  Only host memory is touched:
// device-resident states go through cody_download_state / cody_upload_state.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cody_mctdhx.h"

namespace cody {
int set_error(int code, const char* msg);
}

namespace {

int io_fail(int code, const std::string& msg) { return cody::set_error(code, msg.c_str()); }

std::string in_dir(const char* dir, const char* name) {
  if (!dir || !*dir) return name;
  std::string p(dir);
  if (p.back() != '/') p += '/';
  return p + name;
}

int write_raw(const std::string& path, const void* data, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return io_fail(CODY_ERR_IO, "cannot open " + path + " for writing: " + std::strerror(errno));
  const size_t n = bytes ? std::fwrite(data, 1, bytes, f) : 0;
  const int closed = std::fclose(f);
  if (n != bytes || closed != 0) return io_fail(CODY_ERR_IO, "short write to " + path);
  return CODY_OK;
}

int read_raw(const std::string& path, void* data, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return io_fail(CODY_ERR_IO, "cannot open " + path + ": " + std::strerror(errno));
  const size_t n = bytes ? std::fread(data, 1, bytes, f) : 0;
  std::fclose(f);
  if (n != bytes) return io_fail(CODY_ERR_IO, "more data requested than stored in " + path);
  return CODY_OK;
}

int check_layout(int nC, int nG, int nBodyTypes, const int* nOrbs, const char* who) {
  if (nC < 0 || nG < 1 || nBodyTypes < 1 || nBodyTypes > 99 || !nOrbs)
    return io_fail(CODY_ERR_ARG, std::string(who) + ": bad sizes");
  for (int b = 0; b < nBodyTypes; ++b)
    if (nOrbs[b] < 1 || nOrbs[b] > 99) return io_fail(CODY_ERR_ARG, std::string(who) + ": 1 <= nOrbs <= 99");
  return CODY_OK;
}

std::string orb_name(int bt, int idx) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "orb%02d_%02d.in", bt, idx);
  return buf;
}

}  // namespace

extern "C" {

int cody_save_state_files(const char* dir, int nC, int nG, int nBodyTypes, const int* nOrbs,
                          const double* state) {
  int r = check_layout(nC, nG, nBodyTypes, nOrbs, "cody_save_state_files");
  if (r) return r;
  if (!state) return io_fail(CODY_ERR_ARG, "cody_save_state_files: null state");
  const size_t cplx = 2 * sizeof(double);
  if ((r = write_raw(in_dir(dir, "coeffs.in"), state, cplx * nC))) return r;
  const double* orbs = state + 2 * (size_t)nC;
  size_t k = 0;
  for (int b = 0; b < nBodyTypes; ++b)
    for (int i = 0; i < nOrbs[b]; ++i, ++k)
      if ((r = write_raw(in_dir(dir, orb_name(b + 1, i + 1).c_str()), orbs + 2 * k * nG, cplx * nG)))
        return r;
  return CODY_OK;
}

int cody_load_state_files(const char* dir, int nC, int nG, int nBodyTypes, const int* nOrbs,
                          double* state) {
  int r = check_layout(nC, nG, nBodyTypes, nOrbs, "cody_load_state_files");
  if (r) return r;
  if (!state) return io_fail(CODY_ERR_ARG, "cody_load_state_files: null state");
  const size_t cplx = 2 * sizeof(double);
  if ((r = read_raw(in_dir(dir, "coeffs.in"), state, cplx * nC))) return r;
  double* orbs = state + 2 * (size_t)nC;
  size_t k = 0;
  for (int b = 0; b < nBodyTypes; ++b)
    for (int i = 0; i < nOrbs[b]; ++i, ++k)
      if ((r = read_raw(in_dir(dir, orb_name(b + 1, i + 1).c_str()), orbs + 2 * k * nG, cplx * nG)))
        return r;
  return CODY_OK;
}

int cody_save_two_rdm(const char* dir, int nBodyTypes, const int* nOrbs, const double* rdm2) {
  int r = check_layout(0, 1, nBodyTypes, nOrbs, "cody_save_two_rdm");
  if (r) return r;
  if (!rdm2) return io_fail(CODY_ERR_ARG, "cody_save_two_rdm: null rdm2");
  std::vector<int> start(nBodyTypes + 1, 0);
  for (int b = 0; b < nBodyTypes; ++b) start[b + 1] = start[b] + nOrbs[b];
  const size_t nO = (size_t)start[nBodyTypes];
  // rho2(i1, i2, j1, j2), column-major over the combined orbital index
  auto at = [&](size_t i1, size_t i2, size_t j1, size_t j2) {
    return rdm2 + 2 * (i1 + nO * (i2 + nO * (j1 + nO * j2)));
  };
  for (int b1 = 0; b1 < nBodyTypes; ++b1)
    for (int b2 = 0; b2 < nBodyTypes; ++b2) {
      const size_t n1 = nOrbs[b1], n2 = nOrbs[b2], s1 = start[b1], s2 = start[b2];
      std::vector<double> blk(2 * n1 * n2 * n1 * n2);
      size_t o = 0;
      for (size_t j2 = 0; j2 < n2; ++j2)
        for (size_t j1 = 0; j1 < n1; ++j1)
          for (size_t i2 = 0; i2 < n2; ++i2)
            for (size_t i1 = 0; i1 < n1; ++i1, o += 2) {
              const double* v = at(s1 + i1, s2 + i2, s1 + j1, s2 + j2);
              blk[o] = v[0];
              blk[o + 1] = v[1];
            }
      char name[32];
      std::snprintf(name, sizeof name, "rdm2Bt%02d_%02d.in", b1 + 1, b2 + 1);
      if ((r = write_raw(in_dir(dir, name), blk.data(), sizeof(double) * blk.size()))) return r;
    }
  return CODY_OK;
}

}  // extern "C"
