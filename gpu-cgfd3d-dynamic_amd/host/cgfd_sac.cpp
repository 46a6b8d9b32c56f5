/*
 * cgfd_sac.cpp -- one-component SAC writer; see cgfd_sac.h.
 */
#include "cgfd_sac.h"

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "cgfd_core.h"
#include "cgfd_errors.h"

namespace cgfd {

int write_sac(const std::string &path, const float *data, int npts, float delta, const float ev[4], const float st[4]) {
  float f[70];
  int32_t n[40];
  char k[192];
  for (float &x : f) x = -12345.0f;
  for (int32_t &x : n) x = -12345;
  for (int q = 0; q < 24; q++) std::memcpy(k + 8 * q, "-12345  ", 8);
  std::memset(k + 16, 0, 8);  /* kevnm is 16 characters: "-12345  " and padding */
  f[0] = delta;
  f[5] = 0.0f;                              /* b */
  f[6] = delta * (npts - 1);                /* e */
  for (int q = 0; q < 4; q++) f[31 + q] = st[q];  /* stla, stlo, stel, stdp */
  for (int q = 0; q < 4; q++) f[35 + q] = ev[q];  /* evla, evlo, evel, evdp */
  f[57] = 0.0f;                             /* cmpaz */
  f[58] = 0.0f;                             /* cmpinc */
  n[6] = 6;                                 /* nvhdr */
  n[9] = npts;
  n[15] = 1;                                /* iftype = ITIME */
  n[35] = 1;                                /* leven */
  static const char undef6[8] = {'-', '1', '2', '3', '4', '5', 0, 0};
  std::memcpy(k, undef6, 8);                /* kstnm */
  std::memcpy(k + 8 * 20, undef6, 8);       /* kcmpnm */
  std::FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) return cgfd_set_error(CGFD_ERR_ARG, "cannot create %s", path.c_str());
  bool ok = std::fwrite(f, 4, 70, fp) == 70 && std::fwrite(n, 4, 40, fp) == 40 && std::fwrite(k, 1, 192, fp) == 192 &&
            (npts == 0 || std::fwrite(data, 4, (size_t)npts, fp) == (size_t)npts);
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? CGFD_OK : cgfd_set_error(CGFD_ERR_ARG, "write to %s failed", path.c_str());
}

}  // namespace cgfd
