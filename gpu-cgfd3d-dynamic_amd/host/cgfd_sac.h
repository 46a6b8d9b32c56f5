/*
 * cgfd_sac.h -- one-component SAC file (binary, native byte order), with the
 * header fields the reference's sacExport1C1R sets (src/lib/sacLib.cu:8-93,
 * 96-348): every field undefined (-12345) but delta, b = 0,
 * e = delta (npts - 1), nvhdr = 6, npts, iftype = ITIME, leven = 1, the
 * event and station position fields and cmpaz = cmpinc = 0; kstnm and kcmpnm
 * "-12345". Layout: the published SAC header (70 floats, 40 integers, 24
 * eight-character strings with kevnm taking two), then the samples.
 */
#ifndef CGFD_SAC_H
#define CGFD_SAC_H

#include <string>

namespace cgfd {

/* returns 0 or a negative CGFD_ERR_* code with cgfd_last_error() set */
int write_sac(const std::string &path, const float *data, int npts, float delta, const float ev[4], const float st[4]);

}  // namespace cgfd

#endif
