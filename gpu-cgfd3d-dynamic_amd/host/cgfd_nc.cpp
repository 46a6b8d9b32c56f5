/*
 * cgfd_nc.cpp -- NetCDF classic writer (64-bit offset) and reader; see
 * cgfd_nc.h. Also the reader of the reference's init_stress input.
 */
#include "cgfd_nc.h"

#include <cstring>

#include "cgfd_core.h"
#include "cgfd_errors.h"
#include "cgfd_io.h"

namespace cgfd {

namespace {
constexpr uint32_t kNcDimension = 10, kNcVariable = 11, kNcAttribute = 12;
constexpr uint32_t kNcInt = 4, kNcFloat = 5;

void put_u32(std::vector<unsigned char> &b, uint32_t v) {
  for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
}
void put_u64(std::vector<unsigned char> &b, uint64_t v) {
  for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
}
void put_name(std::vector<unsigned char> &b, const std::string &s) {
  put_u32(b, (uint32_t)s.size());
  b.insert(b.end(), s.begin(), s.end());
  while (b.size() % 4) b.push_back(0);
}
/* big-endian copy of n floats */
void be_floats(const float *src, size_t n, std::vector<unsigned char> &dst) {
  dst.resize(n * 4);
  for (size_t i = 0; i < n; i++) {
    uint32_t u;
    std::memcpy(&u, src + i, 4);
    dst[4 * i] = (unsigned char)(u >> 24);
    dst[4 * i + 1] = (unsigned char)(u >> 16);
    dst[4 * i + 2] = (unsigned char)(u >> 8);
    dst[4 * i + 3] = (unsigned char)u;
  }
}
}  // namespace

NcWriter::~NcWriter() {
  if (fp_) std::fclose(fp_);
}

int NcWriter::create(const std::string &path) {
  fp_ = std::fopen(path.c_str(), "wb+");
  if (!fp_) return cgfd_set_error(CGFD_ERR_ARG, "cannot create %s", path.c_str());
  path_ = path;
  return CGFD_OK;
}

int NcWriter::def_dim(const std::string &name, long len) {
  dims_.push_back({name, len});
  return (int)dims_.size() - 1;
}

int NcWriter::def_var(const std::string &name, const std::vector<int> &dims) {
  Var v;
  v.name = name;
  v.dims = dims;
  v.rec = !dims.empty() && dims_[dims[0]].len == 0;
  int64_t n = 1;
  for (size_t q = v.rec ? 1 : 0; q < dims.size(); q++) n *= dims_[dims[q]].len;
  v.vsize = n * 4;
  vars_.push_back(v);
  return (int)vars_.size() - 1;
}

void NcWriter::att_int(const std::string &name, const std::vector<int> &vals) { gatts_.push_back({name, vals}); }

int NcWriter::enddef() {
  /* header with placeholder offsets, sized first */
  auto header = [&](std::vector<unsigned char> &b) {
    b.clear();
    const char magic[4] = {'C', 'D', 'F', 2};
    b.insert(b.end(), magic, magic + 4);
    put_u32(b, (uint32_t)numrecs_);
    if (dims_.empty()) { put_u32(b, 0); put_u32(b, 0); }
    else {
      put_u32(b, kNcDimension);
      put_u32(b, (uint32_t)dims_.size());
      for (const Dim &d : dims_) { put_name(b, d.name); put_u32(b, (uint32_t)d.len); }
    }
    if (gatts_.empty()) { put_u32(b, 0); put_u32(b, 0); }
    else {
      put_u32(b, kNcAttribute);
      put_u32(b, (uint32_t)gatts_.size());
      for (const Att &a : gatts_) {
        put_name(b, a.name);
        put_u32(b, kNcInt);
        put_u32(b, (uint32_t)a.vals.size());
        for (int v : a.vals) put_u32(b, (uint32_t)v);
      }
    }
    if (vars_.empty()) { put_u32(b, 0); put_u32(b, 0); }
    else {
      put_u32(b, kNcVariable);
      put_u32(b, (uint32_t)vars_.size());
      for (const Var &v : vars_) {
        put_name(b, v.name);
        put_u32(b, (uint32_t)v.dims.size());
        for (int d : v.dims) put_u32(b, (uint32_t)d);
        put_u32(b, 0); put_u32(b, 0);  /* no variable attributes */
        put_u32(b, kNcFloat);
        put_u32(b, (uint32_t)(v.vsize > 0xFFFFFFFFLL ? 0xFFFFFFFFu : (uint32_t)v.vsize));
        put_u64(b, (uint64_t)v.begin);
      }
    }
  };
  std::vector<unsigned char> b;
  header(b);
  int64_t off = (int64_t)b.size();
  for (Var &v : vars_)
    if (!v.rec) { v.begin = off; off += v.vsize; }
  recsize_ = 0;
  for (Var &v : vars_)
    if (v.rec) { v.begin = off + recsize_; recsize_ += v.vsize; }
  header(b);
  defined_ = true;
  return write_at(0, b.data(), b.size());
}

int NcWriter::write_at(int64_t off, const void *p, size_t n) {
  if (!fp_) return cgfd_set_error(CGFD_ERR_STATE, "NetCDF file is not open");
  if (std::fseek(fp_, (long)off, SEEK_SET) != 0 || std::fwrite(p, 1, n, fp_) != n)
    return cgfd_set_error(CGFD_ERR_ARG, "write to %s failed", path_.c_str());
  return CGFD_OK;
}

int NcWriter::put_var(int var, const float *data) {
  if (!defined_ || var < 0 || var >= (int)vars_.size() || vars_[var].rec)
    return cgfd_set_error(CGFD_ERR_ARG, "put_var: bad variable %d", var);
  std::vector<unsigned char> buf;
  be_floats(data, (size_t)vars_[var].vsize / 4, buf);
  return write_at(vars_[var].begin, buf.data(), buf.size());
}

int NcWriter::put_rec(int var, long rec, const float *data) {
  if (!defined_ || var < 0 || var >= (int)vars_.size() || !vars_[var].rec || rec < 0)
    return cgfd_set_error(CGFD_ERR_ARG, "put_rec: bad variable %d or record %ld", var, rec);
  std::vector<unsigned char> buf;
  be_floats(data, (size_t)vars_[var].vsize / 4, buf);
  if (int rc = write_at(vars_[var].begin + (int64_t)rec * recsize_, buf.data(), buf.size())) return rc;
  if (rec + 1 > numrecs_) {
    numrecs_ = rec + 1;
    unsigned char n[4] = {(unsigned char)(numrecs_ >> 24), (unsigned char)(numrecs_ >> 16),
                          (unsigned char)(numrecs_ >> 8), (unsigned char)numrecs_};
    if (int rc = write_at(4, n, 4)) return rc;
  }
  return CGFD_OK;
}

int NcWriter::close() {
  if (!fp_) return CGFD_OK;
  /* a record file ends at its last record */
  const int rc = std::fclose(fp_) == 0 ? CGFD_OK : cgfd_set_error(CGFD_ERR_ARG, "close of %s failed", path_.c_str());
  fp_ = nullptr;
  return rc;
}

/* ---------------------------------------------------------------------- */
namespace {
struct Cursor {
  const std::vector<unsigned char> &b;
  size_t o = 0;
  bool ok = true;
  uint32_t u32() {
    if (o + 4 > b.size()) { ok = false; return 0; }
    uint32_t v = ((uint32_t)b[o] << 24) | ((uint32_t)b[o + 1] << 16) | ((uint32_t)b[o + 2] << 8) | b[o + 3];
    o += 4;
    return v;
  }
  uint64_t u64() { uint64_t hi = u32(); return (hi << 32) | u32(); }
  std::string name() {
    const uint32_t n = u32();
    if (!ok || o + n > b.size()) { ok = false; return std::string(); }
    std::string s((const char *)&b[o], n);
    o += (n + 3) / 4 * 4;
    return s;
  }
  void skip_values(uint32_t type, uint32_t n) {
    const size_t sz = type == 1 || type == 2 ? 1 : (type == 3 ? 2 : (type == 6 ? 8 : 4));
    o += (n * sz + 3) / 4 * 4;
    if (o > b.size()) ok = false;
  }
};
}  // namespace

NcReader::~NcReader() {
  if (fp_) std::fclose(fp_);
}

int NcReader::open(const std::string &path) {
  fp_ = std::fopen(path.c_str(), "rb");
  if (!fp_) return cgfd_set_error(CGFD_ERR_ARG, "cannot open %s", path.c_str());
  path_ = path;
  std::vector<unsigned char> b(1 << 16);
  b.resize(std::fread(b.data(), 1, b.size(), fp_));
  if (b.size() < 8 || b[0] != 'C' || b[1] != 'D' || b[2] != 'F' || (b[3] != 1 && b[3] != 2))
    return cgfd_set_error(CGFD_ERR_ARG, "%s is not a NetCDF classic file", path.c_str());
  const bool off64 = b[3] == 2;
  Cursor c{b};
  c.o = 4;
  c.u32();  /* numrecs */
  /* every count read from the file is bounded by what is left of the header
     buffer, and the loops stop at the first failed read */
  const auto bad = [&]() { return cgfd_set_error(CGFD_ERR_ARG, "%s: header larger than 64 KiB, truncated or malformed",
                                                 path.c_str()); };
  uint32_t tag = c.u32(), n = c.u32();
  if (tag == kNcDimension)
    for (uint32_t q = 0; q < n && c.ok; q++) { c.name(); dimlen_.push_back((long)c.u32()); }
  tag = c.u32(); n = c.u32();
  if (tag == kNcAttribute)
    for (uint32_t q = 0; q < n && c.ok; q++) { c.name(); const uint32_t t = c.u32(), m = c.u32(); c.skip_values(t, m); }
  tag = c.u32(); n = c.u32();
  if (tag == kNcVariable)
    for (uint32_t q = 0; q < n && c.ok; q++) {
      Var v;
      v.name = c.name();
      const uint32_t nd = c.u32();
      if (nd > 64) return bad();
      for (uint32_t d = 0; d < nd && c.ok; d++) {
        const uint32_t id = c.u32();
        if (id >= dimlen_.size()) return bad();  /* a dimension id the header never defined */
        v.dims.push_back((int)id);
      }
      const uint32_t at = c.u32(), na = c.u32();
      if (at == kNcAttribute)
        for (uint32_t a = 0; a < na && c.ok; a++) { c.name(); const uint32_t t = c.u32(), m = c.u32(); c.skip_values(t, m); }
      v.type = c.u32();
      c.u32();  /* vsize */
      v.begin = off64 ? (int64_t)c.u64() : (int64_t)c.u32();
      vars_.push_back(v);
    }
  if (!c.ok) return bad();
  return CGFD_OK;
}

int NcReader::shape(const std::string &name, std::vector<long> &dims) const {
  for (const Var &v : vars_)
    if (v.name == name) {
      dims.clear();
      for (int d : v.dims) dims.push_back(dimlen_[d]);
      return CGFD_OK;
    }
  return cgfd_set_error(CGFD_ERR_ARG, "%s has no variable %s", path_.c_str(), name.c_str());
}

int NcReader::get_block(const std::string &name, const std::vector<long> &count, std::vector<float> &out) {
  for (const Var &v : vars_) {
    if (v.name != name) continue;
    if (count.size() != v.dims.size() || count.empty())
      return cgfd_set_error(CGFD_ERR_ARG, "%s: %s has %zu dimensions, %zu asked", path_.c_str(), name.c_str(),
                            v.dims.size(), count.size());
    for (size_t d = 0; d < count.size(); d++) {
      if (dimlen_[v.dims[d]] == 0)
        return cgfd_set_error(CGFD_ERR_ARG, "%s: %s is a record variable", path_.c_str(), name.c_str());
      if (count[d] < 1 || count[d] > dimlen_[v.dims[d]])
        return cgfd_set_error(CGFD_ERR_ARG, "%s: %s is smaller than the grid along dimension %zu (%ld < %ld)",
                              path_.c_str(), name.c_str(), d, dimlen_[v.dims[d]], count[d]);
    }
    if (v.type != kNcFloat && v.type != 6)
      return cgfd_set_error(CGFD_ERR_ARG, "%s: %s is neither float nor double", path_.c_str(), name.c_str());
    const size_t sz = v.type == kNcFloat ? 4 : 8;
    const size_t nd = count.size();
    const size_t row = (size_t)count[nd - 1], frow = (size_t)dimlen_[v.dims[nd - 1]];
    size_t nrows = 1;
    for (size_t d = 0; d + 1 < nd; d++) nrows *= (size_t)count[d];
    out.resize(nrows * row);
    std::vector<unsigned char> raw(row * sz);
    std::vector<size_t> idx(nd, 0);
    for (size_t r = 0; r < nrows; r++) {
      /* offset of row (idx[0..nd-2], 0) in the file's full dimensions */
      size_t off = 0;
      for (size_t d = 0; d + 1 < nd; d++) off = off * (size_t)dimlen_[v.dims[d]] + idx[d];
      off = off * frow;
      if (std::fseek(fp_, (long)(v.begin + (int64_t)(off * sz)), SEEK_SET) != 0 ||
          std::fread(raw.data(), 1, raw.size(), fp_) != raw.size())
        return cgfd_set_error(CGFD_ERR_ARG, "%s: short read of %s", path_.c_str(), name.c_str());
      for (size_t i = 0; i < row; i++) {
        uint64_t u = 0;
        for (size_t k = 0; k < sz; k++) u = (u << 8) | raw[i * sz + k];
        float *o = &out[r * row + i];
        if (sz == 4) { uint32_t w = (uint32_t)u; std::memcpy(o, &w, 4); }
        else { double dv; std::memcpy(&dv, &u, 8); *o = (float)dv; }
      }
      for (size_t d = nd - 1; d-- > 0;) {  /* next row index */
        if (++idx[d] < (size_t)count[d]) break;
        idx[d] = 0;
      }
    }
    return CGFD_OK;
  }
  return cgfd_set_error(CGFD_ERR_ARG, "%s has no variable %s", path_.c_str(), name.c_str());
}

int NcReader::get_float(const std::string &name, std::vector<float> &out) {
  for (const Var &v : vars_) {
    if (v.name != name) continue;
    size_t n = 1;
    for (int d : v.dims) {
      if (dimlen_[d] == 0) return cgfd_set_error(CGFD_ERR_ARG, "%s: %s is a record variable", path_.c_str(), name.c_str());
      n *= (size_t)dimlen_[d];
    }
    if (v.type != kNcFloat && v.type != 6)
      return cgfd_set_error(CGFD_ERR_ARG, "%s: %s is neither float nor double", path_.c_str(), name.c_str());
    const size_t sz = v.type == kNcFloat ? 4 : 8;
    std::vector<unsigned char> raw(n * sz);
    if (std::fseek(fp_, (long)v.begin, SEEK_SET) != 0 || std::fread(raw.data(), 1, raw.size(), fp_) != raw.size())
      return cgfd_set_error(CGFD_ERR_ARG, "%s: short read of %s", path_.c_str(), name.c_str());
    out.resize(n);
    for (size_t i = 0; i < n; i++) {
      uint64_t u = 0;
      for (size_t k = 0; k < sz; k++) u = (u << 8) | raw[i * sz + k];
      if (sz == 4) { uint32_t w = (uint32_t)u; std::memcpy(&out[i], &w, 4); }
      else { double d; std::memcpy(&d, &u, 8); out[i] = (float)d; }
    }
    return CGFD_OK;
  }
  return cgfd_set_error(CGFD_ERR_ARG, "%s has no variable %s", path_.c_str(), name.c_str());
}

}  // namespace cgfd

namespace {
/* nf float fields of the leading [nk][nj] block of each variable (the
   reference reads that block with nc_get_vara_float from start {0, 0} on
   one process), in order */
int read_planes(const char *path, int nj, int nk, const char *const *names, int nf, float *fields) {
  if (!path || !fields || nj <= 0 || nk <= 0) return cgfd_set_error(CGFD_ERR_ARG, "bad arguments");
  cgfd::NcReader r;
  if (int rc = r.open(path)) return rc;
  std::vector<float> v;
  for (int q = 0; q < nf; q++) {
    if (int rc = r.get_block(names[q], {nk, nj}, v)) return rc;
    std::memcpy(fields + (size_t)q * nj * nk, v.data(), (size_t)nj * nk * 4);
  }
  return CGFD_OK;
}

/* nf float fields of [nk][nj][ni]: the leading block when `block` (coordinate
   import, nc_get_vara_float from start 0), else the variable must be exactly
   that size (metric and media imports read whole variables, nc_get_var) */
int read_volumes(const char *path, int ni, int nj, int nk, const char *const *names, int nf, bool block,
                 float *fields) {
  if (!path || !fields || ni <= 0 || nj <= 0 || nk <= 0) return cgfd_set_error(CGFD_ERR_ARG, "bad arguments");
  cgfd::NcReader r;
  if (int rc = r.open(path)) return rc;
  std::vector<float> v;
  std::vector<long> dims;
  const size_t n = (size_t)ni * nj * nk;
  for (int q = 0; q < nf; q++) {
    if (!block) {
      if (int rc = r.shape(names[q], dims)) return rc;
      if (dims.size() != 3 || dims[0] != nk || dims[1] != nj || dims[2] != ni)
        return cgfd_set_error(CGFD_ERR_ARG, "%s: %s is not [%d][%d][%d] (k, j, i)", path, names[q], nk, nj, ni);
    }
    if (int rc = r.get_block(names[q], {nk, nj, ni}, v)) return rc;
    std::memcpy(fields + q * n, v.data(), n * 4);
  }
  return CGFD_OK;
}
}  // namespace

/* init_stress_<n>.nc (nc_read_init_stress, fault_info.cu:1076-1187) */
extern "C" int cgfd_io_read_init_stress(const char *path, int nj, int nk, float *fields) {
  static const char *names[7] = {"Tx", "Ty", "Tz", "mu_s", "mu_d", "Dc", "C0"};
  return read_planes(path, nj, nk, names, 7, fields);
}

/* fault_coord_<n>.nc (nc_read_fault_geometry, gd_t.cu:548-569) */
extern "C" int cgfd_io_read_fault_coord(const char *path, int nj, int nk, float *xyz) {
  static const char *names[3] = {"x", "y", "z"};
  return read_planes(path, nj, nk, names, 3, xyz);
}

/* coord_<frame>.nc (gd_curv_coord_import, gd_t.cu:855-923) */
extern "C" int cgfd_io_read_coord(const char *path, int ni, int nj, int nk, float *xyz) {
  static const char *names[3] = {"x", "y", "z"};
  return read_volumes(path, ni, nj, nk, names, 3, true, xyz);
}

/* metric_<frame>.nc (gd_curv_metric_import, gd_t.cu:1015-1068; variable
   names of gd_curv_metric_init, gd_t.cu:150-190) */
extern "C" int cgfd_io_read_metric(const char *path, int ni, int nj, int nk, float *metric) {
  static const char *names[10] = {"jac", "xi_x", "xi_y", "xi_z", "eta_x", "eta_y", "eta_z", "zeta_x", "zeta_y", "zeta_z"};
  return read_volumes(path, ni, nj, nk, names, 10, false, metric);
}

/* media_<frame>.nc (md_import, md_t.cu:229-286; isotropic elastic variable
   names of md_init, md_t.cu:80-97) */
extern "C" int cgfd_io_read_media(const char *path, int ni, int nj, int nk, float *media) {
  static const char *names[3] = {"lambda", "mu", "rho"};
  return read_volumes(path, ni, nj, nk, names, 3, false, media);
}
