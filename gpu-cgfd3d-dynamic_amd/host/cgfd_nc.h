/*
 * cgfd_nc.h -- writer (and a small reader) of NetCDF classic files; the writer
 * produces the 64-bit offset variant, the
 * subset the output files of the time loop use: named dimensions (one of
 * them unlimited), float variables, integer global attributes, whole
 * fixed-size variables and one record of a record variable at a time.
 *
 * Format: the published NetCDF classic file format specification (header
 * "CDF\002", big-endian, 8-byte variable offsets; fixed-size variables laid
 * out after the header in definition order, then the records, each record
 * holding one slab of every record variable in definition order). The
 * reference writes the same files through libnetcdf (io_funcs.cu:900-1131,
 * 1827-1879, 1952-2139); this writer needs no library.
 */
#ifndef CGFD_NC_H
#define CGFD_NC_H

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace cgfd {

class NcWriter {
 public:
  ~NcWriter();
  /* returns 0, or a negative CGFD_ERR_* code with cgfd_last_error() set */
  int create(const std::string &path);
  int def_dim(const std::string &name, long len);  /* len 0: unlimited; returns the id */
  int def_var(const std::string &name, const std::vector<int> &dims);  /* NC_FLOAT; returns the id */
  void att_int(const std::string &name, const std::vector<int> &vals);
  int enddef();
  int put_var(int var, const float *data);              /* a fixed-size variable, whole */
  int put_rec(int var, long rec, const float *data);    /* one record of a record variable */
  int close();

 private:
  struct Dim { std::string name; long len; };
  struct Var { std::string name; std::vector<int> dims; bool rec = false; int64_t vsize = 0, begin = 0; };
  struct Att { std::string name; std::vector<int> vals; };
  std::FILE *fp_ = nullptr;
  std::string path_;
  std::vector<Dim> dims_;
  std::vector<Var> vars_;
  std::vector<Att> gatts_;
  int64_t recsize_ = 0;
  long numrecs_ = 0;
  bool defined_ = false;
  int write_at(int64_t off, const void *p, size_t n);
};

/* reader of NetCDF classic files (CDF-1 and the 64-bit offset CDF-2): the
   header, and whole fixed-size NC_FLOAT / NC_DOUBLE variables as float */
class NcReader {
 public:
  ~NcReader();
  int open(const std::string &path);  /* 0 or a CGFD_ERR_* code */
  /* the variable's dimension lengths, or an error if it is absent */
  int shape(const std::string &name, std::vector<long> &dims) const;
  int get_float(const std::string &name, std::vector<float> &out);
  /* the leading block count[0] x count[1] x .. of a fixed-size variable whose
     dimensions are at least that large (nc_get_vara with start 0) */
  int get_block(const std::string &name, const std::vector<long> &count, std::vector<float> &out);

 private:
  struct Var { std::string name; std::vector<int> dims; uint32_t type = 0; int64_t begin = 0; };
  std::FILE *fp_ = nullptr;
  std::string path_;
  std::vector<long> dimlen_;
  std::vector<Var> vars_;
};

}  // namespace cgfd

#endif
