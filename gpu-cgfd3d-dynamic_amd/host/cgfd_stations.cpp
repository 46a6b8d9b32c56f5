/*
 * cgfd_stations.cpp -- station lists and their location on the grid, host
 * only (see include/cgfd_io.h).
 *
 *   station list   io_recv_read_locate, io_funcs.cu:23-320
 *   fault stations io_fault_recv_read_locate, io_funcs.cu:2170-2300
 *   receiver lines io_line_locate, io_funcs.cu:312-425 (SAC names :771)
 *   locator        gd_curv_coord_to_local_indx, gd_t.cu:1365-1503;
 *                  gd_curv_depth_to_axis, gd_t.cu:1522-1612;
 *                  gd_curv_coord2index_sample, gd_t.cu:1694-1775;
 *                  isPointInHexahedron_c, gd_t.cu:1893-1935;
 *                  fdlib_math_isPoint2InQuad / _rdinterp_2d, fdlib_math.cu:163-240
 */
#include "../../include/cgfd_io.h"
#include "../../include/cgfd_setup.h"
#include "cgfd_errors.h"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace {

/* io_get_nextline (io_funcs.cu:1886-1908): the next line that is neither a
   comment ('#' first) nor empty, without its newline */
bool next_line(FILE *fp, char *buf, int n) {
  do {
    if (!fgets(buf, n, fp)) return false;
  } while (buf[0] == '#' || buf[0] == '\n');
  const size_t len = std::strlen(buf);
  if (len > 0 && buf[len - 1] == '\n') buf[len - 1] = '\0';
  return true;
}

struct FileCloser {
  void operator()(FILE *f) const { std::fclose(f); }
};

/* the extended grid of a setup: coordinates with 3 ghosts per side */
struct Ext {
  int ni, nj, nk, nx, ny, nz;
  long siy, siz, n;
  const float *x, *y, *z;
  float xmin, xmax, ymin, ymax, zmin, zmax;  /* over every point, ghosts too */
  int ni1() const { return 3; }
  int ni2() const { return ni + 2; }
  int nj2() const { return nj + 2; }
  int nk2() const { return nk + 2; }
  long at(int i, int j, int k) const { return i + j * siy + k * siz; }
};

int ext_of(const cgfd_setup *s, Ext &g) {
  cgfd_desc d;
  if (int rc = cgfd_setup_get_desc(s, &d)) return rc;
  g.ni = d.ni; g.nj = d.nj; g.nk = d.nk;
  g.nx = g.ni + 6; g.ny = g.nj + 6; g.nz = g.nk + 6;
  g.siy = g.nx; g.siz = (long)g.nx * g.ny; g.n = g.siz * g.nz;
  g.x = cgfd_setup_coords(s);
  if (!g.x) return cgfd_set_error(CGFD_ERR_ARG, "setup has no coordinates");
  g.y = g.x + g.n; g.z = g.y + g.n;
  /* gd_curv_set_minmax, gd_t.cu:1076-1097 */
  g.xmin = g.xmax = g.x[0]; g.ymin = g.ymax = g.y[0]; g.zmin = g.zmax = g.z[0];
  for (long p = 0; p < g.n; p++) {
    g.xmin = std::fmin(g.xmin, g.x[p]); g.xmax = std::fmax(g.xmax, g.x[p]);
    g.ymin = std::fmin(g.ymin, g.y[p]); g.ymax = std::fmax(g.ymax, g.y[p]);
    g.zmin = std::fmin(g.zmin, g.z[p]); g.zmax = std::fmax(g.zmax, g.z[p]);
  }
  return CGFD_OK;
}

float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void cross3(const float *a, const float *b, float *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* isPointInHexahedron_c: inside (or on) every face plane of the cube whose
   vertices are numbered i + 2 j + 4 k; faces as triangles of three vertices
   (gd_t.cu:1899-1906) */
bool in_hexahedron(float px, float py, float pz, const float *vx, const float *vy, const float *vz) {
  static const int faces[6][3] = {{0, 4, 6}, {7, 5, 1}, {5, 4, 0}, {2, 6, 7}, {4, 5, 7}, {3, 1, 0}};
  const float p[3] = {px, py, pz};
  for (const auto &f : faces) {
    const float a[3] = {vx[f[0]], vy[f[0]], vz[f[0]]};
    const float b[3] = {vx[f[1]], vy[f[1]], vz[f[1]]};
    const float c[3] = {vx[f[2]], vy[f[2]], vz[f[2]]};
    const float p2f[3] = {a[0] - p[0], a[1] - p[1], a[2] - p[2]};
    const float A[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float B[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float nrm[3];
    cross3(A, B, nrm);
    const float len = (float)std::sqrt(dot3(nrm, nrm));
    for (float &v : nrm) v /= len;
    float sign = dot3(p2f, nrm);
    const float lp = (float)std::sqrt(dot3(p2f, p2f));
    sign /= lp;
    if (sign < 0.0f) return false;
  }
  return true;
}

/* gd_curv_coord2index_sample: the sample of a 101^3 trilinear lattice of
   the cube nearest to the point, as fractions of a cell */
void sample_in_cube(float sx, float sy, float sz, const float *px, const float *py, const float *pz, float &fi,
                    float &fj, float &fk) {
  const int ns = 100;
  float best = std::sqrt((sx - px[0]) * (sx - px[0]) + (sy - py[0]) * (sy - py[0]) + (sz - pz[0]) * (sz - pz[0]));
  int bi = 0, bj = 0, bk = 0;
  float Lx[2], Ly[2], Lz[2];
  for (int n3 = 0; n3 <= ns; n3++) {
    Lz[1] = (float)n3 / (float)ns;
    Lz[0] = 1.0f - Lz[1];
    for (int n2 = 0; n2 <= ns; n2++) {
      Ly[1] = (float)n2 / (float)ns;
      Ly[0] = 1.0f - Ly[1];
      for (int n1 = 0; n1 <= ns; n1++) {
        Lx[1] = (float)n1 / (float)ns;
        Lx[0] = 1.0f - Lx[1];
        float x = 0.0f, y = 0.0f, z = 0.0f;
        for (int kk = 0; kk < 2; kk++)
          for (int jj = 0; jj < 2; jj++)
            for (int ii = 0; ii < 2; ii++) {
              const int c = ii + 2 * jj + 4 * kk;
              const float w = Lx[ii] * Ly[jj] * Lz[kk];
              x += w * px[c];
              y += w * py[c];
              z += w * pz[c];
            }
        const float d = std::sqrt((sx - x) * (sx - x) + (sy - y) * (sy - y) + (sz - z) * (sz - z));
        if (best > d) { best = d; bi = n1; bj = n2; bk = n3; }
      }
    }
  }
  fi = (float)bi / (float)ns;
  fj = (float)bj / (float)ns;
  fk = (float)bk / (float)ns;
}

/* nearest grid point (every point, ghosts too; the first in k, j, i order
   among equal distances) for every station at once */
void nearest_points(const Ext &g, const std::vector<std::array<float, 3>> &p, std::vector<long> &best) {
  const size_t ns = p.size();
  best.assign(ns, 0);
  std::vector<float> bd(ns, std::numeric_limits<float>::infinity());
#pragma omp parallel
  {
    std::vector<float> d(ns, std::numeric_limits<float>::infinity());
    std::vector<long> b(ns, -1);
#pragma omp for schedule(static)
    for (int k = 0; k < g.nz; k++)
      for (long q = (long)k * g.siz; q < (long)(k + 1) * g.siz; q++) {
        const float x = g.x[q], y = g.y[q], z = g.z[q];
        for (size_t s = 0; s < ns; s++) {
          const float dist = std::sqrt((p[s][0] - x) * (p[s][0] - x) + (p[s][1] - y) * (p[s][1] - y) +
                                       (p[s][2] - z) * (p[s][2] - z));
          if (dist < d[s]) { d[s] = dist; b[s] = q; }
        }
      }
#pragma omp critical
    for (size_t s = 0; s < ns; s++)
      if (b[s] >= 0 && (d[s] < bd[s] || (d[s] == bd[s] && b[s] < best[s]))) { bd[s] = d[s]; best[s] = b[s]; }
  }
}

/* gd_curv_coord_to_local_indx after the nearest point: the cell among the
   eight around it that contains the point, and the sampled position in it.
   Returns false when the nearest point is not physical (the station is
   not on this grid). */
bool cell_of(const Ext &g, float sx, float sy, float sz, long nearest, int &si, int &sj, int &sk, float &di,
             float &dj, float &dk) {
  const int mk = (int)(nearest / g.siz), mj = (int)((nearest % g.siz) / g.siy), mi = (int)(nearest % g.siy);
  if (mi < 3 || mi > g.ni2() || mj < 3 || mj > g.nj2() || mk < 3 || mk > g.nk2()) return false;
  si = mi; sj = mj; sk = mk;
  di = dj = dk = 0.0f;
  for (int kk = 0; kk < 2; kk++)
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++) {
        const int ci = mi - 1 + ii, cj = mj - 1 + jj, ck = mk - 1 + kk;
        float px[8], py[8], pz[8];
        for (int n3 = 0; n3 < 2; n3++)
          for (int n2 = 0; n2 < 2; n2++)
            for (int n1 = 0; n1 < 2; n1++) {
              const long q = g.at(ci + n1, cj + n2, ck + n3);
              const int c = n1 + 2 * n2 + 4 * n3;
              px[c] = g.x[q]; py[c] = g.y[q]; pz[c] = g.z[q];
            }
        if (in_hexahedron(sx, sy, sz, px, py, pz)) {
          float fi, fj, fk;
          sample_in_cube(sx, sy, sz, px, py, pz, fi, fj, fk);
          di = ((float)ci + fi) - (float)mi;
          dj = ((float)cj + fj) - (float)mj;
          dk = ((float)ck + fk) - (float)mk;
          return true;
        }
      }
  return true;  /* in no cell: the nearest point itself */
}

/* fdlib_math_isPoint2InQuad: vertices i + 2 j, taken anticlockwise */
bool in_quad(float px, float py, const float *vx0, const float *vy0) {
  const float vx[4] = {vx0[0], vx0[1], vx0[3], vx0[2]}, vy[4] = {vy0[0], vy0[1], vy0[3], vy0[2]};
  for (int i = 0; i < 4; i++) {
    const float ex = i == 3 ? vx[0] - vx[3] : vx[i + 1] - vx[i];
    const float ey = i == 3 ? vy[0] - vy[3] : vy[i + 1] - vy[i];
    float sign = ex * (py - vy[i]) - ey * (px - vx[i]);
    sign /= (float)std::sqrt(ex * ex + ey * ey);
    if (sign < -1e-6f) return false;
  }
  return true;
}

/* fdlib_math_rdinterp_2d: inverse-distance weights of the 4 corners; a
   point within 1e-9 of a corner takes that corner's value */
float idw4(float x, float y, const float *px, const float *py, const float *pv) {
  float w[4], tot = 0.0f;
  int at = -1;
  for (int i = 0; i < 4; i++) {
    const float d = std::sqrt((x - px[i]) * (x - px[i]) + (y - py[i]) * (y - py[i]));
    if (d < 1e-9f) at = i;
    else { w[i] = 1.0f / d; tot += w[i]; }
  }
  if (at >= 0) {
    for (float &v : w) v = 0.0f;
    w[at] = 1.0f;
    tot = 1.0f;
  }
  float v = 0.0f;
  for (int i = 0; i < 4; i++) {
    w[i] *= 1.0f / tot;
    v += w[i] * pv[i];
  }
  return v;
}

/* gd_curv_depth_to_axis: depth below the free surface -> z, from the first
   top-layer cell (in the reference's 4 x 4 tile order) containing (sx, sy);
   z unchanged when none does */
void depth_to_z(const Ext &g, float sx, float sy, float &sz) {
  if (sx < g.xmin || sx > g.xmax || sy < g.ymin || sy > g.ymax) return;
  const int T = 4, k = g.nk2();
  int is[4], ie[4], js[4], je[4];
  for (int t = 0; t < T; t++) {
    is[t] = t == 0 ? 3 : ie[t - 1] + 1;
    ie[t] = is[t] + g.ni / T - 1 + (t < g.ni % T ? 1 : 0);
    js[t] = t == 0 ? 3 : je[t - 1] + 1;
    je[t] = js[t] + g.nj / T - 1 + (t < g.nj % T ? 1 : 0);
  }
  for (int tj = 0; tj < T; tj++)
    for (int ti = 0; ti < T; ti++)
      for (int j = js[tj]; j <= je[tj]; j++)
        for (int i = is[ti]; i <= ie[ti]; i++) {
          /* the cell's bounding box over its 8 corners (gd_t.cu:1127-1150) */
          float xmn = g.x[g.at(i, j, k)], xmx = xmn, ymn = g.y[g.at(i, j, k)], ymx = ymn;
          for (int n3 = 0; n3 < 2; n3++)
            for (int n2 = 0; n2 < 2; n2++)
              for (int n1 = 0; n1 < 2; n1++) {
                const long q = g.at(i + n1, j + n2, k + n3);
                xmn = std::fmin(xmn, g.x[q]); xmx = std::fmax(xmx, g.x[q]);
                ymn = std::fmin(ymn, g.y[q]); ymx = std::fmax(ymx, g.y[q]);
              }
          if (sx < xmn || sx > xmx || sy < ymn || sy > ymx) continue;
          float px[4], py[4], pz[4];
          for (int n2 = 0; n2 < 2; n2++)
            for (int n1 = 0; n1 < 2; n1++) {
              const long q = g.at(i + n1, j + n2, k);
              px[n1 + 2 * n2] = g.x[q]; py[n1 + 2 * n2] = g.y[q]; pz[n1 + 2 * n2] = g.z[q];
            }
          if (in_quad(sx, sy, px, py)) {
            sz = idw4(sx, sy, px, py, pz) - sz;
            return;
          }
        }
}

int copy_name(char *dst, const char *src, const char *path, int line) {
  if (std::strlen(src) >= CGFD_STATION_NAME_MAX)
    return cgfd_set_error(CGFD_ERR_ARG, "%s: station %d: name longer than %d characters", path, line,
                          CGFD_STATION_NAME_MAX - 1);
  std::strcpy(dst, src);
  return CGFD_OK;
}

}  // namespace

extern "C" int cgfd_io_read_stations(const char *path, const cgfd_setup *setup, cgfd_station *out, int max_n,
                                     int *n_file, int *n_located) {
  if (!path || !setup || !n_file || !n_located || max_n < 0)
    return cgfd_set_error(CGFD_ERR_ARG, "cgfd_io_read_stations: bad arguments");
  *n_file = *n_located = 0;
  std::unique_ptr<FILE, FileCloser> fp(std::fopen(path, "rt"));
  if (!fp) return cgfd_set_error(CGFD_ERR_ARG, "cannot open station file %s", path);
  char line[500];
  int nrec = 0;
  if (!next_line(fp.get(), line, sizeof line) || std::sscanf(line, "%d", &nrec) != 1 || nrec < 0)
    return cgfd_set_error(CGFD_ERR_ARG, "%s: no station count", path);
  *n_file = nrec;
  if (!out) return CGFD_OK;
  if (max_n < nrec) return cgfd_set_error(CGFD_ERR_ARG, "%s: %d stations, room for %d", path, nrec, max_n);
  Ext g;
  if (int rc = ext_of(setup, g)) return rc;
  struct Rec {
    char name[CGFD_STATION_NAME_MAX];
    int flag_coord, flag_depth;
    float c[3];
  };
  std::vector<Rec> rec(nrec);
  for (int r = 0; r < nrec; r++) {
    char nm[600];
    if (!next_line(fp.get(), line, sizeof line) ||
        std::sscanf(line, "%599s %d %d %g %g %g", nm, &rec[r].flag_coord, &rec[r].flag_depth, &rec[r].c[0],
                    &rec[r].c[1], &rec[r].c[2]) != 6)
      return cgfd_set_error(CGFD_ERR_ARG, "%s: station %d: expected 'name flag_coord flag_depth x y z'", path, r + 1);
    if (int rc = copy_name(rec[r].name, nm, path, r + 1)) return rc;
  }
  /* coordinate stations: depth to z, then the nearest grid point of all */
  std::vector<int> coord_ids;
  std::vector<std::array<float, 3>> pts;
  for (int r = 0; r < nrec; r++)
    if (rec[r].flag_coord == 1) {
      if (rec[r].flag_depth == 1) depth_to_z(g, rec[r].c[0], rec[r].c[1], rec[r].c[2]);
      coord_ids.push_back(r);
      pts.push_back({rec[r].c[0], rec[r].c[1], rec[r].c[2]});
    }
  std::vector<long> nearest;
  if (!pts.empty()) nearest_points(g, pts, nearest);
  int located = 0;
  size_t ci = 0;
  for (int r = 0; r < nrec; r++) {
    const Rec &q = rec[r];
    int gi, gj, gk;  /* global physical indices */
    float di, dj, dk;
    float x = q.c[0], y = q.c[1], z = q.c[2];
    if (q.flag_coord == 1) {
      const size_t c = ci++;
      int si, sj, sk;
      const bool inside = x >= g.xmin && x <= g.xmax && y >= g.ymin && y <= g.ymax && z >= g.zmin && z <= g.zmax;
      if (!inside || !cell_of(g, x, y, z, nearest[c], si, sj, sk, di, dj, dk)) continue;  /* outside: dropped */
      gi = si - 3; gj = sj - 3; gk = sk - 3;
      /* a negative shift moves to the previous cell (io_funcs.cu:161-181) */
      if (di < 0.0f) { di = 1.0f + di; gi -= 1; }
      if (dj < 0.0f) { dj = 1.0f + dj; gj -= 1; }
      if (dk < 0.0f) { dk = 1.0f + dk; gk -= 1; }
    } else {
      /* grid indices from 1; depth counted down from the surface from 0
         (io_funcs.cu:184-213) */
      float cx = q.c[0] - 1, cy = q.c[1] - 1, cz = q.flag_depth == 1 ? (float)(g.nk - 1) - q.c[2] : q.c[2] - 1;
      gi = (int)cx; gj = (int)cy; gk = (int)cz;
      di = cx - gi; dj = cy - gj; dk = cz - gk;
      if (gi < 0 || gi > g.ni || gj < 0 || gj > g.nj || gk < 0 || gk > g.nk) continue;
    }
    if (gi < 0 || gi > g.ni - 1 || gj < 0 || gj > g.nj - 1 || gk < 0 || gk > g.nk - 1) continue;
    cgfd_station &s = out[located++];
    std::strcpy(s.name, q.name);
    s.i = gi + 3; s.j = gj + 3; s.k = gk + 3;
    s.di = di; s.dj = dj; s.dk = dk;
    if (q.flag_coord != 1) {
      const long p = g.at(s.i, s.j, s.k);
      x = g.x[p]; y = g.y[p]; z = g.z[p];
    }
    s.x = x; s.y = y; s.z = z;
  }
  *n_located = located;
  return CGFD_OK;
}

extern "C" int cgfd_io_read_fault_stations(const char *path, const cgfd_setup *setup, cgfd_fault_station *out,
                                           int max_n, int *n_file) {
  if (!path || !setup || !n_file || max_n < 0)
    return cgfd_set_error(CGFD_ERR_ARG, "cgfd_io_read_fault_stations: bad arguments");
  *n_file = 0;
  std::unique_ptr<FILE, FileCloser> fp(std::fopen(path, "rt"));
  if (!fp) return cgfd_set_error(CGFD_ERR_ARG, "cannot open fault station file %s", path);
  char line[500];
  int nrec = 0;
  if (!next_line(fp.get(), line, sizeof line) || std::sscanf(line, "%d", &nrec) != 1 || nrec < 0)
    return cgfd_set_error(CGFD_ERR_ARG, "%s: no station count", path);
  *n_file = nrec;
  if (!out) return CGFD_OK;
  if (max_n < nrec) return cgfd_set_error(CGFD_ERR_ARG, "%s: %d stations, room for %d", path, nrec, max_n);
  cgfd_desc d;
  if (int rc = cgfd_setup_get_desc(setup, &d)) return rc;
  Ext g;
  if (int rc = ext_of(setup, g)) return rc;
  for (int r = 0; r < nrec; r++) {
    char nm[600];
    int fid;
    float ry, rz;
    if (!next_line(fp.get(), line, sizeof line) || std::sscanf(line, "%599s %d %g %g", nm, &fid, &ry, &rz) != 4)
      return cgfd_set_error(CGFD_ERR_ARG, "%s: fault station %d: expected 'name fault_id j k'", path, r + 1);
    cgfd_fault_station &s = out[r];
    if (int rc = copy_name(s.name, nm, path, r + 1)) return rc;
    fid -= 1;  /* from 1 */
    if (fid < 0 || fid >= d.number_fault)
      return cgfd_set_error(CGFD_ERR_ARG, "%s: fault station %d: fault %d of %d", path, r + 1, fid + 1,
                            d.number_fault);
    ry -= 1;
    rz -= 1;
    const int iy = (int)std::floor(ry), iz = (int)std::floor(rz);
    /* io_funcs.cu:2283-2293: outside the fault plane is an error */
    if (iy < 0 || iy > g.nj - 1 || iz < 0 || iz > g.nk - 1)
      return cgfd_set_error(CGFD_ERR_ARG, "%s: fault station %d (%s) is outside the calculation area", path, r + 1,
                            s.name);
    s.fault_id = fid;
    s.j = iy + 3;
    s.k = iz + 3;
    s.dj = ry - iy;
    s.dk = rz - iz;
    const long p = g.at(d.fault[fid].fault_x_index + 3, s.j, s.k);
    s.x = g.x[p]; s.y = g.y[p]; s.z = g.z[p];
  }
  return CGFD_OK;
}

extern "C" int cgfd_io_line_stations(const cgfd_setup *setup, const char *name, const int *start, const int *incre,
                                     int count, cgfd_station *out, int *n_out) {
  if (!setup || !name || !start || !incre || !n_out || count < 0 || (count > 0 && !out))
    return cgfd_set_error(CGFD_ERR_ARG, "cgfd_io_line_stations: bad arguments");
  *n_out = 0;
  Ext g;
  if (int rc = ext_of(setup, g)) return rc;
  int n = 0;
  for (int ipt = 0; ipt < count; ipt++) {
    const int gi = start[0] + ipt * incre[0], gj = start[1] + ipt * incre[1], gk = start[2] + ipt * incre[2];
    if (gi < 0 || gi > g.ni - 1 || gj < 0 || gj > g.nj - 1 || gk < 0 || gk > g.nk - 1) continue;
    cgfd_station &s = out[n++];
    const int w = std::snprintf(s.name, CGFD_STATION_NAME_MAX, "%s.no%d", name, ipt);
    if (w < 0 || w >= CGFD_STATION_NAME_MAX)
      return cgfd_set_error(CGFD_ERR_ARG, "receiver line %s: name longer than %d characters", name,
                            CGFD_STATION_NAME_MAX - 1);
    s.i = gi + 3; s.j = gj + 3; s.k = gk + 3;
    s.di = s.dj = s.dk = 0.0f;
    const long p = g.at(s.i, s.j, s.k);
    s.x = g.x[p]; s.y = g.y[p]; s.z = g.z[p];
  }
  *n_out = n;
  return CGFD_OK;
}
