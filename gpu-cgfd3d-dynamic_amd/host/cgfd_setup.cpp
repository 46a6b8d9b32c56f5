/*
 * cgfd_setup.cpp -- host setup of the C++ driver (see include/cgfd_setup.h).
 *
 * Every routine restates a host routine of the reference that runs before the
 * time loop; the comment on each names the reference lines.
 */
#include "../../include/cgfd_setup.h"
#include "cgfd_errors.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include <memory>

namespace {

/* 3x3 helpers with the reference's arithmetic (src/lib/fdlib_math.cu) */
inline float norm3f(const float *A) {
  float x2 = A[0] * A[0], y2 = A[1] * A[1], z2 = A[2] * A[2];
  return sqrtf(x2 + y2 + z2);
}
inline void invert3x3f(float m[][3]) {
  double inv[3][3];
  inv[0][0] = (float)(m[1][1] * m[2][2] - m[2][1] * m[1][2]);
  inv[0][1] = (float)(m[2][1] * m[0][2] - m[0][1] * m[2][2]);
  inv[0][2] = (float)(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
  inv[1][0] = (float)(m[1][2] * m[2][0] - m[1][0] * m[2][2]);
  inv[1][1] = (float)(m[0][0] * m[2][2] - m[2][0] * m[0][2]);
  inv[1][2] = (float)(m[1][0] * m[0][2] - m[0][0] * m[1][2]);
  inv[2][0] = (float)(m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  inv[2][1] = (float)(m[2][0] * m[0][1] - m[0][0] * m[2][1]);
  inv[2][2] = (float)(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
  double det = inv[0][0] * m[0][0] + inv[0][1] * m[1][0] + inv[0][2] * m[2][0];
  det = 1.0 / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = (float)(inv[i][j] * det);
}
inline void matmul3x3f(float A[][3], float B[][3], float C[][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C[i][j] = 0.0f;
      for (int k = 0; k < 3; k++) C[i][j] += A[i][k] * B[k][j];
    }
}
inline void cross3f(const float *A, const float *B, float *C) {
  C[0] = A[1] * B[2] - A[2] * B[1];
  C[1] = A[2] * B[0] - A[0] * B[2];
  C[2] = A[0] * B[1] - A[1] * B[0];
}
inline float dot3f(const float *A, const float *B) {
  float r = 0.0f;
  for (int i = 0; i < 3; i++) r += A[i] * B[i];
  return r;
}

/* 6th-order centred difference, fd_t.h:126-140 (d_4 = 0 is skipped there) */
inline float center6(const float *p, long s) {
  const double d1 = -0.01666667, d2 = 0.15, d3 = -0.75, d5 = 0.75, d6 = -0.15, d7 = 0.01666667;
  return (float)((d1 * p[-3 * s]) + (d2 * p[-2 * s]) + (d3 * p[-s]) + (d5 * p[s]) + (d6 * p[2 * s]) + (d7 * p[3 * s]));
}

/* Elastic traction-strain matrix D_ab of metric rows a, b at a node.
   Entry expressions and their evaluation order are those of
   fault_info.cu:190-369 (all 18 blocks share this pattern). */
inline void dmat(const float *ea, const float *eb, float lam, float mu, float lam2mu, float D[9]) {
  D[0] = lam2mu * ea[0] * eb[0] + mu * (ea[1] * eb[1] + ea[2] * eb[2]);
  D[1] = lam * ea[0] * eb[1] + mu * ea[1] * eb[0];
  D[2] = lam * ea[0] * eb[2] + mu * ea[2] * eb[0];
  D[3] = mu * ea[0] * eb[1] + lam * ea[1] * eb[0];
  D[4] = lam2mu * ea[1] * eb[1] + mu * (ea[0] * eb[0] + ea[2] * eb[2]);
  D[5] = lam * ea[1] * eb[2] + mu * ea[2] * eb[1];
  D[6] = mu * ea[0] * eb[2] + lam * ea[2] * eb[0];
  D[7] = mu * ea[1] * eb[2] + lam * ea[2] * eb[1];
  D[8] = lam2mu * ea[2] * eb[2] + mu * (ea[0] * eb[0] + ea[1] * eb[1]);
}

inline void to33(const float *v, float M[][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) M[i][j] = v[3 * i + j];
}
inline void from33(float M[][3], float *v) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) v[3 * i + j] = M[i][j];
}

struct Grid {
  int ni, nj, nk, nx, ny, nz, ni1, ni2, nj1, nj2, nk1, nk2;
  size_t siy, siz, sicmp, sslice;
  void set(int a, int b, int c) {
    ni = a; nj = b; nk = c;
    nx = ni + 6; ny = nj + 6; nz = nk + 6;
    ni1 = nj1 = nk1 = 3;
    ni2 = ni1 + ni - 1; nj2 = nj1 + nj - 1; nk2 = nk1 + nk - 1;
    siy = nx; siz = (size_t)nx * ny; sicmp = siz * nz; sslice = (size_t)ny * nz;
  }
};

/* geometric_symmetry, gd_t.cu:665-759: linear extrapolation into the ghosts */
void geometric_symmetry(const Grid &g, float *v, int ncmp) {
  for (int c = 0; c < ncmp; c++) {
    float *b = v + c * g.sicmp;
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.ni1; i++) {
          size_t r = j * g.siy + k * g.siz;
          b[i + r] = 2 * b[g.ni1 + r] - b[(2 * g.ni1 - i) + r];
        }
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = g.ni2 + 1; i < g.nx; i++) {
          size_t r = j * g.siy + k * g.siz;
          b[i + r] = 2 * b[g.ni2 + r] - b[(2 * g.ni2 - i) + r];
        }
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.nj1; j++)
        for (int i = 0; i < g.nx; i++) {
          size_t r = i + k * g.siz;
          b[r + j * g.siy] = 2 * b[r + g.nj1 * g.siy] - b[r + (2 * g.nj1 - j) * g.siy];
        }
    for (int k = 0; k < g.nz; k++)
      for (int j = g.nj2 + 1; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
          size_t r = i + k * g.siz;
          b[r + j * g.siy] = 2 * b[r + g.nj2 * g.siy] - b[r + (2 * g.nj2 - j) * g.siy];
        }
    for (int k = 0; k < g.nk1; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
          size_t r = i + j * g.siy;
          b[r + k * g.siz] = 2 * b[r + g.nk1 * g.siz] - b[r + (2 * g.nk1 - k) * g.siz];
        }
    for (int k = g.nk2 + 1; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
          size_t r = i + j * g.siy;
          b[r + k * g.siz] = 2 * b[r + g.nk2 * g.siz] - b[r + (2 * g.nk2 - k) * g.siz];
        }
  }
}

/* mirror_symmetry, gd_t.cu:578-655: even reflection into the ghosts */
void mirror_symmetry(const Grid &g, float *v, int ncmp) {
  for (int c = 0; c < ncmp; c++) {
    float *b = v + c * g.sicmp;
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++) {
        size_t r = j * g.siy + k * g.siz;
        for (int i = 0; i < g.ni1; i++) b[i + r] = b[(2 * g.ni1 - i) + r];
        for (int i = g.ni2 + 1; i < g.nx; i++) b[i + r] = b[(2 * g.ni2 - i) + r];
      }
    for (int k = 0; k < g.nz; k++)
      for (int i = 0; i < g.nx; i++) {
        size_t r = i + k * g.siz;
        for (int j = 0; j < g.nj1; j++) b[r + j * g.siy] = b[r + (2 * g.nj1 - j) * g.siy];
        for (int j = g.nj2 + 1; j < g.ny; j++) b[r + j * g.siy] = b[r + (2 * g.nj2 - j) * g.siy];
      }
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        size_t r = i + j * g.siy;
        for (int k = 0; k < g.nk1; k++) b[r + k * g.siz] = b[r + (2 * g.nk1 - k) * g.siz];
        for (int k = g.nk2 + 1; k < g.nz; k++) b[r + k * g.siz] = b[r + (2 * g.nk2 - k) * g.siz];
      }
  }
}

/* gd_curv_metric_cal, gd_t.cu:209-304 */
void metric_cal(const Grid &g, const float *x3d, const float *y3d, const float *z3d, float *met) {
  float *jac3d = met, *xi_x = met + g.sicmp, *xi_y = xi_x + g.sicmp, *xi_z = xi_y + g.sicmp;
  float *et_x = xi_z + g.sicmp, *et_y = et_x + g.sicmp, *et_z = et_y + g.sicmp;
  float *zt_x = et_z + g.sicmp, *zt_y = zt_x + g.sicmp, *zt_z = zt_y + g.sicmp;
#pragma omp parallel for schedule(static)
  for (int k = g.nk1; k <= g.nk2; k++)
    for (int j = g.nj1; j <= g.nj2; j++)
      for (int i = g.ni1; i <= g.ni2; i++) {
        size_t iptr = i + j * g.siy + k * g.siz;
        float vec1[3], vec2[3], vec3[3], vecg[3];
        vec1[0] = center6(x3d + iptr, 1); vec1[1] = center6(y3d + iptr, 1); vec1[2] = center6(z3d + iptr, 1);
        vec2[0] = center6(x3d + iptr, g.siy); vec2[1] = center6(y3d + iptr, g.siy); vec2[2] = center6(z3d + iptr, g.siy);
        vec3[0] = center6(x3d + iptr, g.siz); vec3[1] = center6(y3d + iptr, g.siz); vec3[2] = center6(z3d + iptr, g.siz);
        cross3f(vec1, vec2, vecg);
        float jac = dot3f(vecg, vec3);
        jac3d[iptr] = jac;
        cross3f(vec2, vec3, vecg);
        xi_x[iptr] = vecg[0] / jac; xi_y[iptr] = vecg[1] / jac; xi_z[iptr] = vecg[2] / jac;
        cross3f(vec3, vec1, vecg);
        et_x[iptr] = vecg[0] / jac; et_y[iptr] = vecg[1] / jac; et_z[iptr] = vecg[2] / jac;
        cross3f(vec1, vec2, vecg);
        zt_x[iptr] = vecg[0] / jac; zt_y[iptr] = vecg[1] / jac; zt_z[iptr] = vecg[2] / jac;
      }
  geometric_symmetry(g, met, 10);
}

/* bdry_t.cu:291-324 */
const float CONSPD = 2.0f, CONSPB = 2.0f, CONSPA = 1.0f;
float pml_cal_R(float num_lay) {
  return (float)(pow(10, -((log10((double)num_lay) - 1.0) / log10(2.0) + 4.0)));
}
float pml_cal_dmax(float L, float Vp, float Rpp) {
  return (float)(-Vp / (2.0 * L) * log(Rpp) * (CONSPD + 1.0));
}
float pml_cal_d(float x, float L, float dmax) { return (x < 0) ? 0.0f : (float)(dmax * pow(x / L, CONSPD)); }
float pml_cal_a(float x, float L, float amax) { return (x < 0) ? 0.0f : (float)(amax * (1.0 - pow(x / L, CONSPA))); }
float pml_cal_b(float x, float L, float bmax) { return (x < 0) ? 1.0f : (float)(1.0 + (bmax - 1.0) * pow(x / L, CONSPB)); }

/* bdry_cal_abl_len_dh, bdry_t.cu:328-445 */
void abl_len_dh(const Grid &g, const float *x3d, const float *y3d, const float *z3d,
                int i1, int i2, int j1, int j2, int k1, int k2, int dim, float *avg_L, float *avg_dh) {
  double L = 0.0;
  int num = 0;
  auto seg = [&](size_t a, size_t b) {
    double x0 = x3d[a], y0 = y3d[a], z0 = z3d[a], x1 = x3d[b], y1 = y3d[b], z1 = z3d[b];
    L += sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0) + (z1 - z0) * (z1 - z0));
    num += 1;
  };
  if (dim == 0) {
    for (int k = k1; k <= k2; k++)
      for (int j = j1; j <= j2; j++)
        for (int i = i1 + 1; i <= i2; i++) seg((i - 1) + j * g.siy + k * g.siz, i + j * g.siy + k * g.siz);
    *avg_dh = (float)(L / num);
    *avg_L = (*avg_dh) * (i2 - i1);
  } else if (dim == 1) {
    for (int k = k1; k <= k2; k++)
      for (int i = i1; i <= i2; i++)
        for (int j = j1 + 1; j <= j2; j++) seg(i + (j - 1) * g.siy + k * g.siz, i + j * g.siy + k * g.siz);
    *avg_dh = (float)(L / num);
    *avg_L = (*avg_dh) * (j2 - j1);
  } else {
    for (int j = j1; j <= j2; j++)
      for (int i = i1; i <= i2; i++)
        for (int k = k1 + 1; k <= k2; k++) seg(i + j * g.siy + (k - 1) * g.siz, i + j * g.siy + k * g.siz);
    *avg_dh = (float)(L / num);
    *avg_L = (*avg_dh) * (k2 - k1);
  }
}

/* fdlib_math_dist_point2plane, src/lib/fdlib_math.cu:87-104 */
float dist_point2plane(const float x0[3], const float x1[3], const float x2[3], const float x3[3]) {
  float x12[3], x13[3], p[3];
  for (int a = 0; a < 3; a++) { x12[a] = x2[a] - x1[a]; x13[a] = x3[a] - x1[a]; }
  cross3f(x12, x13, p);
  float d = dot3f(p, x1);
  float L = (float)fabs(dot3f(p, x0) - d);
  L = L / sqrtf(dot3f(p, p));
  return L;
}

}  // namespace

struct cgfd_setup {
  Grid g;
  cgfd_setup_params p;
  std::vector<float> coords;  /* 3 x sicmp */
  std::vector<float> metric;  /* 10 x sicmp */
  std::vector<float> media;   /* lambda, mu, rho(density) then slowness */
  std::vector<float> slw;
  /* pml */
  int pml_on[3][2];
  int pr[3][2][6];
  std::vector<float> pA[3][2], pB[3][2], pD[3][2];
  /* faults */
  struct F {
    std::vector<float> in[7];  /* T0x T0y T0z mu_s mu_d Dc C0 on ny*nz */
    std::vector<int> united, faultgrid, rupy, rupz;
    std::vector<float> Tn0, Ts10, Ts20;
    std::vector<float> rho_f, lam_f, mu_f, vec_n, vec_s1, vec_s2, x_et, y_et, z_et;
    std::vector<float> D[12], P2M[5], M2P[5], M2[6], Fm[14];
  } f[CGFD_MAX_FAULTS];
  /* further scenarios on the same grid: fault inputs and initial tractions */
  struct Scen {
    std::vector<float> in[CGFD_MAX_FAULTS][7];
    std::vector<float> Tn0[CGFD_MAX_FAULTS], Ts10[CGFD_MAX_FAULTS], Ts20[CGFD_MAX_FAULTS];
  };
  std::vector<std::unique_ptr<Scen>> scen;
};

extern "C" {

float cgfd_pml_cal_R(float num_lay) { return pml_cal_R(num_lay); }
float cgfd_pml_cal_dmax(float L, float Vp, float Rpp) { return pml_cal_dmax(L, Vp, Rpp); }

int cgfd_setup_metric(int ni, int nj, int nk, const float *x3d, const float *y3d, const float *z3d, float *metric) {
  if (ni < 1 || nj < 1 || nk < 1 || !x3d || !y3d || !z3d || !metric) return cgfd_set_error(CGFD_ERR_ARG, "cgfd_setup_metric: bad arguments");
  Grid g;
  g.set(ni, nj, nk);
  metric_cal(g, x3d, y3d, z3d, metric);
  return CGFD_OK;
}

int cgfd_setup_create(const cgfd_setup_params *p, cgfd_setup **out) {
  if (!p || !out) return cgfd_set_error(CGFD_ERR_ARG, "cgfd_setup_create: null argument");
  *out = nullptr;
  if (p->ni < 8 || p->nj < 8 || p->nk < 8)
    return cgfd_set_error(CGFD_ERR_ARG, "grid too small: ni=%d nj=%d nk=%d (each must be >= 8)", p->ni, p->nj, p->nk);
  if (!p->x || !p->y || !p->z || !p->lambda || !p->mu || !p->rho)
    return cgfd_set_error(CGFD_ERR_ARG, "coordinates and media are required");
  if (p->number_fault < 0 || p->number_fault > CGFD_MAX_FAULTS)
    return cgfd_set_error(CGFD_ERR_ARG, "number_fault=%d out of range [0,%d]", p->number_fault, CGFD_MAX_FAULTS);
  if (p->number_fault > 0 && p->imethod != 1 && p->imethod != 2)
    return cgfd_set_error(CGFD_ERR_ARG, "dynamic_method must be 1 or 2 (got %d)", p->imethod);
  for (int id = 0; id < p->number_fault; id++) {
    /* par_t.cu:324-329 */
    if (p->fault_x_index[id] < 0 || p->fault_x_index[id] > p->ni)
      return cgfd_set_error(CGFD_ERR_ARG, "fault index %d out calculate area!", p->fault_x_index[id]);
    if (p->fault_x_index[id] < 4 || p->fault_x_index[id] > p->ni - 5)
      return cgfd_set_error(CGFD_ERR_ARG, "fault %d: x index %d leaves no room for the 7-column fault stencil", id, p->fault_x_index[id]);
    for (int q = 0; q < 7; q++)
      if (!p->init_stress[id][q]) return cgfd_set_error(CGFD_ERR_ARG, "fault %d: init_stress field %d missing", id, q);
  }
  cgfd_setup *s = new (std::nothrow) cgfd_setup();
  if (!s) return cgfd_set_error(CGFD_ERR_NOMEM, "host allocation failed");
  s->p = *p;
  Grid &g = s->g;
  g.set(p->ni, p->nj, p->nk);
  try {
    /* ---- grid: physical coordinates, ghosts by geometric_symmetry
       (gd_curv_gen_fault ends with it, gd_t.cu:542) ---- */
    s->coords.assign(3 * g.sicmp, 0.0f);
    const float *src[3] = {p->x, p->y, p->z};
    for (int c = 0; c < 3; c++) {
      float *dst = s->coords.data() + c * g.sicmp;
      for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++)
          for (int i = 0; i < g.ni; i++)
            dst[(i + 3) + (j + 3) * g.siy + (k + 3) * g.siz] = src[c][i + (size_t)j * g.ni + (size_t)k * g.ni * g.nj];
    }
    geometric_symmetry(g, s->coords.data(), 3);
    const float *x3d = s->coords.data(), *y3d = x3d + g.sicmp, *z3d = y3d + g.sicmp;

    /* ---- metric ---- */
    s->metric.assign(10 * g.sicmp, 0.0f);
    metric_cal(g, x3d, y3d, z3d, s->metric.data());

    /* ---- media: physical values + mirror_symmetry (md_import, md_t.cu:263-278) ---- */
    s->media.assign(3 * g.sicmp, 0.0f);
    const float *msrc[3] = {p->lambda, p->mu, p->rho};
    for (int c = 0; c < 3; c++) {
      float *dst = s->media.data() + c * g.sicmp;
      for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++)
          for (int i = 0; i < g.ni; i++)
            dst[(i + 3) + (j + 3) * g.siy + (k + 3) * g.siz] = msrc[c][i + (size_t)j * g.ni + (size_t)k * g.ni * g.nj];
    }
    mirror_symmetry(g, s->media.data(), 3);
    const float *lam3d = s->media.data(), *mu3d = lam3d + g.sicmp, *rho3d = mu3d + g.sicmp;
    const float *M = s->metric.data();
    const float *jac3d = M, *xi_x = M + g.sicmp, *xi_y = xi_x + g.sicmp, *xi_z = xi_y + g.sicmp;
    const float *et_x = xi_z + g.sicmp, *et_y = et_x + g.sicmp, *et_z = et_y + g.sicmp;
    const float *zt_x = et_z + g.sicmp, *zt_y = zt_x + g.sicmp, *zt_z = zt_y + g.sicmp;

    /* ---- PML: bdry_pml_set, bdry_t.cu:79-251 ---- */
    for (int dim = 0; dim < 3; dim++)
      for (int side = 0; side < 2; side++) {
        int nl = p->pml_layers[dim][side];
        int on = nl > 0;
        if (dim == 2 && side == 1 && p->free_top) on = 0;
        s->pml_on[dim][side] = on;
        int r[6] = {g.ni1, g.ni2, g.nj1, g.nj2, g.nk1, g.nk2};
        if (!on) nl = 0;
        if (dim == 0 && side == 0) r[1] = g.ni1 + nl;
        if (dim == 0 && side == 1) r[0] = g.ni2 - nl;
        if (dim == 1 && side == 0) r[3] = g.nj1 + nl;
        if (dim == 1 && side == 1) r[2] = g.nj2 - nl;
        if (dim == 2 && side == 0) r[5] = g.nk1 + nl;
        if (dim == 2 && side == 1) r[4] = g.nk2 - nl;
        for (int q = 0; q < 6; q++) s->pr[dim][side][q] = r[q];
        if (!on) continue;
        if (nl >= (dim == 0 ? g.ni : dim == 1 ? g.nj : g.nk)) {
          delete s;
          return cgfd_set_error(CGFD_ERR_ARG, "PML of %d layers does not fit axis %d", nl, dim);
        }
        float L0, dh;
        abl_len_dh(g, x3d, y3d, z3d, r[0], r[1], r[2], r[3], r[4], r[5], dim, &L0, &dh);
        int npoints = nl + 1;
        float num_lay = (float)(npoints - 1);
        float Rpp = pml_cal_R(num_lay);
        float dmax = pml_cal_dmax(L0, p->pml_velocity[dim][side], Rpp);
        float amax = p->pml_alpha_max[dim][side];
        float bmax = p->pml_beta_max[dim][side];
        std::vector<float> &A = s->pA[dim][side], &B = s->pB[dim][side], &D = s->pD[dim][side];
        A.assign(npoints, 0.0f); B.assign(npoints, 0.0f); D.assign(npoints, 0.0f);
        for (int n = 0; n < npoints; n++) {
          float L = n * dh;
          int i = (side == 0) ? npoints - 1 - n : n;
          D[i] = pml_cal_d(L, L0, dmax);
          A[i] = pml_cal_a(L, L0, amax);
          B[i] = pml_cal_b(L, L0, bmax);
          D[i] /= B[i];
          A[i] += D[i];
          B[i] = 1.0 / B[i];
        }
      }

    /* ---- faults ---- */
    for (int id = 0; id < p->number_fault; id++) {
      cgfd_setup::F &F = s->f[id];
      int i0 = p->fault_x_index[id] + 3;
      size_t ns = g.sslice;
      F.rho_f.assign(2 * ns, 0.f); F.lam_f.assign(2 * ns, 0.f); F.mu_f.assign(2 * ns, 0.f);
      F.vec_n.assign(3 * ns, 0.f); F.vec_s1.assign(3 * ns, 0.f); F.vec_s2.assign(3 * ns, 0.f);
      F.x_et.assign(ns, 0.f); F.y_et.assign(ns, 0.f); F.z_et.assign(ns, 0.f);
      for (int q = 0; q < 12; q++) F.D[q].assign(9 * ns, 0.f);
      for (int q = 0; q < 5; q++) { F.P2M[q].assign(9 * ns, 0.f); F.M2P[q].assign(9 * ns, 0.f); }
      for (int q = 0; q < 6; q++) F.M2[q].assign(9 * ns, 0.f);
      for (int q = 0; q < 14; q++) F.Fm[q].assign(9 * (size_t)g.ny, 0.f);

      /* fault_coef_cal, fault_info.cu:155-848 */
      for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++) {
          size_t iptr = i0 + j * g.siy + k * g.siz;
          size_t pf = j + (size_t)k * g.ny;
          float rho = rho3d[iptr], mu = mu3d[iptr], lam = lam3d[iptr];
          float lam2mu = lam + 2.0 * mu;
          float jac = jac3d[iptr];
          float e1[3] = {xi_x[iptr], xi_y[iptr], xi_z[iptr]};
          float e2[3] = {et_x[iptr], et_y[iptr], et_z[iptr]};
          float e3[3] = {zt_x[iptr], zt_y[iptr], zt_z[iptr]};
          F.rho_f[pf] = rho; F.lam_f[pf] = lam; F.mu_f[pf] = mu;
          F.rho_f[pf + ns] = rho; F.lam_f[pf + ns] = lam; F.mu_f[pf + ns] = mu;
          float d11[9], d12[9], d13[9];
          dmat(e1, e1, lam, mu, lam2mu, d11);
          dmat(e1, e2, lam, mu, lam2mu, d12);
          dmat(e1, e3, lam, mu, lam2mu, d13);
          const float *rows[3] = {e1, e2, e3};
          /* stored D21..D33 for side 1 ("_1", minus) then side 2 ("_2", plus);
             both sides share the medium on one process */
          float Dst[12][9];
          for (int sd = 0; sd < 2; sd++)
            for (int a = 1; a <= 2; a++)
              for (int b = 0; b < 3; b++) dmat(rows[a], rows[b], lam, mu, lam2mu, Dst[sd * 6 + (a - 1) * 3 + b]);
          for (int q = 0; q < 9; q++) { d11[q] *= jac; d12[q] *= jac; d13[q] *= jac; }
          for (int sd = 0; sd < 12; sd++)
            for (int q = 0; q < 9; q++) Dst[sd][q] *= jac;
          for (int sd = 0; sd < 12; sd++)
            for (int q = 0; q < 9; q++) F.D[sd][pf * 9 + q] = Dst[sd][q];
          float D11[3][3], D12[3][3], D13[3][3], mat1[3][3], mat2[3][3], mat3[3][3];
          to33(d11, D11); to33(d12, D12); to33(d13, D13);
          /* method 1, minus -> plus: inv(D11_2) * {D11_1, D12_1, D13_1, D12_2, D13_2} */
          to33(d11, mat1);
          invert3x3f(mat1);
          float (*Mp[5])[3] = {D11, D12, D13, D12, D13};
          for (int q = 0; q < 5; q++) { matmul3x3f(mat1, Mp[q], mat2); from33(mat2, &F.M2P[q][pf * 9]); }
          /* plus -> minus: inv(D11_1) * {D11_2, D12_2, D13_2, D12_1, D13_1} */
          to33(d11, mat1);
          invert3x3f(mat1);
          for (int q = 0; q < 5; q++) { matmul3x3f(mat1, Mp[q], mat2); from33(mat2, &F.P2M[q][pf * 9]); }
          /* method 2: minus T1,Vy,Vz -> Vx then plus (fault_info.cu:528-607) */
          for (int sd = 0; sd < 2; sd++) {
            to33(d11, mat1);
            invert3x3f(mat1);
            from33(mat1, &F.M2[sd * 3 + 0][pf * 9]);
            matmul3x3f(mat1, D12, mat2); from33(mat2, &F.M2[sd * 3 + 1][pf * 9]);
            matmul3x3f(mat1, D13, mat2); from33(mat2, &F.M2[sd * 3 + 2][pf * 9]);
          }
          /* strike / normal / dip vectors, fault_info.cu:609-654 */
          mat3[0][0] = e1[0]; mat3[0][1] = e1[1]; mat3[0][2] = e1[2];
          mat3[1][0] = e2[0]; mat3[1][1] = e2[1]; mat3[1][2] = e2[2];
          mat3[2][0] = e3[0]; mat3[2][1] = e3[1]; mat3[2][2] = e3[2];
          invert3x3f(mat3);
          F.x_et[pf] = mat3[0][1]; F.y_et[pf] = mat3[1][1]; F.z_et[pf] = mat3[2][1];
          float vs1[3] = {mat3[0][1], mat3[1][1], mat3[2][1]};
          float nrm = norm3f(vs1);
          for (int a = 0; a < 3; a++) vs1[a] /= nrm;
          float vn[3] = {e1[0], e1[1], e1[2]};
          nrm = norm3f(vn);
          for (int a = 0; a < 3; a++) vn[a] /= nrm;
          float vs2[3];
          cross3f(vn, vs1, vs2);
          for (int a = 0; a < 3; a++) {
            F.vec_n[pf * 3 + a] = vn[a]; F.vec_s1[pf * 3 + a] = vs1[a]; F.vec_s2[pf * 3 + a] = vs2[a];
          }
          /* free-surface row, fault_info.cu:656-845 (evaluated whatever the top
             boundary is, as upstream) */
          if (k == g.nk2) {
            float A[3][3], Bm[3][3], Cm[3][3];
            float Vx2Vz1[3][3], Vy2Vz1[3][3], Vx2Vz2[3][3], Vy2Vz2[3][3];
            for (int sd = 0; sd < 2; sd++) {
              const float *D31 = &F.D[sd * 6 + 3][pf * 9], *D32 = &F.D[sd * 6 + 4][pf * 9], *D33 = &F.D[sd * 6 + 5][pf * 9];
              for (int q = 0; q < 9; q++) { A[q / 3][q % 3] = D33[q]; Bm[q / 3][q % 3] = -D31[q]; Cm[q / 3][q % 3] = -D32[q]; }
              invert3x3f(A);
              matmul3x3f(A, Bm, sd == 0 ? Vx2Vz1 : Vx2Vz2);
              matmul3x3f(A, Cm, sd == 0 ? Vy2Vz1 : Vy2Vz2);
            }
            float vn2[3] = {e1[0], e1[1], e1[2]};
            float s1b[3] = {F.x_et[pf], F.y_et[pf], F.z_et[pf]};
            float s2b[3];
            cross3f(vn2, s1b, s2b);
            nrm = norm3f(s2b);
            for (int a = 0; a < 3; a++) s2b[a] /= nrm;
            float g1_2 = 0.0f, g2_2 = 0.0f, g3_2 = 0.0f;
            float h1_3 = s2b[0] * e1[0] + s2b[1] * e1[1] + s2b[2] * e1[2];
            float h2_3 = s2b[0] * e2[0] + s2b[1] * e2[1] + s2b[2] * e2[2];
            float h3_3 = s2b[0] * e3[0] + s2b[1] * e3[1] + s2b[2] * e3[2];
            nrm = norm3f(vn2);
            h1_3 = h1_3 / nrm; h2_3 = h2_3 / nrm; h3_3 = h3_3 / nrm;
            float D11f[3][3], D12f[3][3], D13f[3][3];
            for (int ii = 0; ii < 3; ii++)
              for (int jj = 0; jj < 3; jj++) {
                float To = s1b[ii] * vn2[jj];
                float D11_12 = mu * g1_2 * To, D12_12 = mu * g2_2 * To, D13_12 = mu * g3_2 * To;
                float To3 = s2b[ii] * vn2[jj];
                float D11_13 = mu * h1_3 * To3, D12_13 = mu * h2_3 * To3, D13_13 = mu * h3_3 * To3;
                D11f[ii][jj] = D11[ii][jj] + D11_12 + D11_13;
                D12f[ii][jj] = D12[ii][jj] + D12_12 + D12_13;
                D13f[ii][jj] = D13[ii][jj] + D13_12 + D13_13;
              }
            /* both sides share D11f etc. on one process (D11_1 == D11_2) */
            float m1[3][3], m2[3][3], m3[3][3], m4[3][3];
            matmul3x3f(D13f, Vx2Vz1, m1);
            matmul3x3f(D13f, Vy2Vz1, m2);
            matmul3x3f(D13f, Vx2Vz2, m3);
            matmul3x3f(D13f, Vy2Vz2, m4);
            float Vx1f[3][3], Vy1f[3][3], Vx2f[3][3], Vy2f[3][3];
            for (int ii = 0; ii < 3; ii++)
              for (int jj = 0; jj < 3; jj++) {
                Vx1f[ii][jj] = D11f[ii][jj] + m1[ii][jj];
                Vy1f[ii][jj] = D12f[ii][jj] + m2[ii][jj];
                Vx2f[ii][jj] = D11f[ii][jj] + m3[ii][jj];
                Vy2f[ii][jj] = D12f[ii][jj] + m4[ii][jj];
              }
            float To[3][3], P2M1f[3][3], P2M2f[3][3], P2M3f[3][3], M2P1f[3][3], M2P2f[3][3], M2P3f[3][3];
            for (int ii = 0; ii < 3; ii++)
              for (int jj = 0; jj < 3; jj++) To[ii][jj] = Vx1f[ii][jj];
            invert3x3f(To);
            matmul3x3f(To, Vx2f, P2M1f);
            matmul3x3f(To, Vy2f, P2M2f);
            matmul3x3f(To, Vy1f, P2M3f);
            float *jrow = nullptr;
            (void)jrow;
            size_t jf = (size_t)j * 9;
            from33(To, &F.Fm[10][jf]);     /* matT1toVxf_Min */
            from33(P2M3f, &F.Fm[11][jf]);  /* matVytoVxf_Min */
            for (int ii = 0; ii < 3; ii++)
              for (int jj = 0; jj < 3; jj++) To[ii][jj] = Vx2f[ii][jj];
            invert3x3f(To);
            matmul3x3f(To, Vx1f, M2P1f);
            matmul3x3f(To, Vy1f, M2P2f);
            matmul3x3f(To, Vy2f, M2P3f);
            from33(To, &F.Fm[12][jf]);     /* matT1toVxf_Plus */
            from33(M2P3f, &F.Fm[13][jf]);  /* matVytoVxf_Plus */
            from33(Vx2Vz1, &F.Fm[0][jf]);
            from33(Vy2Vz1, &F.Fm[1][jf]);
            from33(Vx2Vz2, &F.Fm[2][jf]);
            from33(Vy2Vz2, &F.Fm[3][jf]);
            from33(P2M1f, &F.Fm[4][jf]);
            from33(P2M2f, &F.Fm[5][jf]);
            from33(P2M3f, &F.Fm[6][jf]);
            from33(M2P1f, &F.Fm[7][jf]);
            from33(M2P2f, &F.Fm[8][jf]);
            from33(M2P3f, &F.Fm[9][jf]);
          }
        }

      /* fault_set, fault_info.cu:967-1074 */
      for (int q = 0; q < 7; q++) F.in[q].assign(ns, 0.f);
      for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++)
          for (int q = 0; q < 7; q++)
            F.in[q][(j + 3) + (size_t)(k + 3) * g.ny] = p->init_stress[id][q][j + (size_t)k * g.nj];
      F.united.assign(ns, 1); F.faultgrid.assign(ns, 0); F.rupy.assign(ns, 0); F.rupz.assign(ns, 0);
      F.Tn0.assign(ns, 0.f); F.Ts10.assign(ns, 0.f); F.Ts20.assign(ns, 0.f);
      const int *fg = &p->fault_grid[4 * id];
      int margin = p->united_margin;
      for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++) {
          size_t pf = (j + 3) + (size_t)(k + 3) * g.ny;
          const float *vn = &F.vec_n[pf * 3], *s1 = &F.vec_s1[pf * 3], *s2 = &F.vec_s2[pf * 3];
          float T0x = F.in[0][pf], T0y = F.in[1][pf], T0z = F.in[2][pf];
          F.Tn0[pf] = T0x * vn[0] + T0y * vn[1] + T0z * vn[2];
          F.Ts10[pf] = T0x * s1[0] + T0y * s1[1] + T0z * s1[2];
          F.Ts20[pf] = T0x * s2[0] + T0y * s2[1] + T0z * s2[2];
          int gj = j, gk = k;  /* gnj1 = gnk1 = 0 on one process */
          F.rupy[pf] = (gj <= fg[0] + 3 || gj >= fg[1] - 3) ? 0 : 1;
          F.rupz[pf] = (gk <= fg[2] + 3 || gk >= fg[3] - 3) ? 0 : 1;
          if (p->free_top) {
            F.united[pf] = (gj >= margin && gj < g.nj - margin && gk >= margin) ? 0 : 1;
          } else {
            F.united[pf] = (gj >= margin && gj < g.nj - margin && gk >= margin && gk < g.nk - margin) ? 0 : 1;
          }
          F.faultgrid[pf] = (gj >= fg[0] + 3 && gj <= fg[1] - 3 && gk >= fg[2] + 3 && gk <= fg[3] - 3) ? 1 : 0;
        }
    }

    /* ---- md_rho_to_slow, md_t.cu:568-582 (after the fault setup, as in main) ---- */
    s->slw.assign(g.sicmp, 0.f);
    for (size_t i = 0; i < g.sicmp; i++) s->slw[i] = rho3d[i] > 1e-10 ? 1.0 / rho3d[i] : 0.0f;
  } catch (const std::bad_alloc &) {
    delete s;
    return cgfd_set_error(CGFD_ERR_NOMEM, "host allocation failed in setup");
  }
  *out = s;
  return CGFD_OK;
}

int cgfd_setup_destroy(cgfd_setup *s) {
  delete s;
  return CGFD_OK;
}

int cgfd_setup_get_desc(const cgfd_setup *s, cgfd_desc *d) {
  if (!s || !d) return cgfd_set_error(CGFD_ERR_ARG, "cgfd_setup_get_desc: null argument");
  std::memset(d, 0, sizeof(*d));
  const Grid &g = s->g;
  d->ni = g.ni; d->nj = g.nj; d->nk = g.nk;
  d->metric = s->metric.data();
  d->lambda = s->media.data();
  d->mu = s->media.data() + g.sicmp;
  d->rho = s->slw.data();
  d->free_top = s->p.free_top;
  for (int dim = 0; dim < 3; dim++)
    for (int side = 0; side < 2; side++) {
      d->pml_is_sides[dim][side] = s->pml_on[dim][side];
      d->pml_ni1[dim][side] = s->pr[dim][side][0]; d->pml_ni2[dim][side] = s->pr[dim][side][1];
      d->pml_nj1[dim][side] = s->pr[dim][side][2]; d->pml_nj2[dim][side] = s->pr[dim][side][3];
      d->pml_nk1[dim][side] = s->pr[dim][side][4]; d->pml_nk2[dim][side] = s->pr[dim][side][5];
      if (s->pml_on[dim][side]) {
        d->pml_A[dim][side] = s->pA[dim][side].data();
        d->pml_B[dim][side] = s->pB[dim][side].data();
        d->pml_D[dim][side] = s->pD[dim][side].data();
      }
    }
  d->imethod = s->p.imethod;
  d->number_fault = s->p.number_fault;
  for (int id = 0; id < s->p.number_fault; id++) {
    const cgfd_setup::F &F = s->f[id];
    cgfd_fault_desc &fd = d->fault[id];
    fd.fault_x_index = s->p.fault_x_index[id];
    fd.T0x = F.in[0].data(); fd.T0y = F.in[1].data(); fd.T0z = F.in[2].data();
    fd.mu_s = F.in[3].data(); fd.mu_d = F.in[4].data(); fd.Dc = F.in[5].data(); fd.C0 = F.in[6].data();
    fd.united = F.united.data(); fd.faultgrid = F.faultgrid.data();
    fd.rup_index_y = F.rupy.data(); fd.rup_index_z = F.rupz.data();
    fd.Tn0 = F.Tn0.data(); fd.Ts10 = F.Ts10.data(); fd.Ts20 = F.Ts20.data();
    fd.rho_f = F.rho_f.data(); fd.vec_s1 = F.vec_s1.data(); fd.vec_s2 = F.vec_s2.data();
    for (int q = 0; q < 12; q++) fd.D[q] = F.D[q].data();
    for (int q = 0; q < 5; q++) { fd.matPlus2Min[q] = F.P2M[q].data(); fd.matMin2Plus[q] = F.M2P[q].data(); }
    for (int q = 0; q < 6; q++) fd.matM2[q] = F.M2[q].data();
    for (int q = 0; q < 14; q++) fd.matF[q] = F.Fm[q].data();
  }
  d->dt = s->p.dt; d->t0 = s->p.t0; d->nt_total = s->p.nt_total;
  d->device_id = s->p.device_id;
  return CGFD_OK;
}

int cgfd_setup_scenario_desc(cgfd_setup *s, const float *const *init_stress, int nt_total, cgfd_desc *d) {
  if (!s || !d || (s->p.number_fault > 0 && !init_stress))
    return cgfd_set_error(CGFD_ERR_ARG, "cgfd_setup_scenario_desc: null argument");
  const Grid &g = s->g;
  const size_t ns = (size_t)g.ny * g.nz;
  for (int q = 0; q < 7 * s->p.number_fault; q++)
    if (!init_stress[q]) return cgfd_set_error(CGFD_ERR_ARG, "cgfd_setup_scenario_desc: input array %d missing", q);
  std::unique_ptr<cgfd_setup::Scen> sc(new (std::nothrow) cgfd_setup::Scen());
  if (!sc) return cgfd_set_error(CGFD_ERR_NOMEM, "host allocation failed");
  for (int id = 0; id < s->p.number_fault; id++) {
    const cgfd_setup::F &F = s->f[id];
    for (int q = 0; q < 7; q++) sc->in[id][q].assign(ns, 0.f);
    for (int k = 0; k < g.nk; k++)
      for (int j = 0; j < g.nj; j++)
        for (int q = 0; q < 7; q++)
          sc->in[id][q][(j + 3) + (size_t)(k + 3) * g.ny] = init_stress[7 * id + q][j + (size_t)k * g.nj];
    sc->Tn0[id].assign(ns, 0.f); sc->Ts10[id].assign(ns, 0.f); sc->Ts20[id].assign(ns, 0.f);
    /* the same projections as fault_set (fault_info.cu:1015-1023) */
    for (int k = 0; k < g.nk; k++)
      for (int j = 0; j < g.nj; j++) {
        size_t pf = (j + 3) + (size_t)(k + 3) * g.ny;
        const float *vn = &F.vec_n[pf * 3], *s1 = &F.vec_s1[pf * 3], *s2 = &F.vec_s2[pf * 3];
        float T0x = sc->in[id][0][pf], T0y = sc->in[id][1][pf], T0z = sc->in[id][2][pf];
        sc->Tn0[id][pf] = T0x * vn[0] + T0y * vn[1] + T0z * vn[2];
        sc->Ts10[id][pf] = T0x * s1[0] + T0y * s1[1] + T0z * s1[2];
        sc->Ts20[id][pf] = T0x * s2[0] + T0y * s2[1] + T0z * s2[2];
      }
  }
  int rc = cgfd_setup_get_desc(s, d);
  if (rc) return rc;
  for (int id = 0; id < s->p.number_fault; id++) {
    cgfd_fault_desc &fd = d->fault[id];
    fd.T0x = sc->in[id][0].data(); fd.T0y = sc->in[id][1].data(); fd.T0z = sc->in[id][2].data();
    fd.mu_s = sc->in[id][3].data(); fd.mu_d = sc->in[id][4].data(); fd.Dc = sc->in[id][5].data();
    fd.C0 = sc->in[id][6].data();
    fd.Tn0 = sc->Tn0[id].data(); fd.Ts10 = sc->Ts10[id].data(); fd.Ts20 = sc->Ts20[id].data();
  }
  d->nt_total = nt_total;
  s->scen.push_back(std::move(sc));
  return CGFD_OK;
}

const float *cgfd_setup_coords(const cgfd_setup *s) { return s ? s->coords.data() : nullptr; }

const float *cgfd_setup_vec_n(const cgfd_setup *s, int id) {
  if (!s || id < 0 || id >= s->p.number_fault) return nullptr;
  return s->f[id].vec_n.data();
}

int cgfd_setup_dt_max(const cgfd_setup *s, float *dtmax) {
  if (!s || !dtmax) return cgfd_set_error(CGFD_ERR_ARG, "cgfd_setup_dt_max: null argument");
  const Grid &g = s->g;
  const float CFL = 1.3f;  /* fd_t.cu:57 */
  const float *x3d = s->coords.data(), *y3d = x3d + g.sicmp, *z3d = y3d + g.sicmp;
  const float *lam = s->media.data(), *mu = lam + g.sicmp, *rho = mu + g.sicmp;
  float dtm = 1.0e10f;
  for (int k = g.nk1; k <= g.nk2; k++)
    for (int j = g.nj1; j <= g.nj2; j++)
      for (int i = g.ni1; i <= g.ni2; i++) {
        size_t iptr = i + j * g.siy + k * g.siz;
        float Vp = sqrt((lam[iptr] + 2.0 * mu[iptr]) / rho[iptr]);
        float dtLe = 1.0e20f;
        float p0[3] = {x3d[iptr], y3d[iptr], z3d[iptr]};
        for (int kk = -1; kk <= 1; kk++)
          for (int jj = -1; jj <= 1; jj++)
            for (int ii = -1; ii <= 1; ii++)
              if (ii != 0 && jj != 0 && kk != 0) {
                size_t a = iptr - ii, b = iptr - jj * g.siy, c = iptr - kk * g.siz;
                float p1[3] = {x3d[a], y3d[a], z3d[a]};
                float p2[3] = {x3d[b], y3d[b], z3d[b]};
                float p3[3] = {x3d[c], y3d[c], z3d[c]};
                float L = dist_point2plane(p0, p1, p2, p3);
                if (dtLe > L) dtLe = L;
              }
        float dtp = CFL / Vp * dtLe;
        if (dtp < dtm) dtm = dtp;
      }
  *dtmax = dtm;
  return CGFD_OK;
}

}  /* extern "C" */
