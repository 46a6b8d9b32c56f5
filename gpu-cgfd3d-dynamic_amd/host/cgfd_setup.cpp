/*
 * cgfd_setup.cpp -- host setup of the C++ driver (see include/cgfd_setup.h).
 *
 * Every routine restates a host routine of the reference that runs before the
 * time loop; the comment on each names the reference lines.
 */
#include "../../include/cgfd_setup.h"
#include "cgfd_errors.h"
#include "cgfd_setup_node.h"
#include "cgfd_setup_dev.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include <memory>

namespace {

using namespace cgfd_setup_math;

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
#pragma omp parallel for schedule(static)
  for (int k = g.nk1; k <= g.nk2; k++)
    for (int j = g.nj1; j <= g.nj2; j++)
      for (int i = g.ni1; i <= g.ni2; i++) {
        size_t iptr = i + j * g.siy + k * g.siz;
        float o[10];
        metric_point(x3d, y3d, z3d, iptr, (long)g.siy, (long)g.siz, o);
        for (int c = 0; c < 10; c++) met[c * g.sicmp + iptr] = o[c];
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

    /* ---- metric: imported at the physical points (gd_curv_metric_import,
       gd_t.cu:1015-1068, then geometric_symmetry), or computed on the host or
       on the device ---- */
    const bool dev = p->device_setup != 0;
    s->metric.assign(10 * g.sicmp, 0.0f);
    if (p->metric) {
      for (int c = 0; c < 10; c++) {
        float *dst = s->metric.data() + c * g.sicmp;
        const float *src_c = p->metric + (size_t)c * g.ni * g.nj * g.nk;
        for (int k = 0; k < g.nk; k++)
          for (int j = 0; j < g.nj; j++)
            for (int i = 0; i < g.ni; i++)
              dst[(i + 3) + (j + 3) * g.siy + (k + 3) * g.siz] = src_c[i + (size_t)j * g.ni + (size_t)k * g.ni * g.nj];
      }
      geometric_symmetry(g, s->metric.data(), 10);
    } else if (dev) {
      int rc = cgfd_dev_metric(p->device_id, g.ni, g.nj, g.nk, s->coords.data(), s->metric.data());
      if (rc != CGFD_OK) { delete s; return rc; }
      geometric_symmetry(g, s->metric.data(), 10);
    } else {
      metric_cal(g, x3d, y3d, z3d, s->metric.data());
    }

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

      /* fault_coef_cal, fault_info.cu:155-848: per node on the host, or
         the same node arithmetic on the device */
      FaultCoefOut out;
      for (int q = 0; q < 12; q++) out.D[q] = F.D[q].data();
      for (int q = 0; q < 5; q++) { out.P2M[q] = F.P2M[q].data(); out.M2P[q] = F.M2P[q].data(); }
      for (int q = 0; q < 6; q++) out.M2[q] = F.M2[q].data();
      for (int q = 0; q < 14; q++) out.Fm[q] = F.Fm[q].data();
      out.vec_n = F.vec_n.data(); out.vec_s1 = F.vec_s1.data(); out.vec_s2 = F.vec_s2.data();
      out.x_et = F.x_et.data(); out.y_et = F.y_et.data(); out.z_et = F.z_et.data();
      std::vector<float> node_in;
      if (dev) node_in.assign(kFaultNodeIn * ns, 0.f);
      for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++) {
          size_t iptr = i0 + j * g.siy + k * g.siz;
          size_t pf = j + (size_t)k * g.ny;
          float rho = rho3d[iptr], mu = mu3d[iptr], lam = lam3d[iptr];
          float jac = jac3d[iptr];
          float e1[3] = {xi_x[iptr], xi_y[iptr], xi_z[iptr]};
          float e2[3] = {et_x[iptr], et_y[iptr], et_z[iptr]};
          float e3[3] = {zt_x[iptr], zt_y[iptr], zt_z[iptr]};
          F.rho_f[pf] = rho; F.lam_f[pf] = lam; F.mu_f[pf] = mu;
          F.rho_f[pf + ns] = rho; F.lam_f[pf + ns] = lam; F.mu_f[pf + ns] = mu;
          if (!dev) fault_coef_node(out, pf, (size_t)j, e1, e2, e3, jac, lam, mu, k == g.nk2);
          else {
            float *q = &node_in[pf * kFaultNodeIn];
            for (int a = 0; a < 3; a++) { q[a] = e1[a]; q[3 + a] = e2[a]; q[6 + a] = e3[a]; }
            q[9] = jac; q[10] = lam; q[11] = mu;
          }
        }
      if (dev) {
        int rc = cgfd_dev_fault_coef(p->device_id, g.ny, g.nz, g.nk2, node_in.data(), out);
        if (rc != CGFD_OK) { delete s; return rc; }
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
