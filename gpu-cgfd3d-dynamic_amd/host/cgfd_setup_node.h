/*
 * cgfd_setup_node.h -- the per-point and per-fault-node arithmetic of the
 * host setup, shared by the host loops (cgfd_setup.cpp) and the device
 * kernels of the same setup (csrc/k_setup.hip): compiled for both sides
 * when included from HIP code.
 */
#ifndef CGFD_SETUP_NODE_H
#define CGFD_SETUP_NODE_H

#include <math.h>
#include <stddef.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define CGFD_HD __host__ __device__ inline
#else
#define CGFD_HD inline
#endif

namespace cgfd_setup_math {

/* 3x3 helpers with the reference's arithmetic (src/lib/fdlib_math.cu) */
CGFD_HD float norm3f(const float *A) {
  float x2 = A[0] * A[0], y2 = A[1] * A[1], z2 = A[2] * A[2];
  return sqrtf(x2 + y2 + z2);
}
CGFD_HD void invert3x3f(float m[][3]) {
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
CGFD_HD void matmul3x3f(float A[][3], float B[][3], float C[][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C[i][j] = 0.0f;
      for (int k = 0; k < 3; k++) C[i][j] += A[i][k] * B[k][j];
    }
}
CGFD_HD void cross3f(const float *A, const float *B, float *C) {
  C[0] = A[1] * B[2] - A[2] * B[1];
  C[1] = A[2] * B[0] - A[0] * B[2];
  C[2] = A[0] * B[1] - A[1] * B[0];
}
CGFD_HD float dot3f(const float *A, const float *B) {
  float r = 0.0f;
  for (int i = 0; i < 3; i++) r += A[i] * B[i];
  return r;
}

/* 6th-order centred difference, fd_t.h:126-140 (d_4 = 0 is skipped there) */
CGFD_HD float center6(const float *p, long s) {
  const double d1 = -0.01666667, d2 = 0.15, d3 = -0.75, d5 = 0.75, d6 = -0.15, d7 = 0.01666667;
  return (float)((d1 * p[-3 * s]) + (d2 * p[-2 * s]) + (d3 * p[-s]) + (d5 * p[s]) + (d6 * p[2 * s]) + (d7 * p[3 * s]));
}

/* Elastic traction-strain matrix D_ab of metric rows a, b at a node.
   Entry expressions and their evaluation order are those of
   fault_info.cu:190-369 (all 18 blocks share this pattern). */
CGFD_HD void dmat(const float *ea, const float *eb, float lam, float mu, float lam2mu, float D[9]) {
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

CGFD_HD void to33(const float *v, float M[][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) M[i][j] = v[3 * i + j];
}
CGFD_HD void from33(float M[][3], float *v) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) v[3 * i + j] = M[i][j];
}


/* gd_curv_metric_cal at one point (gd_t.cu:253-297): out = jac, xi_x..zeta_z */
CGFD_HD void metric_point(const float *x3d, const float *y3d, const float *z3d, size_t iptr, long siy, long siz,
                          float out[10]) {
  float vec1[3], vec2[3], vec3[3], vecg[3];
  vec1[0] = center6(x3d + iptr, 1); vec1[1] = center6(y3d + iptr, 1); vec1[2] = center6(z3d + iptr, 1);
  vec2[0] = center6(x3d + iptr, siy); vec2[1] = center6(y3d + iptr, siy); vec2[2] = center6(z3d + iptr, siy);
  vec3[0] = center6(x3d + iptr, siz); vec3[1] = center6(y3d + iptr, siz); vec3[2] = center6(z3d + iptr, siz);
  cross3f(vec1, vec2, vecg);
  const float jac = dot3f(vecg, vec3);
  out[0] = jac;
  cross3f(vec2, vec3, vecg);
  out[1] = vecg[0] / jac; out[2] = vecg[1] / jac; out[3] = vecg[2] / jac;
  cross3f(vec3, vec1, vecg);
  out[4] = vecg[0] / jac; out[5] = vecg[1] / jac; out[6] = vecg[2] / jac;
  cross3f(vec1, vec2, vecg);
  out[7] = vecg[0] / jac; out[8] = vecg[1] / jac; out[9] = vecg[2] / jac;
}

/* where fault_coef_cal's results of one node go: [node][9] matrices, [node][3]
   vectors, the free-surface row [j][9] */
struct FaultCoefOut {
  float *D[12], *P2M[5], *M2P[5], *M2[6], *Fm[14];
  float *vec_n, *vec_s1, *vec_s2, *x_et, *y_et, *z_et;
};

/* fault_coef_cal at one node (fault_info.cu:155-845): e1, e2, e3 the xi,
   eta, zeta gradients, both sides sharing the medium (one process). pf is
   the node index j + k ny, jf = j; top: the node is on the free-surface row. */
CGFD_HD void fault_coef_node(const FaultCoefOut &o, size_t pf, size_t jf, const float e1[3], const float e2[3],
                             const float e3[3], float jac, float lam, float mu, bool top) {
  const float lam2mu = lam + 2.0 * mu;
  float d11[9], d12[9], d13[9];
  dmat(e1, e1, lam, mu, lam2mu, d11);
  dmat(e1, e2, lam, mu, lam2mu, d12);
  dmat(e1, e3, lam, mu, lam2mu, d13);
  const float *rows[3] = {e1, e2, e3};
  /* stored D21..D33 for side 1 ("_1", minus) then side 2 ("_2", plus) */
  float Dst[12][9];
  for (int sd = 0; sd < 2; sd++)
    for (int a = 1; a <= 2; a++)
      for (int b = 0; b < 3; b++) dmat(rows[a], rows[b], lam, mu, lam2mu, Dst[sd * 6 + (a - 1) * 3 + b]);
  for (int q = 0; q < 9; q++) { d11[q] *= jac; d12[q] *= jac; d13[q] *= jac; }
  for (int sd = 0; sd < 12; sd++)
    for (int q = 0; q < 9; q++) Dst[sd][q] *= jac;
  for (int sd = 0; sd < 12; sd++)
    for (int q = 0; q < 9; q++) o.D[sd][pf * 9 + q] = Dst[sd][q];
  float D11[3][3], D12[3][3], D13[3][3], mat1[3][3], mat2[3][3], mat3[3][3];
  to33(d11, D11); to33(d12, D12); to33(d13, D13);
  /* method 1, minus -> plus: inv(D11_2) * {D11_1, D12_1, D13_1, D12_2, D13_2} */
  to33(d11, mat1);
  invert3x3f(mat1);
  float (*Mp[5])[3] = {D11, D12, D13, D12, D13};
  for (int q = 0; q < 5; q++) { matmul3x3f(mat1, Mp[q], mat2); from33(mat2, &o.M2P[q][pf * 9]); }
  /* plus -> minus: inv(D11_1) * {D11_2, D12_2, D13_2, D12_1, D13_1} */
  to33(d11, mat1);
  invert3x3f(mat1);
  for (int q = 0; q < 5; q++) { matmul3x3f(mat1, Mp[q], mat2); from33(mat2, &o.P2M[q][pf * 9]); }
  /* method 2: minus T1,Vy,Vz -> Vx then plus (fault_info.cu:528-607) */
  for (int sd = 0; sd < 2; sd++) {
    to33(d11, mat1);
    invert3x3f(mat1);
    from33(mat1, &o.M2[sd * 3 + 0][pf * 9]);
    matmul3x3f(mat1, D12, mat2); from33(mat2, &o.M2[sd * 3 + 1][pf * 9]);
    matmul3x3f(mat1, D13, mat2); from33(mat2, &o.M2[sd * 3 + 2][pf * 9]);
  }
  /* strike / normal / dip vectors, fault_info.cu:609-654 */
  mat3[0][0] = e1[0]; mat3[0][1] = e1[1]; mat3[0][2] = e1[2];
  mat3[1][0] = e2[0]; mat3[1][1] = e2[1]; mat3[1][2] = e2[2];
  mat3[2][0] = e3[0]; mat3[2][1] = e3[1]; mat3[2][2] = e3[2];
  invert3x3f(mat3);
  o.x_et[pf] = mat3[0][1]; o.y_et[pf] = mat3[1][1]; o.z_et[pf] = mat3[2][1];
  float vs1[3] = {mat3[0][1], mat3[1][1], mat3[2][1]};
  float nrm = norm3f(vs1);
  for (int a = 0; a < 3; a++) vs1[a] /= nrm;
  float vn[3] = {e1[0], e1[1], e1[2]};
  nrm = norm3f(vn);
  for (int a = 0; a < 3; a++) vn[a] /= nrm;
  float vs2[3];
  cross3f(vn, vs1, vs2);
  for (int a = 0; a < 3; a++) {
    o.vec_n[pf * 3 + a] = vn[a]; o.vec_s1[pf * 3 + a] = vs1[a]; o.vec_s2[pf * 3 + a] = vs2[a];
  }
  if (!top) return;
  /* free-surface row, fault_info.cu:656-845 (evaluated whatever the top
     boundary is, as upstream) */
  float A[3][3], Bm[3][3], Cm[3][3];
  float Vx2Vz1[3][3], Vy2Vz1[3][3], Vx2Vz2[3][3], Vy2Vz2[3][3];
  for (int sd = 0; sd < 2; sd++) {
    const float *D31 = Dst[sd * 6 + 3], *D32 = Dst[sd * 6 + 4], *D33 = Dst[sd * 6 + 5];
    for (int q = 0; q < 9; q++) { A[q / 3][q % 3] = D33[q]; Bm[q / 3][q % 3] = -D31[q]; Cm[q / 3][q % 3] = -D32[q]; }
    invert3x3f(A);
    matmul3x3f(A, Bm, sd == 0 ? Vx2Vz1 : Vx2Vz2);
    matmul3x3f(A, Cm, sd == 0 ? Vy2Vz1 : Vy2Vz2);
  }
  float vn2[3] = {e1[0], e1[1], e1[2]};
  float s1b[3] = {o.x_et[pf], o.y_et[pf], o.z_et[pf]};
  float s2b[3];
  cross3f(vn2, s1b, s2b);
  nrm = norm3f(s2b);
  for (int a = 0; a < 3; a++) s2b[a] /= nrm;
  const float g1_2 = 0.0f, g2_2 = 0.0f, g3_2 = 0.0f;
  float h1_3 = s2b[0] * e1[0] + s2b[1] * e1[1] + s2b[2] * e1[2];
  float h2_3 = s2b[0] * e2[0] + s2b[1] * e2[1] + s2b[2] * e2[2];
  float h3_3 = s2b[0] * e3[0] + s2b[1] * e3[1] + s2b[2] * e3[2];
  nrm = norm3f(vn2);
  h1_3 = h1_3 / nrm; h2_3 = h2_3 / nrm; h3_3 = h3_3 / nrm;
  float D11f[3][3], D12f[3][3], D13f[3][3];
  for (int ii = 0; ii < 3; ii++)
    for (int jj = 0; jj < 3; jj++) {
      const float To = s1b[ii] * vn2[jj];
      const float D11_12 = mu * g1_2 * To, D12_12 = mu * g2_2 * To, D13_12 = mu * g3_2 * To;
      const float To3 = s2b[ii] * vn2[jj];
      const float D11_13 = mu * h1_3 * To3, D12_13 = mu * h2_3 * To3, D13_13 = mu * h3_3 * To3;
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
  const size_t j9 = jf * 9;
  from33(To, &o.Fm[10][j9]);     /* matT1toVxf_Min */
  from33(P2M3f, &o.Fm[11][j9]);  /* matVytoVxf_Min */
  for (int ii = 0; ii < 3; ii++)
    for (int jj = 0; jj < 3; jj++) To[ii][jj] = Vx2f[ii][jj];
  invert3x3f(To);
  matmul3x3f(To, Vx1f, M2P1f);
  matmul3x3f(To, Vy1f, M2P2f);
  matmul3x3f(To, Vy2f, M2P3f);
  from33(To, &o.Fm[12][j9]);     /* matT1toVxf_Plus */
  from33(M2P3f, &o.Fm[13][j9]);  /* matVytoVxf_Plus */
  from33(Vx2Vz1, &o.Fm[0][j9]);
  from33(Vy2Vz1, &o.Fm[1][j9]);
  from33(Vx2Vz2, &o.Fm[2][j9]);
  from33(Vy2Vz2, &o.Fm[3][j9]);
  from33(P2M1f, &o.Fm[4][j9]);
  from33(P2M2f, &o.Fm[5][j9]);
  from33(P2M3f, &o.Fm[6][j9]);
  from33(M2P1f, &o.Fm[7][j9]);
  from33(M2P2f, &o.Fm[8][j9]);
  from33(M2P3f, &o.Fm[9][j9]);
}

}  // namespace cgfd_setup_math

#endif
