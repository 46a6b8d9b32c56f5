/*
 * k_vol_common.h -- per-point pieces shared by the volume kernels: interior
 * rhs, CFS-PML correction + auxiliary update, fault-column ownership test and
 * the RK stage update.
 */
#ifndef CGFD_K_VOL_COMMON_H
#define CGFD_K_VOL_COMMON_H

#include "cgfd_dev.h"

namespace cgfd {

/* component order Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy (wav_t.cu:56-106) */
enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TYZ, TXZ, TXY };

/* interior rhs at a point (sv_curv_col_el_iso_gpu.cu:349-386) */
DEVI void rhs_interior(const float *Dx, const float *Dy, const float *Dz, float xx, float xy, float xz, float ex,
                       float ey, float ez, float zx, float zy, float zz, float l, float m, float s, float *h) {
  float l2m = l + 2.0f * m;
  h[VX] = s * (xx * Dx[TXX] + xy * Dx[TXY] + xz * Dx[TXZ] + ex * Dy[TXX] + ey * Dy[TXY] + ez * Dy[TXZ] + zx * Dz[TXX] + zy * Dz[TXY] + zz * Dz[TXZ]);
  h[VY] = s * (xx * Dx[TXY] + xy * Dx[TYY] + xz * Dx[TYZ] + ex * Dy[TXY] + ey * Dy[TYY] + ez * Dy[TYZ] + zx * Dz[TXY] + zy * Dz[TYY] + zz * Dz[TYZ]);
  h[VZ] = s * (xx * Dx[TXZ] + xy * Dx[TYZ] + xz * Dx[TZZ] + ex * Dy[TXZ] + ey * Dy[TYZ] + ez * Dy[TZZ] + zx * Dz[TXZ] + zy * Dz[TYZ] + zz * Dz[TZZ]);
  h[TXX] = l2m * (xx * Dx[VX] + ex * Dy[VX] + zx * Dz[VX]) + l * (xy * Dx[VY] + ey * Dy[VY] + zy * Dz[VY] + xz * Dx[VZ] + ez * Dy[VZ] + zz * Dz[VZ]);
  h[TYY] = l2m * (xy * Dx[VY] + ey * Dy[VY] + zy * Dz[VY]) + l * (xx * Dx[VX] + ex * Dy[VX] + zx * Dz[VX] + xz * Dx[VZ] + ez * Dy[VZ] + zz * Dz[VZ]);
  h[TZZ] = l2m * (xz * Dx[VZ] + ez * Dy[VZ] + zz * Dz[VZ]) + l * (xx * Dx[VX] + ex * Dy[VX] + zx * Dz[VX] + xy * Dx[VY] + ey * Dy[VY] + zy * Dz[VY]);
  h[TXY] = m * (xy * Dx[VX] + xx * Dx[VY] + ey * Dy[VX] + ex * Dy[VY] + zy * Dz[VX] + zx * Dz[VY]);
  h[TXZ] = m * (xz * Dx[VX] + xx * Dx[VZ] + ez * Dy[VX] + ex * Dy[VZ] + zz * Dz[VX] + zx * Dz[VZ]);
  h[TYZ] = m * (xz * Dx[VY] + xy * Dx[VZ] + ez * Dy[VY] + ey * Dy[VZ] + zz * Dz[VY] + zy * Dz[VZ]);
}


/* The auxiliary variables of a CFS-PML face follow the wavefield's RK
   bookkeeping (rk_point): stage s reads level pre, la, lb or lc and writes
   the next one; stage 3 writes t_{n+1} over la. */
DEVI const float *aux_input(const PmlFace &f, int istage) {
  return istage == 0 ? f.pre : (istage == 1 ? f.la : (istage == 2 ? f.lb : f.lc));
}

/* av: the stage input's values at slab point ia (already read), ar: the
   auxiliary rhs */
DEVI void aux_update(const PmlFace &f, const VolArgs &a, long poff, long ia, const float *av, const float *ar) {
  const float *pre = f.pre + poff;
  if (a.istage < 3) {
    float *out = (a.istage == 0 ? f.la : (a.istage == 1 ? f.lb : f.lc)) + poff;
#pragma unroll
    for (int c = 0; c < 9; c++) out[c * f.n + ia] = (a.istage == 0 ? av[c] : pre[c * f.n + ia]) + a.ca * ar[c];
  } else {
    float *la = f.la + poff;
    const float *lb = f.lb + poff;
#pragma unroll
    for (int c = 0; c < 9; c++) {
      const long q = c * f.n + ia;
      la[q] = rk_end(pre[q], la[q], lb[q], av[c], a.cb * ar[c]);
    }
  }
}

/* CFS-PML correction of every face containing (i,j,k) plus the auxiliary
   RK update (sv_curv_col_el_iso_gpu.cu:916-1288, drv_rk_curv_col.cu:368-567) */
DEVI void pml_point(const VolArgs &a, int i, int j, int k, long p, const float *Dx, const float *Dy, const float *Dz,
                    float l, float m, float s, float *h, long poff = 0) {
  const long sc = a.g.sicmp;
#pragma unroll 1
  for (int dim = 0; dim < 3; dim++) {
#pragma unroll 1
    for (int side = 0; side < 2; side++) {
      const PmlFace &f = a.pml[dim][side];
      if (!f.on) continue;
      if (i < f.i1 || i > f.i2 || j < f.j1 || j > f.j2 || k < f.k1 || k > f.k2) continue;
      long ia = (long)(k - f.k1) * ((long)f.bn * f.an) + (long)(j - f.j1) * f.an + (i - f.ax0);
      int ic = dim == 0 ? i - f.i1 : (dim == 1 ? j - f.j1 : k - f.k1);
      float cD = f.D[ic], cA = f.A[ic], cB = f.B[ic];
      float cB1 = cB - 1.0f;
      const float *D = dim == 0 ? Dx : (dim == 1 ? Dy : Dz);
      const float *mr = a.met + (long)(1 + 3 * dim) * sc;
      float e1 = mr[p], e2 = mr[p + sc], e3 = mr[p + 2 * sc];
      float l2m = l + 2.0f * m;
      float r[9];
      r[VX] = s * (e1 * D[TXX] + e2 * D[TXY] + e3 * D[TXZ]);
      r[VY] = s * (e1 * D[TXY] + e2 * D[TYY] + e3 * D[TYZ]);
      r[VZ] = s * (e1 * D[TXZ] + e2 * D[TYZ] + e3 * D[TZZ]);
      r[TXX] = l2m * e1 * D[VX] + l * e2 * D[VY] + l * e3 * D[VZ];
      r[TYY] = l * e1 * D[VX] + l2m * e2 * D[VY] + l * e3 * D[VZ];
      r[TZZ] = l * e1 * D[VX] + l * e2 * D[VY] + l2m * e3 * D[VZ];
      r[TXY] = m * (e2 * D[VX] + e1 * D[VY]);
      r[TXZ] = m * (e3 * D[VX] + e1 * D[VZ]);
      r[TYZ] = m * (e3 * D[VY] + e2 * D[VZ]);
      const float *auxc = aux_input(f, a.istage) + poff;
      float av[9], ar[9];
#pragma unroll
      for (int c = 0; c < 9; c++) av[c] = auxc[c * f.n + ia];
#pragma unroll
      for (int c = 0; c < 9; c++) {
        h[c] += cB1 * r[c] - cB * av[c];
        ar[c] = cD * r[c] - cA * av[c];
      }
      if (dim < 2 && a.isfree && k == a.g.nk2) {
        /* free-surface contribution on x/y faces (:1000-1060, :1146-1205) */
        long ij = ((long)i + (long)j * a.g.siy) * 9;
        const float *X = (dim == 0 ? a.matVx2Vz : a.matVy2Vz) + ij;
        float q0 = X[0] * D[VX] + X[1] * D[VY] + X[2] * D[VZ];
        float q1 = X[3] * D[VX] + X[4] * D[VY] + X[5] * D[VZ];
        float q2 = X[6] * D[VX] + X[7] * D[VY] + X[8] * D[VZ];
        float zx = a.met[7 * sc + p], zy = a.met[8 * sc + p], zz = a.met[9 * sc + p];
        float qT[9];
        qT[TXX] = l2m * (zx * q0) + l * (zy * q1 + zz * q2);
        qT[TYY] = l2m * (zy * q1) + l * (zx * q0 + zz * q2);
        qT[TZZ] = l2m * (zz * q2) + l * (zx * q0 + zy * q1);
        qT[TXY] = m * (zy * q0 + zx * q1);
        qT[TXZ] = m * (zz * q0 + zx * q2);
        qT[TYZ] = m * (zz * q1 + zy * q2);
#pragma unroll
        for (int c = 3; c < 9; c++) {
          h[c] = (float)(h[c] + (cB - 1.0) * qT[c]);
          ar[c] += cD * qT[c];
        }
      }
      /* auxiliary RK update; the slab point is owned by this thread only */
      aux_update(f, a, poff, ia, av, ar);
    }
  }
}

DEVI bool fault_column(const FaultSkip &fs, int i, long pf) {
#pragma unroll 1
  for (int q = 0; q < fs.nfault; q++) {
    int d = i - fs.i0[q];
    if (d >= -3 && d <= 3 && fs.united[q][pf] == 0) return true;
  }
  return false;
}

/* RK stage update of the nine wavefield components at p: stages 0..2 write
   the next stage input pre + a_s dt h (wav_update, wav_t.cu:218-226),
   stage 3 the level at t_{n+1} over level A (rk_end) */
DEVI void rk_point(const VolArgs &a, long p, const float *h, long woff = 0) {
  const long sc = a.g.sicmp;
  p += woff;
  if (a.istage < 3) {
#pragma unroll
    for (int c = 0; c < 9; c++) a.tnew[c * sc + p] = a.pre[c * sc + p] + a.ca * h[c];
  } else {
#pragma unroll
    for (int c = 0; c < 9; c++) {
      const long q = c * sc + p;
      a.end[q] = rk_end(a.pre[q], a.end[q], a.lb[q], a.cur[q], a.cb * h[c]);
    }
  }
}

/* rk_point with 32-bit byte offsets (o = 4 p) from the scenario's levels;
   the stage input's own value at stage 3 comes from the cache (the stencil
   has just read it) */
DEVI void rk_point32(const VolArgs &a, unsigned o, const float *h, long woff = 0) {
  const long sc = a.g.sicmp;
  const float *pre = a.pre + woff;
  if (a.istage < 3) {
    float *tnew = a.tnew + woff;
#pragma unroll
    for (int c = 0; c < 9; c++) st32nt(tnew + c * sc, o, ld32nt(pre + c * sc, o) + a.ca * h[c]);
  } else {
    float *end = a.end + woff;
    const float *lb = a.lb + woff, *cur = a.cur + woff;
#pragma unroll
    for (int c = 0; c < 9; c++) {
      const float v = rk_end(ld32nt(pre + c * sc, o), ld32nt(end + c * sc, o), ld32nt(lb + c * sc, o),
                             ld32(cur + c * sc, o), a.cb * h[c]);
      st32nt(end + c * sc, o, v);
    }
  }
}

/* rhs + PML + RK update of one point from its 27 derivatives */
DEVI void vol_point_update(const VolArgs &a, int i, int j, int k, long p, const float *Dx, const float *Dy,
                           const float *Dz) {
  const long sc = a.g.sicmp;
  const float *M = a.met;
  float xx = M[sc + p], xy = M[2 * sc + p], xz = M[3 * sc + p];
  float ex = M[4 * sc + p], ey = M[5 * sc + p], ez = M[6 * sc + p];
  float zx = M[7 * sc + p], zy = M[8 * sc + p], zz = M[9 * sc + p];
  float l = a.lam[p], m = a.mu[p], s = a.slw[p];
  float h[9];
  rhs_interior(Dx, Dy, Dz, xx, xy, xz, ex, ey, ez, zx, zy, zz, l, m, s, h);
  if (a.pml_on) pml_point(a, i, j, k, p, Dx, Dy, Dz, l, m, s, h);
  if (a.fs.nfault > 0 && fault_column(a.fs, i, j + (long)k * a.g.ny)) return;
  rk_point(a, p, h);
}

}  // namespace cgfd

#endif
