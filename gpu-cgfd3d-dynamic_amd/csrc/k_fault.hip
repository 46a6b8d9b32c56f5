/*
 * k_fault.hip -- fault-plane kernels of one RK stage (one thread per fault
 * node; grid z = fault id).
 *
 *  k_wave2fault : welded nodes copy the continuum state onto both split nodes
 *                 (wave2fault_gpu, transform.cu:100-212).
 *  k_trial      : trial traction + linear slip-weakening friction
 *                 (trial_slipweakening_gpu, trial_slipweakening.cu:91-388),
 *                 the stage accumulation of Tn/Ts1/Ts2
 *                 (fault_var_stage_update_gpu, fault_wav_t.cu:290-316), and
 *                 the split-node average written back onto the stage input
 *                 (fault2wave_gpu, transform.cu:292-376).
 *  k_fault_cols<IDIR> : velocity and stress rhs (fault_velo, fault_gpu.cu:
 *                 191-324; fault_stress_F/B, :431-1879) on the six columns
 *                 around the plane, fused with their RK update (one lane per
 *                 node and column).
 *  k_fault_nodes<IDIR> : split-node velocity and traction rates, RK update of
 *                 the 18 split-node fields (fault_wav_update, fault_wav_t.cu:
 *                 318-366), fault2wave onto the stage output, and at the last
 *                 stage the per-step accumulators (fault_var_update_gpu,
 *                 fault_wav_t.cu:197-264).
 * The F and B stress variants are one template: with s the split side whose
 * x-derivative is the one-sided difference into the domain (plus for a
 * forward x-operator) and o the other side, the formulas of the two
 * reference kernels coincide.
 */
#include "cgfd_dev.h"
#include "cgfd_kernels.h"

namespace cgfd {

enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TYZ, TXZ, TXY };

/* split-node state: component c (Vx,Vy,Vz,T2x,T2y,T2z,T3x,T3y,T3z), side m */
DEVI long fidx(int c, int m, long ss) { return (long)(2 * c + m) * ss; }

/* metric matrix [xi eta zeta] as columns (transform.cu:147-149) */
DEVI void metric_cols(const float *met, long sc, long p, float M[3][3]) {
  M[0][0] = met[sc + p];     M[0][1] = met[4 * sc + p]; M[0][2] = met[7 * sc + p];
  M[1][0] = met[2 * sc + p]; M[1][1] = met[5 * sc + p]; M[1][2] = met[8 * sc + p];
  M[2][0] = met[3 * sc + p]; M[2][1] = met[6 * sc + p]; M[2][2] = met[9 * sc + p];
}

/* the same from the fault's copy of the metric on its column, [10][sslice] */
DEVI void metric_cols_f(const float *fm, long ss, long pf, float M[3][3]) {
  M[0][0] = fm[ss + pf];     M[0][1] = fm[4 * ss + pf]; M[0][2] = fm[7 * ss + pf];
  M[1][0] = fm[2 * ss + pf]; M[1][1] = fm[5 * ss + pf]; M[1][2] = fm[8 * ss + pf];
  M[2][0] = fm[3 * ss + pf]; M[2][1] = fm[6 * ss + pf]; M[2][2] = fm[9 * ss + pf];
}

__global__ void __launch_bounds__(256) k_fault_metric(Grid g, const float *__restrict__ met, int i0, float *fmet) {
  const int j = blockIdx.x * 256 + (int)threadIdx.x, k = blockIdx.y;
  if (j >= g.ny) return;
  const long p = i0 + (long)j * g.siy + (long)k * g.siz, pf = j + (long)k * g.ny;
#pragma unroll
  for (int c = 0; c < 10; c++) fmet[c * g.sslice + pf] = met[c * g.sicmp + p];
}

void launch_fault_metric(const Grid &g, const float *met, int i0, float *fmet, hipStream_t st) {
  hipLaunchKernelGGL(k_fault_metric, dim3((unsigned)((g.ny + 255) / 256), (unsigned)g.nz), dim3(256), 0, st, g, met,
                     i0, fmet);
}

/* ---------------------------------------------------------------------- */
DEVI void k_wave2fault_body(const FaultArgs &a, int id) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const FaultDev F = fdev(a, id);
  const long ss = a.g.sslice, sc = a.g.sicmp;
  const long pf = (iy + 3) + (long)(iz + 3) * a.g.ny;
  if (F.united[pf] != 1) return;
  const long p = F.i0 + (long)(iy + 3) * a.g.siy + (long)(iz + 3) * a.g.siz;
  const float *W = a.w_cur;
  float M[3][3], S[3][3], Tr[3][3];
  metric_cols_f(F.fmet, ss, pf, M);
  float jac = F.fmet[pf];
  float txx = W[TXX * sc + p], tyy = W[TYY * sc + p], tzz = W[TZZ * sc + p];
  float tyz = W[TYZ * sc + p], txz = W[TXZ * sc + p], txy = W[TXY * sc + p];
  S[0][0] = txx; S[0][1] = txy; S[0][2] = txz;
  S[1][0] = txy; S[1][1] = tyy; S[1][2] = tyz;
  S[2][0] = txz; S[2][1] = tyz; S[2][2] = tzz;
  mm33(S, M, Tr);
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) Tr[r][c] *= jac;
#pragma unroll
  for (int c = 0; c < 3; c++) F.T1[(long)c * 7 * ss + 3 * ss + pf] = Tr[c][0];
  float *fc = a.f_cur + (long)id * 18 * ss;
  float v[3] = {W[VX * sc + p], W[VY * sc + p], W[VZ * sc + p]};
#pragma unroll
  for (int m = 0; m < 2; m++) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      fc[fidx(c, m, ss) + pf] = v[c];
      fc[fidx(3 + c, m, ss) + pf] = Tr[c][1];
      fc[fidx(6 + c, m, ss) + pf] = Tr[c][2];
    }
  }
}

__global__ void __launch_bounds__(256) k_wave2fault(FaultArgs a, BatArgs b) {
  const int id = (int)blockIdx.z % a.nfault;
  k_wave2fault_body(fault_at(a, b, b.act[blockIdx.z / a.nfault]), id);
}

/* split-node average back onto the wavefield column i0 (transform.cu:331-372),
   writing sigma_xz/sigma_yz into swapped arrays unless fix_q1 (the call at
   transform.cu:274-276 passes Tyz, Txz to parameters named Txz, Tyz) */
DEVI void fault2wave_node(const FaultArgs &a, const FaultDev &F, long pf, long p, float *w, const float *fl) {
  const long sc = a.g.sicmp, ss = a.g.sslice;
  float M[3][3], Tr[3][3], S[3][3];
  metric_cols_f(F.fmet, ss, pf, M);
  float rj = (float)(1.0 / F.fmet[pf]);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    Tr[c][0] = F.T1[(long)c * 7 * ss + 3 * ss + pf];
    Tr[c][1] = (float)((fl[fidx(3 + c, 0, ss) + pf] + fl[fidx(3 + c, 1, ss) + pf]) * 0.5);
    Tr[c][2] = (float)((fl[fidx(6 + c, 0, ss) + pf] + fl[fidx(6 + c, 1, ss) + pf]) * 0.5);
  }
  inv3(M);
  mm33(Tr, M, S);
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) S[r][c] *= rj;
#pragma unroll
  for (int c = 0; c < 3; c++) w[c * sc + p] = (float)((fl[fidx(c, 0, ss) + pf] + fl[fidx(c, 1, ss) + pf]) * 0.5);
  w[TXX * sc + p] = S[0][0];
  w[TYY * sc + p] = S[1][1];
  w[TZZ * sc + p] = S[2][2];
  w[TXY * sc + p] = (float)((S[0][1] + S[1][0]) * 0.5);
  float sxz = (float)((S[0][2] + S[2][0]) * 0.5);
  float syz = (float)((S[1][2] + S[2][1]) * 0.5);
  w[(a.fix_q1 ? TXZ : TYZ) * sc + p] = sxz;
  w[(a.fix_q1 ? TYZ : TXZ) * sc + p] = syz;
}

/* z-derivative of a split-node traction with the odd image across the free
   surface (trial_slipweakening.cu:202-227) */
DEVI float dz_free_image(const float *col, long ny, int n_free, int kdir) {
  float v[7];
#pragma unroll
  for (int l = -3; l <= 3; l++) v[l + 3] = col[l * ny];
  v[n_free] = 0.0f;
  for (int l = n_free + 1; l < 7; l++) v[l] = -v[2 * n_free - l];
  return vecc(v + 3, kdir);
}

/* ---------------------------------------------------------------------- */
/* Traction vectors J (grad xi . sigma) on the three columns each side of
   the plane (the image columns of trial_slipweakening.cu:151-165), written
   into the T1 slots 0-2 and 4-6. Six lanes per node (one per column), ten
   nodes per wavefront, so a node's loads are adjacent addresses. */
constexpr int kTrialNodes = 10;

DEVI void k_trial_t1_body(const FaultArgs &a, int id) {
  const int lane = threadIdx.x;
  if (lane >= 6 * kTrialNodes) return;
  const int iy = blockIdx.x * kTrialNodes + lane / 6;
  const int cl = lane % 6;
  const int iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const FaultDev F = fdev(a, id);
  const long ss = a.g.sslice, sc = a.g.sicmp, ny = a.g.ny;
  const long pf = (iy + 3) + (long)(iz + 3) * ny;
  if (F.united[pf] != 0) return;
  const int m = cl / 3, l = cl % 3 + 1;
  const float *W = a.w_cur;
  const float *met = a.met;
  const long q = F.i0 + (2 * m - 1) * l + (long)(iy + 3) * a.g.siy + (long)(iz + 3) * a.g.siz;
  const float x1 = met[sc + q], jl = met[q];
  const float txx = W[TXX * sc + q], txz = W[TXZ * sc + q], txy = W[TXY * sc + q];
  const long slot = 3 + (2 * m - 1) * l;
  if (a.vcol) {
    /* vertical columns: xi_y and xi_z vanish (see the volume kernels) */
    F.T1[slot * ss + pf] = jl * (x1 * txx);
    F.T1[7 * ss + slot * ss + pf] = jl * (x1 * txy);
    F.T1[14 * ss + slot * ss + pf] = jl * (x1 * txz);
    return;
  }
  const float x2 = met[2 * sc + q], x3 = met[3 * sc + q];
  const float tyy = W[TYY * sc + q], tzz = W[TZZ * sc + q], tyz = W[TYZ * sc + q];
  F.T1[slot * ss + pf] = jl * (x1 * txx + x2 * txy + x3 * txz);
  F.T1[7 * ss + slot * ss + pf] = jl * (x1 * txy + x2 * tyy + x3 * tyz);
  F.T1[14 * ss + slot * ss + pf] = jl * (x1 * txz + x2 * tyz + x3 * tzz);
}

__global__ void __launch_bounds__(256) k_trial_t1(FaultArgs a, BatArgs b) {
  const int id = (int)blockIdx.z % a.nfault;
  k_trial_t1_body(fault_at(a, b, b.act[blockIdx.z / a.nfault]), id);
}

/* ---------------------------------------------------------------------- */
DEVI void k_trial_body(const FaultArgs &a, int id) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const FaultDev F = fdev(a, id);
  const long ss = a.g.sslice, sc = a.g.sicmp, ny = a.g.ny;
  const long pf = (iy + 3) + (long)(iz + 3) * ny;
  if (F.united[pf] != 0) return;
  const int jdir = a.jdir, kdir = a.kdir;
  const float *met = a.met;
  const long row = (long)(iy + 3) * a.g.siy + (long)(iz + 3) * a.g.siz;
  const long p0 = F.i0 + row;
  const float *fc = a.f_cur + (long)id * 18 * ss;
  const float *fp = a.f_pre + (long)id * 18 * ss;
  const int km = a.g.nk - (iz + 1);
  const int n_free = km + 3;
  const int ry = F.rupy[pf], rz = F.rupz[pf];

  float Mrho[2], R[2][3];
#pragma unroll
  for (int m = 0; m < 2; m++) {
    float jac0 = F.fmet[pf];
    float rho = F.rho_f[m * ss + pf];
    Mrho[m] = (float)(0.5 * jac0 * rho);
    float t1[3][3];  /* [comp][l-1]: the traction columns k_trial_t1 wrote */
#pragma unroll
    for (int l = 1; l <= 3; l++) {
      const long slot = 3 + (2 * m - 1) * l;
#pragma unroll
      for (int c = 0; c < 3; c++) t1[c][l - 1] = F.T1[(long)c * 7 * ss + slot * ss + pf];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float *p2 = fc + fidx(3 + c, m, ss) + pf;
      const float *p3 = fc + fidx(6 + c, m, ss) + pf;
      float dy = 0.0f, dz = 0.0f;
      if (ry == 1) dy = mac22(p2, 1, jdir);
      if (ry == 0) dy = macdrp(p2, 1, jdir);
      if (rz == 1) dz = mac22(p3, ny, kdir);
      if (rz == 0) dz = macdrp(p3, ny, kdir);
      if (a.isfree == 1 && km <= 3) dz = dz_free_image(p3, ny, n_free, kdir);
      /* image sums (trial_slipweakening.cu:230-256): slot 2,1,0 minus / 4,5,6 plus */
      float rr = (float)(kA1 * t1[c][0] + kA2 * t1[c][1] + kA3 * t1[c][2]);
      R[m][c] = (float)(0.5 * ((2 * m - 1) * rr + (dy + dz)));
    }
  }
  float Trial[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float dV = fp[fidx(c, 1, ss) + pf] - fp[fidx(c, 0, ss) + pf];
    Trial[c] = (float)(2.0 * (Mrho[0] * Mrho[1] * dV / a.dt + Mrho[0] * R[1][c] - Mrho[1] * R[0][c]) /
                       (kA0 * (Mrho[0] + Mrho[1])));
  }
  if (a.tape) {
    const int slot = *a.it_dev - a.tape_it0;
    if (slot >= 0 && slot < a.tape_nt) {
      float *tp = a.tape + ((long)id * a.tape_nt + slot) * kTapeFloats * ss + pf;
#pragma unroll
      for (int c = 0; c < 3; c++) tp[(3 * a.istage + c) * ss] = Trial[c];
      if (a.istage == 0) tp[12 * ss] = F.out[6 * ss + pf];
    }
  }
  float vs1[3] = {F.vs[pf], F.vs[ss + pf], F.vs[2 * ss + pf]};
  float vs2[3] = {F.vs[3 * ss + pf], F.vs[4 * ss + pf], F.vs[5 * ss + pf]};
  float vn[3] = {F.fmet[ss + pf], F.fmet[2 * ss + pf], F.fmet[3 * ss + pf]};
  float vn0 = norm3(vn);
  float jacvec = F.fmet[pf] * vn0;
#pragma unroll
  for (int c = 0; c < 3; c++) vn[c] /= vn0;
  float T0[3] = {F.T0[pf], F.T0[ss + pf], F.T0[2 * ss + pf]};
  float TL[3], Ts[3];
#pragma unroll
  for (int c = 0; c < 3; c++) TL[c] = Trial[c] / jacvec + T0[c];
  float Tn0 = dot3(TL, vn);
#pragma unroll
  for (int c = 0; c < 3; c++) Ts[c] = TL[c] - vn[c] * Tn0;
  float Ts0 = norm3(Ts);
  int ifchange = 0;
  float Tau_n, Tau_s;
  if (Tn0 >= 0.0f) { Tau_n = 0.0f; ifchange = 1; }  /* no opening */
  else Tau_n = Tn0;
  float mus = F.fric[pf], mud = F.fric[ss + pf], Dc = F.fric[2 * ss + pf], C0 = F.fric[3 * ss + pf];
  float Slip = F.out[6 * ss + pf];
  float fr = (Slip <= Dc) ? mus - (mus - mud) * Slip / Dc : mud;
  float Tau_c = -fr * Tau_n + C0;
  if (Ts0 >= Tau_c) { Tau_s = Tau_c; F.flag_rup[pf] = 1; ifchange = 1; }
  else { Tau_s = Ts0; F.flag_rup[pf] = 0; }
  float Tau[3];
  if (ifchange == 1) {
    if (fabsf(Ts0) < 1e-1) {
#pragma unroll
      for (int c = 0; c < 3; c++) Tau[c] = Tau_n * vn[c];
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) Tau[c] = Tau_n * vn[c] + (Tau_s / Ts0) * Ts[c];
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) Tau[c] = TL[c];
  }
#pragma unroll
  for (int c = 0; c < 3; c++)
    F.T1[(long)c * 7 * ss + 3 * ss + pf] = (ifchange == 1) ? (Tau[c] - T0[c]) * jacvec : Trial[c];
  float tTs1 = dot3(Tau, vs1), tTs2 = dot3(Tau, vs2), tTn = Tau_n;
  if (a.istage == 0) {
    F.out[pf] = a.rkb * tTn; F.out[ss + pf] = a.rkb * tTs1; F.out[2 * ss + pf] = a.rkb * tTs2;
  } else {
    F.out[pf] += a.rkb * tTn; F.out[ss + pf] += a.rkb * tTs1; F.out[2 * ss + pf] += a.rkb * tTs2;
  }
  /* split-node average onto the stage input (transform: fault2wave_onestage(w_cur)) */
  fault2wave_node(a, F, pf, p0, a.w_cur, fc);
}

__global__ void __launch_bounds__(128) k_trial(FaultArgs a, BatArgs b) {
  const int id = (int)blockIdx.z % a.nfault;
  k_trial_body(fault_at(a, b, b.act[blockIdx.z / a.nfault]), id);
}

/* Hooke's law on one column point (sv_curv_col_el_iso_fault_gpu.cu:1126-1154) */
DEVI void hooke(const float *met, long sc, long p, bool vc, float l, float m, const float *Dx, const float *Dy,
                const float *Dz, float *h) {
  float l2m = (float)(l + 2.0 * m);
  /* vc: vertical columns, xi_y, xi_z, eta_x and eta_z vanish (the volume
     kernels' five-term form) and are not read */
  float xx = met[sc + p], xy = vc ? 0.0f : met[2 * sc + p], xz = vc ? 0.0f : met[3 * sc + p];
  float ex = vc ? 0.0f : met[4 * sc + p], ey = met[5 * sc + p], ez = vc ? 0.0f : met[6 * sc + p];
  float zx = met[7 * sc + p], zy = met[8 * sc + p], zz = met[9 * sc + p];
  h[0] = l2m * (xx * Dx[0] + ex * Dy[0] + zx * Dz[0]) + l * (xy * Dx[1] + ey * Dy[1] + zy * Dz[1] + xz * Dx[2] + ez * Dy[2] + zz * Dz[2]);
  h[1] = l2m * (xy * Dx[1] + ey * Dy[1] + zy * Dz[1]) + l * (xx * Dx[0] + ex * Dy[0] + zx * Dz[0] + xz * Dx[2] + ez * Dy[2] + zz * Dz[2]);
  h[2] = l2m * (xz * Dx[2] + ez * Dy[2] + zz * Dz[2]) + l * (xx * Dx[0] + ex * Dy[0] + zx * Dz[0] + xy * Dx[1] + ey * Dy[1] + zy * Dz[1]);
  h[5] = m * (xy * Dx[0] + xx * Dx[1] + ey * Dy[0] + ex * Dy[1] + zy * Dz[0] + zx * Dz[1]); /* Txy */
  h[4] = m * (xz * Dx[0] + xx * Dx[2] + ez * Dy[0] + ex * Dy[2] + zz * Dz[0] + zx * Dz[2]); /* Txz */
  h[3] = m * (xz * Dx[1] + xy * Dx[2] + ez * Dy[1] + ey * Dy[2] + zz * Dz[1] + zy * Dz[2]); /* Tyz */
}

/* RK update of wavefield components [c0, c0+nc) at p with rhs h[], as the
   volume kernels do it (rk_point) */
DEVI void rk_cols(const FaultArgs &a, long p, int c0, int nc, const float *h) {
  const long sc = a.g.sicmp;
  for (int q = 0; q < nc; q++) {
    long o = (long)(c0 + q) * sc + p;
    if (a.istage < 3) a.w_tnew[o] = a.w_pre[o] + a.ca * h[q];
    else a.w_end[o] = rk_end(a.w_pre[o], a.w_end[o], a.w_b[o], a.w_cur[o], a.cb * h[q]);
  }
}

/* ---------------------------------------------------------------------- */
/* The six wavefield columns i0-3..i0-1, i0+1..i0+3 of a split node: velocity
   rhs (fault_gpu.cu:191-324) and stress rhs by Hooke's law (:1065-1155 and
   :1694-1790) with the x differences next to the plane taken against the
   split-node velocity (:518-543, :1020-1063, :1248-1274, :1694-1728), fused
   with the RK update of the nine components. Lane layout: six lanes per
   node, one per column, ten nodes (consecutive j) per wavefront, so the
   loads of a node's x taps are adjacent addresses. */
constexpr int kColNodes = 10;  /* nodes per wavefront in the column kernel */

/* Traction vectors J (grad xi_d . T) a block of kColNodes nodes x kColRows
   rows needs, computed once each into LDS: along x at the 13 columns
   i0-6..i0+6, along y at the 6 columns over rows j-3..j+12, along z at the 6
   columns over planes k-3..k+kColRows+2 (the union of the 7-point stencils
   of the block's lanes, each of which would otherwise evaluate 21 of them).
   Eight rows per block: the z halo planes are 6 per 8 rows instead of 6 per
   4, and the block's traction vectors take 32 KB of LDS. */
constexpr int kColRows = 8;                                       /* blockDim.y */
constexpr int kTvX = 13 * kColNodes * kColRows;                  /* 1040 */
constexpr int kTvY = 6 * (kColNodes + 6) * kColRows;             /* 768 */
constexpr int kTvZ = 6 * kColNodes * (kColRows + 6);             /* 840 */
constexpr int kTv = kTvX + kTvY + kTvZ;

template <int IDIR>
DEVI void k_fault_cols_body(const FaultArgs &a, int id, int j0, int k0) {
  __shared__ float tv[3][kTv];
  const FaultDev F = fdev(a, id);
  const long ss = a.g.sslice, sc = a.g.sicmp, ny = a.g.ny, sy = a.g.siy, sz = a.g.siz;
  const int i0 = F.i0;
  const float *met = a.met;
  const float *W = a.w_cur;
  const float *Txx = W + TXX * sc, *Tyy = W + TYY * sc, *Tzz = W + TZZ * sc;
  const float *Tyz = W + TYZ * sc, *Txz = W + TXZ * sc, *Txy = W + TXY * sc;
  {
    /* blocks whose nodes are all welded (or past the grid) have no work */
    const int ln = threadIdx.x, jb = j0 + ln / 6, kb = k0 + (int)threadIdx.y;
    const bool mine = ln < 6 * kColNodes && jb < a.g.nj && kb < a.g.nk &&
                      F.united[(jb + 3) + (long)(kb + 3) * ny] == 0;
    if (!__syncthreads_or(mine)) return;
  }
  {
    const int tid = threadIdx.x + 64 * threadIdx.y;
    for (int v = tid; v < kTv; v += 64 * kColRows) {
      int x, jg, kg, d;  /* ghost-inclusive point and metric row */
      if (v < kTvX) {
        const int xi = v % 13, r = v / 13;
        x = i0 - 6 + xi; jg = j0 + r % kColNodes + 3; kg = k0 + r / kColNodes + 3; d = 0;
      } else if (v < kTvX + kTvY) {
        const int u = v - kTvX, col = u % 6, r = u / 6;
        x = i0 + (col < 3 ? col - 3 : col - 2); jg = j0 - 3 + r % (kColNodes + 6) + 3;
        kg = k0 + r / (kColNodes + 6) + 3; d = 1;
      } else {
        const int u = v - kTvX - kTvY, col = u % 6, r = u / 6;
        x = i0 + (col < 3 ? col - 3 : col - 2); jg = j0 + r % kColNodes + 3;
        kg = k0 - 3 + r / kColNodes + 3; d = 2;
      }
      float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
      if (x >= 0 && x < a.g.nx && jg >= 0 && jg < a.g.ny && kg >= 0 && kg < a.g.nz) {
        const long q = x + (long)jg * sy + (long)kg * sz;
        const float jq = met[q];
        if (a.vcol && d == 0) {         /* vertical columns: xi_y, xi_z vanish */
          const float a1 = met[sc + q];
          o0 = jq * (a1 * Txx[q]);
          o1 = jq * (a1 * Txy[q]);
          o2 = jq * (a1 * Txz[q]);
        } else if (a.vcol && d == 1) {  /* eta_x, eta_z vanish */
          const float a2 = met[5 * sc + q];
          o0 = jq * (a2 * Txy[q]);
          o1 = jq * (a2 * Tyy[q]);
          o2 = jq * (a2 * Tyz[q]);
        } else {
          const float a1 = met[(1 + 3 * d) * sc + q], a2 = met[(2 + 3 * d) * sc + q],
                      a3 = met[(3 + 3 * d) * sc + q];
          o0 = jq * (a1 * Txx[q] + a2 * Txy[q] + a3 * Txz[q]);
          o1 = jq * (a1 * Txy[q] + a2 * Tyy[q] + a3 * Tyz[q]);
          o2 = jq * (a1 * Txz[q] + a2 * Tyz[q] + a3 * Tzz[q]);
        }
      }
      tv[0][v] = o0;
      tv[1][v] = o1;
      tv[2][v] = o2;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x;
  if (lane >= 6 * kColNodes) return;
  const int jn = lane / 6, kr = threadIdx.y;
  const int iy = j0 + jn;
  const int cl = lane % 6;
  const int iz = k0 + kr;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const long pf = (iy + 3) + (long)(iz + 3) * ny;
  if (F.united[pf] != 0) return;
  const int jdir = a.jdir, kdir = a.kdir;
  const int mm = cl < 3 ? cl - 3 : cl - 2;  /* column offset from the plane */
  const int i = i0 + mm;
  const int n = -mm;
  const long row = (long)(iy + 3) * sy + (long)(iz + 3) * sz;
  const long p = i + row;
  const int km = a.g.nk - (iz + 1);
  const int n_free = km + 3;
  const bool surf = (a.isfree == 1);

  /* ---- velocity rhs ---- */
  float dvx[3], dvy[3], dvz[3];
  {
    float v[3][7];
#pragma unroll
    for (int l = -3; l <= 3; l++) {
      const int e = mm + l + 6 + 13 * (jn + kColNodes * kr);
#pragma unroll
      for (int c = 0; c < 3; c++) v[c][l + 3] = tv[c][e];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float *u = v[c];
      /* the tap on the plane is the fault traction; image beyond it
         (fault_gpu.cu:266-291) */
      const float tc = F.T1[(long)c * 7 * ss + 3 * ss + pf];
#pragma unroll
      for (int l = 0; l < 7; l++)
        if (l == n + 3) u[l] = tc;
      if (n == 2) { u[6] = (float)(2.0 * u[5] - u[4]); }
      if (n == 1) { u[5] = (float)(2.0 * u[4] - u[3]); u[6] = (float)(2.0 * u[4] - u[2]); }
      if (n == -1) { u[0] = (float)(2.0 * u[2] - u[4]); u[1] = (float)(2.0 * u[2] - u[3]); }
      if (n == -2) { u[0] = (float)(2.0 * u[1] - u[2]); }
      dvx[c] = vecc(u + 3, IDIR);
    }
  }
  {
    float v[3][7];
#pragma unroll
    for (int l = -3; l <= 3; l++) {
      const int e = kTvX + cl + 6 * (jn + 3 + l + (kColNodes + 6) * kr);
#pragma unroll
      for (int c = 0; c < 3; c++) v[c][l + 3] = tv[c][e];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) dvy[c] = vecc(v[c] + 3, jdir);
  }
  {
    float v[3][7];
#pragma unroll
    for (int l = -3; l <= 3; l++) {
      const int e = kTvX + kTvY + cl + 6 * (jn + kColNodes * (kr + 3 + l));
#pragma unroll
      for (int c = 0; c < 3; c++) v[c][l + 3] = tv[c][e];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float *u = v[c];
      if (surf && km <= 3) {
        u[n_free] = 0.0f;
        for (int l = n_free + 1; l < 7; l++) u[l] = -u[2 * n_free - l];
      }
      dvz[c] = vecc(u + 3, kdir);
    }
  }
  float h[9];
  const float rrhojac = a.slw[p] / met[p];
#pragma unroll
  for (int c = 0; c < 3; c++) h[c] = (dvx[c] + dvy[c] + dvz[c]) * rrhojac;
  __asm__ volatile("" : "+v"(h[0]), "+v"(h[1]), "+v"(h[2]) : : "memory");

  /* ---- stress rhs ---- */
  const float *V[3] = {W, W + sc, W + 2 * sc};
  const float *fc = a.f_cur + (long)id * 18 * ss;
  float dx[3], dy[3], dz[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float *Vc = V[c] + row;
    if (IDIR == 1) {
      if (mm == 3) dx[c] = macdrp_f(Vc + (i0 + 3), 1);           /* A */
      else if (mm == 2) dx[c] = mac24(Vc + (i0 + 2), 1, 1);      /* B */
      else if (mm == 1) dx[c] = mac22(Vc + (i0 + 1), 1, 1);      /* C */
      else {
        const float vm = fc[fidx(c, 0, ss) + pf];
        if (mm == -1) dx[c] = vm - Vc[i0 - 1];                    /* E */
        else if (mm == -2) {                                       /* F */
          float t3[3] = {Vc[i0 - 2], Vc[i0 - 1], vm};
          dx[c] = vec24_f(t3);
        } else {                                                   /* G */
          float t5[5] = {Vc[i0 - 4], Vc[i0 - 3], Vc[i0 - 2], Vc[i0 - 1], vm};
          dx[c] = vecdrp_f(t5);
        }
      }
    } else {
      if (mm == -3) dx[c] = macdrp_b(Vc + (i0 - 3), 1);          /* G */
      else if (mm == -2) dx[c] = mac24(Vc + (i0 - 2), 1, 0);     /* F */
      else if (mm == -1) dx[c] = mac22(Vc + (i0 - 1), 1, 0);     /* E */
      else {
        const float vp = fc[fidx(c, 1, ss) + pf];
        if (mm == 1) dx[c] = Vc[i0 + 1] - vp;                     /* C */
        else if (mm == 2) {                                        /* B */
          float t3[3] = {vp, Vc[i0 + 1], Vc[i0 + 2]};
          dx[c] = vec24_b(t3);
        } else {                                                   /* A */
          float t5[5] = {vp, Vc[i0 + 1], Vc[i0 + 2], Vc[i0 + 3], Vc[i0 + 4]};
          dx[c] = vecdrp_b(t5);
        }
      }
    }
    dy[c] = macdrp(V[c] + p, sy, jdir);
    dz[c] = macdrp(V[c] + p, sz, kdir);
  }
  if (surf && km == 0) {
    long ij = ((long)i + (long)(iy + 3) * sy) * 9;
    const float *X = a.matVx2Vz + ij, *Y = a.matVy2Vz + ij;
#pragma unroll
    for (int r = 0; r < 3; r++)
      dz[r] = X[3 * r] * dx[0] + X[3 * r + 1] * dx[1] + X[3 * r + 2] * dx[2] + Y[3 * r] * dy[0] + Y[3 * r + 1] * dy[1] +
              Y[3 * r + 2] * dy[2];
  }
  if (surf && km == 1) {
#pragma unroll
    for (int c = 0; c < 3; c++) dz[c] = mac22(V[c] + p, sz, kdir);
  }
  if (surf && km == 2) {
#pragma unroll
    for (int c = 0; c < 3; c++) dz[c] = mac24(V[c] + p, sz, kdir);
  }
  hooke(met, sc, p, a.vcol != 0, a.lam[p], a.mu[p], dx, dy, dz, h + 3);
  rk_cols(a, p, 0, 9, h);
}

template <int IDIR>
__global__ void __launch_bounds__(64 * kColRows) k_fault_cols(FaultArgs a, BatArgs b) {
  const int id = (int)blockIdx.z % a.nfault;
  k_fault_cols_body<IDIR>(fault_at(a, b, b.act[blockIdx.z / a.nfault]), id, (int)blockIdx.x * kColNodes,
                          (int)blockIdx.y * kColRows);
}

/* ---------------------------------------------------------------------- */
/* The split nodes: velocity rhs from the fault traction (fault_gpu.cu:326-425),
   traction rates from the split-node velocity gradient (method 1 or 2, free
   surface rows; :431-1018, :1248-1517), RK update of the 18 split-node
   fields (fault_wav_update, fault_wav_t.cu:318-366), split-node average onto
   the stage output (drv_rk_curv_col.cu:356-359, 541-544) and at the last
   stage the per-step accumulators (fault_var_update_gpu, fault_wav_t.cu:
   197-264). One lane per node, lanes along j. */
template <int IDIR>
DEVI void k_fault_nodes_body(const FaultArgs &a, int id) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const FaultDev F = fdev(a, id);
  const long ss = a.g.sslice, sc = a.g.sicmp, ny = a.g.ny, sy = a.g.siy, sz = a.g.siz;
  const long pf = (iy + 3) + (long)(iz + 3) * ny;
  if (F.united[pf] != 0) return;
  const int jdir = a.jdir, kdir = a.kdir;
  const int i0 = F.i0;
  const long row = (long)(iy + 3) * sy + (long)(iz + 3) * sz;
  const long p0 = i0 + row;
  const int km = a.g.nk - (iz + 1);
  const int n_free = km + 3;
  const bool surf = (a.isfree == 1);
  const float *met = a.met;
  const float *W = a.w_cur;
  const float *fc = a.f_cur + (long)id * 18 * ss;
  const int ry = F.rupy[pf], rz = F.rupz[pf];
  const float *T1 = F.T1;

  /* split-node velocity rhs */
  float fhV[2][3];
#pragma unroll
  for (int m = 0; m < 2; m++) {
    float rrhojac = (float)(1.0 / (F.rho_f[m * ss + pf] * F.fmet[pf]));
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float *t = T1 + (long)c * 7 * ss + pf;
      float dx;
      if (m == 0) dx = (float)(kA0 * t[3 * ss] - kA1 * t[2 * ss] - kA2 * t[1 * ss] - kA3 * t[0]);
      else dx = (float)(-kA0 * t[3 * ss] + kA1 * t[4 * ss] + kA2 * t[5 * ss] + kA3 * t[6 * ss]);
      const float *p2 = fc + fidx(3 + c, m, ss) + pf;
      const float *p3 = fc + fidx(6 + c, m, ss) + pf;
      float dy = 0.0f, dz = 0.0f;
      if (ry == 1) dy = mac22(p2, 1, jdir);
      if (ry == 0) dy = macdrp(p2, 1, jdir);
      if (rz == 1) dz = mac22(p3, ny, kdir);
      if (rz == 0) dz = macdrp(p3, ny, kdir);
      if (surf && km <= 3) dz = dz_free_image(p3, ny, n_free, kdir);
      fhV[m][c] = (dx + dy + dz) * rrhojac;
    }
  }
  /* register fences between the phases keep each phase's loads together */
  __asm__ volatile("" : "+v"(fhV[0][0]), "+v"(fhV[0][1]), "+v"(fhV[0][2]), "+v"(fhV[1][0]), "+v"(fhV[1][1]),
                   "+v"(fhV[1][2]) : : "memory");

  /* split-node velocity gradients: DV[0] minus side, DV[1] plus side */
  const float *V[3] = {W, W + sc, W + 2 * sc};
  const float *fV[3] = {fc + fidx(0, 0, ss), fc + fidx(1, 0, ss), fc + fidx(2, 0, ss)};
  const long pm = pf, pp = pf + ss;
  float DxV[2][3], DyV[2][3], DzV[2][3];
  constexpr int S_ = IDIR == 1 ? 1 : 0;  /* side with the one-sided x difference */
  constexpr int O_ = 1 - S_;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (ry == 1) { DyV[1][c] = mac22(fV[c] + pp, 1, jdir); DyV[0][c] = mac22(fV[c] + pm, 1, jdir); }
    else { DyV[1][c] = macdrp(fV[c] + pp, 1, jdir); DyV[0][c] = macdrp(fV[c] + pm, 1, jdir); }
    if (rz == 1) { DzV[1][c] = mac22(fV[c] + pp, ny, kdir); DzV[0][c] = mac22(fV[c] + pm, ny, kdir); }
    else { DzV[1][c] = macdrp(fV[c] + pp, ny, kdir); DzV[0][c] = macdrp(fV[c] + pm, ny, kdir); }
  }
  const float *Ms = (IDIR == 1 ? F.P2M : F.M2P) + pf; /* 5 matrices, stride 9*ss between them */
  const float *M2 = F.M2 + pf;                        /* T1toVx,VytoVx,VztoVx (minus), then plus */
  const long ms = ss, mq = 9 * ss;
  float dtT1[3] = {0.0f, 0.0f, 0.0f};
  if (F.faultgrid[pf] == 1) {
    dtT1[0] = F.hT1[pf]; dtT1[1] = F.hT1[ss + pf]; dtT1[2] = F.hT1[2 * ss + pf];
  }
  float o1[3], o2[3], o3[3], o4[3], o5[3];
  /* method 2 on both sides with the full matrices (fault_gpu.cu:662-707) */
  auto method2 = [&]() {
    mv3(M2 + 3 * mq, ms, dtT1, o1); mv3(M2 + 4 * mq, ms, DyV[1], o2); mv3(M2 + 5 * mq, ms, DzV[1], o3);
#pragma unroll
    for (int c = 0; c < 3; c++) DxV[1][c] = o1[c] - o2[c] - o3[c];
    mv3(M2 + 0 * mq, ms, dtT1, o1); mv3(M2 + 1 * mq, ms, DyV[0], o2); mv3(M2 + 2 * mq, ms, DzV[0], o3);
#pragma unroll
    for (int c = 0; c < 3; c++) DxV[0][c] = o1[c] - o2[c] - o3[c];
  };
  /* method 1: one-sided difference on side s, continuity maps it to side o */
  auto method1 = [&]() {
    mv3(Ms + 0 * mq, ms, DxV[S_], o1); mv3(Ms + 1 * mq, ms, DyV[S_], o2); mv3(Ms + 2 * mq, ms, DzV[S_], o3);
    mv3(Ms + 3 * mq, ms, DyV[O_], o4); mv3(Ms + 4 * mq, ms, DzV[O_], o5);
#pragma unroll
    for (int c = 0; c < 3; c++) DxV[O_][c] = o1[c] + o2[c] + o3[c] - o4[c] - o5[c];
  };
  if (a.imethod == 1) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if (IDIR == 1) DxV[1][c] = V[c][(i0 + 1) + row] - fV[c][pp];
      else DxV[0][c] = fV[c][pm] - V[c][(i0 - 1) + row];
    }
    method1();
  } else {
    method2();
  }
  if (surf && km == 0) {
    /* split nodes on the free surface (fault_gpu.cu:709-813 / :1427-1517) */
    const long fs_ = ny;  /* SoA stride of the free-surface row */
    const float *Fm = F.Fm + (iy + 3);
    const long fq = 9 * ny;
    if (a.imethod == 1) {
      const float *Mf = Fm + (IDIR == 1 ? 4 : 7) * fq;
      mv3(Mf, fs_, DxV[S_], o1); mv3(Mf + fq, fs_, DyV[S_], o2); mv3(Mf + 2 * fq, fs_, DyV[O_], o3);
#pragma unroll
      for (int c = 0; c < 3; c++) DxV[O_][c] = o1[c] + o2[c] - o3[c];
    } else {
      mv3(Fm + 12 * fq, fs_, dtT1, o1); mv3(Fm + 13 * fq, fs_, DyV[1], o2);
#pragma unroll
      for (int c = 0; c < 3; c++) DxV[1][c] = o1[c] - o2[c];
      mv3(Fm + 10 * fq, fs_, dtT1, o1); mv3(Fm + 11 * fq, fs_, DyV[0], o2);
#pragma unroll
      for (int c = 0; c < 3; c++) DxV[0][c] = o1[c] - o2[c];
    }
    mv3(Fm + 2 * fq, fs_, DxV[1], o1); mv3(Fm + 3 * fq, fs_, DyV[1], o2);
#pragma unroll
    for (int c = 0; c < 3; c++) DzV[1][c] = o1[c] + o2[c];
    mv3(Fm + 0 * fq, fs_, DxV[0], o1); mv3(Fm + 1 * fq, fs_, DyV[0], o2);
#pragma unroll
    for (int c = 0; c < 3; c++) DzV[0][c] = o1[c] + o2[c];
  }
  if (surf && (km == 1 || km == 2)) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if (km == 1) { DzV[1][c] = mac22(fV[c] + pp, ny, kdir); DzV[0][c] = mac22(fV[c] + pm, ny, kdir); }
      else { DzV[1][c] = mac24(fV[c] + pp, ny, kdir); DzV[0][c] = mac24(fV[c] + pm, ny, kdir); }
    }
    if (a.imethod == 1) method1();
    else method2();
  }
#pragma unroll
  for (int m = 0; m < 2; m++)
    __asm__ volatile("" : "+v"(DxV[m][0]), "+v"(DxV[m][1]), "+v"(DxV[m][2]), "+v"(DyV[m][0]), "+v"(DyV[m][1]),
                     "+v"(DyV[m][2]), "+v"(DzV[m][0]), "+v"(DzV[m][1]), "+v"(DzV[m][2]) : : "memory");
  /* traction rates of the split nodes (fault_gpu.cu:936-1018) */
  float fhT2[2][3], fhT3[2][3];
#pragma unroll
  for (int m = 0; m < 2; m++) {
    const float *Db = F.D + (long)(m * 6) * mq + pf;
    float g1[3], g2[3], g3[3];
    mv3(Db + 0 * mq, ms, DxV[m], g1); mv3(Db + 1 * mq, ms, DyV[m], g2); mv3(Db + 2 * mq, ms, DzV[m], g3);
#pragma unroll
    for (int c = 0; c < 3; c++) fhT2[m][c] = g1[c] + g2[c] + g3[c];
    mv3(Db + 3 * mq, ms, DxV[m], g1); mv3(Db + 4 * mq, ms, DyV[m], g2); mv3(Db + 5 * mq, ms, DzV[m], g3);
#pragma unroll
    for (int c = 0; c < 3; c++) fhT3[m][c] = g1[c] + g2[c] + g3[c];
  }

  /* split-node state RK update */
  float *fpre = a.f_pre + (long)id * 18 * ss;
  float *fnew = a.istage < 3 ? a.f_tnew + (long)id * 18 * ss : nullptr;
  float *fend = a.f_end + (long)id * 18 * ss;
  float fe[2][3];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int c = 0; c < 9; c++) {
      float r = c < 3 ? fhV[m][c] : (c < 6 ? fhT2[m][c - 3] : fhT3[m][c - 6]);
      long o = fidx(c, m, ss) + pf;
      float e;
      if (a.istage == 0) {
        float pv = fpre[o];
        fnew[o] = pv + a.ca * r;
        e = pv + a.cb * r;
      } else if (a.istage < 3) {
        fnew[o] = fpre[o] + a.ca * r;
        e = fend[o] + a.cb * r;
      } else {
        e = fend[o] + a.cb * r;
      }
      fend[o] = e;
      if (c < 3) fe[m][c] = e;
    }
  if (a.istage < 3) fault2wave_node(a, F, pf, p0, a.w_tnew, fnew);
  else fault2wave_node(a, F, pf, p0, a.w_end, fend);

  if (a.istage == 3) {
    /* per-step accumulators (fault_var_update_gpu, fault_wav_t.cu:219-262) */
    float dV[3];
#pragma unroll
    for (int c = 0; c < 3; c++) dV[c] = fe[1][c] - fe[0][c];
    if (a.tape) {
      const int slot = *a.it_dev - a.tape_it0;
      if (slot >= 0 && slot < a.tape_nt) {
        float *tp = a.tape + ((long)id * a.tape_nt + slot) * kTapeFloats * ss + pf;
#pragma unroll
        for (int c = 0; c < 3; c++) tp[(13 + c) * ss] = dV[c];
      }
    }
    float Vs1 = dV[0] * F.vs[pf] + dV[1] * F.vs[ss + pf] + dV[2] * F.vs[2 * ss + pf];
    float Vs2 = dV[0] * F.vs[3 * ss + pf] + dV[1] * F.vs[4 * ss + pf] + dV[2] * F.vs[5 * ss + pf];
    float Vs = sqrtf(Vs2 * Vs2 + Vs1 * Vs1);
    float dt = a.dt;
    F.out[3 * ss + pf] = Vs;
    F.out[4 * ss + pf] = Vs1;
    F.out[5 * ss + pf] = Vs2;
    /* Slip, Slip1, Slip2 accumulate in FP64; the outputs (and the Slip the
       friction law reads) are their FP32 values */
    {
      double *sd = F.slip64;
      const double s0 = sd[pf] + (double)(Vs * dt), s1 = sd[ss + pf] + (double)(Vs1 * dt),
                   s2 = sd[2 * ss + pf] + (double)(Vs2 * dt);
      sd[pf] = s0;
      sd[ss + pf] = s1;
      sd[2 * ss + pf] = s2;
      F.out[6 * ss + pf] = (float)s0;
      F.out[7 * ss + pf] = (float)s1;
      F.out[8 * ss + pf] = (float)s2;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float t = T1[(long)c * 7 * ss + 3 * ss + pf];
      F.hT1[c * ss + pf] = (t - F.mT1[c * ss + pf]) / dt;
      F.mT1[c * ss + pf] = t;
    }
    if (Vs > F.out[9 * ss + pf]) F.out[9 * ss + pf] = Vs;
    if (F.init_t0_flag[pf] == 0) {
      if (Vs > 1e-3) {
        F.out[10 * ss + pf] = (float)((*a.it_dev + 1) * (double)dt);
        F.init_t0_flag[pf] = 1;
        F.flag_rup[pf] = 1;
      }
    }
  }
}

template <int IDIR>
__global__ void __launch_bounds__(256) k_fault_nodes(FaultArgs a, BatArgs b) {
  const int id = (int)blockIdx.z % a.nfault;
  k_fault_nodes_body<IDIR>(fault_at(a, b, b.act[blockIdx.z / a.nfault]), id);
}

/* ---------------------------------------------------------------------- */
__global__ void k_step_end(int *it_dev) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *it_dev += 1;
}

/* copy `rows` rows of `width` floats between two pitches */
__global__ void k_repitch(float *dst, long dpitch, const float *src, long spitch, long width, long rows) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = rows * width, stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const long r = i / width, x = i - r * width;
    dst[r * dpitch + x] = src[r * spitch + x];
  }
}

__global__ void k_zero(float *p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0f;
}

__global__ void k_nan_check(const float *p, long n, int *flag) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  int bad = 0;
  for (; i < n; i += stride) bad |= !isfinite(p[i]);
  if (bad) atomicOr(flag, 1);
}

/* receivers (io_recv_keep, drv_rk_curv_col.cu:215): 9 components of level pre */
/* stations: each of the nine components is a sum of nterm terms
   ((v * fx) * fy) * fz in the order of io_recv_keep (io_funcs.cu:452-462);
   one term with unit factors is a station on a node. FMA contraction is off
   so the sums round as the reference's host arithmetic does. */
__global__ void k_record(BatArgs b, const float *w, long sicmp, const long *term, const float *fac, int nterm, int n,
                         float *buf, int nt_max, const int *it_dev, int it0) {
#pragma clang fp contract(off)
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 9) return;
  int s = t / 9, c = t % 9;
  int it = *it_dev - it0;
  if (it < 0 || it >= nt_max) return;
  const long *tm = term + (long)s * nterm;
  const float *f = fac + (long)s * nterm * 3;
  for (int q = 0; q < b.nact; q++) {
    const long sc = b.act[q];
    const float *wc = w + sc * b.wstride + c * sicmp;
    float acc = ((wc[tm[0]] * f[0]) * f[1]) * f[2];
    for (int u = 1; u < nterm; u++) acc = acc + ((wc[tm[u]] * f[3 * u]) * f[3 * u + 1]) * f[3 * u + 2];
    buf[sc * b.ststride + ((long)s * nt_max + it) * 9 + c] = acc;
  }
}

/* ---------------------------------------------------------------------- */
void launch_wave2fault(const FaultArgs &a, const BatArgs &b, hipStream_t st) {
  if (a.nfault <= 0) return;
  dim3 blk(64, 4, 1), grd((a.g.nj + 63) / 64, (a.g.nk + 3) / 4, a.nfault * b.nact);
  hipLaunchKernelGGL(k_wave2fault, grd, blk, 0, st, a, b);
}
void launch_trial(const FaultArgs &a, const BatArgs &b, hipStream_t st) {
  if (a.nfault <= 0) return;
  dim3 tblk(64, 4, 1), tgrd((a.g.nj + kTrialNodes - 1) / kTrialNodes, (a.g.nk + 3) / 4, a.nfault * b.nact);
  hipLaunchKernelGGL(k_trial_t1, tgrd, tblk, 0, st, a, b);
  dim3 blk(64, 2, 1), grd((a.g.nj + 63) / 64, (a.g.nk + 1) / 2, a.nfault * b.nact);
  hipLaunchKernelGGL(k_trial, grd, blk, 0, st, a, b);
}
void launch_fault_update(const FaultArgs &a, const BatArgs &b, hipStream_t st) {
  if (a.nfault <= 0) return;
  dim3 cblk(64, kColRows, 1),
      cgrd((a.g.nj + kColNodes - 1) / kColNodes, (a.g.nk + kColRows - 1) / kColRows, a.nfault * b.nact);
  dim3 nblk(64, 4, 1), ngrd((a.g.nj + 63) / 64, (a.g.nk + 3) / 4, a.nfault * b.nact);
  if (a.idir == 1) {
    hipLaunchKernelGGL(k_fault_cols<1>, cgrd, cblk, 0, st, a, b);
    hipLaunchKernelGGL(k_fault_nodes<1>, ngrd, nblk, 0, st, a, b);
  } else {
    hipLaunchKernelGGL(k_fault_cols<0>, cgrd, cblk, 0, st, a, b);
    hipLaunchKernelGGL(k_fault_nodes<0>, ngrd, nblk, 0, st, a, b);
  }
}
void launch_step_end(int *it_dev, hipStream_t st) { hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, st, it_dev); }
void launch_repitch(float *dst, long dpitch, const float *src, long spitch, long width, long rows, hipStream_t st) {
  if (rows <= 0 || width <= 0) return;
  long blocks = (rows * width + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_repitch, dim3((unsigned)blocks), dim3(256), 0, st, dst, dpitch, src, spitch, width, rows);
}

void launch_zero(float *p, long n, hipStream_t st) {
  if (n <= 0) return;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
}
void launch_nan_check(const float *p, long n, int *flag, hipStream_t st) {
  long blocks = (n + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_nan_check, dim3((unsigned)blocks), dim3(256), 0, st, p, n, flag);
}
/* fault receivers: the nine outputs Tn, Ts1, Ts2, Vs, Vs1, Vs2, Slip, Slip1,
   Slip2 at each receiver, recorded at the step start as the sum of nterm
   terms (v * fy) * fz over (fault id, node) pairs (io_fault_recv_keep,
   io_funcs.cu:2313-2351; one term with unit factors on a node), for every
   active scenario of a batch into its own [n][nt_max][9] block of buf */
__global__ void k_record_fault(BatArgs b, const FaultDev *ftab, const int *rec, const float *fac, int nterm, int n,
                               float *buf, int nt_max, const int *it_dev, int it0, long ss) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 9) return;
  const int r = t / 9, v = t % 9;
  const int it = *it_dev - it0;
  if (it < 0 || it >= nt_max) return;
  const int *tm = rec + 2L * r * nterm;
  const float *f = fac + 2L * r * nterm;
  for (int q = 0; q < b.nact; q++) {
    const long sc = b.act[q], so = sc * b.fpf + (long)v * ss;
    float acc = (ftab[tm[0]].out[so + tm[1]] * f[0]) * f[1];
    for (int u = 1; u < nterm; u++) acc = acc + (ftab[tm[2 * u]].out[so + tm[2 * u + 1]] * f[2 * u]) * f[2 * u + 1];
    buf[sc * n * (long)nt_max * 9 + ((long)r * nt_max + it) * 9 + v] = acc;
  }
}
void launch_record_fault(const BatArgs &b, const FaultDev *ftab, const int *rec, const float *fac, int nterm, int n,
                         float *buf, int nt_max, const int *it_dev, int it0, long ss, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_record_fault, dim3((n * 9 + 255) / 256), dim3(256), 0, st, b, ftab, rec, fac, nterm, n, buf,
                     nt_max, it_dev, it0, ss);
}
void launch_record_stations(const BatArgs &b, const float *w, long sicmp, const long *term, const float *fac,
                            int nterm, int n, float *buf, int nt_max, const int *it_dev, int it0, hipStream_t st) {
  if (n <= 0) return;
  int t = n * 9;
  hipLaunchKernelGGL(k_record, dim3((t + 255) / 256), dim3(256), 0, st, b, w, sicmp, term, fac, nterm, n, buf, nt_max,
                     it_dev, it0);
}

}  // namespace cgfd
