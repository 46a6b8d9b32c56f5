/*
 * k_adj_fault.hip -- discrete adjoint of the fault part of one RK stage
 * (split-node slip-weakening fault, k_fault.hip) and of the per-step fault
 * accumulators, for the reverse sweep of k_adjoint.hip.
 *
 * Forward stage (k_fault.hip, drv_rk_curv_col.cu:215-568):
 *   1 wave2fault   welded nodes copy the stage input onto both split nodes
 *   2 trial_t1     traction columns T1 slots 0-2, 4-6 from the stress
 *   3 trial        trial traction, friction -> T1 slot 3; split-node average
 *                  written onto the stage input at the plane (fault2wave)
 *   4 volume/top   every point but the fault columns of split nodes
 *   5 fault_cols   the six columns next to the plane
 *   6 fault_nodes  split-node rates and RK, fault2wave onto the stage output;
 *                  at the last stage Slip += |V_s| dt, hT1, mT1
 * Everything but the friction law (trial_slipweakening.cu:296-364) and the
 * slip-rate magnitude (fault_wav_t.cu:219-262) is linear with coefficients
 * fixed at setup, so the reverse sweep needs no recomputed forward state:
 * the forward kernels record a tape per step and node (the trial traction of
 * each stage, Slip at the step start, the split-node velocity jump of the
 * last stage), from which the friction branch and its derivatives follow.
 *
 * The value fault2wave writes onto the stage output at the plane is
 * overwritten by the next stage's trial before anything reads it; only the
 * last stage's (the step's end level) is seen, by the station records and
 * the surface maxima at the next step start. The value the trial writes
 * onto the stage input is what the neighbours' stencils read.
 *
 * Reverse of one stage (each kernel the scatter transpose of its forward
 * gather, same lanes and addresses, atomic adds where lanes share a target):
 *   k_adjf_lev   RK structure of the split-node levels (as the wavefield's)
 *   [wave prep / stage / top, fault columns left out]
 *   k_adjf_cols  5^T into the wavefield adjoint, T1 slot 3, split velocity
 *   k_adjf_nodes 6^T into T1, neighbouring split-node state, hT1
 *   k_adjf_trial 3^T (fault2wave, friction on the taped branch, trial) and
 *                2^T; parameter gradients
 *   k_adjf_w2f   1^T at welded nodes
 * and once per step, before the last stage: k_adjf_begin (accumulators
 * and the end-level fault2wave).
 */
#include "cgfd_dev.h"
#include "cgfd_kernels.h"

namespace cgfd {

enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TYZ, TXZ, TXY };

DEVI long fidx(int c, int m, long ss) { return (long)(2 * c + m) * ss; }

/* per-fault adjoint block, offsets in units of sslice */
constexpr int kOffT1 = 72, kOffHT1 = 93, kOffMT1 = 96, kOffSlip = 99, kOffGrad = 100, kOffSlip12 = 107,
              kOffVsr = 109;

DEVI void aadd(float *p, float v) { atomicAdd(p, v); }

/* transposes of the fd_t.h difference macros: add a * d(op)/d(tap) */
DEVI void t_mac22(float *p, long s, int f, float a) {
  if (f == 1) { aadd(p + s, a); aadd(p, -a); }
  else { aadd(p, a); aadd(p - s, -a); }
}
DEVI void t_mac24(float *p, long s, int f, float a) {
  const float b1 = (float)kB1, b2 = (float)kB2, b3 = (float)kB3;
  if (f == 1) { aadd(p, b1 * a); aadd(p + s, b2 * a); aadd(p + 2 * s, b3 * a); }
  else { aadd(p, -b1 * a); aadd(p - s, -b2 * a); aadd(p - 2 * s, -b3 * a); }
}
DEVI void t_macdrp(float *p, long s, int f, float a) {
  const float c1 = (float)kC1, c2 = (float)kC2, c3 = (float)kC3, c4 = (float)kC4, c5 = (float)kC5;
  if (f == 1) {
    aadd(p - s, c1 * a); aadd(p, c2 * a); aadd(p + s, c3 * a); aadd(p + 2 * s, c4 * a); aadd(p + 3 * s, c5 * a);
  } else {
    aadd(p + s, -c1 * a); aadd(p, -c2 * a); aadd(p - s, -c3 * a); aadd(p - 2 * s, -c4 * a); aadd(p - 3 * s, -c5 * a);
  }
}
/* coefficient of tap o (-3..3) in vecc(c, f) */
DEVI float vecc_coef(int o, int f) {
  if (f == 1)
    return o == -1 ? (float)kC1 : o == 0 ? (float)kC2 : o == 1 ? (float)kC3 : o == 2 ? (float)kC4 : o == 3 ? (float)kC5 : 0.0f;
  return o == 1 ? -(float)kC1 : o == 0 ? -(float)kC2 : o == -1 ? -(float)kC3 : o == -2 ? -(float)kC4
                                                                    : o == -3 ? -(float)kC5 : 0.0f;
}
/* odd image across the free surface on a 7-point column (entries above the
   surface replaced), transposed in place on the tap adjoints v */
DEVI void t_free_image(float *v, int n_free) {
  for (int l = 6; l > n_free; l--) {
    v[2 * n_free - l] -= v[l];
    v[l] = 0.0f;
  }
  v[n_free] = 0.0f;
}
/* dz_free_image (k_fault.hip) transposed onto the column at col */
DEVI void t_dz_free_image(float *col, long ny, int n_free, int kdir, float a) {
  float v[7];
#pragma unroll
  for (int l = 0; l < 7; l++) v[l] = a * vecc_coef(l - 3, kdir);
  t_free_image(v, n_free);
#pragma unroll
  for (int l = 0; l < 7; l++)
    if (v[l] != 0.0f) aadd(col + (l - 3) * ny, v[l]);
}
/* x += M^T y for M stored SoA (element e at M[e*stride], row-major) */
DEVI void mv3t_add(const float *__restrict__ M, long stride, const float *y, float *x, float sign = 1.0f) {
#pragma unroll
  for (int j = 0; j < 3; j++)
    x[j] += sign * (M[j * stride] * y[0] + M[(3 + j) * stride] * y[1] + M[(6 + j) * stride] * y[2]);
}

DEVI void metric_cols(const float *met, long sc, long p, float M[3][3]) {
  M[0][0] = met[sc + p];     M[0][1] = met[4 * sc + p]; M[0][2] = met[7 * sc + p];
  M[1][0] = met[2 * sc + p]; M[1][1] = met[5 * sc + p]; M[1][2] = met[8 * sc + p];
  M[2][0] = met[3 * sc + p]; M[2][1] = met[6 * sc + p]; M[2][2] = met[9 * sc + p];
}

/* fault2wave_node (k_fault.hip, transform.cu:331-372) transposed: the
   adjoint aw of the nine wavefield values written at the plane goes to the
   split-node state adjoint af (own node; atomic) and to T1 slot 3 */
DEVI void t_fault2wave(const AdjFaultArgs &a, long pf, long p, const float *aw, float *af, float *aT1) {
  const long sc = a.g.sicmp, ss = a.g.sslice;
  float M[3][3];
  metric_cols(a.met, sc, p, M);
  const float rj = (float)(1.0 / a.met[p]);
  inv3(M);
  float aS[3][3];
  aS[0][0] = aw[TXX];
  aS[1][1] = aw[TYY];
  aS[2][2] = aw[TZZ];
  aS[0][1] = aS[1][0] = 0.5f * aw[TXY];
  const float axz = aw[a.fix_q1 ? TXZ : TYZ], ayz = aw[a.fix_q1 ? TYZ : TXZ];
  aS[0][2] = aS[2][0] = 0.5f * axz;
  aS[1][2] = aS[2][1] = 0.5f * ayz;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float t[3];
#pragma unroll
    for (int k = 0; k < 3; k++) t[k] = rj * (aS[c][0] * M[k][0] + aS[c][1] * M[k][1] + aS[c][2] * M[k][2]);
    aT1[(long)(7 * c + 3) * ss + pf] += t[0];
#pragma unroll
    for (int m = 0; m < 2; m++) {
      aadd(af + fidx(3 + c, m, ss) + pf, 0.5f * t[1]);
      aadd(af + fidx(6 + c, m, ss) + pf, 0.5f * t[2]);
      aadd(af + fidx(c, m, ss) + pf, 0.5f * aw[c]);
    }
  }
}

DEVI float *adj_block(const AdjFaultArgs &a, int id) { return a.adjf + (long)id * kAdjFaultFloats * a.g.sslice; }
DEVI float *adj_lev(const AdjFaultArgs &a, float *blk, int l) { return l < 0 ? nullptr : blk + (long)l * 18 * a.g.sslice; }

/* ---------------------------------------------------------------------- */
/* Once per reverse step, before its last stage: the split-node P level is
   cleared (it accumulates over the step); the last stage's hT1/mT1 update,
   Slip += |V_s| dt and the fault2wave onto the end level are transposed. */
__global__ void __launch_bounds__(256) k_adjf_begin(AdjFaultArgs a) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const int id = blockIdx.z;
  const FaultDev &F = a.ftab[id];
  const long ss = a.g.sslice, sc = a.g.sicmp;
  const long pf = (iy + 3) + (long)(iz + 3) * a.g.ny;
  float *blk = adj_block(a, id);
  float *fP = adj_lev(a, blk, a.lp), *fE = adj_lev(a, blk, a.le);
#pragma unroll
  for (int q = 0; q < 18; q++) fP[q * ss + pf] = 0.0f;
  if (F.united[pf] != 0) return;
  float *aT1 = blk + kOffT1 * ss, *aH = blk + kOffHT1 * ss, *aM = blk + kOffMT1 * ss;
  const float dt = a.dt;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float h = aH[c * ss + pf] / dt;
    aT1[(long)(7 * c + 3) * ss + pf] += h + aM[c * ss + pf];
    aM[c * ss + pf] = -h;
    aH[c * ss + pf] = 0.0f;
  }
  /* Vs1, Vs2 = projections of the velocity jump, Vs = |(Vs1, Vs2)|;
     Slip += Vs dt, Slip1 += Vs1 dt, Slip2 += Vs2 dt; plus the receiver
     residuals of Vs, Vs1, Vs2 of this step */
  const float *tp = a.tape + ((long)id * a.tape_nt + a.n) * kTapeFloats * ss + pf;
  float dV[3] = {tp[13 * ss], tp[14 * ss], tp[15 * ss]};
  const float s1[3] = {F.vs[pf], F.vs[ss + pf], F.vs[2 * ss + pf]};
  const float s2[3] = {F.vs[3 * ss + pf], F.vs[4 * ss + pf], F.vs[5 * ss + pf]};
  const float Vs1 = dV[0] * s1[0] + dV[1] * s1[1] + dV[2] * s1[2];
  const float Vs2 = dV[0] * s2[0] + dV[1] * s2[1] + dV[2] * s2[2];
  const float Vs = sqrtf(Vs2 * Vs2 + Vs1 * Vs1);
  float *vr = blk + kOffVsr * ss + pf;
  const float aVs = vr[0] + blk[kOffSlip * ss + pf] * dt;
  float aV1 = vr[ss] + blk[kOffSlip12 * ss + pf] * dt;
  float aV2 = vr[2 * ss] + blk[(kOffSlip12 + 1) * ss + pf] * dt;
  vr[0] = vr[ss] = vr[2 * ss] = 0.0f;
  if (Vs > 0.0f) {
    aV1 += aVs * Vs1 / Vs;
    aV2 += aVs * Vs2 / Vs;
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float g = aV1 * s1[c] + aV2 * s2[c];
    aadd(fE + fidx(c, 1, ss) + pf, g);
    aadd(fE + fidx(c, 0, ss) + pf, -g);
  }
  /* fault2wave onto the end level */
  const long p0 = F.i0 + (long)(iy + 3) * a.g.siy + (long)(iz + 3) * a.g.siz;
  float aw[9];
#pragma unroll
  for (int f = 0; f < 9; f++) {
    aw[f] = a.wE[f * sc + p0];
    a.wE[f * sc + p0] = 0.0f;
  }
  t_fault2wave(a, pf, p0, aw, fE, aT1);
}

/* fault-receiver residuals of one sample (the outputs after the step being
   reversed): Slip, Slip1, Slip2 into their running adjoints, Vs, Vs1, Vs2
   into this step's; welded nodes output nothing */
__global__ void k_adjf_receivers(AdjFaultArgs a, const int *rec, const float *fac, int nterm, int nrec,
                                 const float *res, int nt, int sample) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrec * nterm) return;
  const int r = t / nterm;
  const int id = rec[2 * t];
  const long pf = rec[2 * t + 1], ss = a.g.sslice;
  const float w = fac[2 * t] * fac[2 * t + 1];
  if (w == 0.0f || a.ftab[id].united[pf] != 0) return;
  float *blk = adj_block(a, id);
  const float *q = res + ((long)r * nt + sample) * 9;
  aadd(blk + kOffVsr * ss + pf, q[3] * w);
  aadd(blk + (kOffVsr + 1) * ss + pf, q[4] * w);
  aadd(blk + (kOffVsr + 2) * ss + pf, q[5] * w);
  aadd(blk + kOffSlip * ss + pf, q[6] * w);
  aadd(blk + kOffSlip12 * ss + pf, q[7] * w);
  aadd(blk + (kOffSlip12 + 1) * ss + pf, q[8] * w);
}

/* RK structure of the split-node levels at reverse stage MODE (the
   wavefield's, k_adjoint.hip): the stage input's adjoint level starts from
   zero (stage 0: P accumulates) */
template <int MODE>
__global__ void __launch_bounds__(256) k_adjf_lev(AdjFaultArgs a) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const long ss = a.g.sslice;
  const long pf = (iy + 3) + (long)(iz + 3) * a.g.ny;
  float *blk = adj_block(a, blockIdx.z);
  float *E = adj_lev(a, blk, a.le), *A = adj_lev(a, blk, 1), *B = adj_lev(a, blk, 2), *P = adj_lev(a, blk, a.lp);
#pragma unroll
  for (int q = 0; q < 18; q++) {
    const long o = q * ss + pf;
    if (MODE == 3) A[o] = 0.0f;
    else if (MODE == 2) { P[o] += E[o] + A[o]; B[o] = 0.0f; }
    else if (MODE == 1) { P[o] += B[o]; A[o] = 0.0f; }
    else P[o] += A[o];
  }
}

/* ---------------------------------------------------------------------- */
/* fault_cols (k_fault.hip) transposed: six lanes per node, one per column */
constexpr int kColNodes = 10;

/* The block's traction-vector adjoints (the entries the forward kernel
   computes once into LDS: 13 x-columns, 16 y-rows and 10 z-planes around
   its kColNodes x kColRows nodes) are accumulated in LDS by the lanes and
   turned into stress adjoints once per entry. */
constexpr int kColRows = 4;
constexpr int kTvX = 13 * kColNodes * kColRows;
constexpr int kTvY = 6 * (kColNodes + 6) * kColRows;
constexpr int kTvZ = 6 * kColNodes * (kColRows + 6);
constexpr int kTv = kTvX + kTvY + kTvZ;

template <int IDIR>
__global__ void __launch_bounds__(256) k_adjf_cols(AdjFaultArgs a) {
  __shared__ float atv[3][kTv];
  const int id = blockIdx.z;
  const FaultDev &F = a.ftab[id];
  const long ss = a.g.sslice, sc = a.g.sicmp, sy = a.g.siy, sz = a.g.siz;
  const int i0 = F.i0;
  const int j0 = blockIdx.x * kColNodes, k0 = blockIdx.y * kColRows;
  const int lane = threadIdx.x, jn = lane / 6, kr = threadIdx.y;
  const int iy = j0 + jn, iz = k0 + kr;
  const long pf = (iy + 3) + (long)(iz + 3) * a.g.ny;
  const bool mine = lane < 6 * kColNodes && iy < a.g.nj && iz < a.g.nk && F.united[pf] == 0;
  if (!__syncthreads_or(mine)) return;  /* blocks of welded nodes only */
  const int tid = threadIdx.x + 64 * threadIdx.y;
  for (int v = tid; v < 3 * kTv; v += 64 * kColRows) (&atv[0][0])[v] = 0.0f;
  __syncthreads();
  const int cl = lane % 6;
  const int mm = cl < 3 ? cl - 3 : cl - 2;
  const int i = i0 + mm, n = -mm;
  const long row = (long)(iy + 3) * sy + (long)(iz + 3) * sz;
  const long p = i + row;
  const int km = a.g.nk - (iz + 1), n_free = km + 3;
  const bool surf = a.isfree == 1;
  const float *met = a.met;
  float *T = a.wT;
  float *blk = adj_block(a, id);
  float *aT1 = blk + kOffT1 * ss, *aC = adj_lev(a, blk, a.lc);
  float hh[9];
  if (mine) {
#pragma unroll
    for (int f = 0; f < 9; f++) hh[f] = a.cb * a.wE[f * sc + p] + (a.wX ? a.ca * a.wX[f * sc + p] : 0.0f);
    /* velocity rhs: (dvx + dvy + dvz) slw / J over the traction vectors at
       the taps */
    const float rrhojac = a.slw[p] / met[p];
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
      const float gc = hh[c] * rrhojac;
      float au[7];
#pragma unroll
      for (int l = 0; l < 7; l++) au[l] = gc * vecc_coef(l - 3, IDIR);
      /* images beyond the plane, then the plane tap (the fault traction) */
      if (n == 2) { au[5] += 2.0f * au[6]; au[4] -= au[6]; au[6] = 0.0f; }
      if (n == 1) {
        au[4] += 2.0f * au[6]; au[2] -= au[6]; au[6] = 0.0f;
        au[4] += 2.0f * au[5]; au[3] -= au[5]; au[5] = 0.0f;
      }
      if (n == -1) {
        au[2] += 2.0f * au[1]; au[3] -= au[1]; au[1] = 0.0f;
        au[2] += 2.0f * au[0]; au[4] -= au[0]; au[0] = 0.0f;
      }
      if (n == -2) { au[1] += 2.0f * au[0]; au[2] -= au[0]; au[0] = 0.0f; }
      aadd(aT1 + (long)(7 * c + 3) * ss + pf, au[n + 3]);
      au[n + 3] = 0.0f;
      float wz[7];
#pragma unroll
      for (int t = 0; t < 7; t++) wz[t] = gc * vecc_coef(t - 3, a.kdir);
      if (surf && km <= 3) t_free_image(wz, n_free);
#pragma unroll
      for (int t = 0; t < 7; t++) {
        if (au[t] != 0.0f) atomicAdd(&atv[c][mm + t + 3 + 13 * (jn + kColNodes * kr)], au[t]);
        const float wy = gc * vecc_coef(t - 3, a.jdir);
        if (wy != 0.0f) atomicAdd(&atv[c][kTvX + cl + 6 * (jn + t + (kColNodes + 6) * kr)], wy);
        if (wz[t] != 0.0f) atomicAdd(&atv[c][kTvX + kTvY + cl + 6 * (jn + kColNodes * (kr + t))], wz[t]);
      }
    }
  }
  __syncthreads();
  /* each entry's stress adjoints: tv_c = J (a1 s_c0 + a2 s_c1 + a3 s_c2) */
  for (int v = tid; v < kTv; v += 64 * kColRows) {
    const float t0 = atv[0][v], t1 = atv[1][v], t2 = atv[2][v];
    if (t0 == 0.0f && t1 == 0.0f && t2 == 0.0f) continue;
    int x, jg, kg, d;
    if (v < kTvX) {
      const int xi = v % 13, r = v / 13;
      x = i0 - 6 + xi; jg = j0 + r % kColNodes + 3; kg = k0 + r / kColNodes + 3; d = 0;
    } else if (v < kTvX + kTvY) {
      const int u = v - kTvX, col = u % 6, r = u / 6;
      x = i0 + (col < 3 ? col - 3 : col - 2); jg = j0 + r % (kColNodes + 6); kg = k0 + r / (kColNodes + 6) + 3; d = 1;
    } else {
      const int u = v - kTvX - kTvY, col = u % 6, r = u / 6;
      x = i0 + (col < 3 ? col - 3 : col - 2); jg = j0 + r % kColNodes + 3; kg = k0 + r / kColNodes; d = 2;
    }
    if (x < 0 || x >= a.g.nx || jg < 0 || jg >= a.g.ny || kg < 0 || kg >= a.g.nz) continue;
    const long q = x + (long)jg * sy + (long)kg * sz;
    const float jq = met[q], a1 = met[(1 + 3 * d) * sc + q], a2 = met[(2 + 3 * d) * sc + q],
                a3 = met[(3 + 3 * d) * sc + q];
    const float u0 = jq * t0, u1 = jq * t1, u2 = jq * t2;
    aadd(T + TXX * sc + q, a1 * u0);
    aadd(T + TXY * sc + q, a2 * u0 + a1 * u1);
    aadd(T + TXZ * sc + q, a3 * u0 + a1 * u2);
    aadd(T + TYY * sc + q, a2 * u1);
    aadd(T + TYZ * sc + q, a3 * u1 + a2 * u2);
    aadd(T + TZZ * sc + q, a3 * u2);
  }
  if (!mine) return;

  /* stress rhs: Hooke's law on the velocity differences */
  const float l = a.lam[p], m = a.mu[p], l2m = (float)(l + 2.0 * m);
  float aG[3][3];
  aG[0][0] = l2m * hh[TXX] + l * hh[TYY] + l * hh[TZZ];
  aG[1][1] = l * hh[TXX] + l2m * hh[TYY] + l * hh[TZZ];
  aG[2][2] = l * hh[TXX] + l * hh[TYY] + l2m * hh[TZZ];
  aG[0][1] = aG[1][0] = m * hh[TXY];
  aG[0][2] = aG[2][0] = m * hh[TXZ];
  aG[1][2] = aG[2][1] = m * hh[TYZ];
  float aDx[3], aDy[3], aDz[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    aDx[c] = met[sc + p] * aG[c][0] + met[2 * sc + p] * aG[c][1] + met[3 * sc + p] * aG[c][2];
    aDy[c] = met[4 * sc + p] * aG[c][0] + met[5 * sc + p] * aG[c][1] + met[6 * sc + p] * aG[c][2];
    aDz[c] = met[7 * sc + p] * aG[c][0] + met[8 * sc + p] * aG[c][1] + met[9 * sc + p] * aG[c][2];
  }
  if (surf && km == 0) {
    const long ij = ((long)i + (long)(iy + 3) * sy) * 9;
    const float *X = a.matVx2Vz + ij, *Y = a.matVy2Vz + ij;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      aDx[c] += X[c] * aDz[0] + X[3 + c] * aDz[1] + X[6 + c] * aDz[2];
      aDy[c] += Y[c] * aDz[0] + Y[3 + c] * aDz[1] + Y[6 + c] * aDz[2];
    }
  }
#pragma unroll 1
  for (int c = 0; c < 3; c++) {
    float *Vc = T + c * sc;
    t_macdrp(Vc + p, sy, a.jdir, aDy[c]);
    if (surf && km == 0) {
    } else if (surf && km == 1) t_mac22(Vc + p, sz, a.kdir, aDz[c]);
    else if (surf && km == 2) t_mac24(Vc + p, sz, a.kdir, aDz[c]);
    else t_macdrp(Vc + p, sz, a.kdir, aDz[c]);
    const float ax = aDx[c];
    float *Vr = Vc + row;
    const float b1 = (float)kB1, b2 = (float)kB2, b3 = (float)kB3;
    const float c1 = (float)kC1, c2 = (float)kC2, c3 = (float)kC3, c4 = (float)kC4, c5 = (float)kC5;
    if (IDIR == 1) {
      if (mm == 3) t_macdrp(Vr + (i0 + 3), 1, 1, ax);
      else if (mm == 2) t_mac24(Vr + (i0 + 2), 1, 1, ax);
      else if (mm == 1) t_mac22(Vr + (i0 + 1), 1, 1, ax);
      else {
        float *vm = aC + fidx(c, 0, ss) + pf;
        if (mm == -1) { aadd(vm, ax); aadd(Vr + (i0 - 1), -ax); }
        else if (mm == -2) { aadd(Vr + (i0 - 2), b1 * ax); aadd(Vr + (i0 - 1), b2 * ax); aadd(vm, b3 * ax); }
        else {
          aadd(Vr + (i0 - 4), c1 * ax); aadd(Vr + (i0 - 3), c2 * ax); aadd(Vr + (i0 - 2), c3 * ax);
          aadd(Vr + (i0 - 1), c4 * ax); aadd(vm, c5 * ax);
        }
      }
    } else {
      if (mm == -3) t_macdrp(Vr + (i0 - 3), 1, 0, ax);
      else if (mm == -2) t_mac24(Vr + (i0 - 2), 1, 0, ax);
      else if (mm == -1) t_mac22(Vr + (i0 - 1), 1, 0, ax);
      else {
        float *vp = aC + fidx(c, 1, ss) + pf;
        if (mm == 1) { aadd(Vr + (i0 + 1), ax); aadd(vp, -ax); }
        else if (mm == 2) { aadd(vp, -b3 * ax); aadd(Vr + (i0 + 1), -b2 * ax); aadd(Vr + (i0 + 2), -b1 * ax); }
        else {
          aadd(vp, -c5 * ax); aadd(Vr + (i0 + 1), -c4 * ax); aadd(Vr + (i0 + 2), -c3 * ax);
          aadd(Vr + (i0 + 3), -c2 * ax); aadd(Vr + (i0 + 4), -c1 * ax);
        }
      }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* fault_nodes (k_fault.hip) transposed: one lane per node */
template <int IDIR>
__global__ void __launch_bounds__(256) k_adjf_nodes(AdjFaultArgs a) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const int id = blockIdx.z;
  const FaultDev &F = a.ftab[id];
  const long ss = a.g.sslice, sc = a.g.sicmp, ny = a.g.ny, sy = a.g.siy, sz = a.g.siz;
  const long pf = (iy + 3) + (long)(iz + 3) * ny;
  if (F.united[pf] != 0) return;
  const int jdir = a.jdir, kdir = a.kdir, i0 = F.i0;
  const long row = (long)(iy + 3) * sy + (long)(iz + 3) * sz;
  const long p0 = i0 + row;
  const int km = a.g.nk - (iz + 1), n_free = km + 3;
  const bool surf = a.isfree == 1;
  const int ry = F.rupy[pf], rz = F.rupz[pf];
  float *blk = adj_block(a, id);
  float *aT1 = blk + kOffT1 * ss, *aC = adj_lev(a, blk, a.lc);
  const float *aE = adj_lev(a, blk, a.le), *aX = adj_lev(a, blk, a.lx);
  float r[2][9];
#pragma unroll
  for (int mside = 0; mside < 2; mside++)
#pragma unroll
    for (int c = 0; c < 9; c++) {
      const long o = fidx(c, mside, ss) + pf;
      r[mside][c] = a.cb * aE[o] + (aX ? a.ca * aX[o] : 0.0f);
    }
  const long ms = ss, mq = 9 * ss;
  /* traction rates fhT2/fhT3 = D (DxV, DyV, DzV) */
  float aDx[2][3] = {}, aDy[2][3] = {}, aDz[2][3] = {};
#pragma unroll
  for (int mside = 0; mside < 2; mside++) {
    const float *Db = F.D + (long)(mside * 6) * mq + pf;
    mv3t_add(Db + 0 * mq, ms, r[mside] + 3, aDx[mside]);
    mv3t_add(Db + 1 * mq, ms, r[mside] + 3, aDy[mside]);
    mv3t_add(Db + 2 * mq, ms, r[mside] + 3, aDz[mside]);
    mv3t_add(Db + 3 * mq, ms, r[mside] + 6, aDx[mside]);
    mv3t_add(Db + 4 * mq, ms, r[mside] + 6, aDy[mside]);
    mv3t_add(Db + 5 * mq, ms, r[mside] + 6, aDz[mside]);
  }
  constexpr int S_ = IDIR == 1 ? 1 : 0, O_ = 1 - S_;
  const float *Ms = (IDIR == 1 ? F.P2M : F.M2P) + pf;
  const float *M2 = F.M2 + pf;
  float aDt[3] = {0.0f, 0.0f, 0.0f};
  auto method1_t = [&]() {
    mv3t_add(Ms + 0 * mq, ms, aDx[O_], aDx[S_]);
    mv3t_add(Ms + 1 * mq, ms, aDx[O_], aDy[S_]);
    mv3t_add(Ms + 2 * mq, ms, aDx[O_], aDz[S_]);
    mv3t_add(Ms + 3 * mq, ms, aDx[O_], aDy[O_], -1.0f);
    mv3t_add(Ms + 4 * mq, ms, aDx[O_], aDz[O_], -1.0f);
#pragma unroll
    for (int c = 0; c < 3; c++) aDx[O_][c] = 0.0f;
  };
  auto method2_t = [&]() {
    mv3t_add(M2 + 3 * mq, ms, aDx[1], aDt);
    mv3t_add(M2 + 4 * mq, ms, aDx[1], aDy[1], -1.0f);
    mv3t_add(M2 + 5 * mq, ms, aDx[1], aDz[1], -1.0f);
    mv3t_add(M2 + 0 * mq, ms, aDx[0], aDt);
    mv3t_add(M2 + 1 * mq, ms, aDx[0], aDy[0], -1.0f);
    mv3t_add(M2 + 2 * mq, ms, aDx[0], aDz[0], -1.0f);
#pragma unroll
    for (int c = 0; c < 3; c++) aDx[0][c] = aDx[1][c] = 0.0f;
  };
  /* the one-sided x difference of method 1 (own split node and the first
     wavefield column on side S) */
  auto diff_t = [&]() {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float g = aDx[S_][c];
      if (IDIR == 1) { aadd(a.wT + c * sc + (i0 + 1) + row, g); aadd(aC + fidx(c, 1, ss) + pf, -g); }
      else { aadd(aC + fidx(c, 0, ss) + pf, g); aadd(a.wT + c * sc + (i0 - 1) + row, -g); }
      aDx[S_][c] = 0.0f;
    }
  };
  const float *fV[3] = {aC + fidx(0, 0, ss), aC + fidx(1, 0, ss), aC + fidx(2, 0, ss)};
  if (surf && km == 0) {
    const float *Fm = F.Fm + (iy + 3);
    const long fs_ = ny, fq = 9 * ny;
    mv3t_add(Fm + 2 * fq, fs_, aDz[1], aDx[1]);
    mv3t_add(Fm + 3 * fq, fs_, aDz[1], aDy[1]);
    mv3t_add(Fm + 0 * fq, fs_, aDz[0], aDx[0]);
    mv3t_add(Fm + 1 * fq, fs_, aDz[0], aDy[0]);
#pragma unroll
    for (int c = 0; c < 3; c++) aDz[0][c] = aDz[1][c] = 0.0f;
    if (a.imethod == 1) {
      const float *Mf = Fm + (IDIR == 1 ? 4 : 7) * fq;
      mv3t_add(Mf, fs_, aDx[O_], aDx[S_]);
      mv3t_add(Mf + fq, fs_, aDx[O_], aDy[S_]);
      mv3t_add(Mf + 2 * fq, fs_, aDx[O_], aDy[O_], -1.0f);
#pragma unroll
      for (int c = 0; c < 3; c++) aDx[O_][c] = 0.0f;
      diff_t();
    } else {
      mv3t_add(Fm + 12 * fq, fs_, aDx[1], aDt);
      mv3t_add(Fm + 13 * fq, fs_, aDx[1], aDy[1], -1.0f);
      mv3t_add(Fm + 10 * fq, fs_, aDx[0], aDt);
      mv3t_add(Fm + 11 * fq, fs_, aDx[0], aDy[0], -1.0f);
    }
  } else {
    if (a.imethod == 1) { method1_t(); diff_t(); }
    else method2_t();
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int mside = 0; mside < 2; mside++) {
        float *q = (float *)fV[c] + (mside ? ss : 0) + pf;
        const float g = aDz[mside][c];
        if (surf && km == 1) t_mac22(q, ny, kdir, g);
        else if (surf && km == 2) t_mac24(q, ny, kdir, g);
        else if (rz == 1) t_mac22(q, ny, kdir, g);
        else t_macdrp(q, ny, kdir, g);
      }
  }
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int mside = 0; mside < 2; mside++) {
      float *q = (float *)fV[c] + (mside ? ss : 0) + pf;
      if (ry == 1) t_mac22(q, 1, jdir, aDy[mside][c]);
      else t_macdrp(q, 1, jdir, aDy[mside][c]);
    }
  if (F.faultgrid[pf] == 1) {
    float *aH = blk + kOffHT1 * ss;
#pragma unroll
    for (int c = 0; c < 3; c++) aH[c * ss + pf] += aDt[c];
  }
  /* split-node velocity rates from T1 and the traction differences */
  const float *met = a.met;
#pragma unroll 1
  for (int mside = 0; mside < 2; mside++) {
    const float rrhojac = (float)(1.0 / (F.rho_f[mside * ss + pf] * met[p0]));
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
      const float g = r[mside][c] * rrhojac;
      float *t = aT1 + (long)c * 7 * ss + pf;
      if (mside == 0) {
        t[3 * ss] += (float)kA0 * g; t[2 * ss] -= (float)kA1 * g; t[ss] -= (float)kA2 * g; t[0] -= (float)kA3 * g;
      } else {
        t[3 * ss] -= (float)kA0 * g; t[4 * ss] += (float)kA1 * g; t[5 * ss] += (float)kA2 * g; t[6 * ss] += (float)kA3 * g;
      }
      float *p2 = aC + fidx(3 + c, mside, ss) + pf, *p3 = aC + fidx(6 + c, mside, ss) + pf;
      if (ry == 1) t_mac22(p2, 1, jdir, g);
      else if (ry == 0) t_macdrp(p2, 1, jdir, g);
      if (surf && km <= 3) t_dz_free_image(p3, ny, n_free, kdir, g);
      else if (rz == 1) t_mac22(p3, ny, kdir, g);
      else if (rz == 0) t_macdrp(p3, ny, kdir, g);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* trial (k_fault.hip, trial_slipweakening.cu:91-388) and trial_t1
   transposed, with the parameter gradients; one lane per split node */
__global__ void __launch_bounds__(128) k_adjf_trial(AdjFaultArgs a) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const int id = blockIdx.z;
  const FaultDev &F = a.ftab[id];
  const long ss = a.g.sslice, sc = a.g.sicmp, ny = a.g.ny;
  const long pf = (iy + 3) + (long)(iz + 3) * ny;
  if (F.united[pf] != 0) return;
  const int jdir = a.jdir, kdir = a.kdir;
  const float *met = a.met;
  const long row = (long)(iy + 3) * a.g.siy + (long)(iz + 3) * a.g.siz;
  const long p0 = F.i0 + row;
  const int km = a.g.nk - (iz + 1), n_free = km + 3;
  const int ry = F.rupy[pf], rz = F.rupz[pf];
  float *blk = adj_block(a, id);
  float *aT1 = blk + kOffT1 * ss, *aC = adj_lev(a, blk, a.lc), *aP = adj_lev(a, blk, a.lp);
  float *gr = blk + kOffGrad * ss;

  /* fault2wave onto the stage input at the plane */
  {
    float aw[9];
#pragma unroll
    for (int f = 0; f < 9; f++) {
      aw[f] = a.wT[f * sc + p0];
      a.wT[f * sc + p0] = 0.0f;
    }
    t_fault2wave(a, pf, p0, aw, aC, aT1);
  }

  /* friction on the taped branch */
  const float *tp = a.tape + ((long)id * a.tape_nt + a.n) * kTapeFloats * ss + pf;
  float Trial[3] = {tp[(3 * a.istage) * ss], tp[(3 * a.istage + 1) * ss], tp[(3 * a.istage + 2) * ss]};
  const float Slip = tp[12 * ss];
  float at[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    at[c] = aT1[(long)(7 * c + 3) * ss + pf];
    aT1[(long)(7 * c + 3) * ss + pf] = 0.0f;
  }
  float vn[3] = {met[sc + p0], met[2 * sc + p0], met[3 * sc + p0]};
  const float vn0 = norm3(vn);
  const float jacvec = met[p0] * vn0;
#pragma unroll
  for (int c = 0; c < 3; c++) vn[c] /= vn0;
  const float T0[3] = {F.T0[pf], F.T0[ss + pf], F.T0[2 * ss + pf]};
  float TL[3], Ts[3];
#pragma unroll
  for (int c = 0; c < 3; c++) TL[c] = Trial[c] / jacvec + T0[c];
  const float Tn0 = dot3(TL, vn);
#pragma unroll
  for (int c = 0; c < 3; c++) Ts[c] = TL[c] - vn[c] * Tn0;
  const float Ts0 = norm3(Ts);
  const bool open = Tn0 >= 0.0f;
  const float Tau_n = open ? 0.0f : Tn0;
  const float mus = F.fric[pf], mud = F.fric[ss + pf], Dc = F.fric[2 * ss + pf], C0 = F.fric[3 * ss + pf];
  const bool weak = Slip <= Dc;
  const float fr = weak ? mus - (mus - mud) * Slip / Dc : mud;
  const float Tau_c = -fr * Tau_n + C0;
  const bool capped = Ts0 >= Tau_c;
  const bool ifchange = open || capped;
  float aTrial[3];
  if (!ifchange) {
#pragma unroll
    for (int c = 0; c < 3; c++) aTrial[c] = at[c];
  } else {
    /* T1_3 = (Tau - T0) jacvec, Tau = Tau_n vn + (Tau_s / Ts0) Ts */
    float aTau[3], aTL[3] = {0.0f, 0.0f, 0.0f}, aT0[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      aTau[c] = at[c] * jacvec;
      aT0[c] = -aTau[c];
    }
    float aTau_n = dot3(aTau, vn), aTs0 = 0.0f, aTau_s = 0.0f;
    float aTs[3] = {0.0f, 0.0f, 0.0f};
    const float Tau_s = capped ? Tau_c : Ts0;
    if (!(fabsf(Ts0) < 1e-1f)) {
      const float kap = Tau_s / Ts0;
      const float akap = dot3(aTau, Ts);
#pragma unroll
      for (int c = 0; c < 3; c++) aTs[c] = kap * aTau[c];
      aTau_s = akap / Ts0;
      aTs0 = -akap * Tau_s / (Ts0 * Ts0);
    }
    float aTau_c = 0.0f;
    if (capped) aTau_c = aTau_s;
    else aTs0 += aTau_s;
    /* Tau_c = -fr Tau_n + C0 */
    const float afr = -aTau_c * Tau_n;
    aTau_n += -fr * aTau_c;
    gr[3 * ss + pf] += aTau_c;
    if (weak) {
      gr[pf] += afr * (1.0f - Slip / Dc);
      gr[ss + pf] += afr * (Slip / Dc);
      gr[2 * ss + pf] += afr * (mus - mud) * Slip / (Dc * Dc);
      blk[kOffSlip * ss + pf] += -afr * (mus - mud) / Dc;
    } else {
      gr[ss + pf] += afr;
    }
    float aTn0 = open ? 0.0f : aTau_n;
    if (Ts0 > 0.0f)
#pragma unroll
      for (int c = 0; c < 3; c++) aTs[c] += aTs0 * Ts[c] / Ts0;
#pragma unroll
    for (int c = 0; c < 3; c++) aTL[c] += aTs[c];
    aTn0 -= dot3(aTs, vn);
#pragma unroll
    for (int c = 0; c < 3; c++) aTL[c] += aTn0 * vn[c];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      aTrial[c] = aTL[c] / jacvec;
      gr[(4 + c) * ss + pf] += aTL[c] + aT0[c];
    }
  }

  /* Trial = 2 (M0 M1 dV / dt + M0 R1 - M1 R0) / (a0 (M0 + M1)) */
  const float jac0 = met[p0];
  const float M0 = (float)(0.5 * jac0 * F.rho_f[pf]), M1 = (float)(0.5 * jac0 * F.rho_f[ss + pf]);
  const float den = (float)(2.0 / (kA0 * (M0 + M1)));
#pragma unroll 1
  for (int c = 0; c < 3; c++) {
    const float g = aTrial[c] * den;
    const float adv = g * M0 * M1 / a.dt;
    aadd(aP + fidx(c, 1, ss) + pf, adv);
    aadd(aP + fidx(c, 0, ss) + pf, -adv);
#pragma unroll 1
    for (int mside = 0; mside < 2; mside++) {
      const float aR = mside ? g * M0 : -g * M1;
      const float ar = 0.5f * (float)(2 * mside - 1) * aR, ad = 0.5f * aR;
      float *t = aT1 + (long)c * 7 * ss + pf;
      t[(3 + (2 * mside - 1)) * ss] += (float)kA1 * ar;
      t[(3 + 2 * (2 * mside - 1)) * ss] += (float)kA2 * ar;
      t[(3 + 3 * (2 * mside - 1)) * ss] += (float)kA3 * ar;
      float *p2 = aC + fidx(3 + c, mside, ss) + pf, *p3 = aC + fidx(6 + c, mside, ss) + pf;
      if (ry == 1) t_mac22(p2, 1, jdir, ad);
      else if (ry == 0) t_macdrp(p2, 1, jdir, ad);
      if (a.isfree == 1 && km <= 3) t_dz_free_image(p3, ny, n_free, kdir, ad);
      else if (rz == 1) t_mac22(p3, ny, kdir, ad);
      else if (rz == 0) t_macdrp(p3, ny, kdir, ad);
    }
  }

  /* traction columns T1 slots 0-2 / 4-6 from the stress (trial_t1) */
#pragma unroll 1
  for (int mside = 0; mside < 2; mside++)
#pragma unroll 1
    for (int l = 1; l <= 3; l++) {
      const long slot = 3 + (2 * mside - 1) * l;
      const long q = F.i0 + (2 * mside - 1) * l + row;
      const float jl = met[q], x1 = met[sc + q], x2 = met[2 * sc + q], x3 = met[3 * sc + q];
      float *t = aT1 + slot * ss + pf;
      const float a0 = jl * t[0], a1 = jl * t[7 * ss], a2 = jl * t[14 * ss];
      t[0] = t[7 * ss] = t[14 * ss] = 0.0f;
      float *W = a.wT;
      W[TXX * sc + q] += x1 * a0;
      W[TXY * sc + q] += x2 * a0 + x1 * a1;
      W[TXZ * sc + q] += x3 * a0 + x1 * a2;
      W[TYY * sc + q] += x2 * a1;
      W[TYZ * sc + q] += x3 * a1 + x2 * a2;
      W[TZZ * sc + q] += x3 * a2;
    }
}

/* wave2fault (k_fault.hip, transform.cu:100-212) transposed at welded nodes */
__global__ void __launch_bounds__(256) k_adjf_w2f(AdjFaultArgs a) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y * blockDim.y + threadIdx.y;
  if (iy >= a.g.nj || iz >= a.g.nk) return;
  const int id = blockIdx.z;
  const FaultDev &F = a.ftab[id];
  const long ss = a.g.sslice, sc = a.g.sicmp;
  const long pf = (iy + 3) + (long)(iz + 3) * a.g.ny;
  if (F.united[pf] != 1) return;
  const long p = F.i0 + (long)(iy + 3) * a.g.siy + (long)(iz + 3) * a.g.siz;
  float *blk = adj_block(a, id);
  float *aT1 = blk + kOffT1 * ss, *aC = adj_lev(a, blk, a.lc);
  float aTr[3][3], aV[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    aTr[c][0] = aT1[(long)(7 * c + 3) * ss + pf];
    aT1[(long)(7 * c + 3) * ss + pf] = 0.0f;
    aTr[c][1] = aTr[c][2] = aV[c] = 0.0f;
#pragma unroll
    for (int m = 0; m < 2; m++) {
      aV[c] += aC[fidx(c, m, ss) + pf];
      aTr[c][1] += aC[fidx(3 + c, m, ss) + pf];
      aTr[c][2] += aC[fidx(6 + c, m, ss) + pf];
    }
  }
#pragma unroll
  for (int q = 0; q < 18; q++) aC[q * ss + pf] = 0.0f;
  float M[3][3];
  metric_cols(a.met, sc, p, M);
  const float jac = a.met[p];
  float aS[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int k = 0; k < 3; k++) aS[r][k] = jac * (aTr[r][0] * M[k][0] + aTr[r][1] * M[k][1] + aTr[r][2] * M[k][2]);
  float *W = a.wT;
#pragma unroll
  for (int c = 0; c < 3; c++) W[c * sc + p] += aV[c];
  W[TXX * sc + p] += aS[0][0];
  W[TYY * sc + p] += aS[1][1];
  W[TZZ * sc + p] += aS[2][2];
  W[TXY * sc + p] += aS[0][1] + aS[1][0];
  W[TXZ * sc + p] += aS[0][2] + aS[2][0];
  W[TYZ * sc + p] += aS[1][2] + aS[2][1];
}

/* ---------------------------------------------------------------------- */
static dim3 node_grid(const AdjFaultArgs &a) { return dim3((a.g.nj + 63) / 64, (a.g.nk + 3) / 4, a.nfault); }

void launch_adjf_begin(const AdjFaultArgs &a, hipStream_t st) {
  if (a.nfault <= 0) return;
  hipLaunchKernelGGL(k_adjf_begin, node_grid(a), dim3(64, 4), 0, st, a);
}

void launch_adjf_receivers(const AdjFaultArgs &a, const int *rec, const float *fac, int nterm, int nrec,
                           const float *res, int nt, int sample, hipStream_t st) {
  if (a.nfault <= 0 || nrec <= 0 || !res) return;
  hipLaunchKernelGGL(k_adjf_receivers, dim3((nrec * nterm + 63) / 64), dim3(64), 0, st, a, rec, fac, nterm, nrec, res,
                     nt, sample);
}

void launch_adjf_stage_pre(const AdjFaultArgs &a, hipStream_t st) {
  if (a.nfault <= 0) return;
  const dim3 g = node_grid(a), b(64, 4);
  switch (a.istage) {
    case 3: hipLaunchKernelGGL(k_adjf_lev<3>, g, b, 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_adjf_lev<2>, g, b, 0, st, a); break;
    case 1: hipLaunchKernelGGL(k_adjf_lev<1>, g, b, 0, st, a); break;
    default: hipLaunchKernelGGL(k_adjf_lev<0>, g, b, 0, st, a); break;
  }
}

void launch_adjf_stage_post(const AdjFaultArgs &a, hipStream_t st) {
  if (a.nfault <= 0) return;
  const dim3 cg((a.g.nj + kColNodes - 1) / kColNodes, (a.g.nk + 3) / 4, a.nfault), b(64, 4);
  if (a.idir == 1) {
    hipLaunchKernelGGL(k_adjf_cols<1>, cg, b, 0, st, a);
    hipLaunchKernelGGL(k_adjf_nodes<1>, node_grid(a), b, 0, st, a);
  } else {
    hipLaunchKernelGGL(k_adjf_cols<0>, cg, b, 0, st, a);
    hipLaunchKernelGGL(k_adjf_nodes<0>, node_grid(a), b, 0, st, a);
  }
  hipLaunchKernelGGL(k_adjf_trial, dim3((a.g.nj + 31) / 32, (a.g.nk + 3) / 4, a.nfault), dim3(32, 4), 0, st, a);
  hipLaunchKernelGGL(k_adjf_w2f, node_grid(a), b, 0, st, a);
}

}  // namespace cgfd
