/*
 * k_adjoint.hip -- discrete adjoint of the RK4 volume update (interior
 * stencil, CFS-PML, free surface); the fault part is in k_adj_fault.hip.
 *
 * One forward stage computes, at every physical point p,
 *   h(p) = H(p) . [D_d f(p)],  D_d f(p) = sum_t c_t f(p + (o_d + t) e_d)
 * (the 5-point one-sided operators of fd_t.cu:85-109 along each axis d,
 * o_d = -1 forward / -3 backward), where H(p) forms the physical gradients
 * g(f,a) = sum_d M[d][a] D_d f with the metric rows M[d] = (xi, eta, zeta)_{x,y,z}
 * and then momentum and Hooke's law (sv_curv_col_el_iso_gpu.cu:349-386).
 * Its transpose applied to an adjoint rhs c is
 *   lambda_f(q) = sum_d sum_t c_t kappa(f,d)(q - (o_d + t) e_d),
 *   kappa(f,d)(p) = sum_a M[d][a](p) gamma(f,a)(p),
 * where gamma are the adjoints of the physical gradients, formed pointwise
 * from c(p) and the media at p:
 *   stress fields: gamma from S = (1/rho) c_V (3 values),
 *   velocity fields: gamma(V,a) from the stress adjoints through Hooke's law
 *   (9 values G).
 * So each reverse stage is two kernels: k_adj_prep writes S and G at the
 * physical points (ghost entries stay zero: the forward never uses h at a
 * ghost), and k_adj_stage gathers the transposed operators with the metric
 * at each tap and does the pointwise RK-adjoint accumulations.
 *
 * CFS-PML (sv_curv_col_el_iso_gpu.cu:916-1288): at a point of face (d, side)
 * the forward adds (B-1) r_d - B a to the rhs and advances the auxiliary a
 * with rhs D r_d - A a, r_d being the part of the rhs from the derivatives
 * along d. Its transpose: the adjoint of r_d gains (B-1) c + D c_a, kept per
 * face as extra S and G (added by the gather at taps inside the face), and
 * the auxiliary's stage adjoint is -B c - A c_a, pointwise, with the same
 * reverse RK structure as the wavefield (c_a: the auxiliary rhs adjoint).
 *
 * Free surface (sv_curv_col_el_iso_gpu.cu:398-747, top four layers): the
 * forward replaces the velocity rhs by the traction-image form (timg) and
 * the stress rhs by Hooke's law on the reduced-order velocity gradient
 * (vlow), and the x/y PML faces add a term through the free-surface
 * matrices at the surface. Transposed: the prep leaves out the base S / G of
 * what was replaced, and k_adj_top scatters the adjoint of timg, vlow and
 * the surface PML term onto the taps (atomic adds; ghosts are not state).
 *
 * Fault columns of split nodes (their rhs comes from the fault kernels) get
 * no S and G here and no top-layer scatter; k_adj_fault.hip transposes them.
 *
 * RK4 (drv_rk_curv_col.cu:329-568) with a = {1/2, 1/2, 1}, b = {1/6, 1/3,
 * 1/3, 1/6}, ca = a dt, cb = b dt, reversed (E = adjoint of the step's
 * output, P = adjoint of its input):
 *   stage 3: A  = L3^T (cb3 E)
 *   stage 2: B  = L2^T (ca2 A + cb2 E),  P  = E + A
 *   stage 1: A  = L1^T (ca1 B + cb1 E),  P += B
 *   stage 0: P += A + L0^T (ca0 A + cb0 E)
 */
#include "cgfd_dev.h"
#include "cgfd_kernels.h"
#include "k_vol_common.h"

namespace cgfd {

/* S (3) and the 6 distinct entries of G of an adjoint rhs c with the media
   at the point: G is symmetric (gamma(Vx, y) = gamma(Vy, x) = mu c_xy and
   so on), so a point stores 9 values, not 12 */
constexpr int kSG = 9;
DEVI void sg_of(const float *c, float l, float m, float s, float *o) {
  const float l2m = l + 2.0f * m;
  o[0] = s * c[VX];
  o[1] = s * c[VY];
  o[2] = s * c[VZ];
  o[3] = l2m * c[TXX] + l * c[TYY] + l * c[TZZ];  /* gamma(Vx, x) */
  o[4] = l * c[TXX] + l2m * c[TYY] + l * c[TZZ];  /* gamma(Vy, y) */
  o[5] = l * c[TXX] + l * c[TYY] + l2m * c[TZZ];  /* gamma(Vz, z) */
  o[6] = m * c[TXY];                              /* gamma(Vx, y) = gamma(Vy, x) */
  o[7] = m * c[TXZ];                              /* gamma(Vx, z) = gamma(Vz, x) */
  o[8] = m * c[TYZ];                              /* gamma(Vy, z) = gamma(Vz, y) */
}

DEVI bool in_face(const AdjFace &F, int i, int j, int k) {
  return F.on && i >= F.i1 && i <= F.i2 && j >= F.j1 && j <= F.j2 && k >= F.k1 && k <= F.k2;
}
DEVI long face_idx(const AdjFace &F, int i, int j, int k) {
  return (long)(k - F.k1) * ((long)F.bn * F.an) + (long)(j - F.j1) * F.an + (i - F.ax0);
}

/* S and G from c = alpha X + beta E at physical points (X may be null);
   at CFS-PML points also the per-face extra S and G and the auxiliaries'
   reverse stage (MODE as k_adj_stage) */
template <int MODE>
__global__ void __launch_bounds__(256) k_adj_prep(Grid g, const float *__restrict__ lam, const float *__restrict__ mu,
                                                  const float *__restrict__ slw, const float *__restrict__ X,
                                                  float alpha, const float *__restrict__ E, float beta,
                                                  float *__restrict__ SG, AdjPml pm, int isfree, int kdir) {
  const int i = blockIdx.x * 64 + (int)threadIdx.x + 3, j = blockIdx.y * 4 + (int)threadIdx.y + 3;
  const int k = (int)blockIdx.z + 3;
  if (i >= g.ni + 3 || j >= g.nj + 3) return;
  const long sc = g.sicmp;
  const long p = i + j * g.siy + (long)k * g.siz;
  if (pm.fs.nfault > 0 && fault_column(pm.fs, i, j + (long)k * g.ny)) {
#pragma unroll
    for (int q = 0; q < kSG; q++) SG[q * sc + p] = 0.0f;
    return;
  }
  float c[9];
#pragma unroll
  for (int f = 0; f < 9; f++) c[f] = beta * E[f * sc + p] + (X ? alpha * X[f * sc + p] : 0.0f);
  const float l = lam[p], m = mu[p], s = slw[p];
  float o[kSG];
  sg_of(c, l, m, s, o);
  if (isfree) {
    /* what timg (velocity rhs) and vlow (stress rhs) replace at the top */
    if (k >= g.nk2 - (kdir ? 3 : 1)) o[0] = o[1] = o[2] = 0.0f;
    if (k >= g.nk2 - 2)
#pragma unroll
      for (int q = 3; q < kSG; q++) o[q] = 0.0f;
  }
#pragma unroll
  for (int q = 0; q < kSG; q++) SG[q * sc + p] = o[q];
  if (!pm.on) return;
#pragma unroll 1
  for (int fc = 0; fc < 6; fc++) {
    const AdjFace &F = pm.f[fc];
    if (!in_face(F, i, j, k)) continue;
    const long ia = face_idx(F, i, j, k), n = F.n;
    const int ic = F.dim == 0 ? i - F.i1 : (F.dim == 1 ? j - F.j1 : k - F.k1);
    const float cA = F.A[ic], cB = F.B[ic], cD = F.D[ic];
    const float *Ea = F.lev + (long)pm.le * 9 * n, *Aa = F.lev + 9 * n, *Ba = F.lev + 18 * n;
    float *Pa = F.lev + (long)pm.lp * 9 * n;
    float *Aw = F.lev + 9 * n, *Bw = F.lev + 18 * n;
    const float *Xa = MODE == 3 ? nullptr : (MODE == 1 ? Ba : Aa);
    float dc[9];
#pragma unroll
    for (int f = 0; f < 9; f++) {
      const long o2 = f * n + ia;
      const float ca_ = beta * Ea[o2] + (Xa ? alpha * Xa[o2] : 0.0f);  /* auxiliary rhs adjoint */
      dc[f] = (cB - 1.0f) * c[f] + cD * ca_;
      const float la = -cB * c[f] - cA * ca_;                          /* auxiliary stage adjoint */
      if (MODE == 3) Aw[o2] = la;
      else if (MODE == 2) { Bw[o2] = la; Pa[o2] = Ea[o2] + Aa[o2]; }
      else if (MODE == 1) { const float b = Ba[o2]; Aw[o2] = la; Pa[o2] += b; }
      else Pa[o2] += Aa[o2] + la;
    }
    sg_of(dc, l, m, s, o);
#pragma unroll
    for (int q = 0; q < kSG; q++) F.dsg[q * n + ia] = o[q];
  }
}

template <int DIR>
DEVI float tap_coef(int t) {
  return DIR ? (t == 0 ? CGFD_CF0 : t == 1 ? CGFD_CF1 : t == 2 ? CGFD_CF2 : t == 3 ? CGFD_CF3 : CGFD_CF4)
             : (t == 0 ? CGFD_CB0 : t == 1 ? CGFD_CB1 : t == 2 ? CGFD_CB2 : t == 3 ? CGFD_CB3 : CGFD_CB4);
}

/* add w * kappa(., d)(p) for all nine fields; M: metric row d at p.
   AX = 0 or 1: only M[AX] is nonzero (the xi and eta rows on a grid of
   vertical columns, see acc_core), and only the S and G entries it
   multiplies are read; AX = -1: the full row */
template <int AX>
DEVI void add_kappa(float *lm, float w, const float *__restrict__ SG, long sc, long p, float M0, float M1, float M2) {
  const float Sx = SG[p], Sy = SG[sc + p], Sz = SG[2 * sc + p];
  const float *G = SG + 3 * sc + p;  /* xx, yy, zz, xy, xz, yz (sg_of) */
  if (AX == 0) {
    lm[VX] += w * (M0 * G[0]);
    lm[VY] += w * (M0 * G[3 * sc]);
    lm[VZ] += w * (M0 * G[4 * sc]);
    lm[TXX] += w * (M0 * Sx);
    lm[TXY] += w * (M0 * Sy);
    lm[TXZ] += w * (M0 * Sz);
  } else if (AX == 1) {
    lm[VX] += w * (M1 * G[3 * sc]);
    lm[VY] += w * (M1 * G[sc]);
    lm[VZ] += w * (M1 * G[5 * sc]);
    lm[TYY] += w * (M1 * Sy);
    lm[TXY] += w * (M1 * Sx);
    lm[TYZ] += w * (M1 * Sz);
  } else {
    const float Gxx = G[0], Gyy = G[sc], Gzz = G[2 * sc], Gxy = G[3 * sc], Gxz = G[4 * sc], Gyz = G[5 * sc];
    lm[VX] += w * (M0 * Gxx + M1 * Gxy + M2 * Gxz);
    lm[VY] += w * (M0 * Gxy + M1 * Gyy + M2 * Gyz);
    lm[VZ] += w * (M0 * Gxz + M1 * Gyz + M2 * Gzz);
    lm[TXX] += w * (M0 * Sx);
    lm[TYY] += w * (M1 * Sy);
    lm[TZZ] += w * (M2 * Sz);
    lm[TXY] += w * (M0 * Sy + M1 * Sx);
    lm[TXZ] += w * (M0 * Sz + M2 * Sx);
    lm[TYZ] += w * (M1 * Sz + M2 * Sy);
  }
}

/* the per-face extra S and G of the faces along `dim` at tap (ti, tj, tk) */
template <int AX>
DEVI void add_face_kappa(float *lm, float w, const AdjPml &pm, int dim, int ti, int tj, int tk, float M0, float M1,
                         float M2) {
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const AdjFace &F = pm.f[2 * dim + side];
    if (!in_face(F, ti, tj, tk)) continue;
    add_kappa<AX>(lm, w, F.dsg, F.n, face_idx(F, ti, tj, tk), M0, M1, M2);
  }
}

/* MODE 3: A = L^T;  2: B = L^T, P = E + A;  1: A = L^T, P += B;  0: P += A + L^T.
   VC: vertical columns (xi_y, xi_z, eta_x, eta_z left out, as the forward
   box split does) */
template <int ID, int JD, int KD, int MODE, bool VC>
__global__ void __launch_bounds__(256) k_adj_stage(Grid g, const float *__restrict__ met,
                                                   const float *__restrict__ SG, const float *__restrict__ E,
                                                   float *A, float *B, float *P, AdjPml pm) {
  const int i = blockIdx.x * 64 + (int)threadIdx.x + 3, j = blockIdx.y * 4 + (int)threadIdx.y + 3;
  const int k = (int)blockIdx.z + 3;
  if (i >= g.ni + 3 || j >= g.nj + 3) return;
  const long sc = g.sicmp, sy = g.siy, sz = g.siz;
  const long q = i + j * sy + (long)k * sz;
  float lm[9];
#pragma unroll
  for (int f = 0; f < 9; f++) lm[f] = 0.0f;
  constexpr int ox = ID ? -1 : -3, oy = JD ? -1 : -3, oz = KD ? -1 : -3;
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const long p = q - (ox + t);
    const float M0 = met[sc + p], M1 = VC ? 0.0f : met[2 * sc + p], M2 = VC ? 0.0f : met[3 * sc + p];
    add_kappa<VC ? 0 : -1>(lm, tap_coef<ID>(t), SG, sc, p, M0, M1, M2);
    if (pm.on) add_face_kappa<VC ? 0 : -1>(lm, tap_coef<ID>(t), pm, 0, i - (ox + t), j, k, M0, M1, M2);
  }
  __asm__ volatile("" : "+v"(lm[0]), "+v"(lm[1]), "+v"(lm[2]), "+v"(lm[3]), "+v"(lm[4]), "+v"(lm[5]),
                   "+v"(lm[6]), "+v"(lm[7]), "+v"(lm[8]) : : "memory");
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const long p = q - (long)(oy + t) * sy;
    const float M0 = VC ? 0.0f : met[4 * sc + p], M1 = met[5 * sc + p], M2 = VC ? 0.0f : met[6 * sc + p];
    add_kappa<VC ? 1 : -1>(lm, tap_coef<JD>(t), SG, sc, p, M0, M1, M2);
    if (pm.on) add_face_kappa<VC ? 1 : -1>(lm, tap_coef<JD>(t), pm, 1, i, j - (oy + t), k, M0, M1, M2);
  }
  __asm__ volatile("" : "+v"(lm[0]), "+v"(lm[1]), "+v"(lm[2]), "+v"(lm[3]), "+v"(lm[4]), "+v"(lm[5]),
                   "+v"(lm[6]), "+v"(lm[7]), "+v"(lm[8]) : : "memory");
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const long p = q - (long)(oz + t) * sz;
    const float M0 = met[7 * sc + p], M1 = met[8 * sc + p], M2 = met[9 * sc + p];
    add_kappa<-1>(lm, tap_coef<KD>(t), SG, sc, p, M0, M1, M2);
    if (pm.on) add_face_kappa<-1>(lm, tap_coef<KD>(t), pm, 2, i, j, k - (oz + t), M0, M1, M2);
  }
#pragma unroll
  for (int f = 0; f < 9; f++) {
    const long o = f * sc + q;
    if (MODE == 3) A[o] = lm[f];
    else if (MODE == 2) { B[o] = lm[f]; P[o] = E[o] + A[o]; }
    else if (MODE == 1) { const float b = B[o]; A[o] = lm[f]; P[o] += b; }
    else P[o] += A[o] + lm[f];
  }
}

/* ------------------------------------------------------------------------
 * Adjoint of the free-surface replacements at the top four layers, added to
 * the stage's L^T output T (the level the stage kernel just wrote: A for
 * stages 3 and 1, B for stage 2, P for stage 0). One thread per point.
 * ---------------------------------------------------------------------- */
DEVI bool phys(const Grid &g, int i, int j, int k) {
  return i >= 3 && i < g.ni + 3 && j >= 3 && j < g.nj + 3 && k >= 3 && k <= g.nk2;
}

template <int ID, int JD, int KD, int MODE>
__global__ void __launch_bounds__(256) k_adj_top(Grid g, const float *__restrict__ met, const float *__restrict__ lam,
                                                 const float *__restrict__ mu, const float *__restrict__ slw,
                                                 const float *__restrict__ mx, const float *__restrict__ my,
                                                 const float *__restrict__ X, float alpha,
                                                 const float *__restrict__ E, float beta, float *T, AdjPml pm) {
  const int i = blockIdx.x * 64 + (int)threadIdx.x + 3, j = blockIdx.y * 4 + (int)threadIdx.y + 3;
  const int nk2 = g.nk2;
  const int k = nk2 - (int)blockIdx.z;  /* planes nk2, nk2-1, nk2-2, nk2-3 */
  if (i >= g.ni + 3 || j >= g.nj + 3 || k < 3) return;
  if (pm.fs.nfault > 0 && fault_column(pm.fs, i, j + (long)k * g.ny)) return;
  const long sc = g.sicmp, sy = g.siy, sz = g.siz;
  const long p = i + j * sy + (long)k * sz;
  float c[9];
#pragma unroll
  for (int f = 0; f < 9; f++) c[f] = beta * E[f * sc + p] + (X ? alpha * X[f * sc + p] : 0.0f);
  const float l = lam[p], m = mu[p], s = slw[p];
  constexpr int fxo = ID ? -1 : -3, fyo = JD ? -1 : -3, fzo = KD ? -1 : -3;
  const int tf[3][3] = {{TXX, TXY, TXZ}, {TXY, TYY, TYZ}, {TXZ, TYZ, TZZ}};

  if (k >= nk2 - (fzo + 4)) {
    /* timg: h_c = s/J(p) sum_taps coef J(q) (xi_d(q) . T_c(q)), odd image above the surface */
    const int n_free = nk2 - k - fzo;
    const float sj = s / met[p];
#pragma unroll 1
    for (int cc = 0; cc < 3; cc++) {
      const float w = c[cc] * sj;
#pragma unroll
      for (int n = 0; n < 5; n++) {
        {
          const int ti = i + fxo + n;
          if (phys(g, ti, j, k)) {
            const long q = p + fxo + n;
            const float u = w * tap_coef<ID>(n) * met[q];
#pragma unroll
            for (int a = 0; a < 3; a++) atomicAdd(T + tf[cc][a] * sc + q, u * met[(1 + a) * sc + q]);
          }
        }
        {
          const int tj = j + fyo + n;
          if (phys(g, i, tj, k)) {
            const long q = p + (long)(fyo + n) * sy;
            const float u = w * tap_coef<JD>(n) * met[q];
#pragma unroll
            for (int a = 0; a < 3; a++) atomicAdd(T + tf[cc][a] * sc + q, u * met[(4 + a) * sc + q]);
          }
        }
        if (n == n_free) continue;
        const int dz = n < n_free ? fzo + n : fzo + n - 2 * (n - n_free);
        const float sign = n < n_free ? 1.0f : -1.0f;
        if (phys(g, i, j, k + dz)) {
          const long q = p + (long)dz * sz;
          const float u = sign * w * tap_coef<KD>(n) * met[q];
#pragma unroll
          for (int a = 0; a < 3; a++) atomicAdd(T + tf[cc][a] * sc + q, u * met[(7 + a) * sc + q]);
        }
      }
    }
  }

  if (k >= nk2 - 2) {
    /* vlow: Hooke's law on dV with x/y one-sided operators and, in z, the
       free-surface matrices (k = nk2), mac22 (nk2-1) or mac24 (nk2-2) */
    const float l2m = (float)(l + 2.0 * m);
    float gm[3][3];  /* adjoint of the physical gradient dV_c/dx_a */
    gm[0][0] = l2m * c[TXX] + l * c[TYY] + l * c[TZZ];
    gm[0][1] = m * c[TXY];
    gm[0][2] = m * c[TXZ];
    gm[1][0] = m * c[TXY];
    gm[1][1] = l * c[TXX] + l2m * c[TYY] + l * c[TZZ];
    gm[1][2] = m * c[TYZ];
    gm[2][0] = m * c[TXZ];
    gm[2][1] = m * c[TYZ];
    gm[2][2] = l * c[TXX] + l * c[TYY] + l2m * c[TZZ];
    float dl[3][3];  /* adjoint of D_d V_c */
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int cc = 0; cc < 3; cc++)
        dl[d][cc] = met[(1 + 3 * d) * sc + p] * gm[cc][0] + met[(2 + 3 * d) * sc + p] * gm[cc][1] +
                    met[(3 + 3 * d) * sc + p] * gm[cc][2];
    if (k == nk2) {
      const long ij = ((long)i + (long)j * sy) * 9;
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        float ax = 0.0f, ay = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; r++) {
          ax += mx[ij + 3 * r + cc] * dl[2][r];
          ay += my[ij + 3 * r + cc] * dl[2][r];
        }
        dl[0][cc] += ax;
        dl[1][cc] += ay;
      }
    } else {
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        float *Tc = T + cc * sc;
        const float w = dl[2][cc];
        if (k == nk2 - 1) {  /* mac22 */
          if (KD) { atomicAdd(Tc + p + sz, w); atomicAdd(Tc + p, -w); }
          else { atomicAdd(Tc + p, w); atomicAdd(Tc + p - sz, -w); }
        } else {             /* mac24 */
          const float b1 = (float)kB1, b2 = (float)kB2, b3 = (float)kB3;
          if (KD) { atomicAdd(Tc + p, b1 * w); atomicAdd(Tc + p + sz, b2 * w); atomicAdd(Tc + p + 2 * sz, b3 * w); }
          else { atomicAdd(Tc + p, -b1 * w); atomicAdd(Tc + p - sz, -b2 * w); atomicAdd(Tc + p - 2 * sz, -b3 * w); }
        }
      }
    }
#pragma unroll
    for (int cc = 0; cc < 3; cc++)
#pragma unroll
      for (int t = 0; t < 5; t++) {
        if (phys(g, i + fxo + t, j, k)) atomicAdd(T + cc * sc + p + fxo + t, tap_coef<ID>(t) * dl[0][cc]);
        if (phys(g, i, j + fyo + t, k)) atomicAdd(T + cc * sc + p + (long)(fyo + t) * sy, tap_coef<JD>(t) * dl[1][cc]);
      }
  }

  if (k == nk2 && pm.on) {
    /* the free-surface term of the x/y PML faces: qT from D_dim V through
       the matrices, added to the stress rhs with (B-1) and to the auxiliary
       rhs with D (sv_curv_col_el_iso_gpu.cu:1000-1060, :1146-1205) */
    const float l2m = l + 2.0f * m;
    const float zx = met[7 * sc + p], zy = met[8 * sc + p], zz = met[9 * sc + p];
#pragma unroll 1
    for (int fc = 0; fc < 4; fc++) {
      const AdjFace &F = pm.f[fc];
      if (!in_face(F, i, j, k)) continue;
      const long ia = face_idx(F, i, j, k), n = F.n;
      const int ic = F.dim == 0 ? i - F.i1 : j - F.j1;
      const float cB = F.B[ic], cD = F.D[ic];
      const float *Ea = F.lev + (long)pm.le * 9 * n;
      const float *Xa = MODE == 3 ? nullptr : (MODE == 1 ? F.lev + 18 * n : F.lev + 9 * n);
      float w[9];
#pragma unroll
      for (int f = 3; f < 9; f++) {
        const long o2 = f * n + ia;
        const float ca_ = beta * Ea[o2] + (Xa ? alpha * Xa[o2] : 0.0f);
        w[f] = (float)(cB - 1.0) * c[f] + cD * ca_;
      }
      /* qT -> q */
      const float q0 = l2m * zx * w[TXX] + l * zx * w[TYY] + l * zx * w[TZZ] + m * zy * w[TXY] + m * zz * w[TXZ];
      const float q1 = l * zy * w[TXX] + l2m * zy * w[TYY] + l * zy * w[TZZ] + m * zx * w[TXY] + m * zz * w[TYZ];
      const float q2 = l * zz * w[TXX] + l * zz * w[TYY] + l2m * zz * w[TZZ] + m * zx * w[TXZ] + m * zy * w[TYZ];
      const long ij = ((long)i + (long)j * sy) * 9;
      const float *Xm = (F.dim == 0 ? mx : my) + ij;
      float dv[3];
#pragma unroll
      for (int cc = 0; cc < 3; cc++) dv[cc] = Xm[cc] * q0 + Xm[3 + cc] * q1 + Xm[6 + cc] * q2;
#pragma unroll
      for (int cc = 0; cc < 3; cc++)
#pragma unroll
        for (int t = 0; t < 5; t++) {
          if (F.dim == 0) {
            if (phys(g, i + fxo + t, j, k)) atomicAdd(T + cc * sc + p + fxo + t, tap_coef<ID>(t) * dv[cc]);
          } else {
            if (phys(g, i, j + fyo + t, k)) atomicAdd(T + cc * sc + p + (long)(fyo + t) * sy, tap_coef<JD>(t) * dv[cc]);
          }
        }
    }
  }
}

template <int ID, int JD, int KD>
static void adj_top_dir(int mode, const Grid &g, const float *met, const float *lam, const float *mu,
                        const float *slw, const float *mx, const float *my, const float *X, float alpha,
                        const float *E, float beta, float *T, const AdjPml &pm, hipStream_t st) {
  dim3 blk(64, 4, 1), grd((g.ni + 63) / 64, (g.nj + 3) / 4, g.nk < 4 ? g.nk : 4);
  switch (mode) {
    case 3:
      hipLaunchKernelGGL((k_adj_top<ID, JD, KD, 3>), grd, blk, 0, st, g, met, lam, mu, slw, mx, my, X, alpha, E,
                         beta, T, pm);
      break;
    case 2:
      hipLaunchKernelGGL((k_adj_top<ID, JD, KD, 2>), grd, blk, 0, st, g, met, lam, mu, slw, mx, my, X, alpha, E,
                         beta, T, pm);
      break;
    case 1:
      hipLaunchKernelGGL((k_adj_top<ID, JD, KD, 1>), grd, blk, 0, st, g, met, lam, mu, slw, mx, my, X, alpha, E,
                         beta, T, pm);
      break;
    default:
      hipLaunchKernelGGL((k_adj_top<ID, JD, KD, 0>), grd, blk, 0, st, g, met, lam, mu, slw, mx, my, X, alpha, E,
                         beta, T, pm);
      break;
  }
}

void launch_adj_top(int mode, int idir, int jdir, int kdir, const Grid &g, const float *met, const float *lam,
                    const float *mu, const float *slw, const float *mx, const float *my, const float *X, float alpha,
                    const float *E, float beta, float *T, const AdjPml &pm, hipStream_t st) {
  switch (idir * 4 + jdir * 2 + kdir) {
    case 0: adj_top_dir<0, 0, 0>(mode, g, met, lam, mu, slw, mx, my, X, alpha, E, beta, T, pm, st); break;
    case 1: adj_top_dir<0, 0, 1>(mode, g, met, lam, mu, slw, mx, my, X, alpha, E, beta, T, pm, st); break;
    case 2: adj_top_dir<0, 1, 0>(mode, g, met, lam, mu, slw, mx, my, X, alpha, E, beta, T, pm, st); break;
    case 3: adj_top_dir<0, 1, 1>(mode, g, met, lam, mu, slw, mx, my, X, alpha, E, beta, T, pm, st); break;
    case 4: adj_top_dir<1, 0, 0>(mode, g, met, lam, mu, slw, mx, my, X, alpha, E, beta, T, pm, st); break;
    case 5: adj_top_dir<1, 0, 1>(mode, g, met, lam, mu, slw, mx, my, X, alpha, E, beta, T, pm, st); break;
    case 6: adj_top_dir<1, 1, 0>(mode, g, met, lam, mu, slw, mx, my, X, alpha, E, beta, T, pm, st); break;
    default: adj_top_dir<1, 1, 1>(mode, g, met, lam, mu, slw, mx, my, X, alpha, E, beta, T, pm, st); break;
  }
}

/* P += residual row n at the stations (all nine components), the transpose
   of the recording: each term's node takes the residual times the term's
   weight fx fy fz (a station on a node: one term of weight 1) */
__global__ void k_adj_inject(const long *__restrict__ term, const float *__restrict__ fac, int nterm, int nsta,
                             const float *__restrict__ res, int nt, int n, long sicmp, float *P) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsta * 9 * nterm) return;
  const int u = t % nterm, sc = t / nterm;
  const int s = sc / 9, c = sc % 9;
  const float *f = fac + ((long)s * nterm + u) * 3;
  const float w = (f[0] * f[1]) * f[2];
  if (w == 0.0f) return;
  atomicAdd(P + c * sicmp + term[(long)s * nterm + u], res[((long)s * nt + n) * 9 + c] * w);
}

void launch_adj_prep(int mode, const Grid &g, const float *lam, const float *mu, const float *slw, const float *X,
                     float alpha, const float *E, float beta, float *SG, const AdjPml &pm, int isfree, int kdir,
                     hipStream_t st) {
  dim3 blk(64, 4, 1), grd((g.ni + 63) / 64, (g.nj + 3) / 4, g.nk);
  switch (mode) {
    case 3:
      hipLaunchKernelGGL(k_adj_prep<3>, grd, blk, 0, st, g, lam, mu, slw, X, alpha, E, beta, SG, pm, isfree, kdir);
      break;
    case 2:
      hipLaunchKernelGGL(k_adj_prep<2>, grd, blk, 0, st, g, lam, mu, slw, X, alpha, E, beta, SG, pm, isfree, kdir);
      break;
    case 1:
      hipLaunchKernelGGL(k_adj_prep<1>, grd, blk, 0, st, g, lam, mu, slw, X, alpha, E, beta, SG, pm, isfree, kdir);
      break;
    default:
      hipLaunchKernelGGL(k_adj_prep<0>, grd, blk, 0, st, g, lam, mu, slw, X, alpha, E, beta, SG, pm, isfree, kdir);
      break;
  }
}

template <int ID, int JD, int KD, bool VC>
static void adj_stage_dir(int mode, const Grid &g, const float *met, const float *SG, const float *E, float *A,
                          float *B, float *P, const AdjPml &pm, hipStream_t st) {
  dim3 blk(64, 4, 1), grd((g.ni + 63) / 64, (g.nj + 3) / 4, g.nk);
  switch (mode) {
    case 3: hipLaunchKernelGGL((k_adj_stage<ID, JD, KD, 3, VC>), grd, blk, 0, st, g, met, SG, E, A, B, P, pm); break;
    case 2: hipLaunchKernelGGL((k_adj_stage<ID, JD, KD, 2, VC>), grd, blk, 0, st, g, met, SG, E, A, B, P, pm); break;
    case 1: hipLaunchKernelGGL((k_adj_stage<ID, JD, KD, 1, VC>), grd, blk, 0, st, g, met, SG, E, A, B, P, pm); break;
    default: hipLaunchKernelGGL((k_adj_stage<ID, JD, KD, 0, VC>), grd, blk, 0, st, g, met, SG, E, A, B, P, pm); break;
  }
}

template <bool VC>
static void adj_stage_vc(int mode, int idir, int jdir, int kdir, const Grid &g, const float *met, const float *SG,
                         const float *E, float *A, float *B, float *P, const AdjPml &pm, hipStream_t st) {
  switch (idir * 4 + jdir * 2 + kdir) {
    case 0: adj_stage_dir<0, 0, 0, VC>(mode, g, met, SG, E, A, B, P, pm, st); break;
    case 1: adj_stage_dir<0, 0, 1, VC>(mode, g, met, SG, E, A, B, P, pm, st); break;
    case 2: adj_stage_dir<0, 1, 0, VC>(mode, g, met, SG, E, A, B, P, pm, st); break;
    case 3: adj_stage_dir<0, 1, 1, VC>(mode, g, met, SG, E, A, B, P, pm, st); break;
    case 4: adj_stage_dir<1, 0, 0, VC>(mode, g, met, SG, E, A, B, P, pm, st); break;
    case 5: adj_stage_dir<1, 0, 1, VC>(mode, g, met, SG, E, A, B, P, pm, st); break;
    case 6: adj_stage_dir<1, 1, 0, VC>(mode, g, met, SG, E, A, B, P, pm, st); break;
    default: adj_stage_dir<1, 1, 1, VC>(mode, g, met, SG, E, A, B, P, pm, st); break;
  }
}

void launch_adj_stage(int mode, int idir, int jdir, int kdir, const Grid &g, const float *met, const float *SG,
                      const float *E, float *A, float *B, float *P, const AdjPml &pm, int vcol, hipStream_t st) {
  if (vcol) adj_stage_vc<true>(mode, idir, jdir, kdir, g, met, SG, E, A, B, P, pm, st);
  else adj_stage_vc<false>(mode, idir, jdir, kdir, g, met, SG, E, A, B, P, pm, st);
}

void launch_adj_inject(const long *term, const float *fac, int nterm, int nsta, const float *res, int nt, int n,
                       long sicmp, float *P, hipStream_t st) {
  const int nthr = nsta * 9 * nterm;
  if (nthr <= 0) return;
  hipLaunchKernelGGL(k_adj_inject, dim3((nthr + 255) / 256), dim3(256), 0, st, term, fac, nterm, nsta, res, nt, n,
                     sicmp, P);
}

}  // namespace cgfd
