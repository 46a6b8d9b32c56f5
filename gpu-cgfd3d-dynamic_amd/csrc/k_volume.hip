/*
 * k_volume.hip -- fused volume kernels of one RK stage.
 *
 * One launch computes, for every point it owns, what the reference computes
 * in four passes per stage:
 *   1. the interior velocity-stress rhs (sv_curv_col_el_iso_rhs_inner_gpu,
 *      sv_curv_col_el_iso_gpu.cu:231-388);
 *   2. on the top four layers with a free surface, the traction-image hV
 *      (timg, :398-600) and the reduced-order hT (vlow, :606-747);
 *   3. the CFS-PML correction and auxiliary rhs of every face containing the
 *      point (:824-1290), in the reference's face order;
 *   4. the RK4 stage update of the wavefield and PML auxiliaries
 *      (wav_update/_end, wav_t.cu:218-234, driven by drv_rk_curv_col.cu:329-568)
 * so the rhs never goes to HBM. Stage input (cur) and stage output (tnew) are
 * different buffers, which makes the in-place stencil hazard of a fused
 * update impossible. Points on the seven fault columns of non-welded fault
 * nodes are left to the fault kernel, which overwrites their rhs in the
 * reference (sv_curv_col_el_iso_fault_gpu.cu:11-189).
 */
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "cgfd_dev.h"
#include "cgfd_kernels.h"
#include "k_vol_common.h"

namespace cgfd {

/* one point of the direct kernels: the shared metric and media are loaded
   once, then for every active scenario 27 derivatives from the cache
   hierarchy, rhs + PML + RK */
/* FENCE: each field's three derivatives are materialised before the next
   field's taps are loaded (fewer loads in flight, far fewer registers) */
template <int ID, int JD, int KD, bool PML = true, bool FENCE = false>
DEVI void direct_point(const VolArgs &a, const BatArgs &b, int i, int j, int k, int sn) {
  const long sc = a.g.sicmp, sy = a.g.siy, sz = a.g.siz;
  const long p = i + j * sy + (long)k * sz;
  const unsigned o = (unsigned)p * 4u;
  const unsigned by = (unsigned)sy * 4u, bz = (unsigned)sz * 4u;
  const float *M = a.met;
  const long woff = (long)sn * b.wstride;
  float Dx[9], Dy[9], Dz[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float *bc = a.cur + woff + c * sc;
    Dx[c] = dfd32<ID>(bc, o, 4u);
    Dy[c] = dfd32<JD>(bc, o, by);
    Dz[c] = dfd32<KD>(bc, o, bz);
    if (FENCE) __asm__ volatile("" : "+v"(Dx[c]), "+v"(Dy[c]), "+v"(Dz[c]) : : "memory");
  }
  const float xx = ld32(M + sc, o), xy = ld32(M + 2 * sc, o), xz = ld32(M + 3 * sc, o);
  const float ex = ld32(M + 4 * sc, o), ey = ld32(M + 5 * sc, o), ez = ld32(M + 6 * sc, o);
  const float zx = ld32(M + 7 * sc, o), zy = ld32(M + 8 * sc, o), zz = ld32(M + 9 * sc, o);
  const float l = ld32(a.lam, o), m = ld32(a.mu, o), s = ld32(a.slw, o);
  float h[9];
  rhs_interior(Dx, Dy, Dz, xx, xy, xz, ex, ey, ez, zx, zy, zz, l, m, s, h);
  if (PML && a.pml_on) pml_point(a, i, j, k, p, Dx, Dy, Dz, l, m, s, h, (long)sn * b.pstride);
  if (a.fs.nfault > 0 && fault_column(a.fs, i, j + (long)k * a.g.ny)) return;
  rk_point32(a, o, h, woff);
}

/* ------------------------------------------------------------------------
 * Batched volume kernel: one 64 x 4 tile of one plane for one scenario per
 * block. Block ids map XCD-contiguously onto (tile, scenario) pairs with the
 * scenario fastest, so the blocks of all scenarios of a tile run together on
 * one XCD and the tile's metric and media lines come from HBM once for the
 * whole batch.
 * ---------------------------------------------------------------------- */
template <int ID, int JD, int KD>
__global__ void __launch_bounds__(256) k_vol_bat(VolArgs a, BatArgs b, long nitems) {
  const long L = blockIdx.x;
  const long per = (long)(gridDim.x / 8);
  const long item = (L % 8) * per + L / 8;
  if (item >= nitems) return;
  const int q = (int)(item % b.nact);
  const long tile = item / b.nact;
  const int ntx = (a.g.ni + 3 + 63) / 64, nty = (a.g.nj + 3) / 4;
  const int bx = (int)(tile % ntx);
  const int by = (int)((tile / ntx) % nty);
  const int kz = (int)(tile / ((long)ntx * nty));
  const int i = bx * 64 + (int)threadIdx.x, iy = by * 4 + (int)threadIdx.y;  /* i ghost-inclusive */
  if (i < 3 || i >= a.g.ni + 3 || iy >= a.g.nj) return;
  direct_point<ID, JD, KD>(a, b, i, iy + 3, a.k_lo + kz + 3, b.act[q]);
}

/* ------------------------------------------------------------------------
 * Split direct kernels: blocks whose tile-plane touches no CFS-PML slab run
 * an interior kernel that turns each field's three derivatives into its
 * physical gradient at once and accumulates the rhs (the sums of
 * sv_curv_col_el_iso_gpu.cu:349-386 grouped per field), so only a handful of
 * values stay live between fields and the kernel fits many waves per SIMD;
 * the other blocks run the full per-point path with CFS-PML. Both launches
 * cover the whole grid (same block order as the direct kernel) and leave
 * the other class's blocks at once.
 * ---------------------------------------------------------------------- */
DEVI bool block_touches_pml(const VolArgs &a, int i0, int i1, int j0, int j1, int k) {
  if (!a.pml_on) return false;
  for (int d = 0; d < 3; d++)
    for (int e = 0; e < 2; e++) {
      const PmlFace &f = a.pml[d][e];
      if (f.on && i1 >= f.i1 && i0 <= f.i2 && j1 >= f.j1 && j0 <= f.j2 && k >= f.k1 && k <= f.k2) return true;
    }
  return false;
}

/* The per-point interior update with each field's three derivatives from
   `deriv(f, dx, dy, dz)`. G > 0: the scheduler may not move instructions
   across the end of every G-th field, which bounds the loads in flight (and
   the registers) per wave.
   VC: the grid's columns are vertical (x depends on i alone, y on j alone):
   xi_y, xi_z, eta_x and eta_z vanish, and the metric computed from such a
   grid holds there only the rounding of its finite differences (about
   1e-15 of xi_x and eta_y; the core checks on the device at create that
   they stay below 2^-32 of them at every point). Those four arrays are not
   read and their products not formed. The gradients are explicit fused
   chains in which a zero term would drop out exactly; the omitted products
   lie 2^-8 of an ulp or more below the kept ones, so both forms agree
   to the last bit at nearly every point and within an ulp elsewhere. */
template <int G, bool VC, class Deriv>
DEVI void acc_core(const VolArgs &a, const BatArgs &b, int i, int j, int k, int sn, Deriv deriv) {
  const long sc = a.g.sicmp, sy = a.g.siy, sz = a.g.siz;
  const long p = i + j * sy + (long)k * sz;
  const unsigned o = (unsigned)p * 4u;
  const float *M = a.met;
  const long woff = (long)sn * b.wstride;
  const float xx = ld32nt(M + sc, o), ey = ld32nt(M + 5 * sc, o);
  const float xy = VC ? 0.0f : ld32nt(M + 2 * sc, o), xz = VC ? 0.0f : ld32nt(M + 3 * sc, o);
  const float ex = VC ? 0.0f : ld32nt(M + 4 * sc, o), ez = VC ? 0.0f : ld32nt(M + 6 * sc, o);
  const float zx = ld32nt(M + 7 * sc, o), zy = ld32nt(M + 8 * sc, o), zz = ld32nt(M + 9 * sc, o);
  float gvx[3], gvy[3], gvz[3];
  float hv0 = 0.f, hv1 = 0.f, hv2 = 0.f;
#pragma unroll
  for (int f = 0; f < 9; f++) {
    float dx, dy, dz;
    deriv(f, o, dx, dy, dz);
    float g0, g1, g2;  /* d/dx, d/dy, d/dz */
    if (VC) {
      g0 = __builtin_fmaf(zx, dz, xx * dx);
      g1 = __builtin_fmaf(zy, dz, ey * dy);
      g2 = zz * dz;
    } else {
      g0 = __builtin_fmaf(zx, dz, __builtin_fmaf(ex, dy, xx * dx));
      g1 = __builtin_fmaf(zy, dz, __builtin_fmaf(ey, dy, xy * dx));
      g2 = __builtin_fmaf(zz, dz, __builtin_fmaf(ez, dy, xz * dx));
    }
    if (G > 0 && (f + 1) % G == 0) {
      /* the gradient exists in registers here and later loads stay below:
         this field's taps are consumed before the next field's are issued */
      __asm__ volatile("" : "+v"(g0), "+v"(g1), "+v"(g2) : : "memory");
    }
    switch (f) {
      case VX: gvx[0] = g0; gvx[1] = g1; gvx[2] = g2; break;
      case VY: gvy[0] = g0; gvy[1] = g1; gvy[2] = g2; break;
      case VZ: gvz[0] = g0; gvz[1] = g1; gvz[2] = g2; break;
      case TXX: hv0 += g0; break;
      case TYY: hv1 += g1; break;
      case TZZ: hv2 += g2; break;
      case TYZ: hv1 += g2; hv2 += g1; break;
      case TXZ: hv0 += g2; hv2 += g0; break;
      case TXY: hv0 += g1; hv1 += g0; break;
    }
  }
  if (a.fs.nfault > 0 && fault_column(a.fs, i, j + (long)k * a.g.ny)) return;
  const float l = ld32nt(a.lam, o), m = ld32nt(a.mu, o), s = ld32nt(a.slw, o);
  const float l2m = l + 2.0f * m;
  float h[9];
  h[VX] = s * hv0;
  h[VY] = s * hv1;
  h[VZ] = s * hv2;
  h[TXX] = l2m * gvx[0] + l * (gvy[1] + gvz[2]);
  h[TYY] = l2m * gvy[1] + l * (gvx[0] + gvz[2]);
  h[TZZ] = l2m * gvz[2] + l * (gvx[0] + gvy[1]);
  h[TXY] = m * (gvx[1] + gvy[0]);
  h[TXZ] = m * (gvx[2] + gvz[0]);
  h[TYZ] = m * (gvy[2] + gvz[1]);
  rk_point32(a, o, h, woff);
}

/* interior point with every tap loaded through the cache hierarchy */
template <int ID, int JD, int KD, int G, bool VC = false>
DEVI void acc_point(const VolArgs &a, const BatArgs &b, int i, int j, int k, int sn) {
  const unsigned by = (unsigned)a.g.siy * 4u, bz = (unsigned)a.g.siz * 4u;
  const float *cur = a.cur + (long)sn * b.wstride;
  const long sc = a.g.sicmp;
  acc_core<G, VC>(a, b, i, j, k, sn, [&](int f, unsigned o, float &dx, float &dy, float &dz) {
    const float *bc = cur + f * sc;
    dx = dfd32<ID>(bc, o, 4u);
    dy = dfd32<JD>(bc, o, by);
    dz = dfd32<KD>(bc, o, bz);
  });
}

/* block id -> (x tile, y tile, plane) in slab order: x tile fastest, then
   the S = zchunk y tiles of a slab, then the plane, then the slab, so the
   blocks resident at one time cover a few planes of one slab of 4 S rows and
   read each other's z and y halos from L2. false: padding item past the last
   slab. */
DEVI bool slab_item(const VolArgs &a, int &bx, int &by, int &kz) {
  const int ntx = (a.g.ni + 3 + 63) / 64, nty = (a.g.nj + 3) / 4;
  const int S = a.zchunk;
  long L = blockIdx.x;
  bx = (int)(L % ntx);
  L /= ntx;
  const int cy = (int)(L % S);
  L /= S;
  kz = (int)(L % a.nk_run);
  by = (int)(L / a.nk_run) * S + cy;
  return by < nty;
}

/* the interior kernel in slab order: 8 waves per SIMD, a register fence
   after every field */
template <int ID, int JD, int KD>
__global__ void __launch_bounds__(256, 8) k_vol_acc_slab(VolArgs a, BatArgs b) {
  int bx, by, kz;
  if (!slab_item(a, bx, by, kz)) return;
  const int k = a.k_lo + kz + 3;
  const int bi0 = bx * 64, bj0 = by * 4 + 3;
  if (block_touches_pml(a, bi0, bi0 + 63, bj0, bj0 + 3, k)) return;
  const int i = bi0 + (int)threadIdx.x, iy = by * 4 + (int)threadIdx.y;
  if (i < 3 || i >= a.g.ni + 3 || iy >= a.g.nj) return;
  acc_point<ID, JD, KD, 1>(a, b, i, iy + 3, k, b.act[0]);
}

/* ------------------------------------------------------------------------
 * Lean CFS-PML point. The correction of face `dim` needs the part of every
 * rhs component that comes from the derivatives along `dim` only (r below;
 * the r[] of sv_curv_col_el_iso_gpu.cu:916-1288). Those parts are
 * accumulated field by field, like the interior kernel's gradient sums, with
 * a register fence per field, so only 27 partial sums, the metric and the
 * media stay live; the full rhs is their sum over the three directions.
 * Then each face containing the point applies its correction and advances
 * its auxiliary variables in the reference's face order. The free-surface
 * term of the x/y faces (k == nk2) never occurs here: with a free surface
 * the top four layers belong to k_vol_top.
 * ---------------------------------------------------------------------- */
/* VC: vertical columns (as in acc_core): e[0][1], e[0][2], e[1][0] and
   e[1][2] vanish, are not read, and their terms are left out of the sums */
template <bool VC>
DEVI constexpr bool met_zero(int d, int c) {
  return VC && ((d == 0 && c != 0) || (d == 1 && c != 1));
}

/* DM: the directions whose CFS-PML faces can contain the point (bit d for
   direction d; the caller's region guarantees no face of another direction
   does). The directional rhs parts are kept for those directions only;
   the others are summed into one remaining part. */
template <int DM>
DEVI constexpr int pml_nslot() {
  return ((DM & 1) + ((DM >> 1) & 1) + ((DM >> 2) & 1)) + (DM == 7 ? 0 : 1);
}
template <int DM>
DEVI constexpr int pml_slot(int d) {
  int s = 0;
  for (int e = 0; e < d; e++) s += (DM >> e) & 1;
  return ((DM >> d) & 1) ? s : pml_nslot<DM>() - 1;
}

template <bool VC, int DM, class Deriv>
DEVI void pml_core(const VolArgs &a, const BatArgs &b, int i, int j, int k, int sn, Deriv deriv) {
  constexpr int NS = pml_nslot<DM>();
  const long sc = a.g.sicmp, sy = a.g.siy, sz = a.g.siz;
  const long p = i + j * sy + (long)k * sz;
  const unsigned o = (unsigned)p * 4u;
  const float *M = a.met;
  const long woff = (long)sn * b.wstride;
  float e[3][3];  /* rows: xi, eta, zeta derivatives (x, y, z components) */
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int c = 0; c < 3; c++) e[d][c] = met_zero<VC>(d, c) ? 0.0f : ld32nt(M + (1 + 3 * d + c) * sc, o);
  const float l = ld32nt(a.lam, o), m = ld32nt(a.mu, o), s = ld32nt(a.slw, o);
  const float l2m = l + 2.0f * m;
  float r[NS][9];
#pragma unroll
  for (int d = 0; d < NS; d++)
#pragma unroll
    for (int c = 0; c < 9; c++) r[d][c] = 0.0f;
#pragma unroll
  for (int f = 0; f < 9; f++) {
    float D0, D1, D2;
    deriv(f, o, D0, D1, D2);
    __asm__ volatile("" : "+v"(D0), "+v"(D1), "+v"(D2) : : "memory");
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const float Dd = d == 0 ? D0 : (d == 1 ? D1 : D2);
      const float e1 = e[d][0], e2 = e[d][1], e3 = e[d][2];
      const bool n1 = !met_zero<VC>(d, 0), n2 = !met_zero<VC>(d, 1), n3 = !met_zero<VC>(d, 2);
      float *const rd = r[pml_slot<DM>(d)];
      switch (f) {
        case TXX: if (n1) rd[VX] += s * e1 * Dd; break;
        case TYY: if (n2) rd[VY] += s * e2 * Dd; break;
        case TZZ: if (n3) rd[VZ] += s * e3 * Dd; break;
        case TXY: if (n2) rd[VX] += s * e2 * Dd; if (n1) rd[VY] += s * e1 * Dd; break;
        case TXZ: if (n3) rd[VX] += s * e3 * Dd; if (n1) rd[VZ] += s * e1 * Dd; break;
        case TYZ: if (n3) rd[VY] += s * e3 * Dd; if (n2) rd[VZ] += s * e2 * Dd; break;
        case VX:
          if (n1) { rd[TXX] += l2m * e1 * Dd; rd[TYY] += l * e1 * Dd; rd[TZZ] += l * e1 * Dd; }
          if (n2) rd[TXY] += m * e2 * Dd;
          if (n3) rd[TXZ] += m * e3 * Dd;
          break;
        case VY:
          if (n2) { rd[TXX] += l * e2 * Dd; rd[TYY] += l2m * e2 * Dd; rd[TZZ] += l * e2 * Dd; }
          if (n1) rd[TXY] += m * e1 * Dd;
          if (n3) rd[TYZ] += m * e3 * Dd;
          break;
        case VZ:
          if (n3) { rd[TXX] += l * e3 * Dd; rd[TYY] += l * e3 * Dd; rd[TZZ] += l2m * e3 * Dd; }
          if (n1) rd[TXZ] += m * e1 * Dd;
          if (n2) rd[TYZ] += m * e2 * Dd;
          break;
      }
    }
  }
  float h[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    h[c] = r[0][c];
#pragma unroll
    for (int q = 1; q < NS; q++) h[c] += r[q][c];
  }
  const long poff = (long)sn * b.pstride;
#pragma unroll
  for (int dim = 0; dim < 3; dim++) {
    if (!((DM >> dim) & 1)) continue;
    const float *const rdim = r[pml_slot<DM>(dim)];
#pragma unroll 1
    for (int side = 0; side < 2; side++) {
      const PmlFace &F = a.pml[dim][side];
      if (!F.on) continue;
      if (i < F.i1 || i > F.i2 || j < F.j1 || j > F.j2 || k < F.k1 || k > F.k2) continue;
      const long ia = (long)(k - F.k1) * ((long)F.bn * F.an) + (long)(j - F.j1) * F.an + (i - F.ax0);
      const int ic = dim == 0 ? i - F.i1 : (dim == 1 ? j - F.j1 : k - F.k1);
      const float cD = F.D[ic], cA = F.A[ic], cB = F.B[ic];
      const float cB1 = cB - 1.0f;
      const float *auxc = aux_input(F, a.istage) + poff;
      float av[9], ar[9];
#pragma unroll
      for (int c = 0; c < 9; c++) {
        av[c] = auxc[c * F.n + ia];
        h[c] += cB1 * rdim[c] - cB * av[c];
        ar[c] = cD * rdim[c] - cA * av[c];
      }
      aux_update(F, a, poff, ia, av, ar);
    }
  }
  if (a.fs.nfault > 0 && fault_column(a.fs, i, j + (long)k * a.g.ny)) return;
  rk_point32(a, o, h, woff);
}

template <int ID, int JD, int KD, bool VC = false, int DM = 7>
DEVI void pml_acc_point(const VolArgs &a, const BatArgs &b, int i, int j, int k, int sn) {
  const unsigned by = (unsigned)a.g.siy * 4u, bz = (unsigned)a.g.siz * 4u;
  const float *cur = a.cur + (long)sn * b.wstride;
  const long sc = a.g.sicmp;
  pml_core<VC, DM>(a, b, i, j, k, sn, [&](int f, unsigned o, float &D0, float &D1, float &D2) {
    const float *bc = cur + f * sc;
    D0 = dfd32<ID>(bc, o, 4u);
    D1 = dfd32<JD>(bc, o, by);
    D2 = dfd32<KD>(bc, o, bz);
  });
}

/* at least 6 waves per SIMD: 71 VGPRs without spills (the compiler's own
   choice, 87 VGPRs, leaves 5 waves and runs the PML tiles 13% slower) */
template <int ID, int JD, int KD>
__global__ void __launch_bounds__(256, 6) k_vol_pml_lean(VolArgs a, BatArgs b) {
  const int k = a.k_lo + (int)blockIdx.z + 3;
  const int bi0 = blockIdx.x * 64, bj0 = blockIdx.y * 4 + 3;
  if (!block_touches_pml(a, bi0, bi0 + 63, bj0, bj0 + 3, k)) return;
  const int i = bi0 + (int)threadIdx.x, iy = blockIdx.y * 4 + (int)threadIdx.y;
  if (i < 3 || i >= a.g.ni + 3 || iy >= a.g.nj) return;
  pml_acc_point<ID, JD, KD>(a, b, i, iy + 3, k, b.act[0]);
}

/* ------------------------------------------------------------------------
 * Box split (variants 23 and 24, one scenario). The interior kernel covers
 * exactly the box of points that no CFS-PML face contains, with its x
 * bounds moved inwards to multiples of 32 columns (whole 128-byte row
 * segments), in TW x 256/TW tiles (64 x 4 for 23, 32 x 8 for 24: fewer
 * y-halo rows per point) and slab order. The band kernel covers the rest,
 * the up to six slabs around the box, in 32 x 8 tiles, so an x-PML band of
 * 21 columns fills 29 of a tile's 32 lanes and no interior column takes the
 * PML path except the few between a PML edge and the next 32-column
 * boundary (the 64-wide tiles of variant 20 send up to 43 interior columns
 * per row through it).
 * ---------------------------------------------------------------------- */
template <int ID, int JD, int KD, int TW, bool VC>
__global__ void __launch_bounds__(256, 8) k_vol_acc_box(VolArgs a, BatArgs b, VolBox x) {
  constexpr int TH = 256 / TW;
  const int S = a.zchunk;
  long L = blockIdx.x;
  const int bx = (int)(L % x.ntx);
  L /= x.ntx;
  const int cy = (int)(L % S);
  L /= S;
  const int nkb = x.k1 - x.k0;
  const int kz = (int)(L % nkb);
  const int by = (int)(L / nkb) * S + cy;
  if (by >= x.nty) return;
  const int t = (int)threadIdx.x;
  const int i = x.x0 + bx * TW + (t % TW), j = x.y0 + by * TH + t / TW;
  if (i >= x.x1 || j >= x.y1) return;
  acc_point<ID, JD, KD, 1, VC>(a, b, i, j, x.k0 + kz, b.act[0]);
}

/* block L0 of the bands -> the thread's point; false: a lane outside its
   region or a padding block past a region's last slab. Within a region the
   blocks run in slab order, so the planes whose z taps a block reads are
   still being worked on by resident blocks when it runs. */
DEVI bool band_point(const VolBands &r, long L0, int &i, int &j, int &k) {
  int q = 0;
  while (q + 1 < r.n && L0 >= r.cum[q + 1]) q++;
  long L = L0 - r.cum[q];
  const int bx = (int)(L % r.ntx[q]);
  L /= r.ntx[q];
  const int cy = (int)(L % r.sy[q]);
  L /= r.sy[q];
  const int nkq = r.k1[q] - r.k0[q];
  const int kz = (int)(L % nkq);
  const int by = (int)(L / nkq) * r.sy[q] + cy;
  const int t = (int)threadIdx.x;
  i = r.ox[q] + bx * 32 + (t & 31);
  j = r.y0[q] + by * 8 + (t >> 5);
  k = r.k0[q] + kz;
  return by < r.nty[q] && i >= r.x0[q] && i < r.x1[q] && j < r.y1[q];
}

/* 7 waves per SIMD: 72 VGPRs and 8-24 bytes of spills per lane; 1.3% faster
   than the 6 waves the compiler picks without the bound (75 VGPRs), while
   8 waves (64 VGPRs, 40-56 bytes of spills) lose it again */
template <int ID, int JD, int KD, bool VC, int DM>
__global__ void __launch_bounds__(256, 7) k_vol_pml_band(VolArgs a, BatArgs b, VolBands r) {
  int i, j, k;
  if (!band_point(r, blockIdx.x, i, j, k)) return;
  pml_acc_point<ID, JD, KD, VC, DM>(a, b, i, j, k, b.act[0]);
}

/* batches on the box split: the interior box's (tile, scenario) pairs with
   the scenario fastest and block ids XCD-contiguous (the shared metric and
   media lines of a tile serve every scenario from one L2), the bands in one
   sweep per scenario (each scenario's halos come back from L2). nitems: the
   box's tiles x scenarios; the grid is padded to a multiple of 8. */
template <int ID, int JD, int KD, int TW, bool VC>
__global__ void __launch_bounds__(256, 8) k_vol_bat_acc_box(VolArgs a, BatArgs b, VolBox x, long nitems) {
  constexpr int TH = 256 / TW;
  const long L = blockIdx.x;
  const long item = (L % 8) * (long)(gridDim.x / 8) + L / 8;
  if (item >= nitems) return;
  const int q = (int)(item % b.nact);
  const long tile = item / b.nact;
  const int bx = (int)(tile % x.ntx);
  const int by = (int)((tile / x.ntx) % x.nty);
  const int kz = (int)(tile / ((long)x.ntx * x.nty));
  const int t = (int)threadIdx.x;
  const int i = x.x0 + bx * TW + (t % TW), j = x.y0 + by * TH + t / TW;
  if (i >= x.x1 || j >= x.y1) return;
  acc_point<ID, JD, KD, 1, VC>(a, b, i, j, x.k0 + kz, b.act[q]);
}

/* variant 32: the interior box in one slab-ordered sweep per scenario, as
   the bands (each scenario's stencil halos come back from L2 as on one
   scenario; the shared metric and media are read once per scenario) */
template <int ID, int JD, int KD, bool VC>
__global__ void __launch_bounds__(256, 8) k_vol_bat_acc_box_slab(VolArgs a, BatArgs b, VolBox x, long per) {
  const int S = a.zchunk;
  const int q = (int)(blockIdx.x / per);
  long L = blockIdx.x % per;
  const int bx = (int)(L % x.ntx);
  L /= x.ntx;
  const int cy = (int)(L % S);
  L /= S;
  const int nkb = x.k1 - x.k0;
  const int kz = (int)(L % nkb);
  const int by = (int)(L / nkb) * S + cy;
  if (by >= x.nty) return;
  const int t = (int)threadIdx.x;
  const int i = x.x0 + (bx << 6) + (t & 63), j = x.y0 + by * 4 + (t >> 6);
  if (i >= x.x1 || j >= x.y1) return;
  acc_point<ID, JD, KD, 1, VC>(a, b, i, j, x.k0 + kz, b.act[q]);
}

template <int ID, int JD, int KD, bool VC, int DM>
__global__ void __launch_bounds__(256, 7) k_vol_bat_pml_band(VolArgs a, BatArgs b, VolBands r) {
  const long per = r.cum[r.n];
  const int sq = (int)(blockIdx.x / per);
  int i, j, k;
  if (!band_point(r, blockIdx.x % per, i, j, k)) return;
  pml_acc_point<ID, JD, KD, VC, DM>(a, b, i, j, k, b.act[sq]);
}

constexpr int kBandSlabBlocks = 192;

/* the bands in groups by the CFS-PML face directions their points can lie
   in (the band kernel's DM): the x bands (any face), the y bands above the
   bottom and below the top z slab (y faces only), the y bands' corners in
   those slabs (y and z faces) and the z slabs inside the box's x and y
   range (z faces only); or, ungrouped, six slabs in one launch: the x
   bands, the y bands over every plane and the z slabs inside the box's x
   and y range */
constexpr int kBandGroups = 4;
constexpr int kBandDM[kBandGroups] = {7, 2, 6, 4};

/* the interior box and the bands around it for one launch's planes */
static void box_split(const VolArgs &a, int tw, bool grouped, VolBox &x, VolBands (&rg)[kBandGroups]) {
  const Grid &g = a.g;
  const int kv0 = a.k_lo + 3, kv1 = a.k_lo + 3 + a.nk_run;
  int xl = 3, xr = g.ni + 3, yl = 3, yr = g.nj + 3, kl = kv0, kr = kv1;
  if (a.pml_on) {
    const PmlFace(&P)[3][2] = a.pml;
    if (P[0][0].on) xl = std::max(xl, P[0][0].i2 + 1);
    if (P[0][1].on) xr = std::min(xr, P[0][1].i1);
    if (P[1][0].on) yl = std::max(yl, P[1][0].j2 + 1);
    if (P[1][1].on) yr = std::min(yr, P[1][1].j1);
    if (P[2][0].on) kl = std::max(kl, P[2][0].k2 + 1);
    if (P[2][1].on) kr = std::min(kr, P[2][1].k1);
  }
  xl = (xl + 31) & ~31;
  xr &= ~31;
  for (int grp = 0; grp < kBandGroups; grp++) rg[grp] = VolBands{};
  auto add = [&](int grp, int x0, int x1, int y0, int y1, int k0, int k1) {
    if (x1 <= x0 || y1 <= y0 || k1 <= k0) return;
    VolBands &r = rg[grp];
    if (r.n == kBandRegions) {  /* the splits below add at most six regions to a group */
      std::fprintf(stderr, "box_split: more than %d band regions in group %d\n", kBandRegions, grp);
      std::abort();
    }
    const int q = r.n++;
    r.x0[q] = x0; r.x1[q] = x1; r.y0[q] = y0; r.y1[q] = y1; r.k0[q] = k0; r.k1[q] = k1;
    r.ox[q] = x0 & ~31;
    r.ntx[q] = (x1 - r.ox[q] + 31) / 32;
    r.nty[q] = (y1 - y0 + 7) / 8;
    /* a plane of a slab holds at most kBandSlabBlocks blocks, so the seven
       planes a z derivative reads fit among the ~1500 blocks resident at 6
       waves per SIMD (a whole plane of the bottom band of a 600 x 1200 grid
       is 2850 blocks) */
    r.sy[q] = std::max(1, std::min(r.nty[q], kBandSlabBlocks / r.ntx[q]));
    const long nslab = (r.nty[q] + r.sy[q] - 1) / r.sy[q];
    r.cum[q + 1] = r.cum[q] + (long)r.ntx[q] * r.sy[q] * nslab * (k1 - k0);
  };
  if (xr <= xl || yr <= yl || kr <= kl) {
    x = VolBox{3, 3, 3, 3, kv0, kv0, 0, 0};
    add(0, 3, g.ni + 3, 3, g.nj + 3, kv0, kv1);
    return;
  }
  x = VolBox{xl, xr, yl, yr, kl, kr, (xr - xl + tw - 1) / tw, (yr - yl + 256 / tw - 1) / (256 / tw)};
  add(0, 3, xl, 3, g.nj + 3, kv0, kv1);
  add(0, xr, g.ni + 3, 3, g.nj + 3, kv0, kv1);
  if (!grouped) {
    add(0, xl, xr, 3, yl, kv0, kv1);
    add(0, xl, xr, yr, g.nj + 3, kv0, kv1);
    add(0, xl, xr, yl, yr, kv0, kl);
    add(0, xl, xr, yl, yr, kr, kv1);
    return;
  }
  add(1, xl, xr, 3, yl, kl, kr);
  add(1, xl, xr, yr, g.nj + 3, kl, kr);
  add(2, xl, xr, 3, yl, kv0, kl);
  add(2, xl, xr, yr, g.nj + 3, kv0, kl);
  add(2, xl, xr, 3, yl, kr, kv1);
  add(2, xl, xr, yr, g.nj + 3, kr, kv1);
  add(3, xl, xr, yl, yr, kv0, kl);
  add(3, xl, xr, yl, yr, kr, kv1);
}

template <int ID, int JD, int KD, bool VC, int G>
static void launch_band_group(const VolBands &r, bool batch, const VolArgs &a, const BatArgs &b, hipStream_t st) {
  if (r.n == 0) return;
  constexpr int DM = kBandDM[G];
  if (batch)
    hipLaunchKernelGGL((k_vol_bat_pml_band<ID, JD, KD, VC, DM>), dim3((unsigned)(r.cum[r.n] * b.nact)), dim3(256), 0,
                       st, a, b, r);
  else
    hipLaunchKernelGGL((k_vol_pml_band<ID, JD, KD, VC, DM>), dim3((unsigned)r.cum[r.n]), dim3(256), 0, st, a, b, r);
}

template <int ID, int JD, int KD, bool VC>
static void launch_bands(const VolBands (&rg)[kBandGroups], bool batch, const VolArgs &a, const BatArgs &b,
                         hipStream_t st) {
  launch_band_group<ID, JD, KD, VC, 0>(rg[0], batch, a, b, st);
  launch_band_group<ID, JD, KD, VC, 1>(rg[1], batch, a, b, st);
  launch_band_group<ID, JD, KD, VC, 2>(rg[2], batch, a, b, st);
  launch_band_group<ID, JD, KD, VC, 3>(rg[3], batch, a, b, st);
}

/* ------------------------------------------------------------------------
 * Batches on the split interior / CFS-PML kernels (variant 20). Order 0,
 * the interior: (tile, scenario) pairs with the scenario fastest,
 * XCD-contiguous, so a tile's metric and media lines serve every scenario
 * from one L2. Order 1, the PML tiles: one slab-ordered sweep per scenario
 * (scenario slowest), so each scenario's x, y and z halos and auxiliaries
 * are reused from L2 as in the single-scenario kernel. false: padding block.
 * ---------------------------------------------------------------------- */
DEVI bool bat_item(const VolArgs &a, const BatArgs &b, int order, long nitems, int &bx, int &by, int &kz,
                   int &q) {
  const int ntx = (a.g.ni + 3 + 63) / 64, nty = (a.g.nj + 3) / 4;
  long L = blockIdx.x;
  if (order == 0) {
    const long item = (L % 8) * (long)(gridDim.x / 8) + L / 8;
    if (item >= nitems) return false;
    q = (int)(item % b.nact);
    const long tile = item / b.nact;
    bx = (int)(tile % ntx);
    by = (int)((tile / ntx) % nty);
    kz = (int)(tile / ((long)ntx * nty));
    return true;
  }
  const int S = a.zchunk;
  const long nslab = (nty + S - 1) / S;
  const long per = (long)ntx * S * a.nk_run * nslab;
  q = (int)(L / per);
  if (q >= b.nact) return false;
  L %= per;
  bx = (int)(L % ntx);
  L /= ntx;
  const int cy = (int)(L % S);
  L /= S;
  kz = (int)(L % a.nk_run);
  by = (int)(L / a.nk_run) * S + cy;
  return by < nty;
}

template <int ID, int JD, int KD>
__global__ void __launch_bounds__(256, 8) k_vol_bat_acc(VolArgs a, BatArgs b, long nitems) {
  int bx, by, kz, q;
  if (!bat_item(a, b, 0, nitems, bx, by, kz, q)) return;
  const int k = a.k_lo + kz + 3;
  const int bi0 = bx * 64, bj0 = by * 4 + 3;
  if (block_touches_pml(a, bi0, bi0 + 63, bj0, bj0 + 3, k)) return;
  const int i = bi0 + (int)threadIdx.x, iy = by * 4 + (int)threadIdx.y;
  if (i < 3 || i >= a.g.ni + 3 || iy >= a.g.nj) return;
  acc_point<ID, JD, KD, 1>(a, b, i, iy + 3, k, b.act[q]);
}

template <int ID, int JD, int KD>
__global__ void __launch_bounds__(256, 6) k_vol_bat_pml(VolArgs a, BatArgs b, long nitems) {
  int bx, by, kz, q;
  if (!bat_item(a, b, 1, nitems, bx, by, kz, q)) return;
  const int k = a.k_lo + kz + 3;
  const int bi0 = bx * 64, bj0 = by * 4 + 3;
  if (!block_touches_pml(a, bi0, bi0 + 63, bj0, bj0 + 3, k)) return;
  const int i = bi0 + (int)threadIdx.x, iy = by * 4 + (int)threadIdx.y;
  if (i < 3 || i >= a.g.ni + 3 || iy >= a.g.nj) return;
  pml_acc_point<ID, JD, KD>(a, b, i, iy + 3, k, b.act[q]);
}

/* ------------------------------------------------------------------------
 * Direct-load volume kernel: one thread per point, neighbours from the
 * cache hierarchy. Handles physical planes k_lo .. k_lo+nk_run-1.
 * ---------------------------------------------------------------------- */
template <int ID, int JD, int KD>
__global__ void __launch_bounds__(256) k_vol_direct(VolArgs a, BatArgs b) {
  /* tiles start at ghost-inclusive column 0, so a 64-wide tile row is two
     whole 128-byte lines of every array */
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 3 || i >= a.g.ni + 3 || iy >= a.g.nj) return;
  const int j = iy + 3, k = a.k_lo + (int)blockIdx.z + 3;
  direct_point<ID, JD, KD>(a, b, i, j, k, b.act[0]);
}

/* ------------------------------------------------------------------------
 * Top-layer kernel (free surface): physical planes nk-4 .. nk-1, reproducing
 * the overwrite order inner -> timg -> vlow -> cfspml of the reference.
 * ---------------------------------------------------------------------- */
__device__ const int kOpIdx[2][5] = {{-3, -2, -1, 0, 1}, {-1, 0, 1, 2, 3}};

template <int ID, int JD, int KD, bool VC>
DEVI void top_point(const VolArgs &a, int i, int j, int k, long woff, long poff) {
  const float *__restrict__ const cur = a.cur + woff;
  const int nk2 = a.g.nk2;
  const long sc = a.g.sicmp, sy = a.g.siy, sz = a.g.siz;
  const long p = i + j * sy + (long)k * sz;
  float Dx[9], Dy[9], Dz[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float *q = cur + c * sc + p;
    Dx[c] = dfd<ID>(q, 1);
    Dy[c] = dfd<JD>(q, sy);
    Dz[c] = dfd<KD>(q, sz);
  }
  const float *M = a.met;
  const float *jac = M, *xi_x = M + sc, *xi_y = M + 2 * sc, *xi_z = M + 3 * sc;
  const float *et_x = M + 4 * sc, *et_y = M + 5 * sc, *et_z = M + 6 * sc;
  const float *zt_x = M + 7 * sc, *zt_y = M + 8 * sc, *zt_z = M + 9 * sc;
  /* VC: vertical columns, xi_y, xi_z, eta_x, eta_z vanish (see acc_core) */
  constexpr bool vc = VC;
  float xx = xi_x[p], xy = vc ? 0.0f : xi_y[p], xz = vc ? 0.0f : xi_z[p];
  float ex = vc ? 0.0f : et_x[p], ey = et_y[p], ez = vc ? 0.0f : et_z[p];
  float zx = zt_x[p], zy = zt_y[p], zz = zt_z[p];
  float l = a.lam[p], m = a.mu[p], s = a.slw[p];
  float h[9];
  rhs_interior(Dx, Dy, Dz, xx, xy, xz, ex, ey, ez, zx, zy, zz, l, m, s, h);
  if (a.isfree) {
    const int *fx = kOpIdx[ID], *fy = kOpIdx[JD], *fz = kOpIdx[KD];
    const int k_min = nk2 - fz[4];
    if (k >= k_min) {
      /* timg: conservative tractions with an odd image above the surface */
      const int n_free = nk2 - k - fz[0];
      const float slwjac = a.slw[p] / jac[p];
      const float *T = cur;
      const float *Ta[3][3] = {{T + TXX * sc, T + TXY * sc, T + TXZ * sc},
                               {T + TXY * sc, T + TYY * sc, T + TYZ * sc},
                               {T + TXZ * sc, T + TYZ * sc, T + TZZ * sc}};
#pragma unroll 1
      for (int c = 0; c < 3; c++) {
        const float *A = Ta[c][0], *B = Ta[c][1], *C = Ta[c][2];
        float vx[5], ve[5], vz[5];
#pragma unroll
        for (int n = 0; n < 5; n++) {
          long q = p + fx[n];
          vx[n] = vc ? jac[q] * (xi_x[q] * A[q]) : jac[q] * (xi_x[q] * A[q] + xi_y[q] * B[q] + xi_z[q] * C[q]);
          q = p + fy[n] * sy;
          ve[n] = vc ? jac[q] * (et_y[q] * B[q]) : jac[q] * (et_x[q] * A[q] + et_y[q] * B[q] + et_z[q] * C[q]);
        }
#pragma unroll
        for (int n = 0; n < 5; n++) {
          if (n < n_free) {
            long q = p + fz[n] * sz;
            vz[n] = jac[q] * (zt_x[q] * A[q] + zt_y[q] * B[q] + zt_z[q] * C[q]);
          } else if (n == n_free) {
            vz[n] = 0.0f;
          } else {
            long q = p + (long)(fz[n] - 2 * (n - n_free)) * sz;
            vz[n] = -jac[q] * (zt_x[q] * A[q] + zt_y[q] * B[q] + zt_z[q] * C[q]);
          }
        }
        h[c] = (vecdrp(vx, ID) + vecdrp(ve, JD) + vecdrp(vz, KD)) * slwjac;
      }
    }
    if (k >= nk2 - 2) {
      /* vlow: reduced-order z derivative of V near the surface */
      const float *Vx = cur + p, *Vy = cur + sc + p, *Vz = cur + 2 * sc + p;
      float dv[3][3];  /* [axis][comp] */
      dv[0][0] = macdrp(Vx, 1, ID); dv[1][0] = macdrp(Vx, sy, JD);
      dv[0][1] = macdrp(Vy, 1, ID); dv[1][1] = macdrp(Vy, sy, JD);
      dv[0][2] = macdrp(Vz, 1, ID); dv[1][2] = macdrp(Vz, sy, JD);
      if (k == nk2) {
        long ij = ((long)i + (long)j * sy) * 9;
        const float *X = a.matVx2Vz + ij, *Y = a.matVy2Vz + ij;
#pragma unroll
        for (int r = 0; r < 3; r++)
          dv[2][r] = X[3 * r] * dv[0][0] + X[3 * r + 1] * dv[0][1] + X[3 * r + 2] * dv[0][2] + Y[3 * r] * dv[1][0] +
                     Y[3 * r + 1] * dv[1][1] + Y[3 * r + 2] * dv[1][2];
      } else if (k == nk2 - 1) {
        dv[2][0] = mac22(Vx, sz, KD); dv[2][1] = mac22(Vy, sz, KD); dv[2][2] = mac22(Vz, sz, KD);
      } else {
        dv[2][0] = mac24(Vx, sz, KD); dv[2][1] = mac24(Vy, sz, KD); dv[2][2] = mac24(Vz, sz, KD);
      }
      float l2m = (float)(l + 2.0 * m);
      float DxVx = dv[0][0], DyVx = dv[1][0], DzVx = dv[2][0];
      float DxVy = dv[0][1], DyVy = dv[1][1], DzVy = dv[2][1];
      float DxVz = dv[0][2], DyVz = dv[1][2], DzVz = dv[2][2];
      h[TXX] = l2m * (xx * DxVx + ex * DyVx + zx * DzVx) + l * (xy * DxVy + ey * DyVy + zy * DzVy + xz * DxVz + ez * DyVz + zz * DzVz);
      h[TYY] = l2m * (xy * DxVy + ey * DyVy + zy * DzVy) + l * (xx * DxVx + ex * DyVx + zx * DzVx + xz * DxVz + ez * DyVz + zz * DzVz);
      h[TZZ] = l2m * (xz * DxVz + ez * DyVz + zz * DzVz) + l * (xx * DxVx + ex * DyVx + zx * DzVx + xy * DxVy + ey * DyVy + zy * DzVy);
      h[TXY] = m * (xy * DxVx + xx * DxVy + ey * DyVx + ex * DyVy + zy * DzVx + zx * DzVy);
      h[TXZ] = m * (xz * DxVx + xx * DxVz + ez * DyVx + ex * DyVz + zz * DzVx + zx * DzVz);
      h[TYZ] = m * (xz * DxVy + xy * DxVz + ez * DyVy + ey * DyVz + zz * DzVy + zy * DzVz);
    }
  }
  if (a.pml_on) pml_point(a, i, j, k, p, Dx, Dy, Dz, l, m, s, h, poff);
  if (a.fs.nfault > 0 && fault_column(a.fs, i, j + (long)k * a.g.ny)) return;
  rk_point(a, p, h, woff);
}

/* grid z = plane x scenario */
template <int ID, int JD, int KD, bool VC>
__global__ void __launch_bounds__(256) k_vol_top(VolArgs a, BatArgs b) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (ix >= a.g.ni || iy >= a.g.nj) return;
  const int kz = (int)blockIdx.z % a.nk_run;
  const long sn = b.act[blockIdx.z / a.nk_run];
  top_point<ID, JD, KD, VC>(a, ix + 3, iy + 3, a.k_lo + kz + 3, sn * b.wstride, sn * b.pstride);
}

/* ------------------------------------------------------------------------
 * free-surface conversion matrices (sv_curv_col_el_iso_dvh2dvz_gpu,
 * sv_curv_col_el_iso_gpu.cu:1296-1405)
 * ---------------------------------------------------------------------- */
__global__ void k_dvh2dvz(Grid g, const float *__restrict__ met, const float *__restrict__ lam,
                          const float *__restrict__ mu, float *matVx2Vz, float *matVy2Vz) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (ix >= g.ni || iy >= g.nj) return;
  const long sc = g.sicmp;
  long p = (ix + 3) + (long)(iy + 3) * g.siy + (long)g.nk2 * g.siz;
  float e11 = met[sc + p], e12 = met[2 * sc + p], e13 = met[3 * sc + p];
  float e21 = met[4 * sc + p], e22 = met[5 * sc + p], e23 = met[6 * sc + p];
  float e31 = met[7 * sc + p], e32 = met[8 * sc + p], e33 = met[9 * sc + p];
  float l = lam[p], m = mu[p];
  float l2m = l + 2.0f * m;
  float A[3][3], B[3][3], C[3][3], AB[3][3], AC[3][3];
  A[0][0] = l2m * e31 * e31 + m * (e32 * e32 + e33 * e33);
  A[0][1] = l * e31 * e32 + m * e32 * e31;
  A[0][2] = l * e31 * e33 + m * e33 * e31;
  A[1][0] = l * e32 * e31 + m * e31 * e32;
  A[1][1] = l2m * e32 * e32 + m * (e31 * e31 + e33 * e33);
  A[1][2] = l * e32 * e33 + m * e33 * e32;
  A[2][0] = l * e33 * e31 + m * e31 * e33;
  A[2][1] = l * e33 * e32 + m * e32 * e33;
  A[2][2] = l2m * e33 * e33 + m * (e31 * e31 + e32 * e32);
  inv3(A);
  B[0][0] = -l2m * e31 * e11 - m * (e32 * e12 + e33 * e13);
  B[0][1] = -l * e31 * e12 - m * e32 * e11;
  B[0][2] = -l * e31 * e13 - m * e33 * e11;
  B[1][0] = -l * e32 * e11 - m * e31 * e12;
  B[1][1] = -l2m * e32 * e12 - m * (e31 * e11 + e33 * e13);
  B[1][2] = -l * e32 * e13 - m * e33 * e12;
  B[2][0] = -l * e33 * e11 - m * e31 * e13;
  B[2][1] = -l * e33 * e12 - m * e32 * e13;
  B[2][2] = -l2m * e33 * e13 - m * (e31 * e11 + e32 * e12);
  C[0][0] = -l2m * e31 * e21 - m * (e32 * e22 + e33 * e23);
  C[0][1] = -l * e31 * e22 - m * e32 * e21;
  C[0][2] = -l * e31 * e23 - m * e33 * e21;
  C[1][0] = -l * e32 * e21 - m * e31 * e22;
  C[1][1] = -l2m * e32 * e22 - m * (e31 * e21 + e33 * e23);
  C[1][2] = -l * e32 * e23 - m * e33 * e22;
  C[2][0] = -l * e33 * e21 - m * e31 * e23;
  C[2][1] = -l * e33 * e22 - m * e32 * e23;
  C[2][2] = -l2m * e33 * e23 - m * (e31 * e21 + e32 * e22);
  mm33(A, B, AB);
  mm33(A, C, AC);
  long ij = ((long)(iy + 3) * g.siy + (ix + 3)) * 9;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      matVx2Vz[ij + 3 * r + c] = AB[r][c];
      matVy2Vz[ij + 3 * r + c] = AC[r][c];
    }
}

/* surface PGV/PGA/PGD (PG_calcu_gpu, wav_t.cu:137-216) */
DEVI void pg_point(const Grid &g, int ix, int iy, const float *__restrict__ wend, const float *__restrict__ wpre,
                   float *PG, float *Dis, float dt) {
  const long sc = g.sicmp, siz = g.siz;
  long p = (ix + 3) + (long)(iy + 3) * g.siy + (long)g.nk2 * siz;
  long q = (ix + 3) + (long)(iy + 3) * g.siy;
  float vx1 = wend[p], vy1 = wend[sc + p], vz1 = wend[2 * sc + p];
  float vx0 = wpre[p], vy0 = wpre[sc + p], vz0 = wpre[2 * sc + p];
  float Ax = fabsf((vx1 - vx0) / dt), Ay = fabsf((vy1 - vy0) / dt), Az = fabsf((vz1 - vz0) / dt);
  float dx = Dis[q], dy = Dis[siz + q], dz = Dis[2 * siz + q];
  dx = (float)(dx + 0.5 * (vx1 + vx0) * dt);
  dy = (float)(dy + 0.5 * (vy1 + vy0) * dt);
  dz = (float)(dz + 0.5 * (vz1 + vz0) * dt);
  Dis[q] = dx; Dis[siz + q] = dy; Dis[2 * siz + q] = dz;
  float V = sqrtf(vx1 * vx1 + vy1 * vy1 + vz1 * vz1);
  float Vh = sqrtf(vx1 * vx1 + vy1 * vy1);
  float A = sqrtf(Ax * Ax + Ay * Ay + Az * Az);
  float Ah = sqrtf(Ax * Ax + Ay * Ay);
  float D = sqrtf(dx * dx + dy * dy + dz * dz);
  float Dh = sqrtf(dx * dx + dy * dy);
  float vals[15] = {V, Vh, fabsf(vx1), fabsf(vy1), fabsf(vz1), A, Ah, Ax, Ay, Az, D, Dh, fabsf(dx), fabsf(dy), fabsf(dz)};
#pragma unroll
  for (int c = 0; c < 15; c++) {
    float o = PG[c * siz + q];
    if (o < vals[c]) PG[c * siz + q] = vals[c];
  }
}

__global__ void k_pg(Grid g, BatArgs b, const float *__restrict__ wend, const float *__restrict__ wpre, float *PG,
                     float *Dis, float dt) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (ix >= g.ni || iy >= g.nj) return;
  const long s = b.act[blockIdx.z];
  pg_point(g, ix, iy, wend + s * b.wstride, wpre + s * b.wstride, PG + s * b.pgstride, Dis + s * b.pgstride, dt);
}

/* ---------------------------------------------------------------------- */
/* host launchers */
/* the box split (variants 23, 64-wide interior tiles, 24, 32-wide, and 33,
   23 with every band in one launch): one scenario or a batch */
template <int ID, int JD, int KD, bool VC>
static void launch_box_split(int tw, bool sweeps, bool grouped, const VolArgs &a, const BatArgs &b, hipStream_t st,
                             hipStream_t st_pml) {
  VolBox x;
  VolBands r[kBandGroups];
  box_split(a, tw, grouped, x, r);
  if (b.nact > 1 && sweeps && x.ntx > 0) {  /* variant 32: per-scenario slab sweeps of the box (64 x 4 tiles) */
    const long nslab = ((long)x.nty + a.zchunk - 1) / a.zchunk;
    const long per = (long)x.ntx * a.zchunk * (x.k1 - x.k0) * nslab;
    hipLaunchKernelGGL((k_vol_bat_acc_box_slab<ID, JD, KD, VC>), dim3((unsigned)(per * b.nact)), dim3(256), 0, st, a,
                       b, x, per);
  }
  if (b.nact > 1) {
    if (x.ntx > 0 && !sweeps) {
      const long items = (long)x.ntx * x.nty * (x.k1 - x.k0) * b.nact, nb = ((items + 7) / 8) * 8;
      if (tw == 64)
        hipLaunchKernelGGL((k_vol_bat_acc_box<ID, JD, KD, 64, VC>), dim3((unsigned)nb), dim3(256), 0, st, a, b, x, items);
      else
        hipLaunchKernelGGL((k_vol_bat_acc_box<ID, JD, KD, 32, VC>), dim3((unsigned)nb), dim3(256), 0, st, a, b, x, items);
    }
    launch_bands<ID, JD, KD, VC>(r, true, a, b, st_pml);
    return;
  }
  if (x.ntx > 0) {
    const long nslab = ((long)x.nty + a.zchunk - 1) / a.zchunk;
    const long nb = (long)x.ntx * a.zchunk * (x.k1 - x.k0) * nslab;
    if (tw == 64) hipLaunchKernelGGL((k_vol_acc_box<ID, JD, KD, 64, VC>), dim3((unsigned)nb), dim3(256), 0, st, a, b, x);
    else hipLaunchKernelGGL((k_vol_acc_box<ID, JD, KD, 32, VC>), dim3((unsigned)nb), dim3(256), 0, st, a, b, x);
  }
  launch_bands<ID, JD, KD, VC>(r, false, a, b, st_pml);
}

template <int ID, int JD, int KD>
static void launch_dir(int variant, const VolArgs &a, const BatArgs &b, int top, hipStream_t st, hipStream_t st_pml) {
  dim3 blk(64, 4, 1);
  dim3 grd((a.g.ni + 63) / 64, (a.g.nj + 3) / 4, a.nk_run);
  dim3 agrd((a.g.ni + 3 + 63) / 64, (a.g.nj + 3) / 4, a.nk_run);  /* aligned tiles */
  if (top) {
    dim3 tgrd(grd.x, grd.y, (unsigned)(a.nk_run * b.nact));
    if (a.vcol) hipLaunchKernelGGL((k_vol_top<ID, JD, KD, true>), tgrd, blk, 0, st, a, b);
    else hipLaunchKernelGGL((k_vol_top<ID, JD, KD, false>), tgrd, blk, 0, st, a, b);
    return;
  }
  if (variant == 23 || variant == 24 || variant == 32 || variant == 33) {
    const int tw = variant == 24 ? 32 : 64;
    const bool grouped = variant != 33;
    if (a.vcol) launch_box_split<ID, JD, KD, true>(tw, variant == 32, grouped, a, b, st, st_pml);
    else launch_box_split<ID, JD, KD, false>(tw, variant == 32, grouped, a, b, st, st_pml);
    return;
  }
  if (b.nact > 1 && variant == 20) {  /* batches, split kernels */
    const long items0 = (long)agrd.x * agrd.y * a.nk_run * b.nact, nb0 = ((items0 + 7) / 8) * 8;
    const long nslab = ((long)agrd.y + a.zchunk - 1) / a.zchunk;
    const long items1 = (long)agrd.x * a.zchunk * a.nk_run * nslab * b.nact;
    hipLaunchKernelGGL((k_vol_bat_acc<ID, JD, KD>), dim3((unsigned)nb0), blk, 0, st, a, b, items0);
    if (a.pml_on) hipLaunchKernelGGL((k_vol_bat_pml<ID, JD, KD>), dim3((unsigned)items1), blk, 0, st_pml, a, b, items1);
    return;
  }
  if (b.nact > 1) {  /* batches: one kernel, tiles x scenarios */
    const long items = (long)agrd.x * agrd.y * a.nk_run * b.nact;
    const long nb = ((items + 7) / 8) * 8;
    hipLaunchKernelGGL((k_vol_bat<ID, JD, KD>), dim3((unsigned)nb), blk, 0, st, a, b, items);
    return;
  }
  if (variant == 20) {
    const long nslab = ((a.g.nj + 3) / 4 + a.zchunk - 1) / a.zchunk;
    const long nb = (long)agrd.x * a.zchunk * a.nk_run * nslab;
    hipLaunchKernelGGL((k_vol_acc_slab<ID, JD, KD>), dim3((unsigned)nb), blk, 0, st, a, b);
    if (a.pml_on) hipLaunchKernelGGL((k_vol_pml_lean<ID, JD, KD>), agrd, blk, 0, st_pml, a, b);
    return;
  }
  hipLaunchKernelGGL((k_vol_direct<ID, JD, KD>), agrd, blk, 0, st, a, b);
}

/* one lane per physical point of a row: x depends on i alone and y on j
   alone when xi_y, xi_z, eta_x and eta_z vanish; computed from such a grid
   they hold the rounding of the finite differences (about 1e-15 of xi_x
   and eta_y), while any tilt of the columns puts them many orders above
   2^-32 (a NaN fails the test too) */
__global__ void __launch_bounds__(256) k_vcol_check(Grid g, const float *__restrict__ met, int *flag) {
  const int i = blockIdx.x * 256 + (int)threadIdx.x;
  if (i >= g.ni) return;
  const long p = (i + 3) + (long)(blockIdx.y + 3) * g.siy + (long)(blockIdx.z + 3) * g.siz;
  const float tiny = 0x1p-32f;
  const float xx = fabsf(met[1 * g.sicmp + p]), ey = fabsf(met[5 * g.sicmp + p]);
  const float a = fabsf(met[2 * g.sicmp + p]), b = fabsf(met[3 * g.sicmp + p]);
  const float c = fabsf(met[4 * g.sicmp + p]), d = fabsf(met[6 * g.sicmp + p]);
  if (!(a <= tiny * xx && b <= tiny * xx && c <= tiny * ey && d <= tiny * ey)) *flag = 1;
}

void launch_vcol_check(const Grid &g, const float *met, int *flag, hipStream_t st) {
  hipLaunchKernelGGL(k_vcol_check, dim3((unsigned)((g.ni + 255) / 256), (unsigned)g.nj, (unsigned)g.nk), dim3(256), 0, st,
                     g, met, flag);
}

void launch_volume(int variant, const VolArgs &a, const BatArgs &b, int idir, int jdir, int kdir, int top,
                   hipStream_t st, hipStream_t st_pml) {
  if (!st_pml) st_pml = st;
  int key = idir * 4 + jdir * 2 + kdir;
  switch (key) {
    case 0: launch_dir<0, 0, 0>(variant, a, b, top, st, st_pml); break;
    case 1: launch_dir<0, 0, 1>(variant, a, b, top, st, st_pml); break;
    case 2: launch_dir<0, 1, 0>(variant, a, b, top, st, st_pml); break;
    case 3: launch_dir<0, 1, 1>(variant, a, b, top, st, st_pml); break;
    case 4: launch_dir<1, 0, 0>(variant, a, b, top, st, st_pml); break;
    case 5: launch_dir<1, 0, 1>(variant, a, b, top, st, st_pml); break;
    case 6: launch_dir<1, 1, 0>(variant, a, b, top, st, st_pml); break;
    default: launch_dir<1, 1, 1>(variant, a, b, top, st, st_pml); break;
  }
}

void launch_dvh2dvz(const Grid &g, const float *met, const float *lam, const float *mu, float *mx, float *my,
                    hipStream_t st) {
  dim3 blk(16, 16, 1);
  dim3 grd((g.ni + 15) / 16, (g.nj + 15) / 16, 1);
  hipLaunchKernelGGL(k_dvh2dvz, grd, blk, 0, st, g, met, lam, mu, mx, my);
}

void launch_pg(const Grid &g, const BatArgs &b, const float *wend, const float *wpre, float *PG, float *Dis, float dt,
               hipStream_t st) {
  dim3 blk(64, 4, 1);
  dim3 grd((g.ni + 63) / 64, (g.nj + 3) / 4, b.nact);
  hipLaunchKernelGGL(k_pg, grd, blk, 0, st, g, b, wend, wpre, PG, Dis, dt);
}

}  // namespace cgfd
