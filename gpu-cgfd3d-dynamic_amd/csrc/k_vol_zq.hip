/*
 * k_vol_zq.hip -- volume stencil with a z register queue and no barriers.
 *
 * Each thread owns one (x, y) column of a TX x TY tile and marches a chunk of
 * z planes, keeping the five z taps of every field (plus one plane of
 * prefetch) in registers, so the plane data are read from HBM once per stage.
 * The x and y taps are plain loads of the neighbours' rows, which the
 * wavefront and its neighbour waves read anyway (L1/L2 hits). Tiles come
 * from a host-built list split into interior tiles (no CFS-PML point) and
 * boundary tiles; interior tiles run an instantiation without the PML code,
 * which keeps its register use low, and accumulate the rhs field by field
 * from the physical gradient of each field (the same sums as the reference's
 * rhs, sv_curv_col_el_iso_gpu.cu:349-386, grouped per field). Block ids map
 * XCD-contiguously onto the tile list.
 */
#include "cgfd_dev.h"
#include "cgfd_kernels.h"
#include "k_vol_common.h"

namespace cgfd {

/* MINW > 0 asks the compiler for at least MINW waves per SIMD (caps VGPRs) */
template <int ID, int JD, int KD, int TX, int TY, bool PML, int MINW>
__global__ void __launch_bounds__(TX *TY, MINW > 0 ? MINW : 1)
    k_vol_zq(VolArgs a, const int *__restrict__ tiles, int ntiles, int ntx, int nty) {
  constexpr int Z0 = KD ? -1 : -3;
  constexpr int ZC = -Z0;
  const int L = blockIdx.x;
  const int per = (int)(gridDim.x / 8);
  const int slot = (L % 8) * per + L / 8;
  if (slot >= ntiles) return;
  const int tile = tiles[slot];
  const int bx = tile % ntx;
  const int by = (tile / ntx) % nty;
  const int bz = tile / (ntx * nty);

  const Grid &g = a.g;
  const long sc = g.sicmp, sy = g.siy, sz = g.siz;
  /* tiles start at ghost-inclusive column 0: whole 128-byte lines per row */
  const int ie = bx * TX + (int)threadIdx.x, gy = by * TY + (int)threadIdx.y;
  if (ie < 3 || ie >= g.ni + 3 || gy >= g.nj) return;  /* no barriers: inactive lanes may leave */
  const long col = (long)ie + (long)(gy + 3) * sy;
  const int k_first = a.k_lo + bz * a.zchunk;
  int k_last = k_first + a.zchunk;
  if (k_last > a.k_lo + a.nk_run) k_last = a.k_lo + a.nk_run;

  const float *__restrict__ cur = a.cur;
  float q[9][6];
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const long o = col + (long)(k_first + Z0 + s + 3) * sz;
#pragma unroll
    for (int f = 0; f < 9; f++) q[f][s] = cur[f * sc + o];
  }
  constexpr int XO = ID ? -1 : -3;  /* first x tap */
  constexpr int YO = JD ? -1 : -3;

  for (int k = k_first; k < k_last; k++) {
    if (k + 1 < k_last) {
      const long o = col + (long)(k + Z0 + 5 + 3) * sz;
#pragma unroll
      for (int f = 0; f < 9; f++) q[f][5] = cur[f * sc + o];
    }
    const long p = col + (long)(k + 3) * sz;
    const float *M = a.met;
    const float xx = M[sc + p], xy = M[2 * sc + p], xz = M[3 * sc + p];
    const float ex = M[4 * sc + p], ey = M[5 * sc + p], ez = M[6 * sc + p];
    const float zx = M[7 * sc + p], zy = M[8 * sc + p], zz = M[9 * sc + p];
    const float l = a.lam[p], m = a.mu[p], s = a.slw[p];
    if (PML) {
      float Dx[9], Dy[9], Dz[9];
#pragma unroll
      for (int f = 0; f < 9; f++) {
        const float *b = cur + f * sc + p;
        float t[5];
#pragma unroll
        for (int u = 0; u < 5; u++) t[u] = (XO + u == 0) ? q[f][ZC] : b[XO + u];
        Dx[f] = dfd5<ID>(t[0], t[1], t[2], t[3], t[4]);
#pragma unroll
        for (int u = 0; u < 5; u++) t[u] = (YO + u == 0) ? q[f][ZC] : b[(long)(YO + u) * sy];
        Dy[f] = dfd5<JD>(t[0], t[1], t[2], t[3], t[4]);
        Dz[f] = dfd5<KD>(q[f][0], q[f][1], q[f][2], q[f][3], q[f][4]);
      }
      float h[9];
      rhs_interior(Dx, Dy, Dz, xx, xy, xz, ex, ey, ez, zx, zy, zz, l, m, s, h);
      pml_point(a, ie, gy + 3, k + 3, p, Dx, Dy, Dz, l, m, s, h);
      if (!(a.fs.nfault > 0 && fault_column(a.fs, ie, (gy + 3) + (long)(k + 3) * g.ny))) rk_point(a, p, h);
    } else {
      /* physical gradient of each field, accumulated into the rhs */
      float gvx[3], gvy[3], gvz[3];
      float hv0 = 0.f, hv1 = 0.f, hv2 = 0.f;
#pragma unroll
      for (int f = 0; f < 9; f++) {
        const float *b = cur + f * sc + p;
        float t[5];
#pragma unroll
        for (int u = 0; u < 5; u++) t[u] = (XO + u == 0) ? q[f][ZC] : b[XO + u];
        const float dx = dfd5<ID>(t[0], t[1], t[2], t[3], t[4]);
#pragma unroll
        for (int u = 0; u < 5; u++) t[u] = (YO + u == 0) ? q[f][ZC] : b[(long)(YO + u) * sy];
        const float dy = dfd5<JD>(t[0], t[1], t[2], t[3], t[4]);
        const float dz = dfd5<KD>(q[f][0], q[f][1], q[f][2], q[f][3], q[f][4]);
        const float g0 = xx * dx + ex * dy + zx * dz;  /* d/dx */
        const float g1 = xy * dx + ey * dy + zy * dz;  /* d/dy */
        const float g2 = xz * dx + ez * dy + zz * dz;  /* d/dz */
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
      if (a.fs.nfault > 0 && fault_column(a.fs, ie, (gy + 3) + (long)(k + 3) * g.ny)) {
        /* owned by the fault kernel */
      } else {
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
        rk_point(a, p, h);
      }
    }
#pragma unroll
    for (int f = 0; f < 9; f++)
#pragma unroll
      for (int u = 0; u < 5; u++) q[f][u] = q[f][u + 1];
  }
}

template <int ID, int JD, int KD, int MINW>
static void launch_t(const VolArgs &a, const ZqTiles &t, hipStream_t st) {
  constexpr int TX = ZqTiles::TX, TY = ZqTiles::TY;
  if (t.n_int > 0) {
    int nb = ((t.n_int + 7) / 8) * 8;
    hipLaunchKernelGGL((k_vol_zq<ID, JD, KD, TX, TY, false, MINW>), dim3(nb), dim3(TX, TY, 1), 0, st, a, t.d_int,
                       t.n_int, t.ntx, t.nty);
  }
  if (t.n_pml > 0) {
    int nb = ((t.n_pml + 7) / 8) * 8;
    /* the PML path spills when capped at 128 VGPRs; allow it three waves */
    constexpr int MW = MINW > 3 ? 3 : MINW;
    hipLaunchKernelGGL((k_vol_zq<ID, JD, KD, TX, TY, true, MW>), dim3(nb), dim3(TX, TY, 1), 0, st, a, t.d_pml,
                       t.n_pml, t.ntx, t.nty);
  }
}

template <int MINW>
static void launch_w(const VolArgs &a, const ZqTiles &t, int idir, int jdir, int kdir, hipStream_t st) {
  switch (idir * 4 + jdir * 2 + kdir) {
    case 0: launch_t<0, 0, 0, MINW>(a, t, st); break;
    case 1: launch_t<0, 0, 1, MINW>(a, t, st); break;
    case 2: launch_t<0, 1, 0, MINW>(a, t, st); break;
    case 3: launch_t<0, 1, 1, MINW>(a, t, st); break;
    case 4: launch_t<1, 0, 0, MINW>(a, t, st); break;
    case 5: launch_t<1, 0, 1, MINW>(a, t, st); break;
    case 6: launch_t<1, 1, 0, MINW>(a, t, st); break;
    default: launch_t<1, 1, 1, MINW>(a, t, st); break;
  }
}

void launch_vol_zq(const VolArgs &a, const ZqTiles &t, int minw, int idir, int jdir, int kdir, hipStream_t st) {
  if (minw >= 4) launch_w<4>(a, t, idir, jdir, kdir, st);
  else if (minw == 3) launch_w<3>(a, t, idir, jdir, kdir, st);
  else launch_w<0>(a, t, idir, jdir, kdir, st);
}

}  // namespace cgfd
