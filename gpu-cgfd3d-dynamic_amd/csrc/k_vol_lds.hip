/*
 * k_vol_lds.hip -- the volume stencil as a z-marching, LDS-tiled kernel.
 *
 * Same arithmetic as k_vol_direct (interior rhs of sv_curv_col_el_iso_gpu.cu:
 * 231-388, CFS-PML of :824-1290 and the fused RK update), organised so every
 * input byte of a stage is read from HBM once:
 *   - a block owns a TX x TY column tile and marches up a chunk of z planes;
 *   - each thread keeps a 6-deep register queue per field along z (the five
 *     taps of the one-sided z operator plus one plane of prefetch);
 *   - the plane being updated is staged in LDS with its one-sided x/y halo
 *     (1 point on one side, 3 on the other), from which the x and y
 *     derivatives are taken; halo values for the next plane are loaded while
 *     the current plane is computed;
 *   - tiles are assigned so that the blocks one XCD runs are spatial
 *     neighbours (workgroups are dispatched round-robin over the 8 XCDs, each
 *     with its own L2), which keeps halo rows in that XCD's L2.
 */
#include "cgfd_dev.h"
#include "cgfd_kernels.h"
#include "k_vol_common.h"

namespace cgfd {

template <int ID, int JD, int KD, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) k_vol_lds(VolArgs a, int ntx, int nty, int ntz, int ntiles) {
  constexpr int HX0 = ID ? 1 : 3;  /* forward taps -1..3, backward -3..1 */
  constexpr int HY0 = JD ? 1 : 3;
  constexpr int LX = TX + 4, LY = TY + 4;
  constexpr int NHALO = LX * LY - TX * TY;
  constexpr int Z0 = KD ? -1 : -3;  /* first z tap */
  constexpr int ZC = -Z0;           /* queue slot of the plane being updated */
  constexpr int NT = TX * TY;
  __shared__ float lds[9][LY][LX];

  /* XCD-aware tile id: block L runs on XCD L % 8; give each XCD a contiguous
     range of tiles (x fastest, then y, then z chunks) */
  const int L = blockIdx.x;
  const int per = (int)((gridDim.x + 7) / 8);
  const int tile = (L % 8) * per + L / 8;
  if (tile >= ntiles) return;
  const int bx = tile % ntx;
  const int by = (tile / ntx) % nty;
  const int bz = tile / (ntx * nty);
  (void)ntz;

  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const Grid &g = a.g;
  const long sc = g.sicmp, sy = g.siy, sz = g.siz;
  const int gx = bx * TX + tx, gy = by * TY + ty;  /* physical x, y */
  const bool active = gx < g.ni && gy < g.nj;
  const bool loadable = (gx + 3) < g.nx && (gy + 3) < g.ny;
  const long col = (long)(gx + 3) + (long)(gy + 3) * sy;
  const int k_first = a.k_lo + bz * a.zchunk;             /* physical */
  int k_last = k_first + a.zchunk;
  if (k_last > a.k_lo + a.nk_run) k_last = a.k_lo + a.nk_run;

  /* halo points owned by this thread (indices tid and tid + NT): LDS
     coordinates and extended-grid column */
  constexpr int NH = (NHALO + NT - 1) / NT;
  int hly[NH], hlx[NH];
  long hcol[NH];
  bool hvalid[NH], hown[NH];
#pragma unroll
  for (int u = 0; u < NH; u++) {
    int h = tid + u * NT;
    hown[u] = h < NHALO;
    hly[u] = 0; hlx[u] = 0; hcol[u] = 0; hvalid[u] = false;
    if (!hown[u]) continue;
    if (h < HY0 * LX) { hly[u] = h / LX; hlx[u] = h % LX; }
    else if ((h -= HY0 * LX) < (LY - TY - HY0) * LX) { hly[u] = HY0 + TY + h / LX; hlx[u] = h % LX; }
    else {
      h -= (LY - TY - HY0) * LX;
      int side = h / TY;            /* 0 .. 3 : left columns then right columns */
      int r = h % TY;
      hly[u] = HY0 + r;
      hlx[u] = side < HX0 ? side : TX + side;
    }
    int ex = bx * TX + (hlx[u] - HX0) + 3;  /* extended coordinates */
    int ey = by * TY + (hly[u] - HY0) + 3;
    hvalid[u] = ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny;
    hcol[u] = (long)ex + (long)ey * sy;
  }

  const float *__restrict__ cur = a.cur;
  float q[9][6];
  float hv[NH][9];
  /* prologue: planes k_first+Z0 .. k_first+Z0+4 into the queue */
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const long zo = (long)(k_first + Z0 + s + 3) * sz;
#pragma unroll
    for (int f = 0; f < 9; f++) q[f][s] = loadable ? cur[f * sc + col + zo] : 0.0f;
  }
  {
    const long zo = (long)(k_first + 3) * sz;
#pragma unroll
    for (int u = 0; u < NH; u++)
#pragma unroll
      for (int f = 0; f < 9; f++) hv[u][f] = hvalid[u] ? cur[f * sc + hcol[u] + zo] : 0.0f;
  }

  for (int k = k_first; k < k_last; k++) {
    /* prefetch the queue plane needed by the next iteration */
    const bool more = (k + 1) < k_last;
    if (more) {
      const long zo = (long)(k + Z0 + 5 + 3) * sz;
#pragma unroll
      for (int f = 0; f < 9; f++) q[f][5] = loadable ? cur[f * sc + col + zo] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 9; f++) lds[f][ty + HY0][tx + HX0] = q[f][ZC];
#pragma unroll
    for (int u = 0; u < NH; u++)
      if (hown[u]) {
#pragma unroll
        for (int f = 0; f < 9; f++) lds[f][hly[u]][hlx[u]] = hv[u][f];
      }
    __syncthreads();
    if (more) {
      const long zo = (long)(k + 1 + 3) * sz;
#pragma unroll
      for (int u = 0; u < NH; u++)
#pragma unroll
        for (int f = 0; f < 9; f++) hv[u][f] = hvalid[u] ? cur[f * sc + hcol[u] + zo] : 0.0f;
    }
    if (active) {
      float Dx[9], Dy[9], Dz[9];
#pragma unroll
      for (int f = 0; f < 9; f++) {
        const float *r = &lds[f][ty + HY0][tx + HX0];
        Dx[f] = dfd5<ID>(r[-HX0], r[-HX0 + 1], r[-HX0 + 2], r[-HX0 + 3], r[-HX0 + 4]);
        Dy[f] = dfd5<JD>(r[-HY0 * LX], r[(1 - HY0) * LX], r[(2 - HY0) * LX], r[(3 - HY0) * LX], r[(4 - HY0) * LX]);
        Dz[f] = dfd5<KD>(q[f][0], q[f][1], q[f][2], q[f][3], q[f][4]);
      }
      const long p = col + (long)(k + 3) * sz;
      vol_point_update(a, gx + 3, gy + 3, k + 3, p, Dx, Dy, Dz);
    }
#pragma unroll
    for (int f = 0; f < 9; f++) {
#pragma unroll
      for (int s = 0; s < 5; s++) q[f][s] = q[f][s + 1];
    }
  }
}

template <int ID, int JD, int KD>
static void launch_t(const VolArgs &a0, hipStream_t st) {
  constexpr int TX = 64, TY = 4;
  VolArgs a = a0;
  if (a.zchunk <= 0) a.zchunk = 32;
  const int ntx = (a.g.ni + TX - 1) / TX;
  const int nty = (a.g.nj + TY - 1) / TY;
  const int ntz = (a.nk_run + a.zchunk - 1) / a.zchunk;
  const int ntiles = ntx * nty * ntz;
  const int nblocks = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k_vol_lds<ID, JD, KD, TX, TY>), dim3(nblocks), dim3(TX, TY, 1), 0, st, a, ntx, nty, ntz, ntiles);
}

void launch_vol_lds(const VolArgs &a, int idir, int jdir, int kdir, hipStream_t st) {
  switch (idir * 4 + jdir * 2 + kdir) {
    case 0: launch_t<0, 0, 0>(a, st); break;
    case 1: launch_t<0, 0, 1>(a, st); break;
    case 2: launch_t<0, 1, 0>(a, st); break;
    case 3: launch_t<0, 1, 1>(a, st); break;
    case 4: launch_t<1, 0, 0>(a, st); break;
    case 5: launch_t<1, 0, 1>(a, st); break;
    case 6: launch_t<1, 1, 0>(a, st); break;
    default: launch_t<1, 1, 1>(a, st); break;
  }
}

}  // namespace cgfd
