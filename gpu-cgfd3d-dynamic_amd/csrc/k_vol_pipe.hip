/*
 * k_vol_pipe.hip -- software-pipelined, LDS-tiled z-marching volume stencil.
 *
 * Same arithmetic as the other volume kernels (interior rhs of
 * sv_curv_col_el_iso_gpu.cu:231-388, CFS-PML of :824-1290, fused RK update of
 * drv_rk_curv_col.cu:329-568). Per block: a TX x TY column tile marching a
 * chunk of z planes. Per iteration (plane k):
 *   - x/y taps come from an LDS copy of plane k (with its one-sided halo),
 *     written at the end of the previous iteration into the other half of a
 *     double buffer (one barrier per plane);
 *   - z taps come from a five-deep register queue per field;
 *   - metric, media, pre and end values of plane k were loaded one iteration
 *     earlier, so the compute of plane k never waits on HBM;
 *   - every global load issued during iteration k is for plane k+2 (queue
 *     head, halo) or k+1 (pointwise arrays), giving each load a full
 *     iteration of latency cover.
 * The stage kind is a template parameter: stage 0 reads no separate pre (the
 * stage input is pre) and no end, stages 1-2 read pre and end, stage 3 reads
 * end only. Interior tiles run without the PML code.
 */
#include "cgfd_dev.h"
#include "cgfd_kernels.h"
#include "k_vol_common.h"

namespace cgfd {

template <int ID, int JD, int KD, int SK, bool PML>
__global__ void __launch_bounds__(256) k_vol_pipe(VolArgs a, const int *__restrict__ tiles, int ntiles, int ntx, int nty) {
  constexpr int TX = 64, TY = 4, NT = TX * TY;
  constexpr int HX0 = ID ? 1 : 3;
  constexpr int HY0 = JD ? 1 : 3;
  constexpr int LX = TX + 4, LY = TY + 4;
  constexpr int NHALO = LX * LY - TX * TY;
  constexpr int NH = (NHALO + NT - 1) / NT;
  constexpr int Z0 = KD ? -1 : -3;
  constexpr int ZC = -Z0;
  constexpr bool RD_PRE = SK == 1;             /* stages 1, 2 */
  constexpr bool RD_END = SK == 1 || SK == 3;  /* stages 1, 2, 3 */
  __shared__ float lds[2][9][LY][LX];

  const int L = blockIdx.x;
  const int per = (int)(gridDim.x / 8);
  const int slot = (L % 8) * per + L / 8;
  if (slot >= ntiles) return;
  const int tile = tiles[slot];
  const int bx = tile % ntx;
  const int by = (tile / ntx) % nty;
  const int bz = tile / (ntx * nty);

  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const Grid &g = a.g;
  const long sc = g.sicmp, sy = g.siy, sz = g.siz;
  const int gx = bx * TX + tx, gy = by * TY + ty;
  const bool active = gx < g.ni && gy < g.nj;
  const bool loadable = (gx + 3) < g.nx && (gy + 3) < g.ny;
  const long col = (long)(gx + 3) + (long)(gy + 3) * sy;
  const int k_first = a.k_lo + bz * a.zchunk;
  int k_last = k_first + a.zchunk;
  if (k_last > a.k_lo + a.nk_run) k_last = a.k_lo + a.nk_run;

  int hly[NH], hlx[NH];
  long hcol[NH];
  bool hok[NH], hown[NH];
#pragma unroll
  for (int u = 0; u < NH; u++) {
    int h = tid + u * NT;
    hown[u] = h < NHALO;
    hly[u] = 0; hlx[u] = 0; hcol[u] = 0; hok[u] = false;
    if (!hown[u]) continue;
    if (h < HY0 * LX) { hly[u] = h / LX; hlx[u] = h % LX; }
    else if ((h -= HY0 * LX) < (LY - TY - HY0) * LX) { hly[u] = HY0 + TY + h / LX; hlx[u] = h % LX; }
    else {
      h -= (LY - TY - HY0) * LX;
      int side = h / TY, r = h % TY;
      hly[u] = HY0 + r;
      hlx[u] = side < HX0 ? side : TX + side;
    }
    int ex = bx * TX + (hlx[u] - HX0) + 3;
    int ey = by * TY + (hly[u] - HY0) + 3;
    hok[u] = ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny;
    hcol[u] = (long)ex + (long)ey * sy;
  }

  const float *__restrict__ cur = a.cur;
  const float *__restrict__ met = a.met;
  /* pointwise values of plane k (met 9 + media 3 + pre 9 + end 9) */
  float pm[12], pp[9], pe[9];
  float q[9][5], qn[9], hn[NH][9];
  float nm[12], np_[9], ne[9];  /* the same for plane k+1 */

  auto load_point = [&](int k, float *m12, float *pr, float *en) {
    if (!active) return;
    const long p = col + (long)(k + 3) * sz;
#pragma unroll
    for (int e = 0; e < 9; e++) m12[e] = met[(long)(1 + e) * sc + p];
    m12[9] = a.lam[p]; m12[10] = a.mu[p]; m12[11] = a.slw[p];
    if (RD_PRE) {
#pragma unroll
      for (int f = 0; f < 9; f++) pr[f] = a.pre[f * sc + p];
    }
    if (RD_END) {
#pragma unroll
      for (int f = 0; f < 9; f++) en[f] = a.end[f * sc + p];
    }
  };
  auto load_halo = [&](int k, float (*hv)[9]) {
    const long zo = (long)(k + 3) * sz;
#pragma unroll
    for (int u = 0; u < NH; u++)
#pragma unroll
      for (int f = 0; f < 9; f++) hv[u][f] = hok[u] ? cur[f * sc + hcol[u] + zo] : 0.0f;
  };
  auto store_plane = [&](int b, const float (*hv)[9]) {
#pragma unroll
    for (int f = 0; f < 9; f++) lds[b][f][ty + HY0][tx + HX0] = q[f][ZC];
#pragma unroll
    for (int u = 0; u < NH; u++)
      if (hown[u]) {
#pragma unroll
        for (int f = 0; f < 9; f++) lds[b][f][hly[u]][hlx[u]] = hv[u][f];
      }
  };

  /* ---- prologue: queue for k_first, plane k_first in LDS, pointwise of k_first,
     prefetch for k_first+1 ---- */
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const long o = col + (long)(k_first + Z0 + s + 3) * sz;
#pragma unroll
    for (int f = 0; f < 9; f++) q[f][s] = loadable ? cur[f * sc + o] : 0.0f;
  }
  load_halo(k_first, hn);
  load_point(k_first, pm, pp, pe);
  store_plane(0, hn);
  if (k_first + 1 < k_last) {
    const long o = col + (long)(k_first + 1 + Z0 + 4 + 3) * sz;
#pragma unroll
    for (int f = 0; f < 9; f++) qn[f] = loadable ? cur[f * sc + o] : 0.0f;
    load_halo(k_first + 1, hn);
    load_point(k_first + 1, nm, np_, ne);
  }

  int buf = 0;
  for (int k = k_first; k < k_last; k++) {
    __syncthreads();
    if (active) {
      float Dx[9], Dy[9], Dz[9];
#pragma unroll
      for (int f = 0; f < 9; f++) {
        const float *r = &lds[buf][f][ty + HY0][tx + HX0];
        Dx[f] = dfd5<ID>(r[-HX0], r[-HX0 + 1], r[-HX0 + 2], r[-HX0 + 3], r[-HX0 + 4]);
        Dy[f] = dfd5<JD>(r[-HY0 * LX], r[(1 - HY0) * LX], r[(2 - HY0) * LX], r[(3 - HY0) * LX], r[(4 - HY0) * LX]);
        Dz[f] = dfd5<KD>(q[f][0], q[f][1], q[f][2], q[f][3], q[f][4]);
      }
      const long p = col + (long)(k + 3) * sz;
      float h[9];
      rhs_interior(Dx, Dy, Dz, pm[0], pm[1], pm[2], pm[3], pm[4], pm[5], pm[6], pm[7], pm[8], pm[9], pm[10],
                   pm[11], h);
      bool own = true;
      if (PML) {
        pml_point(a, gx + 3, gy + 3, k + 3, p, Dx, Dy, Dz, pm[9], pm[10], pm[11], h);
      }
      if (a.fs.nfault > 0 && fault_column(a.fs, gx + 3, (gy + 3) + (long)(k + 3) * g.ny)) own = false;
      if (own) {
#pragma unroll
        for (int f = 0; f < 9; f++) {
          const long o = f * sc + p;
          if (SK == 0) {
            const float pv = q[f][ZC];
            a.tnew[o] = pv + a.ca * h[f];
            a.end[o] = pv + a.cb * h[f];
          } else if (SK == 1) {
            a.tnew[o] = pp[f] + a.ca * h[f];
            a.end[o] = pe[f] + a.cb * h[f];
          } else {
            a.end[o] = pe[f] + a.cb * h[f];
          }
        }
      }
    }
    /* advance: shift the queue, stage plane k+1 into the other buffer,
       rotate the pointwise registers, issue the prefetch for k+2 */
    if (k + 1 < k_last) {
#pragma unroll
      for (int f = 0; f < 9; f++) {
#pragma unroll
        for (int s = 0; s < 4; s++) q[f][s] = q[f][s + 1];
        q[f][4] = qn[f];
      }
      store_plane(buf ^ 1, hn);
#pragma unroll
      for (int e = 0; e < 12; e++) pm[e] = nm[e];
#pragma unroll
      for (int f = 0; f < 9; f++) { pp[f] = np_[f]; pe[f] = ne[f]; }
      if (k + 2 < k_last) {
        const long o = col + (long)(k + 2 + Z0 + 4 + 3) * sz;
#pragma unroll
        for (int f = 0; f < 9; f++) qn[f] = loadable ? cur[f * sc + o] : 0.0f;
        load_halo(k + 2, hn);
        load_point(k + 2, nm, np_, ne);
      }
      buf ^= 1;
    }
  }
}

template <int ID, int JD, int KD, int SK>
static void launch_sk(const VolArgs &a, const ZqTiles &t, hipStream_t st) {
  if (t.n_int > 0) {
    int nb = ((t.n_int + 7) / 8) * 8;
    hipLaunchKernelGGL((k_vol_pipe<ID, JD, KD, SK, false>), dim3(nb), dim3(64, 4, 1), 0, st, a, t.d_int, t.n_int,
                       t.ntx, t.nty);
  }
  if (t.n_pml > 0) {
    int nb = ((t.n_pml + 7) / 8) * 8;
    hipLaunchKernelGGL((k_vol_pipe<ID, JD, KD, SK, true>), dim3(nb), dim3(64, 4, 1), 0, st, a, t.d_pml, t.n_pml,
                       t.ntx, t.nty);
  }
}

template <int ID, int JD, int KD>
static void launch_d(const VolArgs &a, const ZqTiles &t, hipStream_t st) {
  if (a.istage == 0) launch_sk<ID, JD, KD, 0>(a, t, st);
  else if (a.istage < 3) launch_sk<ID, JD, KD, 1>(a, t, st);
  else launch_sk<ID, JD, KD, 3>(a, t, st);
}

void launch_vol_pipe(const VolArgs &a, const ZqTiles &t, int idir, int jdir, int kdir, hipStream_t st) {
  switch (idir * 4 + jdir * 2 + kdir) {
    case 0: launch_d<0, 0, 0>(a, t, st); break;
    case 1: launch_d<0, 0, 1>(a, t, st); break;
    case 2: launch_d<0, 1, 0>(a, t, st); break;
    case 3: launch_d<0, 1, 1>(a, t, st); break;
    case 4: launch_d<1, 0, 0>(a, t, st); break;
    case 5: launch_d<1, 0, 1>(a, t, st); break;
    case 6: launch_d<1, 1, 0>(a, t, st); break;
    default: launch_d<1, 1, 1>(a, t, st); break;
  }
}

}  // namespace cgfd
