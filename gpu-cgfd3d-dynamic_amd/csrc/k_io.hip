/*
 * k_io.hip -- packing of wavefield slices and snapshots for output.
 *
 * One strided gather serves both: a slice is a box one node thick
 * (io_slice_pack_buff_x/y/z, io_funcs.cu:1653-1689), a snapshot a strided box
 * (io_snap_pack_buff, io_funcs.cu:1692-1718). Strain is formed from the
 * stress and the media at the node with the isotropic compliance of
 * io_snap_stress_to_strain_eliso (io_funcs.cu:1585-1651), which the
 * reference evaluates on the host after the copy; here it is packed on the
 * device so only the requested values cross PCIe.
 */
#include "cgfd_dev.h"
#include "cgfd_kernels.h"

namespace cgfd {

enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TYZ, TXZ, TXY };

/* box = {i1, ni, di, j1, nj, dj, k1, nk, dk} (ghost-inclusive starts);
   output groups in order: velocity (3), stress (6), strain (6), each
   [count_k][count_j][count_i] per field */
__global__ void __launch_bounds__(256) k_pack_box(const float *__restrict__ w, long sicmp, long siy, long siz,
                                                  const float *__restrict__ lam, const float *__restrict__ mu,
                                                  IoBox b, int flags, float *__restrict__ out) {
  const long n1 = (long)b.n[0] * b.n[1], n = n1 * b.n[2];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int ix = (int)(t % b.n[0]), iy = (int)((t / b.n[0]) % b.n[1]), iz = (int)(t / n1);
  const long p = (long)(b.s[0] + ix * b.d[0]) + (long)(b.s[1] + iy * b.d[1]) * siy + (long)(b.s[2] + iz * b.d[2]) * siz;
  long o = t;
  if (flags & 1) {
#pragma unroll
    for (int c = 0; c < 3; c++, o += n) out[o] = w[c * sicmp + p];
  }
  if (flags & 2) {
#pragma unroll
    for (int c = 3; c < 9; c++, o += n) out[o] = w[c * sicmp + p];
  }
  if (flags & 4) {
    const float l = lam[p], m = mu[p];
    const float E1 = (float)((l + m) / (m * (3.0 * l + 2.0 * m)));
    const float E2 = (float)(-l / (2.0 * m * (3.0 * l + 2.0 * m)));
    const float E3 = (float)(1.0 / m);
    const float txx = w[TXX * sicmp + p], tyy = w[TYY * sicmp + p], tzz = w[TZZ * sicmp + p];
    const float E0 = E2 * (txx + tyy + tzz);
    out[o] = E0 - (E2 - E1) * txx; o += n;
    out[o] = E0 - (E2 - E1) * tyy; o += n;
    out[o] = E0 - (E2 - E1) * tzz; o += n;
    out[o] = (float)(0.5 * E3 * w[TYZ * sicmp + p]); o += n;
    out[o] = (float)(0.5 * E3 * w[TXZ * sicmp + p]); o += n;
    out[o] = (float)(0.5 * E3 * w[TXY * sicmp + p]);
  }
}

void launch_pack_box(const float *w, const Grid &g, const float *lam, const float *mu, const IoBox &b, int flags,
                     float *out, hipStream_t st) {
  const long n = (long)b.n[0] * b.n[1] * b.n[2];
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pack_box, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, g.sicmp, g.siy, g.siz, lam, mu,
                     b, flags, out);
}

}  // namespace cgfd
