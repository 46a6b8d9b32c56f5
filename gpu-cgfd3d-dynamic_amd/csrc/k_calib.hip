/*
 * k_calib.hip -- the streaming rate of a volume stage's data movement
 * without its stencil: every point of an ni x nj x nk grid (3 ghosts per
 * side, the core's array pitch) reads nread arrays once and writes nwrite
 * arrays once, with the volume kernels' row pitch (the row length rounded
 * up to 32 floats), 64 x 4 tiles starting at ghost-inclusive column 0 and
 * non-temporal accesses. Each RK stage of the volume update reads and writes
 * a fixed set of arrays per point (stage 1 on a grid of vertical columns: the
 * stage input, five metric terms, three media arrays and the step's start
 * level, 26 reads, and its output level, 9 writes), so these rates, one per
 * stage, are what the volume update's bytes could stream at on this GPU
 * with no neighbour reuse at all: the ceiling its roofline fraction is read
 * against, measured in the same process.
 */
#include <hip/hip_runtime.h>

#include "../../include/cgfd_core.h"
#include "cgfd_dev.h"
#include "cgfd_errors.h"

namespace {

__global__ void __launch_bounds__(256) k_calib_stream(const float *__restrict__ src, float *__restrict__ dst, int nread,
                                                      int nwrite, int ni, int nj, int nx, int ny, long sicmp) {
  const int i = blockIdx.x * 64 + (int)threadIdx.x, j = blockIdx.y * 4 + (int)threadIdx.y;
  if (i < 3 || i >= ni + 3 || j >= nj) return;
  const long o = (long)i + (long)(j + 3) * nx + (long)(blockIdx.z + 3) * nx * ny;
  float acc = 0.0f;
#pragma unroll 8
  for (int r = 0; r < nread; r++) acc += __builtin_nontemporal_load(src + r * sicmp + o);
#pragma unroll 6
  for (int w = 0; w < nwrite; w++) __builtin_nontemporal_store(acc + (float)w, dst + w * sicmp + o);
}

/* the stage-1 point's 39 reads and 18 writes with every access unrolled, so
   all loads of a lane can be in flight at once as in the volume kernels */
template <int NR, int NW>
__global__ void __launch_bounds__(256) k_calib_stream_fixed(const float *__restrict__ src, float *__restrict__ dst,
                                                            int ni, int nj, int nx, int ny, long sicmp) {
  const int i = blockIdx.x * 64 + (int)threadIdx.x, j = blockIdx.y * 4 + (int)threadIdx.y;
  if (i < 3 || i >= ni + 3 || j >= nj) return;
  const long o = (long)i + (long)(j + 3) * nx + (long)(blockIdx.z + 3) * nx * ny;
  float v[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) v[r] = __builtin_nontemporal_load(src + r * sicmp + o);
  float acc = 0.0f;
#pragma unroll
  for (int r = 0; r < NR; r++) acc += v[r];
#pragma unroll
  for (int w = 0; w < NW; w++) __builtin_nontemporal_store(acc + (float)w, dst + w * sicmp + o);
}

}  // namespace

extern "C" int cgfd_calib_stream(int device_id, int ni, int nj, int nk, int nread, int nwrite, int reps,
                                 float *gbs) {
  if (!gbs || ni < 1 || nj < 1 || nk < 1 || nread < 1 || nwrite < 0 || reps < 1 || nread > 64 || nwrite > 64)
    return cgfd_set_error(CGFD_ERR_ARG, "cgfd_calib_stream: bad arguments");
  *gbs = 0.0f;
  int old = -1;
  if (hipGetDevice(&old) != hipSuccess) old = -1;
  if (hipSetDevice(device_id) != hipSuccess)
    return cgfd_set_error(CGFD_ERR_HIP, "cgfd_calib_stream: hipSetDevice(%d) failed", device_id);
  const int nx = (ni + 6 + 31) & ~31, ny = nj + 6, nz = nk + 6;
  const long sicmp = (long)nx * ny * nz;
  float *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CGFD_OK;
  hipError_t e = hipMalloc((void **)&src, (size_t)nread * sicmp * 4);
  if (e == hipSuccess && nwrite > 0) e = hipMalloc((void **)&dst, (size_t)nwrite * sicmp * 4);
  if (e != hipSuccess) rc = cgfd_set_error(CGFD_ERR_NOMEM, "cgfd_calib_stream: %s", hipGetErrorString(e));
  if (!rc) {
    e = hipMemset(src, 0, (size_t)nread * sicmp * 4);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
      const dim3 grid((unsigned)((ni + 3 + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)nk), blk(64, 4, 1);
      auto launch = [&]() {
#define CGFD_CALIB_CASE(R, W)                                                                                 \
  if (nread == R && nwrite == W) {                                                                            \
    hipLaunchKernelGGL((k_calib_stream_fixed<R, W>), grid, blk, 0, 0, src, dst, ni, nj, nx, ny, sicmp);      \
    return;                                                                                                   \
  }
        /* the RK stages' array counts (reads of the stage input, metric
           with five or nine terms, media, the levels the stage combines;
           writes of its output level) and the earlier scheme's 39 + 18 */
        CGFD_CALIB_CASE(17, 9) CGFD_CALIB_CASE(26, 9) CGFD_CALIB_CASE(44, 9)
        CGFD_CALIB_CASE(21, 9) CGFD_CALIB_CASE(30, 9) CGFD_CALIB_CASE(48, 9)
        CGFD_CALIB_CASE(39, 18)
#undef CGFD_CALIB_CASE
        hipLaunchKernelGGL(k_calib_stream, grid, blk, 0, 0, src, dst, nread, nwrite, ni, nj, nx, ny, sicmp);
      };
      launch();
      e = hipEventRecord(e0, 0);
      for (int r = 0; r < reps && e == hipSuccess; r++) {
        launch();
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipEventRecord(e1, 0);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e == hipSuccess && ms > 0.0f)
        *gbs = (float)((double)(nread + nwrite) * 4.0 * ni * nj * nk * reps / (ms * 1e-3) / 1e9);
    }
    if (e != hipSuccess) rc = cgfd_set_error(CGFD_ERR_HIP, "cgfd_calib_stream: %s", hipGetErrorString(e));
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (old >= 0) (void)hipSetDevice(old);
  return rc;
}
