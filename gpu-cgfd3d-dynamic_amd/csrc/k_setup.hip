/*
 * k_setup.hip -- the setup's metric and fault-coefficient producers on the
 * device. Each thread runs the same per-point / per-node arithmetic as the
 * host setup (cgfd_setup_node.h), with FMA contraction off so that the
 * device results follow the host's rounding step for step.
 *
 *   k_setup_metric      gd_curv_metric_cal, gd_t.cu:209-304 (6th-order
 *                       centred differences of the coordinates, HBM-bound:
 *                       reads 3 coordinate fields with 6-point halos, writes
 *                       10 metric fields per point)
 *   k_setup_fault_coef  fault_coef_cal, fault_info.cu:155-845 (one node per
 *                       thread: ~30 3x3 products and inverses)
 */
#pragma STDC FP_CONTRACT OFF

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/cgfd_core.h"
#include "cgfd_errors.h"
#include "cgfd_setup_dev.h"

namespace {

using cgfd_setup_math::FaultCoefOut;

__global__ void __launch_bounds__(256) k_setup_metric(const float *__restrict__ x3d, const float *__restrict__ y3d,
                                                      const float *__restrict__ z3d, float *__restrict__ met, int ni,
                                                      int nx, int ny, size_t sicmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  const size_t siy = (size_t)nx, siz = (size_t)nx * ny;
  const size_t iptr = (size_t)(i + 3) + (size_t)(blockIdx.y + 3) * siy + (size_t)(blockIdx.z + 3) * siz;
  float o[10];
  cgfd_setup_math::metric_point(x3d, y3d, z3d, iptr, (long)siy, (long)siz, o);
#pragma unroll
  for (int c = 0; c < 10; c++) met[c * sicmp + iptr] = o[c];
}

__global__ void __launch_bounds__(128) k_setup_fault_coef(const float *__restrict__ node_in, FaultCoefOut out, int ny,
                                                          int nz, int k_top) {
  const size_t pf = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pf >= (size_t)ny * nz) return;
  const int j = (int)(pf % ny), k = (int)(pf / ny);
  const float *q = node_in + pf * kFaultNodeIn;
  const float e1[3] = {q[0], q[1], q[2]}, e2[3] = {q[3], q[4], q[5]}, e3[3] = {q[6], q[7], q[8]};
  cgfd_setup_math::fault_coef_node(out, pf, (size_t)j, e1, e2, e3, q[9], q[10], q[11], k == k_top);
}

/* device buffers freed on every exit path; the caller's current device is
   restored */
struct DevScope {
  int old = -1;
  std::vector<void *> bufs;
  ~DevScope() {
    for (void *b : bufs) (void)hipFree(b);
    if (old >= 0) (void)hipSetDevice(old);
  }
  int enter(int dev, const char *who) {
    if (hipGetDevice(&old) != hipSuccess) old = -1;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return cgfd_set_error(CGFD_ERR_HIP, "%s: hipSetDevice(%d): %s", who, dev, hipGetErrorString(e));
    return CGFD_OK;
  }
  int alloc(float **p, size_t n, const char *who) {
    hipError_t e = hipMalloc((void **)p, n * sizeof(float));
    if (e != hipSuccess)
      return cgfd_set_error(CGFD_ERR_NOMEM, "%s: device allocation of %zu floats: %s", who, n, hipGetErrorString(e));
    bufs.push_back(*p);
    return CGFD_OK;
  }
};

int hip_check(hipError_t e, const char *who, const char *what) {
  if (e == hipSuccess) return CGFD_OK;
  return cgfd_set_error(CGFD_ERR_HIP, "%s: %s: %s", who, what, hipGetErrorString(e));
}

}  // namespace

int cgfd_dev_metric(int device_id, int ni, int nj, int nk, const float *coords, float *metric) {
  static const char *who = "device metric";
  const int nx = ni + 6, ny = nj + 6, nz = nk + 6;
  const size_t sicmp = (size_t)nx * ny * nz;
  DevScope sc;
  int rc = sc.enter(device_id, who);
  if (rc) return rc;
  float *dc = nullptr, *dm = nullptr;
  if ((rc = sc.alloc(&dc, 3 * sicmp, who)) || (rc = sc.alloc(&dm, 10 * sicmp, who))) return rc;
  if ((rc = hip_check(hipMemcpy(dc, coords, 3 * sicmp * sizeof(float), hipMemcpyHostToDevice), who, "upload")))
    return rc;
  if ((rc = hip_check(hipMemset(dm, 0, 10 * sicmp * sizeof(float)), who, "clear"))) return rc;
  dim3 grid((unsigned)((ni + 255) / 256), (unsigned)nj, (unsigned)nk);
  hipLaunchKernelGGL(k_setup_metric, grid, dim3(256), 0, 0, dc, dc + sicmp, dc + 2 * sicmp, dm, ni, nx, ny, sicmp);
  if ((rc = hip_check(hipGetLastError(), who, "launch"))) return rc;
  return hip_check(hipMemcpy(metric, dm, 10 * sicmp * sizeof(float), hipMemcpyDeviceToHost), who, "download");
}

int cgfd_dev_fault_coef(int device_id, int ny, int nz, int k_top, const float *node_in, const FaultCoefOut &out) {
  static const char *who = "device fault coefficients";
  const size_t ns = (size_t)ny * nz;
  DevScope sc;
  int rc = sc.enter(device_id, who);
  if (rc) return rc;
  /* one device array per output, in the order of FaultCoefOut */
  struct Seg {
    float *host;
    size_t n;
    float *dev;
  };
  std::vector<Seg> segs;
  for (int q = 0; q < 12; q++) segs.push_back({out.D[q], 9 * ns, nullptr});
  for (int q = 0; q < 5; q++) segs.push_back({out.P2M[q], 9 * ns, nullptr});
  for (int q = 0; q < 5; q++) segs.push_back({out.M2P[q], 9 * ns, nullptr});
  for (int q = 0; q < 6; q++) segs.push_back({out.M2[q], 9 * ns, nullptr});
  for (int q = 0; q < 14; q++) segs.push_back({out.Fm[q], 9 * (size_t)ny, nullptr});
  for (float *h : {out.vec_n, out.vec_s1, out.vec_s2}) segs.push_back({h, 3 * ns, nullptr});
  for (float *h : {out.x_et, out.y_et, out.z_et}) segs.push_back({h, ns, nullptr});
  size_t total = 0;
  for (const Seg &s : segs) total += s.n;
  float *din = nullptr, *dout = nullptr;
  if ((rc = sc.alloc(&din, kFaultNodeIn * ns, who)) || (rc = sc.alloc(&dout, total, who))) return rc;
  size_t off = 0;
  for (Seg &s : segs) { s.dev = dout + off; off += s.n; }
  FaultCoefOut d;
  int t = 0;
  for (int q = 0; q < 12; q++) d.D[q] = segs[t++].dev;
  for (int q = 0; q < 5; q++) d.P2M[q] = segs[t++].dev;
  for (int q = 0; q < 5; q++) d.M2P[q] = segs[t++].dev;
  for (int q = 0; q < 6; q++) d.M2[q] = segs[t++].dev;
  for (int q = 0; q < 14; q++) d.Fm[q] = segs[t++].dev;
  d.vec_n = segs[t++].dev; d.vec_s1 = segs[t++].dev; d.vec_s2 = segs[t++].dev;
  d.x_et = segs[t++].dev; d.y_et = segs[t++].dev; d.z_et = segs[t++].dev;
  if ((rc = hip_check(hipMemcpy(din, node_in, kFaultNodeIn * ns * sizeof(float), hipMemcpyHostToDevice), who,
                      "upload")))
    return rc;
  /* the free-surface rows are written for k == k_top only */
  if ((rc = hip_check(hipMemset(dout, 0, total * sizeof(float)), who, "clear"))) return rc;
  hipLaunchKernelGGL(k_setup_fault_coef, dim3((unsigned)((ns + 127) / 128)), dim3(128), 0, 0, din, d, ny, nz, k_top);
  if ((rc = hip_check(hipGetLastError(), who, "launch"))) return rc;
  if ((rc = hip_check(hipDeviceSynchronize(), who, "kernel"))) return rc;
  for (const Seg &s : segs)
    if ((rc = hip_check(hipMemcpy(s.host, s.dev, s.n * sizeof(float), hipMemcpyDeviceToHost), who, "download")))
      return rc;
  return CGFD_OK;
}
