/*
 * core.hip -- the C ABI of include/cgfd_core.h: device state, uploads, and
 * the time loop (drv_rk_curv_col_allstep, src/forward/drv_rk_curv_col.cu:28-667)
 * on one GPU without the MPI halo exchange.
 *
 * Device layout: the reference's SoA layout (nx = ni+6 fastest, 3 ghosts per
 * side). The wavefield has four levels used as pre / stage-input ping-pong A /
 * ping-pong B / end: stage 0 reads pre and writes A, stage 1 reads A and
 * writes B, stage 2 reads B and writes A, stage 3 reads A and writes end only.
 * pre and end swap after every step, as the reference's pointers do
 * (drv_rk_curv_col.cu:622-635). The same scheme holds for the split-node
 * fault state; the PML auxiliaries are point-local and keep the reference's
 * pre / tmp / end levels.
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/cgfd_core.h"
#include "../host/cgfd_errors.h"
#include "../host/cgfd_nc.h"
#include "../host/cgfd_sac.h"
#include "../../include/cgfd_io.h"
#include <string>
#include "cgfd_kernels.h"

using namespace cgfd;

#define HIPCHK(call)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return cgfd_set_error(CGFD_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                                  \
  } while (0)

namespace {

/* RK4 (fd_t.cu:38-45) and operator pairs (fd_t.cu:149-159) */
const float kRkA[3] = {0.5f, 0.5f, 1.0f};
const float kRkB[4] = {(float)(1.0 / 6.0), (float)(1.0 / 3.0), (float)(1.0 / 3.0), (float)(1.0 / 6.0)};
const int kPairs[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};

struct DevBuf {
  std::vector<void *> ptrs;
  ~DevBuf() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

}  // namespace

struct cgfd_core {
  int dev = 0;
  hipStream_t st = nullptr;
  Grid g{};
  int ni = 0, nj = 0, nk = 0;
  float dt = 0, t0 = 0;
  int nt_total = 0, isfree = 0, imethod = 0, nfault = 0, fix_q1 = 0, fix_sta = 0, nan_every = 0;
  int vcol_grid = 0;  /* xi_y, xi_z, eta_x, eta_z below 2^-32 of xi_x, eta_y everywhere */
  int vcol_on = 1;    /* use the vertical-column form where the grid has it */
  int variant = 33;
  int zchunk = 4;
  int nvol = 0;
  DevBuf mem;

  float *wave = nullptr;   /* 4 levels x 9 x sicmp */
  /* wavefield levels: the step's start (pre) and the outputs of stages 0,
     1, 2 (A, B, C); stage 3 writes t_{n+1} over A, and pre and A swap */
  int lv_pre = 0, lv_a = 1, lv_b = 2, lv_c = 3;
  float *met = nullptr, *lam = nullptr, *mu = nullptr, *slw = nullptr;
  float *matVx2Vz = nullptr, *matVy2Vz = nullptr;

  PmlFace pml[3][2]{};
  int pml_on = 0;

  float *fwav = nullptr;   /* 4 levels x nfault x 18 x sslice */
  int fl_pre = 0, fl_a = 1, fl_b = 2, fl_end = 3;
  FaultDev flt[5]{};
  FaultDev *d_ftab = nullptr;       /* flt[] in device memory */

  int *it_dev = nullptr;
  int *nan_flag = nullptr;
  int it_host = 0;

  /* scenarios (a plain run is a batch of one) */
  int nscen = 1;
  int scen_nt[kMaxScen]{};           /* steps each scenario runs (batches only) */
  BatArgs bat{};                     /* strides + the active list of the next launch */
  bool active(int s) const { return nscen == 1 || it_host < scen_nt[s]; }
  int refresh_active() {             /* returns the number of active scenarios */
    int n = 0;
    for (int s = 0; s < nscen; s++)
      if (active(s)) bat.act[n++] = s;
    if (n != bat.nact) drop_graph();
    bat.nact = n;
    return n;
  }
  float *PG = nullptr, *Dis = nullptr;
  float *stage = nullptr;  /* nx*ny*nz floats: host-layout staging for volume copies */

  /* stations: base node per station (st_idx), and per station st_nterm
     terms of the recorded sum (node index, factors fx fy fz) */
  long *st_idx = nullptr, *st_term = nullptr;
  float *st_buf = nullptr, *st_fac = nullptr;
  size_t st_idx_cap = 0, st_buf_cap = 0, st_term_cap = 0, st_fac_cap = 0;
  int st_n = 0, st_nt = 0, st_it0 = 0, st_nterm = 1;
  /* fault receivers: fr_nterm (fault id, slice index) pairs and factors
     (fy, fz) per receiver, [n][nt][9] records */
  int *fr_idx = nullptr;
  float *fr_buf = nullptr, *fr_fac = nullptr;
  size_t fr_idx_cap = 0, fr_buf_cap = 0, fr_fac_cap = 0;
  int fr_n = 0, fr_nt = 0, fr_it0 = 0, fr_nterm = 1;

  /* eight steps (one cycle of the operator pairs; an even number of level
     swaps) captured once and replayed */
  static constexpr int kGraphSteps = 8;
  int use_graph = 1;

  /* stage fork: after the trial kernel, the interior volume blocks run on
     st, the CFS-PML blocks and the top layers on side[0], the fault columns
     and split nodes on side[1]; all of them read the stage input level and
     write disjoint points of the output levels, so they only join at the
     end of the stage. All streams have the default priority. Off by
     default: on the 1200 x 600 x 400 rupture the volume blocks fill the GPU
     and the step takes as long either way. */
  int overlap = 0;
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  int fork() {
    if (!overlap) return CGFD_OK;
    HIPCHK(hipEventRecord(ev_fork, st));
    for (int q = 0; q < 2; q++) HIPCHK(hipStreamWaitEvent(side[q], ev_fork, 0));
    return CGFD_OK;
  }
  int join() {
    if (!overlap) return CGFD_OK;
    for (int q = 0; q < 2; q++) {
      HIPCHK(hipEventRecord(ev_join[q], side[q]));
      HIPCHK(hipStreamWaitEvent(st, ev_join[q], 0));
    }
    return CGFD_OK;
  }
  hipStream_t st_pml() const { return overlap ? side[0] : st; }
  hipStream_t st_fault() const { return overlap ? side[1] : st; }
  hipGraphExec_t gexec = nullptr;
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
  }

  /* discrete adjoint: 4 levels of 9 fields (E, A, B, P) and the 12 fields
     of the adjoint gradient terms; allocated by the first backward call */
  float *adj = nullptr, *adj_sg = nullptr;
  float *adj_pml = nullptr;  /* per CFS-PML face: 4 adjoint levels [9][n] + extra S, G [12][n] */
  /* fault part of the adjoint: the friction tape written by the forward
     kernels over the recorded window, and per fault the adjoint blocks of
     kAdjFaultFloats per node (fault state levels E, A, B, P; T1; hT1; mT1;
     Slip; parameter gradients) */
  float *tape = nullptr;
  size_t tape_cap = 0;
  int tape_it0 = 0, tape_nt = 0;
  float *adjf = nullptr;
  FaultDev *d_ftab_adj = nullptr;  /* the fault table as one scenario of a batch sees it */

  float *wlev(int l) { return wave + (long)l * 9 * g.sicmp; }
  float *flev(int l) { return fwav + (long)l * nfault * 18 * g.sslice; }

  template <class T>
  int alloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) return CGFD_OK;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess)
      return cgfd_set_error(CGFD_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    mem.ptrs.push_back(q);
    *p = (T *)q;
    return CGFD_OK;
  }
  /* a buffer that set-up calls may size again: kept while it is large
     enough, else replaced (the old one freed at once, so repeated calls in a
     caller's loop do not accumulate device memory) */
  template <class T>
  int grow(T **p, size_t *cap, size_t n) {
    if (*p && n <= *cap) return CGFD_OK;
    if (*p) {
      for (void *&q : mem.ptrs)
        if (q == (void *)*p) { (void)hipFree(q); q = nullptr; }
      *p = nullptr;
      *cap = 0;
    }
    if (int rc = alloc(p, n)) return rc;
    *cap = n;
    return CGFD_OK;
  }
};

static int check_desc(const cgfd_desc *d) {
  if (!d) return cgfd_set_error(CGFD_ERR_ARG, "null descriptor");
  if (d->ni < 8 || d->nj < 8 || d->nk < 8)
    return cgfd_set_error(CGFD_ERR_ARG, "grid too small: ni=%d nj=%d nk=%d (each must be >= 8)", d->ni, d->nj, d->nk);
  /* the volume kernels address one field array with 32-bit byte offsets */
  if ((double)(((d->ni + 6 + 31) / 32) * 32) * (d->nj + 6) * (d->nk + 6) * 4.0 >= 4294967296.0)
    return cgfd_set_error(CGFD_ERR_ARG, "one field array of %dx%dx%d points exceeds 4 GiB; split the domain",
                          d->ni + 6, d->nj + 6, d->nk + 6);
  if (!d->metric || !d->lambda || !d->mu || !d->rho) return cgfd_set_error(CGFD_ERR_ARG, "metric/media pointers missing");
  if (!(d->dt > 0.0f)) return cgfd_set_error(CGFD_ERR_ARG, "dt must be > 0 (got %g)", (double)d->dt);
  if (d->number_fault < 0 || d->number_fault > CGFD_MAX_FAULTS)
    return cgfd_set_error(CGFD_ERR_ARG, "number_fault=%d out of range", d->number_fault);
  if (d->number_fault > 0 && d->imethod != 1 && d->imethod != 2)
    return cgfd_set_error(CGFD_ERR_ARG, "dynamic_method must be 1 or 2 (got %d)", d->imethod);
  for (int id = 0; id < d->number_fault; id++) {
    const cgfd_fault_desc &f = d->fault[id];
    if (f.fault_x_index < 4 || f.fault_x_index > d->ni - 5)
      return cgfd_set_error(CGFD_ERR_ARG, "fault %d: x index %d outside [4, ni-5]", id, f.fault_x_index);
    if (!f.T0x || !f.T0y || !f.T0z || !f.mu_s || !f.mu_d || !f.Dc || !f.C0 || !f.united || !f.faultgrid ||
        !f.rup_index_y || !f.rup_index_z || !f.rho_f || !f.vec_s1 || !f.vec_s2)
      return cgfd_set_error(CGFD_ERR_ARG, "fault %d: required array missing", id);
    for (int q = 0; q < 12; q++)
      if (!f.D[q]) return cgfd_set_error(CGFD_ERR_ARG, "fault %d: D matrix %d missing", id, q);
    for (int q = 0; q < 5; q++)
      if ((d->imethod == 1) && (!f.matPlus2Min[q] || !f.matMin2Plus[q]))
        return cgfd_set_error(CGFD_ERR_ARG, "fault %d: method-1 matrix %d missing", id, q);
    for (int q = 0; q < 6; q++)
      if (d->imethod == 2 && !f.matM2[q]) return cgfd_set_error(CGFD_ERR_ARG, "fault %d: method-2 matrix %d missing", id, q);
    if (d->free_top)
      for (int q = 0; q < 14; q++)
        if (!f.matF[q]) return cgfd_set_error(CGFD_ERR_ARG, "fault %d: free-surface matrix %d missing", id, q);
  }
  for (int dim = 0; dim < 3; dim++)
    for (int side = 0; side < 2; side++) {
      if (!d->pml_is_sides[dim][side]) continue;
      int lo[3] = {d->pml_ni1[dim][side], d->pml_nj1[dim][side], d->pml_nk1[dim][side]};
      int hi[3] = {d->pml_ni2[dim][side], d->pml_nj2[dim][side], d->pml_nk2[dim][side]};
      int n[3] = {d->ni, d->nj, d->nk};
      for (int a = 0; a < 3; a++)
        if (lo[a] < 3 || hi[a] > n[a] + 2 || hi[a] < lo[a])
          return cgfd_set_error(CGFD_ERR_ARG, "PML face (%d,%d): index range outside the physical grid", dim, side);
      if (!d->pml_A[dim][side] || !d->pml_B[dim][side] || !d->pml_D[dim][side])
        return cgfd_set_error(CGFD_ERR_ARG, "PML face (%d,%d): coefficient arrays missing", dim, side);
      if (d->free_top && dim == 2 && side == 1)
        return cgfd_set_error(CGFD_ERR_ARG, "PML and free surface both requested on the top face");
    }
  return CGFD_OK;
}

/* Volume arrays: host rows of nx floats (the reference's layout), device
   rows of siy floats. Each component goes through a contiguous device
   staging buffer (one full-speed copy) and a re-pitch kernel. */
static int host_to_dev(cgfd_core *c, float *dev, const float *host, int ncmp) {
  const Grid &g = c->g;
  const size_t hc = (size_t)g.nx * g.ny * g.nz;
  for (int q = 0; q < ncmp; q++) {
    HIPCHK(hipMemcpyAsync(c->stage, host + q * hc, hc * 4, hipMemcpyHostToDevice, c->st));
    launch_repitch(dev + q * g.sicmp, g.siy, c->stage, g.nx, g.nx, (long)g.ny * g.nz, c->st);
  }
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipGetLastError());
  return CGFD_OK;
}
/* rows of nx floats from a device array with pitch siy to the host */
static int dev_to_host(cgfd_core *c, float *host, const float *dev, long rows) {
  const Grid &g = c->g;
  const long per = (long)g.ny * g.nz;  /* rows per staging pass */
  for (long r0 = 0; r0 < rows; r0 += per) {
    const long nr = std::min(per, rows - r0);
    launch_repitch(c->stage, g.nx, dev + r0 * g.siy, g.siy, g.nx, nr, c->st);
    HIPCHK(hipMemcpyAsync(host + r0 * g.nx, c->stage, (size_t)nr * g.nx * 4, hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipGetLastError());
  return CGFD_OK;
}

/* host [node][9] -> device SoA [9][node] for a list of matrices */
static void soa9(std::vector<float> &dst, size_t off, const float *src, size_t nodes) {
  for (size_t p = 0; p < nodes; p++)
    for (int e = 0; e < 9; e++) dst[off + e * nodes + p] = src[p * 9 + e];
}


/* fault arrays: the per-scenario block (T0 3, fric 4, out 11, T1 21, hT1 3,
   mT1 3 floats and flag_rup, init_t0_flag ints per node) for every scenario,
   and the shared masks and coefficients from the first descriptor */
static int upload_faults(cgfd_core *c, const cgfd_desc *ds) {
  const size_t ss = c->g.sslice, ny = c->g.ny;
  const cgfd_desc *d = &ds[0];
  const int B = c->nscen;
  for (int id = 0; id < c->nfault; id++) {
    const cgfd_fault_desc &f = d->fault[id];
    FaultDev &F = c->flt[id];
    F.i0 = f.fault_x_index + 3;
    std::vector<float> h;
    float *dp;
    int *ip;
    int rc;
    /* per-scenario blocks */
    float *fb;
    int *ib;
    if ((rc = c->alloc(&fb, (size_t)B * 45 * ss))) return rc;
    if ((rc = c->alloc(&ib, (size_t)B * 2 * ss))) return rc;
    HIPCHK(hipMemset(ib, 0, (size_t)B * 2 * ss * 4));  /* flag_rup = 0 (fault_info.cu:1069), init_t0_flag = 0 (:960) */
    h.assign(45 * ss, 0.f);
    for (int sn = 0; sn < B; sn++) {
      const cgfd_fault_desc &fs = ds[sn].fault[id];
      std::fill(h.begin(), h.end(), 0.f);  /* T1, hT1, mT1 start at zero (fault_wav_t.cu:59-78) */
      for (size_t p = 0; p < ss; p++) {
        h[p] = fs.T0x[p]; h[ss + p] = fs.T0y[p]; h[2 * ss + p] = fs.T0z[p];
        h[3 * ss + p] = fs.mu_s[p]; h[4 * ss + p] = fs.mu_d[p]; h[5 * ss + p] = fs.Dc[p]; h[6 * ss + p] = fs.C0[p];
        h[7 * ss + p] = fs.Tn0 ? fs.Tn0[p] : 0.f;
        h[8 * ss + p] = fs.Ts10 ? fs.Ts10[p] : 0.f;
        h[9 * ss + p] = fs.Ts20 ? fs.Ts20[p] : 0.f;
        h[17 * ss + p] = -9999.9f;  /* Init_t0 (fault_info.cu:1068) */
      }
      HIPCHK(hipMemcpy(fb + (size_t)sn * 45 * ss, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    F.T0 = fb;
    F.fric = fb + 3 * ss;
    F.out = fb + 7 * ss;
    F.T1 = fb + 18 * ss;
    F.hT1 = fb + 39 * ss;
    F.mT1 = fb + 42 * ss;
    F.flag_rup = ib;
    F.init_t0_flag = ib + ss;
    /* FP64 slip accumulators, zero (fault_info.cu:945-947) */
    double *db = nullptr;
    if ((rc = c->alloc(&db, (size_t)B * 3 * ss))) return rc;
    HIPCHK(hipMemset(db, 0, (size_t)B * 3 * ss * sizeof(double)));
    F.slip64 = db;
    /* masks */
    const int *msk[4] = {f.united, f.faultgrid, f.rup_index_y, f.rup_index_z};
    const int *dmsk[4];
    for (int q = 0; q < 4; q++) {
      if ((rc = c->alloc(&ip, ss))) return rc;
      HIPCHK(hipMemcpy(ip, msk[q], ss * 4, hipMemcpyHostToDevice));
      dmsk[q] = ip;
    }
    F.united = dmsk[0]; F.faultgrid = dmsk[1]; F.rupy = dmsk[2]; F.rupz = dmsk[3];
    /* rho_f [2][ss] */
    if ((rc = c->alloc(&dp, 2 * ss))) return rc;
    HIPCHK(hipMemcpy(dp, f.rho_f, 2 * ss * 4, hipMemcpyHostToDevice));
    F.rho_f = dp;
    /* strike / dip vectors [6][ss] */
    h.assign(6 * ss, 0.f);
    for (size_t p = 0; p < ss; p++)
      for (int e = 0; e < 3; e++) { h[e * ss + p] = f.vec_s1[p * 3 + e]; h[(3 + e) * ss + p] = f.vec_s2[p * 3 + e]; }
    if ((rc = c->alloc(&dp, h.size()))) return rc;
    HIPCHK(hipMemcpy(dp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    F.vs = dp;
    /* D [12][9][ss] */
    h.assign(12 * 9 * ss, 0.f);
    for (int q = 0; q < 12; q++) soa9(h, q * 9 * ss, f.D[q], ss);
    if ((rc = c->alloc(&dp, h.size()))) return rc;
    HIPCHK(hipMemcpy(dp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    F.D = dp;
    /* method-1 and method-2 maps (only the ones the method reads are required) */
    if (d->imethod == 1) {
      h.assign(5 * 9 * ss, 0.f);
      for (int q = 0; q < 5; q++) soa9(h, q * 9 * ss, f.matPlus2Min[q], ss);
      if ((rc = c->alloc(&dp, h.size()))) return rc;
      HIPCHK(hipMemcpy(dp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
      F.P2M = dp;
      for (int q = 0; q < 5; q++) soa9(h, q * 9 * ss, f.matMin2Plus[q], ss);
      if ((rc = c->alloc(&dp, h.size()))) return rc;
      HIPCHK(hipMemcpy(dp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
      F.M2P = dp;
    } else {
      h.assign(6 * 9 * ss, 0.f);
      for (int q = 0; q < 6; q++) soa9(h, q * 9 * ss, f.matM2[q], ss);
      if ((rc = c->alloc(&dp, h.size()))) return rc;
      HIPCHK(hipMemcpy(dp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
      F.M2 = dp;
    }
    /* free-surface row [14][9][ny] */
    h.assign(14 * 9 * ny, 0.f);
    if (c->isfree)
      for (int q = 0; q < 14; q++) soa9(h, q * 9 * ny, f.matF[q], ny);
    if ((rc = c->alloc(&dp, h.size()))) return rc;
    HIPCHK(hipMemcpy(dp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    F.Fm = dp;
    /* the metric on the plane's column, contiguous along j */
    if ((rc = c->alloc(&dp, 10 * ss))) return rc;
    launch_fault_metric(c->g, c->met, F.i0, dp, c->st);
    F.fmet = dp;
  }
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipGetLastError());
  return CGFD_OK;
}

/* scenarios of a batch must agree on everything the kernels share */
static int check_batch(const cgfd_desc *ds, int nscen) {
  const cgfd_desc &a = ds[0];
  for (int s = 1; s < nscen; s++) {
    const cgfd_desc &b = ds[s];
    int rc = check_desc(&b);
    if (rc) return rc;
    bool same = a.ni == b.ni && a.nj == b.nj && a.nk == b.nk && a.free_top == b.free_top && a.imethod == b.imethod &&
                a.number_fault == b.number_fault && a.dt == b.dt && a.compat_flags == b.compat_flags &&
                a.device_id == b.device_id;
    for (int dim = 0; dim < 3 && same; dim++)
      for (int side = 0; side < 2; side++)
        same = same && a.pml_is_sides[dim][side] == b.pml_is_sides[dim][side] &&
               a.pml_ni1[dim][side] == b.pml_ni1[dim][side] && a.pml_ni2[dim][side] == b.pml_ni2[dim][side] &&
               a.pml_nj1[dim][side] == b.pml_nj1[dim][side] && a.pml_nj2[dim][side] == b.pml_nj2[dim][side] &&
               a.pml_nk1[dim][side] == b.pml_nk1[dim][side] && a.pml_nk2[dim][side] == b.pml_nk2[dim][side];
    for (int id = 0; id < a.number_fault && same; id++)
      same = a.fault[id].fault_x_index == b.fault[id].fault_x_index;
    if (!same)
      return cgfd_set_error(CGFD_ERR_ARG, "scenario %d differs from scenario 0 in grid, boundaries, method, faults or dt",
                            s);
    if (b.nt_total < 0) return cgfd_set_error(CGFD_ERR_ARG, "scenario %d: nt_total < 0", s);
  }
  return CGFD_OK;
}

static void apply_auto_variant(cgfd_core *c);
static int scratch_setup(int dev);

static int core_create(const cgfd_desc *ds, int nscen, cgfd_core **out) {
  if (!out) return cgfd_set_error(CGFD_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (!ds) return cgfd_set_error(CGFD_ERR_ARG, "null descriptor");
  if (nscen < 1 || nscen > kMaxScen) return cgfd_set_error(CGFD_ERR_ARG, "nscen=%d out of range [1,%d]", nscen, kMaxScen);
  const cgfd_desc *d = &ds[0];
  int rc = check_desc(d);
  if (rc) return rc;
  if ((rc = check_batch(ds, nscen))) return rc;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return cgfd_set_error(CGFD_ERR_NODEV, "no HIP device available");
  if (d->device_id < 0 || d->device_id >= ndev)
    return cgfd_set_error(CGFD_ERR_ARG, "device_id %d out of range (%d devices)", d->device_id, ndev);
  cgfd_core *c = new (std::nothrow) cgfd_core();
  if (!c) return cgfd_set_error(CGFD_ERR_NOMEM, "host allocation failed");
  auto fail = [&](int code) { cgfd_core_destroy(c); return code; };
  c->dev = d->device_id;
  if (hipSetDevice(c->dev) != hipSuccess) return fail(cgfd_set_error(CGFD_ERR_HIP, "hipSetDevice failed"));
  if ((rc = scratch_setup(c->dev))) return fail(rc);
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
    return fail(cgfd_set_error(CGFD_ERR_HIP, "hipStreamCreate failed"));
  for (int q = 0; q < 2; q++)
    if (hipStreamCreateWithFlags(&c->side[q], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join[q], hipEventDisableTiming) != hipSuccess)
      return fail(cgfd_set_error(CGFD_ERR_HIP, "side stream / event creation failed"));
  if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess)
    return fail(cgfd_set_error(CGFD_ERR_HIP, "hipEventCreate failed"));
  Grid &g = c->g;
  c->ni = d->ni; c->nj = d->nj; c->nk = d->nk;
  g.ni = d->ni; g.nj = d->nj; g.nk = d->nk;
  g.nx = g.ni + 6; g.ny = g.nj + 6; g.nz = g.nk + 6;
  g.nk2 = 3 + g.nk - 1;
  g.siy = ((g.nx + 31) / 32) * 32;  /* 128-byte aligned rows on the device */
  g.siz = g.siy * g.ny; g.sicmp = g.siz * g.nz; g.sslice = (long)g.ny * g.nz;
  c->dt = d->dt; c->t0 = d->t0; c->nt_total = d->nt_total;
  c->isfree = d->free_top ? 1 : 0;
  c->imethod = d->imethod;
  c->nfault = d->number_fault;
  c->fix_q1 = (d->compat_flags & CGFD_COMPAT_FIX_F2W_SWAP) ? 1 : 0;
  c->fix_sta = (d->compat_flags & CGFD_COMPAT_FIX_STATION_INTERP) ? 1 : 0;
  c->nan_every = d->qc_check_nan_every;
  const long sc = g.sicmp;
  const int B = nscen;
  c->nscen = B;
  for (int sn = 0; sn < B; sn++) c->scen_nt[sn] = ds[sn].nt_total;
  c->bat.nact = B;
  for (int sn = 0; sn < B; sn++) c->bat.act[sn] = sn;

  /* wavefield: 4 levels per scenario, zeroed (alloc.cu:448-457) */
  c->bat.wstride = 4L * 9 * sc;
  if ((rc = c->alloc(&c->wave, (size_t)B * 4 * 9 * sc))) return fail(rc);
  if (hipMemset(c->wave, 0, (size_t)B * 4 * 9 * sc * 4) != hipSuccess)
    return fail(cgfd_set_error(CGFD_ERR_HIP, "memset failed"));
  /* metric and media */
  if ((rc = c->alloc(&c->met, (size_t)10 * sc))) return fail(rc);
  if ((rc = c->alloc(&c->lam, (size_t)3 * sc))) return fail(rc);
  c->mu = c->lam + sc;
  c->slw = c->lam + 2 * sc;
  /* staging buffer for host <-> device copies of volume arrays */
  if ((rc = c->alloc(&c->stage, (size_t)g.nx * g.ny * g.nz))) return fail(rc);
  if ((rc = host_to_dev(c, c->met, d->metric, 10)) || (rc = host_to_dev(c, c->lam, d->lambda, 1)) ||
      (rc = host_to_dev(c, c->mu, d->mu, 1)) || (rc = host_to_dev(c, c->slw, d->rho, 1)))
    return fail(rc);
  /* free-surface matrices (bdry_free_set calloc, bdry_t.cu:60-71) */
  if ((rc = c->alloc(&c->matVx2Vz, (size_t)2 * 9 * g.siz))) return fail(rc);
  c->matVy2Vz = c->matVx2Vz + 9 * g.siz;
  if (hipMemset(c->matVx2Vz, 0, (size_t)2 * 9 * g.siz * 4) != hipSuccess) return fail(cgfd_set_error(CGFD_ERR_HIP, "memset failed"));
  /* PML: profiles shared; auxiliaries (4 levels x 9 x slab) of all faces in
     one block per scenario */
  long paux = 0;
  for (int dim = 0; dim < 3; dim++)
    for (int side = 0; side < 2; side++) {
      if (!d->pml_is_sides[dim][side]) continue;
      int ax0, an;
      pml_row_layout(dim, d->pml_ni1[dim][side], d->pml_ni2[dim][side], ax0, an);
      long n = (long)an * (d->pml_nj2[dim][side] - d->pml_nj1[dim][side] + 1) *
               (d->pml_nk2[dim][side] - d->pml_nk1[dim][side] + 1);
      paux += 36 * n;
    }
  float *auxblk = nullptr;
  c->bat.pstride = paux;
  if (paux > 0) {
    if ((rc = c->alloc(&auxblk, (size_t)B * paux))) return fail(rc);
    if (hipMemset(auxblk, 0, (size_t)B * paux * 4) != hipSuccess) return fail(cgfd_set_error(CGFD_ERR_HIP, "memset failed"));
  }
  long paoff = 0;
  for (int dim = 0; dim < 3; dim++)
    for (int side = 0; side < 2; side++) {
      PmlFace &f = c->pml[dim][side];
      f.on = d->pml_is_sides[dim][side] ? 1 : 0;
      if (!f.on) continue;
      c->pml_on = 1;
      f.i1 = d->pml_ni1[dim][side]; f.i2 = d->pml_ni2[dim][side];
      f.j1 = d->pml_nj1[dim][side]; f.j2 = d->pml_nj2[dim][side];
      f.k1 = d->pml_nk1[dim][side]; f.k2 = d->pml_nk2[dim][side];
      pml_row_layout(dim, f.i1, f.i2, f.ax0, f.an);
      f.bn = f.j2 - f.j1 + 1;
      int cn = f.k2 - f.k1 + 1;
      f.n = (long)f.an * f.bn * cn;
      int np = dim == 0 ? f.i2 - f.i1 + 1 : (dim == 1 ? f.bn : cn);
      float *cf;
      if ((rc = c->alloc(&cf, (size_t)3 * np))) return fail(rc);
      if (hipMemcpy(cf, d->pml_A[dim][side], np * 4, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(cf + np, d->pml_B[dim][side], np * 4, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(cf + 2 * np, d->pml_D[dim][side], np * 4, hipMemcpyHostToDevice) != hipSuccess)
        return fail(cgfd_set_error(CGFD_ERR_HIP, "PML coefficient upload failed"));
      f.A = cf; f.B = cf + np; f.D = cf + 2 * np;
      float *aux = auxblk + paoff;
      paoff += 36 * f.n;
      f.pre = aux; f.la = aux + 9 * f.n; f.lb = aux + 18 * f.n; f.lc = aux + 27 * f.n;
    }
  /* faults */
  if (c->nfault > 0) {
    c->bat.fwstride = 4L * c->nfault * 18 * g.sslice;
    c->bat.fpf = 45L * g.sslice;
    c->bat.fpi = 2L * g.sslice;
    c->bat.fpd = 3L * g.sslice;
    size_t n = (size_t)B * c->bat.fwstride;
    if ((rc = c->alloc(&c->fwav, n))) return fail(rc);
    if (hipMemset(c->fwav, 0, n * 4) != hipSuccess) return fail(cgfd_set_error(CGFD_ERR_HIP, "memset failed"));
    if ((rc = upload_faults(c, ds))) return fail(rc);
    if ((rc = c->alloc(&c->d_ftab, (size_t)c->nfault))) return fail(rc);
    if (hipMemcpy(c->d_ftab, c->flt, sizeof(FaultDev) * c->nfault, hipMemcpyHostToDevice) != hipSuccess)
      return fail(cgfd_set_error(CGFD_ERR_HIP, "fault table upload failed"));
  }
  {
    const int ntop = c->isfree ? (g.nk < 4 ? g.nk : 4) : 0;
    c->nvol = g.nk - ntop;
  }
  if ((rc = c->alloc(&c->it_dev, 3))) return fail(rc);
  if (hipMemset(c->it_dev, 0, 12) != hipSuccess) return fail(cgfd_set_error(CGFD_ERR_HIP, "memset failed"));
  c->nan_flag = c->it_dev + 1;
  if (c->isfree) {
    c->bat.pgstride = 18L * g.siz;
    if ((rc = c->alloc(&c->PG, (size_t)B * 18 * g.siz))) return fail(rc);
    if (hipMemset(c->PG, 0, (size_t)B * 18 * g.siz * 4) != hipSuccess)
      return fail(cgfd_set_error(CGFD_ERR_HIP, "memset failed"));
    c->Dis = c->PG + 15 * g.siz;
    launch_dvh2dvz(g, c->met, c->lam, c->mu, c->matVx2Vz, c->matVy2Vz, c->st);
  }
  launch_vcol_check(g, c->met, c->it_dev + 2, c->st);
  int not_vertical = 1;
  if (hipMemcpyAsync(&not_vertical, c->it_dev + 2, 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess || hipGetLastError() != hipSuccess)
    return fail(cgfd_set_error(CGFD_ERR_HIP, "setup kernels failed"));
  c->vcol_grid = not_vertical == 0;
  apply_auto_variant(c);  /* the tuned schedule of this shape, if any */
  *out = c;
  return CGFD_OK;
}

extern "C" int cgfd_core_create(const cgfd_desc *d, cgfd_core **out) { return core_create(d, 1, out); }

extern "C" int cgfd_batch_create(const cgfd_desc *descs, int nscen, cgfd_core **out) {
  return core_create(descs, nscen, out);
}

extern "C" int cgfd_core_destroy(cgfd_core *c) {
  if (!c) return CGFD_OK;
  (void)hipSetDevice(c->dev);
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (int q = 0; q < 2; q++)
    if (c->side[q]) (void)hipStreamSynchronize(c->side[q]);
  c->drop_graph();
  for (int q = 0; q < 2; q++) {
    if (c->ev_join[q]) (void)hipEventDestroy(c->ev_join[q]);
    if (c->side[q]) (void)hipStreamDestroy(c->side[q]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return CGFD_OK;
}

namespace {

/* arguments of the volume kernels for one stage */
void fill_vol(cgfd_core *c, VolArgs &v, int istage, const float *cur, float *tnew) {
  v.g = c->g;
  v.cur = cur;
  v.pre = c->wlev(c->lv_pre);
  v.tnew = tnew;
  v.end = c->wlev(c->lv_a);
  v.lb = c->wlev(c->lv_b);
  v.met = c->met; v.lam = c->lam; v.mu = c->mu; v.slw = c->slw;
  v.matVx2Vz = c->matVx2Vz; v.matVy2Vz = c->matVy2Vz;
  for (int dim = 0; dim < 3; dim++)
    for (int side = 0; side < 2; side++) v.pml[dim][side] = c->pml[dim][side];
  v.pml_on = c->pml_on;
  v.isfree = c->isfree;
  v.istage = istage;
  v.ca = istage < 3 ? kRkA[istage] * c->dt : 0.0f;
  v.cb = kRkB[istage] * c->dt;
  v.fs.nfault = c->nfault;
  for (int q = 0; q < c->nfault; q++) { v.fs.i0[q] = c->flt[q].i0; v.fs.united[q] = c->flt[q].united; }
  v.zchunk = c->zchunk;
  v.vcol = c->vcol_grid && c->vcol_on;
}

void fill_fault(cgfd_core *c, FaultArgs &f, int istage, int idir, int jdir, int kdir, float *wcur, float *wnew,
                float *fcur, float *fnew) {
  f.g = c->g;
  f.met = c->met; f.lam = c->lam; f.mu = c->mu; f.slw = c->slw;
  f.matVx2Vz = c->matVx2Vz; f.matVy2Vz = c->matVy2Vz;
  f.w_cur = wcur; f.w_pre = c->wlev(c->lv_pre); f.w_tnew = wnew; f.w_end = c->wlev(c->lv_a);
  f.w_b = c->wlev(c->lv_b);
  f.vcol = c->vcol_grid && c->vcol_on;
  f.f_cur = fcur; f.f_pre = c->flev(c->fl_pre); f.f_tnew = fnew; f.f_end = c->flev(c->fl_end);
  f.nfault = c->nfault;
  f.ftab = c->d_ftab;
  f.sfp = 0;
  f.sfi = 0;
  f.istage = istage; f.idir = idir; f.jdir = jdir; f.kdir = kdir;
  f.isfree = c->isfree; f.imethod = c->imethod; f.fix_q1 = c->fix_q1;
  f.dt = c->dt;
  f.ca = istage < 3 ? kRkA[istage] * c->dt : 0.0f;
  f.cb = kRkB[istage] * c->dt;
  f.rkb = kRkB[istage];
  f.it_dev = c->it_dev;
  f.tape = c->tape;
  f.tape_it0 = c->tape_it0;
  f.tape_nt = c->tape_nt;
  f.tape_sstride = (long)c->nfault * c->tape_nt * kTapeFloats * c->g.sslice;
}

/* volume kernel variants (see cgfd_core_set_variant): 1 direct loads; 20
   an interior kernel in slab order and a lean kernel for the tiles touching
   a CFS-PML slab; 23, 24, 32 and 33 the box split */
void launch_vol(cgfd_core *c, const VolArgs &va, int idir, int jdir, int kdir) {
  launch_volume(c->variant, va, c->bat, idir, jdir, kdir, 0, c->st, c->st_pml());
}

/* one time step; if ev != nullptr, event pairs bracket the volume launches */
int do_step(cgfd_core *c, hipEvent_t *ev, int *nev) {
  const int ipair = c->it_host % 8;
  const int nk = c->g.nk;
  const int ntop = c->isfree ? (nk < 4 ? nk : 4) : 0;
  const int nvol = nk - ntop;
  if (c->refresh_active() == 0) return CGFD_OK;  /* every scenario of the batch has retired */
  if (c->st_n > 0)
    launch_record_stations(c->bat, c->wlev(c->lv_pre), c->g.sicmp, c->st_term, c->st_fac, c->st_nterm, c->st_n,
                           c->st_buf, c->st_nt, c->it_dev, c->st_it0, c->st);
  if (c->fr_n > 0)
    launch_record_fault(c->bat, c->d_ftab, c->fr_idx, c->fr_fac, c->fr_nterm, c->fr_n, c->fr_buf, c->fr_nt, c->it_dev,
                        c->fr_it0, c->g.sslice, c->st);
  for (int s = 0; s < 4; s++) {
    const int idir = (kPairs[ipair][0] + s) % 2;
    const int jdir = (kPairs[ipair][1] + s) % 2;
    const int kdir = (kPairs[ipair][2] + s) % 2;
    const int lcur = s == 0 ? c->lv_pre : (s == 1 ? c->lv_a : (s == 2 ? c->lv_b : c->lv_c));
    const int lnew = s == 0 ? c->lv_a : (s == 1 ? c->lv_b : c->lv_c);
    float *wcur = c->wlev(lcur);
    float *wnew = s < 3 ? c->wlev(lnew) : nullptr;
    float *fcur = nullptr, *fnew = nullptr;
    if (c->nfault > 0) {
      const int fcurl = s == 0 ? c->fl_pre : (s == 2 ? c->fl_b : c->fl_a);
      const int fnewl = s == 0 ? c->fl_a : (s == 1 ? c->fl_b : c->fl_a);
      fcur = c->flev(fcurl);
      fnew = s < 3 ? c->flev(fnewl) : nullptr;
      FaultArgs fa;
      fill_fault(c, fa, s, idir, jdir, kdir, wcur, wnew, fcur, fnew);
      launch_wave2fault(fa, c->bat, c->st);
      launch_trial(fa, c->bat, c->st);
    }
    VolArgs va;
    fill_vol(c, va, s, wcur, wnew);
    /* timed region: the volume kernels; with the stage fork on, from the
       fork to the join (the top-layer and fault kernels run beside them) */
    if (ev && nvol > 0) HIPCHK(hipEventRecord(ev[2 * (*nev)], c->st));
    if (int rc = c->fork()) return rc;
    if (nvol > 0) {
      va.k_lo = 0;
      va.nk_run = nvol;
      launch_vol(c, va, idir, jdir, kdir);
      if (ev && !c->overlap) {
        HIPCHK(hipEventRecord(ev[2 * (*nev) + 1], c->st));
        (*nev)++;
      }
    }
    if (ntop > 0) {
      va.k_lo = nk - ntop;
      va.nk_run = ntop;
      launch_volume(c->variant, va, c->bat, idir, jdir, kdir, 1, c->st_pml());
    }
    if (c->nfault > 0) {
      FaultArgs fa;
      fill_fault(c, fa, s, idir, jdir, kdir, wcur, wnew, fcur, fnew);
      launch_fault_update(fa, c->bat, c->st_fault());
    }
    if (int rc = c->join()) return rc;
    if (ev && nvol > 0 && c->overlap) {
      HIPCHK(hipEventRecord(ev[2 * (*nev) + 1], c->st));
      (*nev)++;
    }
  }
  if (c->isfree) launch_pg(c->g, c->bat, c->wlev(c->lv_a), c->wlev(c->lv_pre), c->PG, c->Dis, c->dt, c->st);
  launch_step_end(c->it_dev, c->st);
  HIPCHK(hipGetLastError());
  /* the new step starts from t_{n+1}, written over level A (the reference
     swaps pre and end, drv_rk_curv_col.cu:622-635) */
  std::swap(c->lv_pre, c->lv_a);
  std::swap(c->fl_pre, c->fl_end);
  for (int dim = 0; dim < 3; dim++)
    for (int side = 0; side < 2; side++) std::swap(c->pml[dim][side].pre, c->pml[dim][side].la);
  c->it_host++;
  if (c->nan_every > 0 && (c->it_host % c->nan_every) == 0) {
    for (int q = 0; q < c->bat.nact; q++)
      launch_nan_check(c->wlev(c->lv_pre) + c->bat.act[q] * c->bat.wstride, 9 * c->g.sicmp, c->nan_flag, c->st);
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, c->nan_flag, 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (flag) return cgfd_set_error(CGFD_ERR_NAN, "non-finite wavefield value after step %d", c->it_host);
  }
  return CGFD_OK;
}

}  // namespace

namespace {
/* Captures kGraphSteps steps starting at a step index that is a multiple of
   kGraphSteps. Capturing only records the launches: the host-side level
   rotation is advanced by the recording and, after an even number of swaps,
   is back where it started; the step counter is restored. */
int capture_graph(cgfd_core *c) {
  const int it_save = c->it_host;
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  int rc = CGFD_OK;
  for (int s = 0; s < cgfd_core::kGraphSteps && rc == CGFD_OK; s++) rc = do_step(c, nullptr, nullptr);
  hipError_t e = hipStreamEndCapture(c->st, &graph);
  c->it_host = it_save;
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return cgfd_set_error(CGFD_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    return cgfd_set_error(CGFD_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  return CGFD_OK;
}
}  // namespace

extern "C" int cgfd_core_run(cgfd_core *c, int nsteps) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  if (nsteps < 0) return cgfd_set_error(CGFD_ERR_ARG, "nsteps must be >= 0");
  HIPCHK(hipSetDevice(c->dev));
  const int G = cgfd_core::kGraphSteps;
  const bool graph = c->use_graph && c->nan_every <= 0;
  int left = nsteps;
  while (left > 0) {
    if (c->refresh_active() == 0) break;  /* every scenario of the batch has retired */
    bool stable = true;                   /* no scenario retires inside the next G steps */
    for (int sn = 0; sn < c->nscen && c->nscen > 1; sn++)
      if (c->it_host < c->scen_nt[sn] && c->scen_nt[sn] < c->it_host + G) stable = false;
    if (graph && stable && left >= G && c->it_host % G == 0) {
      if (!c->gexec) {
        int rc = capture_graph(c);
        if (rc) return rc;
      }
      HIPCHK(hipGraphLaunch(c->gexec, c->st));
      c->it_host += G;
      left -= G;
      continue;
    }
    int rc = do_step(c, nullptr, nullptr);
    if (rc) return rc;
    left--;
  }
  return CGFD_OK;
}

/* ---- step-boundary state (checkpoint / restart) ----
   What a step reads from earlier steps (drv_rk_curv_col.cu:118-127,161-166,
   fault_info.h:97-118, fault_wav_t.h:30-38): the wavefield and fault-state
   "pre" levels, the PML auxiliaries' "pre" level, the fault outputs and
   accumulators (Tn .. Init_t0), hT1 and mT1, the rupture flags, the surface
   PG and displacement, and the step counter. Copied part by part in this
   order into one caller buffer. */
namespace {
struct StatePart {
  void *p;
  size_t bytes;
};
std::vector<StatePart> state_parts(cgfd_core *c) {
  std::vector<StatePart> v;
  const Grid &g = c->g;
  const size_t ss = (size_t)g.sslice;
  v.push_back({c->wlev(c->lv_pre), (size_t)9 * g.sicmp * 4});
  if (c->nfault > 0) {
    v.push_back({c->flev(c->fl_pre), (size_t)c->nfault * 18 * ss * 4});
    for (int id = 0; id < c->nfault; id++) {
      const FaultDev &F = c->flt[id];
      v.push_back({F.out, 11 * ss * 4});
      v.push_back({F.hT1, 3 * ss * 4});
      v.push_back({F.mT1, 3 * ss * 4});
      v.push_back({F.flag_rup, ss * 4});
      v.push_back({F.init_t0_flag, ss * 4});
      v.push_back({F.slip64, 3 * ss * 8});
    }
  }
  for (int dim = 0; dim < 3; dim++)
    for (int side = 0; side < 2; side++)
      if (c->pml[dim][side].on) v.push_back({c->pml[dim][side].pre, (size_t)9 * c->pml[dim][side].n * 4});
  if (c->PG) v.push_back({c->PG, (size_t)18 * g.siz * 4});
  v.push_back({c->it_dev, 4});
  return v;
}
}  // namespace

extern "C" int cgfd_core_state_bytes(cgfd_core *c, size_t *bytes) {
  if (!c || !bytes) return cgfd_set_error(CGFD_ERR_ARG, "bad arguments");
  if (c->nscen > 1) return cgfd_set_error(CGFD_ERR_ARG, "state save/load covers single-scenario cores");
  size_t n = 0;
  for (const StatePart &q : state_parts(c)) n += q.bytes;
  *bytes = n;
  return CGFD_OK;
}

extern "C" int cgfd_core_save_state(cgfd_core *c, void *dst, int dst_on_device) {
  if (!c || !dst) return cgfd_set_error(CGFD_ERR_ARG, "bad arguments");
  if (c->nscen > 1) return cgfd_set_error(CGFD_ERR_ARG, "state save/load covers single-scenario cores");
  HIPCHK(hipSetDevice(c->dev));
  const hipMemcpyKind kind = dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  size_t off = 0;
  for (const StatePart &q : state_parts(c)) {
    HIPCHK(hipMemcpyAsync((char *)dst + off, q.p, q.bytes, kind, c->st));
    off += q.bytes;
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return CGFD_OK;
}

extern "C" int cgfd_core_load_state(cgfd_core *c, const void *src, int src_on_device) {
  if (!c || !src) return cgfd_set_error(CGFD_ERR_ARG, "bad arguments");
  if (c->nscen > 1) return cgfd_set_error(CGFD_ERR_ARG, "state save/load covers single-scenario cores");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  c->drop_graph();
  const hipMemcpyKind kind = src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  size_t off = 0;
  for (const StatePart &q : state_parts(c)) {
    HIPCHK(hipMemcpyAsync(q.p, (const char *)src + off, q.bytes, kind, c->st));
    off += q.bytes;
  }
  int it = 0;
  HIPCHK(hipMemcpyAsync(&it, c->it_dev, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (it < 0) return cgfd_set_error(CGFD_ERR_ARG, "loaded state has a negative step counter");
  c->it_host = it;
  return CGFD_OK;
}

/* ---- discrete adjoint: gradient of J = sum_n sum_s r_n(s) . w_n(s) with
   respect to the state at the first recorded step and, with faults, to the
   friction parameters and initial traction of every split node, for the
   residuals r of the recorded station traces; the gradient of a
   least-squares trace misfit is this with r = recorded - target. Reverse
   sweep over the recorded steps, one reverse RK4 step per recorded step
   (k_adjoint.hip: wave update; k_adj_fault.hip: fault and friction, from the
   tape the forward run wrote). ---- */
extern "C" int cgfd_adj_forward(cgfd_core *c, int nt) {
  if (!c || nt <= 0) return cgfd_set_error(CGFD_ERR_ARG, "bad arguments");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  c->drop_graph();
  c->tape_nt = 0;
  c->tape_it0 = c->it_host;
  if (c->nfault == 0) return CGFD_OK;  /* the wave update needs no tape */
  const size_t n = (size_t)c->nscen * c->nfault * nt * kTapeFloats * c->g.sslice;
  if (int rc = c->grow(&c->tape, &c->tape_cap, n)) return rc;
  HIPCHK(hipMemset(c->tape, 0, n * 4));
  c->tape_nt = nt;
  return CGFD_OK;
}

static int scen_steps(const cgfd_core *c, int s);

namespace {
/* split nodes whose fault columns reach a CFS-PML slab are not covered */
int check_fault_columns(cgfd_core *c) {
  const Grid &g = c->g;
  std::vector<int> un((size_t)g.sslice);
  for (int id = 0; id < c->nfault; id++) {
    HIPCHK(hipMemcpy(un.data(), c->flt[id].united, un.size() * 4, hipMemcpyDeviceToHost));
    const int i0 = c->flt[id].i0;
    for (int k = 3; k < g.nk + 3; k++)
      for (int j = 3; j < g.nj + 3; j++) {
        if (un[(size_t)j + (size_t)k * g.ny] != 0) continue;
        for (int dim = 0; dim < 3; dim++)
          for (int side = 0; side < 2; side++) {
            const PmlFace &f = c->pml[dim][side];
            if (f.on && j >= f.j1 && j <= f.j2 && k >= f.k1 && k <= f.k2 && i0 + 3 >= f.i1 && i0 - 3 <= f.i2)
              return cgfd_set_error(CGFD_ERR_ARG,
                                    "adjoint: split node (j=%d, k=%d) of fault %d has fault columns in a CFS-PML slab", j,
                                    k, id);
          }
      }
  }
  return CGFD_OK;
}

int adj_sweep(cgfd_core *c, int scen, const float *residual, const float *residual_fault, int nt, float *grad,
              float *grad_fault) {
  if (!c || !residual) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  if (scen < 0 || scen >= c->nscen) return cgfd_set_error(CGFD_ERR_ARG, "scenario %d out of range", scen);
  if (c->nscen > 1 && residual_fault)
    return cgfd_set_error(CGFD_ERR_ARG, "adjoint: fault-receiver residuals need a single-scenario core");
  if (scen_steps(c, scen) < c->st_it0 + nt - 1)
    return cgfd_set_error(CGFD_ERR_STATE, "adjoint: scenario %d ran %d of the %d recorded steps", scen,
                          scen_steps(c, scen) - c->st_it0, nt - 1);
  if (c->st_n <= 0) return cgfd_set_error(CGFD_ERR_STATE, "adjoint: no stations set");
  if (nt < 1 || nt > c->st_nt) return cgfd_set_error(CGFD_ERR_ARG, "adjoint: nt %d outside 1..%d", nt, c->st_nt);
  if (c->nfault > 0 && (!c->tape || c->tape_it0 != c->st_it0 || c->tape_nt < nt))
    return cgfd_set_error(CGFD_ERR_STATE,
                          "adjoint: a core with a fault needs the friction tape: call cgfd_adj_forward with at least "
                          "%d steps at the step the stations were set, before running", nt);
  if (residual_fault) {
    if (c->fr_n <= 0 || c->fr_it0 != c->st_it0 || c->fr_nt < nt)
      return cgfd_set_error(CGFD_ERR_STATE,
                            "adjoint: fault-receiver residuals need receivers set at the step the stations were set, "
                            "with at least %d samples", nt);
    for (size_t q = 0; q < (size_t)c->fr_n * nt; q++)
      for (int v = 0; v < 3; v++)
        if (residual_fault[9 * q + v] != 0.0f)
          return cgfd_set_error(CGFD_ERR_ARG,
                                "adjoint: fault-receiver residuals on Tn, Ts1, Ts2 are not supported (must be zero)");
  }
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  const Grid &g = c->g;
  const long sc = g.sicmp, lv = 9 * sc, ss = g.sslice;
  int rc;
  if (c->nfault > 0 && (rc = check_fault_columns(c))) return rc;
  long npml = 0;  /* floats of the PML adjoint block */
  for (int dim = 0; dim < 3; dim++)
    for (int side = 0; side < 2; side++)
      if (c->pml[dim][side].on) npml += 45 * c->pml[dim][side].n;
  if (!c->adj) {
    if ((rc = c->alloc(&c->adj, (size_t)4 * lv))) return rc;
    if ((rc = c->alloc(&c->adj_sg, (size_t)9 * sc))) return rc;
    HIPCHK(hipMemset(c->adj_sg, 0, (size_t)9 * sc * 4));  /* ghost entries stay zero */
    if (npml > 0 && (rc = c->alloc(&c->adj_pml, (size_t)npml))) return rc;
  }
  const size_t nadjf = (size_t)c->nfault * kAdjFaultFloats * ss;
  if (c->nfault > 0 && !c->adjf && (rc = c->alloc(&c->adjf, nadjf))) return rc;
  HIPCHK(hipMemsetAsync(c->adj, 0, (size_t)4 * lv * 4, c->st));
  if (c->nfault > 0) HIPCHK(hipMemsetAsync(c->adjf, 0, nadjf * 4, c->st));
  AdjPml pm{};
  pm.on = c->pml_on;
  pm.le = 0;
  pm.lp = 3;
  pm.fs.nfault = c->nfault;
  for (int q = 0; q < c->nfault; q++) { pm.fs.i0[q] = c->flt[q].i0; pm.fs.united[q] = c->flt[q].united; }
  {
    long off = 0;
    for (int dim = 0; dim < 3; dim++)
      for (int side = 0; side < 2; side++) {
        const PmlFace &f = c->pml[dim][side];
        AdjFace &a = pm.f[2 * dim + side];
        a.on = f.on;
        a.dim = dim;
        if (!f.on) continue;
        a.i1 = f.i1; a.i2 = f.i2; a.j1 = f.j1; a.j2 = f.j2; a.k1 = f.k1; a.k2 = f.k2;
        a.an = f.an; a.ax0 = f.ax0; a.bn = f.bn; a.n = f.n;
        a.A = f.A; a.B = f.B; a.D = f.D;
        a.lev = c->adj_pml + off;
        a.dsg = c->adj_pml + off + 36 * f.n;
        HIPCHK(hipMemsetAsync(a.lev, 0, (size_t)36 * f.n * 4, c->st));
        off += 45 * f.n;
      }
  }
  AdjFaultArgs fa{};
  fa.g = g;
  fa.met = c->met; fa.lam = c->lam; fa.mu = c->mu; fa.slw = c->slw;
  fa.matVx2Vz = c->matVx2Vz; fa.matVy2Vz = c->matVy2Vz;
  fa.ftab = c->d_ftab;
  if (c->nfault > 0 && scen > 0) {
    /* the scenario's own parameters, outputs and traction columns */
    if (!c->d_ftab_adj && (rc = c->alloc(&c->d_ftab_adj, (size_t)c->nfault))) return rc;
    std::vector<FaultDev> t(c->flt, c->flt + c->nfault);
    for (FaultDev &F : t) {
      const long fp = scen * c->bat.fpf, fi = scen * c->bat.fpi;
      F.T0 += fp; F.fric += fp; F.out += fp; F.T1 += fp; F.hT1 += fp; F.mT1 += fp;
      F.flag_rup += fi; F.init_t0_flag += fi;
    }
    HIPCHK(hipMemcpy(c->d_ftab_adj, t.data(), t.size() * sizeof(FaultDev), hipMemcpyHostToDevice));
    fa.ftab = c->d_ftab_adj;
  }
  fa.nfault = c->nfault;
  fa.adjf = c->adjf;
  fa.tape = c->tape ? c->tape + (size_t)scen * c->nfault * c->tape_nt * kTapeFloats * ss : nullptr;
  fa.tape_nt = c->tape_nt;
  fa.isfree = c->isfree; fa.imethod = c->imethod; fa.fix_q1 = c->fix_q1;
  fa.dt = c->dt;
  fa.le = 0;
  fa.lp = 3;
  float *res = nullptr, *resf = nullptr;
  const size_t nres = (size_t)c->st_n * nt * 9, nresf = residual_fault ? (size_t)c->fr_n * nt * 9 : 0;
  HIPCHK(hipMalloc(&res, (nres + nresf) * 4));
  hipError_t e = hipMemcpyAsync(res, residual, nres * 4, hipMemcpyHostToDevice, c->st);
  if (residual_fault && e == hipSuccess) {
    resf = res + nres;
    e = hipMemcpyAsync(resf, residual_fault, nresf * 4, hipMemcpyHostToDevice, c->st);
  }
  float *E = c->adj, *A = c->adj + lv, *B = c->adj + 2 * lv, *P = c->adj + 3 * lv;
  for (int n = nt - 1; n >= 0 && e == hipSuccess; n--) {
    if (n < nt - 1) {
      const int it = c->st_it0 + n;  /* the forward step that took w_n to w_{n+1} */
      const int ipair = it % 8;
      fa.n = n;
      fa.wE = E;
      launch_adjf_receivers(fa, c->fr_idx, c->fr_fac, c->fr_nterm, c->fr_n, resf, nt, n + 1, c->st);
      launch_adjf_begin(fa, c->st);
      for (int s = 3; s >= 0; s--) {
        const int idir = (kPairs[ipair][0] + s) % 2, jdir = (kPairs[ipair][1] + s) % 2, kdir = (kPairs[ipair][2] + s) % 2;
        const float cb = kRkB[s] * c->dt, ca = s < 3 ? kRkA[s] * c->dt : 0.0f;
        float *X = s == 3 ? nullptr : (s == 1 ? B : A);
        float *T = (s == 3 || s == 1) ? A : (s == 2 ? B : P);  /* the level the stage writes L^T into */
        fa.istage = s; fa.idir = idir; fa.jdir = jdir; fa.kdir = kdir;
        fa.ca = ca; fa.cb = cb;
        fa.lx = s == 3 ? -1 : (s == 1 ? 2 : 1);
        fa.lc = s == 0 ? fa.lp : (s == 2 ? 2 : 1);
        fa.wT = T; fa.wX = X; fa.wE = E;
        launch_adjf_stage_pre(fa, c->st);
        launch_adj_prep(s, g, c->lam, c->mu, c->slw, X, ca, E, cb, c->adj_sg, pm, c->isfree, kdir, c->st);
        launch_adj_stage(s, idir, jdir, kdir, g, c->met, c->adj_sg, E, A, B, P, pm, c->vcol_grid && c->vcol_on, c->st);
        if (c->isfree)
          launch_adj_top(s, idir, jdir, kdir, g, c->met, c->lam, c->mu, c->slw, c->matVx2Vz, c->matVy2Vz, X, ca, E,
                         cb, T, pm, c->st);
        launch_adjf_stage_post(fa, c->st);
      }
    }
    launch_adj_inject(c->st_term, c->st_fac, c->st_nterm, c->st_n, res, nt, n, sc, P, c->st);
    std::swap(E, P);
    std::swap(pm.le, pm.lp);
    std::swap(fa.le, fa.lp);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  (void)hipFree(res);
  if (e != hipSuccess) return cgfd_set_error(CGFD_ERR_HIP, "adjoint sweep: %s", hipGetErrorString(e));
  /* E holds the adjoint of the first recorded state */
  const size_t hc = (size_t)g.nx * g.ny * g.nz;
  if (grad)
    for (int f = 0; f < 9; f++)
      if ((rc = dev_to_host(c, grad + f * hc, E + f * sc, (long)g.ny * g.nz))) return rc;
  if (grad_fault)
    for (int id = 0; id < c->nfault; id++)
      for (int q = 0; q < 7; q++) {
        const float *src = c->adjf + ((size_t)id * kAdjFaultFloats + 100 + q) * ss + 3 + 3 * (size_t)g.ny;
        float *dst = grad_fault + ((size_t)id * 7 + q) * g.nj * g.nk;
        HIPCHK(hipMemcpy2D(dst, (size_t)g.nj * 4, src, (size_t)g.ny * 4, (size_t)g.nj * 4, g.nk, hipMemcpyDeviceToHost));
      }
  return CGFD_OK;
}
}  // namespace

extern "C" int cgfd_adj_backward(cgfd_core *c, const float *residual, int nt, float *grad) {
  if (!grad) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  return adj_sweep(c, 0, residual, nullptr, nt, grad, nullptr);
}

extern "C" int cgfd_adj_backward_fault(cgfd_core *c, const float *residual, const float *residual_fault, int nt,
                                       float *grad, float *grad_fault) {
  if (!grad_fault) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  return adj_sweep(c, 0, residual, residual_fault, nt, grad, grad_fault);
}

extern "C" int cgfd_batch_adj_backward(cgfd_core *c, int s, const float *residual, int nt, float *grad,
                                       float *grad_fault) {
  if (!grad && !grad_fault) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  return adj_sweep(c, s, residual, nullptr, nt, grad, grad_fault);
}

/* io_fault_recv_keep (io_funcs.cu:2313-2351): a receiver between nodes is
   the bilinear sum buff0 Ly1 Lz1 + buff1 Ly2 Lz1 + buff2 Ly1 Lz2 +
   buff3 Ly2 Lz2 over (j, k), (j+1, k), (j, k+1), (j+1, k+1), with Ly2 = dj,
   Ly1 = 1 - dj; djk = NULL: on the nodes themselves */
static int set_fault_receivers(cgfd_core *c, int n, const int *fjk, const float *djk, int nt_max) {
  if (!c || n < 0 || (n > 0 && (!fjk || nt_max <= 0))) return cgfd_set_error(CGFD_ERR_ARG, "bad receiver arguments");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  const int nterm = djk ? 4 : 1;
  std::vector<int> idx(2 * (size_t)n * nterm);
  std::vector<float> fac(2 * (size_t)n * nterm);
  for (int r = 0; r < n; r++) {
    const int id = fjk[3 * r], j = fjk[3 * r + 1], k = fjk[3 * r + 2];
    if (id < 0 || id >= c->nfault || j < 3 || j > c->g.nj + 2 || k < 3 || k > c->g.nk + 2)
      return cgfd_set_error(CGFD_ERR_ARG, "fault receiver %d (fault %d, j=%d, k=%d) is outside the fault planes", r, id,
                            j, k);
    const float ly2 = djk ? djk[2 * r] : 0.0f, lz2 = djk ? djk[2 * r + 1] : 0.0f;
    const float ly1 = (float)(1.0 - ly2), lz1 = (float)(1.0 - lz2);
    const int dj[4] = {0, 1, 0, 1}, dk[4] = {0, 0, 1, 1};
    const float fy[4] = {ly1, ly2, ly1, ly2}, fz[4] = {lz1, lz1, lz2, lz2};
    for (int u = 0; u < nterm; u++) {
      const size_t q = (size_t)r * nterm + u;
      idx[2 * q] = id;
      idx[2 * q + 1] = (j + dj[u]) + (k + dk[u]) * c->g.ny;
      fac[2 * q] = nterm == 1 ? 1.0f : fy[u];
      fac[2 * q + 1] = nterm == 1 ? 1.0f : fz[u];
    }
  }
  int rc;
  if (n > 0) {
    if ((rc = c->grow(&c->fr_idx, &c->fr_idx_cap, idx.size()))) return rc;
    HIPCHK(hipMemcpy(c->fr_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
    if ((rc = c->grow(&c->fr_fac, &c->fr_fac_cap, fac.size()))) return rc;
    HIPCHK(hipMemcpy(c->fr_fac, fac.data(), fac.size() * 4, hipMemcpyHostToDevice));
    /* one [n][nt_max][9] block per scenario */
    const size_t nbuf = (size_t)c->nscen * n * nt_max * 9;
    if ((rc = c->grow(&c->fr_buf, &c->fr_buf_cap, nbuf))) return rc;
    HIPCHK(hipMemset(c->fr_buf, 0, nbuf * 4));
  }
  c->fr_n = n;
  c->fr_nt = nt_max;
  c->fr_nterm = nterm;
  c->fr_it0 = c->it_host;
  c->drop_graph();
  return CGFD_OK;
}

extern "C" int cgfd_core_set_fault_receivers(cgfd_core *c, int n, const int *fjk, int nt_max) {
  return set_fault_receivers(c, n, fjk, nullptr, nt_max);
}

extern "C" int cgfd_core_set_fault_receivers_interp(cgfd_core *c, int n, const int *fjk, const float *djk,
                                                    int nt_max) {
  if (n > 0 && !djk) return cgfd_set_error(CGFD_ERR_ARG, "null increments");
  return set_fault_receivers(c, n, fjk, djk, nt_max);
}

static int check_scen(const cgfd_core *c, int s);

extern "C" int cgfd_batch_read_fault_receivers(cgfd_core *c, int s, float *dst, int *nt) {
  if (int rc = check_scen(c, s)) return rc;
  if (!dst) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  int rec = scen_steps(c, s) - c->fr_it0;
  if (rec > c->fr_nt) rec = c->fr_nt;
  if (rec < 0) rec = 0;
  if (nt) *nt = rec;
  if (c->fr_n == 0) return CGFD_OK;
  std::vector<float> h((size_t)c->fr_n * c->fr_nt * 9);
  HIPCHK(hipMemcpy(h.data(), c->fr_buf + (size_t)s * c->fr_n * c->fr_nt * 9, h.size() * 4, hipMemcpyDeviceToHost));
  for (int q = 0; q < c->fr_n; q++)
    std::memcpy(dst + (size_t)q * rec * 9, &h[(size_t)q * c->fr_nt * 9], (size_t)rec * 9 * 4);
  return CGFD_OK;
}

extern "C" int cgfd_core_read_fault_receivers(cgfd_core *c, float *dst, int *nt) {
  return cgfd_batch_read_fault_receivers(c, 0, dst, nt);
}

extern "C" int cgfd_core_set_overlap(cgfd_core *c, int on) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  c->drop_graph();
  c->overlap = on ? 1 : 0;
  return CGFD_OK;
}

extern "C" int cgfd_core_set_graph(cgfd_core *c, int on) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  c->use_graph = on ? 1 : 0;
  if (!on) c->drop_graph();
  return CGFD_OK;
}

extern "C" int cgfd_core_run_timed(cgfd_core *c, int nsteps, float *ms_total, float *ms_volume, int *n_volume) {
  if (!c || nsteps < 0) return cgfd_set_error(CGFD_ERR_ARG, "bad arguments");
  HIPCHK(hipSetDevice(c->dev));
  int nmax = 4 * nsteps;
  std::vector<hipEvent_t> ev(2 * nmax + 2);
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  int nev = 0, rc = CGFD_OK;
  HIPCHK(hipEventRecord(ev[2 * nmax], c->st));
  for (int s = 0; s < nsteps && rc == CGFD_OK; s++) rc = do_step(c, ev.data(), &nev);
  HIPCHK(hipEventRecord(ev[2 * nmax + 1], c->st));
  HIPCHK(hipEventSynchronize(ev[2 * nmax + 1]));
  float tot = 0, vol = 0;
  HIPCHK(hipEventElapsedTime(&tot, ev[2 * nmax], ev[2 * nmax + 1]));
  for (int q = 0; q < nev; q++) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, ev[2 * q], ev[2 * q + 1]));
    vol += t;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (ms_total) *ms_total = tot;
  if (ms_volume) *ms_volume = vol;
  if (n_volume) *n_volume = nev;
  return rc;
}

extern "C" int cgfd_core_step_index(const cgfd_core *c) { return c ? c->it_host : -1; }

extern "C" int cgfd_core_sync(cgfd_core *c) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipGetLastError());
  return CGFD_OK;
}

/* steps scenario s has taken */
static int scen_steps(const cgfd_core *c, int s) {
  return c->nscen == 1 ? c->it_host : std::min(c->it_host, c->scen_nt[s]);
}

/* The levels holding a scenario's current state: "pre" and "end" swap every
   step, so a scenario that retired an odd number of steps ago has its final
   state in today's "end" level. */
static int scen_wlev(const cgfd_core *c, int s) {
  return ((c->it_host - scen_steps(c, s)) % 2 == 0) ? c->lv_pre : c->lv_a;
}
static int scen_flev(const cgfd_core *c, int s) {
  return ((c->it_host - scen_steps(c, s)) % 2 == 0) ? c->fl_pre : c->fl_end;
}

static int check_scen(const cgfd_core *c, int s) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  if (s < 0 || s >= c->nscen) return cgfd_set_error(CGFD_ERR_ARG, "scenario %d out of range [0,%d)", s, c->nscen);
  return CGFD_OK;
}

extern "C" int cgfd_batch_read_wavefield(cgfd_core *c, int s, float *dst) {
  if (int rc = check_scen(c, s)) return rc;
  if (!dst) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  if (int rc = dev_to_host(c, dst, c->wlev(scen_wlev(c, s)) + s * c->bat.wstride, 9L * c->g.ny * c->g.nz)) return rc;
  return CGFD_OK;
}
extern "C" int cgfd_core_read_wavefield(cgfd_core *c, float *dst) { return cgfd_batch_read_wavefield(c, 0, dst); }

extern "C" int cgfd_batch_write_wavefield(cgfd_core *c, int s, const float *src) {
  if (int rc = check_scen(c, s)) return rc;
  if (!src) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  const Grid &g = c->g;
  const size_t hz = (size_t)g.nx * g.ny, hc = hz * g.nz;  /* host strides */
  std::vector<float> h((size_t)9 * hc, 0.0f);
  for (int cmp = 0; cmp < 9; cmp++)
    for (int k = 3; k < 3 + g.nk; k++)
      for (int j = 3; j < 3 + g.nj; j++) {
        size_t o = (size_t)cmp * hc + (size_t)k * hz + (size_t)j * g.nx;
        std::memcpy(&h[o + 3], &src[o + 3], (size_t)g.ni * 4);
      }
  if (int rc = host_to_dev(c, c->wlev(scen_wlev(c, s)) + s * c->bat.wstride, h.data(), 9)) return rc;
  return CGFD_OK;
}
extern "C" int cgfd_core_write_wavefield(cgfd_core *c, const float *src) { return cgfd_batch_write_wavefield(c, 0, src); }

extern "C" int cgfd_batch_read_fault(cgfd_core *c, int s, int id, float *dst) {
  if (int rc = check_scen(c, s)) return rc;
  if (!dst) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  if (id < 0 || id >= c->nfault) return cgfd_set_error(CGFD_ERR_ARG, "fault id %d out of range", id);
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(dst, c->flt[id].out + s * c->bat.fpf, (size_t)11 * c->g.sslice * 4, hipMemcpyDeviceToHost));
  return CGFD_OK;
}
extern "C" int cgfd_core_read_fault(cgfd_core *c, int id, float *dst) { return cgfd_batch_read_fault(c, 0, id, dst); }

extern "C" int cgfd_batch_read_fault_state(cgfd_core *c, int s, int id, float *dst) {
  if (int rc = check_scen(c, s)) return rc;
  if (!dst) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  if (id < 0 || id >= c->nfault) return cgfd_set_error(CGFD_ERR_ARG, "fault id %d out of range", id);
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  const float *src = c->flev(scen_flev(c, s)) + s * c->bat.fwstride + (long)id * 18 * c->g.sslice;
  HIPCHK(hipMemcpy(dst, src, (size_t)18 * c->g.sslice * 4, hipMemcpyDeviceToHost));
  return CGFD_OK;
}
extern "C" int cgfd_core_read_fault_state(cgfd_core *c, int id, float *dst) {
  return cgfd_batch_read_fault_state(c, 0, id, dst);
}

extern "C" int cgfd_batch_write_fault_state(cgfd_core *c, int s, int id, const float *src, const float *t1rec) {
  if (int rc = check_scen(c, s)) return rc;
  if (!src) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  if (id < 0 || id >= c->nfault) return cgfd_set_error(CGFD_ERR_ARG, "fault id %d out of range", id);
  const size_t ss = c->g.sslice;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  float *dst = c->flev(scen_flev(c, s)) + s * c->bat.fwstride + (long)id * 18 * ss;
  HIPCHK(hipMemcpy(dst, src, 18 * ss * 4, hipMemcpyHostToDevice));
  if (t1rec) {
    HIPCHK(hipMemcpy(c->flt[id].hT1 + s * c->bat.fpf, t1rec, 3 * ss * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->flt[id].mT1 + s * c->bat.fpf, t1rec + 3 * ss, 3 * ss * 4, hipMemcpyHostToDevice));
  }
  return CGFD_OK;
}
extern "C" int cgfd_core_write_fault_state(cgfd_core *c, int id, const float *src, const float *t1rec) {
  return cgfd_batch_write_fault_state(c, 0, id, src, t1rec);
}

extern "C" int cgfd_batch_read_pg(cgfd_core *c, int s, float *dst) {
  if (int rc = check_scen(c, s)) return rc;
  if (!dst) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  size_t n = (size_t)15 * c->g.nx * c->g.ny;
  if (!c->isfree) {
    std::memset(dst, 0, n * 4);
    return CGFD_OK;
  }
  if (int rc = dev_to_host(c, dst, c->PG + s * c->bat.pgstride, 15L * c->g.ny)) return rc;
  return CGFD_OK;
}
extern "C" int cgfd_core_read_pg(cgfd_core *c, float *dst) { return cgfd_batch_read_pg(c, 0, dst); }

/* io_recv_keep (io_funcs.cu:427-469): a station between nodes is the sum
   of eight terms buff[t] Lx Ly Lz over the cell's corners t = i + 2 j + 4 k,
   Lx2 = di, Lx1 = 1 - di, summed in corner order, except that the reference
   takes corner 5 for the term of corner 4 (io_funcs.cu:459-460), which
   CGFD_COMPAT_FIX_STATION_INTERP corrects; dijk = NULL: on the nodes */
static int set_stations(cgfd_core *c, int n, const int *ijk, const float *dijk, int nt_max) {
  if (!c || n < 0 || (n > 0 && (!ijk || nt_max <= 0))) return cgfd_set_error(CGFD_ERR_ARG, "bad station arguments");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  const int nterm = dijk ? 8 : 1;
  std::vector<long> idx(n), term((size_t)n * nterm);
  std::vector<float> fac((size_t)n * nterm * 3);
  for (int s = 0; s < n; s++) {
    int i = ijk[3 * s], j = ijk[3 * s + 1], k = ijk[3 * s + 2];
    if (i < 3 || i > c->g.ni + 2 || j < 3 || j > c->g.nj + 2 || k < 3 || k > c->g.nk + 2)
      return cgfd_set_error(CGFD_ERR_ARG, "station %d at (%d,%d,%d) is outside the physical grid", s, i, j, k);
    idx[s] = i + (long)j * c->g.siy + (long)k * c->g.siz;
    if (nterm == 1) {
      term[s] = idx[s];
      fac[3 * s] = fac[3 * s + 1] = fac[3 * s + 2] = 1.0f;
      continue;
    }
    const float lx2 = dijk[3 * s], ly2 = dijk[3 * s + 1], lz2 = dijk[3 * s + 2];
    const float lx[2] = {(float)(1.0 - lx2), lx2}, ly[2] = {(float)(1.0 - ly2), ly2}, lz[2] = {(float)(1.0 - lz2), lz2};
    for (int u = 0; u < 8; u++) {
      const int src = (u == 4 && !c->fix_sta) ? 5 : u;  /* the corner the term reads */
      const int a = src & 1, b = (src >> 1) & 1, e = src >> 2;
      term[(size_t)s * 8 + u] = idx[s] + a + (long)b * c->g.siy + (long)e * c->g.siz;
      fac[((size_t)s * 8 + u) * 3] = lx[u & 1];
      fac[((size_t)s * 8 + u) * 3 + 1] = ly[(u >> 1) & 1];
      fac[((size_t)s * 8 + u) * 3 + 2] = lz[u >> 2];
    }
  }
  int rc;
  c->bat.ststride = (long)n * nt_max * 9;
  if (n > 0) {
    if ((rc = c->grow(&c->st_idx, &c->st_idx_cap, (size_t)n))) return rc;
    HIPCHK(hipMemcpy(c->st_idx, idx.data(), (size_t)n * sizeof(long), hipMemcpyHostToDevice));
    if ((rc = c->grow(&c->st_term, &c->st_term_cap, term.size()))) return rc;
    HIPCHK(hipMemcpy(c->st_term, term.data(), term.size() * sizeof(long), hipMemcpyHostToDevice));
    if ((rc = c->grow(&c->st_fac, &c->st_fac_cap, fac.size()))) return rc;
    HIPCHK(hipMemcpy(c->st_fac, fac.data(), fac.size() * 4, hipMemcpyHostToDevice));
    if ((rc = c->grow(&c->st_buf, &c->st_buf_cap, (size_t)c->nscen * n * nt_max * 9))) return rc;
    HIPCHK(hipMemset(c->st_buf, 0, (size_t)c->nscen * n * nt_max * 9 * 4));
  }
  c->st_n = n;
  c->st_nt = nt_max;
  c->st_nterm = nterm;
  c->st_it0 = c->it_host;
  c->drop_graph();
  return CGFD_OK;
}

extern "C" int cgfd_core_set_stations(cgfd_core *c, int n, const int *ijk, int nt_max) {
  return set_stations(c, n, ijk, nullptr, nt_max);
}

extern "C" int cgfd_core_set_stations_interp(cgfd_core *c, int n, const int *ijk, const float *dijk, int nt_max) {
  if (n > 0 && !dijk) return cgfd_set_error(CGFD_ERR_ARG, "null increments");
  return set_stations(c, n, ijk, dijk, nt_max);
}

extern "C" int cgfd_batch_read_stations(cgfd_core *c, int s, float *dst, int *nt) {
  if (int rc = check_scen(c, s)) return rc;
  if (!dst) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  int rec = scen_steps(c, s) - c->st_it0;
  if (rec > c->st_nt) rec = c->st_nt;
  if (rec < 0) rec = 0;
  if (nt) *nt = rec;
  if (c->st_n == 0) return CGFD_OK;
  std::vector<float> h((size_t)c->st_n * c->st_nt * 9);
  HIPCHK(hipMemcpy(h.data(), c->st_buf + s * c->bat.ststride, h.size() * 4, hipMemcpyDeviceToHost));
  for (int q = 0; q < c->st_n; q++)
    std::memcpy(dst + (size_t)q * rec * 9, &h[(size_t)q * c->st_nt * 9], (size_t)rec * 9 * 4);
  return CGFD_OK;
}
/* validated snapshot box and its size in floats */
static int snapshot_box(const cgfd_core *c, const int *box, int flags, IoBox &b, size_t &n) {
  if (!box || flags <= 0 || flags > 7) return cgfd_set_error(CGFD_ERR_ARG, "bad snapshot arguments");
  const Grid &g = c->g;
  const int lo[3] = {3, 3, 3}, hi[3] = {g.ni + 2, g.nj + 2, g.nk + 2};
  for (int a = 0; a < 3; a++) {
    b.s[a] = box[3 * a];
    b.n[a] = box[3 * a + 1];
    b.d[a] = box[3 * a + 2];
    if (b.n[a] < 1 || b.d[a] < 1 || b.s[a] < lo[a] || b.s[a] + (long)(b.n[a] - 1) * b.d[a] > hi[a])
      return cgfd_set_error(CGFD_ERR_ARG, "snapshot box axis %d (start %d, count %d, stride %d) leaves the physical grid",
                            a, b.s[a], b.n[a], b.d[a]);
  }
  const int groups = (flags & 1 ? 3 : 0) + (flags & 2 ? 6 : 0) + (flags & 4 ? 6 : 0);
  n = (size_t)b.n[0] * b.n[1] * b.n[2] * groups;
  return CGFD_OK;
}

/* pack the box of scenario s into the device buffer buf and copy it to dst */
static hipError_t snapshot_copy(cgfd_core *c, int s, const IoBox &b, int flags, size_t n, float *buf, float *dst) {
  launch_pack_box(c->wlev(scen_wlev(c, s)) + s * c->bat.wstride, c->g, c->lam, c->mu, b, flags, buf, c->st);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(dst, buf, n * 4, hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  return e;
}

extern "C" int cgfd_batch_read_snapshot(cgfd_core *c, int s, const int *box, int flags, float *dst) {
  if (int rc = check_scen(c, s)) return rc;
  if (!dst) return cgfd_set_error(CGFD_ERR_ARG, "bad snapshot arguments");
  IoBox b;
  size_t n = 0;
  if (int rc = snapshot_box(c, box, flags, b, n)) return rc;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  float *buf = nullptr;
  HIPCHK(hipMalloc(&buf, n * 4));
  hipError_t e = snapshot_copy(c, s, b, flags, n, buf, dst);
  (void)hipFree(buf);
  if (e != hipSuccess) return cgfd_set_error(CGFD_ERR_HIP, "snapshot: %s", hipGetErrorString(e));
  return CGFD_OK;
}

/* ---- the allocation-free snapshot reader ----
   The packing buffer is this file-scope device array. Every GPU the library
   runs on holds its own copy (32 MiB) until the process ends; its address on
   a device comes from hipGetSymbolAddress with that device current (at core
   creation), and host code only ever uses that address. All calls on one
   GPU share the array, so calls that overlap on one GPU race; nothing in it
   needs clearing (a call writes every value it copies out). */
constexpr size_t kScratchFloats = (size_t)8 << 20;
static __device__ __align__(256) float scratch[kScratchFloats];
static float *g_scratch[64] = {};
static std::mutex g_scratch_mu;

static int scratch_setup(int dev) {
  if (dev < 0 || dev >= 64) return cgfd_set_error(CGFD_ERR_ARG, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return CGFD_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return cgfd_set_error(CGFD_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch[dev] = (float *)p;
  return CGFD_OK;
}

extern "C" int cgfd_core_read_snapshot_noalloc(cgfd_core *c, const int *box, int flags, float *dst) {
  if (int rc = check_scen(c, 0)) return rc;
  if (!dst) return cgfd_set_error(CGFD_ERR_ARG, "bad snapshot arguments");
  IoBox b;
  size_t n = 0;
  if (int rc = snapshot_box(c, box, flags, b, n)) return rc;
  if (n > kScratchFloats)
    return cgfd_set_error(CGFD_ERR_ARG, "snapshot of %zu values exceeds the allocation-free maximum of %zu (32 MiB)", n,
                          kScratchFloats);
  HIPCHK(hipSetDevice(c->dev));
  float *buf = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    buf = c->dev >= 0 && c->dev < 64 ? g_scratch[c->dev] : nullptr;
  }
  if (!buf) return cgfd_set_error(CGFD_ERR_STATE, "scratch of device %d not set up", c->dev);
  HIPCHK(hipStreamSynchronize(c->st));
  const hipError_t e = snapshot_copy(c, 0, b, flags, n, buf, dst);
  if (e != hipSuccess) return cgfd_set_error(CGFD_ERR_HIP, "snapshot: %s", hipGetErrorString(e));
  return CGFD_OK;
}

extern "C" int cgfd_core_read_snapshot(cgfd_core *c, const int *box, int flags, float *dst) {
  return cgfd_batch_read_snapshot(c, 0, box, flags, dst);
}

extern "C" int cgfd_core_read_slice(cgfd_core *c, int axis, int index, float *dst) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  const Grid &g = c->g;
  int box[9] = {3, g.ni, 1, 3, g.nj, 1, 3, g.nk, 1};
  if (axis < 0 || axis > 2) return cgfd_set_error(CGFD_ERR_ARG, "slice axis %d (0 = x, 1 = y, 2 = z)", axis);
  box[3 * axis] = index;
  box[3 * axis + 1] = 1;
  return cgfd_batch_read_snapshot(c, 0, box, 3, dst);
}

extern "C" int cgfd_core_read_stations(cgfd_core *c, float *dst, int *nt) {
  return cgfd_batch_read_stations(c, 0, dst, nt);
}

extern "C" int cgfd_batch_steps(const cgfd_core *c, int s) {
  if (check_scen(c, s)) return -1;
  return scen_steps(c, s);
}

extern "C" int cgfd_batch_active(cgfd_core *c) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  int n = 0;
  for (int s = 0; s < c->nscen; s++) n += c->active(s) ? 1 : 0;
  return n;
}

/* ---- autotuned schedules, per problem shape ---- */
namespace {
struct TuneKey {
  int ni, nj, nk, nscen, nfault, isfree, pml;
  bool operator<(const TuneKey &o) const {
    const int a[7] = {ni, nj, nk, nscen, nfault, isfree, pml}, b[7] = {o.ni, o.nj, o.nk, o.nscen, o.nfault, o.isfree, o.pml};
    return std::lexicographical_compare(a, a + 7, b, b + 7);
  }
};
std::map<TuneKey, std::pair<int, int>> g_tuned;  /* shape -> (variant, zchunk) */
bool g_tune_file_read = false;
std::mutex g_tune_mu;

TuneKey tune_key(const cgfd_core *c) {
  int pml = 0;
  for (int d = 0; d < 3; d++)
    for (int e = 0; e < 2; e++) pml |= (c->pml[d][e].on ? 1 : 0) << (2 * d + e);
  return TuneKey{c->g.ni, c->g.nj, c->g.nk, c->nscen, c->nfault, c->isfree, pml};
}

bool valid_variant(int v) { return v == 1 || v == 20 || v == 23 || v == 24 || v == 32 || v == 33; }

/* the cache file: one "ni nj nk nscen nfault isfree pml variant zchunk" line
   per tuned shape; later lines win */
void read_tune_file() {
  if (g_tune_file_read) return;
  g_tune_file_read = true;
  const char *path = std::getenv("CGFD_TUNE_CACHE");
  if (!path || !*path) return;
  FILE *f = std::fopen(path, "r");
  if (!f) return;
  TuneKey k;
  int v, z;
  while (std::fscanf(f, "%d %d %d %d %d %d %d %d %d", &k.ni, &k.nj, &k.nk, &k.nscen, &k.nfault, &k.isfree, &k.pml,
                     &v, &z) == 9)
    if (valid_variant(v) && z >= 4 && z <= 4096) g_tuned[k] = {v, z};
  std::fclose(f);
}

void record_tuned(const TuneKey &k, int v, int z) {
  g_tuned[k] = {v, z};
  const char *path = std::getenv("CGFD_TUNE_CACHE");
  if (!path || !*path) return;
  FILE *f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%d %d %d %d %d %d %d %d %d\n", k.ni, k.nj, k.nk, k.nscen, k.nfault, k.isfree, k.pml, v, z);
  std::fclose(f);
}
}  // namespace

/* variant 0: the tuned choice for this shape, else the box split with 8
   y-tiles per slab: 33 (the bands in one launch) on one scenario, 32 (the
   interior box, too, in per-scenario sweeps) on a batch */
static void apply_auto_variant(cgfd_core *c) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  read_tune_file();
  auto it = g_tuned.find(tune_key(c));
  if (it != g_tuned.end()) {
    c->variant = it->second.first;
    c->zchunk = it->second.second;
  } else {
    c->variant = c->nscen > 1 ? 32 : 33;
    c->zchunk = 8;
  }
}

extern "C" int cgfd_core_set_variant(cgfd_core *c, int variant) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  if (variant != 0 && !valid_variant(variant))
    return cgfd_set_error(CGFD_ERR_ARG, "unknown volume kernel variant %d (valid: 0, 1, 20, 23, 24, 32, 33)", variant);
  if (variant == 0) apply_auto_variant(c);
  else c->variant = variant;
  c->drop_graph();
  return CGFD_OK;
}

extern "C" int cgfd_core_get_variant(const cgfd_core *c, int *variant, int *zchunk) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  if (variant) *variant = c->variant;
  if (zchunk) *zchunk = c->zchunk;
  return CGFD_OK;
}

extern "C" int cgfd_core_set_vertical_columns(cgfd_core *c, int on) {
  if (!c || (on != 0 && on != 1)) return cgfd_set_error(CGFD_ERR_ARG, "cgfd_core_set_vertical_columns: bad arguments");
  c->vcol_on = on;
  c->drop_graph();
  return CGFD_OK;
}

extern "C" int cgfd_core_get_vertical_columns(const cgfd_core *c, int *grid, int *in_use) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  if (grid) *grid = c->vcol_grid;
  if (in_use)
    *in_use = c->vcol_grid && c->vcol_on && (c->variant == 23 || c->variant == 24 || c->variant == 32 || c->variant == 33);
  return CGFD_OK;
}

extern "C" int cgfd_core_set_zchunk(cgfd_core *c, int zchunk) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  if (zchunk < 4 || zchunk > 4096) return cgfd_set_error(CGFD_ERR_ARG, "zchunk %d out of range [4, 4096]", zchunk);
  c->zchunk = zchunk;
  c->drop_graph();
  return CGFD_OK;
}

extern "C" int cgfd_core_autotune(cgfd_core *c, int reps, int *variant_out, int *zchunk_out, float *ms_out) {
  if (!c || reps < 1) return cgfd_set_error(CGFD_ERR_ARG, "bad arguments");
  HIPCHK(hipSetDevice(c->dev));
  c->drop_graph();
  if (c->nvol <= 0 || c->refresh_active() == 0) {
    if (variant_out) *variant_out = c->variant;
    if (zchunk_out) *zchunk_out = c->zchunk;
    if (ms_out) *ms_out = 0.0f;
    return CGFD_OK;
  }
  const int ipair = c->it_host % 8;
  const int idir = kPairs[ipair][0], jdir = kPairs[ipair][1], kdir = kPairs[ipair][2];
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  /* 32 differs from 23 on batches only */
  const std::vector<int> variants = c->bat.nact > 1 ? std::vector<int>{23, 32, 33, 24, 20, 1} : std::vector<int>{23, 33, 24, 20, 1};
  const int chunks[4] = {4, 8, 16, 32};
  float best = 1e30f;
  int bv = c->variant, bz = c->zchunk;
  for (int v : variants)
    for (int zi = 0; zi < 4; zi++) {
      const int z = chunks[zi];
      if (v == 1 && zi > 0) continue;  /* no slabs */
      c->variant = v;
      c->zchunk = z;
      /* the four stages' volume launches of one step: they read level pre
         and write only levels A, B and C (and the PML auxiliaries' la, lb, lc), which
         the next real step rewrites before reading */
      VolArgs va[4];
      const int lc[4] = {c->lv_pre, c->lv_a, c->lv_b, c->lv_c};
      const int ln[4] = {c->lv_a, c->lv_b, c->lv_c, -1};
      for (int s = 0; s < 4; s++) {
        fill_vol(c, va[s], s, c->wlev(lc[s]), ln[s] >= 0 ? c->wlev(ln[s]) : nullptr);
        va[s].k_lo = 0;
        va[s].nk_run = c->nvol;
      }
      auto one = [&]() {
        for (int s = 0; s < 4; s++) {
          (void)c->fork();
          launch_vol(c, va[s], (idir + s) % 2, (jdir + s) % 2, (kdir + s) % 2);
          (void)c->join();
        }
      };
      one();
      HIPCHK(hipEventRecord(e0, c->st));
      for (int r = 0; r < reps; r++) one();
      HIPCHK(hipEventRecord(e1, c->st));
      HIPCHK(hipEventSynchronize(e1));
      HIPCHK(hipGetLastError());
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      ms /= 4 * reps;
      if (std::getenv("CGFD_AUTOTUNE_VERBOSE")) std::fprintf(stderr, "[autotune] variant %d zchunk %d: %.3f ms\n", v, z, ms);
      if (ms < best) { best = ms; bv = v; bz = z; }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->variant = bv;
  c->zchunk = bz;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    read_tune_file();
    record_tuned(tune_key(c), bv, bz);
  }
  if (variant_out) *variant_out = bv;
  if (zchunk_out) *zchunk_out = bz;
  if (ms_out) *ms_out = best;
  return CGFD_OK;
}

extern "C" int cgfd_drv_rk_curv_col_allstep(const cgfd_desc *d, float *wave_out, float *fault_out, float *pg_out) {
  cgfd_core *c = nullptr;
  int rc = cgfd_core_create(d, &c);
  if (rc) return rc;
  rc = cgfd_core_run(c, d->nt_total);
  if (!rc && wave_out) rc = cgfd_core_read_wavefield(c, wave_out);
  for (int id = 0; !rc && fault_out && id < c->nfault; id++)
    rc = cgfd_core_read_fault(c, id, fault_out + (size_t)id * 11 * c->g.sslice);
  if (!rc && pg_out) rc = cgfd_core_read_pg(c, pg_out);
  cgfd_core_destroy(c);
  return rc;
}

/* ---- output files in the reference's NetCDF layouts (io_funcs.cu:900-1131,
   1827-1879, 1952-2139), from device-packed data ---- */
struct cgfd_out {
  cgfd::NcWriter w;
  int kind = 0;             /* 0 fault, 1 slice, 2 snapshot */
  int id = 0, axis = 0, index = 0;
  int box[9] = {}, flags = 0, nvar = 0;
  int vtime = -1, vfirst = -1, vpeak = -1, vinit = -1;
};

namespace {
std::string out_name(const char *dir, const char *stem, const char *frame) {
  std::string p = dir && *dir ? std::string(dir) + "/" : std::string();
  return p + stem + "_" + (frame && *frame ? frame : "px0_py0_pz0") + ".nc";
}
}  // namespace

extern "C" int cgfd_io_fault_create(cgfd_core *c, int id, const char *out_dir, const char *frame, cgfd_out **out) {
  if (!c || !out || id < 0 || id >= c->nfault) return cgfd_set_error(CGFD_ERR_ARG, "bad fault output arguments");
  const Grid &g = c->g;
  const int i0 = c->flt[id].i0;
  char stem[64];
  std::snprintf(stem, sizeof stem, "fault_i%d", i0 - 3 + 1);
  cgfd_out *o = new (std::nothrow) cgfd_out;
  if (!o) return cgfd_set_error(CGFD_ERR_NOMEM, "out of host memory");
  o->kind = 0;
  o->id = id;
  if (int rc = o->w.create(out_name(out_dir, stem, frame))) { delete o; return rc; }
  const int dt = o->w.def_dim("time", 0), dk = o->w.def_dim("k", g.nk), dj = o->w.def_dim("j", g.nj);
  o->vtime = o->w.def_var("time", {dt});
  static const char *names[9] = {"Tn", "Ts1", "Ts2", "Vs", "Vs1", "Vs2", "Slip", "Slip1", "Slip2"};
  for (int q = 0; q < 9; q++) {
    const int v = o->w.def_var(names[q], {dt, dk, dj});
    if (q == 0) o->vfirst = v;
  }
  o->vpeak = o->w.def_var("Peak_vs", {dk, dj});
  o->vinit = o->w.def_var("Init_t0", {dk, dj});
  o->w.att_int("i_index_with_ghosts_in_this_thread", {i0});
  o->w.att_int("coords_of_mpi_topo", {0, 0, 0});
  if (int rc = o->w.enddef()) { delete o; return rc; }
  *out = o;
  return CGFD_OK;
}

namespace {
/* physical [nk][nj] of field q of a ghost-inclusive [..][nz][ny] array */
void phys_plane(const Grid &g, const float *src, float *dst) {
  for (int k = 0; k < g.nk; k++)
    std::memcpy(dst + (size_t)k * g.nj, src + (size_t)(k + 3) * g.ny + 3, (size_t)g.nj * 4);
}
}  // namespace

extern "C" int cgfd_io_fault_put(cgfd_out *o, cgfd_core *c, int it_skip, float time) {
  if (!o || !c || o->kind != 0 || it_skip < 0) return cgfd_set_error(CGFD_ERR_ARG, "bad fault output arguments");
  const Grid &g = c->g;
  std::vector<float> f((size_t)11 * g.sslice), plane((size_t)g.nj * g.nk);
  if (int rc = cgfd_core_read_fault(c, o->id, f.data())) return rc;
  if (int rc = o->w.put_rec(o->vtime, it_skip, &time)) return rc;
  for (int q = 0; q < 9; q++) {
    phys_plane(g, f.data() + (size_t)q * g.sslice, plane.data());
    if (int rc = o->w.put_rec(o->vfirst + q, it_skip, plane.data())) return rc;
  }
  return CGFD_OK;
}

extern "C" int cgfd_io_fault_end(cgfd_out *o, cgfd_core *c) {
  if (!o || !c || o->kind != 0) return cgfd_set_error(CGFD_ERR_ARG, "bad fault output arguments");
  const Grid &g = c->g;
  std::vector<float> f((size_t)11 * g.sslice), plane((size_t)g.nj * g.nk);
  int rc = cgfd_core_read_fault(c, o->id, f.data());
  if (!rc) { phys_plane(g, f.data() + 9 * (size_t)g.sslice, plane.data()); rc = o->w.put_var(o->vpeak, plane.data()); }
  if (!rc) { phys_plane(g, f.data() + 10 * (size_t)g.sslice, plane.data()); rc = o->w.put_var(o->vinit, plane.data()); }
  const int rc2 = o->w.close();
  delete o;
  return rc ? rc : rc2;
}

extern "C" int cgfd_io_slice_create(cgfd_core *c, int axis, int index, const char *out_dir, const char *frame,
                                    cgfd_out **out) {
  if (!c || !out || axis < 0 || axis > 2) return cgfd_set_error(CGFD_ERR_ARG, "bad slice output arguments");
  const Grid &g = c->g;
  const int n[3] = {g.ni, g.nj, g.nk};
  if (index < 3 || index > n[axis] + 2) return cgfd_set_error(CGFD_ERR_ARG, "slice index %d outside the grid", index);
  static const char *ax = "xyz", *ix = "ijk";
  char stem[64];
  std::snprintf(stem, sizeof stem, "slice%c_%c%d", ax[axis], ix[axis], index - 3 + 1);
  cgfd_out *o = new (std::nothrow) cgfd_out;
  if (!o) return cgfd_set_error(CGFD_ERR_NOMEM, "out of host memory");
  o->kind = 1;
  o->axis = axis;
  o->index = index;
  if (int rc = o->w.create(out_name(out_dir, stem, frame))) { delete o; return rc; }
  const int dt = o->w.def_dim("time", 0);
  int d1, d2;
  if (axis == 0) { d1 = o->w.def_dim("k", g.nk); d2 = o->w.def_dim("j", g.nj); }
  else if (axis == 1) { d1 = o->w.def_dim("k", g.nk); d2 = o->w.def_dim("i", g.ni); }
  else { d1 = o->w.def_dim("j", g.nj); d2 = o->w.def_dim("i", g.ni); }
  o->vtime = o->w.def_var("time", {dt});
  static const char *names[9] = {"Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"};
  for (int q = 0; q < 9; q++) {
    const int v = o->w.def_var(names[q], {dt, d1, d2});
    if (q == 0) o->vfirst = v;
  }
  static const char *att[3] = {"i_index_with_ghosts_in_this_thread", "j_index_with_ghosts_in_this_thread",
                               "k_index_with_ghosts_in_this_thread"};
  o->w.att_int(att[axis], {index});
  o->w.att_int("coords_of_mpi_topo", {0, 0, 0});
  if (int rc = o->w.enddef()) { delete o; return rc; }
  *out = o;
  return CGFD_OK;
}

extern "C" int cgfd_io_slice_put(cgfd_out *o, cgfd_core *c, int it_skip, float time) {
  if (!o || !c || o->kind != 1 || it_skip < 0) return cgfd_set_error(CGFD_ERR_ARG, "bad slice output arguments");
  const Grid &g = c->g;
  const size_t plane = o->axis == 0 ? (size_t)g.nk * g.nj : (o->axis == 1 ? (size_t)g.nk * g.ni : (size_t)g.nj * g.ni);
  std::vector<float> buf(9 * plane);
  if (int rc = cgfd_core_read_slice(c, o->axis, o->index, buf.data())) return rc;
  if (int rc = o->w.put_rec(o->vtime, it_skip, &time)) return rc;
  for (int q = 0; q < 9; q++)
    if (int rc = o->w.put_rec(o->vfirst + q, it_skip, buf.data() + q * plane)) return rc;
  return CGFD_OK;
}

extern "C" int cgfd_io_snap_create(cgfd_core *c, const char *name, const int *box, int flags, const char *out_dir,
                                   const char *frame, cgfd_out **out) {
  if (!c || !out || !box || !name || !*name || flags <= 0 || flags > 7)
    return cgfd_set_error(CGFD_ERR_ARG, "bad snapshot output arguments");
  const Grid &g = c->g;
  const int hi[3] = {g.ni + 2, g.nj + 2, g.nk + 2};
  for (int a = 0; a < 3; a++)
    if (box[3 * a + 1] < 1 || box[3 * a + 2] < 1 || box[3 * a] < 3 ||
        box[3 * a] + (long)(box[3 * a + 1] - 1) * box[3 * a + 2] > hi[a])
      return cgfd_set_error(CGFD_ERR_ARG, "snapshot box axis %d leaves the physical grid", a);
  cgfd_out *o = new (std::nothrow) cgfd_out;
  if (!o) return cgfd_set_error(CGFD_ERR_NOMEM, "out of host memory");
  o->kind = 2;
  o->flags = flags;
  for (int q = 0; q < 9; q++) o->box[q] = box[q];
  if (int rc = o->w.create(out_name(out_dir, name, frame))) { delete o; return rc; }
  const int dt = o->w.def_dim("time", 0), dk = o->w.def_dim("k", box[7]), dj = o->w.def_dim("j", box[4]),
            di = o->w.def_dim("i", box[1]);
  o->vtime = o->w.def_var("time", {dt});
  /* the reference's variable order (io_funcs.cu:1329-1366) */
  static const char *vn[3] = {"Vx", "Vy", "Vz"}, *tn[6] = {"Txx", "Tyy", "Tzz", "Txz", "Tyz", "Txy"},
                    *en[6] = {"Exx", "Eyy", "Ezz", "Exz", "Eyz", "Exy"};
  o->nvar = 0;
  if (flags & 1) for (int q = 0; q < 3; q++) { int v = o->w.def_var(vn[q], {dt, dk, dj, di}); if (!o->nvar++) o->vfirst = v; }
  if (flags & 2) for (int q = 0; q < 6; q++) { int v = o->w.def_var(tn[q], {dt, dk, dj, di}); if (!o->nvar++) o->vfirst = v; }
  if (flags & 4) for (int q = 0; q < 6; q++) { int v = o->w.def_var(en[q], {dt, dk, dj, di}); if (!o->nvar++) o->vfirst = v; }
  o->w.att_int("first_index_to_snapshot_output", {box[0] - 3, box[3] - 3, box[6] - 3});
  o->w.att_int("first_index_in_this_thread_with_ghosts", {box[0], box[3], box[6]});
  o->w.att_int("index_stride_in_this_thread", {box[2], box[5], box[8]});
  o->w.att_int("coords_of_mpi_topo", {0, 0, 0});
  if (int rc = o->w.enddef()) { delete o; return rc; }
  *out = o;
  return CGFD_OK;
}

extern "C" int cgfd_io_snap_put(cgfd_out *o, cgfd_core *c, int rec, float time) {
  if (!o || !c || o->kind != 2 || rec < 0) return cgfd_set_error(CGFD_ERR_ARG, "bad snapshot output arguments");
  const size_t n = (size_t)o->box[1] * o->box[4] * o->box[7];
  std::vector<float> buf(n * o->nvar);
  if (int rc = cgfd_core_read_snapshot(c, o->box, o->flags, buf.data())) return rc;
  if (int rc = o->w.put_rec(o->vtime, rec, &time)) return rc;
  /* packed groups hold the wavefield order (..., Tyz, Txz, Txy); the file the
     reference's (..., Txz, Tyz, Txy) */
  static const int perm[6] = {0, 1, 2, 4, 3, 5};
  int v = o->vfirst;
  size_t g0 = 0;
  for (int grp = 0; grp < 3; grp++) {
    if (!(o->flags & (1 << grp))) continue;
    const int nf = grp == 0 ? 3 : 6;
    for (int q = 0; q < nf; q++)
      if (int rc = o->w.put_rec(v++, rec, buf.data() + (g0 + (nf == 6 ? perm[q] : q)) * n)) return rc;
    g0 += nf;
  }
  return CGFD_OK;
}

extern "C" int cgfd_io_close(cgfd_out *o) {
  if (!o) return CGFD_OK;
  const int rc = o->w.close();
  delete o;
  return rc;
}

extern "C" int cgfd_io_pg_write(cgfd_core *c, const char *out_dir, const char *frame) {
  if (!c) return cgfd_set_error(CGFD_ERR_ARG, "null core");
  if (!c->PG) return cgfd_set_error(CGFD_ERR_STATE, "PG output needs a free surface");
  const Grid &g = c->g;
  std::vector<float> pg((size_t)15 * g.nx * g.ny);
  if (int rc = cgfd_core_read_pg(c, pg.data())) return rc;
  cgfd::NcWriter w;
  if (int rc = w.create(out_name(out_dir, "PG_V_A_D", frame))) return rc;
  const int dj = w.def_dim("j", g.ny), di = w.def_dim("i", g.nx);
  static const char *names[15] = {"PGV", "PGVh", "PGVx", "PGVy", "PGVz", "PGA", "PGAh", "PGAx",
                                  "PGAy", "PGAz", "PGD", "PGDh", "PGDx", "PGDy", "PGDz"};
  int v0 = -1;
  for (int q = 0; q < 15; q++) {
    const int v = w.def_var(names[q], {dj, di});
    if (q == 0) v0 = v;
  }
  w.att_int("global_index_of_first_physical_points", {0, 0});
  w.att_int("count_index_of_physical_points", {g.ni, g.nj});
  w.att_int("coords_of_mpi_topo", {0, 0, 0});
  int rc = w.enddef();
  for (int q = 0; q < 15 && !rc; q++) rc = w.put_var(v0 + q, pg.data() + (size_t)q * g.nx * g.ny);
  const int rc2 = w.close();
  return rc ? rc : rc2;
}

/* station traces as SAC files (io_recv_output_sac, io_funcs.cu:577-618):
   <out_dir>/<name>.<Vx..Txy>.sac, all nt_max samples of the station window.
   As the reference's call, the event position is zero and the station's
   depth field holds dt (its arguments stop one position short of stdp). */
extern "C" int cgfd_io_stations_sac(cgfd_core *c, const char *out_dir, const char *const *names, const float *xyz,
                                    int with_strain) {
  if (!c || !names || !xyz) return cgfd_set_error(CGFD_ERR_ARG, "null argument");
  if (c->st_n <= 0) return cgfd_set_error(CGFD_ERR_STATE, "no stations set");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  const int n = c->st_n, nt = c->st_nt;
  std::vector<float> h((size_t)n * nt * 9), tr((size_t)nt);
  HIPCHK(hipMemcpy(h.data(), c->st_buf, h.size() * 4, hipMemcpyDeviceToHost));
  static const char *cmp[9] = {"Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"};
  const float ev[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  const std::string dir = out_dir && *out_dir ? std::string(out_dir) + "/" : std::string();
  std::vector<long> idx((size_t)n);
  if (with_strain) HIPCHK(hipMemcpy(idx.data(), c->st_idx, (size_t)n * sizeof(long), hipMemcpyDeviceToHost));
  for (int s = 0; s < n; s++) {
    if (!names[s]) return cgfd_set_error(CGFD_ERR_ARG, "station %d has no name", s);
    const float st4[4] = {xyz[3 * s], xyz[3 * s + 1], xyz[3 * s + 2], c->dt};
    for (int q = 0; q < 9; q++) {
      for (int t = 0; t < nt; t++) tr[t] = h[((size_t)s * nt + t) * 9 + q];
      if (int rc = cgfd::write_sac(dir + names[s] + "." + cmp[q] + ".sac", tr.data(), nt, c->dt, ev, st4)) return rc;
    }
    if (!with_strain) continue;
    /* strain from the stress with the isotropic compliance at the station
       node (io_recv_output_sac_el_iso_strain, io_funcs.cu:626-700) */
    float lm[2];
    HIPCHK(hipMemcpy(&lm[0], c->lam + idx[s], 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&lm[1], c->mu + idx[s], 4, hipMemcpyDeviceToHost));
    const float lam = lm[0], mu = lm[1];
    const float E1 = (lam + mu) / (mu * (3.0 * lam + 2.0 * mu));
    const float E2 = -lam / (2.0 * mu * (3.0 * lam + 2.0 * mu));
    const float E3 = 1.0 / mu;
    static const char *ecmp[6] = {"Exx", "Eyy", "Ezz", "Eyz", "Exz", "Exy"};
    std::vector<float> e((size_t)6 * nt);
    for (int t = 0; t < nt; t++) {
      const float *w = &h[((size_t)s * nt + t) * 9];
      const float E0 = E2 * (w[3] + w[4] + w[5]);
      e[0 * (size_t)nt + t] = E0 - (E2 - E1) * w[3];
      e[1 * (size_t)nt + t] = E0 - (E2 - E1) * w[4];
      e[2 * (size_t)nt + t] = E0 - (E2 - E1) * w[5];
      e[3 * (size_t)nt + t] = 0.5 * E3 * w[6];
      e[4 * (size_t)nt + t] = 0.5 * E3 * w[7];
      e[5 * (size_t)nt + t] = 0.5 * E3 * w[8];
    }
    for (int q = 0; q < 6; q++)
      if (int rc = cgfd::write_sac(dir + names[s] + "." + ecmp[q] + ".sac", e.data() + (size_t)q * nt, nt, c->dt, ev,
                                   st4))
        return rc;
  }
  return CGFD_OK;
}
