/*
 * cgfd_core.h -- C ABI of the MI355X-native time-stepping core.
 *
 * This is the drop-in boundary for the reference's
 *   int drv_rk_curv_col_allstep(fd_t*, gd_t*, gd_metric_t*, md_t*, par_t*,
 *                               bdryfree_t*, bdrypml_t*, bdryexp_t*, wav_t*,
 *                               mympi_t*, fault_coef_t*, fault_t*, fault_wav_t*,
 *                               iorecv_t*, ioline_t*, iofault_t*, ioslice_t*,
 *                               iosnap_t*, io_fault_recv_t*, float dt,
 *                               int nt_total, float t0, char*, char*)
 *   (src/forward/drv_rk_curv_col.h:18-42, body drv_rk_curv_col.cu:28-667),
 * called once from main (src/forward/main_curv_col_el_3d.cu:711-718).
 *
 * Everything crosses the boundary as plain host pointers in the reference's
 * host layouts:
 *   - volume arrays: nx*ny*nz floats, index i + j*nx + k*nx*ny, nx = ni+6
 *     (3 ghost points per side, gd_t.cu:2151-2153);
 *   - fault-plane arrays: ny*nz values, index j + k*ny (fault_info.cu:166);
 *   - 3x3 matrices: 9 consecutive floats per node, row-major
 *     (fault_info.cu:220-228).
 * The library copies what it needs to the GPU in create() and retains no host
 * pointer afterwards. All functions return CGFD_OK (0) or a negative status;
 * cgfd_last_error() gives the message of the last failure on the calling
 * thread. Nothing in the library calls exit().
 *
 * One core object per host thread; a core owns its device memory and its HIP
 * stream. Objects are not internally thread-safe.
 */
#ifndef CGFD_CORE_H
#define CGFD_CORE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGFD_MAX_FAULTS 5   /* fault_t.fault_one[5], fault_info.h:129 */
#define CGFD_NCMP 9         /* Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy (wav_t.cu:56-106) */
#define CGFD_NGHOST 3       /* fd_t.cu:63-65 */
#define CGFD_NFAULT_OUT 11  /* Tn,Ts1,Ts2,Vs,Vs1,Vs2,Slip,Slip1,Slip2,Peak_vs,Init_t0 */
#define CGFD_NPG 15         /* PGV,PGVh,PGVx,PGVy,PGVz,PGA..,PGD.. (wav_t.cu:160-174) */

/* wavefield component order (wav_t.cu:56-106) */
enum { CGFD_VX = 0, CGFD_VY, CGFD_VZ, CGFD_TXX, CGFD_TYY, CGFD_TZZ,
       CGFD_TYZ, CGFD_TXZ, CGFD_TXY };

/* status codes */
#define CGFD_OK          0
#define CGFD_ERR_ARG    -1   /* invalid argument / inconsistent descriptor */
#define CGFD_ERR_HIP    -2   /* HIP runtime error (message has the HIP string) */
#define CGFD_ERR_NOMEM  -3   /* device or host allocation failed */
#define CGFD_ERR_STATE  -4   /* call not valid in the object's state */
#define CGFD_ERR_NAN    -5   /* non-finite value detected by the NaN check */
#define CGFD_ERR_NODEV  -6   /* no usable GPU */

/*
 * One split-node fault: fault_t.fault_one[id] + fault_coef_t.fault_coef_one[id]
 * as fault_set/fault_coef_cal leave them (fault_info.cu:97-1074). Plane arrays
 * have ny*nz entries (ghosts included; only physical nodes are read).
 */
typedef struct cgfd_fault_desc {
  int fault_x_index;        /* 0-based physical x index (par_t.cu:323) */
  /* fault_one_t inputs (fault_info.h:89-95) */
  const float *T0x, *T0y, *T0z, *mu_s, *mu_d, *Dc, *C0;
  /* masks (fault_info.h:113-116) */
  const int *united, *faultgrid, *rup_index_y, *rup_index_z;
  /* initial values of Tn, Ts1, Ts2 (fault_info.cu:1015-1023); may be NULL -> 0 */
  const float *Tn0, *Ts10, *Ts20;
  /* fault_coef_one_t (fault_info.h:11-76) */
  const float *rho_f;             /* [2][ny*nz], side 0 minus, 1 plus (density) */
  const float *vec_s1, *vec_s2;   /* [ny*nz][3] strike / dip unit vectors */
  /* D21_1,D22_1,D23_1,D31_1,D32_1,D33_1,D21_2,D22_2,D23_2,D31_2,D32_2,D33_2 */
  const float *D[12];             /* each [ny*nz][9] */
  const float *matPlus2Min[5];    /* matPlus2Min1..5, [ny*nz][9] */
  const float *matMin2Plus[5];    /* matMin2Plus1..5, [ny*nz][9] */
  /* matT1toVx_Min, matVytoVx_Min, matVztoVx_Min,
     matT1toVx_Plus, matVytoVx_Plus, matVztoVx_Plus */
  const float *matM2[6];          /* each [ny*nz][9] */
  /* free-surface row (k = nk2), each [ny][9]:
     matVx2Vz1, matVy2Vz1, matVx2Vz2, matVy2Vz2,
     matPlus2Min1f, matPlus2Min2f, matPlus2Min3f,
     matMin2Plus1f, matMin2Plus2f, matMin2Plus3f,
     matT1toVxf_Min, matVytoVxf_Min, matT1toVxf_Plus, matVytoVxf_Plus */
  const float *matF[14];
} cgfd_fault_desc;

/*
 * Everything drv_rk_curv_col_allstep reads from its struct arguments on the
 * single-process path.
 */
typedef struct cgfd_desc {
  int ni, nj, nk;           /* gd_t ni,nj,nk (physical points) */
  /* gd_metric_t v4d: 10 x nx*ny*nz in the order jac, xi_x, xi_y, xi_z,
     eta_x, eta_y, eta_z, zeta_x, zeta_y, zeta_z (gd_t.cu:150-190) */
  const float *metric;
  /* md_t: lambda, mu and rho, where rho already holds 1/rho
     (md_rho_to_slow, main_curv_col_el_3d.cu:705) */
  const float *lambda, *mu, *rho;
  /* bdryfree_t is_sides_free[2][1] (only the top is implemented upstream) */
  int free_top;
  /* bdrypml_t (bdry_t.h:97-123): per face [dim][side] */
  int pml_is_sides[3][2];
  int pml_ni1[3][2], pml_ni2[3][2], pml_nj1[3][2], pml_nj2[3][2];
  int pml_nk1[3][2], pml_nk2[3][2];   /* ghost-inclusive index ranges */
  const float *pml_A[3][2], *pml_B[3][2], *pml_D[3][2];  /* nlay+1 each */
  /* par_t */
  int imethod;              /* dynamic_method, 1 or 2 (par_t.cu:313-315) */
  int number_fault;
  cgfd_fault_desc fault[CGFD_MAX_FAULTS];
  float dt, t0;
  int nt_total;             /* number_of_time_steps + 1 (main:479) */
  int device_id;            /* HIP device ordinal */
  int qc_check_nan_every;   /* >0: check the wavefield for NaN every n steps
                               (the reference only prints, drv_rk_curv_col.cu:594-596) */
  int compat_flags;         /* 0 = reproduce the reference exactly; see below */
} cgfd_desc;

/* compat_flags bits */
#define CGFD_COMPAT_FIX_F2W_SWAP 1  /* write sigma_xz into Txz and sigma_yz into Tyz
                                       in fault2wave; the reference swaps them
                                       (transform.cu:274-276 vs :295) */
#define CGFD_COMPAT_FIX_STATION_INTERP 2  /* interpolate stations between nodes
                                             with corner 4 for its term; the
                                             reference reads corner 5 twice
                                             (io_funcs.cu:459-460) */

typedef struct cgfd_core cgfd_core;

/* last error message of the calling thread ("" if none) */
const char *cgfd_last_error(void);

/* library version string */
const char *cgfd_version(void);

/* Allocates device memory, uploads the descriptor's arrays, computes the
   free-surface conversion matrices (drv_rk_curv_col.cu:184-201) and zeroes
   the state exactly as alloc.cu does. */
int cgfd_core_create(const cgfd_desc *desc, cgfd_core **out);

int cgfd_core_destroy(cgfd_core *core);

/* Advances nsteps time steps from the current step counter (the body of the
   time loop, drv_rk_curv_col.cu:209-636, without the I/O hooks). */
int cgfd_core_run(cgfd_core *core, int nsteps);

/* Graph replay (default on): cgfd_core_run captures eight consecutive steps
   (one cycle of the operator-direction pairs) into a HIP graph the first time
   it reaches a step index divisible by 8 and replays it for every further
   block of eight steps; other steps are launched directly. Changing the
   variant, z chunk or stations drops the graph. Off when the NaN check is on.
   No counterpart in the reference, which launches every kernel from the host
   (drv_rk_curv_col.cu:209-636). */
int cgfd_core_set_graph(cgfd_core *core, int on);

/* Step-boundary state, for checkpoint-and-recompute and restart files. No
   counterpart in the reference, which has no restart (its state between
   steps is drv_rk_curv_col.cu:118-127,161-166 plus the fault accumulators of
   fault_info.h:97-118 and hT1/mT1 of fault_wav_t.h:30-38, which is exactly
   what is saved: wavefield and fault-state levels at the step boundary, PML
   auxiliaries, fault outputs and accumulators, hT1, mT1, rupture flags,
   surface PG and displacement, and the step counter).
   cgfd_core_state_bytes: size of the state in bytes.
   cgfd_core_save_state: copies it into dst (device memory if dst_on_device,
   else host memory); the core is unchanged.
   cgfd_core_load_state: overwrites the core's state from src (as written by
   cgfd_core_save_state of a core created from the same descriptor) and
   resets the step index to the saved one; the next steps are bit-identical
   to the steps that followed the save. Single-scenario cores only
   (CGFD_ERR_ARG for batches). */
int cgfd_core_state_bytes(cgfd_core *core, size_t *bytes);
int cgfd_core_save_state(cgfd_core *core, void *dst, int dst_on_device);
int cgfd_core_load_state(cgfd_core *core, const void *src, int src_on_device);

/* Discrete adjoint of the time loop: the interior velocity-stress stencil,
   CFS-PML, the free-surface traction image and reduced-order velocity
   gradient, RK4, and the split-node fault with slip-weakening friction. The
   reference has no adjoint; this is the transpose of the discrete operators
   of sv_curv_col_el_iso_gpu.cu:231-1290, sv_curv_col_el_iso_fault_gpu.cu:
   191-1879, trial_slipweakening.cu:91-388, transform.cu:100-376 and
   fault_wav_t.cu:197-366 composed with the RK4 stages of
   drv_rk_curv_col.cu:329-568.

   Usage: set the stations (cgfd_core_set_stations) and, on a core with a
   fault, call cgfd_adj_forward(core, nt) at the same step; run the nt
   recorded steps; then call cgfd_adj_backward or cgfd_adj_backward_fault.
   residual[nsta][nt][9] (the layout of cgfd_core_read_stations) defines
   J = sum_n sum_s residual[s][n] . w_n(s); for a least-squares trace misfit
   pass residual = recorded - target.

   cgfd_adj_forward: arms the friction tape for the next nt steps (16 floats
   per fault node and step: the trial traction of each stage, Slip at the
   step start, the velocity jump of the last stage), written by the forward
   fault kernels; the friction law is differentiated on the branch the
   forward run took. No-op on a core without a fault. On a batch every
   scenario gets its own tape.
   cgfd_adj_backward: grad[9][nz][ny][nx] (host layout, ghosts zero)
   receives dJ/dw at the first recorded step.
   cgfd_adj_backward_fault: the same (grad may be null) and
   grad_fault[nfault][7][nk][nj] receives dJ/d(mu_s, mu_d, Dc, C0, T0x, T0y,
   T0z) at every fault node (zero at welded nodes). With fault receivers set
   at the same step (cgfd_core_set_fault_receivers), residual_fault
   [nrec][nt][9] (the layout of cgfd_core_read_fault_receivers, may be null)
   adds sum_n sum_r residual_fault[r][n] . out_n(r) to J over the slip-rate
   and slip outputs Vs, Vs1, Vs2, Slip, Slip1, Slip2; its traction entries
   (Tn, Ts1, Ts2) must be zero (CGFD_ERR_ARG otherwise), and sample 0, which
   precedes the recorded steps, carries no gradient.
   Faults whose split nodes have fault columns inside a CFS-PML slab are
   refused (CGFD_ERR_ARG). The core's forward state is not touched. */
int cgfd_adj_forward(cgfd_core *core, int nt);
int cgfd_adj_backward(cgfd_core *core, const float *residual, int nt, float *grad);
int cgfd_adj_backward_fault(cgfd_core *core, const float *residual, const float *residual_fault, int nt, float *grad,
                            float *grad_fault);
/* The same for scenario s of a batch (cgfd_batch_create) with its own
   residual[nsta][nt][9] (cgfd_batch_read_stations of s); grad and/or
   grad_fault may be null; no fault-receiver term. The scenario must have run
   the recorded steps. */
int cgfd_batch_adj_backward(cgfd_core *core, int s, const float *residual, int nt, float *grad, float *grad_fault);

/* Stage fork (default off): within each RK stage, after the trial traction,
   the interior volume blocks, the CFS-PML blocks with the top layers, and
   the fault columns with the split nodes run on three default-priority
   streams and join at the end of the stage; they read the stage input and
   write disjoint points, so results are identical either way. Off: every
   kernel on one stream in the reference's order (sv_curv_col_el_iso_gpu.cu:
   22-225 then the fault kernels, drv_rk_curv_col.cu:329-568). */
int cgfd_core_set_overlap(cgfd_core *core, int on);

/* Current step counter (number of completed steps). */
int cgfd_core_step_index(const cgfd_core *core);

/* Blocks until all queued work of the core has finished. */
int cgfd_core_sync(cgfd_core *core);

/* Copies the current wavefield (level "pre" = state at t0 + it*dt) into
   host_dst[9][nx*ny*nz], reference layout. */
int cgfd_core_read_wavefield(cgfd_core *core, float *host_dst);

/* Overwrites the current wavefield (level "pre") from host_src[9][nx*ny*nz].
   Ghost points are forced to zero (the reference keeps them at zero on one
   process, alloc.cu:448-457). */
int cgfd_core_write_wavefield(cgfd_core *core, const float *host_src);

/* Fault outputs of fault `id`: host_dst[11][ny*nz] in the order
   Tn,Ts1,Ts2,Vs,Vs1,Vs2,Slip,Slip1,Slip2,Peak_vs,Init_t0 (fault_info.cu:865-942). */
int cgfd_core_read_fault(cgfd_core *core, int id, float *host_dst);

/* Fault split-node state of fault `id` at level "pre": host_dst[9][2][ny*nz]
   (Vx,Vy,Vz,T2x,T2y,T2z,T3x,T3y,T3z; side 0 minus, 1 plus). */
int cgfd_core_read_fault_state(cgfd_core *core, int id, float *host_dst);

/* Overwrites that split-node state (same layout) for a restart or a
   prescribed initial condition. t1rec, if not null, is [6][ny*nz]: the T1
   rate hT1 (x,y,z) and the T1 of the last completed step mT1 (x,y,z) that
   method 2 uses (fault_wav_t.cu:242-248); null leaves them. The host-side
   counterpart of the reference's fault_wav_t contents (fault_wav_t.cu:21-78),
   which its driver only ever zero-initialises. */
int cgfd_core_write_fault_state(cgfd_core *core, int id, const float *host_src, const float *t1rec);

/* Surface peak values host_dst[15][nx*ny] (wav_t.cu:160-174); zero if the
   top is not free. */
int cgfd_core_read_pg(cgfd_core *core, float *host_dst);

/* Receivers: n stations at grid indices (ghost-inclusive i,j,k). The core
   records Vx..Txy of level "pre" at every step before the step is taken
   (io_recv_keep, drv_rk_curv_col.cu:215) into a device buffer of nt_max
   samples. */
int cgfd_core_set_stations(cgfd_core *core, int n, const int *ijk, int nt_max);
/* Stations between nodes (io_recv_keep, io_funcs.cu:427-469): base node
   ijk[n][3] (ghost-inclusive) and increments dijk[n][3] towards i+1, j+1,
   k+1 (cgfd_io_read_stations gives both); each component is the trilinear
   sum over the cell's eight corners, in the reference's term order and
   FP32 rounding, with its corner-5 term twice unless compat_flags has
   CGFD_COMPAT_FIX_STATION_INTERP. Increments 0 give the node values. */
int cgfd_core_set_stations_interp(cgfd_core *core, int n, const int *ijk, const float *dijk, int nt_max);
/* host_dst[n][nt_recorded][9]; returns the number of recorded samples in *nt */
int cgfd_core_read_stations(cgfd_core *core, float *host_dst, int *nt);

/* Slices and snapshots of level "pre" (the state at the current step), packed
   on the device so only the requested values are copied to the host.
   cgfd_core_read_slice (io_slice_nc_put, io_funcs.cu:1013-1131, pack kernels
   :1653-1689): the nine components Vx..Txy on the plane at ghost-inclusive
   `index` along `axis` (0 = x: host_dst[9][nk][nj]; 1 = y: [9][nk][ni];
   2 = z: [9][nj][ni]).
   cgfd_core_read_snapshot (io_snap_nc_put, io_funcs.cu:1380-1583):
   box = {i1, ni, di, j1, nj, dj, k1, nk, dk} (ghost-inclusive starts, counts,
   strides, inside the physical grid); flags bit 0 velocity (3 fields), bit 1
   stress (6), bit 2 strain (6, isotropic compliance of
   io_snap_stress_to_strain_eliso, io_funcs.cu:1585-1651); host_dst holds the
   selected groups in that order, each field [nk][nj][ni]. The batch form
   reads scenario s. */
int cgfd_core_read_slice(cgfd_core *core, int axis, int index, float *host_dst);
int cgfd_core_read_snapshot(cgfd_core *core, const int *box, int flags, float *host_dst);
int cgfd_batch_read_snapshot(cgfd_core *core, int s, const int *box, int flags, float *host_dst);
/* cgfd_core_read_snapshot without any device allocation: no counterpart in
   the reference; for callers that cannot allocate device memory (code under
   HIP graph capture, serving loops, harnesses that forbid hipMalloc). The
   packed values go through a fixed 32 MiB device array per GPU (at most
   8 Mi values: a box of nk x nj x ni nodes and 3, 6, 9 or 15 fields);
   larger boxes return CGFD_ERR_ARG. Calls on one GPU share that array, so
   calls that overlap on one GPU race. */
int cgfd_core_read_snapshot_noalloc(cgfd_core *core, const int *box, int flags, float *host_dst);

/* Fault receivers (io_fault_recv_keep, io_funcs.cu:2313-2351): n receivers
   fjk[n][3] = (fault id, ghost-inclusive j, k) on grid nodes of the fault
   planes. The core records the nine fault outputs Tn, Ts1, Ts2, Vs, Vs1,
   Vs2, Slip, Slip1, Slip2 at every step before the step is taken (as the
   reference's call at drv_rk_curv_col.cu:217) into nt_max samples; reading
   gives host_dst[n][nt_recorded][9]. Receivers here sit on nodes (the
   reference's dj = dk = 0 case); cgfd_core_set_fault_receivers_interp
   places them between nodes. On a batch core every active scenario
   records into a block of its own (cgfd_batch_read_fault_receivers);
   cgfd_core_read_fault_receivers reads scenario 0's. */
int cgfd_core_set_fault_receivers(cgfd_core *core, int n, const int *fjk, int nt_max);
/* Fault receivers between nodes: increments djk[n][2] towards j+1, k+1
   (cgfd_io_read_fault_stations); each output is the bilinear sum of
   io_fault_recv_keep over the four nodes. */
int cgfd_core_set_fault_receivers_interp(cgfd_core *core, int n, const int *fjk, const float *djk, int nt_max);
int cgfd_core_read_fault_receivers(cgfd_core *core, float *host_dst, int *nt);

/* Timing support: runs nsteps and returns the total elapsed device time of
   the stream (HIP events) in *ms_total and the summed duration of the volume
   stencil kernel launches in *ms_volume (events bracket each launch), with the
   number of volume launches in *n_volume. */
int cgfd_core_run_timed(cgfd_core *core, int nsteps, float *ms_total,
                        float *ms_volume, int *n_volume);

/* Kernel variant selection for the volume stencil. 0 = automatic: the
   variant and z-chunk cgfd_core_autotune chose for a problem of this shape
   in this process (or in the cache file named by CGFD_TUNE_CACHE), else 33
   (32 for a batch of more than one scenario).
   1 = one direct-load kernel (27 derivatives per point); 20 = an interior
   kernel that accumulates each field's gradient into the rhs (64 x 4 tiles
   in slab order: zchunk y-tiles per slab, then the plane) and a lean kernel
   for the tiles touching a CFS-PML slab that accumulates the directional
   rhs parts field by field; 23 and 24 = the box split: the accumulating
   interior kernel over exactly the points no CFS-PML face contains (x
   bounds on 32-column boundaries; 64 x 4 tiles for 23, 32 x 8 for 24; slab
   order) and the lean PML kernel over the up to six slabs around them in
   32 x 8 tiles, launched in up to four groups by the CFS-PML face
   directions their points can lie in; 32 = 23 on one scenario; 33 = 23
   with every band in one launch that handles all three directions. Batches with more than one active
   scenario run, for 20, the interior tiles with the scenarios back to back
   (XCD-contiguous) and the PML tiles in one slab-ordered sweep per
   scenario; for 23, 24 and 33 the box split the same way (interior tiles with
   the scenarios back to back, bands in one sweep per scenario); for 32 the
   box split with the interior box, too, in one slab-ordered sweep per
   scenario; for 1 one direct tile x scenario kernel. Returns CGFD_ERR_ARG
   for any other value. */
int cgfd_core_set_variant(cgfd_core *core, int variant);

/* The schedule the core will run: variant and z-chunk (after variant 0 has
   been resolved). */
int cgfd_core_get_variant(const cgfd_core *core, int *variant, int *zchunk);

/* y-tiles per slab of the slab-ordered kernels (20, 23, 24, 32, 33), 4..4096. */
int cgfd_core_set_zchunk(cgfd_core *core, int zchunk);

/* Vertical-column metric form, no counterpart in the reference (whose
   kernels read all nine metric terms, sv_curv_col_el_iso_gpu.cu:332-340).
   On grids whose x depends on i alone and y on j alone (flat grids and
   grids of vertical columns under topography) xi_y, xi_z, eta_x and eta_z
   vanish; the metric computed from such a grid holds there only the
   rounding of its finite differences, about 1e-15 of xi_x and eta_y. At
   create the core checks on the device whether they stay below 2^-32 of
   xi_x and eta_y at every physical point. If so, the box-split schedules
   (23, 24, 32, 33) and the adjoint's transposed gather read only the other five
   terms and omit the four products, which agrees with the nine-term form
   to the last bit at nearly every point and within an ulp elsewhere, with
   four fewer arrays per point; the free-surface and fault kernels take the
   same form for their traction vectors. on = 0 makes the core use the
   nine-term form regardless; on = 1 (the default) uses the five-term form
   where the grid has it. *grid: whether the grid has it; *in_use: whether the volume
   update uses it. */
int cgfd_core_set_vertical_columns(cgfd_core *core, int on);
int cgfd_core_get_vertical_columns(const cgfd_core *core, int *grid, int *in_use);

/* Autotuner: times every candidate (variant, z-chunk) pair -- 23, 33, 24,
   20 and 1, and 32 on a batch, z-chunks 4..32 for the slab-ordered ones -- on
   `reps` repetitions
   of one step's four volume launches, keeps
   the fastest and records it for this problem shape (grid, scenarios,
   faults, boundaries) so later cores of the same shape start with it, and
   appends it to the file named by CGFD_TUNE_CACHE if that is set; *ms_best
   is the average per launch. The launches read level "pre" and write only
   levels that the next step rewrites before reading, so the simulation
   state is unchanged. */
int cgfd_core_autotune(cgfd_core *core, int reps, int *variant, int *zchunk, float *ms_best);

/* Measurement support, no counterpart in the reference: the streaming rate
   (GB/s of bytes read + written) of a kernel that has every point of an
   ni x nj x nk grid (the core's padded array layout, 64 x 4 tiles,
   non-temporal accesses) read nread arrays and write nwrite arrays once,
   averaged over reps launches on device_id. With each RK stage's counts
   (17, 26, 26 and 44 reads and 9 writes per point on a grid of vertical
   columns; 21, 30, 30 and 48 with the nine-term metric) this is the data
   movement of the volume update without the stencil: the ceiling its
   roofline fraction is compared with. Those counts (and 39 + 18) run fully
   unrolled kernels, other counts a loop. Allocates (nread + nwrite) x
   (ni+6)(nj+6)(nk+6) floats for the duration. */
int cgfd_calib_stream(int device_id, int ni, int nj, int nk, int nread, int nwrite, int reps, float *gbs);

/* ---------------------------------------------------------------------
 * Batched ensembles: nscen rupture scenarios on one grid. The scenarios
 * share the grid, metric, media, CFS-PML, free surface, fault geometry and
 * fault coefficient matrices (taken from descs[0]; the other descriptors
 * must agree on sizes, boundaries, method, fault indices and dt) and differ
 * in the per-node fault inputs (T0x/y/z, mu_s, mu_d, Dc, C0, Tn0/Ts10/Ts20)
 * and in nt_total. The kernels load the shared data of a point once for all
 * scenarios. cgfd_core_run advances every scenario that has not yet taken
 * its nt_total steps; a scenario that has retires and is no longer touched.
 * All cgfd_core_* functions apply to a batch, reading scenario 0.
 * No counterpart in the reference, which runs one scenario per process.
 * --------------------------------------------------------------------- */
int cgfd_batch_create(const cgfd_desc *descs, int nscen, cgfd_core **out);
int cgfd_batch_active(cgfd_core *core);                 /* scenarios still running */
int cgfd_batch_steps(const cgfd_core *core, int s);     /* steps scenario s has taken */
int cgfd_batch_read_wavefield(cgfd_core *core, int s, float *host_dst);
int cgfd_batch_write_wavefield(cgfd_core *core, int s, const float *host_src);
int cgfd_batch_read_fault(cgfd_core *core, int s, int id, float *host_dst);
int cgfd_batch_read_fault_state(cgfd_core *core, int s, int id, float *host_dst);
int cgfd_batch_write_fault_state(cgfd_core *core, int s, int id, const float *host_src, const float *t1rec);
int cgfd_batch_read_pg(cgfd_core *core, int s, float *host_dst);
int cgfd_batch_read_stations(cgfd_core *core, int s, float *host_dst, int *nt);

/* Fault receivers of scenario s of a batch (cgfd_core_set_fault_receivers
   on the batch core records every active scenario): [nrec][*nt][9] as
   cgfd_core_read_fault_receivers, *nt the steps scenario s recorded. No
   reference counterpart (the reference runs one scenario). */
int cgfd_batch_read_fault_receivers(cgfd_core *core, int s, float *host_dst, int *nt);

/* The one-call drop-in: create + run nt_total steps + copy the final state
   into the optional outputs (NULL to skip) + destroy. */
int cgfd_drv_rk_curv_col_allstep(const cgfd_desc *desc, float *wave_out,
                                 float *fault_out, float *pg_out);

#ifdef __cplusplus
}
#endif

#endif /* CGFD_CORE_H */
