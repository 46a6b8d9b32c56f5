/*
 * cgfd_setup.h -- host-side setup of the C++ driver (CPU only).
 *
 * Restates the reference's host setup that runs in main() before the time
 * loop and produces everything cgfd_desc carries across the boundary:
 *   gd_curv_metric_cal + geometric_symmetry   (src/forward/gd_t.cu:209-304, 665-759)
 *   mirror_symmetry of imported media          (src/forward/gd_t.cu:578-655, md_t.cu:278)
 *   bdry_pml_set + profile functions           (src/forward/bdry_t.cu:79-445)
 *   fault_coef_cal                             (src/forward/fault_info.cu:97-850)
 *   fault_set (masks, initial traction)        (src/forward/fault_info.cu:967-1074)
 *   md_rho_to_slow                             (src/forward/md_t.cu:568-582)
 *   blk_dt_esti_curv                           (src/forward/blk_t.cu:71-151)
 * The setup object owns every array the descriptor points to.
 */
#ifndef CGFD_SETUP_H
#define CGFD_SETUP_H

#include "cgfd_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cgfd_setup_params {
  int ni, nj, nk;                  /* number_of_total_grid_points_x/y/z */
  /* physical coordinates, ni*nj*nk each, index i + j*ni + k*ni*nj
     (the layout of coord/media NetCDF imports, gd_t.cu:855-923) */
  const float *x, *y, *z;
  /* physical media (density, not slowness), same layout */
  const float *lambda, *mu, *rho;
  int free_top;                                  /* boundary_z_top: free */
  int pml_layers[3][2];                          /* cfspml number_of_layers, 0 = none */
  float pml_alpha_max[3][2], pml_beta_max[3][2], pml_velocity[3][2];
  int imethod;                                   /* dynamic_method */
  int number_fault;
  int fault_x_index[CGFD_MAX_FAULTS];            /* 0-based (par_t.cu:323) */
  int fault_grid[4 * CGFD_MAX_FAULTS];           /* 0-based j1,j2,k1,k2 (par_t.cu:337) */
  /* init_stress_<id>.nc variables Tx,Ty,Tz,mu_s,mu_d,Dc,C0 on the physical
     fault plane, nj*nk each, index j + k*nj (fault_info.cu:1094-1180) */
  const float *init_stress[CGFD_MAX_FAULTS][7];
  int united_margin;     /* fault_info.cu:1046-1055 hard-codes 40 */
  float dt, t0;
  int nt_total;
  int device_id;
  /* optional imported metric (metric_<frame>.nc, cgfd_io_read_metric):
     jac, xi_x .. zeta_z, ni*nj*nk each, layout as x; NULL: computed from
     the coordinates (gd_curv_metric_cal). Ghosts come from
     geometric_symmetry either way (gd_curv_metric_import, gd_t.cu:1064). */
  const float *metric;
  /* 1: compute the metric and the fault coefficients with HIP kernels on
     device_id (the same arithmetic; the results land in the setup's host
     arrays as before); 0: on the host */
  int device_setup;
} cgfd_setup_params;

typedef struct cgfd_setup cgfd_setup;

/* Runs the whole host setup. */
int cgfd_setup_create(const cgfd_setup_params *p, cgfd_setup **out);
int cgfd_setup_destroy(cgfd_setup *s);

/* Fills *d with the boundary descriptor (pointers into the setup object). */
int cgfd_setup_get_desc(const cgfd_setup *s, cgfd_desc *d);

/* A further rupture scenario on the setup's grid, for cgfd_batch_create:
   init_stress[7*id + q] holds field q (Tx,Ty,Tz,mu_s,mu_d,Dc,C0, layout as in
   cgfd_setup_params) of fault id. The setup keeps copies of these inputs and
   the initial tractions projected on the fault frame (fault_info.cu:1015-1023)
   and fills *d with a descriptor that shares everything else with the
   setup's own. The descriptor stays valid until cgfd_setup_destroy. */
int cgfd_setup_scenario_desc(cgfd_setup *s, const float *const *init_stress, int nt_total, cgfd_desc *d);

/* Extended-grid coordinates x3d,y3d,z3d (3 x nx*ny*nz) after geometric_symmetry. */
const float *cgfd_setup_coords(const cgfd_setup *s);

/* Fault-plane vec_n ([ny*nz][3]) of fault id (fault_info.cu:637-651). */
const float *cgfd_setup_vec_n(const cgfd_setup *s, int id);

/* Standalone pieces, for tests: metric of an extended grid (coords incl.
   ghosts, nx*ny*nz each) into metric[10][nx*ny*nz]. */
int cgfd_setup_metric(int ni, int nj, int nk, const float *x3d, const float *y3d,
                      const float *z3d, float *metric);

/* Maximum stable dt (blk_dt_esti_curv with CFL 1.3, fd_t.cu:57) of the setup. */
int cgfd_setup_dt_max(const cgfd_setup *s, float *dtmax);

/* PML profile helpers (bdry_t.cu:291-324) */
float cgfd_pml_cal_R(float num_lay);
float cgfd_pml_cal_dmax(float L, float Vp, float Rpp);

#ifdef __cplusplus
}
#endif

#endif /* CGFD_SETUP_H */
