/*
 * cgfd_io.h -- output files of the time loop in the reference's NetCDF
 * layouts, written without libnetcdf (NetCDF classic, 64-bit offset
 * variant, readable by any NetCDF reader). The values are packed on the
 * device (cgfd_core_read_fault / _read_slice / _read_pg); the driver calls
 * these at the reference's cadence (drv_rk_curv_col.cu:219-235, 641-647):
 *
 *   every io_time_skip steps, before the step:   cgfd_io_fault_put(f, core, it / io_time_skip, t)
 *                                                 cgfd_io_slice_put(s, core, it / io_time_skip, t)
 *   after the loop:                               cgfd_io_fault_end(f, core); cgfd_io_close(s);
 *                                                 cgfd_io_pg_write(core, out_dir, frame)
 *
 * frame is the reference's output_fname_part ("px0_py0_pz0" on one GPU;
 * NULL or "" means that). All functions return 0 or a negative CGFD_ERR_*
 * code with cgfd_last_error() set.
 */
#ifndef CGFD_IO_H
#define CGFD_IO_H

#include "cgfd_core.h"
#include "cgfd_setup.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cgfd_out cgfd_out;

/* <out_dir>/fault_i<fault x index + 1>_<frame>.nc (io_fault_nc_create,
   io_funcs.cu:1952-2025): dims time (unlimited), k (nk), j (nj); float time,
   Tn, Ts1, Ts2, Vs, Vs1, Vs2, Slip, Slip1, Slip2 (time, k, j), Peak_vs,
   Init_t0 (k, j); global i_index_with_ghosts_in_this_thread,
   coords_of_mpi_topo. */
int cgfd_io_fault_create(cgfd_core *core, int fault_id, const char *out_dir, const char *frame, cgfd_out **out);
/* record it_skip: time and the nine outputs (io_fault_nc_put, :2027-2099) */
int cgfd_io_fault_put(cgfd_out *f, cgfd_core *core, int it_skip, float time);
/* Peak_vs and Init_t0 (io_fault_end_t_nc_put, :2101-2139), then closes f */
int cgfd_io_fault_end(cgfd_out *f, cgfd_core *core);

/* <out_dir>/slice{x,y,z}_{i,j,k}<index + 1>_<frame>.nc at ghost-inclusive
   `index` along axis 0/1/2 (io_slice_nc_create, :900-1011): dims time and
   (k, j), (k, i) or (j, i); float time and Vx..Txy; global
   {i,j,k}_index_with_ghosts_in_this_thread, coords_of_mpi_topo. */
int cgfd_io_slice_create(cgfd_core *core, int axis, int index, const char *out_dir, const char *frame,
                         cgfd_out **out);
/* record it_skip: time and the nine components (io_slice_nc_put, :1013-1131) */
int cgfd_io_slice_put(cgfd_out *s, cgfd_core *core, int it_skip, float time);
int cgfd_io_close(cgfd_out *s);

/* <out_dir>/<name>_<frame>.nc (io_snap_nc_create, :1260-1378): the
   strided box of cgfd_core_read_snapshot, dims time, k, j, i; float time and,
   by flags (1 velocity, 2 stress, 4 strain), Vx Vy Vz, Txx Tyy Tzz Txz Tyz
   Txy, Exx Eyy Ezz Exz Eyz Exy; global first_index_to_snapshot_output,
   first_index_in_this_thread_with_ghosts, index_stride_in_this_thread,
   coords_of_mpi_topo. cgfd_io_snap_put writes record rec (the reference's
   (it - it1) / dit at steps with (it - it1) % dit == 0, io_funcs.cu:1417-1424). */
int cgfd_io_snap_create(cgfd_core *core, const char *name, const int *box, int flags, const char *out_dir,
                        const char *frame, cgfd_out **out);
int cgfd_io_snap_put(cgfd_out *snap, cgfd_core *core, int rec, float time);

/* <out_dir>/PG_V_A_D_<frame>.nc (PG_slice_output, :1827-1879): dims j (ny),
   i (nx), ghost-inclusive; PGV, PGVh, PGVx..PGDz; global
   global_index_of_first_physical_points, count_index_of_physical_points,
   coords_of_mpi_topo. Needs a free surface. */
int cgfd_io_pg_write(cgfd_core *core, const char *out_dir, const char *frame);

/* Station traces as SAC files (io_recv_output_sac, io_funcs.cu:577-618;
   sacExport1C1R, src/lib/sacLib.cu:8-93): <out_dir>/<names[s]>.<cmp>.sac for
   cmp = Vx, Vy, Vz, Txx, Tyy, Tzz, Tyz, Txz, Txy, with every sample of the
   station window (cgfd_core_set_stations), delta = dt, b = 0, station
   position xyz[s] in stla, stlo, stel and, as in the reference's call, dt in
   stdp; native byte order. With with_strain also <name>.{Exx,Eyy,Ezz,Eyz,
   Exz,Exy}.sac, the strain from the stress with the isotropic compliance at
   the station node (io_recv_output_sac_el_iso_strain, io_funcs.cu:626-700). */
int cgfd_io_stations_sac(cgfd_core *core, const char *out_dir, const char *const *names, const float *xyz,
                         int with_strain);

/* Reads the reference's fault input init_stress_<n>.nc (nc_read_init_stress,
   fault_info.cu:1076-1187; written by the examples' prep_fault/conf_fault_stress.m
   as NetCDF classic): fields[7][nk][nj] = Tx, Ty, Tz, mu_s, mu_d, Dc, C0 (float
   or double in the file), to fill cgfd_fault_desc. Host only; no GPU call. */
int cgfd_io_read_init_stress(const char *path, int nj, int nk, float *fields);
/* Reads the fault-plane coordinates of fault_coord_<n>.nc (nc_read_fault_geometry,
   gd_t.cu:548-569): xyz[3][nk][nj] = x, y, z, the plane a grid generator
   builds the curvilinear grid around. Host only. */
int cgfd_io_read_fault_coord(const char *path, int nj, int nk, float *xyz);

/* The grid and media imports of the reference's main (grid_generation_method
   "import", metric "import", media "import"), host only:
   cgfd_io_read_coord: coord_<frame>.nc x, y, z -> xyz[3][nk][nj][ni] (the
     leading block of larger variables, as gd_curv_coord_import, gd_t.cu:855-923);
   cgfd_io_read_metric: metric_<frame>.nc jac, xi_x .. zeta_z ->
     metric[10][nk][nj][ni] (gd_curv_metric_import, gd_t.cu:1015-1068);
   cgfd_io_read_media: media_<frame>.nc lambda, mu, rho -> media[3][nk][nj][ni]
     (md_import, md_t.cu:229-286; density, not slowness).
   Variables may be float or double; the metric and media variables must be
   exactly [nk][nj][ni], as the reference reads them whole. The outputs feed
   cgfd_setup_params (x, y, z, lambda, mu, rho, metric). */
int cgfd_io_read_coord(const char *path, int ni, int nj, int nk, float *xyz);
int cgfd_io_read_metric(const char *path, int ni, int nj, int nk, float *metric);
int cgfd_io_read_media(const char *path, int ni, int nj, int nk, float *media);

/* Station lists (host only).
   cgfd_io_read_stations reads the reference's station file
   (io_recv_read_locate, io_funcs.cu:23-320): a count, then one line per
   station "name flag_coord flag_depth x y z" ('#' lines and empty lines
   skipped). flag_coord 0: x, y, z are grid indices counted from 1, z counted
   down from the free surface from 0 when flag_depth is 1; the station sits
   at the truncated index with the fractional rest as its increments.
   flag_coord 1: x, y, z are coordinates (z a depth below the surface when
   flag_depth is 1, converted with the top-layer cell that contains x, y,
   gd_t.cu:1522-1612); the station goes to the nearest grid node and, inside
   the hexahedral cell around it that contains the point, to the nearest
   sample of a 101^3 trilinear lattice (gd_t.cu:1365-1503, 1694-1775).
   Stations outside the physical grid are dropped, as the reference drops
   them with a warning. The reference's locating kernels test flag_coord == 0
   where its host code expects 1 (io_funcs.cu:511, 533 against :83), which
   leaves coordinate stations at uninitialised indices; this reads them as
   the host code intends. out = NULL returns the count of the file in
   *n_file only; otherwise out holds room for *n_file stations, and
   *n_located are written, in file order. The increments (di, dj, dk) and
   base nodes feed cgfd_core_set_stations_interp. */
#define CGFD_STATION_NAME_MAX 256
typedef struct cgfd_station {
  char name[CGFD_STATION_NAME_MAX];
  int i, j, k;        /* ghost-inclusive base node */
  float di, dj, dk;   /* increments towards i+1, j+1, k+1 */
  float x, y, z;      /* coordinates: as given, or of the base node */
} cgfd_station;
int cgfd_io_read_stations(const char *path, const cgfd_setup *setup, cgfd_station *out, int max_n, int *n_file,
                          int *n_located);

/* Fault stations (io_fault_recv_read_locate, io_funcs.cu:2170-2300): a
   count, then "name fault_id j k" per line, fault_id, j and k counted from
   1, j and k fractional; the base node is the floor, the rest the
   increments. A station off the fault plane is an error (CGFD_ERR_ARG), as
   in the reference. x, y, z are the coordinates of the base node. */
typedef struct cgfd_fault_station {
  char name[CGFD_STATION_NAME_MAX];
  int fault_id;       /* from 0 */
  int j, k;           /* ghost-inclusive base node on the fault plane */
  float dj, dk;
  float x, y, z;
} cgfd_fault_station;
int cgfd_io_read_fault_stations(const char *path, const cgfd_setup *setup, cgfd_fault_station *out, int max_n,
                                int *n_file);

/* A receiver line of the par file (receiver_line_index_start / _incre /
   _count, io_line_locate, io_funcs.cu:312-425): the nodes start + n incre
   (global physical indices from 0) for n < count that lie in the grid, in
   out (room for count) as stations named "<name>.no<n>" -- the names the
   reference gives their SAC files (io_line_output_sac, :771), so
   cgfd_io_stations_sac writes <name>.no<n>.<cmp>.sac -- with zero
   increments and the node's coordinates. io_line_keep records the nodes'
   values each step, as cgfd_core_set_stations(_interp) does. */
int cgfd_io_line_stations(const cgfd_setup *setup, const char *name, const int *start, const int *incre, int count,
                          cgfd_station *out, int *n_out);

#ifdef __cplusplus
}
#endif

#endif
