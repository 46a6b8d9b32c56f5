/*
 * cgfd_setup_dev.h -- the device half of the setup (csrc/k_setup.hip),
 * called by the host setup when cgfd_setup_params.device_setup is set.
 * Both run the node arithmetic of cgfd_setup_node.h; results land in the
 * host arrays the descriptor points to. Return CGFD_OK or a negative code
 * with the message recorded for cgfd_last_error().
 */
#ifndef CGFD_SETUP_DEV_H
#define CGFD_SETUP_DEV_H

#include "cgfd_setup_node.h"

/* inputs per fault node: grad xi, grad eta, grad zeta, jac, lambda, mu */
static const int kFaultNodeIn = 12;

/* gd_curv_metric_cal (gd_t.cu:209-304) at the physical points of an
   ni x nj x nk grid with 3 ghosts: coords = x, y, z over (nk+6, nj+6, ni+6)
   with the ghosts already extrapolated; metric = 10 components of the same
   shape, written at the physical points only. */
int cgfd_dev_metric(int device_id, int ni, int nj, int nk, const float *coords, float *metric);

/* fault_coef_cal (fault_info.cu:155-845) for every node of one fault plane
   of ny x nz nodes (ghosts included); node_in = kFaultNodeIn floats per node
   j + k ny; k_top = the free-surface row. out = host arrays. */
int cgfd_dev_fault_coef(int device_id, int ny, int nz, int k_top, const float *node_in,
                        const cgfd_setup_math::FaultCoefOut &out);

#endif
