/*
 * cgfd_kernels.h -- host-side launchers of the gfx950 kernels.
 */
#ifndef CGFD_KERNELS_H
#define CGFD_KERNELS_H

#include "cgfd_dev.h"

namespace cgfd {

/* one fault in device memory (fault_one_t + fault_coef_one_t + the T1
   traction columns of fault_wav_t; fault_info.h:11-119, fault_wav_t.h:18-63) */
struct FaultDev {
  /* per scenario (block of 45 floats per node: T0 3, fric 4, out 11, T1 21,
     hT1 3, mT1 3; and 2 ints: flag_rup, init_t0_flag); the rest is shared */
  int i0;                        /* ghost-inclusive x index of the fault plane */
  const float *T0;               /* [3][sslice] T0x,T0y,T0z */
  const float *fric;             /* [4][sslice] mu_s, mu_d, Dc, C0 */
  const int *united, *faultgrid, *rupy, *rupz;
  int *flag_rup, *init_t0_flag;
  float *out;                    /* [11][sslice] Tn..Init_t0 */
  const float *rho_f;            /* [2][sslice] */
  const float *vs;               /* [6][sslice] s1x,s1y,s1z,s2x,s2y,s2z */
  const float *D;                /* [12][9][sslice] D21_1..D33_2 */
  const float *P2M, *M2P;        /* [5][9][sslice] */
  const float *M2;               /* [6][9][sslice] T1toVx_Min,VytoVx_Min,VztoVx_Min,.._Plus */
  const float *Fm;               /* [14][9][ny] free-surface row */
  const float *fmet;             /* [10][sslice] jac and metric on the plane's column i0 (a copy of the
                                    grid's metric there, contiguous along j for the node kernels) */
  float *T1;                     /* [3][7][sslice] */
  float *hT1, *mT1;              /* [3][sslice] */
  double *slip64;                /* per scenario [3][sslice]: Slip, Slip1, Slip2 accumulated in FP64 */
};

struct FaultArgs {
  Grid g;
  const float *met, *lam, *mu, *slw, *matVx2Vz, *matVy2Vz;
  float *w_cur, *w_pre, *w_tnew;
  float *w_end;                             /* level A: t_{n+1} is written over it at stage 3 */
  const float *w_b;                         /* level B (read at stage 3) */
  float *f_cur, *f_pre, *f_tnew, *f_end;   /* level bases; fault id at +id*18*sslice */
  int nfault;
  const FaultDev *ftab;                     /* [nfault] in device memory (scenario 0) */
  long sfp, sfi;                            /* this scenario's offsets into the per-node blocks */
  long sfd;                                 /* and into the FP64 slip accumulators */
  int istage, idir, jdir, kdir;
  int isfree, imethod, fix_q1;
  int vcol;                                 /* vertical columns: xi_y, xi_z, eta_x, eta_z left out */
  float dt, ca, cb, rkb;
  const int *it_dev;                        /* completed steps (device counter) */
  /* friction tape of the adjoint (null: off): per fault [tape_nt][16][sslice],
     per step the trial traction of each stage (12), Slip at the step start
     (1) and the split-node velocity jump of the last stage (3) */
  float *tape;
  int tape_it0, tape_nt;
  long tape_sstride;                        /* floats between scenarios' tapes */
};
constexpr int kTapeFloats = 16;

/* the fault arguments of scenario s: its wavefield and fault-state levels,
   and the offsets of its per-node parameters, outputs and traction columns
   (one block of 45 floats and 2 ints per node and fault, see FaultDev) */
DEVI FaultArgs fault_at(const FaultArgs &a, const BatArgs &b, int s) {
  FaultArgs r = a;
  const long w = (long)s * b.wstride, f = (long)s * b.fwstride;
  r.w_cur += w;
  r.w_pre += w;
  if (r.w_tnew) r.w_tnew += w;
  r.w_end += w;
  if (r.w_b) r.w_b += w;
  r.f_cur += f;
  r.f_pre += f;
  if (r.f_tnew) r.f_tnew += f;
  r.f_end += f;
  r.sfp = (long)s * b.fpf;
  r.sfi = (long)s * b.fpi;
  r.sfd = (long)s * b.fpd;
  if (r.tape) r.tape += (long)s * a.tape_sstride;
  return r;
}

/* fault `id` as seen by the scenario of `a` */
DEVI FaultDev fdev(const FaultArgs &a, int id) {
  FaultDev F = a.ftab[id];
  F.T0 += a.sfp;
  F.fric += a.sfp;
  F.out += a.sfp;
  F.T1 += a.sfp;
  F.hT1 += a.sfp;
  F.mT1 += a.sfp;
  F.flag_rup += a.sfi;
  F.init_t0_flag += a.sfi;
  F.slip64 += a.sfd;
  return F;
}

/* variants with separate interior and CFS-PML block kernels launch the PML
   blocks on st_pml (nullptr = st); the two write disjoint points */
void launch_volume(int variant, const VolArgs &a, const BatArgs &b, int idir, int jdir, int kdir, int top,
                   hipStream_t st, hipStream_t st_pml = nullptr);
void launch_dvh2dvz(const Grid &g, const float *met, const float *lam, const float *mu, float *mx, float *my,
                    hipStream_t st);
/* *flag (device, zeroed by the caller) becomes nonzero if at some physical
   point |xi_y| or |xi_z| exceeds 2^-32 |xi_x|, or |eta_x| or |eta_z| exceeds
   2^-32 |eta_y| (or one of them is NaN) */
void launch_vcol_check(const Grid &g, const float *met, int *flag, hipStream_t st);
void launch_pg(const Grid &g, const BatArgs &b, const float *wend, const float *wpre, float *PG, float *Dis, float dt,
               hipStream_t st);

void launch_wave2fault(const FaultArgs &a, const BatArgs &b, hipStream_t st);
/* fmet[c][j + k ny] = met[c] at (i0, j, k), every ghost-inclusive (j, k) */
void launch_fault_metric(const Grid &g, const float *met, int i0, float *fmet, hipStream_t st);
void launch_trial(const FaultArgs &a, const BatArgs &b, hipStream_t st);
void launch_fault_update(const FaultArgs &a, const BatArgs &b, hipStream_t st);
void launch_step_end(int *it_dev, hipStream_t st);
void launch_zero(float *p, long n, hipStream_t st);
void launch_repitch(float *dst, long dpitch, const float *src, long spitch, long width, long rows, hipStream_t st);
void launch_nan_check(const float *p, long n, int *flag, hipStream_t st);
/* discrete adjoint of the volume update (k_adjoint.hip) */
struct AdjFace {          /* one CFS-PML face of the adjoint */
  int on, dim;
  int i1, i2, j1, j2, k1, k2, an, bn, ax0;  /* row layout as the forward face's (PmlFace) */
  long n;                 /* slab entries */
  const float *A, *B, *D; /* profiles along the face's axis */
  float *lev;             /* auxiliary adjoint levels E, A, B, P (by index), [9][n] each */
  float *dsg;             /* extra S and G of the face, [9][n] */
};
struct AdjPml {
  int on;
  int le, lp;             /* level indices of E and P in every face's lev (A = 1, B = 2) */
  AdjFace f[6];           /* dim-major, side-minor */
  FaultSkip fs;           /* fault columns: their rhs belongs to the fault adjoint */
};
/* fault part of the adjoint (k_adj_fault.hip): per fault a block of
   kAdjFaultFloats x sslice floats: split-node adjoint levels [4][18] (E, A,
   B, P by index), T1 [21], hT1 [3], mT1 [3], Slip [1], parameter gradients
   [7] (mu_s, mu_d, Dc, C0, T0x, T0y, T0z), Slip1 and Slip2 [2], the
   receiver residuals of Vs, Vs1, Vs2 for the step being reversed [3] */
constexpr int kAdjFaultFloats = 112;
struct AdjFaultArgs {
  Grid g;
  const float *met, *lam, *mu, *slw, *matVx2Vz, *matVy2Vz;
  const FaultDev *ftab;
  int nfault;
  float *adjf;
  const float *tape;
  int tape_nt, n;          /* tape steps; slot of the step being reversed */
  int istage, idir, jdir, kdir, isfree, imethod, fix_q1;
  float dt, ca, cb;
  int le, lp, lx, lc;      /* split-node levels: E, P, the stage output (-1 at stage 3), the stage input */
  float *wT;               /* wavefield adjoint of the stage input */
  float *wX, *wE;          /* of the stage output (null at stage 3) and of the end level */
};
void launch_adjf_begin(const AdjFaultArgs &a, hipStream_t st);
/* residuals of fault-receiver sample `sample` ([nrec][nt][9], Tn..Slip2) into
   the adjoint of the outputs the step before it produced */
void launch_adjf_receivers(const AdjFaultArgs &a, const int *rec, const float *fac, int nterm, int nrec,
                           const float *res, int nt, int sample, hipStream_t st);
void launch_adjf_stage_pre(const AdjFaultArgs &a, hipStream_t st);
void launch_adjf_stage_post(const AdjFaultArgs &a, hipStream_t st);
void launch_adj_prep(int mode, const Grid &g, const float *lam, const float *mu, const float *slw, const float *X,
                     float alpha, const float *E, float beta, float *SG, const AdjPml &pm, int isfree, int kdir,
                     hipStream_t st);
void launch_adj_top(int mode, int idir, int jdir, int kdir, const Grid &g, const float *met, const float *lam,
                    const float *mu, const float *slw, const float *mx, const float *my, const float *X, float alpha,
                    const float *E, float beta, float *T, const AdjPml &pm, hipStream_t st);
/* vcol: the grid's columns are vertical (the gather reads the xi and eta
   metric rows' one nonzero term each) */
void launch_adj_stage(int mode, int idir, int jdir, int kdir, const Grid &g, const float *met, const float *SG,
                      const float *E, float *A, float *B, float *P, const AdjPml &pm, int vcol, hipStream_t st);
void launch_adj_inject(const long *term, const float *fac, int nterm, int nsta, const float *res, int nt, int n,
                       long sicmp, float *P, hipStream_t st);
void launch_record_stations(const BatArgs &b, const float *w, long sicmp, const long *term, const float *fac,
                            int nterm, int n, float *buf, int nt_max, const int *it_dev, int it0, hipStream_t st);
void launch_record_fault(const BatArgs &b, const FaultDev *ftab, const int *rec, const float *fac, int nterm, int n,
                         float *buf, int nt_max, const int *it_dev, int it0, long ss, hipStream_t st);
/* output boxes (k_io.hip): start, count, stride per axis, ghost-inclusive */
struct IoBox {
  int s[3], n[3], d[3];
};
void launch_pack_box(const float *w, const Grid &g, const float *lam, const float *mu, const IoBox &b, int flags,
                     float *out, hipStream_t st);

}  // namespace cgfd

#endif
