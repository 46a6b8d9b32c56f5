/*
 * cgfd_dev.h -- device-side types and finite-difference primitives of the
 * gfx950 time-stepping core.
 *
 * Numerics follow the reference operators (src/forward/fd_t.h:9-140,
 * fd_t.cu:71-159): the same coefficients, the same term order, and the same
 * float/double mix where the reference's double literals promote an
 * expression (fault and free-surface paths). The volume stencil uses the FP32
 * operator table exactly as the reference's inner kernel does
 * (sv_curv_col_el_iso_gpu.cu:102-147).
 */
#ifndef CGFD_DEV_H
#define CGFD_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEVI __device__ __forceinline__

namespace cgfd {

/* traction-image coefficients a_0..a_3 (fd_t.h:9-12) */
constexpr double kA0 = 1.2500392377, kA1 = 1.5417647610, kA2 = -0.3334118088, kA3 = 0.0416862855;
/* 4th-order one-sided coefficients b_1..b_3 (fd_t.h:26-28) */
constexpr double kB1 = -7.0 / 6.0, kB2 = 8.0 / 6.0, kB3 = -1.0 / 6.0;
/* DRP/MacCormack coefficients c_1..c_5 (fd_t.h:58-62) */
constexpr double kC1 = -0.291647047, kC2 = -0.666745142, kC3 = 1.250117713, kC4 = -0.333411808, kC5 = 0.041686285;

/* FP32 operator table of the volume stencil (fd_t.cu:106-109, stored as
   float in the reference): forward taps at -1..3, backward taps at -3..1 */
#define CGFD_CF0 ((float)-0.291647047)
#define CGFD_CF1 ((float)-0.666745142)
#define CGFD_CF2 ((float)1.250117713)
#define CGFD_CF3 ((float)-0.333411808)
#define CGFD_CF4 ((float)0.041686285)
#define CGFD_CB0 ((float)-0.041686285)
#define CGFD_CB1 ((float)0.333411808)
#define CGFD_CB2 ((float)-1.250117713)
#define CGFD_CB3 ((float)0.666745142)
#define CGFD_CB4 ((float)0.291647047)

/* ---------------------------------------------------------------------- */
/* grid */
struct Grid {
  int ni, nj, nk, nx, ny, nz;
  int nk2;           /* last physical k (ghost-inclusive) */
  long siy, siz;     /* strides: siy is the device row pitch, nx rounded up to
                        32 floats so that every row starts on a 128-byte line */
  long sicmp;        /* points per component (siy*ny*nz) */
  long sslice;       /* ny*nz */
};

/* one CFS-PML face (bdrypml_t, bdry_t.h:97-123). A face spanning the grid
   in x (the y and z faces) stores its rows from the 32-column boundary at or
   below i1 with a pitch of whole 32-column segments, so a tile row of the
   volume kernels reads its auxiliaries in the same 128-byte lines as its
   wavefield; the narrow x faces store their columns packed. */
__host__ __device__ inline void pml_row_layout(int dim, int i1, int i2, int &ax0, int &an) {
  if (dim == 0) { ax0 = i1; an = i2 - i1 + 1; }
  else { ax0 = i1 & ~31; an = (i2 + 1 - ax0 + 31) & ~31; }
}
struct PmlFace {
  int on;
  int i1, i2, j1, j2, k1, k2;  /* ghost-inclusive index ranges */
  int an, bn;                  /* row pitch (x) and rows per plane (y) of the slab's arrays */
  int ax0;                     /* x index of a row's first stored column */
  long n;                      /* slab entries (an x bn x planes) */
  const float *A, *B, *D;      /* profiles along the face's axis */
  float *pre, *la, *lb, *lc;   /* aux levels, [9][n] each: the step's start and
                                  the outputs of stages 0, 1, 2 (stage 3 writes
                                  t_{n+1} over la), as the wavefield's */
};

/* the fault columns the volume kernels leave to the fault kernel */
struct FaultSkip {
  int nfault;
  int i0[5];
  const int *united[5];
};

/* arguments shared by the volume and top-layer kernels */
struct VolArgs {
  Grid g;
  const float *__restrict__ cur;   /* stage input level, [9][sicmp] */
  const float *__restrict__ pre;   /* level at t_n */
  float *__restrict__ tnew;        /* next stage input (stages 0..2) */
  float *__restrict__ end;         /* level A, the output of stage 0: stage 3
                                      writes the level at t_{n+1} over it */
  const float *__restrict__ lb;    /* level B, the output of stage 1 (read at stage 3) */
  const float *__restrict__ met;   /* [10][sicmp]: jac, xi, eta, zeta */
  const float *__restrict__ lam;
  const float *__restrict__ mu;
  const float *__restrict__ slw;
  const float *__restrict__ matVx2Vz;   /* free-surface conversion, [nx*ny][9] */
  const float *__restrict__ matVy2Vz;
  PmlFace pml[3][2];
  int pml_on;
  int isfree;
  int istage;
  float ca, cb;      /* rk_a*dt, rk_b*dt */
  int k_lo;          /* first physical k offset handled */
  int nk_run;        /* number of k planes handled */
  int zchunk;        /* planes per block (LDS kernel) */
  int vcol;          /* 1: the grid's columns are vertical (xi_y, xi_z, eta_x
                        and eta_z below 2^-32 of xi_x and eta_y at every
                        physical point): the box split skips those four
                        metric arrays */
  FaultSkip fs;
};

/* the box-split schedule of the volume update: the interior box (points of
   no CFS-PML face, x bounds on 32-column boundaries) and the bands around
   it; ghost-inclusive, half-open ranges */
struct VolBox {
  int x0, x1, y0, y1, k0, k1;
  int ntx, nty;  /* tiles of the box; ntx = 0: empty */
};
constexpr int kBandRegions = 6; /* regions per band launch, at most */
struct VolBands {
  int n;  /* regions */
  int x0[kBandRegions], x1[kBandRegions], y0[kBandRegions], y1[kBandRegions], k0[kBandRegions], k1[kBandRegions];
  int ox[kBandRegions];                 /* x origin of a region's 32-wide tiles (32-aligned) */
  int ntx[kBandRegions], nty[kBandRegions]; /* 32 x 8 tiles per plane */
  int sy[kBandRegions];                 /* y tiles per slab: a region's blocks run x tile
                                           fastest, then the sy y tiles of a slab, then the
                                           plane, then the slab */
  long cum[kBandRegions + 1];           /* block prefix sums over the regions (padded slabs) */
};

/* ---------------------------------------------------------------------- */
/* Scenarios of a batch share the grid, metric, media, CFS-PML profiles and
   fault coefficients; each has its own wavefield levels, PML auxiliaries,
   fault state, fault parameters and outputs, stored as equal-sized blocks
   one after another. A kernel loops over the active scenarios, loading the
   shared data of a point once. A plain run is a batch of one. */
constexpr int kMaxScen = 64;

struct BatArgs {
  int nact;              /* active scenarios of this launch */
  int act[kMaxScen];     /* their indices */
  long wstride;          /* floats between scenarios: wavefield levels */
  long pstride;          /* CFS-PML auxiliaries */
  long fwstride;         /* fault split-node state levels */
  long fpf, fpi;         /* fault parameters/outputs (floats), flags (ints) */
  long fpd;              /* FP64 slip accumulators (doubles) */
  long pgstride;         /* PG + displacement */
  long ststride;         /* station records */
};

/* ---------------------------------------------------------------------- */
/* volume operators with the FP32 table; DIR 1 forward, 0 backward */
template <int DIR>
DEVI float dfd(const float *__restrict__ p, long s) {
  if (DIR == 1)
    return CGFD_CF0 * p[-s] + CGFD_CF1 * p[0] + CGFD_CF2 * p[s] + CGFD_CF3 * p[2 * s] + CGFD_CF4 * p[3 * s];
  else
    return CGFD_CB0 * p[-3 * s] + CGFD_CB1 * p[-2 * s] + CGFD_CB2 * p[-s] + CGFD_CB3 * p[0] + CGFD_CB4 * p[s];
}

/* Loads and stores at a 32-bit byte offset from a wave-uniform base pointer.
   The address is base (SGPRs) + zero-extended offset (one VGPR), which the
   compiler emits as the saddr form of global_load/store: one offset register
   serves all nine fields and all metric arrays, instead of a 64-bit address
   pair per load. Valid while one field array is below 4 GiB (checked at
   create). */
/* The RK4 step's end value from its start value p, the inputs of stages 1,
   2 and 3 (A = p + dt h0 / 2, B = p + dt h1 / 2, C = p + dt h2) and
   t = dt h3 / 6: p + dt (h0 + 2 h1 + 2 h2 + h3) / 6, with dt h_s taken back
   from the stage inputs (the subtractions A - p and the like are exact
   whenever the stage increment is at most half the value, Sterbenz). The reference accumulates
   end += b_s dt h_s after every stage (drv_rk_curv_col.cu:329-568); this
   form needs only the stage inputs, so stages 0..2 write one level instead
   of two, and agrees with it to a few ulps of the value. */
DEVI float rk_end(float p, float A, float B, float C, float t) {
  const float s = __builtin_fmaf(2.0f, B - p, (A - p) + (C - p));
  return p + (s / 3.0f + t);
}

DEVI float ld32(const float *base, unsigned off) { return *(const float *)((const char *)base + off); }
DEVI void st32(float *base, unsigned off, float v) { *(float *)((char *)base + off) = v; }
/* the same for data a stage touches once (RK levels, metric, media): marked
   non-temporal so they do not displace the stencil's working set in L2 */
DEVI float ld32nt(const float *base, unsigned off) {
  return __builtin_nontemporal_load((const float *)((const char *)base + off));
}
DEVI void st32nt(float *base, unsigned off, float v) { __builtin_nontemporal_store(v, (float *)((char *)base + off)); }

template <int DIR>
DEVI float dfd32(const float *__restrict__ b, unsigned o, unsigned s) {
  if (DIR == 1)
    return CGFD_CF0 * ld32(b, o - s) + CGFD_CF1 * ld32(b, o) + CGFD_CF2 * ld32(b, o + s) +
           CGFD_CF3 * ld32(b, o + 2 * s) + CGFD_CF4 * ld32(b, o + 3 * s);
  else
    return CGFD_CB0 * ld32(b, o - 3 * s) + CGFD_CB1 * ld32(b, o - 2 * s) + CGFD_CB2 * ld32(b, o - s) +
           CGFD_CB3 * ld32(b, o) + CGFD_CB4 * ld32(b, o + s);
}

/* the same operator on five loaded taps t[0..4] at offsets (-1..3) or (-3..1) */
template <int DIR>
DEVI float dfd5(float t0, float t1, float t2, float t3, float t4) {
  if (DIR == 1) return CGFD_CF0 * t0 + CGFD_CF1 * t1 + CGFD_CF2 * t2 + CGFD_CF3 * t3 + CGFD_CF4 * t4;
  else return CGFD_CB0 * t0 + CGFD_CB1 * t1 + CGFD_CB2 * t2 + CGFD_CB3 * t3 + CGFD_CB4 * t4;
}

/* fd_t.h macros with double constants (fault and boundary paths) */
DEVI float mac22(const float *p, long s, int f) { return f == 1 ? p[s] - p[0] : p[0] - p[-s]; }
DEVI float mac24(const float *p, long s, int f) {
  return f == 1 ? (float)((kB1 * p[0]) + (kB2 * p[s]) + (kB3 * p[2 * s]))
                : (float)(-(kB1 * p[0]) - (kB2 * p[-s]) - (kB3 * p[-2 * s]));
}
DEVI float macdrp_f(const float *p, long s) {
  return (float)((kC1 * p[-s]) + (kC2 * p[0]) + (kC3 * p[s]) + (kC4 * p[2 * s]) + (kC5 * p[3 * s]));
}
DEVI float macdrp_b(const float *p, long s) {
  return (float)(-(kC1 * p[s]) - (kC2 * p[0]) - (kC3 * p[-s]) - (kC4 * p[-2 * s]) - (kC5 * p[-3 * s]));
}
DEVI float macdrp(const float *p, long s, int f) { return f == 1 ? macdrp_f(p, s) : macdrp_b(p, s); }
DEVI float vecdrp_f(const float *v) {
  return (float)((kC1 * v[0]) + (kC2 * v[1]) + (kC3 * v[2]) + (kC4 * v[3]) + (kC5 * v[4]));
}
DEVI float vecdrp_b(const float *v) {
  return (float)(-(kC5 * v[0]) - (kC4 * v[1]) - (kC3 * v[2]) - (kC2 * v[3]) - (kC1 * v[4]));
}
DEVI float vecdrp(const float *v, int f) { return f == 1 ? vecdrp_f(v) : vecdrp_b(v); }
DEVI float vec24_f(const float *v) { return (float)((kB1 * v[0]) + (kB2 * v[1]) + (kB3 * v[2])); }
DEVI float vec24_b(const float *v) { return (float)(-(kB3 * v[0]) - (kB2 * v[1]) - (kB1 * v[2])); }
/* 5-tap operator on a 7-vector whose centre is c (M_FD_VEC) */
DEVI float vecc(const float *c, int f) {
  return f == 1 ? (float)((kC1 * c[-1]) + (kC2 * c[0]) + (kC3 * c[1]) + (kC4 * c[2]) + (kC5 * c[3]))
                : (float)(-(kC1 * c[1]) - (kC2 * c[0]) - (kC3 * c[-1]) - (kC4 * c[-2]) - (kC5 * c[-3]));
}

/* ---------------------------------------------------------------------- */
/* 3x3 helpers, src/lib/fdlib_math.cu:4-84 arithmetic */
DEVI float norm3(const float *A) {
  float x2 = A[0] * A[0], y2 = A[1] * A[1], z2 = A[2] * A[2];
  return sqrtf(x2 + y2 + z2);
}
DEVI void inv3(float m[3][3]) {
  double v[3][3];
  v[0][0] = (float)(m[1][1] * m[2][2] - m[2][1] * m[1][2]);
  v[0][1] = (float)(m[2][1] * m[0][2] - m[0][1] * m[2][2]);
  v[0][2] = (float)(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
  v[1][0] = (float)(m[1][2] * m[2][0] - m[1][0] * m[2][2]);
  v[1][1] = (float)(m[0][0] * m[2][2] - m[2][0] * m[0][2]);
  v[1][2] = (float)(m[1][0] * m[0][2] - m[0][0] * m[1][2]);
  v[2][0] = (float)(m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  v[2][1] = (float)(m[2][0] * m[0][1] - m[0][0] * m[2][1]);
  v[2][2] = (float)(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
  double det = v[0][0] * m[0][0] + v[0][1] * m[1][0] + v[0][2] * m[2][0];
  det = 1.0 / det;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) m[i][j] = (float)(v[i][j] * det);
}
DEVI void mm33(const float A[3][3], const float B[3][3], float C[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      float c = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; k++) c += A[i][k] * B[k][j];
      C[i][j] = c;
    }
}
/* y = M x where M is stored SoA: M[e*stride] for element e (row-major) */
DEVI void mv3(const float *__restrict__ M, long stride, const float *x, float *y) {
  float m0 = M[0], m1 = M[stride], m2 = M[2 * stride], m3 = M[3 * stride], m4 = M[4 * stride];
  float m5 = M[5 * stride], m6 = M[6 * stride], m7 = M[7 * stride], m8 = M[8 * stride];
  y[0] = m0 * x[0] + m1 * x[1] + m2 * x[2];
  y[1] = m3 * x[0] + m4 * x[1] + m5 * x[2];
  y[2] = m6 * x[0] + m7 * x[1] + m8 * x[2];
}
DEVI float dot3(const float *A, const float *B) {
  float r = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; i++) r += A[i] * B[i];
  return r;
}

}  // namespace cgfd

#endif
