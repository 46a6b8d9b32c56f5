/*
 * cgfd_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference time-stepping core
 * (drv_rk_curv_col_allstep, src/forward/drv_rk_curv_col.cu:28-667) and of every
 * per-stage operator it calls, templated on the arithmetic type:
 *   - Oracle<float>  follows the reference's FP32 arithmetic, including the
 *     places where double literals promote an expression to FP64 (the C++
 *     promotion rules are the same as CUDA C++'s, so writing the expressions
 *     as the reference writes them reproduces the promotion);
 *   - Oracle<double> is the same algorithm in FP64 throughout.
 * It deliberately keeps the reference's structure: a separate rhs level, the
 * whole-level RK update kernels, one pass per operator in the reference's
 * order. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it; the product never links or calls it.
 *
 * Each function names the reference lines it restates.
 *
 * What pins it (tests/test_oracle_pins.py, tests/test_oracle_fault_pins.py,
 * tests/test_setup_ref.py, tests/test_oracle_adjoint.py, tests/test_abi.py):
 * the interior rhs, FD operators and RK4 composition by the closed-form step
 * of cubic fields; the metric by affine grids; friction by closed-form
 * values at its kinks; the free surface by sigma.n = 0 after surface
 * rupture; the CFS-PML by absorbed energy; FP32 vs FP64; the trial traction
 * assembly (its factor 2, the side signs, the DyT2/DzT3 terms) and the
 * split-node velocity rows by a locked fault that transmits a 3-D pulse
 * without slip; the slipping path (friction cap, DxT1 rows, fault2wave's
 * side averages, slip/slip-rate/rupture-time accumulators, free-surface
 * rows) by a uniform stress drop against the 1-D solution, strike-slip to
 * the surface and buried dip-slip, both methods; the split-node stress rows
 * with the method-1/2 matrices, method 2's T1 rate and the free-surface
 * rows by a uniform strain rate reproduced exactly on Cartesian and sheared
 * grids; wave2fault/fault2wave by their round trip, with and without the
 * Txz/Tyz swap; the setup arrays the oracle reads (metric, PML profiles,
 * fault matrices, masks) by an independent FP64 restatement
 * (oracle/setup_ref.py) and closed forms, and the oracle runs on that
 * restatement in the parity tests. Each fault pin has a planted-mistake
 * control (plant, below) that fails it.
 * Parity unpinned (line-by-line review against the cited lines only): the
 * one-step lag of hT1 (a constant T1 rate, the only closed form available,
 * cannot tell a lagged rate from a current one), and the values of method
 * 2's top three split-node rows on a fault reaching the surface, where the
 * reference takes no T1 rate outside its fault grid (the tests show the
 * rows differ from the exact solution there, as the rule implies, but no
 * closed form fixes them).
 * oracle_ad_gradient differentiates the FP64 restatement in reverse mode
 * (ad::Var below): the FP64 reference adjoint the GPU adjoint is checked
 * against, itself pinned by FP64 central differences.
 * oracle_read_kink_margin reports how close each node's friction branch
 * decisions came to their switch points over a run (instrumentation of
 * friction_point for the gradient tests, which compare derivatives only
 * where the piecewise-smooth map is smooth).
 */
#include "../include/cgfd_core.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <type_traits>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ---- reverse-mode differentiation of the FP64 restatement ----
   Oracle<ad::Var> runs the same statements on values that record every
   arithmetic operation on a tape (one node per result: up to two parents
   and the local partial derivatives). A reverse sweep over the tape gives
   the exact FP64 gradient of a scalar function of the outputs with respect
   to the marked inputs, following the branches (friction law, guards) the
   FP64 run takes. The tape is global: Oracle<ad::Var> runs on one thread. */
namespace ad {
struct Node {
  int a, b;
  double da, db;
};
std::vector<Node> tape;

struct Var {
  double v = 0.0;
  int i = -1;  /* tape index; -1: a constant */
  Var() = default;
  Var(double x) : v(x), i(-1) {}  /* NOLINT: constants convert implicitly */
};
inline Var mk(double v, int a, double da, int b = -1, double db = 0.0) {
  if (a < 0 && b < 0) return Var(v);
  Var r(v);
  r.i = (int)tape.size();
  tape.push_back({a, b, da, db});
  return r;
}
inline Var leaf(double v) {
  Var r(v);
  r.i = (int)tape.size();
  tape.push_back({-1, -1, 0.0, 0.0});
  return r;
}
inline Var operator+(Var x, Var y) { return mk(x.v + y.v, x.i, 1.0, y.i, 1.0); }
inline Var operator-(Var x, Var y) { return mk(x.v - y.v, x.i, 1.0, y.i, -1.0); }
inline Var operator*(Var x, Var y) { return mk(x.v * y.v, x.i, y.v, y.i, x.v); }
inline Var operator/(Var x, Var y) { return mk(x.v / y.v, x.i, 1.0 / y.v, y.i, -x.v / (y.v * y.v)); }
inline Var operator+(Var x, double y) { return mk(x.v + y, x.i, 1.0); }
inline Var operator+(double x, Var y) { return mk(x + y.v, y.i, 1.0); }
inline Var operator-(Var x, double y) { return mk(x.v - y, x.i, 1.0); }
inline Var operator-(double x, Var y) { return mk(x - y.v, y.i, -1.0); }
inline Var operator*(Var x, double y) { return mk(x.v * y, x.i, y); }
inline Var operator*(double x, Var y) { return mk(x * y.v, y.i, x); }
inline Var operator/(Var x, double y) { return mk(x.v / y, x.i, 1.0 / y); }
inline Var operator/(double x, Var y) { return mk(x / y.v, y.i, -x / (y.v * y.v)); }
inline Var operator-(Var x) { return mk(-x.v, x.i, -1.0); }
inline Var &operator+=(Var &x, Var y) { return x = x + y; }
inline Var &operator-=(Var &x, Var y) { return x = x - y; }
inline Var &operator*=(Var &x, Var y) { return x = x * y; }
inline Var &operator/=(Var &x, Var y) { return x = x / y; }
inline bool operator<(Var x, Var y) { return x.v < y.v; }
inline bool operator>(Var x, Var y) { return x.v > y.v; }
inline bool operator<=(Var x, Var y) { return x.v <= y.v; }
inline bool operator>=(Var x, Var y) { return x.v >= y.v; }
inline bool operator==(Var x, Var y) { return x.v == y.v; }
inline bool operator!=(Var x, Var y) { return x.v != y.v; }
/* sqrt at 0 takes the derivative 0 (|v| of a zero vector) */
inline Var sqrt(Var x) {
  double r = std::sqrt(x.v);
  return mk(r, x.i, r > 0.0 ? 0.5 / r : 0.0);
}
inline Var fabs(Var x) { return mk(std::fabs(x.v), x.i, x.v < 0.0 ? -1.0 : 1.0); }
inline double val(Var x) { return x.v; }
}  // namespace ad

inline float xsqrt(float x) { return std::sqrt(x); }
inline double xsqrt(double x) { return std::sqrt(x); }
inline ad::Var xsqrt(ad::Var x) { return ad::sqrt(x); }
inline float xabs(float x) { return std::fabs(x); }
inline double xabs(double x) { return std::fabs(x); }
inline ad::Var xabs(ad::Var x) { return ad::fabs(x); }
inline double val(double x) { return x; }
inline double val(float x) { return x; }
using ad::val;
/* the type the reference's FP64 intermediates take: double, or the AD type */
template <class T> struct Wide { using type = double; };
template <> struct Wide<ad::Var> { using type = ad::Var; };

/* ---- fd_t.h:9-12 fault traction-image coefficients ---- */
const double a_0 = 1.2500392377;
const double a_1 = 1.5417647610;
const double a_2 = -0.3334118088;
const double a_3 = 0.0416862855;
/* ---- fd_t.h:26-28 (the macros expand "b_1" to "-7.0/6.0") ---- */
const double b_1 = -7.0 / 6.0;
const double b_2 = 8.0 / 6.0;
const double b_3 = -1.0 / 6.0;
/* ---- fd_t.h:58-62 ---- */
const double c_1 = -0.291647047;
const double c_2 = -0.666745142;
const double c_3 = 1.250117713;
const double c_4 = -0.333411808;
const double c_5 = 0.041686285;

/* fd_t.cu:71-89 (drp row) and fd_t.cu:106-109 (drp coefs); row 0 = backward,
   row 1 = forward, as selected by mac_all_*[3][dir]. */
const int op_indx[2][5] = {{-3, -2, -1, 0, 1}, {-1, 0, 1, 2, 3}};
const double op_coef_d[2][5] = {
    {-0.041686285, 0.333411808, -1.250117713, 0.666745142, 0.291647047},
    {-0.291647047, -0.666745142, 1.250117713, -0.333411808, 0.041686285}};

/* fd_t.cu:149-159 */
const int FD_Flags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                            {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};

/* ---- fd_t.h macros as functions; p points at the current point ---- */
template <class T> inline T mac22_f(const T *p, long s) { return p[s] - p[0]; }
template <class T> inline T mac22_b(const T *p, long s) { return p[0] - p[-s]; }
template <class T> inline T mac22(const T *p, long s, int f) { return f == 1 ? mac22_f(p, s) : mac22_b(p, s); }

template <class T> inline T mac24_f(const T *p, long s) {
  return (T)((b_1 * p[0]) + (b_2 * p[s]) + (b_3 * p[2 * s]));
}
template <class T> inline T mac24_b(const T *p, long s) {
  return (T)(-(b_1 * p[0]) - (b_2 * p[-s]) - (b_3 * p[-2 * s]));
}
template <class T> inline T mac24(const T *p, long s, int f) { return f == 1 ? mac24_f(p, s) : mac24_b(p, s); }

template <class T> inline T vec24_f(const T *v) { return (T)((b_1 * v[0]) + (b_2 * v[1]) + (b_3 * v[2])); }
template <class T> inline T vec24_b(const T *v) { return (T)(-(b_3 * v[0]) - (b_2 * v[1]) - (b_1 * v[2])); }

template <class T> inline T macdrp_f(const T *p, long s) {
  return (T)((c_1 * p[-s]) + (c_2 * p[0]) + (c_3 * p[s]) + (c_4 * p[2 * s]) + (c_5 * p[3 * s]));
}
template <class T> inline T macdrp_b(const T *p, long s) {
  return (T)(-(c_1 * p[s]) - (c_2 * p[0]) - (c_3 * p[-s]) - (c_4 * p[-2 * s]) - (c_5 * p[-3 * s]));
}
template <class T> inline T macdrp(const T *p, long s, int f) { return f == 1 ? macdrp_f(p, s) : macdrp_b(p, s); }

template <class T> inline T vecdrp_f(const T *v) {
  return (T)((c_1 * v[0]) + (c_2 * v[1]) + (c_3 * v[2]) + (c_4 * v[3]) + (c_5 * v[4]));
}
template <class T> inline T vecdrp_b(const T *v) {
  return (T)(-(c_5 * v[0]) - (c_4 * v[1]) - (c_3 * v[2]) - (c_2 * v[3]) - (c_1 * v[4]));
}
template <class T> inline T vecdrp(const T *v, int f) { return f == 1 ? vecdrp_f(v) : vecdrp_b(v); }

/* M_FD_VEC: p points at the centre of a 7-vector */
template <class T> inline T vecc_f(const T *p) {
  return (T)((c_1 * p[-1]) + (c_2 * p[0]) + (c_3 * p[1]) + (c_4 * p[2]) + (c_5 * p[3]));
}
template <class T> inline T vecc_b(const T *p) {
  return (T)(-(c_1 * p[1]) - (c_2 * p[0]) - (c_3 * p[-1]) - (c_4 * p[-2]) - (c_5 * p[-3]));
}
template <class T> inline T vecc(const T *p, int f) { return f == 1 ? vecc_f(p) : vecc_b(p); }

/* ---- src/lib/fdlib_math.cu ---- */
template <class T> inline T norm3(const T *A) {  /* :4-11 */
  T x2 = A[0] * A[0], y2 = A[1] * A[1], z2 = A[2] * A[2];
  return xsqrt(x2 + y2 + z2);
}
template <class T> inline void invert3x3(T m[][3]) {  /* :14-43, double cofactors */
  typename Wide<T>::type inv[3][3];
  inv[0][0] = (T)(m[1][1] * m[2][2] - m[2][1] * m[1][2]);
  inv[0][1] = (T)(m[2][1] * m[0][2] - m[0][1] * m[2][2]);
  inv[0][2] = (T)(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
  inv[1][0] = (T)(m[1][2] * m[2][0] - m[1][0] * m[2][2]);
  inv[1][1] = (T)(m[0][0] * m[2][2] - m[2][0] * m[0][2]);
  inv[1][2] = (T)(m[1][0] * m[0][2] - m[0][0] * m[1][2]);
  inv[2][0] = (T)(m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  inv[2][1] = (T)(m[2][0] * m[0][1] - m[0][0] * m[2][1]);
  inv[2][2] = (T)(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
  typename Wide<T>::type det = inv[0][0] * m[0][0] + inv[0][1] * m[1][0] + inv[0][2] * m[2][0];
  det = 1.0 / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = (T)(inv[i][j] * det);
}
template <class T> inline void matmul3x3(T A[][3], T B[][3], T C[][3]) {  /* :46-56 */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C[i][j] = 0.0;
      for (int k = 0; k < 3; k++) C[i][j] += A[i][k] * B[k][j];
    }
}
template <class T> inline void matmul3x1(T A[][3], const T *B, T *C) {  /* :59-65 */
  C[0] = A[0][0] * B[0] + A[0][1] * B[1] + A[0][2] * B[2];
  C[1] = A[1][0] * B[0] + A[1][1] * B[1] + A[1][2] * B[2];
  C[2] = A[2][0] * B[0] + A[2][1] * B[1] + A[2][2] * B[2];
}
template <class T> inline T dot3(const T *A, const T *B) {  /* :77-84 */
  T r = 0.0;
  for (int i = 0; i < 3; i++) r += A[i] * B[i];
  return r;
}

template <class T> inline void load9(T M[][3], const T *src) {
  for (int ii = 0; ii < 3; ii++)
    for (int jj = 0; jj < 3; jj++) M[ii][jj] = src[3 * ii + jj];
}

/* ======================================================================= */

/* Linear slip-weakening friction at one node (trial_slipweakening.cu:296-384):
   trial traction -> local traction -> normal/shear split, no opening,
   friction coefficient, strength cap, |Ts| < 0.1 guard. vn must be the unit
   normal. Writes the fault traction T1 (slot 3) and tTn/tTs1/tTs2; returns
   flag_rup. margin (optional, test instrumentation): the smallest relative
   distance so far of each branch decision from its switch point, [0] shear
   vs strength (|Ts| - Tau_c) / Tau_c, [1] slip vs Dc, [2] normal traction vs
   0 (relative to |traction|), [3] |Ts| vs the 0.1 guard. */
template <class T>
int friction_point(const T *Trial, T jacvec, const T *vn, const T *T0, T mus, T mud, T Slip, T Dc, T C0,
                   const T *vs1, const T *vs2, T *T1c, T &tTn, T &tTs1, T &tTs2, double *margin = nullptr) {
  T TL[3], Ts[3];
  for (int a = 0; a < 3; a++) TL[a] = Trial[a] / jacvec + T0[a];
  T Tn0 = dot3(TL, vn);
  for (int a = 0; a < 3; a++) Ts[a] = TL[a] - vn[a] * Tn0;
  T Ts0 = norm3(Ts);
  T Tau_n, Tau_s;
  int ifchange = 0, flag;
  if (Tn0 >= 0.0) { Tau_n = 0.0; ifchange = 1; }
  else Tau_n = Tn0;
  T friction;
  if (Slip <= Dc) friction = mus - (mus - mud) * Slip / Dc;
  else friction = mud;
  T Tau_c = -friction * Tau_n + C0;
  if (margin) {
    const double tc = val(Tau_c), ts = val(Ts0), tn = val(Tn0), dc = val(Dc), sl = val(Slip);
    const double tl = std::sqrt(val(TL[0]) * val(TL[0]) + val(TL[1]) * val(TL[1]) + val(TL[2]) * val(TL[2]));
    const double m[4] = {std::fabs(ts - tc) / std::max(std::fabs(tc), 1.0), std::fabs(sl - dc) / std::max(dc, 1e-30),
                         std::fabs(tn) / std::max(tl, 1.0), std::fabs(ts - 0.1) / 0.1};
    for (int q = 0; q < 4; q++) margin[q] = std::min(margin[q], m[q]);
  }
  if (Ts0 >= Tau_c) { Tau_s = Tau_c; flag = 1; ifchange = 1; }
  else { Tau_s = Ts0; flag = 0; }
  T Tau[3];
  if (ifchange == 1) {
    if (xabs(Ts0) < 1e-1) {
      for (int a = 0; a < 3; a++) Tau[a] = Tau_n * vn[a];
    } else {
      for (int a = 0; a < 3; a++) Tau[a] = Tau_n * vn[a] + (Tau_s / Ts0) * Ts[a];
    }
  } else {
    for (int a = 0; a < 3; a++) Tau[a] = TL[a];
  }
  for (int a = 0; a < 3; a++) T1c[a] = ifchange == 1 ? (Tau[a] - T0[a]) * jacvec : Trial[a];
  tTs1 = dot3(Tau, vs1);
  tTs2 = dot3(Tau, vs2);
  tTn = Tau_n;
  return flag;
}

template <class T>
struct Oracle {
  /* gd_t */
  int ni, nj, nk, nx, ny, nz;
  int ni1, ni2, nj1, nj2, nk1, nk2;
  size_t siy, siz, sicmp, sslice;
  T dt, t0;
  int imethod, isfree, nfault, it;
  int fix_q1;
  /* plant: a deliberate mistake for the tests that show a pin catches it
     (0 = none; see oracle_create). Never set outside those tests. */
  int plant = 0;
  T rk_a[3], rk_b[4];
  T coef[2][5];

  std::vector<T> met, lam, mu, slw;
  T *jac, *xix, *xiy, *xiz, *etx, *ety, *etz, *ztx, *zty, *ztz;

  /* wav_t: v5d[level 4][cmp 9][k][j][i] */
  std::vector<T> wav;
  T *w_pre, *w_tmp, *w_rhs, *w_end;

  /* bdryfree_t */
  std::vector<T> matVx2Vz, matVy2Vz;

  /* bdrypml_t */
  struct Face {
    int on = 0, i1 = 0, i2 = 0, j1 = 0, j2 = 0, k1 = 0, k2 = 0;
    int an = 0, bn = 0, cn = 0;
    size_t n = 0;
    std::vector<T> A, B, D, var;
    T *pre = nullptr, *tmp = nullptr, *rhs = nullptr, *end = nullptr, *cur = nullptr;
  } pml[3][2];
  int pml_enable;

  /* fault_t + fault_coef_t + fault_wav_t */
  struct Fault {
    int i0;
    std::vector<T> T0x, T0y, T0z, mu_s, mu_d, Dc, C0;
    std::vector<int> united, faultgrid, rupy, rupz, flag_rup, init_t0_flag;
    std::vector<double> kink;  /* [sslice][4] branch margins of friction_point */
    std::vector<T> out;  /* [11][sslice] */
    T *Tn, *Ts1, *Ts2, *Vs, *Vs1, *Vs2, *Slip, *Slip1, *Slip2, *Peak_vs, *Init_t0;
    std::vector<T> tTn, tTs1, tTs2;
    std::vector<T> rho_f, vec_s1, vec_s2;
    std::vector<T> D[12], P2M[5], M2P[5], M2[6], F[14];
    std::vector<T> T1x, T1y, T1z, hT1x, hT1y, hT1z, mT1x, mT1y, mT1z;
  };
  std::vector<Fault> flt;
  size_t f_siz_ilevel;  /* 2*9*sslice */
  std::vector<T> fwav;  /* [4][nfault][9][2][sslice] */
  T *f_pre, *f_tmp, *f_rhs, *f_end;

  /* PG */
  std::vector<T> PG, Dis;

  /* ---------------------------------------------------------------- */
  explicit Oracle(const cgfd_desc *d, int fixq1) {
    fix_q1 = fixq1;
    ni = d->ni; nj = d->nj; nk = d->nk;
    nx = ni + 6; ny = nj + 6; nz = nk + 6;
    ni1 = 3; nj1 = 3; nk1 = 3;
    ni2 = ni1 + ni - 1; nj2 = nj1 + nj - 1; nk2 = nk1 + nk - 1;
    siy = nx; siz = (size_t)nx * ny; sicmp = siz * nz; sslice = (size_t)ny * nz;
    dt = (T)d->dt; t0 = (T)d->t0;
    imethod = d->imethod; isfree = d->free_top; nfault = d->number_fault;
    it = 0;
    /* fd_t.cu:38-45 (float in the reference; for T=double the exact ratios) */
    rk_a[0] = (T)0.5; rk_a[1] = (T)0.5; rk_a[2] = (T)1.0;
    rk_b[0] = (T)(1.0 / 6.0); rk_b[1] = (T)(1.0 / 3.0); rk_b[2] = (T)(1.0 / 3.0); rk_b[3] = (T)(1.0 / 6.0);
    for (int r = 0; r < 2; r++)
      for (int n = 0; n < 5; n++) coef[r][n] = (T)op_coef_d[r][n];

    met.assign(d->metric, d->metric + 10 * sicmp);
    jac = met.data(); xix = jac + sicmp; xiy = xix + sicmp; xiz = xiy + sicmp;
    etx = xiz + sicmp; ety = etx + sicmp; etz = ety + sicmp;
    ztx = etz + sicmp; zty = ztx + sicmp; ztz = zty + sicmp;
    lam.assign(d->lambda, d->lambda + sicmp);
    mu.assign(d->mu, d->mu + sicmp);
    slw.assign(d->rho, d->rho + sicmp);

    wav.assign(4 * 9 * sicmp, (T)0);  /* alloc.cu: zeroed */
    w_pre = wav.data(); w_tmp = w_pre + 9 * sicmp; w_rhs = w_tmp + 9 * sicmp; w_end = w_rhs + 9 * sicmp;

    matVx2Vz.assign(siz * 9, (T)0);
    matVy2Vz.assign(siz * 9, (T)0);

    pml_enable = 0;
    for (int dim = 0; dim < 3; dim++)
      for (int side = 0; side < 2; side++) {
        Face &f = pml[dim][side];
        f.on = d->pml_is_sides[dim][side];
        if (!f.on) continue;
        pml_enable = 1;
        f.i1 = d->pml_ni1[dim][side]; f.i2 = d->pml_ni2[dim][side];
        f.j1 = d->pml_nj1[dim][side]; f.j2 = d->pml_nj2[dim][side];
        f.k1 = d->pml_nk1[dim][side]; f.k2 = d->pml_nk2[dim][side];
        f.an = f.i2 - f.i1 + 1; f.bn = f.j2 - f.j1 + 1; f.cn = f.k2 - f.k1 + 1;
        f.n = (size_t)f.an * f.bn * f.cn;
        int np = (dim == 0 ? f.an : (dim == 1 ? f.bn : f.cn));
        f.A.assign(d->pml_A[dim][side], d->pml_A[dim][side] + np);
        f.B.assign(d->pml_B[dim][side], d->pml_B[dim][side] + np);
        f.D.assign(d->pml_D[dim][side], d->pml_D[dim][side] + np);
        f.var.assign(4 * 9 * f.n, (T)0);  /* bdry_t.cu:283 calloc */
        f.pre = f.var.data(); f.tmp = f.pre + 9 * f.n; f.rhs = f.tmp + 9 * f.n; f.end = f.rhs + 9 * f.n;
      }

    flt.resize(nfault);
    f_siz_ilevel = 2 * 9 * sslice;
    fwav.assign(4 * (size_t)nfault * f_siz_ilevel, (T)0);
    size_t lev = (size_t)nfault * f_siz_ilevel;
    f_pre = fwav.data(); f_tmp = f_pre + lev; f_rhs = f_tmp + lev; f_end = f_rhs + lev;
    for (int id = 0; id < nfault; id++) {
      const cgfd_fault_desc &fd = d->fault[id];
      Fault &F = flt[id];
      F.i0 = fd.fault_x_index + 3;
      auto cp = [&](std::vector<T> &v, const float *src, size_t n) {
        if (src) v.assign(src, src + n); else v.assign(n, (T)0);
      };
      auto cpi = [&](std::vector<int> &v, const int *src, size_t n) {
        if (src) v.assign(src, src + n); else v.assign(n, 0);
      };
      cp(F.T0x, fd.T0x, sslice); cp(F.T0y, fd.T0y, sslice); cp(F.T0z, fd.T0z, sslice);
      cp(F.mu_s, fd.mu_s, sslice); cp(F.mu_d, fd.mu_d, sslice); cp(F.Dc, fd.Dc, sslice); cp(F.C0, fd.C0, sslice);
      cpi(F.united, fd.united, sslice); cpi(F.faultgrid, fd.faultgrid, sslice);
      cpi(F.rupy, fd.rup_index_y, sslice); cpi(F.rupz, fd.rup_index_z, sslice);
      F.flag_rup.assign(sslice, 0);       /* fault_info.cu:1069 */
      F.init_t0_flag.assign(sslice, 0);   /* fault_info.cu:960 */
      F.kink.assign(4 * sslice, 1e300);
      F.out.assign(11 * sslice, (T)0);
      T *o = F.out.data();
      F.Tn = o; F.Ts1 = o + sslice; F.Ts2 = o + 2 * sslice; F.Vs = o + 3 * sslice;
      F.Vs1 = o + 4 * sslice; F.Vs2 = o + 5 * sslice; F.Slip = o + 6 * sslice;
      F.Slip1 = o + 7 * sslice; F.Slip2 = o + 8 * sslice; F.Peak_vs = o + 9 * sslice;
      F.Init_t0 = o + 10 * sslice;
      for (size_t p = 0; p < sslice; p++) {
        F.Tn[p] = fd.Tn0 ? (T)fd.Tn0[p] : (T)0;
        F.Ts1[p] = fd.Ts10 ? (T)fd.Ts10[p] : (T)0;
        F.Ts2[p] = fd.Ts20 ? (T)fd.Ts20[p] : (T)0;
        F.Init_t0[p] = (T)-9999.9f;  /* fault_info.cu:1068 */
      }
      F.tTn.assign(sslice, (T)0); F.tTs1.assign(sslice, (T)0); F.tTs2.assign(sslice, (T)0);
      cp(F.rho_f, fd.rho_f, 2 * sslice);
      cp(F.vec_s1, fd.vec_s1, 3 * sslice);
      cp(F.vec_s2, fd.vec_s2, 3 * sslice);
      for (int q = 0; q < 12; q++) cp(F.D[q], fd.D[q], 9 * sslice);
      for (int q = 0; q < 5; q++) { cp(F.P2M[q], fd.matPlus2Min[q], 9 * sslice); cp(F.M2P[q], fd.matMin2Plus[q], 9 * sslice); }
      for (int q = 0; q < 6; q++) cp(F.M2[q], fd.matM2[q], 9 * sslice);
      for (int q = 0; q < 14; q++) cp(F.F[q], fd.matF[q], 9 * (size_t)ny);
      F.T1x.assign(7 * sslice, (T)0); F.T1y.assign(7 * sslice, (T)0); F.T1z.assign(7 * sslice, (T)0);
      F.hT1x.assign(sslice, (T)0); F.hT1y.assign(sslice, (T)0); F.hT1z.assign(sslice, (T)0);
      F.mT1x.assign(sslice, (T)0); F.mT1y.assign(sslice, (T)0); F.mT1z.assign(sslice, (T)0);
    }

    if (isfree) { PG.assign(15 * siz, (T)0); Dis.assign(3 * siz, (T)0); }
    if (isfree) dvh2dvz();
  }

  /* level pointers of one fault inside a fault level */
  T *fcmp(T *level, int id, int cmp) { return level + (size_t)id * f_siz_ilevel + (size_t)cmp * 2 * sslice; }

  /* ---- sv_curv_col_el_iso_dvh2dvz_gpu, sv_curv_col_el_iso_gpu.cu:1296-1405 ---- */
  void dvh2dvz() {
    int k = nk2;
    for (int iy = 0; iy < nj; iy++)
      for (int ix = 0; ix < ni; ix++) {
        size_t iptr = (ix + ni1) + (iy + nj1) * siy + k * siz;
        T e11 = xix[iptr], e12 = xiy[iptr], e13 = xiz[iptr];
        T e21 = etx[iptr], e22 = ety[iptr], e23 = etz[iptr];
        T e31 = ztx[iptr], e32 = zty[iptr], e33 = ztz[iptr];
        T l = lam[iptr], m = mu[iptr];
        T l2m = l + (T)2.0f * m;
        T A[3][3], B[3][3], C[3][3], AB[3][3], AC[3][3];
        A[0][0] = l2m * e31 * e31 + m * (e32 * e32 + e33 * e33);
        A[0][1] = l * e31 * e32 + m * e32 * e31;
        A[0][2] = l * e31 * e33 + m * e33 * e31;
        A[1][0] = l * e32 * e31 + m * e31 * e32;
        A[1][1] = l2m * e32 * e32 + m * (e31 * e31 + e33 * e33);
        A[1][2] = l * e32 * e33 + m * e33 * e32;
        A[2][0] = l * e33 * e31 + m * e31 * e33;
        A[2][1] = l * e33 * e32 + m * e32 * e33;
        A[2][2] = l2m * e33 * e33 + m * (e31 * e31 + e32 * e32);
        invert3x3(A);
        B[0][0] = -l2m * e31 * e11 - m * (e32 * e12 + e33 * e13);
        B[0][1] = -l * e31 * e12 - m * e32 * e11;
        B[0][2] = -l * e31 * e13 - m * e33 * e11;
        B[1][0] = -l * e32 * e11 - m * e31 * e12;
        B[1][1] = -l2m * e32 * e12 - m * (e31 * e11 + e33 * e13);
        B[1][2] = -l * e32 * e13 - m * e33 * e12;
        B[2][0] = -l * e33 * e11 - m * e31 * e13;
        B[2][1] = -l * e33 * e12 - m * e32 * e13;
        B[2][2] = -l2m * e33 * e13 - m * (e31 * e11 + e32 * e12);
        C[0][0] = -l2m * e31 * e21 - m * (e32 * e22 + e33 * e23);
        C[0][1] = -l * e31 * e22 - m * e32 * e21;
        C[0][2] = -l * e31 * e23 - m * e33 * e21;
        C[1][0] = -l * e32 * e21 - m * e31 * e22;
        C[1][1] = -l2m * e32 * e22 - m * (e31 * e21 + e33 * e23);
        C[1][2] = -l * e32 * e23 - m * e33 * e22;
        C[2][0] = -l * e33 * e21 - m * e31 * e23;
        C[2][1] = -l * e33 * e22 - m * e32 * e23;
        C[2][2] = -l2m * e33 * e23 - m * (e31 * e21 + e32 * e22);
        matmul3x3(A, B, AB);
        matmul3x3(A, C, AC);
        size_t ij = ((iy + nj1) * siy + (ix + ni1)) * 9;
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) {
            matVx2Vz[ij + r * 3 + c] = AB[r][c];
            matVy2Vz[ij + r * 3 + c] = AC[r][c];
          }
      }
  }

  /* M = [grad xi, grad eta, grad zeta] as columns (transform.cu:139-147) */
  void metric_at(size_t iptr, T metric[3][3]) const {
    metric[0][0] = xix[iptr]; metric[0][1] = etx[iptr]; metric[0][2] = ztx[iptr];
    metric[1][0] = xiy[iptr]; metric[1][1] = ety[iptr]; metric[1][2] = zty[iptr];
    metric[2][0] = xiz[iptr]; metric[2][1] = etz[iptr]; metric[2][2] = ztz[iptr];
  }
  /* traction = J (sigma M), transform.cu:150-160 (united == 1) */
  static void w2f_node(T stress[3][3], T metric[3][3], T jj, T traction[3][3]) {
    matmul3x3(stress, metric, traction);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) traction[a][b] *= jj;
  }
  /* sigma = (traction M^-1) / J, symmetrised, transform.cu:330-366 (united == 0);
     out = {xx, yy, zz, xy, "Txz" argument, "Tyz" argument} */
  static void f2w_node(T traction[3][3], T metric[3][3], T jacv, T out[6]) {
    T rj = (T)(1.0 / jacv);
    T stress[3][3];
    invert3x3(metric);
    matmul3x3(traction, metric, stress);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) stress[a][b] *= rj;
    out[0] = stress[0][0]; out[1] = stress[1][1]; out[2] = stress[2][2];
    out[3] = (T)((stress[0][1] + stress[1][0]) * 0.5);
    out[4] = (T)((stress[0][2] + stress[2][0]) * 0.5);
    out[5] = (T)((stress[1][2] + stress[2][1]) * 0.5);
  }

  /* ---- wave2fault_gpu, transform.cu:100-212 ---- */
  void wave2fault(T *w_cur, T *w_rhs, T *fcur, T *frhs) {
    T *Vx = w_cur, *Vy = w_cur + sicmp, *Vz = w_cur + 2 * sicmp;
    T *Txx = w_cur + 3 * sicmp, *Tyy = w_cur + 4 * sicmp, *Tzz = w_cur + 5 * sicmp;
    T *Tyz = w_cur + 6 * sicmp, *Txz = w_cur + 7 * sicmp, *Txy = w_cur + 8 * sicmp;
    T *hTxx = w_rhs + 3 * sicmp, *hTyy = w_rhs + 4 * sicmp, *hTzz = w_rhs + 5 * sicmp;
    T *hTyz = w_rhs + 6 * sicmp, *hTxz = w_rhs + 7 * sicmp, *hTxy = w_rhs + 8 * sicmp;
    for (int id = 0; id < nfault; id++) {
      Fault &F = flt[id];
      int i0 = F.i0;
      T *fVx = fcmp(fcur, id, 0), *fVy = fcmp(fcur, id, 1), *fVz = fcmp(fcur, id, 2);
      T *fT2x = fcmp(fcur, id, 3), *fT2y = fcmp(fcur, id, 4), *fT2z = fcmp(fcur, id, 5);
      T *fT3x = fcmp(fcur, id, 6), *fT3y = fcmp(fcur, id, 7), *fT3z = fcmp(fcur, id, 8);
      T *fhT2x = fcmp(frhs, id, 3), *fhT2y = fcmp(frhs, id, 4), *fhT2z = fcmp(frhs, id, 5);
      T *fhT3x = fcmp(frhs, id, 6), *fhT3y = fcmp(frhs, id, 7), *fhT3z = fcmp(frhs, id, 8);
#pragma omp parallel for schedule(static)
      for (int iz = 0; iz < nk; iz++)
        for (int iy = 0; iy < nj; iy++) {
          size_t pf = (iy + nj1) + (iz + nk1) * ny;
          if (F.united[pf] != 1) continue;
          size_t iptr = i0 + (iy + nj1) * siy + (iz + nk1) * siz;
          T metric[3][3], stress[3][3], traction[3][3];
          metric_at(iptr, metric);
          T jj = jac[iptr];
          stress[0][0] = Txx[iptr]; stress[0][1] = Txy[iptr]; stress[0][2] = Txz[iptr];
          stress[1][0] = Txy[iptr]; stress[1][1] = Tyy[iptr]; stress[1][2] = Tyz[iptr];
          stress[2][0] = Txz[iptr]; stress[2][1] = Tyz[iptr]; stress[2][2] = Tzz[iptr];
          w2f_node(stress, metric, jj, traction);
          size_t p3 = pf + 3 * sslice;
          F.T1x[p3] = traction[0][0]; F.T1y[p3] = traction[1][0]; F.T1z[p3] = traction[2][0];
          for (int m = 0; m < 2; m++) {
            size_t p = pf + m * sslice;
            fVx[p] = Vx[iptr]; fVy[p] = Vy[iptr]; fVz[p] = Vz[iptr];
            fT2x[p] = traction[0][1]; fT2y[p] = traction[1][1]; fT2z[p] = traction[2][1];
            fT3x[p] = traction[0][2]; fT3y[p] = traction[1][2]; fT3z[p] = traction[2][2];
          }
          stress[0][0] = hTxx[iptr]; stress[0][1] = hTxy[iptr]; stress[0][2] = hTxz[iptr];
          stress[1][0] = hTxy[iptr]; stress[1][1] = hTyy[iptr]; stress[1][2] = hTyz[iptr];
          stress[2][0] = hTxz[iptr]; stress[2][1] = hTyz[iptr]; stress[2][2] = hTzz[iptr];
          w2f_node(stress, metric, jj, traction);
          for (int m = 0; m < 2; m++) {
            size_t p = pf + m * sslice;
            fhT2x[p] = traction[0][1]; fhT2y[p] = traction[1][1]; fhT2z[p] = traction[2][1];
            fhT3x[p] = traction[0][2]; fhT3y[p] = traction[1][2]; fhT3z[p] = traction[2][2];
          }
        }
    }
  }

  /* ---- fault2wave_gpu, transform.cu:292-376, incl. the Txz/Tyz swap of the
     call at transform.cu:274-276 (reference behaviour unless fix_q1) ---- */
  void fault2wave(T *w, T *fl) {
    T *Vx = w, *Vy = w + sicmp, *Vz = w + 2 * sicmp;
    T *Txx = w + 3 * sicmp, *Tyy = w + 4 * sicmp, *Tzz = w + 5 * sicmp;
    T *Tyz_arr = w + 6 * sicmp, *Txz_arr = w + 7 * sicmp, *Txy = w + 8 * sicmp;
    /* kernel parameter "Txz" receives the Tyz array and vice versa */
    T *kTxz = fix_q1 ? Txz_arr : Tyz_arr;
    T *kTyz = fix_q1 ? Tyz_arr : Txz_arr;
    for (int id = 0; id < nfault; id++) {
      Fault &F = flt[id];
      int i0 = F.i0;
      T *fVx = fcmp(fl, id, 0), *fVy = fcmp(fl, id, 1), *fVz = fcmp(fl, id, 2);
      T *fT2x = fcmp(fl, id, 3), *fT2y = fcmp(fl, id, 4), *fT2z = fcmp(fl, id, 5);
      T *fT3x = fcmp(fl, id, 6), *fT3y = fcmp(fl, id, 7), *fT3z = fcmp(fl, id, 8);
#pragma omp parallel for schedule(static)
      for (int iz = 0; iz < nk; iz++)
        for (int iy = 0; iy < nj; iy++) {
          size_t pf = (iy + nj1) + (iz + nk1) * ny;
          if (F.united[pf] != 0) continue;
          size_t iptr = i0 + (iy + nj1) * siy + (iz + nk1) * siz;
          T metric[3][3], traction[3][3], sv[6];
          metric_at(iptr, metric);
          size_t p3 = pf + 3 * sslice, p1 = pf + sslice;
          traction[0][0] = F.T1x[p3]; traction[1][0] = F.T1y[p3]; traction[2][0] = F.T1z[p3];
          traction[0][1] = (T)((fT2x[pf] + fT2x[p1]) * 0.5);
          traction[1][1] = (T)((fT2y[pf] + fT2y[p1]) * 0.5);
          traction[2][1] = (T)((fT2z[pf] + fT2z[p1]) * 0.5);
          traction[0][2] = (T)((fT3x[pf] + fT3x[p1]) * 0.5);
          traction[1][2] = (T)((fT3y[pf] + fT3y[p1]) * 0.5);
          traction[2][2] = (T)((fT3z[pf] + fT3z[p1]) * 0.5);
          f2w_node(traction, metric, jac[iptr], sv);
          if (plant == 6) {  /* planted: the plus side's velocity instead of the side average */
            Vx[iptr] = fVx[p1]; Vy[iptr] = fVy[p1]; Vz[iptr] = fVz[p1];
          } else {
            Vx[iptr] = (T)((fVx[pf] + fVx[p1]) * 0.5);
            Vy[iptr] = (T)((fVy[pf] + fVy[p1]) * 0.5);
            Vz[iptr] = (T)((fVz[pf] + fVz[p1]) * 0.5);
          }
          Txx[iptr] = sv[0];
          Tyy[iptr] = sv[1];
          Tzz[iptr] = sv[2];
          Txy[iptr] = sv[3];
          kTxz[iptr] = sv[4];
          kTyz[iptr] = sv[5];
        }
    }
  }

  /* ---- trial_slipweakening_gpu, trial_slipweakening.cu:91-388 ---- */
  void trial(T *w_cur, T *fcur, T *fpre, int jdir, int kdir) {
    T *Txx = w_cur + 3 * sicmp, *Tyy = w_cur + 4 * sicmp, *Tzz = w_cur + 5 * sicmp;
    T *Tyz = w_cur + 6 * sicmp, *Txz = w_cur + 7 * sicmp, *Txy = w_cur + 8 * sicmp;
    for (int id = 0; id < nfault; id++) {
      Fault &F = flt[id];
      int i0 = F.i0;
      T *fT2x = fcmp(fcur, id, 3), *fT2y = fcmp(fcur, id, 4), *fT2z = fcmp(fcur, id, 5);
      T *fT3x = fcmp(fcur, id, 6), *fT3y = fcmp(fcur, id, 7), *fT3z = fcmp(fcur, id, 8);
      T *mVx = fcmp(fpre, id, 0), *mVy = fcmp(fpre, id, 1), *mVz = fcmp(fpre, id, 2);
      T *T1x = F.T1x.data(), *T1y = F.T1y.data(), *T1z = F.T1z.data();
#pragma omp parallel for schedule(static)
      for (int iz = 0; iz < nk; iz++)
        for (int iy = 0; iy < nj; iy++) {
          size_t pf = (iy + nj1) + (iz + nk1) * ny;
          if (F.united[pf] != 0) continue;
          T Mrho[2], Rx[2], Ry[2], Rz[2];
          for (int m = 0; m < 2; m++) {
            size_t iptr = i0 + (iy + nj1) * siy + (iz + nk1) * siz;
            T jj = jac[iptr];
            T rho = F.rho_f[pf + m * sslice];
            Mrho[m] = (T)(0.5 * jj * rho);
            for (int l = 1; l <= 3; l++) {
              size_t ip = (i0 + (2 * m - 1) * l) + (iy + nj1) * siy + (iz + nk1) * siz;
              T x1 = xix[ip], x2 = xiy[ip], x3 = xiz[ip], jl = jac[ip];
              T t1x = jl * (x1 * Txx[ip] + x2 * Txy[ip] + x3 * Txz[ip]);
              T t1y = jl * (x1 * Txy[ip] + x2 * Tyy[ip] + x3 * Tyz[ip]);
              T t1z = jl * (x1 * Txz[ip] + x2 * Tyz[ip] + x3 * Tzz[ip]);
              size_t ps = (3 + (2 * m - 1) * l) * sslice + pf;
              T1x[ps] = t1x; T1y[ps] = t1y; T1z[ps] = t1z;
            }
            T *p2x = fT2x + pf + m * sslice, *p2y = fT2y + pf + m * sslice, *p2z = fT2z + pf + m * sslice;
            T DyT2x = 0, DyT2y = 0, DyT2z = 0, DzT3x = 0, DzT3y = 0, DzT3z = 0;
            if (F.rupy[pf] == 1) {
              DyT2x = mac22(p2x, 1, jdir); DyT2y = mac22(p2y, 1, jdir); DyT2z = mac22(p2z, 1, jdir);
            }
            if (F.rupy[pf] == 0) {
              DyT2x = macdrp(p2x, 1, jdir); DyT2y = macdrp(p2y, 1, jdir); DyT2z = macdrp(p2z, 1, jdir);
            }
            T *p3x = fT3x + pf + m * sslice, *p3y = fT3y + pf + m * sslice, *p3z = fT3z + pf + m * sslice;
            if (F.rupz[pf] == 1) {
              DzT3x = mac22(p3x, ny, kdir); DzT3y = mac22(p3y, ny, kdir); DzT3z = mac22(p3z, ny, kdir);
            }
            if (F.rupz[pf] == 0) {
              DzT3x = macdrp(p3x, ny, kdir); DzT3y = macdrp(p3y, ny, kdir); DzT3z = macdrp(p3z, ny, kdir);
            }
            int km = nk - (iz + 1);
            int n_free = km + 3;
            if (isfree == 1 && km <= 3) {
              T v3x[7], v3y[7], v3z[7];
              for (int l = -3; l <= 3; l++) {
                size_t p = (iy + nj1) + (iz + nk1 + l) * ny + m * sslice;
                v3x[l + 3] = fT3x[p]; v3y[l + 3] = fT3y[p]; v3z[l + 3] = fT3z[p];
              }
              v3x[n_free] = 0.0; v3y[n_free] = 0.0; v3z[n_free] = 0.0;
              for (int l = n_free + 1; l < 7; l++) {
                v3x[l] = -v3x[2 * n_free - l]; v3y[l] = -v3y[2 * n_free - l]; v3z[l] = -v3z[2 * n_free - l];
              }
              DzT3x = vecc(v3x + 3, kdir); DzT3y = vecc(v3y + 3, kdir); DzT3z = vecc(v3z + 3, kdir);
            }
            if (m == 0) {
              Rx[m] = (T)(a_1 * T1x[2 * sslice + pf] + a_2 * T1x[1 * sslice + pf] + a_3 * T1x[0 * sslice + pf]);
              Ry[m] = (T)(a_1 * T1y[2 * sslice + pf] + a_2 * T1y[1 * sslice + pf] + a_3 * T1y[0 * sslice + pf]);
              Rz[m] = (T)(a_1 * T1z[2 * sslice + pf] + a_2 * T1z[1 * sslice + pf] + a_3 * T1z[0 * sslice + pf]);
            } else {
              Rx[m] = (T)(a_1 * T1x[4 * sslice + pf] + a_2 * T1x[5 * sslice + pf] + a_3 * T1x[6 * sslice + pf]);
              Ry[m] = (T)(a_1 * T1y[4 * sslice + pf] + a_2 * T1y[5 * sslice + pf] + a_3 * T1y[6 * sslice + pf]);
              Rz[m] = (T)(a_1 * T1z[4 * sslice + pf] + a_2 * T1z[5 * sslice + pf] + a_3 * T1z[6 * sslice + pf]);
            }
            int sg = plant == 2 ? (1 - 2 * m) : (2 * m - 1);  /* planted: wrong side sign */
            Rx[m] = (T)(0.5 * (sg * Rx[m] + (DyT2x + DzT3x)));
            Ry[m] = (T)(0.5 * (sg * Ry[m] + (DyT2y + DzT3y)));
            Rz[m] = (T)(0.5 * (sg * Rz[m] + (DyT2z + DzT3z)));
          }
          size_t p1 = pf + sslice;
          T dVx = mVx[p1] - mVx[pf];
          T dVy = mVy[p1] - mVy[pf];
          T dVz = mVz[p1] - mVz[pf];
          T Trial[3];
          const double two = plant == 1 ? 1.0 : 2.0;  /* planted: a dropped factor */
          Trial[0] = (T)(two * (Mrho[0] * Mrho[1] * dVx / dt + Mrho[0] * Rx[1] - Mrho[1] * Rx[0]) / (a_0 * (Mrho[0] + Mrho[1])));
          Trial[1] = (T)(two * (Mrho[0] * Mrho[1] * dVy / dt + Mrho[0] * Ry[1] - Mrho[1] * Ry[0]) / (a_0 * (Mrho[0] + Mrho[1])));
          Trial[2] = (T)(two * (Mrho[0] * Mrho[1] * dVz / dt + Mrho[0] * Rz[1] - Mrho[1] * Rz[0]) / (a_0 * (Mrho[0] + Mrho[1])));
          T vn[3], vs1[3], vs2[3];
          vs1[0] = F.vec_s1[pf * 3 + 0]; vs1[1] = F.vec_s1[pf * 3 + 1]; vs1[2] = F.vec_s1[pf * 3 + 2];
          vs2[0] = F.vec_s2[pf * 3 + 0]; vs2[1] = F.vec_s2[pf * 3 + 1]; vs2[2] = F.vec_s2[pf * 3 + 2];
          size_t iptr = i0 + (iy + nj1) * siy + (iz + nk1) * siz;
          vn[0] = xix[iptr]; vn[1] = xiy[iptr]; vn[2] = xiz[iptr];
          T vn0 = norm3(vn);
          T jacvec = jac[iptr] * vn0;
          for (int a = 0; a < 3; a++) vn[a] /= vn0;
          T T0v[3] = {F.T0x[pf], F.T0y[pf], F.T0z[pf]};
          T T1c[3], tn, ts1, ts2;
          F.flag_rup[pf] = friction_point<T>(Trial, jacvec, vn, T0v, F.mu_s[pf], F.mu_d[pf], F.Slip[pf], F.Dc[pf],
                                             F.C0[pf], vs1, vs2, T1c, tn, ts1, ts2, &F.kink[4 * pf]);
          size_t p3 = pf + 3 * sslice;
          T1x[p3] = T1c[0]; T1y[p3] = T1c[1]; T1z[p3] = T1c[2];
          F.tTs1[pf] = ts1;
          F.tTs2[pf] = ts2;
          F.tTn[pf] = tn;
        }
    }
  }

  /* ---- sv_curv_col_el_iso_rhs_inner_gpu, sv_curv_col_el_iso_gpu.cu:231-388 ---- */
  void inner(const T *w, T *rhs, int idir, int jdir, int kdir) {
    const T *Vx = w, *Vy = w + sicmp, *Vz = w + 2 * sicmp;
    const T *Txx = w + 3 * sicmp, *Tyy = w + 4 * sicmp, *Tzz = w + 5 * sicmp;
    const T *Tyz = w + 6 * sicmp, *Txz = w + 7 * sicmp, *Txy = w + 8 * sicmp;
    T *hVx = rhs, *hVy = rhs + sicmp, *hVz = rhs + 2 * sicmp;
    T *hTxx = rhs + 3 * sicmp, *hTyy = rhs + 4 * sicmp, *hTzz = rhs + 5 * sicmp;
    T *hTyz = rhs + 6 * sicmp, *hTxz = rhs + 7 * sicmp, *hTxy = rhs + 8 * sicmp;
    long sx[5], sy[5], sz[5];
    T cx[5], cy[5], cz[5];
    for (int n = 0; n < 5; n++) {
      sx[n] = op_indx[idir][n]; cx[n] = coef[idir][n];
      sy[n] = op_indx[jdir][n] * (long)siy; cy[n] = coef[jdir][n];
      sz[n] = op_indx[kdir][n] * (long)siz; cz[n] = coef[kdir][n];
    }
    auto D = [&](const T *p, const long *s, const T *c) -> T {
      return c[0] * p[s[0]] + c[1] * p[s[1]] + c[2] * p[s[2]] + c[3] * p[s[3]] + c[4] * p[s[4]];
    };
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nk; iz++)
      for (int iy = 0; iy < nj; iy++)
        for (int ix = 0; ix < ni; ix++) {
          size_t iptr = (ix + ni1) + (iy + nj1) * siy + (iz + nk1) * siz;
          T DxVx = D(Vx + iptr, sx, cx), DyVx = D(Vx + iptr, sy, cy), DzVx = D(Vx + iptr, sz, cz);
          T DxVy = D(Vy + iptr, sx, cx), DyVy = D(Vy + iptr, sy, cy), DzVy = D(Vy + iptr, sz, cz);
          T DxVz = D(Vz + iptr, sx, cx), DyVz = D(Vz + iptr, sy, cy), DzVz = D(Vz + iptr, sz, cz);
          T DxTxx = D(Txx + iptr, sx, cx), DyTxx = D(Txx + iptr, sy, cy), DzTxx = D(Txx + iptr, sz, cz);
          T DxTyy = D(Tyy + iptr, sx, cx), DyTyy = D(Tyy + iptr, sy, cy), DzTyy = D(Tyy + iptr, sz, cz);
          T DxTzz = D(Tzz + iptr, sx, cx), DyTzz = D(Tzz + iptr, sy, cy), DzTzz = D(Tzz + iptr, sz, cz);
          T DxTxz = D(Txz + iptr, sx, cx), DyTxz = D(Txz + iptr, sy, cy), DzTxz = D(Txz + iptr, sz, cz);
          T DxTyz = D(Tyz + iptr, sx, cx), DyTyz = D(Tyz + iptr, sy, cy), DzTyz = D(Tyz + iptr, sz, cz);
          T DxTxy = D(Txy + iptr, sx, cx), DyTxy = D(Txy + iptr, sy, cy), DzTxy = D(Txy + iptr, sz, cz);
          T xx = xix[iptr], xy = xiy[iptr], xz = xiz[iptr];
          T ex = etx[iptr], ey = ety[iptr], ez = etz[iptr];
          T zx = ztx[iptr], zy = zty[iptr], zz = ztz[iptr];
          T l = lam[iptr], m = mu[iptr], s = slw[iptr];
          T l2m = (T)(l + 2.0 * m);
          hVx[iptr] = s * (xx * DxTxx + xy * DxTxy + xz * DxTxz + ex * DyTxx + ey * DyTxy + ez * DyTxz + zx * DzTxx + zy * DzTxy + zz * DzTxz);
          hVy[iptr] = s * (xx * DxTxy + xy * DxTyy + xz * DxTyz + ex * DyTxy + ey * DyTyy + ez * DyTyz + zx * DzTxy + zy * DzTyy + zz * DzTyz);
          hVz[iptr] = s * (xx * DxTxz + xy * DxTyz + xz * DxTzz + ex * DyTxz + ey * DyTyz + ez * DyTzz + zx * DzTxz + zy * DzTyz + zz * DzTzz);
          hTxx[iptr] = l2m * (xx * DxVx + ex * DyVx + zx * DzVx) + l * (xy * DxVy + ey * DyVy + zy * DzVy + xz * DxVz + ez * DyVz + zz * DzVz);
          hTyy[iptr] = l2m * (xy * DxVy + ey * DyVy + zy * DzVy) + l * (xx * DxVx + ex * DyVx + zx * DzVx + xz * DxVz + ez * DyVz + zz * DzVz);
          hTzz[iptr] = l2m * (xz * DxVz + ez * DyVz + zz * DzVz) + l * (xx * DxVx + ex * DyVx + zx * DzVx + xy * DxVy + ey * DyVy + zy * DzVy);
          hTxy[iptr] = m * (xy * DxVx + xx * DxVy + ey * DyVx + ex * DyVy + zy * DzVx + zx * DzVy);
          hTxz[iptr] = m * (xz * DxVx + xx * DxVz + ez * DyVx + ex * DyVz + zz * DzVx + zx * DzVz);
          hTyz[iptr] = m * (xz * DxVy + xy * DxVz + ez * DyVy + ey * DyVz + zz * DzVy + zy * DzVz);
        }
  }

  /* ---- sv_curv_col_el_iso_rhs_timg_z2_gpu, sv_curv_col_el_iso_gpu.cu:398-600 ---- */
  void timg(const T *w, T *rhs, int idir, int jdir, int kdir) {
    const T *Txx = w + 3 * sicmp, *Tyy = w + 4 * sicmp, *Tzz = w + 5 * sicmp;
    const T *Tyz = w + 6 * sicmp, *Txz = w + 7 * sicmp, *Txy = w + 8 * sicmp;
    T *hV[3] = {rhs, rhs + sicmp, rhs + 2 * sicmp};
    const int *fx = op_indx[idir], *fy = op_indx[jdir], *fz = op_indx[kdir];
    int k_min = nk2 - fz[4];
    /* traction rows: hVx <- (Txx,Txy,Txz), hVy <- (Txy,Tyy,Tyz), hVz <- (Txz,Tyz,Tzz) */
    const T *Ta[3][3] = {{Txx, Txy, Txz}, {Txy, Tyy, Tyz}, {Txz, Tyz, Tzz}};
    for (int k = k_min; k <= nk2; k++) {
      int n_free = nk2 - k - fz[0];
#pragma omp parallel for schedule(static)
      for (int iy = 0; iy < nj; iy++)
        for (int ix = 0; ix < ni; ix++) {
          size_t iptr = (ix + ni1) + (iy + nj1) * siy + k * siz;
          T slwjac = slw[iptr] / jac[iptr];
          for (int c = 0; c < 3; c++) {
            const T *A = Ta[c][0], *B = Ta[c][1], *C = Ta[c][2];
            T vecxi[5] = {0, 0, 0, 0, 0}, vecet[5] = {0, 0, 0, 0, 0}, veczt[5] = {0, 0, 0, 0, 0};
            for (int n = 0; n < 5; n++) {
              size_t p = iptr + fx[n];
              vecxi[n] = jac[p] * (xix[p] * A[p] + xiy[p] * B[p] + xiz[p] * C[p]);
            }
            for (int n = 0; n < 5; n++) {
              size_t p = iptr + fy[n] * (long)siy;
              vecet[n] = jac[p] * (etx[p] * A[p] + ety[p] * B[p] + etz[p] * C[p]);
            }
            for (int n = 0; n < n_free; n++) {
              size_t p = iptr + fz[n] * (long)siz;
              veczt[n] = jac[p] * (ztx[p] * A[p] + zty[p] * B[p] + ztz[p] * C[p]);
            }
            veczt[n_free] = 0.0;
            for (int n = n_free + 1; n < 5; n++) {
              int n_img = fz[n] - 2 * (n - n_free);
              size_t p = iptr + n_img * (long)siz;
              veczt[n] = -jac[p] * (ztx[p] * A[p] + zty[p] * B[p] + ztz[p] * C[p]);
            }
            T DxTx = vecdrp(vecxi, idir), DyTy = vecdrp(vecet, jdir), DzTz = vecdrp(veczt, kdir);
            hV[c][iptr] = (DxTx + DyTy + DzTz) * slwjac;
          }
        }
    }
  }

  /* ---- sv_curv_col_el_iso_rhs_vlow_z2_gpu, sv_curv_col_el_iso_gpu.cu:606-747 ---- */
  void vlow(const T *w, T *rhs, int idir, int jdir, int kdir) {
    const T *Vx = w, *Vy = w + sicmp, *Vz = w + 2 * sicmp;
    T *hTxx = rhs + 3 * sicmp, *hTyy = rhs + 4 * sicmp, *hTzz = rhs + 5 * sicmp;
    T *hTyz = rhs + 6 * sicmp, *hTxz = rhs + 7 * sicmp, *hTxy = rhs + 8 * sicmp;
    for (int n = 0; n < 3; n++) {
      int k = nk2 - n;
#pragma omp parallel for schedule(static)
      for (int iy = 0; iy < nj; iy++)
        for (int ix = 0; ix < ni; ix++) {
          size_t iptr = (ix + ni1) + (iy + nj1) * siy + k * siz;
          T xx = xix[iptr], xy = xiy[iptr], xz = xiz[iptr];
          T ex = etx[iptr], ey = ety[iptr], ez = etz[iptr];
          T zx = ztx[iptr], zy = zty[iptr], zz = ztz[iptr];
          T l = lam[iptr], m = mu[iptr];
          T l2m = (T)(l + 2.0 * m);
          const T *px = Vx + iptr, *py = Vy + iptr, *pz = Vz + iptr;
          T DxVx = macdrp(px, 1, idir), DyVx = macdrp(px, siy, jdir);
          T DxVy = macdrp(py, 1, idir), DyVy = macdrp(py, siy, jdir);
          T DxVz = macdrp(pz, 1, idir), DyVz = macdrp(pz, siy, jdir);
          T DzVx = 0, DzVy = 0, DzVz = 0;
          if (k == nk2) {
            size_t ij = ((ix + ni1) + (iy + nj1) * siy) * 9;
            const T *X = &matVx2Vz[ij], *Y = &matVy2Vz[ij];
            DzVx = X[0] * DxVx + X[1] * DxVy + X[2] * DxVz + Y[0] * DyVx + Y[1] * DyVy + Y[2] * DyVz;
            DzVy = X[3] * DxVx + X[4] * DxVy + X[5] * DxVz + Y[3] * DyVx + Y[4] * DyVy + Y[5] * DyVz;
            DzVz = X[6] * DxVx + X[7] * DxVy + X[8] * DxVz + Y[6] * DyVx + Y[7] * DyVy + Y[8] * DyVz;
          }
          if (k == nk2 - 1) {
            DzVx = mac22(px, siz, kdir); DzVy = mac22(py, siz, kdir); DzVz = mac22(pz, siz, kdir);
          }
          if (k == nk2 - 2) {
            DzVx = mac24(px, siz, kdir); DzVy = mac24(py, siz, kdir); DzVz = mac24(pz, siz, kdir);
          }
          hTxx[iptr] = l2m * (xx * DxVx + ex * DyVx + zx * DzVx) + l * (xy * DxVy + ey * DyVy + zy * DzVy + xz * DxVz + ez * DyVz + zz * DzVz);
          hTyy[iptr] = l2m * (xy * DxVy + ey * DyVy + zy * DzVy) + l * (xx * DxVx + ex * DyVx + zx * DzVx + xz * DxVz + ez * DyVz + zz * DzVz);
          hTzz[iptr] = l2m * (xz * DxVz + ez * DyVz + zz * DzVz) + l * (xx * DxVx + ex * DyVx + zx * DzVx + xy * DxVy + ey * DyVy + zy * DzVy);
          hTxy[iptr] = m * (xy * DxVx + xx * DxVy + ey * DyVx + ex * DyVy + zy * DzVx + zx * DzVy);
          hTxz[iptr] = m * (xz * DxVx + xx * DxVz + ez * DyVx + ex * DyVz + zz * DzVx + zx * DzVz);
          hTyz[iptr] = m * (xz * DxVy + xy * DxVz + ez * DyVy + ey * DyVz + zz * DzVy + zy * DzVz);
        }
    }
  }

  /* ---- sv_curv_col_el_iso_rhs_cfspml_gpu, sv_curv_col_el_iso_gpu.cu:757-1290 ---- */
  void cfspml(const T *w, T *rhs, int idir, int jdir, int kdir) {
    const T *Vx = w, *Vy = w + sicmp, *Vz = w + 2 * sicmp;
    const T *Txx = w + 3 * sicmp, *Tyy = w + 4 * sicmp, *Tzz = w + 5 * sicmp;
    const T *Tyz = w + 6 * sicmp, *Txz = w + 7 * sicmp, *Txy = w + 8 * sicmp;
    T *hVx = rhs, *hVy = rhs + sicmp, *hVz = rhs + 2 * sicmp;
    T *hTxx = rhs + 3 * sicmp, *hTyy = rhs + 4 * sicmp, *hTzz = rhs + 5 * sicmp;
    T *hTyz = rhs + 6 * sicmp, *hTxz = rhs + 7 * sicmp, *hTxy = rhs + 8 * sicmp;
    for (int dim = 0; dim < 3; dim++)
      for (int side = 0; side < 2; side++) {
        Face &f = pml[dim][side];
        if (!f.on) continue;
        int dir = dim == 0 ? idir : (dim == 1 ? jdir : kdir);
        long stride = dim == 0 ? 1 : (dim == 1 ? (long)siy : (long)siz);
        long sh[5];
        T cf[5];
        for (int n = 0; n < 5; n++) { sh[n] = op_indx[dir][n] * stride; cf[n] = coef[dir][n]; }
        auto D = [&](const T *p) -> T {
          return cf[0] * p[sh[0]] + cf[1] * p[sh[1]] + cf[2] * p[sh[2]] + cf[3] * p[sh[3]] + cf[4] * p[sh[4]];
        };
        /* aux var pointers (cur and rhs levels); component order = wav_t */
        T *pc = f.cur, *pr = f.rhs;
        size_t N = f.n;
        const T *mx = dim == 0 ? xix : (dim == 1 ? etx : ztx);
        const T *my = dim == 0 ? xiy : (dim == 1 ? ety : zty);
        const T *mz = dim == 0 ? xiz : (dim == 1 ? etz : ztz);
        const T *mat = dim == 0 ? matVx2Vz.data() : matVy2Vz.data();
#pragma omp parallel for schedule(static)
        for (int iz = 0; iz < f.cn; iz++)
          for (int iy = 0; iy < f.bn; iy++)
            for (int ix = 0; ix < f.an; ix++) {
              size_t ia = (size_t)iz * (f.bn * f.an) + iy * f.an + ix;
              size_t iptr = (ix + f.i1) + (iy + f.j1) * siy + (iz + f.k1) * siz;
              int ic = dim == 0 ? ix : (dim == 1 ? iy : iz);
              T cD = f.D[ic], cA = f.A[ic], cB = f.B[ic];
              T cB1 = (T)(cB - 1.0);
              T e1 = mx[iptr], e2 = my[iptr], e3 = mz[iptr];
              T l = lam[iptr], m = mu[iptr], s = slw[iptr];
              T l2m = (T)(l + 2.0 * m);
              T DVx = D(Vx + iptr), DVy = D(Vy + iptr), DVz = D(Vz + iptr);
              T DTxx = D(Txx + iptr), DTyy = D(Tyy + iptr), DTzz = D(Tzz + iptr);
              T DTxz = D(Txz + iptr), DTyz = D(Tyz + iptr), DTxy = D(Txy + iptr);
              T rVx = s * (e1 * DTxx + e2 * DTxy + e3 * DTxz);
              T rVy = s * (e1 * DTxy + e2 * DTyy + e3 * DTyz);
              T rVz = s * (e1 * DTxz + e2 * DTyz + e3 * DTzz);
              T rTxx = l2m * e1 * DVx + l * e2 * DVy + l * e3 * DVz;
              T rTyy = l * e1 * DVx + l2m * e2 * DVy + l * e3 * DVz;
              T rTzz = l * e1 * DVx + l * e2 * DVy + l2m * e3 * DVz;
              T rTxy = m * (e2 * DVx + e1 * DVy);
              T rTxz = m * (e3 * DVx + e1 * DVz);
              T rTyz = m * (e3 * DVy + e2 * DVz);
              /* aux component positions: Vx0 Vy1 Vz2 Txx3 Tyy4 Tzz5 Tyz6 Txz7 Txy8 */
              T *aVx = pc, *aVy = pc + N, *aVz = pc + 2 * N, *aTxx = pc + 3 * N, *aTyy = pc + 4 * N;
              T *aTzz = pc + 5 * N, *aTyz = pc + 6 * N, *aTxz = pc + 7 * N, *aTxy = pc + 8 * N;
              T *rhVx = pr, *rhVy = pr + N, *rhVz = pr + 2 * N, *rhTxx = pr + 3 * N, *rhTyy = pr + 4 * N;
              T *rhTzz = pr + 5 * N, *rhTyz = pr + 6 * N, *rhTxz = pr + 7 * N, *rhTxy = pr + 8 * N;
              hVx[iptr] += cB1 * rVx - cB * aVx[ia];
              hVy[iptr] += cB1 * rVy - cB * aVy[ia];
              hVz[iptr] += cB1 * rVz - cB * aVz[ia];
              hTxx[iptr] += cB1 * rTxx - cB * aTxx[ia];
              hTyy[iptr] += cB1 * rTyy - cB * aTyy[ia];
              hTzz[iptr] += cB1 * rTzz - cB * aTzz[ia];
              hTxz[iptr] += cB1 * rTxz - cB * aTxz[ia];
              hTyz[iptr] += cB1 * rTyz - cB * aTyz[ia];
              hTxy[iptr] += cB1 * rTxy - cB * aTxy[ia];
              rhVx[ia] = cD * rVx - cA * aVx[ia];
              rhVy[ia] = cD * rVy - cA * aVy[ia];
              rhVz[ia] = cD * rVz - cA * aVz[ia];
              rhTxx[ia] = cD * rTxx - cA * aTxx[ia];
              rhTyy[ia] = cD * rTyy - cA * aTyy[ia];
              rhTzz[ia] = cD * rTzz - cA * aTzz[ia];
              rhTxz[ia] = cD * rTxz - cA * aTxz[ia];
              rhTyz[ia] = cD * rTyz - cA * aTyz[ia];
              rhTxy[ia] = cD * rTxy - cA * aTxy[ia];
              if (dim < 2 && isfree == 1 && (int)(iz + f.k1) == nk2) {
                size_t ij = ((ix + f.i1) + (iy + f.j1) * siy) * 9;
                const T *X = mat + ij;
                T DzVx_ = X[0] * DVx + X[1] * DVy + X[2] * DVz;
                T DzVy_ = X[3] * DVx + X[4] * DVy + X[5] * DVz;
                T DzVz_ = X[6] * DVx + X[7] * DVy + X[8] * DVz;
                T zx = ztx[iptr], zy = zty[iptr], zz = ztz[iptr];
                T qTxx = l2m * (zx * DzVx_) + l * (zy * DzVy_ + zz * DzVz_);
                T qTyy = l2m * (zy * DzVy_) + l * (zx * DzVx_ + zz * DzVz_);
                T qTzz = l2m * (zz * DzVz_) + l * (zx * DzVx_ + zy * DzVy_);
                T qTxy = m * (zy * DzVx_ + zx * DzVy_);
                T qTxz = m * (zz * DzVx_ + zx * DzVz_);
                T qTyz = m * (zz * DzVy_ + zy * DzVz_);
                hTxx[iptr] += (cB - 1.0) * qTxx;
                hTyy[iptr] += (cB - 1.0) * qTyy;
                hTzz[iptr] += (cB - 1.0) * qTzz;
                hTxz[iptr] += (cB - 1.0) * qTxz;
                hTyz[iptr] += (cB - 1.0) * qTyz;
                hTxy[iptr] += (cB - 1.0) * qTxy;
                rhTxx[ia] += cD * qTxx;
                rhTyy[ia] += cD * qTyy;
                rhTzz[ia] += cD * qTzz;
                rhTxz[ia] += cD * qTxz;
                rhTyz[ia] += cD * qTyz;
                rhTxy[ia] += cD * qTxy;
              }
            }
      }
  }

  /* ---- sv_curv_col_el_iso_rhs_fault_velo_gpu, sv_curv_col_el_iso_fault_gpu.cu:191-429 ---- */
  void fault_velo(const T *w, T *rhs, T *fcur, T *frhs, int idir, int jdir, int kdir) {
    const T *Txx = w + 3 * sicmp, *Tyy = w + 4 * sicmp, *Tzz = w + 5 * sicmp;
    const T *Tyz = w + 6 * sicmp, *Txz = w + 7 * sicmp, *Txy = w + 8 * sicmp;
    T *hVx = rhs, *hVy = rhs + sicmp, *hVz = rhs + 2 * sicmp;
    for (int id = 0; id < nfault; id++) {
      Fault &F = flt[id];
      int i0 = F.i0;
      T *fT2[3] = {fcmp(fcur, id, 3), fcmp(fcur, id, 4), fcmp(fcur, id, 5)};
      T *fT3[3] = {fcmp(fcur, id, 6), fcmp(fcur, id, 7), fcmp(fcur, id, 8)};
      T *fhV[3] = {fcmp(frhs, id, 0), fcmp(frhs, id, 1), fcmp(frhs, id, 2)};
      T *T1[3] = {F.T1x.data(), F.T1y.data(), F.T1z.data()};
#pragma omp parallel for schedule(static)
      for (int iz = 0; iz < nk; iz++)
        for (int iy = 0; iy < nj; iy++) {
          size_t pf = (iy + nj1) + (iz + nk1) * ny;
          if (F.united[pf] != 0) continue;
          int km = nk - (iz + 1);
          int n_free = km + 3;
          for (int i = i0 - 3; i <= i0 + 3; i++) {
            int n = i0 - i;
            if (n == 0) continue;
            T v1[3][7], v2[3][7], v3[3][7];
            for (int l = -3; l <= 3; l++) {
              size_t p = (i + l) + (iy + nj1) * siy + (iz + nk1) * siz;
              v1[0][l + 3] = jac[p] * (xix[p] * Txx[p] + xiy[p] * Txy[p] + xiz[p] * Txz[p]);
              v1[1][l + 3] = jac[p] * (xix[p] * Txy[p] + xiy[p] * Tyy[p] + xiz[p] * Tyz[p]);
              v1[2][l + 3] = jac[p] * (xix[p] * Txz[p] + xiy[p] * Tyz[p] + xiz[p] * Tzz[p]);
              p = i + (iy + nj1 + l) * siy + (iz + nk1) * siz;
              v2[0][l + 3] = jac[p] * (etx[p] * Txx[p] + ety[p] * Txy[p] + etz[p] * Txz[p]);
              v2[1][l + 3] = jac[p] * (etx[p] * Txy[p] + ety[p] * Tyy[p] + etz[p] * Tyz[p]);
              v2[2][l + 3] = jac[p] * (etx[p] * Txz[p] + ety[p] * Tyz[p] + etz[p] * Tzz[p]);
              p = i + (iy + nj1) * siy + (iz + nk1 + l) * siz;
              v3[0][l + 3] = jac[p] * (ztx[p] * Txx[p] + zty[p] * Txy[p] + ztz[p] * Txz[p]);
              v3[1][l + 3] = jac[p] * (ztx[p] * Txy[p] + zty[p] * Tyy[p] + ztz[p] * Tyz[p]);
              v3[2][l + 3] = jac[p] * (ztx[p] * Txz[p] + zty[p] * Tyz[p] + ztz[p] * Tzz[p]);
            }
            size_t p3 = pf + 3 * sslice;
            for (int c = 0; c < 3; c++) v1[c][n + 3] = T1[c][p3];
            for (int c = 0; c < 3; c++) {
              T *v = v1[c];
              if (n == 2) { v[6] = (T)(2.0 * v[5] - v[4]); }
              if (n == 1) { v[5] = (T)(2.0 * v[4] - v[3]); v[6] = (T)(2.0 * v[4] - v[2]); }
              if (n == -1) { v[0] = (T)(2.0 * v[2] - v[4]); v[1] = (T)(2.0 * v[2] - v[3]); }
              if (n == -2) { v[0] = (T)(2.0 * v[1] - v[2]); }
            }
            if (isfree == 1 && km <= 3) {
              for (int c = 0; c < 3; c++) {
                v3[c][n_free] = 0.0;
                for (int l = n_free + 1; l < 7; l++) v3[c][l] = -v3[c][2 * n_free - l];
              }
            }
            size_t iptr = i + (iy + nj1) * siy + (iz + nk1) * siz;
            T rrhojac = slw[iptr] / jac[iptr];
            T *hV[3] = {hVx, hVy, hVz};
            for (int c = 0; c < 3; c++) {
              T DxT1 = vecc(v1[c] + 3, idir), DyT2 = vecc(v2[c] + 3, jdir), DzT3 = vecc(v3[c] + 3, kdir);
              hV[c][iptr] = (DxT1 + DyT2 + DzT3) * rrhojac;
            }
          }
          for (int m = 0; m < 2; m++) {
            T DxT1[3], DyT2[3], DzT3[3];
            for (int c = 0; c < 3; c++) {
              const T *t = T1[c];
              const double a0p = plant == 3 ? 0.0 : a_0;  /* planted: a dropped term */
              if (m == 0)
                DxT1[c] = (T)(a0p * t[pf + 3 * sslice] - a_1 * t[pf + 2 * sslice] - a_2 * t[pf + 1 * sslice] - a_3 * t[pf + 0 * sslice]);
              else
                DxT1[c] = (T)(-a0p * t[pf + 3 * sslice] + a_1 * t[pf + 4 * sslice] + a_2 * t[pf + 5 * sslice] + a_3 * t[pf + 6 * sslice]);
              const T *p2 = fT2[c] + pf + m * sslice;
              if (F.rupy[pf] == 1) DyT2[c] = mac22(p2, 1, jdir);
              if (F.rupy[pf] == 0) DyT2[c] = macdrp(p2, 1, jdir);
              const T *p3 = fT3[c] + pf + m * sslice;
              if (F.rupz[pf] == 1) DzT3[c] = mac22(p3, ny, kdir);
              if (F.rupz[pf] == 0) DzT3[c] = macdrp(p3, ny, kdir);
            }
            if (isfree == 1 && km <= 3) {
              for (int c = 0; c < 3; c++) {
                T v[7];
                for (int l = -3; l <= 3; l++) v[l + 3] = fT3[c][(iy + nj1) + (iz + nk1 + l) * ny + m * sslice];
                v[n_free] = 0.0;
                for (int l = n_free + 1; l < 7; l++) v[l] = -v[2 * n_free - l];
                DzT3[c] = vecc(v + 3, kdir);
              }
            }
            size_t iptr = i0 + (iy + nj1) * siy + (iz + nk1) * siz;
            size_t pm = pf + m * sslice;
            T rrhojac = (T)(1.0 / (F.rho_f[pm] * jac[iptr]));
            for (int c = 0; c < 3; c++) fhV[c][pm] = (DxT1[c] + DyT2[c] + DzT3[c]) * rrhojac;
          }
        }
    }
  }

  /* Hooke's law on a fault column point: sv_curv_col_el_iso_fault_gpu.cu:1125-1154 */
  void hooke_point(size_t iptr, const T *DxV, const T *DyV, const T *DzV, T *rhs) {
    T l = lam[iptr], m = mu[iptr];
    T l2m = (T)(l + 2.0 * m);
    T xx = xix[iptr], xy = xiy[iptr], xz = xiz[iptr];
    T ex = etx[iptr], ey = ety[iptr], ez = etz[iptr];
    T zx = ztx[iptr], zy = zty[iptr], zz = ztz[iptr];
    T *hTxx = rhs + 3 * sicmp, *hTyy = rhs + 4 * sicmp, *hTzz = rhs + 5 * sicmp;
    T *hTyz = rhs + 6 * sicmp, *hTxz = rhs + 7 * sicmp, *hTxy = rhs + 8 * sicmp;
    hTxx[iptr] = l2m * (xx * DxV[0] + ex * DyV[0] + zx * DzV[0]) + l * (xy * DxV[1] + ey * DyV[1] + zy * DzV[1] + xz * DxV[2] + ez * DyV[2] + zz * DzV[2]);
    hTyy[iptr] = l2m * (xy * DxV[1] + ey * DyV[1] + zy * DzV[1]) + l * (xx * DxV[0] + ex * DyV[0] + zx * DzV[0] + xz * DxV[2] + ez * DyV[2] + zz * DzV[2]);
    hTzz[iptr] = l2m * (xz * DxV[2] + ez * DyV[2] + zz * DzV[2]) + l * (xx * DxV[0] + ex * DyV[0] + zx * DzV[0] + xy * DxV[1] + ey * DyV[1] + zy * DzV[1]);
    hTxy[iptr] = m * (xy * DxV[0] + xx * DxV[1] + ey * DyV[0] + ex * DyV[1] + zy * DzV[0] + zx * DzV[1]);
    hTxz[iptr] = m * (xz * DxV[0] + xx * DxV[2] + ez * DyV[0] + ex * DyV[2] + zz * DzV[0] + zx * DzV[2]);
    hTyz[iptr] = m * (xz * DxV[1] + xy * DxV[2] + ez * DyV[1] + ey * DyV[2] + zz * DzV[1] + zy * DzV[2]);
  }

  /* ---- sv_curv_col_el_iso_rhs_fault_stress_F_gpu (:431-1159) and _B_gpu
     (:1161-1879): DxV/DyV/DzV[8][3] indexed by vec grid index 0..7 and
     component; written out per variant, as the reference does ---- */
  void fault_stress(const T *w, T *rhs, T *fcur, T *frhs, int idir, int jdir, int kdir) {
    const T *V[3] = {w, w + sicmp, w + 2 * sicmp};
    for (int id = 0; id < nfault; id++) {
      Fault &F = flt[id];
      int i0 = F.i0;
      T *fV[3] = {fcmp(fcur, id, 0), fcmp(fcur, id, 1), fcmp(fcur, id, 2)};
      T *fhT2[3] = {fcmp(frhs, id, 3), fcmp(frhs, id, 4), fcmp(frhs, id, 5)};
      T *fhT3[3] = {fcmp(frhs, id, 6), fcmp(frhs, id, 7), fcmp(frhs, id, 8)};
#pragma omp parallel for schedule(static)
      for (int iz = 0; iz < nk; iz++)
        for (int iy = 0; iy < nj; iy++) {
          size_t pf = (iy + nj1) + (iz + nk1) * ny;
          if (F.united[pf] != 0) continue;
          int km = nk - (iz + 1);
          size_t row = (iy + nj1) * siy + (iz + nk1) * siz;
          T DxV[8][3], DyV[8][3], DzV[8][3];
          for (int a = 0; a < 8; a++)
            for (int c = 0; c < 3; c++) DxV[a][c] = DyV[a][c] = DzV[a][c] = 0;
          size_t pm = pf, pp = pf + sslice;  /* minus / plus side */
          if (idir == 1) {
            /* points A (i0+3), B (i0+2), C (i0+1): :518-543 */
            for (int c = 0; c < 3; c++) {
              DxV[7][c] = macdrp_f(V[c] + (i0 + 3) + row, 1);
              DxV[6][c] = mac24_f(V[c] + (i0 + 2) + row, 1);
              DxV[5][c] = mac22_f(V[c] + (i0 + 1) + row, 1);
            }
          } else {
            /* points G (i0-3), F (i0-2), E (i0-1): :1248-1274 */
            for (int c = 0; c < 3; c++) {
              DxV[0][c] = macdrp_b(V[c] + (i0 - 3) + row, 1);
              DxV[1][c] = mac24_b(V[c] + (i0 - 2) + row, 1);
              DxV[2][c] = mac22_b(V[c] + (i0 - 1) + row, 1);
            }
          }
          /* split nodes D-, D+: y and z derivatives along the fault plane */
          for (int c = 0; c < 3; c++) {
            if (F.rupy[pf] == 1) { DyV[4][c] = mac22(fV[c] + pp, 1, jdir); DyV[3][c] = mac22(fV[c] + pm, 1, jdir); }
            if (F.rupy[pf] == 0) { DyV[4][c] = macdrp(fV[c] + pp, 1, jdir); DyV[3][c] = macdrp(fV[c] + pm, 1, jdir); }
            if (F.rupz[pf] == 1) { DzV[4][c] = mac22(fV[c] + pp, ny, kdir); DzV[3][c] = mac22(fV[c] + pm, ny, kdir); }
            if (F.rupz[pf] == 0) { DzV[4][c] = macdrp(fV[c] + pp, ny, kdir); DzV[3][c] = macdrp(fV[c] + pm, ny, kdir); }
          }
          size_t idx = pf * 9;
          T P2M[5][3][3], M2P[5][3][3], T1m[3][3], Vym[3][3], Vzm[3][3], T1p[3][3], Vyp[3][3], Vzp[3][3];
          for (int q = 0; q < 5; q++) { load9(P2M[q], &F.P2M[q][idx]); load9(M2P[q], &F.M2P[q][idx]); }
          load9(T1m, &F.M2[0][idx]); load9(Vym, &F.M2[1][idx]); load9(Vzm, &F.M2[2][idx]);
          load9(T1p, &F.M2[3][idx]); load9(Vyp, &F.M2[4][idx]); load9(Vzp, &F.M2[5][idx]);
          T dtT1[3];
          auto get_dtT1 = [&]() {
            if (F.faultgrid[pf] == 1 && plant != 4) { dtT1[0] = F.hT1x[pf]; dtT1[1] = F.hT1y[pf]; dtT1[2] = F.hT1z[pf]; }
            else { dtT1[0] = 0.0; dtT1[1] = 0.0; dtT1[2] = 0.0; }
          };
          T dxV1[3], dyV1[3], dzV1[3], dxV2[3], dyV2[3], dzV2[3];
          T o1[3], o2[3], o3[3], o4[3], o5[3];
          auto setv = [&](T *dst, T (*src)[3], int a) { dst[0] = src[a][0]; dst[1] = src[a][1]; dst[2] = src[a][2]; };
          /* imethod 2 on both sides with full matrices (:662-707 / :1382-1421) */
          auto method2_full = [&]() {
            get_dtT1();
            setv(dyV2, DyV, 4); setv(dzV2, DzV, 4);
            matmul3x1(T1p, dtT1, o1); matmul3x1(Vyp, dyV2, o2); matmul3x1(Vzp, dzV2, o3);
            for (int c = 0; c < 3; c++) DxV[4][c] = o1[c] - o2[c] - o3[c];
            setv(dyV1, DyV, 3); setv(dzV1, DzV, 3);
            matmul3x1(T1m, dtT1, o1); matmul3x1(Vym, dyV1, o2); matmul3x1(Vzm, dzV1, o3);
            for (int c = 0; c < 3; c++) DxV[3][c] = o1[c] - o2[c] - o3[c];
          };
          if (idir == 1) {
            if (imethod == 1) {  /* :636-661 */
              for (int c = 0; c < 3; c++) DxV[4][c] = V[c][(i0 + 1) + row] - fV[c][pp];
              setv(dyV1, DyV, 3); setv(dzV1, DzV, 3);
              setv(dxV2, DxV, 4); setv(dyV2, DyV, 4); setv(dzV2, DzV, 4);
              matmul3x1(P2M[0], dxV2, o1); matmul3x1(P2M[1], dyV2, o2); matmul3x1(P2M[2], dzV2, o3);
              matmul3x1(P2M[3], dyV1, o4); matmul3x1(P2M[4], dzV1, o5);
              if (plant == 5) for (int c = 0; c < 3; c++) o1[c] = 0;  /* planted: a dropped term */
              for (int c = 0; c < 3; c++) DxV[3][c] = o1[c] + o2[c] + o3[c] - o4[c] - o5[c];
            }
            if (imethod == 2) method2_full();
          } else {
            if (imethod == 1) {  /* :1362-1381 */
              for (int c = 0; c < 3; c++) DxV[3][c] = fV[c][pm] - V[c][(i0 - 1) + row];
              setv(dxV1, DxV, 3); setv(dyV1, DyV, 3); setv(dzV1, DzV, 3);
              setv(dyV2, DyV, 4); setv(dzV2, DzV, 4);
              matmul3x1(M2P[0], dxV1, o1); matmul3x1(M2P[1], dyV1, o2); matmul3x1(M2P[2], dzV1, o3);
              matmul3x1(M2P[3], dyV2, o4); matmul3x1(M2P[4], dzV2, o5);
              if (plant == 5) for (int c = 0; c < 3; c++) o1[c] = 0;
              for (int c = 0; c < 3; c++) DxV[4][c] = o1[c] + o2[c] + o3[c] - o4[c] - o5[c];
            }
            if (imethod == 2) method2_full();
          }
          /* free surface at the split nodes */
          if (isfree == 1 && km == 0) {
            size_t jf = (size_t)(iy + 3) * 9;
            T Vx2Vz1[3][3], Vy2Vz1[3][3], Vx2Vz2[3][3], Vy2Vz2[3][3];
            T P2Mf[3][3][3], M2Pf[3][3][3], T1fm[3][3], Vyfm[3][3], T1fp[3][3], Vyfp[3][3];
            load9(Vx2Vz1, &F.F[0][jf]); load9(Vy2Vz1, &F.F[1][jf]);
            load9(Vx2Vz2, &F.F[2][jf]); load9(Vy2Vz2, &F.F[3][jf]);
            for (int q = 0; q < 3; q++) { load9(P2Mf[q], &F.F[4 + q][jf]); load9(M2Pf[q], &F.F[7 + q][jf]); }
            load9(T1fm, &F.F[10][jf]); load9(Vyfm, &F.F[11][jf]);
            load9(T1fp, &F.F[12][jf]); load9(Vyfp, &F.F[13][jf]);
            if (idir == 1) {  /* :729-812 */
              if (imethod == 1) {
                setv(dyV1, DyV, 3); setv(dxV2, DxV, 4); setv(dyV2, DyV, 4);
                matmul3x1(P2Mf[0], dxV2, o1); matmul3x1(P2Mf[1], dyV2, o2); matmul3x1(P2Mf[2], dyV1, o3);
                for (int c = 0; c < 3; c++) DxV[3][c] = o1[c] + o2[c] - o3[c];
              }
              if (imethod == 2) {
                get_dtT1();
                setv(dyV2, DyV, 4);
                matmul3x1(T1fp, dtT1, o1); matmul3x1(Vyfp, dyV2, o2);
                for (int c = 0; c < 3; c++) DxV[4][c] = o1[c] - o2[c];
                setv(dyV1, DyV, 3);
                matmul3x1(T1fm, dtT1, o1); matmul3x1(Vyfm, dyV1, o2);
                for (int c = 0; c < 3; c++) DxV[3][c] = o1[c] - o2[c];
              }
              setv(dxV2, DxV, 4); setv(dyV2, DyV, 4);
              matmul3x1(Vx2Vz2, dxV2, o1); matmul3x1(Vy2Vz2, dyV2, o2);
              for (int c = 0; c < 3; c++) DzV[4][c] = o1[c] + o2[c];
              setv(dxV1, DxV, 3); setv(dyV1, DyV, 3);
              matmul3x1(Vx2Vz1, dxV1, o1); matmul3x1(Vy2Vz1, dyV1, o2);
              for (int c = 0; c < 3; c++) DzV[3][c] = o1[c] + o2[c];
            } else {  /* :1457-1517 */
              if (imethod == 1) {
                setv(dxV1, DxV, 3); setv(dyV1, DyV, 3); setv(dyV2, DyV, 4);
                matmul3x1(M2Pf[0], dxV1, o1); matmul3x1(M2Pf[1], dyV1, o2); matmul3x1(M2Pf[2], dyV2, o3);
                for (int c = 0; c < 3; c++) DxV[4][c] = o1[c] + o2[c] - o3[c];
              }
              if (imethod == 2) {
                get_dtT1();
                setv(dyV1, DyV, 3);
                matmul3x1(T1fm, dtT1, o1); matmul3x1(Vyfm, dyV1, o2);
                for (int c = 0; c < 3; c++) DxV[3][c] = o1[c] - o2[c];
                setv(dyV2, DyV, 4);
                matmul3x1(T1fp, dtT1, o1); matmul3x1(Vyfp, dyV2, o2);
                for (int c = 0; c < 3; c++) DxV[4][c] = o1[c] - o2[c];
              }
              setv(dxV1, DxV, 3);
              matmul3x1(Vx2Vz1, dxV1, o1); matmul3x1(Vy2Vz1, dyV1, o2);
              for (int c = 0; c < 3; c++) DzV[3][c] = o1[c] + o2[c];
              setv(dxV2, DxV, 4);
              matmul3x1(Vx2Vz2, dxV2, o1); matmul3x1(Vy2Vz2, dyV2, o2);
              for (int c = 0; c < 3; c++) DzV[4][c] = o1[c] + o2[c];
            }
          }
          if (isfree == 1 && (km == 1 || km == 2)) {  /* :815-934 / :1519-1625 */
            for (int c = 0; c < 3; c++) {
              if (km == 1) { DzV[4][c] = mac22(fV[c] + pp, ny, kdir); DzV[3][c] = mac22(fV[c] + pm, ny, kdir); }
              if (km == 2) { DzV[4][c] = mac24(fV[c] + pp, ny, kdir); DzV[3][c] = mac24(fV[c] + pm, ny, kdir); }
            }
            if (idir == 1) {
              if (imethod == 1) {
                setv(dyV1, DyV, 3); setv(dzV1, DzV, 3);
                setv(dxV2, DxV, 4); setv(dyV2, DyV, 4); setv(dzV2, DzV, 4);
                matmul3x1(P2M[0], dxV2, o1); matmul3x1(P2M[1], dyV2, o2); matmul3x1(P2M[2], dzV2, o3);
                matmul3x1(P2M[3], dyV1, o4); matmul3x1(P2M[4], dzV1, o5);
                for (int c = 0; c < 3; c++) DxV[3][c] = o1[c] + o2[c] + o3[c] - o4[c] - o5[c];
              }
              if (imethod == 2) method2_full();
            } else {
              if (imethod == 1) {
                setv(dxV1, DxV, 3); setv(dyV1, DyV, 3); setv(dzV1, DzV, 3);
                setv(dyV2, DyV, 4); setv(dzV2, DzV, 4);
                matmul3x1(M2P[0], dxV1, o1); matmul3x1(M2P[1], dyV1, o2); matmul3x1(M2P[2], dzV1, o3);
                matmul3x1(M2P[3], dyV2, o4); matmul3x1(M2P[4], dzV2, o5);
                for (int c = 0; c < 3; c++) DxV[4][c] = o1[c] + o2[c] + o3[c] - o4[c] - o5[c];
              }
              if (imethod == 2) method2_full();
            }
          }
          /* f_hT2, f_hT3 on both sides (:936-1018): side 1 uses D*_2, side 0 D*_1 */
          for (int side = 1; side >= 0; side--) {
            int a = side == 1 ? 4 : 3;
            T vec1[3] = {DxV[a][0], DxV[a][1], DxV[a][2]};
            T vec2[3] = {DyV[a][0], DyV[a][1], DyV[a][2]};
            T vec3[3] = {DzV[a][0], DzV[a][1], DzV[a][2]};
            T m1[3][3], m2[3][3], m3[3][3], g1[3], g2[3], g3[3];
            int base = side == 1 ? 6 : 0;
            load9(m1, &F.D[base + 0][idx]); load9(m2, &F.D[base + 1][idx]); load9(m3, &F.D[base + 2][idx]);
            matmul3x1(m1, vec1, g1); matmul3x1(m2, vec2, g2); matmul3x1(m3, vec3, g3);
            size_t ps = pf + side * sslice;
            for (int c = 0; c < 3; c++) fhT2[c][ps] = g1[c] + g2[c] + g3[c];
            load9(m1, &F.D[base + 3][idx]); load9(m2, &F.D[base + 4][idx]); load9(m3, &F.D[base + 5][idx]);
            matmul3x1(m1, vec1, g1); matmul3x1(m2, vec2, g2); matmul3x1(m3, vec3, g3);
            for (int c = 0; c < 3; c++) fhT3[c][ps] = g1[c] + g2[c] + g3[c];
          }
          if (idir == 1) {
            /* E (i0-1), F (i0-2), G (i0-3) with the minus-side fault velocity (:1020-1063) */
            for (int c = 0; c < 3; c++) {
              DxV[2][c] = fV[c][pm] - V[c][(i0 - 1) + row];
              T v3_[3] = {V[c][(i0 - 2) + row], V[c][(i0 - 1) + row], fV[c][pm]};
              DxV[1][c] = vec24_f(v3_);
              T v5_[5] = {V[c][(i0 - 4) + row], V[c][(i0 - 3) + row], V[c][(i0 - 2) + row], V[c][(i0 - 1) + row], fV[c][pm]};
              DxV[0][c] = vecdrp_f(v5_);
            }
          } else {
            /* C (i0+1), B (i0+2), A (i0+3) with the plus-side fault velocity (:1694-1728) */
            for (int c = 0; c < 3; c++) {
              DxV[5][c] = V[c][(i0 + 1) + row] - fV[c][pp];
              T v3_[3] = {fV[c][pp], V[c][(i0 + 1) + row], V[c][(i0 + 2) + row]};
              DxV[6][c] = vec24_b(v3_);
              T v5_[5] = {fV[c][pp], V[c][(i0 + 1) + row], V[c][(i0 + 2) + row], V[c][(i0 + 3) + row], V[c][(i0 + 4) + row]};
              DxV[7][c] = vecdrp_b(v5_);
            }
          }
          /* the six non-fault columns (:1065-1155) */
          for (int n = 0; n < 8; n++) {
            int mm = n < 4 ? n - 3 : n - 4;
            if (mm == 0) continue;
            size_t iptr = (i0 + mm) + row;
            for (int c = 0; c < 3; c++) {
              DyV[n][c] = macdrp(V[c] + iptr, siy, jdir);
              DzV[n][c] = macdrp(V[c] + iptr, siz, kdir);
            }
            size_t ij = ((i0 + mm) + (iy + nj1) * siy) * 9;
            if (isfree == 1 && km == 0) {
              const T *X = &matVx2Vz[ij], *Y = &matVy2Vz[ij];
              for (int r = 0; r < 3; r++)
                DzV[n][r] = X[3 * r + 0] * DxV[n][0] + X[3 * r + 1] * DxV[n][1] + X[3 * r + 2] * DxV[n][2] +
                            Y[3 * r + 0] * DyV[n][0] + Y[3 * r + 1] * DyV[n][1] + Y[3 * r + 2] * DyV[n][2];
            }
            if (isfree == 1 && km == 1)
              for (int c = 0; c < 3; c++) DzV[n][c] = mac22(V[c] + iptr, siz, kdir);
            if (isfree == 1 && km == 2)
              for (int c = 0; c < 3; c++) DzV[n][c] = mac24(V[c] + iptr, siz, kdir);
            hooke_point(iptr, DxV[n], DyV[n], DzV[n], rhs);
          }
        }
    }
  }

  /* ---- wav_update / wav_update_end, wav_t.cu:218-234 ---- */
  static void upd(size_t n, T c, T *u, const T *a, const T *b) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) u[i] = a[i] + c * b[i];
  }
  static void upd_end(size_t n, T c, T *u, const T *b) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) u[i] += c * b[i];
  }
  /* ---- fault_wav_update(_end), fault_wav_t.cu:318-366 ---- */
  void fupd(T c, T *u, const T *a, const T *b, bool end) {
    for (int id = 0; id < nfault; id++) {
      Fault &F = flt[id];
      size_t off = (size_t)id * f_siz_ilevel;
#pragma omp parallel for schedule(static)
      for (int iz = 0; iz < nk; iz++)
        for (int iy = 0; iy < nj; iy++) {
          size_t pf = (iy + nj1) + (iz + nk1) * ny;
          if (F.united[pf] != 0) continue;
          for (int x = 0; x < 18; x++) {
            size_t p = off + pf + x * sslice;
            if (end) u[p] += c * b[p];
            else u[p] = a[p] + c * b[p];
          }
        }
    }
  }
  void pml_upd(int which, T c) {  /* which: 0 tmp=pre+c*rhs, 1 end=pre+c*rhs, 2 end+=c*rhs */
    for (int dim = 0; dim < 3; dim++)
      for (int side = 0; side < 2; side++) {
        Face &f = pml[dim][side];
        if (!f.on) continue;
        size_t n = 9 * f.n;
        if (which == 0) upd(n, c, f.tmp, f.pre, f.rhs);
        if (which == 1) upd(n, c, f.end, f.pre, f.rhs);
        if (which == 2) upd_end(n, c, f.end, f.rhs);
      }
  }

  /* ---- fault_var_stage_update_gpu, fault_wav_t.cu:290-316 ---- */
  void stage_update(T c, int istage) {
    for (int id = 0; id < nfault; id++) {
      Fault &F = flt[id];
#pragma omp parallel for schedule(static)
      for (int iz = 0; iz < nk; iz++)
        for (int iy = 0; iy < nj; iy++) {
          size_t pf = (iy + nj1) + (iz + nk1) * ny;
          if (F.united[pf] != 0) continue;
          if (istage == 0) { F.Tn[pf] = c * F.tTn[pf]; F.Ts1[pf] = c * F.tTs1[pf]; F.Ts2[pf] = c * F.tTs2[pf]; }
          else { F.Tn[pf] += c * F.tTn[pf]; F.Ts1[pf] += c * F.tTs1[pf]; F.Ts2[pf] += c * F.tTs2[pf]; }
        }
    }
  }

  /* ---- fault_var_update_gpu, fault_wav_t.cu:197-264 ---- */
  void var_update(T *fend, int itt) {
    for (int id = 0; id < nfault; id++) {
      Fault &F = flt[id];
      T *fVx = fcmp(fend, id, 0), *fVy = fcmp(fend, id, 1), *fVz = fcmp(fend, id, 2);
#pragma omp parallel for schedule(static)
      for (int iz = 0; iz < nk; iz++)
        for (int iy = 0; iy < nj; iy++) {
          size_t pf = (iy + nj1) + (iz + nk1) * ny;
          if (F.united[pf] != 0) continue;
          size_t p1 = pf + sslice;
          T dVx = fVx[p1] - fVx[pf], dVy = fVy[p1] - fVy[pf], dVz = fVz[p1] - fVz[pf];
          const T *s1 = &F.vec_s1[pf * 3], *s2 = &F.vec_s2[pf * 3];
          T Vs1 = dVx * s1[0] + dVy * s1[1] + dVz * s1[2];
          T Vs2 = dVx * s2[0] + dVy * s2[1] + dVz * s2[2];
          T Vs = xsqrt(Vs2 * Vs2 + Vs1 * Vs1);
          F.Vs[pf] = Vs; F.Vs1[pf] = Vs1; F.Vs2[pf] = Vs2;
          F.Slip[pf] += Vs * dt; F.Slip1[pf] += Vs1 * dt; F.Slip2[pf] += Vs2 * dt;
          size_t p3 = pf + 3 * sslice;
          F.hT1x[pf] = (F.T1x[p3] - F.mT1x[pf]) / dt;
          F.hT1y[pf] = (F.T1y[p3] - F.mT1y[pf]) / dt;
          F.hT1z[pf] = (F.T1z[p3] - F.mT1z[pf]) / dt;
          F.mT1x[pf] = F.T1x[p3]; F.mT1y[pf] = F.T1y[p3]; F.mT1z[pf] = F.T1z[p3];
          if (Vs > F.Peak_vs[pf]) F.Peak_vs[pf] = Vs;
          if (F.init_t0_flag[pf] == 0) {
            if (Vs > 1e-3) {
              F.Init_t0[pf] = (itt + 1) * dt;
              F.init_t0_flag[pf] = 1;
              F.flag_rup[pf] = 1;
            }
          }
        }
    }
  }

  /* ---- PG_calcu_gpu, wav_t.cu:137-216 ---- */
  void pg(const T *wend, const T *wpre) {
    const T *Vx1 = wend, *Vy1 = wend + sicmp, *Vz1 = wend + 2 * sicmp;
    const T *Vx0 = wpre, *Vy0 = wpre + sicmp, *Vz0 = wpre + 2 * sicmp;
    T *P = PG.data();
    T *PGV = P, *PGVh = P + siz, *PGVx = P + 2 * siz, *PGVy = P + 3 * siz, *PGVz = P + 4 * siz;
    T *PGA = P + 5 * siz, *PGAh = P + 6 * siz, *PGAx = P + 7 * siz, *PGAy = P + 8 * siz, *PGAz = P + 9 * siz;
    T *PGD = P + 10 * siz, *PGDh = P + 11 * siz, *PGDx = P + 12 * siz, *PGDy = P + 13 * siz, *PGDz = P + 14 * siz;
    T *Dx = Dis.data(), *Dy = Dx + siz, *Dz = Dx + 2 * siz;
    for (int iy = 0; iy < nj; iy++)
      for (int ix = 0; ix < ni; ix++) {
        size_t iptr = (ix + ni1) + (iy + nj1) * siy + nk2 * siz;
        size_t i1 = (ix + ni1) + (iy + nj1) * siy;
        T Ax = xabs((Vx1[iptr] - Vx0[iptr]) / dt);
        T Ay = xabs((Vy1[iptr] - Vy0[iptr]) / dt);
        T Az = xabs((Vz1[iptr] - Vz0[iptr]) / dt);
        Dx[i1] += (T)(0.5 * (Vx1[iptr] + Vx0[iptr]) * dt);
        Dy[i1] += (T)(0.5 * (Vy1[iptr] + Vy0[iptr]) * dt);
        Dz[i1] += (T)(0.5 * (Vz1[iptr] + Vz0[iptr]) * dt);
        T Vv = xsqrt(Vx1[iptr] * Vx1[iptr] + Vy1[iptr] * Vy1[iptr] + Vz1[iptr] * Vz1[iptr]);
        T Vh = xsqrt(Vx1[iptr] * Vx1[iptr] + Vy1[iptr] * Vy1[iptr]);
        T A = xsqrt(Ax * Ax + Ay * Ay + Az * Az);
        T Ah = xsqrt(Ax * Ax + Ay * Ay);
        T Dd = xsqrt(Dx[i1] * Dx[i1] + Dy[i1] * Dy[i1] + Dz[i1] * Dz[i1]);
        T Dh = xsqrt(Dx[i1] * Dx[i1] + Dy[i1] * Dy[i1]);
        if (PGVx[i1] < xabs(Vx1[iptr])) PGVx[i1] = xabs(Vx1[iptr]);
        if (PGVy[i1] < xabs(Vy1[iptr])) PGVy[i1] = xabs(Vy1[iptr]);
        if (PGVz[i1] < xabs(Vz1[iptr])) PGVz[i1] = xabs(Vz1[iptr]);
        if (PGAx[i1] < Ax) PGAx[i1] = Ax;
        if (PGAy[i1] < Ay) PGAy[i1] = Ay;
        if (PGAz[i1] < Az) PGAz[i1] = Az;
        if (PGDx[i1] < xabs(Dx[i1])) PGDx[i1] = xabs(Dx[i1]);
        if (PGDy[i1] < xabs(Dy[i1])) PGDy[i1] = xabs(Dy[i1]);
        if (PGDz[i1] < xabs(Dz[i1])) PGDz[i1] = xabs(Dz[i1]);
        if (PGV[i1] < Vv) PGV[i1] = Vv;
        if (PGVh[i1] < Vh) PGVh[i1] = Vh;
        if (PGA[i1] < A) PGA[i1] = A;
        if (PGAh[i1] < Ah) PGAh[i1] = Ah;
        if (PGD[i1] < Dd) PGD[i1] = Dd;
        if (PGDh[i1] < Dh) PGDh[i1] = Dh;
      }
  }

  /* ---- one time step: drv_rk_curv_col.cu:209-636 ---- */
  void step() {
    int ipair = it % 8;
    size_t lev = 9 * sicmp;
    size_t flev = (size_t)nfault * f_siz_ilevel;
    for (int istage = 0; istage < 4; istage++) {
      int idir = (FD_Flags[ipair][0] + istage) % 2;
      int jdir = (FD_Flags[ipair][1] + istage) % 2;
      int kdir = (FD_Flags[ipair][2] + istage) % 2;
      T *w_cur = istage == 0 ? w_pre : w_tmp;
      T *f_cur = istage == 0 ? f_pre : f_tmp;
      for (int dim = 0; dim < 3; dim++)
        for (int side = 0; side < 2; side++)
          if (pml[dim][side].on) pml[dim][side].cur = istage == 0 ? pml[dim][side].pre : pml[dim][side].tmp;

      wave2fault(w_cur, w_rhs, f_cur, f_rhs);
      trial(w_cur, f_cur, f_pre, jdir, kdir);
      fault2wave(w_cur, f_cur);
      inner(w_cur, w_rhs, idir, jdir, kdir);
      if (isfree) {
        timg(w_cur, w_rhs, idir, jdir, kdir);
        vlow(w_cur, w_rhs, idir, jdir, kdir);
      }
      if (pml_enable) cfspml(w_cur, w_rhs, idir, jdir, kdir);
      if (nfault > 0) {
        fault_velo(w_cur, w_rhs, f_cur, f_rhs, idir, jdir, kdir);
        fault_stress(w_cur, w_rhs, f_cur, f_rhs, idir, jdir, kdir);
      }
      if (istage == 0) {
        T ca = rk_a[istage] * dt, cb = rk_b[istage] * dt;
        upd(lev, ca, w_tmp, w_pre, w_rhs);
        fupd(ca, f_tmp, f_pre, f_rhs, false);
        fault2wave(w_tmp, f_tmp);
        pml_upd(0, ca);
        upd(lev, cb, w_end, w_pre, w_rhs);
        fupd(cb, f_end, f_pre, f_rhs, false);
        pml_upd(1, cb);
      } else if (istage < 3) {
        T ca = rk_a[istage] * dt, cb = rk_b[istage] * dt;
        upd(lev, ca, w_tmp, w_pre, w_rhs);
        fupd(ca, f_tmp, f_pre, f_rhs, false);
        fault2wave(w_tmp, f_tmp);
        pml_upd(0, ca);
        upd_end(lev, cb, w_end, w_rhs);
        fupd(cb, f_end, nullptr, f_rhs, true);
        pml_upd(2, cb);
      } else {
        T cb = rk_b[istage] * dt;
        upd_end(lev, cb, w_end, w_rhs);
        fupd(cb, f_end, nullptr, f_rhs, true);
        fault2wave(w_end, f_end);
        pml_upd(2, cb);
      }
      stage_update(rk_b[istage], istage);
    }
    (void)flev;
    if (isfree) pg(w_end, w_pre);
    var_update(f_end, it);
    std::swap(w_pre, w_end);
    std::swap(f_pre, f_end);
    for (int dim = 0; dim < 3; dim++)
      for (int side = 0; side < 2; side++)
        if (pml[dim][side].on) std::swap(pml[dim][side].pre, pml[dim][side].end);
    it++;
  }
};

struct Handle {
  int dbl;
  void *o;
};

template <class F>
int dispatch(void *hp, F &&f) {
  Handle *h = (Handle *)hp;
  if (!h) return -1;
  if (h->dbl) return f((Oracle<double> *)h->o);
  return f((Oracle<float> *)h->o);
}

}  // namespace

extern "C" {

/* flags: bit 0 = use FP64, bit 1 = correct the fault2wave Txz/Tyz swap
   (also taken from desc->compat_flags); bits 8-15 = a planted mistake for
   the tests that show a pin catches it:
     1 trial traction without its factor 2 (trial_slipweakening.cu:273-275)
     2 wrong sign of the side term in R (trial_slipweakening.cu:262-264)
     3 split-node DxT1 without its a_0 term (fault_gpu.cu:326-359)
     4 method 2 without the T1 rate term (fault_gpu.cu:662-707)
     5 method 1 without the inv(D11)D11 DxV term (fault_gpu.cu:636-661)
     6 fault2wave velocity from the plus side only (transform.cu:341-343) */
void *oracle_create(const cgfd_desc *d, int flags) {
  Handle *h = new Handle;
  h->dbl = flags & 1;
  int fq = ((flags >> 1) & 1) | (d->compat_flags & CGFD_COMPAT_FIX_F2W_SWAP);
  int plant = (flags >> 8) & 0xff;
  if (h->dbl) { auto *o = new Oracle<double>(d, fq); o->plant = plant; h->o = o; }
  else { auto *o = new Oracle<float>(d, fq); o->plant = plant; h->o = o; }
  return h;
}

void oracle_destroy(void *hp) {
  Handle *h = (Handle *)hp;
  if (!h) return;
  if (h->dbl) delete (Oracle<double> *)h->o;
  else delete (Oracle<float> *)h->o;
  delete h;
}

int oracle_run(void *hp, int nsteps) {
  return dispatch(hp, [&](auto *o) { for (int s = 0; s < nsteps; s++) o->step(); return 0; });
}

int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int oracle_read_wave(void *hp, double *out) {
  return dispatch(hp, [&](auto *o) { for (size_t i = 0; i < 9 * o->sicmp; i++) out[i] = o->w_pre[i]; return 0; });
}

int oracle_write_wave(void *hp, const double *in) {
  return dispatch(hp, [&](auto *o) {
    using T = std::remove_pointer_t<decltype(o->w_pre)>;
    for (size_t i = 0; i < 9 * o->sicmp; i++) o->w_pre[i] = (T)in[i];
    return 0;
  });
}

int oracle_read_fault(void *hp, int id, double *out) {
  return dispatch(hp, [&](auto *o) {
    if (id < 0 || id >= o->nfault) return -1;
    auto &v = o->flt[id].out;
    for (size_t i = 0; i < v.size(); i++) out[i] = v[i];
    return 0;
  });
}

int oracle_read_fault_state(void *hp, int id, double *out) {
  return dispatch(hp, [&](auto *o) {
    if (id < 0 || id >= o->nfault) return -1;
    auto *p = o->f_pre + id * o->f_siz_ilevel;
    for (size_t i = 0; i < o->f_siz_ilevel; i++) out[i] = p[i];
    return 0;
  });
}

/* Overwrites fault id's split-node state at level pre (in[9][2][sslice]) and,
   if t1rec is not null, hT1 and mT1 (t1rec[6][sslice]). */
int oracle_write_fault_state(void *hp, int id, const double *in, const double *t1rec) {
  return dispatch(hp, [&](auto *o) {
    using T = std::remove_pointer_t<decltype(o->w_pre)>;
    if (id < 0 || id >= o->nfault) return -1;
    auto *p = o->f_pre + id * o->f_siz_ilevel;
    for (size_t i = 0; i < o->f_siz_ilevel; i++) p[i] = (T)in[i];
    if (t1rec) {
      auto &F = o->flt[id];
      size_t ss = o->sslice;
      for (size_t i = 0; i < ss; i++) {
        F.hT1x[i] = (T)t1rec[i]; F.hT1y[i] = (T)t1rec[ss + i]; F.hT1z[i] = (T)t1rec[2 * ss + i];
        F.mT1x[i] = (T)t1rec[3 * ss + i]; F.mT1y[i] = (T)t1rec[4 * ss + i]; F.mT1z[i] = (T)t1rec[5 * ss + i];
      }
    }
    return 0;
  });
}

/* The smallest relative distance of each friction branch decision from its
   switch point over all trial evaluations since creation, out[sslice][4]
   (see friction_point); 1e300 at nodes that never reached the trial. */
int oracle_read_kink_margin(void *hp, int id, double *out) {
  return dispatch(hp, [&](auto *o) {
    if (id < 0 || id >= o->nfault) return -1;
    auto &k = o->flt[id].kink;
    for (size_t i = 0; i < k.size(); i++) out[i] = k[i];
    return 0;
  });
}

int oracle_read_pg(void *hp, double *out) {
  return dispatch(hp, [&](auto *o) { for (size_t i = 0; i < o->PG.size(); i++) out[i] = o->PG[i]; return 0; });
}

int oracle_step_index(void *hp) {
  return dispatch(hp, [&](auto *o) { return o->it; });
}

/* Single-operator probe used by the pinning tests: inner-stencil rhs of an
   arbitrary wavefield for one operator-direction choice. */
int oracle_probe_inner(void *hp, const double *w9, int idir, int jdir, int kdir, double *rhs9) {
  return dispatch(hp, [&](auto *o) {
    using T = std::remove_pointer_t<decltype(o->w_pre)>;
    std::vector<T> w(w9, w9 + 9 * o->sicmp), r(9 * o->sicmp, (T)0);
    o->inner(w.data(), r.data(), idir, jdir, kdir);
    for (size_t i = 0; i < r.size(); i++) rhs9[i] = r[i];
    return 0;
  });
}

/* wave2fault then fault2wave on one node (the oracle's own per-node steps):
   metric9 = {xi_x, xi_y, xi_z, eta_x, .., zeta_z}, stress6 = {xx, yy, zz, yz,
   xz, xy} (the wavefield's component order); out6 = the stress fault2wave
   writes back in the same order, with the Txz/Tyz argument swap unless fix_q1.
   traction9 (may be null) = the traction matrix [a][b] row-major. */
int oracle_probe_w2f_f2w(const double *metric9, double jacv, const double *stress6, int fix_q1, double *out6,
                         double *traction9) {
  using T = double;
  T M[3][3], S[3][3], Tr[3][3], sv[6];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) M[c][r] = metric9[3 * r + c];  /* columns = grad xi, grad eta, grad zeta */
  S[0][0] = stress6[0]; S[1][1] = stress6[1]; S[2][2] = stress6[2];
  S[1][2] = S[2][1] = stress6[3]; S[0][2] = S[2][0] = stress6[4]; S[0][1] = S[1][0] = stress6[5];
  Oracle<T>::w2f_node(S, M, jacv, Tr);
  if (traction9)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) traction9[3 * a + b] = Tr[a][b];
  Oracle<T>::f2w_node(Tr, M, jacv, sv);
  out6[0] = sv[0]; out6[1] = sv[1]; out6[2] = sv[2]; out6[5] = sv[3];
  /* the Txz argument of the kernel receives the Tyz array unless fixed */
  if (fix_q1) { out6[4] = sv[4]; out6[3] = sv[5]; }
  else { out6[3] = sv[4]; out6[4] = sv[5]; }
  return 0;
}

/* FP64 reverse-mode gradient of
     J = sum_t sum_s res[s][t] . w_t(station s)
       + sum_t sum_r res_f[r][t][3..8] . (Vs, Vs1, Vs2, Slip, Slip1, Slip2)_t(receiver r)
   over nt steps from the initial wavefield w0 (host layout [9][nz][ny][nx]),
   where w_t is the wavefield and out_t the fault outputs at the start of
   step t (what io_recv_keep / io_fault_recv_keep record), with respect to
   w0 (g_w0, same layout) and to mu_s, mu_d, Dc, C0, T0x, T0y, T0z of every
   fault node (g_par[nfault][7][sslice]). sta_ijk: ghost-inclusive (i,j,k)
   per station; rec_fjk: (fault id, ghost-inclusive j, k) per receiver;
   res_f may be null. Runs on one thread; the tape needs ~24 bytes per
   arithmetic operation of the run, so only small grids. */
int oracle_ad_gradient(const cgfd_desc *d, int fix_q1, const double *w0, int nt, int nsta, const int *sta_ijk,
                       const double *res, int nrec, const int *rec_fjk, const double *res_f, double *g_w0,
                       double *g_par, double *J) {
#ifdef _OPENMP
  const int nthr = omp_get_max_threads();
  omp_set_num_threads(1);
#endif
  ad::tape.clear();
  {
    Oracle<ad::Var> o(d, fix_q1);
    const size_t nw = 9 * o.sicmp, ss = o.sslice;
    std::vector<int> iw(nw), ip((size_t)o.nfault * 7 * ss);
    for (size_t i = 0; i < nw; i++) {
      o.w_pre[i] = ad::leaf(w0[i]);
      iw[i] = o.w_pre[i].i;
    }
    for (int id = 0; id < o.nfault; id++) {
      auto &F = o.flt[id];
      std::vector<ad::Var> *par[7] = {&F.mu_s, &F.mu_d, &F.Dc, &F.C0, &F.T0x, &F.T0y, &F.T0z};
      for (int q = 0; q < 7; q++)
        for (size_t p = 0; p < ss; p++) {
          (*par[q])[p] = ad::leaf((*par[q])[p].v);
          ip[((size_t)id * 7 + q) * ss + p] = (*par[q])[p].i;
        }
    }
    ad::Var Jv(0.0);
    for (int t = 0; t < nt; t++) {
      for (int sidx = 0; sidx < nsta; sidx++) {
        const size_t iptr = sta_ijk[3 * sidx] + sta_ijk[3 * sidx + 1] * o.siy + (size_t)sta_ijk[3 * sidx + 2] * o.siz;
        for (int c = 0; c < 9; c++) {
          const double r = res[((size_t)sidx * nt + t) * 9 + c];
          if (r != 0.0) Jv += r * o.w_pre[c * o.sicmp + iptr];
        }
      }
      for (int r = 0; res_f && r < nrec; r++) {
        auto &F = o.flt[rec_fjk[3 * r]];
        const size_t pf = rec_fjk[3 * r + 1] + (size_t)rec_fjk[3 * r + 2] * o.ny;
        for (int v = 3; v < 9; v++) {
          const double w = res_f[((size_t)r * nt + t) * 9 + v];
          if (w != 0.0) Jv += w * F.out[v * ss + pf];
        }
      }
      o.step();
    }
    *J = Jv.v;
    std::vector<double> adj(ad::tape.size(), 0.0);
    if (Jv.i >= 0) adj[Jv.i] = 1.0;
    for (long n = (long)ad::tape.size() - 1; n >= 0; n--) {
      const double g = adj[n];
      if (g == 0.0) continue;
      const ad::Node &nd = ad::tape[n];
      if (nd.a >= 0) adj[nd.a] += nd.da * g;
      if (nd.b >= 0) adj[nd.b] += nd.db * g;
    }
    for (size_t i = 0; i < nw; i++) g_w0[i] = adj[iw[i]];
    for (size_t i = 0; i < ip.size(); i++) g_par[i] = adj[ip[i]];
  }
  ad::tape.clear();
  ad::tape.shrink_to_fit();
#ifdef _OPENMP
  omp_set_num_threads(nthr);
#endif
  return 0;
}

/* Free-surface conversion matrices computed at create (bdryfree matVx2Vz2,
   matVy2Vz2): out[2][nx*ny*9] */
int oracle_read_freemats(void *hp, double *out) {
  return dispatch(hp, [&](auto *o) {
    size_t n = o->matVx2Vz.size();
    for (size_t i = 0; i < n; i++) { out[i] = o->matVx2Vz[i]; out[n + i] = o->matVy2Vz[i]; }
    return 0;
  });
}

/* Friction probe: the trial-traction / slip-weakening branch of
   trial_slipweakening.cu:296-384 on given local quantities, for closed-form
   tests. in[]: Trial[3], jacvec, vn[3] (unit), T0[3], mu_s, mu_d, Slip, Dc, C0,
   s1[3], s2[3]; out[]: T1new[3], tTn, tTs1, tTs2, flag_rup */
int oracle_probe_friction(const double *in, double *out, int use_double) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    T Trial[3] = {(T)in[0], (T)in[1], (T)in[2]};
    T jacvec = (T)in[3];
    T vn[3] = {(T)in[4], (T)in[5], (T)in[6]};
    T T0[3] = {(T)in[7], (T)in[8], (T)in[9]};
    T s1[3] = {(T)in[15], (T)in[16], (T)in[17]}, s2[3] = {(T)in[18], (T)in[19], (T)in[20]};
    T T1c[3], tn, ts1, ts2;
    int flag = friction_point<T>(Trial, jacvec, vn, T0, (T)in[10], (T)in[11], (T)in[12], (T)in[13], (T)in[14], s1, s2,
                                 T1c, tn, ts1, ts2);
    for (int a = 0; a < 3; a++) out[a] = (double)T1c[a];
    out[3] = tn; out[4] = ts1; out[5] = ts2; out[6] = flag;
  };
  if (use_double) run(0.0); else run(0.0f);
  return 0;
}

}  /* extern "C" */
