/*
 * fdm_oracle.c -- CPU restatement of FDMAP's FP64 time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP library); see
 * fdm_oracle.h.  Every routine names the reference routine it restates;
 * loops follow the reference's loop structure and every floating-point
 * expression keeps the reference's left-to-right evaluation order.  The
 * recipe in oracle/Makefile compiles this file with -ffp-contract=off so
 * that no multiply-add is fused.
 *
 * Parity status of each part is recorded where the oracle is pinned
 * (tests/test_oracle_*.py).
 */
#include "fdm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* SBP tables: the literal 15-digit values of fd_coeff.f90:79-246 for the    */
/* methods 'SBP2', 'SBP4' and 'SBP6' (init_fd).                             */
/* ---------------------------------------------------------------------- */

typedef struct {
  int order;            /* 2, 4, 6 */
  int nbnd, nbst, pI;   /* fd_coeff.f90:56-70: points of a boundary stencil,
                           boundary stencils, interior half-width */
  double H00i;
  double DI[7], AI[7];  /* DI(s), AI(s) at [pI + s] */
  double DL[6][9];      /* DL[r][s] = DL(s,r) */
  double AL[6][9];      /* AL[r][s] = AL(s,r) */
  double HL[6];
} sbp_t;

/* 'SBP2', fd_coeff.f90:83-96 */
static const sbp_t SBP2_OP = {
    2, 2, 1, 1, 2.0,
    {-0.5, 0.0, 0.5}, {0.5, -1.0, 0.5},
    {{-1.0, 1.0}}, {{-1.0, 1.0}}, {0.5}};

/* 'SBP4', fd_coeff.f90:97-126 */
static const sbp_t SBP4_OP = {
    4, 6, 4, 2, 2.823529411764706,
    {0.083333333333333, -0.666666666666667, 0.000000000000000, 0.666666666666667,
     -0.083333333333333},
    {-0.083333333333333, 0.333333333333333, -0.500000000000000, 0.333333333333333,
     -0.083333333333333},
    {{-1.411764705882353, 1.735294117647059, -0.235294117647059, -0.088235294117647,
      0.000000000000000, 0.000000000000000},
     {-0.500000000000000, 0.000000000000000, 0.500000000000000, 0.000000000000000,
      0.000000000000000, 0.000000000000000},
     {0.093023255813953, -0.686046511627907, 0.000000000000000, 0.686046511627907,
      -0.093023255813953, 0.000000000000000},
     {0.030612244897959, 0.000000000000000, -0.602040816326531, 0.000000000000000,
      0.653061224489796, -0.081632653061224}},
    {{-0.470588235294118, 0.941176470588235, -0.470588235294118, 0.000000000000000,
      0.000000000000000, 0.000000000000000},
     {0.271186440677966, -0.610169491525424, 0.406779661016949, -0.067796610169492,
      0.000000000000000, 0.000000000000000},
     {-0.186046511627907, 0.558139534883721, -0.651162790697674, 0.372093023255814,
      -0.093023255813953, 0.000000000000000},
     {0.000000000000000, -0.081632653061224, 0.326530612244898, -0.489795918367347,
      0.326530612244898, -0.081632653061224}},
    {0.354166666666667, 1.229166666666667, 0.895833333333333, 1.020833333333333}};

/* 'SBP6', fd_coeff.f90:128-173 */
static const sbp_t SBP6_OP = {
    6, 9, 6, 3, 3.165067037878233,
    {-0.016666666666667, 0.150000000000000, -0.750000000000000, 0.000000000000000,
     0.750000000000000, -0.150000000000000, 0.016666666666667},
    {0.016666666666667, -0.100000000000000, 0.250000000000000, -0.333333333333333,
     0.250000000000000, -0.100000000000000, 0.016666666666667},
    {{-1.582533518939116, 2.033378678700676, -0.141512858744873, -0.450398306578272,
      0.104488069284042, 0.036577936277544, 0.000000000000000, 0.000000000000000,
      0.000000000000000},
     {-0.462059195631158, 0.000000000000000, 0.287258622978251, 0.258816087376832,
      -0.069112065532624, -0.014903449191300, 0.000000000000000, 0.000000000000000,
      0.000000000000000},
     {0.071247104721830, -0.636451095137907, 0.000000000000000, 0.606235523609147,
      -0.022902190275815, -0.018129342917256, 0.000000000000000, 0.000000000000000,
      0.000000000000000},
     {0.114713313798970, -0.290087484386815, -0.306681191361148, 0.000000000000000,
      0.520262285050482, -0.051642265516119, 0.013435342414630, 0.000000000000000,
      0.000000000000000},
     {-0.036210680656541, 0.105400944933782, 0.015764336127392, -0.707905442575989,
      0.000000000000000, 0.769199413962647, -0.164529643265203, 0.018281071473911,
      0.000000000000000},
     {-0.011398193015050, 0.020437334208704, 0.011220896474665, 0.063183694641876,
      -0.691649024426814, 0.000000000000000, 0.739709139060752, -0.147941827812150,
      0.016437980868017}},
    {{-0.105502234595941, 0.316506703787823, -0.316506703787823, 0.105502234595941,
      0.000000000000000, 0.000000000000000, 0.000000000000000, 0.000000000000000,
      0.000000000000000},
     {0.071922084408557, -0.227753267293765, 0.251727295429951, -0.107883126612836,
      0.011987014068093, 0.000000000000000, 0.000000000000000, 0.000000000000000,
      0.000000000000000},
     {-0.159350793065290, 0.557727775728513, -0.743637034304685, 0.478052379195869,
      -0.159350793065290, 0.026558465510882, 0.000000000000000, 0.000000000000000,
      0.000000000000000},
     {0.026870684829259, -0.120918081731666, 0.241836163463333, -0.282142190707221,
      0.201530136219444, -0.080612054487778, 0.013435342414630, 0.000000000000000,
      0.000000000000000},
     {0.000000000000000, 0.018281071473911, -0.109686428843468, 0.274216072108671,
      -0.365621429478228, 0.274216072108671, -0.109686428843468, 0.018281071473911,
      0.000000000000000},
     {0.000000000000000, 0.000000000000000, 0.016437980868017, -0.098627885208100,
      0.246569713020251, -0.328759617360334, 0.246569713020251, -0.098627885208100,
      0.016437980868017}},
    {0.315949074074074, 1.390393518518518, 0.627546296296296, 1.240509259259259,
     0.911689814814815, 1.013912037037037}};

static const sbp_t* sbp_of(int order) {
  return order == 2 ? &SBP2_OP : order == 6 ? &SBP6_OP : &SBP4_OP;
}

/* The operator of the context being stepped.  Every entry point that
   evaluates rates sets it from its context (use_op) before any rate is
   formed; the oracle steps one context at a time. */
static const sbp_t* g_op = &SBP4_OP;

#define NBND (g_op->nbnd)
#define NBST (g_op->nbst)
#define PI_I (g_op->pI)

static inline double DIc(int s) { return g_op->DI[g_op->pI + s]; }
static inline double AIc(int s) { return g_op->AI[g_op->pI + s]; }
/* AL(s,r); AR(-s,-r) = AL(s,r) (fd_coeff.f90:242-246, no sign change) */
static inline double ALc(int s, int r) { return g_op->AL[r][s]; }
/* DL(s,r) */
static inline double DLc(int s, int r) { return g_op->DL[r][s]; }
/* DR(-s,-r) = -DL(s,r)  (fd_coeff.f90:242-246) */
static inline double DRc(int s, int r) { return -g_op->DL[r][s]; }

/* the tables of one method: DI(-pI..pI), DL[r][s] (nbst x nbnd), HL(0..nbst-1),
   H00i; returns nbnd*100 + nbst*10 + pI, 0 for an unknown order */
int32_t orc_sbp_tables(int32_t order, double* DI, double* DL, double* HL, double* H00i) {
  if (order != 2 && order != 4 && order != 6) return 0;
  const sbp_t* op = sbp_of(order);
  for (int k = 0; k < 2 * op->pI + 1; k++) DI[k] = op->DI[k];
  for (int r = 0; r < op->nbst; r++)
    for (int s = 0; s < op->nbnd; s++) DL[r * op->nbnd + s] = op->DL[r][s];
  for (int r = 0; r < op->nbst; r++) HL[r] = op->HL[r];
  *H00i = op->H00i;
  return op->nbnd * 100 + op->nbst * 10 + op->pI;
}

void orc_sbp4_tables(double* DI, double* DL, double* HL, double* H00i) {
  (void)orc_sbp_tables(4, DI, DL, HL, H00i);
}

/* fd.f90:42-80, endm = endp = true, unit spacing, with the operator of
   the given order */
void orc_diff_op(int32_t order, int32_t n, const double* f, double* Df) {
  const sbp_t* saved = g_op;
  g_op = sbp_of(order);
  orc_diff(n, f, Df);
  g_op = saved;
}

/* fd.f90:42-80, endm = endp = true, unit spacing (the current operator,
   SBP4 outside a step) */
void orc_diff(int32_t n, const double* f, double* Df) {
  if (n == 1) {
    Df[0] = 0.0;
    return;
  }
  for (int i = 0; i < NBST; i++) {
    double s = 0.0;
    for (int k = 0; k < NBND; k++) s = s + DLc(k, i) * f[k];
    Df[i] = s;
  }
  for (int i = 0; i < NBST; i++) {
    /* DR(:,-i) . f(pg-nbnd+1:pg): DR index runs -(nbnd-1)..0 */
    double s = 0.0;
    for (int k = 0; k < NBND; k++) {
      int sidx = (NBND - 1) - k; /* DR(-sidx,-i) multiplies f(pg-sidx) */
      s = s + DRc(sidx, i) * f[n - 1 - sidx];
    }
    Df[n - 1 - i] = s;
  }
  for (int i = NBST; i <= n - 1 - NBST; i++) {
    double s = 0.0;
    for (int k = -PI_I; k <= PI_I; k++) s = s + DIc(k) * f[i + k];
    Df[i] = s;
  }
}

/* the one-dimensional dissipation operator of the rates with unit J and
   Cdiss: AL rows near the ends, AI in between (the q part of
   rates_dissipation_boundary_mode2 'L'/'R' and of the interior) */
void orc_diss1d(int32_t n, const double* f, double* Af) {
  for (int i = 0; i < n; i++) Af[i] = 0.0;
  if (n < 2 * NBND) return;
  for (int i = 0; i < NBST; i++) {
    double s = 0.0;
    for (int k = 0; k < NBND; k++) s = s + ALc(k, i) * f[k];
    Af[i] = s;
    double t = 0.0;
    for (int k = 0; k < NBND; k++) t = t + ALc(k, i) * f[n - 1 - k];
    Af[n - 1 - i] = t;
  }
  for (int i = NBST; i <= n - 1 - NBST; i++) {
    double s = 0.0;
    for (int k = -PI_I; k <= PI_I; k++) s = s + AIc(k) * f[i + k];
    Af[i] = s;
  }
}

/* time_step.f90:25-66 */
int32_t orc_rk_coeffs(int32_t order, double* A, double* B, double* C) {
  switch (order) {
    case 1:
      A[0] = 0.0;
      B[0] = 1.0;
      C[0] = 0.0;
      C[1] = 1.0;
      return 1;
    case 2:
      A[0] = 0.0;
      A[1] = -1.0;
      B[0] = 1.0;
      B[1] = 1.0 / 2.0;
      C[0] = 0.0;
      C[1] = 1.0;
      C[2] = 1.0;
      return 2;
    case 3:
      A[0] = 0.0;
      A[1] = -5.0 / 9.0;
      A[2] = -153.0 / 128.0;
      B[0] = 1.0 / 3.0;
      B[1] = 15.0 / 16.0;
      B[2] = 8.0 / 15.0;
      C[0] = 0.0;
      C[1] = 1.0 / 3.0;
      C[2] = 3.0 / 4.0;
      C[3] = 1.0;
      return 3;
    case 4:
      A[0] = 0.0;
      A[1] = -567301805773.0 / 1357537059087.0;
      A[2] = -2404267990393.0 / 2016746695238.0;
      A[3] = -3550918686646.0 / 2091501179385.0;
      A[4] = -1275806237668.0 / 842570457699.0;
      B[0] = 1432997174477.0 / 9575080441755.0;
      B[1] = 5161836677717.0 / 13612068292357.0;
      B[2] = 1720146321549.0 / 2090206949498.0;
      B[3] = 3134564353537.0 / 4481467310338.0;
      B[4] = 2277821191437.0 / 14882151754819.0;
      C[0] = 0.0;
      C[1] = 1432997174477.0 / 9575080441755.0;
      C[2] = 2526269341429.0 / 6820363962896.0;
      C[3] = 2006345519317.0 / 3224310063776.0;
      C[4] = 2802321613138.0 / 2924317926251.0;
      C[5] = 1.0;
      return 5;
    default:
      return 0;
  }
}

/* ---------------------------------------------------------------------- */
/* rotations, geometry.f90:285-336 and fields.f90:1312-1356                */
/* ---------------------------------------------------------------------- */

typedef struct {
  double vt, vn, stt, snt, snn, szz;
} nt_t;

static inline nt_t xy2nt(const double* F, double n1, double n2) {
  nt_t r;
  r.vt = n2 * F[0] - n1 * F[1];
  r.vn = n1 * F[0] + n2 * F[1];
  r.stt = -2.0 * n1 * n2 * F[3] + n2 * n2 * F[2] + n1 * n1 * F[4];
  r.snt = n1 * n2 * (F[2] - F[4]) + (n2 * n2 - n1 * n1) * F[3];
  r.snn = 2.0 * n1 * n2 * F[3] + n1 * n1 * F[2] + n2 * n2 * F[4];
  r.szz = F[5];
  return r;
}

static inline void nt2xy(double* F, double n1, double n2, double vt, double vn, double stt,
                         double snt, double snn, double szz) {
  F[0] = n2 * vt + n1 * vn;
  F[1] = -n1 * vt + n2 * vn;
  F[2] = 2.0 * n1 * n2 * snt + n2 * n2 * stt + n1 * n1 * snn;
  F[3] = -n1 * n2 * (stt - snn) + (n2 * n2 - n1 * n1) * snt;
  F[4] = -2.0 * n1 * n2 * snt + n1 * n1 * stt + n2 * n2 * snn;
  F[5] = szz;
}

/* friction.f90:1293-1304 */
static inline double arcsinh_ref(double x) {
  double y = fabs(x);
  double f = log(y + sqrt(y * y + 1.0));
  return copysign(fabs(f), x);
}

/* Fortran SIGN(a,b) */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

/* x**8 as repeated squaring (integer power) */
static inline double pow8(double x) {
  double x2 = x * x;
  double x4 = x2 * x2;
  return x4 * x4;
}

/* ---------------------------------------------------------------------- */
/* manufactured solution, mms.f90:52-234                                   */
/* ---------------------------------------------------------------------- */

double orc_mms_sin(double x, double y, double t, int32_t side, int32_t field) {
  const double G = 32.0, cs = 3.0, cp = 5.0;
  double rho = G / (cs * cs);
  double lambda = rho * (cp * cp) - 2.0 * G;
  double Zs = rho * cs;
  double pi = 4.0 * atan(1.0);
  double kx = 2.0 * pi, ky = 2.0 * pi, mv = 2.0;
  double w = (4.0 / 0.46) * pi;
  double ft = cos(w * t);
  double ft_t = -w * sin(w * t);
  double mf = 0.1, kf = 1.0 * pi;
  double ftpp = mf * kf * cos(kf * x);
  double ftpp_x = -mf * (kf * kf) * sin(kf * x);
  double nx = ftpp / sqrt(1.0 + ftpp * ftpp);
  double ny = -1.0 / sqrt(1.0 + ftpp * ftpp);
  double mx = ny, my = -nx;
  double q = 1.0 + ftpp * ftpp;
  double q3 = q * q * q;
  double nx_x = ftpp_x / sqrt(q3);
  double ny_x = ftpp * ftpp_x / sqrt(q3);
  double mx_x = ny_x, my_x = -nx_x;
  double nx_y = 0.0, ny_y = 0.0;
  double mx_y = ny_y, my_y = -nx_y;
  double sd = (double)side;
  double amp = 2.0 * 12.6 / Zs;
  double ckx = cos(kx * x), sky = sin(ky * y), cky = cos(ky * y), skx = sin(kx * x);

  double vx = mv * (3.0 - 2.0 * sd) * amp * ckx * cky * ft;
  double vx_t = mv * (3.0 - 2.0 * sd) * amp * ckx * cky * ft_t;
  double vx_x = -mv * (3.0 - 2.0 * sd) * kx * amp * skx * cky * ft;
  double vx_y = -mv * (3.0 - 2.0 * sd) * ky * amp * ckx * sky * ft;

  double vy = amp * ckx * cky * ft;
  double vy_t = amp * ckx * cky * ft_t;
  double vy_x = -kx * amp * skx * cky * ft;
  double vy_y = -ky * amp * ckx * sky * ft;

  double syy = -63.0 * ckx * cky * ft - 126.0;
  double syy_t = -63.0 * ckx * cky * ft_t;
  double syy_x = kx * 63.0 * skx * cky * ft;
  double syy_y = ky * 63.0 * ckx * sky * ft;

  double sxy = 0.6 * 126.0 * ckx * cky * ft;
  double sxy_t = 0.6 * 126.0 * ckx * cky * ft_t;
  double sxy_x = -kx * 0.6 * 126.0 * skx * cky * ft;
  double sxy_y = -ky * 0.6 * 126.0 * ckx * sky * ft;

  double sxx = (2.0 * sd - 3.0) * 0.6 * 1260.0 * ckx * cky * ft;
  double sxx_t = (2.0 * sd - 3.0) * 0.6 * 1260.0 * ckx * cky * ft_t;
  double sxx_x = -(2.0 * sd - 3.0) * kx * 0.6 * 1260.0 * skx * cky * ft;
  double sxx_y = -(2.0 * sd - 3.0) * ky * 0.6 * 1260.0 * ckx * sky * ft;

  switch (field) {
    case 5: /* szz */
      return 0.0;
    case 12: /* V */
      return 2.0 * vx;
    case 0: /* vx */
      return ny * vx - my * vy;
    case 1: /* vy */
      return mx * vy - nx * vx;
    case 13: /* S */
      return sxy;
    case 14: /* N */
      return -syy;
    case 2: /* sxx */
      return syy * (my * my) + sxx * (ny * ny) - 2.0 * sxy * my * ny;
    case 4: /* syy */
      return syy * (mx * mx) + sxx * (nx * nx) - 2.0 * sxy * mx * nx;
    case 3: /* sxy */
      return -syy * my * mx - sxx * ny * nx + sxy * (ny * mx + nx * my);
    case 6: { /* s_vx */
      double vxt = ny * vx_t - my * vy_t;
      double sxxx = syy_x * (my * my) + sxx_x * (ny * ny) - 2.0 * sxy_x * my * ny +
                    2.0 * syy * my * my_x + 2.0 * sxx * ny * ny_x -
                    2.0 * sxy * (my_x * ny + my * ny_x);
      double sxyy = -syy_y * my * mx - sxx_y * ny * nx + sxy_y * (ny * mx + nx * my) -
                    syy * (my_y * mx + my * mx_y) - sxx * (ny_y * nx + ny * nx_y) +
                    sxy * (ny_y * mx + ny * mx_y + nx_y * my + nx * my_y);
      return vxt - (1.0 / rho) * (sxxx + sxyy);
    }
    case 7: { /* s_vy */
      double vyt = mx * vy_t - nx * vx_t;
      double sxyx = -syy_x * my * mx - sxx_x * ny * nx + sxy_x * (ny * mx + nx * my) -
                    syy * (my_x * mx + my * mx_x) - sxx * (ny_x * nx + ny * nx_x) +
                    sxy * (ny_x * mx + ny * mx_x + nx_x * my + nx * my_x);
      double syyy = syy_y * (mx * mx) + sxx_y * (nx * nx) - 2.0 * sxy_y * mx * nx +
                    2.0 * syy * mx * mx_y + 2.0 * sxx * nx * nx_y -
                    2.0 * sxy * (mx_y * nx + mx * nx_y);
      return vyt - (1.0 / rho) * (sxyx + syyy);
    }
    case 8: { /* s_sxx */
      double sxxt = syy_t * (my * my) + sxx_t * (ny * ny) - 2.0 * sxy_t * my * ny;
      double vxx = ny_x * vx - my_x * vy + ny * vx_x - my * vy_x;
      double vyy = mx_y * vy - nx_y * vx + mx * vy_y - nx * vx_y;
      return sxxt - (lambda + 2.0 * G) * vxx - lambda * vyy;
    }
    case 10: { /* s_syy */
      double vxx = ny_x * vx - my_x * vy + ny * vx_x - my * vy_x;
      double vyy = mx_y * vy - nx_y * vx + mx * vy_y - nx * vx_y;
      double syyt = syy_t * (mx * mx) + sxx_t * (nx * nx) - 2.0 * sxy_t * mx * nx;
      return syyt - lambda * vxx - (lambda + 2.0 * G) * vyy;
    }
    case 9: { /* s_sxy */
      double vxy = ny_y * vx - my_y * vy + ny * vx_y - my * vy_y;
      double vyx = mx_x * vy - nx_x * vx + mx * vy_x - nx * vx_x;
      double sxyt = -syy_t * my * mx - sxx_t * ny * nx + sxy_t * (ny * mx + nx * my);
      return sxyt - G * (vyx + vxy);
    }
    case 11: { /* s_szz */
      double szzt = 0.0;
      double vxx = ny_x * vx - my_x * vy + ny * vx_x - my * vy_x;
      double vyy = mx_y * vy - nx_y * vx + mx * vy_y - nx * vx_y;
      return szzt - lambda * vxx - lambda * vyy;
    }
    default:
      return 0.0;
  }
}

/* ---------------------------------------------------------------------- */
/* context                                                                 */
/* ---------------------------------------------------------------------- */

typedef struct {
  int32_t bc, n, active;
  double *normal, *M, *F0, *x, *y; /* copies */
  double* F;                       /* [6][n] edge fields (bndF%F)   */
  double* Fhat;                    /* [6][n] hat variables          */
} oside_t;

typedef struct {
  fdm_block b;
  oside_t s[4];
} oblock_t;

typedef struct {
  fdm_iface I;
  int32_t n;
  double *nhat, *rs, *sw, *S0base, *N0base, *ldS, *ldN, *Psi0;
  double* knx; /* distance along the fault, kinematic forcing only */
  int32_t* outside;
  double* f[FDM_IF_NFIELDS];
  double* st1[FDM_IF_NFIELDS];
} oiface_t;

struct orc_ctx {
  const sbp_t* op; /* first-derivative operator (FDmethod) */
  int32_t nx, ny, nb, ni, nstage;
  int32_t mode, nF; /* 2: plane strain, 6 fields; 3: antiplane shear, 3 fields */
  size_t np;
  double rkA[FDM_MAX_STAGES], rkB[FDM_MAX_STAGES], rkC[FDM_MAX_STAGES + 1];
  double *xq, *xr, *yq, *yr, *Ji, *x, *y;
  double *F, *DF;
  double *gammap, *lambda;
  /* heterogeneous medium (material.f90:226-275): rho, G, M, gamma, nu per
     point, or NULL */
  double *mrho, *mG, *mM, *mgam, *mnu;
  /* Jacobian J, when a block has artificial dissipation, or NULL */
  double* J;
  /* PML auxiliary fields and rates, [6][np] each (fields.f90:35, 203-209):
     allocated when some block sets pmlx (Wx) or pmly (Wy), else NULL */
  double *Wx, *DWx, *Wy, *DWy;
  oblock_t* B;
  oiface_t* I;
  fdm_source src;
  double t; /* D%t */
  int newton_fail;
  char err[512];
};

static char g_err[512];
static int g_threads = 1;

int orc_set_threads(int32_t nthreads) {
  g_threads = nthreads > 0 ? nthreads : 1;
  return FDM_OK;
}

static double* dup_arr(const double* p, size_t n) {
  if (!p) return NULL;
  double* q = (double*)malloc(n * sizeof(double));
  if (q) memcpy(q, p, n * sizeof(double));
  return q;
}

static double* alloc_fill(size_t n, double v) {
  double* q = (double*)malloc(n * sizeof(double));
  if (q)
    for (size_t k = 0; k < n; k++) q[k] = v;
  return q;
}

#define IDX(i, j) ((size_t)(j) * (size_t)c->nx + (size_t)(i))
#define FL(l) (c->F + (size_t)(l) * c->np)
#define DFL(l) (c->DF + (size_t)(l) * c->np)

static int side_len(const fdm_block* b, int s) {
  return (s < 2) ? (b->py - b->my + 1) : (b->px - b->mx + 1);
}

int orc_create(const fdm_problem* P, orc_ctx** out) {
  *out = NULL;
  if (!P || P->nx < 1 || P->ny < 1 || P->nblocks < 1 || P->nstage < 1 ||
      P->nstage > FDM_MAX_STAGES) {
    snprintf(g_err, sizeof g_err, "orc_create: malformed problem");
    return FDM_EINVAL;
  }
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->nx = P->nx;
  c->ny = P->ny;
  c->np = (size_t)P->nx * (size_t)P->ny;
  c->nb = P->nblocks;
  c->ni = P->nifaces;
  c->nstage = P->nstage;
  c->mode = P->mode == 3 ? 3 : 2;
  c->nF = c->mode == 3 ? 3 : 6;
  if (P->sbp != 0 && P->sbp != 2 && P->sbp != 4 && P->sbp != 6) {
    snprintf(g_err, sizeof g_err, "orc_create: unknown operator SBP%d", (int)P->sbp);
    free(c);
    return FDM_EINVAL;
  }
  c->op = sbp_of(P->sbp == 0 ? 4 : P->sbp);
  for (int b = 0; b < P->nblocks; b++) {
    const fdm_block* B = &P->blocks[b];
    if (B->px - B->mx + 1 < c->op->nbnd || B->py - B->my + 1 < c->op->nbnd) {
      snprintf(g_err, sizeof g_err, "orc_create: block %d is narrower than the operator's %d-point boundary stencil",
               b, c->op->nbnd);
      free(c);
      return FDM_EINVAL;
    }
  }
  for (int s = 0; s < P->nstage; s++) {
    c->rkA[s] = P->rkA[s];
    c->rkB[s] = P->rkB[s];
  }
  for (int s = 0; s <= P->nstage; s++) c->rkC[s] = P->rkC[s];
  c->xq = dup_arr(P->xq, c->np);
  c->xr = dup_arr(P->xr, c->np);
  c->yq = dup_arr(P->yq, c->np);
  c->yr = dup_arr(P->yr, c->np);
  c->Ji = dup_arr(P->Ji, c->np);
  c->x = dup_arr(P->x, c->np);
  c->y = dup_arr(P->y, c->np);
  if (P->mat_rho && P->mat_cs && P->mat_cp) {
    /* init_medium (material.f90:254-262) */
    c->mrho = dup_arr(P->mat_rho, c->np);
    c->mG = alloc_fill(c->np, 0.0);
    c->mM = alloc_fill(c->np, 0.0);
    c->mgam = alloc_fill(c->np, 0.0);
    c->mnu = alloc_fill(c->np, 0.0);
    if (c->mrho && c->mG && c->mM && c->mgam && c->mnu)
      for (size_t q = 0; q < c->np; q++) {
        double rho = P->mat_rho[q], cs = P->mat_cs[q], cp = P->mat_cp[q];
        c->mG[q] = rho * (cs * cs);
        c->mM[q] = rho * (cp * cp);
        c->mgam[q] = 1.0 - 2.0 * ((cs / cp) * (cs / cp));
        c->mnu[q] = 0.5 * (c->mM[q] - 2.0 * c->mG[q]) / (c->mM[q] - c->mG[q]);
      }
  }
  c->F = alloc_fill(6 * c->np, 0.0);
  c->DF = alloc_fill(6 * c->np, 1e40);
  c->gammap = alloc_fill(c->np, 0.0);
  c->lambda = alloc_fill(c->np, 0.0);
  for (int ib = 0; ib < P->nblocks; ib++) {
    if (P->blocks[ib].pmlx && !c->Wx) {
      c->Wx = alloc_fill(6 * c->np, 0.0);
      c->DWx = alloc_fill(6 * c->np, 0.0);
    }
    if (P->blocks[ib].pmly && !c->Wy) {
      c->Wy = alloc_fill(6 * c->np, 0.0);
      c->DWy = alloc_fill(6 * c->np, 0.0);
    }
  }
  c->src = P->source;
  c->B = (oblock_t*)calloc(c->nb, sizeof(oblock_t));
  for (int ib = 0; ib < c->nb; ib++) {
    const fdm_block* b = &P->blocks[ib];
    c->B[ib].b = *b;
    if (c->mode == 3 && (b->plastic || b->mms || P->source.nplace > 0 || b->pmlx || b->pmly)) {
      snprintf(g_err, sizeof g_err,
               "orc_create: mode 3 covers elastic blocks without sources, plasticity, "
               "manufactured solutions or PML");
      orc_destroy(c);
      return FDM_EINVAL;
    }
    if ((b->pmlx || b->pmly) && b->pml[0] > 0.0 && (!c->x || !c->y)) {
      snprintf(g_err, sizeof g_err, "orc_create: block %d has a graded PML and no coordinates", ib);
      orc_destroy(c);
      return FDM_EINVAL;
    }
    if (b->Cdiss > 0.0 && !c->J) {
      if (!P->J) {
        snprintf(g_err, sizeof g_err, "orc_create: block %d has Cdiss > 0 and no Jacobian", ib);
        orc_destroy(c);
        return FDM_EINVAL;
      }
      c->J = dup_arr(P->J, c->np);
    }
    if (b->px - b->mx + 1 < 8 || b->py - b->my + 1 < 8) {
      snprintf(g_err, sizeof g_err, "orc_create: block %d has fewer than 8 points", ib);
      orc_destroy(c);
      return FDM_EINVAL;
    }
    for (int s = 0; s < 4; s++) {
      oside_t* sd = &c->B[ib].s[s];
      int n = side_len(b, s);
      sd->bc = b->side[s].bc;
      sd->n = n;
      sd->active = 1;
      sd->normal = dup_arr(b->side[s].normal, 2 * (size_t)n);
      sd->M = dup_arr(b->side[s].M, 5 * (size_t)n);
      sd->F0 = dup_arr(b->side[s].F0, 6 * (size_t)n);
      sd->x = dup_arr(b->side[s].x, n);
      sd->y = dup_arr(b->side[s].y, n);
      sd->F = alloc_fill(6 * (size_t)n, 1e40);
      sd->Fhat = alloc_fill(6 * (size_t)n, 1e40);
      if (!sd->normal || !sd->M || !sd->F0) {
        snprintf(g_err, sizeof g_err, "orc_create: block %d side %d arrays missing", ib, s);
        orc_destroy(c);
        return FDM_EINVAL;
      }
    }
  }
  c->I = (oiface_t*)calloc(c->ni > 0 ? c->ni : 1, sizeof(oiface_t));
  for (int ii = 0; ii < c->ni; ii++) {
    oiface_t* I = &c->I[ii];
    I->I = P->ifaces[ii];
    int n = I->I.p - I->I.m + 1;
    I->n = n;
    I->nhat = dup_arr(I->I.nhat, 2 * (size_t)n);
    I->rs = dup_arr(I->I.rs, 7 * (size_t)n);
    I->sw = dup_arr(I->I.sw, 4 * (size_t)n);
    I->S0base = I->I.S0base ? dup_arr(I->I.S0base, n) : alloc_fill(n, 0.0);
    I->N0base = I->I.N0base ? dup_arr(I->I.N0base, n) : alloc_fill(n, 0.0);
    I->ldS = I->I.ldS ? dup_arr(I->I.ldS, n) : alloc_fill(n, 0.0);
    I->ldN = I->I.ldN ? dup_arr(I->I.ldN, n) : alloc_fill(n, 0.0);
    I->Psi0 = I->I.Psi0 ? dup_arr(I->I.Psi0, n) : alloc_fill(n, 0.0);
    I->knx = (I->I.coupling == FDM_COUPLING_FRICTION && I->I.force && I->I.kn_x)
                 ? dup_arr(I->I.kn_x, n)
                 : NULL;
    if (I->I.coupling == FDM_COUPLING_FRICTION && I->I.force && !I->knx) {
      snprintf(g_err, sizeof g_err, "orc_create: interface %d: kinematic forcing without kn_x", ii);
      orc_destroy(c);
      return FDM_EINVAL;
    }
    I->outside = (int32_t*)calloc(n, sizeof(int32_t));
    if (I->I.outside) memcpy(I->outside, I->I.outside, n * sizeof(int32_t));
    /* friction.f90:217-249 initial values */
    I->f[FDM_IF_D] = alloc_fill(n, 0.0);
    I->f[FDM_IF_PSI] = dup_arr(I->Psi0, n);
    I->f[FDM_IF_DPSI] = alloc_fill(n, 1e40);
    I->f[FDM_IF_TRUP] = alloc_fill(n, 1e10);
    I->f[FDM_IF_V] = alloc_fill(n, 1e40);
    I->f[FDM_IF_O] = alloc_fill(n, 1e40);
    I->f[FDM_IF_S] = alloc_fill(n, 1e40);
    I->f[FDM_IF_N] = alloc_fill(n, 1e40);
    I->f[FDM_IF_S0] = alloc_fill(n, 1e40);
    I->f[FDM_IF_N0] = alloc_fill(n, 1e40);
    I->f[FDM_IF_DS] = alloc_fill(n, 0.0);
    I->f[FDM_IF_DN] = alloc_fill(n, 0.0);
    I->f[FDM_IF_DDS] = alloc_fill(n, 0.0);
    I->f[FDM_IF_DDN] = alloc_fill(n, 0.0);
    I->f[FDM_IF_W] = alloc_fill(n, 0.0);
    I->f[FDM_IF_DW] = alloc_fill(n, 1e40);
    I->f[FDM_IF_NREG] = alloc_fill(n, 1e40);
    I->f[FDM_IF_DNREG] = alloc_fill(n, 1e40);
    for (int k = 0; k < FDM_IF_NFIELDS; k++) I->st1[k] = alloc_fill(n, 1e40);
    /* interface sides: block m's R/T side and block p's L/B side */
    int sm = I->I.direction == 0 ? FDM_SIDE_R : FDM_SIDE_T;
    int sp = I->I.direction == 0 ? FDM_SIDE_L : FDM_SIDE_B;
    if (I->I.blockm < 0 || I->I.blockm >= c->nb || I->I.blockp < 0 || I->I.blockp >= c->nb ||
        !I->nhat) {
      snprintf(g_err, sizeof g_err, "orc_create: interface %d malformed", ii);
      orc_destroy(c);
      return FDM_EINVAL;
    }
    (void)sm;
    (void)sp;
    if (I->I.coupling == FDM_COUPLING_FRICTION) {
      int rslaw = I->I.law >= FDM_LAW_SL;
      if ((rslaw && !I->rs) || (I->I.law == FDM_LAW_SW && !I->sw)) {
        snprintf(g_err, sizeof g_err, "orc_create: interface %d friction parameters missing", ii);
        orc_destroy(c);
        return FDM_EINVAL;
      }
    }
  }
  /* every side needs a boundary condition or an interface */
  for (int ib = 0; ib < c->nb; ib++)
    for (int s = 0; s < 4; s++) {
      if (c->B[ib].s[s].bc != FDM_BC_NONE) continue;
      int covered = 0;
      for (int ii = 0; ii < c->ni; ii++) {
        const fdm_iface* I = &c->I[ii].I;
        int sm = I->direction == 0 ? FDM_SIDE_R : FDM_SIDE_T;
        int sp = I->direction == 0 ? FDM_SIDE_L : FDM_SIDE_B;
        if ((I->blockm == ib && sm == s) || (I->blockp == ib && sp == s)) covered = 1;
      }
      if (!covered) {
        snprintf(g_err, sizeof g_err,
                 "orc_create: block %d side %d has bc 'none' and no interface", ib, s);
        orc_destroy(c);
        return FDM_EINVAL;
      }
    }
  *out = c;
  return FDM_OK;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  free(c->xq);
  free(c->xr);
  free(c->yq);
  free(c->yr);
  free(c->Ji);
  free(c->x);
  free(c->y);
  free(c->F);
  free(c->DF);
  free(c->gammap);
  free(c->lambda);
  free(c->Wx);
  free(c->DWx);
  free(c->Wy);
  free(c->DWy);
  free(c->mrho);
  free(c->mG);
  free(c->mM);
  free(c->mgam);
  free(c->mnu);
  free(c->J);
  if (c->B)
    for (int ib = 0; ib < c->nb; ib++)
      for (int s = 0; s < 4; s++) {
        oside_t* sd = &c->B[ib].s[s];
        free(sd->normal);
        free(sd->M);
        free(sd->F0);
        free(sd->x);
        free(sd->y);
        free(sd->F);
        free(sd->Fhat);
      }
  free(c->B);
  if (c->I)
    for (int ii = 0; ii < c->ni; ii++) {
      oiface_t* I = &c->I[ii];
      free(I->nhat);
      free(I->rs);
      free(I->sw);
      free(I->S0base);
      free(I->N0base);
      free(I->ldS);
      free(I->ldN);
      free(I->Psi0);
      free(I->knx);
      free(I->outside);
      for (int k = 0; k < FDM_IF_NFIELDS; k++) {
        free(I->f[k]);
        free(I->st1[k]);
      }
    }
  free(c->I);
  free(c);
}

const char* orc_last_error(const orc_ctx* c) { return c ? c->err : g_err; }

int orc_set_fields(orc_ctx* c, const double* F, const double* DF) {
  size_t nf = (size_t)c->nF * c->np;
  memcpy(c->F, F, nf * sizeof(double));
  if (DF)
    memcpy(c->DF, DF, nf * sizeof(double));
  else
    for (size_t k = 0; k < nf; k++) c->DF[k] = 1e40;
  return FDM_OK;
}

int orc_get_fields(orc_ctx* c, double* F, double* DF) {
  size_t nf = (size_t)c->nF * c->np;
  if (F) memcpy(F, c->F, nf * sizeof(double));
  if (DF) memcpy(DF, c->DF, nf * sizeof(double));
  return FDM_OK;
}

int orc_get_iface(orc_ctx* c, int32_t ii, int32_t field, double* out) {
  if (ii < 0 || ii >= c->ni || field < 0 || field >= FDM_IF_NFIELDS) return FDM_EINVAL;
  memcpy(out, c->I[ii].f[field], c->I[ii].n * sizeof(double));
  return FDM_OK;
}

int orc_set_iface(orc_ctx* c, int32_t ii, int32_t field, const double* in) {
  if (ii < 0 || ii >= c->ni || field < 0 || field >= FDM_IF_NFIELDS) return FDM_EINVAL;
  memcpy(c->I[ii].f[field], in, c->I[ii].n * sizeof(double));
  return FDM_OK;
}

int orc_get_iface_stage1(orc_ctx* c, int32_t ii, int32_t field, double* out) {
  if (ii < 0 || ii >= c->ni || field < 0 || field >= FDM_IF_NFIELDS) return FDM_EINVAL;
  memcpy(out, c->I[ii].st1[field], c->I[ii].n * sizeof(double));
  return FDM_OK;
}

int orc_get_plastic(orc_ctx* c, double* gammap, double* lambda) {
  if (gammap) memcpy(gammap, c->gammap, c->np * sizeof(double));
  if (lambda) memcpy(lambda, c->lambda, c->np * sizeof(double));
  return FDM_OK;
}

int orc_get_pml(orc_ctx* c, double* Wx, double* Wy) {
  if ((Wx && !c->Wx) || (Wy && !c->Wy)) {
    snprintf(c->err, sizeof c->err, "orc_get_pml: no block has that PML direction");
    return FDM_EINVAL;
  }
  if (Wx) memcpy(Wx, c->Wx, 6 * c->np * sizeof(double));
  if (Wy) memcpy(Wy, c->Wy, 6 * c->np * sizeof(double));
  return FDM_OK;
}

int orc_get_pml_rates(orc_ctx* c, double* DWx, double* DWy) {
  if ((DWx && !c->DWx) || (DWy && !c->DWy)) {
    snprintf(c->err, sizeof c->err, "orc_get_pml_rates: no block has that PML direction");
    return FDM_EINVAL;
  }
  if (DWx) memcpy(DWx, c->DWx, 6 * c->np * sizeof(double));
  if (DWy) memcpy(DWy, c->DWy, 6 * c->np * sizeof(double));
  return FDM_OK;
}

int orc_set_pml(orc_ctx* c, const double* Wx, const double* Wy) {
  if ((Wx && !c->Wx) || (Wy && !c->Wy)) {
    snprintf(c->err, sizeof c->err, "orc_set_pml: no block has that PML direction");
    return FDM_EINVAL;
  }
  if (Wx) memcpy(c->Wx, Wx, 6 * c->np * sizeof(double));
  if (Wy) memcpy(c->Wy, Wy, 6 * c->np * sizeof(double));
  return FDM_OK;
}

/* ---------------------------------------------------------------------- */
/* scale / update, fields.f90:1061-1133, 1167-1242; friction.f90:555-590  */
/* ---------------------------------------------------------------------- */

static void scale_rates_all(orc_ctx* c, double A) {
  size_t n6 = (size_t)c->nF * c->np;
  double* DF = c->DF;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (size_t k = 0; k < n6; k++) DF[k] = A * DF[k];
  /* scale_rates_interior (fields.f90:1074-1077) */
  if (c->DWx)
    for (size_t k = 0; k < 6 * c->np; k++) c->DWx[k] = A * c->DWx[k];
  if (c->DWy)
    for (size_t k = 0; k < 6 * c->np; k++) c->DWy[k] = A * c->DWy[k];
  for (int ii = 0; ii < c->ni; ii++) {
    oiface_t* I = &c->I[ii];
    if (I->I.coupling != FDM_COUPLING_FRICTION) continue;
    for (int k = 0; k < I->n; k++) {
      I->f[FDM_IF_DDS][k] = A * I->f[FDM_IF_DDS][k];
      I->f[FDM_IF_DDN][k] = A * I->f[FDM_IF_DDN][k];
      I->f[FDM_IF_DPSI][k] = A * I->f[FDM_IF_DPSI][k];
      I->f[FDM_IF_DW][k] = A * I->f[FDM_IF_DW][k];
      I->f[FDM_IF_DNREG][k] = A * I->f[FDM_IF_DNREG][k];
    }
  }
}

static void update_fields_all(orc_ctx* c, double Bdt) {
  size_t n6 = (size_t)c->nF * c->np;
  double* F = c->F;
  const double* DF = c->DF;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (size_t k = 0; k < n6; k++) F[k] = F[k] + Bdt * DF[k];
  /* update_fields_interior (fields.f90:1181-1184) */
  if (c->Wx)
    for (size_t k = 0; k < 6 * c->np; k++) c->Wx[k] = c->Wx[k] + Bdt * c->DWx[k];
  if (c->Wy)
    for (size_t k = 0; k < 6 * c->np; k++) c->Wy[k] = c->Wy[k] + Bdt * c->DWy[k];
  for (int ii = 0; ii < c->ni; ii++) {
    oiface_t* I = &c->I[ii];
    if (I->I.coupling != FDM_COUPLING_FRICTION) continue;
    for (int k = 0; k < I->n; k++) {
      I->f[FDM_IF_DS][k] = I->f[FDM_IF_DS][k] + Bdt * I->f[FDM_IF_DDS][k];
      I->f[FDM_IF_DN][k] = I->f[FDM_IF_DN][k] + Bdt * I->f[FDM_IF_DDN][k];
      I->f[FDM_IF_D][k] = I->f[FDM_IF_D][k] + Bdt * fabs(I->f[FDM_IF_DDS][k]);
      I->f[FDM_IF_PSI][k] = I->f[FDM_IF_PSI][k] + Bdt * I->f[FDM_IF_DPSI][k];
      I->f[FDM_IF_W][k] = I->f[FDM_IF_W][k] + Bdt * I->f[FDM_IF_DW][k];
      I->f[FDM_IF_NREG][k] = I->f[FDM_IF_NREG][k] + Bdt * I->f[FDM_IF_DNREG][k];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* edges: domain.f90:1325-1346 (interior_to_edges)                          */
/* ---------------------------------------------------------------------- */

static void prepare_edges(orc_ctx* c) {
  for (int ib = 0; ib < c->nb; ib++) {
    fdm_block* b = &c->B[ib].b;
    for (int s = 0; s < 4; s++) {
      oside_t* sd = &c->B[ib].s[s];
      for (int k = 0; k < sd->n; k++) {
        int i, j;
        if (s == FDM_SIDE_L) {
          i = b->mx;
          j = b->my + k;
        } else if (s == FDM_SIDE_R) {
          i = b->px;
          j = b->my + k;
        } else if (s == FDM_SIDE_B) {
          i = b->mx + k;
          j = b->my;
        } else {
          i = b->mx + k;
          j = b->py;
        }
        for (int l = 0; l < c->nF; l++) sd->F[(size_t)l * sd->n + k] = FL(l)[IDX(i, j)];
      }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* boundary conditions, boundaries.f90:216-377 (set_bc_mode2)              */
/* ---------------------------------------------------------------------- */

static void set_bc_mode2(double* Fhat, const double* F, const double* F0, double n1, double n2,
                         int bc, double Zs, double Zp, double gamma, double x, double y,
                         double t, int iblock) {
  double Zsi = (Zs == 0.0) ? 0.0 : 1.0 / Zs;
  double Zpi = 1.0 / Zp;
  double Fd[6];
  int subtract = !(bc == FDM_BC_MMS_SIN_W || bc == FDM_BC_RIGID_0 || bc == FDM_BC_ABSORBING_0 ||
                   bc == FDM_BC_FREE_0);
  for (int l = 0; l < 6; l++) Fd[l] = subtract ? F[l] - F0[l] : F[l];
  nt_t fd = xy2nt(Fd, n1, n2);
  double vt = 0, vn = 0, snt = 0, snn = 0, stt, szz;
  switch (bc) {
    case FDM_BC_MMS_SIN_W: {
      double FEX[6];
      for (int l = 0; l < 5; l++) FEX[l] = orc_mms_sin(x, y, t, iblock, l);
      FEX[5] = 0.0;
      nt_t ex = xy2nt(FEX, n1, n2);
      vt = 0.5 * (ex.vt + fd.vt + Zsi * (ex.snt - fd.snt));
      snt = 0.5 * (ex.snt + fd.snt + Zs * (ex.vt - fd.vt));
      vn = 0.5 * (ex.vn + fd.vn + Zpi * (ex.snn - fd.snn));
      snn = 0.5 * (ex.snn + fd.snn + Zp * (ex.vn - fd.vn));
      break;
    }
    case FDM_BC_ABSORBING:
    case FDM_BC_ABSORBING_0:
      vt = 0.5 * (fd.vt - Zsi * fd.snt);
      snt = 0.5 * (fd.snt - Zs * fd.vt);
      vn = 0.5 * (fd.vn - Zpi * fd.snn);
      snn = 0.5 * (fd.snn - Zp * fd.vn);
      break;
    case FDM_BC_FREE:
    case FDM_BC_FREE_0:
      vt = fd.vt - Zsi * fd.snt;
      snt = 0.0;
      vn = fd.vn - Zpi * fd.snn;
      snn = 0.0;
      break;
    case FDM_BC_RIGID:
    case FDM_BC_RIGID_0:
      vt = 0.0;
      snt = fd.snt - Zs * fd.vt;
      vn = 0.0;
      snn = fd.snn - Zp * fd.vn;
      break;
    case FDM_BC_ABSORBING_VELOCITY:
      vt = -Zsi * fd.snt;
      snt = fd.snt;
      vn = -Zpi * fd.snn;
      snn = fd.snn;
      break;
    case FDM_BC_FREE_VELOCITY:
      vt = fd.vt - Zsi * fd.snt;
      snt = fd.snt;
      vn = fd.vn - Zpi * fd.snn;
      snn = fd.snn;
      break;
    default:
      break;
  }
  stt = fd.stt - gamma * (fd.snn - snn);
  szz = fd.szz - gamma * (fd.snn - snn);
  nt2xy(Fhat, n1, n2, vt, vn, stt, snt, snn, szz);
  if (subtract)
    for (int l = 0; l < 6; l++) Fhat[l] = Fhat[l] + F0[l];
}

/* antiplane shear: rotate_fields_xy2nt_mode3 (fields.f90:1297-1309), vz
   and the vector (sxz, syz) (rotate_xy2nt_vec, geometry.f90:285-295) */
static inline void xy2nt3(const double* F, double n1, double n2, double* vz, double* stz,
                          double* snz) {
  *vz = F[0];
  *stz = n2 * F[1] - n1 * F[2];
  *snz = n1 * F[1] + n2 * F[2];
}

/* rotate_fields_nt2xy_mode3 (fields.f90:1328-1340, geometry.f90:298-308) */
static inline void nt2xy3(double* F, double n1, double n2, double vz, double stz, double snz) {
  F[0] = vz;
  F[1] = n2 * stz + n1 * snz;
  F[2] = -n1 * stz + n2 * snz;
}

/* set_bc_mode3 (boundaries.f90:155-213) */
static void set_bc_mode3(double* Fhat, const double* F, const double* F0, double n1, double n2,
                         int bc, double Zs) {
  if (Zs == 0.0) {
    for (int l = 0; l < 3; l++) Fhat[l] = F[l];
    return;
  }
  double Zsi = 1.0 / Zs;
  double d[3], vzFD, stzFD, snzFD, vz = 0.0, snz = 0.0;
  for (int l = 0; l < 3; l++) d[l] = F[l] - F0[l];
  xy2nt3(d, n1, n2, &vzFD, &stzFD, &snzFD);
  switch (bc) {
    case FDM_BC_ABSORBING:
      vz = 0.5 * (vzFD - Zsi * snzFD);
      snz = 0.5 * (snzFD - Zs * vzFD);
      break;
    case FDM_BC_FREE:
      vz = vzFD - Zsi * snzFD;
      snz = 0.0;
      break;
    case FDM_BC_RIGID:
      vz = 0.0;
      snz = snzFD - Zs * vzFD;
      break;
    default:
      break;
  }
  nt2xy3(Fhat, n1, n2, vz, stzFD, snz);
  for (int l = 0; l < 3; l++) Fhat[l] = Fhat[l] + F0[l];
}

/* boundaries.f90:74-129 */
static void enforce_boundary_conditions(orc_ctx* c, double t) {
  for (int ib = 0; ib < c->nb; ib++) {
    for (int s = 0; s < 4; s++) {
      oside_t* sd = &c->B[ib].s[s];
      if (sd->bc == FDM_BC_NONE) continue;
      int n = sd->n;
      for (int k = 0; k < n; k++) {
        double F[6], F0[6], Fh[6];
        for (int l = 0; l < c->nF; l++) {
          F[l] = sd->F[(size_t)l * n + k];
          F0[l] = sd->F0[(size_t)l * n + k];
        }
        if (c->mode == 3) {
          set_bc_mode3(Fh, F, F0, sd->normal[k], sd->normal[n + k], sd->bc, sd->M[k]);
          for (int l = 0; l < 3; l++) sd->Fhat[(size_t)l * n + k] = Fh[l];
          continue;
        }
        set_bc_mode2(Fh, F, F0, sd->normal[k], sd->normal[n + k], sd->bc, sd->M[k],
                     sd->M[n + k], sd->M[2 * (size_t)n + k], sd->x ? sd->x[k] : 0.0,
                     sd->y ? sd->y[k] : 0.0, t, c->B[ib].b.id);
        for (int l = 0; l < 6; l++) sd->Fhat[(size_t)l * n + k] = Fh[l];
      }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* friction, friction.f90:593-1304                                         */
/* ---------------------------------------------------------------------- */

/* strength (friction.f90:1041-1074) */
static void strength(int law, const double* rs, double V, double Psi, double N, double* S,
                     double* dSdV) {
  double a = rs[0], V0 = rs[2];
  if (law == FDM_LAW_SL || law == FDM_LAW_FL) {
    if (V == 0.0) {
      *S = 0.0;
      *dSdV = 1e10;
    } else {
      double absV = fabs(V);
      double f = a * log(absV / V0) + Psi;
      *S = N * fsign(f, V);
      *dSdV = N * a / absV;
    }
  } else {
    double O = 0.5 / V0 * exp(Psi / a);
    *S = a * N * arcsinh_ref(O * V);
    *dSdV = a * N * O / sqrt(1.0 + (V * O) * (V * O));
  }
}

/* estimate_V (friction.f90:1077-1098) */
static void estimate_V(int law, const double* rs, double* V, double Slock, double N, double Psi) {
  double a = rs[0], V0 = rs[2];
  double SaN = Slock / (a * N);
  if (fabs(SaN) > 10.0) return;
  if (law == FDM_LAW_SL || law == FDM_LAW_FL) {
    *V = fsign(V0 * exp(-Psi / a) * exp(fabs(SaN)), SaN);
  } else if (law == FDM_LAW_RSL || law == FDM_LAW_RFL || law == FDM_LAW_RAL) {
    *V = 2.0 * V0 * exp(-Psi / a) * sinh(SaN);
  }
}

/* newton_solver (friction.f90:909-1038); returns 1 on failure */
static int newton_solver(int law, const double* rs, double* Vio, double* Sio, double Slock,
                         double N, double eta, double Psi) {
  const int imax = 99;
  const double atolR = 2.220446049250313e-16 * 1e4; /* epsilon(1d0)*1d4 */
  double V = *Vio, S = *Sio;
  double Vmin = fmin(0.0, Slock / eta);
  double Vmax = fmax(0.0, Slock / eta);
  /* Fortran MIN/MAX with 0d0 literal: the result for equal arguments is
     either; values are equal, so fmin/fmax agree. */
  estimate_V(law, rs, &V, Slock, N, Psi);
  if (V <= Vmin || Vmax <= V) V = 0.5 * (Vmin + Vmax);
  int i;
  double Sf, dSfdV, R, dRdV, dV, dSdV;
  for (i = 1; i <= imax; i++) {
    S = Slock - eta * V;
    dSdV = -eta;
    strength(law, rs, V, Psi, N, &Sf, &dSfdV);
    R = -S + Sf;
    if (R > 0.0) Vmax = V;
    if (R < 0.0) Vmin = V;
    dRdV = -dSdV + dSfdV;
    dV = -R / dRdV;
    if (fabs(R) <= atolR) break;
    if (i == imax) break;
    V = V + dV;
    if (V <= Vmin || Vmax <= V) {
      if (Vmin == 0.0)
        V = 0.1 * Vmax;
      else if (Vmax == 0.0)
        V = 0.1 * Vmin;
      else
        V = fsign(sqrt(Vmin * Vmax), Slock);
      dV = Vmax - Vmin;
    }
    if (fabs(dV) <= 0.0 + 0.0 * (fabs(V) + fabs(dV))) break;
  }
  *Vio = V;
  *Sio = S;
  return i == imax;
}

/* Slock/Nlock part and law branches of solve_friction (friction.f90:655-826)
   once FR%S0(i), FR%N0(i) are known.  Returns 1 on Newton failure. */
/* kinematic forcing (friction.f90:717-745): the forced friction
   max(fkm, fkp) at distance xx along the fault and time t; kn = c0, fd,
   dfdx, vL, vR, xL, xR, tmax */
static double kinematic_friction(const double* kn, double xx, double t) {
  double xm = kn[5] + kn[3] * fmin(t, kn[7]);
  double fkm = xx < xm ? kn[1] + kn[2] * (xm - xx) : kn[1];
  double xp = kn[6] + kn[4] * fmin(t, kn[7]);
  double fkp = xx > xp ? kn[1] + kn[2] * (xx - xp) : kn[1];
  return fmax(fkm, fkp);
}

/* solve_friction (friction.f90:593-828).  force: the forced strength
   Sk = kc0 + Nreg kf caps the slip-weakening strength (:765-767) and
   replaces a larger rate-and-state strength (:808-815) */
static int solve_friction_core_f(int law, const double* rs, const double* sw, int reg,
                                 int outside, double flock, double S0i, double N0i,
                                 double phip, double phis, double eta, double D, double Psi,
                                 double skempton, double dt, double* V, double* S, double* O,
                                 double* N, double* Nreg, int force, double kc0, double kf) {
  double Slock = S0i + phis;
  double Nlock = N0i - (1.0 - skempton) * phip;
  *O = 0.0;
  *N = Nlock;
  if (Nlock <= 0.0) {
    *S = 0.0;
    *V = Slock / eta;
    return 0;
  }
  if (dt == 0.0) *Nreg = *N;
  if (!reg) *Nreg = *N;
  if (outside) {
    double Sv = flock * (*Nreg);
    if (fabs(Slock) >= Sv) {
      Sv = fsign(Sv, Slock);
      *V = (Slock - Sv) / eta;
      *S = Sv;
    } else {
      *V = 0.0;
      *S = Slock;
    }
    return 0;
  }
  switch (law) {
    case FDM_LAW_FRICTIONLESS:
      *S = 0.0;
      *V = Slock / eta;
      return 0;
    case FDM_LAW_LOCKED:
      *V = 0.0;
      *S = Slock;
      return 0;
    case FDM_LAW_SW: {
      double c0 = sw[0], fs = sw[1], fd = sw[2], Dc = sw[3];
      double Sv = c0 + fmax(*Nreg, 0.0) * ((fs - (fs - fd) * fmin(D, Dc) / Dc));
      if (force) Sv = fmin(Sv, kc0 + *Nreg * kf);
      if (fabs(Slock) >= Sv) {
        Sv = fsign(Sv, Slock);
        *V = (Slock - Sv) / eta;
        *S = Sv;
      } else {
        *V = 0.0;
        *S = Slock;
      }
      return 0;
    }
    default: {
      int fail = newton_solver(law, rs, V, S, Slock, *Nreg, eta, Psi);
      if (fail) fail = newton_solver(law, rs, V, S, Slock, *Nreg, eta, Psi);
      if (force) {
        double Sk = kc0 + *Nreg * kf;
        if (Sk < fabs(*S)) {
          *S = fsign(Sk, Slock);
          *V = (Slock - *S) / eta;
        }
      }
      return fail;
    }
  }
}

static int solve_friction_core(int law, const double* rs, const double* sw, int reg,
                               int outside, double flock, double S0i, double N0i,
                               double phip, double phis, double eta, double D, double Psi,
                               double skempton, double dt, double* V, double* S, double* O,
                               double* N, double* Nreg) {
  return solve_friction_core_f(law, rs, sw, reg, outside, flock, S0i, N0i, phip, phis, eta, D,
                               Psi, skempton, dt, V, S, O, N, Nreg, 0, 0.0, 0.0);
}

int orc_solve_friction_node(int32_t law, const double* rs, const double* sw, double S0,
                            double N0, double phip, double phis, double eta, double D,
                            double Psi, double skempton, double dt, double* V, double* S,
                            double* N, double* O) {
  double Nreg = 1e40;
  int fail = solve_friction_core(law, rs, sw, 0, 0, 1e10, S0, N0, phip, phis, eta, D, Psi,
                                 skempton, dt, V, S, O, N, &Nreg);
  return fail ? FDM_ENEWTON : FDM_OK;
}

/* state_rate (friction.f90:1147-1264) */
double orc_state_rate(int32_t law, const double* rs, double V, double Psi) {
  if (law == FDM_LAW_FRICTIONLESS || law == FDM_LAW_SW || law == FDM_LAW_LOCKED) return 0.0;
  if (V == 0.0) return 0.0;
  double a = rs[0], b = rs[1], V0 = rs[2], f0 = rs[3], L = rs[4], fw = rs[5], Vw = rs[6];
  double absV = fabs(V);
  double frs, fss, fLV, Psiss;
  switch (law) {
    case FDM_LAW_SL:
      frs = a * log(absV / V0) + Psi;
      fss = f0 - (b - a) * log(absV / V0);
      return -(absV / L) * (frs - fss);
    case FDM_LAW_FL:
      frs = a * log(absV / V0) + Psi;
      fLV = f0 - (b - a) * log(absV / V0);
      fss = fw + (fLV - fw) / pow(1.0 + pow8(V / Vw), 0.125);
      return -(absV / L) * (frs - fss);
    case FDM_LAW_RSL:
      fss = f0 - (b - a) * log(absV / V0);
      frs = a * arcsinh_ref(0.5 * absV / V0 * exp(Psi / a));
      return -(absV / L) * (frs - fss);
    case FDM_LAW_RFL:
      fLV = f0 - (b - a) * log(absV / V0);
      fss = fw + (fLV - fw) / pow(1.0 + pow8(V / Vw), 0.125);
      frs = a * arcsinh_ref(0.5 * absV / V0 * exp(Psi / a));
      return (absV / L) * (fss - frs);
    case FDM_LAW_RSF:
      fLV = f0 - (b - a) * log(absV / V0);
      fss = fw + (fLV - fw) / pow(1.0 + pow8(V / Vw), 0.125);
      Psiss = a * log(2.0 * V0 / absV * sinh(fss / a));
      return (absV / L) * (Psiss - Psi);
    case FDM_LAW_RAL:
      return (b * V0 / L) * (exp((f0 - Psi) / b) - absV / V0);
    default:
      return 0.0;
  }
}

/* load_ramp (friction.f90:1455-1484) */
static double load_ramp(double t, double T) {
  if (T <= 0.0) return (t < 0.0) ? 0.0 : 1.0;
  return (t < 0.0) ? 0.0 : 1.0 - exp(-t / T);
}

/* interfaces.f90:715-772 */
static void locked_interface_mode2(double* Fhatm, double* Fhatp, const double* Fm,
                                   const double* Fp, double n1, double n2, double Zsm,
                                   double Zsp, double Zpm, double Zpp, double gammam,
                                   double gammap) {
  nt_t p = xy2nt(Fp, n1, n2);
  nt_t m = xy2nt(Fm, n1, n2);
  double Zsip = 1.0 / Zsp, Zsim = 1.0 / Zsm, Zpip = 1.0 / Zpp, Zpim = 1.0 / Zpm;
  double etap = 1.0 / (Zpip + Zpim);
  double phip = etap * (p.snn * Zpip + m.snn * Zpim + p.vn - m.vn);
  double snnp = phip, snnm = phip;
  double vnp = p.vn + Zpip * (p.snn - snnp);
  double vnm = m.vn - Zpim * (m.snn - snnm);
  double sttp = p.stt - gammap * (p.snn - snnp);
  double sttm = m.stt - gammam * (m.snn - snnm);
  double szzp = p.szz - gammap * (p.snn - snnp);
  double szzm = m.szz - gammam * (m.snn - snnm);
  double etas = 1.0 / (Zsip + Zsim);
  double phis = etas * (p.snt * Zsip + m.snt * Zsim + p.vt - m.vt);
  double sntp = phis, sntm = phis;
  double vtp = p.vt + Zsip * (p.snt - sntp);
  double vtm = m.vt - Zsim * (m.snt - sntm);
  nt2xy(Fhatp, n1, n2, vtp, vnp, sttp, sntp, snnp, szzp);
  nt2xy(Fhatm, n1, n2, vtm, vnm, sttm, sntm, snnm, szzm);
}

/* state for one frictional node */
typedef struct {
  double *V, *S, *O, *N, *Nreg, *S0, *N0;
} frnode_t;

/* interfaces.f90:900-1012 + solve_friction (friction.f90:593-828) with
   load_stress (friction.f90:1335-1374).  Dummy names follow the reference:
   "F0p" receives the minus-side initial fields and "F0m" the plus side
   (the call at interfaces.f90:795 passes F0m(i,:),F0p(i,:) to dummies
   declared F0p,F0m at :900). */
static int frictional_interface_mode2(orc_ctx* c, oiface_t* I, int k, double* Fhatm,
                                      double* Fhatp, const double* Fm, const double* Fp,
                                      const double* F0p, const double* F0m, double n1,
                                      double n2, double Zsm, double Zsp, double Zpm,
                                      double Zpp, double gammam, double gammap, double t,
                                      double dt) {
  int n = I->n;
  nt_t p0 = xy2nt(F0p, n1, n2);
  double dp[6], dm[6];
  for (int l = 0; l < 6; l++) {
    dp[l] = Fp[l] - F0p[l];
    dm[l] = Fm[l] - F0m[l];
  }
  nt_t p = xy2nt(dp, n1, n2);
  nt_t m = xy2nt(dm, n1, n2);
  double Zsip = 1.0 / Zsp, Zsim = 1.0 / Zsm, Zpip = 1.0 / Zpp, Zpim = 1.0 / Zpm;
  double etas = 1.0 / (Zsip + Zsim);
  double phis = etas * (p.snt * Zsip + m.snt * Zsim + p.vt - m.vt);
  double etap = 1.0 / (Zpip + Zpim);
  double phip = etap * (p.snn * Zpip + m.snn * Zpim + p.vn - m.vn);
  double S0 = p0.snt;
  double pp = 0.0;
  double N0 = -p0.snn - pp;
  /* solve_friction: load_stress (friction.f90:1339-1374) */
  double Bl = load_ramp(t, I->I.ld_T);
  double S0l = I->S0base[k] + I->ldS[k] * Bl;
  double N0l = I->N0base[k] + I->ldN[k] * Bl;
  double* fS0 = &I->f[FDM_IF_S0][k];
  double* fN0 = &I->f[FDM_IF_N0][k];
  *fS0 = S0l + S0;
  *fN0 = N0l + N0;
  const double* rs = NULL;
  double rsk[7], swk[4];
  if (I->rs) {
    for (int q = 0; q < 7; q++) rsk[q] = I->rs[(size_t)q * n + k];
    rs = rsk;
  }
  if (I->sw)
    for (int q = 0; q < 4; q++) swk[q] = I->sw[(size_t)q * n + k];
  int force = I->knx != NULL;
  double kf = force ? kinematic_friction(I->I.kn, I->knx[k], t) : 0.0;
  int fail = solve_friction_core_f(I->I.law, rs, swk, I->I.reg, I->outside[k], I->I.flock, *fS0,
                                   *fN0, phip, phis, etas, I->f[FDM_IF_D][k],
                                   I->f[FDM_IF_PSI][k], I->I.skempton, dt, &I->f[FDM_IF_V][k],
                                   &I->f[FDM_IF_S][k], &I->f[FDM_IF_O][k], &I->f[FDM_IF_N][k],
                                   &I->f[FDM_IF_NREG][k], force, I->I.kn[0], kf);
  double skm = 1.0 - I->I.skempton;
  double snnp = -(I->f[FDM_IF_N][k] - *fN0) / skm;
  double snnm = -(I->f[FDM_IF_N][k] - *fN0) / skm;
  double vnp = p.vn + Zpip * (p.snn - snnp);
  double vnm = m.vn - Zpim * (m.snn - snnm);
  double sttp = p.stt - gammap * (p.snn - snnp);
  double sttm = m.stt - gammam * (m.snn - snnm);
  double szzp = p.szz - gammap * (p.snn - snnp);
  double szzm = m.szz - gammam * (m.snn - snnm);
  double sntp = I->f[FDM_IF_S][k] - *fS0;
  double sntm = I->f[FDM_IF_S][k] - *fS0;
  double vtp = p.vt + Zsip * (p.snt - sntp);
  double vtm = m.vt - Zsim * (m.snt - sntm);
  nt2xy(Fhatp, n1, n2, vtp, vnp, sttp, sntp, snnp, szzp);
  nt2xy(Fhatm, n1, n2, vtm, vnm, sttm, sntm, snnm, szzm);
  for (int l = 0; l < 6; l++) {
    Fhatp[l] = Fhatp[l] + F0p[l];
    Fhatm[l] = Fhatm[l] + F0m[l];
  }
  return fail;
}

/* locked_interface_mode3 (interfaces.f90:676-712) */
static void locked_interface_mode3(double* Fhatm, double* Fhatp, const double* Fm,
                                   const double* Fp, double n1, double n2, double Zsm,
                                   double Zsp) {
  double vzFDp, stzFDp, snzFDp, vzFDm, stzFDm, snzFDm;
  xy2nt3(Fp, n1, n2, &vzFDp, &stzFDp, &snzFDp);
  xy2nt3(Fm, n1, n2, &vzFDm, &stzFDm, &snzFDm);
  double Zsip = 1.0 / Zsp, Zsim = 1.0 / Zsm;
  double etas = 1.0 / (Zsip + Zsim);
  double phis = etas * (snzFDp * Zsip + snzFDm * Zsim + vzFDp - vzFDm);
  double snzp = phis, snzm = phis;
  double vzp = vzFDp + Zsip * (snzFDp - snzp);
  double vzm = vzFDm - Zsim * (snzFDm - snzm);
  nt2xy3(Fhatp, n1, n2, vzp, stzFDp, snzp);
  nt2xy3(Fhatm, n1, n2, vzm, stzFDm, snzm);
}

/* frictional_interface_mode3 (interfaces.f90:806-897) with solve_friction.
   The arguments arrive in their own order here (interfaces.f90:798).  The
   normal stress comes from the plus block's in-plane prestress F0(4), F0(5),
   F0(7) resolved on the fault (rotate_xy2nt_mat, geometry.f90:311-322); the
   stress transfer phip is zero. */
static int frictional_interface_mode3(orc_ctx* c, oiface_t* I, int k, double* Fhatm,
                                      double* Fhatp, const double* Fm, const double* Fp,
                                      const double* F0m, const double* F0p, double n1, double n2,
                                      double Zsm, double Zsp, const double* s0p, double t,
                                      double dt) {
  (void)c;
  int n = I->n;
  double vz0p, stz0p, snz0p, vz0m, stz0m, snz0m;
  xy2nt3(F0p, n1, n2, &vz0p, &stz0p, &snz0p);
  xy2nt3(F0m, n1, n2, &vz0m, &stz0m, &snz0m);
  double dp[3], dm[3];
  for (int l = 0; l < 3; l++) {
    dp[l] = Fp[l] - F0p[l];
    dm[l] = Fm[l] - F0m[l];
  }
  double vzFDp, stzFDp, snzFDp, vzFDm, stzFDm, snzFDm;
  xy2nt3(dp, n1, n2, &vzFDp, &stzFDp, &snzFDp);
  xy2nt3(dm, n1, n2, &vzFDm, &stzFDm, &snzFDm);
  double Zsip = 1.0 / Zsp, Zsim = 1.0 / Zsm;
  double etas = 1.0 / (Zsip + Zsim);
  double phis = etas * (snzFDp * Zsip + snzFDm * Zsim + vzFDp - vzFDm);
  double S0 = snz0p;
  double snn = 2.0 * n1 * n2 * s0p[1] + n1 * n1 * s0p[0] + n2 * n2 * s0p[3];
  double pp = 0.0;
  double N0 = -snn - pp;
  double phip = 0.0;
  double Bl = load_ramp(t, I->I.ld_T);
  double S0l = I->S0base[k] + I->ldS[k] * Bl;
  double N0l = I->N0base[k] + I->ldN[k] * Bl;
  double* fS0 = &I->f[FDM_IF_S0][k];
  double* fN0 = &I->f[FDM_IF_N0][k];
  *fS0 = S0l + S0;
  *fN0 = N0l + N0;
  const double* rs = NULL;
  double rsk[7], swk[4];
  if (I->rs) {
    for (int q = 0; q < 7; q++) rsk[q] = I->rs[(size_t)q * n + k];
    rs = rsk;
  }
  if (I->sw)
    for (int q = 0; q < 4; q++) swk[q] = I->sw[(size_t)q * n + k];
  int force = I->knx != NULL;
  double kf = force ? kinematic_friction(I->I.kn, I->knx[k], t) : 0.0;
  int fail = solve_friction_core_f(I->I.law, rs, swk, I->I.reg, I->outside[k], I->I.flock, *fS0,
                                   *fN0, phip, phis, etas, I->f[FDM_IF_D][k],
                                   I->f[FDM_IF_PSI][k], I->I.skempton, dt, &I->f[FDM_IF_V][k],
                                   &I->f[FDM_IF_S][k], &I->f[FDM_IF_O][k], &I->f[FDM_IF_N][k],
                                   &I->f[FDM_IF_NREG][k], force, I->I.kn[0], kf);
  double snzp = I->f[FDM_IF_S][k] - *fS0;
  double snzm = I->f[FDM_IF_S][k] - *fS0;
  double vzp = vzFDp + Zsip * (snzFDp - snzp);
  double vzm = vzFDm - Zsim * (snzFDm - snzm);
  nt2xy3(Fhatp, n1, n2, vzp, stzFDp, snzp);
  nt2xy3(Fhatm, n1, n2, vzm, stzFDm, snzm);
  for (int l = 0; l < 3; l++) {
    Fhatp[l] = Fhatp[l] + F0p[l];
    Fhatm[l] = Fhatm[l] + F0m[l];
  }
  return fail;
}

/* set_rates_friction (friction.f90:1101-1124) for one node */
static void set_rates_friction_node(oiface_t* I, int k, double t) {
  int n = I->n;
  double rsk[7];
  const double* rs = NULL;
  if (I->rs) {
    for (int q = 0; q < 7; q++) rsk[q] = I->rs[(size_t)q * n + k];
    rs = rsk;
  }
  double V = I->f[FDM_IF_V][k];
  I->f[FDM_IF_DDS][k] = I->f[FDM_IF_DDS][k] + V;
  I->f[FDM_IF_DDN][k] = I->f[FDM_IF_DDN][k] + I->f[FDM_IF_O][k];
  I->f[FDM_IF_DPSI][k] =
      I->f[FDM_IF_DPSI][k] + orc_state_rate(I->I.law, rs, V, I->f[FDM_IF_PSI][k]);
  if (I->I.reg) {
    /* regularize_N (friction.f90:1127-1144) */
    double absV = fabs(V);
    double Nr = -(absV / rs[4]) * (I->f[FDM_IF_NREG][k] - fmax(I->f[FDM_IF_N][k], 0.0));
    I->f[FDM_IF_DNREG][k] = I->f[FDM_IF_DNREG][k] + Nr;
  }
  I->f[FDM_IF_DW][k] = I->f[FDM_IF_DW][k] + I->f[FDM_IF_S][k] * V;
  /* rupture_front (friction.f90:1267-1290) */
  int ruptured = I->I.rup_field == 1 ? (I->f[FDM_IF_D][k] > I->I.rup_threshold)
                                     : (fabs(V) > I->I.rup_threshold);
  if (ruptured) I->f[FDM_IF_TRUP][k] = fmin(I->f[FDM_IF_TRUP][k], t);
}

/* enforce_interface_conditions (interfaces.f90:520-649) */
static int enforce_interface_conditions(orc_ctx* c, double t, double dt) {
  int fail = 0;
  for (int ii = 0; ii < c->ni; ii++) {
    oiface_t* I = &c->I[ii];
    int bm = I->I.blockm, bp = I->I.blockp;
    int sm = I->I.direction == 0 ? FDM_SIDE_R : FDM_SIDE_T;
    int sp = I->I.direction == 0 ? FDM_SIDE_L : FDM_SIDE_B;
    oside_t* SM = &c->B[bm].s[sm];
    oside_t* SP = &c->B[bp].s[sp];
    int offm = I->I.direction == 0 ? c->B[bm].b.my : c->B[bm].b.mx;
    int offp = I->I.direction == 0 ? c->B[bp].b.my : c->B[bp].b.mx;
    int n = I->n;
    for (int k = 0; k < n; k++) {
      int km = I->I.m + k - offm, kp = I->I.m + k - offp;
      double Fm[6], Fp[6], F0m[6], F0p[6], Fhm[6], Fhp[6];
      for (int l = 0; l < c->nF; l++) {
        Fm[l] = SM->F[(size_t)l * SM->n + km];
        Fp[l] = SP->F[(size_t)l * SP->n + kp];
        F0m[l] = SM->F0[(size_t)l * SM->n + km];
        F0p[l] = SP->F0[(size_t)l * SP->n + kp];
      }
      double n1 = I->nhat[k], n2 = I->nhat[n + k];
      if (c->mode == 3) {
        double Zsm3 = SM->M[km], Zsp3 = SP->M[kp];
        if (I->I.coupling == FDM_COUPLING_LOCKED)
          locked_interface_mode3(Fhm, Fhp, Fm, Fp, n1, n2, Zsm3, Zsp3);
        else
          fail |= frictional_interface_mode3(c, I, k, Fhm, Fhp, Fm, Fp, F0m, F0p, n1, n2, Zsm3,
                                             Zsp3, c->B[bp].b.s0, t, dt);
        for (int l = 0; l < 3; l++) {
          SM->Fhat[(size_t)l * SM->n + km] = Fhm[l];
          SP->Fhat[(size_t)l * SP->n + kp] = Fhp[l];
        }
        continue;
      }
      double Zsm = SM->M[km], Zsp = SP->M[kp];
      double Zpm = SM->M[SM->n + km], Zpp = SP->M[SP->n + kp];
      double gm = SM->M[2 * (size_t)SM->n + km], gp = SP->M[2 * (size_t)SP->n + kp];
      if (I->I.coupling == FDM_COUPLING_LOCKED) {
        locked_interface_mode2(Fhm, Fhp, Fm, Fp, n1, n2, Zsm, Zsp, Zpm, Zpp, gm, gp);
      } else {
        /* actual F0m goes to dummy F0p and vice versa (interfaces.f90:795,900) */
        fail |= frictional_interface_mode2(c, I, k, Fhm, Fhp, Fm, Fp, F0m, F0p, n1, n2, Zsm, Zsp,
                                           Zpm, Zpp, gm, gp, t, dt);
      }
      for (int l = 0; l < 6; l++) {
        SM->Fhat[(size_t)l * SM->n + km] = Fhm[l];
        SP->Fhat[(size_t)l * SP->n + kp] = Fhp[l];
      }
    }
    if (I->I.coupling == FDM_COUPLING_FRICTION)
      for (int k = 0; k < n; k++) set_rates_friction_node(I, k, t);
  }
  return fail;
}

/* ---------------------------------------------------------------------- */
/* SAT, time_step.f90:754-851                                              */
/* ---------------------------------------------------------------------- */

static void SAT_term(double* DF6, const double* Fhat, const double* F, double n1, double n2,
                     double Ks, double Kp) {
  double d[6], Fs[6], Fp[6];
  for (int l = 0; l < 6; l++) d[l] = Fhat[l] - F[l];
  nt_t r = xy2nt(d, n1, n2);
  nt2xy(Fs, n1, n2, r.vt, 0.0, 0.0, r.snt, 0.0, 0.0);
  nt2xy(Fp, n1, n2, 0.0, r.vn, r.stt, 0.0, r.snn, r.szz);
  for (int l = 0; l < 6; l++) DF6[l] = DF6[l] + Ks * Fs[l];
  for (int l = 0; l < 6; l++) DF6[l] = DF6[l] + Kp * Fp[l];
}

static void set_rates_SAT(orc_ctx* c, int ib) {
  fdm_block* b = &c->B[ib].b;
  for (int s = 0; s < 4; s++) {
    oside_t* sd = &c->B[ib].s[s];
    int n = sd->n;
    for (int k = 0; k < n; k++) {
      int i, j;
      if (s == FDM_SIDE_L) {
        i = b->mx;
        j = b->my + k;
      } else if (s == FDM_SIDE_R) {
        i = b->px;
        j = b->my + k;
      } else if (s == FDM_SIDE_B) {
        i = b->mx + k;
        j = b->my;
      } else {
        i = b->mx + k;
        j = b->py;
      }
      double DF6[6], Fh[6], Fe[6];
      size_t q = IDX(i, j);
      if (c->mode == 3) {
        /* SAT_term, mode 3: DF = DF+Ks*(Fhat-F) (time_step.f90:829-830) */
        double Ks = sd->M[3 * (size_t)n + k];
        for (int l = 0; l < 3; l++)
          DFL(l)[q] = DFL(l)[q] + Ks * (sd->Fhat[(size_t)l * n + k] - sd->F[(size_t)l * n + k]);
        continue;
      }
      for (int l = 0; l < 6; l++) {
        DF6[l] = DFL(l)[q];
        Fh[l] = sd->Fhat[(size_t)l * n + k];
        Fe[l] = sd->F[(size_t)l * n + k];
      }
      SAT_term(DF6, Fh, Fe, sd->normal[k], sd->normal[n + k], sd->M[3 * (size_t)n + k],
               sd->M[4 * (size_t)n + k]);
      for (int l = 0; l < 6; l++) DFL(l)[q] = DF6[l];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* stencils, rates.f90:384-936                                             */
/* ---------------------------------------------------------------------- */

/* field coupling of rates_*_mode2: DF(l) uses F(UA) with metric y? and
   F(UB) with metric x?, gamma on the A or B product */
static const int UA[5] = {2, 3, 0, 1, 0};
static const int UB[5] = {3, 4, 1, 0, 1};
static const int GA[5] = {0, 0, 0, 0, 1};
static const int GB[5] = {0, 0, 1, 0, 0};

static inline double prodg(double a, double f, int g, double gamma) {
  return g ? a * f * gamma : a * f;
}

/* material scaling of the one-sided and central boundary terms:
   ")*D/rho*Ji", ")*D*M*Ji", ")*D*G*Ji" */
static inline double bscale(int l, double t, double D, double rho, double G, double M) {
  if (l < 2) return t * D / rho;
  if (l == 3) return t * D * G;
  return t * D * M;
}

typedef struct {
  const double *F[5], *xq, *xr, *yq, *yr, *Ji;
  double *DF[5];
  double rho, G, M, gamma;
  /* heterogeneous medium: material at each point (NULL: the block's) */
  const double *prho, *pG, *pM, *pgam;
  /* artificial dissipation: Jacobian and the scaled Cdiss (0: none) */
  const double* J;
  double Cdiss;
  int nx;
} sten_t;

/* the material at centre cc: the block's, or the medium's */
static inline void mat_at(const sten_t* S, size_t cc, double* rho, double* G, double* M,
                          double* gamma) {
  if (S->prho) {
    *rho = S->prho[cc];
    *G = S->pG[cc];
    *M = S->pM[cc];
    *gamma = S->pgam[cc];
  } else {
    *rho = S->rho;
    *G = S->G;
    *M = S->M;
    *gamma = S->gamma;
  }
}

/* one-sided q term at neighbour kk (same row) of centre cc */
static inline double q1(const sten_t* S, int l, size_t kk, size_t cc, double g) {
  double ay = 0.5 * (S->yr[kk] + S->yr[cc]);
  double ax = 0.5 * (S->xr[kk] + S->xr[cc]);
  return prodg(ay, S->F[UA[l]][kk], GA[l], g) - prodg(ax, S->F[UB[l]][kk], GB[l], g);
}
/* one-sided r term at neighbour kk (same column) of centre cc */
static inline double r1(const sten_t* S, int l, size_t kk, size_t cc, double g) {
  double by = 0.5 * (S->yq[kk] + S->yq[cc]);
  double bx = 0.5 * (S->xq[kk] + S->xq[cc]);
  return -prodg(by, S->F[UA[l]][kk], GA[l], g) + prodg(bx, S->F[UB[l]][kk], GB[l], g);
}
/* central q (4 terms) */
static inline double qc(const sten_t* S, int l, size_t kp, size_t km, size_t cc, double g) {
  double ayp = 0.5 * (S->yr[kp] + S->yr[cc]), axp = 0.5 * (S->xr[kp] + S->xr[cc]);
  double aym = 0.5 * (S->yr[km] + S->yr[cc]), axm = 0.5 * (S->xr[km] + S->xr[cc]);
  return prodg(ayp, S->F[UA[l]][kp], GA[l], g) - prodg(axp, S->F[UB[l]][kp], GB[l], g) -
         prodg(aym, S->F[UA[l]][km], GA[l], g) + prodg(axm, S->F[UB[l]][km], GB[l], g);
}
/* central r (4 terms) */
static inline double rc(const sten_t* S, int l, size_t kp, size_t km, size_t cc, double g) {
  double byp = 0.5 * (S->yq[kp] + S->yq[cc]), bxp = 0.5 * (S->xq[kp] + S->xq[cc]);
  double bym = 0.5 * (S->yq[km] + S->yq[cc]), bxm = 0.5 * (S->xq[km] + S->xq[cc]);
  return -prodg(byp, S->F[UA[l]][kp], GA[l], g) + prodg(bxp, S->F[UB[l]][kp], GB[l], g) +
         prodg(bym, S->F[UA[l]][km], GA[l], g) - prodg(bxm, S->F[UB[l]][km], GB[l], g);
}

/* coefficient of the elastic sum in the dissipative interior rates:
   a(s) = DI(s)/rho, c(s) = DI(s)*M, b(s) = DI(s)*G (rates.f90:1416-1421),
   per point in a heterogeneous medium (rates_heterogeneous.f90:1409-1450) */
static inline double diss_coef(int l, double D, double rho, double G, double M) {
  if (l < 2) return D / rho;
  if (l == 3) return D * G;
  return D * M;
}

/* rates_dissipation_interior_mode2 (rates.f90:1397-1482) and its
   heterogeneous form (rates_heterogeneous.f90:1376-1458): d(0) F first, then
   per s (a(s) X + d(s) Y) Ji with Y the J-weighted neighbours.  DI(3) of
   the heterogeneous loop lies outside DI(-2:2) for SBP4 (the homogeneous
   routine zeroes its s = 3 coefficient); the s = 3 term is taken as zero. */
static void rates_dissipation_interior(const sten_t* S, int mx, int px, int my, int py) {
  double d[4];
  d[0] = 2.0 * AIc(0) * S->Cdiss;
  for (int s = 1; s <= 3; s++) d[s] = s <= PI_I ? AIc(s) * S->Cdiss : 0.0;
  int nx = S->nx;
  const double* J = S->J;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = my; j <= py; j++)
    for (int i = mx; i <= px; i++) {
      size_t cc = (size_t)j * nx + i;
      double rho, G, M, g;
      mat_at(S, cc, &rho, &G, &M, &g);
      for (int l = 0; l < 5; l++) S->DF[l][cc] = S->DF[l][cc] + d[0] * S->F[l][cc];
      for (int s = 1; s <= PI_I; s++)
        for (int l = 0; l < 5; l++) {
          size_t ip = cc + s, im = cc - s, jp = cc + (size_t)s * nx, jm = cc - (size_t)s * nx;
          double ayp = 0.5 * (S->yr[ip] + S->yr[cc]), axp = 0.5 * (S->xr[ip] + S->xr[cc]);
          double aym = 0.5 * (S->yr[im] + S->yr[cc]), axm = 0.5 * (S->xr[im] + S->xr[cc]);
          double byp = 0.5 * (S->yq[jp] + S->yq[cc]), bxp = 0.5 * (S->xq[jp] + S->xq[cc]);
          double bym = 0.5 * (S->yq[jm] + S->yq[cc]), bxm = 0.5 * (S->xq[jm] + S->xq[cc]);
          const double* FA = S->F[UA[l]];
          const double* FB = S->F[UB[l]];
          const double* Fl = S->F[l];
          double x = prodg(ayp, FA[ip], GA[l], g) - prodg(axp, FB[ip], GB[l], g) -
                     prodg(aym, FA[im], GA[l], g) + prodg(axm, FB[im], GB[l], g) -
                     prodg(byp, FA[jp], GA[l], g) + prodg(bxp, FB[jp], GB[l], g) +
                     prodg(bym, FA[jm], GA[l], g) - prodg(bxm, FB[jm], GB[l], g);
          double y = J[ip] * Fl[ip] + J[im] * Fl[im] + J[jp] * Fl[jp] + J[jm] * Fl[jm];
          double a = diss_coef(l, s <= PI_I ? DIc(s) : 0.0, rho, G, M);
          S->DF[l][cc] = S->DF[l][cc] + (a * x + d[s] * y) * S->Ji[cc];
        }
    }
}

/* rates_interior_mode2, rates.f90:384-450 */
static void rates_interior(const sten_t* S, int mx, int px, int my, int py) {
  if (S->Cdiss > 0.0) {
    rates_dissipation_interior(S, mx, px, my, py);
    return;
  }
  double coef[5][4];
  for (int s = 1; s <= 3; s++) {
    double D = s <= PI_I ? DIc(s) : 0.0;
    coef[0][s] = D / S->rho;
    coef[1][s] = D / S->rho;
    coef[2][s] = D * S->M;
    coef[3][s] = D * S->G;
    coef[4][s] = D * S->M;
  }
  int nx = S->nx;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = my; j <= py; j++)
    for (int i = mx; i <= px; i++) {
      size_t cc = (size_t)j * nx + i;
      double rho, G, M, g;
      mat_at(S, cc, &rho, &G, &M, &g);
      for (int s = 1; s <= PI_I; s++)
        for (int l = 0; l < 5; l++) {
          size_t ip = cc + s, im = cc - s, jp = cc + (size_t)s * nx, jm = cc - (size_t)s * nx;
          double ayp = 0.5 * (S->yr[ip] + S->yr[cc]), axp = 0.5 * (S->xr[ip] + S->xr[cc]);
          double aym = 0.5 * (S->yr[im] + S->yr[cc]), axm = 0.5 * (S->xr[im] + S->xr[cc]);
          double byp = 0.5 * (S->yq[jp] + S->yq[cc]), bxp = 0.5 * (S->xq[jp] + S->xq[cc]);
          double bym = 0.5 * (S->yq[jm] + S->yq[cc]), bxm = 0.5 * (S->xq[jm] + S->xq[cc]);
          const double* FA = S->F[UA[l]];
          const double* FB = S->F[UB[l]];
          double t = prodg(ayp, FA[ip], GA[l], g) - prodg(axp, FB[ip], GB[l], g) -
                     prodg(aym, FA[im], GA[l], g) + prodg(axm, FB[im], GB[l], g) -
                     prodg(byp, FA[jp], GA[l], g) + prodg(bxp, FB[jp], GB[l], g) +
                     prodg(bym, FA[jm], GA[l], g) - prodg(bxm, FB[jm], GB[l], g);
          if (!S->prho) {
            S->DF[l][cc] = S->DF[l][cc] + t * coef[l][s] * S->Ji[cc];
          } else {
            /* rates_h_interior_mode2 (rates_heterogeneous.f90:375-432):
               ")*DI(s)*JacInv/rho", ")*DI(s)*JacInv*M", ")*DI(s)*JacInv*G" */
            double D = s <= PI_I ? DIc(s) : 0.0;
            double u = t * D * S->Ji[cc];
            if (l < 2) u = u / rho;
            else if (l == 3) u = u * G;
            else u = u * M;
            S->DF[l][cc] = S->DF[l][cc] + u;
          }
        }
    }
}

/* dissipation terms of rates_dissipation_boundary_mode2 (rates.f90:1485-
   2021), fields 1..5 (the sixth is overwritten by rates_szz):
   one-sided   DF += (J(k) F(k)) A(s) Ji Cdiss
   diagonal    DF += F AI(0) Cdiss
   central     DF += (J(k+) F(k+) + J(k-) F(k-)) AI(s) Ji Cdiss */
static inline void diss_one(const sten_t* S, size_t k, size_t cc, double A) {
  for (int l = 0; l < 5; l++)
    S->DF[l][cc] = S->DF[l][cc] + S->J[k] * S->F[l][k] * A * S->Ji[cc] * S->Cdiss;
}
static inline void diss_diag(const sten_t* S, size_t cc) {
  for (int l = 0; l < 5; l++) S->DF[l][cc] = S->DF[l][cc] + S->F[l][cc] * AIc(0) * S->Cdiss;
}
static inline void diss_central(const sten_t* S, size_t kp, size_t km, size_t cc, int s) {
  for (int l = 0; l < 5; l++)
    S->DF[l][cc] = S->DF[l][cc] + (S->J[kp] * S->F[l][kp] + S->J[km] * S->F[l][km]) * AIc(s) *
                                      S->Ji[cc] * S->Cdiss;
}
/* the same terms for the three antiplane fields (rates_dissipation_boundary_
   mode3, rates.f90:1002-1394) */
static inline void diss_one3(const sten_t* S, size_t k, size_t cc, double A) {
  for (int l = 0; l < 3; l++)
    S->DF[l][cc] = S->DF[l][cc] + S->J[k] * S->F[l][k] * A * S->Ji[cc] * S->Cdiss;
}
static inline void diss_diag3(const sten_t* S, size_t cc) {
  for (int l = 0; l < 3; l++) S->DF[l][cc] = S->DF[l][cc] + S->F[l][cc] * AIc(0) * S->Cdiss;
}
static inline void diss_central3(const sten_t* S, size_t kp, size_t km, size_t cc, int s) {
  for (int l = 0; l < 3; l++)
    S->DF[l][cc] = S->DF[l][cc] + (S->J[kp] * S->F[l][kp] + S->J[km] * S->F[l][km]) * AIc(s) *
                                      S->Ji[cc] * S->Cdiss;
}

/* rates_boundary_mode2, rates.f90:453-917.  qside: 0 none (central q),
   1 left one-sided, 2 right one-sided; rside likewise for r.  (mx,px,my,py)
   is the loop range; (bx0,bx1,by0,by1) the block limits mgx,pgx,mgy,pgy.
   With Cdiss > 0 the dissipation terms follow each one-sided or central
   term, and strips add the diagonal between their two loops; corners have
   none (rates.f90:1485-2021). */
static void rates_boundary(const sten_t* S, int mx, int px, int my, int py, int qside, int rside,
                           int bx0, int bx1, int by0, int by1) {
  int nx = S->nx;
  for (int j = my; j <= py; j++)
    for (int i = mx; i <= px; i++) {
      size_t cc = (size_t)j * nx + i;
      double Ji = S->Ji[cc];
      /* rates_h_boundary_mode2 (rates_heterogeneous.f90:435-898) has the
         same expressions with the point's material */
      double rho, G, M, g;
      mat_at(S, cc, &rho, &G, &M, &g);
      const int diss = S->Cdiss > 0.0;
      if (qside && rside) {
        for (int s = 0; s < NBND; s++) {
          size_t kq, kr;
          double Dq, Dr, Aq, Ar;
          if (qside == 1) {
            kq = (size_t)j * nx + (bx0 + s);
            Dq = DLc(s, i - bx0);
            Aq = ALc(s, i - bx0);
          } else {
            kq = (size_t)j * nx + (bx1 - s);
            Dq = DRc(s, bx1 - i);
            Aq = ALc(s, bx1 - i);
          }
          if (rside == 1) {
            kr = (size_t)(by0 + s) * nx + i;
            Dr = DLc(s, j - by0);
            Ar = ALc(s, j - by0);
          } else {
            kr = (size_t)(by1 - s) * nx + i;
            Dr = DRc(s, by1 - j);
            Ar = ALc(s, by1 - j);
          }
          for (int l = 0; l < 5; l++)
            S->DF[l][cc] = S->DF[l][cc] + bscale(l, q1(S, l, kq, cc, g), Dq, rho, G, M) * Ji;
          if (diss) diss_one(S, kq, cc, Aq);
          for (int l = 0; l < 5; l++)
            S->DF[l][cc] = S->DF[l][cc] + bscale(l, r1(S, l, kr, cc, g), Dr, rho, G, M) * Ji;
          if (diss) diss_one(S, kr, cc, Ar);
        }
      } else if (qside) {
        for (int s = 0; s < NBND; s++) {
          size_t kq;
          double Dq, Aq;
          if (qside == 1) {
            kq = (size_t)j * nx + (bx0 + s);
            Dq = DLc(s, i - bx0);
            Aq = ALc(s, i - bx0);
          } else {
            kq = (size_t)j * nx + (bx1 - s);
            Dq = DRc(s, bx1 - i);
            Aq = ALc(s, bx1 - i);
          }
          for (int l = 0; l < 5; l++)
            S->DF[l][cc] = S->DF[l][cc] + bscale(l, q1(S, l, kq, cc, g), Dq, rho, G, M) * Ji;
          if (diss) diss_one(S, kq, cc, Aq);
        }
        if (diss) diss_diag(S, cc);
        for (int s = 1; s <= PI_I; s++) {
          size_t jp = cc + (size_t)s * nx, jm = cc - (size_t)s * nx;
          for (int l = 0; l < 5; l++)
            S->DF[l][cc] =
                S->DF[l][cc] + bscale(l, rc(S, l, jp, jm, cc, g), DIc(s), rho, G, M) * Ji;
          if (diss) diss_central(S, jp, jm, cc, s);
        }
      } else {
        for (int s = 0; s < NBND; s++) {
          size_t kr;
          double Dr, Ar;
          if (rside == 1) {
            kr = (size_t)(by0 + s) * nx + i;
            Dr = DLc(s, j - by0);
            Ar = ALc(s, j - by0);
          } else {
            kr = (size_t)(by1 - s) * nx + i;
            Dr = DRc(s, by1 - j);
            Ar = ALc(s, by1 - j);
          }
          for (int l = 0; l < 5; l++)
            S->DF[l][cc] = S->DF[l][cc] + bscale(l, r1(S, l, kr, cc, g), Dr, rho, G, M) * Ji;
          if (diss) diss_one(S, kr, cc, Ar);
        }
        if (diss) diss_diag(S, cc);
        for (int s = 1; s <= PI_I; s++) {
          size_t ip = cc + s, im = cc - s;
          for (int l = 0; l < 5; l++)
            S->DF[l][cc] =
                S->DF[l][cc] + bscale(l, qc(S, l, ip, im, cc, g), DIc(s), rho, G, M) * Ji;
          if (diss) diss_central(S, ip, im, cc, s);
        }
      }
    }
}

/* antiplane shear (rates_interior_mode3, rates_boundary_mode3,
   rates.f90:8-381): DF1 from sxz (with y metrics) and syz (x metrics),
   scaled by DI/rho; DF2 from vz with y metrics, DF3 from vz with x metrics,
   scaled by DI*G.  Terms absent from a field are absent from its sum. */
static inline double q1m3(const sten_t* S, int l, size_t kk, size_t cc) {
  double ay = 0.5 * (S->yr[kk] + S->yr[cc]);
  double ax = 0.5 * (S->xr[kk] + S->xr[cc]);
  if (l == 0) return ay * S->F[1][kk] - ax * S->F[2][kk];
  if (l == 1) return ay * S->F[0][kk];
  return -(ax * S->F[0][kk]);
}
static inline double r1m3(const sten_t* S, int l, size_t kk, size_t cc) {
  double by = 0.5 * (S->yq[kk] + S->yq[cc]);
  double bx = 0.5 * (S->xq[kk] + S->xq[cc]);
  if (l == 0) return -(by * S->F[1][kk]) + bx * S->F[2][kk];
  if (l == 1) return -(by * S->F[0][kk]);
  return bx * S->F[0][kk];
}
static inline double qcm3(const sten_t* S, int l, size_t kp, size_t km, size_t cc) {
  double ayp = 0.5 * (S->yr[kp] + S->yr[cc]), axp = 0.5 * (S->xr[kp] + S->xr[cc]);
  double aym = 0.5 * (S->yr[km] + S->yr[cc]), axm = 0.5 * (S->xr[km] + S->xr[cc]);
  if (l == 0)
    return ayp * S->F[1][kp] - axp * S->F[2][kp] - aym * S->F[1][km] + axm * S->F[2][km];
  if (l == 1) return ayp * S->F[0][kp] - aym * S->F[0][km];
  return -(axp * S->F[0][kp]) + axm * S->F[0][km];
}
static inline double rcm3(const sten_t* S, int l, size_t kp, size_t km, size_t cc) {
  double byp = 0.5 * (S->yq[kp] + S->yq[cc]), bxp = 0.5 * (S->xq[kp] + S->xq[cc]);
  double bym = 0.5 * (S->yq[km] + S->yq[cc]), bxm = 0.5 * (S->xq[km] + S->xq[cc]);
  if (l == 0)
    return -(byp * S->F[1][kp]) + bxp * S->F[2][kp] + bym * S->F[1][km] - bxm * S->F[2][km];
  if (l == 1) return -(byp * S->F[0][kp]) + bym * S->F[0][km];
  return bxp * S->F[0][kp] - bxm * S->F[0][km];
}
/* ")*D/rho*JacInv", ")*D*G*JacInv" */
static inline double bscale3(int l, double t, double D, double rho, double G) {
  return l == 0 ? t * D / rho : t * D * G;
}

/* rates_interior_mode3 (rates.f90:8-56): a(s) = DI(s)/rho, b(s) = DI(s)*G;
   the s = 3 coefficient is zero.  With a heterogeneous medium,
   rates_h_interior_mode3 (rates_heterogeneous.f90:8-48): the point's rho
   and G, as ")*DI(s)*JacInv(i,j)/rho(i,j)" and ")*DI(s)*G(i,j)*JacInv(i,j)",
   evaluated left to right. */
static void rates_dissipation_interior3(const sten_t* S, int mx, int px, int my, int py);

static void rates_interior3(const sten_t* S, int mx, int px, int my, int py) {
  if (S->Cdiss > 0.0) {
    rates_dissipation_interior3(S, mx, px, my, py);
    return;
  }
  double coef[3][4];
  for (int s = 1; s <= PI_I; s++) {
    coef[0][s] = DIc(s) / S->rho;
    coef[1][s] = DIc(s) * S->G;
    coef[2][s] = DIc(s) * S->G;
  }
  int nx = S->nx;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = my; j <= py; j++)
    for (int i = mx; i <= px; i++) {
      size_t cc = (size_t)j * nx + i;
      for (int s = 1; s <= PI_I; s++) {
        size_t ip = cc + s, im = cc - s, jp = cc + (size_t)s * nx, jm = cc - (size_t)s * nx;
        double ayp = 0.5 * (S->yr[ip] + S->yr[cc]), axp = 0.5 * (S->xr[ip] + S->xr[cc]);
        double aym = 0.5 * (S->yr[im] + S->yr[cc]), axm = 0.5 * (S->xr[im] + S->xr[cc]);
        double byp = 0.5 * (S->yq[jp] + S->yq[cc]), bxp = 0.5 * (S->xq[jp] + S->xq[cc]);
        double bym = 0.5 * (S->yq[jm] + S->yq[cc]), bxm = 0.5 * (S->xq[jm] + S->xq[cc]);
        const double *F1 = S->F[0], *F2 = S->F[1], *F3 = S->F[2];
        double t1 = ayp * F2[ip] - axp * F3[ip] - aym * F2[im] + axm * F3[im] - byp * F2[jp] +
                    bxp * F3[jp] + bym * F2[jm] - bxm * F3[jm];
        double t2 = ayp * F1[ip] - aym * F1[im] - byp * F1[jp] + bym * F1[jm];
        double t3 = -(axp * F1[ip]) + axm * F1[im] + bxp * F1[jp] - bxm * F1[jm];
        if (S->prho) {
          const double D = DIc(s), r = S->prho[cc], G = S->pG[cc];
          S->DF[0][cc] = S->DF[0][cc] + t1 * D * S->Ji[cc] / r;
          S->DF[1][cc] = S->DF[1][cc] + t2 * D * G * S->Ji[cc];
          S->DF[2][cc] = S->DF[2][cc] + t3 * D * G * S->Ji[cc];
        } else {
          S->DF[0][cc] = S->DF[0][cc] + t1 * coef[0][s] * S->Ji[cc];
          S->DF[1][cc] = S->DF[1][cc] + t2 * coef[1][s] * S->Ji[cc];
          S->DF[2][cc] = S->DF[2][cc] + t3 * coef[2][s] * S->Ji[cc];
        }
      }
    }
}

/* rates_boundary_mode3 (rates.f90:59-381), the loop structure of the mode-2
   routine; with a heterogeneous medium rates_h_boundary_mode3
   (rates_heterogeneous.f90:51-372) takes the point's rho and G in the same
   expressions */
static void rates_boundary3(const sten_t* S, int mx, int px, int my, int py, int qside,
                            int rside, int bx0, int bx1, int by0, int by1) {
  int nx = S->nx;
  const int diss = S->Cdiss > 0.0;
  for (int j = my; j <= py; j++)
    for (int i = mx; i <= px; i++) {
      size_t cc = (size_t)j * nx + i;
      double Ji = S->Ji[cc];
      const double rho = S->prho ? S->prho[cc] : S->rho, G = S->pG ? S->pG[cc] : S->G;
      if (qside) {
        for (int s = 0; s < NBND; s++) {
          size_t kq = qside == 1 ? (size_t)j * nx + (bx0 + s) : (size_t)j * nx + (bx1 - s);
          double Dq = qside == 1 ? DLc(s, i - bx0) : DRc(s, bx1 - i);
          double Aq = qside == 1 ? ALc(s, i - bx0) : ALc(s, bx1 - i);
          for (int l = 0; l < 3; l++)
            S->DF[l][cc] = S->DF[l][cc] + bscale3(l, q1m3(S, l, kq, cc), Dq, rho, G) * Ji;
          if (diss) diss_one3(S, kq, cc, Aq);
          if (rside) {
            size_t kr = rside == 1 ? (size_t)(by0 + s) * nx + i : (size_t)(by1 - s) * nx + i;
            double Dr = rside == 1 ? DLc(s, j - by0) : DRc(s, by1 - j);
            double Ar = rside == 1 ? ALc(s, j - by0) : ALc(s, by1 - j);
            for (int l = 0; l < 3; l++)
              S->DF[l][cc] = S->DF[l][cc] + bscale3(l, r1m3(S, l, kr, cc), Dr, rho, G) * Ji;
            if (diss) diss_one3(S, kr, cc, Ar);
          }
        }
        if (!rside) {
          if (diss) diss_diag3(S, cc);
          for (int s = 1; s <= PI_I; s++) {
            size_t jp = cc + (size_t)s * nx, jm = cc - (size_t)s * nx;
            for (int l = 0; l < 3; l++)
              S->DF[l][cc] =
                  S->DF[l][cc] + bscale3(l, rcm3(S, l, jp, jm, cc), DIc(s), rho, G) * Ji;
            if (diss) diss_central3(S, jp, jm, cc, s);
          }
        }
      } else {
        for (int s = 0; s < NBND; s++) {
          size_t kr = rside == 1 ? (size_t)(by0 + s) * nx + i : (size_t)(by1 - s) * nx + i;
          double Dr = rside == 1 ? DLc(s, j - by0) : DRc(s, by1 - j);
          double Ar = rside == 1 ? ALc(s, j - by0) : ALc(s, by1 - j);
          for (int l = 0; l < 3; l++)
            S->DF[l][cc] = S->DF[l][cc] + bscale3(l, r1m3(S, l, kr, cc), Dr, rho, G) * Ji;
          if (diss) diss_one3(S, kr, cc, Ar);
        }
        if (diss) diss_diag3(S, cc);
        for (int s = 1; s <= PI_I; s++) {
          size_t ip = cc + s, im = cc - s;
          for (int l = 0; l < 3; l++)
            S->DF[l][cc] =
                S->DF[l][cc] + bscale3(l, qcm3(S, l, ip, im, cc), DIc(s), rho, G) * Ji;
          if (diss) diss_central3(S, ip, im, cc, s);
        }
      }
    }
}

/* rates_dissipation_interior_mode3 (rates.f90:939-999) and its heterogeneous
   form (rates_heterogeneous.f90:920-978): c(0) F first, then per s and
   field (X)*a(s)*JacInv + c(s)*JacInv*(Y), with a(s) = DI(s)/rho for vz and
   DI(s)*G for the stresses, or ")*DI(s)/rho(i,j)*JacInv" and
   ")*DI(s)*G(i,j)*JacInv" with a medium; Y the J-weighted neighbours of
   the field itself.  The s = 3 terms are zero for SBP4 (the only operator
   with dissipation). */
static void rates_dissipation_interior3(const sten_t* S, int mx, int px, int my, int py) {
  double c[4], a[3][4];
  c[0] = 2.0 * AIc(0) * S->Cdiss;
  for (int s = 1; s <= 3; s++) {
    c[s] = s <= PI_I ? AIc(s) * S->Cdiss : 0.0;
    const double D = s <= PI_I ? DIc(s) : 0.0;
    a[0][s] = D / S->rho;
    a[1][s] = D * S->G;
    a[2][s] = D * S->G;
  }
  int nx = S->nx;
  const double* J = S->J;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = my; j <= py; j++)
    for (int i = mx; i <= px; i++) {
      size_t cc = (size_t)j * nx + i;
      const double *F1 = S->F[0], *F2 = S->F[1], *F3 = S->F[2];
      for (int l = 0; l < 3; l++) S->DF[l][cc] = S->DF[l][cc] + c[0] * S->F[l][cc];
      for (int s = 1; s <= PI_I; s++) {
        size_t ip = cc + s, im = cc - s, jp = cc + (size_t)s * nx, jm = cc - (size_t)s * nx;
        double ayp = 0.5 * (S->yr[ip] + S->yr[cc]), axp = 0.5 * (S->xr[ip] + S->xr[cc]);
        double aym = 0.5 * (S->yr[im] + S->yr[cc]), axm = 0.5 * (S->xr[im] + S->xr[cc]);
        double byp = 0.5 * (S->yq[jp] + S->yq[cc]), bxp = 0.5 * (S->xq[jp] + S->xq[cc]);
        double bym = 0.5 * (S->yq[jm] + S->yq[cc]), bxm = 0.5 * (S->xq[jm] + S->xq[cc]);
        double x[3];
        x[0] = ayp * F2[ip] - axp * F3[ip] - aym * F2[im] + axm * F3[im] - byp * F2[jp] +
               bxp * F3[jp] + bym * F2[jm] - bxm * F3[jm];
        x[1] = ayp * F1[ip] - aym * F1[im] - byp * F1[jp] + bym * F1[jm];
        x[2] = -(axp * F1[ip]) + axm * F1[im] + bxp * F1[jp] - bxm * F1[jm];
        for (int l = 0; l < 3; l++) {
          const double* Fl = S->F[l];
          double y = J[ip] * Fl[ip] + J[im] * Fl[im] + J[jp] * Fl[jp] + J[jm] * Fl[jm];
          double u;
          if (S->prho) {
            const double D = DIc(s);
            u = l == 0 ? x[l] * D / S->prho[cc] * S->Ji[cc] : x[l] * D * S->pG[cc] * S->Ji[cc];
          } else {
            u = x[l] * a[l][s] * S->Ji[cc];
          }
          S->DF[l][cc] = S->DF[l][cc] + u + c[s] * S->Ji[cc] * y;
        }
      }
    }
}

/* set_rates, mode-2 homogeneous branch (time_step.f90:457-490) */
/* ---------------------------------------------------------------------- */
/* perfectly matched layer: rates_interior_mode2_pml and                    */
/* rates_boundary_mode2_pml (rates.f90:2024-2619)                           */
/* ---------------------------------------------------------------------- */

/* damping at (x, y), rates.f90:2048-2049, 2063-2066: sigma*|max(0, mx-x,
   x-px)|^n (likewise in y) for n > 0, else sigma */
static void pml_sigma(const fdm_block* b, double x, double y, double* sx, double* sy) {
  const double n = b->pml[0], sigma = b->pml[1];
  if (n > 0.0) {
    *sx = sigma * pow(fabs(fmax(0.0, fmax(b->pml[3] - x, x - b->pml[4]))), n);
    *sy = sigma * pow(fabs(fmax(0.0, fmax(b->pml[5] - y, y - b->pml[6]))), n);
  } else {
    *sx = sigma;
    *sy = sigma;
  }
}

typedef struct {
  double hxi, hyi, sx, sy, rho, G, M, gamma;
  int pmlx, pmly;
} pml_pt;

/* x-derivative term of the boundary forms, v the fields at the one-sided
   neighbour or the central difference, D its coefficient (rates.f90:
   2194-2206 and the like) */
static inline void pml_xterm(const pml_pt* p, const double* v, double D, double* d, double* dwx) {
  if (p->pmlx)
    for (int l = 0; l < 6; l++) dwx[l] = dwx[l] + p->sx * p->hxi * v[l] * D;
  d[0] = d[0] + p->hxi * v[2] * D / p->rho;
  d[1] = d[1] + p->hxi * v[3] * D / p->rho;
  d[2] = d[2] + p->hxi * v[0] * D * p->M;
  d[3] = d[3] + p->hxi * v[1] * D * p->G;
  d[4] = d[4] + p->hxi * v[0] * D * p->M * p->gamma;
}

/* y-derivative term of the boundary forms (rates.f90:2208-2218) */
static inline void pml_yterm(const pml_pt* p, const double* v, double D, double* d, double* dwy) {
  if (p->pmly)
    for (int l = 0; l < 6; l++) dwy[l] = dwy[l] + p->sy * p->hyi * v[l] * D;
  d[0] = d[0] + p->hyi * v[3] * D / p->rho;
  d[1] = d[1] + p->hyi * v[4] * D / p->rho;
  d[2] = d[2] + p->hyi * v[1] * D * p->M * p->gamma;
  d[3] = d[3] + p->hyi * v[0] * D * p->G;
  d[4] = d[4] + p->hyi * v[1] * D * p->M;
}

/* One region of a PML block: the interior (qside = rside = 0), a strip or
   a corner (qside/rside 1: minus side one-sided, 2: plus side).  hxi, hyi
   are Ji*yr and Ji*xq at the region's first point (rates.f90:2060-2061,
   2158-2159): the layer's stencils are Cartesian. */
static void rates_pml(orc_ctx* c, const fdm_block* b, int mx, int px, int my, int py, int qside,
                      int rside) {
  if (mx > px || my > py) return;
  const size_t np = c->np;
  const int mgx = b->mx, pgx = b->px, mgy = b->my, pgy = b->py;
  const size_t q0 = IDX(mx, my);
  pml_pt p;
  p.hxi = c->Ji[q0] * c->yr[q0];
  p.hyi = c->Ji[q0] * c->xq[q0];
  p.rho = b->rho;
  p.G = b->G;
  p.M = b->M;
  p.gamma = b->gamma;
  p.pmlx = b->pmlx;
  p.pmly = b->pmly;
  const double kappa = b->pml[2];
  /* interior coefficients a, b, c (rates.f90:2053-2058) */
  double ca[4], cb[4], cc[4];
  for (int s = 1; s <= PI_I; s++) {
    ca[s] = DIc(s) / p.rho;
    cb[s] = DIc(s) * p.G;
    cc[s] = DIc(s) * p.M;
  }
#define PF(l, k) c->F[(size_t)(l) * np + (k)]
  for (int j = my; j <= py; j++)
    for (int i = mx; i <= px; i++) {
      const size_t q = IDX(i, j);
      pml_sigma(b, c->x[q], c->y[q], &p.sx, &p.sy);
      double d[5], dwx[6], dwy[6], v[6];
      for (int l = 0; l < 5; l++) d[l] = c->DF[(size_t)l * np + q];
      for (int l = 0; l < 6; l++) {
        dwx[l] = p.pmlx ? c->DWx[(size_t)l * np + q] : 0.0;
        dwy[l] = p.pmly ? c->DWy[(size_t)l * np + q] : 0.0;
      }
      if (p.pmlx) {
        const double* W = c->Wx;
        d[0] = d[0] - W[2 * np + q] / p.rho;
        d[1] = d[1] - W[3 * np + q] / p.rho;
        d[2] = d[2] - W[0 * np + q] * p.M;
        d[3] = d[3] - W[1 * np + q] * p.G;
        d[4] = d[4] - W[0 * np + q] * p.gamma * p.M;
        for (int l = 0; l < 6; l++) dwx[l] = dwx[l] - (p.sx + kappa) * W[(size_t)l * np + q];
      }
      if (p.pmly) {
        const double* W = c->Wy;
        d[0] = d[0] - W[3 * np + q] / p.rho;
        d[1] = d[1] - W[4 * np + q] / p.rho;
        d[2] = d[2] - W[1 * np + q] * p.gamma * p.M;
        d[3] = d[3] - W[0 * np + q] * p.G;
        d[4] = d[4] - W[1 * np + q] * p.M;
        for (int l = 0; l < 6; l++) dwy[l] = dwy[l] - (p.sy + kappa) * W[(size_t)l * np + q];
      }
      if (qside == 0 && rside == 0) {
        /* interior: x and y differences in one sum (rates.f90:2088-2124);
           terms with a zero coefficient (s > pI) are left out */
        for (int s = 1; s <= PI_I; s++) {
          double dx[6], dy[6];
          for (int l = 0; l < 6; l++) {
            dx[l] = PF(l, IDX(i + s, j)) - PF(l, IDX(i - s, j));
            dy[l] = PF(l, IDX(i, j + s)) - PF(l, IDX(i, j - s));
          }
          if (p.pmlx)
            for (int l = 0; l < 6; l++) dwx[l] = dwx[l] + p.sx * p.hxi * dx[l] * DIc(s);
          if (p.pmly)
            for (int l = 0; l < 6; l++) dwy[l] = dwy[l] + p.sy * p.hyi * dy[l] * DIc(s);
          d[0] = d[0] + (p.hxi * dx[2] + p.hyi * dy[3]) * ca[s];
          d[1] = d[1] + (p.hxi * dx[3] + p.hyi * dy[4]) * ca[s];
          d[2] = d[2] + (p.hxi * dx[0] + p.hyi * dy[1] * p.gamma) * cc[s];
          d[3] = d[3] + (p.hxi * dx[1] + p.hyi * dy[0]) * cb[s];
          d[4] = d[4] + (p.hxi * dx[0] * p.gamma + p.hyi * dy[1]) * cc[s];
        }
      } else if (rside == 0) {
        /* L or R strip: one-sided x, then central y (rates.f90:2165-2218) */
        for (int s = 0; s < NBND; s++) {
          const size_t k = qside == 1 ? IDX(mgx + s, j) : IDX(pgx - s, j);
          const double D = qside == 1 ? DLc(s, i - mgx) : DRc(s, pgx - i);
          for (int l = 0; l < 6; l++) v[l] = PF(l, k);
          pml_xterm(&p, v, D, d, dwx);
        }
        for (int s = 1; s <= PI_I; s++) {
          for (int l = 0; l < 6; l++) v[l] = PF(l, IDX(i, j + s)) - PF(l, IDX(i, j - s));
          pml_yterm(&p, v, DIc(s), d, dwy);
        }
      } else if (qside == 0) {
        /* B or T strip: central x, then one-sided y (rates.f90:2287-2335) */
        for (int s = 1; s <= PI_I; s++) {
          for (int l = 0; l < 6; l++) v[l] = PF(l, IDX(i + s, j)) - PF(l, IDX(i - s, j));
          pml_xterm(&p, v, DIc(s), d, dwx);
        }
        for (int s = 0; s < NBND; s++) {
          const size_t k = rside == 1 ? IDX(i, mgy + s) : IDX(i, pgy - s);
          const double D = rside == 1 ? DLc(s, j - mgy) : DRc(s, pgy - j);
          for (int l = 0; l < 6; l++) v[l] = PF(l, k);
          pml_yterm(&p, v, D, d, dwy);
        }
      } else {
        /* corner: one-sided x and y, interleaved per s (rates.f90:2399-2441) */
        for (int s = 0; s < NBND; s++) {
          const size_t kx = qside == 1 ? IDX(mgx + s, j) : IDX(pgx - s, j);
          const double Dx = qside == 1 ? DLc(s, i - mgx) : DRc(s, pgx - i);
          for (int l = 0; l < 6; l++) v[l] = PF(l, kx);
          pml_xterm(&p, v, Dx, d, dwx);
          const size_t ky = rside == 1 ? IDX(i, mgy + s) : IDX(i, pgy - s);
          const double Dy = rside == 1 ? DLc(s, j - mgy) : DRc(s, pgy - j);
          for (int l = 0; l < 6; l++) v[l] = PF(l, ky);
          pml_yterm(&p, v, Dy, d, dwy);
        }
      }
      for (int l = 0; l < 5; l++) c->DF[(size_t)l * np + q] = d[l];
      for (int l = 0; l < 6; l++) {
        if (p.pmlx) c->DWx[(size_t)l * np + q] = dwx[l];
        if (p.pmly) c->DWy[(size_t)l * np + q] = dwy[l];
      }
    }
#undef PF
}

static void set_rates(orc_ctx* c, int ib) {
  fdm_block* b = &c->B[ib].b;
  sten_t S;
  for (int l = 0; l < 5; l++) {
    S.F[l] = FL(l);
    S.DF[l] = DFL(l);
  }
  S.xq = c->xq;
  S.xr = c->xr;
  S.yq = c->yq;
  S.yr = c->yr;
  S.Ji = c->Ji;
  S.rho = b->rho;
  S.G = b->G;
  S.M = b->M;
  S.gamma = b->gamma;
  S.prho = c->mrho;
  S.pG = c->mG;
  S.pM = c->mM;
  S.pgam = c->mgam;
  S.J = c->J;
  S.Cdiss = b->Cdiss;
  S.nx = c->nx;
  int mgx = b->mx, pgx = b->px, mgy = b->my, pgy = b->py;
  int mix = mgx + NBST, pix = pgx - NBST, miy = mgy + NBST, piy = pgy - NBST;
  if (c->mode == 3) {
    /* set_rates, mode 3 (time_step.f90:522-547): no szz relation */
    rates_interior3(&S, mix, pix, miy, piy);
    rates_boundary3(&S, mgx, mix - 1, miy, piy, 1, 0, mgx, pgx, mgy, pgy);
    rates_boundary3(&S, pix + 1, pgx, miy, piy, 2, 0, mgx, pgx, mgy, pgy);
    rates_boundary3(&S, mix, pix, mgy, miy - 1, 0, 1, mgx, pgx, mgy, pgy);
    rates_boundary3(&S, mix, pix, piy + 1, pgy, 0, 2, mgx, pgx, mgy, pgy);
    rates_boundary3(&S, mgx, mix - 1, mgy, miy - 1, 1, 1, mgx, pgx, mgy, pgy);
    rates_boundary3(&S, pix + 1, pgx, mgy, miy - 1, 2, 1, mgx, pgx, mgy, pgy);
    rates_boundary3(&S, mgx, mix - 1, piy + 1, pgy, 1, 2, mgx, pgx, mgy, pgy);
    rates_boundary3(&S, pix + 1, pgx, piy + 1, pgy, 2, 2, mgx, pgx, mgy, pgy);
    return;
  }
  if ((b->pmlx || b->pmly) && !(b->Cdiss > 0.0)) {
    /* set_rates with a PML (time_step.f90:354-420); a block with
       dissipation takes set_rates_dissipation instead (time_step.f90:
       241-246) */
    rates_pml(c, b, mix, pix, miy, piy, 0, 0);
    rates_pml(c, b, mgx, mix - 1, miy, piy, 1, 0);
    rates_pml(c, b, pix + 1, pgx, miy, piy, 2, 0);
    rates_pml(c, b, mix, pix, mgy, miy - 1, 0, 1);
    rates_pml(c, b, mix, pix, piy + 1, pgy, 0, 2);
    rates_pml(c, b, mgx, mix - 1, mgy, miy - 1, 1, 1);
    rates_pml(c, b, pix + 1, pgx, mgy, miy - 1, 2, 1);
    rates_pml(c, b, mgx, mix - 1, piy + 1, pgy, 1, 2);
    rates_pml(c, b, pix + 1, pgx, piy + 1, pgy, 2, 2);
  } else {
  rates_interior(&S, mix, pix, miy, piy);
  rates_boundary(&S, mgx, mix - 1, miy, piy, 1, 0, mgx, pgx, mgy, pgy);         /* L  */
  rates_boundary(&S, pix + 1, pgx, miy, piy, 2, 0, mgx, pgx, mgy, pgy);         /* R  */
  rates_boundary(&S, mix, pix, mgy, miy - 1, 0, 1, mgx, pgx, mgy, pgy);         /* B  */
  rates_boundary(&S, mix, pix, piy + 1, pgy, 0, 2, mgx, pgx, mgy, pgy);         /* T  */
  rates_boundary(&S, mgx, mix - 1, mgy, miy - 1, 1, 1, mgx, pgx, mgy, pgy);     /* LB */
  rates_boundary(&S, pix + 1, pgx, mgy, miy - 1, 2, 1, mgx, pgx, mgy, pgy);     /* RB */
  rates_boundary(&S, mgx, mix - 1, piy + 1, pgy, 1, 2, mgx, pgx, mgy, pgy);     /* LT */
  rates_boundary(&S, pix + 1, pgx, piy + 1, pgy, 2, 2, mgx, pgx, mgy, pgy);     /* RT */
  }
  /* rates_szz, rates.f90:920-936 (rates_h_szz, rates_heterogeneous.f90:901-917) */
  double* DF3 = DFL(2);
  double* DF5 = DFL(4);
  double* DF6 = DFL(5);
  for (int j = mgy; j <= pgy; j++)
    for (int i = mgx; i <= pgx; i++) {
      size_t q = IDX(i, j);
      DF6[q] = (c->mnu ? c->mnu[q] : b->nu) * (DF3[q] + DF5[q]);
    }
}

/* source time function (source.f90:212-223) */
static double stf_amplitude(const fdm_source* s, double t) {
  const double pi = 3.14159265358979323846264338327950288419716939937510;
  double tc = s->t_center, td = s->t_duration;
  switch (s->stf) {
    case FDM_STF_RAMP:
      return fmax(fmin((t - tc) / td, 1.0), 0.0);
    case FDM_STF_SINUSOID:
      return sin(2.0 * pi * (t - tc) / td);
    case FDM_STF_GAUSSIAN:
      return exp(-0.5 * ((t - tc) * (t - tc)) / (td * td));
    case FDM_STF_RICKER:
      return (1.0 - 2.0 * (pi * pi) * ((t - tc) * (t - tc)) / (td * td)) *
             exp(-(pi * pi) * ((t - tc) * (t - tc)) / (td * td));
    default:
      return 0.0;
  }
}

/* set_source (source.f90:147-324): singular source and MMS source */
static void set_source(orc_ctx* c, int ib, double t) {
  fdm_block* b = &c->B[ib].b;
  const fdm_source* s = &c->src;
  for (int pl = 0; pl < s->nplace; pl++) {
    const fdm_source_place* P = &s->place[pl];
    if (P->block != ib) continue;
    double A = stf_amplitude(s, t);
    for (int j = b->my; j <= b->py; j++) {
      int dj = j - P->j0 + 2;
      if (dj < 0 || dj > 4) continue;
      for (int i = b->mx; i <= b->px; i++) {
        int di = i - P->i0 + 2;
        if (di < 0 || di > 4) continue;
        double dx = P->dx[di], dy = P->dy[dj];
        size_t q = IDX(i, j);
        DFL(0)[q] = DFL(0)[q] + A * s->Fx * dx * dy / b->rho;
        DFL(1)[q] = DFL(1)[q] + A * s->Fy * dx * dy / b->rho;
        DFL(2)[q] = DFL(2)[q] - A * s->Mxx * dx * dy;
        DFL(3)[q] = DFL(3)[q] - A * s->Mxy * dx * dy;
        DFL(4)[q] = DFL(4)[q] - A * s->Myy * dx * dy;
      }
    }
  }
  if (b->mms) {
    for (int j = b->my; j <= b->py; j++)
      for (int i = b->mx; i <= b->px; i++) {
        size_t q = IDX(i, j);
        double x = c->x[q], y = c->y[q];
        for (int l = 0; l < 6; l++)
          DFL(l)[q] = DFL(l)[q] + orc_mms_sin(x, y, t, b->id, 6 + l);
      }
  }
}

/* set_rates_all (time_step.f90:215-270) */
static void set_rates_all(orc_ctx* c, double t) {
  for (int ib = 0; ib < c->nb; ib++) set_rates_SAT(c, ib);
  for (int ib = 0; ib < c->nb; ib++) set_rates(c, ib);
  for (int ib = 0; ib < c->nb; ib++) set_source(c, ib, t);
}

/* ---------------------------------------------------------------------- */
/* plasticity, plastic.f90:8-261                                           */
/* ---------------------------------------------------------------------- */

void orc_plastic_point(double G, double K, double mu, double beta, double b, double h,
                       double eta, const double* s0, double dt, double* s, double* Ds,
                       double* lambda, double* gammap) {
  static const double delta[6] = {1.0, 0.0, 0.0, 1.0, 0.0, 1.0};
  double sa[6], sd[6], W[6];
  /* stress (plastic.f90:212-242) */
  sa[0] = s0[0] + s[0];
  sa[1] = s0[1] + s[1];
  sa[2] = s0[2];
  sa[3] = s0[3] + s[2];
  sa[4] = s0[4];
  sa[5] = s0[5] + s[3];
  /* invariants (plastic.f90:195-209) */
  double sigma = (sa[0] + sa[3] + sa[5]) / 3.0;
  double J2 = ((sa[0] - sa[3]) * (sa[0] - sa[3]) + (sa[3] - sa[5]) * (sa[3] - sa[5]) +
               (sa[5] - sa[0]) * (sa[5] - sa[0])) /
                  6.0 +
              sa[1] * sa[1] + sa[2] * sa[2] + sa[4] * sa[4];
  double tau = sqrt(J2);
  /* yield (plastic.f90:245-261) */
  double bb = b + *gammap * h;
  double Y = tau - fmax(-mu * sigma + bb, 0.0);
  if (Y <= 0.0) {
    *lambda = 0.0;
    return;
  }
  for (int k = 0; k < 6; k++) sd[k] = sa[k] - sigma * delta[k];
  if (tau * mu * beta * K > (mu * sigma - b) * G)
    *lambda = (tau + mu * sigma - (b + h * *gammap)) / (eta + dt * (h + G + mu * beta * K));
  else
    *lambda = tau / (eta + dt * G);
  double lam = *lambda;
  *gammap = *gammap + dt * lam;
  tau = tau - dt * lam * G;
  sigma = sigma - dt * lam * beta * K;
  for (int k = 0; k < 6; k++) W[k] = K * beta * delta[k] + G * sd[k] / (tau + dt * lam * G);
  for (int k = 0; k < 6; k++) sd[k] = sd[k] * tau / (tau + dt * lam * G);
  for (int k = 0; k < 6; k++) sa[k] = sigma * delta[k] + sd[k];
  s[0] = sa[0] - s0[0];
  s[1] = sa[1] - s0[1];
  s[2] = sa[3] - s0[3];
  s[3] = sa[5] - s0[5];
  /* the reference adds dt*lambda to gammap a second time (plastic.f90:177) */
  *gammap = *gammap + dt * lam;
  Ds[0] = Ds[0] - lam * W[0];
  Ds[1] = Ds[1] - lam * W[1];
  Ds[2] = Ds[2] - lam * W[3];
  Ds[3] = Ds[3] - lam * W[5];
}

static void update_fields_plastic(orc_ctx* c, double dt) {
  for (int ib = 0; ib < c->nb; ib++) {
    fdm_block* b = &c->B[ib].b;
    if (!b->plastic) continue;
    for (int j = b->my; j <= b->py; j++)
      for (int i = b->mx; i <= b->px; i++) {
        size_t q = IDX(i, j);
        double s[4], Ds[4];
        for (int k = 0; k < 4; k++) {
          s[k] = FL(2 + k)[q];
          Ds[k] = DFL(2 + k)[q];
        }
        /* heterogeneous medium: G and K of the point (plastic.f90:37-43) */
        double Gel = c->mG ? c->mG[q] : b->G;
        double Kel = c->mG ? c->mM[q] - 1.333333333333333 * c->mG[q] : b->K;
        orc_plastic_point(Gel, Kel, b->mu, b->beta, b->b, b->h, b->eta, b->s0, dt, s, Ds,
                          &c->lambda[q], &c->gammap[q]);
        for (int k = 0; k < 4; k++) {
          FL(2 + k)[q] = s[k];
          DFL(2 + k)[q] = Ds[k];
        }
      }
  }
}

/* ---------------------------------------------------------------------- */
/* time stepping                                                            */
/* ---------------------------------------------------------------------- */

static void snapshot_stage1(orc_ctx* c) {
  for (int ii = 0; ii < c->ni; ii++) {
    oiface_t* I = &c->I[ii];
    for (int f = 0; f < FDM_IF_NFIELDS; f++)
      memcpy(I->st1[f], I->f[f], I->n * sizeof(double));
  }
}

/* time_step_LS (time_step.f90:71-167) */
static int time_step_LS(orc_ctx* c, double dt, int final_step) {
  g_op = c->op; /* this context's operator for every rate of the step */
  double t0 = c->t;
  int fail = 0;
  for (int stage = 1; stage <= c->nstage; stage++) {
    scale_rates_all(c, c->rkA[stage - 1]);
    prepare_edges(c);
    enforce_boundary_conditions(c, c->t);
    fail |= enforce_interface_conditions(c, c->t, c->rkB[stage - 1] * dt);
    set_rates_all(c, c->t);
    if (stage == 1) snapshot_stage1(c);
    if (final_step) return fail;
    c->t = t0 + c->rkC[stage] * dt;
    update_fields_all(c, c->rkB[stage - 1] * dt);
  }
  update_fields_plastic(c, dt);
  return fail;
}

int orc_init_edges(orc_ctx* c, double t) {
  c->t = t;
  prepare_edges(c);
  enforce_boundary_conditions(c, t);
  int fail = enforce_interface_conditions(c, t, 0.0);
  if (fail) {
    snprintf(c->err, sizeof c->err, "friction Newton solve failed");
    return FDM_ENEWTON;
  }
  return FDM_OK;
}

int orc_step(orc_ctx* c, double t, double dt, int64_t n_first, int32_t nsteps) {
  for (int32_t k = 0; k < nsteps; k++) {
    int64_t n = n_first + k;
    c->t = t + (double)(n - 1) * dt; /* main.f90:176 */
    if (time_step_LS(c, dt, 0)) {
      snprintf(c->err, sizeof c->err, "friction Newton solve failed at step %lld",
               (long long)n);
      return FDM_ENEWTON;
    }
  }
  return FDM_OK;
}

int orc_stage1(orc_ctx* c, double t, double dt, int64_t n) {
  c->t = t + (double)(n - 1) * dt;
  if (time_step_LS(c, dt, 1)) {
    snprintf(c->err, sizeof c->err, "friction Newton solve failed");
    return FDM_ENEWTON;
  }
  return FDM_OK;
}
