// fdm_physics.hpp -- pointwise physics shared by the device kernels:
// rotations, characteristic boundary hats, interface hats and the
// rate-and-state friction solve.  All in FP64; every expression keeps the
// reference's left-to-right evaluation order (the library is compiled with
// -ffp-contract=off so that nothing is fused).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/fdmap.h"

namespace fdm {

// SBP4 interior coefficients DI(1), DI(2) (fd_coeff.f90:101-102)
#define FDM_DI1 0.666666666666667
#define FDM_DI2 (-0.083333333333333)

// DL(s,r) (fd_coeff.f90:108-115); DR(-s,-r) = -DL(s,r) (fd_coeff.f90:242-246)
__device__ constexpr double kDL[4][6] = {
    {-1.411764705882353, 1.735294117647059, -0.235294117647059, -0.088235294117647,
     0.000000000000000, 0.000000000000000},
    {-0.500000000000000, 0.000000000000000, 0.500000000000000, 0.000000000000000,
     0.000000000000000, 0.000000000000000},
    {0.093023255813953, -0.686046511627907, 0.000000000000000, 0.686046511627907,
     -0.093023255813953, 0.000000000000000},
    {0.030612244897959, 0.000000000000000, -0.602040816326531, 0.000000000000000,
     0.653061224489796, -0.081632653061224}};

// The first-derivative operators 'SBP2', 'SBP4' and 'SBP6' (init_fd,
// fd_coeff.f90:41-248): points of a boundary stencil (nbnd), boundary
// stencils (nbst, the width of each block's closure band) and the interior
// half-width.  The tables are the reference's 15-digit literals.
template <int OP>
struct Sbp;
template <>
struct Sbp<2> { static constexpr int nbnd = 2, nbst = 1, half = 1; };
template <>
struct Sbp<4> { static constexpr int nbnd = 6, nbst = 4, half = 2; };
template <>
struct Sbp<6> { static constexpr int nbnd = 9, nbst = 6, half = 3; };

// 'SBP2' DL (fd_coeff.f90:88-94) and 'SBP6' DL (fd_coeff.f90:137-155),
// kDL6[r][s] = DL(s,r)
__device__ constexpr double kDL2[1][2] = {{-1.0, 1.0}};
__device__ constexpr double kDL6[6][9] = {
    {-1.582533518939116, 2.033378678700676, -0.141512858744873, -0.450398306578272,
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
     0.016437980868017}};

// DL(s,r) of the operator (s = 0..nbnd-1, r = 0..nbst-1); DR(-s,-r) = -DL(s,r)
template <int OP>
__device__ __forceinline__ double sbp_dl(int r, int s) {
  if constexpr (OP == 2) return kDL2[r][s];
  else if constexpr (OP == 6) return kDL6[r][s];
  else return kDL[r][s];
}

// interior DI(s), s = 1..half (fd_coeff.f90:86, 101-102, 132-134)
template <int OP>
__host__ __device__ __forceinline__ double sbp_di(int s) {
  if constexpr (OP == 2) return 0.5;
  else if constexpr (OP == 6) return s == 1 ? 0.750000000000000 : (s == 2 ? -0.150000000000000 : 0.016666666666667);
  else return s == 1 ? FDM_DI1 : FDM_DI2;
}

// DI(s) of 'SBP<op>' for s = 1..3, zero beyond the operator's half-width
// (the reference's DI(-3:3) table, fd_coeff.f90:41-50)
inline double sbp_di_host(int op, int s) {
  if (op == 2) return s <= Sbp<2>::half ? sbp_di<2>(s) : 0.0;
  if (op == 6) return s <= Sbp<6>::half ? sbp_di<6>(s) : 0.0;
  return s <= Sbp<4>::half ? sbp_di<4>(s) : 0.0;
}

// artificial dissipation AI(0..2) (fd_coeff.f90:103-104) and AL(s,r)
// (fd_coeff.f90:117-124); AR(-s,-r) = AL(s,r) (fd_coeff.f90:242-246)
#define FDM_AI0 (-0.500000000000000)
#define FDM_AI1 0.333333333333333
#define FDM_AI2 (-0.083333333333333)
__device__ constexpr double kAL[4][6] = {
    {-0.470588235294118, 0.941176470588235, -0.470588235294118, 0.000000000000000,
     0.000000000000000, 0.000000000000000},
    {0.271186440677966, -0.610169491525424, 0.406779661016949, -0.067796610169492,
     0.000000000000000, 0.000000000000000},
    {-0.186046511627907, 0.558139534883721, -0.651162790697674, 0.372093023255814,
     -0.093023255813953, 0.000000000000000},
    {0.000000000000000, -0.081632653061224, 0.326530612244898, -0.489795918367347,
     0.326530612244898, -0.081632653061224}};

struct NT {
  double vt, vn, stt, snt, snn, szz;
};

// rotate_fields_xy2nt_mode2 (fields.f90:1312-1325, geometry.f90:285-322)
__device__ __forceinline__ NT xy2nt(const double* F, double n1, double n2) {
  NT r;
  r.vt = n2 * F[0] - n1 * F[1];
  r.vn = n1 * F[0] + n2 * F[1];
  r.stt = -2.0 * n1 * n2 * F[3] + n2 * n2 * F[2] + n1 * n1 * F[4];
  r.snt = n1 * n2 * (F[2] - F[4]) + (n2 * n2 - n1 * n1) * F[3];
  r.snn = 2.0 * n1 * n2 * F[3] + n1 * n1 * F[2] + n2 * n2 * F[4];
  r.szz = F[5];
  return r;
}

// rotate_fields_nt2xy_mode2 (fields.f90:1343-1356, geometry.f90:298-336)
__device__ __forceinline__ void nt2xy(double* F, double n1, double n2, double vt, double vn,
                                      double stt, double snt, double snn, double szz) {
  F[0] = n2 * vt + n1 * vn;
  F[1] = -n1 * vt + n2 * vn;
  F[2] = 2.0 * n1 * n2 * snt + n2 * n2 * stt + n1 * n1 * snn;
  F[3] = -n1 * n2 * (stt - snn) + (n2 * n2 - n1 * n1) * snt;
  F[4] = -2.0 * n1 * n2 * snt + n1 * n1 * stt + n2 * n2 * snn;
  F[5] = szz;
}

__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }

// arcsinh (friction.f90:1293-1304)
__device__ __forceinline__ double arcsinh_ref(double x) {
  double y = fabs(x);
  double f = log(y + sqrt(y * y + 1.0));
  return fsign(f, x);
}

__device__ __forceinline__ double pow8(double x) {
  double x2 = x * x;
  double x4 = x2 * x2;
  return x4 * x4;
}

// mms_sin (mms.f90:52-234).  field: 0..5 fields, 6..11 sources s_*.
__device__ double mms_sin(double x, double y, double t, int side, int field);

// set_bc_mode2 (boundaries.f90:216-377)
__device__ __forceinline__ void set_bc_mode2(double* Fhat, const double* F, const double* F0,
                                             double n1, double n2, int bc, double Zs, double Zp,
                                             double gamma, double x, double y, double t,
                                             int iblock) {
  double Zsi = (Zs == 0.0) ? 0.0 : 1.0 / Zs;
  double Zpi = 1.0 / Zp;
  bool subtract = !(bc == FDM_BC_MMS_SIN_W || bc == FDM_BC_RIGID_0 || bc == FDM_BC_ABSORBING_0 ||
                    bc == FDM_BC_FREE_0);
  double Fd[6];
#pragma unroll
  for (int l = 0; l < 6; l++) Fd[l] = subtract ? F[l] - F0[l] : F[l];
  NT fd = xy2nt(Fd, n1, n2);
  double vt = 0.0, vn = 0.0, snt = 0.0, snn = 0.0;
  switch (bc) {
    case FDM_BC_MMS_SIN_W: {
      double FEX[6];
      for (int l = 0; l < 5; l++) FEX[l] = mms_sin(x, y, t, iblock, l);
      FEX[5] = 0.0;
      NT ex = xy2nt(FEX, n1, n2);
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
  double stt = fd.stt - gamma * (fd.snn - snn);
  double szz = fd.szz - gamma * (fd.snn - snn);
  nt2xy(Fhat, n1, n2, vt, vn, stt, snt, snn, szz);
  if (subtract) {
#pragma unroll
    for (int l = 0; l < 6; l++) Fhat[l] = Fhat[l] + F0[l];
  }
}

// ---- antiplane shear (mode 3): vz, sxz, syz ----

// rotate_fields_xy2nt_mode3 (fields.f90:1297-1309): vz, and (sxz, syz) as a
// vector (rotate_xy2nt_vec, geometry.f90:285-295)
__device__ __forceinline__ void xy2nt3(const double* F, double n1, double n2, double& vz,
                                       double& stz, double& snz) {
  vz = F[0];
  stz = n2 * F[1] - n1 * F[2];
  snz = n1 * F[1] + n2 * F[2];
}

// rotate_fields_nt2xy_mode3 (fields.f90:1328-1340, geometry.f90:298-308)
__device__ __forceinline__ void nt2xy3(double* F, double n1, double n2, double vz, double stz,
                                       double snz) {
  F[0] = vz;
  F[1] = n2 * stz + n1 * snz;
  F[2] = -n1 * stz + n2 * snz;
}

// set_bc_mode3 (boundaries.f90:155-213): absorbing, free or rigid
__device__ __forceinline__ void set_bc_mode3(double* Fhat, const double* F, const double* F0,
                                             double n1, double n2, int bc, double Zs) {
  if (Zs == 0.0) {
#pragma unroll
    for (int l = 0; l < 3; l++) Fhat[l] = F[l];
    return;
  }
  const double Zsi = 1.0 / Zs;
  double d[3], vzFD, stzFD, snzFD;
#pragma unroll
  for (int l = 0; l < 3; l++) d[l] = F[l] - F0[l];
  xy2nt3(d, n1, n2, vzFD, stzFD, snzFD);
  double vz = 0.0, snz = 0.0;
  if (bc == FDM_BC_ABSORBING) {
    vz = 0.5 * (vzFD - Zsi * snzFD);
    snz = 0.5 * (snzFD - Zs * vzFD);
  } else if (bc == FDM_BC_FREE) {
    vz = vzFD - Zsi * snzFD;
    snz = 0.0;
  } else if (bc == FDM_BC_RIGID) {
    vz = 0.0;
    snz = snzFD - Zs * vzFD;
  }
  nt2xy3(Fhat, n1, n2, vz, stzFD, snz);
#pragma unroll
  for (int l = 0; l < 3; l++) Fhat[l] = Fhat[l] + F0[l];
}

// locked_interface_mode3 (interfaces.f90:676-712)
__device__ __forceinline__ void locked_interface_mode3(double* Fhatm, double* Fhatp, const double* Fm,
                                                       const double* Fp, double n1, double n2,
                                                       double Zsm, double Zsp) {
  double vzFDp, stzFDp, snzFDp, vzFDm, stzFDm, snzFDm;
  xy2nt3(Fp, n1, n2, vzFDp, stzFDp, snzFDp);
  xy2nt3(Fm, n1, n2, vzFDm, stzFDm, snzFDm);
  const double Zsip = 1.0 / Zsp, Zsim = 1.0 / Zsm;
  const double etas = 1.0 / (Zsip + Zsim);
  const double phis = etas * (snzFDp * Zsip + snzFDm * Zsim + vzFDp - vzFDm);
  const double snzp = phis, snzm = phis;
  const double vzp = vzFDp + Zsip * (snzFDp - snzp);
  const double vzm = vzFDm - Zsim * (snzFDm - snzm);
  nt2xy3(Fhatp, n1, n2, vzp, stzFDp, snzp);
  nt2xy3(Fhatm, n1, n2, vzm, stzFDm, snzm);
}

// locked_interface_mode2 (interfaces.f90:715-772)
__device__ __forceinline__ void locked_interface_mode2(double* Fhatm, double* Fhatp,
                                                       const double* Fm, const double* Fp,
                                                       double n1, double n2, double Zsm,
                                                       double Zsp, double Zpm, double Zpp,
                                                       double gammam, double gammap) {
  NT p = xy2nt(Fp, n1, n2);
  NT m = xy2nt(Fm, n1, n2);
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

// strength (friction.f90:1041-1074).  E = exp(Psi/a), evaluated once per
// node and stage by the caller (the same value the reference recomputes
// on every call).
__device__ __forceinline__ void strength(int law, double a, double V0, double V, double Psi,
                                         double E, double N, double& S, double& dSdV) {
  if (law == FDM_LAW_SL || law == FDM_LAW_FL) {
    if (V == 0.0) {
      S = 0.0;
      dSdV = 1e10;
    } else {
      double absV = fabs(V);
      double f = a * log(absV / V0) + Psi;
      S = N * fsign(f, V);
      dSdV = N * a / absV;
    }
  } else {
    double O = 0.5 / V0 * E;
    S = a * N * arcsinh_ref(O * V);
    dSdV = a * N * O / sqrt(1.0 + (V * O) * (V * O));
  }
}

// estimate_V (friction.f90:1077-1098)
__device__ __forceinline__ void estimate_V(int law, double a, double V0, double& V, double Slock,
                                           double N, double Psi) {
  double SaN = Slock / (a * N);
  if (fabs(SaN) > 10.0) return;
  if (law == FDM_LAW_SL || law == FDM_LAW_FL) {
    V = fsign(V0 * exp(-Psi / a) * exp(fabs(SaN)), SaN);
  } else if (law == FDM_LAW_RSL || law == FDM_LAW_RFL || law == FDM_LAW_RAL) {
    V = 2.0 * V0 * exp(-Psi / a) * sinh(SaN);
  }
}

// newton_solver (friction.f90:909-1038): returns true on failure
__device__ __forceinline__ bool newton_solver(int law, double a, double V0, double& Vio,
                                             double& Sio, double Slock, double N, double eta,
                                             double Psi, double E) {
  const int imax = 99;
  const double atolR = 2.220446049250313e-16 * 1e4;
  double V = Vio, S = Sio;
  double Vmin = fmin(0.0, Slock / eta);
  double Vmax = fmax(0.0, Slock / eta);
  estimate_V(law, a, V0, V, Slock, N, Psi);
  if (V <= Vmin || Vmax <= V) V = 0.5 * (Vmin + Vmax);
  int i;
  for (i = 1; i <= imax; i++) {
    S = Slock - eta * V;
    double dSdV = -eta;
    double Sf, dSfdV;
    strength(law, a, V0, V, Psi, E, N, Sf, dSfdV);
    double R = -S + Sf;
    if (R > 0.0) Vmax = V;
    if (R < 0.0) Vmin = V;
    double dRdV = -dSdV + dSfdV;
    double dV = -R / dRdV;
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
    if (fabs(dV) <= 0.0) break;
  }
  Vio = V;
  Sio = S;
  return i == imax;
}

// state_rate (friction.f90:1147-1264); E = exp(Psi/a) as for strength
__device__ __forceinline__ double state_rate(int law, const double* rs, double V, double Psi,
                                             double E) {
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
      frs = a * arcsinh_ref(0.5 * absV / V0 * E);
      return -(absV / L) * (frs - fss);
    case FDM_LAW_RFL:
      fLV = f0 - (b - a) * log(absV / V0);
      fss = fw + (fLV - fw) / pow(1.0 + pow8(V / Vw), 0.125);
      frs = a * arcsinh_ref(0.5 * absV / V0 * E);
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

// solve_friction after load_stress (friction.f90:655-826).  Returns true on
// a second Newton failure (the reference stops the run there).
// kinematic forcing (friction.f90:717-745): max(fkm, fkp) of a node at
// distance xx along the fault at time t; kn = c0, fd, dfdx, vL, vR, xL, xR,
// tmax
__device__ __forceinline__ double kinematic_friction(const double* kn, double xx, double t) {
  const double xm = kn[5] + kn[3] * fmin(t, kn[7]);
  const double fkm = xx < xm ? kn[1] + kn[2] * (xm - xx) : kn[1];
  const double xp = kn[6] + kn[4] * fmin(t, kn[7]);
  const double fkp = xx > xp ? kn[1] + kn[2] * (xx - xp) : kn[1];
  return fmax(fkm, fkp);
}

// force: the forced strength Sk = kc0 + Nreg kf caps the slip-weakening
// strength (friction.f90:765-767) and replaces a larger rate-and-state
// strength (friction.f90:808-815)
__device__ __forceinline__ bool solve_friction_core(int law, const double* rs, const double* sw,
                                                    int reg, int outside, double flock,
                                                    double S0i, double N0i, double phip,
                                                    double phis, double eta, double D,
                                                    double Psi, double E, double skempton,
                                                    double dt, double& V, double& S, double& O,
                                                    double& N, double& Nreg, int force = 0,
                                                    double kc0 = 0.0, double kf = 0.0) {
  double Slock = S0i + phis;
  double Nlock = N0i - (1.0 - skempton) * phip;
  O = 0.0;
  N = Nlock;
  if (Nlock <= 0.0) {
    S = 0.0;
    V = Slock / eta;
    return false;
  }
  if (dt == 0.0) Nreg = N;
  if (!reg) Nreg = N;
  if (outside) {
    double Sv = flock * Nreg;
    if (fabs(Slock) >= Sv) {
      Sv = fsign(Sv, Slock);
      V = (Slock - Sv) / eta;
      S = Sv;
    } else {
      V = 0.0;
      S = Slock;
    }
    return false;
  }
  switch (law) {
    case FDM_LAW_FRICTIONLESS:
      S = 0.0;
      V = Slock / eta;
      return false;
    case FDM_LAW_LOCKED:
      V = 0.0;
      S = Slock;
      return false;
    case FDM_LAW_SW: {
      double c0 = sw[0], fs = sw[1], fd = sw[2], Dc = sw[3];
      double Sv = c0 + fmax(Nreg, 0.0) * ((fs - (fs - fd) * fmin(D, Dc) / Dc));
      if (force) Sv = fmin(Sv, kc0 + Nreg * kf);
      if (fabs(Slock) >= Sv) {
        Sv = fsign(Sv, Slock);
        V = (Slock - Sv) / eta;
        S = Sv;
      } else {
        V = 0.0;
        S = Slock;
      }
      return false;
    }
    default: {
      bool fail = newton_solver(law, rs[0], rs[2], V, S, Slock, Nreg, eta, Psi, E);
      if (fail) fail = newton_solver(law, rs[0], rs[2], V, S, Slock, Nreg, eta, Psi, E);
      if (force) {
        const double Sk = kc0 + Nreg * kf;
        if (Sk < fabs(S)) {
          S = fsign(Sk, Slock);
          V = (Slock - S) / eta;
        }
      }
      return fail;
    }
  }
}

}  // namespace fdm
