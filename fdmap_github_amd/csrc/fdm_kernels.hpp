// fdm_kernels.hpp -- device kernels of the time-stepping path.
//
//   edge_kernel  : one lane per boundary node or interface node.  Boundary
//                  hats (set_bc_mode2), locked interface hats, and for
//                  frictional interfaces the fault-rate scaling, the
//                  rate-and-state solve, hats, state rates and the fault RK
//                  update, all in FP64.
//   bulk_kernel  : one lane per grid point.  DF <- A*DF, SAT penalties,
//                  SBP2/4/6 stencils with curvilinear metrics, szz overwrite,
//                  sources, F <- F + B*dt*DF and (last stage) the
//                  Drucker-Prager return map.  Templated on the storage
//                  type T of fields/rates/metrics (double, or float for the
//                  FP32-bulk mode).
#pragma once

#include "fdm_physics.hpp"

namespace fdm {

struct BlockDev {
  int mx, px, my, py;
  int plastic, mms, id, pad;
  double rho, G, M, gamma, nu;
  double cI[5][3];  // interior coefficients DI(s)/rho, DI(s)*M, DI(s)*G, s = 1..3 (rates.f90:400-405)
  double mu, beta, b, h, eta, K;
  double s0[6];
  int side_off[4], side_n[4], side_bc[4];
  // FP32-bulk stores each field as its deviation from this per-block
  // constant (zero in FP64); totals are formed in FP64 where needed
  double Fb[6];
  // artificial dissipation (rates.f90:1397-2021): Cdiss as the rates use it
  // (0: none) and the interior weights d(0) = 2 AI(0) Cdiss, d(s) =
  // AI(s) Cdiss (rates.f90:1416-1421)
  double Cdiss;
  double dA[3];
  // perfectly matched layer (rates.f90:2024-2619): pmlx/pmly are zero unless
  // the layer is in effect (mode 2, no dissipation in the block); pml = n,
  // sigma, kappa, mx, px, my, py; phx/phy: Ji*yr and Ji*xq at the first
  // point of each of the block's nine regions, index 3*ry + rx with rx, ry
  // 0 (minus-side closure), 1 (interior), 2 (plus-side closure)
  int pmlx, pmly;
  double pml[7];
  double phx[9], phy[9];
  // FP32-delta metrics (FP64 contexts): xq, xr, yq, yr of a point of this
  // block are mbase[q] + (double)d with d the stored FP32 delta, exactly; a
  // NaN delta sends the reader to the FP64 plane
  double mbase[4];
};

struct IfaceDev {
  int direction, bm, bp, m, p, coupling, law, reg, rup_field, node_off, n, pad;
  double rup_threshold, skempton, flock, ld_T;
  // kinematic forcing: c0, fd, dfdx, vL, vR, xL, xR, tmax
  int force, pad2;
  double kn[8];
};

struct SourcePlaceDev {
  int block, i0, j0, pad;
  double dx[5], dy[5];
};

// concatenated per-node arrays of all block sides; component-major, stride NS
struct SidesDev {
  const double* M;       // [5][NS]
  const double* normal;  // [2][NS]
  const double* F0;      // [6][NS]
  const double* x;       // [NS] (may be null)
  const double* y;
  double* Fhat;          // [6][NS]
  int NS;
};

// concatenated per-node arrays of all interfaces; component-major, stride NI
struct IfacesDev {
  const IfaceDev* tab;
  const double* nhat;    // [2][NI]
  const double* rs;      // [7][NI]
  const double* sw;      // [4][NI]
  const double* S0b;     // [NI]
  const double* N0b;
  const double* ldS;
  const double* ldN;
  const int* outside;
  const double* knx;     // [NI] distance along the fault (kinematic forcing)
  double* st;            // [18][NI] fault state (FDM_IF_* order)
  double* st1;           // [18][NI] stage-1 snapshot
  int NI;
};

template <typename T>
struct EdgeArgs {
  const T* F;
  int ld;
  long long ps;
  const int2* tasks;  // .x: 4*block+side for a boundary node, -(1+iface) for an interface node
  int ntasks;
  const BlockDev* blocks;
  SidesDev sd;
  IfacesDev fi;
  double t, Bdt, A, Bload;
  int do_scale, do_update, snapshot;
  int* fail;
  // graph replays: when set, the stage time is formed on the device from the
  // step index *nstep as (t_init + (nstep - 1) dt) + rkc dt, the expression
  // the host uses, instead of t
  const long long* nstep;
  double t_init, dt, rkc;
  // when set, [3][NI]: slip velocity, state variable at the start of the
  // stage and normal stress of every frictional node (adjoint trajectories)
  double* trace;
  // 2: plane strain; 3: antiplane shear (fields vz, sxz, syz)
  int mode;
};

template <typename T>
__device__ __forceinline__ double stage_time(const EdgeArgs<T>& a) {
  if (!a.nstep) return a.t;
  const double t0 = a.t_init + (double)(*a.nstep - 1) * a.dt;
  return t0 + a.rkc * a.dt;
}

// interface node in antiplane shear: locked_interface_mode3
// (interfaces.f90:676-712), or frictional_interface_mode3 (interfaces.f90:
// 806-897) with the fault bookkeeping of the plane-strain path.  The normal
// stress is the plus block's in-plane prestress resolved on the fault
// (rotate_xy2nt_mat, geometry.f90:311-322), the stress transfer phip is zero,
// and the initial fields arrive in their own order (interfaces.f90:798).
template <typename T, bool FORCE>
__device__ __forceinline__ void edge_iface3(const EdgeArgs<T>& a, const IfaceDev& I, const int q,
                                            const BlockDev& Bm, const BlockDev& Bp, const int som,
                                            const int sop, const long long ccm, const long long ccp,
                                            const double tst) {
  const int NS = a.sd.NS, NI = a.fi.NI;
  double Fm[3], Fp[3], F0m[3], F0p[3], Fhm[3], Fhp[3];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    Fm[l] = (double)a.F[l * a.ps + ccm];
    Fp[l] = (double)a.F[l * a.ps + ccp];
    F0m[l] = a.sd.F0[l * NS + som];
    F0p[l] = a.sd.F0[l * NS + sop];
  }
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int l = 0; l < 3; l++) {
      Fm[l] = Fm[l] + Bm.Fb[l];
      Fp[l] = Fp[l] + Bp.Fb[l];
    }
  }
  const double n1 = a.fi.nhat[q], n2 = a.fi.nhat[NI + q];
  const double Zsm = a.sd.M[som], Zsp = a.sd.M[sop];
  if (I.coupling == FDM_COUPLING_LOCKED) {
    locked_interface_mode3(Fhm, Fhp, Fm, Fp, n1, n2, Zsm, Zsp);
  } else {
    double* st = a.fi.st;
#define ST(f) st[(long long)(f) * NI + q]
    if (a.do_scale) {
      ST(FDM_IF_DDS) = a.A * ST(FDM_IF_DDS);
      ST(FDM_IF_DDN) = a.A * ST(FDM_IF_DDN);
      ST(FDM_IF_DPSI) = a.A * ST(FDM_IF_DPSI);
      ST(FDM_IF_DW) = a.A * ST(FDM_IF_DW);
      ST(FDM_IF_DNREG) = a.A * ST(FDM_IF_DNREG);
    }
    double vz0p, stz0p, snz0p, vz0m, stz0m, snz0m;
    xy2nt3(F0p, n1, n2, vz0p, stz0p, snz0p);
    xy2nt3(F0m, n1, n2, vz0m, stz0m, snz0m);
    double dp[3], dm[3];
#pragma unroll
    for (int l = 0; l < 3; l++) {
      dp[l] = Fp[l] - F0p[l];
      dm[l] = Fm[l] - F0m[l];
    }
    double vzFDp, stzFDp, snzFDp, vzFDm, stzFDm, snzFDm;
    xy2nt3(dp, n1, n2, vzFDp, stzFDp, snzFDp);
    xy2nt3(dm, n1, n2, vzFDm, stzFDm, snzFDm);
    const double Zsip = 1.0 / Zsp, Zsim = 1.0 / Zsm;
    const double etas = 1.0 / (Zsip + Zsim);
    const double phis = etas * (snzFDp * Zsip + snzFDm * Zsim + vzFDp - vzFDm);
    const double S0 = snz0p;
    const double snn = 2.0 * n1 * n2 * Bp.s0[1] + n1 * n1 * Bp.s0[0] + n2 * n2 * Bp.s0[3];
    const double N0 = -snn - 0.0;
    const double phip = 0.0;
    const double S0l = a.fi.S0b[q] + a.fi.ldS[q] * a.Bload;
    const double N0l = a.fi.N0b[q] + a.fi.ldN[q] * a.Bload;
    const double fS0 = S0l + S0;
    const double fN0 = N0l + N0;
    ST(FDM_IF_S0) = fS0;
    ST(FDM_IF_N0) = fN0;
    double rs[7], sw[4];
#pragma unroll
    for (int c = 0; c < 7; c++) rs[c] = I.law >= FDM_LAW_SL ? a.fi.rs[(long long)c * NI + q] : 1.0;
#pragma unroll
    for (int c = 0; c < 4; c++) sw[c] = I.law == FDM_LAW_SW ? a.fi.sw[(long long)c * NI + q] : 1.0;
    double V = ST(FDM_IF_V), S = ST(FDM_IF_S), O = ST(FDM_IF_O), N = ST(FDM_IF_N),
           Nreg = ST(FDM_IF_NREG);
    const double D = ST(FDM_IF_D), Psi = ST(FDM_IF_PSI);
    const double E = I.law >= FDM_LAW_RSL ? exp(Psi / rs[0]) : 0.0;
    const int force = FORCE ? I.force : 0;
    const double kf = force ? kinematic_friction(I.kn, a.fi.knx[q], tst) : 0.0;
    const bool fail = solve_friction_core(I.law, rs, sw, I.reg, a.fi.outside[q], I.flock, fS0, fN0,
                                          phip, phis, etas, D, Psi, E, I.skempton, a.Bdt, V, S, O, N,
                                          Nreg, force, I.kn[0], kf);
    if (fail) *a.fail = 1;
    ST(FDM_IF_V) = V;
    ST(FDM_IF_S) = S;
    ST(FDM_IF_O) = O;
    ST(FDM_IF_N) = N;
    ST(FDM_IF_NREG) = Nreg;
    const double snzp = S - fS0, snzm = S - fS0;
    const double vzp = vzFDp + Zsip * (snzFDp - snzp);
    const double vzm = vzFDm - Zsim * (snzFDm - snzm);
    nt2xy3(Fhp, n1, n2, vzp, stzFDp, snzp);
    nt2xy3(Fhm, n1, n2, vzm, stzFDm, snzm);
#pragma unroll
    for (int l = 0; l < 3; l++) {
      Fhp[l] = Fhp[l] + F0p[l];
      Fhm[l] = Fhm[l] + F0m[l];
    }
    // set_rates_friction (friction.f90:1101-1124)
    ST(FDM_IF_DDS) = ST(FDM_IF_DDS) + V;
    ST(FDM_IF_DDN) = ST(FDM_IF_DDN) + O;
    ST(FDM_IF_DPSI) = ST(FDM_IF_DPSI) + state_rate(I.law, rs, V, Psi, E);
    if (I.reg) {
      const double absV = fabs(V);
      const double Nr = -(absV / rs[4]) * (Nreg - fmax(N, 0.0));
      ST(FDM_IF_DNREG) = ST(FDM_IF_DNREG) + Nr;
    }
    ST(FDM_IF_DW) = ST(FDM_IF_DW) + S * V;
    const bool ruptured = I.rup_field == 1 ? (D > I.rup_threshold) : (fabs(V) > I.rup_threshold);
    if (ruptured) ST(FDM_IF_TRUP) = fmin(ST(FDM_IF_TRUP), tst);
    if (a.snapshot) {
      for (int f = 0; f < FDM_IF_NFIELDS; f++) a.fi.st1[(long long)f * NI + q] = st[(long long)f * NI + q];
    }
    if (a.do_update) {
      const double Bdt = a.Bdt;
      ST(FDM_IF_DS) = ST(FDM_IF_DS) + Bdt * ST(FDM_IF_DDS);
      ST(FDM_IF_DN) = ST(FDM_IF_DN) + Bdt * ST(FDM_IF_DDN);
      ST(FDM_IF_D) = ST(FDM_IF_D) + Bdt * fabs(ST(FDM_IF_DDS));
      ST(FDM_IF_PSI) = ST(FDM_IF_PSI) + Bdt * ST(FDM_IF_DPSI);
      ST(FDM_IF_W) = ST(FDM_IF_W) + Bdt * ST(FDM_IF_DW);
      ST(FDM_IF_NREG) = ST(FDM_IF_NREG) + Bdt * ST(FDM_IF_DNREG);
    }
#undef ST
  }
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int l = 0; l < 3; l++) {
      Fhm[l] = Fhm[l] - Bm.Fb[l];
      Fhp[l] = Fhp[l] - Bp.Fb[l];
    }
  }
#pragma unroll
  for (int l = 0; l < 3; l++) {
    a.sd.Fhat[l * NS + som] = Fhm[l];
    a.sd.Fhat[l * NS + sop] = Fhp[l];
  }
}

// MODE 2 or 3: each kernel carries the code of one mode only; FORCE = false
// leaves out kinematic forcing (the host selects it when no interface forces)
template <typename T, int MODE, bool FORCE>
__device__ __forceinline__ void edge_task(const EdgeArgs<T>& a, const int2 task) {
  const double tst = stage_time(a);
  const int NS = a.sd.NS;
  if (task.x >= 0) {
    // boundary node (boundaries.f90:106-129)
    int b = task.x >> 2, s = task.x & 3, k = task.y;
    const BlockDev& B = a.blocks[b];
    int i, j;
    if (s == FDM_SIDE_L) { i = B.mx; j = B.my + k; }
    else if (s == FDM_SIDE_R) { i = B.px; j = B.my + k; }
    else if (s == FDM_SIDE_B) { i = B.mx + k; j = B.my; }
    else { i = B.mx + k; j = B.py; }
    int so = B.side_off[s] + k;
    long long cc = (long long)j * a.ld + i;
    if constexpr (MODE == 3) {
      // set_bc_mode3 (boundaries.f90:148-149, 155-213)
      double F3[3], F03[3], Fh3[3];
#pragma unroll
      for (int l = 0; l < 3; l++) {
        F3[l] = (double)a.F[l * a.ps + cc];
        F03[l] = a.sd.F0[l * NS + so];
      }
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int l = 0; l < 3; l++) F3[l] = F3[l] + B.Fb[l];
      }
      set_bc_mode3(Fh3, F3, F03, a.sd.normal[so], a.sd.normal[NS + so], B.side_bc[s], a.sd.M[so]);
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int l = 0; l < 3; l++) Fh3[l] = Fh3[l] - B.Fb[l];
      }
#pragma unroll
      for (int l = 0; l < 3; l++) a.sd.Fhat[l * NS + so] = Fh3[l];
      return;
    }
    double F[6], F0[6], Fh[6];
#pragma unroll
    for (int l = 0; l < 6; l++) {
      F[l] = (double)a.F[l * a.ps + cc];
      F0[l] = a.sd.F0[l * NS + so];
    }
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int l = 0; l < 6; l++) F[l] = F[l] + B.Fb[l];
    }
    double x = a.sd.x ? a.sd.x[so] : 0.0, y = a.sd.y ? a.sd.y[so] : 0.0;
    set_bc_mode2(Fh, F, F0, a.sd.normal[so], a.sd.normal[NS + so], B.side_bc[s], a.sd.M[so],
                 a.sd.M[NS + so], a.sd.M[2 * NS + so], x, y, tst, B.id);
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int l = 0; l < 6; l++) Fh[l] = Fh[l] - B.Fb[l];
    }
#pragma unroll
    for (int l = 0; l < 6; l++) a.sd.Fhat[l * NS + so] = Fh[l];
    return;
  }
  // interface node (interfaces.f90:520-1012)
  const IfaceDev& I = a.fi.tab[-task.x - 1];
  int k = task.y;
  int g = I.m + k;
  int q = I.node_off + k;
  const int NI = a.fi.NI;
  const BlockDev& Bm = a.blocks[I.bm];
  const BlockDev& Bp = a.blocks[I.bp];
  int im, jm, ip, jp, km, kp, sm, sp;
  if (I.direction == 0) {
    sm = FDM_SIDE_R; sp = FDM_SIDE_L;
    im = Bm.px; jm = g; ip = Bp.mx; jp = g;
    km = g - Bm.my; kp = g - Bp.my;
  } else {
    sm = FDM_SIDE_T; sp = FDM_SIDE_B;
    im = g; jm = Bm.py; ip = g; jp = Bp.my;
    km = g - Bm.mx; kp = g - Bp.mx;
  }
  int som = Bm.side_off[sm] + km, sop = Bp.side_off[sp] + kp;
  long long ccm = (long long)jm * a.ld + im, ccp = (long long)jp * a.ld + ip;
  if constexpr (MODE == 3) {
    edge_iface3<T, FORCE>(a, I, q, Bm, Bp, som, sop, ccm, ccp, tst);
    return;
  }
  double Fm[6], Fp[6], F0m[6], F0p[6], Fhm[6], Fhp[6];
#pragma unroll
  for (int l = 0; l < 6; l++) {
    Fm[l] = (double)a.F[l * a.ps + ccm];
    Fp[l] = (double)a.F[l * a.ps + ccp];
    F0m[l] = a.sd.F0[l * NS + som];
    F0p[l] = a.sd.F0[l * NS + sop];
  }
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int l = 0; l < 6; l++) {
      Fm[l] = Fm[l] + Bm.Fb[l];
      Fp[l] = Fp[l] + Bp.Fb[l];
    }
  }
  double n1 = a.fi.nhat[q], n2 = a.fi.nhat[NI + q];
  double Zsm = a.sd.M[som], Zsp = a.sd.M[sop];
  double Zpm = a.sd.M[NS + som], Zpp = a.sd.M[NS + sop];
  double gm = a.sd.M[2 * NS + som], gp = a.sd.M[2 * NS + sop];
  if (I.coupling == FDM_COUPLING_LOCKED) {
    locked_interface_mode2(Fhm, Fhp, Fm, Fp, n1, n2, Zsm, Zsp, Zpm, Zpp, gm, gp);
  } else {
    double* st = a.fi.st;
#define ST(f) st[(long long)(f) * NI + q]
    if (a.do_scale) {
      // scale_rates_friction (friction.f90:555-573)
      ST(FDM_IF_DDS) = a.A * ST(FDM_IF_DDS);
      ST(FDM_IF_DDN) = a.A * ST(FDM_IF_DDN);
      ST(FDM_IF_DPSI) = a.A * ST(FDM_IF_DPSI);
      ST(FDM_IF_DW) = a.A * ST(FDM_IF_DW);
      ST(FDM_IF_DNREG) = a.A * ST(FDM_IF_DNREG);
    }
    // frictional_interface_mode2 (interfaces.f90:900-1012).  The reference
    // passes the minus-side initial fields to its dummy "F0p" and the
    // plus-side ones to "F0m" (interfaces.f90:795 vs :900); kept as is.
    const double* dF0p = F0m;
    const double* dF0m = F0p;
    NT p0 = xy2nt(dF0p, n1, n2);
    double dp[6], dm[6];
#pragma unroll
    for (int l = 0; l < 6; l++) {
      dp[l] = Fp[l] - dF0p[l];
      dm[l] = Fm[l] - dF0m[l];
    }
    NT p = xy2nt(dp, n1, n2);
    NT m = xy2nt(dm, n1, n2);
    double Zsip = 1.0 / Zsp, Zsim = 1.0 / Zsm, Zpip = 1.0 / Zpp, Zpim = 1.0 / Zpm;
    double etas = 1.0 / (Zsip + Zsim);
    double phis = etas * (p.snt * Zsip + m.snt * Zsim + p.vt - m.vt);
    double etap = 1.0 / (Zpip + Zpim);
    double phip = etap * (p.snn * Zpip + m.snn * Zpim + p.vn - m.vn);
    double S0 = p0.snt;
    double N0 = -p0.snn - 0.0;
    // load_stress (friction.f90:1335-1374) with the ramp B(t) precomputed
    double S0l = a.fi.S0b[q] + a.fi.ldS[q] * a.Bload;
    double N0l = a.fi.N0b[q] + a.fi.ldN[q] * a.Bload;
    double fS0 = S0l + S0;
    double fN0 = N0l + N0;
    ST(FDM_IF_S0) = fS0;
    ST(FDM_IF_N0) = fN0;
    double rs[7], sw[4];
    if (I.law >= FDM_LAW_SL) {
#pragma unroll
      for (int c = 0; c < 7; c++) rs[c] = a.fi.rs[(long long)c * NI + q];
    } else {
#pragma unroll
      for (int c = 0; c < 7; c++) rs[c] = 1.0;
    }
    if (I.law == FDM_LAW_SW) {
#pragma unroll
      for (int c = 0; c < 4; c++) sw[c] = a.fi.sw[(long long)c * NI + q];
    } else {
#pragma unroll
      for (int c = 0; c < 4; c++) sw[c] = 1.0;
    }
    double V = ST(FDM_IF_V), S = ST(FDM_IF_S), O = ST(FDM_IF_O), N = ST(FDM_IF_N),
           Nreg = ST(FDM_IF_NREG);
    double D = ST(FDM_IF_D), Psi = ST(FDM_IF_PSI);
    // exp(Psi/a) of the regularised laws, shared by every strength
    // evaluation of the Newton solve and by the state rate
    const double E = I.law >= FDM_LAW_RSL ? exp(Psi / rs[0]) : 0.0;
    const int force = FORCE ? I.force : 0;
    const double kf = force ? kinematic_friction(I.kn, a.fi.knx[q], tst) : 0.0;
    bool fail = solve_friction_core(I.law, rs, sw, I.reg, a.fi.outside[q], I.flock, fS0, fN0,
                                    phip, phis, etas, D, Psi, E, I.skempton, a.Bdt, V, S, O, N, Nreg,
                                    force, I.kn[0], kf);
    if (fail) *a.fail = 1;
    if (a.trace) {
      a.trace[q] = V;
      a.trace[NI + q] = Psi;
      a.trace[2 * NI + q] = N;
    }
    ST(FDM_IF_V) = V;
    ST(FDM_IF_S) = S;
    ST(FDM_IF_O) = O;
    ST(FDM_IF_N) = N;
    ST(FDM_IF_NREG) = Nreg;
    double skm = 1.0 - I.skempton;
    double snnp = -(N - fN0) / skm;
    double snnm = -(N - fN0) / skm;
    double vnp = p.vn + Zpip * (p.snn - snnp);
    double vnm = m.vn - Zpim * (m.snn - snnm);
    double sttp = p.stt - gp * (p.snn - snnp);
    double sttm = m.stt - gm * (m.snn - snnm);
    double szzp = p.szz - gp * (p.snn - snnp);
    double szzm = m.szz - gm * (m.snn - snnm);
    double sntp = S - fS0;
    double sntm = S - fS0;
    double vtp = p.vt + Zsip * (p.snt - sntp);
    double vtm = m.vt - Zsim * (m.snt - sntm);
    nt2xy(Fhp, n1, n2, vtp, vnp, sttp, sntp, snnp, szzp);
    nt2xy(Fhm, n1, n2, vtm, vnm, sttm, sntm, snnm, szzm);
#pragma unroll
    for (int l = 0; l < 6; l++) {
      Fhp[l] = Fhp[l] + dF0p[l];
      Fhm[l] = Fhm[l] + dF0m[l];
    }
    // set_rates_friction (friction.f90:1101-1124)
    ST(FDM_IF_DDS) = ST(FDM_IF_DDS) + V;
    ST(FDM_IF_DDN) = ST(FDM_IF_DDN) + O;
    ST(FDM_IF_DPSI) = ST(FDM_IF_DPSI) + state_rate(I.law, rs, V, Psi, E);
    if (I.reg) {
      double absV = fabs(V);
      double Nr = -(absV / rs[4]) * (Nreg - fmax(N, 0.0));
      ST(FDM_IF_DNREG) = ST(FDM_IF_DNREG) + Nr;
    }
    ST(FDM_IF_DW) = ST(FDM_IF_DW) + S * V;
    bool ruptured = I.rup_field == 1 ? (D > I.rup_threshold) : (fabs(V) > I.rup_threshold);
    if (ruptured) ST(FDM_IF_TRUP) = fmin(ST(FDM_IF_TRUP), tst);
    if (a.snapshot) {
      for (int f = 0; f < FDM_IF_NFIELDS; f++)
        a.fi.st1[(long long)f * NI + q] = st[(long long)f * NI + q];
    }
    if (a.do_update) {
      // update_fields_friction (friction.f90:576-590)
      double Bdt = a.Bdt;
      ST(FDM_IF_DS) = ST(FDM_IF_DS) + Bdt * ST(FDM_IF_DDS);
      ST(FDM_IF_DN) = ST(FDM_IF_DN) + Bdt * ST(FDM_IF_DDN);
      ST(FDM_IF_D) = ST(FDM_IF_D) + Bdt * fabs(ST(FDM_IF_DDS));
      ST(FDM_IF_PSI) = ST(FDM_IF_PSI) + Bdt * ST(FDM_IF_DPSI);
      ST(FDM_IF_W) = ST(FDM_IF_W) + Bdt * ST(FDM_IF_DW);
      ST(FDM_IF_NREG) = ST(FDM_IF_NREG) + Bdt * ST(FDM_IF_DNREG);
    }
#undef ST
  }
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int l = 0; l < 6; l++) {
      Fhm[l] = Fhm[l] - Bm.Fb[l];
      Fhp[l] = Fhp[l] - Bp.Fb[l];
    }
  }
#pragma unroll
  for (int l = 0; l < 6; l++) {
    a.sd.Fhat[l * NS + som] = Fhm[l];
    a.sd.Fhat[l * NS + sop] = Fhp[l];
  }
}

template <typename T, int MODE, bool FORCE>
__global__ void __launch_bounds__(64) edge_kernel(EdgeArgs<T> a) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= a.ntasks) return;
  edge_task<T, MODE, FORCE>(a, a.tasks[tid]);
}

// Batched edge kernel: tasks[k] = (task.x, task.y, problem, valid) over the
// edge tasks of many problems, every wave's 64 entries of one problem
// (padding entries have valid = 0).
template <typename T, bool FORCE>
__global__ void __launch_bounds__(64) edge_kernel_batch(const EdgeArgs<T>* __restrict__ tab,
                                                        const int4* __restrict__ tasks, int n) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n) return;
  const int4 t = tasks[tid];
  const int m = __builtin_amdgcn_readfirstlane(t.z);
  if (!t.w) return;
  edge_task<T, 2, FORCE>(tab[m], make_int2(t.x, t.y));
}

}  // namespace fdm
