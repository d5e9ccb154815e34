// fdm_bulk.hpp -- the fused bulk kernel: RK scaling, SAT penalties, SBP4
// curvilinear stencils (rates.f90:384-936), szz overwrite, sources, RK
// update and plasticity, one lane per grid point.
#pragma once

#include "fdm_kernels.hpp"

namespace fdm {

template <typename T>
struct BulkArgs {
  const T* Fin;
  T* Fout;
  T* DF;
  const T* xq;
  const T* xr;
  const T* yq;
  const T* yr;
  const T* Ji;
  const double* X;  // coordinates, MMS sources only
  const double* Y;
  int nx, ny, ld;
  long long ps;
  const int* colb;  // [nx] block column of each i
  const int* rowb;  // [ny] block row of each j
  const int* blk;   // [nby][nbx] block index
  int nbx;
  const BlockDev* blocks;
  SidesDev sd;
  double A;
  int a_zero;
  double Bdt;
  int do_update;
  double t;
  const SourcePlaceDev* src;
  int nplace;
  double srcA, Fx, Fy, Mxx, Mxy, Myy;
  int do_plastic;
  double dt;
  double* gammap;
  double* lambda;
  int xcd_swizzle;  // remap workgroups so each XCD (and its L2) owns contiguous row bands
};

// Map the hardware workgroup id to a tile.  Workgroups are dispatched to the
// 8 XCDs round-robin; with the swizzle, XCD x receives a contiguous range of
// tiles (x-fastest), so the rows a tile shares with its neighbours above and
// below are read through the same L2.
__device__ __forceinline__ void tile_of(int swz, int& tx, int& ty) {
  if (!swz) {
    tx = blockIdx.x;
    ty = blockIdx.y;
    return;
  }
  const int nwg = gridDim.x * gridDim.y;
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  const int q = nwg >> 3, r = nwg & 7, x = bid & 7, k = bid >> 3;
  const int t = (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
  tx = t % gridDim.x;
  ty = t / gridDim.x;
}

// field coupling of rates_*_mode2 (rates.f90:411-444): DF(l) is built from
// F(UA) paired with y-metrics and F(UB) paired with x-metrics; gamma
// multiplies the A or the B product.
template <int l>
struct Fld;
template <>
struct Fld<0> { static constexpr int ua = 2, ub = 3, ga = 0, gb = 0, mat = 0; };
template <>
struct Fld<1> { static constexpr int ua = 3, ub = 4, ga = 0, gb = 0, mat = 0; };
template <>
struct Fld<2> { static constexpr int ua = 0, ub = 1, ga = 0, gb = 1, mat = 2; };
template <>
struct Fld<3> { static constexpr int ua = 1, ub = 0, ga = 0, gb = 0, mat = 1; };
template <>
struct Fld<4> { static constexpr int ua = 0, ub = 1, ga = 1, gb = 0, mat = 2; };

template <typename T, int g>
__device__ __forceinline__ T prod(T m, T f, T gamma) {
  if constexpr (g) return m * f * gamma;
  else return m * f;
}

// ")*D/rho*Ji", ")*D*G*Ji", ")*D*M*Ji" of rates_boundary_mode2
template <typename T, int mat>
__device__ __forceinline__ T bscale(T t, T D, T rho, T G, T M) {
  if constexpr (mat == 0) return t * D / rho;
  else if constexpr (mat == 1) return t * D * G;
  else return t * D * M;
}

struct Pt {
  long long cc;  // centre
  int ld;
  long long ps;
};

// two-product one-sided q term at neighbour kk
template <typename T, int l>
__device__ __forceinline__ T q1(const T* __restrict__ F, const T* __restrict__ yr,
                                const T* __restrict__ xr, long long ps, long long kk,
                                long long cc, T gamma) {
  using P = Fld<l>;
  T ay = (T)0.5 * (yr[kk] + yr[cc]);
  T ax = (T)0.5 * (xr[kk] + xr[cc]);
  return prod<T, P::ga>(ay, F[P::ua * ps + kk], gamma) - prod<T, P::gb>(ax, F[P::ub * ps + kk], gamma);
}

// two-product one-sided r term at neighbour kk
template <typename T, int l>
__device__ __forceinline__ T r1(const T* __restrict__ F, const T* __restrict__ yq,
                                const T* __restrict__ xq, long long ps, long long kk,
                                long long cc, T gamma) {
  using P = Fld<l>;
  T by = (T)0.5 * (yq[kk] + yq[cc]);
  T bx = (T)0.5 * (xq[kk] + xq[cc]);
  return -prod<T, P::ga>(by, F[P::ua * ps + kk], gamma) + prod<T, P::gb>(bx, F[P::ub * ps + kk], gamma);
}

// four-product central q term (B/T strips)
template <typename T, int l>
__device__ __forceinline__ T qc(const T* __restrict__ F, const T* __restrict__ yr,
                                const T* __restrict__ xr, long long ps, long long kp,
                                long long km, long long cc, T gamma) {
  using P = Fld<l>;
  T ayp = (T)0.5 * (yr[kp] + yr[cc]), axp = (T)0.5 * (xr[kp] + xr[cc]);
  T aym = (T)0.5 * (yr[km] + yr[cc]), axm = (T)0.5 * (xr[km] + xr[cc]);
  return prod<T, P::ga>(ayp, F[P::ua * ps + kp], gamma) - prod<T, P::gb>(axp, F[P::ub * ps + kp], gamma) -
         prod<T, P::ga>(aym, F[P::ua * ps + km], gamma) + prod<T, P::gb>(axm, F[P::ub * ps + km], gamma);
}

// four-product central r term (L/R strips)
template <typename T, int l>
__device__ __forceinline__ T rc(const T* __restrict__ F, const T* __restrict__ yq,
                                const T* __restrict__ xq, long long ps, long long kp,
                                long long km, long long cc, T gamma) {
  using P = Fld<l>;
  T byp = (T)0.5 * (yq[kp] + yq[cc]), bxp = (T)0.5 * (xq[kp] + xq[cc]);
  T bym = (T)0.5 * (yq[km] + yq[cc]), bxm = (T)0.5 * (xq[km] + xq[cc]);
  return -prod<T, P::ga>(byp, F[P::ua * ps + kp], gamma) + prod<T, P::gb>(bxp, F[P::ub * ps + kp], gamma) +
         prod<T, P::ga>(bym, F[P::ua * ps + km], gamma) - prod<T, P::gb>(bxm, F[P::ub * ps + km], gamma);
}

template <typename T>
__device__ __forceinline__ T matsel(int mat, T rho, T G, T M) {
  return mat == 0 ? rho : (mat == 1 ? G : M);
}

// interior point (rates_interior_mode2, rates.f90:407-448), s = 1, 2; the
// s = 3 term of the reference has a zero coefficient and adds nothing.
template <typename T, int l>
__device__ __forceinline__ void interior_field(T* d, const T* __restrict__ F,
                                               const T* __restrict__ xq, const T* __restrict__ xr,
                                               const T* __restrict__ yq, const T* __restrict__ yr,
                                               long long ps, long long cc, int ld, T gamma,
                                               const T (&cI)[2], T Ji) {
  using P = Fld<l>;
#pragma unroll
  for (int s = 1; s <= 2; s++) {
    long long ip = cc + s, im = cc - s, jp = cc + (long long)s * ld, jm = cc - (long long)s * ld;
    T ayp = (T)0.5 * (yr[ip] + yr[cc]), axp = (T)0.5 * (xr[ip] + xr[cc]);
    T aym = (T)0.5 * (yr[im] + yr[cc]), axm = (T)0.5 * (xr[im] + xr[cc]);
    T byp = (T)0.5 * (yq[jp] + yq[cc]), bxp = (T)0.5 * (xq[jp] + xq[cc]);
    T bym = (T)0.5 * (yq[jm] + yq[cc]), bxm = (T)0.5 * (xq[jm] + xq[cc]);
    const T* FA = F + P::ua * ps;
    const T* FB = F + P::ub * ps;
    T t = prod<T, P::ga>(ayp, FA[ip], gamma) - prod<T, P::gb>(axp, FB[ip], gamma) -
          prod<T, P::ga>(aym, FA[im], gamma) + prod<T, P::gb>(axm, FB[im], gamma) -
          prod<T, P::ga>(byp, FA[jp], gamma) + prod<T, P::gb>(bxp, FB[jp], gamma) +
          prod<T, P::ga>(bym, FA[jm], gamma) - prod<T, P::gb>(bxm, FB[jm], gamma);
    d[l] = d[l] + t * cI[s - 1] * Ji;
  }
}

template <typename T>
__device__ __forceinline__ void sat_term(T* d, const T* Fe, const SidesDev& sd, int so) {
  const int NS = sd.NS;
  double Fh[6];
#pragma unroll
  for (int l = 0; l < 6; l++) Fh[l] = sd.Fhat[l * NS + so];
  T n1 = (T)sd.normal[so], n2 = (T)sd.normal[NS + so];
  T Ks = (T)sd.M[3 * NS + so], Kp = (T)sd.M[4 * NS + so];
  T dd[6];
#pragma unroll
  for (int l = 0; l < 6; l++) dd[l] = (T)Fh[l] - Fe[l];
  // split_sp (time_step.f90:836-851)
  T vt = n2 * dd[0] - n1 * dd[1];
  T vn = n1 * dd[0] + n2 * dd[1];
  T stt = (T)(-2.0) * n1 * n2 * dd[3] + n2 * n2 * dd[2] + n1 * n1 * dd[4];
  T snt = n1 * n2 * (dd[2] - dd[4]) + (n2 * n2 - n1 * n1) * dd[3];
  T snn = (T)2.0 * n1 * n2 * dd[3] + n1 * n1 * dd[2] + n2 * n2 * dd[4];
  T szz = dd[5];
  T z = (T)0.0;
  T Fs[6], Fp[6];
  Fs[0] = n2 * vt + n1 * z;
  Fs[1] = -n1 * vt + n2 * z;
  Fs[2] = (T)2.0 * n1 * n2 * snt + n2 * n2 * z + n1 * n1 * z;
  Fs[3] = -n1 * n2 * (z - z) + (n2 * n2 - n1 * n1) * snt;
  Fs[4] = (T)(-2.0) * n1 * n2 * snt + n1 * n1 * z + n2 * n2 * z;
  Fs[5] = z;
  Fp[0] = n2 * z + n1 * vn;
  Fp[1] = -n1 * z + n2 * vn;
  Fp[2] = (T)2.0 * n1 * n2 * z + n2 * n2 * stt + n1 * n1 * snn;
  Fp[3] = -n1 * n2 * (stt - snn) + (n2 * n2 - n1 * n1) * z;
  Fp[4] = (T)(-2.0) * n1 * n2 * z + n1 * n1 * stt + n2 * n2 * snn;
  Fp[5] = szz;
  // SAT_term (time_step.f90:813-833)
#pragma unroll
  for (int l = 0; l < 6; l++) d[l] = d[l] + Ks * Fs[l];
#pragma unroll
  for (int l = 0; l < 6; l++) d[l] = d[l] + Kp * Fp[l];
}

// plastic_flow (plastic.f90:62-192) in FP64
__device__ __forceinline__ void plastic_point(double G, double K, double mu, double beta,
                                              double b, double h, double eta, const double* s0,
                                              double dt, double* s, double* Ds, double& lambda,
                                              double& gammap) {
  const double delta[6] = {1.0, 0.0, 0.0, 1.0, 0.0, 1.0};
  double sa[6], sd[6], W[6];
  sa[0] = s0[0] + s[0];
  sa[1] = s0[1] + s[1];
  sa[2] = s0[2];
  sa[3] = s0[3] + s[2];
  sa[4] = s0[4];
  sa[5] = s0[5] + s[3];
  double sigma = (sa[0] + sa[3] + sa[5]) / 3.0;
  double J2 = ((sa[0] - sa[3]) * (sa[0] - sa[3]) + (sa[3] - sa[5]) * (sa[3] - sa[5]) +
               (sa[5] - sa[0]) * (sa[5] - sa[0])) / 6.0 +
              sa[1] * sa[1] + sa[2] * sa[2] + sa[4] * sa[4];
  double tau = sqrt(J2);
  double bb = b + gammap * h;
  double Y = tau - fmax(-mu * sigma + bb, 0.0);
  if (Y <= 0.0) {
    lambda = 0.0;
    return;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) sd[k] = sa[k] - sigma * delta[k];
  if (tau * mu * beta * K > (mu * sigma - b) * G)
    lambda = (tau + mu * sigma - (b + h * gammap)) / (eta + dt * (h + G + mu * beta * K));
  else
    lambda = tau / (eta + dt * G);
  double lam = lambda;
  gammap = gammap + dt * lam;
  tau = tau - dt * lam * G;
  sigma = sigma - dt * lam * beta * K;
#pragma unroll
  for (int k = 0; k < 6; k++) W[k] = K * beta * delta[k] + G * sd[k] / (tau + dt * lam * G);
#pragma unroll
  for (int k = 0; k < 6; k++) sd[k] = sd[k] * tau / (tau + dt * lam * G);
#pragma unroll
  for (int k = 0; k < 6; k++) sa[k] = sigma * delta[k] + sd[k];
  s[0] = sa[0] - s0[0];
  s[1] = sa[1] - s0[1];
  s[2] = sa[3] - s0[3];
  s[3] = sa[5] - s0[5];
  gammap = gammap + dt * lam;  // second increment, as plastic.f90:177
  Ds[0] = Ds[0] - lam * W[0];
  Ds[1] = Ds[1] - lam * W[1];
  Ds[2] = Ds[2] - lam * W[3];
  Ds[3] = Ds[3] - lam * W[5];
}

// szz overwrite, sources, RK update, plasticity and stores of one point
// (rates.f90:920-936, source.f90:196-296, fields.f90:1225-1242,
// plastic.f90:8-59).  Fc: the point's fields at the start of the stage.
template <typename T>
__device__ __forceinline__ void finish_point(const BulkArgs<T>& a, const BlockDev& B, const int bi,
                                             const int i, const int j, const long long cc,
                                             T (&d)[6], const T (&Fc)[6]) {
  const long long ps = a.ps;
  // rates_szz (rates.f90:920-936): overwrite
  d[5] = (T)B.nu * (d[2] + d[4]);
  // singular source (source.f90:196-254)
  for (int p = 0; p < a.nplace; p++) {
    const SourcePlaceDev& P = a.src[p];
    int di = i - P.i0 + 2, dj = j - P.j0 + 2;
    if (P.block == bi && di >= 0 && di <= 4 && dj >= 0 && dj <= 4) {
      double dx = P.dx[di], dy = P.dy[dj];
      d[0] = d[0] + (T)(a.srcA * a.Fx * dx * dy / B.rho);
      d[1] = d[1] + (T)(a.srcA * a.Fy * dx * dy / B.rho);
      d[2] = d[2] - (T)(a.srcA * a.Mxx * dx * dy);
      d[3] = d[3] - (T)(a.srcA * a.Mxy * dx * dy);
      d[4] = d[4] - (T)(a.srcA * a.Myy * dx * dy);
    }
  }
  // manufactured-solution source (source.f90:256-269)
  if (B.mms) {
    double x = a.X[(long long)j * a.nx + i], y = a.Y[(long long)j * a.nx + i];
#pragma unroll 1
    for (int l = 0; l < 6; l++) d[l] = d[l] + (T)mms_sin(x, y, a.t, B.id, 6 + l);
  }
  if (a.do_update) {
    // update_fields_interior (fields.f90:1225-1242)
    const T Bdt = (T)a.Bdt;
    T f[6];
#pragma unroll
    for (int l = 0; l < 6; l++) f[l] = Fc[l] + Bdt * d[l];
    if (a.do_plastic && B.plastic) {
      // update_fields_plastic (plastic.f90:8-59), full dt, after the last stage
      double s[4], Ds[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        s[k] = (double)f[2 + k];
        Ds[k] = (double)d[2 + k];
      }
      long long pc = (long long)j * a.nx + i;
      double lam = a.lambda[pc], gp = a.gammap[pc];
      plastic_point(B.G, B.K, B.mu, B.beta, B.b, B.h, B.eta, B.s0, a.dt, s, Ds, lam, gp);
      a.lambda[pc] = lam;
      a.gammap[pc] = gp;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        f[2 + k] = (T)s[k];
        d[2 + k] = (T)Ds[k];
      }
    }
#pragma unroll
    for (int l = 0; l < 6; l++) a.Fout[l * ps + cc] = f[l];
  }
#pragma unroll
  for (int l = 0; l < 6; l++) a.DF[l * ps + cc] = d[l];
}

// Everything that happens at one grid point in one stage, reading its
// neighbours directly from global memory: used for points next to block
// edges (one-sided closures in r) and as the reference path.
template <typename T>
__device__ __forceinline__ void point_generic(const BulkArgs<T>& a, const int i, const int j) {
  const int bi = a.blk[a.rowb[j] * a.nbx + a.colb[i]];
  const BlockDev& B = a.blocks[bi];
  const long long ps = a.ps;
  const int ld = a.ld;
  const long long cc = (long long)j * ld + i;
  const T* __restrict__ F = a.Fin;
  T d[6];
  if (a.a_zero) {
#pragma unroll
    for (int l = 0; l < 6; l++) d[l] = (T)0.0;
  } else {
    const T A = (T)a.A;
#pragma unroll
    for (int l = 0; l < 5; l++) d[l] = A * a.DF[l * ps + cc];
    d[5] = (T)0.0;  // DF6 is overwritten by the szz relation below
  }
  const int dl = i - B.mx, dr = B.px - i, db = j - B.my, dtp = B.py - j;
  // SAT penalties (time_step.f90:754-810), sides in the order L, R, B, T
  if (dl == 0 || dr == 0 || db == 0 || dtp == 0) {
    T Fe[6];
#pragma unroll
    for (int l = 0; l < 6; l++) Fe[l] = F[l * ps + cc];
    if (dl == 0) sat_term<T>(d, Fe, a.sd, B.side_off[FDM_SIDE_L] + db);
    if (dr == 0) sat_term<T>(d, Fe, a.sd, B.side_off[FDM_SIDE_R] + db);
    if (db == 0) sat_term<T>(d, Fe, a.sd, B.side_off[FDM_SIDE_B] + dl);
    if (dtp == 0) sat_term<T>(d, Fe, a.sd, B.side_off[FDM_SIDE_T] + dl);
  }
  const T* __restrict__ xq = a.xq;
  const T* __restrict__ xr = a.xr;
  const T* __restrict__ yq = a.yq;
  const T* __restrict__ yr = a.yr;
  const T Ji = a.Ji[cc];
  const T gamma = (T)B.gamma, rho = (T)B.rho, G = (T)B.G, M = (T)B.M;
  const bool qone = (dl < 4) || (dr < 4);
  const bool rone = (db < 4) || (dtp < 4);
  const long long row = (long long)j * ld;
  const long long colm = (long long)B.my * ld + i, colp = (long long)B.py * ld + i;
  if (!qone && !rone) {
    T c0[2] = {(T)B.cI[0][0], (T)B.cI[0][1]};
    T c1[2] = {(T)B.cI[1][0], (T)B.cI[1][1]};
    T c2[2] = {(T)B.cI[2][0], (T)B.cI[2][1]};
    T c3[2] = {(T)B.cI[3][0], (T)B.cI[3][1]};
    T c4[2] = {(T)B.cI[4][0], (T)B.cI[4][1]};
    interior_field<T, 0>(d, F, xq, xr, yq, yr, ps, cc, ld, gamma, c0, Ji);
    interior_field<T, 1>(d, F, xq, xr, yq, yr, ps, cc, ld, gamma, c1, Ji);
    interior_field<T, 2>(d, F, xq, xr, yq, yr, ps, cc, ld, gamma, c2, Ji);
    interior_field<T, 3>(d, F, xq, xr, yq, yr, ps, cc, ld, gamma, c3, Ji);
    interior_field<T, 4>(d, F, xq, xr, yq, yr, ps, cc, ld, gamma, c4, Ji);
  } else if (qone && rone) {
    // corners LB, RB, LT, RT: one-sided in q and r, interleaved per s
#pragma unroll
    for (int s = 0; s < 6; s++) {
      long long kq = (dl < 4) ? row + B.mx + s : row + B.px - s;
      T Dq = (dl < 4) ? (T)kDL[dl][s] : (T)(-kDL[dr][s]);
      long long kr = (db < 4) ? colm + (long long)s * ld : colp - (long long)s * ld;
      T Dr = (db < 4) ? (T)kDL[db][s] : (T)(-kDL[dtp][s]);
      d[0] = d[0] + bscale<T, 0>(q1<T, 0>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(q1<T, 1>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(q1<T, 2>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(q1<T, 3>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(q1<T, 4>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[0] = d[0] + bscale<T, 0>(r1<T, 0>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(r1<T, 1>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(r1<T, 2>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(r1<T, 3>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(r1<T, 4>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
    }
  } else if (qone) {
    // L or R strip: one-sided q, then central r
#pragma unroll
    for (int s = 0; s < 6; s++) {
      long long kq = (dl < 4) ? row + B.mx + s : row + B.px - s;
      T Dq = (dl < 4) ? (T)kDL[dl][s] : (T)(-kDL[dr][s]);
      d[0] = d[0] + bscale<T, 0>(q1<T, 0>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(q1<T, 1>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(q1<T, 2>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(q1<T, 3>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(q1<T, 4>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
    }
#pragma unroll
    for (int s = 1; s <= 2; s++) {
      T D = (s == 1) ? (T)FDM_DI1 : (T)FDM_DI2;
      long long kp = cc + (long long)s * ld, km = cc - (long long)s * ld;
      d[0] = d[0] + bscale<T, 0>(rc<T, 0>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(rc<T, 1>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(rc<T, 2>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(rc<T, 3>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(rc<T, 4>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
    }
  } else {
    // B or T strip: one-sided r, then central q
#pragma unroll
    for (int s = 0; s < 6; s++) {
      long long kr = (db < 4) ? colm + (long long)s * ld : colp - (long long)s * ld;
      T Dr = (db < 4) ? (T)kDL[db][s] : (T)(-kDL[dtp][s]);
      d[0] = d[0] + bscale<T, 0>(r1<T, 0>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(r1<T, 1>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(r1<T, 2>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(r1<T, 3>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(r1<T, 4>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
    }
#pragma unroll
    for (int s = 1; s <= 2; s++) {
      T D = (s == 1) ? (T)FDM_DI1 : (T)FDM_DI2;
      long long kp = cc + s, km = cc - s;
      d[0] = d[0] + bscale<T, 0>(qc<T, 0>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(qc<T, 1>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(qc<T, 2>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(qc<T, 3>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(qc<T, 4>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
    }
  }
  T Fc[6];
#pragma unroll
  for (int l = 0; l < 6; l++) Fc[l] = F[l * ps + cc];
  finish_point<T>(a, B, bi, i, j, cc, d, Fc);
}

template <typename T>
__global__ void __launch_bounds__(256) bulk_kernel(BulkArgs<T> a) {
  int tx, ty;
  tile_of(a.xcd_swizzle, tx, ty);
  const int i = tx * blockDim.x + threadIdx.x;
  const int j = ty * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  point_generic<T>(a, i, j);
}


// Row-marching bulk kernel.  A workgroup owns a strip of TX columns and
// sweeps a segment of rows.  The arrays differentiated in r (F1..F5, xq,
// yq) live in a five-row register window, so each is read once per row;
// the arrays differentiated in q (F1..F5, xr, yr) of the current row are
// exchanged through a double-buffered LDS row with a halo of 5 columns,
// wide enough for the one-sided closures at block x-edges.  Segments that
// touch the first or last four rows of a block (closures in r) use the
// per-point path.  Each point's arithmetic is that of point_generic.
// seg: x = first row, y = last row, z = block row, w = 1 for closure rows.
template <typename T, int TX>
__global__ void __launch_bounds__(TX) bulk_march(BulkArgs<T> a, const int4* __restrict__ segs) {
  constexpr int HALO = 5;
  constexpr int W = TX + 2 * HALO;
  __shared__ T sh[2][7][W];
  const int4 sg = segs[blockIdx.y];
  const int t = threadIdx.x;
  const int i = blockIdx.x * TX + t;
  const bool valid = i < a.nx;
  if (sg.w == 1) {
    if (!valid) return;
    for (int j = sg.x; j <= sg.y; j++) point_generic<T>(a, i, j);
    return;
  }
  const long long ps = a.ps;
  const int ld = a.ld;
  const int ic = valid ? i : a.nx - 1;
  const int bi = a.blk[sg.z * a.nbx + a.colb[ic]];
  const BlockDev& B = a.blocks[bi];
  const int dl = i - B.mx, dr = B.px - i;
  const T* __restrict__ F = a.Fin;
  const T* __restrict__ xq = a.xq;
  const T* __restrict__ yq = a.yq;
  const T* __restrict__ xr = a.xr;
  const T* __restrict__ yr = a.yr;
  const int i0 = blockIdx.x * TX - HALO;  // global column of LDS slot 0
  int hslot = -1, hcol = 0;
  if (t < 2 * HALO) {
    hslot = (t < HALO) ? t : TX + t;
    hcol = min(max(i0 + hslot, 0), a.nx - 1);
  }
  const int c = HALO + t;
  const T gamma = (T)B.gamma, rho = (T)B.rho, G = (T)B.G, M = (T)B.M;
  T cI[5][2];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    cI[l][0] = (T)B.cI[l][0];
    cI[l][1] = (T)B.cI[l][1];
  }
  // window w[k][r]: array k (F1..F5, xq, yq) at row j-2+r
  T w[7][5];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const long long idx = (long long)(sg.x - 2 + r) * ld + ic;
#pragma unroll
    for (int k = 0; k < 5; k++) w[k][r] = F[k * ps + idx];
    w[5][r] = xq[idx];
    w[6][r] = yq[idx];
  }
  for (int j = sg.x; j <= sg.y; j++) {
    const long long rowj = (long long)j * ld;
    const long long cc = rowj + ic;
    {
      const long long idx = (long long)(j + 2) * ld + ic;
#pragma unroll
      for (int k = 0; k < 5; k++) w[k][4] = F[k * ps + idx];
      w[5][4] = xq[idx];
      w[6][4] = yq[idx];
    }
    const int buf = j & 1;
#pragma unroll
    for (int k = 0; k < 5; k++) sh[buf][k][c] = w[k][2];
    sh[buf][5][c] = xr[cc];
    sh[buf][6][c] = yr[cc];
    if (hslot >= 0) {
      const long long h = rowj + hcol;
#pragma unroll
      for (int k = 0; k < 5; k++) sh[buf][k][hslot] = F[k * ps + h];
      sh[buf][5][hslot] = xr[h];
      sh[buf][6][hslot] = yr[h];
    }
    __syncthreads();
    if (valid) {
      const T(*S)[W] = sh[buf];
      T Fc[6];
#pragma unroll
      for (int k = 0; k < 5; k++) Fc[k] = w[k][2];
      Fc[5] = F[5 * ps + cc];
      T d[6];
      if (a.a_zero) {
#pragma unroll
        for (int l = 0; l < 6; l++) d[l] = (T)0.0;
      } else {
        const T A = (T)a.A;
#pragma unroll
        for (int l = 0; l < 5; l++) d[l] = A * a.DF[l * ps + cc];
        d[5] = (T)0.0;
      }
      // SAT on the block's left/right sides (rows here are not edge rows)
      if (dl == 0) sat_term<T>(d, Fc, a.sd, B.side_off[FDM_SIDE_L] + (j - B.my));
      if (dr == 0) sat_term<T>(d, Fc, a.sd, B.side_off[FDM_SIDE_R] + (j - B.my));
      const T Ji = a.Ji[cc];
      if (dl >= 4 && dr >= 4) {
        // rates_interior_mode2 (rates.f90:407-448)
#define FDM_INTERIOR_FIELD(L)                                                                  \
  {                                                                                            \
    using P = Fld<L>;                                                                          \
    _Pragma("unroll") for (int s = 1; s <= 2; s++) {                                           \
      T ayp = (T)0.5 * (S[6][c + s] + S[6][c]), axp = (T)0.5 * (S[5][c + s] + S[5][c]);        \
      T aym = (T)0.5 * (S[6][c - s] + S[6][c]), axm = (T)0.5 * (S[5][c - s] + S[5][c]);        \
      T byp = (T)0.5 * (w[6][2 + s] + w[6][2]), bxp = (T)0.5 * (w[5][2 + s] + w[5][2]);        \
      T bym = (T)0.5 * (w[6][2 - s] + w[6][2]), bxm = (T)0.5 * (w[5][2 - s] + w[5][2]);        \
      T tt = prod<T, P::ga>(ayp, S[P::ua][c + s], gamma) -                                     \
             prod<T, P::gb>(axp, S[P::ub][c + s], gamma) -                                     \
             prod<T, P::ga>(aym, S[P::ua][c - s], gamma) +                                     \
             prod<T, P::gb>(axm, S[P::ub][c - s], gamma) -                                     \
             prod<T, P::ga>(byp, w[P::ua][2 + s], gamma) +                                     \
             prod<T, P::gb>(bxp, w[P::ub][2 + s], gamma) +                                     \
             prod<T, P::ga>(bym, w[P::ua][2 - s], gamma) -                                     \
             prod<T, P::gb>(bxm, w[P::ub][2 - s], gamma);                                      \
      d[L] = d[L] + tt * cI[L][s - 1] * Ji;                                                    \
    }                                                                                          \
  }
        FDM_INTERIOR_FIELD(0)
        FDM_INTERIOR_FIELD(1)
        FDM_INTERIOR_FIELD(2)
        FDM_INTERIOR_FIELD(3)
        FDM_INTERIOR_FIELD(4)
#undef FDM_INTERIOR_FIELD
      } else {
        // L or R strip (rates.f90:471-591): one-sided q, then central r
#pragma unroll
        for (int s = 0; s < 6; s++) {
          const int k = (dl < 4) ? (B.mx + s - i0) : (B.px - s - i0);
          const T Dq = (dl < 4) ? (T)kDL[dl][s] : (T)(-kDL[dr][s]);
          const T ay = (T)0.5 * (S[6][k] + S[6][c]);
          const T ax = (T)0.5 * (S[5][k] + S[5][c]);
#define FDM_Q1(L)                                                                         \
  {                                                                                       \
    using P = Fld<L>;                                                                     \
    T q = prod<T, P::ga>(ay, S[P::ua][k], gamma) - prod<T, P::gb>(ax, S[P::ub][k], gamma); \
    d[L] = d[L] + bscale<T, P::mat>(q, Dq, rho, G, M) * Ji;                               \
  }
          FDM_Q1(0)
          FDM_Q1(1)
          FDM_Q1(2)
          FDM_Q1(3)
          FDM_Q1(4)
#undef FDM_Q1
        }
#pragma unroll
        for (int s = 1; s <= 2; s++) {
          const T D = (s == 1) ? (T)FDM_DI1 : (T)FDM_DI2;
          const T byp = (T)0.5 * (w[6][2 + s] + w[6][2]), bxp = (T)0.5 * (w[5][2 + s] + w[5][2]);
          const T bym = (T)0.5 * (w[6][2 - s] + w[6][2]), bxm = (T)0.5 * (w[5][2 - s] + w[5][2]);
#define FDM_RC(L)                                                                           \
  {                                                                                         \
    using P = Fld<L>;                                                                       \
    T r = -prod<T, P::ga>(byp, w[P::ua][2 + s], gamma) + prod<T, P::gb>(bxp, w[P::ub][2 + s], gamma) + \
          prod<T, P::ga>(bym, w[P::ua][2 - s], gamma) - prod<T, P::gb>(bxm, w[P::ub][2 - s], gamma);   \
    d[L] = d[L] + bscale<T, P::mat>(r, D, rho, G, M) * Ji;                                  \
  }
          FDM_RC(0)
          FDM_RC(1)
          FDM_RC(2)
          FDM_RC(3)
          FDM_RC(4)
#undef FDM_RC
        }
      }
      finish_point<T>(a, B, bi, i, j, cc, d, Fc);
    }
#pragma unroll
    for (int k = 0; k < 7; k++) {
#pragma unroll
      for (int r = 0; r < 4; r++) w[k][r] = w[k][r + 1];
    }
  }
}

}  // namespace fdm
