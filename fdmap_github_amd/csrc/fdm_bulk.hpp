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
};

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

template <typename T>
__global__ void __launch_bounds__(256) bulk_kernel(BulkArgs<T> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
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
    for (int l = 0; l < 6; l++) d[l] = A * a.DF[l * ps + cc];
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
    for (int l = 0; l < 6; l++) f[l] = F[l * ps + cc] + Bdt * d[l];
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

}  // namespace fdm
