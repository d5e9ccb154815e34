// fdm_bulk.hpp -- the fused bulk kernel: RK scaling, SAT penalties, SBP2/4/6
// curvilinear stencils (rates.f90:384-936), szz overwrite, sources, RK
// update and plasticity, one lane per grid point.
#pragma once

#include <type_traits>

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
  // heterogeneous medium: rho, G, M, gamma, nu per point (null: the block's)
  const T* mrho;
  const T* mG;
  const T* mM;
  const T* mgam;
  const T* mnu;
  // Jacobian J, when a block has artificial dissipation (else null)
  const T* J;
  // PML auxiliary fields: Wx (planes 0-5), Wy (6-11), DWx (12-17), DWy
  // (18-23), plane stride ps (null without a layer)
  T* W;
  // ji_calc: the context's Ji equals 1/(xq yr - xr yq) bit for bit at every
  // point, so the interior tiles may form it from the staged metrics instead
  // of reading the Ji plane (the host then launches their JI variant);
  // skip_df leaves DF1..5 unstored (the last stage of a step that another
  // step of the same call follows: the next stage scales DF by A = 0 and
  // never reads it)
  int ji_calc;
  int skip_df;
  // FP32 deltas of xq, xr, yq, yr from their block's mbase (FP64 contexts
  // whose metrics allow it; read by the interior tiles' MC variant only)
  const float* fxq;
  const float* fxr;
  const float* fyq;
  const float* fyr;
  // inner band points: the band points on no block edge (they need no edge
  // values of the stage), updated by the first nibwg workgroups of the
  // interior launch, one lane per point (nib points; none: nibwg = 0)
  const int2* ibpts;
  int nib, nibwg;
  // receivers: vx, vy of Fin at the nrec points rcc into rout (FP64
  // plane-strain contexts), by nrecwg workgroups ahead of the inner band
  // points in stage 1 of a step; none: nrecwg = 0
  const long long* rcc;
  double* rout;
  int nrec, nrecwg;
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

// Linear tile index of this workgroup in a 1-D grid: with the swizzle, the
// workgroups dispatched to XCD x (ids congruent to x mod 8) take a
// contiguous range of the tile list.
__device__ __forceinline__ int linear_tile(int swz, int off = 0) {
  // off: workgroups ahead of the tiles (the interior launch's inner band)
  const int bid = blockIdx.x - off;
  if (!swz) return bid;
  const int nwg = gridDim.x - off;
  const int q = nwg >> 3, r = nwg & 7, x = bid & 7, k = bid >> 3;
  return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
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
// Acc supplies the point's neighbourhood: f(k, di, dj) = F(k+1) at (i+di,
// j+dj) with one of di, dj zero; xr(di), yr(di) along q; xq(dj), yq(dj)
// along r.  Every kernel evaluates this one expression tree, whatever
// memory the neighbourhood comes from.  The loop runs over s outside the
// five fields: each field still adds its s = 1 term before its s = 2 term,
// as in the reference, while only one set of metric averages is live.
template <typename T, int l, typename Acc, typename Coef>
__device__ __forceinline__ void interior_term(T* d, const Acc& A, int s, T gamma, const Coef& cI,
                                              T Ji, T ayp, T axp, T aym, T axm, T byp, T bxp,
                                              T bym, T bxm) {
  using P = Fld<l>;
  T t = prod<T, P::ga>(ayp, A.f(P::ua, s, 0), gamma) - prod<T, P::gb>(axp, A.f(P::ub, s, 0), gamma) -
        prod<T, P::ga>(aym, A.f(P::ua, -s, 0), gamma) + prod<T, P::gb>(axm, A.f(P::ub, -s, 0), gamma) -
        prod<T, P::ga>(byp, A.f(P::ua, 0, s), gamma) + prod<T, P::gb>(bxp, A.f(P::ub, 0, s), gamma) +
        prod<T, P::ga>(bym, A.f(P::ua, 0, -s), gamma) - prod<T, P::gb>(bxm, A.f(P::ub, 0, -s), gamma);
  d[l] = d[l] + cI.template scale<l>(t, s, Ji);
}

// cI(l, s) returns the interior coefficient of field l and offset s; H is
// the operator's interior half-width (terms of larger s carry a zero
// coefficient in the reference's s = 1..3 loop and add nothing)
template <typename T, typename Acc, typename Coef, int H = 2>
__device__ __forceinline__ void interior_rates(T* d, const Acc& A, T gamma, const Coef& cI, T Ji) {
#pragma unroll
  for (int s = 1; s <= H; s++) {
    const T ayp = (T)0.5 * (A.yr(s) + A.yr(0)), axp = (T)0.5 * (A.xr(s) + A.xr(0));
    const T aym = (T)0.5 * (A.yr(-s) + A.yr(0)), axm = (T)0.5 * (A.xr(-s) + A.xr(0));
    const T byp = (T)0.5 * (A.yq(s) + A.yq(0)), bxp = (T)0.5 * (A.xq(s) + A.xq(0));
    const T bym = (T)0.5 * (A.yq(-s) + A.yq(0)), bxm = (T)0.5 * (A.xq(-s) + A.xq(0));
    interior_term<T, 0>(d, A, s, gamma, cI, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
    interior_term<T, 1>(d, A, s, gamma, cI, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
    interior_term<T, 2>(d, A, s, gamma, cI, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
    interior_term<T, 3>(d, A, s, gamma, cI, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
    interior_term<T, 4>(d, A, s, gamma, cI, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
  }
}

// interior point with artificial dissipation (rates_dissipation_interior_
// mode2, rates.f90:1397-1482; rates_h_dissipation_interior_mode2,
// rates_heterogeneous.f90:1376-1458): d(0) F first, then per s
// (a(s) X + d(s) Y) Ji, where X is the split-form sum and Y the J-weighted
// neighbours of the field itself.  The s = 3 term is zero for SBP4 (the
// heterogeneous loop reads DI(3) outside DI(-2:2); it is taken as zero).
template <typename T, int l, typename Acc, typename Coef>
__device__ __forceinline__ void interior_term_diss(T* d, const Acc& A, int s, T gamma, const Coef& cI,
                                                   T ds, T Ji, T ayp, T axp, T aym, T axm, T byp,
                                                   T bxp, T bym, T bxm) {
  using P = Fld<l>;
  T x = prod<T, P::ga>(ayp, A.f(P::ua, s, 0), gamma) - prod<T, P::gb>(axp, A.f(P::ub, s, 0), gamma) -
        prod<T, P::ga>(aym, A.f(P::ua, -s, 0), gamma) + prod<T, P::gb>(axm, A.f(P::ub, -s, 0), gamma) -
        prod<T, P::ga>(byp, A.f(P::ua, 0, s), gamma) + prod<T, P::gb>(bxp, A.f(P::ub, 0, s), gamma) +
        prod<T, P::ga>(bym, A.f(P::ua, 0, -s), gamma) - prod<T, P::gb>(bxm, A.f(P::ub, 0, -s), gamma);
  T y = A.j(s, 0) * A.f(l, s, 0) + A.j(-s, 0) * A.f(l, -s, 0) + A.j(0, s) * A.f(l, 0, s) +
        A.j(0, -s) * A.f(l, 0, -s);
  d[l] = d[l] + (cI.template coef<l>(s) * x + ds * y) * Ji;
}

template <typename T, typename Acc, typename Coef>
__device__ __forceinline__ void interior_rates_diss(T* d, const Acc& A, T gamma, const Coef& cI,
                                                    const T (&dA)[3], T Ji) {
#pragma unroll
  for (int l = 0; l < 5; l++) d[l] = d[l] + dA[0] * A.f(l, 0, 0);
#pragma unroll
  for (int s = 1; s <= 2; s++) {
    const T ayp = (T)0.5 * (A.yr(s) + A.yr(0)), axp = (T)0.5 * (A.xr(s) + A.xr(0));
    const T aym = (T)0.5 * (A.yr(-s) + A.yr(0)), axm = (T)0.5 * (A.xr(-s) + A.xr(0));
    const T byp = (T)0.5 * (A.yq(s) + A.yq(0)), bxp = (T)0.5 * (A.xq(s) + A.xq(0));
    const T bym = (T)0.5 * (A.yq(-s) + A.yq(0)), bxm = (T)0.5 * (A.xq(-s) + A.xq(0));
    const T ds = dA[s];
    interior_term_diss<T, 0>(d, A, s, gamma, cI, ds, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
    interior_term_diss<T, 1>(d, A, s, gamma, cI, ds, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
    interior_term_diss<T, 2>(d, A, s, gamma, cI, ds, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
    interior_term_diss<T, 3>(d, A, s, gamma, cI, ds, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
    interior_term_diss<T, 4>(d, A, s, gamma, cI, ds, Ji, ayp, axp, aym, axm, byp, bxp, bym, bxm);
  }
}

// interior coefficients of a block, from any address space: t DI(s)/rho Ji
// etc. with DI(s)/rho precomputed (rates.f90:400-405, 411-444); coef() is
// a(s), b(s), c(s) itself, for the dissipative form
template <typename T, typename BlockRef>
struct BlockCoef {
  BlockRef B;
  template <int l>
  __device__ __forceinline__ T scale(T t, int s, T Ji) const {
    return t * (T)B.cI[l][s - 1] * Ji;
  }
  template <int l>
  __device__ __forceinline__ T coef(int s) const {
    return (T)B.cI[l][s - 1];
  }
};

// the material of one point of a heterogeneous medium:
// ")*DI(s)*JacInv/rho", ")*DI(s)*JacInv*M", ")*DI(s)*JacInv*G"
// (rates_heterogeneous.f90:392-428)
template <typename T, int OP = 4>
struct PointCoef {
  T rho, G, M;
  template <int l>
  __device__ __forceinline__ T scale(T t, int s, T Ji) const {
    const T u = t * (T)sbp_di<OP>(s) * Ji;
    if constexpr (l < 2) return u / rho;
    else if constexpr (l == 3) return u * G;
    else return u * M;
  }
  // "DI(s)/rho(i,j)*(", "DI(s)*M(i,j)*(", "DI(s)*G(i,j)*(" of the
  // dissipative form (rates_heterogeneous.f90:1409-1450)
  template <int l>
  __device__ __forceinline__ T coef(int s) const {
    const T D = (T)sbp_di<OP>(s);
    if constexpr (l < 2) return D / rho;
    else if constexpr (l == 3) return D * G;
    else return D * M;
  }
};

// neighbourhood read straight from global memory
template <typename T>
struct GlobalAcc {
  const T* __restrict__ pF;
  const T* __restrict__ pxq;
  const T* __restrict__ pxr;
  const T* __restrict__ pyq;
  const T* __restrict__ pyr;
  long long ps, cc;
  int ld;
  const T* __restrict__ pJ;
  __device__ __forceinline__ T f(int k, int di, int dj) const {
    return pF[k * ps + cc + di + (long long)dj * ld];
  }
  __device__ __forceinline__ T j(int di, int dj) const { return pJ[cc + di + (long long)dj * ld]; }
  __device__ __forceinline__ T xr(int di) const { return pxr[cc + di]; }
  __device__ __forceinline__ T yr(int di) const { return pyr[cc + di]; }
  __device__ __forceinline__ T xq(int dj) const { return pxq[cc + (long long)dj * ld]; }
  __device__ __forceinline__ T yq(int dj) const { return pyq[cc + (long long)dj * ld]; }
};

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
// FULL = false leaves out the singular source, the manufactured-solution
// source and plasticity; the host selects it for launches that need none of
// them.  NTS stores F and DF non-temporally.
template <typename T, bool FULL, bool NTS = false>
__device__ __forceinline__ void finish_point(const BulkArgs<T>& a, const BlockDev& B, const int bi,
                                             const int i, const int j, const long long cc,
                                             T (&d)[6], const T (&Fc)[6]) {
  const long long ps = a.ps;
  // rates_szz (rates.f90:920-936): overwrite
  // (rates_h_szz, rates_heterogeneous.f90:901-917: the point's nu)
  d[5] = (a.mnu ? a.mnu[cc] : (T)B.nu) * (d[2] + d[4]);
  // singular source (source.f90:196-254)
  if constexpr (FULL) for (int p = 0; p < a.nplace; p++) {
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
  if (FULL && B.mms) {
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
    if (FULL && a.do_plastic && B.plastic) {
      // update_fields_plastic (plastic.f90:8-59), full dt, after the last stage
      double s[4], Ds[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        s[k] = (double)f[2 + k];
        Ds[k] = (double)d[2 + k];
      }
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] = s[k] + B.Fb[2 + k];
      }
      long long pc = (long long)j * a.nx + i;
      double lam = a.lambda[pc], gp = a.gammap[pc];
      // a heterogeneous medium gives each point its G and K (plastic.f90:37-43)
      const double Gel = a.mG ? (double)a.mG[cc] : B.G;
      const double Kel = a.mG ? (double)a.mM[cc] - 1.333333333333333 * (double)a.mG[cc] : B.K;
      plastic_point(Gel, Kel, B.mu, B.beta, B.b, B.h, B.eta, B.s0, a.dt, s, Ds, lam, gp);
      a.lambda[pc] = lam;
      a.gammap[pc] = gp;
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] = s[k] - B.Fb[2 + k];
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        f[2 + k] = (T)s[k];
        d[2 + k] = (T)Ds[k];
      }
    }
#pragma unroll
    for (int l = 0; l < 6; l++) {
      if constexpr (NTS) __builtin_nontemporal_store(f[l], &a.Fout[l * ps + cc]);
      else a.Fout[l * ps + cc] = f[l];
    }
  }
#pragma unroll
  for (int l = 0; l < 6; l++) {
    if constexpr (NTS) __builtin_nontemporal_store(d[l], &a.DF[l * ps + cc]);
    else a.DF[l * ps + cc] = d[l];
  }
}

// dissipation terms of rates_dissipation_boundary_mode2 (rates.f90:1485-
// 2021) for fields 1..5 (the szz relation overwrites the sixth), NF = 5; the
// same terms of rates_dissipation_boundary_mode3 (rates.f90:1002-1394) for
// vz, sxz, syz, NF = 3:
// one-sided "(Jac*F)*A*JacInv*Cdiss", the strips' diagonal "F*AI(0)*Cdiss",
// central "(Jac(+)*F(+)+Jac(-)*F(-))*AI(s)*JacInv*Cdiss"
template <typename T, int NF = 5>
__device__ __forceinline__ void diss_one(T (&d)[6], const T* __restrict__ F, const T* __restrict__ J,
                                         long long ps, long long k, T A, T Ji, T Cd) {
  const T Jk = J[k];
#pragma unroll
  for (int l = 0; l < NF; l++) d[l] = d[l] + Jk * F[l * ps + k] * A * Ji * Cd;
}

template <typename T, int NF = 5>
__device__ __forceinline__ void diss_diag(T (&d)[6], const T* __restrict__ F, long long ps,
                                          long long cc, T Cd) {
#pragma unroll
  for (int l = 0; l < NF; l++) d[l] = d[l] + F[l * ps + cc] * (T)FDM_AI0 * Cd;
}

template <typename T, int NF = 5>
__device__ __forceinline__ void diss_central(T (&d)[6], const T* __restrict__ F, const T* __restrict__ J,
                                             long long ps, long long kp, long long km, T A, T Ji,
                                             T Cd) {
  const T Jp = J[kp], Jm = J[km];
#pragma unroll
  for (int l = 0; l < NF; l++) d[l] = d[l] + (Jp * F[l * ps + kp] + Jm * F[l * ps + km]) * A * Ji * Cd;
}

// SAT penalties and one-sided stencils of a point within four points of a
// block edge (time_step.f90:754-810, rates.f90:451-918): L/R strips
// one-sided in q and central in r, B/T strips the converse, corners
// one-sided in both, interleaved per s as in rates_boundary_mode2.  With
// artificial dissipation each one-sided or central term is followed by its
// dissipation term, and strips add the diagonal between their two loops
// (rates_dissipation_boundary_mode2).  The stencil loops are unrolled so
// that all of a point's loads are in flight together: the band is a thin
// set of points, latency-bound.  EXT = false leaves out the heterogeneous
// medium and dissipation code (launches of problems with neither).
template <typename T, bool EXT, int OP = 4>
__device__ __forceinline__ void closure_rates(const BulkArgs<T>& a, const BlockDev& B, const int i,
                                              const int j, const long long cc, T (&d)[6]) {
  const long long ps = a.ps;
  const int ld = a.ld;
  const T* __restrict__ F = a.Fin;
  constexpr int NB = Sbp<OP>::nbst, NS = Sbp<OP>::nbnd, HW = Sbp<OP>::half;
  const int dl = i - B.mx, dr = B.px - i, db = j - B.my, dtp = B.py - j;
  // SAT penalties, sides in the order L, R, B, T
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
  // the point's material in a heterogeneous medium (rates_h_boundary_mode2,
  // rates_heterogeneous.f90:435-898, the same expressions)
  const bool het = EXT && a.mrho != nullptr;
  const T gamma = het ? a.mgam[cc] : (T)B.gamma, rho = het ? a.mrho[cc] : (T)B.rho;
  const T G = het ? a.mG[cc] : (T)B.G, M = het ? a.mM[cc] : (T)B.M;
  // artificial dissipation is built for SBP4 only (fdm_create refuses it
  // with the other operators)
  const bool diss = EXT && OP == 4 && B.Cdiss > 0.0;
  const T Cd = (T)B.Cdiss;
  const T* __restrict__ J = a.J;
  const bool qone = (dl < NB) || (dr < NB);
  const bool rone = (db < NB) || (dtp < NB);
  const long long row = (long long)j * ld;
  const long long colm = (long long)B.my * ld + i, colp = (long long)B.py * ld + i;
  if (qone) {
    // one-sided q: L or R strip, or q part of a corner
#pragma unroll
    for (int s = 0; s < NS; s++) {
      long long kq = (dl < NB) ? row + B.mx + s : row + B.px - s;
      T Dq = (dl < NB) ? (T)sbp_dl<OP>(dl, s) : (T)(-sbp_dl<OP>(dr, s));
      d[0] = d[0] + bscale<T, 0>(q1<T, 0>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(q1<T, 1>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(q1<T, 2>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(q1<T, 3>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(q1<T, 4>(F, yr, xr, ps, kq, cc, gamma), Dq, rho, G, M) * Ji;
      if constexpr (OP == 4) if (diss) diss_one<T>(d, F, J, ps, kq, (T)((dl < 4) ? kAL[dl][s] : kAL[dr][s]), Ji, Cd);
      if (rone) {
        // corner: the r term of the same s follows
        long long kr = (db < NB) ? colm + (long long)s * ld : colp - (long long)s * ld;
        T Dr = (db < NB) ? (T)sbp_dl<OP>(db, s) : (T)(-sbp_dl<OP>(dtp, s));
        d[0] = d[0] + bscale<T, 0>(r1<T, 0>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
        d[1] = d[1] + bscale<T, 0>(r1<T, 1>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
        d[2] = d[2] + bscale<T, 2>(r1<T, 2>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
        d[3] = d[3] + bscale<T, 1>(r1<T, 3>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
        d[4] = d[4] + bscale<T, 2>(r1<T, 4>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
        if constexpr (OP == 4) if (diss) diss_one<T>(d, F, J, ps, kr, (T)((db < 4) ? kAL[db][s] : kAL[dtp][s]), Ji, Cd);
      }
    }
    if (!rone) {
      // L or R strip: central r
      if constexpr (OP == 4) if (diss) diss_diag<T>(d, F, ps, cc, Cd);
#pragma unroll
      for (int s = 1; s <= HW; s++) {
        T D = (T)sbp_di<OP>(s);
        long long kp = cc + (long long)s * ld, km = cc - (long long)s * ld;
        d[0] = d[0] + bscale<T, 0>(rc<T, 0>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
        d[1] = d[1] + bscale<T, 0>(rc<T, 1>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
        d[2] = d[2] + bscale<T, 2>(rc<T, 2>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
        d[3] = d[3] + bscale<T, 1>(rc<T, 3>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
        d[4] = d[4] + bscale<T, 2>(rc<T, 4>(F, yq, xq, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
        if constexpr (OP == 4) if (diss) diss_central<T>(d, F, J, ps, kp, km, (s == 1) ? (T)FDM_AI1 : (T)FDM_AI2, Ji, Cd);
      }
    }
  } else {
    // B or T strip: one-sided r, then central q
#pragma unroll
    for (int s = 0; s < NS; s++) {
      long long kr = (db < NB) ? colm + (long long)s * ld : colp - (long long)s * ld;
      T Dr = (db < NB) ? (T)sbp_dl<OP>(db, s) : (T)(-sbp_dl<OP>(dtp, s));
      d[0] = d[0] + bscale<T, 0>(r1<T, 0>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(r1<T, 1>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(r1<T, 2>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(r1<T, 3>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(r1<T, 4>(F, yq, xq, ps, kr, cc, gamma), Dr, rho, G, M) * Ji;
      if constexpr (OP == 4) if (diss) diss_one<T>(d, F, J, ps, kr, (T)((db < 4) ? kAL[db][s] : kAL[dtp][s]), Ji, Cd);
    }
    if constexpr (OP == 4) if (diss) diss_diag<T>(d, F, ps, cc, Cd);
#pragma unroll
    for (int s = 1; s <= HW; s++) {
      T D = (T)sbp_di<OP>(s);
      long long kp = cc + s, km = cc - s;
      d[0] = d[0] + bscale<T, 0>(qc<T, 0>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[1] = d[1] + bscale<T, 0>(qc<T, 1>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[2] = d[2] + bscale<T, 2>(qc<T, 2>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[3] = d[3] + bscale<T, 1>(qc<T, 3>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      d[4] = d[4] + bscale<T, 2>(qc<T, 4>(F, yr, xr, ps, kp, km, cc, gamma), D, rho, G, M) * Ji;
      if constexpr (OP == 4) if (diss) diss_central<T>(d, F, J, ps, kp, km, (s == 1) ? (T)FDM_AI1 : (T)FDM_AI2, Ji, Cd);
    }
  }
}

// Derivative terms of a PML block's boundary forms (rates.f90:2194-2218
// and the like): v the fields at a one-sided neighbour or a central
// difference, D its coefficient; the auxiliary rate dw takes sigma times
// the same derivative when the layer acts in that direction.
template <typename T>
__device__ __forceinline__ void pml_xterm(T (&d)[6], T (&dw)[6], const T (&v)[6], T D, T hxi, T sx,
                                          bool on, T rho, T G, T M, T gamma) {
  if (on) {
#pragma unroll
    for (int l = 0; l < 6; l++) dw[l] = dw[l] + sx * hxi * v[l] * D;
  }
  d[0] = d[0] + hxi * v[2] * D / rho;
  d[1] = d[1] + hxi * v[3] * D / rho;
  d[2] = d[2] + hxi * v[0] * D * M;
  d[3] = d[3] + hxi * v[1] * D * G;
  d[4] = d[4] + hxi * v[0] * D * M * gamma;
}

template <typename T>
__device__ __forceinline__ void pml_yterm(T (&d)[6], T (&dw)[6], const T (&v)[6], T D, T hyi, T sy,
                                          bool on, T rho, T G, T M, T gamma) {
  if (on) {
#pragma unroll
    for (int l = 0; l < 6; l++) dw[l] = dw[l] + sy * hyi * v[l] * D;
  }
  d[0] = d[0] + hyi * v[3] * D / rho;
  d[1] = d[1] + hyi * v[4] * D / rho;
  d[2] = d[2] + hyi * v[1] * D * M * gamma;
  d[3] = d[3] + hyi * v[0] * D * G;
  d[4] = d[4] + hyi * v[1] * D * M;
}

// A point of a perfectly matched layer block: the SAT penalties, then
// rates_interior_mode2_pml / rates_boundary_mode2_pml (rates.f90:2024-2619):
// the auxiliary fields' coupling and damping (sigma + kappa), and Cartesian
// derivatives with hxi, hyi of the point's region.  Interior points sum the
// x and y differences per s; strips take the one-sided direction's loop and
// the central one's in the reference's order; corners interleave the two
// one-sided loops.  The block's scalar material is used.  The auxiliary
// rates are scaled by A and the fields updated in place (fields.f90:
// 1074-1077, 1181-1184): each lane touches only its own point's W.
// INNER = true compiles the interior form alone (points four or more from
// every edge of their block, where no SAT applies).
template <typename T, bool INNER = false>
__device__ __forceinline__ void pml_rates(const BulkArgs<T>& a, const BlockDev& B, const int i, const int j,
                                          const long long cc, T (&d)[6]) {
  const long long ps = a.ps;
  const int ld = a.ld;
  const T* __restrict__ F = a.Fin;
  const int dl = i - B.mx, dr = B.px - i, db = j - B.my, dtp = B.py - j;
  if (!INNER && (dl == 0 || dr == 0 || db == 0 || dtp == 0)) {
    T Fe[6];
#pragma unroll
    for (int l = 0; l < 6; l++) Fe[l] = F[l * ps + cc];
    if (dl == 0) sat_term<T>(d, Fe, a.sd, B.side_off[FDM_SIDE_L] + db);
    if (dr == 0) sat_term<T>(d, Fe, a.sd, B.side_off[FDM_SIDE_R] + db);
    if (db == 0) sat_term<T>(d, Fe, a.sd, B.side_off[FDM_SIDE_B] + dl);
    if (dtp == 0) sat_term<T>(d, Fe, a.sd, B.side_off[FDM_SIDE_T] + dl);
  }
  const int rx = INNER ? 1 : (dl < 4 ? 0 : (dr < 4 ? 2 : 1));
  const int ry = INNER ? 1 : (db < 4 ? 0 : (dtp < 4 ? 2 : 1));
  const T hxi = (T)B.phx[3 * ry + rx], hyi = (T)B.phy[3 * ry + rx];
  // damping profile (rates.f90:2048-2049, 2063-2066)
  double sxd = B.pml[1], syd = B.pml[1];
  if (B.pml[0] > 0.0) {
    const long long pc = (long long)j * a.nx + i;
    const double x = a.X[pc], y = a.Y[pc];
    sxd = B.pml[1] * pow(fabs(fmax(0.0, fmax(B.pml[3] - x, x - B.pml[4]))), B.pml[0]);
    syd = B.pml[1] * pow(fabs(fmax(0.0, fmax(B.pml[5] - y, y - B.pml[6]))), B.pml[0]);
  }
  const T sx = (T)sxd, sy = (T)syd, kap = (T)B.pml[2];
  const T rho = (T)B.rho, G = (T)B.G, M = (T)B.M, gamma = (T)B.gamma;
  const bool onx = B.pmlx != 0, ony = B.pmly != 0;
  T* __restrict__ W = a.W;
  const T A = (T)a.A;
  T dwx[6], dwy[6];
#pragma unroll
  for (int l = 0; l < 6; l++) {
    // W and its rates stream through once per stage: non-temporal, so that
    // they do not displace the field rows neighbouring points reuse
    dwx[l] = (!onx || a.a_zero) ? (T)0.0 : A * __builtin_nontemporal_load(&W[(12 + l) * ps + cc]);
    dwy[l] = (!ony || a.a_zero) ? (T)0.0 : A * __builtin_nontemporal_load(&W[(18 + l) * ps + cc]);
  }
  T wx[6], wy[6];
  if (onx) {
    T* w = wx;
#pragma unroll
    for (int l = 0; l < 6; l++) w[l] = __builtin_nontemporal_load(&W[l * ps + cc]);
    d[0] = d[0] - w[2] / rho;
    d[1] = d[1] - w[3] / rho;
    d[2] = d[2] - w[0] * M;
    d[3] = d[3] - w[1] * G;
    d[4] = d[4] - w[0] * gamma * M;
#pragma unroll
    for (int l = 0; l < 6; l++) dwx[l] = dwx[l] - (sx + kap) * w[l];
  }
  if (ony) {
    T* w = wy;
#pragma unroll
    for (int l = 0; l < 6; l++) w[l] = __builtin_nontemporal_load(&W[(6 + l) * ps + cc]);
    d[0] = d[0] - w[3] / rho;
    d[1] = d[1] - w[4] / rho;
    d[2] = d[2] - w[1] * gamma * M;
    d[3] = d[3] - w[0] * G;
    d[4] = d[4] - w[1] * M;
#pragma unroll
    for (int l = 0; l < 6; l++) dwy[l] = dwy[l] - (sy + kap) * w[l];
  }
  const long long row = (long long)j * ld;
  const long long colm = (long long)B.my * ld + i, colp = (long long)B.py * ld + i;
  T v[6];
  if (rx == 1 && ry == 1) {
    // interior (rates.f90:2088-2124); the s = 3 terms carry DI(3) = 0
#pragma unroll
    for (int s = 1; s <= 2; s++) {
      const T Dc = (s == 1) ? (T)FDM_DI1 : (T)FDM_DI2;
      const T ca = Dc / rho, cb = Dc * G, cm = Dc * M;
      T dx[6], dy[6];
#pragma unroll
      for (int l = 0; l < 6; l++) {
        dx[l] = F[l * ps + cc + s] - F[l * ps + cc - s];
        dy[l] = F[l * ps + cc + (long long)s * ld] - F[l * ps + cc - (long long)s * ld];
      }
      if (onx) {
#pragma unroll
        for (int l = 0; l < 6; l++) dwx[l] = dwx[l] + sx * hxi * dx[l] * Dc;
      }
      if (ony) {
#pragma unroll
        for (int l = 0; l < 6; l++) dwy[l] = dwy[l] + sy * hyi * dy[l] * Dc;
      }
      d[0] = d[0] + (hxi * dx[2] + hyi * dy[3]) * ca;
      d[1] = d[1] + (hxi * dx[3] + hyi * dy[4]) * ca;
      d[2] = d[2] + (hxi * dx[0] + hyi * dy[1] * gamma) * cm;
      d[3] = d[3] + (hxi * dx[1] + hyi * dy[0]) * cb;
      d[4] = d[4] + (hxi * dx[0] * gamma + hyi * dy[1]) * cm;
    }
  } else if constexpr (INNER) {
  } else if (ry == 1) {
    // L or R strip: one-sided x, then central y
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const long long k = (rx == 0) ? row + B.mx + s : row + B.px - s;
      const T D = (rx == 0) ? (T)kDL[dl][s] : (T)(-kDL[dr][s]);
#pragma unroll
      for (int l = 0; l < 6; l++) v[l] = F[l * ps + k];
      pml_xterm<T>(d, dwx, v, D, hxi, sx, onx, rho, G, M, gamma);
    }
#pragma unroll
    for (int s = 1; s <= 2; s++) {
#pragma unroll
      for (int l = 0; l < 6; l++)
        v[l] = F[l * ps + cc + (long long)s * ld] - F[l * ps + cc - (long long)s * ld];
      pml_yterm<T>(d, dwy, v, (s == 1) ? (T)FDM_DI1 : (T)FDM_DI2, hyi, sy, ony, rho, G, M, gamma);
    }
  } else if (rx == 1) {
    // B or T strip: central x, then one-sided y
#pragma unroll
    for (int s = 1; s <= 2; s++) {
#pragma unroll
      for (int l = 0; l < 6; l++) v[l] = F[l * ps + cc + s] - F[l * ps + cc - s];
      pml_xterm<T>(d, dwx, v, (s == 1) ? (T)FDM_DI1 : (T)FDM_DI2, hxi, sx, onx, rho, G, M, gamma);
    }
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const long long k = (ry == 0) ? colm + (long long)s * ld : colp - (long long)s * ld;
      const T D = (ry == 0) ? (T)kDL[db][s] : (T)(-kDL[dtp][s]);
#pragma unroll
      for (int l = 0; l < 6; l++) v[l] = F[l * ps + k];
      pml_yterm<T>(d, dwy, v, D, hyi, sy, ony, rho, G, M, gamma);
    }
  } else {
    // corner: the two one-sided loops interleaved per s
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const long long kx = (rx == 0) ? row + B.mx + s : row + B.px - s;
      const T Dx = (rx == 0) ? (T)kDL[dl][s] : (T)(-kDL[dr][s]);
#pragma unroll
      for (int l = 0; l < 6; l++) v[l] = F[l * ps + kx];
      pml_xterm<T>(d, dwx, v, Dx, hxi, sx, onx, rho, G, M, gamma);
      const long long ky = (ry == 0) ? colm + (long long)s * ld : colp - (long long)s * ld;
      const T Dy = (ry == 0) ? (T)kDL[db][s] : (T)(-kDL[dtp][s]);
#pragma unroll
      for (int l = 0; l < 6; l++) v[l] = F[l * ps + ky];
      pml_yterm<T>(d, dwy, v, Dy, hyi, sy, ony, rho, G, M, gamma);
    }
  }
  const T Bdt = (T)a.Bdt;
  // W + Bdt*DW from the values loaded above (this lane's own point)
  if (onx) {
#pragma unroll
    for (int l = 0; l < 6; l++) {
      if (a.do_update) __builtin_nontemporal_store(wx[l] + Bdt * dwx[l], &W[l * ps + cc]);
      __builtin_nontemporal_store(dwx[l], &W[(12 + l) * ps + cc]);
    }
  }
  if (ony) {
#pragma unroll
    for (int l = 0; l < 6; l++) {
      if (a.do_update) __builtin_nontemporal_store(wy[l] + Bdt * dwy[l], &W[(6 + l) * ps + cc]);
      __builtin_nontemporal_store(dwy[l], &W[(18 + l) * ps + cc]);
    }
  }
}

// d = A * DF at the start of a stage (scale_rates, time_step.f90:138-141)
template <typename T>
__device__ __forceinline__ void scaled_rates(const BulkArgs<T>& a, const long long cc, T (&d)[6]) {
  if (a.a_zero) {
#pragma unroll
    for (int l = 0; l < 6; l++) d[l] = (T)0.0;
  } else {
    const T A = (T)a.A;
#pragma unroll
    for (int l = 0; l < 5; l++) d[l] = A * a.DF[l * a.ps + cc];
    d[5] = (T)0.0;  // DF6 is overwritten by the szz relation
  }
}

// Everything that happens at one grid point in one stage, reading its
// neighbours straight from global memory: the per-point kernel, and the
// closure points of the tile and marching kernels.  EXT as in closure_rates,
// and also carries the perfectly matched layer.
template <typename T, bool FULL, bool EXT, int OP = 4>
__device__ __forceinline__ void point_generic(const BulkArgs<T>& a, const int i, const int j) {
  const int bi = a.blk[a.rowb[j] * a.nbx + a.colb[i]];
  const BlockDev& B = a.blocks[bi];
  const long long ps = a.ps;
  const int ld = a.ld;
  const long long cc = (long long)j * ld + i;
  constexpr int NB = Sbp<OP>::nbst, HW = Sbp<OP>::half;
  T d[6];
  scaled_rates<T>(a, cc, d);
  bool done = false;
  if constexpr (OP == 4) {
    // the perfectly matched layer and artificial dissipation are built for
    // SBP4 only (fdm_create refuses them with the other operators)
    if (EXT && (B.pmlx | B.pmly)) {
      pml_rates<T>(a, B, i, j, cc, d);
      done = true;
    } else if (EXT && B.Cdiss > 0.0 && i - B.mx >= 4 && B.px - i >= 4 && j - B.my >= 4 &&
               B.py - j >= 4) {
      const GlobalAcc<T> acc{a.Fin, a.xq, a.xr, a.yq, a.yr, ps, cc, ld, a.J};
      const T dA[3] = {(T)B.dA[0], (T)B.dA[1], (T)B.dA[2]};
      if (a.mrho)
        interior_rates_diss<T>(d, acc, a.mgam[cc], PointCoef<T>{a.mrho[cc], a.mG[cc], a.mM[cc]}, dA,
                               a.Ji[cc]);
      else
        interior_rates_diss<T>(d, acc, (T)B.gamma, BlockCoef<T, const BlockDev&>{B}, dA, a.Ji[cc]);
      done = true;
    }
  }
  if (done) {
  } else if (i - B.mx >= NB && B.px - i >= NB && j - B.my >= NB && B.py - j >= NB) {
    const GlobalAcc<T> acc{a.Fin, a.xq, a.xr, a.yq, a.yr, ps, cc, ld, a.J};
    if (EXT && a.mrho) {
      interior_rates<T, GlobalAcc<T>, PointCoef<T, OP>, HW>(
          d, acc, a.mgam[cc], PointCoef<T, OP>{a.mrho[cc], a.mG[cc], a.mM[cc]}, a.Ji[cc]);
    } else {
      interior_rates<T, GlobalAcc<T>, BlockCoef<T, const BlockDev&>, HW>(
          d, acc, (T)B.gamma, BlockCoef<T, const BlockDev&>{B}, a.Ji[cc]);
    }
  } else {
    closure_rates<T, EXT, OP>(a, B, i, j, cc, d);
  }
  T Fc[6];
#pragma unroll
  for (int l = 0; l < 6; l++) Fc[l] = a.Fin[l * ps + cc];
  finish_point<T, FULL>(a, B, bi, i, j, cc, d, Fc);
}

// per-point kernel: one lane per grid point, 64 x blockDim.y tiles
template <typename T, int OP = 4>
__global__ void __launch_bounds__(256) bulk_kernel(BulkArgs<T> a) {
  int tx, ty;
  tile_of(a.xcd_swizzle, tx, ty);
  const int i = tx * blockDim.x + threadIdx.x;
  const int j = ty * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  point_generic<T, true, true, OP>(a, i, j);
}

// global load at a 32-bit byte offset from a wave-uniform base, so that the
// address is formed by the load itself (scalar base + vector offset)
template <typename T>
__device__ __forceinline__ T ld_off(const T* base, unsigned boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}

template <typename T>
__device__ __forceinline__ void st_off(T* base, unsigned boff, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}

// NT: a non-temporal store, for planes that are streamed once per launch
// when the working set exceeds the last-level cache (no allocation there)
// a load at a byte offset, non-temporal when NT
template <bool NT, typename T>
__device__ __forceinline__ T ld_sel(const T* base, unsigned boff) {
  const T* p = reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// a metric value from its FP32 delta: base + d, exact by construction; a
// NaN delta (a value the encoding cannot hold) reads the FP64 plane.  That
// rare load completes inside its branch (s_waitcnt vmcnt(0)): left in
// flight, it would make the compiler wait after the branch, in every wave,
// for all of the tile's loads (the lane's DF and F6 among them).
__device__ __forceinline__ double metric_of(float d, double base, const double* plane, unsigned boff) {
  if (d == d) return base + (double)d;
  const double v = ld_off(plane, boff);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt unchanged
  return v;
}

template <bool NT, typename T>
__device__ __forceinline__ void st_sel(T* base, unsigned boff, T v) {
  T* p = reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff);
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// neighbourhood of the interior kernel: everything from LDS.  sF: F1..F5
// with a halo of 2 in q and r; sR: xr, yr with a halo in q; sQ: xq, yq with
// a halo in r.  (x, y): the lane's point in tile coordinates.
template <typename T, int TX, int TY, int H = 2>
struct TileAcc {
  const T (*sF)[TY + 2 * H][TX + 2 * H];
  const T (*sR)[TY][TX + 2 * H];
  const T (*sQ)[TY + 2 * H][TX];
  int x, y;
  const T* sJ;  // [TY + 2H][TX + 2H], with artificial dissipation only
  __device__ __forceinline__ T f(int k, int di, int dj) const { return sF[k][y + H + dj][x + H + di]; }
  __device__ __forceinline__ T j(int di, int dj) const {
    return sJ[(y + H + dj) * (TX + 2 * H) + x + H + di];
  }
  __device__ __forceinline__ T xr(int di) const { return sR[0][y][x + H + di]; }
  __device__ __forceinline__ T yr(int di) const { return sR[1][y][x + H + di]; }
  __device__ __forceinline__ T xq(int dj) const { return sQ[0][y + H + dj][x]; }
  __device__ __forceinline__ T yq(int dj) const { return sQ[1][y + H + dj][x]; }
};

typedef const __attribute__((address_space(4))) BlockDev ConstBlockDev;

// Interior kernel.  The grid walks a list of TX x TY tiles that cover each
// block's interior rectangle [mx+4, px-4] x [my+4, py-4] from its corner, so
// every lane of a tile that falls inside the rectangle is an interior point
// of the tile's block: no one-sided closure, no SAT, and the block index is
// wave-uniform (its parameters are read as scalars through the constant
// address space).  All of the tile's loads are issued before one barrier:
// F1..F5 with a halo of two points in q and r, xr/yr with the q halo and
// xq/yq with the r halo go to LDS; the lane's own DF, F6 and Ji stay in
// registers.  Halo indices are clamped into the grid; a clamped value never
// reaches an active lane.  tiles[k] = (block, i0, j0, 0).
// JI: form Ji from the staged metrics (a context with ji_calc set) rather
// than read it; a compile-time switch, so that neither path holds the other's
// registers.  MC: read xq, xr, yq, yr as FP32 deltas from their block's base
// (FP64 only; the values staged in LDS are the FP64 metrics to the bit).
template <typename T, int TX, int TY, bool FULL, bool NTS, bool HET, bool DISS, int OP = 4,
          bool JI = false, bool MC = false>
__device__ __forceinline__ void interior_tile(const BulkArgs<T>& a, const int4 td) {
  // H: the operator's interior half-width, the tile's halo; NB: the
  // closure's depth, how far the interior rectangle stays from block edges
  constexpr int H = Sbp<OP>::half, NB = Sbp<OP>::nbst;
  static_assert(!DISS || OP == 4, "artificial dissipation is built for SBP4 only");
  static_assert(!MC || std::is_same<T, double>::value, "FP32-delta metrics serve FP64 fields");
  constexpr int HX = TX + 2 * H, HY = TY + 2 * H, NT = TX * TY;
  constexpr int NF = (HY * HX + NT - 1) / NT;  // staging passes, F1..F5
  constexpr int NR = (TY * HX + NT - 1) / NT;  // xr, yr
  constexpr int NQ = (HY * TX + NT - 1) / NT;  // xq, yq
  __shared__ T sF[5][HY][HX];
  __shared__ T sR[2][TY][HX];
  __shared__ T sQ[2][HY][TX];
  // the Jacobian with the same halo as F, for artificial dissipation
  __shared__ T sJ[DISS ? HY : 1][DISS ? HX : 1];
  const int bu = td.x, i0 = td.y, j0 = td.z;
  ConstBlockDev* B4 = (ConstBlockDev*)a.blocks + bu;
  const int tid = threadIdx.y * TX + threadIdx.x;
  const int ld = a.ld, nx1 = a.nx - 1, ny1 = a.ny - 1;
  const long long ps = a.ps;
  const T* __restrict__ F = a.Fin;
  T vF[NF][5], vR[MC ? 1 : NR][2], vQ[MC ? 1 : NQ][2], vJ[DISS ? NF : 1];
  float fR[MC ? NR : 1][2], fQ[MC ? NQ : 1][2];
  // Every lane loads in every pass: a lane past the end of the staged
  // block (in the last pass) reloads its last element, and only the LDS
  // stores below are conditional.  A load under a lane condition makes the
  // compiler copy its result out of the branch, which waits for every load
  // issued before it.
#pragma unroll
  for (int p = 0; p < NF; p++) {
    const int k = min(tid + p * NT, HY * HX - 1);
    const int r = k / HX, c = k - r * HX;
    const int gj = min(max(j0 - H + r, 0), ny1), gi = min(max(i0 - H + c, 0), nx1);
    const unsigned o = (unsigned)(gj * ld + gi) * (unsigned)sizeof(T);
#pragma unroll
    for (int l = 0; l < 5; l++) vF[p][l] = ld_off(F + l * ps, o);
    if constexpr (DISS) vJ[p] = ld_off(a.J, o);
  }
#pragma unroll
  for (int p = 0; p < NR; p++) {
    const int k = min(tid + p * NT, TY * HX - 1);
    const int r = k / HX, c = k - r * HX;
    const int gj = min(j0 + r, ny1), gi = min(max(i0 - H + c, 0), nx1);
    if constexpr (MC) {
      const unsigned o4 = (unsigned)(gj * ld + gi) * 4u;
      fR[p][0] = ld_off(a.fxr, o4);
      fR[p][1] = ld_off(a.fyr, o4);
    } else {
      const unsigned o = (unsigned)(gj * ld + gi) * (unsigned)sizeof(T);
      vR[p][0] = ld_off(a.xr, o);
      vR[p][1] = ld_off(a.yr, o);
    }
  }
#pragma unroll
  for (int p = 0; p < NQ; p++) {
    const int k = min(tid + p * NT, HY * TX - 1);
    const int r = k / TX, c = k - r * TX;
    const int gj = min(max(j0 - H + r, 0), ny1), gi = min(i0 + c, nx1);
    if constexpr (MC) {
      const unsigned o4 = (unsigned)(gj * ld + gi) * 4u;
      fQ[p][0] = ld_off(a.fxq, o4);
      fQ[p][1] = ld_off(a.fyq, o4);
    } else {
      const unsigned o = (unsigned)(gj * ld + gi) * (unsigned)sizeof(T);
      vQ[p][0] = ld_off(a.xq, o);
      vQ[p][1] = ld_off(a.yq, o);
    }
  }
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const bool valid = i <= B4->px - NB && j <= B4->py - NB;
  const unsigned oc = (unsigned)(min(j, ny1) * ld + min(i, nx1)) * (unsigned)sizeof(T);
  // the lane's own DF1..5 and F6 are read once per launch: with the
  // non-temporal stores (planes beyond the last-level cache) they are read
  // non-temporally too
  T dfa[5];
  if (!a.a_zero) {
#pragma unroll
    for (int l = 0; l < 5; l++) dfa[l] = ld_sel<NTS>(a.DF + l * ps, oc);
  }
  const T f6 = ld_sel<NTS>(F + 5 * ps, oc);
  T Ji = (T)0.0;
  if constexpr (!JI) Ji = ld_off(a.Ji, oc);
  // a heterogeneous medium: the lane's own material (rho, G, M, gamma, nu)
  T mat[5];
  if constexpr (HET) {
    mat[0] = ld_off(a.mrho, oc);
    mat[1] = ld_off(a.mG, oc);
    mat[2] = ld_off(a.mM, oc);
    mat[3] = ld_off(a.mgam, oc);
    mat[4] = ld_off(a.mnu, oc);
  }
#pragma unroll
  for (int p = 0; p < NF; p++) {
    const int k = tid + p * NT;
    if (k < HY * HX) {
      const int r = k / HX, c = k - r * HX;
#pragma unroll
      for (int l = 0; l < 5; l++) sF[l][r][c] = vF[p][l];
      if constexpr (DISS) sJ[r][c] = vJ[p];
    }
  }
#pragma unroll
  for (int p = 0; p < NR; p++) {
    const int k = tid + p * NT;
    if (k < TY * HX) {
      const int r = k / HX, c = k - r * HX;
      if constexpr (MC) {
        const int gj = min(j0 + r, ny1), gi = min(max(i0 - H + c, 0), nx1);
        const unsigned o = (unsigned)(gj * ld + gi) * 8u;
        sR[0][r][c] = metric_of(fR[p][0], B4->mbase[1], a.xr, o);
        sR[1][r][c] = metric_of(fR[p][1], B4->mbase[3], a.yr, o);
      } else {
        sR[0][r][c] = vR[p][0];
        sR[1][r][c] = vR[p][1];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NQ; p++) {
    const int k = tid + p * NT;
    if (k < HY * TX) {
      const int r = k / TX, c = k - r * TX;
      if constexpr (MC) {
        const int gj = min(max(j0 - H + r, 0), ny1), gi = min(i0 + c, nx1);
        const unsigned o = (unsigned)(gj * ld + gi) * 8u;
        sQ[0][r][c] = metric_of(fQ[p][0], B4->mbase[0], a.xq, o);
        sQ[1][r][c] = metric_of(fQ[p][1], B4->mbase[2], a.yq, o);
      } else {
        sQ[0][r][c] = vQ[p][0];
        sQ[1][r][c] = vQ[p][1];
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  if constexpr (JI) {
    // grid.f90:531-534: J = xq yr - xr yq, Ji = 1/J, from the lane's own
    // metrics (no contraction: the host's rounding)
    const T J = sQ[0][threadIdx.y + H][threadIdx.x] * sR[1][threadIdx.y][threadIdx.x + H] -
                sR[0][threadIdx.y][threadIdx.x + H] * sQ[1][threadIdx.y + H][threadIdx.x];
    Ji = (T)1.0 / J;
  }
  T d[6];
  if (a.a_zero) {
#pragma unroll
    for (int l = 0; l < 6; l++) d[l] = (T)0.0;
  } else {
    const T A = (T)a.A;
#pragma unroll
    for (int l = 0; l < 5; l++) d[l] = A * dfa[l];
    d[5] = (T)0.0;  // DF6 is overwritten by the szz relation
  }
  const TileAcc<T, TX, TY, H> acc{sF, sR, sQ, (int)threadIdx.x, (int)threadIdx.y, &sJ[0][0]};
  bool diss_done = false;
  if constexpr (DISS) {
    // a block without dissipation keeps the plain expression
    if (B4->Cdiss > 0.0) {
      const T dA[3] = {(T)B4->dA[0], (T)B4->dA[1], (T)B4->dA[2]};
      if constexpr (HET)
        interior_rates_diss<T>(d, acc, mat[3], PointCoef<T>{mat[0], mat[1], mat[2]}, dA, Ji);
      else
        interior_rates_diss<T>(d, acc, (T)B4->gamma, BlockCoef<T, ConstBlockDev&>{*B4}, dA, Ji);
      diss_done = true;
    }
  }
  if (!diss_done) {
    if constexpr (HET)
      interior_rates<T, TileAcc<T, TX, TY, H>, PointCoef<T, OP>, H>(
          d, acc, mat[3], PointCoef<T, OP>{mat[0], mat[1], mat[2]}, Ji);
    else
      interior_rates<T, TileAcc<T, TX, TY, H>, BlockCoef<T, ConstBlockDev&>, H>(
          d, acc, (T)B4->gamma, BlockCoef<T, ConstBlockDev&>{*B4}, Ji);
  }
  T Fc[6];
#pragma unroll
  for (int k = 0; k < 5; k++) Fc[k] = acc.f(k, 0, 0);
  Fc[5] = f6;
  if constexpr (FULL) {
    finish_point<T, true>(a, a.blocks[bu], bu, i, j, (long long)j * ld + i, d, Fc);
  } else {
    // rates_szz (rates.f90:920-936), update_fields_interior (fields.f90:1225-1242).
    // DF6 is not stored: no stage reads it (each overwrites it with this
    // relation), and fdm_get_fields recomputes it from DF3 and DF5 with the
    // same expression.
    T nu;
    if constexpr (HET) nu = mat[4];
    else nu = (T)B4->nu;
    d[5] = nu * (d[2] + d[4]);
    if (a.do_update) {
      const T Bdt = (T)a.Bdt;
#pragma unroll
      for (int l = 0; l < 6; l++) st_sel<NTS>(a.Fout + l * ps, oc, Fc[l] + Bdt * d[l]);
    }
    if (!a.skip_df) {
#pragma unroll
      for (int l = 0; l < 5; l++) st_sel<NTS>(a.DF + l * ps, oc, d[l]);
    }
  }
}

// Field-lane forms of the coupling terms above: the field's F(UA), F(UB),
// gamma flags and material factor are per-lane values (fld_* of lane field
// l), and each product rounds as the templated form does.
struct FldRt {
  int ua, ub, mat;
  bool ga, gb;
};

__device__ __forceinline__ FldRt fld_rt(int l) {
  FldRt f;
  f.ua = l == 0 ? 2 : (l == 1 ? 3 : (l == 3 ? 1 : 0));
  f.ub = l == 0 ? 3 : (l == 1 ? 4 : (l == 3 ? 0 : 1));
  f.mat = (l <= 1) ? 0 : (l == 3 ? 1 : 2);
  f.ga = l == 4;
  f.gb = l == 2;
  return f;
}

template <typename T>
__device__ __forceinline__ T prod_rt(T m, T f, T gamma, bool g) {
  const T mf = m * f;
  return g ? mf * gamma : mf;
}

template <typename T>
__device__ __forceinline__ T bscale_rt(T t, T D, T rho, T G, T M, int mat) {
  const T td = t * D;
  return mat == 0 ? td / rho : (mat == 1 ? td * G : td * M);
}

// Band kernel with one lane per (point, field) for plane strain in a
// uniform medium without dissipation, lean stages (no sources, MMS or
// plastic update).  Twelve points share a wave, five lanes each (lanes 60-63
// idle); the lane of field l evaluates DF(l) alone, with the terms of
// closure_rates in their order, so that it holds a fifth of the stencil and
// its loads are in flight together.  The field-2 lane also forms szz, DF(6)
// = nu (DF(3) + DF(5)), from the field-4 lane two lanes up.  Every lane runs
// to the shuffle (idle lanes on a clamped point) and only active ones store.
template <typename T, int OP = 4, bool EXT = false>
__device__ __forceinline__ void band_fl_wave(const BulkArgs<T>& a, const int2* __restrict__ pts, int npts,
                                             int wave, int lane) {
  constexpr int NB = Sbp<OP>::nbst, NS = Sbp<OP>::nbnd, HW = Sbp<OP>::half;
  const int grp = lane / 5, l = min(lane - 5 * grp, 4);
  const int k = wave * 12 + grp;
  const bool active = lane < 60 && k < npts;
  const int2 p = pts[min(k, npts - 1)];
  const int i = p.x, j = p.y;
  const int bi = a.blk[a.rowb[j] * a.nbx + a.colb[i]];
  const BlockDev& B = a.blocks[bi];
  const long long ps = a.ps;
  const int ld = a.ld;
  const long long cc = (long long)j * ld + i;
  const T* __restrict__ F = a.Fin;
  const FldRt fr = fld_rt(l);
  T d = a.a_zero ? (T)0.0 : (T)a.A * a.DF[l * ps + cc];
  const int dl = i - B.mx, dr = B.px - i, db = j - B.my, dtp = B.py - j;
  if (dl == 0 || dr == 0 || db == 0 || dtp == 0) {
    // SAT penalties (all six fields, this lane's kept), sides L, R, B, T
    T Fe[6], dd[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
      Fe[m] = F[m * ps + cc];
      dd[m] = d;
    }
    if (dl == 0) sat_term<T>(dd, Fe, a.sd, B.side_off[FDM_SIDE_L] + db);
    if (dr == 0) sat_term<T>(dd, Fe, a.sd, B.side_off[FDM_SIDE_R] + db);
    if (db == 0) sat_term<T>(dd, Fe, a.sd, B.side_off[FDM_SIDE_B] + dl);
    if (dtp == 0) sat_term<T>(dd, Fe, a.sd, B.side_off[FDM_SIDE_T] + dl);
    d = l == 0 ? dd[0] : (l == 1 ? dd[1] : (l == 2 ? dd[2] : (l == 3 ? dd[3] : dd[4])));
  }
  const T* __restrict__ xq = a.xq;
  const T* __restrict__ xr = a.xr;
  const T* __restrict__ yq = a.yq;
  const T* __restrict__ yr = a.yr;
  const T Ji = a.Ji[cc];
  // EXT: a heterogeneous medium (the point's material) and artificial
  // dissipation (SBP4 only), as in closure_rates
  const bool het = EXT && a.mrho != nullptr;
  const T gamma = het ? a.mgam[cc] : (T)B.gamma, rho = het ? a.mrho[cc] : (T)B.rho;
  const T G = het ? a.mG[cc] : (T)B.G, M = het ? a.mM[cc] : (T)B.M;
  const bool diss = EXT && OP == 4 && B.Cdiss > 0.0;
  const T Cd = (T)B.Cdiss;
  const T* __restrict__ Jp = a.J;
  const T* __restrict__ Fl = F + l * ps;
  const T* __restrict__ Fa = F + fr.ua * ps;
  const T* __restrict__ Fb = F + fr.ub * ps;
  const bool qone = (dl < NB) || (dr < NB);
  const bool rone = (db < NB) || (dtp < NB);
  const long long row = (long long)j * ld;
  const long long colm = (long long)B.my * ld + i, colp = (long long)B.py * ld + i;
  // q1: ay F(UA) - ax F(UB); r1: -by F(UA) + bx F(UB) (closure_rates)
  if (qone) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const long long kq = (dl < NB) ? row + B.mx + s : row + B.px - s;
      const T Dq = (dl < NB) ? (T)sbp_dl<OP>(dl, s) : (T)(-sbp_dl<OP>(dr, s));
      const T ay = (T)0.5 * (yr[kq] + yr[cc]), ax = (T)0.5 * (xr[kq] + xr[cc]);
      const T tq = prod_rt(ay, Fa[kq], gamma, fr.ga) - prod_rt(ax, Fb[kq], gamma, fr.gb);
      d = d + bscale_rt(tq, Dq, rho, G, M, fr.mat) * Ji;
      if constexpr (EXT && OP == 4)
        if (diss) d = d + Jp[kq] * Fl[kq] * (T)((dl < 4) ? kAL[dl][s] : kAL[dr][s]) * Ji * Cd;
      if (rone) {
        const long long kr = (db < NB) ? colm + (long long)s * ld : colp - (long long)s * ld;
        const T Dr = (db < NB) ? (T)sbp_dl<OP>(db, s) : (T)(-sbp_dl<OP>(dtp, s));
        const T by = (T)0.5 * (yq[kr] + yq[cc]), bx = (T)0.5 * (xq[kr] + xq[cc]);
        const T tr = -prod_rt(by, Fa[kr], gamma, fr.ga) + prod_rt(bx, Fb[kr], gamma, fr.gb);
        d = d + bscale_rt(tr, Dr, rho, G, M, fr.mat) * Ji;
        if constexpr (EXT && OP == 4)
          if (diss) d = d + Jp[kr] * Fl[kr] * (T)((db < 4) ? kAL[db][s] : kAL[dtp][s]) * Ji * Cd;
      }
    }
    if (!rone) {
      // L or R strip: central r (rc)
      if constexpr (EXT && OP == 4)
        if (diss) d = d + Fl[cc] * (T)FDM_AI0 * Cd;
#pragma unroll
      for (int s = 1; s <= HW; s++) {
        const T D = (T)sbp_di<OP>(s);
        const long long kp = cc + (long long)s * ld, km = cc - (long long)s * ld;
        const T byp = (T)0.5 * (yq[kp] + yq[cc]), bxp = (T)0.5 * (xq[kp] + xq[cc]);
        const T bym = (T)0.5 * (yq[km] + yq[cc]), bxm = (T)0.5 * (xq[km] + xq[cc]);
        const T t = -prod_rt(byp, Fa[kp], gamma, fr.ga) + prod_rt(bxp, Fb[kp], gamma, fr.gb) +
                    prod_rt(bym, Fa[km], gamma, fr.ga) - prod_rt(bxm, Fb[km], gamma, fr.gb);
        d = d + bscale_rt(t, D, rho, G, M, fr.mat) * Ji;
        if constexpr (EXT && OP == 4)
          if (diss) d = d + (Jp[kp] * Fl[kp] + Jp[km] * Fl[km]) * ((s == 1) ? (T)FDM_AI1 : (T)FDM_AI2) * Ji * Cd;
      }
    }
  } else {
    // B or T strip: one-sided r, then central q (qc)
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const long long kr = (db < NB) ? colm + (long long)s * ld : colp - (long long)s * ld;
      const T Dr = (db < NB) ? (T)sbp_dl<OP>(db, s) : (T)(-sbp_dl<OP>(dtp, s));
      const T by = (T)0.5 * (yq[kr] + yq[cc]), bx = (T)0.5 * (xq[kr] + xq[cc]);
      const T tr = -prod_rt(by, Fa[kr], gamma, fr.ga) + prod_rt(bx, Fb[kr], gamma, fr.gb);
      d = d + bscale_rt(tr, Dr, rho, G, M, fr.mat) * Ji;
      if constexpr (EXT && OP == 4)
        if (diss) d = d + Jp[kr] * Fl[kr] * (T)((db < 4) ? kAL[db][s] : kAL[dtp][s]) * Ji * Cd;
    }
    if constexpr (EXT && OP == 4)
      if (diss) d = d + Fl[cc] * (T)FDM_AI0 * Cd;
#pragma unroll
    for (int s = 1; s <= HW; s++) {
      const T D = (T)sbp_di<OP>(s);
      const long long kp = cc + s, km = cc - s;
      const T ayp = (T)0.5 * (yr[kp] + yr[cc]), axp = (T)0.5 * (xr[kp] + xr[cc]);
      const T aym = (T)0.5 * (yr[km] + yr[cc]), axm = (T)0.5 * (xr[km] + xr[cc]);
      const T t = prod_rt(ayp, Fa[kp], gamma, fr.ga) - prod_rt(axp, Fb[kp], gamma, fr.gb) -
                  prod_rt(aym, Fa[km], gamma, fr.ga) + prod_rt(axm, Fb[km], gamma, fr.gb);
      d = d + bscale_rt(t, D, rho, G, M, fr.mat) * Ji;
      if constexpr (EXT && OP == 4)
        if (diss) d = d + (Jp[kp] * Fl[kp] + Jp[km] * Fl[km]) * ((s == 1) ? (T)FDM_AI1 : (T)FDM_AI2) * Ji * Cd;
    }
  }
  // finish_point (lean): szz, the field update, the rates
  const T d4 = __shfl(d, (lane + 2) & 63, 64);
  if (!active) return;
  const T Bdt = (T)a.Bdt;
  if (a.do_update) a.Fout[l * ps + cc] = F[l * ps + cc] + Bdt * d;
  a.DF[l * ps + cc] = d;
  if (l == 2) {
    const T d5 = (a.mnu ? a.mnu[cc] : (T)B.nu) * (d + d4);
    if (a.do_update) a.Fout[5 * ps + cc] = F[5 * ps + cc] + Bdt * d5;
    a.DF[5 * ps + cc] = d5;
  }
}

template <typename T, int OP = 4, bool EXT = false>
__global__ void __launch_bounds__(64) bulk_band_fl(BulkArgs<T> a, const int2* __restrict__ pts, int npts) {
  band_fl_wave<T, OP, EXT>(a, pts, npts, blockIdx.x, threadIdx.x);
}

// a heterogeneous medium holds five more values per lane, dissipation the
// Jacobian's staging: three waves per SIMD keep them in registers
template <typename T, int TX, int TY, bool FULL, bool NTS, bool HET, bool DISS, int OP = 4,
          bool JI = false, bool MC = false>
__global__ void __launch_bounds__(TX * TY)
    __attribute__((amdgpu_waves_per_eu(FULL ? 2 : ((HET || DISS || OP == 6) ? 3 : 4))))
bulk_interior(BulkArgs<T> a, const int4* __restrict__ tiles) {
  if constexpr (!HET && OP == 4) {
    // The first nibwg workgroups update the inner band points, one lane per
    // point.  They need only what the tiles need (the previous stage's band
    // and interior), so they ride in this launch rather than on the second
    // stream after the edge values; dispatched first, their closures (bound
    // by load latency) overlap the tiles.
    if constexpr (std::is_same<T, double>::value) {
      if ((int)blockIdx.x < a.nrecwg) {
        // the receivers of the previous step's fields (this stage's input,
        // which the launch does not write)
        const int r = (int)blockIdx.x * (TX * TY) + (int)(threadIdx.y * TX + threadIdx.x);
        if (r < a.nrec) {
          const long long q = a.rcc[r];
          a.rout[2 * r] = a.Fin[q];
          a.rout[2 * r + 1] = a.Fin[a.ps + q];
        }
        return;
      }
    }
    // receivers ride in FP64 launches only; the FP32 kernel keeps its offsets
    const int wb = (int)blockIdx.x - (std::is_same<T, double>::value ? a.nrecwg : 0);
    if (wb < a.nibwg) {
      // one lane per point (five lanes per point, as in bulk_band_fl, made
      // this launch slower: 600 workgroups ahead of the tiles instead of 112)
      const int k = wb * (TX * TY) + (int)(threadIdx.y * TX + threadIdx.x);
      if (k < a.nib) {
        const int2 p = a.ibpts[k];
        point_generic<T, FULL, DISS, OP>(a, p.x, p.y);
      }
      return;
    }
  }
  interior_tile<T, TX, TY, FULL, NTS, HET, DISS, OP, JI, MC>(
      a, tiles[linear_tile(a.xcd_swizzle, (std::is_same<T, double>::value ? a.nrecwg : 0) + a.nibwg)]);
}

// Batched interior kernel: the tiles of many problems in one launch; tile
// .w selects the problem's arguments in tab.
template <typename T, int TX, int TY, bool FULL, bool NTS>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(FULL ? 2 : 4)))
bulk_interior_batch(const BulkArgs<T>* __restrict__ tab, const int4* __restrict__ tiles, int swz) {
  const int4 td = tiles[linear_tile(swz)];
  interior_tile<T, TX, TY, FULL, NTS, false, false>(tab[td.w], td);
}


// Band kernel: the points within the closure depth of their block's edge
// (one-sided closures, SAT), one lane each, from a list of (i, j).
template <typename T, bool FULL, bool EXT, int OP = 4>
__global__ void __launch_bounds__(64) bulk_band(BulkArgs<T> a, const int2* __restrict__ pts, int npts) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npts) return;
  const int2 p = pts[k];
  point_generic<T, FULL, EXT, OP>(a, p.x, p.y);
}

// Layer kernel: the points of the perfectly matched layer blocks, one lane
// each, from a list that runs row by row through each block (consecutive
// lanes read consecutive addresses; entries with x < 0 pad the inner list
// to aligned runs of 64).  It carries only the layer's code;
// INNER = true serves the points four or more from every block edge with
// the interior form alone, which keeps more waves in flight.  NTS: F and
// DF stored non-temporally (working set beyond the last-level cache).
template <typename T, bool FULL, bool INNER, bool NTS>
__global__ void __launch_bounds__(256) bulk_pml(BulkArgs<T> a, const int2* __restrict__ pts, int npts) {
  // the list runs row by row: with the XCD swizzle each XCD takes a
  // contiguous run of it, so the rows a stencil shares are fetched into one
  // L2 rather than into up to five
  const int k = linear_tile(a.xcd_swizzle) * blockDim.x + threadIdx.x;
  if (k >= npts) return;
  const int2 p = pts[k];
  if (p.x < 0) return;  // padding of an aligned run
  const int i = p.x, j = p.y;
  const int bi = a.blk[a.rowb[j] * a.nbx + a.colb[i]];
  const BlockDev& B = a.blocks[bi];
  const long long cc = (long long)j * a.ld + i;
  T d[6];
  scaled_rates<T>(a, cc, d);
  pml_rates<T, INNER>(a, B, i, j, cc, d);
  T Fc[6];
#pragma unroll
  for (int l = 0; l < 6; l++) Fc[l] = a.Fin[l * a.ps + cc];
  finish_point<T, FULL, NTS>(a, B, bi, i, j, cc, d, Fc);
}

// Batched band kernel: pts[k] = (i, j, problem, valid), every wave's 64
// entries of one problem (padding entries have valid = 0).
template <typename T, bool FULL>
__global__ void __launch_bounds__(64) bulk_band_batch(const BulkArgs<T>* __restrict__ tab,
                                                      const int4* __restrict__ pts, int npts) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npts) return;
  const int4 p = pts[k];
  const int m = __builtin_amdgcn_readfirstlane(p.z);
  if (!p.w) return;
  // ensemble members have neither a heterogeneous medium nor dissipation
  point_generic<T, FULL, false>(tab[m], p.x, p.y);
}

// ---------------------------------------------------------------------
// Antiplane shear (mode 3): fields vz, sxz, syz (rates_interior_mode3,
// rates_boundary_mode3, rates.f90:8-381).  DF1 pairs sxz with the
// y metrics and syz with the x metrics, scaled by DI/rho; DF2 takes vz with
// the y metrics and DF3 vz with the x metrics, scaled by DI*G.  The terms a
// field lacks are absent from its sum; the rest keep the order of the
// plane-strain expressions.
// ---------------------------------------------------------------------

// interior point: a(s) = DI(s)/rho, b(s) = DI(s)*G (rates.f90:24-28), s up
// to the operator's interior half-width H.  HET: a heterogeneous medium
// (rates_h_interior_mode3, rates_heterogeneous.f90:8-48), ")*DI(s)*
// JacInv/rho" and ")*DI(s)*G*JacInv" with the point's rho and G, as(s) then
// holding DI(s), bs(s) unused.
template <typename T, typename Acc, int H = 2, bool HET = false>
__device__ __forceinline__ void interior_rates3(T* d, const Acc& A, const T (&as)[3], const T (&bs)[3],
                                                T Ji, T rho = (T)0, T G = (T)0) {
#pragma unroll
  for (int s = 1; s <= H; s++) {
    const T ayp = (T)0.5 * (A.yr(s) + A.yr(0)), axp = (T)0.5 * (A.xr(s) + A.xr(0));
    const T aym = (T)0.5 * (A.yr(-s) + A.yr(0)), axm = (T)0.5 * (A.xr(-s) + A.xr(0));
    const T byp = (T)0.5 * (A.yq(s) + A.yq(0)), bxp = (T)0.5 * (A.xq(s) + A.xq(0));
    const T bym = (T)0.5 * (A.yq(-s) + A.yq(0)), bxm = (T)0.5 * (A.xq(-s) + A.xq(0));
    const T t1 = ayp * A.f(1, s, 0) - axp * A.f(2, s, 0) - aym * A.f(1, -s, 0) + axm * A.f(2, -s, 0) -
                 byp * A.f(1, 0, s) + bxp * A.f(2, 0, s) + bym * A.f(1, 0, -s) - bxm * A.f(2, 0, -s);
    if constexpr (HET) d[0] = d[0] + t1 * as[s - 1] * Ji / rho;
    else d[0] = d[0] + t1 * as[s - 1] * Ji;
    const T t2 = ayp * A.f(0, s, 0) - aym * A.f(0, -s, 0) - byp * A.f(0, 0, s) + bym * A.f(0, 0, -s);
    if constexpr (HET) d[1] = d[1] + t2 * as[s - 1] * G * Ji;
    else d[1] = d[1] + t2 * bs[s - 1] * Ji;
    const T t3 = -(axp * A.f(0, s, 0)) + axm * A.f(0, -s, 0) + bxp * A.f(0, 0, s) - bxm * A.f(0, 0, -s);
    if constexpr (HET) d[2] = d[2] + t3 * as[s - 1] * G * Ji;
    else d[2] = d[2] + t3 * bs[s - 1] * Ji;
  }
}

// interior point with artificial dissipation (rates_dissipation_interior_
// mode3, rates.f90:939-999; rates_h_dissipation_interior_mode3,
// rates_heterogeneous.f90:920-978): c(0) F first, then per s and field
// (X)*a(s)*JacInv + c(s)*JacInv*(Y), Y the J-weighted neighbours of the
// field itself; a(s) as interior_rates3, or with a medium ")*DI(s)/rho*
// JacInv" and ")*DI(s)*G*JacInv" (as[] then holds DI(s)).  SBP4 only.
template <typename T, typename Acc, bool HET = false>
__device__ __forceinline__ void interior_rates3_diss(T* d, const Acc& A, const T (&as)[3], const T (&bs)[3],
                                                     const T (&dA)[3], T Ji, T rho = (T)0, T G = (T)0) {
#pragma unroll
  for (int l = 0; l < 3; l++) d[l] = d[l] + dA[0] * A.f(l, 0, 0);
#pragma unroll
  for (int s = 1; s <= 2; s++) {
    const T ayp = (T)0.5 * (A.yr(s) + A.yr(0)), axp = (T)0.5 * (A.xr(s) + A.xr(0));
    const T aym = (T)0.5 * (A.yr(-s) + A.yr(0)), axm = (T)0.5 * (A.xr(-s) + A.xr(0));
    const T byp = (T)0.5 * (A.yq(s) + A.yq(0)), bxp = (T)0.5 * (A.xq(s) + A.xq(0));
    const T bym = (T)0.5 * (A.yq(-s) + A.yq(0)), bxm = (T)0.5 * (A.xq(-s) + A.xq(0));
    T x[3];
    x[0] = ayp * A.f(1, s, 0) - axp * A.f(2, s, 0) - aym * A.f(1, -s, 0) + axm * A.f(2, -s, 0) -
           byp * A.f(1, 0, s) + bxp * A.f(2, 0, s) + bym * A.f(1, 0, -s) - bxm * A.f(2, 0, -s);
    x[1] = ayp * A.f(0, s, 0) - aym * A.f(0, -s, 0) - byp * A.f(0, 0, s) + bym * A.f(0, 0, -s);
    x[2] = -(axp * A.f(0, s, 0)) + axm * A.f(0, -s, 0) + bxp * A.f(0, 0, s) - bxm * A.f(0, 0, -s);
#pragma unroll
    for (int l = 0; l < 3; l++) {
      const T y = A.j(s, 0) * A.f(l, s, 0) + A.j(-s, 0) * A.f(l, -s, 0) + A.j(0, s) * A.f(l, 0, s) +
                  A.j(0, -s) * A.f(l, 0, -s);
      T u;
      if constexpr (HET) u = l == 0 ? x[l] * as[s - 1] / rho * Ji : x[l] * as[s - 1] * G * Ji;
      else u = x[l] * (l == 0 ? as[s - 1] : bs[s - 1]) * Ji;
      d[l] = d[l] + u + dA[s] * Ji * y;
    }
  }
}

// one-sided q, one-sided r, central q, central r terms of field l
template <typename T, int l>
__device__ __forceinline__ T q1_3(const T* __restrict__ F, const T* __restrict__ yr, const T* __restrict__ xr,
                                  long long ps, long long kk, long long cc) {
  const T ay = (T)0.5 * (yr[kk] + yr[cc]), ax = (T)0.5 * (xr[kk] + xr[cc]);
  if constexpr (l == 0) return ay * F[ps + kk] - ax * F[2 * ps + kk];
  else if constexpr (l == 1) return ay * F[kk];
  else return -(ax * F[kk]);
}

template <typename T, int l>
__device__ __forceinline__ T r1_3(const T* __restrict__ F, const T* __restrict__ yq, const T* __restrict__ xq,
                                  long long ps, long long kk, long long cc) {
  const T by = (T)0.5 * (yq[kk] + yq[cc]), bx = (T)0.5 * (xq[kk] + xq[cc]);
  if constexpr (l == 0) return -(by * F[ps + kk]) + bx * F[2 * ps + kk];
  else if constexpr (l == 1) return -(by * F[kk]);
  else return bx * F[kk];
}

template <typename T, int l>
__device__ __forceinline__ T qc_3(const T* __restrict__ F, const T* __restrict__ yr, const T* __restrict__ xr,
                                  long long ps, long long kp, long long km, long long cc) {
  const T ayp = (T)0.5 * (yr[kp] + yr[cc]), axp = (T)0.5 * (xr[kp] + xr[cc]);
  const T aym = (T)0.5 * (yr[km] + yr[cc]), axm = (T)0.5 * (xr[km] + xr[cc]);
  if constexpr (l == 0)
    return ayp * F[ps + kp] - axp * F[2 * ps + kp] - aym * F[ps + km] + axm * F[2 * ps + km];
  else if constexpr (l == 1) return ayp * F[kp] - aym * F[km];
  else return -(axp * F[kp]) + axm * F[km];
}

template <typename T, int l>
__device__ __forceinline__ T rc_3(const T* __restrict__ F, const T* __restrict__ yq, const T* __restrict__ xq,
                                  long long ps, long long kp, long long km, long long cc) {
  const T byp = (T)0.5 * (yq[kp] + yq[cc]), bxp = (T)0.5 * (xq[kp] + xq[cc]);
  const T bym = (T)0.5 * (yq[km] + yq[cc]), bxm = (T)0.5 * (xq[km] + xq[cc]);
  if constexpr (l == 0)
    return -(byp * F[ps + kp]) + bxp * F[2 * ps + kp] + bym * F[ps + km] - bxm * F[2 * ps + km];
  else if constexpr (l == 1) return -(byp * F[kp]) + bym * F[km];
  else return bxp * F[kp] - bxm * F[km];
}

// ")*D/rho*JacInv" for DF1, ")*D*G*JacInv" for DF2 and DF3
template <typename T, int l>
__device__ __forceinline__ T bscale3(T t, T D, T rho, T G) {
  if constexpr (l == 0) return t * D / rho;
  else return t * D * G;
}

// SAT (DF = DF+Ks*(Fhat-F), time_step.f90:829-830) and the closures of a
// point within the closure depth of its block's edge
template <typename T, int OP = 4>
__device__ __forceinline__ void closure_rates3(const BulkArgs<T>& a, const BlockDev& B, const int i,
                                               const int j, const long long cc, T (&d)[6]) {
  constexpr int NB = Sbp<OP>::nbst, NS = Sbp<OP>::nbnd, HW = Sbp<OP>::half;
  const long long ps = a.ps;
  const int ld = a.ld;
  const T* __restrict__ F = a.Fin;
  const int dl = i - B.mx, dr = B.px - i, db = j - B.my, dtp = B.py - j;
  if (dl == 0 || dr == 0 || db == 0 || dtp == 0) {
    T Fe[3];
#pragma unroll
    for (int l = 0; l < 3; l++) Fe[l] = F[l * ps + cc];
    const int NS = a.sd.NS;
    const int sides[4] = {dl == 0 ? B.side_off[FDM_SIDE_L] + db : -1,
                          dr == 0 ? B.side_off[FDM_SIDE_R] + db : -1,
                          db == 0 ? B.side_off[FDM_SIDE_B] + dl : -1,
                          dtp == 0 ? B.side_off[FDM_SIDE_T] + dl : -1};
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int so = sides[e];
      if (so < 0) continue;
      const T Ks = (T)a.sd.M[3 * NS + so];
#pragma unroll
      for (int l = 0; l < 3; l++) d[l] = d[l] + Ks * ((T)a.sd.Fhat[l * NS + so] - Fe[l]);
    }
  }
  const T* __restrict__ xq = a.xq;
  const T* __restrict__ xr = a.xr;
  const T* __restrict__ yq = a.yq;
  const T* __restrict__ yr = a.yr;
  const T Ji = a.Ji[cc];
  // a heterogeneous medium: the point's rho and G in the same expressions
  // (rates_h_boundary_mode3, rates_heterogeneous.f90:51-372)
  const T rho = a.mrho ? a.mrho[cc] : (T)B.rho, G = a.mrho ? a.mG[cc] : (T)B.G;
  const bool qone = (dl < NB) || (dr < NB);
  const bool rone = (db < NB) || (dtp < NB);
  const long long row = (long long)j * ld;
  const long long colm = (long long)B.my * ld + i, colp = (long long)B.py * ld + i;
  // artificial dissipation (rates_dissipation_boundary_mode3, rates.f90:
  // 1002-1394; SBP4 only): after each one-sided or central term, and the
  // strips' diagonal between their two loops; corners have none
  const bool diss = OP == 4 && B.Cdiss > 0.0;
  const T* __restrict__ Jp = a.J;
  const T Cd = (T)B.Cdiss;
  if (qone) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const long long kq = (dl < NB) ? row + B.mx + s : row + B.px - s;
      const T Dq = (dl < NB) ? (T)sbp_dl<OP>(dl, s) : (T)(-sbp_dl<OP>(dr, s));
      d[0] = d[0] + bscale3<T, 0>(q1_3<T, 0>(F, yr, xr, ps, kq, cc), Dq, rho, G) * Ji;
      d[1] = d[1] + bscale3<T, 1>(q1_3<T, 1>(F, yr, xr, ps, kq, cc), Dq, rho, G) * Ji;
      d[2] = d[2] + bscale3<T, 2>(q1_3<T, 2>(F, yr, xr, ps, kq, cc), Dq, rho, G) * Ji;
      if constexpr (OP == 4)
        if (diss) diss_one<T, 3>(d, F, Jp, ps, kq, (T)((dl < 4) ? kAL[dl][s] : kAL[dr][s]), Ji, Cd);
      if (rone) {
        const long long kr = (db < NB) ? colm + (long long)s * ld : colp - (long long)s * ld;
        const T Dr = (db < NB) ? (T)sbp_dl<OP>(db, s) : (T)(-sbp_dl<OP>(dtp, s));
        d[0] = d[0] + bscale3<T, 0>(r1_3<T, 0>(F, yq, xq, ps, kr, cc), Dr, rho, G) * Ji;
        d[1] = d[1] + bscale3<T, 1>(r1_3<T, 1>(F, yq, xq, ps, kr, cc), Dr, rho, G) * Ji;
        d[2] = d[2] + bscale3<T, 2>(r1_3<T, 2>(F, yq, xq, ps, kr, cc), Dr, rho, G) * Ji;
        if constexpr (OP == 4)
          if (diss) diss_one<T, 3>(d, F, Jp, ps, kr, (T)((db < 4) ? kAL[db][s] : kAL[dtp][s]), Ji, Cd);
      }
    }
    if (!rone) {
      if constexpr (OP == 4)
        if (diss) diss_diag<T, 3>(d, F, ps, cc, Cd);
#pragma unroll
      for (int s = 1; s <= HW; s++) {
        const T D = (T)sbp_di<OP>(s);
        const long long kp = cc + (long long)s * ld, km = cc - (long long)s * ld;
        d[0] = d[0] + bscale3<T, 0>(rc_3<T, 0>(F, yq, xq, ps, kp, km, cc), D, rho, G) * Ji;
        d[1] = d[1] + bscale3<T, 1>(rc_3<T, 1>(F, yq, xq, ps, kp, km, cc), D, rho, G) * Ji;
        d[2] = d[2] + bscale3<T, 2>(rc_3<T, 2>(F, yq, xq, ps, kp, km, cc), D, rho, G) * Ji;
        if constexpr (OP == 4)
          if (diss) diss_central<T, 3>(d, F, Jp, ps, kp, km, (s == 1) ? (T)FDM_AI1 : (T)FDM_AI2, Ji, Cd);
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const long long kr = (db < NB) ? colm + (long long)s * ld : colp - (long long)s * ld;
      const T Dr = (db < NB) ? (T)sbp_dl<OP>(db, s) : (T)(-sbp_dl<OP>(dtp, s));
      d[0] = d[0] + bscale3<T, 0>(r1_3<T, 0>(F, yq, xq, ps, kr, cc), Dr, rho, G) * Ji;
      d[1] = d[1] + bscale3<T, 1>(r1_3<T, 1>(F, yq, xq, ps, kr, cc), Dr, rho, G) * Ji;
      d[2] = d[2] + bscale3<T, 2>(r1_3<T, 2>(F, yq, xq, ps, kr, cc), Dr, rho, G) * Ji;
      if constexpr (OP == 4)
        if (diss) diss_one<T, 3>(d, F, Jp, ps, kr, (T)((db < 4) ? kAL[db][s] : kAL[dtp][s]), Ji, Cd);
    }
    if constexpr (OP == 4)
      if (diss) diss_diag<T, 3>(d, F, ps, cc, Cd);
#pragma unroll
    for (int s = 1; s <= HW; s++) {
      const T D = (T)sbp_di<OP>(s);
      const long long kp = cc + s, km = cc - s;
      d[0] = d[0] + bscale3<T, 0>(qc_3<T, 0>(F, yr, xr, ps, kp, km, cc), D, rho, G) * Ji;
      d[1] = d[1] + bscale3<T, 1>(qc_3<T, 1>(F, yr, xr, ps, kp, km, cc), D, rho, G) * Ji;
      d[2] = d[2] + bscale3<T, 2>(qc_3<T, 2>(F, yr, xr, ps, kp, km, cc), D, rho, G) * Ji;
      if constexpr (OP == 4)
        if (diss) diss_central<T, 3>(d, F, Jp, ps, kp, km, (s == 1) ? (T)FDM_AI1 : (T)FDM_AI2, Ji, Cd);
    }
  }
}

// one grid point of one stage in antiplane shear, neighbours from global
// memory: A*DF, SAT, rates, update (fields.f90:1117-1133, 1225-1242)
template <typename T, int OP = 4>
__device__ __forceinline__ void point_generic3(const BulkArgs<T>& a, const int i, const int j) {
  constexpr int NB = Sbp<OP>::nbst, HW = Sbp<OP>::half;
  const int bi = a.blk[a.rowb[j] * a.nbx + a.colb[i]];
  const BlockDev& B = a.blocks[bi];
  const long long ps = a.ps;
  const int ld = a.ld;
  const long long cc = (long long)j * ld + i;
  T d[6];
  if (a.a_zero) {
#pragma unroll
    for (int l = 0; l < 3; l++) d[l] = (T)0.0;
  } else {
    const T A = (T)a.A;
#pragma unroll
    for (int l = 0; l < 3; l++) d[l] = A * a.DF[l * ps + cc];
  }
  const bool inner = i - B.mx >= NB && B.px - i >= NB && j - B.my >= NB && B.py - j >= NB;
  bool done = false;
  if constexpr (OP == 4) {
    if (inner && B.Cdiss > 0.0) {
      const GlobalAcc<T> acc{a.Fin, a.xq, a.xr, a.yq, a.yr, ps, cc, ld, a.J};
      const T dA[3] = {(T)B.dA[0], (T)B.dA[1], (T)B.dA[2]};
      if (a.mrho) {
        const T D[3] = {(T)FDM_DI1, (T)FDM_DI2, (T)0};
        interior_rates3_diss<T, GlobalAcc<T>, true>(d, acc, D, D, dA, a.Ji[cc], a.mrho[cc], a.mG[cc]);
      } else {
        const T as[3] = {(T)B.cI[0][0], (T)B.cI[0][1], (T)B.cI[0][2]};
        const T bs[3] = {(T)B.cI[3][0], (T)B.cI[3][1], (T)B.cI[3][2]};
        interior_rates3_diss<T, GlobalAcc<T>>(d, acc, as, bs, dA, a.Ji[cc]);
      }
      done = true;
    }
  }
  if (done) {
  } else if (inner) {
    const GlobalAcc<T> acc{a.Fin, a.xq, a.xr, a.yq, a.yr, ps, cc, ld, nullptr};
    if (a.mrho) {
      const T D[3] = {(T)sbp_di<OP>(1), HW > 1 ? (T)sbp_di<OP>(2) : (T)0, HW > 2 ? (T)sbp_di<OP>(3) : (T)0};
      interior_rates3<T, GlobalAcc<T>, HW, true>(d, acc, D, D, a.Ji[cc], a.mrho[cc], a.mG[cc]);
    } else {
      const T as[3] = {(T)B.cI[0][0], (T)B.cI[0][1], (T)B.cI[0][2]};
      const T bs[3] = {(T)B.cI[3][0], (T)B.cI[3][1], (T)B.cI[3][2]};
      interior_rates3<T, GlobalAcc<T>, HW>(d, acc, as, bs, a.Ji[cc]);
    }
  } else {
    closure_rates3<T, OP>(a, B, i, j, cc, d);
  }
  if (a.do_update) {
    const T Bdt = (T)a.Bdt;
#pragma unroll
    for (int l = 0; l < 3; l++) a.Fout[l * ps + cc] = a.Fin[l * ps + cc] + Bdt * d[l];
  }
#pragma unroll
  for (int l = 0; l < 3; l++) a.DF[l * ps + cc] = d[l];
}

template <typename T, int OP = 4>
__global__ void __launch_bounds__(256) bulk_kernel3(BulkArgs<T> a) {
  int tx, ty;
  tile_of(a.xcd_swizzle, tx, ty);
  const int i = tx * blockDim.x + threadIdx.x;
  const int j = ty * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  point_generic3<T, OP>(a, i, j);
}

// Antiplane band kernel with one lane per (point, field) for blocks
// without dissipation: 21 points per wave, three lanes each (lane 63
// idle).  The lane of field l evaluates DF(l) alone with closure_rates3's
// terms in their order (a medium's rho and G per point as there); the three
// fields are not coupled after the rates, so each lane finishes its own.
// A point of a dissipative block is left to bulk_band3 (the host sends
// only blocks without dissipation here).
template <typename T, int OP = 4>
__global__ void __launch_bounds__(64) bulk_band3_fl(BulkArgs<T> a, const int2* __restrict__ pts, int npts) {
  constexpr int NB = Sbp<OP>::nbst, NS = Sbp<OP>::nbnd, HW = Sbp<OP>::half;
  const int lane = threadIdx.x;
  const int grp = lane / 3, l = lane - 3 * grp;
  const int k = blockIdx.x * 21 + grp;
  if (lane >= 63 || k >= npts) return;
  const int2 p = pts[k];
  const int i = p.x, j = p.y;
  const int bi = a.blk[a.rowb[j] * a.nbx + a.colb[i]];
  const BlockDev& B = a.blocks[bi];
  const long long ps = a.ps;
  const int ld = a.ld;
  const long long cc = (long long)j * ld + i;
  const T* __restrict__ F = a.Fin;
  T d = a.a_zero ? (T)0.0 : (T)a.A * a.DF[l * ps + cc];
  const int dl = i - B.mx, dr = B.px - i, db = j - B.my, dtp = B.py - j;
  if (dl == 0 || dr == 0 || db == 0 || dtp == 0) {
    const T Fe = F[l * ps + cc];
    const int NSd = a.sd.NS;
    const int sides[4] = {dl == 0 ? B.side_off[FDM_SIDE_L] + db : -1,
                          dr == 0 ? B.side_off[FDM_SIDE_R] + db : -1,
                          db == 0 ? B.side_off[FDM_SIDE_B] + dl : -1,
                          dtp == 0 ? B.side_off[FDM_SIDE_T] + dl : -1};
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int so = sides[e];
      if (so < 0) continue;
      const T Ks = (T)a.sd.M[3 * NSd + so];
      d = d + Ks * ((T)a.sd.Fhat[l * NSd + so] - Fe);
    }
  }
  const T* __restrict__ xq = a.xq;
  const T* __restrict__ xr = a.xr;
  const T* __restrict__ yq = a.yq;
  const T* __restrict__ yr = a.yr;
  const T Ji = a.Ji[cc];
  const T rho = a.mrho ? a.mrho[cc] : (T)B.rho, G = a.mrho ? a.mG[cc] : (T)B.G;
  // field 0 (vz) pairs sxz with the y-metrics and syz with the x-metrics;
  // fields 1 and 2 (sxz, syz) take vz alone
  const T* __restrict__ Fa = F + (l == 0 ? ps : 0);
  const T* __restrict__ Fb = F + 2 * ps;
  const bool qone = (dl < NB) || (dr < NB);
  const bool rone = (db < NB) || (dtp < NB);
  const long long row = (long long)j * ld;
  const long long colm = (long long)B.my * ld + i, colp = (long long)B.py * ld + i;
  // q1_3, r1_3, qc_3, rc_3 and bscale3 of field l
  auto q1 = [&](long long kk) {
    const T ay = (T)0.5 * (yr[kk] + yr[cc]), ax = (T)0.5 * (xr[kk] + xr[cc]);
    return l == 0 ? ay * Fa[kk] - ax * Fb[kk] : (l == 1 ? ay * Fa[kk] : -(ax * Fa[kk]));
  };
  auto r1 = [&](long long kk) {
    const T by = (T)0.5 * (yq[kk] + yq[cc]), bx = (T)0.5 * (xq[kk] + xq[cc]);
    return l == 0 ? -(by * Fa[kk]) + bx * Fb[kk] : (l == 1 ? -(by * Fa[kk]) : bx * Fa[kk]);
  };
  auto sc = [&](T t, T D) { return l == 0 ? t * D / rho : t * D * G; };
  if (qone) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const long long kq = (dl < NB) ? row + B.mx + s : row + B.px - s;
      const T Dq = (dl < NB) ? (T)sbp_dl<OP>(dl, s) : (T)(-sbp_dl<OP>(dr, s));
      d = d + sc(q1(kq), Dq) * Ji;
      if (rone) {
        const long long kr = (db < NB) ? colm + (long long)s * ld : colp - (long long)s * ld;
        const T Dr = (db < NB) ? (T)sbp_dl<OP>(db, s) : (T)(-sbp_dl<OP>(dtp, s));
        d = d + sc(r1(kr), Dr) * Ji;
      }
    }
    if (!rone) {
#pragma unroll
      for (int s = 1; s <= HW; s++) {
        const T D = (T)sbp_di<OP>(s);
        const long long kp = cc + (long long)s * ld, km = cc - (long long)s * ld;
        const T byp = (T)0.5 * (yq[kp] + yq[cc]), bxp = (T)0.5 * (xq[kp] + xq[cc]);
        const T bym = (T)0.5 * (yq[km] + yq[cc]), bxm = (T)0.5 * (xq[km] + xq[cc]);
        const T t = l == 0 ? -(byp * Fa[kp]) + bxp * Fb[kp] + bym * Fa[km] - bxm * Fb[km]
                           : (l == 1 ? -(byp * Fa[kp]) + bym * Fa[km] : bxp * Fa[kp] - bxm * Fa[km]);
        d = d + sc(t, D) * Ji;
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const long long kr = (db < NB) ? colm + (long long)s * ld : colp - (long long)s * ld;
      const T Dr = (db < NB) ? (T)sbp_dl<OP>(db, s) : (T)(-sbp_dl<OP>(dtp, s));
      d = d + sc(r1(kr), Dr) * Ji;
    }
#pragma unroll
    for (int s = 1; s <= HW; s++) {
      const T D = (T)sbp_di<OP>(s);
      const long long kp = cc + s, km = cc - s;
      const T ayp = (T)0.5 * (yr[kp] + yr[cc]), axp = (T)0.5 * (xr[kp] + xr[cc]);
      const T aym = (T)0.5 * (yr[km] + yr[cc]), axm = (T)0.5 * (xr[km] + xr[cc]);
      const T t = l == 0 ? ayp * Fa[kp] - axp * Fb[kp] - aym * Fa[km] + axm * Fb[km]
                         : (l == 1 ? ayp * Fa[kp] - aym * Fa[km] : -(axp * Fa[kp]) + axm * Fa[km]);
      d = d + sc(t, D) * Ji;
    }
  }
  if (a.do_update) a.Fout[l * ps + cc] = F[l * ps + cc] + (T)a.Bdt * d;
  a.DF[l * ps + cc] = d;
}

template <typename T, int OP = 4>
__global__ void __launch_bounds__(64) bulk_band3(BulkArgs<T> a, const int2* __restrict__ pts, int npts) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npts) return;
  const int2 p = pts[k];
  point_generic3<T, OP>(a, p.x, p.y);
}

// interior tiles in antiplane shear: vz, sxz, syz with a halo of the
// operator's half-width, the metrics as in the plane-strain kernel, one
// barrier.  HET: a heterogeneous medium, the lane's own rho and G.  DISS:
// artificial dissipation in some block (SBP4), the Jacobian staged with the
// halo of F; a block without it keeps the plain expression.
template <typename T, int TX, int TY, bool NTS, int OP = 4, bool HET = false, bool DISS = false>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(DISS ? 3 : 4)))
bulk_interior3(BulkArgs<T> a, const int4* __restrict__ tiles) {
  static_assert(!DISS || OP == 4, "artificial dissipation is built for SBP4 only");
  constexpr int H = Sbp<OP>::half, NB = Sbp<OP>::nbst;
  constexpr int HX = TX + 2 * H, HY = TY + 2 * H, NT = TX * TY;
  if constexpr (!HET && !DISS && OP == 4) {
    // the inner band points ahead of the tiles, one lane per point, as in
    // the plane-strain launch (band split, on request).  This path takes
    // the FP64 kernel from 76 to 120 VGPRs (six waves per SIMD to four);
    // with the split off, the four-wave kernel made the antiplane step
    // shorter (0.356 against 0.366 ms on the C2-size grid, FP32-bulk 0.215
    // against 0.221).
    if ((int)blockIdx.x < a.nibwg) {
      const int k = (int)blockIdx.x * NT + (int)(threadIdx.y * TX + threadIdx.x);
      if (k < a.nib) {
        const int2 p = a.ibpts[k];
        point_generic3<T, OP>(a, p.x, p.y);
      }
      return;
    }
  }
  constexpr int NF = (HY * HX + NT - 1) / NT;
  constexpr int NR = (TY * HX + NT - 1) / NT;
  constexpr int NQ = (HY * TX + NT - 1) / NT;
  __shared__ T sF[3][HY][HX];
  __shared__ T sR[2][TY][HX];
  __shared__ T sQ[2][HY][TX];
  __shared__ T sJ[DISS ? HY : 1][DISS ? HX : 1];
  const int4 td = tiles[linear_tile(a.xcd_swizzle, a.nibwg)];
  const int bu = td.x, i0 = td.y, j0 = td.z;
  ConstBlockDev* B4 = (ConstBlockDev*)a.blocks + bu;
  const int tid = threadIdx.y * TX + threadIdx.x;
  const int ld = a.ld, nx1 = a.nx - 1, ny1 = a.ny - 1;
  const long long ps = a.ps;
  const T* __restrict__ F = a.Fin;
  T vF[NF][3], vR[NR][2], vQ[NQ][2], vJ[DISS ? NF : 1];
  // The loads stay under their lane conditions here (unlike the plane-strain
  // tile): the compiler then waits for the F halo before the metric loads,
  // which slows this launch a little and leaves the second stream's edge and
  // band kernels, as long as the launch here, room to keep up (C2-sized
  // antiplane step: 0.366 against 0.414 ms with the loads unconditional).
#pragma unroll
  for (int p = 0; p < NF; p++) {
    const int k = tid + p * NT;
    if (k < HY * HX) {
      const int r = k / HX, c = k - r * HX;
      const int gj = min(max(j0 - H + r, 0), ny1), gi = min(max(i0 - H + c, 0), nx1);
      const unsigned o = (unsigned)(gj * ld + gi) * (unsigned)sizeof(T);
#pragma unroll
      for (int l = 0; l < 3; l++) vF[p][l] = ld_off(F + l * ps, o);
      if constexpr (DISS) vJ[p] = ld_off(a.J, o);
    }
  }
#pragma unroll
  for (int p = 0; p < NR; p++) {
    const int k = tid + p * NT;
    if (k < TY * HX) {
      const int r = k / HX, c = k - r * HX;
      const int gj = min(j0 + r, ny1), gi = min(max(i0 - H + c, 0), nx1);
      const unsigned o = (unsigned)(gj * ld + gi) * (unsigned)sizeof(T);
      vR[p][0] = ld_off(a.xr, o);
      vR[p][1] = ld_off(a.yr, o);
    }
  }
#pragma unroll
  for (int p = 0; p < NQ; p++) {
    const int k = tid + p * NT;
    if (k < HY * TX) {
      const int r = k / TX, c = k - r * TX;
      const int gj = min(max(j0 - H + r, 0), ny1), gi = min(i0 + c, nx1);
      const unsigned o = (unsigned)(gj * ld + gi) * (unsigned)sizeof(T);
      vQ[p][0] = ld_off(a.xq, o);
      vQ[p][1] = ld_off(a.yq, o);
    }
  }
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const bool valid = i <= B4->px - NB && j <= B4->py - NB;
  const unsigned oc = (unsigned)(min(j, ny1) * ld + min(i, nx1)) * (unsigned)sizeof(T);
  T dfa[3];
  if (!a.a_zero) {
#pragma unroll
    for (int l = 0; l < 3; l++) dfa[l] = ld_off(a.DF + l * ps, oc);
  }
  const T Ji = ld_off(a.Ji, oc);
  T mrho = (T)0, mG = (T)0;
  if constexpr (HET) {
    mrho = ld_off(a.mrho, oc);
    mG = ld_off(a.mG, oc);
  }
#pragma unroll
  for (int p = 0; p < NF; p++) {
    const int k = tid + p * NT;
    if (k < HY * HX) {
      const int r = k / HX, c = k - r * HX;
#pragma unroll
      for (int l = 0; l < 3; l++) sF[l][r][c] = vF[p][l];
      if constexpr (DISS) sJ[r][c] = vJ[p];
    }
  }
#pragma unroll
  for (int p = 0; p < NR; p++) {
    const int k = tid + p * NT;
    if (k < TY * HX) {
      const int r = k / HX, c = k - r * HX;
      sR[0][r][c] = vR[p][0];
      sR[1][r][c] = vR[p][1];
    }
  }
#pragma unroll
  for (int p = 0; p < NQ; p++) {
    const int k = tid + p * NT;
    if (k < HY * TX) {
      const int r = k / TX, c = k - r * TX;
      sQ[0][r][c] = vQ[p][0];
      sQ[1][r][c] = vQ[p][1];
    }
  }
  __syncthreads();
  if (!valid) return;
  T d[6];
  if (a.a_zero) {
#pragma unroll
    for (int l = 0; l < 3; l++) d[l] = (T)0.0;
  } else {
    const T A = (T)a.A;
#pragma unroll
    for (int l = 0; l < 3; l++) d[l] = A * dfa[l];
  }
  const TileAcc<T, TX, TY, H> acc{sF, sR, sQ, (int)threadIdx.x, (int)threadIdx.y, &sJ[0][0]};
  bool diss_done = false;
  if constexpr (DISS) {
    if (B4->Cdiss > 0.0) {
      const T dA[3] = {(T)B4->dA[0], (T)B4->dA[1], (T)B4->dA[2]};
      if constexpr (HET) {
        const T D[3] = {(T)FDM_DI1, (T)FDM_DI2, (T)0};
        interior_rates3_diss<T, TileAcc<T, TX, TY, H>, true>(d, acc, D, D, dA, Ji, mrho, mG);
      } else {
        const T as[3] = {(T)B4->cI[0][0], (T)B4->cI[0][1], (T)B4->cI[0][2]};
        const T bs[3] = {(T)B4->cI[3][0], (T)B4->cI[3][1], (T)B4->cI[3][2]};
        interior_rates3_diss<T, TileAcc<T, TX, TY, H>>(d, acc, as, bs, dA, Ji);
      }
      diss_done = true;
    }
  }
  if (diss_done) {
  } else if constexpr (HET) {
    const T D[3] = {(T)sbp_di<OP>(1), H > 1 ? (T)sbp_di<OP>(2) : (T)0, H > 2 ? (T)sbp_di<OP>(3) : (T)0};
    interior_rates3<T, TileAcc<T, TX, TY, H>, H, true>(d, acc, D, D, Ji, mrho, mG);
  } else {
    const T as[3] = {(T)B4->cI[0][0], (T)B4->cI[0][1], (T)B4->cI[0][2]};
    const T bs[3] = {(T)B4->cI[3][0], (T)B4->cI[3][1], (T)B4->cI[3][2]};
    interior_rates3<T, TileAcc<T, TX, TY, H>, H>(d, acc, as, bs, Ji);
  }
  if (a.do_update) {
    const T Bdt = (T)a.Bdt;
#pragma unroll
    for (int l = 0; l < 3; l++) st_sel<NTS>(a.Fout + l * ps, oc, acc.f(l, 0, 0) + Bdt * d[l]);
  }
#pragma unroll
  for (int l = 0; l < 3; l++) st_sel<NTS>(a.DF + l * ps, oc, d[l]);
}

}  // namespace fdm
