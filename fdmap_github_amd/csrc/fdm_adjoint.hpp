// fdm_adjoint.hpp -- discrete adjoint of the time-stepping path.
//
// Forward stage s of LSRK (time_step.f90:71-167), per grid point and fault
// node, with A = A_s and b = B_s dt:
//   Fhat   = H(F, Psi)                        boundary / interface hats
//   DPsi  <- A DPsi + G(V, Psi)               state rate at the solved V
//   Psi   <- Psi + b DPsi
//   DF    <- A DF + R(F, Fhat),  DF6 = nu (DF3 + DF5)
//   F     <- F + b DF
// R (rates + SAT) is affine, H is affine except at frictional nodes, where
// V solves Slock - eta V = a N asinh(V exp(Psi/a) / (2 V0)).  The adjoint
// keeps lambda_F and mu = lambda_DF + b lambda_F (the adjoint of the stage's
// new DF including its share of the update); szz feeds nothing but itself,
// so both carry the five fields vx, vy, sxx, sxy, syy.  A backward stage is
//   nu_f   = lambda_DPsi + b lambda_Psi                (fault nodes)
//   edge:  lambda_Fhat = SAT^T mu at edge points, through H^T (friction by
//          the implicit function theorem at the traced V, Psi, N) into
//          contributions gE to lambda_F at edge points, lambda_Psi, and the
//          parameter gradients; lambda_DPsi <- A nu_f
//   bulk:  lambda_F <- lambda_F + R^T mu + gE
//          mu_prev  <- A mu + b_prev lambda_F
// R^T is applied as a gather with the SBP4 matrix entries
// D[row][col] (fd_coeff.f90:101-115, 242-246).
#pragma once

namespace fdm {

// SBP4 first-derivative matrix entry (row r, column c) on n >= 12 points:
// rows 0..3 the closure DL, the last four rows -DL mirrored, else DI.
__device__ __forceinline__ double sbp_entry(int r, int c, int n) {
  if (r < 4) return (c >= 0 && c < 6) ? kDL[r][c] : 0.0;
  if (r > n - 5) {
    const int rr = n - 1 - r, cr = n - 1 - c;
    return (cr >= 0 && cr < 6) ? -kDL[rr][cr] : 0.0;
  }
  const int s = c - r;
  if (s == 1) return FDM_DI1;
  if (s == -1) return -FDM_DI1;
  if (s == 2) return FDM_DI2;
  if (s == -2) return -FDM_DI2;
  return 0.0;
}

struct AdjBulkArgs {
  const double* lin;  // lambda_F of the later state (5 planes)
  double* lout;
  const double* min;  // mu of this stage
  double* mout;       // mu of the stage before
  const double* xq;
  const double* xr;
  const double* yq;
  const double* yr;
  const double* Ji;
  const int* colb;
  const int* rowb;
  const int* blk;
  int nbx;
  const BlockDev* blocks;
  const double* gE;  // [5][NS] edge-point contributions of the edge adjoint
  int NS;
  int nx, ny, ld;
  long long ps;
  double A, Bdt_prev;
};

// the transposed field couplings of rates_*_mode2 (rates.f90:411-444) for one
// stencil row: row kk of the q derivative (Q = true; ay, ax from yr, xr) or
// of the r derivative (Q = false; by, bx from yq, xq, opposite signs)
template <bool Q>
__device__ __forceinline__ void adj_row(double (&g)[5], const double (&w)[5], double ay, double ax,
                                        double gam) {
  if constexpr (Q) {
    g[2] += ay * w[0];
    g[3] -= ax * w[0];
    g[3] += ay * w[1];
    g[4] -= ax * w[1];
    g[0] += ay * w[2];
    g[1] -= ax * gam * w[2];
    g[1] += ay * w[3];
    g[0] -= ax * w[3];
    g[0] += ay * gam * w[4];
    g[1] -= ax * w[4];
  } else {
    g[2] -= ay * w[0];
    g[3] += ax * w[0];
    g[3] -= ay * w[1];
    g[4] += ax * w[1];
    g[0] -= ay * w[2];
    g[1] += ax * gam * w[2];
    g[1] -= ay * w[3];
    g[0] += ax * w[3];
    g[0] -= ay * gam * w[4];
    g[1] += ax * w[4];
  }
}

// the rows of one direction's derivative whose stencil reaches local index
// c of a block side of n points, in increasing row order: a central index
// (6 <= c <= n-7) is reached by interior rows c-2..c+2 only (D = DI(c - r),
// zero at r = c); any other by rows max(0, c-5)..min(n-1, c+5), whose
// entries come from the closures or DI.  Both loops have a fixed trip count
// and no branch, so that a lane's loads are in flight together: a row
// outside the range reads a clamped, valid address and carries D = 0, which
// adds nothing to g.  base: address of row 0's point, step: between rows.
template <bool Q>
__device__ __forceinline__ void adj_dir(const AdjBulkArgs& a, double (&g)[5], const int c, const int n,
                                        const long long base, const long long step, const long long cc,
                                        const double (&cl)[5], const double gam) {
  const double* __restrict__ m1 = Q ? a.yr : a.yq;  // the "y" metric of the pair
  const double* __restrict__ m2 = Q ? a.xr : a.xq;
  const double y0 = m1[cc], x0 = m2[cc];
  const long long ps = a.ps;
  auto row = [&](const int r, const double D) {
    const long long kk = base + (long long)r * step;
    const double Jk = a.Ji[kk];
    double w[5];
#pragma unroll
    for (int l = 0; l < 5; l++) w[l] = a.min[l * ps + kk] * Jk * cl[l];
    const double ay = 0.5 * (m1[kk] + y0) * D, ax = 0.5 * (m2[kk] + x0) * D;
    adj_row<Q>(g, w, ay, ax, gam);
  };
  if (c >= 6 && c <= n - 7) {
    row(c - 2, FDM_DI2);
    row(c - 1, FDM_DI1);
    row(c + 1, -FDM_DI1);
    row(c + 2, -FDM_DI2);
  } else {
    const int r0 = max(0, c - 5), r1 = min(n - 1, c + 5);
#pragma unroll
    for (int t = 0; t < 11; t++) {
      const int r = min(r0 + t, r1);
      const double D = (r0 + t <= r1) ? sbp_entry(r, c, n) : 0.0;
      row(r, D);
    }
  }
}

// R^T mu at (i, j), edge contributions, and the two recurrences
__device__ __forceinline__ void adj_point(const AdjBulkArgs& a, const int i, const int j) {
  const int bi = a.blk[a.rowb[j] * a.nbx + a.colb[i]];
  const BlockDev& B = a.blocks[bi];
  const int nq = B.px - B.mx + 1, nr = B.py - B.my + 1;
  const int il = i - B.mx, jl = j - B.my;
  const long long ps = a.ps;
  const int ld = a.ld;
  const long long cc = (long long)j * ld + i;
  const double gam = B.gamma;
  const double cl[5] = {1.0 / B.rho, 1.0 / B.rho, B.M, B.G, B.M};
  double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  // q: rows of the block's q-derivative whose stencil reaches column il
  adj_dir<true>(a, g, il, nq, (long long)j * ld + B.mx, 1, cc, cl, gam);
  // r: rows of the block's r-derivative whose stencil reaches row jl
  adj_dir<false>(a, g, jl, nr, (long long)B.my * ld + i, ld, cc, cl, gam);
  // SAT and hat contributions of the sides this point lies on
  const int NS = a.NS;
  int so[4], nso = 0;
  if (il == 0) so[nso++] = B.side_off[FDM_SIDE_L] + jl;
  if (il == nq - 1) so[nso++] = B.side_off[FDM_SIDE_R] + jl;
  if (jl == 0) so[nso++] = B.side_off[FDM_SIDE_B] + il;
  if (jl == nr - 1) so[nso++] = B.side_off[FDM_SIDE_T] + il;
  for (int k = 0; k < nso; k++)
#pragma unroll
    for (int l = 0; l < 5; l++) g[l] += a.gE[l * NS + so[k]];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    const double lo = a.lin[l * ps + cc] + g[l];
    a.lout[l * ps + cc] = lo;
    a.mout[l * ps + cc] = a.A * a.min[l * ps + cc] + a.Bdt_prev * lo;
  }
}

__global__ void __launch_bounds__(256) adj_bulk_kernel(AdjBulkArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  adj_point(a, i, j);
}

// The gathers of one direction with the direction a run-time choice (q:
// columns of row j, r: rows of column i): 11 rows from max(0, c-5), D = 0
// where the stencil does not reach (a central point's rows included, which
// leaves its sum unchanged), the r direction's signs folded into ay and ax
// (exact, -(a*b) = (-a)*b), so both lanes of a pair run one code path.
__device__ __forceinline__ void adj_dir_any(const AdjBulkArgs& a, double (&g)[5], const int c, const int n,
                                            const long long base, const long long step,
                                            const long long cc, const double (&cl)[5], const double gam,
                                            const bool q) {
  const double* __restrict__ m1 = q ? a.yr : a.yq;
  const double* __restrict__ m2 = q ? a.xr : a.xq;
  const double sg = q ? 1.0 : -1.0;
  const double y0 = m1[cc], x0 = m2[cc];
  const long long ps = a.ps;
  const int r0 = max(0, c - 5), r1 = min(n - 1, c + 5);
#pragma unroll
  for (int t = 0; t < 11; t++) {
    const int r = min(r0 + t, r1);
    const double D = (r0 + t <= r1) ? sbp_entry(r, c, n) : 0.0;
    const long long kk = base + (long long)r * step;
    const double Jk = a.Ji[kk];
    double w[5];
#pragma unroll
    for (int l = 0; l < 5; l++) w[l] = a.min[l * ps + kk] * Jk * cl[l];
    const double ay = sg * (0.5 * (m1[kk] + y0) * D), ax = sg * (0.5 * (m2[kk] + x0) * D);
    adj_row<true>(g, w, ay, ax, gam);
  }
}

// points within six of their block's edge (closure rows reach them, or
// SAT), two lanes per point: the even lane gathers the q-derivative rows,
// the odd lane the r-derivative rows (so each lane's chain of dependent
// loads is half as long, and twice the waves are in flight); the pair adds
// its sums, and the even lane adds the edge contributions and finishes the
// point as adj_point does
__global__ void __launch_bounds__(64) adj_band_kernel(AdjBulkArgs a, const int2* __restrict__ pts,
                                                      int npts) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 1;
  if (k >= npts) return;  // both lanes of a pair leave together
  const bool q = (t & 1) == 0;
  const int2 p = pts[k];
  const int i = p.x, j = p.y;
  const int bi = a.blk[a.rowb[j] * a.nbx + a.colb[i]];
  const BlockDev& B = a.blocks[bi];
  const int nq = B.px - B.mx + 1, nr = B.py - B.my + 1;
  const int il = i - B.mx, jl = j - B.my;
  const long long ps = a.ps;
  const int ld = a.ld;
  const long long cc = (long long)j * ld + i;
  const double gam = B.gamma;
  const double cl[5] = {1.0 / B.rho, 1.0 / B.rho, B.M, B.G, B.M};
  double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  adj_dir_any(a, g, q ? il : jl, q ? nq : nr, q ? (long long)j * ld + B.mx : (long long)B.my * ld + i,
              q ? 1 : ld, cc, cl, gam, q);
#pragma unroll
  for (int l = 0; l < 5; l++) g[l] = g[l] + __shfl_xor(g[l], 1);
  if (!q) return;
  const int NS = a.NS;
  int so[4], nso = 0;
  if (il == 0) so[nso++] = B.side_off[FDM_SIDE_L] + jl;
  if (il == nq - 1) so[nso++] = B.side_off[FDM_SIDE_R] + jl;
  if (jl == 0) so[nso++] = B.side_off[FDM_SIDE_B] + il;
  if (jl == nr - 1) so[nso++] = B.side_off[FDM_SIDE_T] + il;
  for (int m = 0; m < nso; m++)
#pragma unroll
    for (int l = 0; l < 5; l++) g[l] += a.gE[l * NS + so[m]];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    const double lo = a.lin[l * ps + cc] + g[l];
    a.lout[l * ps + cc] = lo;
    a.mout[l * ps + cc] = a.A * a.min[l * ps + cc] + a.Bdt_prev * lo;
  }
}

// Interior adjoint kernel over TX x TY tiles of each block's rectangle
// [mx+6, px-6] x [my+6, py-6], where only interior SBP rows reach a point
// and no SAT applies: R^T mu is the central stencil of the forward kernel
// with the roles of rows and columns exchanged,
//   sum_s DI(s) [a(P-s) w(P-s) - a(P+s) w(P+s)],  w_l = mu_l Ji c_l,
// a the metric average of the pair.  w over the tile with a halo of two in
// q and r, xr/yr with the q halo and xq/yq with the r halo are staged in
// LDS.  Halo indices are clamped into the grid; a clamped value never
// reaches an active lane.  tiles[k] = (block, i0, j0, 0).
template <int TX, int TY, bool NTS>
__global__ void __launch_bounds__(TX * TY) adj_interior_kernel(AdjBulkArgs a,
                                                               const int4* __restrict__ tiles) {
  constexpr int HX = TX + 4, HY = TY + 4, NT = TX * TY;
  constexpr int NF = (HY * HX + NT - 1) / NT;
  constexpr int NR = (TY * HX + NT - 1) / NT;
  constexpr int NQ = (HY * TX + NT - 1) / NT;
  __shared__ double sW[5][HY][HX];
  __shared__ double sR[2][TY][HX];
  __shared__ double sQ[2][HY][TX];
  const int4 td = tiles[linear_tile(1)];
  const int bu = td.x, i0 = td.y, j0 = td.z;
  ConstBlockDev* B4 = (ConstBlockDev*)a.blocks + bu;
  const double rho = B4->rho, Gm = B4->G, Mm = B4->M, gam = B4->gamma;
  const double cl[5] = {1.0 / rho, 1.0 / rho, Mm, Gm, Mm};
  const int tid = threadIdx.y * TX + threadIdx.x;
  const int ld = a.ld, nx1 = a.nx - 1, ny1 = a.ny - 1;
  const long long ps = a.ps;
  // every load of the tile is issued before the LDS stores and the barrier;
  // every lane loads in every pass (past the end: the last element again),
  // so that no load sits under a lane condition (whose result the compiler
  // would copy out of the branch after waiting for all earlier loads)
  double vW[NF][6], vR[NR][2], vQ[NQ][2];
#pragma unroll
  for (int p = 0; p < NF; p++) {
    const int k = min(tid + p * NT, HY * HX - 1);
    const int r = k / HX, c = k - r * HX;
    const int gj = min(max(j0 - 2 + r, 0), ny1), gi = min(max(i0 - 2 + c, 0), nx1);
    const unsigned o = (unsigned)(gj * ld + gi) * 8u;
    vW[p][5] = ld_off(a.Ji, o);
#pragma unroll
    for (int l = 0; l < 5; l++) vW[p][l] = ld_off(a.min + l * ps, o);
  }
#pragma unroll
  for (int p = 0; p < NR; p++) {
    const int k = min(tid + p * NT, TY * HX - 1);
    const int r = k / HX, c = k - r * HX;
    const int gj = min(j0 + r, ny1), gi = min(max(i0 - 2 + c, 0), nx1);
    const unsigned o = (unsigned)(gj * ld + gi) * 8u;
    vR[p][0] = ld_off(a.xr, o);
    vR[p][1] = ld_off(a.yr, o);
  }
#pragma unroll
  for (int p = 0; p < NQ; p++) {
    const int k = min(tid + p * NT, HY * TX - 1);
    const int r = k / TX, c = k - r * TX;
    const int gj = min(max(j0 - 2 + r, 0), ny1), gi = min(i0 + c, nx1);
    const unsigned o = (unsigned)(gj * ld + gi) * 8u;
    vQ[p][0] = ld_off(a.xq, o);
    vQ[p][1] = ld_off(a.yq, o);
  }
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const bool valid = i <= B4->px - 6 && j <= B4->py - 6;
  const unsigned oc = (unsigned)(min(j, ny1) * ld + min(i, nx1)) * 8u;
  double lin[5], mown[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    lin[l] = ld_off(a.lin + l * ps, oc);
    mown[l] = ld_off(a.min + l * ps, oc);
  }
#pragma unroll
  for (int p = 0; p < NF; p++) {
    const int k = tid + p * NT;
    if (k < HY * HX) {
      const int r = k / HX, c = k - r * HX;
#pragma unroll
      for (int l = 0; l < 5; l++) sW[l][r][c] = vW[p][l] * vW[p][5] * cl[l];
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
  const int x = threadIdx.x, y = threadIdx.y;
  double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 1; s <= 2; s++) {
    const double D = (s == 1) ? FDM_DI1 : FDM_DI2;
    // along q: rows P-s (entry +DI(s)) and P+s (entry -DI(s))
    const double yr0 = sR[1][y][x + 2], xr0 = sR[0][y][x + 2];
    const double aym = 0.5 * (sR[1][y][x + 2 - s] + yr0) * D, ayp = 0.5 * (sR[1][y][x + 2 + s] + yr0) * D;
    const double axm = 0.5 * (sR[0][y][x + 2 - s] + xr0) * D, axp = 0.5 * (sR[0][y][x + 2 + s] + xr0) * D;
    const double yq0 = sQ[1][y + 2][x], xq0 = sQ[0][y + 2][x];
    const double bym = 0.5 * (sQ[1][y + 2 - s][x] + yq0) * D, byp = 0.5 * (sQ[1][y + 2 + s][x] + yq0) * D;
    const double bxm = 0.5 * (sQ[0][y + 2 - s][x] + xq0) * D, bxp = 0.5 * (sQ[0][y + 2 + s][x] + xq0) * D;
    double Qa[5], Qb[5], Ra[5], Rb[5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const double wqm = sW[l][y + 2][x + 2 - s], wqp = sW[l][y + 2][x + 2 + s];
      const double wrm = sW[l][y + 2 - s][x + 2], wrp = sW[l][y + 2 + s][x + 2];
      Qa[l] = aym * wqm - ayp * wqp;
      Qb[l] = axm * wqm - axp * wqp;
      Ra[l] = bym * wrm - byp * wrp;
      Rb[l] = bxm * wrm - bxp * wrp;
    }
    // field couplings of rates_*_mode2 (rates.f90:411-444), transposed
    g[0] += Qa[2] - Qb[3] + gam * Qa[4] - Ra[2] + Rb[3] - gam * Ra[4];
    g[1] += -gam * Qb[2] + Qa[3] - Qb[4] + gam * Rb[2] - Ra[3] + Rb[4];
    g[2] += Qa[0] - Ra[0];
    g[3] += -Qb[0] + Qa[1] + Rb[0] - Ra[1];
    g[4] += -Qb[1] + Rb[1];
  }
#pragma unroll
  for (int l = 0; l < 5; l++) {
    const double lo = lin[l] + g[l];
    st_sel<NTS>(a.lout + l * ps, oc, lo);
    st_sel<NTS>(a.mout + l * ps, oc, a.A * mown[l] + a.Bdt_prev * lo);
  }
}

// ---- pointwise transposes ----

// transpose of xy2nt: lambda of the xy fields from lambda of (vt, vn, stt,
// snt, snn, szz)
__device__ __forceinline__ void xy2nt_T(const NT& l, double n1, double n2, double* lF) {
  lF[0] = n2 * l.vt + n1 * l.vn;
  lF[1] = -n1 * l.vt + n2 * l.vn;
  lF[2] = n2 * n2 * l.stt + n1 * n2 * l.snt + n1 * n1 * l.snn;
  lF[3] = -2.0 * n1 * n2 * l.stt + (n2 * n2 - n1 * n1) * l.snt + 2.0 * n1 * n2 * l.snn;
  lF[4] = n1 * n1 * l.stt - n1 * n2 * l.snt + n2 * n2 * l.snn;
  lF[5] = l.szz;
}

// transpose of nt2xy
__device__ __forceinline__ NT nt2xy_T(const double* lF, double n1, double n2) {
  NT l;
  l.vt = n2 * lF[0] - n1 * lF[1];
  l.vn = n1 * lF[0] + n2 * lF[1];
  l.stt = n2 * n2 * lF[2] - n1 * n2 * lF[3] + n1 * n1 * lF[4];
  l.snt = 2.0 * n1 * n2 * lF[2] + (n2 * n2 - n1 * n1) * lF[3] - 2.0 * n1 * n2 * lF[4];
  l.snn = n1 * n1 * lF[2] + n1 * n2 * lF[3] + n2 * n2 * lF[4];
  l.szz = lF[5];
  return l;
}

// SAT_term / split_sp (time_step.f90:813-851) transposed: mu (six fields,
// szz zero) -> lambda of Fhat - F at the node
__device__ __forceinline__ void sat_T(const double* mu, const SidesDev& sd, int so, double* ldd) {
  const int NS = sd.NS;
  const double n1 = sd.normal[so], n2 = sd.normal[NS + so];
  const double Ks = sd.M[3 * NS + so], Kp = sd.M[4 * NS + so];
  double ls[6], lp[6];
#pragma unroll
  for (int l = 0; l < 6; l++) {
    ls[l] = Ks * mu[l];
    lp[l] = Kp * mu[l];
  }
  const NT a = nt2xy_T(ls, n1, n2), b = nt2xy_T(lp, n1, n2);
  NT l;
  l.vt = a.vt;
  l.snt = a.snt;
  l.vn = b.vn;
  l.stt = b.stt;
  l.snn = b.snn;
  l.szz = b.szz;
  xy2nt_T(l, n1, n2, ldd);
}

// linear part of set_bc_mode2 (boundaries.f90:216-377) per wave type:
// v = cvv fv + cvs Zi fs, s = csv Z fv + css fs
__device__ __forceinline__ void bc_coef(int bc, double& cvv, double& cvs, double& csv, double& css) {
  cvv = cvs = csv = css = 0.0;
  switch (bc) {
    case FDM_BC_MMS_SIN_W:
    case FDM_BC_ABSORBING:
    case FDM_BC_ABSORBING_0:
      cvv = 0.5; cvs = -0.5; csv = -0.5; css = 0.5;
      break;
    case FDM_BC_FREE:
    case FDM_BC_FREE_0:
      cvv = 1.0; cvs = -1.0;
      break;
    case FDM_BC_RIGID:
    case FDM_BC_RIGID_0:
      csv = -1.0; css = 1.0;
      break;
    case FDM_BC_ABSORBING_VELOCITY:
      cvs = -1.0; css = 1.0;
      break;
    case FDM_BC_FREE_VELOCITY:
      cvv = 1.0; cvs = -1.0; css = 1.0;
      break;
    default:
      break;
  }
}

// reverse of the hat assembly shared by locked_interface_mode2 and
// frictional_interface_mode2 (interfaces.f90:715-772, 960-1005): the hats
// of both sides from the rotated fields p, m and the common normal stress X
// and shear stress Y on the interface
__device__ __forceinline__ void hats_T(const NT& lp_h, const NT& lm_h, double Zsip, double Zsim,
                                       double Zpip, double Zpim, double gp, double gm, NT& lp,
                                       NT& lm, double& lX, double& lY) {
  lp.vt += lp_h.vt;
  lp.snt += Zsip * lp_h.vt;
  lY -= Zsip * lp_h.vt;
  lm.vt += lm_h.vt;
  lm.snt -= Zsim * lm_h.vt;
  lY += Zsim * lm_h.vt;
  lY += lp_h.snt + lm_h.snt;
  lp.vn += lp_h.vn;
  lp.snn += Zpip * lp_h.vn;
  lX -= Zpip * lp_h.vn;
  lm.vn += lm_h.vn;
  lm.snn -= Zpim * lm_h.vn;
  lX += Zpim * lm_h.vn;
  lp.stt += lp_h.stt;
  lp.snn -= gp * lp_h.stt;
  lX += gp * lp_h.stt;
  lm.stt += lm_h.stt;
  lm.snn -= gm * lm_h.stt;
  lX += gm * lm_h.stt;
  lp.szz += lp_h.szz;
  lp.snn -= gp * lp_h.szz;
  lX += gp * lp_h.szz;
  lm.szz += lm_h.szz;
  lm.snn -= gm * lm_h.szz;
  lX += gm * lm_h.szz;
  lX += lp_h.snn + lm_h.snn;
}

// partial derivatives of the state rate G(V, Psi; a, b) (friction.f90:
// 1202-1210, 1238-1243, 1254-1258) for the regularised laws
__device__ __forceinline__ void state_rate_grad(int law, const double* rs, double V, double Psi,
                                                double& GV, double& GPsi, double& Ga,
                                                double& Gb) {
  GV = GPsi = Ga = Gb = 0.0;
  if (V == 0.0) return;
  const double a = rs[0], b = rs[1], V0 = rs[2], f0 = rs[3], L = rs[4], fw = rs[5], Vw = rs[6];
  const double u = fabs(V), sg = V > 0.0 ? 1.0 : -1.0;
  if (law == FDM_LAW_RAL) {
    const double e = exp((f0 - Psi) / b);
    GV = -(b / L) * sg;
    GPsi = -(V0 / L) * e;
    Gb = (V0 / L) * (e - u / V0) - (V0 / L) * e * (f0 - Psi) / b;
    return;
  }
  const double lnu = log(u / V0);
  const double z = 0.5 * u / V0 * exp(Psi / a), sq = sqrt(1.0 + z * z);
  const double frs = a * arcsinh_ref(z);
  const double frs_u = a * (z / u) / sq, frs_Psi = z / sq, frs_a = frs / a - z * Psi / (a * sq);
  double fss, fss_u, fss_a, fss_b;
  const double fLV = f0 - (b - a) * lnu;
  if (law == FDM_LAW_RFL) {
    const double q = u / Vw, q8 = pow8(q), Dn = pow(1.0 + q8, 0.125);
    const double Dn_u = Dn * (q8 / u) / (1.0 + q8);
    fss = fw + (fLV - fw) / Dn;
    fss_u = (-(b - a) / u) / Dn - (fLV - fw) * Dn_u / (Dn * Dn);
    fss_a = lnu / Dn;
    fss_b = -lnu / Dn;
  } else {  // RSL
    fss = fLV;
    fss_u = -(b - a) / u;
    fss_a = lnu;
    fss_b = -lnu;
  }
  GV = ((fss - frs) / L + (u / L) * (fss_u - frs_u)) * sg;
  GPsi = -(u / L) * frs_Psi;
  Ga = (u / L) * (fss_a - frs_a);
  Gb = (u / L) * fss_b;
}

struct AdjEdgeArgs {
  const double* mu;  // mu of this stage (5 planes)
  int ld;
  long long ps;
  const int2* tasks;
  int ntasks;
  const BlockDev* blocks;
  SidesDev sd;
  IfacesDev fi;
  const double* trace;  // [3][NI] V, Psi, N of this stage
  double* lpsi;         // [NI]
  double* ldpsi;        // [NI]
  double* grad;         // [FDM_GRAD_NFIELDS][NI]
  double* gE;           // [5][NS]
  double A, Bdt;
};

__device__ __forceinline__ void load_mu(const AdjEdgeArgs& a, long long cc, double* mu) {
#pragma unroll
  for (int l = 0; l < 5; l++) mu[l] = a.mu[l * a.ps + cc];
  mu[5] = 0.0;
}

__device__ __forceinline__ void store_gE(const AdjEdgeArgs& a, int so, const double* lF,
                                         const double* ldd) {
#pragma unroll
  for (int l = 0; l < 5; l++) a.gE[l * a.sd.NS + so] = lF[l] - ldd[l];
}

__global__ void __launch_bounds__(64) adj_edge_kernel(AdjEdgeArgs a) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= a.ntasks) return;
  const int2 task = a.tasks[tid];
  const int NS = a.sd.NS;
  if (task.x >= 0) {
    // boundary node: SAT and set_bc_mode2 transposed
    const int b = task.x >> 2, s = task.x & 3, k = task.y;
    const BlockDev& B = a.blocks[b];
    int i, j;
    if (s == FDM_SIDE_L) { i = B.mx; j = B.my + k; }
    else if (s == FDM_SIDE_R) { i = B.px; j = B.my + k; }
    else if (s == FDM_SIDE_B) { i = B.mx + k; j = B.my; }
    else { i = B.mx + k; j = B.py; }
    const int so = B.side_off[s] + k;
    double mu[6], ldd[6], lF[6];
    load_mu(a, (long long)j * a.ld + i, mu);
    sat_T(mu, a.sd, so, ldd);
    const double n1 = a.sd.normal[so], n2 = a.sd.normal[NS + so];
    const double Zs = a.sd.M[so], Zp = a.sd.M[NS + so], gamma = a.sd.M[2 * NS + so];
    const double Zsi = (Zs == 0.0) ? 0.0 : 1.0 / Zs, Zpi = 1.0 / Zp;
    const NT lh = nt2xy_T(ldd, n1, n2);
    NT lf;
    lf.stt = lh.stt;
    lf.szz = lh.szz;
    const double lsnn = lh.snn + gamma * (lh.stt + lh.szz);
    double cvv, cvs, csv, css;
    bc_coef(B.side_bc[s], cvv, cvs, csv, css);
    lf.vt = cvv * lh.vt + csv * Zs * lh.snt;
    lf.snt = cvs * Zsi * lh.vt + css * lh.snt;
    lf.vn = cvv * lh.vn + csv * Zp * lsnn;
    lf.snn = cvs * Zpi * lh.vn + css * lsnn - gamma * (lh.stt + lh.szz);
    xy2nt_T(lf, n1, n2, lF);
    store_gE(a, so, lF, ldd);
    return;
  }
  // interface node
  const IfaceDev& I = a.fi.tab[-task.x - 1];
  const int k = task.y, g = I.m + k, q = I.node_off + k;
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
  const int som = Bm.side_off[sm] + km, sop = Bp.side_off[sp] + kp;
  double mum[6], mup[6], lddm[6], lddp[6];
  load_mu(a, (long long)jm * a.ld + im, mum);
  load_mu(a, (long long)jp * a.ld + ip, mup);
  sat_T(mum, a.sd, som, lddm);
  sat_T(mup, a.sd, sop, lddp);
  const double n1 = a.fi.nhat[q], n2 = a.fi.nhat[NI + q];
  const double Zsm = a.sd.M[som], Zsp = a.sd.M[sop];
  const double Zpm = a.sd.M[NS + som], Zpp = a.sd.M[NS + sop];
  const double gm = a.sd.M[2 * NS + som], gp = a.sd.M[2 * NS + sop];
  const double Zsip = 1.0 / Zsp, Zsim = 1.0 / Zsm, Zpip = 1.0 / Zpp, Zpim = 1.0 / Zpm;
  const double etas = 1.0 / (Zsip + Zsim), etap = 1.0 / (Zpip + Zpim);
  const NT lhp = nt2xy_T(lddp, n1, n2), lhm = nt2xy_T(lddm, n1, n2);
  NT lp{0, 0, 0, 0, 0, 0}, lm{0, 0, 0, 0, 0, 0};
  double lX = 0.0, lY = 0.0, lphis = 0.0, lphip = 0.0;
  hats_T(lhp, lhm, Zsip, Zsim, Zpip, Zpim, gp, gm, lp, lm, lX, lY);
  if (I.coupling == FDM_COUPLING_LOCKED) {
    lphip = lX;
    lphis = lY;
  } else {
    // frictional node: S = Y + fS0, N from X; V from the friction solve
    double rs[7];
#pragma unroll
    for (int c = 0; c < 7; c++) rs[c] = a.fi.rs[(long long)c * NI + q];
    const double V = a.trace[q], Psi = a.trace[NI + q], N = a.trace[2 * NI + q];
    const double skm = 1.0 - I.skempton;
    double* gr = a.grad;
    // state variable: nu = lambda_DPsi + b lambda_Psi, DPsi <- A DPsi + G
    const double lpsi_next = a.lpsi[q];
    const double nu = a.ldpsi[q] + a.Bdt * lpsi_next;
    double GV, GPsi, Ga, Gb;
    state_rate_grad(I.law, rs, V, Psi, GV, GPsi, Ga, Gb);
    double lPsi = lpsi_next + GPsi * nu;
    double la = Ga * nu, lb = Gb * nu;
    double lV = GV * nu;
    // snn = -(N - fN0)/skm, snt = S - fS0
    double lN = -lX / skm;
    double lS = lY, lfS0 = -lY;
    double lSlock = 0.0;
    if (N <= 0.0) {
      // tensile branch: S = 0, V = Slock / eta (friction.f90:675-683)
      lSlock += lV / etas;
    } else {
      // S = Slock - eta V; R(V) = -(Slock - eta V) + a N asinh(O V) = 0
      lSlock += lS;
      lV -= etas * lS;
      const double aa = rs[0], V0 = rs[2];
      const double O = 0.5 / V0 * exp(Psi / aa), z = O * V, sq = sqrt(1.0 + z * z);
      const double RV = etas + aa * N * O / sq;
      const double ash = arcsinh_ref(z);
      const double RN = aa * ash, RPsi = N * z / sq, Ra = N * ash - N * z * Psi / (aa * sq);
      const double t = lV / RV;
      lSlock += t;
      lN -= t * RN;
      lPsi -= t * RPsi;
      la -= t * Ra;
    }
    // N = fN0 - skm phip, Slock = fS0 + phis
    lphip = -skm * lN;
    lfS0 += lSlock;
    lphis = lSlock;
    a.lpsi[q] = lPsi;
    a.ldpsi[q] = a.A * nu;
    gr[FDM_GRAD_S0 * NI + q] += lfS0;
    gr[FDM_GRAD_A * NI + q] += la;
    gr[FDM_GRAD_B * NI + q] += lb;
  }
  // phis = etas (p.snt Zsip + m.snt Zsim + p.vt - m.vt), phip likewise
  lp.snt += etas * Zsip * lphis;
  lm.snt += etas * Zsim * lphis;
  lp.vt += etas * lphis;
  lm.vt -= etas * lphis;
  lp.snn += etap * Zpip * lphip;
  lm.snn += etap * Zpim * lphip;
  lp.vn += etap * lphip;
  lm.vn -= etap * lphip;
  double lFp[6], lFm[6];
  xy2nt_T(lp, n1, n2, lFp);
  xy2nt_T(lm, n1, n2, lFm);
  store_gE(a, sop, lFp, lddp);
  store_gE(a, som, lFm, lddm);
}

// lambda_F += g and mu += b g at the receivers: one lane per distinct grid
// point, adding the receivers of that point (first[u] .. first[u+1]-1 of
// the receivers ordered by point) in order
__global__ void __launch_bounds__(64) adj_inject_kernel(double* lF, double* mu, long long ps,
                                                        const long long* cc, const int* first,
                                                        const int* order, int nu, const double* g,
                                                        double b) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nu) return;
  const long long k = cc[order[first[u]]];
  double gx = 0.0, gy = 0.0;
  for (int q = first[u]; q < first[u + 1]; q++) {
    const int r = order[q];
    gx += g[2 * r];
    gy += g[2 * r + 1];
  }
  lF[k] += gx;
  lF[ps + k] += gy;
  mu[k] += b * gx;
  mu[ps + k] += b * gy;
}

}  // namespace fdm

struct fdm_adjoint {
  fdm_ctx* c = nullptr;
  int nrec = 0, nsteps = 0, interval = 0, nseg = 0;
  long long* rec_cc = nullptr;   // [nrec] plane offsets of the receivers
  int* rec_first = nullptr;      // [nuniq + 1] receivers grouped by grid point
  int* rec_order = nullptr;      // [nrec]
  int nuniq = 0;
  double* rec = nullptr;         // [nsteps][nrec][2]
  double* rec_bar = nullptr;     // [nsteps][nrec][2]
  double* ckpt = nullptr;        // [nseg][6 ps + 18 NI]
  double* trace = nullptr;       // [interval][nstage][3][NI]
  bool trace_valid = false;      // trace holds the last interval of the latest forward pass
  bool have_forward = false;
  double t = 0.0, dt = 0.0;
  double* planes = nullptr;      // lambda_F x 2, mu x 2 (5 planes each, stride ps)
  double* lF[2] = {nullptr, nullptr};
  double* mu[2] = {nullptr, nullptr};
  double* gE = nullptr;          // [5][NS]
  double* lpsi = nullptr;        // [NI]
  double* ldpsi = nullptr;       // [NI]
  double* grad = nullptr;        // [FDM_GRAD_NFIELDS][NI]
  // bulk adjoint: interior tiles (64 x 4) of [mx+6, px-6] x [my+6, py-6]
  // on the main stream, the other points on the side stream after the edge
  // adjoint; or one per-point kernel (FDM_ADJ_BULK=generic)
  bool tiled = true;
  bool nt_store = false;  // adjoint planes and metrics beyond the last-level cache
  int4* tiles = nullptr;
  int ntiles = 0;
  int2* band = nullptr;
  int nband = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double ms_fwd = 0.0, ms_rec = 0.0, ms_sweep = 0.0;
  std::vector<void*> allocs;
  std::string err;
};

namespace {

int adj_err(fdm_adjoint* a, int rc, const std::string& msg) {
  if (a) a->err = msg;
  else g_create_error = msg;
  return rc;
}

#define ADJCHK(adj, call)                                                              \
  do {                                                                                 \
    hipError_t _st = (call);                                                           \
    if (_st != hipSuccess)                                                             \
      return adj_err(adj, FDM_EDEVICE, std::string(#call ": ") + hipGetErrorString(_st)); \
  } while (0)

template <typename P>
int adj_alloc(fdm_adjoint* a, P** p, size_t bytes) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(bytes, 64)) != hipSuccess)
    return adj_err(a, FDM_ENOMEM, "fdm_adjoint: device allocation of " + std::to_string(bytes) + " bytes failed");
  a->allocs.push_back(q);
  *p = (P*)q;
  return FDM_OK;
}

size_t ckpt_elems(const fdm_ctx* c) { return (size_t)6 * c->ps + (size_t)FDM_IF_NFIELDS * c->NI; }

int adj_checkpoint(fdm_adjoint* a, int seg, bool save) {
  fdm_ctx* c = a->c;
  double* base = a->ckpt + (size_t)seg * ckpt_elems(c);
  double* F = (double*)c->F[c->cur];
  const size_t fb = (size_t)6 * c->ps * 8, sb = (size_t)FDM_IF_NFIELDS * c->NI * 8;
  if (save) {
    ADJCHK(a, hipMemcpyAsync(base, F, fb, hipMemcpyDeviceToDevice, c->stream));
    if (sb) ADJCHK(a, hipMemcpyAsync(base + 6 * c->ps, c->ist, sb, hipMemcpyDeviceToDevice, c->stream));
  } else {
    ADJCHK(a, hipMemcpyAsync(F, base, fb, hipMemcpyDeviceToDevice, c->stream));
    if (sb) ADJCHK(a, hipMemcpyAsync(c->ist, base + 6 * c->ps, sb, hipMemcpyDeviceToDevice, c->stream));
    c->df6_stale = false;
  }
  return FDM_OK;
}

// forward steps n0..n1 (1-based), one fdm_step call per checkpoint
// interval: checkpoints before the first step of every interval when
// save_ckpt, traces of the steps in the trace buffer (slot (n-1) mod
// interval) when record, receivers after every step when receivers
int adj_run_steps(fdm_adjoint* a, int n0, int n1, bool record, bool save_ckpt, bool receivers) {
  fdm_ctx* c = a->c;
  const int K = a->interval;
  int rc = FDM_OK;
  for (int s0 = n0; s0 <= n1 && !rc; ) {
    const int seg = (s0 - 1) / K, s1 = std::min(n1, (seg + 1) * K);
    if (save_ckpt && (s0 - 1) % K == 0) rc = adj_checkpoint(a, seg, true);
    if (rc) break;
    c->trace = record ? a->trace : nullptr;
    c->trace_n0 = (long long)seg * K + 1;
    if (receivers && a->nrec) {
      c->rec_cc = a->rec_cc;
      c->nrec = a->nrec;
      c->rec_out = a->rec;
      c->rec_n0 = 1;
    }
    rc = step_impl<double>(c, a->t, a->dt, s0, s1 - s0 + 1);
    c->trace = nullptr;
    c->rec_out = nullptr;
    c->nrec = 0;
    if (rc) return adj_err(a, rc, c->err);
    s0 = s1 + 1;
  }
  return rc;
}

// backward stage s (1-based) of step n
int adj_stage(fdm_adjoint* a, int n, int s, int& cur) {
  fdm_ctx* c = a->c;
  const int K = a->interval, ns = c->nstage;
  const double A = c->rkA[s - 1], Bdt = c->rkB[s - 1] * a->dt;
  const double Bprev = (s > 1 ? c->rkB[s - 2] : c->rkB[ns - 1]) * a->dt;
  AdjEdgeArgs e;
  if (c->ntasks) {
    e.mu = a->mu[cur];
    e.ld = c->ld;
    e.ps = c->ps;
    e.tasks = c->tasks;
    e.ntasks = c->ntasks;
    e.blocks = c->blocks;
    e.sd = SidesDev{c->sM, c->snormal, c->sF0, c->sx, c->sy, c->sFhat, c->NS};
    e.fi = IfacesDev{c->ifaces, c->inhat, c->irs, c->isw, c->iS0b, c->iN0b, c->ildS, c->ildN,
                     c->iout, c->iknx, c->ist, c->ist1, c->NI};
    e.trace = a->trace + ((size_t)((n - 1) % K) * ns + (s - 1)) * 3 * c->NI;
    e.lpsi = a->lpsi;
    e.ldpsi = a->ldpsi;
    e.grad = a->grad;
    e.gE = a->gE;
    e.A = A;
    e.Bdt = Bdt;
  }
  AdjBulkArgs b;
  b.lin = a->lF[cur];
  b.lout = a->lF[1 - cur];
  b.min = a->mu[cur];
  b.mout = a->mu[1 - cur];
  b.xq = (const double*)c->met[0];
  b.xr = (const double*)c->met[1];
  b.yq = (const double*)c->met[2];
  b.yr = (const double*)c->met[3];
  b.Ji = (const double*)c->met[4];
  b.colb = c->colb;
  b.rowb = c->rowb;
  b.blk = c->blk;
  b.nbx = c->nbx;
  b.blocks = c->blocks;
  b.gE = a->gE;
  b.NS = c->NS;
  b.nx = c->nx;
  b.ny = c->ny;
  b.ld = c->ld;
  b.ps = c->ps;
  b.A = A;
  b.Bdt_prev = Bprev;
  if (a->tiled) {
    // The interior points read mu of this stage within two points (the band
    // of the later stage wrote it there) and overwrite the buffer the later
    // stage's band read; the band reads mu within five points (the later
    // interior launch) and needs this stage's edge adjoint.
    ADJCHK(a, hipStreamWaitEvent(c->stream, c->evB, 0));
    if (a->ntiles) {
      if (a->nt_store)
        hipLaunchKernelGGL((adj_interior_kernel<64, 4, true>), dim3(a->ntiles), dim3(64, 4), 0, c->stream,
                           b, a->tiles);
      else
        hipLaunchKernelGGL((adj_interior_kernel<64, 4, false>), dim3(a->ntiles), dim3(64, 4), 0, c->stream,
                           b, a->tiles);
    }
    if (c->ntasks)
      hipLaunchKernelGGL(adj_edge_kernel, dim3((c->ntasks + 63) / 64), dim3(64), 0, c->side, e);
    ADJCHK(a, hipStreamWaitEvent(c->side, c->evI, 0));
    if (a->nband)
      hipLaunchKernelGGL(adj_band_kernel, dim3((2 * a->nband + 63) / 64), dim3(64), 0, c->side, b,
                         a->band, a->nband);
    ADJCHK(a, hipEventRecord(c->evB, c->side));
    ADJCHK(a, hipEventRecord(c->evI, c->stream));
  } else {
    if (c->ntasks)
      hipLaunchKernelGGL(adj_edge_kernel, dim3((c->ntasks + 63) / 64), dim3(64), 0, c->stream, e);
    hipLaunchKernelGGL(adj_bulk_kernel, dim3((c->nx + 63) / 64, (c->ny + 3) / 4), dim3(64, 4), 0,
                       c->stream, b);
  }
  ADJCHK(a, hipGetLastError());
  cur = 1 - cur;
  return FDM_OK;
}

}  // namespace

extern "C" {

const char* fdm_adjoint_last_error(const fdm_adjoint* a) { return a ? a->err.c_str() : g_create_error.c_str(); }

void fdm_adjoint_destroy(fdm_adjoint* a) {
  if (!a) return;
  if (a->c && a->c->stream) (void)hipStreamSynchronize(a->c->stream);
  for (hipEvent_t e : a->ev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : a->allocs) (void)hipFree(p);
  delete a;
}

int fdm_adjoint_create(fdm_ctx* c, int32_t nrec, const int32_t* rec_i, const int32_t* rec_j,
                       int32_t nsteps, int32_t interval, fdm_adjoint** out) {
  if (!c || !out || nsteps < 1 || nrec < 0 || (nrec && (!rec_i || !rec_j)))
    return adj_err(nullptr, FDM_EINVAL, "fdm_adjoint_create: invalid arguments");
  *out = nullptr;
  if (c->precision != FDM_FP64 || c->any_plastic || c->het || c->diss || c->pml || c->mode != 2 ||
      c->sbp != 4)
    return adj_err(nullptr, FDM_EINVAL,
                   "fdm_adjoint_create: needs an FP64 plane-strain SBP4 context without plasticity, "
                   "heterogeneous medium, artificial dissipation or PML");
  for (const BlockDev& B : c->hblocks)
    if (B.px - B.mx + 1 < 12 || B.py - B.my + 1 < 12)
      return adj_err(nullptr, FDM_EINVAL, "fdm_adjoint_create: blocks must be at least 12 x 12 points");
  if (c->NI) {
    std::vector<int> outside(c->NI);
    if (hipMemcpy(outside.data(), c->iout, outside.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      return adj_err(nullptr, FDM_EDEVICE, "fdm_adjoint_create: reading the interface maps failed");
    for (const IfaceDev& I : c->hifaces) {
      if (I.coupling == FDM_COUPLING_LOCKED) continue;
      const bool law_ok = I.law == FDM_LAW_RSL || I.law == FDM_LAW_RFL || I.law == FDM_LAW_RAL;
      bool out_any = false;
      for (int k = 0; k < I.n; k++) out_any = out_any || outside[I.node_off + k];
      if (I.coupling != FDM_COUPLING_FRICTION || !law_ok || I.reg || out_any || I.force)
        return adj_err(nullptr, FDM_EINVAL,
                       "fdm_adjoint_create: interfaces must be locked or frictional with law RSL, RFL "
                       "or RAL, without normal-stress regularisation, locked region or kinematic "
                       "forcing");
    }
  }
  for (int r = 0; r < nrec; r++)
    if (rec_i[r] < 0 || rec_i[r] >= c->nx || rec_j[r] < 0 || rec_j[r] >= c->ny)
      return adj_err(nullptr, FDM_EINVAL, "fdm_adjoint_create: receiver outside the grid");
  fdm_adjoint* a = new fdm_adjoint();
  a->c = c;
  a->nrec = nrec;
  a->nsteps = nsteps;
  // interval 0: keep the whole fault trace when it fits the trace budget
  // (FDM_ADJ_TRACE_MB, default 2048 MiB), else square-root checkpointing
  const int isqrt = (int)std::ceil(std::sqrt((double)nsteps));
  if (interval > 0) {
    a->interval = std::min(interval, nsteps);
  } else if (interval < 0) {
    a->interval = isqrt;
  } else {
    double budget = 2048.0;
    if (const char* v = getenv("FDM_ADJ_TRACE_MB")) budget = atof(v);
    const double mb = (double)nsteps * c->nstage * 3.0 * std::max(1, c->NI) * 8.0 / (1024.0 * 1024.0);
    a->interval = mb <= budget ? nsteps : isqrt;
  }
  a->nseg = (nsteps + a->interval - 1) / a->interval;
  std::vector<long long> cc(std::max(1, nrec));
  for (int r = 0; r < nrec; r++) cc[r] = (long long)rec_j[r] * c->ld + rec_i[r];
  std::vector<int> order(std::max(1, nrec)), first;
  for (int r = 0; r < nrec; r++) order[r] = r;
  std::stable_sort(order.begin(), order.begin() + nrec, [&](int x, int y) { return cc[x] < cc[y]; });
  for (int q = 0; q < nrec; q++)
    if (q == 0 || cc[order[q]] != cc[order[q - 1]]) first.push_back(q);
  a->nuniq = (int)first.size();
  first.push_back(nrec);
  const size_t NI = std::max(1, c->NI), NS = std::max(1, c->NS);
  int rc = adj_alloc(a, &a->rec_cc, cc.size() * sizeof(long long));
  if (!rc) rc = adj_alloc(a, &a->rec_first, first.size() * sizeof(int));
  if (!rc) rc = adj_alloc(a, &a->rec_order, order.size() * sizeof(int));
  if (!rc && (hipMemcpy(a->rec_cc, cc.data(), cc.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(a->rec_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(a->rec_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess))
    rc = adj_err(a, FDM_EDEVICE, "fdm_adjoint_create: upload failed");
  const size_t nrv = (size_t)nsteps * std::max(1, nrec) * 2;
  if (!rc) rc = adj_alloc(a, &a->rec, nrv * 8);
  if (!rc) rc = adj_alloc(a, &a->rec_bar, nrv * 8);
  if (!rc) rc = adj_alloc(a, &a->ckpt, (size_t)a->nseg * ckpt_elems(c) * 8);
  if (!rc) rc = adj_alloc(a, &a->trace, (size_t)a->interval * c->nstage * 3 * NI * 8);
  if (!rc) rc = adj_alloc(a, &a->planes, ((size_t)20 * c->ps + c->coff + 64) * 8);
  if (!rc) rc = adj_alloc(a, &a->gE, (size_t)5 * NS * 8);
  if (!rc) rc = adj_alloc(a, &a->lpsi, NI * 8);
  if (!rc) rc = adj_alloc(a, &a->ldpsi, NI * 8);
  if (!rc) rc = adj_alloc(a, &a->grad, (size_t)FDM_GRAD_NFIELDS * NI * 8);
  for (int k = 0; k < 4 && !rc; k++)
    if (hipEventCreate(&a->ev[k]) != hipSuccess) rc = adj_err(a, FDM_EDEVICE, "fdm_adjoint_create: event creation failed");
  if (rc) {
    g_create_error = a->err;
    fdm_adjoint_destroy(a);
    return rc;
  }
  // bulk adjoint work lists; per-plane byte offsets of the tile kernel are
  // 32-bit
  if (const char* v = getenv("FDM_ADJ_BULK")) a->tiled = strcmp(v, "generic") != 0;
  a->nt_store = nt_store_choice(25.0 * (double)c->ps * 8.0);
  if ((size_t)c->ps * 8 >= ((size_t)1 << 31)) a->tiled = false;
  if (a->tiled) {
    std::vector<int4> tl;
    std::vector<int2> band;
    for (int b = 0; b < c->nb; b++) {
      const BlockDev& B = c->hblocks[b];
      for (int j0 = B.my + 6; j0 <= B.py - 6; j0 += 4)
        for (int i0 = B.mx + 6; i0 <= B.px - 6; i0 += 64) tl.push_back(make_int4(b, i0, j0, 0));
      for (int j = B.my; j <= B.py; j++)
        for (int i = B.mx; i <= B.px; i++)
          if (i - B.mx < 6 || B.px - i < 6 || j - B.my < 6 || B.py - j < 6) band.push_back(make_int2(i, j));
    }
    a->ntiles = (int)tl.size();
    a->nband = (int)band.size();
    rc = adj_alloc(a, &a->tiles, tl.size() * sizeof(int4));
    if (!rc) rc = adj_alloc(a, &a->band, band.size() * sizeof(int2));
    if (!rc && a->ntiles &&
        hipMemcpy(a->tiles, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
      rc = adj_err(a, FDM_EDEVICE, "fdm_adjoint_create: upload failed");
    if (!rc && a->nband &&
        hipMemcpy(a->band, band.data(), band.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
      rc = adj_err(a, FDM_EDEVICE, "fdm_adjoint_create: upload failed");
    if (rc) {
      g_create_error = a->err;
      fdm_adjoint_destroy(a);
      return rc;
    }
  }
  double* pb = a->planes + c->coff;
  a->lF[0] = pb;
  a->lF[1] = pb + 5 * c->ps;
  a->mu[0] = pb + 10 * c->ps;
  a->mu[1] = pb + 15 * c->ps;
  *out = a;
  return FDM_OK;
}

int fdm_adjoint_info(const fdm_adjoint* a, int32_t* nodes, int32_t* interval) {
  if (!a) return FDM_EINVAL;
  if (nodes) *nodes = a->c->NI;
  if (interval) *interval = a->interval;
  return FDM_OK;
}

int fdm_adjoint_set_param(fdm_adjoint* a, int32_t field, const double* values) {
  if (!a || !values) return FDM_EINVAL;
  fdm_ctx* c = a->c;
  double* dst = nullptr;
  if (field == FDM_GRAD_S0) dst = c->iS0b;
  else if (field == FDM_GRAD_A) dst = c->irs;
  else if (field == FDM_GRAD_B) dst = c->irs + c->NI;
  else return adj_err(a, FDM_EINVAL, "fdm_adjoint_set_param: field must be S0, a or b");
  if (!c->NI) return FDM_OK;
  ADJCHK(a, hipStreamSynchronize(c->stream));
  ADJCHK(a, hipMemcpy(dst, values, (size_t)c->NI * 8, hipMemcpyHostToDevice));
  a->have_forward = false;
  return FDM_OK;
}

int fdm_adjoint_forward(fdm_adjoint* a, double t, double dt, double* rec) {
  if (!a) return FDM_EINVAL;
  fdm_ctx* c = a->c;
  if (c->prof) return adj_err(a, FDM_EINVAL, "fdm_adjoint_forward: disable profiling");
  a->t = t;
  a->dt = dt;
  const int last0 = (a->nseg - 1) * a->interval + 1;
  ADJCHK(a, hipEventRecord(a->ev[0], c->stream));
  int rc = adj_run_steps(a, 1, last0 - 1, false, true, true);
  if (!rc) rc = adj_run_steps(a, last0, a->nsteps, true, true, true);
  if (rc) return rc;
  ADJCHK(a, hipEventRecord(a->ev[1], c->stream));
  ADJCHK(a, hipEventSynchronize(a->ev[1]));
  float ms = 0.0f;
  ADJCHK(a, hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
  a->ms_fwd = ms;
  rc = check_fail(c);
  if (rc) return adj_err(a, rc, c->err);
  if (rec && a->nrec)
    ADJCHK(a, hipMemcpy(rec, a->rec, (size_t)a->nsteps * a->nrec * 2 * 8, hipMemcpyDeviceToHost));
  a->trace_valid = true;
  a->have_forward = true;
  return FDM_OK;
}

int fdm_adjoint_backward(fdm_adjoint* a, const double* rec_bar, double* grad, double* gF) {
  if (!a || (a->nrec && !rec_bar)) return FDM_EINVAL;
  if (!a->have_forward) return adj_err(a, FDM_EINVAL, "fdm_adjoint_backward: no forward pass");
  fdm_ctx* c = a->c;
  const int K = a->interval, ns = c->nstage;
  if (a->nrec)
    ADJCHK(a, hipMemcpy(a->rec_bar, rec_bar, (size_t)a->nsteps * a->nrec * 2 * 8, hipMemcpyHostToDevice));
  ADJCHK(a, hipMemsetAsync(a->planes, 0, ((size_t)20 * c->ps + c->coff + 64) * 8, c->stream));
  ADJCHK(a, hipMemsetAsync(a->gE, 0, (size_t)5 * std::max(1, c->NS) * 8, c->stream));
  const size_t NI = std::max(1, c->NI);
  ADJCHK(a, hipMemsetAsync(a->lpsi, 0, NI * 8, c->stream));
  ADJCHK(a, hipMemsetAsync(a->ldpsi, 0, NI * 8, c->stream));
  ADJCHK(a, hipMemsetAsync(a->grad, 0, (size_t)FDM_GRAD_NFIELDS * NI * 8, c->stream));
  int cur = 0;
  const double b_last = c->rkB[ns - 1] * a->dt;
  float ms_rec = 0.0f, ms_sweep = 0.0f;
  for (int seg = a->nseg - 1; seg >= 0; seg--) {
    const int n0 = seg * K + 1, n1 = std::min(a->nsteps, n0 + K - 1);
    ADJCHK(a, hipEventRecord(a->ev[0], c->stream));
    if (!(seg == a->nseg - 1 && a->trace_valid)) {
      int rc = adj_checkpoint(a, seg, false);
      if (!rc) rc = adj_run_steps(a, n0, n1, true, false, false);
      if (rc) return rc;
    }
    ADJCHK(a, hipEventRecord(a->ev[1], c->stream));
    if (a->tiled) {
      const int rf = fork_side(c);
      if (rf) return adj_err(a, rf, c->err);
    }
    for (int n = n1; n >= n0; n--) {
      if (a->nrec) {
        // after both streams' last writes of the later step, before either
        // reads this step's last stage
        hipStream_t sr = c->stream;
        if (a->tiled) {
          ADJCHK(a, hipStreamWaitEvent(c->side, c->evI, 0));
          sr = c->side;
        }
        hipLaunchKernelGGL(adj_inject_kernel, dim3((a->nuniq + 63) / 64), dim3(64), 0, sr, a->lF[cur],
                           a->mu[cur], c->ps, a->rec_cc, a->rec_first, a->rec_order, a->nuniq,
                           a->rec_bar + (size_t)(n - 1) * a->nrec * 2, b_last);
        ADJCHK(a, hipGetLastError());
        if (a->tiled) ADJCHK(a, hipEventRecord(c->evB, c->side));
      }
      for (int s = ns; s >= 1; s--) {
        int rc = adj_stage(a, n, s, cur);
        if (rc) return rc;
      }
    }
    if (a->tiled) {
      const int rj = join_side(c);
      if (rj) return adj_err(a, rj, c->err);
    }
    ADJCHK(a, hipEventRecord(a->ev[2], c->stream));
    ADJCHK(a, hipEventSynchronize(a->ev[2]));
    float m1 = 0.0f, m2 = 0.0f;
    ADJCHK(a, hipEventElapsedTime(&m1, a->ev[0], a->ev[1]));
    ADJCHK(a, hipEventElapsedTime(&m2, a->ev[1], a->ev[2]));
    ms_rec += m1;
    ms_sweep += m2;
  }
  a->trace_valid = (a->nseg == 1);
  a->ms_rec = ms_rec;
  a->ms_sweep = ms_sweep;
  int rc = check_fail(c);
  if (rc) return adj_err(a, rc, c->err);
  if (c->NI) {
    ADJCHK(a, hipMemcpy(a->grad + (size_t)FDM_GRAD_PSI0 * c->NI, a->lpsi, (size_t)c->NI * 8,
                        hipMemcpyDeviceToDevice));
    if (grad)
      ADJCHK(a, hipMemcpy(grad, a->grad, (size_t)FDM_GRAD_NFIELDS * c->NI * 8, hipMemcpyDeviceToHost));
  }
  if (gF) {
    std::vector<double> h((size_t)5 * c->ps);
    ADJCHK(a, hipMemcpy(h.data(), a->lF[cur], h.size() * 8, hipMemcpyDeviceToHost));
    const size_t np = (size_t)c->nx * c->ny;
    for (int l = 0; l < 6; l++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++)
          gF[l * np + (size_t)j * c->nx + i] = l < 5 ? h[(size_t)l * c->ps + (size_t)j * c->ld + i] : 0.0;
  }
  return FDM_OK;
}

int fdm_adjoint_timing(const fdm_adjoint* a, double* fwd_ms, double* recompute_ms, double* sweep_ms) {
  if (!a) return FDM_EINVAL;
  if (fwd_ms) *fwd_ms = a->ms_fwd;
  if (recompute_ms) *recompute_ms = a->ms_rec;
  if (sweep_ms) *sweep_ms = a->ms_sweep;
  return FDM_OK;
}

}  // extern "C"
