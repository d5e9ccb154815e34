// fdm_api.hip -- C ABI (include/fdmap.h) of the MI355X time-stepping
// library: device-resident storage, kernel launches and host<->device
// transfers.  One HIP stream per context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "fdm_bulk.hpp"
#include "../../include/fdmap_ensemble.h"

namespace fdm {

// mms_sin (mms.f90:52-234) on the device; same expression order as the
// reference.  field: 0 vx,1 vy,2 sxx,3 sxy,4 syy,5 szz, 6..11 sources.
__device__ double mms_sin(double x, double y, double t, int side, int field) {
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
    case 5:
      return 0.0;
    case 0:
      return ny * vx - my * vy;
    case 1:
      return mx * vy - nx * vx;
    case 2:
      return syy * (my * my) + sxx * (ny * ny) - 2.0 * sxy * my * ny;
    case 4:
      return syy * (mx * mx) + sxx * (nx * nx) - 2.0 * sxy * mx * nx;
    case 3:
      return -syy * my * mx - sxx * ny * nx + sxy * (ny * mx + nx * my);
    case 6: {
      double vxt = ny * vx_t - my * vy_t;
      double sxxx = syy_x * (my * my) + sxx_x * (ny * ny) - 2.0 * sxy_x * my * ny +
                    2.0 * syy * my * my_x + 2.0 * sxx * ny * ny_x -
                    2.0 * sxy * (my_x * ny + my * ny_x);
      double sxyy = -syy_y * my * mx - sxx_y * ny * nx + sxy_y * (ny * mx + nx * my) -
                    syy * (my_y * mx + my * mx_y) - sxx * (ny_y * nx + ny * nx_y) +
                    sxy * (ny_y * mx + ny * mx_y + nx_y * my + nx * my_y);
      return vxt - (1.0 / rho) * (sxxx + sxyy);
    }
    case 7: {
      double vyt = mx * vy_t - nx * vx_t;
      double sxyx = -syy_x * my * mx - sxx_x * ny * nx + sxy_x * (ny * mx + nx * my) -
                    syy * (my_x * mx + my * mx_x) - sxx * (ny_x * nx + ny * nx_x) +
                    sxy * (ny_x * mx + ny * mx_x + nx_x * my + nx * my_x);
      double syyy = syy_y * (mx * mx) + sxx_y * (nx * nx) - 2.0 * sxy_y * mx * nx +
                    2.0 * syy * mx * mx_y + 2.0 * sxx * nx * nx_y -
                    2.0 * sxy * (mx_y * nx + mx * nx_y);
      return vyt - (1.0 / rho) * (sxyx + syyy);
    }
    case 8: {
      double sxxt = syy_t * (my * my) + sxx_t * (ny * ny) - 2.0 * sxy_t * my * ny;
      double vxx = ny_x * vx - my_x * vy + ny * vx_x - my * vy_x;
      double vyy = mx_y * vy - nx_y * vx + mx * vy_y - nx * vx_y;
      return sxxt - (lambda + 2.0 * G) * vxx - lambda * vyy;
    }
    case 10: {
      double vxx = ny_x * vx - my_x * vy + ny * vx_x - my * vy_x;
      double vyy = mx_y * vy - nx_y * vx + mx * vy_y - nx * vx_y;
      double syyt = syy_t * (mx * mx) + sxx_t * (nx * nx) - 2.0 * sxy_t * mx * nx;
      return syyt - lambda * vxx - (lambda + 2.0 * G) * vyy;
    }
    case 9: {
      double vxy = ny_y * vx - my_y * vy + ny * vx_y - my * vy_y;
      double vyx = mx_x * vy - nx_x * vx + mx * vy_x - nx * vx_x;
      double sxyt = -syy_t * my * mx - sxx_t * ny * nx + sxy_t * (ny * mx + nx * my);
      return sxyt - G * (vyx + vxy);
    }
    case 11: {
      double szzt = 0.0;
      double vxx = ny_x * vx - my_x * vy + ny * vx_x - my * vy_x;
      double vyy = mx_y * vy - nx_y * vx + mx * vy_y - nx * vx_y;
      return szzt - lambda * vxx - lambda * vyy;
    }
    default:
      return 0.0;
  }
}

}  // namespace fdm

using namespace fdm;

namespace {

std::string g_create_error;

// interior tile shapes (columns x rows) selectable with FDM_TILE
const int kTileShape[3][2] = {{64, 4}, {64, 8}, {128, 4}};

struct EventPair {
  hipEvent_t a, b;
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct fdm_ctx {
  int nx = 0, ny = 0, ld = 0, nb = 0, ni = 0, nbx = 0, nby = 0, nstage = 0;
  int coff = 0;      // column offset of the stored planes (see fdm_create)
  long long ps = 0;  // plane stride in elements
  int precision = FDM_FP64;
  size_t esz = 8;
  double rkA[FDM_MAX_STAGES], rkB[FDM_MAX_STAGES], rkC[FDM_MAX_STAGES + 1];
  hipStream_t stream = nullptr;
  // edge and band kernels run on a second stream (default priority) so that
  // they overlap the interior launches; evI / evB mark the latest interior /
  // band launch, evJ forks the side stream off the main one
  hipStream_t side = nullptr;
  hipEvent_t evI = nullptr, evB = nullptr, evJ = nullptr;
  int device = 0;
  // bulk storage (element type per precision)
  void* F[2] = {nullptr, nullptr};
  int cur = 0;
  void* DF = nullptr;
  void* met[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // xq xr yq yr Ji
  double* X = nullptr;
  double* Y = nullptr;
  int* colb = nullptr;
  int* rowb = nullptr;
  int* blk = nullptr;
  BlockDev* blocks = nullptr;
  std::vector<BlockDev> hblocks;
  std::vector<int> hcolb, hrowb, hblk;  // host copies of the block maps
  // sides
  int NS = 0;
  double* sM = nullptr;
  double* snormal = nullptr;
  double* sF0 = nullptr;
  double* sx = nullptr;
  double* sy = nullptr;
  double* sFhat = nullptr;
  // interfaces
  int NI = 0;
  std::vector<IfaceDev> hifaces;
  IfaceDev* ifaces = nullptr;
  double* inhat = nullptr;
  double* irs = nullptr;
  double* isw = nullptr;
  double* iS0b = nullptr;
  double* iN0b = nullptr;
  double* ildS = nullptr;
  double* ildN = nullptr;
  int* iout = nullptr;
  double* iknx = nullptr;  // [NI] distance along the fault (kinematic forcing)
  double* ist = nullptr;
  double* ist1 = nullptr;
  bool any_friction = false;
  std::vector<double> ld_T;  // per interface
  // edge tasks
  int ntasks = 0;
  int2* tasks = nullptr;
  // source
  fdm_source src{};
  SourcePlaceDev* dsrc = nullptr;
  // plasticity
  bool any_plastic = false;
  bool any_mms = false;
  double* gammap = nullptr;
  double* lambda = nullptr;
  int* fail = nullptr;
  // profiling
  bool prof = false;
  std::vector<EventPair> pool;
  size_t pool_used = 0;
  // kinds: 0 interior (or per-point) bulk, 1 edge, 2 band
  double kind_ms[3] = {0.0, 0.0, 0.0};
  long long kind_n[3] = {0, 0, 0};
  std::vector<int> pair_kind;
  std::string err;
  std::vector<void*> allocs;
  // launch shape of the bulk kernel (tile 64 x tile_y points)
  int tile_y = 4;
  int xcd_swizzle = 1;
  // lean interior launches store F and DF non-temporally (working set beyond
  // the last-level cache)
  bool nt_store = false;
  // heterogeneous medium: rho, G, M, gamma, nu planes (field precision)
  bool het = false;
  void* mat[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // artificial dissipation in some block: the Jacobian plane
  bool diss = false;
  void* Jp = nullptr;
  // kinematic forcing on some interface: the edge kernels carrying it
  bool any_force = false;
  // perfectly matched layer: some block has it in effect (pml); Wx / Wy
  // exist when some block sets pmlx / pmly (pml_wx, pml_wy); W holds Wx,
  // Wy, DWx, DWy, 6 planes each (field precision, plane stride ps)
  bool pml = false, pml_wx = false, pml_wy = false;
  void* W = nullptr;
  // the points of the layer blocks, row-major within each block: those four
  // or more from every edge of their block (lpts[0]) and the rest (lpts[1])
  int2* lpts[2] = {nullptr, nullptr};
  int nlpts[2] = {0, 0};
  long long nlayer_inner = 0;  // lpts[0] entries that are points (not padding)
  // 2: plane strain, 6 fields; 3: antiplane shear, 3 fields
  int mode = 2, nF = 6;
  // first-derivative operator SBP2/4/6 (fd_coeff.f90:41-248) and its
  // closure depth nbst: points nearer a block edge take the one-sided rows
  int sbp = 4, nbst = 4;
  std::vector<double> hnu;  // nu of the medium ([ny][nx]), for restore_df6
  // bulk kernel: 0 per-point, 2 interior tiles (kTileShape[tile_shape]) + band points
  int bulk_mode = 2;
  int tile_shape = 0;  // index into kTileShape
  // interior tiles (block, i0, j0, 0) for the tile shapes 64x4, 64x8, 128x4; band points (i, j)
  int4* tiles[3] = {nullptr, nullptr, nullptr};
  int ntiles[3] = {0, 0, 0};
  int2* band = nullptr;
  int nband = 0;
  // host copies of the tile, band and edge-task lists (ensembles merge them)
  std::vector<int4> htiles[3];
  std::vector<int2> hband, htasks;
  // the latest interior launch was a lean one, which leaves DF6 unwritten at
  // interior points
  bool df6_stale = false;
  // graph capture (ensembles): stage times come from the device step counter
  long long* nstep_dev = nullptr;
  bool graph_mode = false;
  double g_tinit = 0.0, g_dt = 0.0, g_rkc = 0.0;
  long long ninterior = 0;
  // band split (plane strain, SBP4, uniform medium, with or without
  // dissipation; FP64 by default): the band points on no block edge (iband) join the interior
  // launch; the second stream's band kernel takes the edge points (eband)
  bool band_split = false;
  // lean plane-strain band points in a uniform medium: one lane per (point,
  // field) (bulk_band_fl; FP64 by default, FDM_BAND_FL=0|1 overrides)
  bool band_fl = true;
  // antiplane band points without dissipation: one lane per (point, field)
  // (bulk_band3_fl) when FDM_BAND_FL3=1; off by default, since it lengthened
  // the C2-sized antiplane step (0.380 against 0.366 ms in FP64)
  bool band_fl3 = false;
  // plane strain in a medium or with dissipation: field lanes too
  // (bulk_band_fl<..., EXT>): by default only in FP32-bulk in a medium
  // without dissipation, the one case they shortened (FDM_BAND_FL_EXT
  // overrides)
  bool band_fl_ext = false;
  int2* iband = nullptr;
  int2* eband = nullptr;
  int niband = 0, neband = 0;
  // adjoint forward passes: fault trajectories ([nstage][3][NI] per step,
  // step n at trace + (n - trace_n0) nstage 3 NI), the slot of the stage
  // being launched, and vx, vy at receivers after every step (step n at
  // rec_out + (n - rec_n0) nrec 2)
  double* trace = nullptr;
  long long trace_n0 = 0;
  double* trace_cur = nullptr;
  const long long* rec_cc = nullptr;
  int nrec = 0;
  double* rec_out = nullptr;
  long long rec_n0 = 0;
  long long cur_n = 0;       // step being launched
  long long first_n = 0;     // first step of the current fdm_step call
  // the frictional interface nodes (indices into the [.][NI] interface
  // arrays) and the context's reduction buffer of fdm_max_slip_velocity:
  // vred[0] the result, vred[1..] one partial maximum per workgroup
  int* fnode = nullptr;
  int nfnode = 0;
  double* vred = nullptr;
  // the interior kernel forms Ji from the metrics (FP64, when the given Ji
  // equals 1/(xq yr - xr yq) everywhere; FDM_JI_CALC = 0 turns it off)
  bool ji_calc = false;
  // output streaming: a copy stream and the snapshot slots (device staging,
  // pinned host copy, the event that ends the host copy)
  hipStream_t copy = nullptr;
  struct Snapshot {
    void* dev = nullptr;
    void* host = nullptr;
    hipEvent_t staged = nullptr, done = nullptr;
    uint32_t mask = 0;
    bool busy = false;
  } snap[FDM_SNAPSHOT_SLOTS];
  // FP32-delta metrics for the interior tiles (fdm_create decides)
  bool mcomp = false;
  float* mfp[4] = {nullptr, nullptr, nullptr, nullptr};
  long long mc_escapes = 0;
};

// Non-temporal stores pay when a launch's working set (bytes of the planes
// it streams) exceeds MI355X's 256 MB Infinity Cache, and cost when it fits:
// FDM_NT_STORE = 0/1 forces the choice, FDM_NT_MB moves the threshold.
static inline bool nt_store_choice(double working_set_bytes) {
  if (const char* v = getenv("FDM_NT_STORE")) return atoi(v) != 0;
  double mb = 256.0;
  if (const char* v = getenv("FDM_NT_MB")) mb = atof(v);
  return working_set_bytes > mb * 1024.0 * 1024.0;
}

// block index of global point (i, j)
static inline int block_of(const fdm_ctx* c, int i, int j) {
  return c->hblk[(size_t)c->hrowb[j] * c->nbx + c->hcolb[i]];
}


namespace {

int set_err(fdm_ctx* c, int code, const std::string& msg) {
  if (c)
    c->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                            \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return set_err((ctx), FDM_EDEVICE,                                             \
                     std::string(#call) + ": " + hipGetErrorString(e_));             \
  } while (0)

template <typename P>
int dalloc(fdm_ctx* c, P** p, size_t bytes) {
  void* q = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(c, FDM_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  c->allocs.push_back(q);
  *p = (P*)q;
  return FDM_OK;
}

template <typename P>
int dupload(fdm_ctx* c, P** p, const void* host, size_t bytes) {
  int rc = dalloc(c, p, bytes);
  if (rc) return rc;
  if (host && bytes) HIPCHK(c, hipMemcpy(*p, host, bytes, hipMemcpyHostToDevice));
  return FDM_OK;
}

int side_len(const fdm_block& b, int s) { return s < 2 ? (b.py - b.my + 1) : (b.px - b.mx + 1); }

// copy a [ny][nx] double plane into a padded [ny][ld] plane of type T
template <typename T>
void pack_plane(std::vector<T>& out, const double* in, int nx, int ny, int ld) {
  out.assign((size_t)ld * ny, (T)0);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) out[(size_t)j * ld + i] = (T)in[(size_t)j * nx + i];
}

// source time function (source.f90:212-223), evaluated on the host
double stf_amplitude(const fdm_source& s, double t) {
  const double pi = 3.14159265358979323846264338327950288419716939937510;
  double tc = s.t_center, td = s.t_duration;
  switch (s.stf) {
    case FDM_STF_RAMP:
      return std::fmax(std::fmin((t - tc) / td, 1.0), 0.0);
    case FDM_STF_SINUSOID:
      return std::sin(2.0 * pi * (t - tc) / td);
    case FDM_STF_GAUSSIAN:
      return std::exp(-0.5 * ((t - tc) * (t - tc)) / (td * td));
    case FDM_STF_RICKER:
      return (1.0 - 2.0 * (pi * pi) * ((t - tc) * (t - tc)) / (td * td)) *
             std::exp(-(pi * pi) * ((t - tc) * (t - tc)) / (td * td));
    default:
      return 0.0;
  }
}

// load_ramp (friction.f90:1455-1484), evaluated on the host
double load_ramp(double t, double T) {
  if (T <= 0.0) return (t < 0.0) ? 0.0 : 1.0;
  return (t < 0.0) ? 0.0 : 1.0 - std::exp(-t / T);
}

EventPair* prof_begin(fdm_ctx* c, int kind, hipStream_t s) {
  if (!c->prof) return nullptr;
  if (c->pool_used == c->pool.size()) {
    if (c->pool.size() < 65536) {
      EventPair ep;
      // device-scope marks: a system-scope release at the closing mark
      // would write back L2 before its timestamp (C2 FP64 interior
      // launch: 101.2 µs with it, 97.6 without, 96.9 by the kernel's own
      // dispatch timestamps)
      if (hipEventCreateWithFlags(&ep.a, hipEventDisableSystemFence) != hipSuccess ||
          hipEventCreateWithFlags(&ep.b, hipEventDisableSystemFence) != hipSuccess)
        return nullptr;
      c->pool.push_back(ep);
      c->pair_kind.push_back(kind);
    } else {
      // drain: accumulate and recycle
      (void)hipStreamSynchronize(c->stream);
      if (c->side) (void)hipStreamSynchronize(c->side);
      for (size_t k = 0; k < c->pool_used; k++) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->pool[k].a, c->pool[k].b);
        c->kind_ms[c->pair_kind[k]] += ms;
      }
      c->pool_used = 0;
    }
  }
  EventPair* ep = &c->pool[c->pool_used];
  c->pair_kind[c->pool_used] = kind;
  c->pool_used++;
  (void)hipEventRecord(ep->a, s);
  return ep;
}

void prof_end(fdm_ctx* c, EventPair* ep, int kind, hipStream_t s) {
  if (!ep) return;
  (void)hipEventRecord(ep->b, s);
  c->kind_n[kind]++;
}

template <typename T>
EdgeArgs<T> edge_args(fdm_ctx* c, double t, double Bdt, double A, int do_scale, int do_update,
                      int snapshot, double Bload) {
  EdgeArgs<T> a;
  a.F = (const T*)c->F[c->cur];
  a.ld = c->ld;
  a.ps = c->ps;
  a.tasks = c->tasks;
  a.ntasks = c->ntasks;
  a.blocks = c->blocks;
  a.sd = SidesDev{c->sM, c->snormal, c->sF0, c->sx, c->sy, c->sFhat, c->NS};
  a.fi = IfacesDev{c->ifaces, c->inhat, c->irs, c->isw, c->iS0b, c->iN0b, c->ildS, c->ildN,
                   c->iout, c->iknx, c->ist, c->ist1, c->NI};
  a.t = t;
  a.nstep = c->graph_mode ? c->nstep_dev : nullptr;
  a.t_init = c->g_tinit;
  a.dt = c->g_dt;
  a.rkc = c->g_rkc;
  a.Bdt = Bdt;
  a.A = A;
  a.Bload = Bload;
  a.do_scale = do_scale;
  a.do_update = do_update;
  a.snapshot = snapshot;
  a.fail = c->fail;
  a.trace = c->trace_cur;
  a.mode = c->mode;
  return a;
}

template <typename T>
int launch_edges(fdm_ctx* c, hipStream_t s, double t, double Bdt, double A, int do_scale,
                 int do_update, int snapshot, double Bload) {
  if (c->ntasks == 0) return FDM_OK;
  const EdgeArgs<T> a = edge_args<T>(c, t, Bdt, A, do_scale, do_update, snapshot, Bload);
  dim3 grid((c->ntasks + 63) / 64), block(64);
  EventPair* ep = prof_begin(c, 1, s);
  if (c->mode == 3 && c->any_force)
    hipLaunchKernelGGL((edge_kernel<T, 3, true>), grid, block, 0, s, a);
  else if (c->mode == 3)
    hipLaunchKernelGGL((edge_kernel<T, 3, false>), grid, block, 0, s, a);
  else if (c->any_force)
    hipLaunchKernelGGL((edge_kernel<T, 2, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((edge_kernel<T, 2, false>), grid, block, 0, s, a);
  prof_end(c, ep, 1, s);
  HIPCHK(c, hipGetLastError());
  return FDM_OK;
}

template <typename T>
BulkArgs<T> bulk_args(fdm_ctx* c, double A, double Bdt, int do_update, double t, double srcA,
                      int do_plastic, double dt) {
  BulkArgs<T> a;
  a.Fin = (const T*)c->F[c->cur];
  a.Fout = (T*)c->F[1 - c->cur];
  a.DF = (T*)c->DF;
  a.xq = (const T*)c->met[0];
  a.xr = (const T*)c->met[1];
  a.yq = (const T*)c->met[2];
  a.yr = (const T*)c->met[3];
  a.Ji = (const T*)c->met[4];
  a.X = c->X;
  a.Y = c->Y;
  a.nx = c->nx;
  a.ny = c->ny;
  a.ld = c->ld;
  a.ps = c->ps;
  a.colb = c->colb;
  a.rowb = c->rowb;
  a.blk = c->blk;
  a.nbx = c->nbx;
  a.blocks = c->blocks;
  a.sd = SidesDev{c->sM, c->snormal, c->sF0, c->sx, c->sy, c->sFhat, c->NS};
  a.A = A;
  a.a_zero = (A == 0.0);
  a.Bdt = Bdt;
  a.do_update = do_update;
  a.t = t;
  a.src = c->dsrc;
  a.nplace = c->src.nplace;
  a.srcA = srcA;
  a.Fx = c->src.Fx;
  a.Fy = c->src.Fy;
  a.Mxx = c->src.Mxx;
  a.Mxy = c->src.Mxy;
  a.Myy = c->src.Myy;
  a.do_plastic = do_plastic && c->any_plastic;
  a.dt = dt;
  a.gammap = c->gammap;
  a.lambda = c->lambda;
  a.xcd_swizzle = c->xcd_swizzle;
  a.mrho = (const T*)c->mat[0];
  a.mG = (const T*)c->mat[1];
  a.mM = (const T*)c->mat[2];
  a.mgam = (const T*)c->mat[3];
  a.mnu = (const T*)c->mat[4];
  a.J = (const T*)c->Jp;
  a.W = (T*)c->W;
  a.ji_calc = c->ji_calc ? 1 : 0;
  a.skip_df = 0;
  a.fxq = c->mfp[0];
  a.fxr = c->mfp[1];
  a.fyq = c->mfp[2];
  a.fyr = c->mfp[3];
  a.ibpts = c->iband;
  a.nib = c->niband;
  a.nibwg = 0;  // set by the interior launch that carries them
  a.rcc = nullptr;
  a.rout = nullptr;
  a.nrec = 0;
  a.nrecwg = 0;
  return a;
}

// the lean variants serve launches without sources, manufactured solutions
// or a plastic update
template <typename T>
bool full_variant(const fdm_ctx* c, const BulkArgs<T>& a) {
  return c->src.nplace > 0 || c->any_mms || a.do_plastic;
}

template <typename T, int OP>
void launch_band_op(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a, dim3 grid) {
  // with the band split, the edge points only
  const int2* pts = c->band_split ? c->eband : c->band;
  const int n = c->band_split ? c->neband : c->nband;
  if (c->mode == 3 && c->band_fl3 && !c->diss)
    // one lane per (point, field), 21 points per wave
    hipLaunchKernelGGL((bulk_band3_fl<T, OP>), dim3((n + 20) / 21), dim3(64), 0, s, a, pts, n);
  else if (c->mode == 3)
    hipLaunchKernelGGL((bulk_band3<T, OP>), grid, dim3(64), 0, s, a, pts, n);
  else if (full_variant(c, a) && (c->het || c->diss))
    hipLaunchKernelGGL((bulk_band<T, true, true, OP>), grid, dim3(64), 0, s, a, pts, n);
  else if (full_variant(c, a))
    hipLaunchKernelGGL((bulk_band<T, true, false, OP>), grid, dim3(64), 0, s, a, pts, n);
  else if ((c->het || c->diss) && c->band_fl_ext)
    hipLaunchKernelGGL((bulk_band_fl<T, OP, true>), dim3((n + 11) / 12), dim3(64), 0, s, a, pts, n);
  else if (c->het || c->diss)
    hipLaunchKernelGGL((bulk_band<T, false, true, OP>), grid, dim3(64), 0, s, a, pts, n);
  else if (c->band_fl)
    // one lane per (point, field), twelve points per wave
    hipLaunchKernelGGL((bulk_band_fl<T, OP>), dim3((n + 11) / 12), dim3(64), 0, s, a, pts, n);
  else
    hipLaunchKernelGGL((bulk_band<T, false, false, OP>), grid, dim3(64), 0, s, a, pts, n);
}

template <typename T>
void launch_band(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a) {
  const int n = c->band_split ? c->neband : c->nband;
  if (!n) return;
  EventPair* ep = prof_begin(c, 2, s);
  dim3 grid((n + 63) / 64);
  if (c->sbp == 2) launch_band_op<T, 2>(c, s, a, grid);
  else if (c->sbp == 6) launch_band_op<T, 6>(c, s, a, grid);
  else launch_band_op<T, 4>(c, s, a, grid);
  prof_end(c, ep, 2, s);
}

template <typename T, bool INNER>
void launch_layer_list(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a, int q) {
  EventPair* ep = prof_begin(c, INNER ? 0 : 2, s);
  dim3 grid((c->nlpts[q] + 255) / 256), block(256);
  if (full_variant(c, a))
    hipLaunchKernelGGL((bulk_pml<T, true, INNER, false>), grid, block, 0, s, a, c->lpts[q], c->nlpts[q]);
  else if (c->nt_store)
    hipLaunchKernelGGL((bulk_pml<T, false, INNER, true>), grid, block, 0, s, a, c->lpts[q], c->nlpts[q]);
  else
    hipLaunchKernelGGL((bulk_pml<T, false, INNER, false>), grid, block, 0, s, a, c->lpts[q], c->nlpts[q]);
  prof_end(c, ep, INNER ? 0 : 2, s);
}

// the layer points, after the band on the same stream (their boundary
// points need this stage's edge values); the inner points are timed with
// the interior launches, the rest with the band
template <typename T>
void launch_layer(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a) {
  if (c->nlpts[0]) launch_layer_list<T, true>(c, s, a, 0);
  if (c->nlpts[1]) launch_layer_list<T, false>(c, s, a, 1);
}

template <typename T, int OP, bool HET, bool DISS>
void launch_interior3_v(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a, dim3 grid, dim3 block) {
  const int4* tl = c->tiles[c->tile_shape];
  if (c->nt_store)
    hipLaunchKernelGGL((bulk_interior3<T, 64, 4, true, OP, HET, DISS>), grid, block, 0, s, a, tl);
  else
    hipLaunchKernelGGL((bulk_interior3<T, 64, 4, false, OP, HET, DISS>), grid, block, 0, s, a, tl);
}

template <typename T, int OP>
void launch_interior3(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a, dim3 grid, dim3 block) {
  if constexpr (OP == 4) {
    if (c->diss) {
      if (c->het) launch_interior3_v<T, OP, true, true>(c, s, a, grid, block);
      else launch_interior3_v<T, OP, false, true>(c, s, a, grid, block);
      return;
    }
  }
  if (c->het) launch_interior3_v<T, OP, true, false>(c, s, a, grid, block);
  else launch_interior3_v<T, OP, false, false>(c, s, a, grid, block);
}

// one interior tile variant: FULL (sources, MMS, plasticity), or lean with
// non-temporal or normal stores
template <typename T, int TX, int TY, bool HET, bool DISS, int OP, bool JI, bool MC = false>
void launch_interior_v(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a, dim3 grid, dim3 block,
                       bool full) {
  const int4* tl = c->tiles[c->tile_shape];
  if (full)
    hipLaunchKernelGGL((bulk_interior<T, TX, TY, true, false, HET, DISS, OP, JI, MC>), grid, block, 0, s, a, tl);
  else if (c->nt_store)
    hipLaunchKernelGGL((bulk_interior<T, TX, TY, false, true, HET, DISS, OP, JI, MC>), grid, block, 0, s, a, tl);
  else
    hipLaunchKernelGGL((bulk_interior<T, TX, TY, false, false, HET, DISS, OP, JI, MC>), grid, block, 0, s, a, tl);
}

// the JI variants (Ji formed in the kernel) exist for FP64 only: fdm_create
// sets ji_calc in FP64 alone; so do the MC variants (FP32-delta metrics),
// for the 64 x 4 and 64 x 8 tiles
template <typename T, int TX, int TY, bool HET, bool DISS, int OP>
void launch_interior_j(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a, dim3 grid, dim3 block,
                       bool full) {
  if constexpr (std::is_same<T, double>::value) {
    if constexpr (TX == 64 && (TY == 4 || TY == 8)) {
      if (c->mcomp) {
        if (c->ji_calc) launch_interior_v<T, TX, TY, HET, DISS, OP, true, true>(c, s, a, grid, block, full);
        else launch_interior_v<T, TX, TY, HET, DISS, OP, false, true>(c, s, a, grid, block, full);
        return;
      }
    }
    if (c->ji_calc) {
      launch_interior_v<T, TX, TY, HET, DISS, OP, true>(c, s, a, grid, block, full);
      return;
    }
  }
  launch_interior_v<T, TX, TY, HET, DISS, OP, false>(c, s, a, grid, block, full);
}

// the interior tiles of SBP2 or SBP6 (no dissipation): 64 x 4, or 64 x 8
// for SBP6 in a uniform medium (tile_shape 1: its halo of 3 rows then costs
// 6 staged rows per 8 instead of per 4, at the same occupancy)
template <typename T, int OP, bool HET>
void launch_interior_op(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a, dim3 grid, dim3 block,
                        bool full) {
  if constexpr (OP == 6 && !HET) {
    if (c->tile_shape == 1) {
      launch_interior_j<T, 64, 8, false, false, 6>(c, s, a, grid, block, full);
      return;
    }
  }
  launch_interior_j<T, 64, 4, HET, false, OP>(c, s, a, grid, block, full);
}

template <typename T>
void launch_interior(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a) {
  const int q = c->tile_shape;
  if (!c->ntiles[q]) return;
  EventPair* ep = prof_begin(c, 0, s);
  const bool full = full_variant(c, a);
  c->df6_stale = !full;
  dim3 grid(c->ntiles[q]);
  dim3 block(kTileShape[q][0], kTileShape[q][1]);
  if (c->mode == 3) {
    // antiplane shear runs the 64 x 4 tiles (fdm_create keeps tile_shape 0)
    c->df6_stale = false;
    if (c->sbp == 2) launch_interior3<T, 2>(c, s, a, grid, block);
    else if (c->sbp == 6) launch_interior3<T, 6>(c, s, a, grid, block);
    else if (c->band_split) {
      // the inner band points ride ahead of the tiles (uniform medium, no
      // dissipation)
      BulkArgs<T> b = a;
      b.nibwg = (c->niband + 255) / 256;
      grid.x += b.nibwg;
      launch_interior3<T, 4>(c, s, b, grid, block);
    } else launch_interior3<T, 4>(c, s, a, grid, block);
    prof_end(c, ep, 0, s);
    return;
  }
  if (c->sbp != 4) {
    // SBP2 and SBP6 run the 64 x 4 tiles, SBP6 in a uniform medium the
    // 64 x 8 ones by default (fdm_create and fdm_tune keep tile_shape 0 or
    // 1 there); neither has dissipation
    if (c->sbp == 2) {
      if (c->het) launch_interior_op<T, 2, true>(c, s, a, grid, block, full);
      else launch_interior_op<T, 2, false>(c, s, a, grid, block, full);
    } else {
      if (c->het) launch_interior_op<T, 6, true>(c, s, a, grid, block, full);
      else launch_interior_op<T, 6, false>(c, s, a, grid, block, full);
    }
    prof_end(c, ep, 0, s);
    return;
  }
  // a heterogeneous medium and artificial dissipation run the 64 x 4 tiles
  // only (fdm_create and fdm_tune keep tile_shape 0)
  if (c->het && c->diss) launch_interior_j<T, 64, 4, true, true, 4>(c, s, a, grid, block, full);
  else if (c->het) launch_interior_j<T, 64, 4, true, false, 4>(c, s, a, grid, block, full);
  else {
    // the inner band points ride ahead of the tiles (band split), one lane
    // per point
    BulkArgs<T> b = a;
    const int nt = kTileShape[q][0] * kTileShape[q][1];
    b.nibwg = c->band_split ? (c->niband + nt - 1) / nt : 0;
    b.nrecwg = (a.nrec + nt - 1) / nt;
    grid.x += b.nrecwg + b.nibwg;
    if (c->diss) launch_interior_j<T, 64, 4, false, true, 4>(c, s, b, grid, block, full);
    else if (q == 0) launch_interior_j<T, 64, 4, false, false, 4>(c, s, b, grid, block, full);
    else if (q == 1) launch_interior_j<T, 64, 8, false, false, 4>(c, s, b, grid, block, full);
    else launch_interior_j<T, 128, 4, false, false, 4>(c, s, b, grid, block, full);
  }
  prof_end(c, ep, 0, s);
}

template <typename T>
void launch_pointwise(fdm_ctx* c, hipStream_t s, const BulkArgs<T>& a) {
  c->df6_stale = false;
  EventPair* ep = prof_begin(c, 0, s);
  dim3 block(64, c->tile_y);
  dim3 grid((c->nx + 63) / 64, (c->ny + c->tile_y - 1) / c->tile_y);
  if (c->mode == 3) {
    if (c->sbp == 2) hipLaunchKernelGGL((bulk_kernel3<T, 2>), grid, block, 0, s, a);
    else if (c->sbp == 6) hipLaunchKernelGGL((bulk_kernel3<T, 6>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((bulk_kernel3<T, 4>), grid, block, 0, s, a);
  } else {
    if (c->sbp == 2) hipLaunchKernelGGL((bulk_kernel<T, 2>), grid, block, 0, s, a);
    else if (c->sbp == 6) hipLaunchKernelGGL((bulk_kernel<T, 6>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((bulk_kernel<T, 4>), grid, block, 0, s, a);
  }
  prof_end(c, ep, 0, s);
}

// vx, vy of the fields F at the receivers
__global__ void __launch_bounds__(64) record_receivers_kernel(const double* F, long long ps,
                                                              const long long* cc, int nrec,
                                                              double* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  out[2 * r] = F[cc[r]];
  out[2 * r + 1] = F[ps + cc[r]];
}

// receivers of the fields in F[c->cur] after step n, on stream s
void record_receivers(fdm_ctx* c, hipStream_t s, long long n) {
  hipLaunchKernelGGL(record_receivers_kernel, dim3((c->nrec + 63) / 64), dim3(64), 0, s,
                     (const double*)c->F[c->cur], c->ps, c->rec_cc, c->nrec,
                     c->rec_out + (size_t)(n - c->rec_n0) * c->nrec * 2);
}

// Fork the side stream off the main one before a sequence of stages, and
// join it back afterwards, so that work queued on the main stream later (and
// every synchronisation of it) is ordered after both.
int fork_side(fdm_ctx* c) {
  HIPCHK(c, hipEventRecord(c->evJ, c->stream));
  HIPCHK(c, hipStreamWaitEvent(c->side, c->evJ, 0));
  HIPCHK(c, hipEventRecord(c->evB, c->side));
  HIPCHK(c, hipEventRecord(c->evI, c->stream));
  return FDM_OK;
}

int join_side(fdm_ctx* c) {
  HIPCHK(c, hipEventRecord(c->evB, c->side));
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->evB, 0));
  return FDM_OK;
}

// skip_df: the stage is the last of a step that another step of the same
// call follows, so the interior launch may leave DF unstored (the next
// stage has A = 0)
template <typename T>
int run_stage(fdm_ctx* c, int stage, double t0, double dt, int final_step, int snapshot,
              int skip_df = 0) {
  double A = c->rkA[stage - 1];
  double Bdt = c->rkB[stage - 1] * dt;
  double t = t0 + c->rkC[stage - 1] * dt;  // D%t during this stage (time_step.f90:137)
  double Bload = c->ld_T.empty() ? 1.0 : load_ramp(t, c->ld_T[0]);
  double srcA = c->src.nplace ? stf_amplitude(c->src, t) : 0.0;
  c->g_rkc = c->rkC[stage - 1];
  int last = (stage == c->nstage);
  BulkArgs<T> a =
      bulk_args<T>(c, A, Bdt, final_step ? 0 : 1, t, srcA, (!final_step && last) ? 1 : 0, dt);
  a.skip_df = (skip_df && last && !final_step) ? 1 : 0;
  // receivers of the previous step: in stage 1's interior launch where its
  // kernel carries them (FP64 plane strain, SBP4, uniform medium), else a
  // launch of their own
  const bool rec_now = c->rec_out && stage == 1 && !final_step && c->cur_n > c->first_n;
  const bool rec_fused = rec_now && std::is_same<T, double>::value && c->bulk_mode == 2 &&
                         c->mode == 2 && c->sbp == 4 && !c->het && c->ntiles[c->tile_shape] > 0;
  if (rec_fused) {
    a.rcc = c->rec_cc;
    a.rout = c->rec_out + (size_t)(c->cur_n - 1 - c->rec_n0) * c->nrec * 2;
    a.nrec = c->nrec;
  }
  int rc;
  if (c->bulk_mode == 2) {
    // Interior points need neither this stage's edge values nor this
    // stage's band: only the band of the previous stage (halo values, and
    // the buffer this launch overwrites is the one that band read).  The
    // band needs this stage's edge values and the previous interior launch.
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->evB, 0));
    launch_interior<T>(c, c->stream, a);
    c->trace_cur = (c->trace && !final_step)
                       ? c->trace + ((size_t)(c->cur_n - c->trace_n0) * c->nstage + (stage - 1)) * 3 * c->NI
                       : nullptr;
    rc = launch_edges<T>(c, c->side, t, Bdt, A, 1, final_step ? 0 : 1, snapshot, Bload);
    c->trace_cur = nullptr;
    if (rc) return rc;
    HIPCHK(c, hipStreamWaitEvent(c->side, c->evI, 0));
    launch_band<T>(c, c->side, a);
    launch_layer<T>(c, c->side, a);
    // the previous step's fields are this stage's input until stage 2
    // writes them; both streams have finished them here
    if (rec_now && !rec_fused) record_receivers(c, c->side, c->cur_n - 1);
    HIPCHK(c, hipEventRecord(c->evB, c->side));
    HIPCHK(c, hipEventRecord(c->evI, c->stream));
  } else {
    c->trace_cur = (c->trace && !final_step)
                       ? c->trace + ((size_t)(c->cur_n - c->trace_n0) * c->nstage + (stage - 1)) * 3 * c->NI
                       : nullptr;
    rc = launch_edges<T>(c, c->stream, t, Bdt, A, 1, final_step ? 0 : 1, snapshot, Bload);
    c->trace_cur = nullptr;
    if (rc) return rc;
    launch_pointwise<T>(c, c->stream, a);
    if (c->rec_out && stage == 1 && !final_step && c->cur_n > c->first_n) record_receivers(c, c->stream, c->cur_n - 1);
  }
  HIPCHK(c, hipGetLastError());
  if (!final_step) c->cur = 1 - c->cur;
  return FDM_OK;
}

template <typename T>
int step_impl(fdm_ctx* c, double t, double dt, long long n_first, int nsteps) {
  int rc = fork_side(c);
  c->first_n = n_first;
  for (int k = 0; k < nsteps && !rc; k++) {
    long long n = n_first + k;
    c->cur_n = n;
    double t0 = t + (double)(n - 1) * dt;  // main.f90:176
    // DF after the last stage is read by no later stage (A = 0 at stage 1)
    // and is stored for the caller only by the call's last step
    const int skip_df = (k + 1 < nsteps && c->rkA[0] == 0.0) ? 1 : 0;
    for (int stage = 1; stage <= c->nstage && !rc; stage++)
      rc = run_stage<T>(c, stage, t0, dt, 0, stage == 1, skip_df);
  }
  int rj = join_side(c);
  if (!rc && !rj && c->rec_out && nsteps > 0) {
    record_receivers(c, c->stream, n_first + nsteps - 1);
    HIPCHK(c, hipGetLastError());
  }
  return rc ? rc : rj;
}

template <typename T>
int stage1_impl(fdm_ctx* c, double t0, double dt) {
  int rc = fork_side(c);
  if (!rc) rc = run_stage<T>(c, 1, t0, dt, 1, 1);
  int rj = join_side(c);
  return rc ? rc : rj;
}

// DF6 = nu (DF3 + DF5) at the interior points of every block, in the
// precision of the bulk fields: the value the lean interior kernel computes
// and does not store (rates.f90:920-936)
template <typename T>
void restore_df6(const fdm_ctx* c, T* h) {
  const size_t pe = (size_t)c->ps;
  for (int b = 0; b < c->nb; b++) {
    const BlockDev& B = c->hblocks[b];
    for (int j = B.my + c->nbst; j <= B.py - c->nbst; j++)
      for (int i = B.mx + c->nbst; i <= B.px - c->nbst; i++) {
        const size_t k = (size_t)j * c->ld + i;
        const T nu = c->het ? (T)c->hnu[(size_t)j * c->nx + i] : (T)B.nu;
        h[5 * pe + k] = nu * (h[2 * pe + k] + h[4 * pe + k]);
      }
  }
}

// one step of a member context inside a graph capture: advance the device
// step counter after the stages
__global__ void __launch_bounds__(64) step_counter_kernel(long long* n) {
  if (threadIdx.x == 0) n[0] = n[0] + 1;
}

// Maximum slip velocity on the device (check_nucleation, domain.f90:474-499;
// interfaces.f90:506-517): each workgroup of kVmaxWG lanes takes the maximum
// of V over its share of the frictional nodes, starting from 0 as the
// reference's maxval over non-negative V does; one more workgroup reduces
// those partial maxima.  Maxima are exact, so the result does not depend on
// the order; a NaN never replaces a number, as with std::max.
constexpr int kVmaxWG = 256;

__device__ __forceinline__ double vmax_tree(double v) {
  __shared__ double s[kVmaxWG];
  s[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = kVmaxWG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double x = s[threadIdx.x + w];
      if (s[threadIdx.x] < x) s[threadIdx.x] = x;
    }
    __syncthreads();
  }
  return s[0];
}

__global__ void __launch_bounds__(kVmaxWG) vmax_partial_kernel(const double* __restrict__ V,
                                                               const int* __restrict__ node, int n,
                                                               double* __restrict__ part) {
  const int k = blockIdx.x * kVmaxWG + threadIdx.x;
  double v = 0.0;
  if (k < n) {
    const double x = V[node[k]];
    v = (v < x) ? x : v;
  }
  v = vmax_tree(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kVmaxWG) vmax_final_kernel(const double* __restrict__ part, int np,
                                                             double* __restrict__ out) {
  double v = 0.0;
  for (int k = threadIdx.x; k < np; k += kVmaxWG) {
    const double x = part[k];
    v = (v < x) ? x : v;
  }
  v = vmax_tree(v);
  if (threadIdx.x == 0) out[0] = v;
}

// the two reduction launches on the context's stream with the buffer red
// (red[0] the result, red[1..np] the partial maxima), then the result to
// the host
int vmax_reduce(fdm_ctx* c, double* red, double* vmax) {
  if (c->nfnode == 0) {
    *vmax = 0.0;
    return FDM_OK;
  }
  const int np = (c->nfnode + kVmaxWG - 1) / kVmaxWG;
  const double* V = c->ist + (size_t)FDM_IF_V * c->NI;
  hipLaunchKernelGGL(vmax_partial_kernel, dim3(np), dim3(kVmaxWG), 0, c->stream, V, c->fnode,
                     c->nfnode, red + 1);
  hipLaunchKernelGGL(vmax_final_kernel, dim3(1), dim3(kVmaxWG), 0, c->stream, red + 1, np, red);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(vmax, red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDM_OK;
}

// Device memory of fdm_max_slip_velocity_noalloc, which allocates nothing:
// the result and up to kNoallocParts partial maxima (kNoallocParts *
// kVmaxWG frictional nodes at most), 520 bytes per device, held until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: it uses the device address that fdm_create resolves once per device
// (g_scratch_ptr).  Every call on a device shares that device's copy, and
// each call overwrites every element it reads, so no state carries from one
// call to the next; calls that overlap on one device race.
constexpr int kNoallocParts = 64;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[1 + kNoallocParts];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// resolve the device address of scratch on device dev (current on entry)
int resolve_scratch(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) return FDM_EDEVICE;
  if (g_scratch_ptr[dev]) return FDM_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FDM_EDEVICE;
  g_scratch_ptr[dev] = (double*)ptr;
  return FDM_OK;
}

int check_fail(fdm_ctx* c) {
  if (!c->fail) return FDM_OK;
  int h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (h) return set_err(c, FDM_ENEWTON, "friction Newton solve failed (friction.f90:1012-1036)");
  return FDM_OK;
}

}  // namespace

extern "C" {

const char* fdm_last_error(const fdm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

void fdm_destroy(fdm_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->copy) {
    (void)hipStreamSynchronize(c->copy);
    (void)hipStreamDestroy(c->copy);
  }
  for (auto& sn : c->snap) {
    if (sn.host) (void)hipHostFree(sn.host);
    if (sn.staged) (void)hipEventDestroy(sn.staged);
    if (sn.done) (void)hipEventDestroy(sn.done);
  }
  for (auto& ep : c->pool) {
    (void)hipEventDestroy(ep.a);
    (void)hipEventDestroy(ep.b);
  }
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->evI) (void)hipEventDestroy(c->evI);
  if (c->evB) (void)hipEventDestroy(c->evB);
  if (c->evJ) (void)hipEventDestroy(c->evJ);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int fdm_create(const fdm_problem* P, fdm_ctx** out) {
  if (!out) return set_err(nullptr, FDM_EINVAL, "fdm_create: null output pointer");
  *out = nullptr;
  if (!P || P->nx < 1 || P->ny < 1 || P->nblocks < 1 || P->nstage < 1 ||
      P->nstage > FDM_MAX_STAGES || !P->blocks || (P->nifaces > 0 && !P->ifaces) || !P->xq ||
      !P->xr || !P->yq || !P->yr || !P->Ji)
    return set_err(nullptr, FDM_EINVAL, "fdm_create: malformed problem description");
  if (P->precision != FDM_FP64 && P->precision != FDM_FP32_BULK)
    return set_err(nullptr, FDM_EINVAL, "fdm_create: unknown precision");
  if (P->sbp != 0 && P->sbp != 2 && P->sbp != 4 && P->sbp != 6)
    return set_err(nullptr, FDM_EINVAL,
                   "fdm_create: FDmethod must be SBP2, SBP4 or SBP6 ('SBP" + std::to_string(P->sbp) +
                       "' requested)");
  const int sbp = P->sbp == 0 ? 4 : P->sbp;
  // closure depth nbst and closure width nbnd (fd_coeff.f90:56-79)
  const int nbst = sbp == 2 ? 1 : sbp == 4 ? 4 : 6, nbnd = sbp == 2 ? 2 : sbp == 4 ? 6 : 9;
  if (sbp != 4) {
    // the reference forms artificial dissipation and the perfectly matched
    // layer with SBP4's tables (rates.f90:1397-2021, 2024-2619)
    for (int b = 0; b < P->nblocks; b++)
      if (P->blocks[b].Cdiss > 0.0 || P->blocks[b].pmlx || P->blocks[b].pmly)
        return set_err(nullptr, FDM_EINVAL,
                       "fdm_create: artificial dissipation and the perfectly matched layer need "
                       "FDmethod SBP4");
  }
  for (int b = 0; b < P->nblocks; b++) {
    // every point of a block lies in at most one closure per direction, and
    // the closure rows read nbnd points into the block
    const fdm_block& B = P->blocks[b];
    const int n = std::min(B.px - B.mx, B.py - B.my) + 1;
    if (n < std::max(2 * nbst, nbnd))
      return set_err(nullptr, FDM_EINVAL,
                     "fdm_create: block " + std::to_string(b) + " has " + std::to_string(n) +
                         " points across; SBP" + std::to_string(sbp) + " needs at least " +
                         std::to_string(std::max(2 * nbst, nbnd)));
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(nullptr, FDM_EDEVICE, "fdm_create: no HIP device available");
  if (P->source.nplace < 0 || P->source.nplace > FDM_MAX_SOURCE_PLACES)
    return set_err(nullptr, FDM_EINVAL, "fdm_create: bad source placement count");
  if (P->mode != 0 && P->mode != 2 && P->mode != 3)
    return set_err(nullptr, FDM_EINVAL, "fdm_create: mode must be 2 or 3");
  if (P->mode == 3) {
    bool unsupported = P->source.nplace > 0;
    for (int b = 0; b < P->nblocks; b++)
      unsupported |= P->blocks[b].plastic || P->blocks[b].mms || P->blocks[b].pmlx || P->blocks[b].pmly;
    if (unsupported)
      return set_err(nullptr, FDM_EINVAL,
                     "fdm_create: mode 3 covers elastic blocks without sources, plasticity, "
                     "manufactured solutions or PML");
  }

  fdm_ctx* c = new fdm_ctx();
  c->mode = P->mode == 3 ? 3 : 2;
  c->sbp = sbp;
  c->nbst = nbst;
  c->nF = c->mode == 3 ? 3 : 6;
  c->nx = P->nx;
  c->ny = P->ny;
  c->nb = P->nblocks;
  c->ni = P->nifaces;
  c->nstage = P->nstage;
  c->precision = P->precision;
  c->esz = P->precision == FDM_FP64 ? 8 : 4;
  for (int s = 0; s < P->nstage; s++) {
    c->rkA[s] = P->rkA[s];
    c->rkB[s] = P->rkB[s];
  }
  for (int s = 0; s <= P->nstage; s++) c->rkC[s] = P->rkC[s];
  // row pitch: a multiple of 64 elements so every row starts 256/512-byte aligned
  // Column i of a plane is stored at element coff + i, with coff chosen so
  // that the interior tiles of the first block column (which start at column
  // mx + nbst) begin on a 512-byte boundary; every plane pointer below already
  // includes the offset.
  {
    int mx0 = P->nblocks > 0 ? P->blocks[0].mx : 0;
    for (int b = 1; b < P->nblocks; b++) mx0 = std::min(mx0, P->blocks[b].mx);
    c->coff = (64 - (mx0 + nbst) % 64) % 64;
  }
  c->ld = (P->nx + c->coff + 63) / 64 * 64;
  c->ps = (long long)c->ld * P->ny;
  if (const char* e = getenv("FDM_PLANE_PAD")) {
    // extra elements between planes (a multiple of 64 keeps rows aligned)
    long long v = atoll(e);
    if (v >= 0 && v <= (1 << 20)) c->ps += v / 64 * 64;
  }
  // launch-shape overrides (tuning sweeps): FDM_BULK in {tile, generic}, FDM_TILE in
  // {64x4, 64x8, 128x4},
  // FDM_TILE_Y in {1, 2, 4} (per-point kernel), FDM_XCD_SWIZZLE in {0, 1}
  if (const char* e = getenv("FDM_TILE_Y")) {
    int v = atoi(e);
    if (v == 1 || v == 2 || v == 4) c->tile_y = v;
  }
  if (const char* e = getenv("FDM_BULK")) {
    if (!strcmp(e, "generic")) c->bulk_mode = 0;
    else if (!strcmp(e, "tile")) c->bulk_mode = 2;
  }
  if (const char* e = getenv("FDM_TILE")) {
    for (int q = 0; q < 3; q++) {
      char name[16];
      snprintf(name, sizeof name, "%dx%d", kTileShape[q][0], kTileShape[q][1]);
      if (!strcmp(e, name)) c->tile_shape = q;
    }
  }
  if (const char* e = getenv("FDM_XCD_SWIZZLE")) c->xcd_swizzle = atoi(e) != 0;
  (void)hipGetDevice(&c->device);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_err(nullptr, FDM_EDEVICE, "fdm_create: hipStreamCreate failed");
  }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      // the stage marks order work on this device only: no system-scope
      // fence (an event recorded with one costs about 3 us of GPU time per
      // stage on the stream that records it)
      hipEventCreateWithFlags(&c->evI, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess ||
      hipEventCreateWithFlags(&c->evB, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess ||
      hipEventCreateWithFlags(&c->evJ, hipEventDisableTiming) != hipSuccess) {
    fdm_destroy(c);
    return set_err(nullptr, FDM_EDEVICE, "fdm_create: stream/event creation failed");
  }
  auto fail = [&](int rc) {
    g_create_error = c->err.empty() ? g_create_error : c->err;
    fdm_destroy(c);
    return rc;
  };

  // ---- block layout: blocks must tile the grid as a tensor product ----
  std::vector<int> xs, ys;
  for (int b = 0; b < c->nb; b++) {
    const fdm_block& B = P->blocks[b];
    if (B.mx < 0 || B.px >= P->nx || B.my < 0 || B.py >= P->ny || B.px - B.mx + 1 < 8 ||
        B.py - B.my + 1 < 8) {
      set_err(c, FDM_EINVAL, "fdm_create: block " + std::to_string(b) +
                                 " out of range or with fewer than 8 points per direction");
      return fail(FDM_EINVAL);
    }
    xs.push_back(B.mx);
    ys.push_back(B.my);
  }
  std::sort(xs.begin(), xs.end());
  xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
  std::sort(ys.begin(), ys.end());
  ys.erase(std::unique(ys.begin(), ys.end()), ys.end());
  c->nbx = (int)xs.size();
  c->nby = (int)ys.size();
  std::vector<int> blk((size_t)c->nbx * c->nby, -1), colb(P->nx, -1), rowb(P->ny, -1);
  for (int b = 0; b < c->nb; b++) {
    const fdm_block& B = P->blocks[b];
    int cx = int(std::lower_bound(xs.begin(), xs.end(), B.mx) - xs.begin());
    int cy = int(std::lower_bound(ys.begin(), ys.end(), B.my) - ys.begin());
    if (blk[(size_t)cy * c->nbx + cx] != -1) {
      set_err(c, FDM_EINVAL, "fdm_create: two blocks share a corner index");
      return fail(FDM_EINVAL);
    }
    blk[(size_t)cy * c->nbx + cx] = b;
    for (int i = B.mx; i <= B.px; i++) {
      if (colb[i] != -1 && colb[i] != cx) {
        set_err(c, FDM_EINVAL, "fdm_create: blocks overlap in x");
        return fail(FDM_EINVAL);
      }
      colb[i] = cx;
    }
    for (int j = B.my; j <= B.py; j++) {
      if (rowb[j] != -1 && rowb[j] != cy) {
        set_err(c, FDM_EINVAL, "fdm_create: blocks overlap in y");
        return fail(FDM_EINVAL);
      }
      rowb[j] = cy;
    }
  }
  for (int v : blk)
    if (v < 0) {
      set_err(c, FDM_EINVAL, "fdm_create: blocks do not form a tensor-product layout");
      return fail(FDM_EINVAL);
    }
  for (int v : colb)
    if (v < 0) {
      set_err(c, FDM_EINVAL, "fdm_create: grid columns not covered by blocks");
      return fail(FDM_EINVAL);
    }
  for (int v : rowb)
    if (v < 0) {
      set_err(c, FDM_EINVAL, "fdm_create: grid rows not covered by blocks");
      return fail(FDM_EINVAL);
    }
  // every block of one column spans the same x range (and rows the same y range)
  for (int b = 0; b < c->nb; b++) {
    const fdm_block& B = P->blocks[b];
    for (int b2 = 0; b2 < c->nb; b2++) {
      const fdm_block& B2 = P->blocks[b2];
      if ((B.mx == B2.mx && B.px != B2.px) || (B.my == B2.my && B.py != B2.py)) {
        set_err(c, FDM_EINVAL, "fdm_create: blocks do not form a tensor-product layout");
        return fail(FDM_EINVAL);
      }
    }
  }
  int rc;
  if ((rc = dupload(c, &c->blk, blk.data(), blk.size() * sizeof(int)))) return fail(rc);
  if ((rc = dupload(c, &c->colb, colb.data(), colb.size() * sizeof(int)))) return fail(rc);
  c->hcolb = colb;
  c->hrowb = rowb;
  c->hblk = blk;
  {
    // interior tiles: each block's rectangle [mx+nbst, px-nbst] x [my+nbst, py-nbst]
    // covered from its corner, x fastest; band: every other point, row-major.
    // A block with a perfectly matched layer in effect has no tiles and no
    // band: all of its points go to the layer list.
    auto layer = [&](const fdm_block& B) { return (B.pmlx || B.pmly) && !(B.Cdiss > 0.0); };
    std::vector<int4> tl[3];
    for (int b = 0; b < c->nb; b++) {
      const fdm_block& B = P->blocks[b];
      if (layer(B)) continue;
      const int ia = B.mx + nbst, ib = B.px - nbst, ja = B.my + nbst, jb = B.py - nbst;
      if (ia > ib || ja > jb) continue;
      c->ninterior += (long long)(ib - ia + 1) * (jb - ja + 1);
      for (int q = 0; q < 3; q++) {
        const int tx = kTileShape[q][0], ty = kTileShape[q][1];
        for (int j0 = ja; j0 <= jb; j0 += ty)
          for (int i0 = ia; i0 <= ib; i0 += tx) tl[q].push_back(make_int4(b, i0, j0, 0));
      }
    }
    std::vector<int2> band, lpts[2];
    for (int j = 0; j < P->ny; j++)
      for (int i = 0; i < P->nx; i++) {
        const fdm_block& B = P->blocks[block_of(c, i, j)];
        const bool edge = i - B.mx < nbst || B.px - i < nbst || j - B.my < nbst || B.py - j < nbst;
        if (c->mode == 2 && layer(B)) {
          if (edge) lpts[1].push_back(make_int2(i, j));
        } else if (edge) {
          band.push_back(make_int2(i, j));
        }
      }
    // inner layer points: each row of a layer block's interior in runs of
    // 64 that start where the stored planes are 64-element aligned
    // ((i + coff) % 64 == 0), padded with (-1, -1) entries that do nothing
    for (int b = 0; b < c->nb; b++) {
      const fdm_block& B = P->blocks[b];
      if (c->mode != 2 || !layer(B)) continue;
      const int ia = B.mx + 4, ib = B.px - 4;
      if (ia > ib) continue;
      const int i0 = ia - (ia + c->coff) % 64;
      const int i1 = ib + 1 + (64 - (ib + 1 + c->coff) % 64) % 64;
      // count the points as they are pushed: a block with fewer than
      // nine rows has no inner rows at all
      for (int j = B.my + 4; j <= B.py - 4; j++) {
        for (int i = i0; i < i1; i++)
          lpts[0].push_back(i >= ia && i <= ib ? make_int2(i, j) : make_int2(-1, -1));
        c->nlayer_inner += ib - ia + 1;
      }
    }
    for (int q = 0; q < 2; q++) {
      c->nlpts[q] = (int)lpts[q].size();
      if (c->nlpts[q] && (rc = dupload(c, &c->lpts[q], lpts[q].data(), lpts[q].size() * sizeof(int2))))
        return fail(rc);
    }
    for (int q = 0; q < 3; q++) {
      c->ntiles[q] = (int)tl[q].size();
      if (c->ntiles[q] && (rc = dupload(c, &c->tiles[q], tl[q].data(), tl[q].size() * sizeof(int4))))
        return fail(rc);
    }
    c->nband = (int)band.size();
    if (c->nband && (rc = dupload(c, &c->band, band.data(), band.size() * sizeof(int2)))) return fail(rc);
    for (int q = 0; q < 3; q++) c->htiles[q] = std::move(tl[q]);
    c->hband = std::move(band);
  }
  if ((rc = dupload(c, &c->rowb, rowb.data(), rowb.size() * sizeof(int)))) return fail(rc);

  // ---- sides ----
  c->hblocks.resize(c->nb);
  int NS = 0;
  for (int b = 0; b < c->nb; b++)
    for (int s = 0; s < 4; s++) {
      c->hblocks[b].side_off[s] = NS;
      c->hblocks[b].side_n[s] = side_len(P->blocks[b], s);
      NS += c->hblocks[b].side_n[s];
    }
  c->NS = NS;
  std::vector<double> hM((size_t)5 * NS), hN((size_t)2 * NS), hF0((size_t)6 * NS),
      hx(NS, 0.0), hy(NS, 0.0);
  bool have_xy = false;
  for (int b = 0; b < c->nb; b++) {
    const fdm_block& B = P->blocks[b];
    BlockDev& D = c->hblocks[b];
    D.mx = B.mx;
    D.px = B.px;
    D.my = B.my;
    D.py = B.py;
    D.plastic = B.plastic;
    D.mms = B.mms;
    D.id = B.id;
    D.pad = 0;
    D.rho = B.rho;
    D.G = B.G;
    D.M = B.M;
    D.gamma = B.gamma;
    D.nu = B.nu;
    // rates.f90:400-405: a(s) = DI(s)/rho, b(s) = DI(s)*G, c(s) = DI(s)*M;
    // DI(s) of the operator, zero beyond its half-width (fd_coeff.f90:56-79)
    const double DIs[3] = {sbp_di_host(sbp, 1), sbp_di_host(sbp, 2), sbp_di_host(sbp, 3)};
    for (int s = 0; s < 3; s++) {
      D.cI[0][s] = DIs[s] / B.rho;
      D.cI[1][s] = DIs[s] / B.rho;
      D.cI[2][s] = DIs[s] * B.M;
      D.cI[3][s] = DIs[s] * B.G;
      D.cI[4][s] = DIs[s] * B.M;
    }
    D.mu = B.mu;
    D.beta = B.beta;
    D.b = B.b;
    D.h = B.h;
    D.eta = B.eta;
    D.K = B.K;
    for (int k = 0; k < 6; k++) D.s0[k] = B.s0[k];
    // dissipation = (Cdiss > 0) (domain.f90:180); d(0) = 2 AI(0) Cdiss,
    // d(s) = AI(s) Cdiss (rates.f90:1416-1421)
    D.Cdiss = B.Cdiss > 0.0 ? B.Cdiss : 0.0;
    D.dA[0] = 2.0 * FDM_AI0 * D.Cdiss;
    D.dA[1] = FDM_AI1 * D.Cdiss;
    D.dA[2] = FDM_AI2 * D.Cdiss;
    if (D.Cdiss > 0.0) c->diss = true;
    // perfectly matched layer (material.f90:152-223), in effect unless the
    // block has dissipation (set_rates_all, time_step.f90:241-246); hxi =
    // Ji*yr and hyi = Ji*xq at the first point of each region (rates.f90:
    // 2060-2061, 2158-2159; regions as time_step.f90:357-420 bound them)
    D.pmlx = (B.pmlx && D.Cdiss == 0.0) ? 1 : 0;
    D.pmly = (B.pmly && D.Cdiss == 0.0) ? 1 : 0;
    for (int k = 0; k < 7; k++) D.pml[k] = B.pml[k];
    {
      const int i0[3] = {B.mx, B.mx + 4, B.px - 3}, j0[3] = {B.my, B.my + 4, B.py - 3};
      for (int ry = 0; ry < 3; ry++)
        for (int rx = 0; rx < 3; rx++) {
          const size_t q = (size_t)j0[ry] * P->nx + i0[rx];
          D.phx[3 * ry + rx] = P->Ji[q] * P->yr[q];
          D.phy[3 * ry + rx] = P->Ji[q] * P->xq[q];
        }
    }
    if (D.pmlx || D.pmly) c->pml = true;
    if (B.pmlx) c->pml_wx = true;
    if (B.pmly) c->pml_wy = true;
    if ((D.pmlx || D.pmly) && B.pml[0] > 0.0 && (!P->x || !P->y)) {
      set_err(c, FDM_EINVAL, "fdm_create: a graded PML needs coordinates x, y");
      return fail(FDM_EINVAL);
    }
    if (B.plastic) c->any_plastic = true;
    if (B.mms && (!P->x || !P->y)) {
      set_err(c, FDM_EINVAL, "fdm_create: MMS source needs coordinates x, y");
      return fail(FDM_EINVAL);
    }
    for (int s = 0; s < 4; s++) {
      const fdm_side& S = B.side[s];
      int n = D.side_n[s];
      D.side_bc[s] = S.bc;
      if (S.n != n || !S.normal || !S.M || !S.F0) {
        set_err(c, FDM_EINVAL, "fdm_create: block " + std::to_string(b) + " side " +
                                   std::to_string(s) + " arrays missing or wrong length");
        return fail(FDM_EINVAL);
      }
      if (S.bc == FDM_BC_MMS_SIN_W && (!S.x || !S.y)) {
        set_err(c, FDM_EINVAL, "fdm_create: MMS boundary data needs side coordinates");
        return fail(FDM_EINVAL);
      }
      int o = D.side_off[s];
      for (int k = 0; k < n; k++) {
        for (int q = 0; q < 5; q++) hM[(size_t)q * NS + o + k] = S.M[(size_t)q * n + k];
        for (int q = 0; q < 2; q++) hN[(size_t)q * NS + o + k] = S.normal[(size_t)q * n + k];
        for (int q = 0; q < 6; q++)
          hF0[(size_t)q * NS + o + k] = q < c->nF ? S.F0[(size_t)q * n + k] : 0.0;
        if (S.x && S.y) {
          hx[o + k] = S.x[k];
          hy[o + k] = S.y[k];
          have_xy = true;
        }
      }
    }
  }
  if ((rc = dupload(c, &c->sM, hM.data(), hM.size() * 8))) return fail(rc);
  if ((rc = dupload(c, &c->snormal, hN.data(), hN.size() * 8))) return fail(rc);
  if ((rc = dupload(c, &c->sF0, hF0.data(), hF0.size() * 8))) return fail(rc);
  if (have_xy) {
    if ((rc = dupload(c, &c->sx, hx.data(), hx.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->sy, hy.data(), hy.size() * 8))) return fail(rc);
  }
  {
    std::vector<double> fh((size_t)6 * NS, 1e40);
    if ((rc = dupload(c, &c->sFhat, fh.data(), fh.size() * 8))) return fail(rc);
  }
  if ((rc = dupload(c, &c->blocks, c->hblocks.data(), c->hblocks.size() * sizeof(BlockDev))))
    return fail(rc);

  // ---- interfaces ----
  std::vector<int2> tasks;
  for (int b = 0; b < c->nb; b++)
    for (int s = 0; s < 4; s++)
      if (P->blocks[b].side[s].bc != FDM_BC_NONE)
        for (int k = 0; k < c->hblocks[b].side_n[s]; k++) tasks.push_back(make_int2(4 * b + s, k));
  std::vector<std::vector<char>> covered(c->nb, std::vector<char>(4, 0));
  int NI = 0;
  c->hifaces.resize(c->ni);
  for (int ii = 0; ii < c->ni; ii++) {
    const fdm_iface& I = P->ifaces[ii];
    IfaceDev& D = c->hifaces[ii];
    int n = I.p - I.m + 1;
    if (I.blockm < 0 || I.blockm >= c->nb || I.blockp < 0 || I.blockp >= c->nb || n < 1 ||
        !I.nhat || (I.direction != 0 && I.direction != 1)) {
      set_err(c, FDM_EINVAL, "fdm_create: interface " + std::to_string(ii) + " malformed");
      return fail(FDM_EINVAL);
    }
    const fdm_block& Bm = P->blocks[I.blockm];
    const fdm_block& Bp = P->blocks[I.blockp];
    int sm = I.direction == 0 ? FDM_SIDE_R : FDM_SIDE_T;
    int sp = I.direction == 0 ? FDM_SIDE_L : FDM_SIDE_B;
    int lo_m = I.direction == 0 ? Bm.my : Bm.mx, hi_m = I.direction == 0 ? Bm.py : Bm.px;
    int lo_p = I.direction == 0 ? Bp.my : Bp.mx, hi_p = I.direction == 0 ? Bp.py : Bp.px;
    if (I.m < lo_m || I.p > hi_m || I.m < lo_p || I.p > hi_p) {
      set_err(c, FDM_EINVAL, "fdm_create: interface " + std::to_string(ii) +
                                 " range outside the adjacent block sides");
      return fail(FDM_EINVAL);
    }
    if (I.coupling == FDM_COUPLING_FRICTION) {
      c->any_friction = true;
      if ((I.law >= FDM_LAW_SL && !I.rs) || (I.law == FDM_LAW_SW && !I.sw) || I.law < 0 ||
          I.law > FDM_LAW_RAL) {
        set_err(c, FDM_EINVAL, "fdm_create: interface " + std::to_string(ii) +
                                   " friction parameters missing or unknown law");
        return fail(FDM_EINVAL);
      }
    }
    covered[I.blockm][sm] = 1;
    covered[I.blockp][sp] = 1;
    D.direction = I.direction;
    D.bm = I.blockm;
    D.bp = I.blockp;
    D.m = I.m;
    D.p = I.p;
    D.coupling = I.coupling;
    D.law = I.law;
    D.reg = I.reg;
    D.rup_field = I.rup_field;
    D.node_off = NI;
    D.n = n;
    D.pad = 0;
    D.rup_threshold = I.rup_threshold;
    D.skempton = I.skempton;
    D.flock = I.flock;
    D.ld_T = I.ld_T;
    D.force = I.coupling == FDM_COUPLING_FRICTION && I.force ? 1 : 0;
    D.pad2 = 0;
    for (int q = 0; q < 8; q++) D.kn[q] = I.kn[q];
    if (D.force) c->any_force = true;
    if (D.force && !I.kn_x) {
      set_err(c, FDM_EINVAL, "fdm_create: interface " + std::to_string(ii) +
                                 " has kinematic forcing and no kn_x");
      return fail(FDM_EINVAL);
    }
    c->ld_T.push_back(I.ld_T);
    for (int k = 0; k < n; k++) tasks.push_back(make_int2(-(ii + 1), k));
    NI += n;
  }
  // every side without a boundary condition must be an interface side
  for (int b = 0; b < c->nb; b++)
    for (int s = 0; s < 4; s++)
      if (P->blocks[b].side[s].bc == FDM_BC_NONE && !covered[b][s]) {
        set_err(c, FDM_EINVAL, "fdm_create: block " + std::to_string(b) + " side " +
                                   std::to_string(s) + " has bc 'none' and no interface");
        return fail(FDM_EINVAL);
      }
  // one load-ramp time constant is applied to all interfaces per stage
  for (double T : c->ld_T)
    if (T != c->ld_T[0]) {
      set_err(c, FDM_EINVAL, "fdm_create: interfaces with different ld%T are not supported");
      return fail(FDM_EINVAL);
    }
  c->NI = NI;
  {
    size_t NIs = NI > 0 ? NI : 1;
    std::vector<double> nh(2 * NIs, 0.0), rs(7 * NIs, 1.0), sw(4 * NIs, 1.0), S0b(NIs, 0.0),
        N0b(NIs, 0.0), ldS(NIs, 0.0), ldN(NIs, 0.0), st((size_t)FDM_IF_NFIELDS * NIs, 0.0);
    std::vector<int> outv(NIs, 0);
    std::vector<double> knx(NIs, 0.0);
    for (int ii = 0; ii < c->ni; ii++) {
      const fdm_iface& I = P->ifaces[ii];
      int n = c->hifaces[ii].n, o = c->hifaces[ii].node_off;
      for (int k = 0; k < n; k++) {
        nh[o + k] = I.nhat[k];
        nh[NIs + o + k] = I.nhat[(size_t)n + k];
        if (I.rs)
          for (int q = 0; q < 7; q++) rs[(size_t)q * NIs + o + k] = I.rs[(size_t)q * n + k];
        if (I.sw)
          for (int q = 0; q < 4; q++) sw[(size_t)q * NIs + o + k] = I.sw[(size_t)q * n + k];
        if (I.S0base) S0b[o + k] = I.S0base[k];
        if (I.N0base) N0b[o + k] = I.N0base[k];
        if (I.ldS) ldS[o + k] = I.ldS[k];
        if (I.ldN) ldN[o + k] = I.ldN[k];
        if (I.outside) outv[o + k] = I.outside[k];
        if (c->hifaces[ii].force) knx[o + k] = I.kn_x[k];
        // initial values (friction.f90:217-249)
        auto S = [&](int f) -> double& { return st[(size_t)f * NIs + o + k]; };
        S(FDM_IF_D) = 0.0;
        S(FDM_IF_PSI) = I.Psi0 ? I.Psi0[k] : 0.0;
        S(FDM_IF_DPSI) = 1e40;
        S(FDM_IF_TRUP) = 1e10;
        S(FDM_IF_V) = 1e40;
        S(FDM_IF_O) = 1e40;
        S(FDM_IF_S) = 1e40;
        S(FDM_IF_N) = 1e40;
        S(FDM_IF_S0) = 1e40;
        S(FDM_IF_N0) = 1e40;
        S(FDM_IF_DS) = 0.0;
        S(FDM_IF_DN) = 0.0;
        S(FDM_IF_DDS) = 0.0;
        S(FDM_IF_DDN) = 0.0;
        S(FDM_IF_W) = 0.0;
        S(FDM_IF_DW) = 1e40;
        S(FDM_IF_NREG) = 1e40;
        S(FDM_IF_DNREG) = 1e40;
      }
    }
    // component strides use NIs (= NI, or 1 when there is no interface)
    c->NI = (int)NIs;
    for (auto& D : c->hifaces) (void)D;
    if ((rc = dupload(c, &c->ifaces, c->hifaces.data(),
                      std::max<size_t>(1, c->hifaces.size()) * sizeof(IfaceDev))))
      return fail(rc);
    if ((rc = dupload(c, &c->inhat, nh.data(), nh.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->irs, rs.data(), rs.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->isw, sw.data(), sw.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->iS0b, S0b.data(), S0b.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->iN0b, N0b.data(), N0b.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->ildS, ldS.data(), ldS.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->ildN, ldN.data(), ldN.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->iout, outv.data(), outv.size() * sizeof(int)))) return fail(rc);
    if ((rc = dupload(c, &c->iknx, knx.data(), knx.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->ist, st.data(), st.size() * 8))) return fail(rc);
    if ((rc = dupload(c, &c->ist1, st.data(), st.size() * 8))) return fail(rc);
  }
  c->ntasks = (int)tasks.size();
  if ((rc = dupload(c, &c->tasks, tasks.data(), std::max<size_t>(1, tasks.size()) * sizeof(int2))))
    return fail(rc);
  c->htasks = tasks;

  // ---- source ----
  c->src = P->source;
  {
    std::vector<SourcePlaceDev> sp(std::max(1, c->src.nplace));
    for (int k = 0; k < c->src.nplace; k++) {
      const fdm_source_place& S = c->src.place[k];
      if (S.block < 0 || S.block >= c->nb) {
        set_err(c, FDM_EINVAL, "fdm_create: source placement block out of range");
        return fail(FDM_EINVAL);
      }
      sp[k].block = S.block;
      sp[k].i0 = S.i0;
      sp[k].j0 = S.j0;
      sp[k].pad = 0;
      for (int q = 0; q < 5; q++) {
        sp[k].dx[q] = S.dx[q];
        sp[k].dy[q] = S.dy[q];
      }
    }
    if ((rc = dupload(c, &c->dsrc, sp.data(), sp.size() * sizeof(SourcePlaceDev)))) return fail(rc);
  }

  // ---- bulk arrays ----
  size_t pbytes = (size_t)c->ps * c->esz;
  // the tile kernel addresses a plane with 32-bit byte offsets
  if (pbytes >= ((size_t)1 << 31) && c->bulk_mode == 2) c->bulk_mode = 0;
  // one arena of 23 planes with a common stride: F (two buffers of 6), DF
  // (6), metrics xq xr yq yr Ji (5); a heterogeneous medium adds rho, G, M,
  // gamma, nu (5)
  c->het = P->mat_rho && P->mat_cs && P->mat_cp;
  if (c->het || c->diss || c->mode == 3 || c->sbp != 4) {
    // SBP6 in a uniform medium: 64 x 8 unless FDM_TILE asks for 64 x 4
    const bool t8 = c->sbp == 6 && !c->het && !c->diss && c->mode == 2;
    const char* e = getenv("FDM_TILE");
    c->tile_shape = (t8 && !(e && !strcmp(e, "64x4"))) ? 1 : 0;
  }
  if (c->diss && !P->J) {
    set_err(c, FDM_EINVAL, "fdm_create: artificial dissipation needs the Jacobian J");
    return fail(FDM_EINVAL);
  }
  // artificial dissipation adds the Jacobian J (1); antiplane shear has 3
  // fields instead of 6 (14 planes without the extras)
  const int nF = c->nF, pmet = 3 * nF, pmat = pmet + 5, pJ = pmat + (c->het ? 5 : 0);
  const int nplanes = pJ + (c->diss ? 1 : 0);
  void* arena = nullptr;
  const size_t slack = 64 * c->esz;  // the last plane extends coff elements past nplanes * ps
  if ((rc = dalloc(c, &arena, nplanes * pbytes + slack))) return fail(rc);
  HIPCHK(c, hipMemset(arena, 0, nplanes * pbytes + slack));
  char* ab = (char*)arena + (size_t)c->coff * c->esz;
  c->nt_store = nt_store_choice((double)nplanes * (double)pbytes);
  if (c->het) {
    // init_medium (material.f90:254-262), in FP64, then the field precision
    const size_t np = (size_t)c->nx * c->ny;
    std::vector<double> m[5];
    for (auto& v : m) v.resize(np);
    for (size_t q = 0; q < np; q++) {
      const double rho = P->mat_rho[q], cs = P->mat_cs[q], cp = P->mat_cp[q];
      m[0][q] = rho;
      m[1][q] = rho * (cs * cs);
      m[2][q] = rho * (cp * cp);
      m[3][q] = 1.0 - 2.0 * ((cs / cp) * (cs / cp));
      m[4][q] = 0.5 * (m[2][q] - 2.0 * m[1][q]) / (m[2][q] - m[1][q]);
    }
    c->hnu = m[4];
    for (int k = 0; k < 5; k++) {
      c->mat[k] = ab + (size_t)(pmat + k) * pbytes;
      if (c->precision == FDM_FP64) {
        std::vector<double> h;
        pack_plane(h, m[k].data(), c->nx, c->ny, c->ld);
        HIPCHK(c, hipMemcpy(c->mat[k], h.data(), h.size() * 8, hipMemcpyHostToDevice));
      } else {
        std::vector<float> h;
        pack_plane(h, m[k].data(), c->nx, c->ny, c->ld);
        HIPCHK(c, hipMemcpy(c->mat[k], h.data(), h.size() * 4, hipMemcpyHostToDevice));
      }
    }
  }
  if (c->diss) {
    c->Jp = ab + (size_t)pJ * pbytes;
    if (c->precision == FDM_FP64) {
      std::vector<double> h;
      pack_plane(h, P->J, c->nx, c->ny, c->ld);
      HIPCHK(c, hipMemcpy(c->Jp, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    } else {
      std::vector<float> h;
      pack_plane(h, P->J, c->nx, c->ny, c->ld);
      HIPCHK(c, hipMemcpy(c->Jp, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
  }
  c->F[0] = ab;
  c->F[1] = ab + (size_t)nF * pbytes;
  c->DF = ab + (size_t)(2 * nF) * pbytes;
  for (int q = 0; q < 5; q++) c->met[q] = ab + (size_t)(pmet + q) * pbytes;
  const double* mets[5] = {P->xq, P->xr, P->yq, P->yr, P->Ji};
  for (int q = 0; q < 5; q++) {
    if (c->precision == FDM_FP64) {
      std::vector<double> h;
      pack_plane(h, mets[q], c->nx, c->ny, c->ld);
      HIPCHK(c, hipMemcpy(c->met[q], h.data(), h.size() * 8, hipMemcpyHostToDevice));
    } else {
      std::vector<float> h;
      pack_plane(h, mets[q], c->nx, c->ny, c->ld);
      HIPCHK(c, hipMemcpy(c->met[q], h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
  }
  {
    // In FP64 the interior kernel may form Ji itself when the given Ji is
    // 1/(xq yr - xr yq) to the bit (grid.f90:531-534) at every point: the
    // Ji plane is then not read by that kernel (8 bytes per point and
    // stage).  This file is compiled without contraction, as the kernel.
    bool exact = c->precision == FDM_FP64;
    const size_t np = (size_t)c->nx * c->ny;
    for (size_t k = 0; k < np && exact; k++) {
      const double J = P->xq[k] * P->yr[k] - P->xr[k] * P->yq[k];
      const double ji = 1.0 / J;
      exact = std::memcmp(&ji, &P->Ji[k], sizeof(double)) == 0;
    }
    if (const char* v = getenv("FDM_JI_CALC")) exact = exact && atoi(v) != 0;
    c->ji_calc = exact;
  }
  if (c->precision == FDM_FP64 && c->mode == 2 && c->bulk_mode == 2) {
    // FP32-delta metrics: xq, xr, yq, yr of a point stored as d = m - base
    // in FP32, base = the metric at the first interior point of its block,
    // kept where base + (double)d reproduces m to the bit and marked NaN
    // (read from the FP64 plane) elsewhere.  The interior tiles then read 16
    // instead of 32 bytes of metrics per point; worth it when the escapes
    // are rare (grids made of nearly affine blocks), so used when at most
    // 1/64 of the metric values escape.
    const double* mets[4] = {P->xq, P->xr, P->yq, P->yr};
    for (int b = 0; b < c->nb; b++) {
      BlockDev& D = c->hblocks[b];
      const int i0 = std::min(D.mx + c->nbst, D.px), j0 = std::min(D.my + c->nbst, D.py);
      for (int q = 0; q < 4; q++) D.mbase[q] = mets[q][(size_t)j0 * c->nx + i0];
    }
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    std::vector<float> h[4];
    long long esc = 0;
    for (int q = 0; q < 4; q++) h[q].assign((size_t)c->ps, 0.0f);
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        const BlockDev& D = c->hblocks[block_of(c, i, j)];
        for (int q = 0; q < 4; q++) {
          const double m = mets[q][(size_t)j * c->nx + i];
          const double d = m - D.mbase[q];
          const float f = (float)d;
          const double back = D.mbase[q] + (double)f;
          const bool ok = std::isfinite(d) && (double)f == d && std::memcmp(&back, &m, sizeof(double)) == 0;
          h[q][(size_t)j * c->ld + i] = ok ? f : qnan;
          esc += ok ? 0 : 1;
        }
      }
    bool use = esc * 64 <= 4LL * c->nx * c->ny;
    if (const char* v = getenv("FDM_METRIC_F32")) use = use && atoi(v) != 0;
    if (use) {
      void* mb = nullptr;
      const size_t fbytes = (size_t)c->ps * 4;
      if ((rc = dalloc(c, &mb, 4 * fbytes + (size_t)c->coff * 4 + 256))) return fail(rc);
      char* fb = (char*)mb + (size_t)c->coff * 4;
      for (int q = 0; q < 4; q++) {
        c->mfp[q] = (float*)(fb + q * fbytes);
        HIPCHK(c, hipMemcpy(c->mfp[q], h[q].data(), fbytes, hipMemcpyHostToDevice));
      }
      HIPCHK(c, hipMemcpy(c->blocks, c->hblocks.data(), c->hblocks.size() * sizeof(BlockDev),
                          hipMemcpyHostToDevice));
      c->mcomp = true;
      c->mc_escapes = esc;
    }
  }
  for (int b = 0; b < c->nb; b++) c->any_mms |= (P->blocks[b].mms != 0);
  bool graded = false;
  for (int b = 0; b < c->nb; b++) graded |= (c->hblocks[b].pmlx || c->hblocks[b].pmly) && c->hblocks[b].pml[0] > 0.0;
  if (c->any_mms || graded) {
    size_t np = (size_t)c->nx * c->ny;
    if ((rc = dupload(c, &c->X, P->x, np * 8))) return fail(rc);
    if ((rc = dupload(c, &c->Y, P->y, np * 8))) return fail(rc);
  }
  if (c->pml_wx || c->pml_wy) {
    // Wx, Wy start at zero (fields.f90:203-209), and so do their rates; the
    // planes carry the field planes' column offset, so that rows align alike
    const size_t bytes = 24 * (size_t)c->ps * c->esz + 64 * c->esz;
    void* wb = nullptr;
    if ((rc = dalloc(c, &wb, bytes))) return fail(rc);
    HIPCHK(c, hipMemset(wb, 0, bytes));
    c->W = (char*)wb + (size_t)c->coff * c->esz;
  }
  if (c->any_plastic) {
    size_t np = (size_t)c->nx * c->ny;
    std::vector<double> z(np, 0.0);
    if ((rc = dupload(c, &c->gammap, z.data(), np * 8))) return fail(rc);
    if ((rc = dupload(c, &c->lambda, z.data(), np * 8))) return fail(rc);
  }
  {
    int z = 0;
    if ((rc = dupload(c, &c->fail, &z, sizeof(int)))) return fail(rc);
  }
  // Both band forms pay off in FP64 with SBP4, where the interior launch is
  // long enough to carry the inner band points and the field lanes shorten
  // the second stream; in FP32-bulk (an interior launch half as long) and
  // with SBP6 (whose field lanes took longer) the band kernel of one lane
  // per point keeps the step shortest.
  c->band_fl = c->precision == FDM_FP64 && c->sbp == 4;
  if (const char* e = getenv("FDM_BAND_FL")) c->band_fl = atoi(e) != 0;
  if (const char* e = getenv("FDM_BAND_FL3")) c->band_fl3 = atoi(e) != 0;
  c->band_fl_ext = c->precision == FDM_FP32_BULK && c->het && !c->diss;
  if (const char* e = getenv("FDM_BAND_FL_EXT")) c->band_fl_ext = atoi(e) != 0;
  if (c->bulk_mode == 2 && c->sbp == 4 && !c->het && (c->mode == 2 || !c->diss) && c->nband) {
    // band split: the band points on no block edge need no edge values of
    // the stage and join the interior launch (bulk_interior); the second
    // stream's band kernel takes the edge points (FDM_BAND_SPLIT=0|1
    // overrides the default).  Antiplane shear splits only on request: its
    // step took longer with the split (C2-size grid, 0.364 against 0.356
    // ms in FP64, 0.239 against 0.215 in FP32-bulk).
    bool use = c->precision == FDM_FP64 && c->mode == 2;
    if (const char* e = getenv("FDM_BAND_SPLIT")) use = atoi(e) != 0;
    if (use) {
      std::vector<int2> ib, eb;
      for (const int2& p : c->hband) {
        const fdm_block& B = P->blocks[block_of(c, p.x, p.y)];
        const bool edge = p.x == B.mx || p.x == B.px || p.y == B.my || p.y == B.py;
        (edge ? eb : ib).push_back(p);
      }
      c->niband = (int)ib.size();
      c->neband = (int)eb.size();
      if (c->niband && (rc = dupload(c, &c->iband, ib.data(), ib.size() * sizeof(int2)))) return fail(rc);
      if (c->neband && (rc = dupload(c, &c->eband, eb.data(), eb.size() * sizeof(int2)))) return fail(rc);
      c->band_split = true;
    }
  }
  {
    // frictional nodes and the reduction buffer of fdm_max_slip_velocity
    std::vector<int> fn;
    for (const IfaceDev& I : c->hifaces)
      if (I.coupling == FDM_COUPLING_FRICTION)
        for (int k = 0; k < I.n; k++) fn.push_back(I.node_off + k);
    c->nfnode = (int)fn.size();
    if (c->nfnode) {
      if ((rc = dupload(c, &c->fnode, fn.data(), fn.size() * sizeof(int)))) return fail(rc);
      const size_t np = (fn.size() + kVmaxWG - 1) / kVmaxWG;
      if ((rc = dalloc(c, &c->vred, (1 + np) * sizeof(double)))) return fail(rc);
    }
    // the allocation-free reduction's device scratch on this device
    if (resolve_scratch(c->device) != FDM_OK) {
      set_err(c, FDM_EDEVICE, "fdm_create: cannot resolve the device scratch address");
      return fail(FDM_EDEVICE);
    }
  }
  *out = c;
  return FDM_OK;
}

int fdm_set_fields(fdm_ctx* c, const double* F, const double* DF) {
  if (!c || !F) return FDM_EINVAL;
  c->df6_stale = false;
  size_t pe = (size_t)c->ps;
  if (c->precision == FDM_FP64) {
    std::vector<double> h((size_t)c->nF * pe, 0.0);
    for (int l = 0; l < c->nF; l++)
      for (int j = 0; j < c->ny; j++)
        memcpy(&h[l * pe + (size_t)j * c->ld], F + ((size_t)l * c->ny + j) * c->nx, c->nx * 8);
    HIPCHK(c, hipMemcpy(c->F[c->cur], h.data(), h.size() * 8, hipMemcpyHostToDevice));
    for (int l = 0; l < c->nF; l++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++)
          h[l * pe + (size_t)j * c->ld + i] = DF ? DF[((size_t)l * c->ny + j) * c->nx + i] : 1e40;
    HIPCHK(c, hipMemcpy(c->DF, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  } else {
    // FP32-bulk: each block's fields are stored as deviations from the
    // block mean of the fields handed in here, so that FP32 resolves the
    // waves riding on a large prestress
    std::vector<double> sum((size_t)c->nb * 6, 0.0), cnt(c->nb, 0.0);
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        int b = block_of(c, i, j);
        cnt[b] += 1.0;
        for (int l = 0; l < c->nF; l++) sum[(size_t)b * 6 + l] += F[((size_t)l * c->ny + j) * c->nx + i];
      }
    for (int b = 0; b < c->nb; b++)
      for (int l = 0; l < c->nF; l++)
        c->hblocks[b].Fb[l] = cnt[b] > 0 ? (double)(float)(sum[(size_t)b * 6 + l] / cnt[b]) : 0.0;
    HIPCHK(c, hipMemcpy(c->blocks, c->hblocks.data(), c->hblocks.size() * sizeof(BlockDev),
                        hipMemcpyHostToDevice));
    std::vector<float> h((size_t)c->nF * pe, 0.f);
    for (int l = 0; l < c->nF; l++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++)
          h[l * pe + (size_t)j * c->ld + i] =
              (float)(F[((size_t)l * c->ny + j) * c->nx + i] - c->hblocks[block_of(c, i, j)].Fb[l]);
    HIPCHK(c, hipMemcpy(c->F[c->cur], h.data(), h.size() * 4, hipMemcpyHostToDevice));
    for (int l = 0; l < c->nF; l++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++) {
          double v = DF ? DF[((size_t)l * c->ny + j) * c->nx + i] : 3.0e38;
          h[l * pe + (size_t)j * c->ld + i] = (float)std::fmax(std::fmin(v, 3.0e38), -3.0e38);
        }
    HIPCHK(c, hipMemcpy(c->DF, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  }
  return FDM_OK;
}

int fdm_get_fields(fdm_ctx* c, double* F, double* DF) {
  if (!c) return FDM_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  size_t pe = (size_t)c->ps;
  for (int which = 0; which < 2; which++) {
    double* dst = which == 0 ? F : DF;
    if (!dst) continue;
    const void* src = which == 0 ? c->F[c->cur] : c->DF;
    if (c->precision == FDM_FP64) {
      std::vector<double> h((size_t)c->nF * pe);
      HIPCHK(c, hipMemcpy(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost));
      if (which == 1 && c->df6_stale && c->mode == 2) restore_df6(c, h.data());
      for (int l = 0; l < c->nF; l++)
        for (int j = 0; j < c->ny; j++)
          memcpy(dst + ((size_t)l * c->ny + j) * c->nx, &h[l * pe + (size_t)j * c->ld], c->nx * 8);
    } else {
      std::vector<float> h((size_t)c->nF * pe);
      HIPCHK(c, hipMemcpy(h.data(), src, h.size() * 4, hipMemcpyDeviceToHost));
      if (which == 1 && c->df6_stale && c->mode == 2) restore_df6(c, h.data());
      for (int l = 0; l < c->nF; l++)
        for (int j = 0; j < c->ny; j++)
          for (int i = 0; i < c->nx; i++) {
            double v = (double)h[l * pe + (size_t)j * c->ld + i];
            if (which == 0) v = v + c->hblocks[block_of(c, i, j)].Fb[l];
            dst[((size_t)l * c->ny + j) * c->nx + i] = v;
          }
    }
  }
  return check_fail(c);
}

int fdm_snapshot(fdm_ctx* c, uint32_t field_mask, int32_t* slot) {
  if (!c || !slot || field_mask == 0 || (field_mask >> c->nF) != 0) return FDM_EINVAL;
  int k = 0;
  while (k < FDM_SNAPSHOT_SLOTS && c->snap[k].busy) k++;
  if (k == FDM_SNAPSHOT_SLOTS)
    return set_err(c, FDM_EINVAL, "fdm_snapshot: all snapshot slots are waiting to be read");
  auto& sn = c->snap[k];
  const size_t pb = (size_t)c->ps * c->esz;
  if (!c->copy) HIPCHK(c, hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking));
  if (!sn.dev) {
    int rc = dalloc(c, &sn.dev, (size_t)c->nF * pb);
    if (rc) return rc;
    HIPCHK(c, hipHostMalloc(&sn.host, (size_t)c->nF * pb, hipHostMallocDefault));
    HIPCHK(c, hipEventCreateWithFlags(&sn.staged, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&sn.done, hipEventDisableTiming));
  }
  // the selected planes of the current fields, packed, then to the host on
  // the copy stream while the main stream goes on
  int q = 0;
  for (int l = 0; l < c->nF; l++) {
    if (!(field_mask >> l & 1u)) continue;
    HIPCHK(c, hipMemcpyAsync((char*)sn.dev + q * pb, (const char*)c->F[c->cur] + l * pb, pb,
                             hipMemcpyDeviceToDevice, c->stream));
    q++;
  }
  HIPCHK(c, hipEventRecord(sn.staged, c->stream));
  HIPCHK(c, hipStreamWaitEvent(c->copy, sn.staged, 0));
  HIPCHK(c, hipMemcpyAsync(sn.host, sn.dev, q * pb, hipMemcpyDeviceToHost, c->copy));
  HIPCHK(c, hipEventRecord(sn.done, c->copy));
  sn.mask = field_mask;
  sn.busy = true;
  *slot = k;
  return FDM_OK;
}

int fdm_snapshot_read(fdm_ctx* c, int32_t slot, double* out) {
  if (!c || !out || slot < 0 || slot >= FDM_SNAPSHOT_SLOTS || !c->snap[slot].busy) return FDM_EINVAL;
  auto& sn = c->snap[slot];
  HIPCHK(c, hipEventSynchronize(sn.done));
  const size_t pe = (size_t)c->ps, plane = (size_t)c->nx * c->ny;
  int q = 0;
  for (int l = 0; l < c->nF; l++) {
    if (!(sn.mask >> l & 1u)) continue;
    double* dst = out + q * plane;
    if (c->precision == FDM_FP64) {
      const double* h = (const double*)sn.host + q * pe;
      for (int j = 0; j < c->ny; j++) memcpy(dst + (size_t)j * c->nx, h + (size_t)j * c->ld, c->nx * 8);
    } else {
      const float* h = (const float*)sn.host + q * pe;
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++)
          dst[(size_t)j * c->nx + i] =
              (double)h[(size_t)j * c->ld + i] + c->hblocks[block_of(c, i, j)].Fb[l];
    }
    q++;
  }
  sn.busy = false;
  return FDM_OK;
}

int fdm_init_edges(fdm_ctx* c, double t) {
  if (!c) return FDM_EINVAL;
  // prepare_edges + enforce_edge_conditions(D,0d0) (domain.f90:346-356)
  double Bload = c->ld_T.empty() ? 1.0 : load_ramp(t, c->ld_T[0]);
  int rc = c->precision == FDM_FP64 ? launch_edges<double>(c, c->stream, t, 0.0, 1.0, 0, 0, 0, Bload)
                                    : launch_edges<float>(c, c->stream, t, 0.0, 1.0, 0, 0, 0, Bload);
  if (rc) return rc;
  return check_fail(c);
}

int fdm_step(fdm_ctx* c, double t, double dt, int64_t n_first, int32_t nsteps) {
  if (!c || nsteps < 0) return FDM_EINVAL;
  return c->precision == FDM_FP64 ? step_impl<double>(c, t, dt, n_first, nsteps)
                                  : step_impl<float>(c, t, dt, n_first, nsteps);
}

int fdm_stage1(fdm_ctx* c, double t, double dt, int64_t n) {
  if (!c) return FDM_EINVAL;
  double t0 = t + (double)(n - 1) * dt;
  return c->precision == FDM_FP64 ? stage1_impl<double>(c, t0, dt) : stage1_impl<float>(c, t0, dt);
}

static int iface_copy(fdm_ctx* c, int32_t ii, int32_t field, double* out, const double* in,
                      double* base) {
  if (!c || ii < 0 || ii >= c->ni || field < 0 || field >= FDM_IF_NFIELDS) return FDM_EINVAL;
  const IfaceDev& I = c->hifaces[ii];
  double* p = base + (size_t)field * c->NI + I.node_off;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out) HIPCHK(c, hipMemcpy(out, p, I.n * 8, hipMemcpyDeviceToHost));
  if (in) HIPCHK(c, hipMemcpy(p, in, I.n * 8, hipMemcpyHostToDevice));
  return FDM_OK;
}

int fdm_get_iface(fdm_ctx* c, int32_t ii, int32_t field, double* out) {
  if (!out) return FDM_EINVAL;
  int rc = iface_copy(c, ii, field, out, nullptr, c ? c->ist : nullptr);
  return rc ? rc : check_fail(c);
}

int fdm_set_iface(fdm_ctx* c, int32_t ii, int32_t field, const double* in) {
  if (!in) return FDM_EINVAL;
  return iface_copy(c, ii, field, nullptr, in, c ? c->ist : nullptr);
}

int fdm_get_iface_stage1(fdm_ctx* c, int32_t ii, int32_t field, double* out) {
  if (!out) return FDM_EINVAL;
  int rc = iface_copy(c, ii, field, out, nullptr, c ? c->ist1 : nullptr);
  return rc ? rc : check_fail(c);
}

int fdm_get_plastic(fdm_ctx* c, double* gammap, double* lambda) {
  if (!c) return FDM_EINVAL;
  size_t np = (size_t)c->nx * c->ny;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!c->any_plastic) {
    if (gammap) std::fill(gammap, gammap + np, 0.0);
    if (lambda) std::fill(lambda, lambda + np, 0.0);
    return check_fail(c);
  }
  if (gammap) HIPCHK(c, hipMemcpy(gammap, c->gammap, np * 8, hipMemcpyDeviceToHost));
  if (lambda) HIPCHK(c, hipMemcpy(lambda, c->lambda, np * 8, hipMemcpyDeviceToHost));
  return check_fail(c);
}

int fdm_set_plastic(fdm_ctx* c, const double* gammap, const double* lambda) {
  if (!c) return FDM_EINVAL;
  if (!c->any_plastic) return set_err(c, FDM_EINVAL, "fdm_set_plastic: no block is plastic");
  size_t np = (size_t)c->nx * c->ny;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (gammap) HIPCHK(c, hipMemcpy(c->gammap, gammap, np * 8, hipMemcpyHostToDevice));
  if (lambda) HIPCHK(c, hipMemcpy(c->lambda, lambda, np * 8, hipMemcpyHostToDevice));
  return FDM_OK;
}

namespace {

// copy Wx (first = 0) or Wy (first = 6) between [6][ny][nx] doubles and the
// device planes
int pml_copy(fdm_ctx* c, int first, double* out, const double* in) {
  const size_t pe = (size_t)c->ps;
  const size_t n = 6 * pe;
  char* base = (char*)c->W + (size_t)first * pe * c->esz;
  std::vector<double> hd;
  std::vector<float> hf;
  if (c->precision == FDM_FP64) hd.assign(n, 0.0);
  else hf.assign(n, 0.f);
  if (out) {
    if (c->precision == FDM_FP64) HIPCHK(c, hipMemcpy(hd.data(), base, n * 8, hipMemcpyDeviceToHost));
    else HIPCHK(c, hipMemcpy(hf.data(), base, n * 4, hipMemcpyDeviceToHost));
  }
  for (int l = 0; l < 6; l++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        const size_t k = l * pe + (size_t)j * c->ld + i, o = ((size_t)l * c->ny + j) * c->nx + i;
        if (out) out[o] = c->precision == FDM_FP64 ? hd[k] : (double)hf[k];
        else if (c->precision == FDM_FP64) hd[k] = in[o];
        else hf[k] = (float)in[o];
      }
  if (!out) {
    if (c->precision == FDM_FP64) HIPCHK(c, hipMemcpy(base, hd.data(), n * 8, hipMemcpyHostToDevice));
    else HIPCHK(c, hipMemcpy(base, hf.data(), n * 4, hipMemcpyHostToDevice));
  }
  return FDM_OK;
}

}  // namespace

int fdm_get_pml(fdm_ctx* c, double* Wx, double* Wy) {
  if (!c) return FDM_EINVAL;
  if ((Wx && !c->pml_wx) || (Wy && !c->pml_wy))
    return set_err(c, FDM_EINVAL, "fdm_get_pml: no block has that PML direction");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc;
  if (Wx && (rc = pml_copy(c, 0, Wx, nullptr))) return rc;
  if (Wy && (rc = pml_copy(c, 6, Wy, nullptr))) return rc;
  return check_fail(c);
}

int fdm_set_pml(fdm_ctx* c, const double* Wx, const double* Wy) {
  if (!c) return FDM_EINVAL;
  if ((Wx && !c->pml_wx) || (Wy && !c->pml_wy))
    return set_err(c, FDM_EINVAL, "fdm_set_pml: no block has that PML direction");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc;
  if (Wx && (rc = pml_copy(c, 0, nullptr, Wx))) return rc;
  if (Wy && (rc = pml_copy(c, 6, nullptr, Wy))) return rc;
  return FDM_OK;
}

int fdm_max_slip_velocity(fdm_ctx* c, double* vmax) {
  if (!c || !vmax) return FDM_EINVAL;
  const int rc = vmax_reduce(c, c->vred, vmax);
  return rc ? rc : check_fail(c);
}

int fdm_max_slip_velocity_noalloc(fdm_ctx* c, double* vmax) {
  if (!c || !vmax) return FDM_EINVAL;
  if (c->nfnode > kNoallocParts * kVmaxWG)
    return set_err(c, FDM_EINVAL,
                   "fdm_max_slip_velocity_noalloc: " + std::to_string(c->nfnode) +
                       " frictional interface nodes exceed its limit of " +
                       std::to_string(kNoallocParts * kVmaxWG) + " (use fdm_max_slip_velocity)");
  double* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices) ptr = g_scratch_ptr[c->device];
  }
  if (!ptr) return set_err(c, FDM_EDEVICE, "fdm_max_slip_velocity_noalloc: device scratch unresolved");
  const int rc = vmax_reduce(c, ptr, vmax);
  return rc ? rc : check_fail(c);
}

int fdm_sync(fdm_ctx* c) {
  if (!c) return FDM_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_fail(c);
}

int fdm_profile_enable(fdm_ctx* c, int32_t enable) {
  if (!c) return FDM_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->prof = enable != 0;
  c->pool_used = 0;
  for (int k = 0; k < 3; k++) {
    c->kind_ms[k] = 0.0;
    c->kind_n[k] = 0;
  }
  return FDM_OK;
}

int fdm_profile_read(fdm_ctx* c, fdm_profile_t* out) {
  if (!c || !out) return FDM_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < c->pool_used; k++) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->pool[k].a, c->pool[k].b));
    c->kind_ms[c->pair_kind[k]] += ms;
  }
  c->pool_used = 0;
  out->bulk_ms = c->kind_ms[0];
  out->bulk_launches = c->kind_n[0];
  out->edge_ms = c->kind_ms[1];
  out->edge_launches = c->kind_n[1];
  out->band_ms = c->kind_ms[2];
  out->band_launches = c->kind_n[2];
  // the interior launch's points: tiles, inner layer points and, with the
  // band split, the inner band points
  out->bulk_points = c->bulk_mode == 2 ? (int64_t)c->ninterior + c->nlayer_inner + c->niband
                                       : (int64_t)c->nx * c->ny;
  return FDM_OK;
}

int fdm_tune(fdm_ctx* c, int32_t knob, int32_t value) {
  if (!c) return FDM_EINVAL;
  if (knob == FDM_TUNE_TILE && value >= 0 && value < 3) {
    const bool t8 = c->sbp == 6 && !c->het && !c->diss && c->mode == 2;
    if ((c->het || c->diss || c->mode == 3 || c->sbp != 4) && value != 0 && !(t8 && value == 1))
      return set_err(c, FDM_EINVAL,
                     "fdm_tune: a heterogeneous medium, artificial dissipation, antiplane shear and "
                     "SBP2 run 64x4 tiles only, SBP6 64x4 or 64x8");
    c->tile_shape = value;
    return FDM_OK;
  }
  if (knob == FDM_TUNE_SWIZZLE && (value == 0 || value == 1)) {
    c->xcd_swizzle = value;
    return FDM_OK;
  }
  if (knob == FDM_TUNE_NT_STORE && (value == 0 || value == 1)) {
    c->nt_store = value != 0;
    return FDM_OK;
  }
  if (knob == FDM_TUNE_METRIC_F32 && (value == 0 || value == 1)) {
    if (value && !c->mfp[0])
      return set_err(c, FDM_EINVAL, "fdm_tune: this context's metrics have no FP32-delta form");
    c->mcomp = value != 0;
    return FDM_OK;
  }
  return set_err(c, FDM_EINVAL, "fdm_tune: unknown knob or value");
}

}  // extern "C"

// An ensemble steps many contexts with batched launches: every RK stage is
// one interior launch per tile shape in use, one edge launch and one band
// launch over the merged work lists of all members, with the two-stream
// ordering of a single context.  The per-member kernel arguments of each
// (buffer parity, stage) sit in device tables; stage times come from one
// device step counter.  A step is captured once per starting buffer parity
// as a hipGraph (FDM_ENSEMBLE_GRAPH=0 launches the same kernels directly).
struct fdm_ensemble {
  std::vector<fdm_ctx*> m;
  int precision = FDM_FP64, nstage = 0, flip = 0, swz = 1;
  bool any_plastic = false, use_graph = true, nt_store = false, any_force = false;
  hipStream_t master = nullptr, side = nullptr;
  hipEvent_t evJ = nullptr, evI = nullptr, evB = nullptr, done = nullptr;
  std::vector<hipEvent_t> pre;  // each member's earlier work, before a step
  // merged lists: tiles (block, i0, j0, member) per tile shape; band points
  // (i, j, member, valid) and edge tasks (x, y, member, valid), each member's
  // run padded to whole waves so that a wave serves one member
  int4* tiles[3] = {nullptr, nullptr, nullptr};
  int ntiles[3] = {0, 0, 0};
  int4* band = nullptr;
  int nband = 0;
  int4* tasks = nullptr;
  int ntasks = 0;
  void* btab = nullptr;  // BulkArgs<T>[2][nstage][members]
  void* etab = nullptr;  // EdgeArgs<T>[2][nstage][members]
  long long* nstep = nullptr;
  hipGraphExec_t exec[2] = {nullptr, nullptr};  // one RK step from F buffer 0 / 1
  std::vector<double> t, dt;                     // member times the tables hold
  std::vector<void*> allocs;
  std::string err;
};

namespace {

int ens_err(fdm_ensemble* e, int rc, const std::string& msg) {
  if (e) e->err = msg;
  return rc;
}

#define ENSCHK(e, call)                                                              \
  do {                                                                               \
    hipError_t _st = (call);                                                         \
    if (_st != hipSuccess)                                                           \
      return ens_err(e, FDM_EDEVICE, std::string(#call ": ") + hipGetErrorString(_st)); \
  } while (0)

template <typename P>
int ens_upload(fdm_ensemble* e, P** p, const void* host, size_t bytes) {
  ENSCHK(e, hipMalloc((void**)p, std::max<size_t>(bytes, 64)));
  e->allocs.push_back((void*)*p);
  if (host && bytes) ENSCHK(e, hipMemcpy(*p, host, bytes, hipMemcpyHostToDevice));
  return FDM_OK;
}

__global__ void __launch_bounds__(64) step_set_kernel(long long* n, long long v) {
  if (threadIdx.x == 0) n[0] = v;
}

// Per-member arguments of every (parity, stage): the values run_stage
// passes for a member whose current buffer at that stage is the parity.
template <typename T>
int ens_tables(fdm_ensemble* e) {
  const size_t M = e->m.size(), ns = e->nstage;
  std::vector<BulkArgs<T>> bt(2 * ns * M);
  std::vector<EdgeArgs<T>> et(2 * ns * M);
  for (int p = 0; p < 2; p++)
    for (size_t s = 1; s <= ns; s++)
      for (size_t k = 0; k < M; k++) {
        fdm_ctx* c = e->m[k];
        const int cs = p ^ (int)((s - 1) & 1);
        const double A = c->rkA[s - 1], Bdt = c->rkB[s - 1] * e->dt[k];
        // stage time of step 1; the edge kernels form it from the step counter
        const double ts = e->t[k] + c->rkC[s - 1] * e->dt[k];
        const size_t o = ((size_t)p * ns + (s - 1)) * M + k;
        const int saved = c->cur;
        c->cur = cs;
        bt[o] = bulk_args<T>(c, A, Bdt, 1, ts, 0.0, s == ns ? 1 : 0, e->dt[k]);
        et[o] = edge_args<T>(c, ts, Bdt, A, 1, 1, s == 1 ? 1 : 0, 1.0);
        c->cur = saved;
        et[o].nstep = e->nstep;
        et[o].t_init = e->t[k];
        et[o].dt = e->dt[k];
        et[o].rkc = c->rkC[s - 1];
      }
  if (!e->btab && ens_upload(e, (BulkArgs<T>**)&e->btab, nullptr, bt.size() * sizeof(bt[0])))
    return FDM_EDEVICE;
  if (!e->etab && ens_upload(e, (EdgeArgs<T>**)&e->etab, nullptr, et.size() * sizeof(et[0])))
    return FDM_EDEVICE;
  // no launch of an earlier step may still read the tables
  ENSCHK(e, hipStreamSynchronize(e->master));
  ENSCHK(e, hipMemcpy(e->btab, bt.data(), bt.size() * sizeof(bt[0]), hipMemcpyHostToDevice));
  ENSCHK(e, hipMemcpy(e->etab, et.data(), et.size() * sizeof(et[0]), hipMemcpyHostToDevice));
  return FDM_OK;
}

template <typename T, bool FULL, bool NT>
void ens_interior(fdm_ensemble* e, const BulkArgs<T>* tab) {
  for (int q = 0; q < 3; q++) {
    if (!e->ntiles[q]) continue;
    const dim3 grid(e->ntiles[q]), block(kTileShape[q][0], kTileShape[q][1]);
    if (q == 0)
      hipLaunchKernelGGL((bulk_interior_batch<T, 64, 4, FULL, NT>), grid, block, 0, e->master, tab, e->tiles[q], e->swz);
    else if (q == 1)
      hipLaunchKernelGGL((bulk_interior_batch<T, 64, 8, FULL, NT>), grid, block, 0, e->master, tab, e->tiles[q], e->swz);
    else
      hipLaunchKernelGGL((bulk_interior_batch<T, 128, 4, FULL, NT>), grid, block, 0, e->master, tab, e->tiles[q], e->swz);
  }
}

// One RK step of every member from buffer parity p: the stream DAG of
// run_stage (interior on the master stream, edge and band on the side
// stream), then the step counter.
template <typename T>
int ens_step_launches(fdm_ensemble* e, int p) {
  const size_t M = e->m.size(), ns = e->nstage;
  ENSCHK(e, hipEventRecord(e->evJ, e->master));
  ENSCHK(e, hipStreamWaitEvent(e->side, e->evJ, 0));
  ENSCHK(e, hipEventRecord(e->evB, e->side));
  ENSCHK(e, hipEventRecord(e->evI, e->master));
  for (size_t s = 1; s <= ns; s++) {
    const size_t o = ((size_t)p * ns + (s - 1)) * M;
    const BulkArgs<T>* bt = (const BulkArgs<T>*)e->btab + o;
    const EdgeArgs<T>* et = (const EdgeArgs<T>*)e->etab + o;
    const bool full = e->any_plastic && s == ns;
    ENSCHK(e, hipStreamWaitEvent(e->master, e->evB, 0));
    if (full) ens_interior<T, true, false>(e, bt);
    else if (e->nt_store) ens_interior<T, false, true>(e, bt);
    else ens_interior<T, false, false>(e, bt);
    if (e->ntasks && e->any_force)
      hipLaunchKernelGGL((edge_kernel_batch<T, true>), dim3(e->ntasks / 64), dim3(64), 0, e->side, et,
                         e->tasks, e->ntasks);
    else if (e->ntasks)
      hipLaunchKernelGGL((edge_kernel_batch<T, false>), dim3(e->ntasks / 64), dim3(64), 0, e->side, et,
                         e->tasks, e->ntasks);
    ENSCHK(e, hipStreamWaitEvent(e->side, e->evI, 0));
    if (e->nband) {
      if (full)
        hipLaunchKernelGGL((bulk_band_batch<T, true>), dim3(e->nband / 64), dim3(64), 0, e->side, bt, e->band,
                           e->nband);
      else
        hipLaunchKernelGGL((bulk_band_batch<T, false>), dim3(e->nband / 64), dim3(64), 0, e->side, bt,
                           e->band, e->nband);
    }
    ENSCHK(e, hipEventRecord(e->evB, e->side));
    ENSCHK(e, hipEventRecord(e->evI, e->master));
  }
  hipLaunchKernelGGL(step_counter_kernel, dim3(1), dim3(64), 0, e->side, e->nstep);
  ENSCHK(e, hipEventRecord(e->evB, e->side));
  ENSCHK(e, hipStreamWaitEvent(e->master, e->evB, 0));
  ENSCHK(e, hipGetLastError());
  return FDM_OK;
}

int ens_step_any(fdm_ensemble* e, int p) {
  return e->precision == FDM_FP64 ? ens_step_launches<double>(e, p) : ens_step_launches<float>(e, p);
}

int ens_capture(fdm_ensemble* e, int p) {
  hipGraph_t g = nullptr;
  ENSCHK(e, hipStreamBeginCapture(e->master, hipStreamCaptureModeRelaxed));
  const int rc = ens_step_any(e, p);
  hipError_t st = hipStreamEndCapture(e->master, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  ENSCHK(e, st);
  hipError_t si = hipGraphInstantiate(&e->exec[p], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  ENSCHK(e, si);
  return FDM_OK;
}

// merged work lists of all members
int ens_lists(fdm_ensemble* e) {
  const int M = (int)e->m.size();
  std::vector<int4> tl[3], band, tasks;
  for (int k = 0; k < M; k++) {
    const fdm_ctx* c = e->m[k];
    for (const int4& t : c->htiles[c->tile_shape]) tl[c->tile_shape].push_back(make_int4(t.x, t.y, t.z, k));
    for (const int2& b : c->hband) band.push_back(make_int4(b.x, b.y, k, 1));
    while (band.size() % 64) band.push_back(make_int4(0, 0, k, 0));
    for (const int2& t : c->htasks) tasks.push_back(make_int4(t.x, t.y, k, 1));
    while (tasks.size() % 64) tasks.push_back(make_int4(0, 0, k, 0));
  }
  for (int q = 0; q < 3; q++) {
    e->ntiles[q] = (int)tl[q].size();
    if (e->ntiles[q] && ens_upload(e, &e->tiles[q], tl[q].data(), tl[q].size() * sizeof(int4)))
      return FDM_EDEVICE;
  }
  e->nband = (int)band.size();
  if (e->nband && ens_upload(e, &e->band, band.data(), band.size() * sizeof(int4))) return FDM_EDEVICE;
  e->ntasks = (int)tasks.size();
  if (e->ntasks && ens_upload(e, &e->tasks, tasks.data(), tasks.size() * sizeof(int4))) return FDM_EDEVICE;
  return FDM_OK;
}

}  // namespace

extern "C" {

const char* fdm_ensemble_last_error(const fdm_ensemble* e) { return e ? e->err.c_str() : g_create_error.c_str(); }

void fdm_ensemble_destroy(fdm_ensemble* e) {
  if (!e) return;
  if (e->master) (void)hipStreamSynchronize(e->master);
  if (e->side) (void)hipStreamSynchronize(e->side);
  for (int q = 0; q < 2; q++)
    if (e->exec[q]) (void)hipGraphExecDestroy(e->exec[q]);
  for (hipEvent_t ev : e->pre)
    if (ev) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : {e->evJ, e->evI, e->evB, e->done})
    if (ev) (void)hipEventDestroy(ev);
  for (void* p : e->allocs) (void)hipFree(p);
  if (e->master) (void)hipStreamDestroy(e->master);
  if (e->side) (void)hipStreamDestroy(e->side);
  delete e;
}

int fdm_ensemble_create(fdm_ctx* const* members, int32_t n, fdm_ensemble** out) {
  if (!members || n <= 0 || !out) return FDM_EINVAL;
  *out = nullptr;
  for (int k = 0; k < n; k++) {
    const fdm_ctx* c = members[k];
    if (!c) return FDM_EINVAL;
    const fdm_ctx* c0 = members[0];
    if (c->src.nplace > 0 || c->any_mms || (!c->ld_T.empty() && c->ld_T[0] > 0.0) || c->bulk_mode != 2 ||
        c->het || c->diss || c->pml || c->mode != 2 || c->sbp != 4 || c->nstage != c0->nstage ||
        c->precision != c0->precision || c->device != c0->device) {
      g_create_error =
          "fdm_ensemble_create: members must have no point source, no manufactured solution, "
          "no load ramp, no heterogeneous medium, no artificial dissipation, no PML, plane strain, SBP4, "
          "the tile kernels, and the same RK order, precision and device";
      return FDM_EINVAL;
    }
    for (int q = 0; q < k; q++)
      if (members[q] == c) {
        g_create_error = "fdm_ensemble_create: a context appears twice";
        return FDM_EINVAL;
      }
  }
  if (hipSetDevice(members[0]->device) != hipSuccess) {
    g_create_error = "fdm_ensemble_create: hipSetDevice failed";
    return FDM_EDEVICE;
  }
  fdm_ensemble* e = new fdm_ensemble();
  e->m.assign(members, members + n);
  e->precision = members[0]->precision;
  e->nstage = members[0]->nstage;
  e->flip = e->nstage & 1;
  e->swz = members[0]->xcd_swizzle;
  double ws = 0.0;
  for (fdm_ctx* c : e->m) {
    e->any_plastic = e->any_plastic || c->any_plastic;
    e->any_force = e->any_force || c->any_force;
    ws += 23.0 * (double)c->ps * (double)c->esz;
  }
  e->nt_store = nt_store_choice(ws);
  if (const char* v = getenv("FDM_ENSEMBLE_GRAPH")) e->use_graph = atoi(v) != 0;
  e->pre.resize(n, nullptr);
  bool ok = hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) == hipSuccess;
  for (hipEvent_t* ev : {&e->evJ, &e->evI, &e->evB, &e->done})
    ok = ok && hipEventCreateWithFlags(ev, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; k < n && ok; k++) ok = hipEventCreateWithFlags(&e->pre[k], hipEventDisableTiming) == hipSuccess;
  int rc = ok ? FDM_OK : FDM_EDEVICE;
  if (!rc) rc = ens_lists(e);
  if (!rc) rc = ens_upload(e, &e->nstep, nullptr, sizeof(long long));
  if (rc) {
    g_create_error = "fdm_ensemble_create: " + (e->err.empty() ? std::string("stream/event creation failed") : e->err);
    fdm_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return FDM_OK;
}

int fdm_ensemble_step(fdm_ensemble* e, const double* t, const double* dt, int64_t n_first,
                      int32_t nsteps) {
  if (!e || !t || !dt || nsteps < 0) return FDM_EINVAL;
  const size_t n = e->m.size();
  const int cur = e->m[0]->cur;
  for (size_t k = 0; k < n; k++) {
    fdm_ctx* c = e->m[k];
    if (c->prof) return ens_err(e, FDM_EINVAL, "fdm_ensemble_step: disable member profiling");
    if (c->cur != cur) return ens_err(e, FDM_EINVAL, "fdm_ensemble_step: members differ in buffer parity");
    // the load factor of a member without a ramp is 1 from t = 0 on
    if (t[k] < 0.0 || !(dt[k] > 0.0)) return ens_err(e, FDM_EINVAL, "fdm_ensemble_step: need t >= 0, dt > 0");
  }
  if (!nsteps) return FDM_OK;
  ENSCHK(e, hipSetDevice(e->m[0]->device));
  // the master stream follows each member's earlier work
  for (size_t k = 0; k < n; k++) {
    ENSCHK(e, hipEventRecord(e->pre[k], e->m[k]->stream));
    ENSCHK(e, hipStreamWaitEvent(e->master, e->pre[k], 0));
  }
  bool same = e->t.size() == n;
  for (size_t k = 0; k < n && same; k++) same = e->t[k] == t[k] && e->dt[k] == dt[k];
  if (!same) {
    e->t.assign(t, t + n);
    e->dt.assign(dt, dt + n);
    const int rc = e->precision == FDM_FP64 ? ens_tables<double>(e) : ens_tables<float>(e);
    if (rc) {
      e->t.clear();
      return rc;
    }
  }
  hipLaunchKernelGGL(step_set_kernel, dim3(1), dim3(64), 0, e->master, e->nstep, (long long)n_first);
  ENSCHK(e, hipGetLastError());
  for (int s = 0; s < nsteps; s++) {
    const int p = e->m[0]->cur;
    if (e->use_graph) {
      if (!e->exec[p]) {
        const int rc = ens_capture(e, p);
        if (rc) return rc;
      }
      ENSCHK(e, hipGraphLaunch(e->exec[p], e->master));
    } else {
      const int rc = ens_step_any(e, p);
      if (rc) return rc;
    }
    if (e->flip)
      for (fdm_ctx* c : e->m) c->cur = 1 - c->cur;
  }
  // interior points of a lean last-stage launch hold no DF6 (recomputed on
  // read); the full one (any plastic member) stores it, plastic corrections
  // included.  Members' own streams follow the ensemble's work.
  for (fdm_ctx* c : e->m) c->df6_stale = !e->any_plastic;
  ENSCHK(e, hipEventRecord(e->done, e->master));
  for (fdm_ctx* c : e->m) ENSCHK(e, hipStreamWaitEvent(c->stream, e->done, 0));
  return FDM_OK;
}

int fdm_ensemble_sync(fdm_ensemble* e) {
  if (!e) return FDM_EINVAL;
  ENSCHK(e, hipStreamSynchronize(e->master));
  for (size_t k = 0; k < e->m.size(); k++)
    if (check_fail(e->m[k]) != FDM_OK)
      return ens_err(e, FDM_ENEWTON, "member " + std::to_string(k) + ": " + e->m[k]->err);
  return FDM_OK;
}

}  // extern "C"

// discrete adjoint (include/fdmap_adjoint.h)
#include "../../include/fdmap_adjoint.h"
#include "fdm_adjoint.hpp"
