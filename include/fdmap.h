/*
 * fdmap.h -- C ABI of the MI355X time-stepping library for FDMAP problems.
 *
 * The library advances the plane-strain (mode 2) or antiplane (mode 3)
 * elastic wave equation on a multi-block curvilinear grid with SBP4
 * first-derivative operators, SAT boundary/interface penalties, rate-and-
 * state or slip-weakening friction (with kinematic forcing) on block
 * interfaces, a singular moment-tensor source, Drucker-Prager plasticity,
 * heterogeneous media, artificial dissipation, perfectly matched layers and
 * low-storage Runge-Kutta time integration.  Everything lives in device
 * memory; host pointers handed to any call are read or written only during
 * that call.
 *
 * The reference implements this path as Fortran module procedures that take
 * assumed-shape arrays (not C-interoperable).  Each entry point below names
 * the reference routine(s) it replaces.  Paths are relative to the reference
 * repository (jaewooan/fdmap_github).
 *
 * Conventions
 *  - Global grid of nx*ny points, 0-based indices (reference: 1-based).
 *  - 2-D arrays are stored x-fastest with leading dimension nx, exactly the
 *    Fortran column-major order of the reference's F(i,j,l) without ghosts.
 *  - Field planes: l = 0..5 is (vx, vy, sxx, sxy, syy, szz) (fields.f90:166).
 *  - Per-node side/interface arrays are component-major: a "[c][n]" array
 *    holds component c of node k at index c*n+k (Fortran A(m:p,c) order).
 *  - Every call returns FDM_OK (0) or a negative FDM_E* code; the message is
 *    available from fdm_last_error().  Nothing unwinds across the ABI.
 *  - One host thread and one HIP stream per context.  fdm_step may return
 *    before the device finishes; every fdm_get_* call synchronises.
 */
#ifndef FDMAP_H
#define FDMAP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDM_NF 6
#define FDM_MAX_STAGES 8
#define FDM_MAX_SOURCE_PLACES 16

/* return codes */
#define FDM_OK 0
#define FDM_EINVAL (-1)   /* malformed problem description or argument */
#define FDM_EDEVICE (-2)  /* HIP runtime error, or no usable GPU */
#define FDM_ENOMEM (-3)   /* device allocation failed */
#define FDM_ENEWTON (-4)  /* friction Newton solve failed (friction.f90:1012-1036) */

/* storage/arithmetic precision of the bulk fields */
enum { FDM_FP64 = 0, FDM_FP32_BULK = 1 };

/* boundary conditions (boundaries.f90:245-357) */
enum {
  FDM_BC_NONE = 0,            /* 'none': side is an interface or unused     */
  FDM_BC_ABSORBING = 1,       /* 'absorbing'                                */
  FDM_BC_FREE = 2,            /* 'free'                                     */
  FDM_BC_RIGID = 3,           /* 'rigid'                                    */
  FDM_BC_ABSORBING_VELOCITY = 4, /* 'absorbing-velocity'                    */
  FDM_BC_FREE_VELOCITY = 5,   /* 'free-velocity'                            */
  FDM_BC_ABSORBING_0 = 6,     /* 'absorbing-0' (no initial-field removal)   */
  FDM_BC_FREE_0 = 7,          /* 'free-0'                                   */
  FDM_BC_RIGID_0 = 8,         /* 'rigid-0'                                  */
  FDM_BC_MMS_SIN_W = 9        /* 'mms-sin-w' / 'inplane-fault-mms-w'        */
};

enum { FDM_SIDE_L = 0, FDM_SIDE_R = 1, FDM_SIDE_B = 2, FDM_SIDE_T = 3 };

/* interface couplings (interfaces.f90:541-647) */
enum { FDM_COUPLING_LOCKED = 0, FDM_COUPLING_FRICTION = 1 };

/* friction laws (friction.f90:702-826) */
enum {
  FDM_LAW_FRICTIONLESS = 0,
  FDM_LAW_LOCKED = 1,
  FDM_LAW_SW = 2,   /* slip weakening                                  */
  FDM_LAW_SL = 3,   /* slip law, log form                               */
  FDM_LAW_FL = 4,   /* flash-heating law, log form                      */
  FDM_LAW_RSL = 5,  /* regularised slip law                             */
  FDM_LAW_RFL = 6,  /* regularised slip law with strong weakening       */
  FDM_LAW_RSF = 7,  /* regularised, state-form steady state             */
  FDM_LAW_RAL = 8   /* regularised aging law                            */
};

/* source time functions (source.f90:212-223) */
enum { FDM_STF_RAMP = 0, FDM_STF_SINUSOID = 1, FDM_STF_GAUSSIAN = 2, FDM_STF_RICKER = 3 };

/* interface fields, in the order of the reference checkpoint files
   (friction.f90:528-546) */
enum {
  FDM_IF_D = 0, FDM_IF_PSI, FDM_IF_DPSI, FDM_IF_TRUP, FDM_IF_V, FDM_IF_O,
  FDM_IF_S, FDM_IF_N, FDM_IF_S0, FDM_IF_N0, FDM_IF_DS, FDM_IF_DN, FDM_IF_DDS,
  FDM_IF_DDN, FDM_IF_W, FDM_IF_DW, FDM_IF_NREG, FDM_IF_DNREG, FDM_IF_NFIELDS
};

/* One side of a block (the reference's bnd_fields + curve, fields.f90:5-20,
   geometry.f90:9-12). */
typedef struct fdm_side {
  int32_t bc;            /* FDM_BC_*                                          */
  int32_t n;             /* number of nodes along the side                    */
  const double* normal;  /* [2][n] outward unit normal                        */
  const double* M;       /* [5][n] Zs, Zp, gamma, Ks, Kp (domain.f90:1373-1377) */
  const double* F0;      /* [6][n] initial fields on the side                 */
  const double* x;       /* [n] coordinates (MMS boundary data only; may be 0) */
  const double* y;       /* [n]                                               */
} fdm_side;

/* One block (block_grid + block_material, grid.f90:12-24, material.f90:17-24). */
typedef struct fdm_block {
  int32_t mx, px, my, py;        /* inclusive 0-based global index limits     */
  int32_t plastic;               /* 1: Drucker-Prager response 'plastic'      */
  int32_t mms;                   /* 1: 'mms-sin' body source in this block    */
  int32_t id;                    /* reference block number (1-based), MMS only */
  int32_t pad_;
  double rho, G, M, gamma, nu;   /* material.f90:100-120                      */
  double mu, beta, b, h, eta, K; /* plasticity (material.f90:104-115)         */
  double s0[6];                  /* initial stress sxx,sxy,sxz,syy,syz,szz used
                                    by plasticity (fields.f90:1007-1058)      */
  fdm_side side[4];              /* L, R, B, T                                */
  double Cdiss;                  /* artificial dissipation as the rates use it,
                                    Cdiss*CFL/dt (domain.f90:170-180, 316-320);
                                    0: none                                   */
  /* Perfectly matched layer (pml_type, material.f90:9-15, 152-223): with
     pmlx or pmly set (mode 2, no dissipation in the block) every point of
     the block takes the Cartesian rates of rates_interior_mode2_pml and
     rates_boundary_mode2_pml (rates.f90:2024-2619) and the auxiliary fields
     Wx, Wy (6 components each, fields.f90:35, 203-209) of fdm_get_pml.
     pml = n, sigma, kappa, mx, px, my, py: with n > 0 the damping is
     sigma*max(0, mx-x, x-px)^n in x (likewise in y), else sigma. */
  int32_t pmlx, pmly;
  double pml[7];
} fdm_block;

/* One block interface (iface_type + fr_type, interfaces.f90:9-20,
   friction.f90:38-49).  Node k is global index m+k along the interface. */
typedef struct fdm_iface {
  int32_t direction;      /* 0: 'x' (m's R side meets p's L side); 1: 'y' (T/B) */
  int32_t blockm, blockp; /* 0-based block indices                             */
  int32_t m, p;           /* inclusive global index range along the interface  */
  int32_t coupling;       /* FDM_COUPLING_*                                    */
  int32_t law;            /* FDM_LAW_*                                         */
  int32_t reg;            /* regularised normal stress (friction.f90:678-682)  */
  int32_t rup_field;      /* 0: 'V', 1: 'D' (friction.f90:1267-1290)           */
  int32_t pad_;
  double rup_threshold, skempton, flock, ld_T;
  const double* nhat;     /* [2][n] unit normal pointing from block m to p     */
  const double* rs;       /* [7][n] a,b,V0,f0,L,fw,Vw   (rate-and-state laws)  */
  const double* sw;       /* [4][n] c0,fs,fd,Dc         (slip weakening)       */
  const double* S0base;   /* [n] background shear stress from file, else 0     */
  const double* N0base;   /* [n] background normal stress from file, else 0    */
  const double* ldS;      /* [n] ld%S * shape(x) (friction.f90:1373)           */
  const double* ldN;      /* [n] ld%N * shape(x) (friction.f90:1374)           */
  const int32_t* outside; /* [n] 1: node outside [xlockm,xlockp] (:686-698)    */
  const double* Psi0;     /* [n] initial state variable                        */
  /* kinematic forcing (force = T, friction.f90:717-745, 765-767, 808-815):
     kn = c0, fd, dfdx, vL, vR, xL, xR, tmax; kn_x [n] the distance along the
     fault, x cos(angle) + y sin(angle), of every node (NULL without forcing) */
  int32_t force;
  int32_t pad2_;
  double kn[8];
  const double* kn_x;
} fdm_iface;

/* A placement of the singular point source in one block
   (source.f90:196-254); dx[k], dy[k] are the discrete delta weights of
   singular_source (source.f90:327-356) at i0-2+k, j0-2+k. */
typedef struct fdm_source_place {
  int32_t block, i0, j0, pad_;
  double dx[5], dy[5];
} fdm_source_place;

typedef struct fdm_source {
  int32_t nplace;           /* 0: no singular source                         */
  int32_t stf;              /* FDM_STF_*                                     */
  double t_center, t_duration;
  double Fx, Fy, Mxx, Mxy, Myy;
  fdm_source_place place[FDM_MAX_SOURCE_PLACES];
} fdm_source;

/* Complete description of one problem (the reference's domain_type,
   domain.f90:24-39, restricted to the time-stepping path). */
typedef struct fdm_problem {
  int32_t nx, ny;
  int32_t nblocks, nifaces;
  int32_t precision;        /* FDM_FP64 or FDM_FP32_BULK                    */
  int32_t nstage;           /* RK stages (time_step.f90:14-68)              */
  double rkA[FDM_MAX_STAGES], rkB[FDM_MAX_STAGES];
  double rkC[FDM_MAX_STAGES + 1];   /* C(0..nstage)                        */
  const double* xq;         /* [ny][nx] metrics (grid.f90:465-534)          */
  const double* xr;
  const double* yq;
  const double* yr;
  const double* Ji;         /* 1/J                                          */
  const double* x;          /* [ny][nx] coordinates (MMS sources; may be 0) */
  const double* y;
  const fdm_block* blocks;
  const fdm_iface* ifaces;
  fdm_source source;
  /* Heterogeneous medium (material.f90:226-275): density and wave speeds
     at every point ([ny][nx]), or all three NULL.  When given, the rates of
     every block use them (rates_heterogeneous.f90:375-917), as the reference
     does once any block has read a medium file; so do plasticity and the
     szz relation.  Edge coefficients stay those of fdm_side.M, which the
     host fills per block (domain.f90:1379-1456), and sources keep the
     block's density. */
  const double* mat_rho;
  const double* mat_cs;
  const double* mat_cp;
  /* Jacobian J (grid.f90:465-534), [ny][nx]; needed when a block has
     Cdiss > 0 (rates_dissipation_*_mode2, rates.f90:1397-2021), else may be
     NULL. */
  const double* J;
  /* 2: plane strain, fields vx vy sxx sxy syy szz; 3: antiplane shear,
     fields vz sxz syz (domain.f90:220-227, fields.f90:158-175); 0 means 2.
     In mode 3 the field arrays of fdm_set_fields and fdm_get_fields are
     [3][ny][nx], fdm_side.F0 holds 3 components, fdm_side.M uses Zs (M[0])
     and the SAT coefficient Ks (M[3]), and fdm_block.s0 holds the in-plane
     prestress sxx, sxy, -, syy, -, szz that frictional interfaces resolve
     onto the fault (interfaces.f90:859-876).  Mode 3 covers elastic
     homogeneous blocks without sources or dissipation. */
  int32_t mode;
  /* first-derivative operator, FDmethod of domain_list (fd_coeff.f90:
     41-248): 2 'SBP2', 4 'SBP4', 6 'SBP6'; 0 means 4.  The metrics of the
     problem and the SAT coefficients (H00i) must come from the same
     operator. */
  int32_t sbp;
} fdm_problem;

/* Kernel timing gathered with HIP events on the context's stream. */
typedef struct fdm_profile_t {
  double bulk_ms;        /* summed duration of the bulk (stencil+RK) launches
                            that update the block interiors: the interior
                            tiles and, in PML blocks, the layer points four
                            or more from every block edge                  */
  int64_t bulk_launches;
  double edge_ms;        /* summed duration of edge/interface launches      */
  int64_t edge_launches;
  int64_t bulk_points;   /* grid points updated per stage by those launches */
  double band_ms;        /* summed duration of the launches that update the
                            points within four of a block edge, PML blocks'
                            included (0 when the bulk launch covers every
                            point)                                         */
  int64_t band_launches;
} fdm_profile_t;

typedef struct fdm_ctx fdm_ctx;

/* Allocate device storage and upload the problem.  Replaces the device-side
   part of init_domain (domain.f90:45-378): metrics, edge coefficients,
   interface and friction parameters. */
int fdm_create(const fdm_problem* prob, fdm_ctx** ctx);

/* Release everything (finish_domain, domain.f90:381-417). */
void fdm_destroy(fdm_ctx* ctx);

/* Message for the last failed call on ctx (ctx may be NULL: last failed
   fdm_create).  Replaces io.f90:167-206 error(). */
const char* fdm_last_error(const fdm_ctx* ctx);

/* Fields F and rates DF, [6][ny][nx] FP64 (fields.f90:180-182).  DF may be
   NULL: rates are then filled with 1e40 as allocate_array_body does
   (mpi_routines2d.f90:583-587). */
int fdm_set_fields(fdm_ctx* ctx, const double* F, const double* DF);
int fdm_get_fields(fdm_ctx* ctx, double* F, double* DF);

/* Output streaming.  The reference writes its body and block outputs from
   host arrays at every output step, with the time stepping waiting on it
   (domain.f90:502-894, output.f90).  Here fdm_snapshot queues, after the
   work queued so far, a device copy of the fields selected by field_mask
   (bit l: field l of the mode, vx vy sxx sxy syy szz, or vz sxz syz) into a
   staging slot of the context, then an asynchronous copy of that slot to
   pinned host memory on a copy stream of its own, and returns the slot in
   *slot without waiting: later fdm_step calls run while the copy moves.
   fdm_snapshot_read waits for that slot's copy and writes the selected
   fields to out, [nselected][ny][nx] FP64 in field order, as fdm_get_fields
   would have returned them at the snapshot (FP32-bulk adds each block's
   background); the slot is then free.  It does not wait for later work, so
   it does not report a failed Newton solve (fdm_sync does).  There are
   FDM_SNAPSHOT_SLOTS slots; fdm_snapshot fails with FDM_EINVAL when all are
   taken.  A slot's device and pinned host buffers (all fields of the mode)
   are allocated at its first use and kept until fdm_destroy. */
#define FDM_SNAPSHOT_SLOTS 4
int fdm_snapshot(fdm_ctx* ctx, uint32_t field_mask, int32_t* slot);
int fdm_snapshot_read(fdm_ctx* ctx, int32_t slot, double* out);

/* Initial interface solve with dt = 0 at time t: prepare_edges +
   enforce_edge_conditions(D,0d0) at the end of init_domain
   (domain.f90:346-356), including set_rates_friction. */
int fdm_init_edges(fdm_ctx* ctx, double t);

/* nsteps full low-storage RK steps: time_step_LS (time_step.f90:71-167)
   called from the main loop (main.f90:168-231) for steps
   n = n_first .. n_first+nsteps-1, each starting at t + (n-1)*dt
   (main.f90:176).  Includes the plasticity split step.
   The call queues the work and returns.  A failed friction Newton solve
   (the reference stops there, friction.f90:1012-1036) sets a device flag
   that the next synchronising call (fdm_sync, fdm_init_edges, every
   fdm_get_*, the slip-velocity maxima) reports as FDM_ENEWTON, with the
   requested data still copied out; the stages queued after
   the failing one still run, on the failed state.  A caller that must stop
   at the failing step, as the reference does, steps one step per call and
   calls fdm_sync after each. */
int fdm_step(fdm_ctx* ctx, double t, double dt, int64_t n_first, int32_t nsteps);

/* Final step of the main loop (final_step = .true., time_step.f90:104-133):
   stage-1 rates, boundary and interface values at t + (n-1)*dt, no update. */
int fdm_stage1(fdm_ctx* ctx, double t, double dt, int64_t n);

/* Interface arrays (length p-m+1), field = FDM_IF_*. */
int fdm_get_iface(fdm_ctx* ctx, int32_t iface, int32_t field, double* out);
int fdm_set_iface(fdm_ctx* ctx, int32_t iface, int32_t field, const double* in);
/* V, O, S, N, S0, N0 as computed at stage 1 of the most recent step: the
   values the reference writes as output (time_step.f90:129). */
int fdm_get_iface_stage1(fdm_ctx* ctx, int32_t iface, int32_t field, double* out);

/* Plastic strain gammap and plastic strain rate lambda, [ny][nx]
   (fields.f90:194-196). */
int fdm_get_plastic(fdm_ctx* ctx, double* gammap, double* lambda);

/* Restore gammap and lambda, [ny][nx], from a checkpoint
   (checkpoint_fields 'read', fields.f90:380-381); either may be NULL.
   FDM_EINVAL when no block is plastic. */
int fdm_set_plastic(fdm_ctx* ctx, const double* gammap, const double* lambda);

/* PML auxiliary fields Wx and Wy, [6][ny][nx] each (fields.f90:35,
   203-209; zero outside the layer blocks), for checkpoints
   (checkpoint_fields, fields.f90:383-384, 406-407).  Either may be NULL;
   FDM_EINVAL when one is asked for and no block has that direction. */
int fdm_get_pml(fdm_ctx* ctx, double* Wx, double* Wy);
int fdm_set_pml(fdm_ctx* ctx, const double* Wx, const double* Wy);

/* max over all frictional interfaces of V (check_nucleation,
   domain.f90:474-499, interfaces.f90:506-517), reduced on the device in a
   buffer that fdm_create allocates with the context. */
int fdm_max_slip_velocity(fdm_ctx* ctx, double* vmax);

/* The same maximum, bit for bit, from a call that allocates nothing: its
   partial maxima live in a fixed device array of the library (520 bytes per
   device, held until the process ends) instead of context memory.  No
   counterpart in the reference: it serves callers that cannot allocate
   device memory (tight serving loops, harnesses that forbid allocation).
   Limit: 16384 frictional interface nodes per context (FDM_EINVAL beyond,
   use fdm_max_slip_velocity).  Every call on a device shares that device's
   array, so calls that overlap on one device race; consecutive calls do
   not depend on each other. */
int fdm_max_slip_velocity_noalloc(fdm_ctx* ctx, double* vmax);

/* Wait for all queued work. */
int fdm_sync(fdm_ctx* ctx);

/* Kernel timing: enable (1) / disable (0) and reset; read the totals. */
int fdm_profile_enable(fdm_ctx* ctx, int32_t enable);
int fdm_profile_read(fdm_ctx* ctx, fdm_profile_t* out);

#ifdef __cplusplus
}
#endif

#endif /* FDMAP_H */
