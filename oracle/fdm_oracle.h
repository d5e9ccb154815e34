/*
 * fdm_oracle.h -- CPU restatement of the reference's FP64 time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * library declared in include/fdmap.h: it consumes the same fdm_problem
 * description and reproduces the reference's algorithm loop by loop, in the
 * reference's left-to-right evaluation order, compiled without FMA
 * contraction.  It is linked only by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py.  Nothing in the product calls it.
 */
#ifndef FDM_ORACLE_H
#define FDM_ORACLE_H

#include "../include/fdmap.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

int orc_create(const fdm_problem* prob, orc_ctx** ctx);
void orc_destroy(orc_ctx* ctx);
const char* orc_last_error(const orc_ctx* ctx);
int orc_set_fields(orc_ctx* ctx, const double* F, const double* DF);
int orc_get_fields(orc_ctx* ctx, double* F, double* DF);
int orc_init_edges(orc_ctx* ctx, double t);
int orc_step(orc_ctx* ctx, double t, double dt, int64_t n_first, int32_t nsteps);
int orc_stage1(orc_ctx* ctx, double t, double dt, int64_t n);
int orc_get_iface(orc_ctx* ctx, int32_t iface, int32_t field, double* out);
int orc_set_iface(orc_ctx* ctx, int32_t iface, int32_t field, const double* in);
int orc_get_iface_stage1(orc_ctx* ctx, int32_t iface, int32_t field, double* out);
int orc_get_plastic(orc_ctx* ctx, double* gammap, double* lambda);
int orc_get_pml(orc_ctx* ctx, double* Wx, double* Wy);
int orc_set_pml(orc_ctx* ctx, const double* Wx, const double* Wy);
/* the PML rates DWx, DWy, [6][ny][nx] each, as the last stage left them
   (for tests; the device keeps them internal) */
int orc_get_pml_rates(orc_ctx* ctx, double* DWx, double* DWy);
/* OpenMP threads used by orc_step (1 = scalar restatement). */
int orc_set_threads(int32_t nthreads);

/* ---- component entry points used by the oracle's own pinning tests ---- */

/* SBP4 tables (fd_coeff.f90:97-126, 242-246): DI[5] = DI(-2..2),
   DL[4][6] = DL(0..5, r) row r, HL[4], H00i. */
void orc_sbp4_tables(double* DI, double* DL, double* HL, double* H00i);

/* Tables of 'SBP2', 'SBP4' or 'SBP6' (order 2, 4, 6; fd_coeff.f90:79-246):
   DI[2 pI + 1] = DI(-pI..pI), DL[nbst][nbnd] = DL(0..nbnd-1, r) row r,
   HL[nbst], H00i.  Returns nbnd*100 + nbst*10 + pI (0: unknown order). */
int32_t orc_sbp_tables(int32_t order, double* DI, double* DL, double* HL, double* H00i);

/* 1-D SBP4 first derivative with unit spacing over n points (fd.f90:42-80
   with both ends inside the range); orc_diff_op with the operator of the
   given order. */
void orc_diff(int32_t n, const double* f, double* Df);
void orc_diff_op(int32_t order, int32_t n, const double* f, double* Df);
void orc_diss1d(int32_t n, const double* f, double* Af);

/* LSRK coefficients (time_step.f90:25-66): returns nstage, fills A, B,
   C[0..nstage]. */
int32_t orc_rk_coeffs(int32_t order, double* A, double* B, double* C);

/* One frictional-interface node solve (solve_friction, friction.f90:593-828,
   without load_stress): inputs S0 (total shear prestress), N0, phip, phis,
   eta, D, Psi, V (initial guess), dt, law and the 7 rate-and-state / 4
   slip-weakening parameters; outputs V, S, N, O.  Returns 0 or FDM_ENEWTON. */
int orc_solve_friction_node(int32_t law, const double* rs, const double* sw,
                            double S0, double N0, double phip, double phis,
                            double eta, double D, double Psi, double skempton,
                            double dt, double* V, double* S, double* N, double* O);

/* state_rate (friction.f90:1147-1264) for one node. */
double orc_state_rate(int32_t law, const double* rs, double V, double Psi);

/* mms_sin (mms.f90:52-234); field index: 0 vx,1 vy,2 sxx,3 sxy,4 syy,5 szz,
   6..11 the sources s_vx..s_szz, 12 V, 13 S, 14 N. */
double orc_mms_sin(double x, double y, double t, int32_t side, int32_t field);

/* plastic_flow (plastic.f90:62-192) for one point: s[4] = (sxx,sxy,syy,szz)
   stress change, Ds[4] rates, in place. */
void orc_plastic_point(double G, double K, double mu, double beta, double b,
                       double h, double eta, const double* s0, double dt,
                       double* s, double* Ds, double* lambda, double* gammap);

#ifdef __cplusplus
}
#endif

#endif
