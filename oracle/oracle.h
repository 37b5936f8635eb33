/* Oracle: CPU restatement of the reference's UVS2D Newton hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so, and only as the checker of the GPU library
 * (rismo2d_amd/csrc), never as the thing measured or shipped.
 *
 * Every function restates one reference routine (file:line given at each definition in
 * rismo_oracle.c) with the same floating-point operation order, so that its results can
 * be compared bit for bit with the reference program's own outputs (see
 * tests/golden/make_golden.py).  The REALPR-dependent routines exist twice: *_f32 has the
 * reference's stock semantics (REALPR = float matrix storage, Defs.h:92) and *_f64 the
 * FP64 semantics (REALPR = double).
 *
 * Indices are 0-based.  CSR arrays: rowptr[n+1], cols[nnz] with each row's diagonal first
 * and the remaining columns in insertion order (IndexMat.cpp:91).
 */
#ifndef RMX_ORACLE_H
#define RMX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

typedef struct {
  double maxAbs[3];      /* max |dU|, |dV|, |dh| of the Newton correction (Update.cpp:42) */
  double avAbs[3];
  double relaxUs, relaxS;
  int    iters;          /* Krylov iterations */
  double accuracy;       /* Krylov relative accuracy reached */
  int    converged;      /* Krylov converged flag */
  int    nrconv;         /* Newton convergence test of this step */
  double scale;          /* ||B||_2 before scaling */
} orc_step_stats;

orc_model* orc_create(int np, const double* x, const double* y, const double* z,
                      int ne, const int* elems, const int* nnd,
                      const double* manning, const double* vt,
                      int n_in, const int* in_segs, int n_out, const int* out_segs,
                      double Q, double S_in, double S_out);
/* Emergent vegetation per element (stem diameter dp, spacing sp; 0: none), used by the
 * nodal friction (Lindner).  NULL arrays remove it. */
void   orc_set_vegetation(orc_model* m, const double* dp, const double* sp);
double orc_lindner(double dp, double sp, double va, double ha, double cb, double vk, double g);
/* Turbulence model: code as $TM_TURBULENCE (1 constant, 2 algebraic; +10 anisotropic
 * (Elder); +100 type vt is a lower bound, +200 recompute at every Newton step), st, KDest,
 * estx, esty per element (NULL: Type.cpp defaults 1.0, 0.6, 0, 0), k-epsilon constants
 * cm, cd.  Returns -1 for other codes. */
int    orc_set_turbulence(orc_model* m, int code, const double* st, const double* kdest,
                          const double* estx, const double* esty, double cm, double cd);
/* friction, then the model's nodal eddy viscosity vt and normal Reynolds stress uu (= vv;
 * uv = 0) of the current state (constant: Turbulence.cpp:110-165, algebraic: 292-352) and,
 * anisotropic, the nodal Elder tensor (996-1090); outputs nullable */
void   orc_turbulence(orc_model* m, double* vt, double* uu);
/* the nodal vt, uu the last prepare or orc_turbulence left (no recomputation) */
void   orc_get_turbulence(const orc_model* m, double* vt, double* uu);
/* the nodal Elder tensor exx, exy, eyy (global axes) the last prepare left */
void   orc_get_elder(const orc_model* m, double* exx, double* exy, double* eyy);
void   orc_destroy(orc_model* m);
/* relaxMethod (0|2|3|4), dry method (0|2|3 marsh, else clip), dryLimit, maxDeltaUV,
 * maxDeltaS, convUV, convS, maxUs */
void   orc_set_params(orc_model* m, int relaxMethod, int dryMethod, double dryLimit,
                      double maxDeltaUV, double maxDeltaS, double convUV, double convS,
                      double maxUs);
/* relaxation methods 3 and 4: relaxTimeFlow (the relaxed time step a cycle starts from;
 * <= 0 or above the time step: the time step) and the stationary cycle's time interval
 * ($TM_INTERVAL); both restart the cycle's relaxed time.  get: the relaxed time steps of
 * U, V and of h, and the one before the latest update */
void   orc_set_relax_time(orc_model* m, double relaxTimeFlow, double interval);
void   orc_get_relax_time(const orc_model* m, double* dt_UV, double* dt_H, double* last_dt);
void   orc_set_state(orc_model* m, const double* U, const double* V, const double* S);
void   orc_get_state(const orc_model* m, double* U, double* V, double* S, double* z);
void   orc_get_nodal(const orc_model* m, double* cf, int* bckind, int* rot,
                     double* rotm /* np*4: r00 r01 r10 r11 */, double* bcU, double* bcS);

/* Time integration.  orc_set_state is an initial condition: it also sets the previous
 * time step's state to the same values with zero time derivatives.  orc_time_step starts
 * a flow time step: dt <= 0 stationary (no time terms), else the theta scheme with
 * EQS_UVS2D::Predict (EqsUVS2D.cpp:1101-1180; first = first time step of the run); the
 * Newton update then recomputes the time derivatives (EQS_UVS2D::Timegrad, :1183-1230).
 * orc_time_commit ends the time step (Compute.cpp:773-796).  Returns -1 for theta or dt
 * below 1e-10 in magnitude. */
int    orc_time_step(orc_model* m, double dt, double theta, int first);
/* QINLET discharge and reference level for the following Newton steps; hydrograph != 0
 * marks the line as QTINLET (its nodes' kind) */
void   orc_set_inflow(orc_model* m, double Q, double S_in, int hydrograph);
/* hydrograph lookup (BCONLINE::FindInterval + interpolation): Q, S at `time` */
void   orc_hydrograph(int n, const double* t, const double* Q, const double* S, double time,
                      double* Qout, double* Sout);
void   orc_time_commit(orc_model* m);
/* previous state (vo) and time derivatives (v.d*dt, vo.d*dt): set (NULL keeps) / get */
void   orc_set_time_state(orc_model* m, const double* Uo, const double* Vo, const double* So,
                          const double* dUdto, const double* dVdto, const double* dSdto,
                          const double* dUdt, const double* dVdt, const double* dSdt);
void   orc_get_time_state(const orc_model* m, double* Uo, double* Vo, double* So,
                          double* dUdto, double* dVdto, double* dSdto,
                          double* dUdt, double* dVdt, double* dSdt);

/* Newton-step preparation (EqsUVS2D.cpp:300-360): friction, boundary conditions, normals,
 * rotation, and (once) numbering + CSR pattern.  Returns neq. */
int    orc_prepare(orc_model* m);
int    orc_neq(const orc_model* m);
long   orc_nnz(const orc_model* m);
void   orc_get_numbering(const orc_model* m, int* eqno /* 3*np, -1 = none */);
void   orc_get_pattern(const orc_model* m, int* rowptr, int* cols);
int    orc_n_model_elems(const orc_model* m);
/* model element k: returns nnd; nodes[] gets the node indices; *is_bound set */
int    orc_model_elem(const orc_model* m, int k, int* nodes, int* is_bound);
/* drying and rewetting, method 2 (MODEL::DoDryRewet, DoDryRew.cpp:50; GRID::DryRewet,
 * DryRewet.cpp:696): returns the elements that got wet (-1: method not 2), *dried those
 * that got dry */
int    orc_dry_rewet(orc_model* m, int* dried);
void   orc_set_dryrew(orc_model* m, double rewetLimit, int countDown);
/* the slip flow condition on the rotated nodes (GRID::SetSlipFlow, SlipFlow.cpp:43) */
void   orc_slip_flow(orc_model* m);
/* per node the reference's dry (64) and marsh (256) bits; per element 1 dry, 2 marsh */
void   orc_get_wetdry(const orc_model* m, int* nflag, int* eflag);
/* element coefficients of model element k (CoefsUVS2D.cpp:49): estifm[24*24], force[24];
 * returns 0 for a skipped (dry) element */
int    orc_coefs(orc_model* m, int k, double* estifm, double* force);

void   orc_assemble_f64(orc_model* m, double* vals, double* B);
void   orc_assemble_f32(orc_model* m, float* vals, double* B);
/* sum over elements of |contribution| for every CSR entry (FP64 instance) */
void   orc_assemble_abs_f64(orc_model* m, double* vals);

double orc_scale_f64(int n, const int* rowptr, double* vals, double* B);
double orc_scale_f32(int n, const int* rowptr, float* vals, double* B);
int    orc_ilu0_factor_f64(int n, const int* rowptr, const int* cols, const double* A, double* ILU);
int    orc_ilu0_factor_f32(int n, const int* rowptr, const int* cols, const float* A, float* ILU);
void   orc_ilu0_solve_f64(int n, const int* rowptr, const int* cols, const double* ILU,
                          const double* B, double* X);
void   orc_ilu0_solve_f32(int n, const int* rowptr, const int* cols, const float* ILU,
                          const double* B, double* X);
void   orc_mulvec_f64(int n, const int* rowptr, const int* cols, const double* A,
                      const double* x, double* r);
void   orc_mulvec_f32(int n, const int* rowptr, const int* cols, const float* A,
                      const double* x, double* r);
int    orc_bicgstab_f64(int n, const int* rowptr, const int* cols, const double* A,
                        const double* ILU, const double* B, double* X, int maxIter, double eps,
                        int* iters, double* accuracy);
int    orc_bicgstab_f32(int n, const int* rowptr, const int* cols, const float* A,
                        const float* ILU, const double* B, double* X, int maxIter, double eps,
                        int* iters, double* accuracy);
int    orc_fgmres_f64(int n, const int* rowptr, const int* cols, const double* A,
                      const double* ILU, const double* B, double* X, int mkyrl, int maxIter,
                      double eps, int* iters, double* accuracy);
int    orc_fgmres_f32(int n, const int* rowptr, const int* cols, const float* A,
                      const float* ILU, const double* B, double* X, int mkyrl, int maxIter,
                      double eps, int* iters, double* accuracy);

/* One Newton step (EqsUVS2D.cpp:189-760): prepare, assemble, scale, ILU(0), Krylov,
 * statistics, relaxation, correction, midside interpolation.  prec: 32 | 64.
 * solver: 1 = P_BCGSTABD, 2 = P_FGMRESD (with the Hessenberg assignment the serial
 * reference omits).  X_out (nullable) receives the Newton correction. */
int    orc_newton_step(orc_model* m, int prec, int solver, int maxIter, double eps, int mkyrl,
                       orc_step_stats* st, double* X_out);

/* shape tables at GPdeg 5 (Shape.cpp:138-395): kind 0 line3, 1 tri6, 2 quad8.
 * f, dfdx, dfdy: [ngp*nnd]; lf, ldfdx, ldfdy: linear shape [ngp*ncn]; w[ngp] */
int    orc_shape(int kind, double* f, double* dfdx, double* dfdy, double* lf, double* ldfdx,
                 double* ldfdy, double* w);

#ifdef __cplusplus
}
#endif
#endif
