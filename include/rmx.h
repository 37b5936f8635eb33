/* rmx: MI355X (gfx950) library for the Rismo2D UVS2D flow-solve hot path.
 *
 * C ABI, plain pointers and sizes.  Every entry point returns an rmx_status (0 = ok,
 * < 0 = error; rmx_last_error() gives the message of the calling thread's last error).
 * A context owns one HIP stream and all device state of one equation system; calls on
 * one context must not overlap (the reference is single-threaded as well).  Host
 * pointers are caller-owned and only read or written during the call.
 *
 * Two groups of entry points:
 *  - the drop-in group replaces the reference's EQS::Coefs + CRSMAT assembly, scaling
 *    and MulVec, PRECON::Factor / Solve and SOLVER::Iterate (cited at each function);
 *  - the model group is the host-side set-up the reference does in EQS_UVS2D::Execute
 *    around EQS::Solve (friction, boundary conditions, numbering, pattern, Newton update),
 *    used by callers that do not keep the reference's own host code.
 */
#ifndef RMX_H
#define RMX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rmx_ctx rmx_ctx;
typedef int rmx_status;

enum {
  RMX_OK = 0,
  RMX_ERR_ARG = -1,       /* invalid argument or call order */
  RMX_ERR_HIP = -2,       /* HIP runtime error */
  RMX_ERR_PIVOT = -3,     /* ILU(0) met a pivot below 1e-20 that no later row can repair
                             (Preco_ilu0.cpp:95-135) */
  RMX_ERR_NODEVICE = -4,  /* no gfx950 device */
  RMX_ERR_RANGE = -5      /* size outside what the kernels support (row width > 128) */
};

enum { RMX_F64 = 64, RMX_F32 = 32 };           /* matrix / factor storage precision */
enum { RMX_BICGSTAB = 1, RMX_FGMRES = 2 };      /* Krylov method (SOLVER types) */
/* preconditioners: ILU(0) in the reference's equation order; the same factors swept by
 * independent blocks (strip ILU(0)); strip ILU(0) after a coarse-space correction
 * (two-level); the learned message-passing network */
enum { RMX_PREC_NONE = 0, RMX_PREC_ILU0 = 1, RMX_PREC_MPNN = 2, RMX_PREC_SILU0 = 3,
       RMX_PREC_SILU0C = 4 };

const char* rmx_last_error(void);

/* Create a context on HIP device `device` (must be gfx950). */
rmx_status rmx_create(int device, rmx_ctx** out);
rmx_status rmx_destroy(rmx_ctx* ctx);

/* ------------------------------------------------------------------------------------ */
/* model group                                                                           */

typedef struct {
  int np;                     /* region nodes */
  const double *x, *y, *z;    /* coordinates and bed elevation (zor) */
  int ne;                     /* region elements */
  const int* elems;           /* [ne*8], 0-based, -1 padded; tri: c0 c1 c2 m01 m12 m20 */
  const int* nnd;             /* [ne] 6 (triangle) or 8 (quad) */
  const double* manning;      /* [ne] Manning n */
  const double* vt;           /* [ne] constant eddy viscosity */
  int n_inlet;                /* QINLET control-line segments */
  const int* inlet_segs;      /* [n_inlet*3] corner, corner, midside */
  int n_outlet;               /* OUTLET control-line segments */
  const int* outlet_segs;
  double Q, S_in, S_out;      /* inlet discharge, inlet reference level, outlet level */
  const double* veg_dp;       /* [ne] emergent vegetation stem diameter (m), nullable */
  const double* veg_sp;       /* [ne] stem spacing (m), nullable; dp or sp < 1e-4: none */
  const double* turb_st;      /* [ne] st of the algebraic eddy viscosity model, nullable: 1.0
                                 ($TURBULENCE no vt st estx esty KDest, Type.cpp:681, defaults :73) */
  const double* turb_kdest;   /* [ne] KDest, nullable: 0.6 */
  const double* turb_estx;    /* [ne] estx, esty of the anisotropic (Elder) model, nullable:
                                 0 (Turbulence.cpp:1024-1025 divides both by KDest) */
  const double* turb_esty;
} rmx_mesh_desc;

typedef struct {
  int relax_method;           /* 0: none, 2: cap by max_delta_uv / max_delta_s, 3 and 4 (the
                                 reference's default): time-relaxed Newton steps, a relaxed
                                 time step in the Jacobian that starts from relax_time at every
                                 rmx_set_params and rmx_time_step and grows by the ratio of the
                                 maximum change to the largest correction up to the time step
                                 (rmx_time_step's dt, or interval for stationary flow); the
                                 Newton test fails until it gets there (EqsUVS2D.cpp:97-147,
                                 540-627); the correction is capped as by method 2 */
  int dry_method;             /* 0,2,3: marsh bed lowering; else depth clip; 2: the drying and
                                 rewetting method of rmx_dry_rewet */
  double dry_limit, max_delta_uv, max_delta_s, conv_uv, conv_s, max_us;
  int turb;                   /* turbulence model, the $TM_TURBULENCE code (Timeint.cpp:1333):
                                 0 or 1 constant vt of the element, 2 algebraic model
                                 (Turbulence.cpp:292); + 10: anisotropic eddy viscosity
                                 after Elder (Turbulence.cpp:996-1090, the element terms of
                                 CoefsUVS2D_AI.cpp); + 100: the element's vt is a lower
                                 bound (kVtMin); + 200: nodal vt recomputed at every Newton
                                 step (kVtIterat), else at the first step after a mesh,
                                 state or parameter upload or rmx_turbulence */
  double cm, cd;              /* k-epsilon constants ($KDCONST), <= 0: 0.09, 1.0 */
  double relax_time;          /* methods 3 and 4: relaxTimeFlow, the relaxed time step a cycle
                                 starts from ($TM_INTERVAL's second value; <= 0 or above the
                                 time step: the time step) */
  double interval;            /* methods 3 and 4, stationary flow: the cycle's time step
                                 ($TM_INTERVAL's first value) */
  double rewet_limit;         /* drying and rewetting ($DRYREW rewetLimit): the water level a
                                 rewetted marsh node takes is capped at zor + rewet_limit (not
                                 below dry_limit) */
  int count_down;             /* $DRYREW countDown: calls a node stays dry after its depth
                                 last fell below dry_limit (>= 0) */
} rmx_params;

/* Mesh, boundary lines and material data (MODEL::Input, MODEL::Initialize,
 * Model.cpp:121-215; GRID::AreaFactors, ArFact.cpp:41).  Vegetation enters the nodal
 * friction after Lindner (TYPE::bottom, Friction.cpp:101-108; TYPE::lindner,
 * Lindner.cpp:72-262).  Builds the boundary line
 * elements and the model element list on the host and uploads node/element arrays. */
rmx_status rmx_mesh_upload(rmx_ctx* ctx, const rmx_mesh_desc* mesh);
rmx_status rmx_set_params(rmx_ctx* ctx, const rmx_params* p);
rmx_status rmx_state_upload(rmx_ctx* ctx, const double* U, const double* V, const double* S);
rmx_status rmx_state_download(rmx_ctx* ctx, double* U, double* V, double* S, double* z);

/* Per Newton step set-up (EqsUVS2D.cpp:300-360): nodal friction (MODEL::DoFriction), the
 * algebraic model's nodal eddy viscosity when it is due (GRID::Turbulence) and the
 * inlet discharge distribution (BCONSET::InitBcon) on the device; boundary kinds,
 * normals, rotations (MODEL::SetLocation/SetNormal/SetRotation), equation numbering
 * (EQS::SetEqno), CSR pattern (EQS::SetIndexMat), element colouring and ILU level sets on
 * the host on the first call.  Returns the number of equations and non-zeros. */
rmx_status rmx_prepare(rmx_ctx* ctx, int* neq, long long* nnz);
rmx_status rmx_get_numbering(rmx_ctx* ctx, int* eqno /* [3*np] */);
rmx_status rmx_get_pattern(rmx_ctx* ctx, int* rowptr /* [neq+1] */, int* cols /* [nnz] */);
rmx_status rmx_get_nodal(rmx_ctx* ctx, double* cf, double* bcU, int* kind, double* rotm);
rmx_status rmx_n_model_elems(rmx_ctx* ctx, int* n);

/* Nodal friction, then the algebraic model's nodal eddy viscosity vt and normal Reynolds
 * stress uu (= vv; uv = 0) from the current state: replaces MODEL::DoFriction followed by
 * GRID::Turbulence (Turbulence.cpp:292-352), as the flow cycle calls them at its first
 * Newton step (EqsUVS2D.cpp:307-315).  Requires params.turb = 2 (+100/+200); vt, uu are
 * [np] host arrays, nullable. */
rmx_status rmx_turbulence(rmx_ctx* ctx, double* vt, double* uu);

/* Time integration of the flow cycle.  A state upload is an initial condition: the
 * previous time step's state becomes the same, time derivatives zero (GRID::InputInit,
 * Grid.cpp:1119-1129).  rmx_time_step starts a time step: dt <= 0 stationary flow (no
 * time terms, EqsUVS2D.cpp:105-127), else the theta scheme: 1 / (dt theta) on the time
 * terms' diagonal (:130-147) and EQS_UVS2D::Predict (:1101-1180; first = first time step
 * of the run); every Newton update then recomputes the time derivatives
 * (EQS_UVS2D::Timegrad, :1183-1230).  The algebraic eddy viscosity becomes due.
 * rmx_time_commit ends a time step: the state becomes the previous one (Compute.cpp:
 * 773-796). */
rmx_status rmx_time_step(rmx_ctx* ctx, double dt, double theta, int first);

/* Inlet discharge and reference water level of the QINLET line for the following Newton
 * steps: a hydrograph (QTINLET) gives them per time step (BCONLINE::Inlet,
 * BconLine.cpp:263-276, at the time step's end time, EqsUVS2D.cpp:318-320); hydrograph != 0
 * gives the line's nodes the QTINLET kind as well (BconLine.cpp:86). */
rmx_status rmx_set_inflow(rmx_ctx* ctx, double Q, double S_in, int hydrograph);
/* Hydrograph lookup of BCONLINE::FindInterval (BconLine.cpp:167-226) with BconLine.cpp:
 * 268-275: the first interval [t_i, t_i+1] (widened by 0.1 %) that holds `time` gives Q, S
 * by linear interpolation; outside every interval Q[0], S[0].  No context needed. */
rmx_status rmx_hydrograph(int n, const double* t, const double* Q, const double* S, double time,
                          double* Qout, double* Sout);
rmx_status rmx_time_commit(rmx_ctx* ctx);
/* The slip condition at the end of the flow cycle (GRID::SetSlipFlow, SlipFlow.cpp:43-78,
 * called at the end of EQS_UVS2D::Execute, EqsUVS2D.cpp:1078): at rotated nodes (walls,
 * inlet and outlet) the velocity and its time derivative keep only their component along
 * the first column of the node's rotation.  Needs rmx_prepare. */
rmx_status rmx_slip_flow(rmx_ctx* ctx);
/* Drying and rewetting, method 2 (MODEL::DoDryRewet, DoDryRew.cpp:50-238, with
 * GRID::DryRewet, DryRewet.cpp:696-900), on the device: nodes shallower than dry_limit are
 * dry (and stay so for count_down calls), elements with all corners dry are dry, with some
 * marsh; marsh nodes that were dry take the level of their wet neighbours (capped at zor +
 * rewet_limit) and a lowered bed; dry nodes are reset.  When elements got dry or wet the wet
 * model is set up again as MODEL::Initialize, SetNormal and SetRotation do (Model.cpp:
 * 121-215, Transform.cpp): boundary line elements on the wet/dry line, boundary conditions
 * and rotations; the equations are numbered without the dry nodes and the pattern, assembly
 * schedule, ILU plans and coarse space rebuilt, then the slip condition applied
 * (GRID::SetSlipFlow).  The nodal eddy viscosity becomes due.  dried, wetted: the elements
 * that got dry and wet (nullable).  Needs params.dry_method 2; the flow cycle calls it at
 * its first Newton step and every dryRewFreq steps or after a converged one
 * (EqsUVS2D.cpp:200-222).  rmx_time_commit keeps the flags as the previous time step's,
 * which Predict and Timegrad read (Compute.cpp:773-783, EqsUVS2D.cpp:1153, 1193-1202). */
rmx_status rmx_dry_rewet(rmx_ctx* ctx, int* dried, int* wetted);
/* Per node the reference's dry (64) and marsh (256) bits and those of the previous time step
 * (128, 512; Node.h:100-103); per region element 1 dry, 2 marsh.  Either array nullable. */
rmx_status rmx_get_wetdry(rmx_ctx* ctx, int* nflag /* [np] */, int* eflag /* [ne] */);
/* The relaxed time steps dt_UV, dt_H of the running cycle and its time step dt (below
 * relaxation method 3 all three are dt), for the flow cycle's test of the drying and
 * rewetting (EqsUVS2D.cpp:213-214).  Each nullable. */
rmx_status rmx_get_relax_time(rmx_ctx* ctx, double* dt_uv, double* dt_h, double* dt);

/* Host copies of the time-integration state, 9 [np] arrays (NULL entries skipped):
 * Uo, Vo, So, dUdt_o, dVdt_o, dSdt_o (previous time step), dUdt, dVdt, dSdt (current).
 * set = 1 uploads them, 0 downloads. */
rmx_status rmx_time_state(rmx_ctx* ctx, int set, double* const* arrays);

/* ------------------------------------------------------------------------------------ */
/* drop-in group                                                                         */

/* Element coefficients of model elements [first, first+count): replaces EQS::Coefs /
 * EQS_UVS2D::Region + Bound + Rotate2D + inlet rows (CoefsUVS2D.cpp:49-1181).
 * estifm: [count*24*24] row-major per element, force: [count*24]; layout of the
 * reference (row i + j*nnd, j = 0 U, 1 V, 2 H). */
rmx_status rmx_element_coefs(rmx_ctx* ctx, int first, int count, double* estifm, double* force);

/* Assemble the Newton system on the device: replaces CRSMAT::AssembleEqs_im
 * (Assemble.cpp:195) with EQS_UVS2D::Coefs.  prec RMX_F64: FP64 values; RMX_F32: FP32
 * values with each element entry rounded to FP32 before it is added (the reference's
 * stock REALPR = float semantics).  The RHS B is FP64 either way. */
rmx_status rmx_assemble(rmx_ctx* ctx, int prec);

/* B /= ||B||_2 and A /= ||B||_2: replaces CRSMAT::ScaleL2Norm (CRSMat.cpp:319). */
rmx_status rmx_scale_l2(rmx_ctx* ctx, double* scale_out);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc): the scaling of
 * rmx_scale_l2 (B /= ||B||_2, A /= ||B||_2, CRSMAT::ScaleL2Norm) on a system the caller holds
 * in device memory: n rows, rowptr[n + 1] (device), the rowptr[n] values A (device, float
 * for prec RMX_F32, double for RMX_F64) and B (device, double), all scaled in place;
 * ||B||_2 is written to scale (device, one double).  Work is queued on `stream` (a
 * hipStream_t, NULL: the null stream) of HIP device `device` and the call returns without
 * waiting for it.  It allocates nothing: the reduction's partial sums live in a static
 * device array of 520 doubles (4160 bytes) per GPU, held until the process ends, whatever n
 * is, so no input size exceeds it; calls that overlap on one GPU share that array and race.
 * The results are bit-identical to rmx_scale_l2 on the same values (same kernels). */
rmx_status rmx_scale_l2_noalloc(int device, int n, const int* rowptr, int prec, void* A,
                                double* B, double* scale, void* stream);

/* Copy the FP32 matrix to or from the FP64 matrix (rounding to nearest). */
rmx_status rmx_convert_matrix(rmx_ctx* ctx, int to_prec);

/* ILU(0) on the matrix pattern: replaces PRECO_ILU0::Factor (Preco_ilu0.cpp:64), with its
 * pivot repair (:95-135: a pivot below 1e-20 gets the later row of its pattern with the
 * largest entry in its column added).  prec selects which matrix (and factor) storage to
 * use. */
rmx_status rmx_ilu0_factor(rmx_ctx* ctx, int prec);

/* x = U^-1 L^-1 b with the ILU(0) factors: replaces PRECO_ILU0::Solve
 * (Preco_ilu0.cpp:497).  Host pointers, length neq. */
rmx_status rmx_ilu0_apply(rmx_ctx* ctx, int prec, const double* b, double* x);

/* Strip ILU(0) (RMX_PREC_SILU0): factor and apply.  No reference counterpart: the
 * reference's ILU(0) factors (PRECO_ILU0::Factor) applied with block-Jacobi triangular
 * sweeps (PRECO_ILU0::Solve, Preco_ilu0.cpp:497, without the factor entries that couple two
 * blocks).  A block is one of `segments` parts of a strip of consecutive equations, cut
 * along the principal axis of the strip's node coordinates.  Blocks of at most 2047 rows are
 * swept by one wave each, both triangles in one launch (a node's rows solved together);
 * larger blocks by one workgroup per strip and triangle.  rmx_set_strips chooses the
 * number of strips (0 = two per compute unit,
 * the default, and at most one per 2048 rows) and segments (0 = the default: blocks of
 * about 560 rows, at most 32 segments per strip; 1 for a system loaded with
 * rmx_csr_upload, which has no coordinates) and rebuilds the schedule; rmx_get_blocks
 * returns the block of every equation. */
rmx_status rmx_silu0_factor(rmx_ctx* ctx, int prec);
rmx_status rmx_silu0_apply(rmx_ctx* ctx, int prec, const double* b, double* x);
rmx_status rmx_set_strips(rmx_ctx* ctx, int strips, int segments);
rmx_status rmx_get_blocks(rmx_ctx* ctx, int* block /* [neq] */);

/* Two-level preconditioner (RMX_PREC_SILU0C).  No reference counterpart.  Aggregates: one
 * per (strip, coarse segment, variable U/V/H), coarse segments cut like the sweep
 * segments; P is piecewise constant on aggregates and R = P^T.  Apply:
 * z1 = P (R A P)^-1 R r, z = z1 + M^-1 (r - A z1) with M the strip ILU(0).
 * rmx_set_coarse chooses the coarse segments per strip (-1, the default: one per 256 rows of
 * a strip, at most 24; 0: no coarse space; needs a mesh, i.e. not after rmx_csr_upload); rmx_coarse_factor sums R A P from the FP32 or FP64
 * matrix and factors it; rmx_get_aggregates writes the aggregate of every equation and
 * returns the number of aggregates in *nc. */
rmx_status rmx_set_coarse(rmx_ctx* ctx, int segments);
rmx_status rmx_coarse_factor(rmx_ctx* ctx, int prec);
rmx_status rmx_get_aggregates(rmx_ctx* ctx, int* agg /* [neq] */, int* nc);

/* Learned preconditioner (RMX_PREC_MPNN).  No reference counterpart.  A message-passing
 * network on the mesh node graph (an edge joins two nodes of one element) that sees the
 * state, not the residual, so the preconditioner it defines is linear in the residual:
 * node features (u, v, depth/4, bed/4, 100 cf, speed, open-boundary flag, rotated-node
 * flag), an input layer 8 -> 64, 6 layers h = h + relu([h, mean of neighbours' h] W + b)
 * (BF16 MFMA, FP32 accumulate), an output layer 64 -> 18: per node two 3x3 blocks D_i, E_i.
 * The correction of a residual r: q = r / max_j |a_ij| (row-scaled), q_i its (u, v, h)
 * components at node i, qbar_i their neighbour mean, n(r)_i = D_i q_i + E_i qbar_i.  As a
 * preconditioner it corrects the two-level preconditioner M2 (RMX_PREC_SILU0C, which must
 * be factored): z2 = M2^-1 r, r2 = r - A z2, z = z2 + n(r2).  The network runs once per
 * Newton system; at the system's first application the correction is tested on that
 * vector and dropped for the whole system if it does not lower the residual.
 * rmx_mpnn_n_params: number of FP32 weights (layout: W_in [8][64], b_in, then per layer
 * W [128][64], b [64], then W_out [64][18], b_out [18]).
 * rmx_mpnn_set_weights / get_weights: the weights (host, FP32); setting also resets Adam.
 * rmx_mpnn_forward: z = n(r) alone, and the node features [np][8] (nullable).
 * rmx_mpnn_add_sample: store a training sample r (host).  The first sample of a Newton
 * system (the current FP64 matrix A, not stored yet) also stores the system: A, its row
 * maxima and the node features of the current state, shared by the system's later samples.
 * capacity > 0 empties the set and sizes it for `capacity` samples; a mesh of another
 * topology or wet set empties it too, and the next call must then pass a capacity.
 * RMX_ERR_RANGE when the set is full.
 * rmx_mpnn_loss_grad: the preconditioned residual ||r - A n(r)||^2 / ||r||^2 of one
 * stored sample and its gradient (nullable).
 * rmx_mpnn_train: Adam (beta 0.9 / 0.999) over the stored systems, one step per system on
 * the mean loss of its samples, `epochs` passes; *loss = mean sample loss of the last pass. */
rmx_status rmx_mpnn_n_params(int* n);
rmx_status rmx_mpnn_set_weights(rmx_ctx* ctx, const float* theta, int n);
rmx_status rmx_mpnn_get_weights(rmx_ctx* ctx, float* theta, int n);
rmx_status rmx_mpnn_forward(rmx_ctx* ctx, const double* r, double* z, float* feat);
rmx_status rmx_mpnn_add_sample(rmx_ctx* ctx, const double* r, int capacity, int* count);
rmx_status rmx_mpnn_loss_grad(rmx_ctx* ctx, int sample, double* loss, float* grad);
rmx_status rmx_mpnn_train(rmx_ctx* ctx, int epochs, double lr, double* loss);

/* x = M^-1 b for any preconditioner kind (factored beforehand).  Host pointers. */
rmx_status rmx_precond_apply(rmx_ctx* ctx, int kind, int prec, const double* b, double* x);

/* Shape of the sweeps of an ILU(0) instance (kind RMX_PREC_ILU0 or RMX_PREC_SILU0): factor
 * entries each sweep reads and its number of levels (the longest strip's). */
typedef struct {
  long long lower_entries, upper_entries;
  int lower_levels, upper_levels, strips, segments;
} rmx_sweep_info;
rmx_status rmx_get_sweep_info(rmx_ctx* ctx, int kind, rmx_sweep_info* out);

/* r = A x: replaces CRSMAT::MulVec (CRSMat.cpp:466).  Host pointers. */
rmx_status rmx_spmv(rmx_ctx* ctx, int prec, const double* x, double* r);

typedef struct rmx_krylov_opts {
  int method;                 /* RMX_BICGSTAB | RMX_FGMRES */
  int prec;                   /* storage of matrix + factors used by the Krylov loop */
  int precond;                /* RMX_PREC_* */
  int max_iter;               /* SOLVER::maxIter */
  int mkyrl;                  /* FGMRES restart length (SOLVER::mkyrl) */
  double eps;                 /* SOLVER::maxDiff: relative residual tolerance */
  int refine;                 /* 1: FP64 iterative refinement around the solve */
  double refine_tol;          /* true FP64 relative residual to reach when refining; a
                                 refinement sweep's inner tolerance is max(eps, 0.5 refine_tol
                                 / current residual), so the last sweep stops near the target */
  int refine_max;             /* maximum refinement sweeps */
  /* what follows a solve that has not converged (SOLVER::proceed, EQS::Solve,
     Solve.cpp:325-362): 0 keep the unconverged X; -1 set X to zero and report the solve
     interrupted; > 0 continue with the solver `next` (the reference's solver number
     `proceed`) when its tolerance is below the accuracy reached, from X (from zero when
     that accuracy is 1 or more); `next` may chain further */
  int proceed;
  const struct rmx_krylov_opts* next;
} rmx_krylov_opts;

typedef struct {
  int iters;                  /* Krylov iterations (summed over refinement sweeps) */
  int sweeps;                 /* refinement sweeps (0 without refinement) */
  int converged;
  double accuracy;            /* last Krylov relative residual estimate */
  double true_resid;          /* ||B - A64 X||_2 / ||B||_2 in FP64 */
  double seconds;             /* wall time of the solve on the device stream */
  int fallback;               /* 1: a refinement sweep stalled and the remaining sweeps ran
                                 the two-level preconditioner in place of the learned one,
                                 or BiCGStab in place of FGMRES; 2: a sweep raised the
                                 residual and the rest ran FGMRES(30) from the best iterate */
  int solvers;                /* solvers of the proceed chain that ran (1 without it) */
  int interrupted;            /* 1: not converged with proceed = -1 (X is zero) */
} rmx_krylov_out;

/* Solve A X = B for the assembled system: replaces SOLVER::Iterate with P_BCGSTABD
 * (P_bcgstabd.cpp:98) or P_FGMRESD (P_fgmresd.cpp:90) and the ILU(0) preconditioner.
 * X starts at zero.  With refine = 1 the solve runs in the requested storage precision
 * inside an FP64 iterative refinement on the FP64 matrix; a refinement sweep that lowers
 * the true FP64 residual less than tenfold switches the remaining sweeps to the two-level
 * preconditioner in place of the learned one, or from FGMRES to BiCGStab
 * (out->fallback = 1).  A sweep that raises it tenfold or more is undone and the remaining
 * sweeps run FGMRES with restart 30 (out->fallback = 2); if that has run, the refinement
 * ends.  X is the refinement's iterate of the smallest true residual.  FGMRES ends a solve whose restart cycle made no progress (restarted GMRES can stagnate on
 * these indefinite systems).  A solve that does not converge follows opt->proceed. */
rmx_status rmx_solve(rmx_ctx* ctx, const rmx_krylov_opts* opt, rmx_krylov_out* out);

/* Host access to the system (tests, drop-in callers that keep host arrays). */
rmx_status rmx_get_matrix(rmx_ctx* ctx, int prec, void* vals /* nnz */, double* B /* neq */);
rmx_status rmx_get_factors(rmx_ctx* ctx, int prec, void* vals /* nnz */);
rmx_status rmx_get_solution(rmx_ctx* ctx, double* X /* neq */);
rmx_status rmx_set_rhs(rmx_ctx* ctx, const double* B);
/* The Newton correction rmx_newton_update applies: a caller that solved the system itself
 * (as EQS_UVS2D::Execute gets X from EQS::Solve, EqsUVS2D.cpp:365) hands it over. */
rmx_status rmx_set_solution(rmx_ctx* ctx, const double* X /* neq */);
/* Load an arbitrary CSR system (diagonal first in every row): lets a caller that keeps
 * the reference's own EQS/CRSMAT host code hand its pattern and values over. */
rmx_status rmx_csr_upload(rmx_ctx* ctx, int n, const int* rowptr, const int* cols,
                          const double* vals, const double* B);

/* ------------------------------------------------------------------------------------ */
/* Newton step                                                                           */

typedef struct {
  double maxAbs[3];           /* max |dU|, |dV|, |dh| of the correction (EQS::Update) */
  double avAbs[3];
  double relaxUs, relaxS;
  int nrconv;                 /* max corrections below conv_uv / conv_s (methods 3, 4: and the
                                 relaxed time step at the time step) */
  double scale;               /* ||B||_2 before scaling */
  rmx_krylov_out solve;
  double t_prepare, t_assemble, t_factor, t_solve, t_update;   /* seconds */
  double relax_time;          /* methods 3, 4: the relaxed time step after this update */
  double relax_time_last;     /* and the one this step's Jacobian used (last_dt) */
} rmx_step_stats;

/* Relaxation, correction (rotated back for rotated nodes), velocity cap, depth / marsh
 * handling and midside interpolation: EqsUVS2D.cpp:398-760, EQS::Update (Update.cpp:42). */
rmx_status rmx_newton_update(rmx_ctx* ctx, rmx_step_stats* st);

/* One full Newton step on the device: prepare, assemble (FP64), scale, factor, solve,
 * update. */
rmx_status rmx_newton_step(rmx_ctx* ctx, const rmx_krylov_opts* opt, rmx_step_stats* st);

/* Kernel timing with HIP events on the context stream: average duration in seconds of
 * one SpMV and of the forward / backward sweeps of each factored ILU(0), over `reps`
 * runs on the current system; spmv_bytes: algorithmic bytes of that SpMV in the layout
 * it ran on (FP32: the row-group layout when built, else CSR). */
typedef struct {
  double spmv_s, ilu_apply_s, ilu_fwd_s, ilu_bwd_s;   /* reference-order ILU(0) */
  double silu_fwd_s, silu_bwd_s;                      /* strip ILU(0) (0 if not factored) */
  long long n_spmv, n_ilu;
  double spmv_bytes;
} rmx_kernel_times;
rmx_status rmx_profile_kernels(rmx_ctx* ctx, int prec, int reps, rmx_kernel_times* out);

/* Launch timing per kernel class over any sequence of calls (no reference counterpart):
 * with timing enabled, every launch of the classes below is bracketed by HIP events on the
 * context stream (a few microseconds of recording per launch); the sums are read back when
 * the stream synchronises.  Classes: 0 ILU(0) factorisation, 1 lower and 2 upper triangular
 * sweep, 3 Krylov SpMV (the storage precision of the solve, product or residual form),
 * 4 coarse correction (restriction, coarse solve, A P residual), 5 Gram-Schmidt (dots and
 * update), 6 FP64 residual of the refinement, 7 both triangular sweeps of the strip ILU(0)
 * in one launch (block sweeps: one wave per block).  bytes: the algorithmic bytes of each
 * launch summed (factorisation 3 nnz (v + 4); sweep: kept entries (v + 4) + 20 or 24 bytes
 * per row; block sweeps: kept entries of both triangles (v + 2) + (22 + v) bytes per row and
 * 4 per level, + 12 per row with the folded prolongation; SpMV: the bytes of the layout it
 * ran on; Gram-Schmidt: the basis vectors read, FP32 or FP64, and the vector read and
 * written; coarse: the vectors, maps and A P entries it reads). */
typedef struct {
  double seconds[8];
  double bytes[8];
  long long launches[8];
} rmx_kernel_stats;
/* on != 0 resets the sums and starts timing; 0 stops it (the sums stay readable). */
rmx_status rmx_kernel_stats_enable(rmx_ctx* ctx, int on);
rmx_status rmx_kernel_stats_get(rmx_ctx* ctx, rmx_kernel_stats* out);

#ifdef __cplusplus
}
#endif
#endif
