/*
 * festung_hip.h -- C-ABI of the MI355X (gfx950) implementation of FESTUNG's
 * explicit DG advection step on triangles.
 *
 * Conventions
 *  - Every function returns an int status (FESTUNG_OK == 0); the message of
 *    the last failure on the calling thread is festung_last_error().
 *  - Arrays are column-major with the reference's Matlab shapes, so a
 *    Matlab/Octave gateway passes them without copies: a K x N coefficient
 *    matrix cDisc is stored cDisc[k + i*K] (structure-of-arrays), a K x 3 x R
 *    array valOnQuad is valOnQuad[k + n*K + r*3*K].
 *  - Ids are 0-based int32, "none" is -1 (the gateway subtracts 1 from
 *    Matlab's 1-based ids).
 *  - Host pointers are read synchronously and never retained; the caller
 *    owns them.  "_dev" entry points take device pointers instead (same
 *    layout) and enqueue on the problem's stream.
 *  - A problem handle is not re-entrant: calls on one handle must not
 *    overlap.  Handles on different devices are independent.
 */
#ifndef FESTUNG_HIP_H
#define FESTUNG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FESTUNG_OK = 0,
  FESTUNG_ERR_ARG = 1,          /* invalid argument (null, size, range)       */
  FESTUNG_ERR_HIP = 2,          /* HIP runtime error (no device, launch, ...) */
  FESTUNG_ERR_ALLOC = 3,        /* host or device allocation failed           */
  FESTUNG_ERR_STATE = 4,        /* call not valid in the handle's state       */
  FESTUNG_ERR_UNSUPPORTED = 5,  /* p, RK order or option not supported        */
  FESTUNG_ERR_INTERNAL = 6
};

typedef struct festung_grid festung_grid;
typedef struct festung_problem festung_problem;

/* Message of the last failing call on this thread ("" after a success);
 * the text a gateway hands to Matlab's error() (the reference reports
 * failures by error()/assert).  No counterpart in the reference. */
const char* festung_last_error(void);
/* Library version string.  No counterpart in the reference. */
const char* festung_version(void);

/* ---------------------------------------------------------------- grid --
 * Host-only (no device needed). */

/* Friedrichs-Keller triangulation of [a,b]^2 with boundary ids and marks.
 * Replaces core/domainSquare.m:67-105 (+ core/generateGridData.m:130-198,
 * advection/preprocessProblem.m:72-76). */
int festung_domain_square(double hmax, double a, double b, festung_grid** out);

/* Connectivity and geometry of a counter-clockwise triangle mesh.
 * coordV: numV x 2, V0T: numT x 3 (0-based).  Boundary ids are all 0 until
 * festung_grid_set_boundary is called.
 * Replaces core/generateGridData.m:130-198. */
int festung_grid_create(const double* coordV, int32_t numV, const int32_t* V0T, int32_t numT,
                        festung_grid** out);

/* Set edge boundary ids idE (numE, 0 = interior) and optional Neumann marks
 * markE0TbdrN (numT x 3 bytes, may be NULL); recomputes markE0Tint,
 * markE0TbdrD and markV0TbdrD.
 * Replaces advection/configureProblem.m:148-150 and
 * advection/preprocessProblem.m:72-76. */
int festung_grid_set_boundary(festung_grid* g, const int32_t* idE, const uint8_t* markE0TbdrN);

/* Disjoint union of n grids (element, vertex and edge ids offset in order).
 * Used to run many independent problems of one degree as one device
 * problem (the batched form).  No counterpart in the reference, which
 * solves one problem per run. */
int festung_grid_concat(const festung_grid* const* grids, int32_t n, festung_grid** out);

/* g.numT, g.numV, g.numE (core/generateGridData.m:130-198). */
int festung_grid_sizes(const festung_grid* g, int32_t* numT, int32_t* numV, int32_t* numE);

/* Copy a named grid field out.  int32 fields: "V0T" (numT x 3), "V0E",
 * "T0E", "E0E" (numE x 2), "E0T", "idE0T", "nbr", "nbrEdge" (numT x 3),
 * "idE" (numE), "v2tPtr" (numV+1), "v2tIdx" (3 numT), "markE0Tint",
 * "markE0TbdrD", "markE0TbdrN", "markV0TbdrD" (numT x 3, 0/1).
 * double fields: "coordV" (numV x 2), "areaE" (numE), "nuE", "baryE"
 * (numE x 2), "areaT" (numT), "baryT" (numT x 2), "B" (numT x 2 x 2),
 * "areaE0T", "sigE0T" (numT x 3), "nuE0T", "coordV0T" (numT x 3 x 2).
 * Mirrors the struct g of core/generateGridData.m:130-198. */
int festung_grid_field_size(const festung_grid* g, const char* field, int64_t* count, int* is_double);
int festung_grid_get_i32(const festung_grid* g, const char* field, int32_t* out);
int festung_grid_get_f64(const festung_grid* g, const char* field, double* out);

/* Physical coordinates of nq reference points on every element:
 * X1, X2 are numT x nq.  Replaces g.mapRef2Phy (core/generateGridData.m:195-196). */
int festung_grid_map_ref_to_phy(const festung_grid* g, int32_t nq, const double* q1, const double* q2,
                                double* X1, double* X2);

/* Free a grid (the reference's struct g is freed by Matlab). */
void festung_grid_destroy(festung_grid* g);

/* ------------------------------------------------------ reference tables --
 * Host-only.  Named tables for polynomial degree p (0..4):
 *  "quad1D"   (R x 2: nodes, weights) for quadRule1D(ord),   core/quadRule1D.m:48-95
 *  "quad2D"   (R x 3: Q1, Q2, W)      for quadRule2D(ord),   core/quadRule2D.m:64-109
 *  "hatM" (N x N)                     core/integrateRefElemPhiPhi.m:49-59
 *  "hatG" (N x N x N x 2)             core/integrateRefElemDphiPhiPhi.m:55-82
 *  "hatRdiag" (N x N x 3 x R)         core/integrateRefEdgePhiIntPhiIntPerQuad.m:52-66
 *  "hatRoff" (N x N x 3 x 3 x R)      core/integrateRefEdgePhiIntPhiExtPerQuad.m:54-70
 *  "phi2D", "gradPhi2D", "phi1D", "thetaPhi1D" for order ord
 *                                     core/computeBasesOnQuad.m:72-134
 *  "taylorF", "taylorFinv" (N x N)    Taylor change-of-basis tables
 *  "quadG"    (R x 3: Q1, Q2, W)      the element term's rule: quadRule2D(max(2p,1))
 *                                     for p <= 3; at p = 4 a 28-point rule exact to
 *                                     degree 11, which integrates the integrands of
 *                                     hatG exactly, as quadRule2D(8) does
 * `ord` selects the quadrature order for the quad/phi tables (ignored
 * otherwise).  festung_ref_size gives the element count. */
/* The orthonormal basis at reference points: out (npts x N, column-major)
 * = phi_i(X1_r, X2_r), core/phi.m:40-61. */
int festung_phi_at_points(int p, int32_t npts, const double* X1, const double* X2, double* out);
int festung_ref_size(int p, const char* name, int ord, int64_t* count);
int festung_ref_get(int p, const char* name, int ord, double* out);

/* ------------------------------------------------------------- problem --
 * Device-resident advection problem on the current HIP device. */

enum {
  FESTUNG_FLAG_LIMITER = 1,      /* Taylor-linear vertex-based slope limiter (typeSlopeLim 'linear') */
  FESTUNG_FLAG_GRAPH = 2,        /* capture whole steps into a hipGraph in festung_step */
  FESTUNG_FLAG_ADJ_SERIAL = 4,   /* festung_adjoint: recompute checkpointed segments on the problem's
                                    stream (default: on a second stream, overlapping the sweep) */
  FESTUNG_FLAG_ADJ_FUSED = 8,    /* festung_adjoint: velocity sensitivities in the transposed stage
                                    kernel's thread (default on large meshes) */
  FESTUNG_FLAG_ADJ_SPLIT = 16,   /* festung_adjoint: stage, element-term and edge-term sensitivities
                                    on three threads per element */
  FESTUNG_FLAG_ADJ_DEFER = 32,   /* festung_adjoint: transposed stages alone during the sweep, the
                                    sensitivities of each checkpoint segment in one pass after it
                                    (default on small meshes) */
  FESTUNG_FLAG_LIM_HIERARCH = 64, /* limiter type 'hierarch_vert' (core/applySlopeLimiterTaylorHierarchicalVertex.m) */
  FESTUNG_FLAG_LIM_STRICT = 128,  /* limiter type 'strict' (core/applySlopeLimiterTaylorStrict.m); default 'linear' */
  FESTUNG_FLAG_ADJ_BINOMIAL = 256 /* festung_adjoint: binomial checkpointing -- ckptSteps is the number
                                    of state snapshots; every step is recomputed with its stage states
                                    just before it is reversed (fused or split sensitivities) */
};

/* Upload the grid (its connectivity and marks as they are now) and build the
 * device tables for degree p (0..4) and RK order ordRK (1..3 SSP as in
 * core/rungeKuttaExplicit.m:64-78; 4 = classic RK4, an extension).
 * Replaces advection/preprocessProblem.m:48-112. */
int festung_problem_create(const festung_grid* g, int p, int ordRK, int flags, festung_problem** out);
void festung_problem_destroy(festung_problem* h);
int festung_problem_sizes(const festung_problem* h, int32_t* K, int32_t* N, int32_t* RE, int32_t* RF,
                          int32_t* Re);
/* HIP stream the problem enqueues on (hipStream_t as void*). */
int festung_problem_stream(const festung_problem* h, void** stream);

/* Velocity as DG coefficients u1Disc, u2Disc (K x N) and normal velocity at
 * the edge quadrature points vNormalOnQuadEdge (K x 3 x Re), as computed in
 * advection/preprocessSubStep.m:71-79.  Host arrays.  A NULL field keeps
 * its value (so the element and edge assembly calls can each hand over
 * theirs); steps need all three set once. */
int festung_set_velocity_disc(festung_problem* h, const double* u1Disc, const double* u2Disc,
                              const double* vNormalOnQuadEdge);
/* Velocity from point values: u1q, u2q at the element points of
 * quadRule2D(2p) (K x RE, mapped by festung_grid_map_ref_to_phy), u1e, u2e
 * at the edge points of quadRule1D(2p+1) (K x 3 x Re).  The device projects
 * (core/projectFuncCont2DataDisc.m:94-102) and forms u.nu
 * (core/computeFuncContNuOnQuadEdge.m:46-67). */
int festung_set_velocity_points(festung_problem* h, const double* u1q, const double* u2q, const double* u1e,
                                const double* u2e);

/* Right-hand side of the stage: fq = f at the element points of
 * quadRule2D(2p) (K x RE, may be NULL for f = 0), cDe = c_D and gNe = g_N at
 * the edge points of quadRule1D(2p+1) (K x 3 x Re each, may be NULL for
 * zero data).  The device forms globL - globKD - globKN
 * (advection/preprocessSubStep.m:69-98: core/projectFuncCont2DataDisc.m,
 * core/assembleMatElemPhiPhi.m, core/assembleVecEdgePhiIntFuncContVal.m on
 * the Dirichlet edges, core/assembleVecEdgePhiIntFuncCont.m with the upwind
 * part of g_N, (g_N <= 0) g_N, on the Neumann edges of
 * festung_grid_set_boundary).  With all three NULL the source is cleared.
 * Takes effect for the following stages. */
int festung_set_source_points(festung_problem* h, const double* fq, const double* cDe, const double* gNe);

/* Dirichlet data at the vertices for the limiter: vertexD (numV values, NaN
 * where not given), used for vertices with markV0TbdrD
 * (core/computeMinMaxV0TElementPatch.m:81-82). NULL clears (all NaN). */
int festung_set_vertex_dirichlet(festung_problem* h, const double* vertexD);

/* Device-pointer variants of the three set-ups above: the arrays live in
 * device memory, in the same layouts and the reference's element / vertex
 * order (e.g. point values computed by torch on the GPU), so time-dependent
 * data never pass through the host.  The work is enqueued on the problem's
 * stream (festung_problem_stream): the arrays must be ready when that stream
 * reaches it and stay valid until it has run.  The same
 * advection/preprocessSubStep.m:61-98 and core/computeFuncContV0T.m as above. */
int festung_set_velocity_points_dev(festung_problem* h, const double* u1q, const double* u2q, const double* u1e,
                                    const double* u2e);
int festung_set_source_points_dev(festung_problem* h, const double* fq, const double* cDe, const double* gNe);
int festung_set_vertex_dirichlet_dev(festung_problem* h, const double* vertexD);

/* L2 projection of point values fq (K x R of quadRule2D(ord), ord in
 * {2p, 2p+1}) to DG coefficients (K x N): core/projectFuncCont2DataDisc.m:94-102.
 * If set_state != 0 the result also becomes the current state. */
int festung_project(festung_problem* h, int ord, const double* fq, double* cDisc_out, int set_state);

/* Current solution problemData.cDisc (K x N; advection/initializeProblem.m:53,
 * advection/postprocessStep.m:59); the _dev variants take [N][K] device
 * memory on the problem's device. */
int festung_set_state(festung_problem* h, const double* cDisc);
int festung_get_state(festung_problem* h, double* cDisc);
int festung_set_state_dev(festung_problem* h, const double* d_cDisc);
int festung_get_state_dev(festung_problem* h, double* d_cDisc);

/* One RK sub-step nSubStep (1-based) of the time step that started from the
 * state saved by festung_begin_step: rhs, optional dc/dt limiting, stage
 * combination, optional solution limiting.
 * Replaces advection/solveSubStep.m:57-99 (with preprocessSubStep.m's
 * operators applied matrix-free). */
int festung_begin_step(festung_problem* h);
int festung_stage(festung_problem* h, int nSubStep, double tau);

/* nsteps complete time steps of size tau on the device (all RK stages; the
 * source and vertex Dirichlet data stay as set).  Replaces the main loop
 * main.m:151-166 over advection/solveStep.m:57-73 for time-independent data. */
int festung_step(festung_problem* h, int nsteps, double tau);

/* Per-element time step tauK (K values, reference element order) used by
 * the RK combination instead of the scalar tau; NULL restores the scalar.
 * Lets independent problems with different time steps share one device
 * problem (see festung_grid_concat).  No counterpart in the reference. */
int festung_set_tau_field(festung_problem* h, const double* tauK);

/* Elements per workgroup of the element kernels (stage, limiter, adjoint):
 * 64, 128 or 256.  Small meshes fill more compute units with
 * smaller workgroups.  The default is the largest size giving at least 1024
 * workgroups, else 64.  Drops a captured step graph.  No counterpart in the
 * reference (its operators are sparse matrix products). */
int festung_set_elem_block(festung_problem* h, int32_t elemsPerBlock);
int festung_get_elem_block(const festung_problem* h, int32_t* elemsPerBlock);

/* Autotuner: times nsteps steps of size tau with each elements-per-workgroup
 * choice (64, 128, 256; msPerStep[0..2] if not NULL), keeps the fastest
 * (*best) and restores the state the call started from.  No counterpart in
 * the reference. */
int festung_autotune(festung_problem* h, double tau, int32_t nsteps, int32_t* best, double* msPerStep);

/* Tuning table: festung_autotune records its choice under the key (device
 * architecture and compute-unit count, p, ordRK, limiter type or none,
 * ceil(log2 K)); a problem created later under the same key starts with
 * that workgroup size instead of the default.  The table lives for the
 * process; festung_tune_save writes it as text (one entry a line: device p
 * ordRK limiter log2K elemsPerBlock; '#' starts a comment) and
 * festung_tune_load merges such a file (all or nothing: a malformed line
 * leaves the table unchanged, FESTUNG_ERR_ARG names the line).  If the
 * environment variable FESTUNG_TUNE_FILE names a readable file, the first
 * festung_create loads it.  No counterpart in the reference. */
int festung_tune_save(const char* path);
int festung_tune_load(const char* path);
int festung_tune_clear(void);
int festung_tune_count(int32_t* entries);

/* Batched stepping: n problems advanced nsteps steps each (tau[i] for
 * problem i; ignored where a tau field is set) as ONE hipGraph launch whose
 * parallel branches are the problems' step sequences.  The graph is cached
 * while nsteps and tau are unchanged.  festung_batch_step returns after
 * enqueueing on the first problem's stream; festung_batch_sync waits.
 * No counterpart in the reference (one problem per run, main.m:151-166). */
typedef struct festung_batch festung_batch;
int festung_batch_create(festung_problem* const* problems, int32_t n, festung_batch** out);
int festung_batch_step(festung_batch* b, int32_t nsteps, const double* tau);
int festung_batch_sync(festung_batch* b);
void festung_batch_destroy(festung_batch* b);

/* Discrete adjoint (limiter off, f = 0, c_D = 0, SSP orders 1..3).
 * Runs nsteps steps of size tau from the current state (which becomes the
 * final state), the loss J = (c(T) - cObs)^T globM (c(T) - cObs) (M-weighted
 * squared L2 misfit; cObs K x N), and the reverse sweep through the
 * transposed element and edge operators and the reverse RK stages.  Outputs
 * (any may be NULL): J, dJ/dc0 (K x N), dJ/du1Disc, dJ/du2Disc (K x N) and
 * dJ/dvNormalOnQuadEdge (K x 3 x Re) -- the velocity enters only through
 * these fields, so a caller's parameter gradient is their contraction with
 * the fields' derivatives.  Stage states are kept in HBM: checkpoints every
 * ckptSteps steps (<= 0: all steps stored, no recomputation), each segment
 * recomputed once in the reverse sweep; under FESTUNG_FLAG_ADJ_BINOMIAL,
 * ckptSteps state snapshots placed by the binomial schedule
 * (festung_binomial_schedule).  timing_ms (NULL or 2 values): forward (up
 * to the loss) and reverse device time.  No counterpart in the reference. */
int festung_adjoint(festung_problem* h, int32_t nsteps, double tau, const double* cObs, int32_t ckptSteps,
                    double* J, double* gradC0, double* gradU1, double* gradU2, double* gradVn,
                    double* timing_ms);
/* festung_adjoint with device pointers (the problem's device, reference
 * element order, layouts [N][K] and [Re][3][K]) for cObs and the gradient
 * outputs; J and timing_ms stay host values.  Lets a torch caller keep the
 * observation and the gradients on the GPU.  No counterpart in the
 * reference. */
int festung_adjoint_dev(festung_problem* h, int32_t nsteps, double tau, const double* cObs, int32_t ckptSteps,
                        double* J, double* gradC0, double* gradU1, double* gradU2, double* gradVn,
                        double* timing_ms);
/* The binomial checkpointing schedule festung_adjoint runs under
 * FESTUNG_FLAG_ADJ_BINOMIAL: reversing nsteps steps with `snapshots` state
 * snapshots (the initial state included) takes *advances forward steps
 * besides the nsteps recomputations with stored stage states, the least
 * possible (Griewank's binomial bound), in *actions advance / reverse
 * actions.  Host only.  No counterpart in the reference. */
int festung_binomial_schedule(int32_t nsteps, int32_t snapshots, int64_t* advances, int64_t* actions);
/* out = M^-1 A^T mu for the current velocity (K x N): the transposed
 * right-hand-side operator, for dot-product tests of the adjoint. */
int festung_adj_rhs(festung_problem* h, const double* mu, double* out);

/* ------------------------------------------- operator-level entry points --
 * The reference's assembled objects one at a time, for callers that keep
 * advection/solveSubStep.m:62-66 as written (operator objects whose mtimes /
 * mldivide call these).  x, y are K x N column-major (reference element
 * order); the _dev variant takes device memory in the same layout (enqueued
 * on the problem's stream, returns when the result is written). */
enum {
  FESTUNG_OP_G = 0,    /* globG{1} + globG{2} */
  FESTUNG_OP_G1 = 1,   /* globG{1}: core/assembleMatElemDphiPhiFuncDisc.m:181-202 with u1Disc */
  FESTUNG_OP_G2 = 2,   /* globG{2}: the same with u2Disc */
  FESTUNG_OP_R = 3,    /* globR: core/assembleMatEdgePhiPhiValUpwind.m:261-299 with markE0Tbdr =
                          ~markE0TbdrN and vNormalOnQuadEdge (advection/preprocessSubStep.m:83-85) */
  FESTUNG_OP_M = 4,    /* globM: core/assembleMatElemPhiPhi.m:88-100 */
  FESTUNG_OP_MINV = 5  /* globM \ x: the block solve of advection/solveSubStep.m:66 */
};
/* y = A x for A = FESTUNG_OP_* at the problem's current velocity (set by
 * festung_set_velocity_*; G and R need it). */
int festung_apply_operator(festung_problem* h, int op, const double* x, double* y);
int festung_apply_operator_dev(festung_problem* h, int op, const double* x, double* y);

/* festung_apply_operator_dev without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or solver loops, harnesses that forbid
 * hipMalloc).  The operand and result in the device's element order live in
 * a library-owned 64 MiB device array (one per GPU, held until the process
 * ends), so K*N must not exceed 4194304 (2^22; FESTUNG_ERR_ARG otherwise).
 * Enqueued on the problem's stream; returns without waiting.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (handles on
 * different streams) race; calls on one handle are ordered by its stream. */
int festung_apply_operator_dev_noalloc(festung_problem* h, int op, const double* x, double* y);

/* The stage vectors of advection/preprocessSubStep.m:87-98 from point
 * values (as festung_set_source_points takes them), each formed alone and
 * returned K x N: L = globM fDisc from fq (core/projectFuncCont2DataDisc.m,
 * :98), KD = core/assembleVecEdgePhiIntFuncContVal.m:162-185 from cDe on the
 * Dirichlet edges at the current velocity, KN =
 * core/assembleVecEdgePhiIntFuncCont.m:126-135 from the upwind part of gNe
 * on the Neumann edges.  An output may be NULL; a non-NULL output needs its
 * input.  Does not change the stage source. */
int festung_assemble_vectors(festung_problem* h, const double* fq, const double* cDe, const double* gNe, double* L,
                             double* KD, double* KN);

/* core/computeVertexBasedLimiter.m:78-106 with its own inputs: valCentroid
 * (K), valV0T (K x 3), markV0TbdrD (K x 3 bytes) and dataV0T (K x 3; both
 * NULL: no Dirichlet data).  Outputs (any may be NULL): alphaE (K) and
 * minMaxV0T as minV0T, maxV0T (K x 3), with the global shift of
 * core/computeMinMaxV0TElementPatch.m:66-112.  The handle supplies the mesh
 * (vertex patches); its state is not touched. */
int festung_vertex_based_limiter(festung_problem* h, const double* valCentroid, const double* valV0T,
                                 const uint8_t* markV0TbdrD, const double* dataV0T, double* alphaE, double* minV0T,
                                 double* maxV0T);

/* core/applySlopeLimiterDisc.m:77-86: dataDisc (K x N) limited with the
 * handle's limiter type -> dataDiscLim (K x N), minMaxV0T -> minV0T, maxV0T
 * (K x 3, may be NULL).  markV0TbdrD (K x 3 bytes) and dataV0T (K x 3) give
 * the Dirichlet vertex data (both NULL: none); the values of one vertex must
 * agree between its elements (FESTUNG_ERR_ARG otherwise), as
 * core/computeFuncContV0T.m produces them.  The limited data become the
 * handle's state.  Needs FESTUNG_FLAG_LIMITER. */
int festung_apply_slope_limiter_disc(festung_problem* h, const double* dataDisc, const uint8_t* markV0TbdrD,
                                     const double* dataV0T, double* dataDiscLim, double* minV0T, double* maxV0T);

/* core/rungeKuttaExplicit.m:64-78: stage times t[0..S-1] = t0 + c tau and the
 * SSP weights omega for ord 1..3; ord 4 (classic RK4, an extension) gives its
 * stage times and omega = NaN.  t, omega hold at least 4 values; *nStages = S.
 * Host only. */
int festung_runge_kutta_explicit(int ord, double tau, double t0, double* t, double* omega, int32_t* nStages);

/* Apply the slope limiter to the current state.
 * Replaces core/applySlopeLimiterDisc.m:77-86 (the limiter type chosen by
 * the FESTUNG_FLAG_LIM_* flags at creation: 'linear', 'hierarch_vert' or
 * 'strict'). */
int festung_limit(festung_problem* h);

/* Apply the vertex-based limiter of the DG basis' linear part to the
 * current state: vertex values of sum_{j<=3} c_j phi_j against the patch
 * bounds of the centroid values c_1 phi_1(1/3, 1/3) (with the Dirichlet
 * vertex values, if set, as for festung_limit), then
 * [c_1, alpha c_2:3, (alpha == 1) c_4:N].  Needs p >= 1 and a problem
 * created with FESTUNG_FLAG_LIMITER; festung_limiter_info then reports its
 * alpha and bounds.  Replaces core/applySlopeLimiterDiscLinear.m:109-123. */
int festung_limit_disc_linear(festung_problem* h);

/* Evaluate the DG right-hand side cDiscDot = M \ (sysV - sysA c) for the
 * current state without limiting (K x N) (advection/solveSubStep.m:62-66). */
int festung_rhs(festung_problem* h, double* cDiscDot);

/* Per-element alpha and per-(element, vertex) min / max of the last
 * limiter application (K, K x 3, K x 3).
 * Mirrors the outputs of core/computeVertexBasedLimiter.m:78-106. */
int festung_limiter_info(festung_problem* h, double* alphaE, double* minV0T, double* maxV0T);

/* L2 error against point values cq of the exact solution at the element
 * points of quadRule2D(ord) (K x R): core/computeL2Error.m:97-114. */
int festung_l2_error(festung_problem* h, int ord, const double* cq, double* err);

/* Per-launch timing with HIP events on the problem's stream (forces eager
 * launches instead of the step graph while enabled).  festung_profile_read
 * returns, per kernel kind, the summed milliseconds and launch counts since
 * the last read, and clears them.  No counterpart in the reference (its
 * only timing is tic/toc around the loop, main.m:151-166). */
enum {
  FESTUNG_KERNEL_STAGE = 0,      /* stage_kernel: rhs + mass solve + RK (+ Taylor) */
  FESTUNG_KERNEL_VERTEX = 1,     /* vertex patch min / max                          */
  FESTUNG_KERNEL_LIMIT_DOT = 2,  /* dc/dt limiting + RK + Taylor of the new state   */
  FESTUNG_KERNEL_LIMIT = 3,      /* solution limiting                               */
  FESTUNG_KERNEL_COPY = 4,       /* end-of-step copy (ordRK = 1 only)               */
  FESTUNG_KERNEL_KINDS = 5
};
int festung_profile(festung_problem* h, int enable);
int festung_profile_read(festung_problem* h, double* ms, int64_t* count);

/* Synchronise the problem's stream.  No counterpart in the reference
 * (whose calls are synchronous). */
int festung_sync(festung_problem* h);

#ifdef __cplusplus
}
#endif
#endif /* FESTUNG_HIP_H */
