/*
 * hamstr_hip.h -- C ABI of the MI355X implementation of HAMSTR_OOP's
 * line-implicit steady Euler iteration (computeRHS -> ADI -> computeLinearRHS
 * -> update).
 *
 * Every entry point returns an int status (HAMSTR_OK or one of the codes
 * below).  Host pointers are read or written only during the call.  A context
 * owns its device memory and one HIP stream; contexts are independent, and one
 * host thread at a time may use a context.
 *
 * Array layouts follow the reference's memory layout after
 * MESHBLOCK::readMeshData and MESHBLOCK::preprocess:
 *   faces      int    [6*nFace]   n1, n2, leftCell, leftEdge, rightCell, rightEdge
 *                                 (rightCell: >=0 cell, -1 far field, -2 inviscid wall)
 *   chainStart int    [nChain+1]  faceStartPerChain (duplicates already removed)
 *   chainConn  int    [nChainFace] face ids along each chain
 *   normVec    double [2*nFace]   (y2-y1, -(x2-x1))
 *   refMtx     double [4*nFace]   I - 2 n n^T / |n|^2, row-major
 *   q, qt, r, r0, dq double [4*nCell]  AoS (rho, rho u, rho v, E) per cell
 *   sigma            double [nCell]
 *   ff               double [32*nFace] lmat[4][4] then rmat[4][4] per face
 */
#ifndef HAMSTR_HIP_H
#define HAMSTR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define HAMSTR_OK 0
#define HAMSTR_ERR_ARG 1      /* bad argument or size                          */
#define HAMSTR_ERR_HIP 2      /* a HIP runtime call failed                     */
#define HAMSTR_ERR_NAN 3      /* NaN spectral radius (reference: exit(1))      */
#define HAMSTR_ERR_MESH 4     /* chain layout the device path does not accept  */
#define HAMSTR_ERR_NODEV 5    /* no HIP device                                 */

/* LHS precision: the Roe Jacobians and ADI block solves.  The residual, its
 * norms and the solution update are always FP64. */
#define HAMSTR_LHS_FP64 0
#define HAMSTR_LHS_FP32 1

/* ADI block-tridiagonal solve ordering.
 * EXACT: the reference's sequential block Thomas order (bit-identical).
 * FAST : the same systems solved with a segmented (partitioned) elimination,
 *        equal to EXACT up to rounding. */
#define HAMSTR_ADI_EXACT 0
#define HAMSTR_ADI_FAST 1

typedef struct hamstr_ctx hamstr_ctx;

typedef struct hamstr_mesh_desc {
  int nCell;
  int nFace;
  int nChain;
  const int *faces;
  const int *chainStart;
  const int *chainConn;
  const double *normVec;
  const double *refMtx;
  /* Orphan faces: interior faces that lie on no chain (a mesh with
   * triangles, whose cells a chain cannot cover face by face).  An extension
   * with no counterpart in the reference; nOrphan = 0 / orphans = NULL is the
   * reference's all-chain mesh.  Each orphan face gets the first-order Roe
   * flux of its two cell states; its Jacobian blocks join the diagonal of the
   * ADI rows of the last chain listing each of its cells, and its full
   * linearised flux joins computeLinearRHS.  In a batch, list them case by
   * case. */
  int nOrphan;
  const int *orphans;
} hamstr_mesh_desc;

typedef struct hamstr_params {
  double mach;      /* SOLNBLOCK::Mach   */
  double alpha;     /* SOLNBLOCK::alpha, degrees */
  double cfl;       /* SOLNBLOCK::CFL    */
  int order;        /* 1 first order, 3 MUSCL, 5 WENO5 */
  int msweep;       /* ADI sweeps per iteration */
  int lhs;          /* HAMSTR_LHS_FP64 | HAMSTR_LHS_FP32 */
  int adi_mode;     /* HAMSTR_ADI_EXACT | HAMSTR_ADI_FAST */
} hamstr_params;

/* Field ids for hamstr_get_field / hamstr_set_field. */
#define HAMSTR_FIELD_Q 0
#define HAMSTR_FIELD_R 1
#define HAMSTR_FIELD_R0 2
#define HAMSTR_FIELD_DQ 3
#define HAMSTR_FIELD_SIGMA 4
#define HAMSTR_FIELD_FF 5   /* converted to double when the LHS is FP32 */
#define HAMSTR_FIELD_QT 6   /* SOLNBLOCK::qt, rk3's stage-start state (per cell, 4 values) */

/* SOLVER::stepType (solver.cpp:22), which the reference reads from the input
 * deck's scheme= line (io.cpp:30): the branch of stepSolution that
 * hamstr_step runs (march.cpp:64-136). */
#define HAMSTR_SCHEME_ADI 0    /* computeRHS("implicit") + ADI(): the default */
#define HAMSTR_SCHEME_EULER 1  /* computeRHS("explicit") + updateSolution(q, q, 1) */
#define HAMSTR_SCHEME_RK3 2    /* the reference's three-stage Runge-Kutta */

/* ---- set-up ------------------------------------------------------------- */

/* MESHBLOCK::preprocess (preprocess.cpp:56-193): face normals, reflection
 * matrices, and wall tagging of boundary faces (rightCell == -1) whose
 * midpoint radius is below 3.0 (rightCell := -(sol_visc+2)).  Host only. */
int hamstr_preprocess(int nNode, const double *x, int nFace, int *faces,
                      double *normVec, double *refMtx, int sol_visc);

/* SOLVER::SOLVER + SOLVER::initFlow (solver.cpp:13-81, initFlow.cpp:12-93):
 * uploads the mesh, builds the device chain tables and sets q to freestream. */
int hamstr_create(const hamstr_mesh_desc *mesh, const hamstr_params *params,
                  hamstr_ctx **out);
int hamstr_destroy(hamstr_ctx *ctx);

/* A batch of independent flow cases in one context: the disjoint union of
 * their meshes, every kernel launch covering all cases.  Cases share cfl,
 * order, msweep, lhs and adi_mode; mach and alpha are per case.  Fields are
 * the union's arrays, case k owning cells [cell_off[k], cell_off[k+1]). */
int hamstr_batch_create(int ncases, const hamstr_mesh_desc *meshes, const hamstr_params *params,
                        hamstr_ctx **out);
/* *ncases, and cell_off[ncases+1] when not NULL */
int hamstr_batch_info(const hamstr_ctx *ctx, int *ncases, int *cell_off);
/* Up to nsteps iterations of every case.  iters[k] (may be NULL) gets the
 * first step (1-based) at which case k's nonlinear L2 <= tol, or -1; the run
 * stops at the first read-back (every check_every steps) after all cases have
 * converged.  norms[4*ncases] (may be NULL) gets each case's (L2, Linf) of
 * the nonlinear and of the printed residual at the last step (hamstr_step's
 * l2rhs and l2lin).  *done = steps run. */
int hamstr_batch_step(hamstr_ctx *ctx, int nsteps, double tol, int check_every, int *iters,
                      double *norms, int *done);
/* Retirement (no counterpart in the reference, which runs one case for a fixed
 * nsteps, march.cpp:15-59).  With on != 0, hamstr_step and hamstr_batch_step
 * stop iterating a case at the first read-back where its nonlinear L2 <= tol
 * (tol > 0), or once its L2 has not halved within stall_window iterations (a
 * limit cycle; 0 = never).  A retired case keeps its state and drops out of
 * every launch; the calls return once every case has retired.  Calling this
 * again makes every case live and restarts the iteration count. */
int hamstr_set_retire(hamstr_ctx *ctx, int on, int stall_window);
/* hamstr_step and hamstr_batch_step replay a captured HIP graph for each run
 * of up to 64 iterations between norm read-backs (on by default; not while
 * timing).  on = 0 launches every kernel from the host instead. */
int hamstr_set_graph(hamstr_ctx *ctx, int on);
/* Captures now (instead of inside the first step call) the graph that
 * hamstr_step / hamstr_batch_step use for read-backs every check_every steps;
 * runs no iteration. */
int hamstr_prepare_graph(hamstr_ctx *ctx, int check_every);
/* status[k]: 0 live, 1 converged, 2 stalled, 3 diverged, 4 held (with retirement on, a
 * case whose norms stop being finite retires instead of the call failing with
 * HAMSTR_ERR_NAN); iters[k]: the iteration, counted from hamstr_set_retire, at
 * which case k converged or was found stalled or diverged. */
int hamstr_case_status(const hamstr_ctx *ctx, int *status, long *iters);
/* With retirement on: live = 0 holds case k out of every launch (status 4),
 * live = 1 makes it live again from its current state, with a fresh stall
 * window (to continue a case another context gave up on, after
 * hamstr_set_cells).  hamstr_iterations: iterations run since
 * hamstr_set_retire. */
int hamstr_set_case_live(hamstr_ctx *ctx, int k, int live);
long hamstr_iterations(const hamstr_ctx *ctx);

/* Autotuner: times every FAST launch group of the ADI phases at several
 * persistent grid sizes (all chains; 8, 4, 2, 1 workgroups per CU), 0 to 3
 * levels of cyclic reduction before the parallel cyclic reduction and, where
 * the group's chains fit it, both one-row-per-thread kernels (all rows in
 * LDS, or half of them: HAMSTR_KERNEL_*), on the current state, and where
 * the Jacobians can be computed again the source (HAMSTR_JAC_*) by timing
 * the iteration as a captured block runs it; keeps the fastest and restores
 * q, r, r0, dq, sigma and the NaN flag (on an error as well).  Per kernel it times
 * every grid at the group's current depth, then every other depth at the
 * best grid.  best_ms[k] (k < cap) gets group k's best time (-1 for groups
 * with nothing to tune). */
int hamstr_autotune(hamstr_ctx *ctx, int reps, double *best_ms, int cap, int *ngroups);
/* ADI kernels of a launch group */
#define HAMSTR_KERNEL_EXACT 0 /* one thread per chain, the reference's order */
#define HAMSTR_KERNEL_ROW 1   /* FAST: one thread per row, every row in LDS */
#define HAMSTR_KERNEL_HALF 2  /* FAST, FP64: one thread per row, half the rows in LDS
                                 (the rest in global scratch), at most 168 VGPRs:
                                 two 384-row chains per CU, or four of at most
                                 192 rows (a 192-thread build, 28.5 KB of LDS) */
#define HAMSTR_KERNEL_TWO 3   /* FAST: chains longer than a workgroup, two rows per thread */
#define HAMSTR_KERNEL_HALF2 4 /* FAST, open chains (FP64 or FP32): half the rows in LDS, two
                                 chains per 768-thread workgroup (12 waves, 3 per SIMD) */
/* Launch groups, 8 ints each: phase, ADI kernel (HAMSTR_KERNEL_*), final
 * solve (0 open, 1 two-row pairs, 2 odd cyclic), chains, max rows, grid,
 * block, cyclic-reduction levels. */
int hamstr_launch_groups(const hamstr_ctx *ctx, int *out8, int cap, int *n);
/* Sets FAST launch group `group`'s persistent grid and its cyclic-reduction
 * levels (0..3) by hand, as the autotuner would. */
int hamstr_set_launch(hamstr_ctx *ctx, int group, int grid, int cr_levels);
/* FAST launch groups of open chains (not HAMSTR_KERNEL_TWO): the parallel
 * cyclic reduction after the cyclic-reduction levels stops after pcr_levels
 * levels (0..4), and the block Thomas algorithm finishes the 2^pcr_levels
 * interleaved subsystems left, one quad of lanes per subsystem (-1, the
 * default: reduce until every row is decoupled).  The solution equals the
 * other orderings up to rounding; every kernel gives the same bits for the
 * same setting.  hamstr_autotune tries -1, 2, 3 and 4.  No reference
 * counterpart: the reference solves each chain with one sequential block
 * Thomas sweep (triSolvers.cpp:769-847). */
int hamstr_set_thomas(hamstr_ctx *ctx, int group, int pcr_levels);
int hamstr_thomas(const hamstr_ctx *ctx, int group, int *pcr_levels);
/* Chooses a FAST launch group's kernel by hand: HAMSTR_KERNEL_ROW, or
 * HAMSTR_KERNEL_HALF where the group's chains fit it (FP64 left-hand side, at
 * most 384 rows, open chains or closed chains whose final solve is two-row
 * cycles; it always runs at least one cyclic-reduction level), or
 * HAMSTR_KERNEL_HALF2 for open chains of 257 to 384 rows (FP64 or FP32).
 * All give the same solution up to the sign of zero entries.  FP64 contexts
 * start with HAMSTR_KERNEL_HALF2 for open chains of more than 256 rows;
 * everything else starts with HAMSTR_KERNEL_ROW. */
int hamstr_set_adi_kernel(hamstr_ctx *ctx, int group, int kernel);
/* Where the fused FAST iteration's ADI kernels take the chain faces' frozen-
 * Roe Jacobians from (no reference counterpart: the reference always stores
 * them, computeRHS.cpp:362-369, and ADI reads them, ADI.cpp:96-97).
 * HAMSTR_JAC_STORED (the default): k_rhs_faces stores every face's 32
 * entries and the ADI kernels read them.  HAMSTR_JAC_FROM_STATES: it stores
 * each chain face's two reconstructed states (8 values) and the ADI kernels
 * compute the Jacobians from them.  HAMSTR_JAC_FROM_Q: the ADI kernels
 * gather and reconstruct the states again from q.  HAMSTR_JAC_STATES_LINES:
 * phase 0's chain faces keep stored Jacobians and the other phases' use
 * stored states (where phase 0's chains come first, e.g. the loops of an
 * O-grid).  The computed blocks are
 * the stored ones bit for bit; orphan faces and the last faces of open
 * chains keep stored Jacobians.  hamstr_autotune picks the fastest.  The
 * computed sources are refused (HAMSTR_ERR_ARG) where a group's kernel cannot
 * compute them (EXACT, chains longer than a workgroup, FP32 odd cycles).
 * The API's own hamstr_compute_rhs / hamstr_adi always store and read every
 * face's. */
#define HAMSTR_JAC_STORED 0
#define HAMSTR_JAC_FROM_Q 1
#define HAMSTR_JAC_FROM_STATES 2
#define HAMSTR_JAC_STATES_LINES 3
int hamstr_set_jacobian_source(hamstr_ctx *ctx, int source);
int hamstr_jacobian_source(const hamstr_ctx *ctx, int *source);
/* Whether the fused iteration's face kernel stages each workgroup's stencil
 * rows in LDS (no reference counterpart: a launch choice for the
 * reconstruction of computeRHS.cpp:53-298 and computeLinearRHS.cpp:53-298).
 * A workgroup of 256 processed faces reads the consecutive chain slots from
 * its first face's stencil to its last one's; each slot's state, increment
 * and updated state are formed once into LDS instead of gathered by every
 * face that reads them.  On by default where every workgroup's range fits 520
 * slots (chains of a few cells at order 5 may not: then 1 is refused with
 * HAMSTR_ERR_ARG); 0 gathers directly.  The results are the same bit for bit. */
int hamstr_set_face_staging(hamstr_ctx *ctx, int on);
int hamstr_face_staging(const hamstr_ctx *ctx, int *on);

/* SOLVER::initFlow (initFlow.cpp:12-93): q := freestream. */
int hamstr_init_flow(hamstr_ctx *ctx);
/* freestream (rinf, rinf*uinf, rinf*vinf, einf) as initFlow computes it */
int hamstr_freestream(const hamstr_ctx *ctx, double *out4);

/* freestream state of case k of a batch */
int hamstr_case_freestream(const hamstr_ctx *ctx, int k, double *out4);

int hamstr_set_field(hamstr_ctx *ctx, int field, const double *host);
int hamstr_get_field(hamstr_ctx *ctx, int field, double *host);
/* cells [first, first+count) of a per-cell field (q, qt, r, r0, dq, sigma), e.g.
 * one case of a batch */
int hamstr_get_cells(hamstr_ctx *ctx, int field, int first, int count, double *host);
int hamstr_set_cells(hamstr_ctx *ctx, int field, int first, int count, const double *host);

/* Solution files: the intent of SOLNBLOCK::nwrite / outform (solnBlock.h:23,26,
 * read at io.cpp:35,41 and never used by the reference).  A file holds the
 * state q of every case, the iteration count `step` given by the caller, the
 * per-case Mach and alpha, and a fingerprint of the mesh (faces, chainStart,
 * chainConn); little-endian, 64-byte header, 32 bytes per case, then
 * q[4*nCell] doubles.  Writing goes through `path`.tmp and a rename.  Every
 * scheme's step depends on q alone (rk3 rewrites qt before reading it), so a
 * run restarted from a file continues bit for bit.  Reading fails with
 * HAMSTR_ERR_ARG for a file of another mesh or batch. */
int hamstr_write_solution(hamstr_ctx *ctx, const char *path, long step);
int hamstr_read_solution(hamstr_ctx *ctx, const char *path, long *step);

/* ---- the iteration (SOLVER methods) ------------------------------------- */

/* SOLVER::computeRHS(char*) (computeRHS.cpp:17-427); implicit != 0 is
 * "implicit" (also stores the face Jacobians ff).  L2/Linf of the density
 * residual: L2 = sqrt(sum r_rho^2)/nCell.  Either pointer may be NULL. */
int hamstr_compute_rhs(hamstr_ctx *ctx, int implicit, double *L2, double *Linf);

/* SOLVER::ADI (ADI.cpp:15-266), including computeLinearRHS per sweep and the
 * final q += dq.  Returns the linear-residual norms of the last sweep. */
int hamstr_adi(hamstr_ctx *ctx, double *L2lin, double *Linflin);

/* SOLVER::computeLinearRHS (computeLinearRHS.cpp:13-371) on the current
 * q, dq, r0, sigma, ff. */
int hamstr_compute_linear_rhs(hamstr_ctx *ctx, double *L2lin, double *Linflin);

/* SOLVER::updateSolution (march.cpp:142-164) with the current r and sigma:
 * dst = src + (coef*CFL/sigma) r per cell, src and dst each HAMSTR_FIELD_Q or
 * HAMSTR_FIELD_QT (may be the same). */
int hamstr_update_solution(hamstr_ctx *ctx, int src_field, int dst_field, double coef);

/* Selects the stepSolution branch hamstr_step / hamstr_batch_step run
 * (HAMSTR_SCHEME_*; SOLVER::stepType, solver.cpp:22).  The explicit schemes
 * compute no Jacobians, so the LHS precision and ADI mode do not affect them.
 * hamstr_get_scheme returns the current one (-1 for a NULL context). */
int hamstr_set_scheme(hamstr_ctx *ctx, int scheme);
int hamstr_get_scheme(const hamstr_ctx *ctx);

/* nsteps x SOLVER::stepSolution (march.cpp:64-136) with the selected scheme,
 * by default its ADI branch (march.cpp:112-117).  Per-step norms are written
 * to l2rhs[n] / l2lin[n] when the arrays are not NULL: l2rhs the nonlinear L2
 * of the step's first computeRHS (the residual of the state the step starts
 * from), l2lin the norm the reference prints after the step (ADI: the last
 * computeLinearRHS; euler, rk3: the last computeRHS).  When tol > 0 the loop
 * stops after the first step whose l2rhs <= tol; *done gets the number of
 * steps run.  The norms are read back every check_every steps (>= 1); the
 * steps in between run without host synchronisation. */
int hamstr_step(hamstr_ctx *ctx, int nsteps, double tol, int check_every,
                double *l2rhs, double *l2lin, int *done);

/* Norms of the last completed step or call: out4 = (L2, Linf) of the
 * nonlinear residual (computeRHS) and (L2, Linf) of the linear residual
 * (computeLinearRHS). */
int hamstr_last_norms(const hamstr_ctx *ctx, double *out4);

/* Blocks until the context's stream is idle. */
int hamstr_sync(hamstr_ctx *ctx);

/* Turns per-kernel HIP-event timing of hamstr_step on (on != 0) or off. */
int hamstr_set_timing(hamstr_ctx *ctx, int on);

/* Average per-step duration (milliseconds, HIP events recorded on the
 * context's stream around every stage of every step) of the stages of the
 * last hamstr_step call with timing on.  *n gets the number written (at most
 * cap); hamstr_last_timing_names gives their comma-separated names
 * (rhs_faces, rhs_cells, adi_phase0, adi_phase1, ..., accum_dq, lin_faces,
 * lin_cells). */
int hamstr_last_timings(hamstr_ctx *ctx, double *ms, int cap, int *n);
int hamstr_last_timing_names(hamstr_ctx *ctx, char *buf, int cap);

/* ---- per-function entry points, batched over independent problems -------
 * Inputs and outputs are host arrays; each call copies to the device,
 * launches, and copies back.  Used for per-function parity. */

/* SOLVER::fluxRoe2D (interfaceFlux.cpp:24-169), n problems:
 * nvec[2n], ql[4n], qr[4n] -> flux[4n], sr[n].  Returns HAMSTR_ERR_NAN if a
 * spectral radius is NaN. */
int hamstr_flux_roe2d(int n, const double *nvec, const double *ql, const double *qr,
                      double gamma, double *flux, double *sr);

/* SOLVER::jacobianRoe2D (jacobianRoe.cpp:9-70, IMODE 1):
 * -> lmat[16n], rmat[16n] row-major.  lhs selects FP64 or FP32 arithmetic
 * (FP32 results are returned widened to double). */
int hamstr_jacobian_roe2d(int n, const double *nvec, const double *ql, const double *qr,
                          double gamma, int lhs, double *lmat, double *rmat);

/* RECONSTRUCTION::MUSCL / WENO / MUSCL_DERIV / WENO_DERIV
 * (reconstruction.cpp:16-70, 133-160, 168-236, 353-386) on nchain chains
 * of imax rows each: f[nchain*imax*4] (and dq for the derivatives) ->
 * ql, qr [nchain*imax*4].  Rows outside the reference's computed range are
 * left untouched.  is/ie/th/qt/eps as in the reference call sites. */
int hamstr_muscl(int nchain, int imax, int is, int ie, double th, double qt, double eps,
                 const double *f, double *ql, double *qr);
int hamstr_weno(int nchain, int imax, int is, int ie, double eps,
                const double *f, double *ql, double *qr);
int hamstr_muscl_deriv(int nchain, int imax, int is, int ie, double th, double qt, double eps,
                       const double *f, const double *dq, double *ql, double *qr);
int hamstr_weno_deriv(int nchain, int imax, int is, int ie, double eps,
                      const double *f, const double *dq, double *ql, double *qr);

/* TRISOLVER::blockTridag4 / blockTridagPeriodic4 (triSolvers.cpp:769-847,
 * 850-1046) on nsys independent systems of N block rows:
 * a, b, c [nsys*N*16], f [nsys*N*4]; f is overwritten with the solution.
 * a, b, c are not modified. */
int hamstr_block_tridag4(int nsys, int N, const double *a, const double *b, const double *c,
                         double *f);
int hamstr_block_tridag_periodic4(int nsys, int N, const double *a, const double *b,
                                  const double *c, double *f);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the solve of
 * hamstr_block_tridag4 (periodic = 0) or hamstr_block_tridag_periodic4
 * (periodic = 1) on DEVICE pointers a, b, c [nsys*N*16], f [nsys*N*4] (f
 * overwritten), enqueued on `stream` (a hipStream_t; NULL: the null stream)
 * without synchronising and without any allocation.  Its scratch rows live in
 * a library-owned device array of HAMSTR_NOALLOC_MAX_ROWS * 36 doubles
 * (4.5 MiB) on each GPU, held until the process ends and shared by every call
 * on that GPU: calls that overlap on one GPU race, so issue them on one
 * stream.  nsys*N above HAMSTR_NOALLOC_MAX_ROWS is HAMSTR_ERR_ARG. */
#define HAMSTR_NOALLOC_MAX_ROWS 16384
int hamstr_block_tridag4_noalloc(int periodic, int nsys, int N, const double *a, const double *b,
                                 const double *c, double *f, void *stream);

/* ---- the 3D / strand building blocks -----------------------------------------
 * Five conserved variables (rho, rho u, rho v, rho w, E).  The reference's 3D
 * build does not compile (io.cpp:306), but these functions are compiled in its
 * 2D build; each is bit-identical to it.
 * SOLVER::fluxRoe3D (interfaceFlux.cpp:176-319): nxyz[n*3], ql, qr [n*5] ->
 * flux [n*5], spectral radius sr[n]; the unit-normal flux, NOT scaled by the
 * face area (unlike fluxRoe2D). */
int hamstr_flux_roe3d(int n, const double *nxyz, const double *ql, const double *qr, double gamma,
                      double *flux, double *sr);
/* SOLVER::jacobianRoe3D, IMODE 1 (jacobianRoe.cpp:92-620): lmat, rmat
 * [n*25], row-major 5x5, not scaled by the face area. */
int hamstr_jacobian_roe3d(int n, const double *nxyz, const double *ql, const double *qr, double gamma,
                          double *lmat, double *rmat);
/* TRISOLVER::blockTridag / blockTridagPeriodic with nq = 5
 * (triSolvers.cpp:443-536, 540-766): a, b, c [nsys*N*25], f [nsys*N*5];
 * f is overwritten with the solution, a, b, c are not modified. */
int hamstr_block_tridag5(int nsys, int N, const double *a, const double *b, const double *c,
                         double *f);
int hamstr_block_tridag_periodic5(int nsys, int N, const double *a, const double *b,
                                  const double *c, double *f);

/* Text of the last error on this thread. */
const char *hamstr_last_error(void);

/* ---- learned warm start ----------------------------------------------------
 * No counterpart in the reference (which starts every run from the freestream,
 * initFlow.cpp:107-142).  An MLP 6 -> 256 x 4 (SiLU) -> 4 maps normalised
 * per-cell features (Mach, alpha, thickness, centroid x, centroid y, wall
 * distance) to the normalised deviation of the primitive state (rho, u, v, p)
 * from the case's freestream.  The 256x256 layers run on BF16 matrix cores with
 * FP32 accumulation (forward and backward); FP32 master weights, Adam. */
typedef struct hamstr_mlp hamstr_mlp;
int hamstr_mlp_create(unsigned seed, hamstr_mlp **out);
int hamstr_mlp_destroy(hamstr_mlp *mlp);
int hamstr_mlp_num_params(void);                       /* 200196 */
int hamstr_mlp_get_params(hamstr_mlp *mlp, float *host);
int hamstr_mlp_set_params(hamstr_mlp *mlp, const float *host);
/* y[n*4] = MLP(x[n*6]) */
int hamstr_mlp_forward(hamstr_mlp *mlp, int n, const float *x, float *y);
/* mean-squared-error loss and gradient (n a multiple of 32) */
int hamstr_mlp_loss_grad(hamstr_mlp *mlp, int n, const float *x, const float *t, float *loss, float *grad);
/* device-resident training set, then steps of Adam on random batches with a
 * cosine learning-rate decay from lr to lr_final; losses[steps] may be NULL */
int hamstr_mlp_set_data(hamstr_mlp *mlp, long n, const float *x, const float *t);
int hamstr_mlp_train(hamstr_mlp *mlp, int steps, int batch, float lr, float lr_final, unsigned seed,
                     float *losses);
/* hamstr_mlp_train replays one captured training step as a HIP graph (on by
 * default); on = 0 launches its kernels from the host every step. */
int hamstr_mlp_set_graph(hamstr_mlp *mlp, int on);
/* q := conserved form of freestream + the MLP's predicted deviation, from raw
 * per-cell features feat[nCell*6] and the normalisation nrm[20] (feature
 * mean[6], std[6], target mean[4], std[4]); cells whose prediction is not a
 * valid state take their case's freestream (*fallback_cells counts them). */
int hamstr_warm_start(hamstr_ctx *ctx, hamstr_mlp *mlp, const float *feat, const float *nrm,
                      int *fallback_cells);
const char *hamstr_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HAMSTR_HIP_H */
