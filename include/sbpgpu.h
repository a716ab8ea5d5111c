/*
 * sbpgpu.h — C ABI of the MI355X (gfx950) Euler RK4 time-step library.
 *
 * Plain pointers and sizes only. Each entry point names the reference
 * interface it replaces (paths relative to the SummationByParts repository).
 * All arrays are C-order. A field/state is (4, ny, nx): rho, rhou, rhov, e.
 * Metrics are (5, ny, nx): J, J*xi_x, J*xi_y, J*eta_x, J*eta_y
 * (sbp/src/grid.rs:15-21). "Device" pointers are HIP device memory of the
 * context's device; "host" pointers are ordinary CPU memory.
 *
 * Every function returns an sbp_status. Errors map one-to-one to the panics
 * of the reference path; the reference's Rust side would re-panic on them.
 */
#ifndef SBPGPU_H
#define SBPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SBP_OK = 0,
  SBP_ERR_INVALID = 1,     /* bad handle, enum value or pointer */
  SBP_ERR_SHAPE = 2,       /* n < 2M or n < N (sbp/src/operators/algos.rs:118-121) */
  SBP_ERR_DETJ = 3,        /* det J <= 0 (sbp/src/grid.rs:117) */
  SBP_ERR_RHO = 4,         /* rho <= 0 (euler/src/lib.rs:633, euler/src/vortex.rs:74,81,107) */
  SBP_ERR_PRESSURE = 5,    /* p <= 0 (euler/src/lib.rs:642, euler/src/vortex.rs:79) */
  SBP_ERR_GAMMA = 6,       /* GAMMA not set (euler/src/lib.rs:626,686,1166) */
  SBP_ERR_EDGE = 7,        /* edge lengths differ (euler/src/lib.rs:1161-1165) */
  SBP_ERR_HIP = 8,         /* HIP runtime failure */
  SBP_ERR_NOT_UPWIND = 9,  /* "This is not an upwind operator" (euler/src/lib.rs:597) */
  SBP_ERR_OOM = 10,        /* device allocation failed */
  SBP_ERR_GAMMA_SET = 11,  /* GAMMA already set to another value (OnceCell::set, euler/src/lib.rs:15) */
  SBP_ERR_SIZE = 12,       /* a size beyond the entry point's stated maximum */
  SBP_ERR_EXPR = 13        /* an expression does not compile or fails to evaluate (evalexpr errors,
                              multigrid/src/eval/evalexpr.rs: .unwrap() on build and eval) */
} sbp_status;

/* sbp/src/operators.rs:152-165 */
typedef enum {
  SBP_UPWIND4 = 0,
  SBP_UPWIND4H2 = 1,
  SBP_UPWIND9 = 2,
  SBP_UPWIND9H2 = 3,
  SBP_SBP4 = 4,
  SBP_SBP8 = 5
} sbp_operator;

typedef enum { SBP_DIFF = 0, SBP_DISS = 1 } sbp_kind;

/* sbp/src/utils.rs:15-129 (Direction) */
typedef enum { SBP_NORTH = 0, SBP_SOUTH = 1, SBP_EAST = 2, SBP_WEST = 3 } sbp_direction;

/* euler/src/lib.rs:737-745 (BoundaryCharacteristic: This, Grid(j), Vortex,
 * Interpolate(j, operator), MultiGrid(pieces); the last two are set with
 * sbp_system_set_bc_interpolate / sbp_system_set_bc_multigrid) */
typedef enum {
  SBP_BC_THIS = 0,
  SBP_BC_GRID = 1,
  SBP_BC_VORTEX = 2,
  SBP_BC_INTERPOLATE = 3,
  SBP_BC_MULTIGRID = 4,
  SBP_BC_EVAL = 5   /* BoundaryCharacteristic::Eval (euler/src/lib.rs:741): set with sbp_system_set_bc_eval */
} sbp_bc_kind;

/* 2:1 interface interpolation operators (sbp/src/operators/interpolation.rs:1-13) */
typedef enum { SBP_INTERP4 = 0, SBP_INTERP8 = 1, SBP_INTERP9 = 2, SBP_INTERP9H2 = 3 } sbp_interp_op;

/* RHS selection: AUTO follows multigrid/src/system.rs:521-525 (upwind iff both
 * operators are upwind); TRAD = euler::System::advance / RHS_trad;
 * UPWIND = euler::System::advance_upwind / RHS_upwind. */
typedef enum { SBP_RHS_AUTO = 0, SBP_RHS_TRAD = 1, SBP_RHS_UPWIND = 2 } sbp_rhs_kind;

/* FP64: the drop-in. MIXED: FP32 flux/stencil/dissipation arithmetic with
 * FP64 state, metrics and RK4 sum (the stage increments are stored in FP32). */
typedef enum { SBP_FP64 = 0, SBP_MIXED = 1 } sbp_precision;

/* euler/src/vortex.rs:3-21 (Vortice, VortexParameters; at most 5 vortices) */
typedef struct {
  double x0, y0, rstar, eps;
} sbp_vortice;

typedef struct {
  int32_t nvortices;
  int32_t pad_;
  sbp_vortice vortices[5];
  double mach;
} sbp_vortex_params;

typedef struct sbp_context_s* sbp_context;
typedef struct sbp_system_s* sbp_system;
typedef struct sbp_expr_s* sbp_expr;

/* The two expression forms of the multigrid configuration
 * (multigrid/src/input.rs:165-190): conservative variables (rho, rhou, rhov,
 * e) or primitive ones (rho, u, v, p). */
typedef enum { SBP_EXPR_CONSERVATION = 0, SBP_EXPR_PRESSURE = 1 } sbp_expr_form;

/* Human-readable text of a status code. */
const char* sbp_status_string(int status);

/* euler::GAMMA (euler/src/lib.rs:15; set once, multigrid/src/parsing.rs:48).
 * Setting the value it already holds is OK; another value is SBP_ERR_GAMMA_SET. */
int sbp_set_gamma(double gamma);
int sbp_get_gamma(double* gamma);

/* ---- Operator traits (sbp/src/operators.rs:11-106) ---------------------- */

/* Block sizes: M closure rows, N closure columns, D interior width, whether
 * the operator is h2 (SbpOperator1d::is_h2) and upwind (::upwind().is_some()). */
int sbp_operator_info(int op, int* M, int* N, int* D, int* is_h2, int* is_upwind);
/* SbpOperator1d::h (M values). */
int sbp_operator_h(int op, double* h_out);
/* The coefficient blocks as the kernels use them: start (M*N), diag (D), end (M*N). */
int sbp_operator_coefficients(int op, int kind, double* start, double* diag, double* end);

/* ---- Context ------------------------------------------------------------ */

/* One context per device; owns a HIP stream (or uses the caller's). */
int sbp_context_create(int device, sbp_context* out);
int sbp_context_destroy(sbp_context ctx);
/* Use `hip_stream` (a hipStream_t) for all work of this context; NULL = own stream. */
int sbp_context_set_stream(sbp_context ctx, void* hip_stream);
int sbp_context_synchronize(sbp_context ctx);

/* ---- Operator application on device buffers ----------------------------
 * prev, fut: (ny, nx) device arrays with row pitch ld_prev / ld_fut elements.
 * fut is fully overwritten (sbp/src/operators/algos.rs:228,306). */

/* SbpOperator2d::diffxi (sbp/src/operators.rs:56-61; algos.rs:101-187) */
int sbp_diffxi(sbp_context ctx, int op, const double* prev, int64_t ld_prev, double* fut, int64_t ld_fut, int ny, int nx);
/* SbpOperator2d::diffeta (sbp/src/operators.rs:62-64; algos.rs:371-519) */
int sbp_diffeta(sbp_context ctx, int op, const double* prev, int64_t ld_prev, double* fut, int64_t ld_fut, int ny, int nx);
/* UpwindOperator2d::dissxi (sbp/src/operators.rs:91-96) */
int sbp_dissxi(sbp_context ctx, int op, const double* prev, int64_t ld_prev, double* fut, int64_t ld_fut, int ny, int nx);
/* UpwindOperator2d::disseta (sbp/src/operators.rs:98-100) */
int sbp_disseta(sbp_context ctx, int op, const double* prev, int64_t ld_prev, double* fut, int64_t ld_fut, int ny, int nx);

/* InterpolationOperator::fine2coarse (fine2coarse = 1) / coarse2fine
 * (sbp/src/operators.rs:108-113; interpolation.rs:15-50) of a device vector
 * of nin points into nout points: fine = 2 coarse - 1 (2 (coarse - 1) for
 * SBP_INTERP9H2), and every tap of the operator must fall inside the nin
 * input points (short edges make the reference's indexing panic), else
 * SBP_ERR_SHAPE; the lengths are checked before the handle, so a NULL ctx
 * validates a coupling without device work. out is fully overwritten. */
int sbp_interpolate(sbp_context ctx, int interp_op, int fine2coarse, const double* in, int nin, double* out, int nout);
/* Table shapes {block rows, block cols, diag rows, diag cols, jump0, jump1}
 * and coefficients (block row-major, then diag row-major) of one direction of
 * an interpolation operator (host only, no device work). */
int sbp_interp_info(int interp_op, int fine2coarse, int* shape6);
int sbp_interp_coefficients(int interp_op, int fine2coarse, double* block, double* diag);

/* ---- Block-level right-hand side on device buffers ----------------------
 * euler::RHS_trad / RHS_upwind (euler/src/lib.rs:483-515, :571-616) for one
 * block with operator pair (op_eta, op_xi) (the reference's tuple order,
 * multigrid/src/parsing.rs:164). k (out), y (in): (4, ny, nx) contiguous;
 * metrics: (5, ny, nx) contiguous; boundaries[dir]: 4 x n contiguous arrays
 * (n = nx for north/south, ny for east/west) holding the boundary data g of
 * BoundaryTerms (euler/src/lib.rs:760). kind: SBP_RHS_TRAD or SBP_RHS_UPWIND.
 * SBP_ERR_SIZE when 5 x nx x ny doubles reach 2^32 bytes (also for
 * sbp_rhs_no_sat). */
int sbp_rhs(sbp_context ctx, int kind, int op_eta, int op_xi, int ny, int nx, double* k, const double* y,
            const double* metrics, const double* const* boundaries);

/* euler::RHS_no_SAT (euler/src/lib.rs:518-568): the right-hand side without
 * boundary penalties, with the upwind dissipation iff both operators are
 * upwind (op.upwind().is_some(), sbp/src/operators.rs:137-150). Buffers as
 * sbp_rhs. */
int sbp_rhs_no_sat(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, double* k, const double* y, const double* metrics);

/* euler::SAT_characteristics (euler/src/lib.rs:1002-1013): k += the
 * characteristic SATs of the north, south, east and west edges in that
 * order (corners get two), SAT_characteristic :1150-1252. k, y (4, ny, nx),
 * metrics (5, ny, nx) contiguous device buffers; boundaries[dir] as sbp_rhs.
 * k accumulates (the reference adds into its edge views of k). */
int sbp_sat_characteristics(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, double* k, const double* y,
                            const double* metrics, const double* const* boundaries);
/* One of euler::SAT_FUNCTIONS / SAT_north, SAT_south, SAT_east, SAT_west
 * (euler/src/lib.rs:1015-1144): k += the SAT of edge `dir` with boundary
 * data `boundary` (4 x n contiguous device array). */
int sbp_sat_edge(sbp_context ctx, int dir, int op_eta, int op_xi, int ny, int nx, double* k, const double* y,
                 const double* metrics, const double* boundary);

/* Field::h2_err (euler/src/lib.rs:377-432) of two (4, ny, nx) contiguous
 * device fields: sqrt(Σ_c Σ_ij hx_i hy_j (a - b)²) with the operators' H
 * weights, summed compensated in FP64 on the device; *out is a host double.
 * Allocates (and frees) a small device buffer of partial sums per call.
 * SBP_ERR_SHAPE when n < 2M along an axis; SBP_ERR_SIZE above 2^31 - 1
 * points (both checked before the handle). */
int sbp_h2_err(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, const double* a, const double* b, double* out);
/* sbp_h2_err without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same kernels
 * run on a library-owned static device array of 16 KiB per GPU (1024
 * per-workgroup partial sums and the result), which serves every size up to
 * the same maximum of 2^31 - 1 points (SBP_ERR_SIZE beyond it). Calls that
 * overlap on one GPU (other threads or contexts) share that array and race;
 * sequential calls need no clearing between them. */
int sbp_h2_err_noalloc(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, const double* a, const double* b,
                       double* out);

/* sbp::utils::h2linspace (sbp/src/utils.rs:133-140): n >= 2 host values,
 * start, start + h/2, start + 3h/2, ..., end with h = (end - start)/(n - 2). */
int sbp_h2linspace(double start, double end, int n, double* out);

/* Grid::metrics (sbp/src/grid.rs:64-66,95-139) on device buffers: x, y (ny, nx)
 * contiguous in, metrics (5, ny, nx) out. */
int sbp_metrics(sbp_context ctx, int op_eta, int op_xi, int ny, int nx, const double* x, const double* y, double* metrics);

/* ---- Systems: one or many blocks stepped by fused RK4 -------------------
 * Replaces euler::System (euler/src/lib.rs:20-186) for one block with periodic
 * BCs, and multigrid's SingleThreadedSystem / DistributedSystem
 * (multigrid/src/system.rs:456-539, :638-1172) for many. */

/* Blocks with sizes (ny[i], nx[i]), operator pairs (op_eta[i], op_xi[i]) and
 * host coordinate arrays x[i], y[i] of shape (ny[i], nx[i]). Metrics are
 * computed on the device (BaseSystem::create, multigrid/src/system.rs:73-78).
 * Boundary conditions default to SBP_BC_THIS on every edge. At most 32767
 * blocks; SBP_ERR_SHAPE for a block below its operators' closures,
 * SBP_ERR_SIZE for a block whose 5 planes (nx rounded up to a multiple of
 * 16, times ny, doubles) reach 2^32 bytes (about 10 300 x 10 300 points),
 * both checked before the context and the coordinate pointers. */
int sbp_system_create(sbp_context ctx, int nblocks, const int* ny, const int* nx, const int* op_eta, const int* op_xi,
                      const double* const* x, const double* const* y, sbp_system* out);
int sbp_system_destroy(sbp_system sys);
/* BoundaryCharacteristic of one edge: SBP_BC_THIS, SBP_BC_GRID (neighbour),
 * SBP_BC_VORTEX (vortex required). */
int sbp_system_set_bc(sbp_system sys, int block, int dir, int kind, int neighbour, const sbp_vortex_params* vortex);
/* BoundaryCharacteristic::Interpolate(neighbour, operator) (euler/src/lib.rs:743,
 * :816-828): the neighbour's opposite edge through a 2:1 interpolation
 * operator, fine to coarse when this block has fewer points along ξ than the
 * neighbour (the reference compares nx for every edge). SBP_ERR_EDGE unless
 * the edge lengths are the operator's coarse/fine pair and long enough for
 * every tap of its stencil rows (as sbp_interpolate). */
int sbp_system_set_bc_interpolate(sbp_system sys, int block, int dir, int neighbour, int interp_op);
/* BoundaryCharacteristic::MultiGrid(pieces) (euler/src/lib.rs:744, :829-849):
 * npieces triples (block, start, end) of the sources' opposite edges, copied
 * one after another, each starting on the last point of the one before; the
 * first point of every piece but the last is halved. SBP_ERR_EDGE unless the
 * pieces lie inside their sources and cover this edge exactly. */
int sbp_system_set_bc_multigrid(sbp_system sys, int block, int dir, int npieces, const int* pieces);
int sbp_system_set_rhs_kind(sbp_system sys, int kind);
int sbp_system_set_precision(sbp_system sys, int precision);
/* State of one block, (4, ny, nx) host arrays. */
int sbp_system_set_state(sbp_system sys, int block, const double* host);
int sbp_system_get_state(sbp_system sys, int block, double* host);
/* Device pointer and row pitch of a block's current state (4 planes, plane
 * stride returned in *plane_stride); valid until the next call on `sys`. */
int sbp_system_state_device(sbp_system sys, int block, double** dev, int64_t* ld, int64_t* plane_stride);
/* Metrics of one block, (5, ny, nx) host array. */
int sbp_system_get_metrics(sbp_system sys, int block, double* host);
/* Field::vortex on every block (euler/src/lib.rs:319-328, vortex.rs:23-113). */
int sbp_system_vortex(sbp_system sys, double t, const sbp_vortex_params* params);
/* Field::vortex on one block: independent solves batched in one system each
 * start from their own vortex (the reference builds one euler::System per
 * solve, euler/src/lib.rs:29-48,76-80). */
int sbp_system_vortex_block(sbp_system sys, int block, double t, const sbp_vortex_params* params);
/* Block `block` steps with dt*scale and its time is t*scale (independent
 * solves with their own step size in one batched system; each reference
 * System::advance takes its own dt, euler/src/lib.rs:50-74). Default 1. */
int sbp_system_set_dt_scale(sbp_system sys, int block, double scale);
/* One RHS evaluation of the current state at time t (the closure passed to
 * integrate, multigrid/src/system.rs:508-527), k copied to k_host:
 * concatenation of each block's (4, ny, nx). */
int sbp_system_rhs(sbp_system sys, double t, double* k_host);
/* nsteps classical RK4 steps of size dt (integrate::<Rk4>,
 * utils/integrate/src/lib.rs:176-222), fused and graph-captured. */
int sbp_system_advance(sbp_system sys, double dt, int64_t nsteps);
int sbp_system_get_time(sbp_system sys, double* t);
int sbp_system_set_time(sbp_system sys, double t);
/* System::max_dt (multigrid/src/system.rs:542-589). */
int sbp_system_max_dt(sbp_system sys, double* out);
/* System::error against the vortex at the system time: the sum over blocks of
 * Field::h2_err (multigrid/src/system.rs:618-635; euler/src/lib.rs:377-432),
 * compensated FP64 sums on the device. */
int sbp_system_error_vortex(sbp_system sys, const sbp_vortex_params* params, double* out);
/* Field::h2_err of one block against the vortex at the block's time
 * (euler/src/lib.rs:377-432; euler/tests/convergence.rs:48-51). */
int sbp_system_error_vortex_block(sbp_system sys, int block, const sbp_vortex_params* params, double* out);
/* ||q||_H^2 summed over blocks and components (the H-norm energy used for
 * drift reports), compensated FP64 sums on the device. */
int sbp_system_energy(sbp_system sys, double* out);
/* euler::System::advance_adaptive (euler/src/lib.rs:149-185): embedded
 * Bogacki-Shampine steps (integrate_embedded_rk, utils/integrate/src/lib.rs:
 * 138-151, 176-251) of trial size *guess_dt until the loop time reaches dt,
 * accepting a step when the sum over blocks of Field::h2_err between the
 * previous state and the embedded solution is below maxerr (the guess then
 * grows by 1.05, else shrinks by 0.8); *guess_dt returns the last guess. As
 * in the reference, the loop time grows by the trial step on every attempt
 * and once more on acceptance, and boundary data sees the loop's stage times
 * (the reference applies This on every edge); the system time is not
 * advanced. FP64 only (SBP_ERR_INVALID in the mixed form), and every block
 * must step with the system's dt (SBP_ERR_INVALID when a block has a step
 * scale other than 1, see sbp_system_set_dt_scale). */
int sbp_system_advance_adaptive(sbp_system sys, double dt, double* guess_dt, double maxerr);
/* integrate::integrate::<B> (utils/integrate/src/lib.rs:17-24, :176-222) of
 * nsteps steps of size dt for any explicit Butcher tableau given as host
 * arrays: nstages s in [1, 9]; A: s x s row-major, stage i > 0 takes
 * y_n + Σ_{j<i} (A[i*s + j]·dt)·k_j (the reference's A[i-1][j]); B: s
 * weights; C: s - 1 nodes, stage i > 0 evaluated at t + dt·C[i-1] (the
 * reference's indexing; NULL when s = 1). Zero coefficients are skipped and
 * every term is added with two roundings, as the reference's scaled_add.
 * Each stage is one RHS launch per operator group into a work set (s + 1
 * sets of 4 planes per block, allocated on first use); the system time
 * advances by dt per step. FP64 only (SBP_ERR_INVALID in the mixed form). */
int sbp_system_integrate(sbp_system sys, int nstages, const double* A, const double* B, const double* C, double dt,
                         int64_t nsteps);
/* Sticky device error word of the last advance (0 = none); cleared by reading. */
int sbp_system_check(sbp_system sys);

/* ---- Expressions (initial and boundary conditions) ----------------------
 * The evalexpr expressions of a multigrid configuration
 * (multigrid/src/eval/evalexpr.rs; the evalexpr crate 7.2.0): `globals`
 * (statements evaluated once, may be NULL or empty) and four expressions of
 * t, x, y: rho, rhou, rhov, e (conservation form; rhou and rhov may read rho,
 * e may read rho, rhou, rhov) or rho, u, v, p (pressure form; u and v may
 * read rho, p may read rho, u, v; e = p/(GAMMA-1) + rho(u²+v²)/2, as
 * euler/src/eval.rs's EvaluatorPressureWrapper). GAMMA, PI and E are
 * predefined; integers stay integers (5/2 = 2) and `^` is a float power, as
 * in evalexpr. Compiled into a program that runs per point on the device.
 * SBP_ERR_EXPR with a message in errbuf when an expression does not
 * compile (a syntax error, an unknown name or function, a variable read
 * before its assignment); SBP_ERR_GAMMA when GAMMA is not set. */
int sbp_expr_create(int form, const char* globals, const char* const* exprs4, sbp_expr* out, char* errbuf, int errlen);
int sbp_expr_destroy(sbp_expr expr);
/* The four conservative variables at n host points (x[i], y[i]) at time t
 * into out (4 x n, component-major), on the host; SBP_ERR_EXPR when an
 * evaluation fails (the reference panics). */
int sbp_expr_eval_host(sbp_expr expr, double t, int64_t n, const double* x, const double* y, double* out);
/* BoundaryCharacteristic::Eval (euler/src/lib.rs:741, :853-866): the edge's
 * boundary data are the expressions evaluated at every stage time on the
 * edge's coordinates (multigrid/src/parsing.rs:179-184 picks it for edges
 * without an explicit condition when the configuration's boundary
 * conditions are expressions). The system keeps its own copy of the
 * program. */
int sbp_system_set_bc_eval(sbp_system sys, int block, int dir, sbp_expr expr);
/* Initial conditions from expressions on one block at time t, on the device
 * (BaseSystem::create, multigrid/src/system.rs:143-152). */
int sbp_system_eval_block(sbp_system sys, int block, double t, sbp_expr expr);
/* System::error for expression initial conditions (multigrid/src/system.rs:
 * 618-635): the sum over blocks of Field::h2_err against the expressions at
 * the system time. */
int sbp_system_error_eval(sbp_system sys, sbp_expr expr, double* out);

/* ---- Measurement ---------------------------------------------------------
 * Times nsteps RK4 steps (the state is advanced) with HIP events on the
 * context stream into *ms_total; then runs min(nsteps, 8) more steps with
 * every stage (the launches of its edge and interior tiles) bracketed by its
 * own events and returns the mean stage duration in *ms_stage_avg. */
int sbp_system_profile(sbp_system sys, double dt, int64_t nsteps, double* ms_total, double* ms_stage_avg);
/* sbp_system_advance(sys, dt, nsteps) with HIP events recorded on the
 * context stream around the steps; *ms = their elapsed time. A graph capture
 * the steps need happens before the first event. */
int sbp_system_advance_timed(sbp_system sys, double dt, int64_t nsteps, double* ms);
/* Tile shape currently used for block 0 (after autotuning). */
int sbp_system_tile_shape(sbp_system sys, int* tx, int* ty, int* ry);
/* Index into the library's shape list of the tile shape used for block 0
 * (after autotuning); shapes of one (TX, TY, RY) differ in their variant. */
int sbp_system_tile_index(sbp_system sys, int* index);
/* Force a tile shape (index into the library's shape list) or -1 = autotune. */
int sbp_system_set_tile_shape(sbp_system sys, int shape_index);
int sbp_tile_shape_count(void);
int sbp_tile_shape_get(int index, int* tx, int* ty, int* ry);
/* Stage schedule (no counterpart in the reference, whose blocks run one per
 * thread, multigrid/src/system.rs:160-366): 0 = split (the tiles on block
 * edges, which evaluate their edge points' SATs themselves, on a second
 * stream, concurrent with the interior tiles), 1 = merged (one dispatch over
 * all tiles, edge tiles first, SATs evaluated in the edge tiles), 2 = merged
 * with the SATs of every block edge evaluated first by a separate kernel,
 * -1 = chosen by the autotuner at the tile shape (the default). The results
 * do not depend on it. */
int sbp_system_set_schedule(sbp_system sys, int schedule);
/* Schedule used for the first operator group (after autotuning). */
int sbp_system_schedule(sbp_system sys, int* schedule);
/* Tile order of the dispatches (no counterpart in the reference; tuned with
 * the schedule): 0 = each XCD takes a contiguous band of tiles, 1 = row-major
 * order, -1 = chosen by the autotuner (the default). The results do not
 * depend on it. */
int sbp_system_set_tile_order(sbp_system sys, int order);
/* Tile order of the first operator group's dispatches (after autotuning). */
int sbp_system_tile_order(sbp_system sys, int* order);

#ifdef __cplusplus
}
#endif

#endif /* SBPGPU_H */
