/*
 * libpdfp_hip — MI355X (gfx950) PDHG projector and root-LP solver for the
 * PDFP feasibility pump.
 *
 * The library replaces the device side of the reference's two PDHG solvers:
 *   - the projector  PDFPState / pdfp_solve!      (src/pdfp.jl:15-288,
 *                                                   src/pdfp_kernels.jl:1-86)
 *   - the root LP    PDLPState / pdlp_solve!      (src/pdlp.jl:33-372,
 *                                                   src/pdlp_kernels.jl:1-63)
 *   - the ‖A‖₂ power iteration                    (src/preconditioner.jl:108-125)
 * In the reference these are selected by Julia dispatch on CuVector states;
 * here they sit behind a C ABI that a Julia `ccall` (or Python ctypes) binds.
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero pdfp_status
 *     otherwise; the message is in pdfp_last_error(ctx) (or
 *     pdfp_last_error(NULL) when no context exists yet).
 *   - Input arrays are borrowed for the duration of the call and copied into
 *     HBM; output arrays are caller-allocated.
 *   - The matrix comes in compressed-sparse-column form, exactly as Julia's
 *     SparseMatrixCSC stores it (colptr / rowval / nzval with Int64 indices);
 *     `index_base` is 1 for Julia, 0 for C/Python.  The library builds both
 *     CSR(A) and CSR(Aᵀ) in device memory (the reference builds Aᵀ on the host
 *     with sparse(transpose(A)), src/pdfp.jl:68).
 *   - Every call blocks until its results are on the host, matching the
 *     reference's synchronous semantics.  One context belongs to one host
 *     thread; different contexts are independent (own stream, own buffers).
 */
#ifndef PDFP_HIP_H
#define PDFP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status and termination codes ------------------------------------- */

enum pdfp_status {
  PDFP_OK = 0,
  PDFP_ERR_ARGUMENT = 1,  /* bad size, null pointer, malformed CSC          */
  PDFP_ERR_DEVICE = 2,    /* a HIP runtime call failed                      */
  PDFP_ERR_STATE = 3,     /* call out of order (e.g. solve before target)   */
  PDFP_ERR_NOMEM = 4
};

/* Termination reasons: the reference's Symbols (src/pdfp.jl:273-287,
 * src/pdlp.jl:349-371) as integers. */
enum pdfp_termination {
  PDFP_UNSPECIFIED = 0,
  PDFP_OPTIMAL = 1,
  PDFP_ITERATION_LIMIT = 2,
  PDFP_TIME_LIMIT = 3,
  PDFP_PRIMAL_INFEASIBLE = 4,
  PDFP_DUAL_INFEASIBLE = 5
};

/* Arithmetic of the iterates. */
enum pdfp_precision {
  /* FP64 everywhere; the reference's launch structure: SpMV Aᵀz, primal
   * update, y update, SpMV Ax̄, z update — five launches per iteration. */
  PDFP_PREC_FP64 = 0,
  /* FP32 CSR values and iterates, FP64 accumulation inside each row, two
   * fused launches per iteration, FP64 restart/KKT reductions and an FP64
   * polish of the returned point. */
  PDFP_PREC_FP32_MIXED = 1
};

/* Mirrors SolverParameters + RestartParameters (src/pdlp.jl:10-31); the
 * defaults (pdfp_default_params) are the reference's. */
typedef struct pdfp_params {
  double eps_optimal_relative;              /* 1e-4      */
  double eps_feasible_relative;             /* 1e-4      */
  int64_t iteration_limit;                  /* 100000    */
  double time_limit;                        /* 3600 s    */
  int32_t termination_evaluation_frequency; /* 200       */
  int32_t sv_max_iterations;                /* 5000      */
  double reflection_coefficient;            /* 1.0       */
  double sv_tolerance;                      /* 1e-4      */
  double artificial_restart_threshold;      /* 0.36      */
  double sufficient_reduction_for_restart;  /* 0.2       */
  double necessary_reduction_for_restart;   /* 0.5       */
  double k_p;                               /* 0.99      */
  double k_i;                               /* 0.01      */
  double k_d;                               /* 0.0       */
  double i_smooth;                          /* 0.3       */
} pdfp_params;

/* Scalars of the last solve (fields of PDFPState, src/pdfp.jl:31-51). */
typedef struct pdfp_info {
  double relative_primal_residual;
  double relative_dual_residual;
  double relative_l1_gap;
  double l1_distance;        /* Σ_{j∈I} w_j |x̂_j - x̃_j| of the returned point  */
  double fixed_point_error;
  double initial_fixed_point_error;
  double primal_weight;
  double step_size;
  double max_singular_value;
  double solve_seconds;      /* wall time of the last pdfp_solve                */
  int64_t total_count;       /* PDHG iterations of the last solve               */
  int32_t termination_reason;
  int32_t n_restarts;
  int32_t precision;
  int32_t n_evaluations;
  int32_t npb_a;             /* nonzeros per workgroup chosen by the autotuner */
  int32_t npb_at;            /*   for the A pass and for the Aᵀ pass            */
  int32_t persist_a;         /* 0: one workgroup per row block; k: a persistent */
  int32_t persist_at;        /*   grid of k workgroups per CU (autotuned)       */
  int32_t small_path;        /* 1: the last solve ran each block of iterations  */
                             /*   in one workgroup (operators held in LDS)      */
  int32_t lane_rows_a;       /* batch: blocks whose rows hold at most this many */
  int32_t lane_rows_at;      /*   nonzeros are reduced one lane per (row,       */
                             /*   trajectory); 0: none (autotuned)              */
  int32_t block_graph;       /* 1: the last solve replayed each block of the    */
                             /*   general passes as a captured HIP graph        */
  int32_t column_slices;     /* batch: 0, or the nonzeros per workgroup (256,   */
                             /*   512) of the A pass over 8 column slices of A, */
                             /*   one per XCD                                   */
  int32_t slice_min_len;     /* batch, with column slices: the rows longer than */
                             /*   this run over the slices, the others in one   */
                             /*   pass; 0: every row                            */
  int32_t nontemporal;       /* batch: 1 if the passes load their per-row       */
                             /*   operands with non-temporal loads (autotuned)  */
} pdfp_info;

/* Fills `p` with the reference defaults. */
void pdfp_default_params(pdfp_params* p);

/* ---- projector (src/pdfp.jl) ------------------------------------------ */

typedef struct pdfp_ctx pdfp_ctx;

/* Message of the last failure on `ctx`, or of the last failed
 * pdfp_create / pdlp_create when ctx is NULL.  Never NULL. */
const char* pdfp_last_error(const pdfp_ctx* ctx);

/* PDFPState(A, b, is_eq, is_int, var_lb, var_ub, w; use_gpu)
 *   src/pdfp.jl:60-76.  A (m×n, scaled) in CSC; b, is_eq of length m;
 *   is_int, var_lb, var_ub, w of length n (±Inf allowed in the bounds).
 *   Iterates start at zero, the target at zero. */
int pdfp_create(pdfp_ctx** out, int64_t m, int64_t n, int64_t nnz,
                const int64_t* colptr, const int64_t* rowval,
                const double* nzval, int32_t index_base, const double* b,
                const uint8_t* is_eq, const uint8_t* is_int,
                const double* var_lb, const double* var_ub, const double* w,
                int32_t precision, int32_t device);

/* original_primal_residual!(s, D_r)  src/pdfp.jl:85-90. */
int pdfp_set_original_primal_residual(pdfp_ctx* ctx, const double* D_r);

/* copyto!(current_x, x); copyto!(initial_x, x)  src/pump.jl:305-306. */
int pdfp_set_x(pdfp_ctx* ctx, const double* x_scaled);

/* The same for the duals (current and initial y, length n, and z, length
 * m): the reference's pump never sets them (no counterpart); the FP64 polish
 * of the FP32 form starts an FP64 context at an FP32 context's point. */
int pdfp_set_dual(pdfp_ctx* ctx, const double* y, const double* z);

/* set_target!(s, x̃)  src/pdfp.jl:98-101 (copies x̃, then warm_start!). */
int pdfp_set_target(pdfp_ctx* ctx, const double* x_tilde_scaled);

/* The pump's set_target!(projector, x̃ ./ D_c)  src/pump.jl:331 with the
 * division on the device: x̃ in original units, divided (IEEE, as the host
 * would) by the column scaling given once to pdfp_set_column_scale (n
 * positive finite values; PDFP_ERR_STATE before it).  Equal, bit for bit, to
 * pdfp_set_target on the host's x̃ ./ D_c. */
int pdfp_set_column_scale(pdfp_ctx* ctx, const double* D_c);
int pdfp_set_target_original(pdfp_ctx* ctx, const double* x_tilde);

/* warm_start!(s)  src/pdfp.jl:104-114. */
int pdfp_warm_start(pdfp_ctx* ctx);

/* pdfp_solve!(s, p, primal_weight)  src/pdfp.jl:252-288.  `params` may be
 * NULL for the defaults.  termination / total_count may be NULL. */
int pdfp_solve(pdfp_ctx* ctx, const pdfp_params* params, double primal_weight,
               int32_t* termination, int64_t* total_count);

/* Array(s.pdhg_x)  src/pump.jl:338 — the projection in scaled units, as
 * FP64.  In PDFP_PREC_FP32_MIXED the point is polished: a coordinate that
 * the last primal step clamped to a bound takes that bound's FP64 value. */
int pdfp_get_x(pdfp_ctx* ctx, double* x_scaled_out);

/* Other iterates, as FP64 (for checks and warm restarts by the caller).
 * which: 0 pdhg_x, 1 pdhg_y, 2 pdhg_z, 3 current_x, 4 current_y,
 *        5 current_z, 6 dual_slack, 7 x_tilde */
int pdfp_get_vector(pdfp_ctx* ctx, int32_t which, double* out);

/* Scalars of the last solve. */
int pdfp_get_info(pdfp_ctx* ctx, pdfp_info* out);

/* Frees every device buffer of the context. NULL is a no-op. */
void pdfp_destroy(pdfp_ctx* ctx);

/* ---- batch of trajectories on one instance ----------------------------- */
/* The reference runs K concurrent pump trajectories as K OS processes, each
 * with its own PDFPState on the same GPU (scripts/solve.jl:121-148).  A
 * pdfp_batch holds one instance's scaled matrix once and up to 16
 * trajectories (own target, iterates, primal weight, restart state) that
 * iterate together in the same fused launches, in the FP32 form; each
 * trajectory follows pdfp_solve! (src/pdfp.jl:252-288) on its own.  Up to 8
 * trajectories are stored 8 to a column (32 bytes), 9 to 16 sixteen to a
 * column (64 bytes). */

typedef struct pdfp_batch pdfp_batch;

const char* pdfp_batch_last_error(const pdfp_batch* ctx);

/* The continuous form of pdfp_batch_solve (no counterpart in the reference;
 * the batched pump uses it): pdfp_batch_begin starts trajectory k's solve
 * with its own parameters and primal weight (after pdfp_batch_set_target or
 * pdfp_batch_warm_start), as pdfp_solve! does on entry
 * (src/pdfp.jl:220-241).  pdfp_batch_run runs every begun, unfinished
 * trajectory in lockstep blocks of termination_evaluation_frequency
 * iterations (which, with the reflection coefficient, the running
 * trajectories must share), each to its own termination, and returns after
 * the first block in which one ends (until_any = 1), after one block
 * (until_any = 2: a caller with trajectories waiting to start polls between
 * blocks) or when none runs (until_any = 0): termination[K] (0 for those
 * still running),
 * total_count[K], and the number still running.  A trajectory that has
 * ended can be given a new target and begun again while the others run on. */
int pdfp_batch_begin(pdfp_batch* ctx, int32_t k, const pdfp_params* params,
                     double primal_weight);
int pdfp_batch_run(pdfp_batch* ctx, int32_t until_any, int32_t* termination,
                   int64_t* total_count, int32_t* n_running);

/* As pdfp_create, for K ∈ 1..16 trajectories (FP32 form). */
int pdfp_batch_create(pdfp_batch** out, int32_t K, int64_t m, int64_t n,
                      int64_t nnz, const int64_t* colptr, const int64_t* rowval,
                      const double* nzval, int32_t index_base, const double* b,
                      const uint8_t* is_eq, const uint8_t* is_int,
                      const double* var_lb, const double* var_ub,
                      const double* w, int32_t device);
/* original_primal_residual! for every trajectory (src/pdfp.jl:85-90). */
int pdfp_batch_set_original_primal_residual(pdfp_batch* ctx, const double* D_r);
/* Trajectory k's copyto!(current_x/initial_x) (src/pump.jl:305-306),
 * set_target! (src/pdfp.jl:98-101) and warm_start! (src/pdfp.jl:104-114). */
int pdfp_batch_set_x(pdfp_batch* ctx, int32_t k, const double* x_scaled);
int pdfp_batch_set_target(pdfp_batch* ctx, int32_t k, const double* x_tilde_scaled);
/* Trajectory k's set_target!(projector, x̃ ./ D_c)  src/pump.jl:331 with the
 * division on the device, as pdfp_set_target_original. */
int pdfp_batch_set_column_scale(pdfp_batch* ctx, const double* D_c);
int pdfp_batch_set_target_original(pdfp_batch* ctx, int32_t k, const double* x_tilde);
int pdfp_batch_warm_start(pdfp_batch* ctx, int32_t k);
/* pdfp_solve! for every k with active[k] (all when active is NULL), with
 * primal_weight[k]; returns when each of them has terminated.  An inactive
 * trajectory is left unchanged.  termination / total_count: K entries. */
int pdfp_batch_solve(pdfp_batch* ctx, const pdfp_params* params,
                     const double* primal_weight, const uint8_t* active,
                     int32_t* termination, int64_t* total_count);
/* Trajectory k's projection (FP64, polished), vectors (ids as in
 * pdfp_get_vector) and scalars. */
int pdfp_batch_get_x(pdfp_batch* ctx, int32_t k, double* x_scaled_out);
int pdfp_batch_get_vector(pdfp_batch* ctx, int32_t k, int32_t which, double* out);
int pdfp_batch_get_info(pdfp_batch* ctx, int32_t k, pdfp_info* out);
/* Mean milliseconds of one batched iteration (all K trajectories). */
int pdfp_batch_time_iterations(pdfp_batch* ctx, int32_t iterations, double* ms_out);
/* Mean milliseconds of the batched transpose pass and of the batched A pass. */
int pdfp_batch_time_kernels(pdfp_batch* ctx, int32_t iterations, double* ms_at,
                            double* ms_a);
/* Replaces the autotuned schedule of the batched passes (no counterpart in
 * the reference; for tests and tuning): nonzeros per workgroup of the A and
 * Aᵀ passes (256 or 512), persistent workgroups per CU (0: one workgroup per
 * row block), and the row-length bound of the lane-per-(row, trajectory)
 * reduction of each pass (0: off). */
int pdfp_batch_set_schedule(pdfp_batch* ctx, int32_t npb_a, int32_t npb_at, int32_t persist,
                            int32_t lane_rows_a, int32_t lane_rows_at);

/* Block graphs of the batch (no counterpart in the reference), as
 * pdfp_set_block_graph: 1 (the default) replays each block of iterations,
 * with its evaluations and readback, as one captured HIP graph; 0 launches
 * every kernel on the stream.  Both give identical results. */
int pdfp_batch_set_block_graph(pdfp_batch* ctx, int32_t on);
/* Cooperative blocks of the batch (no counterpart in the reference; it
 * replaces the 2 launches per PDHG step of the reference's 200-step block
 * loop, src/pdfp.jl:259-267): wg_per_cu != 0 runs iterations 2 .. freq-1 of
 * each block of the batched passes in one cooperative launch of wg_per_cu
 * workgroups per CU (-1: one fewer than fit), all resident, with a grid barrier
 * between the passes; 0 (the default) launches each pass.  Both give
 * identical results.  PDFP_ERR_ARGUMENT for wg_per_cu outside -1 .. 16 or
 * above what fits; PDFP_ERR_DEVICE if the device has no cooperative launch,
 * or if a barrier wait gives up (the call's iterates are then undefined).
 * No effect on instances that run the one-workgroup path.  On MI355X the
 * grid barriers cost more than the kernel boundaries they replace (75 vs
 * 24 µs per iteration at 1M nonzeros), hence the default. */
int pdfp_batch_set_cooperative(pdfp_batch* ctx, int32_t wg_per_cu);
/* Column slices of the batch's A pass (no counterpart in the reference; it
 * replaces the CSC product A x̄ of src/pdfp_kernels.jl:81 when x̄ outgrows an
 * XCD's L2): 1 (or 256, 512: the nonzeros per workgroup of the slices'
 * pass; 1 keeps the last) runs A x̄ as 8 partial products over A's 8 column slices,
 * each walked by one XCD so that its gathers of x̄ come from that XCD's L2,
 * then adds the 8 partials (FP32) per row in FP64 and updates z; 0 runs A x̄
 * in one pass.  pdfp_batch_create's autotuner turns it on where an
 * iteration is faster with it (x̄ larger than 4 MB and an eighth of it at
 * most 5 MB); pdfp_info's column_slices says which ran.  The partial sums
 * are rounded to FP32, so the iterates differ from the one-pass form's in
 * the last bits. */
int pdfp_batch_set_column_slices(pdfp_batch* ctx, int32_t on);
/* The rows of A that run over the column slices (no counterpart in the
 * reference): those with more than min_len nonzeros; the others run in one
 * pass over A with the long rows emptied, whose epilogue adds each long
 * row's 8 partials.  0 runs every row over the slices.  On power-law rows a
 * few long rows hold most nonzeros: their gathers come from the XCDs' L2s,
 * while the many short rows need no partials.  Takes effect when column
 * slices are on; pdfp_batch_create's autotuner picks the bound with the
 * slices (0, or 32, 64, 128 or 256 where the longer rows hold 20-95% of the
 * nonzeros).  PDFP_ERR_ARGUMENT for min_len < 0. */
int pdfp_batch_set_slice_min_len(pdfp_batch* ctx, int32_t min_len);
/* How the batched passes load their per-row and per-column operands (no
 * counterpart in the reference): 1 with non-temporal loads, which keep them
 * from displacing the gathered vectors' lines in L2 (faster where the
 * iterates outgrow the memory-side cache), 0 with plain loads.  The results
 * are the same; pdfp_batch_create's autotuner picks (pdfp_info's
 * nontemporal).  PDFP_ERR_ARGUMENT unless 0 or 1. */
int pdfp_batch_set_nontemporal(pdfp_batch* ctx, int32_t on);
/* Kernel path of the batch (no counterpart in the reference), as
 * pdfp_set_kernel_path: 0 (the default) runs each block as one launch with
 * one workgroup per running trajectory when the instance fits in LDS (each
 * iterates exactly as the single FP32 projector's one-workgroup path, its
 * σ_max included), else the batched passes; 1 always runs the batched
 * passes.  pdfp_info's small_path says which ran. */
int pdfp_batch_set_kernel_path(pdfp_batch* ctx, int32_t path);
void pdfp_batch_destroy(pdfp_batch* ctx);

/* ---- pack of independent instances ------------------------------------ */
/* The reference solves independent instances in separate processes
 * (scripts/solve.jl:121-148).  A pdfp_pack holds K projector instances of
 * uneven sizes as one block-diagonal A (and Aᵀ) in HBM and iterates them
 * together in the FP32 form's fused launches; each instance has its own
 * target, iterates, step size (its own power iteration), primal weight,
 * restart state and termination, and follows pdfp_solve!
 * (src/pdfp.jl:252-288) on its own.  Instance k is given as pdfp_create's
 * arguments (its CSC and vectors: element k of each pointer array). */

typedef struct pdfp_pack pdfp_pack;

const char* pdfp_pack_last_error(const pdfp_pack* ctx);

/* As K calls of pdfp_create (FP32 form), into one pack. */
int pdfp_pack_create(pdfp_pack** out, int32_t K, const int64_t* m, const int64_t* n,
                     const int64_t* const* colptr, const int64_t* const* rowval,
                     const double* const* nzval, int32_t index_base, const double* const* b,
                     const uint8_t* const* is_eq, const uint8_t* const* is_int,
                     const double* const* var_lb, const double* const* var_ub,
                     const double* const* w, int32_t device);
/* Instance k's original_primal_residual! (src/pdfp.jl:85-90), copyto! of
 * current_x/initial_x (src/pump.jl:305-306), set_target! (src/pdfp.jl:98-101)
 * and warm_start! (src/pdfp.jl:104-114). */
int pdfp_pack_set_original_primal_residual(pdfp_pack* ctx, int32_t k, const double* D_r);
int pdfp_pack_set_x(pdfp_pack* ctx, int32_t k, const double* x_scaled);
int pdfp_pack_set_target(pdfp_pack* ctx, int32_t k, const double* x_tilde_scaled);
int pdfp_pack_warm_start(pdfp_pack* ctx, int32_t k);
/* pdfp_solve! for every k with active[k] (all when active is NULL), with
 * primal_weight[k]; returns when each of them has terminated.  The running
 * instances share the evaluation frequency and the reflection coefficient
 * (one params for all).  termination / total_count: K entries. */
int pdfp_pack_solve(pdfp_pack* ctx, const pdfp_params* params, const double* primal_weight,
                    const uint8_t* active, int32_t* termination, int64_t* total_count);
/* The continuous form (no counterpart in the reference; a pump over the
 * pack's instances uses it), as pdfp_batch_begin / pdfp_batch_run: instance
 * k starts a solve with its own parameters and primal weight; run iterates
 * every begun, unfinished instance until one ends (until_any = 1), for one
 * block (until_any = 2) or until none runs (until_any = 0), and reports termination[K] (0 while running), total_count[K] and
 * the number still running. */
int pdfp_pack_begin(pdfp_pack* ctx, int32_t k, const pdfp_params* params,
                    double primal_weight);
int pdfp_pack_run(pdfp_pack* ctx, int32_t until_any, int32_t* termination,
                  int64_t* total_count, int32_t* n_running);
/* Instance k's projection (FP64, polished), vectors (ids as in
 * pdfp_get_vector) and scalars. */
int pdfp_pack_get_x(pdfp_pack* ctx, int32_t k, double* x_scaled_out);
int pdfp_pack_get_vector(pdfp_pack* ctx, int32_t k, int32_t which, double* out);
int pdfp_pack_get_info(pdfp_pack* ctx, int32_t k, pdfp_info* out);
/* Block graphs of the pack (no counterpart in the reference), as
 * pdfp_set_block_graph: 1 (the default) replays each block of iterations,
 * with its evaluations and readback, as one captured HIP graph; 0 launches
 * every kernel on the stream.  Both give identical results. */
int pdfp_pack_set_block_graph(pdfp_pack* ctx, int32_t on);
/* Kernel path of the pack (no counterpart in the reference), as
 * pdfp_set_kernel_path: 0 (the default) runs each block as one launch with
 * one workgroup per running instance when every instance fits in LDS (each
 * iterates exactly as the single FP32 projector's one-workgroup path, its
 * σ_max included), else the general passes; 1 always runs the general
 * passes.  pdfp_info's small_path says which ran. */
int pdfp_pack_set_kernel_path(pdfp_pack* ctx, int32_t path);
/* Mean milliseconds of one pack iteration (every instance), and of its
 * transpose pass and A pass. */
int pdfp_pack_time_iterations(pdfp_pack* ctx, int32_t iterations, double* ms_out);
int pdfp_pack_time_kernels(pdfp_pack* ctx, int32_t iterations, double* ms_at, double* ms_a);
void pdfp_pack_destroy(pdfp_pack* ctx);

/* ---- root LP (src/pdlp.jl) -------------------------------------------- */

typedef struct pdlp_ctx pdlp_ctx;

/* Message of the last failure on `ctx` (NULL: as pdfp_last_error(NULL)). */
const char* pdlp_last_error(const pdlp_ctx* ctx);

typedef struct pdlp_info {
  double relative_primal_residual;
  double relative_dual_residual;
  double relative_optimality_gap;
  double objective_value;
  double fixed_point_error;
  double primal_weight;
  double step_size;
  double solve_seconds;
  int64_t total_count;
  int32_t termination_reason;
  int32_t n_restarts;
  int32_t small_path;        /* 1: the last solve ran each block of iterations  */
                             /*   in one workgroup (operators held in LDS)      */
} pdlp_info;

/* PDLPState(sp; use_gpu)  src/pdlp.jl:82-96 on an already scaled problem
 * (c, var_lb/ub, con_lb/ub of the ScaledProblem, src/preconditioner.jl:13-24). */
/* Kernel path of the root LP (no counterpart in the reference): 0 (the
 * default) runs each block of iterations in one workgroup when the instance
 * fits in LDS (24 nnz + 52 n + 36 m + 8 bytes ≤ 152 KiB on gfx950: the device's
 * per-workgroup LDS less 8 KiB), else the general
 * passes; 1 always runs the general passes. */
int pdlp_set_kernel_path(pdlp_ctx* ctx, int32_t path);

int pdlp_create(pdlp_ctx** out, int64_t m, int64_t n, int64_t nnz,
                const int64_t* colptr, const int64_t* rowval,
                const double* nzval, int32_t index_base, const double* c,
                const double* var_lb, const double* var_ub,
                const double* con_lb, const double* con_ub,
                double bound_scale, double objective_scale, int32_t device);

/* cupdlpx_termination!(s, D_r, D_c)  src/pdlp.jl:105-110. */
int pdlp_set_cupdlpx_termination(pdlp_ctx* ctx, const double* D_r,
                                 const double* D_c);

/* pdlp_solve!(s, p)  src/pdlp.jl:325-372. */
int pdlp_solve(pdlp_ctx* ctx, const pdfp_params* params, int32_t* termination,
               int64_t* total_count);

/* Array(s.pdhg_primal), Array(s.dual_slack) in scaled units
 * (src/pdlp.jl:391-392). */
int pdlp_get_primal(pdlp_ctx* ctx, double* out);
int pdlp_get_dual_slack(pdlp_ctx* ctx, double* out);
int pdlp_get_info(pdlp_ctx* ctx, pdlp_info* out);
void pdlp_destroy(pdlp_ctx* ctx);

/* ---- building blocks (exposed for measurement and tests) --------------- */

/* ‖A‖₂ by power iteration on AAᵀ on the device
 * (estimate_max_singular_value, src/preconditioner.jl:108-125), from the
 * library's fixed pseudo-random start (pdfp_start_vector).  Uses the
 * projector context's matrix. */
int pdfp_max_singular_value(pdfp_ctx* ctx, int32_t max_iterations,
                            double tolerance, double* sigma_out);

/* The start vector of the power iteration: m standard normals from a
 * SplitMix64 stream seeded with `seed` (Box–Muller). Host only. */
void pdfp_start_vector(int64_t m, uint64_t seed, double* out);

/* y = A x (transpose=0) or y = Aᵀ x (transpose=1) on the context's scaled
 * matrix, in the context's precision, x and y on the host as FP64; used by
 * the SpMV parity tests. */
int pdfp_spmv(pdfp_ctx* ctx, int32_t transpose, const double* x, double* y);

/* Runs `iterations` PDHG iterations with the current scalars and no
 * evaluation, timing them on the device with events on the context's
 * stream; returns the mean milliseconds per iteration.  For measurement. */
int pdfp_time_iterations(pdfp_ctx* ctx, int32_t iterations, double* ms_out);

/* Per-kernel mean duration (ms) of the kernels that pdfp_time_iterations
 * launched, measured with events around each launch on the context's
 * stream: which 0 = transpose (Aᵀz) pass, 1 = A (Ax̄) pass. */
int pdfp_time_kernels(pdfp_ctx* ctx, int32_t iterations, double* ms_at,
                      double* ms_a);

/* No counterpart in the reference (measurement): mean duration in ms of a
 * plain SpMV y = A x (transpose = 1: y = Aᵀ x) over the context's resident
 * CSR operator, in its precision, each launch timed by its own events — the
 * sparse product that cuSPARSE performs at src/pdfp_kernels.jl:26 and 81,
 * without the fused epilogue. */
int pdfp_time_spmv(pdfp_ctx* ctx, int32_t transpose, int32_t iterations, double* ms_out);

/* Device memory held by the context, in bytes. */
int64_t pdfp_device_bytes(const pdfp_ctx* ctx);

/* Kernel path (no counterpart in the reference).  0 (the default): the FP32
 * form runs each block of iterations in one workgroup when the instance fits
 * in LDS (4(m+n+2) + 16 nnz + 36 n + 13 m bytes ≤ 152 KiB on gfx950: the
 * device's per-workgroup LDS less 8 KiB), else the general passes; the FP64
 * form keeps the reference's five launches per iteration.  1: the general
 * passes in either form.  2: one workgroup per block in either form when the
 * instance fits (FP64: 4(m+n+2) + 24 nnz + 72 n + 25 m bytes), as for the
 * FP64 finish of FP32 projections on small instances; the power iteration
 * then runs in one workgroup too.  Set before the first solve: σ_max is
 * computed once. */
int pdfp_set_kernel_path(pdfp_ctx* ctx, int32_t path);

/* Block graphs of the FP32 form (no counterpart in the reference): 1 (the
 * default) captures each block of `termination_evaluation_frequency`
 * iterations of the general passes, with its evaluation and the readback of
 * its sums, as one HIP graph on first use and replays it for every later
 * block of the same shape; 0 launches every kernel on the stream.  Both give
 * identical results.  The one-workgroup path and the FP64 form ignore it. */
int pdfp_set_block_graph(pdfp_ctx* ctx, int32_t on);

/* Replaces the autotuned schedule of the two passes (no counterpart in the
 * reference; for tests and tuning): nonzeros per workgroup of the A and Aᵀ
 * passes (256, 512, 1024 or 2048) and persistent workgroups per CU of each
 * (0: one workgroup per row block; at most 16).  The schedule sets the order
 * of each row's sum, so two states give bit-identical iterates only on the
 * same schedule. */
int pdfp_set_schedule(pdfp_ctx* ctx, int32_t npb_a, int32_t npb_at, int32_t persist_a,
                      int32_t persist_at);

/* ---- the pump's rounding (host) -------------------------------------------
 * Constraint-propagation rounding (FP 2.0's CPR) and the guarded two-phase
 * CPR (src/rounding.jl:100-256) over the bound propagation of
 * src/propagation.jl, in C++ on the host with the reference's arithmetic.
 * The caller ranks the integer variables (src/rounding.jl:105-115, the only
 * random draws) and keeps the batch policy's timing (src/rounding.jl:22-63).
 * A rounder holds PropagationData for one model (src/propagation.jl:19-84). */
typedef struct pdfp_rounder pdfp_rounder;

/* Rounder(milp) (src/rounding.jl:78-98): the model's CSC (m rows), right-hand
 * sides, is_le (1: ≤ row, 0: = row), variable bounds (±Inf allowed), the
 * integer mask and the propagation budget per call (max_pops). */
int pdfp_rounder_create(pdfp_rounder** out, int64_t m, int64_t n, const int64_t* colptr,
                        const int64_t* rowval, const double* nzval, int32_t index_base,
                        const double* b, const uint8_t* is_le, const double* var_lb,
                        const double* var_ub, const uint8_t* is_int, int64_t max_pops);
/* rounder.direction: n flags of the violation-guided columns, or NULL (off). */
int pdfp_rounder_set_direction(pdfp_rounder* r, const uint8_t* direction);
/* round_sequential (src/rounding.jl:191-197): CPR of x_star[n] over the k
 * ranked columns; writes xt[n]. */
int pdfp_round_sequential(pdfp_rounder* r, const double* x_star, const int64_t* ranked,
                          int64_t k, double* xt);
/* round_guarded's batch and rounding (src/rounding.jl:205-256) with schedule
 * regroup (1) or all (0).  *batch_infeasible = 1 when the guard discards the
 * batch: the caller then rounds sequentially (the reference's fallback). */
int pdfp_round_guarded(pdfp_rounder* r, const double* x_star, const int64_t* ranked, int64_t k,
                       int32_t regroup, double* xt, int32_t* batch_infeasible);
const char* pdfp_rounder_last_error(const pdfp_rounder* r);
void pdfp_rounder_destroy(pdfp_rounder* r);

/* ---- the pump's feasibility check on the device ---------------------------
 * max_violation(milp, x) (src/milp.jl:102-114) on the original A kept
 * resident in HBM (FP64 CSR): each call uploads x[n] and returns the largest
 * violation of A x ≤ b (is_le = 1) / A x = b (is_le = 0), bit-identical to
 * the reference's (each row summed in the CSC product's order). */
typedef struct pdfp_checker pdfp_checker;
int pdfp_checker_create(pdfp_checker** out, int64_t m, int64_t n, const int64_t* colptr,
                        const int64_t* rowval, const double* nzval, int32_t index_base,
                        const double* b, const uint8_t* is_le, int32_t device);
int pdfp_checker_max_violation(pdfp_checker* c, const double* x, double* max_violation);
const char* pdfp_checker_last_error(const pdfp_checker* c);
void pdfp_checker_destroy(pdfp_checker* c);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): max_violation (src/milp.jl:102-114) of a device-resident point
 * on a device-resident CSR matrix, with the kernels of
 * pdfp_checker_max_violation, allocating nothing.  All pointers are device
 * pointers: rowptr (m + 1, int32, 0-based), col (int32) and val (FP64) of A
 * in CSR with ascending columns per row, b and is_le (1 = ≤ row, 0 = equality)
 * of length m, x of length n, and out (one double) that receives the result.
 * The launches go on `stream` (a hipStream_t; NULL = the null stream) and the
 * call returns without waiting for them.  The per-workgroup partial maxima
 * live in a static device array of 65,536 doubles (512 KiB) on each GPU, so
 * m may not exceed 16,777,216 rows (PDFP_ERR_ARGUMENT beyond; the message is
 * at pdfp_checker_last_error(NULL)).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads) race. */
int pdfp_max_violation_noalloc(int64_t m, int64_t n, const int32_t* rowptr, const int32_t* col,
                               const double* val, const double* b, const uint8_t* is_le,
                               const double* x, double* out, void* stream);

/* ---- preconditioning ------------------------------------------------------
 * scale_problem (src/preconditioner.jl:26-63), the matrix part, on the
 * device: `ruiz_iterations` Ruiz rounds and one Pock–Chambolle pass
 * (pock_chambolle_alpha must be 1, the reference's value).  Input: the CSC of
 * the m×n matrix (index_base 0 or 1).  Output: the scaled values in the same
 * CSC order (scaled_nzval[nnz]) and the scalings D_r[m], D_c[n], bit-identical
 * to the reference's CPU loop.  The caller forms c·D_c, ℓ/D_c, u/D_c and
 * D_r·(row bounds) (src/preconditioner.jl:65-99).  Error text:
 * pdfp_scale_last_error(). */
int pdfp_scale_matrix(int64_t m, int64_t n, int64_t nnz, const int64_t* colptr,
                      const int64_t* rowval, const double* nzval, int32_t index_base,
                      int32_t ruiz_iterations, double pock_chambolle_alpha,
                      double* scaled_nzval, double* D_r, double* D_c, int32_t device);
const char* pdfp_scale_last_error(void);


#ifdef __cplusplus
}
#endif

#endif /* PDFP_HIP_H */
