/*
 * ORACLE — test infrastructure only.  A plain-C FP64 restatement of the
 * reference's generic CPU code path for the PDHG projector and the root LP.
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load it, and only as the checker; the product (libpdfp_hip) never
 * links, loads or calls it.
 *
 * Followed line by line:
 *   scale_problem                    src/preconditioner.jl:26-99
 *   estimate_max_singular_value      src/preconditioner.jl:108-125
 *   PDFPState and its solver         src/pdfp.jl:15-288
 *   PDLPState and its solver         src/pdlp.jl:33-372
 *   restart rule, PID primal weight  src/pdlp.jl:167-205
 *
 * Julia semantics kept:
 *   - SparseArrays `mul!(y, A, x)` with A in CSC: y = 0, then for each column
 *     j in order, y[rowval[p]] += nzval[p] * x[j]  (no FMA contraction; the
 *     file is compiled with -ffp-contract=off).
 *   - clamp(x, lo, hi) = x > hi ? hi : (x < lo ? lo : x).
 *   - max/min propagate NaN (not needed on finite data; kept as written).
 *   - `norm`, `sum`, `dot`: sequential sums (the reference's BLAS/pairwise
 *     orders differ in the last ulp only).
 * Not reproducible without Julia: randn(MersenneTwister(20240517)) at
 * src/preconditioner.jl:111.  The start vector here is the library's
 * documented SplitMix64/Box–Muller stream (orc_start_vector); ‖A‖₂ converges
 * to the same value from any start with a component on the top singular
 * vector.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define EXPORT __attribute__((visibility("default")))

/* ---- parameters (src/pdlp.jl:10-31) ------------------------------------ */

typedef struct orc_params {
  double eps_optimal_relative;
  double eps_feasible_relative;
  int64_t iteration_limit;
  double time_limit;
  int32_t termination_evaluation_frequency;
  int32_t sv_max_iterations;
  double reflection_coefficient;
  double sv_tolerance;
  double artificial_restart_threshold;
  double sufficient_reduction_for_restart;
  double necessary_reduction_for_restart;
  double k_p, k_i, k_d, i_smooth;
} orc_params;

enum { T_UNSPECIFIED = 0, T_OPTIMAL, T_ITERATION_LIMIT, T_TIME_LIMIT,
       T_PRIMAL_INFEASIBLE, T_DUAL_INFEASIBLE };

EXPORT void orc_default_params(orc_params* p) {
  p->eps_optimal_relative = 1e-4;
  p->eps_feasible_relative = 1e-4;
  p->iteration_limit = 100000;
  p->time_limit = 3600.0;
  p->termination_evaluation_frequency = 200;
  p->sv_max_iterations = 5000;
  p->reflection_coefficient = 1.0;
  p->sv_tolerance = 1e-4;
  p->artificial_restart_threshold = 0.36;
  p->sufficient_reduction_for_restart = 0.2;
  p->necessary_reduction_for_restart = 0.5;
  p->k_p = 0.99;
  p->k_i = 0.01;
  p->k_d = 0.0;
  p->i_smooth = 0.3;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static double jclamp(double x, double lo, double hi) {
  return x > hi ? hi : (x < lo ? lo : x);
}
static double jmax(double a, double b) {  /* Julia max: NaN-propagating */
  if (isnan(a) || isnan(b)) return NAN;
  return a > b ? a : b;
}
static double jmin(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  return a < b ? a : b;
}

/* ---- sparse matrix (CSC, 0-based) ------------------------------------- */

typedef struct csc {
  int64_t m, n, nnz;
  int64_t* colptr; /* n+1 */
  int64_t* rowval; /* nnz */
  double* nzval;   /* nnz */
  int sorted;      /* row indices strictly ascend within every column */
} csc;

static int csc_rows_sorted(const csc* A) {
  for (int64_t j = 0; j < A->n; j++)
    for (int64_t p = A->colptr[j] + 1; p < A->colptr[j + 1]; p++)
      if (A->rowval[p] <= A->rowval[p - 1]) return 0;
  return 1;
}

static csc csc_copy(int64_t m, int64_t n, const int64_t* colptr,
                    const int64_t* rowval, const double* nzval) {
  csc A;
  A.m = m; A.n = n; A.nnz = colptr[n];
  A.colptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  A.rowval = (int64_t*)malloc(sizeof(int64_t) * (size_t)(A.nnz > 0 ? A.nnz : 1));
  A.nzval = (double*)malloc(sizeof(double) * (size_t)(A.nnz > 0 ? A.nnz : 1));
  memcpy(A.colptr, colptr, sizeof(int64_t) * (size_t)(n + 1));
  if (A.nnz) {
    memcpy(A.rowval, rowval, sizeof(int64_t) * (size_t)A.nnz);
    memcpy(A.nzval, nzval, sizeof(double) * (size_t)A.nnz);
  }
  A.sorted = csc_rows_sorted(&A);
  return A;
}

/* sparse(transpose(A)): CSC of Aᵀ with sorted row indices (a counting sort
 * over the rows of A, visiting the columns of A in order). */
static csc csc_transpose(const csc* A) {
  csc T;
  T.m = A->n; T.n = A->m; T.nnz = A->nnz;
  T.colptr = (int64_t*)calloc((size_t)(A->m + 1), sizeof(int64_t));
  T.rowval = (int64_t*)malloc(sizeof(int64_t) * (size_t)(A->nnz > 0 ? A->nnz : 1));
  T.nzval = (double*)malloc(sizeof(double) * (size_t)(A->nnz > 0 ? A->nnz : 1));
  for (int64_t p = 0; p < A->nnz; p++) T.colptr[A->rowval[p] + 1]++;
  for (int64_t i = 0; i < A->m; i++) T.colptr[i + 1] += T.colptr[i];
  int64_t* next = (int64_t*)malloc(sizeof(int64_t) * (size_t)(A->m + 1));
  memcpy(next, T.colptr, sizeof(int64_t) * (size_t)(A->m + 1));
  for (int64_t j = 0; j < A->n; j++)
    for (int64_t p = A->colptr[j]; p < A->colptr[j + 1]; p++) {
      int64_t i = A->rowval[p];
      int64_t q = next[i]++;
      T.rowval[q] = j;
      T.nzval[q] = A->nzval[p];
    }
  free(next);
  T.sorted = csc_rows_sorted(&T);
  return T;
}

static void csc_free(csc* A) {
  free(A->colptr); free(A->rowval); free(A->nzval);
  A->colptr = NULL; A->rowval = NULL; A->nzval = NULL;
}

/* mul!(y, A, x): y (length A.m) = A x, column-major scatter. */
static void spmv(const csc* A, const double* x, double* y) {
  for (int64_t i = 0; i < A->m; i++) y[i] = 0.0;
  for (int64_t j = 0; j < A->n; j++) {
    double xj = x[j];
    for (int64_t p = A->colptr[j]; p < A->colptr[j + 1]; p++)
      y[A->rowval[p]] += A->nzval[p] * xj;
  }
}

/* Host threads for the projector's step (orc_set_threads; 1, the
 * reference's single-threaded CPU path, by default): its two products and
 * its three elementwise updates, each element computed by one thread as the
 * serial loop computes it (the reductions stay serial), so the iterates are
 * bit-identical.  A timing option with no counterpart in the reference. */
static int g_threads = 1;
EXPORT void orc_set_threads(int32_t n) { g_threads = n < 1 ? 1 : n; }

/* The same y = A x with g_threads > 1: each y[i] is summed by one thread
 * from AT = Aᵀ's column i (A's row i, its entries in ascending column order
 * when A's row indices ascend within each column, as csc_transpose and
 * every canonical CSC give them), starting from 0.0 — the order in which
 * the scatter adds them, so the result is bit-identical. */
static void spmv_par(const csc* A, const csc* AT, const double* x, double* y) {
  if (g_threads <= 1 || !AT->sorted) {
    spmv(A, x, y);
    return;
  }
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1024)
  for (int64_t i = 0; i < A->m; i++) {
    double acc = 0.0;
    for (int64_t p = AT->colptr[i]; p < AT->colptr[i + 1]; p++)
      acc += AT->nzval[p] * x[AT->rowval[p]];
    y[i] = acc;
  }
}

static double sumsq(const double* v, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += v[i] * v[i];
  return s;
}
static double dotp(const double* a, const double* b, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static double* zvec(int64_t n) {
  return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
}

/* ---- scale_problem (src/preconditioner.jl:26-99) ------------------------ */

#define SCALING_EPSILON 1e-12

/* In place on nzval (the CSC's values), c, var_lb, var_ub, con_lb, con_ub. */
EXPORT void orc_scale_problem(int64_t m, int64_t n, const int64_t* colptr,
                              const int64_t* rowval, double* vals, double* c,
                              double* var_lb, double* var_ub, double* con_lb,
                              double* con_ub, int32_t ruiz_iterations,
                              double alpha, int32_t bound_objective_rescaling,
                              double* D_r, double* D_c, double* bound_scale_out,
                              double* objective_scale_out) {
  double* row_max = zvec(m);
  double* col_max = zvec(n);
  double* rscale = zvec(m);
  double* cscale = zvec(n);
  for (int64_t i = 0; i < m; i++) D_r[i] = 1.0;
  for (int64_t j = 0; j < n; j++) D_c[j] = 1.0;

  for (int32_t it = 0; it < ruiz_iterations; it++) {
    for (int64_t i = 0; i < m; i++) row_max[i] = 0.0;
    for (int64_t j = 0; j < n; j++) col_max[j] = 0.0;
    for (int64_t j = 0; j < n; j++)
      for (int64_t p = colptr[j]; p < colptr[j + 1]; p++) {
        double v = fabs(vals[p]);
        int64_t i = rowval[p];
        if (v > row_max[i]) row_max[i] = v;
        if (v > col_max[j]) col_max[j] = v;
      }
    for (int64_t i = 0; i < m; i++)
      rscale[i] = row_max[i] < SCALING_EPSILON ? 1.0 : 1.0 / sqrt(row_max[i]);
    for (int64_t j = 0; j < n; j++)
      cscale[j] = col_max[j] < SCALING_EPSILON ? 1.0 : 1.0 / sqrt(col_max[j]);
    for (int64_t j = 0; j < n; j++)
      for (int64_t p = colptr[j]; p < colptr[j + 1]; p++)
        vals[p] *= rscale[rowval[p]] * cscale[j];
    for (int64_t i = 0; i < m; i++) D_r[i] *= rscale[i];
    for (int64_t j = 0; j < n; j++) D_c[j] *= cscale[j];
  }

  /* Pock–Chambolle: α on the rows, 2 - α on the columns. */
  double* row_sum = row_max;
  double* col_sum = col_max;
  for (int64_t i = 0; i < m; i++) row_sum[i] = 0.0;
  for (int64_t j = 0; j < n; j++) col_sum[j] = 0.0;
  for (int64_t j = 0; j < n; j++)
    for (int64_t p = colptr[j]; p < colptr[j + 1]; p++) {
      double v = fabs(vals[p]);
      int64_t i = rowval[p];
      row_sum[i] += pow(v, alpha);
      col_sum[j] += pow(v, 2.0 - alpha);
    }
  for (int64_t i = 0; i < m; i++)
    rscale[i] = row_sum[i] < SCALING_EPSILON ? 1.0 : 1.0 / sqrt(row_sum[i]);
  for (int64_t j = 0; j < n; j++)
    cscale[j] = col_sum[j] < SCALING_EPSILON ? 1.0 : 1.0 / sqrt(col_sum[j]);
  for (int64_t j = 0; j < n; j++)
    for (int64_t p = colptr[j]; p < colptr[j + 1]; p++)
      vals[p] *= rscale[rowval[p]] * cscale[j];
  for (int64_t i = 0; i < m; i++) D_r[i] *= rscale[i];
  for (int64_t j = 0; j < n; j++) D_c[j] *= cscale[j];

  /* Keep the original constraint bounds: the rescaling reads them. */
  double* con_lb0 = zvec(m);
  double* con_ub0 = zvec(m);
  memcpy(con_lb0, con_lb, sizeof(double) * (size_t)m);
  memcpy(con_ub0, con_ub, sizeof(double) * (size_t)m);
  double* var_lb0 = zvec(n);
  double* var_ub0 = zvec(n);
  memcpy(var_lb0, var_lb, sizeof(double) * (size_t)n);
  memcpy(var_ub0, var_ub, sizeof(double) * (size_t)n);

  for (int64_t j = 0; j < n; j++) {
    c[j] = c[j] * D_c[j];
    if (isfinite(var_lb0[j])) var_lb[j] = var_lb0[j] / D_c[j];
    if (isfinite(var_ub0[j])) var_ub[j] = var_ub0[j] / D_c[j];
  }
  for (int64_t i = 0; i < m; i++) {
    if (isfinite(con_lb0[i])) con_lb[i] = con_lb0[i] * D_r[i];
    if (isfinite(con_ub0[i])) con_ub[i] = con_ub0[i] * D_r[i];
  }

  double bound_scale = 1.0, objective_scale = 1.0;
  if (bound_objective_rescaling) {
    double bound_sq = 0.0;
    for (int64_t i = 0; i < m; i++) {
      int finite_lo = isfinite(con_lb0[i]), finite_hi = isfinite(con_ub0[i]);
      double lo = con_lb[i], hi = con_ub[i];
      if (finite_lo && (!finite_hi || fabs(lo - hi) > SCALING_EPSILON))
        bound_sq += lo * lo;
      if (finite_hi) bound_sq += hi * hi;
    }
    bound_scale = 1.0 / (sqrt(bound_sq) + 1.0);
    objective_scale = 1.0 / (sqrt(sumsq(c, n)) + 1.0);
    for (int64_t i = 0; i < m; i++) {
      if (isfinite(con_lb0[i])) con_lb[i] *= bound_scale;
      if (isfinite(con_ub0[i])) con_ub[i] *= bound_scale;
    }
    for (int64_t j = 0; j < n; j++) {
      if (isfinite(var_lb0[j])) var_lb[j] *= bound_scale;
      if (isfinite(var_ub0[j])) var_ub[j] *= bound_scale;
      c[j] *= objective_scale;
    }
  }
  *bound_scale_out = bound_scale;
  *objective_scale_out = objective_scale;
  free(row_max); free(col_max); free(rscale); free(cscale);
  free(con_lb0); free(con_ub0); free(var_lb0); free(var_ub0);
}

/* ---- start vector and power iteration (src/preconditioner.jl:108-125) -- */

static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* m standard normals: Box–Muller on pairs of 53-bit uniforms in (0, 1]. */
EXPORT void orc_start_vector(int64_t m, uint64_t seed, double* out) {
  uint64_t s = seed;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < m; i += 2) {
    double u1 = ((double)(splitmix64(&s) >> 11) + 1.0) * 0x1.0p-53;
    double u2 = (double)(splitmix64(&s) >> 11) * 0x1.0p-53;
    double r = sqrt(-2.0 * log(u1));
    out[i] = r * cos(two_pi * u2);
    if (i + 1 < m) out[i + 1] = r * sin(two_pi * u2);
  }
}

#define SV_SEED 20240517ull

static double max_singular_value(const csc* A, const csc* At,
                                 int32_t max_iterations, double tolerance) {
  int64_t m = A->m, n = A->n;
  double* eig = zvec(m);
  double* nxt = zvec(m);
  double* dual = zvec(n);
  orc_start_vector(m, SV_SEED, eig);
  double sigma_sq = 1.0;
  for (int32_t it = 0; it < max_iterations; it++) {
    memcpy(nxt, eig, sizeof(double) * (size_t)m);
    double nrm = sqrt(sumsq(nxt, m));
    if (nrm == 0.0) break;
    for (int64_t i = 0; i < m; i++) nxt[i] /= nrm;
    spmv(At, nxt, dual);
    spmv(A, dual, eig);
    sigma_sq = dotp(nxt, eig, m);
    for (int64_t i = 0; i < m; i++) nxt[i] = eig[i] - sigma_sq * nxt[i];
    if (sqrt(sumsq(nxt, m)) < tolerance) break;
  }
  free(eig); free(nxt); free(dual);
  return sqrt(jmax(sigma_sq, 0.0));
}

EXPORT double orc_max_singular_value(int64_t m, int64_t n,
                                     const int64_t* colptr,
                                     const int64_t* rowval,
                                     const double* nzval,
                                     int32_t max_iterations, double tolerance) {
  csc A = csc_copy(m, n, colptr, rowval, nzval);
  csc At = csc_transpose(&A);
  double s = max_singular_value(&A, &At, max_iterations, tolerance);
  csc_free(&A); csc_free(&At);
  return s;
}

/* mul!(y, A, x) exported, for the SpMV checks. transpose=1: y = Aᵀ x. */
EXPORT void orc_spmv(int64_t m, int64_t n, const int64_t* colptr,
                     const int64_t* rowval, const double* nzval,
                     int32_t transpose, const double* x, double* y) {
  csc A = csc_copy(m, n, colptr, rowval, nzval);
  if (transpose) {
    csc At = csc_transpose(&A);
    spmv(&At, x, y);
    csc_free(&At);
  } else {
    spmv(&A, x, y);
  }
  csc_free(&A);
}

/* ---- shared restart logic (src/pdlp.jl:167-205) ------------------------ */

typedef struct restart_state {
  double primal_weight, best_primal_weight, best_primal_dual_residual_gap;
  double primal_weight_error_sum, primal_weight_last_error;
  double fixed_point_error, initial_fixed_point_error,
      last_trial_fixed_point_error;
  int64_t inner_count, total_count;
  double relative_primal_residual, relative_dual_residual;
} restart_state;

static int should_restart(restart_state* s, const orc_params* p) {
  int restart = 0;
  if (s->total_count == p->termination_evaluation_frequency) {
    restart = 1;
  } else if (s->total_count > p->termination_evaluation_frequency) {
    if (s->fixed_point_error <=
        p->sufficient_reduction_for_restart * s->initial_fixed_point_error)
      restart = 1;
    if (s->fixed_point_error <=
            p->necessary_reduction_for_restart * s->initial_fixed_point_error &&
        s->fixed_point_error > s->last_trial_fixed_point_error)
      restart = 1;
    if ((double)s->inner_count >=
        p->artificial_restart_threshold * (double)s->total_count)
      restart = 1;
  }
  s->last_trial_fixed_point_error = s->fixed_point_error;
  return restart;
}

#define JULIA_EPS 2.220446049250313e-16

static void update_primal_weight(restart_state* s, const orc_params* rp,
                                 double primal_dist, double dual_dist) {
  double ratio =
      s->relative_dual_residual / jmax(s->relative_primal_residual, JULIA_EPS);
  if (1e-16 < primal_dist && primal_dist < 1e12 && 1e-16 < dual_dist &&
      dual_dist < 1e12 && 1e-8 < ratio && ratio < 1e8) {
    double err = log(dual_dist) - log(primal_dist) - log(s->primal_weight);
    s->primal_weight_error_sum *= rp->i_smooth;
    s->primal_weight_error_sum += err;
    double delta_err = err - s->primal_weight_last_error;
    s->primal_weight *= exp(rp->k_p * err + rp->k_i * s->primal_weight_error_sum +
                            rp->k_d * delta_err);
    s->primal_weight_last_error = err;
  } else {
    s->primal_weight = s->best_primal_weight;
    s->primal_weight_error_sum = 0.0;
    s->primal_weight_last_error = 0.0;
  }
  double gap = fabs(log10(s->relative_dual_residual /
                          jmax(s->relative_primal_residual, JULIA_EPS)));
  if (gap < s->best_primal_dual_residual_gap) {
    s->best_primal_dual_residual_gap = gap;
    s->best_primal_weight = s->primal_weight;
  }
}

/* ---- PDFP projector (src/pdfp.jl) -------------------------------------- */

typedef struct orc_pdfp {
  csc A, At;
  int64_t m, n;
  double *b, *var_lb, *var_ub, *x_tilde, *w;
  uint8_t *is_eq, *is_int;
  double *initial_x, *current_x, *pdhg_x, *reflected_x;
  double *initial_y, *current_y, *pdhg_y, *reflected_y;
  double *initial_z, *current_z, *pdhg_z, *reflected_z;
  double *Ax, *Atz, *tmp_n, *tmp_m, *dual_slack, *delta_x, *delta_y, *delta_z;
  int original_primal_residual;
  double* row_unscale;
  double b_norm_original;
  double max_singular_value, step_size;
  restart_state r;
  double relative_l1_gap, dual_residual_denominator, l1_gap_denominator;
  int32_t termination_reason;
  int32_t n_restarts;
  double solve_seconds;
} orc_pdfp;

typedef struct orc_pdfp_info {
  double relative_primal_residual, relative_dual_residual, relative_l1_gap;
  double l1_distance, fixed_point_error, initial_fixed_point_error;
  double primal_weight, step_size, max_singular_value, solve_seconds;
  int64_t total_count;
  int32_t termination_reason, n_restarts, precision, n_evaluations;
} orc_pdfp_info;

/* PDFPState(A, b, is_eq, is_int, var_lb, var_ub, w)  src/pdfp.jl:60-76. */
EXPORT orc_pdfp* orc_pdfp_create(int64_t m, int64_t n, const int64_t* colptr,
                                 const int64_t* rowval, const double* nzval,
                                 const double* b, const uint8_t* is_eq,
                                 const uint8_t* is_int, const double* var_lb,
                                 const double* var_ub, const double* w) {
  orc_pdfp* s = (orc_pdfp*)calloc(1, sizeof(orc_pdfp));
  s->m = m; s->n = n;
  s->A = csc_copy(m, n, colptr, rowval, nzval);
  s->At = csc_transpose(&s->A);
  s->b = zvec(m); memcpy(s->b, b, sizeof(double) * (size_t)m);
  s->is_eq = (uint8_t*)calloc((size_t)(m > 0 ? m : 1), 1);
  memcpy(s->is_eq, is_eq, (size_t)m);
  s->is_int = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
  memcpy(s->is_int, is_int, (size_t)n);
  s->var_lb = zvec(n); memcpy(s->var_lb, var_lb, sizeof(double) * (size_t)n);
  s->var_ub = zvec(n); memcpy(s->var_ub, var_ub, sizeof(double) * (size_t)n);
  s->w = zvec(n); memcpy(s->w, w, sizeof(double) * (size_t)n);
  s->x_tilde = zvec(n);
  s->initial_x = zvec(n); s->current_x = zvec(n); s->pdhg_x = zvec(n); s->reflected_x = zvec(n);
  s->initial_y = zvec(n); s->current_y = zvec(n); s->pdhg_y = zvec(n); s->reflected_y = zvec(n);
  s->initial_z = zvec(m); s->current_z = zvec(m); s->pdhg_z = zvec(m); s->reflected_z = zvec(m);
  s->Ax = zvec(m); s->Atz = zvec(n); s->tmp_n = zvec(n); s->tmp_m = zvec(m);
  s->dual_slack = zvec(n); s->delta_x = zvec(n); s->delta_y = zvec(n); s->delta_z = zvec(m);
  s->b_norm_original = NAN;
  s->max_singular_value = NAN;
  s->r.primal_weight = 1.0;
  s->r.best_primal_weight = 1.0;
  s->r.best_primal_dual_residual_gap = INFINITY;
  s->r.fixed_point_error = INFINITY;
  s->r.initial_fixed_point_error = INFINITY;
  s->r.last_trial_fixed_point_error = INFINITY;
  s->r.relative_primal_residual = INFINITY;
  s->r.relative_dual_residual = INFINITY;
  s->relative_l1_gap = INFINITY;
  s->dual_residual_denominator = 1.0;
  s->l1_gap_denominator = 1.0;
  return s;
}

EXPORT void orc_pdfp_destroy(orc_pdfp* s) {
  if (!s) return;
  csc_free(&s->A); csc_free(&s->At);
  double* vs[] = {s->b, s->var_lb, s->var_ub, s->x_tilde, s->w,
                  s->initial_x, s->current_x, s->pdhg_x, s->reflected_x,
                  s->initial_y, s->current_y, s->pdhg_y, s->reflected_y,
                  s->initial_z, s->current_z, s->pdhg_z, s->reflected_z,
                  s->Ax, s->Atz, s->tmp_n, s->tmp_m, s->dual_slack,
                  s->delta_x, s->delta_y, s->delta_z, s->row_unscale};
  for (size_t k = 0; k < sizeof(vs) / sizeof(vs[0]); k++) free(vs[k]);
  free(s->is_eq); free(s->is_int);
  free(s);
}

/* original_primal_residual!(s, D_r)  src/pdfp.jl:85-90. */
EXPORT void orc_pdfp_original_primal_residual(orc_pdfp* s, const double* D_r) {
  if (!s->row_unscale) s->row_unscale = zvec(s->m);
  for (int64_t i = 0; i < s->m; i++) s->row_unscale[i] = 1.0 / D_r[i];
  double acc = 0.0;
  for (int64_t i = 0; i < s->m; i++) {
    double v = s->b[i] / D_r[i];
    acc += v * v;
  }
  s->b_norm_original = sqrt(acc);
  s->original_primal_residual = 1;
}

/* copyto!(current_x, x); copyto!(initial_x, x)  src/pump.jl:305-306. */
EXPORT void orc_pdfp_set_x(orc_pdfp* s, const double* x) {
  memcpy(s->current_x, x, sizeof(double) * (size_t)s->n);
  memcpy(s->initial_x, x, sizeof(double) * (size_t)s->n);
}

/* The same for y and z (no counterpart in the reference; the host tests of
 * the FP64 polish start a state at another state's point). */
EXPORT void orc_pdfp_set_dual(orc_pdfp* s, const double* y, const double* z) {
  memcpy(s->current_y, y, sizeof(double) * (size_t)s->n);
  memcpy(s->initial_y, y, sizeof(double) * (size_t)s->n);
  memcpy(s->current_z, z, sizeof(double) * (size_t)s->m);
  memcpy(s->initial_z, z, sizeof(double) * (size_t)s->m);
}

/* warm_start!(s)  src/pdfp.jl:104-114. */
EXPORT void orc_pdfp_warm_start(orc_pdfp* s) {
  memcpy(s->initial_x, s->current_x, sizeof(double) * (size_t)s->n);
  memcpy(s->initial_y, s->current_y, sizeof(double) * (size_t)s->n);
  memcpy(s->initial_z, s->current_z, sizeof(double) * (size_t)s->m);
  s->r.inner_count = 0;
  s->r.total_count = 0;
  s->r.fixed_point_error = INFINITY;
  s->r.last_trial_fixed_point_error = INFINITY;
  s->termination_reason = T_UNSPECIFIED;
}

/* set_target!(s, x̃)  src/pdfp.jl:98-101. */
EXPORT void orc_pdfp_set_target(orc_pdfp* s, const double* xt) {
  memcpy(s->x_tilde, xt, sizeof(double) * (size_t)s->n);
  orc_pdfp_warm_start(s);
}

/* primal_step!  src/pdfp.jl:119-131. */
static void pdfp_primal_step(orc_pdfp* s, int64_t k_offset, double omega) {
  double tau = s->step_size / s->r.primal_weight;
  int64_t k = s->r.inner_count + k_offset;
  double weight = (double)k / ((double)k + 1.0), omw = 1.0 - omega,
         omwt = 1.0 - weight;
  spmv_par(&s->At, &s->A, s->current_z, s->Atz);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (int64_t j = 0; j < s->n; j++) {
    double temp = s->current_x[j] - tau * (s->Atz[j] - s->current_y[j]);
    double px = jclamp(temp, s->var_lb[j], s->var_ub[j]);
    s->pdhg_x[j] = px;
    s->dual_slack[j] = (px - temp) / tau;
    s->reflected_x[j] = 2.0 * px - s->current_x[j];
    s->current_x[j] = weight * (omega * s->reflected_x[j] + omw * s->current_x[j]) +
                      omwt * s->initial_x[j];
  }
}

/* dual_y_step!  src/pdfp.jl:134-143. */
static void pdfp_dual_y_step(orc_pdfp* s, int64_t k_offset, double omega) {
  double sigma = s->step_size * s->r.primal_weight;
  int64_t k = s->r.inner_count + k_offset;
  double weight = (double)k / ((double)k + 1.0), omw = 1.0 - omega,
         omwt = 1.0 - weight;
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (int64_t j = 0; j < s->n; j++) {
    double y = s->is_int[j]
                   ? jclamp(s->current_y[j] + sigma * (s->x_tilde[j] - s->reflected_x[j]),
                            -s->w[j], s->w[j])
                   : 0.0;
    s->pdhg_y[j] = y;
    s->reflected_y[j] = 2.0 * y - s->current_y[j];
    s->current_y[j] = weight * (omega * s->reflected_y[j] + omw * s->current_y[j]) +
                      omwt * s->initial_y[j];
  }
}

/* dual_z_step!  src/pdfp.jl:146-156. */
static void pdfp_dual_z_step(orc_pdfp* s, int64_t k_offset, double omega) {
  double sigma = s->step_size * s->r.primal_weight;
  int64_t k = s->r.inner_count + k_offset;
  double weight = (double)k / ((double)k + 1.0), omw = 1.0 - omega,
         omwt = 1.0 - weight;
  spmv_par(&s->A, &s->At, s->reflected_x, s->Ax);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (int64_t i = 0; i < s->m; i++) {
    double raw = s->current_z[i] + sigma * (s->Ax[i] - s->b[i]);
    double z = s->is_eq[i] ? raw : jmax(raw, 0.0);
    s->pdhg_z[i] = z;
    s->reflected_z[i] = 2.0 * z - s->current_z[i];
    s->current_z[i] = weight * (omega * s->reflected_z[i] + omw * s->current_z[i]) +
                      omwt * s->initial_z[i];
  }
}

static void pdfp_step(orc_pdfp* s, int64_t k_offset, double omega) {
  pdfp_primal_step(s, k_offset, omega);
  pdfp_dual_y_step(s, k_offset, omega);
  pdfp_dual_z_step(s, k_offset, omega);
}

/* compute_fixed_point_error!  src/pdfp.jl:167-178. */
static double pdfp_fixed_point_error(orc_pdfp* s) {
  int64_t n = s->n, m = s->m;
  for (int64_t j = 0; j < n; j++) s->delta_x[j] = s->pdhg_x[j] - s->reflected_x[j];
  for (int64_t j = 0; j < n; j++) s->delta_y[j] = s->pdhg_y[j] - s->reflected_y[j];
  for (int64_t i = 0; i < m; i++) s->delta_z[i] = s->pdhg_z[i] - s->reflected_z[i];
  double movement = sumsq(s->delta_x, n) * s->r.primal_weight +
                    (sumsq(s->delta_y, n) + sumsq(s->delta_z, m)) / s->r.primal_weight;
  spmv_par(&s->At, &s->A, s->delta_z, s->Atz);
  for (int64_t j = 0; j < n; j++) s->tmp_n[j] = s->Atz[j] - s->delta_y[j];
  double interaction = 2.0 * s->step_size * dotp(s->tmp_n, s->delta_x, n);
  s->r.fixed_point_error = sqrt(jmax(movement + interaction, 0.0));
  return s->r.fixed_point_error;
}

/* perform_restart!  src/pdfp.jl:180-192. */
static void pdfp_perform_restart(orc_pdfp* s, const orc_params* p) {
  int64_t n = s->n, m = s->m;
  for (int64_t j = 0; j < n; j++) s->delta_x[j] = s->initial_x[j] - s->pdhg_x[j];
  for (int64_t j = 0; j < n; j++) s->delta_y[j] = s->initial_y[j] - s->pdhg_y[j];
  for (int64_t i = 0; i < m; i++) s->delta_z[i] = s->initial_z[i] - s->pdhg_z[i];
  update_primal_weight(&s->r, p, sqrt(sumsq(s->delta_x, n)),
                       sqrt(sumsq(s->delta_y, n) + sumsq(s->delta_z, m)));
  memcpy(s->initial_x, s->pdhg_x, sizeof(double) * (size_t)n);
  memcpy(s->current_x, s->pdhg_x, sizeof(double) * (size_t)n);
  memcpy(s->initial_y, s->pdhg_y, sizeof(double) * (size_t)n);
  memcpy(s->current_y, s->pdhg_y, sizeof(double) * (size_t)n);
  memcpy(s->initial_z, s->pdhg_z, sizeof(double) * (size_t)m);
  memcpy(s->current_z, s->pdhg_z, sizeof(double) * (size_t)m);
  s->r.inner_count = 0;
  s->r.last_trial_fixed_point_error = INFINITY;
  s->n_restarts++;
}

/* compute_residual! on arbitrary (x, y, z, dual_slack)  src/pdfp.jl:200-218.
 * out[0..2] = rel primal, rel dual, rel ℓ1 gap; out[3] = ℓ1 distance. */
static void pdfp_residual_of(orc_pdfp* s, const double* x, const double* y,
                             const double* z, const double* ds, double* out) {
  int64_t n = s->n, m = s->m;
  spmv_par(&s->A, &s->At, x, s->Ax);
  for (int64_t i = 0; i < m; i++) {
    double d = s->Ax[i] - s->b[i];
    s->tmp_m[i] = s->is_eq[i] ? fabs(d) : jmax(d, 0.0);
  }
  if (s->original_primal_residual)
    for (int64_t i = 0; i < m; i++) s->tmp_m[i] *= s->row_unscale[i];
  double b_norm = s->original_primal_residual ? s->b_norm_original
                                              : sqrt(sumsq(s->b, m));
  out[0] = sqrt(sumsq(s->tmp_m, m)) / (1.0 + b_norm);
  spmv_par(&s->At, &s->A, z, s->Atz);
  for (int64_t j = 0; j < n; j++) s->tmp_n[j] = s->Atz[j] - y[j] - ds[j];
  out[1] = sqrt(sumsq(s->tmp_n, n)) / s->dual_residual_denominator;
  double l1 = 0.0, y_term = 0.0, z_term = 0.0;
  for (int64_t j = 0; j < n; j++)
    l1 += s->is_int[j] ? s->w[j] * fabs(x[j] - s->x_tilde[j]) : 0.0;
  for (int64_t j = 0; j < n; j++)
    y_term += s->is_int[j] ? y[j] * (s->x_tilde[j] - x[j]) : 0.0;
  for (int64_t i = 0; i < m; i++) z_term += z[i] * (s->Ax[i] - s->b[i]);
  out[2] = fabs(y_term + z_term - l1) / s->l1_gap_denominator;
  out[3] = l1;
}

static void pdfp_compute_residual(orc_pdfp* s) {
  double out[4];
  pdfp_residual_of(s, s->pdhg_x, s->pdhg_y, s->pdhg_z, s->dual_slack, out);
  s->r.relative_primal_residual = out[0];
  s->r.relative_dual_residual = out[1];
  s->relative_l1_gap = out[2];
}

/* initialize_step_size_and_primal_weight!  src/pdfp.jl:220-241. */
static void pdfp_initialize(orc_pdfp* s, const orc_params* p, double pw) {
  if (s->A.nnz == 0) {
    s->step_size = 1.0;
  } else {
    if (isnan(s->max_singular_value))
      s->max_singular_value = max_singular_value(&s->A, &s->At,
                                                 p->sv_max_iterations,
                                                 p->sv_tolerance);
    s->step_size = 0.998 / sqrt(1.0 + s->max_singular_value * s->max_singular_value);
  }
  s->r.primal_weight = pw;
  s->r.best_primal_weight = pw;
  s->r.primal_weight_error_sum = 0.0;
  s->r.primal_weight_last_error = 0.0;
  double acc = 0.0;
  for (int64_t j = 0; j < s->n; j++) {
    double v = s->is_int[j] ? s->w[j] : 0.0;
    acc += v * v;
  }
  s->dual_residual_denominator = 1.0 + sqrt(acc);
  acc = 0.0;
  for (int64_t j = 0; j < s->n; j++)
    acc += s->is_int[j] ? s->w[j] * fabs(s->x_tilde[j]) : 0.0;
  s->l1_gap_denominator = 1.0 + acc;
}

/* pdfp_solve!(s, p, primal_weight)  src/pdfp.jl:252-288. */
EXPORT int32_t orc_pdfp_solve(orc_pdfp* s, const orc_params* p,
                              double primal_weight) {
  double t_start = now_s();
  pdfp_initialize(s, p, primal_weight);
  double omega = p->reflection_coefficient;
  int32_t freq = p->termination_evaluation_frequency;
  int do_restart = 0;
  s->n_restarts = 0;
  while (s->r.total_count < p->iteration_limit) {
    pdfp_step(s, 1, omega);
    if (do_restart) {
      pdfp_fixed_point_error(s);
      s->r.initial_fixed_point_error = s->r.fixed_point_error;
      do_restart = 0;
    }
    for (int32_t i = 2; i <= freq; i++) pdfp_step(s, i, omega);
    pdfp_fixed_point_error(s);
    pdfp_compute_residual(s);
    s->r.inner_count += freq;
    s->r.total_count += freq;
    if (s->r.relative_primal_residual <= p->eps_feasible_relative &&
        s->r.relative_dual_residual <= p->eps_optimal_relative &&
        s->relative_l1_gap <= p->eps_optimal_relative) {
      s->termination_reason = T_OPTIMAL;
      break;
    }
    if (now_s() - t_start > p->time_limit) {
      s->termination_reason = T_TIME_LIMIT;
      break;
    }
    do_restart = should_restart(&s->r, p);
    if (do_restart) pdfp_perform_restart(s, p);
  }
  if (s->termination_reason == T_UNSPECIFIED)
    s->termination_reason = T_ITERATION_LIMIT;
  s->solve_seconds = now_s() - t_start;
  return s->termination_reason;
}

/* which: 0 pdhg_x, 1 pdhg_y, 2 pdhg_z, 3 current_x, 4 current_y,
 *        5 current_z, 6 dual_slack, 7 x_tilde (the library's numbering). */
EXPORT void orc_pdfp_get_vector(orc_pdfp* s, int32_t which, double* out) {
  const double* src = NULL;
  int64_t len = s->n;
  switch (which) {
    case 0: src = s->pdhg_x; break;
    case 1: src = s->pdhg_y; break;
    case 2: src = s->pdhg_z; len = s->m; break;
    case 3: src = s->current_x; break;
    case 4: src = s->current_y; break;
    case 5: src = s->current_z; len = s->m; break;
    case 6: src = s->dual_slack; break;
    case 7: src = s->x_tilde; break;
    default: return;
  }
  memcpy(out, src, sizeof(double) * (size_t)len);
}

EXPORT void orc_pdfp_get_info(orc_pdfp* s, orc_pdfp_info* out) {
  memset(out, 0, sizeof(*out));
  out->relative_primal_residual = s->r.relative_primal_residual;
  out->relative_dual_residual = s->r.relative_dual_residual;
  out->relative_l1_gap = s->relative_l1_gap;
  double l1 = 0.0;
  for (int64_t j = 0; j < s->n; j++)
    l1 += s->is_int[j] ? s->w[j] * fabs(s->pdhg_x[j] - s->x_tilde[j]) : 0.0;
  out->l1_distance = l1;
  out->fixed_point_error = s->r.fixed_point_error;
  out->initial_fixed_point_error = s->r.initial_fixed_point_error;
  out->primal_weight = s->r.primal_weight;
  out->step_size = s->step_size;
  out->max_singular_value = s->max_singular_value;
  out->solve_seconds = s->solve_seconds;
  out->total_count = s->r.total_count;
  out->termination_reason = s->termination_reason;
  out->n_restarts = s->n_restarts;
}

/* The reference's compute_residual! evaluated on a caller-given point
 * (x, y, z, dual_slack), e.g. the device's result.  Uses the state's target,
 * weights, denominators (call after a solve so that they are set).
 * out: rel primal, rel dual, rel ℓ1 gap, ℓ1 distance. */
EXPORT void orc_pdfp_residual_of(orc_pdfp* s, const double* x, const double* y,
                                 const double* z, const double* ds,
                                 double* out) {
  pdfp_residual_of(s, x, y, z, ds, out);
}

/* Set the denominators (and the step size) without solving, as
 * initialize_step_size_and_primal_weight! does. */
EXPORT void orc_pdfp_initialize(orc_pdfp* s, const orc_params* p, double pw) {
  pdfp_initialize(s, p, pw);
}

/* ---- PDLP root LP (src/pdlp.jl) ---------------------------------------- */

typedef struct orc_pdlp {
  csc A, At;
  int64_t m, n;
  double *c, *var_lb, *var_ub, *con_lb, *con_ub;
  double *initial_primal, *current_primal, *pdhg_primal, *reflected_primal;
  double *initial_dual, *current_dual, *pdhg_dual, *reflected_dual;
  double *primal_product, *dual_product, *tmp_n, *tmp_m, *dual_slack,
      *delta_primal, *delta_dual;
  double bound_scale, objective_scale;
  int cupdlpx_termination;
  double *row_unscale, *col_unscale;
  double eps_infeasible, constraint_bound_norm, objective_vector_norm;
  double step_size;
  restart_state r;
  double relative_optimality_gap, objective_value;
  double max_primal_ray_infeasibility, primal_ray_linear_objective;
  double max_dual_ray_infeasibility, dual_ray_objective;
  int32_t termination_reason, n_restarts;
  double solve_seconds;
} orc_pdlp;

typedef struct orc_pdlp_info {
  double relative_primal_residual, relative_dual_residual,
      relative_optimality_gap, objective_value, fixed_point_error,
      primal_weight, step_size, solve_seconds;
  int64_t total_count;
  int32_t termination_reason, n_restarts;
} orc_pdlp_info;

/* PDLPState(sp)  src/pdlp.jl:82-96. */
EXPORT orc_pdlp* orc_pdlp_create(int64_t m, int64_t n, const int64_t* colptr,
                                 const int64_t* rowval, const double* nzval,
                                 const double* c, const double* var_lb,
                                 const double* var_ub, const double* con_lb,
                                 const double* con_ub, double bound_scale,
                                 double objective_scale) {
  orc_pdlp* s = (orc_pdlp*)calloc(1, sizeof(orc_pdlp));
  s->m = m; s->n = n;
  s->A = csc_copy(m, n, colptr, rowval, nzval);
  s->At = csc_transpose(&s->A);
  s->c = zvec(n); memcpy(s->c, c, sizeof(double) * (size_t)n);
  s->var_lb = zvec(n); memcpy(s->var_lb, var_lb, sizeof(double) * (size_t)n);
  s->var_ub = zvec(n); memcpy(s->var_ub, var_ub, sizeof(double) * (size_t)n);
  s->con_lb = zvec(m); memcpy(s->con_lb, con_lb, sizeof(double) * (size_t)m);
  s->con_ub = zvec(m); memcpy(s->con_ub, con_ub, sizeof(double) * (size_t)m);
  s->initial_primal = zvec(n); s->current_primal = zvec(n);
  s->pdhg_primal = zvec(n); s->reflected_primal = zvec(n);
  s->initial_dual = zvec(m); s->current_dual = zvec(m);
  s->pdhg_dual = zvec(m); s->reflected_dual = zvec(m);
  s->primal_product = zvec(m); s->dual_product = zvec(n);
  s->tmp_n = zvec(n); s->tmp_m = zvec(m); s->dual_slack = zvec(n);
  s->delta_primal = zvec(n); s->delta_dual = zvec(m);
  s->bound_scale = bound_scale; s->objective_scale = objective_scale;
  s->eps_infeasible = 1e-14;
  s->r.primal_weight = 1.0; s->r.best_primal_weight = 1.0;
  s->r.best_primal_dual_residual_gap = INFINITY;
  s->r.fixed_point_error = INFINITY;
  s->r.initial_fixed_point_error = INFINITY;
  s->r.last_trial_fixed_point_error = INFINITY;
  s->r.relative_primal_residual = INFINITY;
  s->r.relative_dual_residual = INFINITY;
  s->relative_optimality_gap = INFINITY;
  s->objective_value = NAN;
  s->max_primal_ray_infeasibility = NAN;
  s->primal_ray_linear_objective = NAN;
  s->max_dual_ray_infeasibility = NAN;
  s->dual_ray_objective = NAN;
  return s;
}

EXPORT void orc_pdlp_destroy(orc_pdlp* s) {
  if (!s) return;
  csc_free(&s->A); csc_free(&s->At);
  double* vs[] = {s->c, s->var_lb, s->var_ub, s->con_lb, s->con_ub,
                  s->initial_primal, s->current_primal, s->pdhg_primal,
                  s->reflected_primal, s->initial_dual, s->current_dual,
                  s->pdhg_dual, s->reflected_dual, s->primal_product,
                  s->dual_product, s->tmp_n, s->tmp_m, s->dual_slack,
                  s->delta_primal, s->delta_dual, s->row_unscale, s->col_unscale};
  for (size_t k = 0; k < sizeof(vs) / sizeof(vs[0]); k++) free(vs[k]);
  free(s);
}

/* cupdlpx_termination!(s, D_r, D_c)  src/pdlp.jl:105-110. */
EXPORT void orc_pdlp_cupdlpx_termination(orc_pdlp* s, const double* D_r,
                                         const double* D_c) {
  if (!s->row_unscale) s->row_unscale = zvec(s->m);
  if (!s->col_unscale) s->col_unscale = zvec(s->n);
  for (int64_t i = 0; i < s->m; i++) s->row_unscale[i] = 1.0 / D_r[i];
  for (int64_t j = 0; j < s->n; j++) s->col_unscale[j] = 1.0 / D_c[j];
  s->cupdlpx_termination = 1;
}

/* pdhg_step!  src/pdlp.jl:115-153. */
static void pdlp_step(orc_pdlp* s, int64_t k_offset, double omega, int is_major) {
  int64_t n = s->n, m = s->m;
  int64_t k = s->r.inner_count + k_offset;
  double weight = (double)k / ((double)k + 1.0), omw = 1.0 - omega,
         omwt = 1.0 - weight;
  spmv(&s->At, s->current_dual, s->dual_product);
  double step = s->step_size / s->r.primal_weight;
  for (int64_t j = 0; j < n; j++) {
    double temp = s->current_primal[j] - step * (s->c[j] - s->dual_product[j]);
    double px = jclamp(temp, s->var_lb[j], s->var_ub[j]);
    if (is_major) s->dual_slack[j] = (px - temp) / step;
    s->pdhg_primal[j] = px;
    s->reflected_primal[j] = 2.0 * px - s->current_primal[j];
    s->current_primal[j] =
        weight * (omega * s->reflected_primal[j] + omw * s->current_primal[j]) +
        omwt * s->initial_primal[j];
  }
  spmv(&s->A, s->reflected_primal, s->primal_product);
  double dstep = s->step_size * s->r.primal_weight;
  for (int64_t i = 0; i < m; i++) {
    double t = s->current_dual[i] / dstep - s->primal_product[i];
    double d = (t - jclamp(t, -s->con_ub[i], -s->con_lb[i])) * dstep;
    if (is_major) s->pdhg_dual[i] = d;
    s->reflected_dual[i] = 2.0 * d - s->current_dual[i];
    s->current_dual[i] =
        weight * (omega * s->reflected_dual[i] + omw * s->current_dual[i]) +
        omwt * s->initial_dual[i];
  }
}

/* compute_fixed_point_error!  src/pdlp.jl:157-165. */
static double pdlp_fixed_point_error(orc_pdlp* s) {
  int64_t n = s->n, m = s->m;
  for (int64_t j = 0; j < n; j++)
    s->delta_primal[j] = s->pdhg_primal[j] - s->reflected_primal[j];
  for (int64_t i = 0; i < m; i++)
    s->delta_dual[i] = s->pdhg_dual[i] - s->reflected_dual[i];
  double movement = sumsq(s->delta_primal, n) * s->r.primal_weight +
                    sumsq(s->delta_dual, m) / s->r.primal_weight;
  spmv(&s->At, s->delta_dual, s->dual_product);
  double interaction = 2.0 * s->step_size * dotp(s->dual_product, s->delta_primal, n);
  s->r.fixed_point_error = sqrt(jmax(movement + interaction, 0.0));
  return s->r.fixed_point_error;
}

/* perform_restart!  src/pdlp.jl:207-216. */
static void pdlp_perform_restart(orc_pdlp* s, const orc_params* p) {
  int64_t n = s->n, m = s->m;
  for (int64_t j = 0; j < n; j++)
    s->delta_primal[j] = s->initial_primal[j] - s->pdhg_primal[j];
  for (int64_t i = 0; i < m; i++)
    s->delta_dual[i] = s->initial_dual[i] - s->pdhg_dual[i];
  update_primal_weight(&s->r, p, sqrt(sumsq(s->delta_primal, n)),
                       sqrt(sumsq(s->delta_dual, m)));
  memcpy(s->initial_primal, s->pdhg_primal, sizeof(double) * (size_t)n);
  memcpy(s->current_primal, s->pdhg_primal, sizeof(double) * (size_t)n);
  memcpy(s->initial_dual, s->pdhg_dual, sizeof(double) * (size_t)m);
  memcpy(s->current_dual, s->pdhg_dual, sizeof(double) * (size_t)m);
  s->r.inner_count = 0;
  s->r.last_trial_fixed_point_error = INFINITY;
  s->n_restarts++;
}

/* compute_residual!  src/pdlp.jl:220-242. */
static void pdlp_compute_residual(orc_pdlp* s) {
  int64_t n = s->n, m = s->m;
  spmv(&s->A, s->pdhg_primal, s->primal_product);
  for (int64_t i = 0; i < m; i++)
    s->tmp_m[i] = s->primal_product[i] -
                  jclamp(s->primal_product[i], s->con_lb[i], s->con_ub[i]);
  if (s->cupdlpx_termination)
    for (int64_t i = 0; i < m; i++) s->tmp_m[i] *= s->row_unscale[i];
  double primal_residual = sqrt(sumsq(s->tmp_m, m)) / s->bound_scale;
  spmv(&s->At, s->pdhg_dual, s->dual_product);
  for (int64_t j = 0; j < n; j++)
    s->tmp_n[j] = s->c[j] - s->dual_product[j] - s->dual_slack[j];
  if (s->cupdlpx_termination)
    for (int64_t j = 0; j < n; j++) s->tmp_n[j] *= s->col_unscale[j];
  double dual_residual = sqrt(sumsq(s->tmp_n, n)) / s->objective_scale;
  s->r.relative_primal_residual = primal_residual / (1.0 + s->constraint_bound_norm);
  s->r.relative_dual_residual = dual_residual / (1.0 + s->objective_vector_norm);
  double scale = s->bound_scale * s->objective_scale;
  s->objective_value = dotp(s->c, s->pdhg_primal, n) / scale;
  double acc = 0.0;
  for (int64_t i = 0; i < m; i++) {
    double lo = fabs(s->con_lb[i]) < 1e19 ? s->con_lb[i] : 0.0;
    double hi = fabs(s->con_ub[i]) < 1e19 ? s->con_ub[i] : 0.0;
    acc += jmax(s->pdhg_dual[i], 0.0) * lo + jmin(s->pdhg_dual[i], 0.0) * hi;
  }
  double dual_objective = (dotp(s->dual_slack, s->pdhg_primal, n) + acc) / scale;
  s->relative_optimality_gap = fabs(s->objective_value - dual_objective) /
                               (1.0 + fabs(s->objective_value) + fabs(dual_objective));
}

/* compute_infeasibility_information!  src/pdlp.jl:247-285. */
static void pdlp_infeasibility(orc_pdlp* s) {
  int64_t n = s->n, m = s->m;
  double* d = s->delta_primal;
  double* y = s->delta_dual;
  for (int64_t j = 0; j < n; j++) {
    double v = -d[j];
    if (isfinite(s->var_lb[j])) v = jmax(v, 0.0);
    if (isfinite(s->var_ub[j])) v = jmin(v, 0.0);
    d[j] = v;
  }
  for (int64_t i = 0; i < m; i++) {
    double v = -y[i];
    if (!isfinite(s->con_lb[i])) v = jmin(v, 0.0);
    if (!isfinite(s->con_ub[i])) v = jmax(v, 0.0);
    y[i] = v;
  }
  double primal_ray_norm = 0.0;
  for (int64_t j = 0; j < n; j++) primal_ray_norm = jmax(primal_ray_norm, fabs(d[j]));
  if (primal_ray_norm > 0.0)
    for (int64_t j = 0; j < n; j++) d[j] /= primal_ray_norm;
  double dual_ray_norm = 0.0;
  for (int64_t i = 0; i < m; i++) dual_ray_norm = jmax(dual_ray_norm, fabs(y[i]));
  spmv(&s->A, d, s->primal_product);
  spmv(&s->At, y, s->dual_product);
  double scale = s->bound_scale * s->objective_scale;
  s->primal_ray_linear_objective = dotp(s->c, d, n) / scale;
  double primal_slack = 0.0;
  for (int64_t i = 0; i < m; i++) {
    double lo = isfinite(s->con_lb[i]) ? s->con_lb[i] : 0.0;
    double hi = isfinite(s->con_ub[i]) ? s->con_ub[i] : 0.0;
    primal_slack += jmax(y[i], 0.0) * lo + jmin(y[i], 0.0) * hi;
  }
  double dual_slack = 0.0;
  for (int64_t j = 0; j < n; j++) {
    double lo = isfinite(s->var_lb[j]) ? s->var_lb[j] : 0.0;
    double hi = isfinite(s->var_ub[j]) ? s->var_ub[j] : 0.0;
    dual_slack += jmax(-s->dual_product[j], 0.0) * lo + jmin(-s->dual_product[j], 0.0) * hi;
  }
  s->dual_ray_objective = (primal_slack + dual_slack) / scale;
  double max_p = 0.0;
  for (int64_t i = 0; i < m; i++) {
    /* Julia's Bool factor is a strong zero (false * NaN == 0). */
    double v = (isfinite(s->con_lb[i]) ? jmax(0.0, -s->primal_product[i]) : 0.0) +
               (isfinite(s->con_ub[i]) ? jmax(0.0, s->primal_product[i]) : 0.0);
    v *= s->row_unscale[i];
    max_p = jmax(max_p, fabs(v));
  }
  double max_d = 0.0;
  for (int64_t j = 0; j < n; j++) {
    double v = (!isfinite(s->var_lb[j]) ? jmax(0.0, -s->dual_product[j]) : 0.0) -
               (!isfinite(s->var_ub[j]) ? jmin(0.0, -s->dual_product[j]) : 0.0);
    v *= s->col_unscale[j];
    max_d = jmax(max_d, fabs(v));
  }
  s->max_primal_ray_infeasibility = m == 0 ? 0.0 : max_p;
  double dual_scratch_norm = n == 0 ? 0.0 : max_d;
  double scaling = jmax(dual_ray_norm, dual_scratch_norm);
  if (scaling > 0.0) {
    s->max_dual_ray_infeasibility = dual_scratch_norm / scaling;
    s->dual_ray_objective /= scaling;
  } else {
    s->max_dual_ray_infeasibility = 0.0;
    s->dual_ray_objective = 0.0;
  }
}

/* initialize_step_size_and_primal_weight!  src/pdlp.jl:292-316. */
static void pdlp_initialize(orc_pdlp* s, const orc_params* p) {
  int64_t n = s->n, m = s->m;
  if (s->A.nnz == 0) {
    s->step_size = 1.0;
  } else {
    double sv = max_singular_value(&s->A, &s->At, p->sv_max_iterations, p->sv_tolerance);
    s->step_size = (isfinite(sv) && sv > 0) ? 0.998 / sv : 0.998;
  }
  double lo_sq = 0.0, hi_sq = 0.0;
  if (s->cupdlpx_termination) {
    double acc = 0.0;
    for (int64_t j = 0; j < n; j++) {
      double v = s->c[j] * s->col_unscale[j];
      acc += v * v;
    }
    s->objective_vector_norm = sqrt(acc) / s->objective_scale;
    for (int64_t i = 0; i < m; i++) {
      double lo = s->con_lb[i] * s->row_unscale[i] / s->bound_scale;
      double hi = s->con_ub[i] * s->row_unscale[i] / s->bound_scale;
      double lv = (isfinite(lo) && s->con_lb[i] != s->con_ub[i]) ? lo : 0.0;
      double hv = isfinite(hi) ? hi : 0.0;
      lo_sq += lv * lv;
      hi_sq += hv * hv;
    }
  } else {
    s->objective_vector_norm = sqrt(sumsq(s->c, n));
    for (int64_t i = 0; i < m; i++) {
      double lo = s->con_lb[i], hi = s->con_ub[i];
      double lv = (isfinite(lo) && s->con_lb[i] != s->con_ub[i]) ? lo : 0.0;
      double hv = isfinite(hi) ? hi : 0.0;
      lo_sq += lv * lv;
      hi_sq += hv * hv;
    }
  }
  s->constraint_bound_norm = sqrt(lo_sq + hi_sq);
  s->r.primal_weight = 1.0;
  s->r.best_primal_weight = 1.0;
}

static int pdlp_primal_infeasible(const orc_pdlp* s) {
  return s->dual_ray_objective > 0.0 &&
         s->max_dual_ray_infeasibility / s->dual_ray_objective <= s->eps_infeasible;
}
static int pdlp_dual_infeasible(const orc_pdlp* s) {
  return s->primal_ray_linear_objective < 0.0 &&
         s->max_primal_ray_infeasibility / (-s->primal_ray_linear_objective) <=
             s->eps_infeasible;
}

/* pdlp_solve!(s, p)  src/pdlp.jl:325-372. */
EXPORT int32_t orc_pdlp_solve(orc_pdlp* s, const orc_params* p) {
  double t_start = now_s();
  pdlp_initialize(s, p);
  double omega = p->reflection_coefficient;
  int32_t freq = p->termination_evaluation_frequency;
  int do_restart = 0;
  s->n_restarts = 0;
  while (s->r.total_count < p->iteration_limit) {
    pdlp_step(s, 1, omega, 1);
    if (do_restart) {
      pdlp_fixed_point_error(s);
      s->r.initial_fixed_point_error = s->r.fixed_point_error;
      do_restart = 0;
    }
    for (int32_t i = 2; i <= freq - 1; i++) pdlp_step(s, i, omega, 0);
    pdlp_step(s, freq, omega, 1);
    pdlp_fixed_point_error(s);
    pdlp_compute_residual(s);
    if (s->cupdlpx_termination) pdlp_infeasibility(s);
    s->r.inner_count += freq;
    s->r.total_count += freq;
    if (s->r.relative_primal_residual <= p->eps_feasible_relative &&
        s->r.relative_dual_residual <= p->eps_feasible_relative &&
        s->relative_optimality_gap <= p->eps_optimal_relative) {
      s->termination_reason = T_OPTIMAL;
      break;
    }
    if (s->cupdlpx_termination && pdlp_primal_infeasible(s)) {
      s->termination_reason = T_PRIMAL_INFEASIBLE;
      break;
    }
    if (s->cupdlpx_termination && pdlp_dual_infeasible(s)) {
      s->termination_reason = T_DUAL_INFEASIBLE;
      break;
    }
    if (now_s() - t_start > p->time_limit) {
      s->termination_reason = T_TIME_LIMIT;
      break;
    }
    do_restart = should_restart(&s->r, p);
    if (do_restart) pdlp_perform_restart(s, p);
  }
  if (s->termination_reason == T_UNSPECIFIED)
    s->termination_reason = T_ITERATION_LIMIT;
  s->solve_seconds = now_s() - t_start;
  return s->termination_reason;
}

/* which: 0 pdhg_primal, 1 dual_slack, 2 pdhg_dual, 3 current_primal,
 *        4 current_dual. */
EXPORT void orc_pdlp_get_vector(orc_pdlp* s, int32_t which, double* out) {
  switch (which) {
    case 0: memcpy(out, s->pdhg_primal, sizeof(double) * (size_t)s->n); break;
    case 1: memcpy(out, s->dual_slack, sizeof(double) * (size_t)s->n); break;
    case 2: memcpy(out, s->pdhg_dual, sizeof(double) * (size_t)s->m); break;
    case 3: memcpy(out, s->current_primal, sizeof(double) * (size_t)s->n); break;
    case 4: memcpy(out, s->current_dual, sizeof(double) * (size_t)s->m); break;
    default: break;
  }
}

EXPORT void orc_pdlp_get_info(orc_pdlp* s, orc_pdlp_info* out) {
  memset(out, 0, sizeof(*out));
  out->relative_primal_residual = s->r.relative_primal_residual;
  out->relative_dual_residual = s->r.relative_dual_residual;
  out->relative_optimality_gap = s->relative_optimality_gap;
  out->objective_value = s->objective_value;
  out->fixed_point_error = s->r.fixed_point_error;
  out->primal_weight = s->r.primal_weight;
  out->step_size = s->step_size;
  out->solve_seconds = s->solve_seconds;
  out->total_count = s->r.total_count;
  out->termination_reason = s->termination_reason;
  out->n_restarts = s->n_restarts;
}
