// libpdfp_hip: the root-LP solver (src/pdlp.jl) on gfx950 behind a C ABI.
//
// Same structure as the projector: two fused CSR passes per iteration, an
// FP64 evaluation every 200 iterations with one host synchronisation, the
// restart rule and PID primal weight of src/pdlp.jl:167-216 on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pdfp_hip.h"
#include "engine.h"
#include "host_util.h"
#include "pdlp_kernels.h"
#include "small_kernels.h"
#include "small_lp_kernels.h"
#include "solver_common.h"

namespace pdfp {

struct LpBase {
  virtual ~LpBase() = default;
  virtual void set_cupdlpx(const double* D_r, const double* D_c) = 0;
  virtual int solve(const pdfp_params& p) = 0;
  virtual void get(int which, double* out) = 0;
  virtual void info(pdlp_info* out) = 0;
  virtual void set_path(int path) = 0;
};

template <typename T>
struct LpSolver final : LpBase {
  int64_t m, n, nnz;
  hipStream_t stream = nullptr;
  DevCsr<T> A, At;
  DevBuf<T> c, lb, ub, clb, cub;
  DevBuf<T> x, x0, rx, px, ds;  // n
  DevBuf<T> y, y0, py, ry;      // m
  DevBuf<double> row_unscale, col_unscale;
  DevBuf<double> part_t, part_a, sums;
  DevBuf<StepParams> d_sp;
  PinnedBuf<StepParams> h_sp;
  PinnedBuf<double> h_sums;

  std::vector<double> h_c, h_clb, h_cub, h_ru, h_cu;
  double bound_scale, objective_scale;
  bool cupdlpx = false;
  double eps_infeasible = 1e-14;
  double constraint_bound_norm = 0.0, objective_vector_norm = 0.0;
  double step_size = 0.0;
  RestartState r;
  double rel_gap = INFINITY, objective_value = NAN;
  double max_primal_ray_inf = NAN, primal_ray_lin_obj = NAN;
  double max_dual_ray_inf = NAN, dual_ray_obj = NAN;
  int termination = PDFP_UNSPECIFIED;
  int n_restarts = 0;
  double solve_seconds = 0.0;
  // FP64 on a small instance: whole blocks in one workgroup
  // (small_lp_kernels.h); path 1 forces the general passes.
  bool small_ok = false;
  int small_lanes_a = 1, small_lanes_t = 1;
  int path = 0;
  bool small_used = false;
  // General passes: each block of iterations, its evaluation(s) and the
  // readback of its sums captured once as a HIP graph and replayed, one
  // graph per (freq, restart evaluation) shape; the launches read the
  // block's scalars from d_sp, which the host writes before each replay.
  struct BlockGraph {
    hipGraphExec_t exec = nullptr;
    int freq = 0;
  };
  BlockGraph graphs[2];

  LpSolver(int64_t m_, int64_t n_, const int64_t* colptr, const int64_t* rowval,
           const double* nzval, int32_t base, const double* c_, const double* lb_,
           const double* ub_, const double* clb_, const double* cub_, double bs, double os,
           int device)
      : m(m_), n(n_), bound_scale(bs), objective_scale(os) {
    PDFP_HIP(hipSetDevice(device));
    PDFP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HostCsr ha = csc_to_csr(m, n, colptr, rowval, nzval, base);
    HostCsr ht = csc_as_transpose_csr(m, n, colptr, rowval, nzval, base);
    nnz = (int64_t)ha.col.size();
    A.upload(ha, stream);
    At.upload(ht, stream);
    // pick the nonzeros-per-block of each operator's schedule by timing
    autotune_schedule(A, n, stream);
    autotune_schedule(At, m, stream);
    if constexpr (std::is_same<T, double>::value) {
      small_ok = m > 0 && n > 0 && nnz > 0 && small_lp_lds_bytes(m, n, nnz) <= small_lds_max();
      if (small_ok) {
        small_lanes_a = small_lanes(A.h_rowptr, kSmallBlock);
        small_lanes_t = small_lanes(At.h_rowptr, kSmallBlock);
      }
    }
    h_c.assign(c_, c_ + n);
    h_clb.assign(clb_, clb_ + m);
    h_cub.assign(cub_, cub_ + m);
    auto vec = [&](DevBuf<T>& d, int64_t len, const double* src) {
      d.alloc((size_t)std::max<int64_t>(len, 1));
      std::vector<T> tmp((size_t)std::max<int64_t>(len, 1), (T)0);
      if (src)
        for (int64_t k = 0; k < len; ++k) tmp[(size_t)k] = (T)src[k];
      d.upload(tmp.data(), tmp.size(), stream);
      PDFP_HIP(hipStreamSynchronize(stream));
    };
    vec(c, n, c_);
    vec(lb, n, lb_);
    vec(ub, n, ub_);
    vec(clb, m, clb_);
    vec(cub, m, cub_);
    for (DevBuf<T>* v : {&x, &x0, &rx, &px, &ds}) vec(*v, n, nullptr);
    for (DevBuf<T>* v : {&y, &y0, &py, &ry}) vec(*v, m, nullptr);
    part_t.alloc((size_t)std::max(At.nblocks, 1) * kLpColTerms);
    part_a.alloc((size_t)std::max(A.nblocks, 1) * kLpRowTerms);
    sums.alloc(64);
    d_sp.alloc(1);
    h_sp.alloc(1);
    h_sums.alloc(64);
    std::memset(h_sums.p, 0, sizeof(double) * 64);
    sums.upload(h_sums.p, 64, stream);
    PDFP_HIP(hipStreamSynchronize(stream));
  }

  ~LpSolver() override {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      drop_graphs();
      (void)hipStreamDestroy(stream);
    }
  }

  // The captured blocks hold kernel arguments (buffers, flags): dropped
  // whenever one of them changes.
  void drop_graphs() {
    for (BlockGraph& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = BlockGraph{};
    }
  }

  void sync() { PDFP_HIP(hipStreamSynchronize(stream)); }

  // cupdlpx_termination!  src/pdlp.jl:105-110
  void set_cupdlpx(const double* D_r, const double* D_c) override {
    h_ru.resize((size_t)m);
    h_cu.resize((size_t)n);
    drop_graphs();  // the evaluation's epilogues take the unscaling vectors
    for (int64_t i = 0; i < m; ++i) h_ru[(size_t)i] = 1.0 / D_r[i];
    for (int64_t j = 0; j < n; ++j) h_cu[(size_t)j] = 1.0 / D_c[j];
    row_unscale.alloc((size_t)std::max<int64_t>(m, 1));
    col_unscale.alloc((size_t)std::max<int64_t>(n, 1));
    row_unscale.upload(h_ru.data(), (size_t)m, stream);
    col_unscale.upload(h_cu.data(), (size_t)n, stream);
    sync();
    cupdlpx = true;
  }

  template <bool MAJOR>
  void step_t(int k) {
    EpiLpPrimal<T, MAJOR> ep{d_sp.p, k, x.p, x0.p, c.p, lb.p, ub.p, rx.p, px.p, ds.p};
    launch_pass<0>(At, Gather1<T>{y.p}, ep, nullptr, stream);
    EpiLpDual<T, MAJOR> ed{d_sp.p, k, y.p, y0.p, clb.p, cub.p, py.p, ry.p};
    launch_pass<0>(A, Gather1<T>{rx.p}, ed, nullptr, stream);
  }
  void step(int k, bool major) {
    if (major) step_t<true>(k);
    else step_t<false>(k);
  }

  template <bool RAYS>
  void evaluate_t(int slot) {
    const double* cu = cupdlpx ? col_unscale.p : nullptr;
    const double* ru = cupdlpx ? row_unscale.p : nullptr;
    EpiLpEvalCol<T, RAYS> ec{px.p, rx.p, ds.p, c.p, x0.p, lb.p, ub.p, cu};
    launch_pass<kLpColTerms, kLpColMax>(At, GatherLpCol<T, RAYS>{py.p, ry.p, clb.p, cub.p}, ec,
                                        part_t.p, stream);
    EpiLpEvalRow<T, RAYS> er{py.p, ry.p, y0.p, clb.p, cub.p, ru};
    launch_pass<kLpRowTerms, kLpRowMax>(A, GatherLpRow<T, RAYS>{px.p, rx.p, lb.p, ub.p}, er,
                                        part_a.p, stream);
    if (At.nblocks)
      finalize_partials<256><<<1, 256, 0, stream>>>(part_t.p, At.nblocks, kLpColTerms, kLpColMax,
                                                    sums.p + slot * 32);
    if (A.nblocks)
      finalize_partials<256><<<1, 256, 0, stream>>>(part_a.p, A.nblocks, kLpRowTerms, kLpRowMax,
                                                    sums.p + slot * 32 + 16);
  }
  void evaluate(int slot) {
    if (cupdlpx) evaluate_t<true>(slot);
    else evaluate_t<false>(slot);
  }

  void set_path(int p) override {
    if (p != 0 && p != 1) throw Error(PDFP_ERR_ARGUMENT, "kernel path must be 0 or 1");
    path = p;
  }

  // One block of `freq` iterations and its evaluation(s) in one workgroup.
  void small_block(int freq, bool restart_eval) {
    if constexpr (std::is_same<T, double>::value) {
      SmallLpArgs a;
      a.A = SmallCsrT<double>{A.rowptr.p, A.col.p, A.val.p};
      a.At = SmallCsrT<double>{At.rowptr.p, At.col.p, At.val.p};
      a.m = (int)m;
      a.n = (int)n;
      a.nnz = (int)nnz;
      a.lanes_a = small_lanes_a;
      a.lanes_t = small_lanes_t;
      a.freq = freq;
      a.restart_eval = restart_eval ? 1 : 0;
      a.sp = d_sp.p;
      a.x = x.p;
      a.y = y.p;
      a.x0 = x0.p;
      a.y0 = y0.p;
      a.c = c.p;
      a.lb = lb.p;
      a.ub = ub.p;
      a.clb = clb.p;
      a.cub = cub.p;
      a.rx = rx.p;
      a.px = px.p;
      a.ds = ds.p;
      a.py = py.p;
      a.ry = ry.p;
      a.col_unscale = cupdlpx ? col_unscale.p : nullptr;
      a.row_unscale = cupdlpx ? row_unscale.p : nullptr;
      a.sums = sums.p;
      const size_t lds = (size_t)small_lp_lds_bytes(m, n, nnz);
      if (cupdlpx)
        k_small_lp_block<kSmallBlock, true><<<1, kSmallBlock, lds, stream>>>(a);
      else
        k_small_lp_block<kSmallBlock, false><<<1, kSmallBlock, lds, stream>>>(a);
      PDFP_HIP(hipGetLastError());
    }
  }

  // initialize_step_size_and_primal_weight!  src/pdlp.jl:292-316
  void initialize(const pdfp_params& p) {
    if (nnz == 0) {
      step_size = 1.0;
    } else {
      // a small instance runs the whole iteration in one workgroup
      double sv = small_sigma(A, At, m, n, p.sv_max_iterations, p.sv_tolerance, stream);
      if (std::isnan(sv))
        sv = power_iteration(A, At, m, n, p.sv_max_iterations, p.sv_tolerance, stream);
      step_size = (std::isfinite(sv) && sv > 0) ? 0.998 / sv : 0.998;
    }
    double lo_sq = 0.0, hi_sq = 0.0;
    if (cupdlpx) {
      double acc = 0.0;
      for (int64_t j = 0; j < n; ++j) {
        const double v = h_c[(size_t)j] * h_cu[(size_t)j];
        acc += v * v;
      }
      objective_vector_norm = std::sqrt(acc) / objective_scale;
      for (int64_t i = 0; i < m; ++i) {
        const double lo = h_clb[(size_t)i] * h_ru[(size_t)i] / bound_scale;
        const double hi = h_cub[(size_t)i] * h_ru[(size_t)i] / bound_scale;
        const double lv = (std::isfinite(lo) && h_clb[(size_t)i] != h_cub[(size_t)i]) ? lo : 0.0;
        const double hv = std::isfinite(hi) ? hi : 0.0;
        lo_sq += lv * lv;
        hi_sq += hv * hv;
      }
    } else {
      double acc = 0.0;
      for (int64_t j = 0; j < n; ++j) acc += h_c[(size_t)j] * h_c[(size_t)j];
      objective_vector_norm = std::sqrt(acc);
      for (int64_t i = 0; i < m; ++i) {
        const double lo = h_clb[(size_t)i], hi = h_cub[(size_t)i];
        const double lv = (std::isfinite(lo) && lo != hi) ? lo : 0.0;
        const double hv = std::isfinite(hi) ? hi : 0.0;
        lo_sq += lv * lv;
        hi_sq += hv * hv;
      }
    }
    constraint_bound_norm = std::sqrt(lo_sq + hi_sq);
    r.primal_weight = 1.0;
    r.best_primal_weight = 1.0;
  }

  void upload_step_params(double omega) {
    h_sp.p[0].step_size = step_size;
    h_sp.p[0].primal_weight = r.primal_weight;
    h_sp.p[0].omega = omega;
    h_sp.p[0].inner_count = (double)r.inner_count;
    PDFP_HIP(hipMemcpyAsync(d_sp.p, h_sp.p, sizeof(StepParams), hipMemcpyHostToDevice, stream));
  }

  struct Eval {
    double fpe, primal_dist, dual_dist;
  };

  double fpe_of(int slot) const {
    const double* C = h_sums.p + slot * 32;
    const double* R = C + 16;
    const double pw = r.primal_weight;
    const double movement = C[0] * pw + R[0] / pw;
    const double interaction = 2.0 * step_size * C[1];
    return std::sqrt(jmax(movement + interaction, 0.0));
  }

  // compute_residual! (src/pdlp.jl:220-242) and
  // compute_infeasibility_information! (src/pdlp.jl:247-285) from slot 0.
  Eval read_eval() {
    const double* C = h_sums.p;
    const double* R = C + 16;
    Eval e;
    e.fpe = fpe_of(0);
    const double primal_residual = std::sqrt(R[1]) / bound_scale;
    const double dual_residual = std::sqrt(C[2]) / objective_scale;
    r.relative_primal_residual = primal_residual / (1.0 + constraint_bound_norm);
    r.relative_dual_residual = dual_residual / (1.0 + objective_vector_norm);
    const double scale = bound_scale * objective_scale;
    objective_value = C[3] / scale;
    const double dual_objective = (C[4] + R[2]) / scale;
    rel_gap = std::fabs(objective_value - dual_objective) /
              (1.0 + std::fabs(objective_value) + std::fabs(dual_objective));
    e.primal_dist = std::sqrt(C[5]);
    e.dual_dist = std::sqrt(R[3]);
    if (cupdlpx) {
      const double prn = n == 0 ? 0.0 : C[8];
      // d is normalised by max|d| before A d and c·d; both are positively
      // homogeneous in d, so the unnormalised sums are divided here
      primal_ray_lin_obj = (prn > 0.0 ? C[6] / prn : C[6]) / scale;
      dual_ray_obj = (R[4] + C[7]) / scale;
      max_primal_ray_inf = m == 0 ? 0.0 : (prn > 0.0 ? R[6] / prn : R[6]);
      const double dual_ray_norm = m == 0 ? 0.0 : R[5];
      const double dual_scratch_norm = n == 0 ? 0.0 : C[9];
      const double scaling = jmax(dual_ray_norm, dual_scratch_norm);
      if (scaling > 0.0) {
        max_dual_ray_inf = dual_scratch_norm / scaling;
        dual_ray_obj /= scaling;
      } else {
        max_dual_ray_inf = 0.0;
        dual_ray_obj = 0.0;
      }
    }
    return e;
  }

  bool primal_infeasible() const {
    return dual_ray_obj > 0.0 && max_dual_ray_inf / dual_ray_obj <= eps_infeasible;
  }
  bool dual_infeasible() const {
    return primal_ray_lin_obj < 0.0 &&
           max_primal_ray_inf / (-primal_ray_lin_obj) <= eps_infeasible;
  }

  void read_sums() {
    PDFP_HIP(hipMemcpyAsync(h_sums.p, sums.p, sizeof(double) * 64, hipMemcpyDeviceToHost,
                            stream));
  }

  // One block of the general passes: iterations 1 and freq major, the
  // restart evaluation after iteration 1, the block's evaluation and
  // (readback) the copy of the sums to the host.
  void general_block(int freq, bool restart_eval, bool readback = true) {
    step(1, true);
    if (restart_eval) evaluate(1);
    for (int i = 2; i <= freq - 1; ++i) step(i, false);
    step(freq, true);
    evaluate(0);
    if (readback) read_sums();
  }

  // general_block as a captured graph, captured on first use of its shape
  // and replayed after; it holds kernel nodes only (no memset, no copy):
  // the readback follows each replay on the stream.
  void block_graph(int freq, bool restart_eval) {
    BlockGraph& g = graphs[restart_eval ? 1 : 0];
    if (!g.exec || g.freq != freq) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = BlockGraph{};
      hipGraph_t graph = nullptr;
      PDFP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        general_block(freq, restart_eval, false);
      } catch (...) {
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      PDFP_HIP(hipStreamEndCapture(stream, &graph));
      const hipError_t st = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      PDFP_HIP(st);
      g.freq = freq;
    }
    PDFP_HIP(hipGraphLaunch(g.exec, stream));
    read_sums();
  }

  // pdlp_solve!  src/pdlp.jl:325-372
  int solve(const pdfp_params& p) override {
    const double t_start = wall_seconds();
    if (p.termination_evaluation_frequency < 2)
      throw Error(PDFP_ERR_ARGUMENT, "termination_evaluation_frequency must be >= 2");
    initialize(p);
    const double omega = p.reflection_coefficient;
    const int freq = p.termination_evaluation_frequency;
    bool do_restart = false;
    n_restarts = 0;
    termination = PDFP_UNSPECIFIED;
    small_used = small_ok && path == 0;
    while (r.total_count < p.iteration_limit) {
      upload_step_params(omega);
      const bool restart_eval = do_restart;
      if (small_used) {
        small_block(freq, restart_eval);
        read_sums();
      } else {
        block_graph(freq, restart_eval);
      }
      sync();
      if (restart_eval) {
        r.fixed_point_error = fpe_of(1);
        r.initial_fixed_point_error = r.fixed_point_error;
        do_restart = false;
      }
      const Eval e = read_eval();
      r.fixed_point_error = e.fpe;
      r.inner_count += freq;
      r.total_count += freq;
      if (r.relative_primal_residual <= p.eps_feasible_relative &&
          r.relative_dual_residual <= p.eps_feasible_relative &&
          rel_gap <= p.eps_optimal_relative) {
        termination = PDFP_OPTIMAL;
        break;
      }
      if (cupdlpx && primal_infeasible()) {
        termination = PDFP_PRIMAL_INFEASIBLE;
        break;
      }
      if (cupdlpx && dual_infeasible()) {
        termination = PDFP_DUAL_INFEASIBLE;
        break;
      }
      if (wall_seconds() - t_start > p.time_limit) {
        termination = PDFP_TIME_LIMIT;
        break;
      }
      do_restart = should_restart(r, p);
      if (do_restart) {
        // perform_restart!  src/pdlp.jl:207-216
        update_primal_weight(r, p, e.primal_dist, e.dual_dist);
        const int64_t len = std::max(n, m);
        if (len)
          k_lp_restart<T><<<grid_for(len), 256, 0, stream>>>((int)n, (int)m, px.p, py.p, x.p,
                                                             x0.p, y.p, y0.p);
        r.inner_count = 0;
        r.last_trial_fixed_point_error = INFINITY;
        ++n_restarts;
      }
    }
    if (termination == PDFP_UNSPECIFIED) termination = PDFP_ITERATION_LIMIT;
    sync();
    solve_seconds = wall_seconds() - t_start;
    return termination;
  }

  void download(const DevBuf<T>& d, double* out, int64_t len) {
    std::vector<T> tmp((size_t)len);
    d.download(tmp.data(), (size_t)len, stream);
    sync();
    for (int64_t k = 0; k < len; ++k) out[k] = (double)tmp[(size_t)k];
  }

  void get(int which, double* out) override {
    if (which == 0) download(px, out, n);
    else if (which == 1) download(ds, out, n);
    else if (which == 2) download(py, out, m);
    else throw Error(PDFP_ERR_ARGUMENT, "unknown vector id");
  }

  void info(pdlp_info* o) override {
    std::memset(o, 0, sizeof(*o));
    o->relative_primal_residual = r.relative_primal_residual;
    o->relative_dual_residual = r.relative_dual_residual;
    o->relative_optimality_gap = rel_gap;
    o->objective_value = objective_value;
    o->fixed_point_error = r.fixed_point_error;
    o->primal_weight = r.primal_weight;
    o->step_size = step_size;
    o->solve_seconds = solve_seconds;
    o->total_count = r.total_count;
    o->termination_reason = termination;
    o->n_restarts = n_restarts;
    o->small_path = small_used ? 1 : 0;
  }
};

}  // namespace pdfp

struct pdlp_ctx {
  std::unique_ptr<pdfp::LpBase> impl;
  std::string err = "no error";
};

namespace {

template <class F>
int lp_guarded(pdlp_ctx* ctx, F&& f) {
  if (!ctx) return PDFP_ERR_ARGUMENT;
  try {
    f();
    return PDFP_OK;
  } catch (const pdfp::Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    ctx->err = "host allocation failed";
    return PDFP_ERR_NOMEM;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return PDFP_ERR_DEVICE;
  }
}

}  // namespace

extern "C" {

const char* pdlp_last_error(const pdlp_ctx* ctx) {
  if (!ctx) return pdfp::g_create_error.c_str();
  return ctx->err.c_str();
}

int pdlp_set_kernel_path(pdlp_ctx* ctx, int32_t path) {
  return lp_guarded(ctx, [&] {
    if (!ctx->impl) throw pdfp::Error(PDFP_ERR_STATE, "no solver");
    ctx->impl->set_path(path);
  });
}

int pdlp_create(pdlp_ctx** out, int64_t m, int64_t n, int64_t nnz, const int64_t* colptr,
                const int64_t* rowval, const double* nzval, int32_t index_base, const double* c,
                const double* var_lb, const double* var_ub, const double* con_lb,
                const double* con_ub, double bound_scale, double objective_scale,
                int32_t device) {
  if (!out) return PDFP_ERR_ARGUMENT;
  *out = nullptr;
  auto fail = [&](int code, const std::string& msg) {
    pdfp::g_create_error = msg;
    return code;
  };
  if (m < 0 || n < 0 || nnz < 0) return fail(PDFP_ERR_ARGUMENT, "negative size");
  if (!colptr || (nnz > 0 && (!rowval || !nzval)) || (m > 0 && (!con_lb || !con_ub)) ||
      (n > 0 && (!c || !var_lb || !var_ub)))
    return fail(PDFP_ERR_ARGUMENT, "null input array");
  if (index_base != 0 && index_base != 1) return fail(PDFP_ERR_ARGUMENT, "index_base must be 0 or 1");
  if (colptr[n] - index_base != nnz) return fail(PDFP_ERR_ARGUMENT, "colptr[n] does not match nnz");
  if (!(bound_scale > 0.0) || !(objective_scale > 0.0))
    return fail(PDFP_ERR_ARGUMENT, "scales must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PDFP_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(PDFP_ERR_ARGUMENT, "device index out of range");
  auto* ctx = new (std::nothrow) pdlp_ctx();
  if (!ctx) return fail(PDFP_ERR_NOMEM, "host allocation failed");
  try {
    ctx->impl.reset(new pdfp::LpSolver<double>(m, n, colptr, rowval, nzval, index_base, c, var_lb,
                                               var_ub, con_lb, con_ub, bound_scale,
                                               objective_scale, device));
  } catch (const pdfp::Error& e) {
    delete ctx;
    return fail(e.code, e.what());
  } catch (const std::exception& e) {
    delete ctx;
    return fail(PDFP_ERR_DEVICE, e.what());
  }
  *out = ctx;
  return PDFP_OK;
}

int pdlp_set_cupdlpx_termination(pdlp_ctx* ctx, const double* D_r, const double* D_c) {
  return lp_guarded(ctx, [&] {
    if (!D_r || !D_c) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null scaling vector");
    ctx->impl->set_cupdlpx(D_r, D_c);
  });
}

int pdlp_solve(pdlp_ctx* ctx, const pdfp_params* params, int32_t* termination,
               int64_t* total_count) {
  return lp_guarded(ctx, [&] {
    pdfp_params p;
    if (params) p = *params;
    else pdfp_default_params(&p);
    const int t = ctx->impl->solve(p);
    pdlp_info info;
    ctx->impl->info(&info);
    if (termination) *termination = t;
    if (total_count) *total_count = info.total_count;
  });
}

int pdlp_get_primal(pdlp_ctx* ctx, double* out) {
  return lp_guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->get(0, out);
  });
}

int pdlp_get_dual_slack(pdlp_ctx* ctx, double* out) {
  return lp_guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->get(1, out);
  });
}

int pdlp_get_info(pdlp_ctx* ctx, pdlp_info* out) {
  return lp_guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->info(out);
  });
}

void pdlp_destroy(pdlp_ctx* ctx) { delete ctx; }

}  // extern "C"
