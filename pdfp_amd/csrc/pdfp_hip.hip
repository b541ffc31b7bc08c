// libpdfp_hip: the projector (src/pdfp.jl) on gfx950 behind a C ABI.
//
// Host side: the solve loop of pdfp_solve! (src/pdfp.jl:252-288), the
// restart rule and PID primal weight (src/pdlp.jl:167-205), warm starts and
// targets (src/pdfp.jl:98-114).  Device side: everything that touches a
// vector.  Per 200-iteration evaluation block the host reads back one small
// array of FP64 sums (fixed-point error, KKT residuals, ℓ1 gap, restart
// distances), so a block costs one synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pdfp_hip.h"
#include "engine.h"
#include "host_util.h"
#include "pdfp_kernels.h"
#include "small_kernels.h"
#include "solver_common.h"

namespace pdfp {

std::string g_create_error = "no error";

struct ProjectorBase {
  virtual ~ProjectorBase() = default;
  virtual void set_original_primal_residual(const double* D_r) = 0;
  virtual void set_x(const double* x) = 0;
  virtual void set_dual(const double* y, const double* z) = 0;
  virtual void set_target(const double* xt) = 0;
  virtual void set_column_scale(const double* dc) = 0;
  virtual void set_target_original(const double* xt) = 0;
  virtual void warm_start() = 0;
  virtual int solve(const pdfp_params& p, double pw) = 0;
  virtual void get_x(double* out) = 0;
  virtual void get_vector(int which, double* out) = 0;
  virtual void info(pdfp_info* out) = 0;
  virtual double max_singular_value(int iters, double tol) = 0;
  virtual void spmv(int transpose, const double* x, double* y) = 0;
  virtual double time_iterations(int iters) = 0;
  virtual void time_kernels(int iters, double* ms_at, double* ms_a) = 0;
  virtual double time_spmv(int transpose, int iters) = 0;
  virtual int64_t device_bytes() const = 0;
  virtual void set_path(int path) = 0;
  virtual void set_block_graph(int on) = 0;
  virtual void set_schedule(int npb_a, int npb_at, int persist_a, int persist_at) = 0;
};

template <typename T>
struct Projector final : ProjectorBase {
  static constexpr bool kFused = std::is_same<T, float>::value;

  int64_t m, n, nnz;
  int device;
  hipStream_t stream = nullptr;
  DevCsr<T> A, At;  // A: m×n rows, At: n×m rows (columns of A)

  DevBuf<T> b, lb, ub, w, xt;
  DevBuf<T> x, x0, y, y0, rx, px, ds, py, ry;  // n
  DevBuf<T> z, z0, pz, rz;                     // m
  DevBuf<uint8_t> is_eq;
  DevBuf<double> row_unscale;
  DevBuf<double> atz64, ax64;                 // plain FP64 form scratch
  DevBuf<double> lb64, ub64, xt64, w64, x64;  // FP32 form: polish
  DevBuf<double> part_t, part_a, part_v, sums;
  DevBuf<StepParams> d_sp;
  PinnedBuf<StepParams> h_sp;
  PinnedBuf<double> h_sums;
  // x̃ (FP64, scaled) as the caller gave it: the source of its upload and of
  // the ℓ1-gap denominator, which is summed on the host while the first
  // block of a solve runs on the device
  PinnedBuf<double> h_xt;
  bool l1_den_pending = false;
  // D_c for targets given in original units (set_target_original): h_xt
  // then holds x̃ unscaled and the denominator divides as the host would
  DevBuf<double> dc;
  std::vector<double> h_dc;
  bool xt_original = false;

  std::vector<double> h_w, h_b;  // w (0 off I) and b in FP64
  bool orig_primal = false;
  double b_norm = 0.0, b_norm_original = NAN;

  double sv = NAN, step_size = 0.0;
  RestartState r;
  double relative_l1_gap = INFINITY, dual_den = 1.0, l1_den = 1.0;
  double l1_distance = NAN;
  int termination = PDFP_UNSPECIFIED;
  // Whole blocks in one workgroup when the operators fit in LDS
  // (small_kernels.h): path 0 in the FP32 form, path 2 in either form;
  // path 1 forces the general passes (and path 0 keeps the FP64 form on the
  // reference's launch structure).
  bool small_ok = false;
  int small_lanes_a = 1, small_lanes_t = 1;
  int path = 0;
  bool small_used = false;
  // FP32 form, general passes: each block of iterations, its evaluation and
  // the readback of its sums captured once as a HIP graph and replayed
  // (block_graph_on), one graph per (freq, restart evaluation) shape.
  struct BlockGraph {
    hipGraphExec_t exec = nullptr;
    int freq = 0;
  };
  BlockGraph graphs[2];  // [restart evaluation]
  bool block_graph_on = true;
  bool graph_used = false;
  int n_restarts = 0, n_evals = 0;
  double solve_seconds = 0.0;
  bool x64_valid = false;

  Projector(int64_t m_, int64_t n_, const int64_t* colptr, const int64_t* rowval,
            const double* nzval, int32_t base, const double* b_, const uint8_t* is_eq_,
            const uint8_t* is_int_, const double* var_lb, const double* var_ub, const double* w_,
            int dev)
      : m(m_), n(n_), device(dev) {
    PDFP_HIP(hipSetDevice(device));
    PDFP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HostCsr ha = csc_to_csr(m, n, colptr, rowval, nzval, base);
    HostCsr ht = csc_as_transpose_csr(m, n, colptr, rowval, nzval, base);
    nnz = (int64_t)ha.col.size();
    A.upload(ha, stream);
    At.upload(ht, stream);
    small_ok = m > 0 && n > 0 && nnz > 0 && small_lds_bytes<T>(m, n, nnz) <= small_lds_max();
    if (small_ok) {
      small_lanes_a = small_lanes(A.h_rowptr, kSmallBlock);
      small_lanes_t = small_lanes(At.h_rowptr, kSmallBlock);
    }
    if (small_ok && kFused) {
      // the general passes only run when forced: untuned schedules
      A.schedule(512, stream);
      At.schedule(512, stream);
    } else {
      // pick the nonzeros-per-block of each operator's schedule by timing
      autotune_schedule(A, n, stream);
      autotune_schedule(At, m, stream);
    }

    h_w.assign((size_t)n, 0.0);
    for (int64_t j = 0; j < n; ++j) h_w[(size_t)j] = is_int_[j] ? w_[j] : 0.0;
    for (int64_t j = 0; j < n; ++j)
      if (is_int_[j] && !(w_[j] > 0.0))
        throw Error(PDFP_ERR_ARGUMENT, "weights of integer columns must be positive");
    h_xt.alloc((size_t)std::max<int64_t>(n, 1));
    std::memset(h_xt.p, 0, sizeof(double) * (size_t)std::max<int64_t>(n, 1));
    h_b.assign(b_, b_ + m);
    {
      double a2 = 0.0;
      for (int64_t j = 0; j < n; ++j) a2 += h_w[(size_t)j] * h_w[(size_t)j];
      dual_den = 1.0 + std::sqrt(a2);  // 1 + ‖w_I‖₂: w never changes
    }
    double acc = 0.0;
    for (int64_t i = 0; i < m; ++i) acc += b_[i] * b_[i];
    b_norm = std::sqrt(acc);

    auto vec = [&](DevBuf<T>& d, int64_t len, const double* src) {
      d.alloc((size_t)std::max<int64_t>(len, 1));
      std::vector<T> tmp((size_t)std::max<int64_t>(len, 1), (T)0);
      if (src)
        for (int64_t k = 0; k < len; ++k) tmp[(size_t)k] = (T)src[k];
      d.upload(tmp.data(), tmp.size(), stream);
      PDFP_HIP(hipStreamSynchronize(stream));
    };
    vec(b, m, b_);
    vec(lb, n, var_lb);
    vec(ub, n, var_ub);
    vec(w, n, h_w.data());
    vec(xt, n, nullptr);
    for (DevBuf<T>* v : {&x, &x0, &y, &y0, &rx, &px, &ds, &py, &ry}) vec(*v, n, nullptr);
    for (DevBuf<T>* v : {&z, &z0, &pz, &rz}) vec(*v, m, nullptr);
    is_eq.alloc((size_t)std::max<int64_t>(m, 1));
    is_eq.upload(is_eq_, (size_t)m, stream);
    if (!kFused) {
      atz64.alloc((size_t)std::max<int64_t>(n, 1));
      ax64.alloc((size_t)std::max<int64_t>(m, 1));
    } else {
      auto vec64 = [&](DevBuf<double>& d, const double* src) {
        d.alloc((size_t)std::max<int64_t>(n, 1));
        if (src) d.upload(src, (size_t)n, stream);
        PDFP_HIP(hipStreamSynchronize(stream));  // src may be a temporary
      };
      vec64(lb64, var_lb);
      vec64(ub64, var_ub);
      vec64(w64, h_w.data());
      std::vector<double> zeros((size_t)std::max<int64_t>(n, 1), 0.0);
      vec64(xt64, zeros.data());
      vec64(x64, zeros.data());
    }
    part_v.alloc((size_t)std::max<int64_t>(grid_for(std::max(m, n)), 1) * 2 + 8);
    sums.alloc(64);
    d_sp.alloc(1);
    h_sp.alloc(1);
    h_sums.alloc(64);
    // terms of an empty operator are never written: start every slot at 0
    std::memset(h_sums.p, 0, sizeof(double) * 64);
    sums.upload(h_sums.p, 64, stream);
    if constexpr (kFused)
      if (!small_ok) tune_fused();
    part_t.alloc((size_t)std::max(At.nblocks, 1) * kEvalColTerms);
    part_a.alloc((size_t)std::max(A.nblocks, 1) * kEvalRowTerms);
    PDFP_HIP(hipStreamSynchronize(stream));
  }

  // FP32 form: refine each operator's schedule (nonzeros per block, grid
  // form) by timing the fused iteration itself, whose epilogue moves the
  // optimum away from the plain pass's; coordinate-wise, Aᵀ first.  The
  // iterates are zeroed afterwards.
  void tune_fused() {
    if (nnz < kAutotuneMinNnz) return;
    h_sp.p[0] = StepParams{1e-3, 1.0, 1.0, 0.0};
    PDFP_HIP(hipMemcpyAsync(d_sp.p, h_sp.p, sizeof(StepParams), hipMemcpyHostToDevice, stream));
    hipEvent_t e0, e1;
    PDFP_HIP(hipEventCreate(&e0));
    PDFP_HIP(hipEventCreate(&e1));
    auto time_steps = [&] {
      float best = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        step(2, false);
        PDFP_HIP(hipEventRecord(e0, stream));
        for (int i = 0; i < 6; ++i) step(2 + i, false);
        PDFP_HIP(hipEventRecord(e1, stream));
        PDFP_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PDFP_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
      }
      return best;
    };
    constexpr int kPersist[] = {0, 2, 4, 8};
    for (DevCsr<T>* M : {&At, &A}) {
      int bn = M->npb, bp = M->persist;
      float best = time_steps();
      for (int npb : kNpbChoices) {
        M->schedule(npb, stream);
        for (int persist : kPersist) {
          M->persist = persist;
          const float t = time_steps();
          if (t < best) {
            best = t;
            bn = npb;
            bp = persist;
          }
        }
      }
      M->schedule(bn, stream);
      M->persist = bp;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (DevBuf<T>* v : {&x, &x0, &y, &y0, &rx, &px, &ds, &py, &ry, &z, &z0, &pz, &rz})
      PDFP_HIP(hipMemsetAsync(v->p, 0, v->bytes(), stream));
    PDFP_HIP(hipStreamSynchronize(stream));
  }

  ~Projector() override {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      drop_graphs();
      (void)hipStreamDestroy(stream);
    }
  }

  // The captured blocks hold kernel arguments (buffers, schedules, flags):
  // dropped whenever one of them changes.
  void drop_graphs() {
    for (BlockGraph& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = BlockGraph{};
    }
  }

  // Replaces the autotuned schedule of both operators.
  void set_schedule(int npb_a, int npb_at, int persist_a, int persist_at) override {
    auto npb_ok = [](int v) {
      for (int c : kNpbChoices)
        if (v == c) return true;
      return false;
    };
    if (!npb_ok(npb_a) || !npb_ok(npb_at) || persist_a < 0 || persist_a > 16 || persist_at < 0 ||
        persist_at > 16)
      throw Error(PDFP_ERR_ARGUMENT, "bad schedule");
    sync();
    drop_graphs();
    A.schedule(npb_a, stream);
    At.schedule(npb_at, stream);
    A.persist = persist_a;
    At.persist = persist_at;
    part_t.alloc((size_t)std::max(At.nblocks, 1) * kEvalColTerms);
    part_a.alloc((size_t)std::max(A.nblocks, 1) * kEvalRowTerms);
    sync();
  }

  void set_block_graph(int on) override {
    if (on != 0 && on != 1) throw Error(PDFP_ERR_ARGUMENT, "block graph switch must be 0 or 1");
    block_graph_on = on != 0;
  }

  int64_t device_bytes() const override {
    size_t s = A.bytes() + At.bytes();
    for (const DevBuf<T>* v : {&b, &lb, &ub, &w, &xt, &x, &x0, &y, &y0, &rx, &px, &ds, &py, &ry,
                               &z, &z0, &pz, &rz})
      s += v->bytes();
    for (const DevBuf<double>* v : {&row_unscale, &atz64, &ax64, &lb64, &ub64, &xt64, &w64, &x64,
                                    &dc, &part_t, &part_a, &part_v, &sums})
      s += v->bytes();
    return (int64_t)(s + is_eq.bytes());
  }

  void sync() { PDFP_HIP(hipStreamSynchronize(stream)); }

  void upload_T(DevBuf<T>& d, const double* src, int64_t len) {
    std::vector<T> tmp((size_t)len);
    for (int64_t k = 0; k < len; ++k) tmp[(size_t)k] = (T)src[k];
    d.upload(tmp.data(), (size_t)len, stream);
    sync();
  }
  void download_T(const DevBuf<T>& d, double* dst, int64_t len) {
    std::vector<T> tmp((size_t)len);
    d.download(tmp.data(), (size_t)len, stream);
    sync();
    for (int64_t k = 0; k < len; ++k) dst[k] = (double)tmp[(size_t)k];
  }

  void set_original_primal_residual(const double* D_r) override {
    std::vector<double> inv((size_t)m);
    double acc = 0.0;
    for (int64_t i = 0; i < m; ++i) inv[(size_t)i] = 1.0 / D_r[i];
    // b_norm_original = ‖b ./ D_r‖ of the FP64 b (src/pdfp.jl:87)
    for (int64_t i = 0; i < m; ++i) {
      const double v = h_b[(size_t)i] / D_r[i];
      acc += v * v;
    }
    b_norm_original = std::sqrt(acc);
    drop_graphs();  // the evaluation's epilogue takes row_unscale
    row_unscale.alloc((size_t)std::max<int64_t>(m, 1));
    row_unscale.upload(inv.data(), (size_t)m, stream);
    sync();
    orig_primal = true;
  }

  void set_x(const double* xs) override {
    upload_T(x, xs, n);
    upload_T(x0, xs, n);
  }

  // current and initial y (n) and z (m), as set_x does for x.
  void set_dual(const double* ys, const double* zs) override {
    upload_T(y, ys, n);
    upload_T(y0, ys, n);
    upload_T(z, zs, m);
    upload_T(z0, zs, m);
  }

  // set_target!  src/pdfp.jl:98-101: one copy into pinned memory, then an
  // asynchronous upload (and, in the FP32 form, the narrowing on the device).
  void set_target(const double* xts) override {
    sync();  // no copy of an earlier target may still read h_xt
    if (n) {
      std::memcpy(h_xt.p, xts, sizeof(double) * (size_t)n);
      if constexpr (kFused) {
        PDFP_HIP(hipMemcpyAsync(xt64.p, h_xt.p, sizeof(double) * (size_t)n,
                                hipMemcpyHostToDevice, stream));
        k_narrow<T><<<grid_for(n), 256, 0, stream>>>((int)n, xt64.p, xt.p);
        PDFP_HIP(hipGetLastError());
      } else {
        PDFP_HIP(hipMemcpyAsync(xt.p, h_xt.p, sizeof(double) * (size_t)n,
                                hipMemcpyHostToDevice, stream));
      }
    }
    xt_original = false;
    warm_start();
  }

  // The column scaling D_c, kept on the device for set_target_original.
  void set_column_scale(const double* dcs) override {
    for (int64_t j = 0; j < n; ++j)
      if (!(dcs[j] > 0.0) || !std::isfinite(dcs[j]))
        throw Error(PDFP_ERR_ARGUMENT, "column scaling must be positive and finite");
    sync();
    h_dc.assign(dcs, dcs + n);
    if (!dc.p) dc.alloc((size_t)std::max<int64_t>(n, 1));
    if (!xt64.p) xt64.alloc((size_t)std::max<int64_t>(n, 1));  // the FP64 form's staging
    if (n) dc.upload(h_dc.data(), (size_t)n, stream);
    sync();
  }

  // set_target!(s, x̃ ./ D_c)  src/pump.jl:331 with the division on the
  // device: the same copy and upload as set_target, then x̃ ./ D_c into the
  // FP64 target and the iterates' precision in one kernel.
  void set_target_original(const double* xs) override {
    if (!dc.p) throw Error(PDFP_ERR_STATE, "no column scaling (pdfp_set_column_scale)");
    sync();  // no copy of an earlier target may still read h_xt
    if (n) {
      std::memcpy(h_xt.p, xs, sizeof(double) * (size_t)n);
      PDFP_HIP(hipMemcpyAsync(xt64.p, h_xt.p, sizeof(double) * (size_t)n,
                              hipMemcpyHostToDevice, stream));
      k_unscale_target<T><<<grid_for(n), 256, 0, stream>>>((int)n, xt64.p, dc.p, xt.p);
      PDFP_HIP(hipGetLastError());
    }
    xt_original = true;
    warm_start();
  }

  // warm_start!  src/pdfp.jl:104-114
  void warm_start() override {
    if (n) {
      PDFP_HIP(hipMemcpyAsync(x0.p, x.p, sizeof(T) * n, hipMemcpyDeviceToDevice, stream));
      PDFP_HIP(hipMemcpyAsync(y0.p, y.p, sizeof(T) * n, hipMemcpyDeviceToDevice, stream));
    }
    if (m) PDFP_HIP(hipMemcpyAsync(z0.p, z.p, sizeof(T) * m, hipMemcpyDeviceToDevice, stream));
    r.inner_count = 0;
    r.total_count = 0;
    r.fixed_point_error = INFINITY;
    r.last_trial_fixed_point_error = INFINITY;
    termination = PDFP_UNSPECIFIED;
    x64_valid = false;
  }

  // ---- one PDHG iteration ----------------------------------------------------

  template <bool MAJOR>
  void step_fused(int k_offset) {
    EpiPrimalY<T, MAJOR> ex{d_sp.p, k_offset, x.p, x0.p, y.p, y0.p, lb.p, ub.p, xt.p, w.p,
                            rx.p, px.p, ds.p, py.p, ry.p};
    launch_pass<0>(At, Gather1<T>{z.p}, ex, nullptr, stream);
    EpiDualZ<T, MAJOR> ez{d_sp.p, k_offset, z.p, z0.p, b.p, is_eq.p, pz.p, rz.p};
    launch_pass<0>(A, Gather1<T>{rx.p}, ez, nullptr, stream);
  }

  // The reference's structure: SpMV, K1, K2, SpMV, K3 (src/pdfp.jl:158-163).
  void step_plain(int k_offset) {
    if constexpr (!kFused) {
      launch_pass<0>(At, Gather1<double>{z.p}, EpiStore<double>{atz64.p}, nullptr, stream);
      if (n) {
        k_primal<<<grid_for(n), 256, 0, stream>>>(d_sp.p, k_offset, (int)n, px.p, ds.p, rx.p, x.p,
                                                  x0.p, y.p, atz64.p, lb.p, ub.p);
        k_dual_y<<<grid_for(n), 256, 0, stream>>>(d_sp.p, k_offset, (int)n, py.p, y.p, ry.p, y0.p,
                                                  xt.p, rx.p, w.p);
      }
      launch_pass<0>(A, Gather1<double>{rx.p}, EpiStore<double>{ax64.p}, nullptr, stream);
      if (m)
        k_dual_z<<<grid_for(m), 256, 0, stream>>>(d_sp.p, k_offset, (int)m, pz.p, z.p, rz.p, z0.p,
                                                  ax64.p, b.p, is_eq.p);
    }
  }

  void step(int k_offset, bool major) {
    if (kFused) {
      if (major) step_fused<true>(k_offset);
      else step_fused<false>(k_offset);
    } else {
      step_plain(k_offset);
    }
  }

  // Whether blocks run in one workgroup (see small_ok).
  bool small_path() const { return small_ok && (path == 2 || (path == 0 && kFused)); }

  // One block of `freq` iterations and its evaluation(s) in one workgroup.
  void small_block(int freq, bool first_major) {
    {
      SmallArgsV<T> a;
      a.A = SmallCsrT<T>{A.rowptr.p, A.col.p, A.val.p};
      a.At = SmallCsrT<T>{At.rowptr.p, At.col.p, At.val.p};
      a.m = (int)m;
      a.n = (int)n;
      a.nnz = (int)nnz;
      a.lanes_a = small_lanes_a;
      a.lanes_t = small_lanes_t;
      a.freq = freq;
      a.first_major = first_major ? 1 : 0;
      a.sp = d_sp.p;
      a.x = x.p;
      a.y = y.p;
      a.z = z.p;
      a.x0 = x0.p;
      a.y0 = y0.p;
      a.z0 = z0.p;
      a.lb = lb.p;
      a.ub = ub.p;
      a.xt = xt.p;
      a.w = w.p;
      a.b = b.p;
      a.is_eq = is_eq.p;
      a.rx = rx.p;
      a.px = px.p;
      a.ds = ds.p;
      a.py = py.p;
      a.ry = ry.p;
      a.pz = pz.p;
      a.rz = rz.p;
      a.ec = EpiEvalCol<T>{px.p, rx.p, py.p, ry.p, ds.p, xt.p, w.p, x0.p, y0.p};
      a.er = EpiEvalRow<T>{pz.p, rz.p, b.p, z0.p, is_eq.p, orig_primal ? row_unscale.p : nullptr};
      a.sums = sums.p;
      k_small_block<kSmallBlock, T>
          <<<1, kSmallBlock, (size_t)small_lds_bytes<T>(m, n, nnz), stream>>>(a);
      PDFP_HIP(hipGetLastError());
    }
  }

  void read_sums() {
    PDFP_HIP(hipMemcpyAsync(h_sums.p, sums.p, sizeof(double) * 32, hipMemcpyDeviceToHost,
                            stream));
  }

  // One block of the general passes: iteration 1 (major, and evaluated into
  // slot 1 after a restart), iterations 2 .. freq (freq major), the
  // evaluation into slot 0, and (readback) the copy of both slots to the host.
  void general_block(int freq, bool restart_eval, bool readback = true) {
    step(1, restart_eval || freq == 1);
    if (restart_eval) evaluate(1);
    for (int i = 2; i <= freq; ++i) step(i, i == freq);
    evaluate(0);
    if (readback) read_sums();
  }

  // general_block as a captured graph.  Its launches read the block's
  // scalars from d_sp, which the host writes before each replay, so one
  // capture serves every block of its shape.  The graph holds kernel nodes
  // only (no memset, no copy): the readback follows each replay on the
  // stream.
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

  void set_path(int p) override {
    if (p < 0 || p > 2) throw Error(PDFP_ERR_ARGUMENT, "kernel path must be 0, 1 or 2");
    path = p;
  }

  // Evaluation: FP64 sums into sums[slot*16 ...]; see EpiEvalCol/EpiEvalRow.
  void evaluate(int slot) {
    EpiEvalCol<T> ec{px.p, rx.p, py.p, ry.p, ds.p, xt.p, w.p, x0.p, y0.p};
    launch_pass<kEvalColTerms>(At, GatherDelta<T>{pz.p, rz.p}, ec, part_t.p, stream);
    EpiEvalRow<T> er{pz.p, rz.p, b.p, z0.p, is_eq.p, orig_primal ? row_unscale.p : nullptr};
    launch_pass<kEvalRowTerms>(A, Gather1<T>{px.p}, er, part_a.p, stream);
    if (At.nblocks)
      finalize_partials<256><<<1, 256, 0, stream>>>(part_t.p, At.nblocks, kEvalColTerms, 0,
                                                    sums.p + slot * 16);
    if (A.nblocks)
      finalize_partials<256><<<1, 256, 0, stream>>>(part_a.p, A.nblocks, kEvalRowTerms, 0,
                                                    sums.p + slot * 16 + 8);
  }

  void upload_step_params(double omega) {
    h_sp.p[0].step_size = step_size;
    h_sp.p[0].primal_weight = r.primal_weight;
    h_sp.p[0].omega = omega;
    h_sp.p[0].inner_count = (double)r.inner_count;
    PDFP_HIP(hipMemcpyAsync(d_sp.p, h_sp.p, sizeof(StepParams), hipMemcpyHostToDevice, stream));
  }

  double reduce_scalar_vec(int nb, int nacc, double* out) {
    finalize_partials<256><<<1, 256, 0, stream>>>(part_v.p, nb, nacc, 0, sums.p + 48);
    PDFP_HIP(hipMemcpyAsync(h_sums.p + 48, sums.p + 48, sizeof(double) * 16,
                            hipMemcpyDeviceToHost, stream));
    sync();
    for (int k = 0; k < nacc; ++k) out[k] = h_sums.p[48 + k];
    return out[0];
  }

  // estimate_max_singular_value  src/preconditioner.jl:108-125
  double max_singular_value(int max_iterations, double tolerance) override {
    // blocks in one workgroup: the whole power iteration in one too
    if (small_path()) {
      const double s = small_sigma(A, At, m, n, max_iterations, tolerance, stream);
      if (!std::isnan(s)) return s;
    }
    return power_iteration(A, At, m, n, max_iterations, tolerance, stream);
  }

  // initialize_step_size_and_primal_weight!  src/pdfp.jl:220-241
  void initialize(const pdfp_params& p, double pw) {
    if (nnz == 0) {
      step_size = 1.0;
    } else {
      if (std::isnan(sv)) sv = max_singular_value(p.sv_max_iterations, p.sv_tolerance);
      step_size = 0.998 / std::sqrt(1.0 + sv * sv);
    }
    r.primal_weight = pw;
    r.best_primal_weight = pw;
    r.primal_weight_error_sum = 0.0;
    r.primal_weight_last_error = 0.0;
    l1_den_pending = true;  // summed by finish_l1_den, before the first evaluation is read
  }

  // 1 + Σ_I w|x̃| in the reference's order (src/pdfp.jl:230-240).
  void finish_l1_den() {
    if (!l1_den_pending) return;
    double acc = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      const double v = xt_original ? h_xt.p[(size_t)j] / h_dc[(size_t)j] : h_xt.p[(size_t)j];
      acc += h_w[(size_t)j] > 0.0 ? h_w[(size_t)j] * std::fabs(v) : 0.0;
    }
    l1_den = 1.0 + acc;
    l1_den_pending = false;
  }

  // Scalars of one evaluation slot (see evaluate()).
  struct Eval {
    double fpe, rel_primal, rel_dual, rel_gap, l1, primal_dist, dual_dist;
  };
  Eval read_eval(int slot) const {
    const double* S = h_sums.p + slot * 16;
    Eval e;
    const double pw = r.primal_weight;
    const double movement = S[0] * pw + (S[1] + S[8]) / pw;
    const double interaction = 2.0 * step_size * S[2];
    e.fpe = std::sqrt(jmax(movement + interaction, 0.0));
    const double bn = orig_primal ? b_norm_original : b_norm;
    e.rel_primal = std::sqrt(S[9]) / (1.0 + bn);
    e.rel_dual = std::sqrt(S[3]) / dual_den;
    e.rel_gap = std::fabs(S[5] + S[10] - S[4]) / l1_den;
    e.l1 = S[4];
    e.primal_dist = std::sqrt(S[6]);
    e.dual_dist = std::sqrt(S[7] + S[11]);
    return e;
  }

  // pdfp_solve!  src/pdfp.jl:252-288
  int solve(const pdfp_params& p, double primal_weight) override {
    const double t_start = wall_seconds();
    if (p.termination_evaluation_frequency < 1)
      throw Error(PDFP_ERR_ARGUMENT, "termination_evaluation_frequency must be >= 1");
    initialize(p, primal_weight);
    const double omega = p.reflection_coefficient;
    const int freq = p.termination_evaluation_frequency;
    bool do_restart = false;
    n_restarts = 0;
    n_evals = 0;
    x64_valid = false;
    small_used = small_path();
    graph_used = kFused && !small_used && block_graph_on;
    while (r.total_count < p.iteration_limit) {
      upload_step_params(omega);
      const bool restart_eval = do_restart;
      if (small_used) {
        small_block(freq, restart_eval);
        read_sums();
      } else if (graph_used) {
        block_graph(freq, restart_eval);
      } else {
        general_block(freq, restart_eval);
      }
      finish_l1_den();  // on the host while the block runs
      sync();
      ++n_evals;
      if (restart_eval) {
        r.fixed_point_error = read_eval(1).fpe;
        r.initial_fixed_point_error = r.fixed_point_error;
        do_restart = false;
      }
      const Eval e = read_eval(0);
      r.fixed_point_error = e.fpe;
      r.relative_primal_residual = e.rel_primal;
      r.relative_dual_residual = e.rel_dual;
      relative_l1_gap = e.rel_gap;
      l1_distance = e.l1;
      r.inner_count += freq;
      r.total_count += freq;
      if (r.relative_primal_residual <= p.eps_feasible_relative &&
          r.relative_dual_residual <= p.eps_optimal_relative &&
          relative_l1_gap <= p.eps_optimal_relative) {
        termination = PDFP_OPTIMAL;
        break;
      }
      if (wall_seconds() - t_start > p.time_limit) {
        termination = PDFP_TIME_LIMIT;
        break;
      }
      do_restart = should_restart(r, p);
      if (do_restart) {
        // perform_restart!  src/pdfp.jl:180-192
        update_primal_weight(r, p, e.primal_dist, e.dual_dist);
        const int64_t len = std::max(n, m);
        if (len)
          k_restart<T><<<grid_for(len), 256, 0, stream>>>((int)n, (int)m, px.p, py.p, pz.p, x.p,
                                                          x0.p, y.p, y0.p, z.p, z0.p);
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

  void polish() {
    if constexpr (kFused) {
      if (x64_valid || n == 0) return;
      const int g = grid_for(n);
      k_polish<<<g, 256, 0, stream>>>((int)n, px.p, lb.p, ub.p, lb64.p, ub64.p, xt64.p, w64.p,
                                      x64.p, part_v.p);
      double l1;
      reduce_scalar_vec(g, 1, &l1);
      l1_distance = l1;
      x64_valid = true;
    }
  }

  void get_x(double* out) override {
    if (kFused) {
      polish();
      if (n) x64.download(out, (size_t)n, stream);
      sync();
    } else {
      download_T(px, out, n);
    }
  }

  void get_vector(int which, double* out) override {
    switch (which) {
      case 0: get_x(out); break;
      case 1: download_T(py, out, n); break;
      case 2: download_T(pz, out, m); break;
      case 3: download_T(x, out, n); break;
      case 4: download_T(y, out, n); break;
      case 5: download_T(z, out, m); break;
      case 6: download_T(ds, out, n); break;
      case 7: download_T(xt, out, n); break;
      default: throw Error(PDFP_ERR_ARGUMENT, "unknown vector id");
    }
  }

  void info(pdfp_info* o) override {
    if (kFused && termination != PDFP_UNSPECIFIED) polish();
    std::memset(o, 0, sizeof(*o));
    o->relative_primal_residual = r.relative_primal_residual;
    o->relative_dual_residual = r.relative_dual_residual;
    o->relative_l1_gap = relative_l1_gap;
    o->l1_distance = l1_distance;
    o->fixed_point_error = r.fixed_point_error;
    o->initial_fixed_point_error = r.initial_fixed_point_error;
    o->primal_weight = r.primal_weight;
    o->step_size = step_size;
    o->max_singular_value = sv;
    o->solve_seconds = solve_seconds;
    o->total_count = r.total_count;
    o->termination_reason = termination;
    o->n_restarts = n_restarts;
    o->precision = kFused ? PDFP_PREC_FP32_MIXED : PDFP_PREC_FP64;
    o->n_evaluations = n_evals;
    o->npb_a = A.npb;
    o->npb_at = At.npb;
    o->persist_a = A.persist;
    o->persist_at = At.persist;
    o->small_path = small_used ? 1 : 0;
    o->block_graph = graph_used ? 1 : 0;
  }

  void spmv(int transpose, const double* xin, double* yout) override {
    const DevCsr<T>& M = transpose ? At : A;
    const int64_t nin = transpose ? m : n, nout = transpose ? n : m;
    DevBuf<T> xd;
    DevBuf<double> yd;
    xd.alloc((size_t)std::max<int64_t>(nin, 1));
    yd.alloc((size_t)std::max<int64_t>(nout, 1));
    upload_T(xd, xin, nin);
    launch_pass<0>(M, Gather1<T>{xd.p}, EpiStore<double>{yd.p}, nullptr, stream);
    if (nout) yd.download(yout, (size_t)nout, stream);
    sync();
  }

  double time_iterations(int iters) override {
    if (std::isnan(sv)) {
      pdfp_params p;
      pdfp_default_params(&p);
      initialize(p, r.primal_weight);
    }
    upload_step_params(1.0);
    step(1, false);
    sync();
    hipEvent_t e0, e1;
    PDFP_HIP(hipEventCreate(&e0));
    PDFP_HIP(hipEventCreate(&e1));
    PDFP_HIP(hipEventRecord(e0, stream));
    for (int i = 0; i < iters; ++i) step(2 + (i % 198), false);
    PDFP_HIP(hipEventRecord(e1, stream));
    PDFP_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PDFP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / std::max(iters, 1);
  }

  // Mean duration of the transpose pass and of the A pass, each bracketed by
  // its own events on the context's stream (fused form: the two launches of
  // an iteration; plain form: the two SpMVs).
  void time_kernels(int iters, double* ms_at, double* ms_a) override {
    if (std::isnan(sv)) {
      pdfp_params p;
      pdfp_default_params(&p);
      initialize(p, r.primal_weight);
    }
    upload_step_params(1.0);
    std::vector<hipEvent_t> ev((size_t)iters * 4);
    for (auto& e : ev) PDFP_HIP(hipEventCreate(&e));
    for (int i = 0; i < iters; ++i) {
      const int k = 2 + (i % 198);
      if constexpr (kFused) {
        EpiPrimalY<T, false> ex{d_sp.p, k, x.p, x0.p, y.p, y0.p, lb.p, ub.p, xt.p, w.p,
                                rx.p, px.p, ds.p, py.p, ry.p};
        launch_pass<0>(At, Gather1<T>{z.p}, ex, nullptr, stream,
            ev[(size_t)i * 4 + 0], ev[(size_t)i * 4 + 1]);
        EpiDualZ<T, false> ez{d_sp.p, k, z.p, z0.p, b.p, is_eq.p, pz.p, rz.p};
        launch_pass<0>(A, Gather1<T>{rx.p}, ez, nullptr, stream,
            ev[(size_t)i * 4 + 2], ev[(size_t)i * 4 + 3]);
      } else {
        launch_pass<0>(At, Gather1<double>{z.p}, EpiStore<double>{atz64.p}, nullptr, stream,
            ev[(size_t)i * 4 + 0], ev[(size_t)i * 4 + 1]);
        if (n) {
          k_primal<<<grid_for(n), 256, 0, stream>>>(d_sp.p, k, (int)n, px.p, ds.p, rx.p, x.p,
                                                    x0.p, y.p, atz64.p, lb.p, ub.p);
          k_dual_y<<<grid_for(n), 256, 0, stream>>>(d_sp.p, k, (int)n, py.p, y.p, ry.p, y0.p,
                                                    xt.p, rx.p, w.p);
        }
        launch_pass<0>(A, Gather1<double>{rx.p}, EpiStore<double>{ax64.p}, nullptr, stream,
            ev[(size_t)i * 4 + 2], ev[(size_t)i * 4 + 3]);
        if (m)
          k_dual_z<<<grid_for(m), 256, 0, stream>>>(d_sp.p, k, (int)m, pz.p, z.p, rz.p, z0.p,
                                                    ax64.p, b.p, is_eq.p);
      }
    }
    sync();
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < iters; ++i) {
      float t0 = 0.f, t1 = 0.f;
      PDFP_HIP(hipEventElapsedTime(&t0, ev[(size_t)i * 4 + 0], ev[(size_t)i * 4 + 1]));
      PDFP_HIP(hipEventElapsedTime(&t1, ev[(size_t)i * 4 + 2], ev[(size_t)i * 4 + 3]));
      sa += t0;
      sb += t1;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms_at = sa / std::max(iters, 1);
    *ms_a = sb / std::max(iters, 1);
  }

  // Mean duration of a plain SpMV y = A x (transpose: y = Aᵀ x) over the
  // resident operator with its schedule, in the context's precision (FP64
  // accumulation, y stored in T), each launch bracketed by its own events.
  double time_spmv(int transpose, int iters) override {
    const DevCsr<T>& M = transpose ? At : A;
    const int64_t nin = transpose ? m : n, nout = transpose ? n : m;
    DevBuf<T> xd, yd;
    xd.alloc((size_t)std::max<int64_t>(nin, 1));
    yd.alloc((size_t)std::max<int64_t>(nout, 1));
    std::vector<T> ones((size_t)std::max<int64_t>(nin, 1), (T)1);
    xd.upload(ones.data(), ones.size(), stream);
    for (int w = 0; w < 3; ++w)
      launch_pass<0>(M, Gather1<T>{xd.p}, EpiStore<T>{yd.p}, nullptr, stream);
    std::vector<hipEvent_t> ev((size_t)iters * 2);
    for (auto& e : ev) PDFP_HIP(hipEventCreate(&e));
    for (int i = 0; i < iters; ++i)
      launch_pass<0>(M, Gather1<T>{xd.p}, EpiStore<T>{yd.p}, nullptr, stream,
                     ev[(size_t)i * 2], ev[(size_t)i * 2 + 1]);
    sync();
    double sum = 0.0;
    for (int i = 0; i < iters; ++i) {
      float t = 0.f;
      PDFP_HIP(hipEventElapsedTime(&t, ev[(size_t)i * 2], ev[(size_t)i * 2 + 1]));
      sum += t;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return sum / std::max(iters, 1);
  }
};

}  // namespace pdfp

// ---- C ABI ---------------------------------------------------------------------

struct pdfp_ctx {
  std::unique_ptr<pdfp::ProjectorBase> impl;
  std::string err = "no error";
  bool target_set = false;
};

namespace {

template <class F>
int guarded(pdfp_ctx* ctx, F&& f) {
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

void pdfp_default_params(pdfp_params* p) {
  if (!p) return;
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

const char* pdfp_last_error(const pdfp_ctx* ctx) {
  if (!ctx) return pdfp::g_create_error.c_str();
  return ctx->err.c_str();
}

int pdfp_create(pdfp_ctx** out, int64_t m, int64_t n, int64_t nnz, const int64_t* colptr,
                const int64_t* rowval, const double* nzval, int32_t index_base, const double* b,
                const uint8_t* is_eq, const uint8_t* is_int, const double* var_lb,
                const double* var_ub, const double* w, int32_t precision, int32_t device) {
  if (!out) return PDFP_ERR_ARGUMENT;
  *out = nullptr;
  auto fail = [&](int code, const std::string& msg) {
    pdfp::g_create_error = msg;
    return code;
  };
  if (m < 0 || n < 0 || nnz < 0) return fail(PDFP_ERR_ARGUMENT, "negative size");
  if (!colptr || (nnz > 0 && (!rowval || !nzval)) || (m > 0 && (!b || !is_eq)) ||
      (n > 0 && (!is_int || !var_lb || !var_ub || !w)))
    return fail(PDFP_ERR_ARGUMENT, "null input array");
  if (index_base != 0 && index_base != 1) return fail(PDFP_ERR_ARGUMENT, "index_base must be 0 or 1");
  if (colptr[n] - index_base != nnz) return fail(PDFP_ERR_ARGUMENT, "colptr[n] does not match nnz");
  if (precision != PDFP_PREC_FP64 && precision != PDFP_PREC_FP32_MIXED)
    return fail(PDFP_ERR_ARGUMENT, "unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PDFP_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(PDFP_ERR_ARGUMENT, "device index out of range");
  auto* ctx = new (std::nothrow) pdfp_ctx();
  if (!ctx) return fail(PDFP_ERR_NOMEM, "host allocation failed");
  try {
    if (precision == PDFP_PREC_FP64)
      ctx->impl.reset(new pdfp::Projector<double>(m, n, colptr, rowval, nzval, index_base, b,
                                                  is_eq, is_int, var_lb, var_ub, w, device));
    else
      ctx->impl.reset(new pdfp::Projector<float>(m, n, colptr, rowval, nzval, index_base, b,
                                                 is_eq, is_int, var_lb, var_ub, w, device));
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

int pdfp_set_original_primal_residual(pdfp_ctx* ctx, const double* D_r) {
  return guarded(ctx, [&] {
    if (!D_r) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null D_r");
    ctx->impl->set_original_primal_residual(D_r);
  });
}

int pdfp_set_x(pdfp_ctx* ctx, const double* x) {
  return guarded(ctx, [&] {
    if (!x) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null x");
    ctx->impl->set_x(x);
  });
}

int pdfp_set_dual(pdfp_ctx* ctx, const double* y, const double* z) {
  return guarded(ctx, [&] {
    if (!y || !z) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null y or z");
    ctx->impl->set_dual(y, z);
  });
}

int pdfp_set_target(pdfp_ctx* ctx, const double* xt) {
  return guarded(ctx, [&] {
    if (!xt) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null target");
    ctx->impl->set_target(xt);
    ctx->target_set = true;
  });
}

int pdfp_set_column_scale(pdfp_ctx* ctx, const double* D_c) {
  return guarded(ctx, [&] {
    if (!D_c) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null column scaling");
    ctx->impl->set_column_scale(D_c);
  });
}

int pdfp_set_target_original(pdfp_ctx* ctx, const double* xt) {
  return guarded(ctx, [&] {
    if (!xt) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null target");
    ctx->impl->set_target_original(xt);
    ctx->target_set = true;
  });
}

int pdfp_warm_start(pdfp_ctx* ctx) {
  return guarded(ctx, [&] { ctx->impl->warm_start(); });
}

int pdfp_solve(pdfp_ctx* ctx, const pdfp_params* params, double primal_weight,
               int32_t* termination, int64_t* total_count) {
  return guarded(ctx, [&] {
    pdfp_params p;
    if (params) p = *params;
    else pdfp_default_params(&p);
    if (!(primal_weight > 0.0)) throw pdfp::Error(PDFP_ERR_ARGUMENT, "primal_weight must be > 0");
    const int t = ctx->impl->solve(p, primal_weight);
    pdfp_info info;
    ctx->impl->info(&info);
    if (termination) *termination = t;
    if (total_count) *total_count = info.total_count;
  });
}

int pdfp_get_x(pdfp_ctx* ctx, double* out) {
  return guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->get_x(out);
  });
}

int pdfp_get_vector(pdfp_ctx* ctx, int32_t which, double* out) {
  return guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->get_vector(which, out);
  });
}

int pdfp_get_info(pdfp_ctx* ctx, pdfp_info* out) {
  return guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->info(out);
  });
}

void pdfp_destroy(pdfp_ctx* ctx) { delete ctx; }

int pdfp_max_singular_value(pdfp_ctx* ctx, int32_t max_iterations, double tolerance,
                            double* sigma_out) {
  return guarded(ctx, [&] {
    if (!sigma_out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    *sigma_out = ctx->impl->max_singular_value(max_iterations, tolerance);
  });
}

void pdfp_start_vector(int64_t m, uint64_t seed, double* out) {
  if (out && m > 0) pdfp::start_vector(m, seed, out);
}

int pdfp_spmv(pdfp_ctx* ctx, int32_t transpose, const double* x, double* y) {
  return guarded(ctx, [&] {
    if (!x || !y) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null vector");
    ctx->impl->spmv(transpose, x, y);
  });
}

int pdfp_time_iterations(pdfp_ctx* ctx, int32_t iterations, double* ms_out) {
  return guarded(ctx, [&] {
    if (!ms_out || iterations < 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    *ms_out = ctx->impl->time_iterations(iterations);
  });
}

int pdfp_time_kernels(pdfp_ctx* ctx, int32_t iterations, double* ms_at, double* ms_a) {
  return guarded(ctx, [&] {
    if (!ms_at || !ms_a || iterations < 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    ctx->impl->time_kernels(iterations, ms_at, ms_a);
  });
}

int pdfp_time_spmv(pdfp_ctx* ctx, int32_t transpose, int32_t iterations, double* ms_out) {
  return guarded(ctx, [&] {
    if (!ms_out || iterations < 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    *ms_out = ctx->impl->time_spmv(transpose, iterations);
  });
}

int64_t pdfp_device_bytes(const pdfp_ctx* ctx) {
  return ctx && ctx->impl ? ctx->impl->device_bytes() : 0;
}

int pdfp_set_kernel_path(pdfp_ctx* ctx, int32_t path) {
  return guarded(ctx, [&] { ctx->impl->set_path(path); });
}

int pdfp_set_block_graph(pdfp_ctx* ctx, int32_t on) {
  return guarded(ctx, [&] { ctx->impl->set_block_graph(on); });
}

int pdfp_set_schedule(pdfp_ctx* ctx, int32_t npb_a, int32_t npb_at, int32_t persist_a,
                      int32_t persist_at) {
  return guarded(ctx, [&] { ctx->impl->set_schedule(npb_a, npb_at, persist_a, persist_at); });
}

}  // extern "C"
