// libpdfp_hip: a batch of up to 16 projector trajectories sharing one matrix.
//
// The reference runs concurrent trajectories as separate OS processes, each
// with its own PDFPState on the same GPU (scripts/solve.jl:121-148).  Here
// the K trajectories of one instance share the scaled A and Aᵀ in HBM and
// run their PDHG iterations in the same two fused launches per iteration;
// each trajectory keeps its own target, iterates, primal weight, restart
// state and termination, with exactly the per-trajectory logic of
// pdfp_solve! (src/pdfp.jl:252-288).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pdfp_hip.h"
#include "batch_kernels.h"
#include "engine.h"
#include "host_util.h"
#include "pdfp_kernels.h"
#include "small_kernels.h"
#include "solver_common.h"

namespace pdfp {

// The one-workgroup block (small_kernels.h) for the trajectories of a batch
// whose instance fits in LDS: workgroup b runs the trajectory of args[b] on
// its view (stride KB) of the interleaved vectors.
template <int KB>
using SmallBatchArgs = SmallArgsT<EpiBatchEvalCol<KB>, EpiBatchEvalRow<KB>>;

template <int BLOCK, int KB>
__global__ __launch_bounds__(BLOCK) void k_small_batch(const SmallBatchArgs<KB>* __restrict__ args) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[(BLOCK / kWave) * kEvalColTerms];
  const SmallBatchArgs<KB> a = args[blockIdx.x];
  small_block_body<BLOCK, KB>(a, smem, red);
}

// Iterations i0 .. i1 of a block (none of them major: the block's first and
// last iterations run as ordinary launches) in one cooperative launch: each
// iteration is the Aᵀ pass and the A pass of step_t, every workgroup walking
// its share of the row blocks as csr_pass_persistent does, with a grid
// barrier after each pass in place of the kernel boundary.  The two passes
// share one LDS array.  Returns early, with the barrier's error word set,
// if a barrier wait gives up.
template <int KB, int NPB_A, int NPB_T>
__global__ __launch_bounds__(kBlock) void k_batch_coop(Csr<float> A, Csr<float> At,
                                                       EpiBatchPrimalY<KB, false> ex,
                                                       EpiBatchDualZ<KB, false> ez, int i0,
                                                       int i1, GridBarrier* bar,
                                                       unsigned int spin_limit) {
  using LT = PassLds<kBlock, NPB_T, GatherBatch<KB>, EpiBatchPrimalY<KB, false>>;
  using LA = PassLds<kBlock, NPB_A, GatherBatch<KB>, EpiBatchDualZ<KB, false>>;
  __shared__ __align__(16) unsigned char lds[LT::kBytes > LA::kBytes ? LT::kBytes : LA::kBytes];
  const GatherBatch<KB> gz{ez.z};   // the Aᵀ pass gathers z
  const GatherBatch<KB> gx{ex.rx};  // the A pass gathers x̄
  for (int i = i0; i <= i1; ++i) {
    ex.k_offset = i;
    ez.k_offset = i;
    walk_blocks<kBlock, NPB_T, float, GatherBatch<KB>, EpiBatchPrimalY<KB, false>, 0, 0>(
        At, gz, ex, nullptr, lds);
    if (!grid_sync(bar, gridDim.x, spin_limit)) return;
    walk_blocks<kBlock, NPB_A, float, GatherBatch<KB>, EpiBatchDualZ<KB, false>, 0, 0>(
        A, gx, ez, nullptr, lds);
    if (i < i1 && !grid_sync(bar, gridDim.x, spin_limit)) return;
  }
}

struct Trajectory {
  RestartState r;
  bool orig_primal = false;
  double relative_l1_gap = INFINITY, l1_den = 1.0, l1_distance = NAN;
  int termination = PDFP_UNSPECIFIED;
  int n_restarts = 0, n_evals = 0;
  double solve_seconds = 0.0;
  bool do_restart = false;
  bool running = false;  // begun and not yet terminated
  pdfp_params p{};       // the parameters of its current solve
  double t_start = 0.0;
  bool l1_den_pending = false;  // summed on the host while the next block runs
  bool xt_original = false;     // its h_xt holds x̃ unscaled (set_target_original)
};

// What the C ABI sees of a batch; Batch<KB> implements it for a batch width
// KB of 8 or 16 (pdfp_batch_create picks the narrowest that holds K).
struct BatchBase {
  int K = 0;
  std::vector<Trajectory> tr;
  bool graph_on = true;
  bool coop_on = false;
  virtual ~BatchBase() = default;
  virtual void set_coop(int per_cu) = 0;
  virtual void set_slices(int on) = 0;
  virtual void set_slice_min_len(int len) = 0;
  virtual void set_nt(int on) = 0;
  virtual void set_original_primal_residual(const double* D_r) = 0;
  virtual void set_x(int k, const double* xs) = 0;
  virtual void set_target(int k, const double* xts) = 0;
  virtual void set_column_scale(const double* dc) = 0;
  virtual void set_target_original(int k, const double* xs) = 0;
  virtual void warm_start(int k) = 0;
  virtual void begin(int k, const pdfp_params& p, double pw) = 0;
  virtual int run(int mode) = 0;
  virtual void solve(const pdfp_params& p, const double* pw, const uint8_t* act) = 0;
  virtual void get_x(int k, double* out) = 0;
  virtual void get_vector(int k, int which, double* out) = 0;
  virtual void info(int k, pdfp_info* o) = 0;
  virtual void time_kernels(int iters, double* ms_at, double* ms_a) = 0;
  virtual double time_iterations(int iters) = 0;
  virtual void set_schedule(int npb_a, int npb_at, int persist, int lane_a, int lane_at) = 0;
  virtual void set_path(int p) = 0;
  virtual void drop_graphs() = 0;
};

template <int KB>
struct Batch final : BatchBase {
  static constexpr int kBatch = KB;
  static constexpr int kEvalRows = 256 / KB;
  int64_t m, n, nnz;
  hipStream_t stream = nullptr;
  DevCsr<float> A, At;
  DevBuf<float> b, lb, ub, w;                          // shared
  DevBuf<float> xt, x, x0, y, y0, rx, px, ds, py, ry;  // n × 8
  DevBuf<float> z, z0, pz, rz;                         // m × 8
  DevBuf<uint8_t> is_eq;
  DevBuf<double> row_unscale;
  DevBuf<double> lb64, ub64, w64, x64;
  // each trajectory's target x̃ (FP64, scaled) as the caller gave it: in
  // pinned host memory (the upload's source and the ℓ1-gap denominator's
  // input) and on the device (for the polish), n values per trajectory
  PinnedBuf<double> h_xt;
  DevBuf<double> xt64;
  // D_c for targets given in original units (set_target_original)
  DevBuf<double> dc;
  std::vector<double> h_dc;
  DevBuf<float> tmpf;
  DevBuf<double> part_t, part_a, part_v, sums;
  DevBuf<BatchStep> d_sp;
  PinnedBuf<BatchStep> h_sp;
  DevBuf<double2> d_wt;    // Halpern weights [k_offset][8] of the current block
  PinnedBuf<double2> h_wt;
  int wt_rows = 0;
  // the evaluation masks of the current block: [0] the running
  // trajectories, [1] those evaluated after a restart
  DevBuf<int32_t> d_mask;
  PinnedBuf<int32_t> h_mask;
  // Each block of iterations, its evaluations and the readback of its sums
  // captured once as a HIP graph and replayed (graph_on), one graph per
  // (freq, restart evaluation) shape; the launches read the block's scalars,
  // weights and masks from device memory that the host writes before each
  // replay.
  struct BlockGraph {
    hipGraphExec_t exec = nullptr;
    int freq = 0;
  };
  BlockGraph graphs[2];
  // The instance fits in LDS: each block runs as one launch with one
  // workgroup per running trajectory, each iterating as the single FP32
  // projector's one-workgroup path; path 1 forces the batched passes.
  bool small_ok = false;
  int small_lanes_a = 1, small_lanes_t = 1;
  int path = 0;
  bool small_used = false;
  DevBuf<StepParams> d_ssp;
  PinnedBuf<StepParams> h_ssp;
  DevBuf<SmallBatchArgs<KB>> d_sargs;
  PinnedBuf<SmallBatchArgs<KB>> h_sargs;
  PinnedBuf<double> h_sums;
  std::vector<double> h_w, h_b;
  double b_norm = 0.0, b_norm_original = NAN, dual_den = 1.0;
  bool orig_primal = false;
  double sv = NAN, step_size = 0.0;
  int x64_owner = -1;  // the trajectory whose polished point x64 holds, or -1
  std::string err = "no error";

  Batch(int K_, int64_t m_, int64_t n_, const int64_t* colptr, const int64_t* rowval,
        const double* nzval, int32_t base, const double* b_, const uint8_t* is_eq_,
        const uint8_t* is_int_, const double* var_lb, const double* var_ub, const double* w_,
        int device)
      : m(m_), n(n_) {
    K = K_;
    PDFP_HIP(hipSetDevice(device));
    PDFP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HostCsr ha = csc_to_csr(m, n, colptr, rowval, nzval, base);
    HostCsr ht = csc_as_transpose_csr(m, n, colptr, rowval, nzval, base);
    nnz = (int64_t)ha.col.size();
    A.upload(ha, stream);
    At.upload(ht, stream);
    // eight gathered FP32 values and a_ij staged per nonzero (36 bytes), 32
    // rows per block: blocks hold 256 or 512 nonzeros (autotune() picks)
    A.schedule(256, stream, kBlock / kBatch);
    At.schedule(256, stream, kBlock / kBatch);
    tr.resize((size_t)K);
    h_w.assign((size_t)n, 0.0);
    for (int64_t j = 0; j < n; ++j) {
      if (is_int_[j] && !(w_[j] > 0.0))
        throw Error(PDFP_ERR_ARGUMENT, "weights of integer columns must be positive");
      h_w[(size_t)j] = is_int_[j] ? w_[j] : 0.0;
    }
    h_b.assign(b_, b_ + m);
    double acc = 0.0;
    for (int64_t i = 0; i < m; ++i) acc += b_[i] * b_[i];
    b_norm = std::sqrt(acc);
    acc = 0.0;
    for (int64_t j = 0; j < n; ++j) acc += h_w[(size_t)j] * h_w[(size_t)j];
    dual_den = 1.0 + std::sqrt(acc);
    auto vecf = [&](DevBuf<float>& d, int64_t len, const double* src) {
      d.alloc((size_t)std::max<int64_t>(len, 1));
      std::vector<float> tmp((size_t)std::max<int64_t>(len, 1), 0.f);
      if (src)
        for (int64_t k = 0; k < len; ++k) tmp[(size_t)k] = (float)src[k];
      d.upload(tmp.data(), tmp.size(), stream);
      PDFP_HIP(hipStreamSynchronize(stream));
    };
    vecf(b, m, b_);
    vecf(lb, n, var_lb);
    vecf(ub, n, var_ub);
    vecf(w, n, h_w.data());
    for (DevBuf<float>* v : {&xt, &x, &x0, &y, &y0, &rx, &px, &ds, &py, &ry})
      vecf(*v, n * kBatch, nullptr);
    for (DevBuf<float>* v : {&z, &z0, &pz, &rz}) vecf(*v, m * kBatch, nullptr);
    vecf(tmpf, std::max(m, n), nullptr);
    is_eq.alloc((size_t)std::max<int64_t>(m, 1));
    is_eq.upload(is_eq_, (size_t)m, stream);
    auto vec64 = [&](DevBuf<double>& d, const double* src) {
      d.alloc((size_t)std::max<int64_t>(n, 1));
      std::vector<double> tmp((size_t)std::max<int64_t>(n, 1), 0.0);
      if (src) std::copy(src, src + n, tmp.begin());
      d.upload(tmp.data(), tmp.size(), stream);
      PDFP_HIP(hipStreamSynchronize(stream));
    };
    vec64(lb64, var_lb);
    vec64(ub64, var_ub);
    vec64(w64, h_w.data());
    vec64(x64, nullptr);
    xt64.alloc((size_t)std::max<int64_t>(n, 1) * (size_t)K);
    h_xt.alloc((size_t)std::max<int64_t>(n, 1) * (size_t)K);
    std::memset(h_xt.p, 0, sizeof(double) * h_xt.n);
    {
      std::vector<double> zero(h_xt.n, 0.0);
      xt64.upload(zero.data(), zero.size(), stream);
      PDFP_HIP(hipStreamSynchronize(stream));
    }
    part_v.alloc((size_t)std::max(grid_for(std::max(m, n)), 1) + 8);
    sums.alloc((size_t)kBatch * 32 + 16);
    h_sums.alloc((size_t)kBatch * 32 + 16);
    std::memset(h_sums.p, 0, sizeof(double) * ((size_t)kBatch * 32 + 16));
    sums.upload(h_sums.p, (size_t)kBatch * 32 + 16, stream);
    d_sp.alloc(kBatch);
    h_sp.alloc(kBatch);
    d_mask.alloc(2);
    h_mask.alloc(2);
    small_ok = m > 0 && n > 0 && nnz > 0 && small_lds_bytes(m, n, nnz) <= small_lds_max();
    if (small_ok) {
      small_lanes_a = small_lanes(A.h_rowptr, kSmallBlock);
      small_lanes_t = small_lanes(At.h_rowptr, kSmallBlock);
      d_ssp.alloc(kBatch);
      h_ssp.alloc(kBatch);
      d_sargs.alloc(kBatch);
      h_sargs.alloc(kBatch);
    }
    for (int q = 0; q < kBatch; ++q) h_sp.p[q] = BatchStep{1.0, 1.0, 1.0, 0.0, 0.0, {0, 0, 0}};
    autotune();
    part_t.alloc((size_t)std::max<int64_t>((n + kEvalRows - 1) / kEvalRows, 1) * kBatch *
                 kEvalColTerms);
    // the row terms run on A's schedule (set_schedule grows this if a new
    // schedule has more blocks)
    part_a.alloc((size_t)std::max<int32_t>(A.nblocks, 1) * kBatch * kEvalRowTerms);
    PDFP_HIP(hipStreamSynchronize(stream));
  }

  ~Batch() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      drop_graphs();
      (void)hipStreamDestroy(stream);
    }
  }

  // The captured blocks hold kernel arguments (buffers, schedules, flags):
  // dropped whenever one of them changes.
  void drop_graphs() override {
    for (BlockGraph& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = BlockGraph{};
    }
  }

  void sync() { PDFP_HIP(hipStreamSynchronize(stream)); }
  void check_k(int k) const {
    if (k < 0 || k >= K) throw Error(PDFP_ERR_ARGUMENT, "trajectory index out of range");
  }

  // Column k of an interleaved vector from / to FP64 host data.
  void scatter(DevBuf<float>& dst, int64_t len, int k, const double* src) {
    std::vector<float> tmp((size_t)std::max<int64_t>(len, 1));
    for (int64_t i = 0; i < len; ++i) tmp[(size_t)i] = (float)src[i];
    tmpf.upload(tmp.data(), (size_t)len, stream);
    if (len) k_batch_scatter<KB><<<grid_for(len), 256, 0, stream>>>((int)len, k, tmpf.p, dst.p);
    sync();
  }
  void gather(const DevBuf<float>& src, int64_t len, int k, double* out) {
    if (len) k_batch_gather<KB><<<grid_for(len), 256, 0, stream>>>((int)len, k, src.p, tmpf.p);
    std::vector<float> tmp((size_t)std::max<int64_t>(len, 1));
    tmpf.download(tmp.data(), (size_t)len, stream);
    sync();
    for (int64_t i = 0; i < len; ++i) out[i] = (double)tmp[(size_t)i];
  }

  void set_original_primal_residual(const double* D_r) override {
    std::vector<double> inv((size_t)m);
    double acc = 0.0;
    for (int64_t i = 0; i < m; ++i) {
      inv[(size_t)i] = 1.0 / D_r[i];
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

  void set_x(int k, const double* xs) override {
    check_k(k);
    scatter(x, n, k, xs);
    scatter(x0, n, k, xs);
  }

  void warm_start(int k) override {
    check_k(k);
    const int64_t len = std::max(n, m);
    if (len) k_batch_warm<KB><<<grid_for(len), 256, 0, stream>>>((int)n, (int)m, k, x.p, y.p, z.p,
                                                              x0.p, y0.p, z0.p);
    Trajectory& t = tr[(size_t)k];
    t.r.inner_count = 0;
    t.r.total_count = 0;
    t.r.fixed_point_error = INFINITY;
    t.r.last_trial_fixed_point_error = INFINITY;
    t.termination = PDFP_UNSPECIFIED;
    x64_owner = -1;
  }

  // 1 + Σ_I w|x̃| of trajectory k in the reference's order (src/pdfp.jl:230-240).
  void finish_l1_den(int k) {
    Trajectory& t = tr[(size_t)k];
    if (!t.l1_den_pending) return;
    const double* h = h_xt.p + (size_t)k * (size_t)n;
    double acc = 0.0;
    if (t.xt_original) {
      for (int64_t j = 0; j < n; ++j)
        acc += h_w[(size_t)j] > 0.0 ? h_w[(size_t)j] * std::fabs(h[(size_t)j] / h_dc[(size_t)j])
                                    : 0.0;
    } else {
      for (int64_t j = 0; j < n; ++j)
        acc += h_w[(size_t)j] > 0.0 ? h_w[(size_t)j] * std::fabs(h[(size_t)j]) : 0.0;
    }
    t.l1_den = 1.0 + acc;
    t.l1_den_pending = false;
  }

  // set_target!  src/pdfp.jl:98-101: one copy into pinned memory, then an
  // asynchronous upload and the narrowing and interleaving on the device.
  void set_target(int k, const double* xts) override {
    check_k(k);
    sync();  // no copy of an earlier target may still read h_xt
    if (n) {
      double* h = h_xt.p + (size_t)k * (size_t)n;
      double* d = xt64.p + (size_t)k * (size_t)n;
      std::memcpy(h, xts, sizeof(double) * (size_t)n);
      PDFP_HIP(hipMemcpyAsync(d, h, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, stream));
      k_batch_scatter64<KB><<<grid_for(n), 256, 0, stream>>>((int)n, k, d, xt.p);
      PDFP_HIP(hipGetLastError());
    }
    tr[(size_t)k].xt_original = false;
    if (x64_owner == k) x64_owner = -1;
    warm_start(k);
  }

  // The column scaling D_c, kept on the device for set_target_original.
  void set_column_scale(const double* dcs) override {
    for (int64_t j = 0; j < n; ++j)
      if (!(dcs[j] > 0.0) || !std::isfinite(dcs[j]))
        throw Error(PDFP_ERR_ARGUMENT, "column scaling must be positive and finite");
    sync();
    h_dc.assign(dcs, dcs + n);
    if (!dc.p) dc.alloc((size_t)std::max<int64_t>(n, 1));
    if (n) dc.upload(h_dc.data(), (size_t)n, stream);
    sync();
  }

  // set_target!(s, x̃ ./ D_c)  src/pump.jl:331 with the division on the
  // device: the copy and upload of set_target, then x̃ ./ D_c into the FP64
  // target and the interleaved FP32 one in one kernel.
  void set_target_original(int k, const double* xs) override {
    check_k(k);
    if (!dc.p) throw Error(PDFP_ERR_STATE, "no column scaling (pdfp_batch_set_column_scale)");
    sync();  // no copy of an earlier target may still read h_xt
    if (n) {
      double* h = h_xt.p + (size_t)k * (size_t)n;
      double* d = xt64.p + (size_t)k * (size_t)n;
      std::memcpy(h, xs, sizeof(double) * (size_t)n);
      PDFP_HIP(hipMemcpyAsync(d, h, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, stream));
      k_batch_unscale64<KB><<<grid_for(n), 256, 0, stream>>>((int)n, k, d, dc.p, xt.p);
      PDFP_HIP(hipGetLastError());
    }
    tr[(size_t)k].xt_original = true;
    if (x64_owner == k) x64_owner = -1;
    warm_start(k);
  }

  // The batched passes stage 8 FP32 values and a_ij per nonzero: 256 or 512
  // nonzeros per block (about 11 or 21 KB of LDS with the row arrays).
  template <class G, class E>
  void launch_batch(const DevCsr<float>& M, G g, E e, hipEvent_t t0 = nullptr,
                    hipEvent_t t1 = nullptr) {
    if (M.nblocks == 0) {
      for (hipEvent_t t : {t0, t1})
        if (t) PDFP_HIP(hipEventRecord(t, stream));
      return;
    }
    // the lane epilogues serve a block's (row, trajectory) pairs in one pass
    if (M.max_rows > kBlock / kBatch)
      throw Error(PDFP_ERR_STATE, "batch schedule holds too many rows per block");
    if (M.npb == 512) launch_pass_npb<512, 0, 0>(M, g, e, nullptr, stream, t0, t1);
    else launch_pass_npb<256, 0, 0>(M, g, e, nullptr, stream, t0, t1);
  }

  // One iteration's two passes; with ev, ev[0..1] bracket the Aᵀ pass and
  // ev[2..3] the A pass (kernel start and end events).
  template <bool MAJOR, bool NT>
  void step_nt(int k_offset, hipEvent_t* ev) {
    using EX = std::conditional_t<NT, EpiBatchPrimalYNT<KB, MAJOR>, EpiBatchPrimalY<KB, MAJOR>>;
    using EZ = std::conditional_t<NT, EpiBatchDualZNT<KB, MAJOR>, EpiBatchDualZ<KB, MAJOR>>;
    const EX ex{d_sp.p, d_wt.p, k_offset, x.p, x0.p, y.p, y0.p, xt.p, lb.p, ub.p, w.p,
                rx.p, px.p, ds.p, py.p, ry.p};
    launch_batch(At, GatherBatch<KB>{z.p}, ex, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr);
    if (slices_on) {
      launch_sliced_a<MAJOR, NT>(k_offset, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr);
      return;
    }
    const EZ ez{d_sp.p, d_wt.p, k_offset, z.p, z0.p, b.p, is_eq.p, pz.p, rz.p};
    launch_batch(A, GatherBatch<KB>{rx.p}, ez, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr);
  }
  template <bool MAJOR>
  void step_t(int k_offset, hipEvent_t* ev = nullptr) {
    if (epi_nt) step_nt<MAJOR, true>(k_offset, ev);
    else step_nt<MAJOR, false>(k_offset, ev);
  }

  // Schedule autotuner of the batched passes: time whole iterations (the
  // best of two or three timings of each form) for each nonzeros-per-block
  // of A and Aᵀ and each grid form, then for each operator the row-length
  // bounds of the lane-per-(row, trajectory) reduction, then the column
  // slices of the A pass.  A form replaces the one kept so far only if it is
  // at least 2% faster, so that timing noise does not decide between forms
  // that run alike (the first form tried is kept then).
  void autotune() {
    if (nnz == 0) return;
    constexpr int kRows = kBlock / kBatch;
    if (nnz < kAutotuneMinNnz) {  // see kAutotuneMinNnz
      A.schedule(512, stream, kRows, 32);
      At.schedule(512, stream, kRows, 32);
      A.persist = At.persist = 0;
      return;
    }
    // every trajectory active at zero step, as a solve runs (inactive
    // trajectories store no iterate, which times the passes wrong); the
    // iterates are zeroed at the end, as created
    std::vector<bool> none((size_t)K, false);
    upload_steps(1.0, std::vector<bool>((size_t)K, true));
    hipEvent_t e0, e1;
    PDFP_HIP(hipEventCreate(&e0));
    PDFP_HIP(hipEventCreate(&e1));
    auto time8 = [&] {
      for (int w = 0; w < 2; ++w) step(2, false);
      PDFP_HIP(hipEventRecord(e0, stream));
      for (int r = 0; r < 8; ++r) step(2, false);
      PDFP_HIP(hipEventRecord(e1, stream));
      PDFP_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      PDFP_HIP(hipEventElapsedTime(&ms, e0, e1));
      return ms;
    };
    auto best_of = [&](int reps) {
      float t = 1e30f;
      for (int r = 0; r < reps; ++r) t = std::min(t, time8());
      return t;
    };
    constexpr float kGain = 0.98f;  // the margin a form must win by
    float best = 1e30f;
    int bA = 256, bAt = 256, bP = 0;
    for (int na : {256, 512})
      for (int nt : {256, 512})
        for (int persist : {0, 4}) {
          A.schedule(na, stream, kRows);
          At.schedule(nt, stream, kRows);
          A.persist = At.persist = persist;
          const float ms = best_of(2);
          if (ms < kGain * best) {
            best = ms;
            bA = na;
            bAt = nt;
            bP = persist;
          }
        }
    A.schedule(bA, stream, kRows);
    At.schedule(bAt, stream, kRows);
    A.persist = At.persist = bP;
    best = best_of(2);
    for (DevCsr<float>* M : {&A, &At}) {
      int bl = 0;
      for (int len : {32, 128}) {
        M->schedule(M->npb, stream, kRows, len);
        const float ms = best_of(2);
        if (ms < kGain * best) {
          best = ms;
          bl = len;
        }
      }
      M->schedule(M->npb, stream, kRows, bl);
    }
    // the A pass over column slices, where it can pay: every row over the
    // slices, or only the rows longer than a bound whose rows hold 20-95% of
    // the nonzeros (power-law rows), each at both block sizes of the slices'
    // pass; kept only if an iteration is at least 2% faster than in one pass
    // (the best of three timings of each form, taken in turn)
    if (slices_pay(m, n, nnz)) {
      std::vector<int32_t> bounds = {0};
      for (int32_t len : {32, 64, 128, 256}) {
        const double f = (double)nnz_in_rows_over(len) / (double)nnz;
        if (f >= 0.2 && f <= 0.95) bounds.push_back(len);
      }
      const int npbs[2] = {512, 256};
      float t_one = 1e30f, t_best = 1e30f;
      int32_t best_len = 0;
      int best_npb = 512;
      for (int32_t len : bounds) {
        slice_min_len = len;
        build_slices();
        float t_sl[2] = {1e30f, 1e30f};
        for (int rep = 0; rep < 3; ++rep) {
          slices_on = false;
          t_one = std::min(t_one, time8());
          slices_on = true;
          for (int v = 0; v < 2; ++v) {
            slice_schedule(npbs[v]);
            t_sl[v] = std::min(t_sl[v], time8());
          }
        }
        for (int v = 0; v < 2; ++v)
          if (t_sl[v] < kGain * t_best) {
            t_best = t_sl[v];
            best_len = len;
            best_npb = npbs[v];
          }
      }
      if (t_best < kGain * t_one) {
        slice_min_len = best_len;
        build_slices();
        slice_schedule(best_npb);
        slices_on = true;
        if (slice_min_len > 0) {
          // A's schedule now serves the short rows' pass (and the
          // evaluation): its block size and lane-row bound again, with the
          // slices on
          int na_best = A.npb, ln_best = A.lane_len;
          float t = best_of(3);
          for (int na : {256, 512})
            for (int ln : {0, 32, 128}) {
              if (na == A.npb && ln == A.lane_len) continue;
              const int na0 = A.npb, ln0 = A.lane_len;
              A.schedule(na, stream, kRows, ln);
              short_schedule();
              const float ms = best_of(3);
              if (ms < kGain * t) {
                t = ms;
                na_best = na;
                ln_best = ln;
              }
              A.schedule(na0, stream, kRows, ln0);
            }
          A.schedule(na_best, stream, kRows, ln_best);
          short_schedule();
        }
      } else {
        slice_min_len = 0;
        drop_slices();
      }
    }
    // the epilogues' operand loads: non-temporal only if 2% faster
    {
      epi_nt = false;
      const float t_plain = best_of(3);
      epi_nt = true;
      const float t_nt = best_of(3);
      epi_nt = t_nt < kGain * t_plain;
    }
    upload_steps(1.0, none);
    for (DevBuf<float>* d : {&x, &x0, &y, &y0, &rx, &px, &ds, &py, &ry, &z, &z0, &pz, &rz})
      PDFP_HIP(hipMemsetAsync(d->p, 0, d->bytes(), stream));
    sync();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }

  void set_schedule(int npb_a, int npb_at, int persist, int lane_a, int lane_at) override {
    constexpr int kRows = kBlock / kBatch;
    if ((npb_a != 256 && npb_a != 512) || (npb_at != 256 && npb_at != 512) || persist < 0 ||
        persist > 16 || lane_a < 0 || lane_at < 0)
      throw Error(PDFP_ERR_ARGUMENT, "bad schedule");
    sync();
    drop_graphs();
    A.schedule(npb_a, stream, kRows, lane_a);
    At.schedule(npb_at, stream, kRows, lane_at);
    A.persist = At.persist = persist;
    short_schedule();
    const size_t need = (size_t)std::max<int32_t>(A.nblocks, 1) * kBatch * kEvalRowTerms;
    if (part_a.n < need) part_a.alloc(need);
  }
  void step(int k_offset, bool major) {
    if (major) step_t<true>(k_offset);
    else step_t<false>(k_offset);
  }

  // The evaluation of every trajectory in the device mask *mask into its
  // slot, in one pass over each operator (the 8 trajectories share the
  // gathers).
  void evaluate_all(const int32_t* mask, int slot) {
    const int nbt = (int)((n + kEvalRows - 1) / kEvalRows);
    const int nba = (int)((m + kEvalRows - 1) / kEvalRows);
    EpiBatchEvalCol<KB> ec{px.p, rx.p, py.p, ry.p, ds.p, xt.p, w.p, x0.p, y0.p, 0};
    EpiBatchEvalRow<KB> er{pz.p, rz.p, b.p, z0.p, is_eq.p, orig_primal ? row_unscale.p : nullptr, 0};
    if (nbt) {
      k_batch_eval_col<KB><<<nbt, 256, 0, stream>>>((int)n, At.rowptr.p, At.col.p, At.val.p, pz.p,
                                                 rz.p, ec, mask, part_t.p);
      k_batch_eval_finalize<<<kBatch * kEvalColTerms, 256, 0, stream>>>(
          part_t.p, nbt, kEvalColTerms, mask, slot * 16, sums.p);
    }
    if (nba && A.nblocks > 0) {
      if (part_a.n < (size_t)A.nblocks * kBatch * kEvalRowTerms)
        throw Error(PDFP_ERR_STATE, "row-term partials smaller than the schedule");
      EpiBatchEvalRowLanes<KB> el{er, mask};
      if (A.max_rows > kBlock / kBatch)
        throw Error(PDFP_ERR_STATE, "batch schedule holds too many rows per block");
      if (A.npb == 512)
        launch_pass_npb<512, kEvalRowTerms, 0>(A, GatherBatch<KB>{px.p}, el, part_a.p, stream);
      else
        launch_pass_npb<256, kEvalRowTerms, 0>(A, GatherBatch<KB>{px.p}, el, part_a.p, stream);
      k_batch_eval_finalize<<<kBatch * kEvalRowTerms, 256, 0, stream>>>(
          part_a.p, A.nblocks, kEvalRowTerms, mask, slot * 16 + 8, sums.p);
    }
  }

  struct Eval {
    double fpe, rel_primal, rel_dual, rel_gap, l1, primal_dist, dual_dist;
  };
  Eval read_eval(int k, int slot) const {
    const double* S = h_sums.p + (size_t)k * 32 + (size_t)slot * 16;
    const Trajectory& t = tr[(size_t)k];
    const double pw = t.r.primal_weight;
    Eval e;
    const double movement = S[0] * pw + (S[1] + S[8]) / pw;
    e.fpe = std::sqrt(jmax(movement + 2.0 * step_size * S[2], 0.0));
    const double bn = orig_primal ? b_norm_original : b_norm;
    e.rel_primal = std::sqrt(S[9]) / (1.0 + bn);
    e.rel_dual = std::sqrt(S[3]) / dual_den;
    e.rel_gap = std::fabs(S[5] + S[10] - S[4]) / t.l1_den;
    e.l1 = S[4];
    e.primal_dist = std::sqrt(S[6]);
    e.dual_dist = std::sqrt(S[7] + S[11]);
    return e;
  }

  // The block's per-trajectory scalars and Halpern weights for iterations
  // k_offset = 0 .. freq (and at least 0 .. 199, for the timing helpers).
  void upload_steps(double omega, const std::vector<bool>& active, int freq = 200) {
    // the pinned staging and the table are rewritten below: no earlier copy
    // or pass may still read them (the run loop's stream is idle here anyway)
    sync();
    for (int q = 0; q < kBatch; ++q) {
      BatchStep& s = h_sp.p[q];
      const double pw = q < K ? tr[(size_t)q].r.primal_weight : 1.0;
      s.tau = step_size / pw;
      s.sigma = step_size * pw;
      s.omega = omega;
      s.omw = 1.0 - omega;
      s.active = (q < K && active[(size_t)q]) ? 1.0 : 0.0;
    }
    const int rows = std::max(freq, 199) + 1;
    if (rows > wt_rows) {
      drop_graphs();  // the passes take d_wt
      d_wt.alloc((size_t)rows * kBatch);
      h_wt.alloc((size_t)rows * kBatch);
      wt_rows = rows;
    }
    for (int q = 0; q < kBatch; ++q) {
      const double ic = q < K ? (double)tr[(size_t)q].r.inner_count : 0.0;
      for (int ko = 0; ko < rows; ++ko) {
        const double k = ic + (double)ko;
        const double w = k / (k + 1.0);
        h_wt.p[(size_t)ko * kBatch + q] = double2{w, 1.0 - w};
      }
    }
    PDFP_HIP(hipMemcpyAsync(d_sp.p, h_sp.p, sizeof(BatchStep) * kBatch, hipMemcpyHostToDevice,
                            stream));
    PDFP_HIP(hipMemcpyAsync(d_wt.p, h_wt.p, sizeof(double2) * (size_t)rows * kBatch,
                            hipMemcpyHostToDevice, stream));
  }

  // One block of `freq` iterations for every running trajectory: one
  // launch, one workgroup per running trajectory.
  void small_block(int freq, double omega, const std::vector<bool>& active) {
    int na = 0;
    for (int q = 0; q < K; ++q) {
      if (!active[(size_t)q]) continue;
      const Trajectory& t = tr[(size_t)q];
      h_ssp.p[q] = StepParams{step_size, t.r.primal_weight, omega, (double)t.r.inner_count};
      SmallBatchArgs<KB>& a = h_sargs.p[na++];
      a.A = SmallCsr{A.rowptr.p, A.col.p, A.val.p};
      a.At = SmallCsr{At.rowptr.p, At.col.p, At.val.p};
      a.m = (int)m;
      a.n = (int)n;
      a.nnz = (int)nnz;
      a.lanes_a = small_lanes_a;
      a.lanes_t = small_lanes_t;
      a.freq = freq;
      a.first_major = t.do_restart ? 1 : 0;
      a.sp = d_ssp.p + q;
      a.x = x.p + q;
      a.y = y.p + q;
      a.z = z.p + q;
      a.x0 = x0.p + q;
      a.y0 = y0.p + q;
      a.z0 = z0.p + q;
      a.lb = lb.p;
      a.ub = ub.p;
      a.xt = xt.p + q;
      a.w = w.p;
      a.b = b.p;
      a.is_eq = is_eq.p;
      a.rx = rx.p + q;
      a.px = px.p + q;
      a.ds = ds.p + q;
      a.py = py.p + q;
      a.ry = ry.p + q;
      a.pz = pz.p + q;
      a.rz = rz.p + q;
      a.ec = EpiBatchEvalCol<KB>{px.p, rx.p, py.p, ry.p, ds.p, xt.p, w.p, x0.p, y0.p, q};
      a.er = EpiBatchEvalRow<KB>{pz.p, rz.p, b.p, z0.p, is_eq.p,
                             orig_primal ? row_unscale.p : nullptr, q};
      a.sums = sums.p + (size_t)q * 32;
    }
    if (na == 0) return;
    PDFP_HIP(hipMemcpyAsync(d_ssp.p, h_ssp.p, sizeof(StepParams) * kBatch,
                            hipMemcpyHostToDevice, stream));
    PDFP_HIP(hipMemcpyAsync(d_sargs.p, h_sargs.p, sizeof(SmallBatchArgs<KB>) * (size_t)na,
                            hipMemcpyHostToDevice, stream));
    k_small_batch<kSmallBlock, KB>
        <<<na, kSmallBlock, (size_t)small_lds_bytes(m, n, nnz), stream>>>(d_sargs.p);
    PDFP_HIP(hipGetLastError());
    PDFP_HIP(hipMemcpyAsync(h_sums.p, sums.p, sizeof(double) * (size_t)kBatch * 32,
                            hipMemcpyDeviceToHost, stream));
  }

  void set_path(int p) override {
    if (p != 0 && p != 1) throw Error(PDFP_ERR_ARGUMENT, "kernel path must be 0 or 1");
    path = p;
  }

  // One block: iteration 1 (major after a restart; the restarted
  // trajectories evaluated into slot 1), iterations 2 .. freq (freq major),
  // the evaluation of the running trajectories into slot 0, the readback.
  void general_block(int freq, bool any_restart, bool readback = true) {
    step(1, any_restart || freq == 1);
    if (any_restart) evaluate_all(d_mask.p + 1, 1);
    for (int i = 2; i <= freq; ++i) step(i, i == freq);
    evaluate_all(d_mask.p, 0);
    if (readback) read_block_sums();
  }

  void read_block_sums() {
    PDFP_HIP(hipMemcpyAsync(h_sums.p, sums.p, sizeof(double) * (size_t)kBatch * 32,
                            hipMemcpyDeviceToHost, stream));
  }

  // general_block as a captured graph, captured on first use of its shape
  // and replayed after; it holds kernel nodes only (no memset, no copy): the
  // readback follows each replay on the stream.
  void block_graph(int freq, bool any_restart) {
    BlockGraph& g = graphs[any_restart ? 1 : 0];
    if (!g.exec || g.freq != freq) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = BlockGraph{};
      hipGraph_t graph = nullptr;
      PDFP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        general_block(freq, any_restart, false);
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
    read_block_sums();
  }

  // ---- the cooperative form of a block (coop_on) ----
  // Iterations 2 .. freq-1 run in one k_batch_coop launch, the rest of the
  // block as general_block launches it.  The barrier's state is zeroed once;
  // its error word comes back with the block's sums and is checked after the
  // block (check_coop).
  static constexpr unsigned int kCoopSpin = 1u << 20;  // polls before a wait gives up
  DevBuf<GridBarrier> bar;
  PinnedBuf<GridBarrier> h_bar;
  int coop_grid_cache[2][2] = {{0, 0}, {0, 0}};
  int coop_per_cu = 0;  // workgroups per CU of the launch (0: one fewer than fit)
  bool coop_pending = false;

  void set_coop(int per_cu) override {
    if (per_cu < -1 || per_cu > 16)
      throw Error(PDFP_ERR_ARGUMENT, "cooperative workgroups per CU must be -1 .. 16");
    const bool on = per_cu != 0;
    if (on && !bar.p) {
      int dev = 0, coop = 0;
      PDFP_HIP(hipGetDevice(&dev));
      PDFP_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
      if (!coop) throw Error(PDFP_ERR_DEVICE, "the device does not support cooperative launches");
      bar.alloc(1);
      h_bar.alloc(1);
      PDFP_HIP(hipMemsetAsync(bar.p, 0, sizeof(GridBarrier), stream));
      sync();
    }
    coop_on = on;
    coop_per_cu = per_cu > 0 ? per_cu : 0;
    for (auto& r : coop_grid_cache) r[0] = r[1] = 0;
  }

  // Workgroups of the cooperative launch: coop_per_cu per CU, or one per CU
  // fewer than the occupancy of the kernel allows (× the CUs), a multiple of
  // the 8 XCDs.  A grid at the occupancy limit leaves no slack for a CU that
  // holds fewer: at 7 per CU (1792 workgroups) a barrier wait timed out (the
  // bounded wait ended the launch) where 6 per CU had run.
  template <class Kern>
  int coop_grid(Kern kern, int& cache) {
    if (cache == 0) {
      int per_cu = 0, dev = 0, cus = 0;
      PDFP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0));
      if (coop_per_cu > per_cu)
        throw Error(PDFP_ERR_ARGUMENT, "more cooperative workgroups per CU than fit");
      if (coop_per_cu > 0) per_cu = coop_per_cu;
      else if (per_cu > 1) per_cu -= 1;
      PDFP_HIP(hipGetDevice(&dev));
      PDFP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      cache = per_cu * cus / 8 * 8;
      if (cache < 8) throw Error(PDFP_ERR_DEVICE, "the cooperative block does not fit the device");
    }
    return cache;
  }

  template <int NA, int NT>
  void launch_coop_t(int i0, int i1, int& cache) {
    auto kern = k_batch_coop<KB, NA, NT>;
    const int grid = coop_grid(kern, cache);
    Csr<float> av = A.view(), atv = At.view();
    EpiBatchPrimalY<KB, false> ex{d_sp.p, d_wt.p, i0, x.p, x0.p, y.p, y0.p, xt.p, lb.p, ub.p, w.p,
                                  rx.p, px.p, ds.p, py.p, ry.p};
    EpiBatchDualZ<KB, false> ez{d_sp.p, d_wt.p, i0, z.p, z0.p, b.p, is_eq.p, pz.p, rz.p};
    GridBarrier* bp = bar.p;
    unsigned int spin = kCoopSpin;
    void* args[] = {&av, &atv, &ex, &ez, &i0, &i1, &bp, &spin};
    PDFP_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(grid),
                                        dim3(kBlock), args, 0, stream));
  }

  // Iterations i0 .. i1 (non-major) in one cooperative launch.
  void launch_coop(int i0, int i1) {
    if (!bar.p) throw Error(PDFP_ERR_STATE, "cooperative blocks are not enabled");
    if (i1 < i0) return;
    if (A.max_rows > kBlock / kBatch || At.max_rows > kBlock / kBatch)
      throw Error(PDFP_ERR_STATE, "batch schedule holds too many rows per block");
    const int ia = A.npb == 512 ? 1 : 0, it = At.npb == 512 ? 1 : 0;
    int& cache = coop_grid_cache[ia][it];
    if (ia && it) launch_coop_t<512, 512>(i0, i1, cache);
    else if (ia) launch_coop_t<512, 256>(i0, i1, cache);
    else if (it) launch_coop_t<256, 512>(i0, i1, cache);
    else launch_coop_t<256, 256>(i0, i1, cache);
  }

  void coop_block(int freq, bool any_restart) {
    step(1, any_restart || freq == 1);
    if (any_restart) evaluate_all(d_mask.p + 1, 1);
    if (freq > 2) launch_coop(2, freq - 1);
    if (freq >= 2) step(freq, true);
    evaluate_all(d_mask.p, 0);
    PDFP_HIP(hipMemcpyAsync(h_bar.p, bar.p, sizeof(GridBarrier), hipMemcpyDeviceToHost, stream));
    read_block_sums();
    coop_pending = true;
  }

  // After the stream has drained: a barrier wait that gave up leaves the
  // block's iterates undefined; the state is reset and the call fails.
  void check_coop() {
    if (!coop_pending) return;
    coop_pending = false;
    if (h_bar.p->err[0] != 0u) {
      PDFP_HIP(hipMemsetAsync(bar.p, 0, sizeof(GridBarrier), stream));
      sync();
      throw Error(PDFP_ERR_DEVICE, "a grid barrier wait of the cooperative block timed out");
    }
  }

  // ---- the A pass over 8 column slices (slices_on) ----
  // The long rows of A (more than slice_min_len nonzeros; 0: every row) run
  // over A's 8 column slices.  As: their slices stacked (segment s: the long
  // rows with a nonzero in the columns of slice s, in row order), each
  // segment walked by one XCD; part_s: the FP32 partial row sums, KB per
  // (slice, long row), at the position slice_dense gives each stacked row
  // (s·n_long + l for long row l).  With every row long, k_batch_slice_dualz
  // adds a row's 8 partials and updates z; otherwise the short rows run in
  // one pass over Ash (A with the long rows emptied) whose epilogue adds the
  // long rows' partials (EpiBatchDualZLong; long_idx: each row's index among
  // the long rows, -1 for a short row).
  DevCsr<float> As, Ash;
  DevBuf<float> part_s;
  DevBuf<int32_t> slice_dense, long_idx;
  bool slices_on = false;
  int32_t slice_min_len = 0;
  int32_t built_min_len = -1;  // the bound As was built for; -1: not built
  bool epi_nt = false;  // the epilogues' operands by non-temporal loads (ld_epi)
  int64_t n_long = 0;

  // Column slices pay when x̄ (n·KB FP32 values) is larger than an XCD's
  // 4 MB L2 and an eighth of it is not.
  static bool slices_pay(int64_t m, int64_t n, int64_t nnz) {
    const int64_t xbar = n * KB * 4;
    return nnz >= kAutotuneMinNnz && m > 0 && xbar > (int64_t)4 << 20 &&
           xbar / 8 <= (int64_t)5 << 20;
  }

  // Nonzeros of A in rows with more than len nonzeros.
  int64_t nnz_in_rows_over(int32_t len) const {
    const std::vector<int32_t>& rp = A.h_rowptr;
    int64_t c = 0;
    for (int64_t i = 0; i < m; ++i) {
      const int32_t l = rp[(size_t)i + 1] - rp[(size_t)i];
      if (l > len) c += l;
    }
    return c;
  }

  // Builds As (and, for a bound above 0, Ash and long_idx) for slice_min_len
  // from A's CSR (read back from the device once per build).
  void build_slices() {
    if (built_min_len == slice_min_len || nnz == 0) return;
    release_slices();
    std::vector<int32_t> col((size_t)nnz);
    std::vector<float> val((size_t)nnz);
    A.col.download(col.data(), (size_t)nnz, stream);
    A.val.download(val.data(), (size_t)nnz, stream);
    sync();
    const std::vector<int32_t>& rp = A.h_rowptr;
    std::vector<int32_t> li((size_t)m, -1);
    n_long = 0;
    for (int64_t i = 0; i < m; ++i)
      if (slice_min_len == 0 || rp[(size_t)i + 1] - rp[(size_t)i] > slice_min_len)
        li[(size_t)i] = (int32_t)n_long++;
    HostCsr h;
    h.ncols = n;
    h.rowptr.assign(1, 0);
    std::vector<int32_t> cur(rp.begin(), rp.end() - 1);  // next entry of each row
    std::vector<int32_t> srow;                            // long-row index of each stacked row
    struct Piece {
      int32_t l, k0, k1;  // long-row index, its nonzeros [k0, k1) in the slice
    };
    std::vector<Piece> pieces;
    As.segments.assign(1, 0);
    for (int s = 0; s < 8; ++s) {
      const int64_t hi = n * (s + 1) / 8;  // columns [n·s/8, hi) form slice s
      pieces.clear();
      for (int64_t i = 0; i < m; ++i) {
        if (li[(size_t)i] < 0) continue;
        int32_t k = cur[(size_t)i];
        const int32_t end = rp[(size_t)i + 1];
        const int32_t k0 = k;
        while (k < end && col[(size_t)k] < hi) ++k;  // a row's columns ascend
        cur[(size_t)i] = k;
        if (k > k0) pieces.push_back({li[(size_t)i], k0, k});
      }
      // only the long rows sliced: longest first, so a block's rows are alike
      // (its lanes, one per row and trajectory, finish together) and the long
      // ones start early; every row sliced: row order, so a block's partial
      // sums go to neighbouring rows
      if (slice_min_len > 0)
        std::stable_sort(pieces.begin(), pieces.end(), [](const Piece& a, const Piece& b) {
          return a.k1 - a.k0 > b.k1 - b.k0;
        });
      for (const Piece& pc : pieces) {
        for (int32_t k = pc.k0; k < pc.k1; ++k) {
          h.col.push_back(col[(size_t)k]);
          h.val.push_back((double)val[(size_t)k]);
        }
        h.rowptr.push_back((int32_t)h.col.size());
        srow.push_back(pc.l);
      }
      As.segments.push_back((int64_t)srow.size());
    }
    h.nrows = (int64_t)srow.size();
    // each stacked row's (slice, long row) position in the dense partials
    std::vector<int32_t> dense(srow.size());
    for (int s = 0; s < 8; ++s)
      for (int64_t r = As.segments[(size_t)s]; r < As.segments[(size_t)s + 1]; ++r)
        dense[(size_t)r] = (int32_t)(s * n_long + srow[(size_t)r]);
    if (slice_min_len > 0) {
      HostCsr hs;  // A with the long rows emptied
      hs.nrows = m;
      hs.ncols = n;
      hs.rowptr.assign(1, 0);
      for (int64_t i = 0; i < m; ++i) {
        if (li[(size_t)i] < 0)
          for (int32_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) {
            hs.col.push_back(col[(size_t)k]);
            hs.val.push_back((double)val[(size_t)k]);
          }
        hs.rowptr.push_back((int32_t)hs.col.size());
      }
      if ((int64_t)(h.col.size() + hs.col.size()) != nnz)
        throw Error(PDFP_ERR_STATE, "column slices lost nonzeros");
      Ash.upload(hs, stream);
      long_idx.alloc((size_t)std::max<int64_t>(m, 1));
      long_idx.upload(li.data(), (size_t)m, stream);
      short_schedule();
    } else if ((int64_t)h.col.size() != nnz) {
      throw Error(PDFP_ERR_STATE, "column slices lost nonzeros");
    }
    slice_dense.alloc(std::max<size_t>(dense.size(), 1));
    slice_dense.upload(dense.data(), dense.size(), stream);
    As.upload(h, stream);
    slice_schedule(512);
    part_s.alloc((size_t)std::max<int64_t>(8 * n_long * KB, 1));
    PDFP_HIP(hipMemsetAsync(part_s.p, 0, part_s.bytes(), stream));  // the empty pairs stay 0
    sync();
    built_min_len = slice_min_len;
  }

  // The slices' pass: npb nonzeros per block; with every row sliced, up to
  // 256 rows per block reduced by lane groups (EpiSlicePartial); with only
  // the long rows, up to 256/KB rows per block, those of up to 128 nonzeros
  // reduced one lane per (row, trajectory) (EpiSlicePartialLanes).
  void slice_schedule(int npb) {
    if (slice_min_len == 0) As.schedule(npb, stream, kBlock);
    else As.schedule(npb, stream, kBlock / kBatch, 128);
  }

  // The short rows' pass follows A's schedule (block size, lane rows, grid).
  void short_schedule() {
    if (slice_min_len == 0 || Ash.nrows == 0) return;
    Ash.schedule(A.npb, stream, kBlock / kBatch, A.lane_len);
    Ash.persist = A.persist;
  }

  void release_slices() {
    for (DevCsr<float>* M : {&As, &Ash}) {
      M->rowptr.release();
      M->col.release();
      M->val.release();
      M->sched.release();
      M->nrows = M->nblocks = 0;
      M->segments.clear();
      M->seg_blocks.clear();
    }
    part_s.release();
    slice_dense.release();
    long_idx.release();
    n_long = 0;
    built_min_len = -1;
  }

  void drop_slices() {
    release_slices();
    slices_on = false;
  }

  void set_slices(int on) override {
    if (on != 0 && on != 1 && on != 256 && on != 512)
      throw Error(PDFP_ERR_ARGUMENT, "column slices: 0, 1, 256 or 512");
    if (on && nnz == 0) throw Error(PDFP_ERR_ARGUMENT, "no nonzeros to slice");
    sync();
    drop_graphs();
    if (on) build_slices();
    if (on > 1) slice_schedule(on);
    slices_on = on != 0;
  }

  void set_nt(int on) override {
    if (on != 0 && on != 1) throw Error(PDFP_ERR_ARGUMENT, "non-temporal loads: 0 or 1");
    sync();
    drop_graphs();  // the captured launches name the epilogues' kernels
    epi_nt = on != 0;
  }

  void set_slice_min_len(int len) override {
    if (len < 0) throw Error(PDFP_ERR_ARGUMENT, "the long-row bound must be at least 0");
    sync();
    drop_graphs();
    const int npb = As.nrows > 0 ? As.npb : 512;
    slice_min_len = len;
    if (slices_on) {
      build_slices();
      slice_schedule(npb);
    }
  }

  // A x̄ and the z update of iteration k_offset over the column slices; t0 /
  // t1 (optional) bracket both launches.
  template <bool MAJOR, bool NT>
  void launch_sliced_a(int k_offset, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr) {
    SegBlocks sb{};
    int maxb = 0;
    for (int s = 0; s <= 8; ++s) sb.b[s] = As.seg_blocks[(size_t)s];
    for (int s = 0; s < 8; ++s) maxb = std::max(maxb, sb.b[s + 1] - sb.b[s]);
    const GatherBatch<KB> g{rx.p};
    const dim3 grid((unsigned)(8 * std::max(maxb, 1))), block(kBlock);
    auto launch = [&](auto kern, auto ep) {
      if (t0) hipExtLaunchKernelGGL(kern, grid, block, 0, stream, t0, nullptr, 0, As.view(), g, ep,
                                    (double*)nullptr, sb);
      else kern<<<grid, block, 0, stream>>>(As.view(), g, ep, nullptr, sb);
    };
    if (slice_min_len == 0) {
      if (As.max_rows > kBlock) throw Error(PDFP_ERR_STATE, "slice schedule holds too many rows");
      using E = EpiSlicePartial<KB>;
      const E ep{slice_dense.p, part_s.p};
      if (As.npb == 512) launch(csr_pass_xcd_segments<kBlock, 512, float, GatherBatch<KB>, E, 0, 0>, ep);
      else launch(csr_pass_xcd_segments<kBlock, 256, float, GatherBatch<KB>, E, 0, 0>, ep);
    } else {
      if (As.max_rows > kBlock / kBatch)
        throw Error(PDFP_ERR_STATE, "slice schedule holds too many rows per block");
      using E = EpiSlicePartialLanes<KB>;
      const E ep{slice_dense.p, part_s.p};
      if (As.npb == 512) launch(csr_pass_xcd_segments<kBlock, 512, float, GatherBatch<KB>, E, 0, 0>, ep);
      else launch(csr_pass_xcd_segments<kBlock, 256, float, GatherBatch<KB>, E, 0, 0>, ep);
    }
    PDFP_HIP(hipGetLastError());
    using EZ = std::conditional_t<NT, EpiBatchDualZNT<KB, MAJOR>, EpiBatchDualZ<KB, MAJOR>>;
    using EL = std::conditional_t<NT, EpiBatchDualZLongNT<KB, MAJOR>, EpiBatchDualZLong<KB, MAJOR>>;
    const EZ ez{d_sp.p, d_wt.p, k_offset, z.p, z0.p, b.p, is_eq.p, pz.p, rz.p};
    if (slice_min_len > 0) {
      const EL el{ez, long_idx.p, part_s.p, (int32_t)n_long};
      launch_batch(Ash, GatherBatch<KB>{rx.p}, el, nullptr, t1);
      return;
    }
    const int g2 = grid_for(m * KB);
    auto kern = [] {
      if constexpr (NT) return k_batch_slice_dualz_nt<KB, MAJOR>;
      else return k_batch_slice_dualz<KB, MAJOR>;
    }();
    if (t1)
      hipExtLaunchKernelGGL(kern, dim3(g2), dim3(256), 0, stream, nullptr, t1, 0, (int)m,
                            (const float*)part_s.p, ez);
    else
      kern<<<g2, 256, 0, stream>>>((int)m, part_s.p, ez);
    PDFP_HIP(hipGetLastError());
  }

  // initialize_step_size_and_primal_weight! (src/pdfp.jl:220-241) for
  // trajectory k: it starts a solve with its own parameters.
  void begin(int k, const pdfp_params& p, double pw) override {
    check_k(k);
    if (p.termination_evaluation_frequency < 1)
      throw Error(PDFP_ERR_ARGUMENT, "termination_evaluation_frequency must be >= 1");
    if (!(pw > 0.0)) throw Error(PDFP_ERR_ARGUMENT, "primal_weight must be > 0");
    if (nnz > 0 && std::isnan(sv)) {
      // as the single FP32 projector: in one workgroup when the batch's
      // blocks run that way
      if (small_ok && path == 0)
        sv = small_sigma(A, At, m, n, p.sv_max_iterations, p.sv_tolerance, stream);
      if (std::isnan(sv))
        sv = power_iteration(A, At, m, n, p.sv_max_iterations, p.sv_tolerance, stream);
    }
    step_size = nnz == 0 ? 1.0 : 0.998 / std::sqrt(1.0 + sv * sv);
    Trajectory& t = tr[(size_t)k];
    t.p = p;
    t.t_start = wall_seconds();
    t.r.primal_weight = pw;
    t.r.best_primal_weight = pw;
    t.r.primal_weight_error_sum = 0.0;
    t.r.primal_weight_last_error = 0.0;
    t.l1_den_pending = true;  // finish_l1_den, before its first evaluation is read
    t.do_restart = false;
    t.n_restarts = 0;
    t.n_evals = 0;
    t.termination = PDFP_UNSPECIFIED;
    t.running = true;
    x64_owner = -1;
    if (t.r.total_count >= p.iteration_limit) {
      t.termination = PDFP_ITERATION_LIMIT;
      t.running = false;
      t.solve_seconds = 0.0;
    }
  }

  // The block loop of pdfp_solve! (src/pdfp.jl:252-288) for every running
  // trajectory, in lockstep blocks of `freq` iterations, each trajectory with
  // its own termination; returns after the first block in which one ends
  // (until_any) or when none runs.  Returns the number still running.
  // mode: 0 until none runs, 1 until one ends, 2 after one block.
  int run(int mode) override {
    int freq = 0;
    double omega = 0.0;
    for (const Trajectory& t : tr) {
      if (!t.running) continue;
      if (freq == 0) {
        freq = t.p.termination_evaluation_frequency;
        omega = t.p.reflection_coefficient;
      } else if (t.p.termination_evaluation_frequency != freq ||
                 t.p.reflection_coefficient != omega) {
        throw Error(PDFP_ERR_ARGUMENT,
                    "running trajectories must share the evaluation frequency and reflection");
      }
    }
    x64_owner = -1;  // the blocks below move the running trajectories' points
    std::vector<bool> active((size_t)K, false);
    auto refresh = [&] {
      bool any = false;
      for (int k = 0; k < K; ++k) {
        active[(size_t)k] = tr[(size_t)k].running;
        any |= active[(size_t)k];
      }
      return any;
    };
    while (refresh()) {
      upload_steps(omega, active, freq);
      bool any_restart = false;
      for (int k = 0; k < K; ++k) any_restart |= active[(size_t)k] && tr[(size_t)k].do_restart;
      int restart_mask = 0, active_mask = 0;
      for (int k = 0; k < K; ++k) {
        if (!active[(size_t)k]) continue;
        active_mask |= 1 << k;
        if (tr[(size_t)k].do_restart) restart_mask |= 1 << k;
      }
      h_mask.p[0] = active_mask;  // upload_steps synchronised: no copy reads it
      h_mask.p[1] = restart_mask;
      PDFP_HIP(hipMemcpyAsync(d_mask.p, h_mask.p, sizeof(int32_t) * 2, hipMemcpyHostToDevice,
                              stream));
      small_used = small_ok && path == 0;
      if (small_used) small_block(freq, omega, active);
      else if (coop_on && !slices_on) coop_block(freq, any_restart);
      else if (graph_on) block_graph(freq, any_restart);
      else general_block(freq, any_restart);
      for (int k = 0; k < K; ++k)
        if (active[(size_t)k]) finish_l1_den(k);  // on the host while the block runs
      sync();
      check_coop();
      const double now = wall_seconds();
      bool ended = false;
      for (int k = 0; k < K; ++k) {
        if (!active[(size_t)k]) continue;
        Trajectory& t = tr[(size_t)k];
        const pdfp_params& p = t.p;
        ++t.n_evals;
        if (t.do_restart) {
          t.r.fixed_point_error = read_eval(k, 1).fpe;
          t.r.initial_fixed_point_error = t.r.fixed_point_error;
          t.do_restart = false;
        }
        const Eval e = read_eval(k, 0);
        t.r.fixed_point_error = e.fpe;
        t.r.relative_primal_residual = e.rel_primal;
        t.r.relative_dual_residual = e.rel_dual;
        t.relative_l1_gap = e.rel_gap;
        t.l1_distance = e.l1;
        t.r.inner_count += freq;
        t.r.total_count += freq;
        bool stop = false;
        if (t.r.relative_primal_residual <= p.eps_feasible_relative &&
            t.r.relative_dual_residual <= p.eps_optimal_relative &&
            t.relative_l1_gap <= p.eps_optimal_relative) {
          t.termination = PDFP_OPTIMAL;
          stop = true;
        } else if (now - t.t_start > p.time_limit) {
          t.termination = PDFP_TIME_LIMIT;
          stop = true;
        } else {
          t.do_restart = should_restart(t.r, p);
          if (t.do_restart) {
            update_primal_weight(t.r, p, e.primal_dist, e.dual_dist);
            const int64_t len = std::max(n, m);
            if (len)
              k_batch_restart<KB><<<grid_for(len), 256, 0, stream>>>(
                  (int)n, (int)m, k, px.p, py.p, pz.p, x.p, x0.p, y.p, y0.p, z.p, z0.p);
            t.r.inner_count = 0;
            t.r.last_trial_fixed_point_error = INFINITY;
            ++t.n_restarts;
          }
          if (t.r.total_count >= p.iteration_limit) {
            t.termination = PDFP_ITERATION_LIMIT;
            stop = true;
          }
        }
        if (stop) {
          t.running = false;
          t.solve_seconds = now - t.t_start;
          ended = true;
        }
      }
      if (mode == 2 || (mode == 1 && ended)) break;
    }
    sync();
    int n_running = 0;
    for (const Trajectory& t : tr) n_running += t.running ? 1 : 0;
    return n_running;
  }

  // pdfp_solve! (src/pdfp.jl:252-288) for every trajectory with active[k]:
  // each begins with primal weight pw[k], then they run to their ends.
  void solve(const pdfp_params& p, const double* pw, const uint8_t* act) override {
    for (Trajectory& t : tr) t.running = false;
    for (int k = 0; k < K; ++k)
      if (!act || act[k]) begin(k, p, pw[k]);
    run(0);
  }

  void polish(int k) {
    Trajectory& t = tr[(size_t)k];
    if (x64_owner == k || n == 0) return;
    const int g = grid_for(n);
    k_batch_polish<KB><<<g, 256, 0, stream>>>((int)n, k, px.p, lb.p, ub.p, lb64.p, ub64.p,
                                          xt64.p + (size_t)k * (size_t)n, w64.p, x64.p,
                                          part_v.p);
    finalize_partials<256><<<1, 256, 0, stream>>>(part_v.p, g, 1, 0, sums.p + kBatch * 32);
    PDFP_HIP(hipMemcpyAsync(h_sums.p + kBatch * 32, sums.p + kBatch * 32, sizeof(double),
                            hipMemcpyDeviceToHost, stream));
    sync();
    t.l1_distance = h_sums.p[kBatch * 32];
    x64_owner = k;
  }

  void get_x(int k, double* out) override {
    check_k(k);
    polish(k);
    if (n) x64.download(out, (size_t)n, stream);
    sync();
  }

  void get_vector(int k, int which, double* out) override {
    check_k(k);
    switch (which) {
      case 0: get_x(k, out); break;
      case 1: gather(py, n, k, out); break;
      case 2: gather(pz, m, k, out); break;
      case 3: gather(x, n, k, out); break;
      case 4: gather(y, n, k, out); break;
      case 5: gather(z, m, k, out); break;
      case 6: gather(ds, n, k, out); break;
      case 7: gather(xt, n, k, out); break;
      default: throw Error(PDFP_ERR_ARGUMENT, "unknown vector id");
    }
  }

  void info(int k, pdfp_info* o) override {
    check_k(k);
    Trajectory& t = tr[(size_t)k];
    if (t.termination != PDFP_UNSPECIFIED) polish(k);
    std::memset(o, 0, sizeof(*o));
    o->relative_primal_residual = t.r.relative_primal_residual;
    o->relative_dual_residual = t.r.relative_dual_residual;
    o->relative_l1_gap = t.relative_l1_gap;
    o->l1_distance = t.l1_distance;
    o->fixed_point_error = t.r.fixed_point_error;
    o->initial_fixed_point_error = t.r.initial_fixed_point_error;
    o->primal_weight = t.r.primal_weight;
    o->step_size = step_size;
    o->max_singular_value = sv;
    o->solve_seconds = t.solve_seconds;
    o->total_count = t.r.total_count;
    o->termination_reason = t.termination;
    o->n_restarts = t.n_restarts;
    o->precision = PDFP_PREC_FP32_MIXED;
    o->n_evaluations = t.n_evals;
    o->npb_a = A.npb;
    o->npb_at = At.npb;
    o->persist_a = A.persist;
    o->persist_at = At.persist;
    o->lane_rows_a = A.lane_len;
    o->lane_rows_at = At.lane_len;
    o->block_graph = graph_on && !small_used ? 1 : 0;
    o->small_path = small_used ? 1 : 0;
    o->column_slices = slices_on && !small_used ? As.npb : 0;
    o->slice_min_len = slices_on && !small_used ? slice_min_len : 0;
    o->nontemporal = epi_nt && !small_used ? 1 : 0;
  }

  // Mean duration of the transpose pass and of the A pass of a batched
  // iteration, each bracketed by its own events on the batch's stream.
  void time_kernels(int iters, double* ms_at, double* ms_a) override {
    time_iterations(1);
    std::vector<hipEvent_t> ev((size_t)iters * 4);
    for (auto& e : ev) PDFP_HIP(hipEventCreate(&e));
    for (int i = 0; i < iters; ++i) step_t<false>(2 + (i % 198), &ev[(size_t)i * 4]);
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

  double time_iterations(int iters) override {
    if (std::isnan(sv) && nnz > 0) {
      pdfp_params p;
      pdfp_default_params(&p);
      sv = power_iteration(A, At, m, n, p.sv_max_iterations, p.sv_tolerance, stream);
      step_size = 0.998 / std::sqrt(1.0 + sv * sv);
    }
    std::vector<bool> active((size_t)K, true);
    upload_steps(1.0, active);
    step(1, false);
    sync();
    hipEvent_t e0, e1;
    PDFP_HIP(hipEventCreate(&e0));
    PDFP_HIP(hipEventCreate(&e1));
    PDFP_HIP(hipEventRecord(e0, stream));
    if (coop_on && !slices_on) {
      // the cooperative form: runs of up to 198 iterations per launch
      for (int i = 0; i < iters; i += 198) launch_coop(2, 2 + std::min(198, iters - i) - 1);
      PDFP_HIP(hipMemcpyAsync(h_bar.p, bar.p, sizeof(GridBarrier), hipMemcpyDeviceToHost,
                              stream));
      coop_pending = true;
    } else {
      for (int i = 0; i < iters; ++i) step(2 + (i % 198), false);
    }
    PDFP_HIP(hipEventRecord(e1, stream));
    PDFP_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PDFP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    sync();
    check_coop();
    return (double)ms / std::max(iters, 1);
  }
};

}  // namespace pdfp

struct pdfp_batch {
  std::unique_ptr<pdfp::BatchBase> impl;
  std::string err = "no error";
};

namespace {

template <class F>
int batch_guarded(pdfp_batch* ctx, F&& f) {
  if (!ctx || !ctx->impl) return PDFP_ERR_ARGUMENT;
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

const char* pdfp_batch_last_error(const pdfp_batch* ctx) {
  if (!ctx) return pdfp::g_create_error.c_str();
  return ctx->err.c_str();
}

int pdfp_batch_create(pdfp_batch** out, int32_t K, int64_t m, int64_t n, int64_t nnz,
                      const int64_t* colptr, const int64_t* rowval, const double* nzval,
                      int32_t index_base, const double* b, const uint8_t* is_eq,
                      const uint8_t* is_int, const double* var_lb, const double* var_ub,
                      const double* w, int32_t device) {
  if (!out) return PDFP_ERR_ARGUMENT;
  *out = nullptr;
  auto fail = [&](int code, const std::string& msg) {
    pdfp::g_create_error = msg;
    return code;
  };
  if (K < 1 || K > 16) return fail(PDFP_ERR_ARGUMENT, "K must be in 1..16");
  if (m < 0 || n < 0 || nnz < 0) return fail(PDFP_ERR_ARGUMENT, "negative size");
  if (!colptr || (nnz > 0 && (!rowval || !nzval)) || (m > 0 && (!b || !is_eq)) ||
      (n > 0 && (!is_int || !var_lb || !var_ub || !w)))
    return fail(PDFP_ERR_ARGUMENT, "null input array");
  if (index_base != 0 && index_base != 1) return fail(PDFP_ERR_ARGUMENT, "index_base must be 0 or 1");
  if (colptr[n] - index_base != nnz) return fail(PDFP_ERR_ARGUMENT, "colptr[n] does not match nnz");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PDFP_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(PDFP_ERR_ARGUMENT, "device index out of range");
  auto* ctx = new (std::nothrow) pdfp_batch();
  if (!ctx) return fail(PDFP_ERR_NOMEM, "host allocation failed");
  try {
    // the narrowest batch width that holds K trajectories
    if (K <= 8)
      ctx->impl.reset(new pdfp::Batch<8>(K, m, n, colptr, rowval, nzval, index_base, b, is_eq,
                                         is_int, var_lb, var_ub, w, device));
    else
      ctx->impl.reset(new pdfp::Batch<16>(K, m, n, colptr, rowval, nzval, index_base, b, is_eq,
                                          is_int, var_lb, var_ub, w, device));
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

int pdfp_batch_set_original_primal_residual(pdfp_batch* ctx, const double* D_r) {
  return batch_guarded(ctx, [&] {
    if (!D_r) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null D_r");
    ctx->impl->set_original_primal_residual(D_r);
  });
}

int pdfp_batch_set_x(pdfp_batch* ctx, int32_t k, const double* x) {
  return batch_guarded(ctx, [&] {
    if (!x) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null x");
    ctx->impl->set_x(k, x);
  });
}

int pdfp_batch_set_target(pdfp_batch* ctx, int32_t k, const double* xt) {
  return batch_guarded(ctx, [&] {
    if (!xt) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null target");
    ctx->impl->set_target(k, xt);
  });
}

int pdfp_batch_set_column_scale(pdfp_batch* ctx, const double* D_c) {
  return batch_guarded(ctx, [&] {
    if (!D_c) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null column scaling");
    ctx->impl->set_column_scale(D_c);
  });
}

int pdfp_batch_set_target_original(pdfp_batch* ctx, int32_t k, const double* xt) {
  return batch_guarded(ctx, [&] {
    if (!xt) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null target");
    ctx->impl->set_target_original(k, xt);
  });
}

int pdfp_batch_warm_start(pdfp_batch* ctx, int32_t k) {
  return batch_guarded(ctx, [&] { ctx->impl->warm_start(k); });
}

int pdfp_batch_solve(pdfp_batch* ctx, const pdfp_params* params, const double* primal_weight,
                     const uint8_t* active, int32_t* termination, int64_t* total_count) {
  return batch_guarded(ctx, [&] {
    if (!primal_weight) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null primal weights");
    pdfp_params p;
    if (params) p = *params;
    else pdfp_default_params(&p);
    ctx->impl->solve(p, primal_weight, active);
    for (int k = 0; k < ctx->impl->K; ++k) {
      if (termination) termination[k] = ctx->impl->tr[(size_t)k].termination;
      if (total_count) total_count[k] = ctx->impl->tr[(size_t)k].r.total_count;
    }
  });
}

int pdfp_batch_begin(pdfp_batch* ctx, int32_t k, const pdfp_params* params,
                     double primal_weight) {
  return batch_guarded(ctx, [&] {
    pdfp_params p;
    if (params) p = *params;
    else pdfp_default_params(&p);
    ctx->impl->begin(k, p, primal_weight);
  });
}

int pdfp_batch_run(pdfp_batch* ctx, int32_t until_any, int32_t* termination,
                   int64_t* total_count, int32_t* n_running) {
  return batch_guarded(ctx, [&] {
    if (until_any < 0 || until_any > 2)
      throw pdfp::Error(PDFP_ERR_ARGUMENT, "until_any must be 0, 1 or 2");
    const int left = ctx->impl->run(until_any);
    if (n_running) *n_running = left;
    for (int k = 0; k < ctx->impl->K; ++k) {
      if (termination) termination[k] = ctx->impl->tr[(size_t)k].termination;
      if (total_count) total_count[k] = ctx->impl->tr[(size_t)k].r.total_count;
    }
  });
}

int pdfp_batch_get_x(pdfp_batch* ctx, int32_t k, double* out) {
  return batch_guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->get_x(k, out);
  });
}

int pdfp_batch_get_vector(pdfp_batch* ctx, int32_t k, int32_t which, double* out) {
  return batch_guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->get_vector(k, which, out);
  });
}

int pdfp_batch_get_info(pdfp_batch* ctx, int32_t k, pdfp_info* out) {
  return batch_guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->info(k, out);
  });
}

int pdfp_batch_time_iterations(pdfp_batch* ctx, int32_t iterations, double* ms_out) {
  return batch_guarded(ctx, [&] {
    if (!ms_out || iterations < 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    *ms_out = ctx->impl->time_iterations(iterations);
  });
}

int pdfp_batch_time_kernels(pdfp_batch* ctx, int32_t iterations, double* ms_at, double* ms_a) {
  return batch_guarded(ctx, [&] {
    if (!ms_at || !ms_a || iterations < 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    ctx->impl->time_kernels(iterations, ms_at, ms_a);
  });
}

int pdfp_batch_set_schedule(pdfp_batch* ctx, int32_t npb_a, int32_t npb_at, int32_t persist,
                            int32_t lane_rows_a, int32_t lane_rows_at) {
  return batch_guarded(ctx, [&] {
    ctx->impl->set_schedule(npb_a, npb_at, persist, lane_rows_a, lane_rows_at);
  });
}

int pdfp_batch_set_kernel_path(pdfp_batch* ctx, int32_t path) {
  return batch_guarded(ctx, [&] { ctx->impl->set_path(path); });
}

int pdfp_batch_set_block_graph(pdfp_batch* ctx, int32_t on) {
  return batch_guarded(ctx, [&] {
    if (on != 0 && on != 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "block graph switch must be 0 or 1");
    ctx->impl->graph_on = on != 0;
  });
}

int pdfp_batch_set_column_slices(pdfp_batch* ctx, int32_t on) {
  return batch_guarded(ctx, [&] { ctx->impl->set_slices(on); });
}

int pdfp_batch_set_slice_min_len(pdfp_batch* ctx, int32_t min_len) {
  return batch_guarded(ctx, [&] { ctx->impl->set_slice_min_len(min_len); });
}

int pdfp_batch_set_nontemporal(pdfp_batch* ctx, int32_t on) {
  return batch_guarded(ctx, [&] { ctx->impl->set_nt(on); });
}

int pdfp_batch_set_cooperative(pdfp_batch* ctx, int32_t wg_per_cu) {
  return batch_guarded(ctx, [&] { ctx->impl->set_coop(wg_per_cu); });
}

void pdfp_batch_destroy(pdfp_batch* ctx) { delete ctx; }

}  // extern "C"
