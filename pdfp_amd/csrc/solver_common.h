// Host-side pieces shared by the projector and the root-LP solver: the
// restart rule and PID primal weight (src/pdlp.jl:167-205), the device CSR
// with its row-block schedule and the schedule autotuner, and the power
// iteration for ‖A‖₂ (src/preconditioner.jl:108-125).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pdfp_hip.h"

#include "engine.h"
#include "host_util.h"

namespace pdfp {

constexpr int kBlock = 256;  // threads per workgroup of the CSR passes
constexpr uint64_t kSvSeed = 20240517ull;
constexpr double kJuliaEps = 2.220446049250313e-16;
// Nonzeros staged per workgroup: the autotuner picks one per operator.
constexpr int kNpbChoices[] = {256, 512, 1024, 2048};
// Below this many nonzeros a pass is a few microseconds whatever its
// schedule, and timing the choices (a few milliseconds) would cost more than
// any of them saves: such operators keep NPB 512 and one workgroup per block.
constexpr int64_t kAutotuneMinNnz = 20000;

extern std::string g_create_error;  // message of the last failed *_create

inline double jmax(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return NAN;
  return a > b ? a : b;
}

inline int grid_for(int64_t n) { return (int)((n + 255) / 256); }

struct RestartState {
  double primal_weight = 1.0, best_primal_weight = 1.0;
  double best_primal_dual_residual_gap = INFINITY;
  double primal_weight_error_sum = 0.0, primal_weight_last_error = 0.0;
  double fixed_point_error = INFINITY, initial_fixed_point_error = INFINITY;
  double last_trial_fixed_point_error = INFINITY;
  int64_t inner_count = 0, total_count = 0;
  double relative_primal_residual = INFINITY, relative_dual_residual = INFINITY;
};

// should_restart  src/pdlp.jl:167-181
inline bool should_restart(RestartState& s, const pdfp_params& p) {
  bool restart = false;
  if (s.total_count == p.termination_evaluation_frequency) {
    restart = true;
  } else if (s.total_count > p.termination_evaluation_frequency) {
    if (s.fixed_point_error <= p.sufficient_reduction_for_restart * s.initial_fixed_point_error)
      restart = true;
    if (s.fixed_point_error <= p.necessary_reduction_for_restart * s.initial_fixed_point_error &&
        s.fixed_point_error > s.last_trial_fixed_point_error)
      restart = true;
    if ((double)s.inner_count >= p.artificial_restart_threshold * (double)s.total_count)
      restart = true;
  }
  s.last_trial_fixed_point_error = s.fixed_point_error;
  return restart;
}

// update_primal_weight!  src/pdlp.jl:185-205
inline void update_primal_weight(RestartState& s, const pdfp_params& rp, double primal_dist,
                                 double dual_dist) {
  const double ratio = s.relative_dual_residual / jmax(s.relative_primal_residual, kJuliaEps);
  if (1e-16 < primal_dist && primal_dist < 1e12 && 1e-16 < dual_dist && dual_dist < 1e12 &&
      1e-8 < ratio && ratio < 1e8) {
    const double err = std::log(dual_dist) - std::log(primal_dist) - std::log(s.primal_weight);
    s.primal_weight_error_sum *= rp.i_smooth;
    s.primal_weight_error_sum += err;
    const double delta_err = err - s.primal_weight_last_error;
    s.primal_weight *=
        std::exp(rp.k_p * err + rp.k_i * s.primal_weight_error_sum + rp.k_d * delta_err);
    s.primal_weight_last_error = err;
  } else {
    s.primal_weight = s.best_primal_weight;
    s.primal_weight_error_sum = 0.0;
    s.primal_weight_last_error = 0.0;
  }
  const double gap =
      std::fabs(std::log10(s.relative_dual_residual / jmax(s.relative_primal_residual, kJuliaEps)));
  if (gap < s.best_primal_dual_residual_gap) {
    s.best_primal_dual_residual_gap = gap;
    s.best_primal_weight = s.primal_weight;
  }
}

// SplitMix64 + Box–Muller normals: the fixed start of the power iteration
// (the reference draws randn(MersenneTwister(20240517), m); this stream is
// the library's own: u1 ∈ (0,1], u2 ∈ [0,1) from the top 53 bits of
// consecutive SplitMix64 outputs, r = √(-2 ln u1), (r cos 2πu2, r sin 2πu2)).
inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline void start_vector(int64_t m, uint64_t seed, double* out) {
  uint64_t s = seed;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < m; i += 2) {
    const double u1 = ((double)(splitmix64(s) >> 11) + 1.0) * 0x1.0p-53;
    const double u2 = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
    const double r = std::sqrt(-2.0 * std::log(u1));
    out[i] = r * std::cos(two_pi * u2);
    if (i + 1 < m) out[i + 1] = r * std::sin(two_pi * u2);
  }
}

// Device CSR + schedule of one operator (A or Aᵀ).
template <typename TV>
struct DevCsr {
  DevBuf<int32_t> rowptr, col;
  DevBuf<int4> sched;  // per-block {first row, first nonzero, lanes per row, mode}
  DevBuf<TV> val;
  int32_t nblocks = 0, nrows = 0;
  int32_t npb = 512;  // nonzeros per stream-mode block of the current schedule
  int32_t persist = 0;  // 0: one workgroup per block; k: k workgroups per CU, looping
  int32_t max_rows = kBlock;  // rows per block of the current schedule, at most
  int32_t lane_len = 0;  // blocks whose rows all hold at most this many nonzeros
                         // are reduced a lane per (row, value) (kLaneRows); 0: none
  int64_t nnz = 0;
  std::vector<int32_t> h_rowptr;  // kept for re-scheduling
  // Independent row segments (the instances of a block-diagonal pack): row
  // offsets, K + 1 entries, or empty for one segment.  No block of a
  // schedule crosses a segment boundary; seg_blocks[k] is the first block of
  // segment k (seg_blocks[K] = nblocks), so per-block partials sum per segment.
  std::vector<int64_t> segments;
  std::vector<int32_t> seg_blocks;

  void upload(const HostCsr& h, hipStream_t s) {
    nrows = (int32_t)h.nrows;
    nnz = (int64_t)h.col.size();
    h_rowptr = h.rowptr;
    rowptr.alloc(h.rowptr.size());
    rowptr.upload(h.rowptr.data(), h.rowptr.size(), s);
    col.alloc(std::max<size_t>(h.col.size(), 1));
    col.upload(h.col.data(), h.col.size(), s);
    std::vector<TV> v(h.val.size());
    for (size_t k = 0; k < v.size(); ++k) v[k] = (TV)h.val[k];
    val.alloc(std::max<size_t>(v.size(), 1));
    val.upload(v.data(), v.size(), s);
    PDFP_HIP(hipStreamSynchronize(s));  // the host vectors die with this scope
    schedule(npb, s);
  }

  // Row blocks of at most `rows` rows and `max_nnz` nonzeros, with their
  // lanes per row.  `lane_rows` > 0 (lane epilogues over raw staging only)
  // marks the stream-mode blocks whose rows all hold at most that many
  // nonzeros for the lane-per-(row, value) reduction.
  void schedule(int max_nnz, hipStream_t s, int rows = kBlock, int lane_rows = 0) {
    npb = max_nnz;
    max_rows = rows;
    lane_len = lane_rows;
    std::vector<int32_t> blk;
    if (segments.empty()) {
      blk = row_blocks(h_rowptr, rows, max_nnz);
    } else {
      blk.assign(1, (int32_t)segments[0]);
      seg_blocks.assign(1, 0);
      for (size_t k = 0; k + 1 < segments.size(); ++k) {
        row_blocks_range(h_rowptr, segments[k], segments[k + 1], rows, max_nnz, blk);
        seg_blocks.push_back((int32_t)blk.size() - 1);
      }
    }
    std::vector<uint8_t> lanes = block_lanes(h_rowptr, blk, kBlock, max_nnz);
    nblocks = (int32_t)blk.size() - 1;
    std::vector<int4> rec(blk.size());
    for (size_t k = 0; k < blk.size(); ++k) {
      rec[k].x = blk[k];
      rec[k].y = h_rowptr[(size_t)blk[k]];
      rec[k].z = k < lanes.size() ? (int)lanes[k] : 0;
      rec[k].w = 0;
      if (lane_rows > 0 && k < lanes.size() && lanes[k] > 0) {
        int32_t longest = 0;
        for (int32_t r = blk[k]; r < blk[k + 1]; ++r)
          longest = std::max(longest, h_rowptr[(size_t)r + 1] - h_rowptr[(size_t)r]);
        if (longest <= lane_rows) rec[k].w = kLaneRows;
      }
    }
    sched.alloc(rec.size());
    sched.upload(rec.data(), rec.size(), s);
    PDFP_HIP(hipStreamSynchronize(s));
  }

  Csr<TV> view() const { return Csr<TV>{rowptr.p, col.p, val.p, sched.p, nblocks, nrows}; }
  size_t bytes() const { return rowptr.bytes() + col.bytes() + sched.bytes() + val.bytes(); }
};

// Launch one pass of a given NPB: one workgroup per row block, or (when the
// operator's `persist` > 0) a persistent grid of persist × 256 workgroups
// (256 CUs) that walks the blocks.  With events t0/t1 the launch records the
// kernel's own start and end (hipExtLaunchKernel), which times the kernel
// alone, without the gaps that events recorded around it on the stream add;
// either may be null (a timed span of several launches).
template <int NPB, int NACC, int NMAX, typename TV, class G, class E>
inline void launch_pass_npb(const DevCsr<TV>& A, G g, E e, double* partials, hipStream_t s,
                            hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr) {
  const int grid = A.persist * 256;
  const bool persist = A.persist > 0 && grid < A.nblocks;
  if (t0 || t1) {
    if (persist)
      hipExtLaunchKernelGGL(csr_pass_persistent<kBlock, NPB, TV, G, E, NACC, NMAX>, dim3(grid),
                            dim3(kBlock), 0, s, t0, t1, 0, A.view(), g, e, partials);
    else
      hipExtLaunchKernelGGL(csr_pass<kBlock, NPB, TV, G, E, NACC, NMAX>, dim3(A.nblocks),
                            dim3(kBlock), 0, s, t0, t1, 0, A.view(), g, e, partials);
    PDFP_HIP(hipGetLastError());
  } else if (persist) {
    csr_pass_persistent<kBlock, NPB, TV, G, E, NACC, NMAX>
        <<<grid, kBlock, 0, s>>>(A.view(), g, e, partials);
  } else {
    csr_pass<kBlock, NPB, TV, G, E, NACC, NMAX>
        <<<A.nblocks, kBlock, 0, s>>>(A.view(), g, e, partials);
  }
}

// Launch one CSR pass with the operator's schedule (no-op if it is empty).
template <int NACC, int NMAX = 0, typename TV, class G, class E>
inline void launch_pass(const DevCsr<TV>& A, G g, E e, double* partials, hipStream_t s,
                        hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr) {
  if (A.nblocks == 0) {
    if (t0) {
      PDFP_HIP(hipEventRecord(t0, s));
      PDFP_HIP(hipEventRecord(t1, s));
    }
    return;
  }
  switch (A.npb) {
    case 256: launch_pass_npb<256, NACC, NMAX>(A, g, e, partials, s, t0, t1); break;
    case 512: launch_pass_npb<512, NACC, NMAX>(A, g, e, partials, s, t0, t1); break;
    case 1024: launch_pass_npb<1024, NACC, NMAX>(A, g, e, partials, s, t0, t1); break;
    default: launch_pass_npb<2048, NACC, NMAX>(A, g, e, partials, s, t0, t1); break;
  }
}

// Power-iteration helper kernels (src/preconditioner.jl:113-123); static so
// that each translation unit carries its own copy.
static __global__ __launch_bounds__(256) void k_scale_div(int m, double* __restrict__ out,
                                                          const double* __restrict__ in,
                                                          double nrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = in[i] / nrm;
}

// out = in / √(*sumsq): the norm is read from device memory, so that power
// iterations run back to back without a host round trip.
static __global__ __launch_bounds__(256) void k_scale_div_dev(int m, double* __restrict__ out,
                                                              const double* __restrict__ in,
                                                              const double* __restrict__ sumsq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = in[i] / sqrt(sumsq[0]);
}

// hist[0..3) = {σ², ‖nxt‖², ‖eig‖²} of one power iteration
static __global__ void k_pi_record(const double* __restrict__ sums, double* __restrict__ hist) {
  if (threadIdx.x == 0) {
    hist[0] = sums[4];
    hist[1] = sums[0];
    hist[2] = sums[1];
  }
}

// nxt = eig - σ² nxt;  acc0 = ‖nxt‖², acc1 = ‖eig‖²  (σ² read from device)
static __global__ __launch_bounds__(256) void k_pi_resid(int m, double* __restrict__ nxt,
                                                         const double* __restrict__ eig,
                                                         const double* __restrict__ sigma_sq,
                                                         double* __restrict__ partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < m) {
    const double e = eig[i];
    const double v = e - sigma_sq[0] * nxt[i];
    nxt[i] = v;
    acc[0] = v * v;
    acc[1] = e * e;
  }
  block_store_partials<256, 2>(acc, partials + (size_t)blockIdx.x * 2);
}

static __global__ __launch_bounds__(256) void k_sumsq(int m, const double* __restrict__ v,
                                                      double* __restrict__ partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < m) acc[0] = v[i] * v[i];
  block_store_partials<256, 1>(acc, partials + blockIdx.x);
}

// out[row] = s[0]
template <typename TO>
struct EpiStore {
  using Pre = NoPre;
  TO* __restrict__ out;
  __device__ __forceinline__ Pre pre(int) const { return {}; }
  __device__ __forceinline__ void operator()(int r, const Pre&, const double* s, double*) const {
    out[r] = (TO)s[0];
  }
};

// out[row] = s[0];  acc[0] += v[row] * s[0]   (power iteration: eig = A d, ⟨nxt, eig⟩)
struct EpiStoreDot {
  struct Pre {
    double v;
  };
  double* __restrict__ out;
  const double* __restrict__ v;
  __device__ __forceinline__ Pre pre(int r) const { return {v[r]}; }
  __device__ __forceinline__ void operator()(int r, const Pre& p, const double* s,
                                             double* acc) const {
    out[r] = s[0];
    acc[0] += p.v * s[0];
  }
};

// Schedule autotuner: for each nonzeros-per-block choice, rebuild the
// schedule and time a plain pass y = M x (FP64 out) with HIP events; keep
// the fastest.  Returns the chosen value.
template <typename TV>
int autotune_schedule(DevCsr<TV>& M, int64_t ncols, hipStream_t s) {
  if (M.nblocks == 0 || M.nnz == 0) return M.npb;
  if (M.nnz < kAutotuneMinNnz) {
    M.schedule(512, s);
    M.persist = 0;
    return 512;
  }
  DevBuf<TV> x, y;
  x.alloc((size_t)std::max<int64_t>(ncols, 1));
  y.alloc((size_t)std::max<int32_t>(M.nrows, 1));
  std::vector<TV> ones((size_t)std::max<int64_t>(ncols, 1), (TV)1);
  x.upload(ones.data(), ones.size(), s);
  hipEvent_t e0, e1;
  PDFP_HIP(hipEventCreate(&e0));
  PDFP_HIP(hipEventCreate(&e1));
  int best = M.npb, best_persist = 0;
  float best_ms = 1e30f;
  constexpr int kPersist[] = {0, 2, 4, 8};
  for (int npb : kNpbChoices) {
    M.schedule(npb, s);
    for (int persist : kPersist) {
      M.persist = persist;
      for (int w = 0; w < 3; ++w)
        launch_pass<0>(M, Gather1<TV>{x.p}, EpiStore<TV>{y.p}, nullptr, s);
      PDFP_HIP(hipEventRecord(e0, s));
      constexpr int kReps = 10;
      for (int r = 0; r < kReps; ++r)
        launch_pass<0>(M, Gather1<TV>{x.p}, EpiStore<TV>{y.p}, nullptr, s);
      PDFP_HIP(hipEventRecord(e1, s));
      PDFP_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      PDFP_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best_ms) {
        best_ms = ms;
        best = npb;
        best_persist = persist;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  M.schedule(best, s);
  M.persist = best_persist;
  return best;
}

// estimate_max_singular_value(A, A_t, m; max_iterations, tolerance)
// src/preconditioner.jl:108-125, on the device.
template <typename TV>
double power_iteration(const DevCsr<TV>& A, const DevCsr<TV>& At, int64_t m, int64_t n,
                       int max_iterations, double tolerance, hipStream_t stream) {
  if (m == 0) return 0.0;
  DevBuf<double> eig, nxt, dual, part, sums;
  PinnedBuf<double> h;
  eig.alloc((size_t)m);
  nxt.alloc((size_t)m);
  dual.alloc((size_t)std::max<int64_t>(n, 1));
  const int gm = grid_for(m);
  part.alloc((size_t)std::max(gm * 2, std::max(A.nblocks, 1)) + 8);
  sums.alloc(8);
  h.alloc(8);
  std::vector<double> start((size_t)m);
  start_vector(m, kSvSeed, start.data());
  eig.upload(start.data(), (size_t)m, stream);
  // Iterations run in chunks with no host synchronisation inside a chunk:
  // each records {σ², ‖residual‖², ‖eig‖²} on the device, and the host then
  // replays the reference's stopping rule over the records, so the result is
  // the one the iteration-by-iteration loop returns (iterations past the
  // stopping point are computed and discarded).
  constexpr int kMaxChunk = 64;
  DevBuf<double> hist;
  PinnedBuf<double> h_hist;
  hist.alloc(3 * kMaxChunk);
  h_hist.alloc(3 * kMaxChunk);
  k_sumsq<<<gm, 256, 0, stream>>>((int)m, eig.p, part.p);
  finalize_partials<256><<<1, 256, 0, stream>>>(part.p, gm, 1, 0, sums.p + 1);
  PDFP_HIP(hipMemcpyAsync(h.p, sums.p, sizeof(double) * 8, hipMemcpyDeviceToHost, stream));
  PDFP_HIP(hipStreamSynchronize(stream));
  double nrm = std::sqrt(h.p[1]);
  double sigma_sq = 1.0;
  int done = 0, chunk = 4;
  bool stop = false;
  while (!stop && done < max_iterations) {
    const int c = std::min(chunk, max_iterations - done);
    for (int t = 0; t < c; ++t) {
      k_scale_div_dev<<<gm, 256, 0, stream>>>((int)m, nxt.p, eig.p, sums.p + 1);
      launch_pass<0>(At, Gather1<double>{nxt.p}, EpiStore<double>{dual.p}, nullptr, stream);
      launch_pass<1>(A, Gather1<double>{dual.p}, EpiStoreDot{eig.p, nxt.p}, part.p, stream);
      finalize_partials<256><<<1, 256, 0, stream>>>(part.p, A.nblocks, 1, 0, sums.p + 4);
      k_pi_resid<<<gm, 256, 0, stream>>>((int)m, nxt.p, eig.p, sums.p + 4, part.p);
      finalize_partials<256><<<1, 256, 0, stream>>>(part.p, gm, 2, 0, sums.p);
      k_pi_record<<<1, 64, 0, stream>>>(sums.p, hist.p + 3 * t);
    }
    PDFP_HIP(hipMemcpyAsync(h_hist.p, hist.p, sizeof(double) * 3 * (size_t)c,
                            hipMemcpyDeviceToHost, stream));
    PDFP_HIP(hipStreamSynchronize(stream));
    for (int t = 0; t < c && !stop; ++t) {
      if (nrm == 0.0) {
        stop = true;
        break;
      }
      sigma_sq = h_hist.p[3 * t];
      if (std::sqrt(h_hist.p[3 * t + 1]) < tolerance) stop = true;
      nrm = std::sqrt(h_hist.p[3 * t + 2]);
    }
    done += c;
    chunk = std::min(2 * chunk, kMaxChunk);
  }
  return std::sqrt(jmax(sigma_sq, 0.0));
}

}  // namespace pdfp
