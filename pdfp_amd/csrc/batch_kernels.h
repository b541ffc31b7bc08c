// Shared-matrix batch of up to KB projector trajectories (FP32 iterates);
// the batch width KB is 8 or 16.
//
// K ≤ KB trajectories of one instance — the same A, b, bounds and weights,
// each with its own target x̃_k, iterates, primal weight and restart state —
// run their PDHG iterations together: every vector is stored interleaved as
// v[j·KB + k], so a nonzero's gather fetches the KB trajectories' entries of
// column j in one request of 4·KB bytes (32 or 64 of the 128-byte line a
// random gather fills anyway) and the matrix is streamed once for all of
// them.  The per-trajectory scalars (step, primal weight, inner count,
// active flag) live in device memory; an inactive trajectory's iterates are
// left untouched, so it keeps the result of its own last solve.
#pragma once

#include "engine.h"
#include "pdfp_kernels.h"

namespace pdfp {

// Per-trajectory scalars of the current evaluation block (FP64, formed on
// the host with the same operations as step_scalars: τ = step/pw,
// σ = step·pw, 1 - ω), and a table of the Halpern weights of the block's
// iterations: wt[k_offset·KB + q] = {w_k, 1 - w_k}, w_k = k/(k+1),
// k = inner_count_q + k_offset (src/pdfp.jl:120-122).  No division is left
// for the epilogues.
struct BatchStep {
  double tau;
  double sigma;
  double omega;
  double omw;
  double active;  // 1: iterate this trajectory, 0: leave it unchanged
  double pad[3];
};

template <int KB>
__device__ __forceinline__ StepScalars batch_scalars(const BatchStep& sp,
                                                     const double2* __restrict__ wt,
                                                     int k_offset, int q) {
  StepScalars s;
  s.tau = sp.tau;
  s.sigma = sp.sigma;
  s.omega = sp.omega;
  s.omw = sp.omw;
  const double2 w = wt[(size_t)k_offset * KB + q];
  s.weight = w.x;
  s.omwt = w.y;
  return s;
}

// p_k = a · x[c·KB + k]  for the KB trajectories (one gather of 4·KB
// bytes); the pass stages the FP32 values (raw) rather than FP64 products.
template <int KB>
struct GatherBatch {
  static constexpr int kBatch = KB;
  static constexpr int NV = KB;
  static constexpr bool kRawStage = true;
  const float* __restrict__ x;
  __device__ __forceinline__ void operator()(int c, double a, double* p) const {
    const float4* v = reinterpret_cast<const float4*>(x + (size_t)c * kBatch);
#pragma unroll
    for (int h = 0; h < KB / 4; ++h) {
      const float4 f = v[h];
      p[4 * h + 0] = a * (double)f.x;
      p[4 * h + 1] = a * (double)f.y;
      p[4 * h + 2] = a * (double)f.z;
      p[4 * h + 3] = a * (double)f.w;
    }
  }
  __device__ __forceinline__ float4 raw4(int c, int h) const {
    return reinterpret_cast<const float4*>(x + (size_t)c * kBatch)[h];
  }
};

// Trajectory q's step scalars of the launch and its active flag.  The
// epilogues load them together with nothing else in flight behind them and
// test the flag only before their stores, so the loads go out as one batch.
struct BatchScalars {
  StepScalars c;
  double active;  // 1: iterate this trajectory, 0: leave it unchanged
};

template <int KB>
__device__ __forceinline__ BatchScalars load_batch_scalars(const BatchStep* __restrict__ sp,
                                                           const double2* __restrict__ wt,
                                                           int k_offset, int q) {
  const BatchStep& st = sp[q];
  return {batch_scalars<KB>(st, wt, k_offset, q), st.active};
}

// Loads of an epilogue's per-trajectory operands, which a pass reads once:
// plain, or non-temporal (NT), which keeps them from displacing the gathered
// vector's lines in L2.  Non-temporal loads measured 5 % faster per batched
// iteration on the 5M-nonzero power-law workload and 8 % slower on the
// 1M-nonzero one, whose iterates stay in the memory-side cache between
// passes; the autotuner picks per instance (the NT epilogues are the ones
// named ...NT below).
template <bool NT, typename T>
__device__ __forceinline__ T ld_epi(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// Column j, trajectory q (lane q of the row's KB): the fused primal and dual-y
// step (src/pdfp.jl:119-143) with s = (Aᵀ z_q)_j.  An inactive trajectory
// is left untouched.  MAJOR steps also store pdhg x, the dual slack, pdhg y
// and reflected y.
template <int KB, bool MAJOR, bool NT>
struct EpiBatchPrimalYT {
  static constexpr int kBatch = KB;
  static constexpr int kLanes = KB;
  struct Pre {
    float x, y, x0, y0, xt, lb, ub, w;
  };
  const BatchStep* sp;
  const double2* wt;
  int k_offset;
  float* __restrict__ x;
  const float* __restrict__ x0;
  float* __restrict__ y;
  const float* __restrict__ y0;
  const float* __restrict__ xt;
  const float* __restrict__ lb;
  const float* __restrict__ ub;
  const float* __restrict__ w;
  float* __restrict__ rx;
  float* __restrict__ px;
  float* __restrict__ ds;
  float* __restrict__ py;
  float* __restrict__ ry;
  __device__ __forceinline__ Pre pre(int j, int q) const {
    const size_t o = (size_t)j * kBatch + q;
    return {ld_epi<NT>(x + o), ld_epi<NT>(y + o), ld_epi<NT>(x0 + o), ld_epi<NT>(y0 + o),
            ld_epi<NT>(xt + o), lb[j], ub[j], w[j]};
  }
  __device__ __forceinline__ void operator()(int j, int q, const Pre& p, double s,
                                             double*) const {
    const BatchScalars sh = load_batch_scalars<KB>(sp, wt, k_offset, q);
    const size_t o = (size_t)j * kBatch + q;
    const PrimalYOut r = primal_y_math(sh.c, s, (double)p.x, (double)p.y, (double)p.x0,
                                       (double)p.y0, (double)p.lb, (double)p.ub, (double)p.xt,
                                       (double)p.w);
    if (sh.active == 0.0) return;
    // non-temporal: nothing reads these again in this pass, and the next one
    // starts after the kernel boundary's cache write-back (measured 1-2% faster)
    __builtin_nontemporal_store((float)r.x, &x[o]);
    __builtin_nontemporal_store((float)r.y, &y[o]);
    __builtin_nontemporal_store((float)r.rx, &rx[o]);
    if (MAJOR) {
      __builtin_nontemporal_store((float)r.px, &px[o]);
      __builtin_nontemporal_store((float)r.ds, &ds[o]);
      __builtin_nontemporal_store((float)r.py, &py[o]);
      __builtin_nontemporal_store((float)r.ry, &ry[o]);
    }
  }
};

template <int KB, bool MAJOR>
struct EpiBatchPrimalY : EpiBatchPrimalYT<KB, MAJOR, false> {};
template <int KB, bool MAJOR>
struct EpiBatchPrimalYNT : EpiBatchPrimalYT<KB, MAJOR, true> {};

// Row i, trajectory q: the dual-z step (src/pdfp.jl:146-156) with
// s = (A x̄_q)_i.
template <int KB, bool MAJOR, bool NT>
struct EpiBatchDualZT {
  static constexpr int kBatch = KB;
  static constexpr int kLanes = KB;
  struct Pre {
    float z, z0, b;
    uint32_t eq;
  };
  const BatchStep* sp;
  const double2* wt;
  int k_offset;
  float* __restrict__ z;
  const float* __restrict__ z0;
  const float* __restrict__ b;
  const uint8_t* __restrict__ is_eq;
  float* __restrict__ pz;
  float* __restrict__ rz;
  __device__ __forceinline__ Pre pre(int i, int q) const {
    const size_t o = (size_t)i * kBatch + q;
    return {ld_epi<NT>(z + o), ld_epi<NT>(z0 + o), b[i], (uint32_t)is_eq[i]};
  }
  __device__ __forceinline__ void operator()(int i, int q, const Pre& p, double s,
                                             double*) const {
    const BatchScalars sh = load_batch_scalars<KB>(sp, wt, k_offset, q);
    const size_t o = (size_t)i * kBatch + q;
    const DualZOut r = dual_z_math(sh.c, s, (double)p.z, (double)p.z0, (double)p.b, p.eq != 0);
    if (sh.active == 0.0) return;
    __builtin_nontemporal_store((float)r.z, &z[o]);
    if (MAJOR) {
      __builtin_nontemporal_store((float)r.pz, &pz[o]);
      __builtin_nontemporal_store((float)r.rz, &rz[o]);
    }
  }
};

template <int KB, bool MAJOR>
struct EpiBatchDualZ : EpiBatchDualZT<KB, MAJOR, false> {};
template <int KB, bool MAJOR>
struct EpiBatchDualZNT : EpiBatchDualZT<KB, MAJOR, true> {};

// ---- the A pass over column slices -------------------------------------------
//
// A x̄ as 8 partial products: the matrix stacked from A's 8 column slices
// (slice s: A's rows restricted to columns [s·n/8, (s+1)·n/8), only the rows
// with a nonzero there), each slice walked by one XCD
// (csr_pass_xcd_segments), so that its gathers of x̄ come from the eighth of
// x̄ that fits in that XCD's L2 instead of filling a line from memory per
// nonzero.  When only the long rows are sliced, the stacked rows of a slice
// are ordered by length, so that a block's rows are alike and its lanes (one
// per row and trajectory) finish together; otherwise they keep row order.
// Stacked row r (row i of A in slice s) stores its KB sums (FP64
// over the slice's nonzeros in ascending column order), rounded to FP32, at
// part[dense[r]·KB ...]; the (slice, row) pairs with no nonzero keep the
// zeros they were allocated with.  k_batch_slice_dualz (or, when only the
// long rows are sliced, EpiBatchDualZLong) adds row i's 8 partials in slice
// order in FP64 and runs the z update.

// Stacked row r's KB sums, by the row's lane (blocks of up to 256 rows:
// every row sliced, so the slices' rows are short and many).
template <int KB>
struct EpiSlicePartial {
  struct Pre {
    int32_t d;
  };
  const int32_t* __restrict__ dense;
  float* __restrict__ part;
  __device__ __forceinline__ Pre pre(int r) const { return {dense[r]}; }
  __device__ __forceinline__ void operator()(int, const Pre& p, const double* s, double*) const {
    float4* o = reinterpret_cast<float4*>(part + (size_t)p.d * KB);
#pragma unroll
    for (int h = 0; h < KB / 4; ++h)
      o[h] = make_float4((float)s[4 * h], (float)s[4 * h + 1], (float)s[4 * h + 2],
                         (float)s[4 * h + 3]);
  }
};

// The same stores by lane (r, q), one per trajectory (blocks of up to
// 256/KB rows, each reduced one lane per (row, trajectory): only the long
// rows sliced, whose slices' rows are longer and fewer).
template <int KB>
struct EpiSlicePartialLanes {
  static constexpr int kLanes = KB;
  struct Pre {
    int32_t d;
  };
  const int32_t* __restrict__ dense;
  float* __restrict__ part;
  __device__ __forceinline__ Pre pre(int r, int) const { return {dense[r]}; }
  __device__ __forceinline__ void operator()(int, int q, const Pre& p, double s, double*) const {
    part[(size_t)p.d * KB + q] = (float)s;
  }
};

// Row i, trajectory q of the combine: the 8 partials in slice order in
// FP64, then the z update.
template <int KB, class EZ>
__device__ __forceinline__ void slice_dualz(int m, const float* __restrict__ part, const EZ& ez) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)m * KB) return;
  const int i = (int)(e / KB), q = (int)(e % KB);
  const typename EZ::Pre p = ez.pre(i, q);
  float v[8];
#pragma unroll
  for (int sl = 0; sl < 8; ++sl) v[sl] = part[(size_t)sl * (size_t)m * KB + (size_t)e];
  double sum = 0.0;
#pragma unroll
  for (int sl = 0; sl < 8; ++sl) sum += (double)v[sl];
  ez(i, q, p, sum, nullptr);
}

template <int KB, bool MAJOR>
__global__ __launch_bounds__(256) void k_batch_slice_dualz(int m, const float* __restrict__ part,
                                                           EpiBatchDualZ<KB, MAJOR> ez) {
  slice_dualz<KB>(m, part, ez);
}
template <int KB, bool MAJOR>
__global__ __launch_bounds__(256) void k_batch_slice_dualz_nt(int m,
                                                              const float* __restrict__ part,
                                                              EpiBatchDualZNT<KB, MAJOR> ez) {
  slice_dualz<KB>(m, part, ez);
}

// The A pass when only A's long rows run over the column slices: the short
// rows are summed in one pass over A with the long rows emptied, and this
// epilogue adds a long row's 8 slice partials (part[(s·nlong + l)·KB + q],
// l = long_idx[i] ≥ 0; -1 marks a short row) in slice order in FP64 to the
// pass's sum, which is +0.0 for an emptied row, so the sum is formed as
// k_batch_slice_dualz forms it; then the z update.  On power-law rows most
// nonzeros sit in a few long rows: their gathers come from the XCDs' L2s,
// while the many short rows keep one pass and no partials.
template <int KB, bool MAJOR, bool NT>
struct EpiBatchDualZLongT {
  static constexpr int kBatch = KB;
  static constexpr int kLanes = KB;
  using Z = EpiBatchDualZT<KB, MAJOR, NT>;
  struct Pre {
    typename Z::Pre z;
    int32_t l;
  };
  Z ez;
  const int32_t* __restrict__ long_idx;
  const float* __restrict__ part;
  int32_t nlong;
  __device__ __forceinline__ Pre pre(int i, int q) const { return {ez.pre(i, q), long_idx[i]}; }
  __device__ __forceinline__ void operator()(int i, int q, const Pre& p, double s,
                                             double* acc) const {
    if (p.l >= 0) {
      float v[8];
#pragma unroll
      for (int sl = 0; sl < 8; ++sl)
        v[sl] = part[((size_t)sl * (size_t)nlong + (size_t)p.l) * KB + q];
#pragma unroll
      for (int sl = 0; sl < 8; ++sl) s += (double)v[sl];
    }
    ez(i, q, p.z, s, acc);
  }
};

template <int KB, bool MAJOR>
struct EpiBatchDualZLong : EpiBatchDualZLongT<KB, MAJOR, false> {};
template <int KB, bool MAJOR>
struct EpiBatchDualZLongNT : EpiBatchDualZLongT<KB, MAJOR, true> {};

// ---- the evaluation terms of trajectory k (views of the interleaved vectors) --

// The terms of EpiEvalCol (pdfp_kernels.h) for trajectory k.
template <int KB>
struct EpiBatchEvalCol {
  static constexpr int kBatch = KB;
  struct Pre {
    float px, rx, py, ry, ds, xt, w, x0, y0;
  };
  const float *px, *rx, *py, *ry, *ds, *xt, *w, *x0, *y0;
  int k;
  __device__ __forceinline__ Pre pre(int j) const {
    const size_t o = (size_t)j * kBatch + k;
    return {px[o], rx[o], py[o], ry[o], ds[o], xt[o], w[j], x0[o], y0[o]};
  }
  __device__ __forceinline__ void operator()(int, const Pre& p, const double* s,
                                             double* acc) const {
    const double pxj = (double)p.px, pyj = (double)p.py;
    const double dx = pxj - (double)p.rx;
    const double dy = pyj - (double)p.ry;
    acc[0] += dx * dx;
    acc[1] += dy * dy;
    acc[2] += (s[1] - dy) * dx;
    const double r = s[0] - pyj - (double)p.ds;
    acc[3] += r * r;
    const double wj = (double)p.w;
    if (wj > 0.0) {
      const double xtj = (double)p.xt;
      acc[4] += wj * fabs(pxj - xtj);
      acc[5] += pyj * (xtj - pxj);
    }
    const double ex = (double)p.x0 - pxj;
    const double ey = (double)p.y0 - pyj;
    acc[6] += ex * ex;
    acc[7] += ey * ey;
  }
};

// The terms of EpiEvalRow (pdfp_kernels.h) for trajectory k.
template <int KB>
struct EpiBatchEvalRow {
  static constexpr int kBatch = KB;
  struct Pre {
    float pz, rz, b, z0;
    uint8_t eq;
    double ru;
  };
  const float *pz, *rz, *b, *z0;
  const uint8_t* is_eq;
  const double* row_unscale;
  int k;
  __device__ __forceinline__ Pre pre(int i) const {
    const size_t o = (size_t)i * kBatch + k;
    return {pz[o], rz[o], b[i], z0[o], is_eq[i], row_unscale ? row_unscale[i] : 1.0};
  }
  __device__ __forceinline__ void operator()(int, const Pre& p, const double* s,
                                             double* acc) const {
    const double pzi = (double)p.pz;
    const double dz = pzi - (double)p.rz;
    acc[0] += dz * dz;
    const double d = s[0] - (double)p.b;
    double r = p.eq ? fabs(d) : (d > 0.0 ? d : 0.0);
    if (row_unscale) r *= p.ru;
    acc[1] += r * r;
    acc[2] += pzi * d;
    const double ez = (double)p.z0 - pzi;
    acc[3] += ez * ez;
  }
};

// The row terms of every trajectory in the mask (*mask, a device word the
// host writes before the launch, so a captured launch serves any mask) as a
// lane epilogue of the CSR pass over A (gathering the KB trajectories' pdhg
// x): lane (row, q) adds trajectory q's terms; the pass writes
// per-trajectory partials.
template <int KB>
struct EpiBatchEvalRowLanes {
  static constexpr int kLanes = KB;
  using Pre = typename EpiBatchEvalRow<KB>::Pre;
  EpiBatchEvalRow<KB> er;
  const int32_t* mask;
  __device__ __forceinline__ Pre pre(int i, int q) const {
    EpiBatchEvalRow<KB> e = er;
    e.k = q;
    return e.pre(i);
  }
  __device__ __forceinline__ void operator()(int i, int q, const Pre& p, double s,
                                             double* acc) const {
    if (!((*mask >> q) & 1)) return;
    EpiBatchEvalRow<KB> e = er;
    e.k = q;
    e(i, p, &s, acc);
  }
};

// ---- evaluation of several trajectories in one pass ---------------------------
// The column terms: one workgroup serves 256/KB consecutive columns (rows of
// Aᵀ); lane (column, q) = (tid / KB, tid % KB) forms trajectory q's two sums over
// the column in FP64, in ascending row order (the second over the difference
// pz - rz, formed first as the reference does), and adds q's terms
// (EpiBatchEvalCol) when bit q of *mask is set.  Each workgroup writes
// per-trajectory partials partial[(q·NT + t)·nblocks + block]; the KB
// trajectories share every gather.  (The row terms run as a lane
// epilogue of the CSR pass over A, EpiBatchEvalRowLanes, whose schedule
// spreads long rows over the workgroup; columns are short in the pump's
// instances, and a dense column costs one lane loop here.)
template <int KB>
__global__ __launch_bounds__(256) void k_batch_eval_col(
    int nrows, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ pz, const float* __restrict__ rz,
    EpiBatchEvalCol<KB> ec, const int32_t* __restrict__ maskp, double* __restrict__ partial) {
  constexpr int kBatch = KB;
  constexpr int kEvalRows = 256 / KB;
  const int mask = *maskp;
  const int q = threadIdx.x % kBatch;
  const int j = blockIdx.x * kEvalRows + threadIdx.x / kBatch;
  double acc[kEvalColTerms];
#pragma unroll
  for (int t = 0; t < kEvalColTerms; ++t) acc[t] = 0.0;
  if (j < nrows && ((mask >> q) & 1)) {
    double s[2] = {0.0, 0.0};
    for (int k = ptr[j]; k < ptr[j + 1]; ++k) {
      const double a = (double)val[k];
      const size_t o = (size_t)col[k] * kBatch + q;
      const double zc = (double)pz[o];
      s[0] += a * zc;
      s[1] += a * (zc - (double)rz[o]);
    }
    ec.k = q;
    ec(j, ec.pre(j), s, acc);
  }
  lane_store_partials<256, kBatch, kEvalColTerms>(acc, partial, gridDim.x, blockIdx.x);
}

// out[q·32 + offset + t] = Σ_b partial[(q·nt + t)·nb + b] in block order,
// for every trajectory q in *maskp: one workgroup per (q, t), reading its
// series contiguously.
static __global__ __launch_bounds__(256) void k_batch_eval_finalize(
    const double* __restrict__ partial, int nb, int nt, const int32_t* __restrict__ maskp,
    int offset, double* __restrict__ out) {
  __shared__ double red[256 / kWave];
  const int q = blockIdx.x / nt, t = blockIdx.x % nt;
  if (!((*maskp >> q) & 1)) return;  // uniform over the workgroup
  const double* __restrict__ series = partial + (size_t)blockIdx.x * nb;
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) v += series[b];
  v = block_sum_to0<256>(v, red);
  if (threadIdx.x == 0) out[q * 32 + offset + t] = v;
}

// perform_restart! for trajectory k: initial ← pdhg, current ← pdhg.
template <int KB>
__global__ __launch_bounds__(256) void k_batch_restart(int n, int m, int k,
                                                              const float* __restrict__ px,
                                                              const float* __restrict__ py,
                                                              const float* __restrict__ pz,
                                                              float* __restrict__ x,
                                                              float* __restrict__ x0,
                                                              float* __restrict__ y,
                                                              float* __restrict__ y0,
                                                              float* __restrict__ z,
                                                              float* __restrict__ z0) {
  constexpr int kBatch = KB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const size_t o = (size_t)i * kBatch + k;
    x[o] = px[o];
    x0[o] = px[o];
    y[o] = py[o];
    y0[o] = py[o];
  }
  if (i < m) {
    const size_t o = (size_t)i * kBatch + k;
    z[o] = pz[o];
    z0[o] = pz[o];
  }
}

// warm_start! for trajectory k: initial ← current.
template <int KB>
__global__ __launch_bounds__(256) void k_batch_warm(int n, int m, int k,
                                                           const float* __restrict__ x,
                                                           const float* __restrict__ y,
                                                           const float* __restrict__ z,
                                                           float* __restrict__ x0,
                                                           float* __restrict__ y0,
                                                           float* __restrict__ z0) {
  constexpr int kBatch = KB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const size_t o = (size_t)i * kBatch + k;
    x0[o] = x[o];
    y0[o] = y[o];
  }
  if (i < m) {
    const size_t o = (size_t)i * kBatch + k;
    z0[o] = z[o];
  }
}

// Writes column k of an interleaved vector from a dense FP32 vector.
template <int KB>
__global__ __launch_bounds__(256) void k_batch_scatter(int len, int k,
                                                             const float* __restrict__ in,
                                                             float* __restrict__ out) {
  constexpr int kBatch = KB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[(size_t)i * kBatch + k] = in[i];
}

// Writes trajectory k's FP64 target, narrowed to FP32, into the interleaved
// vector (the same rounding as a host conversion).
template <int KB>
__global__ __launch_bounds__(256) void k_batch_scatter64(int len, int k,
                                                               const double* __restrict__ in,
                                                               float* __restrict__ out) {
  constexpr int kBatch = KB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[(size_t)i * kBatch + k] = (float)in[i];
}

// Trajectory k's target in original units to scaled ones (x̃ ./ D_c,
// src/pump.jl:331): io holds x̃ on entry and x̃ ./ D_c (IEEE division, as the
// host's) on exit; out (interleaved) receives it in FP32.
template <int KB>
__global__ __launch_bounds__(256) void k_batch_unscale64(int len, int k,
                                                         double* __restrict__ io,
                                                         const double* __restrict__ dc,
                                                         float* __restrict__ out) {
  constexpr int kBatch = KB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) {
    const double v = io[i] / dc[i];
    io[i] = v;
    out[(size_t)i * kBatch + k] = (float)v;
  }
}

// Reads column k of an interleaved vector into a dense FP32 vector.
template <int KB>
__global__ __launch_bounds__(256) void k_batch_gather(int len, int k,
                                                            const float* __restrict__ in,
                                                            float* __restrict__ out) {
  constexpr int kBatch = KB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[i] = in[(size_t)i * kBatch + k];
}

// FP64 polish of trajectory k's projection (see k_polish).
template <int KB>
__global__ __launch_bounds__(256) void k_batch_polish(int n, int k,
                                                             const float* __restrict__ px,
                                                             const float* __restrict__ lb,
                                                             const float* __restrict__ ub,
                                                             const double* __restrict__ lb64,
                                                             const double* __restrict__ ub64,
                                                             const double* __restrict__ xt64,
                                                             const double* __restrict__ w64,
                                                             double* __restrict__ x64,
                                                             double* __restrict__ partials) {
  constexpr int kBatch = KB;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (j < n) {
    const float p = px[(size_t)j * kBatch + k];
    const double v = p <= lb[j] ? lb64[j] : (p >= ub[j] ? ub64[j] : (double)p);
    x64[j] = v;
    const double wj = w64[j];
    if (wj > 0.0) acc[0] = wj * fabs(v - xt64[j]);
  }
  block_store_partials<256, 1>(acc, partials + blockIdx.x);
}

}  // namespace pdfp
