// Whole evaluation blocks of the projector in one workgroup.
//
// The pump's small instances (the examples: p0201 is 133 × 201 with 1,923
// nonzeros) make a PDHG iteration launch-bound: the reference issues five
// launches per iteration (src/pdfp.jl:158-163), the general FP32 form here
// two, each a few microseconds of launch and drain for well under a
// microsecond of work.  When the instance fits in LDS, one launch of one
// workgroup runs a whole block of `freq` iterations and its evaluation
// (src/pdfp.jl:252-264).  The iterate type V is FP32 (the fused form, and
// the batch's and pack's trajectories) or FP64 (the FP64 form, when its
// caller asks for this path):
//
//   LDS: CSR(Aᵀ) and CSR(A) (int32 pointers and columns, values in V) and
//        every iterate and problem vector the iteration touches (in V): z,
//        z⁰, b per row; x, y, x⁰, y⁰, ℓ, u, x̃, w, x̄ per column; is_eq
//        (bytes).  See small_lds_bytes.
//   per iteration: a pass over the columns, s_j = Σ a·z (FP64), and the x /
//        y steps; a barrier; a pass over the rows, s_i = Σ a·x̄, and the z
//        step; a barrier.  Each pass gives every row (of Aᵀ or A) a group
//        of L lanes (host-chosen per operator from its row count and its
//        longest row: the dependent LDS loads of one long row, not the
//        bandwidth, bound an iteration here); the lanes stride the row and
//        combine by xor-shuffles, and the group's first lane runs the
//        per-element math, which is the general epilogues'
//        (primal_y_math, dual_z_math).
//   major iterations (the first after a restart and the last) also write
//        pdhg x, y, z, reflected x, y, z and the dual slack to HBM, which the
//        evaluation and the restart read.
//   evaluation: the general evaluation epilogues' FP64 terms, reduced in a
//        fixed order by the workgroup into the 16-double slot.
//   end: x, y and z go back to HBM.
#pragma once

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "engine.h"
#include "host_util.h"
#include "pdfp_kernels.h"

namespace pdfp {

constexpr int kSmallBlock = 1024;
// Dynamic LDS one workgroup may take for an instance: the current device's
// per-workgroup maximum (160 KiB on gfx950) less 8 KiB for the kernels'
// static LDS, at most PDFP_SMALL_LDS_MAX.
#ifndef PDFP_SMALL_LDS_MAX
#define PDFP_SMALL_LDS_MAX (152 * 1024)
#endif
inline int64_t small_lds_max() {
  int dev = 0, bytes = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 0;  // unknown: no one-workgroup path
  return std::max<int64_t>(std::min<int64_t>((int64_t)bytes - 8192, PDFP_SMALL_LDS_MAX), 0);
}

// Values of both operators (2 nnz), 9 column and 3 row vectors in V, the
// int32 row pointers and columns, and is_eq: for FP32, 4(n+1) + 4(m+1) +
// 16 nnz + 36 n + 13 m bytes.
template <typename V = float>
inline int64_t small_lds_bytes(int64_t m, int64_t n, int64_t nnz) {
  constexpr int64_t v = sizeof(V);
  return (2 * v + 8) * nnz + 9 * v * n + 3 * v * m + 4 * (n + 1) + 4 * (m + 1) + m;
}

// Lanes per row L ∈ {1, 2, ..., 64} of one operator in the one-workgroup
// kernel: the fewest dependent steps, ⌈rows / (block / L)⌉ rounds of
// ⌈longest / L⌉ loads and log₂ L shuffles each.
inline int small_lanes(const std::vector<int32_t>& rowptr, int block) {
  const int64_t rows = (int64_t)rowptr.size() - 1;
  int64_t longest = 0;
  for (int64_t r = 0; r < rows; ++r)
    longest = std::max<int64_t>(longest, rowptr[(size_t)r + 1] - rowptr[(size_t)r]);
  int best = 1;
  int64_t best_cost = INT64_MAX;
  for (int L = 1, lg = 0; L <= 64; L *= 2, ++lg) {
    const int64_t groups = block / L;
    const int64_t cost = ((rows + groups - 1) / groups) * ((longest + L - 1) / L + lg);
    if (cost < best_cost) {
      best_cost = cost;
      best = L;
    }
  }
  return best;
}

template <typename TV>
struct SmallCsrT {
  const int32_t* __restrict__ ptr;
  const int32_t* __restrict__ col;
  const TV* __restrict__ val;
};
using SmallCsr = SmallCsrT<float>;

// The iterate vectors (x, y, z, their initial, reflected and pdhg copies,
// x̃, the dual slack) are read at index·S: S = 1 for one instance's own
// vectors, S = 8 for one trajectory's view of the batch's interleaved ones
// (the pointers then start at the trajectory's slot).  The problem data (ℓ,
// u, w, b, is_eq) are always dense.
template <class EC, class ER, typename V = float>
struct SmallArgsT {
  SmallCsrT<V> A, At;
  int m, n, nnz;
  int lanes_a, lanes_t;  // lanes per row of A and of Aᵀ (powers of two ≤ 64)
  int freq;         // iterations in this launch
  int first_major;  // iteration 1 is a major step, evaluated into slot 1
  const StepParams* sp;
  // HBM vectors (V)
  V *x, *y, *z;
  const V *x0, *y0, *z0, *lb, *ub, *xt, *w, *b;
  const uint8_t* is_eq;
  V *rx, *px, *ds, *py, *ry, *pz, *rz;
  EC ec;
  ER er;
  double* sums;  // slot 0 at sums[0..16), slot 1 at sums[16..32)
};
template <typename V>
using SmallArgsV = SmallArgsT<EpiEvalCol<V>, EpiEvalRow<V>, V>;
using SmallArgs = SmallArgsV<float>;


template <int BLOCK, int NACC>
__device__ __forceinline__ void small_reduce(double (&acc)[NACC], double* __restrict__ out,
                                             double* red /* LDS, BLOCK/64 × NACC */) {
  constexpr int kW = BLOCK / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[wid * NACC + k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NACC) {
    double s = 0.0;
    for (int w = 0; w < kW; ++w) s += red[w * NACC + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// The evaluation of the last major step from the HBM copies it wrote.
template <int BLOCK, int S, class EC, class ER, typename V>
__device__ __forceinline__ void small_evaluate(const SmallArgsT<EC, ER, V>& a,
                                               double* __restrict__ out, double* red) {
  const int tid = threadIdx.x;
  {
    double acc[kEvalColTerms];
#pragma unroll
    for (int k = 0; k < kEvalColTerms; ++k) acc[k] = 0.0;
    for (int j = tid; j < a.n; j += BLOCK) {
      double s[2] = {0.0, 0.0};
      for (int k = a.At.ptr[j]; k < a.At.ptr[j + 1]; ++k) {
        const double av = (double)a.At.val[k];
        const int c = a.At.col[k];
        const double zc = (double)a.pz[(size_t)c * S];
        s[0] += av * zc;
        s[1] += av * (zc - (double)a.rz[(size_t)c * S]);
      }
      a.ec(j, a.ec.pre(j), s, acc);
    }
    small_reduce<BLOCK, kEvalColTerms>(acc, out, red);
  }
  {
    double acc[kEvalRowTerms];
#pragma unroll
    for (int k = 0; k < kEvalRowTerms; ++k) acc[k] = 0.0;
    for (int i = tid; i < a.m; i += BLOCK) {
      double s = 0.0;
      for (int k = a.A.ptr[i]; k < a.A.ptr[i + 1]; ++k)
        s += (double)a.A.val[k] * (double)a.px[(size_t)a.A.col[k] * S];
      a.er(i, a.er.pre(i), &s, acc);
    }
    small_reduce<BLOCK, kEvalRowTerms>(acc, out + 8, red);
  }
}

// One block of a.freq iterations of one instance in the calling workgroup;
// smem holds small_lds_bytes<V>(m, n, nnz) bytes, red the reduction scratch.
// The V arrays come first, so each starts at a multiple of sizeof(V).
template <int BLOCK, int S, class EC, class ER, typename V>
__device__ __forceinline__ void small_block_body(const SmallArgsT<EC, ER, V>& a,
                                                 unsigned char* smem, double* red) {
  const int tid = threadIdx.x;
  const int n = a.n, m = a.m, nnz = a.nnz;
  V* t_val = reinterpret_cast<V*>(smem);
  V* a_val = t_val + nnz;
  V* s_z = a_val + nnz;
  V* s_z0 = s_z + m;
  V* s_b = s_z0 + m;
  V* s_rx = s_b + m;
  V* s_x = s_rx + n;
  V* s_y = s_x + n;
  V* s_x0 = s_y + n;
  V* s_y0 = s_x0 + n;
  V* s_lb = s_y0 + n;
  V* s_ub = s_lb + n;
  V* s_xt = s_ub + n;
  V* s_w = s_xt + n;
  int32_t* t_ptr = reinterpret_cast<int32_t*>(s_w + n);
  int32_t* a_ptr = t_ptr + (n + 1);
  int32_t* t_col = a_ptr + (m + 1);
  int32_t* a_col = t_col + nnz;
  uint8_t* s_eq = reinterpret_cast<uint8_t*>(a_col + nnz);

  for (int k = tid; k <= n; k += BLOCK) t_ptr[k] = a.At.ptr[k];
  for (int k = tid; k <= m; k += BLOCK) a_ptr[k] = a.A.ptr[k];
  for (int k = tid; k < nnz; k += BLOCK) {
    t_col[k] = a.At.col[k];
    t_val[k] = a.At.val[k];
    a_col[k] = a.A.col[k];
    a_val[k] = a.A.val[k];
  }
  for (int i = tid; i < m; i += BLOCK) {
    s_z[i] = a.z[(size_t)i * S];
    s_z0[i] = a.z0[(size_t)i * S];
    s_b[i] = a.b[i];
    s_eq[i] = a.is_eq[i];
  }
  for (int j = tid; j < n; j += BLOCK) {
    s_x[j] = a.x[(size_t)j * S];
    s_y[j] = a.y[(size_t)j * S];
    s_x0[j] = a.x0[(size_t)j * S];
    s_y0[j] = a.y0[(size_t)j * S];
    s_lb[j] = a.lb[j];
    s_ub[j] = a.ub[j];
    s_xt[j] = a.xt[(size_t)j * S];
    s_w[j] = a.w[j];
  }
  __syncthreads();

  // the block's step parameters, read once (the loop's stores to HBM would
  // otherwise make the compiler reload them every iteration)
  const StepParams sp = *a.sp;
  for (int it = 1; it <= a.freq; ++it) {
    const bool major = it == a.freq || (it == 1 && a.first_major);
    const StepScalars c = step_scalars(&sp, it);
    // columns: s_j = (Aᵀ z)_j, then the x / y steps
    const int Lt = a.lanes_t, gt = BLOCK / Lt, lt = tid & (Lt - 1);
    for (int j0 = 0; j0 < n; j0 += gt) {
      const int j = j0 + tid / Lt;
      double s = 0.0;
      if (j < n) {
        const int k1 = t_ptr[j + 1];
#pragma unroll 4
        for (int k = t_ptr[j] + lt; k < k1; k += Lt) s += (double)t_val[k] * (double)s_z[t_col[k]];
      }
      for (int o = Lt >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, Lt);
      if (j >= n || lt != 0) continue;
      const PrimalYOut o = primal_y_math(c, s, (double)s_x[j], (double)s_y[j], (double)s_x0[j],
                                         (double)s_y0[j], (double)s_lb[j], (double)s_ub[j],
                                         (double)s_xt[j], (double)s_w[j]);
      s_x[j] = (V)o.x;
      s_y[j] = (V)o.y;
      s_rx[j] = (V)o.rx;
      if (major) {
        a.rx[(size_t)j * S] = (V)o.rx;
        a.px[(size_t)j * S] = (V)o.px;
        a.ds[(size_t)j * S] = (V)o.ds;
        a.py[(size_t)j * S] = (V)o.py;
        a.ry[(size_t)j * S] = (V)o.ry;
      }
    }
    __syncthreads();
    // rows: s_i = (A x̄)_i, then the z step
    const int La = a.lanes_a, ga = BLOCK / La, la = tid & (La - 1);
    for (int i0 = 0; i0 < m; i0 += ga) {
      const int i = i0 + tid / La;
      double s = 0.0;
      if (i < m) {
        const int k1 = a_ptr[i + 1];
#pragma unroll 4
        for (int k = a_ptr[i] + la; k < k1; k += La) s += (double)a_val[k] * (double)s_rx[a_col[k]];
      }
      for (int o = La >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, La);
      if (i >= m || la != 0) continue;
      const DualZOut o = dual_z_math(c, s, (double)s_z[i], (double)s_z0[i], (double)s_b[i],
                                     s_eq[i] != 0);
      s_z[i] = (V)o.z;
      if (major) {
        a.pz[(size_t)i * S] = (V)o.pz;
        a.rz[(size_t)i * S] = (V)o.rz;
      }
    }
    __syncthreads();
    if (it == 1 && a.first_major && a.freq > 1) small_evaluate<BLOCK, S>(a, a.sums + 16, red);
  }
  small_evaluate<BLOCK, S>(a, a.sums, red);
  if (a.freq == 1 && a.first_major) {
    // the first step was also the last: both slots hold its evaluation
    if (tid < 16) a.sums[16 + tid] = a.sums[tid];
  }
  for (int i = tid; i < m; i += BLOCK) a.z[(size_t)i * S] = s_z[i];
  for (int j = tid; j < n; j += BLOCK) {
    a.x[(size_t)j * S] = s_x[j];
    a.y[(size_t)j * S] = s_y[j];
  }
}

template <int BLOCK, typename V>
__global__ __launch_bounds__(BLOCK) void k_small_block(SmallArgsV<V> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[(BLOCK / kWave) * kEvalColTerms];
  small_block_body<BLOCK, 1>(a, smem, red);
}

// Independent small instances, one workgroup each (workgroup b runs the
// instance of args[b]); the launch's dynamic LDS is the largest instance's.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_small_pack(const SmallArgs* __restrict__ args) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[(BLOCK / kWave) * kEvalColTerms];
  const SmallArgs a = args[blockIdx.x];
  small_block_body<BLOCK, 1>(a, smem, red);
}

// ---- the power iteration in one workgroup ----------------------------------
// estimate_max_singular_value (src/preconditioner.jl:108-125) for an instance
// that fits in LDS: the iteration runs to its stopping rule inside one launch
// (each iteration is two SpMVs and three reductions, a few microseconds),
// instead of seven launches per iteration.  The sums are FP64, in a fixed
// order.  LDS: small_power_lds_bytes.

template <typename TV>
inline int64_t small_power_lds_bytes(int64_t m, int64_t n, int64_t nnz) {
  return 4 * (n + 1) + 4 * (m + 1) + 8 * nnz + 2 * (int64_t)sizeof(TV) * nnz + 16 * m + 8 * n;
}

// Sum over the workgroup, returned to every thread.
template <int BLOCK>
__device__ __forceinline__ double block_sum_all(double v, double* red /* BLOCK/64 + 1 */) {
  constexpr int kW = BLOCK / kWave;
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kW; ++w) s += red[w];
    red[kW] = s;
  }
  __syncthreads();
  const double s = red[kW];
  __syncthreads();  // red is reused by the next call
  return s;
}

// y[r] = Σ_k val[k]·x[col[k]] over the rows r < nrows of a CSR in LDS, one
// group of L lanes per row.
template <int BLOCK, typename TV>
__device__ __forceinline__ void small_spmv(int nrows, int L, const int32_t* ptr,
                                           const int32_t* col, const TV* val, const double* x,
                                           double* y) {
  const int groups = BLOCK / L, lane = threadIdx.x & (L - 1);
  for (int r0 = 0; r0 < nrows; r0 += groups) {
    const int r = r0 + (int)threadIdx.x / L;
    double s = 0.0;
    if (r < nrows) {
      const int k1 = ptr[r + 1];
#pragma unroll 4
      for (int k = ptr[r] + lane; k < k1; k += L) s += (double)val[k] * x[col[k]];
    }
    for (int o = L >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    if (r < nrows && lane == 0) y[r] = s;
  }
}

template <int BLOCK, typename TV>
__global__ __launch_bounds__(BLOCK) void k_small_power(SmallCsrT<TV> A, SmallCsrT<TV> At, int m,
                                                       int n, int nnz, int lanes_a, int lanes_t,
                                                       const double* __restrict__ start,
                                                       int max_iterations, double tolerance,
                                                       double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[BLOCK / kWave + 1];
  const int tid = threadIdx.x;
  double* eig = reinterpret_cast<double*>(smem);
  double* nxt = eig + m;
  double* dual = nxt + m;
  TV* t_val = reinterpret_cast<TV*>(dual + n);
  TV* a_val = t_val + nnz;
  int32_t* t_ptr = reinterpret_cast<int32_t*>(a_val + nnz);
  int32_t* a_ptr = t_ptr + (n + 1);
  int32_t* t_col = a_ptr + (m + 1);
  int32_t* a_col = t_col + nnz;
  for (int k = tid; k <= n; k += BLOCK) t_ptr[k] = At.ptr[k];
  for (int k = tid; k <= m; k += BLOCK) a_ptr[k] = A.ptr[k];
  for (int k = tid; k < nnz; k += BLOCK) {
    t_col[k] = At.col[k];
    t_val[k] = At.val[k];
    a_col[k] = A.col[k];
    a_val[k] = A.val[k];
  }
  for (int i = tid; i < m; i += BLOCK) eig[i] = start[i];
  __syncthreads();
  double sigma_sq = 1.0;
  int it = 0;
  for (; it < max_iterations; ++it) {
    double acc = 0.0;
    for (int i = tid; i < m; i += BLOCK) acc += eig[i] * eig[i];
    const double nrm = sqrt(block_sum_all<BLOCK>(acc, red));
    if (nrm == 0.0) break;
    for (int i = tid; i < m; i += BLOCK) nxt[i] = eig[i] / nrm;
    __syncthreads();
    small_spmv<BLOCK>(n, lanes_t, t_ptr, t_col, t_val, nxt, dual);
    __syncthreads();
    small_spmv<BLOCK>(m, lanes_a, a_ptr, a_col, a_val, dual, eig);
    __syncthreads();
    acc = 0.0;
    for (int i = tid; i < m; i += BLOCK) acc += nxt[i] * eig[i];
    sigma_sq = block_sum_all<BLOCK>(acc, red);
    acc = 0.0;
    for (int i = tid; i < m; i += BLOCK) {
      const double v = eig[i] - sigma_sq * nxt[i];
      nxt[i] = v;
      acc += v * v;
    }
    const double res = block_sum_all<BLOCK>(acc, red);
    if (sqrt(res) < tolerance) {
      ++it;
      break;
    }
  }
  if (tid == 0) {
    out[0] = sigma_sq;
    out[1] = (double)it;
  }
}

// Host side: σ_max by the one-workgroup power iteration from `start` (m
// values), or NAN when the instance does not fit in LDS.
template <typename TV>
double small_power_iteration(SmallCsrT<TV> A, SmallCsrT<TV> At, const std::vector<int32_t>& a_ptr,
                             const std::vector<int32_t>& t_ptr, int64_t m, int64_t n, int64_t nnz,
                             const double* start, int max_iterations, double tolerance,
                             hipStream_t stream) {
  const int64_t lds = small_power_lds_bytes<TV>(m, n, nnz);
  if (m <= 0 || n <= 0 || nnz <= 0 || lds > small_lds_max()) return NAN;
  DevBuf<double> d_start, d_out;
  PinnedBuf<double> h_out;
  d_start.alloc((size_t)m);
  d_out.alloc(2);
  h_out.alloc(2);
  d_start.upload(start, (size_t)m, stream);
  k_small_power<kSmallBlock, TV><<<1, kSmallBlock, (size_t)lds, stream>>>(
      A, At, (int)m, (int)n, (int)nnz, small_lanes(a_ptr, kSmallBlock),
      small_lanes(t_ptr, kSmallBlock), d_start.p, max_iterations, tolerance, d_out.p);
  PDFP_HIP(hipGetLastError());
  PDFP_HIP(hipMemcpyAsync(h_out.p, d_out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
  PDFP_HIP(hipStreamSynchronize(stream));
  return std::sqrt(std::max(h_out.p[0], 0.0));
}

// The same for an operator pair on the device, from the reference's start
// vector; NAN when it does not fit.
template <typename TV>
double small_sigma(const DevCsr<TV>& A, const DevCsr<TV>& At, int64_t m, int64_t n,
                   int max_iterations, double tolerance, hipStream_t stream) {
  if (m <= 0 || small_power_lds_bytes<TV>(m, n, A.nnz) > small_lds_max()) return NAN;
  std::vector<double> start((size_t)m);
  start_vector(m, kSvSeed, start.data());
  return small_power_iteration<TV>(SmallCsrT<TV>{A.rowptr.p, A.col.p, A.val.p},
                                   SmallCsrT<TV>{At.rowptr.p, At.col.p, At.val.p}, A.h_rowptr,
                                   At.h_rowptr, m, n, A.nnz, start.data(), max_iterations,
                                   tolerance, stream);
}

}  // namespace pdfp
