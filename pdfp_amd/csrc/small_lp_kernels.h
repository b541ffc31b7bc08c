// Whole evaluation blocks of the root-LP PDHG (FP64) in one workgroup.
//
// The pump's root LP on its small instances (p0201: 133 × 201, 1,923
// nonzeros) is launch-bound in the general form: two launches per iteration
// of a few microseconds each, 600 iterations.  When the instance fits in
// LDS, one launch of one workgroup runs a block of `freq` iterations of
// pdlp_solve! (src/pdlp.jl:325-372) and its evaluation(s), as k_small_block
// does for the projector:
//
//   LDS: CSR(Aᵀ) and CSR(A) (int32 pointers and columns, FP64 values) and
//        the iterates and problem vectors the iteration touches (FP64):
//        x, x⁰, c, ℓ, u, x̄ per column; y, y⁰, ℓ_c, u_c per row.  See
//        small_lp_lds_bytes.
//   per iteration: a pass over the columns (s_j = (Aᵀy)_j, K4 by
//        lp_primal_math), a barrier, a pass over the rows (s_i = (A x̄)_i, K5
//        by lp_dual_math), a barrier; each row of an operator is reduced by a
//        host-chosen group of lanes (small_lanes).  Iterations 1 and `freq`
//        are major: they write pdhg x, y, the reflected x, y and the dual
//        slack to HBM, which the evaluation and the restart read.
//   evaluation: the general evaluation's gathers and epilogues
//        (GatherLpCol/Row, EpiLpEvalCol/Row), reduced in a fixed order by the
//        workgroup into the general layout (slot·32: column terms, then row
//        terms at +16; the last kLpColMax / kLpRowMax terms are maxima).
//   end: x and y go back to HBM.
#pragma once

#include "pdlp_kernels.h"
#include "small_kernels.h"

namespace pdfp {

inline int64_t small_lp_lds_bytes(int64_t m, int64_t n, int64_t nnz) {
  return 24 * nnz + 48 * n + 32 * m + 4 * (n + 1) + 4 * (m + 1);
}

struct SmallLpArgs {
  SmallCsrT<double> A, At;
  int m, n, nnz;
  int lanes_a, lanes_t;
  int freq;          // iterations in this launch (≥ 2)
  int restart_eval;  // evaluate iteration 1 into slot 1
  const StepParams* sp;
  double *x, *y;
  const double *x0, *y0, *c, *lb, *ub, *clb, *cub;
  double *rx, *px, *ds, *py, *ry;
  const double* col_unscale;  // null: scaled units
  const double* row_unscale;
  double* sums;  // slot s at sums[32 s ...]
};

// Sums of terms [0, NACC - NMAX) and maxima of the last NMAX over the
// workgroup, in a fixed order, into out[0 .. NACC).
template <int BLOCK, int NACC, int NMAX>
__device__ __forceinline__ void small_reduce_mixed(double (&acc)[NACC], double* __restrict__ out,
                                                   double* red /* BLOCK/64 × NACC */) {
  constexpr int kW = BLOCK / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    const double v = k < NACC - NMAX ? wave_sum(acc[k]) : wave_max(acc[k]);
    if (lane == 0) red[wid * NACC + k] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < NACC) {
    const bool is_max = (int)threadIdx.x >= NACC - NMAX;
    double s = 0.0;
    for (int w = 0; w < kW; ++w) {
      const double v = red[w * NACC + threadIdx.x];
      s = is_max ? fmax(s, v) : s + v;
    }
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// The evaluation of the last major step (evaluate_t of the general form)
// from the HBM copies it wrote.
template <int BLOCK, bool RAYS>
__device__ __forceinline__ void small_lp_evaluate(const SmallLpArgs& a, double* __restrict__ out,
                                                  double* red) {
  const int tid = threadIdx.x;
  {
    const GatherLpCol<double, RAYS> g{a.py, a.ry, a.clb, a.cub};
    const EpiLpEvalCol<double, RAYS> ec{a.px, a.rx, a.ds, a.c, a.x0, a.lb, a.ub, a.col_unscale};
    constexpr int NV = GatherLpCol<double, RAYS>::NV;
    double acc[kLpColTerms];
#pragma unroll
    for (int k = 0; k < kLpColTerms; ++k) acc[k] = 0.0;
    for (int j = tid; j < a.n; j += BLOCK) {
      double s[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) s[q] = 0.0;
      for (int k = a.At.ptr[j]; k < a.At.ptr[j + 1]; ++k) {
        double p[NV];
        g(a.At.col[k], a.At.val[k], p);
#pragma unroll
        for (int q = 0; q < NV; ++q) s[q] += p[q];
      }
      ec(j, ec.pre(j), s, acc);
    }
    small_reduce_mixed<BLOCK, kLpColTerms, kLpColMax>(acc, out, red);
  }
  {
    const GatherLpRow<double, RAYS> g{a.px, a.rx, a.lb, a.ub};
    const EpiLpEvalRow<double, RAYS> er{a.py, a.ry, a.y0, a.clb, a.cub, a.row_unscale};
    constexpr int NV = GatherLpRow<double, RAYS>::NV;
    double acc[kLpRowTerms];
#pragma unroll
    for (int k = 0; k < kLpRowTerms; ++k) acc[k] = 0.0;
    for (int i = tid; i < a.m; i += BLOCK) {
      double s[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) s[q] = 0.0;
      for (int k = a.A.ptr[i]; k < a.A.ptr[i + 1]; ++k) {
        double p[NV];
        g(a.A.col[k], a.A.val[k], p);
#pragma unroll
        for (int q = 0; q < NV; ++q) s[q] += p[q];
      }
      er(i, er.pre(i), s, acc);
    }
    small_reduce_mixed<BLOCK, kLpRowTerms, kLpRowMax>(acc, out + 16, red);
  }
}

template <int BLOCK, bool RAYS>
__global__ __launch_bounds__(BLOCK) void k_small_lp_block(SmallLpArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[(BLOCK / kWave) * kLpColTerms];
  const int tid = threadIdx.x;
  const int n = a.n, m = a.m, nnz = a.nnz;
  double* t_val = reinterpret_cast<double*>(smem);
  double* a_val = t_val + nnz;
  double* s_x = a_val + nnz;
  double* s_x0 = s_x + n;
  double* s_c = s_x0 + n;
  double* s_lb = s_c + n;
  double* s_ub = s_lb + n;
  double* s_rx = s_ub + n;
  double* s_y = s_rx + n;
  double* s_y0 = s_y + m;
  double* s_clb = s_y0 + m;
  double* s_cub = s_clb + m;
  int32_t* t_ptr = reinterpret_cast<int32_t*>(s_cub + m);
  int32_t* a_ptr = t_ptr + (n + 1);
  int32_t* t_col = a_ptr + (m + 1);
  int32_t* a_col = t_col + nnz;

  for (int k = tid; k <= n; k += BLOCK) t_ptr[k] = a.At.ptr[k];
  for (int k = tid; k <= m; k += BLOCK) a_ptr[k] = a.A.ptr[k];
  for (int k = tid; k < nnz; k += BLOCK) {
    t_col[k] = a.At.col[k];
    t_val[k] = a.At.val[k];
    a_col[k] = a.A.col[k];
    a_val[k] = a.A.val[k];
  }
  for (int j = tid; j < n; j += BLOCK) {
    s_x[j] = a.x[j];
    s_x0[j] = a.x0[j];
    s_c[j] = a.c[j];
    s_lb[j] = a.lb[j];
    s_ub[j] = a.ub[j];
  }
  for (int i = tid; i < m; i += BLOCK) {
    s_y[i] = a.y[i];
    s_y0[i] = a.y0[i];
    s_clb[i] = a.clb[i];
    s_cub[i] = a.cub[i];
  }
  __syncthreads();

  const StepParams sp = *a.sp;  // read once (see k_small_block)
  for (int it = 1; it <= a.freq; ++it) {
    const bool major = it == 1 || it == a.freq;
    const StepScalars sc = step_scalars(&sp, it);
    // columns: s_j = (Aᵀ y)_j, then K4
    const int Lt = a.lanes_t, gt = BLOCK / Lt, lt = tid & (Lt - 1);
    for (int j0 = 0; j0 < n; j0 += gt) {
      const int j = j0 + tid / Lt;
      double s = 0.0;
      if (j < n) {
        const int k1 = t_ptr[j + 1];
#pragma unroll 4
        for (int k = t_ptr[j] + lt; k < k1; k += Lt) s += t_val[k] * s_y[t_col[k]];
      }
      for (int o = Lt >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, Lt);
      if (j >= n || lt != 0) continue;
      const LpPrimalOut o = lp_primal_math(sc, s, s_x[j], s_x0[j], s_c[j], s_lb[j], s_ub[j]);
      s_rx[j] = o.r;
      s_x[j] = o.x;
      if (major) {
        a.rx[j] = o.r;
        a.px[j] = o.p;
        a.ds[j] = (o.p - o.temp) / sc.tau;
      }
    }
    __syncthreads();
    // rows: s_i = (A x̄)_i, then K5
    const int La = a.lanes_a, ga = BLOCK / La, la = tid & (La - 1);
    for (int i0 = 0; i0 < m; i0 += ga) {
      const int i = i0 + tid / La;
      double s = 0.0;
      if (i < m) {
        const int k1 = a_ptr[i + 1];
#pragma unroll 4
        for (int k = a_ptr[i] + la; k < k1; k += La) s += a_val[k] * s_rx[a_col[k]];
      }
      for (int o = La >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, La);
      if (i >= m || la != 0) continue;
      const LpDualOut o = lp_dual_math(sc, s, s_y[i], s_y0[i], s_clb[i], s_cub[i]);
      s_y[i] = o.y;
      if (major) {
        a.py[i] = o.p;
        a.ry[i] = o.r;
      }
    }
    __syncthreads();
    if (it == 1 && a.restart_eval) small_lp_evaluate<BLOCK, RAYS>(a, a.sums + 32, red);
  }
  small_lp_evaluate<BLOCK, RAYS>(a, a.sums, red);
  for (int j = tid; j < n; j += BLOCK) a.x[j] = s_x[j];
  for (int i = tid; i < m; i += BLOCK) a.y[i] = s_y[i];
}

}  // namespace pdfp
