// Whole evaluation blocks of the fused FP32 form in one workgroup.
//
// The pump's small instances (the examples: p0201 is 133 × 201 with 1,923
// nonzeros) make a PDHG iteration launch-bound: the reference issues five
// launches per iteration (src/pdfp.jl:158-163), the general form here two,
// each a few microseconds of launch and drain for well under a microsecond
// of work.  When both CSR operators and the two gathered vectors fit in LDS,
// one launch of one workgroup runs a whole block of `freq` iterations and
// its evaluation (src/pdfp.jl:252-264):
//
//   LDS: CSR(Aᵀ) and CSR(A) (int32 pointers and columns, FP32 values), the
//        current z and the reflected x̄ (FP32), i.e. 4(n+1) + 4(m+1) +
//        16·nnz + 4m + 4n bytes.
//   per iteration: thread-per-column pass s_j = Σ a·z (FP64, ascending
//        column order) with the x/y epilogue, which writes x̄ to LDS; a
//        barrier; thread-per-row pass s_i = Σ a·x̄ with the z epilogue,
//        which writes z to LDS; a barrier.
//   major iterations (the first after a restart and the last) also write
//        pdhg x, y, z, reflected y, z, the dual slack and x̄ to HBM, which
//        the evaluation and the restart read.
//   evaluation: the same FP64 terms as the general passes' evaluation
//        epilogues, reduced in a fixed order by the workgroup into the
//        16-double slot.
//
// The other iterates (x, y and the anchors) stay in HBM: each thread reads
// and writes only its own columns and rows, so the L1 serves them.  The
// epilogues are the general passes' (pdfp_kernels.h), only their x̄ and z
// pointers point into LDS.
#pragma once

#include "engine.h"
#include "pdfp_kernels.h"

namespace pdfp {

constexpr int kSmallBlock = 1024;
constexpr int64_t kSmallLdsMax = 64 * 1024;

inline int64_t small_lds_bytes(int64_t m, int64_t n, int64_t nnz) {
  return 4 * (n + 1) + 4 * (m + 1) + 16 * nnz + 4 * m + 4 * n;
}

struct SmallCsr {
  const int32_t* __restrict__ ptr;
  const int32_t* __restrict__ col;
  const float* __restrict__ val;
};

struct SmallArgs {
  SmallCsr A, At;
  int m, n, nnz;
  int freq;        // iterations in this launch
  int first_major; // iteration 1 is a major step (and is evaluated into slot 1)
  EpiPrimalY<float, false> ex;   // k_offset and the x̄ pointer are set in the kernel
  EpiDualZ<float, false> ez;     // k_offset and the z pointer are set in the kernel
  EpiEvalCol<float> ec;
  EpiEvalRow<float> er;
  float* __restrict__ z;         // current z in HBM (read at start, written at end)
  float* __restrict__ rx;        // reflected x in HBM (written on major steps)
  const float* __restrict__ pz;  // pdhg z, reflected z, pdhg x (evaluation gathers)
  const float* __restrict__ rz;
  const float* __restrict__ px;
  double* __restrict__ sums;     // slot 0 at sums[0..16), slot 1 at sums[16..32)
};

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

template <int BLOCK>
__device__ __forceinline__ void small_evaluate(const SmallArgs& a, double* __restrict__ out,
                                               double* red) {
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
        const double zc = (double)a.pz[c];
        s[0] += av * zc;
        s[1] += av * (zc - (double)a.rz[c]);
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
        s += (double)a.A.val[k] * (double)a.px[a.A.col[k]];
      a.er(i, a.er.pre(i), &s, acc);
    }
    small_reduce<BLOCK, kEvalRowTerms>(acc, out + 8, red);
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_small_block(SmallArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[(BLOCK / kWave) * kEvalColTerms];
  const int tid = threadIdx.x;
  const int n = a.n, m = a.m, nnz = a.nnz;
  int32_t* t_ptr = reinterpret_cast<int32_t*>(smem);
  int32_t* a_ptr = t_ptr + (n + 1);
  int32_t* t_col = a_ptr + (m + 1);
  int32_t* a_col = t_col + nnz;
  float* t_val = reinterpret_cast<float*>(a_col + nnz);
  float* a_val = t_val + nnz;
  float* s_z = a_val + nnz;
  float* s_rx = s_z + m;

  for (int k = tid; k <= n; k += BLOCK) t_ptr[k] = a.At.ptr[k];
  for (int k = tid; k <= m; k += BLOCK) a_ptr[k] = a.A.ptr[k];
  for (int k = tid; k < nnz; k += BLOCK) {
    t_col[k] = a.At.col[k];
    t_val[k] = a.At.val[k];
    a_col[k] = a.A.col[k];
    a_val[k] = a.A.val[k];
  }
  for (int i = tid; i < m; i += BLOCK) s_z[i] = a.z[i];
  __syncthreads();

  for (int it = 1; it <= a.freq; ++it) {
    const bool major = it == a.freq || (it == 1 && a.first_major);
    // columns: s_j = (Aᵀ z)_j, then the x / y steps (x̄ into LDS)
    for (int j = tid; j < n; j += BLOCK) {
      double s = 0.0;
      for (int k = t_ptr[j]; k < t_ptr[j + 1]; ++k) s += (double)t_val[k] * (double)s_z[t_col[k]];
      if (major) {
        EpiPrimalY<float, true> e{a.ex.sp, it, a.ex.x, a.ex.x0, a.ex.y, a.ex.y0, a.ex.lb,
                                  a.ex.ub, a.ex.xt, a.ex.w, s_rx, a.ex.px, a.ex.ds, a.ex.py,
                                  a.ex.ry};
        e(j, e.pre(j), &s, nullptr);
        a.rx[j] = s_rx[j];
      } else {
        EpiPrimalY<float, false> e = a.ex;
        e.k_offset = it;
        e.rx = s_rx;
        e(j, e.pre(j), &s, nullptr);
      }
    }
    __syncthreads();
    // rows: s_i = (A x̄)_i, then the z step (z into LDS)
    for (int i = tid; i < m; i += BLOCK) {
      double s = 0.0;
      for (int k = a_ptr[i]; k < a_ptr[i + 1]; ++k) s += (double)a_val[k] * (double)s_rx[a_col[k]];
      if (major) {
        EpiDualZ<float, true> e{a.ez.sp, it, s_z, a.ez.z0, a.ez.b, a.ez.is_eq, a.ez.pz, a.ez.rz};
        e(i, e.pre(i), &s, nullptr);
      } else {
        EpiDualZ<float, false> e = a.ez;
        e.k_offset = it;
        e.z = s_z;
        e(i, e.pre(i), &s, nullptr);
      }
    }
    __syncthreads();
    if (it == 1 && a.first_major && a.freq > 1) small_evaluate<BLOCK>(a, a.sums + 16, red);
  }
  small_evaluate<BLOCK>(a, a.sums, red);
  if (a.freq == 1 && a.first_major) {
    // the first step was also the last: both slots hold its evaluation
    if (tid < 16) a.sums[16 + tid] = a.sums[tid];
  }
  for (int i = tid; i < m; i += BLOCK) a.z[i] = s_z[i];
}

}  // namespace pdfp
