// Kernels of the block-diagonal pack: K independent projector instances of
// uneven sizes whose rows and columns are concatenated (A = diag(A_1..A_K)),
// iterated together in the FP32 form's two fused passes per iteration.
//
// Each column and row carries its instance's index (col_inst / row_inst);
// an epilogue takes that instance's step scalars from a K-entry table that
// the host fills per evaluation block.  The row-block schedule never crosses
// an instance boundary, so per-block partials of the evaluation sum per
// instance (k_pack_finalize).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"
#include "pdfp_kernels.h"

namespace pdfp {

// Instance k's scalars of the current block: τ = step/pw, σ = step·pw and
// 1 - ω formed on the host (the same FP64 operations as step_scalars), the
// inner count at the block's start, and whether the instance iterates.
struct PackStep {
  double tau;
  double sigma;
  double omega;
  double omw;
  double inner_count;
  double active;  // 1: iterate this instance, 0: leave it unchanged
  double pad[2];
};

// step_scalars (pdfp_kernels.h) of instance scalars ps at iteration k_offset
// of the block: w_k = k/(k+1), k = inner_count + k_offset (src/pdfp.jl:120-122).
__device__ __forceinline__ StepScalars pack_scalars(const PackStep& ps, int k_offset) {
  StepScalars s;
  s.tau = ps.tau;
  s.sigma = ps.sigma;
  const double k = ps.inner_count + (double)k_offset;
  s.weight = k / (k + 1.0);
  s.omega = ps.omega;
  s.omw = ps.omw;
  s.omwt = 1.0 - s.weight;
  return s;
}

// Column j (of instance inst_j): the fused primal and dual-y step
// (src/pdfp.jl:119-143, the math of EpiPrimalY) with s = (Aᵀz)_j.
template <bool MAJOR>
struct EpiPackPrimalY {
  struct Pre {
    float x, y, x0, y0, lb, ub, xt, w;
    int32_t inst;
  };
  const PackStep* sp;
  int k_offset;
  const int32_t* __restrict__ col_inst;
  float* __restrict__ x;
  const float* __restrict__ x0;
  float* __restrict__ y;
  const float* __restrict__ y0;
  const float* __restrict__ lb;
  const float* __restrict__ ub;
  const float* __restrict__ xt;
  const float* __restrict__ w;
  float* __restrict__ rx;
  float* __restrict__ px;
  float* __restrict__ ds;
  float* __restrict__ py;
  float* __restrict__ ry;
  __device__ __forceinline__ Pre pre(int j) const {
    return {x[j], y[j], x0[j], y0[j], lb[j], ub[j], xt[j], w[j], col_inst[j]};
  }
  __device__ __forceinline__ void operator()(int j, const Pre& p, const double* s,
                                             double*) const {
    const PackStep& ps = sp[p.inst];
    const StepScalars c = pack_scalars(ps, k_offset);
    const PrimalYOut o = primal_y_math(c, s[0], (double)p.x, (double)p.y, (double)p.x0,
                                       (double)p.y0, (double)p.lb, (double)p.ub, (double)p.xt,
                                       (double)p.w);
    if (ps.active == 0.0) return;
    x[j] = (float)o.x;
    rx[j] = (float)o.rx;
    y[j] = (float)o.y;
    if (MAJOR) {
      px[j] = (float)o.px;
      ds[j] = (float)o.ds;
      py[j] = (float)o.py;
      ry[j] = (float)o.ry;
    }
  }
};

// Row i (of instance inst_i): the dual-z step (src/pdfp.jl:146-156, the math
// of EpiDualZ) with s = (A x̄)_i.
template <bool MAJOR>
struct EpiPackDualZ {
  struct Pre {
    float z, z0, b;
    int32_t inst;
    uint32_t eq;
  };
  const PackStep* sp;
  int k_offset;
  const int32_t* __restrict__ row_inst;
  float* __restrict__ z;
  const float* __restrict__ z0;
  const float* __restrict__ b;
  const uint8_t* __restrict__ is_eq;
  float* __restrict__ pz;
  float* __restrict__ rz;
  __device__ __forceinline__ Pre pre(int i) const {
    return {z[i], z0[i], b[i], row_inst[i], (uint32_t)is_eq[i]};
  }
  __device__ __forceinline__ void operator()(int i, const Pre& p, const double* s,
                                             double*) const {
    const PackStep& ps = sp[p.inst];
    const StepScalars c = pack_scalars(ps, k_offset);
    const DualZOut o = dual_z_math(c, s[0], (double)p.z, (double)p.z0, (double)p.b, p.eq != 0);
    if (ps.active == 0.0) return;
    z[i] = (float)o.z;
    if (MAJOR) {
      pz[i] = (float)o.pz;
      rz[i] = (float)o.rz;
    }
  }
};

// out[k·32 + offset + t] = Σ_{b ∈ instance k's blocks} partial[b·nt + t], in
// block order: one workgroup per (instance, term).
static __global__ __launch_bounds__(256) void k_pack_finalize(const double* __restrict__ partial,
                                                              const int32_t* __restrict__ seg,
                                                              int nt, int offset,
                                                              double* __restrict__ out) {
  __shared__ double red[256 / kWave];
  const int k = blockIdx.x / nt, t = blockIdx.x % nt;
  const int b0 = seg[k], b1 = seg[k + 1];
  double v = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += 256) v += partial[(size_t)b * nt + t];
  v = block_sum_to0<256>(v, red);
  if (threadIdx.x == 0) out[(size_t)k * 32 + offset + t] = v;
}

// perform_restart! (src/pdfp.jl:180-192) for every instance with flag set:
// initial ← pdhg, current ← pdhg.
static __global__ __launch_bounds__(256) void k_pack_restart(
    int n, int m, const int32_t* __restrict__ col_inst, const int32_t* __restrict__ row_inst,
    const uint8_t* __restrict__ flag, const float* __restrict__ px, const float* __restrict__ py,
    const float* __restrict__ pz, float* __restrict__ x, float* __restrict__ x0,
    float* __restrict__ y, float* __restrict__ y0, float* __restrict__ z,
    float* __restrict__ z0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[col_inst[i]]) {
    x[i] = px[i];
    x0[i] = px[i];
    y[i] = py[i];
    y0[i] = py[i];
  }
  if (i < m && flag[row_inst[i]]) {
    z[i] = pz[i];
    z0[i] = pz[i];
  }
}

// warm_start! (src/pdfp.jl:104-114) of one instance's slices: initial ← current.
static __global__ __launch_bounds__(256) void k_pack_warm(int n, int m, float* __restrict__ x0,
                                                          const float* __restrict__ x,
                                                          float* __restrict__ y0,
                                                          const float* __restrict__ y,
                                                          float* __restrict__ z0,
                                                          const float* __restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x0[i] = x[i];
    y0[i] = y[i];
  }
  if (i < m) z0[i] = z[i];
}

}  // namespace pdfp
