// Device kernels and fused epilogues of the PDFP projector.
//
// The saddle-point form of the pump's ℓ1 projection (src/pdfp.jl:1-13):
//   min_{ℓ≤x≤u} max_{y,z}  y'(x̃ - x) + z'(A x - b),
//   |y_j| ≤ w_j (j ∈ I), y_j = 0 (j ∉ I), z_i ≥ 0 on ≤ rows.
// One restarted Halpern PDHG iteration with reflection (src/pdfp.jl:119-163):
//   x⁺ = clamp(x - τ(Aᵀz - y), ℓ, u)                     (primal prox: box)
//   y⁺ = clamp(y + σ(x̃ - x̄), -w, w) on I, 0 elsewhere   (ℓ1 prox, dual form)
//   z⁺ = z + σ(A x̄ - b), projected to z ≥ 0 on ≤ rows
// with x̄ = 2x⁺ - x, every block reflected and Halpern-averaged toward the
// anchor: v ← w_k(ω v̄ + (1-ω) v) + (1 - w_k) v⁰, w_k = k/(k+1).
//
// Non-integer columns carry w_j = 0 in device memory (the reference's is_int
// mask folded into the weight: y_j = 0 exactly where is_int is false).
#pragma once

#include "engine.h"
#include "solver_common.h"

namespace pdfp {

// Scalars of the current evaluation block, in device memory so that a
// captured graph of the block stays valid when they change.
struct StepParams {
  double step_size;
  double primal_weight;
  double omega;        // reflection coefficient
  double inner_count;  // iterations since the last restart (exact in FP64)
};

struct StepScalars {
  double tau, sigma, weight, omega, omw, omwt;
};

// τ = step/pw, σ = step·pw, k = inner_count + k_offset, w_k = k/(k+1)
// (src/pdfp.jl:120-122, 135-137, 147-149) — the same FP64 operations.
__device__ __forceinline__ StepScalars step_scalars(const StepParams* __restrict__ sp,
                                                    int k_offset) {
  StepScalars s;
  s.tau = sp->step_size / sp->primal_weight;
  s.sigma = sp->step_size * sp->primal_weight;
  const double k = sp->inner_count + (double)k_offset;
  s.weight = k / (k + 1.0);
  s.omega = sp->omega;
  s.omw = 1.0 - s.omega;
  s.omwt = 1.0 - s.weight;
  return s;
}

// Julia's clamp(x, lo, hi) = x > hi ? hi : (x < lo ? lo : x).
__device__ __forceinline__ double jclamp(double x, double lo, double hi) {
  return x > hi ? hi : (x < lo ? lo : x);
}

// ---- the per-element math of one PDHG iteration (FP64) ---------------------
// Shared by the fused epilogues below and by the one-workgroup block kernel
// (small_kernels.h), so that both run the same operations.

struct PrimalYOut {
  double x, y, rx, px, ds, py, ry;
};

// Column j: K1 (primal step, pdfp_kernels.jl:4-20) and K2 (dual-y step,
// pdfp_kernels.jl:33-48) given s = (Aᵀz)_j.
__device__ __forceinline__ PrimalYOut primal_y_math(const StepScalars& c, double s, double cx,
                                                   double cy, double x0, double y0, double lb,
                                                   double ub, double xt, double wj) {
  PrimalYOut o;
  const double temp = cx - c.tau * (s - cy);
  o.px = jclamp(temp, lb, ub);
  o.rx = 2.0 * o.px - cx;
  o.x = c.weight * (c.omega * o.rx + c.omw * cx) + c.omwt * x0;
  o.py = wj > 0.0 ? jclamp(cy + c.sigma * (xt - o.rx), -wj, wj) : 0.0;
  o.ry = 2.0 * o.py - cy;
  o.y = c.weight * (c.omega * o.ry + c.omw * cy) + c.omwt * y0;
  o.ds = (o.px - temp) / c.tau;
  return o;
}

struct DualZOut {
  double z, pz, rz;
};

// Row i: K3 (dual-z step, pdfp_kernels.jl:60-75) given s = (A x̄)_i.
__device__ __forceinline__ DualZOut dual_z_math(const StepScalars& c, double s, double cz,
                                               double z0, double b, bool eq) {
  DualZOut o;
  const double raw = cz + c.sigma * (s - b);
  o.pz = eq ? raw : (raw > 0.0 ? raw : 0.0);
  o.rz = 2.0 * o.pz - cz;
  o.z = c.weight * (c.omega * o.rz + c.omw * cz) + c.omwt * z0;
  return o;
}

// ---- fused epilogues (precision-generic iterates T, FP64 arithmetic) ------
// Each epilogue loads its per-row operands in pre() (issued by the engine at
// the start of the pass) and does the math and the stores in operator().

// Column j of A (row j of Aᵀ): s[0] = (Aᵀz)_j.  Fuses K1 (primal step,
// pdfp_kernels.jl:4-20) and K2 (dual-y step, pdfp_kernels.jl:33-48).
// MAJOR steps also store pdhg_x, dual_slack, pdhg_y and reflected_y, which
// only the evaluation and the restart read.
template <typename T, bool MAJOR>
struct EpiPrimalY {
  struct Pre {
    T x, y, x0, y0, lb, ub, xt, w;
  };
  const StepParams* sp;
  int k_offset;
  T* __restrict__ x;
  const T* __restrict__ x0;
  T* __restrict__ y;
  const T* __restrict__ y0;
  const T* __restrict__ lb;
  const T* __restrict__ ub;
  const T* __restrict__ xt;
  const T* __restrict__ w;
  T* __restrict__ rx;
  T* __restrict__ px;
  T* __restrict__ ds;
  T* __restrict__ py;
  T* __restrict__ ry;
  __device__ __forceinline__ Pre pre(int j) const {
    return {x[j], y[j], x0[j], y0[j], lb[j], ub[j], xt[j], w[j]};
  }
  __device__ __forceinline__ void operator()(int j, const Pre& p, const double* s,
                                             double*) const {
    const StepScalars c = step_scalars(sp, k_offset);
    const PrimalYOut o = primal_y_math(c, s[0], (double)p.x, (double)p.y, (double)p.x0,
                                       (double)p.y0, (double)p.lb, (double)p.ub, (double)p.xt,
                                       (double)p.w);
    x[j] = (T)o.x;
    rx[j] = (T)o.rx;
    y[j] = (T)o.y;
    if (MAJOR) {
      px[j] = (T)o.px;
      ds[j] = (T)o.ds;
      py[j] = (T)o.py;
      ry[j] = (T)o.ry;
    }
  }
};

// Row i of A: s[0] = (A x̄)_i.  Fuses K3 (dual-z step, pdfp_kernels.jl:60-75).
template <typename T, bool MAJOR>
struct EpiDualZ {
  struct Pre {
    T z, z0, b;
    uint8_t eq;
  };
  const StepParams* sp;
  int k_offset;
  T* __restrict__ z;
  const T* __restrict__ z0;
  const T* __restrict__ b;
  const uint8_t* __restrict__ is_eq;
  T* __restrict__ pz;
  T* __restrict__ rz;
  __device__ __forceinline__ Pre pre(int i) const { return {z[i], z0[i], b[i], is_eq[i]}; }
  __device__ __forceinline__ void operator()(int i, const Pre& p, const double* s,
                                             double*) const {
    const StepScalars c = step_scalars(sp, k_offset);
    const DualZOut o = dual_z_math(c, s[0], (double)p.z, (double)p.z0, (double)p.b, p.eq != 0);
    z[i] = (T)o.z;
    if (MAJOR) {
      pz[i] = (T)o.pz;
      rz[i] = (T)o.rz;
    }
  }
};

// ---- evaluation epilogues (FP64 reductions) ---------------------------------
// Column pass over Aᵀ with GatherDelta(pdhg_z, reflected_z):
//   s[0] = (Aᵀ pdhg_z)_j, s[1] = (Aᵀ δz)_j, δ = pdhg - reflected.
// Terms: 0 ‖δx‖², 1 ‖δy‖², 2 ⟨Aᵀδz - δy, δx⟩        (src/pdfp.jl:167-178)
//        3 ‖Aᵀz - y - dual_slack‖²                    (src/pdfp.jl:207-209)
//        4 Σ_I w|x - x̃|, 5 Σ_I y(x̃ - x)               (src/pdfp.jl:211-214)
//        6 ‖x⁰ - x‖², 7 ‖y⁰ - y‖²                     (src/pdfp.jl:181-185)
constexpr int kEvalColTerms = 8;
template <typename T>
struct EpiEvalCol {
  struct Pre {
    T px, rx, py, ry, ds, xt, w, x0, y0;
  };
  const T* __restrict__ px;
  const T* __restrict__ rx;
  const T* __restrict__ py;
  const T* __restrict__ ry;
  const T* __restrict__ ds;
  const T* __restrict__ xt;
  const T* __restrict__ w;
  const T* __restrict__ x0;
  const T* __restrict__ y0;
  __device__ __forceinline__ Pre pre(int j) const {
    return {px[j], rx[j], py[j], ry[j], ds[j], xt[j], w[j], x0[j], y0[j]};
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

// Row pass over A with Gather1(pdhg_x): s[0] = (A pdhg_x)_i.
// Terms: 0 ‖δz‖², 1 ‖viol(Ax - b)‖² (scaled by 1/D_r when the primal residual
//        is measured in original units), 2 ⟨z, Ax - b⟩, 3 ‖z⁰ - z‖².
constexpr int kEvalRowTerms = 4;
template <typename T>
struct EpiEvalRow {
  struct Pre {
    T pz, rz, b, z0;
    uint8_t eq;
    double ru;
  };
  const T* __restrict__ pz;
  const T* __restrict__ rz;
  const T* __restrict__ b;
  const T* __restrict__ z0;
  const uint8_t* __restrict__ is_eq;
  const double* __restrict__ row_unscale;  // null: scaled rows
  __device__ __forceinline__ Pre pre(int i) const {
    return {pz[i], rz[i], b[i], z0[i], is_eq[i], row_unscale ? row_unscale[i] : 1.0};
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

// ---- element-wise kernels of the plain FP64 form (the reference's K1-K3) ---

// K1, src/pdfp_kernels.jl:4-20
static __global__ __launch_bounds__(256) void k_primal(const StepParams* __restrict__ sp, int k_offset,
                                                int n, double* __restrict__ px,
                                                double* __restrict__ ds, double* __restrict__ rx,
                                                double* __restrict__ x,
                                                const double* __restrict__ x0,
                                                const double* __restrict__ y,
                                                const double* __restrict__ atz,
                                                const double* __restrict__ lb,
                                                const double* __restrict__ ub) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const StepScalars c = step_scalars(sp, k_offset);
  const double cx = x[j];
  const double temp = cx - c.tau * (atz[j] - y[j]);
  const double p = jclamp(temp, lb[j], ub[j]);
  px[j] = p;
  ds[j] = (p - temp) / c.tau;
  const double r = 2.0 * p - cx;
  rx[j] = r;
  x[j] = c.weight * (c.omega * r + c.omw * cx) + c.omwt * x0[j];
}

// K2, src/pdfp_kernels.jl:33-48 (w_j = 0 off I)
static __global__ __launch_bounds__(256) void k_dual_y(const StepParams* __restrict__ sp, int k_offset,
                                                int n, double* __restrict__ py,
                                                double* __restrict__ y, double* __restrict__ ry,
                                                const double* __restrict__ y0,
                                                const double* __restrict__ xt,
                                                const double* __restrict__ rx,
                                                const double* __restrict__ w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const StepScalars c = step_scalars(sp, k_offset);
  const double cy = y[j];
  const double wj = w[j];
  const double v = wj > 0.0 ? jclamp(cy + c.sigma * (xt[j] - rx[j]), -wj, wj) : 0.0;
  py[j] = v;
  const double r = 2.0 * v - cy;
  ry[j] = r;
  y[j] = c.weight * (c.omega * r + c.omw * cy) + c.omwt * y0[j];
}

// K3, src/pdfp_kernels.jl:60-75
static __global__ __launch_bounds__(256) void k_dual_z(const StepParams* __restrict__ sp, int k_offset,
                                                int m, double* __restrict__ pz,
                                                double* __restrict__ z, double* __restrict__ rz,
                                                const double* __restrict__ z0,
                                                const double* __restrict__ ax,
                                                const double* __restrict__ b,
                                                const uint8_t* __restrict__ is_eq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const StepScalars c = step_scalars(sp, k_offset);
  const double cz = z[i];
  const double raw = cz + c.sigma * (ax[i] - b[i]);
  const double v = is_eq[i] ? raw : (raw > 0.0 ? raw : 0.0);
  pz[i] = v;
  const double r = 2.0 * v - cz;
  rz[i] = r;
  z[i] = c.weight * (c.omega * r + c.omw * cz) + c.omwt * z0[i];
}

// ---- restart / warm start / polish / conversions ----------------------------

// perform_restart!: initial ← pdhg, current ← pdhg for x, y (n) and z (m)
// (src/pdfp.jl:186-188).
template <typename T>
__global__ __launch_bounds__(256) void k_restart(int n, int m, const T* __restrict__ px,
                                                 const T* __restrict__ py,
                                                 const T* __restrict__ pz, T* __restrict__ x,
                                                 T* __restrict__ x0, T* __restrict__ y,
                                                 T* __restrict__ y0, T* __restrict__ z,
                                                 T* __restrict__ z0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const T a = px[i], c = py[i];
    x[i] = a; x0[i] = a;
    y[i] = c; y0[i] = c;
  }
  if (i < m) {
    const T a = pz[i];
    z[i] = a; z0[i] = a;
  }
}

template <typename TO, typename TI>
__global__ __launch_bounds__(256) void k_convert(int n, TO* __restrict__ out,
                                                 const TI* __restrict__ in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (TO)in[i];
}

// out[i] = in[i] in the iterate precision (round to nearest, as the host's
// conversion): the target arrives in FP64 and is narrowed on the device.
template <typename T>
__global__ __launch_bounds__(256) void k_narrow(int n, const double* __restrict__ in,
                                                T* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)in[i];
}

// A target in original units to scaled ones (x̃ ./ D_c, src/pump.jl:331):
// io holds x̃ on entry and x̃ ./ D_c (IEEE division, as the host's) on exit;
// out receives it in the iterates' precision.
template <typename T>
__global__ __launch_bounds__(256) void k_unscale_target(int n, double* __restrict__ io,
                                                        const double* __restrict__ dc,
                                                        T* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = io[i] / dc[i];
    io[i] = v;
    out[i] = (T)v;
  }
}

// FP64 polish of an FP32 projection: a coordinate that the primal prox
// clamped to a bound gets that bound's FP64 value (x = D_c·clamp(x̂, l/D_c,
// u/D_c) then holds exactly as in FP64); interior coordinates are widened.
// acc[0] = Σ_I w|x - x̃| of the polished point, in FP64.
static __global__ __launch_bounds__(256) void k_polish(int n, const float* __restrict__ px,
                                                const float* __restrict__ lb,
                                                const float* __restrict__ ub,
                                                const double* __restrict__ lb64,
                                                const double* __restrict__ ub64,
                                                const double* __restrict__ xt64,
                                                const double* __restrict__ w64,
                                                double* __restrict__ x64,
                                                double* __restrict__ partials) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (j < n) {
    const float p = px[j];
    const double v = p <= lb[j] ? lb64[j] : (p >= ub[j] ? ub64[j] : (double)p);
    x64[j] = v;
    const double wj = w64[j];
    if (wj > 0.0) acc[0] = wj * fabs(v - xt64[j]);
  }
  block_store_partials<256, 1>(acc, partials + blockIdx.x);
}

}  // namespace pdfp
