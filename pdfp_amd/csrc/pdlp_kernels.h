// Device epilogues of the root-LP solver: restarted Halpern PDHG with
// reflection for  min c'x  s.t.  con_lb ≤ A x ≤ con_ub,  var_lb ≤ x ≤ var_ub
// (src/pdlp.jl:115-153, GPU form src/pdlp_kernels.jl:4-63), fused into the
// two CSR passes of an iteration:
//   transpose pass (rows of Aᵀ = columns j):  s = (Aᵀy)_j,
//     x⁺ = clamp(x - τ(c - s), ℓ, u),  x̄ = 2x⁺ - x,  Halpern average
//   A pass (rows i):                          s = (A x̄)_i,
//     t = y/σ - s,  y⁺ = (t - clamp(t, -u_c, -ℓ_c))σ,  ȳ = 2y⁺ - y, Halpern
// The evaluation every 200 iterations computes, in FP64, the fixed-point
// error, the KKT residuals and objectives (src/pdlp.jl:157-165, 220-242) and,
// under cuPDLPx termination, the infeasibility rays (src/pdlp.jl:247-285).
#pragma once

#include "engine.h"
#include "pdfp_kernels.h"
#include "solver_common.h"

namespace pdfp {

__device__ __forceinline__ bool finite(double v) { return fabs(v) < INFINITY; }

// ---- the per-element math of one PDLP iteration (FP64) ---------------------
// Shared by the fused epilogues below and by the one-workgroup block kernel
// (small_lp_kernels.h), so that both run the same operations.

struct LpPrimalOut {
  double temp, p, r, x;  // dual slack (major steps): (p - temp) / τ
};

// K4 (src/pdlp_kernels.jl:4-21) for one column with s = (Aᵀy)_j.
__device__ __forceinline__ LpPrimalOut lp_primal_math(const StepScalars& sc, double s, double cx,
                                                      double x0, double c, double lb, double ub) {
  LpPrimalOut o;
  o.temp = cx - sc.tau * (c - s);
  o.p = jclamp(o.temp, lb, ub);
  o.r = 2.0 * o.p - cx;
  o.x = sc.weight * (sc.omega * o.r + sc.omw * cx) + sc.omwt * x0;
  return o;
}

struct LpDualOut {
  double p, r, y;
};

// K5 (src/pdlp_kernels.jl:35-51) for one row with s = (A x̄)_i.
__device__ __forceinline__ LpDualOut lp_dual_math(const StepScalars& sc, double s, double cy,
                                                  double y0, double clb, double cub) {
  LpDualOut o;
  const double t = cy / sc.sigma - s;
  o.p = (t - jclamp(t, -cub, -clb)) * sc.sigma;
  o.r = 2.0 * o.p - cy;
  o.y = sc.weight * (sc.omega * o.r + sc.omw * cy) + sc.omwt * y0;
  return o;
}

// Transpose pass of one iteration (K4, src/pdlp_kernels.jl:4-21).
template <typename T, bool MAJOR>
struct EpiLpPrimal {
  struct Pre {
    T x, x0, c, lb, ub;
  };
  const StepParams* sp;
  int k_offset;
  T* __restrict__ x;
  const T* __restrict__ x0;
  const T* __restrict__ c;
  const T* __restrict__ lb;
  const T* __restrict__ ub;
  T* __restrict__ rx;
  T* __restrict__ px;
  T* __restrict__ ds;
  __device__ __forceinline__ Pre pre(int j) const { return {x[j], x0[j], c[j], lb[j], ub[j]}; }
  __device__ __forceinline__ void operator()(int j, const Pre& q, const double* s,
                                             double*) const {
    const StepScalars sc = step_scalars(sp, k_offset);
    const LpPrimalOut o = lp_primal_math(sc, s[0], (double)q.x, (double)q.x0, (double)q.c,
                                         (double)q.lb, (double)q.ub);
    rx[j] = (T)o.r;
    x[j] = (T)o.x;
    if (MAJOR) {
      px[j] = (T)o.p;
      ds[j] = (T)((o.p - o.temp) / sc.tau);
    }
  }
};

// A pass of one iteration (K5, src/pdlp_kernels.jl:35-51).
template <typename T, bool MAJOR>
struct EpiLpDual {
  struct Pre {
    T y, y0, clb, cub;
  };
  const StepParams* sp;
  int k_offset;
  T* __restrict__ y;
  const T* __restrict__ y0;
  const T* __restrict__ clb;
  const T* __restrict__ cub;
  T* __restrict__ py;
  T* __restrict__ ry;
  __device__ __forceinline__ Pre pre(int i) const { return {y[i], y0[i], clb[i], cub[i]}; }
  __device__ __forceinline__ void operator()(int i, const Pre& q, const double* s,
                                             double*) const {
    const StepScalars sc = step_scalars(sp, k_offset);
    const LpDualOut o = lp_dual_math(sc, s[0], (double)q.y, (double)q.y0, (double)q.clb,
                                     (double)q.cub);
    y[i] = (T)o.y;
    if (MAJOR) {
      py[i] = (T)o.p;
      ry[i] = (T)o.r;
    }
  }
};

// ---- evaluation ------------------------------------------------------------

// The dual ray component of row i: (ȳ - y⁺) projected by the finiteness of
// the row bounds (src/pdlp.jl:249-253).
template <typename T>
__device__ __forceinline__ double dual_ray(const T* py, const T* ry, const T* clb, const T* cub,
                                           int i) {
  double v = (double)ry[i] - (double)py[i];
  if (!finite((double)clb[i])) v = fmin(v, 0.0);
  if (!finite((double)cub[i])) v = fmax(v, 0.0);
  return v;
}

// The primal ray component of column j: (x̄ - x⁺) projected by the
// finiteness of the variable bounds (src/pdlp.jl:249-251).
template <typename T>
__device__ __forceinline__ double primal_ray(const T* px, const T* rx, const T* lb, const T* ub,
                                             int j) {
  double v = (double)rx[j] - (double)px[j];
  if (finite((double)lb[j])) v = fmax(v, 0.0);
  if (finite((double)ub[j])) v = fmin(v, 0.0);
  return v;
}

// Transpose-pass gather: p0 = a·y⁺_i, p1 = a·(y⁺_i - ȳ_i), p2 = a·ray_i.
template <typename T, bool RAYS>
struct GatherLpCol {
  static constexpr int NV = RAYS ? 3 : 2;
  const T* __restrict__ py;
  const T* __restrict__ ry;
  const T* __restrict__ clb;
  const T* __restrict__ cub;
  __device__ __forceinline__ void operator()(int i, double a, double* p) const {
    const double pi = (double)py[i];
    p[0] = a * pi;
    p[1] = a * (pi - (double)ry[i]);
    if constexpr (RAYS) p[2] = a * dual_ray(py, ry, clb, cub, i);
  }
};

// A-pass gather: p0 = a·x⁺_j, p1 = a·ray_j (unnormalised).
template <typename T, bool RAYS>
struct GatherLpRow {
  static constexpr int NV = RAYS ? 2 : 1;
  const T* __restrict__ px;
  const T* __restrict__ rx;
  const T* __restrict__ lb;
  const T* __restrict__ ub;
  __device__ __forceinline__ void operator()(int j, double a, double* p) const {
    p[0] = a * (double)px[j];
    if constexpr (RAYS) p[1] = a * primal_ray(px, rx, lb, ub, j);
  }
};

// Column terms: 0 ‖δx‖², 1 ⟨Aᵀδy, δx⟩, 2 ‖(c - Aᵀy - ds)·col_unscale‖²,
// 3 c·x, 4 ds·x, 5 ‖x⁰ - x‖², 6 c·d, 7 dual-slack term of the dual ray
// objective; maxima: 8 max|d|, 9 max|dual ray infeasibility|.
constexpr int kLpColTerms = 10;
constexpr int kLpColMax = 2;
template <typename T, bool RAYS>
struct EpiLpEvalCol {
  struct Pre {
    T px, rx, ds, c, x0, lb, ub;
    double cu;
  };
  const T* __restrict__ px;
  const T* __restrict__ rx;
  const T* __restrict__ ds;
  const T* __restrict__ c;
  const T* __restrict__ x0;
  const T* __restrict__ lb;
  const T* __restrict__ ub;
  const double* __restrict__ col_unscale;  // null: scaled units
  __device__ __forceinline__ Pre pre(int j) const {
    return {px[j], rx[j], ds[j], c[j], x0[j], lb[j], ub[j], col_unscale ? col_unscale[j] : 1.0};
  }
  __device__ __forceinline__ void operator()(int, const Pre& q, const double* s,
                                             double* acc) const {
    const double pxj = (double)q.px;
    const double dx = pxj - (double)q.rx;
    acc[0] += dx * dx;
    acc[1] += s[1] * dx;
    const double cj = (double)q.c, dsj = (double)q.ds;
    double r = cj - s[0] - dsj;
    if (col_unscale) r *= q.cu;
    acc[2] += r * r;
    acc[3] += cj * pxj;
    acc[4] += dsj * pxj;
    const double ex = (double)q.x0 - pxj;
    acc[5] += ex * ex;
    if constexpr (RAYS) {
      const double lo = (double)q.lb, hi = (double)q.ub;
      double d = (double)q.rx - pxj;
      if (finite(lo)) d = fmax(d, 0.0);
      if (finite(hi)) d = fmin(d, 0.0);
      acc[6] += cj * d;
      acc[8] = fmax(acc[8], fabs(d));
      const double dp = s[2];
      acc[7] += fmax(-dp, 0.0) * (finite(lo) ? lo : 0.0) + fmin(-dp, 0.0) * (finite(hi) ? hi : 0.0);
      double v = (!finite(lo) ? fmax(0.0, -dp) : 0.0) - (!finite(hi) ? fmin(0.0, -dp) : 0.0);
      v *= q.cu;
      acc[9] = fmax(acc[9], fabs(v));
    }
  }
};

// Row terms: 0 ‖δy‖², 1 ‖(Ax - clamp(Ax, ℓ_c, u_c))·row_unscale‖², 2 the
// bound part of the dual objective, 3 ‖y⁰ - y‖², 4 primal-slack term of the
// dual ray objective; maxima: 5 max|dual ray|, 6 max|primal ray
// infeasibility| (unnormalised).
constexpr int kLpRowTerms = 7;
constexpr int kLpRowMax = 2;
template <typename T, bool RAYS>
struct EpiLpEvalRow {
  struct Pre {
    T py, ry, y0, clb, cub;
    double ru;
  };
  const T* __restrict__ py;
  const T* __restrict__ ry;
  const T* __restrict__ y0;
  const T* __restrict__ clb;
  const T* __restrict__ cub;
  const double* __restrict__ row_unscale;  // null: scaled units
  __device__ __forceinline__ Pre pre(int i) const {
    return {py[i], ry[i], y0[i], clb[i], cub[i], row_unscale ? row_unscale[i] : 1.0};
  }
  __device__ __forceinline__ void operator()(int, const Pre& q, const double* s,
                                             double* acc) const {
    const double pyi = (double)q.py;
    const double dy = pyi - (double)q.ry;
    acc[0] += dy * dy;
    const double lo = (double)q.clb, hi = (double)q.cub;
    double r = s[0] - jclamp(s[0], lo, hi);
    if (row_unscale) r *= q.ru;
    acc[1] += r * r;
    acc[2] += fmax(pyi, 0.0) * (fabs(lo) < 1e19 ? lo : 0.0) +
              fmin(pyi, 0.0) * (fabs(hi) < 1e19 ? hi : 0.0);
    const double ey = (double)q.y0 - pyi;
    acc[3] += ey * ey;
    if constexpr (RAYS) {
      double v = (double)q.ry - pyi;
      if (!finite(lo)) v = fmin(v, 0.0);
      if (!finite(hi)) v = fmax(v, 0.0);
      acc[4] += fmax(v, 0.0) * (finite(lo) ? lo : 0.0) + fmin(v, 0.0) * (finite(hi) ? hi : 0.0);
      acc[5] = fmax(acc[5], fabs(v));
      const double ad = s[1];
      double qv = (finite(lo) ? fmax(0.0, -ad) : 0.0) + (finite(hi) ? fmax(0.0, ad) : 0.0);
      qv *= q.ru;
      acc[6] = fmax(acc[6], fabs(qv));
    }
  }
};

// perform_restart!: initial ← pdhg, current ← pdhg (src/pdlp.jl:211-212).
template <typename T>
__global__ __launch_bounds__(256) void k_lp_restart(int n, int m, const T* __restrict__ px,
                                                    const T* __restrict__ py, T* __restrict__ x,
                                                    T* __restrict__ x0, T* __restrict__ y,
                                                    T* __restrict__ y0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const T a = px[i];
    x[i] = a;
    x0[i] = a;
  }
  if (i < m) {
    const T a = py[i];
    y[i] = a;
    y0[i] = a;
  }
}

}  // namespace pdfp
