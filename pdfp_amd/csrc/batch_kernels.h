// Shared-matrix batch of up to 8 projector trajectories (FP32 iterates).
//
// K trajectories of one instance — the same A, b, bounds and weights, each
// with its own target x̃_k, iterates, primal weight and restart state — run
// their PDHG iterations together: every vector is stored interleaved as
// v[j·8 + k], so a nonzero's gather fetches the 8 trajectories' entries of
// column j in one 32-byte request and the matrix is streamed once for all
// of them.  The per-trajectory scalars (step, primal weight, inner count,
// active flag) live in device memory; an inactive trajectory's iterates are
// left untouched, so it keeps the result of its own last solve.
#pragma once

#include "engine.h"
#include "pdfp_kernels.h"

namespace pdfp {

constexpr int kBatch = 8;

struct BatchStep {
  double step_size;
  double primal_weight;
  double omega;
  double inner_count;
  double active;  // 1: iterate this trajectory, 0: leave it unchanged
  double pad[3];
};

__device__ __forceinline__ StepScalars batch_scalars(const BatchStep& sp, int k_offset) {
  StepScalars s;
  s.tau = sp.step_size / sp.primal_weight;
  s.sigma = sp.step_size * sp.primal_weight;
  const double k = sp.inner_count + (double)k_offset;
  s.weight = k / (k + 1.0);
  s.omega = sp.omega;
  s.omw = 1.0 - s.omega;
  s.omwt = 1.0 - s.weight;
  return s;
}

struct F8 {
  float v[kBatch];
};

__device__ __forceinline__ F8 load8(const float* p) {
  const float4 a = reinterpret_cast<const float4*>(p)[0];
  const float4 b = reinterpret_cast<const float4*>(p)[1];
  return F8{{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}

__device__ __forceinline__ void store8(float* p, const F8& f) {
  reinterpret_cast<float4*>(p)[0] = make_float4(f.v[0], f.v[1], f.v[2], f.v[3]);
  reinterpret_cast<float4*>(p)[1] = make_float4(f.v[4], f.v[5], f.v[6], f.v[7]);
}

// p_k = a · x[c·8 + k]  for the 8 trajectories (one 32-byte gather); the
// pass stages the 8 FP32 values (raw) rather than 8 FP64 products.
struct GatherBatch {
  static constexpr int NV = kBatch;
  static constexpr bool kRawStage = true;
  const float* __restrict__ x;
  __device__ __forceinline__ void operator()(int c, double a, double* p) const {
    const F8 f = load8(x + (size_t)c * kBatch);
#pragma unroll
    for (int q = 0; q < kBatch; ++q) p[q] = a * (double)f.v[q];
  }
  __device__ __forceinline__ void raw(int c, float* v) const {
    const float4* src = reinterpret_cast<const float4*>(x + (size_t)c * kBatch);
    float4* dst = reinterpret_cast<float4*>(v);
    dst[0] = src[0];
    dst[1] = src[1];
  }
};

// Column j: the fused primal and dual-y step of every active trajectory
// (src/pdfp.jl:119-143), s[k] = (Aᵀ z_k)_j.
template <bool MAJOR>
struct EpiBatchPrimalY {
  // x and y load at the start of the pass; x⁰, y⁰ and x̃ load in the
  // epilogue (prefetching them too costs registers, hence occupancy, and
  // measured slower).
  struct Pre {
    F8 x, y;
    float lb, ub, w;
  };
  const BatchStep* sp;
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
  __device__ __forceinline__ Pre pre(int j) const {
    const size_t o = (size_t)j * kBatch;
    return {load8(x + o), load8(y + o), lb[j], ub[j], w[j]};
  }
  __device__ __forceinline__ void operator()(int j, const Pre& p, const double* s,
                                             double*) const {
    const size_t o = (size_t)j * kBatch;
    const F8 px0 = load8(x0 + o), py0 = load8(y0 + o), pxt = load8(xt + o);
    F8 nx = p.x, ny = p.y, nr, npx, nds, npy, nry;
    const double lo = (double)p.lb, hi = (double)p.ub, wj = (double)p.w;
#pragma unroll
    for (int q = 0; q < kBatch; ++q) {
      const BatchStep& st = sp[q];
      const StepScalars c = batch_scalars(st, k_offset);
      const double cx = (double)p.x.v[q], cy = (double)p.y.v[q];
      const double temp = cx - c.tau * (s[q] - cy);
      const double pxq = jclamp(temp, lo, hi);
      const double rxq = 2.0 * pxq - cx;
      const double pyq = wj > 0.0 ? jclamp(cy + c.sigma * ((double)pxt.v[q] - rxq), -wj, wj) : 0.0;
      const double ryq = 2.0 * pyq - cy;
      const bool on = st.active != 0.0;
      nr.v[q] = on ? (float)rxq : p.x.v[q];
      if (on) {
        nx.v[q] = (float)(c.weight * (c.omega * rxq + c.omw * cx) + c.omwt * (double)px0.v[q]);
        ny.v[q] = (float)(c.weight * (c.omega * ryq + c.omw * cy) + c.omwt * (double)py0.v[q]);
      }
      npx.v[q] = (float)pxq;
      nds.v[q] = (float)((pxq - temp) / c.tau);
      npy.v[q] = (float)pyq;
      nry.v[q] = (float)ryq;
    }
    store8(x + o, nx);
    store8(y + o, ny);
    store8(rx + o, nr);
    if (MAJOR) {
      // an inactive trajectory keeps its pdhg values: merge per lane
      F8 opx = load8(px + o), ods = load8(ds + o), opy = load8(py + o), ory = load8(ry + o);
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        if (sp[q].active != 0.0) {
          opx.v[q] = npx.v[q];
          ods.v[q] = nds.v[q];
          opy.v[q] = npy.v[q];
          ory.v[q] = nry.v[q];
        }
      }
      store8(px + o, opx);
      store8(ds + o, ods);
      store8(py + o, opy);
      store8(ry + o, ory);
    }
  }
};

// Row i: the dual-z step of every active trajectory (src/pdfp.jl:146-156),
// s[k] = (A x̄_k)_i.
template <bool MAJOR>
struct EpiBatchDualZ {
  struct Pre {
    F8 z, z0;
    float b;
    uint8_t eq;
  };
  const BatchStep* sp;
  int k_offset;
  float* __restrict__ z;
  const float* __restrict__ z0;
  const float* __restrict__ b;
  const uint8_t* __restrict__ is_eq;
  float* __restrict__ pz;
  float* __restrict__ rz;
  __device__ __forceinline__ Pre pre(int i) const {
    const size_t o = (size_t)i * kBatch;
    return {load8(z + o), load8(z0 + o), b[i], is_eq[i]};
  }
  __device__ __forceinline__ void operator()(int i, const Pre& p, const double* s,
                                             double*) const {
    const size_t o = (size_t)i * kBatch;
    F8 nz = p.z, npz, nrz;
    const double bi = (double)p.b;
#pragma unroll
    for (int q = 0; q < kBatch; ++q) {
      const BatchStep& st = sp[q];
      const StepScalars c = batch_scalars(st, k_offset);
      const double cz = (double)p.z.v[q];
      const double raw = cz + c.sigma * (s[q] - bi);
      const double pzq = p.eq ? raw : (raw > 0.0 ? raw : 0.0);
      const double rzq = 2.0 * pzq - cz;
      if (st.active != 0.0)
        nz.v[q] = (float)(c.weight * (c.omega * rzq + c.omw * cz) + c.omwt * (double)p.z0.v[q]);
      npz.v[q] = (float)pzq;
      nrz.v[q] = (float)rzq;
    }
    store8(z + o, nz);
    if (MAJOR) {
      F8 opz = load8(pz + o), orz = load8(rz + o);
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        if (sp[q].active != 0.0) {
          opz.v[q] = npz.v[q];
          orz.v[q] = nrz.v[q];
        }
      }
      store8(pz + o, opz);
      store8(rz + o, orz);
    }
  }
};

// ---- evaluation of one trajectory k (strided views of the batch) ------------

struct Gather1Strided {
  static constexpr int NV = 1;
  const float* __restrict__ x;
  int k;
  __device__ __forceinline__ void operator()(int c, double a, double* p) const {
    p[0] = a * (double)x[(size_t)c * kBatch + k];
  }
};

struct GatherDeltaStrided {
  static constexpr int NV = 2;
  const float* __restrict__ x;
  const float* __restrict__ r;
  int k;
  __device__ __forceinline__ void operator()(int c, double a, double* p) const {
    const size_t o = (size_t)c * kBatch + k;
    const double xc = (double)x[o];
    p[0] = a * xc;
    p[1] = a * (xc - (double)r[o]);
  }
};

// The terms of EpiEvalCol (pdfp_kernels.h) for trajectory k.
struct EpiBatchEvalCol {
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
struct EpiBatchEvalRow {
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

// perform_restart! for trajectory k: initial ← pdhg, current ← pdhg.
static __global__ __launch_bounds__(256) void k_batch_restart(int n, int m, int k,
                                                              const float* __restrict__ px,
                                                              const float* __restrict__ py,
                                                              const float* __restrict__ pz,
                                                              float* __restrict__ x,
                                                              float* __restrict__ x0,
                                                              float* __restrict__ y,
                                                              float* __restrict__ y0,
                                                              float* __restrict__ z,
                                                              float* __restrict__ z0) {
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
static __global__ __launch_bounds__(256) void k_batch_warm(int n, int m, int k,
                                                           const float* __restrict__ x,
                                                           const float* __restrict__ y,
                                                           const float* __restrict__ z,
                                                           float* __restrict__ x0,
                                                           float* __restrict__ y0,
                                                           float* __restrict__ z0) {
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
static __global__ __launch_bounds__(256) void k_batch_scatter(int len, int k,
                                                             const float* __restrict__ in,
                                                             float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[(size_t)i * kBatch + k] = in[i];
}

// Reads column k of an interleaved vector into a dense FP32 vector.
static __global__ __launch_bounds__(256) void k_batch_gather(int len, int k,
                                                            const float* __restrict__ in,
                                                            float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[i] = in[(size_t)i * kBatch + k];
}

// FP64 polish of trajectory k's projection (see k_polish).
static __global__ __launch_bounds__(256) void k_batch_polish(int n, int k,
                                                             const float* __restrict__ px,
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
    const float p = px[(size_t)j * kBatch + k];
    const double v = p <= lb[j] ? lb64[j] : (p >= ub[j] ? ub64[j] : (double)p);
    x64[j] = v;
    const double wj = w64[j];
    if (wj > 0.0) acc[0] = wj * fabs(v - xt64[j]);
  }
  block_store_partials<256, 1>(acc, partials + blockIdx.x);
}

}  // namespace pdfp
