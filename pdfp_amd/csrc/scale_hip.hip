// libpdfp_hip: diagonal preconditioning of the matrix on the device.
//
// scale_problem (src/preconditioner.jl:26-63), its matrix part: R rounds of
// Ruiz equilibration (each row and column divided by the square root of its
// largest |a_ij|), then one Pock–Chambolle pass with α = 1 (by the square
// root of its sum of |a_ij|), returning the scaled values in the CSC's own
// order and the diagonal scalings D_r, D_c.  The host forms the scaled
// vectors (c, bounds, right-hand sides) from D_r and D_c.
//
// Bit-identical to the reference's CPU loop: every value is scaled as
// v · (r_i · c_j), D_r and D_c take the same sequence of products, the
// maxima are order-free, and each Pock–Chambolle sum is accumulated by one
// thread in the reference's order (row i over ascending columns, column j in
// its stored order).
//
// Layout: the CSC values (column j = row j of Aᵀ) and a CSR copy of A, both
// FP64 and both scaled in place.  A round is two passes, rows of A and
// columns, each of which first applies the previous round's scalings to its
// copy of the values and then takes the maxima; a wavefront serves each row
// or column.  Scalings alternate between two buffers so that a pass reads
// the previous round's while writing its own.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pdfp_hip.h"
#include "engine.h"
#include "host_util.h"

namespace pdfp {
namespace {

constexpr double kScalingEpsilon = 1e-12;  // SCALING_EPSILON, src/preconditioner.jl:11

unsigned blocks_for(int64_t threads) { return (unsigned)((threads + 255) / 256); }

__device__ __forceinline__ double inv_sqrt_or_one(double v) {
  return v < kScalingEpsilon ? 1.0 : 1.0 / sqrt(v);
}

// One wavefront per row of a CSR (rows of A, or columns of A read as rows
// of Aᵀ): scale each value by prev_self[row] · prev_other[col] (when
// `apply`), store it, take max |v|, then scale_out[row] and D[row] *= it.
__global__ __launch_bounds__(256) void k_ruiz_pass(int nrows, const int32_t* __restrict__ ptr,
                                                   const int32_t* __restrict__ idx,
                                                   double* __restrict__ val, int apply,
                                                   const double* __restrict__ prev_self,
                                                   const double* __restrict__ prev_other,
                                                   double* __restrict__ scale_out,
                                                   double* __restrict__ D) {
  const int row = (int)(((int64_t)blockIdx.x * 256 + threadIdx.x) / kWave);
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= nrows) return;  // whole wavefronts leave together
  const int k0 = ptr[row], k1 = ptr[row + 1];
  const double rs = apply ? prev_self[row] : 1.0;
  double mx = 0.0;
  for (int k = k0 + lane; k < k1; k += kWave) {
    double v = val[k];
    if (apply) {
      v *= rs * prev_other[idx[k]];
      val[k] = v;
    }
    mx = fmax(mx, fabs(v));
  }
  mx = wave_max(mx);
  if (lane == 0) {
    const double s = inv_sqrt_or_one(mx);
    scale_out[row] = s;
    D[row] *= s;
  }
}

// Pock–Chambolle with α = 1: one thread per row, sequential sum of |v| in
// stored order (after applying the last Ruiz scalings).
__global__ __launch_bounds__(256) void k_pc_pass(int nrows, const int32_t* __restrict__ ptr,
                                                 const int32_t* __restrict__ idx,
                                                 double* __restrict__ val, int apply,
                                                 const double* __restrict__ prev_self,
                                                 const double* __restrict__ prev_other,
                                                 double* __restrict__ scale_out,
                                                 double* __restrict__ D) {
  const int row = (int)((int64_t)blockIdx.x * 256 + threadIdx.x);
  if (row >= nrows) return;
  const int k0 = ptr[row], k1 = ptr[row + 1];
  const double rs = apply ? prev_self[row] : 1.0;
  double sum = 0.0;
  for (int k = k0; k < k1; ++k) {
    double v = val[k];
    if (apply) {
      v *= rs * prev_other[idx[k]];
      val[k] = v;
    }
    sum += fabs(v);
  }
  const double s = inv_sqrt_or_one(sum);
  scale_out[row] = s;
  D[row] *= s;
}

// Final scaling of the CSC values: v · (r_i · c_j).
__global__ __launch_bounds__(256) void k_pc_apply(int ncols, const int32_t* __restrict__ ptr,
                                                  const int32_t* __restrict__ rowidx,
                                                  double* __restrict__ val,
                                                  const double* __restrict__ rs,
                                                  const double* __restrict__ cs) {
  const int j = (int)(((int64_t)blockIdx.x * 256 + threadIdx.x) / kWave);
  const int lane = threadIdx.x & (kWave - 1);
  if (j >= ncols) return;
  const double c = cs[j];
  for (int k = ptr[j] + lane; k < ptr[j + 1]; k += kWave) val[k] *= rs[rowidx[k]] * c;
}

__global__ __launch_bounds__(256) void k_fill(int len, double v, double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < len) out[i] = v;
}

void scale_matrix(int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowval,
                  const double* nzval, int32_t base, int ruiz_iterations, double alpha,
                  double* scaled, double* D_r_out, double* D_c_out, int device) {
  if (ruiz_iterations < 0) throw Error(PDFP_ERR_ARGUMENT, "ruiz_iterations must be >= 0");
  if (alpha != 1.0)
    throw Error(PDFP_ERR_ARGUMENT, "the device preconditioner implements pock_chambolle_alpha = 1");
  HostCsr ha = csc_to_csr(m, n, colptr, rowval, nzval, base);
  HostCsr ht = csc_as_transpose_csr(m, n, colptr, rowval, nzval, base);
  const int64_t nnz = (int64_t)ha.col.size();
  PDFP_HIP(hipSetDevice(device));
  // buffers first: the guard below is destroyed before them, so the stream
  // has drained when they are freed
  DevBuf<int32_t> a_ptr, a_col, t_ptr, t_row;
  DevBuf<double> a_val, t_val, Dr, Dc, rs[2], cs[2];
  hipStream_t s;
  PDFP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{s};
  auto up_i = [&](DevBuf<int32_t>& d, const std::vector<int32_t>& h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) d.upload(h.data(), h.size(), s);
  };
  auto up_d = [&](DevBuf<double>& d, const std::vector<double>& h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) d.upload(h.data(), h.size(), s);
  };
  up_i(a_ptr, ha.rowptr);
  up_i(a_col, ha.col);
  up_d(a_val, ha.val);
  up_i(t_ptr, ht.rowptr);
  up_i(t_row, ht.col);
  up_d(t_val, ht.val);
  for (DevBuf<double>* d : {&Dr, &rs[0], &rs[1]}) d->alloc((size_t)std::max<int64_t>(m, 1));
  for (DevBuf<double>* d : {&Dc, &cs[0], &cs[1]}) d->alloc((size_t)std::max<int64_t>(n, 1));
  if (m) k_fill<<<blocks_for(m), 256, 0, s>>>((int)m, 1.0, Dr.p);
  if (n) k_fill<<<blocks_for(n), 256, 0, s>>>((int)n, 1.0, Dc.p);
  const unsigned gw_m = blocks_for(m * kWave);  // one wavefront per row
  const unsigned gw_n = blocks_for(n * kWave);
  int cur = 0;  // index of the buffers holding the previous round's scalings
  for (int t = 0; t < ruiz_iterations; ++t) {
    const int apply = t > 0;
    const int nxt = cur ^ 1;
    if (m)
      k_ruiz_pass<<<gw_m, 256, 0, s>>>((int)m, a_ptr.p, a_col.p, a_val.p, apply, rs[cur].p,
                                       cs[cur].p, rs[nxt].p, Dr.p);
    if (n)
      k_ruiz_pass<<<gw_n, 256, 0, s>>>((int)n, t_ptr.p, t_row.p, t_val.p, apply, cs[cur].p,
                                       rs[cur].p, cs[nxt].p, Dc.p);
    cur = nxt;
  }
  {
    const int apply = ruiz_iterations > 0;
    const int nxt = cur ^ 1;
    if (m)
      k_pc_pass<<<blocks_for(m), 256, 0, s>>>((int)m, a_ptr.p, a_col.p, a_val.p, apply,
                                            rs[cur].p, cs[cur].p, rs[nxt].p, Dr.p);
    if (n)
      k_pc_pass<<<blocks_for(n), 256, 0, s>>>((int)n, t_ptr.p, t_row.p, t_val.p, apply,
                                            cs[cur].p, rs[cur].p, cs[nxt].p, Dc.p);
    cur = nxt;
  }
  if (n && nnz) k_pc_apply<<<gw_n, 256, 0, s>>>((int)n, t_ptr.p, t_row.p, t_val.p, rs[cur].p, cs[cur].p);
  PDFP_HIP(hipGetLastError());
  if (nnz) t_val.download(scaled, (size_t)nnz, s);
  if (m) Dr.download(D_r_out, (size_t)m, s);
  if (n) Dc.download(D_c_out, (size_t)n, s);
  PDFP_HIP(hipStreamSynchronize(s));
}

std::string g_scale_error = "no error";

}  // namespace
}  // namespace pdfp

extern "C" {

const char* pdfp_scale_last_error(void) { return pdfp::g_scale_error.c_str(); }

int pdfp_scale_matrix(int64_t m, int64_t n, int64_t nnz, const int64_t* colptr,
                      const int64_t* rowval, const double* nzval, int32_t index_base,
                      int32_t ruiz_iterations, double pock_chambolle_alpha, double* scaled_nzval,
                      double* D_r, double* D_c, int32_t device) {
  try {
    if (!colptr || (nnz > 0 && (!rowval || !nzval || !scaled_nzval)) || (m > 0 && !D_r) ||
        (n > 0 && !D_c))
      throw pdfp::Error(PDFP_ERR_ARGUMENT, "null argument");
    if (colptr[n] - index_base != nnz) throw pdfp::Error(PDFP_ERR_ARGUMENT, "nnz mismatch");
    pdfp::scale_matrix(m, n, colptr, rowval, nzval, index_base, ruiz_iterations,
                       pock_chambolle_alpha, scaled_nzval, D_r, D_c, device);
    return PDFP_OK;
  } catch (const pdfp::Error& e) {
    pdfp::g_scale_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    pdfp::g_scale_error = e.what();
    return PDFP_ERR_DEVICE;
  }
}

}  // extern "C"
