// libpdfp_hip: the pump's feasibility check on the device.
//
// max_violation(milp, x) (src/milp.jl:102-114): the largest violation of
// A x ≤ b / A x = b, called every pump iteration on the rounded point
// (src/pump.jl:321, 345, 367).  The original (unscaled) A stays resident in
// HBM as CSR in FP64; a check uploads x, runs one pass and reads back one
// scalar.  Each row's A x is summed by one thread in ascending column order,
// the order of the reference's CSC product (for each column j, y[i] += a·x_j),
// so the activities are bit-identical to the reference's; the maximum is
// order-free.  As in the reference, a NaN activity never raises the maximum.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdfp_hip.h"
#include "engine.h"
#include "host_util.h"

namespace pdfp {
namespace {

// Per workgroup: the largest violation of its rows.
__global__ __launch_bounds__(256) void k_violation(int m, const int32_t* __restrict__ ptr,
                                                   const int32_t* __restrict__ col,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ b,
                                                   const uint8_t* __restrict__ is_le,
                                                   double* __restrict__ partial) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (i < m) {
    double ax = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) ax += val[k] * x[col[k]];
    const double d = ax - b[i];
    const double r = is_le[i] ? d : fabs(d);
    if (r > v) v = r;
  }
  v = wave_max(v);
  __shared__ double red[256 / kWave];
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < 256 / kWave; ++w) s = red[w] > s ? red[w] : s;
    partial[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(256) void k_max_partials(int nb, const double* __restrict__ partial,
                                                      double* __restrict__ out) {
  double v = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) v = partial[k] > v ? partial[k] : v;
  v = wave_max(v);
  __shared__ double red[256 / kWave];
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < 256 / kWave; ++w) s = red[w] > s ? red[w] : s;
    out[0] = s;
  }
}

struct Checker {
  int64_t m, n;
  hipStream_t stream = nullptr;
  DevBuf<int32_t> ptr, col;
  DevBuf<double> val, b, x, partial, out;
  DevBuf<uint8_t> is_le;
  PinnedBuf<double> h_out, h_x;
  std::string err = "no error";

  Checker(int64_t m_, int64_t n_, const int64_t* colptr, const int64_t* rowval,
          const double* nzval, int32_t base, const double* b_, const uint8_t* is_le_, int device)
      : m(m_), n(n_) {
    HostCsr a = csc_to_csr(m, n, colptr, rowval, nzval, base);
    PDFP_HIP(hipSetDevice(device));
    PDFP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    ptr.alloc(a.rowptr.size());
    ptr.upload(a.rowptr.data(), a.rowptr.size(), stream);
    col.alloc(std::max<size_t>(a.col.size(), 1));
    val.alloc(std::max<size_t>(a.val.size(), 1));
    if (!a.col.empty()) {
      col.upload(a.col.data(), a.col.size(), stream);
      val.upload(a.val.data(), a.val.size(), stream);
    }
    b.alloc((size_t)std::max<int64_t>(m, 1));
    is_le.alloc((size_t)std::max<int64_t>(m, 1));
    if (m) {
      b.upload(b_, (size_t)m, stream);
      is_le.upload(is_le_, (size_t)m, stream);
    }
    x.alloc((size_t)std::max<int64_t>(n, 1));
    partial.alloc((size_t)std::max<int64_t>((m + 255) / 256, 1));
    out.alloc(1);
    h_out.alloc(1);
    h_x.alloc((size_t)std::max<int64_t>(n, 1));
    PDFP_HIP(hipStreamSynchronize(stream));
  }

  ~Checker() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }

  double max_violation(const double* xs) {
    if (m == 0) return 0.0;
    std::copy(xs, xs + n, h_x.p);
    if (n) PDFP_HIP(hipMemcpyAsync(x.p, h_x.p, sizeof(double) * (size_t)n,
                                   hipMemcpyHostToDevice, stream));
    const int nb = (int)((m + 255) / 256);
    k_violation<<<nb, 256, 0, stream>>>((int)m, ptr.p, col.p, val.p, x.p, b.p, is_le.p,
                                        partial.p);
    k_max_partials<<<1, 256, 0, stream>>>(nb, partial.p, out.p);
    PDFP_HIP(hipGetLastError());
    PDFP_HIP(hipMemcpyAsync(h_out.p, out.p, sizeof(double), hipMemcpyDeviceToHost, stream));
    PDFP_HIP(hipStreamSynchronize(stream));
    return h_out.p[0];
  }
};

std::string g_checker_error = "no error";

// Scratch of the allocation-free check: one partial maximum per workgroup of
// k_violation, so at most kNoAllocMaxBlocks · 256 rows.  Every call
// overwrites the first ceil(m / 256) entries before k_max_partials reads
// them, so nothing needs clearing between calls.  Each GPU holds its own copy
// for the life of the process; host code must use the per-device address
// from hipGetSymbolAddress, never the array's name.
constexpr int64_t kNoAllocMaxBlocks = 65536;
constexpr int64_t kNoAllocMaxRows = kNoAllocMaxBlocks * 256;
__device__ __align__(256) double scratch[kNoAllocMaxBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
double* scratch_on_current_device() {
  int dev = 0;
  PDFP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(PDFP_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    PDFP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(PDFP_ERR_DEVICE, "no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace pdfp

struct pdfp_checker {
  std::unique_ptr<pdfp::Checker> impl;
};

extern "C" {

const char* pdfp_checker_last_error(const pdfp_checker* c) {
  return c && c->impl ? c->impl->err.c_str() : pdfp::g_checker_error.c_str();
}

int pdfp_checker_create(pdfp_checker** out, int64_t m, int64_t n, const int64_t* colptr,
                        const int64_t* rowval, const double* nzval, int32_t index_base,
                        const double* b, const uint8_t* is_le, int32_t device) {
  try {
    if (!out || !colptr || (m > 0 && (!b || !is_le)))
      throw pdfp::Error(PDFP_ERR_ARGUMENT, "null argument");
    auto c = std::make_unique<pdfp_checker>();
    c->impl = std::make_unique<pdfp::Checker>(m, n, colptr, rowval, nzval, index_base, b, is_le,
                                              device);
    *out = c.release();
    return PDFP_OK;
  } catch (const pdfp::Error& e) {
    pdfp::g_checker_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    pdfp::g_checker_error = e.what();
    return PDFP_ERR_DEVICE;
  }
}

int pdfp_checker_max_violation(pdfp_checker* c, const double* x, double* out) {
  if (!c || !c->impl) return PDFP_ERR_STATE;
  try {
    if (!x || !out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null argument");
    *out = c->impl->max_violation(x);
    return PDFP_OK;
  } catch (const pdfp::Error& e) {
    c->impl->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c->impl->err = e.what();
    return PDFP_ERR_DEVICE;
  }
}

void pdfp_checker_destroy(pdfp_checker* c) { delete c; }

int pdfp_max_violation_noalloc(int64_t m, int64_t n, const int32_t* rowptr, const int32_t* col,
                               const double* val, const double* b, const uint8_t* is_le,
                               const double* x, double* out, void* stream) {
  try {
    if (m < 0 || n < 0) throw pdfp::Error(PDFP_ERR_ARGUMENT, "negative size");
    if (m > pdfp::kNoAllocMaxRows)
      throw pdfp::Error(PDFP_ERR_ARGUMENT,
                        "pdfp_max_violation_noalloc: m = " + std::to_string(m) +
                            " rows exceeds the maximum of " +
                            std::to_string(pdfp::kNoAllocMaxRows) + " (the scratch array)");
    if (n > INT32_MAX) throw pdfp::Error(PDFP_ERR_ARGUMENT, "n exceeds the int32 range");
    if (!out || (m > 0 && (!rowptr || !b || !is_le)))
      throw pdfp::Error(PDFP_ERR_ARGUMENT, "null argument");
    double* part = pdfp::scratch_on_current_device();
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nb = (int)((m + 255) / 256);
    if (nb > 0)
      pdfp::k_violation<<<nb, 256, 0, s>>>((int)m, rowptr, col, val, x, b, is_le, part);
    pdfp::k_max_partials<<<1, 256, 0, s>>>(nb, part, out);
    PDFP_HIP(hipGetLastError());
    return PDFP_OK;
  } catch (const pdfp::Error& e) {
    pdfp::g_checker_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    pdfp::g_checker_error = e.what();
    return PDFP_ERR_DEVICE;
  }
}

}  // extern "C"
