// libpdfp_probe (beside libpdfp_hip, not part of it): the random-gather
// ceiling of the GPU, for the benchmark's roofline.
//
// A CSR pass over a matrix whose columns are scattered at random gathers one
// record of the vector per nonzero.  Once the vector outgrows an XCD's L2,
// every gather costs a whole cache-line fill however few bytes it uses, and
// the pass is bound by the rate of line fills rather than by HBM bytes.  This
// probe measures that rate for the table size and record size of a pass: G
// gathers of R-byte records at uniformly random rows of a T-byte table, with
// the 4-byte row indices streamed like a matrix's column indices.  It has no
// counterpart in the reference.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <string>

#include "../../include/pdfp_hip.h"
#include "../../include/pdfp_probe.h"
#include "host_util.h"

namespace pdfp {
namespace {

__global__ __launch_bounds__(256) void k_probe_indices(int64_t g, int64_t rows,
                                                       int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g) return;
  uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[i] = (int32_t)(z % (uint64_t)rows);
}

// NF floats per record (1, 2, 4: one load; 8, 16: float4 loads).
template <int NF>
__global__ __launch_bounds__(256) void k_probe_gather(const float* __restrict__ table,
                                                      const int32_t* __restrict__ idx,
                                                      int64_t g, float* __restrict__ sink) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g) return;
  const float* rec = table + (size_t)idx[i] * NF;
  float acc = 0.f;
  if constexpr (NF >= 4) {
#pragma unroll
    for (int q = 0; q < NF / 4; ++q) {
      const float4 v = reinterpret_cast<const float4*>(rec)[q];
      acc += v.x + v.y + v.z + v.w;
    }
  } else if constexpr (NF == 2) {
    const float2 v = *reinterpret_cast<const float2*>(rec);
    acc = v.x + v.y;
  } else {
    acc = rec[0];
  }
  // the table holds zeros: never true, but the loads stay
  if (acc == 1234.5f) sink[0] = acc;
}

template <int NF>
void launch_probe(const float* table, const int32_t* idx, int64_t g, float* sink,
                  hipStream_t s) {
  const int64_t grid = (g + 255) / 256;
  k_probe_gather<NF><<<(unsigned)grid, 256, 0, s>>>(table, idx, g, sink);
}

double gather_ceiling(int64_t table_bytes, int32_t record_bytes, int64_t g, int device) {
  if (table_bytes < record_bytes || table_bytes / record_bytes > INT32_MAX || g <= 0 ||
      g > (int64_t)1 << 30)
    throw Error(PDFP_ERR_ARGUMENT, "gather probe: bad sizes");
  if (record_bytes != 4 && record_bytes != 8 && record_bytes != 16 && record_bytes != 32 &&
      record_bytes != 64)
    throw Error(PDFP_ERR_ARGUMENT, "gather probe: record_bytes must be 4, 8, 16, 32 or 64");
  PDFP_HIP(hipSetDevice(device));
  hipStream_t s;
  PDFP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const int64_t rows = table_bytes / record_bytes;
  DevBuf<float> table, sink;
  DevBuf<int32_t> idx;
  table.alloc((size_t)(rows * record_bytes / 4));
  sink.alloc(1);
  idx.alloc((size_t)g);
  PDFP_HIP(hipMemsetAsync(table.p, 0, (size_t)(rows * record_bytes), s));
  k_probe_indices<<<(unsigned)((g + 255) / 256), 256, 0, s>>>(g, rows, idx.p);
  auto run = [&] {
    switch (record_bytes) {
      case 4: launch_probe<1>(table.p, idx.p, g, sink.p, s); break;
      case 8: launch_probe<2>(table.p, idx.p, g, sink.p, s); break;
      case 16: launch_probe<4>(table.p, idx.p, g, sink.p, s); break;
      case 32: launch_probe<8>(table.p, idx.p, g, sink.p, s); break;
      default: launch_probe<16>(table.p, idx.p, g, sink.p, s); break;
    }
  };
  for (int w = 0; w < 3; ++w) run();
  hipEvent_t e0, e1;
  PDFP_HIP(hipEventCreate(&e0));
  PDFP_HIP(hipEventCreate(&e1));
  constexpr int kReps = 20;
  PDFP_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < kReps; ++r) run();
  PDFP_HIP(hipEventRecord(e1, s));
  PDFP_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  PDFP_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  PDFP_HIP(hipStreamSynchronize(s));
  (void)hipStreamDestroy(s);
  return (double)ms * 1e3 / kReps;
}

std::string g_probe_error = "no error";

}  // namespace
}  // namespace pdfp

extern "C" {

const char* pdfp_probe_last_error(void) { return pdfp::g_probe_error.c_str(); }

int pdfp_gather_ceiling(int64_t table_bytes, int32_t record_bytes, int64_t n_gathers,
                        int32_t device, double* us_per_pass) {
  try {
    if (!us_per_pass) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    *us_per_pass = pdfp::gather_ceiling(table_bytes, record_bytes, n_gathers, device);
    return PDFP_OK;
  } catch (const pdfp::Error& e) {
    pdfp::g_probe_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    pdfp::g_probe_error = e.what();
    return PDFP_ERR_DEVICE;
  }
}

}  // extern "C"
