// Host-side helpers of libpdfp_hip: errors, device buffers, CSC -> CSR
// conversion and the row-block schedule of the CSR pass engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace pdfp {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PDFP_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::pdfp::Error(2, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

inline double wall_seconds() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Owning device allocation.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) PDFP_HIP(hipMalloc(&p, sizeof(T) * count));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return sizeof(T) * n; }
  void upload(const T* src, size_t count, hipStream_t s) {
    if (count) PDFP_HIP(hipMemcpyAsync(p, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
  }
  void download(T* dst, size_t count, hipStream_t s) const {
    if (count) PDFP_HIP(hipMemcpyAsync(dst, p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
  }
};

// Pinned host buffer.
template <typename T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void alloc(size_t count) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = count;
    if (count) PDFP_HIP(hipHostMalloc((void**)&p, sizeof(T) * count, hipHostMallocDefault));
  }
};

// Host CSR built from a CSC (0-based after the index base is removed).
struct HostCsr {
  int64_t nrows = 0, ncols = 0;
  std::vector<int32_t> rowptr, col;
  std::vector<double> val;
};

// CSC of an m×n matrix -> CSR of the same matrix (counting sort; the column
// indices of each row come out ascending).  Validates the CSC.
inline HostCsr csc_to_csr(int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowval,
                          const double* nzval, int32_t base) {
  if (m < 0 || n < 0) throw Error(1, "negative dimension");
  if (m > INT32_MAX - 1 || n > INT32_MAX - 1) throw Error(1, "dimension exceeds int32");
  const int64_t nnz = colptr[n] - base;
  if (colptr[0] != base) throw Error(1, "colptr[0] must equal the index base");
  if (nnz < 0 || nnz > INT32_MAX) throw Error(1, "nnz out of int32 range");
  HostCsr A;
  A.nrows = m;
  A.ncols = n;
  A.rowptr.assign((size_t)m + 1, 0);
  A.col.resize((size_t)nnz);
  A.val.resize((size_t)nnz);
  for (int64_t j = 0; j < n; ++j) {
    const int64_t p0 = colptr[j] - base, p1 = colptr[j + 1] - base;
    if (p1 < p0 || p1 > nnz) throw Error(1, "colptr is not nondecreasing");
    for (int64_t p = p0; p < p1; ++p) {
      const int64_t i = rowval[p] - base;
      if (i < 0 || i >= m) throw Error(1, "row index out of range");
      A.rowptr[(size_t)i + 1]++;
    }
  }
  for (int64_t i = 0; i < m; ++i) A.rowptr[(size_t)i + 1] += A.rowptr[(size_t)i];
  std::vector<int32_t> next(A.rowptr.begin(), A.rowptr.end() - 1);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t p = colptr[j] - base; p < colptr[j + 1] - base; ++p) {
      const int64_t i = rowval[p] - base;
      const int32_t q = next[(size_t)i]++;
      A.col[(size_t)q] = (int32_t)j;
      A.val[(size_t)q] = nzval[p];
    }
  return A;
}

// The CSC itself read as the CSR of Aᵀ (column j of A is row j of Aᵀ).
inline HostCsr csc_as_transpose_csr(int64_t m, int64_t n, const int64_t* colptr,
                                    const int64_t* rowval, const double* nzval, int32_t base) {
  HostCsr T;
  T.nrows = n;
  T.ncols = m;
  const int64_t nnz = colptr[n] - base;
  T.rowptr.resize((size_t)n + 1);
  for (int64_t j = 0; j <= n; ++j) T.rowptr[(size_t)j] = (int32_t)(colptr[j] - base);
  T.col.resize((size_t)nnz);
  T.val.assign(nzval, nzval + nnz);
  for (int64_t p = 0; p < nnz; ++p) T.col[(size_t)p] = (int32_t)(rowval[p] - base);
  // rows of Aᵀ must be sorted for the reference's summation order; Julia
  // CSC (and scipy after sort_indices) already is.  Sort defensively.
  for (int64_t j = 0; j < n; ++j) {
    const int32_t a = T.rowptr[(size_t)j], b = T.rowptr[(size_t)j + 1];
    bool sorted = true;
    for (int32_t p = a + 1; p < b; ++p)
      if (T.col[(size_t)p] < T.col[(size_t)p - 1]) { sorted = false; break; }
    if (!sorted) {
      std::vector<std::pair<int32_t, double>> tmp;
      for (int32_t p = a; p < b; ++p) tmp.emplace_back(T.col[(size_t)p], T.val[(size_t)p]);
      std::stable_sort(tmp.begin(), tmp.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
      for (int32_t p = a; p < b; ++p) {
        T.col[(size_t)p] = tmp[(size_t)(p - a)].first;
        T.val[(size_t)p] = tmp[(size_t)(p - a)].second;
      }
    }
  }
  return T;
}

// Row-block schedule: rows [blocks[k], blocks[k+1]) form block k.  A block
// holds at most `max_rows` rows and at most `max_nnz` nonzeros, except that a
// row with more than `max_nnz` nonzeros forms a block of its own.
// Appends the block ends of rows [r0, r1) to `blocks` (whose last entry is r0).
inline void row_blocks_range(const std::vector<int32_t>& rowptr, int64_t r0, int64_t r1,
                             int max_rows, int max_nnz, std::vector<int32_t>& blocks) {
  int64_t r = r0;
  while (r < r1) {
    const int64_t start = r;
    int64_t acc = 0;
    while (r < r1 && r - start < max_rows) {
      const int64_t len = rowptr[(size_t)r + 1] - rowptr[(size_t)r];
      if (acc + len > max_nnz) break;
      acc += len;
      ++r;
    }
    if (r == start) ++r;  // a single row longer than max_nnz
    blocks.push_back((int32_t)r);
  }
}

inline std::vector<int32_t> row_blocks(const std::vector<int32_t>& rowptr, int max_rows,
                                       int max_nnz) {
  std::vector<int32_t> blocks;
  blocks.push_back(0);
  row_blocks_range(rowptr, 0, (int64_t)rowptr.size() - 1, max_rows, max_nnz, blocks);
  return blocks;
}

// Lanes per row for each block of a schedule: the power of two tpr ≤ 64 that
// minimises the block's longest group, where the BLOCK/tpr groups take the
// rows round-robin and a row costs ⌈len/tpr⌉ adds plus log2(tpr)+1 shuffle
// steps; ties go to the smaller tpr.  0 marks a block that is one row longer
// than max_nnz (the whole workgroup reduces it).
inline std::vector<uint8_t> block_lanes(const std::vector<int32_t>& rowptr,
                                        const std::vector<int32_t>& blocks, int block,
                                        int max_nnz) {
  const size_t nb = blocks.size() - 1;
  std::vector<uint8_t> tpr(nb, 1);
  std::vector<int64_t> group_cost;
  for (size_t k = 0; k < nb; ++k) {
    const int32_t r0 = blocks[k], r1 = blocks[k + 1];
    const int64_t cnt = rowptr[(size_t)r1] - rowptr[(size_t)r0];
    if (r1 - r0 == 1 && cnt > max_nnz) {
      tpr[k] = 0;
      continue;
    }
    int64_t best_cost = -1;
    int best = 1;
    for (int t = 1, lg = 0; t <= 64 && t <= block; t *= 2, ++lg) {
      const int groups = block / t;
      group_cost.assign((size_t)groups, 0);
      for (int32_t r = r0; r < r1; ++r) {
        const int64_t len = rowptr[(size_t)r + 1] - rowptr[(size_t)r];
        group_cost[(size_t)((r - r0) % groups)] += (len + t - 1) / t + lg + 1;
      }
      int64_t c = 0;
      for (int64_t v : group_cost) c = std::max(c, v);
      if (best_cost < 0 || c < best_cost) {
        best_cost = c;
        best = t;
      }
    }
    tpr[k] = (uint8_t)best;
  }
  return tpr;
}

}  // namespace pdfp
