// CSR row-pass engine for gfx950: one launch walks the rows of a CSR matrix
// in a precomputed row-block schedule, forms each row's inner product with a
// gathered vector (or two, or three), and hands the sums to a fused per-row
// epilogue.
//
// Schedule (built on the host by row_blocks in host_util.h): consecutive rows
// are packed into a block while it holds at most BLOCK rows and at most NPB
// nonzeros; a row longer than NPB sits alone in its block.  For every block
// the host also picks `tpr`, the lanes that reduce one row (a power of two,
// at most a wavefront), from the block's row lengths; tpr == 0 marks a
// single long row.  Each block's record carries its first row, its first
// nonzero and tpr, so the streaming starts after one dependent load.
//
//   stream mode  (tpr > 0): the block's contiguous slice of (col, val) is
//     streamed once, coalesced; the products a_ij·x_j are formed in FP64 and
//     staged in LDS together with the row offsets (a gather with many values
//     per nonzero stages the FP32 values and a_ij instead, 4 bytes each, and
//     the reduction forms the same FP64 products: less LDS per workgroup,
//     more workgroups per CU).  Groups of tpr lanes then
//     reduce the rows round-robin (xor-shuffles inside a group) into an LDS
//     row-sum array, and finally lane t runs the epilogue of row t: the
//     epilogue's operands were loaded into registers at the start of the
//     kernel, so their latency overlaps the streaming.
//   vector mode  (tpr == 0): all BLOCK lanes stride the one row and the sum
//     is a workgroup reduction.
//
// Products of FP32 operands are exact in FP64 and every row sum accumulates
// in FP64: on long power-law rows this replaces compensated FP32 summation.
// With tpr == 1 a row is summed in ascending column order, the order of the
// reference's CSC product.
//
// Per-block FP64 partials (NACC terms) of the epilogues' reductions are
// written to partials[block·NACC + k] (a lane epilogue's, per gathered value
// q, to partials[(q·NACC + k)·nblocks + block]) and combined in block order,
// so every reduction is deterministic run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pdfp {

constexpr int kWave = 64;

template <typename TV>
struct Csr {
  const int32_t* __restrict__ rowptr;  // nrows + 1
  const int32_t* __restrict__ col;     // nnz
  const TV* __restrict__ val;          // nnz
  // nblocks + 1 records {first row, first nonzero, lanes per row, mode}; the
  // last record is {nrows, nnz, 0, 0}.  One 16-byte load gives a block
  // everything it needs to start streaming.
  const int4* __restrict__ sched;
  int32_t nblocks;
  int32_t nrows;
};

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD gets a
// contiguous range of row blocks (neighbouring rows share L2 on banded data).
__device__ __forceinline__ int xcd_block(int bid, int nb) {
  constexpr int kXcd = 8;
  const int per = nb / kXcd;
  const int rem = nb % kXcd;
  const int xcd = bid % kXcd;
  const int idx = bid / kXcd;
  // XCDs 0..rem-1 own per+1 blocks, the others per.
  const int start = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
  return start + idx;
}

// Loads of data that a pass reads exactly once (the matrix slice).  Marking
// them non-temporal (PDFP_STREAM_NT=1) measured slower on MI355X at 5M
// nonzeros (101 µs vs 82 µs per iteration), so plain loads are the default.
#ifndef PDFP_STREAM_NT
#define PDFP_STREAM_NT 0
#endif
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
#if PDFP_STREAM_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// s + a·v for FP32 a and v: their product is exact in FP64 (24 + 24
// significand bits), so one fused multiply-add rounds exactly as the
// separate multiply and add do, in one instruction instead of two.
__device__ __forceinline__ double add_exact(double s, float a, float v) {
  return __builtin_fma((double)a, (double)v, s);
}

// Keeps v's words in registers, as loaded, until this point.  The epilogue
// operands are loaded before the row slice streams in; without this the
// compiler converts them right after their loads and so waits for them there,
// before the streaming starts.
template <class T>
__device__ __forceinline__ void hold(T& v) {
  if constexpr (sizeof(T) % 4 == 0) {
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) asm volatile("" : "+v"(w[i]));
    __builtin_memcpy(&v, w, sizeof(T));
  }
}

// Sum of v over the block, returned on thread 0 (other threads: garbage).
template <int BLOCK>
__device__ __forceinline__ double block_sum_to0(double v, double* red) {
  constexpr int kW = BLOCK / kWave;
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kW; ++w) s += red[w];
  }
  return s;
}

// Writes the block's NACC partials: terms [0, NACC-NMAX) are sums, the last
// NMAX terms are maxima (of nonnegative values).
template <int BLOCK, int NACC, int NMAX = 0>
__device__ __forceinline__ void block_store_partials(const double (&acc)[NACC],
                                                     double* __restrict__ out) {
  constexpr int kW = BLOCK / kWave;
  __shared__ double red[kW][NACC];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    const double v = k < NACC - NMAX ? wave_sum(acc[k]) : wave_max(acc[k]);
    if (lane == 0) red[wid][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NACC) {
    const bool is_max = (int)threadIdx.x >= NACC - NMAX;
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kW; ++w)
      s = is_max ? fmax(s, red[w][threadIdx.x]) : s + red[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// Per-value partials of a lane epilogue: thread t's terms belong to gathered
// value q = t % EL, and out[(q·NACC + k)·nb + b] receives the sum of term k
// over the block's threads of class q (sums only), so that a finalize reads
// each (q, k) series contiguously.
template <int BLOCK, int EL, int NACC>
__device__ __forceinline__ void lane_store_partials(const double (&acc)[NACC],
                                                    double* __restrict__ out, int nb, int b) {
  static_assert(kWave % EL == 0, "lane classes must tile a wavefront");
  __shared__ double red[BLOCK / kWave][EL][NACC];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    double v = acc[k];
    for (int o = EL; o < kWave; o <<= 1) v += __shfl_xor(v, o, kWave);
    if (lane < EL) red[wid][lane][k] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < EL * NACC) {
    const int q = threadIdx.x / NACC, k = threadIdx.x % NACC;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / kWave; ++w) v += red[w][q][k];
    out[(size_t)(q * NACC + k) * nb + b] = v;
  }
}

// ---- gathers: the per-nonzero products --------------------------------------

// p0 = a * x[c]
template <typename TX>
struct Gather1 {
  static constexpr int NV = 1;
  const TX* __restrict__ x;
  __device__ __forceinline__ void operator()(int c, double a, double* p) const {
    p[0] = a * (double)x[c];
  }
};

// p0 = a * x[c],  p1 = a * (x[c] - r[c])  (the difference is formed in FP64)
template <typename TX>
struct GatherDelta {
  static constexpr int NV = 2;
  const TX* __restrict__ x;
  const TX* __restrict__ r;
  __device__ __forceinline__ void operator()(int c, double a, double* p) const {
    const double xc = (double)x[c];
    p[0] = a * xc;
    p[1] = a * (xc - (double)r[c]);
  }
};

// A gather that sets `static constexpr bool kRawStage = true` also provides
// raw4(c, h), the gathered FP32 values 4h .. 4h+3 of column c (NV is a
// multiple of 4); the pass then stages them with the FP32 matrix value and
// forms the products a·v[q] in FP64 during the row reduction.  NV/4
// neighbouring lanes load one nonzero's values, so the NV·4 contiguous bytes
// of a column come from a single instruction (one cache-line request).
template <class G, class = void>
struct raw_stage : std::false_type {};
template <class G>
struct raw_stage<G, decltype((void)G::kRawStage)> : std::integral_constant<bool, G::kRawStage> {};

// ---- epilogue interface ------------------------------------------------------
// struct Epi {
//   struct Pre;                                        // per-row operands
//   __device__ Pre pre(int row) const;                 // loads, issued early
//   __device__ void operator()(int row, const Pre&, const double* sums,
//                              double* acc) const;     // math + stores
// };

struct NoPre {};

// An epilogue may instead serve each row with kLanes threads, one per
// gathered value (kLanes == Gather::NV): Pre pre(int row, int q) and
// operator()(int row, int q, const Pre&, double sum_q, double* acc).  The
// batch of trajectories uses it: lane q updates trajectory q of the row.
template <class E, class = void>
struct epi_lanes : std::integral_constant<int, 0> {};
template <class E>
struct epi_lanes<E, decltype((void)E::kLanes)> : std::integral_constant<int, E::kLanes> {};

// Mode field (the record's fourth) of a stream-mode block reduced one lane
// per (row, gathered value): lane (r, q) sums the products of row r with
// value q in ascending column order, then runs the epilogue of (r, q) on its
// own sum, with no shuffles and no second barrier.  Only lane epilogues over
// raw staging read it; every other pass over the same schedule reduces the
// block with its `tpr` lanes per row.  The host sets it on blocks whose rows
// are all short.
constexpr int kLaneRows = 1;

// ---- the pass ------------------------------------------------------------------

// The LDS of one pass, carved from one byte array so that a kernel running
// two passes back to back can give both the same storage:
//   stage   products double[NV][NPB + 1], or raw float[NPB + 1][NV] values
//           then float[NPB + 1] a_ij.  Slot NPB takes the stores of lanes
//           past the end of the slice, so the staging stores need no branch
//           (a branch would let the compiler sink each load into it, one
//           load in flight at a time);
//   rowsum  double[NV][MAXROWS], the reduced row sums;
//   red     double[BLOCK / 64], the workgroup reduction of a long row;
//   off     int[MAXROWS + 2], the row offsets (+ a pad slot for lanes past
//           the last row).
// A lane epilogue serves every (row, value) pair of a block in one pass of
// the workgroup: its schedule holds at most BLOCK / kLanes rows.
template <int BLOCK, int NPB, class Gather, class Epi>
struct PassLds {
  static constexpr int NV = Gather::NV;
  static constexpr bool kRaw = raw_stage<Gather>::value;
  static constexpr int NS = NPB + 1;
  static constexpr int MAXROWS =
      epi_lanes<Epi>::value > 0 ? BLOCK / epi_lanes<Epi>::value : BLOCK;
  static constexpr int kStageBytes = kRaw ? NS * (NV + 1) * 4 : NV * NS * 8;
  static constexpr int kRowsum = (kStageBytes + 15) / 16 * 16;
  static constexpr int kRed = kRowsum + NV * MAXROWS * 8;
  static constexpr int kOff = kRed + (BLOCK / kWave) * 8;
  static constexpr int kBytes = kOff + (MAXROWS + 2) * 4;
};

// One row block b of the schedule, processed by the whole workgroup, with
// the pass's LDS at `lds` (PassLds<...>::kBytes bytes, 16-byte aligned);
// tid is threadIdx.x.
template <int BLOCK, int NPB, typename TV, class Gather, class Epi, int NACC, int NMAX>
__device__ __forceinline__ void pass_block(const Csr<TV>& A, const Gather& g, const Epi& epi,
                                           double* __restrict__ partials, const int b,
                                           unsigned char* __restrict__ lds, const int tid) {
  using L = PassLds<BLOCK, NPB, Gather, Epi>;
  constexpr int NV = Gather::NV;
  constexpr bool kRaw = raw_stage<Gather>::value;
  // sweeps of the slice in flight per lane; a gather of more than 8 values
  // holds twice the registers per sweep, so it keeps half as many in flight
  // (more waves per SIMD instead)
  constexpr int U0 = (NPB / BLOCK) < 4 ? (NPB / BLOCK > 0 ? NPB / BLOCK : 1) : 4;
  constexpr int U = (kRaw && NV > 8) ? (U0 > 1 ? U0 / 2 : 1) : U0;
  // the long-row loop: a raw gather holds NV/4 float4 per nonzero in flight,
  // so it keeps 8 / (NV/4) nonzeros per lane in flight (more would cost the
  // whole kernel waves per SIMD); other gathers keep 4 (2 above 8 values)
  constexpr int UV = kRaw ? (NV >= 16 ? 1 : (NV >= 8 ? 2 : 4)) : (NV > 8 ? 2 : 4);
  constexpr int NA = NACC > 0 ? NACC : 1;
  static_assert(!kRaw || (std::is_same<TV, float>::value && NV % 4 == 0),
                "raw staging: FP32 matrix values, a multiple of 4 gathered values");
  constexpr int NS = L::NS;
  constexpr int MAXROWS = L::MAXROWS;
  double(*stage)[NS] = reinterpret_cast<double(*)[NS]>(lds);
  float* const sraw = reinterpret_cast<float*>(lds);
  float* const sa = sraw + NS * NV;
  double(*rowsum)[MAXROWS] = reinterpret_cast<double(*)[MAXROWS]>(lds + L::kRowsum);
  double* const red = reinterpret_cast<double*>(lds + L::kRed);
  int* const off = reinterpret_cast<int*>(lds + L::kOff);

  const int4 sb = A.sched[b];
  const int4 se = A.sched[b + 1];
  const int r0 = sb.x, r1 = se.x;
  const int base = sb.y;
  const int cnt = se.y - sb.y;
  const int tpr = sb.z;
  const int nrows = r1 - r0;

  // The epilogue's operands: in flight while the row slice streams in.
  constexpr int EL = epi_lanes<Epi>::value;
  static_assert(EL == 0 || EL == NV, "a lane epilogue takes one lane per gathered value");
  typename Epi::Pre pre{};
  if constexpr (EL > 0) {
    if (tid < nrows * EL) pre = epi.pre(r0 + tid / EL, tid % EL);
  } else {
    if (tid < nrows) pre = epi.pre(r0 + tid);
  }

  double acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0.0;

  if (tpr > 0) {
    // row offsets: loaded now, stored to LDS after the slice's loads are in
    // flight (stream mode holds at most MAXROWS ≤ BLOCK rows; off[nrows] = cnt)
    const int row_off = A.rowptr[r0 + (tid < nrows ? tid : nrows)] - base;
    if constexpr (kRaw) {
      // SPL lanes per nonzero, lane h of them staging values 4h .. 4h+3
      constexpr int SPL = NV / 4;
      constexpr int PER = BLOCK / SPL;
      constexpr int UR = U * SPL;
      const int h = tid % SPL;
      const int t = tid / SPL;
      for (int k0 = 0; k0 < cnt; k0 += PER * UR) {
        int cc[UR];
        float aa[UR];
        // every load of the sweep in flight at once: past the end of the
        // slice a lane reloads its last nonzero (never stored)
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          const int k = k0 + u * PER + t;
          const int kc = k < cnt ? k : cnt - 1;
          cc[u] = ld_stream(A.col + base + kc);
          aa[u] = ld_stream(A.val + base + kc);
        }
        float4 gv[UR];
#pragma unroll
        for (int u = 0; u < UR; ++u) gv[u] = g.raw4(cc[u], h);
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          const int k = k0 + u * PER + t;
          const int kw = k < cnt ? k : NPB;
          reinterpret_cast<float4*>(sraw + kw * NV)[h] = gv[u];
          sa[kw] = aa[u];  // the SPL lanes of a nonzero store the same value
        }
      }
    } else {
      for (int k0 = 0; k0 < cnt; k0 += BLOCK * U) {
        int cc[U];
        double vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + u * BLOCK + tid;
          const int kc = k < cnt ? k : cnt - 1;
          cc[u] = ld_stream(A.col + base + kc);
          vv[u] = (double)ld_stream(A.val + base + kc);
        }
        double p[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) g(cc[u], vv[u], p[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + u * BLOCK + tid;
          const int kw = k < cnt ? k : NPB;
#pragma unroll
          for (int q = 0; q < NV; ++q) stage[q][kw] = p[u][q];
        }
      }
    }
    off[tid < nrows ? tid : MAXROWS + 1] = row_off;
    if (tid == 0) off[nrows] = cnt;
    __syncthreads();
    bool lane_rows = false;
    if constexpr (EL > 0 && kRaw) lane_rows = sb.w == kLaneRows;
    if (lane_rows) {
      if constexpr (EL > 0 && kRaw) {
        if (tid < nrows * EL) {
          const int lr = tid / EL, q = tid % EL;
          double s = 0.0;
          const int k1 = off[lr + 1];
          // four nonzeros' LDS loads in flight at a time; the sum keeps the
          // ascending order
          int k = off[lr];
          for (; k + 4 <= k1; k += 4) {
            const float a0 = sa[k], a1 = sa[k + 1], a2 = sa[k + 2], a3 = sa[k + 3];
            const float v0 = sraw[k * NV + q], v1 = sraw[(k + 1) * NV + q];
            const float v2 = sraw[(k + 2) * NV + q], v3 = sraw[(k + 3) * NV + q];
            s = add_exact(s, a0, v0);
            s = add_exact(s, a1, v1);
            s = add_exact(s, a2, v2);
            s = add_exact(s, a3, v3);
          }
          for (; k < k1; ++k) s = add_exact(s, sa[k], sraw[k * NV + q]);
          hold(pre);
          epi(r0 + lr, q, pre, s, acc);
        }
      }
    } else {
      const int groups = BLOCK / tpr;
      const int grp = tid / tpr;
      const int lane = tid & (tpr - 1);
      // Every lane of a group walks the same rows, so the group stays
      // converged around its shuffles.
      for (int lr = grp; lr < nrows; lr += groups) {
        double s[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) s[q] = 0.0;
        const int k1 = off[lr + 1];
        int k = off[lr] + lane;
        if constexpr (!kRaw) {
          // four strides' products in flight at a time; the sum keeps its order
          for (; k + 3 * tpr < k1; k += 4 * tpr) {
            double p[4][NV];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
              for (int q = 0; q < NV; ++q) p[u][q] = stage[q][k + u * tpr];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
              for (int q = 0; q < NV; ++q) s[q] += p[u][q];
          }
        }
#pragma unroll 1
        for (; k < k1; k += tpr) {
          if constexpr (kRaw) {
            const float a = sa[k];
            const float4* v4 = reinterpret_cast<const float4*>(sraw + k * NV);
#pragma unroll
            for (int q4 = 0; q4 < NV / 4; ++q4) {
              const float4 v = v4[q4];
              s[4 * q4 + 0] = add_exact(s[4 * q4 + 0], a, v.x);
              s[4 * q4 + 1] = add_exact(s[4 * q4 + 1], a, v.y);
              s[4 * q4 + 2] = add_exact(s[4 * q4 + 2], a, v.z);
              s[4 * q4 + 3] = add_exact(s[4 * q4 + 3], a, v.w);
            }
          } else {
#pragma unroll
            for (int q = 0; q < NV; ++q) s[q] += stage[q][k];
          }
        }
        for (int o = tpr >> 1; o > 0; o >>= 1) {
#pragma unroll
          for (int q = 0; q < NV; ++q) s[q] += __shfl_xor(s[q], o, tpr);
        }
        if (lane == 0) {
#pragma unroll
          for (int q = 0; q < NV; ++q) rowsum[q][lr] = s[q];
        }
      }
      __syncthreads();
      hold(pre);
      if constexpr (EL > 0) {
        for (int e = tid; e < nrows * EL; e += BLOCK) {
          const int lr = e / EL, q = e % EL;
          const typename Epi::Pre p = e == tid ? pre : epi.pre(r0 + lr, q);
          epi(r0 + lr, q, p, rowsum[q][lr], acc);
        }
      } else if (tid < nrows) {
        double s[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) s[q] = rowsum[q][tid];
        epi(r0 + tid, pre, s, acc);
      }
    }
  } else {
    double s[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = 0.0;
    for (int k0 = 0; k0 < cnt; k0 += BLOCK * UV) {
      int cc[UV];
      TV a[UV];
#pragma unroll
      for (int u = 0; u < UV; ++u) {
        const int k = k0 + u * BLOCK + tid;
        const int kc = k < cnt ? k : cnt - 1;
        cc[u] = ld_stream(A.col + base + kc);
        a[u] = ld_stream(A.val + base + kc);
      }
      // Past the end a lane reloads the last nonzero and takes a_ij = 0: its
      // products are ±0.0, and adding them leaves s unchanged (s starts at
      // +0.0, and a sum from +0.0 never becomes -0.0).  Zeroing the matrix
      // value, not the products, keeps every gathered load unconditional:
      // a select on each product let the compiler sink loads into branches,
      // one gather in flight at a time.
#pragma unroll
      for (int u = 0; u < UV; ++u)
        if (k0 + u * BLOCK + tid >= cnt) a[u] = TV(0);
      if constexpr (kRaw) {
        // the gathered values stay FP32 in registers until their products
        float4 gv[UV][NV / 4];
#pragma unroll
        for (int u = 0; u < UV; ++u)
#pragma unroll
          for (int h = 0; h < NV / 4; ++h) gv[u][h] = g.raw4(cc[u], h);
#pragma unroll
        for (int u = 0; u < UV; ++u) {
#pragma unroll
          for (int h = 0; h < NV / 4; ++h) {
            s[4 * h + 0] = add_exact(s[4 * h + 0], a[u], gv[u][h].x);
            s[4 * h + 1] = add_exact(s[4 * h + 1], a[u], gv[u][h].y);
            s[4 * h + 2] = add_exact(s[4 * h + 2], a[u], gv[u][h].z);
            s[4 * h + 3] = add_exact(s[4 * h + 3], a[u], gv[u][h].w);
          }
        }
      } else {
        double p[UV][NV];
#pragma unroll
        for (int u = 0; u < UV; ++u) g(cc[u], (double)a[u], p[u]);
#pragma unroll
        for (int u = 0; u < UV; ++u)
#pragma unroll
          for (int q = 0; q < NV; ++q) s[q] += p[u][q];
      }
    }
    double t[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) t[q] = block_sum_to0<BLOCK>(s[q], red);
    hold(pre);
    if constexpr (EL > 0) {
      if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) rowsum[q][0] = t[q];
      }
      __syncthreads();
      if (tid < EL) epi(r0, tid, pre, rowsum[tid][0], acc);
    } else if (tid == 0) {
      epi(r0, pre, t, acc);
    }
  }
  if constexpr (NACC > 0) {
    if constexpr (EL > 0) {
      static_assert(NMAX == 0, "lane partials are sums");
      lane_store_partials<BLOCK, EL, NACC>(acc, partials, A.nblocks, b);
    } else {
      block_store_partials<BLOCK, NACC, NMAX>(acc, partials + (size_t)b * NACC);
    }
  }
}

// One workgroup per row block.
template <int BLOCK, int NPB, typename TV, class Gather, class Epi, int NACC, int NMAX = 0>
__global__ __launch_bounds__(BLOCK) void csr_pass(Csr<TV> A, Gather g, Epi epi,
                                                  double* __restrict__ partials) {
  __shared__ __align__(16) unsigned char lds[PassLds<BLOCK, NPB, Gather, Epi>::kBytes];
  pass_block<BLOCK, NPB, TV, Gather, Epi, NACC, NMAX>(A, g, epi, partials,
                                                      xcd_block(blockIdx.x, gridDim.x), lds,
                                                      threadIdx.x);
}

// Persistent walk: a grid of a few workgroups per CU (a multiple of 8) walks
// the row blocks; the workgroups on XCD x take the x-th contiguous eighth of
// the blocks, so each XCD's L2 serves a contiguous row range.
template <int BLOCK, int NPB, typename TV, class Gather, class Epi, int NACC, int NMAX>
__device__ __forceinline__ void walk_blocks(const Csr<TV>& A, const Gather& g, const Epi& epi,
                                            double* __restrict__ partials,
                                            unsigned char* __restrict__ lds) {
  constexpr int kXcd = 8;
  const int nb = A.nblocks;
  const int xcd = blockIdx.x % kXcd;
  const int local = blockIdx.x / kXcd;
  const int nlocal = gridDim.x / kXcd;
  const int per = nb / kXcd, rem = nb % kXcd;
  const int start = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
  const int count = per + (xcd < rem ? 1 : 0);
  for (int i = local; i < count; i += nlocal) {
    // tid passes through an empty asm statement in every block, so that
    // nothing computed from it is hoisted out of the loop (the hoisted values
    // would stay live in registers across every block: 15-30 more VGPRs and
    // one or two waves per SIMD fewer)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    pass_block<BLOCK, NPB, TV, Gather, Epi, NACC, NMAX>(A, g, epi, partials, start + i, lds, tid);
    __syncthreads();  // the next block reuses the LDS arrays
  }
}

template <int BLOCK, int NPB, typename TV, class Gather, class Epi, int NACC, int NMAX = 0>
__global__ __launch_bounds__(BLOCK) void csr_pass_persistent(Csr<TV> A, Gather g, Epi epi,
                                                             double* __restrict__ partials) {
  __shared__ __align__(16) unsigned char lds[PassLds<BLOCK, NPB, Gather, Epi>::kBytes];
  walk_blocks<BLOCK, NPB, TV, Gather, Epi, NACC, NMAX>(A, g, epi, partials, lds);
}

// Segmented form for a matrix stacked from 8 row segments (the column slices
// of one matrix; b[s] .. b[s+1] are segment s's row blocks): the
// workgroups with blockIdx % 8 == s take segment s's blocks.  Workgroups are
// dealt to the 8 XCDs round-robin, so each XCD walks one segment, and the
// gathers of a segment whose columns span an eighth of the vector stay in
// that XCD's L2.
struct SegBlocks {
  int32_t b[9];
};

template <int BLOCK, int NPB, typename TV, class Gather, class Epi, int NACC, int NMAX = 0>
__global__ __launch_bounds__(BLOCK) void csr_pass_xcd_segments(Csr<TV> A, Gather g, Epi epi,
                                                               double* __restrict__ partials,
                                                               SegBlocks sb) {
  __shared__ __align__(16) unsigned char lds[PassLds<BLOCK, NPB, Gather, Epi>::kBytes];
  const int s = blockIdx.x % 8;
  const int b = sb.b[s] + (int)(blockIdx.x / 8);
  if (b < sb.b[s + 1])
    pass_block<BLOCK, NPB, TV, Gather, Epi, NACC, NMAX>(A, g, epi, partials, b, lds, threadIdx.x);
}

// ---- grid barrier of a cooperative launch -------------------------------------

// State of a grid barrier in device memory, zeroed once when it is
// allocated; every word sits on a 256-byte line of its own, so that the
// arrivals of one group, the arrivals of the groups and the polls of the
// waiting workgroups do not contend for one line.  The workgroups arrive in
// 8 groups (blockIdx.x % 8): `group[g]` counts group g's arrivals at the
// current barrier, `top` the groups complete, `gen` the barriers passed.  The
// last workgroup of a group resets its group's count and arrives at `top`;
// the last group resets `top` and advances `gen`, so the state is ready for
// the next barrier and the next launch.  `err` is set when a wait gives up.
struct GridBarrier {
  static constexpr int kLine = 64;  // words per 256-byte line
  unsigned int group[8][kLine];
  unsigned int top[kLine];
  unsigned int gen[kLine];
  unsigned int err[kLine];
};

// Every workgroup of a cooperative launch (all resident at once; a multiple
// of 8 workgroups) waits here until all have arrived.  Each thread's stores
// before the barrier are visible to every thread of the grid after it: the
// workgroup barrier orders them before thread 0's release (agent scope,
// which writes back the XCD's L2) and its arrival; thread 0 acquires (agent
// scope, which invalidates the CU's and the XCD's stale lines) before the
// second workgroup barrier releases the other threads.  The wait is
// bounded: after `spin_limit` polls, or once another workgroup has given
// up, the workgroup sets `err` and returns false; the caller then leaves
// its loop, so every wave of the grid reaches the end of the kernel.
__device__ __forceinline__ bool grid_sync(GridBarrier* bar, unsigned int nwg,
                                          unsigned int spin_limit) {
  __shared__ int ok_s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned int g = blockIdx.x % 8u;
    const unsigned int gen = __hip_atomic_load(&bar->gen[0], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    bool advanced = false;
    const unsigned int a = __hip_atomic_fetch_add(&bar->group[g][0], 1u, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_AGENT);
    if (a == nwg / 8u - 1u) {
      __hip_atomic_store(&bar->group[g][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned int t = __hip_atomic_fetch_add(&bar->top[0], 1u, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_AGENT);
      if (t == 7u) {
        __hip_atomic_store(&bar->top[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&bar->gen[0], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        advanced = true;
      }
    }
    if (!advanced) {
      unsigned int spins = 0;
      // relaxed polls (an acquiring load would invalidate the L2 under the
      // workgroups still working); the acquire fence below orders the reads
      while (__hip_atomic_load(&bar->gen[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             gen) {
        ++spins;
        if (spins > spin_limit ||
            ((spins & 63u) == 0u && __hip_atomic_load(&bar->err[0], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT) != 0u)) {
          __hip_atomic_store(&bar->err[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok_s = ok;
  }
  __syncthreads();
  return ok_s != 0;
}

// Combines per-block partials (nb blocks × nacc terms) in block order into
// out[0..nacc): terms [0, nacc-nmax) are added, the last nmax are maxima.
// One workgroup.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void finalize_partials(const double* __restrict__ partials,
                                                           int nb, int nacc, int nmax,
                                                           double* __restrict__ out) {
  __shared__ double red[BLOCK / kWave];
  constexpr int kW = BLOCK / kWave;
  for (int k = 0; k < nacc; ++k) {
    const bool is_max = k >= nacc - nmax;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += BLOCK) {
      const double v = partials[(size_t)b * nacc + k];
      s = is_max ? fmax(s, v) : s + v;
    }
    if (!is_max) {
      s = block_sum_to0<BLOCK>(s, red);
    } else {
      s = wave_max(s);
      __syncthreads();
      if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = s;
      __syncthreads();
      if (threadIdx.x == 0)
        for (int w = 0; w < kW; ++w) s = fmax(s, red[w]);
    }
    if (threadIdx.x == 0) out[k] = s;
  }
}

}  // namespace pdfp
