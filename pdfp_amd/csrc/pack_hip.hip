// libpdfp_hip: a pack of K independent projector instances of uneven sizes.
//
// The reference solves independent instances one process at a time
// (scripts/solve.jl:121-148 runs them concurrently as separate processes).
// A pack concatenates the instances' rows and columns into one
// block-diagonal A (and Aᵀ) in HBM and runs every instance's PDHG
// iterations in the FP32 form's two fused launches per iteration; each
// instance keeps its own target, iterates, step size, primal weight, restart
// state and termination, with exactly the per-instance logic of pdfp_solve!
// (src/pdfp.jl:252-288).  No row block of the schedule crosses an instance,
// so the evaluation's per-block partials sum per instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pdfp_hip.h"
#include "engine.h"
#include "host_util.h"
#include "pack_kernels.h"
#include "pdfp_kernels.h"
#include "small_kernels.h"
#include "solver_common.h"

namespace pdfp {

struct PackInstance {
  int64_t m = 0, n = 0, nnz = 0;
  int64_t row0 = 0, col0 = 0;  // offsets of its rows and columns in the pack
  RestartState r;
  pdfp_params p{};
  double sv = NAN, step_size = 0.0;
  double b_norm = 0.0, b_norm_original = NAN, dual_den = 1.0, l1_den = 1.0;
  bool orig_primal = false;
  double relative_l1_gap = INFINITY, l1_distance = NAN;
  int termination = PDFP_UNSPECIFIED;
  int n_restarts = 0, n_evals = 0;
  double solve_seconds = 0.0, t_start = 0.0;
  bool do_restart = false;
  bool running = false;
  bool x64_valid = false;
  std::vector<double> h_xt, h_w, h_b;
  HostCsr ha, ht;  // this instance's operators, kept until its σ_max is known
  // its local CSRs in the one-workgroup path's arrays (element offsets)
  int64_t sa_ptr = 0, sa_nz = 0, st_ptr = 0, st_nz = 0;
  int lanes_a = 1, lanes_t = 1;
};

struct Pack {
  int K;
  int64_t M = 0, N = 0, nnz = 0;
  hipStream_t stream = nullptr;
  DevCsr<float> A, At;
  DevBuf<int32_t> col_inst, row_inst, seg_t, seg_a;
  DevBuf<float> b, lb, ub, w;
  DevBuf<float> xt, x, x0, y, y0, rx, px, ds, py, ry;  // N
  DevBuf<float> z, z0, pz, rz;                         // M
  DevBuf<uint8_t> is_eq, d_flag;
  DevBuf<double> row_unscale;
  DevBuf<double> lb64, ub64, w64, xt64, x64;
  DevBuf<double> part_t, part_a, part_v, sums;
  DevBuf<PackStep> d_sp;
  PinnedBuf<PackStep> h_sp;
  PinnedBuf<uint8_t> h_flag;
  PinnedBuf<double> h_sums;
  bool any_orig = false;
  // Each block of iterations, its evaluations and the readback of its sums
  // captured once as a HIP graph and replayed (graph_on), one graph per
  // (freq, restart evaluation) shape; the launches read every instance's
  // scalars from the step table that the host writes before each replay.
  struct BlockGraph {
    hipGraphExec_t exec = nullptr;
    int freq = 0;
  };
  BlockGraph graphs[2];
  bool graph_on = true;
  // Every instance fits in LDS (small_kernels.h): each block runs as one
  // launch with one workgroup per running instance, each iterating its own
  // instance as the single FP32 projector's one-workgroup path does; path 1
  // forces the general passes.
  bool small_ok = false;
  int path = 0;
  bool small_used = false;
  int64_t small_lds = 0;
  DevBuf<int32_t> s_ptr_a, s_col_a, s_ptr_t, s_col_t;
  DevBuf<float> s_val_a, s_val_t;
  DevBuf<StepParams> d_ssp;
  PinnedBuf<StepParams> h_ssp;
  DevBuf<SmallArgs> d_sargs;
  PinnedBuf<SmallArgs> h_sargs;
  std::vector<PackInstance> in;
  std::string err = "no error";

  Pack(int K_, const int64_t* m_, const int64_t* n_, const int64_t* const* colptr,
       const int64_t* const* rowval, const double* const* nzval, int32_t base,
       const double* const* b_, const uint8_t* const* is_eq_, const uint8_t* const* is_int_,
       const double* const* var_lb, const double* const* var_ub, const double* const* w_,
       int device)
      : K(K_), in((size_t)K_) {
    PDFP_HIP(hipSetDevice(device));
    PDFP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    std::vector<int64_t> rseg(1, 0), cseg(1, 0);
    for (int k = 0; k < K; ++k) {
      PackInstance& I = in[(size_t)k];
      I.m = m_[k];
      I.n = n_[k];
      I.row0 = M;
      I.col0 = N;
      M += I.m;
      N += I.n;
      rseg.push_back(M);
      cseg.push_back(N);
    }
    if (M > INT32_MAX - 1 || N > INT32_MAX - 1) throw Error(PDFP_ERR_ARGUMENT, "pack too large");
    // block-diagonal CSR of A (M × N) and of Aᵀ (N × M)
    HostCsr ha, ht;
    ha.nrows = M;
    ha.ncols = N;
    ht.nrows = N;
    ht.ncols = M;
    ha.rowptr.assign(1, 0);
    ht.rowptr.assign(1, 0);
    std::vector<int32_t> h_col_inst((size_t)std::max<int64_t>(N, 1), 0);
    std::vector<int32_t> h_row_inst((size_t)std::max<int64_t>(M, 1), 0);
    for (int k = 0; k < K; ++k) {
      PackInstance& I = in[(size_t)k];
      I.ha = csc_to_csr(I.m, I.n, colptr[k], rowval[k], nzval[k], base);
      I.ht = csc_as_transpose_csr(I.m, I.n, colptr[k], rowval[k], nzval[k], base);
      I.nnz = (int64_t)I.ha.col.size();
      if (nnz + I.nnz > INT32_MAX) throw Error(PDFP_ERR_ARGUMENT, "pack has too many nonzeros");
      auto append = [](HostCsr& dst, const HostCsr& src, int64_t col0) {
        const int32_t base_nz = dst.rowptr.back();
        for (size_t r = 1; r < src.rowptr.size(); ++r)
          dst.rowptr.push_back(base_nz + src.rowptr[r]);
        for (int32_t c : src.col) dst.col.push_back((int32_t)(c + col0));
        dst.val.insert(dst.val.end(), src.val.begin(), src.val.end());
      };
      append(ha, I.ha, I.col0);
      append(ht, I.ht, I.row0);
      nnz += I.nnz;
      for (int64_t j = 0; j < I.n; ++j) h_col_inst[(size_t)(I.col0 + j)] = k;
      for (int64_t i = 0; i < I.m; ++i) h_row_inst[(size_t)(I.row0 + i)] = k;
    }
    A.segments = rseg;
    At.segments = cseg;
    small_ok = K > 0;
    for (const PackInstance& I : in)
      small_ok = small_ok && I.m > 0 && I.n > 0 && I.nnz > 0 &&
                 small_lds_bytes(I.m, I.n, I.nnz) <= small_lds_max();
    if (small_ok) upload_small();
    A.upload(ha, stream);
    At.upload(ht, stream);
    autotune_schedule(A, N, stream);
    autotune_schedule(At, M, stream);

    // per-instance host data and the concatenated vectors
    std::vector<double> h_b((size_t)std::max<int64_t>(M, 1), 0.0);
    std::vector<double> h_lb((size_t)std::max<int64_t>(N, 1), 0.0);
    std::vector<double> h_ub((size_t)std::max<int64_t>(N, 1), 0.0);
    std::vector<double> h_wall((size_t)std::max<int64_t>(N, 1), 0.0);
    std::vector<uint8_t> h_eq((size_t)std::max<int64_t>(M, 1), 0);
    for (int k = 0; k < K; ++k) {
      PackInstance& I = in[(size_t)k];
      I.h_w.assign((size_t)I.n, 0.0);
      for (int64_t j = 0; j < I.n; ++j) {
        if (is_int_[k][j] && !(w_[k][j] > 0.0))
          throw Error(PDFP_ERR_ARGUMENT, "weights of integer columns must be positive");
        I.h_w[(size_t)j] = is_int_[k][j] ? w_[k][j] : 0.0;
        h_lb[(size_t)(I.col0 + j)] = var_lb[k][j];
        h_ub[(size_t)(I.col0 + j)] = var_ub[k][j];
        h_wall[(size_t)(I.col0 + j)] = I.h_w[(size_t)j];
      }
      I.h_b.assign(b_[k], b_[k] + I.m);
      double acc = 0.0;
      for (int64_t i = 0; i < I.m; ++i) {
        h_b[(size_t)(I.row0 + i)] = b_[k][i];
        h_eq[(size_t)(I.row0 + i)] = is_eq_[k][i];
        acc += b_[k][i] * b_[k][i];
      }
      I.b_norm = std::sqrt(acc);
      acc = 0.0;
      for (int64_t j = 0; j < I.n; ++j) acc += I.h_w[(size_t)j] * I.h_w[(size_t)j];
      I.dual_den = 1.0 + std::sqrt(acc);
      I.h_xt.assign((size_t)I.n, 0.0);
    }
    auto vecf = [&](DevBuf<float>& d, int64_t len, const double* src) {
      d.alloc((size_t)std::max<int64_t>(len, 1));
      std::vector<float> tmp((size_t)std::max<int64_t>(len, 1), 0.f);
      if (src)
        for (int64_t i = 0; i < len; ++i) tmp[(size_t)i] = (float)src[i];
      d.upload(tmp.data(), tmp.size(), stream);
      PDFP_HIP(hipStreamSynchronize(stream));
    };
    vecf(b, M, h_b.data());
    vecf(lb, N, h_lb.data());
    vecf(ub, N, h_ub.data());
    vecf(w, N, h_wall.data());
    for (DevBuf<float>* v : {&xt, &x, &x0, &y, &y0, &rx, &px, &ds, &py, &ry}) vecf(*v, N, nullptr);
    for (DevBuf<float>* v : {&z, &z0, &pz, &rz}) vecf(*v, M, nullptr);
    is_eq.alloc(h_eq.size());
    is_eq.upload(h_eq.data(), h_eq.size(), stream);
    col_inst.alloc(h_col_inst.size());
    col_inst.upload(h_col_inst.data(), h_col_inst.size(), stream);
    row_inst.alloc(h_row_inst.size());
    row_inst.upload(h_row_inst.data(), h_row_inst.size(), stream);
    auto vec64 = [&](DevBuf<double>& d, const double* src) {
      d.alloc((size_t)std::max<int64_t>(N, 1));
      std::vector<double> tmp((size_t)std::max<int64_t>(N, 1), 0.0);
      if (src) std::copy(src, src + N, tmp.begin());
      d.upload(tmp.data(), tmp.size(), stream);
      PDFP_HIP(hipStreamSynchronize(stream));
    };
    vec64(lb64, h_lb.data());
    vec64(ub64, h_ub.data());
    vec64(w64, h_wall.data());
    vec64(xt64, nullptr);
    vec64(x64, nullptr);
    part_v.alloc((size_t)std::max(grid_for(N), 1) + 8);
    sums.alloc((size_t)K * 32 + 16);
    h_sums.alloc((size_t)K * 32 + 16);
    std::memset(h_sums.p, 0, sizeof(double) * ((size_t)K * 32 + 16));
    sums.upload(h_sums.p, (size_t)K * 32 + 16, stream);
    d_sp.alloc((size_t)K);
    h_sp.alloc((size_t)K);
    for (int k = 0; k < K; ++k) h_sp.p[k] = PackStep{1.0, 1.0, 1.0, 0.0, 0.0, 0.0, {0, 0}};
    d_flag.alloc((size_t)K);
    h_flag.alloc((size_t)K);
    upload_segments();
    PDFP_HIP(hipStreamSynchronize(stream));
  }

  // The instances' own CSRs (local indices, FP32 values) for the
  // one-workgroup path, concatenated.
  void upload_small() {
    std::vector<int32_t> pa, ca, pt, ct;
    std::vector<float> va, vt;
    for (PackInstance& I : in) {
      I.sa_ptr = (int64_t)pa.size();
      I.sa_nz = (int64_t)ca.size();
      I.st_ptr = (int64_t)pt.size();
      I.st_nz = (int64_t)ct.size();
      pa.insert(pa.end(), I.ha.rowptr.begin(), I.ha.rowptr.end());
      ca.insert(ca.end(), I.ha.col.begin(), I.ha.col.end());
      for (double v : I.ha.val) va.push_back((float)v);
      pt.insert(pt.end(), I.ht.rowptr.begin(), I.ht.rowptr.end());
      ct.insert(ct.end(), I.ht.col.begin(), I.ht.col.end());
      for (double v : I.ht.val) vt.push_back((float)v);
      I.lanes_a = small_lanes(I.ha.rowptr, kSmallBlock);
      I.lanes_t = small_lanes(I.ht.rowptr, kSmallBlock);
      small_lds = std::max(small_lds, small_lds_bytes(I.m, I.n, I.nnz));
    }
    auto up = [&](auto& d, const auto& h) {
      d.alloc(h.size());
      d.upload(h.data(), h.size(), stream);
    };
    up(s_ptr_a, pa);
    up(s_col_a, ca);
    up(s_val_a, va);
    up(s_ptr_t, pt);
    up(s_col_t, ct);
    up(s_val_t, vt);
    d_ssp.alloc((size_t)K);
    h_ssp.alloc((size_t)K);
    d_sargs.alloc((size_t)K);
    h_sargs.alloc((size_t)K);
    sync();
  }

  // One block of `freq` iterations for every running instance: one launch,
  // one workgroup per running instance.
  void small_block(int freq, const std::vector<bool>& active) {
    int na = 0;
    for (int k = 0; k < K; ++k) {
      if (!active[(size_t)k]) continue;
      const PackInstance& I = in[(size_t)k];
      h_ssp.p[k] = StepParams{I.step_size, I.r.primal_weight, I.p.reflection_coefficient,
                              (double)I.r.inner_count};
      const int64_t r0 = I.row0, c0 = I.col0;
      SmallArgs& a = h_sargs.p[na++];
      a.A = SmallCsr{s_ptr_a.p + I.sa_ptr, s_col_a.p + I.sa_nz, s_val_a.p + I.sa_nz};
      a.At = SmallCsr{s_ptr_t.p + I.st_ptr, s_col_t.p + I.st_nz, s_val_t.p + I.st_nz};
      a.m = (int)I.m;
      a.n = (int)I.n;
      a.nnz = (int)I.nnz;
      a.lanes_a = I.lanes_a;
      a.lanes_t = I.lanes_t;
      a.freq = freq;
      a.first_major = I.do_restart ? 1 : 0;
      a.sp = d_ssp.p + k;
      a.x = x.p + c0;
      a.y = y.p + c0;
      a.z = z.p + r0;
      a.x0 = x0.p + c0;
      a.y0 = y0.p + c0;
      a.z0 = z0.p + r0;
      a.lb = lb.p + c0;
      a.ub = ub.p + c0;
      a.xt = xt.p + c0;
      a.w = w.p + c0;
      a.b = b.p + r0;
      a.is_eq = is_eq.p + r0;
      a.rx = rx.p + c0;
      a.px = px.p + c0;
      a.ds = ds.p + c0;
      a.py = py.p + c0;
      a.ry = ry.p + c0;
      a.pz = pz.p + r0;
      a.rz = rz.p + r0;
      a.ec = EpiEvalCol<float>{px.p + c0, rx.p + c0, py.p + c0, ry.p + c0, ds.p + c0,
                               xt.p + c0, w.p + c0, x0.p + c0, y0.p + c0};
      a.er = EpiEvalRow<float>{pz.p + r0, rz.p + r0, b.p + r0, z0.p + r0, is_eq.p + r0,
                               I.orig_primal ? row_unscale.p + r0 : nullptr};
      a.sums = sums.p + (size_t)k * 32;
    }
    if (na == 0) return;
    PDFP_HIP(hipMemcpyAsync(d_ssp.p, h_ssp.p, sizeof(StepParams) * (size_t)K,
                            hipMemcpyHostToDevice, stream));
    PDFP_HIP(hipMemcpyAsync(d_sargs.p, h_sargs.p, sizeof(SmallArgs) * (size_t)na,
                            hipMemcpyHostToDevice, stream));
    k_small_pack<kSmallBlock><<<na, kSmallBlock, (size_t)small_lds, stream>>>(d_sargs.p);
    PDFP_HIP(hipGetLastError());
    PDFP_HIP(hipMemcpyAsync(h_sums.p, sums.p, sizeof(double) * (size_t)K * 32,
                            hipMemcpyDeviceToHost, stream));
  }

  void set_path(int p) {
    if (p != 0 && p != 1) throw Error(PDFP_ERR_ARGUMENT, "kernel path must be 0 or 1");
    path = p;
  }

  ~Pack() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      drop_graphs();
      (void)hipStreamDestroy(stream);
    }
  }

  // The captured blocks hold kernel arguments (buffers, schedules, flags):
  // dropped whenever one of them changes.
  void drop_graphs() {
    for (BlockGraph& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = BlockGraph{};
    }
  }

  void sync() { PDFP_HIP(hipStreamSynchronize(stream)); }
  void check_k(int k) const {
    if (k < 0 || k >= K) throw Error(PDFP_ERR_ARGUMENT, "instance index out of range");
  }

  // The per-instance block ranges of the current schedules, and partial
  // buffers sized to them.
  void upload_segments() {
    seg_a.alloc(A.seg_blocks.size());
    seg_a.upload(A.seg_blocks.data(), A.seg_blocks.size(), stream);
    seg_t.alloc(At.seg_blocks.size());
    seg_t.upload(At.seg_blocks.data(), At.seg_blocks.size(), stream);
    part_t.alloc((size_t)std::max(At.nblocks, 1) * kEvalColTerms);
    part_a.alloc((size_t)std::max(A.nblocks, 1) * kEvalRowTerms);
    sync();
  }

  // A slice of a concatenated FP32 vector from / to FP64 host data.
  void put(DevBuf<float>& dst, int64_t off, int64_t len, const double* src) {
    if (!len) return;
    std::vector<float> tmp((size_t)len);
    for (int64_t i = 0; i < len; ++i) tmp[(size_t)i] = (float)src[i];
    PDFP_HIP(hipMemcpyAsync(dst.p + off, tmp.data(), sizeof(float) * (size_t)len,
                            hipMemcpyHostToDevice, stream));
    sync();
  }
  void get(const DevBuf<float>& src, int64_t off, int64_t len, double* out) {
    if (!len) return;
    std::vector<float> tmp((size_t)len);
    PDFP_HIP(hipMemcpyAsync(tmp.data(), src.p + off, sizeof(float) * (size_t)len,
                            hipMemcpyDeviceToHost, stream));
    sync();
    for (int64_t i = 0; i < len; ++i) out[i] = (double)tmp[(size_t)i];
  }

  // original_primal_residual! (src/pdfp.jl:85-90) for instance k.
  void set_original_primal_residual(int k, const double* D_r) {
    check_k(k);
    PackInstance& I = in[(size_t)k];
    if (!any_orig) {
      drop_graphs();  // the evaluation's epilogue takes row_unscale from now on
      std::vector<double> ones((size_t)std::max<int64_t>(M, 1), 1.0);
      row_unscale.alloc(ones.size());
      row_unscale.upload(ones.data(), ones.size(), stream);
      sync();
      any_orig = true;
    }
    std::vector<double> inv((size_t)std::max<int64_t>(I.m, 1));
    double acc = 0.0;
    for (int64_t i = 0; i < I.m; ++i) {
      inv[(size_t)i] = 1.0 / D_r[i];
      const double v = I.h_b[(size_t)i] / D_r[i];
      acc += v * v;
    }
    I.b_norm_original = std::sqrt(acc);
    if (I.m)
      PDFP_HIP(hipMemcpyAsync(row_unscale.p + I.row0, inv.data(), sizeof(double) * (size_t)I.m,
                              hipMemcpyHostToDevice, stream));
    sync();
    I.orig_primal = true;
  }

  void set_x(int k, const double* xs) {
    check_k(k);
    const PackInstance& I = in[(size_t)k];
    put(x, I.col0, I.n, xs);
    put(x0, I.col0, I.n, xs);
  }

  void warm_start(int k) {
    check_k(k);
    PackInstance& I = in[(size_t)k];
    const int64_t len = std::max(I.n, I.m);
    if (len)
      k_pack_warm<<<grid_for(len), 256, 0, stream>>>(
          (int)I.n, (int)I.m, x0.p + I.col0, x.p + I.col0, y0.p + I.col0, y.p + I.col0,
          z0.p + I.row0, z.p + I.row0);
    sync();
    I.r.inner_count = 0;
    I.r.total_count = 0;
    I.r.fixed_point_error = INFINITY;
    I.r.last_trial_fixed_point_error = INFINITY;
    I.termination = PDFP_UNSPECIFIED;
    I.x64_valid = false;
  }

  void set_target(int k, const double* xts) {
    check_k(k);
    PackInstance& I = in[(size_t)k];
    put(xt, I.col0, I.n, xts);
    std::copy(xts, xts + I.n, I.h_xt.begin());
    warm_start(k);
  }

  template <bool MAJOR>
  void step_t(int k_offset) {
    EpiPackPrimalY<MAJOR> ex{d_sp.p, k_offset, col_inst.p, x.p, x0.p, y.p, y0.p, lb.p, ub.p,
                             xt.p, w.p, rx.p, px.p, ds.p, py.p, ry.p};
    launch_pass<0>(At, Gather1<float>{z.p}, ex, nullptr, stream);
    EpiPackDualZ<MAJOR> ez{d_sp.p, k_offset, row_inst.p, z.p, z0.p, b.p, is_eq.p, pz.p, rz.p};
    launch_pass<0>(A, Gather1<float>{rx.p}, ez, nullptr, stream);
  }
  void step(int k_offset, bool major) {
    if (major) step_t<true>(k_offset);
    else step_t<false>(k_offset);
  }

  // Every instance's evaluation terms into sums[k·32 + slot·16 ...] (the
  // layout of the single projector's evaluate()).
  void evaluate(int slot) {
    EpiEvalCol<float> ec{px.p, rx.p, py.p, ry.p, ds.p, xt.p, w.p, x0.p, y0.p};
    launch_pass<kEvalColTerms>(At, GatherDelta<float>{pz.p, rz.p}, ec, part_t.p, stream);
    EpiEvalRow<float> er{pz.p, rz.p, b.p, z0.p, is_eq.p, any_orig ? row_unscale.p : nullptr};
    launch_pass<kEvalRowTerms>(A, Gather1<float>{px.p}, er, part_a.p, stream);
    k_pack_finalize<<<K * kEvalColTerms, 256, 0, stream>>>(part_t.p, seg_t.p, kEvalColTerms,
                                                           slot * 16, sums.p);
    k_pack_finalize<<<K * kEvalRowTerms, 256, 0, stream>>>(part_a.p, seg_a.p, kEvalRowTerms,
                                                           slot * 16 + 8, sums.p);
  }

  struct Eval {
    double fpe, rel_primal, rel_dual, rel_gap, l1, primal_dist, dual_dist;
  };
  Eval read_eval(int k, int slot) const {
    const double* S = h_sums.p + (size_t)k * 32 + (size_t)slot * 16;
    const PackInstance& I = in[(size_t)k];
    const double pw = I.r.primal_weight;
    Eval e;
    const double movement = S[0] * pw + (S[1] + S[8]) / pw;
    e.fpe = std::sqrt(jmax(movement + 2.0 * I.step_size * S[2], 0.0));
    const double bn = I.orig_primal ? I.b_norm_original : I.b_norm;
    e.rel_primal = std::sqrt(S[9]) / (1.0 + bn);
    e.rel_dual = std::sqrt(S[3]) / I.dual_den;
    e.rel_gap = std::fabs(S[5] + S[10] - S[4]) / I.l1_den;
    e.l1 = S[4];
    e.primal_dist = std::sqrt(S[6]);
    e.dual_dist = std::sqrt(S[7] + S[11]);
    return e;
  }

  void upload_steps(double omega, const std::vector<bool>& active) {
    sync();  // the pinned staging is rewritten below
    for (int k = 0; k < K; ++k) {
      const PackInstance& I = in[(size_t)k];
      PackStep& s = h_sp.p[k];
      s.tau = I.step_size / I.r.primal_weight;
      s.sigma = I.step_size * I.r.primal_weight;
      s.omega = omega;
      s.omw = 1.0 - omega;
      s.inner_count = (double)I.r.inner_count;
      s.active = active[(size_t)k] ? 1.0 : 0.0;
    }
    PDFP_HIP(hipMemcpyAsync(d_sp.p, h_sp.p, sizeof(PackStep) * (size_t)K, hipMemcpyHostToDevice,
                            stream));
  }

  // estimate_max_singular_value (src/preconditioner.jl:108-125) of instance
  // k, on its own operators (the iteration of the single FP32 projector).
  void singular_value(int k, const pdfp_params& p) {
    PackInstance& I = in[(size_t)k];
    if (!std::isnan(I.sv)) return;
    if (I.nnz == 0) {
      I.sv = 0.0;
    } else {
      DevCsr<float> a, t;
      a.upload(I.ha, stream);
      t.upload(I.ht, stream);
      // as the single FP32 projector: in one workgroup when the pack's
      // blocks run that way
      I.sv = small_ok && path == 0
                 ? small_sigma(a, t, I.m, I.n, p.sv_max_iterations, p.sv_tolerance, stream)
                 : NAN;
      if (std::isnan(I.sv))
        I.sv = power_iteration(a, t, I.m, I.n, p.sv_max_iterations, p.sv_tolerance, stream);
      sync();
    }
    I.ha = HostCsr();
    I.ht = HostCsr();
  }

  // initialize_step_size_and_primal_weight! (src/pdfp.jl:220-241) for
  // instance k: it starts a solve with its own parameters.
  void begin(int k, const pdfp_params& p, double pw) {
    check_k(k);
    if (p.termination_evaluation_frequency < 1)
      throw Error(PDFP_ERR_ARGUMENT, "termination_evaluation_frequency must be >= 1");
    if (!(pw > 0.0)) throw Error(PDFP_ERR_ARGUMENT, "primal_weight must be > 0");
    PackInstance& I = in[(size_t)k];
    singular_value(k, p);
    I.step_size = I.nnz == 0 ? 1.0 : 0.998 / std::sqrt(1.0 + I.sv * I.sv);
    I.p = p;
    I.t_start = wall_seconds();
    I.r.primal_weight = pw;
    I.r.best_primal_weight = pw;
    I.r.primal_weight_error_sum = 0.0;
    I.r.primal_weight_last_error = 0.0;
    double acc = 0.0;
    for (int64_t j = 0; j < I.n; ++j)
      acc += I.h_w[(size_t)j] > 0.0 ? I.h_w[(size_t)j] * std::fabs(I.h_xt[(size_t)j]) : 0.0;
    I.l1_den = 1.0 + acc;
    I.do_restart = false;
    I.n_restarts = 0;
    I.n_evals = 0;
    I.termination = PDFP_UNSPECIFIED;
    I.running = true;
    I.x64_valid = false;
    if (I.r.total_count >= p.iteration_limit) {
      I.termination = PDFP_ITERATION_LIMIT;
      I.running = false;
      I.solve_seconds = 0.0;
    }
  }

  // One block: iteration 1 (major after a restart, and every instance
  // evaluated into slot 1), iterations 2 .. freq (freq major), the
  // evaluation into slot 0, the readback of the sums.
  void general_block(int freq, bool any_restart, bool readback = true) {
    step(1, any_restart || freq == 1);
    if (any_restart) evaluate(1);
    for (int i = 2; i <= freq; ++i) step(i, i == freq);
    evaluate(0);
    if (readback) read_block_sums();
  }

  void read_block_sums() {
    PDFP_HIP(hipMemcpyAsync(h_sums.p, sums.p, sizeof(double) * (size_t)K * 32,
                            hipMemcpyDeviceToHost, stream));
  }

  // general_block as a captured graph, captured on first use of its shape
  // and replayed after; it holds kernel nodes only (no memset, no copy): the
  // readback follows each replay on the stream.
  void block_graph(int freq, bool any_restart) {
    BlockGraph& g = graphs[any_restart ? 1 : 0];
    if (!g.exec || g.freq != freq) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = BlockGraph{};
      hipGraph_t graph = nullptr;
      PDFP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        general_block(freq, any_restart, false);
      } catch (...) {
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      PDFP_HIP(hipStreamEndCapture(stream, &graph));
      const hipError_t st = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      PDFP_HIP(st);
      g.freq = freq;
    }
    PDFP_HIP(hipGraphLaunch(g.exec, stream));
    read_block_sums();
  }

  // The block loop of pdfp_solve! (src/pdfp.jl:252-288) for every running
  // instance, in lockstep blocks of `freq` iterations, each instance with
  // its own termination; returns after the first block in which one ends
  // (until_any) or when none runs.  Returns the number still running.
  // mode: 0 until none runs, 1 until one ends, 2 after one block.
  int run(int mode) {
    int freq = 0;
    double omega = 0.0;
    for (const PackInstance& I : in) {
      if (!I.running) continue;
      if (freq == 0) {
        freq = I.p.termination_evaluation_frequency;
        omega = I.p.reflection_coefficient;
      } else if (I.p.termination_evaluation_frequency != freq ||
                 I.p.reflection_coefficient != omega) {
        throw Error(PDFP_ERR_ARGUMENT,
                    "running instances must share the evaluation frequency and reflection");
      }
    }
    std::vector<bool> active((size_t)K, false);
    auto refresh = [&] {
      bool any = false;
      for (int k = 0; k < K; ++k) {
        active[(size_t)k] = in[(size_t)k].running;
        any |= active[(size_t)k];
      }
      return any;
    };
    while (refresh()) {
      upload_steps(omega, active);
      bool any_restart = false;
      for (int k = 0; k < K; ++k) any_restart |= active[(size_t)k] && in[(size_t)k].do_restart;
      small_used = small_ok && path == 0;
      if (small_used) small_block(freq, active);
      else if (graph_on) block_graph(freq, any_restart);
      else general_block(freq, any_restart);
      sync();
      const double now = wall_seconds();
      bool any_flag = false, ended = false;
      for (int k = 0; k < K; ++k) {
        h_flag.p[k] = 0;
        if (!active[(size_t)k]) continue;
        PackInstance& I = in[(size_t)k];
        const pdfp_params& p = I.p;
        ++I.n_evals;
        if (I.do_restart) {
          I.r.fixed_point_error = read_eval(k, 1).fpe;
          I.r.initial_fixed_point_error = I.r.fixed_point_error;
          I.do_restart = false;
        }
        const Eval e = read_eval(k, 0);
        I.r.fixed_point_error = e.fpe;
        I.r.relative_primal_residual = e.rel_primal;
        I.r.relative_dual_residual = e.rel_dual;
        I.relative_l1_gap = e.rel_gap;
        I.l1_distance = e.l1;
        I.r.inner_count += freq;
        I.r.total_count += freq;
        bool stop = false;
        if (I.r.relative_primal_residual <= p.eps_feasible_relative &&
            I.r.relative_dual_residual <= p.eps_optimal_relative &&
            I.relative_l1_gap <= p.eps_optimal_relative) {
          I.termination = PDFP_OPTIMAL;
          stop = true;
        } else if (now - I.t_start > p.time_limit) {
          I.termination = PDFP_TIME_LIMIT;
          stop = true;
        } else {
          I.do_restart = should_restart(I.r, p);
          if (I.do_restart) {
            update_primal_weight(I.r, p, e.primal_dist, e.dual_dist);
            h_flag.p[k] = 1;
            any_flag = true;
            I.r.inner_count = 0;
            I.r.last_trial_fixed_point_error = INFINITY;
            ++I.n_restarts;
          }
          if (I.r.total_count >= p.iteration_limit) {
            I.termination = PDFP_ITERATION_LIMIT;
            stop = true;
          }
        }
        if (stop) {
          I.running = false;
          I.solve_seconds = now - I.t_start;
          ended = true;
        }
      }
      if (any_flag) {
        PDFP_HIP(hipMemcpyAsync(d_flag.p, h_flag.p, (size_t)K, hipMemcpyHostToDevice, stream));
        const int64_t len = std::max(N, M);
        k_pack_restart<<<grid_for(len), 256, 0, stream>>>(
            (int)N, (int)M, col_inst.p, row_inst.p, d_flag.p, px.p, py.p, pz.p, x.p, x0.p, y.p,
            y0.p, z.p, z0.p);
        sync();  // h_flag is rewritten in the next block
      }
      if (mode == 2 || (mode == 1 && ended)) break;
    }
    sync();
    int n_running = 0;
    for (const PackInstance& I : in) n_running += I.running ? 1 : 0;
    return n_running;
  }

  void solve(const pdfp_params& p, const double* pw, const uint8_t* act) {
    for (PackInstance& I : in) I.running = false;
    for (int k = 0; k < K; ++k)
      if (!act || act[k]) begin(k, p, pw[k]);
    run(0);
  }

  // The FP64 polish of instance k's projection and its ℓ1 distance.
  void polish(int k) {
    PackInstance& I = in[(size_t)k];
    if (I.x64_valid || I.n == 0) return;
    PDFP_HIP(hipMemcpyAsync(xt64.p + I.col0, I.h_xt.data(), sizeof(double) * (size_t)I.n,
                            hipMemcpyHostToDevice, stream));
    const int g = grid_for(I.n);
    k_polish<<<g, 256, 0, stream>>>((int)I.n, px.p + I.col0, lb.p + I.col0, ub.p + I.col0,
                                    lb64.p + I.col0, ub64.p + I.col0, xt64.p + I.col0,
                                    w64.p + I.col0, x64.p + I.col0, part_v.p);
    finalize_partials<256><<<1, 256, 0, stream>>>(part_v.p, g, 1, 0, sums.p + (size_t)K * 32);
    PDFP_HIP(hipMemcpyAsync(h_sums.p + (size_t)K * 32, sums.p + (size_t)K * 32, sizeof(double),
                            hipMemcpyDeviceToHost, stream));
    sync();
    I.l1_distance = h_sums.p[(size_t)K * 32];
    I.x64_valid = true;
  }

  void get_x(int k, double* out) {
    check_k(k);
    polish(k);
    const PackInstance& I = in[(size_t)k];
    if (I.n)
      PDFP_HIP(hipMemcpyAsync(out, x64.p + I.col0, sizeof(double) * (size_t)I.n,
                              hipMemcpyDeviceToHost, stream));
    sync();
  }

  void get_vector(int k, int which, double* out) {
    check_k(k);
    const PackInstance& I = in[(size_t)k];
    switch (which) {
      case 0: get_x(k, out); break;
      case 1: get(py, I.col0, I.n, out); break;
      case 2: get(pz, I.row0, I.m, out); break;
      case 3: get(x, I.col0, I.n, out); break;
      case 4: get(y, I.col0, I.n, out); break;
      case 5: get(z, I.row0, I.m, out); break;
      case 6: get(ds, I.col0, I.n, out); break;
      case 7: get(xt, I.col0, I.n, out); break;
      default: throw Error(PDFP_ERR_ARGUMENT, "unknown vector id");
    }
  }

  void info(int k, pdfp_info* o) {
    check_k(k);
    PackInstance& I = in[(size_t)k];
    if (I.termination != PDFP_UNSPECIFIED) polish(k);
    std::memset(o, 0, sizeof(*o));
    o->relative_primal_residual = I.r.relative_primal_residual;
    o->relative_dual_residual = I.r.relative_dual_residual;
    o->relative_l1_gap = I.relative_l1_gap;
    o->l1_distance = I.l1_distance;
    o->fixed_point_error = I.r.fixed_point_error;
    o->initial_fixed_point_error = I.r.initial_fixed_point_error;
    o->primal_weight = I.r.primal_weight;
    o->step_size = I.step_size;
    o->max_singular_value = I.sv;
    o->solve_seconds = I.solve_seconds;
    o->total_count = I.r.total_count;
    o->termination_reason = I.termination;
    o->n_restarts = I.n_restarts;
    o->precision = PDFP_PREC_FP32_MIXED;
    o->n_evaluations = I.n_evals;
    o->npb_a = A.npb;
    o->npb_at = At.npb;
    o->persist_a = A.persist;
    o->persist_at = At.persist;
    o->block_graph = graph_on && !small_used ? 1 : 0;
    o->small_path = small_used ? 1 : 0;
  }

  // Mean milliseconds of one pack iteration (every instance active, step
  // sizes from default-parameter power iterations), and of its two passes.
  void prepare_timing() {
    pdfp_params p;
    pdfp_default_params(&p);
    for (int k = 0; k < K; ++k) {
      PackInstance& I = in[(size_t)k];
      singular_value(k, p);
      I.step_size = I.nnz == 0 ? 1.0 : 0.998 / std::sqrt(1.0 + I.sv * I.sv);
    }
    std::vector<bool> active((size_t)K, true);
    upload_steps(1.0, active);
    step(1, false);
    sync();
  }

  double time_iterations(int iters) {
    prepare_timing();
    hipEvent_t e0, e1;
    PDFP_HIP(hipEventCreate(&e0));
    PDFP_HIP(hipEventCreate(&e1));
    PDFP_HIP(hipEventRecord(e0, stream));
    for (int i = 0; i < iters; ++i) step(2 + (i % 198), false);
    PDFP_HIP(hipEventRecord(e1, stream));
    PDFP_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PDFP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / std::max(iters, 1);
  }

  void time_kernels(int iters, double* ms_at, double* ms_a) {
    prepare_timing();
    std::vector<hipEvent_t> ev((size_t)iters * 4);
    for (auto& e : ev) PDFP_HIP(hipEventCreate(&e));
    for (int i = 0; i < iters; ++i) {
      const int k = 2 + (i % 198);
      EpiPackPrimalY<false> ex{d_sp.p, k, col_inst.p, x.p, x0.p, y.p, y0.p, lb.p, ub.p,
                               xt.p, w.p, rx.p, px.p, ds.p, py.p, ry.p};
      launch_pass<0>(At, Gather1<float>{z.p}, ex, nullptr, stream, ev[(size_t)i * 4 + 0],
                     ev[(size_t)i * 4 + 1]);
      EpiPackDualZ<false> ez{d_sp.p, k, row_inst.p, z.p, z0.p, b.p, is_eq.p, pz.p, rz.p};
      launch_pass<0>(A, Gather1<float>{rx.p}, ez, nullptr, stream, ev[(size_t)i * 4 + 2],
                     ev[(size_t)i * 4 + 3]);
    }
    sync();
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < iters; ++i) {
      float t0 = 0.f, t1 = 0.f;
      PDFP_HIP(hipEventElapsedTime(&t0, ev[(size_t)i * 4 + 0], ev[(size_t)i * 4 + 1]));
      PDFP_HIP(hipEventElapsedTime(&t1, ev[(size_t)i * 4 + 2], ev[(size_t)i * 4 + 3]));
      sa += t0;
      sb += t1;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms_at = sa / std::max(iters, 1);
    *ms_a = sb / std::max(iters, 1);
  }
};

}  // namespace pdfp

struct pdfp_pack {
  std::unique_ptr<pdfp::Pack> impl;
  std::string err = "no error";
};

namespace {

template <class F>
int pack_guarded(pdfp_pack* ctx, F&& f) {
  if (!ctx || !ctx->impl) return PDFP_ERR_ARGUMENT;
  try {
    f();
    return PDFP_OK;
  } catch (const pdfp::Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    ctx->err = "host allocation failed";
    return PDFP_ERR_NOMEM;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return PDFP_ERR_DEVICE;
  }
}

}  // namespace

extern "C" {

const char* pdfp_pack_last_error(const pdfp_pack* ctx) {
  if (!ctx) return pdfp::g_create_error.c_str();
  return ctx->err.c_str();
}

int pdfp_pack_create(pdfp_pack** out, int32_t K, const int64_t* m, const int64_t* n,
                     const int64_t* const* colptr, const int64_t* const* rowval,
                     const double* const* nzval, int32_t index_base, const double* const* b,
                     const uint8_t* const* is_eq, const uint8_t* const* is_int,
                     const double* const* var_lb, const double* const* var_ub,
                     const double* const* w, int32_t device) {
  if (!out) return PDFP_ERR_ARGUMENT;
  *out = nullptr;
  auto fail = [&](int code, const std::string& msg) {
    pdfp::g_create_error = msg;
    return code;
  };
  if (K < 1) return fail(PDFP_ERR_ARGUMENT, "K must be >= 1");
  if (!m || !n || !colptr || !rowval || !nzval || !b || !is_eq || !is_int || !var_lb ||
      !var_ub || !w)
    return fail(PDFP_ERR_ARGUMENT, "null input array");
  if (index_base != 0 && index_base != 1) return fail(PDFP_ERR_ARGUMENT, "index_base must be 0 or 1");
  for (int k = 0; k < K; ++k) {
    if (m[k] < 0 || n[k] < 0) return fail(PDFP_ERR_ARGUMENT, "negative size");
    if (!colptr[k]) return fail(PDFP_ERR_ARGUMENT, "null input array");
    const int64_t nnz_k = colptr[k][n[k]] - index_base;
    if (nnz_k < 0) return fail(PDFP_ERR_ARGUMENT, "colptr[n] before the first index");
    if ((nnz_k > 0 && (!rowval[k] || !nzval[k])) || (m[k] > 0 && (!b[k] || !is_eq[k])) ||
        (n[k] > 0 && (!is_int[k] || !var_lb[k] || !var_ub[k] || !w[k])))
      return fail(PDFP_ERR_ARGUMENT, "null input array");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PDFP_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(PDFP_ERR_ARGUMENT, "device index out of range");
  auto* ctx = new (std::nothrow) pdfp_pack();
  if (!ctx) return fail(PDFP_ERR_NOMEM, "host allocation failed");
  try {
    ctx->impl.reset(new pdfp::Pack(K, m, n, colptr, rowval, nzval, index_base, b, is_eq, is_int,
                                   var_lb, var_ub, w, device));
  } catch (const pdfp::Error& e) {
    delete ctx;
    return fail(e.code, e.what());
  } catch (const std::exception& e) {
    delete ctx;
    return fail(PDFP_ERR_DEVICE, e.what());
  }
  *out = ctx;
  return PDFP_OK;
}

int pdfp_pack_set_original_primal_residual(pdfp_pack* ctx, int32_t k, const double* D_r) {
  return pack_guarded(ctx, [&] {
    if (!D_r) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null D_r");
    ctx->impl->set_original_primal_residual(k, D_r);
  });
}

int pdfp_pack_set_x(pdfp_pack* ctx, int32_t k, const double* x_scaled) {
  return pack_guarded(ctx, [&] {
    if (!x_scaled) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null x");
    ctx->impl->set_x(k, x_scaled);
  });
}

int pdfp_pack_set_target(pdfp_pack* ctx, int32_t k, const double* x_tilde_scaled) {
  return pack_guarded(ctx, [&] {
    if (!x_tilde_scaled) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null target");
    ctx->impl->set_target(k, x_tilde_scaled);
  });
}

int pdfp_pack_warm_start(pdfp_pack* ctx, int32_t k) {
  return pack_guarded(ctx, [&] { ctx->impl->warm_start(k); });
}

int pdfp_pack_solve(pdfp_pack* ctx, const pdfp_params* params, const double* primal_weight,
                    const uint8_t* active, int32_t* termination, int64_t* total_count) {
  return pack_guarded(ctx, [&] {
    if (!params || !primal_weight) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null argument");
    pdfp::Pack& P = *ctx->impl;
    P.solve(*params, primal_weight, active);
    for (int k = 0; k < P.K; ++k) {
      if (termination) termination[k] = P.in[(size_t)k].termination;
      if (total_count) total_count[k] = P.in[(size_t)k].r.total_count;
    }
  });
}

int pdfp_pack_begin(pdfp_pack* ctx, int32_t k, const pdfp_params* params,
                    double primal_weight) {
  return pack_guarded(ctx, [&] {
    if (!params) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null params");
    ctx->impl->begin(k, *params, primal_weight);
  });
}

int pdfp_pack_run(pdfp_pack* ctx, int32_t until_any, int32_t* termination,
                  int64_t* total_count, int32_t* n_running) {
  return pack_guarded(ctx, [&] {
    pdfp::Pack& P = *ctx->impl;
    if (until_any < 0 || until_any > 2)
      throw pdfp::Error(PDFP_ERR_ARGUMENT, "until_any must be 0, 1 or 2");
    const int left = P.run(until_any);
    for (int k = 0; k < P.K; ++k) {
      const pdfp::PackInstance& I = P.in[(size_t)k];
      if (termination) termination[k] = I.running ? PDFP_UNSPECIFIED : I.termination;
      if (total_count) total_count[k] = I.r.total_count;
    }
    if (n_running) *n_running = left;
  });
}

int pdfp_pack_get_x(pdfp_pack* ctx, int32_t k, double* x_scaled_out) {
  return pack_guarded(ctx, [&] {
    if (!x_scaled_out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->get_x(k, x_scaled_out);
  });
}

int pdfp_pack_get_vector(pdfp_pack* ctx, int32_t k, int32_t which, double* out) {
  return pack_guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->get_vector(k, which, out);
  });
}

int pdfp_pack_set_kernel_path(pdfp_pack* ctx, int32_t path) {
  return pack_guarded(ctx, [&] { ctx->impl->set_path(path); });
}

int pdfp_pack_set_block_graph(pdfp_pack* ctx, int32_t on) {
  return pack_guarded(ctx, [&] {
    if (on != 0 && on != 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "block graph switch must be 0 or 1");
    ctx->impl->graph_on = on != 0;
  });
}

int pdfp_pack_get_info(pdfp_pack* ctx, int32_t k, pdfp_info* out) {
  return pack_guarded(ctx, [&] {
    if (!out) throw pdfp::Error(PDFP_ERR_ARGUMENT, "null output");
    ctx->impl->info(k, out);
  });
}

int pdfp_pack_time_iterations(pdfp_pack* ctx, int32_t iterations, double* ms_out) {
  return pack_guarded(ctx, [&] {
    if (!ms_out || iterations < 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    *ms_out = ctx->impl->time_iterations(iterations);
  });
}

int pdfp_pack_time_kernels(pdfp_pack* ctx, int32_t iterations, double* ms_at, double* ms_a) {
  return pack_guarded(ctx, [&] {
    if (!ms_at || !ms_a || iterations < 1) throw pdfp::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    ctx->impl->time_kernels(iterations, ms_at, ms_a);
  });
}

void pdfp_pack_destroy(pdfp_pack* ctx) { delete ctx; }

}  // extern "C"
