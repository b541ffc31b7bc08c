// libpdfp_hip: the pump's rounding on the host, in C++.
//
// Constraint-propagation rounding (FP 2.0's CPR) and the guarded two-phase
// CPR of src/rounding.jl, over the bound propagation of src/propagation.jl.
// This is the host side of the pump: it stays on the CPU with the
// reference's arithmetic (every operation in the reference's order, no FMA
// contraction), so that a rounding is bit-identical to the reference's given
// the same LP point and the same ranking.  The caller ranks the variables
// (the only place the rounding draws random numbers) and keeps the timing
// policy; this file does the propagation-heavy part.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pdfp_hip.h"

namespace pdfp {
namespace rounding {

constexpr double kBoundTol = 1e-6;         // BOUND_TOLERANCE
constexpr double kInfeasTol = 1e-6;        // INFEASIBILITY_TOLERANCE
constexpr int kMaxTightenings = 10;        // wake-up cap per variable and direction
constexpr double kDomainInfinity = 1e12;   // stands in for infinite bounds

enum RowKind : uint8_t { kGeneral = 0, kCardinality = 1, kKnapsack = 2 };

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Julia's round (ties to even, the sign of zero kept) and clamp.
inline double jround(double v) { return std::nearbyint(v); }
inline double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }
inline double pmax(double a, double b) { return b > a ? b : a; }  // Python's max(a, b)
inline double pmin(double a, double b) { return b < a ? b : a; }  // Python's min(a, b)

// Everything about the propagation that depends on the model only
// (src/propagation.jl:19-84).  The propagation visits rows and variables in
// an order set by the model's graph, so it is bound by the latency of its
// scattered accesses: each row's and each variable's fields sit together
// (one or two cache lines per visit), and each column's and row's nonzeros
// are one array of (index, value) pairs.
struct Entry {
  int32_t idx;  // the row of a column's nonzero, the column of a row's
  double val;
};

// A variable's domain, its tightening counts and its integrality.
struct alignas(32) Var {
  double l, u;
  int32_t tlo, thi;
  int32_t is_int, pad;
};

// A row's activity bounds, the largest |a_ij|·(u_j - l_j) seen on its last
// scan, and its right-hand side.
struct alignas(32) Row {
  double L_min, L_max, maxroom, b;
};

enum : uint8_t { kLe = 1 };  // row flags: bit 0 is_le, bits 1-2 the RowKind

struct Data {
  int64_t n = 0, m = 0;
  std::vector<Var> var;             // initial domains, counts zero
  std::vector<Row> row;             // initial activity bounds
  std::vector<uint8_t> flags;       // is_le | kind << 1
  std::vector<double> b, row_maxcoef;
  std::vector<uint8_t> is_le;
  std::vector<int64_t> col_ptr, row_ptr;
  std::vector<Entry> col_nz, row_nz;
};

// Domains of one rounding pass and the FIFO propagation queue
// (src/propagation.jl:87-107).  A row is queued at most once at a time
// (in_queue), so the queue is a ring of m slots.
struct State {
  const Data& d;
  std::vector<Var> v;
  std::vector<Row> r;
  std::vector<uint8_t> in_queue;
  std::vector<int32_t> ring;
  size_t head = 0, count = 0;
  int64_t max_pops;
  bool infeasible = false;

  State(const Data& data, int64_t pops)
      : d(data), v(data.var), r(data.row), in_queue((size_t)data.m, 0),
        ring((size_t)data.m), max_pops(pops) {}

  double lo(int64_t j) const { return v[(size_t)j].l; }
  double hi(int64_t j) const { return v[(size_t)j].u; }

  void enqueue(int32_t i) {
    if (!in_queue[(size_t)i]) {
      size_t tail = head + count;
      if (tail >= ring.size()) tail -= ring.size();
      ring[tail] = i;
      ++count;
      in_queue[(size_t)i] = 1;
    }
  }

  bool is_fixed(int64_t j) const {
    const Var& x = v[(size_t)j];
    return std::fabs(x.u - x.l) < kBoundTol;
  }

  // src/propagation.jl:117-136
  bool tighten_upper(int64_t j, double new_u) {
    Var& x = v[(size_t)j];
    const double u_old = x.u;
    if (!(new_u < u_old - kBoundTol)) return false;
    double nu = x.is_int ? std::floor(new_u + kBoundTol) : new_u;
    if (nu < x.l - kBoundTol) nu = x.l;
    nu = pmax(nu, x.l);
    if (nu >= u_old - kBoundTol) return false;
    x.u = nu;
    x.thi += 1;
    const bool wake = x.thi <= kMaxTightenings || std::fabs(nu - x.l) < kBoundTol;
    const double du = nu - u_old;
    const Entry* e = d.col_nz.data();
    for (int64_t p = d.col_ptr[(size_t)j]; p < d.col_ptr[(size_t)j + 1]; ++p) {
      const int32_t i = e[p].idx;
      const double a = e[p].val;
      if (a > 0) r[(size_t)i].L_max += a * du;
      else r[(size_t)i].L_min += a * du;
      if (wake) enqueue(i);
    }
    return true;
  }

  // src/propagation.jl:138-156
  bool tighten_lower(int64_t j, double new_l) {
    Var& x = v[(size_t)j];
    const double l_old = x.l;
    if (!(new_l > l_old + kBoundTol)) return false;
    double nl = x.is_int ? std::ceil(new_l - kBoundTol) : new_l;
    if (nl > x.u + kBoundTol) nl = x.u;
    nl = pmin(nl, x.u);
    if (nl <= l_old + kBoundTol) return false;
    x.l = nl;
    x.tlo += 1;
    const bool wake = x.tlo <= kMaxTightenings || std::fabs(x.u - nl) < kBoundTol;
    const double dl = nl - l_old;
    const Entry* e = d.col_nz.data();
    for (int64_t p = d.col_ptr[(size_t)j]; p < d.col_ptr[(size_t)j + 1]; ++p) {
      const int32_t i = e[p].idx;
      const double a = e[p].val;
      if (a > 0) r[(size_t)i].L_min += a * dl;
      else r[(size_t)i].L_max += a * dl;
      if (wake) enqueue(i);
    }
    return true;
  }

  // src/propagation.jl:160-177
  void propagate_cardinality_row(int32_t i) {
    const size_t I = (size_t)i;
    const double b = r[I].b;
    const double slack_up = b - r[I].L_min;
    const double slack_dn = r[I].L_max - b;
    const bool is_eq = !(d.flags[I] & kLe);
    if (slack_up < -kInfeasTol || (is_eq && slack_dn < -kInfeasTol)) {
      infeasible = true;
      return;
    }
    const bool up = slack_up < 1.0 - kBoundTol;
    const bool dn = is_eq && slack_dn < 1.0 - kBoundTol;
    if (!(up || dn)) return;
    const Entry* e = d.row_nz.data();
    for (int64_t k = d.row_ptr[I]; k < d.row_ptr[I + 1]; ++k) {
      const int32_t j = e[k].idx;
      if (up) tighten_upper(j, v[(size_t)j].l + slack_up);
      if (dn) tighten_lower(j, v[(size_t)j].u - slack_dn);
    }
  }

  // src/propagation.jl:180-192
  void propagate_knapsack_row(int32_t i) {
    const size_t I = (size_t)i;
    const double slack_up = r[I].b - r[I].L_min;
    if (slack_up < -kInfeasTol) {
      infeasible = true;
      return;
    }
    if (d.row_maxcoef[I] <= slack_up + kBoundTol) return;
    const Entry* e = d.row_nz.data();
    for (int64_t k = d.row_ptr[I]; k < d.row_ptr[I + 1]; ++k) {
      const int32_t j = e[k].idx;
      if (e[k].val > slack_up + kBoundTol) tighten_upper(j, v[(size_t)j].l);
    }
  }

  // src/propagation.jl:194-232
  void propagate_row(int32_t i) {
    const size_t I = (size_t)i;
    const uint8_t f = d.flags[I];
    const uint8_t kind = f >> 1;
    if (kind == kCardinality) return propagate_cardinality_row(i);
    if (kind == kKnapsack) return propagate_knapsack_row(i);
    const bool is_le = f & kLe;
    Row& R = r[I];
    const double b = R.b;
    const double slack_up = b - R.L_min;
    const double slack_dn = R.L_max - b;
    if (slack_up < -kInfeasTol || (!is_le && slack_dn < -kInfeasTol)) {
      infeasible = true;
      return;
    }
    const double threshold = (is_le ? slack_up : pmin(slack_up, slack_dn)) + kBoundTol;
    if (R.maxroom <= threshold) return;
    const Entry* e = d.row_nz.data();
    double max_room = 0.0;
    for (int64_t k = d.row_ptr[I]; k < d.row_ptr[I + 1]; ++k) {
      const Var& x = v[(size_t)e[k].idx];
      const double room = std::fabs(e[k].val) * (x.u - x.l);
      if (room > max_room) max_room = room;
    }
    R.maxroom = max_room;
    if (is_le && max_room <= slack_up + kBoundTol) return;
    for (int64_t k = d.row_ptr[I]; k < d.row_ptr[I + 1]; ++k) {
      const int32_t j = e[k].idx;
      const double a = e[k].val;
      if (a > 0) tighten_upper(j, v[(size_t)j].l + slack_up / a);
      else tighten_lower(j, v[(size_t)j].u + slack_up / a);
      if (!is_le) {
        if (a > 0) tighten_lower(j, v[(size_t)j].u - slack_dn / a);
        else tighten_upper(j, v[(size_t)j].l - slack_dn / a);
      }
    }
  }

  // src/propagation.jl:235-244
  int64_t propagate() {
    int64_t pops = 0;
    while (count > 0 && pops < max_pops) {
      const int32_t i = ring[head];
      if (++head == ring.size()) head = 0;
      --count;
      in_queue[(size_t)i] = 0;
      propagate_row(i);
      ++pops;
    }
    return pops;
  }

  // src/propagation.jl:252-274
  bool fix(int64_t j, double value) {
    Var& x = v[(size_t)j];
    const double l_old = x.l, u_old = x.u;
    if (value < l_old - kBoundTol || value > u_old + kBoundTol) {
      infeasible = true;
      return false;
    }
    const double val = value > u_old ? u_old : (value < l_old ? l_old : value);
    if (std::fabs(val - l_old) < kBoundTol && std::fabs(val - u_old) < kBoundTol) return true;
    x.l = val;
    x.u = val;
    x.tlo += 1;
    x.thi += 1;
    const double dl = val - l_old;
    const double du = val - u_old;
    const Entry* e = d.col_nz.data();
    for (int64_t p = d.col_ptr[(size_t)j]; p < d.col_ptr[(size_t)j + 1]; ++p) {
      const int32_t i = e[p].idx;
      const double a = e[p].val;
      Row& R = r[(size_t)i];
      if (a > 0) {
        R.L_min += a * dl;
        R.L_max += a * du;
      } else {
        R.L_min += a * du;
        R.L_max += a * dl;
      }
      enqueue(i);
    }
    return true;
  }
};

// The model part of the propagation (src/propagation.jl:19-84).
Data build_data(int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowval,
                const double* nzval, int32_t base, const double* b, const uint8_t* is_le,
                const double* lo, const double* hi, const uint8_t* is_int) {
  if (m < 0 || n < 0 || m > INT32_MAX || n > INT32_MAX) throw Error(PDFP_ERR_ARGUMENT, "bad dimensions");
  if (colptr[0] != base) throw Error(PDFP_ERR_ARGUMENT, "colptr[0] must equal the index base");
  Data d;
  d.n = n;
  d.m = m;
  const int64_t nnz = colptr[n] - base;
  d.col_ptr.assign((size_t)n + 1, 0);
  d.col_nz.resize((size_t)nnz);
  for (int64_t j = 0; j <= n; ++j) d.col_ptr[(size_t)j] = colptr[j] - base;
  std::vector<int64_t> count((size_t)m + 1, 0);
  for (int64_t j = 0; j < n; ++j) {
    if (d.col_ptr[(size_t)j + 1] < d.col_ptr[(size_t)j] || d.col_ptr[(size_t)j + 1] > nnz)
      throw Error(PDFP_ERR_ARGUMENT, "colptr is not nondecreasing");
    for (int64_t p = d.col_ptr[(size_t)j]; p < d.col_ptr[(size_t)j + 1]; ++p) {
      const int64_t i = rowval[p] - base;
      if (i < 0 || i >= m) throw Error(PDFP_ERR_ARGUMENT, "row index out of range");
      d.col_nz[(size_t)p] = Entry{(int32_t)i, nzval[p]};
      count[(size_t)i + 1]++;
    }
  }
  // rows with ascending columns (counting sort over the columns in order)
  d.row_ptr.assign((size_t)m + 1, 0);
  for (int64_t i = 0; i < m; ++i) d.row_ptr[(size_t)i + 1] = d.row_ptr[(size_t)i] + count[(size_t)i + 1];
  d.row_nz.resize((size_t)nnz);
  std::vector<int64_t> next(d.row_ptr.begin(), d.row_ptr.end() - 1);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t p = d.col_ptr[(size_t)j]; p < d.col_ptr[(size_t)j + 1]; ++p) {
      const int64_t q = next[(size_t)d.col_nz[(size_t)p].idx]++;
      d.row_nz[(size_t)q] = Entry{(int32_t)j, d.col_nz[(size_t)p].val};
    }
  d.var.resize((size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    Var& x = d.var[(size_t)j];
    x.l = std::isfinite(lo[j]) ? lo[j] : -kDomainInfinity;
    x.u = std::isfinite(hi[j]) ? hi[j] : kDomainInfinity;
    x.tlo = x.thi = 0;
    x.is_int = is_int[j] ? 1 : 0;
    x.pad = 0;
  }
  d.b.assign(b, b + m);
  d.is_le.assign(is_le, is_le + m);
  d.row.resize((size_t)m);
  d.flags.assign((size_t)m, 0);
  d.row_maxcoef.assign((size_t)m, 0.0);
  for (int64_t i = 0; i < m; ++i) {
    const size_t I = (size_t)i;
    const Entry* e = d.row_nz.data();
    double lmin = 0.0, lmax = 0.0;
    for (int64_t k = d.row_ptr[I]; k < d.row_ptr[I + 1]; ++k) {
      const Var& x = d.var[(size_t)e[k].idx];
      const double a = e[k].val;
      if (a > 0) {
        lmin += a * x.l;
        lmax += a * x.u;
      } else {
        lmin += a * x.u;
        lmax += a * x.l;
      }
    }
    bool all_one = true, all_pos = true, all_bin = true;
    double maxcoef = 0.0;
    for (int64_t k = d.row_ptr[I]; k < d.row_ptr[I + 1]; ++k) {
      const int64_t j = e[k].idx;
      const double a = e[k].val;
      if (std::fabs(a) > maxcoef) maxcoef = std::fabs(a);
      if (a != 1.0) all_one = false;
      if (!(a > 0.0)) all_pos = false;
      if (!(is_int[j] && lo[j] == 0.0 && hi[j] == 1.0)) all_bin = false;
      if (!(all_bin && (all_one || all_pos))) break;
    }
    d.row_maxcoef[I] = maxcoef;
    uint8_t kind = kGeneral;
    if (all_bin && all_one) kind = kCardinality;
    else if (all_bin && all_pos && is_le[i]) kind = kKnapsack;
    d.flags[I] = (uint8_t)((is_le[i] ? kLe : 0) | (kind << 1));
    double room = 0.0;
    for (int64_t k = d.row_ptr[I]; k < d.row_ptr[I + 1]; ++k) {
      const int64_t j = e[k].idx;
      const double width = hi[j] - lo[j];
      const double rr = std::isfinite(width) ? std::fabs(e[k].val) * width : INFINITY;
      if (rr > room) room = rr;
    }
    d.row[I] = Row{lmin, lmax, room, b[i]};
  }
  return d;
}

// Rounding state that persists across the calls of one pump run
// (src/rounding.jl:78-98), without the timing policy and the ranking.
struct Rounder {
  Data d;
  int64_t max_pops;
  std::vector<uint8_t> direction;  // violation-guided columns; empty: off
  std::vector<int64_t> int_idx;
  std::string err = "no error";

  double row_violation(size_t i, double activity) const {
    return d.is_le[i] ? pmax(0.0, activity - d.b[i]) : std::fabs(activity - d.b[i]);
  }

  // src/rounding.jl:123-141
  double guided_value(const State& s, int64_t j, double target,
                      const std::vector<double>& activity) const {
    const size_t J = (size_t)j;
    if (!(s.hi(j) - s.lo(j) > kBoundTol)) return target;
    const double lo = std::ceil(s.lo(j) - kBoundTol);
    const double hi = std::floor(s.hi(j) + kBoundTol);
    if (!(hi > lo)) return target;
    auto score = [&](double v) {
      double t = 0.0;
      for (int64_t p = d.col_ptr[J]; p < d.col_ptr[J + 1]; ++p) {
        const size_t i = (size_t)d.col_nz[(size_t)p].idx;
        const double a = d.col_nz[(size_t)p].val;
        t += row_violation(i, activity[i] + a * v) - row_violation(i, activity[i]);
      }
      return t;
    };
    const double score_lo = score(lo);
    const double score_hi = score(hi);
    if (score_lo < score_hi - 1e-12) return lo;
    if (score_hi < score_lo - 1e-12) return hi;
    return target;
  }

  void add_activity(std::vector<double>& activity, int64_t j, double v) const {
    for (int64_t p = d.col_ptr[(size_t)j]; p < d.col_ptr[(size_t)j + 1]; ++p)
      activity[(size_t)d.col_nz[(size_t)p].idx] += d.col_nz[(size_t)p].val * v;
  }

  // src/rounding.jl:152-174
  void round_sequentially(State& s, double* xt, const int64_t* idx, int64_t k,
                          const double* x_star, std::vector<double>* activity) const {
    for (int64_t q = 0; q < k; ++q) {
      const int64_t j = idx[q];
      const size_t J = (size_t)j;
      if (s.is_fixed(j)) {
        xt[J] = s.lo(j);
        continue;
      }
      double target = jround(jclamp(x_star[J], s.lo(j), s.hi(j)));
      if (target < s.lo(j) - kBoundTol || target > s.hi(j) + kBoundTol) {
        const double other =
            jround(jclamp(x_star[J] + (target < s.lo(j) ? 1.0 : -1.0), s.lo(j), s.hi(j)));
        target = jclamp(other, s.lo(j), s.hi(j));
      }
      if (activity && direction[J]) target = guided_value(s, j, target, *activity);
      s.fix(j, target);
      s.propagate();
      if (activity) add_activity(*activity, j, target);
      xt[J] = target;
    }
  }

  void read_integers(const State& s, double* xt, const double* x_star) const {
    for (int64_t j : int_idx) {
      const size_t J = (size_t)j;
      xt[J] = s.is_fixed(j) ? s.lo(j) : jround(jclamp(x_star[J], s.lo(j), s.hi(j)));
    }
  }

  void check_ranked(const int64_t* ranked, int64_t k) const {
    for (int64_t q = 0; q < k; ++q)
      if (ranked[q] < 0 || ranked[q] >= d.n) throw Error(PDFP_ERR_ARGUMENT, "ranked index out of range");
  }

  // FP 2.0's CPR of x_star over the ranked variables (src/rounding.jl:191-197)
  void round_sequential(const double* x_star, const int64_t* ranked, int64_t k, double* xt) const {
    check_ranked(ranked, k);
    State s(d, max_pops);
    std::copy(x_star, x_star + d.n, xt);
    std::vector<double> act;
    if (!direction.empty()) act.assign((size_t)d.m, 0.0);
    round_sequentially(s, xt, ranked, k, x_star, direction.empty() ? nullptr : &act);
    read_integers(s, xt, x_star);
  }

  // The guarded two-phase CPR without its fallback (src/rounding.jl:205-256):
  // returns false when the guard discards the batch.
  bool round_guarded(const double* x_star, const int64_t* ranked, int64_t k, bool regroup,
                     double* xt) const {
    check_ranked(ranked, k);
    State s(d, max_pops);
    std::copy(x_star, x_star + d.n, xt);
    std::vector<double> act;
    std::vector<double>* activity = nullptr;
    if (!direction.empty()) {
      act.assign((size_t)d.m, 0.0);
      activity = &act;
    }
    auto integral = [&](int64_t j) { return std::fabs(x_star[j] - jround(x_star[j])) == 0.0; };
    std::vector<int64_t> before, batch, after;
    if (regroup) {
      int64_t last = 0;
      for (int64_t q = 0; q < k; ++q)
        if (integral(ranked[q])) last = q + 1;
      for (int64_t q = 0; q < last; ++q)
        (integral(ranked[q]) ? batch : before).push_back(ranked[q]);
      after.assign(ranked + last, ranked + k);
    } else {
      for (int64_t q = 0; q < k; ++q) (integral(ranked[q]) ? batch : after).push_back(ranked[q]);
    }
    round_sequentially(s, xt, before.data(), (int64_t)before.size(), x_star, activity);
    for (int64_t j : batch) {
      const size_t J = (size_t)j;
      if (s.is_fixed(j)) {
        xt[J] = s.lo(j);
        continue;
      }
      const double target = jround(jclamp(x_star[J], s.lo(j), s.hi(j)));
      s.fix(j, target);
      if (activity) add_activity(*activity, j, target);
      xt[J] = target;
    }
    s.propagate();
    if (s.infeasible) return false;
    round_sequentially(s, xt, after.data(), (int64_t)after.size(), x_star, activity);
    read_integers(s, xt, x_star);
    return true;
  }
};

std::string g_rounder_error = "no error";

}  // namespace rounding
}  // namespace pdfp

struct pdfp_rounder {
  std::unique_ptr<pdfp::rounding::Rounder> impl;
};

namespace {

template <typename F>
int guarded(pdfp_rounder* r, F&& f) {
  try {
    if (!r || !r->impl) throw pdfp::rounding::Error(PDFP_ERR_STATE, "null rounder");
    f();
    return PDFP_OK;
  } catch (const pdfp::rounding::Error& e) {
    (r && r->impl ? r->impl->err : pdfp::rounding::g_rounder_error) = e.what();
    return e.code;
  } catch (const std::exception& e) {
    (r && r->impl ? r->impl->err : pdfp::rounding::g_rounder_error) = e.what();
    return PDFP_ERR_NOMEM;
  }
}

}  // namespace

extern "C" {

const char* pdfp_rounder_last_error(const pdfp_rounder* r) {
  return r && r->impl ? r->impl->err.c_str() : pdfp::rounding::g_rounder_error.c_str();
}

int pdfp_rounder_create(pdfp_rounder** out, int64_t m, int64_t n, const int64_t* colptr,
                        const int64_t* rowval, const double* nzval, int32_t index_base,
                        const double* b, const uint8_t* is_le, const double* var_lb,
                        const double* var_ub, const uint8_t* is_int, int64_t max_pops) {
  using namespace pdfp::rounding;
  try {
    if (!out || !colptr || (m > 0 && (!b || !is_le)) ||
        (n > 0 && (!var_lb || !var_ub || !is_int)))
      throw Error(PDFP_ERR_ARGUMENT, "null argument");
    if (max_pops < 0) throw Error(PDFP_ERR_ARGUMENT, "max_pops must be >= 0");
    auto r = std::make_unique<pdfp_rounder>();
    r->impl = std::make_unique<Rounder>();
    r->impl->d = build_data(m, n, colptr, rowval, nzval, index_base, b, is_le, var_lb, var_ub,
                            is_int);
    r->impl->max_pops = max_pops;
    for (int64_t j = 0; j < n; ++j)
      if (is_int[j]) r->impl->int_idx.push_back(j);
    *out = r.release();
    return PDFP_OK;
  } catch (const Error& e) {
    g_rounder_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_rounder_error = e.what();
    return PDFP_ERR_NOMEM;
  }
}

int pdfp_rounder_set_direction(pdfp_rounder* r, const uint8_t* direction) {
  return guarded(r, [&] {
    if (direction) r->impl->direction.assign(direction, direction + r->impl->d.n);
    else r->impl->direction.clear();
  });
}

int pdfp_round_sequential(pdfp_rounder* r, const double* x_star, const int64_t* ranked,
                          int64_t k, double* xt) {
  return guarded(r, [&] {
    if (!x_star || !xt || (k > 0 && !ranked) || k < 0)
      throw pdfp::rounding::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    r->impl->round_sequential(x_star, ranked, k, xt);
  });
}

int pdfp_round_guarded(pdfp_rounder* r, const double* x_star, const int64_t* ranked, int64_t k,
                       int32_t regroup, double* xt, int32_t* batch_infeasible) {
  return guarded(r, [&] {
    if (!x_star || !xt || !batch_infeasible || (k > 0 && !ranked) || k < 0)
      throw pdfp::rounding::Error(PDFP_ERR_ARGUMENT, "bad arguments");
    *batch_infeasible = r->impl->round_guarded(x_star, ranked, k, regroup != 0, xt) ? 0 : 1;
  });
}

void pdfp_rounder_destroy(pdfp_rounder* r) { delete r; }

}  // extern "C"
