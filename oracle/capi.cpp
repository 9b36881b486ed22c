// CPU ORACLE -- test infrastructure only (see oracle.hpp).
// extern "C" entry points through which tests/ and bench.py drive the oracle
// (ctypes). Every call returns 0 on success, 1 on a thrown requirement
// (message in or_last_error()).
#include "oracle.hpp"
#include <cstring>
#include <sstream>

using namespace oracle;

namespace {
thread_local std::string last_error;
struct Handle {
  Options opt;
  EulerGas gas;
  Problem p;
  PartialDG dg;
  Hybrid hybrid;
  Condensed condensed;
  bool hybrid_ready = false;
  Handle(const Options &o, double gamma) : opt(o), gas(gamma), p(o, gas), dg(p) {}
};
template <class F> int guard(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    last_error = e.what();
    return 1;
  }
}
Options parse(const char *args) {
  std::istringstream in(args ? args : "");
  std::vector<std::string> words{"oracle"};
  for (std::string w; in >> w;) { words.push_back(w); }
  std::vector<char *> argv;
  for (auto &w : words) { argv.push_back(w.data()); }
  return ParseOptions(int(argv.size()), argv.data());
}
std::vector<double> vec(const double *x, int n) { return std::vector<double>(x, x + n); }
void ensure_hybrid(Handle &h) {
  if (!h.hybrid_ready) {
    h.p.PrepareHybrid(h.hybrid, h.condensed);
    h.hybrid_ready = true;
  }
}
} // namespace

extern "C" {
const char *or_last_error() { return last_error.c_str(); }

int or_create(const char *args, double gamma, void **out) {
  return guard([&] { *out = new Handle(parse(args), gamma); });
}
void or_destroy(void *h) { delete static_cast<Handle *>(h); }

// n, p, m, t, nf, nu, nl, q
int or_sizes(void *hv, int *s) {
  auto &p = static_cast<Handle *>(hv)->p;
  const int v[8] = {p.n, p.p, p.m, p.t, p.nf, p.nu, p.nl, p.q};
  std::memcpy(s, v, sizeof v);
  return 0;
}
// hx, hy, area, lx, ly, dt
int or_geometry(void *hv, double *g) {
  auto &p = static_cast<Handle *>(hv)->p;
  const double v[6] = {p.hx, p.hy, p.area, p.lx, p.ly, p.dt};
  std::memcpy(g, v, sizeof v);
  return 0;
}
// GL nodes x (q), weights (q), d[a*q+i] = l_i'(x_a), l_i(0), l_i(1)
int or_tables(void *hv, double *x, double *w, double *d, double *e0, double *e1) {
  auto &p = static_cast<Handle *>(hv)->p;
  for (int i = 0; i < p.q; ++i) { x[i] = p.x1[i]; w[i] = p.w1[i]; e0[i] = p.end0[i]; e1[i] = p.end1[i]; }
  for (int a = 0; a < p.q; ++a) {
    for (int i = 0; i < p.q; ++i) { d[a * p.q + i] = p.D1(a, i); }
  }
  return 0;
}
int or_metrics(void *hv, double *cell, double *trace) {
  auto &p = static_cast<Handle *>(hv)->p;
  std::memcpy(cell, p.cell_metric.data(), sizeof(double) * p.nu);
  std::memcpy(trace, p.trace_metric.data(), sizeof(double) * p.nl);
  return 0;
}
int or_initial(void *hv, double *u) {
  return guard([&] {
    auto v = static_cast<Handle *>(hv)->p.Initial();
    std::memcpy(u, v.data(), sizeof(double) * v.size());
  });
}
// Freeze HLL bounds at u; sets old=u and dt. frozen: per face point
// (a, b, inverse_sum, left_weight, right_weight, penalty).
int or_freeze(void *hv, const double *u, double dt, double *rate, double *frozen) {
  return guard([&] {
    auto &p = static_cast<Handle *>(hv)->p;
    p.Freeze(vec(u, p.nu), dt, false);
    if (rate) { *rate = p.WaveRate(); }
    if (frozen) {
      for (int f = 0; f < p.nf; ++f) {
        for (int qq = 0; qq < p.q; ++qq) {
          double *o = frozen + 6 * (f * p.q + qq);
          const auto &fc = p.faces[f];
          o[0] = fc.a[qq]; o[1] = fc.b[qq]; o[2] = fc.inverse_sum[qq];
          o[3] = fc.left_weight[qq]; o[4] = fc.right_weight[qq]; o[5] = fc.penalty[qq];
        }
      }
    }
  });
}
int or_set_stage(void *hv, const double *old, double dt) {
  auto &p = static_cast<Handle *>(hv)->p;
  p.old = vec(old, p.nu);
  p.dt = dt;
  return 0;
}
int or_residual(void *hv, const double *u, double *r) {
  return guard([&] { static_cast<Handle *>(hv)->p.Residual(u, r); });
}
double or_cell_norm(void *hv, const double *r) { return static_cast<Handle *>(hv)->p.CellNorm(r); }
int or_admissible(void *hv, const double *u) { return static_cast<Handle *>(hv)->p.Admissible(u) ? 1 : 0; }
int or_integral(void *hv, const double *u, double *out4) {
  auto a = static_cast<Handle *>(hv)->p.Integral(u);
  std::memcpy(out4, a.data(), sizeof a);
  return 0;
}
int or_errors(void *hv, const double *u, double time, double *out4) {
  return guard([&] {
    auto a = static_cast<Handle *>(hv)->p.Errors(u, time);
    std::memcpy(out4, a.data(), sizeof a);
  });
}
int or_differences(void *hv, const double *u, const double *v, double *out4) {
  auto a = static_cast<Handle *>(hv)->p.Differences(u, v);
  std::memcpy(out4, a.data(), sizeof a);
  return 0;
}
// Matrix-free DG Jacobian at state u applied to x.
int or_dg_refresh(void *hv, const double *u) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    h.dg.Refresh(vec(u, h.p.nu));
  });
}
int or_dg_apply(void *hv, const double *x, double *y) {
  return guard([&] { static_cast<Handle *>(hv)->dg.Mult(x, y); });
}
int or_dense_dg(void *hv, const double *u, double *a) {
  return guard([&] {
    auto &p = static_cast<Handle *>(hv)->p;
    Dense d;
    p.DenseDG(u, d);
    std::memcpy(a, d.data(), sizeof(double) * d.d.size());
  });
}
// Unfactored block-Jacobi blocks (column-major, block after block) of the
// current DG Jacobian (form 0) or of the Schur operator (form 1).
int or_diagonal_blocks(void *hv, int form, double *blocks) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    std::vector<Dense> d;
    if (form == 0) { h.dg.DiagonalBlocks(h.p.m, d); }
    else { require(h.hybrid_ready, "condense first"); h.condensed.op->DiagonalBlocks(h.p.t, d); }
    size_t off = 0;
    for (auto &b : d) { std::memcpy(blocks + off, b.data(), sizeof(double) * b.d.size()); off += b.d.size(); }
  });
}
int or_lu(double *a, int *ipiv, int m) { return lu_factor(a, ipiv, m) ? 0 : 1; }
int or_lu_solve(const double *a, const int *ipiv, int m, int nrhs, double *x) {
  lu_solve(a, ipiv, m, nrhs, x);
  return 0;
}
int or_lu_right_solve(const double *a, const int *ipiv, int m, int rows, double *x) {
  lu_right_solve(a, ipiv, m, rows, x);
  return 0;
}
// Preconditioner application x = P^{-1} b for form 0 (DG at the refreshed
// state) or 1 (Schur after condensation); precond 0 block-Jacobi, 1 ILU(0).
// If filled_blocks is non-null the ILU's unfactored BSR values are returned
// there (row-major over the BSR pattern); col/row give the pattern.
int or_precondition(void *hv, int form, int precond, const double *b, double *x) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    Options o = h.opt;
    o.precond = precond ? "block-ilu0" : "block-jacobi";
    const Operator *op = form == 0 ? static_cast<Operator *>(&h.dg) : h.condensed.op.get();
    require(op != nullptr, "condense first");
    const int size = form == 0 ? h.p.nu : h.p.nl;
    LinearSolver s(size, form == 0 ? h.p.m : h.p.t, o, form == 0 ? h.p.cell_metric : h.p.trace_metric);
    s.Prepare(op->BlockGraph(s.block), o);
    if (s.ilu) {
      op->FillBlocks(*s.ilu);
      s.ilu->Factor();
    } else {
      op->DiagonalBlocks(s.block, s.diagonal);
      for (size_t e = 0; e < s.diagonal.size(); ++e) {
        require(lu_factor(s.diagonal[e].data(), s.diagonal_pivots[e].data(), s.block),
                "Singular block Jacobi factor");
      }
    }
    s.Precondition(b, x);
  });
}
// BSR pattern and filled (unfactored) values of the ILU target.
int or_ilu_pattern(void *hv, int form, int *nnzb, int *row, int *col) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    FixedBlockILU ilu(h.p.SolverGraph(form == 1), form == 0 ? h.p.m : h.p.t);
    *nnzb = ilu.col.size();
    if (row) { std::memcpy(row, ilu.row.data(), sizeof(int) * ilu.row.size()); }
    if (col) { std::memcpy(col, ilu.col.data(), sizeof(int) * ilu.col.size()); }
  });
}
// The ILU(0) factor's block row order for the handle's --ilu-ordering.
int or_ilu_order(void *hv, int form, int *order) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    FixedBlockILU ilu(h.p.SolverGraph(form == 1), form == 0 ? h.p.m : h.p.t, h.opt.ilu_ordering);
    std::memcpy(order, ilu.permutation.data(), sizeof(int) * ilu.permutation.size());
  });
}
// The minimum discarded fill order of a graph given as adjacency lists.
int or_mdf_ordering(int n, const int *ptr, const int *adj, int *order) {
  return guard([&] {
    std::vector<std::vector<int>> g(n);
    for (int i = 0; i < n; ++i) { g[i].assign(adj + ptr[i], adj + ptr[i + 1]); }
    const auto o = mdf_ordering(g);
    std::memcpy(order, o.data(), sizeof(int) * o.size());
  });
}
int or_ilu_blocks(void *hv, int form, int factored, double *values) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    const Operator *op = form == 0 ? static_cast<Operator *>(&h.dg) : h.condensed.op.get();
    FixedBlockILU ilu(h.p.SolverGraph(form == 1), form == 0 ? h.p.m : h.p.t);
    op->FillBlocks(ilu);
    if (factored) { ilu.Factor(); }
    size_t off = 0;
    for (auto &b : ilu.data) { std::memcpy(values + off, b.data(), sizeof(double) * b.d.size()); off += b.d.size(); }
  });
}
// HDG: StreamCondense at (u, r); outputs are optional.
int or_condense(void *hv, const double *u, const double *r, double *X, double *y, double *rhs,
                double *D) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    ensure_hybrid(h);
    h.p.StreamCondense(u, r, h.hybrid, h.condensed);
    const auto &p = h.p;
    if (X) { for (int e = 0; e < p.n; ++e) { std::memcpy(X + size_t(e) * p.m * 4 * p.t, h.condensed.X[e].data(), sizeof(double) * p.m * 4 * p.t); } }
    if (y) { for (int e = 0; e < p.n; ++e) { std::memcpy(y + size_t(e) * p.m, h.condensed.y[e].data(), sizeof(double) * p.m); } }
    if (rhs) { std::memcpy(rhs, h.condensed.rhs.data(), sizeof(double) * p.nl); }
    if (D) { std::memcpy(D, h.hybrid.D.data(), sizeof(double) * p.nl); }
  });
}
int or_schur_apply(void *hv, const double *x, double *y) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    require(h.hybrid_ready, "condense first");
    h.condensed.op->Mult(x, y);
  });
}
int or_recover(void *hv, const double *dl, double *du) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    h.p.Recover(h.hybrid, h.condensed, dl, du);
  });
}
int or_recovered_residual(void *hv, const double *eps, double *out) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    h.p.RecoveredResidual(h.hybrid, eps, out);
  });
}
// Full LinearSolver::Solve (refresh=true, zero guess) on the current DG
// Jacobian (form 0) or condensed Schur system (form 1).
int or_linear_solve(void *hv, int form, const double *rhs, double *x, double *eq_res, int *count,
                    double *relres) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    const Operator *op = form == 0 ? static_cast<Operator *>(&h.dg) : h.condensed.op.get();
    const int size = form == 0 ? h.p.nu : h.p.nl;
    LinearSolver s(size, form == 0 ? h.p.m : h.p.t, h.opt, form == 0 ? h.p.cell_metric : h.p.trace_metric);
    if (h.opt.precond == "block-ilu0") { s.Prepare(op->BlockGraph(s.block), h.opt); }
    LinearResult res;
    s.Solve(*op, rhs, h.opt, res);
    std::memcpy(x, res.x.data(), sizeof(double) * size);
    if (eq_res) { std::memcpy(eq_res, res.equation_residual.data(), sizeof(double) * size); }
    if (count) { *count = res.count; }
    if (relres) { *relres = res.residual; }
  });
}
// Whole Run (src/stepper.cpp:28-292). form 0 DG, 1 HDG. u_final (nu),
// stats: updates, time_steps, newton_seconds, final_time, drift, errors[4].
// increments (optional): updates*nu doubles, the accepted du of every update;
// krylov (optional): Krylov iteration count per update.
int or_run(void *hv, int form, const char *prefix, double *u_final, double *stats,
           double *increments, int max_updates, int *krylov) {
  return guard([&] {
    auto &h = *static_cast<Handle *>(hv);
    auto r = Run(h.p, form ? "hdg" : "dg", prefix ? prefix : "", increments != nullptr);
    std::memcpy(u_final, r.u.data(), sizeof(double) * h.p.nu);
    stats[0] = r.updates; stats[1] = r.time_steps; stats[2] = r.newton_seconds;
    stats[3] = r.final_time; stats[4] = r.drift;
    for (int c = 0; c < 4; ++c) { stats[5 + c] = r.errors[c]; }
    const int nrec = std::min<int>(r.updates, max_updates);
    if (increments) {
      for (int i = 0; i < nrec; ++i) {
        std::memcpy(increments + size_t(i) * h.p.nu, r.increments[i].data(), sizeof(double) * h.p.nu);
      }
    }
    if (krylov) { for (int i = 0; i < nrec; ++i) { krylov[i] = r.records[i].krylov; } }
  });
}
// Dense operator with block-diagonal / BSR views, for the weighted-GMRES
// known-answer case of src/verification.cpp:174-254.
namespace {
class DenseOperator : public Operator {
public:
  Dense a;
  explicit DenseOperator(const Dense &m) : a(m) { height = m.h; }
  void Mult(const double *x, double *y) const override {
    for (int i = 0; i < a.h; ++i) {
      double s = 0.;
      for (int j = 0; j < a.w; ++j) { s += a(i, j) * x[j]; }
      y[i] = s;
    }
  }
  void DiagonalBlocks(int size, std::vector<Dense> &blocks) const override {
    blocks.resize(height / size);
    for (int b = 0; b < height / size; ++b) {
      blocks[b].resize(size, size);
      for (int j = 0; j < size; ++j) {
        for (int i = 0; i < size; ++i) { blocks[b](i, j) = a(b * size + i, b * size + j); }
      }
    }
  }
  std::vector<std::vector<int>> BlockGraph(int size) const override {
    std::vector<std::vector<int>> g(height / size);
    for (int r = 0; r < height / size; ++r) {
      for (int c = 0; c < height / size; ++c) {
        bool nz = false;
        for (int j = 0; j < size; ++j) {
          for (int i = 0; i < size; ++i) { nz = nz || a(r * size + i, c * size + j) != 0.; }
        }
        if (nz || r == c) { g[r].push_back(c); }
      }
    }
    return g;
  }
  void FillBlocks(FixedBlockILU &t) const override {
    t.Zero();
    const int b = t.block;
    for (int r = 0; r < height / b; ++r) {
      for (int k = t.row[r]; k < t.row[r + 1]; ++k) {
        const int c = t.col[k];
        for (int j = 0; j < b; ++j) {
          for (int i = 0; i < b; ++i) { t.data[k](i, j) = a(r * b + i, c * b + j); }
        }
      }
    }
  }
};
} // namespace

// src/verification.cpp:174-254: GMRES in the diagonal mass metric and in
// the Cholesky-whitened dense basis must give the same physical correction
// and the same iteration counts; a zero RHS takes 0 iterations.
// out (per preconditioner bj, ilu): error, original count, transformed count,
// zero-rhs count, zero-rhs |x|.
int or_verify_weighted(double *out) {
  return guard([&] {
    constexpr int n = 6;
    Dense k(n, n), t(n, n), kt(n, n), ht(n, n), l(n, n), li(n, n), normalized(n, n);
    const double mass[n] = {.2, .7, 1.1, .3, .8, 1.4};
    for (int i = 0; i < n; ++i) {
      k(i, i) = 4 + .3 * i;
      t(i, i) = .7 + .2 * i;
      if (i + 1 < n) { k(i, i + 1) = .3; k(i + 1, i) = -.2; }
    }
    t(0, 1) = .25; t(1, 2) = -.15; t(3, 5) = .2;
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        for (int a = 0; a < n; ++a) {
          ht(i, j) += t(a, i) * mass[a] * t(a, j);
          for (int b = 0; b < n; ++b) { kt(i, j) += t(a, i) * k(a, b) * t(b, j); }
        }
      }
    }
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j <= i; ++j) {
        double v = ht(i, j);
        for (int a = 0; a < j; ++a) { v -= l(i, a) * l(j, a); }
        if (i == j) { require(v > 0, "Test Gram must be SPD"); l(i, j) = std::sqrt(v); }
        else { l(i, j) = v / l(j, j); }
      }
    }
    for (int j = 0; j < n; ++j) {
      for (int i = 0; i < n; ++i) {
        double v = i == j ? 1. : 0.;
        for (int a = 0; a < i; ++a) { v -= l(i, a) * li(a, j); }
        li(i, j) = v / l(i, i);
      }
    }
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        for (int a = 0; a < n; ++a) {
          for (int b = 0; b < n; ++b) { normalized(i, j) += li(i, a) * kt(a, b) * li(j, b); }
        }
      }
    }
    std::vector<double> rhs(n), trhs(n, 0.), metric(n), ones(n, 1.);
    for (int i = 0; i < n; ++i) { rhs[i] = std::sin(i + 1.); metric[i] = 1 / mass[i]; }
    for (int i = 0; i < n; ++i) {
      for (int a = 0; a < n; ++a) {
        for (int b = 0; b < n; ++b) { trhs[i] += li(i, a) * t(b, a) * rhs[b]; }
      }
    }
    DenseOperator A(k), B(normalized);
    int slot = 0;
    for (const char *pre : {"block-jacobi", "block-ilu0"}) {
      Options o;
      o.precond = pre;
      o.restart = 2;
      o.linear_tol = 1e-10;
      LinearSolver s1(n, 3, o, metric), s2(n, 3, o, ones);
      LinearResult r1, r2, r3;
      s1.Solve(A, rhs.data(), o, r1);
      s2.Solve(B, trhs.data(), o, r2);
      std::vector<double> phys(n, 0.);
      for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
          for (int h = 0; h < n; ++h) { phys[i] += t(i, j) * li(h, j) * r2.x[h]; }
        }
      }
      double num = 0., den = 0.;
      for (int i = 0; i < n; ++i) { num += (phys[i] - r1.x[i]) * (phys[i] - r1.x[i]); den += r1.x[i] * r1.x[i]; }
      std::vector<double> zero(n, 0.);
      LinearSolver s3(n, 3, o, metric);
      s3.Solve(A, zero.data(), o, r3);
      double zx = 0.;
      for (double v : r3.x) { zx += v * v; }
      out[slot++] = std::sqrt(num / den);
      out[slot++] = r1.count;
      out[slot++] = r2.count;
      out[slot++] = r3.count;
      out[slot++] = std::sqrt(zx);
    }
  });
}

// Physics points for known-answer tests.
int or_gas(double gamma, const double *u, int axis, double safety, const double *r, double *out) {
  return guard([&] {
    EulerGas g(gamma);
    State a{u[0], u[1], u[2], u[3]}, b{r[0], r[1], r[2], r[3]};
    out[0] = g.pressure(a);
    out[1] = g.positive(a);
    auto f = g.flux(a, axis);
    for (int c = 0; c < 4; ++c) { out[2 + c] = f[c]; }
    Mat4 j;
    g.jacobian(a, axis, j);
    for (int i = 0; i < 16; ++i) { out[6 + i] = j.a[i]; }
    auto s = g.speeds(a, b, axis, safety);
    out[22] = s[0]; out[23] = s[1];
  });
}
int or_exact(double gamma, double x, double y, double time, double alpha, double *out) {
  return guard([&] {
    auto v = EulerGas(gamma).exact(x, y, time, alpha);
    for (int c = 0; c < 4; ++c) { out[c] = v[c]; }
  });
}
}
