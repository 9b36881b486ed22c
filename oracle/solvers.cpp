// CPU ORACLE -- test infrastructure only (see oracle.hpp).
// Dense LU, matrix-free DG Jacobian, HDG condensation/Schur operator, BSR
// block ILU(0) and weighted restarted GMRES.
#include "oracle.hpp"
#include <algorithm>
#include <cmath>
#include <limits>

namespace oracle {
// ------------------------------------------------------------- dense LU
// Partially pivoted Doolittle LU in place (column-major, unit lower factor),
// the semantics of MFEM LUFactors::Factor used at src/linear.cpp:200-201,
// src/condensation.cpp:137, src/block_ilu.cpp:84. Pivot: first maximum |a|.
bool lu_factor(double *a, int *ipiv, int m) {
  for (int i = 0; i < m; ++i) {
    int piv = i;
    double best = std::abs(a[i + i * m]);
    for (int j = i + 1; j < m; ++j) {
      const double v = std::abs(a[j + i * m]);
      if (v > best) { best = v; piv = j; }
    }
    ipiv[i] = piv;
    if (piv != i) {
      for (int k = 0; k < m; ++k) { std::swap(a[i + k * m], a[piv + k * m]); }
    }
    if (a[i + i * m] == 0.) { return false; }
    const double inv = 1. / a[i + i * m];
    for (int j = i + 1; j < m; ++j) { a[j + i * m] *= inv; }
    for (int k = i + 1; k < m; ++k) {
      const double aik = a[i + k * m];
      for (int j = i + 1; j < m; ++j) { a[j + k * m] -= aik * a[j + i * m]; }
    }
  }
  return true;
}
// X <- A^{-1} X for nrhs columns (LUFactors::Solve).
void lu_solve(const double *a, const int *ipiv, int m, int nrhs, double *x) {
  for (int r = 0; r < nrhs; ++r) {
    double *b = x + size_t(r) * m;
    for (int i = 0; i < m; ++i) { std::swap(b[i], b[ipiv[i]]); }
    for (int j = 0; j < m; ++j) {
      const double bj = b[j];
      for (int i = j + 1; i < m; ++i) { b[i] -= a[i + j * m] * bj; }
    }
    for (int j = m - 1; j >= 0; --j) {
      b[j] /= a[j + j * m];
      const double bj = b[j];
      for (int i = 0; i < j; ++i) { b[i] -= a[i + j * m] * bj; }
    }
  }
}
// X <- X A^{-1} for a rows x m matrix X (LUFactors::RightSolve, used by the
// ILU factor at src/block_ilu.cpp:77-78). With P A = L U: X A^{-1} = X U^{-1} L^{-1} P.
void lu_right_solve(const double *a, const int *ipiv, int m, int rows, double *x) {
  for (int j = 0; j < m; ++j) {   // X <- X U^{-1}
    const double inv = 1. / a[j + j * m];
    for (int i = 0; i < rows; ++i) {
      double s = x[i + size_t(j) * rows];
      for (int k = 0; k < j; ++k) { s -= x[i + size_t(k) * rows] * a[k + j * m]; }
      x[i + size_t(j) * rows] = s * inv;
    }
  }
  for (int j = m - 1; j >= 0; --j) {   // X <- X L^{-1}
    for (int i = 0; i < rows; ++i) {
      double s = x[i + size_t(j) * rows];
      for (int k = j + 1; k < m; ++k) { s -= x[i + size_t(k) * rows] * a[k + j * m]; }
      x[i + size_t(j) * rows] = s;
    }
  }
  for (int j = m - 1; j >= 0; --j) {   // X <- X P
    if (ipiv[j] != j) {
      for (int i = 0; i < rows; ++i) {
        std::swap(x[i + size_t(j) * rows], x[i + size_t(ipiv[j]) * rows]);
      }
    }
  }
}
namespace {
// C -= A B (MFEM AddMult_a(-1,...), src/block_ilu.cpp:80)
void sub_mult(const Dense &a, const Dense &b, Dense &c) {
  for (int j = 0; j < b.w; ++j) {
    for (int k = 0; k < a.w; ++k) {
      const double bkj = b(k, j);
      if (bkj == 0.) { continue; }
      for (int i = 0; i < a.h; ++i) { c(i, j) -= a(i, k) * bkj; }
    }
  }
}
} // namespace

// ------------------------------------------------------------- FaceCoupling
// src/operators.cpp:75-112
void FaceCoupling::Mult(const double *x, double *y) const {
  for (int i = 0; i < Height(); ++i) { y[i] = 0.; }
  for (int side = 0; side < 4; ++side) {
    for (int qq = 0; qq < nodes; ++qq) {
      const auto &restriction = (*basis)[side][qq];
      double value[4] = {}, result[4] = {};
      const int offset = 4 * (side * nodes + qq);
      if (lift) {
        for (int c = 0; c < 4; ++c) { value[c] = x[offset + c]; }
        const double a = scalar_flux[side * nodes + qq];
        for (int c = 0; c < 4; ++c) { result[c] = a * value[c]; }
        for (const auto &entry : restriction) {
          const double weight = entry.second * (*inverse_mass)[entry.first];
          for (int c = 0; c < 4; ++c) { y[4 * entry.first + c] += weight * result[c]; }
        }
      } else {
        for (const auto &entry : restriction) {
          for (int c = 0; c < 4; ++c) { value[c] += entry.second * x[4 * entry.first + c]; }
        }
        const auto &a = flux[side * nodes + qq];
        for (int d = 0; d < 4; ++d) {
          for (int c = 0; c < 4; ++c) { result[c] += a(c, d) * value[d]; }
        }
        for (int c = 0; c < 4; ++c) { y[offset + c] = result[c]; }
      }
    }
  }
}
// src/operators.cpp:121-142
void FaceCoupling::Dense_(Dense &a) const {
  a.resize(Height(), Width());
  for (int side = 0; side < 4; ++side) {
    for (int qq = 0; qq < nodes; ++qq) {
      const int offset = 4 * (side * nodes + qq);
      for (const auto &entry : (*basis)[side][qq]) {
        double weight = entry.second;
        if (lift) {
          weight *= (*inverse_mass)[entry.first];
          const double value = weight * scalar_flux[side * nodes + qq];
          for (int c = 0; c < 4; ++c) { a(4 * entry.first + c, offset + c) = value; }
        } else {
          for (int d = 0; d < 4; ++d) {
            for (int c = 0; c < 4; ++c) {
              a(offset + c, 4 * entry.first + d) = weight * flux[side * nodes + qq](c, d);
            }
          }
        }
      }
    }
  }
}
// src/operators.cpp:7-26
void FaceCoupling::Panel(const Dense &x, int side, int first, int count, Dense &out) const {
  require(!lift, "Only restriction couplings have face output panels");
  out.resize(4 * nodes, count);
  for (int j = 0; j < count; ++j) {
    for (int qq = 0; qq < nodes; ++qq) {
      double value[4] = {}, result[4] = {};
      for (const auto &entry : (*basis)[side][qq]) {
        for (int d = 0; d < 4; ++d) { value[d] += entry.second * x(4 * entry.first + d, first + j); }
      }
      const auto &a = flux[side * nodes + qq];
      for (int d = 0; d < 4; ++d) {
        for (int c = 0; c < 4; ++c) { result[c] += a(c, d) * value[d]; }
      }
      for (int c = 0; c < 4; ++c) { out(4 * qq + c, j) = result[c]; }
    }
  }
}

// ------------------------------------------------------------- PartialDG
// src/partial.cpp:23-30
PartialDG::PartialDG(const Problem &problem) : p(problem) {
  height = p.nu;
  flux_x.resize(p.p);
  flux_y.resize(p.p);
  jx.resize(size_t(p.n) * p.p);
  jy.resize(size_t(p.n) * p.p);
  jl.resize(size_t(p.nf) * p.q);
  jr.resize(size_t(p.nf) * p.q);
}
// src/partial.cpp:31-59
void PartialDG::Refresh(const std::vector<double> &u) {
  state = u;
  for (int e = 0; e < p.n; ++e) {
    for (int qq = 0; qq < p.p; ++qq) {
      const auto v = p.Node(u.data(), e, qq);
      p.gas.jacobian(v, 0, jx[e * p.p + qq]);
      p.gas.jacobian(v, 1, jy[e * p.p + qq]);
    }
  }
  int id = 0;
  for (const auto &f : p.faces) {
    for (int qq = 0; qq < p.q; qq++, ++id) {
      auto &l = jl[id];
      auto &r = jr[id];
      p.gas.jacobian(p.Trace(u.data(), f.l, f.ls, qq), f.axis, l);
      p.gas.jacobian(p.Trace(u.data(), f.r, f.rs, qq), f.axis, r);
      for (int d = 0; d < 4; d++) {
        for (int c = 0; c < 4; c++) {
          l(c, d) = f.left_weight[qq] * l(c, d) + (c == d ? f.penalty[qq] : 0);
          r(c, d) = f.right_weight[qq] * r(c, d) - (c == d ? f.penalty[qq] : 0);
        }
      }
    }
  }
}
// src/partial.cpp:62-135: q^2 pointwise 4x4 products plus two q^3 line
// contractions per cell, then the face terms, the mass inverse and +x.
void PartialDG::Mult(const double *x, double *y) const {
  const int n1 = p.q;
  for (int e = 0; e < p.n; e++) {
    for (int point = 0; point < p.p; ++point) {
      const double *v = x + e * p.m + 4 * point;
      auto &fx = flux_x[point];
      auto &fy = flux_y[point];
      fx = State{};
      fy = State{};
      const int id = e * p.p + point;
      for (int d = 0; d < 4; d++) {
        for (int c = 0; c < 4; c++) {
          fx[c] += jx[id](c, d) * v[d];
          fy[c] += jy[id](c, d) * v[d];
        }
      }
      const double weight = p.dt * p.w[point];
      for (int c = 0; c < 4; ++c) { fx[c] *= weight; fy[c] *= weight; }
    }
    for (int j = 0; j < n1; ++j) {
      for (int i = 0; i < n1; ++i) {
        State sum{};
        for (int a = 0; a < n1; ++a) {
          const double dx = p.D1(a, i) / p.hx, dy = p.D1(a, j) / p.hy;
          for (int c = 0; c < 4; ++c) {
            sum[c] += dx * flux_x[j * n1 + a][c] + dy * flux_y[a * n1 + i][c];
          }
        }
        double *cell = y + e * p.m + 4 * (j * n1 + i);
        for (int c = 0; c < 4; ++c) { cell[c] = -sum[c]; }
      }
    }
  }
  int id = 0;
  for (const auto &f : p.faces) {
    for (int qq = 0; qq < p.q; qq++, id++) {
      const auto l = p.Trace(x, f.l, f.ls, qq), r = p.Trace(x, f.r, f.rs, qq);
      State z{};
      for (int d = 0; d < 4; d++) {
        for (int c = 0; c < 4; c++) { z[c] += jl[id](c, d) * l[d] + jr[id](c, d) * r[d]; }
      }
      const double scale = p.dt * f.length / p.area * p.w1[qq];
      for (const auto &entry : p.face_basis[f.ls][qq]) {
        double *cell = y + f.l * p.m + 4 * entry.first;
        const double weight = scale * entry.second;
        for (int c = 0; c < 4; ++c) { cell[c] += weight * z[c]; }
      }
      for (const auto &entry : p.face_basis[f.rs][qq]) {
        double *cell = y + f.r * p.m + 4 * entry.first;
        const double weight = scale * entry.second;
        for (int c = 0; c < 4; ++c) { cell[c] -= weight * z[c]; }
      }
    }
  }
  p.MassSolve(y);
  for (int i = 0; i < p.nu; ++i) { y[i] += x[i]; }
}
// src/partial.cpp:136-139, src/block_ilu.cpp:110-124
std::vector<std::vector<int>> PartialDG::BlockGraph(int size) const {
  require(size == p.m, "DG BSR requires cell blocks");
  return p.SolverGraph(false);
}
// src/partial.cpp:140-173
void PartialDG::FillBlocks(FixedBlockILU &target) const {
  require(target.block == p.m, "DG BSR requires cell blocks");
  target.Zero(true);
  for (int e = 0; e < p.n; ++e) { p.VolumeJacobianCell(state.data(), e, target.At(e, e), &jx, &jy); }
  for (int fidx = 0; fidx < p.nf; ++fidx) {
    const auto &f = p.faces[fidx];
    Dense *blk[2][2] = {{&target.At(f.l, f.l), &target.At(f.l, f.r)},
                        {&target.At(f.r, f.l), &target.At(f.r, f.r)}};
    for (int qq = 0; qq < p.q; ++qq) {
      const int id = fidx * p.q + qq;
      const double scale = p.dt * f.length / p.area * p.w1[qq];
      for (int side = 0; side < 2; ++side) {
        const int test_side = side ? f.rs : f.ls;
        for (int nb = 0; nb < 2; ++nb) {
          const int trial_side = nb ? f.rs : f.ls;
          Dense &a = *blk[side][nb];
          const Mat4 &flux = nb ? jr[id] : jl[id];
          for (const auto &i : p.face_basis[test_side][qq]) {
            for (const auto &j : p.face_basis[trial_side][qq]) {
              const double weight = (side ? -scale : scale) * p.inv_mass[i.first] * i.second * j.second;
              for (int d = 0; d < 4; ++d) {
                for (int c = 0; c < 4; ++c) { a(4 * i.first + c, 4 * j.first + d) += weight * flux(c, d); }
              }
            }
          }
        }
      }
    }
  }
}
// src/partial.cpp:265-298
void PartialDG::DiagonalBlocks(int size, std::vector<Dense> &diag) const {
  require(size == p.m, "Partial DG expects collocated cell blocks");
  diag.resize(p.n);
  for (int e = 0; e < p.n; ++e) { p.VolumeJacobianCell(state.data(), e, diag[e], &jx, &jy); }
  int id = 0;
  for (const auto &f : p.faces) {
    for (int qq = 0; qq < p.q; qq++, id++) {
      const double scale = p.dt * f.length / p.area * p.w1[qq];
      for (const auto &i : p.face_basis[f.ls][qq]) {
        for (const auto &j : p.face_basis[f.ls][qq]) {
          const double weight = scale * p.inv_mass[i.first] * i.second * j.second;
          for (int d = 0; d < 4; ++d) {
            for (int c = 0; c < 4; ++c) { diag[f.l](4 * i.first + c, 4 * j.first + d) += weight * jl[id](c, d); }
          }
        }
      }
      for (const auto &i : p.face_basis[f.rs][qq]) {
        for (const auto &j : p.face_basis[f.rs][qq]) {
          const double weight = scale * p.inv_mass[i.first] * i.second * j.second;
          for (int d = 0; d < 4; ++d) {
            for (int c = 0; c < 4; ++c) { diag[f.r](4 * i.first + c, 4 * j.first + d) -= weight * jr[id](c, d); }
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------- HDG
// src/condensation.cpp:47-85 (partial assembly: factor entries only)
void Problem::PrepareHybrid(Hybrid &h, Condensed &s) const {
  h.B.resize(n);
  h.C.resize(n);
  h.D.assign(nl, 0.);
  h.ids.assign(n, {});
  s.X.resize(n);
  s.y.assign(n, std::vector<double>(m, 0.));
  s.rhs.assign(nl, 0.);
  s.cell_matrix.resize(m, m);
  s.cell_pivots.assign(m, 0);
  for (int e = 0; e < n; ++e) {
    for (auto *coupling : {&h.B[e], &h.C[e]}) {
      coupling->basis = &face_basis;
      coupling->inverse_mass = &inv_mass;
      coupling->nodes = q;
      coupling->cell_size = m;
    }
    h.B[e].lift = true;
    h.B[e].scalar_flux.assign(4 * q, 0.);
    h.C[e].lift = false;
    h.C[e].flux.assign(4 * q, Mat4{});
    s.X[e].resize(m, 4 * t);
    for (int side = 0; side < 4; ++side) {
      const int f = cell_faces[e][side];
      for (int i = 0; i < t; ++i) { h.ids[e].push_back(f * t + i); }
    }
  }
  s.op = std::make_unique<SchurOperator>(*this, h, s.X);
}
// src/problem.cpp:598-635 (recomputed on every call: the cache is a pure
// function of the frozen bounds and dt)
void Problem::HybridDiagonal(Hybrid &h) const {
  for (int f = 0; f < nf; ++f) {
    for (int qq = 0; qq < q; ++qq) {
      const double diagonal = w1[qq] * (faces[f].a[qq] + faces[f].b[qq]);
      for (int c = 0; c < 4; ++c) { h.D[f * t + 4 * qq + c] = diagonal; }
    }
  }
  for (int e = 0; e < n; ++e) {
    for (int side = 0; side < 4; ++side) {
      const auto &face = faces[cell_faces[e][side]];
      const bool left = face.l == e && face.ls == side;
      for (int qq = 0; qq < q; ++qq) {
        const double scale = dt * face.length / area * w1[qq];
        h.B[e].scalar_flux[side * q + qq] = -scale * (left ? face.a[qq] : face.b[qq]);
      }
    }
  }
}
// src/problem.cpp:642-702
void Problem::HybridCell(const double *u, int e, Hybrid &h, Dense &a) const {
  VolumeJacobianCell(u, e, a);
  for (int side = 0; side < 4; ++side) {
    const auto &face = faces[cell_faces[e][side]];
    const bool left = face.l == e && face.ls == side;
    const double normal_sign = left ? 1. : -1.;
    for (int qq = 0; qq < q; ++qq) {
      Mat4 derivative;
      gas.jacobian(Trace(u, e, side, qq), face.axis, derivative);
      const double stab = left ? face.a[qq] : face.b[qq];
      for (int d = 0; d < 4; ++d) {
        for (int c = 0; c < 4; ++c) {
          derivative(c, d) = normal_sign * derivative(c, d) + (c == d ? stab : 0.);
        }
      }
      const double scale = dt * face.length / area * w1[qq];
      const auto &restriction = face_basis[side][qq];
      for (const auto &test : restriction) {
        const int i = test.first;
        const double row_weight = scale * test.second * inv_mass[i];
        for (const auto &trial : restriction) {
          const int j = trial.first;
          const double value_weight = row_weight * trial.second;
          for (int d = 0; d < 4; ++d) {
            for (int c = 0; c < 4; ++c) { a(4 * i + c, 4 * j + d) += value_weight * derivative(c, d); }
          }
        }
      }
      for (int d = 0; d < 4; ++d) {
        for (int c = 0; c < 4; ++c) { h.C[e].flux[side * q + qq](c, d) = -w1[qq] * derivative(c, d); }
      }
    }
  }
}
// src/condensation.cpp:131-148
void Problem::StreamCondense(const double *u, const double *r, Hybrid &h, Condensed &s) const {
  HybridDiagonal(h);
  std::fill(s.rhs.begin(), s.rhs.end(), 0.);
  std::vector<double> trace(4 * t);
  for (int e = 0; e < n; ++e) {
    HybridCell(u, e, h, s.cell_matrix);
    require(lu_factor(s.cell_matrix.data(), s.cell_pivots.data(), m),
            "Singular streaming HDG cell factor");
    h.B[e].Dense_(s.X[e]);
    lu_solve(s.cell_matrix.data(), s.cell_pivots.data(), m, 4 * t, s.X[e].data());
    std::copy_n(r + size_t(e) * m, m, s.y[e].data());
    lu_solve(s.cell_matrix.data(), s.cell_pivots.data(), m, 1, s.y[e].data());
    h.C[e].Mult(s.y[e].data(), trace.data());
    for (int i = 0; i < 4 * t; ++i) { s.rhs[h.ids[e][i]] += trace[i]; }
  }
}
// src/condensation.cpp:28-41,154-163: du_e = -X_e R_e dl - y_e
void Problem::Recover(const Hybrid &h, const Condensed &s, const double *dl, double *du) const {
  std::vector<double> local(4 * t), cell(m);
  for (int e = 0; e < n; ++e) {
    for (int i = 0; i < 4 * t; ++i) { local[i] = dl[h.ids[e][i]]; }
    const Dense &x = s.X[e];
    for (int i = 0; i < m; ++i) { cell[i] = 0.; }
    for (int j = 0; j < 4 * t; ++j) {
      for (int i = 0; i < m; ++i) { cell[i] += x(i, j) * local[j]; }
    }
    for (int i = 0; i < m; ++i) { du[e * m + i] = -cell[i]; }
    for (int i = 0; i < m; ++i) { du[e * m + i] -= s.y[e][i]; }
  }
}
// src/stepper.cpp:10-26: eps_DG = -B D^{-1} eps_S
void Problem::RecoveredResidual(const Hybrid &h, const double *schur_residual, double *out) const {
  std::vector<double> scaled(nl), local(4 * t), cell(m);
  for (int i = 0; i < nl; ++i) { scaled[i] = schur_residual[i] / h.D[i]; }
  for (int e = 0; e < n; ++e) {
    for (int i = 0; i < 4 * t; ++i) { local[i] = scaled[h.ids[e][i]]; }
    h.B[e].Mult(local.data(), cell.data());
    for (int i = 0; i < m; ++i) { out[e * m + i] = -cell[i]; }
  }
}
// src/block_ilu.cpp:110-124
std::vector<std::vector<int>> Problem::SolverGraph(bool hdg) const {
  std::vector<std::vector<int>> graph(hdg ? nf : n);
  if (hdg) {
    for (const auto &local : cell_faces) {
      for (int i : local) { for (int j : local) { graph[i].push_back(j); } }
    }
  } else {
    for (int e = 0; e < n; ++e) { graph[e].push_back(e); }
    for (const auto &f : faces) { graph[f.l].push_back(f.r); graph[f.r].push_back(f.l); }
  }
  for (auto &r : graph) {
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  return graph;
}

SchurOperator::SchurOperator(const Problem &problem, const Hybrid &hybrid,
                             const std::vector<Dense> &x)
    : p(problem), h(hybrid), X(x) {
  height = p.nl;
}
// src/operators.cpp:147-179 (diagonal plus borrowed-factor entries, sign -1)
void SchurOperator::Mult(const double *x, double *y) const {
  for (int i = 0; i < p.nl; ++i) { y[i] = h.D[i] * x[i]; }
  std::vector<double> local(4 * p.t), mid(p.m), z(4 * p.t);
  for (int e = 0; e < p.n; ++e) {
    for (int i = 0; i < 4 * p.t; ++i) { local[i] = x[h.ids[e][i]]; }
    const Dense &xe = X[e];
    for (int i = 0; i < p.m; ++i) { mid[i] = 0.; }
    for (int j = 0; j < 4 * p.t; ++j) {
      for (int i = 0; i < p.m; ++i) { mid[i] += xe(i, j) * local[j]; }
    }
    h.C[e].Mult(mid.data(), z.data());
    for (int i = 0; i < 4 * p.t; ++i) { y[h.ids[e][i]] += -1. * z[i]; }
  }
}
// src/operators.cpp:275-303
void SchurOperator::DiagonalBlocks(int size, std::vector<Dense> &result) const {
  require(size == p.t, "Schur Jacobi requires complete face blocks");
  result.resize(p.nf);
  for (auto &a : result) { a.resize(size, size); }
  for (int i = 0; i < p.nl; ++i) { result[i / size](i % size, i % size) = h.D[i]; }
  Dense panel;
  for (int e = 0; e < p.n; ++e) {
    for (int side = 0; side < 4; ++side) {
      h.C[e].Panel(X[e], side, side * size, size, panel);
      auto &a = result[h.ids[e][side * size] / size];
      for (size_t i = 0; i < a.d.size(); ++i) { a.d[i] += -1. * panel.d[i]; }
    }
  }
}
// src/operators.cpp:27-47 (factor entries)
std::vector<std::vector<int>> SchurOperator::BlockGraph(int size) const {
  require(size == p.t, "Schur BSR expects face blocks");
  return p.SolverGraph(true);
}
// src/operators.cpp:48-74
void SchurOperator::FillBlocks(FixedBlockILU &target) const {
  const int b = target.block;
  require(b == p.t, "Schur BSR expects face blocks");
  target.Zero();
  for (int r = 0; r < p.nl; r += b) {
    auto &a = target.At(r / b, r / b);
    for (int i = 0; i < b; ++i) { a(i, i) = h.D[r + i]; }
  }
  Dense panel;
  for (int e = 0; e < p.n; ++e) {
    for (int side = 0; side < 4; ++side) {
      for (int other = 0; other < 4; ++other) {
        h.C[e].Panel(X[e], side, other * b, b, panel);
        auto &a = target.At(h.ids[e][side * b] / b, h.ids[e][other * b] / b);
        for (size_t i = 0; i < a.d.size(); ++i) { a.d[i] += -1. * panel.d[i]; }
      }
    }
  }
}

// ------------------------------------------------------------- block ILU(0)
// src/block_ilu.cpp:16-26 calls MFEM's MinimumDiscardedFillOrdering on the
// topology matrix of the graph (every entry 1); MFEM is absent, so this is
// the published rule (Persson and Peraire, SIAM J. Sci. Comput. 30 (2008)
// 2709-2733) on that matrix: eliminate next the uneliminated row whose
// elimination discards the least fill, the weight being the square root of
// the number of ordered pairs (i, j), i != j, of its uneliminated neighbours
// with (i, j) outside the pattern; ties go to the lowest row; the weights of
// the eliminated row's neighbours are then recomputed. Parity with MFEM's
// own tie order is unpinned.
std::vector<int> mdf_ordering(const std::vector<std::vector<int>> &graph) {
  const int n = graph.size();
  std::vector<std::vector<int>> adj(n);
  for (int i = 0; i < n; ++i) {
    adj[i] = graph[i];
    adj[i].push_back(i);
    std::sort(adj[i].begin(), adj[i].end());
    adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
  }
  std::vector<char> eliminated(n, 0);
  auto discarded = [&](int k) {
    int count = 0;
    for (int i : adj[k]) {
      if (i == k || eliminated[i]) { continue; }
      for (int j : adj[k]) {
        if (j == k || j == i || eliminated[j]) { continue; }
        if (!std::binary_search(adj[i].begin(), adj[i].end(), j)) { ++count; }
      }
    }
    return std::sqrt(double(count));
  };
  std::vector<double> w(n);
  for (int k = 0; k < n; ++k) { w[k] = discarded(k); }
  std::vector<int> order;
  for (int step = 0; step < n; ++step) {
    int best = -1;
    for (int k = 0; k < n; ++k) {
      if (!eliminated[k] && (best < 0 || w[k] < w[best])) { best = k; }
    }
    eliminated[best] = 1;
    order.push_back(best);
    for (int j : adj[best]) {
      if (!eliminated[j]) { w[j] = discarded(j); }
    }
  }
  return order;
}
// src/block_ilu.cpp:10-59
FixedBlockILU::FixedBlockILU(const std::vector<std::vector<int>> &graph, int size, const std::string &ordering)
    : block(size) {
  const int nb = graph.size();
  require(block > 0, "Invalid BSR block size");
  permutation.resize(nb);
  inverse.resize(nb);
  for (int i = 0; i < nb; ++i) { permutation[i] = i; }
  if (ordering == "mdf") { permutation = mdf_ordering(graph); }
  for (int i = 0; i < nb; ++i) { inverse[permutation[i]] = i; }
  diag.resize(nb);
  row.push_back(0);
  for (int i = 0; i < nb; ++i) {
    std::vector<int> nbrs{i};
    for (int j : graph[permutation[i]]) { nbrs.push_back(inverse[j]); }
    std::sort(nbrs.begin(), nbrs.end());
    nbrs.erase(std::unique(nbrs.begin(), nbrs.end()), nbrs.end());
    for (int j : nbrs) {
      if (j == i) { diag[i] = col.size(); }
      col.push_back(j);
    }
    row.push_back(col.size());
  }
  data.resize(col.size());
  for (auto &a : data) { a.resize(block, block); }
  pivots.assign(nb, std::vector<int>(block));
  work.assign(size_t(nb) * block, 0.);
  updates.resize(col.size());
  for (int i = 0; i < nb; ++i) {
    for (int ik = row[i]; ik < diag[i]; ++ik) {
      const int k = col[ik];
      int ij = ik + 1, kj = diag[k] + 1;
      while (ij < row[i + 1] && kj < row[k + 1]) {
        if (col[ij] < col[kj]) { ++ij; }
        else if (col[kj] < col[ij]) { ++kj; }
        else { updates[ik].push_back({ij, kj}); ++ij; ++kj; }
      }
    }
  }
}
// src/block_ilu.cpp:60-66
void FixedBlockILU::Zero(bool off_diagonal_only) {
  for (int i = 0; i < int(diag.size()); ++i) {
    for (int j = row[i]; j < row[i + 1]; ++j) {
      if (!off_diagonal_only || j != diag[i]) { data[j].zero(); }
    }
  }
}
// src/block_ilu.cpp:67-72
Dense &FixedBlockILU::At(int original_r, int original_c) {
  const int r = inverse[original_r], c = inverse[original_c];
  auto it = std::lower_bound(col.begin() + row[r], col.begin() + row[r + 1], c);
  require(it != col.begin() + row[r + 1] && *it == c, "Entry outside fixed BSR graph");
  return data[it - col.begin()];
}
// src/block_ilu.cpp:73-86
void FixedBlockILU::Factor() {
  for (int i = 0; i < int(diag.size()); ++i) {
    for (int ik = row[i]; ik < diag[i]; ++ik) {
      const int k = col[ik];
      lu_right_solve(data[diag[k]].data(), pivots[k].data(), block, block, data[ik].data());
      for (const auto &u : updates[ik]) { sub_mult(data[ik], data[u.upper], data[u.target]); }
    }
    require(lu_factor(data[diag[i]].data(), pivots[i].data(), block), "Singular BSR ILU diagonal");
  }
}
// src/block_ilu.cpp:87-109
void FixedBlockILU::Mult(const double *b, double *x) const {
  const int nb = diag.size();
  for (int i = 0; i < nb; ++i) {
    double *wi = work.data() + size_t(i) * block;
    std::copy_n(b + size_t(permutation[i]) * block, block, wi);
    for (int ij = row[i]; ij < diag[i]; ++ij) {
      const double *wj = work.data() + size_t(col[ij]) * block;
      const Dense &a = data[ij];
      for (int jj = 0; jj < block; ++jj) {
        for (int ii = 0; ii < block; ++ii) { wi[ii] -= a(ii, jj) * wj[jj]; }
      }
    }
  }
  for (int i = nb - 1; i >= 0; --i) {
    double *wi = work.data() + size_t(i) * block;
    for (int ij = diag[i] + 1; ij < row[i + 1]; ++ij) {
      const double *wj = work.data() + size_t(col[ij]) * block;
      const Dense &a = data[ij];
      for (int jj = 0; jj < block; ++jj) {
        for (int ii = 0; ii < block; ++ii) { wi[ii] -= a(ii, jj) * wj[jj]; }
      }
    }
    lu_solve(data[diag[i]].data(), pivots[i].data(), block, 1, wi);
    std::copy_n(wi, block, x + size_t(permutation[i]) * block);
  }
}

// ------------------------------------------------------------- GMRES
// src/linear.cpp:27-57
LinearSolver::LinearSolver(int size, int b, const Options &o, const std::vector<double> &weights)
    : n(size), block(b), restart(std::min(size, o.restart)), metric(weights) {
  require(block > 0 && n % block == 0 && int(metric.size()) == n, "Invalid linear workspace dimensions");
  for (double v : metric) { require(std::isfinite(v) && v > 0, "Invalid dual residual metric"); }
  residual.assign(n, 0.); action.assign(n, 0.); preconditioned.assign(n, 0.);
  combination.assign(n, 0.); correction.assign(n, 0.); weighted.assign(n, 0.);
  projection.assign(restart + 1, 0.);
  basis.assign(size_t(n) * (restart + 1), 0.);
  h.assign(size_t(restart + 1) * restart, 0.);
  cosine.assign(restart, 0.); sine.assign(restart, 0.);
  rotated.assign(restart + 1, 0.); coefficients.assign(restart, 0.);
  if (o.solver == "dense" || o.solver == "direct") {
    // src/linear.cpp:46-49 (the reference's direct solver, MKL PARDISO, is
    // an exact solve as well: the same dense LU here)
    require(n <= 5000, "Dense direct solve limited to 5000 unknowns");
    dense_matrix.assign(size_t(n) * n, 0.);
    dense_pivots.assign(n, 0);
  }
  if (o.solver == "gmres" && o.precond == "block-jacobi") {
    diagonal.resize(n / block);
    diagonal_pivots.assign(n / block, std::vector<int>(block));
  }
}
// src/linear.cpp:162-166
void LinearSolver::Prepare(const std::vector<std::vector<int>> &graph, const Options &o) {
  if (!ilu && o.solver == "gmres" && o.precond == "block-ilu0") {
    ilu = std::make_unique<FixedBlockILU>(graph, block, o.ilu_ordering);
  }
}
// src/linear.cpp:58-65
double LinearSolver::Dot(const double *a, const double *b) const {
  double sum = 0.;
  for (int i = 0; i < n; ++i) { sum += metric[i] * a[i] * b[i]; }
  return sum;
}
// src/linear.cpp:66-75
void LinearSolver::Precondition(const double *b, double *x) const {
  if (ilu) { ilu->Mult(b, x); return; }
  for (int e = 0; e < n / block; ++e) {
    std::copy_n(b + size_t(e) * block, block, x + size_t(e) * block);
    lu_solve(diagonal[e].data(), diagonal_pivots[e].data(), block, 1, x + size_t(e) * block);
  }
}
// src/linear.cpp:76-157: right-preconditioned restarted GMRES in the
// metric-weighted inner product, CGS2 Arnoldi, Givens rotations, one P^{-1}
// application per cycle and a true residual per cycle.
double LinearSolver::GMRES(const Operator &a, const double *rhs, double tol, int limit,
                           LinearResult &out, bool zero_guess) {
  auto H = [&](int i, int j) -> double & { return h[i + size_t(j) * (restart + 1)]; };
  if (zero_guess) {
    std::fill(out.x.begin(), out.x.end(), 0.);
    std::copy_n(rhs, n, residual.data());
  } else {
    a.Mult(out.x.data(), action.data());
    for (int i = 0; i < n; ++i) { residual[i] = rhs[i] - action[i]; }
  }
  const double rhs_norm = std::sqrt(Dot(rhs, rhs));
  if (rhs_norm == 0.) { return 0.; }
  int iterations = 0;
  double residual_norm = std::sqrt(Dot(residual.data(), residual.data()));
  while (iterations < limit) {
    const double beta = residual_norm;
    if (beta <= tol * rhs_norm) { break; }
    for (int i = 0; i < n; ++i) { basis[i] = residual[i] * (1 / beta); }
    std::fill(h.begin(), h.end(), 0.);
    std::fill(rotated.begin(), rotated.end(), 0.);
    rotated[0] = beta;
    int count = 0;
    for (int j = 0; j < restart && iterations < limit; ++j) {
      Precondition(basis.data() + size_t(j) * n, preconditioned.data());
      a.Mult(preconditioned.data(), action.data());
      for (int pass = 0; pass < 2; ++pass) {
        for (int k = 0; k < n; ++k) { weighted[k] = metric[k] * action[k]; }
        for (int i = 0; i <= j; ++i) {
          double s = 0.;
          const double *v = basis.data() + size_t(i) * n;
          for (int k = 0; k < n; ++k) { s += v[k] * weighted[k]; }
          projection[i] = s;
          H(i, j) += s;
        }
        for (int k = 0; k < n; ++k) {
          double s = 0.;
          for (int i = 0; i <= j; ++i) { s += basis[size_t(i) * n + k] * projection[i]; }
          combination[k] = s;
        }
        for (int k = 0; k < n; ++k) { action[k] -= combination[k]; }
      }
      const double next = std::sqrt(Dot(action.data(), action.data()));
      H(j + 1, j) = next;
      if (next > 0) {
        double *v = basis.data() + size_t(j + 1) * n;
        for (int k = 0; k < n; ++k) { v[k] = action[k] * (1 / next); }
      }
      for (int i = 0; i < j; ++i) {
        const double u = H(i, j), vr = H(i + 1, j);
        H(i, j) = cosine[i] * u + sine[i] * vr;
        H(i + 1, j) = -sine[i] * u + cosine[i] * vr;
      }
      const double d = std::hypot(H(j, j), H(j + 1, j));
      require(std::isfinite(d) && d > 0, "GMRES Arnoldi breakdown");
      cosine[j] = H(j, j) / d;
      sine[j] = H(j + 1, j) / d;
      H(j, j) = d;
      H(j + 1, j) = 0.;
      rotated[j + 1] = -sine[j] * rotated[j];
      rotated[j] *= cosine[j];
      ++iterations;
      count = j + 1;
      if (std::abs(rotated[j + 1]) <= tol * rhs_norm || next <= std::numeric_limits<double>::min()) {
        break;
      }
    }
    for (int i = count - 1; i >= 0; --i) {
      double value = rotated[i];
      for (int j = i + 1; j < count; ++j) { value -= H(i, j) * coefficients[j]; }
      coefficients[i] = value / H(i, i);
    }
    for (int k = 0; k < n; ++k) {
      double s = 0.;
      for (int i = 0; i < count; ++i) { s += basis[size_t(i) * n + k] * coefficients[i]; }
      combination[k] = s;
    }
    Precondition(combination.data(), correction.data());
    for (int k = 0; k < n; ++k) { out.x[k] += correction[k]; }
    a.Mult(out.x.data(), action.data());
    for (int k = 0; k < n; ++k) { residual[k] = rhs[k] - action[k]; }
    residual_norm = std::sqrt(Dot(residual.data(), residual.data()));
    require(std::isfinite(residual_norm), "Nonfinite GMRES residual");
  }
  out.count = iterations;
  require(residual_norm <= tol * rhs_norm, "GMRES iteration limit reached");
  return residual_norm;
}
// src/linear.cpp:167-249 (gmres, matrix-free)
void LinearSolver::Solve(const Operator &op, const double *rhs, const Options &o,
                         LinearResult &out, bool refresh) {
  out.x.resize(n);
  out.equation_residual.resize(n);
  out.count = 0;
  out.residual = out.assembly = out.setup = out.solve = 0.;
  const double rhs_norm = std::sqrt(Dot(rhs, rhs));
  if (rhs_norm == 0.) {
    std::fill(out.x.begin(), out.x.end(), 0.);
    std::fill(out.equation_residual.begin(), out.equation_residual.end(), 0.);
    return;
  }
  if (o.solver != "gmres") {
    // src/linear.cpp:184-213,224-240: the assembled matrix (column j = A e_j;
    // the reference assembles the same entries from its element blocks),
    // its LU with partial pivoting, the solve and the operator's residual
    double start = clock_s();
    if (refresh) {
      std::vector<double> unit(n, 0.), column(n);
      for (int j = 0; j < n; ++j) {
        unit[j] = 1.;
        op.Mult(unit.data(), column.data());
        unit[j] = 0.;
        std::copy(column.begin(), column.end(), dense_matrix.begin() + size_t(j) * n);
      }
    }
    out.assembly = clock_s() - start;
    start = clock_s();
    if (refresh) { require(lu_factor(dense_matrix.data(), dense_pivots.data(), n), "Singular dense LU factor"); }
    out.setup = clock_s() - start;
    start = clock_s();
    std::copy(rhs, rhs + n, out.x.begin());
    lu_solve(dense_matrix.data(), dense_pivots.data(), n, 1, out.x.data());
    out.solve = clock_s() - start;
    op.Mult(out.x.data(), out.equation_residual.data());
    for (int i = 0; i < n; ++i) { out.equation_residual[i] -= rhs[i]; }
    out.residual = std::sqrt(Dot(out.equation_residual.data(), out.equation_residual.data())) / rhs_norm;
    require(std::isfinite(out.residual) && out.residual <= std::max(10 * o.linear_tol, 1e-12),
            "True inverse-mass weighted linear residual too large");
    return;
  }
  if (o.precond == "block-ilu0" && !ilu) { Prepare(op.BlockGraph(block), o); }
  double start = clock_s();
  if (refresh && ilu) { op.FillBlocks(*ilu); }
  out.assembly = clock_s() - start;
  start = clock_s();
  if (refresh) {
    if (ilu) {
      ilu->Factor();
    } else {
      op.DiagonalBlocks(block, diagonal);
      for (size_t e = 0; e < diagonal.size(); ++e) {
        require(lu_factor(diagonal[e].data(), diagonal_pivots[e].data(), block),
                "Singular block Jacobi factor");
      }
    }
  }
  out.setup = clock_s() - start;
  start = clock_s();
  const double gmres_norm = GMRES(op, rhs, o.linear_tol, o.max_krylov, out, refresh);
  out.solve = clock_s() - start;
  if (o.matrix_free) {
    for (int i = 0; i < n; ++i) { out.equation_residual[i] = -residual[i]; }
    out.residual = gmres_norm / rhs_norm;
  } else {
    // src/linear.cpp:236-240: an assembled run keeps the operator's own residual
    op.Mult(out.x.data(), out.equation_residual.data());
    for (int i = 0; i < n; ++i) { out.equation_residual[i] -= rhs[i]; }
    out.residual = std::sqrt(Dot(out.equation_residual.data(), out.equation_residual.data())) / rhs_norm;
  }
  require(std::isfinite(out.residual) && out.residual <= std::max(10 * o.linear_tol, 1e-12),
          "True inverse-mass weighted linear residual too large");
}
} // namespace oracle
