// CPU ORACLE -- test infrastructure only (see oracle.hpp).
// Physics, discretisation tables, residual and local Jacobian blocks.
#include "oracle.hpp"
#include <algorithm>
#include <chrono>
#include <fstream>
#include <iomanip>

namespace oracle {
double clock_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ---------------------------------------------------------------- physics
// src/physics.cpp (include/euler.hpp:33-35 for the constructor check)
EulerGas::EulerGas(double value) : gamma(value), gamma_minus_one(value - 1.) {
  require(std::isfinite(value) && value > 1., "Euler gamma must be finite and > 1");
}
// src/physics.cpp:5-7
double EulerGas::pressure(const State &u) const {
  return gamma_minus_one * (u[3] - .5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
}
// src/physics.cpp:8-11
bool EulerGas::positive(const State &u) const {
  for (double v : u) { if (!std::isfinite(v)) { return false; } }
  return u[0] > 1e-13 && pressure(u) > 1e-13;
}
// src/physics.cpp:12-17
State EulerGas::flux(const State &u, int axis) const {
  const double vn = u[axis + 1] / u[0], pr = pressure(u);
  State f = {u[axis + 1], u[1] * vn, u[2] * vn, (u[3] + pr) * vn};
  f[axis + 1] += pr;
  return f;
}
// src/physics.cpp:18-39: dF_axis/dU, column-major 4x4.
void EulerGas::jacobian(const State &u, int axis, Mat4 &j) const {
  const double inv = 1. / u[0];
  const double vx = u[1] * inv, vy = u[2] * inv, vn = u[axis + 1] * inv;
  const double enthalpy = u[3] + pressure(u);
  const State dp = {.5 * gamma_minus_one * (vx * vx + vy * vy), -gamma_minus_one * vx,
                    -gamma_minus_one * vy, gamma_minus_one};
  State dv = {-vn * inv, 0, 0, 0};
  dv[axis + 1] = inv;
  j = Mat4{};
  j(0, axis + 1) = 1;
  for (int d = 0; d < 2; d++) {
    for (int c = 0; c < 4; c++) { j(d + 1, c) = u[d + 1] * dv[c] + (axis == d ? dp[c] : 0); }
    j(d + 1, d + 1) += vn;
  }
  for (int c = 0; c < 4; c++) { j(3, c) = vn * dp[c] + enthalpy * dv[c]; }
  j(3, 3) += vn;
}
// src/physics.cpp:40-63: Einfeldt bounds with Roe averages, scaled by safety.
std::array<double, 2> EulerGas::speeds(const State &l, const State &r, int axis,
                                       double safety) const {
  bool finite = true;
  for (int c = 0; c < 4; ++c) { finite = finite && std::isfinite(l[c]) && std::isfinite(r[c]); }
  require(finite && l[0] > 1e-13 && r[0] > 1e-13, "Nonpositive old face state; refine mesh");
  const double pl = pressure(l), pr = pressure(r);
  require(pl > 1e-13 && pr > 1e-13, "Nonpositive old face state; refine mesh");
  const double il = 1. / l[0], ir = 1. / r[0];
  const double sl = std::sqrt(l[0]), sr = std::sqrt(r[0]);
  double v[2];
  for (int d = 0; d < 2; d++) { v[d] = (sl * l[d + 1] * il + sr * r[d + 1] * ir) / (sl + sr); }
  const double hroe = (sl * (l[3] + pl) * il + sr * (r[3] + pr) * ir) / (sl + sr);
  const double c2 = gamma_minus_one * (hroe - .5 * (v[0] * v[0] + v[1] * v[1]));
  require(c2 > 0, "Invalid Roe sound speed");
  const double ct = std::sqrt(c2), cl = std::sqrt(gamma * pl * il), cr = std::sqrt(gamma * pr * ir);
  return {safety * std::max({0., cl - l[axis + 1] * il, ct - v[axis]}),
          safety * std::max({0., cr + r[axis + 1] * ir, ct + v[axis]})};
}
// src/physics.cpp:64-93: rotated simple wave; the left-going characteristic
// map xi + t*amp*sin(pi xi) = phase (mod 2) is solved by safeguarded Newton.
State EulerGas::exact(double x, double y, double time, double alpha) const {
  const double drift = (3. - gamma) / 2., amp = (gamma + 1.) / 4.;
  require(time >= 0 && time * pi * amp < 1, "Exact solution requires T < 4/((gamma+1)*pi)");
  double phase = std::cos(alpha) * x + std::sin(alpha) * y;
  phase -= 2 * std::floor((phase + 1) / 2);
  phase -= time * drift;
  phase -= 2 * std::floor((phase + 1) / 2);
  double lo = phase - time * amp, hi = phase + time * amp, z = phase;
  for (int it = 0; it < 80; it++) {
    const double res = z + time * amp * std::sin(pi * z) - phase;
    if (std::abs(res) < 2e-15) { break; }
    if (res > 0) { hi = z; } else { lo = z; }
    const double cand = z - res / (1 + time * amp * pi * std::cos(pi * z));
    z = (cand >= lo && cand <= hi) ? cand : .5 * (lo + hi);
  }
  const double jm = std::sin(pi * z), v = 1 + .5 * jm;
  const double sound = gamma_minus_one * (2. - jm) / 4.;
  const double rho = std::pow(sound * sound / gamma, 1. / gamma_minus_one);
  const double pr = std::pow(rho, gamma);
  return {rho, rho * v * std::cos(alpha), rho * v * std::sin(alpha),
          pr / gamma_minus_one + .5 * rho * v * v};
}

// ---------------------------------------------------------------- 1-D tables
namespace {
// Gauss-Legendre nodes/weights on [0,1] (ascending): Newton on P_n in long
// double, the textbook construction MFEM's IntRules also uses.
void gauss_legendre(int np, std::vector<double> &x, std::vector<double> &wt) {
  x.assign(np, 0.);
  wt.assign(np, 0.);
  for (int i = 0; i < np; ++i) {
    long double z = std::cos(pi * (np - i - .25L) / (np + .5L));
    long double dp = 1;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1, p1 = z;
      for (int j = 2; j <= np; ++j) {
        const long double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      if (np == 1) { p1 = z; p0 = 1; }
      dp = np * (z * p1 - p0) / (z * z - 1);
      const long double dz = p1 / dp;
      z -= dz;
      if (std::fabs((double)dz) < 1e-19) { break; }
    }
    long double p0 = 1, p1 = z;
    for (int j = 2; j <= np; ++j) {
      const long double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
    dp = np * (z * p1 - p0) / (z * z - 1);
    x[i] = double((1 + z) / 2);
    wt[i] = double(1 / ((1 - z * z) * dp * dp));
  }
  if (np == 1) { x[0] = .5; wt[0] = 1.; }
}
long double lagrange(const std::vector<double> &x, int i, long double at) {
  long double v = 1;
  for (int j = 0; j < (int)x.size(); ++j) {
    if (j != i) { v *= (at - x[j]) / ((long double)x[i] - x[j]); }
  }
  return v;
}
long double lagrange_derivative(const std::vector<double> &x, int i, long double at) {
  long double s = 0;
  for (int k = 0; k < (int)x.size(); ++k) {
    if (k == i) { continue; }
    long double term = 1 / ((long double)x[i] - x[k]);
    for (int j = 0; j < (int)x.size(); ++j) {
      if (j != i && j != k) { term *= (at - x[j]) / ((long double)x[i] - x[j]); }
    }
    s += term;
  }
  return s;
}
} // namespace

// ---------------------------------------------------------------- Problem
// src/problem.cpp:32-223. Nodes and volume points are lexicographic with x
// fastest (MFEM's native tensor ordering, which the reference maps through
// tensor_node/tensor_quad, src/problem.cpp:172-200); cells are e = iy*nx+ix.
Problem::Problem(const Options &o, const EulerGas &g)
    : opt(o), gas(g), alpha(o.alpha * pi / 180), lx(2 / std::cos(alpha)),
      ly(2 / std::sin(alpha)), hx(lx / o.nx), hy(ly / o.ny), area(hx * hy), n(o.nx * o.ny),
      p((o.k + 1) * (o.k + 1)), m(4 * p), t(4 * (o.k + 1)), nf(2 * n), nu(n * m),
      nl(nf * t), q(o.k + 1) {
  old.assign(nu, 0.);
  require(o.steps > 0 || o.end < 4 / ((gas.Gamma() + 1) * pi),
          "Exact solution requires T < 4/((gamma+1)*pi)");
  require(o.alpha > 0 && o.alpha < 90, "Angle must be strictly between 0 and 90 degrees");
  require(o.k >= 0 && o.k <= 8, "Expected native Qk element");
  origx.resize(n);
  origy.resize(n);
  for (int iy = 0; iy < o.ny; ++iy) {
    for (int ix = 0; ix < o.nx; ++ix) {
      origx[iy * o.nx + ix] = double(ix) / o.nx * lx - .5 * lx;
      origy[iy * o.nx + ix] = double(iy) / o.ny * ly - .5 * ly;
    }
  }
  // Periodic faces, src/problem.cpp:73-90.
  cell_faces.assign(n, {-1, -1, -1, -1});
  for (int iy = 0; iy < o.ny; iy++) {
    for (int ix = 0; ix < o.nx; ix++) {
      const int l = iy * o.nx + ix;
      Face fx, fy;
      fx.l = l; fx.r = iy * o.nx + (ix + 1) % o.nx; fx.ls = 1; fx.rs = 0; fx.axis = 0; fx.length = hy;
      fy.l = l; fy.r = ((iy + 1) % o.ny) * o.nx + ix; fy.ls = 3; fy.rs = 2; fy.axis = 1; fy.length = hx;
      for (Face f : {fx, fy}) {
        const int id = faces.size();
        f.a.assign(q, 0.); f.b.assign(q, 0.);
        faces.push_back(f);
        cell_faces[f.l][f.ls] = id;
        cell_faces[f.r][f.rs] = id;
      }
    }
  }
  // GL tables, src/problem.cpp:91-162.
  gauss_legendre(q, x1, w1);
  D1.resize(q, q);
  end0.resize(q);
  end1.resize(q);
  for (int i = 0; i < q; ++i) {
    end0[i] = double(lagrange(x1, i, 0.L));
    end1[i] = double(lagrange(x1, i, 1.L));
    if (std::abs(end0[i]) < 1e-14) { end0[i] = 0; }
    if (std::abs(end1[i]) < 1e-14) { end1[i] = 0; }
    for (int a = 0; a < q; ++a) {
      const double v = double(lagrange_derivative(x1, i, x1[a]));
      D1(a, i) = std::abs(v) < 1e-14 ? 0. : v;
    }
  }
  w.resize(p);
  inv_mass.resize(p);
  for (int iy = 0; iy < q; ++iy) {
    for (int ix = 0; ix < q; ++ix) {
      w[iy * q + ix] = w1[ix] * w1[iy];
      inv_mass[iy * q + ix] = 1. / w[iy * q + ix];
    }
  }
  // Volume test functions per point, src/problem.cpp:163-170 (Gx/Gy of
  // src/problem.cpp:108-111: reference derivative over hx or hy).
  volume_tests.resize(p);
  for (int qy = 0; qy < q; ++qy) {
    for (int qx = 0; qx < q; ++qx) {
      for (int i = 0; i < p; ++i) {
        const int ix = i % q, iy = i / q;
        const double dx = iy == qy ? D1(qx, ix) / hx : 0.;
        const double dy = ix == qx ? D1(qy, iy) / hy : 0.;
        if (dx != 0 || dy != 0) { volume_tests[qy * q + qx].push_back({i, dx, dy}); }
      }
    }
  }
  // Face restriction tables, src/problem.cpp:150-162,201-208. Side 0: x=0,
  // 1: x=1, 2: y=0, 3: y=1; face point qq is the tangential GL node.
  for (int side = 0; side < 4; ++side) {
    face_basis[side].resize(q);
    for (int qq = 0; qq < q; ++qq) {
      for (int i = 0; i < p; ++i) {
        const int ix = i % q, iy = i / q;
        double v = 0;
        if (side < 2) { v = iy == qq ? (side ? end1[ix] : end0[ix]) : 0.; }
        else { v = ix == qq ? (side == 3 ? end1[iy] : end0[iy]) : 0.; }
        if (v != 0) { face_basis[side][qq].push_back({i, v}); }
      }
    }
  }
  // Dual metrics, src/problem.cpp:209-222.
  cell_metric.resize(nu);
  trace_metric.resize(nl);
  for (int e = 0; e < n; ++e) {
    for (int i = 0; i < p; ++i) {
      for (int c = 0; c < 4; ++c) { cell_metric[e * m + 4 * i + c] = area * w[i]; }
    }
  }
  for (int f = 0; f < nf; ++f) {
    for (int qq = 0; qq < q; ++qq) {
      for (int c = 0; c < 4; ++c) { trace_metric[f * t + 4 * qq + c] = faces[f].length / w1[qq]; }
    }
  }
}

// src/problem.cpp:238-258 (collocated volume and tensor face evaluation)
State Problem::Node(const double *u, int e, int node) const {
  return {u[e * m + 4 * node], u[e * m + 4 * node + 1], u[e * m + 4 * node + 2],
          u[e * m + 4 * node + 3]};
}
State Problem::Trace(const double *u, int e, int side, int qq) const {
  State a{};
  for (const auto &entry : face_basis[side][qq]) {
    for (int c = 0; c < 4; ++c) { a[c] += entry.second * u[e * m + 4 * entry.first + c]; }
  }
  return a;
}
// src/problem.cpp:259-275
bool Problem::Admissible(const double *u) const {
  for (int e = 0; e < n; e++) {
    for (int i = 0; i < p; i++) {
      if (!gas.positive(Node(u, e, i))) { return false; }
    }
    for (int s = 0; s < 4; s++) {
      for (int qq = 0; qq < q; qq++) {
        if (!gas.positive(Trace(u, e, s, qq))) { return false; }
      }
    }
  }
  return true;
}
// src/problem.cpp:276-303 (collocated branch)
void Problem::MassSolve(Dense &a) const {
  for (int i = 0; i < p; ++i) {
    const double scale = inv_mass[i];
    for (int j = 0; j < a.w; ++j) {
      for (int c = 0; c < 4; ++c) { a(4 * i + c, j) *= scale; }
    }
  }
}
// src/problem.cpp:304-313
void Problem::MassSolve(double *r) const {
  for (int e = 0; e < n; ++e) {
    for (int i = 0; i < p; ++i) {
      for (int c = 0; c < 4; ++c) { r[e * m + 4 * i + c] *= inv_mass[i]; }
    }
  }
}
// src/problem.cpp:329-349: L2 projection with the 13x13 tensor GL rule that
// MFEM's IntRules returns for order max(24,2k+8)=24 (k<=8).
std::vector<double> Problem::Initial() const {
  std::vector<double> u(nu, 0.), gx, gw;
  const int nq = std::max(24, 2 * opt.k + 8) / 2 + 1;
  gauss_legendre(nq, gx, gw);
  std::vector<double> bx(size_t(nq) * q);
  for (int a = 0; a < nq; ++a) {
    for (int i = 0; i < q; ++i) { bx[a * q + i] = double(lagrange(x1, i, gx[a])); }
  }
  for (int e = 0; e < n; e++) {
    for (int qy = 0; qy < nq; qy++) {
      for (int qx = 0; qx < nq; qx++) {
        const double weight = gw[qx] * gw[qy];
        const auto v = gas.exact(origx[e] + hx * gx[qx], origy[e] + hy * gx[qy], 0, alpha);
        for (int i = 0; i < p; i++) {
          const double sh = bx[qx * q + i % q] * bx[qy * q + i / q];
          for (int c = 0; c < 4; c++) { u[e * m + 4 * i + c] += weight * sh * v[c]; }
        }
      }
    }
  }
  MassSolve(u.data());
  return u;
}
// src/problem.cpp:350-376
void Problem::Freeze(const std::vector<double> &u, double step, bool check_admissible) {
  ++freeze_generation;
  if (check_admissible) { require(Admissible(u.data()), "Old state is inadmissible; refine mesh"); }
  old = u;
  dt = step;
  for (auto &f : faces) {
    f.a.resize(q); f.b.resize(q); f.inverse_sum.resize(q);
    f.left_weight.resize(q); f.right_weight.resize(q); f.penalty.resize(q);
    for (int qq = 0; qq < q; qq++) {
      const auto s = gas.speeds(Trace(u.data(), f.l, f.ls, qq), Trace(u.data(), f.r, f.rs, qq),
                                f.axis, opt.safety);
      f.a[qq] = s[0];
      f.b[qq] = s[1];
      require(s[0] + s[1] > 0, "Degenerate HLL speeds");
      f.inverse_sum[qq] = 1. / (s[0] + s[1]);
      f.left_weight[qq] = s[1] * f.inverse_sum[qq];
      f.right_weight[qq] = s[0] * f.inverse_sum[qq];
      f.penalty[qq] = s[0] * s[1] * f.inverse_sum[qq];
    }
  }
}
// src/problem.cpp:377-384
double Problem::WaveRate() const {
  double speed = 0.;
  for (const auto &f : faces) {
    for (int qq = 0; qq < q; ++qq) { speed = std::max({speed, f.a[qq], f.b[qq]}); }
  }
  return speed / std::min(hx, hy);
}
// src/problem.cpp:385-391
double Problem::CellNorm(const double *r) const {
  double norm = 0.;
  for (int i = 0; i < nu; ++i) { norm += cell_metric[i] * r[i] * r[i]; }
  return std::sqrt(norm);
}
// src/problem.cpp:445-507 (lam == nullptr: the frozen-HLL DG residual)
void Problem::Residual(const double *u, double *r) const {
  std::fill(r, r + nu, 0.);
  for (int e = 0; e < n; e++) {
    for (int qq = 0; qq < p; qq++) {
      const auto v = Node(u, e, qq), fx = gas.flux(v, 0), fy = gas.flux(v, 1);
      for (const auto &test : volume_tests[qq]) {
        double *cell = r + e * m + 4 * test.node;
        const double weight = dt * w[qq];
        for (int c = 0; c < 4; c++) { cell[c] -= weight * (test.dx * fx[c] + test.dy * fy[c]); }
      }
    }
  }
  for (int f = 0; f < nf; f++) {
    const auto &fc = faces[f];
    for (int qq = 0; qq < q; qq++) {
      const auto l = Trace(u, fc.l, fc.ls, qq), v = Trace(u, fc.r, fc.rs, qq);
      const auto fl = gas.flux(l, fc.axis), fr = gas.flux(v, fc.axis);
      State hl, hr;
      for (int c = 0; c < 4; c++) {
        hl[c] = fc.left_weight[qq] * fl[c] + fc.right_weight[qq] * fr[c] +
                fc.penalty[qq] * (l[c] - v[c]);
        hr[c] = -hl[c];
      }
      const double scale = dt * fc.length / area * w1[qq];
      for (const auto &entry : face_basis[fc.ls][qq]) {
        double *cell = r + fc.l * m + 4 * entry.first;
        const double weight = scale * entry.second;
        for (int c = 0; c < 4; ++c) { cell[c] += weight * hl[c]; }
      }
      for (const auto &entry : face_basis[fc.rs][qq]) {
        double *cell = r + fc.r * m + 4 * entry.first;
        const double weight = scale * entry.second;
        for (int c = 0; c < 4; ++c) { cell[c] += weight * hr[c]; }
      }
    }
  }
  MassSolve(r);
  for (int i = 0; i < nu; ++i) { r[i] += u[i]; }
  for (int i = 0; i < nu; ++i) { r[i] -= old[i]; }
}
// src/problem.cpp:519-550 (collocated)
void Problem::VolumeJacobianCell(const double *u, int e, Dense &a, const std::vector<Mat4> *cx,
                                 const std::vector<Mat4> *cy) const {
  a.resize(m, m);
  Mat4 jxl, jyl;
  for (int qq = 0; qq < p; qq++) {
    if (!cx) {
      const auto v = Node(u, e, qq);
      gas.jacobian(v, 0, jxl);
      gas.jacobian(v, 1, jyl);
    }
    const Mat4 &jx = cx ? (*cx)[e * p + qq] : jxl;
    const Mat4 &jy = cy ? (*cy)[e * p + qq] : jyl;
    for (const auto &test : volume_tests[qq]) {
      const int i = test.node, j = qq;
      const double weight = dt * w[qq];
      for (int d = 0; d < 4; ++d) {
        for (int c = 0; c < 4; ++c) {
          a(4 * i + c, 4 * j + d) -= weight * (test.dx * jx(c, d) + test.dy * jy(c, d));
        }
      }
    }
  }
  MassSolve(a);
  for (int i = 0; i < m; i++) { a(i, i) += 1; }
}
// src/problem.cpp:552-590 with add_face (src/problem.cpp:12-30): the dense
// global DG Jacobian, a verification aid for small meshes.
void Problem::DenseDG(const double *u, Dense &a) const {
  a.resize(nu, nu);
  Dense cell;
  for (int e = 0; e < n; ++e) {
    VolumeJacobianCell(u, e, cell);
    for (int j = 0; j < m; ++j) {
      for (int i = 0; i < m; ++i) { a(e * m + i, e * m + j) += cell(i, j); }
    }
  }
  for (const auto &f : faces) {
    for (int qq = 0; qq < q; qq++) {
      Mat4 jl, jr;
      gas.jacobian(Trace(u, f.l, f.ls, qq), f.axis, jl);
      gas.jacobian(Trace(u, f.r, f.rs, qq), f.axis, jr);
      for (int d = 0; d < 4; d++) {
        for (int c = 0; c < 4; c++) {
          jl(c, d) = f.left_weight[qq] * jl(c, d) + (c == d ? f.penalty[qq] : 0);
          jr(c, d) = f.right_weight[qq] * jr(c, d) - (c == d ? f.penalty[qq] : 0);
        }
      }
      const double s = dt * f.length / area * w1[qq];
      const int cells[2] = {f.l, f.r}, sides[2] = {f.ls, f.rs};
      for (int ti = 0; ti < 2; ++ti) {
        for (int tj = 0; tj < 2; ++tj) {
          const Mat4 &jac = tj ? jr : jl;
          const double sign = ti ? -s : s;
          for (const auto &ie : face_basis[sides[ti]][qq]) {
            for (const auto &je : face_basis[sides[tj]][qq]) {
              const double weight = sign * ie.second * je.second * inv_mass[ie.first];
              for (int d = 0; d < 4; ++d) {
                for (int c = 0; c < 4; ++c) {
                  a(cells[ti] * m + 4 * ie.first + c, cells[tj] * m + 4 * je.first + d) +=
                      weight * jac(c, d);
                }
              }
            }
          }
        }
      }
    }
  }
}
// src/problem.cpp:704-726
std::array<double, 4> Problem::Errors(const double *u, double time) const {
  std::array<double, 4> a{};
  std::vector<double> gx, gw;
  const int nq = std::max(24, 2 * opt.k + 8) / 2 + 1;
  gauss_legendre(nq, gx, gw);
  std::vector<double> bx(size_t(nq) * q);
  for (int s = 0; s < nq; ++s) {
    for (int i = 0; i < q; ++i) { bx[s * q + i] = double(lagrange(x1, i, gx[s])); }
  }
  for (int e = 0; e < n; e++) {
    for (int qy = 0; qy < nq; qy++) {
      for (int qx = 0; qx < nq; qx++) {
        const auto v = gas.exact(origx[e] + hx * gx[qx], origy[e] + hy * gx[qy], time, alpha);
        for (int c = 0; c < 4; c++) {
          double z = 0;
          for (int i = 0; i < p; i++) {
            z += bx[qx * q + i % q] * bx[qy * q + i / q] * u[e * m + 4 * i + c];
          }
          a[c] += area * gw[qx] * gw[qy] * (z - v[c]) * (z - v[c]);
        }
      }
    }
  }
  for (auto &v : a) { v = std::sqrt(v); }
  return a;
}
// src/problem.cpp:727-738
std::array<double, 4> Problem::Integral(const double *u) const {
  std::array<double, 4> a{};
  for (int e = 0; e < n; e++) {
    for (int i = 0; i < p; i++) {
      for (int c = 0; c < 4; c++) { a[c] += area * w[i] * u[e * m + 4 * i + c]; }
    }
  }
  return a;
}
// src/problem.cpp:739-761
std::array<double, 4> Problem::Differences(const double *u, const double *v) const {
  std::vector<double> d(nu);
  for (int i = 0; i < nu; ++i) { d[i] = u[i] - v[i]; }
  auto a = std::array<double, 4>{};
  std::vector<double> gx, gw;
  const int nq = std::max(24, 2 * opt.k + 8) / 2 + 1;
  gauss_legendre(nq, gx, gw);
  std::vector<double> bx(size_t(nq) * q);
  for (int s = 0; s < nq; ++s) {
    for (int i = 0; i < q; ++i) { bx[s * q + i] = double(lagrange(x1, i, gx[s])); }
  }
  for (int e = 0; e < n; e++) {
    for (int qy = 0; qy < nq; qy++) {
      for (int qx = 0; qx < nq; qx++) {
        for (int c = 0; c < 4; ++c) {
          double z = 0.;
          for (int i = 0; i < p; ++i) {
            z += bx[qx * q + i % q] * bx[qy * q + i / q] * d[e * m + 4 * i + c];
          }
          a[c] += area * gw[qx] * gw[qy] * z * z;
        }
      }
    }
  }
  for (auto &v : a) { v = std::sqrt(v); }
  return a;
}
// src/problem.cpp:762-790 (coefficient CSV; the MFEM .mesh/.gf writers are
// not restated)
void Problem::Save(const double *u, const std::string &prefix) const {
  std::ofstream cs(prefix + "_coefficients.csv");
  cs << "element,node,rho,rhou,rhov,energy\n" << std::setprecision(17);
  for (int e = 0; e < n; e++) {
    for (int i = 0; i < p; i++) {
      cs << e << ',' << i;
      for (int c = 0; c < 4; c++) { cs << ',' << u[e * m + 4 * i + c]; }
      cs << '\n';
    }
  }
}
} // namespace oracle
