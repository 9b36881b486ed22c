// --verify: the reference's built-in algebra and physics self-test
// (src/verification.cpp:7-169,257-303) run on the device kernels. For every
// degree k = 0..8 on a 3 x 2 mesh at the state .05 Initial + .95
// exact(0,0,0) with frozen bounds at dt = .001, the device results are
// checked against each other and against host references built here:
//  * the matrix-free Jacobian action against a central difference of the
//    device residual (< 1e-7) and against the device block fill (the BSR
//    blocks multiplied on the host, < 1e-13);
//  * the DG increment (block-Jacobi GMRES) against the HDG one (condensed
//    Schur solve and recovery) (< 1e-8);
//  * the device Schur action and its diagonal blocks against the dense Schur
//    complement D - sum_e C_e A_e^-1 B_e formed on the host from the
//    device-built A_e, C_e (host LU) and the face lifting B_e (< 1e-12);
//  * the device block ILU(0) of both systems, filled twice, against a
//    textbook scalar ILU(0) of the dense matrix on the same block pattern
//    (< 1e-10; the reference compares with mfem::BlockILU);
//  * the GL mass matrix is diagonal (< 1e-12).
// Writes <output>_verification.csv and <output>_storage.csv with the
// reference's columns; the off-manifold trace check uses the reference's
// verify-only HDG trace forms, which this build does not have (-1).
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "euler.hpp"

namespace euler {
namespace {
std::vector<double> down(const Problem &p, const double *dev, size_t n) {
  std::vector<double> h(n);
  hip_check(hipMemcpyAsync(h.data(), dev, sizeof(double) * n, hipMemcpyDeviceToHost, p.stream), "download");
  hip_check(hipStreamSynchronize(p.stream), "sync");
  return h;
}
void up(const Problem &p, double *dev, const std::vector<double> &h) {
  hip_check(hipMemcpyAsync(dev, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, p.stream), "upload");
  hip_check(hipStreamSynchronize(p.stream), "sync");
}
double norm2(const std::vector<double> &a) {
  double s = 0.;
  for (double v : a) { s += v * v; }
  return std::sqrt(s);
}
double rel(const std::vector<double> &a, const std::vector<double> &b) {
  std::vector<double> d(a.size());
  for (size_t i = 0; i < a.size(); ++i) { d[i] = a[i] - b[i]; }
  return norm2(d) / std::max(norm2(b), 1e-300);
}
std::vector<double> sines(size_t n, double f, int shift) {
  std::vector<double> v(n);
  for (size_t i = 0; i < n; ++i) { v[i] = std::sin(f * double(int(i) + shift)); }
  return v;
}
// Row-major dense matrix.
struct Dense {
  int n = 0;
  std::vector<double> a;
  explicit Dense(int size) : n(size), a(size_t(size) * size, 0.) {}
  double &operator()(int i, int j) { return a[size_t(i) * n + j]; }
  double operator()(int i, int j) const { return a[size_t(i) * n + j]; }
  std::vector<double> mult(const std::vector<double> &x) const {
    std::vector<double> y(n, 0.);
    for (int i = 0; i < n; ++i) {
      double s = 0.;
      for (int j = 0; j < n; ++j) { s += (*this)(i, j) * x[j]; }
      y[i] = s;
    }
    return y;
  }
};
// Solves a X = b in place for the columns of b (column-major, nrhs of them)
// by LU with partial pivoting of a copy of a (column-major n x n).
void lu_solve(std::vector<double> a, int n, std::vector<double> &b, int nrhs) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i) {
      if (std::abs(a[i + size_t(k) * n]) > std::abs(a[p + size_t(k) * n])) { p = i; }
    }
    require(a[p + size_t(k) * n] != 0., "Singular cell matrix in the verification");
    piv[k] = p;
    if (p != k) {
      for (int j = 0; j < n; ++j) { std::swap(a[k + size_t(j) * n], a[p + size_t(j) * n]); }
    }
    for (int i = k + 1; i < n; ++i) { a[i + size_t(k) * n] /= a[k + size_t(k) * n]; }
    for (int j = k + 1; j < n; ++j) {
      const double akj = a[k + size_t(j) * n];
      for (int i = k + 1; i < n; ++i) { a[i + size_t(j) * n] -= a[i + size_t(k) * n] * akj; }
    }
  }
  for (int r = 0; r < nrhs; ++r) {
    double *x = b.data() + size_t(r) * n;
    for (int k = 0; k < n; ++k) {
      std::swap(x[k], x[piv[k]]);
      for (int i = k + 1; i < n; ++i) { x[i] -= a[i + size_t(k) * n] * x[k]; }
    }
    for (int k = n - 1; k >= 0; --k) {
      x[k] /= a[k + size_t(k) * n];
      for (int i = 0; i < k; ++i) { x[i] -= a[i + size_t(k) * n] * x[k]; }
    }
  }
}
// Dense matrix of column-major BSR blocks on a pattern.
Dense dense_bsr(const BlockPattern &pt, const std::vector<double> &blocks) {
  const int b = pt.block;
  Dense a(pt.nb * b);
  for (int r = 0; r < pt.nb; ++r) {
    for (int s = pt.row[r]; s < pt.row[r + 1]; ++s) {
      const double *blk = blocks.data() + size_t(s) * b * b;
      for (int j = 0; j < b; ++j) {
        for (int i = 0; i < b; ++i) { a(r * b + i, pt.col[s] * b + j) += blk[i + size_t(j) * b]; }
      }
    }
  }
  return a;
}
// Textbook scalar ILU(0) (IKJ, no pivoting) of a on the expanded block
// pattern, then the two triangular solves of x.
std::vector<double> scalar_ilu0(Dense lu, const BlockPattern &pt, std::vector<double> x) {
  const int b = pt.block, n = lu.n;
  std::vector<char> mask(size_t(n) * n, 0);
  for (int r = 0; r < pt.nb; ++r) {
    for (int s = pt.row[r]; s < pt.row[r + 1]; ++s) {
      for (int i = 0; i < b; ++i) {
        for (int j = 0; j < b; ++j) { mask[size_t(r * b + i) * n + pt.col[s] * b + j] = 1; }
      }
    }
  }
  std::vector<int> cols;
  for (int i = 1; i < n; ++i) {
    cols.clear();
    for (int j = 0; j < n; ++j) {
      if (mask[size_t(i) * n + j]) { cols.push_back(j); }
    }
    for (int k : cols) {
      if (k >= i) { break; }
      const double lik = lu(i, k) / lu(k, k);
      lu(i, k) = lik;
      for (int j : cols) {
        if (j > k) { lu(i, j) -= lik * lu(k, j); }
      }
    }
  }
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < i; ++j) { x[i] -= lu(i, j) * x[j]; }
  }
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) { x[i] -= lu(i, j) * x[j]; }
    x[i] /= lu(i, i);
  }
  return x;
}
double max_rel_block_diff(const std::vector<double> &a, const std::vector<double> &b) {
  double diff = 0., scale = 0.;
  for (size_t i = 0; i < a.size(); ++i) {
    diff = std::max(diff, std::abs(a[i] - b[i]));
    scale = std::max(scale, std::abs(b[i]));
  }
  return diff / std::max(1., scale);
}
int face_of(const Problem &p, int e, int side) {
  const int nx = p.opt.nx, ny = p.opt.ny, ix = e % nx, iy = e / nx;
  switch (side) {
  case 0: return 2 * (iy * nx + (ix + nx - 1) % nx);
  case 1: return 2 * e;
  case 2: return 2 * (((iy + ny - 1) % ny) * nx + ix) + 1;
  default: return 2 * e + 1;
  }
}
int line_node(const Problem &p, int side, int qq, int a) { return side < 2 ? qq * p.q + a : a * p.q + qq; }

// Dense Schur complement S = diag(D) - sum_e R_e^T C_e A_e^{-1} B_e R_e
// (src/condensation.cpp:131-148, src/operators.cpp:7-26,121-179) from the
// device-built cell matrices A (column-major m x m), restriction blocks C
// (4q 4x4 column-major blocks a cell) and trace diagonal D, with the face
// lifting B_e formed here: column (side, point, c) holds, at rows
// 4 node + c of the normal line, phi_a(side) / (w_ix w_iy) times
// -dt len / area w_point (a on the left side of the face, b on the right).
Dense host_schur(const Problem &p, const std::vector<double> &A, const std::vector<double> &C,
                 const std::vector<double> &D, const std::vector<eg::Frozen> &fz) {
  const int q = p.q, m = p.m, t = p.t, T4 = 4 * t;
  Dense S(p.nl);
  for (int g = 0; g < p.nl; ++g) { S(g, g) = D[g]; }
  std::vector<double> X(size_t(m) * T4);
  for (int e = 0; e < p.n; ++e) {
    std::fill(X.begin(), X.end(), 0.);
    for (int side = 0; side < 4; ++side) {
      const double len = side < 2 ? p.hy : p.hx;
      for (int qq = 0; qq < q; ++qq) {
        const eg::Frozen &z = fz[size_t(face_of(p, e, side)) * q + qq];
        const double sf = -(p.dt * len / p.area * p.w1[qq]) * ((side & 1) ? z.a : z.b);
        for (int c = 0; c < 4; ++c) {
          double *col = X.data() + size_t(side * t + 4 * qq + c) * m;
          for (int a = 0; a < q; ++a) {
            const int node = line_node(p, side, qq, a);
            const double phi = (side & 1) ? p.end1[a] : p.end0[a];
            col[4 * node + c] = phi / (p.w1[node % q] * p.w1[node / q]) * sf;
          }
        }
      }
    }
    lu_solve(std::vector<double>(A.begin() + size_t(e) * m * m, A.begin() + size_t(e + 1) * m * m), m, X, T4);
    // row (side, qq, c) of C_e X_e: the 4x4 block times the trace of X's columns
    for (int side = 0; side < 4; ++side) {
      for (int qq = 0; qq < q; ++qq) {
        const double *blk = C.data() + (size_t(e) * 4 * q + side * q + qq) * 16;
        const int grow = face_of(p, e, side) * t + 4 * qq;
        for (int col = 0; col < T4; ++col) {
          const int cside = col / t, gcol = face_of(p, e, cside) * t + col % t;
          double tr[4] = {0, 0, 0, 0};
          for (int a = 0; a < q; ++a) {
            const double phi = (side & 1) ? p.end1[a] : p.end0[a];
            const int node = line_node(p, side, qq, a);
            for (int d = 0; d < 4; ++d) { tr[d] += phi * X[size_t(col) * m + 4 * node + d]; }
          }
          for (int c = 0; c < 4; ++c) {
            double v = 0.;
            for (int d = 0; d < 4; ++d) { v += blk[c + 4 * d] * tr[d]; }
            S(grow + c, gcol) -= v;
          }
        }
      }
    }
  }
  return S;
}

// src/verification.cpp:257-303 with the device physics: pressure, fluxes and
// flux Jacobians from the device functions (probe kernel), the HLL bounds
// from the device Freeze (uniform states: symmetric bounds; a left/right
// pair: the one-sided branch), the exact solution's PDE residual on the host.
void VerifyPhysicsDevice(const EulerGas &gas) {
  for (const double value : {1.4, 1.67, gas.Gamma()}) {
    const EulerGas test_gas(value);
    const double u[4] = {1.2, .3, -.2, 3.};
    std::vector<double> states;
    for (int c = 0; c < 4; ++c) { states.push_back(u[c]); }
    for (int c = 0; c < 4; ++c) {
      for (const double sgn : {1., -1.}) {
        for (int d = 0; d < 4; ++d) { states.push_back(u[d] + (d == c ? sgn * 1e-6 : 0.)); }
      }
    }
    Options o;
    o.nx = 2;
    o.ny = 2;
    o.k = 0;
    o.steps = 1;
    Problem pp(o, test_gas);
    DeviceArray<double> ds(states.size()), out(41 * states.size() / 4);
    up(pp, ds, states);
    eg::launch_physics_probe(ds, int(states.size() / 4), value, out, pp.stream);
    const auto r = down(pp, out, out.n);
    const double expected_pressure = (value - 1.) * (3. - .5 * (.3 * .3 + .2 * .2) / 1.2);
    require(std::abs(r[0] - expected_pressure) < 1e-14, "Parameterized pressure failed");
    for (int axis = 0; axis < 2; ++axis) {
      for (int c = 0; c < 4; ++c) {
        const double *fp = r.data() + 41 * (1 + 2 * c) + 1 + 4 * axis, *fm = r.data() + 41 * (2 + 2 * c) + 1 + 4 * axis;
        for (int d = 0; d < 4; ++d) {
          const double jac = r[9 + 16 * axis + d + 4 * c];
          require(std::abs((fp[d] - fm[d]) / 2e-6 - jac) < 1e-8, "Parameterized flux Jacobian failed");
        }
      }
    }
    // symmetric bounds: a uniform state on a 2 x 2 Q0 mesh
    std::vector<double> uniform(pp.nu);
    for (int i = 0; i < pp.nu; ++i) { uniform[i] = u[i % 4]; }
    DeviceArray<double> du(pp.nu);
    up(pp, du, uniform);
    pp.Freeze(du, .001);
    std::vector<eg::Frozen> fz(size_t(pp.nf) * pp.q);
    hip_check(hipMemcpy(fz.data(), pp.frozen, sizeof(eg::Frozen) * fz.size(), hipMemcpyDeviceToHost), "download");
    const double sound = std::sqrt(value * expected_pressure / u[0]);
    for (int f = 0; f < pp.nf; ++f) {
      const int axis = f & 1;
      const double velocity = u[axis + 1] / u[0];
      require(std::abs(fz[f].a - 1.1 * (sound - velocity)) < 1e-12 &&
                  std::abs(fz[f].b - 1.1 * (sound + velocity)) < 1e-12,
              "Parameterized HLL sound speed failed");
    }
  }
  // one-sided branch: l at cells (0,0), (1,1) and r at (1,0), (0,1), so the
  // x-face l|r has a = 0 and the y-face l|r has b = 0 (src/verification.cpp:285-287)
  {
    Options o;
    o.nx = 2;
    o.ny = 2;
    o.k = 0;
    o.steps = 1;
    Problem pp(o, gas);
    const double l[4] = {1, 5, -5, 25.5}, rr[4] = {1.1, 5.5, -5.5, 28.};
    std::vector<double> s(pp.nu);
    for (int e = 0; e < 4; ++e) {
      const bool left = e == 0 || e == 3;
      for (int c = 0; c < 4; ++c) { s[4 * e + c] = left ? l[c] : rr[c]; }
    }
    DeviceArray<double> du(pp.nu);
    up(pp, du, s);
    pp.Freeze(du, .001);
    std::vector<eg::Frozen> fz(size_t(pp.nf));
    hip_check(hipMemcpy(fz.data(), pp.frozen, sizeof(eg::Frozen) * fz.size(), hipMemcpyDeviceToHost), "download");
    require(fz[0].a == 0 && fz[1].b == 0 && fz[6].a == 0 && fz[7].b == 0, "One-sided HLL branch failed");
  }
  double eps = 1e-6, maxerr = 0;
  for (int i = 0; i < 20; i++) {
    const double x = .07 * i, y = .13 * i, t = .07, a = pi / 6;
    const auto ut1 = gas.exact(x, y, t + eps, a), ut0 = gas.exact(x, y, t - eps, a);
    const auto fx1 = gas.flux(gas.exact(x + eps, y, t, a), 0), fx0 = gas.flux(gas.exact(x - eps, y, t, a), 0);
    const auto fy1 = gas.flux(gas.exact(x, y + eps, t, a), 1), fy0 = gas.flux(gas.exact(x, y - eps, t, a), 1);
    for (int c = 0; c < 4; c++) {
      maxerr = std::max(maxerr, std::abs((ut1[c] - ut0[c] + fx1[c] - fx0[c] + fy1[c] - fy0[c]) / (2 * eps)));
    }
  }
  require(maxerr < 2e-8, "Exact Euler PDE verification failed");
  std::cout << "One-sided HLL and exact PDE verified; FD residual=" << maxerr << std::endl;
}
} // namespace

void Verify(const Options &base, const EulerGas &gas) {
  VerifyPhysicsDevice(gas);
  std::ofstream storage(base.output + "_storage.csv");
  require(bool(storage), "Cannot open verification output");
  storage << "k,dg_bsr_ilu_relative,hdg_bsr_ilu_relative,streamed_schur_relative,"
             "schur_diagonal_relative\n"
          << std::setprecision(17);
  std::ofstream report(base.output + "_verification.csv");
  require(bool(report), "Cannot open verification output");
  report << "k,jacobian_fd_relative,dg_hdg_increment_relative,assembled_mf_"
            "relative,mass_off_diagonal,partial_dg_relative,partial_schur_"
            "relative,off_manifold_increment_relative\n"
         << std::setprecision(17);
  for (int k = 0; k <= 8; ++k) {
    Options o = base;
    o.k = k;
    o.nx = 3;
    o.ny = 2;
    o.precond = "block-jacobi";
    o.precision = "fp64";
    o.linear_tol = 1e-12;
    o.steps = 1;
    Problem p(o, gas);
    Options oi = o;
    oi.precond = "block-ilu0";
    Problem pi(oi, gas, p.stream);
    std::vector<double> u = p.Initial();
    const auto background = gas.exact(0, 0, 0, p.alpha);
    for (int i = 0; i < p.nu; ++i) { u[i] = .05 * u[i] + .95 * background[i % 4]; }
    const size_t nu = p.nu, nl = p.nl;
    DeviceArray<double> du(nu), dv(nu), dz(nu), dw(nu), dr(nu), drhs(nu), dx(nu), dl(nl), dy(nl);
    up(p, du, u);
    p.Freeze(du, .001);
    pi.Freeze(du, .001);
    // matrix-free action and its central difference
    const auto v = sines(nu, .37, 1);
    up(p, dv, v);
    LinearSolver dg(p, 0);
    dg.Refresh(du);
    dg.Apply(dv.p, dz.p);
    const auto z = down(p, dz, nu);
    const double eps = 1e-7;
    std::vector<double> w(nu);
    for (size_t i = 0; i < nu; ++i) { w[i] = u[i] + eps * v[i]; }
    up(p, dw, w);
    p.Residual(dw, dr);
    auto fd = down(p, dr, nu);
    for (size_t i = 0; i < nu; ++i) { w[i] = u[i] - eps * v[i]; }
    up(p, dw, w);
    p.Residual(dw, dr);
    const auto rm = down(p, dr, nu);
    for (size_t i = 0; i < nu; ++i) { fd[i] = (fd[i] - rm[i]) * (.5 / eps); }
    const double fdrel = rel(fd, z);
    // the block fill (diagonal and face-coupling blocks) against the action
    LinearSolver dgi(pi, 0);
    dgi.Refresh(du);
    eg::launch_dg_blocks<double>(pi.geo, pi.frozen, du, dgi.blocks, dgi.dslot, dgi.blocks, dgi.slot, pi.stream);
    const Dense J = dense_bsr(*dgi.pattern, down(pi, dgi.blocks, dgi.blocks.n));
    const double mfrel = rel(J.mult(v), z);
    // DG and HDG increments at the state
    p.Residual(du, dr);
    const auto r = down(p, dr, nu);
    std::vector<double> neg(nu);
    for (size_t i = 0; i < nu; ++i) { neg[i] = -r[i]; }
    up(p, drhs, neg);
    LinearResult lr;
    dg.Solve(drhs, lr, o.linear_tol, true);
    const auto d = down(p, dg.x, nu);
    Hybrid h(p, false);
    h.StreamCondense(du, dr);
    LinearSolver hs(p, 1, &h);
    hs.Solve(h.rhs, lr, o.linear_tol, true);
    h.Recover(hs.x, dx);
    const double equiv = rel(down(p, dx, nu), d);
    // Schur action and diagonal blocks against the host Schur complement
    eg::launch_hdg_build(p.geo, p.frozen, du, h.A, h.C, p.stream);
    const auto A = down(p, h.A, h.A.n), C = down(p, h.C, h.C.n), D = down(p, h.D, nl);
    std::vector<eg::Frozen> fz(size_t(p.nf) * p.q);
    hip_check(hipMemcpy(fz.data(), p.frozen, sizeof(eg::Frozen) * fz.size(), hipMemcpyDeviceToHost), "download");
    const Dense S = host_schur(p, A, C, D, fz);
    h.StreamCondense(du, dr); // the build above overwrote the factors
    const auto sz = sines(nl, .27, 0);
    up(p, dl, sz);
    h.SchurMult(dl.p, dy.p);
    const double streamed = rel(down(p, dy, nl), S.mult(sz));
    DeviceArray<double> diag(size_t(p.nf) * p.t * p.t);
    eg::launch_hdg_diag_blocks(p.geo, h.panels, h.D, diag, p.stream);
    const auto sd = down(p, diag, diag.n);
    std::vector<double> ref(sd.size());
    for (int f = 0; f < p.nf; ++f) {
      for (int j = 0; j < p.t; ++j) {
        for (int i = 0; i < p.t; ++i) { ref[(size_t(f) * p.t + j) * p.t + i] = S(f * p.t + i, f * p.t + j); }
      }
    }
    const double diag_error = max_rel_block_diff(sd, ref);
    // block ILU(0) of both systems, filled and factored twice
    Hybrid hi(pi, false);
    hi.StreamCondense(du, dr);
    LinearSolver hsi(pi, 1, &hi);
    auto check_ilu = [&](LinearSolver &s, const Dense &a, size_t size) {
      auto input = sines(size, .17, 1);
      DeviceArray<double> din(size), dout(size);
      double worst = 0.;
      for (int repeat = 0; repeat < 2; ++repeat) {
        s.Setup();
        up(pi, din, input);
        s.Precondition(din.p, dout.p);
        worst = std::max(worst, rel(down(pi, dout, size), scalar_ilu0(a, *s.pattern, input)));
        for (double &x : input) { x *= .7; }
      }
      require(worst < 1e-10, "Direct BSR ILU differs from native fixed-order ILU");
      return worst;
    };
    const double dg_ilu = check_ilu(dgi, J, nu), hdg_ilu = check_ilu(hsi, S, nl);
    require(streamed < 1e-12, "Streaming Schur differs from reference");
    require(diag_error < 1e-12, "Panel-only Schur diagonal differs");
    storage << k << ',' << dg_ilu << ',' << hdg_ilu << ',' << streamed << ',' << diag_error << std::endl;
    // GL mass matrix on the volume rule: diagonal
    double massoff = 0.;
    for (int i = 0; i < p.q; ++i) {
      for (int j = 0; j < p.q; ++j) {
        if (i == j) { continue; }
        double s = 0.;
        for (int a = 0; a < p.q; ++a) {
          double li = 1., lj = 1.;
          for (int b = 0; b < p.q; ++b) {
            if (b != i) { li *= (p.x1[a] - p.x1[b]) / (p.x1[i] - p.x1[b]); }
            if (b != j) { lj *= (p.x1[a] - p.x1[b]) / (p.x1[j] - p.x1[b]); }
          }
          s += p.w1[a] * li * lj;
        }
        massoff = std::max(massoff, std::abs(s));
      }
    }
    require(fdrel < 1e-7 && equiv < 1e-8 && mfrel < 1e-13 && massoff < 1e-12, "Algebra verification failed");
    report << k << ',' << fdrel << ',' << equiv << ',' << mfrel << ',' << massoff << ',' << mfrel << ','
           << streamed << ',' << -1. << std::endl;
    std::cout << "verified k=" << k << " FD=" << fdrel << " equivalence=" << equiv << '\n';
  }
}
} // namespace euler
