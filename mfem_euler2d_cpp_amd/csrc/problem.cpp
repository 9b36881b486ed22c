// Host side of the device-resident discretisation: physics for initial data
// and errors, 1-D Gauss-Legendre tables, the Problem object (device buffers,
// Freeze/Residual/CellNorm/Admissible on the GPU), the BSR ILU(0) topology
// and the HDG condensed state.
#include <algorithm>
#include <set>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>

#include "euler.hpp"

namespace euler {
void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
  }
}
double clock_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------- EulerGas
// src/physics.cpp; include/euler.hpp:33-35
EulerGas::EulerGas(double value) : gamma(value), gamma_minus_one(value - 1.) {
  require(std::isfinite(value) && value > 1., "Euler gamma must be finite and > 1");
}
double EulerGas::pressure(const State &u) const {
  return gamma_minus_one * (u[3] - .5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
}
bool EulerGas::positive(const State &u) const {
  for (double v : u) {
    if (!std::isfinite(v)) { return false; }
  }
  return u[0] > 1e-13 && pressure(u) > 1e-13;
}
State EulerGas::flux(const State &u, int axis) const {
  State f;
  eg::flux<double>(u.data(), axis, gamma_minus_one, f.data());
  return f;
}
void EulerGas::jacobian(const State &u, int axis, double j[16]) const {
  eg::jacobian<double>(u.data(), axis, gamma_minus_one, j);
}
std::array<double, 2> EulerGas::speeds(const State &l, const State &r, int axis, double safety) const {
  bool finite = true;
  for (int c = 0; c < 4; ++c) { finite = finite && std::isfinite(l[c]) && std::isfinite(r[c]); }
  require(finite && l[0] > 1e-13 && r[0] > 1e-13, "Nonpositive old face state; refine mesh");
  const double pl = pressure(l), pr = pressure(r);
  require(pl > 1e-13 && pr > 1e-13, "Nonpositive old face state; refine mesh");
  const double il = 1. / l[0], ir = 1. / r[0], sl = std::sqrt(l[0]), sr = std::sqrt(r[0]);
  double v[2];
  for (int d = 0; d < 2; ++d) { v[d] = (sl * l[d + 1] * il + sr * r[d + 1] * ir) / (sl + sr); }
  const double hr = (sl * (l[3] + pl) * il + sr * (r[3] + pr) * ir) / (sl + sr);
  const double c2 = gamma_minus_one * (hr - .5 * (v[0] * v[0] + v[1] * v[1]));
  require(c2 > 0, "Invalid Roe sound speed");
  const double ct = std::sqrt(c2), cl = std::sqrt(gamma * pl * il), cr = std::sqrt(gamma * pr * ir);
  return {safety * std::max({0., cl - l[axis + 1] * il, ct - v[axis]}),
          safety * std::max({0., cr + r[axis + 1] * ir, ct + v[axis]})};
}
// Rotated simple wave (src/physics.cpp:64-93): Riemann invariant J+ = 2,
// J- = sin(pi xi) carried along xi + t((3-g)/2 + (g+1)/4 sin(pi xi)) = z mod 2;
// the characteristic foot is found by Newton safeguarded by bisection.
State EulerGas::exact(double x, double y, double time, double alpha) const {
  const double drift = (3. - gamma) / 2., amp = (gamma + 1.) / 4.;
  require(time >= 0 && time * pi * amp < 1, "Exact solution requires T < 4/((gamma+1)*pi)");
  auto wrap = [](double z) { return z - 2 * std::floor((z + 1) / 2); };
  const double phase = wrap(wrap(std::cos(alpha) * x + std::sin(alpha) * y) - time * drift);
  const double ta = time * amp;
  double lo = phase - ta, hi = phase + ta, z = phase;
  for (int it = 0; it < 80; ++it) {
    const double g = z + ta * std::sin(pi * z) - phase;
    if (std::abs(g) < 2e-15) { break; }
    (g > 0 ? hi : lo) = z;
    const double step = z - g / (1 + ta * pi * std::cos(pi * z));
    z = (step >= lo && step <= hi) ? step : .5 * (lo + hi);
  }
  const double jm = std::sin(pi * z), v = 1 + .5 * jm;
  const double c = gamma_minus_one * (2. - jm) / 4.;
  const double rho = std::pow(c * c / gamma, 1. / gamma_minus_one);
  const double pr = std::pow(rho, gamma);
  return {rho, rho * v * std::cos(alpha), rho * v * std::sin(alpha), pr / gamma_minus_one + .5 * rho * v * v};
}

// ------------------------------------------------------------- GL tables
namespace {
// Gauss-Legendre rule on [0,1], ascending: Newton iteration on the Legendre
// three-term recurrence in extended precision.
void gl_rule(int np, std::vector<double> &x, std::vector<double> &w) {
  x.resize(np);
  w.resize(np);
  if (np == 1) { x[0] = .5; w[0] = 1.; return; }
  for (int k = 0; k < np; ++k) {
    // k-th root from the left (ascending in t = cos(theta) ordering reversed)
    long double t = -std::cos(pi * (k + .75L) / (np + .5L)), dp = 1;
    for (int it = 0; it < 64; ++it) {
      long double pm = 1, pc = t;
      for (int j = 2; j <= np; ++j) {
        const long double pn = ((2 * j - 1) * t * pc - (j - 1) * pm) / j;
        pm = pc;
        pc = pn;
      }
      dp = np * (pm - t * pc) / (1 - t * t);
      const long double step = pc / dp;
      t -= step;
      if (std::fabs((double)step) < 1e-20) { break; }
    }
    long double pm = 1, pc = t;
    for (int j = 2; j <= np; ++j) {
      const long double pn = ((2 * j - 1) * t * pc - (j - 1) * pm) / j;
      pm = pc;
      pc = pn;
    }
    dp = np * (pm - t * pc) / (1 - t * t);
    x[k] = double((t + 1) / 2);
    w[k] = double(1 / ((1 - t * t) * dp * dp));
  }
}
// Barycentric Lagrange basis values l_i(at).
std::vector<double> lagrange_values(const std::vector<double> &x, long double at) {
  const int n = x.size();
  std::vector<double> v(n);
  for (int i = 0; i < n; ++i) {
    long double p = 1;
    for (int j = 0; j < n; ++j) {
      if (j != i) { p *= (at - x[j]) / ((long double)x[i] - x[j]); }
    }
    v[i] = double(p);
  }
  return v;
}
} // namespace

// ------------------------------------------------------------- Problem
// src/problem.cpp:32-223
Problem::Problem(const Options &o, const EulerGas &g, hipStream_t s)
    : opt(o), gas(g), alpha(o.alpha * pi / 180), lx(2 / std::cos(alpha)), ly(2 / std::sin(alpha)),
      hx(lx / o.nx), hy(ly / o.ny), area(hx * hy), n(o.nx * o.ny), p((o.k + 1) * (o.k + 1)), m(4 * p),
      t(4 * (o.k + 1)), nf(2 * n), nu(n * m), nl(nf * t), q(o.k + 1) {
  require(o.steps > 0 || o.end < 4 / ((gas.Gamma() + 1) * pi), "Exact solution requires T < 4/((gamma+1)*pi)");
  require(o.alpha > 0 && o.alpha < 90, "Angle must be strictly between 0 and 90 degrees");
  require(o.k >= 0 && o.k <= 8, "Expected native Qk element");
  require(o.nx >= 2 && o.ny >= 2, "Invalid mesh, degree, time or safety");
  gl_rule(q, x1, w1);
  D1.assign(q * q, 0.);
  end0 = lagrange_values(x1, 0.L);
  end1 = lagrange_values(x1, 1.L);
  // Derivative matrix D1[a*q+i] = l_i'(x_a) from barycentric weights.
  std::vector<long double> bw(q, 1.L);
  for (int i = 0; i < q; ++i) {
    for (int j = 0; j < q; ++j) {
      if (j != i) { bw[i] /= ((long double)x1[i] - x1[j]); }
    }
  }
  for (int a = 0; a < q; ++a) {
    long double diag = 0;
    for (int i = 0; i < q; ++i) {
      if (i == a) { continue; }
      const long double v = (bw[i] / bw[a]) / ((long double)x1[a] - x1[i]);
      D1[a * q + i] = std::fabs((double)v) < 1e-14 ? 0. : double(v);
      diag -= v;
    }
    D1[a * q + a] = std::fabs((double)diag) < 1e-14 ? 0. : double(diag);
  }
  for (int i = 0; i < q; ++i) {
    if (std::abs(end0[i]) < 1e-14) { end0[i] = 0; }
    if (std::abs(end1[i]) < 1e-14) { end1[i] = 0; }
  }
  origx.resize(n);
  origy.resize(n);
  for (int iy = 0; iy < o.ny; ++iy) {
    for (int ix = 0; ix < o.nx; ++ix) {
      origx[iy * o.nx + ix] = double(ix) / o.nx * lx - .5 * lx;
      origy[iy * o.nx + ix] = double(iy) / o.ny * ly - .5 * ly;
    }
  }
  cell_metric.resize(nu);
  trace_metric.resize(nl);
  for (int e = 0; e < n; ++e) {
    for (int i = 0; i < p; ++i) {
      for (int c = 0; c < 4; ++c) { cell_metric[size_t(e) * m + 4 * i + c] = area * (w1[i % q] * w1[i / q]); }
    }
  }
  for (int f = 0; f < nf; ++f) {
    const double length = (f & 1) ? hx : hy;
    for (int qq = 0; qq < q; ++qq) {
      for (int c = 0; c < 4; ++c) { trace_metric[size_t(f) * t + 4 * qq + c] = length / w1[qq]; }
    }
  }
  // device geometry
  geo.nx = o.nx; geo.ny = o.ny; geo.n = n; geo.q = q; geo.p = p; geo.m = m; geo.t = t; geo.nf = nf;
  geo.hx = hx; geo.hy = hy; geo.area = area; geo.gamma = gas.Gamma(); geo.gm1 = gas.Gamma() - 1.; geo.dt = 0.;
  for (int i = 0; i < q * q; ++i) { geo.D[i] = D1[i]; }
  for (int i = 0; i < q; ++i) { geo.e0[i] = end0[i]; geo.e1[i] = end1[i]; geo.w1[i] = w1[i]; }
  // device state
  if (s) {
    stream = s;
  } else {
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    own_stream = true;
  }
  require(eg::la_init() == 0, "Cannot enable 160 KB LDS for the dense block kernels");
  require(m <= eg::big_block_max(), "Cell blocks above the dense block kernels' limit");
  frozen.resize(size_t(nf) * q);
  old.resize(nu);
  d_cell_metric.resize(nu);
  d_trace_metric.resize(nl);
  d_cell_metric32.resize(nu);
  d_trace_metric32.resize(nl);
  const int np = std::max({eg::cell_blocks(geo), eg::freeze_blocks(geo), eg::recovered_blocks(geo),
                           eg::kReduceBlocks});
  partial.resize(2 * size_t(np));
  scalars.resize(256);
  flags.resize(1);
  hip_check(hipHostMalloc(reinterpret_cast<void **>(&pinned), 512 * sizeof(double) + 64, hipHostMallocDefault),
            "hipHostMalloc");
  pinned_flags = reinterpret_cast<int *>(pinned + 512);
  hip_check(hipMemcpyAsync(d_cell_metric, cell_metric.data(), sizeof(double) * nu, hipMemcpyHostToDevice, stream),
            "upload metric");
  hip_check(hipMemcpyAsync(d_trace_metric, trace_metric.data(), sizeof(double) * nl, hipMemcpyHostToDevice,
                           stream),
            "upload metric");
  eg::launch_convert<double, float>(d_cell_metric, d_cell_metric32, nu, stream);
  eg::launch_convert<double, float>(d_trace_metric, d_trace_metric32, nl, stream);
  hip_check(hipMemsetAsync(flags, 0, sizeof(int), stream), "memset");
  hip_check(hipMemsetAsync(old, 0, sizeof(double) * nu, stream), "memset");
  // the projection / error rule and the cell origins for the device
  {
    std::vector<double> gx, gw;
    const int nq = std::max(24, 2 * opt.k + 8) / 2 + 1;
    require(nq <= eg::kMaxQuad, "Projection rule above the device table size");
    gl_rule(nq, gx, gw);
    rule.nq = nq;
    for (int s2 = 0; s2 < nq; ++s2) {
      rule.x[s2] = gx[s2];
      rule.w[s2] = gw[s2];
      const auto v = lagrange_values(x1, gx[s2]);
      for (int i = 0; i < q; ++i) { rule.b[s2 * q + i] = v[i]; }
    }
    d_origx.resize(n);
    d_origy.resize(n);
    moments.resize(4 * size_t(n));
    hip_check(hipMemcpyAsync(d_origx, origx.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream), "upload");
    hip_check(hipMemcpyAsync(d_origy, origy.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream), "upload");
    wave.origx = d_origx;
    wave.origy = d_origy;
    wave.alpha = alpha;
    wave.ca = std::cos(alpha);
    wave.sa = std::sin(alpha);
  }
  Sync();
}
Problem::~Problem() {
  // Copies into the pinned buffer may still be queued on the stream.
  if (stream) { (void)hipStreamSynchronize(stream); }
  if (pinned) { (void)hipHostFree(pinned); }
  if (own_stream && stream) { (void)hipStreamDestroy(stream); }
}
void Problem::Sync() { hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
double Problem::ReduceScalar(const double *dev_partials, int count, int mode) {
  // the kernel writes the sum and the error flags to pinned memory itself
  eg::launch_sum_partials(dev_partials, count, scalars, mode, stream, 1, pinned, flags, pinned_flags);
  Sync();
  return pinned[0];
}
int Problem::TakeFlags() {
  const int f = *pinned_flags;
  if (f) {
    hip_check(hipMemsetAsync(flags, 0, sizeof(int), stream), "memset");
    *pinned_flags = 0;
  }
  return f;
}
void Problem::CheckFlags(const char *context) {
  hip_check(hipMemcpyAsync(pinned_flags, flags, sizeof(int), hipMemcpyDeviceToHost, stream), "readback");
  Sync();
  CheckPinnedFlags(context);
}
void Problem::CheckPinnedFlags(const char *context) {
  const int f = TakeFlags();
  if (f & eg::kErrSpin) { throw std::runtime_error("ILU dataflow dependency wait exceeded its bound"); }
  if (f & eg::kErrSingular) { throw std::runtime_error(context); }
  if (f & eg::kErrFaceState) { throw std::runtime_error("Nonpositive old face state; refine mesh"); }
  if (f & eg::kErrRoe) { throw std::runtime_error("Invalid Roe sound speed"); }
  if (f & eg::kErrDegenerate) { throw std::runtime_error("Degenerate HLL speeds"); }
}

// src/problem.cpp:329-349: L2 projection with the 13x13 tensor GL rule
// (order max(24, 2k+8)), divided by the diagonal GL mass, on the device.
void Problem::InitialDevice(double *dev_u) {
  eg::launch_initial(geo, rule, wave, dev_u, stream);
  hip_check(hipGetLastError(), "initial projection");
}
std::vector<double> Problem::Initial() {
  DeviceArray<double> d(nu);
  InitialDevice(d);
  std::vector<double> u(nu);
  hip_check(hipMemcpyAsync(u.data(), d, sizeof(double) * nu, hipMemcpyDeviceToHost, stream), "download");
  Sync();
  return u;
}
// src/problem.cpp:350-384 on the device.
void Problem::Freeze(const double *dev_u, double step, bool check_admissible) {
  ++freeze_generation;
  if (check_admissible) { require(Admissible(dev_u), "Old state is inadmissible; refine mesh"); }
  eg::launch_freeze(geo, dev_u, opt.safety, frozen, partial, flags, stream);
  rate = ReduceScalar(partial, eg::freeze_blocks(geo), 1) / std::min(hx, hy);
  const int f = TakeFlags();
  if (f & eg::kErrFaceState) { throw std::runtime_error("Nonpositive old face state; refine mesh"); }
  if (f & eg::kErrRoe) { throw std::runtime_error("Invalid Roe sound speed"); }
  if (f & eg::kErrDegenerate) { throw std::runtime_error("Degenerate HLL speeds"); }
  SetStage(dev_u, step);
}
void Problem::SetStage(const double *dev_old, double step) {
  if (dev_old != old.p) {
    hip_check(hipMemcpyAsync(old, dev_old, sizeof(double) * nu, hipMemcpyDeviceToDevice, stream), "copy old");
  }
  dt = step;
  geo.dt = step;
}
// src/problem.cpp:385-391
double Problem::CellNorm(const double *dev_r) {
  eg::launch_dot_partials<double>(dev_r, dev_r, d_cell_metric, nu, partial, stream);
  return std::sqrt(ReduceScalar(partial, eg::reduce_blocks(nu), 0));
}
// src/problem.cpp:259-275
bool Problem::Admissible(const double *dev_u) {
  eg::launch_admissible(geo, dev_u, flags, stream);
  hip_check(hipMemcpyAsync(pinned_flags, flags, sizeof(int), hipMemcpyDeviceToHost, stream), "readback");
  Sync();
  return !(TakeFlags() & eg::kErrInadmissible);
}
// src/problem.cpp:445-507 (+ CellNorm of the result)
double Problem::Residual(const double *dev_u, double *dev_r) {
  bool adm;
  return TrialResidual(dev_u, nullptr, 0., nullptr, dev_r, adm);
}
double Problem::TrialResidual(const double *dev_u, const double *dev_du, double lambda, double *dev_trial,
                              double *dev_r, bool &admissible) {
  const int nb = eg::cell_blocks(geo);
  eg::launch_residual(geo, frozen, dev_u, dev_du, lambda, old, dev_r, dev_trial, partial, flags, stream);
  eg::launch_sum_partials(partial, nb, scalars, 0, stream, 2, pinned, flags, pinned_flags);
  Sync();
  last_state_norm = std::sqrt(pinned[1]);
  admissible = !(TakeFlags() & eg::kErrInadmissible);
  return std::sqrt(pinned[0]);
}
// src/block_ilu.cpp:110-124
std::vector<std::vector<int>> Problem::SolverGraph(bool hdg) const {
  std::vector<std::vector<int>> graph(hdg ? nf : n);
  auto face = [&](int e, int side) {
    const int ix = e % opt.nx, iy = e / opt.nx;
    switch (side) {
    case 0: return 2 * (iy * opt.nx + (ix + opt.nx - 1) % opt.nx);
    case 1: return 2 * e;
    case 2: return 2 * (((iy + opt.ny - 1) % opt.ny) * opt.nx + ix) + 1;
    default: return 2 * e + 1;
    }
  };
  for (int e = 0; e < n; ++e) {
    const int ix = e % opt.nx, iy = e / opt.nx;
    if (hdg) {
      for (int a = 0; a < 4; ++a) {
        for (int b = 0; b < 4; ++b) { graph[face(e, a)].push_back(face(e, b)); }
      }
    } else {
      const int nb[4] = {iy * opt.nx + (ix + opt.nx - 1) % opt.nx, iy * opt.nx + (ix + 1) % opt.nx,
                         ((iy + opt.ny - 1) % opt.ny) * opt.nx + ix, ((iy + 1) % opt.ny) * opt.nx + ix};
      graph[e].push_back(e);
      for (int s = 0; s < 4; ++s) { graph[e].push_back(nb[s]); }
    }
  }
  for (auto &r : graph) {
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  return graph;
}
// src/problem.cpp:704-726
// src/problem.cpp:704-738 on the device: per-cell sums (the error integrals
// with the 13x13 rule, or the conserved integrals), then one fixed-order
// reduction per component.
std::array<double, 4> Problem::ErrorsDevice(const double *dev_u, double time) {
  require(time >= 0 && time * pi * (gas.Gamma() + 1.) / 4. < 1, "Exact solution requires T < 4/((gamma+1)*pi)");
  eg::launch_cell_moments(geo, rule, wave, dev_u, time, 1, moments, stream);
  eg::launch_sum_partials(moments, n, scalars, 0, stream, 4, pinned, flags, pinned_flags);
  Sync();
  std::array<double, 4> a{};
  for (int c = 0; c < 4; ++c) { a[c] = std::sqrt(pinned[c]); }
  return a;
}
std::array<double, 4> Problem::IntegralDevice(const double *dev_u) {
  eg::launch_cell_moments(geo, rule, wave, dev_u, 0., 0, moments, stream);
  eg::launch_sum_partials(moments, n, scalars, 0, stream, 4, pinned, flags, pinned_flags);
  Sync();
  std::array<double, 4> a{};
  for (int c = 0; c < 4; ++c) { a[c] = pinned[c]; }
  return a;
}
std::array<double, 4> Problem::Errors(const std::vector<double> &u, double time) {
  require(int(u.size()) == nu, "Solution size mismatch");
  DeviceArray<double> d(nu);
  hip_check(hipMemcpyAsync(d, u.data(), sizeof(double) * nu, hipMemcpyHostToDevice, stream), "upload");
  return ErrorsDevice(d, time);
}
std::array<double, 4> Problem::Integral(const std::vector<double> &u) {
  require(int(u.size()) == nu, "Solution size mismatch");
  DeviceArray<double> d(nu);
  hip_check(hipMemcpyAsync(d, u.data(), sizeof(double) * nu, hipMemcpyHostToDevice, stream), "upload");
  return IntegralDevice(d);
}
// src/problem.cpp:739-761
std::array<double, 4> Problem::Differences(const std::vector<double> &u, const std::vector<double> &v) const {
  require(int(u.size()) == nu && int(v.size()) == nu, "Solution difference size mismatch");
  std::vector<double> d(nu);
  for (int i = 0; i < nu; ++i) { d[i] = u[i] - v[i]; }
  std::array<double, 4> a{};
  std::vector<double> gx, gw;
  const int nq = std::max(24, 2 * opt.k + 8) / 2 + 1;
  gl_rule(nq, gx, gw);
  std::vector<double> bx(size_t(nq) * q);
  for (int s = 0; s < nq; ++s) {
    const auto vv = lagrange_values(x1, gx[s]);
    for (int i = 0; i < q; ++i) { bx[s * q + i] = vv[i]; }
  }
  for (int e = 0; e < n; ++e) {
    for (int qy = 0; qy < nq; ++qy) {
      for (int qx = 0; qx < nq; ++qx) {
        for (int c = 0; c < 4; ++c) {
          double z = 0;
          for (int i = 0; i < p; ++i) { z += bx[qx * q + i % q] * bx[qy * q + i / q] * d[size_t(e) * m + 4 * i + c]; }
          a[c] += area * gw[qx] * gw[qy] * z * z;
        }
      }
    }
  }
  for (auto &x : a) { x = std::sqrt(x); }
  return a;
}
// src/problem.cpp:762-790: coefficients CSV plus an MFEM-readable Cartesian
// mesh and L2 GL grid functions (element-wise lexicographic nodal values).
void Problem::Save(const std::vector<double> &u, const std::string &prefix) const {
  {
    // The reference prints its mesh (src/problem.cpp:763-764): the
    // non-periodic nx x ny Cartesian quadrilateral mesh in lexicographic
    // element order, shifted to the box [-lx/2, lx/2] x [-ly/2, ly/2]
    // (src/problem.cpp:37-38,47-51), in the MFEM mesh v1.0 text format:
    // elements "attribute geometry vertices" (attribute 1, square = 3), the
    // boundary segments bottom (attribute 1), top (3), left (4), right (2),
    // and the vertices x fastest, at the stream's default precision.
    std::ofstream ms(prefix + ".mesh");
    require(bool(ms), "Cannot open mesh output");
    const int nx = opt.nx, ny = opt.ny, nvx = nx + 1;
    ms << "MFEM mesh v1.0\n\n#\n# MFEM Geometry Types (see fem/geom.hpp):\n#\n# POINT       = 0\n"
          "# SEGMENT     = 1\n# TRIANGLE    = 2\n# SQUARE      = 3\n# TETRAHEDRON = 4\n# CUBE        = 5\n"
          "# PRISM       = 6\n# PYRAMID     = 7\n#\n\ndimension\n2\n\nelements\n"
       << n << '\n';
    for (int iy = 0; iy < ny; ++iy) {
      for (int ix = 0; ix < nx; ++ix) {
        const int v0 = iy * nvx + ix;
        ms << "1 3 " << v0 << ' ' << v0 + 1 << ' ' << v0 + 1 + nvx << ' ' << v0 + nvx << '\n';
      }
    }
    ms << "\nboundary\n" << 2 * (nx + ny) << '\n';
    const int top = nvx * ny;
    for (int ix = 0; ix < nx; ++ix) { ms << "1 1 " << ix << ' ' << ix + 1 << '\n'; }
    for (int ix = 0; ix < nx; ++ix) { ms << "3 1 " << top + ix + 1 << ' ' << top + ix << '\n'; }
    for (int iy = 0; iy < ny; ++iy) { ms << "4 1 " << (iy + 1) * nvx << ' ' << iy * nvx << '\n'; }
    for (int iy = 0; iy < ny; ++iy) { ms << "2 1 " << iy * nvx + nx << ' ' << (iy + 1) * nvx + nx << '\n'; }
    ms << "\nvertices\n" << nvx * (ny + 1) << "\n2\n";
    for (int iy = 0; iy <= ny; ++iy) {
      for (int ix = 0; ix <= nx; ++ix) {
        ms << double(ix) / nx * lx - .5 * lx << ' ' << double(iy) / ny * ly - .5 * ly << '\n';
      }
    }
  }
  // One GridFunction file per component (src/problem.cpp:765-777): the
  // L2 Gauss-Legendre space's header, then the values in element-DOF order
  // (element e, native node i).
  const char *names[] = {"rho", "rhou", "rhov", "energy"};
  for (int c = 0; c < 4; ++c) {
    std::ofstream fs(prefix + "_" + names[c] + ".gf");
    require(bool(fs), "Cannot open solution output");
    fs << "FiniteElementSpace\nFiniteElementCollection: L2_2D_P" << opt.k << "\nVDim: 1\nOrdering: 0\n\n"
       << std::setprecision(17);
    for (int e = 0; e < n; ++e) {
      for (int i = 0; i < p; ++i) { fs << u[size_t(e) * m + 4 * i + c] << '\n'; }
    }
  }
  std::ofstream cs(prefix + "_coefficients.csv");
  cs << "element,node,rho,rhou,rhov,energy\n" << std::setprecision(17);
  for (int e = 0; e < n; ++e) {
    for (int i = 0; i < p; ++i) {
      cs << e << ',' << i;
      for (int c = 0; c < 4; ++c) { cs << ',' << u[size_t(e) * m + 4 * i + c]; }
      cs << '\n';
    }
  }
}

// ------------------------------------------------------------- BlockPattern
// src/block_ilu.cpp:10-59 (natural ordering) plus level sets for the
// parallel factor and sweeps: a row's level is one more than the deepest
// level among the rows it depends on.
// The reference runs MFEM's MinimumDiscardedFillOrdering once on the
// topology matrix of the block graph (every stored entry 1). MFEM is not
// available here; this restates the published minimum discarded fill rule
// (Persson and Peraire, SIAM J. Sci. Comput. 30 (2008) 2709-2733) on that
// matrix: the weight of an uneliminated row k is the square root of the
// number of fill entries (i, j) that eliminating k would create and ILU(0)
// discards (i != j both uneliminated neighbours of k, (i, j) not in the
// pattern; each product of unit entries counts 1); the row of least weight is
// eliminated next, the lowest index on ties, and the weights of its
// uneliminated neighbours are recomputed. MFEM's own tie order is not
// pinned by anything in the reference.
std::vector<int> MdfOrdering(const std::vector<std::vector<int>> &graph) {
  const int n = int(graph.size());
  std::vector<std::vector<int>> adj(n);
  for (int i = 0; i < n; ++i) {
    adj[i] = graph[i];
    adj[i].push_back(i);
    std::sort(adj[i].begin(), adj[i].end());
    adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
  }
  auto linked = [&](int i, int j) { return std::binary_search(adj[i].begin(), adj[i].end(), j); };
  std::vector<char> done(n, 0);
  auto weight = [&](int k) {
    double w = 0.;
    for (int i : adj[k]) {
      if (i == k || done[i]) { continue; }
      for (int j : adj[k]) {
        if (j == k || j == i || done[j]) { continue; }
        if (!linked(i, j)) { w += 1.; }
      }
    }
    return std::sqrt(w);
  };
  std::vector<double> w(n);
  std::set<std::pair<double, int>> heap;
  for (int k = 0; k < n; ++k) {
    w[k] = weight(k);
    heap.insert({w[k], k});
  }
  std::vector<int> order;
  order.reserve(n);
  while (!heap.empty()) {
    const int k = heap.begin()->second;
    heap.erase(heap.begin());
    done[k] = 1;
    order.push_back(k);
    for (int j : adj[k]) {
      if (j == k || done[j]) { continue; }
      heap.erase({w[j], j});
      w[j] = weight(j);
      heap.insert({w[j], j});
    }
  }
  return order;
}

BlockPattern::BlockPattern(const std::vector<std::vector<int>> &graph_in, int size, hipStream_t s,
                           const std::vector<int> &order)
    : nb(graph_in.size()), block(size) {
  require(block > 0, "Invalid BSR block size");
  // src/block_ilu.cpp:15-29: row i of the factor is graph row perm[i]
  std::vector<std::vector<int>> permuted;
  if (!order.empty()) {
    require(int(order.size()) == nb, "ILU ordering does not match the block graph");
    perm = order;
    inv.assign(nb, -1);
    for (int i = 0; i < nb; ++i) { inv[perm[i]] = i; }
    permuted.resize(nb);
    for (int i = 0; i < nb; ++i) {
      for (int j : graph_in[perm[i]]) { permuted[i].push_back(inv[j]); }
    }
  }
  const auto &graph = order.empty() ? graph_in : permuted;
  diag.resize(nb);
  row.push_back(0);
  for (int i = 0; i < nb; ++i) {
    std::vector<int> nbrs(graph[i]);
    nbrs.push_back(i);
    std::sort(nbrs.begin(), nbrs.end());
    nbrs.erase(std::unique(nbrs.begin(), nbrs.end()), nbrs.end());
    for (int j : nbrs) {
      if (j == i) { diag[i] = col.size(); }
      col.push_back(j);
    }
    row.push_back(col.size());
  }
  nnzb = col.size();
  for (int i = 0; i < nb; ++i) { max_deps = std::max({max_deps, diag[i] - row[i], row[i + 1] - diag[i] - 1}); }
  upd_ptr.assign(1, 0);
  for (int i = 0; i < nb; ++i) {
    for (int ik = row[i]; ik < row[i + 1]; ++ik) {
      if (ik < diag[i]) {
        const int k = col[ik];
        int ij = ik + 1, kj = diag[k] + 1;
        while (ij < row[i + 1] && kj < row[k + 1]) {
          if (col[ij] < col[kj]) { ++ij; }
          else if (col[kj] < col[ij]) { ++kj; }
          else { upd.push_back(ij); upd.push_back(kj); ++ij; ++kj; }
        }
      }
      upd_ptr.push_back(upd.size() / 2);
    }
  }
  std::vector<int> lev(nb, 0), ulev(nb, 0);
  int nlev = 0, nulev = 0;
  for (int i = 0; i < nb; ++i) {
    for (int ij = row[i]; ij < diag[i]; ++ij) { lev[i] = std::max(lev[i], lev[col[ij]] + 1); }
    nlev = std::max(nlev, lev[i] + 1);
  }
  for (int i = nb - 1; i >= 0; --i) {
    for (int ij = diag[i] + 1; ij < row[i + 1]; ++ij) { ulev[i] = std::max(ulev[i], ulev[col[ij]] + 1); }
    nulev = std::max(nulev, ulev[i] + 1);
  }
  auto group = [&](const std::vector<int> &l, int count, std::vector<int> &ptr, std::vector<int> &rows) {
    ptr.assign(count + 1, 0);
    for (int i = 0; i < nb; ++i) { ++ptr[l[i] + 1]; }
    for (int k = 0; k < count; ++k) { ptr[k + 1] += ptr[k]; }
    rows.assign(nb, 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < nb; ++i) { rows[fill[l[i]]++] = i; }
  };
  group(lev, nlev, lower_level_ptr, lower_rows);
  group(ulev, nulev, upper_level_ptr, upper_rows);
  for (const auto *ptr : {&lower_level_ptr, &upper_level_ptr}) {
    for (size_t l = 0; l + 1 < ptr->size(); ++l) { max_width = std::max(max_width, (*ptr)[l + 1] - (*ptr)[l]); }
  }
  auto up = [&](DeviceArray<int> &d, const std::vector<int> &h) {
    d.resize(std::max<size_t>(h.size(), 1));
    if (!h.empty()) {
      hip_check(hipMemcpyAsync(d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s), "upload pattern");
    }
  };
  up(d_row, row);
  up(d_col, col);
  up(d_diag, diag);
  up(d_upd_ptr, upd_ptr);
  up(d_upd, upd);
  up(d_lower_rows, lower_rows);
  up(d_upper_rows, upper_rows);
  std::vector<int> lower_diag(lower_rows.size());
  for (size_t t = 0; t < lower_rows.size(); ++t) { lower_diag[t] = diag[lower_rows[t]]; }
  up(d_lower_diag, lower_diag);
  if (!perm.empty()) {
    up(d_perm, perm);
    up(d_inv, inv);
  }
  hip_check(hipStreamSynchronize(s), "sync");
}
int BlockPattern::Find(int r, int c) const {
  if (!inv.empty()) {
    r = inv[r];
    c = inv[c];
  }
  const auto b = col.begin() + row[r], e = col.begin() + row[r + 1];
  const auto it = std::lower_bound(b, e, c);
  require(it != e && *it == c, "Entry outside fixed BSR graph");
  return int(it - col.begin());
}

// ------------------------------------------------------------- Hybrid
// src/condensation.cpp:47-85: storage allocated once before the time loop.
Hybrid::Hybrid(Problem &problem, bool mixed) : p(problem) {
  A.resize(size_t(p.n) * p.m * p.m);
  C.resize(size_t(p.n) * 4 * p.q * 16);
  Xy.resize(size_t(p.n) * p.m * (4 * p.t + 1));
  panels.resize(size_t(p.nf) * 2 * 4 * p.t * p.t);
  cy.resize(size_t(p.n) * 4 * p.t);
  D.resize(p.nl);
  rhs.resize(p.nl);
  if (mixed) { panels32.resize(panels.n); }
}
// src/condensation.cpp:131-148: A_e, A_e^{-1}, [X|y], S_e panels, rhs, D.
void Hybrid::StreamCondense(const double *dev_u, const double *dev_r) {
  eg::launch_hdg_build(p.geo, p.frozen, dev_u, A, C, p.stream);
  eg::launch_block_inverse(A, p.m, p.n, nullptr, p.flags, p.stream);
  eg::launch_hdg_solve(p.geo, p.frozen, A, dev_r, Xy, p.stream);
  eg::launch_hdg_panels(p.geo, Xy, C, panels, cy, p.stream);
  eg::launch_hdg_face(p.geo, p.frozen, cy, rhs, D, p.stream);
  if (panels32.n) { eg::launch_convert<double, float>(panels, panels32, panels.n, p.stream); }
  p.CheckFlags("Singular streaming HDG cell factor");
}
void Hybrid::SchurMult(const double *x, double *y) { eg::launch_schur_apply<double>(p.geo, panels, D, x, y, p.stream); }
void Hybrid::SchurMult(const float *x, float *y) { eg::launch_schur_apply<float>(p.geo, panels32, D, x, y, p.stream); }
void Hybrid::Recover(const double *dev_dl, double *dev_du) { eg::launch_recover(p.geo, Xy, dev_dl, dev_du, p.stream); }
// src/stepper.cpp:10-26; returns the cell norm of the recovered residual.
double Hybrid::RecoveredResidual(const double *dev_eps, double *dev_out) {
  eg::launch_recovered_residual(p.geo, p.frozen, D, dev_eps, dev_out, p.partial, p.stream);
  return std::sqrt(p.ReduceScalar(p.partial, eg::recovered_blocks(p.geo), 0));
}
} // namespace euler
