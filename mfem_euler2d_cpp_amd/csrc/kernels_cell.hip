// Cell- and face-centric kernels of the implicit DG/HDG Newton update:
// frozen HLL bounds, DG residual (fused with the trial state, admissibility
// and the cell-norm partial sums), matrix-free DG Jacobian action, local
// Jacobian block fills, HDG cell condensation pieces, Schur action, recovery.
//
// Every kernel is atomic-free for its numerical outputs: each face value is
// computed identically by both neighbouring cells (or by one face thread),
// so results do not depend on scheduling.
#include "kernels.hpp"

namespace eg {
namespace {
__device__ __forceinline__ double length_of(const Geo &g, int side) { return side < 2 ? g.hy : g.hx; }

// src/problem.cpp:350-376 (Freeze) and :377-384 (WaveRate partial max)
template <int Q>
__global__ __launch_bounds__(256) void k_freeze(Geo g, const double *__restrict__ u, double safety,
                                                Frozen *__restrict__ fz, double *__restrict__ partial,
                                                int *__restrict__ flags) {
  __shared__ double red[16];
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  double best = 0.;
  if (id < g.nf * Q) {
    const int f = id / Q, qq = id % Q, e = f / 2, axis = f & 1;
    const int ls = axis ? 3 : 1, rs = axis ? 2 : 0;
    const int r = neighbour_of(g, e, ls);
    double l4[4], r4[4];
    trace<Q, double>(g, u + size_t(e) * g.m, ls, qq, l4);
    trace<Q, double>(g, u + size_t(r) * g.m, rs, qq, r4);
    bool finite = true;
    for (int c = 0; c < 4; ++c) { finite = finite && isfinite(l4[c]) && isfinite(r4[c]); }
    int bad = 0;
    if (!(finite && l4[0] > 1e-13 && r4[0] > 1e-13)) { bad |= kErrFaceState; }
    const double pl = pressure(l4, g.gm1), pr = pressure(r4, g.gm1);
    if (!(pl > 1e-13 && pr > 1e-13)) { bad |= kErrFaceState; }
    const double il = 1. / l4[0], ir = 1. / r4[0];
    const double sl = sqrt(l4[0]), sr = sqrt(r4[0]);
    double v[2];
    for (int d = 0; d < 2; ++d) { v[d] = (sl * l4[d + 1] * il + sr * r4[d + 1] * ir) / (sl + sr); }
    const double hroe = (sl * (l4[3] + pl) * il + sr * (r4[3] + pr) * ir) / (sl + sr);
    const double c2 = g.gm1 * (hroe - .5 * (v[0] * v[0] + v[1] * v[1]));
    if (!(c2 > 0)) { bad |= kErrRoe; }
    const double ct = sqrt(c2), cl = sqrt(g.gamma * pl * il), cr = sqrt(g.gamma * pr * ir);
    const double a = safety * fmax(fmax(0., cl - l4[axis + 1] * il), ct - v[axis]);
    const double b = safety * fmax(fmax(0., cr + r4[axis + 1] * ir), ct + v[axis]);
    if (!(a + b > 0)) { bad |= kErrDegenerate; }
    Frozen o;
    o.a = a;
    o.b = b;
    o.inv_sum = 1. / (a + b);
    o.wl = b * o.inv_sum;
    o.wr = a * o.inv_sum;
    o.pen = a * b * o.inv_sum;
    o.pad0 = o.pad1 = 0.;
    fz[id] = o;
    best = fmax(a, b);
    if (bad) { atomicOr(flags, bad); }
  }
  // block max (fixed tree)
  for (int off = 32; off > 0; off >>= 1) { best = fmax(best, __shfl_down(best, off, 64)); }
  if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6] = best; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.;
    for (int i = 0; i < int(blockDim.x >> 6); ++i) { m = fmax(m, red[i]); }
    partial[blockIdx.x] = m;
  }
}

// Rotated simple wave (src/physics.cpp:64-93) at (x, y, time): Riemann
// invariant J+ = 2, J- = sin(pi xi) carried along xi + t((3-g)/2 + (g+1)/4
// sin(pi xi)) = z mod 2; the characteristic foot by Newton safeguarded by
// bisection (at most 80 steps, |residual| < 2e-15), as the host EulerGas.
__device__ void exact_state(double x, double y, double time, const WaveGeo &w, double gamma, double out[4]) {
  constexpr double pi = 3.141592653589793238462643383279502884;
  const double gm1 = gamma - 1., drift = (3. - gamma) / 2., amp = (gamma + 1.) / 4.;
  auto wrap = [](double z) { return z - 2 * floor((z + 1) / 2); };
  const double phase = wrap(wrap(w.ca * x + w.sa * y) - time * drift);
  const double ta = time * amp;
  double lo = phase - ta, hi = phase + ta, z = phase;
  for (int it = 0; it < 80; ++it) {
    const double g = z + ta * sin(pi * z) - phase;
    if (fabs(g) < 2e-15) { break; }
    if (g > 0) { hi = z; } else { lo = z; }
    const double step = z - g / (1 + ta * pi * cos(pi * z));
    z = (step >= lo && step <= hi) ? step : .5 * (lo + hi);
  }
  const double jm = sin(pi * z), v = 1 + .5 * jm;
  const double c = gm1 * (2. - jm) / 4.;
  const double rho = pow(c * c / gamma, 1. / gm1);
  const double pr = pow(rho, gamma);
  out[0] = rho;
  out[1] = rho * v * w.ca;
  out[2] = rho * v * w.sa;
  out[3] = pr / gm1 + .5 * rho * v * v;
}

// Initial (src/problem.cpp:329-349): L2 projection of the exact wave at
// time 0 with the tensor rule r, divided by the diagonal GL mass. One
// workgroup a cell: the exact state at every rule point into LDS, then one
// thread a (node, component) sums over the points (rows, then columns).
template <int Q>
__global__ __launch_bounds__(256) void k_initial(Geo g, QuadRule r, WaveGeo w, double *__restrict__ u) {
  constexpr int P = Q * Q;
  __shared__ double vals[kMaxQuad * kMaxQuad * 4];
  const int e = blockIdx.x, nq = r.nq;
  for (int s = threadIdx.x; s < nq * nq; s += blockDim.x) {
    const int qx = s % nq, qy = s / nq;
    exact_state(w.origx[e] + g.hx * r.x[qx], w.origy[e] + g.hy * r.x[qy], 0., w, g.gamma, vals + 4 * s);
  }
  __syncthreads();
  for (int id = threadIdx.x; id < 4 * P; id += blockDim.x) {
    const int i = id / 4, c = id % 4;
    double acc = 0.;
    for (int qy = 0; qy < nq; ++qy) {
      for (int qx = 0; qx < nq; ++qx) {
        const double wgt = r.w[qx] * r.w[qy];
        const double sh = r.b[qx * Q + i % Q] * r.b[qy * Q + i / Q];
        acc += wgt * sh * vals[4 * (qy * nq + qx) + c];
      }
    }
    u[size_t(e) * 4 * P + id] = acc * (1. / (g.w1[i % Q] * g.w1[i / Q]));
  }
}

// Per-cell moments of a body vector, partials[c * n + e] for component c:
// mode 0: the conserved integral sum_i area w_ix w_iy u (Integral,
// src/problem.cpp:727-738); mode 1: the squared L2 error against the exact
// wave at `time` with the tensor rule r (Errors, src/problem.cpp:704-726).
template <int Q>
__global__ __launch_bounds__(256) void k_cell_moments(Geo g, QuadRule r, WaveGeo w, const double *__restrict__ u,
                                                      double time, int mode, double *__restrict__ partials) {
  constexpr int P = Q * Q;
  __shared__ double red[16 * 4];
  const int e = blockIdx.x;
  const double *cell = u + size_t(e) * 4 * P;
  double acc[4] = {0., 0., 0., 0.};
  if (mode == 0) {
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
      const double wt = g.area * (g.w1[i % Q] * g.w1[i / Q]);
      for (int c = 0; c < 4; ++c) { acc[c] += wt * cell[4 * i + c]; }
    }
  } else {
    const int nq = r.nq;
    for (int s = threadIdx.x; s < nq * nq; s += blockDim.x) {
      const int qx = s % nq, qy = s / nq;
      double v[4];
      exact_state(w.origx[e] + g.hx * r.x[qx], w.origy[e] + g.hy * r.x[qy], time, w, g.gamma, v);
      double z[4] = {0., 0., 0., 0.};
      for (int i = 0; i < P; ++i) {
        const double sh = r.b[qx * Q + i % Q] * r.b[qy * Q + i / Q];
        for (int c = 0; c < 4; ++c) { z[c] += sh * cell[4 * i + c]; }
      }
      for (int c = 0; c < 4; ++c) { acc[c] += g.area * r.w[qx] * r.w[qy] * (z[c] - v[c]) * (z[c] - v[c]); }
    }
  }
  const double s4 = block_sum_k<4>(acc, red);
  if (threadIdx.x < 4) { partials[size_t(threadIdx.x) * g.n + e] = s4; }
}

// Device physics at given states: pressure, the x and y fluxes and their
// Jacobians (column-major), 41 values a state (the --verify known answers).
__global__ __launch_bounds__(64) void k_physics_probe(const double *__restrict__ states, int count, double gamma,
                                                      double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) { return; }
  double u[4];
  for (int c = 0; c < 4; ++c) { u[c] = states[4 * i + c]; }
  const double gm1 = gamma - 1.;
  double *o = out + size_t(i) * 41;
  o[0] = pressure(u, gm1);
  flux(u, 0, gm1, o + 1);
  flux(u, 1, gm1, o + 5);
  jacobian(u, 0, gm1, o + 9);
  jacobian(u, 1, gm1, o + 25);
}

// State value at a node: u + lambda*du (du may be null).
__device__ __forceinline__ void load_state(const double *u, const double *du, double lambda, size_t off,
                                           double s[4]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) { s[c] = u[off + c]; }
  if (du) {
#pragma unroll
    for (int c = 0; c < 4; ++c) { s[c] = s[c] + lambda * du[off + c]; }
  }
}
template <int Q>
__device__ __forceinline__ void trace_state(const Geo &g, const double *u, const double *du, double lambda,
                                            int e, int side, int qq, double out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.;
#pragma unroll
  for (int a = 0; a < Q; ++a) {
    const double phi = side_value<Q, double>(g, side, a);
    double s[4];
    load_state(u, du, lambda, size_t(e) * g.m + 4 * line_node<Q>(side, qq, a), s);
#pragma unroll
    for (int c = 0; c < 4; ++c) { out[c] += phi * s[c]; }
  }
}

// src/problem.cpp:445-507 with :259-275 (Admissible) and :385-391 (CellNorm)
// fused: state = u + lambda*du, r = state - old + M^{-1}[volume + faces],
// admissibility of the state, per-block partial sums of metric*r^2.
template <int Q>
__global__ __launch_bounds__(256) void k_residual(Geo g, const Frozen *__restrict__ fz,
                                                  const double *__restrict__ u, const double *__restrict__ du,
                                                  double lambda, const double *__restrict__ old,
                                                  double *__restrict__ r, double *__restrict__ trial,
                                                  double *__restrict__ partial, int *__restrict__ flags) {
  constexpr int P = Q * Q, CPB = kBlock / P, NF = 4 * Q;
  __shared__ double fx_s[CPB * P * 4], fy_s[CPB * P * 4], h_s[CPB * NF * 4], red[16];
  const int tid = threadIdx.x, lc = tid / P, node = tid % P;
  const int e = blockIdx.x * CPB + lc;
  const bool active = lc < CPB && e < g.n;
  const int ix = node % Q, iy = node / Q;
  bool ok = true;
  double s[4] = {0, 0, 0, 0};
  const size_t off = size_t(e) * g.m + 4 * node;
  if (active) {
    load_state(u, du, lambda, off, s);
    ok = positive(s, g.gm1);
    double fx[4], fy[4];
    flux(s, 0, g.gm1, fx);
    flux(s, 1, g.gm1, fy);
    const double wdt = g.dt * (g.w1[ix] * g.w1[iy]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      fx_s[(lc * P + node) * 4 + c] = wdt * fx[c];
      fy_s[(lc * P + node) * 4 + c] = wdt * fy[c];
    }
  }
  if (lc < CPB && e < g.n) {
    for (int fp = node; fp < NF; fp += P) {
      const int side = fp / Q, qq = fp % Q, axis = side >> 1, left = side & 1;
      const int f = face_of(g, e, side), nb = neighbour_of(g, e, side);
      double own[4], oth[4];
      trace_state<Q>(g, u, du, lambda, e, side, qq, own);
      trace_state<Q>(g, u, du, lambda, nb, side ^ 1, qq, oth);
      ok = ok && positive(own, g.gm1);
      const double *ul = left ? own : oth, *ur = left ? oth : own;
      double fl[4], fr[4];
      flux(ul, axis, g.gm1, fl);
      flux(ur, axis, g.gm1, fr);
      const Frozen z = fz[f * Q + qq];
      const double scale = g.dt * length_of(g, side) / g.area * g.w1[qq];
      const double sign = left ? scale : -scale;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double hh = z.wl * fl[c] + z.wr * fr[c] + z.pen * (ul[c] - ur[c]);
        h_s[(lc * NF + fp) * 4 + c] = sign * hh;
      }
    }
  }
  __syncthreads();
  double contrib = 0., snorm = 0.;
  if (active) {
    double acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < Q; ++a) {
      const double dx = g.D[a * Q + ix] / g.hx, dy = g.D[a * Q + iy] / g.hy;
      const double *px = fx_s + (lc * P + iy * Q + a) * 4, *py = fy_s + (lc * P + a * Q + ix) * 4;
#pragma unroll
      for (int c = 0; c < 4; ++c) { acc[c] -= dx * px[c] + dy * py[c]; }
    }
    const double phis[4] = {g.e0[ix], g.e1[ix], g.e0[iy], g.e1[iy]};
    const int qs[4] = {iy, iy, ix, ix};
#pragma unroll
    for (int side = 0; side < 4; ++side) {
      const double *hh = h_s + (lc * NF + side * Q + qs[side]) * 4;
#pragma unroll
      for (int c = 0; c < 4; ++c) { acc[c] += phis[side] * hh[c]; }
    }
    const double wi = g.w1[ix] * g.w1[iy], im = 1. / wi, metric = g.area * wi;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double rc = (acc[c] * im + s[c]) - old[off + c];
      r[off + c] = rc;
      contrib += metric * rc * rc;
      snorm += metric * s[c] * s[c];
    }
    if (trial) {
#pragma unroll
      for (int c = 0; c < 4; ++c) { trial[off + c] = s[c]; }
    }
  }
  if (!ok) { atomicOr(flags, kErrInadmissible); }
  const double sum = block_sum(contrib, red);
  const double ssum = block_sum(snorm, red);
  if (tid == 0 && partial) {
    partial[blockIdx.x] = sum;
    partial[gridDim.x + blockIdx.x] = ssum;
  }
}

// src/problem.cpp:259-275: positivity at every node and face trace.
template <int Q>
__global__ __launch_bounds__(256) void k_admissible(Geo g, const double *__restrict__ u, int *__restrict__ flags) {
  constexpr int P = Q * Q, PT = P + 4 * Q;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= g.n * PT) { return; }
  const int e = id / PT, k = id % PT;
  double s[4];
  if (k < P) {
    for (int c = 0; c < 4; ++c) { s[c] = u[size_t(e) * g.m + 4 * k + c]; }
  } else {
    const int fp = k - P;
    trace<Q, double>(g, u + size_t(e) * g.m, fp / Q, fp % Q, s);
  }
  if (!positive(s, g.gm1)) { atomicOr(flags, kErrInadmissible); }
}

// src/partial.cpp:31-135: y = x + M^{-1}[-D_x^T(dt w J_x x) - D_y^T(dt w J_y x)
// + faces (jl x_L + jr x_R)], with the flux Jacobians of the Newton state u
// recomputed on the fly (Jacobian-vector products) instead of cached.
template <int Q, class T>
__global__ __launch_bounds__(256) void k_dg_apply(Geo g, const Frozen *__restrict__ fz,
                                                  const double *__restrict__ u, const T *__restrict__ x,
                                                  T *__restrict__ y) {
  constexpr int P = Q * Q, CPB = kBlock / P, NF = 4 * Q;
  __shared__ T fx_s[CPB * P * 4], fy_s[CPB * P * 4], h_s[CPB * NF * 4];
  const int tid = threadIdx.x, lc = tid / P, node = tid % P;
  const int e = blockIdx.x * CPB + lc;
  const bool active = lc < CPB && e < g.n;
  const int ix = node % Q, iy = node / Q;
  const T gm1 = T(g.gm1);
  const size_t off = size_t(e) * g.m + 4 * node;
  T xv[4] = {0, 0, 0, 0};
  if (active) {
    T s[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) { s[c] = T(u[off + c]); xv[c] = x[off + c]; }
    T fx[4], fy[4];
    jvp(s, 0, gm1, xv, fx);
    jvp(s, 1, gm1, xv, fy);
    const T wdt = T(g.dt * (g.w1[ix] * g.w1[iy]));
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      fx_s[(lc * P + node) * 4 + c] = wdt * fx[c];
      fy_s[(lc * P + node) * 4 + c] = wdt * fy[c];
    }
  }
  if (active) {
    for (int fp = node; fp < NF; fp += P) {
      const int side = fp / Q, qq = fp % Q, axis = side >> 1, left = side & 1;
      const int f = face_of(g, e, side), nb = neighbour_of(g, e, side);
      T uo[4], un[4], xo[4], xn[4];
      trace<Q, T>(g, u + size_t(e) * g.m, side, qq, uo);
      trace<Q, T>(g, u + size_t(nb) * g.m, side ^ 1, qq, un);
      trace<Q, T>(g, x + size_t(e) * g.m, side, qq, xo);
      trace<Q, T>(g, x + size_t(nb) * g.m, side ^ 1, qq, xn);
      const T *ul = left ? uo : un, *ur = left ? un : uo, *xl = left ? xo : xn, *xr = left ? xn : xo;
      T jl[4], jr[4];
      jvp(ul, axis, gm1, xl, jl);
      jvp(ur, axis, gm1, xr, jr);
      const Frozen z = fz[f * Q + qq];
      const T wl = T(z.wl), wr = T(z.wr), pen = T(z.pen);
      const double scale = g.dt * length_of(g, side) / g.area * g.w1[qq];
      const T sign = T(left ? scale : -scale);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        h_s[(lc * NF + fp) * 4 + c] = sign * (wl * jl[c] + wr * jr[c] + pen * (xl[c] - xr[c]));
      }
    }
  }
  __syncthreads();
  if (active) {
    T acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < Q; ++a) {
      const T dx = T(g.D[a * Q + ix] / g.hx), dy = T(g.D[a * Q + iy] / g.hy);
      const T *px = fx_s + (lc * P + iy * Q + a) * 4, *py = fy_s + (lc * P + a * Q + ix) * 4;
#pragma unroll
      for (int c = 0; c < 4; ++c) { acc[c] -= dx * px[c] + dy * py[c]; }
    }
    const T phis[4] = {T(g.e0[ix]), T(g.e1[ix]), T(g.e0[iy]), T(g.e1[iy])};
    const int qs[4] = {iy, iy, ix, ix};
#pragma unroll
    for (int side = 0; side < 4; ++side) {
      const T *hh = h_s + (lc * NF + side * Q + qs[side]) * 4;
#pragma unroll
      for (int c = 0; c < 4; ++c) { acc[c] += phis[side] * hh[c]; }
    }
    const T im = T(1. / (g.w1[ix] * g.w1[iy]));
#pragma unroll
    for (int c = 0; c < 4; ++c) { y[off + c] = acc[c] * im + xv[c]; }
  }
}

// Per-cell Jacobian staging shared by the block-fill kernels.
template <int Q> struct CellJac {
  static constexpr int P = Q * Q;
  double jn[P * 32];       // node Jacobians: [node][axis][16]
  double jf[4 * Q * 16];   // own-side face matrices (meaning set by the caller)
  double jo[4 * Q * 16];   // neighbour-side face matrices
};
template <int Q> __device__ void stage_node_jacobians(const Geo &g, const double *cell, CellJac<Q> &cj) {
  for (int node = threadIdx.x; node < Q * Q; node += blockDim.x) {
    double s[4];
    for (int c = 0; c < 4; ++c) { s[c] = cell[4 * node + c]; }
    jacobian(s, 0, g.gm1, cj.jn + node * 32);
    jacobian(s, 1, g.gm1, cj.jn + node * 32 + 16);
  }
}
// Volume part of VolumeJacobianCell (src/problem.cpp:519-550) for entry
// (4i+c, 4j+d), before the mass scaling.
template <int Q>
__device__ __forceinline__ double volume_entry(const Geo &g, const CellJac<Q> &cj, int i, int j, int c, int d) {
  const int ix = i % Q, iy = i / Q, jx = j % Q, jy = j / Q;
  double v = 0.;
  const double *jj = cj.jn + j * 32;
  if (iy == jy) { v += (g.D[jx * Q + ix] / g.hx) * jj[c + 4 * d]; }
  if (ix == jx) { v += (g.D[jy * Q + iy] / g.hy) * jj[16 + c + 4 * d]; }
  return -(g.dt * (g.w1[jx] * g.w1[jy])) * v;
}
// Face term coefficient phi_i*phi_j*scale for test node i on side s and
// trial node j on side sj, both on the normal line through one face point.
// Returns false when the pair is not coupled through that side.
template <int Q>
__device__ __forceinline__ bool face_pair(const Geo &g, int s, int sj, int i, int j, int &qq, double &w) {
  const int ix = i % Q, iy = i / Q, jx = j % Q, jy = j / Q;
  if (s < 2) {
    if (iy != jy) { return false; }
    qq = iy;
    w = side_value<Q, double>(g, s, ix) * side_value<Q, double>(g, sj, jx);
  } else {
    if (ix != jx) { return false; }
    qq = ix;
    w = side_value<Q, double>(g, s, iy) * side_value<Q, double>(g, sj, jy);
  }
  w *= g.dt * length_of(g, s) / g.area * g.w1[qq];
  return true;
}

// DG local blocks (src/partial.cpp:140-173,265-298): the diagonal block
// (volume + own-side face terms, mass-scaled, +I) and, when off != null, the
// four face-coupling blocks into BSR slots slot[e*4+side] (sides sharing a
// neighbour accumulate into the same slot in side order).
// Workgroups a cell of the DG block fill: small cells 4 (one cell a
// workgroup leaves few waves a CU on small grids), larger ones 1.
template <int Q> constexpr int dg_blocks_split() { return Q <= 4 ? 4 : 1; }
// Threads of the per-cell HDG condensation kernels: small cells take 512
// (one cell a workgroup leaves few waves a CU on small grids; more threads a
// cell hide the memory latency; 1024 would spill registers), larger ones 256
// (their register use needs it).
template <int Q> constexpr int cell_threads() { return Q <= 4 ? 512 : 256; }
template <int Q, class T>
__global__ __launch_bounds__(256) void k_dg_blocks(Geo g, const Frozen *__restrict__ fz,
                                                   const double *__restrict__ u, T *__restrict__ diag,
                                                   const int *__restrict__ dslot, T *__restrict__ bsr,
                                                   const int *__restrict__ slot) {
  constexpr int P = Q * Q, M = 4 * P, S = dg_blocks_split<Q>();
  __shared__ CellJac<Q> cj;
  // S workgroups a cell, each filling every S-th run of blockDim.x entries
  // (the node and face Jacobians are formed by each of them)
  const int e = blockIdx.x / S, part = blockIdx.x % S;
  const double *cell = u + size_t(e) * M;
  stage_node_jacobians<Q>(g, cell, cj);
  for (int fp = threadIdx.x; fp < 4 * Q; fp += blockDim.x) {
    const int side = fp / Q, qq = fp % Q, axis = side >> 1, left = side & 1;
    const int f = face_of(g, e, side), nb = neighbour_of(g, e, side);
    const Frozen z = fz[f * Q + qq];
    double uo[4], un[4], jo[16], jn[16];
    trace<Q, double>(g, cell, side, qq, uo);
    trace<Q, double>(g, u + size_t(nb) * M, side ^ 1, qq, un);
    jacobian(uo, axis, g.gm1, jo);
    jacobian(un, axis, g.gm1, jn);
    // own side: left cell uses jl = wl J(U_L) + pen I (sign +); right cell
    // uses jr = wr J(U_R) - pen I (sign -). Neighbour side the opposite.
    const double wown = left ? z.wl : z.wr, woth = left ? z.wr : z.wl;
    const double pown = left ? z.pen : -z.pen, poth = -pown;
    const double sg = left ? 1. : -1.;
    for (int k = 0; k < 16; ++k) {
      const bool dd = (k % 4) == (k / 4);
      cj.jf[fp * 16 + k] = sg * (wown * jo[k] + (dd ? pown : 0.));
      cj.jo[fp * 16 + k] = sg * (woth * jn[k] + (dd ? poth : 0.));
    }
  }
  __syncthreads();
  // entries formed in FP64, stored as T (FP32: the mixed form's factor)
  T *dblk = diag + size_t(dslot ? dslot[e] : e) * M * M;
  for (int idx = part * blockDim.x + threadIdx.x; idx < M * M; idx += S * blockDim.x) {
    const int row = idx % M, col = idx / M;
    const int i = row / 4, c = row % 4, j = col / 4, d = col % 4;
    double v = volume_entry<Q>(g, cj, i, j, c, d);
    const double im = 1. / (g.w1[i % Q] * g.w1[i / Q]);
    for (int s = 0; s < 4; ++s) {
      int qq;
      double w;
      if (face_pair<Q>(g, s, s, i, j, qq, w)) { v += w * cj.jf[(s * Q + qq) * 16 + c + 4 * d]; }
    }
    v *= im;
    if (row == col) { v += 1.; }
    dblk[idx] = T(v);
    if (bsr) {
      for (int s = 0; s < 4; ++s) {
        int qq;
        double w, o = 0.;
        if (face_pair<Q>(g, s, s ^ 1, i, j, qq, w)) { o = im * w * cj.jo[(s * Q + qq) * 16 + c + 4 * d]; }
        const int sl = slot[e * 4 + s];
        bool first = true;
        for (int s2 = 0; s2 < s; ++s2) { first = first && slot[e * 4 + s2] != sl; }
        T *blk = bsr + size_t(sl) * M * M;
        blk[idx] = first ? T(o) : T(blk[idx] + o);
      }
    }
  }
}

// HDG cell system (src/problem.cpp:642-702): A_e (m x m) and the
// restriction blocks C_e = -w_q (+-J(U_e) + stab I) of each face point.
template <int Q>
__global__ __launch_bounds__(cell_threads<Q>()) void k_hdg_build(Geo g, const Frozen *__restrict__ fz,
                                                   const double *__restrict__ u, double *__restrict__ A,
                                                   double *__restrict__ C) {
  constexpr int P = Q * Q, M = 4 * P;
  __shared__ CellJac<Q> cj;
  const int e = blockIdx.x;
  const double *cell = u + size_t(e) * M;
  stage_node_jacobians<Q>(g, cell, cj);
  for (int fp = threadIdx.x; fp < 4 * Q; fp += blockDim.x) {
    const int side = fp / Q, qq = fp % Q, axis = side >> 1, left = side & 1;
    const int f = face_of(g, e, side);
    const Frozen z = fz[f * Q + qq];
    double uo[4], jo[16];
    trace<Q, double>(g, cell, side, qq, uo);
    jacobian(uo, axis, g.gm1, jo);
    const double sg = left ? 1. : -1., stab = left ? z.a : z.b;
    for (int k = 0; k < 16; ++k) {
      const double dv = sg * jo[k] + ((k % 4) == (k / 4) ? stab : 0.);
      cj.jf[fp * 16 + k] = dv;
      C[(size_t(e) * 4 * Q + fp) * 16 + k] = -g.w1[qq] * dv;
    }
  }
  __syncthreads();
  double *ablk = A + size_t(e) * M * M;
  for (int idx = threadIdx.x; idx < M * M; idx += blockDim.x) {
    const int row = idx % M, col = idx / M;
    const int i = row / 4, c = row % 4, j = col / 4, d = col % 4;
    double v = volume_entry<Q>(g, cj, i, j, c, d);
    for (int s = 0; s < 4; ++s) {
      int qq;
      double w;
      if (face_pair<Q>(g, s, s, i, j, qq, w)) { v += w * cj.jf[(s * Q + qq) * 16 + c + 4 * d]; }
    }
    v *= 1. / (g.w1[i % Q] * g.w1[i / Q]);
    if (row == col) { v += 1.; }
    ablk[idx] = v;
  }
}

// X_e = A_e^{-1} Dense(B_e) and y_e = A_e^{-1} r_e (src/condensation.cpp:131-148,
// src/operators.cpp:121-142). Dense(B_e) column (side, qq, c) holds
// phi_a M^{-1} sf at rows 4*node_a + c of the normal line through (side, qq),
// so each X column combines q columns of A_e^{-1}. Output [X | y], m x (4t+1).
// One workgroup a cell takes A_e^{-1} a chunk of SR rows at a time into LDS
// (each entry read from memory once, runs of SR values of a column), forms
// those rows of X there, and sums the rows of y over NP parts (k = p, p +
// NP, ..., then the parts in order). The weights of Dense(B_e) are formed
// once a cell.
template <int Q> constexpr int solve_rows() { return Q <= 7 ? 32 : 16; }
template <int Q>
__global__ __launch_bounds__(256) void k_hdg_solve(Geo g, const Frozen *__restrict__ fz,
                                                   const double *__restrict__ Ainv, const double *__restrict__ res,
                                                   double *__restrict__ Xy) {
  constexpr int P = Q * Q, M = 4 * P, T4 = 16 * Q, SR = solve_rows<Q>(), NT = 256, NP = NT / SR;
  __shared__ double sf[4 * Q];
  __shared__ double bw[4 * Q * Q]; // weight of node a on line (side, qq): bw[(side * Q + qq) * Q + a]
  __shared__ double rs[M];
  __shared__ double ch[M * SR]; // rows r0 .. r0 + SR - 1 of A^{-1}: ch[k * SR + rr]
  __shared__ double yp[NP * SR];
  const int e = blockIdx.x, t = threadIdx.x;
  for (int fp = t; fp < 4 * Q; fp += NT) {
    const int side = fp / Q, qq = fp % Q, left = side & 1;
    const Frozen z = fz[face_of(g, e, side) * Q + qq];
    const double scale = g.dt * length_of(g, side) / g.area * g.w1[qq];
    sf[fp] = -scale * (left ? z.a : z.b);
  }
  for (int i = t; i < M; i += NT) { rs[i] = res[size_t(e) * M + i]; }
  __syncthreads(); // sf is complete
  for (int idx = t; idx < 4 * Q * Q; idx += NT) {
    const int fp = idx / Q, a = idx % Q, side = fp / Q, qq = fp % Q;
    const int node = line_node<Q>(side, qq, a);
    bw[idx] = side_value<Q, double>(g, side, a) * (1. / (g.w1[node % Q] * g.w1[node / Q])) * sf[fp];
  }
  const double *ai = Ainv + size_t(e) * M * M;
  double *xb = Xy + size_t(e) * M * (T4 + 1);
  // the next chunk is loaded into registers while this one is used
  constexpr int LPT = (M * SR + NT - 1) / NT;
  double pf[LPT];
  auto load_chunk = [&](int r0) {
#pragma unroll
    for (int l = 0; l < LPT; ++l) {
      const int idx = t + l * NT, k = idx / SR, rr = idx % SR;
      pf[l] = (idx < M * SR && r0 + rr < M) ? ai[size_t(k) * M + r0 + rr] : 0.;
    }
  };
  load_chunk(0);
  for (int r0 = 0; r0 < M; r0 += SR) {
#pragma unroll
    for (int l = 0; l < LPT; ++l) {
      const int idx = t + l * NT;
      if (idx < M * SR) { ch[idx] = pf[l]; }
    }
    __syncthreads(); // the chunk (and sf, rs, bw) are in LDS
    if (r0 + SR < M) { load_chunk(r0 + SR); }
    {
      const int rr = t % SR, part = t / SR;
      double v = 0.;
      for (int k = part; k < M; k += NP) { v += ch[k * SR + rr] * rs[k]; }
      yp[part * SR + rr] = v;
    }
    for (int idx = t; idx < SR * T4; idx += NT) {
      const int rr = idx % SR, col = idx / SR, row = r0 + rr;
      const int side = col / (4 * Q), qq = (col / 4) % Q, c = col % 4;
      double v = 0.;
      for (int a = 0; a < Q; ++a) {
        const int node = line_node<Q>(side, qq, a);
        v += ch[(4 * node + c) * SR + rr] * bw[(side * Q + qq) * Q + a];
      }
      if (row < M) { xb[row + size_t(M) * col] = v; }
    }
    __syncthreads(); // every read of ch is done; yp is complete
    if (t < SR && r0 + t < M) {
      double v = 0.;
      for (int part = 0; part < NP; ++part) { v += yp[part * SR + t]; }
      xb[r0 + t + size_t(M) * T4] = v;
    }
  }
}

// Condensed products (src/condensation.cpp:131-148, src/operators.cpp:7-26):
// S_e = C_e X_e written as the per-face row panels, and C_e y_e.
// Panel storage: [face][role(0 left cell,1 right cell)][col 4t][row t].
// One workgroup a cell takes [X | y] CC columns at a time into LDS (each
// column read from memory once; the next chunk in registers meanwhile) and
// forms those columns of S_e there, with the sums in the order of the
// per-entry form.
template <int Q>
__global__ __launch_bounds__(256) void k_hdg_panels(Geo g, const double *__restrict__ Xy,
                                                    const double *__restrict__ C, double *__restrict__ panels,
                                                    double *__restrict__ cy) {
  constexpr int P = Q * Q, M = 4 * P, T = 4 * Q, T4 = 16 * Q, CC = 16, NT = 256;
  constexpr int LPT = (M * CC + NT - 1) / NT;
  __shared__ double xs[CC * M];   // columns c0 .. c0 + CC - 1 of [X | y]
  __shared__ double cl[T4 * 4];   // C_e: 16 values a face point
  const int e = blockIdx.x, t = threadIdx.x;
  const double *xb = Xy + size_t(e) * M * (T4 + 1);
  const double *cb = C + size_t(e) * T4 * 4;
  for (int i = t; i < T4 * 4; i += NT) { cl[i] = cb[i]; }
  double pf[LPT];
  auto load_chunk = [&](int c0) {
#pragma unroll
    for (int l = 0; l < LPT; ++l) {
      const int idx = t + l * NT;
      pf[l] = (idx < M * CC && c0 + idx / M <= T4) ? xb[size_t(c0) * M + idx] : 0.;
    }
  };
  load_chunk(0);
  for (int c0 = 0; c0 <= T4; c0 += CC) {
#pragma unroll
    for (int l = 0; l < LPT; ++l) {
      const int idx = t + l * NT;
      if (idx < M * CC) { xs[idx] = pf[l]; }
    }
    __syncthreads(); // the chunk (and C_e) are in LDS
    if (c0 + CC <= T4) { load_chunk(c0 + CC); }
    for (int idx = t; idx < T4 * CC; idx += NT) {
      const int row = idx % T4, cc = idx / T4, col = c0 + cc;
      if (col > T4) { continue; }
      const int side = row / T, qq = (row / 4) % Q, c = row % 4;
      const double *xc = xs + cc * M;
      double val[4] = {0, 0, 0, 0};
      for (int a = 0; a < Q; ++a) {
        const double phi = side_value<Q, double>(g, side, a);
        const double *xv = xc + 4 * line_node<Q>(side, qq, a);
        for (int d = 0; d < 4; ++d) { val[d] += phi * xv[d]; }
      }
      const double *fl = cl + (side * Q + qq) * 16;
      double res = 0.;
      for (int d = 0; d < 4; ++d) { res += fl[c + 4 * d] * val[d]; }
      if (col == T4) {
        cy[size_t(e) * T4 + row] = res;
      } else {
        const int f = face_of(g, e, side);
        const int role = (side & 1) ? 0 : 1;
        panels[((size_t(f) * 2 + role) * T4 + col) * T + (row % T)] = res;
      }
    }
    __syncthreads(); // every read of the chunk is done
  }
}

// Per face: condensed right-hand side (cells in ascending order, as the
// reference accumulates them) and the trace diagonal D = w_q (a + b).
template <int Q>
__global__ void k_hdg_face(Geo g, const Frozen *__restrict__ fz, const double *__restrict__ cy,
                           double *__restrict__ rhs, double *__restrict__ D) {
  constexpr int T = 4 * Q, T4 = 16 * Q;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= g.nf * T) { return; }
  const int f = id / T, r = id % T, qq = r / 4;
  const int e = f / 2, axis = f & 1, ls = axis ? 3 : 1, rs = axis ? 2 : 0;
  const int er = neighbour_of(g, e, ls);
  const double vl = cy[size_t(e) * T4 + ls * T + r], vr = cy[size_t(er) * T4 + rs * T + r];
  rhs[id] = e < er ? (0. + vl) + vr : (0. + vr) + vl;
  const Frozen z = fz[f * Q + qq];
  D[id] = g.w1[qq] * (z.a + z.b);
}

// Schur action (src/operators.cpp:147-179): y_f = D_f x_f - S(first) - S(second),
// one 64-lane wave per face, lane = row of the face block.
// Faces 4 block .. 4 block + 3 of y = S x, one wave a face (g: nx, ny, nf).
template <int Q, class T, class G>
__device__ __forceinline__ void schur_apply_block(const G &g, int block, const T *__restrict__ panels,
                                                  const double *__restrict__ D, const T *__restrict__ x,
                                                  T *__restrict__ y, T (&xs)[4][2][16 * Q]) {
  constexpr int TT = 4 * Q, T4 = 16 * Q;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int f = min(block * 4 + wid, g.nf - 1);
  const bool valid = block * 4 + wid < g.nf;
  const int e = f / 2, axis = f & 1, ls = axis ? 3 : 1;
  const int cells[2] = {e, neighbour_of(g, e, ls)};
  for (int k = lane; k < 2 * T4; k += 64) {
    const int role = k / T4, col = k % T4, side = col / TT;
    xs[wid][role][col] = x[size_t(face_of(g, cells[role], side)) * TT + col % TT];
  }
  __syncthreads();
  if (valid && lane < TT) {
    T acc[2] = {T(0), T(0)};
    for (int role = 0; role < 2; ++role) {
      const T *pn = panels + (size_t(f) * 2 + role) * T4 * TT;
      T s = T(0);
      for (int col = 0; col < T4; ++col) { s += pn[col * TT + lane] * xs[wid][role][col]; }
      acc[role] = s;
    }
    const size_t id = size_t(f) * TT + lane;
    const T dx = T(D[id]) * x[id];
    const bool left_first = cells[0] < cells[1];
    y[id] = left_first ? (dx - acc[0]) - acc[1] : (dx - acc[1]) - acc[0];
  }
}
template <int Q, class T>
__global__ __launch_bounds__(256) void k_schur_apply(Geo g, const T *__restrict__ panels, const double *__restrict__ D,
                                                     const T *__restrict__ x, T *__restrict__ y) {
  __shared__ T xs[4][2][16 * Q];
  schur_apply_block<Q, T>(g, blockIdx.x, panels, D, x, y, xs);
}
// The Schur apply w = S z of every slot of a grouped Arnoldi step.
struct FaceGrid {
  int nx, ny, nf;
};
template <int Q, class T>
__global__ __launch_bounds__(256) void k_group_schur(const GroupTable *__restrict__ tab) {
  __shared__ T xs[4][2][16 * Q];
  const int count = tab->count, total = tab->schur_off[count];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int i = group_slot(tab->schur_off, count, v);
    const ArnoldiSlot &sl = tab->slot[i];
    if (sl.cycle->active == 0) { continue; } // the same for the whole workgroup
    const FaceGrid g{sl.nx, sl.ny, sl.nf};
    __syncthreads(); // xs is free again
    schur_apply_block<Q, T>(g, v - tab->schur_off[i], static_cast<const T *>(sl.panels), sl.D,
                            static_cast<const T *>(sl.z), static_cast<T *>(sl.w), xs);
  }
}

// Recovery (src/condensation.cpp:28-41,154-163): du_e = -X_e R_e dl - y_e.
template <int Q>
__global__ __launch_bounds__(256) void k_recover(Geo g, const double *__restrict__ Xy, const double *__restrict__ dl,
                                                 double *__restrict__ du) {
  constexpr int P = Q * Q, M = 4 * P, TT = 4 * Q, T4 = 16 * Q;
  __shared__ double xs[T4];
  const int e = blockIdx.x;
  for (int k = threadIdx.x; k < T4; k += blockDim.x) {
    xs[k] = dl[size_t(face_of(g, e, k / TT)) * TT + k % TT];
  }
  __syncthreads();
  const double *xb = Xy + size_t(e) * M * (T4 + 1);
  for (int row = threadIdx.x; row < M; row += blockDim.x) {
    double s = 0.;
    for (int col = 0; col < T4; ++col) { s += xb[size_t(col) * M + row] * xs[col]; }
    du[size_t(e) * M + row] = -s - xb[size_t(T4) * M + row];
  }
}

// Recovered DG residual (src/stepper.cpp:10-26): eps_DG = -B D^{-1} eps_S,
// with the cell-norm partial sums fused.
template <int Q>
__global__ __launch_bounds__(256) void k_recovered_residual(Geo g, const Frozen *__restrict__ fz,
                                                            const double *__restrict__ D,
                                                            const double *__restrict__ eps,
                                                            double *__restrict__ out,
                                                            double *__restrict__ partial) {
  constexpr int P = Q * Q, M = 4 * P, TT = 4 * Q;
  __shared__ double red[16];
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.;
  if (id < g.n * M) {
    const int e = id / M, row = id % M, i = row / 4, c = row % 4, ix = i % Q, iy = i / Q;
    const double im = 1. / (g.w1[ix] * g.w1[iy]);
    const int qs[4] = {iy, iy, ix, ix};
    const int as[4] = {ix, ix, iy, iy};
    double v = 0.;
    for (int side = 0; side < 4; ++side) {
      const int qq = qs[side], left = side & 1;
      const int f = face_of(g, e, side);
      const Frozen z = fz[f * Q + qq];
      const double scale = g.dt * length_of(g, side) / g.area * g.w1[qq];
      const double sfv = -scale * (left ? z.a : z.b);
      const size_t tid = size_t(f) * TT + 4 * qq + c;
      const double scaled = eps[tid] / D[tid];
      v += side_value<Q, double>(g, side, as[side]) * im * (sfv * scaled);
    }
    out[id] = -v;
    contrib = g.area * (g.w1[ix] * g.w1[iy]) * v * v;
  }
  const double sum = block_sum(contrib, red);
  if (threadIdx.x == 0) { partial[blockIdx.x] = sum; }
}

// HDG block-Jacobi blocks (src/operators.cpp:275-303): S_ff = diag(D_f) -
// (panel of f's cells restricted to f's own columns), cells in ascending order.
template <int Q, class T>
__global__ void k_hdg_diag_blocks(Geo g, const T *__restrict__ panels, const double *__restrict__ D,
                                  double *__restrict__ blocks) {
  constexpr int TT = 4 * Q, T4 = 16 * Q;
  const int f = blockIdx.x;
  const int e = f / 2, axis = f & 1, ls = axis ? 3 : 1, rs = axis ? 2 : 0;
  const int er = neighbour_of(g, e, ls);
  for (int idx = threadIdx.x; idx < TT * TT; idx += blockDim.x) {
    const int r = idx % TT, c = idx / TT;
    const double dv = r == c ? D[size_t(f) * TT + r] : 0.;
    const double pl = double(panels[((size_t(f) * 2 + 0) * T4 + ls * TT + c) * TT + r]);
    const double pr = double(panels[((size_t(f) * 2 + 1) * T4 + rs * TT + c) * TT + r]);
    blocks[size_t(f) * TT * TT + idx] = e < er ? (dv - pl) - pr : (dv - pr) - pl;
  }
}

// HDG BSR fill (src/operators.cpp:48-74): row f gets -Panel(X_e, side_f,
// other) for each of its two cells (ascending) and each other side, plus D on
// the diagonal. contrib[f*8 + k] gives the BSR slot of contribution k
// (k = cell-order*4 + other side).
template <int Q>
__global__ void k_hdg_fill(Geo g, const double *__restrict__ panels, const double *__restrict__ D,
                           const int *__restrict__ contrib, const int *__restrict__ row_ptr,
                           const int *__restrict__ diag_ptr, double *__restrict__ bsr,
                           const int *__restrict__ row_of) {
  constexpr int TT = 4 * Q, T4 = 16 * Q;
  const int f = blockIdx.x, fr = row_of ? row_of[f] : f;
  const int e = f / 2, axis = f & 1, ls = axis ? 3 : 1;
  const int er = neighbour_of(g, e, ls);
  const int first_role = e < er ? 0 : 1;
  const int b0 = row_ptr[fr], nb = row_ptr[fr + 1] - b0;
  for (int idx = threadIdx.x; idx < TT * TT; idx += blockDim.x) {
    const int r = idx % TT, c = idx / TT;
    double acc[8];
    for (int s = 0; s < nb; ++s) { acc[s] = (b0 + s == diag_ptr[fr] && r == c) ? D[size_t(f) * TT + r] : 0.; }
    for (int k = 0; k < 8; ++k) {
      const int role = k < 4 ? first_role : 1 - first_role, other = k % 4;
      const int s = contrib[f * 8 + k] - b0;
      acc[s] += -1. * panels[((size_t(f) * 2 + role) * T4 + other * TT + c) * TT + r];
    }
    for (int s = 0; s < nb; ++s) { bsr[size_t(b0 + s) * TT * TT + idx] = acc[s]; }
  }
}
} // namespace

// ------------------------------------------------------------------ launchers
#define EG_DISPATCH_Q(q, CALL)                                                                     \
  switch (q) {                                                                                     \
  case 1: { constexpr int Q = 1; CALL; } break;                                                    \
  case 2: { constexpr int Q = 2; CALL; } break;                                                    \
  case 3: { constexpr int Q = 3; CALL; } break;                                                    \
  case 4: { constexpr int Q = 4; CALL; } break;                                                    \
  case 5: { constexpr int Q = 5; CALL; } break;                                                    \
  case 6: { constexpr int Q = 6; CALL; } break;                                                    \
  case 7: { constexpr int Q = 7; CALL; } break;                                                    \
  case 8: { constexpr int Q = 8; CALL; } break;                                                    \
  case 9: { constexpr int Q = 9; CALL; } break;                                                    \
  default: break;                                                                                  \
  }

int cell_blocks(const Geo &g) { return (g.n + kBlock / g.p - 1) / (kBlock / g.p); }

void launch_freeze(const Geo &g, const double *u, double safety, Frozen *fz, double *partial, int *flags,
                   hipStream_t s) {
  const int nb = (g.nf * g.q + 255) / 256;
  EG_DISPATCH_Q(g.q, (k_freeze<Q><<<nb, 256, 0, s>>>(g, u, safety, fz, partial, flags)));
}
int freeze_blocks(const Geo &g) { return (g.nf * g.q + 255) / 256; }

void launch_admissible(const Geo &g, const double *u, int *flags, hipStream_t s) {
  const int nb = (g.n * (g.p + 4 * g.q) + 255) / 256;
  EG_DISPATCH_Q(g.q, (k_admissible<Q><<<nb, 256, 0, s>>>(g, u, flags)));
}
void launch_residual(const Geo &g, const Frozen *fz, const double *u, const double *du, double lambda,
                     const double *old, double *r, double *trial, double *partial, int *flags, hipStream_t s) {
  const int nb = cell_blocks(g);
  EG_DISPATCH_Q(g.q, (k_residual<Q><<<nb, kBlock, 0, s>>>(g, fz, u, du, lambda, old, r, trial, partial, flags)));
}
template <class T>
void launch_dg_apply(const Geo &g, const Frozen *fz, const double *u, const T *x, T *y, hipStream_t s) {
  const int nb = cell_blocks(g);
  EG_DISPATCH_Q(g.q, (k_dg_apply<Q, T><<<nb, kBlock, 0, s>>>(g, fz, u, x, y)));
}
template void launch_dg_apply<double>(const Geo &, const Frozen *, const double *, const double *, double *,
                                      hipStream_t);
template void launch_dg_apply<float>(const Geo &, const Frozen *, const double *, const float *, float *,
                                     hipStream_t);
template <class T>
void launch_dg_blocks(const Geo &g, const Frozen *fz, const double *u, T *diag, const int *dslot, T *bsr,
                      const int *slot, hipStream_t s) {
  EG_DISPATCH_Q(g.q,
                (k_dg_blocks<Q, T><<<g.n * dg_blocks_split<Q>(), 256, 0, s>>>(g, fz, u, diag, dslot, bsr, slot)));
}
template void launch_dg_blocks<double>(const Geo &, const Frozen *, const double *, double *, const int *, double *,
                                       const int *, hipStream_t);
template void launch_dg_blocks<float>(const Geo &, const Frozen *, const double *, float *, const int *, float *,
                                      const int *, hipStream_t);
void launch_hdg_build(const Geo &g, const Frozen *fz, const double *u, double *A, double *C, hipStream_t s) {
  EG_DISPATCH_Q(g.q, (k_hdg_build<Q><<<g.n, cell_threads<Q>(), 0, s>>>(g, fz, u, A, C)));
}
void launch_hdg_solve(const Geo &g, const Frozen *fz, const double *Ainv, const double *res, double *Xy,
                      hipStream_t s) {
  EG_DISPATCH_Q(g.q, (k_hdg_solve<Q><<<g.n, 256, 0, s>>>(g, fz, Ainv, res, Xy)));
}
void launch_hdg_panels(const Geo &g, const double *Xy, const double *C, double *panels, double *cy,
                       hipStream_t s) {
  EG_DISPATCH_Q(g.q, (k_hdg_panels<Q><<<g.n, 256, 0, s>>>(g, Xy, C, panels, cy)));
}
void launch_hdg_face(const Geo &g, const Frozen *fz, const double *cy, double *rhs, double *D, hipStream_t s) {
  const int nb = (g.nf * g.t + 255) / 256;
  EG_DISPATCH_Q(g.q, (k_hdg_face<Q><<<nb, 256, 0, s>>>(g, fz, cy, rhs, D)));
}
template <class T>
void launch_schur_apply(const Geo &g, const T *panels, const double *D, const T *x, T *y, hipStream_t s) {
  const int nb = (g.nf + 3) / 4;
  EG_DISPATCH_Q(g.q, (k_schur_apply<Q, T><<<nb, 256, 0, s>>>(g, panels, D, x, y)));
}
template void launch_schur_apply<double>(const Geo &, const double *, const double *, const double *, double *,
                                         hipStream_t);
template void launch_schur_apply<float>(const Geo &, const float *, const double *, const float *, float *,
                                        hipStream_t);
template <class T> bool launch_group_schur(int q, const GroupTable *tab, int grid, hipStream_t s) {
  bool ok = false;
  EG_DISPATCH_Q(q, ({
    k_group_schur<Q, T><<<grid, 256, 0, s>>>(tab);
    ok = true;
  }));
  return ok;
}
template bool launch_group_schur<double>(int, const GroupTable *, int, hipStream_t);
template bool launch_group_schur<float>(int, const GroupTable *, int, hipStream_t);
void launch_recover(const Geo &g, const double *Xy, const double *dl, double *du, hipStream_t s) {
  EG_DISPATCH_Q(g.q, (k_recover<Q><<<g.n, 256, 0, s>>>(g, Xy, dl, du)));
}
int recovered_blocks(const Geo &g) { return (g.n * g.m + 255) / 256; }
void launch_recovered_residual(const Geo &g, const Frozen *fz, const double *D, const double *eps, double *out,
                               double *partial, hipStream_t s) {
  const int nb = recovered_blocks(g);
  EG_DISPATCH_Q(g.q, (k_recovered_residual<Q><<<nb, 256, 0, s>>>(g, fz, D, eps, out, partial)));
}
void launch_initial(const Geo &g, const QuadRule &r, const WaveGeo &w, double *u, hipStream_t s) {
  EG_DISPATCH_Q(g.q, (k_initial<Q><<<g.n, 256, 0, s>>>(g, r, w, u)));
}
void launch_cell_moments(const Geo &g, const QuadRule &r, const WaveGeo &w, const double *u, double time, int mode,
                         double *partials, hipStream_t s) {
  EG_DISPATCH_Q(g.q, (k_cell_moments<Q><<<g.n, 256, 0, s>>>(g, r, w, u, time, mode, partials)));
}
void launch_physics_probe(const double *states, int count, double gamma, double *out, hipStream_t s) {
  k_physics_probe<<<(count + 63) / 64, 64, 0, s>>>(states, count, gamma, out);
}
void launch_hdg_diag_blocks(const Geo &g, const double *panels, const double *D, double *blocks, hipStream_t s) {
  EG_DISPATCH_Q(g.q, (k_hdg_diag_blocks<Q, double><<<g.nf, 256, 0, s>>>(g, panels, D, blocks)));
}
void launch_hdg_fill(const Geo &g, const double *panels, const double *D, const int *contrib, const int *row_ptr,
                     const int *diag_ptr, double *bsr, hipStream_t s, const int *row_of) {
  EG_DISPATCH_Q(g.q, (k_hdg_fill<Q><<<g.nf, 256, 0, s>>>(g, panels, D, contrib, row_ptr, diag_ptr, bsr, row_of)));
}
} // namespace eg
