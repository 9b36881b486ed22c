// Device-side building blocks shared by the HIP kernels: problem geometry
// passed by value, ideal-gas physics (flux, Jacobian, Jacobian-vector
// product) templated on the arithmetic type, and block reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace eg {
constexpr int kMaxQ = 9;           // k <= 8
constexpr int kBlock = 256;        // threads per block of the cell kernels
constexpr int kReduceBlocks = 512; // largest first-stage width of vector reductions
// First-stage width of a reduction over n entries: one 256-thread block per
// 512 entries, at most kReduceBlocks. Fixed for a given n, so every
// reduction of a vector has the same summation tree.
__host__ __device__ inline int reduce_blocks(long n) {
  const long b = (n + 511) / 512;
  return int(b < 1 ? 1 : (b > kReduceBlocks ? kReduceBlocks : b));
}

// Geometry and 1-D tables of one problem. Node i of a cell is (ix, iy) =
// (i % q, i / q), x fastest; cell e = iy*nx + ix; face 2e is the x-face with
// e on its left (side 1 of e, side 0 of the +x neighbour), face 2e+1 the
// y-face (side 3 of e, side 2 of the +y neighbour).
struct Geo {
  int nx, ny, n, q, p, m, t, nf;
  double hx, hy, area, gamma, gm1, dt;
  double D[kMaxQ * kMaxQ]; // D[a*q+i] = l_i'(x_a) (reference-interval derivative)
  double e0[kMaxQ], e1[kMaxQ], w1[kMaxQ];
};

// Tensor Gauss-Legendre rule of the initial projection and the error
// integrals (order max(24, 2k+8): nq = 13 points a direction for k <= 8) on
// [0,1], with the 1-D basis values b[s*q + i] = l_i(x_s) at its points.
constexpr int kMaxQuad = 13;
struct QuadRule {
  int nq;
  double x[kMaxQuad], w[kMaxQuad];
  double b[kMaxQuad * kMaxQ];
};
// Cell origins and the wave of the exact solution (angle, its cosine and sine).
struct WaveGeo {
  const double *origx, *origy;
  double alpha, ca, sa;
};

// Frozen HLL data of one face point, fixed for one implicit stage solve.
struct Frozen {
  double a, b, inv_sum, wl, wr, pen, pad0, pad1;
};

// ------------------------------------------------------------ grouped launches
// One Arnoldi step of many independent condensed (Schur) solves of one trace
// order in one launch per stage (kernels k_group_*): each instance is a slot
// of the table, and the virtual blocks of stage X of slot i are
// [X_off[i], X_off[i+1]): exactly the blocks its own single-solve launch of
// that stage would have, so every instance computes what its own launches
// compute, bit for bit, whatever else shares the launch. A fixed persistent
// grid walks the virtual blocks (the launches are recorded once in a graph
// and replayed with a new table).
constexpr int kGroupMax = 256;
// GMRES cycle of one solve on the device: the Arnoldi step index (count),
// whether the cycle goes on, and the Givens rotations of the Hessenberg
// columns so far (src/linear.cpp:100-131). The grouped step kernels read the
// step index and skip a solve whose cycle has ended; the last kernel of each
// step (k_group_tail) rotates the new column, decides whether the cycle goes
// on and mirrors the column to the solve's pinned host memory.
constexpr int kCycleMax = 64;
struct CycleState {
  int active, count, jmax, breakdown;
  double target; // tol |rhs|
  double g[kCycleMax + 1], cs[kCycleMax], sn[kCycleMax];
};
// Pinned mirror of a cycle, written by the device: [0] steps done, [1] cycle
// goes on, [2] breakdown, [kMirrorG + i] rotated residual entries,
// [kMirrorH + j (kCycleMax + 1) + i] rotated Hessenberg column j.
constexpr int kMirrorG = 8, kMirrorH = kMirrorG + kCycleMax + 8;
constexpr int kMirrorSize = kMirrorH + kCycleMax * (kCycleMax + 1);
struct ArnoldiSlot {
  void *V, *z, *w;        // Krylov basis (v_j at V + j N), z = M^-1 v_j, w = S z
  const void *metric;     // trace metric
  const void *bsr;        // ILU(0) factors
  void *mid;              // the sweeps' intermediate vector (armed between sweeps)
  const int *row_ptr, *diag_ptr, *col, *lower_rows, *upper_rows;
  int *ticket, *flags;    // the sweeps' ticket counters, the solve's error flags
  const void *panels;     // Schur face panels
  const double *D;        // trace diagonal
  double *partA, *partB, *scal; // reduction partials, projections
  CycleState *cycle;      // the solve's GMRES cycle (step index j = cycle->count)
  double *mirror;         // its pinned mirror
  long N;
  int nb, scaled, nx, ny, nf, pad;
};
struct GroupTable {
  int count;
  int sweep_off[kGroupMax + 1], schur_off[kGroupMax + 1], red_off[kGroupMax + 1], norm_off[kGroupMax + 1];
  ArnoldiSlot slot[kGroupMax];
};
// Slot of virtual block v: off[i] <= v < off[i + 1] (off increasing, count slots).
__device__ __forceinline__ int group_slot(const int *off, int count, int v) {
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= v) { lo = mid; } else { hi = mid; }
  }
  return lo;
}

// ------------------------------------------------------------ physics
template <class T> __host__ __device__ __forceinline__ T pressure(const T u[4], T gm1) {
  return gm1 * (u[3] - T(.5) * (u[1] * u[1] + u[2] * u[2]) / u[0]);
}
template <class T> __host__ __device__ __forceinline__ void flux(const T u[4], int axis, T gm1, T f[4]) {
  const T vn = u[axis + 1] / u[0], pr = pressure(u, gm1);
  f[0] = u[axis + 1];
  f[1] = u[1] * vn;
  f[2] = u[2] * vn;
  f[3] = (u[3] + pr) * vn;
  f[axis + 1] += pr;
}
// Column-major 4x4 dF_axis/dU: j[c + 4*d] = dF_c/dU_d.
template <class T> __host__ __device__ __forceinline__ void jacobian(const T u[4], int axis, T gm1, T j[16]) {
  const T inv = T(1) / u[0];
  const T vx = u[1] * inv, vy = u[2] * inv, vn = u[axis + 1] * inv;
  const T h = u[3] + pressure(u, gm1);
  const T dp[4] = {T(.5) * gm1 * (vx * vx + vy * vy), -gm1 * vx, -gm1 * vy, gm1};
  T dv[4] = {-vn * inv, T(0), T(0), T(0)};
  dv[axis + 1] = inv;
#pragma unroll
  for (int i = 0; i < 16; ++i) { j[i] = T(0); }
  j[0 + 4 * (axis + 1)] = T(1);
#pragma unroll
  for (int d = 0; d < 2; ++d) {
#pragma unroll
    for (int c = 0; c < 4; ++c) { j[(d + 1) + 4 * c] = u[d + 1] * dv[c] + (axis == d ? dp[c] : T(0)); }
    j[(d + 1) + 4 * (d + 1)] += vn;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) { j[3 + 4 * c] = vn * dp[c] + h * dv[c]; }
  j[15] += vn;
}
// y = dF_axis/dU (u) x without forming the matrix.
template <class T>
__device__ __forceinline__ void jvp(const T u[4], int axis, T gm1, const T x[4], T y[4]) {
  const T inv = T(1) / u[0];
  const T vx = u[1] * inv, vy = u[2] * inv, vn = u[axis + 1] * inv;
  const T h = u[3] + pressure(u, gm1);
  const T dpx = gm1 * (T(.5) * (vx * vx + vy * vy) * x[0] - vx * x[1] - vy * x[2] + x[3]);
  const T dvx = (x[axis + 1] - vn * x[0]) * inv;
  y[0] = x[axis + 1];
  y[1] = u[1] * dvx + vn * x[1] + (axis == 0 ? dpx : T(0));
  y[2] = u[2] * dvx + vn * x[2] + (axis == 1 ? dpx : T(0));
  y[3] = vn * dpx + h * dvx + vn * x[3];
}
__device__ __forceinline__ bool positive(const double u[4], double gm1) {
  for (int c = 0; c < 4; ++c) {
    if (!isfinite(u[c])) { return false; }
  }
  return u[0] > 1e-13 && pressure(u, gm1) > 1e-13;
}

// Cell/face topology helpers.
template <class G> __device__ __forceinline__ int face_of(const G &g, int e, int side) {
  const int ix = e % g.nx, iy = e / g.nx;
  switch (side) {
  case 0: return 2 * (iy * g.nx + (ix + g.nx - 1) % g.nx);
  case 1: return 2 * e;
  case 2: return 2 * (((iy + g.ny - 1) % g.ny) * g.nx + ix) + 1;
  default: return 2 * e + 1;
  }
}
template <class G> __device__ __forceinline__ int neighbour_of(const G &g, int e, int side) {
  const int ix = e % g.nx, iy = e / g.nx;
  switch (side) {
  case 0: return iy * g.nx + (ix + g.nx - 1) % g.nx;
  case 1: return iy * g.nx + (ix + 1) % g.nx;
  case 2: return ((iy + g.ny - 1) % g.ny) * g.nx + ix;
  default: return ((iy + 1) % g.ny) * g.nx + ix;
  }
}
// Node index of the a-th node along the normal line through face point qq of
// a side, and the basis value of that node at the side.
template <int Q> __device__ __forceinline__ int line_node(int side, int qq, int a) {
  return side < 2 ? qq * Q + a : a * Q + qq;
}
template <int Q, class T>
__device__ __forceinline__ T side_value(const Geo &g, int side, int a) {
  return T((side & 1) ? g.e1[a] : g.e0[a]);
}
// Trace of a cell vector (4 components per node) at face point qq of a side.
template <int Q, class T, class V>
__device__ __forceinline__ void trace(const Geo &g, const V *cell, int side, int qq, T out[4]) {
  out[0] = out[1] = out[2] = out[3] = T(0);
#pragma unroll
  for (int a = 0; a < Q; ++a) {
    const T phi = side_value<Q, T>(g, side, a);
    const V *v = cell + 4 * line_node<Q>(side, qq, a);
#pragma unroll
    for (int c = 0; c < 4; ++c) { out[c] += phi * T(v[c]); }
  }
}

// ------------------------------------------------------------ reductions
// Sum over the 64 lanes of a wavefront, the same value in every lane, in a
// fixed order: DPP butterflies within rows of 16 lanes (xor 1, xor 2, half-row
// mirror, row mirror; each step adds two equal-order partial sums, so every
// lane of a row holds the same bits), then the four row sums (a + b) + (c + d).
template <int CTRL> __device__ __forceinline__ double dpp_add(double v) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(u)), CTRL, 0xF, 0xF, false));
  const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(u >> 32)), CTRL, 0xF, 0xF, false));
  return v + __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
__device__ __forceinline__ double lane_read(double v, int lane) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(u)), lane));
  const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(u >> 32)), lane));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0xB1>(v);  // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);  // quad_perm [2,3,0,1]
  v = dpp_add<0x141>(v); // row_half_mirror
  v = dpp_add<0x140>(v); // row_mirror
  return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}
// Deterministic block sum (fixed tree); result valid in thread 0.
template <class T> __device__ __forceinline__ T block_sum(T v, T *lds /* >= 16 */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { lds[wid] = v; }
  __syncthreads();
  T s = T(0);
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) { s += lds[i]; }
  }
  __syncthreads();
  return s;
}
// K block sums at once (one pair of barriers): on return thread k < K holds
// the sum of v[k] over the block, in the same fixed tree as block_sum.
template <int K> __device__ __forceinline__ double block_sum_k(double (&v)[K], double *lds /* >= 16 K */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double w = wave_sum(v[k]);
    if (lane == k) { lds[wid * K + k] = w; }
  }
  __syncthreads();
  double s = 0.;
  if (threadIdx.x < K) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) { s += lds[i * K + threadIdx.x]; }
  }
  __syncthreads();
  return s;
}
} // namespace eg
