// Dense-block and Krylov kernels: large-block (b > 100) Gauss-Jordan inverse
// and ILU(0) block products in global memory, block GEMV for block-Jacobi,
// and the vector kernels of the
// weighted restarted GMRES (fused CGS2 passes with deterministic two-stage
// reductions).
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

#include "kernels.hpp"

namespace eg {
namespace {

// ------------------------------------------------------------ large blocks
// Blocks above 100 x 100 (cell blocks of k >= 5: 144, 196, 256, 324) do not
// fit one CU's registers and LDS, so they are worked on in global memory (L2)
// by one 1024-thread workgroup, with the pivot column and row staged in LDS.
constexpr int kBigThreads = 1024;
constexpr int kBigMax = 324; // largest cell block (k = 8)
constexpr int kPanelW = 32;  // columns of a Gauss-Jordan panel
constexpr int kChunkMax = 112; // columns a chunk of the rank-kPanelW update (b = 144)

struct BigLds {
  double pan[kPanelW * kBigMax]; // the panel, [column][row]
  double ak[kPanelW * kChunkMax]; // pivot rows of a column chunk, [panel column][chunk column]
  double col[kBigMax], row[kPanelW];
  unsigned long long key[kBigThreads / 64];
  int piv[kBigMax], perm[kBigMax];
};

__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long v, int mask) {
  const int lo = __shfl_xor(int(unsigned(v)), mask, 64), hi = __shfl_xor(int(unsigned(v >> 32)), mask, 64);
  return (static_cast<unsigned long long>(unsigned(hi)) << 32) | unsigned(lo);
}
// Columns of one chunk of the panel update: (b / 4) x (chunk / 4) thread
// tiles of 4 x 4 fill the 1024 threads (b a multiple of 4).
__device__ __forceinline__ int update_chunk(int b) {
  const int c = 4 * (kBigThreads / (b / 4));
  return c < kChunkMax ? c : kChunkMax;
}

// Partially pivoted in-place Gauss-Jordan inverse of one b x b column-major
// block in global memory (pivot: the row of maximal |a| among rows k..b-1,
// near-ties to the lowest row; explicit row interchanges, undone on the
// columns at the end), blocked by panels of kPanelW columns: the panel is
// eliminated in LDS (step k: swap the pivot row into row k, scale it, and
// subtract its multiples from every other row of the panel, leaving the
// panel's columns of the inverse), then every other column a, read once
// from memory, becomes P a + G (P a)_K with P the panel's row interchanges,
// K its pivot rows and G its final columns minus the identity on K: the
// same transformation as the panel's kPanelW single Gauss-Jordan steps, as a
// rank-kPanelW update (4 x 4 register tiles). Mathematically the inverse of
// the block whose LU the reference uses (src/linear.cpp:198-202,
// src/block_ilu.cpp:83-84, src/condensation.cpp:137-140). Returns true on a
// zero or non-finite pivot.
__device__ bool gj_global(double *__restrict__ A, int b, BigLds &L) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
  const int chunk = update_chunk(b);
  bool singular = false;
  for (int k0 = 0; k0 < b; k0 += kPanelW) {
    const int W = min(kPanelW, b - k0);
    for (int e = tid; e < W * b; e += nt) {
      const int w = e / b, i = e % b;
      L.pan[w * b + i] = A[i + size_t(k0 + w) * b];
    }
    for (int i = tid; i < b; i += nt) { L.perm[i] = i; }
    __syncthreads();
    for (int w = 0; w < W; ++w) {
      const int k = k0 + w;
      double *pc = L.pan + w * b;
      unsigned long long key = 0ull;
      for (int i = k + tid; i < b; i += nt) {
        key = umax64(key, (static_cast<unsigned long long>(__double_as_longlong(pc[i])) & 0x7FFFFFFFFFFFFE00ull) |
                              unsigned(511 - i));
      }
      for (int m = 32; m > 0; m >>= 1) { key = umax64(key, shfl_xor64(key, m)); }
      if (lane == 0) { L.key[wave] = key; }
      __syncthreads();
      unsigned long long best = 0ull;
      for (int v = 0; v < nt / 64; ++v) { best = umax64(best, L.key[v]); }
      const int p = 511 - int(best & 511ull); // rows k..b-1 all carry their index
      if (p != k && tid < W) {
        const double t = L.pan[tid * b + k];
        L.pan[tid * b + k] = L.pan[tid * b + p];
        L.pan[tid * b + p] = t;
      }
      if (tid == 0) {
        L.piv[k] = p;
        const int t = L.perm[k];
        L.perm[k] = L.perm[p];
        L.perm[p] = t;
      }
      __syncthreads();
      for (int i = tid; i < b; i += nt) { L.col[i] = pc[i]; }
      if (tid < W) { L.row[tid] = L.pan[tid * b + k]; }
      __syncthreads();
      const double pivot = L.row[w];
      if (!(fabs(pivot) > 0.) || !isfinite(pivot)) { singular = true; }
      const double inv = 1. / pivot;
      for (int e = tid; e < W * b; e += nt) {
        const int c = e / b, i = e % b;
        double v;
        if (i == k) { v = c == w ? inv : L.row[c] * inv; }
        else if (c == w) { v = -L.col[i] * inv; }
        else { v = L.pan[c * b + i] - L.col[i] * (L.row[c] * inv); }
        L.pan[c * b + i] = v;
      }
      __syncthreads();
    }
    // every other column: a <- P a + G (P a)_K
    const int tr = tid % (b / 4), tc = tid / (b / 4);
    const bool act = tc < chunk / 4;
    for (int c0 = 0; c0 < b; c0 += chunk) {
      const int nc = min(chunk, b - c0);
      for (int e = tid; e < W * nc; e += nt) {
        const int w = e / nc, jj = e % nc;
        L.ak[w * chunk + jj] = A[L.perm[k0 + w] + size_t(c0 + jj) * b];
      }
      __syncthreads();
      double acc[4][4];
      const int i0 = 4 * tr, j0 = 4 * tc;
      if (act) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int src = L.perm[i0 + r];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = c0 + j0 + c;
            acc[r][c] = (j0 + c < nc) ? A[src + size_t(j) * b] : 0.;
          }
        }
        for (int w = 0; w < W; ++w) {
          double g[4], a4[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) { g[r] = L.pan[w * b + i0 + r] - ((i0 + r == k0 + w) ? 1. : 0.); }
#pragma unroll
          for (int c = 0; c < 4; ++c) { a4[c] = L.ak[w * chunk + j0 + c]; }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int c = 0; c < 4; ++c) { acc[r][c] += g[r] * a4[c]; }
          }
        }
      }
      __syncthreads(); // every read of this chunk's columns is done
      if (act) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int j = c0 + j0 + c;
          if (j0 + c < nc && (j < k0 || j >= k0 + W)) {
#pragma unroll
            for (int r = 0; r < 4; ++r) { A[i0 + r + size_t(j) * b] = acc[r][c]; }
          }
        }
      }
    }
    for (int e = tid; e < W * b; e += nt) {
      const int w = e / b, i = e % b;
      A[i + size_t(k0 + w) * b] = L.pan[w * b + i];
    }
    __syncthreads();
  }
  for (int k = b - 1; k >= 0; --k) {
    const int p = L.piv[k];
    if (p != k) {
      for (int i = tid; i < b; i += nt) {
        const double t = A[i + size_t(k) * b];
        A[i + size_t(k) * b] = A[i + size_t(p) * b];
        A[i + size_t(p) * b] = t;
      }
      __syncthreads();
    }
  }
  return singular;
}

__global__ __launch_bounds__(kBigThreads) void k_block_inverse_big(double *__restrict__ blocks, int b,
                                                                   const int *__restrict__ index,
                                                                   int *__restrict__ flags) {
  __shared__ BigLds L;
  const int blk = index ? index[blockIdx.x] : blockIdx.x;
  if (gj_global(blocks + size_t(blk) * b * b, b, L) && threadIdx.x == 0) { atomicOr(flags, kErrSingular); }
}

// C = X Y (mode 0) or C -= X Y (mode 1) for b x b column-major blocks in
// global memory: four 256-thread groups each compute 64 x 64 tiles of C
// (4 x 4 per thread), staging 64 x 16 panels of X and 16 x 64 panels of Y in
// LDS. Every group runs the same number of barriers.
constexpr int kTile = 64, kPanel = 16;
struct ProductLds {
  double x[4][kTile * kPanel], y[4][kPanel * kTile];
};
// The workgroup computes the tiles 4 slice + g, 4 (slice + nslices) + g, ...
// so several workgroups can share one product.
__device__ void tile_product_global(const double *__restrict__ X, const double *__restrict__ Y, double *C, int b,
                                    int mode, ProductLds &S, int slice = 0, int nslices = 1) {
  const int tid = threadIdx.x, g = tid >> 8, lt = tid & 255, tr = lt % 16, tc = lt / 16;
  const int nt = (b + kTile - 1) / kTile, tiles = nt * nt;
  double *xs = S.x[g], *ys = S.y[g];
  for (int t0 = 4 * slice; t0 < tiles; t0 += 4 * nslices) {
    const int tile = t0 + g;
    const bool act = tile < tiles;
    const int r0 = (tile % nt) * kTile, c0 = (tile / nt) * kTile;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < b; k0 += kPanel) {
      for (int e = lt; e < kTile * kPanel; e += 256) {
        const int rr = e % kTile, kk = e / kTile;
        const int r = r0 + rr, kx = k0 + kk;
        xs[e] = (act && r < b && kx < b) ? X[r + size_t(kx) * b] : 0.;
        const int ky = k0 + (e % kPanel), cc = e / kPanel, c = c0 + cc;
        ys[e] = (act && ky < b && c < b) ? Y[ky + size_t(c) * b] : 0.;
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < kPanel; ++kk) {
        double a[4], y[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) { a[i] = xs[tr * 4 + i + kk * kTile]; y[i] = ys[kk + (tc * 4 + i) * kPanel]; }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int i = 0; i < 4; ++i) { acc[j][i] += a[i] * y[j]; }
        }
      }
      __syncthreads();
    }
    if (act) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = r0 + tr * 4 + i, c = c0 + tc * 4 + j;
          if (r < b && c < b) {
            double *cp = C + r + size_t(c) * b;
            *cp = mode == 0 ? acc[j][i] : *cp - acc[j][i];
          }
        }
      }
    }
  }
  __syncthreads();
}

// Block products of one BSR row of ILU(0) (src/block_ilu.cpp:73-86) per
// workgroup for large blocks, rows of one dependency level in parallel:
// L_ik = A_ik inv(A_kk) through the workgroup's scratch block, then
// A_ij -= L_ik A_kj for the cached update pairs. The launcher then replaces
// the level's updated diagonal blocks by their inverses.
__global__ __launch_bounds__(kBigThreads) void k_ilu_products_big(double *__restrict__ bsr, int b,
                                                                const int *__restrict__ rows,
                                                                const int *__restrict__ row_ptr,
                                                                const int *__restrict__ diag_ptr,
                                                                const int *__restrict__ col,
                                                                const int *__restrict__ upd_ptr,
                                                                const int *__restrict__ upd, double *scratch) {
  __shared__ ProductLds S;
  const size_t bb = size_t(b) * b;
  const int i = rows[blockIdx.x];
  double *tmp = scratch + size_t(blockIdx.x) * bb;
  for (int ik = row_ptr[i]; ik < diag_ptr[i]; ++ik) {
    double *aik = bsr + size_t(ik) * bb;
    tile_product_global(aik, bsr + size_t(diag_ptr[col[ik]]) * bb, tmp, b, 0, S);
    for (size_t t = threadIdx.x; t < bb; t += blockDim.x) { aik[t] = tmp[t]; }
    __syncthreads();
    for (int u = upd_ptr[ik]; u < upd_ptr[ik + 1]; ++u) {
      tile_product_global(aik, bsr + size_t(upd[2 * u + 1]) * bb, bsr + size_t(upd[2 * u]) * bb, b, 1, S);
    }
  }
}

// y_blk = A_blk x_blk for count independent b x b blocks (block blk at slot
// slot[blk] when given, else at blk); one wave per block.
template <class T>
__global__ __launch_bounds__(256) void k_block_gemv(const T *__restrict__ blocks, int b, int count,
                                                    const T *__restrict__ x, T *__restrict__ y,
                                                    const int *__restrict__ slot) {
  extern __shared__ unsigned char raw[];
  T *xs = reinterpret_cast<T *>(raw);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + wid;
  const bool valid = blk < count;
  T *xw = xs + wid * b;
  if (valid) { for (int i = lane; i < b; i += 64) { xw[i] = x[size_t(blk) * b + i]; } }
  __syncthreads();
  if (valid) {
    const T *a = blocks + size_t(slot ? slot[blk] : blk) * b * b;
    for (int r = lane; r < b; r += 64) {
      T s = T(0);
      for (int c = 0; c < b; ++c) { s += a[size_t(c) * b + r] * xw[c]; }
      y[size_t(blk) * b + r] = s;
    }
  }
}
// The same product for large blocks (b > 64, b a multiple of 4) with one
// 256-thread workgroup a block and 16-byte loads: thread t takes the V = 16 /
// sizeof(T) rows V p .. V p + V - 1 (p = t mod b / V) of the columns c = t /
// (b / V) mod G, G = 256 / (b / V) column groups, and the G partial sums of a
// row are added in LDS in group order.
template <class T>
__global__ __launch_bounds__(256) void k_block_gemv_wg(const T *__restrict__ blocks, int b,
                                                       const T *__restrict__ x, T *__restrict__ y,
                                                       const int *__restrict__ slot) {
  constexpr int V = 16 / sizeof(T);
  typedef T vec __attribute__((ext_vector_type(V)));
  extern __shared__ unsigned char raw[];
  T *xs = reinterpret_cast<T *>(raw); // b values of x, then G b partial sums
  T *part = xs + b;
  const int blk = blockIdx.x, t = threadIdx.x, np = b / V, ng = 256 / np, p = t % np, grp = t / np;
  for (int i = t; i < b; i += 256) { xs[i] = x[size_t(blk) * b + i]; }
  __syncthreads();
  if (grp < ng) {
    const T *a = blocks + size_t(slot ? slot[blk] : blk) * b * b + V * p;
    vec acc = vec(T(0));
    for (int c = grp; c < b; c += ng) { acc += *reinterpret_cast<const vec *>(a + size_t(c) * b) * xs[c]; }
#pragma unroll
    for (int e = 0; e < V; ++e) { part[grp * b + V * p + e] = acc[e]; }
  }
  __syncthreads();
  for (int r = t; r < b; r += 256) {
    T s = part[r];
    for (int k = 1; k < ng; ++k) { s += part[k * b + r]; }
    y[size_t(blk) * b + r] = s;
  }
}

// ------------------------------------------------------------ Krylov vectors
// Partial sums over a fixed grid of reduce_blocks(n) blocks (grid-stride),
// so every reduction has the same tree regardless of launch timing.
constexpr int kMaxVec = 64;

template <class T>
__device__ __forceinline__ void multidot_block(const T *__restrict__ V, long ldv, int nvec, const T *__restrict__ w,
                                               const T *__restrict__ metric, long n, double *__restrict__ partials,
                                               int vblock, int vgrid, double *red, double *acc_s) {
  for (int i0 = 0; i0 < nvec; i0 += 8) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long k = vblock * long(blockDim.x) + threadIdx.x; k < n; k += long(vgrid) * blockDim.x) {
      const double wk = double(metric[k] * w[k]);
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        if (i0 + t < nvec) { acc[t] += double(V[(i0 + t) * ldv + k]) * wk; }
      }
    }
    const double s = block_sum_k<8>(acc, red);
    if (threadIdx.x < 8 && i0 + threadIdx.x < nvec) { acc_s[i0 + threadIdx.x] = s; }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nvec; i += blockDim.x) { partials[size_t(i) * vgrid + vblock] = acc_s[i]; }
}
template <class T>
__global__ __launch_bounds__(256) void k_multidot(const T *__restrict__ V, long ldv, int nvec,
                                                  const T *__restrict__ w, const T *__restrict__ metric, long n,
                                                  double *__restrict__ partials) {
  __shared__ double red[16 * 8];
  __shared__ double acc_s[kMaxVec];
  multidot_block<T>(V, ldv, nvec, w, metric, n, partials, blockIdx.x, gridDim.x, red, acc_s);
}

// Reduce partials[nvec][nb] -> proj (LDS) redundantly in every block.
__device__ void reduce_partials(const double *__restrict__ partials, int nvec, int nb, double *proj) {
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int i = wid; i < nvec; i += nw) {
    double s = 0.;
    for (int k = lane; k < nb; k += 64) { s += partials[size_t(i) * nb + k]; }
    s = wave_sum(s);
    if (lane == 0) { proj[i] = s; }
  }
  __syncthreads();
}

// w -= V proj (proj from the previous pass's partials), then either the next
// pass's partial dots with V (norm_mode 0) or the partial metric norm of w (1).
// Every thread re-reads only the entries it wrote itself.
template <class T>
__device__ __forceinline__ void cgs_update_block(const T *__restrict__ V, long ldv, int nvec,
                                                 const double *__restrict__ partials_in, T *__restrict__ w,
                                                 const T *__restrict__ metric, long n, double *__restrict__ proj_out,
                                                 double *__restrict__ partials_out, int norm_mode, int vblock,
                                                 int vgrid, double *proj, double *red, double *acc_s) {
  reduce_partials(partials_in, nvec, reduce_blocks(n), proj);
  if (vblock == 0) {
    for (int i = threadIdx.x; i < nvec; i += blockDim.x) { proj_out[i] = proj[i]; }
  }
  const long first = vblock * long(blockDim.x) + threadIdx.x, stride = long(vgrid) * blockDim.x;
  double nacc = 0.;
  for (long k = first; k < n; k += stride) {
    T s = T(0);
    for (int i = 0; i < nvec; ++i) { s += V[i * ldv + k] * T(proj[i]); }
    const T wk = w[k] - s;
    w[k] = wk;
    nacc += double(metric[k] * wk * wk);
  }
  int nout = 1;
  if (norm_mode) {
    const double sn = block_sum(nacc, red);
    if (threadIdx.x == 0) { acc_s[0] = sn; }
  } else {
    nout = nvec;
    for (int i0 = 0; i0 < nvec; i0 += 8) {
      double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (long k = first; k < n; k += stride) {
        const double mw = double(metric[k] * w[k]);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          if (i0 + t < nvec) { acc[t] += double(V[(i0 + t) * ldv + k]) * mw; }
        }
      }
      const double sd = block_sum_k<8>(acc, red);
      if (threadIdx.x < 8 && i0 + threadIdx.x < nvec) { acc_s[i0 + threadIdx.x] = sd; }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nout; i += blockDim.x) { partials_out[size_t(i) * vgrid + vblock] = acc_s[i]; }
}
template <class T>
__global__ __launch_bounds__(256) void k_cgs_update(const T *__restrict__ V, long ldv, int nvec,
                                                    const double *__restrict__ partials_in, T *__restrict__ w,
                                                    const T *__restrict__ metric, long n, double *__restrict__ proj_out,
                                                    double *__restrict__ partials_out, int norm_mode) {
  __shared__ double proj[kMaxVec];
  __shared__ double red[16 * 8];
  __shared__ double acc_s[kMaxVec];
  cgs_update_block<T>(V, ldv, nvec, partials_in, w, metric, n, proj_out, partials_out, norm_mode, blockIdx.x,
                      gridDim.x, proj, red, acc_s);
}

// next = sqrt(sum partials); vnext = w * (1/next) if next > 0 (src/linear.cpp:117-122).
template <class T>
__device__ __forceinline__ void normalize_block(const double *__restrict__ partials, T *__restrict__ w,
                                                T *__restrict__ vnext, long n, double *__restrict__ norm2_out,
                                                const double *__restrict__ proj_in, int nproj,
                                                double *__restrict__ host_out, int vblock, int vgrid, double *proj) {
  reduce_partials(partials, 1, reduce_blocks(n), proj);
  const double norm2 = proj[0];
  if (vblock == 0 && threadIdx.x == 0) { norm2_out[0] = norm2; }
  if (vblock == 0 && host_out) {
    // the Hessenberg column: both projection passes and the squared norm
    for (int i = threadIdx.x; i < nproj; i += blockDim.x) {
      host_out[i] = proj_in[i];
      host_out[64 + i] = proj_in[64 + i];
    }
    if (threadIdx.x == 0) { host_out[128] = norm2; }
  }
  const double next = sqrt(norm2);
  if (next > 0) {
    const T inv = T(1 / next);
    for (long k = vblock * long(blockDim.x) + threadIdx.x; k < n; k += long(vgrid) * blockDim.x) {
      vnext[k] = w[k] * inv;
    }
  }
}
template <class T>
__global__ __launch_bounds__(256) void k_normalize(const double *__restrict__ partials, T *__restrict__ w,
                                                   T *__restrict__ vnext, long n, double *__restrict__ norm2_out,
                                                   const double *__restrict__ proj_in, int nproj,
                                                   double *__restrict__ host_out) {
  __shared__ double proj[1];
  normalize_block<T>(partials, w, vnext, n, norm2_out, proj_in, nproj, host_out, blockIdx.x, gridDim.x, proj);
}

// The CGS2 stages of every slot of a grouped Arnoldi step (stage 0: the
// first projection pass's partial dots, 1: w -= V h1 and the second pass's
// dots, 2: w -= V h2 and the norm, 3: the column to pinned memory and
// v_{j+1} = w / |w|), on each slot's own reduce_blocks(N) partials.
template <class T>
__global__ __launch_bounds__(256) void k_group_krylov(const GroupTable *__restrict__ tab, int stage) {
  __shared__ double proj[kMaxVec];
  __shared__ double red[16 * 8];
  __shared__ double acc_s[kMaxVec];
  const int count = tab->count;
  const int *off = stage == 3 ? tab->norm_off : tab->red_off;
  const int total = off[count];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int i = group_slot(off, count, v);
    const ArnoldiSlot &sl = tab->slot[i];
    if (sl.cycle->active == 0) { continue; } // the same for the whole workgroup
    const int vb = v - off[i], vg = off[i + 1] - off[i];
    const T *V = static_cast<const T *>(sl.V), *metric = static_cast<const T *>(sl.metric);
    T *w = static_cast<T *>(sl.w);
    const long N = sl.N;
    const int nv = sl.cycle->count + 1;
    __syncthreads(); // the shared arrays are free again
    if (stage == 0) {
      multidot_block<T>(V, N, nv, w, metric, N, sl.partA, vb, vg, red, acc_s);
    } else if (stage == 1) {
      cgs_update_block<T>(V, N, nv, sl.partA, w, metric, N, sl.scal, sl.partB, 0, vb, vg, proj, red, acc_s);
    } else if (stage == 2) {
      cgs_update_block<T>(V, N, nv, sl.partB, w, metric, N, sl.scal + 64, sl.partA, 1, vb, vg, proj, red, acc_s);
    } else {
      normalize_block<T>(sl.partA, w, static_cast<T *>(sl.V) + size_t(nv) * N, N, sl.scal + 128, sl.scal, nv,
                         nullptr, vb, vg, proj);
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_lincomb(const T *__restrict__ V, long ldv, int count, Coef coef,
                                                 T *__restrict__ out, long n) {
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < n; k += long(gridDim.x) * blockDim.x) {
    T s = T(0);
    for (int i = 0; i < count; ++i) { s += V[i * ldv + k] * T(coef.c[i]); }
    out[k] = s;
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_dot_partials(const T *__restrict__ a, const T *__restrict__ b,
                                                      const T *__restrict__ metric, long n,
                                                      double *__restrict__ partials) {
  __shared__ double red[16];
  double acc = 0.;
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < n; k += long(gridDim.x) * blockDim.x) {
    acc += double(metric[k] * a[k] * b[k]);
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) { partials[blockIdx.x] = s; }
}

// mode 0: sum, mode 1: max, mode 2: square root of the sum.
// One block per row of count partials: out[row] = sum (mode 0), max (mode 1)
// or sqrt(sum) (mode 2).
// host_out (optional, pinned host memory): the sums are also written there,
// and host_flags[0] = flags[0], so that one stream synchronisation brings
// both to the host without copy launches.
__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ partials, int count,
                                                      double *__restrict__ out, int mode, double *host_out,
                                                      const int *flags, int *host_flags) {
  __shared__ double red[16];
  partials += size_t(blockIdx.x) * count;
  out += blockIdx.x;
  if (host_flags && blockIdx.x == 0 && threadIdx.x == 0) { host_flags[0] = flags[0]; }
  double acc = 0.;
  for (int k = threadIdx.x; k < count; k += blockDim.x) {
    acc = mode == 1 ? fmax(acc, partials[k]) : acc + partials[k];
  }
  if (mode == 1) {
    for (int off = 32; off > 0; off >>= 1) { acc = fmax(acc, __shfl_down(acc, off, 64)); }
    if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6] = acc; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = 0.;
      for (int i = 0; i < int(blockDim.x >> 6); ++i) { m = fmax(m, red[i]); }
      out[0] = m;
      if (host_out) { host_out[blockIdx.x] = m; }
    }
  } else {
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) {
      const double v = mode == 2 ? sqrt(s) : s;
      out[0] = v;
      if (host_out) { host_out[blockIdx.x] = v; }
    }
  }
}

template <class A, class B>
__global__ void k_convert(const A *__restrict__ x, B *__restrict__ y, long n) {
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < n; k += long(gridDim.x) * blockDim.x) {
    y[k] = B(x[k]);
  }
}
template <class T>
__global__ void k_axpby(T *__restrict__ y, double a, const T *__restrict__ x, double b, long n) {
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < n; k += long(gridDim.x) * blockDim.x) {
    y[k] = b == 0. ? T(a) * x[k] : T(a) * x[k] + T(b) * y[k];
  }
}
__global__ void k_sub(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ out, long n) {
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < n; k += long(gridDim.x) * blockDim.x) {
    out[k] = a[k] - b[k];
  }
}
__global__ void k_scale_copy(const double *__restrict__ x, double a, double *__restrict__ y, long n) {
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < n; k += long(gridDim.x) * blockDim.x) {
    y[k] = x[k] * a;
  }
}
struct SlopeSet {
  const double *s[4];
  double c[4];
};
// out = base + sum_j c_j slope_j, accumulated in j order (src/stepper.cpp:99-104,246-250).
__global__ void k_dirk_combine(double *__restrict__ out, const double *__restrict__ base, SlopeSet set, int count,
                               long n) {
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < n; k += long(gridDim.x) * blockDim.x) {
    double v = base[k];
    for (int j = 0; j < count; ++j) { v += set.c[j] * set.s[j][k]; }
    out[k] = v;
  }
}
// K_i = (U_i - base) * (1/(a_ii dt)) (src/stepper.cpp:242-244)
__global__ void k_slope(const double *__restrict__ u, const double *__restrict__ base, double inv_dt,
                        double *__restrict__ slope, long n) {
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < n; k += long(gridDim.x) * blockDim.x) {
    slope[k] = (u[k] - base[k]) * inv_dt;
  }
}
int grid_for(long n) {
  const long b = (n + 255) / 256;
  return int(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}
} // namespace

// Split form of the large-block ILU(0) products of one dependency level, for
// patterns where no update targets a lower block of its own row (cell grids
// of 3 or more cells a side): first every L_ik = A_ik inv(A_kk) of the level
// (one workgroup each, through a scratch block), then every updated block
// A_t -= sum L_ik A_kj over its pairs in ascending k (one workgroup each).
// Same products and per-block order as k_ilu_products_big, on more CUs.
// Products are shared by gridDim.y workgroups (tile slices).
__global__ __launch_bounds__(kBigThreads) void k_ilu_lblocks_big(const double *__restrict__ bsr, int b,
                                                                 const int *__restrict__ a_ik,
                                                                 const int *__restrict__ a_kk, double *scratch) {
  __shared__ ProductLds S;
  const size_t bb = size_t(b) * b;
  tile_product_global(bsr + size_t(a_ik[blockIdx.x]) * bb, bsr + size_t(a_kk[blockIdx.x]) * bb,
                      scratch + size_t(blockIdx.x) * bb, b, 0, S, blockIdx.y, gridDim.y);
}
__global__ void k_copy_lblocks(double *__restrict__ bsr, int b, const int *__restrict__ a_ik,
                               const double *__restrict__ scratch) {
  const size_t bb = size_t(b) * b;
  double *aik = bsr + size_t(a_ik[blockIdx.y]) * bb;
  const double *tmp = scratch + size_t(blockIdx.y) * bb;
  for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < bb; t += size_t(gridDim.x) * blockDim.x) {
    aik[t] = tmp[t];
  }
}
__global__ __launch_bounds__(kBigThreads) void k_ilu_targets_big(double *__restrict__ bsr, int b,
                                                                 const int *__restrict__ t_slot,
                                                                 const int *__restrict__ t_pair_ptr,
                                                                 const int *__restrict__ pairs) {
  __shared__ ProductLds S;
  const size_t bb = size_t(b) * b;
  double *tgt = bsr + size_t(t_slot[blockIdx.x]) * bb;
  for (int q = t_pair_ptr[blockIdx.x]; q < t_pair_ptr[blockIdx.x + 1]; ++q) {
    tile_product_global(bsr + size_t(pairs[2 * q]) * bb, bsr + size_t(pairs[2 * q + 1]) * bb, tgt, b, 1, S,
                        blockIdx.y, gridDim.y);
  }
}

// Clears the error state and lets the LDS-resident factor kernels use the
// whole 160 KB workgroup budget.
int la_init() {
  (void)hipGetLastError();
  return ilu_factor_lds_init();
}
void launch_block_inverse(double *blocks, int b, int count, const int *index, int *flags, hipStream_t s) {
  if (gj_regs_ok(b)) { launch_block_inverse_fast(blocks, b, count, index, flags, s); return; }
  k_block_inverse_big<<<count, kBigThreads, 0, s>>>(blocks, b, index, flags);
}
template <class T>
void launch_block_gemv(const T *blocks, int b, int count, const T *x, T *y, hipStream_t s, const int *slot) {
  if (b > 64 && b % 4 == 0) {
    const int groups = 256 / (b / (16 / int(sizeof(T))));
    k_block_gemv_wg<T><<<count, 256, (groups + 1) * b * sizeof(T), s>>>(blocks, b, x, y, slot);
    return;
  }
  k_block_gemv<T><<<(count + 3) / 4, 256, 4 * b * sizeof(T), s>>>(blocks, b, count, x, y, slot);
}
template void launch_block_gemv<double>(const double *, int, int, const double *, double *, hipStream_t,
                                        const int *);
template void launch_block_gemv<float>(const float *, int, int, const float *, float *, hipStream_t, const int *);

void launch_ilu_factor_level(double *bsr, int b, const int *rows, const int *row_diag, int nrows,
                             const int *row_ptr, const int *diag_ptr, const int *col, const int *upd_ptr,
                             const int *upd, double *scratch, int *flags, hipStream_t s) {
  if (fast_dense_ok(b)) {
    launch_ilu_factor_fast(bsr, b, rows, nrows, row_ptr, diag_ptr, col, upd_ptr, upd, flags, s);
    return;
  }
  k_ilu_products_big<<<nrows, kBigThreads, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, upd_ptr, upd, scratch);
  launch_block_inverse(bsr, b, nrows, row_diag, flags, s);
}
void launch_ilu_factor_level_split(double *bsr, int b, int na, const int *a_ik, const int *a_kk, int nt,
                                   const int *t_slot, const int *t_pair_ptr, const int *pairs, const int *row_diag,
                                   int nrows, double *scratch, int *flags, hipStream_t s) {
  const int tiles = ((b + kTile - 1) / kTile) * ((b + kTile - 1) / kTile);
  const int slices = std::min((tiles + 3) / 4, 16);
  if (na > 0) {
    k_ilu_lblocks_big<<<dim3(na, slices), kBigThreads, 0, s>>>(bsr, b, a_ik, a_kk, scratch);
    k_copy_lblocks<<<dim3(16, na), 256, 0, s>>>(bsr, b, a_ik, scratch);
  }
  if (nt > 0) { k_ilu_targets_big<<<dim3(nt, slices), kBigThreads, 0, s>>>(bsr, b, t_slot, t_pair_ptr, pairs); }
  launch_block_inverse(bsr, b, nrows, row_diag, flags, s);
}
bool big_block(int b) { return !fast_dense_ok(b); }
int big_block_max() { return kBigMax; }

template <class T>
void launch_multidot(const T *V, long ldv, int nvec, const T *w, const T *metric, long n, double *partials,
                     hipStream_t s) {
  k_multidot<T><<<reduce_blocks(n), 256, 0, s>>>(V, ldv, nvec, w, metric, n, partials);
}
template <class T>
void launch_cgs_update(const T *V, long ldv, int nvec, const double *partials_in, T *w, const T *metric, long n,
                       double *proj_out, double *partials_out, int norm_mode, hipStream_t s) {
  k_cgs_update<T><<<reduce_blocks(n), 256, 0, s>>>(V, ldv, nvec, partials_in, w, metric, n, proj_out, partials_out,
                                                norm_mode);
}
template <class T>
void launch_normalize(const double *partials, T *w, T *vnext, long n, double *norm2_out, const double *proj_in,
                      int nproj, double *host_out, hipStream_t s) {
  k_normalize<T><<<grid_for(n), 256, 0, s>>>(partials, w, vnext, n, norm2_out, proj_in, nproj, host_out);
}
template <class T>
void launch_lincomb(const T *V, long ldv, int count, const Coef &coef, T *out, long n, hipStream_t s) {
  k_lincomb<T><<<grid_for(n), 256, 0, s>>>(V, ldv, count, coef, out, n);
}
template <class T>
void launch_dot_partials(const T *a, const T *b, const T *metric, long n, double *partials, hipStream_t s) {
  k_dot_partials<T><<<reduce_blocks(n), 256, 0, s>>>(a, b, metric, n, partials);
}
void launch_sum_partials(const double *partials, int count, double *out, int mode, hipStream_t s, int rows,
                         double *host_out, const int *flags, int *host_flags) {
  k_sum_partials<<<rows, 256, 0, s>>>(partials, count, out, mode, host_out, flags, host_flags);
}
template <class A, class B> void launch_convert(const A *x, B *y, long n, hipStream_t s) {
  k_convert<A, B><<<grid_for(n), 256, 0, s>>>(x, y, n);
}
template <class T> void launch_axpby(T *y, double a, const T *x, double b, long n, hipStream_t s) {
  k_axpby<T><<<grid_for(n), 256, 0, s>>>(y, a, x, b, n);
}
void launch_sub(const double *a, const double *b, double *out, long n, hipStream_t s) {
  k_sub<<<grid_for(n), 256, 0, s>>>(a, b, out, n);
}
void launch_scale_copy(const double *x, double a, double *y, long n, hipStream_t s) {
  k_scale_copy<<<grid_for(n), 256, 0, s>>>(x, a, y, n);
}
void launch_dirk_combine(double *out, const double *base, const double *const *slopes, const double *coef, int count,
                         long n, hipStream_t s) {
  SlopeSet set{};
  for (int j = 0; j < count; ++j) { set.s[j] = slopes[j]; set.c[j] = coef[j]; }
  k_dirk_combine<<<grid_for(n), 256, 0, s>>>(out, base, set, count, n);
}
void launch_slope(const double *u, const double *base, double inv_dt, double *slope, long n, hipStream_t s) {
  k_slope<<<grid_for(n), 256, 0, s>>>(u, base, inv_dt, slope, n);
}

// ------------------------------------------------------------ no allocation
// Scratch of the allocation-free norm (eg_cell_norm_noalloc): the first-stage
// partial sums of one reduction, at most kReduceBlocks of them (4 KiB), one
// copy per device, held until the process ends. Every call on a device shares
// its copy, so calls that overlap on one device race; each call writes every
// partial it then reads, so nothing has to be cleared between calls.
static __device__ __align__(256) double scratch[kReduceBlocks];

namespace {
constexpr int kScratchDevices = 64;
std::mutex scratch_mutex;
double *scratch_address[kScratchDevices] = {};
// Device address of the current device's copy of scratch, resolved once per
// device with that device current (host code never passes the array's name,
// which is only a host-side stand-in, to a kernel or a copy).
double *device_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) {
    (void)hipGetLastError();
    throw std::runtime_error("eg_cell_norm_noalloc: no usable current device");
  }
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (scratch_address[dev] == nullptr) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      (void)hipGetLastError();
      throw std::runtime_error(std::string("eg_cell_norm_noalloc: hipGetSymbolAddress failed: ") +
                               hipGetErrorString(e));
    }
    scratch_address[dev] = static_cast<double *>(ptr);
  }
  return scratch_address[dev];
}
} // namespace

void launch_cell_norm_noalloc(const double *v, const double *metric, long n, double *out, hipStream_t s) {
  if (n < 1 || n > kNoallocMaxValues) {
    throw std::runtime_error("eg_cell_norm_noalloc: n must be in 1 .. 2147483647");
  }
  if (v == nullptr || metric == nullptr || out == nullptr) {
    throw std::runtime_error("eg_cell_norm_noalloc: null device pointer");
  }
  double *partials = device_scratch();
  const int count = reduce_blocks(n); // at most kReduceBlocks, the size of scratch
  k_dot_partials<double><<<count, 256, 0, s>>>(v, v, metric, n, partials);
  k_sum_partials<<<1, 256, 0, s>>>(partials, count, out, 2, nullptr, nullptr, nullptr);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { throw std::runtime_error(std::string("eg_cell_norm_noalloc: ") + hipGetErrorString(e)); }
}

#define EG_INST_T(T)                                                                                          \
  template void launch_multidot<T>(const T *, long, int, const T *, const T *, long, double *, hipStream_t);   \
  template void launch_cgs_update<T>(const T *, long, int, const double *, T *, const T *, long, double *,     \
                                     double *, int, hipStream_t);                                             \
  template void launch_normalize<T>(const double *, T *, T *, long, double *, const double *, int, double *,  \
                                     hipStream_t);                   \
  template void launch_lincomb<T>(const T *, long, int, const Coef &, T *, long, hipStream_t);                \
  template void launch_dot_partials<T>(const T *, const T *, const T *, long, double *, hipStream_t);         \
  template void launch_axpby<T>(T *, double, const T *, double, long, hipStream_t);
EG_INST_T(double)
EG_INST_T(float)
int elementwise_grid(long n) { return grid_for(n); }

// The end of a grouped Arnoldi step j for every slot whose cycle goes on
// (one thread a slot): the Hessenberg column from the two CGS2 passes and
// the norm, the earlier Givens rotations applied to it, its own rotation,
// the rotated residual, and whether the cycle goes on (src/linear.cpp:
// 117-131: |g_{j+1}| above tol |rhs|, a new direction, steps left); the
// column and the state go to the pinned mirror.
__global__ __launch_bounds__(64) void k_group_tail(const GroupTable *__restrict__ tab) {
  const int count = tab->count;
  for (int i = blockIdx.x * 64 + threadIdx.x; i < count; i += gridDim.x * 64) {
    const ArnoldiSlot &sl = tab->slot[i];
    CycleState &c = *sl.cycle;
    if (c.active == 0) { continue; }
    const int j = c.count;
    const double *col = sl.scal;
    double *m = sl.mirror;
    double h[kCycleMax + 1];
    for (int r = 0; r <= j; ++r) { h[r] = (0. + col[r]) + col[64 + r]; }
    const double next = sqrt(col[128]);
    h[j + 1] = next;
    for (int r = 0; r < j; ++r) {
      const double u = h[r], vr = h[r + 1];
      h[r] = c.cs[r] * u + c.sn[r] * vr;
      h[r + 1] = -c.sn[r] * u + c.cs[r] * vr;
    }
    const double d = hypot(h[j], h[j + 1]);
    if (!(isfinite(d) && d > 0)) {
      c.active = 0;
      c.breakdown = 1;
      m[1] = 0.;
      m[2] = 1.;
      continue;
    }
    c.cs[j] = h[j] / d;
    c.sn[j] = h[j + 1] / d;
    h[j] = d;
    c.g[j + 1] = -c.sn[j] * c.g[j];
    c.g[j] *= c.cs[j];
    c.count = j + 1;
    const bool more = !(fabs(c.g[j + 1]) <= c.target || next <= 2.2250738585072014e-308) && j + 1 < c.jmax;
    c.active = more ? 1 : 0;
    double *hc = m + kMirrorH + size_t(j) * (kCycleMax + 1);
    for (int r = 0; r <= j; ++r) { hc[r] = h[r]; }
    m[kMirrorG + j] = c.g[j];
    m[kMirrorG + j + 1] = c.g[j + 1];
    m[0] = double(j + 1);
    m[1] = more ? 1. : 0.;
  }
}
__global__ void k_cycle_init(CycleState *c, double *mirror, double beta, double target, int jmax) {
  if (threadIdx.x == 0) {
    c->active = jmax > 0 ? 1 : 0;
    c->count = 0;
    c->jmax = jmax;
    c->breakdown = 0;
    c->target = target;
    c->g[0] = beta;
    mirror[0] = 0.;
    mirror[1] = c->active;
    mirror[2] = 0.;
    mirror[kMirrorG] = beta;
  }
}
void launch_group_tail(const GroupTable *tab, hipStream_t s) { k_group_tail<<<kGroupMax / 64, 64, 0, s>>>(tab); }
void launch_cycle_init(CycleState *c, double *mirror, double beta, double target, int jmax, hipStream_t s) {
  k_cycle_init<<<1, 64, 0, s>>>(c, mirror, beta, target, jmax);
}
template <class T> void launch_group_krylov(int stage, const GroupTable *tab, int grid, hipStream_t s) {
  k_group_krylov<T><<<grid, 256, 0, s>>>(tab, stage);
}
template void launch_group_krylov<double>(int, const GroupTable *, int, hipStream_t);
template void launch_group_krylov<float>(int, const GroupTable *, int, hipStream_t);
template void launch_convert<double, float>(const double *, float *, long, hipStream_t);
template void launch_convert<float, double>(const float *, double *, long, hipStream_t);
template void launch_convert<double, double>(const double *, double *, long, hipStream_t);
} // namespace eg
