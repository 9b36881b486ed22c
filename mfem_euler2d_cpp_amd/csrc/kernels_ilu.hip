// Dense block kernels for the block-Jacobi / ILU(0) hot spots, specialised on
// the block size B (B = 4q for face blocks, 4q^2 <= 100 for cell blocks):
//  * Gauss-Jordan inverse with the block resident in registers, implicit
//    partial pivoting (no row swaps) and a branch-free update; 1024-thread
//    workgroups for the 64 and 100 blocks so each SIMD has 4 waves to hide
//    the FP64 latency of the elimination chain,
//  * ILU(0) factor rows with LDS-staged, 4x4 register-tiled block products,
//  * ILU(0) sweeps whose block GEMVs split each row's dot product across
//    threads, per dependency level or as one dataflow launch.
#include <algorithm>

#include "kernels.hpp"

namespace eg {
namespace {
// Relaxed agent-scope atomic accesses: coherent across XCDs (sc1) without
// cache write-back/invalidate; used for data that another workgroup of the
// same launch writes or reads.
template <class T> __device__ __forceinline__ T ld_coherent(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void st_coherent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool COH> __device__ __forceinline__ double ld(const double *p) { return COH ? ld_coherent(p) : *p; }
template <bool COH> __device__ __forceinline__ void st(double *p, double v) {
  if (COH) { st_coherent(p, v); } else { *p = v; }
}
// Complete this thread's outstanding memory operations (ordering before a
// workgroup barrier that precedes a completion flag).
__device__ __forceinline__ void complete_memory_ops() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int B> struct Shape {
  static constexpr int R = std::min((B >= 64 ? 1024 : 256) / B, B); // row groups
  static constexpr int NT = (B * R + 63) / 64 * 64;                  // threads per workgroup
  static constexpr int NE = (B + R - 1) / R;                         // register entries per thread
};
// LDS scratch of the Gauss-Jordan (doubles): column and pivot row (double
// buffered), R candidates (double buffered), then ints.
template <int B> constexpr int gj_lds() { return 4 * B + 2 * Shape<B>::R + (2 * Shape<B>::R + 2 * B + 1) / 2 + 1; }

// In-place inverse of the B x B column-major block A by one workgroup of
// Shape<B>::NT threads. Thread (tc, tr) = (tid % B, tid / B) owns column tc
// at rows tr, tr + R, ... in registers. Column k is eliminated with pivot
// row p = the first unused row of maximal |S(p,k)| (NaN counts as 0 so the
// index stays in range); the stored matrix S then satisfies
// A^{-1}(pinv[i], piv[j]) = S(i,j). The owners of column k publish it with
// one pivot candidate per row group, so the choice costs R LDS reads.
// Mathematically the inverse of the block whose LU the reference uses
// (src/linear.cpp:198-202, src/block_ilu.cpp:83-84, src/condensation.cpp:137-140).
template <int B, bool COH>
__device__ __forceinline__ bool gj_inverse(double *__restrict__ A, double *lds) {
  constexpr int R = Shape<B>::R, NE = Shape<B>::NE;
  double *colk = lds, *rowp = lds + 2 * B, *cbest = lds + 4 * B;
  int *carg = reinterpret_cast<int *>(cbest + 2 * R), *piv = carg + 2 * R, *pinv = piv + B;
  const int tid = threadIdx.x, tc = tid % B, tr = tid / B;
  const bool active = tr < R;
  double v[NE];
  unsigned long long used = 0ull;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int i = tr + e * R;
    v[e] = (active && i < B) ? ld<COH>(A + i + tc * B) : 0.;
  }
  __syncthreads();
  bool singular = false;
  for (int k = 0; k < B; ++k) {
    double *ck = colk + (k & 1) * B, *rp = rowp + (k & 1) * B, *cb = cbest + (k & 1) * R;
    int *ca = carg + (k & 1) * R;
    if (active && tc == k) {
      double best = -1.;
      int arg = B;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const int i = tr + e * R;
        if (i < B) {
          ck[i] = v[e];
          const double a = isnan(v[e]) ? 0. : fabs(v[e]);
          if (!((used >> e) & 1ull) && a > best) { best = a; arg = i; }
        }
      }
      cb[tr] = best;
      ca[tr] = arg;
    }
    __syncthreads();
    double best = -1.;
    int p = B;
#pragma unroll 4
    for (int g = 0; g < R; ++g) {
      const double ob = cb[g];
      const int oa = ca[g];
      if (ob > best || (ob == best && oa < p)) { best = ob; p = oa; }
    }
    const double ap = ck[p];
    if (!(fabs(ap) > 0.) || !isfinite(ap)) { singular = true; }
    const double inv = 1. / ap;
    const int ep = p / R;
    if (active && tr == p - ep * R) {
      used |= 1ull << ep;
      double sel = 0.;
#pragma unroll
      for (int e = 0; e < NE; ++e) { sel = e == ep ? v[e] : sel; }
      rp[tc] = sel;
    }
    __syncthreads();
    if (tid == 0) { piv[k] = p; pinv[p] = k; }
    if (active) {
      const double rpj = rp[tc] * inv;
      const bool is_col = tc == k;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const int i = tr + e * R;
        const double c = ck[i < B ? i : 0];
        const bool is_row = i == p;
        double nv = v[e] - c * rpj;
        nv = is_row ? v[e] * inv : nv;
        nv = is_col ? (is_row ? inv : -c * inv) : nv;
        v[e] = nv;
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int i = tr + e * R;
      if (i < B) { st<COH>(A + pinv[i] + piv[tc] * B, v[e]); }
    }
  }
  __syncthreads();
  return singular;
}

template <int B>
__global__ __launch_bounds__(Shape<B>::NT) void k_block_inverse_b(double *__restrict__ blocks,
                                                                  const int *__restrict__ index,
                                                                  int *__restrict__ flags) {
  __shared__ double lds[gj_lds<B>()];
  const int blk = index ? index[blockIdx.x] : blockIdx.x;
  if (gj_inverse<B, false>(blocks + size_t(blk) * B * B, lds) && threadIdx.x == 0) { atomicOr(flags, kErrSingular); }
}

// Stage a global column-major block into LDS, optionally transposed.
template <int B, bool COH> __device__ __forceinline__ void stage(const double *__restrict__ src, double *dst, bool tr) {
  for (int t = threadIdx.x; t < B * B; t += blockDim.x) {
    const double v = ld<COH>(src + t);
    if (tr) { dst[(t / B) + (t % B) * B] = v; }
    else { dst[t] = v; }
  }
}
// 4x4-tiled product of X (LDS, column-major) and Y (LDS, transposed):
// mode 0: C = X Y, mode 1: C -= X Y, C in global memory.
template <int B, bool COH>
__device__ __forceinline__ void tile_product(const double *__restrict__ X, const double *__restrict__ Yt,
                                             double *__restrict__ C, int mode) {
  constexpr int nt4 = B / 4, tiles = nt4 * nt4;
  for (int tile = threadIdx.x; tile < tiles; tile += blockDim.x) {
    const int r0 = (tile % nt4) * 4, c0 = (tile / nt4) * 4;
    double acc[4][4] = {};
#pragma unroll 4
    for (int kk = 0; kk < B; ++kk) {
      double a[4], y[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) { a[i] = X[r0 + i + kk * B]; y[i] = Yt[c0 + i + kk * B]; }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int i = 0; i < 4; ++i) { acc[j][i] += a[i] * y[j]; }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        double *cp = C + (r0 + i) + (c0 + j) * B;
        st<COH>(cp, mode == 0 ? acc[j][i] : ld<COH>(cp) - acc[j][i]);
      }
    }
  }
}
// ILU(0) row i (src/block_ilu.cpp:73-86): for each lower block, L_ik = A_ik
// inv(A_kk) and A_ij -= L_ik A_kj over the cached update pairs; then the
// updated diagonal block is replaced by its inverse. LDS: 2 B^2 doubles.
template <int B, bool COH>
__device__ __forceinline__ bool factor_row(double *bsr, int i, const int *__restrict__ row_ptr,
                                           const int *__restrict__ diag_ptr, const int *__restrict__ col,
                                           const int *__restrict__ upd_ptr, const int *__restrict__ upd, double *sm) {
  constexpr size_t bb = size_t(B) * B;
  double *X = sm, *Y = sm + bb;
  for (int ik = row_ptr[i]; ik < diag_ptr[i]; ++ik) {
    double *aik = bsr + ik * bb;
    stage<B, COH>(aik, X, false);                        // X <- A_ik
    stage<B, COH>(bsr + diag_ptr[col[ik]] * bb, Y, true); // Y <- inv(A_kk)^T
    __syncthreads();
    tile_product<B, COH>(X, Y, aik, 0);                   // A_ik <- L_ik
    __syncthreads();
    stage<B, COH>(aik, X, false);                        // X <- L_ik
    for (int u = upd_ptr[ik]; u < upd_ptr[ik + 1]; ++u) {
      __syncthreads();
      stage<B, COH>(bsr + size_t(upd[2 * u + 1]) * bb, Y, true); // Y <- A_kj^T
      __syncthreads();
      tile_product<B, COH>(X, Y, bsr + size_t(upd[2 * u]) * bb, 1);
    }
    __syncthreads();
  }
  return gj_inverse<B, COH>(bsr + diag_ptr[i] * bb, sm);
}
template <int B> constexpr size_t factor_lds() {
  return std::max<size_t>(2 * size_t(B) * B, size_t(gj_lds<B>())) * sizeof(double);
}
// One row per workgroup, rows of one dependency level in parallel.
template <int B>
__global__ __launch_bounds__(Shape<B>::NT) void k_ilu_factor_rows(double *__restrict__ bsr, const int *__restrict__ rows,
                                                                  const int *__restrict__ row_ptr,
                                                                  const int *__restrict__ diag_ptr,
                                                                  const int *__restrict__ col,
                                                                  const int *__restrict__ upd_ptr,
                                                                  const int *__restrict__ upd, int *__restrict__ flags) {
  extern __shared__ double sm[];
  if (factor_row<B, false>(bsr, rows[blockIdx.x], row_ptr, diag_ptr, col, upd_ptr, upd, sm) && threadIdx.x == 0) {
    atomicOr(flags, kErrSingular);
  }
}

// ------------------------------------------------------------ sweeps
// Dot products of block row r with x, split over P = blockDim/b threads
// (columns c = part, part + P, ...).
template <class T>
__device__ __forceinline__ T row_dot(const T *__restrict__ a, const T *__restrict__ x, int b, int r, int part, int P) {
  T acc0 = T(0), acc1 = T(0);
  int c = part;
  for (; c + P < b; c += 2 * P) {
    acc0 += a[size_t(c) * b + r] * x[c];
    acc1 += a[size_t(c + P) * b + r] * x[c + P];
  }
  if (c < b) { acc0 += a[size_t(c) * b + r] * x[c]; }
  return acc0 + acc1;
}
constexpr int kSweepThreads = 256;
// Forward sweep of one level (src/block_ilu.cpp:87-99): w_i = b_i - sum_{j<i} L_ij w_j.
template <class T>
__global__ __launch_bounds__(kSweepThreads) void k_ilu_lower_split(const T *__restrict__ bsr, int b,
                                                                   const int *__restrict__ rows,
                                                                   const int *__restrict__ row_ptr,
                                                                   const int *__restrict__ diag_ptr,
                                                                   const int *__restrict__ col,
                                                                   const T *__restrict__ rhs, T *__restrict__ w) {
  __shared__ T red[kSweepThreads];
  const int i = rows[blockIdx.x];
  const size_t bb = size_t(b) * b;
  const int P = blockDim.x / b, r = threadIdx.x % b, part = threadIdx.x / b;
  T acc = T(0);
  if (part < P) {
    for (int ij = row_ptr[i]; ij < diag_ptr[i]; ++ij) { acc += row_dot(bsr + ij * bb, w + size_t(col[ij]) * b, b, r, part, P); }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (part == 0) {
    T s = T(0);
    for (int q = 0; q < P; ++q) { s += red[q * b + r]; }
    w[size_t(i) * b + r] = rhs[size_t(i) * b + r] - s;
  }
}
// Backward sweep of one level (src/block_ilu.cpp:100-108):
// w_i = inv(U_ii) (w_i - sum_{j>i} U_ij w_j).
template <class T>
__global__ __launch_bounds__(kSweepThreads) void k_ilu_upper_split(const T *__restrict__ bsr, int b,
                                                                   const int *__restrict__ rows,
                                                                   const int *__restrict__ row_ptr,
                                                                   const int *__restrict__ diag_ptr,
                                                                   const int *__restrict__ col, T *__restrict__ w) {
  __shared__ T red[kSweepThreads];
  __shared__ T ts[kSweepThreads];
  const int i = rows[blockIdx.x];
  const size_t bb = size_t(b) * b;
  const int P = blockDim.x / b, r = threadIdx.x % b, part = threadIdx.x / b;
  T acc = T(0);
  if (part < P) {
    for (int ij = diag_ptr[i] + 1; ij < row_ptr[i + 1]; ++ij) { acc += row_dot(bsr + ij * bb, w + size_t(col[ij]) * b, b, r, part, P); }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (part == 0) {
    T s = T(0);
    for (int q = 0; q < P; ++q) { s += red[q * b + r]; }
    ts[r] = w[size_t(i) * b + r] - s;
  }
  __syncthreads();
  acc = part < P ? row_dot(bsr + diag_ptr[i] * bb, ts, b, r, part, P) : T(0);
  red[threadIdx.x] = acc;
  __syncthreads();
  if (part == 0) {
    T s = T(0);
    for (int q = 0; q < P; ++q) { s += red[q * b + r]; }
    w[size_t(i) * b + r] = s;
  }
}

// ------------------------------------------------------------ dataflow forms
// One launch per sweep (or per factorisation) instead of one per level: the
// grid is co-resident (cooperative launch) and each workgroup handles its
// rows in dependency order (increasing for the lower sweep and the factor,
// decreasing for the upper sweep). Before a row it waits until every row it
// depends on has published done[j] >= epoch; after the row it completes its
// stores and publishes done[i] = epoch. As the rows waited on precede the
// waiting row in that order and all workgroups are resident, every wait
// ends; each wait is nevertheless bounded by spin_limit and then raises
// kErrSpin instead of hanging. Data crossing workgroups is read and written
// with coherent (sc1) accesses only.
__device__ __forceinline__ void wait_rows(const int *__restrict__ col, int j0, int j1, const int *done, int epoch,
                                          int spin_limit, int *flags) {
  for (int ij = j0 + int(threadIdx.x); ij < j1; ij += blockDim.x) {
    const int j = col[ij];
    int spins = 0;
    while (ld_coherent(done + j) < epoch) {
      if (++spins > spin_limit) {
        atomicOr(flags, kErrSpin);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}
__device__ __forceinline__ void publish_row(int *done, int i, int epoch) {
  complete_memory_ops();
  __syncthreads();
  if (threadIdx.x == 0) { st_coherent(done + i, epoch); }
}
// Gather the dependency vectors of row i (coherently, one element per thread)
// into LDS: xs[(ij - j0) * b + c] = w[col[ij] * b + c].
template <class T>
__device__ __forceinline__ void gather_deps(const T *w, const int *__restrict__ col, int j0, int j1, int b, T *xs) {
  const int total = (j1 - j0) * b;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int d = t / b, c = t % b;
    xs[t] = ld_coherent(w + size_t(col[j0 + d]) * b + c);
  }
  __syncthreads();
}
constexpr int kMaxDeps = 8;
template <class T, bool LOWER>
__global__ __launch_bounds__(kSweepThreads) void k_ilu_sweep_df(const T *__restrict__ bsr, int b, int nb,
                                                                const int *__restrict__ row_ptr,
                                                                const int *__restrict__ diag_ptr,
                                                                const int *__restrict__ col, const T *__restrict__ rhs,
                                                                T *w, int *done, int epoch, int *flags, int spin_limit) {
  __shared__ T red[kSweepThreads];
  __shared__ T ts[kSweepThreads];
  __shared__ T xs[kMaxDeps * kSweepThreads];
  const size_t bb = size_t(b) * b;
  const int P = blockDim.x / b, r = threadIdx.x % b, part = threadIdx.x / b;
  for (int idx = blockIdx.x; idx < nb; idx += gridDim.x) {
    const int i = LOWER ? idx : nb - 1 - idx;
    const int j0 = LOWER ? row_ptr[i] : diag_ptr[i] + 1, j1 = LOWER ? diag_ptr[i] : row_ptr[i + 1];
    wait_rows(col, j0, j1, done, epoch, spin_limit, flags);
    gather_deps(w, col, j0, j1, b, xs);
    T acc = T(0);
    if (part < P) {
      for (int ij = j0; ij < j1; ++ij) { acc += row_dot(bsr + ij * bb, xs + (ij - j0) * b, b, r, part, P); }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (part == 0) {
      T s = T(0);
      for (int q = 0; q < P; ++q) { s += red[q * b + r]; }
      if (LOWER) { st_coherent(w + size_t(i) * b + r, rhs[size_t(i) * b + r] - s); }
      else { ts[r] = w[size_t(i) * b + r] - s; }
    }
    if (!LOWER) {
      __syncthreads();
      acc = part < P ? row_dot(bsr + diag_ptr[i] * bb, ts, b, r, part, P) : T(0);
      red[threadIdx.x] = acc;
      __syncthreads();
      if (part == 0) {
        T s = T(0);
        for (int q = 0; q < P; ++q) { s += red[q * b + r]; }
        st_coherent(w + size_t(i) * b + r, s);
      }
    }
    publish_row(done, i, epoch);
  }
}
// Dataflow ILU(0) factor: factor_row per row, rows in increasing order per
// workgroup, waiting on the rows of the lower pattern.
template <int B>
__global__ __launch_bounds__(Shape<B>::NT) void k_ilu_factor_df(double *bsr, int nb, const int *__restrict__ row_ptr,
                                                                const int *__restrict__ diag_ptr,
                                                                const int *__restrict__ col,
                                                                const int *__restrict__ upd_ptr,
                                                                const int *__restrict__ upd, int *done, int epoch,
                                                                int *flags, int spin_limit) {
  extern __shared__ double sm[];
  for (int i = blockIdx.x; i < nb; i += gridDim.x) {
    wait_rows(col, row_ptr[i], diag_ptr[i], done, epoch, spin_limit, flags);
    if (factor_row<B, true>(bsr, i, row_ptr, diag_ptr, col, upd_ptr, upd, sm) && threadIdx.x == 0) {
      atomicOr(flags, kErrSingular);
    }
    publish_row(done, i, epoch);
  }
}

// Largest co-resident grid of a kernel (all CUs x blocks per CU).
int coresident(const void *fn, int threads, size_t lds) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return cus * per;
}
void allow_lds(const void *fn) {
  hipFuncAttributes attr{};
  if (hipFuncGetAttributes(&attr, fn) == hipSuccess) {
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 163840 - int(attr.sharedSizeBytes));
  }
  (void)hipGetLastError();
}
constexpr int kSpinLimit = 1 << 20;
} // namespace

#define EG_DISPATCH_B(b, CALL)                                                                     \
  switch (b) {                                                                                     \
  case 4: { constexpr int B = 4; CALL; } break;                                                    \
  case 8: { constexpr int B = 8; CALL; } break;                                                    \
  case 12: { constexpr int B = 12; CALL; } break;                                                  \
  case 16: { constexpr int B = 16; CALL; } break;                                                  \
  case 20: { constexpr int B = 20; CALL; } break;                                                  \
  case 24: { constexpr int B = 24; CALL; } break;                                                  \
  case 28: { constexpr int B = 28; CALL; } break;                                                  \
  case 32: { constexpr int B = 32; CALL; } break;                                                  \
  case 36: { constexpr int B = 36; CALL; } break;                                                  \
  case 64: { constexpr int B = 64; CALL; } break;                                                  \
  case 100: { constexpr int B = 100; CALL; } break;                                                \
  default: break;                                                                                  \
  }

bool fast_dense_ok(int b) {
  return (b % 4 == 0 && b >= 4 && b <= 36) || b == 64 || b == 100;
}
int ilu_factor_lds_init() {
#define EG_ALLOW(B)                                                                                \
  allow_lds(reinterpret_cast<const void *>(k_ilu_factor_rows<B>));                                 \
  allow_lds(reinterpret_cast<const void *>(k_ilu_factor_df<B>));
  EG_ALLOW(4) EG_ALLOW(8) EG_ALLOW(12) EG_ALLOW(16) EG_ALLOW(20) EG_ALLOW(24) EG_ALLOW(28) EG_ALLOW(32)
  EG_ALLOW(36) EG_ALLOW(64) EG_ALLOW(100)
#undef EG_ALLOW
  return 0;
}
void launch_block_inverse_fast(double *blocks, int b, int count, const int *index, int *flags, hipStream_t s) {
  EG_DISPATCH_B(b, (k_block_inverse_b<B><<<count, Shape<B>::NT, 0, s>>>(blocks, index, flags)));
}
void launch_ilu_factor_fast(double *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                            const int *col, const int *upd_ptr, const int *upd, int *flags, hipStream_t s) {
  EG_DISPATCH_B(b, (k_ilu_factor_rows<B><<<nrows, Shape<B>::NT, factor_lds<B>(), s>>>(bsr, rows, row_ptr, diag_ptr,
                                                                                       col, upd_ptr, upd, flags)));
}
template <class T>
void launch_ilu_lower_fast(const T *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                           const int *col, const T *rhs, T *w, hipStream_t s) {
  k_ilu_lower_split<T><<<nrows, kSweepThreads, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, rhs, w);
}
template <class T>
void launch_ilu_upper_fast(const T *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                           const int *col, T *w, hipStream_t s) {
  k_ilu_upper_split<T><<<nrows, kSweepThreads, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, w);
}
template <class T>
bool launch_ilu_sweep_df(bool lower, const T *bsr, int b, int nb, int max_deps, const int *row_ptr,
                         const int *diag_ptr, const int *col, const T *rhs, T *w, int *done, int epoch, int *flags,
                         hipStream_t s) {
  if (b > kSweepThreads || max_deps * b > kMaxDeps * kSweepThreads) { return false; }
  const void *fn = lower ? reinterpret_cast<const void *>(k_ilu_sweep_df<T, true>)
                         : reinterpret_cast<const void *>(k_ilu_sweep_df<T, false>);
  const int cap = coresident(fn, kSweepThreads, 0);
  if (cap <= 0) { return false; }
  const int grid = std::min(nb, cap);
  int spin = kSpinLimit;
  void *args[] = {&bsr, &b, &nb, &row_ptr, &diag_ptr, &col, &rhs, &w, &done, &epoch, &flags, &spin};
  if (hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kSweepThreads), args, 0, s) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}
bool launch_ilu_factor_df(double *bsr, int b, int nb, const int *row_ptr, const int *diag_ptr, const int *col,
                          const int *upd_ptr, const int *upd, int *done, int epoch, int *flags, hipStream_t s) {
  if (!fast_dense_ok(b)) { return false; }
  bool ok = false;
  EG_DISPATCH_B(b, ({
    const void *fn = reinterpret_cast<const void *>(k_ilu_factor_df<B>);
    const size_t lds = factor_lds<B>();
    const int cap = coresident(fn, Shape<B>::NT, lds);
    if (cap > 0) {
      const int grid = std::min(nb, cap);
      int spin = kSpinLimit;
      void *args[] = {&bsr, &nb, &row_ptr, &diag_ptr, &col, &upd_ptr, &upd, &done, &epoch, &flags, &spin};
      ok = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(Shape<B>::NT), args, lds, s) == hipSuccess;
      if (!ok) { (void)hipGetLastError(); }
    }
  }));
  return ok;
}

template void launch_ilu_lower_fast<double>(const double *, int, const int *, int, const int *, const int *,
                                            const int *, const double *, double *, hipStream_t);
template void launch_ilu_lower_fast<float>(const float *, int, const int *, int, const int *, const int *,
                                           const int *, const float *, float *, hipStream_t);
template void launch_ilu_upper_fast<double>(const double *, int, const int *, int, const int *, const int *,
                                            const int *, double *, hipStream_t);
template void launch_ilu_upper_fast<float>(const float *, int, const int *, int, const int *, const int *,
                                           const int *, float *, hipStream_t);
template bool launch_ilu_sweep_df<double>(bool, const double *, int, int, int, const int *, const int *, const int *,
                                          const double *, double *, int *, int, int *, hipStream_t);
template bool launch_ilu_sweep_df<float>(bool, const float *, int, int, int, const int *, const int *, const int *,
                                         const float *, float *, int *, int, int *, hipStream_t);
} // namespace eg
