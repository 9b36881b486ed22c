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

// Gauss-Jordan thread layout for a B x B block: column c is owned by R
// consecutive lanes of one wavefront (R a power of two <= 16, so the group
// is one aligned DPP row or a part of one), lane r holding rows r, r + R, ...
template <int B> struct Shape {
  static_assert(B >= 1 && B <= 511, "pivot keys hold the row index in 9 bits");
  static constexpr int R = B <= 4 ? 4 : (B <= 8 ? 8 : (16 * B <= 1024 ? 16 : (8 * B <= 1024 ? 8 : 4))); // lanes per column
  static constexpr int NT = (B * R + 63) / 64 * 64;                                // threads per workgroup
  static constexpr int NE = (B + R - 1) / R;                                       // register entries per thread
  static_assert(NE <= 64 && NT <= 1024, "one bit per register entry, one workgroup");
};
template <int B> struct GjLds {
  double col[2][B];              // pivot column, double buffered
  double pval[2];                // pivot value
  unsigned long long key[2];     // winning pivot key
  unsigned long long wkey[2][16]; // per-wave keys of the workgroup-wide search
  double scale[B];               // row scale: true row = scale * stored row
  int piv[B], pinv[B];           // step -> pivot row, pivot row -> step
  int used[B];                   // row already a pivot row (workgroup-wide search)
};
template <int B> constexpr int gj_lds() { return int((sizeof(GjLds<B>) + 7) / 8); } // in doubles

// Max of a 64-bit key over an aligned group of R lanes (R = 4, 8 or 16)
// with DPP row permutations: xor 1, xor 2, half-row mirror, row mirror.
template <int CTRL> __device__ __forceinline__ unsigned long long dpp_max(unsigned long long k) {
  const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(k)), CTRL, 0xF, 0xF, false));
  const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(k >> 32)), CTRL, 0xF, 0xF, false));
  const unsigned long long o = (static_cast<unsigned long long>(hi) << 32) | lo;
  return o > k ? o : k;
}
template <int R> __device__ __forceinline__ unsigned long long group_max(unsigned long long k) {
  k = dpp_max<0xB1>(k); // quad_perm [1,0,3,2]
  k = dpp_max<0x4E>(k); // quad_perm [2,3,0,1]
  if (R >= 8) { k = dpp_max<0x141>(k); }
  if (R >= 16) { k = dpp_max<0x140>(k); }
  return k;
}
// Double from lane `src` of the wavefront.
__device__ __forceinline__ double lane_value(double v, int src) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = unsigned(__builtin_amdgcn_ds_bpermute(src * 4, int(unsigned(u))));
  const unsigned hi = unsigned(__builtin_amdgcn_ds_bpermute(src * 4, int(unsigned(u >> 32))));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Pivot key of row i: |value| bits with the low 9 replaced by 511 - i, so a
// larger key is a larger |value| and near-ties go to the lowest row (a NaN
// ranks high and is reported through its pivot value).
__device__ __forceinline__ unsigned long long pivot_key(double v, int i) {
  return (static_cast<unsigned long long>(__double_as_longlong(v)) & 0x7FFFFFFFFFFFFE00ull) | unsigned(511 - i);
}
// Max over the 64 lanes of a wavefront: DPP within rows of 16, then xor
// shuffles across rows.
__device__ __forceinline__ unsigned long long wave_max(unsigned long long k) {
  k = group_max<16>(k);
  for (int m = 16; m < 64; m <<= 1) {
    const unsigned lo = unsigned(__shfl_xor(int(unsigned(k)), m, 64));
    const unsigned hi = unsigned(__shfl_xor(int(unsigned(k >> 32)), m, 64));
    const unsigned long long o = (static_cast<unsigned long long>(hi) << 32) | lo;
    k = o > k ? o : k;
  }
  return k;
}

// In-place inverse of the B x B column-major block A by one workgroup of
// Shape<B>::NT threads; thread (tc, tr) = (tid / R, tid % R) keeps column tc
// at rows tr + e R in registers. Step k eliminates column k with pivot row
// p = an unused row of maximal |S(p,k)| (pivot_key). For up to 16 entries a
// lane, the column's own lane group reduces the keys in registers (DPP), one
// barrier a step; above that, the column is written to LDS and the whole
// workgroup reduces one key a thread, two barriers a step. The pivot row's
// value for each column comes from a lane of the same group (ds_bpermute).
// Rows are stored unscaled: every entry, the pivot row's included, takes
// S(i,j) -= S(i,k) S(p,j) / S(p,k), after which the pivot row's owners put
// back its old values and record scale[p] = 1 / S(p,k); column k is replaced
// by -S(i,k) / S(p,k), and 1 at row p. With the scales applied at the end,
// A^{-1}(pinv[i], piv[j]) = scale[i] S(i,j). Register indices that depend on
// the pivot are resolved with uniform branches, not per-entry selects.
// Mathematically the inverse of the block whose LU the reference uses
// (src/linear.cpp:198-202, src/block_ilu.cpp:83-84, src/condensation.cpp:137-140).
// Returns true on a zero or non-finite pivot.
template <int B, bool COH>
__device__ __forceinline__ bool gj_inverse(double *__restrict__ A, double *lds) {
  constexpr int R = Shape<B>::R, NE = Shape<B>::NE, NT = Shape<B>::NT;
  constexpr bool WIDE = NE > 16;
  static_assert(!WIDE || (NT / 64 <= 16 && B <= NT), "one key a thread, at most 16 waves");
  GjLds<B> &L = *reinterpret_cast<GjLds<B> *>(lds);
  const int tid = threadIdx.x, tc = tid / R, tr = tid % R;
  const bool active = tc < B;
  const int group = (tid & 63) & ~(R - 1); // first lane of this column's group
  double v[NE];
  unsigned long long used = 0ull; // bit e: row tr + e R is a pivot row already
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int i = tr + e * R;
    v[e] = (active && i < B) ? ld<COH>(A + i + tc * B) : 0.;
  }
  if (WIDE) {
    for (int i = tid; i < B; i += blockDim.x) { L.used[i] = 0; }
  }
  __syncthreads();
  bool singular = false;
  for (int k = 0; k < B; ++k) {
    const int buf = k & 1;
    double *ck = L.col[buf];
    int p;
    double ap;
    if (!WIDE) {
      if (tc == k) {
        unsigned long long key = 0ull;
        double bv = 0.;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
          const int i = tr + e * R;
          if (i < B) {
            ck[i] = v[e];
            const unsigned long long kk = pivot_key(v[e], i);
            if (!((used >> e) & 1ull) && kk > key) { key = kk; bv = v[e]; }
          }
        }
        const unsigned long long best = group_max<R>(key);
        if (key == best) {
          L.key[buf] = best;
          L.pval[buf] = bv;
        }
      }
      __syncthreads();
      p = 511 - int(L.key[buf] & 511ull);
      ap = L.pval[buf];
    } else {
      if (tc == k) {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
          const int i = tr + e * R;
          if (i < B) { ck[i] = v[e]; }
        }
      }
      __syncthreads();
      const unsigned long long key = wave_max((tid < B && !L.used[tid]) ? pivot_key(ck[tid], tid) : 0ull);
      if ((tid & 63) == 0) { L.wkey[buf][tid >> 6] = key; }
      __syncthreads();
      unsigned long long best = 0ull;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) { best = L.wkey[buf][w] > best ? L.wkey[buf][w] : best; }
      p = 511 - int(best & 511ull);
      ap = ck[p];
    }
    if (!(fabs(ap) > 0.) || !isfinite(ap)) { singular = true; }
    const int ep = p / R, lp = p % R;
    const bool mine = tr == lp;
    double sel = 0.;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      if (e == ep) {
        asm volatile(""); // uniform branch, no per-entry select
        sel = v[e];
      }
    }
    if (mine) { used |= 1ull << ep; }
    const double rpv = lane_value(sel, group + lp);
    const double inv = 1. / ap;
    if (tid == 0) {
      L.piv[k] = p;
      L.pinv[p] = k;
      L.scale[p] = inv;
      if (WIDE) { L.used[p] = 1; }
    }
    const double rpj = rpv * inv;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int i = tr + e * R;
      v[e] = fma(-ck[i < B ? i : B - 1], rpj, v[e]);
    }
    if (tc == k) {
      asm volatile(""); // keep this a branch taken by one lane group only
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const int i = tr + e * R;
        v[e] = -ck[i < B ? i : B - 1] * inv;
      }
    }
    const double keep = tc == k ? 1. : sel;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      if (e == ep) {
        asm volatile(""); // uniform branch: restore the pivot row
        v[e] = mine ? keep : v[e];
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int i = tr + e * R;
      if (i < B) { st<COH>(A + L.pinv[i] + L.piv[tc] * B, v[e] * L.scale[i]); }
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
// (columns c = part, part + P, ...), four independent loads in flight.
template <class T>
__device__ __forceinline__ T row_dot(const T *__restrict__ a, const T *__restrict__ x, int b, int r, int part, int P) {
  T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  int c = part;
  for (; c + 3 * P < b; c += 4 * P) {
    a0 += a[size_t(c) * b + r] * x[c];
    a1 += a[size_t(c + P) * b + r] * x[c + P];
    a2 += a[size_t(c + 2 * P) * b + r] * x[c + 2 * P];
    a3 += a[size_t(c + 3 * P) * b + r] * x[c + 3 * P];
  }
  for (; c < b; c += P) { a0 += a[size_t(c) * b + r] * x[c]; }
  return (a0 + a1) + (a2 + a3);
}
constexpr int kSweepMaxBlock = 511;
// Forward sweep of one level (src/block_ilu.cpp:87-99): w_i = b_i - sum_{j<i} L_ij w_j.
// NT threads (256 up to 64 x 64 blocks, 1024 above: more loads in flight).
template <class T, int NT>
__global__ __launch_bounds__(NT) void k_ilu_lower_split(const T *__restrict__ bsr, int b, const int *__restrict__ rows,
                                                       const int *__restrict__ row_ptr,
                                                       const int *__restrict__ diag_ptr, const int *__restrict__ col,
                                                       const T *__restrict__ rhs, T *__restrict__ w) {
  __shared__ T red[NT];
  const int i = rows[blockIdx.x];
  const size_t bb = size_t(b) * b;
  const int P = NT / b, r = threadIdx.x % b, part = threadIdx.x / b;
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
template <class T, int NT>
__global__ __launch_bounds__(NT) void k_ilu_upper_split(const T *__restrict__ bsr, int b, const int *__restrict__ rows,
                                                       const int *__restrict__ row_ptr,
                                                       const int *__restrict__ diag_ptr, const int *__restrict__ col,
                                                       T *__restrict__ w) {
  __shared__ T red[NT];
  __shared__ T ts[kSweepMaxBlock];
  const int i = rows[blockIdx.x];
  const size_t bb = size_t(b) * b;
  const int P = NT / b, r = threadIdx.x % b, part = threadIdx.x / b;
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
// decreasing for the upper sweep). As the rows waited on precede the waiting
// row in that order and all workgroups are resident, every wait ends; each
// wait is nevertheless bounded by spin_limit and then raises kErrSpin
// instead of hanging. Data crossing workgroups is read and written with
// coherent (sc1) accesses only.
//
// Factor: before a row the workgroup waits until every row it depends on has
// published done[j] >= epoch; after the row it completes its stores and
// publishes done[i] = epoch.
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

// Sweeps: the output vector is armed with a signalling-NaN sentinel before
// the sweep (arithmetic never produces a signalling NaN), and a row polls the
// values of the rows it depends on until none is the sentinel any more: the
// data is its own completion flag, one coherent round trip per dependency.
// The row's matrix blocks are loaded into registers before the wait.
template <class T> struct Sentinel;
template <> struct Sentinel<double> {
  using U = unsigned long long;
  static constexpr U bits = 0x7FF4DEADBEEF0001ull;
};
template <> struct Sentinel<float> {
  using U = unsigned;
  static constexpr U bits = 0x7FA0DEADu;
};
template <class T> __device__ __forceinline__ T sentinel() {
  return __builtin_bit_cast(T, Sentinel<T>::bits);
}
template <class T> __device__ __forceinline__ bool armed(T v) {
  return __builtin_bit_cast(typename Sentinel<T>::U, v) == Sentinel<T>::bits;
}
// Threads per sweep workgroup and the split of a block row over them:
// thread (r, part) = (tid % B, tid / B) holds columns part + v P of row r of
// every block of its row (VB values per block).
template <int B> struct SweepShape {
  static constexpr int NT = B > 64 ? 1024 : 256;
  static constexpr int P = NT / B;
  static constexpr int VB = (B + P - 1) / P;
};
// LOWER: w_i = b_i - sum_{j<i} L_ij w_j (src/block_ilu.cpp:87-99); src = b,
// dst = w, and the kernel arms `arm` (the upper sweep's output).
// UPPER: y_i = inv(U_ii) (w_i - sum_{j>i} U_ij y_j) (src/block_ilu.cpp:100-108);
// src = w, re-armed row by row after it is read, dst = y.
template <class T, int B, int KD, bool LOWER>
__global__ __launch_bounds__(SweepShape<B>::NT) void k_ilu_sweep_sn(const T *__restrict__ bsr, int nb,
                                                                    const int *__restrict__ row_ptr,
                                                                    const int *__restrict__ diag_ptr,
                                                                    const int *__restrict__ col, T *src, T *dst,
                                                                    T *arm, int *flags, int spin_limit) {
  constexpr int NT = SweepShape<B>::NT, P = SweepShape<B>::P, VB = SweepShape<B>::VB;
  constexpr int KB = LOWER ? KD : KD + 1; // blocks held per row (upper: + inv(U_ii))
  constexpr size_t bb = size_t(B) * B;
  __shared__ T xs[KD * B];
  __shared__ T red[NT];
  __shared__ T ts[B];
  const int tid = threadIdx.x, r = tid % B, part = tid / B;
  const bool worker = part < P;
  if (LOWER) {
    for (size_t t = size_t(blockIdx.x) * NT + tid; t < size_t(nb) * B; t += size_t(gridDim.x) * NT) {
      arm[t] = sentinel<T>();
    }
  }
  for (int idx = blockIdx.x; idx < nb; idx += gridDim.x) {
    const int i = LOWER ? idx : nb - 1 - idx;
    const int j0 = LOWER ? row_ptr[i] : diag_ptr[i] + 1, j1 = LOWER ? diag_ptr[i] : row_ptr[i + 1];
    const int nd = j1 - j0;
    T a[KB][VB];
#pragma unroll
    for (int d = 0; d < KB; ++d) {
      const int slot = (!LOWER && d == KD) ? diag_ptr[i] : j0 + d;
      const bool use = (!LOWER && d == KD) || d < nd;
#pragma unroll
      for (int v = 0; v < VB; ++v) {
        const int c = part + v * P;
        a[d][v] = (use && worker && c < B) ? bsr[size_t(slot) * bb + size_t(c) * B + r] : T(0);
      }
    }
    T own = T(0);
    if (part == 0) {
      own = src[size_t(i) * B + r];
      if (!LOWER) { src[size_t(i) * B + r] = sentinel<T>(); }
    }
    for (int t = tid; t < nd * B; t += NT) {
      const T *p = dst + size_t(col[j0 + t / B]) * B + t % B;
      T x = ld_coherent(p);
      int spins = 0;
      while (armed(x)) {
        if (++spins > spin_limit) {
          atomicOr(flags, kErrSpin);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        x = ld_coherent(p);
      }
      xs[t] = x;
    }
    __syncthreads();
    T acc = T(0);
#pragma unroll
    for (int d = 0; d < KD; ++d) {
      if (d < nd) {
#pragma unroll
        for (int v = 0; v < VB; ++v) {
          const int c = part + v * P;
          if (c < B) { acc += a[d][v] * xs[d * B + c]; }
        }
      }
    }
    red[tid] = acc;
    __syncthreads();
    if (part == 0) {
      T sum = T(0);
      for (int q = 0; q < P; ++q) { sum += red[q * B + r]; }
      if (LOWER) { st_coherent(dst + size_t(i) * B + r, own - sum); }
      else { ts[r] = own - sum; }
    }
    if (!LOWER) {
      __syncthreads();
      acc = T(0);
#pragma unroll
      for (int v = 0; v < VB; ++v) {
        const int c = part + v * P;
        if (c < B) { acc += a[KD][v] * ts[c]; }
      }
      red[tid] = acc;
      __syncthreads();
      if (part == 0) {
        T sum = T(0);
        for (int q = 0; q < P; ++q) { sum += red[q * B + r]; }
        st_coherent(dst + size_t(i) * B + r, sum);
      }
    }
    __syncthreads();
  }
}
template <class T> __global__ void k_arm(T *v, size_t n) {
  for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += size_t(gridDim.x) * blockDim.x) {
    v[t] = sentinel<T>();
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

#define EG_DISPATCH_B_SMALL(b, CALL)                                                               \
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
  default: break;                                                                                  \
  }
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

#define EG_DISPATCH_GJ(b, CALL)                                                                    \
  switch (b) {                                                                                     \
  case 144: { constexpr int B = 144; CALL; } break;                                                \
  case 196: { constexpr int B = 196; CALL; } break;                                                \
  default: EG_DISPATCH_B(b, CALL) break;                                                           \
  }

bool gj_regs_ok(int b) { return fast_dense_ok(b) || b == 144 || b == 196; }
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
  EG_DISPATCH_GJ(b, (k_block_inverse_b<B><<<count, Shape<B>::NT, 0, s>>>(blocks, index, flags)));
}
void launch_ilu_factor_fast(double *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                            const int *col, const int *upd_ptr, const int *upd, int *flags, hipStream_t s) {
  EG_DISPATCH_B(b, (k_ilu_factor_rows<B><<<nrows, Shape<B>::NT, factor_lds<B>(), s>>>(bsr, rows, row_ptr, diag_ptr,
                                                                                       col, upd_ptr, upd, flags)));
}
// threads: 256 or 1024 per row (0: 256 up to 64 x 64 blocks, else 1024).
template <class T>
void launch_ilu_lower_level(const T *bsr, int b, const int *rows, int nrows, const int *row_ptr,
                            const int *diag_ptr, const int *col, const T *rhs, T *w, int threads, hipStream_t s) {
  if (threads == 0) { threads = b <= 64 ? 256 : 1024; }
  if (threads == 256 && b <= 256) { k_ilu_lower_split<T, 256><<<nrows, 256, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, rhs, w); }
  else { k_ilu_lower_split<T, 1024><<<nrows, 1024, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, rhs, w); }
}
template <class T>
void launch_ilu_upper_level(const T *bsr, int b, const int *rows, int nrows, const int *row_ptr,
                            const int *diag_ptr, const int *col, T *w, int threads, hipStream_t s) {
  if (threads == 0) { threads = b <= 64 ? 256 : 1024; }
  if (threads == 256 && b <= 256) { k_ilu_upper_split<T, 256><<<nrows, 256, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, w); }
  else { k_ilu_upper_split<T, 1024><<<nrows, 1024, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, w); }
}
template <class T>
bool launch_ilu_sweep_df(bool lower, const T *bsr, int b, int nb, int max_deps, const int *row_ptr,
                         const int *diag_ptr, const int *col, T *src, T *dst, T *arm, int *flags, hipStream_t s) {
  if (max_deps > 8) { return false; }
  const int kd = max_deps <= 4 ? 4 : 8;
  bool ok = false, known = false;
  auto go = [&](const void *fn, int threads) {
    known = true;
    const int cap = coresident(fn, threads, 0);
    if (cap <= 0) { return; }
    const int grid = std::min(nb, cap);
    int spin = kSpinLimit;
    void *args[] = {&bsr, &nb, &row_ptr, &diag_ptr, &col, &src, &dst, &arm, &flags, &spin};
    ok = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(threads), args, 0, s) == hipSuccess;
    if (!ok) { (void)hipGetLastError(); }
  };
#define EG_SWEEP(DISPATCH, KD)                                                                     \
  DISPATCH(b, (go(lower ? reinterpret_cast<const void *>(k_ilu_sweep_sn<T, B, KD, true>)           \
                        : reinterpret_cast<const void *>(k_ilu_sweep_sn<T, B, KD, false>),         \
                  SweepShape<B>::NT)))
  // Cell blocks (DG, at most 4 lower or upper neighbours) or face blocks
  // (HDG, at most 7).
  if (kd == 4) { EG_SWEEP(EG_DISPATCH_B, 4) } else { EG_SWEEP(EG_DISPATCH_B_SMALL, 8) }
#undef EG_SWEEP
  return known && ok;
}
template <class T> void launch_arm(T *v, size_t n, hipStream_t s) {
  k_arm<T><<<int(std::min<size_t>((n + 255) / 256, 4096)), 256, 0, s>>>(v, n);
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

template void launch_ilu_lower_level<double>(const double *, int, const int *, int, const int *, const int *,
                                             const int *, const double *, double *, int, hipStream_t);
template void launch_ilu_lower_level<float>(const float *, int, const int *, int, const int *, const int *,
                                            const int *, const float *, float *, int, hipStream_t);
template void launch_ilu_upper_level<double>(const double *, int, const int *, int, const int *, const int *,
                                             const int *, double *, int, hipStream_t);
template void launch_ilu_upper_level<float>(const float *, int, const int *, int, const int *, const int *,
                                            const int *, float *, int, hipStream_t);
template bool launch_ilu_sweep_df<double>(bool, const double *, int, int, int, const int *, const int *, const int *,
                                          double *, double *, double *, int *, hipStream_t);
template bool launch_ilu_sweep_df<float>(bool, const float *, int, int, int, const int *, const int *, const int *,
                                         float *, float *, float *, int *, hipStream_t);
template void launch_arm<double>(double *, size_t, hipStream_t);
template void launch_arm<float>(float *, size_t, hipStream_t);
} // namespace eg
