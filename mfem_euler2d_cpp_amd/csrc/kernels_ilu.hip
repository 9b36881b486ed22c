// Dense block kernels tuned for the ILU(0)/block-Jacobi hot spots:
//  * Gauss-Jordan inverse with the matrix resident in registers and implicit
//    partial pivoting (no row swaps; two workgroup barriers per column),
//  * one-row-per-workgroup ILU(0) factor with LDS-staged, 4x4 register-tiled
//    block products,
//  * level sweeps whose block GEMVs split every row's dot product over
//    several threads (independent loads in flight, LDS reduction).
#include "kernels.hpp"

namespace eg {
namespace {
constexpr int kThreads = 256;

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

// In-place inverse of the b x b column-major block A (global memory) by one
// workgroup (b <= blockDim.x). Thread (tc, tr) = (tid % b, tid / b) owns
// column tc at rows tr, tr + R, ... (R = blockDim.x / b) in registers.
// Column k is eliminated with pivot row p = first unused row of maximal
// |S(p,k)|; the stored matrix S then satisfies A^{-1}(pinv[i], piv[j]) = S(i,j).
// The owners of column k publish it together with one pivot candidate each
// (their largest unused |entry|, computed in registers), so choosing the
// pivot costs R LDS reads instead of a column scan and wave reduction.
// lds: 4b + 2*64 doubles, then 2*64 + 2b ints. Returns true on a zero pivot.
constexpr int kGjLds = 4 * 128 + 2 * 64 + (2 * 64 + 2 * 128 + 1) / 2 + 1;
template <int NE, bool COH = false>
__device__ bool gj_inverse_regs(double *__restrict__ A, int b, double *lds) {
  const int R = blockDim.x / b;
  double *colk = lds, *rowp = lds + 2 * b, *cbest = lds + 4 * b;
  int *carg = reinterpret_cast<int *>(cbest + 2 * 64), *piv = carg + 2 * 64, *pinv = piv + b;
  const int tid = threadIdx.x;
  const int tc = tid % b, tr = tid / b;
  const bool active = tr < R;
  double v[NE];
  unsigned long long used = 0ull; // bit e: row tr + e*R already a pivot row
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int i = tr + e * R;
    v[e] = (active && i < b) ? ld<COH>(A + i + size_t(tc) * b) : 0.;
  }
  __syncthreads();
  bool singular = false;
  for (int k = 0; k < b; ++k) {
    double *ck = colk + (k & 1) * b, *rp = rowp + (k & 1) * b, *cb = cbest + (k & 1) * 64;
    int *ca = carg + (k & 1) * 64;
    if (active && tc == k) {
      double best = -1.;
      int arg = b;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const int i = tr + e * R;
        if (i < b) {
          ck[i] = v[e];
          const double a = fabs(v[e]);
          if (!((used >> e) & 1ull) && a > best) { best = a; arg = i; }
        }
      }
      cb[tr] = best;
      ca[tr] = arg;
    }
    __syncthreads();
    double best = -1.;
    int p = b;
    for (int g = 0; g < R; ++g) {
      const double ob = cb[g];
      const int oa = ca[g];
      if (ob > best || (ob == best && oa < p)) { best = ob; p = oa; }
    }
    const double ap = ck[p];
    if (ap == 0.) { singular = true; }
    const double inv = 1. / ap;
    if (active && tr == p % R) {
      const int ep = p / R;
      used |= 1ull << ep;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        if (e == ep) { rp[tc] = v[e]; }
      }
    }
    __syncthreads();
    if (tid == 0) { piv[k] = p; pinv[p] = k; }
    if (active) {
      const double rpj = rp[tc] * inv;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const int i = tr + e * R;
        if (i < b) {
          if (i == p) { v[e] = tc == k ? inv : v[e] * inv; }
          else if (tc == k) { v[e] = -ck[i] * inv; }
          else { v[e] = v[e] - ck[i] * rpj; }
        }
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int i = tr + e * R;
      if (i < b) { st<COH>(A + pinv[i] + size_t(piv[tc]) * b, v[e]); }
    }
  }
  __syncthreads();
  return singular;
}

template <int NE>
__global__ __launch_bounds__(kThreads) void k_block_inverse_regs(double *__restrict__ blocks, int b,
                                                                 const int *__restrict__ index, int *__restrict__ flags) {
  __shared__ double lds[kGjLds];
  const int blk = index ? index[blockIdx.x] : blockIdx.x;
  if (gj_inverse_regs<NE>(blocks + size_t(blk) * b * b, b, lds) && threadIdx.x == 0) {
    atomicOr(flags, kErrSingular);
  }
}

// C (-)= X Y for b x b blocks: X and Yt (Y transposed) in LDS, C in global.
// Each thread computes 4x4 tiles; mode 0 stores C = X Y (and into Xout in
// LDS when given), mode 1 subtracts from C.
template <bool COH = false>
__device__ void tile_product(const double *__restrict__ X, const double *__restrict__ Yt, double *__restrict__ C,
                             int b, int mode, double *__restrict__ Xout) {
  const int nt4 = b / 4, tiles = nt4 * nt4;
  for (int tile = threadIdx.x; tile < tiles; tile += blockDim.x) {
    const int r0 = (tile % nt4) * 4, c0 = (tile / nt4) * 4;
    double acc[4][4] = {};
    for (int kk = 0; kk < b; ++kk) {
      double a[4], y[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) { a[i] = X[r0 + i + kk * b]; y[i] = Yt[c0 + i + kk * b]; }
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
        const size_t idx = (r0 + i) + size_t(c0 + j) * b;
        if (mode == 0) {
          st<COH>(C + idx, acc[j][i]);
          if (Xout) { Xout[idx] = acc[j][i]; }
        } else {
          st<COH>(C + idx, ld<COH>(C + idx) - acc[j][i]);
        }
      }
    }
  }
}
// Stage a global column-major block into LDS, optionally transposed.
template <bool COH = false>
__device__ void stage(const double *__restrict__ src, double *__restrict__ dst, int b, bool transpose) {
  for (int t = threadIdx.x; t < b * b; t += blockDim.x) {
    const int r = t % b, c = t / b;
    const double v = ld<COH>(src + t);
    if (transpose) { dst[c + r * b] = v; }
    else { dst[t] = v; }
  }
}

// One BSR row of ILU(0) per workgroup (src/block_ilu.cpp:73-86), rows of one
// dependency level in parallel. L_ik = A_ik inv(A_kk) and A_ij -= L_ik A_kj
// as LDS-staged tile products; the updated diagonal block is replaced by its
// inverse. LDS: X (L_ik) and Y (second operand, transposed): 2 b^2 doubles.
template <int NE>
__global__ __launch_bounds__(kThreads) void k_ilu_factor_rows(double *__restrict__ bsr, int b,
                                                              const int *__restrict__ rows,
                                                              const int *__restrict__ row_ptr,
                                                              const int *__restrict__ diag_ptr,
                                                              const int *__restrict__ col,
                                                              const int *__restrict__ upd_ptr,
                                                              const int *__restrict__ upd, int *__restrict__ flags) {
  extern __shared__ double sm[];
  const size_t bb = size_t(b) * b;
  double *X = sm, *Y = sm + bb;
  const int i = rows[blockIdx.x];
  for (int ik = row_ptr[i]; ik < diag_ptr[i]; ++ik) {
    const int k = col[ik];
    double *aik = bsr + ik * bb;
    stage(aik, Y, b, false);                         // Y <- A_ik (untransposed: used as X operand)
    stage(bsr + diag_ptr[k] * bb, X, b, true);       // X <- inv(A_kk)^T
    __syncthreads();
    // L = A_ik * inv(A_kk): operands (Y, X^T) -> write to A_ik; keep L in Y.
    {
      const int nt4 = b / 4, tiles = nt4 * nt4;
      double *L = aik;
      for (int tile = threadIdx.x; tile < tiles; tile += blockDim.x) {
        const int r0 = (tile % nt4) * 4, c0 = (tile / nt4) * 4;
        double acc[4][4] = {};
        for (int kk = 0; kk < b; ++kk) {
          double a[4], y[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) { a[t] = Y[r0 + t + kk * b]; y[t] = X[c0 + t + kk * b]; }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int t = 0; t < 4; ++t) { acc[j][t] += a[t] * y[j]; }
          }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int t = 0; t < 4; ++t) { L[(r0 + t) + size_t(c0 + j) * b] = acc[j][t]; }
        }
      }
    }
    __syncthreads();
    stage(aik, X, b, false);  // X <- L_ik
    for (int u = upd_ptr[ik]; u < upd_ptr[ik + 1]; ++u) {
      __syncthreads();
      stage(bsr + size_t(upd[2 * u + 1]) * bb, Y, b, true);  // Y <- A_kj^T
      __syncthreads();
      tile_product(X, Y, bsr + size_t(upd[2 * u]) * bb, b, 1, nullptr);
    }
    __syncthreads();
  }
  __threadfence_block();
  if (gj_inverse_regs<NE>(bsr + diag_ptr[i] * bb, b, sm) && threadIdx.x == 0) { atomicOr(flags, kErrSingular); }
}

// Level sweeps: every row's dot products are split over P = blockDim/b
// threads (columns c = part, part + P, ...) and reduced through LDS in a
// fixed order.
template <class T>
__device__ __forceinline__ T row_dot(const T *__restrict__ a, const T *__restrict__ x, int b, int r, int part,
                                     int P) {
  T acc0 = T(0), acc1 = T(0);
  int c = part;
  for (; c + P < b; c += 2 * P) {
    acc0 += a[size_t(c) * b + r] * x[c];
    acc1 += a[size_t(c + P) * b + r] * x[c + P];
  }
  if (c < b) { acc0 += a[size_t(c) * b + r] * x[c]; }
  return acc0 + acc1;
}
// Forward sweep of one level (src/block_ilu.cpp:87-99): w_i = b_i - sum_{j<i} L_ij w_j.
template <class T>
__global__ __launch_bounds__(kThreads) void k_ilu_lower_split(const T *__restrict__ bsr, int b,
                                                              const int *__restrict__ rows,
                                                              const int *__restrict__ row_ptr,
                                                              const int *__restrict__ diag_ptr,
                                                              const int *__restrict__ col, const T *__restrict__ rhs,
                                                              T *__restrict__ w) {
  __shared__ T red[kThreads];
  const int i = rows[blockIdx.x];
  const size_t bb = size_t(b) * b;
  const int P = blockDim.x / b, r = threadIdx.x % b, part = threadIdx.x / b;
  T acc = T(0);
  if (part < P) {
    for (int ij = row_ptr[i]; ij < diag_ptr[i]; ++ij) {
      acc += row_dot(bsr + ij * bb, w + size_t(col[ij]) * b, b, r, part, P);
    }
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
__global__ __launch_bounds__(kThreads) void k_ilu_upper_split(const T *__restrict__ bsr, int b,
                                                              const int *__restrict__ rows,
                                                              const int *__restrict__ row_ptr,
                                                              const int *__restrict__ diag_ptr,
                                                              const int *__restrict__ col, T *__restrict__ w) {
  __shared__ T red[kThreads];
  __shared__ T ts[kThreads];
  const int i = rows[blockIdx.x];
  const size_t bb = size_t(b) * b;
  const int P = blockDim.x / b, r = threadIdx.x % b, part = threadIdx.x / b;
  T acc = T(0);
  if (part < P) {
    for (int ij = diag_ptr[i] + 1; ij < row_ptr[i + 1]; ++ij) {
      acc += row_dot(bsr + ij * bb, w + size_t(col[ij]) * b, b, r, part, P);
    }
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
// grid is co-resident (cooperative launch) and every workgroup handles its
// rows in dependency order (increasing for the lower sweep and the factor,
// decreasing for the upper sweep). Before a row it waits until each row it
// depends on has published done[j] >= epoch; after it, it releases its
// results and publishes done[i] = epoch. Since the rows a workgroup waits on
// precede its own row in that order and all workgroups are resident, every
// wait ends; a wait is nevertheless bounded by spin_limit and then raises
// kErrSpin instead of hanging.
//
// Cross-workgroup data (the sweep vector and the flags) is accessed only with
// relaxed agent-scope atomics, which the gfx950 memory model makes coherent
// across XCDs without L2 write-back/invalidate; the writer completes its
// stores (s_waitcnt) before the workgroup barrier that precedes the flag.
__device__ __forceinline__ void complete_memory_ops() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
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
  complete_memory_ops();
  __syncthreads();
}
__device__ __forceinline__ void publish_row(int *done, int i, int epoch) {
  complete_memory_ops();
  __syncthreads();
  if (threadIdx.x == 0) { st_coherent(done + i, epoch); }
}
// row_dot with the vector entries read coherently (written by other
// workgroups of the same launch).
template <class T>
__device__ __forceinline__ T row_dot_coherent(const T *__restrict__ a, const T *x, int b, int r, int part, int P) {
  T acc0 = T(0), acc1 = T(0);
  int c = part;
  for (; c + P < b; c += 2 * P) {
    acc0 += a[size_t(c) * b + r] * ld_coherent(x + c);
    acc1 += a[size_t(c + P) * b + r] * ld_coherent(x + c + P);
  }
  if (c < b) { acc0 += a[size_t(c) * b + r] * ld_coherent(x + c); }
  return acc0 + acc1;
}

template <class T, bool LOWER>
__global__ __launch_bounds__(kThreads) void k_ilu_sweep_df(const T *__restrict__ bsr, int b, int nb,
                                                           const int *__restrict__ row_ptr,
                                                           const int *__restrict__ diag_ptr,
                                                           const int *__restrict__ col, const T *__restrict__ rhs,
                                                           T *w, int *done, int epoch, int *flags, int spin_limit) {
  __shared__ T red[kThreads];
  __shared__ T ts[kThreads];
  const size_t bb = size_t(b) * b;
  const int P = blockDim.x / b, r = threadIdx.x % b, part = threadIdx.x / b;
  for (int idx = blockIdx.x; idx < nb; idx += gridDim.x) {
    const int i = LOWER ? idx : nb - 1 - idx;
    const int j0 = LOWER ? row_ptr[i] : diag_ptr[i] + 1, j1 = LOWER ? diag_ptr[i] : row_ptr[i + 1];
    wait_rows(col, j0, j1, done, epoch, spin_limit, flags);
    T acc = T(0);
    if (part < P) {
      for (int ij = j0; ij < j1; ++ij) {
        acc += row_dot_coherent(bsr + ij * bb, w + size_t(col[ij]) * b, b, r, part, P);
      }
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

// Dataflow ILU(0) factor: the body of k_ilu_factor_rows per row, rows in
// increasing order per workgroup, waiting on the rows of the lower pattern.
template <int NE>
__global__ __launch_bounds__(kThreads) void k_ilu_factor_df(double *bsr, int b, int nb, const int *__restrict__ row_ptr,
                                                            const int *__restrict__ diag_ptr,
                                                            const int *__restrict__ col, const int *__restrict__ upd_ptr,
                                                            const int *__restrict__ upd, int *done, int epoch,
                                                            int *flags, int spin_limit) {
  extern __shared__ double sm[];
  const size_t bb = size_t(b) * b;
  double *X = sm, *Y = sm + bb;
  for (int i = blockIdx.x; i < nb; i += gridDim.x) {
    wait_rows(col, row_ptr[i], diag_ptr[i], done, epoch, spin_limit, flags);
    for (int ik = row_ptr[i]; ik < diag_ptr[i]; ++ik) {
      const int k = col[ik];
      double *aik = bsr + ik * bb;
      stage<true>(aik, Y, b, false);
      stage<true>(bsr + diag_ptr[k] * bb, X, b, true);
      __syncthreads();
      {
        const int nt4 = b / 4, tiles = nt4 * nt4;
        for (int tile = threadIdx.x; tile < tiles; tile += blockDim.x) {
          const int r0 = (tile % nt4) * 4, c0 = (tile / nt4) * 4;
          double acc[4][4] = {};
          for (int kk = 0; kk < b; ++kk) {
            double a[4], y[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) { a[t] = Y[r0 + t + kk * b]; y[t] = X[c0 + t + kk * b]; }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
#pragma unroll
              for (int t = 0; t < 4; ++t) { acc[j][t] += a[t] * y[j]; }
            }
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int t = 0; t < 4; ++t) { st_coherent(aik + (r0 + t) + size_t(c0 + j) * b, acc[j][t]); }
          }
        }
      }
      __syncthreads();
      stage<true>(aik, X, b, false);
      for (int u = upd_ptr[ik]; u < upd_ptr[ik + 1]; ++u) {
        __syncthreads();
        stage<true>(bsr + size_t(upd[2 * u + 1]) * bb, Y, b, true);
        __syncthreads();
        tile_product<true>(X, Y, bsr + size_t(upd[2 * u]) * bb, b, 1, nullptr);
      }
      __syncthreads();
    }
    if (gj_inverse_regs<NE, true>(bsr + diag_ptr[i] * bb, b, sm) && threadIdx.x == 0) { atomicOr(flags, kErrSingular); }
    publish_row(done, i, epoch);
  }
}

int ne_for(int b) {
  const int R = kThreads / b, need = (b + R - 1) / R;
  int ne = 1;
  while (ne < need) { ne *= 2; }
  return ne;
}
} // namespace

#define EG_DISPATCH_NE(ne, CALL)                                                                   \
  switch (ne) {                                                                                    \
  case 1: { constexpr int NE = 1; CALL; } break;                                                   \
  case 2: { constexpr int NE = 2; CALL; } break;                                                   \
  case 4: { constexpr int NE = 4; CALL; } break;                                                   \
  case 8: { constexpr int NE = 8; CALL; } break;                                                   \
  case 16: { constexpr int NE = 16; CALL; } break;                                                 \
  case 32: { constexpr int NE = 32; CALL; } break;                                                 \
  default: { constexpr int NE = 64; CALL; } break;                                                 \
  }

bool fast_dense_ok(int b) { return b <= 128 && b % 4 == 0; }
void launch_block_inverse_fast(double *blocks, int b, int count, const int *index, int *flags, hipStream_t s) {
  EG_DISPATCH_NE(ne_for(b), (k_block_inverse_regs<NE><<<count, kThreads, 0, s>>>(blocks, b, index, flags)));
}
int ilu_factor_lds_init() {
  int bad = 0;
#define EG_SET(NE)                                                                                  \
  {                                                                                                \
    hipFuncAttributes attr{};                                                                      \
    const void *fn = reinterpret_cast<const void *>(k_ilu_factor_rows<NE>);                        \
    if (hipFuncGetAttributes(&attr, fn) != hipSuccess ||                                           \
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,                        \
                            163840 - int(attr.sharedSizeBytes)) != hipSuccess) {                   \
      bad = 1;                                                                                     \
    }                                                                                              \
  }
  EG_SET(1) EG_SET(2) EG_SET(4) EG_SET(8) EG_SET(16) EG_SET(32) EG_SET(64)
#undef EG_SET
  (void)hipGetLastError();
  return bad;
}
void launch_ilu_factor_fast(double *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                            const int *col, const int *upd_ptr, const int *upd, int *flags, hipStream_t s) {
  const size_t lds = std::max<size_t>(2 * size_t(b) * b * 8, size_t(kGjLds) * 8);
  EG_DISPATCH_NE(ne_for(b), (k_ilu_factor_rows<NE><<<nrows, kThreads, lds, s>>>(bsr, b, rows, row_ptr, diag_ptr, col,
                                                                                 upd_ptr, upd, flags)));
}
template <class T>
void launch_ilu_lower_fast(const T *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                           const int *col, const T *rhs, T *w, hipStream_t s) {
  k_ilu_lower_split<T><<<nrows, kThreads, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, rhs, w);
}
template <class T>
void launch_ilu_upper_fast(const T *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                           const int *col, T *w, hipStream_t s) {
  k_ilu_upper_split<T><<<nrows, kThreads, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, w);
}
// Largest co-resident grid of a kernel (all CUs x blocks per CU).
static int coresident(const void *fn, size_t lds) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, kThreads, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return cus * per;
}
constexpr int kSpinLimit = 1 << 20;

template <class T>
bool launch_ilu_sweep_df(bool lower, const T *bsr, int b, int nb, const int *row_ptr, const int *diag_ptr,
                         const int *col, const T *rhs, T *w, int *done, int epoch, int *flags, hipStream_t s) {
  if (b > kThreads) { return false; }
  const void *fn = lower ? reinterpret_cast<const void *>(k_ilu_sweep_df<T, true>)
                         : reinterpret_cast<const void *>(k_ilu_sweep_df<T, false>);
  static int cap_l = -1, cap_u = -1;
  int &cap = lower ? cap_l : cap_u;
  if (cap < 0) { cap = coresident(fn, 0); }
  if (cap <= 0) { return false; }
  const int grid = std::min(nb, cap);
  int spin = kSpinLimit;
  void *args[] = {&bsr, &b, &nb, &row_ptr, &diag_ptr, &col, &rhs, &w, &done, &epoch, &flags, &spin};
  if (hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kThreads), args, 0, s) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}
template bool launch_ilu_sweep_df<double>(bool, const double *, int, int, const int *, const int *, const int *,
                                          const double *, double *, int *, int, int *, hipStream_t);
template bool launch_ilu_sweep_df<float>(bool, const float *, int, int, const int *, const int *, const int *,
                                         const float *, float *, int *, int, int *, hipStream_t);

bool launch_ilu_factor_df(double *bsr, int b, int nb, const int *row_ptr, const int *diag_ptr, const int *col,
                          const int *upd_ptr, const int *upd, int *done, int epoch, int *flags, hipStream_t s) {
  if (!fast_dense_ok(b) || b > 100) { return false; }
  const size_t lds = std::max<size_t>(2 * size_t(b) * b * 8, size_t(kGjLds) * 8);
  const void *fn = nullptr;
  EG_DISPATCH_NE(ne_for(b), (fn = reinterpret_cast<const void *>(k_ilu_factor_df<NE>)));
  static int lds_set[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int slot = 0;
  for (int ne = ne_for(b); ne > 1; ne >>= 1) { ++slot; }
  if (!lds_set[slot]) {
    hipFuncAttributes attr{};
    if (hipFuncGetAttributes(&attr, fn) == hipSuccess) {
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 163840 - int(attr.sharedSizeBytes));
    }
    (void)hipGetLastError();
    lds_set[slot] = 1;
  }
  const int cap = coresident(fn, lds);
  if (cap <= 0) { return false; }
  const int grid = std::min(nb, cap);
  int spin = kSpinLimit;
  void *args[] = {&bsr, &b, &nb, &row_ptr, &diag_ptr, &col, &upd_ptr, &upd, &done, &epoch, &flags, &spin};
  if (hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kThreads), args, lds, s) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

template void launch_ilu_lower_fast<double>(const double *, int, const int *, int, const int *, const int *,
                                            const int *, const double *, double *, hipStream_t);
template void launch_ilu_lower_fast<float>(const float *, int, const int *, int, const int *, const int *,
                                           const int *, const float *, float *, hipStream_t);
template void launch_ilu_upper_fast<double>(const double *, int, const int *, int, const int *, const int *,
                                            const int *, double *, hipStream_t);
template void launch_ilu_upper_fast<float>(const float *, int, const int *, int, const int *, const int *,
                                           const int *, float *, hipStream_t);
} // namespace eg
