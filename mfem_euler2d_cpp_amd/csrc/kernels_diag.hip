// Kernels of the reference's assembled and direct diagnostics
// (src/operators.cpp:180-268, src/partial.cpp:174-258, src/linear.cpp:46-56,
// 204-213,224-230, src/block_ilu.cpp:10-59,87-109): the assembled block-sparse
// operator, its dense form and LU with partial pivoting (the dense / direct
// solvers), and the block permutation of a reordered ILU(0).
#include <algorithm>

#include "kernels.hpp"

namespace eg {
namespace {
constexpr int kDiagThreads = 256;
int diag_grid(long n) { return int(std::min<long>((n + kDiagThreads - 1) / kDiagThreads, 1L << 20)); }

// y = A x for a BSR matrix of b x b column-major blocks: one thread an output
// entry, its row's blocks in column order (a fixed summation order).
template <class T>
__global__ __launch_bounds__(kDiagThreads) void k_bsr_mult(const T *__restrict__ bsr, int b, int nb,
                                                           const int *__restrict__ row_ptr,
                                                           const int *__restrict__ col, const T *__restrict__ x,
                                                           T *__restrict__ y) {
  const long N = long(nb) * b;
  for (long id = long(blockIdx.x) * blockDim.x + threadIdx.x; id < N; id += long(gridDim.x) * blockDim.x) {
    const int i = int(id / b), r = int(id % b);
    T acc = T(0);
    for (int jb = row_ptr[i]; jb < row_ptr[i + 1]; ++jb) {
      const T *blk = bsr + size_t(jb) * b * b + r;
      const T *xv = x + size_t(col[jb]) * b;
      for (int c = 0; c < b; ++c) { acc += blk[size_t(c) * b] * xv[c]; }
    }
    y[id] = acc;
  }
}
// Dense column-major N x N form of a BSR matrix (zeros elsewhere).
__global__ __launch_bounds__(kDiagThreads) void k_fill_zero(double *__restrict__ a, size_t n) {
  for (size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x; id < n; id += size_t(gridDim.x) * blockDim.x) {
    a[id] = 0.;
  }
}
__global__ __launch_bounds__(kDiagThreads) void k_fill_zero_int(int *__restrict__ a, size_t n) {
  for (size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x; id < n; id += size_t(gridDim.x) * blockDim.x) {
    a[id] = 0;
  }
}
__global__ __launch_bounds__(kDiagThreads) void k_bsr_to_dense(const double *__restrict__ bsr, int b, int nb,
                                                               const int *__restrict__ row_ptr,
                                                               const int *__restrict__ col, double *__restrict__ a) {
  const long N = long(nb) * b;
  // one thread a (block row i, entry (r, c)) over every block of the row
  for (long id = long(blockIdx.x) * blockDim.x + threadIdx.x; id < N * b; id += long(gridDim.x) * blockDim.x) {
    const int i = int(id / (long(b) * b)), e = int(id % (long(b) * b)), r = e % b, c = e / b;
    for (int jb = row_ptr[i]; jb < row_ptr[i + 1]; ++jb) {
      a[size_t(long(col[jb]) * b + c) * N + size_t(i) * b + r] = bsr[size_t(jb) * b * b + e];
    }
  }
}

// LU with partial pivoting of a column-major N x N matrix, without row
// interchanges (rows keep their place; piv[k] is the pivot row of column k).
// Step k, pivot search (one workgroup): the unused row of largest |A(i,k)|
// (lowest row on ties) becomes piv[k]; mult(i) = A(i,k) / A(p,k) for the
// unused rows.
constexpr int kLuThreads = 1024;
__global__ __launch_bounds__(kLuThreads) void k_lu_pivot(double *__restrict__ a, int N, int k,
                                                         int *__restrict__ used, int *__restrict__ piv,
                                                         double *__restrict__ mult, int *__restrict__ flags) {
  __shared__ double best_v[kLuThreads / 64];
  __shared__ int best_i[kLuThreads / 64];
  __shared__ int pivot;
  const double *ak = a + size_t(k) * N;
  double bv = -1.;
  int bi = N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    if (!used[i]) {
      const double v = fabs(ak[i]);
      const double key = v != v ? INFINITY : v;
      if (key > bv) { bv = key; bi = i; } // i ascends per thread: the first of equal keys is kept
    }
  }
  // wavefront then workgroup arg-max: larger key, then lower row
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_down(bv, off, 64);
    const int oi = __shfl_down(bi, off, 64);
    if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { best_v[wid] = bv; best_i[wid] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = best_v[0];
    int p = best_i[0];
    for (int w = 1; w < int(blockDim.x >> 6); ++w) {
      if (best_v[w] > v || (best_v[w] == v && best_i[w] < p)) { v = best_v[w]; p = best_i[w]; }
    }
    if (p >= N) { p = 0; }
    pivot = p;
    piv[k] = p;
    used[p] = 1;
    const double d = ak[p];
    if (!(d != 0. && isfinite(d))) { atomicOr(flags, kErrSingular); }
  }
  __syncthreads();
  const int p = pivot;
  const double d = ak[p];
  for (int i = threadIdx.x; i < N; i += blockDim.x) { mult[i] = used[i] ? 0. : ak[i] / d; }
}
// Step k, elimination: for every unused row i, A(i,j) -= mult(i) A(p,j) for
// j > k and A(i,k) = mult(i) (the L factor).
__global__ __launch_bounds__(kDiagThreads) void k_lu_update(double *__restrict__ a, int N, int k,
                                                            const int *__restrict__ used,
                                                            const int *__restrict__ piv,
                                                            const double *__restrict__ mult) {
  const int p = piv[k];
  const long cols = N - k;
  for (long id = long(blockIdx.x) * blockDim.x + threadIdx.x; id < cols * N; id += long(gridDim.x) * blockDim.x) {
    const int i = int(id % N), j = k + int(id / N);
    if (used[i]) { continue; }
    double *aij = a + size_t(j) * N + i;
    if (j == k) { *aij = mult[i]; }
    else { *aij = fma(-mult[i], a[size_t(j) * N + p], *aij); }
  }
}
// x = A^{-1} b from the factors above (one workgroup): forward elimination
// in pivot order, then back substitution; work: 2 N doubles.
__global__ __launch_bounds__(kLuThreads) void k_lu_solve(const double *__restrict__ a, int N,
                                                         const int *__restrict__ piv, const double *__restrict__ b,
                                                         double *__restrict__ x, double *__restrict__ work,
                                                         int *__restrict__ step) {
  // step[i]: the column at which row i was the pivot row
  double *z = work, *zr = work + N;
  for (int k = threadIdx.x; k < N; k += blockDim.x) { step[piv[k]] = k; }
  for (int i = threadIdx.x; i < N; i += blockDim.x) { z[i] = b[i]; }
  __syncthreads();
  // forward: z(piv[k]) final at column k; rows pivoted later take -L(i,k) z
  for (int k = 0; k < N; ++k) {
    const double zk = z[piv[k]];
    const double *ak = a + size_t(k) * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      if (step[i] > k) { z[i] = fma(-ak[i], zk, z[i]); }
    }
    __syncthreads();
  }
  // back substitution on U(k, l) = A(piv[k], l), l >= k
  for (int k = threadIdx.x; k < N; k += blockDim.x) { zr[k] = z[piv[k]]; }
  __syncthreads();
  for (int k = N - 1; k >= 0; --k) {
    const double *ak = a + size_t(k) * N;
    const double xk = zr[k] / ak[piv[k]];
    __syncthreads(); // every thread has read zr[k]
    if (threadIdx.x == 0) { x[k] = xk; }
    for (int l = threadIdx.x; l < k; l += blockDim.x) { zr[l] = fma(-ak[piv[l]], xk, zr[l]); }
    __syncthreads();
  }
}
// Block permutation of a vector of nb blocks of b: gather dst(i) =
// src(perm(i)), or scatter dst(perm(i)) = src(i).
template <class T>
__global__ __launch_bounds__(kDiagThreads) void k_block_permute(const T *__restrict__ src, T *__restrict__ dst,
                                                                const int *__restrict__ perm, int b, int nb,
                                                                int scatter) {
  const long N = long(nb) * b;
  for (long id = long(blockIdx.x) * blockDim.x + threadIdx.x; id < N; id += long(gridDim.x) * blockDim.x) {
    const int i = int(id / b), r = int(id % b);
    const size_t other = size_t(perm[i]) * b + r;
    if (scatter) { dst[other] = src[id]; }
    else { dst[id] = src[other]; }
  }
}
} // namespace

template <class T>
void launch_bsr_mult(const T *bsr, int b, int nb, const int *row_ptr, const int *col, const T *x, T *y,
                     hipStream_t s) {
  k_bsr_mult<T><<<diag_grid(long(nb) * b), kDiagThreads, 0, s>>>(bsr, b, nb, row_ptr, col, x, y);
}
template void launch_bsr_mult<double>(const double *, int, int, const int *, const int *, const double *, double *,
                                      hipStream_t);
template void launch_bsr_mult<float>(const float *, int, int, const int *, const int *, const float *, float *,
                                     hipStream_t);
void launch_bsr_to_dense(const double *bsr, int b, int nb, const int *row_ptr, const int *col, double *dense,
                         hipStream_t s) {
  const long N = long(nb) * b;
  k_fill_zero<<<diag_grid(N * N), kDiagThreads, 0, s>>>(dense, size_t(N) * N);
  k_bsr_to_dense<<<diag_grid(N * b), kDiagThreads, 0, s>>>(bsr, b, nb, row_ptr, col, dense);
}
void launch_dense_lu(double *a, int N, int *used, int *piv, double *mult, int *flags, hipStream_t s) {
  k_fill_zero_int<<<diag_grid(N), kDiagThreads, 0, s>>>(used, size_t(N));
  for (int k = 0; k < N; ++k) {
    k_lu_pivot<<<1, kLuThreads, 0, s>>>(a, N, k, used, piv, mult, flags);
    k_lu_update<<<diag_grid(long(N - k) * N), kDiagThreads, 0, s>>>(a, N, k, used, piv, mult);
  }
}
void launch_dense_solve(const double *a, int N, const int *piv, const double *b, double *x, double *work, int *step,
                        hipStream_t s) {
  k_lu_solve<<<1, kLuThreads, 0, s>>>(a, N, piv, b, x, work, step);
}
template <class T>
void launch_block_permute(const T *src, T *dst, const int *perm, int b, int nb, bool scatter, hipStream_t s) {
  k_block_permute<T><<<diag_grid(long(nb) * b), kDiagThreads, 0, s>>>(src, dst, perm, b, nb, scatter ? 1 : 0);
}
template void launch_block_permute<double>(const double *, double *, const int *, int, int, bool, hipStream_t);
template void launch_block_permute<float>(const float *, float *, const int *, int, int, bool, hipStream_t);
} // namespace eg
