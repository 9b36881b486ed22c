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
#include <stdexcept>
#include <cstdlib>
#include <string>

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

// Gauss-Jordan inverse for 64 <= B <= 196 with one wavefront per column
// group: wave g keeps columns g + NCG j (j < NE) of every row, lane l rows
// l + 64 i (i < RW), all in registers. Column k = s + NCG jb belongs to
// wave s. Each step is one workgroup barrier: after it every wave reads the
// pivot row p of step k, 1 / S(p,k) and column k from LDS. The wave that
// owns column k + 1 first updates that column alone and searches its pivot
// (a register max over its rows, then wave_argmax over the lanes' float
// magnitudes), and publishes p, 1 / S(p,k+1) and the column for
// the next step while the other waves are still updating; then every wave
// finishes the rank-1 update S(i,j) -= S(i,k) S(p,j) / S(p,k) of its columns
// with the pivot row's values taken from lane p % 64 (readlane). Rows are
// kept unscaled: the pivot row keeps its values and records scale[p] =
// 1 / S(p,k), column k becomes -S(i,k) / S(p,k) and 1 at row p, and
// A^{-1}(pinv[i], piv[j]) = scale[i] S(i,j) at the end (piv[k] = p,
// pinv[p] = k; no row interchanges). After each group of NCG steps every
// wave rotates its register columns by one, so the column of the current
// step is register 0 of its owner. Mathematically the inverse of the block
// whose LU the reference uses (src/linear.cpp:198-202,
// src/block_ilu.cpp:83-84, src/condensation.cpp:137-140).
template <int B, int NCG> struct Gj2 {
  static constexpr int RW = (B + 63) / 64;
  static constexpr int NE = (B + NCG - 1) / NCG;
  static constexpr int NT = 64 * NCG;
  static_assert(RW <= 4 && NT <= 1024, "register-resident block");
};
template <int B, int NCG, class T> struct Gj2Lds {
  T col[2][64 * Gj2<B, NCG>::RW];
  T inv[2];
  int prow[2];
  T scale[B];
  int piv[B], pinv[B];
};
// 1 / a from the hardware reciprocal and two Newton steps (about 1 ulp;
// zero, infinite or NaN a give a NaN, which the caller reports as a
// singular pivot).
__device__ __forceinline__ double recip(double a) {
  double x = __builtin_amdgcn_rcp(a);
  double e = fma(-a, x, 1.);
  x = fma(x, e, x);
  e = fma(-a, x, 1.);
  return fma(x, e, x);
}
// 1 / a in single precision: the hardware reciprocal and one Newton step.
__device__ __forceinline__ float recip(float a) {
  const float x = __builtin_amdgcn_rcpf(a);
  return fmaf(x, fmaf(-a, x, 1.f), x);
}
__device__ __forceinline__ float read_lane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double read_lane(double v, int lane) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(u)), lane));
  const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(u >> 32)), lane));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
template <int CTRL, int ROWS> __device__ __forceinline__ int dpp_imax(int x) {
  // rows outside ROWS read INT_MIN, the identity of max
  return max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, CTRL, ROWS, 0xF, false));
}
// Pivot lane of a wavefront: the largest float magnitude mag (sign clear; NaN
// counts as infinity) over the lanes with used false, the lowest lane among
// magnitudes that differ only in the six low bits of the float significand
// (a relative difference below 2^-17). The key carries 63 - lane in the
// magnitude's six low bits, so one integer max names the lane: DPP within
// rows of 16 lanes, then row_bcast:15 and row_bcast:31 carry the row maxima
// up to lane 63. At least one lane must have used false.
__device__ __forceinline__ int wave_argmax(float mag, bool used, int lane) {
  const int bits = min(__float_as_int(mag), 0x7F800000);
  int x = used ? -1 : ((bits & ~63) | (63 - lane));
  x = dpp_imax<0xB1, 0xF>(x);  // quad_perm [1,0,3,2]
  x = dpp_imax<0x4E, 0xF>(x);  // quad_perm [2,3,0,1]
  x = dpp_imax<0x141, 0xF>(x); // row_half_mirror
  x = dpp_imax<0x140, 0xF>(x); // row_mirror
  x = dpp_imax<0x142, 0xA>(x); // row_bcast:15 into rows 1 and 3
  x = dpp_imax<0x143, 0xC>(x); // row_bcast:31 into rows 2 and 3
  return 63 - (__builtin_amdgcn_readlane(x, 63) & 63);
}
// Pivot search of the owning wave over the register column x (rows lane +
// 64 i not yet used); publishes the pivot row, its reciprocal and the column.
template <int B, int NCG, class T>
__device__ __forceinline__ void gj2_publish(const T (&x)[Gj2<B, NCG>::RW], unsigned used, int lane,
                                            Gj2Lds<B, NCG, T> &L, int buf) {
  constexpr int RW = Gj2<B, NCG>::RW;
  float key = -1.f;
  int bi = 0;
  T bv = T(0);
#pragma unroll
  for (int i = 0; i < RW; ++i) {
    const int r = lane + 64 * i;
    const float a = float(x[i]);
    const float kk = (r < B && !((used >> i) & 1u)) ? (a != a ? __int_as_float(0x7F800000) : fabsf(a)) : -1.f;
    if (kk > key) { key = kk; bi = i; bv = x[i]; }
    L.col[buf][r] = x[i];
  }
  const int pl = wave_argmax(key, key < 0.f, lane);
  const int pi = __builtin_amdgcn_readlane(bi, pl);
  const T ap = read_lane(bv, pl);
  if (lane == 0) {
    L.prow[buf] = pl + 64 * pi;
    L.inv[buf] = recip(ap);
  }
}
template <int B, int NCG, class T>
__device__ __forceinline__ bool gj2_inverse(T *__restrict__ A, Gj2Lds<B, NCG, T> &L) {
  using G = Gj2<B, NCG>;
  constexpr int RW = G::RW, NE = G::NE;
  const int t = threadIdx.x, lane = t & 63, cg = t >> 6;
  T v[RW][NE];
#pragma unroll
  for (int i = 0; i < RW; ++i) {
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int r = lane + 64 * i, c = cg + NCG * j;
      v[i][j] = (r < B && c < B) ? A[r + size_t(c) * B] : T(0);
    }
  }
  unsigned used = 0u; // bit i: row lane + 64 i is a pivot row already
  bool singular = false;
  if (cg == 0) {
    T x[RW];
#pragma unroll
    for (int i = 0; i < RW; ++i) { x[i] = v[i][0]; }
    gj2_publish<B, NCG, T>(x, used, lane, L, 0);
  }
  for (int jb = 0; jb < NE; ++jb) {
#pragma unroll
    for (int s = 0; s < NCG; ++s) {
      const int k = s + NCG * jb;
      if (k >= B) { break; }
      const int buf = k & 1;
      const int s1 = (s + 1) % NCG;
      const int j1 = s + 1 < NCG ? 0 : 1; // register of column k + 1 in its owner (unrolled: a constant)
      __syncthreads();
      const int p = __builtin_amdgcn_readfirstlane(L.prow[buf]);
      const T inv = L.inv[buf];
      if (!(isfinite(inv) && inv != T(0))) { singular = true; }
      const int ip = p >> 6, lp = p & 63;
      T m[RW];
#pragma unroll
      for (int i = 0; i < RW; ++i) { m[i] = (lane + 64 * i == p) ? T(0) : L.col[buf][lane + 64 * i] * inv; }
      const unsigned used_next = used | (lane == lp ? 1u << ip : 0u);
      const bool ahead = cg == s1 && k + 1 < B;
      if (ahead && j1 < NE) {
        __builtin_amdgcn_s_setprio(3); // the next pivot is the critical path
        T pr = T(0);
#pragma unroll
        for (int i = 0; i < RW; ++i) {
          if (ip == i) { pr = read_lane(v[i][j1], lp); }
        }
        T x[RW];
#pragma unroll
        for (int i = 0; i < RW; ++i) {
          v[i][j1] = fma(-m[i], pr, v[i][j1]);
          x[i] = v[i][j1];
        }
        gj2_publish<B, NCG, T>(x, used_next, lane, L, buf ^ 1);
        __builtin_amdgcn_s_setprio(0);
      }
      // each pivot-row value of this wave's columns read (readlane) just
      // before its column's update: two scalar registers live, not 2 NE
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        if (ip == i) {
#pragma unroll
          for (int j = 0; j < NE; ++j) {
            if ((ahead && j == j1) || (cg == s && j == 0)) { continue; }
            const T pj = read_lane(v[i][j], lp);
#pragma unroll
            for (int r = 0; r < RW; ++r) { v[r][j] = fma(-m[r], pj, v[r][j]); }
          }
        }
      }
      if (cg == s) {
#pragma unroll
        for (int i = 0; i < RW; ++i) { v[i][0] = (lane + 64 * i == p) ? T(1) : -m[i]; }
      }
      used = used_next;
      if (t == 0) {
        L.piv[k] = p;
        L.pinv[p] = k;
        L.scale[p] = inv;
      }
    }
#pragma unroll
    for (int i = 0; i < RW; ++i) {
      const T first = v[i][0];
#pragma unroll
      for (int j = 0; j + 1 < NE; ++j) { v[i][j] = v[i][j + 1]; }
      v[i][NE - 1] = first;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RW; ++i) {
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int r = lane + 64 * i, c = cg + NCG * j;
      if (r < B && c < B) { A[L.pinv[r] + size_t(L.piv[c]) * B] = L.scale[r] * v[i][j]; }
    }
  }
  __syncthreads();
  return singular;
}
// Gauss-Jordan inverse of a B x B block (B <= 64) by one wavefront, lane i
// holding row i in registers: no workgroup barrier and no LDS traffic, each
// step a wavefront arg-max of float pivot magnitudes (wave_argmax), the
// pivot row's values by readlane and a rank-1 update. Rows stay unscaled and
// pivoting is implicit, as in gj2_inverse: A^{-1}(step of row i, piv[j]) =
// scale_i S(i, j). A may point to LDS or global memory.
template <int B>
__device__ __forceinline__ bool gj_wave(double *A) {
  static_assert(B <= 64, "one row a lane");
  const int lane = threadIdx.x & 63;
  double v[B];
#pragma unroll
  for (int j = 0; j < B; ++j) { v[j] = lane < B ? A[lane + j * B] : 0.; }
  bool used = lane >= B;
  int step = 0, piv[B];
  double scale = 1.;
  bool singular = false;
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const int p = wave_argmax(fabsf(float(v[k])), used, lane);
    const double inv = recip(read_lane(v[k], p));
    if (!(isfinite(inv) && inv != 0.)) { singular = true; }
    const bool pivot = lane == p;
    const double m = pivot ? 0. : v[k] * inv;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      if (j != k) { v[j] = fma(-m, read_lane(v[j], p), v[j]); }
    }
    v[k] = pivot ? 1. : -m;
    if (pivot) {
      used = true;
      step = k;
      scale = inv;
    }
    piv[k] = p;
  }
  if (lane < B) {
#pragma unroll
    for (int j = 0; j < B; ++j) { A[step + piv[j] * B] = scale * v[j]; }
  }
  return singular;
}
// The same Gauss-Jordan inverse for B <= 16 with a quad of lanes per row:
// lane 4 r + q holds columns 4 q .. 4 q + 3 of row r, so a step updates four
// values a lane. The pivot search is gj_wave's (float magnitudes of column k
// over the unused rows by wave_argmax), S(r, k) reaches the lanes of row
// r by a quad broadcast and the pivot row by a lane permute; every value is
// computed by the same operations as in gj_wave, so the results are the same.
template <int KQ> __device__ __forceinline__ double quad_bcast(double v) {
  constexpr int ctrl = KQ | (KQ << 2) | (KQ << 4) | (KQ << 6);
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u)), ctrl, 0xF, 0xF, false));
  const unsigned hi = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u >> 32)), ctrl, 0xF, 0xF, false));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
template <int B>
__device__ __forceinline__ bool gj_quad(double *A) {
  static_assert(B <= 16, "one row a quad of lanes");
  const int lane = threadIdx.x & 63, q = lane & 3, r = lane >> 2;
  double v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) { v[c] = (r < B && 4 * q + c < B) ? A[r + (4 * q + c) * B] : 0.; }
  bool used = r >= B;
  int step = 0, piv[B];
  double scale = 1.;
  bool singular = false;
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const int kq = k >> 2, kc = k & 3;
    const int pl = wave_argmax(fabsf(float(v[kc])), used || q != kq, lane);
    const int p = pl >> 2;
    const double inv = recip(read_lane(v[kc], pl));
    if (!(isfinite(inv) && inv != 0.)) { singular = true; }
    double srk;
    switch (kq) {
    case 0: srk = quad_bcast<0>(v[kc]); break;
    case 1: srk = quad_bcast<1>(v[kc]); break;
    case 2: srk = quad_bcast<2>(v[kc]); break;
    default: srk = quad_bcast<3>(v[kc]); break;
    }
    const bool pivot = r == p;
    const double m = pivot ? 0. : srk * inv;
    const int src = 4 * p + q;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double pr = __shfl(v[c], src, 64);
      v[c] = (c == kc && q == kq) ? (pivot ? 1. : -m) : fma(-m, pr, v[c]);
    }
    if (pivot) {
      used = true;
      step = k;
      scale = inv;
    }
    piv[k] = p;
  }
  if (r < B) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int pc = 0;
#pragma unroll
      for (int j = c; j < B; j += 4) { pc = (j == 4 * q + c) ? piv[j] : pc; }
      if (4 * q + c < B) { A[step + pc * B] = scale * v[c]; }
    }
  }
  return singular;
}
// Gauss-Jordan inverse of a 64 x 64 block by B / PW wavefronts with panel
// ownership: wave w keeps columns PW w .. PW w + PW - 1 of every row (lane =
// row) and owns steps k of its panel. The owner searches the pivot of column
// k (wave_argmax over float magnitudes, as in gj_wave), publishes
// the multipliers of step k (with 1 / S(p,k) in the pivot row's slot) and p
// in LDS and raises the step counter; every other wave applies step k to its
// columns as soon as the counter passes k. No workgroup barrier in the step
// loop: the owner's pivot recurrence never waits for the other waves, and a
// wave that becomes owner has applied every earlier step to its panel. Each
// value is computed by the same operations as in gj_wave.
template <int B> struct GjPanelLds {
  double m[B][64];
  int p[B];
  int step;
};
template <int B, int PW>
__device__ __forceinline__ bool gj_panel(double *__restrict__ A, GjPanelLds<B> &L, int *flags) {
  static_assert(B == 64 && B % PW == 0, "one row a lane");
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double v[PW];
#pragma unroll
  for (int c = 0; c < PW; ++c) { v[c] = A[lane + (PW * w + c) * B]; }
  if (t == 0) { L.step = 0; }
  __syncthreads();
  bool used = false, singular = false;
  int step = 0;
  double scale = 1.;
  for (int q = 0; q < B / PW; ++q) {
    if (w == q) {
      // the owner's pivot chain is the critical path: its instructions go
      // first on the SIMD it shares with a follower
      __builtin_amdgcn_s_setprio(3);
#pragma unroll
      for (int kc = 0; kc < PW; ++kc) {
        const int k = PW * q + kc;
        const int p = wave_argmax(fabsf(float(v[kc])), used, lane);
        const double inv = recip(read_lane(v[kc], p));
        if (!(isfinite(inv) && inv != 0.)) { singular = true; }
        const bool pivot = lane == p;
        const double m = pivot ? 0. : v[kc] * inv;
        if (kc + 1 < PW) { v[kc + 1] = fma(-m, read_lane(v[kc + 1], p), v[kc + 1]); }
        L.m[k][lane] = pivot ? inv : m;
        L.p[k] = p;
#pragma unroll
        for (int c = 0; c < PW; ++c) {
          if (c != kc && c != kc + 1) { v[c] = fma(-m, read_lane(v[c], p), v[c]); }
        }
        // The LDS operations of one wavefront complete in issue order, so the
        // multipliers are visible before the counter that announces them; a
        // compiler fence keeps that order without waiting for the stores.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __hip_atomic_store(&L.step, k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        v[kc] = pivot ? 1. : -m;
        used = used || pivot;
        step = pivot ? k : step;
        scale = pivot ? inv : scale;
      }
      __builtin_amdgcn_s_setprio(0);
    } else {
      for (int kc = 0; kc < PW; ++kc) {
        const int k = PW * q + kc;
        int spins = 0;
        while (__hip_atomic_load(&L.step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) <= k) {
          if (++spins > (1 << 22)) {
            atomicOr(flags, kErrSpin);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        const int p = __builtin_amdgcn_readfirstlane(L.p[k]);
        const double mv = L.m[k][lane];
        const bool pivot = lane == p;
        const double m = pivot ? 0. : mv;
#pragma unroll
        for (int c = 0; c < PW; ++c) { v[c] = fma(-m, read_lane(v[c], p), v[c]); }
        used = used || pivot;
        step = pivot ? k : step;
        scale = pivot ? mv : scale;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < PW; ++c) { A[step + L.p[PW * w + c] * B] = scale * v[c]; }
  __syncthreads();
  return singular;
}
template <int B, int PW>
__global__ __launch_bounds__(64 * B / PW) void k_block_inverse_p(double *__restrict__ blocks,
                                                                 const int *__restrict__ index,
                                                                 int *__restrict__ flags) {
  __shared__ GjPanelLds<B> L;
  const int blk = index ? index[blockIdx.x] : blockIdx.x;
  if (gj_panel<B, PW>(blocks + size_t(blk) * B * B, L, flags) && threadIdx.x == 0) { atomicOr(flags, kErrSingular); }
}
template <int B> struct Gj2Pick;
template <> struct Gj2Pick<100> { static constexpr int NCG = 8; };
template <> struct Gj2Pick<144> { static constexpr int NCG = 8; };
template <> struct Gj2Pick<196> { static constexpr int NCG = 8; };
template <int B> constexpr int gj2_threads() { return 64 * Gj2Pick<B>::NCG; }
template <int B, class T = double>
__global__ __launch_bounds__(gj2_threads<B>()) void k_block_inverse_w(T *__restrict__ blocks,
                                                                      const int *__restrict__ index,
                                                                      int *__restrict__ flags) {
  constexpr int NCG = Gj2Pick<B>::NCG;
  __shared__ Gj2Lds<B, NCG, T> L;
  const int blk = index ? index[blockIdx.x] : blockIdx.x;
  if (gj2_inverse<B, NCG, T>(blocks + size_t(blk) * B * B, L) && threadIdx.x == 0) { atomicOr(flags, kErrSingular); }
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

// Row-resident ILU(0) factor of one row (small blocks): every block of the
// row and every block it reads from earlier rows (inv(A_kk) of each lower
// block, A_kj of each update pair) is loaded into LDS in one batch of
// independent loads, the row is factored in LDS (src/block_ilu.cpp:73-86:
// ascending k, L_ik = A_ik inv(A_kk), then A_ij -= L_ik A_kj over the update
// pairs of ik in parallel), the diagonal block is inverted in place and the
// row is written back as one contiguous run. The per-row program (built on
// the host from the fixed pattern) is
//   [len, first slot, blocks in row, diagonal position, dependencies,
//    dependency slots..., per lower block: (inverse dependency, updates,
//    (target position, dependency) x updates)].
// Entry o of a B x B block product in LDS (column-major), the even and odd
// terms in two accumulators (half the dependent chain) added at the end.
template <int B>
__device__ __forceinline__ double block_dot(const double *X, const double *Y, int o) {
  static_assert(B % 2 == 0, "even block size");
  const int r = o % B, c = o / B;
  double acc0 = 0., acc1 = 0.;
#pragma unroll 4
  for (int kk = 0; kk < B; kk += 2) {
    acc0 += X[r + kk * B] * Y[kk + c * B];
    acc1 += X[r + (kk + 1) * B] * Y[kk + 1 + c * B];
  }
  return acc0 + acc1;
}
template <int B>
__device__ __forceinline__ void block_dot2(const double *X, const double *Y0, const double *Y1, int o, double &a0,
                                           double &a1) {
  static_assert(B % 2 == 0, "even block size");
  const int r = o % B, c = o / B;
  double a00 = 0., a01 = 0., a10 = 0., a11 = 0.;
#pragma unroll 4
  for (int kk = 0; kk < B; kk += 2) {
    const double x0 = X[r + kk * B], x1 = X[r + (kk + 1) * B];
    a00 += x0 * Y0[kk + c * B];
    a01 += x1 * Y0[kk + 1 + c * B];
    a10 += x0 * Y1[kk + c * B];
    a11 += x1 * Y1[kk + 1 + c * B];
  }
  a0 = a00 + a01;
  a1 = a10 + a11;
}
// Blocks b0 .. b1 - 1 of a row program's LDS image (the row's own blocks,
// then its dependencies) loaded with all loads of a batch in flight before
// the first LDS store.
template <int B, bool COH>
__device__ __forceinline__ void resident_load(const double *bsr, const int *P, double *sm, int b0, int b1) {
  constexpr int bb = B * B, NT = Shape<B>::NT, U = 16;
  const int t = threadIdx.x, j0 = P[1], nrow = P[2];
  const int base = b0 * bb, total = (b1 - b0) * bb;
  for (int e0 = 0; e0 < total; e0 += U * NT) {
    double tmp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * NT + t;
      tmp[u] = 0.;
      if (e < total) {
        const int blk = (base + e) / bb, o = base + e - blk * bb;
        const int slot = blk < nrow ? j0 + blk : P[5 + blk - nrow];
        tmp[u] = ld<COH>(bsr + size_t(slot) * bb + o);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * NT + t;
      if (e < total) { sm[base + e] = tmp[u]; }
    }
  }
}
// One 16 x 16 output tile (tr, tc) of the B x B block product X Y (B <= 36,
// a multiple of 4) on the FP64 matrix cores (v_mfma_f64_16x16x4f64): lane
// (g, j) supplies X(16 tr + j, 4 s + g) and Y(4 s + g, 16 tc + j) at step s;
// entry q of the result is (X Y)(16 tr + g + 4 q, 16 tc + j).
typedef double rf_d4 __attribute__((ext_vector_type(4)));
template <int B> __device__ __forceinline__ rf_d4 block_mm_tile(const double *X, const double *Y, int tr, int tc) {
  static_assert(B % 4 == 0, "whole k steps");
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  const int r = 16 * tr + j, c = 16 * tc + j;
  rf_d4 acc = {0., 0., 0., 0.};
#pragma unroll
  for (int s = 0; s < B / 4; ++s) {
    const int k = 4 * s + g;
    const double a = r < B ? X[r + k * B] : 0.;
    const double b = c < B ? Y[k + c * B] : 0.;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  return acc;
}
// Factor of one row program in LDS (src/block_ilu.cpp:73-86): for each lower
// block in ascending k, L_ik = A_ik inv(U_kk) and then A_ij -= L_ik U_kj for
// its update pairs; then the diagonal block's inverse by one wavefront.
// Returns true for a singular pivot. Blocks from 20 x 20 form each product as
// 16 x 16 matrix-core tiles, one tile a wavefront (the face factor at 48x48
// Q4 / Q6: 3.96 -> 3.32 / 6.77 -> 4.85 ms); smaller ones, where a tile is a
// latency chain on one wavefront, keep one entry a thread (block_dot; 1.63 /
// 1.83 ms at Q2 / Q3 against 1.87 / 2.22 with tiles).
template <int B>
__device__ __forceinline__ bool resident_factor(const int *P, double *sm) {
  constexpr int bb = B * B, NT = Shape<B>::NT, T = (B + 15) / 16, NW = NT / 64, NQ = (bb + NT - 1) / NT;
  constexpr bool tiles = B >= 20;
  static_assert(!tiles || T * T <= NW, "one product tile a wavefront");
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, g = lane >> 4, j = lane & 15;
  const int tr = w % T, tc = w / T;
  const bool tiler = w < T * T;
  const int nrow = P[2], dpos = P[3], ndep = P[4];
  double *Arow = sm, *Dep = sm + size_t(nrow) * bb;
  int pc = 5 + ndep;
  for (int l = 0; l < dpos; ++l) {
    const int inv = P[pc], nupd = P[pc + 1];
    const int ops = pc + 2;
    pc = ops + 2 * nupd;
    double *X = Arow + size_t(l) * bb;
    const double *Y = Dep + size_t(inv) * bb;
    if constexpr (!tiles) {
      double res[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int o = t + q * NT;
        res[q] = o < bb ? block_dot<B>(X, Y, o) : 0.;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int o = t + q * NT;
        if (o < bb) { X[o] = res[q]; }
      }
      __syncthreads();
      // the update pairs of this lower block two at a time (independent
      // accumulators, one read of L_ik for both)
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int o = t + q * NT;
        if (o < bb) {
          int u = 0;
          for (; u + 1 < nupd; u += 2) {
            double a0, a1;
            block_dot2<B>(X, Dep + size_t(P[ops + 2 * u + 1]) * bb, Dep + size_t(P[ops + 2 * u + 3]) * bb, o, a0,
                          a1);
            Arow[size_t(P[ops + 2 * u]) * bb + o] -= a0;
            Arow[size_t(P[ops + 2 * u + 2]) * bb + o] -= a1;
          }
          if (u < nupd) {
            Arow[size_t(P[ops + 2 * u]) * bb + o] -= block_dot<B>(X, Dep + size_t(P[ops + 2 * u + 1]) * bb, o);
          }
        }
      }
      __syncthreads();
      continue;
    }
    rf_d4 res = {0., 0., 0., 0.};
    if (tiler) { res = block_mm_tile<B>(X, Y, tr, tc); }
    __syncthreads(); // every read of A_ik is done
    if (tiler) {
      const int c = 16 * tc + j;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = 16 * tr + g + 4 * q;
        if (r < B && c < B) { X[r + c * B] = res[q]; }
      }
    }
    __syncthreads();
    if (tiler) {
      const int c = 16 * tc + j;
      for (int u = 0; u < nupd; ++u) {
        const rf_d4 prod = block_mm_tile<B>(X, Dep + size_t(P[ops + 2 * u + 1]) * bb, tr, tc);
        double *tgt = Arow + size_t(P[ops + 2 * u]) * bb;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 16 * tr + g + 4 * q;
          if (r < B && c < B) { tgt[r + c * B] -= prod[q]; }
        }
      }
    }
    __syncthreads();
  }
  bool singular = false;
  if (t < 64) {
    if constexpr (B <= 16) { singular = gj_quad<B>(Arow + size_t(dpos) * bb); }
    else { singular = gj_wave<B>(Arow + size_t(dpos) * bb); }
  }
  __syncthreads();
  return singular;
}
template <int B>
__global__ __launch_bounds__(Shape<B>::NT) void k_ilu_factor_resident(double *__restrict__ bsr,
                                                                      const int *__restrict__ prog_off,
                                                                      const int *__restrict__ prog,
                                                                      int *__restrict__ flags) {
  extern __shared__ double sm[];
  __shared__ int P[kIluProgMax];
  constexpr int bb = B * B, NT = Shape<B>::NT;
  const int t = threadIdx.x;
  const int *pg = prog + prog_off[blockIdx.x];
  for (int e = t; e < kIluProgMax; e += NT) { P[e] = pg[e]; } // the program array is padded by kIluProgMax
  __syncthreads();
  const int j0 = P[1], nrow = P[2];
  resident_load<B, false>(bsr, P, sm, 0, nrow + P[4]);
  __syncthreads();
  const bool singular = resident_factor<B>(P, sm);
  for (int e = t; e < nrow * bb; e += NT) { bsr[size_t(j0) * bb + e] = sm[e]; }
  if (singular && t == 0) { atomicOr(flags, kErrSingular); }
}
template <int B> constexpr size_t resident_gj_lds() { return 0; }

// ILU(0) block products of one dependency level by row slabs: rows
// [r0, r0 + R) of the blocks of row i depend only on the same rows of row i's
// blocks and on whole blocks of earlier rows, so every (row, slab) pair is an
// independent workgroup (src/block_ilu.cpp:73-86, same order of products per
// slab: ascending k, L_ik then its update pairs). The diagonal blocks are
// inverted by a second launch. Thread (tc, tr) = (tid % 64, tid / 64) of the
// 256 computes rows tr RT .. tr RT + RT - 1 of the slab at columns tc + 64 j;
// the slab of the left operand sits in LDS transposed ([kk][r], broadcast
// reads), the right operand is staged in panels of KP rows ([kk][c], padded
// stride), prefetched into registers one panel ahead.
// Workgroup barrier that orders LDS accesses only: outstanding global loads
// (operand prefetches) and stores stay in flight across it.
// Workgroup w of a launch is dispatched to XCD w % 8. This maps the n
// workgroups to logical ids so that each XCD receives one run of consecutive
// ids (n / 8 or n / 8 + 1 of them): the slabs of a block row, which read the
// same operand blocks, then share one XCD's L2.
__device__ __forceinline__ int xcd_run_id(int w, int n) {
  constexpr int X = 8;
  const int per = n / X, rem = n % X, x = w % X, q = w / X;
  return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + q;
}
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
template <int B> struct Slab {
  static constexpr int R = 16;
  static constexpr int LDX = R + 1; // padded slab stride (the column-wise stores of L do not conflict)
  static constexpr int RT = R / 4;
  static constexpr int CJ = (B + 63) / 64;
  static constexpr int KP = B <= 64 ? B : 16; // larger blocks: panels of 16 rows (two workgroups a CU)
  static constexpr int LDY = B + 1;
  static constexpr int NS = (B + R - 1) / R;
  static constexpr int NP = (B + KP - 1) / KP;
  static constexpr int PER = (KP * B + 255) / 256;
  static constexpr size_t lds() { return (size_t(B) * LDX + size_t(KP) * LDY) * sizeof(double); }
};
template <int B>
__device__ __forceinline__ void slab_fetch(const double *__restrict__ Y, int k0, double (&pre)[Slab<B>::PER]) {
  using S = Slab<B>;
  const int t = threadIdx.x, kp = min(S::KP, B - k0);
#pragma unroll
  for (int q = 0; q < S::PER; ++q) {
    const int e = t + q * 256;
    const int kk = e % kp, c = e / kp;
    pre[q] = e < kp * B ? Y[k0 + kk + size_t(c) * B] : 0.;
  }
}
// acc = Xs Y: Xs the slab in LDS ([kk][r]), Y streamed from global memory in
// panels through Ys; pre holds Y's first panel on entry and, on return, the
// first panel of Ynext (when given).
template <int B>
__device__ __forceinline__ void slab_product(const double *__restrict__ Y, const double *__restrict__ Ynext,
                                             const double *Xs, double *Ys, double (&acc)[Slab<B>::RT][Slab<B>::CJ],
                                             double (&pre)[Slab<B>::PER]) {
  using S = Slab<B>;
  const int t = threadIdx.x, tc = t & 63, tr = t >> 6;
#pragma unroll
  for (int a = 0; a < S::RT; ++a) {
#pragma unroll
    for (int j = 0; j < S::CJ; ++j) { acc[a][j] = 0.; }
  }
  for (int p = 0; p < S::NP; ++p) {
    const int k0 = p * S::KP, kp = min(S::KP, B - k0);
    __syncthreads(); // the previous panel (and the caller's LDS stores) are done
#pragma unroll
    for (int q = 0; q < S::PER; ++q) {
      const int e = t + q * 256;
      const int kk = e % kp, c = e / kp;
      if (e < kp * B) { Ys[kk * S::LDY + c] = pre[q]; }
    }
    __syncthreads();
    if (p + 1 < S::NP) { slab_fetch<B>(Y, k0 + S::KP, pre); }
    else if (Ynext) { slab_fetch<B>(Ynext, 0, pre); }
#pragma unroll 4
    for (int kk = 0; kk < kp; ++kk) {
      double x[S::RT];
#pragma unroll
      for (int a = 0; a < S::RT; ++a) { x[a] = Xs[(k0 + kk) * S::LDX + tr * S::RT + a]; }
#pragma unroll
      for (int j = 0; j < S::CJ; ++j) {
        const int c = tc + 64 * j;
        const double y = (j + 1 < S::CJ || c < B) ? Ys[kk * S::LDY + c] : 0.;
#pragma unroll
        for (int a = 0; a < S::RT; ++a) { acc[a][j] += x[a] * y; }
      }
    }
  }
}
template <int B>
__global__ __launch_bounds__(256) void k_ilu_slab(double *__restrict__ bsr, const int *__restrict__ rows,
                                                  const int *__restrict__ row_ptr, const int *__restrict__ diag_ptr,
                                                  const int *__restrict__ col, const int *__restrict__ upd_ptr,
                                                  const int *__restrict__ upd) {
  using S = Slab<B>;
  extern __shared__ double sm[];
  constexpr size_t bb = size_t(B) * B;
  double *Xs = sm, *Ys = sm + B * S::LDX;
  const int bid = xcd_run_id(blockIdx.x, gridDim.x);
  const int i = rows[bid / S::NS], r0 = (bid % S::NS) * S::R;
  const int t = threadIdx.x, tc = t & 63, tr = t >> 6;
  double acc[S::RT][S::CJ];
  for (int ik = row_ptr[i]; ik < diag_ptr[i]; ++ik) {
    double *aik = bsr + size_t(ik) * bb;
    __syncthreads(); // Xs is free
    for (int e = t; e < B * S::R; e += 256) {
      const int r = e % S::R, kk = e / S::R;
      Xs[kk * S::LDX + r] = r0 + r < B ? aik[r0 + r + size_t(kk) * B] : 0.;
    }
    double pre[S::PER];
    slab_fetch<B>(bsr + size_t(diag_ptr[col[ik]]) * bb, 0, pre);
    slab_product<B>(bsr + size_t(diag_ptr[col[ik]]) * bb, nullptr, Xs, Ys, acc, pre); // L_ik = A_ik inv(A_kk)
    __syncthreads(); // every read of Xs is done
#pragma unroll
    for (int a = 0; a < S::RT; ++a) {
      const int r = tr * S::RT + a;
#pragma unroll
      for (int j = 0; j < S::CJ; ++j) {
        const int c = tc + 64 * j;
        if (c < B) {
          Xs[c * S::LDX + r] = acc[a][j];
          if (r0 + r < B) { aik[r0 + r + size_t(c) * B] = acc[a][j]; }
        }
      }
    }
    for (int u = upd_ptr[ik]; u < upd_ptr[ik + 1]; ++u) {
      slab_fetch<B>(bsr + size_t(upd[2 * u + 1]) * bb, 0, pre);
      slab_product<B>(bsr + size_t(upd[2 * u + 1]) * bb, nullptr, Xs, Ys, acc, pre); // L_ik A_kj
      double *tgt = bsr + size_t(upd[2 * u]) * bb;
#pragma unroll
      for (int a = 0; a < S::RT; ++a) {
        const int r = r0 + tr * S::RT + a;
#pragma unroll
        for (int j = 0; j < S::CJ; ++j) {
          const int c = tc + 64 * j;
          if (c < B && r < B) { tgt[r + size_t(c) * B] -= acc[a][j]; }
        }
      }
    }
  }
}

// The same slab products on the matrix cores (v_mfma_f64_16x16x4f64, and
// v_mfma_f32_16x16x4f32 for the single-precision factor): the 16-row slab
// is the M dimension of the tile, each wavefront owns the 16-column output
// tiles w, w + 4, ...; the right operand is read straight from global
// memory, one chunk of loads ahead of the MFMAs, the left operand (the
// slab, k-major) sits in LDS. SlabM: the layout of the two-stage kernel of
// 64 x 64 blocks (k_ilu_slab_star), lane (g, j) = (lane / 16, lane % 16)
// supplying k = g KS + s at step s (s < KS, 4 KS >= B, k >= B zero), the
// contiguous run Y[g KS .. g KS + KS) of column 16 ct + j. SlabV (below):
// the layout of k_ilu_slab_m, 16-byte loads.
constexpr int slab_chunk(int k4) {
  int best = 16, waste = 1 << 30;
  for (int c = 16; c >= 8; --c) {
    const int w = (k4 + c - 1) / c * c - k4;
    if (w < waste) { waste = w; best = c; }
  }
  return k4 < 16 ? k4 : best;
}
template <int B> struct SlabM {
  static constexpr int NS = (B + 15) / 16;           // slabs of a block row
  static constexpr int CT = (B + 15) / 16;           // 16-column output tiles
  static constexpr int TW = (CT + 3) / 4;            // tiles per wavefront (at most)
  static constexpr int K4 = (B + 3) / 4;
  static constexpr int S = slab_chunk(K4);           // steps per operand chunk
  static constexpr int KS = (K4 + S - 1) / S * S;    // steps per tile
  static constexpr int NH = KS / S;                  // chunks per tile
  static constexpr size_t lds() { return size_t(4) * KS * 16 * sizeof(double); }
};
typedef double slab_d4 __attribute__((ext_vector_type(4)));
typedef float slab_f4 __attribute__((ext_vector_type(4)));
// One 16 x 16 x 4 MFMA step per value type. Lane (g, j) = (lane / 16,
// lane % 16) supplies A(j, g) and B(g, j); entry q of its accumulator is
// D(row(g, q), j): row g + 4 q for v_mfma_f64_16x16x4f64, 4 g + q for
// v_mfma_f32_16x16x4f32.
template <class T> struct Mfma16;
template <> struct Mfma16<double> {
  typedef slab_d4 acc;
  static __device__ __forceinline__ acc mac(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return g + 4 * q; }
};
template <> struct Mfma16<float> {
  typedef slab_f4 acc;
  static __device__ __forceinline__ acc mac(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return 4 * g + q; }
};
// The operand layout of k_ilu_slab_m: lane (g, j) loads V = 16 / sizeof(T)
// consecutive values of its column with one 16-byte load, k = 4 V t + V g +
// e (e < V) at load t, so the four lane groups read one 64-byte run of each
// column (16 cache lines a load instruction, a load every V MFMAs). The
// MFMA step (t, e) reduces over k = 4 V t + V g + e, g < 4; k >= B is zero
// (B is a multiple of V: a load lies wholly inside or outside the block).
// The loads come in chunks of S, one chunk ahead of the MFMAs.
constexpr int slabv_chunk(int kl, int cap) {
  int best = cap, waste = 1 << 30;
  for (int c = cap; c >= 4; --c) {
    const int w = (kl + c - 1) / c * c - kl;
    if (w < waste) { waste = w; best = c; }
  }
  return kl < 4 ? kl : best;
}
template <int B, class T> struct SlabV {
  static constexpr int V = 16 / int(sizeof(T));
  static_assert(B % V == 0, "whole 16-byte loads");
  static constexpr int NS = (B + 15) / 16;                          // slabs of a block row
  static constexpr int CT = (B + 15) / 16;                          // 16-column output tiles
  static constexpr int TW = (CT + 3) / 4;                           // tiles per wavefront (at most)
  static constexpr int KL = (B + 4 * V - 1) / (4 * V);              // loads per tile
  static constexpr int S = slabv_chunk(KL, sizeof(T) == 8 ? 6 : 8); // loads per chunk
  static constexpr int NH = (KL + S - 1) / S;                       // chunks per tile
  static constexpr int KP = 4 * V * S * NH;                         // k extent of the slab in LDS
  static constexpr size_t lds() { return size_t(KP) * 16 * sizeof(T); }
};
template <class T> struct Vec16;
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct Vec16<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <int B, class T>
__device__ __forceinline__ void slabv_load(const T *__restrict__ Y, int ct, int h,
                                           typename Vec16<T>::type (&y)[SlabV<B, T>::S]) {
  using M = SlabV<B, T>;
  using vec = typename Vec16<T>::type;
  const int lane = threadIdx.x & 63, g = lane >> 4, c = 16 * ct + (lane & 15);
#pragma unroll
  for (int s = 0; s < M::S; ++s) {
    const int k = M::V * (4 * (h * M::S + s) + g);
    y[s] = (k < B && c < B) ? *reinterpret_cast<const vec *>(Y + k + size_t(c) * B) : vec(T(0));
  }
}
// acc[tw] = Xs Y for the tiles of this wavefront; pre holds the first chunk
// of Y on entry and the first chunk of Ynext (when given) on return.
template <int B, class T>
__device__ __forceinline__ void slabv_mfma(const T *__restrict__ Y, const T *__restrict__ Ynext, const T *Xs,
                                           typename Mfma16<T>::acc (&acc)[SlabV<B, T>::TW],
                                           typename Vec16<T>::type (&pre)[SlabV<B, T>::S], int w) {
  using M = SlabV<B, T>;
  using F = Mfma16<T>;
  using vec = typename Vec16<T>::type;
  // w: the wave (0..3) of the slab's workgroup
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  vec cur[M::S];
#pragma unroll
  for (int s = 0; s < M::S; ++s) { cur[s] = pre[s]; }
#pragma unroll
  for (int tw = 0; tw < M::TW; ++tw) {
    const int ct = w + 4 * tw;
    acc[tw] = typename F::acc{T(0), T(0), T(0), T(0)};
    if (ct < M::CT) {
#pragma unroll
      for (int h = 0; h < M::NH; ++h) {
        vec nxt[M::S];
        if (h + 1 < M::NH) { slabv_load<B, T>(Y, ct, h + 1, nxt); }
        else if (tw + 1 < M::TW && ct + 4 < M::CT) { slabv_load<B, T>(Y, ct + 4, 0, nxt); }
        else if (Ynext) { slabv_load<B, T>(Ynext, w, 0, nxt); }
        else {
#pragma unroll
          for (int s = 0; s < M::S; ++s) { nxt[s] = vec(T(0)); }
        }
        typename F::acc a4 = acc[tw];
#pragma unroll
        for (int s = 0; s < M::S; ++s) {
#pragma unroll
          for (int e = 0; e < M::V; ++e) {
            const T a = Xs[(M::V * (4 * (h * M::S + s) + g) + e) * 16 + j];
            a4 = F::mac(a, cur[s][e], a4);
          }
        }
        acc[tw] = a4;
#pragma unroll
        for (int s = 0; s < M::S; ++s) { cur[s] = nxt[s]; }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < M::S; ++s) { pre[s] = cur[s]; }
}
template <int B, class T>
__device__ __forceinline__ void slab_chunk_load(const T *__restrict__ Y, int ct, int h, T (&y)[SlabM<B>::S]) {
  using M = SlabM<B>;
  const int lane = threadIdx.x & 63, g = lane >> 4, c = 16 * ct + (lane & 15);
#pragma unroll
  for (int s = 0; s < M::S; ++s) {
    const int k = g * M::KS + h * M::S + s;
    y[s] = (k < B && c < B) ? Y[k + size_t(c) * B] : T(0);
  }
}
// The products of rows [r0, r0 + 16) of block row i by the 256 threads of a
// workgroup (t the thread index), the slab Xs in LDS. The wave index is
// passed to the MFMA chains as a value (fewer live registers than recomputing
// it there: 171 instead of 195 VGPRs at b = 196 in FP64, 135 instead of 153
// in FP32).
template <int B, class T>
__device__ __forceinline__ void slab_m_row(T *__restrict__ bsr, int i, int r0, const int *__restrict__ row_ptr,
                                           const int *__restrict__ diag_ptr, const int *__restrict__ col,
                                           const int *__restrict__ upd_ptr, const int *__restrict__ upd, T *Xs,
                                           int t) {
  using M = SlabV<B, T>;
  using F = Mfma16<T>;
  constexpr size_t bb = size_t(B) * B;
  const int w = t >> 6, lane = t & 63, g = lane >> 4, j = lane & 15;
  for (int e = B * 16 + t; e < M::KP * 16; e += 256) { Xs[e] = T(0); } // k >= B
  const int ik0 = row_ptr[i], ik1 = diag_ptr[i];
  typename F::acc acc[M::TW];
  typename Vec16<T>::type pre[M::S];
  if (ik0 < ik1) { slabv_load<B, T>(bsr + size_t(diag_ptr[col[ik0]]) * bb, w, 0, pre); }
  for (int ik = ik0; ik < ik1; ++ik) {
    T *aik = bsr + size_t(ik) * bb;
    if (ik > ik0) { __syncthreads(); } // earlier updates of A_ik are visible; Xs is free
    for (int e = t; e < B * 16; e += 256) {
      const int r = e & 15, k = e >> 4;
      Xs[e] = r0 + r < B ? aik[r0 + r + size_t(k) * B] : T(0);
    }
    lds_barrier();
    const int u0 = upd_ptr[ik], u1 = upd_ptr[ik + 1];
    const T *dnext = ik + 1 < ik1 ? bsr + size_t(diag_ptr[col[ik + 1]]) * bb : nullptr;
    slabv_mfma<B, T>(bsr + size_t(diag_ptr[col[ik]]) * bb, u0 < u1 ? bsr + size_t(upd[2 * u0 + 1]) * bb : dnext, Xs,
                     acc, pre, w); // L_ik = A_ik inv(A_kk)
    lds_barrier(); // every read of Xs is done
#pragma unroll
    for (int tw = 0; tw < M::TW; ++tw) {
      const int c = 16 * (w + 4 * tw) + j;
      if (c < B) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = F::row(g, q);
          Xs[c * 16 + r] = acc[tw][q];
          if (r0 + r < B) { aik[r0 + r + size_t(c) * B] = acc[tw][q]; }
        }
      }
    }
    lds_barrier();
    for (int u = u0; u < u1; ++u) {
      T *tgt = bsr + size_t(upd[2 * u]) * bb;
      T tv[M::TW][4];
#pragma unroll
      for (int tw = 0; tw < M::TW; ++tw) {
        const int c = 16 * (w + 4 * tw) + j;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = r0 + F::row(g, q);
          tv[tw][q] = (c < B && r < B) ? tgt[r + size_t(c) * B] : T(0);
        }
      }
      slabv_mfma<B, T>(bsr + size_t(upd[2 * u + 1]) * bb, u + 1 < u1 ? bsr + size_t(upd[2 * u + 3]) * bb : dnext,
                       Xs, acc, pre, w); // L_ik A_kj
#pragma unroll
      for (int tw = 0; tw < M::TW; ++tw) {
        const int c = 16 * (w + 4 * tw) + j;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = r0 + F::row(g, q);
          if (c < B && r < B) { tgt[r + size_t(c) * B] = tv[tw][q] - acc[tw][q]; }
        }
      }
    }
  }
}

// Up to 196 x 196 blocks three waves a SIMD (at most 168 VGPRs): a level's
// 13 slabs a row of 64 rows at 128x128 Q6 fit about one round of the device.
template <int B, class T = double>
__global__ __launch_bounds__(256, B <= 196 ? 3 : 1) void k_ilu_slab_m(T *__restrict__ bsr,
                                                                       const int *__restrict__ rows,
                                                    const int *__restrict__ row_ptr, const int *__restrict__ diag_ptr,
                                                    const int *__restrict__ col, const int *__restrict__ upd_ptr,
                                                    const int *__restrict__ upd) {
  extern __shared__ __attribute__((aligned(16))) unsigned char slab_lds[];
  const int bid = xcd_run_id(blockIdx.x, gridDim.x);
  slab_m_row<B, T>(bsr, rows[bid / SlabV<B, T>::NS], (bid % SlabV<B, T>::NS) * 16, row_ptr, diag_ptr, col, upd_ptr,
                   upd, reinterpret_cast<T *>(slab_lds), threadIdx.x);
}

// Panel-blocked Gauss-Jordan inverse on the matrix cores (B = 100, 144, 196).
// The pivots and the stored form are those of gj2_inverse (the largest float
// magnitude of the pivot column among the rows not used yet, unscaled rows,
// A^{-1}(pinv[i], piv[j]) = scale_i S(i, j)), but the steps are applied four
// at a time. One panel wave factors the four columns k0 .. k0 + 3 alone (rows
// lane + 64 i in registers, column k replaced by -S(i,k) / S(p,k) and 1 at
// row p as it is done) and hands over M' = M L^{-1}: M holds the four steps'
// multipliers S(i,k) / S(p,k) (0 at the step's own pivot row), L is unit
// lower with L(s,t) = M(p_s, t), the coupling of the steps through their
// pivot rows. The four steps then take every other column c from S(:, c) to
// S(:, c) - M' S(P, c), S(P, :) the four pivot rows before the panel: the
// tile waves apply this as one v_mfma 16x16x4 a 16 x 16 tile and take the
// panel's own columns from the panel wave. Tile wave w keeps tile rows TRW w
// .. TRW w + TRW - 1 of the NT x NT tiles (zero padded) as MFMA accumulators
// and rotates them by one tile column after each tile column, so that the
// panel's tile column is register 0. After the update of panel np the tile
// waves hand the columns of panel np + 2 to the panel wave, which applies
// the update of np + 1 to them itself (the pivot rows' values are in those
// columns) and factors them while the tile waves write the pivot rows of np +
// 1, wait for each other (an LDS count, not a barrier) and update every
// tile. One barrier a panel of four steps.
template <int B, class T> struct GjM {
  static constexpr int NT = (B + 15) / 16;
  static constexpr int TRW = 2;
  static constexpr int NW = (NT + TRW - 1) / TRW; // tile waves; wave NW factors the panels
  static constexpr int BP = 16 * NT;
  static constexpr int RW = (BP + 63) / 64;
  static constexpr int threads() { return 64 * (NW + 1); }
  static_assert(B % 4 == 0, "panels of four columns");
};
template <int B, class T> struct GjMLds {
  alignas(16) int pv[2][4];                 // the panel's pivot rows (by panel parity)
  int cnt;                                  // tile waves past their pivot-row writes
  alignas(16) T mp[2][GjM<B, T>::BP][4];    // -M', the MFMA A operand (by panel parity)
  alignas(16) T pan[3][4][GjM<B, T>::BP];   // a panel's columns (by panel index mod 3)
  alignas(16) T prow[4][GjM<B, T>::BP];     // the pivot rows before the panel, the B operand
  T scale[B];
  int piv[B], pinv[B];
};
// Wait of the tile waves for each other (the panel wave takes no part): each
// counts itself once its LDS writes are done, then waits for target counts.
// The wait is bounded (kErrSpin past 2^20 polls); every tile wave counts
// itself once a panel, so the count always arrives.
__device__ __forceinline__ int gjm_tile_sync(int *cnt, int target) {
  if ((threadIdx.x & 63) == 0) { __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP); }
  int spins = 0;
  while (__builtin_amdgcn_readfirstlane(__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) <
         target) {
    if (++spins > (1 << 20)) { return kErrSpin; }
    __builtin_amdgcn_s_sleep(1);
  }
  return 0;
}
template <int B, class T> using GjMTiles = typename Mfma16<T>::acc[GjM<B, T>::TRW][GjM<B, T>::NT];
// tile row r16 (0..15) of an accumulator: lane group and register
template <class T> __device__ __forceinline__ int gjm_grp(int r16) { return sizeof(T) == 8 ? r16 & 3 : r16 >> 2; }
template <class T> __device__ __forceinline__ int gjm_reg(int r16) { return sizeof(T) == 8 ? r16 >> 2 : r16 & 3; }
// lanes j in [j0, j0 + 4) of register tile column TC hold a panel's columns
template <int B, class T, int TC>
__device__ __forceinline__ void gjm_gather(const GjMTiles<B, T> &v, T (&pan)[4][GjM<B, T>::BP], int w, int j0) {
  using G = GjM<B, T>;
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  if (j >= j0 && j < j0 + 4) {
#pragma unroll
    for (int a = 0; a < G::TRW; ++a) {
      const int tr = G::TRW * w + a;
      if (tr < G::NT) {
#pragma unroll
        for (int q = 0; q < 4; ++q) { pan[j - j0][16 * tr + Mfma16<T>::row(g, q)] = v[a][TC][q]; }
      }
    }
  }
}
// the panel's factored columns into lanes j in [j0, j0 + 4) of register tile column 0
template <int B, class T>
__device__ __forceinline__ void gjm_replace(GjMTiles<B, T> &v, const T (&pan)[4][GjM<B, T>::BP], int w, int j0) {
  using G = GjM<B, T>;
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  if (j >= j0 && j < j0 + 4) {
#pragma unroll
    for (int a = 0; a < G::TRW; ++a) {
      const int tr = G::TRW * w + a;
      if (tr < G::NT) {
#pragma unroll
        for (int q = 0; q < 4; ++q) { v[a][0][q] = pan[j - j0][16 * tr + Mfma16<T>::row(g, q)]; }
      }
    }
  }
}
// S -= M' S(P, :) on every register tile column (prow holds the pivot rows
// by register tile column). The B operand of the next tile column is read
// from LDS while the MFMAs of this one issue. A tile row past the last (the
// last wave's, when NT is odd) has a zero A operand and stays zero; it costs
// no time, as that wave shares its SIMD with a wave of two tile rows.
template <int B, class T>
__device__ __forceinline__ void gjm_update(GjMTiles<B, T> &v, const T (&am)[GjM<B, T>::TRW],
                                           const T (&prow)[4][GjM<B, T>::BP], int w) {
  using G = GjM<B, T>;
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  const T *pr = &prow[g][j];
  T bn = pr[0];
#pragma unroll
  for (int tc = 0; tc < G::NT; ++tc) {
    const T bo = bn;
    if (tc + 1 < G::NT) { bn = pr[16 * (tc + 1)]; }
#pragma unroll
    for (int a = 0; a < G::TRW; ++a) { v[a][tc] = Mfma16<T>::mac(am[a], bo, v[a][tc]); }
  }
}
// x[pi] of lane pl (pi, pl uniform): every register read from lane pl, the
// choice made among the scalar values (no dynamically indexed array)
template <class T, int RW> __device__ __forceinline__ T gjm_pick(const T (&x)[RW], int pi, int pl) {
  T r = read_lane(x[0], pl);
#pragma unroll
  for (int i = 1; i < RW; ++i) {
    const T c = read_lane(x[i], pl);
    r = pi == i ? c : r;
  }
  return r;
}
template <int B, class T>
__device__ __forceinline__ int gjm_inverse(T *__restrict__ A, GjMLds<B, T> &L) {
  using G = GjM<B, T>;
  using F = Mfma16<T>;
  constexpr int NT = G::NT, TRW = G::TRW, RW = G::RW, BP = G::BP;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
  int err = 0;
  if (threadIdx.x == 0) { L.cnt = 0; }
  if (w < G::NW) {
    GjMTiles<B, T> v;
#pragma unroll
    for (int a = 0; a < TRW; ++a) {
      const int tr = TRW * w + a;
#pragma unroll
      for (int tc = 0; tc < NT; ++tc) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 16 * tr + F::row(g, q), c = 16 * tc + j;
          v[a][tc][q] = (r < B && c < B) ? A[r + size_t(c) * B] : T(0);
        }
      }
    }
    // panels 0 and 1 before any update
    gjm_gather<B, T, 0>(v, L.pan[0], w, 0);
    gjm_gather<B, T, 0>(v, L.pan[1], w, 4);
    lds_barrier(); // panels 0 and 1 are in pan[0], pan[1]; cnt is zero
    for (int tcb = 0; tcb < NT; ++tcb) {
#pragma unroll
      for (int pp = 0; pp < 4; ++pp) {
        const int k0 = 16 * tcb + 4 * pp;
        if (k0 >= B) { break; }
        const int np = k0 >> 2, par = np & 1, b3 = np % 3;
        lds_barrier(); // the panel wave has published pv[par], mp[par] and the factored panel in pan[b3]
        // the pivot rows before this panel's update
        const int4 pvv = *reinterpret_cast<const int4 *>(L.pv[par]);
        const int ps[4] = {__builtin_amdgcn_readfirstlane(pvv.x), __builtin_amdgcn_readfirstlane(pvv.y),
                           __builtin_amdgcn_readfirstlane(pvv.z), __builtin_amdgcn_readfirstlane(pvv.w)};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int trp = ps[s] >> 4;
          if (trp / TRW == w) {
            const int ap = trp % TRW, gp = gjm_grp<T>(ps[s] & 15), qp = gjm_reg<T>(ps[s] & 15);
#pragma unroll
            for (int a = 0; a < TRW; ++a) {
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                if (a == ap && q == qp && g == gp) {
#pragma unroll
                  for (int tc = 0; tc < NT; ++tc) { L.prow[s][16 * tc + j] = v[a][tc][q]; }
                }
              }
            }
          }
        }
        err |= gjm_tile_sync(&L.cnt, G::NW * (np + 1)); // every tile wave's pivot rows are in prow
        T am[TRW];
#pragma unroll
        for (int a = 0; a < TRW; ++a) { am[a] = TRW * w + a < NT ? L.mp[par][16 * (TRW * w + a) + j][g] : T(0); }
        gjm_update<B, T>(v, am, L.prow, w);
        gjm_replace<B, T>(v, L.pan[b3], w, 4 * pp);
        // the columns of panel np + 2 after this update, for the panel wave
        // (which applies the update of panel np + 1 to them itself)
        if (k0 + 8 < B) {
          if (pp < 2) { gjm_gather<B, T, 0>(v, L.pan[(np + 2) % 3], w, 4 * pp + 8); }
          else { gjm_gather<B, T, 1>(v, L.pan[(np + 2) % 3], w, 4 * pp - 8); }
        }
      }
#pragma unroll
      for (int a = 0; a < TRW; ++a) {
        const typename F::acc first = v[a][0];
#pragma unroll
        for (int tc = 0; tc + 1 < NT; ++tc) { v[a][tc] = v[a][tc + 1]; }
        v[a][NT - 1] = first;
      }
    }
    lds_barrier(); // piv, pinv and scale are complete
#pragma unroll
    for (int a = 0; a < TRW; ++a) {
      const int tr = TRW * w + a;
#pragma unroll
      for (int tc = 0; tc < NT; ++tc) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 16 * tr + F::row(g, q), c = 16 * tc + j;
          if (r < B && c < B) { A[L.pinv[r] + size_t(L.piv[c]) * B] = L.scale[r] * v[a][tc][q]; }
        }
      }
    }
  } else {
    bool singular = false;
    unsigned used = 0u; // bit i: row lane + 64 i is a pivot row already
    T mq[4][RW];        // -M' of the previous panel
    int pq[4] = {0, 0, 0, 0}; // its pivot rows
    lds_barrier();      // panels 0 and 1 are in pan[0], pan[1]
    for (int np = 0; np < B / 4; ++np) {
      const int par = np & 1, b3 = np % 3, k0 = 4 * np;
      T x[4][RW], m[4][RW], lc[4][4];
      int ps[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int i = 0; i < RW; ++i) {
          const int r = lane + 64 * i;
          x[c][i] = r < BP ? L.pan[b3][c][r] : T(0);
        }
      }
      if (np > 0) {
        // the previous panel's update of these columns, x -= M' S(P, cols),
        // with the pivot rows' values taken from these columns themselves
        // (they are the values before that update)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const T u0 = L.pan[b3][c][pq[0]], u1 = L.pan[b3][c][pq[1]], u2 = L.pan[b3][c][pq[2]],
                  u3 = L.pan[b3][c][pq[3]];
#pragma unroll
          for (int i = 0; i < RW; ++i) {
            x[c][i] = fma(mq[3][i], u3, fma(mq[2][i], u2, fma(mq[1][i], u1, fma(mq[0][i], u0, x[c][i]))));
          }
        }
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        float key = -1.f;
        int bi = 0;
#pragma unroll
        for (int i = 0; i < RW; ++i) {
          const int r = lane + 64 * i;
          const float a = float(x[s][i]);
          const float kk =
              (r < B && !((used >> i) & 1u)) ? (a != a ? __int_as_float(0x7F800000) : fabsf(a)) : -1.f;
          if (kk > key) { key = kk; bi = i; }
        }
        const int pl = wave_argmax(key, key < 0.f, lane);
        const int pi = __builtin_amdgcn_readlane(bi, pl);
        const int p = pl + 64 * pi;
        // the pivot row's values in the four columns and in m[t] (t < s),
        // read from lane pl under one uniform branch on its register pi
        T pc[4], lm[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) { pc[c] = lm[c] = T(0); }
#pragma unroll
        for (int i = 0; i < RW; ++i) {
          if (pi == i) {
#pragma unroll
            for (int c = 0; c < 4; ++c) { pc[c] = read_lane(x[c][i], pl); }
#pragma unroll
            for (int t = 0; t < s; ++t) { lm[t] = read_lane(m[t][i], pl); }
          }
        }
        const T inv = recip(pc[s]);
        if (!(isfinite(inv) && inv != T(0))) { singular = true; }
#pragma unroll
        for (int t = 0; t < s; ++t) { lc[s][t] = lm[t]; }
#pragma unroll
        for (int i = 0; i < RW; ++i) { m[s][i] = lane + 64 * i == p ? T(0) : x[s][i] * inv; }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (c == s) { continue; }
#pragma unroll
          for (int i = 0; i < RW; ++i) { x[c][i] = fma(-m[s][i], pc[c], x[c][i]); }
        }
#pragma unroll
        for (int i = 0; i < RW; ++i) { x[s][i] = lane + 64 * i == p ? T(1) : -m[s][i]; }
        used |= lane == pl ? 1u << pi : 0u;
        ps[s] = p;
        if (lane == s) {
          L.piv[k0 + s] = p;
          L.pinv[p] = k0 + s;
          L.scale[p] = inv;
        }
      }
      if (lane == 0) { *reinterpret_cast<int4 *>(L.pv[par]) = int4{ps[0], ps[1], ps[2], ps[3]}; }
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const int r = lane + 64 * i;
        const T m3 = m[3][i];
        const T m2 = fma(-m3, lc[3][2], m[2][i]);
        const T m1 = fma(-m3, lc[3][1], fma(-m2, lc[2][1], m[1][i]));
        const T m0 = fma(-m3, lc[3][0], fma(-m2, lc[2][0], fma(-m1, lc[1][0], m[0][i])));
        mq[0][i] = -m0;
        mq[1][i] = -m1;
        mq[2][i] = -m2;
        mq[3][i] = -m3;
        if (r < BP) {
          using vec = typename Vec16<T>::type;
          if constexpr (sizeof(T) == 8) {
            *reinterpret_cast<vec *>(&L.mp[par][r][0]) = vec{-m0, -m1};
            *reinterpret_cast<vec *>(&L.mp[par][r][2]) = vec{-m2, -m3};
          } else {
            *reinterpret_cast<vec *>(&L.mp[par][r][0]) = vec{-m0, -m1, -m2, -m3};
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) { L.pan[b3][c][r] = x[c][i]; }
        }
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) { pq[s] = ps[s]; }
      lds_barrier(); // pv[par], mp[par] and the factored panel in pan[b3] are published
    }
    lds_barrier();
    err = singular ? kErrSingular : 0;
  }
  return err;
}
template <int B, class T = double>
__global__ __launch_bounds__((GjM<B, T>::threads())) void k_block_inverse_m(T *__restrict__ blocks,
                                                                            const int *__restrict__ index,
                                                                            int *__restrict__ flags) {
  __shared__ GjMLds<B, T> L;
  const int blk = index ? index[blockIdx.x] : blockIdx.x;
  const int err = gjm_inverse<B, T>(blocks + size_t(blk) * B * B, L);
  if (err && (threadIdx.x & 63) == 0) { atomicOr(flags, err); }
}

// M_ij = inv(U_ii) U_ij in place for every block right of the diagonal
// (blocks up to 64, after the factor), one workgroup a block: both blocks in
// LDS, 4 x 4 output tiles a thread, the k sum in ascending order.
template <int B>
__global__ __launch_bounds__(256) void k_ilu_scale_upper(double *__restrict__ bsr, const int *__restrict__ slot,
                                                         const int *__restrict__ dslot) {
  static_assert(B % 4 == 0 && B <= 64, "4 x 4 tiles of a block up to 64");
  constexpr int bb = B * B, NQ = (bb + 255) / 256, T4 = B / 4;
  extern __shared__ double sh[];
  double *Ds = sh, *Us = sh + bb;
  const int t = threadIdx.x;
  const double *D = bsr + size_t(dslot[blockIdx.x]) * bb;
  double *U = bsr + size_t(slot[blockIdx.x]) * bb;
  double dv[NQ], uv[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int e = t + 256 * q;
    dv[q] = e < bb ? D[e] : 0.;
    uv[q] = e < bb ? U[e] : 0.;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int e = t + 256 * q;
    if (e < bb) {
      Ds[e] = dv[q];
      Us[e] = uv[q];
    }
  }
  __syncthreads();
  for (int tile = t; tile < T4 * T4; tile += 256) {
    const int r0 = 4 * (tile % T4), c0 = 4 * (tile / T4);
    double acc[4][4] = {};
    for (int k = 0; k < B; ++k) {
      double d[4], u[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        d[a] = Ds[r0 + a + k * B];
        u[a] = Us[k + (c0 + a) * B];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int c = 0; c < 4; ++c) { acc[a][c] = fma(d[a], u[c], acc[a][c]); }
      }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
      for (int c = 0; c < 4; ++c) { U[r0 + a + (c0 + c) * B] = acc[a][c]; }
    }
  }
}

// The same scaling for 144 x 144 and 196 x 196 blocks on the FP64 matrix
// cores. Column c of inv(U_ii) U_ij needs only column c of U_ij, so a block
// is cut into CG column groups, one 1024-thread workgroup each, which
// overwrites its own columns in place once it has read them: the product is
// accumulated in registers (16 x 16 output tiles, wavefront w the tiles w,
// w + 16, ..), its operands staged through LDS in panels of 16 k (A =
// inv(U_ii) k-major, B = the group's columns of U_ij), the next panel
// loaded into registers while the current one is multiplied.
template <int B> struct ScaleM {
  static constexpr int BP = (B + 15) / 16 * 16, TN = BP / 16;
  static constexpr int CG = B > 160 ? 2 : 1;        // column groups a block
  static constexpr int TC = (TN + CG - 1) / CG;     // tile columns a group (at most)
  static constexpr int TW = (TN * TC + 15) / 16;    // tiles a wavefront (at most)
  static constexpr int KP = 16, NP = BP / KP;       // k panels
  static constexpr int PA = KP * BP, PB = KP * 16 * TC; // values of one A / B panel
  static constexpr int LA = (PA + 1023) / 1024, LB = (PB + 1023) / 1024;
  static constexpr int KS = KP + 1;                 // B panel column stride (no bank conflicts across columns)
};
template <int B, class T = double>
__global__ __launch_bounds__(1024) void k_ilu_scale_upper_mfma(T *__restrict__ bsr, const int *__restrict__ slot,
                                                               const int *__restrict__ dslot) {
  using M = ScaleM<B>;
  using F = Mfma16<T>;
  __shared__ T As[M::PA];           // [k][r]
  __shared__ T Bs[16 * M::TC * M::KS]; // [c][k], column stride KS
  constexpr size_t bb = size_t(B) * B;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, g = lane >> 4, j = lane & 15;
  // the column groups of a block and the blocks of a row (consecutive
  // slots) read the same inv(U_ii): one XCD each run
  const int bid = xcd_run_id(blockIdx.x, gridDim.x), blk = bid / M::CG, cg = bid % M::CG;
  const int cstart = 16 * M::TC * cg, tcols = min(M::TC, M::TN - M::TC * cg);
  const T *D = bsr + size_t(dslot[blk]) * bb;
  T *U = bsr + size_t(slot[blk]) * bb;
  auto load = [&](int p, T (&av)[M::LA], T (&bv)[M::LB]) {
#pragma unroll
    for (int q = 0; q < M::LA; ++q) {
      const int e = t + 1024 * q, ka = e / M::BP, ra = e % M::BP, k = p * M::KP + ka;
      av[q] = (e < M::PA && ra < B && k < B) ? D[ra + size_t(k) * B] : T(0);
    }
#pragma unroll
    for (int q = 0; q < M::LB; ++q) {
      const int e = t + 1024 * q, cb = e / M::KP, kb = e % M::KP, k = p * M::KP + kb, c = cstart + cb;
      bv[q] = (e < M::PB && c < B && k < B) ? U[k + size_t(c) * B] : T(0);
    }
  };
  typename F::acc acc[M::TW];
#pragma unroll
  for (int tw = 0; tw < M::TW; ++tw) { acc[tw] = typename F::acc{T(0), T(0), T(0), T(0)}; }
  T av[M::LA], bv[M::LB];
  load(0, av, bv);
  for (int p = 0; p < M::NP; ++p) {
    if (p > 0) { __syncthreads(); } // every read of the previous panel is done
#pragma unroll
    for (int q = 0; q < M::LA; ++q) {
      const int e = t + 1024 * q;
      if (e < M::PA) { As[e] = av[q]; }
    }
#pragma unroll
    for (int q = 0; q < M::LB; ++q) {
      const int e = t + 1024 * q;
      if (e < M::PB) { Bs[(e / M::KP) * M::KS + e % M::KP] = bv[q]; }
    }
    __syncthreads();
    if (p + 1 < M::NP) { load(p + 1, av, bv); }
#pragma unroll
    for (int tw = 0; tw < M::TW; ++tw) {
      const int tile = w + 16 * tw;
      if (tile < M::TN * tcols) {
        const int r0 = 16 * (tile % M::TN), c0 = 16 * (tile / M::TN);
        typename F::acc a4 = acc[tw];
#pragma unroll
        for (int s4 = 0; s4 < M::KP / 4; ++s4) {
          const int k = 4 * s4 + g;
          a4 = F::mac(As[k * M::BP + r0 + j], Bs[(c0 + j) * M::KS + k], a4);
        }
        acc[tw] = a4;
      }
    }
  }
  __syncthreads(); // every read of this group's columns of U_ij is done
#pragma unroll
  for (int tw = 0; tw < M::TW; ++tw) {
    const int tile = w + 16 * tw;
    if (tile < M::TN * tcols) {
      const int r0 = 16 * (tile % M::TN), c = cstart + 16 * (tile / M::TN) + j;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = r0 + F::row(g, q);
        if (r < B && c < B) { U[r + size_t(c) * B] = acc[tw][q]; }
      }
    }
  }
}

// The same products for rows whose every lower block ik has one update
// pair, the diagonal (A_ii -= L_ik U_ki; the cell stencil), 64 x 64 blocks
// (each wavefront one 16-column tile, one chunk): the L products of all
// lower blocks are independent, so their operands are loaded in one batch
// and their MFMA chains run together; then the update products, whose
// operands are loaded while the L blocks are written. A_ii is updated in
// ascending k as in k_ilu_slab_m (the same products, the same bits).
constexpr int kStarLower = 4; // most lower blocks a row
template <int B, int NL>
__device__ __forceinline__ void slab_star_row(double *__restrict__ bsr, int i, int r0, int j0,
                                              const int *__restrict__ diag_ptr, const int *__restrict__ col,
                                              const int *__restrict__ upd_ptr, const int *__restrict__ upd,
                                              double *Xs) {
  using M = SlabM<B>;
  constexpr int XS = 4 * M::KS * 16;
  constexpr size_t bb = size_t(B) * B;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, g = lane >> 4, j = lane & 15;
  const int c = 16 * w + j; // this lane's output column (w < CT = 4)
  // stage 1 operands: the inverses of the lower blocks' rows, and the slabs
  double y[NL > 0 ? NL : 1][M::S];
#pragma unroll
  for (int l = 0; l < NL; ++l) { slab_chunk_load<B>(bsr + size_t(diag_ptr[col[j0 + l]]) * bb, w, 0, y[l]); }
  constexpr int XP = (B * 16 + 255) / 256;
  double xv[NL > 0 ? NL : 1][XP];
#pragma unroll
  for (int l = 0; l < NL; ++l) {
#pragma unroll
    for (int q = 0; q < XP; ++q) {
      const int e = t + 256 * q, r = e & 15, k = e >> 4;
      xv[l][q] = (e < B * 16 && r0 + r < B) ? bsr[size_t(j0 + l) * bb + r0 + r + size_t(k) * B] : 0.;
    }
  }
  for (int e = B * 16 + t; e < 4 * M::KS * 16; e += 256) {
#pragma unroll
    for (int l = 0; l < NL; ++l) { Xs[l * XS + e] = 0.; } // k >= B
  }
#pragma unroll
  for (int l = 0; l < NL; ++l) {
#pragma unroll
    for (int q = 0; q < XP; ++q) {
      const int e = t + 256 * q;
      if (e < B * 16) { Xs[l * XS + e] = xv[l][q]; }
    }
  }
  lds_barrier();
  slab_d4 acc[NL > 0 ? NL : 1];
#pragma unroll
  for (int l = 0; l < NL; ++l) { acc[l] = slab_d4{0., 0., 0., 0.}; }
#pragma unroll
  for (int s = 0; s < M::S; ++s) {
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      acc[l] = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[l * XS + (g * M::KS + s) * 16 + j], y[l][s], acc[l], 0, 0, 0);
    }
  }
  // stage 2 operands (U_ki of each lower block's row) while L is written
#pragma unroll
  for (int l = 0; l < NL; ++l) { slab_chunk_load<B>(bsr + size_t(upd[2 * upd_ptr[j0 + l] + 1]) * bb, w, 0, y[l]); }
  double *dg = bsr + size_t(diag_ptr[i]) * bb;
  double dv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = r0 + g + 4 * q;
    dv[q] = (c < B && r < B) ? dg[r + size_t(c) * B] : 0.;
  }
  lds_barrier(); // every read of the slabs is done
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    if (c < B) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = g + 4 * q;
        Xs[l * XS + c * 16 + r] = acc[l][q];
        if (r0 + r < B) { bsr[size_t(j0 + l) * bb + r0 + r + size_t(c) * B] = acc[l][q]; }
      }
    }
  }
  lds_barrier();
#pragma unroll
  for (int l = 0; l < NL; ++l) { acc[l] = slab_d4{0., 0., 0., 0.}; }
#pragma unroll
  for (int s = 0; s < M::S; ++s) {
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      acc[l] = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[l * XS + (g * M::KS + s) * 16 + j], y[l][s], acc[l], 0, 0, 0);
    }
  }
#pragma unroll
  for (int l = 0; l < NL; ++l) {
#pragma unroll
    for (int q = 0; q < 4; ++q) { dv[q] -= acc[l][q]; }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = r0 + g + 4 * q;
    if (c < B && r < B) { dg[r + size_t(c) * B] = dv[q]; }
  }
}
template <int B>
__global__ __launch_bounds__(256) void k_ilu_slab_star(double *__restrict__ bsr, const int *__restrict__ rows,
                                                       const int *__restrict__ row_ptr,
                                                       const int *__restrict__ diag_ptr,
                                                       const int *__restrict__ col, const int *__restrict__ upd_ptr,
                                                       const int *__restrict__ upd) {
  using M = SlabM<B>;
  static_assert(M::TW == 1 && M::NH == 1, "one tile a wavefront, one chunk a tile");
  extern __shared__ double Xs[]; // [l][k][16]
  const int bid = xcd_run_id(blockIdx.x, gridDim.x);
  const int i = rows[bid / M::NS], r0 = (bid % M::NS) * 16;
  const int j0 = row_ptr[i], nl = diag_ptr[i] - j0;
  switch (nl) { // the number of lower blocks as a constant: no guards in the MFMA chains
  case 1: slab_star_row<B, 1>(bsr, i, r0, j0, diag_ptr, col, upd_ptr, upd, Xs); break;
  case 2: slab_star_row<B, 2>(bsr, i, r0, j0, diag_ptr, col, upd_ptr, upd, Xs); break;
  case 3: slab_star_row<B, 3>(bsr, i, r0, j0, diag_ptr, col, upd_ptr, upd, Xs); break;
  case 4: slab_star_row<B, 4>(bsr, i, r0, j0, diag_ptr, col, upd_ptr, upd, Xs); break;
  default: break; // no lower block: nothing to update
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
// With scaled upper blocks (M_ij = inv(U_ii) U_ij, stored in place of U_ij
// after the factor): w_i = inv(U_ii) w_i - sum_{j>i} M_ij w_j.
template <class T, int NT>
__global__ __launch_bounds__(NT) void k_ilu_upper_split(const T *__restrict__ bsr, int b, const int *__restrict__ rows,
                                                       const int *__restrict__ row_ptr,
                                                       const int *__restrict__ diag_ptr, const int *__restrict__ col,
                                                       T *__restrict__ w, int scaled) {
  __shared__ T red[NT];
  __shared__ T ts[kSweepMaxBlock];
  const int i = rows[blockIdx.x];
  const size_t bb = size_t(b) * b;
  const int P = NT / b, r = threadIdx.x % b, part = threadIdx.x / b;
  if (scaled) {
    T c = part < P ? row_dot(bsr + diag_ptr[i] * bb, w + size_t(i) * b, b, r, part, P) : T(0);
    red[threadIdx.x] = c;
    __syncthreads();
    if (part == 0) {
      T s = T(0);
      for (int q = 0; q < P; ++q) { s += red[q * b + r]; }
      ts[r] = s;
    }
    __syncthreads();
    T acc = T(0);
    if (part < P) {
      for (int ij = diag_ptr[i] + 1; ij < row_ptr[i + 1]; ++ij) { acc += row_dot(bsr + ij * bb, w + size_t(col[ij]) * b, b, r, part, P); }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (part == 0) {
      T s = T(0);
      for (int q = 0; q < P; ++q) { s += red[q * b + r]; }
      w[size_t(i) * b + r] = ts[r] - s;
    }
    return;
  }
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
// One launch per sweep (or per factorisation) instead of one per level. The
// workgroups take their rows from a ticket counter in dependency order
// (increasing for the lower sweep and the factor, decreasing for the upper
// sweep), so every row waited on has been taken before the waiting row by a
// workgroup that is running: every wait ends whether or not the whole grid is
// resident at once (other solves may share the device), and the grid is only
// sized to what the device can hold, for speed. A workgroup takes its next
// ticket when it starts a row (the atomic's latency is hidden behind the
// row's loads); holding one ticket ahead keeps the order argument: the
// smallest unfinished ticket is always held, as current row, by a running
// workgroup whose dependencies are all finished. Each wait is nevertheless
// bounded by spin_limit and then raises kErrSpin instead of hanging. A plain
// launch is used: a cooperative launch costs about 20 us more on this stack,
// as much as a whole sweep of a small case. Data crossing workgroups is read
// and written with coherent (sc1) accesses only.
//
// ticket[0] counts the tickets taken and ticket[1] the workgroups that have
// taken their last one; the last of those resets both to zero for the next
// launch on the stream.
__device__ __forceinline__ void release_tickets(int *ticket, int grid) {
  if (threadIdx.x == 0 && atomicAdd(ticket + 1, 1) == grid - 1) {
    atomicExch(ticket, 0);
    atomicExch(ticket + 1, 0);
  }
}
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
  static constexpr int NT = B >= 64 ? 1024 : 256;
  static constexpr int P = NT / B;
  static constexpr int VB = (B + P - 1) / P;
};
// LOWER: w_i = b_i - sum_{j<i} L_ij w_j (src/block_ilu.cpp:87-99); src = b,
// dst = w, and the kernel arms `arm` (the upper sweep's output).
// UPPER: y_i = inv(U_ii) (w_i - sum_{j>i} U_ij y_j) (src/block_ilu.cpp:100-108);
// src = w, re-armed row by row after it is read, dst = y.
// SCALED (upper, blocks up to 64): the blocks right of the diagonal hold
// M_ij = inv(U_ii) U_ij, so y_i = c_i - sum_{j>i} M_ij y_j with c_i =
// inv(U_ii) w_i formed before the wait: one product after it, as in the
// lower sweep.
// Sum over each DPP row of 16 lanes (the same bits in every lane of the row).
template <class T> __device__ __forceinline__ T row16_sum(T v);
template <> __device__ __forceinline__ double row16_sum<double>(double v) {
  v = dpp_add<0xB1>(v);  // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);  // quad_perm [2,3,0,1]
  v = dpp_add<0x141>(v); // row_half_mirror
  return dpp_add<0x140>(v); // row_mirror
}
template <int CTRL> __device__ __forceinline__ float dpp_addf(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <> __device__ __forceinline__ float row16_sum<float>(float v) {
  v = dpp_addf<0xB1>(v);
  v = dpp_addf<0x4E>(v);
  v = dpp_addf<0x141>(v);
  return dpp_addf<0x140>(v);
}
// Shared memory of a dataflow sweep workgroup.
template <class T, int B, int KD> struct SweepLds {
  T xs[KD * B];
  T red[SweepShape<B>::NT];
  T ts[B];
  int next_row;
};
// The rows one workgroup of a dataflow sweep takes: workgroup vblock of a
// launch of vgrid workgroups (the ticket protocol counts vgrid exits).
template <class T, int B, int KD, bool LOWER>
__device__ __forceinline__ void sweep_sn_rows(const T *__restrict__ bsr, int nb, const int *__restrict__ row_ptr,
                                              const int *__restrict__ diag_ptr, const int *__restrict__ col,
                                              T *src, T *dst, T *arm, int *flags, int spin_limit, int scaled,
                                              const int *__restrict__ order, int *ticket, int vblock, int vgrid,
                                              SweepLds<T, B, KD> &S) {
  constexpr int NT = SweepShape<B>::NT, P = SweepShape<B>::P, VB = SweepShape<B>::VB;
  constexpr int KB = LOWER ? KD : KD + 1; // blocks held per row (upper: + inv(U_ii))
  constexpr size_t bb = size_t(B) * B;
  // 16 x 16 blocks (P == 16): the parts of a row are one DPP row of lanes
  // (part fastest) and the row's partial sums are added by DPP, with no LDS
  // round trip or workgroup barrier; otherwise rows fastest (coalesced block
  // loads; measured faster for 64 x 64) and a sum over LDS.
  constexpr bool RR = P == 16 && B == 16;
  const int tid = threadIdx.x, r = RR ? tid / P : tid % B, part = RR ? tid % P : tid / B;
  const bool worker = part < P;
  if (tid == 0) {
    const int first = atomicAdd(ticket, 1);
    S.next_row = first < nb ? order[first] : -1;
  }
  if (LOWER) {
    for (size_t t = size_t(vblock) * NT + tid; t < size_t(nb) * B; t += size_t(vgrid) * NT) {
      arm[t] = sentinel<T>();
    }
  }
  __syncthreads();
  // Rows from the ticket counter (release_tickets): ticket t is row order[t],
  // the rows of the sweep in its dependency-level order (a natural-order
  // window of tickets would hold only a few rows of one level).
  // the row index is the same in every lane: read it into a scalar register
  // (the row's pattern and block addresses then stay scalar)
  for (int i = __builtin_amdgcn_readfirstlane(S.next_row); i >= 0;) {
    // With a workgroup for every row the first tickets cover all rows, so no
    // workgroup takes a second one (and none delays its row for it).
    int ahead = nb, ahead_row = -1;
    if (tid == 0 && vgrid < nb) { ahead = atomicAdd(ticket, 1); }
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
    if (!LOWER && scaled) {
      T cacc = T(0);
#pragma unroll
      for (int v = 0; v < VB; ++v) {
        const int c = part + v * P;
        if (worker && c < B) { cacc += a[KD][v] * src[size_t(i) * B + c]; }
      }
      if constexpr (RR) {
        own = row16_sum<T>(cacc); // w_i is re-armed after the barrier below (every read of it done)
      } else {
        S.red[tid] = cacc;
        __syncthreads(); // every read of w_i is done
        if (part == 0) {
          for (int q = 0; q < P; ++q) { own += S.red[q * B + r]; }
          src[size_t(i) * B + r] = sentinel<T>();
        }
        __syncthreads(); // red is free again
      }
    } else if (part == 0) {
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
      S.xs[t] = x;
    }
    __syncthreads();
    // the ticket has arrived by now: look up its row while this one is summed
    if (tid == 0 && ahead < nb) { ahead_row = order[ahead]; }
    if (RR && !LOWER && scaled && part == 0) { src[size_t(i) * B + r] = sentinel<T>(); }
    T acc = T(0);
#pragma unroll
    for (int d = 0; d < KD; ++d) {
      if (d < nd) {
#pragma unroll
        for (int v = 0; v < VB; ++v) {
          const int c = part + v * P;
          if (c < B) { acc += a[d][v] * S.xs[d * B + c]; }
        }
      }
    }
    T sum = T(0);
    if constexpr (RR) {
      sum = row16_sum<T>(acc);
    } else {
      S.red[tid] = acc;
      __syncthreads();
      if (part == 0) {
        for (int q = 0; q < P; ++q) { sum += S.red[q * B + r]; }
      }
    }
    if (part == 0) {
      if (LOWER || scaled) { st_coherent(dst + size_t(i) * B + r, own - sum); }
      else { S.ts[r] = own - sum; }
    }
    if (!LOWER && !scaled) {
      __syncthreads();
      acc = T(0);
#pragma unroll
      for (int v = 0; v < VB; ++v) {
        const int c = part + v * P;
        if (c < B) { acc += a[KD][v] * S.ts[c]; }
      }
      T sum2 = T(0);
      if constexpr (RR) {
        sum2 = row16_sum<T>(acc);
      } else {
        S.red[tid] = acc;
        __syncthreads();
        if (part == 0) {
          for (int q = 0; q < P; ++q) { sum2 += S.red[q * B + r]; }
        }
      }
      if (part == 0) { st_coherent(dst + size_t(i) * B + r, sum2); }
    }
    // every thread read S.next_row at the top of this row, before the barrier
    // after the dependency wait
    if (tid == 0) { S.next_row = ahead_row; }
    __syncthreads();
    i = __builtin_amdgcn_readfirstlane(S.next_row);
  }
  release_tickets(ticket, vgrid);
}
template <class T, int B, int KD, bool LOWER>
__global__ __launch_bounds__(SweepShape<B>::NT) void k_ilu_sweep_sn(const T *__restrict__ bsr, int nb,
                                                                    const int *__restrict__ row_ptr,
                                                                    const int *__restrict__ diag_ptr,
                                                                    const int *__restrict__ col, T *src, T *dst,
                                                                    T *arm, int *flags, int spin_limit, int scaled,
                                                                    const int *__restrict__ order, int *ticket) {
  __shared__ SweepLds<T, B, KD> S;
  sweep_sn_rows<T, B, KD, LOWER>(bsr, nb, row_ptr, diag_ptr, col, src, dst, arm, flags, spin_limit, scaled, order,
                                 ticket, blockIdx.x, gridDim.x, S);
}
// The sweeps of every slot of a grouped Arnoldi step (z = M^-1 v_j: lower
// src v_j, dst mid, arm z; upper src mid, dst z). Slot i has
// sweep_off[i+1] - sweep_off[i] virtual workgroups on its own ticket
// counters; a workgroup of the persistent grid finishes each virtual
// workgroup (all rows it took) before the next, so the order argument of the
// single-solve sweep holds per slot.
template <class T, int B, int KD, bool LOWER>
__global__ __launch_bounds__(SweepShape<B>::NT) void k_group_sweep(const GroupTable *__restrict__ tab,
                                                                   int spin_limit) {
  __shared__ SweepLds<T, B, KD> S;
  const int count = tab->count, total = tab->sweep_off[count];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int i = group_slot(tab->sweep_off, count, v);
    const ArnoldiSlot &sl = tab->slot[i];
    if (sl.cycle->active == 0) { continue; } // the same for the whole workgroup
    const int first = tab->sweep_off[i];
    T *vj = static_cast<T *>(sl.V) + size_t(sl.cycle->count) * sl.N;
    T *mid = static_cast<T *>(sl.mid), *z = static_cast<T *>(sl.z);
    __syncthreads(); // S is free again
    sweep_sn_rows<T, B, KD, LOWER>(static_cast<const T *>(sl.bsr), sl.nb, sl.row_ptr, sl.diag_ptr, sl.col,
                                   LOWER ? vj : mid, LOWER ? mid : z, z, sl.flags, spin_limit,
                                   LOWER ? 0 : sl.scaled, LOWER ? sl.lower_rows : sl.upper_rows, sl.ticket,
                                   v - first, tab->sweep_off[i + 1] - first, S);
  }
}
template <class T> __global__ void k_arm(T *v, size_t n) {
  for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += size_t(gridDim.x) * blockDim.x) {
    v[t] = sentinel<T>();
  }
}

// Dataflow sweeps for large blocks (b > 100, whose rows do not fit in
// registers): same row order, sentinel polling and coherent exchange as
// k_ilu_sweep_sn, with the block GEMVs streamed from memory after the wait by
// 1024 threads (thread (r, part) = (tid % b, tid / b), columns part + v P).
// The first PRE columns of each block are loaded before the wait.
// sum over columns c = c0, c0 + P, ... < b of a[c b] x[c], in column order,
// with 16 loads in flight (the memory-level parallelism a CU needs to stream
// at its share of the HBM bandwidth).
template <class T>
__device__ __forceinline__ T stream_dot(const T *__restrict__ a, const T *x, int b, int c0, int P) {
  constexpr int U = 16;
  T acc = T(0);
  int c = c0;
  for (; c + (U - 1) * P < b; c += U * P) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { v[u] = a[size_t(c + u * P) * b]; }
#pragma unroll
    for (int u = 0; u < U; ++u) { acc += v[u] * x[c + u * P]; }
  }
  for (; c < b; c += P) { acc += a[size_t(c) * b] * x[c]; }
  return acc;
}
template <class T, bool LOWER>
__global__ __launch_bounds__(1024) void k_ilu_sweep_stream(const T *__restrict__ bsr, int b, int nb, int kd,
                                                           const int *__restrict__ row_ptr,
                                                           const int *__restrict__ diag_ptr,
                                                           const int *__restrict__ col, T *src, T *dst, T *arm,
                                                           int *flags, int spin_limit,
                                                           const int *__restrict__ order, int *ticket) {
  constexpr int NT = 1024, PRE = 4, KDM = 4; // cell blocks: at most 4 neighbours a triangle
  // 16-byte aligned: the dynamic LDS follows the static next_row, and its
  // doubles must stay 8-byte aligned for ds_read_b64
  extern __shared__ __attribute__((aligned(16))) unsigned char stream_lds[];
  __shared__ int next_row;
  T *xs = reinterpret_cast<T *>(stream_lds); // kd x b dependency values
  T *red = xs + size_t(kd) * b;               // NT partial sums
  T *ts = red + NT;                           // b
  const int tid = threadIdx.x, P = NT / b, r = tid % b, part = tid / b;
  const bool worker = part < P;
  const size_t bb = size_t(b) * b;
  if (tid == 0) {
    const int first = atomicAdd(ticket, 1);
    next_row = first < nb ? order[first] : -1;
  }
  if (LOWER) {
    for (size_t t = size_t(blockIdx.x) * NT + tid; t < size_t(nb) * b; t += size_t(gridDim.x) * NT) {
      arm[t] = sentinel<T>();
    }
  }
  __syncthreads();
  // rows from the ticket counter in level order, as in k_ilu_sweep_sn
  // the row index is the same in every lane: read it into a scalar register
  // (the row's pattern and block addresses then stay scalar)
  for (int i = __builtin_amdgcn_readfirstlane(next_row); i >= 0;) {
    // With a workgroup for every row the first tickets cover all rows, so no
    // workgroup takes a second one (and none delays its row for it).
    int ahead = nb, ahead_row = -1;
    if (tid == 0 && int(gridDim.x) < nb) { ahead = atomicAdd(ticket, 1); }
    const int j0 = LOWER ? row_ptr[i] : diag_ptr[i] + 1, j1 = LOWER ? diag_ptr[i] : row_ptr[i + 1];
    const int nd = j1 - j0;
    T pre[KDM][PRE];
#pragma unroll
    for (int d = 0; d < KDM; ++d) {
#pragma unroll
      for (int v = 0; v < PRE; ++v) {
        const int c = part + v * P;
        pre[d][v] = (d < nd && worker && c < b) ? bsr[size_t(j0 + d) * bb + size_t(c) * b + r] : T(0);
      }
    }
    T own = T(0);
    if (part == 0) {
      own = src[size_t(i) * b + r];
      if (!LOWER) { src[size_t(i) * b + r] = sentinel<T>(); }
    }
    for (int t = tid; t < nd * b; t += NT) {
      const T *p = dst + size_t(col[j0 + t / b]) * b + t % b;
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
    if (tid == 0 && ahead < nb) { ahead_row = order[ahead]; }
    T acc = T(0);
    if (worker) {
#pragma unroll
      for (int d = 0; d < KDM; ++d) {
        if (d < nd) {
          const T *a = bsr + size_t(j0 + d) * bb + r;
          const T *x = xs + d * b;
#pragma unroll
          for (int v = 0; v < PRE; ++v) {
            const int c = part + v * P;
            if (c < b) { acc += pre[d][v] * x[c]; }
          }
          acc += stream_dot<T>(a, x, b, part + PRE * P, P);
        }
      }
    }
    red[tid] = acc;
    __syncthreads();
    if (LOWER) {
      if (part == 0) {
        T sum = T(0);
        for (int q = 0; q < P; ++q) { sum += red[q * b + r]; }
        st_coherent(dst + size_t(i) * b + r, own - sum);
      }
    } else {
      if (part == 0) {
        T sum = T(0);
        for (int q = 0; q < P; ++q) { sum += red[q * b + r]; }
        ts[r] = own - sum;
      }
      __syncthreads();
      acc = T(0);
      if (worker) {
        acc = stream_dot<T>(bsr + size_t(diag_ptr[i]) * bb + r, ts, b, part, P);
      }
      red[tid] = acc;
      __syncthreads();
      if (part == 0) {
        T sum = T(0);
        for (int q = 0; q < P; ++q) { sum += red[q * b + r]; }
        st_coherent(dst + size_t(i) * b + r, sum);
      }
    }
    if (tid == 0) { next_row = ahead_row; } // read by every thread before the barrier after the wait
    __syncthreads();
    i = __builtin_amdgcn_readfirstlane(next_row);
  }
  release_tickets(ticket, int(gridDim.x));
}

// Dataflow sweeps for 144 x 144 and 196 x 196 blocks with scaled upper
// blocks, each block row split over `slices` workgroups by rows of its
// blocks (one workgroup streams a block row at a few tens of GB/s, so the
// level chain, not HBM, bounded the unsplit form): ticket t is slice
// t % slices of row order[t / slices], which forms rows [r0, r0 + rs) of the
// row's output from the whole dependency vectors, so the slices of a row
// never wait on each other and the ticket order argument holds.
// LOWER: y_i = b_i - sum L_ij y_j (src = b, dst = w; arms `arm`, the
// upper sweep's output). UPPER (scaled): y_i = c_i - sum M_ij y_j, with
// c_i = inv(U_ii) w_i given (CDIAG false: src = c, formed between the two
// sweeps by a block-diagonal product) or formed here before the wait from
// the whole w_i (CDIAG: src = w, complete: the lower sweep is an earlier
// launch), dst = y; it arms `arm`, the next application's intermediate
// vector (the intermediate vectors alternate: a slice cannot re-arm w_i
// while the others still read it).
template <class T, bool LOWER, bool CDIAG = !LOWER>
__global__ __launch_bounds__(1024) void k_ilu_sweep_slice(const T *__restrict__ bsr, int b, int nb, int kd,
                                                          int slices, const int *__restrict__ row_ptr,
                                                          const int *__restrict__ diag_ptr,
                                                          const int *__restrict__ col, const T *src, T *dst,
                                                          T *arm, long narm, int *flags, int spin_limit,
                                                          const int *__restrict__ order, int *ticket) {
  // cell blocks: at most 4 neighbours a triangle. PRE: columns of each
  // dependency block loaded before the wait (5 covers a 196-row slice of 8:
  // no load after the wait; the FP64 upper sweep keeps 4, as 5 spills there)
  constexpr int NT = 1024, PRE = (sizeof(T) == 8 && CDIAG) ? 4 : 5, KDM = 4;
  extern __shared__ __attribute__((aligned(16))) unsigned char slice_lds[];
  __shared__ int next_t;
  T *xs = reinterpret_cast<T *>(slice_lds); // kd x b dependency values
  T *red = xs + size_t(kd) * b;             // NT partial sums
  T *ws = red + NT;                         // b: w_i (upper)
  const int rs = (b + slices - 1) / slices, total = nb * slices;
  const int tid = threadIdx.x, P = NT / rs, r = tid % rs, part = tid / rs;
  const size_t bb = size_t(b) * b;
  if (tid == 0) { next_t = atomicAdd(ticket, 1); }
  for (size_t t = size_t(blockIdx.x) * NT + tid; t < size_t(narm); t += size_t(gridDim.x) * NT) {
    arm[t] = sentinel<T>();
  }
  __syncthreads();
  for (int tk = __builtin_amdgcn_readfirstlane(next_t); tk < total;) {
    int ahead = total;
    if (tid == 0 && int(gridDim.x) < total) { ahead = atomicAdd(ticket, 1); }
    const int i = order[tk / slices], rr = (tk % slices) * rs + r;
    const bool worker = part < P && rr < b;
    const int j0 = LOWER ? row_ptr[i] : diag_ptr[i] + 1, j1 = LOWER ? diag_ptr[i] : row_ptr[i + 1];
    const int nd = j1 - j0;
    T pre[KDM][PRE];
#pragma unroll
    for (int d = 0; d < KDM; ++d) {
#pragma unroll
      for (int v = 0; v < PRE; ++v) {
        const int c = part + v * P;
        pre[d][v] = (d < nd && worker && c < b) ? bsr[size_t(j0 + d) * bb + size_t(c) * b + rr] : T(0);
      }
    }
    T own = T(0);
    if (!CDIAG) {
      if (part == 0 && rr < b) { own = src[size_t(i) * b + rr]; }
    } else {
      for (int c = tid; c < b; c += NT) { ws[c] = src[size_t(i) * b + c]; }
      __syncthreads();
      red[tid] = worker ? stream_dot<T>(bsr + size_t(diag_ptr[i]) * bb + rr, ws, b, part, P) : T(0);
      __syncthreads();
      if (part == 0 && rr < b) {
        for (int q = 0; q < P; ++q) { own += red[q * rs + r]; }
      }
      // red is written again only after the barrier that ends the wait
    }
    for (int t = tid; t < nd * b; t += NT) {
      const T *p = dst + size_t(col[j0 + t / b]) * b + t % b;
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
    if (worker) {
#pragma unroll
      for (int d = 0; d < KDM; ++d) {
        if (d < nd) {
          const T *x = xs + d * b;
#pragma unroll
          for (int v = 0; v < PRE; ++v) {
            const int c = part + v * P;
            if (c < b) { acc += pre[d][v] * x[c]; }
          }
          acc += stream_dot<T>(bsr + size_t(j0 + d) * bb + rr, x, b, part + PRE * P, P);
        }
      }
    }
    red[tid] = acc;
    __syncthreads();
    if (part == 0 && rr < b) {
      T sum = T(0);
      for (int q = 0; q < P; ++q) { sum += red[q * rs + r]; }
      st_coherent(dst + size_t(i) * b + rr, own - sum);
    }
    if (tid == 0) { next_t = ahead; } // read by every thread before the barrier after the wait
    __syncthreads();
    tk = __builtin_amdgcn_readfirstlane(next_t);
  }
  release_tickets(ticket, int(gridDim.x));
}

// Dataflow ILU(0) factor: factor_row per row, rows taken from the ticket
// counter in level order (order = the rows by dependency level), waiting on
// the rows of the lower pattern.
template <int B>
__global__ __launch_bounds__(Shape<B>::NT) void k_ilu_factor_df(double *bsr, int nb, const int *__restrict__ row_ptr,
                                                                const int *__restrict__ diag_ptr,
                                                                const int *__restrict__ col,
                                                                const int *__restrict__ upd_ptr,
                                                                const int *__restrict__ upd, int *done, int epoch,
                                                                int *flags, int spin_limit,
                                                                const int *__restrict__ order, int *ticket) {
  extern __shared__ double sm[];
  __shared__ int row;
  for (;;) {
    __syncthreads(); // row is no longer read
    if (threadIdx.x == 0) {
      const int t = atomicAdd(ticket, 1);
      row = t < nb ? order[t] : -1;
    }
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(row);
    if (i < 0) { break; }
    wait_rows(col, row_ptr[i], diag_ptr[i], done, epoch, spin_limit, flags);
    if (factor_row<B, true>(bsr, i, row_ptr, diag_ptr, col, upd_ptr, upd, sm) && threadIdx.x == 0) {
      atomicOr(flags, kErrSingular);
    }
    publish_row(done, i, epoch);
  }
  release_tickets(ticket, int(gridDim.x));
}

// Dependencies of a row program read from the published copy of the factor
// (pub, armed with the sentinel before the launch): every value is polled
// with coherent loads until it is no longer the sentinel, so the data is its
// own completion flag. Blocks b0 .. b1 - 1 of the program's LDS image.
template <int B>
__device__ __forceinline__ void resident_poll(const double *pub, const int *P, double *sm, int b0, int b1, int *flags,
                                              int spin_limit) {
  constexpr int bb = B * B, NT = Shape<B>::NT, U = 16;
  const int t = threadIdx.x, nrow = P[2];
  const int base = b0 * bb, total = (b1 - b0) * bb;
  bool waiting = true; // false once a wait has exceeded its bound
  for (int e0 = 0; e0 < total; e0 += U * NT) {
    double tmp[U];
    const double *src[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * NT + t;
      src[u] = nullptr;
      tmp[u] = 0.;
      if (e < total) {
        const int blk = (base + e) / bb, o = base + e - blk * bb;
        src[u] = pub + size_t(P[5 + blk - nrow]) * bb + o;
        tmp[u] = ld_coherent(src[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int spins = 0;
      while (waiting && src[u] && armed(tmp[u])) {
        if (++spins > spin_limit) {
          atomicOr(flags, kErrSpin);
          waiting = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        tmp[u] = ld_coherent(src[u]);
      }
      const int e = e0 + u * NT + t;
      if (e < total) { sm[base + e] = tmp[u]; }
    }
  }
}
// Dataflow form of the row-resident factor: one launch for all rows. A
// workgroup takes the next row program in level order from a ticket counter,
// loads the row's own blocks (written before the launch), reads the blocks of
// the earlier rows it depends on from pub as they are published
// (resident_poll), factors the row exactly as k_ilu_factor_resident does,
// writes it to bsr and publishes its diagonal inverse and upper blocks to pub
// with coherent stores. Rows are taken in dependency order by running
// workgroups, so every wait ends whether or not the whole grid is resident
// at once (ticket counters as in release_tickets).
template <int B>
__global__ __launch_bounds__(Shape<B>::NT) void k_ilu_factor_resident_df(double *__restrict__ bsr, double *pub,
                                                                         const int *__restrict__ prog_off,
                                                                         const int *__restrict__ prog, int nrows,
                                                                         int *ticket, int *flags, int spin_limit) {
  extern __shared__ double sm[];
  __shared__ int P[kIluProgMax];
  __shared__ int row;
  constexpr int bb = B * B, NT = Shape<B>::NT;
  const int t = threadIdx.x;
  for (;;) {
    __syncthreads(); // the previous row's LDS image and index are no longer read
    if (t == 0) { row = atomicAdd(ticket, 1); }
    __syncthreads();
    const int r = __builtin_amdgcn_readfirstlane(row); // uniform: keep it scalar
    if (r >= nrows) { break; }
    const int *pg = prog + prog_off[r];
    for (int e = t; e < kIluProgMax; e += NT) { P[e] = pg[e]; }
    __syncthreads();
    const int j0 = P[1], nrow = P[2], dpos = P[3];
    resident_load<B, false>(bsr, P, sm, 0, nrow);
    resident_poll<B>(pub, P, sm, nrow, nrow + P[4], flags, spin_limit);
    __syncthreads();
    const bool singular = resident_factor<B>(P, sm);
    for (int e = t; e < nrow * bb; e += NT) {
      bsr[size_t(j0) * bb + e] = sm[e];
      if (e >= dpos * bb) { st_coherent(pub + size_t(j0) * bb + e, sm[e]); }
    }
    if (singular && t == 0) { atomicOr(flags, kErrSingular); }
  }
  release_tickets(ticket, int(gridDim.x));
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
#define EG_ALLOW_R(B)                                                                              \
  allow_lds(reinterpret_cast<const void *>(k_ilu_factor_resident<B>));                             \
  allow_lds(reinterpret_cast<const void *>(k_ilu_factor_resident_df<B>));
  EG_ALLOW_R(4) EG_ALLOW_R(8) EG_ALLOW_R(12) EG_ALLOW_R(16) EG_ALLOW_R(20) EG_ALLOW_R(24) EG_ALLOW_R(28)
  EG_ALLOW_R(32) EG_ALLOW_R(36)
#undef EG_ALLOW_R
#define EG_ALLOW_S(B)                                                                              \
  allow_lds(reinterpret_cast<const void *>(k_ilu_slab<B>));                                        \
  allow_lds(reinterpret_cast<const void *>(k_ilu_slab_m<B>));
  EG_ALLOW_S(64) EG_ALLOW_S(100) EG_ALLOW_S(144) EG_ALLOW_S(196) EG_ALLOW_S(256) EG_ALLOW_S(324)
#undef EG_ALLOW_S
  allow_lds(reinterpret_cast<const void *>(k_ilu_scale_upper<64>));
  allow_lds(reinterpret_cast<const void *>(k_ilu_slab_star<64>));
  return 0;
}
bool ilu_resident_ok(int b) { return b % 4 == 0 && b >= 4 && b <= 36; }
size_t ilu_resident_gj_lds(int b) {
  size_t n = 0;
  EG_DISPATCH_B_SMALL(b, (n = resident_gj_lds<B>()));
  return n;
}
void launch_ilu_factor_resident(double *bsr, int b, const int *prog_off, const int *prog, int nrows, size_t lds,
                                int *flags, hipStream_t s) {
  EG_DISPATCH_B_SMALL(b, (k_ilu_factor_resident<B><<<nrows, Shape<B>::NT, lds, s>>>(bsr, prog_off, prog, flags)));
}
#define EG_DISPATCH_SLAB(b, CALL)                                                                  \
  switch (b) {                                                                                     \
  case 64: { constexpr int B = 64; CALL; } break;                                                  \
  case 100: { constexpr int B = 100; CALL; } break;                                                \
  case 144: { constexpr int B = 144; CALL; } break;                                                \
  case 196: { constexpr int B = 196; CALL; } break;                                                \
  case 256: { constexpr int B = 256; CALL; } break;                                                \
  case 324: { constexpr int B = 324; CALL; } break;                                                \
  default: break;                                                                                  \
  }
bool ilu_scale_upper_ok(int b) {
  return b % 4 == 0 && b >= 4 && (b <= 36 || b == 64 || b == 100 || b == 144 || b == 196);
}
void launch_ilu_scale_upper(double *bsr, int b, const int *slot, const int *dslot, int count, hipStream_t s) {
  if (count <= 0) { return; }
  if (b == 100) {
    k_ilu_scale_upper_mfma<100><<<count * ScaleM<100>::CG, 1024, 0, s>>>(bsr, slot, dslot);
    return;
  }
  if (b == 144) {
    k_ilu_scale_upper_mfma<144><<<count * ScaleM<144>::CG, 1024, 0, s>>>(bsr, slot, dslot);
    return;
  }
  if (b == 196) {
    k_ilu_scale_upper_mfma<196><<<count * ScaleM<196>::CG, 1024, 0, s>>>(bsr, slot, dslot);
    return;
  }
  const size_t lds = 2 * size_t(b) * b * sizeof(double);
  switch (b) {
#define EG_SC(BB)                                                                                  \
  case BB: k_ilu_scale_upper<BB><<<count, 256, lds, s>>>(bsr, slot, dslot); return;
    EG_SC(4) EG_SC(8) EG_SC(12) EG_SC(16) EG_SC(20) EG_SC(24) EG_SC(28) EG_SC(32) EG_SC(36) EG_SC(64)
#undef EG_SC
  default: break;
  }
  throw std::runtime_error("scaled upper blocks: unsupported block size");
}
bool ilu_slab_ok(int b) { return b == 64 || b == 100 || b == 144 || b == 196 || b == 256 || b == 324; }
int ilu_star_max_lower() { return kStarLower; }
void launch_ilu_slab(double *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                     const int *col, const int *upd_ptr, const int *upd, bool star, hipStream_t s) {
  const char *mode = std::getenv("EG_SLAB"); // read per call: a solver may be set up under another value
  const bool vector = mode && std::string(mode) == "vector";
  if (vector) {
    EG_DISPATCH_SLAB(b, (k_ilu_slab<B><<<nrows * Slab<B>::NS, 256, Slab<B>::lds(), s>>>(bsr, rows, row_ptr, diag_ptr,
                                                                                       col, upd_ptr, upd)));
  } else {
    if (star && b == 64) {
      k_ilu_slab_star<64><<<nrows * SlabM<64>::NS, 256, kStarLower * SlabM<64>::lds(), s>>>(
          bsr, rows, row_ptr, diag_ptr, col, upd_ptr, upd);
      return;
    }
    EG_DISPATCH_SLAB(b, (k_ilu_slab_m<B><<<nrows * SlabV<B, double>::NS, 256, SlabV<B, double>::lds(), s>>>(
                            bsr, rows, row_ptr, diag_ptr, col, upd_ptr, upd)));
  }
}
bool ilu_factor32_ok(int b) { return b == 100 || b == 144 || b == 196; }
void launch_ilu_slab32(float *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                       const int *col, const int *upd_ptr, const int *upd, hipStream_t s) {
#define EG_SLAB32(BB)                                                                              \
  case BB:                                                                                         \
    k_ilu_slab_m<BB, float><<<nrows * SlabV<BB, float>::NS, 256, SlabV<BB, float>::lds(), s>>>(   \
        bsr, rows, row_ptr, diag_ptr, col, upd_ptr, upd);                                          \
    return;
  switch (b) { EG_SLAB32(100) EG_SLAB32(144) EG_SLAB32(196) default: break; }
#undef EG_SLAB32
  throw std::runtime_error("FP32 ILU factor: unsupported block size");
}
// The panel-blocked MFMA inverse for 100 / 144 / 196 blocks unless
// EG_GJ_MFMA=0 (then the lane-per-row one), read per call.
static bool gjm_on() {
  const char *e = std::getenv("EG_GJ_MFMA");
  return !(e && std::string(e) == "0");
}
void launch_block_inverse32(float *blocks, int b, int count, const int *index, int *flags, hipStream_t s) {
  if (gjm_on()) {
    switch (b) {
    case 100: k_block_inverse_m<100, float><<<count, GjM<100, float>::threads(), 0, s>>>(blocks, index, flags); return;
    case 144: k_block_inverse_m<144, float><<<count, GjM<144, float>::threads(), 0, s>>>(blocks, index, flags); return;
    case 196: k_block_inverse_m<196, float><<<count, GjM<196, float>::threads(), 0, s>>>(blocks, index, flags); return;
    default: break;
    }
  }
  switch (b) {
  case 100: k_block_inverse_w<100, float><<<count, gj2_threads<100>(), 0, s>>>(blocks, index, flags); return;
  case 144: k_block_inverse_w<144, float><<<count, gj2_threads<144>(), 0, s>>>(blocks, index, flags); return;
  case 196: k_block_inverse_w<196, float><<<count, gj2_threads<196>(), 0, s>>>(blocks, index, flags); return;
  default: break;
  }
  throw std::runtime_error("FP32 block inverse: unsupported block size");
}
void launch_ilu_scale_upper32(float *bsr, int b, const int *slot, const int *dslot, int count, hipStream_t s) {
  if (count <= 0) { return; }
  switch (b) {
  case 100: k_ilu_scale_upper_mfma<100, float><<<count * ScaleM<100>::CG, 1024, 0, s>>>(bsr, slot, dslot); return;
  case 144: k_ilu_scale_upper_mfma<144, float><<<count * ScaleM<144>::CG, 1024, 0, s>>>(bsr, slot, dslot); return;
  case 196: k_ilu_scale_upper_mfma<196, float><<<count * ScaleM<196>::CG, 1024, 0, s>>>(bsr, slot, dslot); return;
  default: break;
  }
  throw std::runtime_error("FP32 scaled upper blocks: unsupported block size");
}
void launch_block_inverse_fast(double *blocks, int b, int count, const int *index, int *flags, hipStream_t s) {
  if (gjm_on()) {
    switch (b) {
    case 100: k_block_inverse_m<100><<<count, GjM<100, double>::threads(), 0, s>>>(blocks, index, flags); return;
    case 144: k_block_inverse_m<144><<<count, GjM<144, double>::threads(), 0, s>>>(blocks, index, flags); return;
    case 196: k_block_inverse_m<196><<<count, GjM<196, double>::threads(), 0, s>>>(blocks, index, flags); return;
    default: break;
    }
  }
  switch (b) {
  case 64: k_block_inverse_p<64, 8><<<count, 512, 0, s>>>(blocks, index, flags); return;
  case 100: k_block_inverse_w<100><<<count, gj2_threads<100>(), 0, s>>>(blocks, index, flags); return;
  case 144: k_block_inverse_w<144><<<count, gj2_threads<144>(), 0, s>>>(blocks, index, flags); return;
  case 196: k_block_inverse_w<196><<<count, gj2_threads<196>(), 0, s>>>(blocks, index, flags); return;
  default: break;
  }
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
                            const int *diag_ptr, const int *col, T *w, int threads, bool scaled, hipStream_t s) {
  if (threads == 0) { threads = b <= 64 ? 256 : 1024; }
  const int sc = scaled ? 1 : 0;
  if (threads == 256 && b <= 256) {
    k_ilu_upper_split<T, 256><<<nrows, 256, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, w, sc);
  } else {
    k_ilu_upper_split<T, 1024><<<nrows, 1024, 0, s>>>(bsr, b, rows, row_ptr, diag_ptr, col, w, sc);
  }
}
template <class T>
bool launch_ilu_sweep_df(bool lower, const T *bsr, int b, int nb, int max_deps, const int *row_ptr,
                         const int *diag_ptr, const int *col, T *src, T *dst, T *arm, int *flags, bool scaled,
                         const int *order, int *ticket, int max_grid, hipStream_t s) {
  const int limit = max_grid > 0 ? std::min(nb, max_grid) : nb;
  if (max_deps > 8) { return false; }
  const int kd = max_deps <= 4 ? 4 : 8;
  bool ok = false, known = false;
  auto go = [&](const void *fn, int threads) {
    known = true;
    // more workgroups than the device holds at once would only wait for a slot
    const int cap = coresident(fn, threads, 0);
    const int grid = std::max(1, std::min(limit, cap > 0 ? cap : limit));
    int spin = kSpinLimit, sc = scaled ? 1 : 0;
    void *args[] = {&bsr, &nb, &row_ptr, &diag_ptr, &col, &src, &dst, &arm, &flags, &spin, &sc, &order, &ticket};
    ok = hipLaunchKernel(fn, dim3(grid), dim3(threads), args, 0, s) == hipSuccess;
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
  if (!known && b > 100 && b <= 1024 && kd == 4) {
    // large cell blocks: streamed GEMVs
    known = true;
    const void *fn = lower ? reinterpret_cast<const void *>(k_ilu_sweep_stream<T, true>)
                           : reinterpret_cast<const void *>(k_ilu_sweep_stream<T, false>);
    const size_t lds = (size_t(kd) * b + 1024 + b) * sizeof(T);
    const int cap = coresident(fn, 1024, lds);
    const int grid = std::max(1, std::min(limit, cap > 0 ? cap : limit));
    int spin = kSpinLimit, bv = b, kdv = kd;
    void *args[] = {&bsr, &bv, &nb, &kdv, &row_ptr, &diag_ptr, &col, &src, &dst, &arm, &flags, &spin, &order,
                    &ticket};
    ok = hipLaunchKernel(fn, dim3(grid), dim3(1024), args, lds, s) == hipSuccess;
    if (!ok) { (void)hipGetLastError(); }
  }
  return known && ok;
}
template <class T> bool launch_group_sweep(bool lower, int b, const GroupTable *tab, int grid, hipStream_t s) {
  bool ok = false;
  EG_DISPATCH_B_SMALL(b, ({
    const int spin = kSpinLimit;
    if (lower) { k_group_sweep<T, B, 8, true><<<grid, SweepShape<B>::NT, 0, s>>>(tab, spin); }
    else { k_group_sweep<T, B, 8, false><<<grid, SweepShape<B>::NT, 0, s>>>(tab, spin); }
    ok = hipGetLastError() == hipSuccess;
  }));
  return ok;
}
template <class T> int group_sweep_capacity(int b) {
  int cap = 0;
  EG_DISPATCH_B_SMALL(b, ({
    cap = std::min(coresident(reinterpret_cast<const void *>(k_group_sweep<T, B, 8, true>), SweepShape<B>::NT, 0),
                   coresident(reinterpret_cast<const void *>(k_group_sweep<T, B, 8, false>), SweepShape<B>::NT, 0));
  }));
  return cap;
}
template bool launch_group_sweep<double>(bool, int, const GroupTable *, int, hipStream_t);
template bool launch_group_sweep<float>(bool, int, const GroupTable *, int, hipStream_t);
template int group_sweep_capacity<double>(int);
template int group_sweep_capacity<float>(int);
bool ilu_sweep_slice_ok(int b, int max_deps) { return (b == 100 || b == 144 || b == 196) && max_deps <= 4; }
template <class T>
bool launch_ilu_sweep_slice(bool lower, const T *bsr, int b, int nb, int slices, const int *row_ptr,
                            const int *diag_ptr, const int *col, const T *src, T *dst, T *arm, long narm, int *flags,
                            const int *order, int *ticket, int max_grid, hipStream_t s, bool given_c) {
  if (slices < 1 || slices > 8) { return false; }
  const int kd = 4, total = nb * slices;
  const void *fn = lower     ? reinterpret_cast<const void *>(k_ilu_sweep_slice<T, true>)
                   : given_c ? reinterpret_cast<const void *>(k_ilu_sweep_slice<T, false, false>)
                             : reinterpret_cast<const void *>(k_ilu_sweep_slice<T, false, true>);
  const size_t lds = (size_t(kd) * b + 1024 + b) * sizeof(T);
  const int cap = coresident(fn, 1024, lds);
  int grid = max_grid > 0 ? std::min(total, max_grid * slices) : total;
  grid = std::max(1, std::min(grid, cap > 0 ? cap : grid));
  int spin = kSpinLimit, bv = b, kdv = kd, sl = slices;
  void *args[] = {&bsr, &bv, &nb, &kdv, &sl, &row_ptr, &diag_ptr, &col, &src, &dst, &arm, &narm, &flags, &spin,
                  &order, &ticket};
  const bool ok = hipLaunchKernel(fn, dim3(grid), dim3(1024), args, lds, s) == hipSuccess;
  if (!ok) { (void)hipGetLastError(); }
  return ok;
}
template bool launch_ilu_sweep_slice<double>(bool, const double *, int, int, int, const int *, const int *,
                                             const int *, const double *, double *, double *, long, int *,
                                             const int *, int *, int, hipStream_t, bool);
template bool launch_ilu_sweep_slice<float>(bool, const float *, int, int, int, const int *, const int *,
                                            const int *, const float *, float *, float *, long, int *, const int *,
                                            int *, int, hipStream_t, bool);
template <class T> void launch_arm(T *v, size_t n, hipStream_t s) {
  k_arm<T><<<int(std::min<size_t>((n + 255) / 256, 4096)), 256, 0, s>>>(v, n);
}
bool launch_ilu_factor_df(double *bsr, int b, int nb, const int *row_ptr, const int *diag_ptr, const int *col,
                          const int *upd_ptr, const int *upd, int *done, int epoch, const int *order, int *ticket,
                          int *flags, int max_grid, hipStream_t s) {
  if (!fast_dense_ok(b)) { return false; }
  bool ok = false;
  EG_DISPATCH_B(b, ({
    const void *fn = reinterpret_cast<const void *>(k_ilu_factor_df<B>);
    const size_t lds = factor_lds<B>();
    const int cap = coresident(fn, Shape<B>::NT, lds);
    if (cap > 0) {
      const int grid = std::min(max_grid > 0 ? std::min(nb, max_grid) : nb, cap);
      int spin = kSpinLimit;
      void *args[] = {&bsr, &nb, &row_ptr, &diag_ptr, &col, &upd_ptr, &upd, &done, &epoch, &flags, &spin, &order,
                      &ticket};
      ok = hipLaunchKernel(fn, dim3(grid), dim3(Shape<B>::NT), args, lds, s) == hipSuccess;
      if (!ok) { (void)hipGetLastError(); }
    }
  }));
  return ok;
}

bool launch_ilu_factor_resident_df(double *bsr, double *pub, int b, const int *prog_off, const int *prog, int nrows,
                                   int *ticket, size_t lds, int *flags, int max_grid, hipStream_t s) {
  if (!ilu_resident_ok(b)) { return false; }
  bool ok = false;
  EG_DISPATCH_B_SMALL(b, ({
    const void *fn = reinterpret_cast<const void *>(k_ilu_factor_resident_df<B>);
    const int cap = coresident(fn, Shape<B>::NT, lds);
    if (cap > 0) {
      const int grid = std::min(max_grid > 0 ? std::min(nrows, max_grid) : nrows, cap);
      int spin = kSpinLimit;
      void *args[] = {&bsr, &pub, &prog_off, &prog, &nrows, &ticket, &flags, &spin};
      ok = hipLaunchKernel(fn, dim3(grid), dim3(Shape<B>::NT), args, lds, s) == hipSuccess;
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
                                             const int *, double *, int, bool, hipStream_t);
template void launch_ilu_upper_level<float>(const float *, int, const int *, int, const int *, const int *,
                                            const int *, float *, int, bool, hipStream_t);
template bool launch_ilu_sweep_df<double>(bool, const double *, int, int, int, const int *, const int *, const int *,
                                          double *, double *, double *, int *, bool, const int *, int *, int,
                                          hipStream_t);
template bool launch_ilu_sweep_df<float>(bool, const float *, int, int, int, const int *, const int *, const int *,
                                         float *, float *, float *, int *, bool, const int *, int *, int,
                                         hipStream_t);
template void launch_arm<double>(double *, size_t, hipStream_t);
template void launch_arm<float>(float *, size_t, hipStream_t);
} // namespace eg
