// Host-side launcher declarations of the HIP kernels (gfx950).
#pragma once
#include "device.hpp"

namespace eg {
// Device error flags (OR-ed by kernels, read by the host after a sync).
enum : int {
  kErrInadmissible = 1,  // nonpositive / nonfinite node or face trace state
  kErrFaceState = 2,     // Freeze: nonpositive old face state
  kErrRoe = 4,           // Freeze: invalid Roe sound speed
  kErrDegenerate = 8,    // Freeze: a + b <= 0
  kErrSingular = 16,     // zero pivot in a dense block inverse
  kErrNonfinite = 32,
  kErrSpin = 64,         // a dataflow dependency wait exceeded its spin bound
};

int cell_blocks(const Geo &g);
int freeze_blocks(const Geo &g);
int recovered_blocks(const Geo &g);
void launch_freeze(const Geo &g, const double *u, double safety, Frozen *fz, double *partial, int *flags,
                   hipStream_t s);
void launch_residual(const Geo &g, const Frozen *fz, const double *u, const double *du, double lambda,
                     const double *old, double *r, double *trial, double *partial, int *flags, hipStream_t s);
template <class T>
void launch_dg_apply(const Geo &g, const Frozen *fz, const double *u, const T *x, T *y, hipStream_t s);
// DG cell blocks, formed in FP64 and stored as T (double, or float for the
// single-precision factor of the mixed form).
template <class T>
void launch_dg_blocks(const Geo &g, const Frozen *fz, const double *u, T *diag, const int *dslot, T *bsr,
                      const int *slot, hipStream_t s);
void launch_hdg_build(const Geo &g, const Frozen *fz, const double *u, double *A, double *C, hipStream_t s);
void launch_hdg_solve(const Geo &g, const Frozen *fz, const double *Ainv, const double *res, double *Xy,
                      hipStream_t s);
void launch_hdg_panels(const Geo &g, const double *Xy, const double *C, double *panels, double *cy,
                       hipStream_t s);
void launch_hdg_face(const Geo &g, const Frozen *fz, const double *cy, double *rhs, double *D, hipStream_t s);
template <class T>
void launch_schur_apply(const Geo &g, const T *panels, const double *D, const T *x, T *y, hipStream_t s);
void launch_recover(const Geo &g, const double *Xy, const double *dl, double *du, hipStream_t s);
void launch_recovered_residual(const Geo &g, const Frozen *fz, const double *D, const double *eps, double *out,
                               double *partial, hipStream_t s);
void launch_hdg_diag_blocks(const Geo &g, const double *panels, const double *D, double *blocks, hipStream_t s);
// Initial projection of the exact wave into the body vector u (device).
void launch_initial(const Geo &g, const QuadRule &r, const WaveGeo &w, double *u, hipStream_t s);
// Per-cell conserved integrals (mode 0) or squared L2 errors at `time`
// (mode 1) of u, partials[c * n + e] (4 n doubles).
void launch_cell_moments(const Geo &g, const QuadRule &r, const WaveGeo &w, const double *u, double time, int mode,
                         double *partials, hipStream_t s);
// Pressure, fluxes and flux Jacobians of count states (41 doubles each).
void launch_physics_probe(const double *states, int count, double gamma, double *out, hipStream_t s);
// row_of (optional): BSR row of face f (a reordered pattern), else f.
void launch_hdg_fill(const Geo &g, const double *panels, const double *D, const int *contrib, const int *row_ptr,
                     const int *diag_ptr, double *bsr, hipStream_t s, const int *row_of = nullptr);

// Assembled and direct diagnostics (kernels_diag.hip): y = A x for a BSR
// matrix; its dense column-major N x N form; LU with partial pivoting of a
// dense matrix without row interchanges (piv[k]: pivot row of column k; used,
// mult: N ints / doubles of scratch; a zero pivot sets kErrSingular) and the
// solve x = A^{-1} b from it (work: 2N doubles, step: N ints); the block
// gather (dst(i) = src(perm(i))) or scatter (dst(perm(i)) = src(i)).
template <class T>
void launch_bsr_mult(const T *bsr, int b, int nb, const int *row_ptr, const int *col, const T *x, T *y,
                     hipStream_t s);
void launch_bsr_to_dense(const double *bsr, int b, int nb, const int *row_ptr, const int *col, double *dense,
                         hipStream_t s);
void launch_dense_lu(double *a, int N, int *used, int *piv, double *mult, int *flags, hipStream_t s);
void launch_dense_solve(const double *a, int N, const int *piv, const double *b, double *x, double *work, int *step,
                        hipStream_t s);
template <class T>
void launch_block_permute(const T *src, T *dst, const int *perm, int b, int nb, bool scatter, hipStream_t s);

// Dense / block linear algebra.
int la_init(); // nonzero if the large-LDS attribute could not be set (clears the error state)
void launch_block_inverse(double *blocks, int b, int count, const int *index, int *flags, hipStream_t s);
template <class T>
void launch_block_gemv(const T *blocks, int b, int count, const T *x, T *y, hipStream_t s, const int *slot = nullptr);
// rows: the level's rows; row_diag: their diagonal block slots; scratch:
// nrows b x b blocks, used for large blocks (big_block(b)) only.
void launch_ilu_factor_level(double *bsr, int b, const int *rows, const int *row_diag, int nrows,
                             const int *row_ptr, const int *diag_ptr, const int *col, const int *upd_ptr,
                             const int *upd, double *scratch, int *flags, hipStream_t s);
// Split large-block form of one level: na lower blocks (slots a_ik, inverse
// diagonal slots a_kk, scratch of na blocks), nt updated blocks t_slot with
// pairs [t_pair_ptr[t], t_pair_ptr[t+1]) of (L slot, U slot), then the
// inverses of the level's nrows diagonal blocks row_diag.
void launch_ilu_factor_level_split(double *bsr, int b, int na, const int *a_ik, const int *a_kk, int nt,
                                   const int *t_slot, const int *t_pair_ptr, const int *pairs, const int *row_diag,
                                   int nrows, double *scratch, int *flags, hipStream_t s);
// Blocks handled by the global-memory large-block kernels, and their limit.
bool big_block(int b);
int big_block_max();
// Per-level sweeps; threads per row 256 or 1024 (0: 256 up to 64 x 64
// blocks, else 1024).
template <class T>
void launch_ilu_lower_level(const T *bsr, int b, const int *rows, int nrows, const int *row_ptr,
                            const int *diag_ptr, const int *col, const T *rhs, T *w, int threads, hipStream_t s);
template <class T>
void launch_ilu_upper_level(const T *bsr, int b, const int *rows, int nrows, const int *row_ptr,
                            const int *diag_ptr, const int *col, T *w, int threads, bool scaled, hipStream_t s);

// Register-resident / split-dot variants (kernels_ilu.hip) used by the
// launchers above when the block size allows.
bool fast_dense_ok(int b); // LDS-staged ILU factor rows and register sweeps
bool gj_regs_ok(int b);    // register-resident Gauss-Jordan inverse
int ilu_factor_lds_init();
void launch_block_inverse_fast(double *blocks, int b, int count, const int *index, int *flags, hipStream_t s);
void launch_ilu_factor_fast(double *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                            const int *col, const int *upd_ptr, const int *upd, int *flags, hipStream_t s);
// Row-resident factor of small blocks (b <= 36): one workgroup per row of a
// level, driven by a per-row program (prog_off[r] into prog, which is padded
// by kIluProgMax entries); lds: the largest row's blocks and dependencies
// (b * b doubles each) plus ilu_resident_gj_lds(b).
constexpr int kIluProgMax = 128;
bool ilu_resident_ok(int b);
size_t ilu_resident_gj_lds(int b);
void launch_ilu_factor_resident(double *bsr, int b, const int *prog_off, const int *prog, int nrows, size_t lds,
                                int *flags, hipStream_t s);
// The same programs as one launch (programs in level order, taken from a
// ticket counter): a row reads the blocks of earlier rows from pub, which
// must hold the sweep sentinel (launch_arm) at launch, as they are published
// there, and publishes its diagonal inverse and upper blocks to pub as well
// as writing bsr. ticket: two counters, zero before the first launch and
// zero again after each. False when the kernel cannot be launched.
bool launch_ilu_factor_resident_df(double *bsr, double *pub, int b, const int *prog_off, const int *prog, int nrows,
                                   int *ticket, size_t lds, int *flags, int max_grid, hipStream_t s);
// Row-slab block products of one level for b in {64, 100, 144, 196, 256,
// 324} (the diagonal inverses follow with launch_block_inverse).
bool ilu_slab_ok(int b);
// star: every lower block of every row has exactly one update pair, the
// row's diagonal, and no row has more than ilu_star_max_lower() lower
// blocks (the two-stage slab kernel for 64 x 64 blocks).
int ilu_star_max_lower();
// In place M_ij = inv(U_ii) U_ij for the upper blocks (slot[i], diagonal
// dslot[i]) after the factor, for block sizes where ilu_scale_upper_ok.
bool ilu_scale_upper_ok(int b);
void launch_ilu_scale_upper(double *bsr, int b, const int *slot, const int *dslot, int count, hipStream_t s);

void launch_ilu_slab(double *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                     const int *col, const int *upd_ptr, const int *upd, bool star, hipStream_t s);
// The single-precision factor of the mixed form, for b in {100, 144, 196}
// (ilu_factor32_ok): the MFMA slab products of one level, the diagonal
// inverses of its rows, and the upper scaling, on FP32 blocks.
bool ilu_factor32_ok(int b);
void launch_ilu_slab32(float *bsr, int b, const int *rows, int nrows, const int *row_ptr, const int *diag_ptr,
                       const int *col, const int *upd_ptr, const int *upd, hipStream_t s);
void launch_block_inverse32(float *blocks, int b, int count, const int *index, int *flags, hipStream_t s);
void launch_ilu_scale_upper32(float *bsr, int b, const int *slot, const int *dslot, int count, hipStream_t s);
// Single-launch dataflow forms: rows taken from a ticket counter in
// dependency order (per-row completion flags or sentinel-armed values), so
// they finish whether or not the whole grid is resident; return false when
// the block size has no such kernel (the caller uses the per-level launches).
// Sweeps: src/dst/arm as documented at k_ilu_sweep_sn (lower: src = rhs,
// dst = intermediate, arm = final output; upper: src = intermediate, dst =
// final output). The intermediate must be armed before the first lower sweep.
// order: the nb rows in the sweep's dependency-level order (the lower or
// upper level lists of the pattern); ticket: two ints, zero before the first
// launch and left zero by each launch; launches sharing them must be ordered
// on one stream.
// max_grid (> 0): at most that many workgroups (more than the rows that can
// be in flight at once would only poll).
template <class T>
bool launch_ilu_sweep_df(bool lower, const T *bsr, int b, int nb, int max_deps, const int *row_ptr,
                         const int *diag_ptr, const int *col, T *src, T *dst, T *arm, int *flags, bool scaled,
                         const int *order, int *ticket, int max_grid, hipStream_t s);
template <class T> void launch_arm(T *v, size_t n, hipStream_t s);
// Row-sliced dataflow sweeps for 144 / 196 cell blocks with scaled upper
// blocks (k_ilu_sweep_slice): each block row over `slices` workgroups. Lower:
// src = rhs, dst = intermediate (armed), arm = final output; upper: src =
// intermediate (complete), dst = final output, arm = the other intermediate
// vector (narm values each). False when the kernel cannot be launched.
bool ilu_sweep_slice_ok(int b, int max_deps);
template <class T>
// given_c (upper sweep): src holds c = inv(U_ii) w_i for every row already
// (launch_block_gemv over the diagonal slots), not w.
bool launch_ilu_sweep_slice(bool lower, const T *bsr, int b, int nb, int slices, const int *row_ptr,
                            const int *diag_ptr, const int *col, const T *src, T *dst, T *arm, long narm, int *flags,
                            const int *order, int *ticket, int max_grid, hipStream_t s, bool given_c = false);

bool launch_ilu_factor_df(double *bsr, int b, int nb, const int *row_ptr, const int *diag_ptr, const int *col,
                          const int *upd_ptr, const int *upd, int *done, int epoch, const int *order, int *ticket,
                          int *flags, int max_grid, hipStream_t s);

// Allocation-free weighted norm sqrt(sum metric v^2) into the device double
// out: its partial sums go to a fixed per-device array of kReduceBlocks
// doubles (no allocation, no host synchronisation); n in 1 .. kNoallocMaxValues.
constexpr long kNoallocMaxValues = 2147483647L;
void launch_cell_norm_noalloc(const double *v, const double *metric, long n, double *out, hipStream_t s);

// Vector kernels (Krylov). nb = kReduceBlocks partials per reduction.
struct Coef {
  double c[64];
};
template <class T>
void launch_multidot(const T *V, long ldv, int nvec, const T *w, const T *metric, long n, double *partials,
                     hipStream_t s);
template <class T>
void launch_cgs_update(const T *V, long ldv, int nvec, const double *partials_in, T *w, const T *metric, long n,
                       double *proj_out, double *partials_out, int norm_mode, hipStream_t s);
template <class T>
void launch_normalize(const double *partials, T *w, T *vnext, long n, double *norm2_out, const double *proj_in,
                      int nproj, double *host_out, hipStream_t s);
template <class T>
void launch_lincomb(const T *V, long ldv, int count, const Coef &coef, T *out, long n, hipStream_t s);
template <class T>
void launch_dot_partials(const T *a, const T *b, const T *metric, long n, double *partials, hipStream_t s);
// host_out / flags / host_flags (optional): the sums also to pinned host
// memory, and flags[0] to host_flags[0], by the kernel itself.
void launch_sum_partials(const double *partials, int count, double *out, int mode, hipStream_t s, int rows = 1,
                         double *host_out = nullptr, const int *flags = nullptr, int *host_flags = nullptr);
template <class A, class B>
void launch_convert(const A *x, B *y, long n, hipStream_t s);
template <class T>
void launch_axpby(T *y, double a, const T *x, double b, long n, hipStream_t s); // y = a*x + b*y
void launch_sub(const double *a, const double *b, double *out, long n, hipStream_t s); // out = a - b
void launch_scale_copy(const double *x, double a, double *y, long n, hipStream_t s); // y = a*x
void launch_dirk_combine(double *out, const double *base, const double *const *slopes, const double *coef,
                         int count, long n, hipStream_t s);
void launch_slope(const double *u, const double *base, double inv_dt, double *slope, long n, hipStream_t s);
void launch_admissible(const Geo &g, const double *u, int *flags, hipStream_t s);

// Grouped Arnoldi steps (GroupTable in device memory; grid: the persistent
// grid, any size >= 1). The stages of one step in order: lower sweep, upper
// sweep (launch_group_sweep, face blocks b = 4 q), Schur apply
// (launch_group_schur, trace order q), CGS2 stages 0..3
// (launch_group_krylov). False when the block size or order has no kernel.
template <class T> bool launch_group_sweep(bool lower, int b, const GroupTable *tab, int grid, hipStream_t s);
template <class T> int group_sweep_capacity(int b); // workgroups of the sweeps the device holds at once
template <class T> bool launch_group_schur(int q, const GroupTable *tab, int grid, hipStream_t s);
template <class T> void launch_group_krylov(int stage, const GroupTable *tab, int grid, hipStream_t s);
int elementwise_grid(long n); // blocks of the elementwise kernels (k_normalize) over n entries
// The last stage of a grouped Arnoldi step: Givens rotation of each slot's
// new Hessenberg column, the cycle's continuation and its pinned mirror.
void launch_group_tail(const GroupTable *tab, hipStream_t s);
// Starts a device GMRES cycle: g = (beta, 0, ...), at most jmax steps, ends
// when |g_{j+1}| <= target.
void launch_cycle_init(CycleState *c, double *mirror, double beta, double target, int jmax, hipStream_t s);
} // namespace eg
