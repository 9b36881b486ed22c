// Host API of the MI355X implicit DG/HDG Newton update. It keeps the names,
// argument meaning and error behaviour of the reference C++ API
// (include/euler.hpp:29-258 of YueWu2002/mfem_euler2d_cpp): EulerGas,
// Options/ParseOptions, Problem, LinearSolver, Run. Vectors named dev_* are
// device pointers; all numerical work of the Newton update runs in the HIP
// kernels of kernels_cell.hip / kernels_la.hip on the problem's stream.
#pragma once
#include <algorithm>
#include <array>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "kernels.hpp"

namespace euler {
inline constexpr double pi = 3.141592653589793238462643383279502884;
using State = std::array<double, 4>;

// include/euler.hpp:19-23
inline void require(bool ok, const char *msg) {
  if (!ok) { throw std::runtime_error(msg); }
}
inline void require(bool ok, const std::string &msg) {
  if (!ok) { throw std::runtime_error(msg); }
}
void hip_check(hipError_t e, const char *what);
double clock_s();

// include/euler.hpp:29-43 (host physics: initial data, errors, known-answer tests)
class EulerGas {
  double gamma, gamma_minus_one;
public:
  explicit EulerGas(double value);
  double Gamma() const { return gamma; }
  double pressure(const State &u) const;
  bool positive(const State &u) const;
  State flux(const State &u, int axis) const;
  void jacobian(const State &u, int axis, double j[16]) const;
  std::array<double, 2> speeds(const State &l, const State &r, int axis, double safety) const;
  State exact(double x, double y, double time, double alpha) const;
};

// include/euler.hpp:46-57, plus --precision fp64|mixed and --inner-tol for
// the mixed-precision form.
struct Options {
  int nx = 100, ny = 100, k = 1, volume_order = -1, max_newton = 30;
  int restart = 20, max_krylov = 1500, newton_steps = -1, steps = -1;
  double alpha = 45, dt = .001, end = .2, cfl = -1, safety = 1.1, tol = 1e-12, linear_tol = 1e-12;
  std::string assembly = "partial";
  std::string form = "both", solver = "gmres", precond = "block-jacobi", integrator = "ssp-dirk3",
              output = "code_test/results/run";
  bool matrix_free = true, verify = false, save_solution = true, hdg_first = false;
  bool check_update = false, timing_only = false;
  std::string ilu_ordering = "none";
  std::string precision = "fp64";
  double inner_tol = 1e-4;
  // ILU(0) launch schedule: "levels", "dataflow" or "dataflow-all"; empty
  // takes EG_ILU_SCHEDULE from the environment, else "dataflow".
  std::string ilu_schedule;
  // Threads per row of the per-level ILU sweeps: 256 or 1024 (0: by block size).
  int sweep_threads = 0;
};
Options ParseOptions(int argc, char **argv);
Options ParseOptions(const std::string &args);

// Device memory made by this thread's DeviceArrays: live and peak bytes,
// and, inside an AllocScope, a stream for stream-ordered allocation, so that
// a solve that ends frees its memory without a device-wide synchronisation
// (concurrent solves on other streams keep running).
struct AllocStats {
  size_t live = 0, peak = 0;
  hipStream_t stream = nullptr;
};
AllocStats &thread_allocs();
class AllocScope {
public:
  explicit AllocScope(hipStream_t s) : prev_(thread_allocs().stream) { thread_allocs().stream = s; }
  ~AllocScope() { thread_allocs().stream = prev_; }
  AllocScope(const AllocScope &) = delete;
  AllocScope &operator=(const AllocScope &) = delete;
private:
  hipStream_t prev_;
};

template <class T> class DeviceArray {
public:
  T *p = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  explicit DeviceArray(size_t count) { resize(count); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() { release(); }
  void resize(size_t count) {
    if (count == n) { return; }
    release();
    if (count) {
      auto &a = thread_allocs();
      stream_ = a.stream;
      if (stream_) { hip_check(hipMallocAsync(reinterpret_cast<void **>(&p), count * sizeof(T), stream_), "hipMallocAsync"); }
      else { hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc"); }
      a.live += count * sizeof(T);
      a.peak = std::max(a.peak, a.live);
    }
    n = count;
  }
  void release() {
    if (p) {
      if (stream_) { (void)hipFreeAsync(p, stream_); }
      else { (void)hipFree(p); }
      auto &a = thread_allocs();
      a.live -= std::min(a.live, n * sizeof(T));
    }
    p = nullptr;
    n = 0;
    stream_ = nullptr;
  }
  operator T *() const { return p; }
private:
  hipStream_t stream_ = nullptr;
};

// Fixed BSR topology of a block ILU(0): pattern, update schedule and the
// dependency levels of the forward and backward sweeps (src/block_ilu.cpp:10-59).
struct BlockPattern {
  int nb = 0, block = 0, nnzb = 0;
  int max_deps = 0; // most off-diagonal blocks in one triangle of a row
  int max_width = 0; // most rows in one dependency level (lower or upper)
  std::vector<int> row, col, diag, upd_ptr, upd;
  std::vector<int> lower_level_ptr, lower_rows, upper_level_ptr, upper_rows;
  DeviceArray<int> d_row, d_col, d_diag, d_upd_ptr, d_upd, d_lower_rows, d_upper_rows;
  DeviceArray<int> d_lower_diag; // diagonal block slot of each row of lower_rows
  // Reordered pattern (src/block_ilu.cpp:10-29): BSR row i is graph row
  // perm[i], and inv[perm[i]] = i; both empty for the natural order.
  std::vector<int> perm, inv;
  DeviceArray<int> d_perm, d_inv;
  BlockPattern(const std::vector<std::vector<int>> &graph, int size, hipStream_t s,
               const std::vector<int> &order = {});
  // BSR slot of block (r, c) of the graph (rows and columns in the graph's
  // numbering, mapped through the order).
  int Find(int r, int c) const;
};
// Minimum discarded fill ordering of an unweighted block graph (the
// reference's --ilu-ordering mdf, src/block_ilu.cpp:16-26): order[i] is the
// graph row eliminated i-th.
std::vector<int> MdfOrdering(const std::vector<std::vector<int>> &graph);

class KrylovBatcher;

// include/euler.hpp:161-230: device-resident discretisation of one case.
class Problem {
public:
  Options opt;
  const EulerGas gas;
  double alpha, lx, ly, hx, hy, area, dt = 0;
  int n, p, m, t, nf, nu, nl, q;
  std::vector<double> x1, w1, D1, end0, end1, origx, origy;
  std::vector<double> cell_metric, trace_metric;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  eg::Geo geo{};
  eg::QuadRule rule{};  // the projection / error rule (src/problem.cpp:329-349,704-726)
  eg::WaveGeo wave{};   // cell origins (device) and the wave angle
  DeviceArray<double> d_origx, d_origy, moments;
  DeviceArray<eg::Frozen> frozen;
  DeviceArray<double> old, d_cell_metric, d_trace_metric, partial, scalars;
  DeviceArray<float> d_cell_metric32, d_trace_metric32;
  DeviceArray<int> flags;
  double *pinned = nullptr; // 512 doubles (two GMRES Hessenberg slots), then the flags
  int *pinned_flags = nullptr;
  int freeze_generation = 0;
  double rate = 0;
  double last_state_norm = 0; // CellNorm of the state of the last (Trial)Residual
  // Grouped Arnoldi steps with other solves (set by the batched form before
  // the solvers of this problem are made; null: every launch is its own).
  KrylovBatcher *batcher = nullptr;

  Problem(const Options &o, const EulerGas &gas, hipStream_t s = nullptr);
  ~Problem();
  Problem(const Problem &) = delete;
  std::vector<double> Initial();
  void Freeze(const double *dev_u, double step, bool check_admissible = false);
  double WaveRate() const { return rate; }
  void SetStage(const double *dev_old, double step);
  double CellNorm(const double *dev_r);
  bool Admissible(const double *dev_u);
  double Residual(const double *dev_u, double *dev_r);
  double TrialResidual(const double *dev_u, const double *dev_du, double lambda, double *dev_trial,
                       double *dev_r, bool &admissible);
  std::vector<std::vector<int>> SolverGraph(bool hdg) const;
  // Initial, Errors and Integral run on the device: the *Device forms on a
  // device body vector, the others on a host vector (copied through).
  void InitialDevice(double *dev_u);
  std::array<double, 4> ErrorsDevice(const double *dev_u, double time);
  std::array<double, 4> IntegralDevice(const double *dev_u);
  std::array<double, 4> Errors(const std::vector<double> &u, double time);
  std::array<double, 4> Integral(const std::vector<double> &u);
  std::array<double, 4> Differences(const std::vector<double> &u, const std::vector<double> &v) const;
  void Save(const std::vector<double> &u, const std::string &prefix) const;
  // scalar plumbing
  void Sync();
  int TakeFlags();
  void CheckFlags(const char *context);
  // The same checks on the flags last copied to pinned memory (by
  // ReduceScalar), without a readback of their own.
  void CheckPinnedFlags(const char *context);
  double ReduceScalar(const double *dev_partials, int count, int mode);
};

// Condensed HDG state of one Newton update (include/euler.hpp:138-160):
// A_e^{-1} workspace, restriction blocks, [X | y], face panels of S_e, the
// condensed right-hand side and the trace diagonal.
class Hybrid {
public:
  Problem &p;
  DeviceArray<double> A, C, Xy, panels, cy, D, rhs;
  DeviceArray<float> panels32;
  explicit Hybrid(Problem &problem, bool mixed);
  void StreamCondense(const double *dev_u, const double *dev_r);
  void SchurMult(const double *x, double *y);
  void SchurMult(const float *x, float *y);
  void Recover(const double *dev_dl, double *dev_du);
  double RecoveredResidual(const double *dev_eps, double *dev_out);
};

// Result of one linear solve (include/euler.hpp:231-235).
struct LinearResult {
  int count = 0, outer = 0;
  double residual = 0, assembly = 0, setup = 0, solve = 0, residual_norm = 0;
};

// Persistent device GMRES with block-Jacobi or block ILU(0) preconditioning
// (include/euler.hpp:240-249, src/linear.cpp). form 0: DG body system (the
// matrix-free Jacobian at the refreshed state); form 1: condensed Schur system.
class LinearSolver {
public:
  Problem &p;
  int form, N, block, restart;
  bool mixed, ilu;
  const double *state = nullptr; // DG Newton state (device)
  Hybrid *hybrid = nullptr;
  DeviceArray<double> V, w, z, comb, corr, res, x, act, partA, partB, blocks;
  DeviceArray<float> V32, w32, z32, comb32, corr32, res32, x32, b32, blocks32;
  std::unique_ptr<BlockPattern> pattern;
  DeviceArray<int> dslot, slot, contrib;
  // The reference's diagnostics (src/linear.cpp:46-56,184-213,224-240):
  // assembled: the operator is the assembled block matrix (--assembly full,
  // or element for DG; the HDG element form is the Schur panels' own
  // action), applied by launch_bsr_mult; direct: --solver dense|direct, one
  // dense LU with partial pivoting of the assembled matrix a set-up and a
  // solve with it (no preconditioner, no Krylov iterations).
  bool assembled = false, direct = false;
  std::unique_ptr<BlockPattern> op_pattern;
  DeviceArray<double> op_blocks;
  DeviceArray<float> op_blocks32;
  DeviceArray<int> op_dslot, op_slot, op_contrib;
  DeviceArray<double> dense, dense_mult, dense_work;
  DeviceArray<int> dense_used, dense_piv, dense_step;
  // Reordered ILU(0) (--ilu-ordering mdf): the preconditioner's input and
  // output in the factor's row order (src/block_ilu.cpp:87-109).
  DeviceArray<double> perm_in, perm_out;
  DeviceArray<float> perm_in32, perm_out32;
  // Dataflow ILU schedules (EG_ILU_SCHEDULE): the factor's per-row
  // completion flags and epoch, and the sweeps' intermediate vector, kept
  // armed with the sentinel between applications.
  DeviceArray<int> done_factor;
  int epoch_factor = 0;
  DeviceArray<double> sweep;
  DeviceArray<float> sweep32;
  // Row-sliced sweeps (144 / 196 cell blocks with scaled upper blocks):
  // workgroups a block row, the second intermediate vectors (the two
  // alternate between applications) and which one is next (FP64, FP32).
  int slices = 0;
  DeviceArray<double> sweep_alt;
  DeviceArray<float> sweep32_alt;
  int parity[2] = {0, 0};
  // c = inv(U_ii) w_i of every row, formed between the two sliced sweeps by
  // a block-diagonal product (given_c; EG_SWEEP_GIVEN_C=0: in the upper
  // sweep, before each slice's wait).
  bool given_c = false;
  DeviceArray<double> sweep_c;
  DeviceArray<float> sweep32_c;
  DeviceArray<double> factor_scratch; // large blocks: one block per row (or lower block) of the widest level
  // Large-block split factor plan, per dependency level l: lower blocks
  // [a_ptr[l], a_ptr[l+1]) and updated blocks [t_ptr[l], t_ptr[l+1]).
  bool split_factor = false;
  std::vector<int> a_ptr, t_ptr;
  DeviceArray<int> d_a_ik, d_a_kk, d_t_slot, d_t_pair_ptr, d_pairs;
  bool dataflow = true, dataflow_factor = false;
  // Row-resident factor as one dataflow launch (EG_RESIDENT_DF=0: per level);
  // factor_pub: the published diagonal inverses and upper blocks it reads.
  bool resident_df = true;
  DeviceArray<double> factor_pub;
  // Ticket counters of the dataflow sweeps (two ints, zero between launches).
  DeviceArray<int> sweep_ticket;
  // Most workgroups of a dataflow launch: EG_DF_WIDTH (default below) times
  // the widest dependency level; 0: as many as the device holds.
  int df_grid = 0;
  // Per-level factor launches recorded once into a graph (EG_FACTOR_GRAPH=0:
  // launched directly).
  bool use_graph = true;
  hipGraphExec_t factor_exec = nullptr;
  // Mixed precision with 100, 144 or 196 cell blocks: the blocks are filled
  // in FP32 into blocks32 and factored there (factor32, per-level launches
  // in their own graph); the FP64 blocks are filled and factored only when
  // an FP64 preconditioner apply asks for them (factor64_pending).
  bool factor32 = false;
  bool factor64_pending = false;
  hipGraphExec_t factor_exec32 = nullptr;
  // Row-resident factor of small blocks: per-row programs in level order.
  bool resident = false;
  size_t resident_lds = 0;
  DeviceArray<int> d_prog, d_prog_off;
  // Blocks up to 64: the factor's upper blocks are scaled in place to
  // inv(U_ii) U_ij after each factorisation (one product a row in the
  // backward sweep); the upper block slots and their rows' diagonal slots.
  bool scaled_upper = false;
  bool star_rows = false; // every lower block updates only its row's diagonal
  DeviceArray<int> d_up_slot, d_up_diag;
  std::vector<double> h, cosine, sine, rotated, coefficients;
  // Completion of the Arnoldi step whose Hessenberg column sits in pinned
  // slot j % 2 (pinned + 256 (j % 2)).
  hipEvent_t step_done[2] = {nullptr, nullptr};
  hipEvent_t setup_mark[2] = {nullptr, nullptr}; // around the set-up on the stream (its device time)
  // The batched form: the grouped Arnoldi steps' batcher (from the problem)
  // and the event marking this solve's work before a step it hands over.
  KrylovBatcher *batcher = nullptr;
  hipEvent_t batch_ready = nullptr;
  DeviceArray<eg::CycleState> cycle; // the device GMRES cycle of the grouped steps
  double *mirror = nullptr;          // its pinned mirror (eg::kMirrorSize doubles, from the batcher)
  bool joined = false;               // a member of the batcher's rendezvous
  // Whether the Arnoldi steps in T can be handed to the batcher (condensed
  // system, single-launch ILU(0) sweeps of face blocks).
  bool Batchable() const;

  LinearSolver(Problem &p, int form, Hybrid *hybrid = nullptr);
  ~LinearSolver();
  LinearSolver(const LinearSolver &) = delete;
  LinearSolver &operator=(const LinearSolver &) = delete;
  void PlanBigFactor();
  void PlanResident();
  // BSR slots the DG / HDG block fills write for pattern pt (DG: diagonal
  // and the four face-neighbour slots of each cell; HDG: the eight
  // contributions of each face row).
  void FillMaps(const BlockPattern &pt, DeviceArray<int> &ds, DeviceArray<int> &sl, DeviceArray<int> &ct);
  void FactorLevels();
  void FactorLevels32();
  // The per-level launches of levels, recorded once into exec and replayed.
  void ReplayLevels(hipGraphExec_t &exec, void (LinearSolver::*levels)());
  // The FP64 factor (or block-Jacobi inverses) of the filled blocks.
  void Factor64();
  void Refresh(const double *dev_u) { state = dev_u; }
  // Fill + factor the preconditioner at the current state. check_flags
  // false leaves the singular-factor check to the next readback of Solve
  // (pending_check).
  void Setup(bool check_flags = true);
  const char *pending_check = nullptr;
  void Apply(const double *x, double *y);
  void Apply(const float *x, float *y);
  void Precondition(const double *b, double *y);
  void Precondition(const float *b, float *y);
  // x (member) is the solution; res holds rhs - A x on return.
  void Solve(const double *dev_rhs, LinearResult &out, double tol, bool refresh);
  double Norm(const double *v);
  [[noreturn]] void Fail(const char *what);
  template <class T> const T *Factors() const;
  template <class T> T *SweepBuffer();
private:
  friend class KrylovBatcher;
  template <class T> void EnqueueArnoldi(T *Vb, T *zb, T *wb, const T *metric, int j);
  template <class T> int DeviceCycle(T *Vb, T *zb, T *wb, const T *metric, double beta, double target, int jmax);
  template <class T> double GMRES(const T *rhs, T *xv, double tol, int limit, int &iterations, bool zero_guess,
                                  double *true_norm,
                                    double known_rhs_norm = -1.);
  double MixedSolve(const double *rhs, double tol, int limit, int &iterations, int &outer, bool zero_guess,
                    double rhs_norm);
};

// include/euler.hpp:250-256
struct UpdateRecord {
  int step, stage, newton, krylov;
  double rms, damping, linear_residual, dg_relative, seconds;
};
struct RunResult {
  std::vector<double> u;
  std::array<double, 4> errors{};
  double seconds = 0, drift = 0, newton_seconds = 0, final_time = 0;
  int updates = 0, time_steps = 0;
  std::vector<UpdateRecord> records;
  std::vector<std::vector<double>> increments;
};
// keep_state false: the final state stays on the device (RunResult::u is
// left empty unless the solution is saved).
RunResult Run(Problem &p, const std::string &form, const std::string &prefix, bool keep_increments = false,
              bool keep_state = true);
// include/euler.hpp:257: the --verify self-test (src/verification.cpp:7-169),
// run on the device kernels; throws on a failed check.
void Verify(const Options &base, const EulerGas &gas);

// ------------------------------------------------------------ many solves
// Batched and served execution of independent solves (no counterpart in the
// reference, which runs one case per process): each solve is a whole Run on
// its own stream and context, device-resident from set-up to result.
size_t SolveFootprint(const Options &o, int form); // device bytes one solve may hold at its peak
struct SolveRequest {
  std::string args; // reference command line of the case
  int form = 1;     // 0 dg, 1 hdg
  double arrival = 0.; // seconds after the start of Serve
  double gamma = 3.;
};
struct SolveOutcome {
  int status = 0; // 0 done, 1 failed or rejected (error says why)
  std::string error;
  double arrival = 0., start = 0., finish = 0., newton_seconds = 0.;
  int updates = 0, time_steps = 0, worker = -1;
  size_t footprint = 0, measured_peak = 0;
  std::array<double, 4> errors{};
  std::vector<double> u; // final state when kept
};
struct ServeConfig {
  int workers = 8;         // concurrent solves (one host thread and stream each)
  double hbm_budget = 0.;  // bytes of device memory the solves may hold; 0: 80% of the free memory
  bool keep_states = false;
  // Batched form: the Arnoldi steps of the condensed solves that run at the
  // same time are grouped (KrylovBatcher); window_us: how long the first
  // step of a group waits for the other solves' steps.
  bool batched = false;
  double window_us = 150.;
};
// Grouped Arnoldi steps of concurrent solves: counts, and per trace order
// q (index q) the chosen virtual sweep workgroups per row of a solve's
// widest dependency level (0: none seen).
struct BatchReport {
  long flushes = 0, chunks = 0, graph_launches = 0; // chunks of steps handed over; graph replays
  int max_group = 0;
  int tile[eg::kMaxQ + 1] = {};
  double launch_s = 0., window_s = 0.; // the batcher thread's time launching, and gathering steps
};
struct ServeReport {
  double wall = 0.;
  int max_concurrent = 0;
  size_t max_reserved = 0; // largest sum of admitted footprints
  BatchReport batch;
};

// Collects the Arnoldi steps that concurrent condensed solves hand over
// (Submit, from each solve's own thread) and launches those of one trace
// order and arithmetic together: one replay of a recorded graph of the
// grouped kernels (two ILU(0) sweeps, Schur apply, four CGS2 stages) over a
// table of all of them, on the group's stream, ordered after each solve's
// earlier work and before its later work by events. Each solve's results
// are those of its own launches, bit for bit. The virtual sweep workgroups
// per solve (the per-order tile) are chosen online: the first flushes of
// each group time the candidates (any choice gives the same results).
class KrylovBatcher {
public:
  explicit KrylovBatcher(double window_us = 150.);
  ~KrylovBatcher();
  KrylovBatcher(const KrylovBatcher &) = delete;
  KrylovBatcher &operator=(const KrylovBatcher &) = delete;
  // A solve enters / leaves its Krylov iteration (a flush waits for the
  // steps of the solves inside one, up to the window).
  void Enter();
  void Leave();
  // Condensed solves in the batch (Join when made, Depart when done), and
  // their rendezvous before each linear solve (up to EG_BATCH_ALIGN_US, 2000
  // by default; 0: none), so that their Krylov iterations overlap and their
  // steps share launches.
  void Join();
  void Depart();
  void Align();
  // Hands over at least `chunk` Arnoldi steps of s's device cycle (s.cycle;
  // a launch may take more, up to the cycle's end); returns once they are
  // launched, their completion then recorded in `done` on s's stream.
  template <class T>
  void Submit(LinearSolver &s, T *V, T *z, T *w, const T *metric, int chunk, hipEvent_t done);
  // Pinned cycle mirrors (eg::kMirrorSize doubles), kept until the batcher ends.
  double *AcquireMirror();
  void ReleaseMirror(double *m);
  BatchReport Report();
private:
  struct Entry;
  struct Group;
  void Loop();
  void Flush(std::vector<Entry *> &batch);
  void Launch(Group &g, Entry *const *entries, int count);
  Group &GroupFor(int q, bool f32);
  std::mutex mu_;
  std::condition_variable wake_, launched_;
  std::vector<Entry *> queue_;
  std::vector<std::unique_ptr<Group>> groups_;
  std::vector<double *> mirrors_, free_mirrors_;
  int in_krylov_ = 0, members_ = 0, arrived_ = 0;
  long align_gen_ = 0;
  double align_us_ = 2000.;
  std::condition_variable align_;
  bool stop_ = false;
  double window_us_;
  bool use_graph_ = true;
  BatchReport report_;
  std::thread thread_;
};
// Requests are admitted in arrival order when a worker is free and their
// footprint fits in the budget next to the solves already running (a request
// larger than the whole budget is rejected); results in request order.
std::vector<SolveOutcome> Serve(const std::vector<SolveRequest> &requests, const ServeConfig &cfg,
                                ServeReport *report = nullptr);
} // namespace euler
