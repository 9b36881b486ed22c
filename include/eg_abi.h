/*
 * C ABI of the MI355X implicit DG/HDG Newton update (libeg_hip.so).
 *
 * The reference (YueWu2002/mfem_euler2d_cpp) has no FFI: its hot path is the
 * C++ API of include/euler.hpp driven by Run (src/stepper.cpp:28-292). Each
 * entry point below replaces one of those C++ calls; the cited file:line is
 * the reference interface it stands in for. All functions return 0 on success
 * and 1 on failure, with the reference's error text in eg_last_error() (the
 * reference throws std::runtime_error from require(), include/euler.hpp:19-23).
 *
 * Vectors named dev_* are device pointers (HIP global memory) in the
 * reference layouts: body vectors u[e*m + 4*i + c] (cell e, GL node i with x
 * fastest, component rho, rho*u, rho*v, E) and trace vectors
 * l[f*t + 4*j + c] (face f, face GL point j) (include/euler.hpp:44-45).
 * Work is enqueued on the context's HIP stream; calls that return scalars
 * synchronise that stream. One context per host thread, as the reference
 * Problem is single-threaded (include/euler.hpp:97,120,188).
 */
#ifndef EG_ABI_H
#define EG_ABI_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct eg_ctx eg_ctx;

typedef struct {
  int iterations;            /* Krylov iterations (inner FP32 ones in mixed precision) */
  int outer;                 /* FP64 refinement steps (mixed precision), else 0 */
  double relative_residual;  /* ||rhs - A x||_W / ||rhs||_W (src/linear.cpp:240) */
  double residual_norm;      /* ||rhs - A x||_W */
  double setup_s, solve_s;   /* src/linear.cpp:185-231 timers */
} eg_linear_stats;

typedef struct {
  int updates, time_steps;   /* RunResult (include/euler.hpp:250-255) */
  double newton_seconds, final_time, drift, seconds;
  double errors[4];
} eg_run_stats;

/* Problem construction from the reference command line (ParseOptions,
 * src/options.cpp:8-150, then Problem::Problem, src/problem.cpp:32-223).
 * stream: a hipStream_t to run on, or NULL for a context-owned stream. */
int eg_create(const char *args, double gamma, void *stream, eg_ctx **out);
void eg_destroy(eg_ctx *ctx);
/* Message of the last failure on ctx (or of the last failed eg_create in
 * this thread when ctx is NULL). */
const char *eg_last_error(const eg_ctx *ctx);
/* n, p, m, t, nf, nu, nl, q (include/euler.hpp:166-168). */
int eg_sizes(const eg_ctx *ctx, int *out8);
/* hx, hy, area, lx, ly, dt (include/euler.hpp:165). */
int eg_geometry(const eg_ctx *ctx, double *out6);
/* 1-D GL nodes/weights on [0,1] (q), D[a*q+i] = l_i'(x_a), l_i(0), l_i(1)
 * (src/problem.cpp:91-162). */
int eg_tables(const eg_ctx *ctx, double *x, double *w, double *d, double *e0, double *e1);
int eg_sync(eg_ctx *ctx);

/* Problem::Initial (src/problem.cpp:329-349), into host memory. */
int eg_initial(eg_ctx *ctx, double *host_u);
/* Problem::Freeze + WaveRate (src/problem.cpp:350-384): frozen HLL bounds at
 * dev_u, old = dev_u, dt = dt. host_frozen (optional, nf*q*6 doubles): per
 * face point a, b, 1/(a+b), b/(a+b), a/(a+b), ab/(a+b). */
int eg_freeze(eg_ctx *ctx, const double *dev_u, double dt, double *rate, double *host_frozen);
/* p.old = dev_old; p.dt = dt (src/stepper.cpp:109-110). */
int eg_set_stage(eg_ctx *ctx, const double *dev_old, double dt);
/* Problem::Residual + CellNorm (src/problem.cpp:445-507,385-391). */
int eg_residual(eg_ctx *ctx, const double *dev_u, double *dev_r, double *norm);
/* One line-search trial (src/stepper.cpp:205-210): trial = u + lambda du,
 * Admissible(trial), Residual(trial), CellNorm. */
int eg_trial_residual(eg_ctx *ctx, const double *dev_u, const double *dev_du, double lambda, double *dev_trial,
                      double *dev_r, double *norm, int *admissible);
/* Problem::CellNorm (src/problem.cpp:385-391). */
int eg_cell_norm(eg_ctx *ctx, const double *dev_v, double *out);
/* No counterpart in the reference: an allocation-free form of eg_cell_norm
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops). Writes sqrt(sum_i dev_metric[i] dev_v[i]^2)
 * over n values (with the CellNorm weights area * w_i of the context's
 * metric it is CellNorm, src/problem.cpp:385-391) to the device double
 * dev_out, enqueued on stream (NULL: the null stream) with no context, no
 * allocation and no host synchronisation. Its partial sums go to a fixed
 * 4 KiB device array per GPU that every call on that GPU shares: calls that
 * overlap on one GPU race. n must be in 1 .. 2147483647 (else an error,
 * with the message in eg_last_error(NULL)). */
int eg_cell_norm_noalloc(const double *dev_v, const double *dev_metric, long n, void *stream, double *dev_out);
/* Problem::Admissible (src/problem.cpp:259-275). */
int eg_admissible(eg_ctx *ctx, const double *dev_u, int *ok);

/* PartialDGOperator::Refresh (src/partial.cpp:31-59): sets the Newton state. */
int eg_dg_refresh(eg_ctx *ctx, const double *dev_u);
/* PartialDGOperator::Mult (src/partial.cpp:62-135). */
int eg_dg_apply(eg_ctx *ctx, const double *dev_x, double *dev_y);
/* The same action in FP32 (mixed-precision inner solves). */
int eg_dg_apply_f32(eg_ctx *ctx, const float *dev_x, float *dev_y);

/* Preconditioner fill + factor at the current state (src/linear.cpp:189-203
 * with FillBlocks/DiagonalBlocks and FixedBlockILU::Factor), form 0 DG,
 * form 1 Schur; the kind is the context's --precond. */
int eg_pc_setup(eg_ctx *ctx, int form);
/* LinearSolver::Impl::Precondition (src/linear.cpp:66-75). */
int eg_pc_apply(eg_ctx *ctx, int form, const double *dev_x, double *dev_y);
/* The same preconditioner applied in FP32 (the mixed-precision inner solves;
 * the context needs --precision mixed and a set-up preconditioner). */
int eg_pc_apply_f32(eg_ctx *ctx, int form, const float *dev_x, float *dev_y);

/* Problem::StreamCondense (src/condensation.cpp:131-148). */
int eg_hdg_condense(eg_ctx *ctx, const double *dev_u, const double *dev_r);
/* Condensed right-hand side and trace diagonal D (Condensed::rhs, Hybrid::D). */
int eg_hdg_rhs(eg_ctx *ctx, double *dev_rhs, double *dev_D);
/* [X_e | y_e] per cell, column-major m x (4t+1) (Condensed::X, ::y). */
int eg_hdg_xy(eg_ctx *ctx, double *dev_xy);
/* ElementOperator::Mult with borrowed factors (src/operators.cpp:147-179). */
int eg_schur_apply(eg_ctx *ctx, const double *dev_x, double *dev_y);
/* Problem::Recover (src/condensation.cpp:149-163). */
int eg_recover(eg_ctx *ctx, const double *dev_dl, double *dev_du);
/* RecoveredResidual + CellNorm (src/stepper.cpp:10-26). */
int eg_recovered_residual(eg_ctx *ctx, const double *dev_eps, double *dev_out, double *norm);

/* LinearSolver::Solve (src/linear.cpp:167-249) on form 0 (DG at the
 * refreshed state) or 1 (condensed Schur system): refresh != 0 rebuilds the
 * preconditioner and starts from zero, else warm-starts from the last x.
 * dev_x receives the solution, dev_res (optional) rhs - A x. */
int eg_linear_solve(eg_ctx *ctx, int form, const double *dev_rhs, double *dev_x, double *dev_res, double tol,
                    int refresh, eg_linear_stats *stats);

/* The block row order of the ILU(0) factor of form 0 (cell graph) or 1 (face
 * graph) that --ilu-ordering selects (FixedBlockILU's permutation,
 * src/block_ilu.cpp:14-29): order[i] = the graph row factored i-th (n or nf
 * ints); the natural order for none, the minimum discarded fill order for
 * mdf. */
int eg_ilu_order(eg_ctx *ctx, int form, int *order);

/* Run (src/stepper.cpp:28-292) for form 0 (dg) or 1 (hdg). prefix: CSV/solution
 * output prefix or NULL. host_u_final: nu doubles. host_increments (optional):
 * up to max_updates accepted Newton increments (nu doubles each); krylov
 * (optional): Krylov iterations per update. */
int eg_run(eg_ctx *ctx, int form, const char *prefix, double *host_u_final, eg_run_stats *stats,
           double *host_increments, int max_updates, int *krylov);

/* Measurement hook (no counterpart in the reference, which only times whole
 * updates with steady_clock, include/euler.hpp:24-28): sets up the context's
 * problem at its initial state and times `reps` back-to-back calls of one
 * operation with HIP events on the context's stream. kernel: "residual",
 * "dg_apply", "dg_apply_f32", "schur_apply", "pc_dg", "pc_hdg",
 * "hdg_condense", "pc_setup_dg", "pc_setup_hdg" (the setups are the block
 * fill plus the factor's launches). avg_ms: mean duration per call; bytes:
 * algorithmic HBM bytes of one call (each array counted once); flops (may be
 * NULL): algorithmic FP64 flops of one call for the dense block work
 * (condensation and preconditioner setup, LU-based counts), else 0. */
int eg_kernel_time(eg_ctx *ctx, const char *kernel, int reps, double *avg_ms, double *bytes, double *flops);

/* ---------------------------------------------------------------- many solves
 * No counterpart in the reference, which runs one case per process
 * (src/main.cpp:17-29); these batch and serve independent whole solves
 * (Run, src/stepper.cpp:28-292), each on its own stream and context. */
typedef struct {
  double arrival, start, finish; /* seconds from the start of eg_serve */
  double newton_seconds;         /* timed Newton updates (src/stepper.cpp:117,226) */
  double footprint;              /* device bytes reserved for admission (eg_solve_footprint) */
  double measured_peak;          /* device bytes the solve actually held at its peak */
  double errors[4];              /* L2 errors against the exact wave (Problem::Errors) */
  int status;                    /* 0 done; 1 failed or rejected (message in errors buffer) */
  int updates, time_steps, worker;
} eg_solve_outcome;
typedef struct {
  double wall;         /* seconds from the first arrival window to the last result */
  double max_reserved; /* largest sum of admitted footprints, bytes */
  int max_concurrent;  /* most solves on the device at once */
} eg_serve_report;
/* Upper bound of the device bytes one solve (args, form 0 dg / 1 hdg) holds. */
int eg_solve_footprint(const char *args, int form, double *bytes);
/* Runs n independent solves: args[i] a reference command line, forms[i]
 * 0 (dg) or 1 (hdg), arrivals[i] seconds after the call (NULL: all at once).
 * Requests are admitted in arrival order when one of `workers` slots is free
 * and their footprint fits in hbm_budget bytes (0: 80% of the free device
 * memory) next to the solves already running; results overlap on the device.
 * out[n] receives the outcomes; states (optional) n host pointers that
 * receive each final state (nu doubles of that case, or NULL to skip);
 * errors (optional) n * 256 chars for failure messages. Returns 0 when the
 * loop ran; per-request failures are in out[i].status. */
int eg_serve(const char *const *args, const int *forms, const double *arrivals, int n, int workers,
             double hbm_budget, eg_solve_outcome *out, double *const *states, char *errors,
             eg_serve_report *report);

/* The batched form of eg_serve (the same requests, admission and outcomes;
 * up to 256 workers): the Arnoldi steps of the condensed (form 1) solves
 * with block ILU(0) that run at the same time are launched together, those
 * of one trace order k + 1 and arithmetic as one replay of a recorded graph
 * of persistent grouped kernels (ILU(0) sweeps, Schur apply, CGS2) over all
 * of them, instead of each solve's own launches; every result is the same,
 * bit for bit, as that solve's own launches give. window_us (<= 0: 150): how
 * long the first waiting step waits for the steps of the other solves inside
 * a Krylov iteration. batch (optional) receives the grouping counts. */
typedef struct {
  long long flushes;        /* grouped launches (one or more graph replays each) */
  long long chunks;         /* chunks of Arnoldi steps the solves handed over */
  long long graph_launches; /* graph replays, one step of every solve in the flush each */
  int max_group;            /* most solves in one flush */
  int tile[10];             /* per trace order q = k + 1: virtual sweep workgroups per row of a solve's widest
                               dependency level, chosen online (0: none) */
  double launch_s;          /* host seconds the batcher spent launching */
  double window_s;          /* host seconds it spent gathering steps */
} eg_batch_report;
int eg_serve_batched(const char *const *args, const int *forms, const double *arrivals, int n, int workers,
                     double hbm_budget, double window_us, eg_solve_outcome *out, double *const *states,
                     char *errors, eg_serve_report *report, eg_batch_report *batch);
/* ILU(0) launch-shape autotuner: for each k in ks, times the preconditioner
 * setup plus 16 applications (form 0 dg / 1 hdg, precision "fp64" or
 * "mixed") on a 16x16 case of that k under three shapes: "levels-256" and
 * "levels-1024" (one launch per dependency level, 256 or 1024 threads per
 * block row of the sweeps) and "dataflow" (one launch per sweep, rows taken
 * in dependency order from a ticket counter). out + 16*i receives the
 * fastest name (NUL-terminated); times (optional, 3*nk) the three times in
 * ms, in that order. */
int eg_autotune_ilu(const int *ks, int nk, int form, const char *precision, char *out, double *times);

#ifdef __cplusplus
}
#endif
#endif
