// extern "C" boundary (include/eg_abi.h) over the euler:: host API.
#include <cstring>
#include <functional>
#include <memory>

#include "../../include/eg_abi.h"
#include "euler.hpp"

using namespace euler;

struct eg_ctx {
  Options opt;
  EulerGas gas;
  std::unique_ptr<Problem> problem;
  std::unique_ptr<Hybrid> hybrid;
  std::unique_ptr<LinearSolver> solver[2];
  const double *state = nullptr;
  std::string error;
  eg_ctx(const Options &o, double gamma) : opt(o), gas(gamma) {}
  Hybrid &hyb() {
    if (!hybrid) { hybrid = std::make_unique<Hybrid>(*problem, opt.precision == "mixed"); }
    return *hybrid;
  }
  LinearSolver &lin(int form) {
    require(form == 0 || form == 1, "form must be 0 (dg) or 1 (hdg)");
    if (!solver[form]) {
      solver[form] = std::make_unique<LinearSolver>(*problem, form, form ? &hyb() : nullptr);
    }
    if (form == 0) { solver[0]->Refresh(state); }
    return *solver[form];
  }
};

namespace {
thread_local std::string create_error;
template <class F> int guard(eg_ctx *ctx, F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    if (ctx) { ctx->error = e.what(); }
    else { create_error = e.what(); }
    return 1;
  }
}
} // namespace

extern "C" {
int eg_create(const char *args, double gamma, void *stream, eg_ctx **out) {
  *out = nullptr;
  return guard(nullptr, [&] {
    auto ctx = std::make_unique<eg_ctx>(ParseOptions(std::string(args ? args : "")), gamma);
    ctx->problem = std::make_unique<Problem>(ctx->opt, ctx->gas, static_cast<hipStream_t>(stream));
    *out = ctx.release();
  });
}
void eg_destroy(eg_ctx *ctx) { delete ctx; }
const char *eg_last_error(const eg_ctx *ctx) { return ctx ? ctx->error.c_str() : create_error.c_str(); }
int eg_sizes(const eg_ctx *ctx, int *s) {
  const auto &p = *ctx->problem;
  const int v[8] = {p.n, p.p, p.m, p.t, p.nf, p.nu, p.nl, p.q};
  std::memcpy(s, v, sizeof v);
  return 0;
}
int eg_geometry(const eg_ctx *ctx, double *g) {
  const auto &p = *ctx->problem;
  const double v[6] = {p.hx, p.hy, p.area, p.lx, p.ly, p.dt};
  std::memcpy(g, v, sizeof v);
  return 0;
}
int eg_tables(const eg_ctx *ctx, double *x, double *w, double *d, double *e0, double *e1) {
  const auto &p = *ctx->problem;
  for (int i = 0; i < p.q; ++i) { x[i] = p.x1[i]; w[i] = p.w1[i]; e0[i] = p.end0[i]; e1[i] = p.end1[i]; }
  std::memcpy(d, p.D1.data(), sizeof(double) * p.q * p.q);
  return 0;
}
int eg_sync(eg_ctx *ctx) { return guard(ctx, [&] { ctx->problem->Sync(); }); }
int eg_initial(eg_ctx *ctx, double *host_u) {
  return guard(ctx, [&] {
    const auto u = ctx->problem->Initial();
    std::memcpy(host_u, u.data(), sizeof(double) * u.size());
  });
}
int eg_freeze(eg_ctx *ctx, const double *dev_u, double dt, double *rate, double *host_frozen) {
  return guard(ctx, [&] {
    auto &p = *ctx->problem;
    p.Freeze(dev_u, dt, false);
    if (rate) { *rate = p.WaveRate(); }
    if (host_frozen) {
      std::vector<eg::Frozen> fz(size_t(p.nf) * p.q);
      hip_check(hipMemcpy(fz.data(), p.frozen, sizeof(eg::Frozen) * fz.size(), hipMemcpyDeviceToHost), "download");
      for (size_t i = 0; i < fz.size(); ++i) {
        const double v[6] = {fz[i].a, fz[i].b, fz[i].inv_sum, fz[i].wl, fz[i].wr, fz[i].pen};
        std::memcpy(host_frozen + 6 * i, v, sizeof v);
      }
    }
  });
}
int eg_set_stage(eg_ctx *ctx, const double *dev_old, double dt) {
  return guard(ctx, [&] { ctx->problem->SetStage(dev_old, dt); });
}
int eg_residual(eg_ctx *ctx, const double *dev_u, double *dev_r, double *norm) {
  return guard(ctx, [&] {
    const double v = ctx->problem->Residual(dev_u, dev_r);
    if (norm) { *norm = v; }
  });
}
int eg_trial_residual(eg_ctx *ctx, const double *dev_u, const double *dev_du, double lambda, double *dev_trial,
                      double *dev_r, double *norm, int *admissible) {
  return guard(ctx, [&] {
    bool adm = false;
    const double v = ctx->problem->TrialResidual(dev_u, dev_du, lambda, dev_trial, dev_r, adm);
    if (norm) { *norm = v; }
    if (admissible) { *admissible = adm ? 1 : 0; }
  });
}
int eg_cell_norm(eg_ctx *ctx, const double *dev_v, double *out) {
  return guard(ctx, [&] { *out = ctx->problem->CellNorm(dev_v); });
}
int eg_cell_norm_noalloc(const double *dev_v, const double *dev_metric, long n, void *stream, double *dev_out) {
  return guard(nullptr, [&] {
    eg::launch_cell_norm_noalloc(dev_v, dev_metric, n, dev_out, static_cast<hipStream_t>(stream));
  });
}
int eg_admissible(eg_ctx *ctx, const double *dev_u, int *ok) {
  return guard(ctx, [&] { *ok = ctx->problem->Admissible(dev_u) ? 1 : 0; });
}
int eg_dg_refresh(eg_ctx *ctx, const double *dev_u) {
  return guard(ctx, [&] {
    ctx->state = dev_u;
    if (ctx->solver[0]) { ctx->solver[0]->Refresh(dev_u); }
  });
}
int eg_dg_apply(eg_ctx *ctx, const double *dev_x, double *dev_y) {
  return guard(ctx, [&] {
    require(ctx->state != nullptr, "DG operator has no refreshed state");
    auto &p = *ctx->problem;
    eg::launch_dg_apply<double>(p.geo, p.frozen, ctx->state, dev_x, dev_y, p.stream);
  });
}
int eg_dg_apply_f32(eg_ctx *ctx, const float *dev_x, float *dev_y) {
  return guard(ctx, [&] {
    require(ctx->state != nullptr, "DG operator has no refreshed state");
    auto &p = *ctx->problem;
    eg::launch_dg_apply<float>(p.geo, p.frozen, ctx->state, dev_x, dev_y, p.stream);
  });
}
int eg_pc_setup(eg_ctx *ctx, int form) { return guard(ctx, [&] { ctx->lin(form).Setup(); }); }
int eg_pc_apply(eg_ctx *ctx, int form, const double *dev_x, double *dev_y) {
  return guard(ctx, [&] { ctx->lin(form).Precondition(dev_x, dev_y); });
}
int eg_pc_apply_f32(eg_ctx *ctx, int form, const float *dev_x, float *dev_y) {
  return guard(ctx, [&] {
    require(ctx->opt.precision == "mixed", "FP32 preconditioner needs --precision mixed");
    ctx->lin(form).Precondition(dev_x, dev_y);
  });
}
int eg_hdg_condense(eg_ctx *ctx, const double *dev_u, const double *dev_r) {
  return guard(ctx, [&] { ctx->hyb().StreamCondense(dev_u, dev_r); });
}
int eg_hdg_rhs(eg_ctx *ctx, double *dev_rhs, double *dev_D) {
  return guard(ctx, [&] {
    auto &h = ctx->hyb();
    auto &p = *ctx->problem;
    if (dev_rhs) { hip_check(hipMemcpyAsync(dev_rhs, h.rhs, sizeof(double) * p.nl, hipMemcpyDeviceToDevice, p.stream), "copy"); }
    if (dev_D) { hip_check(hipMemcpyAsync(dev_D, h.D, sizeof(double) * p.nl, hipMemcpyDeviceToDevice, p.stream), "copy"); }
  });
}
int eg_hdg_xy(eg_ctx *ctx, double *dev_xy) {
  return guard(ctx, [&] {
    auto &h = ctx->hyb();
    hip_check(hipMemcpyAsync(dev_xy, h.Xy, sizeof(double) * h.Xy.n, hipMemcpyDeviceToDevice, ctx->problem->stream),
              "copy");
  });
}
int eg_schur_apply(eg_ctx *ctx, const double *dev_x, double *dev_y) {
  return guard(ctx, [&] { ctx->hyb().SchurMult(dev_x, dev_y); });
}
int eg_recover(eg_ctx *ctx, const double *dev_dl, double *dev_du) {
  return guard(ctx, [&] { ctx->hyb().Recover(dev_dl, dev_du); });
}
int eg_recovered_residual(eg_ctx *ctx, const double *dev_eps, double *dev_out, double *norm) {
  return guard(ctx, [&] {
    const double v = ctx->hyb().RecoveredResidual(dev_eps, dev_out);
    if (norm) { *norm = v; }
  });
}
int eg_ilu_order(eg_ctx *ctx, int form, int *order) {
  return guard(ctx, [&] {
    require(form == 0 || form == 1, "form must be 0 (dg) or 1 (hdg)");
    const auto graph = ctx->problem->SolverGraph(form == 1);
    std::vector<int> o(graph.size());
    if (ctx->opt.ilu_ordering == "mdf") { o = MdfOrdering(graph); }
    else { for (size_t i = 0; i < o.size(); ++i) { o[i] = int(i); } }
    std::copy(o.begin(), o.end(), order);
  });
}
int eg_linear_solve(eg_ctx *ctx, int form, const double *dev_rhs, double *dev_x, double *dev_res, double tol,
                    int refresh, eg_linear_stats *stats) {
  return guard(ctx, [&] {
    auto &s = ctx->lin(form);
    auto &p = *ctx->problem;
    LinearResult r;
    s.Solve(dev_rhs, r, tol, refresh != 0);
    hip_check(hipMemcpyAsync(dev_x, s.x, sizeof(double) * s.N, hipMemcpyDeviceToDevice, p.stream), "copy");
    if (dev_res) { hip_check(hipMemcpyAsync(dev_res, s.res, sizeof(double) * s.N, hipMemcpyDeviceToDevice, p.stream), "copy"); }
    p.Sync();
    if (stats) {
      stats->iterations = r.count;
      stats->outer = r.outer;
      stats->relative_residual = r.residual;
      stats->residual_norm = r.residual_norm;
      stats->setup_s = r.setup;
      stats->solve_s = r.solve;
    }
  });
}
int eg_kernel_time(eg_ctx *ctx, const char *kernel, int reps, double *avg_ms, double *bytes, double *flops) {
  return guard(ctx, [&] {
    auto &p = *ctx->problem;
    const std::string k(kernel);
    require(reps > 0, "reps must be positive");
    DeviceArray<double> u(p.nu), r(p.nu), x(p.nu), y(p.nu), tl(p.nl), tl2(p.nl);
    DeviceArray<float> x32(p.nu), y32(p.nu);
    const auto init = p.Initial();
    hip_check(hipMemcpyAsync(u, init.data(), sizeof(double) * p.nu, hipMemcpyHostToDevice, p.stream), "upload");
    p.Freeze(u, ctx->opt.dt, false);
    const double dt = (ctx->opt.cfl > 0 ? ctx->opt.cfl / p.WaveRate() : ctx->opt.dt) * 0.24169426078821;
    p.SetStage(u, dt);
    p.Residual(u, r);
    eg::launch_scale_copy(r, -1., x, p.nu, p.stream);
    eg::launch_convert<double, float>(x, x32, p.nu, p.stream);
    ctx->state = u;
    const double nu = p.nu, nl = p.nl, nfq = double(p.nf) * p.q;
    std::function<void()> fn;
    double b = 0., fl = 0.;
    // LU-based count of the ILU(0) factor (src/block_ilu.cpp:73-86): 2 b^3
    // per lower block (right solve) and per update product, 2/3 b^3 per
    // diagonal LU; block Jacobi: 2/3 b^3 per block.
    auto setup_flops = [](const LinearSolver &s) {
      const double b3 = double(s.block) * s.block * s.block;
      if (!s.pattern) { return 2. / 3. * b3 * (s.N / s.block); }
      const auto &pt = *s.pattern;
      return 2. * b3 * (double(pt.nnzb - pt.nb) / 2. + double(pt.upd.size() / 2)) + 2. / 3. * b3 * pt.nb;
    };
    if (k == "residual") {
      fn = [&] { eg::launch_residual(p.geo, p.frozen, u, nullptr, 0., p.old, y, nullptr, p.partial, p.flags, p.stream); };
      b = 3 * nu * 8 + nfq * 5 * 8;
    } else if (k == "dg_apply" || k == "dg_apply_f32") {
      const bool f32 = k == "dg_apply_f32";
      fn = [&, f32] {
        if (f32) { eg::launch_dg_apply<float>(p.geo, p.frozen, u, x32, y32, p.stream); }
        else { eg::launch_dg_apply<double>(p.geo, p.frozen, u, x, y, p.stream); }
      };
      b = 2 * nu * (f32 ? 4 : 8) + nu * 8 + nfq * 5 * 8;
    } else if (k == "hdg_condense") {
      fn = [&] { ctx->hyb().StreamCondense(u, r); };
      b = double(p.n) * p.m * (4 * p.t + 1) * 8 + 2 * nu * 8 + 2 * nl * 8 + double(p.nf) * 8 * p.t * p.t * 8;
      // src/condensation.cpp:131-148: LU of A_e, X_e = A_e^{-1} B_e (4t
      // columns) and y_e, then the panels C_e X_e (C_e: 4q nonzeros a row).
      const double m = p.m, t = p.t;
      fl = double(p.n) * (2. / 3. * m * m * m + 2. * m * m * (4. * t + 1.) + 2. * (4. * t) * (4. * t) * (4. * p.q));
    } else if (k == "schur_apply") {
      ctx->hyb().StreamCondense(u, r);
      hip_check(hipMemcpyAsync(tl, ctx->hyb().rhs, sizeof(double) * p.nl, hipMemcpyDeviceToDevice, p.stream), "copy");
      fn = [&] { ctx->hyb().SchurMult(tl.p, tl2.p); };
      b = double(p.nf) * 8 * p.t * p.t * 8 + 3 * nl * 8;
    } else if (k == "pc_dg" || k == "pc_setup_dg") {
      auto &s = ctx->lin(0);
      s.Setup();
      if (k == "pc_dg") { fn = [&] { s.Precondition(x.p, y.p); }; }
      else { fn = [&] { s.Setup(); }; }
      b = double(s.blocks.n) * 8 + 2 * nu * 8;
      if (k == "pc_setup_dg") { fl = setup_flops(s); }
    } else if (k == "pc_hdg" || k == "pc_setup_hdg") {
      ctx->hyb().StreamCondense(u, r);
      hip_check(hipMemcpyAsync(tl, ctx->hyb().rhs, sizeof(double) * p.nl, hipMemcpyDeviceToDevice, p.stream), "copy");
      auto &s = ctx->lin(1);
      s.Setup();
      if (k == "pc_hdg") { fn = [&] { s.Precondition(tl.p, tl2.p); }; }
      else { fn = [&] { s.Setup(); }; }
      b = double(s.blocks.n) * 8 + 2 * nl * 8;
      if (k == "pc_setup_hdg") { fl = setup_flops(s); }
    } else {
      throw std::runtime_error("Unknown kernel name: " + k);
    }
    fn(); // warm
    p.Sync();
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, p.stream), "event");
    for (int i = 0; i < reps; ++i) { fn(); }
    hip_check(hipEventRecord(e1, p.stream), "event");
    hip_check(hipEventSynchronize(e1), "event");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p.TakeFlags();
    ctx->state = nullptr;
    if (ctx->solver[0]) { ctx->solver[0]->Refresh(nullptr); }
    *avg_ms = double(ms) / reps;
    *bytes = b;
    if (flops) { *flops = fl; }
  });
}
int eg_run(eg_ctx *ctx, int form, const char *prefix, double *host_u_final, eg_run_stats *stats,
           double *host_increments, int max_updates, int *krylov) {
  return guard(ctx, [&] {
    auto &p = *ctx->problem;
    auto r = Run(p, form ? "hdg" : "dg", prefix ? prefix : "", host_increments != nullptr);
    std::memcpy(host_u_final, r.u.data(), sizeof(double) * p.nu);
    if (stats) {
      stats->updates = r.updates;
      stats->time_steps = r.time_steps;
      stats->newton_seconds = r.newton_seconds;
      stats->final_time = r.final_time;
      stats->drift = r.drift;
      stats->seconds = r.seconds;
      for (int c = 0; c < 4; ++c) { stats->errors[c] = r.errors[c]; }
    }
    const int nrec = std::min<int>(r.updates, max_updates);
    if (host_increments) {
      for (int i = 0; i < nrec; ++i) {
        std::memcpy(host_increments + size_t(i) * p.nu, r.increments[i].data(), sizeof(double) * p.nu);
      }
    }
    if (krylov) { for (int i = 0; i < nrec; ++i) { krylov[i] = r.records[i].krylov; } }
  });
}

int eg_solve_footprint(const char *args, int form, double *bytes) {
  return guard(nullptr, [&] {
    require(form == 0 || form == 1, "form must be 0 (dg) or 1 (hdg)");
    *bytes = double(SolveFootprint(ParseOptions(std::string(args ? args : "")), form));
  });
}
namespace {
void serve_into(const char *const *args, const int *forms, const double *arrivals, int n, const ServeConfig &cfg,
                eg_solve_outcome *out, double *const *states, char *errors, eg_serve_report *report,
                eg_batch_report *batch) {
  require(n >= 0 && args && forms && out, "eg_serve needs args, forms and out");
  std::vector<SolveRequest> reqs(n);
  for (int i = 0; i < n; ++i) {
    reqs[i].args = args[i] ? args[i] : "";
    reqs[i].form = forms[i];
    reqs[i].arrival = arrivals ? arrivals[i] : 0.;
  }
  ServeConfig c = cfg;
  c.keep_states = states != nullptr;
  ServeReport rep;
  auto res = Serve(reqs, c, &rep);
  for (int i = 0; i < n; ++i) {
    const auto &r = res[i];
    eg_solve_outcome o{};
    o.arrival = r.arrival;
    o.start = r.start;
    o.finish = r.finish;
    o.newton_seconds = r.newton_seconds;
    o.footprint = double(r.footprint);
    o.measured_peak = double(r.measured_peak);
    for (int k = 0; k < 4; ++k) { o.errors[k] = r.errors[k]; }
    o.status = r.status;
    o.updates = r.updates;
    o.time_steps = r.time_steps;
    o.worker = r.worker;
    out[i] = o;
    if (states && states[i] && !r.u.empty()) { std::memcpy(states[i], r.u.data(), sizeof(double) * r.u.size()); }
    if (errors) {
      std::strncpy(errors + size_t(i) * 256, r.error.c_str(), 255);
      errors[size_t(i) * 256 + 255] = 0;
    }
  }
  if (report) {
    report->wall = rep.wall;
    report->max_reserved = double(rep.max_reserved);
    report->max_concurrent = rep.max_concurrent;
  }
  if (batch) {
    *batch = eg_batch_report{};
    batch->flushes = rep.batch.flushes;
    batch->chunks = rep.batch.chunks;
    batch->graph_launches = rep.batch.graph_launches;
    batch->max_group = rep.batch.max_group;
    batch->launch_s = rep.batch.launch_s;
    batch->window_s = rep.batch.window_s;
    for (int q = 0; q <= eg::kMaxQ; ++q) { batch->tile[q] = rep.batch.tile[q]; }
  }
}
} // namespace

int eg_serve(const char *const *args, const int *forms, const double *arrivals, int n, int workers,
             double hbm_budget, eg_solve_outcome *out, double *const *states, char *errors,
             eg_serve_report *report) {
  return guard(nullptr, [&] {
    ServeConfig cfg;
    cfg.workers = workers;
    cfg.hbm_budget = hbm_budget;
    serve_into(args, forms, arrivals, n, cfg, out, states, errors, report, nullptr);
  });
}
int eg_serve_batched(const char *const *args, const int *forms, const double *arrivals, int n, int workers,
                     double hbm_budget, double window_us, eg_solve_outcome *out, double *const *states,
                     char *errors, eg_serve_report *report, eg_batch_report *batch) {
  return guard(nullptr, [&] {
    ServeConfig cfg;
    cfg.workers = workers;
    cfg.hbm_budget = hbm_budget;
    cfg.batched = true;
    if (window_us > 0.) { cfg.window_us = window_us; }
    serve_into(args, forms, arrivals, n, cfg, out, states, errors, report, batch);
  });
}
int eg_autotune_ilu(const int *ks, int nk, int form, const char *precision, char *out, double *times) {
  return guard(nullptr, [&] {
    require(form == 0 || form == 1, "form must be 0 (dg) or 1 (hdg)");
    const std::string prec = precision ? precision : "fp64";
    const char *names[3] = {"levels-256", "levels-1024", "dataflow"};
    const char *flags[3] = {"--ilu-schedule levels --sweep-threads 256", "--ilu-schedule levels --sweep-threads 1024",
                            "--ilu-schedule dataflow"};
    for (int i = 0; i < nk; ++i) {
      double best = 0.;
      int pick = 0;
      for (int c = 0; c < 3; ++c) {
        const std::string args = "--nx 16 --ny 16 --k " + std::to_string(ks[i]) +
                                 " --cfl 1 --precond block-ilu0 --precision " + prec + " " + flags[c];
        eg_ctx *ctx = nullptr;
        if (eg_create(args.c_str(), 3., nullptr, &ctx) != 0) { throw std::runtime_error(create_error); }
        std::unique_ptr<eg_ctx> hold(ctx);
        double setup = 0., apply = 0., bytes = 0.;
        if (eg_kernel_time(ctx, form ? "pc_setup_hdg" : "pc_setup_dg", 3, &setup, &bytes, nullptr) != 0 ||
            eg_kernel_time(ctx, form ? "pc_hdg" : "pc_dg", 16, &apply, &bytes, nullptr) != 0) {
          throw std::runtime_error(ctx->error);
        }
        const double t = setup + 16. * apply;
        if (times) { times[3 * i + c] = t; }
        if (c == 0 || t < best) { best = t; pick = c; }
      }
      std::strncpy(out + 16 * i, names[pick], 15);
      out[16 * i + 15] = 0;
    }
  });
}
}
