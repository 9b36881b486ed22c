// Batched and served execution of independent solves on one GPU.
//
// Each admitted solve runs a whole Run (Newton/DIRK with its preconditioner,
// Krylov basis and, for HDG, condensed factors resident on the device) on a
// worker thread that owns a HIP stream, so solves of different sizes overlap
// on the device. A scheduler thread admits arriving requests in arrival order
// when a worker is free and the request's device footprint fits in the HBM
// budget next to the solves already running (continuous batching: a request
// that arrives while others run starts as soon as a slot and memory free up).
// Device memory of a solve is allocated and freed stream-ordered on its
// worker's stream, so a finishing solve does not synchronise the device.
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <mutex>
#include <thread>

#include "euler.hpp"

namespace euler {

AllocStats &thread_allocs() {
  thread_local AllocStats stats;
  return stats;
}

// Upper bound of the device bytes one Run holds at its peak (Problem, the
// Newton vectors of Run, the HDG condensed factors, the Krylov solver and its
// preconditioner blocks), from the same sizes the constructors use.
size_t SolveFootprint(const Options &o, int form) {
  const double n = double(o.nx) * o.ny, q = o.k + 1, p = q * q, m = 4 * p, t = 4 * q, nf = 2 * n;
  const double nu = n * m, nl = nf * t;
  double bytes = 0.;
  bytes += nf * q * 64 + nu * (8 + 12) + nl * 12 + 2 * 8 * 4096 + 4096; // Problem
  bytes += 15 * nu * 8;                                                   // Run vectors
  const bool mixed = o.precision == "mixed", ilu = o.precond == "block-ilu0";
  if (form == 1) {
    bytes += n * m * m * 8 + n * 4 * q * 16 * 8 + n * m * (4 * t + 1) * 8 + nf * 8 * t * t * (mixed ? 12 : 8) +
             n * 4 * t * 8 + 2 * nl * 8; // Hybrid
  }
  const double N = form == 0 ? nu : nl, b = form == 0 ? m : t;
  const double restart = std::min<double>(N, o.restart);
  bytes += (restart + 1) * N * (mixed ? 12 : 8) + 8 * N * (mixed ? 12 : 8) + 2 * 64 * 512 * 8; // Krylov
  const double nnzb = ilu ? (form == 0 ? 5 * n : 8 * nf) : N / b;
  bytes += nnzb * b * b * (mixed ? 12 : 8);
  if (ilu) {
    bytes += N * (mixed ? 12 : 8) + 64 * nnzb;                                  // sweep vectors, pattern
    if (b > 100) { bytes += 2 * std::max(o.nx, o.ny) * 2 * b * b * 8; }         // large-block scratch
  }
  if (o.check_update && form == 1) { bytes += (o.restart + 9) * nu * 8 + 5 * n * m * m * 8; }
  // assembled operator (--assembly full, element for DG), dense LU (--solver dense|direct)
  const bool direct = o.solver != "gmres";
  if (direct || o.assembly == "full" || (form == 0 && o.assembly == "element")) {
    bytes += (form == 0 ? 5 * n : 8 * nf) * b * b * (mixed ? 12 : 8);
  }
  if (direct) { bytes += N * N * 8 + 6 * N * 8; }
  if (o.ilu_ordering == "mdf") { bytes += 2 * N * (mixed ? 12 : 8); }
  return size_t(bytes * 1.05) + (size_t(16) << 20);
}

std::vector<SolveOutcome> Serve(const std::vector<SolveRequest> &requests, const ServeConfig &cfg,
                                ServeReport *report) {
  const int nreq = int(requests.size());
  std::vector<SolveOutcome> out(nreq);
  require(cfg.workers >= 1 && cfg.workers <= (cfg.batched ? eg::kGroupMax : 64),
          cfg.batched ? "The batched form needs 1 to 256 workers" : "Serve needs 1 to 64 workers");
  double budget = cfg.hbm_budget;
  if (budget <= 0.) {
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    budget = 0.8 * double(free_b);
  }
  std::vector<Options> opts(nreq);
  std::vector<int> order;
  for (int i = 0; i < nreq; ++i) {
    out[i].arrival = requests[i].arrival;
    try {
      require(requests[i].form == 0 || requests[i].form == 1, "form must be 0 (dg) or 1 (hdg)");
      opts[i] = ParseOptions(requests[i].args);
      out[i].footprint = SolveFootprint(opts[i], requests[i].form);
      if (double(out[i].footprint) > budget) {
        out[i].status = 1;
        out[i].error = "Request exceeds the HBM budget";
      } else {
        order.push_back(i);
      }
    } catch (const std::exception &e) {
      out[i].status = 1;
      out[i].error = e.what();
    }
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return requests[a].arrival < requests[b].arrival; });

  std::mutex mu;
  std::condition_variable wake_sched, wake_work;
  std::deque<int> ready;
  int idle = cfg.workers, running = 0, done = 0, max_running = 0;
  double reserved = 0., max_reserved = 0.;
  bool stop = false;
  // EG_SERVE_PROGRESS=1: one line on stderr per finished request
  const char *progress_env = std::getenv("EG_SERVE_PROGRESS");
  const bool progress = progress_env && std::string(progress_env) == "1";
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };

  // The workers' streams are made here, before any thread starts, so that a
  // failure ends Serve with an error instead of inside a worker thread.
  std::vector<hipStream_t> streams(cfg.workers, nullptr);
  for (auto &st : streams) {
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      for (auto &made : streams) {
        if (made) { (void)hipStreamDestroy(made); }
      }
      throw std::runtime_error("Serve could not create its worker streams");
    }
  }
  std::unique_ptr<KrylovBatcher> batcher;
  if (cfg.batched) { batcher = std::make_unique<KrylovBatcher>(cfg.window_us); }
  auto worker = [&](int wid) {
    hipStream_t stream = streams[wid];
    for (;;) {
      int i = -1;
      {
        std::unique_lock<std::mutex> lk(mu);
        wake_work.wait(lk, [&] { return stop || !ready.empty(); });
        if (ready.empty()) { break; }
        i = ready.front();
        ready.pop_front();
      }
      SolveOutcome &r = out[i];
      r.worker = wid;
      r.start = now();
      {
        AllocScope scope(stream);
        auto &acc = thread_allocs();
        const size_t base = acc.live;
        acc.peak = base;
        try {
          EulerGas gas(requests[i].gamma);
          Problem p(opts[i], gas, stream);
          p.batcher = batcher.get();
          RunResult res = Run(p, requests[i].form ? "hdg" : "dg", "", false, cfg.keep_states);
          r.updates = res.updates;
          r.time_steps = res.time_steps;
          r.newton_seconds = res.newton_seconds;
          r.errors = res.errors;
          if (cfg.keep_states) { r.u = std::move(res.u); }
        } catch (const std::exception &e) {
          r.status = 1;
          r.error = e.what();
        }
        r.measured_peak = acc.peak - base;
      }
      (void)hipStreamSynchronize(stream);
      r.finish = now();
      {
        std::lock_guard<std::mutex> lk(mu);
        reserved -= double(r.footprint);
        ++idle;
        --running;
        ++done;
        if (progress) {
          std::fprintf(stderr, "[serve] %d/%d done at %.1f s (%d running)%s%s\n", done, int(order.size()), r.finish,
                       running, r.status ? ": failed: " : "", r.status ? r.error.c_str() : "");
        }
      }
      wake_sched.notify_one();
    }
  };
  std::vector<std::thread> pool;
  for (int w = 0; w < cfg.workers; ++w) { pool.emplace_back(worker, w); }

  size_t next = 0; // position in arrival order
  std::deque<int> pending;
  {
    std::unique_lock<std::mutex> lk(mu);
    while (done < int(order.size())) {
      const double t = now();
      while (next < order.size() && requests[order[next]].arrival <= t) { pending.push_back(order[next++]); }
      while (!pending.empty()) {
        const int i = pending.front();
        if (idle == 0 || reserved + double(out[i].footprint) > budget) { break; }
        pending.pop_front();
        reserved += double(out[i].footprint);
        max_reserved = std::max(max_reserved, reserved);
        --idle;
        ++running;
        max_running = std::max(max_running, running);
        ready.push_back(i);
        wake_work.notify_one();
      }
      if (done >= int(order.size())) { break; }
      if (next < order.size() && pending.empty()) {
        const double wait = requests[order[next]].arrival - t;
        wake_sched.wait_for(lk, std::chrono::duration<double>(std::max(wait, 0.)));
      } else {
        wake_sched.wait_for(lk, std::chrono::milliseconds(50));
      }
    }
    stop = true;
  }
  wake_work.notify_all();
  for (auto &th : pool) { th.join(); }
  if (batcher) {
    if (report) { report->batch = batcher->Report(); }
    batcher.reset();
  }
  for (auto &st : streams) { (void)hipStreamDestroy(st); }
  // Give the memory the solves cached in the stream-ordered pool back.
  int dev = 0;
  hipMemPool_t mempool = nullptr;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&mempool, dev) == hipSuccess) {
    (void)hipDeviceSynchronize();
    (void)hipMemPoolTrimTo(mempool, 0);
  }
  (void)hipGetLastError();
  if (report) {
    report->wall = now();
    report->max_concurrent = max_running;
    report->max_reserved = size_t(max_reserved);
  }
  return out;
}

} // namespace euler
