// Grouped Arnoldi steps of concurrent condensed solves (the batched form; no
// counterpart in the reference, which runs one case per process,
// src/main.cpp:17-29). Each solve keeps its own Newton/DIRK and GMRES control
// flow on its own host thread and stream (Run, src/stepper.cpp:28-292;
// GMRES, src/linear.cpp:100-131); its Arnoldi steps go to the batcher, which
// launches the steps of all solves of one trace order that are waiting at
// that moment as one replay of a recorded graph of persistent grouped
// kernels over a table of them.
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstdlib>
#include <cstring>

#include "euler.hpp"

namespace euler {

// Candidates of the virtual sweep workgroups per solve (the per-order tile),
// per row of the solve's widest dependency level, and the timed flushes per
// candidate before the choice.
static constexpr int kTileCand[] = {1, 2, 4};
static constexpr int kNumCand = int(sizeof(kTileCand) / sizeof(kTileCand[0]));
static constexpr int kTrialsPerCand = 3;
// Persistent grids of the Schur apply and the CGS2 stages (workgroups of 256
// threads walking the table's virtual blocks).
static constexpr int kWalkGrid = 2048;

struct KrylovBatcher::Entry {
  LinearSolver *s = nullptr;
  eg::ArnoldiSlot slot{};
  int q = 0, chunk = 1, width = 1;
  hipEvent_t done = nullptr; // recorded on the solve's stream after its steps
  bool f32 = false, launched = false;
  std::string error;
};

struct KrylovBatcher::Group {
  int q = 0, b = 0;
  bool f32 = false;
  hipStream_t stream = nullptr;
  eg::GroupTable *table = nullptr;        // device
  static constexpr int kRing = 8;
  eg::GroupTable *staging[kRing] = {};  // pinned host, in turn
  hipEvent_t staged[kRing] = {}, done = nullptr, t0 = nullptr, t1 = nullptr;
  int next_stage = 0;
  hipGraphExec_t exec = nullptr;
  int sweep_grid = 0;
  // online tile choice: trials so far, their cost (ms per row of the
  // table), the trial awaiting its timing, and the choice (0: exploring)
  int tile = 0, trials = 0, timing_cand = -1;
  double timing_rows = 0.;
  std::vector<double> cost[kNumCand];
  ~Group() {
    if (exec) { (void)hipGraphExecDestroy(exec); }
    for (hipEvent_t e : staged) {
      if (e) { (void)hipEventDestroy(e); }
    }
    for (hipEvent_t e : {done, t0, t1}) {
      if (e) { (void)hipEventDestroy(e); }
    }
    for (auto *h : staging) {
      if (h) { (void)hipHostFree(h); }
    }
    if (table) { (void)hipFree(table); }
    if (stream) { (void)hipStreamDestroy(stream); }
  }
};

KrylovBatcher::KrylovBatcher(double window_us) : window_us_(window_us) {
  const char *g = std::getenv("EG_GROUP_GRAPH");
  use_graph_ = !(g && std::string(g) == "0");
  const char *a = std::getenv("EG_BATCH_ALIGN_US");
  align_us_ = a ? std::atof(a) : 2000.;
  const char *wnd = std::getenv("EG_BATCH_WINDOW_US");
  if (wnd) { window_us_ = std::atof(wnd); }
  thread_ = std::thread([this] { Loop(); });
}

KrylovBatcher::~KrylovBatcher() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  wake_.notify_all();
  thread_.join();
  for (auto &g : groups_) {
    if (g->stream) { (void)hipStreamSynchronize(g->stream); }
  }
  groups_.clear();
  for (double *m : mirrors_) { (void)hipHostFree(m); }
  (void)hipGetLastError();
}

double *KrylovBatcher::AcquireMirror() {
  std::lock_guard<std::mutex> lk(mu_);
  if (free_mirrors_.empty()) {
    double *m = nullptr;
    hip_check(hipHostMalloc(&m, sizeof(double) * eg::kMirrorSize, hipHostMallocDefault), "hipHostMalloc");
    mirrors_.push_back(m);
    return m;
  }
  double *m = free_mirrors_.back();
  free_mirrors_.pop_back();
  return m;
}
void KrylovBatcher::ReleaseMirror(double *m) {
  std::lock_guard<std::mutex> lk(mu_);
  free_mirrors_.push_back(m);
}

void KrylovBatcher::Join() {
  std::lock_guard<std::mutex> lk(mu_);
  ++members_;
}
void KrylovBatcher::Depart() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    --members_;
  }
  align_.notify_all();
}
void KrylovBatcher::Align() {
  if (align_us_ <= 0.) { return; }
  std::unique_lock<std::mutex> lk(mu_);
  const long gen = align_gen_;
  if (++arrived_ >= members_) {
    arrived_ = 0;
    ++align_gen_;
    align_.notify_all();
    return;
  }
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double, std::micro>(align_us_);
  while (align_gen_ == gen) {
    if (arrived_ >= members_) { // a member departed meanwhile
      arrived_ = 0;
      ++align_gen_;
      align_.notify_all();
      break;
    }
    if (align_.wait_until(lk, until) == std::cv_status::timeout) {
      if (align_gen_ == gen) { --arrived_; }
      break;
    }
  }
}

void KrylovBatcher::Enter() {
  std::lock_guard<std::mutex> lk(mu_);
  ++in_krylov_;
}
void KrylovBatcher::Leave() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    --in_krylov_;
  }
  wake_.notify_all();
}

BatchReport KrylovBatcher::Report() {
  std::lock_guard<std::mutex> lk(mu_);
  return report_;
}

template <class T>
void KrylovBatcher::Submit(LinearSolver &s, T *V, T *z, T *w, const T *metric, int chunk, hipEvent_t done) {
  Entry e;
  e.s = &s;
  e.done = done;
  e.width = std::max(1, s.pattern->max_width);
  e.q = s.p.q;
  e.chunk = std::max(1, chunk);
  e.f32 = sizeof(T) == 4;
  auto &sl = e.slot;
  const auto &pt = *s.pattern;
  sl.V = V;
  sl.z = z;
  sl.w = w;
  sl.metric = metric;
  sl.bsr = s.Factors<T>();
  sl.mid = s.SweepBuffer<T>();
  sl.row_ptr = pt.d_row;
  sl.diag_ptr = pt.d_diag;
  sl.col = pt.d_col;
  sl.lower_rows = pt.d_lower_rows;
  sl.upper_rows = pt.d_upper_rows;
  sl.ticket = s.sweep_ticket;
  sl.flags = s.p.flags;
  sl.panels = e.f32 ? static_cast<const void *>(s.hybrid->panels32.p) : static_cast<const void *>(s.hybrid->panels.p);
  sl.D = s.hybrid->D;
  sl.partA = s.partA;
  sl.partB = s.partB;
  sl.scal = s.p.scalars;
  sl.cycle = s.cycle;
  sl.mirror = s.mirror;
  sl.N = s.N;
  sl.nb = pt.nb;
  sl.scaled = s.scaled_upper ? 1 : 0;
  sl.nx = s.p.opt.nx;
  sl.ny = s.p.opt.ny;
  sl.nf = s.p.nf;
  hip_check(hipEventRecord(s.batch_ready, s.p.stream), "hipEventRecord");
  std::unique_lock<std::mutex> lk(mu_);
  queue_.push_back(&e);
  wake_.notify_all();
  launched_.wait(lk, [&] { return e.launched; });
  if (!e.error.empty()) { throw std::runtime_error(e.error); }
}
template void KrylovBatcher::Submit<double>(LinearSolver &, double *, double *, double *, const double *, int,
                                            hipEvent_t);
template void KrylovBatcher::Submit<float>(LinearSolver &, float *, float *, float *, const float *, int, hipEvent_t);

void KrylovBatcher::Loop() {
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    wake_.wait(lk, [&] { return stop_ || !queue_.empty(); });
    if (queue_.empty()) {
      if (stop_) { break; }
      continue;
    }
    // wait, up to the window, for the steps of the other solves inside a
    // Krylov iteration
    const double w0 = clock_s();
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double, std::micro>(window_us_);
    while (!stop_ && int(queue_.size()) < in_krylov_) {
      if (wake_.wait_until(lk, until) == std::cv_status::timeout) { break; }
    }
    std::vector<Entry *> batch;
    batch.swap(queue_);
    lk.unlock();
    const double w1 = clock_s();
    Flush(batch);
    const double w2 = clock_s();
    lk.lock();
    report_.window_s += w1 - w0;
    report_.launch_s += w2 - w1;
    for (Entry *e : batch) { e->launched = true; }
    launched_.notify_all();
  }
}

KrylovBatcher::Group &KrylovBatcher::GroupFor(int q, bool f32) {
  for (auto &g : groups_) {
    if (g->q == q && g->f32 == f32) { return *g; }
  }
  auto g = std::make_unique<Group>();
  g->q = q;
  g->b = 4 * q;
  g->f32 = f32;
  hip_check(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipMalloc(&g->table, sizeof(eg::GroupTable)), "hipMalloc");
  for (auto *&h : g->staging) { hip_check(hipHostMalloc(&h, sizeof(eg::GroupTable), hipHostMallocDefault), "hipHostMalloc"); }
  for (hipEvent_t &e : g->staged) { hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"); }
  hip_check(hipEventCreateWithFlags(&g->done, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreate(&g->t0), "hipEventCreate");
  hip_check(hipEventCreate(&g->t1), "hipEventCreate");
  g->sweep_grid = std::max(1, f32 ? eg::group_sweep_capacity<float>(g->b) : eg::group_sweep_capacity<double>(g->b));
  groups_.push_back(std::move(g));
  return *groups_.back();
}

void KrylovBatcher::Flush(std::vector<Entry *> &batch) {
  // group by (trace order, arithmetic), in submission order
  std::vector<Entry *> rest = batch;
  while (!rest.empty()) {
    const int q = rest.front()->q;
    const bool f32 = rest.front()->f32;
    std::vector<Entry *> take, keep;
    for (Entry *e : rest) {
      ((e->q == q && e->f32 == f32 && int(take.size()) < eg::kGroupMax) ? take : keep).push_back(e);
    }
    rest.swap(keep);
    try {
      Launch(GroupFor(q, f32), take.data(), int(take.size()));
    } catch (const std::exception &ex) {
      (void)hipGetLastError();
      for (Entry *e : take) { e->error = ex.what(); }
    }
  }
}

void KrylovBatcher::Launch(Group &g, Entry *const *entries, int count) {
  // the tile: the timed trial of a candidate while exploring, else the
  // choice; a trial whose timing is not in yet is read at a later flush
  // (this thread never waits for the device)
  bool timing_free = g.timing_cand < 0;
  if (g.timing_cand >= 0) {
    const hipError_t q = hipEventQuery(g.t1);
    if (q == hipSuccess) { timing_free = true; }
    else { (void)hipGetLastError(); }
  }
  if (g.timing_cand >= 0 && timing_free) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, g.t0, g.t1), "hipEventElapsedTime");
    g.cost[g.timing_cand].push_back(double(ms) / std::max(1., g.timing_rows)); // per row and step
    g.timing_cand = -1;
    if (g.trials == kNumCand * kTrialsPerCand) {
      double best = 0.;
      for (int c = 0; c < kNumCand; ++c) {
        auto v = g.cost[c];
        std::sort(v.begin(), v.end());
        const double med = v[v.size() / 2];
        if (c == 0 || med < best) {
          best = med;
          g.tile = kTileCand[c];
        }
      }
    }
  }
  const bool explore = g.tile == 0 && timing_free;
  int replays = 1;
  for (int i = 0; i < count; ++i) { replays = std::max(replays, entries[i]->chunk); }
  const int cand = explore ? g.trials % kNumCand : -1;
  // while a trial's timing is pending: the candidate of that trial
  const int tile = explore ? kTileCand[cand] : (g.tile ? g.tile : kTileCand[g.timing_cand]);
  for (int i = 0; i < count; ++i) {
    // a solve whose earlier work is already done needs no wait (in the
    // steady state its stream holds only the previous grouped step)
    const hipError_t q = hipEventQuery(entries[i]->s->batch_ready);
    if (q == hipSuccess) { continue; }
    (void)hipGetLastError(); // hipErrorNotReady is not an error here
    hip_check(hipStreamWaitEvent(g.stream, entries[i]->s->batch_ready, 0), "hipStreamWaitEvent");
  }
  const int k = g.next_stage;
  g.next_stage = (k + 1) % Group::kRing;
  hip_check(hipEventSynchronize(g.staged[k]), "hipEventSynchronize"); // the copy that last read this buffer
  eg::GroupTable &t = *g.staging[k];
  t.count = count;
  t.sweep_off[0] = t.schur_off[0] = t.red_off[0] = t.norm_off[0] = 0;
  double rows = 0., want = 0.;
  for (int i = 0; i < count; ++i) { want += std::min(entries[i]->slot.nb, tile * entries[i]->width); }
  // a virtual workgroup holds its physical one for the whole sweep of its
  // solve: keep them all resident at once (scaled down to the grid)
  const double scale = want > g.sweep_grid ? g.sweep_grid / want : 1.;
  for (int i = 0; i < count; ++i) {
    const eg::ArnoldiSlot &sl = entries[i]->slot;
    t.slot[i] = sl;
    const int vw = int(scale * std::min(sl.nb, tile * entries[i]->width));
    t.sweep_off[i + 1] = t.sweep_off[i] + std::max(1, vw);
    t.schur_off[i + 1] = t.schur_off[i] + (sl.nf + 3) / 4;
    t.red_off[i + 1] = t.red_off[i] + eg::reduce_blocks(sl.N);
    t.norm_off[i + 1] = t.norm_off[i] + eg::elementwise_grid(sl.N);
    rows += sl.nb;
  }
  const size_t bytes = offsetof(eg::GroupTable, slot) + sizeof(eg::ArnoldiSlot) * size_t(count);
  hip_check(hipMemcpyAsync(g.table, &t, bytes, hipMemcpyHostToDevice, g.stream), "hipMemcpyAsync");
  hip_check(hipEventRecord(g.staged[k], g.stream), "hipEventRecord");
  auto record = [&](hipStream_t st) {
    bool ok = true;
    if (g.f32) {
      ok = ok && eg::launch_group_sweep<float>(true, g.b, g.table, g.sweep_grid, st);
      ok = ok && eg::launch_group_sweep<float>(false, g.b, g.table, g.sweep_grid, st);
      ok = ok && eg::launch_group_schur<float>(g.q, g.table, kWalkGrid, st);
      for (int stage = 0; stage < 4 && ok; ++stage) { eg::launch_group_krylov<float>(stage, g.table, kWalkGrid, st); }
      eg::launch_group_tail(g.table, st);
    } else {
      ok = ok && eg::launch_group_sweep<double>(true, g.b, g.table, g.sweep_grid, st);
      ok = ok && eg::launch_group_sweep<double>(false, g.b, g.table, g.sweep_grid, st);
      ok = ok && eg::launch_group_schur<double>(g.q, g.table, kWalkGrid, st);
      for (int stage = 0; stage < 4 && ok; ++stage) { eg::launch_group_krylov<double>(stage, g.table, kWalkGrid, st); }
      eg::launch_group_tail(g.table, st);
    }
    require(ok, "Launch of a grouped Arnoldi step failed");
  };
  if (use_graph_ && !g.exec) {
    hipGraph_t graph = nullptr;
    hip_check(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      record(g.stream);
    } catch (...) {
      (void)hipStreamEndCapture(g.stream, &graph);
      if (graph) { (void)hipGraphDestroy(graph); }
      throw;
    }
    hip_check(hipStreamEndCapture(g.stream, &graph), "hipStreamEndCapture");
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    hip_check(e, "hipGraphInstantiate");
  }
  if (explore) { hip_check(hipEventRecord(g.t0, g.stream), "hipEventRecord"); }
  // one replay per step: a solve whose cycle has ended is skipped by every kernel
  for (int r = 0; r < replays; ++r) {
    if (use_graph_) {
      hip_check(hipGraphLaunch(g.exec, g.stream), "hipGraphLaunch");
    } else {
      record(g.stream);
    }
  }
  if (explore) {
    hip_check(hipEventRecord(g.t1, g.stream), "hipEventRecord");
    g.timing_cand = cand;
    g.timing_rows = rows * replays;
    ++g.trials;
  }
  hip_check(hipEventRecord(g.done, g.stream), "hipEventRecord");
  for (int i = 0; i < count; ++i) {
    LinearSolver &s = *entries[i]->s;
    hip_check(hipStreamWaitEvent(s.p.stream, g.done, 0), "hipStreamWaitEvent");
    hip_check(hipEventRecord(entries[i]->done, s.p.stream), "hipEventRecord");
  }
  std::lock_guard<std::mutex> lk(mu_);
  ++report_.flushes;
  report_.chunks += count;
  report_.graph_launches += use_graph_ ? replays : 0;
  report_.max_group = std::max(report_.max_group, count);
  if (g.q <= eg::kMaxQ) { report_.tile[g.q] = tile; }
}

} // namespace euler
