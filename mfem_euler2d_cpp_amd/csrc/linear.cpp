// Device-resident weighted restarted GMRES (src/linear.cpp:76-157) with
// block-Jacobi or block ILU(0) preconditioning, for the DG body system and
// the condensed Schur system. Krylov vectors, operator actions, the
// preconditioner and all reductions run on the GPU; only the (j+1)-vectors
// of CGS2 projections and the norms cross to the host, where the Hessenberg
// Givens rotations are applied exactly as in the reference.
//
// precision "mixed": an FP64 refinement loop (true residual with the FP64
// operator) around an inner GMRES whose Krylov basis, operator action and
// preconditioner run in FP32.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <map>

#include "euler.hpp"

namespace euler {
// Workgroups of a dataflow ILU(0) launch per row of its widest dependency
// level (0: as many as the device holds). More only poll; fewer leave the
// device to concurrent solves: measured at 4, +31% updates/s for 12
// concurrent small condensed solves and -5% / -7% of the face-block ILU(0)
// application time alone at 16x16 Q3 / 128x128 Q6.
constexpr int kDataflowWidth = 4;
// Workgroups a block row of the row-sliced sweeps (100, 144 and 196 cell
// blocks; EG_SWEEP_SLICES overrides, 0: for 100 the register-resident sweep,
// above it the streamed sweep with unscaled upper blocks). Measured at
// 128x128 Q6 (196 blocks, FP64): unsliced 12.97 ms, two 13.0, four 8.34,
// eight 8.11 ms an application; 64x64 Q5: 3.19, 2.80, 1.48, 1.47 ms; 64x64
// Q4 (100 blocks): register-resident 0.854, four 0.764, eight 1.41 ms.
constexpr int kSweepSlices = 8, kSweepSlices100 = 4;

// src/linear.cpp:27-57,162-166
LinearSolver::LinearSolver(Problem &prob, int f, Hybrid *hyb)
    : p(prob), form(f), N(f == 0 ? prob.nu : prob.nl), block(f == 0 ? prob.m : prob.t),
      restart(std::min(N, prob.opt.restart)), mixed(prob.opt.precision == "mixed"),
      ilu(prob.opt.precond == "block-ilu0"), hybrid(hyb) {
  direct = p.opt.solver == "dense" || p.opt.solver == "direct";
  require(p.opt.solver == "gmres" || direct, "Invalid solver");
  assembled = direct || p.opt.assembly == "full" || (form == 0 && p.opt.assembly == "element");
  if (direct) {
    ilu = false;
    mixed = false;
    // the reference's dense solver is limited to 5000 unknowns
    // (src/linear.cpp:46-49); its direct solver (PARDISO) runs the same
    // dense LU here and keeps the limit
    require(N <= 5000, p.opt.solver == "dense" ? "Dense direct solve limited to 5000 unknowns"
                                               : "Direct solve on the device (dense LU) limited to 5000 unknowns");
  }
  require(restart + 1 <= 64, "GMRES restart above 63 is not supported by the device solver");
  require(form == 0 || hybrid != nullptr, "Schur solver needs its condensed state");
  V.resize(size_t(N) * (restart + 1));
  for (hipEvent_t &e : step_done) { hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"); }
  for (hipEvent_t &e : setup_mark) { hip_check(hipEventCreate(&e), "hipEventCreate"); }
  batcher = p.batcher;
  if (batcher && form == 1) {
    hip_check(hipEventCreateWithFlags(&batch_ready, hipEventDisableTiming), "hipEventCreate");
    cycle.resize(1);
    mirror = batcher->AcquireMirror();
  }
  for (auto *v : {&w, &z, &comb, &corr, &res, &x, &act}) { v->resize(N); }
  partA.resize(size_t(64) * eg::kReduceBlocks);
  partB.resize(size_t(64) * eg::kReduceBlocks);
  if (mixed) {
    V32.resize(size_t(N) * (restart + 1));
    for (auto *v : {&w32, &z32, &comb32, &corr32, &res32, &x32, &b32}) { v->resize(N); }
  }
  const int nblk = N / block;
  if (assembled) {
    op_pattern = std::make_unique<BlockPattern>(p.SolverGraph(form == 1), block, p.stream);
    op_blocks.resize(size_t(op_pattern->nnzb) * block * block);
    if (mixed) { op_blocks32.resize(op_blocks.n); }
    FillMaps(*op_pattern, op_dslot, op_slot, op_contrib);
  }
  if (direct) {
    dense.resize(size_t(N) * N);
    dense_mult.resize(N);
    dense_work.resize(2 * size_t(N));
    dense_used.resize(N);
    dense_piv.resize(N);
    dense_step.resize(N);
  }
  if (ilu) {
    const auto block_graph = p.SolverGraph(form == 1);
    pattern = std::make_unique<BlockPattern>(
        block_graph, block, p.stream, p.opt.ilu_ordering == "mdf" ? MdfOrdering(block_graph) : std::vector<int>{});
    if (!pattern->perm.empty()) {
      perm_in.resize(N);
      perm_out.resize(N);
      if (mixed) {
        perm_in32.resize(N);
        perm_out32.resize(N);
      }
    }
    blocks.resize(size_t(pattern->nnzb) * block * block);
    // EG_ILU_SCHEDULE: "levels" (one launch per dependency level for the
    // factor and the sweeps), "dataflow" (default: single-launch sweeps,
    // per-level factor) or "dataflow-all" (single-launch factor as well).
    const char *sched = std::getenv("EG_ILU_SCHEDULE");
    const std::string mode = !p.opt.ilu_schedule.empty() ? p.opt.ilu_schedule : (sched ? sched : "dataflow");
    require(mode == "levels" || mode == "dataflow" || mode == "dataflow-all", "Invalid EG_ILU_SCHEDULE");
    dataflow = mode != "levels";
    dataflow_factor = mode == "dataflow-all";
    const char *graph = std::getenv("EG_FACTOR_GRAPH");
    use_graph = !(graph && std::string(graph) == "0");
    const char *rdf = std::getenv("EG_RESIDENT_DF");
    resident_df = !(rdf && std::string(rdf) == "0");
    done_factor.resize(pattern->nb + 2); // row flags, then the single-launch factors' ticket counters
    hip_check(hipMemsetAsync(done_factor, 0, sizeof(int) * done_factor.n, p.stream), "memset");
    sweep.resize(N);
    if (mixed) { sweep32.resize(N); }
    if (eg::ilu_resident_ok(block)) { PlanResident(); }
    if (resident && dataflow && resident_df) { factor_pub.resize(blocks.n); }
    const char *width = std::getenv("EG_DF_WIDTH");
    const int factor = width ? std::atoi(width) : kDataflowWidth;
    df_grid = factor > 0 ? factor * std::max(1, pattern->max_width) : 0;
    if (df_grid > 0 && !width) {
      // at most one workgroup a CU by default (128x128 Q6 face sweeps:
      // 1.57 -> 1.47 ms an application; the 16x16 Q3 grids are smaller)
      int dev = 0, cus = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) {
        df_grid = std::min(df_grid, cus);
      }
      (void)hipGetLastError();
    }
    sweep_ticket.resize(2);
    hip_check(hipMemsetAsync(sweep_ticket, 0, sizeof(int) * 2, p.stream), "memset");
    if (eg::big_block(block) && !eg::ilu_slab_ok(block)) { PlanBigFactor(); }
    {
      // the cell stencil: each lower block updates only the diagonal
      const auto &pt = *pattern;
      star_rows = true;
      for (int i = 0; i < pt.nb && star_rows; ++i) {
        star_rows = pt.diag[i] - pt.row[i] <= eg::ilu_star_max_lower();
        for (int ik = pt.row[i]; ik < pt.diag[i] && star_rows; ++ik) {
          star_rows = pt.upd_ptr[ik + 1] - pt.upd_ptr[ik] == 1 && pt.upd[2 * pt.upd_ptr[ik]] == pt.diag[i];
        }
      }
      const char *st = std::getenv("EG_SLAB_STAR");
      if (st && std::string(st) == "0") { star_rows = false; }
    }
    // EG_SCALED_UPPER=0 keeps U_ij unscaled (two products a row in the
    // backward sweep, as the reference writes it).
    const char *su = std::getenv("EG_SCALED_UPPER");
    scaled_upper = eg::ilu_scale_upper_ok(block) && !(su && std::string(su) == "0");
    if (scaled_upper && dataflow && eg::ilu_sweep_slice_ok(block, pattern->max_deps)) {
      const char *sl = std::getenv("EG_SWEEP_SLICES");
      slices = sl ? std::atoi(sl) : (block > 100 ? kSweepSlices : kSweepSlices100);
    }
    // the single-launch sweeps of blocks above 100 take scaled upper blocks
    // only in their row-sliced form (the streamed form uses U_ij as it is)
    if (block > 100 && dataflow && slices <= 0) { scaled_upper = false; }
    if (scaled_upper) {
      const auto &pt = *pattern;
      std::vector<int> us, ud;
      for (int i = 0; i < pt.nb; ++i) {
        for (int ij = pt.diag[i] + 1; ij < pt.row[i + 1]; ++ij) {
          us.push_back(ij);
          ud.push_back(pt.diag[i]);
        }
      }
      d_up_slot.resize(us.size());
      d_up_diag.resize(ud.size());
      if (!us.empty()) {
        hip_check(hipMemcpyAsync(d_up_slot, us.data(), sizeof(int) * us.size(), hipMemcpyHostToDevice, p.stream),
                  "upload");
        hip_check(hipMemcpyAsync(d_up_diag, ud.data(), sizeof(int) * ud.size(), hipMemcpyHostToDevice, p.stream),
                  "upload");
      }
      p.Sync(); // the host vectors go out of scope
    }
    if (slices > 0) {
      sweep_alt.resize(N);
      if (mixed) { sweep32_alt.resize(N); }
      const char *gc = std::getenv("EG_SWEEP_GIVEN_C");
      given_c = !(gc && std::string(gc) == "0");
      if (given_c) {
        sweep_c.resize(N);
        if (mixed) { sweep32_c.resize(N); }
      }
    }
    // EG_FACTOR32=0: the mixed form factors in FP64 and converts the factor
    const char *f32 = std::getenv("EG_FACTOR32");
    const char *slab = std::getenv("EG_SLAB");
    factor32 = mixed && eg::ilu_factor32_ok(block) && !resident && !dataflow_factor &&
               !(slab && std::string(slab) == "vector") && !(f32 && std::string(f32) == "0");
    FillMaps(*pattern, dslot, slot, contrib);
  } else if (!direct) {
    blocks.resize(size_t(nblk) * block * block);
  }
  if (mixed) { blocks32.resize(blocks.n); }
  h.assign(size_t(restart + 1) * restart, 0.);
  cosine.assign(restart, 0.);
  sine.assign(restart, 0.);
  rotated.assign(restart + 1, 0.);
  coefficients.assign(restart, 0.);
  p.Sync();
}

void LinearSolver::FillMaps(const BlockPattern &pt, DeviceArray<int> &ds_dev, DeviceArray<int> &sl_dev,
                            DeviceArray<int> &ct_dev) {
  const int nx = p.opt.nx, ny = p.opt.ny;
  if (form == 0) {
    std::vector<int> ds(p.n), sl(4 * p.n);
    for (int e = 0; e < p.n; ++e) {
      const int ix = e % nx, iy = e / nx;
      const int nb[4] = {iy * nx + (ix + nx - 1) % nx, iy * nx + (ix + 1) % nx, ((iy + ny - 1) % ny) * nx + ix,
                         ((iy + 1) % ny) * nx + ix};
      ds[e] = pt.Find(e, e);
      for (int s = 0; s < 4; ++s) { sl[4 * e + s] = pt.Find(e, nb[s]); }
    }
    ds_dev.resize(ds.size());
    sl_dev.resize(sl.size());
    hip_check(hipMemcpyAsync(ds_dev, ds.data(), sizeof(int) * ds.size(), hipMemcpyHostToDevice, p.stream), "upload");
    hip_check(hipMemcpyAsync(sl_dev, sl.data(), sizeof(int) * sl.size(), hipMemcpyHostToDevice, p.stream), "upload");
    p.Sync(); // the host vectors go out of scope
    return;
  }
  auto face = [&](int e, int side) {
    const int ix = e % nx, iy = e / nx;
    switch (side) {
    case 0: return 2 * (iy * nx + (ix + nx - 1) % nx);
    case 1: return 2 * e;
    case 2: return 2 * (((iy + ny - 1) % ny) * nx + ix) + 1;
    default: return 2 * e + 1;
    }
  };
  std::vector<int> ct(8 * p.nf);
  for (int f = 0; f < p.nf; ++f) {
    const int e = f / 2, ix = e % nx, iy = e / nx;
    const int er = (f & 1) ? ((iy + 1) % ny) * nx + ix : iy * nx + (ix + 1) % nx;
    const int first = std::min(e, er), second = std::max(e, er);
    for (int o = 0; o < 4; ++o) {
      ct[8 * f + o] = pt.Find(f, face(first, o));
      ct[8 * f + 4 + o] = pt.Find(f, face(second, o));
    }
  }
  ct_dev.resize(ct.size());
  hip_check(hipMemcpyAsync(ct_dev, ct.data(), sizeof(int) * ct.size(), hipMemcpyHostToDevice, p.stream), "upload");
  p.Sync(); // the host vector goes out of scope
}

// Large blocks: one workgroup per row is too few at a dependency level, so
// the level's products are split per lower block and per updated block when
// no update of a row targets one of its own lower blocks.
void LinearSolver::PlanBigFactor() {
  const auto &pt = *pattern;
  const int nlev = int(pt.lower_level_ptr.size()) - 1;
  bool independent = true;
  for (int i = 0; i < pt.nb && independent; ++i) {
    for (int ik = pt.row[i]; ik < pt.diag[i]; ++ik) {
      for (int u = pt.upd_ptr[ik]; u < pt.upd_ptr[ik + 1]; ++u) {
        if (pt.col[pt.upd[2 * u]] < i) { independent = false; }
      }
    }
  }
  int widest = 0;
  std::vector<int> a_ik, a_kk, t_slot, t_pair_ptr{0}, pairs;
  a_ptr.assign(1, 0);
  t_ptr.assign(1, 0);
  for (int l = 0; l < nlev; ++l) {
    const int first = pt.lower_level_ptr[l], last = pt.lower_level_ptr[l + 1];
    int na = 0;
    for (int r = first; r < last; ++r) {
      const int i = pt.lower_rows[r];
      std::map<int, std::vector<int>> by_target; // target slot -> (L, U) pairs, ascending k
      for (int ik = pt.row[i]; ik < pt.diag[i]; ++ik) {
        a_ik.push_back(ik);
        a_kk.push_back(pt.diag[pt.col[ik]]);
        ++na;
        for (int u = pt.upd_ptr[ik]; u < pt.upd_ptr[ik + 1]; ++u) {
          auto &v = by_target[pt.upd[2 * u]];
          v.push_back(ik);
          v.push_back(pt.upd[2 * u + 1]);
        }
      }
      for (const auto &[slot, v] : by_target) {
        t_slot.push_back(slot);
        pairs.insert(pairs.end(), v.begin(), v.end());
        t_pair_ptr.push_back(int(pairs.size() / 2));
      }
    }
    a_ptr.push_back(int(a_ik.size()));
    t_ptr.push_back(int(t_slot.size()));
    widest = std::max(widest, independent ? na : last - first);
  }
  split_factor = independent;
  factor_scratch.resize(size_t(std::max(widest, 1)) * block * block);
  if (!split_factor) { return; }
  auto up = [&](DeviceArray<int> &d, const std::vector<int> &h) {
    d.resize(std::max<size_t>(h.size(), 1));
    if (!h.empty()) {
      hip_check(hipMemcpyAsync(d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, p.stream), "upload");
    }
  };
  up(d_a_ik, a_ik);
  up(d_a_kk, a_kk);
  up(d_t_slot, t_slot);
  up(d_t_pair_ptr, t_pair_ptr);
  up(d_pairs, pairs);
}

// Programs of the row-resident factor (small blocks), one per row in the
// order of the dependency levels: [length, first slot, blocks in the row,
// diagonal position, dependencies, dependency slots..., then per lower block
// ik (ascending k): dependency of inv(A_kk), update count, (target position
// in the row, dependency of A_kj) per update pair]. Rows whose blocks and
// dependencies do not fit in LDS leave the factor on the LDS-staged kernel.
void LinearSolver::PlanResident() {
  const auto &pt = *pattern;
  const size_t bb = size_t(block) * block;
  std::vector<int> prog, off;
  size_t most = 0;
  for (int r = 0; r < pt.nb; ++r) {
    const int i = pt.lower_rows[r];
    std::vector<int> deps, ops;
    for (int ik = pt.row[i]; ik < pt.diag[i]; ++ik) {
      ops.push_back(int(deps.size()));
      deps.push_back(pt.diag[pt.col[ik]]);
      ops.push_back(pt.upd_ptr[ik + 1] - pt.upd_ptr[ik]);
      for (int u = pt.upd_ptr[ik]; u < pt.upd_ptr[ik + 1]; ++u) {
        ops.push_back(pt.upd[2 * u] - pt.row[i]);
        ops.push_back(int(deps.size()));
        deps.push_back(pt.upd[2 * u + 1]);
      }
    }
    const int nrow = pt.row[i + 1] - pt.row[i];
    const int len = 5 + int(deps.size() + ops.size());
    if (len > eg::kIluProgMax) { return; }
    off.push_back(int(prog.size()));
    prog.insert(prog.end(), {len, pt.row[i], nrow, pt.diag[i] - pt.row[i], int(deps.size())});
    prog.insert(prog.end(), deps.begin(), deps.end());
    prog.insert(prog.end(), ops.begin(), ops.end());
    most = std::max(most, (nrow + deps.size()) * bb * sizeof(double));
  }
  const size_t lds = most + eg::ilu_resident_gj_lds(block);
  if (lds + sizeof(int) * eg::kIluProgMax > 163840) { return; }
  prog.resize(prog.size() + eg::kIluProgMax, 0);
  d_prog.resize(prog.size());
  d_prog_off.resize(off.size());
  hip_check(hipMemcpyAsync(d_prog, prog.data(), sizeof(int) * prog.size(), hipMemcpyHostToDevice, p.stream), "upload");
  hip_check(hipMemcpyAsync(d_prog_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, p.stream),
            "upload");
  resident_lds = lds;
  resident = true;
}

// Fill and factor the preconditioner at the current Newton state
// (src/linear.cpp:182-203; src/partial.cpp:140-173,265-298;
// src/operators.cpp:48-74,269-303; src/block_ilu.cpp:73-86).
void LinearSolver::Setup(bool check_flags) {
  const int b = block;
  if (form == 0) { require(state != nullptr, "DG solver has no refreshed state"); }
  if (assembled) {
    // src/linear.cpp:184-187: the assembled operator at the current state
    const auto &op = *op_pattern;
    if (form == 0) {
      eg::launch_dg_blocks<double>(p.geo, p.frozen, state, op_blocks, op_dslot, op_blocks, op_slot, p.stream);
    } else {
      eg::launch_hdg_fill(p.geo, hybrid->panels, hybrid->D, op_contrib, op.d_row, op.d_diag, op_blocks, p.stream);
    }
    if (mixed) { eg::launch_convert<double, float>(op_blocks, op_blocks32, op_blocks.n, p.stream); }
  }
  if (direct) {
    // src/linear.cpp:204-206: the dense matrix and its LU
    const auto &op = *op_pattern;
    eg::launch_bsr_to_dense(op_blocks, b, op.nb, op.d_row, op.d_col, dense, p.stream);
    eg::launch_dense_lu(dense, N, dense_used, dense_piv, dense_mult, p.flags, p.stream);
    if (check_flags) { p.CheckFlags("Singular dense LU factor"); }
    else { pending_check = "Singular dense LU factor"; }
    return;
  }
  if (form == 0) {
    if (factor32) {
      // the single-precision factor's blocks, stored in FP32 directly
      eg::launch_dg_blocks<float>(p.geo, p.frozen, state, blocks32, dslot, blocks32, slot, p.stream);
    } else if (ilu) {
      eg::launch_dg_blocks<double>(p.geo, p.frozen, state, blocks, dslot, blocks, slot, p.stream);
    } else {
      eg::launch_dg_blocks<double>(p.geo, p.frozen, state, blocks, nullptr, nullptr, nullptr, p.stream);
    }
  } else {
    if (ilu) {
      eg::launch_hdg_fill(p.geo, hybrid->panels, hybrid->D, contrib, pattern->d_row, pattern->d_diag, blocks,
                          p.stream, pattern->perm.empty() ? nullptr : pattern->d_inv.p);
    } else {
      eg::launch_hdg_diag_blocks(p.geo, hybrid->panels, hybrid->D, blocks, p.stream);
    }
  }
  if (factor32) {
    // mixed precision: the FP32 factor of the FP32 blocks; the FP64 blocks
    // and their factor wait until an FP64 apply asks for them
    ReplayLevels(factor_exec32, &LinearSolver::FactorLevels32);
    if (scaled_upper) {
      eg::launch_ilu_scale_upper32(blocks32, b, d_up_slot, d_up_diag, static_cast<int>(d_up_slot.n), p.stream);
    }
    factor64_pending = true;
  } else {
    Factor64();
    if (mixed) { eg::launch_convert<double, float>(blocks, blocks32, blocks.n, p.stream); }
  }
  if (ilu && dataflow) {
    eg::launch_arm<double>(sweep, N, p.stream);
    if (mixed) { eg::launch_arm<float>(sweep32, N, p.stream); }
    if (slices > 0) {
      eg::launch_arm<double>(sweep_alt, N, p.stream);
      if (mixed) { eg::launch_arm<float>(sweep32_alt, N, p.stream); }
      parity[0] = parity[1] = 0;
    }
  }
  const char *context = ilu ? "Singular BSR ILU diagonal" : "Singular block Jacobi factor";
  if (check_flags) { p.CheckFlags(context); }
  else { pending_check = context; }
}
void LinearSolver::Factor64() {
  const int b = block;
  if (ilu) {
    const auto &pt = *pattern;
    bool df = dataflow_factor && eg::launch_ilu_factor_df(blocks, b, pt.nb, pt.d_row, pt.d_diag, pt.d_col, pt.d_upd_ptr,
                                                   pt.d_upd, done_factor, ++epoch_factor, pt.d_lower_rows,
                                                   done_factor.p + pt.nb, p.flags, df_grid, p.stream);
    // row-resident programs: one launch for all levels (the rows wait on
    // each other's completion flags) when the sweeps are single-launch too
    if (!df && resident && dataflow && resident_df) {
      eg::launch_arm<double>(factor_pub, factor_pub.n, p.stream);
      df = eg::launch_ilu_factor_resident_df(blocks, factor_pub, b, d_prog_off, d_prog, pt.nb, done_factor.p + pt.nb,
                                             resident_lds, p.flags, df_grid, p.stream);
    }
    if (!df) { ReplayLevels(factor_exec, &LinearSolver::FactorLevels); }
  } else {
    eg::launch_block_inverse(blocks, b, N / b, nullptr, p.flags, p.stream);
  }
  if (scaled_upper) {
    eg::launch_ilu_scale_upper(blocks, b, d_up_slot, d_up_diag, static_cast<int>(d_up_slot.n), p.stream);
  }
}
// The per-level launches are recorded once into a graph and replayed (their
// arguments are fixed for the solver's lifetime), which removes most of the
// launch gaps between the levels.
void LinearSolver::ReplayLevels(hipGraphExec_t &exec, void (LinearSolver::*levels)()) {
  if (exec == nullptr && use_graph) {
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(p.stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
      (this->*levels)();
      ok = hipStreamEndCapture(p.stream, &graph) == hipSuccess && graph != nullptr;
    }
    ok = ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) { (void)hipGraphDestroy(graph); }
    if (!ok) {
      (void)hipGetLastError();
      exec = nullptr;
      use_graph = false;
    }
  }
  if (exec) { hip_check(hipGraphLaunch(exec, p.stream), "hipGraphLaunch"); }
  else { (this->*levels)(); }
}
// The FP32 factor of the mixed form, level by level (src/block_ilu.cpp:73-86
// in single precision; the same products as the FP64 slab kernels).
void LinearSolver::FactorLevels32() {
  const auto &pt = *pattern;
  for (size_t l = 0; l + 1 < pt.lower_level_ptr.size(); ++l) {
    const int first = pt.lower_level_ptr[l], count = pt.lower_level_ptr[l + 1] - first;
    eg::launch_ilu_slab32(blocks32, block, pt.d_lower_rows.p + first, count, pt.d_row, pt.d_diag, pt.d_col,
                          pt.d_upd_ptr, pt.d_upd, p.stream);
    eg::launch_block_inverse32(blocks32, block, count, pt.d_lower_diag.p + first, p.flags, p.stream);
  }
}

// ILU(0) factor, one launch (or two) per dependency level (src/block_ilu.cpp:73-86).
void LinearSolver::FactorLevels() {
  const int b = block;
  const auto &pt = *pattern;
  {
    {
      for (size_t l = 0; l + 1 < pt.lower_level_ptr.size(); ++l) {
        const int first = pt.lower_level_ptr[l], count = pt.lower_level_ptr[l + 1] - first;
        if (resident) {
          eg::launch_ilu_factor_resident(blocks, b, d_prog_off.p + first, d_prog, count, resident_lds, p.flags,
                                         p.stream);
          continue;
        }
        if (eg::ilu_slab_ok(b)) {
          eg::launch_ilu_slab(blocks, b, pt.d_lower_rows.p + first, count, pt.d_row, pt.d_diag, pt.d_col,
                              pt.d_upd_ptr, pt.d_upd, star_rows, p.stream);
          eg::launch_block_inverse(blocks, b, count, pt.d_lower_diag.p + first, p.flags, p.stream);
          continue;
        }
        if (split_factor) {
          const int a0 = a_ptr[l], t0 = t_ptr[l];
          eg::launch_ilu_factor_level_split(blocks, b, a_ptr[l + 1] - a0, d_a_ik.p + a0, d_a_kk.p + a0,
                                            t_ptr[l + 1] - t0, d_t_slot.p + t0, d_t_pair_ptr.p + t0, d_pairs.p,
                                            pt.d_lower_diag.p + first, count, factor_scratch.p, p.flags, p.stream);
          continue;
        }
        eg::launch_ilu_factor_level(blocks, b, pt.d_lower_rows.p + first, pt.d_lower_diag.p + first, count,
                                    pt.d_row, pt.d_diag, pt.d_col, pt.d_upd_ptr, pt.d_upd, factor_scratch.p, p.flags,
                                    p.stream);
      }
    }
  }
}
LinearSolver::~LinearSolver() {
  if (factor_exec) { (void)hipGraphExecDestroy(factor_exec); }
  if (factor_exec32) { (void)hipGraphExecDestroy(factor_exec32); }
  for (hipEvent_t e : step_done) {
    if (e) { (void)hipEventDestroy(e); }
  }
  for (hipEvent_t e : setup_mark) {
    if (e) { (void)hipEventDestroy(e); }
  }
  if (batch_ready) { (void)hipEventDestroy(batch_ready); }
  if (mirror) { batcher->ReleaseMirror(mirror); }
  if (joined) { batcher->Depart(); }
}
bool LinearSolver::Batchable() const {
  return batcher && mirror && form == 1 && restart <= eg::kCycleMax && ilu && dataflow && pattern &&
         pattern->perm.empty() && !assembled && pattern->max_deps <= 8 && block >= 4 && block <= 36 &&
         block % 4 == 0;
}
template <> const double *LinearSolver::Factors<double>() const {
  require(!factor64_pending, "FP64 factor of a mixed-precision solver not computed");
  return blocks.p;
}
template <> const float *LinearSolver::Factors<float>() const { return blocks32.p; }
template <> double *LinearSolver::SweepBuffer<double>() { return sweep.p; }
template <> float *LinearSolver::SweepBuffer<float>() { return sweep32.p; }

void LinearSolver::Apply(const double *xv, double *yv) {
  if (assembled) {
    const auto &op = *op_pattern;
    eg::launch_bsr_mult<double>(op_blocks, block, op.nb, op.d_row, op.d_col, xv, yv, p.stream);
  } else if (form == 0) {
    eg::launch_dg_apply<double>(p.geo, p.frozen, state, xv, yv, p.stream);
  } else {
    hybrid->SchurMult(xv, yv);
  }
}
void LinearSolver::Apply(const float *xv, float *yv) {
  if (assembled) {
    const auto &op = *op_pattern;
    eg::launch_bsr_mult<float>(op_blocks32, block, op.nb, op.d_row, op.d_col, xv, yv, p.stream);
  } else if (form == 0) {
    eg::launch_dg_apply<float>(p.geo, p.frozen, state, xv, yv, p.stream);
  } else {
    hybrid->SchurMult(xv, yv);
  }
}
template <class T> static void sweeps(LinearSolver &s, const T *bsr, const T *b, T *y);
// src/block_ilu.cpp:87-109: a reordered factor takes its input in its own
// row order (b(perm(i))) and returns y(perm(i)).
template <class T> static void precondition(LinearSolver &s, const T *bsr, const T *b, T *y) {
  if (s.ilu && !s.pattern->perm.empty()) {
    T *in = sizeof(T) == 8 ? reinterpret_cast<T *>(s.perm_in.p) : reinterpret_cast<T *>(s.perm_in32.p);
    T *out = sizeof(T) == 8 ? reinterpret_cast<T *>(s.perm_out.p) : reinterpret_cast<T *>(s.perm_out32.p);
    eg::launch_block_permute<T>(b, in, s.pattern->d_perm, s.block, s.pattern->nb, false, s.p.stream);
    sweeps<T>(s, bsr, in, out);
    eg::launch_block_permute<T>(out, y, s.pattern->d_perm, s.block, s.pattern->nb, true, s.p.stream);
    return;
  }
  sweeps<T>(s, bsr, b, y);
}
template <class T> static void sweeps(LinearSolver &s, const T *bsr, const T *b, T *y) {
  if (!s.ilu) {
    eg::launch_block_gemv<T>(bsr, s.block, s.N / s.block, b, y, s.p.stream);
    return;
  }
  const auto &pt = *s.pattern;
  if (s.dataflow) {
    require(b != y, "ILU apply needs distinct input and output vectors");
    T *mid = sizeof(T) == 8 ? reinterpret_cast<T *>(s.sweep.p) : reinterpret_cast<T *>(s.sweep32.p);
    if (s.slices > 0) {
      // the intermediate vectors alternate: the upper sweep arms the next one
      const int pi = sizeof(T) == 8 ? 0 : 1;
      T *alt = sizeof(T) == 8 ? reinterpret_cast<T *>(s.sweep_alt.p) : reinterpret_cast<T *>(s.sweep32_alt.p);
      T *cur = s.parity[pi] ? alt : mid, *next = s.parity[pi] ? mid : alt;
      T *cv = sizeof(T) == 8 ? reinterpret_cast<T *>(s.sweep_c.p) : reinterpret_cast<T *>(s.sweep32_c.p);
      if (!eg::launch_ilu_sweep_slice<T>(true, bsr, s.block, pt.nb, s.slices, pt.d_row, pt.d_diag, pt.d_col, b, cur,
                                         y, s.N, s.p.flags, pt.d_lower_rows, s.sweep_ticket, s.df_grid,
                                         s.p.stream)) {
        throw std::runtime_error("Launch of a row-sliced ILU sweep failed");
      }
      // c_i = inv(U_ii) w_i for every row at once (off the upper sweep's chain)
      if (s.given_c) { eg::launch_block_gemv<T>(bsr, s.block, pt.nb, cur, cv, s.p.stream, pt.d_diag); }
      if (!eg::launch_ilu_sweep_slice<T>(false, bsr, s.block, pt.nb, s.slices, pt.d_row, pt.d_diag, pt.d_col,
                                         s.given_c ? cv : cur, y, next, s.N, s.p.flags, pt.d_upper_rows,
                                         s.sweep_ticket, s.df_grid, s.p.stream, s.given_c)) {
        throw std::runtime_error("Launch of a row-sliced ILU sweep failed");
      }
      s.parity[pi] ^= 1;
      return;
    }
    if (eg::launch_ilu_sweep_df<T>(true, bsr, s.block, pt.nb, pt.max_deps, pt.d_row, pt.d_diag, pt.d_col,
                                   const_cast<T *>(b), mid, y, s.p.flags, false, pt.d_lower_rows, s.sweep_ticket, s.df_grid,
                                   s.p.stream)) {
      if (!eg::launch_ilu_sweep_df<T>(false, bsr, s.block, pt.nb, pt.max_deps, pt.d_row, pt.d_diag, pt.d_col, mid,
                                      y, y, s.p.flags, s.scaled_upper, pt.d_upper_rows, s.sweep_ticket, s.df_grid,
                                      s.p.stream)) {
        throw std::runtime_error("Launch of the upper ILU sweep failed");
      }
      return;
    }
    s.dataflow = false; // no single-launch sweep for this block size: per-level launches from now on
  }
  for (size_t l = 0; l + 1 < pt.lower_level_ptr.size(); ++l) {
    const int first = pt.lower_level_ptr[l], count = pt.lower_level_ptr[l + 1] - first;
    eg::launch_ilu_lower_level<T>(bsr, s.block, pt.d_lower_rows.p + first, count, pt.d_row, pt.d_diag, pt.d_col, b,
                                  y, s.p.opt.sweep_threads, s.p.stream);
  }
  for (size_t l = 0; l + 1 < pt.upper_level_ptr.size(); ++l) {
    const int first = pt.upper_level_ptr[l], count = pt.upper_level_ptr[l + 1] - first;
    eg::launch_ilu_upper_level<T>(bsr, s.block, pt.d_upper_rows.p + first, count, pt.d_row, pt.d_diag, pt.d_col, y,
                                  s.p.opt.sweep_threads, s.scaled_upper, s.p.stream);
  }
}
void LinearSolver::Precondition(const double *b, double *y) {
  if (factor64_pending) {
    // the FP64 blocks and factor of a mixed-precision solver, on their first
    // FP64 use after a set-up (an FP64 fallback solve or an FP64 apply); the
    // refreshed state is still the set-up's
    factor64_pending = false;
    eg::launch_dg_blocks<double>(p.geo, p.frozen, state, blocks, dslot, blocks, slot, p.stream);
    Factor64();
    p.CheckFlags("Singular BSR ILU diagonal");
  }
  precondition<double>(*this, blocks, b, y);
}
void LinearSolver::Precondition(const float *b, float *y) { precondition<float>(*this, blocks32, b, y); }

double LinearSolver::Norm(const double *v) {
  const double *metric = form == 0 ? p.d_cell_metric.p : p.d_trace_metric.p;
  eg::launch_dot_partials<double>(v, v, metric, N, p.partial, p.stream);
  return std::sqrt(p.ReduceScalar(p.partial, eg::reduce_blocks(N), 0));
}
// A Krylov failure (breakdown, nonfinite residual) is first checked against
// the device error flags, so that its cause is named: a dataflow dependency
// wait that exceeded its bound (its output is the NaN sentinel) or a zero
// pivot of the preconditioner, rather than the Krylov symptom.
void LinearSolver::Fail(const char *what) {
  p.CheckFlags(ilu ? "Singular BSR ILU diagonal" : "Singular block Jacobi factor");
  throw std::runtime_error(what);
}

// Arnoldi step j of GMRES (src/linear.cpp:100-131): z = M^-1 v_j, w = A z,
// CGS2 against v_0..v_j (two fused projection passes), the norm and
// v_{j+1} = w / |w|. The Hessenberg column (both passes and |w|^2) is written
// by the last kernel straight into pinned slot j % 2; step_done[j % 2] marks
// its completion.
template <class T>
void LinearSolver::EnqueueArnoldi(T *Vb, T *zb, T *wb, const T *metric, int j) {
  hipStream_t s = p.stream;
  double *scal = p.scalars.p;
  Precondition(Vb + size_t(j) * N, zb);
  Apply(zb, wb);
  eg::launch_multidot<T>(Vb, N, j + 1, wb, metric, N, partA, s);
  eg::launch_cgs_update<T>(Vb, N, j + 1, partA, wb, metric, N, scal, partB, 0, s);
  eg::launch_cgs_update<T>(Vb, N, j + 1, partB, wb, metric, N, scal + 64, partA, 1, s);
  eg::launch_normalize<T>(partA, wb, Vb + size_t(j + 1) * N, N, scal + 128, scal, j + 1, p.pinned + 256 * (j & 1), s);
  hip_check(hipEventRecord(step_done[j & 1], s), "hipEventRecord");
}

// One GMRES cycle through the batcher (grouped steps with other solves): the
// rotations of the Hessenberg columns run on the device (k_group_tail) with
// the same arithmetic as the host loop of GMRES, so the host waits once per
// handed-over chunk of steps instead of once per step. Leaves the rotated
// columns in h and the rotated residual in rotated; returns the steps done.
template <class T>
int LinearSolver::DeviceCycle(T *Vb, T *zb, T *wb, const T *metric, double beta, double target, int jmax) {
  auto H = [&](int i, int j) -> double & { return h[i + size_t(j) * (restart + 1)]; };
  eg::launch_cycle_init(cycle, mirror, beta, target, jmax, p.stream);
  int count = 0;
  // (Handing a second chunk over before the first is read was measured
  // slower: the extra flushes queue on the group's stream before other
  // solves' steps.)
  for (;;) {
    // a guess of the steps left from the last contraction of |g|, 2 to 8
    int chunk = 4;
    if (count > 0) {
      const double gc = std::abs(mirror[eg::kMirrorG + count]), gp = std::abs(mirror[eg::kMirrorG + count - 1]);
      const double rate = gp > 0. ? gc / gp : 1.;
      if (rate > 0. && rate < 1. && gc > target) {
        chunk = int(std::ceil(std::log(target / gc) / std::log(rate)));
      }
    }
    chunk = std::max(2, std::min({chunk, 8, jmax - count}));
    batcher->Submit<T>(*this, Vb, zb, wb, metric, chunk, step_done[0]);
    hip_check(hipEventSynchronize(step_done[0]), "hipEventSynchronize");
    count = int(mirror[0]);
    if (mirror[2] != 0.) { Fail("GMRES Arnoldi breakdown"); }
    if (mirror[1] == 0.) { break; }
  }
  for (int j = 0; j < count; ++j) {
    for (int i = 0; i <= j; ++i) { H(i, j) = mirror[eg::kMirrorH + size_t(j) * (eg::kCycleMax + 1) + i]; }
  }
  for (int i = 0; i <= count; ++i) { rotated[i] = mirror[eg::kMirrorG + i]; }
  return count;
}

// One restarted GMRES solve in arithmetic T on the buffers of that type.
// On return x holds the solution and res = rhs - A x (its norm returned).
template <class T>
double LinearSolver::GMRES(const T *rhs, T *xv, double tol, int limit, int &iterations, bool zero_guess,
                           double *true_norm, double known_rhs_norm) {
  const bool f64 = sizeof(T) == 8;
  T *Vb = f64 ? reinterpret_cast<T *>(V.p) : reinterpret_cast<T *>(V32.p);
  T *wb = f64 ? reinterpret_cast<T *>(w.p) : reinterpret_cast<T *>(w32.p);
  T *zb = f64 ? reinterpret_cast<T *>(z.p) : reinterpret_cast<T *>(z32.p);
  T *cb = f64 ? reinterpret_cast<T *>(comb.p) : reinterpret_cast<T *>(comb32.p);
  T *kb = f64 ? reinterpret_cast<T *>(corr.p) : reinterpret_cast<T *>(corr32.p);
  T *rb = f64 ? reinterpret_cast<T *>(res.p) : reinterpret_cast<T *>(res32.p);
  T *ab = wb; // the action buffer doubles as w outside the Arnoldi step
  const T *metric = f64 ? reinterpret_cast<const T *>(form == 0 ? p.d_cell_metric.p : p.d_trace_metric.p)
                        : reinterpret_cast<const T *>(form == 0 ? p.d_cell_metric32.p : p.d_trace_metric32.p);
  hipStream_t s = p.stream;
  auto norm = [&](const T *v) {
    eg::launch_dot_partials<T>(v, v, metric, N, p.partial, s);
    return std::sqrt(p.ReduceScalar(p.partial, eg::reduce_blocks(N), 0));
  };
  auto H = [&](int i, int j) -> double & { return h[i + size_t(j) * (restart + 1)]; };
  // the batcher's count of solves inside a Krylov iteration
  struct Inside {
    KrylovBatcher *b;
    explicit Inside(KrylovBatcher *kb) : b(kb) { if (b) { b->Enter(); } }
    ~Inside() { if (b) { b->Leave(); } }
  } inside(Batchable() ? batcher : nullptr);
  if (zero_guess) {
    hip_check(hipMemsetAsync(xv, 0, sizeof(T) * N, s), "memset");
    hip_check(hipMemcpyAsync(rb, rhs, sizeof(T) * N, hipMemcpyDeviceToDevice, s), "copy");
  } else {
    Apply(xv, ab);
    hip_check(hipMemcpyAsync(rb, rhs, sizeof(T) * N, hipMemcpyDeviceToDevice, s), "copy");
    eg::launch_axpby<T>(rb, -1., ab, 1., N, s);
  }
  const double rhs_norm = known_rhs_norm >= 0. ? known_rhs_norm : norm(rhs);
  iterations = 0;
  if (rhs_norm == 0.) {
    if (true_norm) { *true_norm = 0.; }
    return 0.;
  }
  double residual_norm = zero_guess ? rhs_norm : norm(rb);
  while (iterations < limit) {
    const double beta = residual_norm;
    if (beta <= tol * rhs_norm) { break; }
    eg::launch_axpby<T>(Vb, 1 / beta, rb, 0., N, s);
    std::fill(h.begin(), h.end(), 0.);
    std::fill(rotated.begin(), rotated.end(), 0.);
    rotated[0] = beta;
    int count = 0;
    bool queued = false; // step j already enqueued (ahead of the host)
    if (Batchable()) {
      // the cycle's steps on the device, handed to the batcher in chunks
      count = DeviceCycle<T>(Vb, zb, wb, metric, beta, tol * rhs_norm, std::min(restart, limit - iterations));
      iterations += count;
    }
    for (int j = 0; j < restart && iterations < limit && !Batchable(); ++j) {
      if (!queued) { EnqueueArnoldi(Vb, zb, wb, metric, j); }
      // Enqueue step j + 1 before waiting for step j's column when the
      // residual history says step j will not end the cycle: the device then
      // runs on while the host applies the rotations. Step j + 1 depends only
      // on v_{j+1}, which step j computes on the device; if step j does end
      // the cycle, its results are never read.
      queued = false;
      if (j + 1 < restart && iterations + 1 < limit) {
        bool ahead = true;
        if (j > 0) {
          const double gj = std::abs(rotated[j]), gp = std::abs(rotated[j - 1]);
          const double rho = gp > 0. ? std::min(1., gj / gp) : 1.;
          ahead = gj * rho > 3. * tol * rhs_norm;
        }
        if (ahead) {
          EnqueueArnoldi(Vb, zb, wb, metric, j + 1);
          queued = true;
        }
      }
      hip_check(hipEventSynchronize(step_done[j & 1]), "hipEventSynchronize");
      const double *col = p.pinned + 256 * (j & 1);
      for (int i = 0; i <= j; ++i) { H(i, j) = (0. + col[i]) + col[64 + i]; }
      const double next = std::sqrt(col[128]);
      H(j + 1, j) = next;
      for (int i = 0; i < j; ++i) {
        const double u = H(i, j), vr = H(i + 1, j);
        H(i, j) = cosine[i] * u + sine[i] * vr;
        H(i + 1, j) = -sine[i] * u + cosine[i] * vr;
      }
      const double d = std::hypot(H(j, j), H(j + 1, j));
      if (!(std::isfinite(d) && d > 0)) { Fail("GMRES Arnoldi breakdown"); }
      cosine[j] = H(j, j) / d;
      sine[j] = H(j + 1, j) / d;
      H(j, j) = d;
      H(j + 1, j) = 0.;
      rotated[j + 1] = -sine[j] * rotated[j];
      rotated[j] *= cosine[j];
      ++iterations;
      count = j + 1;
      if (std::abs(rotated[j + 1]) <= tol * rhs_norm || next <= std::numeric_limits<double>::min()) { break; }
    }
    eg::Coef coef{};
    for (int i = count - 1; i >= 0; --i) {
      double value = rotated[i];
      for (int j = i + 1; j < count; ++j) { value -= H(i, j) * coefficients[j]; }
      coefficients[i] = value / H(i, i);
      coef.c[i] = coefficients[i];
    }
    eg::launch_lincomb<T>(Vb, N, count, coef, cb, N, s);
    Precondition(cb, kb);
    eg::launch_axpby<T>(xv, 1., kb, 1., N, s);
    Apply(xv, ab);
    hip_check(hipMemcpyAsync(rb, rhs, sizeof(T) * N, hipMemcpyDeviceToDevice, s), "copy");
    eg::launch_axpby<T>(rb, -1., ab, 1., N, s);
    residual_norm = norm(rb);
    require(!(*p.pinned_flags & eg::kErrSpin), "ILU dataflow dependency wait exceeded its bound");
    if (!std::isfinite(residual_norm)) { Fail("Nonfinite GMRES residual"); }
  }
  if (true_norm) { *true_norm = residual_norm; }
  return residual_norm / rhs_norm;
}

// FP64 refinement around the FP32 inner GMRES (precision "mixed").
double LinearSolver::MixedSolve(const double *rhs, double tol, int limit, int &iterations, int &outer,
                                bool zero_guess, double rhs_norm) {
  hipStream_t s = p.stream;
  if (zero_guess) {
    hip_check(hipMemsetAsync(x, 0, sizeof(double) * N, s), "memset");
    hip_check(hipMemcpyAsync(res, rhs, sizeof(double) * N, hipMemcpyDeviceToDevice, s), "copy");
  } else {
    Apply(x.p, act.p);
    eg::launch_sub(rhs, act, res, N, s);
  }
  iterations = 0;
  outer = 0;
  if (rhs_norm == 0.) { return 0.; }
  double residual_norm = zero_guess ? rhs_norm : Norm(res);
  while (residual_norm > tol * rhs_norm && iterations < limit) {
    const double before = residual_norm;
    // The correction equation is solved for res / ||res||, so that residuals
    // near the rounding floor do not underflow in single precision.
    eg::launch_scale_copy(res, 1. / residual_norm, corr, N, s);
    eg::launch_convert<double, float>(corr, b32, N, s);
    int inner = 0;
    double inner_norm = 0.;
    const double inner_tol = std::max(p.opt.inner_tol, 0.5 * tol * rhs_norm / residual_norm);
    // A few restart cycles per refinement step: an inner solve that needs
    // more has stalled in single precision (caught below).
    const int inner_limit = std::min(limit - iterations, 4 * restart);
    GMRES<float>(b32.p, x32.p, inner_tol, inner_limit, inner, true, &inner_norm);
    iterations += inner;
    ++outer;
    eg::launch_convert<float, double>(x32, corr, N, s);
    eg::launch_axpby<double>(x, before, corr, 1., N, s);
    Apply(x.p, act.p);
    eg::launch_sub(rhs, act, res, N, s);
    residual_norm = Norm(res);
    if (!std::isfinite(residual_norm)) { Fail("Nonfinite GMRES residual"); }
    if (inner == 0) { break; }
    if (residual_norm > 0.5 * before && residual_norm > tol * rhs_norm) {
      // The FP32 inner solves stall (the system is too ill-conditioned for
      // single precision): finish in FP64 from the current iterate.
      int rest = 0;
      double true_norm = 0.;
      GMRES<double>(rhs, x.p, tol, limit - iterations, rest, false, &true_norm, rhs_norm);
      iterations += rest;
      ++outer;
      return true_norm;
    }
  }
  return residual_norm;
}

// src/linear.cpp:167-249 (gmres, matrix-free). tol is the relative GMRES
// tolerance (o.linear_tol, tightened by the caller for HDG refinement).
void LinearSolver::Solve(const double *dev_rhs, LinearResult &out, double tol, bool refresh) {
  out = LinearResult{};
  const double start = clock_s();
  // The set-up is not waited for on its own: its device time is taken from
  // events and the solve time is the rest of the host time.
  if (refresh) {
    hip_check(hipEventRecord(setup_mark[0], p.stream), "hipEventRecord");
    Setup(false);
    hip_check(hipEventRecord(setup_mark[1], p.stream), "hipEventRecord");
  }
  int its = 0, outer = 0;
  double true_norm = 0.;
  const double rhs_norm = Norm(dev_rhs); // its readback also brings the factor's flags
  if (refresh) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, setup_mark[0], setup_mark[1]), "hipEventElapsedTime");
    out.setup = 1e-3 * ms;
  }
  if (pending_check) {
    const char *context = pending_check;
    pending_check = nullptr;
    p.CheckPinnedFlags(context);
  }
  if (rhs_norm == 0.) {
    hip_check(hipMemsetAsync(x, 0, sizeof(double) * N, p.stream), "memset");
    hip_check(hipMemsetAsync(res, 0, sizeof(double) * N, p.stream), "memset");
    out.solve = std::max(0., clock_s() - start - out.setup);
    return;
  }
  if (direct) {
    // src/linear.cpp:224-240: x from the LU, then the residual of the
    // assembled operator
    eg::launch_dense_solve(dense, N, dense_piv, dev_rhs, x, dense_work, dense_step, p.stream);
    Apply(x.p, act.p);
    eg::launch_sub(dev_rhs, act, res, N, p.stream);
    true_norm = Norm(res);
    out.solve = std::max(0., clock_s() - start - out.setup);
    out.residual_norm = true_norm;
    out.residual = true_norm / rhs_norm;
    require(std::isfinite(out.residual) && out.residual <= std::max(10 * tol, 1e-12),
            "True inverse-mass weighted linear residual too large");
    return;
  }
  if (Batchable()) {
    if (!joined) {
      batcher->Join();
      joined = true;
    }
    batcher->Align();
  }
  if (mixed) {
    true_norm = MixedSolve(dev_rhs, tol, p.opt.max_krylov, its, outer, refresh, rhs_norm);
  } else {
    GMRES<double>(dev_rhs, x.p, tol, p.opt.max_krylov, its, refresh, &true_norm, rhs_norm);
  }
  out.solve = std::max(0., clock_s() - start - out.setup);
  out.count = its;
  out.outer = outer;
  out.residual_norm = true_norm;
  out.residual = true_norm / rhs_norm;
  require(true_norm <= tol * rhs_norm, "GMRES iteration limit reached");
  require(std::isfinite(out.residual) && out.residual <= std::max(10 * tol, 1e-12),
          "True inverse-mass weighted linear residual too large");
}
template void LinearSolver::EnqueueArnoldi<double>(double *, double *, double *, const double *, int);
template void LinearSolver::EnqueueArnoldi<float>(float *, float *, float *, const float *, int);
template double LinearSolver::GMRES<double>(const double *, double *, double, int, int &, bool, double *, double);
template double LinearSolver::GMRES<float>(const float *, float *, double, int, int &, bool, double *, double);
template int LinearSolver::DeviceCycle<double>(double *, double *, double *, const double *, double, double, int);
template int LinearSolver::DeviceCycle<float>(float *, float *, float *, const float *, double, double, int);
} // namespace euler
