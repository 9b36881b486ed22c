// SSP-DIRK3 / backward-Euler time loop with the per-stage Newton iteration
// (src/stepper.cpp:28-292) over device vectors, and the reference CLI option
// parser (src/options.cpp:8-150). Only scalars (norms, admissibility, GMRES
// projections) cross to the host inside a Newton update.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>

#include "euler.hpp"

namespace euler {
namespace {
struct Vec {
  DeviceArray<double> d;
  explicit Vec(size_t n) : d(n) {}
  double *p() { return d.p; }
};
void download(const Problem &p, const double *dev, std::vector<double> &host) {
  host.resize(p.nu);
  hip_check(hipMemcpyAsync(host.data(), dev, sizeof(double) * p.nu, hipMemcpyDeviceToHost, p.stream), "download");
  hip_check(hipStreamSynchronize(p.stream), "sync");
}
} // namespace

RunResult Run(Problem &p, const std::string &form, const std::string &prefix, bool keep_increments,
              bool keep_state) {
  const auto &o = p.opt;
  RunResult out;
  const int nu = p.nu;
  hipStream_t s = p.stream;
  // Storage allocation happens before the time loop (src/stepper.cpp:45-65).
  Vec u(nu), residual(nu), rhs(nu), er(nu), du(nu), trial(nu), trial_residual(nu), step_state(nu),
      stage_base(nu), neg(nu), check_y(nu);
  Vec slope0(nu), slope1(nu), slope2(nu), slope3(nu);
  double *slope[4] = {slope0.p(), slope1.p(), slope2.p(), slope3.p()};
  // the initial projection and its integrals on the device (src/stepper.cpp:31-36)
  p.InitialDevice(u.p());
  require(p.Admissible(u.p()), "Initial projection not positive; refine mesh");
  const auto initial = p.IntegralDevice(u.p());
  std::ofstream log;
  if (!o.timing_only && !prefix.empty()) {
    log.open(prefix + "_iterations.csv");
    require(bool(log), "Cannot open iteration output");
    log << "form,matrix_free,step,stage,newton,time,stage_time,dt,cfl,"
           "residual_physical_rms,damping,"
           "krylov_iterations,linear_relative_residual,dg_linear_relative_residual,"
           "residual_s,validation_s,jacobian_s,condensation_s,assembly_s,setup_s,"
           "solve_s,recovery_s,line_search_s,iteration_s,update_check_relative\n"
        << std::setprecision(17);
  }
  const bool hdg = form == "hdg";
  std::unique_ptr<Hybrid> hybrid;
  if (hdg) { hybrid = std::make_unique<Hybrid>(p, o.precision == "mixed"); }
  LinearSolver solver(p, hdg ? 1 : 0, hybrid.get());
  std::unique_ptr<LinearSolver> check_solver;
  if (o.check_update && hdg) { check_solver = std::make_unique<LinearSolver>(p, 0, nullptr); }
  double time = 0, begin = clock_s(), dt_sum = 0., max_cfl = 0.;
  int step = 0;
  // SSP-DIRK3(4,3,3) tableau, src/stepper.cpp:71-80
  constexpr double alpha = 0.24169426078821;
  constexpr double beta = alpha / 4.;
  constexpr double eta = 0.12915286960590;
  const double ssp_a[4][4] = {
      {alpha, 0., 0., 0.}, {-alpha, alpha, 0., 0.}, {0., 1. - alpha, alpha, 0.}, {beta, eta, .5 - beta - eta - alpha, alpha}};
  const double ssp_b[4] = {0., 1. / 6., 1. / 6., 2. / 3.};
  const double ssp_c[4] = {alpha, 0., 1., .5};
  bool final_step = time >= o.end;
  while (o.steps > 0 ? step < o.steps : !final_step) {
    hip_check(hipMemcpyAsync(step_state.p(), u.p(), sizeof(double) * nu, hipMemcpyDeviceToDevice, s), "copy");
    p.Freeze(step_state.p(), o.dt, false);
    const double rate = p.WaveRate();
    double dt = o.cfl > 0 ? o.cfl / rate : o.dt;
    if (o.steps < 0) {
      final_step = time + dt >= o.end;
      if (final_step) { dt = o.end - time; }
    }
    require(std::isfinite(dt) && dt > 0 && time + dt < 4 / ((p.gas.Gamma() + 1) * pi),
            "Step leaves the smooth simple-wave interval T < 4/((gamma+1)*pi)");
    const int stage_count = o.integrator == "be" ? 1 : 4;
    for (int stage = 0; stage < stage_count; ++stage) {
      const double diagonal = o.integrator == "be" ? 1. : ssp_a[stage][stage];
      if (o.integrator != "be" && stage > 0) {
        double coef[4];
        for (int j = 0; j < stage; ++j) { coef[j] = dt * ssp_a[stage][j]; }
        eg::launch_dirk_combine(stage_base.p(), step_state.p(), slope, coef, stage, nu, s);
      } else {
        hip_check(hipMemcpyAsync(stage_base.p(), step_state.p(), sizeof(double) * nu, hipMemcpyDeviceToDevice, s),
                  "copy");
      }
      if (stage > 0) { p.Freeze(u.p(), diagonal * dt, false); }
      p.SetStage(stage_base.p(), diagonal * dt);
      const double physical_scale = 1 / (p.dt * std::sqrt(p.lx * p.ly));
      const double old_norm = p.CellNorm(stage_base.p());
      double r0 = -1., residual_norm = 0., u_norm = 0.;
      bool converged = false, residual_ready = false;
      const int limit = o.newton_steps > 0 ? o.newton_steps : o.max_newton;
      for (int it = 0; it <= limit; ++it) {
        const double istart = clock_s();
        if (!residual_ready) {
          residual_norm = p.Residual(u.p(), residual.p());
          u_norm = p.last_state_norm;
        }
        const double rms = physical_scale * residual_norm;
        if (r0 < 0) { r0 = rms; }
        if (o.newton_steps < 0 && rms <= o.tol * std::max(1., r0)) {
          converged = true;
          break;
        }
        if (it == limit) { break; }
        const double residual_time = clock_s() - istart;
        double astart = clock_s(), jt = 0, ct = 0, rt = 0, vt = 0;
        double er_norm = 0, update_check = -1., diagnostic_time = 0.;
        LinearResult lr, attempt;
        const double allowance =
            128 * std::numeric_limits<double>::epsilon() * (u_norm + old_norm + residual_norm);
        const double target = std::max(10 * o.linear_tol * residual_norm, allowance);
        if (!hdg) {
          solver.Refresh(u.p());
          jt = clock_s() - astart;
          eg::launch_scale_copy(residual.p(), -1., rhs.p(), nu, s);
          solver.Solve(rhs.p(), lr, o.linear_tol, true);
          hip_check(hipMemcpyAsync(du.p(), solver.x, sizeof(double) * nu, hipMemcpyDeviceToDevice, s), "copy");
          // The solve's true residual is measured in the cell metric already.
          er_norm = lr.residual_norm;
        } else {
          hybrid->StreamCondense(u.p(), residual.p());
          ct = clock_s() - astart;
          double tol = o.linear_tol;
          bool valid = false;
          for (int a = 0; a < 4; ++a) {
            solver.Solve(hybrid->rhs, attempt, tol, a == 0);
            lr.count += attempt.count;
            lr.outer += attempt.outer;
            lr.setup += attempt.setup;
            lr.solve += attempt.solve;
            lr.residual = attempt.residual;
            astart = clock_s();
            hybrid->Recover(solver.x, du.p());
            rt += clock_s() - astart;
            astart = clock_s();
            // eps_S = S dl - b_S = -res; the recovered norm is sign-invariant.
            er_norm = hybrid->RecoveredResidual(solver.res, er.p());
            valid = er_norm <= target;
            vt += clock_s() - astart;
            if (valid) { break; }
            tol *= .01;
          }
          require(valid, "Recovered DG residual criterion failed after Schur refinement");
        }
        const double dgrel = residual_norm > 0 ? er_norm / residual_norm : 0.;
        require(er_norm <= target, "DG Newton linear residual too large");
        if (check_solver) {
          // src/stepper.cpp:184-201 (diagnostic, excluded from the timing)
          astart = clock_s();
          check_solver->Refresh(u.p());
          check_solver->Apply(du.p(), check_y.p());
          eg::launch_axpby<double>(check_y.p(), 1., residual.p(), 1., nu, s);
          require(p.CellNorm(check_y.p()) <= std::max(4 * target, 1e-12),
                  "Independent recovered DG equation check failed");
          eg::launch_scale_copy(residual.p(), -1., neg.p(), nu, s);
          LinearResult ref;
          check_solver->Solve(neg.p(), ref, o.linear_tol, true);
          eg::launch_sub(du.p(), check_solver->x, check_y.p(), nu, s);
          const double rnorm = p.CellNorm(check_solver->x);
          const double dnorm = p.CellNorm(check_y.p());
          update_check = dnorm / std::max(rnorm, 1e-30);
          require(dnorm <= std::max(1e-8 * rnorm, 4 * allowance), "DG/HDG one-update equivalence failed");
          diagnostic_time = clock_s() - astart;
        }
        const double ls = clock_s();
        double damping = 1.;
        bool accept = false;
        double trial_norm = 0.;
        for (int back = 0; back < 30; ++back) {
          bool admissible = false;
          trial_norm = p.TrialResidual(u.p(), du.p(), damping, trial.p(), trial_residual.p(), admissible);
          if (admissible && trial_norm <= (1 - 1e-4 * damping) * residual_norm + allowance) {
            accept = true;
            break;
          }
          damping *= .5;
        }
        require(accept, "Newton line search failed");
        std::swap(u.d.p, trial.d.p);
        std::swap(residual.d.p, trial_residual.d.p);
        residual_norm = trial_norm;
        u_norm = p.last_state_norm;
        residual_ready = true;
        const double lst = clock_s() - ls;
        const double iteration_time = clock_s() - istart - diagnostic_time;
        ++out.updates;
        out.newton_seconds += iteration_time;
        out.records.push_back({step, stage, it, lr.count, rms, damping, lr.residual, dgrel, iteration_time});
        if (keep_increments) {
          std::vector<double> h;
          download(p, du.p(), h);
          out.increments.push_back(std::move(h));
        }
        if (log.is_open()) {
          const double stage_time = time + (o.integrator == "be" ? 1. : ssp_c[stage]) * dt;
          log << form << ',' << o.matrix_free << ',' << step << ',' << stage << ',' << it << ',' << time + dt << ','
              << stage_time << ',' << dt << ',' << rate * dt << ',' << rms << ',' << damping << ',' << lr.count << ','
              << lr.residual << ',' << dgrel << ',' << residual_time << ',' << vt << ',' << jt << ',' << ct << ','
              << lr.assembly << ',' << lr.setup << ',' << lr.solve << ',' << rt << ',' << lst << ',' << iteration_time
              << ',' << update_check << '\n';
        }
      }
      require(converged || o.newton_steps > 0, "Newton failed to converge");
      eg::launch_slope(u.p(), stage_base.p(), 1. / p.dt, slope[stage], nu, s);
    }
    if (o.integrator != "be") {
      double coef[4];
      for (int j = 0; j < 4; ++j) { coef[j] = dt * ssp_b[j]; }
      eg::launch_dirk_combine(u.p(), step_state.p(), slope, coef, 4, nu, s);
      require(p.Admissible(u.p()), "SSP-DIRK3 step produced an inadmissible state");
    }
    time = o.steps < 0 && final_step ? o.end : time + dt;
    dt_sum += dt;
    max_cfl = std::max(max_cfl, rate * dt);
    ++step;
  }
  p.Sync();
  out.seconds = clock_s() - begin;
  out.final_time = time;
  out.time_steps = step;
  // errors and conservation on the device; the state is copied to the host
  // only when it is returned or saved
  out.errors = p.ErrorsDevice(u.p(), time);
  const auto final_integral = p.IntegralDevice(u.p());
  for (int c = 0; c < 4; ++c) { out.drift = std::max(out.drift, std::abs(final_integral[c] - initial[c])); }
  const bool save = o.save_solution && !prefix.empty();
  if (keep_state || save) { download(p, u.p(), out.u); }
  if (save) { p.Save(out.u, prefix); }
  if (!prefix.empty()) {
    std::ofstream summary(prefix + "_summary.csv");
    require(bool(summary), "Cannot open summary output");
    summary << "form,nx,ny,k,gamma,integrator,stages,T,dt,cfl,actual_max_cfl,mean_dt,"
               "volume_dofs,trace_dofs,matrix_free,assembly,solver,preconditioner,"
               "ilu_ordering,requested_steps,fixed_newton_steps,time_steps,"
               "newton_updates,newton_s,mean_newton_s,time_loop_s,rho_L2,rhou_L2,"
               "rhov_L2,energy_L2,max_conservation_drift,check_update\n"
            << std::setprecision(17);
    const double mean = out.updates ? out.newton_seconds / out.updates : 0.;
    summary << form << ',' << o.nx << ',' << o.ny << ',' << o.k << ',' << p.gas.Gamma() << ',' << o.integrator << ','
            << (o.integrator == "be" ? 1 : 4) << ',' << time << ',' << o.dt << ',' << o.cfl << ',' << max_cfl << ','
            << (step ? dt_sum / step : 0.) << ',' << p.nu << ',' << p.nl << ',' << o.matrix_free << ',' << o.assembly
            << ',' << o.solver << ',' << o.precond << ',' << o.ilu_ordering << ',' << o.steps << ','
            << o.newton_steps << ',' << step << ',' << out.updates << ',' << out.newton_seconds << ',' << mean << ','
            << out.seconds;
    for (double x : out.errors) { summary << ',' << x; }
    summary << ',' << out.drift << ',' << o.check_update << '\n';
    std::cout << form << " k=" << o.k << " " << o.nx << "x" << o.ny << " updates=" << out.updates
              << " mean_newton_s=" << mean << " rho_L2=" << out.errors[0] << " drift=" << out.drift << '\n';
  }
  return out;
}

// src/options.cpp:8-150, with the reference's diagnostics: --solver
// dense|direct (a dense LU on the device; the reference's direct solver is
// MKL PARDISO), --assembly element|full and --ilu-ordering mdf.
Options ParseOptions(int argc, char **argv) {
  Options o;
  bool requested_matrix_free = false, assembly_selected = false;
  for (int i = 1; i < argc; i++) {
    const std::string s = argv[i];
    auto value = [&]() {
      require(i + 1 < argc, "Missing option value");
      return std::string(argv[++i]);
    };
    if (s == "--hdg-first") { o.hdg_first = true; }
    else if (s == "--nx") { o.nx = std::stoi(value()); }
    else if (s == "--ny") { o.ny = std::stoi(value()); }
    else if (s == "--k") { o.k = std::stoi(value()); }
    else if (s == "--cfl") { o.cfl = std::stod(value()); require(o.cfl > 0, "--cfl must be positive"); }
    else if (s == "--steps") { o.steps = std::stoi(value()); require(o.steps > 0, "--steps must be positive"); }
    else if (s == "--newton-steps") {
      o.newton_steps = std::stoi(value());
      require(o.newton_steps > 0, "--newton-steps must be positive");
    }
    else if (s == "--check-update") { o.check_update = true; }
    else if (s == "--timing-only") { o.timing_only = true; }
    else if (s == "--ilu-ordering") { o.ilu_ordering = value(); }
    else if (s == "--dt") { o.dt = std::stod(value()); }
    else if (s == "--T") { o.end = std::stod(value()); }
    else if (s == "--alpha-deg") { o.alpha = std::stod(value()); }
    else if (s == "--safety") { o.safety = std::stod(value()); }
    else if (s == "--volume-order") { o.volume_order = std::stoi(value()); }
    else if (s == "--assembly") {
      o.assembly = value();
      assembly_selected = true;
    }
    else if (s == "--form") { o.form = value(); }
    else if (s == "--solver") { o.solver = value(); }
    else if (s == "--integrator") { o.integrator = value(); }
    else if (s == "--precond") { o.precond = value(); }
    else if (s == "--output") { o.output = value(); }
    else if (s == "--matrix-free") { requested_matrix_free = true; }
    else if (s == "--no-save-solution") { o.save_solution = false; }
    else if (s == "--linear-tol") { o.linear_tol = std::stod(value()); }
    else if (s == "--newton-tol") { o.tol = std::stod(value()); }
    else if (s == "--restart") { o.restart = std::stoi(value()); }
    else if (s == "--max-krylov") { o.max_krylov = std::stoi(value()); }
    else if (s == "--max-newton") { o.max_newton = std::stoi(value()); }
    else if (s == "--verify") { o.verify = true; }
    else if (s == "--precision") { o.precision = value(); }
    else if (s == "--inner-tol") { o.inner_tol = std::stod(value()); }
    else if (s == "--ilu-schedule") { o.ilu_schedule = value(); }
    else if (s == "--sweep-threads") { o.sweep_threads = std::stoi(value()); }
    else if (s == "--help") {
      std::cout << "--nx N --ny N --k 0..8 --alpha-deg A (default 45)\n"
                   "--dt DT | --cfl C (CFL wins; h=min(hx,hy), speed includes safety)\n"
                   "--T T (default .2) | --steps N (steps wins; no T clipping)\n"
                   "--max-newton N --newton-tol E | --newton-steps N (fixed updates/stage)\n"
                   "--form dg|hdg|both --integrator ssp-dirk3|be\n"
                   "--solver gmres|direct|dense (direct and dense: dense LU on the device)\n"
                   "--precond block-jacobi|block-ilu0 --ilu-ordering none|mdf\n"
                   "--assembly full|element|partial --matrix-free --safety 1.1\n"
                   "--linear-tol E --restart N --max-krylov N --output PREFIX\n"
                   "--precision fp64|mixed --inner-tol E --ilu-schedule levels|dataflow|dataflow-all --sweep-threads 0|256|1024\n"
                   "--check-update --timing-only --no-save-solution --hdg-first\n";
      std::exit(EXIT_SUCCESS);
    } else {
      throw std::runtime_error("Unknown option: " + s);
    }
  }
  require(o.nx >= 2 && o.ny >= 2 && o.k >= 0 && o.k <= 8 && o.dt > 0 && o.end >= 0 && o.safety >= 1 && o.tol > 0 &&
              o.linear_tol > 0,
          "Invalid mesh, degree, time or safety");
  const long double cells = static_cast<long double>(o.nx) * o.ny, qq = o.k + 1.L;
  require(4 * cells * qq * qq <= std::numeric_limits<int>::max() && 8 * cells * qq <= std::numeric_limits<int>::max(),
          "Mesh/degree exceeds MFEM int DOF capacity");
  require(std::isfinite(o.alpha) && std::isfinite(o.dt) && std::isfinite(o.end) && std::isfinite(o.safety) &&
              std::isfinite(o.tol) && std::isfinite(o.linear_tol),
          "Nonfinite option");
  require(o.form == "dg" || o.form == "hdg" || o.form == "both", "Invalid form");
  require(o.integrator == "ssp-dirk3" || o.integrator == "be", "Invalid time integrator");
  require(o.solver == "gmres" || o.solver == "direct" || o.solver == "dense", "Invalid solver");
  require(o.restart > 0 && o.max_krylov > 0 && o.max_newton > 0, "Iteration limits must be positive");
  require(o.alpha > 0 && o.alpha < 90, "Require 0 < alpha-deg < 90");
  require(o.ilu_ordering == "none" || o.ilu_ordering == "mdf", "Invalid ILU ordering");
  require(o.volume_order == -1 || o.volume_order == 2 * o.k + 1, "Volume rule must be (k+1)^2 GL points (order 2k+1)");
  o.volume_order = -1;
  if (o.solver == "gmres") {
    require(o.precond == "block-jacobi" || o.precond == "block-ilu0", "Invalid preconditioner");
    if (requested_matrix_free) { o.assembly = "partial"; }
    o.matrix_free = o.assembly != "full";
  } else {
    require(!requested_matrix_free, "Direct LU requires an assembled matrix");
    if (!assembly_selected) { o.assembly = "full"; }
    require(o.assembly == "full", "Direct LU requires --assembly full");
    o.matrix_free = false;
    o.precond = "none";
  }
  require(o.assembly == "full" || o.assembly == "element" || o.assembly == "partial", "Invalid assembly");
  require(o.precision == "fp64" || o.precision == "mixed", "Invalid precision");
  require(o.solver == "gmres" || o.precision == "fp64", "--precision mixed needs --solver gmres");
  require(o.inner_tol > 0 && o.inner_tol < 1, "Invalid inner tolerance");
  require(o.ilu_schedule.empty() || o.ilu_schedule == "levels" || o.ilu_schedule == "dataflow" ||
              o.ilu_schedule == "dataflow-all",
          "Invalid ILU schedule");
  require(o.sweep_threads == 0 || o.sweep_threads == 256 || o.sweep_threads == 1024, "Invalid sweep threads");
  if (o.timing_only) { o.save_solution = false; }
  return o;
}
Options ParseOptions(const std::string &args) {
  std::istringstream in(args);
  std::vector<std::string> words{"euler2d"};
  for (std::string w; in >> w;) { words.push_back(w); }
  std::vector<char *> argv;
  for (auto &w : words) { argv.push_back(w.data()); }
  return ParseOptions(int(argv.size()), argv.data());
}
} // namespace euler
