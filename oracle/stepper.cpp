// CPU ORACLE -- test infrastructure only (see oracle.hpp).
// SSP-DIRK3 / backward-Euler time loop with the per-stage Newton iteration,
// DG and condensed-HDG linearised solves, line search and CSV outputs.
#include "oracle.hpp"
#include <algorithm>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>

namespace oracle {
// src/stepper.cpp:28-292
RunResult Run(Problem &p, const std::string &form, const std::string &prefix,
              bool keep_increments) {
  const auto &o = p.opt;
  RunResult out;
  out.u = p.Initial();
  require(p.Admissible(out.u.data()), "Initial projection not positive; refine mesh");
  const auto initial = p.Integral(out.u.data());
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
  const int nu = p.nu;
  std::vector<double> residual(nu), rhs(nu), er(nu), du(nu), trial(nu), trial_residual(nu),
      step_state(nu), stage_base(nu);
  std::array<std::vector<double>, 4> slope;
  for (auto &v : slope) { v.assign(nu, 0.); }
  Hybrid hybrid;
  Condensed condensed;
  if (form == "hdg") { p.PrepareHybrid(hybrid, condensed); }
  std::unique_ptr<PartialDG> dg_jacobian;
  if (form == "dg") { dg_jacobian = std::make_unique<PartialDG>(p); }
  const int global_size = form == "dg" ? p.nu : p.nl;
  LinearSolver solver(global_size, form == "dg" ? p.m : p.t, o,
                      form == "dg" ? p.cell_metric : p.trace_metric);
  if (o.precond == "block-ilu0") { solver.Prepare(p.SolverGraph(form == "hdg"), o); }
  LinearResult lr, attempt;
  Options linear = o;
  double time = 0, begin = clock_s(), dt_sum = 0., max_cfl = 0.;
  int step = 0;
  // SSP-DIRK3(4,3,3) tableau, src/stepper.cpp:71-80
  constexpr double alpha = 0.24169426078821;
  constexpr double beta = alpha / 4.;
  constexpr double eta = 0.12915286960590;
  const double ssp_a[4][4] = {{alpha, 0., 0., 0.},
                              {-alpha, alpha, 0., 0.},
                              {0., 1. - alpha, alpha, 0.},
                              {beta, eta, .5 - beta - eta - alpha, alpha}};
  const double ssp_b[4] = {0., 1. / 6., 1. / 6., 2. / 3.};
  const double ssp_c[4] = {alpha, 0., 1., .5};
  bool final_step = time >= o.end;
  while (o.steps > 0 ? step < o.steps : !final_step) {
    step_state = out.u;
    p.Freeze(step_state, o.dt, false);
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
      stage_base = step_state;
      if (o.integrator != "be") {
        for (int j = 0; j < stage; ++j) {
          const double c = dt * ssp_a[stage][j];
          for (int i = 0; i < nu; ++i) { stage_base[i] += c * slope[j][i]; }
        }
      }
      if (stage > 0) { p.Freeze(out.u, diagonal * dt, false); }
      p.dt = diagonal * dt;
      p.old = stage_base;
      const double physical_scale = 1 / (p.dt * std::sqrt(p.lx * p.ly));
      const double old_norm = p.CellNorm(stage_base.data());
      double r0 = -1.;
      bool converged = false, residual_ready = false;
      const int limit = o.newton_steps > 0 ? o.newton_steps : o.max_newton;
      for (int it = 0; it <= limit; ++it) {
        const double istart = clock_s();
        if (!residual_ready) { p.Residual(out.u.data(), residual.data()); }
        const double residual_norm = p.CellNorm(residual.data());
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
        lr.count = 0;
        lr.residual = lr.assembly = lr.setup = lr.solve = 0.;
        const double allowance = 128 * std::numeric_limits<double>::epsilon() *
                                 (p.CellNorm(out.u.data()) + old_norm + residual_norm);
        const double target = std::max(10 * o.linear_tol * residual_norm, allowance);
        if (form == "dg") {
          dg_jacobian->Refresh(out.u);
          jt = clock_s() - astart;
          for (int i = 0; i < nu; ++i) { rhs[i] = -residual[i]; }
          solver.Solve(*dg_jacobian, rhs.data(), o, lr);
          du = lr.x;
          astart = clock_s();
          er_norm = p.CellNorm(lr.equation_residual.data());
          vt = clock_s() - astart;
        } else {
          auto &h = hybrid;
          auto &s = condensed;
          p.StreamCondense(out.u.data(), residual.data(), h, s);
          ct = clock_s() - astart;
          linear.linear_tol = o.linear_tol;
          bool valid = false;
          for (int a = 0; a < 4; ++a) {
            solver.Solve(*s.op, s.rhs.data(), linear, attempt, a == 0);
            lr.count += attempt.count;
            lr.assembly += attempt.assembly;
            lr.setup += attempt.setup;
            lr.solve += attempt.solve;
            lr.residual = attempt.residual;
            astart = clock_s();
            p.Recover(h, s, attempt.x.data(), du.data());
            rt += clock_s() - astart;
            astart = clock_s();
            p.RecoveredResidual(h, attempt.equation_residual.data(), er.data());
            er_norm = p.CellNorm(er.data());
            valid = er_norm <= target;
            vt += clock_s() - astart;
            if (valid) { break; }
            linear.linear_tol *= .01;
          }
          require(valid, "Recovered DG residual criterion failed after Schur refinement");
        }
        const double dgrel = residual_norm > 0 ? er_norm / residual_norm : 0.;
        require(er_norm <= target, "DG Newton linear residual too large");
        if (o.check_update && form == "hdg") {
          // src/stepper.cpp:184-201 (diagnostic, excluded from the timing)
          astart = clock_s();
          PartialDG dg(p);
          dg.Refresh(out.u);
          std::vector<double> independent(nu), neg(nu);
          dg.Mult(du.data(), independent.data());
          for (int i = 0; i < nu; ++i) { independent[i] += residual[i]; neg[i] = -residual[i]; }
          require(p.CellNorm(independent.data()) <= std::max(4 * target, 1e-12),
                  "Independent recovered DG equation check failed");
          LinearSolver check(p.nu, p.m, o, p.cell_metric);
          LinearResult reference;
          check.Solve(dg, neg.data(), o, reference);
          std::vector<double> difference(nu);
          for (int i = 0; i < nu; ++i) { difference[i] = du[i] - reference.x[i]; }
          const double rnorm = p.CellNorm(reference.x.data());
          update_check = p.CellNorm(difference.data()) / std::max(rnorm, 1e-30);
          require(p.CellNorm(difference.data()) <= std::max(1e-8 * rnorm, 4 * allowance),
                  "DG/HDG one-update equivalence failed");
          diagnostic_time = clock_s() - astart;
        }
        const double ls = clock_s();
        double damping = 1.;
        bool accept = false;
        for (int back = 0; back < 30; ++back) {
          for (int i = 0; i < nu; ++i) { trial[i] = out.u[i] + damping * du[i]; }
          if (p.Admissible(trial.data())) {
            p.Residual(trial.data(), trial_residual.data());
            if (p.CellNorm(trial_residual.data()) <= (1 - 1e-4 * damping) * residual_norm + allowance) {
              accept = true;
              break;
            }
          }
          damping *= .5;
        }
        require(accept, "Newton line search failed");
        out.u = trial;
        residual = trial_residual;
        residual_ready = true;
        const double lst = clock_s() - ls;
        const double iteration_time = clock_s() - istart - diagnostic_time;
        ++out.updates;
        out.newton_seconds += iteration_time;
        out.records.push_back({step, stage, it, lr.count, rms, damping, lr.residual, dgrel,
                               iteration_time});
        if (keep_increments) { out.increments.push_back(du); }
        if (log.is_open()) {
          const double stage_time = time + (o.integrator == "be" ? 1. : ssp_c[stage]) * dt;
          log << form << ',' << o.matrix_free << ',' << step << ',' << stage << ',' << it << ','
              << time + dt << ',' << stage_time << ',' << dt << ',' << rate * dt << ',' << rms
              << ',' << damping << ',' << lr.count << ',' << lr.residual << ',' << dgrel << ','
              << residual_time << ',' << vt << ',' << jt << ',' << ct << ',' << lr.assembly
              << ',' << lr.setup << ',' << lr.solve << ',' << rt << ',' << lst << ','
              << iteration_time << ',' << update_check << '\n';
        }
      }
      require(converged || o.newton_steps > 0, "Newton failed to converge");
      for (int i = 0; i < nu; ++i) { slope[stage][i] = (out.u[i] - stage_base[i]) * (1. / p.dt); }
    }
    if (o.integrator != "be") {
      out.u = step_state;
      for (int stage = 0; stage < 4; ++stage) {
        const double c = dt * ssp_b[stage];
        for (int i = 0; i < nu; ++i) { out.u[i] += c * slope[stage][i]; }
      }
      require(p.Admissible(out.u.data()), "SSP-DIRK3 step produced an inadmissible state");
    }
    time = o.steps < 0 && final_step ? o.end : time + dt;
    dt_sum += dt;
    max_cfl = std::max(max_cfl, rate * dt);
    ++step;
  }
  out.seconds = clock_s() - begin;
  out.final_time = time;
  out.time_steps = step;
  out.errors = p.Errors(out.u.data(), time);
  const auto final_integral = p.Integral(out.u.data());
  for (int c = 0; c < 4; ++c) { out.drift = std::max(out.drift, std::abs(final_integral[c] - initial[c])); }
  if (o.save_solution && !prefix.empty()) { p.Save(out.u.data(), prefix); }
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
    summary << form << ',' << o.nx << ',' << o.ny << ',' << o.k << ',' << p.gas.Gamma() << ','
            << o.integrator << ',' << (o.integrator == "be" ? 1 : 4) << ',' << time << ','
            << o.dt << ',' << o.cfl << ',' << max_cfl << ',' << (step ? dt_sum / step : 0.)
            << ',' << p.nu << ',' << p.nl << ',' << o.matrix_free << ',' << o.assembly << ','
            << o.solver << ',' << o.precond << ',' << o.ilu_ordering << ',' << o.steps << ','
            << o.newton_steps << ',' << step << ',' << out.updates << ',' << out.newton_seconds
            << ',' << mean << ',' << out.seconds;
    for (double x : out.errors) { summary << ',' << x; }
    summary << ',' << out.drift << ',' << o.check_update << '\n';
    std::cout << form << " k=" << o.k << " " << o.nx << "x" << o.ny << " updates=" << out.updates
              << " mean_newton_s=" << mean << " rho_L2=" << out.errors[0]
              << " drift=" << out.drift << '\n';
  }
  return out;
}

// src/options.cpp:8-150 (--solver direct, MKL PARDISO in the reference, is
// restated as the same dense LU as --solver dense)
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
    else { throw std::runtime_error("Unknown option: " + s); }
  }
  require(o.nx >= 2 && o.ny >= 2 && o.k >= 0 && o.k <= 8 && o.dt > 0 && o.end >= 0 &&
              o.safety >= 1 && o.tol > 0 && o.linear_tol > 0,
          "Invalid mesh, degree, time or safety");
  const long double cells = static_cast<long double>(o.nx) * o.ny, qq = o.k + 1.L;
  require(4 * cells * qq * qq <= std::numeric_limits<int>::max() &&
              8 * cells * qq <= std::numeric_limits<int>::max(),
          "Mesh/degree exceeds MFEM int DOF capacity");
  require(std::isfinite(o.alpha) && std::isfinite(o.dt) && std::isfinite(o.end) &&
              std::isfinite(o.safety) && std::isfinite(o.tol) && std::isfinite(o.linear_tol),
          "Nonfinite option");
  require(o.form == "dg" || o.form == "hdg" || o.form == "both", "Invalid form");
  require(o.integrator == "ssp-dirk3" || o.integrator == "be", "Invalid time integrator");
  require(o.solver == "gmres" || o.solver == "direct" || o.solver == "dense", "Invalid solver");
  require(o.restart > 0 && o.max_krylov > 0 && o.max_newton > 0, "Iteration limits must be positive");
  require(o.alpha > 0 && o.alpha < 90, "Require 0 < alpha-deg < 90");
  require(o.ilu_ordering == "none" || o.ilu_ordering == "mdf", "Invalid ILU ordering");
  require(o.volume_order == -1 || o.volume_order == 2 * o.k + 1,
          "Volume rule must be (k+1)^2 GL points (order 2k+1)");
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
  if (o.timing_only) { o.save_solution = false; }
  return o;
}
} // namespace oracle
