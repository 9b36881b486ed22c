// euler2d: the reference command-line program (src/main.cpp:4-54) on the GPU.
// Runs DG and/or HDG, writes <output>_comparison.csv and requires the final
// DG/HDG states to agree to 1e-7; --verify runs the self-test on the device
// kernels instead; errors print "ERROR: ..." and exit 1.
#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "euler.hpp"

int main(int argc, char **argv) {
  using namespace euler;
  try {
    Options o = ParseOptions(argc, argv);
    const EulerGas gas(3.);
    const auto parent = std::filesystem::path(o.output).parent_path();
    if (!parent.empty()) { std::filesystem::create_directories(parent); }
    if (o.verify) {
      Verify(o, gas);
      return 0;
    }
    Problem p(o, gas);
    RunResult dg, hdg;
    if (o.form == "both" && o.hdg_first) {
      hdg = Run(p, "hdg", o.output + "_hdg");
      dg = Run(p, "dg", o.output + "_dg");
    } else {
      if (o.form != "hdg") { dg = Run(p, "dg", o.output + "_dg"); }
      if (o.form != "dg") { hdg = Run(p, "hdg", o.output + "_hdg"); }
    }
    if (o.form == "both") {
      double diff = 0.;
      for (int i = 0; i < p.nu; ++i) { diff = std::max(diff, std::abs(dg.u[i] - hdg.u[i])); }
      const auto l2 = p.Differences(dg.u, hdg.u);
      double combined = 0.;
      for (double v : l2) { combined += v * v; }
      combined = std::sqrt(combined);
      std::ofstream f(o.output + "_comparison.csv");
      f << "max_nodal_difference,rho_L2_difference,rhou_L2_difference,"
           "rhov_L2_difference,energy_L2_difference,solution_L2_difference,"
           "dg_over_hdg_mean_newton\n"
        << std::setprecision(17) << diff << ',' << l2[0] << ',' << l2[1] << ',' << l2[2] << ',' << l2[3] << ','
        << combined << ','
        << (dg.updates && hdg.updates && hdg.newton_seconds > 0
                ? (dg.newton_seconds / dg.updates) / (hdg.newton_seconds / hdg.updates)
                : 0.)
        << '\n';
      std::cout << "DG/HDG max nodal difference=" << diff << '\n';
      require(diff < 1e-7, "DG/HDG final solutions disagree");
    }
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "ERROR: " << e.what() << '\n';
    return 1;
  }
}
