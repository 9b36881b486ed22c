// CPU ORACLE -- test infrastructure only.
//
// A plain serial C++17 restatement of the reference implicit SSP-DIRK3 Newton
// update (DG body form and condensed frozen-HLL HDG face form) of
// YueWu2002/mfem_euler2d_cpp, written without MFEM. It exists to check the
// HIP product path: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and never as the thing measured or shipped.
//
// Every function cites the reference file:line it restates. MFEM pieces the
// reference relies on (Gauss-Legendre tables, Lagrange basis values and
// derivatives, partially pivoted LU, RightSolve, GEMM) are re-derived from the
// mathematics; their rounding is not MFEM's, so bitwise parity with the MFEM
// build is not claimed (tolerance parity only).
#pragma once
#include <array>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace oracle {
inline constexpr double pi = 3.141592653589793238462643383279502884;
using State = std::array<double, 4>;

// include/euler.hpp:19-23
inline void require(bool ok, const char *msg) {
  if (!ok) { throw std::runtime_error(msg); }
}
double clock_s();

// Column-major dense matrix, the MFEM DenseMatrix convention
// (src/problem.cpp:18-19, src/partial.cpp:74-75).
struct Dense {
  int h = 0, w = 0;
  std::vector<double> d;
  Dense() = default;
  Dense(int r, int c) : h(r), w(c), d(size_t(r) * c, 0.) {}
  void resize(int r, int c) { h = r; w = c; d.assign(size_t(r) * c, 0.); }
  void zero() { std::fill(d.begin(), d.end(), 0.); }
  double &operator()(int i, int j) { return d[i + size_t(j) * h]; }
  double operator()(int i, int j) const { return d[i + size_t(j) * h]; }
  double *data() { return d.data(); }
  const double *data() const { return d.data(); }
};
struct Mat4 {
  double a[16] = {};
  double &operator()(int i, int j) { return a[i + 4 * j]; }
  double operator()(int i, int j) const { return a[i + 4 * j]; }
};

// include/euler.hpp:29-43, src/physics.cpp
class EulerGas {
  double gamma, gamma_minus_one;
public:
  explicit EulerGas(double value);
  double Gamma() const { return gamma; }
  double pressure(const State &u) const;
  bool positive(const State &u) const;
  State flux(const State &u, int axis) const;
  void jacobian(const State &u, int axis, Mat4 &j) const;
  std::array<double, 2> speeds(const State &l, const State &r, int axis, double safety) const;
  State exact(double x, double y, double time, double alpha) const;
};

// include/euler.hpp:46-57
struct Options {
  int nx = 100, ny = 100, k = 1, volume_order = -1, max_newton = 30;
  int restart = 20, max_krylov = 1500, newton_steps = -1, steps = -1;
  double alpha = 45, dt = .001, end = .2, cfl = -1, safety = 1.1, tol = 1e-12,
         linear_tol = 1e-12;
  std::string assembly = "partial";
  std::string form = "both", solver = "gmres", precond = "block-jacobi",
              integrator = "ssp-dirk3", output = "code_test/results/run";
  bool matrix_free = true, verify = false, save_solution = true, hdg_first = false;
  bool check_update = false, timing_only = false;
  std::string ilu_ordering = "none";
};
Options ParseOptions(int argc, char **argv);

// include/euler.hpp:62-68
struct Face {
  int l = -1, r = -1, ls = -1, rs = -1, axis = 0;
  double length = 0.;
  std::vector<double> a, b, inverse_sum, left_weight, right_weight, penalty;
};

using Basis = std::vector<std::vector<std::pair<int, double>>>;

// Lift (B) or restriction (C) coupling of one cell to its 4t trace values,
// include/euler.hpp:71-88, src/operators.cpp:7-26,75-142.
struct FaceCoupling {
  const std::array<Basis, 4> *basis = nullptr;
  const std::vector<double> *inverse_mass = nullptr;
  int nodes = 0, cell_size = 0;
  bool lift = false;
  std::vector<Mat4> flux;
  std::vector<double> scalar_flux;
  int Height() const { return lift ? cell_size : 16 * nodes; }
  int Width() const { return lift ? 16 * nodes : cell_size; }
  void Mult(const double *x, double *y) const;
  void Dense_(Dense &a) const;
  void Panel(const Dense &x, int side, int first, int count, Dense &out) const;
};

// Fixed-pattern BSR block ILU(0), natural ordering, include/euler.hpp:91-106,
// src/block_ilu.cpp:10-109.
class FixedBlockILU {
public:
  struct Update { int target, upper; };
  std::vector<std::vector<Update>> updates;
  std::vector<int> row, col, diag;
  std::vector<Dense> data;
  std::vector<std::vector<int>> pivots;
  std::vector<int> permutation, inverse; // factor row i is graph row permutation[i]
  mutable std::vector<double> work;
  int block;
  FixedBlockILU(const std::vector<std::vector<int>> &graph, int size, const std::string &ordering = "none");
  void Zero(bool off_diagonal_only = false);
  Dense &At(int r, int c);
  void Factor();
  void Mult(const double *b, double *x) const;
};

class Problem;
// Operators of the linear Newton systems, include/euler.hpp:119-136.
class Operator {
public:
  int height = 0;
  virtual ~Operator() = default;
  virtual void Mult(const double *x, double *y) const = 0;
  virtual void DiagonalBlocks(int size, std::vector<Dense> &blocks) const = 0;
  virtual std::vector<std::vector<int>> BlockGraph(int size) const = 0;
  virtual void FillBlocks(FixedBlockILU &target) const = 0;
};

// Matrix-free sum-factorised DG Jacobian, src/partial.cpp:8-302.
class PartialDG : public Operator {
public:
  const Problem &p;
  std::vector<double> state;
  std::vector<Mat4> jx, jy, jl, jr;
  mutable std::vector<State> flux_x, flux_y;
  explicit PartialDG(const Problem &problem);
  void Refresh(const std::vector<double> &u);
  void Mult(const double *x, double *y) const override;
  void DiagonalBlocks(int size, std::vector<Dense> &blocks) const override;
  std::vector<std::vector<int>> BlockGraph(int size) const override;
  void FillBlocks(FixedBlockILU &target) const override;
};

// include/euler.hpp:138-160
struct Hybrid {
  std::vector<double> D;
  std::vector<FaceCoupling> B, C;
  std::vector<std::vector<int>> ids;
};
// Schur complement with borrowed factors: y = D x - sum_e R_e^T C_e X_e R_e x,
// src/operators.cpp:147-179,269-303,48-74.
class SchurOperator : public Operator {
public:
  const Problem &p;
  const Hybrid &h;
  const std::vector<Dense> &X;
  SchurOperator(const Problem &problem, const Hybrid &hybrid, const std::vector<Dense> &x);
  void Mult(const double *x, double *y) const override;
  void DiagonalBlocks(int size, std::vector<Dense> &blocks) const override;
  std::vector<std::vector<int>> BlockGraph(int size) const override;
  void FillBlocks(FixedBlockILU &target) const override;
};
struct Condensed {
  std::vector<Dense> X;
  std::vector<std::vector<double>> y;
  std::vector<double> rhs;
  Dense cell_matrix;
  std::vector<int> cell_pivots;
  std::unique_ptr<SchurOperator> op;
};

struct VolumeTest { int node; double dx, dy; };

// include/euler.hpp:161-230, src/problem.cpp
class Problem {
public:
  Options opt;
  EulerGas gas;
  double alpha, lx, ly, hx, hy, area, dt = 0;
  int n, p, m, t, nf, nu, nl, q;
  std::vector<double> x1, w1;   // 1-D GL nodes and weights on [0,1]
  Dense D1;                     // D1(a,i) = l_i'(x_a)
  std::vector<double> end0, end1; // l_i(0), l_i(1)
  std::vector<double> w;        // tensor weights per node (lexicographic, x fastest)
  std::vector<double> inv_mass; // diagonal GL mass inverse on the reference square
  std::array<Basis, 4> face_basis;
  std::vector<std::vector<VolumeTest>> volume_tests;
  std::vector<double> cell_metric, trace_metric, origx, origy;
  std::vector<Face> faces;
  std::vector<std::array<int, 4>> cell_faces;
  std::vector<double> old;
  int freeze_generation = 0;

  Problem(const Options &o, const EulerGas &gas);
  State Node(const double *u, int e, int node) const;
  State Trace(const double *u, int e, int side, int qq) const;
  bool Admissible(const double *u) const;
  std::vector<double> Initial() const;
  void Freeze(const std::vector<double> &u, double step, bool check_admissible = true);
  double WaveRate() const;
  double CellNorm(const double *r) const;
  void Residual(const double *u, double *r) const;
  void MassSolve(Dense &a) const;
  void MassSolve(double *r) const;
  void VolumeJacobianCell(const double *u, int e, Dense &a, const std::vector<Mat4> *cx = nullptr,
                          const std::vector<Mat4> *cy = nullptr) const;
  void DenseDG(const double *u, Dense &a) const; // full nu x nu Jacobian (test aid)
  void PrepareHybrid(Hybrid &h, Condensed &s) const;
  void HybridDiagonal(Hybrid &h) const;
  void HybridCell(const double *u, int e, Hybrid &h, Dense &a) const;
  void StreamCondense(const double *u, const double *r, Hybrid &h, Condensed &s) const;
  void Recover(const Hybrid &h, const Condensed &s, const double *dl, double *du) const;
  void RecoveredResidual(const Hybrid &h, const double *schur_residual, double *out) const;
  std::vector<std::vector<int>> SolverGraph(bool hdg) const;
  std::array<double, 4> Errors(const double *u, double time) const;
  std::array<double, 4> Integral(const double *u) const;
  std::array<double, 4> Differences(const double *u, const double *v) const;
  void Save(const double *u, const std::string &prefix) const;
};

// Dense LU with partial pivoting (MFEM LUFactors semantics).
bool lu_factor(double *a, int *ipiv, int m);
// Minimum discarded fill ordering of an unweighted block graph (the
// reference's --ilu-ordering mdf, src/block_ilu.cpp:16-26).
std::vector<int> mdf_ordering(const std::vector<std::vector<int>> &graph);
void lu_solve(const double *a, const int *ipiv, int m, int nrhs, double *x);
void lu_right_solve(const double *a, const int *ipiv, int m, int rows, double *x);

// include/euler.hpp:231-249, src/linear.cpp
struct LinearResult {
  std::vector<double> x, equation_residual;
  int count = 0;
  double residual = 0, assembly = 0, setup = 0, solve = 0;
};
class LinearSolver {
public:
  int n, block, restart;
  std::vector<double> metric, residual, action, preconditioned, combination, correction,
      weighted, projection, basis, h, cosine, sine, rotated, coefficients;
  std::vector<Dense> diagonal;
  std::vector<std::vector<int>> diagonal_pivots;
  std::unique_ptr<FixedBlockILU> ilu;
  std::vector<double> dense_matrix; // --solver dense|direct: the assembled matrix and its LU
  std::vector<int> dense_pivots;
  LinearSolver(int n, int block, const Options &o, const std::vector<double> &metric);
  void Prepare(const std::vector<std::vector<int>> &graph, const Options &o);
  double Dot(const double *a, const double *b) const;
  void Precondition(const double *b, double *x) const;
  double GMRES(const Operator &a, const double *rhs, double tol, int limit, LinearResult &out,
               bool zero_guess);
  void Solve(const Operator &op, const double *rhs, const Options &o, LinearResult &out,
             bool refresh = true);
};

// include/euler.hpp:250-256, src/stepper.cpp
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
  std::vector<std::vector<double>> increments; // du per accepted update (optional)
};
RunResult Run(Problem &p, const std::string &form, const std::string &prefix,
              bool keep_increments = false);
} // namespace oracle
