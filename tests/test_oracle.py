"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed block) fails it: textbook closed forms,
finite differences, brute-force dense algebra on tiny meshes, and the
reference's own known-answer tests and thresholds
(src/verification.cpp:7-303, code_test/test_solver.py).
"""
import math

import numpy as np
import pytest


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


# --------------------------------------------------------------- physics KATs
@pytest.mark.parametrize("gamma", [1.4, 1.67, 3.0])
def test_physics_known_answers(oracle_lib, gamma):
    """src/verification.cpp:257-284."""
    u = [1.2, .3, -.2, 3.]
    p_exp = (gamma - 1.) * (3. - .5 * (.3 * .3 + .2 * .2) / 1.2)
    for axis in (0, 1):
        g = oracle_lib.gas(gamma, u, axis, 1.1, u)
        assert abs(g["pressure"] - p_exp) < 1e-14
        for c in range(4):
            up, um = list(u), list(u)
            up[c] += 1e-6
            um[c] -= 1e-6
            fd = (oracle_lib.gas(gamma, up, axis, 1.1, up)["flux"] - oracle_lib.gas(gamma, um, axis, 1.1, um)["flux"]) / 2e-6
            assert np.max(np.abs(fd - g["jacobian"][:, c])) < 1e-8
        sound = math.sqrt(gamma * p_exp / u[0])
        vel = u[axis + 1] / u[0]
        assert abs(g["speeds"][0] - 1.1 * (sound - vel)) < 1e-12
        assert abs(g["speeds"][1] - 1.1 * (sound + vel)) < 1e-12


def test_one_sided_hll_branch(oracle_lib):
    """src/verification.cpp:285-287."""
    l, r = [1, 5, -5, 25.5], [1.1, 5.5, -5.5, 28.]
    assert oracle_lib.gas(3., l, 0, 1.1, r)["speeds"][0] == 0
    assert oracle_lib.gas(3., l, 1, 1.1, r)["speeds"][1] == 0


def test_exact_solution_satisfies_pde(oracle_lib):
    """src/verification.cpp:288-300: central FD residual of U_t + F_x + G_y."""
    eps, worst = 1e-6, 0.
    flux = lambda v, ax: oracle_lib.gas(3., v, ax, 1.1, v)["flux"]
    for i in range(20):
        x, y, t, a = .07 * i, .13 * i, .07, math.pi / 6
        ex = lambda xx, yy, tt: oracle_lib.exact(3., xx, yy, tt, a)
        r = (ex(x, y, t + eps) - ex(x, y, t - eps) + flux(ex(x + eps, y, t), 0) - flux(ex(x - eps, y, t), 0)
             + flux(ex(x, y + eps, t), 1) - flux(ex(x, y - eps, t), 1)) / (2 * eps)
        worst = max(worst, float(np.max(np.abs(r))))
    assert worst < 2e-8


def test_exact_solution_closed_form_at_t0(oracle_lib):
    """At t=0: xi = phase, v = 1 + sin(pi xi)/2, c = (g-1)(2 - sin)/4, p = rho^g."""
    g, a = 3., math.radians(30)
    for x, y in [(.1, .2), (-.4, .9), (1.3, -2.)]:
        z = math.cos(a) * x + math.sin(a) * y
        z -= 2 * math.floor((z + 1) / 2)
        j = math.sin(math.pi * z)
        v, c = 1 + j / 2, (g - 1) * (2 - j) / 4
        rho = (c * c / g) ** (1 / (g - 1))
        want = [rho, rho * v * math.cos(a), rho * v * math.sin(a), rho ** g / (g - 1) + .5 * rho * v * v]
        assert np.allclose(oracle_lib.exact(g, x, y, 0., a), want, rtol=1e-14, atol=1e-15)


# --------------------------------------------------------------- GL tables
@pytest.mark.parametrize("k", range(9))
def test_gl_tables(oracle_lib, k):
    o = oracle_lib.Oracle(f"--nx 2 --ny 2 --k {k}")
    x, w, d, e0, e1 = o.tables()
    q = k + 1
    xr, wr = np.polynomial.legendre.leggauss(q)
    assert np.allclose(x, (xr + 1) / 2, rtol=0, atol=2e-16 * 4)
    assert np.allclose(w, wr / 2, rtol=0, atol=2e-15)  # numpy's leggauss weights carry ~1e-15 error
    assert abs(sum(w) - 1) < 1e-15
    D = d.reshape(q, q)  # D[a, i] = l_i'(x_a)
    for deg in range(q):  # derivative of x^deg is exact for degree <= k
        f = x ** deg
        df = deg * x ** max(deg - 1, 0) if deg else np.zeros(q)
        assert np.allclose(D @ f, df, atol=1e-11)
    for deg in range(q):  # boundary values
        assert abs(e0 @ x ** deg - (1. if deg == 0 else 0.)) < 1e-12
        assert abs(e1 @ x ** deg - 1.) < 1e-12
    for deg in range(2 * q):  # quadrature exactness 2q-1
        assert abs(w @ x ** deg - 1 / (deg + 1)) < 1e-14


def test_gl_textbook_values(oracle_lib):
    x, w, *_ = oracle_lib.Oracle("--nx 2 --ny 2 --k 2").tables()
    assert np.allclose(x, [.5 - math.sqrt(.15), .5, .5 + math.sqrt(.15)], atol=1e-16)
    assert np.allclose(w, [5 / 18, 4 / 9, 5 / 18], atol=1e-16)


# --------------------------------------------------------------- residual & Jacobian
def state(o, mix=.05):
    """src/verification.cpp:26-30: .05 Initial + .95 exact(0,0,0)."""
    u = o.initial()
    bg = oracle_lib_exact_background(o)
    return mix * u + (1 - mix) * np.tile(bg, o.nu // 4)


def oracle_lib_exact_background(o):
    import oracle
    g = o.geometry()
    alpha = math.atan2(2 / g[4], 2 / g[3])  # lx = 2/cos a, ly = 2/sin a
    return oracle.exact(3., 0., 0., 0., alpha)


@pytest.mark.parametrize("k", range(9))
def test_jacobian_fd_and_dense(oracle_lib, k):
    """src/verification.cpp:33-50,81-84: FD of the residual vs the matrix-free
    action (< 1e-7), dense assembled vs matrix-free (< 1e-12)."""
    o = oracle_lib.Oracle(f"--nx 3 --ny 2 --k {k}")
    u = state(o)
    o.freeze(u, .001)
    v = np.sin(.37 * (np.arange(o.nu) + 1))
    o.dg_refresh(u)
    z = o.dg_apply(v)
    eps = 1e-7
    fd = (o.residual(u + eps * v) - o.residual(u - eps * v)) * (.5 / eps)
    assert relerr(fd, z) < 1e-7
    if o.nu <= 2500:
        a = o.dense_dg(u)
        assert relerr(a @ v, z) < 1e-12
        # block-Jacobi blocks are the dense diagonal blocks
        blocks = o.diagonal_blocks(0)
        for e in range(o.n):
            s = slice(e * o.m, (e + 1) * o.m)
            assert np.max(np.abs(blocks[e] - a[s, s])) < 1e-12 * max(1, np.max(np.abs(a[s, s])))


def test_free_stream_preservation(oracle_lib):
    """A constant admissible state has r = u - old: GL quadrature integrates the
    constant flux exactly and HLL with U_L = U_R returns F(U)."""
    o = oracle_lib.Oracle("--nx 4 --ny 3 --k 3 --alpha-deg 30")
    u = np.tile([1.1, .3, -.2, 2.9], o.nu // 4)
    o.freeze(u, .01)
    o.set_stage(.97 * u, .01)
    r = o.residual(u)
    assert np.max(np.abs(r - (u - .97 * u))) < 1e-13


def test_conservation_of_residual(oracle_lib):
    """Sum_cells M (r - u + old) telescopes to zero over the periodic faces."""
    o = oracle_lib.Oracle("--nx 4 --ny 3 --k 2")
    u = state(o, .5)
    o.freeze(u, .01)
    old = 0.9 * u
    o.set_stage(old, .01)
    r = o.residual(u)
    metric, _ = o.metrics()
    for c in range(4):
        s = np.sum((metric * (r - u + old))[c::4])
        assert abs(s) < 1e-14 * np.sum(np.abs(metric * (r - u + old)))


# --------------------------------------------------------------- preconditioners
def scalar_ilu0_apply(a, pattern_mask, b):
    """Textbook scalar ILU(0) (IKJ variant, no pivoting) on the expanded block
    pattern, then the L U triangular solves."""
    n = a.shape[0]
    lu = a.copy()
    for i in range(1, n):
        for k in range(i):
            if not pattern_mask[i, k]:
                continue
            lu[i, k] /= lu[k, k]
            cols = np.nonzero(pattern_mask[i, k + 1:])[0] + k + 1
            lu[i, cols] -= lu[i, k] * lu[k, cols]
    y = b.copy()
    for i in range(n):
        y[i] -= lu[i, :i] @ y[:i]
    for i in range(n - 1, -1, -1):
        y[i] = (y[i] - lu[i, i + 1:] @ y[i + 1:]) / lu[i, i]
    return y


def expand_mask(row, col, b, nb):
    mask = np.zeros((nb * b, nb * b), dtype=bool)
    for r in range(nb):
        for k in range(row[r], row[r + 1]):
            c = col[k]
            mask[r * b:(r + 1) * b, c * b:(c + 1) * b] = True
    return mask


@pytest.mark.parametrize("args", ["--nx 3 --ny 2 --k 1", "--nx 3 --ny 3 --k 0", "--nx 2 --ny 3 --k 1"])
def test_dg_preconditioners_brute_force(oracle_lib, args):
    """BSR ILU(0) and block-Jacobi vs scalar ILU(0) / dense block solves on the
    assembled Jacobian (replaces the mfem::BlockILU check of
    src/verification.cpp:91-107)."""
    o = oracle_lib.Oracle(args)
    u = state(o)
    o.freeze(u, .001)
    o.set_stage(u, .05)
    o.dg_refresh(u)
    a = o.dense_dg(u)
    b = np.sin(.17 * (np.arange(o.nu) + 1))
    row, col = o.ilu_pattern(0)
    mask = expand_mask(row, col, o.m, o.n)
    assert not np.any(a[~mask])  # the block graph covers every coupling
    assert relerr(o.precondition(0, 1, b), scalar_ilu0_apply(a, mask, b)) < 1e-10
    bj = np.concatenate([np.linalg.solve(a[e * o.m:(e + 1) * o.m, e * o.m:(e + 1) * o.m], b[e * o.m:(e + 1) * o.m])
                         for e in range(o.n)])
    assert relerr(o.precondition(0, 0, b), bj) < 1e-12


def condensed(o, u, r):
    X, y, rhs, D = o.condense(u, r)
    S = np.column_stack([o.schur_apply(np.eye(o.nl)[:, j]) for j in range(o.nl)])
    return X, y, rhs, D, S


@pytest.mark.parametrize("args", ["--nx 3 --ny 2 --k 1", "--nx 2 --ny 2 --k 0", "--nx 3 --ny 3 --k 2"])
def test_hdg_algebra(oracle_lib, args):
    """DG/HDG increment equivalence (src/verification.cpp:51-61, < 1e-8), the
    recovered-residual identity eps_DG = -B D^{-1} eps_S
    (src/stepper.cpp:7-26), Schur preconditioners vs brute force."""
    o = oracle_lib.Oracle(args)
    u = state(o)
    o.freeze(u, .001)
    r = o.residual(u)
    o.dg_refresh(u)
    a = o.dense_dg(u)
    du_dg = np.linalg.solve(a, -r)
    X, y, rhs, D, S = condensed(o, u, r)
    dl = np.linalg.solve(S, rhs)
    du_hdg = o.recover(dl)
    assert relerr(du_hdg, du_dg) < 1e-8
    # identity for an inexact trace increment
    dl_bad = dl + 1e-3 * np.sin(.23 * np.arange(o.nl))
    du_bad = o.recover(dl_bad)
    eps_s = S @ dl_bad - rhs
    assert relerr(o.recovered_residual(eps_s), a @ du_bad + r) < 1e-9
    # preconditioners
    b = np.sin(.19 * (np.arange(o.nl) + 1))
    row, col = o.ilu_pattern(1)
    mask = expand_mask(row, col, o.t, o.nf)
    assert not np.any(S[~mask])
    assert relerr(o.precondition(1, 1, b), scalar_ilu0_apply(S, mask, b)) < 1e-10
    bj = np.concatenate([np.linalg.solve(S[f * o.t:(f + 1) * o.t, f * o.t:(f + 1) * o.t], b[f * o.t:(f + 1) * o.t])
                         for f in range(o.nf)])
    assert relerr(o.precondition(1, 0, b), bj) < 1e-12


@pytest.mark.parametrize("precond", [0, 1])
def test_gmres_matches_direct(oracle_lib, precond):
    o = oracle_lib.Oracle(f"--nx 3 --ny 3 --k 2 --precond {'block-ilu0' if precond else 'block-jacobi'}")
    u = state(o)
    o.freeze(u, .01)
    r = o.residual(u)
    o.dg_refresh(u)
    x, eq, cnt, rel = o.linear_solve(0, -r)
    assert rel <= 1e-12 and cnt > 0
    assert relerr(x, np.linalg.solve(o.dense_dg(u), -r)) < 1e-9


def test_weighted_gmres_invariance(oracle_lib):
    """src/verification.cpp:174-254 known-answer case."""
    import ctypes
    out = np.zeros(10)
    lib = oracle_lib.load()
    assert lib.or_verify_weighted(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    for k in range(2):
        err, c1, c2, c0, zx = out[5 * k:5 * k + 5]
        assert err < 1e-8 and c1 == c2 and c0 == 0 and zx == 0


# --------------------------------------------------------------- whole runs
def test_reference_integration_smoke(oracle_lib, tmp_path):
    """code_test/test_solver.py:47-56 on the oracle CLI."""
    import csv
    import subprocess
    out = tmp_path / "smoke"
    subprocess.run([str(oracle_lib.CLI), "--output", str(out), "--newton-tol", "1e-12", "--linear-tol", "1e-12",
                    "--nx", "8", "--ny", "8", "--k", "1", "--cfl", ".1", "--steps", "1", "--form", "both",
                    "--no-save-solution"], check=True, capture_output=True)
    comp = next(csv.DictReader(open(f"{out}_comparison.csv")))
    assert float(comp["max_nodal_difference"]) < 1e-8


def test_reference_integration_backward_euler(oracle_lib, tmp_path):
    """code_test/test_solver.py:58-74 on the oracle CLI."""
    import csv
    import subprocess
    out = tmp_path / "be"
    subprocess.run([str(oracle_lib.CLI), "--output", str(out), "--newton-tol", "1e-12", "--linear-tol", "1e-12",
                    "--nx", "8", "--ny", "8", "--k", "2", "--cfl", ".1", "--steps", "1", "--integrator", "be",
                    "--form", "both", "--check-update", "--no-save-solution"], check=True, capture_output=True)
    for r in csv.DictReader(open(f"{out}_hdg_iterations.csv")):
        assert float(r["update_check_relative"]) < 1e-8
    comp = next(csv.DictReader(open(f"{out}_comparison.csv")))
    assert float(comp["max_nodal_difference"]) < 1e-8 and float(comp["solution_L2_difference"]) < 1e-8


# --------------------------------------------------------------- anisotropic geometry and published values
def iterations_csv(path):
    import csv
    return list(csv.DictReader(open(path)))


@pytest.mark.parametrize("alpha", [30., 60.])
def test_wave_rate_closed_form_anisotropic(oracle_lib, alpha):
    """WaveRate = max HLL bound / min(hx, hy) (src/problem.cpp:377-384). For
    the initial simple wave at gamma = 3 (v = 1 + j/2, c = (2 - j)/2,
    j = sin(pi xi) in [-1, 1]) the largest bound is c + v n.d at j = -1:
    1.1 (1.5 + cos(a)/2) on x-faces, 1.1 (1.5 + sin(a)/2) on y-faces. On a
    12 x 8 mesh at 30 or 60 degrees hx != hy and the two face families
    differ, so max(hx, hy), a swapped axis or a swapped angle all fail."""
    a = math.radians(alpha)
    o = oracle_lib.Oracle(f"--nx 12 --ny 8 --k 2 --alpha-deg {alpha}")
    hx, hy, area, lx, ly, _ = o.geometry()
    assert abs(lx - 2 / math.cos(a)) < 1e-14 and abs(ly - 2 / math.sin(a)) < 1e-14
    assert abs(hx - lx / 12) < 1e-15 and abs(hy - ly / 8) < 1e-15 and abs(area - hx * hy) < 1e-15
    rate = o.freeze(o.initial(), 1e-3)
    want = 1.1 * (1.5 + max(math.cos(a), math.sin(a)) / 2) / min(hx, hy)
    assert abs(rate - want) < 0.01 * want, (rate, want)
    for wrong in (1.1 * (1.5 + max(math.cos(a), math.sin(a)) / 2) / max(hx, hy),
                  1.1 * (1.5 + min(math.cos(a), math.sin(a)) / 2) / min(hx, hy)):
        assert abs(rate - wrong) > 0.03 * wrong


# RESULTS.md:93-100,124-128,141-145 of the reference: k = 2, SSP-DIRK3,
# CFL 2, T = 0.2, Newton and GMRES tolerances 1e-13 (DG = HDG to the digits
# shown). Published before the per-solve frozen-bound policy of the current
# code (RESULTS.md:30-33), which moves the errors by about 0.1 % and the
# Newton update counts (52 published for N = 32, 39 now), so the errors are
# pinned to 0.5 %, the step counts and mean dt exactly.
PUBLISHED_L2 = {32: [1.00539387e-3, 7.41333053e-4, 7.41333053e-4, 6.88933246e-4],
                64: [1.51106794e-4, 1.12145056e-4, 1.12145056e-4, 9.26221620e-5]}
PUBLISHED_STEPS = {32: (3, 0.0666666667), 64: (5, 0.04)}


def published_case(n):
    return f"--nx {n} --ny {n} --k 2 --cfl 2 --T .2 --newton-tol 1e-13 --linear-tol 1e-13"


def test_published_l2_n32_both_forms(oracle_lib, tmp_path):
    o = oracle_lib.Oracle(published_case(32))
    metric, _ = o.metrics()
    finals = {}
    for form in ("dg", "hdg"):
        u, s, *_ = o.run(form, prefix=str(tmp_path / f"n32_{form}"))
        finals[form] = u
        assert np.allclose(s[5:9], PUBLISHED_L2[32], rtol=5e-3, atol=0), (form, s[5:9])
        steps, mean_dt = PUBLISHED_STEPS[32]
        assert int(s[1]) == steps and abs(s[3] - .2) < 1e-15
        row = iterations_csv(tmp_path / f"n32_{form}_summary.csv")[0]
        assert int(row["time_steps"]) == steps and abs(float(row["mean_dt"]) - mean_dt) < 1e-9
        assert abs(float(row["actual_max_cfl"]) - 2.0) < 1e-12
        assert s[4] < 3e-13  # conservation drift, RESULTS.md:145 (1.2e-13)
    d = finals["dg"] - finals["hdg"]
    joint = math.sqrt(float(np.sum(metric * d * d)))
    assert joint < 1e-13  # RESULTS.md:124-128 (7.8e-15)


def test_published_l2_n64_dg(oracle_lib):
    u, s, *_ = oracle_lib.Oracle(published_case(64)).run("dg")
    assert np.allclose(s[5:9], PUBLISHED_L2[64], rtol=5e-3, atol=0), s[5:9]
    assert int(s[1]) == PUBLISHED_STEPS[64][0] and abs(s[3] - .2) < 1e-15 and s[4] < 3e-13


def test_convergence_at_30_degrees(oracle_lib):
    """The alpha = 30 degree wave (lx != ly, hx != hy, the wave not x <-> y
    symmetric) converges at the scheme's order: an x <-> y slip in Initial,
    Errors or the geometry leaves O(0.1) errors that do not converge."""
    errs = {}
    for n in (16, 32):
        u, s, *_ = oracle_lib.Oracle(f"--nx {n} --ny {n} --k 2 --alpha-deg 30 --cfl 1 --T .1").run("dg")
        errs[n] = np.array(s[5:9])
    rates = np.log2(errs[16] / errs[32])
    assert np.all(rates > 2.3), rates
    assert np.all(errs[16] < 1.2e-3) and np.all(errs[32] < 2e-4), errs


# A workload whose first update of stage 2 takes the damped branch of the
# line search (src/stepper.cpp:202-217): lambda = 1/2 is accepted.
LINE_SEARCH_CASE = "--nx 16 --ny 16 --k 2 --cfl 3 --steps 1"


def test_line_search_damping_branch(oracle_lib, tmp_path):
    o = oracle_lib.Oracle(LINE_SEARCH_CASE)
    finals = {}
    for form in ("dg", "hdg"):
        u, s, *_ = o.run(form, prefix=str(tmp_path / f"ls_{form}"))
        finals[form] = u
        damp = [float(r["damping"]) for r in iterations_csv(tmp_path / f"ls_{form}_iterations.csv")]
        assert len(damp) == int(s[0]) and min(damp) == 0.5 and damp.count(0.5) >= 1, damp
        assert all(d in (1.0, 0.5) for d in damp)
    assert np.max(np.abs(finals["dg"] - finals["hdg"])) < 1e-7  # src/main.cpp:47


# --------------------------------------------------------------- diagnostics
def cell_graph(nx, ny):
    g = []
    for e in range(nx * ny):
        ix, iy = e % nx, e // nx
        g.append(sorted({e, iy * nx + (ix + 1) % nx, iy * nx + (ix - 1) % nx, ((iy + 1) % ny) * nx + ix,
                         ((iy - 1) % ny) * nx + ix}))
    return g


def face_graph(nx, ny):
    def face(e, side):
        ix, iy = e % nx, e // nx
        return [2 * (iy * nx + (ix - 1) % nx), 2 * e, 2 * (((iy - 1) % ny) * nx + ix) + 1, 2 * e + 1][side]
    g = [set() for _ in range(2 * nx * ny)]
    for e in range(nx * ny):
        for a in range(4):
            g[face(e, a)].update(face(e, b) for b in range(4))
    return [sorted(s) for s in g]


def mdf_by_definition(graph):
    """The minimum discarded fill rule written out directly: repeatedly take
    the uneliminated row whose elimination would create the fewest fill
    entries (i, j) outside the pattern among its uneliminated neighbours
    (ordered pairs, i != j; lowest row on ties)."""
    n = len(graph)
    nbr = [set(g) | {i} for i, g in enumerate(graph)]
    left = set(range(n))
    order = []
    while left:
        def fill(k):
            live = [i for i in nbr[k] if i != k and i in left]
            return sum(1 for i in live for j in live if i != j and j not in nbr[i])
        k = min(sorted(left), key=fill)
        order.append(k)
        left.remove(k)
    return order


@pytest.mark.parametrize("graph", [cell_graph(4, 3), face_graph(3, 3), face_graph(2, 3), face_graph(4, 2)],
                         ids=["cells 4x3", "faces 3x3", "faces 2x3", "faces 4x2"])
def test_mdf_ordering_rule(oracle_lib, graph):
    """--ilu-ordering mdf (src/block_ilu.cpp:16-26; MFEM's routine is absent,
    its tie order unpinned): the oracle's order is the stated rule, and a
    permutation."""
    got = list(oracle_lib.mdf_ordering(graph))
    assert sorted(got) == list(range(len(graph)))
    assert got == mdf_by_definition(graph)


def test_mdf_permuted_ilu_brute_force(oracle_lib):
    """A reordered block ILU(0) (src/block_ilu.cpp:14-29,87-109) is the
    scalar ILU(0) of the permuted Schur matrix P S P^T on the permuted
    pattern, applied to P b and permuted back."""
    o = oracle_lib.Oracle("--nx 3 --ny 3 --k 1 --precond block-ilu0 --ilu-ordering mdf")
    u = state(o)
    o.freeze(u, .001)
    r = o.residual(u)
    *_, S = condensed(o, u, r)
    order = o.ilu_order(1)
    assert list(order) != list(range(o.nf))  # a nontrivial reordering of the face graph
    idx = np.concatenate([np.arange(f * o.t, (f + 1) * o.t) for f in order])
    row, col = o.ilu_pattern(1)  # the natural pattern
    mask = expand_mask(row, col, o.t, o.nf)
    b = np.sin(.19 * (np.arange(o.nl) + 1))
    yp = scalar_ilu0_apply(S[np.ix_(idx, idx)], mask[np.ix_(idx, idx)], b[idx])
    y = np.zeros(o.nl)
    y[idx] = yp
    assert relerr(o.precondition(1, 1, b), y) < 1e-10


@pytest.mark.parametrize("form", ["dg", "hdg"])
def test_diagnostic_runs_solve_the_same_systems(oracle_lib, form):
    """--solver dense|direct, --assembly element|full and --ilu-ordering mdf
    change how each Newton system is solved, not the system: the final states
    agree with the default run to the GMRES tolerance, and the dense solver
    runs no Krylov iterations."""
    base = "--nx 5 --ny 5 --k 2 --alpha-deg 30 --cfl .5 --steps 1 --newton-steps 2"
    u0, s0, _, _ = oracle_lib.Oracle(base).run(form)
    for extra in ("--solver dense", "--solver direct", "--assembly full --precond block-ilu0",
                  "--assembly element", "--precond block-ilu0 --ilu-ordering mdf"):
        u, s, _, kry = oracle_lib.Oracle(f"{base} {extra}").run(form, increments=8)
        assert int(s[0]) == int(s0[0]), extra
        assert relerr(u, u0) < 1e-10, extra
        if "--solver" in extra:
            assert all(k == 0 for k in kry[:int(s[0])])
