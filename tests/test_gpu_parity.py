"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances: kernels that evaluate the same formulas in a different summation
order agree to ~1e-13 relative; dense inverses (Gauss-Jordan on the GPU vs the
oracle's pivoted LU solves) to 1e-10; Newton increments and final states of
whole runs at GMRES tolerance 1e-12 to 1e-9 relative (the north-star bar is
1e-8 per update).
"""
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = [f"--nx 3 --ny 2 --k {k}" for k in range(5)] + ["--nx 4 --ny 5 --k 2 --alpha-deg 30"]
# cell blocks above 100 x 100 (k >= 5) take the large-block kernels
BIG = ["--nx 3 --ny 2 --k 5", "--nx 2 --ny 3 --k 6", "--nx 2 --ny 2 --k 7", "--nx 2 --ny 2 --k 8"]
CFG1 = "--nx 8 --ny 8 --k 2 --alpha-deg 45 --steps 1 --dt 1e-4 --newton-steps 1 --precond block-jacobi"
CFG2 = "--nx 16 --ny 16 --k 3 --cfl 1 --steps 3 --newton-steps 3 --precond block-ilu0"


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def perturbed(o):
    """.05 Initial + .95 exact(0,0,0) as in src/verification.cpp:26-30 (coarse
    projections are not positive at every face trace), then a 2% ripple."""
    import math

    import oracle
    g = o.geometry()
    bg = oracle.exact(3., 0., 0., 0., math.atan2(2 / g[4], 2 / g[3]))
    u = .05 * o.initial() + .95 * np.tile(bg, o.nu // 4)
    i = np.arange(o.nu)
    return u * (1 + 0.02 * np.sin(0.37 * (i + 1)))


def make(gpu, oracle_lib, args, dt=1e-3):
    import torch
    print(f"case: {args}", flush=True)
    o = oracle_lib.Oracle(args)
    c = gpu.Context(args)
    u = perturbed(o)
    du = torch.tensor(u, device="cuda")
    return o, c, u, du


@pytest.mark.parametrize("args", SMALL + [CFG1, CFG2])
def test_tables_initial_freeze(gpu, oracle_lib, args):
    o, c, u, du = make(gpu, oracle_lib, args)
    for a, b in zip(c.tables(), o.tables()):
        assert np.allclose(a, b, rtol=0, atol=1e-14), (a, b)
    assert rel(c.initial(), o.initial()) < 1e-13
    rg, fg = c.freeze(du, 1e-3, frozen=True)
    ro, fo = o.freeze(u, 1e-3, frozen=True)
    assert abs(rg - ro) <= 1e-14 * ro
    assert rel(fg, fo) < 1e-14


@pytest.mark.parametrize("args", SMALL + BIG + [CFG1, CFG2])
def test_residual_and_trial(gpu, oracle_lib, args):
    import torch
    o, c, u, du = make(gpu, oracle_lib, args)
    c.freeze(du, 1e-3)
    o.freeze(u, 1e-3)
    old = 0.999 * u
    o.set_stage(old, 2e-3)
    c.set_stage(torch.tensor(old, device="cuda"), 2e-3)
    r = torch.zeros_like(du)
    ng = c.residual(du, r)
    ro = o.residual(u)
    assert rel(r.cpu().numpy(), ro) < 1e-12
    assert abs(ng - o.cell_norm(ro)) <= 1e-12 * ng
    step = torch.tensor(np.cos(0.11 * np.arange(o.nu)) * 1e-3 * np.abs(u), device="cuda")
    trial = torch.zeros_like(du)
    nt, adm = c.trial_residual(du, step, 0.5, trial, r)
    ut = u + 0.5 * step.cpu().numpy()
    assert adm and o.admissible(ut)
    assert rel(trial.cpu().numpy(), ut) < 1e-15
    assert rel(r.cpu().numpy(), o.residual(ut)) < 1e-12
    # a large negative step makes the density negative: inadmissible
    _, adm = c.trial_residual(du, du, -2.0, trial, r)
    assert not adm
    assert c.admissible(du) and not c.admissible(-du)


@pytest.mark.parametrize("args", SMALL + BIG + [CFG2])
def test_dg_apply(gpu, oracle_lib, args):
    import torch
    o, c, u, du = make(gpu, oracle_lib, args)
    c.freeze(du, 1e-3)
    o.freeze(u, 1e-3)
    c.set_stage(du, 0.05)
    o.set_stage(u, 0.05)
    o.dg_refresh(u)
    c.call("eg_dg_refresh", gpu._ptr(du))
    x = np.sin(0.37 * (np.arange(o.nu) + 1))
    y = torch.zeros_like(du)
    c.call("eg_dg_apply", gpu._ptr(torch.tensor(x, device="cuda")), gpu._ptr(y))
    assert rel(y.cpu().numpy(), o.dg_apply(x)) < 1e-12
    y32 = torch.zeros(o.nu, dtype=torch.float32, device="cuda")
    c.call("eg_dg_apply_f32", gpu._ptr(torch.tensor(x, dtype=torch.float32, device="cuda")), gpu._ptr(y32))
    assert rel(y32.cpu().numpy(), o.dg_apply(x)) < 1e-5


@pytest.mark.parametrize("precond", ["block-jacobi", "block-ilu0", "block-ilu0 --ilu-ordering mdf"])
@pytest.mark.parametrize("args", SMALL + BIG + [CFG2.replace(" --precond block-ilu0", "")])
def test_dg_preconditioner(gpu, oracle_lib, args, precond):
    import torch
    args = f"{args} --precond {precond}"
    o, c, u, du = make(gpu, oracle_lib, args)
    c.freeze(du, 1e-3)
    o.freeze(u, 1e-3)
    c.set_stage(du, 0.05)
    o.set_stage(u, 0.05)
    o.dg_refresh(u)
    c.call("eg_dg_refresh", gpu._ptr(du))
    c.call("eg_pc_setup", 0)
    b = np.sin(0.17 * (np.arange(o.nu) + 1))
    y = torch.zeros_like(du)
    c.call("eg_pc_apply", 0, gpu._ptr(torch.tensor(b, device="cuda")), gpu._ptr(y))
    ref = o.precondition(0, 1 if precond.startswith("block-ilu0") else 0, b)
    assert rel(y.cpu().numpy(), ref) < 1e-10


def hdg_case(gpu, oracle_lib, args):
    import torch
    o, c, u, du = make(gpu, oracle_lib, args)
    c.freeze(du, 1e-3)
    o.freeze(u, 1e-3)
    c.set_stage(du, 0.05)
    o.set_stage(u, 0.05)
    r = 1e-3 * np.sin(0.29 * (np.arange(o.nu) + 1))
    dr = torch.tensor(r, device="cuda")
    c.call("eg_hdg_condense", gpu._ptr(du), gpu._ptr(dr))
    X, y, rhs, D = o.condense(u, r)
    return o, c, u, du, X, y, rhs, D


@pytest.mark.parametrize("args", SMALL + BIG + [CFG2])
def test_hdg_condense_schur_recover(gpu, oracle_lib, args):
    import torch
    o, c, u, du, X, y, rhs, D = hdg_case(gpu, oracle_lib, args)
    n, m, t = o.n, o.m, o.t
    grhs, gD = torch.zeros(o.nl, dtype=torch.float64, device="cuda"), torch.zeros(o.nl, dtype=torch.float64,
                                                                                  device="cuda")
    c.call("eg_hdg_rhs", gpu._ptr(grhs), gpu._ptr(gD))
    assert rel(gD.cpu().numpy(), D) < 1e-14
    assert rel(grhs.cpu().numpy(), rhs) < 1e-10
    xy = torch.zeros(n * m * (4 * t + 1), dtype=torch.float64, device="cuda")
    c.call("eg_hdg_xy", gpu._ptr(xy))
    xy = xy.cpu().numpy().reshape(n, 4 * t + 1, m)
    assert rel(xy[:, :4 * t, :].reshape(-1), X.reshape(n, 4 * t, m).reshape(-1)) < 1e-10
    assert rel(xy[:, 4 * t, :].reshape(-1), y) < 1e-10
    x = np.sin(0.27 * np.arange(o.nl))
    dx = torch.tensor(x, device="cuda")
    out = torch.zeros(o.nl, dtype=torch.float64, device="cuda")
    c.call("eg_schur_apply", gpu._ptr(dx), gpu._ptr(out))
    assert rel(out.cpu().numpy(), o.schur_apply(x)) < 1e-10
    dun = torch.zeros_like(du)
    c.call("eg_recover", gpu._ptr(dx), gpu._ptr(dun))
    assert rel(dun.cpu().numpy(), o.recover(x)) < 1e-10
    er = torch.zeros_like(du)
    import ctypes
    nrm = ctypes.c_double()
    c.call("eg_recovered_residual", gpu._ptr(dx), gpu._ptr(er), ctypes.byref(nrm))
    ref = o.recovered_residual(x)
    assert rel(er.cpu().numpy(), ref) < 1e-13
    assert abs(nrm.value - o.cell_norm(ref)) <= 1e-12 * nrm.value


@pytest.mark.parametrize("precond", ["block-jacobi", "block-ilu0", "block-ilu0 --ilu-ordering mdf"])
@pytest.mark.parametrize("args", SMALL + BIG + [CFG2.replace(" --precond block-ilu0", "")])
def test_hdg_preconditioner(gpu, oracle_lib, args, precond):
    import torch
    o, c, u, du, *_ = hdg_case(gpu, oracle_lib, f"{args} --precond {precond}")
    c.call("eg_pc_setup", 1)
    b = np.sin(0.17 * (np.arange(o.nl) + 1))
    y = torch.zeros(o.nl, dtype=torch.float64, device="cuda")
    c.call("eg_pc_apply", 1, gpu._ptr(torch.tensor(b, device="cuda")), gpu._ptr(y))
    ref = o.precondition(1, 1 if precond.startswith("block-ilu0") else 0, b)
    assert rel(y.cpu().numpy(), ref) < 1e-10


@pytest.mark.parametrize("precond", ["block-jacobi", "block-ilu0"])
@pytest.mark.parametrize("args", ["--nx 3 --ny 2 --k 2", "--nx 8 --ny 8 --k 2", "--nx 16 --ny 16 --k 3"])
def test_linear_solves(gpu, oracle_lib, args, precond):
    import torch
    args = f"{args} --precond {precond}"
    o, c, u, du, *_ = hdg_case(gpu, oracle_lib, args)
    # DG system
    o.dg_refresh(u)
    c.call("eg_dg_refresh", gpu._ptr(du))
    b = -1e-3 * np.sin(0.31 * (np.arange(o.nu) + 1))
    xo, eqo, cnt, relres = o.linear_solve(0, b)
    xg = torch.zeros_like(du)
    st = c.linear_solve(0, torch.tensor(b, device="cuda"), xg, tol=1e-12)
    assert rel(xg.cpu().numpy(), xo) < 1e-9
    assert abs(st.iterations - cnt) <= 1 and st.relative_residual <= 1e-11
    # condensed Schur system
    X, y, rhs, D = o.condense(u, 1e-3 * np.sin(0.29 * (np.arange(o.nu) + 1)), want_x=False)
    xo, eqo, cnt, relres = o.linear_solve(1, rhs)
    xg = torch.zeros(o.nl, dtype=torch.float64, device="cuda")
    st = c.linear_solve(1, torch.tensor(rhs, device="cuda"), xg, tol=1e-12)
    assert rel(xg.cpu().numpy(), xo) < 1e-9
    assert abs(st.iterations - cnt) <= 1


def increment_ok(o, inc_gpu, inc_ref, u_scale, bound=1e-8):
    """The reference's DG/HDG one-update equivalence criterion
    (src/stepper.cpp:197-198): ||du - du_ref||_M <= max(bound ||du_ref||_M,
    4*allowance), allowance = 128 eps (||u||_M + ||old||_M + ||r||_M). Near
    Newton convergence the residual itself is rounding noise of size
    ~eps ||u||, so the increments only agree to the allowance there."""
    metric, _ = o.metrics()
    nrm = lambda v: float(np.sqrt(np.sum(metric * v * v)))
    allowance = 128 * np.finfo(float).eps * 2 * u_scale
    d, r = nrm(inc_gpu - inc_ref), nrm(inc_ref)
    return d <= max(bound * r, 4 * allowance), d / max(r, 1e-300)


@pytest.mark.parametrize("form", ["dg", "hdg"])
@pytest.mark.parametrize("args", [CFG1, CFG2.replace("--steps 3", "--steps 1"),
                                  "--nx 12 --ny 12 --k 1 --cfl .5 --steps 1 --newton-steps 2 --precond block-ilu0",
                                  "--nx 6 --ny 6 --k 4 --cfl .5 --steps 1 --newton-steps 2",
                                  "--nx 4 --ny 4 --k 6 --cfl .5 --steps 1 --newton-steps 1 --precond block-ilu0"])
def test_run_parity(gpu, oracle_lib, args, form):
    o = oracle_lib.Oracle(args)
    c = gpu.Context(args)
    print(f"case: {args} form={form}", flush=True)
    ugo, so, inco, kro = o.run(form, increments=64)
    ug, sg, incg, krg = c.run(form, increments=64)
    nupd = int(so[0])
    assert sg.updates == nupd and sg.time_steps == int(so[1])
    metric, _ = o.metrics()
    u_scale = float(np.sqrt(np.sum(metric * ugo * ugo)))
    worst = 0.
    for i in range(nupd):
        ok, e = increment_ok(o, incg[i], inco[i], u_scale, bound=1e-9)
        worst = max(worst, e)
        assert ok, (i, e)
    print(f"max per-update relative increment error {worst:.2e}")
    assert rel(ug, ugo) < 1e-10
    assert abs(sg.drift) < 1e-12
    assert np.allclose(list(sg.errors), so[5:9], rtol=1e-8)


def test_mixed_precision_run(gpu, oracle_lib):
    args = "--nx 16 --ny 16 --k 3 --alpha-deg 30 --cfl 1 --steps 1 --newton-steps 3 --restart 30 --precond block-ilu0"
    for form in ("dg", "hdg"):
        print(f"case: mixed {form}", flush=True)
        c64 = gpu.Context(args)
        c32 = gpu.Context(args + " --precision mixed")
        u64, s64, i64, _ = c64.run(form, increments=16)
        u32, s32, i32, _ = c32.run(form, increments=16)
        o = oracle_lib.Oracle(args)
        metric, _ = o.metrics()
        u_scale = float(np.sqrt(np.sum(metric * u64 * u64)))
        for i in range(s64.updates):
            ok, e = increment_ok(o, i32[i], i64[i], u_scale, bound=1e-8)
            assert ok, (form, i, e)
        assert rel(u32, u64) < 1e-9


LARGE = {"64x64 Q4": "--nx 64 --ny 64 --k 4 --alpha-deg 30 --cfl 1 --steps 1 --newton-steps 1 --restart 30 "
                      "--precond block-ilu0",
         "128x128 Q6": "--nx 128 --ny 128 --k 6 --alpha-deg 30 --cfl 1 --steps 1 --newton-steps 1 --restart 30 "
                       "--precond block-ilu0"}


@pytest.mark.parametrize("size", list(LARGE))
@pytest.mark.parametrize("form", ["dg", "hdg"])
def test_mixed_precision_multirow_sweeps(gpu, form, size):
    """64x64 Q4 and 128x128 Q6 (the mixed-precision workloads), more block
    rows than dataflow-sweep workgroups (several rows a workgroup, FP64 and
    FP32 sweeps): the mixed-precision increments against the FP64 ones of the
    same run (relative 2-norm 1e-9; about 1e-12 is measured). Too large for
    the oracle; the FP64 path is pinned against it at the smaller sizes."""
    args = LARGE[size]
    print(f"case: {size} mixed {form}", flush=True)
    c64 = gpu.Context(args)
    c32 = gpu.Context(args + " --precision mixed")
    u64, s64, i64, _ = c64.run(form, increments=4)
    u32, s32, i32, _ = c32.run(form, increments=4)
    assert s64.updates == s32.updates > 0
    for i in range(s64.updates):
        assert np.all(np.isfinite(i32[i])) and rel(i32[i], i64[i]) < 1e-9, (form, i)
    assert rel(u32, u64) < 1e-9


def test_cli_reference_smoke(gpu, tmp_path):
    """code_test/test_solver.py:47-56 thresholds on the GPU euler2d CLI."""
    import csv
    out = tmp_path / "smoke"
    subprocess.run([str(gpu.CLI), "--output", str(out), "--newton-tol", "1e-12", "--linear-tol", "1e-12",
                    "--nx", "8", "--ny", "8", "--k", "1", "--cfl", ".1", "--steps", "1", "--form", "both",
                    "--no-save-solution"], check=True, timeout=300)
    for form in ("dg", "hdg"):
        row = next(csv.DictReader(open(f"{out}_{form}_summary.csv")))
        assert row["integrator"] == "ssp-dirk3" and int(row["time_steps"]) == 1 and int(row["newton_updates"]) > 0
    comp = next(csv.DictReader(open(f"{out}_comparison.csv")))
    assert float(comp["max_nodal_difference"]) < 1e-8


def test_cli_backward_euler_update_check(gpu, tmp_path):
    """code_test/test_solver.py:58-74: converged BE, --check-update."""
    import csv
    out = tmp_path / "be"
    subprocess.run([str(gpu.CLI), "--output", str(out), "--newton-tol", "1e-12", "--linear-tol", "1e-12",
                    "--nx", "8", "--ny", "8", "--k", "2", "--cfl", ".1", "--steps", "1", "--integrator", "be",
                    "--form", "both", "--check-update", "--no-save-solution"], check=True, timeout=300)
    for r in csv.DictReader(open(f"{out}_hdg_iterations.csv")):
        v = float(r["update_check_relative"])
        assert np.isfinite(v) and v < 1e-8
    comp = next(csv.DictReader(open(f"{out}_comparison.csv")))
    assert float(comp["max_nodal_difference"]) < 1e-8 and float(comp["solution_L2_difference"]) < 1e-8


@pytest.mark.parametrize("args", ["--nx 16 --ny 16 --k 3", "--nx 5 --ny 3 --k 2", "--nx 4 --ny 6 --k 0",
                                  "--nx 3 --ny 3 --k 5", "--nx 4 --ny 3 --k 6"])
def test_ilu_dataflow_matches_levels(gpu, oracle_lib, args, monkeypatch):
    """The single-launch dataflow ILU(0) sweeps (sentinel-armed dependency
    polling) and factor (per-row completion flags) give the preconditioner of
    the one-launch-per-level schedule. The sweeps sum in the same order
    (1e-13); the single-launch factor uses other dense kernels than the
    per-level one (1e-12)."""
    import torch
    out = {}
    for mode in ("levels", "dataflow", "dataflow-all"):
        monkeypatch.setenv("EG_ILU_SCHEDULE", mode)
        o, c, u, du, *_ = hdg_case(gpu, oracle_lib, args + " --precond block-ilu0")
        res = []
        for form, size in ((1, o.nl), (0, o.nu)):
            if form == 0:
                c.call("eg_dg_refresh", gpu._ptr(du))
                o.dg_refresh(u)
            c.call("eg_pc_setup", form)
            b = np.sin(0.13 * (np.arange(size) + 1))
            y = torch.zeros(size, dtype=torch.float64, device="cuda")
            c.call("eg_pc_apply", form, gpu._ptr(torch.tensor(b, device="cuda")), gpu._ptr(y))
            res.append(y.cpu().numpy())
            assert rel(res[-1], o.precondition(form, 1, b)) < 1e-10
        out[mode] = res
    for mode, tol in (("dataflow", 1e-13), ("dataflow-all", 1e-12)):
        for a, b in zip(out["levels"], out[mode]):
            assert rel(a, b) < tol, mode


@pytest.mark.parametrize("args", ["--nx 6 --ny 6 --k 4", "--nx 6 --ny 5 --k 5", "--nx 5 --ny 4 --k 6"])
@pytest.mark.parametrize("slices", ["0", "1", "3", "8"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_row_sliced_sweeps(gpu, oracle_lib, args, slices, precision, monkeypatch):
    """DG ILU(0) application with 100 / 144 / 196 cell blocks: the
    row-sliced dataflow sweeps (upper blocks scaled by the MFMA kernel, each
    block row over 1, 3 or 8 workgroups) and the unsliced ones (0: for 100
    the register-resident sweep with scaled blocks, above 100 the streamed
    sweep with unscaled ones), three applications in a row (the intermediate
    vectors alternate), in FP64 and FP32 against the oracle."""
    import torch
    monkeypatch.setenv("EG_SWEEP_SLICES", slices)
    base = f"{args} --precond block-ilu0"
    print(f"case: {base} --precision {precision}, {slices} slices", flush=True)
    o = oracle_lib.Oracle(base)  # the oracle has no precision option
    c = gpu.Context(f"{base} --precision {precision}")
    u = perturbed(o)
    du = torch.tensor(u, device="cuda")
    c.freeze(du, 1e-3)
    o.freeze(u, 1e-3)
    c.set_stage(du, 0.05)
    o.set_stage(u, 0.05)
    c.call("eg_dg_refresh", gpu._ptr(du))
    o.dg_refresh(u)
    c.call("eg_pc_setup", 0)
    f32 = precision == "mixed"
    for rep in range(3):
        b = np.sin((0.11 + 0.07 * rep) * (np.arange(o.nu) + 1))
        dt = torch.float32 if f32 else torch.float64
        y = torch.zeros(o.nu, dtype=dt, device="cuda")
        c.call("eg_pc_apply_f32" if f32 else "eg_pc_apply", 0, gpu._ptr(torch.tensor(b, dtype=dt, device="cuda")),
               gpu._ptr(y))
        assert rel(y.double().cpu().numpy(), o.precondition(0, 1, b)) < (2e-4 if f32 else 1e-10), rep


@pytest.mark.parametrize("args", ["--nx 6 --ny 6 --k 4", "--nx 6 --ny 5 --k 5", "--nx 5 --ny 4 --k 6"])
def test_fp32_factor(gpu, oracle_lib, args, monkeypatch):
    """The mixed form's FP32 ILU(0) factor of 100 / 144 / 196 cell blocks
    (slab products on v_mfma_f32_16x16x4f32, FP32 Gauss-Jordan inverses and
    upper scaling): its preconditioned vector against the oracle's FP64 one
    and against the FP64 factor converted to FP32 (EG_FACTOR32=0), both to a
    single-precision bound, on three applications in a row; and the FP64
    apply of the same mixed-precision solver (its FP64 factor, computed on
    first use) bitwise equal to an FP64 solver's."""
    import torch
    base = f"{args} --precond block-ilu0"
    print(f"case: FP32 factor {base}", flush=True)
    o = oracle_lib.Oracle(base)
    u = perturbed(o)
    du = torch.tensor(u, device="cuda")
    o.freeze(u, 1e-3)
    o.set_stage(u, 0.05)
    o.dg_refresh(u)
    ctx = {}
    for name, extra, f32 in (("fp32", " --precision mixed", None), ("converted", " --precision mixed", "0"),
                             ("fp64", "", None)):
        if f32 is None:
            monkeypatch.delenv("EG_FACTOR32", raising=False)
        else:
            monkeypatch.setenv("EG_FACTOR32", f32)
        c = gpu.Context(base + extra)
        c.freeze(du, 1e-3)
        c.set_stage(du, 0.05)
        c.call("eg_dg_refresh", gpu._ptr(du))
        c.call("eg_pc_setup", 0)
        ctx[name] = c
    for rep in range(3):
        b = np.sin((0.11 + 0.07 * rep) * (np.arange(o.nu) + 1))
        ref = o.precondition(0, 1, b)
        y = {}
        for name in ("fp32", "converted"):
            y[name] = torch.zeros(o.nu, dtype=torch.float32, device="cuda")
            ctx[name].call("eg_pc_apply_f32", 0, gpu._ptr(torch.tensor(b, dtype=torch.float32, device="cuda")),
                           gpu._ptr(y[name]))
        e_oracle = rel(y["fp32"].double().cpu().numpy(), ref)
        e_conv = rel(y["fp32"].double().cpu().numpy(), y["converted"].double().cpu().numpy())
        print(f"rep {rep}: FP32 factor vs oracle {e_oracle:.2e}, vs converted FP64 factor {e_conv:.2e}")
        assert e_oracle < 2e-4 and e_conv < 2e-4, rep
    b = torch.tensor(np.cos(0.13 * (np.arange(o.nu) + 1)), device="cuda")
    y64 = [torch.zeros(o.nu, dtype=torch.float64, device="cuda") for _ in range(2)]
    ctx["fp32"].call("eg_pc_apply", 0, gpu._ptr(b), gpu._ptr(y64[0]))
    ctx["fp64"].call("eg_pc_apply", 0, gpu._ptr(b), gpu._ptr(y64[1]))
    assert torch.equal(y64[0], y64[1])


@pytest.mark.parametrize("args", ["--nx 6 --ny 6 --k 4 --steps 1 --newton-steps 2",
                                  "--nx 4 --ny 4 --k 5 --steps 1 --newton-steps 2",
                                  "--nx 4 --ny 4 --k 6 --steps 1 --newton-steps 1"])
def test_fp32_factor_run(gpu, oracle_lib, args):
    """DG runs of the mixed form with the FP32 factor (k = 4, 5, 6): every
    update within 1e-8 of the oracle's FP64 update, the final state within
    1e-9."""
    args = f"{args} --alpha-deg 30 --cfl .5 --restart 30 --precond block-ilu0"
    print(f"case: mixed dg {args}", flush=True)
    o = oracle_lib.Oracle(args)
    ugo, so, inco, _ = o.run("dg", increments=16)
    ug, sg, incg, _ = gpu.Context(args + " --precision mixed").run("dg", increments=16)
    assert sg.updates == int(so[0])
    metric, _ = o.metrics()
    u_scale = float(np.sqrt(np.sum(metric * ugo * ugo)))
    for i in range(sg.updates):
        ok, e = increment_ok(o, incg[i], inco[i], u_scale, bound=1e-8)
        assert ok, (i, e)
    assert rel(ug, ugo) < 1e-9


@pytest.mark.parametrize("args", ["--nx 16 --ny 16 --k 3", "--nx 5 --ny 3 --k 2", "--nx 4 --ny 4 --k 4",
                                  "--nx 3 --ny 3 --k 6", "--nx 64 --ny 64 --k 1"])
@pytest.mark.parametrize("env", [("EG_SCALED_UPPER", "0"), ("EG_SLAB", "vector"), ("EG_FACTOR_GRAPH", "0"),
                                 ("EG_SLAB_STAR", "0"), ("EG_RESIDENT_DF", "0")])
def test_ilu_variants_match_default(gpu, oracle_lib, args, env, monkeypatch):
    """The ILU(0) preconditioner with the factor's upper blocks unscaled (the
    reference's inv(U_ii) (w - sum U y) backward step), with the vector-FMA
    slab products instead of the MFMA ones, with the per-level factor issued
    without a graph, with the general slab kernel instead of the two-stage
    one of the cell stencil, and with the face factor per level instead of
    one dataflow launch (64x64 k = 1: more face rows than workgroups, so a
    workgroup factors several rows): each against the oracle (1e-10) and
    against the default path (1e-12: other summation orders)."""
    import torch
    out = {}
    for variant in ("default", "env"):
        if variant == "env":
            monkeypatch.setenv(*env)
        else:
            monkeypatch.delenv(env[0], raising=False)
        o, c, u, du, *_ = hdg_case(gpu, oracle_lib, args + " --precond block-ilu0")
        res = []
        for form, size in ((1, o.nl), (0, o.nu)):
            if form == 0:
                c.call("eg_dg_refresh", gpu._ptr(du))
                o.dg_refresh(u)
            c.call("eg_pc_setup", form)
            b = np.sin(0.11 * (np.arange(size) + 1))
            y = torch.zeros(size, dtype=torch.float64, device="cuda")
            c.call("eg_pc_apply", form, gpu._ptr(torch.tensor(b, device="cuda")), gpu._ptr(y))
            res.append(y.cpu().numpy())
            assert rel(res[-1], o.precondition(form, 1, b)) < 1e-10, (variant, form)
        out[variant] = res
    for a, b in zip(out["default"], out["env"]):
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("size", list(LARGE))
def test_large_dataflow_matches_levels(gpu, size, monkeypatch):
    """The single-launch dataflow sweeps (rows from a ticket counter, several
    rows a workgroup at these sizes) against the one-launch-per-level
    schedule at 64x64 Q4 and 128x128 Q6, both systems, FP64 and FP32: the
    same preconditioned vector (same summation order; 1e-13 relative)."""
    import torch
    args = LARGE[size].replace(" --steps 1 --newton-steps 1", "") + " --precision mixed"
    out = {}
    for mode in ("levels", "dataflow"):
        monkeypatch.setenv("EG_ILU_SCHEDULE", mode)
        print(f"case: {size} {mode}", flush=True)
        c = gpu.Context(args)
        u = torch.tensor(c.initial(), device="cuda")
        c.freeze(u, 1e-3)
        c.set_stage(u, 0.01)
        r = torch.zeros_like(u)
        c.residual(u, r)
        c.call("eg_hdg_condense", gpu._ptr(u), gpu._ptr(r))
        c.call("eg_dg_refresh", gpu._ptr(u))
        res = []
        for form, size_ in ((1, c.nl), (0, c.nu)):
            c.call("eg_pc_setup", form)
            b = torch.sin(0.13 * (torch.arange(size_, device="cuda", dtype=torch.float64) + 1))
            y = torch.zeros_like(b)
            c.call("eg_pc_apply", form, gpu._ptr(b), gpu._ptr(y))
            y32 = torch.zeros(size_, dtype=torch.float32, device="cuda")
            c.call("eg_pc_apply_f32", form, gpu._ptr(b.float()), gpu._ptr(y32))
            res += [y.cpu().numpy(), y32.double().cpu().numpy()]
            assert np.all(np.isfinite(res[-2])) and np.all(np.isfinite(res[-1]))
            assert rel(res[-1], res[-2]) < 1e-4, (mode, form)  # FP32 against FP64
        out[mode] = res
        del c
    # FP64: the same summation order; FP32: the per-level kernels reduce a
    # row's products in another order (single-precision rounding, carried
    # through up to 255 dependency levels at 128x128)
    for a, b, what in zip(out["levels"], out["dataflow"], ("hdg fp64", "hdg fp32", "dg fp64", "dg fp32")):
        assert rel(b, a) < (1e-13 if "fp64" in what else 1e-4), what
