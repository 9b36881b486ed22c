"""GPU counterparts of the oracle pins of tests/test_oracle.py: the
anisotropic geometry (alpha != 45 degrees, hx != hy), the reference's
published accuracy values, the damped branch of the Newton line search and
the mixed-precision form, each run through the C ABI and compared with the
oracle on the same inputs and with the published or closed-form values.
"""
import math

import numpy as np
import pytest

from test_gpu_parity import increment_ok, rel
from test_oracle import LINE_SEARCH_CASE, PUBLISHED_L2, PUBLISHED_STEPS, iterations_csv, published_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alpha", [30., 60.])
def test_wave_rate_anisotropic(gpu, oracle_lib, alpha):
    """Frozen bounds and WaveRate on a 12 x 8 mesh at 30 / 60 degrees: the
    oracle's to 1e-14 and the closed form 1.1 (1.5 + max(cos a, sin a)/2) /
    min(hx, hy) of the initial wave to 1 %."""
    import torch
    args = f"--nx 12 --ny 8 --k 2 --alpha-deg {alpha}"
    print(f"case: {args}", flush=True)
    o = oracle_lib.Oracle(args)
    c = gpu.Context(args)
    hx, hy, *_ = c.geometry()
    u = o.initial()
    assert rel(c.initial(), u) < 1e-13
    rg, fg = c.freeze(torch.tensor(u, device="cuda"), 1e-3, frozen=True)
    ro, fo = o.freeze(u, 1e-3, frozen=True)
    assert abs(rg - ro) <= 1e-14 * ro and rel(fg, fo) < 1e-14
    a = math.radians(alpha)
    want = 1.1 * (1.5 + max(math.cos(a), math.sin(a)) / 2) / min(hx, hy)
    assert abs(rg - want) < 0.01 * want


def test_published_l2_n32(gpu, oracle_lib, tmp_path):
    """RESULTS.md:93-100,124-128,141-145 of the reference on the GPU: L2
    errors within 0.5 % of the published values (and 1e-8 of the oracle's),
    3 steps of mean dt 1/15, drift < 3e-13, DG - HDG joint L2 < 1e-13."""
    args = published_case(32)
    print(f"case: {args}", flush=True)
    o = oracle_lib.Oracle(args)
    c = gpu.Context(args)
    metric, _ = o.metrics()
    finals = {}
    for form in ("dg", "hdg"):
        u, st, *_ = c.run(form, prefix=str(tmp_path / f"n32_{form}"))
        uo, so, *_ = o.run(form)
        finals[form] = u
        assert np.allclose(list(st.errors), PUBLISHED_L2[32], rtol=5e-3, atol=0), (form, list(st.errors))
        assert np.allclose(list(st.errors), so[5:9], rtol=1e-8, atol=0)
        assert st.updates == int(so[0]) and st.time_steps == PUBLISHED_STEPS[32][0]
        row = iterations_csv(tmp_path / f"n32_{form}_summary.csv")[0]
        assert abs(float(row["mean_dt"]) - PUBLISHED_STEPS[32][1]) < 1e-9
        assert st.drift < 3e-13
        assert rel(u, uo) < 1e-10
    d = finals["dg"] - finals["hdg"]
    assert math.sqrt(float(np.sum(metric * d * d))) < 1e-13


@pytest.mark.parametrize("form", ["dg", "hdg"])
def test_convergence_at_30_degrees(gpu, oracle_lib, form):
    """The alpha = 30 degree wave on the GPU: errors equal to the oracle's
    (1e-8 relative) and converging at rate > 2.3 from N = 16 to 32."""
    errs = {}
    for n in (16, 32):
        args = f"--nx {n} --ny {n} --k 2 --alpha-deg 30 --cfl 1 --T .1 --precond block-ilu0"
        print(f"case: {args} {form}", flush=True)
        u, st, *_ = gpu.Context(args).run(form)
        uo, so, *_ = oracle_lib.Oracle(args).run(form)
        errs[n] = np.array(list(st.errors))
        assert np.allclose(errs[n], so[5:9], rtol=1e-8, atol=0)
        assert rel(u, uo) < 1e-10
    assert np.all(np.log2(errs[16] / errs[32]) > 2.3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("form", ["dg", "hdg"])
def test_line_search_damping(gpu, oracle_lib, form, precision, tmp_path):
    """The damped branch of the line search (src/stepper.cpp:202-217): the
    first update of stage 2 accepts lambda = 1/2. The GPU takes the same
    damping factors as the oracle in every update and its increments pass the
    reference's one-update bound against the oracle's."""
    args = LINE_SEARCH_CASE + f" --precision {precision}"
    print(f"case: {args} {form}", flush=True)
    o = oracle_lib.Oracle(LINE_SEARCH_CASE)
    c = gpu.Context(args)
    ug, sg, incg, _ = c.run(form, prefix=str(tmp_path / "gpu"), increments=64)
    uo, so, inco, _ = o.run(form, prefix=str(tmp_path / "cpu"), increments=64)
    dg_ = [float(r["damping"]) for r in iterations_csv(tmp_path / "gpu_iterations.csv")]
    do_ = [float(r["damping"]) for r in iterations_csv(tmp_path / "cpu_iterations.csv")]
    assert dg_ == do_ and min(dg_) == 0.5, (dg_, do_)
    metric, _ = o.metrics()
    u_scale = float(np.sqrt(np.sum(metric * uo * uo)))
    for i in range(int(so[0])):
        ok, e = increment_ok(o, incg[i], inco[i], u_scale, bound=1e-8)
        assert ok, (i, e)
    assert rel(ug, uo) < 1e-9


@pytest.mark.parametrize("form", ["dg", "hdg"])
def test_mixed_precision_vs_oracle(gpu, oracle_lib, form):
    """The mixed-precision form (FP32 Krylov basis, operator and
    preconditioner inside FP64 refinement) against the FP64 oracle on the
    same workload: every accepted increment within the reference's
    one-update bound (1e-8 relative, src/stepper.cpp:197-198), the final state
    within 1e-9."""
    args = "--nx 16 --ny 16 --k 3 --alpha-deg 30 --cfl 1 --steps 1 --newton-steps 3 --restart 30 --precond block-ilu0"
    print(f"case: mixed {form} {args}", flush=True)
    o = oracle_lib.Oracle(args)
    u32, s32, i32, _ = gpu.Context(args + " --precision mixed").run(form, increments=16)
    uo, so, io, _ = o.run(form, increments=16)
    assert s32.updates == int(so[0]) == 12
    metric, _ = o.metrics()
    u_scale = float(np.sqrt(np.sum(metric * uo * uo)))
    for i in range(s32.updates):
        ok, e = increment_ok(o, i32[i], io[i], u_scale, bound=1e-8)
        assert ok, (form, i, e)
    assert rel(u32, uo) < 1e-9


def test_save_outputs(gpu, tmp_path):
    """Save (src/problem.cpp:762-790): _coefficients.csv holds the final state
    (element, node, four components at 17 digits), each component's .gf
    holds the same values in element-DOF order after the L2 Gauss-Legendre
    space header, and the .mesh is the non-periodic Cartesian quadrilateral
    mesh of the shifted box: its vertices span [-lx/2, lx/2] x [-ly/2, ly/2]
    and each boundary attribute's segments lie on their side of the box."""
    import csv
    args = "--nx 6 --ny 4 --k 2 --alpha-deg 30 --dt 1e-3 --steps 1 --newton-steps 1"
    c = gpu.Context(args)
    prefix = str(tmp_path / "save")
    u, st, *_ = c.run("hdg", prefix=prefix)
    n, p, m = c.n, c.p, c.m
    rows = list(csv.DictReader(open(prefix + "_coefficients.csv")))
    assert len(rows) == n * p
    for r in rows:
        e, i = int(r["element"]), int(r["node"])
        got = [float(r[k]) for k in ("rho", "rhou", "rhov", "energy")]
        assert got == [u[e * m + 4 * i + cc] for cc in range(4)]
    for cc, name in enumerate(("rho", "rhou", "rhov", "energy")):
        lines = open(f"{prefix}_{name}.gf").read().split("\n")
        assert lines[:4] == ["FiniteElementSpace", "FiniteElementCollection: L2_2D_P2", "VDim: 1", "Ordering: 0"]
        vals = [float(v) for v in lines[5:] if v.strip()]
        assert vals == [u[e * m + 4 * i + cc] for e in range(n) for i in range(p)]
    hx, hy, area, lx, ly, _ = c.geometry()
    text = open(prefix + ".mesh").read().split("\n")
    assert text[0] == "MFEM mesh v1.0"
    sec = {name: text.index(name) for name in ("dimension", "elements", "boundary", "vertices")}
    assert text[sec["dimension"] + 1] == "2"
    ne = int(text[sec["elements"] + 1])
    elems = [list(map(int, text[sec["elements"] + 2 + j].split())) for j in range(ne)]
    nb = int(text[sec["boundary"] + 1])
    bdr = [list(map(int, text[sec["boundary"] + 2 + j].split())) for j in range(nb)]
    nv = int(text[sec["vertices"] + 1])
    assert text[sec["vertices"] + 2] == "2"
    xy = np.array([list(map(float, text[sec["vertices"] + 3 + j].split())) for j in range(nv)])
    assert ne == n and nb == 2 * (6 + 4) and nv == 7 * 5
    assert np.allclose(xy.min(0), [-lx / 2, -ly / 2], rtol=1e-5) and np.allclose(xy.max(0), [lx / 2, ly / 2], rtol=1e-5)
    for e, el in enumerate(elems):  # lexicographic cells of size hx x hy, counter-clockwise
        assert el[:2] == [1, 3]
        v = xy[el[2:]]
        ix, iy = e % 6, e // 6
        assert np.allclose(v[0], [-lx / 2 + ix * hx, -ly / 2 + iy * hy], rtol=1e-5, atol=1e-5)
        assert np.allclose(v[2] - v[0], [hx, hy], rtol=1e-4)
    side = {1: (1, -ly / 2), 3: (1, ly / 2), 4: (0, -lx / 2), 2: (0, lx / 2)}
    for attr, geom, a, b in bdr:
        axis, value = side[attr]
        assert geom == 1 and abs(xy[a, axis] - value) < 1e-5 and abs(xy[b, axis] - value) < 1e-5


def test_cli_verify_self_test(gpu, tmp_path):
    """euler2d --verify (src/verification.cpp:7-169 on the device kernels):
    exit 0, one 'verified k=' line a degree, and the reference's CSV columns
    with every check within the reference's thresholds."""
    import csv
    import subprocess
    out = tmp_path / "verify"
    r = subprocess.run([str(gpu.CLI), "--verify", "--output", str(out)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert [line.split()[1] for line in r.stdout.splitlines() if line.startswith("verified k=")] == \
        [f"k={k}" for k in range(9)]
    rows = list(csv.DictReader(open(f"{out}_verification.csv")))
    assert [int(x["k"]) for x in rows] == list(range(9))
    for x in rows:
        assert float(x["jacobian_fd_relative"]) < 1e-7 and float(x["dg_hdg_increment_relative"]) < 1e-8
        assert float(x["assembled_mf_relative"]) < 1e-13 and float(x["mass_off_diagonal"]) < 1e-12
        assert float(x["partial_schur_relative"]) < 1e-12
    for x in csv.DictReader(open(f"{out}_storage.csv")):
        assert float(x["dg_bsr_ilu_relative"]) < 1e-10 and float(x["hdg_bsr_ilu_relative"]) < 1e-10
        assert float(x["streamed_schur_relative"]) < 1e-12 and float(x["schur_diagonal_relative"]) < 1e-12
