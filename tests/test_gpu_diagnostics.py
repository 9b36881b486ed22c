"""The reference's assembled and direct diagnostics on the GPU against the
CPU oracle (src/options.cpp:124-147, src/linear.cpp:46-56,184-240,
src/operators.cpp:180-268, src/partial.cpp:174-258, src/block_ilu.cpp:10-29,
87-109): --solver dense|direct (a dense LU with partial pivoting of the
assembled matrix), --assembly element|full (the assembled block matrix as
the operator) and --ilu-ordering mdf (the reordered block ILU(0)).

Each run is compared with the oracle run of the same options update by
update, with the reference's one-update acceptance bound as in
test_run_parity, and with the GPU run of the default options (the same
Newton systems solved to the same tolerance: final states within 1e-9).
"""
import numpy as np
import pytest

from test_gpu_parity import increment_ok, rel

pytestmark = pytest.mark.gpu

RUN = "--cfl .5 --steps 1 --newton-steps 2"
CASES = [
    f"--nx 6 --ny 6 --k 2 {RUN} --solver dense",
    f"--nx 5 --ny 5 --k 2 --alpha-deg 30 {RUN} --solver direct",
    "--nx 2 --ny 3 --k 3 --dt 1e-4 --steps 1 --newton-steps 2 --solver dense",
    f"--nx 5 --ny 4 --k 3 --alpha-deg 30 {RUN} --solver dense",
    f"--nx 6 --ny 6 --k 2 {RUN} --assembly full --precond block-ilu0",
    f"--nx 5 --ny 4 --k 3 --alpha-deg 30 {RUN} --assembly full --precond block-jacobi",
    f"--nx 6 --ny 6 --k 2 {RUN} --assembly element --precond block-jacobi",
    f"--nx 6 --ny 5 --k 3 {RUN} --precond block-ilu0 --ilu-ordering mdf",
    f"--nx 6 --ny 6 --k 4 {RUN} --precond block-ilu0 --ilu-ordering mdf",
    f"--nx 4 --ny 4 --k 5 --steps 1 --newton-steps 1 --cfl .5 --precond block-ilu0 --ilu-ordering mdf",
]


def default_options(args):
    """args without the diagnostic options (--solver, --assembly, --ilu-ordering)."""
    words, out = args.split(), []
    skip = False
    for w in words:
        if skip:
            skip = False
        elif w in ("--solver", "--assembly", "--ilu-ordering"):
            skip = True
        else:
            out.append(w)
    return " ".join(out)


@pytest.mark.parametrize("form", ["dg", "hdg"])
@pytest.mark.parametrize("args", CASES)
def test_diagnostic_run_parity(gpu, oracle_lib, args, form):
    print(f"case: {args} form={form}", flush=True)
    o = oracle_lib.Oracle(args)
    c = gpu.Context(args)
    ugo, so, inco, _ = o.run(form, increments=64)
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
    if "--solver" in args:
        assert all(int(k) == 0 for k in krg[:nupd]) # no Krylov iterations in a direct solve
    # the default options solve the same Newton systems
    ud, sd, _, _ = gpu.Context(default_options(args)).run(form)
    assert sd.updates == nupd
    assert rel(ug, ud) < 1e-9


@pytest.mark.parametrize("form", ["dg", "hdg"])
def test_assembled_and_mdf_mixed_precision(gpu, oracle_lib, form):
    """The FP32 inner solves on the assembled operator and through the
    reordered factor: each update within 1e-8 of the oracle's FP64 update."""
    for extra in ("--assembly full --precond block-ilu0", "--precond block-ilu0 --ilu-ordering mdf"):
        args = f"--nx 8 --ny 8 --k 3 --alpha-deg 30 --cfl 1 --steps 1 --newton-steps 2 --restart 30 {extra}"
        print(f"case: mixed {form} {extra}", flush=True)
        o = oracle_lib.Oracle(args)
        ugo, so, inco, _ = o.run(form, increments=16)
        ug, sg, incg, _ = gpu.Context(args + " --precision mixed").run(form, increments=16)
        assert sg.updates == int(so[0])
        metric, _ = o.metrics()
        u_scale = float(np.sqrt(np.sum(metric * ugo * ugo)))
        for i in range(sg.updates):
            ok, e = increment_ok(o, incg[i], inco[i], u_scale, bound=1e-8)
            assert ok, (extra, i, e)
        assert rel(ug, ugo) < 1e-9


@pytest.mark.parametrize("mesh", ["--nx 3 --ny 3 --k 1", "--nx 4 --ny 6 --k 2", "--nx 2 --ny 5 --k 1",
                                  "--nx 16 --ny 16 --k 3"])
def test_mdf_order_matches_oracle(gpu, oracle_lib, mesh):
    """The solver's minimum discarded fill order (host code of the library)
    is the oracle's, on the cell and the face graphs."""
    args = f"{mesh} --precond block-ilu0 --ilu-ordering mdf"
    c, o = gpu.Context(args), oracle_lib.Oracle(args)
    for form in (0, 1):
        got, want = c.ilu_order(form), o.ilu_order(form)
        assert np.array_equal(got, want), form
        assert np.array_equal(np.sort(got), np.arange(len(got)))


def test_dense_solver_size_limit(gpu):
    """src/linear.cpp:46-49: the dense solver is limited to 5000 unknowns."""
    c = gpu.Context("--nx 8 --ny 8 --k 4 --steps 1 --newton-steps 1 --solver dense")  # DG: 6400 unknowns
    with pytest.raises(RuntimeError, match="limited to 5000 unknowns"):
        c.run("dg")
