"""Golden fixture checks (tests/golden/cfg1.npz, written by
tests/golden/make_golden.py from the CPU oracle) for the reference CPU case
(8x8 Q2 simple wave, gamma 3, alpha 45 deg, dt 1e-4).

CPU: the oracle still reproduces the fixture bit for bit, and the fixture
satisfies what the mathematics fixes (weights summing to one, derivative of a
constant zero, nodal interpolation of the end values, positivity, linearity
of the Jacobian action). GPU: the HIP library, called through the C ABI,
reproduces the fixture (initial projection and GL tables to 1e-15, Jacobian
action to 1e-12 and the pure spatial residual to 1e-10 relative)."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "cfg1.npz"
ARGS = "--nx 8 --ny 8 --k 2 --alpha-deg 45 --steps 1 --dt 1e-4 --newton-steps 1 --precond block-jacobi"
DT = 1e-4


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_oracle_reproduces_fixture(golden, oracle_lib):
    import sys
    sys.path.insert(0, str(GOLDEN.parent))
    import make_golden
    now = make_golden.compute()
    assert sorted(now) == sorted(golden)
    for k in golden:
        assert np.array_equal(now[k], golden[k]), k


def test_fixture_invariants(golden, oracle_lib):
    for k in range(9):
        q = k + 1
        x, w = golden[f"gl_k{k}_x"], golden[f"gl_k{k}_w"]
        dd = golden[f"gl_k{k}_d"].reshape(q, q)
        e0, e1 = golden[f"gl_k{k}_e0"], golden[f"gl_k{k}_e1"]
        assert abs(w.sum() - 1) < 1e-14 and np.all(np.diff(x) > 0)
        assert np.max(np.abs(dd.sum(axis=1))) < 1e-11          # d/dx of the constant 1
        assert abs(e0.sum() - 1) < 1e-13 and abs(e1.sum() - 1) < 1e-13
        if k >= 1:  # the polynomial x is in the space: d/dx x = 1, x(0) = 0, x(1) = 1
            assert np.max(np.abs(dd @ x - 1)) < 1e-11
            assert abs(e0 @ x) < 1e-13 and abs(e1 @ x - 1) < 1e-13
    u0 = golden["u0"].reshape(-1, 9, 4)
    assert np.all(u0[..., 0] > 0)
    o = oracle_lib.Oracle(ARGS)
    o.freeze(golden["u0"], DT)
    o.dg_refresh(golden["u0"])
    v = np.sin(0.37 * (np.arange(o.nu) + 1))
    w = np.cos(0.11 * np.arange(o.nu))
    assert rel(o.dg_apply(2 * v - 3 * w), 2 * golden["jv"] - 3 * o.dg_apply(w)) < 1e-13


@pytest.mark.gpu
def test_gpu_matches_fixture(gpu, golden):
    import torch
    c = gpu.Context(ARGS)
    for k in (2,):
        x, w, dd, e0, e1 = c.tables()
        for name, a in zip("x w d e0 e1".split(), (x, w, dd, e0, e1)):
            assert rel(a, golden[f"gl_k{k}_{name}"]) < 1e-15, name
    u0 = c.initial()
    assert rel(u0, golden["u0"]) < 1e-15
    u = torch.tensor(u0, device="cuda")
    rate = c.freeze(u, DT)
    assert abs(rate - golden["rate"][0]) <= 1e-15 * abs(golden["rate"][0])
    c.set_stage(u, DT)
    r = torch.zeros_like(u)
    c.residual(u, r)
    # u = old: r is the spatial term alone, dt M^{-1}(volume + faces), whose
    # volume and face parts are each ~1e2 times larger than their sum here.
    assert rel(r.cpu().numpy(), golden["r0"]) < 1e-10
    c.call("eg_dg_refresh", gpu._ptr(u))
    v = torch.tensor(np.sin(0.37 * (np.arange(c.nu) + 1)), device="cuda")
    y = torch.zeros_like(v)
    c.call("eg_dg_apply", gpu._ptr(v), gpu._ptr(y))
    assert rel(y.cpu().numpy(), golden["jv"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(9))
def test_gpu_tables_match_fixture(gpu, golden, k):
    c = gpu.Context(f"--nx 2 --ny 2 --k {k}")
    for name, a in zip("x w d e0 e1".split(), c.tables()):
        assert rel(a, golden[f"gl_k{k}_{name}"]) < 1e-14, name
