"""CPU checks of the drop-in boundary: the HIP library loads and exports every
entry point include/eg_abi.h declares, and the euler2d CLI keeps the
reference's option errors and exit convention (src/options.cpp:8-147,
src/main.cpp:50-53). No compute call is made, so no GPU is needed."""
import ctypes
import subprocess

import pytest

import mfem_euler2d_cpp_amd as eg

CLI = eg.LIB.parent / "euler2d"


@pytest.fixture(scope="module")
def lib():
    if not eg.LIB.exists():
        eg.build(force=False)
    return ctypes.CDLL(str(eg.LIB))


def test_header_declares_the_boundary():
    names = eg.declared_symbols()
    for required in ("eg_create", "eg_destroy", "eg_last_error", "eg_freeze", "eg_residual", "eg_cell_norm",
                     "eg_admissible", "eg_dg_refresh", "eg_dg_apply", "eg_pc_setup", "eg_pc_apply",
                     "eg_hdg_condense", "eg_schur_apply", "eg_recover", "eg_recovered_residual", "eg_linear_solve",
                     "eg_run"):
        assert required in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in eg.declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_loader_binds_every_declared_symbol():
    assert sorted(eg._SIGNATURES) == eg.declared_symbols()


@pytest.mark.parametrize("args, message", [
    (["--nx", "0"], "Invalid mesh, degree, time or safety"),
    (["--form", "cg"], "Invalid form"),
    (["--integrator", "rk4"], "Invalid time integrator"),
    (["--solver", "pardiso"], "Invalid solver"),
    (["--precond", "amg"], "Invalid preconditioner"),
    (["--assembly", "csr"], "Invalid assembly"),
    (["--ilu-ordering", "rcm"], "Invalid ILU ordering"),
    (["--solver", "dense", "--matrix-free"], "Direct LU requires an assembled matrix"),
    (["--solver", "direct", "--assembly", "partial"], "Direct LU requires --assembly full"),
    (["--solver", "dense", "--precision", "mixed"], "--precision mixed needs --solver gmres"),
    (["--precision", "half"], "Invalid precision"),
    (["--nx"], "Missing option value"),
])
def test_cli_option_errors(args, message):
    if not CLI.exists():
        eg.build(force=False)
    r = subprocess.run([str(CLI)] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 1
    assert r.stderr.startswith("ERROR: ") and message in r.stderr
