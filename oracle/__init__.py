"""CPU ORACLE -- test infrastructure only.

ctypes loader for the C++17 restatement of the reference algorithm
(oracle/*.cpp, built into oracle/_build/liboracle.so by oracle/Makefile).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
as the checker; the product path (mfem_euler2d_cpp_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
CLI = HERE / "_build" / "euler2d_oracle"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None


def build(force: bool = False) -> None:
    cmd = ["make", "-C", str(HERE), f"-j{min(16, os.cpu_count() or 1)}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = ctypes.CDLL(str(LIB))
        lib.or_last_error.restype = ctypes.c_char_p
        lib.or_cell_norm.restype = ctypes.c_double
        _lib = lib
    return _lib


def d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class OracleError(RuntimeError):
    pass


class Oracle:
    """One oracle Problem built from reference CLI flags."""

    def __init__(self, args: str, gamma: float = 3.0):
        self.lib = load()
        self.h = _vp()
        self._check(self.lib.or_create(args.encode(), ctypes.c_double(gamma), ctypes.byref(self.h)))
        s = (ctypes.c_int * 8)()
        self.lib.or_sizes(self.h, s)
        self.n, self.p, self.m, self.t, self.nf, self.nu, self.nl, self.q = list(s)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc:
            raise OracleError(self.lib.or_last_error().decode())

    def call(self, name, *args):
        self._check(getattr(self.lib, name)(self.h, *args))

    def tables(self):
        q = self.q
        x, w, dd, e0, e1 = (np.zeros(q), np.zeros(q), np.zeros(q * q), np.zeros(q), np.zeros(q))
        self.call("or_tables", d(x), d(w), d(dd), d(e0), d(e1))
        return x, w, dd, e0, e1

    def geometry(self):
        g = np.zeros(6)
        self.call("or_geometry", d(g))
        return g

    def metrics(self):
        c, t = np.zeros(self.nu), np.zeros(self.nl)
        self.call("or_metrics", d(c), d(t))
        return c, t

    def initial(self):
        u = np.zeros(self.nu)
        self.call("or_initial", d(u))
        return u

    def freeze(self, u, dt, frozen=False):
        rate = ctypes.c_double()
        fz = np.zeros(self.nf * self.q * 6) if frozen else None
        self.call("or_freeze", d(np.ascontiguousarray(u)), ctypes.c_double(dt), ctypes.byref(rate), d(fz))
        return (rate.value, fz) if frozen else rate.value

    def set_stage(self, old, dt):
        self.call("or_set_stage", d(np.ascontiguousarray(old)), ctypes.c_double(dt))

    def residual(self, u):
        r = np.zeros(self.nu)
        self.call("or_residual", d(np.ascontiguousarray(u)), d(r))
        return r

    def cell_norm(self, r):
        return self.lib.or_cell_norm(self.h, d(np.ascontiguousarray(r)))

    def admissible(self, u):
        return bool(self.lib.or_admissible(self.h, d(np.ascontiguousarray(u))))

    def integral(self, u):
        o = np.zeros(4)
        self.call("or_integral", d(np.ascontiguousarray(u)), d(o))
        return o

    def errors(self, u, time):
        o = np.zeros(4)
        self.call("or_errors", d(np.ascontiguousarray(u)), ctypes.c_double(time), d(o))
        return o

    def dg_refresh(self, u):
        self.call("or_dg_refresh", d(np.ascontiguousarray(u)))

    def dg_apply(self, x):
        y = np.zeros(self.nu)
        self.call("or_dg_apply", d(np.ascontiguousarray(x)), d(y))
        return y

    def dense_dg(self, u):
        a = np.zeros((self.nu, self.nu), order="F")
        self.call("or_dense_dg", d(np.ascontiguousarray(u)), a.ctypes.data_as(_dp))
        return a

    def diagonal_blocks(self, form):
        b = self.m if form == 0 else self.t
        nb = self.n if form == 0 else self.nf
        out = np.zeros(nb * b * b)
        self.call("or_diagonal_blocks", ctypes.c_int(form), d(out))
        return out.reshape(nb, b, b).transpose(0, 2, 1)  # row-major view of column-major blocks

    def precondition(self, form, precond, b):
        x = np.zeros_like(b)
        self.call("or_precondition", ctypes.c_int(form), ctypes.c_int(precond), d(np.ascontiguousarray(b)), d(x))
        return x

    def ilu_pattern(self, form):
        nnzb = ctypes.c_int()
        self.call("or_ilu_pattern", ctypes.c_int(form), ctypes.byref(nnzb), None, None)
        nrow = (self.n if form == 0 else self.nf) + 1
        row = np.zeros(nrow, dtype=np.int32)
        col = np.zeros(nnzb.value, dtype=np.int32)
        self.call("or_ilu_pattern", ctypes.c_int(form), ctypes.byref(nnzb), row.ctypes.data_as(_ip),
                  col.ctypes.data_as(_ip))
        return row, col

    def ilu_order(self, form):
        """The ILU(0) factor's block row order for this handle's --ilu-ordering."""
        order = np.zeros(self.n if form == 0 else self.nf, dtype=np.int32)
        self.call("or_ilu_order", ctypes.c_int(form), order.ctypes.data_as(_ip))
        return order

    def ilu_blocks(self, form, factored):
        row, col = self.ilu_pattern(form)
        b = self.m if form == 0 else self.t
        vals = np.zeros(len(col) * b * b)
        self.call("or_ilu_blocks", ctypes.c_int(form), ctypes.c_int(int(factored)), d(vals))
        return row, col, vals.reshape(len(col), b, b).transpose(0, 2, 1)

    def condense(self, u, r, want_x=True):
        n, m, t = self.n, self.m, self.t
        X = np.zeros(n * m * 4 * t) if want_x else None
        y, rhs, D = np.zeros(self.nu), np.zeros(self.nl), np.zeros(self.nl)
        self.call("or_condense", d(np.ascontiguousarray(u)), d(np.ascontiguousarray(r)), d(X), d(y), d(rhs), d(D))
        return X, y, rhs, D

    def schur_apply(self, x):
        y = np.zeros(self.nl)
        self.call("or_schur_apply", d(np.ascontiguousarray(x)), d(y))
        return y

    def recover(self, dl):
        du = np.zeros(self.nu)
        self.call("or_recover", d(np.ascontiguousarray(dl)), d(du))
        return du

    def recovered_residual(self, eps):
        out = np.zeros(self.nu)
        self.call("or_recovered_residual", d(np.ascontiguousarray(eps)), d(out))
        return out

    def linear_solve(self, form, rhs):
        size = self.nu if form == 0 else self.nl
        x, eq = np.zeros(size), np.zeros(size)
        cnt, rel = ctypes.c_int(), ctypes.c_double()
        self.call("or_linear_solve", ctypes.c_int(form), d(np.ascontiguousarray(rhs)), d(x), d(eq),
                  ctypes.byref(cnt), ctypes.byref(rel))
        return x, eq, cnt.value, rel.value

    def run(self, form, prefix=None, increments=0):
        u = np.zeros(self.nu)
        stats = np.zeros(9)
        inc = np.zeros((increments, self.nu)) if increments else None
        kry = (ctypes.c_int * max(increments, 1))()
        self.call("or_run", ctypes.c_int(1 if form in (1, "hdg") else 0), prefix.encode() if prefix else None,
                  d(u), d(stats), inc.ctypes.data_as(_dp) if inc is not None else None, ctypes.c_int(increments),
                  kry if increments else None)
        return u, stats, inc, list(kry)[:increments]


def mdf_ordering(graph):
    """The oracle's minimum discarded fill order of a graph (adjacency lists)."""
    lib = load()
    ptr = np.zeros(len(graph) + 1, dtype=np.int32)
    for i, g in enumerate(graph):
        ptr[i + 1] = ptr[i] + len(g)
    adj = np.array([j for g in graph for j in g] or [0], dtype=np.int32)
    order = np.zeros(len(graph), dtype=np.int32)
    if lib.or_mdf_ordering(ctypes.c_int(len(graph)), ptr.ctypes.data_as(_ip), adj.ctypes.data_as(_ip),
                           order.ctypes.data_as(_ip)):
        raise OracleError(lib.or_last_error().decode())
    return order


def gas(gamma, u, axis, safety, r):
    lib = load()
    out = np.zeros(24)
    rc = lib.or_gas(ctypes.c_double(gamma), d(np.asarray(u, float)), ctypes.c_int(axis), ctypes.c_double(safety),
                    d(np.asarray(r, float)), d(out))
    if rc:
        raise OracleError(lib.or_last_error().decode())
    return {"pressure": out[0], "positive": bool(out[1]), "flux": out[2:6],
            "jacobian": out[6:22].reshape(4, 4).T, "speeds": out[22:24]}


def exact(gamma, x, y, time, alpha):
    lib = load()
    out = np.zeros(4)
    rc = lib.or_exact(ctypes.c_double(gamma), ctypes.c_double(x), ctypes.c_double(y), ctypes.c_double(time),
                      ctypes.c_double(alpha), d(out))
    if rc:
        raise OracleError(lib.or_last_error().decode())
    return out
