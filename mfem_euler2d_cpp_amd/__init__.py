"""MI355X-native implicit SSP-DIRK3 Newton update (DG and condensed HDG) for
the 2-D Euler simple wave.

The product is ``libeg_hip.so`` (HIP kernels for gfx950 + the C++ host API
``euler::Problem/LinearSolver/Run`` + the C ABI of ``include/eg_abi.h``) and
the ``euler2d`` command-line program with the reference options. This module
only loads the library through ctypes and offers a thin wrapper used by the
tests and the benchmark; it never falls back to CPU code: a missing library
raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libeg_hip.so"
CLI = PKG / "euler2d"
HEADER = ROOT / "include" / "eg_abi.h"

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class LinearStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("outer", ctypes.c_int),
                ("relative_residual", ctypes.c_double), ("residual_norm", ctypes.c_double),
                ("setup_s", ctypes.c_double), ("solve_s", ctypes.c_double)]


class RunStats(ctypes.Structure):
    _fields_ = [("updates", ctypes.c_int), ("time_steps", ctypes.c_int),
                ("newton_seconds", ctypes.c_double), ("final_time", ctypes.c_double),
                ("drift", ctypes.c_double), ("seconds", ctypes.c_double),
                ("errors", ctypes.c_double * 4)]


class SolveOutcome(ctypes.Structure):
    _fields_ = [("arrival", ctypes.c_double), ("start", ctypes.c_double), ("finish", ctypes.c_double),
                ("newton_seconds", ctypes.c_double), ("footprint", ctypes.c_double),
                ("measured_peak", ctypes.c_double), ("errors", ctypes.c_double * 4), ("status", ctypes.c_int),
                ("updates", ctypes.c_int), ("time_steps", ctypes.c_int), ("worker", ctypes.c_int)]


class ServeReport(ctypes.Structure):
    _fields_ = [("wall", ctypes.c_double), ("max_reserved", ctypes.c_double), ("max_concurrent", ctypes.c_int)]


class BatchReport(ctypes.Structure):
    _fields_ = [("flushes", ctypes.c_longlong), ("chunks", ctypes.c_longlong), ("graph_launches", ctypes.c_longlong),
                ("max_group", ctypes.c_int), ("tile", ctypes.c_int * 10), ("launch_s", ctypes.c_double),
                ("window_s", ctypes.c_double)]


_SIGNATURES = {
    "eg_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_double, _vp, ctypes.POINTER(_vp)]),
    "eg_destroy": (None, [_vp]),
    "eg_last_error": (ctypes.c_char_p, [_vp]),
    "eg_sizes": (ctypes.c_int, [_vp, _c_int_p]),
    "eg_geometry": (ctypes.c_int, [_vp, _c_double_p]),
    "eg_tables": (ctypes.c_int, [_vp, _c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_double_p]),
    "eg_sync": (ctypes.c_int, [_vp]),
    "eg_initial": (ctypes.c_int, [_vp, _c_double_p]),
    "eg_freeze": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _c_double_p, _c_double_p]),
    "eg_set_stage": (ctypes.c_int, [_vp, _vp, ctypes.c_double]),
    "eg_residual": (ctypes.c_int, [_vp, _vp, _vp, _c_double_p]),
    "eg_trial_residual": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, _vp, _vp, _c_double_p, _c_int_p]),
    "eg_cell_norm": (ctypes.c_int, [_vp, _vp, _c_double_p]),
    "eg_cell_norm_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_long, _vp, _vp]),
    "eg_admissible": (ctypes.c_int, [_vp, _vp, _c_int_p]),
    "eg_dg_refresh": (ctypes.c_int, [_vp, _vp]),
    "eg_dg_apply": (ctypes.c_int, [_vp, _vp, _vp]),
    "eg_dg_apply_f32": (ctypes.c_int, [_vp, _vp, _vp]),
    "eg_pc_setup": (ctypes.c_int, [_vp, ctypes.c_int]),
    "eg_pc_apply": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "eg_pc_apply_f32": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "eg_hdg_condense": (ctypes.c_int, [_vp, _vp, _vp]),
    "eg_hdg_rhs": (ctypes.c_int, [_vp, _vp, _vp]),
    "eg_hdg_xy": (ctypes.c_int, [_vp, _vp]),
    "eg_schur_apply": (ctypes.c_int, [_vp, _vp, _vp]),
    "eg_recover": (ctypes.c_int, [_vp, _vp, _vp]),
    "eg_recovered_residual": (ctypes.c_int, [_vp, _vp, _vp, _c_double_p]),
    "eg_linear_solve": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(LinearStats)]),
    "eg_ilu_order": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p]),
    "eg_kernel_time": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, _c_double_p, _c_double_p, _c_double_p]),
    "eg_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_char_p, _c_double_p, ctypes.POINTER(RunStats),
                              _c_double_p, ctypes.c_int, _c_int_p]),
    "eg_solve_footprint": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _c_double_p]),
    "eg_serve": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), _c_int_p, _c_double_p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.POINTER(SolveOutcome), ctypes.POINTER(_c_double_p),
                                ctypes.c_char_p, ctypes.POINTER(ServeReport)]),
    "eg_serve_batched": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), _c_int_p, _c_double_p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(SolveOutcome),
                                        ctypes.POINTER(_c_double_p), ctypes.c_char_p, ctypes.POINTER(ServeReport),
                                        ctypes.POINTER(BatchReport)]),
    "eg_autotune_ilu": (ctypes.c_int, [_c_int_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
                                       _c_double_p]),
}

_lib = None


def build(force: bool = True) -> None:
    """Compile libeg_hip.so and euler2d for gfx950 (hipcc, in-tree)."""
    cmd = ["make", "-C", str(PKG / "csrc"), f"-j{min(16, os.cpu_count() or 1)}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)


def load():
    """Load libeg_hip.so (raises FileNotFoundError when it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        raise FileNotFoundError(f"{LIB} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(str(LIB))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols() -> list[str]:
    """Names of the entry points declared in include/eg_abi.h."""
    import re
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(eg_[a-z0-9_]+)\s*\(", text)))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dptr(a):
    return a.ctypes.data_as(_c_double_p) if a is not None else None


class EulerError(RuntimeError):
    pass


class Context:
    """One device-resident problem (eg_ctx) built from reference CLI flags.

    Device vectors are torch CUDA tensors (plumbing only); every numerical
    call goes to the HIP library on the tensor's current stream."""

    def __init__(self, args: str, gamma: float = 3.0, stream=None):
        import torch
        self.lib = load()
        self.torch = torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self.ctx = _vp()
        rc = self.lib.eg_create(args.encode(), gamma, ctypes.c_void_p(stream), ctypes.byref(self.ctx))
        if rc:
            raise EulerError(self.lib.eg_last_error(None).decode())
        s = (ctypes.c_int * 8)()
        self.lib.eg_sizes(self.ctx, s)
        self.n, self.p, self.m, self.t, self.nf, self.nu, self.nl, self.q = list(s)

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib.eg_destroy(self.ctx)
            self.ctx = None

    def _check(self, rc):
        if rc:
            raise EulerError(self.lib.eg_last_error(self.ctx).decode())

    def call(self, name, *args):
        # The context runs on its own stream: order it after torch's work
        # and make its outputs visible to torch before returning.
        self.torch.cuda.synchronize()
        self._check(getattr(self.lib, name)(self.ctx, *args))
        self._check(self.lib.eg_sync(self.ctx))

    # convenience wrappers -------------------------------------------------
    def vec(self, n=None, dtype=None):
        t = self.torch
        return t.zeros(n or self.nu, dtype=dtype or t.float64, device="cuda")

    def geometry(self):
        g = (ctypes.c_double * 6)()
        self.call("eg_geometry", g)
        return list(g)

    def tables(self):
        import numpy as np
        q = self.q
        x, w, d, e0, e1 = (np.zeros(q), np.zeros(q), np.zeros(q * q), np.zeros(q), np.zeros(q))
        self.call("eg_tables", _dptr(x), _dptr(w), _dptr(d), _dptr(e0), _dptr(e1))
        return x, w, d, e0, e1

    def initial(self):
        import numpy as np
        u = np.zeros(self.nu)
        self.call("eg_initial", _dptr(u))
        return u

    def freeze(self, u, dt, frozen=False):
        import numpy as np
        rate = ctypes.c_double()
        fz = np.zeros(self.nf * self.q * 6) if frozen else None
        self.call("eg_freeze", _ptr(u), dt, ctypes.byref(rate), _dptr(fz))
        return (rate.value, fz) if frozen else rate.value

    def set_stage(self, old, dt):
        self.call("eg_set_stage", _ptr(old), dt)

    def residual(self, u, r):
        nrm = ctypes.c_double()
        self.call("eg_residual", _ptr(u), _ptr(r), ctypes.byref(nrm))
        return nrm.value

    def trial_residual(self, u, du, lam, trial, r):
        nrm, adm = ctypes.c_double(), ctypes.c_int()
        self.call("eg_trial_residual", _ptr(u), _ptr(du), lam, _ptr(trial), _ptr(r), ctypes.byref(nrm),
                  ctypes.byref(adm))
        return nrm.value, bool(adm.value)

    def cell_norm(self, v):
        out = ctypes.c_double()
        self.call("eg_cell_norm", _ptr(v), ctypes.byref(out))
        return out.value

    def admissible(self, u):
        ok = ctypes.c_int()
        self.call("eg_admissible", _ptr(u), ctypes.byref(ok))
        return bool(ok.value)

    def linear_solve(self, form, rhs, x, res=None, tol=1e-12, refresh=True):
        st = LinearStats()
        self.call("eg_linear_solve", form, _ptr(rhs), _ptr(x), _ptr(res), tol, int(refresh), ctypes.byref(st))
        return st

    def kernel_time(self, kernel, reps=50):
        """(mean ms per call, algorithmic bytes, algorithmic FP64 flops) of one
        operation timed with HIP events on the library stream."""
        ms, nbytes, flops = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.call("eg_kernel_time", kernel.encode(), reps, ctypes.byref(ms), ctypes.byref(nbytes), ctypes.byref(flops))
        return ms.value, nbytes.value, flops.value

    def ilu_order(self, form):
        """The ILU(0) factor's block row order (eg_ilu_order)."""
        import numpy as np
        count = self.n if form in (0, "dg") else 2 * self.n
        order = np.zeros(count, dtype=np.int32)
        self.call("eg_ilu_order", 1 if form in (1, "hdg") else 0, order.ctypes.data_as(_c_int_p))
        return order

    def run(self, form, prefix=None, increments=0):
        import numpy as np
        u = np.zeros(self.nu)
        st = RunStats()
        inc = np.zeros((increments, self.nu)) if increments else None
        kry = (ctypes.c_int * max(increments, 1))()
        self.call("eg_run", 1 if form in (1, "hdg") else 0, prefix.encode() if prefix else None, _dptr(u),
                  ctypes.byref(st), inc.ctypes.data_as(_c_double_p) if inc is not None else None, increments,
                  kry if increments else None)
        return u, st, inc, list(kry)[:increments]


def body_size(args: str) -> int:
    """nu = nx ny 4 (k+1)^2 of a command line (reference defaults nx = ny =
    100, k = 1, src/options.cpp)."""
    tok = args.split()
    val = {"--nx": 100, "--ny": 100, "--k": 1}
    for i, t in enumerate(tok[:-1]):
        if t in val:
            val[t] = int(tok[i + 1])
    return val["--nx"] * val["--ny"] * 4 * (val["--k"] + 1) ** 2


def _check(rc):
    if rc != 0:
        raise RuntimeError(load().eg_last_error(None).decode())


def solve_footprint(args: str, form) -> float:
    """Device bytes one solve of `args` (form "dg"/"hdg") holds at most."""
    out = ctypes.c_double()
    _check(load().eg_solve_footprint(args.encode(), 1 if form in (1, "hdg") else 0, ctypes.byref(out)))
    return out.value


def serve(requests, workers=8, hbm_budget=0.0, keep_states=False, batched=False, window_us=0.0):
    """Run independent solves concurrently (eg_serve, or eg_serve_batched
    when batched: the Arnoldi steps of concurrent condensed solves launched
    together). requests: list of (args, form, arrival_seconds). Returns
    (outcomes, report, states): one dict per request in request order, the
    serve report (with the grouping counts under "batch" when batched), and
    the final states (numpy arrays, or None) when keep_states."""
    import numpy as np
    lib = load()
    n = len(requests)
    args = (ctypes.c_char_p * n)(*[r[0].encode() for r in requests])
    forms = (ctypes.c_int * n)(*[1 if r[1] in (1, "hdg") else 0 for r in requests])
    arrivals = (ctypes.c_double * n)(*[float(r[2]) for r in requests])
    out = (SolveOutcome * n)()
    errors = ctypes.create_string_buffer(256 * max(n, 1))
    rep = ServeReport()
    states = None
    holders = []
    if keep_states:
        ptrs = (_c_double_p * n)()
        for i, r in enumerate(requests):
            a = np.zeros(body_size(r[0]))
            holders.append(a)
            ptrs[i] = a.ctypes.data_as(_c_double_p)
        states = ptrs
    brep = BatchReport()
    if batched:
        _check(lib.eg_serve_batched(args, forms, arrivals, n, workers, hbm_budget, window_us, out, states, errors,
                                    ctypes.byref(rep), ctypes.byref(brep)))
    else:
        _check(lib.eg_serve(args, forms, arrivals, n, workers, hbm_budget, out, states, errors, ctypes.byref(rep)))
    res = []
    for i in range(n):
        o = out[i]
        res.append({"arrival": o.arrival, "start": o.start, "finish": o.finish, "latency": o.finish - o.arrival,
                    "newton_seconds": o.newton_seconds, "footprint": o.footprint, "measured_peak": o.measured_peak,
                    "errors": list(o.errors), "status": o.status, "updates": o.updates,
                    "time_steps": o.time_steps, "worker": o.worker,
                    "error": errors.raw[256 * i:256 * (i + 1)].split(b"\0", 1)[0].decode()})
    report = {"wall": rep.wall, "max_reserved": rep.max_reserved, "max_concurrent": rep.max_concurrent}
    if batched:
        report["batch"] = {"flushes": brep.flushes, "chunks": brep.chunks, "graph_launches": brep.graph_launches,
                           "max_group": brep.max_group, "launch_s": brep.launch_s, "window_s": brep.window_s,
                           "tile": {q - 1: brep.tile[q] for q in range(1, 10) if brep.tile[q]}}
    return res, report, (holders if keep_states else None)


AUTOTUNE_SHAPES = ("levels-256", "levels-1024", "dataflow")


def autotune_ilu(ks, form="hdg", precision="fp64"):
    """Per-k ILU(0) launch shape measured on the device (eg_autotune_ilu):
    returns ({k: fastest shape}, {k: {shape: ms}})."""
    lib = load()
    ks = list(ks)
    arr = (ctypes.c_int * len(ks))(*ks)
    out = ctypes.create_string_buffer(16 * len(ks))
    times = (ctypes.c_double * (3 * len(ks)))()
    _check(lib.eg_autotune_ilu(arr, len(ks), 1 if form in (1, "hdg") else 0, precision.encode(), out, times))
    pick = {k: out.raw[16 * i:16 * (i + 1)].split(b"\0", 1)[0].decode() for i, k in enumerate(ks)}
    return pick, {k: dict(zip(AUTOTUNE_SHAPES, times[3 * i:3 * i + 3])) for i, k in enumerate(ks)}
