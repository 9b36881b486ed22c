"""Benchmark: mean seconds per accepted Newton update (DG and HDG) of the
implicit SSP-DIRK3 update on one MI355X.

Workload (plain FP64 GPU timing case): rotated simple wave, gamma=3,
alpha=45 deg, 16x16 Q3 cells, CFL 1 with dt = C h / a_max, 3 DIRK steps of
exactly 3 Newton updates per stage (36 accepted updates per form), matrix-free
(partial) Jacobian, block ILU(0)-preconditioned restarted GMRES (restart 20,
tolerance 1e-12), DG and HDG forms. One bench step = one full run of each form
(72 accepted updates). The state is synthetic: the simple-wave initial
projection computed by the library.

value: mean wall seconds per accepted Newton update over both forms, timed by
the library with the reference's update timer boundary (src/stepper.cpp:117,
226). The per-update error against the CPU oracle of the same workload is
stated as the reference's own acceptance test does (src/stepper.cpp:135-137,
197-198): relative error for the updates above the rounding-allowance floor,
absolute error relative to the state for the floor updates.

The same JSON line also carries, as separate workloads: the roofline of the
dominant operations at 128x128 Q6; the CPU baseline under the reference's
timing protocol (speed_test/benchmark.py:37-75); the mixed-precision form at
64x64 Q4 and 128x128 Q6 (3 DIRK steps x 3 updates) against plain FP64; the
batched form (256 seeded HDG mixed-precision instances, a documented
reject-and-redraw rule); and a seeded sample of the served request stream
(the full request distribution, admission by HBM footprint, p50/p99).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

WORKLOAD = "--nx 16 --ny 16 --k 3 --alpha-deg 45 --cfl 1 --steps 3 --newton-steps 3 --precond block-ilu0"
LARGE = "--nx 128 --ny 128 --k 6 --alpha-deg 30 --cfl 1 --restart 30 --precond block-ilu0"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (matrix and vector)
EPS = 2.220446049250313e-16
# Operations timed for the roofline entry; the setups and the condensation are
# dense FP64 block work (flop-bound), the rest streams vectors (HBM-bound).
KERNELS = ["pc_dg", "dg_apply", "pc_hdg", "schur_apply", "residual", "pc_setup_dg", "pc_setup_hdg", "hdg_condense"]
FLOP_BOUND = {"pc_setup_dg", "pc_setup_hdg", "hdg_condense"}


def counter_kib(counter: str, kernel: str, workload: str, names: tuple[str, ...]) -> float | None:
    """Per-call KiB of one rocprofv3 counter (FETCH_SIZE or WRITE_SIZE) summed
    over the dispatches of `kernel` run alone through --only-kernel, for the
    dispatches whose names start with one of `names` (the operation's own
    kernels). Returns None when the profiler is unavailable or fails."""
    import csv
    import shutil
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None
    reps = 5
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        cmd = ["timeout", "-s", "KILL", "90", "rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", tmp,
               "-o", "p", "--", sys.executable, str(ROOT / "bench.py"), "--only-kernel", kernel, "--reps",
               str(reps), "--workload", workload]
        env = dict(os.environ, TMPDIR="/tmp")
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        files = list(Path(tmp).rglob("*counter_collection.csv"))
        if r.returncode != 0 or not files:
            return None
        total = 0.0
        for row in csv.DictReader(open(files[0])):
            if row["Counter_Name"] == counter and any(n in row["Kernel_Name"] for n in names):
                total += float(row["Counter_Value"])
        # the hook makes one warm call before the reps timed ones, and its set-up
        # already calls the setups and the residual once; copy: reps calls
        calls = reps if kernel == "copy" else reps + (2 if kernel.startswith("pc_setup") or kernel == "residual" else 1)
        return total / calls


# kernels (name prefixes as the profiler reports them) making up each operation
OP_KERNELS = {
    "pc_setup_dg": ("k_dg_blocks", "k_ilu_factor", "k_block_inverse", "k_arm", "k_convert"),
    "pc_setup_hdg": ("k_hdg_fill", "k_hdg_diag_blocks", "k_ilu_factor", "k_block_inverse", "k_arm", "k_convert"),
    "hdg_condense": ("k_hdg_build", "k_hdg_solve", "k_hdg_panels", "k_hdg_face", "k_block_inverse"),
    "pc_dg": ("k_ilu_sweep", "k_ilu_lower", "k_ilu_upper", "k_block_gemv"),
    "pc_hdg": ("k_ilu_sweep", "k_ilu_lower", "k_ilu_upper", "k_block_gemv"),
    "dg_apply": ("k_dg_apply",), "schur_apply": ("k_schur_apply",), "residual": ("k_residual",),
    "copy": ("copyBuffer",),
}


def measured_traffic(kernel: str, workload: str) -> dict | None:
    """HBM bytes per call of `kernel` from FETCH_SIZE and WRITE_SIZE passes,
    each scaled by what the counter reads for a device copy of known size."""
    mib = 256.0
    raw = {c: counter_kib(c, kernel, workload, OP_KERNELS[kernel]) for c in ("FETCH_SIZE", "WRITE_SIZE")}
    cal = {c: counter_kib(c, "copy", workload, OP_KERNELS["copy"]) for c in ("FETCH_SIZE", "WRITE_SIZE")}
    if any(v is None for v in raw.values()) or any(not v for v in cal.values()):
        return None
    scale = {c: mib * 1024 / cal[c] for c in cal}  # copy reads and writes mib MiB
    read = raw["FETCH_SIZE"] * scale["FETCH_SIZE"] * 1024
    write = raw["WRITE_SIZE"] * scale["WRITE_SIZE"] * 1024
    return {"bytes": read + write, "read_bytes": read, "write_bytes": write,
            "fetch_scale": round(scale["FETCH_SIZE"], 4), "write_scale": round(scale["WRITE_SIZE"], 4)}


def roofline_entry(name, ms, nbytes, flops):
    """Achieved rate of one operation against the MI355X peak of its bound."""
    if name in FLOP_BOUND:
        achieved = flops / (ms * 1e-3) / 1e12
        return {"bound": "mfma", "kernel": name, "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 6), "traffic": None,
                "avg_launch_us": round(ms * 1e3, 3), "algorithmic_flops": flops}
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "avg_launch_us": round(ms * 1e3, 3),
            "algorithmic_bytes": nbytes}


# --------------------------------------------------------------- error statement
def floor_split_errors(metric, incs, refs, u_ref):
    """Per-update error of increments `incs` against `refs` in the CellNorm
    metric, as the reference's one-update acceptance test splits it
    (src/stepper.cpp:135-137,197-198): an update passes when ||du - du_ref||
    <= max(1e-8 ||du_ref||, 4 allowance), allowance = 128 eps (||u|| + ||old||
    + ||r||) ~ 256 eps ||u||. Updates with ||du_ref|| above 4 allowance / 1e-8
    are held to the relative bound: their largest relative error is reported
    (must be <= 1e-8); the rest are rounding-floor increments: their largest
    error relative to ||u|| is reported."""
    import numpy as np
    nrm = lambda v: float(np.sqrt(np.sum(metric * v * v)))
    u_norm = nrm(u_ref)
    floor = 4 * 128 * EPS * 2 * u_norm / 1e-8
    above, below, rel_all = [], [], []
    for a, b in zip(incs, refs):
        d, r = nrm(a - b), nrm(b)
        rel_all.append(d / max(r, 1e-300))
        (above if r > floor else below).append((d / max(r, 1e-300), d / u_norm))
    return {"updates": len(rel_all), "above_floor": len(above),
            "max_rel_err_above_floor": max((x[0] for x in above), default=None),
            "floor_updates": len(below), "max_err_rel_to_state_floor": max((x[1] for x in below), default=None),
            "median_rel_err": float(np.median(rel_all)) if rel_all else None,
            "floor_increment_norm": floor}


# --------------------------------------------------------------- CPU baseline
def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def cpu_baseline(workload, repeats=5):
    """The reference's timing protocol (speed_test/benchmark.py:37-75) on the
    CPU oracle CLI built for this host (-march=native, CMakeLists.txt:77-87):
    threads 1 (OMP/BLAS variables), pinned to one core, one warm-up run,
    DG/HDG order alternating per repeat, median over `repeats` of each form's
    mean_newton_s, the update count asserted."""
    import csv
    import tempfile
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "native", "-j8"], check=True, stdout=subprocess.DEVNULL)
    exe = ROOT / "oracle" / "_build_native" / "euler2d_oracle"
    env = dict(os.environ)
    for name in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "BLIS_NUM_THREADS",
                 "VECLIB_MAXIMUM_THREADS"):
        env[name] = "1"
    core = sorted(os.sched_getaffinity(0))[-1]
    values = {"dg": [], "hdg": []}
    expected = 3 * 3 * 4
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for repeat in range(-1, repeats):
            forms = ["dg", "hdg"] if repeat % 2 == 0 else ["hdg", "dg"]
            for form in forms:
                prefix = Path(tmp) / f"r{repeat}_{form}"
                cmd = [str(exe), "--form", form, *workload.split(), "--newton-tol", "1e-12", "--linear-tol", "1e-12",
                       "--timing-only", "--output", str(prefix)]
                subprocess.run(cmd, check=True, env=env, stdout=subprocess.DEVNULL,
                               preexec_fn=lambda: os.sched_setaffinity(0, {core}))
                row = next(csv.DictReader(open(f"{prefix}_{form}_summary.csv")))
                assert int(row["newton_updates"]) == expected, row["newton_updates"]
                if repeat >= 0:
                    values[form].append(float(row["mean_newton_s"]))
    dg, hdg = statistics.median(values["dg"]), statistics.median(values["hdg"])
    model, nproc, avail = cpu_info()
    return {"value": (dg + hdg) / 2, "unit": "s/update", "cores": 1, "kind": "port",
            "dg_s_per_update": dg, "hdg_s_per_update": hdg,
            "sample": f"the same workload ({workload}), reference timing protocol: 1 warm-up, {repeats} repeats "
                      f"alternating DG/HDG, median of mean_newton_s per form, value = mean of the two medians; "
                      f"serial C++ oracle restating the reference algorithm, -O3 -march=native, one thread "
                      f"pinned to core {core}",
            "cpu_model": model, "nproc": nproc, "cores_available": avail}


# --------------------------------------------------------------- batched and served
SHOCK_TIME = 1.0 / math.pi  # 4/((gamma+1) pi) at gamma = 3 (src/stepper.cpp:94-95)


def wave_dt(nx, ny, alpha_deg, cfl):
    """dt = CFL h / a_max of the simple wave (v + c = 2, HLL safety 1.1) on
    the rotated periodic rectangle [-1/cos a, 1/cos a] x [-1/sin a, 1/sin a]."""
    a = math.radians(alpha_deg)
    h = min(2 / math.cos(a) / nx, 2 / math.sin(a) / ny)
    return cfl * h / (1.1 * 2.0)


def batched_draws(seed=0xE1E12D):
    """Endless seeded stream of batched-form instances: nx, ny ~ U{8..48},
    k ~ U{1..6}, alpha ~ U(10, 80) deg, CFL ~ U[0.5, 2], HDG, 2 DIRK steps x
    3 Newton updates. Draws whose two steps would come within 10% of the shock
    time 1/pi are skipped (the step uses the frozen wave rate of the current
    state, which can be below the initial one)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    while True:
        nx, ny = (int(v) for v in rng.integers(8, 49, size=2))
        k = int(rng.integers(1, 7))
        alpha = float(rng.uniform(10, 80))
        cfl = float(rng.uniform(0.5, 2.0))
        if 2.2 * wave_dt(nx, ny, alpha, cfl) >= SHOCK_TIME:
            continue
        yield (f"--nx {nx} --ny {ny} --k {k} --alpha-deg {alpha:.6f} --cfl {cfl:.6f} --steps 2 "
               f"--newton-steps 3 --precond block-ilu0")


def batched_cases(count, screen, seed=0xE1E12D, rounds=6):
    """`count` instances from the seeded stream with the reject-and-redraw
    rule: an instance whose plain FP64 solve the reference algorithm rejects
    (its own error: an inadmissible state after a DIRK step with three fixed
    Newton updates, the shock-time guard, a failed line search) is replaced
    by the next draw of the same stream; `screen(cases)` runs the solves and
    returns the outcomes. Returns (cases, rejected draws with their reasons)."""
    gen = batched_draws(seed)
    cases = [next(gen) for _ in range(count)]
    rejected = []
    todo = list(range(count))
    for _ in range(rounds):
        res = screen([cases[i] for i in todo])
        bad = []
        for i, o in zip(todo, res):
            if o["status"] != 0:
                rejected.append((cases[i], o["error"]))
                cases[i] = next(gen)
                bad.append(i)
        todo = bad
        if not todo:
            break
    return cases, rejected


def run_batched(count, workers, seq_sample=32, group_workers=64):
    """The batched form: `count` independent HDG mixed-precision instances
    run concurrently (stream-ordered, admission by HBM footprint), against
    the same instances in plain FP64 concurrently (the precision gain) and a
    sample of them in plain FP64 one at a time (the batching gain); and the
    same mixed instances with their Arnoldi steps grouped across solves
    (eg_serve_batched: graph replays of the grouped kernels over all solves
    waiting at that moment, `group_workers` solves at once)."""
    import mfem_euler2d_cpp_amd as eg
    kof = lambda c: int(c.split("--k ")[1].split()[0])
    screen_walls = []

    def screen(cases):
        out, rep, _ = eg.serve([(c, "hdg", 0.) for c in cases], workers=workers)
        screen_walls.append(rep["wall"])
        return out

    cases, rejected = batched_cases(count, screen)
    out_c, rep_c, _ = eg.serve([(c, "hdg", 0.) for c in cases], workers=workers)
    out_m, rep_m, _ = eg.serve([(c + " --precision mixed", "hdg", 0.) for c in cases], workers=workers)
    out_g, rep_g, _ = eg.serve([(c + " --precision mixed", "hdg", 0.) for c in cases], workers=group_workers,
                               batched=True)
    both = [i for i in range(count) if out_c[i]["status"] == 0 and out_m[i]["status"] == 0
            and out_g[i]["status"] == 0]
    upd = lambda out: sum(out[i]["updates"] for i in both)
    rate_m, rate_c, rate_g = upd(out_m) / rep_m["wall"], upd(out_c) / rep_c["wall"], upd(out_g) / rep_g["wall"]
    err_g = [max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(out_g[i]["errors"], out_m[i]["errors"]))
             for i in both]
    sample = both[:seq_sample]
    out_s, rep_s, _ = eg.serve([(cases[i], "hdg", 0.) for i in sample], workers=1)
    rate_s = sum(o["updates"] for o in out_s if o["status"] == 0) / rep_s["wall"]
    err = [max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(out_m[i]["errors"], out_c[i]["errors"]))
           for i in both]
    return {"instances": count, "workers": workers, "completed": len(both),
            "rejected_and_redrawn": len(rejected), "rejected": [{"case": c, "error": e} for c, e in rejected],
            "updates": upd(out_m), "mixed_concurrent_updates_per_s": rate_m, "wall_s": rep_m["wall"],
            "fp64_concurrent_updates_per_s": rate_c,
            "precision_speedup_concurrent": rate_m / max(rate_c, 1e-30),
            "fp64_sequential_updates_per_s_sample": rate_s, "sequential_sample": len(sample),
            "mixed_concurrent_vs_fp64_sequential": rate_m / max(rate_s, 1e-30),
            "max_rel_diff_L2_error_mixed_vs_fp64": max(err) if err else None,
            "max_concurrent": rep_m["max_concurrent"],
            "grouped": {"workers": group_workers, "updates_per_s": rate_g, "wall_s": rep_g["wall"],
                        "vs_own_launches_concurrent": rate_g / max(rate_m, 1e-30),
                        "max_rel_diff_L2_error_vs_own_launches": max(err_g) if err_g else None,
                        "max_concurrent": rep_g["max_concurrent"], **rep_g["batch"]},
            "k_histogram": {k: sum(kof(cases[i]) == k for i in both) for k in range(1, 7)},
            "screening_wall_s": screen_walls}


def served_cases(count, seed=0x5E12ED):
    """The served request stream: nx = ny log-uniform in [8, 128], k ~
    U{0..8}, alpha ~ U(5, 85) deg, final time T ~ U[0.01, 0.2], CFL 1,
    Newton to convergence (tol 1e-12, at most 30), DG or HDG 50/50; the
    first `count` requests of the seeded stream."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(count):
        n = int(round(float(np.exp(rng.uniform(np.log(8), np.log(128))))))
        k = int(rng.integers(0, 9))
        alpha = float(rng.uniform(5, 85))
        T = float(rng.uniform(0.01, 0.2))
        form = "dg" if rng.uniform() < 0.5 else "hdg"
        cases.append((f"--nx {n} --ny {n} --k {k} --alpha-deg {alpha:.6f} --cfl 1 --T {T:.6f} "
                      f"--precond block-ilu0", form))
    return cases


def run_served(count, workers, load=0.8, seed=0x5E12ED):
    """The served form on the first `count` requests of the seeded stream:
    the requests are first run all at once to measure the capacity
    (requests/s); then a Poisson stream at `load` x capacity is served with
    admission by HBM footprint; p50/p99 latency and throughput."""
    import numpy as np
    import mfem_euler2d_cpp_amd as eg
    cases = served_cases(count)
    burst, rep0, _ = eg.serve([(a, f, 0.) for a, f in cases], workers=workers)
    capacity = sum(o["status"] == 0 for o in burst) / rep0["wall"]
    rate = load * capacity
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    out, rep, _ = eg.serve([(a, f, float(t)) for (a, f), t in zip(cases, arrivals)], workers=workers)
    ok = [o for o in out if o["status"] == 0]
    lat = np.array([o["latency"] for o in ok])
    ks = [int(a.split("--k ")[1].split()[0]) for a, _ in cases]
    ns = [int(a.split("--nx ")[1].split()[0]) for a, _ in cases]
    return {"requests": count, "of_stream": 2000, "workers": workers, "completed": len(ok),
            "failures": sorted({o["error"] for o in out if o["status"] != 0}),
            "rate_per_s": rate, "capacity_per_s": capacity, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "throughput_per_s": len(ok) / rep["wall"],
            "updates_per_s": sum(o["updates"] for o in ok) / rep["wall"], "max_concurrent": rep["max_concurrent"],
            "max_reserved_GB": rep["max_reserved"] / 1e9, "nx_max": max(ns), "k_max": max(ks),
            "updates": sum(o["updates"] for o in ok)}


# --------------------------------------------------------------- larger sizes
MIXED_CASES = {
    "64x64 Q4": "--nx 64 --ny 64 --k 4 --alpha-deg 30 --cfl 1 --steps 3 --newton-steps 3 --restart 30 "
                "--precond block-ilu0",
    "128x128 Q6": LARGE + " --steps 3 --newton-steps 3",
}


def run_mixed_sizes():
    """The mixed-precision form at 64x64 Q4 and 128x128 Q6 (3 DIRK steps x 3
    Newton updates per stage, 36 updates per form) against the plain FP64 form
    on the same cases: seconds per update and the per-update error of the
    mixed increments against the FP64 GPU ones (the FP64 GPU path is pinned
    against the CPU oracle at the smaller sizes)."""
    import numpy as np
    import mfem_euler2d_cpp_amd as eg
    res = {}
    for name, args in MIXED_CASES.items():
        entry = {}
        for form in ("dg", "hdg"):
            runs = {}
            metric = None
            for prec in ("fp64", "mixed"):
                c = eg.Context(args + f" --precision {prec}")
                u, st, inc, kr = c.run(form, increments=36)
                runs[prec] = (st.newton_seconds / st.updates, inc[:st.updates], list(kr[:st.updates]), u)
                if metric is None:
                    _, w, *_ = c.tables()
                    area, q = c.geometry()[2], c.q
                    node = np.array([area * (w[i % q] * w[i // q]) for i in range(q * q)])
                    metric = np.tile(np.repeat(node, 4), c.n)
                del c
            entry[form] = {"fp64_s_per_update": runs["fp64"][0], "mixed_s_per_update": runs["mixed"][0],
                           "speedup": runs["fp64"][0] / runs["mixed"][0],
                           "krylov_fp64_mean": float(np.mean(runs["fp64"][2])),
                           "krylov_mixed_mean": float(np.mean(runs["mixed"][2])),
                           "error_vs_fp64": floor_split_errors(metric, runs["mixed"][1], runs["fp64"][1],
                                                               runs["fp64"][3])}
        res[name] = entry
    return res


def large_roofline():
    """Roofline of the operations of one Newton update at 128x128 Q6 (FP64):
    live HIP-event timing per call, algorithmic bytes (HBM-bound operations)
    or LU-count FP64 flops (dense block work) per call."""
    import mfem_euler2d_cpp_amd as eg
    ctx = eg.Context(LARGE + " --steps 1")
    out = {}
    for k, reps in (("pc_setup_dg", 2), ("pc_dg", 5), ("dg_apply", 10), ("hdg_condense", 3), ("pc_setup_hdg", 3),
                    ("pc_hdg", 10), ("schur_apply", 10)):
        ms, nbytes, flops = ctx.kernel_time(k, reps=reps)
        out[k] = roofline_entry(k, ms, nbytes, flops)
    return {"workload": "simple wave 128x128 Q6, alpha 30, FP64, one call of each operation at the initial state",
            "kernels": out}


def only_kernel(args) -> int:
    """Time one operation alone (for profiler counter passes): `copy` is a
    device-to-device copy of --copy-mib MiB used to calibrate the HBM byte
    counters; any other name goes through the library's timing hook."""
    import torch
    import mfem_euler2d_cpp_amd as eg
    if args.only_kernel == "copy":
        n = int(args.copy_mib * (1 << 20)) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(args.reps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"kernel": "copy", "bytes_read": 8 * n, "bytes_written": 8 * n, "reps": args.reps}))
        return 0
    ctx = eg.Context(args.workload)
    ms, nbytes, flops = ctx.kernel_time(args.only_kernel, reps=args.reps)
    print(json.dumps({"kernel": args.only_kernel, "ms": ms, "algorithmic_bytes": nbytes, "flops": flops,
                      "reps": args.reps}))
    return 0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--kernel", default="auto", help="kernel for the roofline entry")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--only-kernel", default=None, help="time one operation alone (profiling) and exit")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--copy-mib", type=float, default=256.0)
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--batched", type=int, default=256, help="instances of the batched form (0: skip)")
    ap.add_argument("--served", type=int, default=12, help="requests of the served stream sample (0: skip)")
    ap.add_argument("--workers", type=int, default=12, help="concurrent solves of the batched/served forms")
    ap.add_argument("--extras-only", action="store_true", help="run only the batched/served forms")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 64x64 Q4 / 128x128 Q6 runs")
    ap.add_argument("--no-large-roofline", action="store_true", help="skip the 128x128 Q6 roofline")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: independent solves shard across GPUs without any exchange, "
              "and this build has no multi-GPU path; run with --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import mfem_euler2d_cpp_amd as eg
    if not eg.LIB.exists():
        eg.build(force=False)
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.only_kernel:
        return only_kernel(args)
    if args.extras_only:
        os.environ["EG_SERVE_PROGRESS"] = "1"
        print(json.dumps({"batched": run_batched(args.batched, args.workers) if args.batched else None,
                          "served": run_served(args.served, args.workers) if args.served else None}))
        return 0
    os.environ["EG_SERVE_PROGRESS"] = "1"  # the batched/served loops report each finished solve
    ctx = eg.Context(args.workload)

    log("headline workload: warm-up")
    for _ in range(args.warmup):
        for form in ("dg", "hdg"):
            ctx.run(form)

    newton = {"dg": 0.0, "hdg": 0.0}
    updates = {"dg": 0, "hdg": 0}
    krylov = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for form in ("dg", "hdg"):
            u, st, _, _ = ctx.run(form)
            newton[form] += st.newton_seconds
            updates[form] += st.updates
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dg_s = newton["dg"] / updates["dg"]
    hdg_s = newton["hdg"] / updates["hdg"]
    value = (newton["dg"] + newton["hdg"]) / (updates["dg"] + updates["hdg"])
    log(f"headline: {value * 1e3:.3f} ms/update")

    # per-update increments of one GPU run per form, for the error statement
    gpu_inc = {}
    for form in ("dg", "hdg"):
        u, st, inc, kry = ctx.run(form, increments=64)
        gpu_inc[form] = (u, inc[:st.updates])
        krylov[form] = kry[:st.updates]

    # roofline of the dominant operation (live HIP-event timing on the library
    # stream): the one with the largest time per accepted update, counting one
    # preconditioner setup (and, for HDG, one condensation) per update and one
    # preconditioner application and operator action per Krylov iteration
    # (+2 per update for the restart and the true residual).
    timed = {k: ctx.kernel_time(k, reps=20 if k in FLOP_BOUND else 50) for k in KERNELS}
    it_dg = float(np.mean(krylov["dg"])) + 2
    it_hdg = float(np.mean(krylov["hdg"])) + 2
    per_update = {"pc_dg": it_dg, "dg_apply": it_dg, "pc_hdg": it_hdg, "schur_apply": it_hdg, "residual": 2,
                  "pc_setup_dg": 1, "pc_setup_hdg": 1, "hdg_condense": 1}
    share = {k: timed[k][0] * per_update[k] for k in KERNELS}
    kname = max(share, key=share.get) if args.kernel == "auto" else args.kernel
    if kname not in timed:
        timed[kname] = ctx.kernel_time(kname, reps=50)
    ms, nbytes, flops = timed[kname]
    roofline = roofline_entry(kname, ms, nbytes, flops)
    if not args.no_traffic:
        log("traffic counter passes")
        t = measured_traffic(kname, args.workload)
        if t is not None:
            roofline["traffic"] = t["bytes"]
            roofline["traffic_detail"] = t
    roofline["note"] = ("16x16 Q3 is cache-resident and latency-bound (dependency levels); the target-size "
                        "roofline is roofline_128x128_q6")
    roofline["all_kernels"] = {k: {"us": round(v[0] * 1e3, 3), "us_per_update": round(share[k] * 1e3, 1),
                                   "GB/s": round(v[1] / (v[0] * 1e-3) / 1e9, 2),
                                   "TFLOP/s": round(v[2] / (v[0] * 1e-3) / 1e12, 4)} for k, v in timed.items()}
    del ctx

    err = {}
    cpu = None
    if not args.no_cpu:
        import oracle
        log("oracle increments (error statement)")
        o = oracle.Oracle(args.workload)
        metric, _ = o.metrics()
        for form in ("dg", "hdg"):
            uo, so, inco, _ = o.run(form, increments=64)
            ug, incg = gpu_inc[form]
            e = floor_split_errors(metric, incg, inco[:int(so[0])], uo)
            e["final_state_max_rel_err"] = float(np.max(np.abs(ug - uo)) / np.max(np.abs(uo)))
            err[form] = e
        log("CPU baseline (reference timing protocol)")
        cpu = cpu_baseline(args.workload)

    large = None
    if not args.no_large_roofline:
        log("128x128 Q6 roofline")
        large = large_roofline()
    mixed = None
    if not args.no_mixed:
        log("mixed precision 64x64 Q4 / 128x128 Q6")
        mixed = run_mixed_sizes()
    batched = None
    if args.batched:
        log(f"batched form: {args.batched} instances")
        batched = run_batched(args.batched, args.workers)
    served = None
    if args.served:
        log(f"served form: {args.served} requests")
        served = run_served(args.served, args.workers)

    line = {
        "metric": "mean s/accepted Newton update (DG,HDG)",
        "value": value,
        "unit": "s/update",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (simple-wave initial projection)",
        "config": {"workload": "simple wave 16x16 Q3, CFL 1, 3 DIRK3 steps x 3 Newton updates/stage, "
                               "partial assembly, block ILU(0), DG and HDG", "args": args.workload,
                   "updates_per_step": updates["dg"] // args.steps + updates["hdg"] // args.steps},
        "dg_s_per_update": dg_s,
        "hdg_s_per_update": hdg_s,
        "updates_per_s": 1.0 / value,
        "krylov_mean": {"dg": float(np.mean(krylov["dg"])), "hdg": float(np.mean(krylov["hdg"]))},
        "error_vs_oracle": err,
        "roofline": roofline,
        "roofline_128x128_q6": large,
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
