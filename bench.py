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
measured in the same run (the oracle run is also the CPU baseline).

The same JSON line also carries, as separate workloads: the mixed-precision
form at 64x64 Q4 and 128x128 Q6 against plain FP64 ("mixed_precision"), a
seeded sample of the batched form (independent HDG mixed-precision instances
on concurrent streams, against the same instances in plain FP64 one at a
time and concurrently: "batched"), and a seeded Poisson stream of solve
requests with admission by HBM footprint ("served": p50/p99 latency,
throughput).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

WORKLOAD = "--nx 16 --ny 16 --k 3 --alpha-deg 45 --cfl 1 --steps 3 --newton-steps 3 --precond block-ilu0"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (matrix and vector)
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
    import subprocess
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
        total, calls = 0.0, 0
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


SHOCK_TIME = 1.0 / 3.141592653589793  # 4/((gamma+1) pi) at gamma = 3 (src/stepper.cpp:94-95)


def wave_dt(nx, ny, alpha_deg, cfl):
    """dt = CFL h / a_max of the simple wave (v + c = 2, HLL safety 1.1) on
    the rotated periodic rectangle [-1/cos a, 1/cos a] x [-1/sin a, 1/sin a]."""
    import math
    a = math.radians(alpha_deg)
    h = min(2 / math.cos(a) / nx, 2 / math.sin(a) / ny)
    return cfl * h / (1.1 * 2.0)


def batched_cases(count, seed=0xE1E12D):
    """Independent HDG mixed-precision instances: nx, ny ~ U{8..48}, k ~
    U{1..6}, alpha ~ U(10, 80) deg, CFL ~ U[0.5, 2], 2 DIRK steps x 3 Newton
    updates; draws whose two steps would come within 10% of the shock time
    are redrawn (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cases = []
    while len(cases) < count:
        nx, ny = (int(v) for v in rng.integers(8, 49, size=2))
        k = int(rng.integers(1, 7))
        alpha = float(rng.uniform(10, 80))
        cfl = float(rng.uniform(0.5, 2.0))
        # the step is set from the frozen wave rate of the current state, which
        # can be below the initial one: a 10% margin keeps both steps short
        # of the shock time
        if 2.2 * wave_dt(nx, ny, alpha, cfl) >= SHOCK_TIME:
            continue
        cases.append(f"--nx {nx} --ny {ny} --k {k} --alpha-deg {alpha:.6f} --cfl {cfl:.6f} --steps 2 "
                     f"--newton-steps 3 --precond block-ilu0")
    return cases


def served_cases(count, seed=0x5E12ED, max_n=32, max_k=6):
    """Solve requests: nx = ny log-uniform in [8, max_n], k ~ U{0..max_k},
    alpha ~ U(5, 85) deg, final time T ~ U[0.01, 0.05], CFL 1, Newton to
    convergence (tol 1e-12, at most 30), DG or HDG 50/50."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(count):
        n = int(round(float(np.exp(rng.uniform(np.log(8), np.log(max_n))))))
        k = int(rng.integers(0, max_k + 1))
        alpha = float(rng.uniform(5, 85))
        T = float(rng.uniform(0.01, 0.05))
        form = "dg" if rng.uniform() < 0.5 else "hdg"
        cases.append((f"--nx {n} --ny {n} --k {k} --alpha-deg {alpha:.6f} --cfl 1 --T {T:.6f} "
                      f"--precond block-ilu0", form))
    return cases


def run_batched(count, workers):
    """The batched form (many independent HDG mixed-precision instances on
    concurrent streams) against the same instances run one at a time in plain
    FP64."""
    import numpy as np
    import mfem_euler2d_cpp_amd as eg
    cases = batched_cases(count)
    kof = lambda c: int(c.split("--k ")[1].split()[0])
    # concurrent solves run the ILU one launch per level; the autotuner picks
    # the sweep's threads per block row for each k
    _, times = eg.autotune_ilu(sorted({kof(c) for c in cases}), form="hdg", precision="mixed")
    threads = {k: (256 if t["levels-256"] <= t["levels-1024"] else 1024) for k, t in times.items()}
    mixed = [(c + f" --precision mixed --sweep-threads {threads[kof(c)]}", "hdg", 0.) for c in cases]
    out_b, rep_b, _ = eg.serve(mixed, workers=workers)
    plain = [(c, "hdg", 0.) for c in cases]
    out_c, rep_c, _ = eg.serve(plain, workers=workers)
    out_p, rep_p, _ = eg.serve(plain, workers=1)
    both = [i for i in range(count) if out_b[i]["status"] == 0 and out_p[i]["status"] == 0]
    upd = lambda out: sum(out[i]["updates"] for i in both)
    rate = lambda out, rep: upd(out) / rep["wall"]
    err = [max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(out_b[i]["errors"], out_p[i]["errors"]))
           for i in both]
    return {"instances": count, "workers": workers, "completed": len(both),
            "failed": {"mixed": sum(o["status"] != 0 for o in out_b), "plain": sum(o["status"] != 0 for o in out_p)},
            "failure_reasons": sorted({o["error"] for o in out_b + out_p if o["status"] != 0}),
            "failed_cases": [cases[i] for i in range(count) if out_b[i]["status"] != 0 or out_p[i]["status"] != 0],
            "updates": upd(out_b), "updates_per_s": rate(out_b, rep_b), "wall_s": rep_b["wall"],
            "plain_fp64_sequential_updates_per_s": rate(out_p, rep_p),
            "fp64_concurrent_updates_per_s": rate(out_c, rep_c),
            "speedup_vs_plain": rate(out_b, rep_b) / max(rate(out_p, rep_p), 1e-30),
            "max_rel_diff_L2_error_vs_fp64": max(err) if err else None,
            "max_concurrent": rep_b["max_concurrent"], "sweep_threads_per_k": threads,
            "autotune_ms": {k: {n: round(v, 3) for n, v in t.items()} for k, t in times.items()}}


def run_served(count, workers, load=0.8, seed=0x5E12ED):
    """The served form: the requests are first run all at once to measure the
    capacity (requests/s); then a Poisson stream at `load` x capacity is
    served with admission by HBM footprint; p50/p99 latency and throughput."""
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
    return {"requests": count, "workers": workers, "completed": len(ok), "rate_per_s": rate,
            "capacity_per_s": capacity, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "throughput_per_s": len(ok) / rep["wall"],
            "updates_per_s": sum(o["updates"] for o in ok) / rep["wall"], "max_concurrent": rep["max_concurrent"],
            "max_reserved_GB": rep["max_reserved"] / 1e9}


MIXED_CASES = {
    "64x64 Q4": "--nx 64 --ny 64 --k 4 --alpha-deg 30 --cfl 1 --steps 3 --newton-steps 3 --restart 30 --precond block-ilu0",
    "128x128 Q6": "--nx 128 --ny 128 --k 6 --alpha-deg 30 --cfl 1 --steps 3 --newton-steps 3 --restart 30 "
                  "--precond block-ilu0",
}


def run_mixed_sizes():
    """The mixed-precision form at 64x64 Q4 and 128x128 Q6 (one DIRK step x
    one Newton update per stage, 4 updates per form) against the plain FP64
    form on the same cases: seconds per update and the per-update relative
    difference of the increments (2-norm)."""
    import numpy as np
    import mfem_euler2d_cpp_amd as eg
    res = {}
    for name, args in MIXED_CASES.items():
        entry = {}
        for form in ("dg", "hdg"):
            runs = {}
            for prec in ("fp64", "mixed"):
                c = eg.Context(args + f" --precision {prec}")
                u, st, inc, kr = c.run(form, increments=36)
                runs[prec] = (st.newton_seconds / st.updates, inc[:st.updates], list(kr[:st.updates]))
                del c
            d = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(runs["mixed"][1], runs["fp64"][1])]
            entry[form] = {"fp64_s_per_update": runs["fp64"][0], "mixed_s_per_update": runs["mixed"][0],
                           "speedup": runs["fp64"][0] / runs["mixed"][0], "krylov_fp64": runs["fp64"][2],
                           "krylov_mixed": runs["mixed"][2], "max_update_rel_diff_vs_fp64": max(d)}
        res[name] = entry
    return res


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
    ap.add_argument("--batched", type=int, default=32, help="instances of the batched form (0: skip)")
    ap.add_argument("--served", type=int, default=64, help="requests of the served form (0: skip)")
    ap.add_argument("--workers", type=int, default=12, help="concurrent solves of the batched/served forms")
    ap.add_argument("--extras-only", action="store_true", help="run only the batched/served forms")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision 64x64 Q4 / 128x128 Q6 runs")
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
        print(json.dumps({"batched": run_batched(args.batched, args.workers) if args.batched else None,
                          "served": run_served(args.served, args.workers) if args.served else None}))
        return 0
    ctx = eg.Context(args.workload)

    # warm-up (also JIT/first-touch of every kernel)
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
    if kname in FLOP_BOUND:
        achieved = flops / (ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "kernel": kname, "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 6), "traffic": None,
                    "avg_launch_us": round(ms * 1e3, 3), "algorithmic_flops": flops}
    else:
        achieved = nbytes / (ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                    "avg_launch_us": round(ms * 1e3, 3), "algorithmic_bytes": nbytes}
    if not args.no_traffic:
        t = measured_traffic(kname, args.workload)
        if t is not None:
            roofline["traffic"] = t["bytes"]
            roofline["traffic_detail"] = t
    roofline["all_kernels"] = {k: {"us": round(v[0] * 1e3, 3), "us_per_update": round(share[k] * 1e3, 1),
                                   "GB/s": round(v[1] / (v[0] * 1e-3) / 1e9, 2),
                                   "TFLOP/s": round(v[2] / (v[0] * 1e-3) / 1e12, 4)} for k, v in timed.items()}

    cpu = None
    err = {}
    if not args.no_cpu:
        import oracle
        o = oracle.Oracle(args.workload)
        metric, _ = o.metrics()
        nrm = lambda v: float(np.sqrt(np.sum(metric * v * v)))
        cpu_s, cpu_upd = 0.0, 0
        for form in ("dg", "hdg"):
            uo, so, inco, _ = o.run(form, increments=64)
            cpu_s += so[2]
            cpu_upd += int(so[0])
            ug, incg = gpu_inc[form]
            rel = [nrm(incg[i] - inco[i]) / nrm(inco[i]) for i in range(int(so[0]))]
            err[form] = {"max_update_rel_err": max(rel), "median_update_rel_err": float(np.median(rel)),
                         "final_state_max_rel_err": float(np.max(np.abs(ug - uo)) / np.max(np.abs(uo)))}
        cpu = {"value": cpu_s / cpu_upd, "unit": "s/update", "cores": 1, "kind": "port",
               "sample": f"the same workload ({WORKLOAD}), one run per form ({cpu_upd} updates), "
                         "serial C++ oracle restating the reference algorithm, -O3 -march=x86-64-v3"}

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
        "cpu_baseline": cpu,
        "mixed_precision": run_mixed_sizes() if not args.no_mixed else None,
        "batched": run_batched(args.batched, args.workers) if args.batched else None,
        "served": run_served(args.served, args.workers) if args.served else None,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
