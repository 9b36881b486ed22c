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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--kernel", default="auto", help="kernel for the roofline entry")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--workload", default=WORKLOAD)
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

    # roofline of the dominant kernel (live HIP-event timing on the library stream)
    kernels = ["pc_dg", "dg_apply", "schur_apply", "pc_hdg", "residual"]
    timed = {}
    for k in kernels:
        ms, nbytes = ctx.kernel_time(k, reps=50)
        timed[k] = (ms, nbytes)
    # dominant = largest per-update time share: launches per update estimated
    # from the Krylov counts (one PC + one apply per iteration)
    it_dg = float(np.mean(krylov["dg"])) + 2
    it_hdg = float(np.mean(krylov["hdg"])) + 2
    share = {"pc_dg": timed["pc_dg"][0] * it_dg, "dg_apply": timed["dg_apply"][0] * it_dg,
             "pc_hdg": timed["pc_hdg"][0] * it_hdg, "schur_apply": timed["schur_apply"][0] * it_hdg,
             "residual": timed["residual"][0] * 2}
    kname = max(share, key=share.get) if args.kernel == "auto" else args.kernel
    if kname not in timed:
        timed[kname] = ctx.kernel_time(kname, reps=50)
    ms, nbytes = timed[kname]
    achieved = nbytes / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "avg_launch_us": round(ms * 1e3, 3), "algorithmic_bytes": nbytes,
                "all_kernels_us": {k: round(v[0] * 1e3, 3) for k, v in timed.items()}}

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
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
