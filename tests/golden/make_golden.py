"""Writes tests/golden/cfg1.npz from the CPU oracle (run from the repository
root: python tests/golden/make_golden.py).

Contents, for the reference CPU case (8x8 Q2 simple wave, gamma 3, alpha 45
deg; src/problem.cpp:329-349, :350-384, :445-507, src/partial.cpp:31-135):
  u0       initial L2 projection (nu = 2304 doubles)
  r0       residual at u0 after Freeze(u0, dt=1e-4) with old = u0, dt = 1e-4
  jv       DG Jacobian action at u0 on v_i = sin(0.37 (i+1))
  rate     WaveRate after the freeze
  gl_k*    1-D GL tables on [0,1] for k = 0..8: nodes, weights, D[a*q+i] =
           l_i'(x_a), l_i(0), l_i(1) (src/problem.cpp:91-162)
The oracle these come from is pinned by tests/test_oracle.py (closed forms,
finite differences, brute force); the fixture freezes its outputs so that the
oracle and the GPU library are both checked against fixed numbers.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

ARGS = "--nx 8 --ny 8 --k 2 --alpha-deg 45 --steps 1 --dt 1e-4 --newton-steps 1 --precond block-jacobi"
DT = 1e-4


def compute():
    import oracle
    o = oracle.Oracle(ARGS)
    out = {}
    u0 = o.initial()
    rate = o.freeze(u0, DT)
    o.set_stage(u0, DT)
    out["u0"] = u0
    out["r0"] = o.residual(u0)
    out["rate"] = np.array([rate])
    o.dg_refresh(u0)
    v = np.sin(0.37 * (np.arange(o.nu) + 1))
    out["jv"] = o.dg_apply(v)
    for k in range(9):
        x, w, dd, e0, e1 = oracle.Oracle(f"--nx 2 --ny 2 --k {k}").tables()
        out[f"gl_k{k}_x"], out[f"gl_k{k}_w"], out[f"gl_k{k}_d"] = x, w, dd
        out[f"gl_k{k}_e0"], out[f"gl_k{k}_e1"] = e0, e1
    return out


if __name__ == "__main__":
    np.savez(Path(__file__).with_name("cfg1.npz"), **compute())
