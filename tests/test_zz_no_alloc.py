"""The allocation-free norm eg_cell_norm_noalloc (partial sums in a fixed
per-device array, result in device memory) against eg_cell_norm of a context
on the same vectors: three calls in a row, each on a smaller vector than the
one before, so later calls run with partial sums of earlier, larger calls
left in the shared array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def cell_metric(c):
    """The CellNorm weights area * (w_ix w_iy) of every node and component
    (src/problem.cpp:211-215), from the context's GL weights and geometry."""
    _, w, *_ = c.tables()
    area = c.geometry()[2]
    q = c.q
    node = np.array([area * (w[i % q] * w[i // q]) for i in range(q * q)])
    return np.tile(np.repeat(node, 4), c.n)


def test_cell_norm_noalloc_matches_cell_norm(gpu):
    import torch
    lib = gpu.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    for args in ("--nx 64 --ny 64 --k 4 --alpha-deg 30", "--nx 16 --ny 16 --k 3", "--nx 8 --ny 8 --k 2"):
        print(f"case: {args}", flush=True)
        c = gpu.Context(args)
        u = c.initial()
        v = torch.tensor(u * (1 + 0.1 * np.sin(0.37 * np.arange(c.nu))), device="cuda")
        metric = torch.tensor(cell_metric(c), device="cuda")
        assert metric.numel() == v.numel() == c.nu
        ref = c.cell_norm(v)
        torch.cuda.synchronize()
        rc = lib.eg_cell_norm_noalloc(gpu._ptr(v), gpu._ptr(metric), c.nu, stream, gpu._ptr(out))
        assert rc == 0, lib.eg_last_error(None).decode()
        torch.cuda.synchronize()
        got = out.item()
        # same partial sums and tree as eg_cell_norm; the square root taken
        # on the device instead of the host (both correctly rounded)
        assert abs(got - ref) <= 4e-16 * ref, (args, got, ref)
        del c


def test_cell_norm_noalloc_rejects_bad_sizes(gpu):
    import torch
    lib = gpu.load()
    v = torch.ones(8, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    for n in (0, -5, 1 << 31):
        assert lib.eg_cell_norm_noalloc(gpu._ptr(v), gpu._ptr(v), n, None, gpu._ptr(out)) == 1
        assert "n must be in" in lib.eg_last_error(None).decode()
    assert lib.eg_cell_norm_noalloc(None, gpu._ptr(v), 8, None, gpu._ptr(out)) == 1
