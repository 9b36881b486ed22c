"""CPU checks of bench.py's workload generators: the batched form's seeded
reject-and-redraw rule and the served stream's distribution (the requests
themselves run on the GPU in tests/test_serve.py and in bench.py)."""
import math

import bench


def k_of(case):
    return int(case.split("--k ")[1].split()[0])


def test_batched_redraw_rule_is_seeded_and_replaces_rejections():
    calls = []

    def screen(cases):  # a stand-in that rejects every k = 1 draw
        calls.append(list(cases))
        return [{"status": int(k_of(c) == 1), "error": "SSP-DIRK3 step produced an inadmissible state"}
                for c in cases]

    cases, rejected = bench.batched_cases(24, screen)
    again, rejected2 = bench.batched_cases(24, screen)
    assert cases == again and rejected == rejected2  # deterministic for the seed
    assert len(cases) == 24 and all(k_of(c) != 1 for c in cases)
    assert rejected and all(k_of(c) == 1 for c, _ in rejected)
    # the replacements are the next draws of the same stream, in order
    stream = bench.batched_draws()
    first = [next(stream) for _ in range(24 + len(rejected))]
    assert set(cases) <= set(first)
    # second screening round only re-runs the replaced instances
    assert len(calls[1]) == len(rejected) or len(calls) == 2


def test_batched_draws_respect_ranges_and_shock_margin():
    gen = bench.batched_draws()
    for _ in range(200):
        c = next(gen)
        f = c.split()
        val = lambda name: float(f[f.index(name) + 1])
        nx, ny, k, a, cfl = val("--nx"), val("--ny"), val("--k"), val("--alpha-deg"), val("--cfl")
        assert 8 <= nx <= 48 and 8 <= ny <= 48 and 1 <= k <= 6 and 10 <= a <= 80 and .5 <= cfl <= 2
        assert 2.2 * bench.wave_dt(nx, ny, a, cfl) < bench.SHOCK_TIME
        assert "--steps 2 --newton-steps 3" in c


def test_served_stream_covers_the_stated_distribution():
    cases = bench.served_cases(2000)
    ns = [int(a.split("--nx ")[1].split()[0]) for a, _ in cases]
    ks = [k_of(a) for a, _ in cases]
    ts = [float(a.split("--T ")[1].split()[0]) for a, _ in cases]
    assert min(ns) >= 8 and max(ns) <= 128 and max(ns) >= 120
    assert set(ks) == set(range(9))
    assert min(ts) >= .01 and max(ts) <= .2 < 1 / math.pi
    forms = [f for _, f in cases]
    assert 900 < forms.count("dg") < 1100
    assert bench.served_cases(24) == cases[:24]  # the bench sample is the stream's start
