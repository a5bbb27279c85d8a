"""The benchmark's derived figures, on the host: the whole-step roofline
(264 + 4 T/N bytes per particle-substep, the reference's stage traffic) and
the split of a serving leg's step-latency spread from the server's tick log."""
import importlib.util
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_step_roofline_bytes_and_fraction(bench):
    # T = N: 268 B per particle-substep; T = 2N (the 1M scene's 2^21 table): 272 B
    r = bench.step_roofline(1.0e9, 3, 1.0)
    assert r["alg_bytes_per_particle_substep"] == 268.0
    assert r["achieved"] == pytest.approx(268.0 * 3.0e9 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert bench.step_roofline(1.0e9, 3, 2.0)["alg_bytes_per_particle_substep"] == 272.0


def test_tick_spread_separates_load_from_jitter(bench):
    # two memberships: 20 ticks of a small load at 1 ms (one at 1.5 ms), then
    # 20 ticks of a large load at 10 ms; the overall p99/p50 is the load's
    ms = np.array([1.0] * 19 + [1.5] + [10.0] * 20, np.float32)
    parts = np.array([100_000] * 20 + [5_000_000] * 20, np.uint64)
    layout = np.array([0] * 20 + [1] * 20, np.uint32)
    s = bench.tick_spread(ms, parts, layout)
    assert s["ticks"] == 40 and s["layouts"] == 2
    assert s["within_membership"]["p50"] == pytest.approx(1.0)
    assert 1.0 < s["within_membership"]["p99"] <= 1.5
    # consecutive ticks are compared only inside one membership: 38 pairs
    assert s["tick_to_tick"]["pairs"] == 38
    assert s["tick_to_tick"]["p50"] == pytest.approx(1.0)
    loads = {tuple(b["particles"]): b for b in s["by_load"]}
    small = loads[(0, 1 << 18)]
    large = loads[(1 << 22, 1 << 24)]
    assert small["ticks"] == 20 and large["ticks"] == 20
    assert large["step_ms_p50"] == pytest.approx(30.0)  # a step is three ticks
    assert small["step_ms_p50"] == pytest.approx(3.0)


def test_tick_spread_skips_untimed_ticks(bench):
    ms = np.array([1.0, 0.0, 1.1, 2.0, 2.2], np.float32)
    s = bench.tick_spread(ms, np.full(5, 1000, np.uint64), np.array([0, 0, 0, 1, 1], np.uint32))
    assert s["tick_to_tick"]["pairs"] == 1
    assert s["tick_to_tick"]["p50"] == pytest.approx(1.1)
    assert bench.tick_spread(np.zeros(0, np.float32), np.zeros(0, np.uint64), np.zeros(0, np.uint32)) is None
