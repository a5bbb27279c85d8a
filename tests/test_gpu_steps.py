"""Steps of several substeps.  After the first update of a run, a step's
updates are steady (buckets recorded by the previous update, same dt) and the
engine launches them as one captured graph; while profiling, only the first
update of each step carries the stage events.  Neither may change a result:
  - a plain run, a profiled run and a run of single updates give the same
    bits, and equal the oracle;
  - profiling over K steps reports K profiled updates, a zero predict stage
    (the previous update recorded the buckets) and positive kernel stages.
"""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

N = 4096
STEPS = 4


def _scene():
    return W.dam_break(n=N, seed=3, box=(2.6, 2.6, 2.6))


def test_graph_steps_profiled_steps_and_single_updates_agree():
    sc = _scene()
    announce("steps", sc.name, N, STEPS)
    sub = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    results = []
    for mode in ("steps", "profiled", "updates"):
        sim = S.Sim(sc.params, sc.n, S.TABLE_GRID)
        try:
            sim.upload(sc.pos, sc.vel)
            if mode == "profiled":
                sim.profile_begin(STEPS * W.SUBSTEPS)
            for _ in range(STEPS):
                if mode == "updates":
                    for _ in range(W.SUBSTEPS):
                        sim.step(sub, 1)
                else:
                    sim.step_async(W.DT, W.SUBSTEPS)
            if mode == "profiled":
                ms, used = sim.profile_end()
                # the first update of each step is profiled
                assert used == STEPS, used
                assert all(ms[i] > 0.0 for i in (1, 2, 3, 4)), ms
            sim.sync()
            results.append(sim.download())
        finally:
            sim.close()
    for mode, got in zip(("profiled", "updates"), results[1:]):
        for i, what in enumerate(("pos", "vel", "dens")):
            assert_bitexact(got[i], results[0][i], f"{mode} {what}")
    pos_o, vel_o, dens_o = O.update(orc_params(sc.params), sub, sc.pos, sc.vel, T=sc.n,
                                    substeps=STEPS * W.SUBSTEPS)
    assert_bitexact(results[0][0][:, :3], pos_o[:, :3], "pos")
    assert_bitexact(results[0][1][:, :3], vel_o[:, :3], "vel")
    assert_bitexact(results[0][2], dens_o, "dens")


def test_profiled_steady_steps_have_no_predict_stage():
    sc = _scene()
    announce("profile predict stage", sc.name, N)
    sim = S.Sim(sc.params, sc.n, S.TABLE_GRID)
    try:
        sim.upload(sc.pos, sc.vel)
        sim.step(W.DT, W.SUBSTEPS)  # the first update predicts on its own
        sim.profile_begin(3 * W.SUBSTEPS)
        for _ in range(3):
            sim.step_async(W.DT, W.SUBSTEPS)
        ms, used = sim.profile_end()
    finally:
        sim.close()
    assert used == 3, used
    assert ms[0] == 0.0, ms
