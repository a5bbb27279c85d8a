"""The fused steady update -- one persistent launch per Update, grid barriers
between its stages (count scan, scatter, reorder, density, pressure,
viscosity+integrate) -- gives bit-identical results to the separate launches
of the same variants, for a mixed batch (waves spanning scene boundaries
included) and for single scenes, and the batch equals the oracle."""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

STEPS = 4


def _rows(fused, block=256):
    v = {"work": False, "list": False, "block": block, "rows": True}
    return {"density": dict(v), "pressure": dict(v), "viscosity": dict(v), "nbr": True, "fused": fused}


def _scenes():
    # uneven counts, most not a multiple of 64, so waves span scene boundaries
    spec = [(4096, 0.1, 0.0), (5003, 0.3, 0.2), (7777, 0.2, 0.05), (4501, 0.15, 0.1),
            (6007, 0.25, 0.02), (4099, 0.12, 0.19), (30011, 0.2, 0.1), (12345, 0.18, 0.07)]
    return [W.batch_scene(n, h, mu, 900 + seed) for seed, (n, h, mu) in enumerate(spec)]


def _run_batch(scenes, fused, steps=STEPS):
    with S.Batch([sc.params for sc in scenes], [sc.n for sc in scenes]) as b:
        for s, sc in enumerate(scenes):
            b.upload(s, sc.pos, sc.vel)
        b.variants = _rows(fused)
        assert b.variants["fused"] == fused
        for _ in range(steps):
            b.step(W.DT, W.SUBSTEPS)
        return [b.download(s) for s in range(len(scenes))]


def _run_sim(sc, fused, steps=STEPS):
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.variants = _rows(fused)
        assert sim.variants["fused"] == fused
        for _ in range(steps):
            sim.step(W.DT, W.SUBSTEPS)
        return sim.download()


def test_fused_batch_equals_separate_launches_and_oracle():
    scenes = _scenes()
    announce("fused batch", len(scenes), [sc.n for sc in scenes], "steps", STEPS)
    sep = _run_batch(scenes, False)
    fus = _run_batch(scenes, True)
    for s in range(len(scenes)):
        for a, g, what in zip(sep[s], fus[s], ("pos", "vel", "dens")):
            assert_bitexact(g, a, f"scene {s} {what} fused vs separate")
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    for s in (0, 3):
        sc = scenes[s]
        announce("oracle", s, sc.n)
        pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel,
                                        substeps=STEPS * W.SUBSTEPS)
        assert_bitexact(fus[s][0][:, :3], pos_o[:, :3], f"scene {s} pos vs oracle")
        assert_bitexact(fus[s][1][:, :3], vel_o[:, :3], f"scene {s} vel vs oracle")
        assert_bitexact(fus[s][2], dens_o, f"scene {s} dens vs oracle")


@pytest.mark.parametrize("name", ["dam_break_16k", "uniform_64k"])
def test_fused_single_scene_equals_separate_launches(name):
    sc = W.dam_break(n=16384) if name == "dam_break_16k" else W.uniform_lower(n=65536, box=(8.0, 2.0, 8.0))
    announce("fused single scene", name, sc.n)
    sep = _run_sim(sc, False)
    fus = _run_sim(sc, True)
    for a, g, what in zip(sep, fus, ("pos", "vel", "dens")):
        assert_bitexact(g, a, f"{name} {what} fused vs separate")


def test_fused_needs_the_row_mode_everywhere():
    sc = W.dam_break(n=4096)
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        v = _rows(True)
        v["pressure"]["rows"] = False
        with pytest.raises(S.SphError) as e:
            sim.variants = v
        assert e.value.code == S.SPH_ERR_INVALID


def test_autotune_reports_a_fused_choice():
    sc = W.dam_break(n=16384)
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.step(W.DT, W.SUBSTEPS)
        sim.autotune(float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS))), updates=3)
        v = sim.variants
        announce("autotuned", v)
        assert isinstance(v["fused"], bool)
        sim.step(W.DT, W.SUBSTEPS)
