"""Batching parity: every scene of a mixed batch (one launch per kernel for all
scenes) gives bit-identical results to the same scene run alone, and both
equal the oracle."""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

STEPS = 3


def _mixed():
    scenes = [W.batch_scene(n, h, mu, seed) for seed, (n, h, mu) in enumerate(
        [(4096, 0.1, 0.0), (5000, 0.3, 0.2), (7777, 0.2, 0.05), (4500, 0.15, 0.1),
         (6000, 0.25, 0.02), (4096, 0.12, 0.19)])]
    scenes[2].table = 3001           # table size other than the particle count
    scenes[4].table = 2 * scenes[4].n
    return scenes


def test_mixed_batch_equals_scenes_alone_and_oracle():
    scenes = _mixed()
    announce("batch", len(scenes), [sc.n for sc in scenes])
    with S.Batch([sc.params for sc in scenes], [sc.n for sc in scenes],
                 [sc.table for sc in scenes]) as b:
        for s, sc in enumerate(scenes):
            b.upload(s, sc.pos, sc.vel)
        for _ in range(STEPS):
            b.step(W.DT, W.SUBSTEPS)
        got = [b.download(s) for s in range(len(scenes))]
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    for s, sc in enumerate(scenes):
        announce("alone", s)
        with S.Sim(sc.params, sc.n, sc.table) as sim:
            sim.upload(sc.pos, sc.vel)
            for _ in range(STEPS):
                sim.step(W.DT, W.SUBSTEPS)
            alone = sim.download()
        for a, g, what in zip(alone, got[s], ("pos", "vel", "dens")):
            assert_bitexact(g, a, f"scene {s} {what} batch vs alone")
        pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel, T=sc.table,
                                        substeps=STEPS * W.SUBSTEPS)
        assert_bitexact(got[s][0][:, :3], pos_o[:, :3], f"scene {s} pos vs oracle")
        assert_bitexact(got[s][1][:, :3], vel_o[:, :3], f"scene {s} vel vs oracle")
        assert_bitexact(got[s][2], dens_o, f"scene {s} dens vs oracle")


@pytest.mark.parametrize("w,l,nb", [(0, 0, 0), (1, 1, 0), (1, 0, 0), (0, 1, 0), (1, 0, 1)])
def test_batch_variants_bitexact(w, l, nb):
    scenes = _mixed()[:4]
    announce("batch variants", w, l, nb)
    with S.Batch([sc.params for sc in scenes], [sc.n for sc in scenes],
                 [sc.table for sc in scenes]) as b:
        v = {"work": w, "list": l, "block": 128, "rows": bool(nb)}
        b.variants = {"density": v, "pressure": v, "viscosity": v, "nbr": nb}
        assert b.variants["pressure"] == {"work": bool(w), "list": bool(l), "block": 128,
                                          "rows": bool(nb)}
        assert b.variants["nbr"] == bool(nb)
        for s, sc in enumerate(scenes):
            b.upload(s, sc.pos, sc.vel)
        b.step(W.DT, W.SUBSTEPS)
        got = [b.download(s) for s in range(len(scenes))]
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    for s, sc in enumerate(scenes):
        pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel, T=sc.table,
                                        substeps=W.SUBSTEPS)
        assert_bitexact(got[s][1][:, :3], vel_o[:, :3], f"scene {s} vel {v}")
        assert_bitexact(got[s][2], dens_o, f"scene {s} dens {v}")


def test_batch_autotune_keeps_state():
    scenes = _mixed()
    with S.Batch([sc.params for sc in scenes], [sc.n for sc in scenes],
                 [sc.table for sc in scenes]) as b:
        for s, sc in enumerate(scenes):
            b.upload(s, sc.pos, sc.vel)
        dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
        ms = b.autotune(dt, updates=2)
        assert np.all(ms > 0)
        for s, sc in enumerate(scenes):
            pos, vel, _ = b.download(s)
            assert_bitexact(pos, sc.pos, f"scene {s} pos after autotune")
            assert_bitexact(vel, sc.vel, f"scene {s} vel after autotune")


def test_many_scenes_grid_key():
    """80 small scenes under the cell-grid key: their bucket tables add up to
    more than 64 chunks of the count scan, so its look-back crosses several
    64-chunk windows; every scene equals the oracle."""
    rng = np.random.default_rng(11)
    scenes = [W.batch_scene(int(n), float(h), float(mu), 100 + s) for s, (n, h, mu) in enumerate(
        zip(rng.integers(200, 3000, 80), rng.uniform(0.08, 0.3, 80), rng.uniform(0.0, 0.2, 80)))]
    announce("many scenes", len(scenes), sum(sc.n for sc in scenes))
    with S.Batch([sc.params for sc in scenes], [sc.n for sc in scenes]) as b:
        for s, sc in enumerate(scenes):
            b.upload(s, sc.pos, sc.vel)
        # each scene's grid bucket count, from the server's byte accounting
        # (8 bytes per bucket): the count scan runs over more than 64 chunks
        tables = [1 + (S.request_bytes(sc.params, sc.n) - S.request_bytes(sc.params, sc.n, 1)) // 8
                  for sc in scenes]
        assert sum(tables) > 64 * 2048, sum(tables)
        for _ in range(2):
            b.step(W.DT, W.SUBSTEPS)
        got = [b.download(s) for s in range(len(scenes))]
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    for s, sc in enumerate(scenes):
        pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel, T=sc.n,
                                        substeps=2 * W.SUBSTEPS)
        assert_bitexact(got[s][0][:, :3], pos_o[:, :3], f"scene {s} pos")
        assert_bitexact(got[s][1][:, :3], vel_o[:, :3], f"scene {s} vel")
        assert_bitexact(got[s][2], dens_o, f"scene {s} dens")
