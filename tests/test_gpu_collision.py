"""A fast collision under the cell-grid key: two 32,768-particle blocks meeting
at +-16 m/s in a 6 x 3 x 6 box compress the fluid within a few steps, so a
large share of the neighbour-mask records overflow (row runs past 64 slots,
streams past 512 bits) while the rest use every word of the stream.  Those
particles walk again in pressure and viscosity, staging their rows over the
wave's finished mask table.  The rollout is bit-exact against the oracle."""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

N = 65536
STEPS = 15


def _sub_dt():
    return float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))


def _overflow_share(pos, h):
    """Share of particles whose 27-cell neighbourhood has a row of three
    z-cells holding more than 64 particles or more than 512 in all (the
    record's limits), from the cell counts of the given positions."""
    c = np.floor(pos[:, :3] / h).astype(np.int64)
    c -= c.min(0) - 1
    dims = c.max(0) + 2
    cnt = np.bincount((c[:, 0] * dims[1] + c[:, 1]) * dims[2] + c[:, 2],
                      minlength=int(dims.prod())).reshape(dims)
    tot = np.zeros(len(pos), np.int64)
    longest = np.zeros(len(pos), np.int64)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            run = sum(cnt[c[:, 0] + dx, c[:, 1] + dy, c[:, 2] + dz] for dz in (-1, 0, 1))
            tot += run
            longest = np.maximum(longest, run)
    return float(((longest > 64) | (tot > 512)).mean())


@pytest.mark.parametrize("blocks", [None, (128, 64, 256), "fused"])
def test_colliding_blocks_rollout_bitexact(blocks):
    plan = W.two_blocks_plan(n=N, box=(6.0, 3.0, 6.0), speed=16.0)
    sc = plan["scene"]
    kind, seed, box, spacing, jitter, speed = plan["init"]
    pos, vel = S.generate(kind, seed, N, box, spacing, jitter, speed)
    announce("colliding blocks", N, "steps", STEPS, "blocks", blocks)
    with S.Sim(sc.params, N, S.TABLE_GRID) as sim:
        if blocks is not None:
            v = sim.variants
            fused = blocks == "fused"  # one persistent launch per steady Update
            for k, b in zip(("density", "pressure", "viscosity"), (256, 256, 256) if fused else blocks):
                v[k] = {"work": False, "list": False, "block": b, "rows": True}
            v["nbr"] = True
            v["fused"] = fused
            sim.variants = v
            assert sim.variants["fused"] == fused
        assert sim.variants["nbr"] and sim.variants["pressure"]["rows"]
        sim.upload(pos, vel)
        for _ in range(STEPS):
            sim.step_async(W.DT, W.SUBSTEPS)
        sim.sync()
        pos_g, vel_g, dens_g = sim.download()
    O.set_threads(16)
    pos_o, vel_o, dens_o = O.update(orc_params(sc.params), _sub_dt(), pos, vel, T=N,
                                    substeps=STEPS * W.SUBSTEPS)
    assert_bitexact(pos_g[:, :3], pos_o[:, :3], "positions after the collision")
    assert_bitexact(vel_g[:, :3], vel_o[:, :3], "velocities after the collision")
    assert_bitexact(dens_g, dens_o, "densities after the collision")
    # the case covers overflowing records (about 45% of the particles here)
    assert _overflow_share(pos_o, sc.params.smoothing_radius) > 0.2


def test_colliding_blocks_in_a_batch_bitexact():
    """The same collision as one scene of a multi-scene batch, beside a dam
    break and a uniform scene: the multi-scene kernels take each wave's
    scenes in turn, and the overflowing records' walks stage their rows
    inside that loop."""
    plan = W.two_blocks_plan(n=N, box=(6.0, 3.0, 6.0), speed=16.0)
    kind, seed, box, spacing, jitter, speed = plan["init"]
    pos, vel = S.generate(kind, seed, N, box, spacing, jitter, speed)
    scenes = [(plan["scene"].params, pos, vel),
              (lambda d: (d.params, d.pos, d.vel))(W.dam_break(n=16384)),
              (lambda u: (u.params, u.pos, u.vel))(W.uniform_lower(n=32768, table=65536, box=(4.0, 2.0, 4.0)))]
    steps = 10
    announce("colliding blocks in a batch", [len(p) for _, p, _ in scenes], "steps", steps)
    b = S.Batch([p for p, _, _ in scenes], [len(x) for _, x, _ in scenes])
    try:
        v = b.variants
        for k, blk in zip(("density", "pressure", "viscosity"), (256, 256, 128)):
            v[k] = {"work": False, "list": False, "block": blk, "rows": True}
        v["nbr"] = True
        b.variants = v
        assert b.variants["nbr"] and b.variants["viscosity"]["rows"]
        for q, (_, x, u) in enumerate(scenes):
            b.upload(q, x, u)
        for _ in range(steps):
            b.step_async(W.DT, W.SUBSTEPS)
        b.sync()
        got = [b.download(q) for q in range(len(scenes))]
    finally:
        b.close()
    O.set_threads(16)
    for q, (p, x, u) in enumerate(scenes):
        pos_o, vel_o, dens_o = O.update(orc_params(p), _sub_dt(), x, u, T=len(x), substeps=steps * W.SUBSTEPS)
        assert_bitexact(got[q][0][:, :3], pos_o[:, :3], f"scene {q} positions")
        assert_bitexact(got[q][1][:, :3], vel_o[:, :3], f"scene {q} velocities")
        assert_bitexact(got[q][2], dens_o, f"scene {q} densities")
