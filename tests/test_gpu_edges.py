"""Edge cases of the step on the GPU, bit-exact against the oracle under both
bucket keys: particle counts around the wave and workgroup sizes (1 to 257),
every particle at one point (d = 0: direction (0, 1, 0), cc:414), particles
far outside the box (their cells wrap onto the grid), and a batch of
one-particle scenes."""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))


def _check(p, pos, vel, what, updates=3):
    n = pos.shape[0]
    po, vo, do = O.update(orc_params(p), DT, pos, vel, T=n, substeps=updates)
    for table in (S.TABLE_GRID, S.TABLE_REFERENCE):
        announce(what, "n", n, "table", table)
        with S.Sim(p, n, table) as sim:
            sim.upload(pos, vel)
            for _ in range(updates):
                sim.update(DT)
            pg, vg, dg = sim.download()
        assert_bitexact(pg[:, :3], po[:, :3], f"{what} positions (table {table})")
        assert_bitexact(vg[:, :3], vo[:, :3], f"{what} velocities (table {table})")
        assert_bitexact(dg, do, f"{what} densities (table {table})")


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 127, 255, 257])
def test_small_counts(n):
    p = W.scene_params(h=0.2, box=(2.0, 2.0, 2.0))
    pos, vel = S.generate(S.INIT_CENTRED_BLOCK, 3, n, (2.0, 2.0, 2.0), 0.08, 0.01)
    _check(p, pos, vel, f"small n={n}")


def test_all_particles_at_one_point():
    n = 500
    p = W.scene_params(h=0.2, box=(2.0, 2.0, 2.0))
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.float32(0.123)
    pos[:, 3] = np.float32(0.34)
    vel = np.zeros((n, 4), np.float32)
    _check(p, pos, vel, "coincident")


def test_particles_far_outside_the_box():
    n = 4096
    p = W.scene_params(h=0.2, box=(4.0, 4.0, 4.0))
    pos, vel = S.generate(S.INIT_DAM_BREAK, 4, n, (4.0, 4.0, 4.0), 0.1, 0.002)
    rng = np.random.default_rng(1)
    far = rng.choice(n, 300, replace=False)
    pos[far, :3] += rng.uniform(-3e4, 3e4, (300, 3)).astype(np.float32)
    # a tight group far away keeps some of them within h of each other
    pos[far[:40], :3] = np.float32([1234.5, -987.25, 4321.0]) + \
        rng.uniform(-0.1, 0.1, (40, 3)).astype(np.float32)
    _check(p, pos, vel, "far outside", updates=1)


def test_batch_of_one_particle_scenes():
    params = [W.scene_params(h=0.1 + 0.02 * s, box=(1.0, 1.0, 1.0)) for s in range(5)]
    pos = [np.array([[0.01 * s, -0.02 * s, 0.03, 0.34]], np.float32) for s in range(5)]
    vel = [np.array([[0.1, 0.2 * s, -0.3, 0.0]], np.float32) for s in range(5)]
    announce("batch of one-particle scenes")
    with S.Batch(params, [1] * 5) as b:
        for s in range(5):
            b.upload(s, pos[s], vel[s])
        b.step(W.DT, W.SUBSTEPS)
        got = [b.download(s) for s in range(5)]
    for s in range(5):
        po, vo, do = O.update(orc_params(params[s]), DT, pos[s], vel[s], T=1, substeps=W.SUBSTEPS)
        assert_bitexact(got[s][0][:, :3], po[:, :3], f"scene {s} position")
        assert_bitexact(got[s][1][:, :3], vo[:, :3], f"scene {s} velocity")
        assert_bitexact(got[s][2], do, f"scene {s} density")


def test_coordinates_near_zero_and_near_coincident_pairs():
    """Zero, tiny and subnormal coordinates (differences below 2^-84) and
    pairs closer than 2^-26 take the pressure pass's IEEE divisions by d;
    every other pair its reciprocal quotients.  Both stay bit-exact."""
    n = 600
    rng = np.random.default_rng(5)
    p = W.scene_params(h=0.2, box=(2.0, 2.0, 2.0))
    pos = np.zeros((n, 4), np.float32)
    pos[:, 3] = np.float32(0.34)
    pos[:, :3] = rng.uniform(-0.15, 0.15, (n, 3)).astype(np.float32)
    pos[:100, 0] = rng.choice(np.array([0.0, 1e-30, -3e-35, 1e-38, 1.4e-45], np.float32), 100)
    pos[100:200, 1] = rng.choice(np.array([0.0, 1e-25, -1e-29], np.float32), 100)
    # one float32 ulp apart near 0.1 (about 7.5e-9 < 2^-26)
    pos[200:300, :3] = np.nextafter(pos[300:400, :3], np.float32(1.0))
    vel = np.zeros((n, 4), np.float32)
    _check(p, pos, vel, "near zero")


def test_reciprocal_equals_ieee_on_the_whole_divisor_range():
    """The pressure kernel's RN(1/d) (hardware estimate + one Newton step) is
    IEEE 1/d for every float d in [2^-26, 2^26]: 436M values on the GPU."""
    announce("reciprocal sweep")
    count, first = S.debug_reciprocal_mismatches(0x32800000, 0x4C800000)
    assert count == 0, f"{count} mismatches, first at bits {first:#010x}"
