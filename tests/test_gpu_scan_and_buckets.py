"""The counting sort under stress: a count scan whose bounded wait gives up
must be reported, and buckets of thousands of records must stay exact.

  - the scan's look-back wait (the prefix sum behind the bucket table of
    UpdateSpatialLookup, physicsWorld.cc:486-496) is bounded; the host sets
    the bound to 0 so that any wait gives up at once.  Every kernel still
    exits normally, every synchronising call then reports SPH_ERR_DEVICE or
    the update was exact (no wait was needed), the state stays invalid until
    it is uploaded again, and afterwards the handle is exact again;
  - a reference table of 8 buckets for 32,768 particles (about 4,096 records
    per bucket) and a scene with every particle at the same point (one bucket
    of all of them), under both keys: bit-exact against the oracle.  The
    reorder ranks records of buckets over 64 records through LDS tiles
    shared by the workgroup.
"""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu


def _sub_dt():
    return float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))


@pytest.fixture(scope="module", autouse=True)
def _threads():
    O.set_threads(16)


def _expect_exact(sim, sc, dt, what):
    pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel, T=sc.table)
    pos, vel, dens = sim.download()
    assert_bitexact(pos[:, :3], pos_o[:, :3], f"{what} pos")
    assert_bitexact(vel[:, :3], vel_o[:, :3], f"{what} vel")
    assert_bitexact(dens, dens_o, f"{what} dens")


def test_scan_give_up_is_reported_and_recoverable():
    # 2^24 buckets: 4096 chunks of the count scan, so with a zero bound some
    # workgroup finds a predecessor's chunk unpublished
    sc = W.uniform_lower(n=65536, table=1 << 24, box=(8.0, 2.0, 8.0))
    dt = _sub_dt()
    announce("scan give-up", sc.n, "table", sc.table)
    with S.Sim(sc.params, sc.n, sc.table) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.debug_set_scan_spin(0)
        reported = 0
        for attempt in range(12):
            sim.upload(sc.pos, sc.vel)
            try:
                sim.update(dt)
            except S.SphError as e:
                assert e.code == S.SPH_ERR_DEVICE, e
                reported += 1
                # the state stays invalid until it is replaced
                with pytest.raises(S.SphError) as e2:
                    sim.download()
                assert e2.value.code == S.SPH_ERR_DEVICE
                continue
            _expect_exact(sim, sc, dt, f"attempt {attempt} without a give-up")
        announce("give-ups reported", reported, "of 12")
        assert reported > 0, "no wait gave up with a zero bound"
        sim.debug_set_scan_spin(65536)
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        _expect_exact(sim, sc, dt, "after recovery")


@pytest.mark.parametrize("table", [8, 64])
def test_large_reference_buckets_exact(table):
    sc = W.uniform_lower(n=32768, table=table, box=(4.0, 2.0, 4.0))
    dt = _sub_dt()
    announce("large buckets", sc.n, "table", table, "records per bucket", sc.n // table)
    with S.Sim(sc.params, sc.n, table) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        sp_g = sim.debug_spatial()
        _expect_exact(sim, sc, dt, f"table {table}")
    v1, xp = O.predict(orc_params(sc.params), dt, sc.pos, sc.vel)
    sp = O.spatial(xp, table, sc.params.smoothing_radius)
    np.testing.assert_array_equal(sp_g["sorted_idx"], sp.sorted_idx)
    np.testing.assert_array_equal(sp_g["start"], sp.start)
    if table == 8:
        assert np.bincount(sp.key_id, minlength=table).max() >= 4000


@pytest.mark.parametrize("table", [S.TABLE_GRID, S.TABLE_REFERENCE])
def test_all_particles_at_one_point(table):
    n = 8192
    p = W.scene_params(box=(4.0, 4.0, 4.0))
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.float32(0.3)
    vel = np.zeros((n, 4), np.float32)
    sc = W.Scene("one_point", p, n, n, pos, vel, 1)
    dt = _sub_dt()
    announce("one point", n, "table", table)
    with S.Sim(p, n, table) as sim:
        sim.upload(pos, vel)
        sim.update(dt)
        _expect_exact(sim, sc, dt, "one point")
