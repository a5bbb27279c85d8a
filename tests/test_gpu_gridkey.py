"""GPU parity under the cell-grid bucket key (sph_create with SPH_TABLE_GRID).

The grid key changes the spatial lookup tables, never a result: positions,
velocities and densities after an Update must be bit-identical to the
oracle's reference-key step (hash % T, physicsWorld.cc:513-516).  The lookup
tables themselves are checked against the oracle run with the same grid key
(its test hook), bit for bit.  Cases include particles outside the grid's
box (cells wrap around the grid) and a smoothing radius so small that the
grid is capped (every axis wraps)."""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu


def _sub_dt():
    return float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))


def _outside():
    """Particles spread over 8 x 2 x 8 while the box is 2 x 2 x 2: most cells
    lie outside the grid and wrap onto it (seams inside the fluid)."""
    sc = W.uniform_lower(n=32768, table=32768, box=(8.0, 2.0, 8.0))
    p = sc.params
    p.bounds[0], p.bounds[1], p.bounds[2] = 2.0, 2.0, 2.0
    sc.params = p
    return sc


def _tiny_h():
    """h = 0.012 over a 16 x 8 x 16 box: 10^9 cells, capped to 2^20 buckets."""
    n = 20000
    pos, vel = S.generate(S.INIT_CENTRED_BLOCK, 7, n, (16.0, 8.0, 16.0), 0.006, 0.0005)
    p = W.scene_params(h=0.012, box=(16.0, 8.0, 16.0))
    return W.Scene("tiny_h", p, n, n, pos, vel, 1)


def _cases():
    return {
        "dam_break_4k": lambda: W.dam_break(n=4096),
        "dam_break_16k": lambda: W.dam_break(n=16384),
        "uniform_64k": lambda: W.uniform_lower(n=65536, table=131072, box=(8.0, 2.0, 8.0)),
        "outside_box_wrap": _outside,
        "tiny_h_capped": _tiny_h,
    }


def _oracle(sc, dt):
    op = orc_params(sc.params)
    v1, xp = O.predict(op, dt, sc.pos, sc.vel)
    sp = O.spatial(xp, sc.table, op.h)
    dens = O.density(op, xp, sp)
    v2 = O.pressure(op, dt, xp, sp, dens, v1)
    v3 = O.viscosity(op, dt, xp, sp, v2, gauss_seidel=False)
    pos, vel = O.integrate(op, dt, sc.pos, v3)
    return dict(op=op, xp=xp, dens=dens, pos=pos, vel=vel)


@pytest.fixture(scope="module", autouse=True)
def _threads():
    O.set_threads(16)


@pytest.mark.parametrize("case", list(_cases()))
def test_grid_key_update_bitexact(case):
    sc = _cases()[case]()
    dt = _sub_dt()
    announce("grid key update", case, "n", sc.n)
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        grid, gdim, lo = sim.key()
        assert grid
        T = sim.table
        assert int(np.prod(gdim)) <= T
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        sp_g = sim.debug_spatial()
        pos_g, vel_g, dens_g = sim.download()
    ref = _oracle(sc, dt)
    assert_bitexact(dens_g, ref["dens"], "density")
    assert_bitexact(vel_g[:, :3], ref["vel"][:, :3], "velocity after Update")
    assert_bitexact(pos_g[:, :3], ref["pos"][:, :3], "position after Update")
    # the lookup tables equal the oracle's under the same grid key
    with O.grid_key(gdim, lo):
        sp = O.spatial(ref["xp"], T, ref["op"].h)
    np.testing.assert_array_equal(sp_g["hash_id"], sp.hash_id)
    np.testing.assert_array_equal(sp_g["key_id"], sp.key_id)
    np.testing.assert_array_equal(sp_g["sorted_idx"], sp.sorted_idx)
    np.testing.assert_array_equal(sp_g["start"], sp.start)
    counts = np.bincount(sp.key_id, minlength=T)
    nz = counts > 0
    np.testing.assert_array_equal(sp_g["end"][nz], sp.start[nz] + counts[nz])


def test_grid_key_capped_table_wraps():
    sc = _tiny_h()
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        _, gdim, _ = sim.key()
        assert sim.table == 1 << 20 and int(np.prod(gdim)) <= 1 << 20
        assert max(gdim) < 16.0 / 0.012  # shrunk: cells wrap


@pytest.mark.parametrize("case", ["dam_break_4k", "outside_box_wrap"])
def test_grid_key_neighbour_lists(case):
    sc = _cases()[case]()
    dt = _sub_dt()
    announce("grid key neighbours", case)
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        offs_g, list_g = sim.debug_neighbours()
    ref = _oracle(sc, dt)
    sp = O.spatial(ref["xp"], sc.table, ref["op"].h)
    offs_o, list_o = O.neighbours(ref["op"], ref["xp"], sp)
    np.testing.assert_array_equal(offs_g, offs_o)
    np.testing.assert_array_equal(list_g, list_o)


def _dense():
    """~97 neighbours per particle: row runs longer than 64 slots, so the
    neighbour masks cannot hold them and those particles walk instead."""
    n = 8192
    pos, vel = S.generate(S.INIT_DAM_BREAK, 3, n, (4.0, 4.0, 4.0), 0.07, 0.001)
    return W.Scene("dense_8192", W.scene_params(), n, n, pos, vel, 1)


def _lattice_08():
    """Lattice spacing 0.08 (h / 2.5): row runs of 48-64 slots, streams of
    about 300-500 bits, so a third of the records use the stream's fourth
    128-bit word."""
    n = 8192
    pos, vel = S.generate(S.INIT_DAM_BREAK, 3, n, (4.0, 4.0, 4.0), 0.08, 0.001)
    return W.Scene("lattice_08", W.scene_params(), n, n, pos, vel, 1)


def _stream_over_512():
    """2 x 2 x 5 particles in every cell (spacing 0.1, 0.1, 0.04 at fixed
    offsets inside the cells): interior rows run 60 slots, within the 64-slot
    limit, but nine of them make 540 bits, past the stream's 512, so those
    records are unusable while the block's boundary layers keep usable ones."""
    n = 16 * 16 * 32
    pos, vel = S.generate(S.INIT_DAM_BREAK, 3, n, (4.0, 4.0, 4.0), 0.1, 0.0)
    i, j, k = np.meshgrid(np.arange(16), np.arange(16), np.arange(32), indexing="ij")
    pos[:, 0] = (-2.0 + 0.05 + 0.1 * i).reshape(-1)
    pos[:, 1] = (-2.0 + 0.05 + 0.1 * j).reshape(-1)
    pos[:, 2] = (-2.0 + 0.02 + 0.04 * k).reshape(-1)
    return W.Scene("stream_over_512", W.scene_params(), n, n, pos, vel, 1)


# (list, block, neighbour masks, pressure reads masks, viscosity reads masks)
GRID_VARIANTS = [(l, b, 0, 0, 0) for l in (0, 1) for b in (64, 128, 256)] + \
    [(l, b, 1, 1, 1) for l in (0, 1) for b in (64, 128, 256)] + [(0, 128, 1, 1, 0), (1, 128, 1, 0, 1)]


@pytest.mark.parametrize("case", ["dam_break_16k", "uniform_64k", "outside_box_wrap", "dense",
                                  "lattice_08", "stream_over_512"])
def test_grid_key_every_variant_bitexact(case):
    sc = {"dense": _dense, "lattice_08": _lattice_08,
          "stream_over_512": _stream_over_512}.get(case, lambda: _cases()[case]())()
    dt = _sub_dt()
    ref = _oracle(sc, dt)
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        for (l, b, nb, rp, rv) in GRID_VARIANTS:
            announce("grid variant", case, "list", l, "block", b, "masks", nb, rp, rv)
            v = {"work": 0, "list": l, "block": b, "rows": False}
            sim.variants = {"density": v, "pressure": dict(v, rows=bool(rp)),
                            "viscosity": dict(v, rows=bool(rv)), "nbr": nb}
            got = sim.variants
            assert got["nbr"] == bool(nb) and got["pressure"]["rows"] == bool(rp)
            sim.upload(sc.pos, sc.vel)
            sim.update(dt)
            pos_g, vel_g, dens_g = sim.download()
            assert_bitexact(dens_g, ref["dens"], f"density {v} masks {nb}")
            assert_bitexact(vel_g[:, :3], ref["vel"][:, :3], f"velocity {v} masks {nb}{rp}{rv}")
            assert_bitexact(pos_g[:, :3], ref["pos"][:, :3], f"position {v} masks {nb}{rp}{rv}")
        # the grid key has no Morton work order
        sim.variants = {"density": {"work": 1, "list": 0, "block": 64, "rows": False},
                        "pressure": {"work": 1, "list": 1, "block": 64, "rows": False},
                        "viscosity": {"work": 0, "list": 0, "block": 64, "rows": False}, "nbr": 0}
        got = sim.variants
        assert not got["density"]["work"] and not got["pressure"]["work"]


def test_grid_key_rollout_and_autotune():
    sc = W.dam_break(n=4096)
    steps = 10
    announce("grid key rollout", sc.name, steps)
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.autotune(_sub_dt(), updates=2)
        for _ in range(steps):
            sim.step(W.DT, W.SUBSTEPS)
        pos_g, vel_g, dens_g = sim.download()
    pos_o, vel_o, dens_o = O.update(orc_params(sc.params), _sub_dt(), sc.pos, sc.vel,
                                    T=sc.table, substeps=steps * W.SUBSTEPS)
    assert_bitexact(pos_g[:, :3], pos_o[:, :3], "positions after 10 steps")
    assert_bitexact(vel_g[:, :3], vel_o[:, :3], "velocities after 10 steps")
    assert_bitexact(dens_g, dens_o, "densities after 10 steps")


def test_grid_key_uniform_1m():
    sc = W.uniform_lower()
    dt = _sub_dt()
    announce("grid key uniform_1m", sc.n)
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        pos_g, vel_g, dens_g = sim.download()
    ref = _oracle(sc, dt)
    assert_bitexact(dens_g, ref["dens"], "density uniform_1m")
    assert_bitexact(vel_g[:, :3], ref["vel"][:, :3], "velocity uniform_1m")
    assert_bitexact(pos_g[:, :3], ref["pos"][:, :3], "position uniform_1m")


def test_grid_key_batch_equals_oracle():
    scenes = [W.batch_scene(n, h, mu, seed) for seed, (n, h, mu) in enumerate(
        [(4096, 0.1, 0.0), (5000, 0.3, 0.2), (7777, 0.2, 0.05), (4500, 0.15, 0.1)])]
    announce("grid key batch", [sc.n for sc in scenes])
    with S.Batch([sc.params for sc in scenes], [sc.n for sc in scenes]) as b:
        for s, sc in enumerate(scenes):
            b.upload(s, sc.pos, sc.vel)
        for l, nb in ((0, 0), (1, 0), (0, 1)):
            v = {"work": 0, "list": l, "block": 128, "rows": bool(nb)}
            b.variants = {"density": v, "pressure": v, "viscosity": v, "nbr": nb}
            for s, sc in enumerate(scenes):
                b.upload(s, sc.pos, sc.vel)
            b.step(W.DT, W.SUBSTEPS)
            got = [b.download(s) for s in range(len(scenes))]
            for s, sc in enumerate(scenes):
                pos_o, vel_o, dens_o = O.update(orc_params(sc.params), _sub_dt(), sc.pos, sc.vel,
                                                T=sc.table, substeps=W.SUBSTEPS)
                assert_bitexact(got[s][0][:, :3], pos_o[:, :3], f"scene {s} pos list {l} masks {nb}")
                assert_bitexact(got[s][1][:, :3], vel_o[:, :3], f"scene {s} vel list {l} masks {nb}")
                assert_bitexact(got[s][2], dens_o, f"scene {s} dens list {l} masks {nb}")


def test_grid_key_settings_change_mid_run():
    """setInteractionRadius / setBound between updates (physicsWorld.h:56-77):
    the grid periods are recomputed inside the handle's fixed bucket count, so
    a smaller h makes cells wrap.  Results stay bit-exact."""
    sc = W.dam_break(n=8192)
    dt = _sub_dt()
    op0 = orc_params(sc.params)
    p1 = W.scene_params(h=0.15, box=(5.0, 4.0, 6.0))  # dam_break's settings otherwise
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        g0 = sim.key()[1]
        sim.upload(sc.pos, sc.vel)
        sim.step(W.DT, W.SUBSTEPS)
        sim.params = p1
        g1 = sim.key()[1]
        sim.step(W.DT, W.SUBSTEPS)
        pos_g, vel_g, dens_g = sim.download()
    assert g0 != g1
    pos_o, vel_o, _ = O.update(op0, dt, sc.pos, sc.vel, T=sc.n, substeps=W.SUBSTEPS)
    pos_o, vel_o, dens_o = O.update(orc_params(p1), dt, pos_o, vel_o, T=sc.n, substeps=W.SUBSTEPS)
    assert_bitexact(pos_g[:, :3], pos_o[:, :3], "positions after a settings change")
    assert_bitexact(vel_g[:, :3], vel_o[:, :3], "velocities after a settings change")
    assert_bitexact(dens_g, dens_o, "densities after a settings change")
