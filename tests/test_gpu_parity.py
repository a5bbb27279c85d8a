"""GPU parity: the gfx950 step through the C ABI vs the CPU oracle on identical inputs.

Bar (stated per quantity):
  - predicted positions, hashes, keys, sorted key sequence, sorted particle
    order, bucket start/end tables, neighbour lists: bit-exact;
  - density / near density, and positions / velocities after a full Update:
    bit-exact against the oracle's FP32 restatement (same operation order,
    same (key, index) summation order, Jacobi viscosity), and density within
    1e-5 relative of the FP64 restatement.
"""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

DENSITY_RTOL_FP64 = 1e-5


def _sub_dt(dt=W.DT, substeps=W.SUBSTEPS):
    return float(np.float32(np.float32(dt) / np.float32(substeps)))


def _cases():
    return {
        "dam_break_4k": lambda: W.dam_break(n=4096),
        "dam_break_16k": lambda: W.dam_break(n=16384),
        "uniform_64k_T2N": lambda: W.uniform_lower(n=65536, table=131072, box=(8.0, 2.0, 8.0)),
        "uniform_64k_T1000": lambda: W.uniform_lower(n=65536, table=1000, box=(8.0, 2.0, 8.0)),
        "uniform_64k_T2pow16_h035": lambda: _h035(),
        "dense_lattice_overflow": lambda: _dense(),
    }


def _dense():
    """Lattice spacing 0.07 at h = 0.2: ~97 neighbours per particle, more than
    the 64-entry neighbour list holds, so the walk fallback runs."""
    n = 8192
    pos, vel = S.generate(S.INIT_DAM_BREAK, 3, n, (4.0, 4.0, 4.0), 0.07, 0.001)
    return W.Scene("dense_8192", W.scene_params(), n, n, pos, vel, 1)


def _h035():
    sc = W.uniform_lower(n=65536, table=65536, box=(8.0, 2.0, 8.0))
    p = sc.params
    p.smoothing_radius = 0.35  # the reference default: prefilter radius == h
    p.target_density = 99.7
    sc.params = p
    return sc


def _oracle_update(sc, dt):
    """One Update(dt) of the oracle, stage by stage, keeping the intermediates."""
    op = orc_params(sc.params)
    v1, xp = O.predict(op, dt, sc.pos, sc.vel)
    sp = O.spatial(xp, sc.table, op.h)
    dens = O.density(op, xp, sp)
    v2 = O.pressure(op, dt, xp, sp, dens, v1)
    v3 = O.viscosity(op, dt, xp, sp, v2, gauss_seidel=False)
    pos, vel = O.integrate(op, dt, sc.pos, v3)
    return dict(op=op, xp=xp, sp=sp, dens=dens, pos=pos, vel=vel)


@pytest.fixture(scope="module", autouse=True)
def _threads():
    O.set_threads(16)


@pytest.mark.parametrize("case", list(_cases()))
def test_one_update_bitexact(case):
    sc = _cases()[case]()
    dt = _sub_dt()
    announce("one update", case, "n", sc.n, "table", sc.table)
    with S.Sim(sc.params, sc.n, sc.table) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        xp_g = sim.debug_predicted()
        sp_g = sim.debug_spatial()
        pos_g, vel_g, dens_g = sim.download()
    ref = _oracle_update(sc, dt)
    sp = ref["sp"]
    assert_bitexact(xp_g[:, :3], ref["xp"][:, :3], "predicted positions")
    np.testing.assert_array_equal(sp_g["hash_id"], sp.hash_id)
    np.testing.assert_array_equal(sp_g["key_id"], sp.key_id)
    np.testing.assert_array_equal(sp_g["sorted_key"], sp.sorted_key)
    np.testing.assert_array_equal(sp_g["sorted_idx"], sp.sorted_idx)
    np.testing.assert_array_equal(sp_g["start"], sp.start)
    # end table: one past the last slot of every non-empty bucket
    counts = np.bincount(sp.key_id, minlength=sc.table)
    nz = counts > 0
    np.testing.assert_array_equal(sp_g["end"][nz], sp.start[nz] + counts[nz])
    assert_bitexact(dens_g, ref["dens"], "density")
    assert_bitexact(vel_g[:, :3], ref["vel"][:, :3], "velocity after Update")
    assert_bitexact(pos_g[:, :3], ref["pos"][:, :3], "position after Update")
    assert np.all(pos_g[:, 3] == np.float32(0.34))
    d64 = O.density64(ref["op"], ref["xp"], sp)
    rel = np.abs(dens_g.astype(np.float64) - d64) / np.abs(d64)
    assert rel.max() <= DENSITY_RTOL_FP64, rel.max()


@pytest.mark.parametrize("case", ["dam_break_4k", "uniform_64k_T1000"])
def test_neighbour_lists_exact(case):
    sc = _cases()[case]()
    dt = _sub_dt()
    announce("neighbours", case)
    with S.Sim(sc.params, sc.n, sc.table) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        offs_g, list_g = sim.debug_neighbours()
    ref = _oracle_update(sc, dt)
    offs_o, list_o = O.neighbours(ref["op"], ref["xp"], ref["sp"])
    np.testing.assert_array_equal(offs_g, offs_o)
    np.testing.assert_array_equal(list_g, list_o)


def test_rollout_bitexact_dam_break_4k():
    sc = W.dam_break(n=4096)
    steps = 10
    announce("rollout", sc.name, "steps", steps)
    with S.Sim(sc.params, sc.n, sc.table) as sim:
        sim.upload(sc.pos, sc.vel)
        for _ in range(steps):
            sim.step(W.DT, W.SUBSTEPS)
        pos_g, vel_g, dens_g = sim.download()
    pos_o, vel_o, dens_o = O.update(orc_params(sc.params), _sub_dt(), sc.pos, sc.vel,
                                    T=sc.table, substeps=steps * W.SUBSTEPS)
    assert_bitexact(pos_g[:, :3], pos_o[:, :3], "positions after 10 steps")
    assert_bitexact(vel_g[:, :3], vel_o[:, :3], "velocities after 10 steps")
    assert_bitexact(dens_g, dens_o, "densities after 10 steps")


def test_uniform_1m_tables_and_density():
    """The 1,048,576-particle scene under the reference key with its 2^21
    table, step 0: hashes, keys, sorted keys and order, the start table, every
    particle's neighbour list (151,572,666 entries) and the densities equal the
    oracle's."""
    sc = W.uniform_lower()
    dt = _sub_dt()
    announce("uniform_1m step-0 tables", sc.n, sc.table)
    with S.Sim(sc.params, sc.n, sc.table) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        sp_g = sim.debug_spatial()
        offs_g, list_g = sim.debug_neighbours()
        _, _, dens_g = sim.download()
    op = orc_params(sc.params)
    _, xp = O.predict(op, dt, sc.pos, sc.vel)
    sp = O.spatial(xp, sc.table, op.h)
    np.testing.assert_array_equal(sp_g["hash_id"], sp.hash_id)
    np.testing.assert_array_equal(sp_g["key_id"], sp.key_id)
    np.testing.assert_array_equal(sp_g["sorted_key"], sp.sorted_key)
    np.testing.assert_array_equal(sp_g["sorted_idx"], sp.sorted_idx)
    np.testing.assert_array_equal(sp_g["start"], sp.start)
    offs_o, list_o = O.neighbours(op, xp, sp)
    announce("uniform_1m neighbour entries", int(offs_o[-1]))
    np.testing.assert_array_equal(offs_g, offs_o)
    assert np.array_equal(list_g, list_o), "neighbour lists differ"
    del list_g, list_o
    dens_o = O.density(op, xp, sp)
    assert_bitexact(dens_g, dens_o, "density uniform_1m")


def test_initialize_data_matches_reference_lattice():
    n = 10000  # the reference app's particleAmount (gameApp.cc:41)
    p = S.default_params()
    announce("initialize_data", n)
    with S.Sim(p, n) as sim:
        sim.initialize_data()
        pos_g, vel_g, dens_g = sim.download()
    pos_o = O.grid_arrangement(n)
    assert_bitexact(pos_g, pos_o, "GridArrangement positions")
    assert np.all(vel_g == 0)
    op = orc_params(p)
    sp = O.spatial(pos_o, n, op.h)
    assert_bitexact(dens_g, O.density(op, pos_o, sp), "InitializeData densities")


# (work, list, block, row mode); in row mode pressure reads rows and
# viscosity alternates between rows and its walk
VARIANTS = [(w, l, b, 0) for w in (0, 1) for l in (0, 1) for b in (64, 128, 256)] + \
    [(1, 0, b, 1) for b in (64, 128, 256)]


@pytest.mark.parametrize("case", ["dam_break_16k", "uniform_64k_T1000", "dense_lattice_overflow"])
def test_every_gather_variant_bitexact(case):
    """All work-order / LDS-list / workgroup-size variants give the oracle's bits."""
    sc = _cases()[case]()
    dt = _sub_dt()
    ref = _oracle_update(sc, dt)
    with S.Sim(sc.params, sc.n, sc.table) as sim:
        for (w, l, b, nb) in VARIANTS:
            announce("variant", case, "work", w, "list", l, "block", b, "nbr", nb)
            v = {"work": w, "list": l, "block": b, "rows": bool(nb)}
            vv = dict(v, rows=bool(nb) and b != 128)
            sim.variants = {"density": v, "pressure": v, "viscosity": vv, "nbr": nb}
            assert sim.variants["nbr"] == bool(nb)
            sim.upload(sc.pos, sc.vel)
            sim.update(dt)
            pos_g, vel_g, dens_g = sim.download()
            assert_bitexact(dens_g, ref["dens"], f"density {v}")
            assert_bitexact(vel_g[:, :3], ref["vel"][:, :3], f"velocity {v}")
            assert_bitexact(pos_g[:, :3], ref["pos"][:, :3], f"position {v}")


def test_autotune_keeps_state_and_results():
    sc = W.dam_break(n=16384)
    dt = _sub_dt()
    announce("autotune", sc.name)
    with S.Sim(sc.params, sc.n, sc.table) as sim:
        sim.upload(sc.pos, sc.vel)
        ms = sim.autotune(dt, updates=2)
        assert np.all(ms > 0)
        pos_a, vel_a, _ = sim.download()
        assert_bitexact(pos_a, sc.pos, "positions after autotune")
        assert_bitexact(vel_a, sc.vel, "velocities after autotune")
        for _ in range(2):
            sim.step(W.DT, W.SUBSTEPS)
        pos_g, vel_g, dens_g = sim.download()
    pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel, T=sc.table,
                                    substeps=2 * W.SUBSTEPS)
    assert_bitexact(pos_g[:, :3], pos_o[:, :3], "positions after autotune + 2 steps")
    assert_bitexact(vel_g[:, :3], vel_o[:, :3], "velocities after autotune + 2 steps")
    assert_bitexact(dens_g, dens_o, "densities after autotune + 2 steps")
