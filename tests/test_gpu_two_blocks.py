"""The 16,777,216-particle two-block scene (two lattice blocks in a 40 x 20 x 40
box, moving together at 2 m/s) at full size: one Update bit-exact against the
oracle under both bucket keys (the cell grid, and the reference's hash % N
with N = 2^24), then ten steps that stay finite and move the blocks toward
each other."""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_two_blocks_16m_full_size():
    sc = W.two_blocks()
    assert sc.n == 1 << 24
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    O.set_threads(16)
    announce("two_blocks_16m oracle update (16 threads)")
    pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel, T=sc.table)
    for table in (S.TABLE_GRID, sc.table):
        announce("two_blocks_16m GPU update, table", table)
        with S.Sim(sc.params, sc.n, table) as sim:
            sim.upload(sc.pos, sc.vel)
            sim.update(dt)
            pos_g, vel_g, dens_g = sim.download()
            assert_bitexact(dens_g, dens_o, f"densities (table {table})")
            assert_bitexact(vel_g[:, :3], vel_o[:, :3], f"velocities (table {table})")
            assert_bitexact(pos_g[:, :3], pos_o[:, :3], f"positions (table {table})")
            if table == S.TABLE_GRID:
                half = sc.n // 2
                gap0 = pos_g[half:, 0].min() - pos_g[:half, 0].max()
                for _ in range(10):
                    sim.step(W.DT, W.SUBSTEPS)
                assert sim.count_nonfinite() == 0
                pos_t, _, _ = sim.download()
                gap1 = pos_t[half:, 0].min() - pos_t[:half, 0].max()
                assert gap1 < gap0  # the blocks approach each other
