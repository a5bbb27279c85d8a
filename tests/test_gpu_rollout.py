"""Rollouts: the 16,384-particle dam break for 200 steps (600 Update calls).

  - against the oracle's FP32 step with Jacobi viscosity: bit-exact after all
    200 steps (every Update is bit-exact, so the trajectory is too);
  - against the reference's own update order (Gauss-Seidel viscosity, what
    the reference computes with a serial PSTL backend): bounded position
    drift, kinetic energy within 5% and density histograms within a total
    variation distance of 0.05 (a 4,096-particle block, because the
    Gauss-Seidel sweep is serial on the CPU).
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


def _gpu(sc, steps):
    with S.Sim(sc.params, sc.n, sc.table) as sim:
        sim.upload(sc.pos, sc.vel)
        for _ in range(steps):
            sim.step_async(W.DT, W.SUBSTEPS)
        sim.sync()
        return sim.download()


def test_dam_break_16k_200_steps_bitexact():
    sc = W.dam_break(n=16384)
    announce("rollout", sc.name, 200)
    pos_g, vel_g, dens_g = _gpu(sc, 200)
    O.set_threads(16)
    pos_o, vel_o, dens_o = O.update(orc_params(sc.params), _sub_dt(), sc.pos, sc.vel,
                                    T=sc.table, substeps=200 * W.SUBSTEPS)
    assert_bitexact(pos_g[:, :3], pos_o[:, :3], "positions after 200 steps")
    assert_bitexact(vel_g[:, :3], vel_o[:, :3], "velocities after 200 steps")
    assert_bitexact(dens_g, dens_o, "densities after 200 steps")


def test_dam_break_statistics_vs_reference_order():
    sc = W.dam_break(n=4096)
    steps = 200
    announce("rollout statistics", sc.name, steps)
    pos_g, vel_g, dens_g = _gpu(sc, steps)
    O.set_threads(16)
    pos_r, vel_r, dens_r = O.update(orc_params(sc.params), _sub_dt(), sc.pos, sc.vel,
                                    T=sc.table, gauss_seidel=True, substeps=steps * W.SUBSTEPS)
    ke_g = 0.5 * float((vel_g[:, :3].astype(np.float64) ** 2).sum())
    ke_r = 0.5 * float((vel_r[:, :3].astype(np.float64) ** 2).sum())
    assert abs(ke_g - ke_r) <= 0.05 * ke_r, (ke_g, ke_r)
    # centre of mass and spread agree; individual particles may part ways
    com_g, com_r = pos_g[:, :3].mean(0), pos_r[:, :3].mean(0)
    assert np.abs(com_g - com_r).max() < 0.02
    assert np.abs(pos_g[:, :3].std(0) - pos_r[:, :3].std(0)).max() < 0.02
    bins = np.linspace(0, 2000, 41)
    hg = np.histogram(dens_g[:, 0], bins)[0] / sc.n
    hr = np.histogram(dens_r[:, 0], bins)[0] / sc.n
    assert 0.5 * np.abs(hg - hr).sum() < 0.05
