"""Forces of one GPU update against the FP64 oracle.

The pressure stage (physicsWorld.cc:367-422) and the viscosity stage
(cc:424-464, Jacobi) of one GPU Update are read back per particle:
  pressure term   v_press - v_ext          (sph_debug_velocities)
  viscosity term  v_final - v_press        (particles that touched no wall)
and compared per component with orc64_pressure / orc64_viscosity, the FP64
sums (pinned on the CPU against closed forms and an all-pairs FP64 sum in
test_oracle_forces.py).  The FP64 sums take the GPU's own FP32 inputs of the
stage -- predicted positions, densities, post-pressure velocities -- so the
check measures the FP32 force arithmetic itself; densities are compared with
their FP64 sums separately (test_gpu_parity.py, 1e-5 relative).

Tolerance, per particle and component:
    |term_gpu - term_64| <= 1e-5 * mag + 2^-23 * |v|
mag is the scale of the FP64 sum (orc64's mag3: the sum of the magnitudes of
the pair terms, divided by rho_i and times dt), and 2^-23 |v| covers the one
FP32 rounding of the velocity each term is added to (cc:421, 463).

Workloads: the 16,384-particle dam break at its first update (gravity on,
surface particles below the target density) and the 1,048,576-particle
uniform scene (C1 of the benchmark set).
"""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ULP_V = 2.0**-23


def _sub_dt():
    return float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))


@pytest.fixture(scope="module", autouse=True)
def _threads():
    O.set_threads(16)


def _check(sc, updates_before=0):
    dt = _sub_dt()
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        sim.upload(sc.pos, sc.vel)
        for _ in range(updates_before):
            sim.update(dt)
        sim.update(dt)
        xp = sim.debug_predicted()
        v_ext, v_press = sim.debug_velocities()
        pos, vel, dens = sim.download()
    op = orc_params(sc.params)
    sp = O.spatial(xp, sc.n, op.h)
    # pressure: FP64 sum from the GPU's FP32 positions and densities
    dv64, mag64 = O.pressure64(op, dt, xp, sp, dens.astype(np.float64))
    got = v_press[:, :3].astype(np.float64) - v_ext[:, :3]
    err = np.abs(got - dv64)
    bound = FORCE_RTOL * mag64 + ULP_V * np.abs(v_press[:, :3])
    bad = np.argwhere(err > bound)
    assert bad.size == 0, (f"pressure: {len(bad)} components beyond the bound; first "
                           f"{bad[:3].tolist()}, err {err[tuple(bad[0])]}, bound {bound[tuple(bad[0])]}")
    assert np.abs(dv64).max() > 1e3 * ULP_V * np.abs(v_press[:, :3]).max(), "pressure terms vanish"
    worst_p = float((err[bound > 0] / bound[bound > 0]).max())
    # viscosity (Jacobi): FP64 sum from the GPU's post-pressure velocities
    dvv, magv = O.viscosity64(op, dt, xp, sp, v_press[:, :3].astype(np.float64))
    half = np.asarray(sc.params.bounds, np.float32) * np.float32(0.5)
    free = ~np.any(np.abs(pos[:, :3]) == half[None, :], axis=1)   # cc:90-105 left v alone
    assert free.mean() > 0.9
    gotv = vel[free, :3].astype(np.float64) - v_press[free, :3]
    errv = np.abs(gotv - dvv[free])
    boundv = FORCE_RTOL * magv[free] + ULP_V * np.abs(vel[free, :3])
    badv = np.argwhere(errv > boundv)
    assert badv.size == 0, f"viscosity: {len(badv)} components beyond the bound"
    assert np.abs(dvv[free]).max() > 0, "viscosity terms vanish"
    worst_v = float((errv[boundv > 0] / boundv[boundv > 0]).max())
    announce("forces", sc.name, "max err/bound: pressure", f"{worst_p:.3f}", "viscosity", f"{worst_v:.3f}")


def test_forces_fp64_dam_break_16k():
    sc = W.dam_break(n=16384)
    announce("forces fp64", sc.name, "first update")
    _check(sc)


@pytest.mark.parametrize("mu", [0.05, 1.0])
def test_forces_fp64_dam_break_16k_after_30_updates(mu):
    # the block has started to fall and spread: uneven densities, viscosity
    # acting on a nonuniform velocity field (at mu = 1 its terms are 20x
    # those of the benchmark setting, well above the rounding of v)
    sc = W.dam_break(n=16384)
    p = sc.params
    p.viscosity = mu
    sc.params = p
    announce("forces fp64", sc.name, "update 31, viscosity", mu)
    _check(sc, updates_before=30)


def test_forces_fp64_uniform_1m():
    sc = W.uniform_lower()
    announce("forces fp64", sc.name, sc.n)
    _check(sc)
