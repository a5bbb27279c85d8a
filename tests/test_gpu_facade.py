"""The C++ drop-in (Physics::Fluid::FluidSimulation over the C ABI) gives the
same state as the C-ABI path for the same calls, and its FluidSimBase-shaped
backend returns the render colours of FluidSimCPU::updateColors
(fluidSimCPU.cc:100-125) computed on the GPU."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_check")


def test_facade_matches_c_abi(tmp_path):
    assert os.path.exists(EXE), "tests/cpp/facade_check is not built (run __graft_entry__.build())"
    out = tmp_path / "facade.bin"
    announce("facade program")
    r = subprocess.run([EXE, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    both = np.fromfile(out, np.float32).reshape(2, -1, 12)
    got, lifted = both[0], both[1]
    p = S.default_params(gravity_on=1, gravity_scale=9.81, viscosity=0.3, bounds=(6.0, 6.0, 6.0))
    with S.Sim(S.default_params(), 10000) as sim:
        sim.initialize_data()
        sim.params = p
        for _ in range(5):
            sim.update(1.0 / 60.0)
        pos, vel, dens = sim.download()
    assert_bitexact(got[:, 0:3], pos[:, :3], "positions")
    assert_bitexact(got[:, 3:6], vel[:, :3], "velocities")
    assert_bitexact(got[:, 6:8], dens, "densities")
    assert_bitexact(got[:, 8:12], O.speed_colors(vel), "render colours")
    # the oracle from the reference's GridArrangement lattice (cc:112-147,
    # 518-557), zero velocity, five Update(1/60) calls, then the caller's
    # write to `positions` (y + 0.05) and one more Update
    dt = float(np.float32(1.0 / 60.0))
    op = O.from_sph(p)
    pos0 = O.grid_arrangement(10000)
    pos_o, vel_o, dens_o = O.update(op, dt, pos0, np.zeros_like(pos0), T=10000, substeps=5)
    assert_bitexact(got[:, 0:3], pos_o[:, :3], "positions vs oracle")
    assert_bitexact(got[:, 3:6], vel_o[:, :3], "velocities vs oracle")
    assert_bitexact(got[:, 6:8], dens_o, "densities vs oracle")
    pos_o[:, 1] = pos_o[:, 1] + np.float32(0.05)
    pos_o2, vel_o2, dens_o2 = O.update(op, dt, pos_o, vel_o, T=10000)
    assert_bitexact(lifted[:, 0:3], pos_o2[:, :3], "positions after writing positions")
    assert_bitexact(lifted[:, 3:6], vel_o2[:, :3], "velocities after writing positions")
    assert_bitexact(lifted[:, 6:8], dens_o2, "densities after writing positions")
    assert_bitexact(lifted[:, 8:12], O.speed_colors(vel_o2), "colours after writing positions")
