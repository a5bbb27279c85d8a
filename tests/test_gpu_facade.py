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
    got = np.fromfile(out, np.float32).reshape(-1, 12)
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
