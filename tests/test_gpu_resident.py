"""The particle state stays on the device in the sorted order of the last
update (slot k holds particle ids[k]); every interface still speaks particle
order.  After updates have permuted the resident order:
  - downloads, partial and full uploads and the device generator give the
    oracle's results (bit-exact), so the internal order never leaks;
  - the render buffers (positions, colours: fluidSimCPU.cc:100-125, cc:107)
    hold particle order, read straight from device memory.
"""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            lib = ctypes.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("libamdhip64.so not found")
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemcpy.restype = ctypes.c_int
    return lib


def _read_f4(ptr, n):
    out = np.empty((n, 4), np.float32)
    rc = _hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)  # device to host
    assert rc == 0, f"hipMemcpy failed: {rc}"
    return out


def _scene():
    return W.dam_break(n=8000)


def test_partial_upload_after_reordering_updates():
    sc = _scene()
    op = orc_params(sc.params)
    announce("partial upload", sc.name)
    with S.Sim(sc.params, sc.n) as sim:
        sim.upload(sc.pos, sc.vel)
        for _ in range(4):
            sim.update(DT)
        pos_g, vel_g, _ = sim.download()
        po, vo, _ = O.update(op, DT, sc.pos, sc.vel, T=sc.n, substeps=4)
        assert_bitexact(pos_g[:, :3], po[:, :3], "positions after 4 updates")
        # new velocities only: positions keep the simulated ones
        rng = np.random.default_rng(3)
        nv = vel_g.copy()
        nv[:, :3] = rng.uniform(-0.5, 0.5, (sc.n, 3)).astype(np.float32)
        sim.upload(None, nv)
        sim.update(DT)
        pos2, vel2, dens2 = sim.download()
        po2, vo2, do2 = O.update(op, DT, pos_g, nv, T=sc.n)
        assert_bitexact(pos2[:, :3], po2[:, :3], "positions after a velocity upload")
        assert_bitexact(vel2[:, :3], vo2[:, :3], "velocities after a velocity upload")
        assert_bitexact(dens2, do2, "densities after a velocity upload")
        # new positions only
        npos = pos2.copy()
        npos[:, :3] = npos[:, :3] + np.float32(0.01)
        sim.upload(npos, None)
        sim.update(DT)
        pos3, vel3, _ = sim.download()
        po3, vo3, _ = O.update(op, DT, npos, vel2, T=sc.n)
        assert_bitexact(pos3[:, :3], po3[:, :3], "positions after a position upload")
        assert_bitexact(vel3[:, :3], vo3[:, :3], "velocities after a position upload")


def test_full_upload_and_generate_reset_the_order():
    sc = _scene()
    op = orc_params(sc.params)
    announce("full upload", sc.name)
    with S.Sim(sc.params, sc.n) as sim:
        sim.upload(sc.pos, sc.vel)
        for _ in range(3):
            sim.update(DT)
        sim.upload(sc.pos, sc.vel)  # the initial state again
        sim.update(DT)
        pos_g, vel_g, dens_g = sim.download()
        po, vo, do = O.update(op, DT, sc.pos, sc.vel, T=sc.n)
        assert_bitexact(pos_g[:, :3], po[:, :3], "positions after re-upload")
        assert_bitexact(dens_g, do, "densities after re-upload")
        box = (4.0, 4.0, 4.0)
        sim.initialize(S.INIT_DAM_BREAK, 11, box, 0.1, 0.002, 0.0)
        pos_h, vel_h = S.generate(S.INIT_DAM_BREAK, 11, sc.n, box, 0.1, 0.002, 0.0)
        pos_i, vel_i, _ = sim.download()
        assert_bitexact(pos_i, pos_h, "generated positions")
        assert_bitexact(vel_i, vel_h, "generated velocities")


def test_render_buffers_hold_particle_order():
    sc = _scene()
    announce("render buffers", sc.name)
    with S.Sim(sc.params, sc.n) as sim:
        sim.upload(sc.pos, sc.vel)
        for _ in range(2):
            sim.update(DT)  # resident order is now the sorted one
        pos_ptr, col_ptr = sim.render_buffers()  # created from the current state
        assert pos_ptr and not col_ptr
        pos_g, _, _ = sim.download()
        assert_bitexact(_read_f4(pos_ptr, sc.n), pos_g, "render positions at creation")
        sim.set_colors(enable=True)
        pos_ptr2, col_ptr = sim.render_buffers()
        assert pos_ptr2 == pos_ptr and col_ptr
        for _ in range(3):
            sim.update(DT)
        pos_g, vel_g, _ = sim.download()
        assert_bitexact(_read_f4(pos_ptr, sc.n), pos_g, "render positions after updates")
        assert_bitexact(_read_f4(col_ptr, sc.n), O.speed_colors(vel_g), "render colours")
        assert np.all(pos_g[:, 3] == np.float32(0.34))
        # an upload rewrites the render positions at once
        sim.upload(sc.pos, sc.vel)
        assert_bitexact(_read_f4(pos_ptr, sc.n), sc.pos, "render positions after upload")


def test_changing_dt_and_settings_between_updates():
    """Each update's viscosity pass records the next update's buckets with its
    own dt; a different dt, new settings or a new state must not reuse them."""
    sc = _scene()
    announce("dt and settings changes", sc.name)
    dts = [DT, DT, 0.5 * DT, 0.5 * DT, DT, 0.0, DT]
    p2 = S.default_params(gravity_on=1, gravity_scale=3.0, viscosity=0.5, bounds=(4.0, 4.0, 4.0))
    with S.Sim(sc.params, sc.n) as sim:
        sim.upload(sc.pos, sc.vel)
        pos, vel = sc.pos, sc.vel
        for dt in dts:
            sim.update(dt)
            pos, vel, dens = O.update(orc_params(sc.params), dt, pos, vel, T=sc.n)
        pg, vg, dg = sim.download()
        assert_bitexact(pg[:, :3], pos[:, :3], "positions over changing dt")
        assert_bitexact(vg[:, :3], vel[:, :3], "velocities over changing dt")
        assert_bitexact(dg, dens, "densities over changing dt")
        sim.params = p2  # new settings between two updates with the same dt
        for _ in range(2):
            sim.update(DT)
            pos, vel, dens = O.update(orc_params(p2), DT, pos, vel, T=sc.n)
        pg, vg, dg = sim.download()
        assert_bitexact(pg[:, :3], pos[:, :3], "positions after a settings change")
        assert_bitexact(dg, dens, "densities after a settings change")
