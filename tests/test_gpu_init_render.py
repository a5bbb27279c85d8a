"""Device-side scene initialisers and the render hand-off, through the C ABI.

  - sph_initialize generates a scene on the GPU: positions and velocities must
    equal sph_generate's host bytes for every kind (bit-exact), and the
    densities of the InitializeData tail (cc:144-145) must equal the oracle's
    on that state;
  - sph_initialize_data (InitializeData, cc:112-147) equals the reference
    lattice (GridArrangement, cc:518-557);
  - batches and the server's generated requests reproduce the host-uploaded
    runs bit for bit;
  - the speed colours fused into integration equal the oracle's restatement of
    FluidSimCPU::updateColors (fluidSimCPU.cc:100-125) bit for bit, with the
    reference stops and with custom stops, and the render buffers expose
    OutPositions (w = 0.34) and the colours in particle order.
"""
import ctypes

import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

KINDS = [
    (S.INIT_GRID_ARRANGEMENT, 5000, (1.0, 1.0, 1.0), 0.0, 0.0, 0.0),
    (S.INIT_DAM_BREAK, 6000, (4.0, 4.0, 4.0), 0.1, 0.002, 0.0),
    (S.INIT_UNIFORM_LOWER, 7000, (6.0, 3.0, 6.0), 0.0, 0.0, 0.0),
    (S.INIT_TWO_BLOCKS, 8001, (8.0, 4.0, 8.0), 0.1, 0.002, 2.0),
    (S.INIT_CENTRED_BLOCK, 4913, (4.0, 4.0, 4.0), 0.1, 0.003, 0.0),
    # moving blocks that straddle |x| = 2: the densities of predictedPositions =
    # positions (cc:550) differ from those of positions + velocity / 120
    (S.INIT_TWO_BLOCKS, 6000, (4.4, 4.0, 4.4), 0.1, 0.002, 1.7),
]


@pytest.mark.parametrize("kind,n,box,spacing,jitter,speed", KINDS)
def test_device_initialise_equals_host_generator(kind, n, box, spacing, jitter, speed):
    announce("initialise", kind, n)
    p = W.scene_params(box=box)
    pos_h, vel_h = S.generate(kind, 17, n, box, spacing, jitter, speed)
    with S.Sim(p, n) as sim:
        sim.initialize(kind, 17, box, spacing, jitter, speed)
        pos_g, vel_g, dens_g = sim.download()
    assert_bitexact(pos_g, pos_h, f"positions kind {kind}")
    assert_bitexact(vel_g, vel_h, f"velocities kind {kind}")
    op = orc_params(p)
    sp = O.spatial(pos_h, n, op.h)
    assert_bitexact(dens_g, O.density(op, pos_h, sp), f"InitializeData densities kind {kind}")


def test_initialize_data_is_the_reference_lattice():
    n = 4096
    with S.Sim(S.default_params(), n) as sim:
        sim.initialize_data()
        pos_g, vel_g, _ = sim.download()
    assert_bitexact(pos_g, O.grid_arrangement(n), "GridArrangement")
    assert np.all(vel_g == 0)


def test_batch_initialise_equals_upload():
    specs = [KINDS[1], KINDS[3], KINDS[4]]
    params = [W.scene_params(box=k[2]) for k in specs]
    counts = [k[1] for k in specs]
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    got, want = [], []
    with S.Batch(params, counts) as b:
        for s, (kind, n, box, sp_, jit, spd) in enumerate(specs):
            b.initialize(s, kind, 5 + s, box, sp_, jit, spd)
        b.step(W.DT, W.SUBSTEPS)
        got = [b.download(s) for s in range(len(specs))]
    for s, (kind, n, box, sp_, jit, spd) in enumerate(specs):
        pos, vel = S.generate(kind, 5 + s, n, box, sp_, jit, spd)
        po, vo, do = O.update(orc_params(params[s]), dt, pos, vel, T=n, substeps=W.SUBSTEPS)
        assert_bitexact(got[s][0][:, :3], po[:, :3], f"scene {s} pos")
        assert_bitexact(got[s][1][:, :3], vo[:, :3], f"scene {s} vel")
        assert_bitexact(got[s][2], do, f"scene {s} dens")


def test_server_generated_requests_equal_uploaded_ones():
    kind, n, box, sp_, jit, spd = KINDS[1]
    p = W.scene_params(box=box)
    pos, vel = S.generate(kind, 9, n, box, sp_, jit, spd)
    announce("server generated", n)
    with S.Server(max_running=4, initial_capacity=1 << 15) as srv:
        a = srv.submit_generated(p, n, 3, kind, 9, box, sp_, jit, spd)
        b = srv.submit(p, pos, vel, 3)
        for _ in range(200):
            if all(srv.status(i)[0] == S.REQ_DONE for i in (a, b)):
                break
            srv.tick()
        ra, rb = srv.fetch(a, n), srv.fetch(b, n)
    for x, y, what in zip(ra, rb, ("pos", "vel", "dens")):
        assert_bitexact(x, y, f"generated vs uploaded {what}")


def test_speed_colours_fused_into_integration():
    sc = W.dam_break(n=16384)
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    announce("colours", sc.name)
    stops = np.arange(16, dtype=np.float32) / np.float32(16)
    with S.Sim(sc.params, sc.n) as sim:
        pos_ptr, col_ptr = sim.render_buffers()
        assert pos_ptr != 0 and col_ptr == 0
        sim.upload(sc.pos, sc.vel)
        sim.set_colors(enable=True)
        _, col_ptr = sim.render_buffers()
        assert col_ptr != 0
        for _ in range(5):
            sim.step(W.DT, W.SUBSTEPS)
        pos_g, vel_g, _ = sim.download()
        col_g = sim.download_colors()
        assert_bitexact(col_g, O.speed_colors(vel_g), "colours, reference stops")
        # custom stops, then colours off (the buffer keeps the last values) and on again
        sim.set_colors(stops, enable=True)
        sim.update(dt)
        _, vel2, _ = sim.download()
        assert_bitexact(sim.download_colors(), O.speed_colors(vel2, stops), "colours, custom stops")
        sim.set_colors(enable=False)
        sim.update(dt)
        assert_bitexact(sim.download_colors(), O.speed_colors(vel2, stops), "colours kept while off")
        sim.set_colors(enable=True)
        sim.update(dt)
        pos3, vel3, _ = sim.download()
        assert_bitexact(sim.download_colors(), O.speed_colors(vel3, stops), "colours on again")
        assert np.all(pos3[:, 3] == np.float32(0.34))
    # the speeds of a dam break after a few steps span more than one gradient segment
    speeds = np.linalg.norm(vel_g[:, :3], axis=1)
    assert speeds.max() > 0.33 * 1.5
