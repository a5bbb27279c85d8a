"""The batching and serving workloads at their own sizes.

  - batch_256 (256 scenes, 4,096-262,144 particles each, h in [0.1, 0.3],
    viscosity in [0, 0.2]; 15.25M particles in one engine): 3 steps batched;
    every scene equals the same scene stepped alone on its own handle, bit
    for bit, and a seeded sample of 8 scenes, the largest among them, equals
    the oracle.
  - serving: 16 bake requests drawn from the serving distribution (Poisson
    arrivals, 8,192-2,097,152 particles log-uniform, dam-break or centred
    block), submitted in two waves to a server with the 200 GiB budget and
    drained; steps are the drawn ones capped at 40 to bound the test's time.
    Every result equals the same request stepped alone, and the 4 smallest
    equal the oracle.
"""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu

BATCH_STEPS = 3
SERVE_STEPS_CAP = 40


def _sub_dt():
    return float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))


@pytest.fixture(scope="module", autouse=True)
def _threads():
    O.set_threads(16)


def _alone(p, pos, vel, steps, table=S.TABLE_GRID):
    with S.Sim(p, pos.shape[0], table) as sim:
        sim.upload(pos, vel)
        for _ in range(steps):
            sim.step(W.DT, W.SUBSTEPS)
        return sim.download()


def test_batch_256_equals_alone_and_oracle():
    bs = W.batch()
    scenes = bs.scenes
    assert len(scenes) == 256
    announce("batch_256", bs.total, "particles,", BATCH_STEPS, "steps")
    with S.Batch([sc.params for sc in scenes], [sc.n for sc in scenes]) as b:
        for s, sc in enumerate(scenes):
            b.upload(s, sc.pos, sc.vel)
        for _ in range(BATCH_STEPS):
            b.step(bs.dt, bs.substeps)
        got = [b.download(s) for s in range(len(scenes))]
    for s, sc in enumerate(scenes):
        alone = _alone(sc.params, sc.pos, sc.vel, BATCH_STEPS)
        for g, a, what in zip(got[s], alone, ("pos", "vel", "dens")):
            assert_bitexact(g, a, f"scene {s} (n={sc.n}) {what}: batch vs alone")
    announce("batch_256: all 256 scenes equal their runs alone")
    rng = np.random.default_rng(256)
    largest = int(np.argmax([sc.n for sc in scenes]))
    sample = [largest] + [int(x) for x in rng.choice(
        [s for s in range(len(scenes)) if s != largest], 7, replace=False)]
    dt = _sub_dt()
    for s in sample:
        sc = scenes[s]
        announce("oracle", "scene", s, "n", sc.n, "h", sc.meta["h"])
        pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel, T=sc.table,
                                        substeps=BATCH_STEPS * bs.substeps)
        assert_bitexact(got[s][0][:, :3], pos_o[:, :3], f"scene {s} pos vs oracle")
        assert_bitexact(got[s][1][:, :3], vel_o[:, :3], f"scene {s} vel vs oracle")
        assert_bitexact(got[s][2], dens_o, f"scene {s} dens vs oracle")


def test_serving_distribution_drained():
    arrivals = W.serving_requests(60.0, rate=20.0, seed=4)[:16]
    reqs = []
    for (_, n, steps, kind, seed) in arrivals:
        p, pos, vel = W.serving_scene(n, kind, seed)
        reqs.append((p, pos, vel, min(steps, SERVE_STEPS_CAP), kind))
    sizes = [r[1].shape[0] for r in reqs]
    kinds = {r[4] for r in reqs}
    assert kinds == {S.INIT_DAM_BREAK, S.INIT_CENTRED_BLOCK}
    assert min(sizes) >= 8192 and max(sizes) <= 2097152 and max(sizes) > 500000
    announce("serving", len(reqs), "requests", sizes)
    with S.Server(hbm_budget_bytes=200 << 30, max_running=64, initial_capacity=1 << 22) as srv:
        ids = [srv.submit(p, pos, vel, st) for (p, pos, vel, st, _) in reqs[:8]]
        for _ in range(5):
            srv.tick()
        ids += [srv.submit(p, pos, vel, st) for (p, pos, vel, st, _) in reqs[8:]]
        for _ in range(100000):
            if all(srv.status(i)[0] == S.REQ_DONE for i in ids):
                break
            srv.tick()
        else:
            raise AssertionError("requests did not finish")
        st = srv.stats()
        assert st["finished"] == len(reqs) and st["running"] == 0 and st["queued"] == 0
        results = [srv.fetch(i, r[1].shape[0]) for i, r in zip(ids, reqs)]
    for q, (p, pos, vel, steps, _) in enumerate(reqs):
        ref = _alone(p, pos, vel, steps)
        for g, a, what in zip(results[q], ref, ("pos", "vel", "dens")):
            assert_bitexact(g, a, f"request {q} (n={pos.shape[0]}) {what}: served vs alone")
    announce("serving: all", len(reqs), "requests equal their runs alone")
    dt = _sub_dt()
    for q in np.argsort(sizes)[:4]:
        p, pos, vel, steps, _ = reqs[q]
        announce("oracle", "request", int(q), "n", pos.shape[0], "steps", steps)
        pos_o, vel_o, dens_o = O.update(orc_params(p), dt, pos, vel, T=pos.shape[0],
                                        substeps=steps * W.SUBSTEPS)
        assert_bitexact(results[q][0][:, :3], pos_o[:, :3], f"request {q} pos vs oracle")
        assert_bitexact(results[q][1][:, :3], vel_o[:, :3], f"request {q} vel vs oracle")
        assert_bitexact(results[q][2], dens_o, f"request {q} dens vs oracle")
