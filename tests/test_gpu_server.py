"""Serving parity and admission: every request's result equals the same scene
run alone (bit for bit), whatever else is admitted, retired or cancelled
around it; admission respects slots and the HBM budget."""
import numpy as np
import pytest

import sph_amd as S
from helpers import announce, assert_bitexact
from sph_amd import workloads as W

pytestmark = pytest.mark.gpu


def _scene(n, seed, h=0.2, kind=S.INIT_DAM_BREAK):
    p, pos, vel = W.serving_scene(n, kind, seed, h)
    return p, pos, vel


def _alone(p, pos, vel, steps, table=0):
    with S.Sim(p, pos.shape[0], table) as sim:
        sim.upload(pos, vel)
        for _ in range(steps):
            sim.step(W.DT, W.SUBSTEPS)
        return sim.download()


def _drain(srv, ids, limit=2000):
    for _ in range(limit):
        if all(srv.status(i)[0] in (S.REQ_DONE, S.REQ_CANCELLED) for i in ids):
            return
        srv.tick()
    raise AssertionError("requests did not finish")


def test_requests_alone_equal_mixed():
    specs = [(6000, 3, 11, S.INIT_DAM_BREAK), (9000, 5, 12, S.INIT_CENTRED_BLOCK),
             (4200, 2, 13, S.INIT_DAM_BREAK), (12000, 4, 14, S.INIT_CENTRED_BLOCK)]
    scenes = [_scene(n, seed, kind=kind) for (n, _, seed, kind) in specs]
    announce("server mix", [s[0] for s in specs])
    with S.Server(max_running=3, initial_capacity=1 << 16) as srv:
        ids = [srv.submit(scenes[0][0], scenes[0][1], scenes[0][2], specs[0][1])]
        srv.tick()
        srv.tick()
        ids.append(srv.submit(scenes[1][0], scenes[1][1], scenes[1][2], specs[1][1]))
        ids.append(srv.submit(scenes[2][0], scenes[2][1], scenes[2][2], specs[2][1], table=777))
        srv.tick()
        ids.append(srv.submit(scenes[3][0], scenes[3][1], scenes[3][2], specs[3][1]))
        assert srv.status(ids[3])[0] == S.REQ_QUEUED      # three slots are taken
        _drain(srv, ids)
        results = [srv.fetch(i, specs[q][0]) for q, i in enumerate(ids)]
        st = srv.stats()
        assert st["finished"] == 4 and st["running"] == 0 and st["queued"] == 0
    for q, (n, steps, seed, kind) in enumerate(specs):
        announce("alone", q)
        ref = _alone(*scenes[q], steps, table=777 if q == 2 else 0)
        for g, a, what in zip(results[q], ref, ("pos", "vel", "dens")):
            assert_bitexact(g, a, f"request {q} {what}")


def test_budget_admission_and_refusal():
    a = _scene(5000, 21)
    b = _scene(7000, 22)
    budget = S.request_bytes(a[0], 5000) + S.request_bytes(b[0], 7000) - 1
    with S.Server(hbm_budget_bytes=budget, max_running=8, initial_capacity=1 << 16) as srv:
        ia = srv.submit(a[0], a[1], a[2], 2)
        ib = srv.submit(b[0], b[1], b[2], 2)
        srv.tick()
        assert srv.status(ia)[0] == S.REQ_RUNNING
        assert srv.status(ib)[0] == S.REQ_QUEUED          # does not fit beside a
        _drain(srv, [ia])
        srv.tick()
        assert srv.status(ib)[0] in (S.REQ_RUNNING, S.REQ_DONE)
        _drain(srv, [ib])
        big = _scene(20000, 23)
        with pytest.raises(S.SphError) as e:
            srv.submit(big[0], big[1], big[2], 1)
        assert e.value.code == S.SPH_ERR_BUDGET
        rb = srv.fetch(ib, 7000)
    ref = _alone(*b, 2)
    assert_bitexact(rb[0], ref[0], "b after waiting for budget")


def test_cancel_leaves_others_exact():
    sc = [_scene(n, 30 + i) for i, n in enumerate((5000, 6000, 7000))]
    with S.Server(max_running=2, initial_capacity=1 << 16) as srv:
        ids = [srv.submit(p, pos, vel, 3) for (p, pos, vel) in sc]
        srv.tick()
        srv.cancel(ids[0])                              # running
        srv.tick()
        _drain(srv, ids[1:])
        assert srv.stats()["cancelled"] == 1
        r1 = srv.fetch(ids[1], 6000)
        r2 = srv.fetch(ids[2], 7000)
        with pytest.raises(S.SphError):
            srv.status(ids[0])
    for r, s_ in ((r1, sc[1]), (r2, sc[2])):
        ref = _alone(*s_, 3)
        assert_bitexact(r[0], ref[0], "pos after a cancellation")
        assert_bitexact(r[2], ref[2], "dens after a cancellation")


def test_capacity_growth_keeps_results():
    sc = [_scene(n, 40 + i) for i, n in enumerate((50000, 60000, 70000))]
    with S.Server(max_running=4, initial_capacity=1 << 16) as srv:
        ids = [srv.submit(p, pos, vel, 2) for (p, pos, vel) in sc]
        _drain(srv, ids)
        assert srv.stats()["capacity"] >= 180000
        res = [srv.fetch(i, s_[1].shape[0]) for i, s_ in zip(ids, sc)]
    for r, s_ in zip(res, sc):
        ref = _alone(*s_, 2)
        assert_bitexact(r[1], ref[1], "vel after growth")
