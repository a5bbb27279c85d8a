"""Serving parity and admission: every request's result equals the same scene
run alone (bit for bit), whatever else is admitted, retired or cancelled
around it; admission respects slots and the HBM budget."""
import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
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
    # and the served results equal the oracle (requests admitted mid-run and
    # queued behind the slots, one with its own bucket count)
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    for q in (1, 2):
        p, pos, vel = scenes[q]
        pos_o, vel_o, dens_o = O.update(orc_params(p), dt, pos, vel, T=specs[q][0],
                                        substeps=specs[q][1] * W.SUBSTEPS)
        assert_bitexact(results[q][0][:, :3], pos_o[:, :3], f"request {q} pos vs oracle")
        assert_bitexact(results[q][1][:, :3], vel_o[:, :3], f"request {q} vel vs oracle")
        assert_bitexact(results[q][2], dens_o, f"request {q} dens vs oracle")


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


def test_failed_growth_leaves_admission_undone():
    """Admissions that need more device slots than the layout may grow to
    (max_capacity, as when device memory runs out) change nothing: the
    request waits at the head of the queue with the budget uncharged, the
    running request goes on; once it is done the queued one fits and runs,
    and both equal their runs alone.  With nothing running, a request that
    cannot fit fails the tick with SPH_ERR_OOM."""
    a = _scene(40000, 51)
    b = _scene(50000, 52)
    c = _scene(70000, 53)
    with S.Server(max_running=4, initial_capacity=1 << 16, max_capacity=1 << 16) as srv:
        ia = srv.submit(a[0], a[1], a[2], 2)
        srv.tick()
        assert srv.status(ia)[0] == S.REQ_RUNNING
        used = srv.stats()["used_bytes"]
        ib = srv.submit(b[0], b[1], b[2], 2)
        done_before = srv.status(ia)[1]
        srv.tick()                                        # 90,000 particles > 65,536 slots
        assert srv.status(ib)[0] == S.REQ_QUEUED
        assert srv.status(ia)[1] == done_before + 1       # a went on
        st = srv.stats()
        assert st["admission_failures"] >= 1
        assert st["used_bytes"] == used and st["admitted"] == 1 and st["queued"] == 1
        _drain(srv, [ia, ib])                             # b runs once a has left
        ra = srv.fetch(ia, 40000)
        with pytest.raises(S.SphError) as e3:
            srv.fetch(ia, 40000)                          # forgotten once fetched
        assert e3.value.code == S.SPH_ERR_NOT_FOUND
        with pytest.raises(S.SphError) as e4:
            srv.fetch(ib, 49999)                          # wrong particle count: refused
        assert e4.value.code == S.SPH_ERR_INVALID
        rb = srv.fetch(ib, 50000)
        ic = srv.submit(c[0], c[1], c[2], 1)              # larger than the layout can ever be
        with pytest.raises(S.SphError) as e5:
            srv.tick()
        assert e5.value.code == S.SPH_ERR_OOM
        assert srv.status(ic)[0] == S.REQ_QUEUED
        srv.cancel(ic)
        assert srv.stats()["queued"] == 0
    assert_bitexact(ra[0], _alone(*a, 2)[0], "a beside a failed admission")
    assert_bitexact(rb[1], _alone(*b, 2)[1], "b after waiting for room")


def test_generated_requests_admitted_midrun_equal_alone():
    # device-generated requests admitted at staggered ticks while others run
    # and leave: every repack is device-only (moves, generation, the scene
    # map filled on the device, descriptors through the staging ring, which
    # wraps), with no stream synchronisation
    specs = [(4096, 3, S.INIT_DAM_BREAK, 0), (6001, 2, S.INIT_CENTRED_BLOCK, 0),
             (9000, 5, S.INIT_DAM_BREAK, 2), (4500, 2, S.INIT_CENTRED_BLOCK, 3),
             (12000, 4, S.INIT_CENTRED_BLOCK, 5), (5003, 3, S.INIT_DAM_BREAK, 7),
             (7777, 2, S.INIT_DAM_BREAK, 8), (20000, 3, S.INIT_CENTRED_BLOCK, 11),
             (4099, 4, S.INIT_DAM_BREAK, 13), (8192, 2, S.INIT_CENTRED_BLOCK, 17)]
    announce("generated mid-run", len(specs), [s[0] for s in specs])
    plans = [W.serving_plan(n, kind, 500 + q) for q, (n, _, kind, _) in enumerate(specs)]
    ids = [None] * len(specs)
    with S.Server(hbm_budget_bytes=8 << 30, max_running=16, initial_capacity=1 << 17) as srv:
        for tick in range(200):
            for q, (n, steps, kind, at) in enumerate(specs):
                if at == tick:
                    p, gen = plans[q]
                    ids[q] = srv.submit_generated(p, n, steps, kind, 500 + q, *gen)
            if tick > max(s[3] for s in specs) and all(srv.status(i)[0] == S.REQ_DONE for i in ids):
                break
            srv.tick()
        else:
            raise AssertionError("requests did not finish")
        got = [srv.fetch(i, n) for i, (n, _, _, _) in zip(ids, specs)]
    for q, (n, steps, kind, _) in enumerate(specs):
        p, gen = plans[q]
        with S.Sim(p, n, S.TABLE_GRID) as sim:
            sim.initialize(kind, 500 + q, *gen)
            for _ in range(steps):
                sim.step(W.DT, W.SUBSTEPS)
            alone = sim.download()
        for g, a, what in zip(got[q], alone, ("pos", "vel", "dens")):
            assert_bitexact(g, a, f"request {q} (n={n}) {what}: served vs alone")


def test_tick_log_counts_particles_and_membership():
    a = _scene(5000, 31)
    b = _scene(7000, 32)
    announce("tick log", (5000, 7000))
    with S.Server(max_running=4, initial_capacity=1 << 16) as srv:
        ia = srv.submit(a[0], a[1], a[2], 2)    # 6 updates
        ib = srv.submit(b[0], b[1], b[2], 4)    # 12 updates
        _drain(srv, [ia, ib])
        ms, parts, lay = srv.tick_log()
        assert srv.stats()["ticks"] == len(ms) == 12
    assert (ms > 0).all()
    assert parts.tolist() == [12000] * 6 + [7000] * 6
    # both admitted by one repack; the departure of the first is the next
    assert len(set(lay[:6].tolist())) == 1 and len(set(lay[6:].tolist())) == 1
    assert lay[6] == lay[0] + 1
