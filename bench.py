#!/usr/bin/env python3
"""Benchmark of the gfx950 SPH step (one MI355X).

Headline (`metric`, `value`): particle-steps/s of the 256-scene batching
workload batch_256 -- 256 independent scenes, 4,096-262,144 particles each
(log-uniform from a seed; 15,253,915 in all), h in [0.1, 0.3], viscosity in
[0, 0.2], dam-break blocks at 40% fill -- stepped batched: one launch per
kernel for all scenes, graph-replayed.  Beside it, on the same workload, the
plain one-scene-per-launch GPU path (every scene on its own handle and
stream, all enqueued before any is waited for) and the same launches one
scene at a time, with p50/p99 step latency.  A step is one sph_step(1/120,
3): three Update(dt/3) calls of the reference (physicsWorld.cc:39-111).

Further legs (each measured here, on the same box, in the same run):
  single_scene        uniform_1m: 1,048,576 particles uniform in the lower
                      half of a 16x8x16 box, cell-grid key, and the same
                      under the reference key with its 2^21-bucket table;
  large_scene         two_blocks_16m: 16,777,216 particles in two colliding
                      blocks in a 40x20x40 box;
  dam_break_16k       16,384 particles, 200 steps (the CPU-reference config);
  serve               Poisson bake requests (20/s, 8,192-2,097,152 particles,
                      100-2,000 steps) by continuous batching, beside the
                      plain path on the same arrivals (handles' steps
                      overlapped, and one request at a time), and a
                      sustainable load.

`roofline` is computed for the headline's dominant kernel from its mean
duration, measured live over the timed region with HIP events recorded on the
batch's own stream.  `cpu_baseline` times the CPU oracle (a restatement of the
reference step, stage B as the reference runs it: std::sort of 12-byte float
records) on a bounded sample of the same workload, its scenes one after
another as the singleton reference runs them; `cpu_baseline_c0` follows the
reference CPU protocol on the 16,384-particle dam break (5 warm-up steps, 20
timed, the six stage timers).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluid-simulation-3d_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import sph_amd as S  # noqa: E402
from sph_amd import workloads as W  # noqa: E402

METRIC = "particle-steps/s batched vs single-scene on 1 MI355X; p50/p99 step latency"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = 157.3
# Algorithmic HBM bytes per particle per launch of each timed stage: each
# array read or written once, neighbour re-reads counted as on-chip.
#   predict:   R pos, vel (16+16), W key, ticket (4+4)
#   spatial:   scatter (R key, ticket, id; W order pair: 20) + reorder (R order
#              pair, key, pos, vel; W sids, xps, vs, ps: 96) -- counted per kernel
#   density:   R xps (16), W rho, rho_n, 1/rho, 1/rho_n (16); bucket table reads ~0
#   pressure:  R xps, dens, v (48), W v (16)
#   viscosity+integrate: R xps, v, pos, id (52), W pos, vel, id (36), and the
#     next update's bucket key and ticket (8)
ALG_BYTES = {"predict": 40.0, "density": 32.0, "pressure": 64.0, "viscosity_integrate": 96.0}
STAGES = ["predict", "spatial", "density", "pressure", "viscosity_integrate"]
GATHER = ["density", "pressure", "viscosity_integrate"]
PROFILES = os.path.join(ROOT, "profiles")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sub_dt(dt=W.DT, substeps=W.SUBSTEPS):
    return float(np.float32(np.float32(dt) / np.float32(substeps)))


def kernel_name(stage, variants, grid, multi=False):
    """Template instantiation that runs `stage` under the gather variants `variants`
    (kernel<MULTI, WORK, LIST, NBR, GRID>)."""
    m = "true" if multi else "false"
    b = lambda x: "true" if x else "false"  # noqa: E731
    if stage == "density" and variants["nbr"] and not grid:
        return f"k_density_rows<{m}>"
    base = {"density": "k_density", "pressure": "k_pressure",
            "viscosity_integrate": "k_viscosity_integrate"}[stage]
    v = variants[stage.split("_")[0]]
    if grid:
        if stage == "density":
            return f"{base}<{m}, false, {b(v['list'])}, {b(variants['nbr'])}, true>"
        if variants["nbr"] and v["rows"]:
            return f"{base}<{m}, false, false, true, true>"
        return f"{base}<{m}, false, {b(v['list'])}, false, true>"
    if variants["nbr"] and v["rows"]:
        return f"{base}<{m}, true, false, true, false>"
    return f"{base}<{m}, {b(v['work'])}, {b(v['list'])}, false, false>"


# committed counter summaries are named for the workload's short name
PROFILE_FILES = {"batch_256": "batch_256", "uniform_1048576": "uniform_1m",
                 "two_blocks_16777216": "two_blocks_16m"}


def _profile(name):
    try:
        with open(os.path.join(PROFILES, name)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    `workload` (profiles/<short name>_traffic.json), or None."""
    t = _profile(f"{PROFILE_FILES.get(workload, workload)}_traffic.json")
    if not t or t.get("workload") != workload:
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else float(k["traffic_bytes"])


def valu_issue(workload, kernel, ms):
    """VALU issue fraction of `kernel` at duration `ms`, from its committed
    instruction count (rocprofv3 SQ_INSTS_VALU): instructions x 2 cycles (a
    wave64 VALU instruction on a 32-lane SIMD) / (1024 SIMDs x 2.4 GHz x ms)."""
    t = _profile(f"{PROFILE_FILES.get(workload, workload)}_valu.json")
    if not t or t.get("workload") != workload:
        return None
    k = t.get("kernels", {}).get(kernel)
    if k is None:
        return None
    instr = float(k["valu_instr_per_launch"])
    cyc = t["cycles_per_wave64_valu"]
    return {"kernel": kernel, "instr_per_launch": instr,
            "issue_frac": instr * cyc / (t["simds"] * t["clock_ghz"] * 1e9 * ms * 1e-3),
            "lane_use": k.get("lane_use"), "wait_frac": k.get("wait_frac"),
            "mfma_instr_per_launch": k.get("mfma_instr_per_launch"),
            "model": f"{cyc} cycles per wave64 VALU instruction, {t['simds']} SIMDs, "
                     f"{t['clock_ghz']} GHz"}


def gather_fractions(workload, n, variants, grid, per, multi):
    """HBM and VALU fractions of each gather pass at its live duration:
    algorithmic bytes / duration / 8 TB/s, the committed measured traffic /
    duration / 8 TB/s, and the VALU issue share from the committed
    instruction count."""
    out = {}
    for stage in GATHER:
        kern = kernel_name(stage, variants, grid, multi)
        ms = per[stage]
        traffic = measured_traffic(workload, kern)
        valu = valu_issue(workload, kern, ms)
        out[stage] = {"kernel": kern, "ms": ms,
                      "alg_hbm_frac": ALG_BYTES[stage] * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "traffic_hbm_frac": None if traffic is None
                      else traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "traffic_over_alg": None if traffic is None else traffic / (ALG_BYTES[stage] * n),
                      "valu_issue_frac": None if valu is None else valu["issue_frac"]}
    return out


def roofline(workload, n, variants, grid, per, multi):
    """Roofline of the dominant gather kernel against HBM.  No MFMA work
    exists in this path (no dense contraction; the committed counters show
    zero MFMA instructions), so the roofline is HBM's; `limiter` names what
    the committed counters show bounds the kernel at its live duration: the
    larger of its VALU issue share and its measured-traffic share of HBM."""
    dom = max(GATHER, key=lambda k: per[k])
    kern = kernel_name(dom, variants, grid, multi)
    ms = per[dom]
    ach = ALG_BYTES[dom] * n / (ms * 1e-3) / 1e9
    traffic = measured_traffic(workload, kern)
    valu = valu_issue(workload, kern, ms)
    shares = {}
    if valu is not None:
        shares["valu_issue"] = valu["issue_frac"]
    if traffic is not None:
        shares["hbm_traffic"] = traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    mfma = None if valu is None else valu.get("mfma_instr_per_launch")
    return {"bound": "mfma" if mfma else "hbm", "kernel": kern, "stage": dom,
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": traffic, "ms_per_launch": ms,
            "alg_bytes_per_launch": ALG_BYTES[dom] * n, "alg_bytes_per_particle": ALG_BYTES[dom],
            "limiter": max(shares, key=shares.get) if shares else None, "shares": shares,
            "valu": valu}


def step_roofline(particle_steps_per_s, substeps, buckets_per_particle):
    """The whole step against HBM with the reference's stage traffic per
    particle-substep (each array read or written once per stage A-F,
    neighbour re-reads on chip): 264 + 4 T/N bytes, T/N the bucket count per
    particle (stage B1 clears the table)."""
    b = 264.0 + 4.0 * buckets_per_particle
    ach = b * particle_steps_per_s * substeps / 1e9
    return {"alg_bytes_per_particle_substep": b, "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}


# ---------------------------------------------------------------------------
# GPU legs
# ---------------------------------------------------------------------------
def bench_batch(steps, warmup, plain_steps, latency_steps, plain=True):
    """Batched (one launch per kernel for all scenes, graph-replayed) vs plain
    (one handle per scene) on batch_256."""
    bs = W.batch()
    total = bs.total
    log(f"[bench] batch_256: {len(bs.scenes)} scenes, {total} particles")
    b = S.Batch([sc.params for sc in bs.scenes], [sc.n for sc in bs.scenes])
    for s_, sc in enumerate(bs.scenes):
        b.upload(s_, sc.pos, sc.vel)
    for _ in range(max(1, warmup)):
        b.step_async(bs.dt, bs.substeps)
    b.sync()
    tuned = b.autotune(sub_dt(bs.dt, bs.substeps), updates=3)
    log(f"[bench] batch autotune: {b.variants} ms {np.round(tuned, 4)}")
    for _ in range(max(1, warmup)):
        b.step_async(bs.dt, bs.substeps)
    b.sync()
    # timed region: K steps, stage events on the batch's stream
    b.profile_begin(steps * bs.substeps)
    b.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step_async(bs.dt, bs.substeps)
    b.sync()
    el_b = time.perf_counter() - t0
    stage_ms, _ = b.profile_end()
    lat = []
    for _ in range(latency_steps):
        t = time.perf_counter()
        b.step(bs.dt, bs.substeps)
        lat.append((time.perf_counter() - t) * 1e3)
    variants = b.variants
    b.close()
    log(f"[bench] batched {steps} steps in {el_b:.3f} s")

    per = {STAGES[i]: float(stage_ms[i]) for i in range(5)}
    v_b = total * steps / el_b
    out = {"workload": bs.name, "scenes": len(bs.scenes), "particles": total,
           "value": v_b, "ms_per_step": el_b * 1e3 / steps, "steps": steps,
           "step_latency_ms": {"p50": float(np.percentile(lat, 50)),
                               "p99": float(np.percentile(lat, 99)), "n": len(lat)},
           "stage_ms_per_update": per, "gather_variants": variants,
           "roofline": roofline(bs.name, total, variants, True, per, True),
           "gather_kernels": gather_fractions(bs.name, total, variants, True, per, True)}
    if not plain:
        return out
    sims = []
    for sc in bs.scenes:
        sim = S.Sim(sc.params, sc.n, S.TABLE_GRID)
        sim.upload(sc.pos, sc.vel)
        sims.append(sim)
    for sim in sims:
        sim.step_async(bs.dt, bs.substeps)
    for sim in sims:
        sim.sync()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for sim in sims:
            sim.step_async(bs.dt, bs.substeps)
        for sim in sims:
            sim.sync()
    el_p = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for sim in sims:
            sim.step(bs.dt, bs.substeps)
    el_s = time.perf_counter() - t0
    for sim in sims:
        sim.close()
    log(f"[bench] plain {plain_steps} steps in {el_p:.3f} s (one at a time {el_s:.3f} s)")
    v_p = total * plain_steps / el_p
    v_s = total * plain_steps / el_s
    out.update({
            "plain_one_scene_per_launch": {
                "value": v_p, "ms_per_step": el_p * 1e3 / plain_steps, "steps": plain_steps,
                "note": "every scene on its own handle and stream, all enqueued before any is "
                        "waited for (scenes overlap)"},
            "plain_one_scene_at_a_time": {
                "value": v_s, "ms_per_step": el_s * 1e3 / plain_steps, "steps": plain_steps,
                "note": "the same launches, each scene's step finished before the next starts"},
            "speedup_vs_plain": v_b / v_p, "speedup_vs_one_at_a_time": v_b / v_s})
    return out


def bench_scene(sc, steps, warmup, key="grid", latency_steps=20, autotune=True, generated=None):
    """One scene on its own handle: K timed steps with stage events."""
    table = S.TABLE_GRID if key == "grid" else sc.table
    sim = S.Sim(sc.params, sc.n, table)
    grid, gdim, _ = sim.key()
    if generated is not None:
        sim.initialize(*generated)
    else:
        sim.upload(sc.pos, sc.vel)
    for _ in range(max(1, warmup)):
        sim.step_async(sc.dt, sc.substeps)
    sim.sync()
    if autotune:
        tuned = sim.autotune(sub_dt(sc.dt, sc.substeps), updates=3)
        log(f"[bench] {sc.name} ({key}) autotune: ms {np.round(tuned, 4)}")
        for _ in range(2):
            sim.step_async(sc.dt, sc.substeps)
        sim.sync()
    sim.profile_begin(steps * sc.substeps)
    sim.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step_async(sc.dt, sc.substeps)
    sim.sync()
    el = time.perf_counter() - t0
    stage_ms, _ = sim.profile_end()
    lat = []
    for _ in range(latency_steps):
        t = time.perf_counter()
        sim.step(sc.dt, sc.substeps)
        lat.append((time.perf_counter() - t) * 1e3)
    nonfinite = sim.count_nonfinite()
    variants = sim.variants
    buckets = sim.table
    sim.close()
    per = {STAGES[i]: float(stage_ms[i]) for i in range(5)}
    log(f"[bench] {sc.name} ({key}): {steps} steps in {el:.3f} s")
    out = {"workload": sc.name, "particles": sc.n, "value": sc.n * steps / el,
           "ms_per_step": el * 1e3 / steps, "steps": steps,
           "bucket_key": ("cell grid %dx%dx%d" % tuple(gdim)) if grid
           else "reference hash %% %d" % sc.table, "buckets": buckets,
           "step_latency_ms": {"p50": float(np.percentile(lat, 50)),
                               "p99": float(np.percentile(lat, 99)), "n": len(lat)},
           "stage_ms_per_update": per, "gather_variants": variants,
           "nonfinite_components": nonfinite}
    out["gather_kernels"] = gather_fractions(sc.name, sc.n, variants, grid, per, False)
    out["step_roofline"] = step_roofline(out["value"], sc.substeps, buckets / sc.n)
    return out


def bench_serve(duration, rate, max_running, budget_gib, seed=4):
    """Poisson arrivals of bake requests for `duration` seconds of wall time,
    served by continuous batching (re-batched every substep).  Requests are
    generated on the device at admission (the server's generator writes the
    bytes sph_generate would).  Requests still queued or running at the end
    count towards nothing but the particle-steps they already advanced."""
    arrivals = W.serving_requests(duration, rate=rate, seed=seed)
    srv = S.Server(hbm_budget_bytes=budget_gib << 30, max_running=max_running,
                   initial_capacity=1 << 25, updates_per_tick=1)
    log(f"[bench] serve: {len(arrivals)} arrivals over {duration} s at {rate}/s, slots {max_running}")
    nxt, submitted = 0, 0
    offered = sum(n * st for (_, n, st, _, _) in arrivals) / duration
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        if now >= duration:
            break
        while nxt < len(arrivals) and arrivals[nxt][0] <= now:
            _, n, steps, kind, sd = arrivals[nxt]
            p, gen = W.serving_plan(n, kind, sd)
            srv.submit_generated(p, n, steps, kind, sd, *gen)
            submitted += 1
            nxt += 1
        info = srv.tick()
        if not info["particles"] and not info["queued"] and nxt < len(arrivals):
            time.sleep(max(0.0, min(0.01, arrivals[nxt][0] - (time.perf_counter() - t0))))
    st = srv.stats()  # waits for the enqueued ticks
    wall = time.perf_counter() - t0
    spread = tick_spread(*srv.tick_log())
    srv.close()
    psteps = st["particle_updates"] / 3.0
    return {"arrivals_per_s": rate, "window_s": duration, "submitted": submitted,
            "admitted": st["admitted"], "finished": st["finished"], "queued_at_end": st["queued"],
            "running_at_end": st["running"], "slots": max_running, "hbm_budget_gib": budget_gib,
            "value": psteps / wall, "unit": "particle-steps/s", "wall_s": wall,
            "offered_particle_steps_per_s": offered, "gpu_busy_frac": st["busy_s"] / wall,
            # one step = 3 ticks (each tick is one Update of every running request)
            "step_latency_ms": {"p50": 3 * st["tick_ms_p50"], "p99": 3 * st["tick_ms_p99"]},
            "substep_tick_ms": {"p50": st["tick_ms_p50"], "p99": st["tick_ms_p99"],
                                "n": st["ticks"]},
            "request_latency_ms": {"p50": st["request_ms_p50"] if st["finished"] else None,
                                   "p99": st["request_ms_p99"] if st["finished"] else None,
                                   "n": st["finished"]},
            "queue_wait_ms": {"p50": st["wait_ms_p50"], "p99": st["wait_ms_p99"]},
            "step_latency_spread": spread,
            "device_capacity_particles": st["capacity"]}


def tick_spread(ms, particles, layout):
    """Where the spread of step latency comes from, from the server's tick
    log.  within_membership: each tick over the median tick of its layout
    (ticks between two membership changes advance the same requests), p50
    and p99 of that ratio over layouts of 10 ticks or more.  tick_to_tick:
    the factor between consecutive ticks of one layout (larger over smaller),
    p50 and p99; it separates jitter from the slow change of a scene's cost as
    it evolves.  by_load: p50/p99 step latency (3 ticks) of the ticks in each
    range of particles advanced."""
    if len(ms) == 0:
        return None
    ms = ms.astype(np.float64)
    same = (layout[1:] == layout[:-1]) & (ms[1:] > 0) & (ms[:-1] > 0)
    step = np.maximum(ms[1:], ms[:-1])[same] / np.minimum(ms[1:], ms[:-1])[same]
    rel = []
    for q in np.unique(layout):
        t = ms[layout == q]
        if len(t) >= 10:
            rel.append(t / np.median(t))
    rel = np.concatenate(rel) if rel else np.zeros(0)
    by_load = []
    edges = [0, 1 << 18, 1 << 20, 1 << 22, 1 << 24, 1 << 62]
    for lo, hi in zip(edges[:-1], edges[1:]):
        t = ms[(particles >= lo) & (particles < hi)]
        if len(t):
            by_load.append({"particles": [lo, hi if hi < (1 << 62) else None], "ticks": int(len(t)),
                            "step_ms_p50": 3 * float(np.percentile(t, 50)),
                            "step_ms_p99": 3 * float(np.percentile(t, 99))})
    return {"ticks": int(len(ms)), "layouts": int(len(np.unique(layout))),
            "within_membership": None if not len(rel) else
            {"ticks": int(len(rel)), "p50": float(np.percentile(rel, 50)),
             "p99": float(np.percentile(rel, 99))},
            "tick_to_tick": None if not len(step) else
            {"pairs": int(len(step)), "p50": float(np.percentile(step, 50)),
             "p99": float(np.percentile(step, 99))},
            "particles_per_tick": {"p50": float(np.percentile(particles, 50)),
                                   "p99": float(np.percentile(particles, 99))},
            "by_load": by_load}


def bench_serve_plain(duration, rate, max_running, budget_gib, seed=4, overlap=True):
    """The plain one-scene-per-launch path on the same arrivals: each admitted
    request on its own handle (created at admission, state generated on the
    device), the same FIFO admission by slots and HBM budget.  overlap: every
    running handle's next step enqueued before any is waited for (the
    handles' streams overlap); else each request's step runs to completion
    before the next request's is enqueued (one scene at a time)."""
    arrivals = W.serving_requests(duration, rate=rate, seed=seed)
    budget = budget_gib << 30
    queue, running = [], []  # running: [sim, steps_left, n, bytes]
    used, nxt, psteps, finished = 0, 0, 0.0, 0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        if now >= duration:
            break
        while nxt < len(arrivals) and arrivals[nxt][0] <= now:
            queue.append(arrivals[nxt])
            nxt += 1
        while queue and len(running) < max_running:
            _, n, steps, kind, sd = queue[0]
            p, gen = W.serving_plan(n, kind, sd)
            nb = S.request_bytes(p, n)
            if used + nb > budget:
                break
            queue.pop(0)
            sim = S.Sim(p, n, S.TABLE_GRID)
            sim.initialize(kind, sd, *gen)
            running.append([sim, steps, n, nb])
            used += nb
        if not running:
            if nxt < len(arrivals):
                time.sleep(max(0.0, min(0.01, arrivals[nxt][0] - (time.perf_counter() - t0))))
            continue
        if overlap:
            for r in running:
                r[0].step_async(W.DT, W.SUBSTEPS)
        for r in running:
            if not overlap:
                r[0].step_async(W.DT, W.SUBSTEPS)
            r[0].sync()
            r[1] -= 1
            psteps += r[2]
        for r in [r for r in running if r[1] <= 0]:
            r[0].download()
            r[0].close()
            used -= r[3]
            finished += 1
            running.remove(r)
    wall = time.perf_counter() - t0
    for r in running:
        r[0].close()
    return {"arrivals_per_s": rate, "window_s": duration, "finished": finished,
            "running_at_end": len(running), "queued_at_end": len(queue),
            "value": psteps / wall, "unit": "particle-steps/s", "wall_s": wall,
            "note": "each request on its own handle, one step of every running request enqueued "
                    "per round" + ("" if overlap else ", each waited for before the next is enqueued")}


# ---------------------------------------------------------------------------
# CPU baselines
# ---------------------------------------------------------------------------
def host_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_batch(n_scenes, threads, seed=7):
    """The oracle on a seeded sample of batch_256's scenes, one after another
    (the reference simulates one scene per process), one step (3 Update
    calls) each after one warm-up Update, stage B as the reference runs it."""
    import pyoracle as O
    bs = W.batch()
    rng = np.random.default_rng(seed)
    pick = sorted(int(x) for x in rng.choice(len(bs.scenes), n_scenes, replace=False))
    O.set_threads(threads)
    O.set_reference_sort(True)
    el, parts = 0.0, 0
    try:
        for s in pick:
            sc = bs.scenes[s]
            op = O.from_sph(sc.params)
            dt = sub_dt(sc.dt, sc.substeps)
            pos, vel, _ = O.update(op, dt, sc.pos, sc.vel, T=sc.table)      # warm-up
            t0 = time.perf_counter()
            O.update(op, dt, pos, vel, T=sc.table, substeps=sc.substeps)
            el += time.perf_counter() - t0
            parts += sc.n
    finally:
        O.set_reference_sort(False)
        O.set_threads(1)
    where = "1 host thread" if threads == 1 else f"{threads} host threads"
    return {"value": parts / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{n_scenes} scenes of batch_256 (seed {seed}: {parts} particles), one after "
                      f"another, 1 step = 3 Update(dt/3) each after 1 warm-up Update, on {where}, "
                      f"{el:.1f} s timed; stage B as std::sort of 12-byte float records"}


def cpu_baseline_c0(threads, warmup_steps=5, steps=20):
    """The reference CPU protocol on dam_break_16k: 5 warm-up steps, 20 timed
    steps of 3 Update(dt/3), the six stage timers (physicsWorld.h:45-50)."""
    import pyoracle as O
    sc = W.dam_break(n=16384)
    op = O.from_sph(sc.params)
    dt = sub_dt()
    O.set_threads(threads)
    O.set_reference_sort(True)
    try:
        pos, vel, _ = O.update(op, dt, sc.pos, sc.vel, T=sc.n, substeps=warmup_steps * 3)
        ms = np.zeros(6, np.float64)
        t0 = time.perf_counter()
        O.update(op, dt, pos, vel, T=sc.n, substeps=steps * 3, stage_ms=ms)
        el = time.perf_counter() - t0
    finally:
        O.set_reference_sort(False)
        O.set_threads(1)
    names = ["gravity", "spatial", "density", "pressure", "viscosity", "position_collision"]
    return {"value": sc.n * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "workload": sc.name, "steps": steps, "warmup_steps": warmup_steps,
            "stage_ms_per_update": {k: float(v) / (steps * 3) for k, v in zip(names, ms)},
            "sample": f"{steps} steps (x3 Update) of {sc.name} after {warmup_steps} warm-up steps "
                      f"on {threads} host thread(s), {el:.1f} s; stage B as std::sort of 12-byte "
                      f"float records"}


LEGS = ["uniform_1m_grid", "uniform_1m_reference_key", "two_blocks_16m", "dam_break_16k"]


def run_serve(args):
    """The serving workload, configured and plain on the same arrivals, then
    the sustainable rate (--serve-sustained-seconds 0 skips it)."""
    configured = bench_serve(args.serve_seconds, args.serve_rate, args.serve_slots, 200)
    plain = None if args.no_serve_plain else bench_serve_plain(args.serve_seconds, args.serve_rate,
                                                                args.serve_slots, 200)
    serial = None if args.no_serve_plain else bench_serve_plain(args.serve_seconds, args.serve_rate,
                                                                 args.serve_slots, 200, overlap=False)
    return {
        "workload": "serving (Poisson bake requests, 8,192-2,097,152 particles, 100-2,000 steps)",
        "configured": configured, "plain_same_arrivals": plain,
        "plain_one_at_a_time_same_arrivals": serial,
        "speedup_vs_plain": configured["value"] / plain["value"] if plain and plain["value"] else None,
        "speedup_vs_one_at_a_time": configured["value"] / serial["value"] if serial and serial["value"]
        else None,
        "sustained": None if args.serve_sustained_seconds <= 0 else
        bench_serve(args.serve_sustained_seconds, args.serve_sustained_rate, args.serve_slots, 200, seed=5)}


def run_legs(names, args):
    """The single-scene legs, each on its own handle, for the step counts of
    their configurations (uniform_1m 1,000, two_blocks_16m 500, dam_break_16k
    200) unless --scene-steps sets the first three."""
    legs = {}
    for name in names:
        if name == "uniform_1m_grid":
            sc = W.uniform_lower()
            legs[name] = bench_scene(sc, args.scene_steps or sc.steps, args.warmup)
        elif name == "uniform_1m_reference_key":
            sc = W.uniform_lower()
            legs[name] = bench_scene(sc, args.scene_steps or sc.steps, args.warmup, key="reference")
        elif name == "two_blocks_16m":
            c2 = W.two_blocks_plan()
            legs[name] = bench_scene(c2["scene"], args.scene_steps or c2["scene"].steps, 2,
                                     generated=c2["init"])
        elif name == "dam_break_16k":
            c0 = W.dam_break(n=16384)
            legs[name] = bench_scene(c0, c0.steps, args.warmup, autotune=False)
    return legs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed batched steps (headline)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain one-scene-per-launch paths (profiling runs)")
    ap.add_argument("--latency-steps", type=int, default=20)
    ap.add_argument("--scene-steps", type=int, default=0,
                    help="timed steps of the uniform_1m and two_blocks_16m legs (0: their "
                         "configured 1,000 and 500)")
    ap.add_argument("--no-legs", action="store_true", help="headline only")
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma-separated single-scene legs to run (default: all)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched headline (profiling runs of the legs; value is then null)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-scenes", type=int, default=6)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--serve-sustained-rate", type=float, default=0.4)
    ap.add_argument("--serve-sustained-seconds", type=float, default=20.0)
    ap.add_argument("--no-serve-plain", action="store_true", help="skip the plain serving comparison")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: "
              "the path runs on a single MI355X here (scenes or slabs would shard across GPUs)",
              file=sys.stderr)
        return 2
    if S.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    if args.copy_calibration_mib:
        ms = S.copy_calibration(args.copy_calibration_mib << 20, 3)
        log(f"[bench] copy calibration: {args.copy_calibration_mib} MiB in {ms:.4f} ms")

    legs_wanted = [x for x in args.legs.split(",") if x]
    unknown = [x for x in legs_wanted if x not in LEGS]
    if unknown:
        print(f"unknown legs {unknown}; known: {LEGS}", file=sys.stderr)
        return 2
    if args.no_batch:
        out = {"metric": METRIC, "value": None, "unit": "particle-steps/s", "n_gpus": 1,
               "note": "batched headline skipped (--no-batch)",
               "legs": {} if args.no_legs else run_legs(legs_wanted, args)}
        if not args.no_serve:
            out["serve"] = run_serve(args)
        print(json.dumps(out), flush=True)
        return 0
    batch = bench_batch(args.steps, args.warmup, args.plain_steps, args.latency_steps,
                        plain=not args.no_plain)
    out = {
        "metric": METRIC,
        "value": batch["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batch["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch_256", "scenes": batch["scenes"], "particles": batch["particles"],
                   "particles_per_scene": "4,096-262,144 log-uniform (seed 3)",
                   "h": "0.1-0.3", "viscosity": "0-0.2", "fill": 0.4, "bucket_key": "cell grid",
                   "dt": W.DT, "substeps": W.SUBSTEPS, "parallelism": "single GPU"},
        "step_latency_ms": batch["step_latency_ms"],
        "batched_vs_single_scene": None if args.no_plain else {
            "batched": batch["value"],
            "plain_one_scene_per_launch": batch["plain_one_scene_per_launch"]["value"],
            "plain_one_scene_at_a_time": batch["plain_one_scene_at_a_time"]["value"],
            "speedup_vs_plain": batch["speedup_vs_plain"],
            "speedup_vs_one_at_a_time": batch["speedup_vs_one_at_a_time"]},
        "roofline": batch["roofline"],
        # batch_256's scenes are configured with table = particle count
        "step_roofline": step_roofline(batch["value"], W.SUBSTEPS, 1.0),
        "batch": batch,
    }
    if not args.no_legs:
        out["legs"] = run_legs(legs_wanted, args)
    if not args.no_serve:
        out["serve"] = run_serve(args)
    if not args.no_cpu:
        log("[bench] cpu baselines ...")
        # the reference's PSTL `par` loops run serially without TBB (none in
        # this image): one thread is its baseline; every core of this host's
        # share (at most 16) is shown beside it
        threads = host_threads()
        out["cpu_baseline"] = cpu_baseline_batch(args.cpu_scenes, 1)
        if threads > 1:
            out["cpu_baseline_all_cores"] = cpu_baseline_batch(args.cpu_scenes, threads)
        out["cpu_baseline_c0"] = cpu_baseline_c0(1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
