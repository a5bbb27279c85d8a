#!/usr/bin/env python3
"""Benchmark of the gfx950 SPH step (one MI355X).

Default workload: uniform_1m -- 1,048,576 particles uniform in the lower half
of a 16x8x16 box, h = 0.2, dt = 1/120 in 3 substeps, bucketed by the cell-grid
key (--key reference: the reference's hash % T with the workload's T = 2^21;
results are identical).
A step is one sph_step(1/120, 3): three Update(dt/3) calls of the reference
(physicsWorld.cc:39-111), i.e. three particle-substeps per particle.

Output: one JSON line (rank 0).  `value` is particle-steps/s with the state
resident in HBM when the timed region starts.  `roofline` is computed for the
dominant kernel from its mean duration, measured live over the timed region
with HIP events recorded on the simulation's own stream.  `cpu_baseline` times
the CPU oracle (a restatement of the reference step, the reference's `par`
loops run serially as they do without TBB) on a bounded sample;
`cpu_baseline_all_cores` the same with its per-particle loops on every core
of the host's share (at most 16).
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

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per particle per launch of each timed stage: each
# array read or written once, neighbour re-reads counted as on-chip.
#   predict:   R pos, vel (16+16), W key, ticket (4+4)
#   density:   R xps (16), W rho, rho_n, 1/rho, 1/rho_n (16); bucket table reads ~0
#   pressure:  R xps, dens, v (48), W v (16)
#   viscosity+integrate: R xps, v, pos, id (52), W pos, vel, id (36), and the
#     next update's bucket key and ticket (8)
#   (predict runs on its own only when the previous update did not record
#   the buckets: first update, changed state, settings or dt)
ALG_BYTES = {"predict": 40.0, "density": 32.0, "pressure": 64.0, "viscosity_integrate": 96.0}
STAGES = ["predict", "spatial", "density", "pressure", "viscosity_integrate"]
# Per-dispatch HBM traffic of each kernel instantiation on uniform_1m, from two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this script, each scaled by
# what the same counter reads for a device copy of known size.
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "uniform_1m_traffic.json")


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


VALU_JSON = os.path.join(ROOT, "profiles", "uniform_1m_valu.json")


def valu_issue(workload, kernel, ms):
    """VALU issue fraction of `kernel` at duration `ms`, from its committed
    instruction count (rocprofv3 SQ_INSTS_VALU): instructions x 2 cycles (a
    wave64 VALU instruction on a 32-lane SIMD) / (1024 SIMDs x 2.4 GHz x ms)."""
    try:
        with open(VALU_JSON) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if t.get("workload") != workload or k is None:
        return None
    instr = float(k["valu_instr_per_launch"])
    cyc = t["cycles_per_wave64_valu"]
    return {"kernel": kernel, "instr_per_launch": instr,
            "issue_frac": instr * cyc / (t["simds"] * t["clock_ghz"] * 1e9 * ms * 1e-3),
            "model": f"{cyc} cycles per wave64 VALU instruction, {t['simds']} SIMDs, "
                     f"{t['clock_ghz']} GHz"}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary, or None."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload:
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else float(k["traffic_bytes"])


def gather_fractions(sc, variants, grid, per):
    """HBM and VALU fractions of each gather pass (density, pressure,
    viscosity+integrate) at its live duration: algorithmic bytes / duration /
    8 TB/s, the committed measured traffic / duration / 8 TB/s, and the VALU
    issue share from the committed instruction count."""
    out = {}
    for stage in ("density", "pressure", "viscosity_integrate"):
        kern = kernel_name(stage, variants, grid)
        ms = per[stage]
        traffic = measured_traffic(sc.name, kern)
        valu = valu_issue(sc.name, kern, ms)
        out[stage] = {"kernel": kern, "ms": ms,
                      "alg_hbm_frac": ALG_BYTES[stage] * sc.n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "traffic_hbm_frac": None if traffic is None
                      else traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "valu_issue_frac": None if valu is None else valu["issue_frac"]}
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sc, substeps, threads=1):
    """Time the oracle on `substeps` Update calls of the same input, its
    per-particle loops on `threads` host threads (OpenMP; the sort stays serial)."""
    import pyoracle as O
    O.set_threads(threads)
    dt = float(np.float32(np.float32(sc.dt) / np.float32(sc.substeps)))
    op = O.from_sph(sc.params)
    t0 = time.perf_counter()
    O.update(op, dt, sc.pos, sc.vel, T=sc.table, substeps=substeps)
    el = time.perf_counter() - t0
    O.set_threads(1)
    psub = sc.n * substeps / el
    where = "1 host thread" if threads == 1 else f"{threads} host threads"
    return {"value": psub / sc.substeps, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{substeps} Update(dt/3) calls of {sc.name} (n={sc.n}, table={sc.table}) "
                      f"on {where}, {el:.1f} s"}


def copy_calibration(mib):
    """A plain device copy kernel of known size (k_copy), so that per-dispatch
    FETCH_SIZE / WRITE_SIZE counters can be scaled against known bytes."""
    ms = S.copy_calibration(mib << 20, 3)
    log(f"[bench] copy calibration: {mib} MiB copy in {ms:.4f} ms "
        f"({2 * (mib << 20) / (ms * 1e-3) / 1e9:.0f} GB/s read+write)")


def bench_batch(steps, plain_steps, warmup, key="grid"):
    """Batched (one launch per kernel for all scenes, graph-replayed) vs plain
    (one handle per scene, each kernel launch covering one scene, all handles
    enqueued asynchronously) on the same 256-scene workload."""
    bs = W.batch()
    total = bs.total
    log(f"[bench] batch: {len(bs.scenes)} scenes, {total} particles")
    grid = key == "grid"
    b = S.Batch([sc.params for sc in bs.scenes], [sc.n for sc in bs.scenes],
                None if grid else [sc.table for sc in bs.scenes])
    for s_, sc in enumerate(bs.scenes):
        b.upload(s_, sc.pos, sc.vel)
    for _ in range(warmup):
        b.step_async(bs.dt, bs.substeps)
    b.sync()
    sub_dt = float(np.float32(np.float32(bs.dt) / np.float32(bs.substeps)))
    b.autotune(sub_dt, updates=3)
    b.step(bs.dt, bs.substeps)
    b.profile_begin(steps * bs.substeps)
    b.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step_async(bs.dt, bs.substeps)
    b.sync()
    el_b = time.perf_counter() - t0
    stage_ms, _ = b.profile_end()
    lat = []
    for _ in range(20):
        t = time.perf_counter()
        b.step(bs.dt, bs.substeps)
        lat.append((time.perf_counter() - t) * 1e3)
    variants = b.variants
    b.close()
    log(f"[bench] batched {steps} steps in {el_b:.3f} s")

    sims = []
    for sc in bs.scenes:
        sim = S.Sim(sc.params, sc.n, S.TABLE_GRID if grid else sc.table)
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
    # the same launches one scene after another: each scene's step finishes
    # before the next scene's starts
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for sim in sims:
            sim.step(bs.dt, bs.substeps)
    el_s = time.perf_counter() - t0
    for sim in sims:
        sim.close()
    log(f"[bench] plain {plain_steps} steps in {el_p:.3f} s (serial {el_s:.3f} s)")
    v_b = total * steps / el_b
    v_p = total * plain_steps / el_p
    v_s = total * plain_steps / el_s
    return {"workload": bs.name, "scenes": len(bs.scenes), "particles": total,
            "bucket_key": key, "value": v_b, "unit": "particle-steps/s", "ms_per_step": el_b * 1e3 / steps,
            "steps": steps,
            "plain_one_scene_per_launch": {"value": v_p, "ms_per_step": el_p * 1e3 / plain_steps,
                                           "steps": plain_steps,
                                           "note": "every scene's own handle and stream, all enqueued "
                                                   "before any is waited for (scenes overlap)"},
            "plain_one_scene_at_a_time": {"value": v_s, "ms_per_step": el_s * 1e3 / plain_steps,
                                          "steps": plain_steps,
                                          "note": "the same launches, each scene's step finished "
                                                  "before the next scene's starts"},
            "speedup_vs_plain": v_b / v_p,
            "speedup_vs_one_at_a_time": v_b / v_s,
            "step_latency_ms": {"p50": float(np.percentile(lat, 50)),
                                "p99": float(np.percentile(lat, 99)), "n": len(lat)},
            "stage_ms_per_update": {STAGES[i]: float(stage_ms[i]) for i in range(5)},
            "gather_variants": variants}


def bench_serve(duration, rate, max_running, budget_gib, seed=4):
    """Poisson arrivals of bake requests for `duration` seconds of wall time,
    served by continuous batching (re-batched every substep).  Request inputs
    are generated before the window; the window is the whole duration, and
    requests still queued or running at its end count towards nothing but
    the particle-steps they already advanced."""
    arrivals = W.serving_requests(duration, rate=rate, seed=seed)
    inputs = [W.serving_scene(n, kind, sd) for (_, n, _, kind, sd) in arrivals]
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
            p, pos, vel = inputs[nxt]
            srv.submit(p, pos, vel, arrivals[nxt][2])
            submitted += 1
            nxt += 1
        info = srv.tick()
        if not info["particles"] and not info["queued"] and nxt < len(arrivals):
            time.sleep(max(0.0, min(0.01, arrivals[nxt][0] - (time.perf_counter() - t0))))
    st = srv.stats()     # waits for the enqueued ticks
    wall = time.perf_counter() - t0
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
            "device_capacity_particles": st["capacity"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="uniform_1m")
    ap.add_argument("--key", choices=["grid", "reference"], default="grid",
                    help="bucket key of the spatial lookup (results are identical): the "
                         "cell-grid key, or the reference's hash %% T with the workload's T")
    ap.add_argument("--latency-steps", type=int, default=100)
    ap.add_argument("--cpu-substeps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=30)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=15.0)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--serve-sustained-rate", type=float, default=0.4)
    ap.add_argument("--serve-sustained-seconds", type=float, default=30.0)
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
        copy_calibration(args.copy_calibration_mib)
    sc = W.by_name(args.workload)
    log(f"[bench] workload {sc.name} n={sc.n} table={sc.table} steps={args.steps} warmup={args.warmup}")
    sim = S.Sim(sc.params, sc.n, S.TABLE_GRID if args.key == "grid" else sc.table)
    key_grid, key_gdim, _ = sim.key()
    sim.upload(sc.pos, sc.vel)
    for _ in range(args.warmup):
        sim.step_async(sc.dt, sc.substeps)
    sim.sync()
    log("[bench] warmup done")
    if not args.no_autotune:
        sub_dt = float(np.float32(np.float32(sc.dt) / np.float32(sc.substeps)))
        tuned = sim.autotune(sub_dt, updates=3)
        log(f"[bench] autotune: {sim.variants} ms {np.round(tuned, 4)}")
        for _ in range(2):
            sim.step_async(sc.dt, sc.substeps)
        sim.sync()

    # timed region: K steps, stage events recorded on the simulation stream
    sim.profile_begin(args.steps * sc.substeps)
    sim.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step_async(sc.dt, sc.substeps)
    sim.sync()
    elapsed = time.perf_counter() - t0
    stage_ms, n_upd = sim.profile_end()
    log(f"[bench] timed {args.steps} steps in {elapsed:.3f} s; stage ms {np.round(stage_ms[:5], 4)}")

    # per-step latency (blocking steps, graph replay), outside the timed region
    lat = []
    for _ in range(args.latency_steps):
        t = time.perf_counter()
        sim.step(sc.dt, sc.substeps)
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat)
    nonfinite = sim.count_nonfinite()
    sim_table = sim.table
    variants = sim.variants
    sim.close()

    value = sc.n * args.steps / elapsed
    per = {STAGES[i]: float(stage_ms[i]) for i in range(5)}
    gather = ["density", "pressure", "viscosity_integrate"]
    dom = max(gather, key=lambda k: per[k])
    ach = ALG_BYTES[dom] * sc.n / (per[dom] * 1e-3) / 1e9
    dom_kernel = kernel_name(dom, variants, key_grid)
    out = {
        "metric": "particle-steps/s (single scene, 1 MI355X)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sc.name, "particles": sc.n, "reference_table_size": sc.table,
                   "bucket_key": ("cell grid %dx%dx%d" % tuple(key_gdim)) if key_grid
                   else "reference hash %% %d" % sc.table, "buckets": sim_table,
                   "h": sc.params.smoothing_radius, "dt": sc.dt, "substeps": sc.substeps,
                   "parallelism": "single GPU"},
        "step_latency_ms": {"p50": float(np.percentile(lat, 50)),
                            "p99": float(np.percentile(lat, 99)), "n": int(lat.size)},
        "stage_ms_per_update": per,
        "gather_variants": variants,
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic(sc.name, dom_kernel),
                     "alg_bytes_per_launch": ALG_BYTES[dom] * sc.n,
                     "alg_bytes_per_particle": ALG_BYTES[dom]},
        "valu": valu_issue(sc.name, dom_kernel, per[dom]),
        "gather_kernels": gather_fractions(sc, variants, key_grid, per),
        "nonfinite_components": nonfinite,
    }
    if not args.no_batch:
        out["batch"] = bench_batch(args.batch_steps, args.plain_steps, 3, args.key)
    if not args.no_serve:
        out["serve"] = {"workload": "serving (Poisson bake requests, 8,192-2,097,152 particles, "
                                    "100-2,000 steps)",
                        "configured": bench_serve(args.serve_seconds, args.serve_rate,
                                                  args.serve_slots, 200),
                        "sustained": bench_serve(args.serve_sustained_seconds,
                                                 args.serve_sustained_rate, args.serve_slots, 200,
                                                 seed=5)}
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        # the reference's PSTL `par` loops run serially without TBB (none in
        # this image): one thread is its baseline; every core of this host's
        # share (at most 16) is shown beside it
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_substeps)
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        if threads > 1:
            out["cpu_baseline_all_cores"] = cpu_baseline(sc, args.cpu_substeps, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
