"""The benchmark scenes, as (params, n, table_size, pos4, vel4, dt, substeps, steps).

Scene definitions (all FP32, positions from sph_generate's counter-based
generator, so every consumer sees identical bytes):

  dam_break_16k   16,384 particles, dam-break block 16x32x32 on a lattice of
                  spacing 0.1 (= 0.5 h) at the -x,-y,-z corner of a 4x4x4 box,
                  jitter +-0.002 (= 0.01 h), v = 0, h = 0.2, gravity 9.81,
                  dt = 1/120 in 3 substeps, viscosity 0.05, 200 steps.
  uniform_1m      1,048,576 particles uniform in the lower half of a 16x8x16
                  box, table 2^21, same settings, 1,000 steps.
  batch_256       256 independent scenes, particle counts log-uniform in
                  [4,096, 262,144] from the seed, h uniform in [0.1, 0.3],
                  viscosity uniform in [0, 0.2], each a dam-break block on a
                  lattice of spacing 0.5 h in a cubic box sized for 40% fill
                  (particle volume (0.5 h)^3 each), target density 8 / h^3
                  (the lattice's number density), table = count, 300 steps.
  serving         bake requests with Poisson arrivals (20/s), particle counts
                  log-uniform in [8,192, 2,097,152], 100-2,000 steps each,
                  a centred or dam-break block (even odds) at 40% fill,
                  h = 0.2; admission by slots and a 200 GiB HBM budget.
  two_blocks_16m  16,777,216 particles in two lattice blocks moving at +-2 m/s
                  along x in a 40x20x40 box, h = 0.2, 500 steps.

Settings shared by all three: pressure multiplier 300 and near-pressure
multiplier 20 (the reference's defaults, physicsWorld.h:99-100), target
density 1000 (the number density of a lattice of spacing 0.1, against which
the normalised kernels sum to about 1000), the reference's fixed lookahead
1/120, damping 0.95 and prefilter 0.35^2 (physicsWorld.cc:47, 87; h:96).
"""
from dataclasses import dataclass, field

import numpy as np

from . import (INIT_CENTRED_BLOCK, INIT_DAM_BREAK, INIT_TWO_BLOCKS, INIT_UNIFORM_LOWER,
               default_params,
               generate)

DT = 1.0 / 120.0
SUBSTEPS = 3


def scene_params(h=0.2, rho0=1000.0, mu=0.05, box=(4.0, 4.0, 4.0), gravity=True):
    return default_params(smoothing_radius=h, target_density=rho0, pressure_mult=300.0,
                          near_pressure_mult=20.0, viscosity=mu, gravity_on=1 if gravity else 0,
                          gravity_scale=9.81, bounds=box)


@dataclass
class Scene:
    name: str
    params: object
    n: int
    table: int
    pos: np.ndarray
    vel: np.ndarray
    steps: int
    dt: float = DT
    substeps: int = SUBSTEPS
    meta: dict = field(default_factory=dict)


def dam_break(n=16384, seed=0, h=0.2, steps=200, box=(4.0, 4.0, 4.0)):
    pos, vel = generate(INIT_DAM_BREAK, seed, n, box, 0.5 * h, 0.01 * h)
    return Scene("dam_break_%d" % n, scene_params(h=h, box=box), n, n, pos, vel, steps)


def uniform_lower(n=1 << 20, table=1 << 21, seed=1, h=0.2, steps=1000, box=(16.0, 8.0, 16.0)):
    pos, vel = generate(INIT_UNIFORM_LOWER, seed, n, box)
    return Scene("uniform_%d" % n, scene_params(h=h, box=box), n, table, pos, vel, steps)


def two_blocks(n=1 << 24, seed=2, h=0.2, steps=500, box=(40.0, 20.0, 40.0), speed=2.0):
    pos, vel = generate(INIT_TWO_BLOCKS, seed, n, box, 0.5 * h, 0.01 * h, speed)
    return Scene("two_blocks_%d" % n, scene_params(h=h, box=box), n, n, pos, vel, steps)


def two_blocks_plan(n=1 << 24, seed=2, h=0.2, steps=500, box=(40.0, 20.0, 40.0), speed=2.0):
    """two_blocks without host arrays: the scene (pos, vel None) and the
    generator arguments (kind, seed, box, spacing, jitter, speed) with which
    sph_initialize writes the same bytes on the device."""
    sc = Scene("two_blocks_%d" % n, scene_params(h=h, box=box), n, n, None, None, steps)
    return {"scene": sc, "init": (INIT_TWO_BLOCKS, seed, box, 0.5 * h, 0.01 * h, speed)}


@dataclass
class BatchScenes:
    name: str
    scenes: list
    steps: int
    dt: float = DT
    substeps: int = SUBSTEPS

    @property
    def total(self):
        return sum(sc.n for sc in self.scenes)


def batch_scene(n, h, mu, seed):
    L = (n * (0.5 * h) ** 3 / 0.4) ** (1.0 / 3.0)
    box = (L, L, L)
    pos, vel = generate(INIT_DAM_BREAK, seed, n, box, 0.5 * h, 0.01 * h)
    p = scene_params(h=h, rho0=8.0 / h**3, mu=mu, box=box)
    return Scene("scene_%d" % seed, p, n, n, pos, vel, 0, meta={"h": h, "mu": mu, "box": L})


def batch(n_scenes=256, seed=3, lo=4096, hi=262144, steps=300):
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), n_scenes)).astype(np.int64)
    hs = rng.uniform(0.1, 0.3, n_scenes)
    mus = rng.uniform(0.0, 0.2, n_scenes)
    scenes = [batch_scene(int(counts[s]), float(hs[s]), float(mus[s]), seed * 100000 + s)
              for s in range(n_scenes)]
    return BatchScenes("batch_%d" % n_scenes, scenes, steps)


def serving_requests(duration_s, rate=20.0, seed=4, lo=8192, hi=2097152, steps_lo=100,
                     steps_hi=2000, h=0.2):
    """Poisson arrivals (rate per second) over duration_s: (arrival_s, n, steps,
    kind, seed) tuples; scenes are built lazily by serving_scene."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration_s:
            return out
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        kind = INIT_DAM_BREAK if rng.random() < 0.5 else INIT_CENTRED_BLOCK
        out.append((t, n, steps, kind, seed * 1000003 + len(out)))


def serving_plan(n, kind, seed, h=0.2):
    """serving_scene without host arrays: (params, (box, spacing, jitter,
    speed)), the generator arguments of the request's initial state."""
    L = (n * (0.5 * h) ** 3 / 0.4) ** (1.0 / 3.0)
    box = (L, L, L)
    return scene_params(h=h, box=box), (box, 0.5 * h, 0.01 * h, 0.0)


def serving_scene(n, kind, seed, h=0.2):
    """A bake request's scene: a block (centred, or dam-break at the corner) on
    a lattice of spacing 0.5 h in a cubic box at 40% fill, h = 0.2, density
    target 1000, the shared settings of the single-scene workloads."""
    L = (n * (0.5 * h) ** 3 / 0.4) ** (1.0 / 3.0)
    box = (L, L, L)
    pos, vel = generate(kind, seed, n, box, 0.5 * h, 0.01 * h)
    return scene_params(h=h, box=box), pos, vel


def by_name(name):
    return {"dam_break_16k": dam_break, "uniform_1m": uniform_lower,
            "two_blocks_16m": two_blocks, "batch_256": batch}[name]()
