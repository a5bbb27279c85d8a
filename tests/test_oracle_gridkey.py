"""The step's results do not depend on the bucket key (CPU, oracle only).

The reference buckets particles by hash % T (physicsWorld.cc:513-516) and,
for each of the 27 neighbour cells, walks that cell's bucket keeping the
records whose float-rounded hash matches (cc:331-363).  The GPU library can
bucket by a cell-grid key instead, (wx * Gy + wy) * Gz + wz with
w = (u32(cell) - u32(lo)) mod G per axis.  These tests run the oracle with
both keys on the same inputs and require bit-identical positions,
velocities, densities and neighbour lists: grids large enough for no cell to
wrap, grids so small that far cells share buckets, and degenerate periods
(G = 1 or 2 on an axis, where the three cells of a neighbour row wrap onto
each other).  They also pin the oracle's grid key against a numpy
restatement, and check a Python restatement of the GPU's row walk (three
z-adjacent buckets walked as one run, a record passing when its float hash
equals any of the row's three cell hashes) against the reference walk's
neighbour lists, with and without wrapping.
"""
import numpy as np
import pytest

import pyoracle as O
from sph_amd import workloads as W


def _ideal_grid(p, margin=4):
    """Grid with no wrap inside the box (+margin cells): (gdim, lo)."""
    h = np.float32(p.h)
    lo, dim = [], []
    for c in range(3):
        half = np.float32(p.bounds[c]) * np.float32(0.5)
        a = int(np.floor(-half / h)) - margin
        b = int(np.floor(half / h)) + margin
        lo.append(a)
        dim.append(b - a + 1)
    return dim, lo


def _np_grid_key(xp, h, gdim, lo):
    cells = np.floor(xp[:, :3] / np.float32(h)).astype(np.int64)
    w = [((cells[:, c] - lo[c]) % (1 << 32)) % gdim[c] for c in range(3)]
    return ((w[0] * gdim[1] + w[1]) * gdim[2] + w[2]).astype(np.uint32)


def _dam(n=4096):
    sc = W.dam_break(n)
    return O.from_sph(sc.params), sc.pos, sc.vel, float(np.float32(np.float32(1 / 120) / np.float32(3)))


def _uniform(n=3000, seed=3):
    rng = np.random.default_rng(seed)
    p = O.params(h=0.2)
    p.gravity_on = 1
    p.gravity_scale = 9.81
    p.rho0 = 1000.0
    for c in range(3):
        p.bounds[c] = 3.0
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    return p, pos, vel, float(np.float32(1 / 360))


@pytest.mark.parametrize("gdim_lo", [((5, 3, 7), (-2, 4, -9)), ((13, 11, 17), (0, 0, 0))])
def test_grid_key_matches_numpy(gdim_lo):
    gdim, lo = gdim_lo
    p, pos, vel, dt = _uniform()
    xp = pos.copy()
    T = int(np.prod(gdim)) + 5
    with O.grid_key(gdim, lo):
        s = O.spatial(xp, T, p.h)
    np.testing.assert_array_equal(s.key_id, _np_grid_key(xp, p.h, gdim, lo))
    order = np.argsort(s.key_id, kind="stable")
    np.testing.assert_array_equal(s.sorted_idx, order)


def _grids(p, n):
    ideal, lo = _ideal_grid(p)
    return [
        ("ideal", ideal, lo),
        ("wrap", [7, 5, 6], lo),              # far cells share buckets
        ("shifted", ideal, [l + 3 for l in lo]),  # seam inside the box
        ("gz1", [ideal[0], ideal[1], 1], lo),  # z rows collapse onto one bucket
        ("gz2", [9, 4, 2], lo),
        ("g111", [1, 1, 1], lo),               # one bucket for everything
    ]


@pytest.mark.parametrize("scene", ["dam", "uniform"])
@pytest.mark.parametrize("gauss_seidel", [False, True])
def test_update_does_not_depend_on_key(scene, gauss_seidel):
    p, pos, vel, dt = _dam() if scene == "dam" else _uniform()
    n = pos.shape[0]
    want = O.update(p, dt, pos, vel, T=n, gauss_seidel=gauss_seidel, substeps=3)
    for name, gdim, lo in _grids(p, n):
        T = max(int(np.prod(gdim)), 1)
        with O.grid_key(gdim, lo):
            got = O.update(p, dt, pos, vel, T=T, gauss_seidel=gauss_seidel, substeps=3)
        for a, b, what in zip(got, want, ("pos", "vel", "dens")):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{scene}/{name}: {what} differs"


def test_neighbour_lists_do_not_depend_on_key():
    p, pos, vel, dt = _uniform(2000, seed=5)
    xp = pos.copy()
    n = xp.shape[0]
    s = O.spatial(xp, n, p.h)
    want = O.neighbours(p, xp, s)
    for name, gdim, lo in _grids(p, n):
        T = max(int(np.prod(gdim)), 1)
        with O.grid_key(gdim, lo):
            sg = O.spatial(xp, T, p.h)
            got = O.neighbours(p, xp, sg)
        np.testing.assert_array_equal(got[0], want[0], err_msg=name)
        np.testing.assert_array_equal(got[1], want[1], err_msg=name)


def test_neighbour_cells_never_float_alias():
    """The fact the key independence rests on: no two cells of a 3x3x3 block
    have hashes that round to the same float (differences are >= 15823 away
    from 0 mod 2^32; u32 values rounding to one float are < 256 apart)."""
    m = 1 << 32
    worst = m
    for dx in range(-2, 3):
        for dy in range(-2, 3):
            for dz in range(-2, 3):
                if (dx, dy, dz) == (0, 0, 0):
                    continue
                d = (dx * 15823 + dy * 9737333 + dz * 440817757) % m
                worst = min(worst, d, m - d)
    assert worst >= 256
    # and spot-check the float rounding claim on hashes near the top of u32
    rng = np.random.default_rng(0)
    base = rng.integers(0, m, 10000, dtype=np.uint64)
    other = (base + np.uint64(worst)) % np.uint64(m)
    assert not np.any(base.astype(np.float32) == other.astype(np.float32))


def _row_walk_lists(xp, p, gdim, lo, T):
    """Python restatement of the GPU's grid-key walk (sph_kernels.hip,
    walk_grid / walk_run): 9 rows (dx outer, dy inner); where the three
    z-cells of a row are consecutive buckets their slots form one run and a
    record passes if d2 < thr and its float hash equals ANY of the three row
    cells' float hashes -- no hash test at all when no particle's cell lies
    outside the grid's box and this particle's 27 cells are inside it; at the
    wrap seam the buckets are walked one by one.
    Returns per-particle lists of passing particle indices, in walk order,
    self included, with the prefilter d2 <= r2 only (the neighbour list)."""
    h = np.float32(p.h)
    with O.grid_key(gdim, lo):
        s = O.spatial(xp, T, p.h)
    n = xp.shape[0]
    start = s.start
    end = np.zeros(T, np.int64)
    counts = np.bincount(s.key_id, minlength=T)
    nz = counts > 0
    end[nz] = start[nz].astype(np.int64) + counts[nz]
    srt = s.sorted_idx
    rec = xp[srt, :3].astype(np.float32)
    rech = s.hash_id[srt].astype(np.float32)
    r2 = np.float32(p.r2)
    cells = np.floor(xp[:, :3] / h).astype(np.int64)
    M = 1 << 32
    G = [int(g) for g in gdim]

    def wrap(u, g):
        return (u % M) % g

    # no particle's cell outside the grid's box: the key is injective on them
    uall = [((cells[:, a] - int(lo[a])) % M) for a in range(3)]
    clean = all(bool(np.all(uall[a] < G[a])) for a in range(3))
    out = []
    for i in range(n):
        c = cells[i]
        u = [(int(c[a]) - int(lo[a])) % M for a in range(3)]
        # the GPU skips the hash test for such a particle whose 27 cells are inside
        inner = clean and all(1 <= u[a] <= G[a] - 2 for a in range(3))
        wx = [wrap(u[0] + d - 1, G[0]) for d in range(3)]
        wy = [wrap(u[1] + d - 1, G[1]) for d in range(3)]
        wz = [wrap(u[2] + d - 1, G[2]) for d in range(3)]
        run = wz[1] == wz[0] + 1 and wz[2] == wz[1] + 1
        h0 = (15823 * c[0] + 9737333 * c[1] + 440817757 * c[2]) % M
        lst = []
        pi = xp[i, :3].astype(np.float32)
        for r in range(9):
            dx, dy = r // 3 - 1, r % 3 - 1
            base = (wx[r // 3] * G[1] + wy[r % 3]) * G[2]
            keys = [base + wz[q] for q in range(3)]
            hr = (h0 + dx * 15823 + dy * 9737333) % M
            f = [np.float32((hr + (q - 1) * 440817757) % M) for q in range(3)]
            if run:
                segs = []
                ne = [k for k in keys if start[k] != O.START_EMPTY]
                if ne:
                    segs.append((int(start[ne[0]]), int(end[ne[-1]]), f))
            else:
                segs = [(int(start[k]), int(end[k]), [f[q]] * 3)
                        for q, k in enumerate(keys) if start[k] != O.START_EMPTY]
            for b, e, fs in segs:
                for m in range(b, e):
                    o = rec[m] - pi
                    d2 = np.float32(np.float32(o[0] * o[0]) + np.float32(o[1] * o[1])) + np.float32(o[2] * o[2])
                    if d2 <= r2 and (inner or rech[m] in fs):
                        lst.append(int(srt[m]))
        out.append(lst)
    return out


@pytest.mark.parametrize("grid", ["ideal", "wrap", "shifted", "gz2"])
def test_row_walk_matches_reference_walk(grid):
    p, pos, vel, dt = _uniform(1500, seed=9)
    xp = pos.copy()
    n = xp.shape[0]
    s = O.spatial(xp, n, p.h)
    offs, lst = O.neighbours(p, xp, s)
    want = [list(lst[offs[i]:offs[i + 1]]) for i in range(n)]
    name, gdim, lo = next(g for g in _grids(p, n) if g[0] == grid)
    got = _row_walk_lists(xp, p, gdim, lo, max(int(np.prod(gdim)), 1))
    assert got == want
