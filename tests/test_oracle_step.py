"""Pins the oracle's stages (physicsWorld.cc:39-111, 304-498) by brute force,
closed forms and invariants on small inputs.

  - spatial lookup (cc:466-498): against numpy's stable argsort and a
    first-occurrence table; INT_MAX for empty buckets;
  - neighbour sets (cc:331-363): against an O(N^2) search over the 27
    adjacent cells (integer cell coordinates) with the squared-distance
    prefilter, including tables small enough to force bucket collisions;
  - density (cc:325-365): against an FP64 brute-force sum, relative 1e-5;
  - pressure (cc:367-422): two-particle closed form, antisymmetry,
    repulsion, d = 0 direction (0, 1, 0);
  - viscosity (cc:424-464): Jacobi conserves momentum; two-particle closed
    forms distinguish the Jacobi and Gauss-Seidel orders;
  - integration and wall collision (cc:80-108): closed forms;
  - GridArrangement (cc:518-557): lattice, order and centring.
"""
import math

import numpy as np
import pytest

import pyoracle as O


def _random_scene(n, box, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.5, 0.5, (n, 3)) * np.asarray(box, np.float32)
    return pos


def _w2_64(d, h):
    return np.where(d < h, 15.0 / (2 * math.pi * h**5) * (h - d) ** 2, 0.0)


def _w3_64(d, h):
    return np.where(d < h, 15.0 / (math.pi * h**6) * (h - d) ** 3, 0.0)


@pytest.mark.parametrize("T", [97, 1000, 4096])
def test_spatial_lookup_is_stable_counting_sort(T):
    p = O.params(h=0.2)
    xp = _random_scene(3000, (3, 3, 3), 1)
    s = O.spatial(xp, T, p.h)
    cells = np.floor(xp[:, :3] / np.float32(p.h)).astype(np.int64)
    hashes = (15823 * cells[:, 0] + 9737333 * cells[:, 1] + 440817757 * cells[:, 2]) % (1 << 32)
    np.testing.assert_array_equal(s.hash_id, hashes.astype(np.uint32))
    np.testing.assert_array_equal(s.key_id, (hashes % T).astype(np.uint32))
    order = np.argsort(s.key_id, kind="stable")
    np.testing.assert_array_equal(s.sorted_idx, order)
    np.testing.assert_array_equal(s.sorted_key, s.key_id[order])
    want_start = np.full(T, O.START_EMPTY, np.uint32)
    first = np.unique(s.sorted_key, return_index=True)
    want_start[first[0]] = first[1]
    np.testing.assert_array_equal(s.start, want_start)


def _brute_neighbours(xp, h, r2):
    x = xp[:, :3].astype(np.float32)
    cells = np.floor(x / np.float32(h)).astype(np.int64)
    out = []
    for i in range(x.shape[0]):
        o = x - x[i]
        d2 = (o[:, 0] * o[:, 0] + o[:, 1] * o[:, 1]) + o[:, 2] * o[:, 2]   # float32, dot order
        adj = np.all(np.abs(cells - cells[i]) <= 1, axis=1)
        out.append(set(np.nonzero(adj & (d2 <= np.float32(r2)))[0].tolist()))
    return out


@pytest.mark.parametrize("h,T", [(0.2, 500), (0.2, 37), (0.35, 800), (0.1, 1000)])
def test_neighbour_sets_brute_force(h, T):
    p = O.params(h=h)
    xp = _random_scene(500, (1.5, 1.5, 1.5), 2)
    s = O.spatial(xp, T, p.h)
    offs, lst = O.neighbours(p, xp, s)
    brute = _brute_neighbours(xp, h, p.r2)
    for i in range(xp.shape[0]):
        got = lst[offs[i]:offs[i + 1]]
        assert len(set(got.tolist())) == len(got), "a neighbour listed twice"
        assert set(got.tolist()) == brute[i]


def test_density_against_fp64_brute_force():
    p = O.params(h=0.2)
    xp = _random_scene(1500, (1.2, 1.2, 1.2), 3)
    s = O.spatial(xp, 1500, p.h)
    dens = O.density(p, xp, s).astype(np.float64)
    x = xp[:, :3].astype(np.float64)
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    h = float(np.float32(0.2))
    rho = _w2_64(d, h).sum(1)
    rhon = _w3_64(d, h).sum(1)
    assert np.max(np.abs(dens[:, 0] - rho) / rho) < 1e-5
    assert np.max(np.abs(dens[:, 1] - rhon) / rhon) < 1e-5
    d64 = O.density64(p, xp, s)
    assert np.max(np.abs(d64[:, 0] - rho) / rho) < 1e-12


def test_single_particle_density_and_no_force():
    p = O.params(h=0.2, rho0=1000.0)
    pos = np.array([[0.3, -0.2, 0.1, 0]], np.float32)
    vel = np.array([[0.5, -1.0, 2.0, 0]], np.float32)
    _, xp = O.predict(p, 0.01, pos, vel)
    s = O.spatial(xp, 1, p.h)
    dens = O.density(p, xp, s)
    assert dens[0, 0] == np.float32(O.W2(0.0, 0.2))
    assert dens[0, 1] == np.float32(O.W3(0.0, 0.2))
    v2 = O.pressure(p, 0.01, xp, s, dens, vel)
    v3 = O.viscosity(p, 0.01, xp, s, v2)
    np.testing.assert_array_equal(v3, vel)


def _pair(d, axis=0, vel=None):
    pos = np.zeros((2, 4), np.float32)
    pos[1, axis] = d
    return pos, (np.zeros((2, 4), np.float32) if vel is None else vel)


@pytest.mark.parametrize("d", [0.05, 0.11, 0.19])
def test_two_particle_pressure_closed_form(d):
    h, rho0, k, kn, dt = 0.2, 500.0, 300.0, 20.0, 0.01
    p = O.params(h=h, rho0=rho0, k=k, k_near=kn)
    xp, _ = _pair(d)
    s = O.spatial(xp, 2, h)
    dens = O.density(p, xp, s)
    v = O.pressure(p, dt, xp, s, dens, np.zeros((2, 4), np.float32))
    hh = float(np.float32(h))
    dd = float(np.float32(d))
    rho = float(_w2_64(0.0, hh) + _w2_64(dd, hh))
    rhon = float(_w3_64(0.0, hh) + _w3_64(dd, hh))
    P, Pn = (rho - rho0) * k, rhon * kn
    dw2 = -15.0 / (math.pi * hh**5) * (hh - dd)
    dw3 = -45.0 / (math.pi * hh**6) * (hh - dd) ** 2
    f0 = 1.0 * (dw2 * P / rho + dw3 * Pn / rhon)   # dir from 0 to 1 is +x
    assert v[0, 0] == pytest.approx(f0 / rho * dt, rel=1e-5)
    assert v[1, 0] == pytest.approx(-v[0, 0], rel=1e-6)        # antisymmetric
    assert v[0, 1] == 0 and v[0, 2] == 0
    if P > 0:
        assert v[0, 0] < 0                                      # pushed away from particle 1
    # the FP64 restatement (used for tolerances) meets the same closed form tightly
    d64 = O.density64(p, xp, s)
    assert d64[0, 0] == pytest.approx(rho, rel=1e-12) and d64[0, 1] == pytest.approx(rhon, rel=1e-12)
    dv64, mag64 = O.pressure64(p, dt, xp, s, d64)
    assert dv64[0, 0] == pytest.approx(f0 / rho * dt, rel=1e-9)
    assert dv64[1, 0] == pytest.approx(-dv64[0, 0], rel=1e-12)
    assert dv64[0, 1] == 0 and dv64[0, 2] == 0
    # the scale of the sum: |pressure term| + |near-pressure term|
    assert mag64[0, 0] == pytest.approx((abs(dw2 * P / rho) + abs(dw3 * Pn / rhon)) / rho * dt,
                                        rel=1e-9)


def test_coincident_particles_push_along_plus_y():
    p = O.params(h=0.2, rho0=0.0)
    xp = np.zeros((2, 4), np.float32)
    s = O.spatial(xp, 2, 0.2)
    dens = O.density(p, xp, s)
    v = O.pressure(p, 0.01, xp, s, dens, np.zeros((2, 4), np.float32))
    # dir = (0, 1, 0) when d == 0 (cc:414); dW < 0 and pressures > 0
    assert v[0, 0] == 0 and v[0, 2] == 0 and v[0, 1] < 0
    assert v[1, 1] == v[0, 1]


def test_viscosity_jacobi_conserves_momentum():
    p = O.params(h=0.2, mu=0.3)
    rng = np.random.default_rng(4)
    xp = _random_scene(800, (1.0, 1.0, 1.0), 5)
    vel = np.zeros((800, 4), np.float32)
    vel[:, :3] = rng.normal(0, 1, (800, 3))
    s = O.spatial(xp, 800, 0.2)
    v = O.viscosity(p, 0.01, xp, s, vel, gauss_seidel=False)
    dp = (v[:, :3].astype(np.float64) - vel[:, :3]).sum(0)
    scale = np.abs(v[:, :3].astype(np.float64) - vel[:, :3]).sum(0)
    assert np.all(np.abs(dp) <= 1e-5 * scale)
    vin = vel[:, :3].astype(np.float64)
    dv64, mag = O.viscosity64(p, 0.01, xp, s, vin)
    assert np.abs(dv64.sum(0)).max() < 1e-9 * mag.sum()


def test_viscosity_two_particle_jacobi_and_gauss_seidel():
    h, mu, dt, d = 0.2, 0.5, 0.01, 0.1
    p = O.params(h=h, mu=mu)
    vel = np.zeros((2, 4), np.float32)
    vel[0, 0], vel[1, 0] = 1.0, -1.0
    xp, _ = _pair(d, vel=vel)
    s = O.spatial(xp, 2, h)
    w = 315.0 / (64 * math.pi * h**9) * (h * h - d * d) ** 3
    c = mu * dt * w
    vj = O.viscosity(p, dt, xp, s, vel, gauss_seidel=False)
    assert vj[0, 0] == pytest.approx(1.0 + c * (-2.0), rel=1e-5)
    assert vj[1, 0] == pytest.approx(-1.0 + c * 2.0, rel=1e-5)
    vg = O.viscosity(p, dt, xp, s, vel, gauss_seidel=True)
    v0 = 1.0 - 2.0 * c                  # particle 0 first (index order)
    assert vg[0, 0] == pytest.approx(v0, rel=1e-5)
    assert vg[1, 0] == pytest.approx(-1.0 + c * (v0 + 1.0), rel=1e-5)


def test_integrate_and_wall_collision():
    p = O.params(bounds=(2.0, 4.0, 6.0), damp=0.95)
    pos = np.array([[0.0, 0.0, 0.0, 0], [0.99, -1.0, 2.0, 0], [0.5, -1.99, -2.99, 0]], np.float32)
    vel = np.array([[1.0, 2.0, 3.0, 0], [2.0, 0.0, 0.0, 0], [0.0, -2.0, -2.0, 0]], np.float32)
    dt = np.float32(0.01)
    x, v = O.integrate(p, dt, pos, vel)
    np.testing.assert_allclose(x[0, :3], [0.01, 0.02, 0.03], rtol=1e-6)
    assert x[1, 0] == 1.0 and v[1, 0] == np.float32(2.0) * np.float32(-0.95)   # hit +x wall
    assert x[2, 1] == -2.0 and v[2, 1] == np.float32(-2.0) * np.float32(-0.95)  # hit -y wall
    assert x[2, 2] == -3.0 and v[2, 2] == np.float32(-2.0) * np.float32(-0.95)  # hit -z wall
    assert np.all(x[:, 3] == np.float32(0.34))


def test_predict_gravity_and_lookahead():
    p = O.params(gravity_on=1, gravity_scale=9.81)
    pos = _random_scene(10, (1, 1, 1), 6)
    vel = np.zeros_like(pos)
    dt = 0.01
    v, xp = O.predict(p, dt, pos, vel)
    assert np.all(v[:, 1] == np.float32(-9.81) * np.float32(dt))
    assert np.all(v[:, 0] == 0) and np.all(v[:, 2] == 0)
    np.testing.assert_allclose(xp[:, 1] - pos[:, 1], v[:, 1] / 120.0, rtol=1e-5, atol=1e-9)
    p.gravity_on = 0
    v, xp = O.predict(p, dt, pos, vel)
    assert np.all(v == 0) and np.all(xp[:, :3] == pos[:, :3])


def test_grid_arrangement_lattice():
    pos = O.grid_arrangement(27)   # 3 per axis, gap 0.215, centred (cc:518-557)
    gap = np.float32(0.215)
    x = np.unique(np.round(pos[:, 0], 5))
    np.testing.assert_allclose(x, [-gap, 0, gap], atol=1e-6)
    # loop order y (top down), x, z: z varies fastest, then x, y slowest
    assert pos[0, 2] < pos[1, 2] < pos[2, 2]
    assert pos[0, 0] < pos[3, 0]
    assert pos[0, 1] > pos[9, 1]
    assert np.all(pos[:, 3] == np.float32(0.34))
    big = O.grid_arrangement(10000)   # ceil(cbrt) = 22 per axis; all placed
    assert np.all(big[:, 3] == np.float32(0.34))


def test_update_composes_its_stages():
    p = O.params(h=0.2, rho0=1000.0, mu=0.05, gravity_on=1, gravity_scale=9.81, bounds=(2, 2, 2))
    pos = _random_scene(2000, (1.2, 1.2, 1.2), 7)
    vel = np.zeros_like(pos)
    dt = float(np.float32(1 / 360))
    v1, xp = O.predict(p, dt, pos, vel)
    s = O.spatial(xp, 2000, p.h)
    dens = O.density(p, xp, s)
    v2 = O.pressure(p, dt, xp, s, dens, v1)
    v3 = O.viscosity(p, dt, xp, s, v2)
    x4, v4 = O.integrate(p, dt, pos, v3)
    x, v, d = O.update(p, dt, pos, vel, T=2000)
    np.testing.assert_array_equal(x, x4)
    np.testing.assert_array_equal(v, v4)
    np.testing.assert_array_equal(d, dens)


def test_reference_sort_mode_keeps_keys_and_bucket_sets():
    """The CPU baseline's stage B2 (std::sort of 12-byte float records, as
    cc:480-484 run it) gives the same key sequence, start table and per-bucket
    index sets as the stable order, and an Update close to the stable one
    (only the within-bucket summation order differs)."""
    p = O.params(h=0.2, rho0=1000.0, mu=0.05, gravity_on=1, gravity_scale=9.81, bounds=(2, 2, 2))
    pos = _random_scene(3000, (1.2, 1.2, 1.2), 8)
    vel = np.zeros_like(pos)
    _, xp = O.predict(p, 0.01, pos, vel)
    stable = O.spatial(xp, 500, p.h)
    O.set_reference_sort(True)
    try:
        ref = O.spatial(xp, 500, p.h)
        x_r, v_r, d_r = O.update(p, float(np.float32(1 / 360)), pos, vel, T=3000)
    finally:
        O.set_reference_sort(False)
    np.testing.assert_array_equal(ref.sorted_key, stable.sorted_key)
    np.testing.assert_array_equal(ref.start, stable.start)
    for k in np.unique(stable.sorted_key):
        sel = stable.sorted_key == k
        assert set(ref.sorted_idx[sel].tolist()) == set(stable.sorted_idx[sel].tolist())
    assert not np.array_equal(ref.sorted_idx, stable.sorted_idx)  # an unstable order indeed
    x_s, v_s, d_s = O.update(p, float(np.float32(1 / 360)), pos, vel, T=3000)
    np.testing.assert_allclose(d_r, d_s, rtol=1e-5)
    np.testing.assert_allclose(x_r, x_s, rtol=1e-5, atol=1e-6)
