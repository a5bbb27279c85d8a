"""Pins the oracle's pressure (physicsWorld.cc:367-422) and viscosity
(cc:424-464) stages against sums written here independently of oracle.c.

  - three particles on a line at unequal spacings: every particle sees
    neighbours of different density and pressure, so the closed form
    separates the reference's terms -- dividing by the NEIGHBOUR's density
    (cc:416-417), the SHARED pressure (P_i + P_j) / 2 (cc:410-411), the
    direction from the particle to the neighbour (cc:414), the division of
    the sum by the particle's own density (cc:421).  Each plausible misreading
    is evaluated beside the closed form and must lie far outside the
    tolerance that the oracle meets;
  - the same for Jacobi viscosity with three different velocities (cc:460,
    463): the sign of (v_j - v_i) and the poly6 weight of each pair;
  - a random scene with unequal densities: an O(N^2) FP64 sum over all pairs
    (no buckets, no walk) against the FP32 oracle stages and their FP64
    twins orc64_pressure / orc64_viscosity.

Tolerances are relative to the scale of each sum, mag = sum over pair terms
of |term| (divided by rho_i and times dt, as the velocity change is; for
pressure with |P_i| + |P_j| in place of |P_i + P_j|, since a shared pressure
that cancels still carries the rounding of its two inputs): FP32
accumulation of about 30 terms stays within 30 * 2^-24 < 2e-6 of it.
"""
import math

import numpy as np
import pytest

import pyoracle as O

TOL = 1e-5


def _dw2(d, h):
    return np.where(d <= h, -15.0 / (math.pi * h**5) * (h - d), 0.0)


def _dw3(d, h):
    return np.where(d <= h, -45.0 / (math.pi * h**6) * (h - d) ** 2, 0.0)


def _w2(d, h):
    return np.where(d < h, 15.0 / (2 * math.pi * h**5) * (h - d) ** 2, 0.0)


def _w3(d, h):
    return np.where(d < h, 15.0 / (math.pi * h**6) * (h - d) ** 3, 0.0)


def _poly6(d, h):
    return np.where(d < h, 315.0 / (64 * math.pi * abs(h) ** 9) * (h * h - d * d) ** 3, 0.0)


def pressure_terms(x, rho, rhon, h, rho0, k, kn, r2, variant="reference"):
    """Per-pair pressure terms of cc:395-418 in FP64 over all pairs j != i
    with d^2 <= r2: returns (F[n,3], M[n,3]) with F the force sum and M the
    sum of |terms| per component.  `variant` evaluates a misreading instead."""
    n = x.shape[0]
    o = x[None, :, :] - x[:, None, :]                 # offset from i to j (cc:402)
    d2 = (o**2).sum(-1)
    d = np.sqrt(d2)
    use = (d2 <= r2) & ~np.eye(n, dtype=bool)         # cc:396, 405
    P = (rho - rho0) * k                              # cc:371, 407
    Pn = rhon * kn                                    # cc:372, 408
    sp = 0.5 * (P[:, None] + P[None, :])              # cc:410
    snp = 0.5 * (Pn[:, None] + Pn[None, :])           # cc:411
    rj, rnj = np.broadcast_to(rho[None, :], (n, n)), np.broadcast_to(rhon[None, :], (n, n))
    if variant == "own_density":                      # rho_i in place of rho_j
        rj, rnj = np.broadcast_to(rho[:, None], (n, n)), np.broadcast_to(rhon[:, None], (n, n))
    elif variant == "own_pressure":                   # P_i in place of (P_i + P_j) / 2
        sp = np.broadcast_to(P[:, None], (n, n))
        snp = np.broadcast_to(Pn[:, None], (n, n))
    with np.errstate(invalid="ignore", divide="ignore"):
        dirv = np.where(d[..., None] > 0, o / d[..., None], np.array([0.0, 1.0, 0.0]))  # cc:414
    if variant == "reversed_direction":
        dirv = -dirv
    a = (_dw2(d, h) * sp / rj)[..., None] * dirv      # cc:416
    b = (_dw3(d, h) * snp / rnj)[..., None] * dirv    # cc:417
    t = np.where(use[..., None], a, 0.0), np.where(use[..., None], b, 0.0)
    F = t[0].sum(1) + t[1].sum(1)
    # scale: |P_i| + |P_j| in place of |P_i + P_j| (a shared pressure that
    # cancels keeps the size, and the rounding, of its two inputs)
    up = 0.5 * (np.abs(P)[:, None] + np.abs(P)[None, :])
    unp = 0.5 * (np.abs(Pn)[:, None] + np.abs(Pn)[None, :])
    ma = (np.abs(_dw2(d, h)) * up / np.abs(rj))[..., None] * np.abs(dirv)
    mb = (np.abs(_dw3(d, h)) * unp / np.abs(rnj))[..., None] * np.abs(dirv)
    M = np.where(use[..., None], ma + mb, 0.0).sum(1)
    return F, M


def viscosity_terms(x, v, h, r2, variant="reference"):
    """Jacobi viscosity terms of cc:442-461 in FP64: (F[n,3], M[n,3])."""
    n = x.shape[0]
    o = x[None, :, :] - x[:, None, :]
    d2 = (o**2).sum(-1)
    use = (d2 <= r2) & ~np.eye(n, dtype=bool)
    w = _poly6(np.sqrt(d2), h)
    dv = v[None, :, :] - v[:, None, :]                # v_j - v_i (cc:460)
    if variant == "reversed_difference":
        dv = -dv
    elif variant == "weight_of_squared_distance":     # poly6(d^2) in place of poly6(d)
        w = _poly6(d2, h)
    t = np.where(use[..., None], dv * w[..., None], 0.0)
    return t.sum(1), np.abs(t).sum(1)


def _line(spacings, h):
    x = np.zeros((len(spacings) + 1, 4), np.float32)
    x[1:, 0] = np.cumsum(spacings)
    return x


def _dens64(x, h, r2):
    d2 = ((x[None, :, :] - x[:, None, :]) ** 2).sum(-1)
    d = np.sqrt(d2)
    use = d2 <= r2
    return np.where(use, _w2(d, h), 0).sum(1), np.where(use, _w3(d, h), 0).sum(1)


@pytest.mark.parametrize("spacings", [(0.06, 0.11), (0.13, 0.04), (0.05, 0.05, 0.09)])
def test_pressure_closed_form_with_unequal_neighbours(spacings):
    h, rho0, k, kn, dt = 0.2, 60.0, 300.0, 20.0, 0.01
    p = O.params(h=h, rho0=rho0, k=k, k_near=kn)
    xp = _line(spacings, h)
    n = xp.shape[0]
    s = O.spatial(xp, n, h)
    dens = O.density(p, xp, s)
    v = O.pressure(p, dt, xp, s, dens, np.zeros((n, 4), np.float32))
    hh, r2 = float(np.float32(h)), float(np.float32(p.r2))
    x64 = xp[:, :3].astype(np.float64)
    rho, rhon = dens[:, 0].astype(np.float64), dens[:, 1].astype(np.float64)
    # the FP32 densities themselves are pinned by their own closed form
    r64, rn64 = _dens64(x64, hh, r2)
    np.testing.assert_allclose(rho, r64, rtol=1e-6)
    np.testing.assert_allclose(rhon, rn64, rtol=1e-6)
    assert len(set(np.round(rho, 3))) == n, "the densities must differ"
    F, M = pressure_terms(x64, rho, rhon, hh, rho0, k, kn, r2)
    want = F / rho[:, None] * dt                      # cc:421
    scale = M / np.abs(rho)[:, None] * dt
    err = np.abs(v[:, :3] - want)
    assert np.all(err <= TOL * scale + 1e-30), err / np.maximum(scale, 1e-30)
    assert np.all(v[:, 1:3] == 0)
    # each misreading misses by far more than the tolerance
    for variant in ("own_density", "own_pressure", "reversed_direction"):
        Fv, _ = pressure_terms(x64, rho, rhon, hh, rho0, k, kn, r2, variant)
        off = np.abs(Fv / rho[:, None] * dt - want)[:, 0] / scale[:, 0]
        assert off.max() > 1e3 * TOL, variant
    # dividing by the particle's own density vs the neighbour's at the end
    wrong = F / rho.mean() * dt
    assert (np.abs(wrong - want)[:, 0] / scale[:, 0]).max() > 1e3 * TOL
    # the FP64 twin meets the same closed form to rounding
    d64 = O.density64(p, xp, s)
    np.testing.assert_allclose(d64[:, 0], r64, rtol=1e-12)
    F64, M64 = pressure_terms(x64, d64[:, 0], d64[:, 1], hh, rho0, k, kn, r2)
    dv64, mag64 = O.pressure64(p, dt, xp, s, d64)
    np.testing.assert_allclose(dv64, F64 / d64[:, :1] * dt, rtol=1e-10, atol=1e-12 * mag64.max())
    np.testing.assert_allclose(mag64, M64 / d64[:, :1] * dt, rtol=1e-10)


@pytest.mark.parametrize("spacings", [(0.06, 0.11), (0.13, 0.04)])
def test_viscosity_closed_form_with_three_velocities(spacings):
    h, dt = 0.2, 0.01
    p = O.params(h=h, mu=0.3)
    mu = float(np.float32(p.mu))
    xp = _line(spacings, h)
    n = xp.shape[0]
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = np.array([[1.0, -0.5, 0.25], [-2.0, 0.75, 1.5], [0.5, 2.0, -1.0]], np.float32)[:n]
    s = O.spatial(xp, n, h)
    v = O.viscosity(p, dt, xp, s, vel, gauss_seidel=False)
    hh, r2 = float(np.float32(h)), float(np.float32(p.r2))
    x64, v64 = xp[:, :3].astype(np.float64), vel[:, :3].astype(np.float64)
    F, M = viscosity_terms(x64, v64, hh, r2)
    want = v64 + F * mu * dt                          # cc:463
    scale = M * mu * dt
    err = np.abs(v[:, :3] - want)
    # the new velocity is rounded once more to FP32: half an ulp of it
    assert np.all(err <= TOL * scale + 2.0**-24 * np.abs(want)), err
    for variant in ("reversed_difference", "weight_of_squared_distance"):
        Fv, _ = viscosity_terms(x64, v64, hh, r2, variant)
        off = np.abs(Fv - F) * mu * dt / np.maximum(scale, 1e-30)
        assert off.max() > 1e3 * TOL, variant
    dv64, mag64 = O.viscosity64(p, dt, xp, s, v64)
    np.testing.assert_allclose(dv64, F * mu * dt, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(mag64, scale, rtol=1e-10)


def _random_scene(n, box, seed, rng_v=None):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.5, 0.5, (n, 3)) * np.asarray(box, np.float32)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0.0, 1.0, (n, 3))
    return pos, vel


@pytest.mark.parametrize("seed,n,box", [(11, 1200, (1.0, 1.0, 1.0)), (12, 900, (1.6, 0.6, 1.2))])
def test_forces_against_all_pairs_fp64(seed, n, box):
    h, dt = 0.2, 0.01
    xp, vel = _random_scene(n, box, seed)
    hh = float(np.float32(h))
    x64 = xp[:, :3].astype(np.float64)
    p = O.params(h=h, rho0=0.0, k=300.0, k_near=20.0, mu=0.2)
    mu = float(np.float32(p.mu))
    r2 = float(np.float32(p.r2))
    r64, rn64 = _dens64(x64, hh, r2)
    # a target density inside the range of the scene's densities: pressures
    # of both signs, neighbours of unequal density
    p.rho0 = float(np.float32(np.median(r64)))
    assert r64.min() < p.rho0 < r64.max()
    s = O.spatial(xp, n, p.h)
    dens = O.density(p, xp, s)
    # FP32 pressure stage vs the FP64 all-pairs sum of the same inputs
    rho, rhon = dens[:, 0].astype(np.float64), dens[:, 1].astype(np.float64)
    F, M = pressure_terms(x64, rho, rhon, hh, p.rho0, p.k, p.k_near, r2)
    v0 = np.zeros((n, 4), np.float32)
    vp = O.pressure(p, dt, xp, s, dens, v0)
    want = F / rho[:, None] * dt
    scale = M / rho[:, None] * dt
    assert np.all(np.abs(vp[:, :3] - want) <= TOL * scale + 1e-30)
    # the FP64 twin on the walk equals the all-pairs sum
    d64 = O.density64(p, xp, s)
    np.testing.assert_allclose(d64[:, 0], r64, rtol=1e-12)
    np.testing.assert_allclose(d64[:, 1], rn64, rtol=1e-12)
    dv64, mag64 = O.pressure64(p, dt, xp, s, dens.astype(np.float64))
    assert np.all(np.abs(dv64 - want) <= 1e-10 * scale + 1e-300)
    np.testing.assert_allclose(mag64, scale, rtol=1e-10)
    # Jacobi viscosity (FP32) vs the FP64 all-pairs sum from the same velocities
    vv = O.viscosity(p, dt, xp, s, vel, gauss_seidel=False)
    v64 = vel[:, :3].astype(np.float64)
    Fv, Mv = viscosity_terms(x64, v64, hh, r2)
    wantv = v64 + Fv * mu * dt
    scalev = Mv * mu * dt
    assert np.all(np.abs(vv[:, :3] - wantv) <= TOL * scalev + 2.0**-24 * np.abs(wantv))
    dvv, magv = O.viscosity64(p, dt, xp, s, v64)
    assert np.all(np.abs(dvv - Fv * mu * dt) <= 1e-10 * scalev + 1e-300)
    np.testing.assert_allclose(magv, scalev, rtol=1e-10)
