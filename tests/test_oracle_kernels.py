"""Pins the oracle's smoothing kernels and spatial hash against mathematics,
not against a retyped copy of its formulas.

Reference: engine/physics/kernels.h:25-82 (kernels), physicsWorld.cc:499-516
(cell coordinate, hash, key).  Checks:
  - normalisation: each density-type kernel integrates to 1 over the ball of
    radius h (a wrong constant, power of h or factor of 2 fails);
  - derivative relations: dW2 = d/dd W2 and dW3 = d/dd W3 numerically (a wrong
    sign or exponent fails);
  - support: zero at and beyond h (strict < for W2, W3, poly6; <= for dW);
  - known-answer values at d = 0 listed with the reference analysis;
  - hash: known answers, two's-complement wrap of negative cells, linearity
    mod 2^32, and the product's integer path (sph_hash_cell / sph_bucket_of)
    against Python's arbitrary-precision arithmetic.
"""
import math

import numpy as np
import pytest

import pyoracle as O
import sph_amd as S

H_VALUES = [0.1, 0.2, 0.35, 1.0]


def _ball_integral(f, h, n=20000):
    # 4 pi int_0^h r^2 f(r) dr, midpoint rule in float64 on the FP32 kernel
    r = (np.arange(n) + 0.5) * (h / n)
    vals = np.array([f(float(x), h) for x in r], dtype=np.float64)
    return 4.0 * math.pi * float(np.sum(r * r * vals)) * (h / n)


@pytest.mark.parametrize("h", H_VALUES)
@pytest.mark.parametrize("name", ["W2", "W3", "poly6"])
def test_density_kernels_normalised(name, h):
    f = getattr(O, name)
    assert _ball_integral(f, h) == pytest.approx(1.0, rel=2e-4)


@pytest.mark.parametrize("h", H_VALUES)
def test_derivative_kernels_are_derivatives(h):
    for d in np.linspace(0.05 * h, 0.95 * h, 7):
        eps = 1e-3 * h
        for W, dW in ((O.W2, O.dW2), (O.W3, O.dW3)):
            num = (W(d + eps, h) - W(d - eps, h)) / (2 * eps)
            assert dW(d, h) == pytest.approx(num, rel=2e-3)


@pytest.mark.parametrize("h", H_VALUES)
def test_kernel_support(h):
    hf = float(np.float32(h))
    for W in (O.W2, O.W3, O.poly6):
        assert W(hf, hf) == 0.0
        assert W(hf * 1.5, hf) == 0.0
        assert W(0.0, hf) > 0.0
    for dW in (O.dW2, O.dW3):
        assert dW(hf, hf) == 0.0          # (h - h) * scale: zero, inside the <= branch
        assert dW(hf * 1.5, hf) == 0.0
        assert dW(0.5 * hf, hf) < 0.0     # repulsive slope


def test_kernel_known_answers():
    # values at d = 0 (FP64 closed forms, listed with the reference analysis)
    kat = {(0.2, "W2"): 298.415518, (0.2, "W3"): 596.831037, (0.2, "poly6"): 195.835184,
           (0.2, "dW2"): -2984.15518, (0.2, "dW3"): -8952.46555,
           (0.35, "W2"): 55.6810297, (0.35, "W3"): 111.362059}
    for (h, name), want in kat.items():
        assert getattr(O, name)(0.0, h) == pytest.approx(want, rel=2e-6)


def _hash_py(cx, cy, cz):
    return (15823 * cx + 9737333 * cy + 440817757 * cz) % (1 << 32)


def test_hash_known_answers():
    kat = {(1, 1, 1): 450570913, (-1, 0, 0): 4294951473, (-1, -3, -20): 4039318926,
           (-10, -20, -10): 3986852132}
    for c, want in kat.items():
        assert _hash_py(*c) == want
        assert O.hash_cell(np.array(c, np.float32)) == want    # float cell path (cc:505-511)
        assert S.hash_cell(*c) == want                         # integer path of the kernels


def test_hash_wraps_and_is_linear():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        c = rng.integers(-5000, 5000, 3)
        d = rng.integers(-3, 4, 3)
        h0 = O.hash_cell(c.astype(np.float32))
        assert h0 == _hash_py(*c)
        assert S.hash_cell(*(c + d)) == (h0 + _hash_py(*d)) % (1 << 32)


def test_bucket_of_matches_modulo():
    rng = np.random.default_rng(1)
    tables = [1, 2, 3, 7, 1000, 4096, 16384, 65536, 131072, 1 << 21, 1 << 24, 2**31 - 1,
              2**32 - 1] + list(rng.integers(1, 2**32, 40))
    hashes = [0, 1, 2**31, 2**32 - 1] + list(rng.integers(0, 2**32, 300))
    for T in tables:
        for h in hashes:
            assert S.bucket_of(int(h), int(T)) == int(h) % int(T)
            assert O.key(int(h), int(T)) == int(h) % int(T)


def test_cell_coordinates_floor_semantics():
    h = 0.2
    cases = {(0.0, 0.0, 0.0): (0, 0, 0), (-0.1, 0.1, 0.39): (-1, 0, 1),
             (-1e-9, 0.2, -0.2): (-1, 1, -1), (3.99, -4.01, 0.6): (19, -21, 3)}
    for xp, want in cases.items():
        got = O.cell_coord(np.array(xp, np.float32), h)
        xp32 = np.array(xp, np.float32)
        # numpy float32 division is IEEE; floor is exact
        assert tuple(int(v) for v in got) == tuple(int(v) for v in np.floor(xp32 / np.float32(h)))
        assert tuple(int(v) for v in got) == want


def test_no_small_cell_offset_aliases_a_hash():
    """Two distinct cells within 4 cells of each other never share a full hash,
    so the float-hash filter of the bucket walk only removes foreign cells."""
    r = np.arange(-8, 9)
    dx, dy, dz = np.meshgrid(r, r, r, indexing="ij")
    hv = (15823 * dx.astype(np.int64) + 9737333 * dy + 440817757 * dz) % (1 << 32)
    nz = (dx != 0) | (dy != 0) | (dz != 0)
    assert not np.any(hv[nz] == 0)
