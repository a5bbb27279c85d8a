"""C-ABI library checks that need no GPU: it loads, exports every entry point
include/sph_amd.h declares, its host-side helpers are right, and it fails
cleanly where a GPU is required but absent."""
import ctypes
import re

import numpy as np
import pytest

import sph_amd as S


def _declared():
    text = open(S.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_entry_point():
    lib = ctypes.CDLL(S.LIB_PATH)
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_default_params_are_the_reference_defaults():
    p = S.default_params()
    assert p.smoothing_radius == np.float32(0.35)   # physicsWorld.h:97
    assert p.target_density == np.float32(99.7)     # :98
    assert p.pressure_mult == 300.0 and p.near_pressure_mult == 20.0 and p.viscosity == 0.5
    assert p.gravity_on == 0 and p.gravity_scale == 10.0
    assert tuple(p.bounds) == (20.0, 20.0, 20.0)
    assert p.lookahead == np.float32(1.0 / 120.0) and p.damp == np.float32(0.95)
    assert p.prefilter_r2 == np.float32(np.float32(0.35) * np.float32(0.35))


@pytest.mark.parametrize("kind", [S.INIT_GRID_ARRANGEMENT, S.INIT_DAM_BREAK, S.INIT_UNIFORM_LOWER,
                                  S.INIT_TWO_BLOCKS, S.INIT_CENTRED_BLOCK])
def test_generate_is_deterministic_and_inside_the_box(kind):
    box = (4.0, 4.0, 4.0)
    a = S.generate(kind, 7, 5000, box, 0.1, 0.002, 2.0)
    b = S.generate(kind, 7, 5000, box, 0.1, 0.002, 2.0)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    if kind != S.INIT_GRID_ARRANGEMENT:
        assert np.all(np.abs(a[0][:, :3]) <= np.asarray(box, np.float32) / 2)
    c = S.generate(kind, 8, 5000, box, 0.1, 0.002, 2.0)
    if kind in (S.INIT_UNIFORM_LOWER, S.INIT_DAM_BREAK):
        assert not np.array_equal(a[0], c[0])   # the seed matters where there is randomness


def test_dam_break_16k_layout():
    pos, vel = S.generate(S.INIT_DAM_BREAK, 0, 16384, (4.0, 4.0, 4.0), 0.1, 0.0)
    # 16 x 32 x 32 block, z fastest, cell-centred at the -x,-y,-z corner
    np.testing.assert_allclose(pos[0, :3], [-1.95, -1.95, -1.95], atol=1e-6)
    np.testing.assert_allclose(pos[1, :3], [-1.95, -1.95, -1.85], atol=1e-6)
    np.testing.assert_allclose(pos[-1, :3], [-1.95 + 1.5, -1.95 + 3.1, -1.95 + 3.1], atol=1e-5)
    assert np.all(vel == 0)


def test_two_blocks_move_towards_each_other():
    pos, vel = S.generate(S.INIT_TWO_BLOCKS, 2, 20000, (40.0, 20.0, 40.0), 0.1, 0.0, 2.0)
    left = vel[:, 0] > 0
    assert left.sum() == 10000
    assert pos[left, 0].max() < pos[~left, 0].min()
    assert np.all(vel[~left, 0] == -2.0)


def test_generate_rejects_bad_input():
    with pytest.raises(S.SphError):
        S.generate(S.INIT_DAM_BREAK, 0, 10, (4.0, 4.0, 4.0), -1.0)
    with pytest.raises(S.SphError):
        S.generate(99, 0, 10)


def test_create_without_device_fails_cleanly():
    if S.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(S.SphError) as e:
        S.Sim(S.default_params(), 100)
    assert e.value.code == S.SPH_ERR_HIP
    assert "no HIP device" in S.last_error()


def test_null_handle_is_rejected():
    lib = S.lib()
    assert lib.sph_update(None, 0.01) == S.SPH_ERR_INVALID
    assert lib.sph_destroy(None) == S.SPH_OK


def test_kernel_quotient_equals_ieee_division():
    """The pressure kernel's quotient (reciprocal RN(1/b) with one correction,
    Markstein) against IEEE float32 division: every significand of b in [1, 2)
    with fixed numerators, every significand of a with sampled b, scaled
    operands, zeros, and the out-of-range fallbacks."""
    rng = np.random.default_rng(7)
    mb = np.arange(1 << 23, dtype=np.uint32)
    b_all = (np.uint32(0x3F800000) | mb).view(np.float32)
    for a0 in (1.0, 1.9999999, 1.5, -1.25, 1.0000001):
        a = np.full(b_all.shape, a0, np.float32)
        got = S.debug_quotients(a, b_all)
        want = (a / b_all).astype(np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), a0
    a_all = (np.uint32(0x3F800000) | mb).view(np.float32)
    for b0 in rng.uniform(1.0, 2.0, 4).astype(np.float32):
        b = np.full(a_all.shape, b0, np.float32)
        got = S.debug_quotients(a_all, b)
        assert np.array_equal(got.view(np.uint32), (a_all / b).view(np.uint32)), b0
    # scaled operands across the whole range, including outside the fast range
    e_a = rng.integers(-140, 127, 1 << 20)
    e_b = rng.integers(-30, 30, 1 << 20)
    a = (rng.uniform(1, 2, 1 << 20) * np.exp2(e_a.astype(np.float64))).astype(np.float32)
    a *= np.where(rng.random(1 << 20) < 0.5, -1, 1).astype(np.float32)
    b = (rng.uniform(1, 2, 1 << 20) * np.exp2(e_b.astype(np.float64))).astype(np.float32)
    with np.errstate(over="ignore", under="ignore"):
        want = (a / b).astype(np.float32)
    got = S.debug_quotients(a, b)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the edges of the fast range and the values outside it (tiny, huge,
    # subnormal, infinite, NaN numerators; divisors at and beyond 2^-26, 2^26)
    def bits(u):
        return np.array(u, np.uint32).view(np.float32)
    edge_a = bits([0x12000000, 0x11FFFFFF, 0x71800000, 0x71800001, 0x00000001, 0x007FFFFF,
                   0x7F800000, 0x7FC00000, 0x92000000, 0xF1800000, 0x3F800000])
    edge_b = bits([0x32800000, 0x327FFFFF, 0x4C800000, 0x4C800001, 0x3F800000, 0x3FFFFFFF])
    a, b = np.meshgrid(edge_a, edge_b)
    a, b = a.ravel(), b.ravel()
    with np.errstate(over="ignore", under="ignore", invalid="ignore"):
        want = (a / b).astype(np.float32)
    got = S.debug_quotients(a, b)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all(), list(zip(a[~same], b[~same]))
    # zeros: +0 exact; -0 gives +0 (a zero term leaves a sum from +0 unchanged)
    z = S.debug_quotients(np.array([0.0, -0.0], np.float32), np.array([3.0, 3.0], np.float32))
    assert z[0] == 0 and not np.signbit(z[0]) and z[1] == 0
