"""sph_update_noalloc, the allocation-free entry point, against the handle
path (sph_create + sph_upload + sph_update + sph_download) that runs the same
kernels with buffers the library allocates: bit-exact positions, velocities
and densities, on two calls in a row, the second with a smaller scene (the
scratch array keeps the first call's contents), and against the oracle.

The caller's device buffers come from the HIP runtime the library itself
uses (hipMalloc through ctypes): the caller allocates, the entry point does
not.  The size check runs on any host (no device call precedes it).
"""
import ctypes

import numpy as np
import pytest

import pyoracle as O
import sph_amd as S
from helpers import announce, assert_bitexact, orc_params
from sph_amd import workloads as W

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        S.lib()  # loads the runtime the library links against
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.malloc(arr.nbytes)
        assert self.L.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self, *ps):
        for p in ps:
            self.L.hipFree(p)


def _entry():
    fn = S.lib().sph_update_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.POINTER(S.Params), ctypes.c_uint32, ctypes.c_float] + [ctypes.c_void_p] * 5 + \
        [ctypes.c_void_p, ctypes.c_int]
    return fn


def _noalloc(hip, sc, dt, wait=1):
    fn = _entry()
    n = sc.n
    pin, vin = hip.put(sc.pos), hip.put(sc.vel)
    pout, vout, dout = hip.malloc(16 * n), hip.malloc(16 * n), hip.malloc(8 * n)
    try:
        p = sc.params
        rc = fn(ctypes.byref(p), n, dt, pin, vin, pout, vout, dout, None, wait)
        assert rc == 0, S.last_error()
        if not wait:
            assert hip.L.hipDeviceSynchronize() == 0
        return (hip.get(pout, (n, 4), np.float32), hip.get(vout, (n, 4), np.float32),
                hip.get(dout, (n, 2), np.float32))
    finally:
        hip.free(pin, vin, pout, vout, dout)


def _handle(sc, dt):
    with S.Sim(sc.params, sc.n, S.TABLE_GRID) as sim:
        sim.upload(sc.pos, sc.vel)
        sim.update(dt)
        return sim.download()


@pytest.mark.gpu
def test_noalloc_equals_handle_path_twice_smaller_second():
    hip = _Hip()
    dt = float(np.float32(np.float32(W.DT) / np.float32(W.SUBSTEPS)))
    first = W.dam_break(n=16384)
    second = W.uniform_lower(n=9000, table=9000, box=(2.0, 1.0, 2.0))
    for sc in (first, second):
        announce("noalloc", sc.name, sc.n)
        got = _noalloc(hip, sc, dt)
        want = _handle(sc, dt)
        for g, w, what in zip(got, want, ("pos", "vel", "dens")):
            assert_bitexact(g, w, f"{sc.name} {what}: noalloc vs handle")
        pos_o, vel_o, dens_o = O.update(orc_params(sc.params), dt, sc.pos, sc.vel, T=sc.n)
        assert_bitexact(got[0][:, :3], pos_o[:, :3], f"{sc.name} pos vs oracle")
        assert_bitexact(got[1][:, :3], vel_o[:, :3], f"{sc.name} vel vs oracle")
        assert_bitexact(got[2], dens_o, f"{sc.name} dens vs oracle")
    # an asynchronous call (wait = 0) on the null stream, then the first scene again
    got = _noalloc(hip, first, dt, wait=0)
    want = _handle(first, dt)
    assert_bitexact(got[1], want[1], "async call vel")


def test_noalloc_rejects_sizes_beyond_its_maximum():
    fn = _entry()
    p = W.scene_params()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = fn(ctypes.byref(p), 65537, 0.001, dummy, dummy, dummy, dummy, None, None, 1)
    assert rc == S.SPH_ERR_INVALID and "65536" in S.last_error()
    rc = fn(ctypes.byref(p), 0, 0.001, dummy, dummy, dummy, dummy, None, None, 1)
    assert rc == S.SPH_ERR_INVALID
