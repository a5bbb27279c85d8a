"""ctypes wrapper of liboracle.so -- the CPU restatement of the reference SPH step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always to check results.  The product library
(fluid-simulation-3d_amd/) never imports it.

Each wrapper names the reference function it restates (Allkams/Fluid-Simulation-3D,
engine/physics/physicsWorld.cc and kernels.h); see oracle.h for the full map.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
START_EMPTY = 0x7FFFFFFF  # INT_MAX, physicsWorld.cc:481


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("h", ctypes.c_float),
        ("rho0", ctypes.c_float),
        ("k", ctypes.c_float),
        ("k_near", ctypes.c_float),
        ("mu", ctypes.c_float),
        ("gravity_on", ctypes.c_int32),
        ("gravity_scale", ctypes.c_float),
        ("bounds", ctypes.c_float * 3),
        ("lookahead", ctypes.c_float),
        ("damp", ctypes.c_float),
        ("r2", ctypes.c_float),
    ]


def build():
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        f, u32, i32, p = ctypes.c_float, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p
        pp = ctypes.POINTER(OrcParams)
        sig = {
            "orc_default_params": (None, [pp]),
            "orc_W2": (f, [f, f]), "orc_W3": (f, [f, f]), "orc_dW2": (f, [f, f]),
            "orc_dW3": (f, [f, f]), "orc_poly6": (f, [f, f]),
            "orc_cell_coord": (None, [p, f, p]),
            "orc_hash_cell": (u32, [p]),
            "orc_key": (u32, [u32, u32]),
            "orc_set_grid_key": (None, [p, p]),
            "orc_set_reference_sort": (None, [i32]),
            "orc_stage_predict": (None, [u32, pp, f, p, p, p]),
            "orc_stage_spatial": (i32, [u32, u32, f, p, p, p, p, p, p]),
            "orc_stage_density": (None, [u32, u32, pp, p, p, p, p, p, p]),
            "orc_stage_pressure": (None, [u32, u32, pp, f, p, p, p, p, p, p, p]),
            "orc_stage_viscosity": (None, [u32, u32, pp, f, p, p, p, p, p, p, i32]),
            "orc_stage_integrate": (None, [u32, pp, f, p, p]),
            "orc_update": (i32, [u32, u32, pp, f, p, p, p, i32, p]),
            "orc_neighbours": (ctypes.c_size_t, [u32, u32, pp, p, p, p, p, p, p, p]),
            "orc64_density": (None, [u32, u32, pp, p, p, p, p, p, p]),
            "orc64_pressure": (None, [u32, u32, pp, ctypes.c_double, p, p, p, p, p, p, p, p]),
            "orc64_viscosity": (None, [u32, u32, pp, ctypes.c_double, p, p, p, p, p, p, p, p]),
            "orc_grid_arrangement": (None, [u32, p]),
            "orc_speed_colors": (None, [u32, p, p, p]),
            "orc_set_threads": (None, [i32]),
            "orc_get_threads": (i32, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def params(**kw):
    """Reference defaults (physicsWorld.h:96-106,145) overridden by kw."""
    p = OrcParams()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "bounds":
            p.bounds[:] = [float(x) for x in v]
        else:
            setattr(p, k, v)
    return p


def from_sph(p):
    """Oracle params from an object with the sph_params field names."""
    return params(h=p.smoothing_radius, rho0=p.target_density, k=p.pressure_mult,
                  k_near=p.near_pressure_mult, mu=p.viscosity, gravity_on=p.gravity_on,
                  gravity_scale=p.gravity_scale, bounds=tuple(p.bounds),
                  lookahead=p.lookahead, damp=p.damp, r2=p.prefilter_r2)


def set_threads(n):
    lib().orc_set_threads(int(n))


def set_reference_sort(on):
    """Timing hook: stage B2 as std::sort of 12-byte float records (the
    reference's cc:484) instead of the stable counting sort."""
    lib().orc_set_reference_sort(1 if on else 0)


def W2(d, h): return lib().orc_W2(d, h)
def W3(d, h): return lib().orc_W3(d, h)
def dW2(d, h): return lib().orc_dW2(d, h)
def dW3(d, h): return lib().orc_dW3(d, h)
def poly6(d, h): return lib().orc_poly6(d, h)


def cell_coord(xp, h):
    xp = np.ascontiguousarray(xp, dtype=np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_cell_coord(_ptr(xp), h, _ptr(out))
    return out


def hash_cell(cell):
    c = np.ascontiguousarray(cell, dtype=np.float32)
    return lib().orc_hash_cell(_ptr(c))


def key(hash_, T):
    return lib().orc_key(hash_, T)


def set_grid_key(gdim=None, lo=None):
    """Test hook: bucket by the cell-grid key (gdim, lo) instead of hash % T;
    no arguments restore the reference key."""
    if gdim is None:
        lib().orc_set_grid_key(None, None)
        return
    g = np.ascontiguousarray(gdim, dtype=np.uint32)
    o = np.ascontiguousarray(lo, dtype=np.int32)
    lib().orc_set_grid_key(_ptr(g), _ptr(o))


class grid_key:
    """Context manager form of set_grid_key."""

    def __init__(self, gdim, lo):
        self.gdim, self.lo = gdim, lo

    def __enter__(self):
        set_grid_key(self.gdim, self.lo)
        return self

    def __exit__(self, *exc):
        set_grid_key()
        return False


class Spatial:
    """Output of stage B (UpdateSpatialLookup, physicsWorld.cc:466-498)."""

    def __init__(self, n, T):
        self.n, self.T = n, T
        self.hash_id = np.zeros(n, np.uint32)
        self.key_id = np.zeros(n, np.uint32)
        self.sorted_key = np.zeros(n, np.uint32)
        self.sorted_idx = np.zeros(n, np.uint32)
        self.start = np.zeros(T, np.uint32)

    def args(self):
        return (_ptr(self.sorted_key), _ptr(self.sorted_idx), _ptr(self.hash_id), _ptr(self.start))


def predict(p, dt, pos4, vel4):
    """Stage A (cc:42-48). Returns (vel4_new, xp4)."""
    n = pos4.shape[0]
    vel = np.ascontiguousarray(vel4, dtype=np.float32).copy()
    xp = np.zeros((n, 4), np.float32)
    lib().orc_stage_predict(n, ctypes.byref(p), dt, _ptr(np.ascontiguousarray(pos4, np.float32)),
                            _ptr(vel), _ptr(xp))
    return vel, xp


def spatial(xp4, T, h):
    n = xp4.shape[0]
    s = Spatial(n, T)
    rc = lib().orc_stage_spatial(n, T, h, _ptr(np.ascontiguousarray(xp4, np.float32)),
                                 _ptr(s.hash_id), _ptr(s.key_id), _ptr(s.sorted_key),
                                 _ptr(s.sorted_idx), _ptr(s.start))
    if rc != 0:
        raise ValueError("orc_stage_spatial rejected its input")
    return s


def density(p, xp4, s):
    out = np.zeros((s.n, 2), np.float32)
    lib().orc_stage_density(s.n, s.T, ctypes.byref(p), _ptr(xp4), *s.args(), _ptr(out))
    return out


def pressure(p, dt, xp4, s, dens2, vel4):
    v = np.ascontiguousarray(vel4, np.float32).copy()
    lib().orc_stage_pressure(s.n, s.T, ctypes.byref(p), dt, _ptr(xp4), *s.args(),
                             _ptr(np.ascontiguousarray(dens2, np.float32)), _ptr(v))
    return v


def viscosity(p, dt, xp4, s, vel4, gauss_seidel=False):
    v = np.ascontiguousarray(vel4, np.float32).copy()
    lib().orc_stage_viscosity(s.n, s.T, ctypes.byref(p), dt, _ptr(xp4), *s.args(), _ptr(v),
                              1 if gauss_seidel else 0)
    return v


def integrate(p, dt, pos4, vel4):
    pos = np.ascontiguousarray(pos4, np.float32).copy()
    vel = np.ascontiguousarray(vel4, np.float32).copy()
    lib().orc_stage_integrate(pos.shape[0], ctypes.byref(p), dt, _ptr(pos), _ptr(vel))
    return pos, vel


def update(p, dt, pos4, vel4, T=None, gauss_seidel=False, substeps=1, stage_ms=None):
    """`substeps` calls of Update(dt) (cc:39-111), in place on copies.  Returns (pos, vel, dens)."""
    n = pos4.shape[0]
    T = n if T is None else T
    pos = np.ascontiguousarray(pos4, np.float32).copy()
    vel = np.ascontiguousarray(vel4, np.float32).copy()
    dens = np.zeros((n, 2), np.float32)
    ms = np.zeros(6, np.float64)
    for _ in range(substeps):
        rc = lib().orc_update(n, T, ctypes.byref(p), dt, _ptr(pos), _ptr(vel), _ptr(dens),
                              1 if gauss_seidel else 0, _ptr(ms))
        if rc != 0:
            raise ValueError("orc_update rejected its input")
        if stage_ms is not None:
            stage_ms += ms
    return pos, vel, dens


def neighbours(p, xp4, s):
    """CSR (offs, list) of the reference walk's in-prefilter neighbours, self included."""
    offs = np.zeros(s.n + 1, np.uint64)
    total = lib().orc_neighbours(s.n, s.T, ctypes.byref(p), _ptr(xp4), *s.args(), _ptr(offs), None)
    lst = np.zeros(max(total, 1), np.uint32)
    lib().orc_neighbours(s.n, s.T, ctypes.byref(p), _ptr(xp4), *s.args(), _ptr(offs), _ptr(lst))
    return offs, lst[:total]


def density64(p, xp4, s):
    out = np.zeros((s.n, 2), np.float64)
    lib().orc64_density(s.n, s.T, ctypes.byref(p), _ptr(xp4), *s.args(), _ptr(out))
    return out


def pressure64(p, dt, xp4, s, dens2_64):
    dv = np.zeros((s.n, 3), np.float64)
    mag = np.zeros((s.n, 3), np.float64)
    lib().orc64_pressure(s.n, s.T, ctypes.byref(p), dt, _ptr(xp4), *s.args(),
                         _ptr(np.ascontiguousarray(dens2_64, np.float64)), _ptr(dv), _ptr(mag))
    return dv, mag


def viscosity64(p, dt, xp4, s, vin3):
    dv = np.zeros((s.n, 3), np.float64)
    mag = np.zeros((s.n, 3), np.float64)
    lib().orc64_viscosity(s.n, s.T, ctypes.byref(p), dt, _ptr(xp4), *s.args(),
                          _ptr(np.ascontiguousarray(vin3, np.float64)), _ptr(dv), _ptr(mag))
    return dv, mag


def grid_arrangement(n):
    pos = np.zeros((n, 4), np.float32)
    lib().orc_grid_arrangement(n, _ptr(pos))
    return pos


REFERENCE_COLOR_STOPS = np.array([0.0, 0.75, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0,
                                  1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0], np.float32)


def speed_colors(vel4, stops=None):
    """FluidSimCPU::updateColors (fluidSimCPU.cc:100-125): RGBA per particle."""
    v = np.ascontiguousarray(vel4, np.float32)
    st = REFERENCE_COLOR_STOPS if stops is None else np.ascontiguousarray(stops, np.float32)
    out = np.zeros((v.shape[0], 4), np.float32)
    lib().orc_speed_colors(v.shape[0], _ptr(v), _ptr(st), _ptr(out))
    return out
