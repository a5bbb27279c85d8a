"""Python binding of libsph_amd.so (the C ABI in include/sph_amd.h).

The SPH step itself runs in the hand-written gfx950 kernels of libsph_amd.so;
this module only marshals numpy arrays across the C ABI.  There is no CPU
fallback: every call that needs the GPU fails with SphError when the library or
the device is missing.

Mirrors the reference's FluidSimulation interface (Allkams/Fluid-Simulation-3D,
engine/physics/physicsWorld.h:28-155): `Sim.update(dt)` is Update(dt)
(physicsWorld.cc:39-111), `Sim.initialize_data()` is InitializeData
(cc:112-147), the settings fields are the get/set pairs of physicsWorld.h:52-77.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_DIR, "libsph_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "sph_amd.h")

SPH_OK = 0
SPH_ERR_INVALID = -1
SPH_ERR_HIP = -2
SPH_ERR_OOM = -3
SPH_ERR_BUDGET = -4
SPH_ERR_STATE = -5
SPH_ERR_NOT_FOUND = -6
SPH_ERR_NONFINITE = -7
SPH_ERR_DEVICE = -8

# bucket keys of sph_create's table_size (sph_amd.h)
TABLE_GRID = 0
TABLE_REFERENCE = 0xFFFFFFFF
START_EMPTY = 0x7FFFFFFF  # INT_MAX, physicsWorld.cc:481

INIT_GRID_ARRANGEMENT = 0
INIT_DAM_BREAK = 1
INIT_UNIFORM_LOWER = 2
INIT_TWO_BLOCKS = 3
INIT_CENTRED_BLOCK = 4


class SphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sph error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """sph_params: the reference's settings (physicsWorld.h:96-106,145)."""

    _fields_ = [
        ("smoothing_radius", ctypes.c_float),
        ("target_density", ctypes.c_float),
        ("pressure_mult", ctypes.c_float),
        ("near_pressure_mult", ctypes.c_float),
        ("viscosity", ctypes.c_float),
        ("gravity_on", ctypes.c_int32),
        ("gravity_scale", ctypes.c_float),
        ("bounds", ctypes.c_float * 3),
        ("lookahead", ctypes.c_float),
        ("damp", ctypes.c_float),
        ("prefilter_r2", ctypes.c_float),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["bounds"] = tuple(self.bounds)
        return d


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_CANCELLED = 0, 1, 2, 3


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_running", ctypes.c_uint32),
                ("initial_capacity", ctypes.c_uint32), ("dt", ctypes.c_float),
                ("substeps", ctypes.c_int32), ("updates_per_tick", ctypes.c_int32),
                ("max_capacity", ctypes.c_uint32)]


class TickInfo(ctypes.Structure):
    _fields_ = [("running", ctypes.c_uint32), ("queued", ctypes.c_uint32),
                ("admitted", ctypes.c_uint32), ("finished", ctypes.c_uint32),
                ("particles", ctypes.c_uint64), ("ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("admitted", "finished", "rejected", "cancelled",
                                               "queued", "running", "used_bytes",
                                               "particle_updates", "ticks", "capacity")] + \
               [(k, ctypes.c_double) for k in ("busy_s", "tick_ms_p50", "tick_ms_p99",
                                               "request_ms_p50", "request_ms_p99", "wait_ms_p50",
                                               "wait_ms_p99")] + \
               [("admission_failures", ctypes.c_uint64)]


def _struct_dict(st):
    return {k: getattr(st, k) for k, _ in st._fields_}


def build():
    """Compile libsph_amd.so (hipcc, gfx950) with the package Makefile."""
    subprocess.run(["make", "-s", "-j8", "-C", PKG_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SphError(SPH_ERR_STATE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    i32, u32, u64, f, p = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_void_p
    H = ctypes.c_void_p
    pp = ctypes.POINTER(Params)
    sig = {
        "sph_default_params": (None, [pp]),
        "sph_last_error": (ctypes.c_char_p, []),
        "sph_version": (ctypes.c_char_p, []),
        "sph_device_count": (i32, []),
        "sph_create": (i32, [pp, u32, u32, ctypes.POINTER(H)]),
        "sph_destroy": (i32, [H]),
        "sph_set_params": (i32, [H, pp]),
        "sph_get_params": (i32, [H, pp]),
        "sph_get_size": (i32, [H, p, p]),
        "sph_upload": (i32, [H, p, p]),
        "sph_initialize_data": (i32, [H]),
        "sph_initialize": (i32, [H, i32, u64, p, f, f, f]),
        "sph_set_colors": (i32, [H, p, i32]),
        "sph_download_colors": (i32, [H, p]),
        "sph_render_buffers": (i32, [H, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]),
        "sph_batch_initialize": (i32, [H, u32, i32, u64, p, f, f, f]),
        "sph_server_submit_generated": (i32, [H, pp, u32, u32, i32, u64, p, f, f, f, u32, p]),
        "sph_update": (i32, [H, f]),
        "sph_step": (i32, [H, f, i32]),
        "sph_step_async": (i32, [H, f, i32]),
        "sph_sync": (i32, [H]),
        "sph_download": (i32, [H, p, p, p]),
        "sph_set_timing": (i32, [H, i32]),
        "sph_stage_ms": (i32, [H, p]),
        "sph_count_nonfinite": (i32, [H, p]),
        "sph_profile_begin": (i32, [H, u32]),
        "sph_autotune": (i32, [H, f, i32, p]),
        "sph_get_variants": (i32, [H, p]),
        "sph_set_variants": (i32, [H, p]),
        "sph_profile_end": (i32, [H, p, p]),
        "sph_debug_spatial": (i32, [H, p, p, p, p, p, p]),
        "sph_debug_predicted": (i32, [H, p]),
        "sph_debug_velocities": (i32, [H, p, p]),
        "sph_debug_set_scan_spin": (i32, [H, u32]),
        "sph_debug_neighbours": (i32, [H, p, p, u64, p]),
        "sph_generate": (i32, [i32, u64, u32, p, f, f, f, p, p]),
        "sph_server_default_config": (None, [ctypes.POINTER(ServerConfig)]),
        "sph_server_request_bytes": (u64, [ctypes.POINTER(Params), u32, u32]),
        "sph_get_key": (i32, [H, p, p, p]),
        "sph_server_create": (i32, [ctypes.POINTER(ServerConfig), ctypes.POINTER(H)]),
        "sph_server_destroy": (i32, [H]),
        "sph_server_submit": (i32, [H, pp, u32, u32, p, p, u32, p]),
        "sph_server_tick": (i32, [H, ctypes.POINTER(TickInfo)]),
        "sph_server_status": (i32, [H, u64, p, p, p]),
        "sph_server_fetch": (i32, [H, u64, u32, p, p, p]),
        "sph_server_cancel": (i32, [H, u64]),
        "sph_server_get_stats": (i32, [H, ctypes.POINTER(ServerStats)]),
        "sph_server_tick_log": (i32, [H, u64, u64, p, p, p, p]),
        "sph_batch_create": (i32, [u32, p, p, p, ctypes.POINTER(H)]),
        "sph_batch_destroy": (i32, [H]),
        "sph_batch_upload": (i32, [H, u32, p, p]),
        "sph_batch_set_params": (i32, [H, u32, pp]),
        "sph_batch_step": (i32, [H, f, i32]),
        "sph_batch_step_async": (i32, [H, f, i32]),
        "sph_batch_sync": (i32, [H]),
        "sph_batch_download": (i32, [H, u32, p, p, p]),
        "sph_batch_autotune": (i32, [H, f, i32, p]),
        "sph_batch_profile_begin": (i32, [H, u32]),
        "sph_batch_profile_end": (i32, [H, p, p]),
        "sph_batch_get_variants": (i32, [H, p]),
        "sph_batch_set_variants": (i32, [H, p]),
        "sph_copy_calibration": (i32, [u64, i32, p]),
        "sph_hash_cell": (u32, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
        "sph_bucket_of": (u32, [u32, u32]),
        "sph_debug_quotients": (i32, [p, p, p, u64]),
        "sph_debug_reciprocal_mismatches": (i32, [u32, u32, p, p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error():
    return lib().sph_last_error().decode()


def _check(rc):
    if rc != SPH_OK:
        raise SphError(rc, last_error())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def default_params(**kw):
    p = Params()
    lib().sph_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "bounds":
            p.bounds[:] = [float(x) for x in v]
        else:
            setattr(p, k, v)
    return p


def device_count():
    return lib().sph_device_count()


def generate(kind, seed, n, box=(1.0, 1.0, 1.0), spacing=0.1, jitter=0.0, speed=0.0):
    """Deterministic synthetic scene: (pos4, vel4) float32 arrays of shape (n, 4)."""
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    b = np.asarray(box, np.float32)
    _check(lib().sph_generate(kind, seed, n, _ptr(b), spacing, jitter, speed, _ptr(pos), _ptr(vel)))
    return pos, vel


def copy_calibration(nbytes, reps=3):
    """Runs a plain device copy kernel of nbytes; returns ms per copy."""
    ms = np.zeros(1, np.float64)
    _check(lib().sph_copy_calibration(int(nbytes), int(reps), _ptr(ms)))
    return float(ms[0])


def hash_cell(cx, cy, cz):
    """HashCell (physicsWorld.cc:505-511) as the kernels compute it."""
    return lib().sph_hash_cell(int(cx), int(cy), int(cz))


def debug_reciprocal_mismatches(lo_bits, hi_bits):
    """(count, first) of floats in the bit range whose GPU reciprocal differs from IEEE 1/x."""
    cnt = np.zeros(1, np.uint64)
    first = np.zeros(1, np.uint32)
    _check(lib().sph_debug_reciprocal_mismatches(int(lo_bits), int(hi_bits), _ptr(cnt), _ptr(first)))
    return int(cnt[0]), int(first[0])


def debug_quotients(a, b):
    """a / b as the pressure kernel computes it (host build of the same code)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.empty_like(a)
    _check(lib().sph_debug_quotients(_ptr(a), _ptr(b), _ptr(out), a.size))
    return out


def bucket_of(hash_, table):
    """GetKeyFromHash (physicsWorld.cc:513-516) as the kernels compute it."""
    return lib().sph_bucket_of(hash_, table)


_KERNELS = ("density", "pressure", "viscosity")


def _variants_dict(v):
    d = {k: {"work": bool(v[4 * q]), "list": bool(v[4 * q + 1]), "block": int(v[4 * q + 2]),
             "rows": bool(v[4 * q + 3])} for q, k in enumerate(_KERNELS)}
    d["nbr"] = bool(v[12])
    d["fused"] = bool(v[13])
    return d


def _variants_array(d):
    v = np.zeros(14, np.int32)
    for q, k in enumerate(_KERNELS):
        v[4 * q:4 * q + 4] = [int(d[k]["work"]), int(d[k]["list"]), int(d[k]["block"]),
                              int(d[k].get("rows", False))]
    v[12] = int(d.get("nbr", False))
    v[13] = int(d.get("fused", False))
    return v


class Sim:
    """One scene on the GPU: the FluidSimulation of physicsWorld.h:28-155."""

    def __init__(self, params, n, table_size=TABLE_GRID):
        """table_size: TABLE_GRID (cell-grid key), TABLE_REFERENCE (hash % n) or
        T (hash % T); results are identical under every key."""
        self._h = ctypes.c_void_p()
        self.n = int(n)
        _check(lib().sph_create(ctypes.byref(params), self.n, table_size, ctypes.byref(self._h)))
        t = np.zeros(1, np.uint32)
        lib().sph_get_size(self._h, None, _ptr(t))
        self.table = int(t[0])

    def close(self):
        if self._h:
            lib().sph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = Params()
        _check(lib().sph_get_params(self._h, ctypes.byref(p)))
        return p

    @params.setter
    def params(self, p):
        _check(lib().sph_set_params(self._h, ctypes.byref(p)))

    def upload(self, pos4=None, vel4=None):
        pos = None if pos4 is None else np.ascontiguousarray(pos4, np.float32)
        vel = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        for a in (pos, vel):
            if a is not None and a.shape != (self.n, 4):
                raise ValueError(f"expected shape {(self.n, 4)}, got {a.shape}")
        _check(lib().sph_upload(self._h, _ptr(pos), _ptr(vel)))

    def initialize_data(self):
        _check(lib().sph_initialize_data(self._h))

    def initialize(self, kind, seed=0, box=(1.0, 1.0, 1.0), spacing=0.0, jitter=0.0, speed=0.0):
        """Initial state generated on the device (bytes of generate() for the
        same arguments), then the spatial lookup and densities."""
        b = np.asarray(box, np.float32)
        _check(lib().sph_initialize(self._h, int(kind), int(seed), _ptr(b), spacing, jitter, speed))

    def set_colors(self, stops=None, enable=True):
        """Render hand-off: speed colours written by every update."""
        st = None if stops is None else np.ascontiguousarray(stops, np.float32).reshape(16)
        _check(lib().sph_set_colors(self._h, _ptr(st), 1 if enable else 0))

    def download_colors(self):
        c = np.zeros((self.n, 4), np.float32)
        _check(lib().sph_download_colors(self._h, _ptr(c)))
        return c

    def render_buffers(self):
        """(positions, colors) device pointers (SSBO bindings 0 and 1); colors 0 if off."""
        pp_, cp_ = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().sph_render_buffers(self._h, ctypes.byref(pp_), ctypes.byref(cp_)))
        return pp_.value or 0, cp_.value or 0

    def update(self, dt):
        _check(lib().sph_update(self._h, dt))

    def step(self, dt, substeps=1):
        _check(lib().sph_step(self._h, dt, substeps))

    def step_async(self, dt, substeps=1):
        _check(lib().sph_step_async(self._h, dt, substeps))

    def sync(self):
        _check(lib().sph_sync(self._h))

    def download(self):
        pos = np.zeros((self.n, 4), np.float32)
        vel = np.zeros((self.n, 4), np.float32)
        dens = np.zeros((self.n, 2), np.float32)
        _check(lib().sph_download(self._h, _ptr(pos), _ptr(vel), _ptr(dens)))
        return pos, vel, dens

    def set_timing(self, on):
        _check(lib().sph_set_timing(self._h, 1 if on else 0))

    def stage_ms(self):
        ms = np.zeros(6, np.float64)
        _check(lib().sph_stage_ms(self._h, _ptr(ms)))
        return ms

    def profile_begin(self, max_updates):
        _check(lib().sph_profile_begin(self._h, int(max_updates)))

    def profile_end(self):
        ms = np.zeros(6, np.float64)
        u = np.zeros(1, np.uint32)
        _check(lib().sph_profile_end(self._h, _ptr(ms), _ptr(u)))
        return ms, int(u[0])

    def autotune(self, dt, updates=3):
        ms = np.zeros(3, np.float64)
        _check(lib().sph_autotune(self._h, dt, int(updates), _ptr(ms)))
        return ms

    @property
    def variants(self):
        v = np.zeros(14, np.int32)
        _check(lib().sph_get_variants(self._h, _ptr(v)))
        return _variants_dict(v)

    @variants.setter
    def variants(self, d):
        _check(lib().sph_set_variants(self._h, _ptr(_variants_array(d))))

    def debug_set_scan_spin(self, sleeps):
        """Bound of the count scan's look-back wait (64-clock sleeps); 0 makes
        every wait give up, forcing the SPH_ERR_DEVICE path."""
        _check(lib().sph_debug_set_scan_spin(self._h, int(sleeps)))

    def count_nonfinite(self):
        c = np.zeros(1, np.uint64)
        _check(lib().sph_count_nonfinite(self._h, _ptr(c)))
        return int(c[0])

    def key(self):
        """(grid, gdim, lo): the bucket key in use (grid False: hash % table)."""
        g = np.zeros(1, np.int32)
        gd = np.zeros(3, np.uint32)
        lo = np.zeros(3, np.int32)
        _check(lib().sph_get_key(self._h, _ptr(g), _ptr(gd), _ptr(lo)))
        return bool(g[0]), [int(x) for x in gd], [int(x) for x in lo]

    def debug_spatial(self):
        n, T = self.n, self.table
        out = {k: np.zeros(n, np.uint32) for k in ("hash_id", "key_id", "sorted_key", "sorted_idx")}
        out["start"] = np.zeros(T, np.uint32)
        out["end"] = np.zeros(T, np.uint32)
        _check(lib().sph_debug_spatial(self._h, _ptr(out["hash_id"]), _ptr(out["key_id"]),
                                       _ptr(out["sorted_key"]), _ptr(out["sorted_idx"]),
                                       _ptr(out["start"]), _ptr(out["end"])))
        return out

    def debug_predicted(self):
        xp = np.zeros((self.n, 4), np.float32)
        _check(lib().sph_debug_predicted(self._h, _ptr(xp)))
        return xp

    def debug_velocities(self):
        """(v_ext, v_press): the last update's velocities after its external
        force (cc:45) and after its pressure stage (cc:421), particle order."""
        ve = np.zeros((self.n, 4), np.float32)
        vp = np.zeros((self.n, 4), np.float32)
        _check(lib().sph_debug_velocities(self._h, _ptr(ve), _ptr(vp)))
        return ve, vp

    def debug_neighbours(self):
        offs = np.zeros(self.n + 1, np.uint64)
        total = np.zeros(1, np.uint64)
        _check(lib().sph_debug_neighbours(self._h, _ptr(offs), None, 0, _ptr(total)))
        lst = np.zeros(max(int(total[0]), 1), np.uint32)
        _check(lib().sph_debug_neighbours(self._h, _ptr(offs), _ptr(lst), lst.size, _ptr(total)))
        return offs, lst[: int(total[0])]


class Batch:
    """Many independent scenes stepped by one launch per kernel (sph_batch_*)."""

    def __init__(self, params_list, counts, table_sizes=None):
        self._h = ctypes.c_void_p()
        self.counts = [int(c) for c in counts]
        n = len(self.counts)
        if len(params_list) != n:
            raise ValueError("one Params per scene")
        arr = (Params * n)(*params_list)
        cnt = np.asarray(self.counts, np.uint32)
        tab = None if table_sizes is None else np.asarray(table_sizes, np.uint32)
        _check(lib().sph_batch_create(n, _ptr(cnt), _ptr(tab), ctypes.cast(arr, ctypes.c_void_p),
                                      ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().sph_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene, pos4=None, vel4=None):
        n = self.counts[scene]
        pos = None if pos4 is None else np.ascontiguousarray(pos4, np.float32)
        vel = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        for a in (pos, vel):
            if a is not None and a.shape != (n, 4):
                raise ValueError(f"expected shape {(n, 4)}, got {a.shape}")
        _check(lib().sph_batch_upload(self._h, scene, _ptr(pos), _ptr(vel)))

    def initialize(self, scene, kind, seed=0, box=(1.0, 1.0, 1.0), spacing=0.0, jitter=0.0,
                   speed=0.0):
        b = np.asarray(box, np.float32)
        _check(lib().sph_batch_initialize(self._h, scene, int(kind), int(seed), _ptr(b), spacing,
                                          jitter, speed))

    def step(self, dt, substeps=1):
        _check(lib().sph_batch_step(self._h, dt, substeps))

    def step_async(self, dt, substeps=1):
        _check(lib().sph_batch_step_async(self._h, dt, substeps))

    def sync(self):
        _check(lib().sph_batch_sync(self._h))

    def download(self, scene):
        n = self.counts[scene]
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        dens = np.zeros((n, 2), np.float32)
        _check(lib().sph_batch_download(self._h, scene, _ptr(pos), _ptr(vel), _ptr(dens)))
        return pos, vel, dens

    def autotune(self, dt, updates=3):
        ms = np.zeros(3, np.float64)
        _check(lib().sph_batch_autotune(self._h, dt, int(updates), _ptr(ms)))
        return ms

    def profile_begin(self, max_updates):
        _check(lib().sph_batch_profile_begin(self._h, int(max_updates)))

    def profile_end(self):
        ms = np.zeros(6, np.float64)
        u = np.zeros(1, np.uint32)
        _check(lib().sph_batch_profile_end(self._h, _ptr(ms), _ptr(u)))
        return ms, int(u[0])

    @property
    def variants(self):
        v = np.zeros(14, np.int32)
        _check(lib().sph_batch_get_variants(self._h, _ptr(v)))
        return _variants_dict(v)

    @variants.setter
    def variants(self, d):
        _check(lib().sph_batch_set_variants(self._h, _ptr(_variants_array(d))))


def request_bytes(params, n, table=0):
    """Device bytes a server request holds while running (0 for invalid settings)."""
    return int(lib().sph_server_request_bytes(ctypes.byref(params), int(n), int(table)))


class Server:
    """Continuous batching of bake requests (sph_server_*)."""

    def __init__(self, **cfg):
        c = ServerConfig()
        lib().sph_server_default_config(ctypes.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.config = c
        self._h = ctypes.c_void_p()
        _check(lib().sph_server_create(ctypes.byref(c), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().sph_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_generated(self, params, n, steps, kind, seed=0, box=(1.0, 1.0, 1.0), spacing=0.0,
                         jitter=0.0, speed=0.0, table=0):
        """A request whose initial state is generated on the device at admission."""
        b = np.asarray(box, np.float32)
        rid = np.zeros(1, np.uint64)
        _check(lib().sph_server_submit_generated(self._h, ctypes.byref(params), int(n), int(table),
                                                 int(kind), int(seed), _ptr(b), spacing, jitter,
                                                 speed, int(steps), _ptr(rid)))
        return int(rid[0])

    def submit(self, params, pos4, vel4, steps, table=0):
        pos = np.ascontiguousarray(pos4, np.float32)
        vel = np.ascontiguousarray(vel4, np.float32)
        if pos.ndim != 2 or pos.shape[1] != 4 or pos.shape[0] == 0 or vel.shape != pos.shape:
            raise ValueError(f"pos and vel must both have shape (n, 4), got {pos.shape} and {vel.shape}")
        rid = np.zeros(1, np.uint64)
        _check(lib().sph_server_submit(self._h, ctypes.byref(params), pos.shape[0], int(table),
                                       _ptr(pos), _ptr(vel), int(steps), _ptr(rid)))
        return int(rid[0])

    def tick(self):
        t = TickInfo()
        _check(lib().sph_server_tick(self._h, ctypes.byref(t)))
        return _struct_dict(t)

    def status(self, rid):
        st = np.zeros(1, np.int32)
        done = np.zeros(1, np.uint64)
        lat = np.zeros(1, np.float64)
        _check(lib().sph_server_status(self._h, int(rid), _ptr(st), _ptr(done), _ptr(lat)))
        return int(st[0]), int(done[0]), float(lat[0])

    def fetch(self, rid, n):
        """Results of a finished request of n particles (the library checks n)."""
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        dens = np.zeros((n, 2), np.float32)
        _check(lib().sph_server_fetch(self._h, int(rid), int(n), _ptr(pos), _ptr(vel), _ptr(dens)))
        return pos, vel, dens

    def cancel(self, rid):
        _check(lib().sph_server_cancel(self._h, int(rid)))

    def stats(self):
        st = ServerStats()
        _check(lib().sph_server_get_stats(self._h, ctypes.byref(st)))
        return _struct_dict(st)

    def tick_log(self):
        """(ms, particles, layout) arrays of every timed tick, in the order
        they ran (stats() finishes timing the enqueued ones first)."""
        cnt = ctypes.c_uint64(0)
        total = self.stats()["ticks"]
        ms = np.zeros(total, np.float32)
        parts = np.zeros(total, np.uint64)
        lay = np.zeros(total, np.uint32)
        if total:
            _check(lib().sph_server_tick_log(self._h, 0, total, ms.ctypes.data, parts.ctypes.data,
                                             lay.ctypes.data, ctypes.byref(cnt)))
        n = int(cnt.value)
        return ms[:n], parts[:n], lay[:n]
