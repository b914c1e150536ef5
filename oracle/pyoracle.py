"""ctypes loader for the CPU oracle (test infrastructure only).

Only tests/, the smoke check and the bench's CPU-baseline leg import this
module.  It exposes:

* ``RefSim`` - the oracle's restatement of the reference's scatter /
  Gauss-Seidel step (src/hash.c, src/fluid.c of AdamSimpson/SPH);
* ``GatherSim`` - the same step in gather / Jacobi form, the definition the
  device kernels are checked against;
* ``CompiledRef`` - the reference's own step functions compiled from its
  sources into oracle/_ref (present only where it was built).

Particle state crosses the boundary as numpy structured arrays with the
reference's ``fluid_particle`` layout (52 bytes, src/fluid.h:56-70) for f32,
and the same fields in float64 for the f64 build.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PH_GRAVITY, PH_VISCOSITY, PH_PREDICT, PH_HASH_DENSITY, PH_RELAX, PH_UPDATE_VELOCITIES, PH_HASH = range(7)
PHASES = ("gravity", "viscosity", "predict", "hash_density", "relax", "update_velocities", "hash")

_FIELDS = ("x_prev", "y_prev", "x", "y", "v_x", "v_y", "a_x", "a_y",
           "density", "density_near", "pressure", "pressure_near")


def particle_dtype(real=np.float32):
    """numpy dtype of one particle record (fluid_particle for float32)."""
    fields = [(f, real) for f in _FIELDS] + [("id", np.int32)]
    if real == np.float64:
        return np.dtype(fields, align=True)  # 12*8 + 4 -> padded to 104
    return np.dtype(fields)  # 52 bytes, packed like the C struct


P32 = particle_dtype(np.float32)
P64 = particle_dtype(np.float64)
assert P32.itemsize == 52


class Tunables(ctypes.Structure):
    """tunable_parameters, src/fluid.h:78-97."""
    _fields_ = [(n, ctypes.c_float) for n in (
        "rest_density", "smoothing_radius", "g", "k", "k_near", "k_spring", "sigma", "beta",
        "time_step", "node_start_x", "node_end_x", "mover_center_x", "mover_center_y",
        "mover_width", "mover_height")] + [
        ("mover_type", ctypes.c_char), ("kill_sim", ctypes.c_char), ("active", ctypes.c_char)]


class AABB(ctypes.Structure):
    """AABB_t, src/geometry.h:33-40."""
    _fields_ = [(n, ctypes.c_float) for n in ("min_x", "max_x", "min_y", "max_y", "min_z", "max_z")]


assert ctypes.sizeof(Tunables) == 64 and ctypes.sizeof(AABB) == 24


def _tun(t) -> Tunables:
    """Accept any 64-byte tunable_parameters ctypes record."""
    return t if isinstance(t, Tunables) else Tunables.from_buffer_copy(bytes(t))


def _box(b) -> AABB:
    return b if isinstance(b, AABB) else AABB.from_buffer_copy(bytes(b))


def make_tunables(**kw) -> Tunables:
    t = Tunables()
    for k, v in kw.items():
        if k == "mover_type":
            t.mover_type = bytes([int(v)])
        elif k in ("kill_sim", "active"):
            setattr(t, k, bytes([int(v)]))
        else:
            setattr(t, k, float(v))
    return t


def make_box(max_x, max_y) -> AABB:
    return AABB(0.0, float(max_x), 0.0, float(max_y), 0.0, 0.0)


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (run `make -C oracle`)")
    return ctypes.CDLL(path)


_libs = {}


def lib(real="f32"):
    if real not in _libs:
        name = {"f32": "liboracle_f32.so", "f64": "liboracle_f64.so",
                "fast_f32": "liboracle_fast_f32.so"}[real]
        L = _load(name)
        sfx = "_f64" if real == "f64" else "_f32"
        vp, ip, up = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint
        for fn, res, args in (
            ("ref_create", vp, [ctypes.POINTER(AABB), ctypes.POINTER(Tunables), ip, up, up]),
            ("ref_destroy", None, [vp]),
            ("ref_set_tunables", None, [vp, ctypes.POINTER(Tunables)]),
            ("ref_phase", ip, [vp, vp, ip]),
            ("ref_step", None, [vp, vp, ip]),
            ("ref_counters", None, [vp, vp]),
            ("ref_lists", None, [vp, vp, vp, ip]),
            ("gather_create", vp, [ctypes.POINTER(AABB), ctypes.POINTER(Tunables), ip]),
            ("gather_destroy", None, [vp]),
            ("gather_set_tunables", None, [vp, ctypes.POINTER(Tunables)]),
            ("gather_set_cap", None, [vp, up]),
            ("gather_set_state", None, [vp, vp]),
            ("gather_get_state", None, [vp, vp]),
            ("gather_phase", ip, [vp, ip]),
            ("gather_step", None, [vp, ip]),
            ("gather_cells", None, [vp, vp, vp, vp]),
            ("gather_counters", None, [vp, vp]),
            ("gather_neighbor_sets", None, [vp, vp, vp, ip]),
        ):
            f = getattr(L, "oracle_" + fn + sfx)
            f.restype = res
            f.argtypes = args
            setattr(L, fn, f)
        _libs[real] = L
    return _libs[real]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def grid_size(box: AABB, h: float):
    """src/fluid.c:214-215 with float arithmetic."""
    sx = int(np.ceil(np.float32(box.max_x - box.min_x) / np.float32(h)))
    sy = int(np.ceil(np.float32(box.max_y - box.min_y) / np.float32(h)))
    return sx, sy


class RefSim:
    """Reference semantics (scatter, reverse order, in place), restated."""

    def __init__(self, box, tun, n, real="f32", max_bucket=100, max_neighbors=400):
        self.L = lib(real)
        self.dtype = P64 if real == "f64" else P32
        self.n = n
        self.h = self.L.ref_create(ctypes.byref(_box(box)), ctypes.byref(_tun(tun)), n, max_bucket, max_neighbors)
        if not self.h:
            raise MemoryError("oracle ref_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def set_tunables(self, tun):
        self.L.ref_set_tunables(self.h, ctypes.byref(_tun(tun)))

    def phase(self, parts, ph):
        assert parts.dtype == self.dtype and parts.flags.c_contiguous
        self.L.ref_phase(self.h, _ptr(parts), ph)

    def step(self, parts, nsteps=1):
        assert parts.dtype == self.dtype and parts.flags.c_contiguous
        self.L.ref_step(self.h, _ptr(parts), nsteps)

    def counters(self):
        out = np.zeros(3, np.uint64)
        self.L.ref_counters(self.h, _ptr(out))
        return dict(bucket_overflow=int(out[0]), neighbor_overflow=int(out[1]), key_out_of_range=int(out[2]))

    def lists(self, stride=400):
        counts = np.zeros(self.n, np.int32)
        lists = np.zeros((self.n, stride), np.int32)
        self.L.ref_lists(self.h, _ptr(counts), _ptr(lists), stride)
        return counts, lists


class GatherSim:
    """Gather / Jacobi form: the definition the device is checked against."""

    def __init__(self, box, tun, n, real="f32"):
        self.L = lib(real)
        self.dtype = P64 if real == "f64" else P32
        self.real = np.float64 if real == "f64" else np.float32
        self.n = n
        self.sx, self.sy = grid_size(_box(box), tun.smoothing_radius)
        self.h = self.L.gather_create(ctypes.byref(_box(box)), ctypes.byref(_tun(tun)), n)
        if not self.h:
            raise MemoryError("oracle gather_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.gather_destroy(self.h)
            self.h = None

    def set_tunables(self, tun):
        self.L.gather_set_tunables(self.h, ctypes.byref(_tun(tun)))

    def set_cap(self, cap):
        """The reference's bucket cap (0: off): particles ranked cap or more
        by id in their cell are in no bucket of that hash."""
        self.L.gather_set_cap(self.h, int(cap))

    def set_state(self, parts):
        p = np.ascontiguousarray(parts.astype(self.dtype) if parts.dtype != self.dtype else parts)
        self.L.gather_set_state(self.h, _ptr(p))

    def get_state(self):
        out = np.zeros(self.n, self.dtype)
        self.L.gather_get_state(self.h, _ptr(out))
        return out

    def phase(self, ph):
        self.L.gather_phase(self.h, ph)

    def step(self, nsteps=1):
        self.L.gather_step(self.h, nsteps)

    def cells(self):
        keys = np.zeros(self.n, np.uint32)
        order = np.zeros(self.n, np.int32)
        start = np.zeros(self.sx * self.sy + 1, np.uint32)
        self.L.gather_cells(self.h, _ptr(keys), _ptr(order), _ptr(start))
        return keys, order, start

    def counters(self):
        out = np.zeros(6, np.uint64)
        self.L.gather_counters(self.h, _ptr(out))
        return dict(bucket_overflow=int(out[0]), neighbor_overflow=int(out[1]),
                    key_out_of_range=int(out[2]), max_cell_count=int(out[3]),
                    max_neighbors=int(out[4]), impulse_clamps=int(out[5]))

    def neighbor_sets(self, stride=128):
        counts = np.zeros(self.n, np.int32)
        lists = np.full((self.n, stride), -1, np.int32)
        self.L.gather_neighbor_sets(self.h, _ptr(counts), _ptr(lists), stride)
        return counts, lists


def compiled_ref_path(fast=False):
    return os.path.join(HERE, "_ref", "libtinysph_ref_fast.so" if fast else "libtinysph_ref.so")


def have_compiled_ref(fast=False):
    return os.path.exists(compiled_ref_path(fast))


class CompiledRef:
    """The reference's own step functions (oracle/_ref), same interface as RefSim."""

    _cache = {}

    def __init__(self, box, tun, n, max_bucket=100, fast=False):
        path = compiled_ref_path(fast)
        if path not in CompiledRef._cache:
            L = ctypes.CDLL(path)
            vp, ip = ctypes.c_void_p, ctypes.c_int
            L.refh_create.restype = vp
            L.refh_create.argtypes = [ctypes.POINTER(AABB), ctypes.POINTER(Tunables), ip, ctypes.c_uint]
            L.refh_destroy.argtypes = [vp]
            L.refh_set_tunables.argtypes = [vp, ctypes.POINTER(Tunables)]
            L.refh_load.argtypes = [vp, vp]
            L.refh_load.restype = ip
            L.refh_store.argtypes = [vp, vp]
            L.refh_phase.argtypes = [vp, ip]
            L.refh_phase.restype = ip
            L.refh_step.argtypes = [vp, ip]
            L.refh_lists.argtypes = [vp, vp, vp, ip]
            L.refh_bucket_counts.argtypes = [vp, vp]
            CompiledRef._cache[path] = L
        self.L = CompiledRef._cache[path]
        self.n = n
        self.sx, self.sy = grid_size(_box(box), tun.smoothing_radius)
        self.h = self.L.refh_create(ctypes.byref(_box(box)), ctypes.byref(_tun(tun)), n, max_bucket)
        if not self.h:
            raise MemoryError("refh_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.refh_destroy(self.h)
            self.h = None

    def set_tunables(self, tun):
        self.L.refh_set_tunables(self.h, ctypes.byref(_tun(tun)))

    def load(self, parts):
        p = np.ascontiguousarray(parts, dtype=P32)
        if self.L.refh_load(self.h, _ptr(p)) != 0:
            raise ValueError("particle i must carry id i")

    def store(self):
        out = np.zeros(self.n, P32)
        self.L.refh_store(self.h, _ptr(out))
        return out

    def phase(self, ph):
        self.L.refh_phase(self.h, ph)

    def step(self, nsteps=1):
        self.L.refh_step(self.h, nsteps)

    def lists(self, stride=400):
        counts = np.zeros(self.n, np.int32)
        lists = np.zeros((self.n, stride), np.int32)
        self.L.refh_lists(self.h, _ptr(counts), _ptr(lists), stride)
        return counts, lists

    def bucket_counts(self):
        out = np.zeros(self.sx * self.sy, np.uint32)
        self.L.refh_bucket_counts(self.h, _ptr(out))
        return out
