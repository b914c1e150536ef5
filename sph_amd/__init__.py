"""sph_amd - Python host binding of libsph_hip.so, the MI355X (gfx950)
implementation of TinySPH's step (AdamSimpson/SPH).

The library's C ABI (include/sph_abi.h) is the product boundary; this module
is a thin ctypes mirror of it for tests, the bench and Python callers.  It
loads the in-tree ``sph_amd/libsph_hip.so`` and raises if that library is
missing: there is no CPU fallback.

Records keep the reference's layouts: particles are numpy structured arrays
of ``PARTICLE`` (= fluid_particle, src/fluid.h:56-70, 52 bytes), tunables are
``Tunables`` (src/fluid.h:78-97), boxes ``AABB`` (src/geometry.h:33-40).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = ["PARTICLE", "Tunables", "Param", "AABB", "SceneDesc", "Counters", "Solver", "SphError",
           "Surrogate", "Trajectory", "mfma_selftest", "make_tunables", "make_box", "scene_lattice", "lib",
           "LIB_PATH"]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


TUNED = os.path.join(PKG_DIR, "configs", "tuned.json")


def tuned_library(tuned=None):
    """Path of the library variant that configs/tuned.json (written by
    sph_amd.autotune) records for its default workload, or None."""
    import json
    tuned = tuned or TUNED
    if not os.path.exists(tuned):
        return None
    with open(tuned) as f:
        cfg = json.load(f)
    best = cfg.get("workloads", {}).get(cfg.get("default", ""), {}).get("best")
    path = os.path.join(PKG_DIR, "variants", best) if best else None
    return path if path and os.path.exists(path) else None


def _lib_path():
    """SPH_AMD_LIB if set; else the tuned variant; else the default build."""
    return (os.environ.get("SPH_AMD_LIB") or tuned_library()
            or os.path.join(PKG_DIR, "libsph_hip.so"))


LIB_PATH = _lib_path()

_FIELDS = ("x_prev", "y_prev", "x", "y", "v_x", "v_y", "a_x", "a_y",
           "density", "density_near", "pressure", "pressure_near")
PARTICLE = np.dtype([(f, np.float32) for f in _FIELDS] + [("id", np.int32)])
assert PARTICLE.itemsize == 52

REF_MAX_BUCKET = 100  # the reference's bucket cap (src/fluid.c:174)

SPH_FIELD = dict(x=0, y=1, v_x=2, v_y=3, x_prev=4, y_prev=5, density=6, density_near=7,
                 id=8, key=9)


class Tunables(ctypes.Structure):
    """tunable_parameters (src/fluid.h:78-97)."""
    _fields_ = [(n, ctypes.c_float) for n in (
        "rest_density", "smoothing_radius", "g", "k", "k_near", "k_spring", "sigma", "beta",
        "time_step", "node_start_x", "node_end_x", "mover_center_x", "mover_center_y",
        "mover_width", "mover_height")] + [
        ("mover_type", ctypes.c_char), ("kill_sim", ctypes.c_char), ("active", ctypes.c_char)]


class Param(ctypes.Structure):
    """param (src/fluid.h:100-106)."""
    _fields_ = [("tunable_params", Tunables), ("number_fluid_particles_global", ctypes.c_int),
                ("number_fluid_particles_local", ctypes.c_int),
                ("max_fluid_particle_index", ctypes.c_int),
                ("number_halo_particles", ctypes.c_int)]


class AABB(ctypes.Structure):
    """AABB_t (src/geometry.h:33-40)."""
    _fields_ = [(n, ctypes.c_float) for n in ("min_x", "max_x", "min_y", "max_y", "min_z", "max_z")]


class SceneDesc(ctypes.Structure):
    """sph_scene_desc: one scene of a batched context."""
    _fields_ = [("tunables", Tunables), ("boundary", AABB), ("capacity", ctypes.c_int)]


class Counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "steps", "hashes", "bucket_overflow", "neighbor_overflow", "key_out_of_range",
        "nonfinite", "max_cell_count", "max_neighbors")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


assert ctypes.sizeof(Tunables) == 64 and ctypes.sizeof(Param) == 80 and ctypes.sizeof(AABB) == 24


class SphError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libsph_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SphError(f"{LIB_PATH} is missing: build it with `python -m sph_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p
        sigs = {
            "sph_create": (ip, [ctypes.POINTER(vp), ctypes.POINTER(Param), ctypes.POINTER(AABB), ip]),
            "sph_destroy": (None, [vp]),
            "sph_create_batch": (ip, [ctypes.POINTER(vp), ctypes.POINTER(SceneDesc), ip]),
            "sph_num_scenes": (ip, [vp]),
            "sph_scene_info": (ip, [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(ctypes.c_uint),
                                    ctypes.POINTER(ctypes.c_uint)]),
            "sph_upload_batch": (ip, [vp, vp, vp]),
            "sph_set_scene_tunables": (ip, [vp, ip, ctypes.POINTER(Tunables)]),
            "sph_set_scene_forcing": (ip, [vp, ip, ctypes.c_float, ctypes.c_float]),
            "sph_set_scene_frozen": (ip, [vp, vp]),
            "sph_set_bucket_cap": (ip, [vp, ip]),
            "sph_set_gather_lanes": (ip, [vp, ip]),
            "sph_upload": (ip, [vp, vp, ip]),
            "sph_set_tunables": (ip, [vp, ctypes.POINTER(Tunables)]),
            "sph_step": (ip, [vp, ip]),
            "sph_step_eager": (ip, [vp, ip]),
            "sph_step_tunables2": (ip, [vp, ctypes.POINTER(Tunables), ctypes.POINTER(Tunables)]),
            "sph_download_positions": (ip, [vp, vp]),
            "sph_shim_set_sync_interval": (ip, [ip]),
            "sph_shim_set_sync_fields": (ip, [ip]),
            "sph_shim_sync": (ip, []),
            "sph_shim_mark_host_dirty": (ip, []),
            "sph_shim_release": (None, []),
            "sph_shim_context": (vp, []),
            "sph_sync": (ip, [vp]),
            "sph_download": (ip, [vp, vp]),
            "sph_pack_coords_i16": (ip, [vp, vp]),
            "sph_get_counters": (ip, [vp, ctypes.POINTER(Counters)]),
            "sph_reset_counters": (ip, [vp]),
            "sph_num_particles": (ip, [vp]),
            "sph_last_error": (cp, []),
            "sph_phase_gravity": (ip, [vp]),
            "sph_phase_viscosity": (ip, [vp]),
            "sph_phase_predict": (ip, [vp]),
            "sph_phase_hash": (ip, [vp, ip]),
            "sph_phase_relax": (ip, [vp]),
            "sph_phase_update_velocities": (ip, [vp]),
            "sph_read_sorted": (ip, [vp, ip, vp]),
            "sph_read_cells": (ip, [vp, vp, vp]),
            "sph_read_neighbors": (ip, [vp, ip, vp, vp, vp]),
            "sph_num_cells": (ip, [vp, ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint)]),
            "sph_time_kernels": (ip, [vp, ip, ip, ctypes.POINTER(cp), ctypes.POINTER(ctypes.c_float),
                                      ctypes.POINTER(ip)]),
            "sph_scene_lattice": (ip, [vp, ip, ip, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_uint64]),
            "sph_set_scene_time_offset": (ip, [vp, ip, ctypes.c_uint]),
            "sph_snapshot_save": (ip, [vp]),
            "sph_snapshot_restore": (ip, [vp, vp]),
            "sph_gather_state_dev": (ip, [vp, vp]),
            "sph_load_state_dev": (ip, [vp, vp]),
            "sph_scene_density": (ip, [vp, vp, vp]),
            "sph_scene_density_noalloc": (ip, [vp, vp, vp]),
            "sph_surrogate_last_error": (cp, []),
            "sph_surrogate_num_params": (ip, []),
            "sph_surrogate_create": (ip, [ctypes.POINTER(vp), ctypes.c_uint64]),
            "sph_surrogate_destroy": (None, [vp]),
            "sph_surrogate_get_params": (ip, [vp, vp]),
            "sph_surrogate_set_params": (ip, [vp, vp]),
            "sph_surrogate_get_grads": (ip, [vp, vp]),
            "sph_surrogate_zero_grads": (ip, [vp]),
            "sph_surrogate_predict": (ip, [vp, vp, vp]),
            "sph_surrogate_backward": (ip, [vp, vp, vp, ctypes.POINTER(ctypes.c_float)]),
            "sph_surrogate_adam": (ip, [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float]),
            "sph_surrogate_apply": (ip, [vp, vp]),
            "sph_surrogate_macro_step": (ip, [vp, vp, ctypes.c_float, vp, ctypes.POINTER(ip)]),
            "sph_surrogate_mfma_selftest": (ip, [vp]),
            "sph_surrogate_set_shadow": (ip, [vp, ip]),
            "sph_surrogate_shadow_stats": (ip, [vp, vp]),
            "sph_surrogate_backward_trajectory": (ip, [vp, vp, vp, ctypes.POINTER(ctypes.c_float)]),
            "sph_trajectory_create": (ip, [ctypes.POINTER(vp), ip, ip, ip]),
            "sph_selftest_sqrt": (ip, [ctypes.c_uint, ctypes.c_uint,
                                       ctypes.POINTER(ctypes.c_uint64)]),
            "sph_selftest_rcp": (ip, [ctypes.c_uint, ctypes.c_uint,
                                       ctypes.POINTER(ctypes.c_uint64)]),
            "sph_trajectory_destroy": (None, [vp]),
            "sph_trajectory_record": (ip, [vp, vp, ip]),
            "sph_trajectory_read": (ip, [vp, ip, vp]),
            "sph_trajectory_load": (ip, [vp, vp, ip, vp, vp]),
        }
        lenient = bool(os.environ.get("SPH_AMD_LIB"))  # another build may predate some calls
        for name, (res, args) in sigs.items():
            if lenient and not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().sph_last_error().decode(errors="replace")
        raise SphError(f"{what} failed ({rc}): {msg}")


def _check_m(rc, what):
    if rc != 0:
        msg = lib().sph_surrogate_last_error().decode(errors="replace")
        raise SphError(f"{what} failed ({rc}): {msg}")


def make_tunables(**kw) -> Tunables:
    t = Tunables()
    for k, v in kw.items():
        if k in ("mover_type", "kill_sim", "active"):
            setattr(t, k, bytes([int(v)]))
        else:
            setattr(t, k, float(v))
    return t


def make_box(max_x, max_y) -> AABB:
    return AABB(0.0, float(max_x), 0.0, float(max_y), 0.0, 0.0)


def scene_lattice(nx, ny, spacing, jitter=0.0, seed=0, x0=0.0, y0=0.0) -> np.ndarray:
    """The reference's lattice (src/geometry.c:29-67) + seeded jitter."""
    out = np.zeros(nx * ny, PARTICLE)
    _check(lib().sph_scene_lattice(out.ctypes.data, nx, ny, x0, y0, spacing, jitter, seed),
           "sph_scene_lattice")
    return out


class Solver:
    """One device-resident scene (sph_create ... sph_destroy)."""

    def __init__(self, tunables: Tunables, box: AABB, capacity: int):
        self._h = ctypes.c_void_p()
        p = Param()
        p.tunable_params = tunables
        p.number_fluid_particles_global = capacity
        p.number_fluid_particles_local = capacity
        p.max_fluid_particle_index = capacity - 1
        self.box = box
        self.tunables = tunables
        _check(lib().sph_create(ctypes.byref(self._h), ctypes.byref(p), ctypes.byref(box), capacity),
               "sph_create")

    @classmethod
    def batch(cls, scenes):
        """A batched context: scenes = [(tunables, box, capacity), ...]."""
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p()
        arr = (SceneDesc * len(scenes))()
        for i, (t, b, cap) in enumerate(scenes):
            arr[i].tunables = t
            arr[i].boundary = b
            arr[i].capacity = int(cap)
        self.box = None
        self.tunables = None
        _check(lib().sph_create_batch(ctypes.byref(self._h), arr, len(scenes)), "sph_create_batch")
        return self

    @property
    def nscenes(self):
        return lib().sph_num_scenes(self._h)

    def upload_batch(self, particles: np.ndarray, counts):
        p = np.ascontiguousarray(particles, dtype=PARTICLE)
        c = np.ascontiguousarray(counts, dtype=np.int32)
        _check(lib().sph_upload_batch(self._h, p.ctypes.data, c.ctypes.data), "sph_upload_batch")

    def set_scene_forcing(self, scene: int, amplitude: float, period: float):
        _check(lib().sph_set_scene_forcing(self._h, scene, amplitude, period), "sph_set_scene_forcing")

    def set_scene_time_offset(self, scene: int, steps: int):
        _check(lib().sph_set_scene_time_offset(self._h, scene, int(steps)),
               "sph_set_scene_time_offset")

    def snapshot_save(self):
        _check(lib().sph_snapshot_save(self._h), "sph_snapshot_save")

    def snapshot_restore(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        _check(lib().sph_snapshot_restore(self._h, f.ctypes.data), "sph_snapshot_restore")

    def scene_density(self):
        ns = self.nscenes
        mean = np.zeros(ns, np.float32)
        mx = np.zeros(ns, np.float32)
        _check_m(lib().sph_scene_density(self._h, mean.ctypes.data, mx.ctypes.data),
                 "sph_scene_density")
        return mean, mx

    def set_frozen(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        _check(lib().sph_set_scene_frozen(self._h, f.ctypes.data), "sph_set_scene_frozen")

    def set_bucket_cap(self, cap: int = REF_MAX_BUCKET):
        """The reference's per-cell bucket cap (src/hash.c:153-164); 0 turns
        it off (the default)."""
        _check(lib().sph_set_bucket_cap(self._h, int(cap)), "sph_set_bucket_cap")

    def set_gather_lanes(self, lanes: int = 0):
        """Lanes per particle of the fused step's gathers: 0 chooses by size
        (4 for small contexts), 1 or 4 force a shape; results are the same."""
        _check(lib().sph_set_gather_lanes(self._h, int(lanes)), "sph_set_gather_lanes")

    def set_scene_tunables(self, scene: int, t: Tunables):
        _check(lib().sph_set_scene_tunables(self._h, scene, ctypes.byref(t)), "sph_set_scene_tunables")

    def close(self):
        if self._h:
            lib().sph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        return lib().sph_num_particles(self._h)

    def upload(self, particles: np.ndarray):
        p = np.ascontiguousarray(particles, dtype=PARTICLE)
        _check(lib().sph_upload(self._h, p.ctypes.data, len(p)), "sph_upload")

    def set_tunables(self, t: Tunables):
        _check(lib().sph_set_tunables(self._h, ctypes.byref(t)), "sph_set_tunables")
        self.tunables = t

    def step(self, nsteps=1, graph=True):
        f = lib().sph_step if graph else lib().sph_step_eager
        _check(f(self._h, nsteps), "sph_step")

    def sync(self):
        _check(lib().sph_sync(self._h), "sph_sync")

    def step_tunables2(self, first: Tunables, second: Tunables):
        """One step with `first` for gravity / viscosity / predict and
        `second` for the rest (sph_step_tunables2)."""
        _check(lib().sph_step_tunables2(self._h, ctypes.byref(first), ctypes.byref(second)),
               "sph_step_tunables2")
        self.tunables = second

    def download_positions(self) -> np.ndarray:
        out = np.zeros((self.n, 2), np.float32)
        _check(lib().sph_download_positions(self._h, out.ctypes.data), "sph_download_positions")
        return out

    def download(self) -> np.ndarray:
        out = np.zeros(self.n, PARTICLE)
        _check(lib().sph_download(self._h, out.ctypes.data), "sph_download")
        return out

    def pack_coords(self) -> np.ndarray:
        out = np.zeros((self.n, 2), np.int16)
        _check(lib().sph_pack_coords_i16(self._h, out.ctypes.data), "sph_pack_coords_i16")
        return out

    def phase(self, name: str, **kw):
        L = lib()
        if name == "gravity":
            rc = L.sph_phase_gravity(self._h)
        elif name == "viscosity":
            rc = L.sph_phase_viscosity(self._h)
        elif name == "predict":
            rc = L.sph_phase_predict(self._h)
        elif name == "hash_density":
            rc = L.sph_phase_hash(self._h, 1)
        elif name == "hash":
            rc = L.sph_phase_hash(self._h, 0)
        elif name == "relax":
            rc = L.sph_phase_relax(self._h)
        elif name == "update_velocities":
            rc = L.sph_phase_update_velocities(self._h)
        else:
            raise ValueError(name)
        _check(rc, "sph_phase_" + name)

    def read_sorted(self, field: str) -> np.ndarray:
        dt = {"id": np.int32, "key": np.uint32}.get(field, np.float32)
        out = np.zeros(self.n, dt)
        _check(lib().sph_read_sorted(self._h, SPH_FIELD[field], out.ctypes.data), "sph_read_sorted")
        return out

    def cells(self):
        sx, sy = ctypes.c_uint(), ctypes.c_uint()
        _check(lib().sph_num_cells(self._h, ctypes.byref(sx), ctypes.byref(sy)), "sph_num_cells")
        nc = sx.value * sy.value
        start = np.zeros(nc, np.uint32)
        count = np.zeros(nc, np.uint32)
        _check(lib().sph_read_cells(self._h, start.ctypes.data, count.ctypes.data), "sph_read_cells")
        return start, count, (sx.value, sy.value)

    def neighbors(self, stride=64):
        """(full count, half-list length, first `stride` neighbour ids), by id."""
        n = self.n
        full = np.zeros(n, np.int32)
        half = np.zeros(n, np.int32)
        ids = np.full((n, stride), -1, np.int32)
        _check(lib().sph_read_neighbors(self._h, stride, full.ctypes.data, half.ctypes.data,
                                        ids.ctypes.data), "sph_read_neighbors")
        return full, half, ids

    def counters(self) -> dict:
        c = Counters()
        _check(lib().sph_get_counters(self._h, ctypes.byref(c)), "sph_get_counters")
        return c.as_dict()

    def reset_counters(self):
        _check(lib().sph_reset_counters(self._h), "sph_reset_counters")

    def time_kernels(self, nsteps: int) -> dict:
        K = 16
        names = (ctypes.c_char_p * K)()
        ms = (ctypes.c_float * K)()
        nk = ctypes.c_int()
        _check(lib().sph_time_kernels(self._h, nsteps, K, names, ms, ctypes.byref(nk)),
               "sph_time_kernels")
        return {names[i].decode(): float(ms[i]) for i in range(nk.value)}


class Surrogate:
    """The learned 4-step macro-step (sph_surrogate_*)."""

    def __init__(self, seed: int = 0):
        self._h = ctypes.c_void_p()
        _check_m(lib().sph_surrogate_create(ctypes.byref(self._h), seed), "sph_surrogate_create")
        self.nparams = lib().sph_surrogate_num_params()

    def close(self):
        if self._h:
            lib().sph_surrogate_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        _check_m(lib().sph_surrogate_get_params(self._h, out.ctypes.data), "get_params")
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check_m(lib().sph_surrogate_set_params(self._h, p.ctypes.data), "set_params")

    def grads(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        _check_m(lib().sph_surrogate_get_grads(self._h, out.ctypes.data), "get_grads")
        return out

    def zero_grads(self):
        _check_m(lib().sph_surrogate_zero_grads(self._h), "zero_grads")

    def predict(self, solver: Solver) -> np.ndarray:
        out = np.zeros((solver.n, 2), np.float32)
        _check_m(lib().sph_surrogate_predict(self._h, solver._h, out.ctypes.data), "predict")
        return out

    def backward(self, solver: Solver, target) -> float:
        """target: n x 2 float32 displacements (caller order)."""
        t = np.ascontiguousarray(target, dtype=np.float32)
        assert t.shape == (solver.n, 2)
        loss = ctypes.c_float()
        _check_m(lib().sph_surrogate_backward(self._h, solver._h, t.ctypes.data,
                                              ctypes.byref(loss)), "backward")
        return float(loss.value)

    def backward_trajectory(self, solver: Solver, traj: "Trajectory") -> float:
        loss = ctypes.c_float()
        _check_m(lib().sph_surrogate_backward_trajectory(self._h, solver._h, traj._h,
                                                         ctypes.byref(loss)), "backward")
        return float(loss.value)

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _check_m(lib().sph_surrogate_adam(self._h, lr, beta1, beta2, eps), "adam")

    def apply(self, solver: Solver):
        _check_m(lib().sph_surrogate_apply(self._h, solver._h), "apply")

    def set_shadow(self, every: int):
        """Shadow judging of one scene in every `every` per macro-step (0: off)."""
        _check_m(lib().sph_surrogate_set_shadow(self._h, int(every)), "set_shadow")

    def shadow_stats(self) -> dict:
        o = np.zeros(8, np.float64)
        _check_m(lib().sph_surrogate_shadow_stats(self._h, o.ctypes.data), "shadow_stats")
        acc = max(o[1], 1.0)
        return {"shadowed": int(o[0]), "accepted": int(o[1]), "accepted_over_tol": int(o[2]),
                "mean_abs_error": o[3] / acc, "max_abs_error": float(o[4]),
                "mean_signed_error": o[5] / acc, "mean_rms_error": o[6] / acc,
                "max_rms_error": float(o[7])}

    def macro_step(self, solver: Solver, tol=0.01):
        flags = np.zeros(solver.nscenes, np.int32)
        nf = ctypes.c_int()
        _check_m(lib().sph_surrogate_macro_step(self._h, solver._h, tol, flags.ctypes.data,
                                                ctypes.byref(nf)), "macro_step")
        return flags


class Trajectory:
    """Device-resident trajectory store (sph_trajectory_*)."""

    def __init__(self, nscenes: int, n: int, frames: int):
        self._h = ctypes.c_void_p()
        self.nscenes, self.n, self.frames = nscenes, n, frames
        _check_m(lib().sph_trajectory_create(ctypes.byref(self._h), nscenes, n, frames),
                 "sph_trajectory_create")

    def close(self):
        if self._h:
            lib().sph_trajectory_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def record(self, solver: Solver, frame: int):
        _check_m(lib().sph_trajectory_record(self._h, solver._h, frame), "sph_trajectory_record")

    def read(self, frame: int) -> np.ndarray:
        out = np.zeros((self.nscenes, self.n, 4), np.float32)
        _check_m(lib().sph_trajectory_read(self._h, frame, out.ctypes.data), "sph_trajectory_read")
        return out

    def load(self, solver: Solver, scenes, frames):
        s = np.ascontiguousarray(scenes, dtype=np.int32)
        f = np.ascontiguousarray(frames, dtype=np.int32)
        assert s.shape == f.shape
        _check_m(lib().sph_trajectory_load(self._h, solver._h, len(s), s.ctypes.data, f.ctypes.data),
                 "sph_trajectory_load")


def selftest_sqrt(lo_bits: int, hi_bits: int) -> int:
    """Mismatches of the kernels' square root against IEEE sqrtf over the
    float bit patterns [lo_bits, hi_bits)."""
    bad = ctypes.c_uint64()
    _check(lib().sph_selftest_sqrt(lo_bits, hi_bits, ctypes.byref(bad)), "sph_selftest_sqrt")
    return int(bad.value)


def selftest_rcp(lo_bits: int, hi_bits: int) -> int:
    """Mismatches of the kernels' reciprocal against IEEE 1.0f / x over the
    float bit patterns [lo_bits, hi_bits)."""
    bad = ctypes.c_uint64()
    _check(lib().sph_selftest_rcp(lo_bits, hi_bits, ctypes.byref(bad)), "sph_selftest_rcp")
    return int(bad.value)


def mfma_selftest() -> np.ndarray:
    out = np.zeros(256, np.float32)
    _check_m(lib().sph_surrogate_mfma_selftest(out.ctypes.data), "mfma_selftest")
    return out.reshape(16, 16)
