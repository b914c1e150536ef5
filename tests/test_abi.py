"""The C ABI boundary, checked without a GPU.

* libsph_hip.so loads and exports every function include/sph_abi.h declares;
* the header's record layouts compile (its static asserts) with gcc and
  match the Python mirror;
* the host-side helpers with reference names compute what the reference's
  own functions compute (compiled reference, where it was built);
* the scene builder is the documented SplitMix64 lattice;
* without a device, creating a context fails loudly with SPH_ENODEV.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import sph_amd as sa
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sph_abi.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    src = "\n".join(l for l in src.splitlines() if not l.strip().startswith("#"))
    src = re.sub(r"struct\s+\w+\s*\{.*?\};", "", src, flags=re.S)
    src = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", "", src, flags=re.S)
    names = []
    for stmt in src.split(";"):
        stmt = " ".join(stmt.split())
        if "(" not in stmt or stmt.startswith("typedef") or "SPH_STATIC_ASSERT" in stmt:
            continue
        stmt = stmt.replace('extern "C" {', "").strip()
        m = re.match(r"^[\w\s\*]*?\b(\w+)\s*\(", stmt)
        if m:
            names.append(m.group(1))
    return names


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_header_declares_reference_entry_points():
    names = set(declared_functions())
    ref = {"boundaryConditions", "initParticles", "calculate_density", "apply_gravity",
           "viscosity_impluses", "predict_positions", "double_density_relaxation",
           "updateVelocity", "updateVelocities", "checkVelocity", "identify_oob_particles",
           "hash_val", "hash_fluid", "hash_halo"}
    assert ref <= names, ref - names
    assert {"sph_create", "sph_upload", "sph_step", "sph_download", "sph_destroy"} <= names


def test_library_exports_every_declared_function():
    assert os.path.exists(sa.LIB_PATH), "build the library first"
    exported = exported_symbols(sa.LIB_PATH)
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    L = sa.lib()
    for n in declared_functions():
        getattr(L, n)


def test_every_exported_c_symbol_is_declared():
    """The reverse direction: every C-linkage entry point the library exports
    (sph_* and the reference-named ones) is declared in the header, so the
    header is the whole boundary."""
    exported = exported_symbols(sa.LIB_PATH)
    declared = set(declared_functions())
    c_symbols = {s for s in exported if not s.startswith("_") and "@" not in s}
    undeclared = sorted(c_symbols - declared)
    assert not undeclared, undeclared


def test_every_tuning_variant_exports_the_abi():
    """The autotuner's library variants (other launch shapes of the gather
    kernels) are drop-in replacements: same exported entry points."""
    from sph_amd import build as bd
    declared = declared_functions()
    for cap, run in bd.VARIANTS + bd.CHECK_VARIANTS:
        path = os.path.join(bd.VARIANT_DIR, bd.variant_name(cap, run))
        assert os.path.exists(path), f"{path}: build the variants first"
        exported = exported_symbols(path)
        assert not [n for n in declared if n not in exported], path


def test_tuned_library_choice(tmp_path):
    """sph_amd.lib() takes SPH_AMD_LIB, else the variant tuned.json records
    for the default workload, else the default build."""
    t = tmp_path / "tuned.json"
    assert sa.tuned_library(str(t)) is None
    t.write_text('{"default": "dam_break_1m", "workloads": '
                 '{"dam_break_1m": {"best": "libsph_hip_t2048_x16.so"}}}')
    assert sa.tuned_library(str(t)).endswith("variants/libsph_hip_t2048_x16.so")
    t.write_text('{"default": "dam_break_1m", "workloads": {"dam_break_1m": {"best": "no_such.so"}}}')
    assert sa.tuned_library(str(t)) is None


def test_header_layout_compiles_and_matches_python(tmp_path):
    c = tmp_path / "t.c"
    c.write_text('#include "sph_abi.h"\n#include <stdio.h>\nint main(void){printf("%zu %zu %zu %zu\\n",'
                 ' sizeof(fluid_particle), sizeof(tunable_parameters), sizeof(param), sizeof(AABB_t));'
                 ' return 0;}\n')
    exe = tmp_path / "t"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", f"-I{ROOT}/include", str(c), "-o",
                    str(exe)], check=True)
    sizes = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert sizes == ["52", "64", "80", "24"]
    assert sa.PARTICLE.itemsize == 52
    assert ctypes.sizeof(sa.Tunables) == 64 and ctypes.sizeof(sa.Param) == 80


def test_header_compiles_as_cxx(tmp_path):
    c = tmp_path / "t.cpp"
    c.write_text('#include "sph_abi.h"\nint main(){ return sizeof(fluid_particle) == 52 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT}/include", str(c), "-o",
                    str(tmp_path / "t")], check=True)


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & (2**64 - 1)
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
    return state, z ^ (z >> 31)


def test_scene_lattice_is_documented_splitmix64():
    p = sa.scene_lattice(5, 3, 0.5, 0.01, seed=42, x0=1.0, y0=2.0)
    s = 42
    f32 = np.float32
    for i in range(15):
        s, zx = splitmix64(s)
        s, zy = splitmix64(s)
        ux = f32(zx >> 40) * f32(2.0 ** -24)
        uy = f32(zy >> 40) * f32(2.0 ** -24)
        jx = f32(0.01) * (f32(2.0) * ux - f32(1.0))
        jy = f32(0.01) * (f32(2.0) * uy - f32(1.0))
        x = f32(1.0) + f32(i % 5) * f32(0.5)
        y = f32(2.0) + f32(i // 5) * f32(0.5)
        assert p["x"][i] == x + jx and p["y"][i] == y + jy, i
        assert p["id"][i] == i
    q = sa.scene_lattice(5, 3, 0.5, 0.01, seed=42, x0=1.0, y0=2.0)
    assert p.tobytes() == q.tobytes()


def _ref_fns():
    L = ctypes.CDLL(po.compiled_ref_path())
    return L


def test_host_helpers_match_reference_functions():
    """boundaryConditions / checkVelocity / updateVelocity / calculate_density
    / hash_val of libsph_hip.so against the same-named reference functions."""
    if not po.have_compiled_ref():
        pytest.skip("oracle/_ref not built here")
    R, L = _ref_fns(), sa.lib()
    rng = np.random.default_rng(0)
    box = sa.make_box(10, 8)
    for mover in (0, 1):
        prm = sa.Param()
        prm.tunable_params = sa.make_tunables(smoothing_radius=1.0, time_step=1 / 60.0,
                                              mover_type=mover, mover_center_x=5.0,
                                              mover_center_y=4.0, mover_width=3.0,
                                              mover_height=2.0)
        for _ in range(300):
            a = np.zeros(1, sa.PARTICLE)
            a["x"] = rng.uniform(-1, 11)
            a["y"] = rng.uniform(-1, 9)
            a["x_prev"] = rng.uniform(0, 10)
            a["y_prev"] = rng.uniform(0, 8)
            b = a.copy()
            for lib_, arr in ((R, a), (L, b)):
                lib_.boundaryConditions(ctypes.c_void_p(arr.ctypes.data), ctypes.byref(box),
                                        ctypes.byref(prm))
                lib_.updateVelocity(ctypes.c_void_p(arr.ctypes.data), ctypes.byref(prm))
            assert a.tobytes() == b.tobytes()
    for _ in range(200):
        v = (ctypes.c_float * 2)(*rng.uniform(-9, 9, 2))
        w = (ctypes.c_float * 2)(*v)
        R.checkVelocity(ctypes.byref(v, 0), ctypes.byref(v, 4))
        L.checkVelocity(ctypes.byref(w, 0), ctypes.byref(w, 4))
        assert list(v) == list(w)
    pq_a = np.zeros(2, sa.PARTICLE)
    pq_b = np.zeros(2, sa.PARTICLE)
    for ratio in (0.0, 0.3, 0.99, 1.0, 1.5):
        R.calculate_density(ctypes.c_void_p(pq_a.ctypes.data), ctypes.c_void_p(pq_a.ctypes.data + 52),
                            ctypes.c_float(ratio))
        L.calculate_density(ctypes.c_void_p(pq_b.ctypes.data), ctypes.c_void_p(pq_b.ctypes.data + 52),
                            ctypes.c_float(ratio))
    assert pq_a.tobytes() == pq_b.tobytes()

    class Grid(ctypes.Structure):
        _fields_ = [("spacing", ctypes.c_float), ("size_x", ctypes.c_uint), ("size_y", ctypes.c_uint),
                    ("neighbors", ctypes.c_void_p), ("grid_buckets", ctypes.c_void_p),
                    ("max_neighbors", ctypes.c_uint), ("max_bucket_size", ctypes.c_uint)]
    g = Grid(0.7, 15, 12, None, None, 400, 100)
    R.hash_val.restype = L.hash_val.restype = ctypes.c_uint
    for _ in range(500):
        x, y = (ctypes.c_float(v) for v in rng.uniform(0, 10, 2))
        assert R.hash_val(x, y, ctypes.byref(g), None) == L.hash_val(x, y, ctypes.byref(g), None)


def test_no_device_fails_loudly():
    """Without a GPU the library reports SPH_ENODEV; there is no CPU path."""
    try:
        s = sa.Solver(sa.make_tunables(smoothing_radius=1.0, time_step=1 / 60.0),
                      sa.make_box(10, 10), 8)
    except sa.SphError as e:
        assert "(-5)" in str(e)
        return
    s.close()  # a GPU is present: nothing to check here


def test_driver_builds():
    from sph_amd import build
    exe = build.build_driver()
    if exe is None:
        pytest.skip("driver source absent")
    assert os.access(exe, os.X_OK)


def test_gather_kernels_use_no_scratch_memory():
    """The gather kernels keep their walk state (the accepted-chunk queue,
    the row bounds, the staged records) in registers and LDS.  A queue
    written as selects between captured variables was once folded into
    loads from a selected address: scratch memory, and twice the
    relaxation kernel's time.  hipcc's resource report must show no scratch
    for any instance of the three gather kernels."""
    from sph_amd import build as bd
    src = os.path.join(bd.CSRC, "sph_kernels.hip")
    r = subprocess.run([bd.HIPCC, *bd.HIP_FLAGS, *bd.SRC_FLAGS.get("sph_kernels.hip", []),
                        "--cuda-device-only", "-c", src, "-o", os.devnull,
                        "-Rpass-analysis=kernel-resource-usage"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    name, scratch = None, {}
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            scratch[name] = int(m.group(1))
    hot = {k: v for k, v in scratch.items()
           if any(t in k for t in ("k_visc_predict", "k_density", "k_relax"))}
    assert len(hot) >= 4, sorted(scratch)
    assert all(v == 0 for v in hot.values()), hot
