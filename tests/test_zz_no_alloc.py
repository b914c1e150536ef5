"""sph_scene_density_noalloc - the entry point that allocates nothing -
against sph_scene_density, through the C ABI: two calls in a row, the second
on a smaller input (fewer scenes), bit-equal results; and the scene limit of
its fixed scratch array is rejected with a clear error."""
import numpy as np
import pytest

import sph_amd as sa
from sph_amd import scenes as sc

pytestmark = pytest.mark.gpu


def densities(s, name):
    ns = s.nscenes
    mean = np.zeros(ns, np.float32)
    mx = np.zeros(ns, np.float32)
    rc = getattr(sa.lib(), name)(s._h, mean.ctypes.data, mx.ctypes.data)
    assert rc == 0, sa.lib().sph_surrogate_last_error()
    return mean, mx


def batch(nscenes, seed):
    rng = np.random.default_rng(seed)
    scenes, parts = [], []
    for i in range(nscenes):
        nx, ny = int(rng.integers(6, 30)), int(rng.integers(6, 30))
        box = sa.make_box(float(nx) * 0.9 + 4.0, float(ny) * 0.9 + 4.0)
        p = sa.scene_lattice(nx, ny, 0.5, 0.01, seed=seed * 100 + i, x0=0.25, y0=0.25)
        if i == 1:
            p = p[:0].copy()  # an empty scene
        scenes.append((sc.preset_tunables(i % 4), box, max(len(p), 1)))
        parts.append(p)
    s = sa.Solver.batch(scenes)
    s.upload_batch(np.concatenate(parts), [len(p) for p in parts])
    s.step(5)
    return s


def test_noalloc_scene_density_equals_allocating_call():
    for nscenes, seed in ((9, 1), (3, 2)):  # the second call on a smaller input
        s = batch(nscenes, seed)
        print(f"{nscenes} scenes", flush=True)
        m1, x1 = densities(s, "sph_scene_density")
        m2, x2 = densities(s, "sph_scene_density_noalloc")
        assert np.array_equal(m1.view(np.uint32), m2.view(np.uint32)), (m1, m2)
        assert np.array_equal(x1.view(np.uint32), x2.view(np.uint32)), (x1, x2)
        assert m1[1] == 0.0 and np.all(m1[m1 != 0] > 0)
        s.close()


def test_noalloc_scene_limit_rejected():
    t = sc.preset_tunables(0)
    box = sa.make_box(1.0, 1.0)
    s = sa.Solver.batch([(t, box, 1)] * 4097)
    mean = np.zeros(4097, np.float32)
    mx = np.zeros(4097, np.float32)
    rc = sa.lib().sph_scene_density_noalloc(s._h, mean.ctypes.data, mx.ctypes.data)
    assert rc == -1
    assert b"4096" in sa.lib().sph_surrogate_last_error()
    s.close()
