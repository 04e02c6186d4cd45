"""Pin the per-step FLOP figures the benchmark's roofline uses against the
counting build of the oracle (every + - * / sqrt of one SDF evaluation)."""
import ctypes

import numpy as np
import pytest

import oracle_py
from snelly_amd import scenes


def count(scene, material, p):
    L = oracle_py.lib()
    L.oracle_sdf_count.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                   ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    oracle_py.Oracle(scene, 4, 4)
    f, o = ctypes.c_longlong(), ctypes.c_longlong()
    L.oracle_sdf_count(material, oracle_py.fp(np.asarray(p, np.float32)), ctypes.byref(f), ctypes.byref(o))
    return f.value


@pytest.mark.parametrize("it", [2, 3, 4, 5])
def test_menger_flops(it):
    s = scenes.menger_metal(4, 4, 1, iterations=it)
    for p in [(0.3, 0.2, 0.1), (2.0, -1.0, 0.5)]:
        assert count(s, 1, p) == scenes.MENGER_SDF_FLOPS[it]
    assert s.flops_per_step == 8 + scenes.MENGER_SDF_FLOPS[it]


def test_sphere_plane_flops():
    s = scenes.sphere_plane(4, 4, 1)
    assert count(s, 2, (0.3, 0.2, 0.1)) == scenes.SPHERE_PLANE_SDF_FLOPS
