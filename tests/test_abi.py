"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol include/snelly.h declares, and the scene translator accepts the
reference's scene snippets (no GPU calls)."""
import ctypes
import os
import re

import pytest

from snelly_amd import host, native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "snelly.h")).read()
    return sorted(set(re.findall(r"\b(snl_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    assert native.lib().snl_version().startswith(b"snelly-mi355x")


@pytest.mark.parametrize("make", [scenes.sphere_plane, scenes.menger_metal, scenes.mandelbox_dielectric])
def test_scene_compiles_for_gfx950(make):
    """translate + hipRTC compile of the benchmark scenes (cached code objects)"""
    s = make(8, 8, 1)
    sh, ior, d = host.assemble(s.glsl, s.renderer, s.materials)
    r = s.renderer
    native.precompile(sh, ior, d, r.maxMarchSteps, r.maxBounces, r.maxAtmosphereScatters, r.maxSamplesPerFrame)


def test_compile_error_is_reported():
    with pytest.raises(native.SnellyError) as e:
        native.precompile("float SDF_SURFACE(vec3 X) { return lenght(X) - 1.0; }", "", "#define HAS_SURFACE\n",
                          64, 1, 1, 1)
    assert "lenght" in str(e.value)


def test_scene_without_geometry_fails_like_the_reference():
    s = scenes.sphere_plane(8, 8, 1)
    with pytest.raises(native.SnellyError):
        host.assemble("float foo(vec3 X) { return 1.0; }", s.renderer, s.materials)
