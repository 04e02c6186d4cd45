"""Scenes whose golden vectors come from the reference's own GLSL
(shaders/pathtracer-fragment-shader.glsl), executed by a GLES 3 implementation
through reference_gles/gles_render.  make_reference_fixtures.py writes
tests/golden/ref_<name>.npz; the tests rebuild the same scene from this table.

Each case: builder from snelly_amd.scenes, its keyword arguments, frames,
RNG seed, and an optional tweak applied to the built scene.
"""
import math

from snelly_amd import host, scenes


def _sun_and_specular(s):
    """sphere over plane lit by sky (tinted) and sun, with a specular lobe"""
    r = s.renderer
    r.sunPower = 1.0
    r.sunAngularSize = 8.0
    r.sunLatitude = 40.0
    r.sunLongitude = 30.0
    r.sunColor = [1.0, 0.9, 0.8]
    r.skyTintUp = [0.6, 0.8, 1.0]
    r.skyTintDown = [0.9, 0.8, 0.7]
    su = s.materials.surface
    su.specAlbedo = [0.6, 0.6, 0.6]
    su.ior = 1.5
    su.roughness = 0.2


def _tabulated_metal(s):
    """a metal torus of tabulated gold (the reference's TabulatedMetal),
    whose K_METAL samples texture unit 0 (js/materials.js:665)"""
    s.glsl = scenes.TORUS_GLSL.replace("@MAT@", "METAL")
    s.materials.metal = host.metal("Gold")
    s.materials.metal.roughness = 0.1
    s.oracle_sdfs = [(1, 5, 1.0, 0.3, 0, 0)]  # torus R = 1, r = 0.3


def _bk7_sphere(s):
    """a BK7 sphere (Sellmeier IOR, dispersion) over a diffuse plane, sun on"""
    s.glsl = """
float SDF_DIELECTRIC(vec3 X) { return length(X) - 1.0; }
float SDF_SURFACE(vec3 X) { return X.y + 1.0; }
"""
    s.materials.dielectric = host.DIELECTRICS["Glass (BK7)"]()
    s.materials.dielectric.roughness = 0.02
    s.materials.dielectric.absorptionScale = 10.0
    s.renderer.sunPower = 0.5
    s.renderer.sunLongitude = 60.0
    s.materials.surface.diffuseAlbedo = [0.7, 0.7, 0.7]
    s.oracle_sdfs = [(0, 1, 1.0, 0, 0, 0), (2, 6, 1.0, 0, 0, 0)]  # unit sphere; plane y = -1
    s.oracle_albedo_hook = 0


CASES = {
    "sphere_plane": dict(builder="sphere_plane", kwargs=dict(width=64, height=64, spp=16), frames=16, seed=1234),
    "furnace_rgb": dict(builder="furnace", kwargs=dict(width=32, height=32, spp=8, dispersive=False), frames=8,
                        seed=7),
    "menger_metal": dict(builder="menger_metal", kwargs=dict(width=48, height=48, spp=4), frames=4, seed=1234),
    "mandelbox_dielectric": dict(builder="mandelbox_dielectric", kwargs=dict(width=48, height=36, spp=4), frames=4,
                                 seed=99),
    "sun_specular": dict(builder="sphere_plane", kwargs=dict(width=48, height=48, spp=8), frames=8, seed=5,
                         tweak=_sun_and_specular),
    "tabulated_metal": dict(builder="sphere_plane", kwargs=dict(width=40, height=32, spp=4), frames=4, seed=11,
                            tweak=_tabulated_metal),
    "bk7_sphere": dict(builder="sphere_plane", kwargs=dict(width=48, height=40, spp=4), frames=4, seed=3,
                       tweak=_bk7_sphere),
}


def build(name):
    c = CASES[name]
    s = getattr(scenes, c["builder"])(**c["kwargs"])
    if c.get("tweak"):
        c["tweak"](s)
    return s, c["kwargs"]["width"], c["kwargs"]["height"], c["frames"], c["seed"]
