"""ctypes front of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline use it, always as the checker of the HIP path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
# the counting build of the same source (counting.h): identical results, plus
# operation counts per category
LIB_COUNT = os.path.join(HERE, "_build", "liboracle_count.so")
_libs = {}

DIELECTRIC_MODELS = {"AbbeDielectric": 0, "SellmeierDielectric": 1, "Sellmeier2Dielectric": 2,
                     "PolyanskiyDielectric": 3, "CauchyDielectric": 4}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(counting=False):
    if counting not in _libs:
        path = LIB_COUNT if counting else LIB
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        fp = ctypes.POINTER(ctypes.c_float)
        L.oracle_set.argtypes = [ctypes.c_char_p, fp, ctypes.c_int]
        L.oracle_config.argtypes = [ctypes.c_int] * 6
        L.oracle_sdf.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 4
        L.oracle_albedo_hook.argtypes = [ctypes.c_int]
        L.oracle_volume.argtypes = [ctypes.c_int] * 3
        L.oracle_emission_hook.argtypes = [ctypes.c_int]
        L.oracle_mode.argtypes = [ctypes.c_int]
        L.oracle_envmap.argtypes = [ctypes.c_int, ctypes.c_int, fp]
        L.oracle_dielectric.argtypes = [ctypes.c_int, fp, ctypes.c_int]
        L.oracle_metal.argtypes = [ctypes.c_int, fp, fp]
        L.oracle_luts.argtypes = [fp, fp]
        L.oracle_rand.argtypes = [fp]
        L.oracle_rand.restype = ctypes.c_float
        for n in ["oracle_ior_diele", "oracle_ior_metal", "oracle_k_metal"]:
            getattr(L, n).argtypes = [ctypes.c_float]
            getattr(L, n).restype = ctypes.c_float
        L.oracle_fresnel_dielectric.argtypes = [ctypes.c_float, ctypes.c_float]
        L.oracle_fresnel_dielectric.restype = ctypes.c_float
        L.oracle_fresnel_metal.argtypes = [ctypes.c_float] * 3
        L.oracle_fresnel_metal.restype = ctypes.c_float
        L.oracle_microfacet_eval.argtypes = [fp, ctypes.c_float]
        L.oracle_microfacet_eval.restype = ctypes.c_float
        L.oracle_smith_g1.argtypes = [fp, fp, ctypes.c_float]
        L.oracle_smith_g1.restype = ctypes.c_float
        L.oracle_sdf_eval.argtypes = [ctypes.c_int, fp]
        L.oracle_sdf_eval.restype = ctypes.c_float
        L.oracle_bsdf.argtypes = [ctypes.c_int, ctypes.c_float, fp, fp, fp, fp, fp]
        L.oracle_bsdf_sample.argtypes = [ctypes.c_int, ctypes.c_float, fp, fp, fp, fp, fp]
        L.oracle_seed.argtypes = [fp, ctypes.c_int, ctypes.c_uint]
        L.oracle_render.argtypes = [fp, fp] + [ctypes.c_int] * 6
        L.oracle_normals.argtypes = [fp, fp] + [ctypes.c_int] * 4
        L.oracle_pick.argtypes = [fp, ctypes.c_int, fp]
        L.oracle_tonemap.argtypes = [fp, ctypes.c_longlong, fp]
        L.oracle_medium.argtypes = [ctypes.c_int] + [ctypes.c_float] * 4
        for n in ["oracle_value_noise", "oracle_fbm3", "oracle_medium_density", "oracle_medium_bound"]:
            getattr(L, n).argtypes = [fp]
            getattr(L, n).restype = ctypes.c_float
        L.oracle_lattice_value.argtypes = [ctypes.c_int]
        L.oracle_lattice_value.restype = ctypes.c_float
        for n in ["oracle_medium_free_flight", "oracle_medium_transmittance"]:
            getattr(L, n).argtypes = [fp, fp, ctypes.c_float, fp]
            getattr(L, n).restype = ctypes.c_float
        L.oracle_sample_hg.argtypes = [fp, ctypes.c_float, fp, fp, fp]
        L.oracle_phase_function.argtypes = [ctypes.c_float, ctypes.c_float]
        L.oracle_phase_function.restype = ctypes.c_float
        if counting:
            L.oracle_count_get.argtypes = [ctypes.POINTER(ctypes.c_longlong)]
        _libs[counting] = L
    return _libs[counting]


def fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def f3(v):
    return np.ascontiguousarray(v, np.float32)


class Oracle:
    """The oracle configured for one scene (a snelly_amd.scenes.Scene)."""

    def __init__(self, scene, width, height, counting=False):
        from snelly_amd import host  # plumbing: uniform values, LUTs
        self.scene, self.width, self.height = scene, width, height
        self.L = L = lib(counting)
        L.oracle_reset()
        r, mats = scene.renderer, scene.materials
        scene.camera.aspect = width / height
        u = host.frame_uniforms(r, scene.camera, scene.scales, mats, width, height)
        env = getattr(scene, "env_map", None)
        u["haveEnvMap"] = int(env is not None)
        for k, v in u.items():
            arr = np.atleast_1d(np.asarray(v, np.float32)).ravel()
            L.oracle_set(k.encode(), fp(arr), arr.size)
        if env is not None:
            self.env = f3(host.srgb_to_linear(np.asarray(env[2], np.uint8)).reshape(-1))
            L.oracle_envmap(env[0], env[1], fp(self.env))
        sphere_light = 1 if r.sphereLightPower > -7.0 else 0
        L.oracle_config(int(r.dispersive), int(r.interactive), int(round(r.maxBounces)),
                        int(round(r.maxMarchSteps)), int(round(r.maxSamplesPerFrame)), sphere_light)
        v = mats.volume
        # a scene with SDF_VOLUME has the heterogeneous medium instead of the atmosphere
        has_medium = host.has_hook(scene.glsl, "SDF_VOLUME")
        has_atm = not has_medium and (any(c > 0.0 for c in v.scatteringColor) or any(c > 0.0 for c in v.absorptionColor))
        L.oracle_volume(int(has_atm), int(bool(v.fogEnable)), int(round(r.maxAtmosphereScatters)))
        medium = getattr(scene, "oracle_medium", None)
        assert has_medium == (medium is not None), "a scene with SDF_VOLUME needs its oracle restatement"
        L.oracle_medium(*(medium or (0, 0.0, 0.0, 0.0, 0.0)))
        for (m, kind, a, b, c, d) in scene.oracle_sdfs:
            L.oracle_sdf(m, kind, a, b, c, d)
        L.oracle_albedo_hook(scene.oracle_albedo_hook)
        L.oracle_emission_hook(getattr(scene, "oracle_emission_hook", 0))
        L.oracle_mode({"ao": 1, "ptsimple": 2}.get(r.renderMode, 0))
        diel = mats.dielectric
        model = DIELECTRIC_MODELS[type(diel).__name__]
        coeffs = f3([diel.iorVal, diel.abbe] if model == 0 else diel.coeffs)
        L.oracle_dielectric(model, fp(coeffs), coeffs.size)
        met = mats.metal
        if isinstance(met, host.TabulatedMetal):
            L.oracle_metal(0, fp(met.n), fp(met.k))
        else:
            L.oracle_metal(1, None, None)
        self.xyz = f3(host.xyz_lut())
        self.icdf = f3(host.inverse_cdf(lambda wl: host.blackbody(wl, r.skyTemperature)))
        L.oracle_luts(fp(self.xyz), fp(self.icdf))

    @staticmethod
    def seed(npix, seed):
        rng = np.empty((npix, 4), np.float32)
        lib().oracle_seed(fp(rng), npix, seed)
        return rng

    # the counting build's totals (Oracle(..., counting=True)): FLOPs,
    # transcendentals and other operations of each category, SDF evaluations of
    # the sphere tracer, normals and camera paths since the last reset
    CATEGORIES = ("shade", "march", "normal", "medium", "init")

    def count_reset(self):
        self.L.oracle_count_reset()

    def counts(self):
        out = (ctypes.c_longlong * 18)()
        self.L.oracle_count_get(out)
        v = list(out)
        cats = self.CATEGORIES
        return {"flops": dict(zip(cats, v[0:5])), "transcendentals": dict(zip(cats, v[5:10])),
                "other_ops": dict(zip(cats, v[10:15])), "steps": v[15], "normals": v[16], "paths": v[17]}

    def new_state(self, seed=1234):
        npix = self.width * self.height
        rad = np.zeros((npix, 4), np.float32)
        rad[:, 3] = 1.0
        return rad, self.seed(npix, seed)

    def render(self, rad, rng, frames, stride=1, offset=0, threads=0):
        self.L.oracle_render(fp(rad), fp(rng), self.width, self.height, frames, stride, offset, threads)

    def normals(self, rad, rng, frames, threads=0):
        self.L.oracle_normals(fp(rad), fp(rng), self.width, self.height, frames, threads)

    def pick(self, xy):
        xy = f3(xy).reshape(-1, 2)
        out = np.empty_like(xy)
        self.L.oracle_pick(fp(xy), xy.shape[0], fp(out))
        return out

    def tonemap(self, rad):
        rad = f3(rad).reshape(-1, 4)
        out = np.empty((rad.shape[0], 3), np.float32)
        self.L.oracle_tonemap(fp(rad), rad.shape[0], fp(out))
        return out
