"""Ahead-of-time compilation of scene programs into the code-object cache
(snelly_amd/_kcache), so that a process that renders them later -- on a
machine without a compiler toolchain, or under a time limit -- only loads them.
hipRTC compiles one program per call on one core; programs are spread over a
small process pool."""
import os
from concurrent.futures import ProcessPoolExecutor

from . import host, native


def program_key(scene, flags=0):
    r = scene.renderer
    shader, ior, defines = host.assemble(scene.glsl, r, scene.materials)
    return (shader, ior, defines, int(round(r.maxMarchSteps)), int(round(r.maxBounces)),
            int(round(r.maxAtmosphereScatters)), int(round(r.maxSamplesPerFrame)), int(flags) | host.mode_flags(r))


def _compile(key):
    native.precompile(*key)
    return key[3]


def precompile_scenes(scenes, flags=0, workers=None):
    """compile each distinct program of `scenes` once; returns how many"""
    keys = list(dict.fromkeys(program_key(s, flags) for s in scenes))
    workers = workers or max(1, min(len(keys), (os.cpu_count() or 2) - 2, 8))
    if workers == 1 or len(keys) == 1:
        for k in keys:
            _compile(k)
    else:
        with ProcessPoolExecutor(max_workers=workers) as ex:
            list(ex.map(_compile, keys))
    return len(keys)


def precompile_batch_module(scenes, flags=0):
    """compile the batch module (snl_batch_module) of the distinct programs of
    `scenes` into the cache; returns how many kinds it holds"""
    keys = list(dict.fromkeys(program_key(s, flags) for s in scenes))
    mods = [(k[0], k[1], k[2], k[3], k[5]) for k in keys]
    native.batch_module_precompile(mods, int(flags))
    return len(mods)
