"""ctypes binding of libsnelly_hip.so (include/snelly.h).

The library is built in-tree by snelly_amd/build.py.  Importing this module never
falls back to anything else: a missing or unloadable library raises at once.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsnelly_hip.so")

T_FLOAT, T_VEC2, T_VEC3, T_VEC4, T_INT, T_MAT2, T_MAT3, T_MAT4, T_SAMPLER2D, T_SAMPLER3D = range(1, 11)
BUILD_PRECISE = 1
BUILD_FAST = 2
BUILD_AO = 4
BUILD_PTSIMPLE = 8

# every symbol include/snelly.h declares
EXPORTS = [
    "snl_ctx_create", "snl_ctx_destroy", "snl_last_error", "snl_version", "snl_program_build",
    "snl_program_precompile", "snl_program_destroy", "snl_uniform_index", "snl_uniform_set",
    "snl_uniform_texture", "snl_texture_create", "snl_texture_upload", "snl_texture_destroy",
    "snl_state_create", "snl_state_destroy", "snl_state_seed", "snl_state_clear", "snl_state_read",
    "snl_state_write", "snl_render", "snl_render_normals", "snl_pick", "snl_tonemap", "snl_sync",
    "snl_get_stats", "snl_set_tuning", "snl_set_profiling", "snl_set_engine", "snl_render_batch",
    "snl_state_create_chains", "snl_state_error", "snl_ctx_mem_info", "snl_state_bytes", "snl_set_march_burst",
    "snl_mega_variants", "snl_set_mega_variant",
    "snl_render_noalloc", "snl_program_state_bytes", "snl_state_info", "snl_texture_info",
    "snl_batch_submit", "snl_job_poll", "snl_batch_wait", "snl_batch_job_stats", "snl_batch_release",
    "snl_batch_module", "snl_batch_module_precompile", "snl_ctx_set_batch_fusion",
    "snl_batch_graph_create", "snl_batch_graph_launch", "snl_batch_graph_destroy",
]
ENGINE_WAVEFRONT, ENGINE_PERSISTENT, ENGINE_MEGAKERNEL = 0, 1, 2


class SnellyError(RuntimeError):
    pass


class Stats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_trace", ctypes.c_double), ("trace_steps", ctypes.c_longlong),
                ("paths", ctypes.c_longlong), ("trace_launches", ctypes.c_int), ("normal_evals", ctypes.c_longlong)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SnellyError(f"{LIB_PATH} is missing: run python3 snelly_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, cp, fp = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_float)
    sig = {
        "snl_ctx_create": (ip, [ip, ctypes.POINTER(vp)]),
        "snl_ctx_destroy": (ip, [vp]),
        "snl_last_error": (cp, [vp]),
        "snl_version": (cp, []),
        "snl_program_build": (ip, [vp, cp, cp, cp, ip, ip, ip, ip, ctypes.c_uint, ctypes.POINTER(vp), cp, ip]),
        "snl_program_precompile": (ip, [cp, cp, cp, ip, ip, ip, ip, ctypes.c_uint, cp, ip]),
        "snl_program_destroy": (ip, [vp]),
        "snl_uniform_index": (ip, [vp, cp]),
        "snl_uniform_set": (ip, [vp, ip, ip, vp, ip]),
        "snl_uniform_texture": (ip, [vp, ip, vp]),
        "snl_texture_create": (ip, [vp, ip, ip, ip, ip, ip, ip, fp, ctypes.POINTER(vp)]),
        "snl_texture_upload": (ip, [vp, fp]),
        "snl_texture_destroy": (ip, [vp]),
        "snl_state_create": (ip, [vp, ip, ip, ctypes.POINTER(vp)]),
        "snl_state_destroy": (ip, [vp]),
        "snl_state_seed": (ip, [vp, ctypes.c_uint]),
        "snl_state_clear": (ip, [vp]),
        "snl_state_read": (ip, [vp, ip, fp]),
        "snl_state_write": (ip, [vp, ip, fp]),
        "snl_render": (ip, [vp, vp, ip]),
        "snl_render_noalloc": (ip, [vp, vp, ip]),
        "snl_render_normals": (ip, [vp, vp, ip]),
        "snl_pick": (ip, [vp, fp, ip, fp]),
        "snl_tonemap": (ip, [vp, vp, fp, ctypes.POINTER(ctypes.c_ubyte)]),
        "snl_sync": (ip, [vp]),
        "snl_get_stats": (ip, [vp, ctypes.POINTER(Stats)]),
        "snl_set_tuning": (ip, [vp, ip, ip, ip]),
        "snl_set_profiling": (ip, [vp, ip]),
        "snl_set_engine": (ip, [vp, ip, ip, ip, ip, ip]),
        "snl_render_batch": (ip, [ip, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(ip),
                                  ctypes.POINTER(Stats)]),
        "snl_state_create_chains": (ip, [vp, ip, ip, ip, ctypes.POINTER(vp)]),
        "snl_state_error": (ip, [vp, ctypes.POINTER(ctypes.c_double)]),
        "snl_ctx_mem_info": (ip, [vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]),
        "snl_state_bytes": (ctypes.c_ulonglong, [ip, ip, ip]),
        "snl_program_state_bytes": (ctypes.c_ulonglong, [vp, ip, ip, ip]),
        "snl_state_info": (ip, [vp] + [ctypes.POINTER(ip)] * 3),
        "snl_texture_info": (ip, [vp] + [ctypes.POINTER(ip)] * 4),
        "snl_set_march_burst": (ip, [vp, ip]),
        "snl_mega_variants": (ip, [vp]),
        "snl_set_mega_variant": (ip, [vp, ip]),
        "snl_batch_submit": (ip, [ip, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(ip),
                                  ctypes.POINTER(vp)]),
        "snl_job_poll": (ip, [vp, ip]),
        "snl_batch_wait": (ip, [vp, ctypes.POINTER(Stats)]),
        "snl_batch_job_stats": (ip, [vp, ip, ctypes.POINTER(Stats)]),
        "snl_batch_release": (ip, [vp]),
        "snl_batch_module": (ip, [vp, ip, ctypes.POINTER(vp)]),
        "snl_batch_module_precompile": (ip, [ip, ctypes.POINTER(cp), ctypes.POINTER(cp), ctypes.POINTER(cp),
                                             ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.c_uint]),
        "snl_ctx_set_batch_fusion": (ip, [vp, ip]),
        "snl_batch_graph_create": (ip, [ip, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(ip),
                                        ctypes.POINTER(vp)]),
        "snl_batch_graph_launch": (ip, [vp, ctypes.POINTER(Stats)]),
        "snl_batch_graph_destroy": (ip, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def check(rc, ctx=None):
    if rc != 0:
        msg = lib().snl_last_error(ctx).decode(errors="replace")
        raise SnellyError(f"snelly error {rc}: {msg}")


def state_error(st):
    """(sum of the variance of the chains' mean, sum of its square) for a state
    of two or more chains"""
    out = (ctypes.c_double * 2)()
    st.ctx.check(lib().snl_state_error(st.h, out))
    return out[0], out[1]


def mem_info(ctx):
    f, t = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    ctx.check(lib().snl_ctx_mem_info(ctx.h, ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def state_bytes(width, height, chains=1, prog=None):
    """device bytes of a state; with `prog`, those of a state rendered by that
    program (its texture-unit-0 snapshot included)"""
    if prog is not None:
        return int(lib().snl_program_state_bytes(prog.h, width, height, chains))
    return int(lib().snl_state_bytes(width, height, chains))


def render_batch(jobs):
    """snl_render_batch over [(Program, State, frames), ...]; returns the Stats"""
    n = len(jobs)
    progs = (ctypes.c_void_p * n)(*[p.h for p, _, _ in jobs])
    states = (ctypes.c_void_p * n)(*[s.h for _, s, _ in jobs])
    frames = (ctypes.c_int * n)(*[int(f) for _, _, f in jobs])
    st = Stats()
    if n == 0:
        return st
    jobs[0][0].ctx.check(lib().snl_render_batch(n, progs, states, frames, ctypes.byref(st)))
    return st


class Batch:
    """snl_batch_submit over [(Program, State, frames), ...]: the jobs' launches
    are queued and the call returns; poll(j) / wait() / job_stats(j) follow
    them, release() (or close()) frees the handle.  The jobs' programs and
    states must stay open until the batch is released."""

    def __init__(self, jobs):
        n = len(jobs)
        if n == 0:
            raise SnellyError("an empty batch")
        self.ctx = jobs[0][0].ctx
        self.n = n
        progs = (ctypes.c_void_p * n)(*[p.h for p, _, _ in jobs])
        states = (ctypes.c_void_p * n)(*[s.h for _, s, _ in jobs])
        frames = (ctypes.c_int * n)(*[int(f) for _, _, f in jobs])
        h = ctypes.c_void_p()
        self.ctx.check(lib().snl_batch_submit(n, progs, states, frames, ctypes.byref(h)))
        self.h = h

    def poll(self, job=-1):
        """True once job `job` (-1: every job) has completed"""
        rc = lib().snl_job_poll(self.h, job)
        if rc < 0:
            self.ctx.check(rc)
        return rc == 1

    def wait(self):
        st = Stats()
        self.ctx.check(lib().snl_batch_wait(self.h, ctypes.byref(st)))
        return st

    def job_stats(self, job):
        st = Stats()
        self.ctx.check(lib().snl_batch_job_stats(self.h, job, ctypes.byref(st)))
        return st

    def release(self):
        if self.h:
            rc = lib().snl_batch_release(self.h)
            self.h = None
            self.ctx.check(rc)

    close = release

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().snl_batch_release(self.h)
        except Exception:
            pass


class Graph:
    """snl_batch_graph_create over [(Program, State, frames), ...]: the batch's
    launch sequence captured once; launch() replays it (frames more per job)"""

    def __init__(self, jobs):
        n = len(jobs)
        self.ctx = jobs[0][0].ctx
        progs = (ctypes.c_void_p * n)(*[p.h for p, _, _ in jobs])
        states = (ctypes.c_void_p * n)(*[s.h for _, s, _ in jobs])
        frames = (ctypes.c_int * n)(*[int(f) for _, _, f in jobs])
        h = ctypes.c_void_p()
        self.ctx.check(lib().snl_batch_graph_create(n, progs, states, frames, ctypes.byref(h)))
        self.h = h

    def launch(self):
        st = Stats()
        self.ctx.check(lib().snl_batch_graph_launch(self.h, ctypes.byref(st)))
        return st

    def close(self):
        if self.h:
            lib().snl_batch_graph_destroy(self.h)
            self.h = None


def batch_module(ctx, progs):
    """snl_batch_module: load (building if needed) the batch module of these
    Programs' kinds"""
    n = len(progs)
    arr = (ctypes.c_void_p * n)(*[p.h for p in progs])
    ctx.check(lib().snl_batch_module(ctx.h, n, arr))


def set_batch_fusion(ctx, on):
    ctx.check(lib().snl_ctx_set_batch_fusion(ctx.h, 1 if on else 0))


def batch_module_precompile(keys, flags):
    """snl_batch_module_precompile over program keys (shader, ior, defines,
    max_march, max_atm)"""
    n = len(keys)
    enc = lambda i: (ctypes.c_char_p * n)(*[k[i].encode() for k in keys])
    rc = lib().snl_batch_module_precompile(n, enc(0), enc(1), enc(2), (ctypes.c_int * n)(*[k[3] for k in keys]),
                                           (ctypes.c_int * n)(*[k[4] for k in keys]), flags)
    if rc != 0:
        raise SnellyError(f"batch module precompile failed ({rc}): {lib().snl_last_error(None).decode()[:4000]}")


def precompile(shader, ior, defines, max_march, max_bounces, max_atm, spf, flags=0):
    log = ctypes.create_string_buffer(1 << 16)
    rc = lib().snl_program_precompile(shader.encode(), ior.encode(), defines.encode(), max_march, max_bounces,
                                      max_atm, spf, flags, log, len(log))
    if rc != 0:
        raise SnellyError(f"precompile failed ({rc}): {lib().snl_last_error(None).decode()}\n{log.value.decode()[:4000]}")


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().snl_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise SnellyError(f"snl_ctx_create failed ({rc}): {lib().snl_last_error(None).decode()}")
        self.h = h

    def check(self, rc):
        check(rc, self.h)

    def sync(self):
        self.check(lib().snl_sync(self.h))

    def close(self):
        if self.h:
            lib().snl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Texture:
    def __init__(self, ctx, data, width, height=1, depth=1, channels=1, linear=True, clamp=True):
        self.ctx = ctx
        h = ctypes.c_void_p()
        arr = None if data is None else np.ascontiguousarray(data, dtype=np.float32)
        ptr = _fptr(arr) if arr is not None else None
        ctx.check(lib().snl_texture_create(ctx.h, width, height, depth, channels, int(linear), int(clamp), ptr,
                                           ctypes.byref(h)))
        self.h = h

    def upload(self, data):
        arr = np.ascontiguousarray(data, dtype=np.float32)
        self.ctx.check(lib().snl_texture_upload(self.h, _fptr(arr)))

    def close(self):
        if self.h:
            lib().snl_texture_destroy(self.h)
            self.h = None


class State:
    def __init__(self, ctx, width, height, seed=1234, chains=1):
        self.ctx, self.width, self.height, self.chains = ctx, width, height, chains
        h = ctypes.c_void_p()
        ctx.check(lib().snl_state_create_chains(ctx.h, width, height, chains, ctypes.byref(h)))
        self.h = h
        self.seed(seed)

    def seed(self, seed):
        self.ctx.check(lib().snl_state_seed(self.h, seed & 0xFFFFFFFF))

    def clear(self):
        self.ctx.check(lib().snl_state_clear(self.h))

    def read(self, which=0):
        """(height, width, 4) float32, or (chains, height, width, 4) for several chains"""
        shape = (self.height, self.width, 4) if self.chains == 1 else (self.chains, self.height, self.width, 4)
        out = np.empty(shape, np.float32)
        self.ctx.check(lib().snl_state_read(self.h, which, _fptr(out)))
        return out

    def write(self, which, data):
        arr = np.ascontiguousarray(data, dtype=np.float32).reshape(self.chains * self.height, self.width, 4)
        self.ctx.check(lib().snl_state_write(self.h, which, _fptr(arr)))

    def close(self):
        if self.h:
            lib().snl_state_destroy(self.h)
            self.h = None


class Program:
    def __init__(self, ctx, shader, ior, defines, max_march=256, max_bounces=3, max_atm=1, spf=1, flags=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        log = ctypes.create_string_buffer(1 << 16)
        rc = lib().snl_program_build(ctx.h, shader.encode(), ior.encode(), defines.encode(), max_march, max_bounces,
                                     max_atm, spf, flags, ctypes.byref(h), log, len(log))
        if rc != 0:
            raise SnellyError(f"program build failed ({rc}): {lib().snl_last_error(ctx.h).decode()}\n"
                              f"{log.value.decode()[:4000]}")
        self.h = h
        self.engine_id = ENGINE_MEGAKERNEL  # a program starts with the megakernel

    def index(self, name):
        return lib().snl_uniform_index(self.h, name.encode())

    def set(self, name, value, type_code=None):
        idx = self.index(name)
        if idx < 0:
            return False
        if type_code == T_INT or (type_code is None and isinstance(value, (bool, np.bool_, np.integer))):
            arr = np.atleast_1d(np.asarray(value, dtype=np.int32))
            t = T_INT
            count = arr.size
        else:
            arr = np.atleast_1d(np.asarray(value, dtype=np.float32)).ravel()
            t = type_code or {1: T_FLOAT, 2: T_VEC2, 3: T_VEC3, 4: T_VEC4, 9: T_MAT3, 16: T_MAT4}.get(arr.size, T_FLOAT)
            per = {T_FLOAT: 1, T_VEC2: 2, T_VEC3: 3, T_VEC4: 4, T_MAT2: 4, T_MAT3: 9, T_MAT4: 16}[t]
            count = max(1, arr.size // per)
        self.ctx.check(lib().snl_uniform_set(self.h, idx, t, arr.ctypes.data_as(ctypes.c_void_p), count))
        return True

    def set_texture(self, name, tex):
        idx = self.index(name)
        if idx >= 0:
            self.ctx.check(lib().snl_uniform_texture(self.h, idx, tex.h if tex is not None else None))

    def render(self, state, frames):
        self.ctx.check(lib().snl_render(self.h, state.h, frames))

    def render_normals(self, state, frames):
        self.ctx.check(lib().snl_render_normals(self.h, state.h, frames))

    def pick(self, xy):
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1, 2)
        out = np.empty((xy.shape[0], 2), np.float32)
        self.ctx.check(lib().snl_pick(self.h, _fptr(xy), xy.shape[0], _fptr(out)))
        return out

    def tonemap(self, state):
        rgb = np.empty((state.height, state.width, 3), np.float32)
        rgba8 = np.empty((state.height, state.width, 4), np.uint8)
        self.ctx.check(lib().snl_tonemap(self.h, state.h, _fptr(rgb),
                                         rgba8.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))))
        return rgb, rgba8

    def stats(self):
        s = Stats()
        self.ctx.check(lib().snl_get_stats(self.h, ctypes.byref(s)))
        return s

    def tune(self, trace_blocks=0, trace_chunk=0, shade_blocks=0):
        self.ctx.check(lib().snl_set_tuning(self.h, trace_blocks, trace_chunk, shade_blocks))

    def march_burst(self, steps):
        self.ctx.check(lib().snl_set_march_burst(self.h, steps))

    def mega_variants(self):
        """2 when the megakernel has a four-wave build that spills beside the default"""
        n = lib().snl_mega_variants(self.h)
        self.ctx.check(min(n, 0))
        return n

    def mega_variant(self, variant):
        self.ctx.check(lib().snl_set_mega_variant(self.h, variant))

    def engine(self, engine, tiles=0, persist_blocks=0, shade_thresh=0, chunk=0):
        self.ctx.check(lib().snl_set_engine(self.h, engine, tiles, persist_blocks, shade_thresh, chunk))
        self.engine_id = engine

    def profiling(self, on):
        self.ctx.check(lib().snl_set_profiling(self.h, int(on)))

    def close(self):
        if self.h:
            lib().snl_program_destroy(self.h)
            self.h = None
