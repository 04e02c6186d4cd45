"""Autotuner for the engine and launch shape of pixel-queue renders.

Knobs: the engine (the persistent state machine or the megakernel, which
take pixels from the same queues), the megakernel's build where it has two
(four waves per SIMD with spills, or fewer waves without), pixel tiles
pipelined on separate streams,
resident blocks per launch
(waves per CU = 4 x blocks / CUs), the shading threshold (lanes of a wave that
must wait before the wave runs the shading code), the number of pixels a
wave takes from its tile's queue per atomic and the march steps a lane takes
between two checks of the shading gate.  The tuner times short renders of
the scene being rendered (the state is re-seeded afterwards by the caller) and
keeps the fastest shape; coordinate descent from the defaults keeps the number
of trial renders small.
"""
import itertools
import time

from . import native


def time_render(sess, frames):
    sess.ctx.sync()
    t = time.perf_counter()
    sess.prog.render(sess.state, frames)
    sess.ctx.sync()
    return time.perf_counter() - t


def tune(sess, frames=4, cu=256, verbose=False, engines=(native.ENGINE_PERSISTENT, native.ENGINE_MEGAKERNEL),
         tiles=(1, 2, 4), variants=(0, 1)):
    """coordinate descent over the knobs; `tiles` (1) for jobs rendered by
    snl_render_batch, which launches one kernel per job; `variants` the
    megakernel builds to consider where the program has two (a single value
    fixes it)"""
    prog = sess.prog
    space = {
        "engine": list(engines),
        "tiles": list(tiles),
        "blocks": [cu * 2, cu * 3, cu * 4, cu * 6, cu * 8],
        "thresh": [8, 16, 24, 32, 48],
        "chunk": [32, 64, 128, 256],
        "burst": [1, 2, 4, 8, 16],
    }
    best = {"engine": engines[0], "tiles": 1, "blocks": cu * 6, "thresh": 32, "chunk": 32, "burst": 4}
    if native.ENGINE_MEGAKERNEL in engines and prog.mega_variants() == 2:
        space["variant"] = list(variants)
        best["variant"] = variants[0]

    def run(cfg):
        prog.engine(cfg["engine"], cfg["tiles"], cfg["blocks"], cfg["thresh"], cfg["chunk"])
        prog.march_burst(cfg["burst"])
        if "variant" in cfg:
            prog.mega_variant(cfg["variant"])
        return time_render(sess, frames)

    run(best)  # warm the code object and caches
    best_t = run(best)
    trials = [(dict(best), best_t)]
    for _ in range(2):
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cfg = dict(best, **{knob: v})
                t = run(cfg)
                trials.append((cfg, t))
                if verbose:
                    print("autotune", cfg, "%.2f ms" % (t * 1e3), flush=True)
                if t < best_t:
                    best, best_t = cfg, t
    if "variant" in best and best["engine"] == native.ENGINE_MEGAKERNEL and len(variants) == 2:
        # the two megakernel builds differ by a few per cent, about the spread
        # of one short trial: decide between them on the best of three each
        other = dict(best, variant=1 - best["variant"])
        tb = min(run(best) for _ in range(3))
        to = min(run(other) for _ in range(3))
        trials += [(dict(best), tb), (other, to)]
        if to < tb:
            best, best_t = other, to
    prog.engine(best["engine"], best["tiles"], best["blocks"], best["thresh"], best["chunk"])
    prog.march_burst(best["burst"])
    if "variant" in best:
        prog.mega_variant(best["variant"])
    return best, best_t, trials


def pick_engine(sess, frames=4, engines=(native.ENGINE_MEGAKERNEL, native.ENGINE_PERSISTENT)):
    """the engine alone (default launch shape) that renders `frames` frames of
    the session's job fastest; set on its program and returned with its time.  For the many
    jobs of a batch, where a full tune() per job would cost more than the
    batch: heavy fractal metals favour the persistent kernel, most other jobs
    the megakernel.  The caller re-seeds the state afterwards."""
    best, best_t = engines[0], None
    for e in engines:
        sess.prog.engine(e)
        time_render(sess, 1)  # load the kernel, warm the caches
        t = time_render(sess, frames)
        if best_t is None or t < best_t:
            best, best_t = e, t
    sess.prog.engine(best)
    return best, best_t


ENGINE_NAMES = {native.ENGINE_WAVEFRONT: "wavefront", native.ENGINE_PERSISTENT: "persistent",
                native.ENGINE_MEGAKERNEL: "megakernel"}
KERNEL_NAMES = {native.ENGINE_PERSISTENT: "snl_persist", native.ENGINE_MEGAKERNEL: "snl_mega"}
