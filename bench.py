"""Benchmark: path samples/s of Snelly's spectral SDF path tracer on one MI355X.

Workload (one step): the Menger-sponge metal scene -- 4 iterations, half-size 1,
analytic complex IOR n = 0.2 + 0.8 t, k = 3.5 - 1.5 t, Beckmann roughness 0.05,
spectral (hero wavelength), 8 bounces -- rendered at 1024 x 1024 with 64 path
samples per pixel from a freshly seeded accumulation state, i.e. 67,108,864
path samples.  Scene data are synthetic (an analytic SDF); nothing is read from
disk.  Prints one JSON line.

Also measured: the 64-scene batched workload rendered plain (one launch per
scene and frame) and as one batch ("batched"), and a served run -- jobs
arriving at 20/s, rendered to a target error by continuous batching, with
p50/p99 time from arrival to target ("served").

  python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--cpu-seconds 15] [--no-batch] [--served-jobs 200]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# hardware queues of this process: the batch runs up to this many jobs'
# launches at once (set before the HIP runtime starts)
os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("SNL_BENCH_HW_QUEUES", "16")

METRIC = "path samples/s (plain vs batched/served) at RMSE≤1%; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per pixel per render of the queue kernels, from the FETCH_SIZE /
# WRITE_SIZE counter runs scaled by the copy yardstick
# (profiles/<kernel>_menger4_1024_traffic_valu.md): the persistent kernel's one
# launch, the megakernel's two (frame 1 with the tile-cost atomics, then the
# other frames in cost order)
TRAFFIC_B_PER_PX = {"snl_persist": 67.1, "snl_mega": 144.3}
# served: path samples per tick beyond each job's smallest quantum, given to the
# jobs with the least remaining work first; chains per job: the fewest (up to
# 128) that give it 2^18 lanes
TICK_SAMPLES = 1 << 24
CHAIN_LANES, MAX_CHAINS = 1 << 18, 128


def unbiased(rad):
    rad = rad.reshape(-1, 4)
    return rad[:, :3] * (rad[:, 3:4] / (rad[:, 3:4] - 1.0))


def cpu_baseline(scene, width, height, spp, budget_s):
    """The CPU oracle (a C++ restatement of the reference shaders, OpenMP) on a
    bounded, strided sample of the same workload's pixels."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    o = oracle_py.Oracle(scene, width, height)
    npix = width * height
    # calibrate on a sparse sample, then size the stride to the time budget
    stride = 4096
    rad, rng = o.new_state(99)
    t = time.perf_counter()
    o.render(rad, rng, spp, stride, 0, threads)
    dt = max(time.perf_counter() - t, 1e-3)
    per_pixel = dt / math.ceil(npix / stride)
    stride = max(1, int(math.ceil(npix * per_pixel / budget_s)))
    rad, rng = o.new_state(99)
    t = time.perf_counter()
    o.render(rad, rng, spp, stride, 0, threads)
    dt = time.perf_counter() - t
    pixels = math.ceil(npix / stride)
    return {"value": pixels * spp / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"every {stride}th pixel of the {width}x{height} image ({pixels} pixels) x {spp} spp, "
                      f"{dt:.1f} s on {threads} threads"}


def batched(ctx, steps, warmup):
    """The batched workload (64 seeded scenes, scenes.batch_jobs): the same jobs
    rendered plain -- one scene per launch and one launch per frame, job after
    job, as the reference draws -- and as one snl_render_batch call."""
    from snelly_amd import autotune, host, native, scenes
    jobs = scenes.batch_jobs()
    sessions = [host.Session(s, s.width, s.height, ctx, seed=500 + j, flags=native.BUILD_FAST)
                for j, s in enumerate(jobs)]
    # each job's engine from a short trial of both alone (untimed); the plain
    # and the batched runs then use the same programs and engines.  (Tuning the
    # whole launch shape of the heaviest jobs as well sped the plain run up more
    # than the batch.)
    for j, (se, s) in enumerate(zip(sessions, jobs)):
        autotune.pick_engine(se, min(4, s.spp))
    chosen = {}
    for j, se in enumerate(sessions):
        name = autotune.ENGINE_NAMES[se.prog.engine_id]
        chosen[name] = chosen.get(name, 0) + 1
        se.state.seed(500 + j)
    samples = sum(s.width * s.height * s.spp for s in jobs)
    ctx.sync()
    t = time.perf_counter()
    for se, s in zip(sessions, jobs):
        for _ in range(s.spp):
            se.render(1)
    ctx.sync()
    plain_s = time.perf_counter() - t
    batch = [(se.prog, se.state, s.spp) for se, s in zip(sessions, jobs)]
    for w in range(warmup):
        native.render_batch(batch)
    times, gpu_ms = [], []
    for k in range(steps):
        for j, se in enumerate(sessions):
            se.state.seed(900 + 64 * k + j)
        ctx.sync()
        t = time.perf_counter()
        st = native.render_batch(batch)
        ctx.sync()
        times.append(time.perf_counter() - t)
        gpu_ms.append(st.ms_total)
    for se in sessions:
        se.close()
    batch_s = sum(times) / len(times)
    fams = {}
    for s in jobs:
        fams[s.name] = fams.get(s.name, 0) + 1
    return {"workload": "batch64_mixed_sdf", "jobs": len(jobs), "path_samples_per_batch": samples,
            "distinct_programs": len({(s.glsl, s.name) for s in jobs}),
            "plain_path_samples_per_s": round(samples / plain_s, 1), "plain_s": round(plain_s, 3),
            "batched_path_samples_per_s": round(samples / batch_s, 1), "batched_s": round(batch_s, 3),
            "batched_gpu_ms": round(sum(gpu_ms) / len(gpu_ms), 2), "speedup_vs_plain": round(plain_s / batch_s, 2),
            "build": "fast", "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), "engines": chosen,
            "jobs_by_scene": fams}


def served(ctx, n_jobs):
    """The served workload (scenes.served_jobs): jobs arriving as a Poisson
    process at 20/s, each rendered until its estimated relative RMS error
    reaches its target (10-25%), by the continuous-batching server."""
    from snelly_amd import scenes, serve
    arrivals = scenes.served_jobs(n_jobs, target=(0.10, 0.25))
    srv = serve.Server(ctx, quantum=(2, 256), tick_samples=TICK_SAMPLES, chain_lanes=CHAIN_LANES,
                       max_chains=MAX_CHAINS)
    t = time.perf_counter()
    done, left = srv.run(arrivals, time_limit=90.0)
    wall = time.perf_counter() - t
    rep = serve.Server.report(done, wall, srv.samples)
    rep.update({"workload": "served_poisson20_mixed_sdf", "jobs": len(arrivals), "unfinished": len(left),
                "unfinished_scenes": sorted({j.scene.name for j in left}),
                "quantum_frames": [2, 256], "tick_samples": TICK_SAMPLES, "chain_lanes": CHAIN_LANES,
                "max_chains": MAX_CHAINS,
                "arrival_rate_per_s": 20.0, "sides": [64, 512], "target_rrmse": [0.10, 0.25], "wall_s": round(wall, 2),
                "ticks": srv.ticks, "gpu_busy_frac": round(srv.gpu_ms * 1e-3 / wall, 3),
                "error_estimate": "spread of 2-128 independently seeded chains per job, reduced on the GPU"})
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--served-jobs", type=int, default=200)
    ap.add_argument("--engine", choices=["auto", "persistent", "megakernel"], default="auto",
                    help="pixel-queue engine of the timed renders (auto: the faster one, tuned)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path is per-image data parallel "
              f"and is run on a single MI355X here", file=sys.stderr)
        sys.exit(2)

    from snelly_amd import autotune, host, native, scenes

    W, H, spp = args.width, args.height, args.spp
    scene = scenes.menger_metal(W, H, spp)
    ctx = native.Context(0)
    # FAST: FMA contraction and 1-ulp divide/sqrt outside the RNG (parity of this
    # build with the oracle is statistical: tests/test_gpu_parity.py)
    sess = host.Session(scene, W, H, ctx, flags=native.BUILD_FAST)
    # the plain wavefront engine (one launch sequence per bounce), for comparison
    sess.prog.engine(native.ENGINE_WAVEFRONT)
    sess.prog.profiling(True)
    sess.state.seed(7)
    sess.render(spp)
    wf = sess.prog.stats()
    wavefront = {"path_samples_per_s": round(W * H * spp / (wf.ms_total * 1e-3), 1), "ms": round(wf.ms_total, 2),
                 "trace_ms": round(wf.ms_trace, 2),
                 "trace_tflops": round(wf.trace_steps * scene.flops_per_step / (wf.ms_trace * 1e-3) / 1e12, 3)}
    # the megakernel (cameraPath per lane, rays traced in place), for comparison
    sess.prog.engine(native.ENGINE_MEGAKERNEL)
    sess.state.seed(7)
    sess.render(spp)
    mk = sess.prog.stats()
    megakernel = {"path_samples_per_s": round(W * H * spp / (mk.ms_total * 1e-3), 1), "ms": round(mk.ms_total, 2)}
    # the pixel-queue engine (persistent state machine or megakernel) and its
    # launch shape, autotuned untimed
    engines = {"auto": (native.ENGINE_PERSISTENT, native.ENGINE_MEGAKERNEL),
               "persistent": (native.ENGINE_PERSISTENT,), "megakernel": (native.ENGINE_MEGAKERNEL,)}[args.engine]
    best, _, _ = autotune.tune(sess, frames=8, engines=engines)
    for w in range(args.warmup):
        sess.state.seed(1000 + w)
        sess.render(spp)
    ctx.sync()
    images = []
    trace_ms = 0.0
    trace_steps = 0
    normal_evals = 0
    launches = 0
    step_ms = []
    elapsed = 0.0
    for k in range(args.steps):
        ctx.sync()
        ts = time.perf_counter()
        sess.state.seed(1234 + k)
        sess.render(spp)
        ctx.sync()
        dt = time.perf_counter() - ts
        elapsed += dt
        step_ms.append(dt * 1e3)
        st = sess.prog.stats()
        trace_ms += st.ms_trace
        trace_steps += st.trace_steps
        normal_evals += st.normal_evals
        launches += st.trace_launches
        if k < 2:  # read back outside the timed region
            images.append(unbiased(sess.radiance()))
    samples = W * H * spp * args.steps
    value = samples / elapsed

    # the queue kernel is the whole render; its launches on the tile streams overlap,
    # so the rate is total algorithmic FLOPs over the GPU time of the renders.
    # Counted: every SDF march step (SDF + 8 for the step itself) and every
    # central-difference normal (6 SDF evaluations); shading is not counted.
    flops = trace_steps * scene.flops_per_step + normal_evals * 6 * scene.sdf_flops
    achieved = flops / (trace_ms * 1e-3) / 1e12
    # per render (one step): a megakernel render is two launches (frame 1, then
    # the other frames in cost order), each reading and writing every pixel once
    renders = args.steps
    per_render = launches / renders
    kname = autotune.KERNEL_NAMES[best["engine"]]
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": (None if TRAFFIC_B_PER_PX[kname] is None else round(TRAFFIC_B_PER_PX[kname] * W * H)),
                "traffic_algorithmic": 64 * W * H * per_render, "per": "render",
                "kernel": kname, "launches_per_render": per_render, "gpu_ms": round(trace_ms, 2),
                "flops_per_render": flops / renders, "flops_per_step": scene.flops_per_step,
                "steps_per_render": trace_steps / renders, "normal_evals_per_render": normal_evals / renders,
                "sdf_flops": scene.sdf_flops,
                "flops_counted": "SDF march steps (SDF + 8) and normals (6 SDF); shading excluded"}
    rr = None
    if len(images) == 2:
        a, b = images
        rr = float(np.linalg.norm(a - b) / math.sqrt(2.0) / np.linalg.norm(0.5 * (a + b)))
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "path samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (analytic SDF scene, PCG-seeded per-pixel RNG)",
        "config": {"workload": "menger4_metal_spectral", "width": W, "height": H, "spp": spp, "bounces": 8,
                   "mode": "plain", "engine": autotune.ENGINE_NAMES[best["engine"]], "build": "fast",
                   "launch_shape": {k: v for k, v in best.items() if k != "engine"},
                   "tile_order": ("cost (frame 1, then the 8x8 tiles costliest first)"
                                  if os.environ.get("SNL_TILE_ORDER", "1") != "0" and best["tiles"] == 1 else "queue"),
                   "path_samples_per_step": W * H * spp},
        "engines": {"wavefront": wavefront, "megakernel": megakernel},
        "roofline": roofline,
        "rrmse_per_step": None if rr is None else round(rr, 5),
        "spp_for_rrmse_1pct": None if rr is None else int(math.ceil(spp * (rr / 0.01) ** 2)),
    }
    sess.close()
    if not args.no_batch:
        out["batched"] = batched(ctx, max(1, min(args.steps, 3)), min(args.warmup, 1))
    if args.served_jobs > 0:
        out["served"] = served(ctx, args.served_jobs)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scenes.menger_metal(W, H, spp), W, H, spp, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
