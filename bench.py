"""Benchmark: path samples/s of Snelly's spectral SDF path tracer on one MI355X.

Headline (`value`), config 3: the Mandelbox dielectric (scale 2.5, 12
iterations, bailout 10; Cauchy dispersion n = 1.5 + 0.0042/lambda^2 um, Beer
absorption (0.1, 0.3, 0.5)) with a nested Surface sphere of radius 0.5, in a
heterogeneous medium (density max(0, 1 - |p| + 0.3 fbm3(p)), 5 octaves,
albedo 0.9, Henyey-Greenstein g = 0.3), 1920 x 1080, 256 spp, 16 bounces,
spectral (hero wavelength), rendered from a freshly seeded state: 530,841,600
path samples per step.  Scene data are synthetic (analytic SDFs and noise);
nothing is read from disk.  Prints one JSON line.

Also measured, in the same line:
  secondary  config 2: Menger-4 metal with analytic complex IOR, 1024^2 x 64
             spp, 8 bounces, and its first-hit depth and normal passes;
  batched    config 4: 64 seeded scenes rendered by the plain path (each
             job's render, job after job; and one launch per frame, as the
             reference draws) and as one batch;
  served     config 5: 2000 jobs arriving at 20/s, sides 64..2048, target
             relative RMSE 0.5..5%, served by continuous batching for a fixed
             time; p50/p99 time to target, unfinished jobs censored;
  cpu_baseline  the oracle (the CPU restatement) on the host's cores.

  python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--cpu-seconds 12] [--no-batch]
                  [--served-jobs 2000] [--served-seconds 100] [--no-secondary] [--mega-variant 0|1]
  python bench.py --copy-yardstick [MiB]   a plain device copy, for scaling
                                           HBM traffic counters
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
FP32_PEAK_TFLOPS = 157.3           # MI355X FP32 vector, packed (dual-issue) FP32
# FP32 FMAs without packed math (as every code object here is built) measured
# by bench.py --valu-yardstick: 101.8 TFLOP/s at the 1.95 GHz the card holds
# under that load (0.399 wave64 VALU instructions per cycle per SIMD)
FP32_YARDSTICK_TFLOPS = 101.8
# HBM bytes per pixel per render of the megakernel, from the FETCH_SIZE /
# WRITE_SIZE counter runs scaled by the copy yardstick (profiles/*_traffic*.md):
# two launches per render (frame 1 with the tile-cost atomics, then the other
# frames in cost order), each reading and writing the pixel's radiance and RNG
# (per program build: 0 the default, 1 the four-wave build with spills)
TRAFFIC_B_PER_PX = {("mandelbox_volume", 0): 140.6, ("mandelbox_volume", 1): 150.5, ("menger_metal", 0): 144.3}
# served: path samples per tick (every job's smallest quantum included, in
# round-robin order); chains per job: the fewest (up to 128) that give it 2^18 lanes
TICK_SAMPLES = 1 << 26
CHAIN_LANES, MAX_CHAINS = 1 << 18, 128


def unbiased(rad):
    rad = rad.reshape(-1, 4)
    return rad[:, :3] * (rad[:, 3:4] / (rad[:, 3:4] - 1.0))


def host_threads():
    """every core this process may run on (the box's CPU share), unless
    OMP_NUM_THREADS says otherwise"""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return n or avail, avail, os.cpu_count() or 1


def cpu_baseline(scene, width, height, spp, budget_s):
    """The CPU oracle (a C++ restatement of the reference shaders, OpenMP) on a
    bounded, strided sample of the same workload's pixels, on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    threads, avail, ncpu = host_threads()
    o = oracle_py.Oracle(scene, width, height)
    npix = width * height
    # calibrate on a sparse sample with enough pixels to keep every thread
    # busy, then size the stride to the time budget
    stride = max(1, npix // (256 * threads))
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
            "nproc": avail, "cpu_count": ncpu,
            "sample": f"every {stride}th pixel of the {width}x{height} image ({pixels} pixels) x {spp} spp, "
                      f"{dt:.1f} s on {threads} threads"}


def path_counts(scene, width, height, spp, budget_s):
    """The oracle's counting build (every float operation of the restated
    path counted) over a strided sample of the workload's pixels: FLOPs and
    transcendentals per sphere-tracer SDF step, per normal, and per camera
    path for the rest (camera, RNG, BSDFs, lights, MIS, medium tracking,
    colour).  The GPU's own step, normal and path counts then price a render."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    threads = host_threads()[0]
    o = oracle_py.Oracle(scene, width, height, counting=True)
    npix = width * height
    stride = 16381
    rad, rng = o.new_state(77)
    o.count_reset()
    t = time.perf_counter()
    o.render(rad, rng, spp, stride, 3, threads)
    dt = max(time.perf_counter() - t, 1e-3)
    stride = max(1, int(math.ceil(npix * (dt / math.ceil(npix / stride)) / budget_s)))
    rad, rng = o.new_state(77)
    o.count_reset()
    o.render(rad, rng, spp, stride, 3, threads)
    c = o.counts()
    f, tr = c["flops"], c["transcendentals"]
    paths = max(c["paths"], 1)
    return {"flops_per_step": f["march"] / max(c["steps"], 1),
            "flops_per_normal": f["normal"] / max(c["normals"], 1),
            "other_flops_per_path": (f["shade"] + f["medium"]) / paths,
            "medium_flops_per_path": f["medium"] / paths,
            "trans_per_step": tr["march"] / max(c["steps"], 1),
            "trans_per_normal": tr["normal"] / max(c["normals"], 1),
            "other_trans_per_path": (tr["shade"] + tr["medium"]) / paths,
            "oracle_steps_per_path": c["steps"] / paths, "oracle_normals_per_path": c["normals"] / paths,
            "sample": f"counting build over every {stride}th pixel x {spp} spp ({c['paths']} paths)"}


def roofline(counts, stats, renders, gpu_ms, kname, width, height, traffic_key, variant=0):
    """FP32 VALU roofline of the render kernel: algorithmic FLOPs per render
    (the GPU's sphere-tracer steps, normals and paths priced by the counting
    build) over the kernel's GPU time per render (hipEvents on its stream)"""
    steps, normals, paths = stats["steps"] / renders, stats["normals"] / renders, stats["paths"] / renders
    flops = steps * counts["flops_per_step"] + normals * counts["flops_per_normal"] + \
        paths * counts["other_flops_per_path"]
    trans = steps * counts["trans_per_step"] + normals * counts["trans_per_normal"] + \
        paths * counts["other_trans_per_path"]
    ms = gpu_ms / renders
    achieved = flops / (ms * 1e-3) / 1e12
    per_render = stats["launches"] / renders
    bpp = TRAFFIC_B_PER_PX.get((traffic_key, variant))
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "frac_of_yardstick": round(achieved / FP32_YARDSTICK_TFLOPS, 4),
            "yardstick": "unpacked v_fma_f32 kernel, 101.8 TFLOP/s measured (bench.py --valu-yardstick)",
            "traffic": None if bpp is None else round(bpp * width * height),
            "traffic_algorithmic": 64 * width * height * per_render, "per": "render", "kernel": kname,
            "launches_per_render": per_render, "gpu_ms_per_render": round(ms, 2),
            "flops_per_render": flops, "transcendentals_per_render": trans,
            "steps_per_render": steps, "normal_evals_per_render": normals, "paths_per_render": paths,
            "flops_counted": "every + - * / sqrt of the restated path (counting build of the oracle): "
                             "sphere-tracer steps, normals, camera, RNG, BSDFs, lights, MIS, medium tracking; "
                             "transcendentals reported apart",
            "counts": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in counts.items()}}


def plain(ctx, scene, steps, warmup, engines, tune_frames, variants=(0, 1)):
    """the plain one-scene renderer: the fast build, engine and launch shape
    autotuned (untimed), then `steps` renders of scene.spp frames each from a
    freshly seeded state, timed between device synchronisations"""
    from snelly_amd import autotune, host, native
    W, H, spp = scene.width, scene.height, scene.spp
    sess = host.Session(scene, W, H, ctx, flags=native.BUILD_FAST)
    best, _, _ = autotune.tune(sess, frames=tune_frames, engines=engines, variants=variants)
    for w in range(warmup):
        sess.state.seed(1000 + w)
        sess.render(spp)
    ctx.sync()
    images, elapsed = [], 0.0
    stats = {"steps": 0, "normals": 0, "paths": 0, "launches": 0}
    gpu_ms = 0.0
    for k in range(steps):
        sess.state.seed(1234 + k)
        ctx.sync()
        ts = time.perf_counter()
        sess.render(spp)
        ctx.sync()
        elapsed += time.perf_counter() - ts
        st = sess.prog.stats()
        gpu_ms += st.ms_total
        stats["steps"] += st.trace_steps
        stats["normals"] += st.normal_evals
        stats["paths"] += st.paths
        stats["launches"] += st.trace_launches
        if k < 2:  # read back outside the timed region
            images.append(unbiased(sess.radiance()))
    rr = None
    if len(images) == 2:
        a, b = images
        rr = float(np.linalg.norm(a - b) / math.sqrt(2.0) / np.linalg.norm(0.5 * (a + b)))
    return sess, best, elapsed, gpu_ms, stats, rr


def first_hit_pass(ctx, scene, reps=3):
    """config 2's first-hit depth and normal passes over every pixel: the pick
    program (depth, material) and one frame of the normals program"""
    from snelly_amd import host, native
    W, H = scene.width, scene.height
    sess = host.Session(scene, W, H, ctx, flags=native.BUILD_FAST)
    xy = np.stack(np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5), -1).reshape(-1, 2).astype(np.float32)
    sess.prog.pick(xy)
    sess.prog.render_normals(sess.state, 1)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        sess.prog.pick(xy)
    pick_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        sess.prog.render_normals(sess.state, 1)
    ctx.sync()
    normals_s = (time.perf_counter() - t0) / reps
    sess.close()
    return {"pixels": W * H, "depth_ms": round(pick_s * 1e3, 2), "normals_ms": round(normals_s * 1e3, 2),
            "depth_includes": "pixel positions up and hit distances back over PCIe (16 B/px)"}


def batched(ctx, steps, warmup):
    """config 4 (64 seeded scenes, scenes.batch_jobs): rendered by the plain
    path -- each job's own render of all its frames, job after job, and one
    launch per frame as the reference draws -- and as one snl_render_batch"""
    from snelly_amd import autotune, host, native, scenes
    jobs = scenes.batch_jobs()
    sessions = [host.Session(s, s.width, s.height, ctx, seed=500 + j, flags=native.BUILD_FAST)
                for j, s in enumerate(jobs)]
    # each job's engine from a short trial of both alone (untimed); all legs
    # then use the same programs and engines
    for se, s in zip(sessions, jobs):
        autotune.pick_engine(se, min(4, s.spp))
    chosen = {}
    for se in sessions:
        name = autotune.ENGINE_NAMES[se.prog.engine_id]
        chosen[name] = chosen.get(name, 0) + 1
    samples = sum(s.width * s.height * s.spp for s in jobs)

    def seed_all(base):
        for j, se in enumerate(sessions):
            se.state.seed(base + j)

    seed_all(500)
    ctx.sync()
    t = time.perf_counter()
    for se, s in zip(sessions, jobs):
        se.render(s.spp)
        ctx.sync()
    plain_own_s = time.perf_counter() - t
    seed_all(600)
    ctx.sync()
    t = time.perf_counter()
    for se, s in zip(sessions, jobs):
        for _ in range(s.spp):
            se.render(1)
    ctx.sync()
    plain_frame_s = time.perf_counter() - t
    batch = [(se.prog, se.state, s.spp) for se, s in zip(sessions, jobs)]
    # the batch module of the jobs' programs (compiled ahead into the code-object
    # cache): with fusion on, the megakernel jobs run in one persistent launch
    native.batch_module(ctx, [se.prog for se in sessions])

    def timed(fusion):
        native.set_batch_fusion(ctx, fusion)
        for w in range(warmup):
            native.render_batch(batch)
        times, gpu_ms, launches = [], [], 0
        for k in range(steps):
            seed_all(900 + 64 * k)
            ctx.sync()
            t = time.perf_counter()
            st = native.render_batch(batch)
            ctx.sync()
            times.append(time.perf_counter() - t)
            gpu_ms.append(st.ms_total)
            launches = st.trace_launches
        return sum(times) / len(times), sum(gpu_ms) / len(gpu_ms), launches

    streams_s, streams_ms, streams_launches = timed(False)
    fused_s, fused_ms, fused_launches = timed(True)
    # every job on the megakernel, so that all of them run in the one launch
    tuned = [se.prog.engine_id for se in sessions]
    for se in sessions:
        se.prog.engine(native.ENGINE_MEGAKERNEL)
    all_s, all_ms, all_launches = timed(True)
    # the fastest is the batched figure (and the context's setting)
    modes = {"streams": (streams_s, streams_ms), "one_launch": (fused_s, fused_ms),
             "one_launch_all_megakernel": (all_s, all_ms)}
    mode = min(modes, key=lambda m: modes[m][0])
    if mode != "one_launch_all_megakernel":
        for se, e in zip(sessions, tuned):
            se.prog.engine(e)
    fusion = mode != "streams"
    native.set_batch_fusion(ctx, fusion)
    batch_s, batch_ms = modes[mode]
    # the same launch sequence captured once as a hipGraph and replayed
    graph = native.Graph(batch)
    graph.launch()
    gtimes = []
    for k in range(steps):
        seed_all(900 + 64 * k)
        ctx.sync()
        t = time.perf_counter()
        graph.launch()
        gtimes.append(time.perf_counter() - t)
    graph.close()
    graph_s = sum(gtimes) / len(gtimes)
    gpu_ms = [batch_ms]
    for se in sessions:
        se.close()
    fams = {}
    for s in jobs:
        fams[s.name] = fams.get(s.name, 0) + 1
    return {"workload": "batch64_mixed_sdf (config 4)", "jobs": len(jobs), "path_samples_per_batch": samples,
            "distinct_programs": len({(s.glsl, s.name) for s in jobs}),
            "plain_path_samples_per_s": round(samples / plain_own_s, 1), "plain_s": round(plain_own_s, 3),
            "plain_per_frame_path_samples_per_s": round(samples / plain_frame_s, 1),
            "plain_per_frame_s": round(plain_frame_s, 3),
            "batched_path_samples_per_s": round(samples / batch_s, 1), "batched_s": round(batch_s, 3),
            "batched_mode": mode,
            "streams_s": round(streams_s, 3), "streams_launches": streams_launches,
            "fused_s": round(fused_s, 3), "fused_launches": fused_launches,
            "fused_all_megakernel_s": round(all_s, 3), "fused_all_megakernel_launches": all_launches,
            "graph_replay_s": round(graph_s, 3),
            "graph": "the chosen mode's launch sequence captured as one hipGraph (snl_batch_graph_create) and "
                     "replayed per step",
            "fused": "megakernel jobs in one snl_mega_fused launch, heaviest first; persistent-engine jobs on "
                     "streams beside it",
            "batched_gpu_ms": round(sum(gpu_ms) / len(gpu_ms), 2),
            "speedup_vs_plain": round(plain_own_s / batch_s, 2),
            "speedup_vs_plain_per_frame": round(plain_frame_s / batch_s, 2),
            "plain": "each job's snl_render of all its frames, job after job (tile cost order)",
            "plain_per_frame": "one snl_render(1) per job and frame, as the reference draws one frame per canvas",
            "mandelbox_scales": sorted(round(s.user_uniforms["Scale"], 3) for s in jobs if "Scale" in s.user_uniforms),
            "build": "fast", "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), "engines": chosen,
            "jobs_by_scene": fams}


def served(ctx, n_jobs, seconds, module=True):
    """config 5 (scenes.served_jobs): jobs arriving as a Poisson process at
    20/s, sides log-uniform 64..2048, each rendered until its estimated
    relative RMS error reaches its target (0.5..5%), by the continuous-batching
    server, for `seconds` of arrivals plus a fifth of that to drain; with
    `module`, the server loads the batch module of the workload's 21 program
    kinds"""
    from snelly_amd import scenes, serve
    arrivals = [a for a in scenes.served_jobs(n_jobs) if a[0] <= seconds]
    cutoff = 1.2 * seconds
    srv = serve.Server(ctx, quantum=(2, 256), tick_samples=TICK_SAMPLES, chain_lanes=CHAIN_LANES,
                       max_chains=MAX_CHAINS, module_scenes=scenes.batch_kinds() if module else None)
    t = time.perf_counter()
    done, left = srv.run(arrivals, time_limit=cutoff)
    wall = time.perf_counter() - t
    rep = serve.Server.report(done, left, wall, srv.samples, cutoff_s=cutoff)
    # offered load: the path samples each job needs for its target, from its
    # error trend (samples_needed ~ samples_done * (error / target)^2), over
    # the arrival span; jobs never estimated are left out
    need = 0.0
    for j in done + left:
        if j.frames > 0 and math.isfinite(j.error):
            samples = j.frames * j.pixels * j.chains
            need += samples * max(1.0, (j.error / j.target) ** 2)
    span = max(a[0] for a in arrivals) if arrivals else 1.0
    fams = {}
    for j in done:
        fams[j.scene.name] = fams.get(j.scene.name, 0) + 1
    rep.update({"workload": "served_poisson20 (config 5)", "jobs_arrived": len(arrivals),
                "arrival_rate_per_s": 20.0, "sides": [64, 2048], "target_rrmse": [0.005, 0.05],
                "cutoff_s": round(cutoff, 1), "wall_s": round(wall, 2),
                "gpu_busy_frac": round(srv.busy_s / wall, 3), "submissions": srv.ticks,
                "submission": "snl_batch_submit without blocking; quanta polled with snl_job_poll" +
                              ("; the megakernel quanta of a submission in one persistent launch (batch module "
                               "of the 21 program kinds)" if module else "; each quantum its own launch"),
                "launches_per_quantum": round(srv.launches / max(1, srv.quanta), 3),
                "offered_path_samples_per_s_est": round(need / span, 1),
                "tick_samples": TICK_SAMPLES, "chain_lanes": CHAIN_LANES, "max_chains": MAX_CHAINS,
                "quantum_frames": [2, 256], "done_by_scene": fams,
                "error_estimate": "spread of 2-128 independently seeded chains per job, reduced on the GPU",
                "latency": "arrival to target; unfinished jobs censored at the cutoff"})
    return rep


def copy_yardstick(mib):
    """A plain device copy of known size, the yardstick for HBM traffic
    counters: profiled under the same counters as a render, its FETCH_SIZE /
    WRITE_SIZE per dispatch against the bytes it moves give the counters'
    scale factors."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x 5: each dispatch reads {mib * 1024} KiB and writes {mib * 1024} KiB", flush=True)


VALU_YARDSTICK_SRC = r"""
extern "C" __global__ void __launch_bounds__(256) valu_yardstick(float* out, int iters) {
  // eight independent FMA chains per lane: issue-bound, not latency-bound
  float a0 = threadIdx.x * 1e-3f, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
  float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
  const float m = 0.999999f, c = 1e-7f;
  for (int i = 0; i < iters; ++i) {
    a0 = __builtin_fmaf(a0, m, c); a1 = __builtin_fmaf(a1, m, c); a2 = __builtin_fmaf(a2, m, c);
    a3 = __builtin_fmaf(a3, m, c); a4 = __builtin_fmaf(a4, m, c); a5 = __builtin_fmaf(a5, m, c);
    a6 = __builtin_fmaf(a6, m, c); a7 = __builtin_fmaf(a7, m, c);
  }
  float s = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
  if (s == 12345.678f) out[blockIdx.x] = s;
}
"""


def valu_yardstick(iters=20000, blocks=2048, launches=5):
    """A kernel of known VALU work, the yardstick for the SQ counters: every
    lane runs `iters` x 8 independent v_fma_f32 (built without packed FP32, as
    every kernel here).  Prints its FMA rate; profiled under the SQ counters,
    SQ_INSTS_VALU against the 8 * iters * waves FMAs it issues and
    SQ_BUSY_CYCLES against its time give the counters' meaning and the clock."""
    import ctypes
    rtc = ctypes.CDLL("libhiprtc.so")
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    prog = vp()
    assert rtc.hiprtcCreateProgram(ctypes.byref(prog), VALU_YARDSTICK_SRC.encode(), b"valu.hip", 0, None, None) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    arr = (ctypes.c_char_p * len(opts))(*opts)
    if rtc.hiprtcCompileProgram(prog, len(opts), arr) != 0:
        n = ctypes.c_size_t()
        rtc.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
        log = ctypes.create_string_buffer(n.value + 1)
        rtc.hiprtcGetProgramLog(prog, log)
        raise RuntimeError(log.value.decode())
    n = ctypes.c_size_t()
    rtc.hiprtcGetCodeSize(prog, ctypes.byref(n))
    code = ctypes.create_string_buffer(n.value)
    rtc.hiprtcGetCode(prog, code)
    mod, fn, out, e0, e1 = vp(), vp(), vp(), vp(), vp()
    assert hip.hipModuleLoadData(ctypes.byref(mod), code) == 0
    assert hip.hipModuleGetFunction(ctypes.byref(fn), mod, b"valu_yardstick") == 0
    assert hip.hipMalloc(ctypes.byref(out), ctypes.c_size_t(4 * blocks)) == 0
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    a_out, a_it = vp(out.value), ctypes.c_int(iters)
    args = (vp * 2)(ctypes.addressof(a_out), ctypes.addressof(a_it))
    u = ctypes.c_uint
    launch = lambda: hip.hipModuleLaunchKernel(fn, u(blocks), u(1), u(1), u(256), u(1), u(1), u(0), None, args, None)
    assert launch() == 0 and hip.hipDeviceSynchronize() == 0  # load and warm
    hip.hipEventRecord(e0, None)
    for _ in range(launches):
        assert launch() == 0
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    waves = blocks * 256 // 64
    fma = 8.0 * iters * blocks * 256 * launches
    print(json.dumps({"valu_yardstick": {"iters": iters, "blocks": blocks, "launches": launches,
                                         "waves_per_launch": waves, "fma_wave_instructions_per_launch": 8 * iters * waves,
                                         "ms_per_launch": round(ms.value / launches, 4),
                                         "tflops_fp32": round(2.0 * fma / (ms.value * 1e-3) / 1e12, 2)}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--count-seconds", type=float, default=6.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--served-jobs", type=int, default=2000)
    ap.add_argument("--served-seconds", type=float, default=100.0)
    ap.add_argument("--copy-yardstick", type=int, nargs="?", const=1024, default=None, metavar="MIB")
    ap.add_argument("--batch-only", action="store_true", help="only the batched leg (config 4), as one JSON line")
    ap.add_argument("--valu-yardstick", action="store_true", help="a kernel of known FMA count (SQ counter scale)")
    ap.add_argument("--mega-variant", type=int, choices=(0, 1), default=None,
                    help="fix the headline's megakernel build (0 default, 1 four waves with spills) instead of "
                         "letting the autotuner choose")
    args = ap.parse_args()
    if args.copy_yardstick is not None:
        copy_yardstick(args.copy_yardstick)
        return
    if args.valu_yardstick:
        valu_yardstick()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path is per-image data parallel "
              f"and is run on a single MI355X here", file=sys.stderr)
        sys.exit(2)

    from snelly_amd import autotune, native, scenes

    ctx = native.Context(0)
    if args.batch_only:
        print(json.dumps({"batched": batched(ctx, max(1, min(args.steps, 3)), min(args.warmup, 1))}), flush=True)
        return
    # ---- headline: config 3 (megakernel: the medium's tracking runs inside a vertex)
    scene = scenes.mandelbox_volume(1920, 1080, 256)
    W, H, spp = scene.width, scene.height, scene.spp
    variants = (0, 1) if args.mega_variant is None else (args.mega_variant,)
    sess, best, elapsed, gpu_ms, stats, rr = plain(ctx, scene, args.steps, args.warmup,
                                                   (native.ENGINE_MEGAKERNEL,), 4, variants)
    sess.close()
    samples = W * H * spp * args.steps
    counts = path_counts(scenes.mandelbox_volume(W, H, spp), W, H, spp, args.count_seconds)
    kname = autotune.KERNEL_NAMES[best["engine"]]
    out = {
        "metric": METRIC, "value": round(samples / elapsed, 1), "unit": "path samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (analytic SDFs and value noise, PCG-seeded per-pixel RNG)",
        "config": {"workload": "mandelbox_dielectric_sphere_cloud (config 3)", "width": W, "height": H, "spp": spp,
                   "bounces": 16, "mode": "plain", "engine": autotune.ENGINE_NAMES[best["engine"]],
                   "build": "fast", "launch_shape": {k: v for k, v in best.items() if k != "engine"},
                   "tile_order": "cost (frame 1, then the 8x8 tiles costliest first)",
                   "path_samples_per_step": W * H * spp},
        "roofline": roofline(counts, stats, args.steps, gpu_ms, kname, W, H, "mandelbox_volume",
                             best.get("variant", 0)),
        "rrmse_per_step": None if rr is None else round(rr, 5),
        "spp_for_rrmse_1pct": None if rr is None else int(math.ceil(spp * (rr / 0.01) ** 2)),
    }
    # ---- config 2 and its first-hit pass
    if not args.no_secondary:
        m = scenes.menger_metal(1024, 1024, 64)
        msess, mbest, mel, mgpu, mstats, mrr = plain(ctx, m, 3, 1, (native.ENGINE_PERSISTENT,
                                                                   native.ENGINE_MEGAKERNEL), 8)
        msess.close()
        mcounts = path_counts(scenes.menger_metal(1024, 1024, 64), 1024, 1024, 64, args.count_seconds / 2)
        mk = autotune.KERNEL_NAMES[mbest["engine"]]
        out["secondary"] = {
            "workload": "menger4_metal_spectral (config 2)", "width": 1024, "height": 1024, "spp": 64, "bounces": 8,
            "path_samples_per_s": round(1024 * 1024 * 64 * 3 / mel, 1), "ms_per_step": round(mel * 1e3 / 3, 3),
            "engine": autotune.ENGINE_NAMES[mbest["engine"]],
            "launch_shape": {k: v for k, v in mbest.items() if k != "engine"},
            "roofline": roofline(mcounts, mstats, 3, mgpu, mk, 1024, 1024, "menger_metal", mbest.get("variant", 0)),
            "rrmse_per_step": None if mrr is None else round(mrr, 5),
            "first_hit_pass": first_hit_pass(ctx, scenes.menger_metal(1024, 1024, 1))}
    if not args.no_batch:
        out["batched"] = batched(ctx, max(1, min(args.steps, 3)), min(args.warmup, 1))
    if args.served_jobs > 0:
        out["served"] = served(ctx, args.served_jobs, args.served_seconds)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scenes.mandelbox_volume(W, H, spp), W, H, spp, args.cpu_seconds)
        if not args.no_secondary:
            out["secondary"]["cpu_baseline"] = cpu_baseline(scenes.menger_metal(1024, 1024, 64), 1024, 1024, 64,
                                                            args.cpu_seconds / 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
