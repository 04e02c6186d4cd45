"""Served rendering (snelly_amd/serve.py) and its GPU-side error estimate."""
import math

import numpy as np
import pytest

from snelly_amd import host, native, scenes, serve


class FakeJob:
    def __init__(self, frames, error, target, last_tick=-1, arrival=0.0):
        self.frames, self.error, self.target = frames, error, target
        self.last_tick, self.arrival = last_tick, arrival


def test_quantum_follows_error_trend():
    s = serve.Server.__new__(serve.Server)
    s.min_frames, s.quantum = 4, (2, 64)
    assert s._quantum(FakeJob(0, math.inf, 0.1)) == 4          # up to the first estimate
    assert s._quantum(FakeJob(8, 0.2, 0.1)) == 24              # 8 * (0.2/0.1)^2 - 8
    assert s._quantum(FakeJob(8, 0.101, 0.1)) == 2             # nearly there: the smallest quantum
    assert s._quantum(FakeJob(64, 1.0, 0.1)) == 64             # capped


def test_plan_budget_goes_to_least_remaining_work():
    s = serve.Server.__new__(serve.Server)
    s.min_frames, s.quantum, s.tick_samples = 4, (2, 256), 1000
    new = FakeJob(0, math.inf, 0.1)          # before its first estimate: its min_frames
    near = FakeJob(8, 0.11, 0.1)             # wants 2 frames (8 * 1.21 - 8 -> 2)
    cheap = FakeJob(100, 0.2, 0.1)           # wants 256 (capped); remaining 300 x 1 samples
    costly = FakeJob(100, 0.2, 0.1)          # same frames, 10 samples per frame: remaining 3000
    jobs, spf = [new, near, cheap, costly], [10, 10, 1, 10]
    q = s._plan(jobs, spf)
    # first quanta 4 + 2 + 2 + 2 frames (82 samples of the 1000); cheap takes 254
    # more (254 samples), costly the rest (664 // 10 = 66)
    assert q == [4, 2, 256, 68]
    s.tick_samples = None
    assert s._plan(jobs, spf) == [4, 2, 256, 256]


def test_plan_budget_bounds_the_tick_round_robin():
    """a budget below the jobs' smallest quanta: the longest-unserved jobs go
    first, at least one per tick, and the others wait for a later tick"""
    s = serve.Server.__new__(serve.Server)
    s.min_frames, s.quantum, s.tick_samples = 4, (2, 64), 45
    a = FakeJob(8, 0.5, 0.1, last_tick=3, arrival=0.0)
    b = FakeJob(8, 0.5, 0.1, last_tick=1, arrival=1.0)
    c = FakeJob(8, 0.5, 0.1, last_tick=1, arrival=0.5)
    assert s._plan([a, b, c], [10, 10, 10]) == [0, 2, 2]
    s.tick_samples = 5   # smaller than any quantum: still one job
    assert s._plan([a, b, c], [10, 10, 10]) == [0, 0, 2]


def test_served_workload():
    """config 5: Poisson arrivals at 20/s, sides log-uniform 64..2048, target
    relative RMSE uniform in [0.5%, 5%], scenes of the batched families"""
    a = scenes.served_jobs(400, seed=3)
    b = scenes.served_jobs(400, seed=3)
    assert [x[0] for x in a] == [x[0] for x in b]
    times = np.array([x[0] for x in a])
    assert np.all(np.diff(times) > 0)
    assert abs(len(a) / times[-1] - 20.0) < 3.0
    sides = np.array([s.width for _, s, _ in a])
    assert sides.min() >= 64 and sides.max() <= 2048 and sides.max() > 1500 and sides.min() < 90
    assert all(s.width == s.height for _, s, _ in a)
    assert all(0.005 <= t <= 0.05 for _, _, t in a)


class J:
    def __init__(self, a, f=None):
        self.arrival, self.finished = a, f


def test_report_percentiles():
    r = serve.Server.report([J(0.0, float(k)) for k in range(1, 101)], [], 10.0, 1e9)
    assert r["jobs_done"] == 100 and r["jobs_unfinished"] == 0
    assert r["p50_s"] == 50.0 and r["p99_s"] == 99.0
    assert not r["p50_censored"] and not r["p99_censored"]
    assert r["path_samples_per_s"] == 1e8


def test_report_counts_unfinished_jobs_as_censored():
    """jobs still running at the cutoff enter the percentiles at cutoff -
    arrival, and a percentile they can reach is flagged as a lower bound"""
    done = [J(0.0, float(k)) for k in range(1, 91)]       # latencies 1..90 s
    left = [J(5.0 * k) for k in range(10)]                # arrived at 0..45 s, unfinished
    r = serve.Server.report(done, left, 100.0, 1e9, cutoff_s=100.0)
    assert r["jobs_done"] == 90 and r["jobs_unfinished"] == 10
    # censored latencies 100, 95, ..., 55: p99 is the 99th of 100 -> 95 s, a lower bound
    assert r["p99_s"] == 95.0 and r["p99_censored"]
    # p50: the 50th value (50 s) -- censored jobs start at 55 s, so it is exact
    assert r["p50_s"] == 50.0 and not r["p50_censored"]
    # jobs arriving after the cutoff are not counted
    r2 = serve.Server.report(done, left + [J(150.0)], 100.0, 1e9, cutoff_s=100.0)
    assert r2["jobs_unfinished"] == 10


def test_chains_for_small_images():
    assert serve.Job.chains_for(64 * 64) == 16
    assert serve.Job.chains_for(512 * 512) == 2
    assert serve.Job.chains_for(8 * 8) == 32


@pytest.mark.gpu
def test_state_error_sums(ctx):
    W, H, K = 19, 7, 3
    rs = np.random.RandomState(0)
    st = native.State(ctx, W, H, 1, chains=K)
    r = rs.uniform(0, 2, (K, H, W, 4)).astype(np.float32)
    r[..., 3] = rs.randint(2, 9, (K, H, W))
    st.write(0, r)
    assert np.array_equal(st.read(0), r)
    v2, m2 = native.state_error(st)
    f = r.astype(np.float64)
    u = f[..., :3] * (f[..., 3:] / (f[..., 3:] - 1.0))
    m = u.mean(axis=0)
    assert v2 == pytest.approx((u.var(axis=0, ddof=1) / K).sum(), rel=1e-9)
    assert m2 == pytest.approx((m ** 2).sum(), rel=1e-12)
    st.close()


@pytest.mark.gpu
def test_chain_zero_is_the_single_chain_render(ctx):
    """chain 0 of a multi-chain state starts from the same seeds as a plain
    state and renders the same image: bit for bit"""
    s = scenes.batch_job("menger3", "METAL", 40, 3, 5, np.random.RandomState(2))
    a = host.Session(s, 40, 40, ctx, seed=77, flags=native.BUILD_PRECISE, chains=4)
    b = host.Session(s, 40, 40, ctx, seed=77, flags=native.BUILD_PRECISE)
    a.render(3)
    b.render(3)
    ra = a.state.read(0)
    assert ra.shape == (4, 40, 40, 4)
    assert np.array_equal(ra[0], b.state.read(0))
    assert not np.array_equal(ra[1], ra[0])
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tick_samples", [None, 1 << 16], ids=["no_budget", "budget"])
def test_served_jobs_finish_and_match_alone(ctx, tick_samples):
    """a few small jobs served together reach their targets, and each job's
    images are what the job renders alone with the same seeds and frames"""
    arr = scenes.served_jobs(6, seed=11, rate=200.0, min_side=24, max_side=48, target=(0.2, 0.35))
    srv = serve.Server(ctx, flags=native.BUILD_PRECISE, reserve=1 << 30, keep_images=True, quantum=(2, 256),
                       tick_samples=tick_samples)
    done, left = srv.run(arr, time_limit=120.0)
    assert not left and len(done) == 6
    for job in done:
        assert job.error <= job.target and job.frames >= srv.min_frames
        s = job.scene
        alone = host.Session(s, s.width, s.height, ctx, seed=1000003 * (job.id + 1), flags=native.BUILD_PRECISE,
                             chains=job.chains)
        alone.render(job.frames)
        assert np.array_equal(alone.radiance(), job.image), s.name
        alone.close()


@pytest.mark.gpu
def test_served_2048_job_matches_alone(ctx):
    """a 2048 x 2048 job arriving among small ones, all submitted without
    blocking: every job's image equals the job rendered alone with the same
    seeds and frames"""
    rs = np.random.RandomState(5)
    big = scenes.batch_job("sphere", "SURFACE", 2048, 0, 4, rs)
    arr = [(0.0, scenes.batch_job("menger2", "METAL", 48, 0, 4, rs), 0.3),
           (0.0, big, 0.3),
           (0.01, scenes.batch_job("torus", "DIELECTRIC", 40, 0, 5, rs), 0.3),
           (0.02, scenes.batch_job("menger3", "SURFACE", 64, 0, 4, rs), 0.3)]
    srv = serve.Server(ctx, flags=native.BUILD_PRECISE, reserve=1 << 30, keep_images=True, quantum=(2, 16),
                       tick_samples=1 << 22)
    done, left = srv.run(arr, time_limit=120.0)
    assert not left and len(done) == 4
    assert any(j.scene.width == 2048 for j in done)
    for job in done:
        s = job.scene
        alone = host.Session(s, s.width, s.height, ctx, seed=1000003 * (job.id + 1), flags=native.BUILD_PRECISE,
                             chains=job.chains)
        alone.render(job.frames)
        assert np.array_equal(alone.radiance(), job.image), s.name
        alone.close()


@pytest.mark.gpu
def test_served_jobs_in_one_launch_match_alone(ctx):
    """with the batch module of the jobs' program kinds loaded, the quanta of a
    submission run in one persistent launch (fewer launches than quanta), and
    every job's image -- multi-chain states included -- equals the job
    rendered alone with the same seeds and frames"""
    # the kinds of scenes.batch_sample(), whose precise-build module is precompiled
    rs = np.random.RandomState(9)
    arr = [(0.002 * k, scenes.batch_job(fam, scenes.BATCH_MATERIALS[k % 3], side, 0, 5 + k % 4, rs), 0.3)
           for k, (fam, side) in enumerate(zip(scenes.BATCH_FAMILIES, [40, 24, 64, 32, 48, 56, 28]))]
    srv = serve.Server(ctx, flags=native.BUILD_PRECISE, reserve=1 << 30, keep_images=True, quantum=(2, 16),
                       module_scenes=scenes.batch_sample())
    done, left = srv.run(arr, time_limit=120.0)
    assert not left and len(done) == len(arr)
    assert srv.launches < srv.quanta
    for job in done:
        s = job.scene
        alone = host.Session(s, s.width, s.height, ctx, seed=1000003 * (job.id + 1), flags=native.BUILD_PRECISE,
                             chains=job.chains)
        alone.render(job.frames)
        assert np.array_equal(alone.radiance(), job.image), s.name
        alone.close()
