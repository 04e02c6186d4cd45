"""Served rendering (snelly_amd/serve.py) and its GPU-side error estimate."""
import math

import numpy as np
import pytest

from snelly_amd import host, native, scenes, serve


class FakeJob:
    def __init__(self, frames, error, target):
        self.frames, self.error, self.target = frames, error, target


def test_quantum_follows_error_trend():
    s = serve.Server.__new__(serve.Server)
    s.min_frames, s.quantum = 4, (2, 64)
    assert s._quantum(FakeJob(0, math.inf, 0.1)) == 4          # up to the first estimate
    assert s._quantum(FakeJob(8, 0.2, 0.1)) == 24              # 8 * (0.2/0.1)^2 - 8
    assert s._quantum(FakeJob(8, 0.101, 0.1)) == 2             # nearly there: the smallest quantum
    assert s._quantum(FakeJob(64, 1.0, 0.1)) == 64             # capped


def test_served_workload():
    a = scenes.served_jobs(50, seed=3)
    b = scenes.served_jobs(50, seed=3)
    assert [x[0] for x in a] == [x[0] for x in b]
    times = np.array([x[0] for x in a])
    assert np.all(np.diff(times) > 0)
    assert all(64 <= s.width <= 512 and 0.05 <= t <= 0.2 for _, s, t in a)


def test_report_percentiles():
    class J:
        def __init__(self, a, f):
            self.arrival, self.finished = a, f
    r = serve.Server.report([J(0.0, float(k)) for k in range(1, 101)], 10.0, 1e9)
    assert r["jobs_done"] == 100 and r["p50_s"] == 50.5 and abs(r["p99_s"] - 99.01) < 1e-6
    assert r["path_samples_per_s"] == 1e8


@pytest.mark.gpu
def test_state_compare_sums(ctx):
    W, H = 19, 7
    rs = np.random.RandomState(0)
    a, b = native.State(ctx, W, H, 1), native.State(ctx, W, H, 2)
    ra = rs.uniform(0, 2, (W * H, 4)).astype(np.float32)
    rb = rs.uniform(0, 2, (W * H, 4)).astype(np.float32)
    ra[:, 3] = rs.randint(2, 9, W * H)
    rb[:, 3] = rs.randint(2, 9, W * H)
    a.write(0, ra)
    b.write(0, rb)
    d2, m2 = native.compare_states(a, b)
    ua = ra[:, :3].astype(np.float64) * (ra[:, 3:] / (ra[:, 3:] - 1.0))
    ub = rb[:, :3].astype(np.float64) * (rb[:, 3:] / (rb[:, 3:] - 1.0))
    assert d2 == pytest.approx(((ua - ub) ** 2).sum(), rel=1e-12)
    assert m2 == pytest.approx((((ua + ub) / 2) ** 2).sum(), rel=1e-12)
    a.close()
    b.close()


@pytest.mark.gpu
def test_served_jobs_finish_and_match_alone(ctx):
    """a few small jobs served together reach their targets, and each job's
    images are what the job renders alone with the same seeds and frames"""
    arr = scenes.served_jobs(6, seed=11, rate=200.0, min_side=24, max_side=48, target=(0.2, 0.35))
    srv = serve.Server(ctx, flags=native.BUILD_PRECISE, reserve=1 << 30, keep_images=True)
    done, left = srv.run(arr, time_limit=120.0)
    assert not left and len(done) == 6
    for job in done:
        assert job.error <= job.target and job.frames >= srv.min_frames
        s = job.scene
        seed = 1000003 * (job.id + 1)
        alone = host.Session(s, s.width, s.height, ctx, seed=seed, flags=native.BUILD_PRECISE)
        alone.render(job.frames)
        b = native.State(ctx, s.width, s.height, seed ^ 0x5EED5EED)
        alone.prog.render(b, job.frames)
        assert np.array_equal(alone.radiance(), job.images[0]), s.name
        assert np.array_equal(b.read(0), job.images[1]), s.name
        b.close()
        alone.close()
