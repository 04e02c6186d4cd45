"""The three parity scenes at their full sizes: the GPU renders the whole
image (precise build, default engine) and the oracle renders a strided subset
of its pixels from the same seeds (the oracle renders only those), compared
per pixel as in the small-size tests.  The Mandelbox dielectric is chaotic
after a few internal bounces, so it compares first hits per pixel and, for
radiance, the paired per-pixel differences' mean."""
import numpy as np
import pytest

import oracle_py
from snelly_amd import host, native, scenes

pytestmark = pytest.mark.gpu


def gpu_render(ctx, s, W, H, frames):
    sess = host.Session(s, W, H, ctx, flags=native.BUILD_PRECISE)
    sess.render(frames)
    rad = sess.radiance().reshape(-1, 4).copy()
    rng = sess.state.read(1).reshape(-1, 4).copy()
    return sess, rad, rng


def oracle_subset(s, W, H, frames, stride, offset=0):
    o = oracle_py.Oracle(s, W, H)
    rad, rng = o.new_state()
    o.render(rad, rng, frames, stride, offset)
    return o, rad, rng


def close(g, r):
    return np.all(np.abs(g[:, :3] - r[:, :3]) <= 1e-4 * np.maximum(np.abs(r[:, :3]), 1e-3), axis=1)


def test_sphere_plane_256_16spp_every_pixel(ctx):
    """the CPU-runnable configuration: 256 x 256, 16 spp, 4 bounces"""
    W = H = 256
    s = scenes.sphere_plane(W, H, 16)
    sess, g, gr = gpu_render(ctx, s, W, H, 16)
    sess.close()
    _, r, rr = oracle_subset(s, W, H, 16, 1)
    assert np.array_equal(g[:, 3], r[:, 3])
    assert close(g, r).mean() >= 0.99
    assert np.all(gr == rr, axis=1).mean() >= 0.99


_MENGER = {}


def menger_oracle():
    """config 2's oracle render of every 1021st pixel (1027 of 1048576), from
    the sessions' default seed, computed once for the tests below"""
    if not _MENGER:
        W = H = 1024
        s = scenes.menger_metal(W, H, 64)
        _, r, rr = oracle_subset(s, W, H, 64, 1021, 7)
        _MENGER.update(scene=s, rad=r, rng=rr, idx=np.arange(7, W * H, 1021))
    return _MENGER


def unbiased(rad):
    """the running mean's radiance estimate: (X, Y, Z) N / (N - 1)"""
    return rad[:, :3].astype(np.float64) * (rad[:, 3:] / (rad[:, 3:] - 1.0))


def test_menger_metal_1024_64spp_strided_pixels(ctx):
    """the benchmark scene at its size in the precise build: every 1021st pixel"""
    W = H = 1024
    m = menger_oracle()
    sess, g, gr = gpu_render(ctx, m["scene"], W, H, 64)
    sess.close()
    idx, r, rr = m["idx"], m["rad"], m["rng"]
    assert np.array_equal(g[idx, 3], r[idx, 3])
    assert close(g[idx], r[idx]).mean() >= 0.95
    assert np.all(gr[idx] == rr[idx], axis=1).mean() >= 0.95


def test_menger_metal_1024_64spp_fast_build_strided_pixels(ctx):
    """the build the benchmark times (SNL_BUILD_FAST: contraction, 1-ulp
    reciprocals in the scene code, subnormals flushed) at config 2's size, on
    the same pixels.  Roundings differ, so a path that grazes an edge can take
    another branch; required: the same sample counts, no systematic
    difference (every channel's mean paired difference within 4 standard
    errors) and a paired difference well below that of an independent-seed
    render -- the same estimator, mostly the same paths"""
    W = H = 1024
    m = menger_oracle()
    idx, r = m["idx"], m["rad"]
    out = []
    for seed in (1234, 4321):
        sess = host.Session(m["scene"], W, H, ctx, seed=seed, flags=native.BUILD_FAST)
        sess.render(64)
        out.append(sess.radiance().reshape(-1, 4)[idx].copy())
        sess.close()
    g, g2 = out
    assert np.array_equal(g[:, 3], r[idx, 3])
    ug, ug2, ur = unbiased(g), unbiased(g2), unbiased(r[idx])
    d = ug - ur
    se = d.std(axis=0, ddof=1) / np.sqrt(len(idx))
    assert np.all(np.abs(d.mean(axis=0)) <= 4.0 * se + 1e-6 * np.abs(ur).mean(axis=0)), (d.mean(axis=0), se)

    def rrmse(a, b):
        return np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))
    paired, independent = rrmse(ug, ur), rrmse(ug2, ur)
    print(f"fast build, config 2: paired rrmse {paired:.4g}, independent seeds {independent:.4g}, "
          f"close {close(g, r[idx]).mean():.4f}")
    assert paired < 0.2 * independent, (paired, independent)
    assert close(g, r[idx]).mean() >= 0.85


def test_menger_metal_1024_first_hit_every_pixel(ctx):
    """config 2's first-hit passes at 1024 x 1024 on every pixel, in the
    precise build: the pick program's distance bit-identical to the oracle's
    with the same material, and one frame of the normals program within 1e-4
    with identical RNG states"""
    W = H = 1024
    s = scenes.menger_metal(W, H, 64)
    xy = np.stack(np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5), -1).reshape(-1, 2).astype(np.float32)
    sess = host.Session(s, W, H, ctx, flags=native.BUILD_PRECISE)
    pk = sess.prog.pick(xy)
    sess.prog.render_normals(sess.state, 1)
    g = sess.radiance().reshape(-1, 4).copy()
    gr = sess.state.read(1).reshape(-1, 4).copy()
    sess.close()
    o = oracle_py.Oracle(s, W, H)
    rp = o.pick(xy)
    r, rr = o.new_state()
    o.normals(r, rr, 1)
    hit = rp[:, 1] >= 0
    assert 0.3 < hit.mean() < 0.7
    assert np.mean(pk[:, 1] == rp[:, 1]) >= 0.9999
    same = pk[:, 1] == rp[:, 1]
    assert np.mean(pk[:, 0] == rp[:, 0]) >= 0.999
    assert np.all(np.abs(pk[same & hit, 0] - rp[same & hit, 0]) <= 1e-4 * rp[same & hit, 0])
    assert np.array_equal(gr, rr)
    near = np.all(np.abs(g[:, :3] - r[:, :3]) <= 1e-4 * np.maximum(np.abs(r[:, :3]), 1e-3), axis=1)
    print(f"config 2 first hits: pick exact {np.mean(pk[:, 0] == rp[:, 0]):.6f}, normals 1e-4 {near.mean():.6f}")
    assert near.mean() >= 0.999


def test_mandelbox_dielectric_1920x1080_256spp_strided_pixels(ctx):
    """1920 x 1080, 256 spp, 16 bounces: first hits per pixel (pick) on a
    strided subset, and the radiance of every 4099th pixel as paired
    differences whose mean lies within 4 standard errors of zero"""
    W, H = 1920, 1080
    s = scenes.mandelbox_dielectric(W, H, 256)
    sess, g, gr = gpu_render(ctx, s, W, H, 256)
    idx = np.arange(11, W * H, 4099)
    xy = np.stack([idx % W + 0.5, idx // W + 0.5], -1).astype(np.float32)
    pk = sess.prog.pick(xy)
    sess.close()
    o, r, _ = oracle_subset(s, W, H, 256, 4099, 11)
    assert np.array_equal(g[idx, 3], r[idx, 3])
    rp = o.pick(xy)
    same = pk[:, 1] == rp[:, 1]
    assert same.mean() >= 0.995
    assert np.mean(np.abs(pk[same, 0] - rp[same, 0]) <= 1e-4 * np.abs(rp[same, 0])) >= 0.995
    ug = g[idx, :3] * (g[idx, 3:] / (g[idx, 3:] - 1.0))
    ur = r[idx, :3] * (r[idx, 3:] / (r[idx, 3:] - 1.0))
    d = (ug - ur).astype(np.float64)
    se = d.std(axis=0, ddof=1) / np.sqrt(len(idx))
    assert np.all(np.abs(d.mean(axis=0)) <= 4.0 * se + 1e-6 * np.abs(ur).mean(axis=0)), (d.mean(axis=0), se)


def test_mandelbox_volume_1920x1080_256spp_strided_pixels(ctx):
    """config 3 at its size -- 1920 x 1080, 256 spp, 16 bounces, the
    heterogeneous medium in the glass -- in the precise build and in the fast
    build the benchmark times: first hits per pixel on a strided subset, and
    per build the radiance of every 4099th pixel against the oracle as paired
    differences whose mean lies within 4 standard errors of zero"""
    W, H = 1920, 1080
    s = scenes.mandelbox_volume(W, H, 256)
    idx = np.arange(11, W * H, 4099)
    xy = np.stack([idx % W + 0.5, idx // W + 0.5], -1).astype(np.float32)
    o, r, _ = oracle_subset(s, W, H, 256, 4099, 11)
    ur = r[idx, :3] * (r[idx, 3:] / (r[idx, 3:] - 1.0))
    for flags in (native.BUILD_PRECISE, native.BUILD_FAST):
        sess = host.Session(s, W, H, ctx, flags=flags)
        sess.render(256)
        g = sess.radiance().reshape(-1, 4).copy()
        pk = sess.prog.pick(xy)
        sess.close()
        assert np.array_equal(g[idx, 3], r[idx, 3])
        if flags == native.BUILD_PRECISE:
            rp = o.pick(xy)
            same = pk[:, 1] == rp[:, 1]
            assert same.mean() >= 0.995
            assert np.mean(np.abs(pk[same, 0] - rp[same, 0]) <= 1e-4 * np.abs(rp[same, 0])) >= 0.995
        ug = g[idx, :3] * (g[idx, 3:] / (g[idx, 3:] - 1.0))
        d = (ug - ur).astype(np.float64)
        se = d.std(axis=0, ddof=1) / np.sqrt(len(idx))
        print(f"config 3 flags {flags}: paired mean {d.mean(axis=0)} se {se}")
        assert np.all(np.abs(d.mean(axis=0)) <= 4.0 * se + 1e-6 * np.abs(ur).mean(axis=0)), (flags, d.mean(axis=0), se)

