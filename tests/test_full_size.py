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


def test_menger_metal_1024_64spp_strided_pixels(ctx):
    """the benchmark scene at its size: every 1021st pixel (1027 of 1048576)"""
    W = H = 1024
    s = scenes.menger_metal(W, H, 64)
    sess, g, gr = gpu_render(ctx, s, W, H, 64)
    sess.close()
    _, r, rr = oracle_subset(s, W, H, 64, 1021, 7)
    idx = np.arange(7, W * H, 1021)
    assert np.array_equal(g[idx, 3], r[idx, 3])
    assert close(g[idx], r[idx]).mean() >= 0.95
    assert np.all(gr[idx] == rr[idx], axis=1).mean() >= 0.95


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

