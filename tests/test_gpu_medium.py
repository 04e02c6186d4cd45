"""The heterogeneous medium (a scene's SDF_VOLUME) on the GPU against the
oracle, which tests/test_medium.py pins to closed forms and to the reference's
homogeneous absorber.  Precise builds follow the oracle's float operations, so
per pixel the same random decisions are made (delta and ratio tracking draw
from the pixel's RNG in a fixed order) and radiance and RNG states agree; the
Mandelbox scene's glass paths separate after a few internal bounces, so it is
held to the bars of the dielectric scene without the medium."""
import math

import numpy as np
import pytest

import oracle_py
from snelly_amd import host, native, scenes
from test_oracle import unbiased

pytestmark = pytest.mark.gpu


def render_pair(ctx, s, W, H, frames, flags=native.BUILD_PRECISE, seed=1234, masks=False):
    sess = host.Session(s, W, H, ctx, seed=seed, flags=flags)
    sess.render(frames)
    g = sess.radiance().reshape(-1, 4).copy()
    gr = sess.state.read(1).reshape(-1, 4).copy()
    sess.close()
    o = oracle_py.Oracle(s, W, H)
    r, rr = o.new_state(seed)
    o.render(r, rr, frames)
    assert np.array_equal(g[:, 3], r[:, 3])
    close = np.all(np.abs(g[:, :3] - r[:, :3]) <= 1e-4 * np.maximum(np.abs(r[:, :3]), 1e-3), axis=1)
    same = np.all(gr == rr, axis=1)
    if masks:
        return g, r, close, same
    return g, r, close.mean(), same.mean()


@pytest.mark.parametrize("albedo", [0.0, 0.7, 1.0])
def test_medium_ball_matches_oracle(ctx, albedo):
    """the medium alone (absorbing, scattering, white), 32 x 32, 16 spp, RGB"""
    s = scenes.medium_ball(32, 32, 16, albedo=albedo, extinction=2.0)
    g, r, close, same = render_pair(ctx, s, 32, 32, 16)
    print(f"medium_ball albedo {albedo}: close {close:.4f} rng {same:.4f}")
    assert close >= 0.99 and same >= 0.99


def test_medium_ball_spectral_matches_oracle(ctx):
    s = scenes.medium_ball(32, 32, 8, albedo=0.9, extinction=2.0, dispersive=True)
    g, r, close, same = render_pair(ctx, s, 32, 32, 8)
    print(f"medium_ball spectral: close {close:.4f} rng {same:.4f}")
    assert close >= 0.99 and same >= 0.99


def test_absorber_box_matches_oracle(ctx):
    """the grey absorbing box around the sphere and plane: surface shading,
    shadow rays with ratio tracking, collisions that end paths"""
    s = scenes.absorber_box(40, 32, 8, medium=True)
    g, r, close, same = render_pair(ctx, s, 40, 32, 8)
    print(f"absorber_box: close {close:.4f} rng {same:.4f}")
    assert close >= 0.99 and same >= 0.99


def test_mandelbox_volume_matches_oracle(ctx):
    """config 3's scene at 48 x 32, 4 spp, 8 bounces.  Where a pixel's RNG
    state ends equal the GPU and the oracle made the same random decisions
    (free flights, scatters, glass choices), and its radiance must agree; the
    rest are paths sent elsewhere by a last-bit difference in a transcendental,
    whose fireflies make an image-level norm meaningless at 4 spp"""
    s = scenes.mandelbox_volume(48, 32, 4)
    s.renderer.maxBounces = 8
    g, r, close, same = render_pair(ctx, s, 48, 32, 4, masks=True)
    print(f"mandelbox_volume: close {close.mean():.4f} rng {same.mean():.4f} "
          f"close where rng equal {close[same].mean():.4f}")
    assert close.mean() >= 0.9 and same.mean() >= 0.8
    assert close[same].mean() >= 0.99


def test_white_medium_furnace_fast_build(ctx):
    """the fast (timed) build on the albedo-1 medium under a white sky: white
    in expectation"""
    W = H = 32
    s = scenes.medium_ball(W, H, 64, albedo=1.0, extinction=2.0)
    sess = host.Session(s, W, H, ctx, flags=native.BUILD_FAST)
    sess.render(64)
    img = unbiased(sess.radiance())
    sess.close()
    white_y = 0.2126729 + 0.7151522 + 0.0721750
    rel = img[:, 1] / white_y - 1.0
    se = rel.std() / math.sqrt(rel.size)
    print(f"fast furnace: mean {rel.mean():.4g} se {se:.3g}")
    assert abs(rel.mean()) < 4.0 * se + 2e-3


def test_pick_reaches_the_medium_bound(ctx):
    """the pick program's HAS_VOLUME branch (pick-fragment-shader.glsl:74-100):
    the medium's bound is hit as material 3; along the view axis of the ball
    scene at distance 3 (camera at z = 4, bound radius 1)"""
    W = H = 32
    s = scenes.medium_ball(W, H, 1)
    sess = host.Session(s, W, H, ctx, flags=native.BUILD_PRECISE)
    xy = np.stack(np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5), -1).reshape(-1, 2).astype(np.float32)
    g = sess.prog.pick(xy)
    centre = sess.prog.pick(np.array([[W / 2.0, H / 2.0]], np.float32))
    sess.close()
    r = oracle_py.Oracle(s, W, H).pick(xy)
    assert np.array_equal(g[:, 1], r[:, 1])
    assert set(np.unique(g[:, 1]).tolist()) == {-1.0, 3.0}
    hit = g[:, 1] == 3.0
    assert np.all(np.abs(g[hit, 0] - r[hit, 0]) <= 1e-4 * r[hit, 0])
    assert centre[0, 1] == 3.0 and abs(centre[0, 0] - 3.0) < 2e-4


def test_megakernel_variants_render_the_same(ctx):
    """config 3's fast-build program spills at four waves per SIMD, so it has
    two megakernel builds (the default without the bound, the four-wave one
    with spills); the autotuner may pick either: both render the same image
    and RNG states, bit for bit"""
    s = scenes.mandelbox_volume(48, 32, 4)
    out = []
    for variant in (0, 1):
        sess = host.Session(s, 48, 32, ctx, seed=77, flags=native.BUILD_FAST)
        assert sess.prog.mega_variants() == 2
        sess.prog.mega_variant(variant)
        sess.render(4)
        out.append((sess.radiance().copy(), sess.state.read(1).copy()))
        sess.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
