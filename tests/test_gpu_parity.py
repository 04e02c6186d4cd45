"""Parity of the HIP path (libsnelly_hip.so through its C ABI) with the CPU oracle
on the same seeded inputs.

Integer work (RNG seeding, RNG state evolution) must be bit-exact.  Float work
is compared per pixel: programs built with SNL_BUILD_PRECISE (no FMA
contraction) follow the oracle's float operations, so almost every pixel's
path makes identical random decisions and agrees to 1e-4 relative; the rest
(a transcendental rounding flips one branch of a path) are bounded by the
image-level relative RMSE.  The default (FMA) build is held to the same
image-level bound.
"""
import math

import numpy as np
import pytest

import oracle_py
from snelly_amd import host, native, scenes
from test_oracle import furnace_check, silhouette, unbiased

pytestmark = pytest.mark.gpu

PRECISE = native.BUILD_PRECISE


ENGINES = [native.ENGINE_WAVEFRONT, native.ENGINE_PERSISTENT, native.ENGINE_MEGAKERNEL]
ENGINE_IDS = ["wavefront", "persistent", "megakernel"]


def session(ctx, scene, w, h, flags=PRECISE, engine=native.ENGINE_PERSISTENT):
    sess = host.Session(scene, w, h, ctx, flags=flags)
    sess.prog.engine(engine)
    return sess


def rrmse(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def assert_same_estimator(g, r, g_other_seed, max_paired_fraction=0.5):
    """g and r render the same pixels from the same RNG states; where a rounding
    difference sends a path down another branch the two samples are
    independent.  Require (1) no systematic difference: every channel's mean
    difference within 4 standard errors, and (2) the paired difference well
    below the difference between two independent-seed renders."""
    d = g - r
    n = d.shape[0]
    se = d.std(axis=0) / math.sqrt(n)
    assert np.all(np.abs(d.mean(axis=0)) <= 4.0 * se + 1e-6 * np.abs(r.mean(axis=0)))
    assert rrmse(g, r) < max_paired_fraction * rrmse(g_other_seed, r)


def pixel_grid(w, h):
    return np.stack(np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5), -1).reshape(-1, 2).astype(np.float32)


@pytest.mark.parametrize("w,h,seed", [(64, 64, 1234), (37, 23, 0), (1, 1, 42), (128, 96, 0xDEADBEEF)])
def test_seed_bit_exact(ctx, w, h, seed):
    st = native.State(ctx, w, h, seed)
    got = st.read(1).reshape(-1, 4)
    rad = st.read(0).reshape(-1, 4)
    st.close()
    assert np.array_equal(got, oracle_py.Oracle.seed(w * h, seed))
    assert np.array_equal(rad, np.tile([0.0, 0.0, 0.0, 1.0], (w * h, 1)).astype(np.float32))


@pytest.mark.parametrize("make", [scenes.sphere_plane, scenes.menger_metal, scenes.mandelbox_dielectric])
def test_first_hit_depth(ctx, make):
    """pick program per pixel: distance to 1e-4 relative, same material"""
    W, H = 64, 48
    s = make(W, H, 1)
    sess = session(ctx, s, W, H)
    o = oracle_py.Oracle(s, W, H)
    xy = pixel_grid(W, H)
    g, r = sess.prog.pick(xy), o.pick(xy)
    sess.close()
    same = g[:, 1] == r[:, 1]
    assert same.mean() >= 0.995
    rel = np.abs(g[same, 0] - r[same, 0]) / np.abs(r[same, 0])
    assert np.mean(rel <= 1e-4) >= 0.995
    assert np.median(rel) < 1e-6


@pytest.mark.parametrize("make", [scenes.sphere_plane, scenes.menger_metal])
def test_first_hit_normals(ctx, make):
    """normals render mode (jittered primary ray, SDF gradient) per pixel"""
    W, H = 64, 48
    s = make(W, H, 1)
    sess = session(ctx, s, W, H)
    sess.prog.render_normals(sess.state, 4)
    g = sess.radiance().reshape(-1, 4)
    rng_g = sess.state.read(1).reshape(-1, 4)
    sess.close()
    o = oracle_py.Oracle(s, W, H)
    rad, rng = o.new_state()
    o.normals(rad, rng, 4)
    assert np.array_equal(rng_g, rng)
    sil = silhouette(o, W, H)
    rel = np.max(np.abs(g[:, :3] - rad[:, :3]) / np.maximum(np.abs(rad[:, :3]), 1e-3), axis=1)
    assert np.mean(rel[~sil] <= 1e-4) >= 0.99
    assert np.array_equal(g[:, 3], rad[:, 3])


def compare_render(ctx, s, W, H, frames, flags=PRECISE, engine=native.ENGINE_PERSISTENT):
    sess = session(ctx, s, W, H, flags, engine)
    sess.render(frames)
    g = sess.radiance().reshape(-1, 4)
    rng_g = sess.state.read(1).reshape(-1, 4)
    sess.close()
    o = oracle_py.Oracle(s, W, H)
    rad, rng = o.new_state()
    o.render(rad, rng, frames)
    close = np.all(np.abs(g[:, :3] - rad[:, :3]) <= 1e-4 * np.maximum(np.abs(rad[:, :3]), 1e-3), axis=1)
    rng_same = np.all(rng_g == rng, axis=1)
    assert np.array_equal(g[:, 3], rad[:, 3])
    return g, rad, close.mean(), rng_same.mean(), o


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_pathtrace_sphere_plane(ctx, engine):
    """config-1 scene at 64x64, 16 spp: per-pixel agreement with the oracle"""
    s = scenes.sphere_plane(64, 64, 16)
    g, r, close, rng_same, _ = compare_render(ctx, s, 64, 64, 16, engine=engine)
    assert close >= 0.99 and rng_same >= 0.99
    assert rrmse(unbiased(g), unbiased(r)) < 2e-3


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_tiny_ragged_images_zero_frames_and_bounces(ctx, engine):
    """edge cases of the launch shapes: 1x1, sizes that are not multiples of
    the 8x8 tile or of a wave (3x5, 9x7, 17x8, 65x3), no bounce beyond the
    first vertex, and a render of zero frames, which leaves the state as
    seeded; every pixel's sample count equals the oracle's and all but a rare
    rounding-flipped path agree per pixel"""
    closes, rngs = [], []
    for (W, H, bounces) in [(1, 1, 4), (3, 5, 4), (9, 7, 0), (17, 8, 4), (65, 3, 1)]:
        s = scenes.sphere_plane(W, H, 3)
        s.renderer.maxBounces = bounces
        g, r, close, rng_same, _ = compare_render(ctx, s, W, H, 3, engine=engine)
        closes.append(close * W * H)
        rngs.append(rng_same * W * H)
    n = 1 + 15 + 63 + 136 + 195
    assert sum(closes) >= 0.97 * n and sum(rngs) >= 0.97 * n
    s = scenes.sphere_plane(9, 7, 3)
    sess = session(ctx, s, 9, 7, engine=engine)
    rad0, rng0 = sess.radiance().copy(), sess.state.read(1).copy()
    sess.render(0)
    assert np.array_equal(sess.radiance(), rad0) and np.array_equal(sess.state.read(1), rng0)
    assert np.all(rad0.reshape(-1, 4) == np.array([0, 0, 0, 1], np.float32))
    sess.close()


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_pathtrace_menger_metal(ctx, engine):
    s = scenes.menger_metal(64, 64, 8)
    g, r, close, rng_same, _ = compare_render(ctx, s, 64, 64, 8, engine=engine)
    assert close >= 0.95 and rng_same >= 0.95
    assert rrmse(unbiased(g), unbiased(r)) < 1e-2


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_pathtrace_mandelbox_dielectric(ctx, engine):
    s = scenes.mandelbox_dielectric(48, 32, 4, iterations=8)
    s.renderer.maxBounces = 8
    g, r, close, rng_same, _ = compare_render(ctx, s, 48, 32, 4, engine=engine)
    assert close >= 0.8
    assert rrmse(unbiased(g), unbiased(r)) < 5e-2


# the scene families of the batched and served workloads (scenes.batch_job), each
# with the oracle's restatement of its SDF; the last Mandelbox has the negative
# scale of served job 52 (scale -1.897, rough dielectric with Beer absorption)
BATCH_FAMILY_CASES = [("sphere", "DIELECTRIC", None), ("torus", "METAL", None), ("torus", "SURFACE", None),
                      ("menger3", "DIELECTRIC", None), ("menger5", "SURFACE", None), ("mandelbox", "METAL", None),
                      ("mandelbox", "DIELECTRIC", -1.8974167769271963)]


@pytest.mark.parametrize("engine", [native.ENGINE_PERSISTENT, native.ENGINE_MEGAKERNEL], ids=["persistent", "megakernel"])
@pytest.mark.parametrize("family,material,scale", BATCH_FAMILY_CASES,
                         ids=[f"{f}_{m.lower()}" for f, m, _ in BATCH_FAMILY_CASES])
def test_batch_families_match_oracle(ctx, family, material, scale, engine):
    """the batch and served families against the oracle, 40x40 at 4 spp,
    6 bounces: per-pixel agreement where the paths stay the same"""
    s = scenes.batch_job(family, material, 40, 4, 6, np.random.RandomState(3))
    if scale is not None:
        s.user_uniforms["Scale"] = scale
        s.oracle_sdfs = [(0, 4, scale, 12.0, 10.0, 0.25)]
    g, r, close, rng_same, _ = compare_render(ctx, s, 40, 40, 4, engine=engine)
    assert close >= (0.8 if material == "DIELECTRIC" else 0.95), close
    assert rng_same >= (0.8 if material == "DIELECTRIC" else 0.95), rng_same


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_pathtrace_fast_build_statistics(ctx, engine):
    """the default FMA build against the oracle: contraction changes roundings,
    a few per cent of paths take another branch; the image stays within the
    1% relative-RMSE bound (the Monte-Carlo noise itself is ~30% at 16 spp)"""
    s = scenes.sphere_plane(64, 64, 16)
    g, r, close, _, _ = compare_render(ctx, s, 64, 64, 16, flags=0, engine=engine)
    assert close >= 0.9
    assert rrmse(unbiased(g), unbiased(r)) < 1e-2


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_pathtrace_fast_math_build_statistics(ctx, engine):
    """SNL_BUILD_FAST (1-ulp divide and sqrt outside the RNG): same 1% bound"""
    s = scenes.menger_metal(64, 64, 16)
    g, r, close, _, _ = compare_render(ctx, s, 64, 64, 16, flags=native.BUILD_FAST, engine=engine)
    sess = session(ctx, s, 64, 64, native.BUILD_FAST, engine)
    sess.state.seed(4321)
    sess.render(16)
    g2 = sess.radiance().reshape(-1, 4)
    sess.close()
    assert close >= 0.5
    assert_same_estimator(unbiased(g), unbiased(r), unbiased(g2))


def test_fast_math_keeps_rng_exact(ctx):
    """the RNG's divisions stay correctly rounded in the fast build: furnace
    paths draw a fixed number of random numbers, so the states match exactly"""
    s = scenes.furnace(32, 32, 16)
    _, _, _, rng_same, _ = compare_render(ctx, s, 32, 32, 16, flags=native.BUILD_FAST)
    assert rng_same == 1.0


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_furnace(ctx, engine):
    W = H = 32
    s = scenes.furnace(W, H, 64)
    g, r, close, rng_same, o = compare_render(ctx, s, W, H, 64, engine=engine)
    assert rng_same == 1.0  # every path draws the same number of random numbers
    furnace_check(unbiased(g), silhouette(o, W, H), 64)


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_samples_per_frame_and_interactive_skip(ctx, engine):
    """maxSamplesPerFrame = 2 (sample loop 1954-1972) and the interactive skip
    (main 1994-2005, skipProbability 0.5) against the oracle"""
    s = scenes.sphere_plane(32, 32, 4)
    s.renderer.maxSamplesPerFrame = 2
    s.renderer.interactive = True
    s.renderer.skipProbability = 0.5
    g, r, close, rng_same, _ = compare_render(ctx, s, 32, 32, 4, engine=engine)
    assert close >= 0.99 and rng_same >= 0.99


def test_tonemap(ctx):
    W, H = 32, 32
    s = scenes.sphere_plane(W, H, 4)
    s.renderer.exposure = 0.5
    s.renderer.saturation = 1.2
    s.renderer.contrast = 1.1
    sess = session(ctx, s, W, H)
    sess.render(4)
    rgb, rgba8 = sess.prog.tonemap(sess.state)
    rad = sess.radiance()
    sess.close()
    o = oracle_py.Oracle(s, W, H)
    ref = o.tonemap(rad)
    assert np.allclose(rgb.reshape(-1, 3), ref, rtol=1e-5, atol=1e-6)
    assert np.array_equal(rgba8[..., :3].reshape(-1, 3), np.floor(np.clip(ref, 0, 1) * 255 + 0.5).astype(np.uint8))


def test_uniform_semantics(ctx):
    """unknown names and type-mismatched setters are silent no-ops (WebGL)"""
    s = scenes.sphere_plane(8, 8, 1)
    sess = session(ctx, s, 8, 8)
    assert sess.prog.index("noSuchUniform") == -1
    assert not sess.prog.set("noSuchUniform", 1.0)
    sess.prog.set("maxStepsIsMiss", 2.5)          # float into an int uniform: ignored
    sess.prog.set("camFovy", np.int32(10))         # int into a float uniform: ignored
    sess.render(1)
    ref = session(ctx, s, 8, 8)
    ref.render(1)
    assert np.array_equal(sess.radiance(), ref.radiance())
    sess.close()
    ref.close()


def _case_scene(name, W, H, spp):
    """a reference fixture case (tests/golden/reference_cases.py) at another size"""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import reference_cases
    c = reference_cases.CASES[name]
    s = getattr(scenes, c["builder"])(W, H, spp)
    c["tweak"](s)
    return s


@pytest.mark.parametrize("name", ["atmosphere", "atmosphere_fog_light"])
def test_pathtrace_atmosphere(ctx, name):
    """homogeneous atmosphere (in-scattering estimates, up to two real scatters,
    fog, sphere light through the medium): megakernel against the oracle"""
    s = _case_scene(name, 64, 48, 8)
    g, r, close, rng_same, _ = compare_render(ctx, s, 64, 48, 8, engine=native.ENGINE_MEGAKERNEL)
    assert close >= 0.97 and rng_same >= 0.97, (close, rng_same)
    assert rrmse(unbiased(g), unbiased(r)) < 1e-2


def test_atmosphere_ignores_engine_choice(ctx):
    """atmosphere programs always take the megakernel: the engine knob does not
    change a bit"""
    s = _case_scene("atmosphere", 32, 24, 2)
    out = []
    for e in ENGINES:
        sess = session(ctx, s, 32, 24, engine=e)
        sess.render(2)
        out.append((sess.radiance().copy(), sess.state.read(1).copy()))
        sess.close()
    for rad, rng in out[1:]:
        assert np.array_equal(rad, out[0][0]) and np.array_equal(rng, out[0][1])


@pytest.mark.parametrize("engine", [native.ENGINE_WAVEFRONT, native.ENGINE_MEGAKERNEL], ids=["wavefront", "megakernel"])
def test_pathtrace_subsurface(ctx, engine):
    """sub-surface random walk (randomwalk_SSS, SSS_exit_radiance) against the oracle"""
    s = _case_scene("subsurface", 64, 48, 8)
    g, r, close, rng_same, _ = compare_render(ctx, s, 64, 48, 8, engine=engine)
    assert close >= 0.97 and rng_same >= 0.97, (close, rng_same)
    assert rrmse(unbiased(g), unbiased(r)) < 1e-2


@pytest.mark.parametrize("name", ["ao", "ao_menger", "ptsimple", "ptsimple_subsurface"])
def test_render_modes(ctx, name):
    """renderMode 'ao' and 'ptsimple' (their own integrators) against the oracle"""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import reference_cases
    c = reference_cases.CASES[name]
    s = getattr(scenes, c["builder"])(64, 48, 8)
    c["tweak"](s)
    g, r, close, rng_same, _ = compare_render(ctx, s, 64, 48, 8)
    assert close >= 0.97 and rng_same >= 0.97, (close, rng_same)
    assert rrmse(unbiased(g), unbiased(r)) < 1e-2


@pytest.mark.parametrize("engine", [native.ENGINE_MEGAKERNEL, native.ENGINE_PERSISTENT], ids=["megakernel", "persistent"])
def test_tile_cost_order_does_not_change_results(ctx, engine, monkeypatch):
    """a render of several frames takes its 8x8 tiles costliest first after
    frame 1; each pixel still runs its frames in order in one lane, so the
    radiance and RNG states are bit-identical to a render in queue order
    (SNL_TILE_ORDER=0, read when the program is built)"""
    s = scenes.menger_metal(64, 48, 8)
    out = []
    for order in ("1", "0"):
        monkeypatch.setenv("SNL_TILE_ORDER", order)
        sess = session(ctx, s, 64, 48, native.BUILD_PRECISE, engine)
        sess.render(8)
        out.append((sess.radiance().copy(), sess.state.read(1).copy(), sess.prog.stats().trace_launches))
        sess.close()
    (ra, ga, la), (rb, gb, lb) = out
    assert la == 2 and lb == 1  # frame 1, then the rest in cost order; one launch in queue order
    assert np.array_equal(ra, rb) and np.array_equal(ga, gb)
