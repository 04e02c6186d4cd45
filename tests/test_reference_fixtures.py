"""Parity against the reference's own shaders.

tests/golden/ref_<case>.npz hold what shaders/pathtracer-fragment-shader.glsl,
pick-fragment-shader.glsl and normals-fragment-shader.glsl compute for the
scenes of tests/golden/reference_cases.py, from integer RNG seeds
(make_reference_fixtures.py: a GLES 3 implementation with WebGL's zero
initialisation of locals and accurate transcendentals).

The oracle (CPU) and the HIP path (GPU) are held to the same bars:
  * the sample count channel is identical;
  * the per-pixel RNG states after all frames are bit-identical for >= 99% of
    pixels -- every random decision of those paths was the same;
  * per pixel, radiance within 1e-3 relative for >= 99.5% of pixels and within
    1e-4 for >= 98.5%.  What remains are last-ulp differences of the
    transcendentals (the fixture's are Cephes', 1-2 ulp) that near-specular
    lobes (roughness 0.02) at grazing incidence amplify (the BK7 sphere's
    silhouette);
  * first hits: the pick distance bit-identical and the same material for
    >= 99.5% of pixels (within 1e-4 relative for all of them); the normals
    pass within 1e-4 for >= 99.5% of pixels.  This needs the camera rays to be
    bit-identical: the fixture's tan, normalize and mix follow their
    definitions (the frustum height tan(fovy/2) correctly rounded, normalize
    as x / length(x), mix as x (1 - a) + y a), because central-difference
    normals (eps = 2e-4) turn one ulp of a hit point into ~1e-4.
The Mandelbox dielectric at 16 bounces is chaotic: paths separate after a few
internal bounces, so its radiance is compared as statistics (mean per channel
within 4 standard errors) beside floors on the per-pixel agreement; at
maxBounces 3 (mandelbox_dielectric_b3) it meets the per-pixel bars.
The environment-map case holds every bar but the 1e-4 ones on pixels that see
the map: an SRGB8 texture's decode and filter precision are the
implementation's (the GLES implementation that wrote the fixture decodes texel
centres to within 1.4e-3 relative and filters to within a median 2e-4 of exact
float bilinear filtering of the exactly decoded texels, which is what the HIP
path and the oracle compute); its radiance is held to 1e-3 and its normals to
1e-4 on the pixels whose first ray hits the sphere.
"""
import json
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import reference_cases  # noqa: E402

CASES = list(reference_cases.CASES)
CHAOTIC = {"mandelbox_dielectric"}
TEXTURE_PRECISION = {"env_map"}
ORACLE_CASES = [c for c in CASES if reference_cases.build(c)[0].oracle_sdfs]


def fixture(name):
    z = np.load(os.path.join(HERE, "golden", f"ref_{name}.npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    return {k: z[k] for k in z.files if k != "meta"}, meta


def rel_close(a, b, tol):
    return np.all(np.abs(a[:, :3] - b[:, :3]) <= tol * np.maximum(np.abs(b[:, :3]), 1e-3), axis=1)


def check_pathtrace(name, rad, rng, ref):
    """the bars of the module docstring for one render against its fixture"""
    assert np.array_equal(rad[:, 3], ref["radiance"][:, 3])
    same_rng = np.all(rng == ref["rng"], axis=1).mean()
    c3, c4 = (rel_close(rad, ref["radiance"], t).mean() for t in (1e-3, 1e-4))
    if name in CHAOTIC:
        d = rad[:, :3].astype(np.float64) - ref["radiance"][:, :3]
        se = d.std(axis=0) / math.sqrt(d.shape[0])
        assert np.all(np.abs(d.mean(axis=0)) <= 4.0 * se + 1e-6), (d.mean(axis=0), se)
        assert same_rng >= 0.85 and c4 >= 0.95, (same_rng, c4)
        return
    assert same_rng >= 0.99, same_rng
    assert c3 >= (0.99 if name in TEXTURE_PRECISION else 0.995), (c3, c4)
    if name not in TEXTURE_PRECISION:
        assert c4 >= 0.985, (c3, c4)


def check_pick(pk, ref):
    rp = ref["pick"]
    same = pk[:, 1] == rp[:, 1]
    assert same.mean() >= 0.995
    assert np.mean(same & (pk[:, 0] == rp[:, 0])) >= 0.995
    rel = np.abs(pk[same, 0] - rp[same, 0]) / np.abs(rp[same, 0])
    assert np.all(rel <= 1e-4)


def check_normals(name, nrad, nrng, ref):
    assert np.array_equal(nrng, ref["normals_rng"])
    if name in TEXTURE_PRECISION:
        hit = ref["pick"][:, 1] >= 0
        assert rel_close(nrad, ref["normals"], 1e-3).mean() >= 0.995
        assert rel_close(nrad[hit], ref["normals"][hit], 1e-4).mean() >= 0.98
    else:
        assert rel_close(nrad, ref["normals"], 1e-4).mean() >= 0.995


def pixel_grid(w, h):
    return np.stack(np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5), -1).reshape(-1, 2).astype(np.float32)


# ---------------------------------------------------------------- CPU: fixtures and oracle

@pytest.mark.parametrize("name", CASES)
def test_fixture_shapes(name):
    ref, meta = fixture(name)
    s, W, H, frames, seed = reference_cases.build(name)
    assert (meta["width"], meta["height"], meta["frames"], meta["seed"]) == (W, H, frames, seed)
    n = W * H
    for k in ("radiance", "rng", "normals", "normals_rng"):
        assert ref[k].shape == (n, 4) and ref[k].dtype == np.float32
    assert ref["pick"].shape == (n, 2)
    n_acc = ref["radiance"][:, 3]
    if s.renderer.interactive:  # skipped frames leave N as it is
        assert np.all((n_acc >= 1) & (n_acc <= frames + 1)) and np.any(n_acc < frames + 1)
    else:
        assert np.all(n_acc == frames + 1)  # cleared to N = 1, one sample per frame
    assert np.all(np.isfinite(ref["radiance"]))


@pytest.mark.parametrize("name", ORACLE_CASES)
def test_oracle_pathtrace_matches_reference(name):
    import oracle_py
    ref, _ = fixture(name)
    s, W, H, frames, seed = reference_cases.build(name)
    o = oracle_py.Oracle(s, W, H)
    rad, rng = o.new_state(seed)
    o.render(rad, rng, frames)
    check_pathtrace(name, rad, rng, ref)


@pytest.mark.parametrize("name", ORACLE_CASES)
def test_oracle_first_hit_matches_reference(name):
    import oracle_py
    ref, _ = fixture(name)
    s, W, H, frames, seed = reference_cases.build(name)
    o = oracle_py.Oracle(s, W, H)
    check_pick(o.pick(pixel_grid(W, H)), ref)
    nrad, nrng = o.new_state(seed)
    o.normals(nrad, nrng, 1)
    check_normals(name, nrad, nrng, ref)


# ---------------------------------------------------------------- GPU: the HIP path

ENGINE_OF = {"persistent": 1, "wavefront": 0, "megakernel": 2}  # native.ENGINE_*


@pytest.mark.gpu
@pytest.mark.parametrize("engine", list(ENGINE_OF))
@pytest.mark.parametrize("name", CASES)
def test_gpu_pathtrace_matches_reference(ctx, name, engine):
    """atmosphere programs render with the megakernel whichever engine is
    asked for; sub-surface programs with the wavefront unless the megakernel is"""
    from snelly_amd import host, native
    ref, _ = fixture(name)
    s, W, H, frames, seed = reference_cases.build(name)
    sess = host.Session(s, W, H, ctx, seed=seed, flags=native.BUILD_PRECISE)
    sess.prog.engine(ENGINE_OF[engine])
    sess.render(frames)
    rad = sess.radiance().reshape(-1, 4)
    rng = sess.state.read(1).reshape(-1, 4)
    sess.close()
    check_pathtrace(name, rad, rng, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_first_hit_matches_reference(ctx, name):
    from snelly_amd import host, native
    ref, _ = fixture(name)
    s, W, H, frames, seed = reference_cases.build(name)
    sess = host.Session(s, W, H, ctx, seed=seed, flags=native.BUILD_PRECISE)
    check_pick(sess.prog.pick(pixel_grid(W, H)), ref)
    sess.prog.render_normals(sess.state, 1)
    nrad = sess.radiance().reshape(-1, 4)
    nrng = sess.state.read(1).reshape(-1, 4)
    sess.close()
    check_normals(name, nrad, nrng, ref)


def tonemap_scene(name, meta):
    """the case's scene with the tonemapper settings its fixture was made with"""
    s, W, H, _, _ = reference_cases.build(name)
    for k, v in meta["tonemap"].items():
        setattr(s.renderer, k, v)
    return s, W, H


def check_tonemap(rgb, rgba8, ref):
    """float output within 1e-5 of the pixel's largest channel (the hue
    rotation mixes channels) plus 1e-6 (a channel that the out-of-gamut shift
    sets to 0 stays 0 here and comes out 2^-25 from the reference's inexact
    pow(x, 1.0)); the 8-bit store off by at most one step, and only where the
    float sits at a rounding boundary"""
    t = ref["tonemap"][:, :3]
    scale = np.abs(t).max(axis=1, keepdims=True)
    assert np.all(np.abs(rgb - t) <= 1e-5 * scale + 1e-6)
    if rgba8 is not None:
        want = np.floor(np.clip(t, 0.0, 1.0) * 255.0 + 0.5)
        diff = np.abs(rgba8.astype(np.int32) - want)
        assert diff.max() <= 1 and np.mean(diff > 0) <= 1e-3


@pytest.mark.parametrize("name", CASES)
def test_oracle_tonemap_matches_reference(name):
    """the oracle's tonemap over the fixture's radiance against the reference's
    tonemapper-fragment-shader.glsl (three settings of exposure, gamma,
    contrast, saturation and hue shift across the cases)"""
    import oracle_py
    ref, meta = fixture(name)
    s, W, H = tonemap_scene(name, meta)
    o = oracle_py.Oracle(s, W, H)
    check_tonemap(o.tonemap(ref["radiance"]), None, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_tonemap_matches_reference(ctx, name):
    from snelly_amd import host, native
    ref, meta = fixture(name)
    s, W, H = tonemap_scene(name, meta)
    sess = host.Session(s, W, H, ctx, seed=1, flags=native.BUILD_PRECISE)
    sess.state.write(0, ref["radiance"])
    rgb, rgba8 = sess.prog.tonemap(sess.state)
    sess.close()
    check_tonemap(rgb.reshape(-1, 3), rgba8[..., :3].reshape(-1, 3), ref)
