"""The heterogeneous medium of a scene that defines SDF_VOLUME, pinned in the
oracle.  The reference declares the medium's hooks (js/sceneAPI.js:124-131,
js/renderer.js:258-263) but its path tracer never calls them, so what pins the
oracle here is what the mathematics fixes:

* ratio tracking and delta tracking against the closed-form transmittance of
  the linear ball (density e (1 - |X|)), and the delta-tracked collision
  distance's distribution against its closed-form CDF;
* Henyey-Greenstein sampling against the phase function's CDF and mean cosine;
* the config-3 cloud below its majorant and zero outside its bound, and the
  value noise behind it at lattice points and across cell faces;
* renders: a pure absorber's pixels are the transmittance of their ray; an
  albedo-1 medium under a uniform white sky is white (furnace);
* a grey absorbing box equals the reference's homogeneous absorbing
  atmosphere, rendered by the reference's own shaders (tests/golden).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_py
from snelly_amd import host, scenes
from test_oracle import unbiased

HERE = os.path.dirname(os.path.abspath(__file__))
fp = oracle_py.fp


def f32(*v):
    return np.array(v, np.float32)


def linear_ball_tau(b, e):
    """optical depth of a full chord at impact parameter b through the ball of
    density e (1 - r): e (h - b^2 ln((1 + h) / b)), h = sqrt(1 - b^2)"""
    if b >= 1.0:
        return 0.0
    h = math.sqrt(1.0 - b * b)
    return e * (h - (b * b * math.log((1.0 + h) / b) if b > 0.0 else 0.0))


@pytest.fixture
def ball_oracle():
    s = scenes.medium_ball(8, 8, 4, albedo=0.0, extinction=2.0)
    return oracle_py.Oracle(s, 8, 8)


@pytest.mark.parametrize("b", [0.0, 0.35, 0.7, 0.95])
def test_ratio_tracking_matches_closed_form(ball_oracle, b):
    L = oracle_py.lib()
    rnd = oracle_py.Oracle.seed(1, 100 + int(b * 100))[0].copy()
    o, d = f32(b, 0.0, 4.0), f32(0.0, 0.0, -1.0)
    n = 30000
    T = np.array([L.oracle_medium_transmittance(fp(o), fp(d), 100.0, fp(rnd)) for _ in range(n)])
    want = math.exp(-linear_ball_tau(b, 2.0))
    se = T.std() / math.sqrt(n) + 1e-9
    assert abs(T.mean() - want) < 4.0 * se + 1e-6
    assert np.all((T >= 0.0) & (T <= 1.0))


@pytest.mark.parametrize("b", [0.0, 0.5, 0.9])
def test_delta_tracking_escape_probability(ball_oracle, b):
    L = oracle_py.lib()
    rnd = oracle_py.Oracle.seed(1, 200 + int(b * 100))[0].copy()
    o, d = f32(b, 0.0, 4.0), f32(0.0, 0.0, -1.0)
    n = 30000
    t = np.array([L.oracle_medium_free_flight(fp(o), fp(d), 100.0, fp(rnd)) for _ in range(n)])
    p = math.exp(-linear_ball_tau(b, 2.0))
    esc = (t >= 100.0).mean()
    assert abs(esc - p) < 4.0 * math.sqrt(p * (1.0 - p) / n)
    # collisions only inside the ball's chord
    hit = t[t < 100.0]
    h = math.sqrt(max(0.0, 1.0 - b * b))
    assert np.all((hit >= 4.0 - h - 1e-4) & (hit <= 4.0 + h + 1e-4))


def test_delta_tracking_collision_cdf(ball_oracle):
    """along the axis from z = 4 toward -z the optical depth to distance t is
    e (1 - s)^2 / 2 for s = 4 - t in [0, 1], e (1/2 - s - s^2/2) for s in
    [-1, 0]; P(collision before t) = 1 - exp(-depth)"""
    L = oracle_py.lib()
    rnd = oracle_py.Oracle.seed(1, 300)[0].copy()
    o, d = f32(0.0, 0.0, 4.0), f32(0.0, 0.0, -1.0)
    n = 40000
    t = np.array([L.oracle_medium_free_flight(fp(o), fp(d), 100.0, fp(rnd)) for _ in range(n)])
    for tt in [3.4, 3.8, 4.0, 4.3, 4.7]:
        s = 4.0 - tt
        depth = 2.0 * ((1.0 - s) ** 2 / 2.0 if s >= 0.0 else 0.5 - s - s * s / 2.0)
        p = 1.0 - math.exp(-depth)
        got = (t < tt).mean()
        assert abs(got - p) < 4.0 * math.sqrt(p * (1.0 - p) / n) + 1e-4, (tt, got, p)


def hg_cdf(mu, g):
    """P(cos theta <= mu) of the Henyey-Greenstein distribution"""
    if abs(g) < 1e-6:
        return (mu + 1.0) / 2.0
    return (1.0 - g * g) / (2.0 * g) * ((1.0 + g * g - 2.0 * g * mu) ** -0.5 - 1.0 / (1.0 + g))


@pytest.mark.parametrize("g", [0.3, -0.5, 0.0, 0.85])
def test_henyey_greenstein_sampling(g):
    L = oracle_py.lib()
    rnd = oracle_py.Oracle.seed(1, 400 + int(abs(g) * 100))[0].copy()
    d = f32(0.3, -0.5, 0.8)
    d /= np.linalg.norm(d)
    d = d.astype(np.float32)
    n = 40000
    w = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    out, pp = np.empty(3, np.float32), np.empty(1, np.float32)
    for k in range(n):
        L.oracle_sample_hg(fp(d), g, fp(rnd), fp(out), fp(pp))
        w[k], pdf[k] = out, pp[0]
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-5)
    mu = w @ d
    # histogram of cos(theta) against the CDF (chi-square, 20 bins)
    edges = np.linspace(-1.0, 1.0, 21)
    counts, _ = np.histogram(np.clip(mu, -1.0, 1.0), edges)
    expect = n * np.diff([hg_cdf(e, g) for e in edges])
    chi2 = float(np.sum((counts - expect) ** 2 / np.maximum(expect, 1e-9)))
    assert chi2 < 60.0, chi2
    # the mean direction is g d (mean cosine g, azimuth uniform around d)
    assert np.allclose(w.mean(axis=0), g * d, atol=4.0 / math.sqrt(n) + 1e-3)
    # the returned density is the phase function at the sampled cosine
    ph = np.array([L.oracle_phase_function(float(m), g) for m in mu[:200]])
    assert np.allclose(pdf[:200], ph, rtol=1e-4)


def test_phase_function_normalised():
    from scipy.integrate import quad
    L = oracle_py.lib()
    for g in [0.0, 0.3, -0.6, 0.9]:
        peak = 1.0 if g > 0 else -1.0
        integral = 2.0 * math.pi * quad(lambda m: L.oracle_phase_function(m, g), -1.0, 1.0, points=[peak], limit=200)[0]
        assert abs(integral - 1.0) < 1e-5, (g, integral)


def corner_value(c):
    """the noise's lattice value at integer corner c, from the oracle's hash
    levels (latticeValue(x) = hash(x) / 2^15 exactly)"""
    L = oracle_py.lib()

    def H(x):
        return int(round(L.oracle_lattice_value(int(x)) * 32768.0))
    return L.oracle_lattice_value(H(H(c[0] & 4095) + (c[1] & 4095)) + (c[2] & 4095))


def test_lattice_hash_exact_and_uniform():
    """hash values: exact integers (no int32 overflow over the whole input
    range the noise feeds it), in [0, 1) after the 2^-15 scale, and close to
    uniform"""
    L = oracle_py.lib()
    xs = np.arange(0, 32749 + 4096)
    v = np.array([L.oracle_lattice_value(int(x)) for x in xs[::7]])
    h = [((int(x) % 32749) ** 2 + 7919) % 32749 for x in xs[::7]]
    assert np.array_equal(v * 32768.0, np.array(h, float))
    assert v.min() >= 0.0 and v.max() < 1.0
    # corner values over a block of the lattice: near uniform
    cv = np.array([corner_value((i, j, k)) for i in range(-8, 8) for j in range(-8, 8) for k in range(-4, 4)])
    counts, _ = np.histogram(cv, np.linspace(0, 1, 9))
    expect = cv.size / 8.0
    assert np.sum((counts - expect) ** 2 / expect) < 40.0


def test_value_noise_lattice_points_and_continuity():
    L = oracle_py.lib()
    rs = np.random.RandomState(5)
    for c in rs.randint(-60, 60, size=(40, 3)):
        p = f32(*c)
        assert L.oracle_value_noise(fp(p)) == np.float32(corner_value(tuple(int(x) for x in c)))
    # across each kind of cell face the noise is continuous (a wrong corner
    # offset makes it jump by a lattice-value difference, ~0.3)
    for _ in range(60):
        base = rs.uniform(-20, 20, 3)
        for axis in range(3):
            p1 = base.copy()
            p1[axis] = math.floor(p1[axis]) + 1.0
            p0 = p1.copy()
            p0[axis] -= 1e-4
            a = L.oracle_value_noise(fp(p0.astype(np.float32)))
            b = L.oracle_value_noise(fp(p1.astype(np.float32)))
            assert abs(a - b) < 3e-4
    # the centre of a cell: the mean of its eight corners (smoothstep(1/2) = 1/2)
    for c in rs.randint(-30, 30, size=(20, 3)):
        corners = [corner_value((c[0] + i, c[1] + j, c[2] + k)) for i in (0, 1) for j in (0, 1) for k in (0, 1)]
        v = L.oracle_value_noise(fp(f32(*(c + 0.5))))
        assert abs(v - np.mean(corners)) < 1e-6


def test_cloud_density_under_majorant_and_inside_bound():
    s = scenes.mandelbox_volume(8, 8, 4)
    oracle_py.Oracle(s, 8, 8)
    L = oracle_py.lib()
    rs = np.random.RandomState(3)
    X = rs.uniform(-6.4, 6.4, (30000, 3)).astype(np.float32)
    P = (np.float32(0.25) * X).astype(np.float32)   # cloud coordinates
    dens = np.array([L.oracle_medium_density(fp(x)) for x in X])
    bound = np.array([L.oracle_medium_bound(fp(x)) for x in X])
    fb = np.array([L.oracle_fbm3(fp(p)) for p in P[:5000]])
    assert np.all(np.abs(fb) < 31.0 / 32.0)
    assert dens.max() <= 1.3                       # VOLUME_EXTINCTION_MAX(1) = 1.3
    assert np.all(dens[bound > 0.0] == 0.0)         # nothing outside SDF_VOLUME
    assert np.allclose(bound, np.linalg.norm(X, axis=1) - 5.2, atol=1e-5)
    inside = dens[np.linalg.norm(P, axis=1) < 0.8]
    assert inside.std() > 0.05                      # heterogeneous, not a plain ramp
    # the fbm term: density - max(0, 1 - |p|) is 0.3 fbm3(p) where both are positive
    r = np.linalg.norm(P[:5000], axis=1)
    pos = (1.0 - r + 0.3 * fb) > 0.0
    assert np.allclose(dens[:5000][pos], (1.0 - r + 0.3 * fb)[pos], atol=2e-6)


def ray_impact(scene, W, H):
    """impact parameter of each pixel's (unjittered, pinhole) camera ray about
    the origin, from the uniforms Renderer.render sets"""
    scene.camera.aspect = W / H
    u = host.frame_uniforms(scene.renderer, scene.camera, scene.scales, scene.materials, W, H)
    pos, cd, cx, cy = (np.array(u[k], float) for k in ("camPos", "camDir", "camX", "camY"))
    fh = math.tan(0.5 * math.radians(u["camFovy"]))
    fw = u["camAspect"] * fh
    ys, xs = np.mgrid[0:H, 0:W]
    ndc_x = -1.0 + 2.0 * (xs.ravel() + 0.5) / W
    ndc_y = -1.0 + 2.0 * (ys.ravel() + 0.5) / H
    d = cd[None, :] - fw * ndc_x[:, None] * cx[None, :] + fh * ndc_y[:, None] * cy[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.linalg.norm(np.cross(pos[None, :], d), axis=1)


def test_absorber_render_is_ray_transmittance():
    """a pure absorber (albedo 0) under a white sky: each sample is 1 if the
    ray escapes the ball without a collision, else 0, so a pixel's mean is the
    closed-form transmittance exp(-tau(b)) of its ray"""
    W = H = 16
    spp = 512
    s = scenes.medium_ball(W, H, spp, albedo=0.0, extinction=2.0)
    o = oracle_py.Oracle(s, W, H)
    rad, rng = o.new_state(7)
    o.render(rad, rng, spp)
    Y = unbiased(rad)[:, 1]
    b = ray_impact(s, W, H)
    T = np.exp(-np.array([linear_ball_tau(x, 2.0) for x in b]))
    white_y = 0.2126729 + 0.7151522 + 0.0721750
    se = np.sqrt(T * (1.0 - T) / spp) + 1e-7
    z = (Y / white_y - T) / se
    assert np.all(np.abs(z) < 5.0), np.abs(z).max()
    assert abs(z.mean()) < 4.0 / math.sqrt(z.size)
    assert np.all(Y[b >= 1.0] == np.float32(white_y))


@pytest.mark.parametrize("dispersive", [False, True])
def test_white_medium_furnace(dispersive):
    """an albedo-1 medium (g = 0.3) under a uniform white sky: every path's
    expected radiance is the sky's, whatever the medium does (next-event
    estimation and phase sampling combined by the balance heuristic)"""
    W = H = 12
    spp = 256
    s = scenes.medium_ball(W, H, spp, albedo=1.0, extinction=2.0, anisotropy=0.3, dispersive=dispersive)
    o = oracle_py.Oracle(s, W, H)
    rad, rng = o.new_state(9)
    o.render(rad, rng, spp)
    img = unbiased(rad)
    b = ray_impact(s, W, H)
    through, miss = b < 0.999, b > 1.001
    assert through.sum() > 40 and miss.sum() > 20
    if dispersive:
        # E over the hero wavelength of xyz(lambda): the image of the rays that
        # miss the ball, which see the sky alone
        want = img[miss].mean(axis=0)
    else:
        want = np.array([0.4124564 + 0.3575761 + 0.1804375, 0.2126729 + 0.7151522 + 0.0721750,
                         0.0193339 + 0.1191920 + 0.9503041])
        assert np.allclose(img[miss], want, rtol=1e-5)  # running-mean rounding over 256 frames
    rel = img[through, 1] / want[1] - 1.0
    se = rel.std() / math.sqrt(rel.size)
    assert abs(rel.mean()) < 4.0 * se + 2e-3, (rel.mean(), se)
    assert rel.std() < 0.25 + (0.5 if dispersive else 0.0)


def test_absorbing_box_equals_reference_atmosphere():
    """the reference's homogeneous atmosphere as a grey absorber in a box
    (rendered by its own shaders: tests/golden/ref_absorber_box.npz) and the
    heterogeneous medium's box of the same extinction describe the same
    transport; the oracle's medium at many samples per pixel must agree with
    the reference image within its Monte-Carlo error"""
    ref = np.load(os.path.join(HERE, "golden", "ref_absorber_box.npz"))
    meta = json.loads(bytes(ref["meta"]).decode())
    W, H, frames = meta["width"], meta["height"], meta["frames"]
    ref_img = unbiased(ref["radiance"])
    spp = 1024
    s = scenes.absorber_box(W, H, spp, medium=True)
    o = oracle_py.Oracle(s, W, H)
    rad, rng = o.new_state(3)
    o.render(rad, rng, spp)
    med = unbiased(rad)
    # the reference's own per-pixel spread: its atmosphere restated in the
    # oracle over independent seeds (pinned to the reference by the fixtures)
    a = scenes.absorber_box(W, H, frames, medium=False)
    oa = oracle_py.Oracle(a, W, H)
    reps = []
    for k in range(6):
        r2, g2 = oa.new_state(500 + k)
        oa.render(r2, g2, frames)
        reps.append(unbiased(r2))
    reps = np.array(reps)
    se = reps.std(axis=0, ddof=1) + 1e-4 * np.abs(med) + 1e-6
    z = (ref_img - med) / se
    assert np.median(np.abs(z)) < 1.5
    # no systematic difference: the image means agree
    d = (ref_img - med).mean(axis=0)
    sd = (ref_img - med).std(axis=0) / math.sqrt(W * H)
    assert np.all(np.abs(d) < 4.0 * sd + 1e-5), (d, sd)
    # and the medium's absorption is actually there: darker than no medium
    plain = scenes.sphere_plane(W, H, frames)
    op = oracle_py.Oracle(plain, W, H)
    r3, g3 = op.new_state(41)
    op.render(r3, g3, frames)
    assert unbiased(r3)[:, 1].mean() > 1.2 * med[:, 1].mean()
