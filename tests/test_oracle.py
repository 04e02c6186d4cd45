"""Pins of the CPU oracle (oracle/oracle.cpp) against closed forms, invariants,
known-answer values and brute force -- each chosen so that a plausible slip in
the restatement (dropped term, wrong sign or index, transposed operand) fails.
"""
import math
import os

import numpy as np
import pytest

import oracle_py
from oracle_py import fp, f3
from snelly_amd import host, scenes

REF = "/root/reference"


def L():
    return oracle_py.lib()


# ---------------------------------------------------------------- RNG
M = [4194287, 4194277, 4194191, 4194167]
A = [3423, 2646, 1707, 1999]


def test_rand_is_exact_mlcg():
    """The float Schrage recurrence must reproduce s <- a*s mod m exactly on
    integer states (pathtracer-fragment-shader.glsl:119-130)."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        s = [int(rng.integers(1, m)) for m in M]
        st = f3(s)
        for _ in range(50):
            u = L().oracle_rand(fp(st))
            s = [(a * x) % m for a, x, m in zip(A, s, M)]
            assert [int(v) for v in st] == s
            exact = sum(sg * x / m for sg, x, m in zip([1, -1, 1, -1], s, M))
            exact -= math.floor(exact)
            assert abs(u - exact) < 2e-7 or abs(abs(u - exact) - 1.0) < 2e-7


def test_rand_uniform():
    st = f3([12345, 54321, 99999, 777])
    u = np.array([L().oracle_rand(fp(st)) for _ in range(200000)])
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3
    assert abs(u.var() - 1.0 / 12.0) < 2e-3
    hist, _ = np.histogram(u, bins=20, range=(0, 1))
    assert (np.abs(hist - 10000) < 600).all()


def test_pcg_seed_known_answer():
    """PCG32 with initstate 42, initseq 54 draws 0xa15c02b7 0x7b47f409 0xba1d3330
    0x83d2f293 (the pcg32 reference demo); pixel 0 with seed 42 maps each draw k
    to 1 + u mod (m_k - 1)."""
    rng = oracle_py.Oracle.seed(1, 42)
    draws = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293]
    assert [int(v) for v in rng[0]] == [1 + d % (m - 1) for d, m in zip(draws, M)]


# ---------------------------------------------------------------- Fresnel / microfacet
def fresnel_sp(cosi, eta):
    """textbook unpolarised Fresnel reflectance from Snell's law"""
    sint2 = (1 - cosi * cosi) / (eta * eta)
    if sint2 >= 1:
        return 1.0
    cost = math.sqrt(1 - sint2)
    rs = (cosi - eta * cost) / (cosi + eta * cost)
    rp = (eta * cosi - cost) / (eta * cosi + cost)
    return 0.5 * (rs * rs + rp * rp)


@pytest.mark.parametrize("eta", [1.0 / 1.5, 1.2, 1.5, 2.4])
def test_fresnel_dielectric_matches_snell(eta):
    for c in np.linspace(0.01, 1.0, 40):
        assert abs(L().oracle_fresnel_dielectric(float(c), eta) - fresnel_sp(c, eta)) < 2e-5


@pytest.mark.parametrize("n,k", [(0.2, 3.5), (1.0, 2.0), (1.5, 0.0), (0.05, 5.0)])
def test_fresnel_metal_limits(n, k):
    # normal incidence: exact complex-index reflectance ((n-1)^2+k^2)/((n+1)^2+k^2)
    r0 = ((n - 1) ** 2 + k * k) / ((n + 1) ** 2 + k * k)
    assert abs(L().oracle_fresnel_metal(1.0, n, k) - r0) < 1e-5
    # grazing incidence -> 1
    assert abs(L().oracle_fresnel_metal(0.0, n, k) - 1.0) < 1e-6
    if k == 0.0:  # a lossless "metal" is the dielectric formula at normal incidence
        assert abs(L().oracle_fresnel_metal(1.0, n, 0.0) - ((n - 1) / (n + 1)) ** 2) < 1e-6


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
def test_beckmann_normalisation(alpha):
    """integral of D(m) cos(theta_m) over the hemisphere = 1"""
    th = np.linspace(0.0, math.pi / 2, 4001)[:-1] + (math.pi / 2) / 8000
    dth = (math.pi / 2) / 4000
    tot = 0.0
    for t in th:
        m = f3([math.sin(t), 0.0, math.cos(t)])
        tot += L().oracle_microfacet_eval(fp(m), alpha) * math.cos(t) * math.sin(t) * dth * 2 * math.pi
    assert abs(tot - 1.0) < 2e-3


def test_smith_g1_close_to_exact_beckmann():
    """Walter et al.'s rational fit vs the erf form 2/(1 + erf(a) + exp(-a^2)/(a sqrt(pi)))"""
    for alpha in [0.05, 0.2, 0.5]:
        for t in np.linspace(0.05, 1.5, 30):
            v = f3([math.sin(t), 0.0, math.cos(t)])
            m = f3([0.0, 0.0, 1.0])
            a = 1.0 / (alpha * math.tan(t))
            exact = 2.0 / (1.0 + math.erf(a) + math.exp(-a * a) / (a * math.sqrt(math.pi)))
            g = L().oracle_smith_g1(fp(v), fp(m), alpha)
            assert abs(g - exact) <= 0.0035 * exact + 1e-6


# ---------------------------------------------------------------- dispersion models
def test_ior_models_at_d_line():
    """n at 587.56 nm: Fused Silica and Diamond agree with catalogue values
    (Malitson 1.4585, 2.4175); BK7 and N-FK51A carry Schott model-2 coefficients
    in a model-1 formula in the reference, giving 1.50723 / 1.47931 as coded."""
    for name, ref, tol in [("Fused Silica", 1.4585, 2e-4), ("Diamond", 2.4175, 2e-4),
                           ("Glass (BK7)", 1.50723, 2e-5), ("Glass (N-FK51A)", 1.47931, 2e-5)]:
        d = host.DIELECTRICS[name]()
        s = scenes.sphere_plane(4, 4, 1)
        s.materials.dielectric = d
        oracle_py.Oracle(s, 4, 4)
        assert abs(L().oracle_ior_diele(587.56) - ref) < tol, name


def test_cauchy_config3():
    s = scenes.mandelbox_dielectric(4, 4, 1)
    oracle_py.Oracle(s, 4, 4)
    for wl in [400.0, 550.0, 700.0]:
        assert abs(L().oracle_ior_diele(wl) - (1.5 + 0.0042 / (wl * 1e-3) ** 2)) < 1e-6


def test_analytic_metal_config2():
    s = scenes.menger_metal(4, 4, 1)
    oracle_py.Oracle(s, 4, 4)
    for wl in [390.0, 570.0, 750.0]:
        t = (wl - 390.0) / 360.0
        assert abs(L().oracle_ior_metal(wl) - (0.2 + 0.8 * t)) < 1e-6
        assert abs(L().oracle_k_metal(wl) - (3.5 - 1.5 * t)) < 1e-6


# ---------------------------------------------------------------- spectra
def test_blackbody_icdf_inverts_cdf():
    icdf = host.inverse_cdf(lambda wl: host.blackbody(wl, 6000.0))
    assert np.all(np.diff(icdf) >= 0) and 0.0 <= icdf[0] and icdf[-1] <= 1.0 + 1e-6
    # independent float64 inversion by interpolation of the cumulative spectrum
    n = 1024
    wl = 390.0 + 360.0 / n * np.arange(n)
    lp = wl / 360.0
    spec = 1.0 / (lp ** 5 * (np.exp(1.43877737467e7 / (wl * 6000.0)) - 1.0))
    cdf = np.concatenate([[0.0], np.cumsum(spec)])
    cdf /= cdf[-1]
    xi = np.minimum((np.arange(4 * n) + 1) / (4 * n), 1.0)
    expect = np.interp(xi, cdf, np.arange(n + 1) / n)
    assert np.max(np.abs(icdf - expect)) < 2e-5


@pytest.mark.skipif(not os.path.exists(REF), reason="needs the reference's CIE CSV (this container only)")
def test_xyz_lut_regenerates_from_cie_csv():
    """scripts/color/cie.py: numpy.interp of the CIE-2006 2-degree CMFs at
    390 + n*360/1023 nm, printed with %f"""
    data = np.loadtxt(os.path.join(REF, "scripts/color/lin2012xyz2e_fine_7sf.csv"), delimiter=",")
    l = 390.0 + (750.0 - 390.0) / 1023.0 * np.arange(1024)
    regen = np.stack([np.interp(l, data[:, 0], data[:, k]) for k in (1, 2, 3)], 1)
    regen = np.array([[float("%f" % v) for v in row] for row in regen])
    assert np.array_equal(np.asarray(host.tables()["xyz_lut"]), regen)


# ---------------------------------------------------------------- images
def unbiased(rad):
    """running mean (X,Y,Z,N) with N starting at 1 -> mean of the N-1 samples"""
    rad = rad.reshape(-1, 4)
    return rad[:, :3] * (rad[:, 3:4] / (rad[:, 3:4] - 1.0))


def furnace_check(img, sil, spp):
    """White furnace: a convex Lambertian sphere (albedo 1, ior 1) under a
    uniform sky returns rgbToXyz(1,1,1) in expectation.  NEE contributes exactly
    1/2 per vertex; the escaping BSDF ray is weighted by the reference's MIS
    weight p_bsdf(w_b)/(p_bsdf(w_b) + p_sky(w_nee)) -- two independent
    cosine-distributed pdfs -- whose mean is 1/2 by symmetry.  Its spread gives
    the per-pixel Monte-Carlo bound."""
    white = np.array([0.4124564 + 0.3575761 + 0.1804375, 0.2126729 + 0.7151522 + 0.0721750,
                      0.0193339 + 0.1191920 + 0.9503041])
    rel = img / white
    assert np.allclose(rel[:, 0], rel[:, 1], atol=1e-4) and np.allclose(rel[:, 1], rel[:, 2], atol=1e-4)
    x, y = np.sqrt(np.random.default_rng(0).random((2, 200000)))
    sigma = np.std(x / (x + y)) / math.sqrt(spp)
    dev = rel[~sil, 1] - 1.0
    assert abs(dev.mean()) < 4 * sigma / math.sqrt(dev.size)
    assert np.all(np.abs(dev) < 6 * sigma)
    return dev


def test_furnace_white_in_expectation():
    s = scenes.furnace(32, 32, 64, dispersive=False)
    o = oracle_py.Oracle(s, 32, 32)
    rad, rng = o.new_state()
    o.render(rad, rng, 64)
    furnace_check(unbiased(rad), silhouette(o, 32, 32), 64)


def silhouette(o, W, H):
    """pixels whose 3x3 neighbourhood mixes hit and miss first-hit materials"""
    xy = np.stack(np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5), -1).reshape(-1, 2)
    mat = o.pick(xy)[:, 1].reshape(H, W)
    pad = np.pad(mat, 1, mode="edge")
    mixed = np.zeros((H, W), bool)
    for dy in (0, 1, 2):
        for dx in (0, 1, 2):
            mixed |= pad[dy:dy + H, dx:dx + W] != mat
    return mixed.ravel()


def test_first_hit_depth_matches_ray_intersection():
    """sphere_plane pick distances vs analytic ray/sphere and ray/plane hits"""
    W = H = 32
    s = scenes.sphere_plane(W, H, 1)
    o = oracle_py.Oracle(s, W, H)
    xy = np.stack(np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5), -1).reshape(-1, 2)
    got = o.pick(xy)
    pos, camDir, camX, camY = s.camera.basis()
    fh = math.tan(0.5 * math.radians(45.0))
    for (px, py), (d, mat) in zip(xy, got):
        ndc = (-1 + 2 * px / W, -1 + 2 * py / H)
        ray = camDir - fh * ndc[0] * camX + fh * ndc[1] * camY
        ray /= np.linalg.norm(ray)
        b = np.dot(pos, ray)
        c = np.dot(pos, pos) - 1.0
        disc = b * b - c
        ts = -b - math.sqrt(disc) if disc > 0 else math.inf
        tp = (-1.0 - pos[1]) / ray[1] if ray[1] < 0 else math.inf
        t = min(ts, tp)
        if t > 100.0:
            assert d == pytest.approx(100.0) and mat == -1
        elif mat == -1:
            # a grazing plane hit that sphere tracing cannot reach within 256
            # steps (maxStepsIsMiss): the height above the plane shrinks by a
            # factor (1 - sin(theta)) per step from 1 + pos.y down to 1e-4
            assert tp < ts
            steps = math.log(1e-4 / (1.0 + pos[1])) / math.log(1.0 + ray[1])
            assert steps > 200
        else:
            assert mat == 2
            assert abs(d - t) < 3e-4 * max(1.0, t) / max(0.05, abs(ray[1]) if tp < ts else 1.0)
