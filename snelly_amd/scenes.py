"""Scene definitions of the benchmark workloads, written the way a Snelly scene
script writes them: a GLSL snippet plus renderer / camera / material settings
(compare exampleScenes/fractals/menger.html and mandelbox.html in the reference).
"""
import math

import numpy as np

from . import host

SPHERE_PLANE_GLSL = """
float SDF_SURFACE(vec3 X)
{
    return min(length(X) - 1.0, X.y + 1.0);
}
vec3 SURFACE_DIFFUSE_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V)
{
    if (X.y > -1.0 + 1.0e-3) return vec3(0.8);
    return vec3(0.5);
}
"""

# the furnace: one convex Lambertian sphere (albedo 1) under a uniform sky
FURNACE_GLSL = """
float SDF_SURFACE(vec3 X) { return length(X) - 1.0; }
"""

MENGER_GLSL = """
uniform float size;
uniform float scale;
float sdBox(vec3 X, vec3 bounds)
{
    vec3 d = abs(X) - bounds;
    return min(max(d.x,max(d.y,d.z)),0.0) + length(max(d,0.0));
}
float menger(vec3 X)
{
    float sd = sdBox(X, vec3(size));
    float s = scale;
    const int iter = int(__ITER__);
    for (int i=0; i<iter; i++)
    {
        vec3 a = mod(X*s, 2.0)-1.0;
        s *= 3.0;
        vec3 r = abs(1.0 - 3.0*abs(a));
        float da = max(r.x,r.y);
        float db = max(r.y,r.z);
        float dc = max(r.z,r.x);
        float c = (min(da,min(db,dc))-1.0)/s;
        sd = max(sd, c);
    }
    return sd;
}
float SDF_METAL(in vec3 X)
{
    return menger(X);
}
"""

MANDELBOX_GLSL = """
uniform float Scale;
uniform float MinRad2;
vec4 scale;
float absScalem1;
float AbsScaleRaisedTo1mIters;
void INIT()
{
    scale = vec4(Scale, Scale, Scale, abs(Scale)) / MinRad2;
    absScalem1 = abs(Scale - 1.0);
    AbsScaleRaisedTo1mIters = pow(abs(Scale), float(1-__ITER__));
}
float mandelbox(vec3 pos)
{
    vec4 p = vec4(pos,1);
    vec4 p0 = p;
    for (int i=0; i<__ITER__; i++)
    {
        p.xyz = clamp(p.xyz, -1.0, 1.0)*2.0 - p.xyz;
        float r2 = dot(p.xyz, p.xyz);
        p *= clamp(max(MinRad2/r2, MinRad2), 0.0, 1.0);
        p = p*scale + p0;
        if (r2>__BAILOUT__) break;
    }
    return ((length(p.xyz) - absScalem1) / p.w - AbsScaleRaisedTo1mIters);
}
float SDF_DIELECTRIC(vec3 X) { return mandelbox(X); }
float SDF_SURFACE(vec3 X) { return length(X) - 0.5; }
"""

TORUS_GLSL = """
float SDF_@MAT@(vec3 X)
{
    vec2 q = vec2(length(X.xz) - 1.0, X.y);
    return length(q) - 0.3;
}
"""

SPHERE_GLSL = """
float SDF_@MAT@(vec3 X) { return length(X) - 1.0; }
"""


# FLOPs of one Menger SDF evaluation: 10 for the box, 24 per iteration
MENGER_SDF_FLOPS = {n: 10 + 24 * n for n in range(1, 8)}
SPHERE_PLANE_SDF_FLOPS = 8


# a volume emission field (the VOLUME_EMISSION hook, pathtracer-fragment-shader.glsl:863-869)
VOLUME_EMISSION_GLSL = """
vec3 VOLUME_EMISSION(vec3 emission, vec3 X) { return X.y > 0.0 ? emission : 0.25 * emission; }
"""


class Scene:
    def __init__(self, name, glsl, width, height, spp):
        self.name = name
        self.glsl = glsl
        self.width, self.height, self.spp = width, height, spp
        self.renderer = host.Renderer()
        self.materials = host.Materials()
        self.camera = host.Camera()
        self.scales = host.Scales()
        self.user_uniforms = {}
        # Scene.envMap(): (width, height, uint8 RGB image rows, top row first) or None
        self.env_map = None
        # the same geometry written by hand for the CPU oracle:
        # (material id 0 dielectric / 1 metal / 2 surface, kind, a, b, c, d)
        self.oracle_sdfs = []
        self.oracle_albedo_hook = 0
        self.oracle_emission_hook = 0  # 1: the VOLUME_EMISSION of VOLUME_EMISSION_GLSL
        # algorithmic FLOPs of one march step (SDF evaluations of the enabled
        # materials + 8 for p = o + t*d and t += sign*sdf), from the counting
        # build of the oracle (pinned by tests/test_flops.py); None = varies
        self.flops_per_step = None
        self.sdf_flops = None  # one SDF evaluation; a normal costs 6
        r = self.renderer
        r.interactive = False
        r.sunLongitude = 0.0

    @property
    def path_samples(self):
        return self.width * self.height * self.spp * int(round(self.renderer.maxSamplesPerFrame))


def sphere_plane(width=256, height=256, spp=16):
    """Unit sphere (albedo 0.8) over the plane y=-1 (albedo 0.5), Lambertian;
    constant white sky 1.0; pinhole camera at (0,0,4), fov 45; 4 bounces;
    spectral (hero wavelength); per-pixel RNG from PCG(pixel ^ 1234)."""
    s = Scene("sphere_plane", SPHERE_PLANE_GLSL, width, height, spp)
    r = s.renderer
    r.dispersive = True
    r.maxBounces = 4
    r.maxMarchSteps = 256
    r.skyPower = 0.0        # 10^0 = 1
    r.sunPower = -7.0       # the reference's "off"
    r.filterRadius = 2.0
    r.radianceClamp = 3.0
    s.camera.position = [0.0, 0.0, 4.0]
    s.camera.target = [0.0, 0.0, 0.0]
    s.camera.fov = 45.0
    s.camera.aperture = -math.inf   # pinhole
    su = s.materials.surface
    su.diffuseAlbedo = [1.0, 1.0, 1.0]
    su.specAlbedo = [0.0, 0.0, 0.0]
    su.ior = 1.0
    su.roughness = 0.1
    s.oracle_sdfs = [(2, 2, 0, 0, 0, 0)]
    s.oracle_albedo_hook = 1
    s.flops_per_step = 8 + SPHERE_PLANE_SDF_FLOPS
    s.sdf_flops = SPHERE_PLANE_SDF_FLOPS
    return s


def furnace(width=64, height=64, spp=64, dispersive=False):
    """White furnace: convex Lambertian sphere, diffuse albedo 1, uniform sky 1."""
    s = Scene("furnace", FURNACE_GLSL, width, height, spp)
    r = s.renderer
    r.dispersive = dispersive
    r.maxBounces = 3
    r.skyPower = 0.0
    r.sunPower = -7.0
    r.radianceClamp = 3.0
    s.camera.position = [0.0, 0.0, 4.0]
    s.camera.fov = 45.0
    s.camera.aperture = -math.inf
    su = s.materials.surface
    su.diffuseAlbedo = [1.0, 1.0, 1.0]
    su.specAlbedo = [0.0, 0.0, 0.0]
    su.ior = 1.0
    s.oracle_sdfs = [(2, 1, 1.0, 0, 0, 0)]
    return s


def menger_metal(width=1024, height=1024, spp=64, iterations=4):
    """Menger sponge (4 iterations, half-size 1) as Metal with analytic complex
    IOR n = 0.2 + 0.8 t, k = 3.5 - 1.5 t, t = (lambda - 390)/360, Beckmann
    roughness 0.05; 8 bounces; spectral."""
    s = Scene("menger_metal", MENGER_GLSL.replace("__ITER__", str(iterations)), width, height, spp)
    r = s.renderer
    r.dispersive = True
    r.maxBounces = 8
    r.maxMarchSteps = 256
    r.skyPower = 0.0
    r.sunPower = -7.0
    r.radianceClamp = 3.0
    s.camera.position = [2.6, 2.0, 3.4]
    s.camera.target = [0.0, 0.0, 0.0]
    s.camera.fov = 45.0
    s.camera.aperture = -math.inf
    m = host.GlslMetal("analytic", host.ANALYTIC_METAL_GLSL)
    m.roughness = 0.05
    s.materials.metal = m
    s.user_uniforms = {"size": 1.0, "scale": 1.0}
    s.oracle_sdfs = [(1, 3, 1.0, 1.0, float(iterations), 0)]
    s.flops_per_step = 8 + MENGER_SDF_FLOPS[iterations]
    s.sdf_flops = MENGER_SDF_FLOPS[iterations]
    return s


def mandelbox_dielectric(width=1920, height=1080, spp=256, scale=2.5, iterations=12, bailout=10.0):
    """Mandelbox (scale 2.5, 12 iterations, bailout r^2 > 10) as a Cauchy
    dielectric n = 1.5 + 0.0042/lambda^2 (um) with Beer absorption (0.1, 0.3, 0.5),
    embedding a Surface sphere of radius 0.5; 16 bounces; spectral."""
    glsl = MANDELBOX_GLSL.replace("__ITER__", str(iterations)).replace("__BAILOUT__", repr(float(bailout)))
    s = Scene("mandelbox_dielectric", glsl, width, height, spp)
    r = s.renderer
    r.dispersive = True
    r.maxBounces = 16
    r.maxMarchSteps = 512
    r.skyPower = 0.0
    r.sunPower = -7.0
    s.camera.position = [0.0, 2.0, 9.0]
    s.camera.fov = 45.0
    s.camera.aperture = -math.inf
    d = host.CauchyDielectric("cauchy", [1.5, 0.0042, -2.0])
    d.roughness = 0.005
    d.absorptionScale = 1.0
    d.absorptionColor = [0.9, 0.7, 0.5]
    s.materials.dielectric = d
    s.materials.surface.diffuseAlbedo = [0.8, 0.8, 0.8]
    s.materials.surface.specAlbedo = [0.0, 0.0, 0.0]
    s.materials.surface.ior = 1.0
    s.user_uniforms = {"Scale": float(scale), "MinRad2": 0.25}
    s.oracle_sdfs = [(0, 4, float(scale), float(iterations), float(bailout), 0.25), (2, 1, 0.5, 0, 0, 0)]
    return s


# ---------------------------------------------------------------- batched workload
MANDELBOX_SDF_GLSL = """
uniform float Scale;
uniform float MinRad2;
float mandelbox(vec3 pos)
{
    vec4 scale = vec4(Scale, Scale, Scale, abs(Scale)) / MinRad2;
    float absScalem1 = abs(Scale - 1.0);
    float AbsScaleRaisedTo1mIters = pow(abs(Scale), float(1-12));
    vec4 p = vec4(pos,1);
    vec4 p0 = p;
    for (int i=0; i<12; i++)
    {
        p.xyz = clamp(p.xyz, -1.0, 1.0)*2.0 - p.xyz;
        float r2 = dot(p.xyz, p.xyz);
        p *= clamp(max(MinRad2/r2, MinRad2), 0.0, 1.0);
        p = p*scale + p0;
        if (r2>10.0) break;
    }
    return ((length(p.xyz) - absScalem1) / p.w - AbsScaleRaisedTo1mIters);
}
float SDF_@MAT@(vec3 X) { return mandelbox(X); }
"""

BATCH_FAMILIES = ["sphere", "torus", "menger2", "menger3", "menger4", "menger5", "mandelbox"]
BATCH_MATERIALS = ["METAL", "DIELECTRIC", "SURFACE"]


def batch_job(family, material, side, spp, bounces, rs):
    """One job of the batched workload: `family` in BATCH_FAMILIES as `material`,
    a side x side image, `spp` frames, `bounces`; other parameters from rs."""
    if family == "sphere":
        glsl, extent, flops = SPHERE_GLSL, 1.0, 8 + 6
    elif family == "torus":
        glsl, extent, flops = TORUS_GLSL, 1.3, None
    elif family.startswith("menger"):
        n = int(family[-1])
        glsl = MENGER_GLSL.replace("__ITER__", str(n)).replace("float SDF_METAL(in vec3 X)", "float SDF_@MAT@(in vec3 X)")
        extent, flops = 1.0, 8 + MENGER_SDF_FLOPS[n]
    else:
        glsl, extent, flops = MANDELBOX_SDF_GLSL, None, None
    glsl = glsl.replace("@MAT@", material)
    s = Scene(f"{family}_{material.lower()}", glsl, side, side, spp)
    mid = {"DIELECTRIC": 0, "METAL": 1, "SURFACE": 2}[material]
    r = s.renderer
    r.dispersive = True
    r.maxBounces = bounces
    r.maxMarchSteps = 256
    r.skyPower = 0.0
    r.sunPower = -7.0
    s.camera.fov = 45.0
    s.camera.aperture = -math.inf
    # the same SDFs restated in the oracle (material id, kind, parameters)
    if family == "sphere":
        s.oracle_sdfs = [(mid, 1, 1.0, 0, 0, 0)]
    elif family == "torus":
        s.oracle_sdfs = [(mid, 5, 1.0, 0.3, 0, 0)]
    if family.startswith("menger"):
        s.user_uniforms = {"size": 1.0, "scale": 1.0}
        s.oracle_sdfs = [(mid, 3, 1.0, 1.0, float(family[-1]), 0)]
    if family == "mandelbox":
        # |scale| < 1.5 gives degenerate boxes; draw again
        sc = 0.0
        while abs(sc) < 1.5:
            sc = float(rs.uniform(-2.0, 2.5))
        s.user_uniforms = {"Scale": sc, "MinRad2": 0.25}
        s.oracle_sdfs = [(mid, 4, sc, 12.0, 10.0, 0.25)]
        extent = 2.0 * (abs(sc) + 1.0) / (abs(sc) - 1.0) if sc > 0 else 2.0
        r.maxMarchSteps = 512
    d = 2.6 * extent
    s.camera.position = [0.55 * d, 0.45 * d, 0.75 * d]
    s.camera.target = [0.0, 0.0, 0.0]
    if material == "METAL":
        m = host.GlslMetal("analytic", host.ANALYTIC_METAL_GLSL)
        m.roughness = float(rs.uniform(0.02, 0.2))
        s.materials.metal = m
    elif material == "DIELECTRIC":
        dl = host.CauchyDielectric("cauchy", [1.5, 0.0042, -2.0])
        dl.roughness = float(rs.uniform(0.005, 0.05))
        dl.absorptionScale = 1.0
        dl.absorptionColor = [0.9, 0.7, 0.5]
        s.materials.dielectric = dl
    else:
        su = s.materials.surface
        su.diffuseAlbedo = [float(x) for x in rs.uniform(0.3, 0.9, 3)]
        su.specAlbedo = [0.2, 0.2, 0.2]
        su.ior = 1.5
        su.roughness = float(rs.uniform(0.05, 0.3))
    s.flops_per_step = flops
    return s


def batch_jobs(n=64, seed=2024, min_side=128, max_side=1024):
    """The batched workload: n scenes drawn from a fixed seed -- family in
    BATCH_FAMILIES, material Metal / Dielectric / Surface, side log-uniform in
    [min_side, max_side], spp in {4, 16, 64}, bounces 4..12."""
    rs = np.random.RandomState(seed)
    jobs = []
    for _ in range(n):
        family = BATCH_FAMILIES[rs.randint(len(BATCH_FAMILIES))]
        material = BATCH_MATERIALS[rs.randint(len(BATCH_MATERIALS))]
        side = int(round(math.exp(rs.uniform(math.log(min_side), math.log(max_side)))))
        spp = int([4, 16, 64][rs.randint(3)])
        bounces = int(rs.randint(4, 13))
        jobs.append(batch_job(family, material, side, spp, bounces, rs))
    return jobs


def served_jobs(n=200, seed=7, rate=20.0, min_side=64, max_side=512, target=(0.05, 0.20)):
    """The served workload: n jobs arriving as a Poisson process of `rate` per
    second, scenes drawn like batch_jobs (family, material, bounces), side
    log-uniform in [min_side, max_side], target relative RMS error uniform in
    `target`.  Returns [(arrival_s, scene, target)]."""
    rs = np.random.RandomState(seed)
    t, out = 0.0, []
    for _ in range(n):
        t += float(rs.exponential(1.0 / rate))
        family = BATCH_FAMILIES[rs.randint(len(BATCH_FAMILIES))]
        material = BATCH_MATERIALS[rs.randint(len(BATCH_MATERIALS))]
        side = int(round(math.exp(rs.uniform(math.log(min_side), math.log(max_side)))))
        bounces = int(rs.randint(4, 13))
        s = batch_job(family, material, side, 0, bounces, rs)
        out.append((t, s, float(rs.uniform(*target))))
    return out
