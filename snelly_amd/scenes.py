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


# Heterogeneous media, written with the hooks the reference's scene API names
# for them (SDF_VOLUME, VOLUME_EXTINCTION, VOLUME_EXTINCTION_MAX;
# js/sceneAPI.js:124-131, js/renderer.js:258-263).  The oracle restates each
# (oracle/oracle.cpp, kinds MEDIUM_*) with the same float operations.
#
# CLOUD_GLSL: density max(0, 1 - |p| + 0.3 fbm3(p)) in cloud coordinates
# p = X / 4, with fbm3 five octaves of value noise -- lattice values from an
# integer hash (nested quadratic maps mod the prime 32749; every product stays
# below 2^31, so GLSL int arithmetic is exact) scaled by 2^-15 into [0, 1),
# smoothstep-blended -- weighted 1/2, 1/4, ..., 1/32 at doubling frequencies
# (base 2 per unit of p, offsets between octaves), so |fbm3| < 31/32: the
# density stays below 1.3 (the majorant) and vanishes outside |p| < 1.3 (the
# bound |X| < 5.2).  The extinction is volume.extinction (1/4) per world unit
# times the density, the optical depth of the unit-scale cloud (about 1 along
# a diameter).  At this scale the cloud envelops the Mandelbox of scale 2.5
# (half-size 4.67); at unit scale it would sit in the fractal's hollow centre,
# behind a shell that paths of 16 bounces through the glass do not cross.
CLOUD_GLSL = """
int latticeHash(int x)
{
    int h = x % 32749;
    return (h * h + 7919) % 32749;
}
float latticeValue(int x)
{
    return float(latticeHash(x)) * (1.0 / 32768.0);
}
float valueNoise(vec3 p)
{
    vec3 i = floor(p);
    vec3 f = p - i;
    vec3 u = f * f * (3.0 - 2.0 * f);
    ivec3 c = ivec3(i);
    int hx0 = latticeHash(c.x & 4095);
    int hx1 = latticeHash((c.x + 1) & 4095);
    int y0 = c.y & 4095;
    int y1 = (c.y + 1) & 4095;
    int z0 = c.z & 4095;
    int z1 = (c.z + 1) & 4095;
    int h00 = latticeHash(hx0 + y0);
    int h10 = latticeHash(hx1 + y0);
    int h01 = latticeHash(hx0 + y1);
    int h11 = latticeHash(hx1 + y1);
    float x00 = mix(latticeValue(h00 + z0), latticeValue(h10 + z0), u.x);
    float x10 = mix(latticeValue(h01 + z0), latticeValue(h11 + z0), u.x);
    float x01 = mix(latticeValue(h00 + z1), latticeValue(h10 + z1), u.x);
    float x11 = mix(latticeValue(h01 + z1), latticeValue(h11 + z1), u.x);
    return mix(mix(x00, x10, u.y), mix(x01, x11, u.y), u.z);
}
float fbm3(vec3 p)
{
    float sum = 0.0;
    float amp = 0.5;
    vec3 q = 2.0 * p;
    for (int i = 0; i < 5; i++)
    {
        sum += amp * (2.0 * valueNoise(q) - 1.0);
        amp *= 0.5;
        q = 2.0 * q + vec3(1.7, 9.2, 3.4);
    }
    return sum;
}
float SDF_VOLUME(vec3 X) { return length(X) - 5.2; }
float VOLUME_EXTINCTION(float extinction_ui, vec3 X)
{
    vec3 p = 0.25 * X;
    return extinction_ui * max(0.0, 1.0 - length(p) + 0.3 * fbm3(p));
}
float VOLUME_EXTINCTION_MAX(float extinction_ui) { return 1.3 * extinction_ui; }
"""

# density max(0, 1 - |X|) in the unit ball (its line integrals have a closed form)
LINEAR_BALL_GLSL = """
float SDF_VOLUME(vec3 X) { return length(X) - 1.0; }
float VOLUME_EXTINCTION(float extinction_ui, vec3 X) { return extinction_ui * max(0.0, 1.0 - length(X)); }
"""

# density 1 in the box |X_k| < h_k
BOX_MEDIUM_GLSL = """
float sdBoxMedium(vec3 X, vec3 h)
{
    vec3 d = abs(X) - h;
    return min(max(d.x, max(d.y, d.z)), 0.0) + length(max(d, 0.0));
}
float SDF_VOLUME(vec3 X) { return sdBoxMedium(X, vec3(@H@)); }
float VOLUME_EXTINCTION(float extinction_ui, vec3 X) { return sdBoxMedium(X, vec3(@H@)) <= 0.0 ? extinction_ui : 0.0; }
"""
MEDIUM_CLOUD, MEDIUM_LINEAR_BALL, MEDIUM_BOX = 1, 2, 3


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
        # the SDF_VOLUME medium restated for the oracle: (kind, a, b, c, d) or None
        self.oracle_medium = None
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


def mandelbox_volume(width=1920, height=1080, spp=256):
    """The Mandelbox dielectric scene of mandelbox_dielectric (scale 2.5, 12
    iterations, bailout 10, Cauchy dispersion, Beer absorption, nested Surface
    sphere of radius 0.5) in a heterogeneous medium: density
    max(0, 1 - |p| + 0.3 fbm3(p)) (5 octaves) in cloud coordinates p = X / 4
    (CLOUD_GLSL), albedo 0.9, Henyey-Greenstein g = 0.3; 1920 x 1080, 256 spp,
    16 bounces, spectral."""
    s = mandelbox_dielectric(width, height, spp)
    s.name = "mandelbox_volume"
    s.glsl = s.glsl + CLOUD_GLSL
    v = s.materials.volume
    v.extinction = 0.25
    v.scatteringColor = [0.9, 0.9, 0.9]
    v.absorptionColor = [0.1, 0.1, 0.1]
    v.anisotropy = 0.3
    s.oracle_medium = (MEDIUM_CLOUD, 5.2, 0.25, 0.0, 0.0)
    return s


def medium_ball(width=32, height=32, spp=16, albedo=0.0, extinction=2.0, anisotropy=0.3, dispersive=False,
                bounces=24):
    """A medium alone -- density `extinction` * max(0, 1 - |X|) in the unit
    ball -- under a uniform white sky 1, pinhole camera at (0, 0, 4), fov 45,
    no pixel jitter (every sample of a pixel takes the same ray).  albedo 0:
    a pure absorber, whose pixels have the closed-form transmittance of their
    ray as expectation; albedo 1: a white furnace, every pixel 1."""
    s = Scene("medium_ball", LINEAR_BALL_GLSL, width, height, spp)
    r = s.renderer
    r.dispersive = dispersive
    r.maxBounces = bounces
    r.skyPower = 0.0
    r.sunPower = -7.0
    r.filterRadius = 0.0
    r.radianceClamp = 3.0
    s.camera.position = [0.0, 0.0, 4.0]
    s.camera.fov = 45.0
    s.camera.aperture = -math.inf
    v = s.materials.volume
    v.extinction = extinction
    v.scatteringColor = [albedo] * 3
    v.absorptionColor = [1.0 - albedo] * 3
    v.anisotropy = anisotropy
    s.oracle_medium = (MEDIUM_LINEAR_BALL, 0.0, 0.0, 0.0, 0.0)
    return s


def absorber_box(width=40, height=32, spp=16, half=1.5, absorption=0.5, medium=True):
    """sphere_plane inside a grey absorbing box |X_k| < half of extinction
    `absorption` and no scattering: as the heterogeneous medium (SDF_VOLUME,
    medium=True) or as the reference's homogeneous atmosphere (medium=False,
    atmosphere bounds +-half, mfp 1).  The two describe the same transport."""
    s = sphere_plane(width, height, spp)
    v = s.materials.volume
    if medium:
        s.name = "absorber_box_medium"
        h = ", ".join([repr(float(half))] * 3)
        s.glsl = s.glsl + BOX_MEDIUM_GLSL.replace("@H@", h)
        v.extinction = absorption
        v.scatteringColor = [0.0, 0.0, 0.0]
        v.absorptionColor = [1.0, 1.0, 1.0]
        s.oracle_medium = (MEDIUM_BOX, float(half), float(half), float(half), 0.0)
    else:
        s.name = "absorber_box_atmosphere"
        v.atmosphereMinX = v.atmosphereMinY = v.atmosphereMinZ = -half
        v.atmosphereMaxX = v.atmosphereMaxY = v.atmosphereMaxZ = half
        v.lof10_mfp = 0.0
        v.scatteringColor = [0.0, 0.0, 0.0]
        v.absorptionColor = [absorption] * 3
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
        # scale uniform in [-2, 2.5] as drawn; the camera distance follows the
        # box's extent 2 (s + 1) / (s - 1) for s > 1, at most 6
        sc = float(rs.uniform(-2.0, 2.5))
        s.user_uniforms = {"Scale": sc, "MinRad2": 0.25}
        s.oracle_sdfs = [(mid, 4, sc, 12.0, 10.0, 0.25)]
        extent = min(6.0, 2.0 * (sc + 1.0) / (sc - 1.0)) if sc > 1.0 else (2.0 if sc < 0.0 else 6.0)
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


def batch_kinds():
    """one small scene of every family and material: the program kinds of the
    batched and served workloads (their uniforms differ per job, their code
    does not), from which a batch module holding all of them is built"""
    rs = np.random.RandomState(5)
    return [batch_job(fam, mat, 64, 1, 4, rs) for fam in BATCH_FAMILIES for mat in BATCH_MATERIALS]


def batch_sample():
    """one small job of every family (materials in turn, 3-5 frames, 4-10
    bounces): the programs of the precise-build batch module the tests use"""
    rs = np.random.RandomState(11)
    sides = [40, 48, 32, 56, 24, 64, 40]
    return [batch_job(fam, BATCH_MATERIALS[k % 3], sides[k], 3 + k % 3, 4 + k, rs)
            for k, fam in enumerate(BATCH_FAMILIES)]


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


def served_jobs(n=2000, seed=7, rate=20.0, min_side=64, max_side=2048, target=(0.005, 0.05)):
    """The served workload (config 5 by default): n jobs arriving as a
    Poisson process of `rate` per second, scenes drawn like batch_jobs
    (family, material, bounces), side log-uniform in [min_side, max_side],
    target relative RMS error uniform in `target`.
    Returns [(arrival_s, scene, target)]."""
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
