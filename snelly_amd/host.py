"""Python mirror of the reference's JS host for one render path.

What Snelly's JavaScript does around its shader programs -- renderer defaults
(js/renderer.js:78-126), shader assembly (Renderer.compileShaders,
js/renderer.js:217-359), the per-frame uniform upload (Renderer.render,
js/renderer.js:462-611), materials and their IOR code (js/materials.js), the
blackbody emission spectrum ICDF (js/spectra.js:28-61, 171-184) and the
three.js camera basis -- restated so tests and the benchmark can drive
libsnelly_hip.so from Python with the same inputs the JS host would send.
The JS host (snelly_amd/js) does the same for scene scripts.
"""
import json
import math
import os
import struct
import zlib

import numpy as np

from . import native

HERE = os.path.dirname(os.path.abspath(__file__))
_TABLES = None


def tables():
    global _TABLES
    if _TABLES is None:
        with open(os.path.join(HERE, "data", "spectral_tables.json")) as f:
            _TABLES = json.load(f)
    return _TABLES


def f32(x):
    return float(np.float32(x))


# ---------------------------------------------------------------- spectra
def blackbody(wavelength_nm, temperature):
    """BlackbodySpectrum.spectrum (js/spectra.js:179-184)."""
    boltzmann_factor = 1.43877737467e7 / (wavelength_nm * temperature)
    l = wavelength_nm / 360.0
    return 1.0 / (l * l * l * l * l * (math.exp(boltzmann_factor) - 1.0))


def inverse_cdf(spectrum, lmin=390.0, lmax=750.0, n=1024):
    """Spectrum.inverseCDF (js/spectra.js:28-61) with its Float32Array storage:
    every stored CDF/ICDF value is rounded to float32 as the JS typed arrays do."""
    cdf = np.zeros(n + 1, np.float32)
    dl = (lmax - lmin) / n
    for i in range(n):
        cdf[i + 1] = np.float32(float(cdf[i]) + spectrum(lmin + dl * i))
    s = float(cdf[n])
    for i in range(n):
        cdf[i + 1] = np.float32(float(cdf[i + 1]) / s)
    cdf[n] = 1.0
    icdf = np.zeros(4 * n, np.float32)
    idx = 0
    for k in range(4 * n):
        xi = min((k + 1) / (4 * n), 1.0)
        while float(cdf[idx]) < xi:
            idx += 1
        idx = max(1, idx)
        lo, hi = float(cdf[idx - 1]), float(cdf[idx])
        icdf[k] = np.float32((idx - 1.0) / n + (1.0 / n) * (xi - lo) / (hi - lo))
    return icdf


# ---------------------------------------------------------------- environment map
def srgb_to_linear(rgb8):
    """the SRGB8 texture decode of GLU.loadImageAndCreateTextureInfo
    (js/gl.js:532-559): the sRGB transfer function of each 8-bit channel,
    rounded to float32"""
    c = np.asarray(rgb8, np.float64) / 255.0
    lin = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    return lin.astype(np.float32)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def load_png(path):
    """decode an 8-bit RGB or RGBA, non-interlaced PNG (the env-map images
    scenes name; the browser decodes them in the reference) into
    (width, height, uint8 RGB array of shape (h, w, 3)), top row first"""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        chunk = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", chunk)
        elif kind == b"IDAT":
            idat += chunk
        elif kind == b"IEND":
            break
        pos += 12 + n
    w, h, depth, ctype, _, _, interlace = hdr
    if depth != 8 or ctype not in (2, 6) or interlace:
        raise ValueError(f"{path}: only 8-bit RGB/RGBA non-interlaced PNGs are read")
    bpp = 3 if ctype == 2 else 4
    raw = zlib.decompress(idat)
    stride = w * bpp
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        ft = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for x in range(stride):
            a = cur[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            pred = (0, a, b, (a + b) // 2, _paeth(a, b, c))[ft]
            cur[x] = (line[x] + pred) & 255
        out[y] = cur
        prev = cur
    return w, h, out.reshape(h, w, bpp)[:, :, :3].copy()


def save_png(path, rgba8):
    """write an (h, w, 4) uint8 image whose row 0 is the bottom row (as the
    tonemapper and gl.readPixels return it) as an RGBA PNG, top row first --
    the canvas.toBlob('image/png') of the 'P' key (js/snelly.js:791-806)"""
    img = np.asarray(rgba8, np.uint8)
    h, w = img.shape[:2]
    rows = img[::-1].reshape(h, w * 4)
    raw = b"".join(b"\x00" + rows[y].tobytes() for y in range(h))

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xffffffff)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(raw)) +
                chunk(b"IEND", b""))


def xyz_lut():
    """1024 x 4 RGBA32F wavelength -> XYZ table (js/color.js:24-283)."""
    t = np.asarray(tables()["xyz_lut"], np.float32)
    return np.concatenate([t, np.zeros((1024, 1), np.float32)], axis=1)


# ---------------------------------------------------------------- materials
class Surface:
    """js/materials.js:39-83"""

    def __init__(self):
        self.diffuseAlbedo = [1.0, 1.0, 1.0]
        self.specAlbedo = [1.0, 1.0, 1.0]
        self.roughness = 0.1
        self.ior = 1.5
        self.subsurface = 0.0
        self.subsurfaceAlbedoRGB = [0.5, 0.5, 0.5]
        self.subsurfaceMFP = 0.0
        self.subsurfaceAnisotropy = 0.0
        self.subsurfaceDiffuseWeight = 1.0

    def uniforms(self, s):
        return {"surfaceDiffuseAlbedoRGB": self.diffuseAlbedo, "surfaceSpecAlbedoRGB": self.specAlbedo,
                "surfaceRoughness": self.roughness, "surfaceIor": self.ior, "subsurface": self.subsurface,
                "subsurfaceAlbedoRGB": self.subsurfaceAlbedoRGB, "subsurfaceMFP": self.subsurfaceMFP,
                "subsurfaceAnisotropy": self.subsurfaceAnisotropy,
                "subsurfaceDiffuseWeight": self.subsurfaceDiffuseWeight}


class Volume:
    """js/materials.js:193-275.  `extinction` is the grey extinction of a
    scene's SDF_VOLUME medium (the property the reference's volume scenes set,
    e.g. exampleScenes/volumes/pyroclastic.html: volume.extinction), passed to
    VOLUME_EXTINCTION / VOLUME_EXTINCTION_MAX as the uniform volumeExtinction;
    scatteringColor / absorptionColor / anisotropy then describe that medium."""

    def __init__(self):
        self.atmosphereMinX, self.atmosphereMaxX = -10.0, 10.0
        self.atmosphereMinY, self.atmosphereMaxY = 0.0, 10.0
        self.atmosphereMinZ, self.atmosphereMaxZ = -10.0, 10.0
        self.lof10_mfp = 1.0
        self.scatteringColor = [0.0, 0.0, 0.0]
        self.absorptionColor = [0.0, 0.0, 0.0]
        self.anisotropy = 0.0
        self.emission = 0.0
        self.emissionColor = [1.0, 1.0, 1.0]
        self.fogEnable = False
        self.fogMFP = 0.0
        self.fogMaxOpticalDepth = 5.0
        self.fogEmission = [1.0, 1.0, 1.0]
        self.fogTint = [0.5, 0.5, 0.5]
        self.extinction = 1.0

    def uniforms(self, s):
        ext = [a + b for a, b in zip(self.absorptionColor, self.scatteringColor)]
        return {"atmosphereMFP": 10.0 ** self.lof10_mfp, "atmosphereExtinctionCoeffRGB": ext,
                "atmosphereScatteringCoeffRGB": list(self.scatteringColor),
                "atmosphereAnisotropy": self.anisotropy,
                "atmosphereBoundsMin": [self.atmosphereMinX, self.atmosphereMinY, self.atmosphereMinZ],
                "atmosphereBoundsMax": [self.atmosphereMaxX, self.atmosphereMaxY, self.atmosphereMaxZ],
                "volumeEmission": self.emission, "volumeEmissionColorRGB": self.emissionColor,
                "fogEnable": int(bool(self.fogEnable)), "fogMFP": 10.0 ** self.fogMFP,
                "fogMaxOpticalDepth": self.fogMaxOpticalDepth, "fogEmission": self.fogEmission,
                "fogTint": self.fogTint, "volumeExtinction": self.extinction}


class Metal:
    """Metal base (js/materials.js:479-501)."""

    def __init__(self, name):
        self.name = name
        self.roughness = 0.02
        self.specAlbedo = [1.0, 1.0, 1.0]

    def ior(self):
        raise NotImplementedError

    def uniforms(self, s):
        return {"metalRoughness": self.roughness, "metalSpecAlbedoRGB": self.specAlbedo}

    def textures(self):
        return {}


class TabulatedMetal(Metal):
    """js/materials.js:631-668: n and k from 64-texel linear textures.  The
    reference binds k under the name "k_tex" while the shader samples "kTex"
    (js/materials.js:665), so K_METAL reads texture unit 0; this host keeps that
    binding: the runtime points a sampler the host never bound at a per-frame
    snapshot of the radiance buffer, as GL texture unit 0 holds it."""

    def __init__(self, name, n, k):
        super().__init__(name)
        self.n = np.asarray(n, np.float32)
        self.k = np.asarray(k, np.float32)

    def ior(self):
        delta = repr(1.0 / (750.0 - 390.0))
        return f"""
float IOR_METAL(float wavelength_nm)
{{
    float u = (wavelength_nm - 390.0) * {delta};
    return texture(iorTex, vec2(u, 0.5)).r;
}}
float K_METAL(float wavelength_nm)
{{
    float u = (wavelength_nm - 390.0) * {delta};
    return texture(kTex, vec2(u, 0.5)).r;
}}
    """

    def textures(self):
        return {"iorTex": (self.n, 64, 1), "k_tex": (self.k, 64, 1)}


class GlslMetal(Metal):
    """A metal whose IOR_METAL / K_METAL are given as GLSL, the way a scene adds
    its own Metal subclass through Materials.addMetal (js/materials.js:1245-1248)."""

    def __init__(self, name, glsl):
        super().__init__(name)
        self.glsl = glsl

    def ior(self):
        return self.glsl


ANALYTIC_METAL_GLSL = """
float IOR_METAL(float wavelength_nm) { float t = (wavelength_nm - 390.0)/360.0; return 0.2 + 0.8*t; }
float K_METAL(float wavelength_nm) { float t = (wavelength_nm - 390.0)/360.0; return 3.5 - 1.5*t; }
"""


class Dielectric:
    """Dielectric base (js/materials.js:707-762)."""

    def __init__(self, name):
        self.name = name
        self.roughness = 0.005
        self.absorptionScale = -1.0
        self.absorptionColor = [1.0, 1.0, 1.0]
        self.specAlbedo = [1.0, 1.0, 1.0]

    def uniforms(self, s):
        sc = s.lengthScale / max(self.absorptionScale, 1.0e-3)
        absorb = [sc * max(0.0, 1.0 - c) for c in self.absorptionColor]
        return {"dieleRoughness": self.roughness, "dieleSpecAlbedoRGB": self.specAlbedo, "dieleAbsorptionRGB": absorb}


def _coeff_uniforms(coeffs):
    return "".join(f"uniform float _C{n};\n" for n in range(1, len(coeffs) + 1))


class AbbeDielectric(Dielectric):
    """js/materials.js:833-868"""

    def __init__(self, name, iorVal, abbe):
        super().__init__(name)
        self.iorVal, self.abbe = iorVal, abbe

    def ior(self):
        return """
uniform float _iorVal;
uniform float _abbe;
float IOR_DIELE(float wavelength_nm)
{
    if (_abbe <= 0.0)
        return _iorVal;
    float nD2 = 0.5893; nD2 *= nD2;
    float nF2 = 0.4861; nF2 *= nF2;
    float nC2 = 0.6563; nC2 *= nC2;
    float cauchyC = (_iorVal - 1.0)/_abbe * (nC2*nF2)/(nC2 - nF2);
    float cauchyB = _iorVal - cauchyC/nD2;
    float wavelength_um = 1.0e-3*wavelength_nm;
    float l = wavelength_um;
    return cauchyB + cauchyC/(l*l);
}
"""

    def uniforms(self, s):
        u = super().uniforms(s)
        u.update({"_iorVal": self.iorVal, "_abbe": self.abbe})
        return u


class SellmeierDielectric(Dielectric):
    """Model-1 Sellmeier with squared poles (js/materials.js:897-939)."""

    squared = True

    def __init__(self, name, coeffs):
        super().__init__(name)
        self.coeffs = list(coeffs)

    def ior(self):
        terms = (len(self.coeffs) - 1) // 2
        f = "1.0 + _C1 "
        for t in range(1, terms + 1):
            pole = f"_C{2*t+1}*_C{2*t+1}" if self.squared else f"_C{2*t+1}"
            f += f"+ _C{2*t}*l2/(l2 - {pole})"
        return _coeff_uniforms(self.coeffs) + f"""
float IOR_DIELE(float wavelength_nm)
{{
    float wavelength_um = 1.0e-3*wavelength_nm;
    float l2 = wavelength_um*wavelength_um;
    float n2 = {f};
    return max(sqrt(abs(n2)), 1.0e-3);
}}"""

    def uniforms(self, s):
        u = super().uniforms(s)
        u.update({f"_C{n}": c for n, c in enumerate(self.coeffs, 1)})
        return u


class Sellmeier2Dielectric(SellmeierDielectric):
    """Model-2 Sellmeier, poles in um^2 (js/materials.js:948-990)."""

    squared = False


class PolyanskiyDielectric(Dielectric):
    """js/materials.js:1002-1044"""

    def __init__(self, name, coeffs):
        super().__init__(name)
        self.coeffs = list(coeffs)

    def ior(self):
        return _coeff_uniforms(self.coeffs) + """
float IOR_DIELE(float wavelength_nm)
{
    float wavelength_um = 1.0e-3*wavelength_nm;
    float l = wavelength_um;
    float n2 =  _C1 + _C2*pow(l, _C3)/(l*l - pow(_C4, _C5));
    return max(sqrt(abs(n2)), 1.0e-3);
}"""

    def uniforms(self, s):
        u = super().uniforms(s)
        u.update({f"_C{n}": c for n, c in enumerate(self.coeffs, 1)})
        return u


class CauchyDielectric(Dielectric):
    """js/materials.js:1056-1098"""

    def __init__(self, name, coeffs):
        super().__init__(name)
        self.coeffs = list(coeffs)

    def ior(self):
        terms = (len(self.coeffs) - 1) // 2
        f = "_C1"
        for t in range(1, terms + 1):
            f += f" + _C{2*t}*pow(l, _C{2*t+1})"
        return _coeff_uniforms(self.coeffs) + f"""
float IOR_DIELE(float wavelength_nm)
{{
    float wavelength_um = 1.0e-3*wavelength_nm;
    float l = wavelength_um;
    float n = {f};
    return max(n, 1.0e-3);
}}"""

    def uniforms(self, s):
        u = super().uniforms(s)
        u.update({f"_C{n}": c for n, c in enumerate(self.coeffs, 1)})
        return u


# the registry of js/materials.js:1125-1147 (coefficients as coded there)
DIELECTRICS = {
    "Abbe dielectric": lambda: AbbeDielectric("Abbe dielectric", 1.5, 0.0),
    "Glass (BK7)": lambda: SellmeierDielectric("Glass (BK7)", [0.0, 1.03961212, 0.00600069867, 0.231792344, 0.0200179144, 1.01046945, 103.560653]),
    "Glass (K7)": lambda: Sellmeier2Dielectric("Glass (K7)", [0.0, 1.1273555, 0.00720341707, 0.124412303, 0.0269835916, 0.827100531, 100.384588]),
    "Glass (F5)": lambda: Sellmeier2Dielectric("Glass (F5)", [0.0, 1.3104463, 0.00958633048, 0.19603426, 0.0457627627, 0.96612977, 115.011883]),
    "Glass (LAFN7)": lambda: Sellmeier2Dielectric("Glass (LAFN7)", [0.0, 1.66842615, 0.0103159999, 0.298512803, 0.0469216348, 1.0774376, 82.5078509]),
    "Glass (LASF35)": lambda: Sellmeier2Dielectric("Glass (LASF35)", [0.0, 2.45505861, 0.0135670404, 0.453006077, 0.054580302, 2.3851308, 167.904715]),
    "Glass (N-LAK33A)": lambda: Sellmeier2Dielectric("Glass (N-LAK33A)", [0.0, 1.44116999, 0.00680933877, 0.571749501, 0.0222291824, 1.16605226, 80.9379555]),
    "Glass (N-FK51A)": lambda: SellmeierDielectric("Glass (N-FK51A)", [0.0, 0.97124781, 0.00472301995, 0.216901417, 0.0153575612, 0.90465166, 168.68133]),
    "Glass (SF4)": lambda: Sellmeier2Dielectric("Glass (SF4)", [0.0, 1.61957826, 0.0125502104, 0.339493189, 0.0544559822, 1.02566931, 117.652222]),
    "Glass (SF67)": lambda: Sellmeier2Dielectric("Glass (SF67)", [0.0, 1.97464225, 0.0145772324, 0.467095921, 0.0669790359, 2.43154209, 157.444895]),
    "Water": lambda: Sellmeier2Dielectric("Water", [0.0, 5.67252e-1, 5.08555046e-3, 1.736581e-1, 1.8149386e-2, 2.12153e-2, 2.61726e-2, 1.1384932e-1, 1.073888e1]),
    "Ethanol": lambda: Sellmeier2Dielectric("Ethanol", [0.0, 0.83189, 0.00930, -0.15582, -49.45200]),
    "Polycarbonate": lambda: Sellmeier2Dielectric("Polycarbonate", [0.0, 0.83189, 0.00930, -0.15582, -49.45200]),
    "Glycerol": lambda: CauchyDielectric("Glycerol", [1.45797, 0.00598, -2, -0.00036, -4]),
    "Liquid Crystal (E7)": lambda: CauchyDielectric("Liquid Crystal (E7)", [1.4990, 0.0072, -2, 0.0003, -4]),
    "Diamond": lambda: SellmeierDielectric("Diamond", [0.0, 0.3306, 0.175, 4.3356, 0.1060]),
    "Quartz": lambda: SellmeierDielectric("Quartz", [0.0, 0.6961663, 0.0684043, 0.4079426, 0.1162414, 0.8974794, 9.896161]),
    "Fused Silica": lambda: SellmeierDielectric("Fused Silica", [0.0, 0.6961663, 0.0684043, 0.4079426, 0.1162414, 0.8974794, 9.896161]),
    "Sapphire": lambda: SellmeierDielectric("Sapphire", [0.0, 1.5039759, 0.0740288, 0.55069141, 0.1216529, 6.5927379, 20.072248]),
    "Sodium Chloride": lambda: SellmeierDielectric("Sodium Chloride", [0.00055, 0.19800, 0.050, 0.48398, 0.100, 0.38696, 0.128]),
    "Proustite": lambda: PolyanskiyDielectric("Proustite", [7.483, 0.474, 0.0, 0.09, 1.0]),
    "Rutile": lambda: PolyanskiyDielectric("Rutile", [5.913, 0.2441, 0.0, 0.0803, 1.0]),
    "Silver Chloride": lambda: PolyanskiyDielectric("Silver Chloride", [4.00804, 0.079086, 0.0, 0.04584, 1.0]),
}


def metal(name):
    t = tables()["metals"][name]
    return TabulatedMetal(name, t["n"], t["k"])


class Materials:
    """js/materials.js:1117-1341 (defaults: BK7 + Aluminium)."""

    def __init__(self):
        self.surface = Surface()
        self.volume = Volume()
        self.dielectric = DIELECTRICS["Glass (BK7)"]()
        self.metal = metal("Aluminium")

    def loadDielectric(self, name):
        self.dielectric = DIELECTRICS[name]()
        return self.dielectric

    def loadMetal(self, name):
        self.metal = metal(name)
        return self.metal


# ---------------------------------------------------------------- camera
class Camera:
    """The parts of THREE.PerspectiveCamera + OrbitControls the renderer reads."""

    def __init__(self):
        self.position = [1.0, 1.0, 1.0]
        self.target = [0.0, 0.0, 0.0]
        self.up = [0.0, 1.0, 0.0]
        self.fov = 45.0
        self.aspect = 1.0
        self.aperture = -10.0
        self.focalDistance = 1.0

    def controls_position(self):
        """The position the reference renders its first frame from.  Its
        OrbitControls.update (js/thirdparty/three/controls/OrbitControls.js:
        269-321) re-derives the camera position from spherical coordinates
        about the target (polar angle kept within [EPS, pi - EPS], EPS = 1e-6),
        and a page runs it twice before that frame: in Snelly._resize (or the
        constructor's Snelly.resize) and at the end of Snelly.initScene
        (js/snelly.js:80, 274-279, 291, 605).  The controls' up-space rotation
        is the identity: they are made while camera.up is still (0, 1, 0)."""
        eps = 0.000001
        p = [float(x) for x in self.position]
        t = [float(x) + 0.0 for x in self.target]  # target.add(pan) with pan = 0
        for _ in range(2):
            ox, oy, oz = p[0] - t[0], p[1] - t[1], p[2] - t[2]
            theta = math.atan2(ox, oz)
            phi = math.atan2(math.sqrt(ox * ox + oz * oz), oy)
            phi = max(eps, min(math.pi - eps, max(0.0, min(math.pi, phi))))
            radius = math.sqrt(ox * ox + oy * oy + oz * oz)
            p = [t[0] + radius * math.sin(phi) * math.sin(theta), t[1] + radius * math.cos(phi),
                 t[2] + radius * math.sin(phi) * math.cos(theta)]
        return p

    def basis(self, position=None):
        """(camPos, camDir, camX, camY) as js/renderer.js:520-527 derives them
        from the THREE.PerspectiveCamera that OrbitControls points at its
        target, through three.js r79's chain (js/thirdparty/three/three.js):
        Camera.lookAt sets the quaternion from Matrix4.lookAt's rotation;
        getWorldDirection rebuilds the world matrix from it (compose), takes the
        quaternion back out (decompose) and turns (0, 0, -1) by it; camY is `up`
        carried through the world matrix and normalised (transformDirection),
        camX = camY x camDir.  Vector arithmetic is in doubles, as JavaScript
        computes it, but every Matrix4 element is stored in a Float32Array, so
        each matrix write rounds to float32."""
        p = [float(c) for c in (self.position if position is None else position)]
        t = [float(c) for c in self.target]
        up = [float(c) for c in self.up]
        q = _three_quaternion_from_rotation(_three_look_at(p, t, up))
        world = _three_compose(q)
        cam_dir = _three_apply_quaternion([0.0, 0.0, -1.0], _three_decompose_quaternion(world))
        cam_y = _three_transform_direction(up, world)
        cam_x = _three_cross(cam_y, cam_dir)
        return (np.array(p, np.float64), np.array(cam_dir, np.float64), np.array(cam_x, np.float64),
                np.array(cam_y, np.float64))


# three.js r79's camera chain (js/thirdparty/three/three.js).  Matrix4.elements
# is a Float32Array: `_f32` is a store into it.
def _f32(v):
    return float(np.float32(v))


def _three_normalize(v):
    """Vector3.normalize (2460-2464): times 1 / length (NaN for a zero vector)"""
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    with np.errstate(divide="ignore", invalid="ignore"):
        s = 1.0 / n if n != 0.0 else math.inf
    return [c * s if not (math.isinf(s) and c == 0.0) else math.nan for c in v]


def _three_cross(a, b):
    """Vector3.crossVectors (2393-2402)"""
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _three_look_at(eye, target, up):
    """Matrix4.lookAt (5105-5150): the rotation part of a fresh matrix"""
    z = _three_normalize([eye[k] - target[k] for k in range(3)])
    if z[0] * z[0] + z[1] * z[1] + z[2] * z[2] == 0:
        z[2] = 1.0
    x = _three_normalize(_three_cross(up, z))
    if x[0] * x[0] + x[1] * x[1] + x[2] * x[2] == 0:
        z[2] += 0.0001
        x = _three_normalize(_three_cross(up, z))
    y = _three_cross(z, x)
    te = [1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0]
    te[0], te[4], te[8] = _f32(x[0]), _f32(y[0]), _f32(z[0])
    te[1], te[5], te[9] = _f32(x[1]), _f32(y[1]), _f32(z[1])
    te[2], te[6], te[10] = _f32(x[2]), _f32(y[2]), _f32(z[2])
    return te


def _three_quaternion_from_rotation(te):
    """Quaternion.setFromRotationMatrix (1048-1105)"""
    m11, m12, m13 = te[0], te[4], te[8]
    m21, m22, m23 = te[1], te[5], te[9]
    m31, m32, m33 = te[2], te[6], te[10]
    trace = m11 + m22 + m33
    if trace > 0:
        s = 0.5 / math.sqrt(trace + 1.0)
        return [(m32 - m23) * s, (m13 - m31) * s, (m21 - m12) * s, 0.25 / s]
    if m11 > m22 and m11 > m33:
        s = 2.0 * math.sqrt(1.0 + m11 - m22 - m33)
        return [0.25 * s, (m12 + m21) / s, (m13 + m31) / s, (m32 - m23) / s]
    if m22 > m33:
        s = 2.0 * math.sqrt(1.0 + m22 - m11 - m33)
        return [(m12 + m21) / s, 0.25 * s, (m23 + m32) / s, (m13 - m31) / s]
    s = 2.0 * math.sqrt(1.0 + m33 - m11 - m22)
    return [(m13 + m31) / s, (m23 + m32) / s, 0.25 * s, (m21 - m12) / s]


def _three_compose(q):
    """Object3D.updateMatrix (8999-9005) = Matrix4.compose (5585-5593):
    makeRotationFromQuaternion (5068-5100), scale by (1, 1, 1), setPosition.
    The translation column is not read by the chain."""
    x, y, z, w = q
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz = x * x2, x * y2, x * z2
    yy, yz, zz = y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    te = [0.0] * 16
    te[0], te[4], te[8] = _f32(1 - (yy + zz)), _f32(xy - wz), _f32(xz + wy)
    te[1], te[5], te[9] = _f32(xy + wz), _f32(1 - (xx + zz)), _f32(yz - wx)
    te[2], te[6], te[10] = _f32(xz - wy), _f32(yz + wx), _f32(1 - (xx + yy))
    te[15] = 1.0
    return te


def _three_decompose_quaternion(te):
    """the quaternion of Matrix4.decompose (5595-5650); the sign test is the
    n44 term of Matrix4.determinant (5290-5336), the others being zero"""
    sx = math.sqrt(te[0] * te[0] + te[1] * te[1] + te[2] * te[2])
    sy = math.sqrt(te[4] * te[4] + te[5] * te[5] + te[6] * te[6])
    sz = math.sqrt(te[8] * te[8] + te[9] * te[9] + te[10] * te[10])
    n11, n12, n13 = te[0], te[4], te[8]
    n21, n22, n23 = te[1], te[5], te[9]
    n31, n32, n33 = te[2], te[6], te[10]
    det = (-n13 * n22 * n31 - n11 * n23 * n32 + n11 * n22 * n33 + n13 * n21 * n32 - n12 * n21 * n33
           + n12 * n23 * n31)
    if det < 0:
        sx = -sx
    m = list(te)
    for k, inv in ((0, 1 / sx), (4, 1 / sy), (8, 1 / sz)):
        m[k], m[k + 1], m[k + 2] = _f32(m[k] * inv), _f32(m[k + 1] * inv), _f32(m[k + 2] * inv)
    return _three_quaternion_from_rotation(m)


def _three_apply_quaternion(v, q):
    """Vector3.applyQuaternion (2240-2260)"""
    x, y, z = v
    qx, qy, qz, qw = q
    ix = qw * x + qy * z - qz * y
    iy = qw * y + qz * x - qx * z
    iz = qw * z + qx * y - qy * x
    iw = -qx * x - qy * y - qz * z
    return [ix * qw + iw * -qx + iy * -qz - iz * -qy,
            iy * qw + iw * -qy + iz * -qx - ix * -qz,
            iz * qw + iw * -qz + ix * -qy - iy * -qx]


def _three_transform_direction(v, te):
    """Vector3.transformDirection (2292-2306)"""
    x, y, z = v
    return _three_normalize([te[0] * x + te[4] * y + te[8] * z, te[1] * x + te[5] * y + te[9] * z,
                             te[2] * x + te[6] * y + te[10] * z])


# ---------------------------------------------------------------- renderer
class Renderer:
    """Renderer defaults (js/renderer.js:78-126)."""

    def __init__(self):
        self.renderMode = "pt"
        self.dispersive = False
        self.maxSamplesPerFrame = 1
        self.maxSpp = 1024
        self.maxBounces = 3
        self.maxAtmosphereScatters = 1
        self.maxMarchSteps = 256
        self.maxSSSSteps = 128
        self.radianceClamp = 3.0
        self.wavelengthSamples = 256
        self.filterRadius = 2.0
        self.goalFPS = 20.0
        self.shadowStrength = 1.0
        self.maxStepsIsMiss = True
        self.interactive = True
        self.skipProbability = 0.0
        self.skyPower = 0.0
        self.skyTemperature = 6000.0
        self.skyTintUp = [1.0, 1.0, 1.0]
        self.skyTintDown = [1.0, 1.0, 1.0]
        self.envMapVisible = True
        self.envMapPhiRotation = 0.0
        self.envMapThetaRotation = 0.0
        self.envMapTransitionAngle = 135.0
        self.sunPower = -7.0
        self.sunAngularSize = 5.0
        self.sunColor = [1.0, 1.0, 1.0]
        self.sunLatitude = 50.0
        # the reference initialises `sunLongitudef` (js/renderer.js:113), leaving
        # sunLongitude undefined (NaN in the sun direction) unless a scene sets it
        self.sunLongitude = float("nan")
        self.sunVisibleDirectly = True
        self.sphereLightRadius = 1.0
        self.sphereLightPower = -7.0
        self.sphereLightColor = [1.0, 1.0, 1.0]
        self.sphereLightPosition = [0.0, 0.0, 0.0]
        self.exposure = 0.0
        self.gamma = 2.2
        self.contrast = 1.0
        self.saturation = 1.0
        self.hueShift = 0.0

    def sun_dir(self):
        """Renderer.updateSunDir (js/renderer.js:371-383)"""
        lat_theta = (90.0 - self.sunLatitude) * math.pi / 180.0
        lon_phi = self.sunLongitude * math.pi / 180.0
        st, ct = math.sin(lat_theta), math.cos(lat_theta)
        return [st * math.cos(lon_phi), ct, st * math.sin(lon_phi)]


class Scales:
    """Snelly.initScene length scales (js/snelly.js:218-237)"""

    def __init__(self, lengthScale=1.0, minLengthScale=1.0e-4, maxLengthScale=1.0e2):
        self.lengthScale = lengthScale
        self.minLengthScale = max(1.0e-12, minLengthScale)
        self.maxLengthScale = min(1.0e9, maxLengthScale)


def has_hook(shader, name):
    return shader.find(name + "(") != -1


def reads_texture_unit0(scene):
    """whether the scene's program samples texture unit 0 -- the radiance
    being rendered -- and so needs a per-frame snapshot of it: a tabulated
    metal's K_METAL reads kTex, which the reference never binds
    (js/materials.js:664-665), and only a scene with a metal calls it"""
    return isinstance(scene.materials.metal, TabulatedMetal) and has_hook(scene.glsl, "SDF_METAL")


def assemble(scene_glsl, r, mats):
    """Renderer.compileShaders (js/renderer.js:217-316): inject default hooks,
    derive the feature #defines, return (shader, ior_code, defines)."""
    shader = scene_glsl
    if shader.find("INIT(") == -1:
        shader += "\n void INIT() {}\n"
    custom_cam = has_hook(shader, "CONSTRUCT_PRIMARY_RAY")
    has_surface = has_hook(shader, "SDF_SURFACE")
    if not has_hook(shader, "SURFACE_DIFFUSE_REFLECTANCE"):
        shader += "\n vec3 SURFACE_DIFFUSE_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V) { return C; }\n"
    if not has_hook(shader, "SURFACE_SPECULAR_REFLECTANCE"):
        shader += "\n vec3 SURFACE_SPECULAR_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V) { return C; }\n"
    if not has_hook(shader, "SURFACE_ROUGHNESS"):
        shader += "\n float SURFACE_ROUGHNESS(in float roughness, in vec3 X, in vec3 N) { return roughness; }\n"
    surface_nm = has_hook(shader, "SURFACE_NORMAL_MAP")
    if not has_hook(shader, "SUBSURFACE_ALBEDO"):
        shader += "\n vec3 SUBSURFACE_ALBEDO(in vec3 C, in vec3 X, in vec3 N) { return C; }\n"
    has_metal = has_hook(shader, "SDF_METAL")
    if not has_hook(shader, "METAL_SPECULAR_REFLECTANCE"):
        shader += "\n vec3 METAL_SPECULAR_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V) { return C; }\n"
    if not has_hook(shader, "METAL_ROUGHNESS"):
        shader += "\n float METAL_ROUGHNESS(in float roughness, in vec3 X, in vec3 N) { return roughness; }\n"
    metal_nm = has_hook(shader, "METAL_NORMAL_MAP")
    has_diele = has_hook(shader, "SDF_DIELECTRIC")
    if not has_hook(shader, "DIELECTRIC_SPECULAR_REFLECTANCE"):
        shader += "\n vec3 DIELECTRIC_SPECULAR_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V) { return C; }\n"
    if not has_hook(shader, "DIELECTRIC_ROUGHNESS"):
        shader += "\n float DIELECTRIC_ROUGHNESS(in float roughness, in vec3 X, in vec3 N) { return roughness; }\n"
    diele_nm = has_hook(shader, "DIELECTRIC_NORMAL_MAP")
    if not has_hook(shader, "DIELECTRIC_ABSORPTION"):
        shader += "\n vec3 DIELECTRIC_ABSORPTION(vec3 absorption_ui, vec3 X) { return absorption_ui; }\n"
    has_vol_emission = True
    for hook, code in [
        ("VOLUME_EXTINCTION", "float VOLUME_EXTINCTION(float extinction_ui, vec3 X) { return extinction_ui; }"),
        ("VOLUME_EXTINCTION_MAX", "float VOLUME_EXTINCTION_MAX(float extinction_ui) { return extinction_ui; }"),
        ("VOLUME_SCATTERING_COLOR", "vec3 VOLUME_SCATTERING_COLOR(vec3 scattering_color_ui, vec3 X) { return scattering_color_ui; }"),
        ("VOLUME_ABSORPTION_COLOR", "vec3 VOLUME_ABSORPTION_COLOR(vec3 absorption_color_ui, vec3 X) { return absorption_color_ui; }"),
        ("VOLUME_ANISOTROPY", "float VOLUME_ANISOTROPY(float anisotropy, vec3 X) { return anisotropy; }"),
    ]:
        if not has_hook(shader, hook):
            shader += f"\n {code}\n"
    if not has_hook(shader, "VOLUME_EMISSION"):
        has_vol_emission = False
        shader += "\n vec3 VOLUME_EMISSION(vec3 emission, vec3 X) { return emission; }\n"
    has_sphere_light = r.sphereLightPower > -7.0
    has_geometry = has_surface or has_metal or has_diele
    # the reference tests `hasGeometry || hasVolume` without declaring hasVolume
    # (js/renderer.js:267-271); here a scene that defines SDF_VOLUME has the
    # heterogeneous medium (HAS_VOLUME), which replaces the homogeneous
    # atmosphere: the Volume material's colours then describe the medium
    has_volume = has_hook(shader, "SDF_VOLUME")
    if not (has_geometry or has_volume):
        raise native.SnellyError("Scene must define at least one of: SDF_SURFACE, SDF_METAL, SDF_DIELECTRIC, or SDF_VOLUME")
    v = mats.volume
    has_atm = not has_volume and (any(c > 0.0 for c in v.scatteringColor) or any(c > 0.0 for c in v.absorptionColor))
    defines = ""
    flags = [(custom_cam, "HAS_CUSTOM_CAMERA"), (has_surface, "HAS_SURFACE"), (has_metal, "HAS_METAL"),
             (has_diele, "HAS_DIELECTRIC"), (has_atm, "HAS_ATMOSPHERE"), (has_volume, "HAS_VOLUME"),
             (v.fogEnable, "HAS_FOG"),
             (has_vol_emission, "HAS_VOLUME_EMISSION"), (has_geometry, "HAS_GEOMETRY"),
             (surface_nm or metal_nm or diele_nm, "HAS_NORMALMAP"), (surface_nm, "HAS_SURFACE_NORMALMAP"),
             (metal_nm, "HAS_METAL_NORMALMAP"), (diele_nm, "HAS_DIELECTRIC_NORMALMAP"),
             (r.interactive, "INTERACTIVE_MODE"), (r.dispersive, "DISPERSION_ENABLED"),
             (has_sphere_light, "HAS_SPHERE_LIGHT")]
    for on, name in flags:
        if on:
            defines += f"\n#define {name}\n"
    ior = mats.dielectric.ior() + "\n" + mats.metal.ior()
    return shader, ior, defines


def frame_uniforms(r, cam, s, mats, width, height):
    """The uniform values Renderer.render uploads each frame
    (js/renderer.js:517-584 and Materials.syncShader, js/materials.js:1335-1341)."""
    pos, camDir, camX, camY = cam.basis(cam.controls_position())
    ls = s.lengthScale
    u = {
        "camPos": pos, "camDir": camDir, "camX": camX, "camY": camY,
        "camFovy": cam.fov, "camAspect": cam.aspect,
        "camAperture": ls * 1.4142135 ** cam.aperture if cam.aperture != -math.inf else 0.0,
        "camFocalDistance": ls * 10.0 ** cam.focalDistance,
        "resolution": [width, height],
        "filterRadius": r.filterRadius, "radianceClamp": 10.0 ** r.radianceClamp,
        "skipProbability": r.skipProbability,
        "lengthScale": max(ls, 1e-6), "maxLengthScale": max(s.maxLengthScale, 1e-6),
        "minLengthScale": max(s.minLengthScale, 1e-6), "shadowStrength": r.shadowStrength,
        "maxSSSSteps": int(r.maxSSSSteps), "maxStepsIsMiss": int(bool(r.maxStepsIsMiss)),
        "wavelengthSamples": int(r.wavelengthSamples),
        "skyPower": 10.0 ** r.skyPower, "skyTintUp": r.skyTintUp, "skyTintDown": r.skyTintDown,
        "envMapVisible": int(bool(r.envMapVisible)),
        "envMapPhiRotation": min(max(r.envMapPhiRotation, 0.0), 360.0),
        "envMapThetaRotation": min(max(r.envMapThetaRotation, 0.0), 180.0),
        "envMapTransitionAngle": min(max(r.envMapTransitionAngle, 0.0), 180.0),
        "sunPower": 10.0 ** r.sunPower, "sunAngularSize": r.sunAngularSize, "sunLatitude": r.sunLatitude,
        "sunLongitude": r.sunLongitude, "sunColor": r.sunColor, "sunDir": r.sun_dir(),
        "sunVisibleDirectly": int(bool(r.sunVisibleDirectly)),
        "sphereLightPosition": [x * ls for x in r.sphereLightPosition],
        "sphereLightRadius": r.sphereLightRadius * ls,
        "sphereLightPower": 10.0 ** r.sphereLightPower / (4.0 * math.pi * r.sphereLightRadius ** 2),
        "sphereLightColor": r.sphereLightColor,
        "haveEnvMap": 0,
        "exposure": r.exposure, "invGamma": 1.0 / r.gamma, "contrast": r.contrast,
        "saturation": r.saturation, "hueShift": r.hueShift,
    }
    u.update(mats.metal.uniforms(s))
    u.update(mats.dielectric.uniforms(s))
    u.update(mats.surface.uniforms(s))
    u.update(mats.volume.uniforms(s))
    return u


INT_UNIFORMS = {"maxSSSSteps", "maxStepsIsMiss", "wavelengthSamples", "envMapVisible", "sunVisibleDirectly",
                "haveEnvMap", "fogEnable"}


def mode_flags(renderer):
    """build flags of Renderer.renderMode (js/renderer.js:334-352): 'ao' and
    'ptsimple' select their integrators; 'pt' and 'normals' use the path
    tracer's program ('normals' renders through snl_render_normals)"""
    return {"ao": native.BUILD_AO, "ptsimple": native.BUILD_PTSIMPLE}.get(renderer.renderMode, 0)


class Session:
    """One scene on one GPU: program + accumulation state + LUT textures."""

    def __init__(self, scene, width, height, ctx=None, seed=1234, flags=0, chains=1):
        self.scene = scene
        self.ctx = ctx or native.Context(0)
        self.width, self.height = width, height
        r, mats = scene.renderer, scene.materials
        self.shader, self.ior, self.defines = assemble(scene.glsl, r, mats)
        self.prog = native.Program(self.ctx, self.shader, self.ior, self.defines, int(round(r.maxMarchSteps)),
                                   int(round(r.maxBounces)), int(round(r.maxAtmosphereScatters)),
                                   int(round(r.maxSamplesPerFrame)), flags | mode_flags(r))
        self.state = native.State(self.ctx, width, height, seed, chains)
        self.textures = {}
        self.textures["WavelengthToXYZ"] = native.Texture(self.ctx, xyz_lut(), 1024, 1, 1, 4, True, True)
        icdf = inverse_cdf(lambda wl: blackbody(wl, r.skyTemperature))
        self.textures["ICDF"] = native.Texture(self.ctx, icdf, 4096, 1, 1, 1, True, True)
        for name, (data, w, h) in mats.metal.textures().items():
            self.textures[name] = native.Texture(self.ctx, data, w, h, 1, 1, True, True)
        # the scene's environment map (Scene.envMap(), js/renderer.js:143-159):
        # (width, height, uint8 RGB rows, top row first as decoded); texture row
        # 0 is the image's first row, as WebGL uploads it without a flip
        if getattr(scene, "env_map", None) is not None:
            ew, eh, rgb8 = scene.env_map
            lin = srgb_to_linear(np.asarray(rgb8, np.uint8).reshape(eh, ew, 3))
            self.textures["envMap"] = native.Texture(self.ctx, lin, ew, eh, 1, 3, True, True)
        self.sync()

    def sync(self):
        sc = self.scene
        sc.camera.aspect = self.width / self.height
        u = frame_uniforms(sc.renderer, sc.camera, sc.scales, sc.materials, self.width, self.height)
        u["haveEnvMap"] = int("envMap" in self.textures)
        u.update(sc.user_uniforms)
        for k, v in u.items():
            self.prog.set(k, v, native.T_INT if k in INT_UNIFORMS else None)
        for k, t in self.textures.items():
            self.prog.set_texture(k, t)
        return u

    def render(self, frames):
        if self.scene.renderer.renderMode == "normals":
            self.prog.render_normals(self.state, frames)
        else:
            self.prog.render(self.state, frames)

    def radiance(self):
        return self.state.read(0)

    def save_png(self, path):
        """the tonemapped image as a PNG file"""
        _, rgba8 = self.prog.tonemap(self.state)
        save_png(path, rgba8.reshape(self.height, self.width, 4))

    def close(self):
        for t in self.textures.values():
            t.close()
        self.state.close()
        self.prog.close()
