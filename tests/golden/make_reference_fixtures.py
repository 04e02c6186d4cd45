"""Writes tests/golden/ref_<case>.npz: the reference's path tracer
(shaders/pathtracer-fragment-shader.glsl, assembled as js/renderer.js:217-359
and scripts/build do) run by a GLES 3 implementation over the cases of
reference_cases.py, from integer RNG seeds.

Development-container tooling: it reads the reference's shader sources from
/root/reference and drives reference_gles/_build/gles_render, which dlopens the
SwiftShader GLES libraries installed with the kaleido package.  Three steps
make the run follow the reference's arithmetic rather than the driver's
liberties: uninitialised locals are zero-initialised, as under WebGL; the
transcendental builtins are replaced by accurate ones, and min/max/clamp by
the GLSL definitions (reference_gles/precise_math.glsl); for atmosphere
programs cameraPath's two `continue` statements are restated as a skip flag
(without_continue).  Each step is checked under the GLES implementation before
any fixture is written.  The fixtures it writes are
data (initial state is implied by the seed; outputs are the
accumulated radiance and the final RNG state).

    python tests/golden/make_reference_fixtures.py [case ...]
    python tests/golden/make_reference_fixtures.py --time   # the reference's CPU rate
"""
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

from snelly_amd import host  # noqa: E402
import oracle_py  # noqa: E402
import reference_cases  # noqa: E402

REF_SHADERS = "/root/reference/shaders"
GLES_DIR = "/usr/local/lib/python3.10/dist-packages/kaleido/executable/bin/swiftshader"
HARNESS = os.path.join(HERE, "reference_gles", "_build", "gles_render")


def shader_source(name):
    """scripts/build: '#version 300 es' + newline + the stripped file's lines"""
    with open(os.path.join(REF_SHADERS, name + ".glsl")) as f:
        lines = f.read().strip().split("\n")
    return "#version 300 es\n" + "".join(l.lstrip("\n") + "\n" for l in lines)


ZERO_INIT_TYPES = ("float", "int", "bool", "vec2", "vec3", "vec4", "ivec2", "ivec3", "ivec4", "mat2", "mat3", "mat4",
                   "RadianceType")
_DECL = re.compile(r"^(\s*)(" + "|".join(ZERO_INIT_TYPES) + r")\s+([A-Za-z_]\w*(?:\s*,\s*[A-Za-z_]\w*)*)\s*;(.*)$")


def zero_init_locals(src):
    """Give every uninitialised local variable of a built-in type an explicit
    zero initialiser, as WebGL implementations do (ANGLE's shader translator
    initialises uninitialised locals for WebGL contexts).  The reference reads
    some before writing them -- woutputL and pdfOut in sampleDielectric's
    total-internal-reflection exit (pathtracer-fragment-shader.glsl:590-594) --
    and a GLES driver is free to leave them undefined."""
    out, depth, struct_depth = [], 0, None
    for line in src.split("\n"):
        code = line.split("//", 1)[0]
        m = _DECL.match(line)
        if m and depth >= 1 and struct_depth is None:
            ind, ty, names, rest = m.groups()
            decl = ", ".join(f"{n.strip()} = {ty}(0.0)" for n in names.split(","))
            line = f"{ind}{ty} {decl};{rest}"
        if re.search(r"\bstruct\b", code) and struct_depth is None:
            struct_depth = depth
        depth += code.count("{") - code.count("}")
        if struct_depth is not None and depth <= struct_depth and "}" in code:
            struct_depth = None
        out.append(line)
    return "\n".join(out)


def without_continue(src):
    """Restate cameraPath's two `continue` statements (an atmosphere scatter,
    pathtracer-fragment-shader.glsl:1706 and 1767) as a flag that skips the
    rest of the loop body.  Each `continue` is the last statement of blocks
    that end right after it, so `skip = true` there, `if (!skip) break` for the
    miss branch's closing `break` and `if (!skip) { ... }` around the code that
    follows are the same control flow.  SwiftShader 4.1 mis-executes a
    `continue` in a loop whose body has a `break` after it (the loop ends
    early, or runs on with wrong state); check_loop_control shows it on a
    minimal loop and checks the flag form before any fixture is written."""
    import re
    head = re.search(r"for \(int vertex=0; vertex<=\d+; \+\+vertex\)\n    \{\n", src)
    assert head, "cameraPath loop not found"
    body = src[head.end():]
    pieces = [
        ("                    continue;\n", "                    snl_skip = true;\n"),
        ("            // Ray escapes to infinity\n            break;\n",
         "            // Ray escapes to infinity\n            if (!snl_skip) break;\n"),
        ("        RadianceType Tr = RadianceType(1.0);", "        if (!snl_skip) {\n        RadianceType Tr = RadianceType(1.0);"),
        ("                continue;\n", "                snl_skip = true;\n"),
        ("        if (hitSphereLight)\n            break;", "        if (!snl_skip) {\n        if (hitSphereLight)\n            break;"),
        ("        if (maxComponent(throughput) < THROUGHPUT_EPSILON)\n            break;\n    }\n    return L;",
         "        if (maxComponent(throughput) < THROUGHPUT_EPSILON)\n            break;\n        }}\n    }\n    return L;"),
    ]
    out, rest = [], body
    for old, new in pieces:
        at = rest.find(old)
        assert at >= 0, f"anchor not found: {old!r}"
        out.append(rest[:at] + new)
        rest = rest[at + len(old):]
    assert "continue;" not in "".join(out)
    return src[:head.end()] + "        bool snl_skip = false;\n" + "".join(out) + rest


def program_source(scene, program="pathtracer"):
    """Renderer.compileShaders' replacements (js/renderer.js:286-316), applied in
    the order the replacements object lists them, to the vertex and fragment
    shaders of `program` (pathtracer, normals or pick)"""
    r, mats = scene.renderer, scene.materials
    shader, ior, defines = host.assemble(scene.glsl, r, mats)
    rep = [("__SHADER__", shader), ("__IOR_FUNC__", ior),
           ("__MAX_MARCH_STEPS__", str(int(round(r.maxMarchSteps)))),
           ("__MAX_BOUNCES__", str(int(round(r.maxBounces)))),
           ("__MAX_ATMOSPHERE_SCATTERS__", str(int(round(r.maxAtmosphereScatters)))),
           ("__MAX_SAMPLES_PER_FRAME__", str(int(round(r.maxSamplesPerFrame)))),
           ("__DEFINES__", defines)]
    v, f = shader_source(program + "-vertex-shader"), shader_source(program + "-fragment-shader")
    for k, val in rep:
        v, f = v.replace(k, val), f.replace(k, val)
    if program == "pathtracer" and "#define HAS_ATMOSPHERE" in defines:
        f = without_continue(f)
    return v, with_precise_math(zero_init_locals(f))


PRECISE_MATH = os.path.join(HERE, "reference_gles", "precise_math.glsl")


def with_precise_math(src):
    """accurate transcendentals (reference_gles/precise_math.glsl) ahead of the
    shader body: the GLES implementation's own are not accurate enough for
    per-pixel comparisons"""
    head, body = src.split("\n", 1)
    assert head.startswith("#version")
    with open(PRECISE_MATH) as fp:
        pm = fp.read()
    return head + "\nprecision highp float;\nprecision highp int;\n" + pm + "\n" + body


def check_precise_math(td):
    """measure the math block, as executed by the GLES implementation, against
    float64 numpy over wide argument ranges; refuse to write fixtures if it is
    not within a few ulp"""
    W = H = 64
    n = W * H
    rs = np.random.RandomState(0)
    x = np.zeros((n, 4), np.float32)
    x[:, 0] = rs.uniform(-60, 60, n)
    x[:, 1] = rs.uniform(-80, 80, n)
    x[:, 2] = np.exp(rs.uniform(-60, 60, n)).astype(np.float32)
    x[:, 3] = rs.uniform(-1, 1, n)
    vs = ("#version 300 es\nprecision highp float;\nin vec3 Position;\nin vec2 TexCoord;\nout vec2 vTexCoord;\n"
          "void main(){gl_Position=vec4(Position,1.0);vTexCoord=TexCoord;}\n")
    fs = with_precise_math("#version 300 es\n" + """
uniform sampler2D Radiance; uniform sampler2D RngData;
in vec2 vTexCoord;
layout(location=0) out vec4 o0; layout(location=1) out vec4 o1;
void main(){
  vec4 r = texture(RngData, vTexCoord);
  o0 = vec4(sin(r.x), cos(r.x), exp(r.y), log(r.z));
  o1 = vec4(tan(r.x*0.02), pow(r.z, 0.37), atan(r.w*20.0), acos(r.w));
}
""")
    f = {k: os.path.join(td, "pm_" + k) for k in ("vs", "fs", "rad", "rng", "o0", "o1", "spec")}
    open(f["vs"], "w").write(vs)
    open(f["fs"], "w").write(fs)
    np.tile(np.float32([0, 0, 0, 1]), (n, 1)).tofile(f["rad"])
    x.tofile(f["rng"])
    open(f["spec"], "w").write(f"size {W} {H}\nframes 1\nvertex {f['vs']}\nfragment {f['fs']}\n"
                               f"state {f['rad']} {f['rng']}\nout {f['o0']} {f['o1']}\n")
    subprocess.run([HARNESS, GLES_DIR, f["spec"]], check=True, capture_output=True)
    o0 = np.fromfile(f["o0"], np.float32).reshape(-1, 4).astype(np.float64)
    o1 = np.fromfile(f["o1"], np.float32).reshape(-1, 4).astype(np.float64)
    xd = x.astype(np.float64)

    def ulp(a, b):
        return float(np.max(np.abs(a - b) / np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)))
    t_arg = (x[:, 0] * np.float32(0.02)).astype(np.float64)
    a_arg = (x[:, 3] * np.float32(20.0)).astype(np.float64)
    errs = {"sin": ulp(o0[:, 0], np.sin(xd[:, 0])), "cos": ulp(o0[:, 1], np.cos(xd[:, 0])),
            "exp": ulp(o0[:, 2], np.exp(xd[:, 1])), "log": ulp(o0[:, 3], np.log(xd[:, 2])),
            "tan": ulp(o1[:, 0], np.tan(t_arg)), "pow": ulp(o1[:, 1], xd[:, 2] ** 0.37),
            "atan": ulp(o1[:, 2], np.arctan(a_arg)), "acos": ulp(o1[:, 3], np.arccos(xd[:, 3]))}
    limits = {"tan": 4.0, "pow": 64.0}  # pow = exp(y log x): error grows with |y log x| (here up to 22)
    bad = {k: v for k, v in errs.items() if v > limits.get(k, 3.0)}
    if bad:
        raise RuntimeError(f"precise_math.glsl inaccurate under this GLES implementation: {bad}")
    return errs


LOOP_CONTINUE = """
    for (int v = 0; v <= 3; ++v) {
      count += 1;
      if (!(r.x > 0.5)) {
        acc += 0.5 * r.z;
        if (sc + 1 < 3) { if (r.y > 0.0) { sc++; acc += 1.0; continue; } }
        break;
      }
      acc += 10.0;
      if (sc + 1 < 3) { if (r.w > 0.0) { sc++; continue; } else acc += 100.0; }
      if (r.z > 0.7) break;
      acc += 1000.0;
    }"""
LOOP_FLAG = """
    for (int v = 0; v <= 3; ++v) {
      bool skip = false;
      count += 1;
      if (!(r.x > 0.5)) {
        acc += 0.5 * r.z;
        if (sc + 1 < 3) { if (r.y > 0.0) { sc++; acc += 1.0; skip = true; } }
        if (!skip) break;
      }
      if (!skip) {
      acc += 10.0;
      if (sc + 1 < 3) { if (r.w > 0.0) { sc++; skip = true; } else acc += 100.0; }
      if (!skip) {
      if (r.z > 0.7) break;
      acc += 1000.0;
      }}
    }"""


def loop_expected(r):
    """the GLSL semantics of LOOP_CONTINUE, in float32"""
    count = sc = 0
    acc = np.float32(0.0)
    for _ in range(4):
        count += 1
        if not r[0] > 0.5:
            acc = np.float32(acc + np.float32(0.5) * r[2])
            if sc + 1 < 3 and r[1] > 0.0:
                sc += 1
                acc = np.float32(acc + 1)
                continue
            break
        acc = np.float32(acc + 10)
        if sc + 1 < 3:
            if r[3] > 0.0:
                sc += 1
                continue
            acc = np.float32(acc + 100)
        if r[2] > 0.7:
            break
        acc = np.float32(acc + 1000)
    return count, sc, acc


def check_loop_control(td):
    """run LOOP_CONTINUE and its flag form LOOP_FLAG over 256 random inputs;
    the flag form must give the GLSL result everywhere (it is what
    without_continue produces).  Returns how many pixels the `continue` form
    got wrong, for the record."""
    W = H = 16
    n = W * H
    rs = np.random.RandomState(1)
    x = np.zeros((n, 4), np.float32)
    x[:, 0] = rs.uniform(0, 1, n)
    x[:, 1] = np.where(rs.uniform(0, 1, n) > 0.2, 1.0, -1.0)
    x[:, 2] = rs.uniform(0, 1, n)
    x[:, 3] = np.where(rs.uniform(0, 1, n) > 0.5, 1.0, -1.0)
    want = np.array([loop_expected(r) for r in x], np.float64)
    wrong = {}
    for tag, loop in (("continue", LOOP_CONTINUE), ("flag", LOOP_FLAG)):
        body = ("  int count = 0; int sc = 0; float acc = 0.0;" + loop +
                "\n  o0 = vec4(float(count), float(sc), acc, 0.0); o1 = r;")
        got = run_tiny(td, "lc_" + tag, body, x, W, H)[0][:, :3].astype(np.float64)
        wrong[tag] = int(np.sum(np.any(np.abs(got - want) > 1e-4 * np.maximum(1.0, np.abs(want)), axis=1)))
    if wrong["flag"]:
        raise RuntimeError(f"the flag form of the loop is mis-executed too: {wrong}")
    return wrong["continue"]


def run_tiny(td, tag, body, x, W, H, math=False):
    """run a fragment program `body` (reads vec4 r = the RNG texel, writes
    o0 and o1) over a W x H grid whose RNG state is x; returns (o0, o1)"""
    n = W * H
    vs = ("#version 300 es\nprecision highp float;\nin vec3 Position;\nin vec2 TexCoord;\nout vec2 vTexCoord;\n"
          "void main(){gl_Position=vec4(Position,1.0);vTexCoord=TexCoord;}\n")
    fs = ("#version 300 es\nprecision highp float;\nuniform sampler2D Radiance; uniform sampler2D RngData;\n"
          "in vec2 vTexCoord;\nlayout(location=0) out vec4 o0; layout(location=1) out vec4 o1;\n"
          "void main(){\n  vec4 r = texture(RngData, vTexCoord);\n" + body + "\n}\n")
    if math:
        fs = with_precise_math(fs)
    f = {k: os.path.join(td, f"{tag}_{k}") for k in ("vs", "fs", "rad", "rng", "o0", "o1", "spec")}
    open(f["vs"], "w").write(vs)
    open(f["fs"], "w").write(fs)
    np.tile(np.float32([0, 0, 0, 1]), (n, 1)).tofile(f["rad"])
    np.asarray(x, np.float32).tofile(f["rng"])
    open(f["spec"], "w").write(f"size {W} {H}\nframes 1\nvertex {f['vs']}\nfragment {f['fs']}\n"
                               f"state {f['rad']} {f['rng']}\nout {f['o0']} {f['o1']}\n")
    subprocess.run([HARNESS, GLES_DIR, f["spec"]], check=True, capture_output=True, timeout=120)
    return (np.fromfile(f["o0"], np.float32).reshape(-1, 4), np.fromfile(f["o1"], np.float32).reshape(-1, 4))


def check_mix_normalize(td):
    """mix and normalize of the math block are bit-exact against their
    definitions evaluated in float32 (x * (1 - a) + y * a; x / sqrt(x.x x.x +
    x.y x.y + x.z x.z) per component), as the oracle computes them"""
    W = H = 32
    rs = np.random.RandomState(3)
    x = rs.uniform(-3, 3, (W * H, 4)).astype(np.float32)
    body = ("  o0 = vec4(mix(r.x, r.y, r.z), mix(r.xyz, r.yzw, r.w).xy, mix(r.xy, r.zw, r.yx).y);\n"
            "  o1 = vec4(normalize(r.xyz), normalize(r.zw).x);")
    o0, o1 = run_tiny(td, "mixnorm", body, x, W, H, math=True)
    f = np.float32
    one = f(1)

    def mix(a, b, t):
        return (a * (one - t) + b * t).astype(np.float32)
    X, Y, Z, Wc = (x[:, k] for k in range(4))
    want0 = np.stack([mix(X, Y, Z), mix(X, Y, Wc), mix(Y, Z, Wc), mix(Y, Wc, X)], -1)
    l3 = np.sqrt(X * X + Y * Y + Z * Z).astype(np.float32)
    l2 = np.sqrt(Z * Z + Wc * Wc).astype(np.float32)
    want1 = np.stack([X / l3, Y / l3, Z / l3, Z / l2], -1).astype(np.float32)
    if not (np.array_equal(o0, want0) and np.array_equal(o1, want1)):
        raise RuntimeError("mix/normalize of the math block differ from their float32 definitions: "
                           f"{np.mean(o0 == want0, axis=0)} {np.mean(o1 == want1, axis=0)}")


def check_camera_tan(td, names):
    """the camera's frustum height tan(0.5 * radians(camFovy)), as the math
    block computes it under the GLES implementation, is the correctly rounded
    float for every case's field of view -- the value the oracle and the HIP
    path take (through double); one ulp there tilts every primary ray"""
    fovs = sorted({float(np.float32(host.frame_uniforms(s.renderer, s.camera, s.scales, s.materials, W, H)["camFovy"]))
                   for s, W, H, _, _ in (reference_cases.build(n) for n in names)})
    x = np.zeros((len(fovs), 4), np.float32)
    x[:, 0] = fovs
    o0, _ = run_tiny(td, "camtan", "  o0 = vec4(tan(0.5 * radians(r.x)), 0.0, 0.0, 0.0);\n  o1 = vec4(0.0);",
                     x, len(fovs), 1, math=True)
    f = np.float32
    want = np.tan((f(0.5) * (x[:, 0] * f(0.017453292519943295))).astype(np.float64)).astype(np.float32)
    if not np.array_equal(o0[:, 0], want):
        raise RuntimeError(f"camera tan not correctly rounded for fovy {fovs}: {o0[:, 0]} vs {want}")


def check_min_max(td):
    """min, max and clamp of the math block follow the GLSL definitions with a
    NaN operand: min(3, NaN) = 3, max(1e-6, NaN) = 1e-6, min(NaN, 3) = NaN,
    and agree with the builtins on ordinary values"""
    x = np.float32([[-1, 4, 0, 0], [-2, 9, 0, 0], [4, -3, 0, 0], [0.25, 0.5, 0, 0]])
    body = ("  float a = sqrt(r.x);\n"
            "  o0 = vec4(min(3.0, a), max(1e-6, a), min(a, 3.0), clamp(a, 0.0, 1.0));\n"
            "  vec3 v = min(vec3(3.0), vec3(a)); vec3 w = max(vec3(r.y), 0.5);\n"
            "  o1 = vec4(v.x, w.x, clamp(r.y, -1.0, 1.0), float(min(int(r.y), 2)));")
    o0, o1 = run_tiny(td, "minmax", body, x, 2, 2, math=True)
    a = np.sqrt(np.where(x[:, 0] >= 0, x[:, 0], np.nan)).astype(np.float32)
    nan = np.isnan(a)
    ok = (np.array_equal(o0[nan, 0], np.full(nan.sum(), 3.0, np.float32)) and
          np.allclose(o0[nan, 1], 1e-6) and np.all(np.isnan(o0[nan, 2])) and np.all(np.isnan(o0[nan, 3])) and
          np.allclose(o0[~nan, 0], np.minimum(3.0, a[~nan])) and np.allclose(o0[~nan, 3], np.clip(a[~nan], 0, 1)) and
          np.array_equal(o1[nan, 0], np.full(nan.sum(), 3.0, np.float32)) and
          np.allclose(o1[:, 1], np.maximum(x[:, 1], 0.5)) and np.allclose(o1[:, 2], np.clip(x[:, 1], -1, 1)) and
          np.allclose(o1[:, 3], np.minimum(np.trunc(x[:, 1]), 2)))
    if not ok:
        raise RuntimeError(f"min/max/clamp of the math block do not follow the GLSL definitions:\n{o0}\n{o1}")


def uniform_lines(u):
    out = []
    for k, val in u.items():
        if k in host.INT_UNIFORMS:
            out.append(f"u1i {k} {int(val)}")
            continue
        if isinstance(val, (list, tuple, np.ndarray)):
            vals = [float(x) for x in val]
            if len(vals) in (2, 3, 4):
                out.append(f"u{len(vals)}f {k} " + " ".join(repr(x) for x in vals))
            else:
                out.append(f"u1fv {k} {len(vals)} " + " ".join(repr(x) for x in vals))
        else:
            out.append(f"u1f {k} {float(val)!r}")
    return out


def run_reference(scene, W, H, frames, seed, td, name="case", program="pathtracer", radiance=None):
    """render `frames` frames of `scene` with the reference's `program`
    (pathtracer or normals), or its pick program once per pixel, from a fresh
    state or from `radiance` (the tonemapper reads the pathtracer's result);
    returns (radiance, rng, gl_info, seconds)"""
    scene.camera.aspect = W / H
    u = host.frame_uniforms(scene.renderer, scene.camera, scene.scales, scene.materials, W, H)
    env = getattr(scene, "env_map", None)
    u["haveEnvMap"] = int(env is not None)
    u.update(scene.user_uniforms)
    if program == "pick":  # Renderer.pick sets the length scales unclamped (js/renderer.js:436-438)
        u.update(lengthScale=scene.scales.lengthScale, minLengthScale=scene.scales.minLengthScale,
                 maxLengthScale=scene.scales.maxLengthScale)
    vs, fs = program_source(scene, program)
    name = f"{name}_{program}"
    paths = {k: os.path.join(td, f"{name}_{k}") for k in ("vs", "fs", "rad", "rng", "rad_out", "rng_out", "spec")}
    open(paths["vs"], "w").write(vs)
    open(paths["fs"], "w").write(fs)
    rad = np.tile(np.float32([0, 0, 0, 1]), (W * H, 1)) if radiance is None else np.float32(radiance)
    rng = oracle_py.Oracle.seed(W * H, seed).astype(np.float32)
    rad.tofile(paths["rad"])
    rng.tofile(paths["rng"])
    spec = [f"size {W} {H}", f"frames {frames}", f"vertex {paths['vs']}", f"fragment {paths['fs']}",
            f"state {paths['rad']} {paths['rng']}", f"out {paths['rad_out']} {paths['rng_out']}"]
    if program == "pick":
        spec.append("pick")
    # textures at the units Renderer.render and the materials bind them to
    tex = [(2, "WavelengthToXYZ", host.xyz_lut(), 1024, 4),
           (3, "ICDF", host.inverse_cdf(lambda wl: host.blackbody(wl, scene.renderer.skyTemperature)), 4096, 1)]
    if env is not None:  # the env map at unit 6 (js/renderer.js:595-602), uploaded as SRGB8 bytes
        p = os.path.join(td, f"{name}_envmap")
        np.asarray(env[2], np.uint8).tofile(p)
        spec.append(f"texture_srgb8 6 envMap {env[0]} {env[1]} {p}")
    mt = scene.materials.metal.textures()
    if "iorTex" in mt:
        tex.append((4, "iorTex", mt["iorTex"][0], 64, 1))
        tex.append((5, "k_tex", mt["k_tex"][0], 64, 1))
    for unit, nm, data, w, ch in tex:
        p = os.path.join(td, f"{name}_tex{unit}")
        np.asarray(data, np.float32).tofile(p)
        spec.append(f"texture {unit} {nm} {w} 1 {ch} 1 1 {p}")
    spec += uniform_lines(u)
    open(paths["spec"], "w").write("\n".join(spec) + "\n")
    t0 = time.time()
    r = subprocess.run([HARNESS, GLES_DIR, paths["spec"]], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{name}: gles_render failed\n{r.stderr[-4000:]}")
    gl_info = [l for l in r.stderr.splitlines() if l.startswith("GL:")]
    out_rad = np.fromfile(paths["rad_out"], np.float32).reshape(-1, 4)
    out_rng = np.fromfile(paths["rng_out"], np.float32).reshape(-1, 4) if program != "pick" else None
    return out_rad, out_rng, (gl_info[0][4:] if gl_info else ""), time.time() - t0


# tonemapper settings (Renderer's exposure, gamma, contrast, saturation, hueShift),
# one per case in turn: the defaults and two others
TONEMAP_SETTINGS = [dict(exposure=0.0, gamma=2.2, contrast=1.0, saturation=1.0, hueShift=0.0),
                    dict(exposure=0.8, gamma=2.0, contrast=1.3, saturation=1.4, hueShift=0.25),
                    dict(exposure=-1.5, gamma=2.4, contrast=0.8, saturation=0.6, hueShift=-0.4)]


def make(name, td):
    scene, W, H, frames, seed = reference_cases.build(name)
    # Renderer.compileShaders' program per renderMode (js/renderer.js:334-352)
    program = {"ao": "ao", "ptsimple": "simplepathtracer"}.get(scene.renderer.renderMode, "pathtracer")
    out_rad, out_rng, gl, secs = run_reference(scene, W, H, frames, seed, td, name, program)
    # first-hit passes: the pick program per pixel, one frame of the normals program
    pick, _, _, _ = run_reference(scene, W, H, 1, seed, td, name, "pick")
    nrm_rad, nrm_rng, _, _ = run_reference(scene, W, H, 1, seed, td, name, "normals")
    # the tonemapper over the final radiance (js/renderer.js:622-642), read back
    # as floats before the canvas's 8-bit store
    tm = TONEMAP_SETTINGS[list(reference_cases.CASES).index(name) % len(TONEMAP_SETTINGS)]
    for k, v in tm.items():
        setattr(scene.renderer, k, v)
    tone, _, _, _ = run_reference(scene, W, H, 1, seed, td, name, "tonemapper", radiance=out_rad)
    meta = dict(case=name, width=W, height=H, frames=frames, seed=seed, gl=gl, seconds=round(secs, 2),
                shaders=[f"shaders/{program}-fragment-shader.glsl", "shaders/pick-fragment-shader.glsl",
                         "shaders/normals-fragment-shader.glsl", "shaders/tonemapper-fragment-shader.glsl"],
                tonemap=tm)
    np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), radiance=out_rad, rng=out_rng,
                        pick=pick[:, :2].copy(), normals=nrm_rad, normals_rng=nrm_rng, tonemap=tone.copy(),
                        meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(json.dumps(meta))


def time_plain_case(td, W=256, H=256, frames=16):
    """the reference's pathtracer on this container's CPUs through the GLES
    implementation, on the CPU-runnable case (unit sphere over a plane, 256 x
    256, 16 spp, 4 bounces): one frame and 1 + `frames` frames are timed, and
    their difference, which leaves out program compilation and process start,
    gives the per-frame rate; the children's CPU time over the wall time gives
    the cores the implementation kept busy"""
    import resource
    from snelly_amd import scenes
    out = {}
    for n in (1, 1 + frames):
        c0 = resource.getrusage(resource.RUSAGE_CHILDREN)
        _, _, gl, secs = run_reference(scenes.sphere_plane(W, H, frames), W, H, n, 1234, td, "timing")
        c1 = resource.getrusage(resource.RUSAGE_CHILDREN)
        out[n] = (secs, (c1.ru_utime - c0.ru_utime) + (c1.ru_stime - c0.ru_stime))
    wall = out[1 + frames][0] - out[1][0]
    cpu = out[1 + frames][1] - out[1][1]
    return dict(workload=f"sphere over plane {W}x{H}, {frames} spp, 4 bounces", gl=gl, seconds_per_frame=wall / frames,
                path_samples_per_s=W * H * frames / wall, cores_busy=round(cpu / wall, 2),
                nproc=os.cpu_count())


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "reference_gles")], check=True)
    if sys.argv[1:] == ["--time"]:
        with tempfile.TemporaryDirectory() as td:
            print(json.dumps(time_plain_case(td)))
        return
    names = sys.argv[1:] or list(reference_cases.CASES)
    with tempfile.TemporaryDirectory() as td:
        print("precise_math ulp:", json.dumps({k: round(v, 2) for k, v in check_precise_math(td).items()}))
        print("loop with continue: wrong results at", check_loop_control(td), "of 256 pixels; flag form exact")
        check_min_max(td)
        check_mix_normalize(td)
        check_camera_tan(td, names)
        for n in names:
            make(n, td)


if __name__ == "__main__":
    main()
