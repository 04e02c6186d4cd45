"""Writes tests/golden/ref_<case>.npz: the reference's path tracer
(shaders/pathtracer-fragment-shader.glsl, assembled as js/renderer.js:217-359
and scripts/build do) run by a GLES 3 implementation over the cases of
reference_cases.py, from integer RNG seeds.

Development-container tooling: it reads the reference's shader sources from
/root/reference and drives reference_gles/_build/gles_render, which dlopens the
SwiftShader GLES libraries installed with the kaleido package.  Two steps
make the run follow the reference's arithmetic rather than the driver's
liberties: uninitialised locals are zero-initialised, as under WebGL, and the
transcendental builtins are replaced by accurate ones
(reference_gles/precise_math.glsl, checked before any fixture is written).  The fixtures it writes are
data (initial state is implied by the seed; outputs are the
accumulated radiance and the final RNG state).

    python tests/golden/make_reference_fixtures.py [case ...]
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


def run_reference(scene, W, H, frames, seed, td, name="case", program="pathtracer"):
    """render `frames` frames of `scene` with the reference's `program`
    (pathtracer or normals), or its pick program once per pixel; returns
    (radiance, rng, gl_info, seconds)"""
    scene.camera.aspect = W / H
    u = host.frame_uniforms(scene.renderer, scene.camera, scene.scales, scene.materials, W, H)
    u.update(scene.user_uniforms)
    if program == "pick":  # Renderer.pick sets the length scales unclamped (js/renderer.js:436-438)
        u.update(lengthScale=scene.scales.lengthScale, minLengthScale=scene.scales.minLengthScale,
                 maxLengthScale=scene.scales.maxLengthScale)
    vs, fs = program_source(scene, program)
    name = f"{name}_{program}"
    paths = {k: os.path.join(td, f"{name}_{k}") for k in ("vs", "fs", "rad", "rng", "rad_out", "rng_out", "spec")}
    open(paths["vs"], "w").write(vs)
    open(paths["fs"], "w").write(fs)
    rad = np.tile(np.float32([0, 0, 0, 1]), (W * H, 1))
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


def make(name, td):
    scene, W, H, frames, seed = reference_cases.build(name)
    out_rad, out_rng, gl, secs = run_reference(scene, W, H, frames, seed, td, name)
    # first-hit passes: the pick program per pixel, one frame of the normals program
    pick, _, _, _ = run_reference(scene, W, H, 1, seed, td, name, "pick")
    nrm_rad, nrm_rng, _, _ = run_reference(scene, W, H, 1, seed, td, name, "normals")
    meta = dict(case=name, width=W, height=H, frames=frames, seed=seed, gl=gl, seconds=round(secs, 2),
                shaders=["shaders/pathtracer-fragment-shader.glsl", "shaders/pick-fragment-shader.glsl",
                         "shaders/normals-fragment-shader.glsl"])
    np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), radiance=out_rad, rng=out_rng,
                        pick=pick[:, :2].copy(), normals=nrm_rad, normals_rng=nrm_rng,
                        meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(json.dumps(meta))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "reference_gles")], check=True)
    names = sys.argv[1:] or list(reference_cases.CASES)
    with tempfile.TemporaryDirectory() as td:
        print("precise_math ulp:", json.dumps({k: round(v, 2) for k, v in check_precise_math(td).items()}))
        for n in names:
            make(n, td)


if __name__ == "__main__":
    main()
