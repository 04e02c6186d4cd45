"""The two hosts against the reference's own host JavaScript.

tests/golden/host_vectors.json holds what js/renderer.js, js/materials.js,
js/spectra.js, js/color.js and js/snelly.js hand to WebGL when they run each
scene of tests/golden/reference_cases.py and each example scene of the
reference (made by tests/golden/make_host_vectors.py, which runs those files
under Node with stand-ins for the DOM and WebGL):

* snelly_amd/host.py (assemble, frame_uniforms, the LUTs and the ICDF) must
  produce the same program text (digests), #defines, __MAX_*__ tokens, uniform
  values (as float32, the precision WebGL stores them in) and texture texels
  for the fixture scenes;
* snelly_amd/js/snelly_host.js, run over a recording stand-in for the addon,
  must do the same for the fixture scenes and the example scenes.

Deliberate differences, each stated where it is allowed: a scene that defines
SDF_VOLUME gets HAS_VOLUME here (the reference throws on an undeclared
`hasVolume`, js/renderer.js:268), and the medium's extinction is one more
uniform (volumeExtinction).
"""
import hashlib
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_host_vectors as mhv  # noqa: E402
import reference_cases  # noqa: E402
from snelly_amd import host  # noqa: E402

VECTORS = os.path.join(HERE, "golden", "host_vectors.json")
NODE = shutil.which("node")
REF_EXAMPLES = "/root/reference/exampleScenes"

# uniforms this build sets that the reference never does
EXTRA_UNIFORMS = {"volumeExtinction"}
# example scenes that the stand-ins of make_host_vectors.py cannot run in the
# reference (they decode packed mesh data from images that never load there)
STANDIN_LIMITS = {"fractals/female-mandel.html", "fractals/female-menger.html", "meshes/female.html"}
# samplers bound in the load callback of GLU.loadImageAndCreateTextureInfo: the
# images load here (local files) but never under make_host_vectors.py's stand-ins
IMAGE_SAMPLERS = {"example:fractals/IFS/damaged2.html": {"albedoTex", "displacementTex", "roughnessTex"},
                  "example:meshes/PBR_setup.html": {"albedoTex", "displacementTex", "roughnessTex"}}


def load():
    with open(VECTORS) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


def sha(text):
    return hashlib.sha256(text.encode("utf-8")).hexdigest()


def f32(vals):
    out = []
    for v in vals:
        if isinstance(v, str):
            v = {"NaN": math.nan, "Infinity": math.inf, "-Infinity": -math.inf, "undefined": math.nan}[v]
        out.append(np.float32(v))
    return np.array(out, np.float32)


def ref_uniforms(rec, integrator):
    """name -> values of the render's uploads: the integrator's, then the
    tonemapper's (later calls win, as in GL)"""
    u = {}
    for prog, fn, name, vals in rec["uniforms"]:
        if fn == "uniformTexture" or prog not in (integrator, "tonemapper"):
            continue
        if fn == "uniform1i":
            # WebGL's uniform1i takes a GLint: the number is converted by
            # ToInt32 (truncation toward zero), as scenes pass non-integers
            vals = [int(math.trunc(v)) if not isinstance(v, str) else 0 for v in vals]
        u[name] = vals
    return u


INTEGRATOR = {"pt": "pathtracer", "ao": "ao", "ptsimple": "simplepathtracer", "normals": "normals"}


def same(a, b):
    a, b = f32(a), f32(np.atleast_1d(np.asarray(b, np.float64)).tolist())
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def check_uniforms(ref, ours, label, extra=()):
    bad = []
    for name, vals in ref.items():
        if name not in ours:
            bad.append(f"{name}: not set (reference {vals})")
        elif not same(vals, ours[name]):
            bad.append(f"{name}: reference {vals}, here {list(np.atleast_1d(ours[name]))}")
    for name in ours:
        if name not in ref and name not in EXTRA_UNIFORMS and name not in extra:
            bad.append(f"{name}: set here, never by the reference")
    assert not bad, f"{label}: " + "; ".join(bad)


def check_program(rep, shader, ior, defines, tokens, label):
    assert rep is not None, label
    assert sha(shader) == rep["__SHADER__"]["sha256"], f"{label}: scene code with default hooks differs"
    # the IOR generators' text up to trailing blanks on a line
    assert mhv.sha_norm(ior) == rep["__IOR_FUNC__"]["sha256_norm"], f"{label}: IOR code differs"
    assert defines == rep["__DEFINES__"], f"{label}: defines {defines!r} vs {rep['__DEFINES__']!r}"
    for k, v in tokens.items():
        assert str(v) == rep[k], f"{label}: {k} {v} vs {rep[k]}"


def texel_sha(arr):
    return hashlib.sha256(np.ascontiguousarray(arr, np.float32).tobytes()).hexdigest()


def test_vectors_cover_cases_and_examples():
    v = load()
    cases = [n for n in v if n.startswith("case:")]
    examples = [n for n in v if n.startswith("example:")]
    assert {"case:" + n for n in reference_cases.CASES} == set(cases)
    assert len(examples) >= 50
    ok = [n for n in examples if "error" not in v[n]]
    assert len(ok) >= 45
    # the scenes the reference itself cannot run: undeclared hasVolume
    # (volume-only scenes) and the TextSDF font scenes
    for n in examples:
        if "error" in v[n] and n[len("example:"):] not in STANDIN_LIMITS:
            assert "hasVolume" in v[n]["error"] or "TextSDF" in v[n]["error"], (n, v[n]["error"])


@pytest.mark.parametrize("name", sorted(reference_cases.CASES))
def test_python_host_matches_reference_js(name):
    rec = load()["case:" + name]
    s, W, H, _, _ = reference_cases.build(name)
    # the vectors were made from this very scene script
    assert sha(mhv.scene_script(s, W, H)) == rec["scene_sha256"]
    r, mats = s.renderer, s.materials
    shader, ior, defines = host.assemble(s.glsl, r, mats)
    tokens = {"__MAX_MARCH_STEPS__": int(round(r.maxMarchSteps)), "__MAX_BOUNCES__": int(round(r.maxBounces)),
              "__MAX_ATMOSPHERE_SCATTERS__": int(round(r.maxAtmosphereScatters)),
              "__MAX_SAMPLES_PER_FRAME__": int(round(r.maxSamplesPerFrame))}
    check_program(rec["integrator"], shader, ior, defines, tokens, name)
    s.camera.aspect = W / H
    u = host.frame_uniforms(r, s.camera, s.scales, mats, W, H)
    u.update(s.user_uniforms)
    u["haveEnvMap"] = 0 if rec["envMapUnloaded"] else u["haveEnvMap"]
    check_uniforms(ref_uniforms(rec, INTEGRATOR[rec["renderMode"]]), u, name)
    # textures the render binds
    tex = rec["textures"]
    # Renderer.render binds the LUTs for 'pt' and 'ao' only (js/renderer.js:540-546)
    assert ("ICDF" in tex) == (r.renderMode in ("pt", "ao"))
    if "ICDF" in tex:
        assert texel_sha(host.xyz_lut()) == tex["WavelengthToXYZ"]["texels"]["sha256"]
        icdf = host.inverse_cdf(lambda wl: host.blackbody(wl, r.skyTemperature))
        assert texel_sha(icdf) == tex["ICDF"]["texels"]["sha256"]
    ours = mats.metal.textures()
    for k, (data, w, h) in ours.items():
        assert texel_sha(data) == tex[k]["texels"]["sha256"], k
        assert (w, h) == (tex[k]["width"], tex[k]["height"])
    assert set(ours) == {k for k in tex if k not in ("WavelengthToXYZ", "ICDF")}


def run_js_recorder(cases):
    with tempfile.TemporaryDirectory() as td:
        cj, oj = os.path.join(td, "cases.json"), os.path.join(td, "out.json")
        with open(cj, "w") as f:
            json.dump(cases, f)
        r = subprocess.run([NODE, os.path.join(HERE, "js", "record_js_host.js"), cj, oj], capture_output=True,
                           text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        with open(oj) as f:
            return {c["name"]: c for c in json.load(f)}


def check_js(rec, got):
    name = rec["name"]
    assert "error" not in got, (name, got.get("error"))
    b = got["build"]
    tokens = {"__MAX_MARCH_STEPS__": b["march"], "__MAX_BOUNCES__": b["bounces"],
              "__MAX_ATMOSPHERE_SCATTERS__": b["atm"], "__MAX_SAMPLES_PER_FRAME__": b["spf"]}
    defines = b["defines"]
    rep = dict(rec["integrator"])
    if "\n#define HAS_VOLUME\n" in defines:
        # the deliberate difference: a scene with SDF_VOLUME and a surface SDF
        # has the heterogeneous medium here, where the reference ignores
        # SDF_VOLUME (it reaches the undeclared hasVolume only without one) and
        # may take the Volume material's colours for the homogeneous atmosphere
        assert "SDF_VOLUME(" in b["shader"], name
        drop = ("\n#define HAS_VOLUME\n", "\n#define HAS_ATMOSPHERE\n")
        for d in drop:
            defines = defines.replace(d, "")
            rep["__DEFINES__"] = rep["__DEFINES__"].replace(d, "")
    check_program(rep, b["shader"], b["ior"], defines, tokens, name)
    ours = {n: v for n, t, v in got["sets"] if t != "texture"}
    if rec["envMapUnloaded"]:
        ours["haveEnvMap"] = [0]
    check_uniforms(ref_uniforms(rec, INTEGRATOR[rec["renderMode"]]), ours, name, IMAGE_SAMPLERS.get(name, ()))
    for k, t in rec["textures"].items():
        if k in got["textures"] and t["texels"] is not None:
            raw = bytes.fromhex(got["textures"][k]["texels"])
            assert hashlib.sha256(raw).hexdigest() == t["texels"]["sha256"], (name, k)


@pytest.mark.skipif(NODE is None, reason="node is missing")
def test_js_host_matches_reference_js_on_fixture_scenes():
    v = load()
    cases = []
    for name in sorted(reference_cases.CASES):
        s, W, H, _, _ = reference_cases.build(name)
        cases.append({"name": "case:" + name, "scene_js": mhv.scene_script(s, W, H), "width": W, "height": H})
    got = run_js_recorder(cases)
    for c in cases:
        check_js(v[c["name"]], got[c["name"]])


@pytest.mark.skipif(NODE is None or not os.path.isdir(REF_EXAMPLES),
                    reason="node or the reference's example scenes are missing")
def test_js_host_matches_reference_js_on_example_scenes():
    v = load()
    names = [n for n in v if n.startswith("example:") and "error" not in v[n]]
    cases = [{"name": n, "file": os.path.join(REF_EXAMPLES, n[len("example:"):]), "width": 256, "height": 256}
             for n in names]
    got = run_js_recorder(cases)
    bad = []
    for n in names:
        try:
            check_js(v[n], got[n])
        except AssertionError as e:
            bad.append(str(e)[:600])
    assert not bad, "\n".join(bad)
