"""Writes tests/golden/host_vectors.json: what the reference's own host
JavaScript (js/renderer.js, js/materials.js, js/spectra.js, js/color.js,
js/snelly.js, with three.js r79) hands to WebGL for

* each scene of tests/golden/reference_cases.py (written out as a scene script
  with the same settings), and
* each example scene the reference ships (exampleScenes/**/*.html).

For every scene: the program replacements of Renderer.compileShaders (the
assembled scene code with its default hooks, the IOR code, the #defines, the
__MAX_*__ tokens), the uniforms of the first Renderer.render in call order, and
the texels of the textures that render binds (the emission ICDF of
Snelly.loadSpectrum, the XYZ table, tabulated metals' n and k).  Long strings
and texel arrays are kept as SHA-256 digests.

tests/test_host_vectors.py compares snelly_amd/host.py and
snelly_amd/js/snelly_host.js against this file.  It is made in the development
container, where the reference and node are present:

    python tests/golden/make_host_vectors.py [--reference /root/reference]

The node side is reference_host.js, next to this file.
"""
import argparse
import hashlib
import json
import math
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

OUT = os.path.join(HERE, "host_vectors.json")

# the uniforms of the tonemapper draw of Renderer.render (js/renderer.js:622-642)
TONEMAP_UNIFORMS = ("exposure", "invGamma", "contrast", "saturation", "hueShift")


def sha(text):
    return hashlib.sha256(text.encode("utf-8")).hexdigest()


def sha_norm(text):
    """digest of the text without trailing blanks and empty lines (GLSL does
    not see them; the reference's generators leave runs of them)"""
    return sha("\n".join(line.rstrip() for line in text.split("\n") if line.strip()))


def js_value(v):
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (list, tuple)):
        return "[" + ", ".join(js_value(x) for x in v) + "]"
    if hasattr(v, "tolist"):
        return js_value(v.tolist())
    if isinstance(v, str):
        return json.dumps(v)
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    return repr(v)


def _dielectric_js(d):
    """the scene-script lines that make d the loaded dielectric"""
    from snelly_amd import host
    ref = host.DIELECTRICS.get(d.name)
    same = ref is not None and type(ref()) is type(d) and vars(ref()).get("coeffs") == vars(d).get("coeffs") \
        and getattr(ref(), "iorVal", None) == getattr(d, "iorVal", None)
    lines = []
    if not same:
        if isinstance(d, host.AbbeDielectric):
            ctor = f"new AbbeDielectric({json.dumps(d.name)}, {js_value(d.iorVal)}, {js_value(d.abbe)})"
        else:
            ctor = f"new {type(d).__name__}({json.dumps(d.name)}, {js_value(d.coeffs)})"
        lines.append(f"materials.addDielectric({ctor});")
    lines.append(f"let dielectric = materials.loadDielectric({json.dumps(d.name)});")
    for k in ("roughness", "absorptionScale", "absorptionColor", "specAlbedo"):
        lines.append(f"dielectric.{k} = {js_value(getattr(d, k))};")
    return lines


def _metal_js(m):
    from snelly_amd import host
    lines = []
    if isinstance(m, host.GlslMetal):
        # a scene's own Metal subclass (Materials.addMetal, js/materials.js:1245-1248)
        lines.append(f"let customMetal = new Metal({json.dumps(m.name)});")
        lines.append(f"customMetal.ior = function() {{ return {json.dumps(m.glsl)}; }};")
        lines.append("materials.addMetal(customMetal);")
    lines.append(f"let metal = materials.loadMetal({json.dumps(m.name)});")
    for k in ("roughness", "specAlbedo"):
        lines.append(f"metal.{k} = {js_value(getattr(m, k))};")
    return lines


def scene_script(s, width, height):
    """a Snelly scene script (js/sceneAPI.js) with the settings of the
    scene object s"""
    r, cam, mats, sc = s.renderer, s.camera, s.materials, s.scales
    init = ["let renderer = snelly.getRenderer();", "let camera = snelly.getCamera();",
            "let controls = snelly.getControls();", "let materials = snelly.getMaterials();",
            f"renderer.width = {width};", f"renderer.height = {height};"]
    for k, v in vars(r).items():
        if k == "sunLongitude" and isinstance(v, float) and math.isnan(v):
            continue  # left undefined, as the reference's defaults leave it
        init.append(f"renderer.{k} = {js_value(v)};")
    init.append(f"camera.up.set({', '.join(js_value(x) for x in cam.up)});")
    init.append(f"camera.position.set({', '.join(js_value(x) for x in cam.position)});")
    init.append(f"controls.target.set({', '.join(js_value(x) for x in cam.target)});")
    for k in ("fov", "aperture", "focalDistance"):
        init.append(f"camera.{k} = {js_value(getattr(cam, k))};")
    init.append("let surface = materials.loadSurface();")
    init += [f"surface.{k} = {js_value(v)};" for k, v in vars(mats.surface).items()]
    init.append("let volume = materials.loadVolume();")
    init += [f"volume.{k} = {js_value(v)};" for k, v in vars(mats.volume).items()]
    init += _dielectric_js(mats.dielectric)
    init += _metal_js(mats.metal)
    sync = [f"shader.{'uniformF' if not isinstance(v, (list, tuple)) else f'uniform{len(v)}Fv'}"
            f"({json.dumps(k)}, {js_value(v)});" for k, v in s.user_uniforms.items()]
    env = ""
    if getattr(s, "env_map", None) is not None:
        env = "Scene.prototype.envMap = function() { return 'env_map.png'; };\n"
    return ("function Scene() {}\n"
            "Scene.prototype.init = function(snelly) {\n  " + "\n  ".join(init) + "\n};\n"
            f"Scene.prototype.shader = function() {{ return {json.dumps(s.glsl)}; }};\n"
            f"Scene.prototype.getLengthScale = function() {{ return {js_value(sc.lengthScale)}; }};\n"
            f"Scene.prototype.getMinLengthScale = function() {{ return {js_value(sc.minLengthScale)}; }};\n"
            f"Scene.prototype.getMaxLengthScale = function() {{ return {js_value(sc.maxLengthScale)}; }};\n"
            "Scene.prototype.syncShader = function(snelly, shader) {\n  " + "\n  ".join(sync) + "\n};\n" + env)


def extract_scene_script(html):
    """the inline scripts of an example page that define the scene (the
    remote snelly.min.js include and a bare onLoad() block are left out)"""
    blocks = []
    for m in re.finditer(r"<script(?![^>]*\bsrc=)[^>]*>([\s\S]*?)</script>", html, re.I):
        code = m.group(1)
        if re.search(r"function\s+onLoad\s*\(", code) and not re.search(r"Scene\.prototype", code):
            continue
        blocks.append(code)
    return "\n".join(blocks)


def example_files(ref):
    base = os.path.join(ref, "exampleScenes")
    out = []
    for d, _, files in os.walk(base):
        for f in files:
            if f.endswith(".html"):
                out.append(os.path.relpath(os.path.join(d, f), base))
    return sorted(out)


def compact(rec):
    """long strings and texel arrays become digests"""
    if "error" in rec:
        return rec

    def reps(x):
        if x is None:
            return None
        out = {}
        for k, v in x.items():
            out[k] = {"sha256": sha(v), "sha256_norm": sha_norm(v), "len": len(v)} \
                if k in ("__SHADER__", "__IOR_FUNC__") else v
        return out

    tex = {}
    for name, t in rec["textures"].items():
        d = dict(t)
        if d["texels"] is not None:
            raw = bytes.fromhex(d["texels"])
            d["texels"] = {"sha256": hashlib.sha256(raw).hexdigest(), "n": len(raw) // 4}
        tex[name] = d
    rec = dict(rec, integrator=reps(rec["integrator"]), pick=reps(rec["pick"]), textures=tex)
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--out", default=OUT)
    a = ap.parse_args()
    import reference_cases
    cases = []
    for name in reference_cases.CASES:
        s, W, H, _, _ = reference_cases.build(name)
        cases.append({"name": "case:" + name, "scene_js": scene_script(s, W, H), "width": W, "height": H})
    for f in example_files(a.reference):
        with open(os.path.join(a.reference, "exampleScenes", f)) as fh:
            cases.append({"name": "example:" + f, "scene_js": extract_scene_script(fh.read()),
                          "width": 256, "height": 256})
    with tempfile.TemporaryDirectory() as td:
        cj, oj = os.path.join(td, "cases.json"), os.path.join(td, "out.json")
        with open(cj, "w") as fh:
            json.dump(cases, fh)
        subprocess.run(["node", os.path.join(HERE, "reference_host.js"), a.reference, cj, oj], check=True)
        with open(oj) as fh:
            raw = json.load(fh)
    scripts = {c["name"]: c["scene_js"] for c in cases if c["name"].startswith("case:")}
    out = {"about": "what the reference's host JS hands to WebGL per scene; made by make_host_vectors.py",
           "reference_files": raw["reference_files"], "cases": []}
    for rec in raw["cases"]:
        rec = compact(rec)
        if rec["name"] in scripts:
            rec["scene_sha256"] = sha(scripts[rec["name"]])
        out["cases"].append(rec)
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=0, sort_keys=True)
    bad = [c["name"] + ": " + c["error"] for c in out["cases"] if "error" in c]
    print(f"{len(out['cases'])} scenes, {len(bad)} raise in the reference host")
    for b in bad:
        print("  ", b)


if __name__ == "__main__":
    main()
