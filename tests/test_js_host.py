"""The JavaScript host (snelly_amd/js): the reference's scene API (Snelly,
Renderer, Materials, GLU.Shader, a WebGL stand-in) over the Node-API addon.

CPU: the addon loads and exports its functions; a scene file in the
reference's format assembles and compiles; every example scene the reference
ships compiles, or fails exactly where the reference itself fails.
GPU: a scene file rendered by node matches the CPU oracle on the same seeds.
"""
import json
import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "snelly_amd", "snelly_napi.node")
RUN = os.path.join(ROOT, "snelly_amd", "js", "run_scene.js")
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")
REF_EXAMPLES = "/root/reference/exampleScenes"

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the addon is missing")

# Example scenes the reference itself cannot run:
#  - meshmenger uses `size`, which it never declares (fractals/IFS/meshmenger.html:429);
#  - texture_rocks defines SURFACE_NORMAL_MAP(vec3) while the shader calls it
#    with (X, N) (pathtracer-fragment-shader.glsl:948);
#  - the font scenes construct TextSDF, a browser font rasteriser that is not
#    part of the reference.
REFERENCE_FAILS = {
    "fractals/IFS/meshmenger.html", "textures/texture_rocks.html",
    "fonts/font_sdf.html", "fonts/font_sdf2.html", "textures/texture_flesh.html",
}
# Volume-only scenes: the reference's compileShaders tests an undeclared
# `hasVolume` when no surface SDF exists (js/renderer.js:267-271) and throws.
# This host declares it (SDF_VOLUME defines the heterogeneous medium), so they
# compile here.
VOLUME_ONLY = {"physics/orbitals.html", "volumes/pyroclastic.html", "volumes/volumeBunny.html",
               "volumes/volumeCheckerboard.html"}


def run_node(args, timeout=300, env=None):
    return subprocess.run([NODE, RUN] + args, capture_output=True, text=True, timeout=timeout, env=env)


@needs_node
def test_addon_exports():
    names = ["ctxCreate", "ctxDestroy", "sync", "programBuild", "programDestroy", "uniformIndex", "uniformSet",
             "uniformTexture", "textureCreate", "textureUpload", "textureDestroy", "stateCreate", "stateDestroy",
             "stateSeed", "stateClear", "stateRead", "stateWrite", "render", "renderNormals", "pick", "tonemap",
             "tonemapRGBA8", "stats", "setEngine", "version"]
    code = f"const a = require({json.dumps(ADDON)}); console.log(JSON.stringify([Object.keys(a), a.version()]));"
    out = subprocess.run([NODE, "-e", code], capture_output=True, text=True, check=True).stdout
    keys, version = json.loads(out.strip().splitlines()[-1])
    assert sorted(keys) == sorted(names)
    assert "gfx950" in version


@needs_node
def test_scene_file_compiles():
    r = run_node([os.path.join(SCENES, "sphere_plane.html"), "--compile-only", "--precise", "--json"])
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout.strip().splitlines()[-1])["compile_ms"] >= 0


@needs_node
def test_shader_assembly_matches_python_host():
    """Renderer.compileShaders in JS and host.assemble in Python splice the same
    hooks and #defines for the same scene"""
    from snelly_amd import host, scenes
    code = f"""
const h = require({json.dumps(os.path.join(ROOT, 'snelly_amd', 'js', 'snelly_host.js'))});
const {{ Scene }} = h.loadScene({json.dumps(os.path.join(SCENES, 'sphere_plane.html'))});
const sn = new h.Snelly(new Scene(), {{ compileOnly: true, width: 64, height: 64 }});
console.log(JSON.stringify(sn.getRenderer().assembleShaders()));
"""
    r = subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])
    s = scenes.sphere_plane(64, 64)
    shader, ior, defines = host.assemble(s.glsl, s.renderer, s.materials)
    norm = lambda t: " ".join(t.split())
    assert norm(js["shader"]) == norm(shader)
    assert sorted(js["defines"].split()) == sorted(defines.split())


@needs_node
def test_gui_and_webgl_standins():
    """scene code that drives the GUI and raw WebGL textures runs unchanged"""
    code = f"""
const h = require({json.dumps(os.path.join(ROOT, 'snelly_amd', 'js', 'snelly_host.js'))});
const g = new h.GUI();
g.getSceneFolder().add({{}}, 'x', 0, 1).name('x').onChange(() => 0);
g.addSlider({{}}, {{name: 'x', min: 0, max: 1}}).step(0.1);
const gl = h.sceneGlobals().gl;
const t = gl.createTexture();
gl.activeTexture(gl.TEXTURE0 + 7);
gl.bindTexture(gl.TEXTURE_3D, t);
gl.texImage3D(gl.TEXTURE_3D, 0, gl.R32F, 2, 3, 4, 0, gl.RED, gl.FLOAT, new Float32Array(24).fill(0.5));
gl.texParameteri(gl.TEXTURE_3D, gl.TEXTURE_MAG_FILTER, gl.NEAREST);
gl.texParameteri(gl.TEXTURE_3D, gl.TEXTURE_WRAP_S, gl.CLAMP_TO_EDGE);
console.log(JSON.stringify([t.w, t.h, t.d, t.ch, t.linear, t.clamp, t.data.length, t.data[5]]));
"""
    r = subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout.strip()) == [2, 3, 4, 1, False, True, 24, 0.5]


def _example_list():
    out = []
    for d, _, files in os.walk(REF_EXAMPLES):
        for f in files:
            if f.endswith(".html"):
                out.append(os.path.relpath(os.path.join(d, f), REF_EXAMPLES))
    return sorted(out)


@needs_node
@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="the reference's example scenes are not present")
def test_reference_example_scenes_compile():
    """every example scene compiles (translation + hipRTC for gfx950), except
    those the reference cannot run either"""
    names = _example_list()
    assert len(names) >= 50
    # their code objects go to a cache of their own, not the package's (which
    # holds the programs that the GPU tests and the benchmark load)
    env = dict(os.environ, SNL_KCACHE=os.path.join(ROOT, ".pytest_cache", "snl_kcache_examples"))
    os.makedirs(env["SNL_KCACHE"], exist_ok=True)

    def one(n):
        r = run_node([os.path.join(REF_EXAMPLES, n), "--compile-only", "--json"], timeout=600, env=env)
        return n, r.returncode, (r.stderr or "")[-400:]

    with ThreadPoolExecutor(max_workers=6) as ex:
        results = list(ex.map(one, names))
    unexpected = [(n, err) for n, rc, err in results if (rc != 0) != (n in REFERENCE_FAILS)]
    assert not unexpected, unexpected
    assert VOLUME_ONLY <= set(names) and not (VOLUME_ONLY & REFERENCE_FAILS)


@needs_node
@pytest.mark.gpu
def test_js_render_matches_oracle():
    """node renders the scene file through the addon; the CPU oracle renders the
    same scene (snelly_amd.scenes.sphere_plane) from the same seeds"""
    import oracle_py
    from snelly_amd import scenes
    W, H, frames = 64, 48, 4
    with tempfile.TemporaryDirectory() as td:
        rad_f, rng_f = os.path.join(td, "rad.f32"), os.path.join(td, "rng.f32")
        r = run_node([os.path.join(SCENES, "sphere_plane.html"), "--precise", "--width", str(W), "--height", str(H),
                      "--frames", str(frames), "--seed", "1234", "--out", rad_f, "--rng-out", rng_f, "--json"])
        assert r.returncode == 0, r.stderr
        g = np.fromfile(rad_f, np.float32).reshape(-1, 4)
        g_rng = np.fromfile(rng_f, np.float32).reshape(-1, 4)
    s = scenes.sphere_plane(W, H, frames)
    o = oracle_py.Oracle(s, W, H)
    rad, rng = o.new_state()
    o.render(rad, rng, frames)
    assert np.array_equal(g[:, 3], rad[:, 3])
    close = np.all(np.abs(g[:, :3] - rad[:, :3]) <= 1e-4 * np.maximum(np.abs(rad[:, :3]), 1e-3), axis=1)
    assert close.mean() >= 0.99
    assert np.all(g_rng == rng, axis=1).mean() >= 0.99


HOST_JS = os.path.join(ROOT, "snelly_amd", "js", "snelly_host.js")


@needs_node
def test_dump_scene_round_trip():
    """Snelly.dumpScene (the 'O' key) prints the settings as scene code; run
    against a fresh Snelly it restores them"""
    code = f"""
const h = require({json.dumps(HOST_JS)});
const {{ Scene }} = h.loadScene({json.dumps(os.path.join(SCENES, 'sphere_plane.html'))});
const a = new h.Snelly(new Scene(), {{ compileOnly: true, width: 64, height: 48 }});
const r = a.getRenderer(), c = a.getCamera(), m = a.getMaterials();
r.skyPower = 0.7; r.renderMode = 'ao'; r.sunLatitude = 33.5; r.skyTintUp = [0.1, 0.2, 0.3]; r.maxBounces = 7;
c.fov = 33; c.aperture = -2.5; c.position.set(1.5, 2.5, 3.5); a.getControls().target.set(0.1, 0.2, 0.3);
m.loadSurface().roughness = 0.27; m.loadSurface().diffuseAlbedo = [0.4, 0.5, 0.6];
m.loadVolume().fogEnable = true; m.loadVolume().fogTint = [0.3, 0.2, 0.1];
const code = a.dumpScene();
const b = new h.Snelly(new Scene(), {{ compileOnly: true, width: 64, height: 48 }});
new Function('snelly', code)(b);
const pick = (s) => {{ const r = s.getRenderer(), c = s.getCamera(), m = s.getMaterials();
  return [r.skyPower, r.renderMode, r.sunLatitude, r.skyTintUp, r.maxBounces, c.fov, c.aperture,
          c.position.toArray(), s.getControls().target.toArray(), m.loadSurface().roughness,
          m.loadSurface().diffuseAlbedo, m.loadVolume().fogEnable, m.loadVolume().fogTint]; }};
console.log(JSON.stringify([pick(a), pick(b), code.includes('copy-pasted console output')]));
"""
    r = subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    a, b, marker = json.loads(r.stdout.strip().splitlines()[-1])
    assert a == b and marker
    assert a[1] == "ao" and a[3] == [0.1, 0.2, 0.3]


@needs_node
@pytest.mark.gpu
def test_js_autofocus_and_png():
    """camera.auto_focus focuses on the first hit under the image centre
    (Snelly.auto_focus); saveImage writes the tonemapped frame as a PNG that
    decodes to the tonemapper's UNORM8 output, top row first"""
    from snelly_amd import host
    with tempfile.TemporaryDirectory() as td:
        png = os.path.join(td, "frame.png")
        code = f"""
const h = require({json.dumps(HOST_JS)});
const {{ Scene }} = h.loadScene({json.dumps(os.path.join(SCENES, 'sphere_plane.html'))});
const sn = new h.Snelly(new Scene(), {{ width: 40, height: 30, override: (s) => {{ s.getCamera().auto_focus = true; }} }});
const r = sn.getRenderer();
const p = r.pick(20, 15);
sn.render(3);
sn.saveImage({json.dumps(png)});
r.syncUniforms(r.shaderObj);
const rgba = h.addon.tonemapRGBA8(r.program, r.state, 40 * 30);
console.log(JSON.stringify([sn.getCamera().focalDistance, p.distance, sn.lengthScale, Array.from(rgba)]));
"""
        out = subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr
        fd, dist, ls, rgba = json.loads(out.stdout.strip().splitlines()[-1])
        assert 0.0 < dist < 100.0
        assert abs(fd - np.log10(max(1e-6, abs(dist)) / ls)) < 1e-12
        w, h, rgb = host.load_png(png)
        assert (w, h) == (40, 30)
        want = np.array(rgba, np.uint8).reshape(30, 40, 4)[::-1, :, :3]
        assert np.array_equal(rgb, want)


STUB_ADDON = """
let next = 1; const log = [];
module.exports = {
  _log: log,
  textureCreate: (...a) => { const h = { id: next++ }; log.push(['create', h.id, a.slice(1, 5).join('x')]); return h; },
  textureUpload: (h) => log.push(['upload', h.id]),
  textureDestroy: (h) => log.push(['destroy', h.id]),
  uniformTexture: (p, idx, h) => log.push(['bind', p.name, idx, h ? h.id : null]),
  uniformIndex: (p, name) => (name === 'a' ? 0 : (name === 'b' ? 1 : -1)),
  uniformSet: () => {},
  version: () => 'recording stand-in',
};
"""


@needs_node
def test_texture_object_keeps_bindings_when_respecified():
    """the WebGL stand-in, over a recording stand-in for the addon: re-uploading
    an image of the same shape updates the device texture in place; a new shape
    makes a new device texture, bound again to every sampler that held the old
    one before the old one is freed; deleting a texture frees it"""
    with tempfile.TemporaryDirectory() as td:
        stub = os.path.join(td, "stub_addon.js")
        with open(stub, "w") as f:
            f.write(STUB_ADDON)
        code = f"""
const host = require({json.dumps(HOST_JS)});
const stub = require({json.dumps(stub)});
host.setContextForTests({{}});
const gl = new host.WebGLStandIn();
const P = {{ name: 'P' }}, Q = {{ name: 'Q' }};
const t = gl.createTexture(), u = gl.createTexture();
gl.activeTexture(gl.TEXTURE0 + 7); gl.bindTexture(gl.TEXTURE_2D, t);
gl.texImage2D(gl.TEXTURE_2D, 0, gl.RGBA32F, 2, 2, 0, gl.RGBA, gl.FLOAT, new Float32Array(16));
gl.uniform1i(gl.getUniformLocation(P, 'a'), 7);
gl.uniform1i(gl.getUniformLocation(Q, 'a'), 7);
gl.texImage2D(gl.TEXTURE_2D, 0, gl.RGBA32F, 2, 2, 0, gl.RGBA, gl.FLOAT, new Float32Array(16));
gl.uniform1i(gl.getUniformLocation(P, 'a'), 7);
gl.texImage2D(gl.TEXTURE_2D, 0, gl.RGBA32F, 4, 2, 0, gl.RGBA, gl.FLOAT, new Float32Array(32));
gl.uniform1i(gl.getUniformLocation(P, 'a'), 7);
// Q's slot moves to texture u: a later re-specification of t leaves it alone
gl.activeTexture(gl.TEXTURE0 + 8); gl.bindTexture(gl.TEXTURE_2D, u);
gl.texImage2D(gl.TEXTURE_2D, 0, gl.RGBA32F, 1, 1, 0, gl.RGBA, gl.FLOAT, new Float32Array(4));
gl.uniform1i(gl.getUniformLocation(Q, 'a'), 8);
gl.activeTexture(gl.TEXTURE0 + 7);
gl.texImage2D(gl.TEXTURE_2D, 0, gl.RGBA32F, 8, 2, 0, gl.RGBA, gl.FLOAT, new Float32Array(64));
gl.uniform1i(gl.getUniformLocation(P, 'a'), 7);
gl.deleteTexture(u);
console.log(JSON.stringify(stub._log));
"""
        r = subprocess.run([NODE, "-e", code], capture_output=True, text=True,
                           env=dict(os.environ, SNL_NAPI_ADDON=stub))
        assert r.returncode == 0, r.stderr
        log = json.loads(r.stdout.strip().splitlines()[-1])
    assert log == [["create", 1, "2x2x1x4"], ["bind", "P", 0, 1], ["bind", "Q", 0, 1],
                   ["upload", 1], ["bind", "P", 0, 1],
                   ["create", 2, "4x2x1x4"], ["bind", "P", 0, 2], ["bind", "Q", 0, 2], ["destroy", 1],
                   ["bind", "P", 0, 2],
                   ["create", 3, "1x1x1x4"], ["bind", "Q", 0, 3],
                   ["create", 4, "8x2x1x4"], ["bind", "P", 0, 4], ["destroy", 2], ["bind", "P", 0, 4],
                   ["destroy", 3]]


@needs_node
def test_addon_checks_buffer_types_and_lengths():
    """the addon checks each TypedArray against what the call reads before it
    calls the library (these calls fail before any GPU work)"""
    code = f"""
const a = require({json.dumps(ADDON)});
const out = [];
function t(f) {{ try {{ f(); out.push('ok'); }} catch (e) {{ out.push(e.constructor.name + ': ' + e.message); }} }}
t(() => a.uniformSet(null, 0, 3, new Float32Array(5), 2));       // vec3 x 2 needs 6 floats
t(() => a.uniformSet(null, 0, 5, new Float32Array(1), 1));       // int needs an Int32Array
t(() => a.uniformSet(null, 0, 1, [1.0], 1));                     // not a TypedArray
t(() => a.textureCreate(null, 2, 2, 1, 4, 1, 1, new Float32Array(15)));
t(() => a.stateWrite(null, 0, new Float32Array(4)));             // no state: no size to check against
console.log(JSON.stringify(out));
"""
    r = subprocess.run([NODE, "-e", code], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out[0].startswith("RangeError") and "6 needed" in out[0]
    assert out[1].startswith("TypeError") and "Int32Array" in out[1]
    assert out[2].startswith("TypeError")
    assert out[3].startswith("RangeError") and "16 needed" in out[3]
    assert out[4].startswith("Error") and "state is NULL" in out[4]


def gl_linear(tex, coords):
    """GL LINEAR + CLAMP_TO_EDGE sampling of a texture array (..., C) whose
    leading axes are (depth,) height, width -- texel centres at (i + 0.5) / n"""
    dims = tex.shape[:-1][::-1]                 # width, height[, depth]
    idx0, frac = [], []
    for k, n in enumerate(dims):
        x = coords[:, k] * n - 0.5
        f = np.floor(x)
        idx0.append(f.astype(int))
        frac.append(x - f)
    out = np.zeros((coords.shape[0], tex.shape[-1]))
    import itertools
    for corner in itertools.product((0, 1), repeat=len(dims)):
        w = np.ones(coords.shape[0])
        ii = []
        for k, n in enumerate(dims):
            w = w * (frac[k] if corner[k] else 1.0 - frac[k])
            ii.append(np.clip(idx0[k] + corner[k], 0, n - 1))
        out += w[:, None] * tex[tuple(ii[::-1])]
    return out


@needs_node
@pytest.mark.gpu
def test_js_user_textures_on_gpu():
    """a scene that uploads a sampler3D (RGB32F) and a sampler2D (RGBA8) through
    raw WebGL calls at the user texture units and samples them in its albedo
    hook.  A convex Lambertian sphere under a white sky: a pixel's expected
    colour is the albedo at its hit point, and every sample scales all three
    channels alike, so each pixel's colour ratios match the texture value
    restated here (GL LINEAR / CLAMP_TO_EDGE) with no Monte-Carlo noise"""
    W = H = 48
    frames = 64
    with tempfile.TemporaryDirectory() as td:
        rad_f = os.path.join(td, "rad.f32")
        r = run_node([os.path.join(SCENES, "user_textures.html"), "--precise", "--width", str(W), "--height", str(H),
                      "--frames", str(frames), "--seed", "3", "--out", rad_f, "--json"])
        assert r.returncode == 0, r.stderr
        g = np.fromfile(rad_f, np.float32).reshape(-1, 4).astype(np.float64)
    img = g[:, :3] * (g[:, 3:] / (g[:, 3:] - 1.0))
    # the textures of the scene file
    t3 = np.array([[[[0.2 + 0.7 * (((i * 7 + j * 13 + k * 29 + c * 3) % 17) / 16.0) for c in range(3)]
                     for i in range(4)] for j in range(4)] for k in range(4)])
    t2 = np.array([[[(40 + ((i * 11 + j * 23 + c * 5) % 9) * 25) / 255.0 for c in range(4)]
                    for i in range(8)] for j in range(4)])
    # each pixel's (unjittered) camera ray and its hit on the unit sphere
    fh = np.tan(0.5 * np.radians(45.0))
    ys, xs = np.mgrid[0:H, 0:W]
    nx = -1.0 + 2.0 * (xs.ravel() + 0.5) / W
    ny = -1.0 + 2.0 * (ys.ravel() + 0.5) / H
    d = np.stack([fh * nx, fh * ny, -np.ones_like(nx)], -1)   # camera at z = 4 looking down -z: camX = -x
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.array([0.0, 0.0, 4.0])
    b = d @ o
    disc = b * b - (o @ o - 1.0)
    hit = disc > 1e-3
    t = -b - np.sqrt(np.maximum(disc, 0.0))
    M = np.array([[0.4124564, 0.3575761, 0.1804375], [0.2126729, 0.7151522, 0.0721750],
                  [0.0193339, 0.1191920, 0.9503041]])

    def expected(X):
        albedo = np.zeros((W * H, 3))
        pos = hit & (X[:, 0] > 0.0)
        neg = hit & (X[:, 0] <= 0.0)
        albedo[pos] = gl_linear(t3, 0.5 * X[pos] + 0.5)
        uv = np.stack([np.arctan2(X[neg, 2], X[neg, 0]) / (2 * np.pi) + 0.5, 0.5 * X[neg, 1] + 0.5], -1)
        albedo[neg] = gl_linear(t2, uv)[:, :3]
        return albedo @ M.T

    # Where the albedo is read: the sphere tracer stops at the first point whose
    # SDF is below minLengthScale (1e-4), on the ray between the analytic hit
    # and 1e-4 / cos(incidence) before it (cos(incidence) on the unit sphere is
    # sqrt(disc)); the sampled bounce reads it there, and the direct-light
    # estimate reads it at that point pushed 3 minLengthScale out along the
    # normal.  A pixel's colour is a positive mix of those reads, so its
    # chromaticity lies within theirs.
    X = o + t[:, None] * d
    Xe = o + (t - 1.5e-4 / np.sqrt(np.maximum(disc, 1e-6)))[:, None] * d
    reads = [X, Xe, X + 3e-4 * X / np.linalg.norm(X, axis=1, keepdims=True),
             Xe + 3e-4 * Xe / np.linalg.norm(Xe, axis=1, keepdims=True)]
    want = expected(X)
    # pixels whose branch or seam a last-bit difference in the hit point could flip
    safe = hit & (np.abs(X[:, 0]) > 2e-3) & ~((X[:, 0] < 0) & (np.abs(X[:, 2]) < 2e-3)) & (disc > 0.05)
    assert safe.sum() > 600

    def chroma(v):
        return v / v.sum(axis=1, keepdims=True)
    c = chroma(img[safe])
    cs = np.stack([chroma(expected(r)[safe]) for r in reads])
    assert np.all(c >= cs.min(0) - 2e-6) and np.all(c <= cs.max(0) + 2e-6)
    # every sample scales the three channels alike: the colour is the albedo's
    s = img[safe].sum(axis=1) / want[safe].sum(axis=1)
    # and that common scale is 1 in expectation (the sky is white)
    assert abs(s.mean() - 1.0) < 4.0 * s.std() / np.sqrt(s.size) + 1e-3
    # the rays that miss see the sky: rgbToXyz(1, 1, 1)
    miss = ~hit & (disc < -1e-3)
    assert np.allclose(img[miss], M.sum(axis=1), rtol=1e-5)
