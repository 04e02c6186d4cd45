// reference_host.js -- runs the reference's own host JavaScript under Node to
// record what it hands to WebGL, so that this build's two hosts (host.py and
// snelly_amd/js/snelly_host.js) can be checked against it.
//
//   node reference_host.js REFERENCE_ROOT CASES.json OUT.json
//
// Loaded unchanged from REFERENCE_ROOT, in the order of scripts/build:
// js/thirdparty/three/three.js (r79), controls/OrbitControls.js, js/color.js,
// js/materials.js, js/spectra.js, js/renderer.js, js/snelly.js.  Stood in for:
// the DOM (canvases, window events), WebGL and js/gl.js's GLU (a recording
// GLU.Shader whose replacements and uniform calls are kept, a GLU.Texture that
// keeps its texels), js/gui.js (no GUI), dat.GUI and Stats.  Each case builds
// `new Snelly(new Scene())` -- which runs Snelly.initScene, the scene's init,
// Renderer.compileShaders and loadSpectrum -- then calls Renderer.render once
// and records:
//   * the replacements of the integrator program and of the pick program
//     (__SHADER__ with the injected default hooks, __IOR_FUNC__, __DEFINES__,
//     the __MAX_*__ tokens);
//   * every uniform the first render() uploads to the integrator program, by
//     name, in call order (the scene's syncShader and Materials.syncShader
//     included), with the GL entry point used;
//   * the emission ICDF texels of Snelly.loadSpectrum (spectra.js inverseCDF)
//     and the XYZ table (color.js), as float32 hex;
//   * the texels of every texture a material binds (tabulated metals).
// A case is a JS source text defining `Scene` (a scene script) plus a size.
// Nothing here is shipped or run at test time: the tests read OUT.json.

const fs = require('fs');
const path = require('path');
const vm = require('vm');

const [REF, CASES, OUT] = process.argv.slice(2);
if (!OUT) throw new Error('usage: node reference_host.js REFERENCE_ROOT CASES.json OUT.json');

// ------------------------------------------------------------ stand-ins
// an object that accepts any property access or call (GUI folders, event
// targets, 2D contexts): reads give another such object, calls return one
function sink() {
  const f = function () { return proxy; };
  const proxy = new Proxy(f, {
    get(t, k) {
      if (k === Symbol.toPrimitive) return () => 0;
      if (k === 'then') return undefined;
      if (k in t) return t[k];
      return proxy;
    },
    set(t, k, v) { t[k] = v; return true; },
    apply() { return proxy; },
    construct() { return proxy; },
  });
  return proxy;
}

function canvas() {
  return { width: 0, height: 0, style: {}, getContext: () => sink(), addEventListener() {},
           removeEventListener() {}, appendChild() {}, setAttribute() {},
           getBoundingClientRect: () => ({ left: 0, top: 0, width: 0, height: 0 }) };
}

let CUR = null;           // the record of the case being run
const textures = [];      // every GLU.Texture created

function f32hex(arr) {
  return Buffer.from(Float32Array.from(arr).buffer).toString('hex');
}

function makeContext(width, height) {
  const g = {};
  g.window = g;
  g.console = { log() {}, info() {}, warn() {}, debug() {}, error: (...a) => process.stderr.write(a.join(' ') + '\n') };
  g.innerWidth = width;
  g.innerHeight = height;
  g.devicePixelRatio = 1;
  g.addEventListener = () => {};
  g.removeEventListener = () => {};
  g.alert = (m) => { throw new Error('alert: ' + m); };
  const els = { 'render-canvas': canvas(), 'text-canvas': canvas(), container: canvas() };
  g.document = {
    getElementById: (id) => (els[id] || (els[id] = canvas())),
    addEventListener() {}, removeEventListener() {},
    createElement: () => canvas(),
    getElementsByTagName: () => [canvas()],
    body: canvas(),
  };
  // a real clock: some scenes busy-wait on it
  g.performance = require('perf_hooks').performance;
  // images never finish loading (no network, no decoder): onload is not called
  g.Image = function () { return Object.assign(canvas(), { complete: false }); };
  g.requestAnimationFrame = () => 0;
  g.Stats = function () { return sink(); };
  g.dat = { GUI: function () { return sink(); } };
  g.GUI = function () { return sink(); };
  g.$ = g.jQuery = sink();

  const gl = new Proxy({}, {
    get(t, k) {
      if (k in t) return t[k];
      if (typeof k === 'string' && /^[A-Z0-9_]+$/.test(k)) return 0;  // enumerants
      return () => sink();
    },
  });

  class Shader {
    constructor(name, sources, replacements) {
      this.name = name;
      this.calls = [];
      const rep = {};
      if (replacements) for (const k of Object.keys(replacements)) rep[k] = String(replacements[k]);
      CUR.programs.push({ name, replacements: replacements ? rep : null });
      this.record = CUR.programs[CUR.programs.length - 1];
    }
    bind() { CUR.bound = this; }
    getProgram() { return this; }
    _rec(fn, name, vals) {
      // the uniforms of the render under test, with the program's name (the
      // integrator, then the tonemapper); NaN and infinities as strings
      const enc = (x) => {
        if (typeof x === 'boolean') return x ? 1 : 0;
        if (x === undefined) return 'undefined';
        if (typeof x === 'number' && !Number.isFinite(x)) return String(x);
        return x;
      };
      if (CUR.recording) CUR.uniforms.push([this.name, fn, name, vals.map(enc)]);
    }
    uniformIndex(name) { return name; }
    uniformTexture(name, tex) { this._rec('uniformTexture', name, [tex ? tex.id : -1]); }
    uniformI(n, i) { this._rec('uniform1i', n, [i]); }
    uniformF(n, f) { this._rec('uniform1f', n, [f]); }
    uniform2F(n, a, b) { this._rec('uniform2f', n, [a, b]); }
    uniform1Fv(n, v) { this._rec('uniform1fv', n, Array.from(v)); }
    uniform2Fv(n, v) { this._rec('uniform2fv', n, Array.from(v)); }
    uniform3F(n, a, b, c) { this._rec('uniform3f', n, [a, b, c]); }
    uniform3Fv(n, v) { this._rec('uniform3fv', n, Array.from(v)); }
    uniform4F(n, a, b, c, d) { this._rec('uniform4f', n, [a, b, c, d]); }
    uniform4Fv(n, v) { this._rec('uniform4fv', n, Array.from(v)); }
    uniformMatrix4fv(n, m) { this._rec('uniformMatrix4fv', n, Array.from(m)); }
  }

  class Texture {
    constructor(width, height, channels, isFloat, isLinear, isClamped, texels) {
      this.id = textures.length;
      textures.push(this);
      this.width = width; this.height = height; this.channels = channels;
      this.isFloat = !!isFloat; this.isLinear = !!isLinear; this.isClamped = !!isClamped;
      this.texels = texels ? Array.from(texels) : null;
      this.boundUnit = -1;
    }
    copy(texels) { this.texels = texels ? Array.from(texels) : null; }
    setFilter() {}
    bind(unit) { this.boundUnit = unit; }
  }

  g.GLU = {
    gl,
    Shader,
    Texture,
    resolveShaderSource: (d) => d,
    RenderTarget: function () { return sink(); },
    VertexBuffer: function () { return sink(); },
    fail: (m) => { throw new Error('GLU.fail: ' + m); },
    // no network: a scene's env map stays unloaded (the load callback never runs)
    loadImageAndCreateTextureInfo: () => {},
  };
  return vm.createContext(g);
}

const FILES = ['js/thirdparty/three/three.js', 'js/thirdparty/three/controls/OrbitControls.js', 'js/color.js',
               'js/materials.js', 'js/spectra.js', 'js/renderer.js', 'js/snelly.js'];
const SOURCES = FILES.map((f) => new vm.Script(fs.readFileSync(path.join(REF, f), 'utf8'), { filename: f }));

function runCase(c) {
  CUR = { name: c.name, programs: [], uniforms: [], recording: null, bound: null };
  textures.length = 0;
  // a fresh global object per case: the reference's files, then the scene
  // script, which defines Scene (and perhaps onLoad, which is not called)
  const ctx = makeContext(c.width, c.height);
  // every step under a time limit: a scene script may loop on what a stand-in returns
  const lim = { timeout: 20000 };
  for (const s of SOURCES) s.runInContext(ctx, lim);
  vm.runInContext(c.scene_js, ctx, { filename: c.name, timeout: 20000 });
  const sn = vm.runInContext('snelly = new Snelly(new Scene()); snelly', ctx, lim);
  const r = sn.getRenderer();
  // an env map named by URL never loads without a network: render as the page
  // would before the load completes, but without returning early
  const envMapUnloaded = !r.loaded;
  r.loaded = true;
  CUR.recording = true;
  vm.runInContext('snelly.getRenderer().render()', ctx, lim);
  CUR.recording = null;
  const last = (name) => {
    for (let i = CUR.programs.length - 1; i >= 0; --i) if (CUR.programs[i].name === name) return CUR.programs[i];
    return null;
  };
  const integrator = { pt: 'pathtracer', ao: 'ao', ptsimple: 'simplepathtracer', normals: 'normals' }[r.renderMode]
    || 'pathtracer';
  const tex = {};
  for (const [, fn, name, v] of CUR.uniforms) {
    if (fn !== 'uniformTexture') continue;
    const t = textures[v[0]];
    if (t && name !== 'Radiance' && name !== 'RngData')
      tex[name] = { width: t.width, height: t.height, channels: t.channels, isFloat: t.isFloat,
                    isLinear: t.isLinear, isClamped: t.isClamped, texels: t.texels ? f32hex(t.texels) : null };
  }
  return {
    name: c.name,
    width: r._width, height: r._height,
    renderMode: r.renderMode,
    envMapUnloaded,
    lengthScale: sn.lengthScale, minLengthScale: sn.minLengthScale, maxLengthScale: sn.maxLengthScale,
    integrator: last(integrator) ? last(integrator).replacements : null,
    pick: last('pick') ? last('pick').replacements : null,
    uniforms: CUR.uniforms,
    textures: tex,
  };
}

function main() {
  const cases = JSON.parse(fs.readFileSync(CASES, 'utf8'));
  const out = { reference_files: FILES, cases: [] };
  for (const c of cases) {
    process.stderr.write(c.name + "\n");
    try {
      out.cases.push(runCase(c));
    } catch (e) {
      out.cases.push({ name: c.name, error: String(e && e.message ? e.message : e) });
    }
  }
  fs.writeFileSync(OUT, JSON.stringify(out));
}

main();
