// snelly_host.js -- the Snelly scene API for Node, over libsnelly_hip.so.
//
// A scene script written for the reference (a `Scene` object with init, shader,
// syncShader, initGui, ... -- js/sceneAPI.js:16-408) runs unchanged: this
// module provides the objects it talks to -- Snelly (js/snelly.js), Renderer
// (js/renderer.js), Materials and the IOR models (js/materials.js), spectra
// (js/spectra.js), the GLU.Shader uniform setters (js/gl.js:255-406), a small
// THREE camera / OrbitControls stand-in and a no-op GUI -- and renders through
// the N-API addon (snelly_napi.node) instead of WebGL.
'use strict';

const fs = require('fs');
const path = require('path');

const HERE = __dirname;
// SNL_NAPI_ADDON: a module standing in for the addon, for tests that record
// the calls this host makes
const addon = require(process.env.SNL_NAPI_ADDON || path.join(HERE, '..', 'snelly_napi.node'));
const TABLES = JSON.parse(fs.readFileSync(path.join(HERE, '..', 'data', 'spectral_tables.json'), 'utf8'));

const T = { FLOAT: 1, VEC2: 2, VEC3: 3, VEC4: 4, INT: 5, MAT2: 6, MAT3: 7, MAT4: 8 };
const BUILD = { PRECISE: 1, FAST: 2 };

// ------------------------------------------------------------------ minimal THREE
class Vector3 {
  constructor(x = 0, y = 0, z = 0) { this.x = x; this.y = y; this.z = z; }
  set(x, y, z) { this.x = x; this.y = y; this.z = z; return this; }
  clone() { return new Vector3(this.x, this.y, this.z); }
  copy(v) { this.x = v.x; this.y = v.y; this.z = v.z; return this; }
  sub(v) { this.x -= v.x; this.y -= v.y; this.z -= v.z; return this; }
  add(v) { this.x += v.x; this.y += v.y; this.z += v.z; return this; }
  multiplyScalar(s) { this.x *= s; this.y *= s; this.z *= s; return this; }
  length() { return Math.sqrt(this.x * this.x + this.y * this.y + this.z * this.z); }
  normalize() { const l = this.length() || 1; return this.multiplyScalar(1 / l); }
  dot(v) { return this.x * v.x + this.y * v.y + this.z * v.z; }
  crossVectors(a, b) {
    return this.set(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
  }
  cross(v) { return this.crossVectors(this.clone(), v); }
  toArray() { return [this.x, this.y, this.z]; }
}

class PerspectiveCamera {
  constructor(fov, aspect, near, far) {
    this.fov = fov; this.aspect = aspect; this.near = near; this.far = far;
    this.position = new Vector3(0, 0, 0);
    this.up = new Vector3(0, 1, 0);
    // the orientation, [x, y, z, w]: set by lookAt only (as Object3D.quaternion),
    // so moving the camera or its up vector afterwards keeps the old rotation
    this.quaternion = [0, 0, 0, 1];
    this.aperture = -10.0;
    this.focalDistance = 1.0;
    this.auto_focus = false;
  }
  // Camera.lookAt (js/thirdparty/three/three.js:17025-17036)
  lookAt(v) { this.quaternion = threeQuaternionFromRotation(threeLookAt(this.position, v, this.up)); }
  updateProjectionMatrix() {}
  // (camPos, camDir, camX, camY) as js/renderer.js:520-527 derives them,
  // through three.js r79's chain (js/thirdparty/three/three.js):
  // getWorldDirection rebuilds the world matrix from the quaternion that the
  // last lookAt set (compose), takes the quaternion back out (decompose) and
  // turns (0, 0, -1) by it; camY is the current `up` carried through the world
  // matrix and normalised (transformDirection); camX = camY x camDir.
  // Matrix4.elements is a Float32Array, so every matrix write rounds to float32.
  basis() {
    const world = threeCompose(this.quaternion);
    const dir = threeApplyQuaternion(new Vector3(0, 0, -1), threeDecomposeQuaternion(world));
    const camY = threeTransformDirection(this.up.clone(), world);
    const camX = new Vector3().crossVectors(camY, dir);
    return { pos: this.position.clone(), dir, camX, camY };
  }
  getWorldDirection() { return this.basis().dir; }
}

// three.js r79's camera chain, Matrix4 elements as Float32Array
// Vector3.normalize (2460-2464): times 1 / length, NaN for a zero vector
function threeNormalize(v) { return v.multiplyScalar(1 / v.length()); }
// Matrix4.lookAt (5105-5150) into a fresh matrix
function threeLookAt(eye, target, up) {
  const te = Float32Array.of(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
  const z = threeNormalize(eye.clone().sub(target));
  if (z.dot(z) === 0) z.z = 1;
  let x = threeNormalize(new Vector3().crossVectors(up, z));
  if (x.dot(x) === 0) { z.z += 0.0001; x = threeNormalize(new Vector3().crossVectors(up, z)); }
  const y = new Vector3().crossVectors(z, x);
  te[0] = x.x; te[4] = y.x; te[8] = z.x;
  te[1] = x.y; te[5] = y.y; te[9] = z.y;
  te[2] = x.z; te[6] = y.z; te[10] = z.z;
  return te;
}
// Quaternion.setFromRotationMatrix (1048-1105): [x, y, z, w]
function threeQuaternionFromRotation(te) {
  const m11 = te[0], m12 = te[4], m13 = te[8], m21 = te[1], m22 = te[5], m23 = te[9];
  const m31 = te[2], m32 = te[6], m33 = te[10];
  const trace = m11 + m22 + m33;
  let s;
  if (trace > 0) {
    s = 0.5 / Math.sqrt(trace + 1.0);
    return [(m32 - m23) * s, (m13 - m31) * s, (m21 - m12) * s, 0.25 / s];
  } else if (m11 > m22 && m11 > m33) {
    s = 2.0 * Math.sqrt(1.0 + m11 - m22 - m33);
    return [0.25 * s, (m12 + m21) / s, (m13 + m31) / s, (m32 - m23) / s];
  } else if (m22 > m33) {
    s = 2.0 * Math.sqrt(1.0 + m22 - m11 - m33);
    return [(m12 + m21) / s, 0.25 * s, (m23 + m32) / s, (m13 - m31) / s];
  }
  s = 2.0 * Math.sqrt(1.0 + m33 - m11 - m22);
  return [(m13 + m31) / s, (m23 + m32) / s, 0.25 * s, (m21 - m12) / s];
}
// Object3D.updateMatrix = Matrix4.compose (5585-5593): makeRotationFromQuaternion
// (5068-5100), scale (1, 1, 1); the translation column is not read below
function threeCompose(q) {
  const [x, y, z, w] = q;
  const x2 = x + x, y2 = y + y, z2 = z + z;
  const xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
  const wx = w * x2, wy = w * y2, wz = w * z2;
  const te = new Float32Array(16);
  te[0] = 1 - (yy + zz); te[4] = xy - wz; te[8] = xz + wy;
  te[1] = xy + wz; te[5] = 1 - (xx + zz); te[9] = yz - wx;
  te[2] = xz - wy; te[6] = yz + wx; te[10] = 1 - (xx + yy);
  te[15] = 1;
  return te;
}
// the quaternion of Matrix4.decompose (5595-5650); the determinant's sign
// (5290-5336) is its n44 term, the others being zero
function threeDecomposeQuaternion(te) {
  let sx = Math.sqrt(te[0] * te[0] + te[1] * te[1] + te[2] * te[2]);
  const sy = Math.sqrt(te[4] * te[4] + te[5] * te[5] + te[6] * te[6]);
  const sz = Math.sqrt(te[8] * te[8] + te[9] * te[9] + te[10] * te[10]);
  const n11 = te[0], n12 = te[4], n13 = te[8], n21 = te[1], n22 = te[5], n23 = te[9];
  const n31 = te[2], n32 = te[6], n33 = te[10];
  const det = -n13 * n22 * n31 - n11 * n23 * n32 + n11 * n22 * n33 + n13 * n21 * n32 - n12 * n21 * n33
    + n12 * n23 * n31;
  if (det < 0) sx = -sx;
  const m = Float32Array.from(te);
  const inv = [1 / sx, 1 / sy, 1 / sz];
  for (let c = 0; c < 3; ++c) for (let r = 0; r < 3; ++r) m[4 * c + r] *= inv[c];
  return threeQuaternionFromRotation(m);
}
// Vector3.applyQuaternion (2240-2260)
function threeApplyQuaternion(v, q) {
  const x = v.x, y = v.y, z = v.z, [qx, qy, qz, qw] = q;
  const ix = qw * x + qy * z - qz * y, iy = qw * y + qz * x - qx * z;
  const iz = qw * z + qx * y - qy * x, iw = -qx * x - qy * y - qz * z;
  return v.set(ix * qw + iw * -qx + iy * -qz - iz * -qy, iy * qw + iw * -qy + iz * -qx - ix * -qz,
               iz * qw + iw * -qz + ix * -qy - iy * -qx);
}
// Vector3.transformDirection (2292-2306)
function threeTransformDirection(v, te) {
  const x = v.x, y = v.y, z = v.z;
  return threeNormalize(v.set(te[0] * x + te[4] * y + te[8] * z, te[1] * x + te[5] * y + te[9] * z,
                              te[2] * x + te[6] * y + te[10] * z));
}

class OrbitControls {
  constructor(camera) {
    this.camera = camera;
    this.target = new Vector3(0, 0, 0);
    this.zoomSpeed = 1.0; this.flySpeed = 0.01; this.keyPanSpeed = 1.0; this.enabled = true;
  }
  // OrbitControls.update (js/thirdparty/three/controls/OrbitControls.js:269-321)
  // without user input: the position is re-derived from spherical coordinates
  // about the target (the polar angle kept within [EPS, pi - EPS]), then the
  // camera looks at the target.  The controls' up-space rotation is the identity
  // (they are made while camera.up is still (0, 1, 0)).
  update() {
    const EPS = 0.000001;
    const p = this.camera.position, t = this.target;
    const ox = p.x - t.x, oy = p.y - t.y, oz = p.z - t.z;
    const theta = Math.max(-Infinity, Math.min(Infinity, Math.atan2(ox, oz)));
    let phi = Math.atan2(Math.sqrt(ox * ox + oz * oz), oy);
    phi = Math.max(0, Math.min(Math.PI, phi));
    phi = Math.max(EPS, Math.min(Math.PI - EPS, phi));
    const radius = Math.max(0, Math.min(Infinity, Math.sqrt(ox * ox + oy * oy + oz * oz) * 1));
    t.set(t.x + 0, t.y + 0, t.z + 0);
    const sx = radius * Math.sin(phi) * Math.sin(theta), sy = radius * Math.cos(phi);
    const sz = radius * Math.sin(phi) * Math.cos(theta);
    p.set(t.x + sx, t.y + sy, t.z + sz);
    this.camera.lookAt(t);
  }
  addEventListener() {}
}

const THREE = { Vector3, PerspectiveCamera, OrbitControls };

// ------------------------------------------------------------------ GLU shim
// GLU.Shader over a compiled program; unknown uniform names are ignored as
// WebGL ignores a null location (js/gl.js:255-260).
class Shader {
  constructor(prog) { this.prog = prog; this.uniforms = {}; }
  getProgram() { return this.prog; }
  uniformIndex(name) {
    if (!(name in this.uniforms)) this.uniforms[name] = addon.uniformIndex(this.prog, name);
    return this.uniforms[name];
  }
  _set(name, type, arr, count) {
    const idx = this.uniformIndex(name);
    if (idx < 0) return;
    addon.uniformSet(this.prog, idx, type, arr, count);
  }
  uniformI(name, i) { this._set(name, T.INT, Int32Array.of(Number(i) | 0), 1); }
  uniformF(name, f) { this._set(name, T.FLOAT, Float32Array.of(f), 1); }
  uniform2F(name, a, b) { this._set(name, T.VEC2, Float32Array.of(a, b), 1); }
  uniform3F(name, a, b, c) { this._set(name, T.VEC3, Float32Array.of(a, b, c), 1); }
  uniform4F(name, a, b, c, d) { this._set(name, T.VEC4, Float32Array.of(a, b, c, d), 1); }
  uniform1Fv(name, v) { this._set(name, T.FLOAT, Float32Array.from(v), v.length); }
  uniform2Fv(name, v) { this._set(name, T.VEC2, Float32Array.from(v), Math.max(1, v.length / 2)); }
  uniform3Fv(name, v) { this._set(name, T.VEC3, Float32Array.from(v), Math.max(1, v.length / 3)); }
  uniform4Fv(name, v) { this._set(name, T.VEC4, Float32Array.from(v), Math.max(1, v.length / 4)); }
  uniformMatrix4fv(name, m) { this._set(name, T.MAT4, Float32Array.from(m), 1); }
  uniformTexture(name, tex) {
    const idx = this.uniformIndex(name);
    if (idx >= 0 && tex && tex.handle) addon.uniformTexture(this.prog, idx, tex.handle);
  }
  bind() {}
}

// ------------------------------------------------------------------ env-map images
// The reference's GLU.loadImageAndCreateTextureInfo (js/gl.js:532-559) lets the
// browser decode the image and uploads it as SRGB8.  Here a local 8-bit RGB or
// RGBA PNG is decoded (zlib from Node) and its sRGB transfer function applied.
function loadPng(file) {
  const zlib = require('zlib');
  const buf = fs.readFileSync(file);
  const sig = [0x89, 0x50, 0x4e, 0x47, 0x0d, 0x0a, 0x1a, 0x0a];
  if (!sig.every((b, i) => buf[i] === b)) throw new Error(file + ': not a PNG file');
  let pos = 8, w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
  const idat = [];
  while (pos < buf.length) {
    const n = buf.readUInt32BE(pos), kind = buf.toString('latin1', pos + 4, pos + 8);
    const chunk = buf.subarray(pos + 8, pos + 8 + n);
    if (kind === 'IHDR') { w = chunk.readUInt32BE(0); h = chunk.readUInt32BE(4); depth = chunk[8]; ctype = chunk[9]; interlace = chunk[12]; }
    else if (kind === 'IDAT') idat.push(chunk);
    else if (kind === 'IEND') break;
    pos += 12 + n;
  }
  if (depth !== 8 || (ctype !== 2 && ctype !== 6) || interlace) throw new Error(file + ': only 8-bit RGB/RGBA non-interlaced PNGs are read');
  const bpp = ctype === 2 ? 3 : 4, stride = w * bpp;
  const raw = zlib.inflateSync(Buffer.concat(idat));
  const rgb = new Uint8Array(w * h * 3);
  let prev = new Uint8Array(stride);
  for (let y = 0; y < h; ++y) {
    const ft = raw[y * (stride + 1)], line = raw.subarray(y * (stride + 1) + 1, (y + 1) * (stride + 1));
    const cur = new Uint8Array(stride);
    for (let x = 0; x < stride; ++x) {
      const a = x >= bpp ? cur[x - bpp] : 0, b = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
      let pred = 0;
      if (ft === 1) pred = a;
      else if (ft === 2) pred = b;
      else if (ft === 3) pred = (a + b) >> 1;
      else if (ft === 4) {
        const p = a + b - c, pa = Math.abs(p - a), pb = Math.abs(p - b), pc = Math.abs(p - c);
        pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
      }
      cur[x] = (line[x] + pred) & 255;
    }
    for (let x = 0; x < w; ++x) for (let k = 0; k < 3; ++k) rgb[3 * (y * w + x) + k] = cur[x * bpp + k];
    prev = cur;
  }
  return { width: w, height: h, rgb };
}
// an RGBA8 image (rows bottom first, as gl.readPixels returns them) as a PNG
// file, top row first -- the canvas.toBlob('image/png') of the 'P' key
// (js/snelly.js:791-806)
const CRC_TABLE = (() => {
  const t = new Uint32Array(256);
  for (let n = 0; n < 256; ++n) {
    let c = n;
    for (let k = 0; k < 8; ++k) c = (c & 1) ? (0xedb88320 ^ (c >>> 1)) : (c >>> 1);
    t[n] = c >>> 0;
  }
  return t;
})();
function crc32(buf) {
  let c = 0xffffffff;
  for (let i = 0; i < buf.length; ++i) c = CRC_TABLE[(c ^ buf[i]) & 255] ^ (c >>> 8);
  return (c ^ 0xffffffff) >>> 0;
}
function savePng(file, width, height, rgba) {
  const zlib = require('zlib');
  const raw = Buffer.alloc((width * 4 + 1) * height);
  for (let y = 0; y < height; ++y) {
    const src = (height - 1 - y) * width * 4;
    raw[y * (width * 4 + 1)] = 0;
    for (let i = 0; i < width * 4; ++i) raw[y * (width * 4 + 1) + 1 + i] = rgba[src + i];
  }
  const chunk = (kind, data) => {
    const len = Buffer.alloc(4); len.writeUInt32BE(data.length);
    const body = Buffer.concat([Buffer.from(kind, 'latin1'), data]);
    const crc = Buffer.alloc(4); crc.writeUInt32BE(crc32(body));
    return Buffer.concat([len, body, crc]);
  };
  const ihdr = Buffer.alloc(13);
  ihdr.writeUInt32BE(width, 0); ihdr.writeUInt32BE(height, 4);
  ihdr[8] = 8; ihdr[9] = 6; ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0;
  fs.writeFileSync(file, Buffer.concat([Buffer.from([0x89, 0x50, 0x4e, 0x47, 0x0d, 0x0a, 0x1a, 0x0a]),
    chunk('IHDR', ihdr), chunk('IDAT', zlib.deflateSync(raw)), chunk('IEND', Buffer.alloc(0))]));
}
function srgbToLinear(bytes) {
  const out = new Float32Array(bytes.length);
  for (let i = 0; i < bytes.length; ++i) {
    const c = bytes[i] / 255.0;
    out[i] = c <= 0.04045 ? c / 12.92 : Math.pow((c + 0.055) / 1.055, 2.4);
  }
  return out;
}

class Texture {
  constructor(ctx, width, height, channels, isFloat, isLinear, isClamped, texels, depth = 1) {
    this.width = width; this.height = height; this.channels = channels;
    this.handle = addon.textureCreate(ctx, width, height, depth, channels, isLinear ? 1 : 0, isClamped ? 1 : 0,
                                      texels ? Float32Array.from(texels) : null);
  }
  copy(texels) { addon.textureUpload(this.handle, Float32Array.from(texels)); }
  bind() {}
}

// ------------------------------------------------------------------ WebGL stand-in
// Scenes that bring their own textures drive WebGL2 directly through
// snelly.getGLContext() (exampleScenes/meshes/bunny.html:179-258): create,
// bind to a unit, texImage2D/3D, texParameteri, then in syncShader bind the
// unit to a sampler uniform with uniform1i.  This object keeps that state and
// turns a sampler binding into a device texture bound to the program.
const GLC = {
  TEXTURE0: 0x84C0, TEXTURE_2D: 0x0DE1, TEXTURE_3D: 0x806F, TEXTURE_MIN_FILTER: 0x2801,
  TEXTURE_MAG_FILTER: 0x2800, TEXTURE_WRAP_S: 0x2802, TEXTURE_WRAP_T: 0x2803, TEXTURE_WRAP_R: 0x8072,
  NEAREST: 0x2600, LINEAR: 0x2601, LINEAR_MIPMAP_LINEAR: 0x2703, NEAREST_MIPMAP_NEAREST: 0x2700,
  CLAMP_TO_EDGE: 0x812F, REPEAT: 0x2901, MIRRORED_REPEAT: 0x8370, FLOAT: 0x1406, HALF_FLOAT: 0x140B,
  UNSIGNED_BYTE: 0x1401, RED: 0x1903, RG: 0x8227, RGB: 0x1907, RGBA: 0x1908, LUMINANCE: 0x1909,
  R8: 0x8229, R16F: 0x822D, R32F: 0x822E, RG32F: 0x8230, RGB8: 0x8051, RGBA8: 0x8058, RGB32F: 0x8815,
  RGBA32F: 0x8814, RGB16F: 0x881B, RGBA16F: 0x881A, SRGB8: 0x8C41, SRGB8_ALPHA8: 0x8C43,
  UNPACK_FLIP_Y_WEBGL: 0x9240, UNPACK_ALIGNMENT: 0x0CF5,
};
const FORMAT_CHANNELS = { [GLC.RED]: 1, [GLC.LUMINANCE]: 1, [GLC.RG]: 2, [GLC.RGB]: 3, [GLC.RGBA]: 4 };

class GLTexture {
  constructor() {
    this.w = 1; this.h = 1; this.d = 1; this.ch = 4; this.data = null;
    this.linear = true; this.clamp = false; this.handle = null; this.dirty = true;
    this.shape = '';              // w x h x d x ch, linear, clamp of the device texture in `handle`
    this.bindings = new Map();    // program -> sampler uniform indices bound to this texture
  }
}

class WebGLStandIn {
  constructor() {
    Object.assign(this, GLC);
    this.unit = 0;
    this.units = {};
    this.slotTex = new Map();   // program -> (sampler uniform index -> the GLTexture bound there)
  }
  createTexture() { return new GLTexture(); }
  // the runtime resets samplers still bound to a deleted texture to an empty
  // unit, which reads (0, 0, 0, 1) as in WebGL
  deleteTexture(t) {
    if (t && t.handle) { addon.textureDestroy(t.handle); t.handle = null; }
    if (t) t.bindings.clear();
  }
  activeTexture(u) { this.unit = u - GLC.TEXTURE0; }
  bindTexture(target, t) { this.units[this.unit] = t; }
  pixelStorei() {}
  generateMipmap() {}
  texParameteri(target, pname, param) {
    const t = this.units[this.unit];
    if (!t) return;
    if (pname === GLC.TEXTURE_MAG_FILTER) t.linear = param !== GLC.NEAREST;
    if (pname === GLC.TEXTURE_WRAP_S || pname === GLC.TEXTURE_WRAP_T || pname === GLC.TEXTURE_WRAP_R)
      t.clamp = param === GLC.CLAMP_TO_EDGE;
    t.dirty = true;
  }
  _store(t, w, h, d, format, type, data) {
    t.w = Math.max(1, w | 0); t.h = Math.max(1, h | 0); t.d = Math.max(1, d | 0);
    t.ch = FORMAT_CHANNELS[format] || 4;
    const n = t.w * t.h * t.d * t.ch;
    const f = new Float32Array(n);
    if (data && data.length !== undefined) {
      const byte = type === GLC.UNSIGNED_BYTE;
      for (let i = 0; i < Math.min(n, data.length); ++i) f[i] = byte ? data[i] / 255.0 : data[i];
    }
    t.data = f;
    t.dirty = true;
  }
  // texImage2D(target, level, ifmt, w, h, border, format, type, data) or the
  // image-source form texImage2D(target, level, ifmt, format, type, image)
  texImage2D(target, level, ifmt, a, b, c, d, e, f) {
    const t = this.units[this.unit];
    if (!t || level !== 0) return;
    if (f === undefined && e === undefined) {
      const img = c || {};
      this._store(t, img.width || 1, img.height || 1, 1, a, b, img.data || null);
    } else {
      this._store(t, a, b, 1, d, e, f);
    }
  }
  texImage3D(target, level, ifmt, w, h, depth, border, format, type, data) {
    const t = this.units[this.unit];
    if (t && level === 0) this._store(t, w, h, depth, format, type, data);
  }
  getUniformLocation(prog, name) { return { prog, name, idx: addon.uniformIndex(prog, name) }; }
  // The device copy of a texture object.  A texture object keeps its
  // bindings when its image is re-specified (WebGL): the same shape updates the
  // device texture in place; a new shape or filter gets a new device texture,
  // bound again to every sampler that held the old one before that is freed.
  _upload(t) {
    const ctx = SNELLY ? SNELLY.ctx : null;
    if (!ctx) return null;
    if (t.dirty || !t.handle) {
      const data = t.data || new Float32Array(t.w * t.h * t.d * t.ch);
      const shape = `${t.w}x${t.h}x${t.d}x${t.ch}:${t.linear ? 1 : 0}${t.clamp ? 1 : 0}`;
      if (t.handle && shape === t.shape) {
        addon.textureUpload(t.handle, data);
      } else {
        const old = t.handle;
        t.handle = addon.textureCreate(ctx, t.w, t.h, t.d, t.ch, t.linear ? 1 : 0, t.clamp ? 1 : 0, data);
        t.shape = shape;
        for (const [prog, idxs] of t.bindings) for (const idx of idxs) addon.uniformTexture(prog, idx, t.handle);
        if (old) addon.textureDestroy(old);
      }
      t.dirty = false;
    }
    return t.handle;
  }
  uniform1i(loc, v) {
    if (!loc || loc.idx < 0) return;
    const t = this.units[v];
    if (t instanceof GLTexture) {
      const h = this._upload(t);
      if (h) {
        addon.uniformTexture(loc.prog, loc.idx, h);
        if (!this.slotTex.has(loc.prog)) this.slotTex.set(loc.prog, new Map());
        const slots = this.slotTex.get(loc.prog);
        const prev = slots.get(loc.idx);
        if (prev && prev !== t && prev.bindings.has(loc.prog)) prev.bindings.get(loc.prog).delete(loc.idx);
        slots.set(loc.idx, t);
        if (!t.bindings.has(loc.prog)) t.bindings.set(loc.prog, new Set());
        t.bindings.get(loc.prog).add(loc.idx);
      }
      return;
    }
    addon.uniformSet(loc.prog, loc.idx, T.INT, Int32Array.of(v), 1);
  }
  uniform1f(loc, x) { if (loc && loc.idx >= 0) addon.uniformSet(loc.prog, loc.idx, T.FLOAT, Float32Array.of(x), 1); }
  uniform2f(loc, x, y) { if (loc && loc.idx >= 0) addon.uniformSet(loc.prog, loc.idx, T.VEC2, Float32Array.of(x, y), 1); }
  uniform3f(loc, x, y, z) { if (loc && loc.idx >= 0) addon.uniformSet(loc.prog, loc.idx, T.VEC3, Float32Array.of(x, y, z), 1); }
  uniform4f(loc, x, y, z, w) { if (loc && loc.idx >= 0) addon.uniformSet(loc.prog, loc.idx, T.VEC4, Float32Array.of(x, y, z, w), 1); }
  uniform1fv(loc, v) { if (loc && loc.idx >= 0) addon.uniformSet(loc.prog, loc.idx, T.FLOAT, Float32Array.from(v), v.length); }
  uniform3fv(loc, v) { if (loc && loc.idx >= 0) addon.uniformSet(loc.prog, loc.idx, T.VEC3, Float32Array.from(v), v.length / 3); }
}

// an Image that never loads: scene assets referenced by URL are not
// available offline, so a scene waiting on one stays not-ready (isReady)
class OfflineImage {
  constructor() { this.width = 1; this.height = 1; this.src = ''; this.crossOrigin = ''; }
  addEventListener() {}
}

// the one WebGL context of the process (the global `gl` of a scene script)
const GL = new WebGLStandIn();

const GLU = {
  T, BUILD, Shader, Texture,
  fail(msg) { throw new Error('[snelly] ' + msg); },
  // js/gl.js:532-559: a 1x1 black texture, replaced when the image loads
  // (never, offline)
  loadImageAndCreateTextureInfo(url, callback) {
    const gl = GL;
    const tex = gl.createTexture();
    gl.bindTexture(GLC.TEXTURE_2D, tex);
    gl.texImage2D(GLC.TEXTURE_2D, 0, GLC.SRGB8, 1, 1, 0, GLC.RGB, GLC.UNSIGNED_BYTE, new Uint8Array([0, 0, 0, 0]));
    return { width: 1, height: 1, texture: tex, url };
  },
  applyShaderReplacements(source, replacements) {
    let s = source;
    for (const k of Object.keys(replacements)) s = s.replace(new RegExp(k, 'g'), replacements[k]);
    return s;
  },
};

// ------------------------------------------------------------------ spectra (js/spectra.js)
class Spectrum {
  constructor(name, desc) { this._name = name; this._desc = desc; }
  getName() { return this._name; }
  // Float32Array storage as the reference: every CDF/ICDF value is rounded to float
  inverseCDF(lmin, lmax, n) {
    const cdf = new Float32Array(n + 1);
    const dl = (lmax - lmin) / n;
    for (let i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + this.spectrum(lmin + dl * i);
    const sum = cdf[n];
    for (let i = 0; i < n; ++i) cdf[i + 1] /= sum;
    cdf[n] = 1.0;
    const icdf = new Float32Array(4 * n);
    let idx = 0;
    for (let k = 0; k < 4 * n; ++k) {
      const xi = Math.min((k + 1) / (4 * n), 1.0);
      while (cdf[idx] < xi) idx++;
      idx = Math.max(1, idx);
      const lo = cdf[idx - 1], hi = cdf[idx];
      icdf[k] = (idx - 1.0) / n + (1.0 / n) * (xi - lo) / (hi - lo);
    }
    return icdf;
  }
}
class BlackbodySpectrum extends Spectrum {
  constructor(name, desc, temperature) { super(name, desc); this.skyTemperature = temperature; }
  spectrum(wl) {
    const b = 1.43877737467e7 / (wl * this.skyTemperature);
    const l = wl / 360.0;
    return 1.0 / (l * l * l * l * l * (Math.exp(b) - 1.0));
  }
}
class FlatSpectrum extends Spectrum {
  constructor(name, desc, lo, hi) { super(name, desc); this['Minimum wavelength'] = lo; this['Maximum wavelength'] = hi; }
  spectrum(wl) { return (wl < this['Minimum wavelength'] || wl > this['Maximum wavelength']) ? 0.0 : 1.0; }
}
class MonochromaticSpectrum extends Spectrum {
  constructor(name, desc, wl) { super(name, desc); this.wavelength = wl; }
  inverseCDF(lmin, lmax, n) {
    const wl = Math.max(lmin, Math.min(this.wavelength, lmax));
    return new Float32Array(4 * n).fill((wl - lmin) / (lmax - lmin));
  }
}

function wavelengthToXYZTable() {
  const out = new Float32Array(4096);
  TABLES.xyz_lut.forEach((v, i) => { out[4 * i] = v[0]; out[4 * i + 1] = v[1]; out[4 * i + 2] = v[2]; });
  return out;
}

// ------------------------------------------------------------------ materials (js/materials.js)
class Material { constructor(name) { this._name = name; } getName() { return this._name; } }
// The repr() of the materials and Snelly.dumpScene print settings as scene
// code (js/materials.js:56-70, 222-243, 488-495, 740-749; js/snelly.js:296-406):
// the user pastes it into a scene, so the lines keep the reference's form.
const v3s = (v) => `[${v[0]}, ${v[1]}, ${v[2]}]`;

class Surface extends Material {
  repr() {
    return `
    surface.roughness = ${this.roughness};
    surface.ior = ${this.ior};
    surface.diffuseAlbedo = ${v3s(this.diffuseAlbedo)};
    surface.specAlbedo = ${v3s(this.specAlbedo)};
    surface.subsurface = ${this.subsurface};
    surface.subsurfaceAlbedoRGB = ${v3s(this.subsurfaceAlbedoRGB)};
    surface.subsurfaceMFP = ${this.subsurfaceMFP};
    surface.subsurfaceAnisotropy = ${this.subsurfaceAnisotropy};
    surface.subsurfaceDiffuseWeight = ${this.subsurfaceDiffuseWeight};
    `;
  }
  constructor(name) {
    super(name);
    this.diffuseAlbedo = [1, 1, 1]; this.specAlbedo = [1, 1, 1]; this.roughness = 0.1; this.ior = 1.5;
    this.subsurface = 0.0; this.subsurfaceAlbedoRGB = [0.5, 0.5, 0.5]; this.subsurfaceMFP = 0.0;
    this.subsurfaceAnisotropy = 0.0; this.subsurfaceDiffuseWeight = 1.0;
  }
  syncShader(s) {
    s.uniform3Fv('surfaceDiffuseAlbedoRGB', this.diffuseAlbedo);
    s.uniform3Fv('surfaceSpecAlbedoRGB', this.specAlbedo);
    s.uniformF('surfaceRoughness', this.roughness);
    s.uniformF('surfaceIor', this.ior);
    s.uniformF('subsurface', this.subsurface);
    s.uniform3Fv('subsurfaceAlbedoRGB', this.subsurfaceAlbedoRGB);
    s.uniformF('subsurfaceMFP', this.subsurfaceMFP);
    s.uniformF('subsurfaceAnisotropy', this.subsurfaceAnisotropy);
    s.uniformF('subsurfaceDiffuseWeight', this.subsurfaceDiffuseWeight);
  }
  initGui() {}
}

class Volume extends Material {
  repr() {
    return `
    volume.atmosphereMinX = ${this.atmosphereMinX};
    volume.atmosphereMinY = ${this.atmosphereMinY};
    volume.atmosphereMinZ = ${this.atmosphereMinZ};
    volume.atmosphereMaxX = ${this.atmosphereMaxX};
    volume.atmosphereMaxY = ${this.atmosphereMaxY};
    volume.atmosphereMaxZ = ${this.atmosphereMaxZ};
    volume.lof10_mfp = ${this.lof10_mfp};
    volume.scatteringColor = ${v3s(this.scatteringColor)};
    volume.absorptionColor = ${v3s(this.absorptionColor)};
    volume.emission = ${this.emission};
    volume.emissionColor = ${v3s(this.emissionColor)};
    volume.fogEmission = ${v3s(this.fogEmission)};
    volume.fogTint = ${v3s(this.fogTint)};
    volume.fogMFP = ${this.fogMFP};
    volume.fogMaxOpticalDepth = ${this.fogMaxOpticalDepth};
    volume.fogEnable = ${this.fogEnable};
    `;
  }
  constructor(name) {
    super(name);
    this.atmosphereMinX = -10; this.atmosphereMaxX = 10; this.atmosphereMinY = 0; this.atmosphereMaxY = 10;
    this.atmosphereMinZ = -10; this.atmosphereMaxZ = 10; this.lof10_mfp = 1.0;
    this.scatteringColor = [0, 0, 0]; this.absorptionColor = [0, 0, 0]; this.anisotropy = 0.0;
    this.emission = 0.0; this.emissionColor = [1, 1, 1];
    this.fogEnable = false; this.fogMFP = 0.0; this.fogMaxOpticalDepth = 5.0;
    this.fogEmission = [1, 1, 1]; this.fogTint = [0.5, 0.5, 0.5];
    // grey extinction of a scene's SDF_VOLUME medium (the property the
    // reference's volume scenes set, e.g. exampleScenes/volumes/pyroclastic.html),
    // passed to VOLUME_EXTINCTION / VOLUME_EXTINCTION_MAX
    this.extinction = 1.0;
  }
  syncShader(s) {
    const ext = [0, 1, 2].map(c => this.absorptionColor[c] + this.scatteringColor[c]);
    s.uniformF('atmosphereMFP', Math.pow(10, this.lof10_mfp));
    s.uniform3Fv('atmosphereExtinctionCoeffRGB', ext);
    s.uniform3Fv('atmosphereScatteringCoeffRGB', this.scatteringColor.slice());
    s.uniformF('atmosphereAnisotropy', this.anisotropy);
    s.uniform3Fv('atmosphereBoundsMin', [this.atmosphereMinX, this.atmosphereMinY, this.atmosphereMinZ]);
    s.uniform3Fv('atmosphereBoundsMax', [this.atmosphereMaxX, this.atmosphereMaxY, this.atmosphereMaxZ]);
    s.uniformF('volumeEmission', this.emission);
    s.uniform3Fv('volumeEmissionColorRGB', this.emissionColor);
    s.uniformI('fogEnable', this.fogEnable);
    s.uniformF('fogMFP', Math.pow(10.0, this.fogMFP));
    s.uniformF('fogMaxOpticalDepth', this.fogMaxOpticalDepth);
    s.uniform3Fv('fogEmission', this.fogEmission);
    s.uniform3Fv('fogTint', this.fogTint);
    s.uniformF('volumeExtinction', this.extinction);
  }
}

class Metal extends Material {
  repr() {
    return `
    metal.roughness = ${this.roughness};
    metal.specAlbedo = ${v3s(this.specAlbedo)};
    `;
  }
  constructor(name) { super(name); this.roughness = 0.02; this.specAlbedo = [1, 1, 1]; }
  syncShader(s) {
    s.uniformF('metalRoughness', this.roughness);
    s.uniform3Fv('metalSpecAlbedoRGB', this.specAlbedo);
  }
  initGui() {}
}

class TabulatedMetal extends Metal {
  constructor(name, nk) { super(name); this.nk = nk; this.ior_tex = null; this.k_tex = null; }
  ior() {
    const delta = 1.0 / (750.0 - 390.0);
    return `
float IOR_METAL(float wavelength_nm)
{
    float u = (wavelength_nm - 390.0) * ${delta};
    return texture(iorTex, vec2(u, 0.5)).r;
}
float K_METAL(float wavelength_nm)
{
    float u = (wavelength_nm - 390.0) * ${delta};
    return texture(kTex, vec2(u, 0.5)).r;
}
    `;
  }
  textures(ctx) {
    if (!this.ior_tex) {
      this.ior_tex = new Texture(ctx, 64, 1, 1, true, true, true, this.nk.n);
      this.k_tex = new Texture(ctx, 64, 1, 1, true, true, true, this.nk.k);
    }
  }
  syncShader(s) {
    if (this.ior_tex) s.uniformTexture('iorTex', this.ior_tex);
    // the reference binds k under "k_tex" while the shader samples "kTex" (js/materials.js:665)
    if (this.k_tex) s.uniformTexture('k_tex', this.k_tex);
    super.syncShader(s);
  }
}

class Dielectric extends Material {
  repr() {
    return `
    dielectric.absorptionColor = ${v3s(this.absorptionColor)};
    dielectric.absorptionScale = ${this.absorptionScale}; // mfp in multiples of scene scale
    dielectric.roughness = ${this.roughness};
    dielectric.iorVal = ${this.iorVal};
    dielectric.abbe = ${this.abbe};
    `;
  }
  constructor(name) {
    super(name);
    this.roughness = 0.005; this.absorptionScale = -1.0; this.absorptionColor = [1, 1, 1];
    this.absorptionRGB = [0, 0, 0]; this.specAlbedo = [1, 1, 1];
  }
  syncShader(s) {
    s.uniformF('dieleRoughness', this.roughness);
    s.uniform3Fv('dieleSpecAlbedoRGB', this.specAlbedo);
    for (let c = 0; c < 3; ++c)
      this.absorptionRGB[c] = snellyGlobal().lengthScale / Math.max(this.absorptionScale, 1.0e-3) *
                              Math.max(0.0, 1.0 - this.absorptionColor[c]);
    s.uniform3Fv('dieleAbsorptionRGB', this.absorptionRGB);
  }
  initGui() { if (this.absorptionScale < 0.0) this.absorptionScale = snellyGlobal().maxLengthScale; }
}

const coeffUniforms = (c) => c.map((_, i) => `uniform float _C${i + 1};\n`).join('');
const syncCoeffs = (s, c) => c.forEach((v, i) => s.uniformF(`_C${i + 1}`, v));

class AbbeDielectric extends Dielectric {
  constructor(name, iorVal, abbe) { super(name); this.iorVal = iorVal; this.abbe = abbe; }
  ior() {
    return `
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
`;
  }
  syncShader(s) { s.uniformF('_iorVal', this.iorVal); s.uniformF('_abbe', this.abbe); super.syncShader(s); }
}

class SellmeierDielectric extends Dielectric {
  constructor(name, coeffs, squared = true) { super(name); this.coeffs = coeffs; this.squared = squared; }
  ior() {
    const terms = (this.coeffs.length - 1) / 2;
    let f = '1.0 + _C1 ';
    for (let t = 1; t <= terms; ++t) {
      const pole = this.squared ? `_C${2 * t + 1}*_C${2 * t + 1}` : `_C${2 * t + 1}`;
      f += `+ _C${2 * t}*l2/(l2 - ${pole})`;
    }
    return coeffUniforms(this.coeffs) + `
float IOR_DIELE(float wavelength_nm)
{
    float wavelength_um = 1.0e-3*wavelength_nm;
    float l2 = wavelength_um*wavelength_um;
    float n2 = ${f};
    return max(sqrt(abs(n2)), 1.0e-3);
}`;
  }
  syncShader(s) { syncCoeffs(s, this.coeffs); super.syncShader(s); }
}
class Sellmeier2Dielectric extends SellmeierDielectric { constructor(name, coeffs) { super(name, coeffs, false); } }

class PolyanskiyDielectric extends Dielectric {
  constructor(name, coeffs) { super(name); this.coeffs = coeffs; }
  ior() {
    return coeffUniforms(this.coeffs) + `
float IOR_DIELE(float wavelength_nm)
{
    float wavelength_um = 1.0e-3*wavelength_nm;
    float l = wavelength_um;
    float n2 =  _C1 + _C2*pow(l, _C3)/(l*l - pow(_C4, _C5));
    return max(sqrt(abs(n2)), 1.0e-3);
}`;
  }
  syncShader(s) { syncCoeffs(s, this.coeffs); super.syncShader(s); }
}

class CauchyDielectric extends Dielectric {
  constructor(name, coeffs) { super(name); this.coeffs = coeffs; }
  ior() {
    const terms = (this.coeffs.length - 1) / 2;
    let f = '_C1';
    for (let t = 1; t <= terms; ++t) f += ` + _C${2 * t}*pow(l, _C${2 * t + 1})`;
    return coeffUniforms(this.coeffs) + `
float IOR_DIELE(float wavelength_nm)
{
    float wavelength_um = 1.0e-3*wavelength_nm;
    float l = wavelength_um;
    float n = ${f};
    return max(n, 1.0e-3);
}`;
  }
  syncShader(s) { syncCoeffs(s, this.coeffs); super.syncShader(s); }
}

// the registry of js/materials.js:1125-1171 (coefficients as coded there)
const DIELECTRIC_REGISTRY = [
  ['Abbe dielectric', () => new AbbeDielectric('Abbe dielectric', 1.5, 0.0)],
  ['Glass (BK7)', () => new SellmeierDielectric('Glass (BK7)', [0.0, 1.03961212, 0.00600069867, 0.231792344, 0.0200179144, 1.01046945, 103.560653])],
  ['Glass (K7)', () => new Sellmeier2Dielectric('Glass (K7)', [0.0, 1.1273555, 0.00720341707, 0.124412303, 0.0269835916, 0.827100531, 100.384588])],
  ['Glass (F5)', () => new Sellmeier2Dielectric('Glass (F5)', [0.0, 1.3104463, 0.00958633048, 0.19603426, 0.0457627627, 0.96612977, 115.011883])],
  ['Glass (LAFN7)', () => new Sellmeier2Dielectric('Glass (LAFN7)', [0.0, 1.66842615, 0.0103159999, 0.298512803, 0.0469216348, 1.0774376, 82.5078509])],
  ['Glass (LASF35)', () => new Sellmeier2Dielectric('Glass (LASF35)', [0.0, 2.45505861, 0.0135670404, 0.453006077, 0.054580302, 2.3851308, 167.904715])],
  ['Glass (N-LAK33A)', () => new Sellmeier2Dielectric('Glass (N-LAK33A)', [0.0, 1.44116999, 0.00680933877, 0.571749501, 0.0222291824, 1.16605226, 80.9379555])],
  ['Glass (N-FK51A)', () => new SellmeierDielectric('Glass (N-FK51A)', [0.0, 0.97124781, 0.00472301995, 0.216901417, 0.0153575612, 0.90465166, 168.68133])],
  ['Glass (SF4)', () => new Sellmeier2Dielectric('Glass (SF4)', [0.0, 1.61957826, 0.0125502104, 0.339493189, 0.0544559822, 1.02566931, 117.652222])],
  ['Glass (SF67)', () => new Sellmeier2Dielectric('Glass (SF67)', [0.0, 1.97464225, 0.0145772324, 0.467095921, 0.0669790359, 2.43154209, 157.444895])],
  ['Water', () => new Sellmeier2Dielectric('Water', [0.0, 5.67252e-1, 5.08555046e-3, 1.736581e-1, 1.8149386e-2, 2.12153e-2, 2.61726e-2, 1.1384932e-1, 1.073888e1])],
  ['Ethanol', () => new Sellmeier2Dielectric('Ethanol', [0.0, 0.83189, 0.00930, -0.15582, -49.45200])],
  ['Polycarbonate', () => new Sellmeier2Dielectric('Polycarbonate', [0.0, 0.83189, 0.00930, -0.15582, -49.45200])],
  ['Glycerol', () => new CauchyDielectric('Glycerol', [1.45797, 0.00598, -2, -0.00036, -4])],
  ['Liquid Crystal (E7)', () => new CauchyDielectric('Liquid Crystal (E7)', [1.4990, 0.0072, -2, 0.0003, -4])],
  ['Diamond', () => new SellmeierDielectric('Diamond', [0.0, 0.3306, 0.175, 4.3356, 0.1060])],
  ['Quartz', () => new SellmeierDielectric('Quartz', [0.0, 0.6961663, 0.0684043, 0.4079426, 0.1162414, 0.8974794, 9.896161])],
  ['Fused Silica', () => new SellmeierDielectric('Fused Silica', [0.0, 0.6961663, 0.0684043, 0.4079426, 0.1162414, 0.8974794, 9.896161])],
  ['Sapphire', () => new SellmeierDielectric('Sapphire', [0.0, 1.5039759, 0.0740288, 0.55069141, 0.1216529, 6.5927379, 20.072248])],
  ['Sodium Chloride', () => new SellmeierDielectric('Sodium Chloride', [0.00055, 0.19800, 0.050, 0.48398, 0.100, 0.38696, 0.128])],
  ['Proustite', () => new PolyanskiyDielectric('Proustite', [7.483, 0.474, 0.0, 0.09, 1.0])],
  ['Rutile', () => new PolyanskiyDielectric('Rutile', [5.913, 0.2441, 0.0, 0.0803, 1.0])],
  ['Silver Chloride', () => new PolyanskiyDielectric('Silver Chloride', [4.00804, 0.079086, 0.0, 0.04584, 1.0])],
];

class Materials {
  constructor() {
    this.dielectrics = {}; this.metals = {};
    for (const [name, make] of DIELECTRIC_REGISTRY) this.dielectrics[name] = make();
    for (const name of Object.keys(TABLES.metals)) this.metals[name] = new TabulatedMetal(name, TABLES.metals[name]);
    this.surfaceObj = new Surface('Surface');
    this.volumeObj = new Volume('Volume');
    this.loadDielectric('Glass (BK7)');
    this.loadMetal('Aluminium');
  }
  addDielectric(m) { this.dielectrics[m.getName()] = m; }
  getDielectrics() { return this.dielectrics; }
  loadDielectric(name) { this.dielectricObj = this.dielectrics[name]; return this.dielectricObj; }
  getLoadedDielectric() { return this.dielectricObj; }
  getDielectric() { return this.dielectricObj; }
  addMetal(m) { this.metals[m.getName()] = m; }
  getMetals() { return this.metals; }
  loadMetal(name) { this.metalObj = this.metals[name]; return this.metalObj; }
  getLoadedMetal() { return this.metalObj; }
  getMetal() { return this.metalObj; }
  loadSurface() { return this.surfaceObj; }
  getSurface() { return this.surfaceObj; }
  loadVolume() { return this.volumeObj; }
  getVolume() { return this.volumeObj; }
  syncShader(s) {
    if (this.metalObj) this.metalObj.syncShader(s);
    if (this.dielectricObj) this.dielectricObj.syncShader(s);
    if (this.surfaceObj) this.surfaceObj.syncShader(s);
    if (this.volumeObj) this.volumeObj.syncShader(s);
  }
}

// ------------------------------------------------------------------ renderer (js/renderer.js)
let SNELLY = null;
function snellyGlobal() { return SNELLY; }

class Renderer {
  constructor(snelly) {
    this.snelly = snelly;
    this._width = 64; this._height = 64;
    this.renderMode = 'pt'; this.dispersive = false; this.maxSamplesPerFrame = 1; this.maxSpp = 1024;
    this.maxBounces = 3; this.maxAtmosphereScatters = 1; this.maxMarchSteps = 256; this.maxSSSSteps = 128;
    this.radianceClamp = 3.0; this.wavelengthSamples = 256; this.filterRadius = 2.0; this.goalFPS = 20.0;
    this.minsSPPToRedraw = 0.0; this.shadowStrength = 1.0; this.maxStepsIsMiss = true; this.interactive = true;
    this.skyPower = 0.0; this.skyTemperature = 6000.0; this.skyTintUp = [1, 1, 1]; this.skyTintDown = [1, 1, 1];
    this.envMapVisible = true; this.envMapPhiRotation = 0.0; this.envMapThetaRotation = 0.0;
    this.envMapTransitionAngle = 135.0;
    this.sunPower = -7.0; this.sunAngularSize = 5.0; this.sunColor = [1, 1, 1]; this.sunLatitude = 50.0;
    this.sunLongitudef = 0.0;  // (sic: js/renderer.js:113 leaves sunLongitude undefined)
    this.sunVisibleDirectly = true;
    this.sphereLightRadius = 1.0; this.sphereLightPower = -7.0; this.sphereLightColor = [1, 1, 1];
    this.sphereLightPosition = [0, 0, 0];
    this.exposure = 0.0; this.gamma = 2.2; this.contrast = 1.0; this.saturation = 1.0; this.hueShift = 0.0;
    this.numSamples = 0; this.spp = 0; this.numFramesSinceReset = 0; this.skipProbability = 0.0;
    this.frametime_measure_ms = 0.0;
    this.program = null; this.state = null; this.buildFlags = 0; this.engine = 1;
    this.envMap = null;
  }
  // the scene's env map (js/renderer.js:140-159): a local PNG file, resolved
  // against the scene file's folder; remote URLs are not fetched (the map then
  // stays unloaded, haveEnvMap 0, as before a browser's load completes)
  loadEnvMap(baseDir) {
    const sceneObj = this.snelly.getScene();
    if (typeof sceneObj.envMap === 'undefined') return;
    const url = sceneObj.envMap();
    if (typeof url !== 'string' || url === '' || /^[a-z]+:\/\//i.test(url)) return;
    const file = path.isAbsolute(url) ? url : path.join(baseDir || '.', url);
    if (!fs.existsSync(file)) return;
    try {
      const img = loadPng(file);
      this.envMap = new Texture(this.snelly.ctx, img.width, img.height, 3, false, true, true, srgbToLinear(img.rgb));
    } catch (e) {
      console.warn('[snelly] env map not loaded: ' + e.message);
    }
  }
  // Renderer.compileShaders (js/renderer.js:217-316)
  assembleShaders() {
    const sceneObj = this.snelly.getScene();
    if (typeof sceneObj.shader === 'undefined') GLU.fail('Scene must define a "shader" function!');
    let shader = sceneObj.shader();
    const has = (n) => shader.indexOf(n + '(') !== -1;
    if (!has('INIT')) shader += '\n void INIT() {}\n';
    const hasCustomCamera = has('CONSTRUCT_PRIMARY_RAY');
    const hasSurface = has('SDF_SURFACE');
    if (!has('SURFACE_DIFFUSE_REFLECTANCE')) shader += '\n vec3 SURFACE_DIFFUSE_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V) { return C; }\n';
    if (!has('SURFACE_SPECULAR_REFLECTANCE')) shader += '\n vec3 SURFACE_SPECULAR_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V) { return C; }\n';
    if (!has('SURFACE_ROUGHNESS')) shader += '\n float SURFACE_ROUGHNESS(in float roughness, in vec3 X, in vec3 N) { return roughness; }\n';
    const hasSurfaceNM = has('SURFACE_NORMAL_MAP');
    if (!has('SUBSURFACE_ALBEDO')) shader += '\n vec3 SUBSURFACE_ALBEDO(in vec3 C, in vec3 X, in vec3 N) { return C; }\n';
    const hasMetal = has('SDF_METAL');
    if (!has('METAL_SPECULAR_REFLECTANCE')) shader += '\n vec3 METAL_SPECULAR_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V) { return C; }\n';
    if (!has('METAL_ROUGHNESS')) shader += '\n float METAL_ROUGHNESS(in float roughness, in vec3 X, in vec3 N) { return roughness; }\n';
    const hasMetalNM = has('METAL_NORMAL_MAP');
    const hasDielectric = has('SDF_DIELECTRIC');
    if (!has('DIELECTRIC_SPECULAR_REFLECTANCE')) shader += '\n vec3 DIELECTRIC_SPECULAR_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V) { return C; }\n';
    if (!has('DIELECTRIC_ROUGHNESS')) shader += '\n float DIELECTRIC_ROUGHNESS(in float roughness, in vec3 X, in vec3 N) { return roughness; }\n';
    const hasDielectricNM = has('DIELECTRIC_NORMAL_MAP');
    if (!has('DIELECTRIC_ABSORPTION')) shader += '\n vec3 DIELECTRIC_ABSORPTION(vec3 absorption_ui, vec3 X) { return absorption_ui; }\n';
    let hasVolumeEmission = true;
    if (!has('VOLUME_EXTINCTION')) shader += '\n float VOLUME_EXTINCTION(float extinction_ui, vec3 X) { return extinction_ui; }\n';
    if (!has('VOLUME_EXTINCTION_MAX')) shader += '\n float VOLUME_EXTINCTION_MAX(float extinction_ui) { return extinction_ui; }\n';
    if (!has('VOLUME_SCATTERING_COLOR')) shader += '\n vec3 VOLUME_SCATTERING_COLOR(vec3 scattering_color_ui, vec3 X) { return scattering_color_ui; }\n';
    if (!has('VOLUME_ABSORPTION_COLOR')) shader += '\n vec3 VOLUME_ABSORPTION_COLOR(vec3 absorption_color_ui, vec3 X) { return absorption_color_ui; }\n';
    if (!has('VOLUME_ANISOTROPY')) shader += '\n float VOLUME_ANISOTROPY(float anisotropy, vec3 X) { return anisotropy; }\n';
    if (!has('VOLUME_EMISSION')) { hasVolumeEmission = false; shader += '\n vec3 VOLUME_EMISSION(vec3 emission, vec3 X) { return emission; }\n'; }
    const hasSphereLight = this.sphereLightPower > -7.0;
    const hasGeometry = hasSurface || hasMetal || hasDielectric;
    // the reference tests `hasGeometry || hasVolume` without declaring
    // hasVolume (js/renderer.js:267-271); here a scene defining SDF_VOLUME has
    // the heterogeneous medium (HAS_VOLUME) in place of the homogeneous
    // atmosphere, described by the Volume material's colours and extinction
    const hasVolume = has('SDF_VOLUME');
    if (!(hasGeometry || hasVolume)) GLU.fail('Scene must define at least one of: SDF_SURFACE, SDF_METAL, SDF_DIELECTRIC, or SDF_VOLUME');
    const vol = this.snelly.getVolume();
    const nz = (c) => c[0] > 0 || c[1] > 0 || c[2] > 0;
    const hasAtmosphere = !hasVolume && (nz(vol.scatteringColor) || nz(vol.absorptionColor));
    const hasNM = hasSurfaceNM || hasMetalNM || hasDielectricNM;
    const d = [[hasCustomCamera, 'HAS_CUSTOM_CAMERA'], [hasSurface, 'HAS_SURFACE'], [hasMetal, 'HAS_METAL'],
      [hasDielectric, 'HAS_DIELECTRIC'], [hasAtmosphere, 'HAS_ATMOSPHERE'], [hasVolume, 'HAS_VOLUME'],
      [vol.fogEnable, 'HAS_FOG'],
      [hasVolumeEmission, 'HAS_VOLUME_EMISSION'], [hasGeometry, 'HAS_GEOMETRY'], [hasNM, 'HAS_NORMALMAP'],
      [hasSurfaceNM, 'HAS_SURFACE_NORMALMAP'], [hasMetalNM, 'HAS_METAL_NORMALMAP'],
      [hasDielectricNM, 'HAS_DIELECTRIC_NORMALMAP'], [this.interactive, 'INTERACTIVE_MODE'],
      [this.dispersive, 'DISPERSION_ENABLED'], [hasSphereLight, 'HAS_SPHERE_LIGHT']];
    const defines = d.filter(x => x[0]).map(x => `\n#define ${x[1]}\n`).join('');
    const mats = this.snelly.getMaterials();
    const ior = mats.getLoadedDielectric().ior() + '\n' + mats.getLoadedMetal().ior();
    return { shader, ior, defines };
  }
  // renderMode 'ao' / 'ptsimple' select those integrators (js/renderer.js:334-352)
  modeFlags() { return { ao: 4, ptsimple: 8 }[this.renderMode] || 0; }
  compileShaders(compileOnly = false) {
    const src = this.assembleShaders();
    const args = [src.shader, src.ior, src.defines, Math.round(this.maxMarchSteps), Math.round(this.maxBounces),
                  Math.round(this.maxAtmosphereScatters), Math.round(this.maxSamplesPerFrame),
                  this.buildFlags | this.modeFlags()];
    if (compileOnly) { addon.programBuild(null, ...args); return; }
    if (this.program) addon.programDestroy(this.program);
    this.program = addon.programBuild(this.snelly.ctx, ...args);
    addon.setEngine(this.program, this.engine);
    this.shaderObj = new Shader(this.program);
  }
  reset(no_recompile = false) {
    this.numSamples = 0; this.spp = 0; this.numFramesSinceReset = 0;
    if (!no_recompile) this.compileShaders();
    if (this.state) addon.stateClear(this.state);
  }
  resize(width, height) {
    this._width = width; this._height = height;
    if (this.state) addon.stateDestroy(this.state);
    this.state = addon.stateCreate(this.snelly.ctx, width, height);
    addon.stateSeed(this.state, this.snelly.rngSeed >>> 0);
    this.reset(true);
  }
  updateSunDir() {
    const lat = (90.0 - this.sunLatitude) * Math.PI / 180.0;
    const lon = this.sunLongitude * Math.PI / 180.0;
    this.sunDir = [Math.sin(lat) * Math.cos(lon), Math.cos(lat), Math.sin(lat) * Math.sin(lon)];
  }
  // the uniform upload of Renderer.render (js/renderer.js:517-611)
  syncUniforms(P) {
    const sn = this.snelly;
    const cam = sn.getCamera();
    const b = cam.basis();
    P.uniform3Fv('camPos', b.pos.toArray());
    P.uniform3Fv('camDir', b.dir.toArray());
    P.uniform3Fv('camX', b.camX.toArray());
    P.uniform3Fv('camY', b.camY.toArray());
    P.uniformF('camFovy', cam.fov);
    P.uniformF('camAspect', cam.aspect);
    P.uniformF('camAperture', sn.lengthScale * Math.pow(1.4142135, cam.aperture));
    P.uniformF('camFocalDistance', sn.lengthScale * Math.pow(10.0, cam.focalDistance));
    P.uniform2Fv('resolution', [this._width, this._height]);
    P.uniformTexture('WavelengthToXYZ', sn.wavelengthToXYZ);
    P.uniformTexture('ICDF', sn.emissionIcdf);
    P.uniformF('filterRadius', this.filterRadius);
    P.uniformF('radianceClamp', Math.pow(10.0, this.radianceClamp));
    P.uniformF('skipProbability', this.skipProbability);
    P.uniformF('lengthScale', Math.max(sn.lengthScale, 1.0e-6));
    P.uniformF('maxLengthScale', Math.max(sn.maxLengthScale, 1.0e-6));
    P.uniformF('minLengthScale', Math.max(sn.minLengthScale, 1.0e-6));
    P.uniformF('shadowStrength', this.shadowStrength);
    P.uniformI('maxSSSSteps', this.maxSSSSteps);
    P.uniformI('maxStepsIsMiss', Boolean(this.maxStepsIsMiss) ? 1 : 0);
    P.uniformI('wavelengthSamples', this.wavelengthSamples);
    P.uniformF('skyPower', Math.pow(10.0, this.skyPower));
    P.uniform3Fv('skyTintUp', this.skyTintUp);
    P.uniform3Fv('skyTintDown', this.skyTintDown);
    P.uniformI('envMapVisible', Boolean(this.envMapVisible) ? 1 : 0);
    P.uniformF('envMapPhiRotation', Math.min(Math.max(this.envMapPhiRotation, 0.0), 360.0));
    P.uniformF('envMapThetaRotation', Math.min(Math.max(this.envMapThetaRotation, 0.0), 180.0));
    P.uniformF('envMapTransitionAngle', Math.min(Math.max(this.envMapTransitionAngle, 0.0), 180.0));
    P.uniformF('sunPower', Math.pow(10.0, this.sunPower));
    P.uniformF('sunAngularSize', this.sunAngularSize);
    P.uniformF('sunLatitude', this.sunLatitude);
    P.uniformF('sunLongitude', this.sunLongitude);
    P.uniform3Fv('sunColor', this.sunColor);
    this.updateSunDir();
    P.uniform3Fv('sunDir', this.sunDir);
    P.uniformI('sunVisibleDirectly', this.sunVisibleDirectly ? 1 : 0);
    P.uniform3Fv('sphereLightPosition', this.sphereLightPosition.map(x => x * sn.lengthScale));
    P.uniformF('sphereLightRadius', this.sphereLightRadius * sn.lengthScale);
    P.uniformF('sphereLightPower', Math.pow(10.0, this.sphereLightPower) / (4.0 * Math.PI * Math.pow(this.sphereLightRadius, 2.0)));
    P.uniform3Fv('sphereLightColor', this.sphereLightColor);
    P.uniformI('haveEnvMap', this.envMap ? 1 : 0);
    if (this.envMap) P.uniformTexture('envMap', this.envMap);
    P.uniformF('exposure', this.exposure);
    P.uniformF('invGamma', 1.0 / this.gamma);
    P.uniformF('contrast', this.contrast);
    P.uniformF('saturation', this.saturation);
    P.uniformF('hueShift', this.hueShift);
    const sceneObj = sn.getScene();
    if (typeof sceneObj.syncShader !== 'undefined') sceneObj.syncShader(sn, P);
    sn.materials.syncShader(P);
  }
  // one frame of Renderer.render (js/renderer.js:462-673); `frames` > 1 renders
  // several frames with the uniforms of the first
  render(frames = 1) {
    const sn = this.snelly;
    const sceneObj = sn.getScene();
    if (this.spp > this.maxSpp) return;
    if (typeof sceneObj.preframeCallback !== 'undefined') sceneObj.preframeCallback(sn, null);
    const npx = this._width * this._height;
    this.numSamples += (this.interactive ? (1.0 - this.skipProbability) : 1.0) * Math.round(this.maxSamplesPerFrame) * npx * frames;
    this.spp = this.numSamples / npx;
    this.syncUniforms(this.shaderObj);
    if (this.renderMode === 'normals') addon.renderNormals(this.program, this.state, frames);
    else addon.render(this.program, this.state, frames);
    const st = addon.stats(this.program);
    const ms = st.ms_total / Math.max(1, frames);
    this.frametime_measure_ms = this.numFramesSinceReset === 0 ? ms : 0.5 * ms + 0.5 * this.frametime_measure_ms;
    const goalMs = Math.min(1.0e3, Math.max(1.0, 1.0e3 / this.goalFPS));
    this.skipProbability = Math.max(0.0, 1.0 - goalMs / Math.max(1e-6, this.frametime_measure_ms));
    this.numFramesSinceReset += frames;
    if (typeof sceneObj.postframeCallback !== 'undefined') sceneObj.postframeCallback(sn, null);
  }
  pick(xPick, yPick) {
    this.syncUniforms(this.shaderObj);
    const out = addon.pick(this.program, Float32Array.of(xPick, yPick));
    return { distance: out[0], material: out[1] };
  }
  getSPP() { return this.spp; }
  radiance() { return addon.stateRead(this.state, 0, this._width * this._height); }
  rngState() { return addon.stateRead(this.state, 1, this._width * this._height); }
  tonemap() { this.syncUniforms(this.shaderObj); return addon.tonemap(this.program, this.state, this._width * this._height); }
}

// ------------------------------------------------------------------ Snelly (js/snelly.js)
// The GUI has no counterpart here: every method (addSlider, addColor,
// getSceneFolder, add(...).onChange(...), ...) accepts its arguments and
// returns another no-op, so scene initGui() code runs unchanged.
function noopChain() {
  const target = function () {};
  const proxy = new Proxy(target, {
    get(t, k) {
      if (k === Symbol.toPrimitive) return () => 0;
      if (k === 'then' || k === 'toJSON') return undefined;
      return proxy;
    },
    apply() { return proxy; },
    set() { return true; },
  });
  return proxy;
}
const GUI = function GUI() { return noopChain(); };

class Snelly {
  constructor(sceneObj, opts = {}) {
    SNELLY = this;
    this.sceneObj = sceneObj;
    this.initialized = false; this.terminated = false;
    this.width = opts.width || 256; this.height = opts.height || 256;
    this.rngSeed = opts.seed === undefined ? 1234 : opts.seed;
    // a compile-only Snelly (opts.compileOnly) assembles and compiles the
    // scene's program without a device
    this.ctx = opts.compileOnly ? null : (opts.ctx || addon.ctxCreate(opts.device || 0));
    this.camera = new PerspectiveCamera(45, this.width / this.height, 0.05, 1000);
    this.camera.position.set(1, 1, 1);
    this.camControls = new OrbitControls(this.camera);
    this.gui = new GUI();
    this.gl = GL;
    this.guiVisible = true; this.statusText = '';
    this.materials = new Materials();
    this.pathtracer = new Renderer(this);
    if (opts.buildFlags !== undefined) this.pathtracer.buildFlags = opts.buildFlags;
    if (opts.engine !== undefined) this.pathtracer.engine = opts.engine;
    this.LAMBDA_MIN = 390.0; this.LAMBDA_MAX = 750.0; this.SPECTRUM_SAMPLES = 1024;
    if (this.ctx) {
      this.wavelengthToXYZ = new Texture(this.ctx, 1024, 1, 4, true, true, true, wavelengthToXYZTable());
      this.emissionIcdf = new Texture(this.ctx, 4 * this.SPECTRUM_SAMPLES, 1, 1, true, true, true, null);
    }
    this.spectra = {};
    if (this.ctx) this.pathtracer.loadEnvMap(opts.baseDir);
    this.initScene(opts);
    this.initialized = true;
    if (this.camera.auto_focus && !opts.compileOnly) this.auto_focus();
  }
  getVersion() { return [1, 19, 0]; }
  getRenderer() { return this.pathtracer; }
  getGUI() { return this.gui; }
  getCamera() { return this.camera; }
  getControls() { return this.camControls; }
  getScene() { return this.sceneObj; }
  getGLContext() { return this.gl; }
  getMaterials() { return this.materials; }
  getSurface() { return this.materials.loadSurface(); }
  getVolume() { return this.materials.loadVolume(); }
  getDielectrics() { return this.materials.getDielectrics(); }
  getMetals() { return this.materials.getMetals(); }
  getLoadedDielectric() { return this.materials.getLoadedDielectric(); }
  getLoadedMetal() { return this.materials.getLoadedMetal(); }
  loadDielectric(name) { this.materials.loadDielectric(name); this.reset(); }
  loadMetal(name) { this.materials.loadMetal(name); this.reset(); }
  showGUI(show) { this.guiVisible = show; }
  enableGUI() {}
  setStatus(t) { this.statusText = t; }
  getUserTextureUnitStart() { return 7; }
  getSpectra() { return this.spectra; }
  addSpectrum(s) { this.spectra[s.getName()] = s; }
  loadSpectrum(name) {
    this.spectrumObj = this.spectra[name];
    this.emissionIcdf.copy(this.spectrumObj.inverseCDF(this.LAMBDA_MIN, this.LAMBDA_MAX, this.SPECTRUM_SAMPLES));
  }
  getLoadedSpectrum() { return this.spectrumObj; }
  // Snelly.initScene (js/snelly.js:209-294)
  initScene(opts = {}) {
    const s = this.sceneObj;
    if (typeof s.shader === 'undefined') GLU.fail('Scene must define a "shader" function!');
    this.lengthScale = typeof s.getLengthScale !== 'undefined' ? s.getLengthScale() : 1.0;
    this.minLengthScale = Math.max(1.0e-12, typeof s.getMinLengthScale !== 'undefined' ? s.getMinLengthScale() : 1.0e-4);
    this.maxLengthScale = Math.min(1.0e9, typeof s.getMaxLengthScale !== 'undefined' ? s.getMaxLengthScale() : 1.0e2);
    const po = 2.0 * this.lengthScale;
    this.camera.position.set(po, po, po);
    this.camControls.target.set(0, 0, 0);
    this.camera.aperture = -10.0;
    this.camera.focalDistance = 1.0;
    if (typeof s.init !== 'undefined') s.init(this);
    if (opts.override) opts.override(this);
    if (typeof s.initGui !== 'undefined') s.initGui(this.gui);
    this.materials.getLoadedDielectric().initGui();
    const r = this.pathtracer;
    // the reference's page updates the controls twice before its first frame:
    // Snelly._resize (for a scene that fixes the size, js/snelly.js:274-279,
    // 605) or the constructor's Snelly.resize (js/snelly.js:80, 627), and the
    // end of initScene (js/snelly.js:291)
    this.camControls.update();
    this.camControls.update();
    const w = r.width || this.width, h = r.height || this.height;
    this.width = w; this.height = h;
    this.camera.aspect = w / h;
    if (opts.compileOnly) { r.compileShaders(true); return; }
    for (const m of Object.values(this.materials.getMetals())) if (m.textures && m === this.materials.getLoadedMetal()) m.textures(this.ctx);
    r.compileShaders();
    r.resize(w, h);
    this.spectra = {};
    this.addSpectrum(new BlackbodySpectrum('blackbody', 'Blackbody spectrum', r.skyTemperature));
    this.loadSpectrum('blackbody');
  }
  reset(no_recompile = false) {
    if (!this.initialized || this.terminated) return;
    if (this.camera.auto_focus) this.auto_focus();
    this.pathtracer.reset(no_recompile);
  }
  // Snelly.auto_focus (js/snelly.js:707-722): focus on the first hit under the
  // centre of the image
  auto_focus() {
    const r = this.pathtracer;
    const pickData = r.pick(0.5 * r._width, 0.5 * r._height);
    if (pickData == null) return;
    const fd = Math.max(1.0e-6, Math.abs(pickData.distance));
    const newFocalDistance = Math.log10(fd / this.lengthScale);
    if (newFocalDistance === this.camera.focalDistance) return;
    this.camera.focalDistance = newFocalDistance;
    if (this.gui) this.gui.sync();
  }
  // the tonemapped image as a PNG file (the 'P' key, js/snelly.js:791-806)
  saveImage(file) {
    const r = this.pathtracer;
    r.syncUniforms(r.shaderObj);
    savePng(file, r._width, r._height, addon.tonemapRGBA8(r.program, r.state, r._width * r._height));
  }
  // Snelly.dumpScene (js/snelly.js:296-406): the current settings as scene code
  dumpScene() {
    const camera = this.camera, controls = this.camControls, renderer = this.pathtracer, materials = this.materials;
    let code = `
{
    /******* copy-pasted console output on 'O', begin *******/\n`;
    code += `
    let renderer  = snelly.getRenderer();
    let camera    = snelly.getCamera();
    let controls  = snelly.getControls();
    let materials = snelly.getMaterials();
    `;
    if (typeof this.sceneObj.initGenerator !== 'undefined') code += this.sceneObj.initGenerator();
    code += this.guiVisible ? `\nsnelly.showGUI(true);\n` : `\nsnelly.showGUI(false);\n`;
    const r = renderer;
    code += `
    /** Camera settings **/
    camera.fov = ${camera.fov};
    camera.aperture = ${camera.aperture};
    camera.focalDistance = ${camera.focalDistance};
    camera.auto_focus = ${camera.auto_focus};
    camera.up.set(${camera.up.x}, ${camera.up.y}, ${camera.up.z});
    camera.position.set(${camera.position.x}, ${camera.position.y}, ${camera.position.z});
    controls.target.set(${controls.target.x}, ${controls.target.y}, ${controls.target.z});

    /** Renderer settings **/
    // General rendering settings
    renderer.renderMode = '${r.renderMode}';
    renderer.dispersive = ${r.dispersive};
    renderer.maxSamplesPerFrame = ${r.maxSamplesPerFrame};
    renderer.maxSpp = ${r.maxSpp};
    renderer.maxBounces = ${r.maxBounces};
    renderer.maxAtmosphereScatters = ${r.maxAtmosphereScatters};
    renderer.maxMarchSteps = ${r.maxMarchSteps};
    renderer.maxStepsIsMiss = ${r.maxStepsIsMiss};
    renderer.interactive = ${r.interactive};
    renderer.goalFPS = ${r.goalFPS};
    renderer.minsSPPToRedraw = ${r.minsSPPToRedraw};
    renderer.filterRadius = ${r.filterRadius};
    renderer.radianceClamp = ${r.radianceClamp};
    renderer.wavelengthSamples = ${r.wavelengthSamples};
    renderer.shadowStrength = ${r.shadowStrength};
    // Tone-mapping
    renderer.exposure = ${r.exposure};
    renderer.gamma = ${r.gamma};
    renderer.contrast = ${r.contrast};
    renderer.saturation = ${r.saturation};
    renderer.hueShift = ${r.hueShift};
    // Lights
        // sky light
        renderer.skyPower = ${r.skyPower};
        renderer.skyTintUp = ${v3s(r.skyTintUp)};
        renderer.skyTintDown = ${v3s(r.skyTintDown)};
        renderer.envMapVisible = ${r.envMapVisible};
        renderer.envMapPhiRotation = ${r.envMapPhiRotation};
        renderer.envMapThetaRotation = ${r.envMapThetaRotation};
        renderer.envMapTransitionAngle = ${r.envMapTransitionAngle};
        // sun light
        renderer.sunPower = ${r.sunPower};
        renderer.sunColor = ${v3s(r.sunColor)};
        renderer.sunAngularSize = ${r.sunAngularSize};
        renderer.sunLatitude = ${r.sunLatitude};
        renderer.sunLongitude = ${r.sunLongitude};
        renderer.sunVisibleDirectly = ${r.sunVisibleDirectly};
        // sphere light
        renderer.sphereLightPosition = ${v3s(r.sphereLightPosition)};
        renderer.sphereLightRadius = ${r.sphereLightRadius};
        renderer.sphereLightPower = ${r.sphereLightPower};
        renderer.sphereLightColor = ${v3s(r.sphereLightColor)};
`;
    code += `
    /** Material settings **/`;
    const shader = this.sceneObj.shader();
    if (shader.indexOf('SDF_SURFACE(') !== -1) {
      code += `
    let surface = materials.loadSurface();`;
      code += materials.loadSurface().repr();
    }
    if (shader.indexOf('SDF_METAL(') !== -1) {
      code += `
    let metal = materials.loadMetal('${materials.getLoadedMetal().getName()}');`;
      code += materials.loadMetal(materials.getLoadedMetal().getName()).repr();
    }
    if (shader.indexOf('SDF_DIELECTRIC(') !== -1) {
      code += `
    let dielectric = materials.loadDielectric('${materials.getLoadedDielectric().getName()}');`;
      code += materials.loadDielectric(materials.getLoadedDielectric().getName()).repr();
    }
    code += `
    let volume = materials.loadVolume();`;
    code += materials.loadVolume().repr();
    code += `
    /******* copy-pasted console output on 'O', end *******/
}`;
    return code;
  }
  render(frames = 1) {
    if (!this.initialized || this.terminated) return;
    let ready = true;
    if (typeof this.sceneObj.isReady !== 'undefined') ready = this.sceneObj.isReady(this);
    if (ready) this.pathtracer.render(frames);
  }
  destroy() {
    const r = this.pathtracer;
    if (r.state) addon.stateDestroy(r.state);
    if (r.program) addon.programDestroy(r.program);
    r.state = null; r.program = null;
  }
}

// ------------------------------------------------------------------ scene scripts
// The globals a scene script may reference: the API classes, THREE, GLU, a
// no-op `window`/`document`, and the `snelly` instance once constructed.
function sceneGlobals() {
  const headElement = { appendChild() {} };
  const g = {
    gl: GL, Image: OfflineImage, innerWidth: 256, innerHeight: 256, addEventListener() {},
    requestAnimationFrame() {}, setTimeout, clearTimeout,
    Snelly, Renderer, Materials, Material, Surface, Volume, Metal, TabulatedMetal, Dielectric,
    AbbeDielectric, SellmeierDielectric, Sellmeier2Dielectric, PolyanskiyDielectric, CauchyDielectric,
    Spectrum, BlackbodySpectrum, FlatSpectrum, MonochromaticSpectrum, GUI, GLU, THREE, console, Math,
    Float32Array, Uint8Array, Int32Array, Array, Object, JSON, parseFloat, parseInt, isNaN,
    performance: { now: () => Number(process.hrtime.bigint()) / 1e6 },
    document: { getElementById: () => ({ width: 256, height: 256, style: {}, getContext: () => null, addEventListener() {} }),
                createElement: () => ({ getContext: () => null, style: {}, appendChild() {} }),
                getElementsByTagName: () => [headElement], body: headElement },
  };
  return g;
}

// the scene definition of an HTML scene file: the inline <script> blocks that
// define Scene (those without src=, except the onLoad/animateLoop driver)
function extractSceneScript(html) {
  const blocks = [];
  const re = /<script(?![^>]*\bsrc=)[^>]*>([\s\S]*?)<\/script>/gi;
  let m;
  while ((m = re.exec(html)) !== null) {
    const code = m[1];
    if (/function\s+onLoad\s*\(/.test(code) && !/Scene\.prototype/.test(code)) continue;
    blocks.push(code);
  }
  return blocks.join('\n');
}

function loadScene(file) {
  const vm = require('vm');
  const html = fs.readFileSync(file, 'utf8');
  const code = file.endsWith('.js') ? html : extractSceneScript(html);
  const sandbox = sceneGlobals();
  vm.createContext(sandbox);
  sandbox.window = sandbox;  // browser scripts reach globals through window
  vm.runInContext(code + '\n;this.__Scene = (typeof Scene !== "undefined") ? Scene : null;', sandbox, { filename: file });
  if (!sandbox.__Scene) throw new Error('no Scene defined in ' + file);
  return { Scene: sandbox.__Scene, sandbox };
}

module.exports = {
  loadPng, srgbToLinear, savePng,
  addon, T, BUILD, THREE, GLU, Shader, Texture, Spectrum, BlackbodySpectrum, FlatSpectrum, MonochromaticSpectrum,
  Material, Surface, Volume, Metal, TabulatedMetal, Dielectric, AbbeDielectric, SellmeierDielectric,
  Sellmeier2Dielectric, PolyanskiyDielectric, CauchyDielectric, Materials, Renderer, Snelly, GUI,
  extractSceneScript, loadScene, sceneGlobals, wavelengthToXYZTable, WebGLStandIn, GLTexture,
  // the context the stand-in uploads textures to (set by Snelly; tests may set it)
  setContextForTests(ctx) { SNELLY = { ctx }; },
};
