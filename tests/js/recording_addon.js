// recording_addon.js -- stands in for snelly_napi.node (SNL_NAPI_ADDON) in CPU
// tests: no device work; program builds, uniform uploads and texture texels
// are kept in `_log` for the test to read.
'use strict';

const log = { builds: [], sets: [], textures: {}, recording: false };
let nextTex = 1;
const hex = (a) => (a ? Buffer.from(a.buffer, a.byteOffset, a.byteLength).toString('hex') : null);

module.exports = {
  _log: log,
  _record(on) { log.recording = on; if (on) log.sets = []; },
  version: () => 'recording stand-in',
  ctxCreate: () => ({ ctx: 1 }),
  ctxDestroy() {},
  programBuild(ctx, shader, ior, defines, march, bounces, atm, spf, flags) {
    log.builds.push({ shader, ior, defines, march, bounces, atm, spf, flags });
    return { id: log.builds.length, names: {}, byIdx: [] };
  },
  programDestroy() {},
  setEngine() {},
  stateCreate: () => ({}),
  stateSeed() {}, stateClear() {}, stateDestroy() {},
  uniformIndex(p, name) {
    if (!(name in p.names)) { p.names[name] = p.byIdx.length; p.byIdx.push(name); }
    return p.names[name];
  },
  uniformSet(p, idx, type, arr) { if (log.recording) log.sets.push([p.byIdx[idx], type, Array.from(arr)]); },
  uniformTexture(p, idx, h) { if (log.recording) log.sets.push([p.byIdx[idx], 'texture', h ? h.id : -1]); },
  textureCreate(ctx, w, h, d, ch, lin, clamp, data) {
    const t = { id: nextTex++ };
    log.textures[t.id] = { width: w, height: h, depth: d, channels: ch, linear: lin, clamp, texels: hex(data) };
    return t;
  },
  textureUpload(h, data) { log.textures[h.id].texels = hex(data); },
  textureDestroy() {},
  render() {}, renderNormals() {},
  stats: () => ({ ms_total: 0 }),
  stateRead: (s, which, n) => new Float32Array(n * 4),
  pick: () => new Float32Array(4),
};
