// record_js_host.js -- runs scenes through snelly_amd/js/snelly_host.js over
// recording_addon.js and prints, per scene, the program build arguments, the
// uniforms of the first render() and the texels of the textures it binds.
//
//   node record_js_host.js CASES.json OUT.json
// CASES.json: [{name, file | scene_js, width, height}]
'use strict';

const fs = require('fs');
const os = require('os');
const path = require('path');

process.env.SNL_NAPI_ADDON = path.join(__dirname, 'recording_addon.js');
const host = require('../../snelly_amd/js/snelly_host.js');
const addon = require(process.env.SNL_NAPI_ADDON);

const [CASES, OUT] = process.argv.slice(2);
const cases = JSON.parse(fs.readFileSync(CASES, 'utf8'));
const tmp = fs.mkdtempSync(path.join(os.tmpdir(), 'snl-rec-'));
const out = [];
for (const c of cases) {
  try {
    let file = c.file;
    if (!file) { file = path.join(tmp, 'scene.js'); fs.writeFileSync(file, c.scene_js); }
    const { Scene } = host.loadScene(file);
    addon._log.builds.length = 0;
    const sn = new host.Snelly(new Scene(), { width: c.width, height: c.height, seed: 1,
                                              baseDir: path.dirname(file) });
    addon._record(true);
    sn.getRenderer().render(1);
    addon._record(false);
    const tex = {};
    for (const [name, type, v] of addon._log.sets)
      if (type === 'texture' && v >= 0) tex[name] = addon._log.textures[v];
    const r = sn.getRenderer();
    out.push({ name: c.name, width: r._width, height: r._height, build: addon._log.builds[addon._log.builds.length - 1],
               sets: addon._log.sets, textures: tex });
  } catch (e) {
    out.push({ name: c.name, error: String(e && e.message ? e.message : e) });
  }
}
fs.rmdirSync(tmp, { recursive: true });
fs.writeFileSync(OUT, JSON.stringify(out));
