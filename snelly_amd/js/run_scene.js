#!/usr/bin/env node
// run_scene.js -- run a Snelly scene file (HTML with an inline Scene
// definition, or a .js file defining Scene) through the MI355X renderer.
//
//   node run_scene.js SCENE [--width W] [--height H] [--frames N] [--seed S]
//                     [--compile-only] [--engine wavefront|persistent]
//                     [--precise|--fast] [--out FILE.f32] [--rng-out FILE.f32] [--json]
//
// --compile-only translates and compiles the scene's program (no GPU needed).
// --out writes the accumulated radiance (W*H*4 float32) to FILE, --rng-out the
// per-pixel RNG states.
'use strict';

const fs = require('fs');
const host = require('./snelly_host.js');

function parseArgs(argv) {
  const o = { width: 256, height: 256, frames: 16, seed: 1234, compileOnly: false, engine: null, flags: 0,
              out: null, rngOut: null, json: false, scene: null };
  for (let i = 0; i < argv.length; ++i) {
    const a = argv[i];
    if (a === '--width') o.width = parseInt(argv[++i], 10);
    else if (a === '--height') o.height = parseInt(argv[++i], 10);
    else if (a === '--frames') o.frames = parseInt(argv[++i], 10);
    else if (a === '--seed') o.seed = parseInt(argv[++i], 10);
    else if (a === '--compile-only') o.compileOnly = true;
    else if (a === '--engine') o.engine = argv[++i] === 'wavefront' ? 0 : 1;
    else if (a === '--precise') o.flags = host.BUILD.PRECISE;
    else if (a === '--fast') o.flags = host.BUILD.FAST;
    else if (a === '--out') o.out = argv[++i];
    else if (a === '--rng-out') o.rngOut = argv[++i];
    else if (a === '--json') o.json = true;
    else o.scene = a;
  }
  if (!o.scene) throw new Error('usage: run_scene.js SCENE [options]');
  return o;
}

function main() {
  const o = parseArgs(process.argv.slice(2));
  const { Scene } = host.loadScene(o.scene);
  const t0 = Date.now();
  const opts = { width: o.width, height: o.height, seed: o.seed, buildFlags: o.flags,
                 compileOnly: o.compileOnly, baseDir: require('path').dirname(o.scene) };
  if (o.engine !== null) opts.engine = o.engine;
  const snelly = new host.Snelly(new Scene(), opts);
  const tCompile = Date.now() - t0;
  const res = { scene: o.scene, compile_ms: tCompile };
  if (!o.compileOnly) {
    const r = snelly.getRenderer();
    const t1 = Date.now();
    snelly.render(o.frames);
    const rad = r.radiance();
    res.render_ms = Date.now() - t1;
    res.frames = o.frames;
    res.width = r._width; res.height = r._height;
    res.spp = r.getSPP();
    let sum = 0;
    for (let i = 0; i < rad.length; i += 4) sum += rad[i + 1];
    res.mean_Y = sum / (rad.length / 4);
    if (o.out) fs.writeFileSync(o.out, Buffer.from(rad.buffer));
    if (o.rngOut) fs.writeFileSync(o.rngOut, Buffer.from(r.rngState().buffer));
    snelly.destroy();
  }
  if (o.json) console.log(JSON.stringify(res));
  else console.log(res);
}

main();
