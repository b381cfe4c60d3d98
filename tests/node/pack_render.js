'use strict'
// packScene of a caustic-scene document by the Node mirror, written out as
// raw arrays, and renderImage of it called with the Scene and with the
// PackedScene.  argv: dir (scene.json, frame.bin in; outputs out), w, h, spp, seed.
const fs = require('fs')
const path = require('path')
const ct = require(path.join(__dirname, '..', '..', 'caustic-tracer_amd', 'node', 'index.js'))

const [dir, w, h, spp, seed] = [process.argv[2], +process.argv[3], +process.argv[4], +process.argv[5], +process.argv[6]]
const scene = JSON.parse(fs.readFileSync(path.join(dir, 'scene.json'), 'utf8'))
const fb = fs.readFileSync(path.join(dir, 'frame.bin'))
const frame = new Float64Array(fb.buffer, fb.byteOffset, 16)
const opts = { width: w, height: h, spp, seed, maxBounces: 6, russianRoulette: true, nee: true }
const ps = ct.packScene(scene)
for (const k of ['prims', 'nodes', 'mats', 'lights', 'primObject'])
  fs.writeFileSync(path.join(dir, k + '.bin'), Buffer.from(ps[k].buffer, ps[k].byteOffset, ps[k].byteLength))
fs.writeFileSync(path.join(dir, 'meta.json'), JSON.stringify({ primCount: ps.primCount, nodeCount: ps.nodeCount }))
const a = ct.renderImage(scene, frame, opts)
const b = ct.renderImage(ps, frame, opts)
fs.writeFileSync(path.join(dir, 'img_scene.bin'), Buffer.from(a.buffer))
fs.writeFileSync(path.join(dir, 'img_packed.bin'), Buffer.from(b.buffer))
console.log('ok')
