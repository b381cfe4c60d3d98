'use strict'
// Renders a packed scene through the Node mirror (caustic-tracer_amd/node) and
// writes the linear image.  argv: <dir with packed .bin files> <w> <h> <spp> <seed>
const fs = require('fs')
const path = require('path')
const ct = require(path.join(__dirname, '..', '..', 'caustic-tracer_amd', 'node', 'index.js'))
const dir = process.argv[2]
const [w, h, spp, seed] = process.argv.slice(3, 7).map(Number)
const rd = (name, T) => { const b = fs.readFileSync(path.join(dir, name)); return new T(b.buffer, b.byteOffset, b.byteLength / T.BYTES_PER_ELEMENT) }
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'))
const ps = {
  prims: rd('prims.bin', Float32Array), primCount: meta.primCount,
  nodes: rd('nodes.bin', Float32Array), nodeCount: meta.nodeCount,
  mats: rd('mats.bin', Float32Array), matCount: meta.matCount,
  lights: rd('lights.bin', Int32Array), primObject: rd('primObject.bin', Uint32Array),
  sky: { kind: 'none' },
}
const frame = rd('frame.bin', Float64Array)
const img = ct.renderImage(ps, frame, { width: w, height: h, spp, seed, maxBounces: 8, russianRoulette: true, nee: true })
fs.writeFileSync(path.join(dir, 'out.bin'), Buffer.from(img.buffer))
// the same image through the wavefront launches (the default runs this small scene per path)
ct.setPathKernel(1)
const imgW = ct.renderImage(ps, frame, { width: w, height: h, spp, seed, maxBounces: 8, russianRoulette: true, nee: true })
ct.setPathKernel(0)
const sameWave = Buffer.compare(Buffer.from(img.buffer), Buffer.from(imgW.buffer)) === 0
// single-ray API through the same addon
const hit = { t: Infinity, prim: -1, u: 0, v: 0 }
const r = ct.intersectClosest(ps, 0, 0.5, 1.6, 0, 0, -1, 1e-4, Infinity, hit)
const blocked = ct.occluded(ps, 0, 0.5, 1.6, 0, 0, -1, 1e-4, 100)
const rgba = ct.tonemapBuffer(img, w, h, 1, 0, 'aces')
fs.writeFileSync(path.join(dir, 'rgba.bin'), Buffer.from(rgba.buffer))
const picked = ct.pick(ps, frame, 0.74, 0.15)
fs.writeFileSync(path.join(dir, 'hit.json'), JSON.stringify({ r, hit, blocked, picked, sameWave }))
