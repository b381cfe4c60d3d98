'use strict'
// radianceBatch / radianceInto of the Node mirror on the packed arrays in
// argv[2]: rays.bin (6 float64 per ray), states.bin (uint32 per ray); writes
// the radiance and the states after the call, and radianceInto of ray 0 with
// Rng.fromPixel(5, 9).
const fs = require('fs')
const path = require('path')
const ct = require(path.join(__dirname, '..', '..', 'caustic-tracer_amd', 'node', 'index.js'))

const dir = process.argv[2]
const rd = (name, T) => { const b = fs.readFileSync(path.join(dir, name)); return new T(b.buffer, b.byteOffset, b.byteLength / T.BYTES_PER_ELEMENT) }
const bytes = a => Buffer.from(a.buffer, a.byteOffset, a.byteLength)
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'))
const ps = {
  prims: rd('prims.bin', Float32Array), primCount: meta.primCount,
  nodes: rd('nodes.bin', Float32Array), nodeCount: meta.nodeCount,
  mats: rd('mats.bin', Float32Array), matCount: meta.matCount,
  lights: rd('lights.bin', Int32Array), primObject: rd('primObject.bin', Uint32Array),
  sky: { kind: 'none' },
}
const rays = rd('rays.bin', Float64Array)
const states = rd('states.bin', Uint32Array)
const L = ct.radianceBatch(ps, rays, states, meta.opts)
fs.writeFileSync(path.join(dir, 'L.bin'), bytes(L))
fs.writeFileSync(path.join(dir, 'states_after.bin'), bytes(states))
const rng = ct.Rng.fromPixel(5, 9)
const out = new Float64Array(3)
const r = ct.radianceInto(ps, rays[0], rays[1], rays[2], rays[3], rays[4], rays[5], rng, meta.opts, out)
fs.writeFileSync(path.join(dir, 'one.json'),
  JSON.stringify({ same: r === out, L: Array.from(out), state: rng.state }))
