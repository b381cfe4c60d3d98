'use strict'
// Submits three renderImage jobs of one packed scene to the serving loop
// through the Node mirror, awaits them (Promises from N-API async work) and
// writes each image.  argv: <dir with packed .bin files>
const fs = require('fs')
const path = require('path')
const ct = require(path.join(__dirname, '..', '..', 'caustic-tracer_amd', 'node', 'index.js'))
const dir = process.argv[2]
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
const jobs = meta.jobs
const srv = new ct.Server({ maxPaths: 4096, slicePasses: 1 })
const ids = jobs.map(j => srv.submit(ps, frame, j))
let ticks = 0
const timer = setInterval(() => { ticks++ }, 1)
Promise.all(ids.map(id => srv.wait(id))).then(statuses => {
  clearInterval(timer)
  ids.forEach((id, k) => {
    const img = srv.read(id, jobs[k].width, jobs[k].height)
    fs.writeFileSync(path.join(dir, `job${k}.bin`), Buffer.from(img.buffer))
    srv.release(id)
  })
  fs.writeFileSync(path.join(dir, 'serve.json'), JSON.stringify({ statuses, stats: srv.stats(), ticks }))
}).catch(e => { console.error(e); process.exit(1) })
