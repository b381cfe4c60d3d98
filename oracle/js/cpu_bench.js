'use strict'
/**
 * The reference's tracer (trace.js) on the host's cores through Node worker
 * threads: renderImage's passes (tracer.ts:348-355, frame index s + 7919*seed)
 * over a band of rows of one frame, every row's pixels on one thread, the
 * band's rows dealt out to the threads in turn.  The protocol of
 * scripts/bench.ts: a warm-up, then timed trials, every trial reported.
 *
 *   node cpu_bench.js <dir>
 *
 * <dir>/job.json names the packed arrays (raw little-endian files in <dir>),
 * the frame, the render options, the band [y0, y1), the thread count and the
 * trial count; the last trial's float32 sums of the band are written to
 * <dir>/accum.f32 (width*height*3 floats, rows outside the band zero).  Prints
 * one JSON line: seconds and rays (extension + NEE) per trial.
 *
 * Test infrastructure: bench.py's CPU-baseline leg and the tests run it.
 */
const fs = require('fs')
const os = require('os')
const path = require('path')
const { Worker, isMainThread, parentPort, workerData } = require('worker_threads')
const T = require('./trace.js')

const TYPES = { f32: Float32Array, i32: Int32Array, u32: Uint32Array }

function sharedArray(file, type) {
  const buf = fs.readFileSync(file)
  const sab = new SharedArrayBuffer(buf.length)
  new Uint8Array(sab).set(buf)
  return { sab, type }
}

function view(a) {
  return a ? new TYPES[a.type](a.sab) : null
}

function passFrames(job) {
  const out = []
  for (let s = 0; s < job.spp; s++) out.push((s + Math.imul(job.seed >>> 0, 7919)) >>> 0)
  return out
}

if (isMainThread) {
  const dir = process.argv[2]
  const job = JSON.parse(fs.readFileSync(path.join(dir, 'job.json'), 'utf8'))
  const arrays = {}
  for (const [name, type] of Object.entries(job.arrays))
    arrays[name] = sharedArray(path.join(dir, name + '.' + type), type)
  const n = job.width * job.height * 3
  const accumSab = new SharedArrayBuffer(n * 4)
  const accum = new Float32Array(accumSab)
  const threads = Math.max(1, job.threads)
  const rows = []
  for (let y = job.rows[0]; y < job.rows[1]; y++) rows.push(y)
  const share = []
  for (let k = 0; k < threads; k++) share.push(rows.filter((_, i) => i % threads === k))
  const workers = []
  for (let k = 0; k < threads; k++)
    workers.push(new Worker(__filename, { workerData: { arrays, accumSab, job } }))

  const runAll = (rowsOf) => Promise.all(workers.map((w, k) => new Promise((resolve, reject) => {
    w.once('message', resolve)
    w.once('error', reject)
    w.postMessage({ rows: rowsOf(k) })
  })))

  ;(async () => {
    // warm-up: every thread renders its first row once
    await runAll((k) => share[k].slice(0, 1))
    const trials = []
    for (let t = 0; t < job.trials; t++) {
      accum.fill(0)
      const t0 = process.hrtime.bigint()
      const res = await runAll((k) => share[k])
      const sec = Number(process.hrtime.bigint() - t0) / 1e9
      let ext = 0, shadow = 0
      for (const r of res) {
        ext += r.extension
        shadow += r.shadow
      }
      trials.push({ seconds: sec, extension: ext, shadow })
    }
    fs.writeFileSync(path.join(dir, 'accum.f32'), Buffer.from(accumSab))
    for (const w of workers) await w.terminate()
    const cpus = os.cpus()
    process.stdout.write(JSON.stringify({
      threads, trials, node: process.version, cpu_model: cpus.length ? cpus[0].model : '',
      cpus: cpus.length,
    }) + '\n')
  })().catch((e) => {
    process.stderr.write(String(e && e.stack ? e.stack : e) + '\n')
    process.exit(1)
  })
} else {
  const { arrays, accumSab, job } = workerData
  const scene = new T.Scene({
    prims: view(arrays.prims), nodes: view(arrays.nodes), nodeCount: job.nodeCount,
    mats: view(arrays.mats), lights: view(arrays.lights), lightCount: job.lightCount,
    primObject: view(arrays.primObject), leafWords: view(arrays.leafWords), sky: job.sky,
  })
  const accum = new Float32Array(accumSab)
  const frames = passFrames(job)
  const opts = { maxBounces: job.maxBounces, russianRoulette: job.russianRoulette, nee: job.nee }
  parentPort.on('message', (msg) => {
    scene.extension = 0
    scene.shadow = 0
    for (const fi of frames)
      T.accumulateRows(scene, job.frame, job.width, job.height, fi, opts, accum, msg.rows)
    parentPort.postMessage({ extension: scene.extension, shadow: scene.shadow })
  })
}
