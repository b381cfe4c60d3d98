'use strict'
/**
 * Node mirror of the reference tracer's public API for the accelerated path
 * (src/tracer/index.ts:1-20): the functions keep the reference's names,
 * arguments and return values, and run on the MI355X through the N-API addon.
 * A PackedScene is the object packScene (pack.ts:195) returns; it is uploaded
 * once and cached per object.  Plain JS (no TypeScript-only syntax), so Node 12
 * loads it directly.
 */
const path = require('path')
const addon = require(path.join(__dirname, 'caustic_amd.node'))
const pack = require(path.join(__dirname, 'pack.js'))

const WALK_STACKLESS = 0
const WALK_ORDERED = 1
const T_MIN = 1e-4

const scenes = new WeakMap()
function deviceScene(ps) {
  let h = scenes.get(ps)
  if (h === undefined) {
    if (!('primCount' in ps)) throw new TypeError('caustic_amd: expected a PackedScene (packScene output)')
    h = addon.createScene(ps.prims, ps.primCount, ps.nodes, ps.nodeCount, ps.mats, ps.matCount,
      ps.lights, ps.primObject, ps.sky)
    scenes.set(ps, h)
  }
  return h
}

/** CameraFrame (camera.ts:8-22) -> the 16 doubles of the C ABI. */
function frameArray(f) {
  if (f instanceof Float64Array) return f
  return Float64Array.from([f.origin.x, f.origin.y, f.origin.z, f.right.x, f.right.y, f.right.z,
    f.up.x, f.up.y, f.up.z, f.forward.x, f.forward.y, f.forward.z, f.tanHalfFov, f.aspect,
    f.lensRadius, f.focusDistance])
}

/** buildBvh (bvh.ts:299) on the GPU; same {nodes, nodeCount, order}. */
function buildBvh(pb, bins) {
  return addon.buildBvh(pb.min, pb.max, pb.centroid, bins === undefined ? 12 : bins)
}

/** packScene (pack.ts:195) with the BVH built on the GPU. */
function packScene(scene, bins) {
  return pack.packScene(scene, buildBvh, bins)
}

// renderImage and accumulatePass take a Scene or a PackedScene (tracer.ts:349
// packs a Scene on demand); a packed Scene is kept per Scene object.
const packedScenes = new WeakMap()
function packedOf(sceneOrPacked) {
  if ('primCount' in sceneOrPacked) return sceneOrPacked
  let ps = packedScenes.get(sceneOrPacked)
  if (ps === undefined) {
    ps = packScene(sceneOrPacked)
    packedScenes.set(sceneOrPacked, ps)
  }
  return ps
}

function single(ps, walk, ox, oy, oz, dx, dy, dz, tMin, tMax, hit) {
  const r = addon.closestHit(deviceScene(ps), Float64Array.of(ox, oy, oz, dx, dy, dz), tMin,
    tMax === Infinity ? null : Float64Array.of(tMax), walk)
  hit.prim = r.prim[0]
  hit.t = r.t[0]
  if (hit.prim >= 0) {
    hit.u = r.u[0]
    hit.v = r.v[0]
  }
  return hit.prim >= 0
}

/** intersectClosest (intersect.ts:202). */
function intersectClosest(ps, ox, oy, oz, dx, dy, dz, tMin, tMax, hit) {
  return single(ps, WALK_ORDERED, ox, oy, oz, dx, dy, dz, tMin, tMax, hit)
}

/** intersectClosestStackless (intersect.ts:290). */
function intersectClosestStackless(ps, ox, oy, oz, dx, dy, dz, tMin, tMax, hit) {
  return single(ps, WALK_STACKLESS, ox, oy, oz, dx, dy, dz, tMin, tMax, hit)
}

/** occluded (intersect.ts:368). */
function occluded(ps, ox, oy, oz, dx, dy, dz, tMin, tMax) {
  return addon.occluded(deviceScene(ps), Float64Array.of(ox, oy, oz, dx, dy, dz), tMin,
    Float64Array.of(tMax))[0] === 1
}

/** Batched closest hit: rays = Float64Array(6n); tMax = Float64Array(n) or null. */
function closestHitBatch(ps, rays, tMin, tMax, walk) {
  return addon.closestHit(deviceScene(ps), rays, tMin === undefined ? T_MIN : tMin,
    tMax || null, walk === undefined ? WALK_ORDERED : walk)
}

/** Batched occluded: Uint8Array of 0/1. */
function occludedBatch(ps, rays, tMin, tMax) {
  return addon.occluded(deviceScene(ps), rays, tMin === undefined ? T_MIN : tMin, tMax)
}

/** accumulatePass (tracer.ts:361): adds one sample per pixel into accum. */
function accumulatePass(ps, frame, width, height, frameIndex, opts, accum) {
  addon.accumulatePass(deviceScene(packedOf(ps)), frameArray(frame), width, height, frameIndex,
    opts.maxBounces, opts.russianRoulette ? 1 : 0, opts.nee ? 1 : 0, accum)
}

/** renderImage (tracer.ts:348): linear RGB averaged over opts.spp; scene is a Scene or a PackedScene. */
function renderImage(scene, frame, opts) {
  const seed = opts.seed === undefined ? 0 : opts.seed
  return addon.renderImage(deviceScene(packedOf(scene)), frameArray(frame), opts.width, opts.height, opts.spp,
    seed, opts.maxBounces, opts.russianRoulette ? 1 : 0, opts.nee ? 1 : 0)
}

/** tonemapBuffer (tonemap.ts:41-62) on the GPU: float32 sum -> RGBA8. */
function tonemapBuffer(accum, width, height, samples, exposureEv, mode, out) {
  const rgba = addon.tonemap(accum, width, height, samples, exposureEv,
    mode === 'reinhard' ? 1 : 0)
  if (out) {
    out.set(rgba)
    return out
  }
  return rgba
}

/** pick (picking.ts:21-33): primaryRay (camera.ts:75) in binary64, closest hit on the GPU. */
function pick(ps, frame, sx, sy) {
  const f = frameArray(frame)
  const px = (2 * sx - 1) * f[12] * f[13]
  const py = (2 * sy - 1) * f[12]
  const o = [f[0], f[1], f[2]]
  const d = [0, 1, 2].map(k => o[k] + ((f[3 + k] * px + f[6 + k] * py) + f[9 + k]) * f[15] - o[k])
  const len = Math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
  const inv = len > 0 ? 1 / len : 0
  const dir = d.map(x => x * inv)
  const hit = { t: Infinity, prim: -1, u: 0, v: 0 }
  if (!intersectClosest(ps, o[0], o[1], o[2], dir[0], dir[1], dir[2], T_MIN, Infinity, hit)) return null
  return {
    objectIndex: ps.primObject[hit.prim],
    prim: hit.prim,
    t: hit.t,
    point: { x: o[0] + dir[0] * hit.t, y: o[1] + dir[1] * hit.t, z: o[2] + dir[2] * hit.t },
  }
}

// rng.ts:10-64 -- the per-path random stream the radiance estimates draw from:
// a PCG output permutation turns (pixel, frame) into a seed, xorshift32 steps
// produce the numbers; all arithmetic in uint32.
function pcgHash(input) {
  const lcg = (Math.imul(input >>> 0, 747796405) + 2891336453) >>> 0
  const x = ((lcg >>> ((lcg >>> 28) + 4)) ^ lcg) >>> 0
  const w = Math.imul(x, 277803737) >>> 0
  return ((w >>> 22) ^ w) >>> 0
}

function seedFor(pixelIndex, frame) {
  const h = pcgHash(pcgHash(pixelIndex >>> 0) ^ Math.imul(pcgHash((frame >>> 0) ^ 0x9e3779b9), 0x85ebca6b))
  return h === 0 ? 0x1234567 : h
}

class Rng {
  constructor(seed) {
    const h = pcgHash(seed >>> 0)
    this.state = h === 0 ? 0x9e3779b9 : h
  }
  static fromPixel(pixelIndex, frame) {
    const r = new Rng(0)
    r.state = seedFor(pixelIndex, frame)
    return r
  }
  nextU32() {
    let x = this.state
    x = (x ^ (x << 13)) >>> 0
    x = (x ^ (x >>> 17)) >>> 0
    x = (x ^ (x << 5)) >>> 0
    this.state = x
    return x
  }
  next() { return this.nextU32() / 4294967296 }
  range(min, max) { return min + (max - min) * this.next() }
  int(n) { return Math.min(n - 1, Math.floor(this.next() * n)) }
}

/**
 * radianceInto (tracer.ts:191-328) for many rays on the GPU: rays is a
 * Float64Array of 6 values per ray (origin, unit direction), states a
 * Uint32Array of xorshift32 states (one Rng per ray), advanced in place as the
 * reference advances each Rng.  Returns a Float64Array of 3 values per ray.
 */
function radianceBatch(ps, rays, states, opts) {
  return addon.radiance(deviceScene(packedOf(ps)), rays, states, opts.maxBounces,
    opts.russianRoulette ? 1 : 0, opts.nee ? 1 : 0)
}

/** radianceInto (tracer.ts:191): one ray, drawing from rng (an Rng); out gets linear RGB. */
function radianceInto(ps, ox, oy, oz, dx, dy, dz, rng, opts, out) {
  const st = new Uint32Array([rng.state >>> 0])
  const L = radianceBatch(ps, new Float64Array([ox, oy, oz, dx, dy, dz]), st, opts)
  rng.state = st[0]
  const o = out || new Float64Array(3)
  o[0] = L[0]
  o[1] = L[1]
  o[2] = L[2]
  return o
}

/** Arithmetic of accumulatePass / renderImage: 0 binary64 (reference), 1 FP32 precision form. */
function setPrecision(p) {
  addon.setPrecision(p)
}

/**
 * How renders are launched: 0 auto (one per-path launch for scenes of at most
 * 32768 primitives, the wavefront above), 1 the wavefront, 2 per path.  The
 * images are the same under each.
 */
function setPathKernel(mode) {
  addon.setPathKernel(mode)
}

/**
 * Serving loop over resident scenes: jobs are renderImage calls that run with
 * continuous batching on the library's own thread; wait() returns a Promise
 * (N-API async work) so the event loop keeps accepting requests meanwhile.
 */
class Server {
  constructor(opts) {
    this.handle = addon.createServer(opts || {})
  }
  /** Queues renderImage(ps, frame, opts); returns the job id. */
  submit(ps, frame, opts) {
    return addon.submitJob(this.handle, deviceScene(ps), frameArray(frame), opts.width,
      opts.height, opts.spp, opts.seed === undefined ? 0 : opts.seed, opts.maxBounces,
      opts.russianRoulette ? 1 : 0, opts.nee ? 1 : 0)
  }
  poll(id) { return addon.pollJob(this.handle, id) }
  wait(id) { return addon.waitJob(this.handle, id) }
  read(id, width, height) { return addon.readJob(this.handle, id, width, height) }
  release(id) { addon.releaseJob(this.handle, id) }
  stats() { return addon.serverStats(this.handle) }
}

module.exports = {
  WALK_STACKLESS, WALK_ORDERED, T_MIN, buildBvh, packScene, intersectClosest, intersectClosestStackless,
  occluded, closestHitBatch, occludedBatch, accumulatePass, renderImage, frameArray,
  setPrecision, setPathKernel, Server, tonemapBuffer, pick, pcgHash, seedFor, Rng, radianceInto,
  radianceBatch,
}
