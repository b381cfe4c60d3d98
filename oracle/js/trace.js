'use strict'
/**
 * The reference's path tracer (rng.ts, camera.ts, intersect.ts, materials.ts,
 * tracer.ts) restated in plain JS for Node 12: the algorithm in the
 * reference's own arithmetic -- JS numbers over the float32 packed arrays,
 * the engine's Math.sin/cos/cbrt -- so that it can be timed on the host's
 * cores through worker threads (cpu_bench.js).  Each function cites the
 * reference lines it follows; the C oracle (o64.c) restates the same lines,
 * and tests/test_oracle_js.py checks the two against each other bit for bit.
 *
 * Test infrastructure: only the CPU-baseline leg of bench.py and the tests
 * run it.
 */

const PRIM_STRIDE = 16
const MAT_STRIDE = 16
const NODE_STRIDE = 8
const PRIM_SPHERE = 0
const T_MIN = 1e-4 // tracer.ts:35
const SHADOW_SHORTEN = 1 - 1e-3 // tracer.ts:36
const TWO_PI = 2 * Math.PI // materials.ts:11, tracer.ts:37
const INV_PI = 1 / Math.PI // materials.ts:10
const ORDERED_STACK = 512

// ---------------------------------------------------------------------------
// rng.ts

/** rng.ts:10-15: PCG RXS-M-XS over one LCG step, uint32 arithmetic. */
function pcgHash(input) {
  const state = (Math.imul(input, 747796405) + 2891336453) >>> 0
  const shifted = ((state >>> ((state >>> 28) + 4)) ^ state) >>> 0
  const word = Math.imul(shifted, 277803737) >>> 0
  return ((word >>> 22) ^ word) >>> 0
}

/** rng.ts:18-23 */
function seedFor(pixel, frame) {
  const a = pcgHash(pixel >>> 0)
  const b = pcgHash((frame ^ 0x9e3779b9) >>> 0)
  const s = pcgHash((a ^ Math.imul(b, 0x85ebca6b)) >>> 0)
  return s === 0 ? 0x1234567 : s
}

/** rng.ts:33-64: xorshift32 (13, 17, 5); a draw is u32 / 2^32. */
class Rng {
  constructor(state) {
    this.state = state >>> 0
  }
  next() {
    let x = this.state
    x ^= x << 13
    x ^= x >>> 17
    x ^= x << 5
    this.state = x >>> 0
    return this.state / 4294967296
  }
  int(n) {
    return Math.min(n - 1, Math.floor(this.next() * n))
  }
}

// ---------------------------------------------------------------------------
// intersect.ts

/** intersect.ts:20-39 */
function raySphere(ox, oy, oz, dx, dy, dz, cx, cy, cz, r, tMin, tMax) {
  const lx = ox - cx, ly = oy - cy, lz = oz - cz
  const a = dx * dx + dy * dy + dz * dz
  const halfB = lx * dx + ly * dy + lz * dz
  const c = lx * lx + ly * ly + lz * lz - r * r
  const disc = halfB * halfB - a * c
  if (disc < 0) return -1
  const s = Math.sqrt(disc)
  let t = (-halfB - s) / a
  if (t <= tMin || t >= tMax) {
    t = (-halfB + s) / a
    if (t <= tMin || t >= tMax) return -1
  }
  return t
}

/** intersect.ts:52-85; the hit's t, u, v go to out[0..2]. */
function rayTriangle(ox, oy, oz, dx, dy, dz, ax, ay, az, bx, by, bz, cx, cy, cz, tMin, tMax, out) {
  const e1x = bx - ax, e1y = by - ay, e1z = bz - az
  const e2x = cx - ax, e2y = cy - ay, e2z = cz - az
  const px = dy * e2z - dz * e2y
  const py = dz * e2x - dx * e2z
  const pz = dx * e2y - dy * e2x
  const det = e1x * px + e1y * py + e1z * pz
  if (det > -1e-12 && det < 1e-12) return false
  const invDet = 1 / det
  const tx = ox - ax, ty = oy - ay, tz = oz - az
  const u = (tx * px + ty * py + tz * pz) * invDet
  if (u < 0 || u > 1) return false
  const qx = ty * e1z - tz * e1y
  const qy = tz * e1x - tx * e1z
  const qz = tx * e1y - ty * e1x
  const v = (dx * qx + dy * qy + dz * qz) * invDet
  if (v < 0 || u + v > 1) return false
  const t = (e2x * qx + e2y * qy + e2z * qz) * invDet
  if (t <= tMin || t >= tMax) return false
  out[0] = t
  out[1] = u
  out[2] = v
  return true
}

/** intersect.ts:92-123: entry distance, or Infinity for a miss. */
function rayAabb(ox, oy, oz, idx, idy, idz, nodes, o, tMax) {
  let t0 = (nodes[o] - ox) * idx
  let t1 = (nodes[o + 4] - ox) * idx
  let tNear = t0 < t1 ? t0 : t1
  let tFar = t0 < t1 ? t1 : t0
  t0 = (nodes[o + 1] - oy) * idy
  t1 = (nodes[o + 5] - oy) * idy
  if (t0 < t1) {
    if (t0 > tNear) tNear = t0
    if (t1 < tFar) tFar = t1
  } else {
    if (t1 > tNear) tNear = t1
    if (t0 < tFar) tFar = t0
  }
  t0 = (nodes[o + 2] - oz) * idz
  t1 = (nodes[o + 6] - oz) * idz
  if (t0 < t1) {
    if (t0 > tNear) tNear = t0
    if (t1 < tFar) tFar = t1
  } else {
    if (t1 > tNear) tNear = t1
    if (t0 < tFar) tFar = t0
  }
  if (tNear > tFar || tFar < 0 || tNear > tMax) return Infinity
  return tNear < 0 ? 0 : tNear
}

/** intersect.ts:142 */
function safeInv(d) {
  return 1 / (d === 0 ? 1e-20 : d)
}

/**
 * A packed scene (pack.ts:195-223 arrays) with what the walks keep between
 * calls: the stack, the closest hit and the light-sample scratch.
 */
class Scene {
  constructor(a) {
    this.prims = a.prims
    this.nodes = a.nodes
    this.nodeCount = a.nodeCount
    this.mats = a.mats
    this.lights = a.lights
    this.lightCount = a.lightCount
    this.primObject = a.primObject
    // exact leaf words of the layout extension (from 2^21 primitives), else null
    this.leafWords = a.leafWords || null
    // sky (scene.ts Sky): gradient colours as [r, g, b] or {r, g, b}
    const sky = a.sky || { kind: 'none' }
    const rgb = (c) => (Array.isArray(c) ? c : [c.r, c.g, c.b])
    this.skyGradient = sky.kind === 'gradient'
    this.zenith = this.skyGradient ? rgb(sky.zenith) : [0, 0, 0]
    this.horizon = this.skyGradient ? rgb(sky.horizon) : [0, 0, 0]
    this.skyStrength = this.skyGradient ? sky.strength : 0
    this.stackNode = new Int32Array(ORDERED_STACK)
    this.stackEntry = new Float64Array(ORDERED_STACK)
    this.tri = new Float64Array(3)
    this.hitT = Infinity
    this.hitPrim = -1
    this.hitU = 0
    this.hitV = 0
    this.ls = new Float64Array(8)
    this.dir = new Float64Array(3)
    this.basisOut = new Float64Array(6)
    this.extension = 0
    this.shadow = 0
  }

  leaf(node) {
    if (this.leafWords) {
      const w = this.leafWords[node]
      return w === 0xffffffff ? -1 : w
    }
    return this.nodes[node * NODE_STRIDE + 7]
  }

  /** intersect.ts:155-193 (the leaf test of both walks). */
  testLeaf(leaf, ox, oy, oz, dx, dy, dz, tMin, closest) {
    const prims = this.prims
    const li = leaf | 0
    const count = li & 7
    const start = li >> 3
    const tri = this.tri
    for (let p = start; p < start + count; p++) {
      const r = p * PRIM_STRIDE
      if (prims[r + 3] === PRIM_SPHERE) {
        const t = raySphere(ox, oy, oz, dx, dy, dz, prims[r], prims[r + 1], prims[r + 2],
          prims[r + 4], tMin, closest)
        if (t > 0) {
          closest = t
          this.hitPrim = p
        }
      } else if (rayTriangle(ox, oy, oz, dx, dy, dz, prims[r], prims[r + 1], prims[r + 2],
        prims[r + 4], prims[r + 5], prims[r + 6], prims[r + 8], prims[r + 9], prims[r + 10],
        tMin, closest, tri)) {
        closest = tri[0]
        this.hitPrim = p
        this.hitU = tri[1]
        this.hitV = tri[2]
      }
    }
    return closest
  }

  /** intersect.ts:202-282: the ordered walk; the hit in hitT/hitPrim/hitU/hitV. */
  intersectClosest(ox, oy, oz, dx, dy, dz, tMin, tMax) {
    const nodes = this.nodes
    this.hitPrim = -1
    this.hitU = 0
    this.hitV = 0
    this.hitT = Infinity
    if (this.nodeCount === 0) return false
    const idx = safeInv(dx), idy = safeInv(dy), idz = safeInv(dz)
    let closest = tMax
    if (rayAabb(ox, oy, oz, idx, idy, idz, nodes, 0, closest) === Infinity) return false
    const stackNode = this.stackNode, stackEntry = this.stackEntry
    let sp = 0
    let node = 0
    for (;;) {
      const leaf = this.leaf(node)
      if (leaf >= 0) {
        closest = this.testLeaf(leaf, ox, oy, oz, dx, dy, dz, tMin, closest)
      } else {
        const left = node + 1
        const right = nodes[left * NODE_STRIDE + 3]
        const tl = rayAabb(ox, oy, oz, idx, idy, idz, nodes, left * NODE_STRIDE, closest)
        const tr = rayAabb(ox, oy, oz, idx, idy, idz, nodes, right * NODE_STRIDE, closest)
        if (tl !== Infinity) {
          if (tr !== Infinity) {
            if (sp >= ORDERED_STACK) throw new Error('intersectClosest: stack overflow')
            if (tl <= tr) {
              stackNode[sp] = right
              stackEntry[sp++] = tr
              node = left
            } else {
              stackNode[sp] = left
              stackEntry[sp++] = tl
              node = right
            }
          } else {
            node = left
          }
          continue
        }
        if (tr !== Infinity) {
          node = right
          continue
        }
      }
      do {
        if (sp === 0) {
          if (this.hitPrim < 0) return false
          this.hitT = closest
          return true
        }
        sp--
      } while (stackEntry[sp] >= closest)
      node = stackNode[sp]
    }
  }

  /** intersect.ts:368-420: the stackless any-hit walk. */
  occluded(ox, oy, oz, dx, dy, dz, tMin, tMax) {
    const nodes = this.nodes, prims = this.prims, nodeCount = this.nodeCount
    const idx = safeInv(dx), idy = safeInv(dy), idz = safeInv(dz)
    const tri = this.tri
    let node = 0
    while (node < nodeCount) {
      const entry = rayAabb(ox, oy, oz, idx, idy, idz, nodes, node * NODE_STRIDE, tMax)
      if (entry === Infinity) {
        node = nodes[node * NODE_STRIDE + 3]
        continue
      }
      const leaf = this.leaf(node)
      if (leaf >= 0) {
        const li = leaf | 0
        const count = li & 7
        const start = li >> 3
        for (let p = start; p < start + count; p++) {
          const r = p * PRIM_STRIDE
          if (prims[r + 3] === PRIM_SPHERE) {
            if (raySphere(ox, oy, oz, dx, dy, dz, prims[r], prims[r + 1], prims[r + 2],
              prims[r + 4], tMin, tMax) > 0) return true
          } else if (rayTriangle(ox, oy, oz, dx, dy, dz, prims[r], prims[r + 1], prims[r + 2],
            prims[r + 4], prims[r + 5], prims[r + 6], prims[r + 8], prims[r + 9], prims[r + 10],
            tMin, tMax, tri)) {
            return true
          }
        }
      }
      node++
    }
    return false
  }

  // -------------------------------------------------------------------------
  // tracer.ts

  /** tracer.ts:61-90 */
  lightPdf(px, py, pz, hx, hy, hz, hnx, hny, hnz, prim) {
    const lightCount = this.lightCount
    if (lightCount === 0) return 0
    const prims = this.prims
    const r0 = prim * PRIM_STRIDE
    const tx = hx - px, ty = hy - py, tz = hz - pz
    const dist2 = tx * tx + ty * ty + tz * tz
    if (dist2 <= 0) return 0
    const inv = 1 / Math.sqrt(dist2)
    const wx = tx * inv, wy = ty * inv, wz = tz * inv
    if (prims[r0 + 3] === PRIM_SPHERE) {
      const cx = prims[r0] - px, cy = prims[r0 + 1] - py, cz = prims[r0 + 2] - pz
      const r = prims[r0 + 4]
      const d2 = cx * cx + cy * cy + cz * cz
      if (d2 > r * r) {
        const cosMax = Math.sqrt(Math.max(0, 1 - (r * r) / d2))
        const solid = TWO_PI * (1 - cosMax)
        return solid > 0 ? 1 / (solid * lightCount) : 0
      }
    }
    const cosL = Math.abs(hnx * wx + hny * wy + hnz * wz)
    if (cosL < 1e-6) return 0
    return dist2 / (prims[r0 + 11] * cosL * lightCount)
  }

  /** tracer.ts:93-180; ls = [dir3, dist, pdf, emission3]. */
  sampleLight(px, py, pz, rng) {
    const lightCount = this.lightCount
    if (lightCount === 0) return false
    const prim = this.lights[rng.int(lightCount)]
    const prims = this.prims, ls = this.ls
    const r0 = prim * PRIM_STRIDE
    const m = this.primObject[prim] * MAT_STRIDE
    ls[5] = this.mats[m + 12]
    ls[6] = this.mats[m + 13]
    ls[7] = this.mats[m + 14]
    if (prims[r0 + 3] === PRIM_SPHERE) {
      const cx = prims[r0] - px, cy = prims[r0 + 1] - py, cz = prims[r0 + 2] - pz
      const r = prims[r0 + 4]
      const d2 = cx * cx + cy * cy + cz * cz
      if (d2 > r * r) {
        const cosMax = Math.sqrt(Math.max(0, 1 - (r * r) / d2))
        const cosT = 1 - rng.next() * (1 - cosMax)
        const sinT = Math.sqrt(Math.max(0, 1 - cosT * cosT))
        const phi = TWO_PI * rng.next()
        const invD = 1 / Math.sqrt(d2)
        const ax = cx * invD, ay = cy * invD, az = cz * invD
        const b = basis(ax, ay, az, this.basisOut)
        const sx = sinT * Math.cos(phi), sy = sinT * Math.sin(phi)
        let wx = b[0] * sx + b[3] * sy + ax * cosT
        let wy = b[1] * sx + b[4] * sy + ay * cosT
        let wz = b[2] * sx + b[5] * sy + az * cosT
        const invW = 1 / Math.sqrt(wx * wx + wy * wy + wz * wz)
        wx *= invW
        wy *= invW
        wz *= invW
        const halfB = -(cx * wx + cy * wy + cz * wz)
        const disc = halfB * halfB - (d2 - r * r)
        if (disc < 0) return false
        const dist = -halfB - Math.sqrt(disc)
        if (dist <= 0) return false
        ls[0] = wx
        ls[1] = wy
        ls[2] = wz
        ls[3] = dist
        ls[4] = 1 / (TWO_PI * (1 - cosMax) * lightCount)
        return true
      }
      const u1 = rng.next()
      const u2 = rng.next()
      const nv = unitVector(u1, u2, this.dir)
      const nx = nv[0], ny = nv[1], nz = nv[2]
      const tx = cx + nx * r, ty = cy + ny * r, tz = cz + nz * r
      const dist2 = tx * tx + ty * ty + tz * tz
      if (dist2 <= 0) return false
      const dist = Math.sqrt(dist2)
      const wx = tx / dist, wy = ty / dist, wz = tz / dist
      const cosL = Math.abs(nx * wx + ny * wy + nz * wz)
      if (cosL < 1e-6) return false
      ls[0] = wx
      ls[1] = wy
      ls[2] = wz
      ls[3] = dist
      ls[4] = dist2 / (prims[r0 + 11] * cosL * lightCount)
      return true
    }
    let u1 = rng.next()
    let u2 = rng.next()
    if (u1 + u2 > 1) {
      u1 = 1 - u1
      u2 = 1 - u2
    }
    const ax = prims[r0], ay = prims[r0 + 1], az = prims[r0 + 2]
    const qx = ax + (prims[r0 + 4] - ax) * u1 + (prims[r0 + 8] - ax) * u2
    const qy = ay + (prims[r0 + 5] - ay) * u1 + (prims[r0 + 9] - ay) * u2
    const qz = az + (prims[r0 + 6] - az) * u1 + (prims[r0 + 10] - az) * u2
    const tx = qx - px, ty = qy - py, tz = qz - pz
    const dist2 = tx * tx + ty * ty + tz * tz
    if (dist2 <= 0) return false
    const dist = Math.sqrt(dist2)
    const wx = tx / dist, wy = ty / dist, wz = tz / dist
    const cosL = Math.abs(prims[r0 + 12] * wx + prims[r0 + 13] * wy + prims[r0 + 14] * wz)
    if (cosL < 1e-6) return false
    ls[0] = wx
    ls[1] = wy
    ls[2] = wz
    ls[3] = dist
    ls[4] = dist2 / (prims[r0 + 11] * cosL * lightCount)
    return true
  }

  /** tracer.ts:191-328: radiance of one camera ray into out[0..2]. */
  radiance(ox, oy, oz, dx, dy, dz, rng, maxBounces, russianRoulette, nee, out) {
    let Lr = 0, Lg = 0, Lb = 0
    let br = 1, bg = 1, bb = 1
    let specularBounce = true
    let prevPdf = 0
    let prevX = ox, prevY = oy, prevZ = oz
    const prims = this.prims, mats = this.mats, ls = this.ls, dir = this.dir
    const zen = this.zenith, hor = this.horizon, skyS = this.skyStrength
    const useNee = nee && this.lightCount > 0
    for (let bounce = 0; bounce <= maxBounces; bounce++) {
      this.extension++
      if (!this.intersectClosest(ox, oy, oz, dx, dy, dz, T_MIN, Infinity)) {
        if (this.skyGradient) { // tracer.ts:211-216
          const t = 0.5 * (dy + 1)
          Lr += br * (hor[0] + (zen[0] - hor[0]) * t) * skyS
          Lg += bg * (hor[1] + (zen[1] - hor[1]) * t) * skyS
          Lb += bb * (hor[2] + (zen[2] - hor[2]) * t) * skyS
        }
        break
      }
      const prim = this.hitPrim
      const t = this.hitT
      const r0 = prim * PRIM_STRIDE
      const px = ox + dx * t, py = oy + dy * t, pz = oz + dz * t
      let nx, ny, nz
      if (prims[r0 + 3] === PRIM_SPHERE) {
        const invR = 1 / prims[r0 + 4]
        nx = (px - prims[r0]) * invR
        ny = (py - prims[r0 + 1]) * invR
        nz = (pz - prims[r0 + 2]) * invR
      } else {
        nx = prims[r0 + 12]
        ny = prims[r0 + 13]
        nz = prims[r0 + 14]
      }
      const frontFace = dx * nx + dy * ny + dz * nz < 0
      const sx = frontFace ? nx : -nx
      const sy = frontFace ? ny : -ny
      const sz = frontFace ? nz : -nz
      const m = this.primObject[prim] * MAT_STRIDE
      const type = mats[m + 3]
      const er = mats[m + 12], eg = mats[m + 13], eb = mats[m + 14]
      if (er > 0 || eg > 0 || eb > 0) {
        let w = 1
        if (useNee && !specularBounce) {
          const pl = this.lightPdf(prevX, prevY, prevZ, px, py, pz, nx, ny, nz, prim)
          w = powerHeuristic(prevPdf, pl)
        }
        Lr += br * er * w
        Lg += bg * eg * w
        Lb += bb * eb * w
      }
      if (type === 3) break // emissive
      let ndx, ndy, ndz
      if (type === 0) { // lambertian
        let ar = mats[m], ag = mats[m + 1], ab = mats[m + 2]
        if (mats[m + 6] > 0.5 && !checkerPrimary(px, py, pz, mats[m + 5])) {
          ar = mats[m + 8]
          ag = mats[m + 9]
          ab = mats[m + 10]
        }
        if (useNee && this.sampleLight(px, py, pz, rng)) {
          const wx = ls[0], wy = ls[1], wz = ls[2]
          const pdfL = ls[4]
          const cosS = sx * wx + sy * wy + sz * wz
          if (cosS > 0 && pdfL > 0) {
            this.shadow++
            if (!this.occluded(px, py, pz, wx, wy, wz, T_MIN, ls[3] * SHADOW_SHORTEN)) {
              const pdfB = cosS * INV_PI
              const k = (cosS * INV_PI * powerHeuristic(pdfL, pdfB)) / pdfL
              Lr += br * ar * ls[5] * k
              Lg += bg * ag * ls[6] * k
              Lb += bb * ab * ls[7] * k
            }
          }
        }
        const u1 = rng.next()
        const u2 = rng.next()
        cosineHemisphere(sx, sy, sz, u1, u2, dir, this.basisOut)
        ndx = dir[0]
        ndy = dir[1]
        ndz = dir[2]
        prevPdf = Math.max(0, sx * ndx + sy * ndy + sz * ndz) * INV_PI
        br *= ar
        bg *= ag
        bb *= ab
        specularBounce = false
      } else if (type === 1) { // metal
        const u1 = rng.next()
        const u2 = rng.next()
        const u3 = rng.next()
        metal(dx, dy, dz, sx, sy, sz, mats[m + 4], u1, u2, u3, dir)
        ndx = dir[0]
        ndy = dir[1]
        ndz = dir[2]
        if (ndx * sx + ndy * sy + ndz * sz <= 0) break
        br *= mats[m]
        bg *= mats[m + 1]
        bb *= mats[m + 2]
        specularBounce = true
      } else { // dielectric
        const u = rng.next()
        dielectric(dx, dy, dz, sx, sy, sz, mats[m + 4], frontFace, u, dir)
        ndx = dir[0]
        ndy = dir[1]
        ndz = dir[2]
        br *= mats[m]
        bg *= mats[m + 1]
        bb *= mats[m + 2]
        specularBounce = true
      }
      if (russianRoulette && bounce >= 3) {
        const maxB = br > bg ? (br > bb ? br : bb) : bg > bb ? bg : bb
        const pSurvive = maxB > 1 ? 1 : maxB < 0.05 ? 0.05 : maxB
        if (rng.next() > pSurvive) break
        br /= pSurvive
        bg /= pSurvive
        bb /= pSurvive
      }
      prevX = px
      prevY = py
      prevZ = pz
      ox = px
      oy = py
      oz = pz
      dx = ndx
      dy = ndy
      dz = ndz
    }
    out[0] = Lr
    out[1] = Lg
    out[2] = Lb
  }
}

// ---------------------------------------------------------------------------
// materials.ts

/** materials.ts:20-42: tangent and bitangent of n into out[0..5]. */
function basis(nx, ny, nz, out) {
  let hx = 1, hy = 0
  const hz = 0
  if (Math.abs(nx) > 0.9) {
    hx = 0
    hy = 1
  }
  let tx = hy * nz - hz * ny
  let ty = hz * nx - hx * nz
  let tz = hx * ny - hy * nx
  const inv = 1 / Math.sqrt(tx * tx + ty * ty + tz * tz)
  tx *= inv
  ty *= inv
  tz *= inv
  out[0] = tx
  out[1] = ty
  out[2] = tz
  out[3] = ny * tz - nz * ty
  out[4] = nz * tx - nx * tz
  out[5] = nx * ty - ny * tx
  return out
}

/** materials.ts:47-64 */
function cosineHemisphere(nx, ny, nz, u1, u2, out, scratch) {
  const r = Math.sqrt(u1)
  const phi = TWO_PI * u2
  const x = r * Math.cos(phi)
  const y = r * Math.sin(phi)
  const z = Math.sqrt(u1 < 1 ? 1 - u1 : 0)
  const b = basis(nx, ny, nz, scratch)
  const dx = b[0] * x + b[3] * y + nx * z
  const dy = b[1] * x + b[4] * y + ny * z
  const dz = b[2] * x + b[5] * y + nz * z
  const inv = 1 / Math.sqrt(dx * dx + dy * dy + dz * dz)
  out[0] = dx * inv
  out[1] = dy * inv
  out[2] = dz * inv
}

/** materials.ts:67-74 */
function unitVector(u1, u2, out) {
  const z = 1 - 2 * u1
  const r = Math.sqrt(z * z < 1 ? 1 - z * z : 0)
  const phi = TWO_PI * u2
  out[0] = r * Math.cos(phi)
  out[1] = r * Math.sin(phi)
  out[2] = z
  return out
}

/** materials.ts:77-81 */
function schlick(cosTheta, etaI, etaT) {
  const q = (etaI - etaT) / (etaI + etaT)
  const r0 = q * q
  const c = Math.max(0, Math.min(1, cosTheta))
  const m = 1 - c
  return r0 + (1 - r0) * m * m * m * m * m
}

/** materials.ts:89-95 */
function dielectricReflectance(cosI, etaI, etaT) {
  const eta = etaI / etaT
  const sin2T = eta * eta * (1 - cosI * cosI)
  if (sin2T >= 1) return 1
  if (etaI > etaT) return schlick(Math.sqrt(1 - sin2T), etaI, etaT)
  return schlick(cosI, etaI, etaT)
}

/** materials.ts:102-124 */
function metal(dx, dy, dz, nx, ny, nz, roughness, u1, u2, u3, out) {
  const k = 2 * (dx * nx + dy * ny + dz * nz)
  let rx = dx - k * nx
  let ry = dy - k * ny
  let rz = dz - k * nz
  if (roughness > 0) {
    unitVector(u1, u2, out)
    const rad = Math.cbrt(u3) * roughness
    rx += out[0] * rad
    ry += out[1] * rad
    rz += out[2] * rad
  }
  const inv = 1 / Math.sqrt(rx * rx + ry * ry + rz * rz)
  out[0] = rx * inv
  out[1] = ry * inv
  out[2] = rz * inv
}

/** materials.ts:131-166 */
function dielectric(dx, dy, dz, nx, ny, nz, ior, frontFace, u, out) {
  const etaI = frontFace ? 1 : ior
  const etaT = frontFace ? ior : 1
  const eta = etaI / etaT
  let cosI = -(dx * nx + dy * ny + dz * nz)
  if (cosI > 1) cosI = 1
  const reflectance = dielectricReflectance(cosI, etaI, etaT)
  const sin2T = eta * eta * (1 - cosI * cosI)
  if (u < reflectance || sin2T > 1) {
    const k = 2 * (dx * nx + dy * ny + dz * nz)
    out[0] = dx - k * nx
    out[1] = dy - k * ny
    out[2] = dz - k * nz
    return
  }
  const cosT = Math.sqrt(1 - sin2T)
  const s = eta * cosI - cosT
  const tx = dx * eta + nx * s
  const ty = dy * eta + ny * s
  const tz = dz * eta + nz * s
  const inv = 1 / Math.sqrt(tx * tx + ty * ty + tz * tz)
  out[0] = tx * inv
  out[1] = ty * inv
  out[2] = tz * inv
}

/** materials.ts:169-174 */
function powerHeuristic(a, b) {
  const a2 = a * a
  const b2 = b * b
  const denom = a2 + b2
  return denom > 0 ? a2 / denom : 0
}

/** materials.ts:177-181 */
function checkerPrimary(px, py, pz, s) {
  const sum = Math.floor(px * s + 0.5) + Math.floor(py * s + 0.5) + Math.floor(pz * s + 0.5)
  return ((sum % 2) + 2) % 2 === 0
}

// ---------------------------------------------------------------------------
// camera.ts / tracer.ts:348-395

/** camera.ts:62-72: the ray through (sx, sy) from lens point (lx, ly), into o and d. */
function generateRay(f, sx, sy, lx, ly, o, d) {
  const px = (2 * sx - 1) * f[12] * f[13]
  const py = (2 * sy - 1) * f[12]
  const lr = f[14], fd = f[15]
  for (let k = 0; k < 3; k++) {
    const focus = f[k] + (f[3 + k] * px + f[6 + k] * py + f[9 + k]) * fd
    o[k] = f[k] + (f[3 + k] * (lx * lr) + f[6 + k] * (ly * lr))
    d[k] = focus - o[k]
  }
  const len = Math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
  if (len > 0) {
    const s = 1 / len
    d[0] *= s
    d[1] *= s
    d[2] *= s
  } else {
    d[0] = d[1] = d[2] = 0
  }
}

/**
 * accumulatePass (tracer.ts:361-395) over rows [y0, y1) (or the listed rows):
 * per pixel the camera sample (jitter, lens draws) and its radiance, added to
 * the float32 sums when finite.  Returns nothing; ray counts accumulate on the
 * scene object.
 */
function accumulateRows(scene, frame, width, height, frameIndex, opts, accum, rows) {
  const o = new Float64Array(3), d = new Float64Array(3), c = new Float64Array(3)
  const maxBounces = opts.maxBounces, rr = opts.russianRoulette !== false
  const nee = opts.nee !== false
  const rng = new Rng(1)
  for (let k = 0; k < rows.length; k++) {
    const y = rows[k]
    for (let x = 0; x < width; x++) {
      const pixel = y * width + x
      rng.state = seedFor(pixel, frameIndex)
      const sx = (x + rng.next()) / width
      const sy = 1 - (y + rng.next()) / height
      let lx = 0, ly = 0
      if (frame[14] > 0) {
        const r = Math.sqrt(rng.next())
        const theta = TWO_PI * rng.next()
        lx = r * Math.cos(theta)
        ly = r * Math.sin(theta)
      }
      generateRay(frame, sx, sy, lx, ly, o, d)
      scene.radiance(o[0], o[1], o[2], d[0], d[1], d[2], rng, maxBounces, rr, nee, c)
      if (Number.isFinite(c[0]) && Number.isFinite(c[1]) && Number.isFinite(c[2])) {
        const i = pixel * 3
        accum[i] += c[0]
        accum[i + 1] += c[1]
        accum[i + 2] += c[2]
      }
    }
  }
}

module.exports = { Scene, Rng, pcgHash, seedFor, accumulateRows, generateRay }
