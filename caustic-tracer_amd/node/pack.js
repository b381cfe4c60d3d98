'use strict'
/**
 * packScene (pack.ts:195-223) for the Node mirror: a Scene in the reference's
 * object format (scene.ts:5-74; Vec3 {x, y, z}, Color {r, g, b}) to the
 * PackedScene typed arrays, with the BVH built by the GPU builder through the
 * addon.  The records, normals, areas and PrimBounds are the binary64
 * arithmetic of pack.ts:131-192 and scene.ts:85-127 in the same order; the
 * records are then put in BVH leaf order (pack.ts:195-204).  Plain JS for
 * Node 12.
 */
const PRIM_STRIDE = 16
const MAT_STRIDE = 16
const PRIM_SPHERE = 0
const PRIM_TRIANGLE = 1
const MAT_CODE = { lambertian: 0, metal: 1, dielectric: 2, emissive: 3 }
// from this many primitives the float32 leaf words (bvh.ts:60-62) are inexact;
// the Python packer carries the exact words then, this one refuses
const EXACT_LEAF_LIMIT = 1 << 21

/** materialEmission (scene.ts:176-185). */
function materialEmission(m) {
  if (m.type === 'emissive') return { r: m.color.r * m.strength, g: m.color.g * m.strength, b: m.color.b * m.strength }
  if (m.type === 'lambertian' && m.emission !== undefined && m.emission !== null) {
    const e = m.emission
    if (e.r > 0 || e.g > 0 || e.b > 0) return e
  }
  return null
}

/** writeMaterial (pack.ts:75-119). */
function writeMaterial(out, i, m) {
  const o = i * MAT_STRIDE
  out[o + 3] = MAT_CODE[m.type]
  if (m.type === 'lambertian' || m.type === 'metal') {
    out[o] = m.albedo.r; out[o + 1] = m.albedo.g; out[o + 2] = m.albedo.b
    if (m.type === 'metal') {
      out[o + 4] = m.roughness
    } else if (m.checker) {
      out[o + 5] = m.checker.scale
      out[o + 6] = 1
      out[o + 8] = m.checker.color2.r; out[o + 9] = m.checker.color2.g; out[o + 10] = m.checker.color2.b
    }
  } else if (m.type === 'dielectric') {
    const t = m.tint || { r: 1, g: 1, b: 1 }
    out[o] = t.r; out[o + 1] = t.g; out[o + 2] = t.b
    out[o + 4] = m.ior
  } else {
    out[o] = m.color.r; out[o + 1] = m.color.g; out[o + 2] = m.color.b
    out[o + 4] = m.strength
  }
  const e = materialEmission(m)
  if (e) {
    out[o + 12] = e.r; out[o + 13] = e.g; out[o + 14] = e.b
  }
}

/** Primitives in scene order: [kind, a, b, c, radius, object] (scene.ts:85-118). */
function enumerate(scene) {
  const prims = []
  scene.objects.forEach((ob, oi) => {
    const s = ob.shape
    if (s.kind === 'sphere') {
      prims.push([PRIM_SPHERE, s.center, null, null, s.radius, oi])
    } else if (s.kind === 'triangle') {
      prims.push([PRIM_TRIANGLE, s.a, s.b, s.c, 0, oi])
    } else if (s.kind === 'quad') {
      const o = s.origin
      const p1 = { x: o.x + s.u.x, y: o.y + s.u.y, z: o.z + s.u.z }
      const p2 = { x: p1.x + s.v.x, y: p1.y + s.v.y, z: p1.z + s.v.z }
      const p3 = { x: o.x + s.v.x, y: o.y + s.v.y, z: o.z + s.v.z }
      prims.push([PRIM_TRIANGLE, o, p1, p2, 0, oi], [PRIM_TRIANGLE, o, p2, p3, 0, oi])
    } else if (s.kind === 'mesh') {
      const V = s.vertices
      for (let t = 0; t + 2 < s.indices.length; t += 3)
        prims.push([PRIM_TRIANGLE, V[s.indices[t]], V[s.indices[t + 1]], V[s.indices[t + 2]], 0, oi])
    } else {
      throw new TypeError('packScene: unknown shape kind ' + s.kind)
    }
  })
  return prims
}

/** packScene(scene, bins?) -> PackedScene, BVH built on the GPU (buildBvh). */
function packScene(scene, buildBvh, bins) {
  const list = enumerate(scene)
  const n = list.length
  if (n >= EXACT_LEAF_LIMIT)
    throw new RangeError('packScene: ' + n + ' primitives; from 2^21 the leaf words need the ' +
      'exact-word layout extension (use the Python packer)')
  const raw = new Float32Array(Math.max(1, n) * PRIM_STRIDE)
  const min = new Float64Array(3 * n), max = new Float64Array(3 * n), cen = new Float64Array(3 * n)
  const obj = new Uint32Array(Math.max(1, n))
  for (let i = 0; i < n; i++) {
    const [kind, a, b, c, r, oi] = list[i]
    const o = i * PRIM_STRIDE
    obj[i] = oi
    raw[o] = a.x; raw[o + 1] = a.y; raw[o + 2] = a.z
    raw[o + 3] = kind
    raw[o + 7] = oi
    let x0, y0, z0, x1, y1, z1
    if (kind === PRIM_SPHERE) {
      raw[o + 4] = r
      raw[o + 11] = 4 * Math.PI * r * r
      const rp = r + 1e-6 + Math.abs(r) * 1e-6 // pack.ts:153
      x0 = a.x - rp; y0 = a.y - rp; z0 = a.z - rp
      x1 = a.x + rp; y1 = a.y + rp; z1 = a.z + rp
    } else {
      const e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z
      const e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z
      const cx = e1y * e2z - e1z * e2y // vec3.ts:22
      const cy = e1z * e2x - e1x * e2z
      const cz = e1x * e2y - e1y * e2x
      const ln = Math.sqrt(cx * cx + cy * cy + cz * cz)
      raw[o + 4] = b.x; raw[o + 5] = b.y; raw[o + 6] = b.z
      raw[o + 8] = c.x; raw[o + 9] = c.y; raw[o + 10] = c.z
      raw[o + 11] = 0.5 * Math.sqrt(cx * cx + cy * cy + cz * cz) // scene.ts:124-127
      if (ln > 0) { // normalize (vec3.ts:28-31): scaled by 1/len, else (0, 0, 0)
        const inv = 1 / ln
        raw[o + 12] = cx * inv; raw[o + 13] = cy * inv; raw[o + 14] = cz * inv
      }
      const pad = 1e-6 // pack.ts:177
      x0 = Math.min(a.x, b.x, c.x) - pad; y0 = Math.min(a.y, b.y, c.y) - pad; z0 = Math.min(a.z, b.z, c.z) - pad
      x1 = Math.max(a.x, b.x, c.x) + pad; y1 = Math.max(a.y, b.y, c.y) + pad; z1 = Math.max(a.z, b.z, c.z) + pad
    }
    min[3 * i] = x0; min[3 * i + 1] = y0; min[3 * i + 2] = z0
    max[3 * i] = x1; max[3 * i + 1] = y1; max[3 * i + 2] = z1
    cen[3 * i] = 0.5 * (x0 + x1); cen[3 * i + 1] = 0.5 * (y0 + y1); cen[3 * i + 2] = 0.5 * (z0 + z1)
  }
  const bvh = buildBvh({ min, max, centroid: cen, count: n }, bins)
  const prims = new Float32Array(Math.max(1, n) * PRIM_STRIDE)
  const primObject = new Uint32Array(Math.max(1, n))
  for (let i = 0; i < n; i++) {
    const src = bvh.order[i]
    prims.set(raw.subarray(src * PRIM_STRIDE, (src + 1) * PRIM_STRIDE), i * PRIM_STRIDE)
    primObject[i] = obj[src]
  }
  const objs = scene.objects
  const mats = new Float32Array(Math.max(1, objs.length) * MAT_STRIDE)
  const emits = objs.map((ob, i) => {
    writeMaterial(mats, i, ob.material)
    return materialEmission(ob.material) !== null
  })
  const lights = []
  for (let i = 0; i < n; i++) if (emits[primObject[i]]) lights.push(i)
  return {
    prims, primCount: n, nodes: bvh.nodes, nodeCount: bvh.nodeCount, mats, matCount: objs.length,
    lights: Int32Array.from(lights), primObject, sky: scene.sky || { kind: 'none' }, order: bvh.order,
  }
}

module.exports = { packScene, materialEmission }
