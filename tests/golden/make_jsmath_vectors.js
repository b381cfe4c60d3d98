// Writes jsmath_vectors.json: inputs and outputs of the JS engine's Math.sin,
// Math.cos, Math.cbrt and Math.hypot (three arguments, as bvh.test.ts:97 and
// scripts/bench.ts:86 call it), as 16-hex-digit binary64 bit patterns, on the
// arguments the tracer passes them (2*PI*u and u with u = k / 2^32, the draws
// of rng.ts:52-54) and on a spread of other arguments.  Run with node.
'use strict'
const fs = require('fs')
const path = require('path')

const dv = new DataView(new ArrayBuffer(8))
function hex(x) {
  dv.setFloat64(0, x)
  return dv.getUint32(0).toString(16).padStart(8, '0') + dv.getUint32(4).toString(16).padStart(8, '0')
}
let state = 0x2545f491
function nextU32() {  // xorshift32, any fixed sequence of draws will do
  state ^= state << 13; state >>>= 0
  state ^= state >>> 17
  state ^= state << 5; state >>>= 0
  return state
}
const TWO_PI = 2 * Math.PI
const trig = [], roots = []
for (let i = 0; i < 2500; i++) {
  const u = nextU32() / 4294967296
  trig.push(TWO_PI * u)
  roots.push(u)
}
// other arguments: near multiples of pi/2 (reduction with cancellation), the
// kernels' thresholds, signs, tiny and large values
for (let k = -40; k <= 40; k++) {
  const c = k * Math.PI / 2
  for (const d of [0, 1e-12, -1e-12, 1e-6, -3e-9]) trig.push(c + d)
}
for (const x of [0, -0, 1e-300, -1e-300, 2 ** -28, 0.29, 0.31, 0.78, 0.79, 1, -1, 2.35, 2.36, 3.14, 100, -1e5, 4e5])
  trig.push(x)
for (let i = 0; i < 600; i++) trig.push((nextU32() / 4294967296 - 0.5) * 60)
for (const x of [0, -0, 5e-324, 1e-310, -1e-310, 1e-300, 8, -27, 0.001, 1e300, -1e300, 3, 0.5])
  roots.push(x)
for (let i = 0; i < 600; i++) roots.push((nextU32() / 4294967296 - 0.5) * Math.pow(10, (nextU32() % 600) - 300))

const hyp = []
for (let i = 0; i < 1500; i++) {
  const e = (nextU32() % 40) - 20
  hyp.push([0, 1, 2].map(() => (nextU32() / 4294967296 - 0.5) * Math.pow(2, e + (nextU32() % 5))))
}
hyp.push([0, 0, 0], [-0, 3, 4], [1e300, 1e300, 1e-300], [5e-324, 0, 0], [3, 4, 12], [Infinity, 1, 2], [1, -Infinity, 0])
for (let i = 0; i < 500; i++) {  // the shape of the ray set: targets minus origins
  const o = [0, 1, 2].map(() => -14 + 28 * (nextU32() / 4294967296))
  const t = [0, 1, 2].map(() => -8 + 16 * (nextU32() / 4294967296))
  hyp.push([t[0] - o[0], t[1] - o[1], t[2] - o[2]])
}

const out = {
  engine: 'node ' + process.version + ' (V8 ' + process.versions.v8 + ')',
  sin: trig.map((x) => [hex(x), hex(Math.sin(x))]),
  cos: trig.map((x) => [hex(x), hex(Math.cos(x))]),
  cbrt: roots.map((x) => [hex(x), hex(Math.cbrt(x))]),
  hypot3: hyp.map((v) => [hex(v[0]), hex(v[1]), hex(v[2]), hex(Math.hypot(v[0], v[1], v[2]))]),
}
fs.writeFileSync(path.join(__dirname, 'jsmath_vectors.json'), JSON.stringify(out))
console.log('wrote', out.sin.length, out.cos.length, out.cbrt.length, out.hypot3.length, 'vectors from', out.engine)
