// Writes tests/golden/rng_js.json: the first values of the two seeded
// generators the reference uses, evaluated by a JavaScript engine, so the
// oracle's C restatement is checked against JS number semantics (int32
// coercion in `<<`/`>>`, uint32 in `>>>`, Math.imul) rather than against a
// second reading of the same formula.
//   xorshift32 (src/sht/reference.ts:122-136) as stored in a Float32Array
//   mulberry32 (src/mgpu/noise.ts:15-24)
// Run: node tests/golden/make_rng_golden.js
'use strict';
const fs = require('fs');
const path = require('path');

function xorshiftStream(seed, n) {
  let s = seed >>> 0;
  const out = new Float32Array(n);
  for (let k = 0; k < n; k++) {
    s ^= s << 13; s >>>= 0;
    s ^= s >> 17;
    s ^= s << 5; s >>>= 0;
    out[k] = (s / 4294967296) * 2 - 1;
  }
  return Array.from(out);
}

function mulberryStream(seed, n) {
  let s = seed >>> 0;
  const out = [];
  for (let k = 0; k < n; k++) {
    s = (s + 0x6d2b79f5) >>> 0;
    let t = s;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    out.push(((t ^ (t >>> 14)) >>> 0) / 4294967296);
  }
  return out;
}

const golden = {
  engine: `node ${process.version}`,
  xorshift_f32: {},
  mulberry32: {},
};
for (const seed of [42, 12345, 1, 0x9e3779b9]) golden.xorshift_f32[seed] = xorshiftStream(seed, 512);
for (const seed of [0, 1, 42, 0xdeadbeef]) golden.mulberry32[seed] = mulberryStream(seed, 512);
fs.writeFileSync(path.join(__dirname, 'rng_js.json'), JSON.stringify(golden));
console.log('wrote rng_js.json with', golden.engine);
