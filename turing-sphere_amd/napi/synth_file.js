// Drive a synthesis and an analysis through the N-API binding the way the
// reference's ShtPlan would (create, bind device buffers, encode, read back):
//   node synth_file.js <lmax> <nlat> <nphi> <in.f32> <spat_out.f32> <q_out.f32>
// in.f32 holds 2*nlm float32 coefficients; the outputs are float32 files.
'use strict';
const fs = require('fs');
const path = require('path');
const ts = require(path.join(__dirname, 'tsphere_napi.node'));

const [lmax, nlat, nphi] = process.argv.slice(2, 5).map(Number);
const [inFile, spatFile, qFile] = process.argv.slice(5, 8);
const plan = ts.planCreate({ lmax, mmax: lmax, nlat, nphi });
const info = ts.planQuery(plan);
const q = new Float32Array(fs.readFileSync(inFile).buffer.slice(0));
if (q.length !== 2 * info.nlm) throw new Error(`expected ${2 * info.nlm} coefficients`);
const stream = ts.streamCreate();
const dq = ts.malloc(8 * info.nlm);
const ds = ts.malloc(4 * nlat * nphi);
const dq2 = ts.malloc(8 * info.nlm);
ts.upload(dq, q, stream);
ts.synth(plan, dq, ds, stream);
ts.analys(plan, ds, dq2, stream);
const spat = ts.download(ds, nlat * nphi, stream);
const back = ts.download(dq2, 2 * info.nlm, stream);
fs.writeFileSync(spatFile, Buffer.from(spat.buffer));
fs.writeFileSync(qFile, Buffer.from(back.buffer));
for (const p of [dq, ds, dq2]) ts.free(p);
ts.planDestroy(plan);
console.log(JSON.stringify({ nlm: info.nlm, fourierMode: info.fourierMode, legLat: info.legLat }));
