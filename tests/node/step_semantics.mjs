// NBodyGPU.step() of the JS drop-in is the reference's step (NBodyGPU.js:321-329):
// the `main` dispatch alone, with encodeRefresh (:308-319) its own call.
//  1. encodeRefresh(); step() after sort() equals stepN(1, 1) (rebuild +
//     refresh + walk in one call), bit for bit.
//  2. updateCentralMass() and then step() without a refresh walks the tree's
//     old COMs: the script writes the particles before the step, the node
//     words and the particles after it to the folder in argv[2], and
//     tests/test_gpu_node.py replays that walk on the oracle.
// Run by tests/test_gpu_node.py:  node tests/node/step_semantics.mjs OUTDIR
import fs from 'fs';
import path from 'path';
import { NBodyGPU, requestGPUDevice } from '../../galaxy-sim_amd/node/NBodyGPU.js';

const golden = path.join(path.dirname(new URL(import.meta.url).pathname), '..', 'golden');
const out = process.argv[2];
function load(name) {
  const b = fs.readFileSync(path.join(golden, `${name}.in.f32`));
  return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}
function eqBits(a, b, what) {
  const x = new Uint32Array(a.buffer, a.byteOffset, a.length);
  const y = new Uint32Array(b.buffer, b.byteOffset, b.length);
  if (x.length !== y.length) throw new Error(`${what}: length ${x.length} vs ${y.length}`);
  for (let i = 0; i < x.length; i++) if (x[i] !== y[i]) throw new Error(`${what}: differs at ${i}`);
}
const params = { dt: 0.005, G: 1, epsilon: 0.02, theta: 0.6, gasDrag: 0.25 };

async function main() {
  const device = await requestGPUDevice();
  const data = load('cluster_3k');
  const opts = { exactForce: true };

  const a = new NBodyGPU(device, data, null, opts);
  await a.initPromise;
  a.updateParams(params);
  a.sort();
  a.encodeRefresh(null);
  a.step(null);
  const b = new NBodyGPU(device, data, null, opts);
  await b.initPromise;
  b.updateParams(params);
  b.stepN(1, 1);
  eqBits(a.positionBuffer, b.positionBuffer, 'encodeRefresh + step vs one rebuild step');

  // a second step() without a refresh walks the COMs of the first refresh
  const c = new NBodyGPU(device, data, null, opts);
  await c.initPromise;
  c.updateParams(params);
  c.sort();
  c.encodeRefresh(null);
  c.updateCentralMass(25.0, 7);
  const before = c.positionBuffer;
  const nodes = new Uint8Array(c.downloadNodes());
  c.step(null);
  const after = c.positionBuffer;
  fs.writeFileSync(path.join(out, 'before.f32'), Buffer.from(before.buffer));
  fs.writeFileSync(path.join(out, 'nodes.u32'), Buffer.from(nodes.buffer, nodes.byteOffset, nodes.byteLength));
  fs.writeFileSync(path.join(out, 'after.f32'), Buffer.from(after.buffer));
  a.destroy(); b.destroy(); c.destroy();
}

main().then(() => {
  console.log('node step semantics ok');
}).catch((e) => { console.error(e); process.exit(1); });
