// A GalaxyPool slot against the same galaxy stepped alone on an NBodyGPU
// (JS drop-in, N-API addon): bit-identical particles, energies and
// reductions.  Run by tests/test_gpu_node.py:  node tests/node/pool_parity.mjs
import fs from 'fs';
import path from 'path';
import { NBodyGPU, requestGPUDevice } from '../../galaxy-sim_amd/node/NBodyGPU.js';
import { GalaxyPool } from '../../galaxy-sim_amd/node/GalaxyPool.js';

const golden = path.join(path.dirname(new URL(import.meta.url).pathname), '..', 'golden');
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

async function main() {
  const device = await requestGPUDevice();
  const jobs = [
    { name: 'plummer_4k', params: { dt: 0.005, G: 1, epsilon: 0.01, theta: 0.6 }, bucket: 16 },
    { name: 'uniform_1k', params: { dt: 0.002, G: 1, epsilon: 0.02, theta: 0.7 }, bucket: 8 },
    { name: 'cluster_3k', params: { dt: 0.005, G: 1, epsilon: 0.005, theta: 0.5, haloMass: 1, haloA: 0.5, haloCenters: [[0, 0, 0]] }, bucket: 32 },
  ];
  const datas = jobs.map((j) => load(j.name));
  const pool = new GalaxyPool(device, datas.map((d) => d.length / 8 + 500),
    { buckets: jobs.map((j) => j.bucket) });
  datas.forEach((d, k) => pool.upload(k, d, jobs[k].params));
  pool.step(3, 1);
  for (let k = 0; k < jobs.length; k++) {
    const sim = new NBodyGPU(device, datas[k], null, { bucket: jobs[k].bucket });
    await sim.initPromise;
    sim.updateParams(jobs[k].params);
    sim.stepN(3, 1);
    eqBits(pool.particles(k), sim.positionBuffer, `${jobs[k].name} particles`);
    const red = sim.reduce();
    const pr = pool.reduce(k);
    for (let i = 0; i < red.length; i++) {
      if (red[i] !== pr[i]) throw new Error(`${jobs[k].name} reduction ${i}: ${pr[i]} vs ${red[i]}`);
    }
    sim.destroy();
  }
  pool.destroy();
  console.log(`node pool parity ok (${jobs.length} slots)`);
}
main().catch((e) => { console.error(e); process.exit(1); });
