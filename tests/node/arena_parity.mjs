// GalaxyArena slots (JS drop-in, N-API addon) against the same galaxies
// stepped alone on an NBodyGPU: bit-identical particles after galaxies are
// placed, stepped, removed and a new one placed in the freed range; an
// upload that no free range holds throws with code GS_E_FULL.
// Run by tests/test_gpu_node.py:  node tests/node/arena_parity.mjs
import fs from 'fs';
import path from 'path';
import { NBodyGPU, requestGPUDevice } from '../../galaxy-sim_amd/node/NBodyGPU.js';
import { GalaxyArena } from '../../galaxy-sim_amd/node/GalaxyPool.js';

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
// bytes the arena takes for n particles (GS_ARENA_PARTICLE_BYTES, GS_ARENA_NODE_BYTES)
function charge(n) { return Math.ceil(n / 1024) * 1024 * 112 + (2 * n + 4096) * 36; }

async function alone(device, data, params, schedule) {
  const sim = new NBodyGPU(device, data, null, { bucket: 16 });
  await sim.initPromise;
  sim.updateParams(params);
  for (const [n, every] of schedule) sim.stepN(n, every);
  const out = sim.positionBuffer.slice();
  sim.destroy();
  return out;
}

async function main() {
  const device = await requestGPUDevice();
  const params = { dt: 0.005, G: 1, epsilon: 0.01, theta: 0.6 };
  const a = load('plummer_4k'), b = load('cluster_3k'), c = load('uniform_1k');
  const na = a.length / 8, nb = b.length / 8, nc = c.length / 8;
  // room for a and b, with a margin for the arena's tables
  const arena = new GalaxyArena(device, charge(na) + charge(nb) + (2 << 20), 3);
  arena.upload(0, a, params);
  arena.upload(1, b, params);
  // one particle more than the largest galaxy that still fits
  const over = arena.arenaInfo().largestFit + 1;
  const big = new Float32Array(over * 8);
  for (let i = 0; i < big.length; i += a.length) big.set(a.subarray(0, Math.min(a.length, big.length - i)), i);
  let full = false;
  try {
    arena.upload(2, big, params);
  } catch (e) {
    if (e.code !== 'GS_E_FULL') throw e;
    full = true;
  }
  if (!full) throw new Error('upload beyond the budget was accepted');
  if (arena.counts[2] !== 0) throw new Error('refused slot not empty');
  arena.step(2, 1);
  eqBits(arena.particles(0), await alone(device, a, params, [[2, 1]]), 'slot 0');
  eqBits(arena.particles(1), await alone(device, b, params, [[2, 1]]), 'slot 1');
  arena.clear(0);
  arena.upload(2, c, params);
  arena.step(1, 1);
  eqBits(arena.particles(1), await alone(device, b, params, [[3, 1]]), 'slot 1 after reuse');
  eqBits(arena.particles(2), await alone(device, c, params, [[1, 1]]), 'slot 2 in freed range');
  const info = arena.arenaInfo();
  if (info.usedBytes !== charge(nb) + charge(nc)) throw new Error(`usedBytes ${info.usedBytes}`);
  arena.destroy();
  console.log('node arena parity ok');
}
main().catch((e) => { console.error(e); process.exit(1); });
