// The JS drop-in's stellar surface: stellarBuffer follows applyPermutation on
// the device, and encodeStarFormation turns certain-to-form dense gas into
// stars (mass in the IMF range, birth = time * agingRate).
// Run by tests/test_gpu_node.py:  node tests/node/stellar_flow.mjs
import { NBodyGPU, requestGPUDevice } from '../../galaxy-sim_amd/node/NBodyGPU.js';

async function main() {
  const device = await requestGPUDevice();
  const N = 5000;
  const parts = new Float32Array(N * 8), stellar = new Float32Array(N * 4);
  for (let i = 0; i < N; i++) {
    parts[i * 8] = Math.cos(i) * (1 + i / N); parts[i * 8 + 1] = Math.sin(i) * (1 + i / N);
    parts[i * 8 + 2] = ((i * 7919) % 101) / 1000; parts[i * 8 + 3] = 1 / N;
    parts[i * 8 + 7] = i % 2 ? -(1 + 1000) : 0;       // odd: dense gas, even: stars
    stellar[i * 4] = i % 2 ? 0 : 1 + (i % 5);          // low-mass stars never recycle
    stellar[i * 4 + 2] = i;                            // tag
  }
  const sim = new NBodyGPU(device, parts, stellar);
  await sim.initPromise;
  const perm = new Uint32Array(N);
  for (let i = 0; i < N; i++) perm[i] = (i * 2957) % N;  // 2957 is prime to 5000
  sim.applyPermutation(null, perm);
  let st = sim.stellarBuffer;
  for (let i = 0; i < N; i++) {
    if (st[i * 4 + 2] !== perm[i]) throw new Error(`stellar tag ${i}: ${st[i * 4 + 2]} vs ${perm[i]}`);
  }
  sim.updateParams({ dt: 0.005, G: 1, epsilon: 0.01, time: 2, agingRate: 25, sfDensity: 10, sfEff: 1, sfDt: 1 });
  sim.encodeStarFormation(null);
  st = sim.stellarBuffer;
  const p = sim.positionBuffer;
  let formed = 0;
  for (let i = 0; i < N; i++) {
    const wasGas = perm[i] % 2 === 1;
    if (wasGas) {
      if (p[i * 8 + 7] !== 0) throw new Error(`gas ${i} did not form a star`);
      if (!(st[i * 4] >= 0.08 && st[i * 4] <= 100)) throw new Error(`mass ${st[i * 4]}`);
      if (st[i * 4 + 1] !== 50) throw new Error(`birth ${st[i * 4 + 1]}`);
      formed++;
    } else if (st[i * 4 + 2] !== perm[i]) {
      throw new Error(`star ${i} changed`);
    }
  }
  sim.destroy();
  console.log(`node stellar flow ok (${formed} stars formed)`);
}
main().catch((e) => { console.error(e); process.exit(1); });
