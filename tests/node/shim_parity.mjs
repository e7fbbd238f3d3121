// Node-side parity of the JS drop-in (galaxy-sim_amd/node) against the golden
// vectors written from the reference's mortonTree.js.  Run by
// tests/test_gpu_node.py on a GPU machine:  node tests/node/shim_parity.mjs
import fs from 'fs';
import path from 'path';
import { computeMortonCodes, radixSortIndices, MortonTree, sortAndBuild }
  from '../../galaxy-sim_amd/node/mortonTree.js';
import { NBodyGPU, requestGPUDevice } from '../../galaxy-sim_amd/node/NBodyGPU.js';

const golden = path.join(path.dirname(new URL(import.meta.url).pathname), '..', 'golden');
const manifest = JSON.parse(fs.readFileSync(path.join(golden, 'manifest.json'), 'utf8'));
function load(name, suffix, T) {
  const b = fs.readFileSync(path.join(golden, `${name}.${suffix}`));
  return new T(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}
function eq(a, b, what) {
  if (a.length !== b.length) throw new Error(`${what}: length ${a.length} vs ${b.length}`);
  for (let i = 0; i < a.length; i++) if (a[i] !== b[i]) throw new Error(`${what}: differs at ${i}`);
}

let checked = 0;
for (const c of manifest.cases) {
  if (!c.stored || c.codesOnly || c.maxNodes !== 2 * c.N + 4096) continue;
  const data = load(c.input || c.name, 'in.f32', Float32Array);
  const codes = new Uint32Array(c.N);
  const rs = computeMortonCodes(data, c.N, 8, codes);
  eq(codes, load(c.name, 'codes.u32', Uint32Array), `${c.name} codes`);
  if (rs !== c.rootSize && !(Number.isNaN(rs) && c.rootSize === null)) throw new Error(`${c.name} rootSize ${rs} vs ${c.rootSize}`);
  const perm = radixSortIndices(codes, new Uint32Array(c.N), null, null, c.N);
  eq(perm, load(c.name, 'perm.u32', Uint32Array), `${c.name} perm`);
  const r = sortAndBuild(data, null, c.N, c.trackIdx2);
  eq(r.perm, perm, `${c.name} sortAndBuild perm`);
  eq(new Uint32Array(r.tree._buf, 0, r.tree.numNodes * 16), load(c.name, 'nodes.u32', Uint32Array), `${c.name} nodes`);
  if (JSON.stringify(r.tree.levelRanges) !== JSON.stringify(c.levelRanges)) throw new Error(`${c.name} levelRanges`);
  if (c.centralIdx >= 0 && (r.centralIdx !== c.centralIdx || r.centralIdx2 !== c.centralIdx2)) throw new Error(`${c.name} central`);
  // MortonTree.build on host arrays from the sorted snapshot
  const sp = new Float32Array(c.N * 3), sm = new Float32Array(c.N), sc = new Uint32Array(c.N);
  for (let i = 0; i < c.N; i++) {
    sp[i * 3] = r.sorted[i * 8]; sp[i * 3 + 1] = r.sorted[i * 8 + 1]; sp[i * 3 + 2] = r.sorted[i * 8 + 2];
    sm[i] = r.sorted[i * 8 + 3]; sc[i] = codes[perm[i]];
  }
  const t = new MortonTree(c.N);
  t.build(sc, sp, sm, c.N, rs);
  eq(new Uint32Array(t._buf, 0, t.numNodes * 16), load(c.name, 'nodes.u32', Uint32Array), `${c.name} build`);
  checked++;
}

async function flows() {
  // NBodyGPU: reference flow (host tree -> uploadBHTree) vs on-device sort
  const dev = await requestGPUDevice();
  const c = manifest.cases.find((x) => x.name === 'cluster_3k');
  const data = load(c.name, 'in.f32', Float32Array);
  const r = sortAndBuild(data, null, c.N, -1);
  const a = new NBodyGPU(dev, r.sorted, null);
  await a.initPromise;
  a.updateParams({ dt: 0.005, G: 1, epsilon: 0.02, theta: 0.7 });
  a.uploadBHTree(r.tree._buf, r.tree.numNodes, r.tree.levelRanges);
  a.encodeRefresh(null);
  a.step(null);
  const b = new NBodyGPU(dev, data, null);
  await b.initPromise;
  b.updateParams({ dt: 0.005, G: 1, epsilon: 0.02, theta: 0.7 });
  b.sort();
  b.encodeRefresh(null);
  b.step(null);
  eq(a.positionBuffer, b.positionBuffer, 'uploadBHTree flow vs on-device sort flow');
  const pos = new Float32Array(c.N * 4);
  await b.readPositions(pos);
  if (!(pos[3] !== 0)) throw new Error('readPositions energy missing');
  a.destroy(); b.destroy();
  return checked;
}

flows().then(() => {
  console.log(`node shim parity ok: ${checked} golden cases, NBodyGPU flows agree`);
}).catch((e) => { console.error(e); process.exit(1); });
