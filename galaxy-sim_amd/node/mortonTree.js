// Drop-in replacement of src/physics/mortonTree.js: the same exports with the
// same arguments and results, executed on the GPU through the N-API addon.
// Device state is resident: one simulation handle per particle count (and
// node-pool size, when a tree asks for more than 2N + 4096 nodes), reused by
// every call, so a worker rebuilding each frame allocates nothing per call.
import { addon } from './binding.js';

const MAX_HANDLES = 4;  // least recently used handles beyond this are freed
const handles = new Map();
function sim(N, maxNodes) {
  const mn = maxNodes && maxNodes > 2 * N + 4096 ? maxNodes : 0;
  const key = N + ':' + mn;
  let h = handles.get(key);
  if (h) {
    handles.delete(key);  // re-insert: the Map's order is the use order
  } else {
    h = addon.create({ n: N, maxNodes: mn });
    while (handles.size >= MAX_HANDLES) {
      const [oldKey, old] = handles.entries().next().value;
      addon.destroy(old);
      handles.delete(oldKey);
    }
  }
  handles.set(key, h);
  return h;
}

/** computeMortonCodes (mortonTree.js:43-62) -> root cube size. */
export function computeMortonCodes(pos, N, stride, codes) {
  return addon.computeMortonCodes(sim(N), pos, N, stride, codes);
}

/** radixSortIndices (mortonTree.js:68-84): returns `indices` holding the
 * stable sorted index array (the reference returns one of its two arrays). */
export function radixSortIndices(codes, indices, scratchArr, counts, N) {
  addon.radixSortIndices(sim(N), codes, N, indices);
  return indices;
}

export class MortonTree {
  constructor(maxParticles, bucket) {
    this.maxNodes = maxParticles * 2 + 4096;
    this.bucket = bucket === undefined ? 16 : bucket;
    this._buf = new ArrayBuffer(this.maxNodes * 64);
    this.f32 = new Float32Array(this._buf);
    this.i32 = new Int32Array(this._buf);
    this.numNodes = 0;
    this.levelRanges = [];
  }

  /** build (mortonTree.js:107-182) from Morton-sorted codes / pos / masses. */
  build(codes, pos, masses, N, rootSize) {
    const h = sim(N, this.maxNodes);
    addon.slotSetTree(h, 0, this.bucket, this.maxNodes);  // BUCKET and the pool rule
    const r = addon.treeBuild(h, codes, pos, masses, N, rootSize, this._buf);
    this.numNodes = r.numNodes;
    this.levelRanges = r.levelRanges;
  }

  get byteLength() { return this.numNodes * 64; }
}

/** sortAndBuild (mortonTree.js:192-224): one on-device rebuild (codes,
 * stable sort, gather, BFS build, FP64 COM).  The 16-byte stellar render
 * record is not held on the device; it is gathered on the host by the same
 * permutation. */
export function sortAndBuild(particleData, stellarData, N, trackIdx2) {
  if (trackIdx2 === undefined) trackIdx2 = -1;
  const h = sim(N, 0);
  addon.slotSetTree(h, 0, 16, 0);  // BUCKET = 16, maxNodes = 2N + 4096 (new MortonTree(N))
  addon.upload(h, particleData);
  addon.sort(h);
  const perm = new Uint32Array(N);
  addon.downloadPerm(h, perm);
  const sorted = new Float32Array(N * 8);
  addon.downloadParticles(h, sorted);
  addon.refresh(h);
  const info = addon.treeInfo(h);
  const tree = new MortonTree(N);
  new Uint8Array(tree._buf).set(new Uint8Array(addon.downloadNodes(h)));
  tree.numNodes = info.numNodes;
  tree.levelRanges = info.levelRanges;
  const masses = new Float32Array(N);
  let centralIdx = 0, centralIdx2 = -1;
  for (let i = 0; i < N; i++) {
    masses[i] = sorted[i * 8 + 3];
    if (perm[i] === 0) centralIdx = i;
    if (perm[i] === trackIdx2) centralIdx2 = i;
  }
  const sortedStellar = new Float32Array(N * 4);
  if (stellarData) {
    for (let i = 0; i < N; i++) sortedStellar.set(stellarData.subarray(perm[i] * 4, perm[i] * 4 + 4), i * 4);
  }
  return { sorted, sortedStellar, tree, perm, centralIdx, centralIdx2, masses };
}
