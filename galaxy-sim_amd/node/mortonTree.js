// Drop-in replacement of src/physics/mortonTree.js: the same exports with the
// same arguments and results, executed on the GPU through the N-API addon
// (one resident scratch simulation per particle count).
import { addon } from './binding.js';

let scratch = null, scratchN = -1;
function sim(N) {
  if (scratchN !== N) {
    if (scratch) addon.destroy(scratch);
    scratch = addon.create({ n: N });
    scratchN = N;
  }
  return scratch;
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
    const h = addon.create({ n: N, bucket: this.bucket, maxNodes: this.maxNodes });
    try {
      const r = addon.treeBuild(h, codes, pos, masses, N, rootSize, this._buf);
      this.numNodes = r.numNodes;
      this.levelRanges = r.levelRanges;
    } finally {
      addon.destroy(h);
    }
  }

  get byteLength() { return this.numNodes * 64; }
}

/** sortAndBuild (mortonTree.js:192-224): one on-device rebuild (codes,
 * stable sort, gather, BFS build, FP64 COM).  The 16-byte stellar render
 * record is not held on the device; it is gathered on the host by the same
 * permutation. */
export function sortAndBuild(particleData, stellarData, N, trackIdx2) {
  if (trackIdx2 === undefined) trackIdx2 = -1;
  const h = addon.create({ n: N });
  try {
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
  } finally {
    addon.destroy(h);
  }
}
