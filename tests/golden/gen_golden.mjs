// Writes golden vectors for the Morton / sort / BFS-tree path by importing the
// reference's own src/physics/mortonTree.js (run in the build container only:
//   node tests/golden/gen_golden.mjs
// ).  Outputs, per case, a JSON manifest plus raw little-endian arrays.  Small
// cases store inputs and outputs; large cases store inputs generated by an
// integer-exact PRNG recipe (re-created bit-for-bit by tests/golden/cases.py)
// and SHA-256 digests of the outputs.
import fs from 'fs';
import path from 'path';
import crypto from 'crypto';
import { computeMortonCodes, radixSortIndices, MortonTree, sortAndBuild }
  from '/root/reference/src/physics/mortonTree.js';

const outDir = path.dirname(new URL(import.meta.url).pathname);

// xorshift32 — the same integer recipe lives in tests/golden/cases.py
function rng(seed) {
  let s = seed >>> 0 || 1;
  return () => {
    s ^= s << 13; s >>>= 0;
    s ^= s >>> 17;
    s ^= s << 5; s >>>= 0;
    return s;
  };
}

function sha(buf) {
  return crypto.createHash('sha256').update(Buffer.from(buf.buffer, buf.byteOffset, buf.byteLength)).digest('hex');
}
function hexDouble(d) {
  const b = Buffer.alloc(8); b.writeDoubleLE(d); return b.toString('hex');
}
function write(name, arr) {
  fs.writeFileSync(path.join(outDir, name), Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));
}

// ── input recipes ───────────────────────────────────────────────────────────
// "exact" recipe: every coordinate is a dyadic rational whose f32 rounding is
// a single RNE step from an exactly computed double (reproducible in numpy).
function exactCluster(N, seed, stride) {
  const r = rng(seed);
  const a = new Float32Array(N * stride);
  const centres = [];
  for (let c = 0; c < 6; c++) {
    centres.push([((r() >>> 24) - 128) / 16, ((r() >>> 24) - 128) / 16, ((r() >>> 26) - 32) / 16]);
  }
  for (let i = 0; i < N; i++) {
    const c = centres[r() % 6];
    const sc = ((r() >>> 29) + 1) * ((r() >>> 29) + 1);
    for (let k = 0; k < 3; k++) {
      const off = (((r() >>> 16) - 32768) * sc) / 1048576;
      a[i * stride + k] = c[k] + off;
    }
    a[i * stride + 3] = ((r() >>> 20) + 1) / 4096;
    if (stride === 8) {
      for (let k = 4; k < 7; k++) a[i * stride + k] = (((r() >>> 16) - 32768)) / 65536;
      a[i * stride + 7] = (r() % 10 < 3) ? -1 : 0;
    }
  }
  return a;
}

function uniformCube(N, seed, stride, scale, shift) {
  const r = rng(seed);
  const a = new Float32Array(N * stride);
  for (let i = 0; i < N; i++) {
    for (let k = 0; k < 3; k++) a[i * stride + k] = (r() / 4294967296) * scale + shift;
    a[i * stride + 3] = (r() >>> 8) / 16777216 + 0.001;
    if (stride === 8) for (let k = 4; k < 8; k++) a[i * stride + k] = 0;
  }
  return a;
}

function plummer(N, seed, stride) {
  const r = rng(seed);
  const a = new Float32Array(N * stride);
  for (let i = 0; i < N; i++) {
    let u = r() / 4294967296; if (u < 1e-6) u = 1e-6;
    const rad = 1 / Math.sqrt(Math.pow(u, -2 / 3) - 1);
    const ct = 2 * (r() / 4294967296) - 1, ph = 2 * Math.PI * (r() / 4294967296);
    const st = Math.sqrt(1 - ct * ct);
    a[i * stride] = rad * st * Math.cos(ph);
    a[i * stride + 1] = rad * st * Math.sin(ph);
    a[i * stride + 2] = rad * ct;
    a[i * stride + 3] = 1 / N;
  }
  return a;
}

// ── runners ─────────────────────────────────────────────────────────────────
const manifest = { generator: 'tests/golden/gen_golden.mjs', source: 'src/physics/mortonTree.js', cases: [] };

function runSortAndBuild(name, data, N, opts = {}) {
  const { store = true, maxNodes = null, stellar = null, trackIdx2 = -1, recipe = null, input = null } = opts;
  const st = stellar || new Float32Array(N * 4);
  let res;
  if (maxNodes === null) {
    res = sortAndBuild(data, st, N, trackIdx2);
  } else {
    // Same composition as sortAndBuild (mortonTree.js:192-224) with the pool
    // size overridden on the MortonTree instance before build().
    const codes = new Uint32Array(N), idx = new Uint32Array(N), scr = new Uint32Array(N);
    const rootSize = computeMortonCodes(data, N, 8, codes);
    const perm = radixSortIndices(codes, idx, scr, new Uint32Array(1024), N);
    const sortedCodes = new Uint32Array(N), sp = new Float32Array(N * 3), sm = new Float32Array(N);
    const sorted = new Float32Array(N * 8);
    for (let i = 0; i < N; i++) {
      const s = perm[i];
      sorted.set(data.subarray(s * 8, s * 8 + 8), i * 8);
      sortedCodes[i] = codes[s];
      sp[i * 3] = data[s * 8]; sp[i * 3 + 1] = data[s * 8 + 1]; sp[i * 3 + 2] = data[s * 8 + 2];
      sm[i] = data[s * 8 + 3];
    }
    const tree = new MortonTree(N);
    tree.maxNodes = maxNodes;
    tree.build(sortedCodes, sp, sm, N, rootSize);
    res = { sorted, tree, perm: new Uint32Array(perm), centralIdx: -1, centralIdx2: -1 };
  }
  // raw codes in input order
  const codes = new Uint32Array(N);
  const rootSize = computeMortonCodes(data, N, 8, codes);
  const nodes = new Uint32Array(res.tree._buf, 0, res.tree.numNodes * 16);
  const c = {
    name, N, stride: 8, bucket: 16, maxNodes: maxNodes === null ? 2 * N + 4096 : maxNodes,
    rootSizeHex: hexDouble(rootSize), rootSize,
    numNodes: res.tree.numNodes, levelRanges: res.tree.levelRanges,
    centralIdx: res.centralIdx, centralIdx2: res.centralIdx2, trackIdx2,
    sha: { codes: sha(codes), perm: sha(res.perm), nodes: sha(nodes), sorted: sha(res.sorted) },
    stored: store, recipe, input: input || name,
  };
  if (store) {
    if (!input) write(`${name}.in.f32`, data);
    write(`${name}.codes.u32`, codes);
    write(`${name}.perm.u32`, res.perm);
    write(`${name}.nodes.u32`, new Uint32Array(nodes));
    if (stellar) write(`${name}.stellar.f32`, stellar);
  }
  manifest.cases.push(c);
  return c;
}

function runCodesOnly(name, data, N, stride) {
  const codes = new Uint32Array(N);
  const rootSize = computeMortonCodes(data, N, stride, codes);
  write(`${name}.in.f32`, data);
  write(`${name}.codes.u32`, codes);
  manifest.cases.push({ name, N, stride, codesOnly: true, rootSizeHex: hexDouble(rootSize), rootSize,
    sha: { codes: sha(codes) }, stored: true });
}

// 1. generic distributions
runSortAndBuild('uniform_1k', uniformCube(1000, 11, 8, 20, -10), 1000);
runSortAndBuild('plummer_4k', plummer(4096, 12, 8), 4096);
runSortAndBuild('cluster_3k', exactCluster(3000, 13, 8), 3000, { trackIdx2: 1234 });

// 2. edge cases
runSortAndBuild('n1', uniformCube(1, 21, 8, 1, 0), 1);
runSortAndBuild('n16', uniformCube(16, 22, 8, 1, 0), 16);
runSortAndBuild('n17', uniformCube(17, 23, 8, 1, 0), 17);
{
  const N = 200, a = new Float32Array(N * 8);
  for (let i = 0; i < N; i++) { a[i * 8] = 1.5; a[i * 8 + 1] = -2.25; a[i * 8 + 2] = 0.125; a[i * 8 + 3] = 0.01 * (1 + (i % 7)); }
  runSortAndBuild('coincident_200', a, N);
}
{
  // points on the 1024-cell lattice of their own bbox, incl. both bbox faces
  const N = 729, a = new Float32Array(N * 8); let i = 0;
  for (let x = 0; x < 9; x++) for (let y = 0; y < 9; y++) for (let z = 0; z < 9; z++, i++) {
    a[i * 8] = x * 128; a[i * 8 + 1] = y * 128 - 512; a[i * 8 + 2] = z * 0.125; a[i * 8 + 3] = 1;
  }
  runSortAndBuild('lattice_729', a, N);
}
{
  // non-finite particles: skipped by the bbox, still encoded (ToInt32 semantics)
  const N = 300, a = uniformCube(N, 24, 8, 4, -2);
  a[5 * 8] = NaN; a[17 * 8 + 1] = Infinity; a[29 * 8 + 2] = -Infinity;
  a[40 * 8] = NaN; a[40 * 8 + 1] = 1e30;                // finite y far outside the bbox
  a[41 * 8 + 2] = NaN; a[41 * 8] = -3e9;               // finite x below the bbox
  a[42 * 8 + 1] = NaN; a[42 * 8] = 7e12;                // ToInt32 wrap-around
  runSortAndBuild('nonfinite_300', a, N);
}
{
  const N = 2048, a = uniformCube(N, 25, 8, 3e30, -1e30);
  runSortAndBuild('huge_2k', a, N);
  const b = uniformCube(N, 26, 8, 3e-30, 1e-29);
  runSortAndBuild('tiny_2k', b, N);
}
{
  // zero masses: leaves with m = 0 get COM 0 (m > 0 ? 1/m : 0)
  const N = 500, a = uniformCube(N, 27, 8, 1, 0);
  for (let i = 0; i < 120; i++) a[i * 8 + 3] = 0;
  runSortAndBuild('zeromass_500', a, N);
}
{
  // node-pool cap: maxNodes forces leaves early (mortonTree.js:126-127)
  const N = 4096, a = plummer(N, 28, 8);
  const free = runSortAndBuild('cap_free', a, N);
  for (const cap of [60, 200, 500]) {
    runSortAndBuild(`cap_${cap}`, a, N, { maxNodes: cap, store: true, input: 'cap_free' });
  }
}
{
  const N = 1500, a = exactCluster(N, 29, 8);
  const stellar = new Float32Array(N * 4);
  for (let i = 0; i < N * 4; i++) stellar[i] = i * 0.25;
  runSortAndBuild('stellar_1500', a, N, { stellar, trackIdx2: 777 });
}
// 3. worker stride-4 codes (bhWorker.js:58)
runCodesOnly('stride4_2k', exactCluster(2000, 30, 4), 2000, 4);

// 4. large cases: inputs re-created from the integer recipe, digests only
for (const [name, N, seed] of [['cluster_64k', 65536, 41], ['cluster_262k', 262144, 42]]) {
  runSortAndBuild(name, exactCluster(N, seed, 8), N, { store: false, recipe: { kind: 'exactCluster', seed, stride: 8 } });
}

fs.writeFileSync(path.join(outDir, 'manifest.json'), JSON.stringify(manifest, null, 1));
console.log(`wrote ${manifest.cases.length} cases`);
