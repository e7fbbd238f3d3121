// Drop-in replacement of src/physics/NBodyGPU.js on MI355X: the same class,
// constructor, methods and getters, backed by libgalaxy_hip.so through the
// N-API addon instead of WebGPU.
//
// Differences a caller sees:
//   - `device` is the object returned by requestGPUDevice() ({ordinal}).
//   - Command encoders are not needed: every method takes the reference's
//     `commandEncoder` argument and ignores it; work is queued on the
//     simulation's own HIP stream in call order, which is the order a WebGPU
//     queue would execute the encoded passes.
//   - stellarBuffer returns a host copy of the stellar records; they live on
//     the device and follow every permutation (sort or applyPermutation).
//   - sort() and stepN() expose the on-device rebuild (codes, sort, build).
import { addon } from './binding.js';

/** requestGPUDevice (NBodyGPU.js:2-12): first gfx950 device. */
export async function requestGPUDevice() {
  if (addon.deviceCount() < 1) throw new Error('No MI355X (gfx950) HIP device found.');
  return { ordinal: 0, kind: 'hip-gfx950' };
}

/** Halo-centre EMA of the frame loop (main.js:449-457), in place. */
export function trackHaloCenters(haloCenters, targets, k) {
  const kk = k === undefined ? 0.1 : k;
  for (let h = 0; h < haloCenters.length && h < targets.length; h++) {
    for (let c = 0; c < 3; c++) haloCenters[h][c] += (targets[h][c] - haloCenters[h][c]) * kk;
  }
  return haloCenters;
}

export class NBodyGPU {
  constructor(device, particleData, stellarData, options) {
    this.device = device;
    this.N = particleData.length / 8;
    this.currentBuffer = 0;
    this.stellarCurrent = 0;
    this.ready = false;
    this.options = options || {};
    this.initPromise = this._init(particleData, stellarData);
  }

  async _init(particleData, stellarData) {
    const o = this.options;
    this.h = addon.create({
      n: this.N,
      bucket: o.bucket === undefined ? 16 : o.bucket,
      maxNodes: o.maxNodes === undefined ? 0 : o.maxNodes,
      exactForce: o.exactForce ? 1 : 0,
      timing: o.timing ? 1 : 0,
      device: this.device && typeof this.device.ordinal === 'number' ? this.device.ordinal : -1,
    });
    addon.upload(this.h, particleData);
    // stellar records live on the device next to the particles (NBodyGPU.js:83-90)
    this.hasStellar = !!stellarData;
    if (stellarData) addon.uploadStellar(this.h, new Float32Array(stellarData));
    this.levelRanges = null;
    this.ready = true;
  }

  /** applyPermutation (NBodyGPU.js:225-238): gather by a host permutation. */
  applyPermutation(commandEncoder, perm) {
    if (!this.ready) return;
    addon.applyPermutation(this.h, perm);  // particles and stellar records
    this.currentBuffer ^= 1;
    this.stellarCurrent ^= 1;
  }

  /** stellarBuffer (NBodyGPU.js:241-243): N x 4 f32 in the current order (host copy). */
  get stellarBuffer() {
    if (!this.hasStellar) return null;
    const out = new Float32Array(this.N * 4);
    addon.downloadStellar(this.h, out);
    return out;
  }

  updateCentralMass(mass, idx) {
    addon.updateCentralMass(this.h, mass, idx === undefined ? 0 : idx);
  }

  /** updateParams (NBodyGPU.js:253-275): same fields and defaults. */
  updateParams(p) {
    addon.setParams(this.h, {
      dt: p.dt, G: p.G, epsilon: p.epsilon,
      theta: p.theta === undefined ? 0.5 : p.theta,
      haloMass: p.haloMass === undefined ? 0 : p.haloMass,
      haloA: p.haloA === undefined ? 1 : p.haloA,
      gasDrag: p.gasDrag === undefined ? 0 : p.gasDrag,
      time: p.time === undefined ? 0 : p.time,
      agingRate: p.agingRate === undefined ? 25 : p.agingRate,
      sfDensity: p.sfDensity === undefined ? 1e9 : p.sfDensity,
      sfEff: p.sfEff === undefined ? 0 : p.sfEff,
      sfDt: p.sfDt === undefined ? 0 : p.sfDt,
      haloCenters: p.haloCenters === undefined ? [] : p.haloCenters,
    });
  }

  /** encodeStarFormation (NBodyGPU.js:277-285): star_formation in place. */
  encodeStarFormation(commandEncoder) {
    if (!this.ready || !this.hasStellar) return;
    addon.starFormation(this.h);
  }

  /** uploadBHTree (NBodyGPU.js:288-301): tree built by mortonTree.build. */
  uploadBHTree(treeBuffer, numNodes, levelRanges) {
    if (!this.ready) return;
    addon.uploadTree(this.h, treeBuffer, numNodes, levelRanges || []);
    if (levelRanges) this.levelRanges = levelRanges;
  }

  /** encodeRefresh (NBodyGPU.js:308-319): FP64 COM refresh of the tree. */
  encodeRefresh(commandEncoder) {
    if (!this.ready || !this.hasTree()) return;
    addon.refresh(this.h);
  }

  /** step (NBodyGPU.js:321-329): the `main` dispatch alone -- walk, halos,
   * kick-drift and gas drag on the tree's current COMs, then flip.  The
   * refresh is encodeRefresh's, as in the reference frame loop. */
  step(commandEncoder) {
    if (!this.ready) return;
    addon.walkStep(this.h);
    this.currentBuffer ^= 1;
  }

  /** On-device topology rebuild: codes, stable sort, gather, BFS build. */
  sort() {
    addon.sort(this.h);
    const info = addon.treeInfo(this.h);
    this.levelRanges = info.levelRanges;
    this.currentBuffer ^= 1;
    return info;
  }

  /** n x (refresh + step) with an on-device rebuild every sortEvery steps. */
  stepN(n, sortEvery) {
    addon.step(this.h, n, sortEvery === undefined ? 1 : sortEvery);
  }

  hasTree() {
    try { addon.treeInfo(this.h); return true; } catch (e) { return false; }
  }

  encodePositionReadback(commandEncoder) { /* readPositions copies on demand */ }

  /** readPositions (NBodyGPU.js:346-350): vec4 (x, y, z, specific energy). */
  readPositions(out) { return addon.readPositionsAsync(this.h, out); }

  /** Current particle records, N x 8 f32 (the buffer a renderer would read). */
  get positionBuffer() {
    const out = new Float32Array(this.N * 8);
    addon.downloadParticles(this.h, out);
    return out;
  }

  /** remapIndex (main.js:436-440) after an on-device rebuild: {index, pos}. */
  locate(oldIndex) { return addon.locate(this.h, oldIndex); }

  treeInfo() { return addon.treeInfo(this.h); }
  downloadNodes() { return addon.downloadNodes(this.h); }
  downloadPerm() { const p = new Uint32Array(this.N); addon.downloadPerm(this.h, p); return p; }
  reduce() { return addon.reduce(this.h); }

  destroy() {
    if (this.h) addon.destroy(this.h);
    this.h = null;
    this.ready = false;
  }
}
