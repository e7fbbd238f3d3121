// Many galaxies on one MI355X, stepped together: the slot pool of
// libgalaxy_hip.so (gs_create_pool / gs_slot_*) for a Node.js host.  The
// reference runs one galaxy per NBodyGPU; a pool keeps many resident in
// fixed-capacity slots and runs every stage of a step (codes, sort, build,
// refresh, walk) as one launch over all of them, replayed as one captured
// graph per step.  A slot's results are bit-identical to the same galaxy
// stepped alone on an NBodyGPU.
import { addon } from './binding.js';

function paramsOf(p) {
  // updateParams defaults (NBodyGPU.js:253-255)
  return {
    dt: p.dt, G: p.G, epsilon: p.epsilon,
    theta: p.theta === undefined ? 0.5 : p.theta,
    haloMass: p.haloMass === undefined ? 0 : p.haloMass,
    haloA: p.haloA === undefined ? 1 : p.haloA,
    gasDrag: p.gasDrag === undefined ? 0 : p.gasDrag,
    time: p.time === undefined ? 0 : p.time,
    haloCenters: p.haloCenters === undefined ? [] : p.haloCenters,
  };
}

export class GalaxyPool {
  /**
   * capacities: particles per slot (Array or Uint32Array, 1 .. 256 slots).
   * options: {buckets, exactForce, device, walkWg, radixBits, graph}.
   */
  constructor(device, capacities, options, arena) {
    const o = options || {};
    if (arena) {  // GalaxyArena: slots without fixed capacities
      this.capacities = new Array(arena.slots).fill(0);
      this.counts = new Array(arena.slots).fill(0);
      this.h = arena.create();
      return;
    }
    this.capacities = Array.from(capacities);
    this.counts = new Array(this.capacities.length).fill(0);
    this.h = addon.createPool({
      capacities: this.capacities,
      buckets: o.buckets,
      exactForce: o.exactForce ? 1 : 0,
      device: device && typeof device.ordinal === 'number' ? device.ordinal : -1,
      walkWg: o.walkWg || 0,
      radixBits: o.radixBits || 0,
      graph: o.graph === false ? 0 : 1,
    });
  }

  get nslots() { return this.capacities.length; }
  info() { return addon.poolInfo(this.h); }

  /** Place a galaxy (N x 8 f32 records) in a slot, with its updateParams. */
  upload(slot, particleData, params) {
    addon.slotUpload(this.h, slot, particleData);
    this.counts[slot] = particleData.length / 8;
    if (params) this.setParams(slot, params);
  }

  clear(slot) { addon.slotClear(this.h, slot); this.counts[slot] = 0; }
  setParams(slot, params) { addon.slotSetParams(this.h, slot, paramsOf(params)); }
  /** Leaf capacity (mortonTree.js BUCKET) and leaf-rule pool (0 -> 2n + 4096). */
  setTree(slot, bucket, maxNodes) { addon.slotSetTree(this.h, slot, bucket, maxNodes || 0); }

  /** n steps of every occupied slot, rebuilding topology every sortEvery steps. */
  step(n, sortEvery) { addon.step(this.h, n === undefined ? 1 : n, sortEvery === undefined ? 1 : sortEvery); }
  sync() { addon.sync(this.h); }

  /** A slot's particle records in its current (sorted) order. */
  particles(slot) {
    const out = new Float32Array(this.counts[slot] * 8);
    if (this.counts[slot]) addon.slotDownloadParticles(this.h, slot, out);
    return out;
  }

  energy(slot) {
    const out = new Float32Array(this.counts[slot]);
    if (this.counts[slot]) addon.slotDownloadEnergy(this.h, slot, out);
    return out;
  }

  /** FP64 [ekin, epot, ehalo, px, py, pz, espec, mass] of a slot's last step. */
  reduce(slot) { return addon.slotReduce(this.h, slot); }

  destroy() {
    if (this.h) addon.destroy(this.h);
    this.h = null;
  }
}

/**
 * A pool whose device state is sized by a byte budget (gs_create_arena):
 * `slots` slots start empty; upload() places a galaxy in a free range of the
 * arena and throws an Error with code 'GS_E_FULL' (the slot left empty) when
 * no range holds it; clear() gives the range back.  Steps launch over the
 * occupied slots only.  No stellar records.
 */
export class GalaxyArena extends GalaxyPool {
  /** options: {exactForce, device, walkWg, radixBits, graph}. */
  constructor(device, budgetBytes, slots, options) {
    const o = options || {};
    super(null, [], null, {
      create: () => addon.createArena({
        budgetBytes,
        slots,
        exactForce: o.exactForce ? 1 : 0,
        device: device && typeof device.ordinal === 'number' ? device.ordinal : -1,
        walkWg: o.walkWg || 0,
        radixBits: o.radixBits || 0,
        graph: o.graph === false ? 0 : 1,
      }),
      slots,
    });
    this.budgetBytes = budgetBytes;
  }

  upload(slot, particleData, params) {
    try {
      super.upload(slot, particleData, params);
    } catch (e) {
      this.counts[slot] = 0;
      throw e;
    }
  }

  /** {deviceBytes, usedBytes, largestFit}: the arena, its resident ranges, the largest galaxy that fits now. */
  arenaInfo() { return addon.arenaInfo(this.h); }
}
