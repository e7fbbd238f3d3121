/*
 * galaxy_hip — MI355X (gfx950) Barnes-Hut gravity step behind a plain C ABI.
 *
 * The step is the gravity path of FlaccidNoise/galaxy-sim:
 *   30-bit Morton encode -> stable LSD radix sort + permutation gather ->
 *   bucketed breadth-first octree build -> bottom-up COM refresh ->
 *   theta-criterion tree walk (Plummer softening, sub-softening monopole) ->
 *   halos + kick-drift + gas drag.
 * Each entry point below names the reference interface it replaces.  The
 * reference is JavaScript (WebGPU + a Web Worker), so the "interface" is the
 * JS call surface of src/physics/NBodyGPU.js and src/physics/mortonTree.js;
 * the N-API addon and JS shim in galaxy-sim_amd/node re-export it unchanged.
 *
 * Conventions
 *   - Every function returns GS_OK (0) or a negative GS_E* status; the
 *     handle's gs_last_error() string says what went wrong.  No C++ exception
 *     crosses this boundary.
 *   - A handle owns one HIP stream and all device state of one galaxy, or of
 *     a pool of galaxy slots (gs_create_pool).  It is single-threaded;
 *     distinct handles may be driven from distinct threads.
 *   - Host buffers are caller-owned and copied at call time (the reference's
 *     writeBuffer / mapAsync semantics).  Calls are stream-ordered and return
 *     once work is queued, except the gs_download_* / gs_read_* / gs_sync /
 *     gs_reduce / host-buffer mortonTree calls, which wait for the stream.
 *   - Particle record (nbody_bh.wgsl:10-17): 8 x f32
 *       [x, y, z, mass, vx, vy, vz, pad]   pad = 0 star, < -0.5 gas.
 *   - Node record (mortonTree.js:12-21): 16 x 32-bit words
 *       f32 com[3], f32 mass, f32 size, i32 isLeaf, i32 first, i32 count,
 *       i32 children[8] (-1 = empty).
 */
#ifndef GALAXY_HIP_H
#define GALAXY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_OK 0
#define GS_E_INVALID (-1)   /* bad argument or size */
#define GS_E_DEVICE (-2)    /* HIP runtime / no usable gfx950 device */
#define GS_E_NOMEM (-3)     /* device allocation failed */
#define GS_E_NODE_POOL (-4) /* tree needs more nodes than the pool holds */
#define GS_E_STATE (-5)     /* call order: e.g. force before any sort */
#define GS_E_STACK (-6)     /* walk stack overflow (unreachable, depth <= 10) */
#define GS_E_FULL (-7)      /* arena pool: no free range holds the galaxy */

/* gs_config.flags */
#define GS_FLAG_EXACT_FORCE 0x1u /* IEEE sqrt and '/' in every interaction:   \
                                    forces bit-identical to the transliterated \
                                    WGSL; default is the rsqrt fast path */
#define GS_FLAG_TIMING 0x2u      /* time stages with HIP events; gs_step  \
                                    then replays two graphs per step        \
                                    (rebuild + refresh, walk) and times the \
                                    walk and the whole step */
#define GS_FLAG_NO_GRAPH 0x4u    /* gs_step launches kernels one by one
                                    instead of replaying a captured graph */

typedef struct gs_sim gs_sim;

typedef struct gs_config {
  uint32_t n;         /* particles, N (NBodyGPU.js:17); 1 .. 2^27 - 1 */
  uint32_t bucket;    /* leaf capacity BUCKET (mortonTree.js:25); 0 -> 16 */
  uint64_t max_nodes; /* leaf-rule pool size maxNodes (mortonTree.js:88,127);
                         0 -> 2n + 4096.  The device pool holds
                         min(max(2n + 4096, max_nodes), 2^26) nodes (the
                         walk addresses them by 32-bit byte offsets);
                         max_nodes at most 2^26. */
  uint32_t flags;     /* GS_FLAG_* */
  int32_t device;     /* HIP device ordinal, -1 = current device */
  uint32_t walk_wg;   /* walk workgroup size, a power of two in [64, 1024];
                         0 -> tuned default */
  uint32_t radix_bits; /* sort digit width, 8 (4 passes) or 10 (3 passes);
                          0 -> 8.  The permutation is the same. */
  uint32_t reserved[4];
} gs_config;

/* NBodyGPU.updateParams({dt, G, epsilon, theta, haloMass, haloA, gasDrag,
 * time, agingRate, sfDensity, sfEff, sfDt, haloCenters}) (NBodyGPU.js:253-275).
 * Fields are JS Numbers (doubles); gs_set_params rounds each to f32 exactly
 * as the Float32Array store does, and eps2 = f32(epsilon * epsilon). */
typedef struct gs_params {
  double dt, G, epsilon, theta;
  double halo_mass, halo_a, gas_drag, time;
  double aging_rate, sf_density, sf_eff, sf_dt;
  uint32_t halo_count; /* haloCenters.length, clamped to 2 */
  uint32_t pad_;
  double halo_centers[2][3];
} gs_params;

/* Reductions returned by gs_reduce (FP64, over all particles, in the current
 * state after the last walk). */
enum {
  GS_RED_EKIN = 0,   /* sum 1/2 m |v|^2 */
  GS_RED_EPOT = 1,   /* sum 1/2 m phi_pairs */
  GS_RED_EHALO = 2,  /* sum m phi_halo */
  GS_RED_PX = 3,     /* sum m v */
  GS_RED_PY = 4,
  GS_RED_PZ = 5,
  GS_RED_ESPEC = 6,  /* sum of per-particle specific energy, the reference's
                        drift diagnostic (nbody_bh.wgsl:180, bhWorker.js:74) */
  GS_RED_MASS = 7,
  GS_NRED = 8
};

/* Stage timers returned by gs_stage_times (milliseconds, summed). */
enum {
  GS_T_ENCODE = 0, GS_T_SORT = 1, GS_T_BUILD = 2, GS_T_REFRESH = 3,
  GS_T_WALK = 4, GS_T_STEP = 5, GS_NT = 6
};

const char* gs_version(void);
/* Number of usable gfx950 devices (0 when none). */
int gs_device_count(int* out);
/* JS defaults of updateParams (NBodyGPU.js:253-255). */
void gs_params_default(gs_params* p);

/* new NBodyGPU(device, particleData, stellarData) (NBodyGPU.js:15-22, _init
 * 24-203): allocates every device buffer of one galaxy. */
int gs_create(const gs_config* cfg, gs_sim** out);
/* NBodyGPU.destroy() (NBodyGPU.js:358-368). */
void gs_destroy(gs_sim* sim);
const char* gs_last_error(const gs_sim* sim);

/* writeBuffer of the particle records (NBodyGPU.js:41-42): n x 8 f32. */
int gs_upload(gs_sim* sim, const float* particles, uint64_t n);
/* updateParams (NBodyGPU.js:253-275). */
int gs_set_params(gs_sim* sim, const gs_params* p);
/* updateCentralMass(mass, idx) (NBodyGPU.js:245-251). */
int gs_update_central_mass(gs_sim* sim, float mass, uint32_t idx);

/* computeMortonCodes(pos, N, 8, codes) on the current particles
 * (mortonTree.js:43-62); keeps the codes for gs_download_codes(.., 0). */
int gs_encode(gs_sim* sim);
/* applyPermutation(enc, perm) (NBodyGPU.js:225-238): gather the particle
 * records through a host permutation (perm[new] = old, a bijection of
 * [0, N)) into the other ping-pong buffer. */
int gs_apply_permutation(gs_sim* sim, const uint32_t* perm, uint64_t n);
/* uploadBHTree(treeBuffer, numNodes, levelRanges) (NBodyGPU.js:288-301): a
 * tree in the 64-byte layout built on the host (e.g. by mortonTree.js) for the
 * current particle order.  Children of each node must be contiguous in octant
 * order, as every breadth-first MortonTree.build emits them. */
int gs_upload_tree(gs_sim* sim, const uint32_t* nodes, uint64_t num_nodes,
                   const uint32_t* level_ranges, uint32_t num_levels);
/* One topology rebuild: the worker's computeMortonCodes + radixSortIndices +
 * gather (bhWorker.js:58-80), then applyPermutation (NBodyGPU.js:225-238)
 * and MortonTree.build + uploadBHTree (mortonTree.js:107-182,
 * NBodyGPU.js:288-301).  Node COMs are filled by the next refresh. */
int gs_sort(gs_sim* sim);
/* encodeRefresh (NBodyGPU.js:308-319): node COM/mass re-summed from the
 * current positions in FP64, in the build's order (mortonTree.js:129-181). */
int gs_refresh(gs_sim* sim);
/* The walk of `main` alone (nbody_bh.wgsl:69-142): acceleration, pair
 * potential and leaf density per particle, no integration. */
int gs_force(gs_sim* sim);
/* nsteps x (encodeRefresh + step) (NBodyGPU.js:308-329), with a topology
 * rebuild (gs_sort) before every step whose running index is a multiple of
 * sort_every (0 = never rebuild; a tree must exist). */
int gs_step(gs_sim* sim, uint32_t nsteps, uint32_t sort_every);
/* step(commandEncoder) (NBodyGPU.js:321-329): the `main` dispatch alone --
 * walk, halos, kick-drift and gas drag on the node COMs the tree holds now,
 * then the ping-pong flip.  It neither refreshes nor rebuilds: encodeRefresh
 * (gs_refresh) is its own call, as in the frame loop (main.js:471-477), so a
 * step after updateCentralMass without a refresh walks the old COMs, as the
 * reference does.  Needs a tree whose COMs were filled (gs_refresh,
 * gs_upload_tree or a gs_step since the last gs_sort); GS_E_STATE otherwise. */
int gs_walk_step(gs_sim* sim);
int gs_sync(gs_sim* sim);

/* Current particle records in current (sorted) order, n x 8 f32. */
int gs_download_particles(gs_sim* sim, float* dst, uint64_t n);
/* encodePositionReadback + readPositions (NBodyGPU.js:335-350,
 * compact_positions nbody_bh.wgsl:186-195): n x vec4 (x, y, z, energy). */
int gs_read_positions(gs_sim* sim, float* dst4, uint64_t n);
/* Morton codes: sorted = 0 -> codes of the last gs_encode in input order;
 * sorted = 1 -> codes in sorted order after the last gs_sort. */
int gs_download_codes(gs_sim* sim, uint32_t* dst, uint64_t n, int sorted);
/* The gather permutation of the last gs_sort (perm[new] = old). */
int gs_download_perm(gs_sim* sim, uint32_t* dst, uint64_t n);
/* Tree in the reference's 64-byte node layout; cap is in nodes. */
int gs_download_nodes(gs_sim* sim, uint32_t* dst, uint64_t cap,
                      uint64_t* num_nodes);
/* numNodes, levelRanges ([start, count] x up to 16) and rootSize. */
int gs_tree_info(gs_sim* sim, uint64_t* num_nodes, uint32_t* level_ranges,
                 uint32_t* num_levels, double* root_size);
/* energyOut (nbody_bh.wgsl:180) of the last step. */
int gs_download_energy(gs_sim* sim, float* dst, uint64_t n);
/* Result of the last gs_force: n x (ax, ay, az, pot), and leaf density
 * (either pointer may be NULL). */
int gs_download_accel(gs_sim* sim, float* acc4, float* density, uint64_t n);
/* Stellar records (NBodyGPU.js:83-90, stellarBuffer :241-243): n x 4 f32
 * [stellar mass, birth time (Myr), seed, dust] in the current particle order.
 * Allocated on first upload; the sort's gather and gs_apply_permutation carry
 * them along (apply_permutation_stellar, nbody_bh.wgsl:336-340); a new
 * particle upload into a slot clears that slot's records. */
int gs_upload_stellar(gs_sim* sim, const float* stellar4, uint64_t n);
int gs_download_stellar(gs_sim* sim, float* dst4, uint64_t n);
/* encodeStarFormation (NBodyGPU.js:277-285, star_formation
 * nbody_bh.wgsl:229-268): gas denser than sfDensity turns into stars, massive
 * stars past their supernova return to gas with a kick; every occupied slot,
 * in place, with each slot's time / agingRate / sfDensity / sfEff / sfDt. */
int gs_star_formation(gs_sim* sim);
/* remapIndex (main.js:436-440) on the device: the current index of the
 * particle that had index old_index before the last gs_sort (perm[new] =
 * old), and its (x, y, z, m) now (pos4 may be NULL).  The frame loop uses it
 * to follow the central black holes and drift the halo centres toward them
 * (main.js:449-457). */
int gs_locate(gs_sim* sim, uint32_t old_index, uint32_t* new_index,
              float pos4[4]);
/* FP64 reductions of the last step (GS_RED_*). */
int gs_reduce(gs_sim* sim, double out[GS_NRED]);
/* Summed stage times (GS_T_*) since the last reset; needs GS_FLAG_TIMING.
 * counts (may be NULL) receives the number of launches timed per stage. */
int gs_stage_times(gs_sim* sim, double ms[GS_NT], uint64_t counts[GS_NT],
                   int reset);
/* Interactions (monopole acceptances + leaf pairs) of the current tree and
 * positions, counted on the device by a counting walk; per-particle counts
 * go to per_particle when it is not NULL. */
int gs_count_interactions(gs_sim* sim, uint64_t* total, uint32_t* per_particle);

/* ---- slot pools: many galaxies stepped by the same launches -------------
 * No counterpart in the reference, which runs one galaxy per NBodyGPU.  A
 * pool holds nslots slots with fixed capacities; each slot holds 0 .. cap
 * particles with its own params and tree shape, and every stage (encode,
 * sort, build, refresh, walk) is one launch over all slots, replayed as one
 * captured graph per step.  Slot results equal those of the same galaxy on a
 * handle of its own.  The single-galaxy calls above act on slot 0 (gs_upload
 * also requires n == its capacity); gs_encode, gs_sort, gs_refresh, gs_force,
 * gs_step and gs_count_interactions act on every occupied slot. */
typedef struct gs_pool_config {
  uint32_t nslots;           /* 1 .. 256 */
  const uint32_t* capacity;  /* nslots capacities, each 1 .. 2^27 - 1 */
  const uint32_t* bucket;    /* nslots leaf capacities, NULL or 0 -> 16 */
  uint32_t flags;            /* GS_FLAG_* */
  int32_t device;            /* HIP device ordinal, -1 = current device */
  uint32_t walk_wg;          /* as gs_config.walk_wg */
  uint32_t radix_bits;       /* as gs_config.radix_bits */
  uint32_t reserved[4];
} gs_pool_config;

int gs_create_pool(const gs_pool_config* cfg, gs_sim** out);
/* Slot count and device bytes held by the pool. */
int gs_pool_info(gs_sim* sim, uint32_t* nslots, uint64_t* device_bytes);

/* Arena pools: a pool whose device state is sized by a byte budget instead
 * of fixed slot capacities.  No counterpart in the reference (a serving
 * loop's admission control).  Slots start empty; gs_slot_upload of n > 0
 * records takes a particle range of align_up(n, 1024) and a node range of
 * 2n + 4096 (mortonTree.js:88) from the arena's free lists (best fit), and
 * GS_E_FULL when none holds them (the budget is spent: the caller queues the
 * galaxy); n = 0 or gs_slot_clear gives the ranges back.  Launches cover the
 * occupied slots only (their tables are rebuilt, and the step graphs
 * captured again, before the first launch after a change), so a step costs
 * what the resident galaxies need whatever the budget.  Every slot call
 * above works on an arena slot (gs_slot_reduce until the next launch after
 * a slot changes); gs_upload_tree, gs_apply_permutation, stellar records
 * and the mortonTree.js calls below need a plain handle (GS_E_INVALID). */
#define GS_ARENA_PARTICLE_BYTES 112u /* per particle of a slot's range */
#define GS_ARENA_NODE_BYTES 36u      /* per node of a slot's range */
typedef struct gs_arena_config {
  uint64_t budget_bytes; /* device bytes the whole arena may hold */
  uint32_t nslots;       /* 1 .. 256 slots (galaxies resident at once) */
  uint32_t flags;        /* GS_FLAG_* */
  int32_t device;        /* HIP device ordinal, -1 = current device */
  uint32_t walk_wg;      /* as gs_config.walk_wg */
  uint32_t radix_bits;   /* as gs_config.radix_bits */
  uint32_t reserved[5];
} gs_arena_config;

int gs_create_arena(const gs_arena_config* cfg, gs_sim** out);
/* device_bytes: held by the arena (<= the budget); used_bytes: taken by the
 * occupied slots' ranges (GS_ARENA_PARTICLE_BYTES per particle of each
 * range, GS_ARENA_NODE_BYTES per node); largest_fit: the most particles one
 * more galaxy may have now.  Any output may be NULL. */
int gs_arena_info(gs_sim* sim, uint64_t* device_bytes, uint64_t* used_bytes,
                  uint64_t* largest_fit);
/* gs_upload into one slot: n (0 .. capacity) records; n = 0 empties it. */
int gs_slot_upload(gs_sim* sim, uint32_t slot, const float* particles,
                   uint64_t n);
int gs_slot_clear(gs_sim* sim, uint32_t slot);
int gs_slot_count(gs_sim* sim, uint32_t slot, uint32_t* n);
/* gs_set_params for one slot. */
int gs_slot_set_params(gs_sim* sim, uint32_t slot, const gs_params* p);
/* Leaf capacity and leaf-rule pool size of one slot (max_nodes <= the
 * slot's node pool, 2 capacity + 4096; 0 -> 2n + 4096 of each upload, as
 * new MortonTree(n) sizes it). */
int gs_slot_set_tree(gs_sim* sim, uint32_t slot, uint32_t bucket,
                     uint64_t max_nodes);
/* Per-slot forms of the downloads above. */
int gs_slot_download_particles(gs_sim* sim, uint32_t slot, float* dst,
                               uint64_t n);
int gs_slot_download_perm(gs_sim* sim, uint32_t slot, uint32_t* dst,
                          uint64_t n);
int gs_slot_tree_info(gs_sim* sim, uint32_t slot, uint64_t* num_nodes,
                      uint32_t* level_ranges, uint32_t* num_levels,
                      double* root_size);
int gs_slot_download_nodes(gs_sim* sim, uint32_t slot, uint32_t* dst,
                           uint64_t cap, uint64_t* num_nodes);
int gs_slot_download_energy(gs_sim* sim, uint32_t slot, float* dst,
                            uint64_t n);
int gs_slot_reduce(gs_sim* sim, uint32_t slot, double out[GS_NRED]);
int gs_slot_locate(gs_sim* sim, uint32_t slot, uint32_t old_index,
                   uint32_t* new_index, float pos4[4]);
int gs_slot_upload_stellar(gs_sim* sim, uint32_t slot, const float* stellar4,
                           uint64_t n);
int gs_slot_download_stellar(gs_sim* sim, uint32_t slot, float* dst4,
                             uint64_t n);

/* ---- mortonTree.js on host arrays (device-executed, copies in and out) --- */
/* computeMortonCodes(pos, N, stride, codes) -> rootSize (mortonTree.js:43). */
int gs_compute_morton_codes(gs_sim* sim, const float* pos, uint32_t n,
                            uint32_t stride, uint32_t* codes,
                            double* root_size);
/* radixSortIndices(codes, indices, scratch, counts, N) (mortonTree.js:68):
 * writes the stable sorted index array. */
int gs_radix_sort_indices(gs_sim* sim, const uint32_t* codes, uint32_t n,
                          uint32_t* perm);
/* new MortonTree(N).build(codes, pos, masses, N, rootSize)
 * (mortonTree.js:86-182): sorted codes, stride-3 positions, masses ->
 * nodes (64-byte layout, cap nodes), numNodes, levelRanges. */
int gs_tree_build(gs_sim* sim, const uint32_t* sorted_codes, const float* pos3,
                  const float* masses, uint32_t n, double root_size,
                  uint32_t* nodes, uint64_t cap, uint64_t* num_nodes,
                  uint32_t* level_ranges, uint32_t* num_levels);

/* ---- allocation-free form -------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): radixSortIndices (mortonTree.js:68-84, the same stable
 * permutation as gs_radix_sort_indices) on DEVICE arrays, with no handle and
 * no allocation.  codes: n Morton codes (30-bit values), perm: n outputs
 * (perm[new] = old), both device pointers on the current device; stream: a
 * hipStream_t of the current device, or NULL for the null stream.  Work is
 * queued on the stream (kernels only, no memset, no host synchronisation).
 * Its scratch is a static device array of the library, about 13 MB for
 * GS_NOALLOC_MAX_KEYS keys, held on each device from its first use until the
 * process ends.  n > GS_NOALLOC_MAX_KEYS returns GS_E_INVALID before any work
 * (n = 0 does nothing); a failed launch returns GS_E_DEVICE.  Every call on
 * one device shares that device's scratch: calls that overlap on a device
 * (other streams, other threads) race and must be serialised by the caller. */
#define GS_NOALLOC_MAX_KEYS (1u << 20)
int gs_radix_sort_indices_noalloc(const uint32_t* codes, uint32_t n, uint32_t* perm,
                                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GALAXY_HIP_H */
