// Internal device-side layout of a slot pool (gs_sim) and the kernel
// launchers shared by the translation units of libgalaxy_hip.so.
//
// A gs_sim holds S galaxy slots (S = 1 for a plain galaxy).  Slot s owns a
// fixed particle range [poff_s, poff_s + cap_s) (poff a multiple of
// GS_SLOT_ALIGN, so no walk workgroup spans two slots) and a node pool
// [noff_s, noff_s + ncap_s); it holds n_s <= cap_s active particles
// (n_s = 0: empty).  Every stage is ONE launch over all slots: particle
// kernels map blocks to slots with static tables, tree kernels map work to
// slots with an in-LDS prefix over the slots' device-side counts.  Sizes that
// change at run time (n_s, tree shape) live in device memory, so a captured
// step graph stays valid as jobs enter and leave slots.
//
// HBM layout (structure of arrays, all 16-byte aligned), P = sum of caps:
//   posm[2][P]  float4 (x, y, z, m)      ping-pong particle positions + mass
//   velp[2][P]  float4 (vx, vy, vz, pad) ping-pong velocities + gas flag
//   keys[2][P], vals[2][P]   u32         radix-sort ping-pong (code, index)
//   node[C]     32-byte record (one s_load_dwordx8 in the walk):
//               float4 cm   (comx, comy, comz, mass)
//               int4   info (opening threshold bits, first | lastChild, meta,
//                            skip); the threshold T of the node's size for
//                            the slot's theta^2: accept <=> d^2 > T
//                            (gs_open_threshold)
//               (lastChild and skip are byte offsets of 32-byte records,
//               i.e. 32 x the node index, so the walk loads a record without
//               an index shift; -1 ends the walk)
//               meta: bit 31 leaf, bits 27..30 depth, then
//                     leaf: bits 0..26 particle count
//                     internal: bits 0..3 child count, bits 8..15 octant mask
//               skip: the node that follows this node's subtree in the
//                     walk's pre-order (previous sibling, else the parent's
//                     skip; -1 ends the walk)
//   nsize[C]    float node size (read for the density of a particle's own
//               leaf, the s^2 < eps^2/4 test of trees that need it, downloads
//               and threshold rewrites)
// Node and particle indices inside a slot are slot-local (0-based).  A slot's
// node pool holds min(2 cap + 4096, 2^27) nodes (mortonTree.js:88).  Children
// of a BFS node are contiguous, so an internal node stores (lastChild, nChild,
// mask) instead of 8 slots (the walk descends to the last child first);
// gs_download_nodes expands to the 64-byte layout.  A childless internal node
// (only an uploaded tree can hold one) stores its skip as lastChild.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/galaxy_hip.h"
#include "arena.h"

#define GS_MAX_LEVEL 10   // mortonTree.js:26
#define GS_MAX_LEVELS 16  // level slots (NBodyGPU.js:171-174)

// status bits raised by kernels, checked by the host at sync points
#define GS_ST_NODE_POOL 0x1u
#define GS_ST_STACK 0x2u
#define GS_ST_SPIN 0x4u    // a look-back wait hit its bound (never expected)
#define GS_ST_LAYOUT 0x8u  // an uploaded tree's children are not contiguous
#define GS_SPIN_LIMIT (1u << 24)

#define GS_SLOT_ALIGN 1024  // slot particle ranges start on multiples of this
#define GS_MAX_SLOTS 256
#define GS_CHUNK 4096       // particles per block in bbox / encode

// Params uniform (nbody_bh.wgsl:19-42) in f32 plus derived per-step values.
struct DevParams {
  float dt, G, eps2;
  uint32_t n;
  float theta, halo_mass, halo_a, gas_drag;
  float time, aging_rate, sf_density, sf_eff;
  float sf_dt;
  uint32_t halo_count;
  float theta2;     // f32(theta * theta) as the WGSL evaluates it
  float drag_damp;  // f32 exp(-gasDrag * dt), evaluated once on the host
  // bits of the smallest size with f32(size * size) >= f32(0.25 * eps2): a
  // node is accepted by s^2 < eps^2/4 iff |size| bits are below this
  uint32_t qsize_bits;
  uint32_t pad_[3];
  float4 halo_c1, halo_c2;
};

// Sort digit widths (8 -> 4 passes, 10 -> 3 passes over the 30 code bits)
// and tiling.
#define GS_HIST_WORDS 3072  // max over widths of passes x radix (4 x 256, 3 x 1024)
__host__ __device__ constexpr int gs_sort_passes(int bits) { return (30 + bits - 1) / bits; }
#define GS_SORT_BLOCK 256
#define GS_SORT_ITEMS 16
#define GS_SORT_TILE (GS_SORT_BLOCK * GS_SORT_ITEMS)
// the sort's tile table is built from the encode chunks: a sort tile is one
// encode chunk
static_assert(GS_SORT_TILE == GS_CHUNK, "sort tiles and encode chunks must coincide");

#define GS_BUILD_TILE 256

// Static layout and active size of one slot.
struct DevSlot {
  uint32_t n;          // active particles (0 = empty slot)
  uint32_t cap;        // particle capacity
  uint32_t poff;       // first particle (multiple of GS_SLOT_ALIGN)
  uint32_t noff;       // first node of the slot's pool
  uint32_t ncap;       // node pool size
  uint32_t bucket;     // leaf capacity (mortonTree.js:25)
  uint64_t max_nodes;  // leaf-rule pool size (mortonTree.js:88,127)
  uint32_t chunk0;     // first bbox/encode block of the slot
  uint32_t tile0;      // first sort tile of the slot
  uint32_t btoff;      // first build look-back word of the slot
  uint32_t wblock0;    // first walk block of the slot
  uint32_t nwblocks;   // walk blocks of the slot
  uint32_t nchunks;    // bbox/encode blocks = sort tiles of the slot (GS_CHUNK each)
  uint32_t pad_[2];
};

// Device-resident scalars of one slot.
struct DevScalars {
  uint64_t min_bits[3];  // bbox minimum (f32 widened to binary64 bits)
  uint64_t s_bits;       // 1024 / size  (binary64 bits)
  uint64_t size_bits;    // rootSize     (binary64 bits)
  float min_f32[3];      // bbox minimum (f32), for the encode's f32 fast path
  float s32;             // f32(1024 / size)
  uint32_t nfinite;      // particles that entered the bbox
  uint32_t num_nodes;
  uint32_t num_levels;
  float min_s2;          // smallest size * size over the tree's nodes (NaN ignored)
  // opening thresholds of each depth's node size for the slot's theta^2
  // (gs_open_threshold), written by the root kernel for the build
  uint32_t level_thr[GS_MAX_LEVELS + 1];
  uint32_t level_start[GS_MAX_LEVELS + 1];
  uint32_t level_count[GS_MAX_LEVELS + 1];
  float level_size[GS_MAX_LEVELS + 1];
  uint32_t sort_hist[GS_HIST_WORDS];  // this slot's digit histograms [pass][digit]
};

// Pool-wide bookkeeping.
struct DevBatch {
  uint32_t status;       // GS_ST_* bits
  uint32_t tickets[32];  // [0..3] sort passes, [4..14] build levels
  uint64_t interactions;
  uint32_t big_count;    // large leaves queued for the wave-wide refresh pass
  uint32_t located;      // gs_locate's answer
};

// leaves larger than this are refreshed by a whole wavefront each
#define GS_BIG_LEAF 32

struct __align__(32) NodeRec {
  float4 cm;
  int4 info;
};

// f32(size * size) as the walk's opening test evaluates it
// (nbody_bh.wgsl:99), one rounding, as bits
__host__ __device__ __forceinline__ int gs_s2_bits(float size) {
  const float s2 = size * size;
  union {
    float f;
    int i;
  } u;
  u.f = s2;
  return u.i;
}

// The walk's opening test s2 < f32(theta2 * d2) (nbody_bh.wgsl:97-112) as one
// compare d2 > T per node.  accept(x) = s2 < f32(theta2 * x) is monotone in x
// (rounding is monotone; theta2 >= 0 or NaN, s2 >= 0 or NaN), and false at
// x = +0, so T = max{x in [+0, +inf] : !accept(x)} (bits returned) gives
// accept(x) <=> x > T for every x >= 0, and false for NaN on both sides.
// Non-negative floats order as their bit patterns: bisection over the bits.
__host__ __device__ inline bool gs_accepts(float s2, float theta2, uint32_t xbits) {
  union {
    uint32_t u;
    float f;
  } x;
  x.u = xbits;
  return s2 < theta2 * x.f;
}
__host__ __device__ inline uint32_t gs_open_threshold(float s2, float theta2) {
  const uint32_t inf = 0x7f800000u;
  if (!gs_accepts(s2, theta2, inf)) return inf;  // never accepted
  uint32_t lo = 0u, hi = inf;                    // !accept(lo), accept(hi)
  while (hi - lo > 1u) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (gs_accepts(s2, theta2, mid)) hi = mid;
    else lo = mid;
  }
  return lo;
}
// smallest size bits with f32(size * size) >= q: a node passes s^2 < q
// (q = eps^2/4) iff its |size| bits are below it (NaN sizes never pass)
__host__ __device__ inline uint32_t gs_quarter_eps_size(float q) {
  if (!(q > 0.0f)) return 0u;  // q NaN or <= 0: no size passes
  union {
    uint32_t u;
    float f;
  } x;
  uint32_t lo = 0u, hi = 0x7f800000u;  // (lo * lo) < q, (hi * hi) >= q
  while (hi - lo > 1u) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    x.u = mid;
    if (x.f * x.f >= q) hi = mid;
    else lo = mid;
  }
  return hi;
}

// node meta word
#define GS_META_LEAF 0x80000000u
#define GS_MAX_N ((1u << 27) - 1)  // a leaf count must fit the 27-bit field
// a slot's node pool holds at most this many records (2^26 x 32 bytes =
// 2 GiB), so node links are non-negative 32-bit byte offsets; slots above
// 2^25 - 2048 particles get this cap instead of 2N + 4096 records
#define GS_NODE32_MAX (1u << 26)
// node index <-> node link (byte offset of the 32-byte record)
__host__ __device__ __forceinline__ int32_t gs_link(uint32_t node) { return (int32_t)(node << 5); }
__host__ __device__ __forceinline__ uint32_t gs_node(int32_t link) { return (uint32_t)link >> 5; }
__host__ __device__ __forceinline__ uint32_t gs_meta_leaf(uint32_t depth, uint32_t count) {
  return GS_META_LEAF | (depth << 27) | count;
}
__host__ __device__ __forceinline__ uint32_t gs_meta_internal(uint32_t depth, uint32_t mask,
                                                             uint32_t nc) {
  return (depth << 27) | (mask << 8) | nc;
}
// a node emitted by its parent, not yet classified: depth + particle count
__host__ __device__ __forceinline__ uint32_t gs_meta_open(uint32_t depth, uint32_t count) {
  return (depth << 27) | count;
}
__host__ __device__ __forceinline__ bool gs_meta_is_leaf(uint32_t m) { return (m >> 31) != 0; }
__host__ __device__ __forceinline__ uint32_t gs_meta_depth(uint32_t m) { return (m >> 27) & 0xfu; }
__host__ __device__ __forceinline__ uint32_t gs_meta_count(uint32_t m) { return m & 0x7ffffffu; }
__host__ __device__ __forceinline__ uint32_t gs_meta_nc(uint32_t m) { return m & 0xfu; }
__host__ __device__ __forceinline__ uint32_t gs_meta_mask(uint32_t m) { return (m >> 8) & 0xffu; }

// Block-wide exclusive prefix over per-slot values (thread t holds slot t's
// value; nslots <= GS_MAX_SLOTS <= blockDim.x, blockDim.x a multiple of 64);
// result in s_pre[0..nslots], s_pre[nslots] = total.  Wavefront scans of the
// first GS_MAX_SLOTS / 64 waves, then the wave totals.  All threads must call
// it; s_pre needs GS_MAX_SLOTS + 1 + GS_MAX_SLOTS / 64 words.
__device__ __forceinline__ void gs_slot_prefix(uint32_t v, uint32_t nslots, uint32_t* s_pre) {
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  uint32_t* s_wave = s_pre + GS_MAX_SLOTS + 1;
  uint32_t x = t < nslots ? v : 0u;
  if (t < GS_MAX_SLOTS) {
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[w] = x;
  }
  __syncthreads();
  if (t < GS_MAX_SLOTS) {
    uint32_t off = 0;
    for (uint32_t k = 0; k < w; k++) off += s_wave[k];
    if (t < nslots) s_pre[t + 1] = off + x;
    if (t == 0) s_pre[0] = 0;
  }
  __syncthreads();
}

// slot k with s_pre[k] <= x < s_pre[k + 1] (s_pre ascending, x < s_pre[n])
__device__ __forceinline__ uint32_t gs_find_slot(const uint32_t* s_pre, uint32_t nslots,
                                                 uint32_t x) {
  uint32_t lo = 0, hi = nslots;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s_pre[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ping-pong buffer that holds the sorted (key, index) pairs after a sort
inline int gs_sorted_buf(const gs_sim* s);

struct StageTimer {
  int stage;
  hipEvent_t a, b;
};

// Host mirror of one slot.
struct HostSlot {
  DevSlot d;
  bool have_params = false, have_tree = false, have_codes = false, have_force = false,
       have_step = false;
  // node COM/mass words hold sums of particles (a refresh, a tree upload or a
  // step ran since the last topology build, which writes only the info words)
  bool have_com = false;
  bool fixed_max_nodes = false;  // else the leaf rule's pool follows n: 2n + 4096
  // the per-block reduction partials of the last step are where this slot's
  // walk blocks put them (an arena's launch tables were not rebuilt since)
  bool have_red = false;
};

struct gs_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t flags = 0, walk_wg = 256;
  uint32_t radix_bits = 8;  // sort digit width: 8 or 10
  uint32_t nslots = 1;
  uint64_t ptot = 0;         // particle records (sum of aligned slot capacities, or the arena's)
  uint64_t ntot = 0;         // node records (sum of node pools, or the arena's)
  uint32_t chunks = 0;       // bbox/encode blocks
  uint32_t sort_tiles = 0;   // sum over slots of capacity tiles
  uint32_t build_words = 0;  // look-back words per build buffer
  uint32_t walk_blocks = 0;
  uint32_t max_ncap = 0;
  // allocated sizes of the per-block arrays and tables (the launch sizes above
  // stay within them; equal to them for a pool of fixed slots)
  uint32_t cap_chunks = 0, cap_build_words = 0, cap_walk_blocks = 0;
  uint64_t node64_cap = 0;  // nodes the download / upload staging holds

  // arena pools: slot ranges come from these free lists (particles in units
  // of GS_SLOT_ALIGN, nodes in nodes) when a slot is filled and go back when
  // it is emptied; the launch tables cover the occupied slots and are rebuilt
  // before the next launch after a change (layout_dirty)
  bool arena = false;
  bool layout_dirty = false;
  gs::RangeList pfree, nfree;

  float4* posm[2] = {nullptr, nullptr};
  float4* velp[2] = {nullptr, nullptr};
  int cur = 0;
  // stellar records (mass, birth Myr, seed, dust), allocated on first upload;
  // their ping-pong flips only when the particles are permuted
  float4* stel[2] = {nullptr, nullptr};
  int stel_cur = 0;
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* codes_raw = nullptr;
  uint32_t* sort_status = nullptr;  // passes x sort_tiles x radix (<= GS_HIST_WORDS x tiles)
  uint64_t* build_status[2] = {nullptr, nullptr};
  float* bbox_part = nullptr;       // chunks x 8
  NodeRec* nodes = nullptr;         // ntot records
  float* nsize = nullptr;           // ntot node sizes
  // per slot: nonzero when the tree may hold a node of zero mass (set by the
  // refresh and by tree uploads; a fresh handle starts at 1)
  uint32_t* massless = nullptr;
  float* energy = nullptr;
  float4* acc = nullptr;
  float* dens = nullptr;
  uint32_t* inter = nullptr;
  double* red_part = nullptr;       // walk blocks x GS_NRED
  double* red_out = nullptr;        // GS_NRED
  uint32_t* node64 = nullptr;       // download / upload staging (largest pool)
  DevSlot* slots = nullptr;         // nslots
  DevScalars* sc = nullptr;         // nslots
  DevParams* prm = nullptr;         // nslots
  DevBatch* batch = nullptr;
  uint2* chunk_tab = nullptr;       // chunks: (slot, local start)
  uint32_t* tile_tab = nullptr;     // sort_tiles: slot << 16 | local tile
  uint32_t* wblock_slot = nullptr;  // walk_blocks: slot of each walk block
  // passed to the next walk launch (direct or captured): the tree was built
  // on the device from the current positions of every slot (set only for the
  // walk of a step that rebuilds)
  bool walk_fresh = false;
  uint2* big_leaf = nullptr;        // (slot, node) of large leaves
  uint32_t big_cap = 0;             // ptot / GS_BIG_LEAF + 64 entries

  std::vector<HostSlot> hs;
  std::vector<DevParams> host_prm;
  uint64_t step_count = 0;

  std::vector<StageTimer> timers;
  std::vector<hipEvent_t> spare_events;  // timing events reused across steps
  bool capturing = false;                 // a step graph is being captured
  double t_ms[GS_NT] = {0};
  uint64_t t_cnt[GS_NT] = {0};

  // step graphs [particle parity][stellar parity][sort][part]: part 0 the
  // whole step, 1 rebuild + refresh, 2 walk (timing handles replay 1 then 2)
  hipGraphExec_t graph[2][2][2][3] = {};
  bool graph_dirty = true;

  std::string err;
};

inline int gs_sorted_buf(const gs_sim* s) { return gs_sort_passes((int)s->radix_bits) & 1; }
inline uint64_t gs_sort_status_words(const gs_sim* s) {
  return (uint64_t)gs_sort_passes((int)s->radix_bits) * s->sort_tiles * (1u << s->radix_bits);
}

// Device buffers of one radix sort without the particle gather: a handle's
// own, or the static scratch of the allocation-free entry point.
struct SortBufs {
  uint32_t* keys[2];           // pass p reads keys[p & 1], writes the other
  uint32_t* vals[2];
  uint32_t* status;            // passes x tiles x radix look-back words
  DevBatch* batch;             // tile tickets
  const DevSlot* slots;
  DevScalars* sc;              // digit histograms (accumulated by the hist pass)
  const uint32_t* tile_tab;
  uint32_t tiles;              // grid of each pass
  uint32_t radix_bits;         // 8 or 10
  hipStream_t stream;
};

// ---- launchers (stream-ordered; return hipError_t of the launch) ----------
hipError_t gs_launch_bbox(gs_sim* s, const float4* posm);
hipError_t gs_launch_root(gs_sim* s);
hipError_t gs_launch_encode(gs_sim* s, const float4* posm, bool keep_raw);
hipError_t gs_launch_reset(gs_sim* s);
hipError_t gs_launch_hist(gs_sim* s);  // histograms of host-supplied keys (slot 0)
// slot 0 of b: keys[0][i] = src[i], vals[0][i] = i and the digit histograms,
// over the slot's capacity (src may be keys[0])
hipError_t gs_launch_hist_bufs(const SortBufs& b, const uint32_t* src, uint32_t cap);
hipError_t gs_launch_sort(gs_sim* s, bool gather);
hipError_t gs_launch_sort_bufs(const SortBufs& b);  // every pass, no gather
// top_from_hist: the digit histograms belong to the sorted keys (an encode
// counted them), so levels 0 and 1 come from them
hipError_t gs_launch_build(gs_sim* s, bool top_from_hist);
hipError_t gs_launch_refresh(gs_sim* s, const float4* posm);
hipError_t gs_launch_walk(gs_sim* s, bool integrate, bool count);
hipError_t gs_launch_reduce_final(gs_sim* s, uint32_t slot, double* out_dev);
hipError_t gs_launch_node64(gs_sim* s, uint32_t slot);
hipError_t gs_launch_import64(gs_sim* s, uint32_t slot, const uint32_t* words, const int2* links,
                              uint32_t num);
// rewrite the opening thresholds of one slot's tree for its current theta^2
hipError_t gs_launch_rethreshold(gs_sim* s, uint32_t slot);
hipError_t gs_launch_gather(gs_sim* s, uint32_t slot, const uint32_t* perm);
hipError_t gs_launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t st);
hipError_t gs_launch_star_formation(gs_sim* s);  // every slot, in place
// 32-byte host records <-> (posm, velp) of one slot, staged contiguously in
// the slot's range of the non-current ping-pong buffers (stream-ordered)
hipError_t gs_upload_records(gs_sim* s, uint32_t slot, const float* host, uint32_t n);
hipError_t gs_download_records(gs_sim* s, uint32_t slot, float* host, uint32_t n);
hipError_t gs_download_positions(gs_sim* s, uint32_t slot, float* host4, uint32_t n);
hipError_t gs_launch_locate(gs_sim* s, uint32_t slot, uint32_t old, uint32_t* out_dev);
