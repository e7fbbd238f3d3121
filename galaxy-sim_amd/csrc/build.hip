// Breadth-first bucketed octree build — the device form of MortonTree.build
// (mortonTree.js:107-162), one kernel per depth level, bit-identical node
// order and contents.
//
// Level d holds nodes [start_d, start_d + count_d), each a segment [lo, hi)
// of the Morton-sorted particles.  Per node (mortonTree.js:119-156):
//   leaf   iff count <= bucket || d >= 10 || idx + 9 > maxNodes
//          (numNodes after its own increment is idx + 1, so the pool rule is
//          a function of the node index alone)
//   else   children = the non-empty runs of the 3-bit digit at shift 27 - 3d,
//          found by binary search (codes in a node share all higher digits,
//          so the digit is monotone over [lo, hi)); 8 lanes search the 8
//          octant boundaries of one node in parallel.
// Children are numbered in BFS order = parent order x octant order, i.e. an
// exclusive scan of child counts over the level; the scan runs in the same
// kernel with a decoupled look-back over 256-node tiles taken in ticket
// order.  A node's children are therefore contiguous: internal nodes store
// (lastChild, nChild, octant mask) and the 64-byte layout is produced on
// download.  The last tile of a level writes count_{d+1}.
//
// One launch per level serves every slot of the pool: a block's ticket is
// mapped to (slot, tile of the slot) through a prefix over the slots' tile
// counts at this level; each slot has its own look-back chain and indices.
#include "gs_internal.h"

namespace {

// phase 1 gives each node of a tile 8 lanes (one per octant) and covers the
// tile in rows of 32 nodes, and gs_slot_prefix needs a thread per slot: the
// tile is exactly 256 nodes, one per thread
static_assert(GS_BUILD_TILE == 256, "build phase 1 maps 8 lanes x 32 nodes per row");
static_assert(GS_BUILD_TILE >= GS_MAX_SLOTS, "slot prefix needs a thread per slot");

constexpr uint64_t kAgg = 1ull << 32;
constexpr uint64_t kPre = 2ull << 32;

__device__ __forceinline__ uint64_t ld_agent64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Nodes per tile: a level whose GS_BUILD_TILE-node tiles over all slots
// number at most kFewTiles (a few blocks' work on a GPU of 256 CUs) takes
// 32-node tiles (one row of 8 lanes x 32 nodes) instead, which spreads its
// searches over 8 times as many blocks.  Small tiles need one look-back
// word each, and a slot has ceil(ncap / GS_BUILD_TILE) of them.
constexpr uint32_t kSmallTile = 32;
constexpr uint32_t kFewTiles = 64;
__device__ __forceinline__ uint32_t tile_nodes(uint32_t cnt, uint32_t ncap, bool small) {
  const uint32_t words = (ncap + GS_BUILD_TILE - 1) / GS_BUILD_TILE;
  return small && (cnt + kSmallTile - 1) / kSmallTile <= words ? kSmallTile
                                                               : (uint32_t)GS_BUILD_TILE;
}

// Every loaded value passes through one empty asm: every load of the round
// issues before the first wait.  Left alone, the compiler turns "select
// between two loaded values" into "load at the selected address" (which
// waits for one step before the next step's load) and interleaves the waits
// of one search with the loads of the next.
#define GS_V3(r) "+v"(v[r][0]), "+v"(v[r][1]), "+v"(v[r][2])
__device__ __forceinline__ void loads_in_flight(uint32_t (&v)[8][3]) {
  asm("" : GS_V3(0), GS_V3(1), GS_V3(2), GS_V3(3), GS_V3(4), GS_V3(5), GS_V3(6), GS_V3(7));
}
#undef GS_V3
#define GS_V5(a) "+v"(v[0][a]), "+v"(v[0][a + 1]), "+v"(v[0][a + 2]), "+v"(v[0][a + 3]), \
                 "+v"(v[0][a + 4])
__device__ __forceinline__ void loads_in_flight(uint32_t (&v)[1][15]) {
  asm("" : GS_V5(0), GS_V5(5), GS_V5(10));
}
#undef GS_V5

// Phase 1 of a tile: octant boundaries b_0..b_7 (+ hi in s_b[j][8]) of its
// nodes, 8 lanes per node (lane k of a node finds the first code with digit
// >= k below the node's common prefix: a lower_bound over the node's
// particles, whose digit is monotone).  Each thread owns one octant of R
// nodes (rows of 32 nodes) and advances their R searches in lockstep, STEPS
// bisection steps per round: the probes of all 2^STEPS - 1 decision-tree
// nodes of every search are loaded together, so a round costs one memory
// latency for STEPS steps of R searches.  Every probe lies inside the
// node's particle range; a finished search (len 0) probes index 0, which
// exists whenever a lane is still in the loop (its node is non-empty).
template <int R, int STEPS>
__device__ __forceinline__ void octant_bounds(const NodeRec* __restrict__ nodes,
                                              const uint32_t* __restrict__ codes, uint32_t start,
                                              uint32_t count, uint32_t first_local, uint32_t n,
                                              uint32_t bucket, uint64_t max_nodes, int d,
                                              uint32_t (*s_b)[9], int4* s_info) {
  constexpr int P = (1 << STEPS) - 1;
  const int shift = 27 - 3 * d;
  const uint32_t k = threadIdx.x & 7;
  uint32_t sfirst[R], slen[R];
  bool more = false;
  // the R node records are loaded together (rows past the level's end load
  // the level's first node); the first code of each node (nodes without a
  // valid range load code 0 of a non-empty slot) is consumed only after the
  // first round's probes have issued, so set-up and first round share a
  // memory latency
  int4 infr[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t local = first_local + (uint32_t)r * 32 + (threadIdx.x >> 3);
    infr[r] = nodes[start + (local < count ? local : 0u)].info;
  }
  uint32_t c0[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t lo = (uint32_t)infr[r].y;
    c0[r] = 0;
    if (n > 0) c0[r] = codes[lo < n ? lo : 0u];  // n is uniform over the block
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t j = (uint32_t)r * 32 + (threadIdx.x >> 3);
    const uint32_t local = first_local + j;
    sfirst[r] = 0;
    slen[r] = 0;
    if (local < count) {
      const uint32_t idx = start + local;
      const int4 inf = infr[r];
      const uint32_t lo = (uint32_t)inf.y, cnt = gs_meta_count((uint32_t)inf.z);
      const bool leaf = cnt <= bucket || d >= GS_MAX_LEVEL || (uint64_t)idx + 9 > max_nodes;
      sfirst[r] = lo;
      if (!leaf && k > 0 && lo < n && cnt <= n - lo) {
        slen[r] = cnt;
        more = true;
      }
      if (k == 0) {
        s_b[j][8] = leaf ? 0xffffffffu : lo + cnt;
        s_info[j] = inf;  // phase 2 reads the record from here
      }
    }
  }
  while (more) {
    uint32_t v[R][P];
#pragma unroll
    for (int r = 0; r < R; r++) {
      // (first, len) of every decision-tree node in heap order
      uint32_t pf[P], pl[P];
      pf[0] = sfirst[r];
      pl[0] = slen[r];
#pragma unroll
      for (int q = 0; q < P; q++) {
        const uint32_t l = pl[q], h = l >> 1;
        v[r][q] = codes[l > 0 ? pf[q] + h : 0u];
        if (2 * q + 2 < P) {
          pf[2 * q + 1] = pf[q];
          pl[2 * q + 1] = h;
          pf[2 * q + 2] = pf[q] + h + 1;
          pl[2 * q + 2] = l > 0 ? l - h - 1 : 0u;
        }
      }
    }
    loads_in_flight(v);
    // the STEPS steps as selects (a finished search keeps len 0 and its first)
    more = false;
#pragma unroll
    for (int r = 0; r < R; r++) {
      // v[r] holds the remaining decision tree in heap order; each step
      // moves the taken child's subtree to the front (level lv's half to
      // level lv - 1, in place), so every index is a constant
      // the target: the node's common prefix above this level's digit, then k
      const uint32_t tgt = ((c0[r] >> (shift + 3)) << (shift + 3)) | (k << shift);
      uint32_t f = sfirst[r], l = slen[r];
#pragma unroll
      for (int st = 0; st < STEPS; st++) {
        const uint32_t h = l >> 1;
        const bool right = l > 0 && v[r][0] < tgt;
        f = right ? f + h + 1 : f;
        l = right ? l - h - 1 : h;
#pragma unroll
        for (int lv = 1; lv < STEPS - st; lv++) {
#pragma unroll
          for (int i = 0; i < (1 << (lv - 1)); i++) {
            const int src = (1 << lv) - 1 + i;
            v[r][(1 << (lv - 1)) - 1 + i] = right ? v[r][src + (1 << (lv - 1))] : v[r][src];
          }
        }
      }
      sfirst[r] = f;
      slen[r] = l;
      more |= l > 0;
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t j = (uint32_t)r * 32 + (threadIdx.x >> 3);
    if (first_local + j < count) s_b[j][k] = sfirst[r];
  }
}

__global__ __launch_bounds__(GS_BUILD_TILE) void build_level_kernel(
    const uint32_t* __restrict__ keys, NodeRec* __restrict__ pool, float* __restrict__ nsize_pool,
    const DevSlot* __restrict__ slots, DevScalars* scs, DevBatch* batch, uint32_t nslots, int d,
    uint64_t* status, uint64_t* status_clear, uint32_t clear_n) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_b[GS_BUILD_TILE][9];
  __shared__ int4 s_info[GS_BUILD_TILE];
  __shared__ uint32_t s_wsum[GS_BUILD_TILE / 64];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_pre[GS_MAX_SLOTS + 1 + GS_MAX_SLOTS / 64];

  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t k = gtid; k < clear_n; k += gridDim.x * blockDim.x) status_clear[k] = 0;

  // nodes, then tiles per slot at this level; slots without nodes here end
  // their tree
  uint32_t cnt = 0, ncap = 0;
  if (threadIdx.x < nslots) {
    cnt = scs[threadIdx.x].level_count[d];
    ncap = slots[threadIdx.x].ncap;
    if (cnt == 0 && blockIdx.x == 0) {
      scs[threadIdx.x].level_start[d + 1] = scs[threadIdx.x].level_start[d];
      scs[threadIdx.x].level_count[d + 1] = 0;
    }
  }
  gs_slot_prefix((cnt + GS_BUILD_TILE - 1) / GS_BUILD_TILE, nslots, s_pre);
  // the level's tile total, the same in every block (read before the next
  // prefix, which rewrites s_pre only after a barrier every thread reaches)
  const bool small = s_pre[nslots] <= kFewTiles;
  if (small) {
    const uint32_t my_tiles = cnt == 0 ? 0u : (cnt - 1) / tile_nodes(cnt, ncap, true) + 1;
    gs_slot_prefix(my_tiles, nslots, s_pre);
  }
  const uint32_t total_tiles = s_pre[nslots];
  // the grid is sized for the largest level: blocks beyond this level's tile
  // count leave without touching the ticket counter (tickets keep the
  // look-back order; the blocks that stay still take every tile)
  if (blockIdx.x >= total_tiles) return;
  // when the grid covers every tile, the blocks that stay are exactly
  // total_tiles and their first tickets are 0 .. total_tiles - 1: each block
  // builds one tile and leaves without asking for another
  const bool one_tile = total_tiles <= gridDim.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;

  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(&batch->tickets[4 + d], 1u);
    __syncthreads();
    const uint32_t ticket = s_tile;
    if (ticket >= total_tiles) break;
    const uint32_t slot = gs_find_slot(s_pre, nslots, ticket);
    const uint32_t tile = ticket - s_pre[slot];
    const uint32_t ntiles = s_pre[slot + 1] - s_pre[slot];
    const DevSlot sl = slots[slot];
    DevScalars* sc = scs + slot;
    NodeRec* nodes = pool + sl.noff;
    float* nsize = nsize_pool + sl.noff;
    const uint32_t* codes = keys + sl.poff;
    const uint32_t n = sl.n, bucket = sl.bucket, cap = sl.ncap;
    const uint64_t max_nodes = sl.max_nodes;
    uint64_t* st_base = status + sl.btoff;
    const uint32_t count = sc->level_count[d];
    const uint32_t start = sc->level_start[d];
    const uint32_t next_start = start + count;
    // records hold the opening threshold of their depth's size (the root
    // kernel's table for the slot's theta^2); the sizes go to nsize
    const int s2_d = (int)sc->level_thr[d];
    const float size_c = sc->level_size[d + 1];
    const int s2_c = (int)sc->level_thr[d + 1];

    const uint32_t tn = tile_nodes(count, cap, small);
    if (tn == GS_BUILD_TILE)
      octant_bounds<GS_BUILD_TILE / 32, 2>(nodes, codes, start, count, tile * tn, n, bucket,
                                            max_nodes, d, s_b, s_info);
    else
      octant_bounds<1, 4>(nodes, codes, start, count, tile * tn, n, bucket, max_nodes, d, s_b,
                          s_info);
    __syncthreads();

    // phase 2: one node per thread
    const uint32_t local = tile * tn + threadIdx.x;
    const bool valid = threadIdx.x < tn && local < count;
    const uint32_t idx = start + local;
    uint32_t mask = 0, nchild = 0;
    bool leaf = true;
    if (valid) {
      const uint32_t hi = s_b[threadIdx.x][8];
      leaf = hi == 0xffffffffu;
      if (!leaf) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const uint32_t b0 = s_b[threadIdx.x][k];
          const uint32_t b1 = k < 7 ? s_b[threadIdx.x][k + 1] : hi;
          if (b1 > b0) mask |= 1u << k;
        }
        nchild = __popc(mask);
      }
    }
    // block exclusive scan of child counts
    uint32_t x = nchild;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_wsum[w] = x;
    __syncthreads();
    uint32_t woff = 0, tile_total = 0;
    for (int k = 0; k < GS_BUILD_TILE / 64; k++) {
      if (k < w) woff += s_wsum[k];
      tile_total += s_wsum[k];
    }
    const uint32_t excl_local = woff + x - nchild;

    if (w == 0) {
      // decoupled look-back, 64 predecessors per load round: lane l reads
      // tile j - l; the nearest inclusive prefix (kPre) ends the walk once
      // every tile before it in the window has published its aggregate
      uint64_t* st = st_base + tile;
      uint64_t prefix = 0;
      if (tile == 0) {
        if (lane == 0) st_agent64(st, kPre | tile_total);
      } else {
        if (lane == 0) st_agent64(st, kAgg | tile_total);
        int32_t j = (int32_t)tile - 1;
        uint32_t spins = 0;
        for (;;) {
          const int32_t jj = j - lane;
          // tile 0 always publishes kPre, so lanes below it are never summed
          const uint64_t v = jj >= 0 ? ld_agent64(st_base + jj) : kPre;
          const uint64_t f = v & 0xffffffff00000000ull;
          const uint64_t pre = __ballot(f == kPre), none = __ballot(f == 0);
          const uint64_t upto = pre ? (pre & (0 - pre)) * 2 - 1 : ~0ull;  // lanes <= first kPre
          if (none & upto) {
            if (++spins > GS_SPIN_LIMIT) {  // bounded wait: flag and stop, never hang
              if (lane == 0) atomicOr(&batch->status, GS_ST_SPIN);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint64_t x = ((upto >> lane) & 1) ? (v & 0xffffffffull) : 0;
#pragma unroll
          for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
          prefix += x;
          if (pre) break;
          j -= 64;
        }
        if (lane == 0) st_agent64(st, kPre | (prefix + tile_total));
      }
      if (lane == 0) s_prefix = prefix;
      if (lane == 0 && tile == ntiles - 1) {
        uint64_t total = prefix + tile_total;
        uint64_t cnt_next = total;
        if ((uint64_t)next_start + total > cap) {
          atomicOr(&batch->status, GS_ST_NODE_POOL);
          cnt_next = cap > next_start ? cap - next_start : 0;
        }
        sc->level_start[d + 1] = next_start;
        sc->level_count[d + 1] = (uint32_t)cnt_next;
        sc->num_nodes = next_start + (uint32_t)cnt_next;
        sc->num_levels = (uint32_t)d + 1;
      }
    }
    __syncthreads();

    if (valid) {
      const int4 inf = s_info[threadIdx.x];
      const int32_t skip = inf.w;
      if (leaf) {
        nodes[idx].info = make_int4(s2_d, inf.y,
                                     (int)gs_meta_leaf((uint32_t)d, gs_meta_count((uint32_t)inf.z)),
                                     skip);
      } else {
        const uint32_t fc = next_start + (uint32_t)s_prefix + excl_local;
        const uint32_t hi = s_b[threadIdx.x][8];
        uint32_t c = fc;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          if (mask & (1u << k)) {
            const uint32_t b0 = s_b[threadIdx.x][k];
            const uint32_t b1 = k < 7 ? s_b[threadIdx.x][k + 1] : hi;
            if (c < cap) {
              // pre-order with octant 7 first: after child c comes c - 1
              nsize[c] = size_c;
              nodes[c].info = make_int4(s2_c, (int)b0,
                                         (int)gs_meta_open((uint32_t)d + 1, b1 - b0),
                                         c > fc ? gs_link(c - 1) : skip);
            }
            c++;
          }
        }
        nodes[idx].info = make_int4(s2_d, gs_link(fc + nchild - 1),  // last child
                                     (int)gs_meta_internal((uint32_t)d, mask, nchild), skip);
      }
    }
    if (one_tile) break;
    __syncthreads();
  }
}

// ---- levels 0 and 1 from the encode kernel's digit histogram -------------
// The encode kernel counts the digits of every key for every radix pass.  The
// top pass's digits hold key bits 24..29 (8-bit digits: bits 24..31, whose
// two highest bits are 0 for every 30-bit code; the padding keys all fall in
// digit 255; 10-bit digits: bits 20..29, padding in digit 1023), so the
// number of keys per value of their two highest 3-bit digits -- 64 cells --
// follows from the histogram, and a prefix over the cells gives the sorted
// position of every cell's first key.  The root's octant boundaries are the
// prefixes at cells 8k and those of the level-1 node of octant k the
// prefixes at cells 8k + m: exactly the lower_bound results of
// build_level_kernel for d = 0 and 1 (mortonTree.js:130-150), found without
// reading a key.  One wavefront per slot writes the records, level ranges and
// counts that those two launches write; blocks past the slots only clear.
template <int BITS>
__global__ __launch_bounds__(64) void build_top_kernel(NodeRec* __restrict__ pool,
                                                       float* __restrict__ nsize_pool,
                                                       const DevSlot* __restrict__ slots,
                                                       DevScalars* scs, DevBatch* batch,
                                                       uint32_t nslots, uint64_t* status_clear,
                                                       uint32_t clear_n) {
  __shared__ uint32_t s_ex[65];
  // the level-2 launch's look-back words
  for (uint32_t k = blockIdx.x * 64u + threadIdx.x; k < clear_n; k += gridDim.x * 64u)
    status_clear[k] = 0;
  if (blockIdx.x >= nslots) return;
  DevScalars* sc = scs + blockIdx.x;
  // an empty slot: the root kernel left every level empty, as the level
  // launches would
  if (sc->level_count[0] == 0) return;
  const DevSlot sl = slots[blockIdx.x];
  NodeRec* nodes = pool + sl.noff;
  float* nsize = nsize_pool + sl.noff;
  const uint32_t n = sl.n, bucket = sl.bucket, cap = sl.ncap;
  const uint64_t max_nodes = sl.max_nodes;
  const uint32_t lane = threadIdx.x;

  // keys per cell; cell 63 is what the other cells leave of n (with 10-bit
  // digits the padding keys share its last digit)
  constexpr int top = gs_sort_passes(BITS) - 1;
  uint32_t c = 0;
  if (lane < 63) {
    if (BITS == 8) {
      c = sc->sort_hist[top * 256 + lane];
    } else {
#pragma unroll
      for (int i = 0; i < 16; i++) c += sc->sort_hist[top * 1024 + lane * 16 + i];
    }
  }
  uint32_t x = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if ((int)lane >= o) x += y;
  }
  s_ex[lane] = x - c;  // first sorted position of the cell's keys
  if (lane == 0) s_ex[64] = n;
  __syncthreads();

  const int thr0 = (int)sc->level_thr[0], thr1 = (int)sc->level_thr[1];
  const int thr2 = (int)sc->level_thr[2];
  const float size1 = sc->level_size[1], size2 = sc->level_size[2];
  uint32_t st = 0;  // status flags

  // level 0: the root record of the root kernel covers [0, n) with skip -1
  const int32_t rskip = -1;
  const bool rleaf = n <= bucket || 0 >= GS_MAX_LEVEL || 9ull > max_nodes;
  uint32_t mask0 = 0;
  if (!rleaf) {
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (s_ex[8 * k + 8] > s_ex[8 * k]) mask0 |= 1u << k;
  }
  const uint32_t nch0 = __popc(mask0), fc0 = 1;  // level 1 starts after the root
  uint32_t cnt1 = nch0;
  if (1ull + nch0 > cap) {
    st |= GS_ST_NODE_POOL;
    cnt1 = cap > 1 ? cap - 1 : 0;
  }
  if (lane < 8 && ((mask0 >> lane) & 1u)) {
    const uint32_t c1 = fc0 + __popc(mask0 & ((1u << lane) - 1u));
    if (c1 < cap) {
      // pre-order with octant 7 first: after child c comes c - 1
      nsize[c1] = size1;
      nodes[c1].info = make_int4(thr1, (int)s_ex[8 * lane],
                                 (int)gs_meta_open(1, s_ex[8 * lane + 8] - s_ex[8 * lane]),
                                 c1 > fc0 ? gs_link(c1 - 1) : rskip);
    }
  }
  if (lane == 0)
    nodes[0].info = rleaf ? make_int4(thr0, 0, (int)gs_meta_leaf(0, n), rskip)
                          : make_int4(thr0, gs_link(fc0 + nch0 - 1),
                                      (int)gs_meta_internal(0, mask0, nch0), rskip);

  // level 1: the node of root octant k is the rank of k among the root's
  // children (those that fitted the pool)
  const uint32_t next1 = fc0 + cnt1;  // level 2 starts here
  uint32_t mask1[8], excl1[8], total1 = 0;
  bool leaf1[8], valid1[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t rank = __popc(mask0 & ((1u << k) - 1u));
    valid1[k] = ((mask0 >> k) & 1u) && rank < cnt1;
    const uint32_t idx = fc0 + rank, cnt = s_ex[8 * k + 8] - s_ex[8 * k];
    leaf1[k] = cnt <= bucket || 1 >= GS_MAX_LEVEL || (uint64_t)idx + 9 > max_nodes;
    mask1[k] = 0;
    if (valid1[k] && !leaf1[k]) {
#pragma unroll
      for (int m = 0; m < 8; m++)
        if (s_ex[8 * k + m + 1] > s_ex[8 * k + m]) mask1[k] |= 1u << m;
    }
    excl1[k] = total1;
    total1 += __popc(mask1[k]);
  }
  {
    // this lane's level-1 node (lanes 0..7) and level-2 child (cell = lane)
    const uint32_t k = lane >> 3, m = lane & 7u;
    uint32_t mk = 0, ek = 0;
    bool vk = false, lk = true;
#pragma unroll
    for (int q = 0; q < 8; q++)
      if ((uint32_t)q == k) {
        mk = mask1[q];
        ek = excl1[q];
        vk = valid1[q];
        lk = leaf1[q];
      }
    const uint32_t idx = fc0 + __popc(mask0 & ((1u << k) - 1u));
    const int32_t skip_k = idx > fc0 ? gs_link(idx - 1) : rskip;
    const uint32_t fc = next1 + ek, nch = __popc(mk);
    if (vk && ((mk >> m) & 1u)) {
      const uint32_t c2 = fc + __popc(mk & ((1u << m) - 1u));
      if (c2 < cap) {
        nsize[c2] = size2;
        nodes[c2].info = make_int4(thr2, (int)s_ex[lane],
                                   (int)gs_meta_open(2, s_ex[lane + 1] - s_ex[lane]),
                                   c2 > fc ? gs_link(c2 - 1) : skip_k);
      }
    }
    if (m == 0 && vk) {
      const uint32_t lo = s_ex[8 * k], cnt = s_ex[8 * k + 8] - lo;
      nodes[idx].info = lk ? make_int4(thr1, (int)lo, (int)gs_meta_leaf(1, cnt), skip_k)
                           : make_int4(thr1, gs_link(fc + nch - 1),
                                       (int)gs_meta_internal(1, mk, nch), skip_k);
    }
  }
  if (lane == 0) {
    if (cnt1 == 0) {  // level 1 is empty: the tree ends at the root
      sc->level_start[1] = fc0;
      sc->level_count[1] = 0;
      sc->level_start[2] = fc0;
      sc->level_count[2] = 0;
      sc->num_nodes = fc0;
      sc->num_levels = 1;
    } else {
      uint32_t cnt2 = total1;
      if ((uint64_t)next1 + total1 > cap) {
        st |= GS_ST_NODE_POOL;
        cnt2 = cap > next1 ? cap - next1 : 0;
      }
      sc->level_start[1] = fc0;
      sc->level_count[1] = cnt1;
      sc->level_start[2] = next1;
      sc->level_count[2] = cnt2;
      sc->num_nodes = next1 + cnt2;
      sc->num_levels = 2;
    }
    if (st) atomicOr(&batch->status, st);
  }
}

}  // namespace

hipError_t gs_launch_build(gs_sim* s, bool top_from_hist) {
  uint32_t grid = s->build_words < 1024 ? s->build_words : 1024;
  if (grid < 1) grid = 1;
  int d0 = 0;
  if (top_from_hist) {
    // levels 0 and 1 in one launch (the level-2 launch uses build_status[0])
    const uint32_t tg = s->nslots > 32 ? s->nslots : 32;
    if (s->radix_bits == 10)
      build_top_kernel<10><<<tg, 64, 0, s->stream>>>(s->nodes, s->nsize, s->slots, s->sc,
                                                     s->batch, s->nslots, s->build_status[0],
                                                     s->build_words);
    else
      build_top_kernel<8><<<tg, 64, 0, s->stream>>>(s->nodes, s->nsize, s->slots, s->sc,
                                                    s->batch, s->nslots, s->build_status[0],
                                                    s->build_words);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    d0 = 2;
  }
  for (int d = d0; d <= GS_MAX_LEVEL; d++) {
    build_level_kernel<<<grid, GS_BUILD_TILE, 0, s->stream>>>(
        s->keys[gs_sorted_buf(s)], s->nodes, s->nsize, s->slots, s->sc, s->batch, s->nslots, d,
        s->build_status[d & 1],
        s->build_status[(d + 1) & 1], s->build_words);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ---- expansion to the reference's 64-byte node layout (mortonTree.js:12-21)
namespace {
__global__ __launch_bounds__(256) void node64_kernel(const NodeRec* __restrict__ nodes,
                                                     const float* __restrict__ nsize,
                                                     const DevScalars* __restrict__ sc,
                                                     uint32_t* __restrict__ out) {
  const uint32_t num = sc->num_nodes;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < num; i += gridDim.x * blockDim.x) {
    const float4 c = nodes[i].cm;
    const int4 inf = nodes[i].info;
    uint32_t w[16];
    w[0] = __float_as_uint(c.x);
    w[1] = __float_as_uint(c.y);
    w[2] = __float_as_uint(c.z);
    w[3] = __float_as_uint(c.w);
    w[4] = __float_as_uint(nsize[i]);
    const uint32_t meta = (uint32_t)inf.z;
    const bool leaf = gs_meta_is_leaf(meta);
    w[5] = leaf ? 1u : 0u;
    w[6] = leaf ? (uint32_t)inf.y : 0u;
    w[7] = leaf ? gs_meta_count(meta) : 0u;
    const uint32_t mask = leaf ? 0u : gs_meta_mask(meta);
    uint32_t c0 = gs_node(inf.y) + 1u - gs_meta_nc(meta);  // internal: y = last child
    for (int k = 0; k < 8; k++) {
      if (mask & (1u << k)) {
        w[8 + k] = c0++;
      } else {
        w[8 + k] = 0xffffffffu;
      }
    }
    uint4* o = (uint4*)(out + (size_t)i * 16);
    o[0] = make_uint4(w[0], w[1], w[2], w[3]);
    o[1] = make_uint4(w[4], w[5], w[6], w[7]);
    o[2] = make_uint4(w[8], w[9], w[10], w[11]);
    o[3] = make_uint4(w[12], w[13], w[14], w[15]);
  }
}

// ---- opening thresholds after a change of theta (updateParams) -----------
// Every record of the slot's tree gets the threshold of its own size for the
// new theta^2 (trees imported from the host may hold any sizes), and the
// depth table follows for the next build.
__global__ __launch_bounds__(256) void rethreshold_kernel(NodeRec* __restrict__ nodes,
                                                          const float* __restrict__ nsize,
                                                          DevScalars* sc, float theta2) {
  const uint32_t num = sc->num_nodes;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g <= GS_MAX_LEVELS)
    sc->level_thr[g] = gs_open_threshold(__int_as_float(gs_s2_bits(sc->level_size[g])), theta2);
  for (uint32_t i = g; i < num; i += gridDim.x * blockDim.x)
    nodes[i].info.x = (int)gs_open_threshold(__int_as_float(gs_s2_bits(nsize[i])), theta2);
}

// ---- import of a tree in the reference's 64-byte layout (uploadBHTree) ----
__global__ __launch_bounds__(256) void import64_kernel(NodeRec* __restrict__ nodes,
                                                       float* __restrict__ nsize,
                                                       DevBatch* batch,
                                                       const uint32_t* __restrict__ in,
                                                       const int2* __restrict__ links,
                                                       uint32_t num, float theta2) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < num; i += gridDim.x * blockDim.x) {
    const uint4* r = (const uint4*)(in + (size_t)i * 16);
    const uint4 a = r[0], b = r[1], c = r[2], d = r[3];
    nodes[i].cm = make_float4(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z),
                              __uint_as_float(a.w));
    const int32_t skip = links[i].x < 0 ? -1 : gs_link((uint32_t)links[i].x);
    const uint32_t depth = (uint32_t)links[i].y;
    const float size = __uint_as_float(b.x);
    const int s2 = (int)gs_open_threshold(__int_as_float(gs_s2_bits(size)), theta2);
    nsize[i] = size;
    if (b.y == 1u) {
      nodes[i].info = make_int4(s2, (int)b.z, (int)gs_meta_leaf(depth, b.w), skip);
      continue;
    }
    const uint32_t ch[8] = {c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
    uint32_t mask = 0, fc = 0xffffffffu, nc = 0;
    for (int k = 0; k < 8; k++) {
      if ((int32_t)ch[k] >= 0) {
        if (fc == 0xffffffffu) fc = ch[k];
        // BFS trees keep a node's children contiguous in octant order
        if (ch[k] != fc + nc) atomicOr(&batch->status, GS_ST_LAYOUT);
        mask |= 1u << k;
        nc++;
      }
    }
    // internal: y = last child (a childless node continues at its skip)
    nodes[i].info = make_int4(s2, nc ? gs_link(fc + nc - 1) : skip,
                              (int)gs_meta_internal(depth, mask, nc), skip);
  }
}

// gather through a host permutation for one slot; every other slot is copied
// unchanged so the whole pool moves to the other ping-pong buffer (stellar
// records too, when present)
__global__ __launch_bounds__(256) void gather_kernel(const uint32_t* __restrict__ perm,
                                                     uint32_t target,
                                                     const DevSlot* __restrict__ slots,
                                                     const uint32_t* __restrict__ wblock_slot,
                                                     uint32_t wg, uint64_t total,
                                                     const float4* __restrict__ pin,
                                                     const float4* __restrict__ qin,
                                                     float4* __restrict__ pout,
                                                     float4* __restrict__ qout,
                                                     const float4* __restrict__ sin,
                                                     float4* __restrict__ sout) {
  // q runs over the walk blocks' particles (each slot's aligned span, slot
  // after slot in table order)
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t blk = (uint32_t)(q / wg);
    const uint32_t slot = wblock_slot[blk];
    const DevSlot sl = slots[slot];
    const uint32_t local = (blk - sl.wblock0) * wg + (uint32_t)(q % wg);
    uint32_t src = local;
    if (slot == target && local < sl.n) src = perm[local];
    const uint64_t p = (uint64_t)sl.poff + local;
    pout[p] = pin[sl.poff + src];
    qout[p] = qin[sl.poff + src];
    if (sin) sout[p] = sin[sl.poff + src];
  }
}
}  // namespace

hipError_t gs_launch_node64(gs_sim* s, uint32_t slot) {
  const HostSlot& h = s->hs[slot];
  uint32_t grid = (h.d.ncap + 255) / 256;
  if (grid > 4096) grid = 4096;
  node64_kernel<<<grid, 256, 0, s->stream>>>(s->nodes + h.d.noff, s->nsize + h.d.noff,
                                              s->sc + slot, s->node64);
  return hipGetLastError();
}

hipError_t gs_launch_rethreshold(gs_sim* s, uint32_t slot) {
  const HostSlot& h = s->hs[slot];
  uint32_t grid = (h.d.ncap + 255) / 256;
  if (grid > 2048) grid = 2048;
  rethreshold_kernel<<<grid, 256, 0, s->stream>>>(s->nodes + h.d.noff, s->nsize + h.d.noff,
                                                  s->sc + slot, s->host_prm[slot].theta2);
  return hipGetLastError();
}

hipError_t gs_launch_import64(gs_sim* s, uint32_t slot, const uint32_t* words, const int2* links,
                              uint32_t num) {
  uint32_t grid = (num + 255) / 256;
  if (grid > 4096) grid = 4096;
  if (grid < 1) grid = 1;
  import64_kernel<<<grid, 256, 0, s->stream>>>(s->nodes + s->hs[slot].d.noff,
                                                s->nsize + s->hs[slot].d.noff, s->batch, words,
                                                links, num, s->host_prm[slot].theta2);
  return hipGetLastError();
}

hipError_t gs_launch_gather(gs_sim* s, uint32_t slot, const uint32_t* perm) {
  const uint64_t total = (uint64_t)s->walk_blocks * s->walk_wg;
  uint64_t g = (total + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) return hipSuccess;
  gather_kernel<<<(uint32_t)g, 256, 0, s->stream>>>(perm, slot, s->slots, s->wblock_slot,
                                                     s->walk_wg, total, s->posm[s->cur],
                                                     s->velp[s->cur], s->posm[s->cur ^ 1],
                                                     s->velp[s->cur ^ 1], s->stel[s->stel_cur],
                                                     s->stel[s->stel_cur ^ 1]);
  return hipGetLastError();
}
