// Stable LSD radix sort of (Morton code, particle index) pairs — the device
// form of radixSortIndices (mortonTree.js:68-84).  A stable sort by key of the
// identity index array is unique, so 4 passes of 8-bit digits or 3 passes of
// 10-bit digits give exactly the permutation of the reference's 3 passes of
// 10-bit digits.  (10-bit passes look at bits 0..29 only; padding keys
// 0xffffffff then tie with code 0x3fffffff and stay behind it, because their
// indices are larger and the sort is stable.)
//
// One kernel per pass ("onesweep") over every slot of the pool; each slot is
// sorted on its own (slot-local keys, values and output positions):
//   1. a block takes the next 4096-key tile in launch order (atomic ticket)
//      and maps it to (slot, tile of the slot) through a static table; tiles
//      past the slot's active particles exit;
//   2. each wavefront ranks its 16 x 64 keys in order: per 64-key round, one
//      wave ballot per digit bit gives every lane the mask of lanes holding
//      the same digit (peers), rank = running per-wave digit count +
//      popcount(peers below);
//   3. per digit, the tile publishes its count and looks back over earlier
//      tiles of its slot (decoupled look-back; a slot's first tile starts the
//      chain) to get its output offset, added to the digit's base from the
//      slot's histograms the encode kernel accumulated;
//   4. keys are staged in LDS in tile-sorted order and written out as
//      contiguous per-digit runs.
// The last pass also gathers the particle records through the permutation
// (apply_permutation, nbody_bh.wgsl:319-326), so sorted positions land in the
// other ping-pong buffer without a separate pass over the index array.
//
// HBM per pass: 8 B read + 8 B written per key (4 B read in the first pass,
// which pairs each key with its slot-local index itself; + 32 B gathered +
// 32 B written per particle in the last pass).  8-bit digits write 256 runs of
// ~16 keys per tile; 10-bit digits save a pass but write 1024 runs of ~4.
#include "gs_internal.h"

namespace {

constexpr uint32_t kAgg = 1u << 30;
constexpr uint32_t kPre = 2u << 30;
constexpr uint32_t kVal = (1u << 30) - 1;

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// exclusive scan over the 256 threads of a block (one value each)
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* s_wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wsum[w] = x;
  __syncthreads();
  uint32_t off = 0;
  for (int k = 0; k < w; k++) off += s_wsum[k];
  __syncthreads();
  return off + x - v;
}

template <int BITS, bool GATHER>
__global__ __launch_bounds__(GS_SORT_BLOCK) void onesweep_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int pass, uint32_t* status,
    DevBatch* batch, const DevSlot* __restrict__ slots, const DevScalars* __restrict__ scs,
    const uint32_t* __restrict__ tile_tab, const float4* __restrict__ pin,
    const float4* __restrict__ qin, float4* __restrict__ pout, float4* __restrict__ qout,
    const float4* __restrict__ sin, float4* __restrict__ sout) {
  constexpr int ITEMS = GS_SORT_ITEMS;
  constexpr uint32_t RADIX = 1u << BITS;
  constexpr int DPT = (int)RADIX / GS_SORT_BLOCK;  // digits per thread
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wh[GS_SORT_BLOCK / 64][RADIX];
  __shared__ uint32_t s_dstart[RADIX];
  __shared__ uint32_t s_gbase[RADIX];
  __shared__ uint32_t s_wsum[GS_SORT_BLOCK / 64];
  __shared__ uint32_t s_k[GS_SORT_TILE];
  __shared__ uint32_t s_v[GS_SORT_TILE];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int shift = pass * BITS;
  if (tid == 0) s_tile = atomicAdd(&batch->tickets[pass], 1u);
  for (uint32_t k = tid; k < (GS_SORT_BLOCK / 64) * RADIX; k += GS_SORT_BLOCK)
    (&s_wh[0][0])[k] = 0;
  __syncthreads();
  const uint32_t gtile = s_tile;
  const uint32_t tt = tile_tab[gtile];
  const uint32_t slot = tt >> 16, tile = tt & 0xffffu;
  const DevSlot sl = slots[slot];
  const uint32_t active = (sl.n + GS_SORT_TILE - 1) / GS_SORT_TILE;
  if (tile >= active) return;  // no later tile of this slot is active either
  // keys exist up to the end of the slot's last active tile (inside the cap)
  const uint32_t n = min(active * (uint32_t)GS_SORT_TILE, sl.cap);
  kin += sl.poff;
  vin += sl.poff;
  kout += sl.poff;
  vout += sl.poff;
  const uint32_t base = tile * GS_SORT_TILE;
  const uint64_t below_mask = (1ull << lane) - 1;

  // every key and value load of the tile issues before the first is used:
  // one memory latency per tile (ranking item r before loading item r + 1
  // would wait for each load in turn, as the loads are conditional and
  // cannot be moved above the previous item's ballots), and the values,
  // needed only after the look-back, add no latency of their own
  uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    const uint32_t idx = base + (uint32_t)w * (64 * ITEMS) + (uint32_t)r * 64 + lane;
    const bool valid = idx < n;
    key[r] = valid ? kin[idx] : 0u;
    val[r] = pass != 0 && valid ? vin[idx] : idx;  // the first pass pairs keys with their index
  }
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    const uint32_t idx = base + (uint32_t)w * (64 * ITEMS) + (uint32_t)r * 64 + lane;
    const bool valid = idx < n;
    const uint32_t k = key[r];
    const uint32_t d = (k >> shift) & (RADIX - 1);
    // the lanes whose digit equals this lane's: per bit, the lanes that
    // agree with this lane's bit, ~(ballot(bit) ^ (bit ? ~0 : 0)), the
    // 0 / -1 word of a signed one-bit extract in both halves (a select per
    // half, as `bit ? bb : ~bb` compiles, costs twice the instructions)
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; b++) {
      const uint32_t sb = (uint32_t)__builtin_amdgcn_sbfe((int32_t)d, (uint32_t)b, 1u);
      const uint64_t bb = __ballot(sb != 0u);
      peers &= ~(bb ^ (((uint64_t)sb << 32) | sb));
    }
    const uint32_t below = (uint32_t)__popcll(peers & below_mask);
    uint32_t prior = 0;
    if (valid) prior = s_wh[w][d];
    rank[r] = prior + below;
    if (valid && below == 0) s_wh[w][d] = prior + (uint32_t)__popcll(peers);
  }
  __syncthreads();

  // digits tid*DPT .. tid*DPT+DPT-1: wave-exclusive offsets and tile counts
  uint32_t total[DPT], hist[DPT];
  uint32_t tsum = 0, hsum = 0;
  uint32_t* st[DPT];
#pragma unroll
  for (int j = 0; j < DPT; j++) {
    const uint32_t d = (uint32_t)tid * DPT + j;
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < GS_SORT_BLOCK / 64; k++) {
      const uint32_t c = s_wh[k][d];
      s_wh[k][d] = t;
      t += c;
    }
    total[j] = t;
    st[j] = status + (size_t)gtile * RADIX + d;
    st_agent(st[j], (tile == 0 ? kPre : kAgg) | t);
    hist[j] = scs[slot].sort_hist[pass * RADIX + d];
    tsum += t;
    hsum += hist[j];
  }
  uint32_t hist_excl = block_excl_scan256(hsum, s_wsum);
  uint32_t tile_excl = block_excl_scan256(tsum, s_wsum);

#pragma unroll
  for (int j = 0; j < DPT; j++) {
    const uint32_t d = (uint32_t)tid * DPT + j;
    uint32_t excl = 0;
    // a digit that no key of the slot holds (most top digits of the 30-bit
    // codes) needs no prefix: no tile reads its status or its offset
    if (tile > 0 && hist[j] != 0) {
      // a slot's tiles are consecutive; LB predecessors are read per round
      // (independent loads in flight together) and consumed nearest first
      // up to the first inclusive prefix or the first unpublished tile
      constexpr int LB = 8;
      const int32_t first_gt = (int32_t)(gtile - tile);
      int32_t jt = (int32_t)gtile - 1;
      uint32_t spins = 0;
      for (;;) {
        uint32_t v[LB];
#pragma unroll
        for (int q = 0; q < LB; q++)
          v[q] = jt - q >= first_gt ? ld_agent(status + (size_t)(jt - q) * RADIX + d) : kPre;
        int q = 0;
        bool done = false;
        for (; q < LB; q++) {
          const uint32_t f = v[q] & ~kVal;
          if (f == 0) break;
          excl += v[q] & kVal;
          if (f == kPre) {
            done = true;
            break;
          }
        }
        if (done) break;
        jt -= q;
        if (q == 0) {  // the nearest unread tile has not published yet
          if (++spins > GS_SPIN_LIMIT) {  // bounded wait: flag and stop, never hang
            atomicOr(&batch->status, GS_ST_SPIN);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      st_agent(st[j], kPre | (excl + total[j]));
    }
    s_gbase[d] = hist_excl + excl;
    s_dstart[d] = tile_excl;
    hist_excl += hist[j];
    tile_excl += total[j];
  }
  __syncthreads();

#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    const uint32_t idx = base + (uint32_t)w * (64 * ITEMS) + (uint32_t)r * 64 + lane;
    if (idx < n) {
      const uint32_t dd = (key[r] >> shift) & (RADIX - 1);
      const uint32_t pos = s_dstart[dd] + s_wh[w][dd] + rank[r];
      s_k[pos] = key[r];
      s_v[pos] = val[r];
    }
  }
  __syncthreads();

  const uint32_t tile_n = n - base < (uint32_t)GS_SORT_TILE ? n - base : (uint32_t)GS_SORT_TILE;
  if (!GATHER) {
    for (uint32_t i = tid; i < tile_n; i += GS_SORT_BLOCK) {
      const uint32_t k = s_k[i];
      const uint32_t dd = (k >> shift) & (RADIX - 1);
      const uint32_t out = s_gbase[dd] + (i - s_dstart[dd]);
      if (out < n) {
        kout[out] = k;
        vout[out] = s_v[i];
      }
    }
    return;
  }
  // the gather: the records of GB items are loaded before the first is
  // stored (one memory latency per GB items, not one per item).  Every
  // item loads (one without a record to move, past the tile or the
  // particles, loads record 0 of the slot, which exists in an active tile)
  // so the loads run unconditionally and their values stay in registers.
  constexpr int GB = 8;
  for (uint32_t i0 = tid; i0 < tile_n; i0 += GS_SORT_BLOCK * GB) {
    uint32_t src[GB], dst[GB];
    bool move[GB];
#pragma unroll
    for (int b = 0; b < GB; b++) {
      const uint32_t i = i0 + (uint32_t)b * GS_SORT_BLOCK;
      src[b] = 0;
      dst[b] = 0;
      move[b] = false;
      if (i < tile_n) {
        const uint32_t k = s_k[i];
        const uint32_t dd = (k >> shift) & (RADIX - 1);
        const uint32_t out = s_gbase[dd] + (i - s_dstart[dd]);
        const uint32_t v = s_v[i];
        if (out < n) {
          kout[out] = k;
          vout[out] = v;
          if (v < n) {
            src[b] = v;
            dst[b] = out;
            move[b] = true;
          }
        }
      }
    }
    float4 pr[GB], qr[GB];
#pragma unroll
    for (int b = 0; b < GB; b++) {
      pr[b] = pin[sl.poff + src[b]];
      qr[b] = qin[sl.poff + src[b]];
    }
    // the values pass through empty asm here, so the compiler cannot sink
    // each load into the conditional store that uses it (which would wait
    // for each load in turn)
#pragma unroll
    for (int b = 0; b < GB; b++)
      asm volatile("" : "+v"(pr[b].x), "+v"(pr[b].y), "+v"(pr[b].z), "+v"(pr[b].w),
                   "+v"(qr[b].x), "+v"(qr[b].y), "+v"(qr[b].z), "+v"(qr[b].w));
#pragma unroll
    for (int b = 0; b < GB; b++)
      if (move[b]) {
        pout[sl.poff + dst[b]] = pr[b];
        qout[sl.poff + dst[b]] = qr[b];
      }
    if (sin) {  // stellar records, when present
      float4 sr[GB];
#pragma unroll
      for (int b = 0; b < GB; b++) sr[b] = sin[sl.poff + src[b]];
#pragma unroll
      for (int b = 0; b < GB; b++)
        asm volatile("" : "+v"(sr[b].x), "+v"(sr[b].y), "+v"(sr[b].z), "+v"(sr[b].w));
#pragma unroll
      for (int b = 0; b < GB; b++)
        if (move[b]) sout[sl.poff + dst[b]] = sr[b];
    }
  }
}

// every pass over the buffers of b; the last pass gathers the particle
// records of s when s is given
template <int BITS>
hipError_t launch_sort(const SortBufs& b, gs_sim* s) {
  constexpr int RADIX = 1 << BITS, PASSES = gs_sort_passes(BITS);
  for (int p = 0; p < PASSES; p++) {
    uint32_t* status = b.status + (size_t)p * b.tiles * RADIX;
    const int x = p & 1, y = x ^ 1;
    if (s && p == PASSES - 1)
      onesweep_kernel<BITS, true><<<b.tiles, GS_SORT_BLOCK, 0, b.stream>>>(
          b.keys[x], b.vals[x], b.keys[y], b.vals[y], p, status, b.batch, b.slots, b.sc,
          b.tile_tab, s->posm[s->cur], s->velp[s->cur], s->posm[s->cur ^ 1],
          s->velp[s->cur ^ 1], s->stel[s->stel_cur], s->stel[s->stel_cur ^ 1]);
    else
      onesweep_kernel<BITS, false><<<b.tiles, GS_SORT_BLOCK, 0, b.stream>>>(
          b.keys[x], b.vals[x], b.keys[y], b.vals[y], p, status, b.batch, b.slots, b.sc,
          b.tile_tab, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

SortBufs bufs_of(gs_sim* s) {
  SortBufs b;
  for (int k = 0; k < 2; k++) {
    b.keys[k] = s->keys[k];
    b.vals[k] = s->vals[k];
  }
  b.status = s->sort_status;
  b.batch = s->batch;
  b.slots = s->slots;
  b.sc = s->sc;
  b.tile_tab = s->tile_tab;
  b.tiles = s->sort_tiles;
  b.radix_bits = s->radix_bits;
  b.stream = s->stream;
  return b;
}

}  // namespace

hipError_t gs_launch_sort(gs_sim* s, bool gather) {
  const SortBufs b = bufs_of(s);
  return s->radix_bits == 10 ? launch_sort<10>(b, gather ? s : nullptr)
                             : launch_sort<8>(b, gather ? s : nullptr);
}

hipError_t gs_launch_sort_bufs(const SortBufs& b) {
  return b.radix_bits == 10 ? launch_sort<10>(b, nullptr) : launch_sort<8>(b, nullptr);
}
