// Bounding box, root cube and 30-bit Morton encode (mortonTree.js:43-62),
// one launch each over every slot of the pool.
//
//   bbox_kernel   per-block min/max over active particles whose x+y+z is
//                 finite (mortonTree.js:46-52); a block covers one GS_CHUNK
//                 of one slot (static chunk table).  Also clears this step's
//                 sort tickets and onesweep status words.
//   root_kernel   one block per slot: reduces the slot's partials, then one
//                 lane evaluates size = max(dx,dy,dz)*1.0001 + 1e-9 and
//                 s = 1024/size in integer soft-double (bit-exact with the JS
//                 Number math; no binary64 hardware arithmetic), the per-level
//                 node sizes f32(size * 2^-d), and
//                 the slot's root node covering [0, n); clears the slot's
//                 digit histograms.
//   encode_kernel q = min(ToInt32((p - min) * s), 1023) per axis, bit-exact
//                 with binary64 (quantize.h: an f32 evaluation with a proven
//                 error margin, integer soft-double within the margin of a
//                 cell boundary or outside the box), code =
//                 part1by2(qx) | qy<<1 | qz<<2; writes the
//                 sort's keys, padding keys 0xffffffff (sorted last) up to
//                 the end of the slot's last active chunk, and the slot's 4
//                 digit histograms.  (The sort's first pass pairs each key
//                 with its slot-local index itself.)
// All three are HBM-streaming: 16 B read per particle for bbox, 16 B read +
// 4 B written for encode.
#include "gs_internal.h"
#include "quantize.h"
#include "softdouble.h"

namespace {

// bbox runs 256 threads per block, one block per GS_CHUNK chunk
constexpr int kPerThread = GS_CHUNK / 256;
static_assert(GS_CHUNK % 256 == 0 && kPerThread <= 32, "a thread's particles of a chunk");
// encode runs kEncodeBlock threads per chunk: its per-particle work (the
// quantization, with the rare soft-double path, and four digit counts) is
// issue-bound, and 16 wavefronts per chunk keep more of it in flight than 4
// (43.6 against 48.4 us at N = 4,194,304)
constexpr int kEncodeBlock = 1024;

__device__ __forceinline__ uint32_t part1by2(uint32_t v) {  // mortonTree.js:29-36
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x30000ffu;
  v = (v | (v << 8)) & 0x300f00fu;
  v = (v | (v << 4)) & 0x30c30c3u;
  v = (v | (v << 2)) & 0x9249249u;
  return v;
}

__device__ __forceinline__ bool finite3(float4 p) {
  // JS isFinite(x + y + z) on f32 inputs summed in double: the sum of three
  // finite f32 values never overflows binary64, so this is "all finite".
  return __builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z);
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ void clear_step_state(DevBatch* batch, uint32_t* sort_status,
                                                 uint64_t status_words) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = g; k < status_words; k += stride) sort_status[k] = 0;
  if (g < 32) batch->tickets[g] = 0;
}

__global__ __launch_bounds__(256) void reset_kernel(DevBatch* batch, uint32_t* sort_status,
                                                    uint64_t status_words, DevScalars* sc,
                                                    uint32_t nslots) {
  clear_step_state(batch, sort_status, status_words);
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t per = GS_HIST_WORDS;
  for (uint64_t k = g; k < (uint64_t)nslots * per; k += stride) sc[k / per].sort_hist[k % per] = 0;
}

__global__ __launch_bounds__(256) void bbox_kernel(const float4* __restrict__ posm,
                                                   const DevSlot* __restrict__ slots,
                                                   const uint2* __restrict__ chunk_tab,
                                                   float* __restrict__ part, DevBatch* batch,
                                                   uint32_t* sort_status, uint64_t status_words) {
  clear_step_state(batch, sort_status, status_words);
  const uint2 ct = chunk_tab[blockIdx.x];
  const DevSlot sl = slots[ct.x];
  const uint32_t end = min(ct.y + (uint32_t)GS_CHUNK, sl.n);
  float mnx = INFINITY, mny = INFINITY, mnz = INFINITY;
  float mxx = -INFINITY, mxy = -INFINITY, mxz = -INFINITY;
  uint32_t cnt = 0;
  // the thread's GS_CHUNK / 256 particles are loaded before the first is
  // used: one memory latency per chunk, not one per particle
  float4 pv[kPerThread];
#pragma unroll
  for (int r = 0; r < kPerThread; r++) {
    const uint32_t i = ct.y + threadIdx.x + (uint32_t)r * 256u;
    pv[r] = i < end ? posm[sl.poff + i] : make_float4(NAN, NAN, NAN, 0.f);
  }
#pragma unroll
  for (int r = 0; r < kPerThread; r++) {
    const float4 p = pv[r];
    if (finite3(p)) {  // (past the chunk's end: NaN)
      mnx = fminf(mnx, p.x); mxx = fmaxf(mxx, p.x);
      mny = fminf(mny, p.y); mxy = fmaxf(mxy, p.y);
      mnz = fminf(mnz, p.z); mxz = fmaxf(mxz, p.z);
      cnt++;
    }
  }
  mnx = wave_min(mnx); mny = wave_min(mny); mnz = wave_min(mnz);
  mxx = wave_max(mxx); mxy = wave_max(mxy); mxz = wave_max(mxz);
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  __shared__ float s[4][7];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    s[w][0] = mnx; s[w][1] = mny; s[w][2] = mnz;
    s[w][3] = mxx; s[w][4] = mxy; s[w][5] = mxz;
    s[w][6] = __uint_as_float(cnt);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float r[6] = {s[0][0], s[0][1], s[0][2], s[0][3], s[0][4], s[0][5]};
    uint32_t c = __float_as_uint(s[0][6]);
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
      for (int j = 0; j < 3; j++) r[j] = fminf(r[j], s[k][j]);
      for (int j = 3; j < 6; j++) r[j] = fmaxf(r[j], s[k][j]);
      c += __float_as_uint(s[k][6]);
    }
    float* o = part + (size_t)blockIdx.x * 8;
    for (int j = 0; j < 6; j++) o[j] = r[j];
    o[6] = __uint_as_float(c);
  }
}

__global__ __launch_bounds__(256) void root_kernel(const float* __restrict__ part,
                                                   const DevSlot* __restrict__ slots,
                                                   const DevParams* __restrict__ prm,
                                                   DevScalars* scs, NodeRec* nodes,
                                                   float* nsize) {
  const uint32_t slot = blockIdx.x;
  const DevSlot sl = slots[slot];
  DevScalars* sc = scs + slot;
  __shared__ uint64_t s_size;
  for (int k = threadIdx.x; k < GS_HIST_WORDS; k += blockDim.x) sc->sort_hist[k] = 0;
  float r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  uint32_t c = 0;
  for (uint32_t b = threadIdx.x; b < sl.nchunks; b += blockDim.x) {
    if ((uint64_t)b * GS_CHUNK >= sl.n) break;  // chunks past n hold no particle
    const float* o = part + (size_t)(sl.chunk0 + b) * 8;
    for (int j = 0; j < 3; j++) r[j] = fminf(r[j], o[j]);
    for (int j = 3; j < 6; j++) r[j] = fmaxf(r[j], o[j]);
    c += __float_as_uint(o[6]);
  }
  for (int j = 0; j < 3; j++) r[j] = wave_min(r[j]);
  for (int j = 3; j < 6; j++) r[j] = wave_max(r[j]);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  __shared__ float s[4][7];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    for (int j = 0; j < 6; j++) s[w][j] = r[j];
    s[w][6] = __uint_as_float(c);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
      for (int j = 0; j < 3; j++) r[j] = fminf(r[j], s[k][j]);
      for (int j = 3; j < 6; j++) r[j] = fmaxf(r[j], s[k][j]);
    }
    c = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) c += __float_as_uint(s[k][6]);
    sc->nfinite = c;
    uint64_t size;
    if (c == 0) {
      // no finite particle: minX = +Inf, size = -Inf, s = -0 -> every code is 0
      size = 0xfff0000000000000ull;
      for (int j = 0; j < 3; j++) {
        sc->min_bits[j] = 0x7ff0000000000000ull;
        sc->min_f32[j] = INFINITY;
      }
      sc->s_bits = 0x8000000000000000ull;
      sc->s32 = -0.0f;
    } else {
      // the JS Number arithmetic in integer soft-double, each operation
      // correctly rounded: d = max - min per axis (>= +0), size =
      // max(d) * 1.0001 + 1e-9, s = 1024 / size
      uint64_t d[3];
      for (int j = 0; j < 3; j++) {
        const uint64_t mn = sd::from_f32(__float_as_uint(r[j]));
        sc->min_bits[j] = mn;
        sc->min_f32[j] = r[j];
        d[j] = sd::sub(sd::from_f32(__float_as_uint(r[j + 3])), mn);
      }
      uint64_t m = d[0];
      if (sd::lt(m, d[1])) m = d[1];
      if (sd::lt(m, d[2])) m = d[2];
      const uint64_t k10001 = 0x3FF00068DB8BAC71ull;  // 1.0001
      const uint64_t k1em9 = 0x3E112E0BE826D695ull;   // 1e-9
      size = sd::add(sd::mul(m, k10001), k1em9);
      sc->s_bits = sd::div(0x4090000000000000ull /* 1024.0 */, size);
      sc->s32 = __uint_as_float(sd::to_f32(sc->s_bits));
    }
    sc->size_bits = size;
    s_size = size;
    // level 0 holds the root alone (none for an empty slot)
    for (int k = 0; k <= GS_MAX_LEVELS; k++) {
      sc->level_start[k] = 0;
      sc->level_count[k] = 0;
    }
    sc->level_count[0] = sl.n ? 1u : 0u;
    sc->num_nodes = sl.n ? 1u : 0u;
    sc->num_levels = 0;
  }
  __syncthreads();
  // node sizes per depth, f32(rootSize * 0.5^d) (mortonTree.js:124,161), and
  // their opening thresholds for the slot's theta^2, one lane per depth.  A
  // halving of a binary64 whose exponent field is 2 .. 0x7fe is exact and
  // lowers the field by one; Inf and NaN stay; below that the soft-double
  // multiply rounds as the JS multiplication does.
  float s2 = INFINITY;  // size^2 of this lane's depth (lanes past the depths: +Inf)
  if (threadIdx.x <= GS_MAX_LEVELS) {
    uint64_t sz = s_size;
    for (uint32_t k = 0; k < threadIdx.x; k++) {
      const uint32_t e = (uint32_t)(sz >> 52) & 0x7ffu;
      if (e == 0x7ffu) break;
      sz = e >= 2u ? sz - (1ull << 52) : sd::mul(sz, 0x3FE0000000000000ull /* 0.5 */);
    }
    const float f = __uint_as_float(sd::to_f32(sz));
    s2 = __int_as_float(gs_s2_bits(f));
    sc->level_size[threadIdx.x] = f;
    sc->level_thr[threadIdx.x] = gs_open_threshold(s2, prm[slot].theta2);
  }
  // smallest size^2 over the depths that hold nodes (0 .. GS_MAX_LEVEL);
  // fminf drops a NaN
  if (threadIdx.x < 64) {
    float v = threadIdx.x <= GS_MAX_LEVEL ? s2 : INFINITY;
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    if (threadIdx.x == 0) sc->min_s2 = v;
  }
  __syncthreads();
  // level 0 = the root, covering [0, n); an empty slot has no tree
  if (threadIdx.x == 0 && sl.n) {
    nsize[sl.noff] = sc->level_size[0];
    nodes[sl.noff].info = make_int4((int)sc->level_thr[0], 0, (int)gs_meta_open(0, sl.n), -1);
  }
}

// digit histograms of one key for every pass, into LDS hist[pass * radix + digit].
// Keys of neighbouring particles mostly share their high digits (the
// particles are in the previous step's Morton order): when every active lane
// of the wavefront has the same digit, one lane adds the count, instead of
// 64 atomics to one LDS word.
template <int BITS>
__device__ __forceinline__ void count_digits(uint32_t* hist, uint32_t code) {
  constexpr uint32_t RADIX = 1u << BITS;
  const uint64_t live = __builtin_amdgcn_read_exec();
#pragma unroll
  for (int ps = 0; ps < gs_sort_passes(BITS); ps++) {
    const uint32_t d = (code >> (ps * BITS)) & (RADIX - 1);
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    if (__ballot(d == d0) == live) {
      if (__builtin_amdgcn_mbcnt_hi((uint32_t)(live >> 32),
                                    __builtin_amdgcn_mbcnt_lo((uint32_t)live, 0u)) == 0u)
        atomicAdd(&hist[ps * RADIX + d0], (uint32_t)__popcll(live));
    } else {
      atomicAdd(&hist[ps * RADIX + d], 1u);
    }
  }
}

template <int BITS>
__device__ __forceinline__ void flush_hist(const uint32_t* hist, DevScalars* sc) {
  __syncthreads();
  for (int k = threadIdx.x; k < gs_sort_passes(BITS) << BITS; k += blockDim.x) {
    const uint32_t v = hist[k];
    if (v) atomicAdd(&sc->sort_hist[k], v);
  }
}

template <int BITS, bool KEEP_RAW>
__global__ __launch_bounds__(kEncodeBlock) void encode_kernel(const float4* __restrict__ posm,
                                                     const DevSlot* __restrict__ slots,
                                                     const uint2* __restrict__ chunk_tab,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ raw,
                                                     DevScalars* scs) {
  const uint2 ct = chunk_tab[blockIdx.x];
  const DevSlot sl = slots[ct.x];
  if (ct.y >= sl.n) return;  // the whole chunk is past the slot's particles
  DevScalars* sc = scs + ct.x;
  constexpr int HW = gs_sort_passes(BITS) << BITS;
  __shared__ uint32_t hist[HW];
  for (int k = threadIdx.x; k < HW; k += blockDim.x) hist[k] = 0;
  __syncthreads();
  const uint64_t mxb = sc->min_bits[0], myb = sc->min_bits[1], mzb = sc->min_bits[2];
  const uint64_t sb = sc->s_bits;
  const float mx = sc->min_f32[0], my = sc->min_f32[1], mz = sc->min_f32[2], s32 = sc->s32;
  const bool empty = sc->nfinite == 0;
  const uint32_t end = min(ct.y + (uint32_t)GS_CHUNK, sl.cap);
  for (uint32_t i = ct.y + threadIdx.x; i < end; i += blockDim.x) {
    uint32_t code = 0xffffffffu;  // padding: sorts after every 30-bit code
    if (i < sl.n) {
      const float4 p = posm[sl.poff + i];
      code = 0;
      if (!empty)
        code = part1by2(gsq::quantize(p.x, mx, s32, mxb, sb)) |
               (part1by2(gsq::quantize(p.y, my, s32, myb, sb)) << 1) |
               (part1by2(gsq::quantize(p.z, mz, s32, mzb, sb)) << 2);
      if (KEEP_RAW) raw[sl.poff + i] = code;
    }
    keys[sl.poff + i] = code;
    count_digits<BITS>(hist, code);
  }
  flush_hist<BITS>(hist, sc);
}

// digit histograms of caller-supplied keys in slot 0 (gs_radix_sort_indices
// and its allocation-free form): keys[i] = src[i] below n, padding keys
// 0xffffffff up to the end of the last active chunk (inside the capacity).
// src may equal keys.
template <int BITS>
__global__ __launch_bounds__(256) void hist_kernel(const uint32_t* src,
                                                   uint32_t* keys,
                                                   const DevSlot* __restrict__ slots,
                                                   DevScalars* scs) {
  const DevSlot sl = slots[0];
  constexpr int HW = gs_sort_passes(BITS) << BITS;
  __shared__ uint32_t hist[HW];
  for (int k = threadIdx.x; k < HW; k += blockDim.x) hist[k] = 0;
  __syncthreads();
  const uint32_t end = min(((sl.n + GS_CHUNK - 1) / GS_CHUNK) * GS_CHUNK, sl.cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < end; i += gridDim.x * blockDim.x) {
    const uint32_t code = i < sl.n ? src[sl.poff + i] : 0xffffffffu;
    keys[sl.poff + i] = code;
    count_digits<BITS>(hist, code);
  }
  flush_hist<BITS>(hist, scs);
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (uint32_t)(g < cap ? g : cap);
}

inline uint64_t status_words(const gs_sim* s) { return gs_sort_status_words(s); }

template <int BITS>
void launch_encode(gs_sim* s, const float4* posm, bool keep_raw) {
  if (keep_raw)
    encode_kernel<BITS, true><<<s->chunks, kEncodeBlock, 0, s->stream>>>(
        posm, s->slots, s->chunk_tab, s->keys[0], s->codes_raw, s->sc);
  else
    encode_kernel<BITS, false><<<s->chunks, kEncodeBlock, 0, s->stream>>>(
        posm, s->slots, s->chunk_tab, s->keys[0], s->codes_raw, s->sc);
}

}  // namespace

hipError_t gs_launch_reset(gs_sim* s) {
  reset_kernel<<<grid_for(status_words(s), 256, 2048), 256, 0, s->stream>>>(
      s->batch, s->sort_status, status_words(s), s->sc, s->nslots);
  return hipGetLastError();
}

hipError_t gs_launch_bbox(gs_sim* s, const float4* posm) {
  bbox_kernel<<<s->chunks, 256, 0, s->stream>>>(posm, s->slots, s->chunk_tab, s->bbox_part,
                                                s->batch, s->sort_status, status_words(s));
  return hipGetLastError();
}

hipError_t gs_launch_root(gs_sim* s) {
  root_kernel<<<s->nslots, 256, 0, s->stream>>>(s->bbox_part, s->slots, s->prm, s->sc,
                                                 s->nodes, s->nsize);
  return hipGetLastError();
}

hipError_t gs_launch_encode(gs_sim* s, const float4* posm, bool keep_raw) {
  if (s->radix_bits == 10)
    launch_encode<10>(s, posm, keep_raw);
  else
    launch_encode<8>(s, posm, keep_raw);
  return hipGetLastError();
}

hipError_t gs_launch_hist_bufs(const SortBufs& b, const uint32_t* src, uint32_t cap) {
  const uint32_t g = grid_for(cap, 256, 4096);
  if (b.radix_bits == 10)
    hist_kernel<10><<<g, 256, 0, b.stream>>>(src, b.keys[0], b.slots, b.sc);
  else
    hist_kernel<8><<<g, 256, 0, b.stream>>>(src, b.keys[0], b.slots, b.sc);
  return hipGetLastError();
}

hipError_t gs_launch_hist(gs_sim* s) {
  SortBufs b;
  b.keys[0] = s->keys[0];
  b.vals[0] = s->vals[0];
  b.slots = s->slots;
  b.sc = s->sc;
  b.radix_bits = s->radix_bits;
  b.stream = s->stream;
  return gs_launch_hist_bufs(b, s->keys[0], s->hs[0].d.cap);
}

hipError_t gs_launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t st) {
  fill_u32_kernel<<<grid_for(n, 256, 4096), 256, 0, st>>>(p, v, n);
  return hipGetLastError();
}
