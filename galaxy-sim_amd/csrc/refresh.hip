// Bottom-up node COM / mass refresh in FP64 — encodeRefresh
// (NBodyGPU.js:308-319) with the summation order of the build's COM pass
// (mortonTree.js:129-137 for leaves, 166-181 for internal nodes):
//   leaf      m += pm; cx += x*pm ... over members in index order
//   internal  m += cm; cx += cx_c*cm ... over children in octant order
//   then inv = m > 0 ? 1/m : 0 and store f32(cx*inv), ..., f32(m).
// Every product and sum is an explicitly rounded FP64 op (no contraction),
// so the stored f32 words equal the reference build's bit-for-bit.
//
// Launch plan: one pass over every leaf of every level (a leaf only reads
// particles; leaves larger than a wavefront are queued and then summed by a
// wavefront each), then one launch per internal level, deepest first, each
// thread summing one node's children (contiguous, octant order); the four
// smallest levels share one launch of a block per slot.  Kernel boundaries
// order the levels; no device-wide fences or atomics are needed, which on a
// multi-XCD GPU (8 non-coherent L2s) is far cheaper than per-node
// release/acquire.  Node counts per level come from device scalars, so the
// launch sequence is fixed and graph-capturable.
//
// HBM: 16 B read per particle, 16 B written per node (+ child re-reads from
// L2).
#include "gs_internal.h"

namespace {

constexpr int LEAF_BATCH = 8;

// an empty asm that takes a loaded value in and out: the load issues before
// this point, and the value's uses cannot pull it past it
__device__ __forceinline__ void pin_value(float4& v) {
  asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
}

// a node stored with zero mass marks its slot, so that the walk runs its
// variant with the massless-node test (nbody_bh.wgsl:86)
__device__ __forceinline__ void store_com(NodeRec* nodes, uint32_t idx, double m, double cx,
                                          double cy, double cz, uint32_t* massless) {
  if ((float)m == 0.0f) *massless = 1u;  // the stored mass, after rounding
  const double inv = m > 0.0 ? __ddiv_rn(1.0, m) : 0.0;
  nodes[idx].cm = make_float4((float)__dmul_rn(cx, inv), (float)__dmul_rn(cy, inv),
                              (float)__dmul_rn(cz, inv), (float)m);
}

// FP64 sums over members [j, last) in index order, continuing (m, cx, cy, cz)
__device__ __forceinline__ void sum_members(const float4* __restrict__ pp, uint32_t j,
                                            uint32_t last, double& m, double& cx, double& cy,
                                            double& cz) {
  // members are loaded LEAF_BATCH at a time ahead of the (sequential,
  // order-fixed) FP64 sums, so the loads of a batch are in flight together:
  // every load is unconditional (past the leaf it re-reads the last member,
  // whose value is not used) and the values pass through empty asm, so the
  // compiler can neither make a load conditional nor sink it to its use
  // (either would wait for the loads one by one)
  for (uint32_t j0 = j; j0 < last; j0 += LEAF_BATCH) {
    float4 q[LEAF_BATCH];
#pragma unroll
    for (int k = 0; k < LEAF_BATCH; k++) q[k] = pp[min(j0 + (uint32_t)k, last - 1u)];
#pragma unroll
    for (int k = 0; k < LEAF_BATCH; k++) pin_value(q[k]);
#pragma unroll
    for (int k = 0; k < LEAF_BATCH; k++) {
      if (j0 + k >= last) break;
      const double pm = (double)q[k].w;
      m = __dadd_rn(m, pm);
      cx = __dadd_rn(cx, __dmul_rn((double)q[k].x, pm));
      cy = __dadd_rn(cy, __dmul_rn((double)q[k].y, pm));
      cz = __dadd_rn(cz, __dmul_rn((double)q[k].z, pm));
    }
  }
}

// every leaf of every slot: FP64 sum over its members in index order, one
// thread per node.  Leaves of more than GS_BIG_LEAF members (dense cores at
// the depth limit hold hundreds, and they sit next to each other in node
// order) are queued for refresh_big_kernel instead.
__global__ __launch_bounds__(256) void refresh_leaves_kernel(const float4* __restrict__ posm,
                                                             NodeRec* __restrict__ pool,
                                                             const DevSlot* __restrict__ slots,
                                                             const DevScalars* __restrict__ scs,
                                                             uint32_t nslots, DevBatch* batch,
                                                             uint2* __restrict__ big,
                                                             uint32_t big_cap,
                                                             uint32_t* __restrict__ massless) {
  __shared__ uint32_t s_pre[GS_MAX_SLOTS + 1 + GS_MAX_SLOTS / 64];
  gs_slot_prefix(threadIdx.x < nslots ? scs[threadIdx.x].num_nodes : 0u, nslots, s_pre);
  const uint32_t total = s_pre[nslots];
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < total;
       f += gridDim.x * blockDim.x) {
    const uint32_t slot = gs_find_slot(s_pre, nslots, f);
    const DevSlot sl = slots[slot];
    NodeRec* nodes = pool + sl.noff;
    const uint32_t idx = f - s_pre[slot];
    const int4 inf = nodes[idx].info;
    if (!gs_meta_is_leaf((uint32_t)inf.z)) continue;
    const uint32_t first = (uint32_t)inf.y, count = gs_meta_count((uint32_t)inf.z);
    if (count > GS_BIG_LEAF) {
      // disjoint leaves of > GS_BIG_LEAF particles fit the list; overlapping
      // ranges (only an uploaded tree could have them) are flagged
      const uint32_t k = atomicAdd(&batch->big_count, 1u);
      if (k < big_cap) big[k] = make_uint2(slot, idx);
      else atomicOr(&batch->status, GS_ST_LAYOUT);
      continue;
    }
    double m = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    sum_members(posm + sl.poff, first, first + count, m, cx, cy, cz);
    store_com(nodes, idx, m, cx, cy, cz, massless + slot);
  }
}

// the queued large leaves, one wavefront each, 256 members per round: every
// lane converts and multiplies 4 members (pm, x*pm, y*pm, z*pm in FP64, the
// same roundings as the thread-per-leaf sum) into LDS, then lanes 0..3 each
// carry one of the four running sums (m, cx, cy, cz) over the members in
// index order, so the four sequential chains advance in one instruction
__global__ __launch_bounds__(256) void refresh_big_kernel(const float4* __restrict__ posm,
                                                          NodeRec* __restrict__ pool,
                                                          const DevSlot* __restrict__ slots,
                                                          const DevBatch* __restrict__ batch,
                                                          const uint2* __restrict__ big,
                                                          uint32_t big_cap,
                                                          uint32_t* __restrict__ massless) {
  constexpr uint32_t CH = 256;  // members per round
  __shared__ double s_val[4][CH][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t nbig = min(batch->big_count, big_cap);
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t e = blockIdx.x * (blockDim.x >> 6) + (uint32_t)w; e < nbig; e += waves) {
    const uint2 b = big[e];
    const DevSlot sl = slots[b.x];
    NodeRec* nodes = pool + sl.noff;
    const int4 inf = nodes[b.y].info;
    const uint32_t first = (uint32_t)inf.y, last = first + gs_meta_count((uint32_t)inf.z);
    const float4* pp = posm + sl.poff;
    double acc = 0.0;  // lane k < 4: sum k of (m, cx, cy, cz)
    for (uint32_t c = first; c < last; c += CH) {
      // the round's loads first (unconditional, pinned as in sum_members;
      // past the leaf they re-read its last member and weigh 0)
      float4 qv[CH / 64];
#pragma unroll
      for (uint32_t k = 0; k < CH / 64; k++) qv[k] = pp[min(c + k * 64 + (uint32_t)lane, last - 1u)];
#pragma unroll
      for (uint32_t k = 0; k < CH / 64; k++) pin_value(qv[k]);
#pragma unroll
      for (uint32_t k = 0; k < CH / 64; k++) {
        const uint32_t t = k * 64 + (uint32_t)lane;
        const float4 q = c + t < last ? qv[k] : make_float4(0.f, 0.f, 0.f, 0.f);
        const double pm = (double)q.w;
        s_val[w][t][0] = pm;
        s_val[w][t][1] = __dmul_rn((double)q.x, pm);
        s_val[w][t][2] = __dmul_rn((double)q.y, pm);
        s_val[w][t][3] = __dmul_rn((double)q.z, pm);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (lane < 4) {
        const uint32_t nk = last - c < CH ? last - c : CH;
        uint32_t t = 0;
        for (; t + 8 <= nk; t += 8) {
          double v[8];
#pragma unroll
          for (int k = 0; k < 8; k++) v[k] = s_val[w][t + k][lane];
#pragma unroll
          for (int k = 0; k < 8; k++) acc = __dadd_rn(acc, v[k]);
        }
        for (; t < nk; t++) acc = __dadd_rn(acc, s_val[w][t][lane]);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const double m = __shfl(acc, 0), cx = __shfl(acc, 1), cy = __shfl(acc, 2),
                 cz = __shfl(acc, 3);
    if (lane == 0) store_com(nodes, b.y, m, cx, cy, cz, massless + b.x);
  }
}

// one internal node (a leaf is left alone): FP64 sum over its children in
// octant order
__device__ __forceinline__ void refresh_node(NodeRec* __restrict__ nodes, uint32_t idx,
                                             uint32_t* massless) {
  const int4 inf = nodes[idx].info;
  if (gs_meta_is_leaf((uint32_t)inf.z)) return;
  double m = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
  const uint32_t nc = gs_meta_nc((uint32_t)inf.z);
  if (nc > 0) {
    const uint32_t fc = gs_node(inf.y) + 1u - nc;  // y = last child
    // the (at most 8) children are loaded together, as in sum_members
    float4 q[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) q[k] = nodes[fc + min(k, nc - 1u)].cm;
#pragma unroll
    for (int k = 0; k < 8; k++) pin_value(q[k]);
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
      if (k >= nc) break;
      const double qm = (double)q[k].w;
      m = __dadd_rn(m, qm);
      cx = __dadd_rn(cx, __dmul_rn((double)q[k].x, qm));
      cy = __dadd_rn(cy, __dmul_rn((double)q[k].y, qm));
      cz = __dadd_rn(cz, __dmul_rn((double)q[k].z, qm));
    }
  }
  store_com(nodes, idx, m, cx, cy, cz, massless);
}

// levels below GS_REFRESH_TOP hold at most 8^d nodes per slot: one block
// per slot refreshes them all, deepest first, a block barrier ordering the
// levels (the children of level d were written by this block or by an
// earlier launch)
constexpr int GS_REFRESH_TOP = 4;

__global__ __launch_bounds__(256) void refresh_top_kernel(NodeRec* __restrict__ pool,
                                                          const DevSlot* __restrict__ slots,
                                                          const DevScalars* __restrict__ scs,
                                                          uint32_t* __restrict__ massless) {
  const uint32_t slot = blockIdx.x;
  NodeRec* nodes = pool + slots[slot].noff;
  for (int d = GS_REFRESH_TOP - 1; d >= 0; d--) {
    const uint32_t start = scs[slot].level_start[d], count = scs[slot].level_count[d];
    for (uint32_t f = threadIdx.x; f < count; f += blockDim.x)
      refresh_node(nodes, start + f, massless + slot);
    __syncthreads();
  }
}

// internal nodes of one level of every slot: FP64 sum over children in
// octant order
__global__ __launch_bounds__(256) void refresh_level_kernel(NodeRec* __restrict__ pool,
                                                            const DevSlot* __restrict__ slots,
                                                            const DevScalars* __restrict__ scs,
                                                            uint32_t nslots, int d,
                                                            DevBatch* batch,
                                                            uint32_t* __restrict__ massless) {
  __shared__ uint32_t s_pre[GS_MAX_SLOTS + 1 + GS_MAX_SLOTS / 64];
  // the large-leaf queue was consumed by the previous launch: empty it for
  // the next refresh
  if (d == GS_MAX_LEVEL - 1 && blockIdx.x == 0 && threadIdx.x == 0) batch->big_count = 0;
  gs_slot_prefix(threadIdx.x < nslots ? scs[threadIdx.x].level_count[d] : 0u, nslots, s_pre);
  const uint32_t total = s_pre[nslots];
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < total;
       f += gridDim.x * blockDim.x) {
    const uint32_t slot = gs_find_slot(s_pre, nslots, f);
    NodeRec* nodes = pool + slots[slot].noff;
    const uint32_t idx = scs[slot].level_start[d] + (f - s_pre[slot]);
    refresh_node(nodes, idx, massless + slot);
  }
}

}  // namespace

hipError_t gs_launch_refresh(gs_sim* s, const float4* posm) {
  // one thread per node, grid-stride: nodes are ~0.1-0.3 of the pool, and
  // 2048 blocks of 256 fill the GPU
  uint64_t g = (s->ntot / 4 + 255) / 256;
  if (g > 2048) g = 2048;
  if (g < 64) g = 64;
  hipError_t e = gs_launch_fill_u32(s->massless, 0u, s->nslots, s->stream);
  if (e != hipSuccess) return e;
  refresh_leaves_kernel<<<(uint32_t)g, 256, 0, s->stream>>>(posm, s->nodes, s->slots, s->sc,
                                                              s->nslots, s->batch, s->big_leaf,
                                                              s->big_cap, s->massless);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  refresh_big_kernel<<<256, 256, 0, s->stream>>>(posm, s->nodes, s->slots, s->batch, s->big_leaf,
                                                 s->big_cap, s->massless);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // internal nodes exist only at depths 0..9; a level's node count is bounded
  // by nslots * 8^d and by the pools.  The levels below GS_REFRESH_TOP share
  // one launch.
  for (int d = GS_MAX_LEVEL - 1; d >= GS_REFRESH_TOP; d--) {
    uint64_t bound = (uint64_t)s->nslots << (3 * d);
    if (bound > s->ntot) bound = s->ntot;
    uint32_t gl = (uint32_t)((bound + 255) / 256);
    if (gl > 2048) gl = 2048;
    refresh_level_kernel<<<gl, 256, 0, s->stream>>>(s->nodes, s->slots, s->sc, s->nslots, d,
                                                     s->batch, s->massless);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  refresh_top_kernel<<<s->nslots, 256, 0, s->stream>>>(s->nodes, s->slots, s->sc, s->massless);
  return hipGetLastError();
}
