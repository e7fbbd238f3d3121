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
// particles), then one launch per internal level, deepest first, each thread
// summing one node's children (contiguous, octant order).  Kernel boundaries
// order the levels; no device-wide fences or atomics are needed, which on a
// multi-XCD GPU (8 non-coherent L2s) is far cheaper than per-node
// release/acquire.  Node counts per level come from device scalars, so the
// launch sequence is fixed and graph-capturable.
//
// HBM: 16 B read per particle, 16 B written per node (+ child re-reads from
// L2).
#include "gs_internal.h"

namespace {

__device__ __forceinline__ void store_com(NodeRec* nodes, uint32_t idx, double m, double cx,
                                          double cy, double cz) {
  const double inv = m > 0.0 ? __ddiv_rn(1.0, m) : 0.0;
  nodes[idx].cm = make_float4((float)__dmul_rn(cx, inv), (float)__dmul_rn(cy, inv),
                              (float)__dmul_rn(cz, inv), (float)m);
}

// every leaf of every slot: FP64 sum over its members in index order
__global__ __launch_bounds__(256) void refresh_leaves_kernel(const float4* __restrict__ posm,
                                                             NodeRec* __restrict__ pool,
                                                             const DevSlot* __restrict__ slots,
                                                             const DevScalars* __restrict__ scs,
                                                             uint32_t nslots) {
  __shared__ uint32_t s_pre[GS_MAX_SLOTS + 1];
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
    double m = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    const uint32_t first = (uint32_t)inf.y, last = first + gs_meta_count((uint32_t)inf.z);
    const float4* pp = posm + sl.poff;
    for (uint32_t j = first; j < last; j++) {
      const float4 p = pp[j];
      const double pm = (double)p.w;
      m = __dadd_rn(m, pm);
      cx = __dadd_rn(cx, __dmul_rn((double)p.x, pm));
      cy = __dadd_rn(cy, __dmul_rn((double)p.y, pm));
      cz = __dadd_rn(cz, __dmul_rn((double)p.z, pm));
    }
    store_com(nodes, idx, m, cx, cy, cz);
  }
}

// internal nodes of one level of every slot: FP64 sum over children in
// octant order
__global__ __launch_bounds__(256) void refresh_level_kernel(NodeRec* __restrict__ pool,
                                                            const DevSlot* __restrict__ slots,
                                                            const DevScalars* __restrict__ scs,
                                                            uint32_t nslots, int d) {
  __shared__ uint32_t s_pre[GS_MAX_SLOTS + 1];
  gs_slot_prefix(threadIdx.x < nslots ? scs[threadIdx.x].level_count[d] : 0u, nslots, s_pre);
  const uint32_t total = s_pre[nslots];
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < total;
       f += gridDim.x * blockDim.x) {
    const uint32_t slot = gs_find_slot(s_pre, nslots, f);
    NodeRec* nodes = pool + slots[slot].noff;
    const uint32_t idx = scs[slot].level_start[d] + (f - s_pre[slot]);
    const int4 inf = nodes[idx].info;
    if (gs_meta_is_leaf((uint32_t)inf.z)) continue;
    double m = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    const uint32_t fc = (uint32_t)inf.y, nc = gs_meta_nc((uint32_t)inf.z);
    for (uint32_t c = fc; c < fc + nc; c++) {
      const float4 q = nodes[c].cm;
      const double qm = (double)q.w;
      m = __dadd_rn(m, qm);
      cx = __dadd_rn(cx, __dmul_rn((double)q.x, qm));
      cy = __dadd_rn(cy, __dmul_rn((double)q.y, qm));
      cz = __dadd_rn(cz, __dmul_rn((double)q.z, qm));
    }
    store_com(nodes, idx, m, cx, cy, cz);
  }
}

}  // namespace

hipError_t gs_launch_refresh(gs_sim* s, const float4* posm) {
  uint64_t g = (s->ntot / 4 + 255) / 256;  // nodes are ~0.1-0.3 of the pool
  if (g > 8192) g = 8192;
  if (g < 64) g = 64;
  refresh_leaves_kernel<<<(uint32_t)g, 256, 0, s->stream>>>(posm, s->nodes, s->slots, s->sc,
                                                              s->nslots);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // internal nodes exist only at depths 0..9; a level's node count is bounded
  // by nslots * 8^d and by the pools
  for (int d = GS_MAX_LEVEL - 1; d >= 0; d--) {
    uint64_t bound = (uint64_t)s->nslots << (3 * d);
    if (bound > s->ntot) bound = s->ntot;
    uint32_t gl = (uint32_t)((bound + 255) / 256);
    if (gl > 2048) gl = 2048;
    refresh_level_kernel<<<gl, 256, 0, s->stream>>>(s->nodes, s->slots, s->sc, s->nslots, d);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
