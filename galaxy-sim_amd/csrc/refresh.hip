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

// every leaf of the tree: FP64 sum over its members in index order
__global__ __launch_bounds__(256) void refresh_leaves_kernel(const float4* __restrict__ posm,
                                                             NodeRec* __restrict__ nodes,
                                                             const DevScalars* __restrict__ sc) {
  const uint32_t num = sc->num_nodes;
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < num;
       idx += gridDim.x * blockDim.x) {
    const int4 inf = nodes[idx].info;
    if (!gs_meta_is_leaf((uint32_t)inf.z)) continue;
    double m = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    const uint32_t first = (uint32_t)inf.y, last = first + gs_meta_count((uint32_t)inf.z);
    for (uint32_t j = first; j < last; j++) {
      const float4 p = posm[j];
      const double pm = (double)p.w;
      m = __dadd_rn(m, pm);
      cx = __dadd_rn(cx, __dmul_rn((double)p.x, pm));
      cy = __dadd_rn(cy, __dmul_rn((double)p.y, pm));
      cz = __dadd_rn(cz, __dmul_rn((double)p.z, pm));
    }
    store_com(nodes, idx, m, cx, cy, cz);
  }
}

// internal nodes of one level: FP64 sum over children in octant order
__global__ __launch_bounds__(256) void refresh_level_kernel(NodeRec* __restrict__ nodes,
                                                            const DevScalars* __restrict__ sc,
                                                            int d) {
  const uint32_t start = sc->level_start[d], count = sc->level_count[d];
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count;
       t += gridDim.x * blockDim.x) {
    const uint32_t idx = start + t;
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
  uint32_t grid = (s->tree.cap + 255) / 256;
  if (grid > 8192) grid = 8192;
  refresh_leaves_kernel<<<grid, 256, 0, s->stream>>>(posm, s->tree.node, s->sc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // internal nodes exist only at depths 0..9; a level's node count is bounded
  // by 8^d and by the pool
  for (int d = GS_MAX_LEVEL - 1; d >= 0; d--) {
    uint64_t bound = 1ull << (3 * d);
    if (bound > s->tree.cap) bound = s->tree.cap;
    uint32_t g = (uint32_t)((bound + 255) / 256);
    if (g > 2048) g = 2048;
    refresh_level_kernel<<<g, 256, 0, s->stream>>>(s->tree.node, s->sc, d);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
