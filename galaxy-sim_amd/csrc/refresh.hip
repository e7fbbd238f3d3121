// Bottom-up node COM / mass refresh in FP64 — encodeRefresh
// (NBodyGPU.js:308-319) with the summation order of the build's COM pass
// (mortonTree.js:129-137 for leaves, 166-181 for internal nodes):
//   leaf      m += pm; cx += x*pm ... over members in index order
//   internal  m += cm; cx += cx_c*cm ... over children in octant order
//   then inv = m > 0 ? 1/m : 0 and store f32(cx*inv), ..., f32(m).
// Every product and sum is an explicitly rounded FP64 op (no contraction),
// so the stored f32 words equal the reference build's bit-for-bit.
//
// One launch covers the whole tree instead of one dispatch per level: a
// thread sums its leaf, then climbs; at each parent it increments the
// parent's arrival counter (acq_rel, agent scope) and only the last child to
// arrive sums the parent (reading its children in octant order, so the
// result does not depend on arrival order) and continues upward.  The last
// arrival resets the counter, leaving every counter at zero for the next
// refresh.
//
// HBM: 16 B read per particle, 16 B written per node (+ child re-reads from
// L2).
#include "gs_internal.h"

namespace {

__device__ __forceinline__ void store_com(float4* cm, uint32_t idx, double m, double cx, double cy,
                                          double cz) {
  const double inv = m > 0.0 ? __ddiv_rn(1.0, m) : 0.0;
  cm[idx] = make_float4((float)__dmul_rn(cx, inv), (float)__dmul_rn(cy, inv),
                        (float)__dmul_rn(cz, inv), (float)m);
}

__global__ __launch_bounds__(256) void refresh_kernel(const float4* __restrict__ posm,
                                                      float4* cm, const int4* __restrict__ info,
                                                      const int32_t* __restrict__ parent,
                                                      uint32_t* arrivals,
                                                      const DevScalars* __restrict__ sc) {
  const uint32_t num = sc->num_nodes;
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < num;
       idx += gridDim.x * blockDim.x) {
    const int4 inf = info[idx];
    if ((inf.w & 1) == 0) continue;
    double m = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    const uint32_t first = (uint32_t)inf.y, last = first + (uint32_t)inf.z;
    for (uint32_t j = first; j < last; j++) {
      const float4 p = posm[j];
      const double pm = (double)p.w;
      m = __dadd_rn(m, pm);
      cx = __dadd_rn(cx, __dmul_rn((double)p.x, pm));
      cy = __dadd_rn(cy, __dmul_rn((double)p.y, pm));
      cz = __dadd_rn(cz, __dmul_rn((double)p.z, pm));
    }
    store_com(cm, idx, m, cx, cy, cz);

    int32_t node = (int32_t)idx;
    for (;;) {
      const int32_t par = parent[node];
      if (par < 0) break;
      const int4 pinf = info[par];
      const uint32_t nc = (uint32_t)pinf.z;
      const uint32_t old = __hip_atomic_fetch_add(&arrivals[par], 1u, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_AGENT);
      if (old + 1 != nc) break;
      arrivals[par] = 0;
      m = cx = cy = cz = 0.0;
      const uint32_t fc = (uint32_t)pinf.y;
      for (uint32_t c = fc; c < fc + nc; c++) {
        const float4 q = cm[c];
        const double qm = (double)q.w;
        m = __dadd_rn(m, qm);
        cx = __dadd_rn(cx, __dmul_rn((double)q.x, qm));
        cy = __dadd_rn(cy, __dmul_rn((double)q.y, qm));
        cz = __dadd_rn(cz, __dmul_rn((double)q.z, qm));
      }
      store_com(cm, (uint32_t)par, m, cx, cy, cz);
      node = par;
    }
  }
}

}  // namespace

hipError_t gs_launch_refresh(gs_sim* s, const float4* posm) {
  uint32_t grid = (s->tree.cap + 255) / 256;
  if (grid > 8192) grid = 8192;
  refresh_kernel<<<grid, 256, 0, s->stream>>>(posm, s->tree.cm, s->tree.info, s->tree.parent,
                                              s->tree.arrivals, s->sc);
  return hipGetLastError();
}
