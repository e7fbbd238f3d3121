// Barnes-Hut walk + integrate — the `main` kernel of nbody_bh.wgsl:64-181.
//
// Wavefront-coherent, stackless traversal with per-lane exactness.  The 64
// lanes of a wavefront own 64 consecutive Morton-sorted particles and walk
// ONE node sequence together: the pre-order of the union of their trees, in
// the order the reference's per-thread stack pops (children pushed octant 0
// first, so octant 7 is visited first).  Each node record carries its depth
// and a skip link (the node that follows its subtree).  A lane is active at a
// node iff it opened every ancestor of the node; in pre-order this is one
// integer per lane, the active depth limit L: the lane is active at depth d
// iff d <= L.  (Descending from a node at depth d sets L = d + 1 for the
// lanes that open it and L = min(L, d) for the others; the ancestors above a
// later node reached by a skip link are unchanged, so the test stays exact.)
// At a node, the active lanes evaluate the reference's own per-particle
// decision (nbody_bh.wgsl:97-141): monopole if s^2 < theta^2 d^2 or
// s^2 < eps^2/4, else leaf direct sum (skipping j == i), else open.  If any
// lane opens, the wave descends to the last child; otherwise it follows the
// skip link.  Every lane therefore visits exactly the nodes its own stack
// would pop, in the same order: its interaction list and its summation order
// equal the reference walk's.  The node index is
// wave-uniform, so node records and leaf members are scalar (SMEM) broadcast
// loads, there is no stack in memory, and divergence costs only the lanes
// masked off.  The active, monopole and opening lanes are 64-bit wave masks
// (two vector compares per node, combined on the scalar unit) that set the
// execution mask of the monopole and leaf branches directly.
//
// Arithmetic: the opening test uses the exact reference expression order
// d2 = ((rx*rx + ry*ry) + rz*rz) + eps2 with no contraction (the library is
// built with -ffp-contract=off), so decisions match the transliteration bit
// for bit.  EXACT evaluates f = G*m / (d2*sqrt(d2)) with IEEE sqrt and
// division (bit-identical forces); the default path uses the hardware
// reciprocal square root, f = m * rsq(d2)^3, fused multiply-add accumulation
// and one multiplication by G at the end (same interaction list, <= 1e-5).
// Node records are one 32-byte scalar load; leaf members are read four at a
// time (one 64-byte scalar load).
//
// After the walk (STEP): Hernquist halos, kick-drift, gas drag, pad density
// and specific energy (nbody_bh.wgsl:144-180), then per-wavefront FP64
// partial sums of kinetic, pair-potential and halo energy, momentum, the
// reference's specific-energy sum and mass (fixed-order, deterministic).
#include "gs_internal.h"

namespace {

typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));

// EXACT: the reference expression f = G*m / (d2*sqrt(d2)), IEEE sqrt and '/',
// separate multiply and add (bit-identical to the transliteration).
__device__ __forceinline__ void interact_exact(float gm, float rx, float ry, float rz, float d2,
                                               float& ax, float& ay, float& az, float& pot) {
  const float f = gm / (d2 * __builtin_sqrtf(d2));
  ax = ax + rx * f;
  ay = ay + ry * f;
  az = az + rz * f;
  pot = pot - f * d2;
}

// Default path: m * rsq(d2)^3 with fused accumulation; G multiplies the sums
// once at the end of the walk.
__device__ __forceinline__ void interact_fast(float m, float rx, float ry, float rz, float d2,
                                              float& ax, float& ay, float& az, float& pot) {
  const float inv = __builtin_amdgcn_rsqf(d2);
  const float mi = m * inv;
  const float f = mi * (inv * inv);
  ax = __builtin_fmaf(rx, f, ax);
  ay = __builtin_fmaf(ry, f, ay);
  az = __builtin_fmaf(rz, f, az);
  pot = pot - mi;
}

// leaf member j on particle i (a lane that opened the leaf), skipping j == i
template <bool EXACT, bool COUNT>
__device__ __forceinline__ void leaf_pair(float4 q, uint32_t j, uint32_t i, float4 p, float G,
                                          float eps2, float& ax, float& ay, float& az,
                                          float& pot, uint32_t& cnt) {
  const bool take = j != i;
  if (take) {
    const float qx = q.x - p.x, qy = q.y - p.y, qz = q.z - p.z;
    if (EXACT) {
      const float dj2 = ((qx * qx + qy * qy) + qz * qz) + eps2;
      interact_exact(G * q.w, qx, qy, qz, dj2, ax, ay, az, pot);
    } else {
      const float dj2 =
          __builtin_fmaf(qx, qx, __builtin_fmaf(qy, qy, __builtin_fmaf(qz, qz, eps2)));
      interact_fast(q.w, qx, qy, qz, dj2, ax, ay, az, pot);
    }
    if (COUNT) cnt++;
  }
}

// default-path leaf member, no index test (eps > 0: j == i adds zero force)
__device__ __forceinline__ void pair_fast(float4 q, float4 p, float eps2, float& ax, float& ay,
                                          float& az, float& pot) {
  const float qx = q.x - p.x, qy = q.y - p.y, qz = q.z - p.z;
  const float dj2 = __builtin_fmaf(qx, qx, __builtin_fmaf(qy, qy, __builtin_fmaf(qz, qz, eps2)));
  interact_fast(q.w, qx, qy, qz, dj2, ax, ay, az, pot);
}

// ---- transposed leaves (default path) --------------------------------------
// A leaf opened by few lanes wastes most of the broadcast member loop (one
// member per iteration, lanes that did not open the leaf idle).  Instead the
// wavefront takes the (opening lane, member) pairs themselves: with the member
// count padded to cp = 2^LCP <= 32, lane L holds member j = L mod cp of the
// leaf (padding members weigh 0) and, per round, the particle of opening lane
// number base + L / cp (positions staged in LDS by rank); each segment of cp
// lanes sums its pairs with row-shift DPP adds into its last lane, which
// stores the sum in LDS by rank; after the rounds every opening lane adds its
// own sum.  64 / cp opening lanes per round.  The summation order differs
// from the reference's, so only the default (tolerance) path uses it;
// FASTLEAF semantics (the j == i pair adds r = 0) hold.
struct WaveLds {
  float4 pos[64];   // positions of the opening lanes, by rank
  float4 sum[64];   // their sums (ax, ay, az, pot), by rank
  float4 junk[64];  // the stores of lanes that do not end a segment
  float4 zero;      // read by the lanes that did not open the leaf
};

template <int SH>
__device__ __forceinline__ float row_shr_add(float v) {
  // v + the value of the lane SH below in the same 16-lane row (0 past the
  // row's first lane)
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 | SH, 0xf,
                                                        0xf, true));
}

__device__ __forceinline__ float row_bcast15_add(float v) {
  // rows 1 and 3 add the last lane of the row below (rows 0 and 2 add 0)
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xa, 0xf,
                                                        false));
}

// inclusive sums over segments of 2^LCP lanes: each segment's last lane ends
// with the sum of its segment
template <int LCP>
__device__ __forceinline__ void segment_sums(float& a, float& b, float& c, float& d) {
  if (LCP > 0) {
    a = row_shr_add<1>(a); b = row_shr_add<1>(b); c = row_shr_add<1>(c); d = row_shr_add<1>(d);
  }
  if (LCP > 1) {
    a = row_shr_add<2>(a); b = row_shr_add<2>(b); c = row_shr_add<2>(c); d = row_shr_add<2>(d);
  }
  if (LCP > 2) {
    a = row_shr_add<4>(a); b = row_shr_add<4>(b); c = row_shr_add<4>(c); d = row_shr_add<4>(d);
  }
  if (LCP > 3) {
    a = row_shr_add<8>(a); b = row_shr_add<8>(b); c = row_shr_add<8>(c); d = row_shr_add<8>(d);
  }
  if (LCP > 4) {
    a = row_bcast15_add(a); b = row_bcast15_add(b); c = row_bcast15_add(c);
    d = row_bcast15_add(d);
  }
}

// the transposed form pays off when its rounds cost less than the broadcast
// loop's count iterations (per-round and per-leaf costs in vector
// instructions, measured on the headline workload)
#define GS_TR_ROUND 20u
#define GS_TR_LCP 4u
#define GS_TR_LEAF 12u
// The rule as a table, one entry per lane (read with v_readlane, so the
// choice costs the scalar unit one compare): lane c holds the most opening
// lanes k for which a leaf of c members is summed transposed (0: never).  A
// leaf of c <= 32 members padded to 2^lcp, opened by k lanes, takes
// rounds(k) = ceil(k / per) rounds of 64 / 2^lcp = per opening lanes, and
//   rounds(k) * (GS_TR_ROUND + GS_TR_LCP lcp) + GS_TR_LEAF < 14 c
//   <=>  ceil(k / per) <= R  <=>  k <= R per,
// R = floor((14 c - GS_TR_LEAF - 1) / (GS_TR_ROUND + GS_TR_LCP lcp)).
__device__ __forceinline__ uint32_t tr_kmax(uint32_t c) {
  if (c < 2u || c > 32u) return 0u;
  const uint32_t lcp = 32u - (uint32_t)__builtin_clz(c - 1u);
  const uint32_t per = 64u >> lcp;
  const uint32_t r = (14u * c - GS_TR_LEAF - 1u) / (GS_TR_ROUND + GS_TR_LCP * lcp);
  const uint32_t k = r * per;
  return k > 64u ? 64u : k;
}

// the rule for a leaf of count members opened by the lanes `open`; trk is
// tr_kmax of the lane's own index
__device__ __forceinline__ bool leaf_transposes(uint64_t open, uint32_t count, uint32_t trk) {
  const uint32_t kmax = (uint32_t)__builtin_amdgcn_readlane((int)trk, (int)(count <= 32u ? count : 0u));
  return (uint32_t)__builtin_popcountll(open) <= kmax;
}

// the rounds of one leaf [first, first + count) for the opening lanes
// `open` (every lane of the wavefront executes this; 2 <= count <= 2^LCP <=
// 32): each opening lane's sum ends in L->sum[its rank]
template <int LCP>
__device__ __forceinline__ void leaf_rounds(const float4* __restrict__ posm, uint32_t first,
                                            uint32_t count, uint32_t k, uint32_t lane,
                                            float eps2, WaveLds* L) {
  constexpr uint32_t CP = 1u << LCP, PER = 64u >> LCP;
  const uint32_t j = lane & (CP - 1u), g = lane >> LCP;
  float4 q = posm[first + (j < count ? j : count - 1u)];
  if (j >= count) q.w = 0.f;  // padding member: no force, no potential
  // base + g <= 63: base is a multiple of PER below k <= 64.  Segments past
  // the last opening lane read stale positions; nobody reads their sums.
  for (uint32_t base = 0; base < k; base += PER) {
    const float4 pi = L->pos[base + g];
    float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f;
    pair_fast(q, pi, eps2, sx, sy, sz, sp);
    segment_sums<LCP>(sx, sy, sz, sp);
    // every lane stores, with no branch, so the sums stay live in every lane
    // and each DPP stage is one fused add: a segment's last lane into the
    // sums, the others into their own junk slot
    float4* dst = j == CP - 1u ? &L->sum[base + g] : &L->junk[lane];
    *dst = make_float4(sx, sy, sz, sp);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// One particle per lane, the lanes of a wavefront walking one node sequence.
// FASTLEAF (default path with eps > 0 and a finite self term): leaf members
// run without the j == i test, since that pair adds an exact zero to the
// acceleration (r = 0, f finite); the self potential -m_i / eps is taken back
// once per leaf.  The
// kernel picks the instance once per block, so the node loop carries no
// test of it.
template <bool EXACT, bool COUNT, bool FASTLEAF, bool GENERIC, bool OWNOPEN>
__device__ __forceinline__ void walk_particle(const NodeRec* __restrict__ nodes,
                                              const float* __restrict__ nsize,
                                              const float4* __restrict__ posm, float4 p,
                                              uint32_t i, bool valid, float G,
                                              float eps2, uint32_t qsize,
                                              float& ax, float& ay, float& az, float& pot,
                                              float& dens, uint32_t& cnt, WaveLds* wl) {
  int32_t alim = valid ? 0 : -1;  // active depth limit: active at depth d iff d <= alim
  // this lane's entry of the transposed-leaf table
  const uint32_t trk =
      tr_kmax(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
  const float selfpot = FASTLEAF ? p.w * __builtin_amdgcn_rsqf(eps2) : 0.f;
  int32_t nd = __builtin_amdgcn_ballot_w64(valid) ? 0 : -1;  // node link (record byte offset)
  // the wavefront's execution mask in the node loop (every lane of the
  // block's wavefronts; the loop holds no divergent branch around a visit)
  const uint64_t full = __builtin_amdgcn_read_exec();
  // the kernel's own scalar loads (parameters, slot) complete here, so the
  // loop below carries no wait for them
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  while (nd >= 0) {
    // one 32-byte scalar load at the node link, a 32-bit byte offset (a
    // slot's node pool holds at most GS_NODE32_MAX = 2^26 records): the load
    // takes the link register directly, with no shift or 64-bit address
    // arithmetic on the scalar unit, which the walk keeps nearly as busy as
    // the vector unit
    // (written as one instruction that waits for its data in the same asm
    // block: left to itself the compiler may split it into narrower loads
    // and re-load words inside the branches)
    u32x8 w;
    asm("s_load_dwordx8 %0, %1, %2\n\ts_waitcnt lgkmcnt(0)"
        : "=s"(w)
        : "s"(nodes), "s"(nd));
    const float4 c = make_float4(__uint_as_float(w[0]), __uint_as_float(w[1]),
                                 __uint_as_float(w[2]), __uint_as_float(w[3]));
    const int4 inf = make_int4((int)w[4], (int)w[5], (int)w[6], (int)w[7]);
    const uint32_t meta = (uint32_t)inf.z;
    const uint32_t depth = gs_meta_depth(meta);
    const int32_t skip = inf.w;
    // a massless node is skipped by every lane (nbody_bh.wgsl:86): its depth
    // test is made to fail on the scalar unit ((mass bits << 1) == 0).  Only
    // the GENERIC variant tests it: a tree without massless nodes has none.
    int32_t dtest = (int32_t)depth;
    if (GENERIC) {
      uint32_t mbits;
      asm("s_lshl_b32 %0, %1, 1" : "=s"(mbits) : "s"(__float_as_uint(c.w)));
      dtest = mbits != 0u ? (int32_t)depth : 64;
    }
    const float thr = __int_as_float(inf.x);  // opening threshold: accept iff d2 > thr
    const bool leaf = gs_meta_is_leaf(meta);
    const uint32_t first = (uint32_t)inf.y;  // leaf: first member; internal: last child
    const float rx = c.x - p.x, ry = c.y - p.y, rz = c.z - p.z;
    const float d2 = ((rx * rx + ry * ry) + rz * rz) + eps2;  // exact: drives the decision
    // monopole iff s2 < theta2*d2 || s2 < eps2/4 (nbody_bh.wgsl:97-112).  The
    // first test is one compare per lane, d2 > thr, with the record's exact
    // threshold of s2 and theta2 (false for a NaN d2, as in the reference);
    // the second depends on the node alone and is an integer compare of
    // float bits on the scalar unit, made only when some node of the tree is
    // that small (GENERIC).  Every other active lane opens the node.
    uint64_t act, need_mask;
    if (!EXACT && !COUNT && !GENERIC) {
      // the two compares write the execution mask themselves (v_cmpx writes
      // its result to an SGPR pair and to EXEC): EXEC = active lanes, then
      // EXEC = active lanes accepting the monopole, whose interaction
      // (interact_fast, same instructions and order) runs under that mask;
      // EXEC is restored before the block ends.  Saves the scalar AND and
      // save-exec of the compiled branch on every visit.
      uint64_t mono;
      float inv, mi;
      asm volatile(
          "v_cmpx_le_i32_e64 %[act], %[dep], %[alim]\n\t"
          "v_cmpx_lt_f32_e64 %[mono], %[thr], %[d2]\n\t"
          "s_cbranch_execz 1f\n\t"
          "v_rsq_f32_e32 %[inv], %[d2]\n\t"
          "s_nop 0\n\t"
          "v_mul_f32_e32 %[mi], %[m], %[inv]\n\t"
          "v_mul_f32_e32 %[inv], %[inv], %[inv]\n\t"
          "v_mul_f32_e32 %[inv], %[mi], %[inv]\n\t"
          "v_fmac_f32_e32 %[ax], %[rx], %[inv]\n\t"
          "v_fmac_f32_e32 %[ay], %[ry], %[inv]\n\t"
          "v_fmac_f32_e32 %[az], %[rz], %[inv]\n\t"
          "v_sub_f32_e32 %[pot], %[pot], %[mi]\n"
          "1:\n\t"
          "s_mov_b64 exec, %[full]"
          : [act] "=&s"(act), [mono] "=&s"(mono), [inv] "=&v"(inv), [mi] "=&v"(mi),
            [ax] "+v"(ax), [ay] "+v"(ay), [az] "+v"(az), [pot] "+v"(pot)
          : [dep] "s"(dtest), [alim] "v"(alim), [thr] "s"(thr), [d2] "v"(d2),
            [m] "s"(c.w), [rx] "v"(rx), [ry] "v"(ry), [rz] "v"(rz), [full] "s"(full)
          : "scc");
      need_mask = act & ~mono;
    } else {
      // the lanes that opened every ancestor, as a wave mask (one compare)
      act = __builtin_amdgcn_ballot_w64(dtest <= alim);
      uint64_t mono_raw = __builtin_amdgcn_ballot_w64(thr < d2);
      if (GENERIC && (__float_as_uint(nsize[gs_node(nd)]) & 0x7fffffffu) < qsize)
        mono_raw = ~0ull;
      const uint64_t mono_mask = act & mono_raw;
      need_mask = act & ~mono_raw;
      if (__builtin_amdgcn_inverse_ballot_w64(mono_mask)) {
        if (EXACT)
          interact_exact(G * c.w, rx, ry, rz, d2, ax, ay, az, pot);
        else
          interact_fast(c.w, rx, ry, rz, d2, ax, ay, az, pot);
        if (COUNT) cnt++;
      }
    }
    // lanes that descend into the node's children (none for a leaf)
    uint64_t desc_mask = need_mask;
    if (leaf) {
      desc_mask = 0;
      const uint32_t count = gs_meta_count(meta);
      // the leaf holding a lane's own particle sets its density.  FASTLEAF: a
      // lane that opens its own leaf takes back the self potential that the
      // member loop below adds.  OWNOPEN (no lane can accept its own leaf, see
      // walk_kernel): only a leaf that some lane opens can be a lane's own.
      auto own_leaf = [&]() {
        const uint64_t own = act & __builtin_amdgcn_ballot_w64(i - first < count);
        if (own != 0) {  // a scalar test first: most leaves hold no lane's particle
          if (__builtin_amdgcn_inverse_ballot_w64(own)) {
            const float sz = nsize[gs_node(nd)];
            dens = (float)count / ((sz * sz) * sz);
            if (FASTLEAF && __builtin_amdgcn_inverse_ballot_w64(need_mask)) pot = pot + selfpot;
          }
        }
      };
      if (!OWNOPEN) own_leaf();
      uint32_t lcp = 0;
      if (need_mask == 0) {
        // no lane opens the leaf
      } else {
      if (OWNOPEN) own_leaf();
      // FASTLEAF: the transposed form's sums come back in d, added below
      // where the forms meet (d = 0 after the broadcast loop, which adds into
      // the accumulators member by member): the accumulators keep one
      // register on every path, so the forms merge without copies
      float dx = 0.f, dy = 0.f, dz = 0.f, dp = 0.f;
      if (FASTLEAF && leaf_transposes(need_mask, count, trk)) {
        lcp = 32u - (uint32_t)__builtin_clz(count - 1u);  // ceil(log2(count)), 1..5
        const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need_mask >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)need_mask, 0u));
        const bool mine = __builtin_amdgcn_inverse_ballot_w64(need_mask);
        const uint32_t k = (uint32_t)__builtin_popcountll(need_mask);
        // the opening lanes stage their positions by rank, the others write
        // their junk slot (an address select instead of a branch)
        float4* pdst = mine ? &wl->pos[rank] : &wl->junk[lane];
        *pdst = p;
        __builtin_amdgcn_wave_barrier();
#define GS_TR(L) leaf_rounds<L>(posm, first, count, k, lane, eps2, wl)
        if (lcp == 1) GS_TR(1);
        else if (lcp == 2) GS_TR(2);
        else if (lcp == 3) GS_TR(3);
        else if (lcp == 4) GS_TR(4);
        else GS_TR(5);
#undef GS_TR
        __builtin_amdgcn_wave_barrier();
        // each opening lane adds its own sum; the others add the zero slot (x +
        // 0 leaves x, up to the sign of a zero)
        const float4 t = *(mine ? &wl->sum[rank] : &wl->zero);
        dx = t.x;
        dy = t.y;
        dz = t.z;
        dp = t.w;
        __builtin_amdgcn_wave_barrier();
      } else if (__builtin_amdgcn_inverse_ballot_w64(need_mask)) {  // lanes opening the leaf
        const uint32_t last = first + count;
        uint32_t j = first;
        if (FASTLEAF) {
          // eps > 0: the j == i term adds an exact zero to the acceleration
          // (r = 0, f finite), so members run without a per-pair index test
          // and the self potential is taken back once per leaf.  Members are
          // addressed by 32-bit byte offsets (slot-local index < 2^27), so each
          // batch of 4 scalar loads shares one offset register.
          const char* pb = (const char*)posm;
          uint32_t off = first << 4;
          const uint32_t end = last << 4;
          for (; off + 64 <= end; off += 64) {
            const float4 q0 = *(const float4*)(pb + off), q1 = *(const float4*)(pb + off + 16);
            const float4 q2 = *(const float4*)(pb + off + 32), q3 = *(const float4*)(pb + off + 48);
            pair_fast(q0, p, eps2, ax, ay, az, pot);
            pair_fast(q1, p, eps2, ax, ay, az, pot);
            pair_fast(q2, p, eps2, ax, ay, az, pot);
            pair_fast(q3, p, eps2, ax, ay, az, pot);
          }
          const uint32_t rem = end - off;  // 0-3 more members
          if (rem >= 16) pair_fast(*(const float4*)(pb + off), p, eps2, ax, ay, az, pot);
          if (rem >= 32) pair_fast(*(const float4*)(pb + off + 16), p, eps2, ax, ay, az, pot);
          if (rem >= 48) pair_fast(*(const float4*)(pb + off + 32), p, eps2, ax, ay, az, pot);
        } else {
          for (; j + 4 <= last; j += 4) {
            const float4 q0 = posm[j], q1 = posm[j + 1], q2 = posm[j + 2], q3 = posm[j + 3];
            leaf_pair<EXACT, COUNT>(q0, j, i, p, G, eps2, ax, ay, az, pot, cnt);
            leaf_pair<EXACT, COUNT>(q1, j + 1, i, p, G, eps2, ax, ay, az, pot, cnt);
            leaf_pair<EXACT, COUNT>(q2, j + 2, i, p, G, eps2, ax, ay, az, pot, cnt);
            leaf_pair<EXACT, COUNT>(q3, j + 3, i, p, G, eps2, ax, ay, az, pot, cnt);
          }
          for (; j < last; j++)
            leaf_pair<EXACT, COUNT>(posm[j], j, i, p, G, eps2, ax, ay, az, pot, cnt);
        }
      }
      if (FASTLEAF) {
        ax += dx;
        ay += dy;
        az += dz;
        pot += dp;
      }
      }
    }
    // next node: an opened internal node continues at its last child
    // (visited first), leaves and unopened nodes after the subtree.  Only a
    // descent updates the depth limit, alim = min(alim, depth) + desc: the
    // opening lanes become active at the children's depth and the others
    // only above this node; after any other visit the next node is no deeper
    // than this one, where the update would change no lane's activity.
    uint64_t carry;
    asm("s_cmp_eq_u64 %[dm], 0\n\t"
        "s_cselect_b32 %[nd], %[skip], %[first]\n\t"
        "s_cbranch_scc1 1f\n\t"
        "v_min_i32_e32 %[alim], %[depth], %[alim]\n\t"
        "v_addc_co_u32_e64 %[alim], %[carry], 0, %[alim], %[dm]\n"
        "1:"
        : [nd] "=&s"(nd), [alim] "+v"(alim), [carry] "=&s"(carry)
        : [dm] "s"(desc_mask), [skip] "s"(skip), [first] "s"(first), [depth] "s"(depth)
        : "scc");
  }
}

template <bool EXACT, bool STEP, bool COUNT>
__global__ __launch_bounds__(1024) void walk_kernel(
    const float4* __restrict__ posm_all, const float4* __restrict__ velp_all,
    float4* __restrict__ posm_out_all, float4* __restrict__ velp_out_all,
    float* __restrict__ energy_all, float4* __restrict__ acc_all, float* __restrict__ dens_all,
    uint32_t* __restrict__ inter_all, const NodeRec* __restrict__ pool,
    const float* __restrict__ nsize_all, const DevSlot* __restrict__ slots, const uint32_t* __restrict__ wblock_slot,
    const DevParams* __restrict__ prms, const DevScalars* __restrict__ scs,
    const uint32_t* __restrict__ massless, DevBatch* batch, double* __restrict__ red,
    uint32_t fresh_tree) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  // transposed-leaf staging of this wavefront (dynamic LDS sized to the
  // block's wavefronts; only the default force path launches it)
  extern __shared__ WaveLds s_leaf[];
  WaveLds* wl = (EXACT || COUNT) ? nullptr : &s_leaf[w];
  if (wl && lane == 0) wl->zero = make_float4(0.f, 0.f, 0.f, 0.f);

  // a workgroup never spans two slots (slot ranges are GS_SLOT_ALIGN-aligned).
  // The hardware deals workgroups to the 8 XCDs round-robin (XCD = id % 8);
  // the work index is remapped so that each XCD takes runs of XCD_RUN
  // consecutive blocks (neighbouring particles share most of their tree, and
  // each XCD has its own L2), while the runs still interleave over the whole
  // particle range, so every XCD walks a mix of dense and sparse regions.
  // The remap is a permutation of the grid (a partial last group keeps the
  // identity order).  The work is then taken from the top of the particle
  // range down (the first workgroup walks the last block): measured faster
  // than ascending order on every workload larger than the launch latency
  // (N = 500,000 disk 3.7 %, 64-galaxy pool 2.0 %, refresh-only steps of
  // the 4,194,304 collision 2.5 %, its rebuilding steps 0.6 %), with the same
  // results (each block's walk is independent of the order).
  constexpr uint32_t XCD_RUN = 64;
  uint32_t blk = blockIdx.x;
  if (blk < gridDim.x / (8 * XCD_RUN) * (8 * XCD_RUN)) {
    const uint32_t r = blk % (8 * XCD_RUN);
    blk = blk - r + (r % 8) * XCD_RUN + r / 8;
  }
  blk = gridDim.x - 1u - blk;
  const uint32_t slot = wblock_slot[blk];
  const DevSlot sl = slots[slot];
  const DevParams* __restrict__ prm = prms + slot;
  const uint32_t n = sl.n;
  const float4* __restrict__ posm = posm_all + sl.poff;
  const NodeRec* __restrict__ nodes = pool + sl.noff;
  const uint32_t i = (blk - sl.wblock0) * blockDim.x + threadIdx.x;  // slot-local particle
  const bool valid = i < n;
  if ((blk - sl.wblock0) * blockDim.x >= n) {  // a block of empty slot capacity
    if (STEP && lane < GS_NRED) red[((size_t)blk * waves + w) * GS_NRED + lane] = 0.0;
    return;
  }
  const float4 p = valid ? posm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float G = prm->G, eps2 = prm->eps2;
  const float qeps2 = 0.25f * eps2;
  // the node sizes that pass s^2 < eps^2/4 are those below qsize (as bits)
  const uint32_t qsize = prm->qsize_bits;
  const float* __restrict__ nsize = nsize_all + sl.noff;

  // the eps^2/4 test can only accept a node if some node of the tree has
  // size^2 below eps^2/4 (never at depth <= 10 for the usual eps), and the
  // massless-node test only matters if the tree holds one: the walk runs
  // without both tests unless the slot needs either
  const bool generic = scs[slot].min_s2 < qeps2 || massless[slot] != 0u;

  // OWNOPEN: no lane can accept the leaf holding its own particle as a
  // monopole, so only a leaf that some lane opens can be a lane's own (the
  // own-leaf density test then runs on opened leaves only).  When the tree
  // was built from the current positions (fresh_tree), a particle lies in
  // its own leaf's cell, at most sqrt(3) size from the leaf's COM, and
  // accepting needs s^2 < theta^2 (3 s^2 + eps^2); no node of size^2 >=
  // min_s2 passes that when 3 theta^2 < 1 and min_s2 (1 - 3 theta^2) >=
  // theta^2 eps^2.  The margins (3.03, 1.01) cover the f32 roundings of d^2,
  // of the COM and of the differences, the last bounded by requiring
  // coordinates |x| <= M with M^2 <= 1e9 min_s2 (rounding of order M 2^-24
  // against a cell of size >= sqrt(min_s2)).  NaN anywhere fails a compare.
  bool ownopen = false;
  if (!EXACT && !COUNT && fresh_tree && !generic) {
    const DevScalars* sc = scs + slot;
    const float th2 = prm->theta2, t3 = 3.03f * th2, rs = sc->level_size[0];
    float m = 0.f;
    for (int a = 0; a < 3; a++)
      m = fmaxf(m, fmaxf(fabsf(sc->min_f32[a]), fabsf(sc->min_f32[a] + rs)));
    ownopen = t3 < 1.0f && sc->min_s2 * (1.0f - t3) >= 1.01f * th2 * eps2 &&
              m * m <= 1e9f * sc->min_s2;
  }

  // FASTLEAF adds a lane's own j == i pair as r * f with r = 0, an exact zero
  // only while f = m rsq(eps2)^3 (and m rsq(eps2)) is finite; a huge mass or
  // a tiny eps makes it Inf and the zero a NaN, where the reference skips the
  // pair.  The block takes FASTLEAF only if every lane's self term is finite.
  bool fastleaf = false;
  if (!EXACT && !COUNT && eps2 > 0.0f) {
    const float inv = __builtin_amdgcn_rsqf(eps2);
    const float mi = p.w * inv;
    const float f = mi * (inv * inv);
    fastleaf = __syncthreads_and(__builtin_isfinite(mi) && __builtin_isfinite(f)) != 0;
  }

  float ax = 0.f, ay = 0.f, az = 0.f, pot = 0.f, dens = 0.f;
  uint32_t cnt = 0;
#define GS_WALK(FL, QC, OO)                                                                  \
  walk_particle<EXACT, COUNT, FL, QC, OO>(nodes, nsize, posm, p, i, valid, G, eps2, qsize,   \
                                          ax, ay, az, pot, dens, cnt, wl)
  if (fastleaf) {
    if (generic) GS_WALK(true, true, false);
    else if (ownopen) GS_WALK(true, false, true);
    else GS_WALK(true, false, false);
  } else {
    if (generic) GS_WALK(false, true, false);
    else if (ownopen) GS_WALK(false, false, true);
    else GS_WALK(false, false, false);
  }
#undef GS_WALK
  if (!EXACT) {
    ax *= G;
    ay *= G;
    az *= G;
    pot *= G;
  }

  if (COUNT) {
    if (valid) inter_all[sl.poff + i] = cnt;
    uint64_t s = cnt;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0 && s)
      atomicAdd((unsigned long long*)&batch->interactions, (unsigned long long)s);
  }

  if (!STEP) {
    if (valid) {
      acc_all[sl.poff + i] = make_float4(ax, ay, az, pot);
      dens_all[sl.poff + i] = dens;
    }
    return;
  }

  // ---- halos, kick-drift, gas drag, energy (nbody_bh.wgsl:144-180) --------
  double r_ekin = 0, r_epot = 0, r_ehalo = 0, r_px = 0, r_py = 0, r_pz = 0, r_espec = 0, r_m = 0;
  if (valid) {
    const float4 v = velp_all[sl.poff + i];
    float pot_halo = 0.f;
    if (prm->halo_mass > 0.0f) {
      const float gm = G * prm->halo_mass;
      for (uint32_t h = 0; h < prm->halo_count; h++) {
        const float4 ctr = h == 0 ? prm->halo_c1 : prm->halo_c2;
        const float dx = p.x - ctr.x, dy = p.y - ctr.y, dz = p.z - ctr.z;
        const float len = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
        const float rlen = fmaxf(len, 1e-6f);
        const float ra = rlen + prm->halo_a;
        const float t = gm / ((ra * ra) * rlen);
        ax = ax - dx * t;
        ay = ay - dy * t;
        az = az - dz * t;
        pot_halo = pot_halo - gm / ra;
      }
    }
    const float dt = prm->dt;
    float vx = v.x + ax * dt, vy = v.y + ay * dt, vz = v.z + az * dt;
    const bool gas = v.w < -0.5f;
    if (gas && prm->gas_drag > 0.0f) {
      float hx = -p.y, hy = p.x, hz = 1e-6f;
      const float hl = __builtin_sqrtf((hx * hx + hy * hy) + hz * hz);
      hx = hx / hl;
      hy = hy / hl;
      hz = hz / hl;
      const float s = (vx * hx + vy * hy) + vz * hz;
      const float fx = hx * s, fy = hy * s, fz = hz * s;
      const float k = prm->drag_damp;
      vx = fx + (vx - fx) * k;
      vy = fy + (vy - fy) * k;
      vz = fz + (vz - fz) * k;
    }
    posm_out_all[sl.poff + i] = make_float4(p.x + vx * dt, p.y + vy * dt, p.z + vz * dt, p.w);
    velp_out_all[sl.poff + i] = make_float4(vx, vy, vz, gas ? -(1.0f + dens) : 0.0f);
    const float v2 = (vx * vx + vy * vy) + vz * vz;
    const float e = (0.5f * v2 + 0.5f * pot) + pot_halo;
    energy_all[sl.poff + i] = e;
    const double m = (double)p.w;
    r_ekin = 0.5 * m * (double)v2;
    r_epot = 0.5 * m * (double)pot;
    r_ehalo = m * (double)pot_halo;
    r_px = m * (double)vx;
    r_py = m * (double)vy;
    r_pz = m * (double)vz;
    r_espec = (double)e;
    r_m = m;
  }
  // one set of partial sums per wavefront (fixed order: butterfly within the
  // wavefront, then the wavefronts in index order in reduce_final_kernel), so
  // no wavefront waits for the others of its workgroup
  double acc[GS_NRED] = {r_ekin, r_epot, r_ehalo, r_px, r_py, r_pz, r_espec, r_m};
#pragma unroll
  for (int k = 0; k < GS_NRED; k++) acc[k] = wave_sum(acc[k]);
  if (lane == 0) {
    double* dst = red + ((size_t)blk * waves + w) * GS_NRED;
#pragma unroll
    for (int k = 0; k < GS_NRED; k++) dst[k] = acc[k];
  }
}

__global__ __launch_bounds__(256) void reduce_final_kernel(const double* __restrict__ red,
                                                           uint32_t b0, uint32_t nblocks,
                                                           double* out) {
  __shared__ double s[4][GS_NRED];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a[GS_NRED] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x)
    for (int k = 0; k < GS_NRED; k++) a[k] += red[(size_t)(b0 + b) * GS_NRED + k];
  for (int k = 0; k < GS_NRED; k++) a[k] = wave_sum(a[k]);
  if (lane == 0)
    for (int k = 0; k < GS_NRED; k++) s[w][k] = a[k];
  __syncthreads();
  if (threadIdx.x < GS_NRED) out[threadIdx.x] = ((s[0][threadIdx.x] + s[1][threadIdx.x]) +
                                                 s[2][threadIdx.x]) + s[3][threadIdx.x];
}

template <bool EXACT, bool STEP, bool COUNT>
hipError_t launch(gs_sim* s) {
  const int c = s->cur;
  const size_t lds = (EXACT || COUNT) ? 0 : (size_t)(s->walk_wg / 64) * sizeof(WaveLds);
  walk_kernel<EXACT, STEP, COUNT><<<s->walk_blocks, s->walk_wg, lds, s->stream>>>(
      s->posm[c], s->velp[c], s->posm[c ^ 1], s->velp[c ^ 1], s->energy, s->acc, s->dens,
      s->inter, s->nodes, s->nsize, s->slots, s->wblock_slot, s->prm, s->sc, s->massless,
      s->batch, s->red_part, s->walk_fresh ? 1u : 0u);
  return hipGetLastError();
}

}  // namespace

hipError_t gs_launch_walk(gs_sim* s, bool integrate, bool count) {
  const bool exact = (s->flags & GS_FLAG_EXACT_FORCE) != 0;
  if (count) return exact ? launch<true, false, true>(s) : launch<false, false, true>(s);
  if (integrate) return exact ? launch<true, true, false>(s) : launch<false, true, false>(s);
  return exact ? launch<true, false, false>(s) : launch<false, false, false>(s);
}

hipError_t gs_launch_reduce_final(gs_sim* s, uint32_t slot, double* out_dev) {
  const HostSlot& h = s->hs[slot];
  // the slot's wavefronts: walk_wg / 64 per walk block
  const uint32_t wpb = s->walk_wg / 64;
  reduce_final_kernel<<<1, 256, 0, s->stream>>>(s->red_part, h.d.wblock0 * wpb, h.d.nwblocks * wpb,
                                                out_dev);
  return hipGetLastError();
}
