// Star formation and stellar recycling — star_formation (nbody_bh.wgsl:
// 197-268), run once per frame in place on the current particle and stellar
// records of every slot (encodeStarFormation, NBodyGPU.js:277-285).
//
// Stellar record per particle: float4 (stellar mass, birth time in Myr, seed,
// dust), kept in its own ping-pong pair that flips only when the particles
// are permuted (the sort's gather and gs_apply_permutation carry it along).
//
//   gas (pad < -0.5) denser than sfDensity: becomes a star with probability
//     sfEff * sfDt * min(sqrt(density / max(sfDensity, 1e-6)), 3); the mass is
//     a Kroupa-IMF draw, birth = time * agingRate, seed = a fresh uniform
//   star of >= 8 solar masses, older than main sequence + giant phase + 45 Myr
//     (tMS = 10000 m^-2.5, tGiant = 0.12 tMS + 15): back to gas with a dusty
//     stellar record and a velocity kick uniform in [-1, 1)^3
// Randomness is a counter hash of (particle index, salt = u32(time * 977) + 1
// + stream offset), the PCG-RXS-M-XS 32-bit output permutation.
//
// One thread per particle; HBM: 48 B read + at most 48 B written per particle.
#include <math.h>

#include "gs_internal.h"

namespace {

__device__ __forceinline__ uint32_t pcg_hash(uint32_t v0) {
  const uint32_t v = v0 * 747796405u + 2891336453u;
  const uint32_t w = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
  return (w >> 22u) ^ w;
}

// f32(hash) / 4294967295.0, whose f32 value is 2^32
__device__ __forceinline__ float rand01(uint32_t i, uint32_t salt) {
  return (float)pcg_hash(i ^ (salt * 0x9E3779B9u)) / 4294967296.0f;
}

// Kroupa (2001) inverse CDF: slope 1.3 on [0.08, 0.5], 2.3 on [0.5, 100]
// (nbody_bh.wgsl:216-227).  The shader's constant powers and exponents are
// constant expressions: their f32 roundings of the exact values are used
// here; the remaining arithmetic is f32, left to right.
__device__ __forceinline__ float imf_sample(float u1, float u2) {
  const float a0 = 0x1.11135ep+1f;   // f32(0.08^-0.3)
  const float b0 = 0x1.3b2c48p+0f;   // f32(0.5^-0.3)
  const float a1 = 0x1.3b2c48p+1f;   // f32(0.5^-1.3)
  const float b1 = 0x1.493cecp-9f;   // f32(100^-1.3)
  const float e_low = -0x1.aaaaaap+1f;  // f32(-1 / 0.3)
  const float e_hi = -0x1.89d89ep-1f;   // f32(-1 / 1.3)
  const float low = (a0 - b0) / 0.3f;
  const float hi = 0.5f * (a1 - b1) / 1.3f;
  if (u1 < low / (low + hi)) return powf(a0 - u2 * (a0 - b0), e_low);
  return powf(a1 - u2 * (a1 - b1), e_hi);
}

__global__ __launch_bounds__(256) void star_formation_kernel(
    float4* __restrict__ velp_all, float4* __restrict__ stel_all,
    const DevSlot* __restrict__ slots, const uint32_t* __restrict__ wblock_slot, uint32_t wg,
    const DevParams* __restrict__ prms) {
  // the walk blocks' particles: wg per block, each slot's aligned span
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t blk = q / wg;
  const uint32_t slot = wblock_slot[blk];
  const DevSlot sl = slots[slot];
  const uint32_t i = (blk - sl.wblock0) * wg + q % wg;  // slot-local index
  if (i >= sl.n) return;
  const DevParams& prm = prms[slot];
  float4* velp = velp_all + sl.poff;
  float4* stel = stel_all + sl.poff;
  const uint32_t salt = (uint32_t)(prm.time * 977.0f) + 1u;
  const float4 v = velp[i];
  const float4 st = stel[i];
  if (v.w < -0.5f) {
    const float density = -v.w - 1.0f;
    if (density > prm.sf_density && prm.sf_eff > 0.0f) {
      const float prob = prm.sf_eff * prm.sf_dt *
                         fminf(__builtin_sqrtf(density / fmaxf(prm.sf_density, 1e-6f)), 3.0f);
      if (rand01(i, salt) < prob) {
        const float m = imf_sample(rand01(i, salt + 7u), rand01(i, salt + 13u));
        const float birth = prm.time * prm.aging_rate;
        stel[i] = make_float4(m, birth, rand01(i, salt + 29u), 0.0f);
        velp[i].w = 0.0f;  // now a star
      }
    }
  } else if (st.x >= 8.0f) {
    const float age = prm.time * prm.aging_rate - st.y;
    const float tms = 10000.0f * powf(st.x, -2.5f);
    const float tgiant = tms * 0.12f + 15.0f;
    if (age > ((tms + tgiant) + 15.0f) + 30.0f) {
      stel[i] = make_float4(-1.0f, 0.0f, st.z, 0.7f);  // gas again, dusty
      const float kx = (rand01(i, salt + 41u) - 0.5f) * 2.0f;
      const float ky = (rand01(i, salt + 43u) - 0.5f) * 2.0f;
      const float kz = (rand01(i, salt + 47u) - 0.5f) * 2.0f;
      velp[i] = make_float4(v.x + kx, v.y + ky, v.z + kz, -1.0f);
    }
  }
}

}  // namespace

hipError_t gs_launch_star_formation(gs_sim* s) {
  if (!s->stel[0]) return hipSuccess;
  // the walk blocks cover whole aligned spans (multiples of 1024 particles)
  const uint32_t blocks = (uint32_t)((uint64_t)s->walk_blocks * s->walk_wg / 256);
  if (!blocks) return hipSuccess;
  star_formation_kernel<<<blocks, 256, 0, s->stream>>>(
      s->velp[s->cur], s->stel[s->stel_cur], s->slots, s->wblock_slot,
      s->walk_wg, s->prm);
  return hipGetLastError();
}
