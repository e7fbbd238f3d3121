// Host <-> device movement of 32-byte particle records for one slot.
//
// The device keeps particles as two 16-byte arrays (posm, velp); the host
// side uses interleaved records [x,y,z,m, vx,vy,vz,pad] (nbody_bh.wgsl:10-17).
// A strided host copy with 16-byte rows is one tiny transfer per row, so
// records instead cross PCIe as one contiguous block staged in the slot's
// range of the non-current ping-pong buffers (dead data between steps: every
// step writes them before reading), and a kernel splits or packs them:
//   staging float4 j of the byte stream lives in sa[j] for j < cap, else in
//   sb[j - cap]  (sa, sb = the slot's ranges of posm[cur^1], velp[cur^1]);
//   record i = stream float4s 2i (position + mass) and 2i+1 (velocity + pad).
#include "gs_internal.h"

namespace {

__device__ __forceinline__ float4 stream_ld(const float4* sa, const float4* sb, uint32_t cap,
                                            uint32_t j) {
  return j < cap ? sa[j] : sb[j - cap];
}

__global__ __launch_bounds__(256) void split_records_kernel(const float4* __restrict__ sa,
                                                            const float4* __restrict__ sb,
                                                            uint32_t cap, uint32_t n,
                                                            float4* __restrict__ posm,
                                                            float4* __restrict__ velp) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    posm[i] = stream_ld(sa, sb, cap, 2 * i);
    velp[i] = stream_ld(sa, sb, cap, 2 * i + 1);
  }
}

__global__ __launch_bounds__(256) void pack_records_kernel(const float4* __restrict__ posm,
                                                           const float4* __restrict__ velp,
                                                           uint32_t cap, uint32_t n,
                                                           float4* __restrict__ sa,
                                                           float4* __restrict__ sb) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t j = 2 * i;
    const float4 p = posm[i], v = velp[i];
    if (j < cap) sa[j] = p;
    else sb[j - cap] = p;
    if (j + 1 < cap) sa[j + 1] = v;
    else sb[j + 1 - cap] = v;
  }
}

// compact_positions (nbody_bh.wgsl:186-195): (x, y, z, energy)
__global__ __launch_bounds__(256) void pack_positions_kernel(const float4* __restrict__ posm,
                                                             const float* __restrict__ energy,
                                                             uint32_t n,
                                                             float4* __restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = posm[i];
    out[i] = make_float4(p.x, p.y, p.z, energy[i]);
  }
}

// where the particle that had index `old` before the last sort is now:
// perm[new] = old (remapIndex of the frame loop, main.js:436-440)
__global__ __launch_bounds__(256) void locate_kernel(const uint32_t* __restrict__ perm,
                                                     uint32_t n, uint32_t old,
                                                     uint32_t* __restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (perm[i] == old) *out = i;
}

inline uint32_t grid_of(uint32_t n) {
  uint32_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  return g ? g : 1;
}

}  // namespace

hipError_t gs_upload_records(gs_sim* s, uint32_t slot, const float* host, uint32_t n) {
  const DevSlot& d = s->hs[slot].d;
  if (n == 0) return hipSuccess;
  float4* sa = s->posm[s->cur ^ 1] + d.poff;
  float4* sb = s->velp[s->cur ^ 1] + d.poff;
  const size_t bytes = 32ull * n, first = bytes < 16ull * d.cap ? bytes : 16ull * d.cap;
  hipError_t e = hipMemcpyAsync(sa, host, first, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess && bytes > first)
    e = hipMemcpyAsync(sb, (const char*)host + first, bytes - first, hipMemcpyHostToDevice,
                       s->stream);
  if (e != hipSuccess) return e;
  split_records_kernel<<<grid_of(n), 256, 0, s->stream>>>(sa, sb, d.cap, n, s->posm[s->cur] + d.poff,
                                                          s->velp[s->cur] + d.poff);
  return hipGetLastError();
}

hipError_t gs_download_records(gs_sim* s, uint32_t slot, float* host, uint32_t n) {
  const DevSlot& d = s->hs[slot].d;
  if (n == 0) return hipSuccess;
  float4* sa = s->posm[s->cur ^ 1] + d.poff;
  float4* sb = s->velp[s->cur ^ 1] + d.poff;
  pack_records_kernel<<<grid_of(n), 256, 0, s->stream>>>(s->posm[s->cur] + d.poff,
                                                         s->velp[s->cur] + d.poff, d.cap, n, sa, sb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t bytes = 32ull * n, first = bytes < 16ull * d.cap ? bytes : 16ull * d.cap;
  e = hipMemcpyAsync(host, sa, first, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess && bytes > first)
    e = hipMemcpyAsync((char*)host + first, sb, bytes - first, hipMemcpyDeviceToHost, s->stream);
  return e;
}

hipError_t gs_download_positions(gs_sim* s, uint32_t slot, float* host4, uint32_t n) {
  const DevSlot& d = s->hs[slot].d;
  if (n == 0) return hipSuccess;
  float4* sa = s->posm[s->cur ^ 1] + d.poff;
  pack_positions_kernel<<<grid_of(n), 256, 0, s->stream>>>(s->posm[s->cur] + d.poff,
                                                           s->energy + d.poff, n, sa);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(host4, sa, 16ull * n, hipMemcpyDeviceToHost, s->stream);
}

hipError_t gs_launch_locate(gs_sim* s, uint32_t slot, uint32_t old, uint32_t* out_dev) {
  const DevSlot& d = s->hs[slot].d;
  hipError_t e = gs_launch_fill_u32(out_dev, 0xffffffffu, 1, s->stream);
  if (e != hipSuccess) return e;
  locate_kernel<<<grid_of(d.n), 256, 0, s->stream>>>(s->vals[gs_sorted_buf(s)] + d.poff, d.n, old,
                                                      out_dev);
  return hipGetLastError();
}
