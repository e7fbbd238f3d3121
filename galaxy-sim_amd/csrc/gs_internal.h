// Internal device-side layout of one galaxy (gs_sim) and the kernel
// launchers shared by the translation units of libgalaxy_hip.so.
//
// HBM layout (structure of arrays, all 16-byte aligned):
//   posm[2][N]  float4 (x, y, z, m)      ping-pong particle positions + mass
//   velp[2][N]  float4 (vx, vy, vz, pad) ping-pong velocities + gas flag
//   keys[2][N], vals[2][N]   u32         radix-sort ping-pong (code, index)
//   node[C]     32-byte record (one s_load_dwordx8 in the walk):
//               float4 cm   (comx, comy, comz, mass)
//               int4   info (size bits, first | firstChild, meta, skip)
//               meta: bit 31 leaf, bits 27..30 depth, then
//                     leaf: bits 0..26 particle count
//                     internal: bits 0..3 child count, bits 8..15 octant mask
//               skip: the node that follows this node's subtree in the
//                     walk's pre-order (previous sibling, else the parent's
//                     skip; -1 ends the walk)
// with C = node capacity >= 2N + 4096 (mortonTree.js:88).  Children of a BFS
// node are contiguous, so an internal node stores (firstChild, nChild, mask)
// instead of 8 slots; gs_download_nodes expands to the 64-byte layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/galaxy_hip.h"

#define GS_MAX_LEVEL 10   // mortonTree.js:26
#define GS_MAX_LEVELS 16  // level slots (NBodyGPU.js:171-174)

// status bits raised by kernels, checked by the host at sync points
#define GS_ST_NODE_POOL 0x1u
#define GS_ST_STACK 0x2u
#define GS_ST_SPIN 0x4u   // a look-back wait hit its bound (never expected)
#define GS_SPIN_LIMIT (1u << 24)

// Params uniform (nbody_bh.wgsl:19-42) in f32 plus derived per-step values.
struct DevParams {
  float dt, G, eps2;
  uint32_t n;
  float theta, halo_mass, halo_a, gas_drag;
  float time, aging_rate, sf_density, sf_eff;
  float sf_dt;
  uint32_t halo_count;
  float theta2;     // f32(theta * theta) as the WGSL evaluates it
  float drag_damp;  // f32 exp(-gasDrag * dt), evaluated once on the host
  float4 halo_c1, halo_c2;
};

// Sort digit width and tiling.
#define GS_RADIX_BITS 8
#define GS_RADIX 256
#define GS_SORT_PASSES 4
#define GS_SORT_BLOCK 256
#define GS_SORT_ITEMS 16
#define GS_SORT_TILE (GS_SORT_BLOCK * GS_SORT_ITEMS)

#define GS_BUILD_TILE 256
#define GS_BBOX_BLOCKS 1024

// Device-resident scalars of one galaxy.
struct DevScalars {
  uint64_t min_bits[3];   // bbox minimum (f32 widened to binary64 bits)
  uint64_t s_bits;        // 1024 / size  (binary64 bits)
  uint64_t size_bits;     // rootSize     (binary64 bits)
  uint32_t nfinite;       // particles that entered the bbox
  uint32_t status;        // GS_ST_* bits
  uint32_t num_nodes;
  uint32_t num_levels;
  uint32_t level_start[GS_MAX_LEVELS + 1];
  uint32_t level_count[GS_MAX_LEVELS + 1];
  float level_size[GS_MAX_LEVELS + 1];
  uint32_t tickets[32];   // [0..3] sort passes, [4..20] build levels
  uint32_t sort_hist[GS_SORT_PASSES][GS_RADIX];
  uint64_t interactions;
};

struct __align__(32) NodeRec {
  float4 cm;
  int4 info;
};

// node meta word
#define GS_META_LEAF 0x80000000u
#define GS_MAX_N ((1u << 27) - 1)  // a leaf count must fit the 27-bit field
__host__ __device__ __forceinline__ uint32_t gs_meta_leaf(uint32_t depth, uint32_t count) {
  return GS_META_LEAF | (depth << 27) | count;
}
__host__ __device__ __forceinline__ uint32_t gs_meta_internal(uint32_t depth, uint32_t mask,
                                                             uint32_t nc) {
  return (depth << 27) | (mask << 8) | nc;
}
// a node emitted by its parent, not yet classified: depth + particle count
__host__ __device__ __forceinline__ uint32_t gs_meta_open(uint32_t depth, uint32_t count) {
  return (depth << 27) | count;
}
__host__ __device__ __forceinline__ bool gs_meta_is_leaf(uint32_t m) { return (m >> 31) != 0; }
__host__ __device__ __forceinline__ uint32_t gs_meta_depth(uint32_t m) { return (m >> 27) & 0xfu; }
__host__ __device__ __forceinline__ uint32_t gs_meta_count(uint32_t m) { return m & 0x7ffffffu; }
__host__ __device__ __forceinline__ uint32_t gs_meta_nc(uint32_t m) { return m & 0xfu; }
__host__ __device__ __forceinline__ uint32_t gs_meta_mask(uint32_t m) { return (m >> 8) & 0xffu; }

struct DevTree {
  NodeRec* node;
  uint32_t cap;
};

struct StageTimer {
  int stage;
  hipEvent_t a, b;
};

struct gs_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n = 0, bucket = 16, flags = 0, walk_wg = 256;
  uint64_t max_nodes = 0;
  uint32_t sort_tiles = 0, build_tiles = 0;

  float4* posm[2] = {nullptr, nullptr};
  float4* velp[2] = {nullptr, nullptr};
  int cur = 0;
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* codes_raw = nullptr;
  uint32_t* sort_status = nullptr;  // GS_SORT_PASSES x tiles x radix
  uint64_t* build_status[2] = {nullptr, nullptr};
  float* bbox_part = nullptr;       // GS_BBOX_BLOCKS x 8
  DevTree tree{};
  float* energy = nullptr;
  float4* acc = nullptr;
  float* dens = nullptr;
  uint32_t* inter = nullptr;
  double* red_part = nullptr;       // walk blocks x GS_NRED
  uint32_t red_blocks = 0;
  uint32_t* node64 = nullptr;       // download staging
  DevScalars* sc = nullptr;
  DevParams* prm = nullptr;
  DevParams host_prm{};

  bool have_params = false, have_tree = false, have_codes = false,
       have_force = false, have_step = false;
  uint64_t step_count = 0;

  hipGraphExec_t graph[2] = {nullptr, nullptr};  // full step by parity
  bool graph_with_sort[2] = {false, false};

  std::vector<StageTimer> timers;
  double t_ms[GS_NT] = {0};
  uint64_t t_cnt[GS_NT] = {0};
  bool capturing = false;

  std::string err;
};

// ---- launchers (stream-ordered; return hipError_t of the launch) ----------
hipError_t gs_launch_reset(gs_sim* s);
hipError_t gs_launch_bbox(gs_sim* s, const float4* posm);
hipError_t gs_launch_root(gs_sim* s);
hipError_t gs_launch_encode(gs_sim* s, const float4* posm, bool keep_raw);
hipError_t gs_launch_hist(gs_sim* s);  // histograms of keys[0] (host-supplied codes)
hipError_t gs_launch_sort(gs_sim* s, bool gather);
hipError_t gs_launch_build(gs_sim* s);
hipError_t gs_launch_refresh(gs_sim* s, const float4* posm);
hipError_t gs_launch_walk(gs_sim* s, bool integrate, bool count, bool reduce);
hipError_t gs_launch_reduce_final(gs_sim* s, double* out_dev);
hipError_t gs_launch_node64(gs_sim* s);
hipError_t gs_launch_import64(gs_sim* s, const uint32_t* words, const int2* links, uint32_t num);
hipError_t gs_launch_gather(gs_sim* s, const uint32_t* perm);
hipError_t gs_launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t st);
hipError_t gs_launch_pack_soa(gs_sim* s, const float* aos, int which);
hipError_t gs_launch_unpack_soa(gs_sim* s, float* aos, int which);
