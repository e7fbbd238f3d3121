// radixSortIndices (mortonTree.js:68-84) on device arrays without any device
// allocation: gs_radix_sort_indices_noalloc.
//
// The sort's scratch -- two key buffers and one value buffer of the 8-bit LSD
// ping-pong, the onesweep look-back words of 4 passes, the tile tickets, the
// digit histograms, a one-slot descriptor and the tile table -- lives in the
// file-scope device array `scratch`, sized for GS_NOALLOC_MAX_KEYS keys.  Each
// device holds its own copy from the library's first use of this code object
// on it until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): the
// first call on a device resolves the device's address with
// hipGetSymbolAddress, with that device current, and keeps it per device.
//
// Per call, stream-ordered on the caller's stream, kernels only (no memset,
// no allocation, no host synchronisation):
//   1. clear_kernel: zero the look-back words of the tiles in use, the
//      tickets and the histograms; write the slot descriptor (n, capacity n)
//      and the tile table;
//   2. hist_kernel (encode.hip): keys[0] = codes, digit histograms;
//   3. onesweep_kernel x 4 (sort.hip): keys[0] -> keys[1] -> keys[0] ...,
//      values (the first pass pairs each key with its index) -> vals ->
//      perm -> vals -> perm, so the last pass leaves the gather permutation
//      in the caller's perm.
// Every call on a device shares that device's scratch: calls that overlap on
// one device (different streams, or threads) race, and must be serialised by
// the caller.
#include <mutex>

#include "gs_internal.h"

namespace {

constexpr uint32_t kMax = GS_NOALLOC_MAX_KEYS;
constexpr uint32_t kTiles = kMax / GS_SORT_TILE;
constexpr uint32_t kBits = 8, kRadix = 1u << kBits;
constexpr uint32_t kPasses = (uint32_t)gs_sort_passes((int)kBits);
static_assert(kMax % GS_SORT_TILE == 0, "whole tiles");
static_assert(kPasses == 4, "perm ends in the caller's buffer after an even number of passes");

struct Layout {
  uint32_t keys[2][kMax];
  uint32_t vals[kMax];
  uint32_t status[kPasses * kTiles * kRadix];
  uint32_t tile_tab[kTiles];
  DevBatch batch;
  DevSlot slot;
  DevScalars sc;
};

}  // namespace

static __device__ __align__(256) uint8_t scratch[sizeof(Layout)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
Layout* g_scratch[kMaxDevices] = {};

// the current device's address of `scratch` (resolved once per device)
int scratch_of_current_device(Layout** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return GS_E_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return GS_E_DEVICE;
    }
    g_scratch[dev] = static_cast<Layout*>(p);
  }
  *out = g_scratch[dev];
  return GS_OK;
}

__global__ __launch_bounds__(256) void clear_kernel(Layout* L, uint32_t n, uint32_t tiles) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  // the look-back words of the tiles in use: pass p's block is
  // status[p * tiles * radix ...], as the sort launcher addresses them
  for (uint32_t k = g; k < kPasses * tiles * kRadix; k += stride) L->status[k] = 0;
  for (uint32_t k = g; k < GS_HIST_WORDS; k += stride) L->sc.sort_hist[k] = 0;
  for (uint32_t k = g; k < tiles; k += stride) L->tile_tab[k] = k;  // slot 0, tile k
  if (g < 32) L->batch.tickets[g] = 0;
  if (g == 0) {
    L->batch.status = 0;
    DevSlot d = {};
    d.n = n;
    d.cap = n;  // no padding keys: every tile reads inside [0, n)
    d.nchunks = tiles;
    L->slot = d;
  }
}

}  // namespace

extern "C" int gs_radix_sort_indices_noalloc(const uint32_t* codes, uint32_t n, uint32_t* perm,
                                             void* stream) {
  if (n == 0) return GS_OK;
  if (!codes || !perm) return GS_E_INVALID;
  if (n > kMax) return GS_E_INVALID;  // larger than the static scratch holds
  Layout* L = nullptr;
  const int rc = scratch_of_current_device(&L);
  if (rc != GS_OK) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t tiles = (n + GS_SORT_TILE - 1) / GS_SORT_TILE;
  uint32_t grid = (kPasses * tiles * kRadix + 255) / 256;
  if (grid > 1024) grid = 1024;
  clear_kernel<<<grid, 256, 0, st>>>(L, n, tiles);
  if (hipGetLastError() != hipSuccess) return GS_E_DEVICE;
  SortBufs b;
  b.keys[0] = L->keys[0];
  b.keys[1] = L->keys[1];
  b.vals[0] = perm;
  b.vals[1] = L->vals;
  b.status = L->status;
  b.batch = &L->batch;
  b.slots = &L->slot;
  b.sc = &L->sc;
  b.tile_tab = L->tile_tab;
  b.tiles = tiles;
  b.radix_bits = kBits;
  b.stream = st;
  if (gs_launch_hist_bufs(b, codes, n) != hipSuccess) return GS_E_DEVICE;
  if (gs_launch_sort_bufs(b) != hipSuccess) return GS_E_DEVICE;
  return GS_OK;
}
