// C ABI of libgalaxy_hip.so (include/galaxy_hip.h): pool layout, uploads,
// stage sequencing on the handle's stream (direct launches or a replayed
// step graph), downloads, and the mortonTree.js functions on host arrays.
// No exception crosses the ABI: every entry point catches, records the
// message on the handle and returns a GS_E* status.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>

#include "gs_internal.h"
#include "tree_check.h"

#define GS_VERSION "galaxy_hip 0.2 (gfx950)"

namespace {

struct Fail {
  int code;
  std::string msg;
};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Fail{GS_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

template <class T>
void dalloc(T** p, size_t count, const char* what) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    throw Fail{GS_E_NOMEM, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e)};
  }
}

template <class F>
int guarded(gs_sim* s, F&& f) {
  if (!s) return GS_E_INVALID;
  try {
    if (s->device >= 0) check(hipSetDevice(s->device), "hipSetDevice");
    f();
    return GS_OK;
  } catch (const Fail& x) {
    s->err = x.msg;
    return x.code;
  } catch (const std::bad_alloc&) {
    s->err = "host allocation failed";
    return GS_E_NOMEM;
  } catch (...) {
    s->err = "unexpected C++ exception";
    return GS_E_INVALID;
  }
}

void need(bool ok, int code, const char* msg) {
  if (!ok) throw Fail{code, msg};
}

HostSlot& slot_of(gs_sim* s, uint32_t slot) {
  need(slot < s->nslots, GS_E_INVALID, "slot index out of range");
  return s->hs[slot];
}

void sync(gs_sim* s) { check(hipStreamSynchronize(s->stream), "hipStreamSynchronize"); }

// ---- stage timing (direct launches only) -----------------------------------
struct Stage {
  gs_sim* s;
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  // events come from the handle's spare list (created once, reused every
  // step), so timing adds no event creation to a step
  static hipEvent_t take(gs_sim* s) {
    if (!s->spare_events.empty()) {
      hipEvent_t e = s->spare_events.back();
      s->spare_events.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  Stage(gs_sim* sim, int stage) : s(sim), id(stage) {
    if (!(s->flags & GS_FLAG_TIMING) || s->capturing) return;
    a = take(s);
    b = take(s);
    check(hipEventRecord(a, s->stream), "hipEventRecord");
  }
  ~Stage() {
    if (!a) return;
    if (hipEventRecord(b, s->stream) == hipSuccess) {
      s->timers.push_back(StageTimer{id, a, b});
    } else {
      s->spare_events.push_back(a);
      s->spare_events.push_back(b);
    }
  }
};

void collect_timers(gs_sim* s) {
  if (s->timers.empty()) return;
  sync(s);
  for (auto& t : s->timers) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      s->t_ms[t.stage] += ms;
      s->t_cnt[t.stage] += 1;
    }
    s->spare_events.push_back(t.a);
    s->spare_events.push_back(t.b);
  }
  s->timers.clear();
}

void check_status(gs_sim* s) {
  uint32_t st = 0;
  check(hipMemcpyAsync(&st, &s->batch->status, sizeof st, hipMemcpyDeviceToHost, s->stream),
        "read status");
  sync(s);
  if (st & GS_ST_NODE_POOL)
    throw Fail{GS_E_NODE_POOL, "octree needs more nodes than the pool holds (raise max_nodes)"};
  if (st & GS_ST_LAYOUT)
    throw Fail{GS_E_INVALID,
               "uploaded tree: children not contiguous, or leaf ranges overlap"};
  if (st & GS_ST_STACK) throw Fail{GS_E_STACK, "walk stack overflow"};
  if (st & GS_ST_SPIN) throw Fail{GS_E_DEVICE, "a look-back wait timed out (scan state corrupt)"};
}

void destroy_graphs(gs_sim* s) {
  for (auto& a : s->graph)
    for (auto& b : a)
      for (auto& c : b)
        for (hipGraphExec_t& g : c)
          if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
          }
}

void free_all(gs_sim* s) {
  destroy_graphs(s);
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (int k = 0; k < 2; k++) {
    f(s->posm[k]);
    f(s->velp[k]);
    f(s->keys[k]);
    f(s->vals[k]);
    f(s->build_status[k]);
    f(s->stel[k]);
  }
  f(s->codes_raw);
  f(s->sort_status);
  f(s->bbox_part);
  f(s->nodes);
  f(s->nsize);
  f(s->massless);
  f(s->energy);
  f(s->acc);
  f(s->dens);
  f(s->inter);
  f(s->red_part);
  f(s->red_out);
  f(s->node64);
  f(s->slots);
  f(s->sc);
  f(s->prm);
  f(s->batch);
  f(s->chunk_tab);
  f(s->tile_tab);
  f(s->wblock_slot);
  f(s->big_leaf);
  for (auto& t : s->timers) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  s->timers.clear();
  for (hipEvent_t e : s->spare_events) (void)hipEventDestroy(e);
  s->spare_events.clear();
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

void pack_params(const gs_params* p, uint32_t n, DevParams* d) {
  // NBodyGPU.updateParams (NBodyGPU.js:253-275): Float32Array stores
  memset(d, 0, sizeof *d);
  d->dt = (float)p->dt;
  d->G = (float)p->G;
  d->eps2 = (float)(p->epsilon * p->epsilon);
  d->n = n;
  d->theta = (float)p->theta;
  d->halo_mass = (float)p->halo_mass;
  d->halo_a = (float)p->halo_a;
  d->gas_drag = (float)p->gas_drag;
  d->time = (float)p->time;
  d->aging_rate = (float)p->aging_rate;
  d->sf_density = (float)p->sf_density;
  d->sf_eff = (float)p->sf_eff;
  d->sf_dt = (float)p->sf_dt;
  d->halo_count = p->halo_count > 2 ? 2 : p->halo_count;
  d->halo_c1 = make_float4(d->halo_count > 0 ? (float)p->halo_centers[0][0] : 0.f,
                           d->halo_count > 0 ? (float)p->halo_centers[0][1] : 0.f,
                           d->halo_count > 0 ? (float)p->halo_centers[0][2] : 0.f, 0.f);
  d->halo_c2 = make_float4(d->halo_count > 1 ? (float)p->halo_centers[1][0] : 0.f,
                           d->halo_count > 1 ? (float)p->halo_centers[1][1] : 0.f,
                           d->halo_count > 1 ? (float)p->halo_centers[1][2] : 0.f, 0.f);
  d->theta2 = d->theta * d->theta;
  d->drag_damp = expf(-d->gas_drag * d->dt);
  d->qsize_bits = gs_quarter_eps_size(0.25f * d->eps2);
}

void push_slot(gs_sim* s, uint32_t slot) {
  check(hipMemcpyAsync(&s->slots[slot], &s->hs[slot].d, sizeof(DevSlot), hipMemcpyHostToDevice,
                       s->stream),
        "slot descriptor");
}

void push_params(gs_sim* s, uint32_t slot) {
  s->host_prm[slot].n = s->hs[slot].d.n;
  check(hipMemcpyAsync(&s->prm[slot], &s->host_prm[slot], sizeof(DevParams),
                       hipMemcpyHostToDevice, s->stream),
        "params");
}

// ---- stage sequences --------------------------------------------------------
void launch_sort(gs_sim* s, bool keep_raw) {
  {
    Stage t(s, GS_T_ENCODE);
    check(gs_launch_bbox(s, s->posm[s->cur]), "bbox");
    check(gs_launch_root(s), "root");
    check(gs_launch_encode(s, s->posm[s->cur], keep_raw), "encode");
  }
  {
    Stage t(s, GS_T_SORT);
    check(gs_launch_sort(s, true), "radix sort");
  }
  s->cur ^= 1;
  if (s->stel[0]) s->stel_cur ^= 1;  // the gather moved the stellar records too
  {
    Stage t(s, GS_T_BUILD);
    check(gs_launch_build(s, true), "build");
  }
}

void launch_refresh(gs_sim* s) {
  {
    Stage t(s, GS_T_REFRESH);
    check(gs_launch_refresh(s, s->posm[s->cur]), "refresh");
  }
}

void launch_walk_step(gs_sim* s) {
  {
    Stage t(s, GS_T_WALK);
    check(gs_launch_walk(s, true, false), "walk");
  }
  s->cur ^= 1;
}

void mark_sorted(gs_sim* s) {
  for (auto& h : s->hs) {
    h.have_tree = h.d.n > 0;
    h.have_com = false;  // the build writes the info words only
    h.have_force = false;
  }
}

void mark_refreshed(gs_sim* s) {
  for (auto& h : s->hs) h.have_com = h.d.n > 0 && h.have_tree;
}

void mark_stepped(gs_sim* s) {
  for (auto& h : s->hs) {
    h.have_step = h.have_red = h.d.n > 0;
    h.have_force = false;
  }
}

// the walk reads node COMs: every occupied slot needs a tree whose COM words
// were filled (refresh, tree upload or a full step)
void need_com(gs_sim* s, const char* msg) {
  for (auto& h : s->hs) need(!h.d.n || (h.have_tree && h.have_com), GS_E_STATE, msg);
}

bool use_graph(const gs_sim* s) { return !(s->flags & GS_FLAG_NO_GRAPH); }

// One step (optional rebuild + refresh + walk) as a graph for the current
// ping-pong parity; the launch sequence depends only on static sizes, so one
// capture serves every later step and every change of slot occupancy.
// part 0: the whole step; 1: rebuild (if sort) + refresh; 2: the walk
hipGraphExec_t step_graph(gs_sim* s, bool sort, int part) {
  hipGraphExec_t& slot = s->graph[s->cur][s->stel_cur][sort ? 1 : 0][part];
  if (slot) return slot;
  const int cur0 = s->cur, stel0 = s->stel_cur;
  hipGraph_t g = nullptr;
  check(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  s->capturing = true;  // no stage events inside a capture
  try {
    if (part != 2) {
      if (sort) launch_sort(s, false);
      launch_refresh(s);
    }
    if (part != 1) launch_walk_step(s);
  } catch (...) {
    s->capturing = false;
    hipGraph_t dead = nullptr;
    (void)hipStreamEndCapture(s->stream, &dead);
    if (dead) (void)hipGraphDestroy(dead);
    s->cur = cur0;
    s->stel_cur = stel0;
    throw;
  }
  s->capturing = false;
  check(hipStreamEndCapture(s->stream, &g), "end capture");
  s->cur = cur0;
  s->stel_cur = stel0;
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  check(e, "instantiate step graph");
  slot = ex;
  return ex;
}

// Marks the walks launched or captured in its scope as walking a tree built
// from the current positions; every other walk sees walk_fresh false, also
// when the scope is left by an error.
struct FreshWalk {
  gs_sim* s;
  FreshWalk(gs_sim* sim, bool fresh) : s(sim) { s->walk_fresh = fresh; }
  ~FreshWalk() { s->walk_fresh = false; }
};

void run_step(gs_sim* s, bool sort) {
  // the walk of a rebuilding step sees the tree built from its positions
  // (captured into the step graphs of that kind)
  FreshWalk fresh(s, sort);
  if (use_graph(s) && (s->flags & GS_FLAG_TIMING)) {
    // two replays with events between them: the walk and the whole step are
    // timed on the stream that runs them
    Stage t(s, GS_T_STEP);
    check(hipGraphLaunch(step_graph(s, sort, 1), s->stream), "graph launch");
    if (sort) {  // the sort's gather flipped the buffers
      s->cur ^= 1;
      if (s->stel[0]) s->stel_cur ^= 1;
    }
    {
      Stage w(s, GS_T_WALK);
      check(hipGraphLaunch(step_graph(s, sort, 2), s->stream), "graph launch");
    }
    s->cur ^= 1;
  } else if (use_graph(s)) {
    hipGraphExec_t ex = step_graph(s, sort, 0);
    check(hipGraphLaunch(ex, s->stream), "graph launch");
    s->cur ^= sort ? 0 : 1;  // the sort's gather flips once, the walk once
    if (sort && s->stel[0]) s->stel_cur ^= 1;
  } else {
    Stage t(s, GS_T_STEP);
    if (sort) launch_sort(s, false);
    launch_refresh(s);
    launch_walk_step(s);
  }
  if (sort) mark_sorted(s);
  mark_refreshed(s);
  mark_stepped(s);
}

uint32_t align_up(uint64_t v, uint32_t a) { return (uint32_t)((v + a - 1) / a * a); }

void push_tables(gs_sim* s, const std::vector<uint2>& chunk_tab,
                 const std::vector<uint32_t>& tile_tab, const std::vector<uint32_t>& wblock_slot) {
  if (!chunk_tab.empty())
    check(hipMemcpyAsync(s->chunk_tab, chunk_tab.data(), sizeof(uint2) * chunk_tab.size(),
                         hipMemcpyHostToDevice, s->stream),
          "tables");
  if (!tile_tab.empty())
    check(hipMemcpyAsync(s->tile_tab, tile_tab.data(), sizeof(uint32_t) * tile_tab.size(),
                         hipMemcpyHostToDevice, s->stream),
          "tables");
  if (!wblock_slot.empty())
    check(hipMemcpyAsync(s->wblock_slot, wblock_slot.data(), sizeof(uint32_t) * wblock_slot.size(),
                         hipMemcpyHostToDevice, s->stream),
          "tables");
}

// Device bytes of a pool: P particle records, C node records, the per-block
// arrays and tables at their allocated sizes, the per-slot records, stellar
// records when present and the node staging.
uint64_t pool_bytes(uint64_t P, uint64_t C, uint64_t chunks, uint64_t build_words,
                    uint64_t walk_blocks, uint64_t walk_waves, uint32_t nslots, bool stellar,
                    uint64_t node64) {
  return P * GS_ARENA_PARTICLE_BYTES + C * GS_ARENA_NODE_BYTES +
         2 * build_words * 8 +                   // build look-back words (ping-pong)
         chunks * (GS_HIST_WORDS * 4 + 8 * 4 +   // sort look-back words, bbox partials
                   sizeof(uint2) + 4) +          // chunk and tile tables
         walk_blocks * 4 +                       // walk block table
         walk_waves * GS_NRED * 8 +              // reduction partials (one set per wavefront)
         (P / GS_BIG_LEAF + 64) * sizeof(uint2) +  // large-leaf list
         (uint64_t)nslots * (sizeof(DevSlot) + sizeof(DevScalars) + sizeof(DevParams) + 4) +
         sizeof(DevBatch) + GS_NRED * 8 + (stellar ? P * 32 : 0) + node64 * 64;
}

// node download / upload staging for pools of up to ncap nodes
void ensure_node64(gs_sim* s, uint64_t ncap) {
  if (s->node64 && s->node64_cap >= ncap) return;
  if (s->node64) {
    sync(s);
    (void)hipFree(s->node64);
    s->node64 = nullptr;
    s->node64_cap = 0;
  }
  dalloc(&s->node64, (size_t)ncap * 16, "node staging");
  s->node64_cap = ncap;
}

// Rebuilds an arena's launch tables over its occupied slots (slot order):
// bbox/encode chunks = sort tiles, build look-back words and walk blocks of
// each slot follow those of the slots before it.  The build's look-back words
// start cleared, the step graphs (captured with the old launch sizes) are
// dropped, and the reduction partials of the last step no longer sit where
// the slots' walk blocks now point.
void apply_layout(gs_sim* s) {
  std::vector<uint2> chunk_tab;
  std::vector<uint32_t> tile_tab, wblock_slot;
  uint32_t chunk = 0, bt = 0, wb = 0;
  const uint32_t wg = s->walk_wg;
  for (uint32_t k = 0; k < s->nslots; k++) {
    DevSlot& d = s->hs[k].d;
    d.nchunks = (d.cap + GS_CHUNK - 1) / GS_CHUNK;
    d.chunk0 = chunk;
    d.tile0 = chunk;
    for (uint32_t c = 0; c < d.nchunks; c++) {
      chunk_tab.push_back(make_uint2(k, c * GS_CHUNK));
      tile_tab.push_back((k << 16) | c);
    }
    chunk += d.nchunks;
    d.btoff = bt;
    bt += (d.ncap + GS_BUILD_TILE - 1) / GS_BUILD_TILE;
    d.wblock0 = wb;
    d.nwblocks = align_up(d.cap, GS_SLOT_ALIGN) / wg;
    for (uint32_t b = 0; b < d.nwblocks; b++) wblock_slot.push_back(k);
    wb += d.nwblocks;
    s->hs[k].have_red = false;
  }
  need(chunk <= s->cap_chunks && bt <= s->cap_build_words && wb <= s->cap_walk_blocks,
       GS_E_INVALID, "arena launch tables overflow");
  s->chunks = chunk;
  s->sort_tiles = chunk;
  s->build_words = bt;
  s->walk_blocks = wb;
  push_tables(s, chunk_tab, tile_tab, wblock_slot);
  std::vector<DevSlot> ds(s->nslots);
  for (uint32_t k = 0; k < s->nslots; k++) ds[k] = s->hs[k].d;
  check(hipMemcpyAsync(s->slots, ds.data(), sizeof(DevSlot) * s->nslots, hipMemcpyHostToDevice,
                       s->stream),
        "slot descriptors");
  // the build's look-back words of the new layout start cleared (each level
  // clears the other buffer for the next); the sort's are cleared by every
  // step's first kernel
  for (int k = 0; k < 2; k++)
    if (bt)
      check(gs_launch_fill_u32((uint32_t*)s->build_status[k], 0, 2ull * bt, s->stream), "fill");
  sync(s);
  destroy_graphs(s);
  s->layout_dirty = false;
}

// Before every launch over all slots: an arena's tables follow its slots.
// False when no slot holds a galaxy (an empty arena has nothing to launch).
bool ready(gs_sim* s) {
  if (s->arena && s->layout_dirty) apply_layout(s);
  return s->walk_blocks > 0;
}

// An arena slot's ranges back to the free lists; the slot holds nothing.
void arena_release(gs_sim* s, HostSlot& h) {
  if (!h.d.cap) return;
  const bool ok = s->pfree.give(h.d.poff / GS_SLOT_ALIGN,
                                align_up(h.d.cap, GS_SLOT_ALIGN) / GS_SLOT_ALIGN) &&
                  s->nfree.give(h.d.noff, h.d.ncap);
  need(ok, GS_E_INVALID, "arena free lists corrupt");
  // the slot's tree scalars go to zero (no nodes, no levels): a refresh or
  // walk before its next rebuild finds nothing to do in the ranges given away
  const size_t k = (size_t)(&h - s->hs.data());
  check(gs_launch_fill_u32((uint32_t*)(s->sc + k), 0, sizeof(DevScalars) / 4, s->stream),
        "clear slot scalars");
  h.d.cap = h.d.ncap = h.d.poff = h.d.noff = 0;
  h.d.max_nodes = 0;
  h.d.nchunks = h.d.nwblocks = 0;  // no blocks until the tables are rebuilt
  s->layout_dirty = true;
}

// Ranges for n particles in an arena slot: align_up(n, GS_SLOT_ALIGN)
// particles and 2n + 4096 nodes (mortonTree.js:88), best fit; a slot that
// already holds n keeps its ranges.  GS_E_FULL (the slot left empty) when
// no free range holds them.
void arena_place(gs_sim* s, HostSlot& h, uint64_t n) {
  if (n && n == h.d.cap) return;
  arena_release(s, h);
  if (!n) return;
  need(n <= GS_MAX_N, GS_E_INVALID, "more particles than a slot may hold");
  const uint64_t units = align_up(n, GS_SLOT_ALIGN) / GS_SLOT_ALIGN;
  const uint64_t nodes = std::min<uint64_t>(2 * n + 4096, GS_NODE32_MAX);
  uint64_t pu = 0, nu = 0;
  need(s->pfree.take(units, &pu), GS_E_FULL, "arena full: no free particle range holds the galaxy");
  if (!s->nfree.take(nodes, &nu)) {
    s->pfree.give(pu, units);
    throw Fail{GS_E_FULL, "arena full: no free node range holds the galaxy's tree"};
  }
  h.d.cap = (uint32_t)n;
  h.d.poff = (uint32_t)(pu * GS_SLOT_ALIGN);
  h.d.noff = (uint32_t)nu;
  h.d.ncap = (uint32_t)nodes;
  h.d.max_nodes = nodes;
  h.fixed_max_nodes = false;  // a new galaxy starts with the leaf rule of new MortonTree(n)
  h.d.nchunks = h.d.nwblocks = 0;
  s->max_ncap = std::max(s->max_ncap, h.d.ncap);
  s->layout_dirty = true;
}

// Allocates a pool of nslots slots (caps[k] particles, buckets[k] leaf size,
// a node pool of max(2 caps[k] + 4096, max_nodes) nodes), or, with
// arena_p > 0, an arena of arena_p particles (a multiple of GS_SLOT_ALIGN)
// and arena_c nodes whose nslots slots start empty.
gs_sim* make_pool(const uint32_t* caps, const uint32_t* buckets, uint32_t nslots, uint32_t flags,
                  int device, uint32_t wg, uint32_t radix_bits, uint64_t max_nodes,
                  uint64_t arena_p = 0, uint64_t arena_c = 0) {
  gs_sim* s = new gs_sim();
  s->flags = flags;
  s->walk_wg = wg;
  s->radix_bits = radix_bits;
  s->nslots = nslots;
  s->hs.resize(nslots);
  s->host_prm.resize(nslots);
  uint64_t poff = 0, noff = 0;
  uint32_t chunk = 0, tile = 0, bt = 0, wb = 0;
  std::vector<uint2> chunk_tab;
  std::vector<uint32_t> tile_tab, wblock_slot;
  for (uint32_t k = 0; k < nslots; k++) {
    DevSlot& d = s->hs[k].d;
    memset(&d, 0, sizeof d);
    d.bucket = buckets && buckets[k] ? buckets[k] : 16;
    if (arena_p) continue;  // arena slots get their ranges when filled
    const uint32_t cap = caps[k];
    d.n = 0;
    d.cap = cap;
    d.poff = (uint32_t)poff;
    // mortonTree.js:88, capped at GS_NODE32_MAX records
    d.ncap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(2ull * cap + 4096, max_nodes),
                                          GS_NODE32_MAX);
    d.noff = (uint32_t)noff;
    d.max_nodes = d.ncap;
    d.nchunks = (cap + GS_CHUNK - 1) / GS_CHUNK;
    d.chunk0 = chunk;
    d.tile0 = tile;
    for (uint32_t c = 0; c < d.nchunks; c++) {
      chunk_tab.push_back(make_uint2(k, c * GS_CHUNK));
      tile_tab.push_back((k << 16) | c);
    }
    chunk += d.nchunks;
    tile += d.nchunks;
    d.btoff = bt;
    bt += (d.ncap + GS_BUILD_TILE - 1) / GS_BUILD_TILE;
    const uint32_t span = align_up(cap, GS_SLOT_ALIGN);
    d.wblock0 = wb;
    d.nwblocks = span / wg;
    for (uint32_t b = 0; b < d.nwblocks; b++) wblock_slot.push_back(k);
    wb += d.nwblocks;
    poff += span;
    noff += d.ncap;
    s->max_ncap = std::max(s->max_ncap, d.ncap);
  }
  s->ptot = poff;
  s->ntot = noff;
  s->chunks = chunk;
  s->sort_tiles = tile;
  s->build_words = bt;
  s->walk_blocks = wb;
  s->cap_chunks = chunk;
  s->cap_build_words = bt;
  s->cap_walk_blocks = wb;
  if (arena_p) {
    // an arena's slots take ranges of the arena: each slot's blocks are at
    // most its aligned span / GS_CHUNK + 1 chunks, ncap / GS_BUILD_TILE + 1
    // look-back words and span / wg walk blocks
    s->arena = true;
    s->ptot = arena_p;
    s->ntot = arena_c;
    s->cap_chunks = (uint32_t)(arena_p / GS_CHUNK + nslots);
    s->cap_build_words = (uint32_t)(arena_c / GS_BUILD_TILE + nslots);
    s->cap_walk_blocks = (uint32_t)(arena_p / wg);
    s->pfree.reset(arena_p / GS_SLOT_ALIGN);
    s->nfree.reset(arena_c);
  }
  int rc = guarded(s, [&] {
    need(s->ptot < (1ull << 32) && s->ntot < (1ull << 32), GS_E_INVALID,
         "pool larger than 2^32 particles or nodes");
    int ndev = 0;
    check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    need(ndev > 0, GS_E_DEVICE, "no HIP device");
    int dev = device;
    if (dev < 0) check(hipGetDevice(&dev), "hipGetDevice");
    need(dev < ndev, GS_E_INVALID, "device ordinal out of range");
    s->device = dev;
    check(hipSetDevice(dev), "hipSetDevice");
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    need(strncmp(prop.gcnArchName, "gfx950", 6) == 0, GS_E_DEVICE,
         "libgalaxy_hip is built for gfx950 (MI355X) only");
    check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    const size_t P = s->ptot;
    for (int k = 0; k < 2; k++) {
      dalloc(&s->posm[k], P, "posm");
      dalloc(&s->velp[k], P, "velp");
      dalloc(&s->keys[k], P, "keys");
      dalloc(&s->vals[k], P, "vals");
      dalloc(&s->build_status[k], s->cap_build_words, "build status");
    }
    dalloc(&s->codes_raw, P, "codes");
    dalloc(&s->sort_status, (size_t)GS_HIST_WORDS * s->cap_chunks, "sort status");
    dalloc(&s->bbox_part, (size_t)s->cap_chunks * 8, "bbox");
    dalloc(&s->nodes, s->ntot, "nodes");
    dalloc(&s->nsize, s->ntot, "node sizes");
    dalloc(&s->massless, nslots, "massless flags");
    dalloc(&s->energy, P, "energy");
    dalloc(&s->acc, P, "acc");
    dalloc(&s->dens, P, "density");
    dalloc(&s->inter, P, "interactions");
    // one set of FP64 partials per walk wavefront
    dalloc(&s->red_part, (size_t)s->cap_walk_blocks * (s->walk_wg / 64) * GS_NRED, "reductions");
    dalloc(&s->red_out, GS_NRED, "reduction");
    dalloc(&s->slots, nslots, "slots");
    dalloc(&s->sc, nslots, "scalars");
    dalloc(&s->prm, nslots, "params");
    dalloc(&s->batch, 1, "batch");
    dalloc(&s->chunk_tab, s->cap_chunks, "chunk table");
    dalloc(&s->tile_tab, s->cap_chunks, "tile table");
    dalloc(&s->wblock_slot, s->cap_walk_blocks, "walk block table");
    // a leaf of more than GS_BIG_LEAF members holds that many particles
    s->big_cap = (uint32_t)(P / GS_BIG_LEAF + 64);
    dalloc(&s->big_leaf, s->big_cap, "large-leaf list");
    // zero state with fill kernels (graph-safe, no memset nodes)
    check(gs_launch_fill_u32((uint32_t*)s->sc, 0, sizeof(DevScalars) / 4 * nslots, s->stream),
          "fill");
    check(gs_launch_fill_u32((uint32_t*)s->batch, 0, sizeof(DevBatch) / 4, s->stream), "fill");
    check(gs_launch_fill_u32(s->massless, 1u, nslots, s->stream), "fill");
    check(gs_launch_fill_u32((uint32_t*)s->energy, 0, P, s->stream), "fill");
    for (int k = 0; k < 2; k++) {
      check(gs_launch_fill_u32((uint32_t*)s->build_status[k], 0, 2ull * s->cap_build_words,
                               s->stream),
            "fill");
      check(gs_launch_fill_u32((uint32_t*)s->posm[k], 0, 4 * P, s->stream), "fill");
      check(gs_launch_fill_u32((uint32_t*)s->velp[k], 0, 4 * P, s->stream), "fill");
    }
    push_tables(s, chunk_tab, tile_tab, wblock_slot);
    gs_params p;
    gs_params_default(&p);
    for (uint32_t k = 0; k < nslots; k++) {
      pack_params(&p, 0, &s->host_prm[k]);
      push_slot(s, k);
      push_params(s, k);
    }
    sync(s);
  });
  if (rc != GS_OK) {
    free_all(s);
    delete s;
    return nullptr;
  }
  return s;
}

void download_slot(gs_sim* s, uint32_t slot, void* dst, const void* dev_base, size_t elem,
                   uint64_t n, const char* what) {
  const HostSlot& h = s->hs[slot];
  check(hipMemcpyAsync(dst, (const char*)dev_base + (size_t)h.d.poff * elem, elem * n,
                       hipMemcpyDeviceToHost, s->stream),
        what);
  sync(s);
}

// f32(size * size), the value a node record holds
float host_s2(float size) {
  const int b = gs_s2_bits(size);
  float f;
  memcpy(&f, &b, 4);
  return f;
}

DevScalars read_scalars(gs_sim* s, uint32_t slot) {
  DevScalars h;
  check(hipMemcpyAsync(&h, s->sc + slot, sizeof h, hipMemcpyDeviceToHost, s->stream), "scalars");
  sync(s);
  return h;
}

void upload_particles(gs_sim* s, uint32_t slot, const float* particles, uint64_t n) {
  HostSlot& h = s->hs[slot];
  need(particles != nullptr || n == 0, GS_E_INVALID, "null particles");
  if (s->arena) arena_place(s, h, n);
  need(n <= h.d.cap, GS_E_INVALID, "more particles than the slot holds");
  check(gs_upload_records(s, slot, particles, (uint32_t)n), "upload particles");
  if (s->stel[0])  // a new galaxy starts without stellar records until it uploads its own
    for (int k = 0; k < 2; k++)
      check(gs_launch_fill_u32((uint32_t*)(s->stel[k] + h.d.poff), 0, 4ull * h.d.cap, s->stream),
            "clear stellar");
  h.d.n = (uint32_t)n;
  if (!h.fixed_max_nodes)  // new MortonTree(N): maxNodes = 2N + 4096 (mortonTree.js:88)
    h.d.max_nodes = std::min<uint64_t>(2 * n + 4096, h.d.ncap);
  push_slot(s, slot);
  push_params(s, slot);
  sync(s);
  h.have_tree = h.have_force = h.have_step = h.have_codes = h.have_com = h.have_red = false;
}

// Stellar records are allocated on first upload; step graphs captured
// without them are dropped so the next capture gathers them too.
void ensure_stellar(gs_sim* s) {
  if (s->stel[0]) return;
  for (int k = 0; k < 2; k++) {
    dalloc(&s->stel[k], s->ptot, "stellar records");
    check(gs_launch_fill_u32((uint32_t*)s->stel[k], 0, 4 * s->ptot, s->stream), "fill");
  }
  s->stel_cur = 0;
  destroy_graphs(s);
}

}  // namespace

extern "C" {

const char* gs_version(void) { return GS_VERSION; }

int gs_device_count(int* out) {
  if (!out) return GS_E_INVALID;
  *out = 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return GS_OK;
  }
  int good = 0;
  for (int d = 0; d < n; d++) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, d) == hipSuccess && strncmp(p.gcnArchName, "gfx950", 6) == 0)
      good++;
  }
  *out = good;
  return GS_OK;
}

void gs_params_default(gs_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->dt = 0.01;  // Controls.js:34 defaults for the required fields
  p->G = 1.0;
  p->epsilon = 1.5;
  p->theta = 0.5;  // NBodyGPU.js:253-255 defaults
  p->halo_mass = 0.0;
  p->halo_a = 1.0;
  p->gas_drag = 0.0;
  p->time = 0.0;
  p->aging_rate = 25.0;
  p->sf_density = 1e9;
  p->sf_eff = 0.0;
  p->sf_dt = 0.0;
  p->halo_count = 0;
}

static bool valid_wg(uint32_t wg) { return wg >= 64 && wg <= 1024 && (wg & (wg - 1)) == 0; }

int gs_create(const gs_config* cfg, gs_sim** out) {
  if (!cfg || !out) return GS_E_INVALID;
  *out = nullptr;
  if (cfg->n == 0 || cfg->n > GS_MAX_N) return GS_E_INVALID;
  const uint32_t wg = cfg->walk_wg ? cfg->walk_wg : 64;
  if (!valid_wg(wg)) return GS_E_INVALID;
  if (cfg->radix_bits != 0 && cfg->radix_bits != 8 && cfg->radix_bits != 10) return GS_E_INVALID;
  const uint32_t bucket = cfg->bucket ? cfg->bucket : 16;
  gs_sim* s = nullptr;
  try {
    if (cfg->max_nodes > GS_NODE32_MAX) return GS_E_INVALID;
    s = make_pool(&cfg->n, &bucket, 1, cfg->flags, cfg->device, wg,
                  cfg->radix_bits ? cfg->radix_bits : 8, cfg->max_nodes);
  } catch (...) {
    s = nullptr;
  }
  if (!s) return GS_E_DEVICE;
  if (cfg->max_nodes) {
    // the leaf rule's pool; the device pool holds max(2N + 4096, max_nodes)
    if (gs_slot_set_tree(s, 0, bucket, cfg->max_nodes) != GS_OK) {
      gs_destroy(s);
      return GS_E_INVALID;
    }
  }
  *out = s;
  return GS_OK;
}

int gs_create_pool(const gs_pool_config* cfg, gs_sim** out) {
  if (!cfg || !out || !cfg->capacity) return GS_E_INVALID;
  *out = nullptr;
  if (cfg->nslots == 0 || cfg->nslots > GS_MAX_SLOTS) return GS_E_INVALID;
  for (uint32_t k = 0; k < cfg->nslots; k++)
    if (cfg->capacity[k] == 0 || cfg->capacity[k] > GS_MAX_N) return GS_E_INVALID;
  const uint32_t wg = cfg->walk_wg ? cfg->walk_wg : 64;
  if (!valid_wg(wg)) return GS_E_INVALID;
  if (cfg->radix_bits != 0 && cfg->radix_bits != 8 && cfg->radix_bits != 10) return GS_E_INVALID;
  gs_sim* s = nullptr;
  try {
    s = make_pool(cfg->capacity, cfg->bucket, cfg->nslots, cfg->flags, cfg->device, wg,
                  cfg->radix_bits ? cfg->radix_bits : 8, 0);
  } catch (...) {
    s = nullptr;
  }
  if (!s) return GS_E_DEVICE;
  *out = s;
  return GS_OK;
}

void gs_destroy(gs_sim* s) {
  if (!s) return;
  if (s->device >= 0) (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
}

const char* gs_last_error(const gs_sim* s) { return s ? s->err.c_str() : "null handle"; }

int gs_pool_info(gs_sim* s, uint32_t* nslots, uint64_t* device_bytes) {
  return guarded(s, [&] {
    if (nslots) *nslots = s->nslots;
    if (device_bytes)
      *device_bytes = pool_bytes(s->ptot, s->ntot, s->cap_chunks, s->cap_build_words,
                                 s->cap_walk_blocks, (uint64_t)s->cap_walk_blocks * (s->walk_wg / 64),
                                 s->nslots, s->stel[0] != nullptr,
                                 s->node64 ? s->node64_cap : 0);
  });
}

int gs_create_arena(const gs_arena_config* cfg, gs_sim** out) {
  if (!cfg || !out) return GS_E_INVALID;
  *out = nullptr;
  if (cfg->nslots == 0 || cfg->nslots > GS_MAX_SLOTS) return GS_E_INVALID;
  const uint32_t wg = cfg->walk_wg ? cfg->walk_wg : 64;
  if (!valid_wg(wg)) return GS_E_INVALID;
  if (cfg->radix_bits != 0 && cfg->radix_bits != 8 && cfg->radix_bits != 10) return GS_E_INVALID;
  // the most particles P (a multiple of GS_SLOT_ALIGN) whose arena, with
  // C = 2P + 4096 nslots nodes (every slot's 2n + 4096 fits), stays within
  // the budget; P and C stay below 2^32
  const uint64_t ns = cfg->nslots;
  auto bytes = [&](uint64_t P) {
    const uint64_t C = 2 * P + 4096 * ns;
    return pool_bytes(P, C, P / GS_CHUNK + ns, C / GS_BUILD_TILE + ns, P / wg, P / 64, cfg->nslots, false,
                      0);
  };
  const uint64_t per = GS_ARENA_PARTICLE_BYTES + 2 * GS_ARENA_NODE_BYTES;
  const uint64_t pmax = ((0xffffffffull - 4096 * ns) / 2) / GS_SLOT_ALIGN * GS_SLOT_ALIGN;
  uint64_t P = std::min<uint64_t>(cfg->budget_bytes / per / GS_SLOT_ALIGN * GS_SLOT_ALIGN, pmax);
  while (P && bytes(P) > cfg->budget_bytes) {
    const uint64_t step = ((bytes(P) - cfg->budget_bytes) / per / GS_SLOT_ALIGN + 1) * GS_SLOT_ALIGN;
    P = P > step ? P - step : 0;
  }
  if (!P) return GS_E_INVALID;  // the budget holds no aligned range
  gs_sim* s = nullptr;
  try {
    s = make_pool(nullptr, nullptr, cfg->nslots, cfg->flags, cfg->device, wg,
                  cfg->radix_bits ? cfg->radix_bits : 8, 0, P, 2 * P + 4096 * ns);
  } catch (...) {
    s = nullptr;
  }
  if (!s) return GS_E_DEVICE;
  *out = s;
  return GS_OK;
}

int gs_arena_info(gs_sim* s, uint64_t* device_bytes, uint64_t* used_bytes, uint64_t* largest_fit) {
  return guarded(s, [&] {
    need(s->arena, GS_E_INVALID, "not an arena pool");
    if (device_bytes)
      *device_bytes = pool_bytes(s->ptot, s->ntot, s->cap_chunks, s->cap_build_words,
                                 s->cap_walk_blocks, (uint64_t)s->cap_walk_blocks * (s->walk_wg / 64),
                                 s->nslots, false,
                                 s->node64 ? s->node64_cap : 0);
    if (used_bytes) {
      uint64_t u = 0;
      for (const auto& h : s->hs)
        if (h.d.cap)
          u += (uint64_t)align_up(h.d.cap, GS_SLOT_ALIGN) * GS_ARENA_PARTICLE_BYTES +
               (uint64_t)h.d.ncap * GS_ARENA_NODE_BYTES;
      *used_bytes = u;
    }
    if (largest_fit) {
      const uint64_t by_p = s->pfree.largest() * GS_SLOT_ALIGN;
      const uint64_t nl = s->nfree.largest();
      const uint64_t by_n = nl >= 4096 ? (nl - 4096) / 2 : 0;
      *largest_fit = std::min<uint64_t>(std::min(by_p, by_n), GS_MAX_N);
    }
  });
}

int gs_upload(gs_sim* s, const float* particles, uint64_t n) {
  return guarded(s, [&] {
    need(n == s->hs[0].d.cap, GS_E_INVALID, "upload needs exactly N records");
    upload_particles(s, 0, particles, n);
  });
}

int gs_slot_upload(gs_sim* s, uint32_t slot, const float* particles, uint64_t n) {
  return guarded(s, [&] {
    slot_of(s, slot);
    upload_particles(s, slot, particles, n);
  });
}

int gs_slot_clear(gs_sim* s, uint32_t slot) {
  return guarded(s, [&] {
    slot_of(s, slot);
    upload_particles(s, slot, nullptr, 0);
  });
}

int gs_slot_count(gs_sim* s, uint32_t slot, uint32_t* n) {
  return guarded(s, [&] {
    need(n != nullptr, GS_E_INVALID, "null output");
    *n = slot_of(s, slot).d.n;
  });
}

int gs_set_params(gs_sim* s, const gs_params* p) { return gs_slot_set_params(s, 0, p); }

int gs_slot_set_params(gs_sim* s, uint32_t slot, const gs_params* p) {
  return guarded(s, [&] {
    HostSlot& h = slot_of(s, slot);
    need(p != nullptr, GS_E_INVALID, "null params");
    pack_params(p, h.d.n, &s->host_prm[slot]);
    push_params(s, slot);
    // the tree's records hold opening thresholds for the old theta
    if (h.have_tree) check(gs_launch_rethreshold(s, slot), "rethreshold");
    sync(s);
    h.have_params = true;
  });
}

int gs_slot_set_tree(gs_sim* s, uint32_t slot, uint32_t bucket, uint64_t max_nodes) {
  return guarded(s, [&] {
    HostSlot& h = slot_of(s, slot);
    need(bucket >= 1, GS_E_INVALID, "bucket must be >= 1");
    const uint64_t mn = max_nodes ? max_nodes : h.d.ncap;
    need(mn <= h.d.ncap, GS_E_INVALID, "max_nodes larger than the slot's node pool");
    h.d.bucket = bucket;
    h.d.max_nodes = mn;
    h.fixed_max_nodes = max_nodes != 0;
    push_slot(s, slot);
    sync(s);
    h.have_tree = false;
  });
}

int gs_update_central_mass(gs_sim* s, float mass, uint32_t idx) {
  return guarded(s, [&] {
    const HostSlot& h = s->hs[0];
    need(idx < h.d.n, GS_E_INVALID, "index out of range");
    // mass is the w of posm in both ping-pong buffers (NBodyGPU.js:248-250)
    for (int k = 0; k < 2; k++)
      check(hipMemcpyAsync((char*)(s->posm[k] + h.d.poff + idx) + 12, &mass, 4,
                           hipMemcpyHostToDevice, s->stream),
            "central mass");
    sync(s);
  });
}

int gs_apply_permutation(gs_sim* s, const uint32_t* perm, uint64_t n) {
  return guarded(s, [&] {
    need(!s->arena, GS_E_INVALID, "applyPermutation needs a plain handle, not an arena");
    const HostSlot& h = s->hs[0];
    need(perm != nullptr && n == h.d.n, GS_E_INVALID, "permutation needs N entries");
    const char* why = gs::check_permutation(perm, n);
    need(why == nullptr, GS_E_INVALID, why ? why : "");
    uint32_t* dperm = nullptr;
    dalloc(&dperm, n, "perm");
    hipError_t e = hipMemcpyAsync(dperm, perm, 4 * n, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = gs_launch_gather(s, 0, dperm);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dperm);
    check(e, "gather");
    s->cur ^= 1;
    if (s->stel[0]) s->stel_cur ^= 1;
    for (auto& hh : s->hs) hh.have_force = false;
  });
}

int gs_upload_tree(gs_sim* s, const uint32_t* nodes, uint64_t num_nodes,
                   const uint32_t* level_ranges, uint32_t num_levels) {
  return guarded(s, [&] {
    need(!s->arena, GS_E_INVALID, "tree uploads need a plain handle, not an arena");
    HostSlot& h = s->hs[0];
    need(nodes != nullptr && num_nodes >= 1 && num_nodes <= h.d.ncap, GS_E_INVALID,
         "node count outside [1, pool]");
    need(num_levels <= GS_MAX_LEVELS && (num_levels == 0 || level_ranges), GS_E_INVALID,
         "bad level ranges");
    // host validation: every index the device will follow stays in bounds;
    // the walk's (skip, depth) links follow from the BFS numbering
    std::vector<gs::TreeLink> links;
    float min_s2 = INFINITY;  // the walk's eps^2/4 opening test needs it
    uint32_t massless = 0;    // and its massless-node test this
    const char* why = gs::check_tree64(nodes, num_nodes, h.d.n, level_ranges, num_levels,
                                       GS_MAX_LEVEL, links, min_s2, massless);
    need(why == nullptr, GS_E_INVALID, why ? why : "");
    ensure_node64(s, s->max_ncap);
    check(hipMemcpyAsync(s->node64, nodes, num_nodes * 64, hipMemcpyHostToDevice, s->stream),
          "nodes");
    DevScalars sc = read_scalars(s, 0);
    sc.num_nodes = (uint32_t)num_nodes;
    sc.num_levels = num_levels;
    sc.min_s2 = min_s2;
    for (uint32_t k = 0; k <= GS_MAX_LEVELS; k++) {
      sc.level_start[k] = k < num_levels ? level_ranges[2 * k] : 0;
      sc.level_count[k] = k < num_levels ? level_ranges[2 * k + 1] : 0;
    }
    check(hipMemcpyAsync(s->sc, &sc, sizeof sc, hipMemcpyHostToDevice, s->stream), "scalars");
    check(hipMemcpyAsync(s->massless, &massless, 4, hipMemcpyHostToDevice, s->stream), "flag");
    static_assert(sizeof(gs::TreeLink) == sizeof(int2), "links are copied as int2 (skip, depth)");
    int2* dlinks = nullptr;
    dalloc(&dlinks, num_nodes, "links");
    hipError_t e = hipMemcpyAsync(dlinks, links.data(), sizeof(int2) * num_nodes,
                                  hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = gs_launch_import64(s, 0, s->node64, dlinks, (uint32_t)num_nodes);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dlinks);
    check(e, "import");
    check_status(s);  // non-contiguous children are rejected here
    h.have_tree = true;
    h.have_com = true;  // the uploaded words carry the caller's COMs
    h.have_force = false;
  });
}

int gs_encode(gs_sim* s) {
  return guarded(s, [&] {
    if (!ready(s)) return;
    Stage t(s, GS_T_ENCODE);
    check(gs_launch_bbox(s, s->posm[s->cur]), "bbox");
    check(gs_launch_root(s), "root");
    check(gs_launch_encode(s, s->posm[s->cur], true), "encode");
    for (auto& h : s->hs) {
      h.have_codes = h.d.n > 0;
      h.have_tree = h.have_com = false;  // the root node slot was rewritten
    }
  });
}

int gs_sort(gs_sim* s) {
  return guarded(s, [&] {
    if (!ready(s)) return;
    launch_sort(s, true);
    mark_sorted(s);
    for (auto& h : s->hs) h.have_codes = h.d.n > 0;
  });
}

int gs_refresh(gs_sim* s) {
  return guarded(s, [&] {
    for (auto& h : s->hs) need(!h.d.n || h.have_tree, GS_E_STATE, "refresh before any sort");
    if (!ready(s)) return;
    launch_refresh(s);
    mark_refreshed(s);
  });
}

int gs_force(gs_sim* s) {
  return guarded(s, [&] {
    need_com(s, "force needs a tree with COMs (sort, then refresh or upload a tree)");
    if (!ready(s)) return;
    Stage t(s, GS_T_WALK);
    check(gs_launch_walk(s, false, false), "walk");
    for (auto& h : s->hs) h.have_force = h.d.n > 0;
  });
}

int gs_step(gs_sim* s, uint32_t nsteps, uint32_t sort_every) {
  return guarded(s, [&] {
    if (!ready(s)) return;
    for (uint32_t k = 0; k < nsteps; k++) {
      bool missing = false;  // a slot without a tree forces a rebuild
      for (auto& h : s->hs) missing |= h.d.n && !h.have_tree;
      need(!missing || sort_every > 0, GS_E_STATE, "step without a tree and sort_every = 0");
      const bool sort = missing || (sort_every > 0 && s->step_count % sort_every == 0);
      run_step(s, sort);
      s->step_count++;
    }
  });
}

int gs_walk_step(gs_sim* s) {
  return guarded(s, [&] {
    need_com(s, "step needs a tree with COMs (refresh, upload a tree or gs_step first)");
    if (!ready(s)) return;
    if (use_graph(s)) {
      Stage w(s, GS_T_WALK);
      check(hipGraphLaunch(step_graph(s, false, 2), s->stream), "graph launch");
      s->cur ^= 1;
    } else {
      launch_walk_step(s);
    }
    mark_stepped(s);
  });
}

int gs_sync(gs_sim* s) {
  return guarded(s, [&] {
    check_status(s);
    collect_timers(s);  // the stream is idle: fold the finished stage events in
  });
}

int gs_download_particles(gs_sim* s, float* dst, uint64_t n) {
  return gs_slot_download_particles(s, 0, dst, n);
}

int gs_slot_download_particles(gs_sim* s, uint32_t slot, float* dst, uint64_t n) {
  return guarded(s, [&] {
    const HostSlot& h = slot_of(s, slot);
    need(dst != nullptr && n == h.d.n, GS_E_INVALID, "download needs exactly n records");
    if (!n) return;
    check(gs_download_records(s, slot, dst, (uint32_t)n), "download particles");
    sync(s);
  });
}

int gs_read_positions(gs_sim* s, float* dst4, uint64_t n) {
  return guarded(s, [&] {
    const HostSlot& h = s->hs[0];
    need(dst4 != nullptr && n == h.d.n, GS_E_INVALID, "read_positions needs N");
    // compact_positions (nbody_bh.wgsl:186-195): xyz of posm, w = energy
    check(gs_download_positions(s, 0, dst4, (uint32_t)n), "positions");
    sync(s);
  });
}

int gs_download_codes(gs_sim* s, uint32_t* dst, uint64_t n, int sorted) {
  return guarded(s, [&] {
    const HostSlot& h = s->hs[0];
    need(dst != nullptr && n == h.d.n, GS_E_INVALID, "codes need N");
    need(sorted ? h.have_tree : h.have_codes, GS_E_STATE, "no codes yet");
    download_slot(s, 0, dst, sorted ? s->keys[gs_sorted_buf(s)] : s->codes_raw, 4, n, "codes");
  });
}

int gs_download_perm(gs_sim* s, uint32_t* dst, uint64_t n) {
  return gs_slot_download_perm(s, 0, dst, n);
}

int gs_slot_download_perm(gs_sim* s, uint32_t slot, uint32_t* dst, uint64_t n) {
  return guarded(s, [&] {
    const HostSlot& h = slot_of(s, slot);
    need(dst != nullptr && n == h.d.n, GS_E_INVALID, "perm needs n");
    need(h.have_tree, GS_E_STATE, "no sort yet");
    download_slot(s, slot, dst, s->vals[gs_sorted_buf(s)], 4, n, "perm");
  });
}

int gs_tree_info(gs_sim* s, uint64_t* num_nodes, uint32_t* level_ranges, uint32_t* num_levels,
                 double* root_size) {
  return gs_slot_tree_info(s, 0, num_nodes, level_ranges, num_levels, root_size);
}

int gs_slot_tree_info(gs_sim* s, uint32_t slot, uint64_t* num_nodes, uint32_t* level_ranges,
                      uint32_t* num_levels, double* root_size) {
  return guarded(s, [&] {
    const HostSlot& hs = slot_of(s, slot);
    need(hs.have_tree, GS_E_STATE, "no tree yet");
    check_status(s);
    const DevScalars h = read_scalars(s, slot);
    if (num_nodes) *num_nodes = h.num_nodes;
    if (num_levels) *num_levels = h.num_levels;
    if (level_ranges)
      for (uint32_t k = 0; k < h.num_levels && k < GS_MAX_LEVELS; k++) {
        level_ranges[2 * k] = h.level_start[k];
        level_ranges[2 * k + 1] = h.level_count[k];
      }
    if (root_size) memcpy(root_size, &h.size_bits, 8);
  });
}

int gs_download_nodes(gs_sim* s, uint32_t* dst, uint64_t cap, uint64_t* num_nodes) {
  return gs_slot_download_nodes(s, 0, dst, cap, num_nodes);
}

int gs_slot_download_nodes(gs_sim* s, uint32_t slot, uint32_t* dst, uint64_t cap,
                           uint64_t* num_nodes) {
  return guarded(s, [&] {
    const HostSlot& hs = slot_of(s, slot);
    need(hs.have_tree, GS_E_STATE, "no tree yet");
    need(hs.have_com, GS_E_STATE, "node COMs not filled since the last sort (refresh first)");
    check_status(s);
    const uint32_t nn = read_scalars(s, slot).num_nodes;
    if (num_nodes) *num_nodes = nn;
    need(dst != nullptr && cap >= nn, GS_E_INVALID, "node buffer too small");
    ensure_node64(s, hs.d.ncap);
    check(gs_launch_node64(s, slot), "node64");
    check(hipMemcpyAsync(dst, s->node64, (size_t)nn * 64, hipMemcpyDeviceToHost, s->stream),
          "nodes");
    sync(s);
  });
}

int gs_download_energy(gs_sim* s, float* dst, uint64_t n) {
  return gs_slot_download_energy(s, 0, dst, n);
}

int gs_slot_download_energy(gs_sim* s, uint32_t slot, float* dst, uint64_t n) {
  return guarded(s, [&] {
    const HostSlot& h = slot_of(s, slot);
    need(dst != nullptr && n == h.d.n, GS_E_INVALID, "energy needs n");
    need(h.have_step, GS_E_STATE, "no step yet");
    download_slot(s, slot, dst, s->energy, 4, n, "energy");
  });
}

int gs_download_accel(gs_sim* s, float* acc4, float* density, uint64_t n) {
  return guarded(s, [&] {
    const HostSlot& h = s->hs[0];
    need(n == h.d.n, GS_E_INVALID, "accel needs N");
    need(h.have_force, GS_E_STATE, "no force yet");
    check_status(s);
    if (acc4) download_slot(s, 0, acc4, s->acc, 16, n, "acc");
    if (density) download_slot(s, 0, density, s->dens, 4, n, "density");
  });
}

int gs_reduce(gs_sim* s, double out[GS_NRED]) { return gs_slot_reduce(s, 0, out); }

int gs_slot_reduce(gs_sim* s, uint32_t slot, double out[GS_NRED]) {
  return guarded(s, [&] {
    const HostSlot& h = slot_of(s, slot);
    need(out != nullptr, GS_E_INVALID, "null output");
    need(h.have_step, GS_E_STATE, "no step yet");
    need(h.have_red, GS_E_STATE, "slots changed since the last step: its reductions are gone");
    check(gs_launch_reduce_final(s, slot, s->red_out), "reduce");
    check(hipMemcpyAsync(out, s->red_out, sizeof(double) * GS_NRED, hipMemcpyDeviceToHost,
                         s->stream),
          "reduce");
    sync(s);
  });
}

int gs_upload_stellar(gs_sim* s, const float* stellar4, uint64_t n) {
  return gs_slot_upload_stellar(s, 0, stellar4, n);
}

int gs_slot_upload_stellar(gs_sim* s, uint32_t slot, const float* stellar4, uint64_t n) {
  return guarded(s, [&] {
    need(!s->arena, GS_E_INVALID, "an arena holds no stellar records");
    const HostSlot& h = slot_of(s, slot);
    need(stellar4 != nullptr && n == h.d.n, GS_E_INVALID,
         "stellar records must match the slot's particles (n x 4 f32)");
    ensure_stellar(s);
    if (n)
      check(hipMemcpyAsync(s->stel[s->stel_cur] + h.d.poff, stellar4, 16 * n,
                           hipMemcpyHostToDevice, s->stream),
            "stellar upload");
    sync(s);
  });
}

int gs_download_stellar(gs_sim* s, float* dst4, uint64_t n) {
  return gs_slot_download_stellar(s, 0, dst4, n);
}

int gs_slot_download_stellar(gs_sim* s, uint32_t slot, float* dst4, uint64_t n) {
  return guarded(s, [&] {
    const HostSlot& h = slot_of(s, slot);
    need(dst4 != nullptr && n == h.d.n, GS_E_INVALID, "stellar download needs n x 4 f32");
    need(s->stel[0] != nullptr, GS_E_STATE, "no stellar records uploaded");
    download_slot(s, slot, dst4, s->stel[s->stel_cur], 16, n, "stellar download");
  });
}

int gs_locate(gs_sim* s, uint32_t old_index, uint32_t* new_index, float pos4[4]) {
  return gs_slot_locate(s, 0, old_index, new_index, pos4);
}

int gs_slot_locate(gs_sim* s, uint32_t slot, uint32_t old_index, uint32_t* new_index,
                   float pos4[4]) {
  return guarded(s, [&] {
    const HostSlot& h = slot_of(s, slot);
    need(h.have_tree, GS_E_STATE, "no sort yet");
    need(old_index < h.d.n, GS_E_INVALID, "index out of range");
    check(gs_launch_locate(s, slot, old_index, &s->batch->located), "locate");
    uint32_t ni = 0xffffffffu;
    check(hipMemcpyAsync(&ni, &s->batch->located, 4, hipMemcpyDeviceToHost, s->stream), "locate");
    sync(s);
    need(ni < h.d.n, GS_E_STATE, "index not found in the last permutation");
    if (new_index) *new_index = ni;
    if (pos4) {
      check(hipMemcpyAsync(pos4, s->posm[s->cur] + h.d.poff + ni, 16, hipMemcpyDeviceToHost,
                           s->stream),
            "locate position");
      sync(s);
    }
  });
}

int gs_star_formation(gs_sim* s) {
  return guarded(s, [&] {
    need(s->stel[0] != nullptr, GS_E_STATE, "star formation needs stellar records");
    if (!ready(s)) return;
    check(gs_launch_star_formation(s), "star formation");
  });
}

int gs_stage_times(gs_sim* s, double ms[GS_NT], uint64_t counts[GS_NT], int reset) {
  return guarded(s, [&] {
    collect_timers(s);
    for (int k = 0; k < GS_NT; k++) {
      if (ms) ms[k] = s->t_ms[k];
      if (counts) counts[k] = s->t_cnt[k];
      if (reset) {
        s->t_ms[k] = 0;
        s->t_cnt[k] = 0;
      }
    }
  });
}

int gs_count_interactions(gs_sim* s, uint64_t* total, uint32_t* per_particle) {
  return guarded(s, [&] {
    need_com(s, "counting needs a tree with COMs (sort, then refresh or upload a tree)");
    if (total) *total = 0;
    if (!ready(s)) return;
    uint64_t zero = 0;
    check(hipMemcpyAsync(&s->batch->interactions, &zero, 8, hipMemcpyHostToDevice, s->stream),
          "zero");
    check(gs_launch_walk(s, false, true), "count walk");
    uint64_t t = 0;
    check(hipMemcpyAsync(&t, &s->batch->interactions, 8, hipMemcpyDeviceToHost, s->stream),
          "total");
    if (per_particle) download_slot(s, 0, per_particle, s->inter, 4, s->hs[0].d.n, "counts");
    sync(s);
    if (total) *total = t;
    for (auto& h : s->hs) h.have_force = h.d.n > 0;
  });
}

int gs_compute_morton_codes(gs_sim* s, const float* pos, uint32_t n, uint32_t stride,
                            uint32_t* codes, double* root_size) {
  return guarded(s, [&] {
    HostSlot& h = s->hs[0];
    need(pos && codes && n >= 1 && n <= h.d.cap && stride >= 3, GS_E_INVALID,
         "bad morton arguments");
    need(s->nslots == 1 && !s->arena, GS_E_INVALID, "host-array functions need a one-slot handle");
    // strided host positions -> posm of the current buffer (mass unused)
    std::vector<float> tmp((size_t)n * 4);
    for (size_t i = 0; i < n; i++) {
      tmp[4 * i] = pos[i * stride];
      tmp[4 * i + 1] = pos[i * stride + 1];
      tmp[4 * i + 2] = pos[i * stride + 2];
      tmp[4 * i + 3] = 0.f;
    }
    check(hipMemcpyAsync(s->posm[s->cur] + h.d.poff, tmp.data(), 16ull * n, hipMemcpyHostToDevice,
                         s->stream),
          "pos");
    h.d.n = n;
    push_slot(s, 0);
    check(gs_launch_bbox(s, s->posm[s->cur]), "bbox");
    check(gs_launch_root(s), "root");
    check(gs_launch_encode(s, s->posm[s->cur], true), "encode");
    uint64_t sz = 0;
    check(hipMemcpyAsync(codes, s->codes_raw + h.d.poff, 4ull * n, hipMemcpyDeviceToHost,
                         s->stream),
          "codes");
    check(hipMemcpyAsync(&sz, &s->sc->size_bits, 8, hipMemcpyDeviceToHost, s->stream), "size");
    sync(s);
    if (root_size) memcpy(root_size, &sz, 8);
    h.have_tree = h.have_com = false;
    h.have_codes = true;
  });
}

int gs_radix_sort_indices(gs_sim* s, const uint32_t* codes, uint32_t n, uint32_t* perm) {
  return guarded(s, [&] {
    HostSlot& h = s->hs[0];
    need(codes && perm && n >= 1 && n <= h.d.cap, GS_E_INVALID, "bad sort arguments");
    need(s->nslots == 1 && !s->arena, GS_E_INVALID, "host-array functions need a one-slot handle");
    h.d.n = n;
    push_slot(s, 0);
    check(hipMemcpyAsync(s->keys[0] + h.d.poff, codes, 4ull * n, hipMemcpyHostToDevice, s->stream),
          "codes");
    check(gs_launch_reset(s), "reset");
    check(gs_launch_hist(s), "hist");
    check(gs_launch_sort(s, false), "sort");
    check(hipMemcpyAsync(perm, s->vals[gs_sorted_buf(s)] + h.d.poff, 4ull * n,
                         hipMemcpyDeviceToHost, s->stream),
          "perm");
    sync(s);
    h.have_tree = h.have_com = false;
  });
}

int gs_tree_build(gs_sim* s, const uint32_t* sorted_codes, const float* pos3, const float* masses,
                  uint32_t n, double root_size, uint32_t* nodes, uint64_t cap, uint64_t* num_nodes,
                  uint32_t* level_ranges, uint32_t* num_levels) {
  return guarded(s, [&] {
    HostSlot& h = s->hs[0];
    need(sorted_codes && pos3 && masses && n >= 1 && n <= h.d.cap, GS_E_INVALID,
         "bad build arguments");
    need(s->nslots == 1 && !s->arena, GS_E_INVALID, "host-array functions need a one-slot handle");
    std::vector<float> pm((size_t)n * 4);
    for (size_t i = 0; i < n; i++) {
      pm[4 * i] = pos3[3 * i];
      pm[4 * i + 1] = pos3[3 * i + 1];
      pm[4 * i + 2] = pos3[3 * i + 2];
      pm[4 * i + 3] = masses[i];
    }
    h.d.n = n;
    push_slot(s, 0);
    check(hipMemcpyAsync(s->posm[s->cur] + h.d.poff, pm.data(), 16ull * n, hipMemcpyHostToDevice,
                         s->stream),
          "pos");
    check(hipMemcpyAsync(s->keys[gs_sorted_buf(s)] + h.d.poff, sorted_codes, 4ull * n,
                         hipMemcpyHostToDevice, s->stream),
          "codes");
    // root segment and per-level sizes from the caller's rootSize
    DevScalars sc = read_scalars(s, 0);
    double sz = root_size;
    float min_s2 = INFINITY;
    for (int d = 0; d <= GS_MAX_LEVELS; d++) {
      sc.level_size[d] = (float)sz;
      if (d <= GS_MAX_LEVEL) min_s2 = fminf(min_s2, host_s2(sc.level_size[d]));
      sz *= 0.5;
    }
    sc.min_s2 = min_s2;
    for (int d = 0; d <= GS_MAX_LEVELS; d++)
      sc.level_thr[d] = gs_open_threshold(host_s2(sc.level_size[d]), s->host_prm[0].theta2);
    memcpy(&sc.size_bits, &root_size, 8);
    for (int d = 0; d <= GS_MAX_LEVELS; d++) sc.level_start[d] = sc.level_count[d] = 0;
    sc.level_count[0] = 1;
    sc.num_nodes = 1;
    sc.num_levels = 0;
    check(hipMemcpyAsync(s->sc, &sc, sizeof sc, hipMemcpyHostToDevice, s->stream), "scalars");
    DevBatch b;
    memset(&b, 0, sizeof b);
    check(hipMemcpyAsync(s->batch, &b, sizeof b, hipMemcpyHostToDevice, s->stream), "batch");
    const float f0 = sc.level_size[0];
    const int4 root = make_int4((int)sc.level_thr[0], 0, (int)gs_meta_open(0, n), -1);
    check(hipMemcpyAsync(&s->nodes[h.d.noff].info, &root, sizeof root, hipMemcpyHostToDevice,
                         s->stream),
          "root");
    check(hipMemcpyAsync(&s->nsize[h.d.noff], &f0, sizeof f0, hipMemcpyHostToDevice, s->stream),
          "root size");
    // the codes come from the caller: no histogram of them
    check(gs_launch_build(s, false), "build");
    check(gs_launch_refresh(s, s->posm[s->cur]), "refresh");
    h.have_tree = true;
    h.have_com = true;
    check_status(s);
    const DevScalars r = read_scalars(s, 0);
    if (num_nodes) *num_nodes = r.num_nodes;
    if (num_levels) *num_levels = r.num_levels;
    if (level_ranges)
      for (uint32_t k = 0; k < r.num_levels && k < GS_MAX_LEVELS; k++) {
        level_ranges[2 * k] = r.level_start[k];
        level_ranges[2 * k + 1] = r.level_count[k];
      }
    if (nodes) {
      need(cap >= r.num_nodes, GS_E_INVALID, "node buffer too small");
      ensure_node64(s, s->max_ncap);
      check(gs_launch_node64(s, 0), "node64");
      check(hipMemcpyAsync(nodes, s->node64, (size_t)r.num_nodes * 64, hipMemcpyDeviceToHost,
                           s->stream),
            "nodes");
      sync(s);
    }
  });
}

}  // extern "C"
