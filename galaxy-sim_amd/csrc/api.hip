// C ABI of libgalaxy_hip.so (include/galaxy_hip.h): handle lifetime, uploads,
// stage sequencing on the handle's stream, downloads, and the mortonTree.js
// functions on host arrays.  No exception crosses the ABI: every entry point
// catches, records the message on the handle and returns a GS_E* status.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>

#include "gs_internal.h"

#define GS_VERSION "galaxy_hip 0.1 (gfx950)"

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

// ---- stage timing ---------------------------------------------------------
struct Stage {
  gs_sim* s;
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  Stage(gs_sim* sim, int stage) : s(sim), id(stage) {
    if (!(s->flags & GS_FLAG_TIMING)) return;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    check(hipEventRecord(a, s->stream), "hipEventRecord");
  }
  ~Stage() {
    if (!a) return;
    if (hipEventRecord(b, s->stream) == hipSuccess) {
      s->timers.push_back(StageTimer{id, a, b});
    } else {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  }
};

void collect_timers(gs_sim* s) {
  if (s->timers.empty()) return;
  check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  for (auto& t : s->timers) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      s->t_ms[t.stage] += ms;
      s->t_cnt[t.stage] += 1;
    }
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  s->timers.clear();
}

void check_status(gs_sim* s) {
  uint32_t st = 0;
  check(hipMemcpyAsync(&st, &s->sc->status, sizeof st, hipMemcpyDeviceToHost, s->stream),
        "read status");
  check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  if (st & GS_ST_NODE_POOL)
    throw Fail{GS_E_NODE_POOL, "octree needs more nodes than the pool holds (raise max_nodes)"};
  if (st & GS_ST_STACK) throw Fail{GS_E_STACK, "walk stack overflow"};
  if (st & GS_ST_SPIN) throw Fail{GS_E_DEVICE, "a look-back wait timed out (scan state corrupt)"};
}

void free_all(gs_sim* s) {
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (int k = 0; k < 2; k++) {
    f(s->posm[k]);
    f(s->velp[k]);
    f(s->keys[k]);
    f(s->vals[k]);
    f(s->build_status[k]);
  }
  f(s->codes_raw);
  f(s->sort_status);
  f(s->bbox_part);
  f(s->tree.node);
  f(s->energy);
  f(s->acc);
  f(s->dens);
  f(s->inter);
  f(s->red_part);
  f(s->node64);
  f(s->sc);
  f(s->prm);
  for (auto& t : s->timers) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  s->timers.clear();
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
}

// ---- stage sequences --------------------------------------------------------
void do_sort(gs_sim* s, bool keep_raw) {
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
  {
    Stage t(s, GS_T_BUILD);
    check(gs_launch_build(s), "build");
  }
  s->have_tree = true;
  s->have_force = false;
}

void do_refresh(gs_sim* s) {
  Stage t(s, GS_T_REFRESH);
  check(gs_launch_refresh(s, s->posm[s->cur]), "refresh");
}

void do_walk_step(gs_sim* s) {
  {
    Stage t(s, GS_T_WALK);
    check(gs_launch_walk(s, true, false, true), "walk");
  }
  s->cur ^= 1;
  s->have_step = true;
  s->have_force = false;
}

}  // namespace

hipError_t gs_launch_node64(gs_sim* s);

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

int gs_create(const gs_config* cfg, gs_sim** out) {
  if (!cfg || !out) return GS_E_INVALID;
  *out = nullptr;
  if (cfg->n == 0 || cfg->n > GS_MAX_N) return GS_E_INVALID;
  uint32_t wg = cfg->walk_wg ? cfg->walk_wg : 256;
  if (wg < 64 || wg > 1024 || (wg & 63)) return GS_E_INVALID;
  if (cfg->radix_bits != 0 && cfg->radix_bits != GS_RADIX_BITS) return GS_E_INVALID;
  gs_sim* s = new (std::nothrow) gs_sim();
  if (!s) return GS_E_NOMEM;
  s->n = cfg->n;
  s->bucket = cfg->bucket ? cfg->bucket : 16;
  s->flags = cfg->flags;
  s->walk_wg = wg;
  const uint64_t pool = 2ull * cfg->n + 4096;  // mortonTree.js:88
  s->max_nodes = cfg->max_nodes ? cfg->max_nodes : pool;
  const uint64_t cap = std::max<uint64_t>(pool, std::min<uint64_t>(s->max_nodes, 0xfffffff0ull));
  s->tree.cap = (uint32_t)cap;
  s->sort_tiles = (s->n + GS_SORT_TILE - 1) / GS_SORT_TILE;
  s->build_tiles = (uint32_t)((cap + GS_BUILD_TILE - 1) / GS_BUILD_TILE);
  s->red_blocks = (s->n + wg - 1) / wg;
  int dev = cfg->device;
  int rc = guarded(s, [&] {
    int ndev = 0;
    check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    need(ndev > 0, GS_E_DEVICE, "no HIP device");
    if (dev < 0) check(hipGetDevice(&dev), "hipGetDevice");
    need(dev < ndev, GS_E_INVALID, "device ordinal out of range");
    s->device = dev;
    check(hipSetDevice(dev), "hipSetDevice");
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    need(strncmp(prop.gcnArchName, "gfx950", 6) == 0, GS_E_DEVICE,
         "libgalaxy_hip is built for gfx950 (MI355X) only");
    check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    const size_t n = s->n;
    for (int k = 0; k < 2; k++) {
      dalloc(&s->posm[k], n, "posm");
      dalloc(&s->velp[k], n, "velp");
      dalloc(&s->keys[k], n, "keys");
      dalloc(&s->vals[k], n, "vals");
      dalloc(&s->build_status[k], s->build_tiles, "build status");
    }
    dalloc(&s->codes_raw, n, "codes");
    dalloc(&s->sort_status, (size_t)GS_SORT_PASSES * s->sort_tiles * GS_RADIX, "sort status");
    dalloc(&s->bbox_part, (size_t)GS_BBOX_BLOCKS * 8, "bbox");
    dalloc(&s->tree.node, cap, "nodes");
    dalloc(&s->energy, n, "energy");
    dalloc(&s->acc, n, "acc");
    dalloc(&s->dens, n, "density");
    dalloc(&s->inter, n, "interactions");
    dalloc(&s->red_part, (size_t)s->red_blocks * GS_NRED, "reductions");
    dalloc(&s->sc, 1, "scalars");
    dalloc(&s->prm, 1, "params");
    // zero state with fill kernels (graph-safe, no memset nodes)
    check(gs_launch_fill_u32((uint32_t*)s->sc, 0, sizeof(DevScalars) / 4, s->stream), "fill");
    check(gs_launch_fill_u32(s->energy ? (uint32_t*)s->energy : nullptr, 0, n, s->stream), "fill");
    for (int k = 0; k < 2; k++)
      check(gs_launch_fill_u32(s->build_status[k] ? (uint32_t*)s->build_status[k] : nullptr, 0,
                               2ull * s->build_tiles, s->stream),
            "fill");
    check(gs_launch_fill_u32(s->sort_status, 0, (size_t)GS_SORT_PASSES * s->sort_tiles * GS_RADIX,
                             s->stream),
          "fill");
    gs_params p;
    gs_params_default(&p);
    pack_params(&p, s->n, &s->host_prm);
    check(hipMemcpyAsync(s->prm, &s->host_prm, sizeof(DevParams), hipMemcpyHostToDevice,
                         s->stream),
          "params");
    check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  });
  if (rc != GS_OK) {
    std::string msg = s->err;
    free_all(s);
    delete s;
    (void)msg;
    return rc;
  }
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

int gs_upload(gs_sim* s, const float* particles, uint64_t n) {
  return guarded(s, [&] {
    need(particles != nullptr && n == s->n, GS_E_INVALID, "upload needs exactly N records");
    // 32-byte records -> (posm, velp): two strided copies, no staging buffer
    check(hipMemcpy2DAsync(s->posm[s->cur], 16, particles, 32, 16, n, hipMemcpyHostToDevice,
                           s->stream),
          "upload positions");
    check(hipMemcpy2DAsync(s->velp[s->cur], 16, particles + 4, 32, 16, n, hipMemcpyHostToDevice,
                           s->stream),
          "upload velocities");
    check(hipStreamSynchronize(s->stream), "upload");
    s->have_tree = s->have_force = s->have_step = false;
  });
}

int gs_set_params(gs_sim* s, const gs_params* p) {
  return guarded(s, [&] {
    need(p != nullptr, GS_E_INVALID, "null params");
    pack_params(p, s->n, &s->host_prm);
    check(hipMemcpyAsync(s->prm, &s->host_prm, sizeof(DevParams), hipMemcpyHostToDevice,
                         s->stream),
          "params");
    check(hipStreamSynchronize(s->stream), "params");
    s->have_params = true;
  });
}

int gs_update_central_mass(gs_sim* s, float mass, uint32_t idx) {
  return guarded(s, [&] {
    need(idx < s->n, GS_E_INVALID, "index out of range");
    // mass is the w of posm in both ping-pong buffers (NBodyGPU.js:248-250)
    for (int k = 0; k < 2; k++)
      check(hipMemcpyAsync((char*)(s->posm[k] + idx) + 12, &mass, 4, hipMemcpyHostToDevice,
                           s->stream),
            "central mass");
    check(hipStreamSynchronize(s->stream), "central mass");
  });
}

int gs_apply_permutation(gs_sim* s, const uint32_t* perm, uint64_t n) {
  return guarded(s, [&] {
    need(perm != nullptr && n == s->n, GS_E_INVALID, "permutation needs N entries");
    std::vector<uint8_t> seen(n, 0);
    for (uint64_t i = 0; i < n; i++) {
      need(perm[i] < n && !seen[perm[i]], GS_E_INVALID, "not a permutation of [0, N)");
      seen[perm[i]] = 1;
    }
    check(hipMemcpyAsync(s->vals[0], perm, 4 * n, hipMemcpyHostToDevice, s->stream), "perm");
    check(gs_launch_gather(s, s->vals[0]), "gather");
    check(hipStreamSynchronize(s->stream), "gather");
    s->cur ^= 1;
    s->have_force = false;
  });
}

int gs_upload_tree(gs_sim* s, const uint32_t* nodes, uint64_t num_nodes,
                   const uint32_t* level_ranges, uint32_t num_levels) {
  return guarded(s, [&] {
    need(nodes != nullptr && num_nodes >= 1 && num_nodes <= s->tree.cap, GS_E_INVALID,
         "node count outside [1, pool]");
    need(num_levels <= GS_MAX_LEVELS && (num_levels == 0 || level_ranges), GS_E_INVALID,
         "bad level ranges");
    // host validation: every index the device will follow stays in bounds;
    // the walk's (skip, depth) links follow from the BFS numbering
    std::vector<int2> links(num_nodes, make_int2(-1, 0));
    for (uint64_t i = 0; i < num_nodes; i++) {
      const uint32_t* w = nodes + i * 16;
      if (w[5] == 1u) {
        need(w[6] <= s->n && w[7] <= s->n - w[6], GS_E_INVALID, "leaf range outside [0, N)");
      } else {
        int32_t prev = -1;
        for (int k = 0; k < 8; k++) {
          const int32_t c = (int32_t)w[8 + k];
          need(c < 0 || ((uint64_t)c > i && (uint64_t)c < num_nodes), GS_E_INVALID,
               "child index must follow its parent and lie inside the tree");
          if (c >= 0) {
            need(links[c].y == 0, GS_E_INVALID, "node has two parents");
            links[c] = make_int2(prev >= 0 ? prev : links[i].x, links[i].y + 1);
            need(links[c].y <= GS_MAX_LEVEL, GS_E_INVALID, "tree deeper than 10 levels");
            prev = c;
          }
        }
      }
    }
    if (!s->node64) dalloc(&s->node64, (size_t)s->tree.cap * 16, "node staging");
    check(hipMemcpyAsync(s->node64, nodes, num_nodes * 64, hipMemcpyHostToDevice, s->stream),
          "nodes");
    DevScalars h;
    check(hipMemcpyAsync(&h, s->sc, sizeof h, hipMemcpyDeviceToHost, s->stream), "scalars");
    check(hipStreamSynchronize(s->stream), "sync");
    h.num_nodes = (uint32_t)num_nodes;
    h.num_levels = num_levels;
    for (uint32_t k = 0; k <= GS_MAX_LEVELS; k++) {
      h.level_start[k] = k < num_levels ? level_ranges[2 * k] : 0;
      h.level_count[k] = k < num_levels ? level_ranges[2 * k + 1] : 0;
    }
    h.status = 0;
    check(hipMemcpyAsync(s->sc, &h, sizeof h, hipMemcpyHostToDevice, s->stream), "scalars");
    int2* dlinks = nullptr;
    dalloc(&dlinks, num_nodes, "links");
    hipError_t e = hipMemcpyAsync(dlinks, links.data(), sizeof(int2) * num_nodes,
                                  hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = gs_launch_import64(s, s->node64, dlinks, (uint32_t)num_nodes);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dlinks);
    check(e, "import");
    check_status(s);  // non-contiguous children are rejected here
    s->have_tree = true;
    s->have_force = false;
  });
}

int gs_encode(gs_sim* s) {
  return guarded(s, [&] {
    Stage t(s, GS_T_ENCODE);
    check(gs_launch_bbox(s, s->posm[s->cur]), "bbox");
    check(gs_launch_root(s), "root");
    check(gs_launch_encode(s, s->posm[s->cur], true), "encode");
    s->have_codes = true;
  });
}

int gs_sort(gs_sim* s) {
  return guarded(s, [&] {
    do_sort(s, true);
    s->have_codes = true;
  });
}

int gs_refresh(gs_sim* s) {
  return guarded(s, [&] {
    need(s->have_tree, GS_E_STATE, "refresh before any sort");
    do_refresh(s);
  });
}

int gs_force(gs_sim* s) {
  return guarded(s, [&] {
    need(s->have_tree, GS_E_STATE, "force before any sort");
    Stage t(s, GS_T_WALK);
    check(gs_launch_walk(s, false, false, false), "walk");
    s->have_force = true;
  });
}

int gs_step(gs_sim* s, uint32_t nsteps, uint32_t sort_every) {
  return guarded(s, [&] {
    need(s->have_tree || sort_every > 0, GS_E_STATE, "step without a tree and sort_every = 0");
    for (uint32_t k = 0; k < nsteps; k++) {
      Stage t(s, GS_T_STEP);
      if (sort_every > 0 && (s->step_count % sort_every == 0 || !s->have_tree)) do_sort(s, false);
      do_refresh(s);
      do_walk_step(s);
      s->step_count++;
    }
  });
}

int gs_sync(gs_sim* s) {
  return guarded(s, [&] { check_status(s); });
}

int gs_download_particles(gs_sim* s, float* dst, uint64_t n) {
  return guarded(s, [&] {
    need(dst != nullptr && n == s->n, GS_E_INVALID, "download needs exactly N records");
    check(hipMemcpy2DAsync(dst, 32, s->posm[s->cur], 16, 16, n, hipMemcpyDeviceToHost, s->stream),
          "download positions");
    check(hipMemcpy2DAsync(dst + 4, 32, s->velp[s->cur], 16, 16, n, hipMemcpyDeviceToHost,
                           s->stream),
          "download velocities");
    check(hipStreamSynchronize(s->stream), "download");
  });
}

int gs_read_positions(gs_sim* s, float* dst4, uint64_t n) {
  return guarded(s, [&] {
    need(dst4 != nullptr && n == s->n, GS_E_INVALID, "read_positions needs N");
    std::vector<float> pm((size_t)n * 4), en(n);
    check(hipMemcpyAsync(pm.data(), s->posm[s->cur], 16 * n, hipMemcpyDeviceToHost, s->stream),
          "positions");
    check(hipMemcpyAsync(en.data(), s->energy, 4 * n, hipMemcpyDeviceToHost, s->stream), "energy");
    check(hipStreamSynchronize(s->stream), "sync");
    for (size_t i = 0; i < n; i++) {
      dst4[4 * i] = pm[4 * i];
      dst4[4 * i + 1] = pm[4 * i + 1];
      dst4[4 * i + 2] = pm[4 * i + 2];
      dst4[4 * i + 3] = en[i];
    }
  });
}

int gs_download_codes(gs_sim* s, uint32_t* dst, uint64_t n, int sorted) {
  return guarded(s, [&] {
    need(dst != nullptr && n == s->n, GS_E_INVALID, "codes need N");
    need(sorted ? s->have_tree : s->have_codes, GS_E_STATE, "no codes yet");
    check(hipMemcpyAsync(dst, sorted ? s->keys[0] : s->codes_raw, 4 * n, hipMemcpyDeviceToHost,
                         s->stream),
          "codes");
    check(hipStreamSynchronize(s->stream), "sync");
  });
}

int gs_download_perm(gs_sim* s, uint32_t* dst, uint64_t n) {
  return guarded(s, [&] {
    need(dst != nullptr && n == s->n, GS_E_INVALID, "perm needs N");
    need(s->have_tree, GS_E_STATE, "no sort yet");
    check(hipMemcpyAsync(dst, s->vals[0], 4 * n, hipMemcpyDeviceToHost, s->stream), "perm");
    check(hipStreamSynchronize(s->stream), "sync");
  });
}

int gs_tree_info(gs_sim* s, uint64_t* num_nodes, uint32_t* level_ranges, uint32_t* num_levels,
                 double* root_size) {
  return guarded(s, [&] {
    need(s->have_tree, GS_E_STATE, "no tree yet");
    check_status(s);
    DevScalars h;
    check(hipMemcpyAsync(&h, s->sc, sizeof h, hipMemcpyDeviceToHost, s->stream), "scalars");
    check(hipStreamSynchronize(s->stream), "sync");
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
  return guarded(s, [&] {
    need(s->have_tree, GS_E_STATE, "no tree yet");
    check_status(s);
    uint32_t nn = 0;
    check(hipMemcpyAsync(&nn, &s->sc->num_nodes, 4, hipMemcpyDeviceToHost, s->stream), "nodes");
    check(hipStreamSynchronize(s->stream), "sync");
    if (num_nodes) *num_nodes = nn;
    need(dst != nullptr && cap >= nn, GS_E_INVALID, "node buffer too small");
    if (!s->node64) dalloc(&s->node64, (size_t)s->tree.cap * 16, "node staging");
    check(gs_launch_node64(s), "node64");
    check(hipMemcpyAsync(dst, s->node64, (size_t)nn * 64, hipMemcpyDeviceToHost, s->stream),
          "nodes");
    check(hipStreamSynchronize(s->stream), "sync");
  });
}

int gs_download_energy(gs_sim* s, float* dst, uint64_t n) {
  return guarded(s, [&] {
    need(dst != nullptr && n == s->n, GS_E_INVALID, "energy needs N");
    need(s->have_step, GS_E_STATE, "no step yet");
    check(hipMemcpyAsync(dst, s->energy, 4 * n, hipMemcpyDeviceToHost, s->stream), "energy");
    check(hipStreamSynchronize(s->stream), "sync");
  });
}

int gs_download_accel(gs_sim* s, float* acc4, float* density, uint64_t n) {
  return guarded(s, [&] {
    need(n == s->n, GS_E_INVALID, "accel needs N");
    need(s->have_force, GS_E_STATE, "no force yet");
    check_status(s);
    if (acc4)
      check(hipMemcpyAsync(acc4, s->acc, 16 * n, hipMemcpyDeviceToHost, s->stream), "acc");
    if (density)
      check(hipMemcpyAsync(density, s->dens, 4 * n, hipMemcpyDeviceToHost, s->stream), "dens");
    check(hipStreamSynchronize(s->stream), "sync");
  });
}

int gs_reduce(gs_sim* s, double out[GS_NRED]) {
  return guarded(s, [&] {
    need(out != nullptr, GS_E_INVALID, "null output");
    need(s->have_step, GS_E_STATE, "no step yet");
    double* d = nullptr;
    dalloc(&d, GS_NRED, "reduce");
    hipError_t e = gs_launch_reduce_final(s, d);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out, d, sizeof(double) * GS_NRED, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    check(e, "reduce");
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
    need(s->have_tree, GS_E_STATE, "no tree yet");
    uint64_t zero = 0;
    check(hipMemcpyAsync(&s->sc->interactions, &zero, 8, hipMemcpyHostToDevice, s->stream),
          "zero");
    check(gs_launch_walk(s, false, true, false), "count walk");
    uint64_t t = 0;
    check(hipMemcpyAsync(&t, &s->sc->interactions, 8, hipMemcpyDeviceToHost, s->stream), "total");
    if (per_particle)
      check(hipMemcpyAsync(per_particle, s->inter, 4ull * s->n, hipMemcpyDeviceToHost, s->stream),
            "counts");
    check(hipStreamSynchronize(s->stream), "sync");
    if (total) *total = t;
    s->have_force = true;
  });
}

int gs_compute_morton_codes(gs_sim* s, const float* pos, uint32_t n, uint32_t stride,
                            uint32_t* codes, double* root_size) {
  return guarded(s, [&] {
    need(pos && codes && n == s->n && stride >= 3, GS_E_INVALID, "bad morton arguments");
    // strided host positions -> posm of the spare buffer (mass unused)
    std::vector<float> tmp((size_t)n * 4);
    for (size_t i = 0; i < n; i++) {
      tmp[4 * i] = pos[i * stride];
      tmp[4 * i + 1] = pos[i * stride + 1];
      tmp[4 * i + 2] = pos[i * stride + 2];
      tmp[4 * i + 3] = 0.f;
    }
    float4* work = s->posm[s->cur ^ 1];
    check(hipMemcpyAsync(work, tmp.data(), 16ull * n, hipMemcpyHostToDevice, s->stream), "pos");
    check(gs_launch_bbox(s, work), "bbox");
    check(gs_launch_root(s), "root");
    check(gs_launch_encode(s, work, true), "encode");
    uint64_t sz = 0;
    check(hipMemcpyAsync(codes, s->codes_raw, 4ull * n, hipMemcpyDeviceToHost, s->stream), "codes");
    check(hipMemcpyAsync(&sz, &s->sc->size_bits, 8, hipMemcpyDeviceToHost, s->stream), "size");
    check(hipStreamSynchronize(s->stream), "sync");
    if (root_size) memcpy(root_size, &sz, 8);
    s->have_tree = false;  // the root node slot was rewritten
  });
}

int gs_radix_sort_indices(gs_sim* s, const uint32_t* codes, uint32_t n, uint32_t* perm) {
  return guarded(s, [&] {
    need(codes && perm && n == s->n, GS_E_INVALID, "bad sort arguments");
    check(hipMemcpyAsync(s->keys[0], codes, 4ull * n, hipMemcpyHostToDevice, s->stream), "codes");
    check(gs_launch_reset(s), "reset");
    check(gs_launch_hist(s), "hist");
    check(gs_launch_sort(s, false), "sort");
    check(hipMemcpyAsync(perm, s->vals[0], 4ull * n, hipMemcpyDeviceToHost, s->stream), "perm");
    check(hipStreamSynchronize(s->stream), "sync");
    s->have_tree = false;
  });
}

int gs_tree_build(gs_sim* s, const uint32_t* sorted_codes, const float* pos3, const float* masses,
                  uint32_t n, double root_size, uint32_t* nodes, uint64_t cap, uint64_t* num_nodes,
                  uint32_t* level_ranges, uint32_t* num_levels) {
  return guarded(s, [&] {
    need(sorted_codes && pos3 && masses && n == s->n, GS_E_INVALID, "bad build arguments");
    std::vector<float> pm((size_t)n * 4);
    for (size_t i = 0; i < n; i++) {
      pm[4 * i] = pos3[3 * i];
      pm[4 * i + 1] = pos3[3 * i + 1];
      pm[4 * i + 2] = pos3[3 * i + 2];
      pm[4 * i + 3] = masses[i];
    }
    check(hipMemcpyAsync(s->posm[s->cur], pm.data(), 16ull * n, hipMemcpyHostToDevice, s->stream),
          "pos");
    check(hipMemcpyAsync(s->keys[0], sorted_codes, 4ull * n, hipMemcpyHostToDevice, s->stream),
          "codes");
    // root segment and per-level sizes from the caller's rootSize
    DevScalars h;
    check(hipMemcpyAsync(&h, s->sc, sizeof h, hipMemcpyDeviceToHost, s->stream), "scalars");
    check(hipStreamSynchronize(s->stream), "sync");
    double sz = root_size;
    for (int d = 0; d <= GS_MAX_LEVELS; d++) {
      h.level_size[d] = (float)sz;
      sz *= 0.5;
    }
    memcpy(&h.size_bits, &root_size, 8);
    h.level_start[0] = 0;
    h.level_count[0] = 1;
    for (int d = 1; d <= GS_MAX_LEVELS; d++) h.level_start[d] = h.level_count[d] = 0;
    h.num_nodes = 1;
    h.num_levels = 0;
    h.status = 0;
    for (int k = 0; k < 32; k++) h.tickets[k] = 0;
    check(hipMemcpyAsync(s->sc, &h, sizeof h, hipMemcpyHostToDevice, s->stream), "scalars");
    int4 root = make_int4(0, 0, (int)gs_meta_open(0, n), -1);
    float f0 = h.level_size[0];
    memcpy(&root.x, &f0, 4);
    check(hipMemcpyAsync(&s->tree.node[0].info, &root, sizeof root, hipMemcpyHostToDevice,
                         s->stream),
          "root");
    check(gs_launch_build(s), "build");
    check(gs_launch_refresh(s, s->posm[s->cur]), "refresh");
    s->have_tree = true;
    uint64_t nn = 0;
    uint32_t nl = 0;
    check_status(s);
    DevScalars r;
    check(hipMemcpyAsync(&r, s->sc, sizeof r, hipMemcpyDeviceToHost, s->stream), "scalars");
    check(hipStreamSynchronize(s->stream), "sync");
    nn = r.num_nodes;
    nl = r.num_levels;
    if (num_nodes) *num_nodes = nn;
    if (num_levels) *num_levels = nl;
    if (level_ranges)
      for (uint32_t k = 0; k < nl && k < GS_MAX_LEVELS; k++) {
        level_ranges[2 * k] = r.level_start[k];
        level_ranges[2 * k + 1] = r.level_count[k];
      }
    if (nodes) {
      need(cap >= nn, GS_E_INVALID, "node buffer too small");
      if (!s->node64) dalloc(&s->node64, (size_t)s->tree.cap * 16, "node staging");
      check(gs_launch_node64(s), "node64");
      check(hipMemcpyAsync(nodes, s->node64, nn * 64, hipMemcpyDeviceToHost, s->stream), "nodes");
      check(hipStreamSynchronize(s->stream), "sync");
    }
  });
}

}  // extern "C"
