// Thin N-API (v8) binding of libgalaxy_hip.so for Node.js.  Every function
// maps one-to-one onto a C-ABI entry point of include/galaxy_hip.h; typed
// arrays are passed zero-copy (napi_get_typedarray_info) and copied by the
// library.  A non-zero status becomes a thrown JS Error carrying
// gs_last_error().  readPositionsAsync runs gs_read_positions on the libuv
// pool and resolves a Promise, like NBodyGPU.readPositions (NBodyGPU.js:346).
#include <node_api.h>
#include <stdint.h>
#include <string.h>

#include <string>

#include "galaxy_hip.h"

namespace {

#define CALL(env, expr)                                   \
  do {                                                    \
    if ((expr) != napi_ok) {                              \
      napi_throw_error((env), nullptr, "N-API failure");  \
      return nullptr;                                     \
    }                                                     \
  } while (0)

// the thrown Error's code names the status (err.code === 'GS_E_FULL' when an
// arena has no range for a galaxy)
const char* gs_code_name(int rc) {
  switch (rc) {
    case GS_E_INVALID: return "GS_E_INVALID";
    case GS_E_DEVICE: return "GS_E_DEVICE";
    case GS_E_NOMEM: return "GS_E_NOMEM";
    case GS_E_NODE_POOL: return "GS_E_NODE_POOL";
    case GS_E_STATE: return "GS_E_STATE";
    case GS_E_STACK: return "GS_E_STACK";
    case GS_E_FULL: return "GS_E_FULL";
    default: return "GS_E_UNKNOWN";
  }
}

napi_value throw_gs(napi_env env, gs_sim* s, int rc, const char* what) {
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " +
                    (s ? gs_last_error(s) : "no handle");
  napi_throw_error(env, gs_code_name(rc), msg.c_str());
  return nullptr;
}

struct Args {
  napi_value v[8];
  size_t n = 8;
};

bool get_args(napi_env env, napi_callback_info info, Args& a) {
  return napi_get_cb_info(env, info, &a.n, a.v, nullptr, nullptr) == napi_ok;
}

// handle objects are {_h: External}; destroy() clears the external's pointer
struct Holder {
  gs_sim* sim;
};

void finalize_holder(napi_env, void* data, void*) {
  Holder* h = static_cast<Holder*>(data);
  if (h) {
    if (h->sim) gs_destroy(h->sim);
    delete h;
  }
}

gs_sim* sim_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a galaxy_hip handle");
    return nullptr;
  }
  Holder* h = static_cast<Holder*>(p);
  if (!h->sim) {
    napi_throw_error(env, nullptr, "handle already destroyed");
    return nullptr;
  }
  return h->sim;
}

template <class T>
T* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t* len) {
  napi_typedarray_type t;
  size_t length = 0;
  void* data = nullptr;
  napi_value ab;
  size_t off = 0;
  if (napi_get_typedarray_info(env, v, &t, &length, &data, &ab, &off) != napi_ok || t != want) {
    napi_throw_type_error(env, nullptr, "wrong typed array type");
    return nullptr;
  }
  if (len) *len = length;
  return static_cast<T*>(data);
}

double num_prop(napi_env env, napi_value obj, const char* key, double dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  napi_valuetype ty;
  napi_typeof(env, v, &ty);
  if (ty != napi_number) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

napi_value num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

napi_value level_ranges_js(napi_env env, const uint32_t* lr, uint32_t nl) {
  napi_value arr;
  napi_create_array_with_length(env, nl, &arr);
  for (uint32_t k = 0; k < nl; k++) {
    napi_value pair;
    napi_create_array_with_length(env, 2, &pair);
    napi_set_element(env, pair, 0, num(env, lr[2 * k]));
    napi_set_element(env, pair, 1, num(env, lr[2 * k + 1]));
    napi_set_element(env, arr, k, pair);
  }
  return arr;
}

// create({n, bucket, maxNodes, exactForce, timing, device, walkWg}) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) {
    napi_throw_type_error(env, nullptr, "create(options)");
    return nullptr;
  }
  gs_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.n = (uint32_t)num_prop(env, a.v[0], "n", 0);
  cfg.bucket = (uint32_t)num_prop(env, a.v[0], "bucket", 16);
  cfg.max_nodes = (uint64_t)num_prop(env, a.v[0], "maxNodes", 0);
  cfg.flags = (num_prop(env, a.v[0], "exactForce", 0) != 0 ? GS_FLAG_EXACT_FORCE : 0u) |
              (num_prop(env, a.v[0], "timing", 0) != 0 ? GS_FLAG_TIMING : 0u);
  cfg.device = (int32_t)num_prop(env, a.v[0], "device", -1);
  cfg.walk_wg = (uint32_t)num_prop(env, a.v[0], "walkWg", 0);
  cfg.radix_bits = (uint32_t)num_prop(env, a.v[0], "radixBits", 0);
  gs_sim* s = nullptr;
  int rc = gs_create(&cfg, &s);
  if (rc != GS_OK) return throw_gs(env, nullptr, rc, "gs_create");
  Holder* h = new Holder{s};
  napi_value ext;
  CALL(env, napi_create_external(env, h, finalize_holder, nullptr, &ext));
  return ext;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return nullptr;
  void* p = nullptr;
  if (napi_get_value_external(env, a.v[0], &p) == napi_ok && p) {
    Holder* h = static_cast<Holder*>(p);
    if (h->sim) gs_destroy(h->sim);
    h->sim = nullptr;
  }
  return nullptr;
}

#define SIM_OR_RETURN(var, val)      \
  gs_sim* var = sim_of(env, (val));  \
  if (!var) return nullptr;

napi_value Upload(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t len = 0;
  float* p = typed<float>(env, a.v[1], napi_float32_array, &len);
  if (!p) return nullptr;
  int rc = gs_upload(s, p, len / 8);
  return rc ? throw_gs(env, s, rc, "gs_upload") : nullptr;
}

// updateParams dictionary (NBodyGPU.js:253-275) -> gs_params
void parse_params(napi_env env, napi_value o, gs_params& p) {
  gs_params_default(&p);
  p.dt = num_prop(env, o, "dt", p.dt);
  p.G = num_prop(env, o, "G", p.G);
  p.epsilon = num_prop(env, o, "epsilon", p.epsilon);
  p.theta = num_prop(env, o, "theta", p.theta);
  p.halo_mass = num_prop(env, o, "haloMass", p.halo_mass);
  p.halo_a = num_prop(env, o, "haloA", p.halo_a);
  p.gas_drag = num_prop(env, o, "gasDrag", p.gas_drag);
  p.time = num_prop(env, o, "time", p.time);
  p.aging_rate = num_prop(env, o, "agingRate", p.aging_rate);
  p.sf_density = num_prop(env, o, "sfDensity", p.sf_density);
  p.sf_eff = num_prop(env, o, "sfEff", p.sf_eff);
  p.sf_dt = num_prop(env, o, "sfDt", p.sf_dt);
  bool has = false;
  napi_has_named_property(env, o, "haloCenters", &has);
  if (has) {
    napi_value hc;
    napi_get_named_property(env, o, "haloCenters", &hc);
    bool isarr = false;
    napi_is_array(env, hc, &isarr);
    if (isarr) {
      uint32_t len = 0;
      napi_get_array_length(env, hc, &len);
      p.halo_count = len > 2 ? 2 : len;
      for (uint32_t k = 0; k < p.halo_count; k++) {
        napi_value c;
        napi_get_element(env, hc, k, &c);
        for (uint32_t j = 0; j < 3; j++) {
          napi_value e;
          double d = 0;
          if (napi_get_element(env, c, j, &e) == napi_ok) napi_get_value_double(env, e, &d);
          p.halo_centers[k][j] = d;
        }
      }
    }
  }
}

napi_value SetParams(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  gs_params p;
  parse_params(env, a.v[1], p);
  int rc = gs_set_params(s, &p);
  return rc ? throw_gs(env, s, rc, "gs_set_params") : nullptr;
}

napi_value UpdateCentralMass(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  double m = 0, idx = 0;
  napi_get_value_double(env, a.v[1], &m);
  napi_get_value_double(env, a.v[2], &idx);
  int rc = gs_update_central_mass(s, (float)m, (uint32_t)idx);
  return rc ? throw_gs(env, s, rc, "gs_update_central_mass") : nullptr;
}

#define SIMPLE(NAME, FN)                                   \
  napi_value NAME(napi_env env, napi_callback_info info) { \
    Args a;                                                \
    get_args(env, info, a);                                \
    SIM_OR_RETURN(s, a.v[0]);                              \
    int rc = FN(s);                                        \
    return rc ? throw_gs(env, s, rc, #FN) : nullptr;       \
  }
SIMPLE(Encode, gs_encode)
SIMPLE(Sort, gs_sort)
SIMPLE(Refresh, gs_refresh)
SIMPLE(Force, gs_force)
SIMPLE(WalkStep, gs_walk_step)
SIMPLE(Sync, gs_sync)

napi_value Step(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  double n = 1, every = 0;
  if (a.n > 1) napi_get_value_double(env, a.v[1], &n);
  if (a.n > 2) napi_get_value_double(env, a.v[2], &every);
  int rc = gs_step(s, (uint32_t)n, (uint32_t)every);
  return rc ? throw_gs(env, s, rc, "gs_step") : nullptr;
}

napi_value ApplyPermutation(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t len = 0;
  uint32_t* p = typed<uint32_t>(env, a.v[1], napi_uint32_array, &len);
  if (!p) return nullptr;
  int rc = gs_apply_permutation(s, p, len);
  return rc ? throw_gs(env, s, rc, "gs_apply_permutation") : nullptr;
}

// uploadTree(h, Uint32Array|ArrayBuffer words, numNodes, levelRanges)
napi_value UploadTree(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  void* data = nullptr;
  size_t bytes = 0;
  bool isab = false;
  napi_is_arraybuffer(env, a.v[1], &isab);
  if (isab) {
    napi_get_arraybuffer_info(env, a.v[1], &data, &bytes);
  } else {
    size_t len = 0;
    data = typed<uint32_t>(env, a.v[1], napi_uint32_array, &len);
    if (!data) return nullptr;
    bytes = len * 4;
  }
  double nn = 0;
  napi_get_value_double(env, a.v[2], &nn);
  if ((size_t)nn * 64 > bytes) {
    napi_throw_range_error(env, nullptr, "tree buffer shorter than numNodes * 64");
    return nullptr;
  }
  uint32_t lr[32] = {0};
  uint32_t nl = 0;
  if (a.n > 3) {
    bool isarr = false;
    napi_is_array(env, a.v[3], &isarr);
    if (isarr) {
      napi_get_array_length(env, a.v[3], &nl);
      if (nl > 16) nl = 16;
      for (uint32_t k = 0; k < nl; k++) {
        napi_value pr, e0, e1;
        napi_get_element(env, a.v[3], k, &pr);
        napi_get_element(env, pr, 0, &e0);
        napi_get_element(env, pr, 1, &e1);
        double d0 = 0, d1 = 0;
        napi_get_value_double(env, e0, &d0);
        napi_get_value_double(env, e1, &d1);
        lr[2 * k] = (uint32_t)d0;
        lr[2 * k + 1] = (uint32_t)d1;
      }
    }
  }
  int rc = gs_upload_tree(s, static_cast<uint32_t*>(data), (uint64_t)nn, lr, nl);
  return rc ? throw_gs(env, s, rc, "gs_upload_tree") : nullptr;
}

#define DOWNLOAD(NAME, TYPE, ENUM, CALLEXPR)                         \
  napi_value NAME(napi_env env, napi_callback_info info) {           \
    Args a;                                                          \
    get_args(env, info, a);                                          \
    SIM_OR_RETURN(s, a.v[0]);                                        \
    size_t len = 0;                                                  \
    TYPE* p = typed<TYPE>(env, a.v[1], ENUM, &len);                  \
    if (!p) return nullptr;                                          \
    int rc = CALLEXPR;                                               \
    return rc ? throw_gs(env, s, rc, #NAME) : nullptr;               \
  }
DOWNLOAD(DownloadParticles, float, napi_float32_array, gs_download_particles(s, p, len / 8))
DOWNLOAD(ReadPositions, float, napi_float32_array, gs_read_positions(s, p, len / 4))
DOWNLOAD(DownloadPerm, uint32_t, napi_uint32_array, gs_download_perm(s, p, len))
DOWNLOAD(DownloadEnergy, float, napi_float32_array, gs_download_energy(s, p, len))
DOWNLOAD(DownloadAccel, float, napi_float32_array, gs_download_accel(s, p, nullptr, len / 4))

napi_value DownloadCodes(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t len = 0;
  uint32_t* p = typed<uint32_t>(env, a.v[1], napi_uint32_array, &len);
  if (!p) return nullptr;
  bool sorted = false;
  if (a.n > 2) napi_get_value_bool(env, a.v[2], &sorted);
  int rc = gs_download_codes(s, p, len, sorted ? 1 : 0);
  return rc ? throw_gs(env, s, rc, "gs_download_codes") : nullptr;
}

// treeInfo(h) -> {numNodes, numLevels, levelRanges, rootSize}
napi_value TreeInfo(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  uint64_t nn = 0;
  uint32_t lr[32] = {0}, nl = 0;
  double rs = 0;
  int rc = gs_tree_info(s, &nn, lr, &nl, &rs);
  if (rc) return throw_gs(env, s, rc, "gs_tree_info");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "numNodes", num(env, (double)nn));
  napi_set_named_property(env, o, "numLevels", num(env, nl));
  napi_set_named_property(env, o, "rootSize", num(env, rs));
  napi_set_named_property(env, o, "levelRanges", level_ranges_js(env, lr, nl));
  return o;
}

// downloadNodes(h) -> ArrayBuffer of numNodes * 64 bytes
napi_value DownloadNodes(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  uint64_t nn = 0;
  int rc = gs_tree_info(s, &nn, nullptr, nullptr, nullptr);
  if (rc) return throw_gs(env, s, rc, "gs_tree_info");
  void* data = nullptr;
  napi_value ab;
  CALL(env, napi_create_arraybuffer(env, nn * 64, &data, &ab));
  uint64_t got = 0;
  rc = gs_download_nodes(s, static_cast<uint32_t*>(data), nn, &got);
  if (rc) return throw_gs(env, s, rc, "gs_download_nodes");
  return ab;
}

napi_value Reduce(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  void* data = nullptr;
  napi_value ab, arr;
  CALL(env, napi_create_arraybuffer(env, GS_NRED * 8, &data, &ab));
  int rc = gs_reduce(s, static_cast<double*>(data));
  if (rc) return throw_gs(env, s, rc, "gs_reduce");
  CALL(env, napi_create_typedarray(env, napi_float64_array, GS_NRED, ab, 0, &arr));
  return arr;
}

napi_value CountInteractions(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  uint64_t t = 0;
  int rc = gs_count_interactions(s, &t, nullptr);
  if (rc) return throw_gs(env, s, rc, "gs_count_interactions");
  return num(env, (double)t);
}

// computeMortonCodes(h, pos, N, stride, codes) -> rootSize
napi_value ComputeMortonCodes(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t npos = 0, ncodes = 0;
  float* pos = typed<float>(env, a.v[1], napi_float32_array, &npos);
  if (!pos) return nullptr;
  double n = 0, stride = 0;
  napi_get_value_double(env, a.v[2], &n);
  napi_get_value_double(env, a.v[3], &stride);
  uint32_t* codes = typed<uint32_t>(env, a.v[4], napi_uint32_array, &ncodes);
  if (!codes) return nullptr;
  // the last particle reads pos[(N - 1) * stride + 0..2]
  if (!(n >= 1) || !(stride >= 3) || ncodes < (size_t)n ||
      npos < ((size_t)n - 1) * (size_t)stride + 3) {
    napi_throw_range_error(env, nullptr, "computeMortonCodes: pos or codes shorter than N");
    return nullptr;
  }
  double rs = 0;
  int rc = gs_compute_morton_codes(s, pos, (uint32_t)n, (uint32_t)stride, codes, &rs);
  if (rc) return throw_gs(env, s, rc, "gs_compute_morton_codes");
  return num(env, rs);
}

// radixSortIndices(h, codes, N, out)
napi_value RadixSortIndices(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t nc = 0, no = 0;
  uint32_t* codes = typed<uint32_t>(env, a.v[1], napi_uint32_array, &nc);
  if (!codes) return nullptr;
  double n = 0;
  napi_get_value_double(env, a.v[2], &n);
  uint32_t* out = typed<uint32_t>(env, a.v[3], napi_uint32_array, &no);
  if (!out) return nullptr;
  if (!(n >= 0) || nc < (size_t)n || no < (size_t)n) {
    napi_throw_range_error(env, nullptr, "radixSortIndices: codes or indices shorter than N");
    return nullptr;
  }
  int rc = gs_radix_sort_indices(s, codes, (uint32_t)n, out);
  return rc ? throw_gs(env, s, rc, "gs_radix_sort_indices") : nullptr;
}

// treeBuild(h, codes, pos3, masses, N, rootSize, outBuffer) -> {numNodes, levelRanges}
napi_value TreeBuild(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t nc = 0, np3 = 0, nm = 0;
  uint32_t* codes = typed<uint32_t>(env, a.v[1], napi_uint32_array, &nc);
  float* pos = typed<float>(env, a.v[2], napi_float32_array, &np3);
  float* masses = typed<float>(env, a.v[3], napi_float32_array, &nm);
  if (!codes || !pos || !masses) return nullptr;
  double n = 0, rs = 0;
  napi_get_value_double(env, a.v[4], &n);
  napi_get_value_double(env, a.v[5], &rs);
  if (!(n >= 0) || nc < (size_t)n || np3 < 3 * (size_t)n || nm < (size_t)n) {
    napi_throw_range_error(env, nullptr, "treeBuild: codes, pos (3N) or masses shorter than N");
    return nullptr;
  }
  void* out = nullptr;
  size_t bytes = 0;
  if (napi_get_arraybuffer_info(env, a.v[6], &out, &bytes) != napi_ok) {
    napi_throw_type_error(env, nullptr, "treeBuild needs an output ArrayBuffer");
    return nullptr;
  }
  uint64_t nn = 0;
  uint32_t lr[32] = {0}, nl = 0;
  int rc = gs_tree_build(s, codes, pos, masses, (uint32_t)n, rs, static_cast<uint32_t*>(out),
                         bytes / 64, &nn, lr, &nl);
  if (rc) return throw_gs(env, s, rc, "gs_tree_build");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "numNodes", num(env, (double)nn));
  napi_set_named_property(env, o, "levelRanges", level_ranges_js(env, lr, nl));
  return o;
}

// ---- slot pools (gs_create_pool / gs_slot_*) --------------------------------
uint32_t u32_arg(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return (uint32_t)d;
}

// capacities / buckets: Uint32Array or Array of numbers
bool u32_list(napi_env env, napi_value v, std::string& err, uint32_t* out, uint32_t cap,
              uint32_t* n) {
  bool isarr = false;
  napi_is_array(env, v, &isarr);
  if (isarr) {
    uint32_t len = 0;
    napi_get_array_length(env, v, &len);
    if (len > cap) {
      err = "too many slots";
      return false;
    }
    for (uint32_t k = 0; k < len; k++) {
      napi_value e;
      napi_get_element(env, v, k, &e);
      out[k] = u32_arg(env, e);
    }
    *n = len;
    return true;
  }
  napi_typedarray_type t;
  size_t length = 0;
  void* data = nullptr;
  napi_value ab;
  size_t off = 0;
  if (napi_get_typedarray_info(env, v, &t, &length, &data, &ab, &off) != napi_ok ||
      t != napi_uint32_array || length > cap) {
    err = "expected an Array or Uint32Array of at most 256 values";
    return false;
  }
  memcpy(out, data, 4 * length);
  *n = (uint32_t)length;
  return true;
}

// createPool({capacities, buckets, exactForce, device, walkWg, radixBits, graph}) -> handle
napi_value CreatePool(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) {
    napi_throw_type_error(env, nullptr, "createPool(options)");
    return nullptr;
  }
  napi_value o = a.v[0];
  uint32_t caps[256] = {0}, buckets[256] = {0}, nc = 0, nb = 0;
  std::string err;
  bool has = false;
  napi_has_named_property(env, o, "capacities", &has);
  if (!has) {
    napi_throw_type_error(env, nullptr, "createPool: capacities required");
    return nullptr;
  }
  napi_value v;
  napi_get_named_property(env, o, "capacities", &v);
  if (!u32_list(env, v, err, caps, 256, &nc)) {
    napi_throw_type_error(env, nullptr, err.c_str());
    return nullptr;
  }
  napi_has_named_property(env, o, "buckets", &has);
  if (has) {
    napi_get_named_property(env, o, "buckets", &v);
    if (!u32_list(env, v, err, buckets, 256, &nb) || nb != nc) {
      napi_throw_type_error(env, nullptr, "createPool: buckets must match capacities");
      return nullptr;
    }
  }
  gs_pool_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.nslots = nc;
  cfg.capacity = caps;
  cfg.bucket = nb ? buckets : nullptr;
  cfg.flags = (num_prop(env, o, "exactForce", 0) != 0 ? GS_FLAG_EXACT_FORCE : 0u) |
              (num_prop(env, o, "graph", 1) == 0 ? GS_FLAG_NO_GRAPH : 0u);
  cfg.device = (int32_t)num_prop(env, o, "device", -1);
  cfg.walk_wg = (uint32_t)num_prop(env, o, "walkWg", 0);
  cfg.radix_bits = (uint32_t)num_prop(env, o, "radixBits", 0);
  gs_sim* s = nullptr;
  int rc = gs_create_pool(&cfg, &s);
  if (rc != GS_OK) return throw_gs(env, nullptr, rc, "gs_create_pool");
  Holder* h = new Holder{s};
  napi_value ext;
  CALL(env, napi_create_external(env, h, finalize_holder, nullptr, &ext));
  return ext;
}

// slotUpload(h, slot, Float32Array records)
napi_value SlotUpload(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t len = 0;
  float* p = typed<float>(env, a.v[2], napi_float32_array, &len);
  if (!p) return nullptr;
  int rc = gs_slot_upload(s, u32_arg(env, a.v[1]), p, len / 8);
  return rc ? throw_gs(env, s, rc, "gs_slot_upload") : nullptr;
}

napi_value SlotClear(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  int rc = gs_slot_clear(s, u32_arg(env, a.v[1]));
  return rc ? throw_gs(env, s, rc, "gs_slot_clear") : nullptr;
}

napi_value SlotCount(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  uint32_t n = 0;
  int rc = gs_slot_count(s, u32_arg(env, a.v[1]), &n);
  return rc ? throw_gs(env, s, rc, "gs_slot_count") : num(env, n);
}

// slotSetParams(h, slot, updateParams dictionary)
napi_value SlotSetParams(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  gs_params p;
  parse_params(env, a.v[2], p);
  int rc = gs_slot_set_params(s, u32_arg(env, a.v[1]), &p);
  return rc ? throw_gs(env, s, rc, "gs_slot_set_params") : nullptr;
}

// slotSetTree(h, slot, bucket, maxNodes)
napi_value SlotSetTree(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  double mn = 0;
  if (a.n > 3) napi_get_value_double(env, a.v[3], &mn);
  int rc = gs_slot_set_tree(s, u32_arg(env, a.v[1]), u32_arg(env, a.v[2]), (uint64_t)mn);
  return rc ? throw_gs(env, s, rc, "gs_slot_set_tree") : nullptr;
}

// slotDownloadParticles(h, slot, Float32Array(8 n))
napi_value SlotDownloadParticles(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t len = 0;
  float* p = typed<float>(env, a.v[2], napi_float32_array, &len);
  if (!p) return nullptr;
  int rc = gs_slot_download_particles(s, u32_arg(env, a.v[1]), p, len / 8);
  return rc ? throw_gs(env, s, rc, "gs_slot_download_particles") : nullptr;
}

// slotDownloadEnergy(h, slot, Float32Array(n))
napi_value SlotDownloadEnergy(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t len = 0;
  float* p = typed<float>(env, a.v[2], napi_float32_array, &len);
  if (!p) return nullptr;
  int rc = gs_slot_download_energy(s, u32_arg(env, a.v[1]), p, len);
  return rc ? throw_gs(env, s, rc, "gs_slot_download_energy") : nullptr;
}

// slotReduce(h, slot) -> Float64Array(GS_NRED)
napi_value SlotReduce(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  void* data = nullptr;
  napi_value ab, arr;
  CALL(env, napi_create_arraybuffer(env, GS_NRED * 8, &data, &ab));
  int rc = gs_slot_reduce(s, u32_arg(env, a.v[1]), static_cast<double*>(data));
  if (rc) return throw_gs(env, s, rc, "gs_slot_reduce");
  CALL(env, napi_create_typedarray(env, napi_float64_array, GS_NRED, ab, 0, &arr));
  return arr;
}

// poolInfo(h) -> {nslots, deviceBytes}
napi_value PoolInfo(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  uint32_t ns = 0;
  uint64_t bytes = 0;
  int rc = gs_pool_info(s, &ns, &bytes);
  if (rc) return throw_gs(env, s, rc, "gs_pool_info");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "nslots", num(env, ns));
  napi_set_named_property(env, o, "deviceBytes", num(env, (double)bytes));
  return o;
}

// createArena({budgetBytes, slots, exactForce, device, walkWg, radixBits, graph}) -> handle
napi_value CreateArena(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) {
    napi_throw_type_error(env, nullptr, "createArena(options)");
    return nullptr;
  }
  napi_value o = a.v[0];
  gs_arena_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.budget_bytes = (uint64_t)num_prop(env, o, "budgetBytes", 0);
  cfg.nslots = (uint32_t)num_prop(env, o, "slots", 0);
  cfg.flags = (num_prop(env, o, "exactForce", 0) != 0 ? GS_FLAG_EXACT_FORCE : 0u) |
              (num_prop(env, o, "graph", 1) == 0 ? GS_FLAG_NO_GRAPH : 0u);
  cfg.device = (int32_t)num_prop(env, o, "device", -1);
  cfg.walk_wg = (uint32_t)num_prop(env, o, "walkWg", 0);
  cfg.radix_bits = (uint32_t)num_prop(env, o, "radixBits", 0);
  gs_sim* s = nullptr;
  int rc = gs_create_arena(&cfg, &s);
  if (rc != GS_OK) return throw_gs(env, nullptr, rc, "gs_create_arena");
  Holder* h = new Holder{s};
  napi_value ext;
  CALL(env, napi_create_external(env, h, finalize_holder, nullptr, &ext));
  return ext;
}

// arenaInfo(h) -> {deviceBytes, usedBytes, largestFit}
napi_value ArenaInfo(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  uint64_t dev = 0, used = 0, fit = 0;
  int rc = gs_arena_info(s, &dev, &used, &fit);
  if (rc) return throw_gs(env, s, rc, "gs_arena_info");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "deviceBytes", num(env, (double)dev));
  napi_set_named_property(env, o, "usedBytes", num(env, (double)used));
  napi_set_named_property(env, o, "largestFit", num(env, (double)fit));
  return o;
}

// uploadStellar(h, Float32Array(4 n)) / slotUploadStellar(h, slot, Float32Array)
napi_value UploadStellar(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  const bool slotted = a.n > 2;
  size_t len = 0;
  float* p = typed<float>(env, a.v[slotted ? 2 : 1], napi_float32_array, &len);
  if (!p) return nullptr;
  int rc = slotted ? gs_slot_upload_stellar(s, u32_arg(env, a.v[1]), p, len / 4)
                   : gs_upload_stellar(s, p, len / 4);
  return rc ? throw_gs(env, s, rc, "gs_upload_stellar") : nullptr;
}

// downloadStellar(h, Float32Array(4 n)) / slotDownloadStellar(h, slot, Float32Array)
napi_value DownloadStellar(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  const bool slotted = a.n > 2;
  size_t len = 0;
  float* p = typed<float>(env, a.v[slotted ? 2 : 1], napi_float32_array, &len);
  if (!p) return nullptr;
  int rc = slotted ? gs_slot_download_stellar(s, u32_arg(env, a.v[1]), p, len / 4)
                   : gs_download_stellar(s, p, len / 4);
  return rc ? throw_gs(env, s, rc, "gs_download_stellar") : nullptr;
}

SIMPLE(StarFormation, gs_star_formation)

// locate(h, oldIndex) -> {index, pos: Float32Array(4)}
napi_value Locate(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  uint32_t ni = 0;
  void* data = nullptr;
  napi_value ab, pos, o;
  CALL(env, napi_create_arraybuffer(env, 16, &data, &ab));
  int rc = gs_locate(s, u32_arg(env, a.v[1]), &ni, static_cast<float*>(data));
  if (rc) return throw_gs(env, s, rc, "gs_locate");
  CALL(env, napi_create_typedarray(env, napi_float32_array, 4, ab, 0, &pos));
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "index", num(env, ni));
  napi_set_named_property(env, o, "pos", pos);
  return o;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_string_utf8(env, gs_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

napi_value DeviceCount(napi_env env, napi_callback_info) {
  int n = 0;
  gs_device_count(&n);
  return num(env, n);
}

// ---- readPositionsAsync(h, Float32Array) -> Promise -------------------------
struct ReadWork {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref keep;
  gs_sim* sim;
  float* dst;
  size_t n;
  int rc;
};

void read_execute(napi_env, void* data) {
  ReadWork* w = static_cast<ReadWork*>(data);
  w->rc = gs_read_positions(w->sim, w->dst, w->n);
}

void read_complete(napi_env env, napi_status, void* data) {
  ReadWork* w = static_cast<ReadWork*>(data);
  if (w->rc == GS_OK) {
    napi_value undef;
    napi_get_undefined(env, &undef);
    napi_resolve_deferred(env, w->deferred, undef);
  } else {
    napi_value msg, err;
    std::string m = std::string("gs_read_positions failed: ") + gs_last_error(w->sim);
    napi_create_string_utf8(env, m.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  }
  napi_delete_reference(env, w->keep);
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value ReadPositionsAsync(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  SIM_OR_RETURN(s, a.v[0]);
  size_t len = 0;
  float* p = typed<float>(env, a.v[1], napi_float32_array, &len);
  if (!p) return nullptr;
  ReadWork* w = new ReadWork{};
  w->sim = s;
  w->dst = p;
  w->n = len / 4;
  napi_value promise, name;
  CALL(env, napi_create_promise(env, &w->deferred, &promise));
  CALL(env, napi_create_reference(env, a.v[1], 1, &w->keep));
  napi_create_string_utf8(env, "galaxy_hip.readPositions", NAPI_AUTO_LENGTH, &name);
  CALL(env, napi_create_async_work(env, nullptr, name, read_execute, read_complete, w, &w->work));
  CALL(env, napi_queue_async_work(env, w->work));
  return promise;
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback cb;
  } fns[] = {
      {"create", Create}, {"destroy", Destroy}, {"upload", Upload}, {"setParams", SetParams},
      {"updateCentralMass", UpdateCentralMass}, {"encode", Encode}, {"sort", Sort},
      {"refresh", Refresh}, {"force", Force}, {"step", Step}, {"walkStep", WalkStep},
      {"sync", Sync},
      {"applyPermutation", ApplyPermutation}, {"uploadTree", UploadTree},
      {"downloadParticles", DownloadParticles}, {"readPositions", ReadPositions},
      {"readPositionsAsync", ReadPositionsAsync}, {"downloadPerm", DownloadPerm},
      {"downloadCodes", DownloadCodes}, {"downloadEnergy", DownloadEnergy},
      {"downloadAccel", DownloadAccel}, {"downloadNodes", DownloadNodes},
      {"treeInfo", TreeInfo}, {"reduce", Reduce}, {"countInteractions", CountInteractions},
      {"computeMortonCodes", ComputeMortonCodes}, {"radixSortIndices", RadixSortIndices},
      {"treeBuild", TreeBuild}, {"version", Version}, {"deviceCount", DeviceCount},
      {"createPool", CreatePool}, {"poolInfo", PoolInfo}, {"slotUpload", SlotUpload},
      {"createArena", CreateArena}, {"arenaInfo", ArenaInfo},
      {"slotClear", SlotClear}, {"slotCount", SlotCount}, {"slotSetParams", SlotSetParams},
      {"slotSetTree", SlotSetTree}, {"slotDownloadParticles", SlotDownloadParticles},
      {"slotDownloadEnergy", SlotDownloadEnergy}, {"slotReduce", SlotReduce},
      {"uploadStellar", UploadStellar}, {"slotUploadStellar", UploadStellar},
      {"downloadStellar", DownloadStellar}, {"slotDownloadStellar", DownloadStellar},
      {"starFormation", StarFormation}, {"locate", Locate},
  };
  for (auto& f : fns) {
    napi_value fn;
    CALL(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    CALL(env, napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
