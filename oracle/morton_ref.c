/*
 * Oracle (test infrastructure only): C restatement of
 * src/physics/mortonTree.js.  JS Numbers are IEEE doubles; this file is
 * built with -ffp-contract=off so every `a * b + c` rounds twice, as V8 does.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* mortonTree.js:29-36 */
uint32_t or_part1by2(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x30000ffu;
  v = (v | (v << 8)) & 0x300f00fu;
  v = (v | (v << 4)) & 0x30c30c3u;
  v = (v | (v << 2)) & 0x9249249u;
  return v;
}

/* ECMAScript ToInt32, the semantics of `expr | 0` (mortonTree.js:56-58). */
int32_t or_js_toint32(double d) {
  if (!isfinite(d)) return 0;
  double t = trunc(d);
  double m = fmod(t, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return (int32_t)(uint32_t)m;
}

static double js_max3(double a, double b, double c) {
  if (isnan(a) || isnan(b) || isnan(c)) return NAN;
  double m = a > b ? a : b;
  return m > c ? m : c;
}

static uint32_t quant(double v) {
  int32_t q = or_js_toint32(v);
  if (q > 1023) q = 1023; /* Math.min(q, 1023); negatives pass through */
  return (uint32_t)q;     /* part1by2 masks to 10 bits */
}

/* mortonTree.js:43-62 */
double or_compute_morton_codes(const float* pos, uint32_t n, uint32_t stride,
                               uint32_t* codes) {
  double minX = INFINITY, minY = INFINITY, minZ = INFINITY;
  double maxX = -INFINITY, maxY = -INFINITY, maxZ = -INFINITY;
  for (uint32_t i = 0; i < n; i++) {
    double x = pos[(size_t)i * stride], y = pos[(size_t)i * stride + 1],
           z = pos[(size_t)i * stride + 2];
    if (!isfinite(x + y + z)) continue;
    if (x < minX) minX = x;
    if (x > maxX) maxX = x;
    if (y < minY) minY = y;
    if (y > maxY) maxY = y;
    if (z < minZ) minZ = z;
    if (z > maxZ) maxZ = z;
  }
  double size = js_max3(maxX - minX, maxY - minY, maxZ - minZ) * 1.0001 + 1e-9;
  double s = 1024 / size;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t x = quant(((double)pos[(size_t)i * stride] - minX) * s);
    uint32_t y = quant(((double)pos[(size_t)i * stride + 1] - minY) * s);
    uint32_t z = quant(((double)pos[(size_t)i * stride + 2] - minZ) * s);
    codes[i] = or_part1by2(x) | (or_part1by2(y) << 1) | (or_part1by2(z) << 2);
  }
  return size;
}

/* mortonTree.js:68-84: 3 stable passes of 10-bit digits. */
void or_radix_sort_indices(const uint32_t* codes, uint32_t* perm, uint32_t n) {
  uint32_t* scratch = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t counts[1024];
  uint32_t *src = perm, *dst = scratch;
  for (uint32_t i = 0; i < n; i++) src[i] = i;
  for (int pass = 0; pass < 3; pass++) {
    int shift = pass * 10;
    memset(counts, 0, sizeof counts);
    for (uint32_t i = 0; i < n; i++) counts[(codes[src[i]] >> shift) & 1023]++;
    uint32_t sum = 0;
    for (int b = 0; b < 1024; b++) {
      uint32_t c = counts[b];
      counts[b] = sum;
      sum += c;
    }
    for (uint32_t i = 0; i < n; i++) {
      uint32_t v = src[i];
      dst[counts[(codes[v] >> shift) & 1023]++] = v;
    }
    uint32_t* t = src;
    src = dst;
    dst = t;
  }
  /* odd pass count: the result sits in scratch */
  if (src != perm) memcpy(perm, src, sizeof(uint32_t) * n);
  free(scratch);
}

typedef struct {
  uint32_t lo, hi;
  int64_t parent_slot; /* word index of the parent's child slot, or -1 */
} item_t;

static inline float* fw(uint32_t* w, uint64_t i) { return (float*)(w + i); }

/* mortonTree.js:107-182, with BUCKET (:25) and maxNodes (:88) as arguments. */
int64_t or_tree_build(const uint32_t* codes, const float* pos,
                      const float* masses, uint32_t n, double root_size,
                      uint32_t bucket, uint64_t max_nodes, uint32_t* w,
                      uint64_t cap, uint32_t* level_ranges,
                      uint32_t* num_levels) {
  const int MAX_LEVEL = 10;
  uint64_t num_nodes = 0;
  uint32_t nlev = 0;
  item_t* queue = (item_t*)malloc(sizeof(item_t));
  size_t qlen = 1;
  queue[0].lo = 0;
  queue[0].hi = n;
  queue[0].parent_slot = -1;
  int depth = 0;
  double size = root_size;
  while (qlen) {
    item_t* next = (item_t*)malloc(sizeof(item_t) * qlen * 8);
    size_t nlen = 0;
    uint64_t level_start = num_nodes;
    for (size_t q = 0; q < qlen; q++) {
      item_t it = queue[q];
      uint64_t idx = num_nodes++;
      if (idx >= cap) {
        free(next);
        free(queue);
        return -1;
      }
      if (it.parent_slot >= 0) w[it.parent_slot] = (uint32_t)idx;
      uint64_t b = idx * 16;
      uint32_t count = it.hi - it.lo;
      *fw(w, b + 4) = (float)size;
      int leaf = count <= bucket || depth >= MAX_LEVEL || num_nodes + 8 > max_nodes;
      if (leaf) {
        double m = 0, cx = 0, cy = 0, cz = 0;
        for (uint32_t p = it.lo; p < it.hi; p++) {
          double pm = masses[p];
          m += pm;
          cx += (double)pos[(size_t)p * 3] * pm;
          cy += (double)pos[(size_t)p * 3 + 1] * pm;
          cz += (double)pos[(size_t)p * 3 + 2] * pm;
        }
        double inv = m > 0 ? 1 / m : 0;
        *fw(w, b) = (float)(cx * inv);
        *fw(w, b + 1) = (float)(cy * inv);
        *fw(w, b + 2) = (float)(cz * inv);
        *fw(w, b + 3) = (float)m;
        w[b + 5] = 1;
        w[b + 6] = it.lo;
        w[b + 7] = count;
        for (int c = 0; c < 8; c++) w[b + 8 + c] = (uint32_t)-1;
      } else {
        w[b + 5] = 0;
        w[b + 6] = 0;
        w[b + 7] = 0;
        int shift = 27 - 3 * depth;
        uint32_t start = it.lo;
        for (uint32_t k = 0; k < 8; k++) {
          uint32_t end = start;
          while (end < it.hi && ((codes[end] >> shift) & 7) == k) end++;
          if (end > start) {
            next[nlen].lo = start;
            next[nlen].hi = end;
            next[nlen].parent_slot = (int64_t)(b + 8 + k);
            nlen++;
          } else {
            w[b + 8 + k] = (uint32_t)-1;
          }
          start = end;
        }
      }
    }
    if (nlev < 16) {
      level_ranges[2 * nlev] = (uint32_t)level_start;
      level_ranges[2 * nlev + 1] = (uint32_t)(num_nodes - level_start);
    }
    nlev++;
    free(queue);
    queue = next;
    qlen = nlen;
    depth++;
    size *= 0.5;
  }
  free(queue);
  /* bottom-up COM, reverse index sweep (mortonTree.js:166-181) */
  for (int64_t nd = (int64_t)num_nodes - 1; nd >= 0; nd--) {
    uint64_t b = (uint64_t)nd * 16;
    if (w[b + 5] == 1) continue;
    double m = 0, cx = 0, cy = 0, cz = 0;
    for (int c = 0; c < 8; c++) {
      int32_t ch = (int32_t)w[b + 8 + c];
      if (ch >= 0) {
        uint64_t cb = (uint64_t)ch * 16;
        double cm = *fw(w, cb + 3);
        m += cm;
        cx += (double)*fw(w, cb) * cm;
        cy += (double)*fw(w, cb + 1) * cm;
        cz += (double)*fw(w, cb + 2) * cm;
      }
    }
    double inv = m > 0 ? 1 / m : 0;
    *fw(w, b) = (float)(cx * inv);
    *fw(w, b + 1) = (float)(cy * inv);
    *fw(w, b + 2) = (float)(cz * inv);
    *fw(w, b + 3) = (float)m;
  }
  *num_levels = nlev;
  return (int64_t)num_nodes;
}

/* mortonTree.js:192-224 */
int64_t or_sort_and_build(const float* particles, const float* stellar,
                          uint32_t n, int64_t track_idx2, uint32_t bucket,
                          uint64_t max_nodes, float* sorted,
                          float* sorted_stellar, uint32_t* perm,
                          uint32_t* sorted_codes, uint32_t* nodes,
                          uint64_t cap, uint32_t* level_ranges,
                          uint32_t* num_levels, double* root_size,
                          int64_t* central_idx, int64_t* central_idx2) {
  uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  float* spos = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
  float* smass = (float*)malloc(sizeof(float) * (n ? n : 1));
  double rs = or_compute_morton_codes(particles, n, 8, codes);
  or_radix_sort_indices(codes, perm, n);
  int64_t c1 = 0, c2 = -1;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t s = perm[i];
    if (s == 0) c1 = i;
    if ((int64_t)s == track_idx2) c2 = i;
    memcpy(sorted + (size_t)i * 8, particles + (size_t)s * 8, 32);
    if (stellar && sorted_stellar)
      memcpy(sorted_stellar + (size_t)i * 4, stellar + (size_t)s * 4, 16);
    sorted_codes[i] = codes[s];
    spos[(size_t)i * 3] = particles[(size_t)s * 8];
    spos[(size_t)i * 3 + 1] = particles[(size_t)s * 8 + 1];
    spos[(size_t)i * 3 + 2] = particles[(size_t)s * 8 + 2];
    smass[i] = particles[(size_t)s * 8 + 3];
  }
  int64_t nn = or_tree_build(sorted_codes, spos, smass, n, rs, bucket,
                             max_nodes, nodes, cap, level_ranges, num_levels);
  *root_size = rs;
  *central_idx = c1;
  *central_idx2 = c2;
  free(codes);
  free(spos);
  free(smass);
  return nn;
}
