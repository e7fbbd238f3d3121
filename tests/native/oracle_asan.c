/* The oracle's C code (oracle/morton_ref.c, oracle/walk_ref.c) under
 * AddressSanitizer and UBSan: sortAndBuild, both refreshes, the walk on all
 * particles and on a subset, the permutation gather, the FP64 direct sum and
 * star formation, over inputs that reach the edge paths: N = 1, 16, 17,
 * coincident particles, NaN / ±Inf positions, 1e±30 scales, zero masses,
 * bucket 1 and node-pool caps.  Built and run by tests/test_cpu_abi.py; any
 * sanitizer report fails the run, and each case checks that every particle
 * lands in exactly one leaf. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/oracle.h"

static uint64_t rs = 0x9E3779B97F4A7C15ull;
static double urand(void) {
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return (double)(rs >> 11) * (1.0 / 9007199254740992.0);
}

static int bad = 0;

static void run_case(const char* name, float* part, uint32_t n, uint32_t bucket,
                     uint64_t max_nodes) {
  const uint64_t cap = 2ull * n + 4096 > max_nodes ? 2ull * n + 4096 : max_nodes;
  float* sorted = malloc(sizeof(float) * 8 * n);
  float* sst = malloc(sizeof(float) * 4 * n);
  float* stel = calloc(4 * (size_t)n, sizeof(float));
  uint32_t* perm = malloc(sizeof(uint32_t) * n);
  uint32_t* codes = malloc(sizeof(uint32_t) * n);
  uint32_t* nodes = malloc(sizeof(uint32_t) * 16 * cap);
  uint32_t lr[32] = {0}, nl = 0;
  double rsz = 0;
  int64_t c1 = 0, c2 = 0;
  for (uint32_t i = 0; i < n; i++) stel[4 * i] = (float)(i % 13);
  const int64_t nn = or_sort_and_build(part, stel, n, n > 1 ? 1 : -1, bucket, max_nodes, sorted,
                                       sst, perm, codes, nodes, cap, lr, &nl, &rsz, &c1, &c2);
  if (nn <= 0) {
    printf("%s: sortAndBuild returned %lld\n", name, (long long)nn);
    bad++;
    goto out;
  }
  {
    /* every particle in exactly one leaf */
    uint32_t* hits = calloc(n, sizeof(uint32_t));
    for (int64_t v = 0; v < nn; v++) {
      const uint32_t* w = nodes + 16 * v;
      if (w[5] == 1u)
        for (uint32_t j = w[6]; j < w[6] + w[7] && j < n; j++) hits[j]++;
    }
    for (uint32_t i = 0; i < n; i++)
      if (hits[i] != 1) {
        printf("%s: particle %u in %u leaves\n", name, i, hits[i]);
        bad++;
        break;
      }
    free(hits);
    or_refresh_f64(sorted, nodes, (uint64_t)nn);
    or_refresh_f32(sorted, nodes, lr, nl);
    or_params p;
    const double cc1[3] = {0.1, 0.0, 0.0}, cc2[3] = {-0.2, 0.1, 0.0};
    or_pack_params(&p, 0.005, 1.0, 0.02, 0.6, 2.0, 1.5, 0.25, n, 2, cc1, cc2);
    p.sf_density = 0.5f;
    p.sf_eff = 1.0f;
    p.sf_dt = 0.01f;
    p.time = 3.0f;
    float* pout = malloc(sizeof(float) * 8 * n);
    float* energy = malloc(sizeof(float) * n);
    float* acc4 = malloc(sizeof(float) * 4 * n);
    float* dens = malloc(sizeof(float) * n);
    uint32_t* inter = malloc(sizeof(uint32_t) * n);
    or_walk(sorted, n, nodes, &p, pout, energy, acc4, dens, inter, 4);
    uint32_t idx[3] = {0, n / 2, n - 1};
    float sub[12];
    uint32_t sint[3];
    or_walk_subset(sorted, n, nodes, &p, idx, 3, sub, sint, 2);
    for (int k = 0; k < 3; k++)
      if (sint[k] != inter[idx[k]]) {
        printf("%s: subset walk differs at %u\n", name, idx[k]);
        bad++;
      }
    float* back = malloc(sizeof(float) * 8 * n);
    or_apply_permutation(sorted, perm, n, 8, back);
    double ds[9];
    or_direct_sum_f64(sorted, n, 1.0, 4e-4, idx, 3, ds, 3);
    or_star_formation(pout, sst, n, &p);
    free(pout);
    free(energy);
    free(acc4);
    free(dens);
    free(inter);
    free(back);
  }
out:
  free(sorted);
  free(sst);
  free(stel);
  free(perm);
  free(codes);
  free(nodes);
}

static float* plummer(uint32_t n, double scale) {
  float* p = calloc(8 * (size_t)n, sizeof(float));
  for (uint32_t i = 0; i < n; i++) {
    const double r = scale / sqrt(pow(urand() * 0.99 + 0.005, -2.0 / 3.0) - 1.0);
    const double z = 2 * urand() - 1, ph = 6.283185307179586 * urand();
    const double s = sqrt(1 - z * z);
    p[8 * i] = (float)(r * s * cos(ph));
    p[8 * i + 1] = (float)(r * s * sin(ph));
    p[8 * i + 2] = (float)(r * z);
    p[8 * i + 3] = 1.0f / (float)n;
    p[8 * i + 4] = (float)(0.1 * (urand() - 0.5));
    p[8 * i + 7] = (i % 3 == 0) ? -1.0f : 0.0f; /* a third is gas */
  }
  return p;
}

int main(void) {
  const uint32_t sizes[] = {1, 16, 17, 1000, 20000};
  for (int k = 0; k < 5; k++) {
    float* p = plummer(sizes[k], 1.0);
    char name[64];
    snprintf(name, sizeof name, "plummer n=%u", sizes[k]);
    run_case(name, p, sizes[k], 16, 0);
    run_case("  bucket 1", p, sizes[k], 1, 0);
    free(p);
  }
  float* p = plummer(3000, 1.0);
  run_case("pool cap 60", p, 3000, 16, 60);
  run_case("pool cap 200", p, 3000, 8, 200);
  free(p);
  p = plummer(2000, 1e30);
  run_case("scale 1e30", p, 2000, 16, 0);
  free(p);
  p = plummer(2000, 1e-30);
  run_case("scale 1e-30", p, 2000, 16, 0);
  free(p);
  p = plummer(500, 1.0);
  for (uint32_t i = 0; i < 200; i++) memcpy(p + 8 * i, p, 3 * sizeof(float)); /* coincident */
  for (uint32_t i = 200; i < 260; i++) p[8 * i + 3] = 0.0f;                    /* massless */
  run_case("coincident + massless", p, 500, 16, 0);
  p[8 * 300] = NAN;
  p[8 * 301 + 1] = INFINITY;
  p[8 * 302 + 2] = -INFINITY;
  p[8 * 303] = 3e9f; /* past ToInt32 */
  run_case("NaN / Inf / wrap", p, 500, 16, 0);
  free(p);
  printf("bad=%d\n", bad);
  return bad ? 1 : 0;
}
