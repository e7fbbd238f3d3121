/*
 * CPU oracle for the Barnes-Hut step — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker or the timed CPU baseline; the HIP
 * library never links or calls it.
 *
 * Two parts:
 *   morton_ref.c  restatement of src/physics/mortonTree.js (JS Number = IEEE
 *                 double, built with -ffp-contract=off): Morton encode, stable
 *                 LSD radix sort, BFS bucketed octree build with FP64 COM,
 *                 sortAndBuild.  Pinned by golden fixtures written from the
 *                 real mortonTree.js plus brute force and invariants.
 *   walk_ref.c    FP32 transliteration of src/physics/shaders/nbody_bh.wgsl
 *                 `main` (walk + halo + kick-drift), `refresh_level` and
 *                 `apply_permutation`, with the implementation-defined parts
 *                 fixed: dot = (x*x + y*y) + z*z, no FMA, IEEE sqrtf and
 *                 division.  Pinned by closed forms (see tests).
 *
 * Node words (16 x 32-bit, mortonTree.js:12-21):
 *   f32[0..2] com, f32[3] mass, f32[4] size, i32[5] isLeaf, i32[6] first,
 *   i32[7] count, i32[8..15] children (-1 = empty).
 * Particle record (nbody_bh.wgsl:10-17): f32 [x,y,z,m, vx,vy,vz,pad].
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mortonTree.js ------------------------------------------------------ */
uint32_t or_part1by2(uint32_t v);
int32_t or_js_toint32(double d);
double or_compute_morton_codes(const float* pos, uint32_t n, uint32_t stride,
                               uint32_t* codes);
/* Writes the sorted index array (gather permutation) into perm. */
void or_radix_sort_indices(const uint32_t* codes, uint32_t* perm, uint32_t n);
/* BFS build.  nodes must hold cap*16 words; returns numNodes or -1 when cap
 * is too small.  level_ranges holds up to 16 [start,count] pairs. */
int64_t or_tree_build(const uint32_t* codes, const float* pos3,
                      const float* masses, uint32_t n, double root_size,
                      uint32_t bucket, uint64_t max_nodes, uint32_t* nodes,
                      uint64_t cap, uint32_t* level_ranges,
                      uint32_t* num_levels);
/* sortAndBuild (mortonTree.js:192-224) on stride-8 particles (+ stride-4
 * stellar, may be NULL).  Outputs sorted particles, sorted stellar, perm,
 * sorted codes, tree, central indices. */
int64_t or_sort_and_build(const float* particles, const float* stellar,
                          uint32_t n, int64_t track_idx2, uint32_t bucket,
                          uint64_t max_nodes, float* sorted,
                          float* sorted_stellar, uint32_t* perm,
                          uint32_t* sorted_codes, uint32_t* nodes,
                          uint64_t cap, uint32_t* level_ranges,
                          uint32_t* num_levels, double* root_size,
                          int64_t* central_idx, int64_t* central_idx2);

/* ---- nbody_bh.wgsl ------------------------------------------------------ */
typedef struct or_params {      /* Params (nbody_bh.wgsl:19-42), f32 words */
  float dt, G, eps2;
  uint32_t n;
  float theta, halo_mass, halo_a, gas_drag, time, aging_rate, sf_density,
      sf_eff, sf_dt;
  uint32_t halo_count;
  float pad0, pad1;
  float halo_c1[4], halo_c2[4];
  /* exp(-gasDrag*dt) evaluated once in f32 (the WGSL evaluates the same
   * uniform expression per thread). */
  float drag_damp;
} or_params;

/* Pack Params exactly as NBodyGPU.updateParams (NBodyGPU.js:253-275). */
void or_pack_params(or_params* p, double dt, double G, double epsilon,
                    double theta, double halo_mass, double halo_a,
                    double gas_drag, uint32_t n, uint32_t halo_count,
                    const double* c1, const double* c2);

/* `main`: walk + halos + kick-drift.  Any of pout/energy/acc4/dens/inter
 * may be NULL.  acc4 = (ax, ay, az, pot) before the halo term; inter counts
 * monopole acceptances + leaf pairs per particle. */
void or_walk(const float* pin, uint32_t n, const uint32_t* nodes,
             const or_params* p, float* pout, float* energy, float* acc4,
             float* dens, uint32_t* inter, int nthreads);
/* Same walk over a subset of particle indices (for sampled baselines). */
void or_walk_subset(const float* pin, uint32_t n, const uint32_t* nodes,
                    const or_params* p, const uint32_t* idx, uint32_t m,
                    float* acc4, uint32_t* inter, int nthreads);
/* refresh_level in f32 (nbody_bh.wgsl:288-313), all levels deepest first. */
void or_refresh_f32(const float* pin, uint32_t* nodes,
                    const uint32_t* level_ranges, uint32_t num_levels);
/* COM refresh in FP64 with the build's summation order (mortonTree.js:129-181):
 * the refresh the HIP path implements. */
void or_refresh_f64(const float* pin, uint32_t* nodes, uint64_t num_nodes);
void or_apply_permutation(const float* in, const uint32_t* perm, uint32_t n,
                          uint32_t stride, float* out);
/* FP64 direct sum (Plummer softened) on selected particles: acc3 doubles. */
void or_direct_sum_f64(const float* pin, uint32_t n, double G, double eps2,
                       const uint32_t* idx, uint32_t m, double* acc3,
                       int nthreads);

/* star_formation (nbody_bh.wgsl:197-268) */
uint32_t or_pcg(uint32_t v0);
float or_rand01(uint32_t i, uint32_t salt);
float or_imf_sample(float u1, float u2);
void or_star_formation(float* particles, float* stellar, uint32_t n, const or_params* p);

#ifdef __cplusplus
}
#endif
#endif
