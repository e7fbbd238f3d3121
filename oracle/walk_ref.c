/*
 * Oracle (test infrastructure only): FP32 transliteration of
 * src/physics/shaders/nbody_bh.wgsl.  Built with -ffp-contract=off; sqrtf and
 * '/' are IEEE-correct on the host, which is the fixed meaning this project
 * gives WGSL's implementation-defined sqrt / division precision.  Vector
 * expressions are expanded component-wise, dot(a,b) = (ax*bx + ay*by) + az*bz.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define STACK_SIZE 96 /* nbody_bh.wgsl:62 */

static inline float nf(const uint32_t* w, uint64_t i) {
  float f;
  memcpy(&f, w + i, 4);
  return f;
}
static inline void nfs(uint32_t* w, uint64_t i, float f) { memcpy(w + i, &f, 4); }

/* NBodyGPU.updateParams (NBodyGPU.js:253-275): every field goes through a
 * Float32Array store; eps2 is epsilon*epsilon in double, then rounded. */
void or_pack_params(or_params* p, double dt, double G, double epsilon,
                    double theta, double halo_mass, double halo_a,
                    double gas_drag, uint32_t n, uint32_t halo_count,
                    const double* c1, const double* c2) {
  memset(p, 0, sizeof *p);
  p->dt = (float)dt;
  p->G = (float)G;
  p->eps2 = (float)(epsilon * epsilon);
  p->n = n;
  p->theta = (float)theta;
  p->halo_mass = (float)halo_mass;
  p->halo_a = (float)halo_a;
  p->gas_drag = (float)gas_drag;
  p->time = 0.0f;
  p->aging_rate = 25.0f;
  p->sf_density = 1e9f;
  p->sf_eff = 0.0f;
  p->sf_dt = 0.0f;
  p->halo_count = halo_count > 2 ? 2 : halo_count;
  for (int c = 0; c < 4; c++) {
    p->halo_c1[c] = c1 && c < 3 ? (float)c1[c] : 0.0f;
    p->halo_c2[c] = c2 && c < 3 ? (float)c2[c] : 0.0f;
  }
  p->drag_damp = expf(-p->gas_drag * p->dt);
}

typedef struct {
  float ax, ay, az, pot, dens;
  uint32_t inter;
} walk_out;

/* nbody_bh.wgsl:69-142: the stack walk for particle i. */
static walk_out walk_one(const float* pin, uint32_t i, const uint32_t* w,
                         const or_params* p) {
  const float* pi = pin + (size_t)i * 8;
  float px = pi[0], py = pi[1], pz = pi[2];
  float ax = 0.0f, ay = 0.0f, az = 0.0f, pot = 0.0f, density = 0.0f;
  uint32_t inter = 0;
  float theta2 = p->theta * p->theta;
  float eps2 = p->eps2;
  float G = p->G;
  int32_t stack[STACK_SIZE];
  int sp = 1;
  stack[0] = 0;
  while (sp > 0) {
    sp -= 1;
    uint64_t b = (uint64_t)stack[sp] * 16;
    float mass = nf(w, b + 3);
    if (mass == 0.0f) continue;
    float cx = nf(w, b), cy = nf(w, b + 1), cz = nf(w, b + 2);
    float size = nf(w, b + 4);
    int leaf = w[b + 5] == 1u;
    uint32_t first = w[b + 6], count = w[b + 7];
    if (leaf && i >= first && i < first + count)
      density = (float)count / ((size * size) * size);
    float rx = cx - px, ry = cy - py, rz = cz - pz;
    float geom = (rx * rx + ry * ry) + rz * rz;
    float d2 = geom + eps2;
    float s2 = size * size;
    if (s2 < theta2 * d2 || s2 < 0.25f * eps2) {
      float f = (G * mass) / (d2 * sqrtf(d2));
      ax = ax + rx * f;
      ay = ay + ry * f;
      az = az + rz * f;
      pot = pot - f * d2;
      inter++;
    } else if (leaf) {
      for (uint32_t j = first; j < first + count; j++) {
        if (j == i) continue;
        const float* pj = pin + (size_t)j * 8;
        float qx = pj[0] - px, qy = pj[1] - py, qz = pj[2] - pz;
        float dj2 = ((qx * qx + qy * qy) + qz * qz) + eps2;
        float f = (G * pj[3]) / (dj2 * sqrtf(dj2));
        ax = ax + qx * f;
        ay = ay + qy * f;
        az = az + qz * f;
        pot = pot - f * dj2;
        inter++;
      }
    } else {
      for (int c = 0; c < 8; c++) {
        int32_t ch = (int32_t)w[b + 8 + c];
        if (ch >= 0 && sp < STACK_SIZE) stack[sp++] = ch;
      }
    }
  }
  walk_out o = {ax, ay, az, pot, density, inter};
  return o;
}

/* nbody_bh.wgsl:144-180: halos, kick-drift, gas drag, pad, energy. */
static void integrate_one(const float* pin, uint32_t i, const or_params* p,
                          walk_out wo, float* pout, float* energy) {
  const float* pi = pin + (size_t)i * 8;
  float px = pi[0], py = pi[1], pz = pi[2];
  float ax = wo.ax, ay = wo.ay, az = wo.az;
  float pot_halo = 0.0f;
  if (p->halo_mass > 0.0f) {
    for (uint32_t h = 0; h < p->halo_count; h++) {
      const float* c = h == 0 ? p->halo_c1 : p->halo_c2;
      float dx = px - c[0], dy = py - c[1], dz = pz - c[2];
      float len = sqrtf((dx * dx + dy * dy) + dz * dz);
      float rlen = fmaxf(len, 1e-6f);
      float ra = rlen + p->halo_a;
      float gm = p->G * p->halo_mass;
      float t = gm / ((ra * ra) * rlen);
      ax = ax - dx * t;
      ay = ay - dy * t;
      az = az - dz * t;
      pot_halo = pot_halo - gm / ra;
    }
  }
  float dt = p->dt;
  float vx = pi[4] + ax * dt, vy = pi[5] + ay * dt, vz = pi[6] + az * dt;
  int gas = pi[7] < -0.5f;
  if (gas && p->gas_drag > 0.0f) {
    float hx = -py, hy = px, hz = 1e-6f;
    float hl = sqrtf((hx * hx + hy * hy) + hz * hz);
    hx = hx / hl;
    hy = hy / hl;
    hz = hz / hl;
    float s = (vx * hx + vy * hy) + vz * hz;
    float fx = hx * s, fy = hy * s, fz = hz * s;
    float k = p->drag_damp;
    vx = fx + (vx - fx) * k;
    vy = fy + (vy - fy) * k;
    vz = fz + (vz - fz) * k;
  }
  if (pout) {
    float* po = pout + (size_t)i * 8;
    po[0] = px + vx * dt;
    po[1] = py + vy * dt;
    po[2] = pz + vz * dt;
    po[3] = pi[3];
    po[4] = vx;
    po[5] = vy;
    po[6] = vz;
    po[7] = gas ? -(1.0f + wo.dens) : 0.0f;
  }
  if (energy)
    energy[i] = (0.5f * ((vx * vx + vy * vy) + vz * vz) + 0.5f * wo.pot) + pot_halo;
}

typedef struct {
  const float* pin;
  uint32_t n;
  const uint32_t* nodes;
  const or_params* p;
  const uint32_t* idx;
  float *pout, *energy, *acc4, *dens;
  uint32_t* inter;
  uint32_t count;
  uint32_t* next; /* shared work counter (chunks of WALK_CHUNK) */
} job_t;

#define WALK_CHUNK 256u

static void* walk_worker(void* arg) {
  job_t* j = (job_t*)arg;
  for (;;) {
    uint32_t begin = __atomic_fetch_add(j->next, WALK_CHUNK, __ATOMIC_RELAXED);
    if (begin >= j->count) break;
    uint32_t end = begin + WALK_CHUNK < j->count ? begin + WALK_CHUNK : j->count;
    for (uint32_t k = begin; k < end; k++) {
      uint32_t i = j->idx ? j->idx[k] : k;
      walk_out wo = walk_one(j->pin, i, j->nodes, j->p);
      if (j->acc4) {
        j->acc4[(size_t)k * 4] = wo.ax;
        j->acc4[(size_t)k * 4 + 1] = wo.ay;
        j->acc4[(size_t)k * 4 + 2] = wo.az;
        j->acc4[(size_t)k * 4 + 3] = wo.pot;
      }
      if (j->dens) j->dens[k] = wo.dens;
      if (j->inter) j->inter[k] = wo.inter;
      if (!j->idx && (j->pout || j->energy))
        integrate_one(j->pin, i, j->p, wo, j->pout, j->energy);
    }
  }
  return NULL;
}

static void run_jobs(job_t proto, uint32_t count, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  uint32_t next = 0;
  proto.count = count;
  proto.next = &next;
  for (int t = 0; t < nthreads; t++)
    pthread_create(&th[t], NULL, walk_worker, &proto);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
}

void or_walk(const float* pin, uint32_t n, const uint32_t* nodes,
             const or_params* p, float* pout, float* energy, float* acc4,
             float* dens, uint32_t* inter, int nthreads) {
  job_t j;
  memset(&j, 0, sizeof j);
  j.pin = pin;
  j.n = n;
  j.nodes = nodes;
  j.p = p;
  j.pout = pout;
  j.energy = energy;
  j.acc4 = acc4;
  j.dens = dens;
  j.inter = inter;
  run_jobs(j, n, nthreads);
}

void or_walk_subset(const float* pin, uint32_t n, const uint32_t* nodes,
                    const or_params* p, const uint32_t* idx, uint32_t m,
                    float* acc4, uint32_t* inter, int nthreads) {
  job_t j;
  memset(&j, 0, sizeof j);
  j.pin = pin;
  j.n = n;
  j.nodes = nodes;
  j.p = p;
  j.idx = idx;
  j.acc4 = acc4;
  j.inter = inter;
  run_jobs(j, m, nthreads);
}

/* nbody_bh.wgsl:288-313, dispatched deepest level first (NBodyGPU.js:312). */
void or_refresh_f32(const float* pin, uint32_t* w, const uint32_t* lr,
                    uint32_t num_levels) {
  for (int lvl = (int)num_levels - 1; lvl >= 0; lvl--) {
    uint32_t start = lr[2 * lvl], count = lr[2 * lvl + 1];
    for (uint32_t t = 0; t < count; t++) {
      uint64_t b = (uint64_t)(start + t) * 16;
      float cx = 0.0f, cy = 0.0f, cz = 0.0f, m = 0.0f;
      if (w[b + 5] == 1u) {
        uint32_t first = w[b + 6], cnt = w[b + 7];
        for (uint32_t j = first; j < first + cnt; j++) {
          const float* pj = pin + (size_t)j * 8;
          cx = cx + pj[0] * pj[3];
          cy = cy + pj[1] * pj[3];
          cz = cz + pj[2] * pj[3];
          m = m + pj[3];
        }
      } else {
        for (int c = 0; c < 8; c++) {
          int32_t ch = (int32_t)w[b + 8 + c];
          if (ch >= 0) {
            uint64_t cb = (uint64_t)ch * 16;
            float cm = nf(w, cb + 3);
            cx = cx + nf(w, cb) * cm;
            cy = cy + nf(w, cb + 1) * cm;
            cz = cz + nf(w, cb + 2) * cm;
            m = m + cm;
          }
        }
      }
      int ok = m > 0.0f;
      nfs(w, b, ok ? cx / m : 0.0f);
      nfs(w, b + 1, ok ? cy / m : 0.0f);
      nfs(w, b + 2, ok ? cz / m : 0.0f);
      nfs(w, b + 3, m);
    }
  }
}

/* FP64 refresh in the build's order (mortonTree.js:129-137, 166-181). */
void or_refresh_f64(const float* pin, uint32_t* w, uint64_t num_nodes) {
  for (int64_t nd = (int64_t)num_nodes - 1; nd >= 0; nd--) {
    uint64_t b = (uint64_t)nd * 16;
    double m = 0, cx = 0, cy = 0, cz = 0;
    if (w[b + 5] == 1u) {
      uint32_t first = w[b + 6], cnt = w[b + 7];
      for (uint32_t j = first; j < first + cnt; j++) {
        const float* pj = pin + (size_t)j * 8;
        double pm = pj[3];
        m += pm;
        cx += (double)pj[0] * pm;
        cy += (double)pj[1] * pm;
        cz += (double)pj[2] * pm;
      }
    } else {
      for (int c = 0; c < 8; c++) {
        int32_t ch = (int32_t)w[b + 8 + c];
        if (ch >= 0) {
          uint64_t cb = (uint64_t)ch * 16;
          double cm = nf(w, cb + 3);
          m += cm;
          cx += (double)nf(w, cb) * cm;
          cy += (double)nf(w, cb + 1) * cm;
          cz += (double)nf(w, cb + 2) * cm;
        }
      }
    }
    double inv = m > 0 ? 1 / m : 0;
    nfs(w, b, (float)(cx * inv));
    nfs(w, b + 1, (float)(cy * inv));
    nfs(w, b + 2, (float)(cz * inv));
    nfs(w, b + 3, (float)m);
  }
}

/* nbody_bh.wgsl:321-326: out[i] = in[perm[i]] over `stride` floats. */
void or_apply_permutation(const float* in, const uint32_t* perm, uint32_t n,
                          uint32_t stride, float* out) {
  for (uint32_t i = 0; i < n; i++)
    memcpy(out + (size_t)i * stride, in + (size_t)perm[i] * stride,
           sizeof(float) * stride);
}

typedef struct {
  const float* pin;
  uint32_t n;
  double G, eps2;
  const uint32_t* idx;
  double* acc3;
  uint32_t begin, end;
} ds_job;

/* Samples go in tiles of DS_TILE per pass over the sources, so each source
 * record is read once per tile; every sample's sum still runs over q in
 * ascending order, so the results do not depend on the tiling. */
enum { DS_TILE = 8 };

static void* ds_worker(void* arg) {
  ds_job* j = (ds_job*)arg;
  for (uint32_t k0 = j->begin; k0 < j->end; k0 += DS_TILE) {
    uint32_t t = j->end - k0 < DS_TILE ? j->end - k0 : DS_TILE;
    uint32_t ii[DS_TILE];
    double px[DS_TILE], py[DS_TILE], pz[DS_TILE];
    double ax[DS_TILE], ay[DS_TILE], az[DS_TILE];
    for (uint32_t u = 0; u < t; u++) {
      ii[u] = j->idx[k0 + u];
      const float* pi = j->pin + (size_t)ii[u] * 8;
      px[u] = pi[0];
      py[u] = pi[1];
      pz[u] = pi[2];
      ax[u] = ay[u] = az[u] = 0;
    }
    for (uint32_t q = 0; q < j->n; q++) {
      const float* pq = j->pin + (size_t)q * 8;
      double qx = pq[0], qy = pq[1], qz = pq[2], mq = j->G * pq[3];
      for (uint32_t u = 0; u < t; u++) {
        if (q == ii[u]) continue;
        double dx = qx - px[u], dy = qy - py[u], dz = qz - pz[u];
        double d2 = dx * dx + dy * dy + dz * dz + j->eps2;
        double f = mq / (d2 * sqrt(d2));
        ax[u] += dx * f;
        ay[u] += dy * f;
        az[u] += dz * f;
      }
    }
    for (uint32_t u = 0; u < t; u++) {
      j->acc3[(size_t)(k0 + u) * 3] = ax[u];
      j->acc3[(size_t)(k0 + u) * 3 + 1] = ay[u];
      j->acc3[(size_t)(k0 + u) * 3 + 2] = az[u];
    }
  }
  return NULL;
}

void or_direct_sum_f64(const float* pin, uint32_t n, double G, double eps2,
                       const uint32_t* idx, uint32_t m, double* acc3,
                       int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if ((uint32_t)nthreads > m) nthreads = m ? (int)m : 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  ds_job* jobs = (ds_job*)malloc(sizeof(ds_job) * nthreads);
  for (int t = 0; t < nthreads; t++) {
    ds_job jj = {pin, n, G, eps2, idx, acc3,
                 (uint32_t)((uint64_t)m * t / nthreads),
                 (uint32_t)((uint64_t)m * (t + 1) / nthreads)};
    jobs[t] = jj;
    pthread_create(&th[t], NULL, ds_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(jobs);
}

/* ---- star_formation (nbody_bh.wgsl:197-268) ------------------------------
 * pcg (nbody_bh.wgsl:206-210): the PCG-RXS-M-XS 32-bit hash — one step of
 * the 32-bit PCG LCG (multiplier 747796405, increment 2891336453) followed by
 * the RXS-M-XS output permutation (multiplier 277803737). */
uint32_t or_pcg(uint32_t v0) {
  uint32_t v = v0 * 747796405u + 2891336453u;
  v = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
  return (v >> 22u) ^ v;
}

/* rand01 (nbody_bh.wgsl:211-213): f32(pcg(i ^ salt * 0x9E3779B9)) / f32(4294967295.0) */
float or_rand01(uint32_t i, uint32_t salt) {
  return (float)or_pcg(i ^ (salt * 0x9E3779B9u)) / (float)4294967295.0;
}

/* imfSample (nbody_bh.wgsl:216-227).  The shader's pow(0.08, -0.3) etc. are
 * constant expressions: evaluated here in double and rounded to f32 once; the
 * rest is f32 arithmetic, left to right, with the library powf. */
float or_imf_sample(float u1, float u2) {
  const float a0 = (float)pow(0.08, -0.3);
  const float b0 = (float)pow(0.5, -0.3);
  const float lowInt = (a0 - b0) / 0.3f;
  const float a1 = (float)pow(0.5, -1.3);
  const float b1 = (float)pow(100.0, -1.3);
  const float hiInt = 0.5f * (a1 - b1) / 1.3f;
  if (u1 < lowInt / (lowInt + hiInt)) return powf(a0 - u2 * (a0 - b0), (float)(-1.0 / 0.3));
  return powf(a1 - u2 * (a1 - b1), (float)(-1.0 / 1.3));
}

/* star_formation (nbody_bh.wgsl:229-268), in place on n particle records
 * (8 f32) and stellar records (4 f32). */
void or_star_formation(float* particles, float* stellar, uint32_t n, const or_params* p) {
  const uint32_t salt = (uint32_t)(p->time * 977.0f) + 1u;
  for (uint32_t i = 0; i < n; i++) {
    float* q = particles + (size_t)i * 8;
    float* st = stellar + (size_t)i * 4;
    if (q[7] < -0.5f) {
      const float density = -q[7] - 1.0f;
      if (density > p->sf_density && p->sf_eff > 0.0f) {
        const float prob = p->sf_eff * p->sf_dt *
                           fminf(sqrtf(density / fmaxf(p->sf_density, 1e-6f)), 3.0f);
        if (or_rand01(i, salt) < prob) {
          const float m = or_imf_sample(or_rand01(i, salt + 7u), or_rand01(i, salt + 13u));
          const float birthMyr = p->time * p->aging_rate;
          st[0] = m;
          st[1] = birthMyr;
          st[2] = or_rand01(i, salt + 29u);
          st[3] = 0.0f;
          q[7] = 0.0f;
        }
      }
    } else if (st[0] >= 8.0f) {
      const float ageMyr = p->time * p->aging_rate - st[1];
      const float tMS = 10000.0f * powf(st[0], -2.5f);
      const float tGiant = tMS * 0.12f + 15.0f;
      if (ageMyr > tMS + tGiant + 15.0f + 30.0f) {
        const float seed = st[2];
        st[0] = -1.0f;
        st[1] = 0.0f;
        st[2] = seed;
        st[3] = 0.7f;
        q[7] = -1.0f;
        q[4] += (or_rand01(i, salt + 41u) - 0.5f) * 2.0f;
        q[5] += (or_rand01(i, salt + 43u) - 0.5f) * 2.0f;
        q[6] += (or_rand01(i, salt + 47u) - 0.5f) * 2.0f;
      }
    }
  }
}
