// Host check of galaxy-sim_amd/csrc/quantize.h (the encode's Morton
// quantization without binary64 hardware arithmetic) against native binary64
// arithmetic evaluating mortonTree.js:53-58 as JS does:
//   Math.min(((p - min) * s) | 0, 1023),   s = 1024 / (maxExtent * 1.0001 + 1e-9)
// Boxes of many scales and offsets; coordinates uniform in the box, placed
// within a few ulps of cell boundaries (where the f32 fast path must defer to
// the exact path), on the box edges, and outside the box (ToInt32 wraps).
// Prints the count of mismatches and the share of coordinates that took the
// exact path.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../galaxy-sim_amd/csrc/quantize.h"

static uint64_t bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

static uint64_t st = 0x243F6A8885A308D3ull;
static uint64_t next() {  // splitmix64
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unit() { return (double)(next() >> 11) / 9007199254740992.0; }

// ECMAScript ToInt32 of a binary64 value, then Math.min(., 1023), low bits
static uint32_t js_quant(float p, float mn, double s) {
  if (!std::isfinite(p)) return 0;
  volatile double d = (double)p - (double)mn;
  volatile double q = d * s;
  double t = trunc(q);
  double m = fmod(t, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  int32_t v = std::isfinite(q) ? (int32_t)(uint32_t)m : 0;
  return (uint32_t)(v < 1023 ? v : 1023);
}

int main(int argc, char** argv) {
  long boxes = argc > 1 ? atol(argv[1]) : 20000;
  long bad = 0, total = 0, slow = 0;
  for (long b = 0; b < boxes; b++) {
    // box: minimum and extent over many scales (1e-30 .. 1e30), some boxes
    // centred far from the origin relative to their extent
    const double scale = pow(10.0, -30.0 + 60.0 * unit());
    const double off = (next() & 1 ? -1.0 : 1.0) * scale * pow(10.0, 6.0 * unit() - 3.0);
    const float mn = (float)(off - scale * unit());
    float mx = (float)((double)mn + scale * (0.1 + unit()));
    if (!std::isfinite(mn) || !std::isfinite(mx) || !(mx >= mn)) continue;
    if (next() % 50 == 0) mx = mn;  // every particle at one point: size = 1e-9
    volatile double ext = (double)mx - (double)mn;
    volatile double size = ext * 1.0001 + 1e-9;
    volatile double s = 1024.0 / size;
    const float s32 = (float)s;
    const uint64_t mnb = bits((double)mn), sb = bits(s);
    for (int k = 0; k < 64; k++) {
      float p;
      const uint64_t kind = next() % 8;
      if (kind < 3) {  // uniform in the box
        p = (float)((double)mn + ext * unit());
        if (p > mx) p = mx;
      } else if (kind < 6) {  // near a cell boundary, nudged by a few ulps
        const double c = (double)(next() % 1025) / s;
        p = (float)((double)mn + c);
        const int nud = (int)(next() % 9) - 4;
        for (int u = 0; u < (nud < 0 ? -nud : nud); u++)
          p = nextafterf(p, nud < 0 ? -INFINITY : INFINITY);
      } else if (kind == 6) {  // the box edges and their neighbours
        p = next() & 1 ? mn : mx;
        if (next() & 1) p = nextafterf(p, next() & 1 ? -INFINITY : INFINITY);
      } else {  // outside the box: a particle left out of the bounding box
        p = (float)((double)mn + (unit() * 4.0 - 2.0) * scale * pow(10.0, 12.0 * unit()));
      }
      if (!std::isfinite(p)) continue;
      bool ex = false;
      const uint32_t got = gsq::quantize(p, mn, s32, mnb, sb, &ex);
      const uint32_t want = js_quant(p, mn, s);
      total++;
      slow += ex;
      if (got != want) {
        if (bad++ < 10)
          printf("mismatch p=%a mn=%a s=%a got %u want %u\n", (double)p, (double)mn, s, got, want);
      }
    }
  }
  // non-finite coordinates
  const float nf[3] = {NAN, INFINITY, -INFINITY};
  for (float p : nf) {
    total++;
    if (gsq::quantize(p, 0.0f, 1.0f, 0, bits(1.0)) != 0) bad++;
  }
  printf("bad=%ld total=%ld exact_path=%.4f\n", bad, total, (double)slow / (double)total);
  return bad != 0;
}
