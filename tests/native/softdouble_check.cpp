// Host check of galaxy-sim_amd/csrc/softdouble.h against native IEEE double:
// random and edge-case operands, every result compared bit-for-bit.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include "../../galaxy-sim_amd/csrc/softdouble.h"

static uint64_t bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
static double dbl(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float flt(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next() {  // splitmix64
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double rand_double(int mode) {
  uint64_t r = next();
  switch (mode) {
    case 0: { // any finite
      uint64_t b = r; if (((b >> 52) & 0x7ff) == 0x7ff) b &= ~(1ull << 62); return dbl(b);
    }
    case 1: { // moderate exponents
      int e = (int)(next() % 120) - 60;
      return ldexp((double)(r >> 11) / 9007199254740992.0 + 0.5, e) * ((r & 1) ? -1 : 1);
    }
    case 2: { // close to each other / integers
      return (double)(int64_t)(r % 2049) - 1024 + ((r >> 40) % 8) * 0.125;
    }
    default: { // widened f32
      uint32_t f = (uint32_t)r; if (((f >> 23) & 0xff) == 0xff) f &= ~(1u << 30); return (double)flt(f);
    }
  }
}

int main(int argc, char** argv) {
  long iters = argc > 1 ? atol(argv[1]) : 2000000;
  long bad = 0;
  for (long i = 0; i < iters; i++) {
    int ma = (int)(next() % 4), mb = (int)(next() % 4);
    volatile double a = rand_double(ma), b = rand_double(mb);
    volatile double sum = a + b, prod = a * b, quo = b != 0 ? a / b : 0.0;
    if (bits(sum) != sd::add(bits(a), bits(b))) { if (bad++ < 10) printf("add %a %a\n", (double)a, (double)b); }
    if (bits(a - b) != sd::sub(bits(a), bits(b))) { if (bad++ < 10) printf("sub %a %a\n", (double)a, (double)b); }
    if (bits(prod) != sd::mul(bits(a), bits(b))) { if (bad++ < 10) printf("mul %a %a\n", (double)a, (double)b); }
    if (b != 0 && bits(quo) != sd::div(bits(a), bits(b))) { if (bad++ < 10) printf("div %a %a\n", (double)a, (double)b); }
    volatile float fa = (float)a;
    if (fbits(fa) != sd::to_f32(bits(a))) { if (bad++ < 10) printf("to_f32 %a\n", (double)a); }
    uint32_t fr = (uint32_t)next();
    if (((fr >> 23) & 0xff) != 0xff) {
      if (bits((double)flt(fr)) != sd::from_f32(fr)) { if (bad++ < 10) printf("from_f32 %08x\n", fr); }
    }
    // ToInt32 reference: trunc, mod 2^32
    double t = trunc(a);
    double m = fmod(t, 4294967296.0); if (m < 0) m += 4294967296.0;
    int32_t want = std::isfinite((double)a) ? (int32_t)(uint32_t)m : 0;
    if (want != sd::to_int32(bits(a))) { if (bad++ < 10) printf("toint32 %a\n", (double)a); }
  }
  // the Morton constants path: size = max*1.0001 + 1e-9, s = 1024/size
  for (long i = 0; i < iters / 4; i++) {
    float lo = flt((uint32_t)next() & 0xbfffffffu), hi = flt((uint32_t)next() & 0xbfffffffu);
    if (!std::isfinite(lo) || !std::isfinite(hi)) continue;
    volatile double d = (double)hi - (double)lo;
    volatile double size = d * 1.0001 + 1e-9;
    volatile double sc = 1024 / size;
    uint64_t sd_d = sd::sub(sd::from_f32(fbits(hi)), sd::from_f32(fbits(lo)));
    uint64_t sd_size = sd::add(sd::mul(sd_d, bits(1.0001)), bits(1e-9));
    uint64_t sd_s = sd::div(bits(1024.0), sd_size);
    if (bits(d) != sd_d || bits(size) != sd_size || bits(sc) != sd_s) { if (bad++ < 10) printf("chain %a %a\n", lo, hi); }
  }
  printf("bad=%ld\n", bad);
  return bad != 0;
}
