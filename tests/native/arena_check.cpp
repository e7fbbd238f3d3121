// Randomised check of gs::RangeList (csrc/arena.h) against a unit bitmap:
// after every take/give the free ranges are exactly the bitmap's free runs
// (coalesced, disjoint, inside the resource), a take returns the smallest run
// that holds it (lowest start among equals) or fails only when no run does,
// and a give of a range that is not wholly taken is refused without change.
// Built with -fsanitize=address,undefined by the test that runs it.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../../galaxy-sim_amd/csrc/arena.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}

struct Model {
  std::vector<char> used;
  explicit Model(uint64_t n) : used(n, 0) {}
  // free runs as (start, len)
  std::vector<std::pair<uint64_t, uint64_t>> runs() const {
    std::vector<std::pair<uint64_t, uint64_t>> r;
    for (uint64_t i = 0; i < used.size();) {
      if (used[i]) {
        i++;
        continue;
      }
      uint64_t j = i;
      while (j < used.size() && !used[j]) j++;
      r.push_back({i, j - i});
      i = j;
    }
    return r;
  }
};

static int bad = 0;
#define EXPECT(c)                                              \
  do {                                                         \
    if (!(c)) {                                                \
      if (bad < 10) printf("FAIL line %d: %s\n", __LINE__, #c); \
      bad++;                                                   \
    }                                                          \
  } while (0)

static void compare(const gs::RangeList& L, const Model& m) {
  auto r = m.runs();
  EXPECT(L.pieces() == r.size());
  size_t k = 0;
  for (const auto& f : L.ranges()) {
    if (k >= r.size()) break;
    EXPECT(f.first == r[k].first && f.second == r[k].second);
    k++;
  }
}

int main(int argc, char** argv) {
  const int rounds = argc > 1 ? atoi(argv[1]) : 2000;
  uint64_t ops = 0;
  for (int round = 0; round < rounds; round++) {
    const uint64_t total = 1 + rnd() % 300;
    gs::RangeList L;
    L.reset(total);
    Model m(total);
    std::vector<std::pair<uint64_t, uint64_t>> taken;
    for (int op = 0; op < 200; op++, ops++) {
      const uint64_t kind = rnd() % 10;
      if (kind < 5) {  // take
        const uint64_t len = rnd() % 40;  // 0 is refused
        uint64_t start = ~0ull;
        const bool ok = L.take(len, &start);
        auto runs = m.runs();
        // expected: best fit
        int64_t best = -1;
        for (size_t i = 0; i < runs.size(); i++)
          if (len && runs[i].second >= len && (best < 0 || runs[i].second < runs[best].second))
            best = (int64_t)i;
        EXPECT(ok == (best >= 0));
        if (ok && best >= 0) {
          EXPECT(start == runs[best].first);
          for (uint64_t u = start; u < start + len; u++) m.used[u] = 1;
          taken.push_back({start, len});
        }
      } else if (kind < 9 && !taken.empty()) {  // give back a taken range
        const size_t i = rnd() % taken.size();
        const auto t = taken[i];
        EXPECT(L.give(t.first, t.second));
        for (uint64_t u = t.first; u < t.first + t.second; u++) m.used[u] = 0;
        taken[i] = taken.back();
        taken.pop_back();
      } else {  // a give that must be refused: outside, or touching a free unit
        const uint64_t start = rnd() % (total + 5), len = 1 + rnd() % 20;
        bool all_used = start < total && len <= total - start;
        if (all_used)
          for (uint64_t u = start; u < start + len; u++) all_used &= m.used[u] != 0;
        if (!all_used) EXPECT(!L.give(start, len));
      }
      compare(L, m);
      uint64_t nfree = 0;
      for (char c : m.used) nfree += !c;
      EXPECT(L.free_units() == nfree);
    }
  }
  printf("ops=%llu bad=%d\n", (unsigned long long)ops, bad);
  return bad ? 1 : 0;
}
