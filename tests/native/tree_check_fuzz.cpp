// Randomised check of the host validation in csrc/tree_check.h, built with
// -fsanitize=address,undefined by the test that runs it:
//   * random BFS trees in the 64-byte layout (with their level ranges) pass,
//     and the links hold each node's depth and pre-order skip;
//   * random word mutations of those trees are accepted only when an
//     independent brute-force check of the same rules accepts them;
//   * check_permutation accepts exactly the permutations.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../galaxy-sim_amd/csrc/tree_check.h"

static uint64_t rs = 0x2545F4914F6CDD1Dull;
static uint64_t rnd() {
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return rs;
}

static int bad = 0;
#define EXPECT(c)                                               \
  do {                                                          \
    if (!(c)) {                                                 \
      if (bad < 10) printf("FAIL line %d: %s\n", __LINE__, #c); \
      bad++;                                                    \
    }                                                           \
  } while (0)

struct Tree {
  std::vector<uint32_t> w;     // 16 words per node
  std::vector<uint32_t> lr;    // level ranges
  std::vector<int> depth, parent;
  uint32_t n;
};

// a random BFS tree: nodes numbered level by level, children contiguous and
// in octant order, leaves holding disjoint ranges of [0, n)
static Tree make_tree(uint32_t n) {
  Tree t;
  t.n = n;
  std::vector<uint32_t> level = {0};
  t.depth.push_back(0);
  t.parent.push_back(-1);
  uint32_t next = 1, cursor = 0;
  int d = 0;
  std::vector<std::vector<int>> kids(1);
  while (!level.empty() && d < 10) {
    std::vector<uint32_t> nl;
    for (uint32_t v : level) {
      const int nc = (rnd() % 3 == 0 || next > 400) ? 0 : (int)(1 + rnd() % 8);
      for (int k = 0; k < nc; k++) {
        kids[v].push_back((int)next);
        kids.emplace_back();
        t.depth.push_back(d + 1);
        t.parent.push_back((int)v);
        nl.push_back(next++);
      }
    }
    level = nl;
    d++;
  }
  const uint32_t nn = next;
  t.w.assign((size_t)nn * 16, 0);
  for (uint32_t v = 0; v < nn; v++) {
    uint32_t* w = &t.w[(size_t)v * 16];
    float size = 1.0f / (float)(1 << t.depth[v]), mass = (rnd() % 7) ? 1.0f : 0.0f;
    memcpy(&w[3], &mass, 4);
    memcpy(&w[4], &size, 4);
    if (kids[v].empty()) {
      w[5] = 1;
      const uint32_t cnt = cursor < n ? (uint32_t)(rnd() % (n - cursor + 1)) % 9 : 0;
      w[6] = cursor;
      w[7] = cnt;
      cursor += cnt;
    } else {
      w[5] = 0;
      for (int k = 0; k < 8; k++) w[8 + k] = 0xffffffffu;
      // occupied octants in increasing order
      int oct = 0;
      for (size_t c = 0; c < kids[v].size(); c++) {
        oct += (int)(rnd() % std::max(1, (int)(8 - oct - (kids[v].size() - c) + 1)));
        if (oct > 7) oct = 7;
        w[8 + oct] = (uint32_t)kids[v][c];
        oct++;
      }
    }
  }
  int maxd = 0;
  for (int x : t.depth) maxd = std::max(maxd, x);
  uint32_t start = 0;
  for (int k = 0; k <= maxd; k++) {
    uint32_t cnt = 0;
    for (uint32_t v = 0; v < nn; v++) cnt += t.depth[v] == k;
    t.lr.push_back(start);
    t.lr.push_back(cnt);
    start += cnt;
  }
  return t;
}

// the same rules, stated directly
static bool brute_ok(const std::vector<uint32_t>& w, uint64_t nn, uint32_t n,
                     const std::vector<uint32_t>& lr) {
  std::vector<int> par(nn, -1), dep(nn, 0);
  for (uint64_t i = 0; i < nn; i++) {
    const uint32_t* r = &w[i * 16];
    if (r[5] == 1u) {
      if ((uint64_t)r[6] + r[7] > n) return false;
      continue;
    }
    for (int k = 0; k < 8; k++) {
      const int32_t c = (int32_t)r[8 + k];
      if (c < 0) continue;
      if ((uint64_t)c <= i || (uint64_t)c >= nn) return false;
      if (par[c] >= 0) return false;
      par[c] = (int)i;
      dep[c] = dep[i] + 1;
      if (dep[c] > 10) return false;
    }
  }
  uint64_t next = 0;
  for (size_t k = 0; k < lr.size() / 2; k++) {
    const uint64_t s = lr[2 * k], c = lr[2 * k + 1];
    if (s != next || s + c > nn) return false;
    for (uint64_t i = s; i < s + c; i++)
      if (dep[i] != (int)k) return false;
    next = s + c;
  }
  return lr.empty() || next == nn;
}

int main(int argc, char** argv) {
  const int rounds = argc > 1 ? atoi(argv[1]) : 300;
  uint64_t checked = 0, accepted_mut = 0;
  for (int r = 0; r < rounds; r++) {
    const uint32_t n = 1 + (uint32_t)(rnd() % 3000);
    Tree t = make_tree(n);
    const uint64_t nn = t.w.size() / 16;
    std::vector<gs::TreeLink> links;
    float ms = 0;
    uint32_t ml = 0;
    const char* why = gs::check_tree64(t.w.data(), nn, n, t.lr.data(), (uint32_t)(t.lr.size() / 2),
                                       10, links, ms, ml);
    EXPECT(why == nullptr);
    if (!why) {
      for (uint64_t v = 0; v < nn; v++) EXPECT(links[v].depth == t.depth[v]);
      EXPECT(links[0].skip == -1);
    }
    EXPECT(brute_ok(t.w, nn, n, t.lr));
    // mutations
    for (int m = 0; m < 40; m++) {
      std::vector<uint32_t> w = t.w, lr = t.lr;
      const int kind = (int)(rnd() % 4);
      if (kind == 0) {
        w[(rnd() % nn) * 16 + 8 + rnd() % 8] = (uint32_t)(rnd() % (nn + 3)) - 1u;
      } else if (kind == 1) {
        const uint64_t v = rnd() % nn;
        w[v * 16 + 5] = 1;
        w[v * 16 + 6] = (uint32_t)(rnd() % (n + 4));
        w[v * 16 + 7] = (uint32_t)(rnd() % 12);
      } else if (kind == 2 && !lr.empty()) {
        lr[rnd() % lr.size()] += (uint32_t)(rnd() % 3) - 1u;
      } else {
        w[(rnd() % nn) * 16 + 6 + rnd() % 10] = (uint32_t)rnd();
      }
      const char* why2 = gs::check_tree64(w.data(), nn, n, lr.data(), (uint32_t)(lr.size() / 2),
                                          10, links, ms, ml);
      const bool ok = brute_ok(w, nn, n, lr);
      EXPECT((why2 == nullptr) == ok);
      accepted_mut += why2 == nullptr;
      checked++;
    }
    // permutations
    std::vector<uint32_t> p(n);
    for (uint32_t i = 0; i < n; i++) p[i] = i;
    for (uint32_t i = n; i > 1; i--) std::swap(p[i - 1], p[rnd() % i]);
    EXPECT(gs::check_permutation(p.data(), n) == nullptr);
    std::vector<uint32_t> q = p;
    q[rnd() % n] = (uint32_t)(rnd() % (n + 2));
    std::vector<uint32_t> sq = q;
    std::sort(sq.begin(), sq.end());
    bool isperm = true;
    for (uint32_t i = 0; i < n; i++) isperm &= sq[i] == i;
    EXPECT((gs::check_permutation(q.data(), n) == nullptr) == isperm);
  }
  printf("mutations=%llu accepted=%llu bad=%d\n", (unsigned long long)checked,
         (unsigned long long)accepted_mut, bad);
  return bad ? 1 : 0;
}
