// Host validation of what a caller hands the library before any device work:
// a tree in the reference's 64-byte node layout (mortonTree.js:12-21) with
// its levelRanges (mortonTree.js:158), and a host permutation.  Every index
// the device will follow is checked here.  No HIP: the CPU tests build this
// header alone under AddressSanitizer and UBSan.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace gs {

// per node: the walk's skip link (the node after its subtree in pre-order,
// octant 7 first; -1 ends the walk) and its depth
struct TreeLink {
  int32_t skip;
  int32_t depth;
};

// Validates num_nodes 64-byte records (16 words: com.xyz, mass, size,
// isLeaf, first, count, child[8]) of a galaxy of n particles:
//   * a leaf's member range lies inside [0, n);
//   * a child index follows its parent and lies inside the tree, and no node
//     has two parents; depths stay <= max_depth;
//   * level ranges (num_levels pairs [start, count]) are contiguous from node
//     0, cover every node, and each node sits at the depth of its range.
// Fills links, the smallest f32(size * size) over the nodes (NaN ignored)
// and whether some node has zero mass.  Returns nullptr for a valid tree,
// else the reason.
inline const char* check_tree64(const uint32_t* nodes, uint64_t num_nodes, uint32_t n,
                                const uint32_t* level_ranges, uint32_t num_levels,
                                int32_t max_depth, std::vector<TreeLink>& links,
                                float& min_s2, uint32_t& massless) {
  links.assign(num_nodes, TreeLink{-1, 0});
  min_s2 = INFINITY;
  massless = 0;
  for (uint64_t i = 0; i < num_nodes; i++) {
    const uint32_t* w = nodes + i * 16;
    float size, mass;
    memcpy(&size, &w[4], 4);
    memcpy(&mass, &w[3], 4);
    min_s2 = fminf(min_s2, size * size);
    if (mass == 0.0f) massless = 1;
    if (w[5] == 1u) {
      if (!(w[6] <= n && w[7] <= n - w[6])) return "leaf range outside [0, N)";
    } else {
      int32_t prev = -1;
      for (int k = 0; k < 8; k++) {
        const int32_t c = (int32_t)w[8 + k];
        if (!(c < 0 || ((uint64_t)c > i && (uint64_t)c < num_nodes)))
          return "child index must follow its parent and lie inside the tree";
        if (c >= 0) {
          if (links[c].depth != 0) return "node has two parents";
          links[c] = TreeLink{prev >= 0 ? prev : links[i].skip, links[i].depth + 1};
          if (links[c].depth > max_depth) return "tree deeper than 10 levels";
          prev = c;
        }
      }
    }
  }
  if (num_levels > 0) {
    if (!level_ranges) return "bad level ranges";
    uint64_t next = 0;
    for (uint32_t k = 0; k < num_levels; k++) {
      const uint64_t start = level_ranges[2 * k], count = level_ranges[2 * k + 1];
      if (!(start == next && count <= num_nodes - start))
        return "level ranges must be contiguous from node 0 and inside the tree";
      for (uint64_t i = start; i < start + count; i++)
        if ((uint32_t)links[i].depth != k)
          return "a node's depth differs from the level range holding it";
      next = start + count;
    }
    if (next != num_nodes) return "level ranges must cover every node";
  }
  return nullptr;
}

// perm[0 .. n) is a permutation of [0, n)
inline const char* check_permutation(const uint32_t* perm, uint64_t n) {
  std::vector<uint8_t> seen(n, 0);
  for (uint64_t i = 0; i < n; i++) {
    if (perm[i] >= n || seen[perm[i]]) return "not a permutation of [0, N)";
    seen[perm[i]] = 1;
  }
  return nullptr;
}

}  // namespace gs
