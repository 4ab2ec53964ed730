// Tile scheduler of the batched persistent launch (host code, shared by the
// library and the CPU emulator).
//
// Policies (values of SFX_BATCH_* in include/sfx.h):
//   0 LPT          longest-processing-time first: tiles by decreasing
//                  estimated cost, each to the currently least-loaded list;
//   1 CONTIGUOUS   tiles sorted by grid size (largest first) then problem,
//                  cut into equal-cost contiguous ranges, so neighbouring
//                  elements of one problem stay on one work-group;
//   2 ROUND_ROBIN  tile i to list i mod n_wg, in problem order;
//   3 BY_SIZE      one list per grid size (smallest first), tiles in problem
//                  order: each list runs as an ordinary grid of its own
//                  (one work-group per tile), with the wave count best for
//                  that grid size; n_wg is ignored.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <functional>
#include <numeric>
#include <queue>
#include <vector>

namespace sfx {

struct BatchTile {
  int32_t prob;    // problem index
  int32_t block;   // element block within the problem
};

inline int batch_log2(int v) { int k = 0; while ((1 << k) < v) ++k; return k; }

// Estimated cost of one tile: every tile holds 7*256 lane-values; the
// transforms add log2(MP) radix passes to a fixed interpolation/product cost.
inline double batch_tile_cost(int mp) { return 6.0 + batch_log2(mp); }

// Natural tiles (problem order) of problems with `me` elements and grid size `mp`.
inline void batch_tiles(const std::vector<int64_t>& me, const std::vector<int>& mp,
                        std::vector<BatchTile>& tiles, std::vector<double>& cost) {
  tiles.clear();
  cost.clear();
  for (size_t i = 0; i < me.size(); ++i) {
    const int E = 256 / mp[i];
    const int64_t nblk = (me[i] + E - 1) / E;
    for (int64_t k = 0; k < nblk; ++k) {
      tiles.push_back({(int32_t)i, (int32_t)k});
      cost.push_back(batch_tile_cost(mp[i]));
    }
  }
}

// Per-work-group lists: out_tiles in execution order, out_ptr (n_wg + 1).
// Returns the number of lists actually used (n_wg clamped to [1, #tiles]),
// or -1 for an unknown policy.
inline int32_t batch_schedule(const std::vector<BatchTile>& tiles, const std::vector<double>& cost,
                              const std::vector<int>& mp, int policy, int32_t n_wg,
                              std::vector<BatchTile>& out_tiles, std::vector<int32_t>& out_ptr) {
  const int64_t nt = (int64_t)tiles.size();
  if (n_wg > nt) n_wg = (int32_t)std::max<int64_t>(nt, 1);
  if (n_wg < 1) n_wg = 1;
  std::vector<std::vector<int64_t>> lists(n_wg);
  std::vector<int64_t> order(nt);
  std::iota(order.begin(), order.end(), 0);
  if (policy == 0) {
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return cost[x] > cost[y]; });
    using Item = std::pair<double, int32_t>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    for (int32_t w = 0; w < n_wg; ++w) heap.push({0.0, w});
    for (int64_t t : order) {
      Item it = heap.top();
      heap.pop();
      lists[it.second].push_back(t);
      heap.push({it.first + cost[t], it.second});
    }
  } else if (policy == 1) {
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
      const int mx = mp[tiles[x].prob], my = mp[tiles[y].prob];
      if (mx != my) return mx > my;
      return tiles[x].prob < tiles[y].prob;
    });
    double total = 0;
    for (double c : cost) total += c;
    const double per = total / n_wg;
    double acc = 0;
    int32_t w = 0;
    for (int64_t t : order) {
      while (acc >= per * (w + 1) && w + 1 < n_wg) ++w;
      lists[w].push_back(t);
      acc += cost[t];
    }
  } else if (policy == 2) {
    for (int64_t t = 0; t < nt; ++t) lists[t % n_wg].push_back(t);
  } else if (policy == 3) {
    std::vector<int> sizes;
    for (int64_t t = 0; t < nt; ++t) sizes.push_back(mp[tiles[t].prob]);
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    n_wg = (int32_t)std::max<size_t>(sizes.size(), 1);
    lists.assign(n_wg, {});
    for (int64_t t = 0; t < nt; ++t) {
      const int c = (int)(std::lower_bound(sizes.begin(), sizes.end(), mp[tiles[t].prob]) - sizes.begin());
      lists[c].push_back(t);
    }
  } else {
    return -1;
  }
  out_tiles.clear();
  out_ptr.assign(1, 0);
  out_tiles.reserve(nt);
  for (const auto& l : lists) {
    for (int64_t t : l) out_tiles.push_back(tiles[t]);
    out_ptr.push_back((int32_t)out_tiles.size());
  }
  return n_wg;
}

}  // namespace sfx
