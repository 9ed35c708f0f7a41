// qd_pack.hpp -- host-side packing of the R pass's mixing units into bundles
// (the work items of k_rpass_mx).  Header-only, no HIP, so it is also built
// and checked on the CPU.
//
// A unit is the g lines (the K blocks) of one (J, j) group at one r; every
// group has n2 units, taken in r order.  A bundle holds at most nl_max lines
// and the units of at most two groups (one unit matrix each).  Each bundle is
// filled with units of the largest group that still has units, as many as fit,
// then the free lines with units of the largest group that fits.
#pragma once

#include <algorithm>
#include <utility>
#include <vector>

namespace qdp {

// One run of units of one group inside a bundle: units r0 .. r0 + count - 1.
struct Run {
  int group, r0, count;
};

// bundles[b] = the runs of bundle b, in line order.  group_size[i] = lines per
// unit of group i (1 .. nl_max).  Returns false if a group does not fit.
inline bool pack_bundles(const std::vector<int> &group_size, int n2, int nl_max,
                         std::vector<std::vector<Run>> &bundles) {
  bundles.clear();
  const int ng = (int)group_size.size();
  for (int g : group_size)
    if (g < 1 || g > nl_max) return false;
  std::vector<int> order(ng), next_r(ng, 0);
  for (int i = 0; i < ng; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return group_size[a] > group_size[b]; });
  size_t head = 0;
  while (true) {
    while (head < order.size() && next_r[order[head]] >= n2) head++;
    if (head == order.size()) break;
    std::vector<Run> runs;
    const int big = order[head], gb = group_size[big];
    const int k = std::min(n2 - next_r[big], nl_max / gb);
    runs.push_back({big, next_r[big], k});
    next_r[big] += k;
    const int free_lines = nl_max - k * gb;
    if (free_lines > 0) {
      // first group in size order with at most free_lines lines
      size_t lo = head, hi = order.size();
      while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (group_size[order[mid]] > free_lines) lo = mid + 1;
        else hi = mid;
      }
      for (size_t oi = lo; oi < order.size(); oi++) {
        const int gi = order[oi];
        if (gi == big || next_r[gi] >= n2) continue;
        const int k2 = std::min(n2 - next_r[gi], free_lines / group_size[gi]);
        runs.push_back({gi, next_r[gi], k2});
        next_r[gi] += k2;
        break;
      }
    }
    bundles.push_back(runs);
  }
  return true;
}

}  // namespace qdp
