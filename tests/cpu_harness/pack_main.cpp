// CPU harness over the R pass's bundle packing (qd_pack.hpp).  Arguments:
// J_max j_max n2 nl_max.  Builds the mixing groups of J = 0..J_max with every
// K block (Coriolis-coupled: one group of min(J, j) + 1 lines per (J, j); J = 0
// and j = 0 give single lines), packs them, checks that every unit (group, r)
// lands in exactly one bundle, that no bundle exceeds nl_max lines or holds more
// than two groups, and that each group's units stay in r order; prints
// "ok <bundles> <lines> <fill>" or "fail <reason>".
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../quantum-dynamics_amd/csrc/qd_pack.hpp"

int main(int argc, char **argv) {
  if (argc != 5) return 2;
  const int Jmax = atoi(argv[1]), jmax = atoi(argv[2]), n2 = atoi(argv[3]), nl = atoi(argv[4]);
  std::vector<int> sizes;
  for (int J = 0; J <= Jmax; J++)
    for (int j = 0; j <= jmax; j++) {
      const int g = std::min(J, j) + 1;
      if (g > 1) sizes.push_back(g);
      else sizes.push_back(1);
    }
  std::vector<std::vector<qdp::Run>> plan;
  if (!qdp::pack_bundles(sizes, n2, nl, plan)) {
    printf("fail pack_bundles refused\n");
    return 0;
  }
  std::vector<int> next(sizes.size(), 0);
  long long lines = 0;
  for (const auto &runs : plan) {
    int nlb = 0;
    if (runs.empty() || runs.size() > 2) {
      printf("fail bundle with %d groups\n", (int)runs.size());
      return 0;
    }
    for (const auto &r : runs) {
      if (r.count < 1 || r.r0 != next[r.group]) {
        printf("fail units of group %d out of order\n", r.group);
        return 0;
      }
      next[r.group] += r.count;
      nlb += r.count * sizes[r.group];
    }
    if (runs.size() == 2 && runs[0].group == runs[1].group) {
      printf("fail one group twice in a bundle\n");
      return 0;
    }
    if (nlb > nl) {
      printf("fail bundle of %d lines\n", nlb);
      return 0;
    }
    lines += nlb;
  }
  for (size_t i = 0; i < sizes.size(); i++)
    if (next[i] != n2) {
      printf("fail group %d has %d of %d units\n", (int)i, next[i], n2);
      return 0;
    }
  printf("ok %d %lld %.4f\n", (int)plan.size(), lines, (double)lines / ((double)plan.size() * nl));
  return 0;
}
