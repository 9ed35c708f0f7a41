// CPU harness over the host-side angular tables (qd_tables.hpp): for each
// (J, j) on the command line prints g, the largest |U^T U - I| entry, and per
// rounding (plain nearest, norm-preserving) the largest column term
// |sum_K U[K][m] (Q[K][m] - U[K][m])| and the largest |Q - U| in units of the
// float spacing.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../quantum-dynamics_amd/csrc/qd_tables.hpp"

int main(int argc, char **argv) {
  for (int a = 1; a + 1 < argc; a += 2) {
    const int J = atoi(argv[a]), j = atoi(argv[a + 1]);
    std::vector<double> U;
    std::vector<int> l;
    qdt::coriolis_eigen(J, j, U, l);
    const int g = (int)l.size();
    double orth = 0;
    for (int m = 0; m < g; m++)
      for (int n = 0; n < g; n++) {
        double s = 0;
        for (int K = 0; K < g; K++) s += U[(size_t)K * g + m] * U[(size_t)K * g + n];
        orth = std::max(orth, std::fabs(s - (m == n ? 1.0 : 0.0)));
      }
    std::vector<float> plain(U.begin(), U.end()), fixed;
    qdt::round_orthogonal_columns(U.data(), g, fixed);
    double col[2] = {0, 0}, ulp[2] = {0, 0};
    for (int v = 0; v < 2; v++) {
      const std::vector<float> &Q = v ? fixed : plain;
      for (int m = 0; m < g; m++) {
        double s = 0;
        for (int K = 0; K < g; K++) {
          const double u = U[(size_t)K * g + m], q = Q[(size_t)K * g + m];
          s += u * (q - u);
          const double sp = std::fabs((double)std::nextafter((float)u, INFINITY) - (double)(float)u);
          if (sp > 0) ulp[v] = std::max(ulp[v], std::fabs(q - u) / sp);
        }
        col[v] = std::max(col[v], std::fabs(s));
      }
    }
    printf("%d %d %d %.3e %.3e %.3e %.3f %.3f\n", J, j, g, orth, col[0], col[1], ulp[0], ulp[1]);
  }
  return 0;
}
