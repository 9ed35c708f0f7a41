// CPU harness over the work records of the mixing R pass and the fused theta
// kernel (qd_records.hpp) for one batched workload.  Arguments:
//   n1 n2 jmax coriolis J...
// Builds the batch's FBR columns, mixing groups and blocks as the library
// does for (J, K) blocks K = 0..min(J, jmax) with jmax - K + 1 theta points,
// builds the records, runs the library's own bounds checks, and then checks
// independently, by walking every index the kernels form:
//  * every line of the FBR array (column, r) lies in exactly one bundle;
//  * every unit-matrix element a unit reads (16-column parts for units above
//    32 lines) lies inside the matrix array, and in its own matrix for the
//    elements it keeps;
//  * for several tiles-per-item settings, every (block, 16-row tile) lies in
//    exactly one fused-theta work item, whose blocks share matrices, sizes
//    and V.
// Prints "ok <bundles32> <bundles48> <items>" or "fail <reason>".
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "../../quantum-dynamics_amd/csrc/qd_records.hpp"

static int fail(const char *what, long long a = -1, long long b = -1) {
  printf("fail %s %lld %lld\n", what, a, b);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 6) return 2;
  const int n1 = atoi(argv[1]), n2 = atoi(argv[2]), jmax = atoi(argv[3]), coriolis = atoi(argv[4]);
  std::vector<int> Js;
  for (int i = 5; i < argc; i++) Js.push_back(atoi(argv[i]));
  // blocks and FBR columns: block (J, K) holds columns j = K..jmax
  struct Blk { int J, K, n, fcol0, dcol0; };
  std::vector<Blk> blocks;
  std::map<std::pair<int, int>, int> fcol0;
  int fc = 0;
  for (int J : Js)
    for (int K = 0; K <= std::min(J, jmax); K++) {
      const int n = jmax - K + 1;
      blocks.push_back({J, K, n, fc, fc});
      fcol0[{J, K}] = fc;
      fc += n;
    }
  const long long nfcol = fc, n12 = (long long)n1 * n2, nlp = nfcol * n12;
  // mixing groups (J, j): columns of K = 0..min(J, j); the D_R row of each
  // column is the index of its l(l+1) (Coriolis, J + j + l even) or of the
  // centrifugal-sudden value among the distinct values
  std::vector<std::vector<int>> cols;
  std::vector<int> u0;
  std::vector<double> umat;
  std::vector<long long> lamval(nfcol, 0);
  for (int J : Js)
    for (int j = 0; j <= jmax; j++) {
      const int g = std::min(J, j) + 1;
      std::vector<int> c;
      for (int K = 0; K < g; K++) c.push_back(fcol0[{J, K}] + (j - K));
      if (coriolis && g > 1) {
        u0.push_back((int)umat.size());
        for (int a = 0; a < g; a++)
          for (int b = 0; b < g; b++) umat.push_back(a == b ? 1.0 : 0.0);
        const int l0 = std::abs(J - j);
        for (int m = 0; m < g; m++) lamval[c[m]] = (long long)(l0 + 2 * m) * (l0 + 2 * m + 1);
        cols.push_back(c);
      } else {
        for (int K = 0; K < g; K++) {
          lamval[c[K]] = (long long)J * (J + 1) + (long long)j * (j + 1) - 2LL * K * K;
          cols.push_back({c[K]});
          u0.push_back(-1);
        }
      }
    }
  std::map<long long, int> lam_index;
  std::vector<int> fcol_lam(nfcol);
  for (long long c = 0; c < nfcol; c++) {
    auto it = lam_index.emplace(lamval[c], (int)lam_index.size()).first;
    fcol_lam[c] = it->second;
  }
  const int nlam = (int)lam_index.size();

  std::vector<float> upad;
  qd::MxRecords rec[2];
  if (!qd::build_mx_records(cols, u0, umat, fcol_lam, n1, n2, upad, rec)) return fail("packing refused");
  // QD_PLANT: corrupt one record on purpose (the tests check that the checks catch it)
  const char *plant_env = getenv("QD_PLANT");
  const int plant = plant_env ? atoi(plant_env) : 0;
  if (plant == 1) rec[0].loff[5] = nlp;                                  // a line past the FBR array
  if (plant == 2) rec[0].uni[0] = (rec[0].uni[0] & ~0xff00) | (40 << 8);  // a unit larger than its bucket
  if (plant == 3) rec[0].llam[0] = nlam;                                 // a D_R row past the table
  for (int c = 0; c < 2; c++) {
    const std::string err = qd::check_mx_records(rec[c], (size_t)nlp, n1, nlam, upad.size());
    if (!err.empty()) return fail(err.c_str());
  }
  // independent walk: lines exactly once, matrix reads inside the array
  std::vector<unsigned char> seen((size_t)nfcol * n2, 0);
  for (int c = 0; c < 2; c++) {
    const qd::MxRecords &R = rec[c];
    for (size_t b = 0; b < R.heads.size(); b++) {
      const qd::MxHead &h = R.heads[b];
      for (int L = 0; L < h.nl; L++) {
        const long long off = R.loff[b * R.nl + L];
        if (off < 0 || off + n1 > nlp) return fail("line offset", (long long)b, off);
        if (seen[off / n1]++) return fail("line in two bundles", (long long)b, off / n1);
      }
      for (int u = 0; u < h.nu; u++) {
        const int w = R.uni[b * R.nl + u], line0 = w & 255, g = (w >> 8) & 255, slot = w >> 16;
        if (!slot) continue;
        const long long base = h.msrc[slot - 1], g4 = h.mg4[slot - 1];
        // K -> l reads U rows k < g, columns < width; l -> K the same of U^T
        const long long width = R.gb > 32 ? 48 : g4;
        for (int part = 0; part < 2; part++) {
          const long long m0 = base + part * g4 * g4;
          const long long last = m0 + (long long)(g - 1) * g4 + width - 1;
          if (base < 0 || last >= (long long)upad.size()) return fail("matrix read outside", (long long)b, last);
          // the kept outputs (columns < g) stay inside this matrix
          if (g > g4 || m0 + (long long)(g - 1) * g4 + g - 1 >= m0 + g4 * g4) return fail("kept column outside", b, u);
        }
        if (line0 + g > h.nl) return fail("unit lines outside the bundle", (long long)b, u);
      }
    }
  }
  for (size_t i = 0; i < seen.size(); i++)
    if (seen[i] != 1) return fail("line not in any bundle", (long long)i);

  // fused theta items: blocks in K order (stable), groups of equal matrices, sizes and V
  std::vector<int> order(blocks.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return blocks[a].K < blocks[b].K; });
  std::map<int, long long> vset;  // V columns per K
  long long nV = 0;
  for (const Blk &b : blocks)
    if (!vset.count(b.K)) {
      vset[b.K] = nV;
      nV += (long long)b.n * n12;
    }
  std::vector<qd::TfuBlock> tb;
  for (int i : order) {
    const Blk &b = blocks[i];
    const int mat = b.J == 0 ? -1 : b.K;  // J = 0 uses the reference Legendre matrices
    tb.push_back({b.n, b.n, b.dcol0, b.fcol0, mat, vset[b.K]});
  }
  std::vector<std::pair<int, int>> groups;
  for (size_t i = 0; i < tb.size();) {
    size_t j = i + 1;
    while (j < tb.size() && tb[j].mat == tb[i].mat && tb[j].nd == tb[i].nd && tb[j].nf == tb[i].nf &&
           tb[j].vidx == tb[i].vidx)
      j++;
    groups.push_back({(int)i, (int)(j - i)});
    i = j;
  }
  const long long ntiles = (n12 + 15) / 16;
  size_t nitems = 0;
  // tiles side by side per work group: 4 (256 threads, or 512 as wave pairs), 8, 12 (768 threads)
  for (int lanes : {4, 8, 12})
  for (int tpw : {1, 9, 64, 150, 1000, 100000}) {
    std::vector<qd::TfuItem> items = qd::build_tfu_items(groups, ntiles, tpw, lanes);
    if (plant == 4) items.back().t1 += 1;                                // a tile past the grid
    if (plant == 5 && items.size() > 1) items.pop_back();                // tiles left uncovered
    const std::string err = qd::check_tfu_items(items, tb, n12, nfcol, nfcol, nV);
    if (!err.empty()) return fail(err.c_str(), tpw);
    std::vector<unsigned char> cov(tb.size() * (size_t)ntiles, 0);
    for (const qd::TfuItem &it : items)
      for (int k = 0; k < it.gnb; k++)
        for (int t = it.t0; t < it.t1; t++)
          if (cov[(size_t)(it.gb0 + k) * ntiles + t]++) return fail("tile in two items", tpw, t);
    for (size_t i = 0; i < cov.size(); i++)
      if (cov[i] != 1) return fail("tile in no item", tpw, (long long)i);
    nitems = items.size();
  }
  printf("ok %zu %zu %zu\n", rec[0].heads.size(), rec[1].heads.size(), nitems);
  return 0;
}
