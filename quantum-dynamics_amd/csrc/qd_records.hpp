// qd_records.hpp -- the work records of the batched step's two table-driven
// kernels, built on the host, and the bounds checks of every index those
// kernels form from them.  Header-only and free of HIP, so the CPU tests
// build the records of every workload and check them without a GPU; the
// library runs the same checks before it uploads the records and refuses a
// handle whose records would leave an array.
//
//   k_rpass_mx     (qd_rfft.hpp)   bundles of mixing units: per bundle a header
//                                  (MxHead), per line its element offset and
//                                  D_R row, per unit a packed word
//   k_theta_fused  (qd_kernels.hpp) work items: a group of blocks sharing the
//                                  theta matrices and V, and a range of 16-row
//                                  tiles
#pragma once

#include <algorithm>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "qd_pack.hpp"
#include "qd_tables.hpp"

namespace qd {

// Header of one bundle of the mixing R pass (read by the kernel as two int4).
struct MxHead {
  int nl, nu;    // lines and units of the bundle
  int msrc[2];   // offsets in umat of its matrices (U, then U^T at + g4^2)
  int mg4[2];    // their padded orders (multiples of 4; 0 = unused)
  int pad[2];
};

// floats of zero padding after the last unit matrix: the 16-column parts of
// the 48-line kernel read up to 12 floats past a row (into the next row, or
// into this padding after the last matrix)
constexpr int MX_UMAT_PAD = 64;

// One class of bundles: NL lines per bundle, unit sizes up to GB.
struct MxRecords {
  int nl = 0, gb = 0;
  std::vector<MxHead> heads;
  std::vector<long long> loff;  // [bundle][NL] element offset of each line in the FBR array
  std::vector<int> llam;        // [bundle][NL] D_R row of each line
  std::vector<int> uni;         // [bundle][NL] line0 | g << 8 | slot << 16 (slot 0: no mixing)
};

inline int mx_bucket(int gmax) { return gmax <= 8 ? 8 : gmax <= 16 ? 16 : gmax <= 24 ? 24 : gmax <= 32 ? 32 : 48; }

// The mixing groups of a batch: group i has the FBR columns cols[i] (one
// line per column at every r), and its g x g unit matrix at u0[i] in umat
// (row K, column m), or u0[i] < 0 for no mixing.  Groups of at most 32 lines
// go to 32-line bundles (rec[0]), larger ones to 48-line bundles (rec[1]).
// upad receives the matrices, each padded to g4 = 4 ceil(g / 4), U then U^T,
// rounded to FP32 with unit column lengths, then MX_UMAT_PAD zeros.
// Returns false when the groups cannot be packed.
inline bool build_mx_records(const std::vector<std::vector<int>> &cols, const std::vector<int> &u0,
                             const std::vector<double> &umat, const std::vector<int> &fcol_lam, int n1, int n2,
                             std::vector<float> &upad, MxRecords rec[2]) {
  upad.clear();
  std::map<int, std::pair<int, int>> mat_of;  // u0 -> (offset in upad, g4)
  for (size_t gi = 0; gi < cols.size(); gi++) {
    const int gs = (int)cols[gi].size();
    if (u0[gi] < 0 || gs < 2 || mat_of.count(u0[gi])) continue;
    const int g4 = (gs + 3) & ~3;
    std::vector<float> q;
    qdt::round_orthogonal_columns(umat.data() + u0[gi], gs, q);
    const int off = (int)upad.size();
    upad.resize(upad.size() + 2 * (size_t)g4 * g4, 0.f);
    for (int K = 0; K < gs; K++)
      for (int m = 0; m < gs; m++) {
        upad[off + (size_t)K * g4 + m] = q[(size_t)K * gs + m];                  // U[K][m]
        upad[off + (size_t)g4 * g4 + (size_t)m * g4 + K] = q[(size_t)K * gs + m];  // U^T[m][K]
      }
    mat_of[u0[gi]] = {off, g4};
  }
  upad.resize(upad.size() + MX_UMAT_PAD, 0.f);
  std::vector<int> sel[2];
  for (size_t i = 0; i < cols.size(); i++) sel[cols[i].size() > 32 ? 1 : 0].push_back((int)i);
  for (int c = 0; c < 2; c++) {
    MxRecords &R = rec[c];
    R = MxRecords{};
    if (sel[c].empty()) continue;
    int gm = 1;
    for (int gi : sel[c]) gm = std::max(gm, (int)cols[gi].size());
    R.nl = c == 0 ? 32 : 48;
    R.gb = mx_bucket(gm);
    const int NL = R.nl;
    MxHead cur{};
    // units r0 .. r0 + count - 1 of group gi into the open bundle
    auto add_units = [&](int gi, int r0, int count) {
      const int gs = (int)cols[gi].size();
      const bool mixed = u0[gi] >= 0 && gs > 1;
      int slot = 0;
      if (mixed) {
        const int q = cur.mg4[0] ? 1 : 0;
        cur.msrc[q] = mat_of[u0[gi]].first;
        cur.mg4[q] = mat_of[u0[gi]].second;
        slot = q + 1;
      }
      const size_t bb = R.heads.size() * (size_t)NL;
      for (int r = r0; r < r0 + count; r++) {
        R.uni[bb + cur.nu] = cur.nl | gs << 8 | slot << 16;
        for (int k = 0; k < gs; k++) {
          const int col = cols[gi][k];
          R.loff[bb + cur.nl + k] = ((long long)col * n2 + r) * n1;
          R.llam[bb + cur.nl + k] = fcol_lam[col];
        }
        cur.nl += gs;
        cur.nu += 1;
      }
    };
    std::vector<int> gsize(sel[c].size());
    for (size_t i = 0; i < sel[c].size(); i++) gsize[i] = (int)cols[sel[c][i]].size();
    std::vector<std::vector<qdp::Run>> plan;
    if (!qdp::pack_bundles(gsize, n2, NL, plan)) return false;
    for (const auto &runs : plan) {
      R.loff.resize(R.loff.size() + NL, 0);
      R.llam.resize(R.llam.size() + NL, 0);
      R.uni.resize(R.uni.size() + NL, 0);
      for (const qdp::Run &run : runs) add_units(sel[c][run.group], run.r0, run.count);
      if (cur.nl) R.heads.push_back(cur);
      cur = MxHead{};
    }
  }
  return true;
}

// Every index k_rpass_mx forms from one class of records, against the sizes
// of what it reads and writes: the FBR array (nlp complex values), the D_R
// table (nlam rows of n1) and the unit matrices (upad_size floats).  Also
// checks that the units of a bundle partition its lines (the mixing works in
// place, one unit per thread and point).  Returns "" when everything is in
// range, else the first violation.
inline std::string check_mx_records(const MxRecords &R, size_t nlp, int n1, int nlam, size_t upad_size) {
  char msg[256];
  auto bad = [&](const char *what, size_t b) {
    snprintf(msg, sizeof msg, "mixing R pass records (%d-line bundles), bundle %zu: %s", R.nl, b, what);
    return std::string(msg);
  };
  const int NL = R.nl;
  if (R.heads.empty()) return "";
  if (NL != 32 && NL != 48) return bad("lines per bundle is not 32 or 48", 0);
  if (R.gb % 4 || R.gb > NL || R.gb < 4) return bad("unit-size bucket out of range", 0);
  const size_t nb = R.heads.size();
  if (R.loff.size() != nb * NL || R.llam.size() != nb * NL || R.uni.size() != nb * NL)
    return bad("record arrays are not NL entries per bundle", 0);
  if (upad_size < (size_t)MX_UMAT_PAD) return bad("unit matrices lack their padding", 0);
  const size_t mat_end = upad_size - MX_UMAT_PAD;
  for (size_t b = 0; b < nb; b++) {
    const MxHead &h = R.heads[b];
    if (h.nl < 1 || h.nl > NL) return bad("line count out of range", b);
    if (h.nu < 1 || h.nu > h.nl) return bad("unit count out of range", b);
    for (int q = 0; q < 2; q++) {
      if (h.mg4[q] == 0) continue;
      if (h.mg4[q] % 4 || h.mg4[q] < 4) return bad("matrix order not a positive multiple of 4", b);
      if (h.msrc[q] < 0 || (size_t)h.msrc[q] + 2 * (size_t)h.mg4[q] * h.mg4[q] > mat_end)
        return bad("unit matrix outside the matrix array", b);
    }
    for (int L = 0; L < h.nl; L++) {
      const long long off = R.loff[b * NL + L];
      if (off < 0 || off % n1 || (size_t)off + (size_t)n1 > nlp) return bad("line outside the FBR array", b);
      const int lam = R.llam[b * NL + L];
      if (lam < 0 || lam >= nlam) return bad("D_R row out of range", b);
    }
    int next = 0;
    for (int u = 0; u < h.nu; u++) {
      const int w = R.uni[b * NL + u];
      const int line0 = w & 255, g = (w >> 8) & 255, slot = w >> 16;
      if (line0 != next) return bad("units do not partition the lines in order", b);
      if (g < 1 || line0 + g > h.nl) return bad("unit lines outside the bundle", b);
      if (g > R.gb || (R.gb <= 32 && g > 32)) return bad("unit larger than the kernel's bucket", b);
      if (slot < 0 || slot > 2) return bad("matrix slot out of range", b);
      if (slot) {
        const int g4 = h.mg4[slot - 1];
        if (g < 2 || g4 != ((g + 3) & ~3)) return bad("unit matrix order does not match the unit", b);
        // the 48-line kernel reads 16 columns at a time: up to column 47 of the
        // last row of U^T, past its end when g4 < 48
        if (R.gb > 32) {
          const size_t last = (size_t)h.msrc[slot - 1] + (size_t)g4 * g4 + (size_t)(g - 1) * g4 + 47;
          if (last >= upad_size) return bad("16-column reads leave the matrix array", b);
        }
      }
      next = line0 + g;
    }
    if (next != h.nl) return bad("units do not cover the bundle's lines", b);
  }
  return "";
}

// One work item of the fused theta kernel (read by the kernel as an int4):
// the blocks gb0 .. gb0 + gnb - 1 (in K order) of one group and the 16-row
// tiles t0 .. t1 - 1.
struct TfuItem {
  int gb0, gnb, t0, t1;
};

// Each group's tiles in ranges of max(W, ceil(tpw / blocks)) tiles, so that
// every item holds about tpw tile-blocks and at least one tile per wave.
inline std::vector<TfuItem> build_tfu_items(const std::vector<std::pair<int, int>> &groups, long long ntiles,
                                            int tpw, int W) {
  std::vector<TfuItem> items;
  for (const auto &g : groups) {
    const long long tc = std::min(ntiles, std::max((long long)W, ((long long)tpw + g.second - 1) / g.second));
    for (long long t = 0; t < ntiles; t += tc)
      items.push_back({g.first, g.second, (int)t, (int)std::min(ntiles, t + tc)});
  }
  return items;
}

// What the fused theta kernel needs to know of a block (in K order).
struct TfuBlock {
  int nd, nf, dcol0, fcol0, mat;
  long long vidx;  // first element of its V columns
};

// Every index k_theta_fused forms from its items: the blocks of an item share
// matrices, sizes and V; its tiles lie in the grid; each block's columns lie in
// the DVR / FBR arrays (ndcol / nfcol columns of n12 points) and V (nV
// values); the per-block byte counts of its buffer descriptors stay below 2^31.
// Also checks that each block gets as many tiles as it has.
inline std::string check_tfu_items(const std::vector<TfuItem> &items, const std::vector<TfuBlock> &blocks,
                                   long long n12, long long ndcol, long long nfcol, long long nV) {
  char msg[256];
  auto bad = [&](const char *what, size_t i) {
    snprintf(msg, sizeof msg, "fused theta work item %zu: %s", i, what);
    return std::string(msg);
  };
  const long long ntiles = (n12 + 15) / 16;
  const long long nb = (long long)blocks.size();
  std::vector<long long> covered(blocks.size(), 0);
  for (size_t i = 0; i < items.size(); i++) {
    const TfuItem &it = items[i];
    if (it.gb0 < 0 || it.gnb < 1 || (long long)it.gb0 + it.gnb > nb) return bad("blocks out of range", i);
    if (it.t0 < 0 || it.t0 >= it.t1 || it.t1 > ntiles) return bad("tiles out of range", i);
    const TfuBlock &b0 = blocks[it.gb0];
    for (int k = 0; k < it.gnb; k++) {
      const TfuBlock &b = blocks[it.gb0 + k];
      if (b.mat != b0.mat || b.nd != b0.nd || b.nf != b0.nf || b.vidx != b0.vidx)
        return bad("blocks of one item do not share matrices, sizes and V", i);
      if (b.nd < 1 || b.nf < 1 || b.dcol0 < 0 || b.fcol0 < 0 || b.dcol0 + b.nd > ndcol || b.fcol0 + b.nf > nfcol)
        return bad("block columns outside the DVR or FBR array", i);
      if (b.vidx < 0 || b.vidx + (long long)b.nd * n12 > nV) return bad("V columns outside V", i);
      // below 2^31: masked lanes use offsets from 2^31 up, which must read as
      // out of range
      if ((unsigned long long)b.nd * n12 * 8 >= (1ULL << 31) || (unsigned long long)b.nf * n12 * 8 >= (1ULL << 31))
        return bad("a block's byte count reaches 2^31, the kernels' out-of-range offset", i);
      covered[it.gb0 + k] += it.t1 - it.t0;
    }
  }
  for (long long b = 0; b < nb; b++)
    if (covered[b] != ntiles) return bad("a block does not get each of its tiles once", (size_t)b);
  return "";
}

}  // namespace qd
