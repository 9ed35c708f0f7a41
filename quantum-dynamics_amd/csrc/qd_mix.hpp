// qd_mix.hpp -- the mixing launches of the split R pass: K <-> l mixing of
// whole R lines in a kernel of its own, for mixing groups of any size (the
// Coriolis coupling of the K blocks of one (J, j); the reference propagates
// J = 0 only, TimeEvolutionMex.cpp, and has no counterpart).
#pragma once
#include "qd_kernels.hpp"

namespace qd {

// K <-> l mixing of whole R lines, out of place, for the R pass done in three
// launches (mixing, the FFT pass without mixing, mixing back).  Each wave takes
// 64 points of one unit (wpu waves per unit), so the unit and its line offsets
// are wave-uniform.  K -> l (to_l): out_m = sum_K U[K][m] in_K; l -> K:
// out_K = sum_m U[K][m] in_m.  Tiles of 16 outputs x 16 inputs: the wave
// copies the tile's weights to its own LDS slice (zeros past the unit), then
// issues the 16 input loads together (indices clamped into the unit), so a
// unit of up to 16 lines costs one memory latency; larger units reread their
// inputs (from the caches) per 16 outputs.  Lanes past the line's end take
// part in the weight copies and store nothing.  Units without a matrix
// (u0 < 0) are copied.
template <typename T>
__global__ __launch_bounds__(256) void k_mix_lines(const typename cx_of<T>::type *__restrict__ in,
                                                   typename cx_of<T>::type *__restrict__ out,
                                                   const RUnit *__restrict__ units, const int *__restrict__ ucols,
                                                   const T *__restrict__ umat, int nunits, int n1, int n2, int wpu,
                                                   int to_l) {
  typedef typename cx_of<T>::type C;
  constexpr int CH = 16;
  __shared__ T s_w[4][CH * CH];
  const int lane = (int)threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)blockIdx.x * 4 + ((int)threadIdx.x >> 6));
  const int un = w / wpu;
  if (un >= nunits) return;  // wave-uniform
  const int i0 = (w - un * wpu) * 64 + lane;
  const bool act = i0 < n1;
  const int i = act ? i0 : n1 - 1;
  const RUnit u = units[un];
  const int g = u.g;
  auto line = [&](int k) { return ((size_t)ucols[u.col0 + k] * n2 + u.r) * n1; };
  if (u.u0 < 0) {
    if (act)
      for (int k = 0; k < g; k++) out[line(k) + i] = in[line(k) + i];
    return;
  }
  const T *U = umat + u.u0;
  T *tw = s_w[threadIdx.x >> 6];
  for (int m0 = 0; m0 < g; m0 += CH) {
    C acc[CH];
#pragma unroll
    for (int c = 0; c < CH; c++) acc[c].x = acc[c].y = (T)0;
    for (int k0 = 0; k0 < g; k0 += CH) {
      // the previous tile's reads are complete before the slice is rewritten
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int q = 0; q < CH * CH / 64; q++) {
        const int e = lane + 64 * q, k = k0 + e / CH, m = m0 + e % CH;
        tw[e] = (k < g && m < g) ? (to_l ? U[k * g + m] : U[m * g + k]) : (T)0;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // the tile's 16 line indices: one vector load, then broadcast per line
      const int colk = ucols[u.col0 + min(k0 + (lane & (CH - 1)), g - 1)];
      C x[CH];
#pragma unroll
      for (int kk = 0; kk < CH; kk++)
        x[kk] = in[((size_t)__builtin_amdgcn_readlane(colk, kk) * n2 + u.r) * n1 + i];
#pragma unroll
      for (int kk = 0; kk < CH; kk++) {
#pragma unroll
        for (int c = 0; c < CH; c++) {
          const T wt = tw[kk * CH + c];
          acc[c].x += wt * x[kk].x;
          acc[c].y += wt * x[kk].y;
        }
      }
    }
    const int colm = ucols[u.col0 + min(m0 + (lane & (CH - 1)), g - 1)];
    if (act) {
#pragma unroll
      for (int c = 0; c < CH; c++)
        if (m0 + c < g) out[((size_t)__builtin_amdgcn_readlane(colm, c) * n2 + u.r) * n1 + i] = acc[c];
    }
  }
}

}  // namespace qd
