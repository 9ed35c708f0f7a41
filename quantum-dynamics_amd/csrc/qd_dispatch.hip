// qd_dispatch.hip -- definitions of the getters declared in qd_dispatch.hpp.
//
// The file is compiled once per part (-DQD_PART=n), each part instantiating
// one family of kernels, so that the template instantiations -- the bulk of
// the build -- compile in parallel:
//   0  Stockham passes, FP32      1  Stockham passes, FP64
//   2  register-FFT passes, FP32  3  register-FFT passes, FP64 (no mixing)
//   4  mixed theta GEMM
//   5  mixing R pass with prefetch, 256-point lines
//   6  mixing R pass with prefetch, 64- and 128-point lines
//   7  register-FFT R pass with LDS-staged mixing, FP32, lengths up to 96
//   8  register-FFT R pass with LDS-staged mixing, FP64, lengths up to 96
//   9  register-FFT R pass with LDS-staged mixing, FP32, lengths 128..512
//  10  register-FFT R pass with LDS-staged mixing, FP64, lengths 128..512
// Every kernel instantiation lives in exactly one part.
#include <hip/hip_runtime.h>

#include "qd_dispatch.hpp"
#include "qd_fft.hpp"
#include "qd_kernels.hpp"
#include "qd_rfft.hpp"

#ifndef QD_PART
#error "compile with -DQD_PART=<0..10>"
#endif

namespace qd {

#define QD_CT_LENGTHS(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

#if QD_PART == 0
bool ct_length(int n) {
  switch (n) {
#define QD_CASE(N) case N:
    QD_CT_LENGTHS(QD_CASE)
#undef QD_CASE
    return true;
    default:
      return false;
  }
}

#endif

#if QD_PART == 0 || QD_PART == 1
template <typename T, int N, int ML>
const void *rpass_ct_fn(bool mix) {
  if constexpr (256 / ML * FFT_BPT >= ct_maxnb(N))
    return mix ? (const void *)k_rpass_ct<T, N, ML, true> : (const void *)k_rpass_ct<T, N, ML, false>;
  else return nullptr;
}
template <typename T, int N, int LPB>
const void *rline_ct_fn() {
  if constexpr (256 / LPB * FFT_BPT >= ct_maxnb(N)) return (const void *)k_rline_ct<T, N, LPB>;
  else return nullptr;
}
// kernel for (length, lines per group); nullptr -> generic kernel
template <typename T>
const void *rpass_kernel(int n, int ml, bool mix) {
  switch (n) {
#define QD_CASE(N) \
  case N:          \
    return ml == 8 ? rpass_ct_fn<T, N, 8>(mix) : ml == 32 ? rpass_ct_fn<T, N, 32>(mix) : nullptr;
    QD_CT_LENGTHS(QD_CASE)
#undef QD_CASE
    default:
      return nullptr;
  }
}
template <typename T>
const void *rline_kernel(int n, int lpb) {
  switch (n) {
#define QD_CASE(N) \
  case N:          \
    return lpb == 16 ? rline_ct_fn<T, N, 16>() : lpb == 8 ? rline_ct_fn<T, N, 8>() : nullptr;
    QD_CT_LENGTHS(QD_CASE)
#undef QD_CASE
    default:
      return nullptr;
  }
}

#if QD_PART == 0
template const void *rpass_kernel<float>(int, int, bool);
template const void *rline_kernel<float>(int, int);
#else
template const void *rpass_kernel<double>(int, int, bool);
template const void *rline_kernel<double>(int, int);
#endif
#endif  // Stockham parts

#define QD_RF_LENGTHS(X) X(24, 8) X(32, 8) X(48, 8) X(64, 8) X(96, 16) X(128, 16) X(192, 16) X(256, 16) X(384, 32) X(512, 32)

#if QD_PART == 2 || QD_PART == 3
// register-FFT passes: lengths with a factorisation N1 x N2 (rf_n1)
template <typename T, int N1, int N2>
const void *rpass_rf_fn(int lw, size_t *lds) {
  constexpr int LW0 = 256 / N1;
  if (lw != LW0) return nullptr;
  *lds = rpass_rf_lds<T, N1, N2, false, LW0>();
  return (const void *)k_rpass_rf<T, N1, N2, false, LW0>;
}
template <typename T>
const void *rpass_rf_kernel(int n, bool mix, int lw, size_t *lds) {
  if (mix) return rpass_rf_mix_kernel<T>(n, lw, lds);
  switch (n) {
#define QD_CASE(N, N1) \
  case N:              \
    return rpass_rf_fn<T, N1, N / N1>(lw, lds);
    QD_RF_LENGTHS(QD_CASE)
#undef QD_CASE
    default:
      return nullptr;
  }
}
// nth: threads per work group, 256 or 512 (batches of nth / N1 consecutive R points)
template <typename T>
const void *rline_rf_kernel(int n, int nth, size_t *lds, int *lpb) {
  switch (n) {
#define QD_CASE(N, N1)                                                   \
  case N:                                                                \
    *lpb = nth / N1;                                                     \
    if (nth == 256) {                                                    \
      *lds = rline_rf_lds<T, N1, N / N1, 256>();                         \
      return (const void *)k_rline_rf<T, N1, N / N1, 256>;               \
    }                                                                    \
    if (nth == 512) {                                                    \
      *lds = rline_rf_lds<T, N1, N / N1, 512>();                         \
      return (const void *)k_rline_rf<T, N1, N / N1, 512>;               \
    }                                                                    \
    return nullptr;
    QD_RF_LENGTHS(QD_CASE)
#undef QD_CASE
    default:
      return nullptr;
  }
}
#if QD_PART == 2
template const void *rpass_rf_kernel<float>(int, bool, int, size_t *);
template const void *rline_rf_kernel<float>(int, int, size_t *, int *);
#else
template const void *rpass_rf_kernel<double>(int, bool, int, size_t *);
template const void *rline_rf_kernel<double>(int, int, size_t *, int *);
#endif
#endif  // register-FFT parts

#if QD_PART >= 7 && QD_PART <= 10
// register-FFT R pass with K <-> l mixing staged in LDS (lw lines per work group);
// parts 7 / 8 (FP32 / FP64) hold lengths up to 96, parts 9 / 10 the longer ones
template <typename T, int N1, int N2>
const void *rpass_rf_mix_fn(int lw, size_t *lds) {
  constexpr int LW0 = 256 / N1;
  if (lw == LW0) {
    *lds = rpass_rf_lds<T, N1, N2, true, LW0>();
    return (const void *)k_rpass_rf<T, N1, N2, true, LW0>;
  }
  if (lw == 2 * LW0) {  // 512 threads: mixing groups of up to 2 * LW0 lines
    *lds = rpass_rf_lds<T, N1, N2, true, 2 * LW0>();
    return (const void *)k_rpass_rf<T, N1, N2, true, 2 * LW0>;
  }
  return nullptr;
}
#define QD_RF_LENGTHS_LO(X) X(24, 8) X(32, 8) X(48, 8) X(64, 8) X(96, 16)
#define QD_RF_LENGTHS_HI(X) X(128, 16) X(192, 16) X(256, 16) X(384, 32) X(512, 32)
#define QD_CASE(N, N1) \
  case N:              \
    return rpass_rf_mix_fn<T, N1, N / N1>(lw, lds);
#if QD_PART == 7 || QD_PART == 8
template <typename T>
const void *rpass_rf_mix_kernel(int n, int lw, size_t *lds) {
  switch (n) {
    QD_RF_LENGTHS_LO(QD_CASE)
    default:
      return rpass_rf_mix_kernel_long<T>(n, lw, lds);
  }
}
#if QD_PART == 7
template const void *rpass_rf_mix_kernel<float>(int, int, size_t *);
#else
template const void *rpass_rf_mix_kernel<double>(int, int, size_t *);
#endif
#else
template <typename T>
const void *rpass_rf_mix_kernel_long(int n, int lw, size_t *lds) {
  switch (n) {
    QD_RF_LENGTHS_HI(QD_CASE)
    default:
      return nullptr;
  }
}
#if QD_PART == 9
template const void *rpass_rf_mix_kernel_long<float>(int, int, size_t *);
#else
template const void *rpass_rf_mix_kernel_long<double>(int, int, size_t *);
#endif
#endif
#undef QD_CASE
#endif  // register-FFT mixing parts

#if QD_PART == 4
// matrix-stationary mixed theta kernel for KS k-steps (KP = NP = 32 KS)
const void *theta_mx_kernel(int ks, int pro, int epi) {
#define QD_TMX(KS_)                                                                    \
  if (ks == KS_) {                                                                     \
    if (pro == 0 && epi == 0) return reinterpret_cast<const void *>(&k_theta_mx<KS_, 2 * KS_, 0, 0>); \
    if (pro == 1 && epi == 0) return reinterpret_cast<const void *>(&k_theta_mx<KS_, 2 * KS_, 1, 0>); \
    if (pro == 2 && epi == 0) return reinterpret_cast<const void *>(&k_theta_mx<KS_, 2 * KS_, 2, 0>); \
    if (pro == 0 && epi == 1) return reinterpret_cast<const void *>(&k_theta_mx<KS_, 2 * KS_, 0, 1>); \
    if (pro == 0 && epi == 2) return reinterpret_cast<const void *>(&k_theta_mx<KS_, 2 * KS_, 0, 2>); \
  }
  QD_TMX(1)
  QD_TMX(2)
  QD_TMX(3)
  QD_TMX(4)
#undef QD_TMX
  return nullptr;
}

// fused FBR -> DVR -> FBR kernel for KS k-steps (both matrices in LDS)
const void *theta_fused_kernel(int ks, int nth, size_t *lds) {
#define QD_TFU(KS_, NTH_)                                                \
  if (ks == KS_ && nth == NTH_) {                                        \
    *lds = theta_fused_lds<KS_, NTH_>();                                 \
    return reinterpret_cast<const void *>(&k_theta_fused<KS_, NTH_>);    \
  }
  // (768 threads, three waves per SIMD, fits the LDS but not the 168-register
  // budget: it spills and measured 3.35 ms against 2.74 ms on the batched step)
  QD_TFU(1, 512)
  QD_TFU(2, 512)
#undef QD_TFU
  return nullptr;
}

// fused kernel on the parity halves: half bases of at most 32 (ksh 1) or 64
// (ksh 2, the tile one half at a time) functions
const void *theta_fused_par_kernel(int ksh, int nth, size_t *lds) {
#define QD_TFP(KSH_, NTH_)                                                  \
  if (ksh == KSH_ && nth == NTH_) {                                         \
    *lds = theta_fused_par_lds<KSH_, NTH_>();                               \
    return reinterpret_cast<const void *>(&k_theta_fused_par<KSH_, NTH_>);  \
  }
  QD_TFP(1, 512)
  QD_TFP(1, 768)
  QD_TFP(2, 256)
#undef QD_TFP
  return nullptr;
}

// fused kernel on the parity halves with each tile shared by a pair of waves
// (one per component): half bases of 33..64
const void *theta_fused_pair_kernel(int nth, bool psync, size_t *lds) {
  if (nth != 512) return nullptr;
  *lds = theta_fused_pair_lds<512>();
  return psync ? reinterpret_cast<const void *>(&k_theta_fused_pair<512, true>)
               : reinterpret_cast<const void *>(&k_theta_fused_pair<512, false>);
}
#endif  // theta part

#if QD_PART == 5 || QD_PART == 6
template <int N1, int N2, int NL, int GB>
const void *rpass_mx_fn(bool two, size_t *lds) {
  *lds = rpass_mx_lds<N1, N2, NL>();
  if constexpr (NL == 32)
    if (two) return (const void *)k_rpass_mx<N1, N2, NL, GB, true>;
  return (const void *)k_rpass_mx<N1, N2, NL, GB, false>;
}
template <int N1, int N2>
const void *rpass_mx_len(int nl, int gb, bool two, size_t *lds) {
  if (nl == 32) {
    if (gb == 8) return rpass_mx_fn<N1, N2, 32, 8>(two, lds);
    if (gb == 16) return rpass_mx_fn<N1, N2, 32, 16>(two, lds);
    if (gb == 24) return rpass_mx_fn<N1, N2, 32, 24>(two, lds);
    if (gb == 32) return rpass_mx_fn<N1, N2, 32, 32>(two, lds);
  } else if (nl == 48 && gb == 48 && !two) {
    return rpass_mx_fn<N1, N2, 48, 48>(false, lds);
  }
  return nullptr;
}
#endif
#if QD_PART == 5
const void *rpass_mx_kernel(int n, int nl, int gb, bool two, size_t *lds) {
  if (n == 256) return rpass_mx_len<16, 16>(nl, gb, two, lds);
  return rpass_mx_kernel_small(n, nl, gb, two, lds);
}
#endif
#if QD_PART == 6
const void *rpass_mx_kernel_small(int n, int nl, int gb, bool two, size_t *lds) {
  if (n == 64) return rpass_mx_len<8, 8>(nl, gb, two, lds);
  if (n == 128) return rpass_mx_len<16, 8>(nl, gb, two, lds);
  return nullptr;
}
#endif

}  // namespace qd
