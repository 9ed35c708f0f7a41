// qd_dispatch.hpp -- kernel selection for the compile-time specialised
// passes.  Each getter returns the address of one kernel instantiation (or
// nullptr when no specialisation fits) and is defined in qd_dispatch.hip,
// which is compiled in several parts so the instantiations build in parallel.
#pragma once

#include <stddef.h>

namespace qd {

// lengths with compile-time Stockham kernels (qd_fft.hpp)
bool ct_length(int n);
// Stockham R pass for (length, lines per mixing group, mixing on/off)
template <typename T>
const void *rpass_kernel(int n, int ml, bool mix);
// Stockham r pass for (length, lines per work group)
template <typename T>
const void *rline_kernel(int n, int lpb);
// register-FFT R pass (qd_rfft.hpp): lw lines per work group; *lds = dynamic LDS bytes
template <typename T>
const void *rpass_rf_kernel(int n, bool mix, int lw, size_t *lds);
// the same with K <-> l mixing staged in LDS (lw lines per work group)
template <typename T>
const void *rpass_rf_mix_kernel(int n, int lw, size_t *lds);
// its lengths above 96 (a separate build part)
template <typename T>
const void *rpass_rf_mix_kernel_long(int n, int lw, size_t *lds);
// register-FFT r pass: *lds = dynamic LDS bytes, *lpb = lines per work group
template <typename T>
const void *rline_rf_kernel(int n, int nth, size_t *lds, int *lpb);
// FP32 mixing R pass with prefetch (k_rpass_mx) for line length n, nl lines
// per bundle and unit sizes up to gb, sized for one or (two) two resident work
// groups per CU; *lds = its dynamic LDS bytes.  nullptr when no instantiation
// exists.
const void *rpass_mx_kernel(int n, int nl, int gb, bool two, size_t *lds);
const void *rpass_mx_kernel_small(int n, int nl, int gb, bool two, size_t *lds);
// matrix-stationary mixed theta GEMM for ks k-steps, V-phase prologue pro, dump epilogue epi
// (epi 2: FBR -> DVR times conj(exp(-i V dt)), the DVR state from the fused step's FBR state)
const void *theta_mx_kernel(int ks, int pro, int epi);
// fused FBR -> DVR -> FBR theta kernel (k_theta_fused) for ks k-steps and nth
// threads per work group (512); *lds = dynamic LDS bytes
const void *theta_fused_kernel(int ks, int nth, size_t *lds);
const void *theta_fused_par_kernel(int ksh, int nth, size_t *lds);
const void *theta_fused_pair_kernel(int nth, bool psync, size_t *lds);

}  // namespace qd
