// kernels_fft_ct4.hip - instantiations of the compile-time FFT passes (fft_ct_impl.h)
// for lengths 256.
#include "fft_ct_impl.h"

namespace qd {

QD_FFT_CT_DEFINE(256)

}  // namespace qd
