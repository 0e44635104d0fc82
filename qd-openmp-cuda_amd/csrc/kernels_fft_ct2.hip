// kernels_fft_ct2.hip - instantiations of the compile-time FFT passes (fft_ct_impl.h)
// for lengths 96, 128.
#include "fft_ct_impl.h"

namespace qd {

QD_FFT_CT_DEFINE(96)
QD_FFT_CT_DEFINE(128)

}  // namespace qd
