// kernels_fft_ct3.hip - instantiations of the compile-time FFT passes (fft_ct_impl.h)
// for lengths 192.
#include "fft_ct_impl.h"

namespace qd {

QD_FFT_CT_DEFINE(192)

}  // namespace qd
