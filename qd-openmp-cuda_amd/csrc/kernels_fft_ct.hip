// kernels_fft_ct.hip - instantiations of the compile-time FFT passes (fft_ct_impl.h)
// for lengths 32, 48, 64.
#include <cstdlib>

#include "fft_ct_impl.h"

namespace qd {

QD_FFT_CT_DEFINE(32)
QD_FFT_CT_DEFINE(48)
QD_FFT_CT_DEFINE(64)

void launch_fft_R_ct_32(const StepArgs &a, int units, int inverse, int acf, int nb_max, hipStream_t s);
void launch_fft_r_ct_32(const StepArgs &a, int units, hipStream_t s);
void launch_fft_R_ct_48(const StepArgs &a, int units, int inverse, int acf, int nb_max, hipStream_t s);
void launch_fft_r_ct_48(const StepArgs &a, int units, hipStream_t s);
void launch_fft_R_ct_64(const StepArgs &a, int units, int inverse, int acf, int nb_max, hipStream_t s);
void launch_fft_r_ct_64(const StepArgs &a, int units, hipStream_t s);
void launch_fft_R_ct_96(const StepArgs &a, int units, int inverse, int acf, int nb_max, hipStream_t s);
void launch_fft_r_ct_96(const StepArgs &a, int units, hipStream_t s);
void launch_fft_R_ct_128(const StepArgs &a, int units, int inverse, int acf, int nb_max, hipStream_t s);
void launch_fft_r_ct_128(const StepArgs &a, int units, hipStream_t s);
void launch_fft_R_ct_192(const StepArgs &a, int units, int inverse, int acf, int nb_max, hipStream_t s);
void launch_fft_r_ct_192(const StepArgs &a, int units, hipStream_t s);
void launch_fft_R_ct_256(const StepArgs &a, int units, int inverse, int acf, int nb_max, hipStream_t s);
void launch_fft_r_ct_256(const StepArgs &a, int units, hipStream_t s);
void launch_fft_R_kin_32(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s);
void launch_fft_r_mix_32(const StepArgs &a, int units, int nb_max, hipStream_t s);
void launch_fft_R_kin_48(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s);
void launch_fft_r_mix_48(const StepArgs &a, int units, int nb_max, hipStream_t s);
void launch_fft_R_kin_64(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s);
void launch_fft_r_mix_64(const StepArgs &a, int units, int nb_max, hipStream_t s);
void launch_fft_R_kin_96(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s);
void launch_fft_r_mix_96(const StepArgs &a, int units, int nb_max, hipStream_t s);
void launch_fft_R_kin_128(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s);
void launch_fft_r_mix_128(const StepArgs &a, int units, int nb_max, hipStream_t s);
void launch_fft_R_kin_192(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s);
void launch_fft_r_mix_192(const StepArgs &a, int units, int nb_max, hipStream_t s);
void launch_fft_R_kin_256(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s);
void launch_fft_r_mix_256(const StepArgs &a, int units, int nb_max, hipStream_t s);

bool fft_ct_supported(int n) {
  switch (n) {
    case 32: case 48: case 64: case 96: case 128: case 192: case 256: return true;
    default: return false;
  }
}
int fft_ct_rows(int n1, int n_blocks, bool inverse) { return ct_rows(n1, n_blocks, inverse); }
int fft_ct_cols(int n2) { return ct_cols(n2); }

// QD_FFT_TWO_PASS=0 keeps the three-pass form (read per job; jobs of either
// form never share an engine or a server class).
int fft_mix_cols(int n2, int n_blocks) {
  const char *e = std::getenv("QD_FFT_TWO_PASS");
  const bool on = !(e && std::atoi(e) == 0);
  return on && fft_ct_supported(n2) ? mix_cols(n2, n_blocks) : 0;
}

void launch_fft_R_ct(const StepArgs &a, int total_units, int inverse, int acf, int n1, int nb_max, hipStream_t s) {
  switch (n1) {
#define QD_R_CASE(N) case N: launch_fft_R_ct_##N(a, total_units, inverse, acf, nb_max, s); break;
    QD_R_CASE(32) QD_R_CASE(48) QD_R_CASE(64) QD_R_CASE(96) QD_R_CASE(128) QD_R_CASE(192) QD_R_CASE(256)
#undef QD_R_CASE
    default: break;
  }
}

void launch_fft_r_ct(const StepArgs &a, int total_units, int n2, hipStream_t s) {
  switch (n2) {
#define QD_r_CASE(N) case N: launch_fft_r_ct_##N(a, total_units, s); break;
    QD_r_CASE(32) QD_r_CASE(48) QD_r_CASE(64) QD_r_CASE(96) QD_r_CASE(128) QD_r_CASE(192) QD_r_CASE(256)
#undef QD_r_CASE
    default: break;
  }
}

void launch_fft_r_mix(const StepArgs &a, int total_units, int n2, int nb_max, hipStream_t s) {
  switch (n2) {
#define QD_m_CASE(N) case N: launch_fft_r_mix_##N(a, total_units, nb_max, s); break;
    QD_m_CASE(32) QD_m_CASE(48) QD_m_CASE(64) QD_m_CASE(96) QD_m_CASE(128) QD_m_CASE(192) QD_m_CASE(256)
#undef QD_m_CASE
    default: break;
  }
}

void launch_fft_R_kin(const StepArgs &a, int total_units, int acf, int n1, int nb_max, hipStream_t s) {
  switch (n1) {
#define QD_k_CASE(N) case N: launch_fft_R_kin_##N(a, total_units, acf, nb_max, s); break;
    QD_k_CASE(32) QD_k_CASE(48) QD_k_CASE(64) QD_k_CASE(96) QD_k_CASE(128) QD_k_CASE(192) QD_k_CASE(256)
#undef QD_k_CASE
    default: break;
  }
}

}  // namespace qd
