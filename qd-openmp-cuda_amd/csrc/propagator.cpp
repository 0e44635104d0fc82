// propagator.cpp - input validation, device residency of a partial-wave job
// and the step engine.
//
// Reference counterparts:
//   validate_inputs ... the insist() checks spread over omegawavepacket.cu:46,
//                       :142-143, wavepackets1device.cu:276-284, CoriolisMatrix.C:65-71,
//                       cudaMath.h:119 (odd n), matlabStructures.C:19-22 (unchecked
//                       classes) and the norm check of cudaOpenMP2.cu:437-440
//   DeviceJob ......... OmegaWavepacketsOnSingleDevice::setup_data_on_device
//                       (wavepackets1device.cu:58-150) + OmegaWavepacket::setup_device_data
//                       (omegawavepacket.cu:66-198) + setup_coriolis_matrices_and_copy_to_device
//                       (cudaOpenMP2.cu:158-209) + setup_constant_memory_on_device
//                       (wavepackets1device.cu:265-287)
//   Engine::step ...... one iteration of CudaOpenMPQMMD::test_multiple_cards
//                       (cudaOpenMP2.cu:338-387)
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>

#include "host_internal.h"

namespace qd {

namespace {

constexpr size_t kLdsLimit = 160 * 1024;
constexpr int kRing = 1024;
constexpr int kFluxWindow = 64;   // steps buffered per energy transform

bool factors_235(int n) {
  if (n < 2) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }


size_t gemm_lds_bytes(int kp, int np, bool mixed) { return leg_lds_bytes(kp, np, mixed ? 1 : 0); }

// bfloat16 bits of a finite float, rounded to nearest even
uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Compensation table of the paired-row mixed kernels for B ([nO][kp][np]
// row-major FP64): per block, entry (4 c + g, n) holds 8 bfloat16,
// B_lo(k_q, n) for q < 4 then B_hi(k_q, n) for q < 4, k_q = 16 c + g + 4 q,
// B_hi = fp32(B), B_lo = fp32(B - B_hi); zero for k_q >= kp.
std::vector<uint4> bf16_compensation_table(const std::vector<double> &B, int nO, int kp, int np) {
  const int xrows = 4 * ((kp + 15) / 16);
  std::vector<uint4> out((size_t)nO * xrows * np);
  for (int b = 0; b < nO; b++)
    for (int r = 0; r < xrows; r++)
      for (int n = 0; n < np; n++) {
        uint16_t h[8];
        for (int q = 0; q < 4; q++) {
          const int k = 16 * (r / 4) + (r % 4) + 4 * q;
          const double v = k < kp ? B[((size_t)b * kp + k) * np + n] : 0.0;
          const float hi = (float)v, lo = (float)(v - (double)hi);
          h[q] = bf16_bits(lo);
          h[4 + q] = bf16_bits(hi);
        }
        uint4 e;
        e.x = h[0] | ((uint32_t)h[1] << 16);
        e.y = h[2] | ((uint32_t)h[3] << 16);
        e.z = h[4] | ((uint32_t)h[5] << 16);
        e.w = h[6] | ((uint32_t)h[7] << 16);
        out[((size_t)b * xrows + r) * np + n] = e;
      }
  return out;
}

constexpr size_t kLegLdsTwoGroups = 80 * 1024;
constexpr long kCUs = 256;   // compute units of an MI355X

int pick_divisor(int n, const int *cands, int ncand, int limit_elems, int other) {
  for (int q = 0; q < ncand; q++)
    if (n % cands[q] == 0 && cands[q] * other <= limit_elems) return cands[q];
  return 1;
}

}  // namespace

void init_attrs_fft();
void init_attrs_gemm();
void init_attrs_leg_step();
void init_attrs_leg_split();

void init_kernel_attributes() {
  // once per process, also when several host threads create handles at once
  static std::once_flag once;
  std::call_once(once, [] {
    init_attrs_fft();
    init_attrs_gemm();
    init_attrs_leg_step();
    init_attrs_leg_split();
    (void)hipGetLastError();   // an unsupported attribute must not leak into later error checks
  });
}

JobShape validate_inputs(const qd_inputs *in) {
  QD_REQUIRE(in != nullptr, "inputs pointer is NULL");
  JobShape s{};
  const qd_omega_states &os = in->omega_states;
  s.n1 = in->r1.n; s.n2 = in->r2.n; s.nt = in->theta.n;
  QD_REQUIRE(s.n1 > 0 && s.n2 > 0 && s.nt > 0, "grid sizes must be positive");
  QD_REQUIRE(s.n1 % 2 == 0 && s.n2 % 2 == 0,
             "r1.n and r2.n must be even (the kinetic table is undefined for odd n, cudaMath.h:119)");
  QD_REQUIRE(factors_235(s.n1) && factors_235(s.n2) && s.n1 <= kFftMaxN && s.n2 <= kFftMaxN,
             "r1.n and r2.n must factor into 2, 3, 5 and not exceed 512");
  QD_REQUIRE(std::isfinite(in->r1.dr) && in->r1.dr > 0 && std::isfinite(in->r2.dr) && in->r2.dr > 0,
             "r1.dr and r2.dr must be positive");
  QD_REQUIRE(in->r1.mass > 0 && in->r2.mass > 0, "masses must be positive");
  QD_REQUIRE(in->r1.left > 0 && in->r2.left > 0, "radial grids must start at r > 0 (1/r^2 terms)");
  QD_REQUIRE(in->theta.x && in->theta.w, "theta.x and theta.w are required");
  QD_REQUIRE(in->pot, "pot is required");
  QD_REQUIRE(in->time.steps, "time.steps is required (in/out)");
  QD_REQUIRE(in->time.total_steps >= 0, "time.total_steps must be >= 0");
  QD_REQUIRE(std::isfinite(in->time.time_step), "time.time_step must be finite");
  QD_REQUIRE(*in->time.steps >= 0, "time.steps must be >= 0");
  s.lmax = os.l_max; s.L = os.l_max + 1;
  QD_REQUIRE(os.l_max >= 0, "l_max must be >= 0");
  QD_REQUIRE(s.nt > s.lmax + 1, "theta.n must exceed l_max + 1 (omegawavepacket.cu:46)");
  s.nO = os.n_omegas;
  QD_REQUIRE(s.nO >= 1 && os.omegas, "at least one omega block is required");
  QD_REQUIRE(s.nO <= kMaxOmega, "at most 24 omega blocks per job");
  QD_REQUIRE(os.n_omegas_max >= s.nO, "n_omegas_max < number of omegas (matlabStructures.C:108)");
  s.J = os.J; s.parity = os.parity;
  QD_REQUIRE(s.J >= 0, "J must be >= 0");
  QD_REQUIRE(s.J != 0 || s.parity == 0, "J = 0 requires parity 0 (CoriolisMatrix.C:65)");
  QD_REQUIRE(s.parity == 0 || s.parity == 1, "parity must be 0 or 1");
  const int omin_rule = (s.J + s.parity) % 2 == 0 ? 0 : 1;
  s.omin = os.omegas[0];
  QD_REQUIRE(s.omin == (s.J == 0 ? 0 : omin_rule),
             "omegas[0] must be 0 when J + p is even, 1 when odd (OmegaList.m:4-16)");
  for (int b = 1; b < s.nO; b++)
    QD_REQUIRE(os.omegas[b] == s.omin + b, "omegas must be consecutive integers (omegawavepacket.cu:142)");
  const int kcap = std::min(s.J, s.lmax);
  s.kmax = in->k_max >= 0 ? std::min(kcap, (int)in->k_max) : kcap;
  QD_REQUIRE(s.omin + s.nO - 1 == s.kmax,
             "the omega list must run to min(J, l_max, k_max): the Coriolis exponential mixes all of them");
  QD_REQUIRE(os.legendre_n_max == s.lmax - s.omin + 1,
             "associated_legendres must be (theta.n, l_max - omegas[0] + 1, n_omegas)");
  QD_REQUIRE(os.associated_legendres, "associated_legendres is required");
  QD_REQUIRE(os.wave_packets, "wave_packets is required");
  s.legendre_n_max = os.legendre_n_max;
  QD_REQUIRE(s.nt <= kMaxTheta, "theta.n must not exceed 512");
  // The propagation relies on sum_k w_k P_l(x_k) P_l'(x_k) = delta_ll' (the
  // forward/backward Legendre pair of omegawavepacket.cu:134-240 is then a
  // projector, and norms can be taken on either side of it).
  for (int b = 0; b < s.nO; b++) {
    const int nl = s.lmax - os.omegas[b] + 1;
    const double *P = os.associated_legendres + (size_t)b * s.nt * s.legendre_n_max;
    for (int c = 0; c < nl; c++)
      for (int c2 = c; c2 < nl; c2++) {
        double acc = 0.0;
        for (int k = 0; k < s.nt; k++) acc += in->theta.w[k] * P[(size_t)c * s.nt + k] * P[(size_t)c2 * s.nt + k];
        QD_REQUIRE(std::fabs(acc - (c == c2 ? 1.0 : 0.0)) < 1e-9,
                   "associated_legendres are not orthonormal under theta.w (sum_k w_k P_l P_l' != delta)");
      }
  }
  s.M = s.n1 * s.n2;
  s.dt = in->time.time_step;
  s.dr1 = in->r1.dr; s.dr2 = in->r2.dr; s.left1 = in->r1.left; s.left2 = in->r2.left;
  s.m1 = in->r1.mass; s.m2 = in->r2.mass;
  s.dump = in->dump1.dump && in->dump2.dump;
  s.mixed = in->precision == QD_PREC_MIXED;
  QD_REQUIRE(in->precision == QD_PREC_FP64 || in->precision == QD_PREC_MIXED, "unknown precision");
  s.steps0 = *in->time.steps;
  s.acf = in->autocorrelation != nullptr;   // the initial state is kept and read every step only on request
  const qd_crp &c = in->crp;
  s.crp = c.calculate_CRP != 0 && c.n_energies > 0;
  if (s.crp) {
    QD_REQUIRE(c.energies && c.eta_sq && c.CRP, "CRP.energies, eta_sq and CRP are required");
    QD_REQUIRE(c.n_gradient_points >= 3 && c.n_gradient_points % 2 == 1,
               "CRP.n_gradient_points must be odd and >= 3");
    const int h = c.n_gradient_points / 2;
    QD_REQUIRE(c.n_dividing_surface - h >= 0 && c.n_dividing_surface + h < s.n2,
               "CRP dividing surface stencil leaves the r grid");
  }
  const bool copy = in->options.wave_to_matlab || in->options.wave_callback;
  QD_REQUIRE(!copy || in->options.steps_to_copy_psi_from_device_to_host > 0,
             "steps_to_copy_psi_from_device_to_host must be > 0 when copy-back is enabled");
  // LDS limits of the Legendre GEMMs (B operand staged whole)
  const int nl_max = s.lmax - s.omin + 1;
  QD_REQUIRE(!s.acf || round_up(nl_max, 16) <= 128,
             "the autocorrelation needs l_max - omegas[0] + 1 <= 128");
  QD_REQUIRE(gemm_lds_bytes(round_up(s.nt, 4), round_up(nl_max, 16), s.mixed) <= kLdsLimit &&
                 gemm_lds_bytes(round_up(nl_max, 4), round_up(s.nt, 16), s.mixed) <= kLdsLimit,
             "theta.n / l_max too large for the LDS-resident Legendre tables");
  QD_REQUIRE(leg_tiles_supported(round_up(nl_max, 16) / 16) && leg_tiles_supported(round_up(s.nt, 16) / 16),
             "no Legendre GEMM instantiation for this theta.n / l_max (column tiles of 16)");
  // initial norm (cudaOpenMP2.cu:437-440 would abort at the first step)
  // The rows (b, k) are summed on up to 8 host threads and combined in row
  // order, so the result does not depend on the thread count.  One thread
  // reads host memory at about 8 GB/s: 85 ms for the 680 MB of 9 blocks of
  // 256 x 192 x 96, which a burst of submissions to the server paid before
  // its first step.
  const long G = (long)s.M * s.nt;
  const int rows = s.nO * s.nt;
  std::vector<double> rsum(rows);
  auto sum_rows = [&](int r0, int r1) {
    for (int r = r0; r < r1; r++) {
      const double *p = os.wave_packets + 2 * ((long)(r / s.nt) * G + (long)(r % s.nt) * s.M);
      double a = 0.0;
      for (long m = 0; m < 2L * s.M; m++) a += p[m] * p[m];
      rsum[r] = a;
    }
  };
  // one thread per MB, at most 8 (and the host's count): a 6 MB job of
  // 128 x 64 x 48 takes 0.8 ms on one thread, which a burst of them paid
  // before the server's first step
  const long bytes = 16L * G * s.nO;
  const int nthr = (int)std::min<long>(std::min(8, std::max(1, (int)std::thread::hardware_concurrency())),
                                       std::max(1L, bytes >> 20));
  if (nthr == 1) {
    sum_rows(0, rows);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < nthr; t++) pool.emplace_back(sum_rows, rows * t / nthr, rows * (t + 1) / nthr);
    for (auto &th : pool) th.join();
  }
  double norm = 0.0;
  for (int r = 0; r < rows; r++) norm += in->theta.w[r % s.nt] * rsum[r];
  norm *= s.dr1 * s.dr2;
  QD_REQUIRE(std::isfinite(norm), "initial wavepacket is not finite");
  QD_REQUIRE(norm <= 1.0 + 1e-10, "initial wavepacket norm exceeds 1");
  return s;
}

// ---------------------------------------------------------------------------
// DeviceJob
// ---------------------------------------------------------------------------
void *DeviceMemory::allocate(size_t bytes) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(QD_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  return p;
}

void DeviceMemory::release(void *p, size_t) { (void)hipFree(p); }

void DeviceMemory::h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
  QD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}

void DeviceMemory::d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
  QD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  QD_HIP(hipStreamSynchronize(s));
}

DeviceMemory &DeviceMemory::plain() {
  static DeviceMemory m;
  return m;
}

void *DeviceJob::alloc(size_t bytes) {
  if (bytes == 0) bytes = 16;
  bytes = (bytes + 255) / 256 * 256;
  void *p = mem_->allocate(bytes);
  allocs_.emplace_back(p, bytes);
  footprint += bytes;
  return p;
}

// The host copy is kept in staging_ until the constructor's final stream
// synchronisation: an asynchronous copy from pageable memory may still read
// its source after the call returns.
template <class T>
T *DeviceJob::upload(std::vector<T> v, hipStream_t s) {
  T *p = (T *)alloc(v.size() * sizeof(T));
  if (v.empty()) return p;
  auto held = std::make_shared<std::vector<T>>(std::move(v));
  staging_.push_back(held);
  QD_HIP(hipMemcpyAsync(p, held->data(), held->size() * sizeof(T), hipMemcpyHostToDevice, s));
  return p;
}

DeviceJob::~DeviceJob() {
  for (auto &a : allocs_) mem_->release(a.first, a.second);
}

// zeroed slabs past the last block that the GEMM operand reads may touch:
// psi rows nt..fwd_kp-1 of the last block; fbr rows l = omega_last + k for
// k < bwd_kp of the last block, which run past slab L - 1 of that block
static size_t gemm_pad_slabs_psi(const JobShape &s, const JobDev &d) {
  (void)s;
  return (size_t)(d.fwd_kp - d.nt);
}
static size_t gemm_pad_slabs_fbr(const JobShape &s, const JobDev &d) {
  const int om_last = s.omin + s.nO - 1;
  return (size_t)std::max(0, om_last + d.bwd_kp - s.L);
}

// QD_LEG_SPLIT=0 keeps the direct form of the fused Legendre step (tests).
static bool leg_split_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("QD_LEG_SPLIT");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

// Whether the caller's theta grid and Legendre tables have the symmetry the
// parity-split Legendre products use: nodes x_k' = -x_k and weights w_k' = w_k
// (k' = nt - 1 - k), and in every block P_l(x_k') = (-1)^(l - omega) P_l(x_k),
// all to 1e-12 of the largest value (Gauss-Legendre nodes and normalised
// associated Legendre functions satisfy it to rounding).
static bool legendre_symmetric(const qd_inputs *in, const JobShape &s) {
  const int nt = s.nt;
  const double *x = in->theta.x, *w = in->theta.w;
  double wmax = 0.0;
  for (int k = 0; k < nt; k++) wmax = std::max(wmax, std::fabs(w[k]));
  for (int k = 0; k < nt / 2; k++) {
    const int kq = nt - 1 - k;
    if (!(std::fabs(x[k] + x[kq]) <= 1e-12) || !(std::fabs(w[k] - w[kq]) <= 1e-12 * wmax)) return false;
  }
  const double *Pin = in->omega_states.associated_legendres;
  for (int b = 0; b < s.nO; b++) {
    const int nl = s.lmax - (s.omin + b) + 1;
    const double *pb = Pin + (size_t)b * nt * s.legendre_n_max;
    for (int c = 0; c < nl; c++) {
      const double *p = pb + (size_t)c * nt;
      double pmax = 0.0;
      for (int k = 0; k < nt; k++) pmax = std::max(pmax, std::fabs(p[k]));
      const double sg = c % 2 == 0 ? 1.0 : -1.0;
      for (int k = 0; k < nt / 2; k++)
        if (!(std::fabs(p[nt - 1 - k] - sg * p[k]) <= 1e-12 * pmax)) return false;
    }
  }
  return true;
}

// two-pass form where both lengths have compile-time kernels and the r pass
// with the first half step has a tiling for nO blocks
int fft_passes(const JobShape &sh, int *cols) {
  const bool ct = fft_ct_supported(sh.n1) && sh.n2 % fft_ct_rows(sh.n1, sh.nO, true) == 0;
  const int q = ct ? fft_mix_cols(sh.n2, sh.nO) : 0;
  const bool two = q > 0 && sh.n1 % q == 0;
  if (cols) *cols = two ? q : 0;
  return two ? 2 : 3;
}

int legendre_step_form(const qd_inputs *in, const JobShape &sh) {
  const int nt = sh.nt, nl_max = sh.lmax - sh.omin + 1;
  const int theta_tiles = round_up(nt, 16) / 16, l_tiles = round_up(nl_max, 16) / 16;
  const int th = (nt / 2 + 15) / 16, te = ((nl_max + 1) / 2 + 15) / 16, to = (nl_max / 2 + 15) / 16;
  if (leg_step_enabled() && leg_split_enabled() && nt % 2 == 0 && nl_max >= 2 &&
      leg_split_supported(th, te, to, sh.mixed) && leg_split_lds_bytes(th, te, to, sh.mixed) <= kLdsLimit &&
      legendre_symmetric(in, sh))
    return 2;
  if (leg_step_supported(theta_tiles, l_tiles, sh.mixed) && leg_step_lds_bytes(theta_tiles, l_tiles) <= kLdsLimit)
    return 1;
  return 0;
}

size_t DeviceJob::estimate_footprint(const JobShape &s, int n_energies) {
  const size_t G = (size_t)s.M * s.nt;
  size_t b = 0;
  b += 16 * (G + 4 * (size_t)s.M) * s.nO; // psi, with the GEMM padding
  b += 16 * (size_t)s.M * s.L * s.nO * (s.acf ? 1 : 0);   // FBR of the initial state
  b += 16 * (size_t)s.M * std::max((s.L + 4) * s.nO, s.nt + 4);   // fbr, with the GEMM padding
  b += 24 * G;                            // vphase + pot
  b += 8 * (G + 16 * (size_t)s.M);        // weighted potential (padded to 16 theta rows)
  b += 16 * (size_t)s.L * s.n1 * s.nO * s.nO;
  b += 16 * (size_t)(n_energies + kFluxWindow) * s.nO * s.nt * s.n1 * 2;   // flux accumulators, window
  b += 4 << 20;                           // tables, partials, ring
  return b;
}

DeviceJob::DeviceJob(const qd_inputs *in, const JobShape &sh, hipStream_t st, DeviceMemory *mem)
    : shape(sh), mem_(mem ? mem : &DeviceMemory::plain()) {
  const int n1 = sh.n1, n2 = sh.n2, nt = sh.nt, M = sh.M, L = sh.L, nO = sh.nO;
  const long G = (long)M * nt;
  const double dt = sh.dt;
  JobDev &d = dev;
  d.n1 = n1; d.n2 = n2; d.nt = nt; d.M = M; d.L = L; d.lmax = sh.lmax; d.nO = nO;
  d.omin = sh.omin; d.kmax = sh.kmax; d.J = sh.J; d.mixed = sh.mixed ? 1 : 0;
  d.dt = dt; d.dr1 = sh.dr1; d.dr2 = sh.dr2;
  steps = sh.steps0;

  // ---- work decomposition
  const int nl_max = sh.lmax - sh.omin + 1;
  d.fwd_kp = round_up(nt, 4);
  d.fwd_np = round_up(nl_max, 16);
  d.bwd_kp = round_up(nl_max, 4);
  d.bwd_np = round_up(nt, 16);
  const long strips = (2L * M + 15) / 16;
  // long-lived workgroups: about two per CU in one round, so each wave runs
  // many pipelined tiles after paying the B staging and pipeline fill once
  // Waves per GEMM workgroup.  8 waves share one LDS copy of B per CU (two
  // per SIMD, as two 4-wave workgroups would be, with half the B staging):
  // always when B is over 80 KB (one workgroup per CU fits), and otherwise
  // when there are strips enough to give every CU one 8-wave workgroup
  // (the J=20 workload: 72 / 74 us instead of 74 / 80 us).  Not where the
  // 8-wave build spills (more than 8 column tiles; mixed: more than 6).  The
  // strips per wave make the workgroups of all blocks fit one round of the
  // CUs (2 four-wave or 1 eight-wave workgroup per CU): each wave then runs
  // many pipelined tiles after paying the B staging and pipeline fill once.
  // the paired-row passes take strips of 16 complex points (32 rows)
  const bool fwd_paired = leg_fwd_paired(d.fwd_np / 16, sh.mixed, nO);
  const bool bwd_paired = leg_bwd_paired(d.bwd_np / 16);
  d.leg_fwd_paired = fwd_paired ? 1 : 0;
  d.leg_bwd_paired = bwd_paired ? 1 : 0;
  const long fwd_strips = fwd_paired ? ((long)M + 15) / 16 : strips;
  const long bwd_strips = bwd_paired ? ((long)M + 15) / 16 : strips;
  auto pick_waves = [&](int kp, int np, long nstrips) {
    if (np / 16 > (sh.mixed ? 6 : 8)) return 4;
    if (gemm_lds_bytes(kp, np, sh.mixed) > kLegLdsTwoGroups) return 8;
    return nstrips * nO >= 8L * kCUs ? 8 : 4;
  };
  d.gemm_fwd_waves = pick_waves(d.fwd_kp, d.fwd_np, fwd_strips);
  d.gemm_bwd_waves = pick_waves(d.bwd_kp, d.bwd_np, bwd_strips);
  if (const char *e = std::getenv("QD_LEG_WAVES")) {   // tuning sweeps: 4 or 8 for both GEMMs
    const int w = std::atoi(e);
    if (w == 4 || w == 8) d.gemm_fwd_waves = d.gemm_bwd_waves = w;
  }
  auto strips_per_wave = [&](int waves, long nstrips) {
    const long slots = waves == 8 ? kCUs : 2L * kCUs;
    long S = std::max(1L, (nstrips * nO + waves * slots - 1) / (waves * slots));
    while (nO * ((nstrips + waves * S - 1) / (waves * S)) > slots) S++;
    return (int)S;
  };
  d.gemm_strips = strips_per_wave(d.gemm_fwd_waves, fwd_strips);
  d.gemm_bwd_strips = strips_per_wave(d.gemm_bwd_waves, bwd_strips);
  d.gemm_fwd_groups = (int)((fwd_strips + (long)d.gemm_fwd_waves * d.gemm_strips - 1) /
                            ((long)d.gemm_fwd_waves * d.gemm_strips));
  d.gemm_bwd_groups = (int)((bwd_strips + (long)d.gemm_bwd_waves * d.gemm_bwd_strips - 1) /
                            ((long)d.gemm_bwd_waves * d.gemm_bwd_strips));
  // Fused Legendre step (FP64, paired rows, both tables in LDS): its waves
  // hold both products' accumulators and run one per SIMD, so one 4-wave
  // workgroup per CU, all in one round.  Its parity-split form where the
  // theta grid and the tables are symmetric under x -> -x.
  const int form = legendre_step_form(in, sh);
  d.sp_th = (nt / 2 + 15) / 16;
  d.sp_te = ((nl_max + 1) / 2 + 15) / 16;
  d.sp_to = (nl_max / 2 + 15) / 16;
  d.leg_split = form == 2 ? 1 : 0;
  // 192 x 128 x 64 with 5 blocks (126 MB): plain stores, read back from the
  // cache (R inverse pass 56 us against 62 us streamed); 256 x 192 x 96 with
  // 9 blocks (680 MB): streamed (r pass 275 -> 261 us, R inverse 373 -> 364 us)
  d.stream_out = 16.0 * (double)G * nO > 256.0 * (1 << 20) ? 1 : 0;
  d.leg_step = form >= 1 ? 1 : 0;
  d.step_waves = 4;
  {
    const long pstrips = ((long)M + 15) / 16, slots = kCUs;
    long S = std::max(1L, (pstrips * nO + d.step_waves * slots - 1) / (d.step_waves * slots));
    while (nO * ((pstrips + d.step_waves * S - 1) / (d.step_waves * S)) > slots) S++;
    d.step_strips = (int)S;
    d.step_groups = (int)((pstrips + d.step_waves * S - 1) / (d.step_waves * S));
  }
  static const int rows_c[] = {8, 4, 2, 1};
  static const int cols_c[] = {16, 8, 4, 2, 1};
  // compile-time FFT kernels where the lengths allow; otherwise the runtime-size
  // R pass keeps nb(l) <= nO blocks of fft_rows rows plus one scratch tile in LDS
  if (fft_ct_supported(n1) && n2 % fft_ct_rows(n1, nO, false) == 0 && n2 % fft_ct_rows(n1, nO, true) == 0) {
    d.fft_rows = fft_ct_rows(n1, nO, false);
    d.fft_rows_inv = fft_ct_rows(n1, nO, true);
  } else {
    d.fft_rows = d.fft_rows_inv = pick_divisor(n2, rows_c, 4, 5120 / (nO + 1), n1);
  }
  if (fft_ct_supported(n2) && n1 % fft_ct_cols(n2) == 0) d.fft_cols = fft_ct_cols(n2);
  else d.fft_cols = pick_divisor(n1, cols_c, 5, 2048, n2);
  d.fft2 = fft_passes(sh, &d.mix_cols) == 2 ? 1 : 0;
  long slabs = 0;
  for (int b = 0; b < nO; b++) slabs += L - (sh.omin + b);
  d.units[KK_LEG_FWD] = nO * d.gemm_fwd_groups;
  d.units[KK_LEG_BWD] = nO * d.gemm_bwd_groups;
  d.units[KK_FFT_R] = d.fft2 ? 0 : (L - sh.omin) * (n2 / d.fft_rows);
  d.units[KK_FFT_RI] = (L - sh.omin) * (n2 / d.fft_rows_inv);
  d.units[KK_FFT_r] = d.fft2 ? (L - sh.omin) * (n1 / d.mix_cols) : (int)(slabs * (n1 / d.fft_cols));
  d.units[KK_REDUCE] = nO * kRedFields;
  d.units[KK_FLUX] = 0;
  d.units[KK_LEG_STEP] = d.leg_step ? nO * d.step_groups : 0;

  // ---- host tables
  std::vector<double> r1(n1), r2(n2), I1(n1), I2(n2), kinE1(n1), kinE2(n2);
  for (int i = 0; i < n1; i++) {
    r1[i] = sh.left1 + i * sh.dr1;
    I1[i] = 1.0 / (2 * sh.m1 * r1[i] * r1[i]);                 // cudaMath.h:150-157
  }
  for (int j = 0; j < n2; j++) {
    r2[j] = sh.left2 + j * sh.dr2;
    I2[j] = 1.0 / (2 * sh.m2 * r2[j] * r2[j]);
  }
  auto kin_table = [](int n, double dr, double mass, std::vector<double> &out) {   // cudaMath.h:117-130
    const double two_pi_xl = 2 * M_PI / (n * dr);
    for (int i = 0; i < n; i++) {
      const double k = (i <= n / 2) ? two_pi_xl * i : two_pi_xl * (-n + i);
      out[i] = k * k / (mass + mass);
    }
  };
  kin_table(n1, sh.dr1, sh.m1, kinE1);
  kin_table(n2, sh.dr2, sh.m2, kinE2);
  auto phase = [](double x) { return make_double2(std::cos(x), std::sin(x)); };
  std::vector<double2> kin1(n1), kin2(n2);
  for (int i = 0; i < n1; i++) kin1[i] = phase(-dt * kinE1[i]);
  for (int j = 0; j < n2; j++) kin2[j] = phase(-dt * kinE2[j]);
  auto twiddles = [](int n) {
    std::vector<double2> w(n);
    for (int e = 0; e < n; e++) {
      const long double a = -2.0L * 3.141592653589793238462643383279502884L * e / n;
      w[e] = make_double2((double)cosl(a), (double)sinl(a));
    }
    return w;
  };
  // Legendre tables (omegawavepacket.cu:134-181), padded row-major [k][c]
  std::vector<double> wP((size_t)nO * d.fwd_kp * d.fwd_np, 0.0), Pt((size_t)nO * d.bwd_kp * d.bwd_np, 0.0);
  const double *Pin = in->omega_states.associated_legendres;
  for (int b = 0; b < nO; b++) {
    const int nl = sh.lmax - (sh.omin + b) + 1;
    const double *pb = Pin + (size_t)b * nt * sh.legendre_n_max;
    for (int c = 0; c < nl; c++)
      for (int k = 0; k < nt; k++) {
        const double p = pb[(size_t)c * nt + k];
        wP[((size_t)b * d.fwd_kp + k) * d.fwd_np + c] = in->theta.w[k] * p;
        Pt[((size_t)b * d.bwd_kp + c) * d.bwd_np + k] = p;
      }
  }
  // Coriolis eigenpairs and phases: exp(-i dt/2 C_l/(2 m1 R_i^2)) = V^T diag(exp(-i dt/2 e_a I1_i)) V,
  // with the R part of the two rotational quarter steps folded into the phases:
  //   corPhi(l, i, a) = exp(-i dt/2 l(l+1) I1_i) exp(-i dt/2 e_a I1_i)
  std::vector<double> corV((size_t)L * nO * nO, 0.0);
  std::vector<double2> corPhi((size_t)L * n1 * nO, make_double2(0.0, 0.0));
  std::vector<double2> cmat((size_t)L * nO, make_double2(0.0, 0.0));
  std::vector<double2> rot2((size_t)L * n2);
  const double tau = dt / 2;
  for (int l = 0; l < L; l++)
    for (int j = 0; j < n2; j++) rot2[(size_t)l * n2 + j] = phase(-tau * l * (l + 1) * I2[j]);
  for (int l = sh.omin; l < L; l++) {
    const int n = std::min(sh.kmax, l) - sh.omin + 1;
    std::vector<double> diag(n), off(std::max(1, n - 1)), z(n * n);
    coriolis_tridiagonal(sh.J, l, sh.omin, n, diag.data(), off.data());
    for (int b = 0; b < n; b++)
      cmat[(size_t)l * nO + b] = make_double2(diag[b], b + 1 < n ? off[b] : 0.0);
    tridiag_eigen(n, diag.data(), off.data(), z.data());
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++) corV[((size_t)l * nO + a) * nO + b] = z[a * n + b];
    for (int i = 0; i < n1; i++) {
      for (int a = 0; a < n; a++)
        corPhi[((size_t)l * n1 + i) * nO + a] = phase(-tau * l * (l + 1) * I1[i] - tau * I1[i] * diag[a]);
    }
  }

  // ---- device state
  // The Legendre GEMMs read operand rows k < kp (K rounded up to 4) without
  // clamping: psi and fbr carry zeroed slabs past the last block, so those
  // reads stay inside the allocations (rows k >= K are masked where they are
  // used).
  const size_t psi_pad = gemm_pad_slabs_psi(sh, d), fbr_pad = gemm_pad_slabs_fbr(sh, d);
  d.psi = (double2 *)alloc(16 * (G * nO + psi_pad * M));
  QD_HIP(hipMemsetAsync(d.psi, 0, 16 * (G * nO + psi_pad * M), st));
  mem_->h2d(d.psi, in->omega_states.wave_packets, 16 * G * nO, st);
  // fbr also serves once as the staging block of the initial state's FBR
  // coefficients below, so it holds at least fwd_kp slabs
  const size_t fbr_slabs = std::max<size_t>((size_t)L * nO + fbr_pad, (size_t)d.fwd_kp);
  d.fbr = (double2 *)alloc(16 * (size_t)M * fbr_slabs);
  d.c0 = nullptr;
  {
    double *pp = (double *)alloc(sizeof(double) * (size_t)G);
    mem_->h2d(pp, in->pot, sizeof(double) * (size_t)G, st);
    d.pot = pp;
  }
  {
    // potential phase with the absorber folded in, built on the device from pot
    const double *dd1 = nullptr, *dd2 = nullptr;
    if (sh.dump) {
      dd1 = upload(std::vector<double>(in->dump1.dump, in->dump1.dump + n1), st);
      dd2 = upload(std::vector<double>(in->dump2.dump, in->dump2.dump + n2), st);
    }
    double2 *vp = (double2 *)alloc(16 * (size_t)G);
    launch_vphase(vp, d.pot, dd1, dd2, n1, n2, (long)G, dt, st);
    QD_HIP(hipGetLastError());
    d.vphase = vp;
  }
  d.w = upload(std::vector<double>(in->theta.w, in->theta.w + nt), st);
  {
    // quadrature-weighted potential of the potential-energy epilogue,
    // w_k V(k, m), padded with zero rows to the bwd_np output columns
    double *wp = (double *)alloc(sizeof(double) * (size_t)d.bwd_np * M);
    launch_wpot(wp, d.pot, d.w, nt, (long)M, d.bwd_np, st);
    QD_HIP(hipGetLastError());
    d.wpot = wp;
  }
  d.wP = upload(wP, st);
  d.Pt = upload(Pt, st);
  if (d.leg_split) {
    // parity-split tables (kernels_leg_split.hip), from the first half of the
    // theta grid; l = omega_b + 2 m (even) and omega_b + 2 m + 1 (odd)
    const int nh = nt / 2, TH = 16 * d.sp_th, TE = 16 * d.sp_te, TO = 16 * d.sp_to;
    std::vector<double> pe((size_t)nO * TE * TH, 0.0), po((size_t)nO * TO * TH, 0.0);
    std::vector<double> wpe((size_t)nO * TH * TE, 0.0), wpo((size_t)nO * TH * TO, 0.0);
    for (int b = 0; b < nO; b++) {
      const int nl = sh.lmax - (sh.omin + b) + 1;
      const double *pb = Pin + (size_t)b * nt * sh.legendre_n_max;
      for (int c = 0; c < nl; c++) {
        const int m = c / 2;
        for (int k = 0; k < nh; k++) {
          const double p = pb[(size_t)c * nt + k], wp = in->theta.w[k] * p;
          if (c % 2 == 0) {
            pe[((size_t)b * TE + m) * TH + k] = p;
            wpe[((size_t)b * TH + k) * TE + m] = wp;
          } else {
            po[((size_t)b * TO + m) * TH + k] = p;
            wpo[((size_t)b * TH + k) * TO + m] = wp;
          }
        }
      }
    }
    d.sp_PE = upload(pe, st);
    d.sp_PO = upload(po, st);
    d.sp_wPE = upload(wpe, st);
    d.sp_wPO = upload(wpo, st);
    if (sh.mixed) {
      // forward tables in the FP32 MFMA's row order (row 16 c + 4 q + g holds
      // node 16 c + 4 g + q), backward tables as they are
      auto permuted = [&](const std::vector<double> &t, int cols) {
        std::vector<double> o(t.size());
        for (int b = 0; b < nO; b++)
          for (int r = 0; r < TH; r++) {
            const int k = (r & ~15) + 4 * (r & 3) + ((r >> 2) & 3);
            std::memcpy(&o[((size_t)b * TH + r) * cols], &t[((size_t)b * TH + k) * cols], sizeof(double) * cols);
          }
        return o;
      };
      auto hi = [](const std::vector<double> &t) {
        std::vector<float> h(t.size());
        for (size_t q = 0; q < t.size(); q++) h[q] = (float)t[q];
        return h;
      };
      const std::vector<double> wpe_p = permuted(wpe, TE), wpo_p = permuted(wpo, TO);
      d.sp_hPE = upload(hi(pe), st);
      d.sp_hPO = upload(hi(po), st);
      d.sp_hwPE = upload(hi(wpe_p), st);
      d.sp_hwPO = upload(hi(wpo_p), st);
      d.sp_xPE = upload(bf16_compensation_table(pe, nO, TE, TH), st);
      d.sp_xPO = upload(bf16_compensation_table(po, nO, TO, TH), st);
      d.sp_xwPE = upload(bf16_compensation_table(wpe_p, nO, TH, TE), st);
      d.sp_xwPO = upload(bf16_compensation_table(wpo_p, nO, TH, TO), st);
    }
  }
  if (sh.mixed) {
    // FP32 hi parts for the hi*hi products, bfloat16 tables for the compensation terms
    std::vector<float> h(wP.size()), h2(Pt.size());
    for (size_t q = 0; q < wP.size(); q++) h[q] = (float)wP[q];
    for (size_t q = 0; q < Pt.size(); q++) h2[q] = (float)Pt[q];
    d.wP_hi = upload(h, st);
    d.Pt_hi = upload(h2, st);
    d.wP_x = upload(bf16_compensation_table(wP, nO, d.fwd_kp, d.fwd_np), st);
    d.Pt_x = upload(bf16_compensation_table(Pt, nO, d.bwd_kp, d.bwd_np), st);
  }
  d.corV = upload(corV, st);
  d.corPhi = upload(corPhi, st);
  d.rot2 = upload(rot2, st);
  d.kin1 = upload(kin1, st);
  d.kin2 = upload(kin2, st);
  d.kinE1 = upload(kinE1, st);
  d.kinE2 = upload(kinE2, st);
  d.I1 = upload(I1, st);
  d.I2 = upload(I2, st);
  d.cmat = upload(cmat, st);
  d.twR = upload(twiddles(n1), st);
  d.twr = upload(twiddles(n2), st);
  d.part_cor = (double *)alloc(sizeof(double) * (size_t)d.units[KK_FFT_RI] * nO * 5);
  d.part_ke = (double *)alloc(sizeof(double) * (size_t)(d.fft2 ? nO * d.units[KK_FFT_RI] : d.units[KK_FFT_r]));
  d.part_ke2 = d.fft2 ? (double *)alloc(sizeof(double) * (size_t)nO * d.units[KK_FFT_r]) : nullptr;
  // potential-energy partials of the backward pass or of the fused step
  d.part_obs = (double *)alloc(sizeof(double) * (size_t)nO * std::max(d.gemm_bwd_groups, d.step_groups));
  if (sh.acf) {
    // Autocorrelation <psi(t_0)|psi(t_n)>.  The state stored after a step
    // lacks the trailing V(dt/2) of the leapfrogged potential
    // (cudaOpenMP2.cu:338-387), so psi(t_n) = exp(-i dt/2 V) psi_stored(n),
    // and on a fresh start psi(t_0) is the caller's psi_0; on a resume
    // (steps != 0) the caller's array is itself a stored state, psi(t_0) =
    // exp(-i dt/2 V) psi_in.  As V is real and diagonal,
    //   <psi(t_0)|psi(t_n)> = <a|psi_stored(n)>,  a = psi_0 exp(+i dt/2 V) (fresh) or psi_in (resume),
    // and since psi_stored(n) is a Legendre expansion (the output of the
    // backward transform) the DVR inner product equals the FBR one on the
    // forward-transformed a.  c0 holds those coefficients; the FFT pass
    // forms the products.  Each block of a is staged in fbr (zeroed after).
    d.c0 = (double2 *)alloc(16 * (size_t)M * L * nO);
    for (int b = 0; b < nO; b++) {
      const int om = sh.omin + b, nl = sh.lmax - om + 1;
      QD_HIP(hipMemcpyAsync(d.fbr, d.psi + (size_t)b * G, 16 * (size_t)G, hipMemcpyDeviceToDevice, st));
      if (sh.steps0 == 0) launch_prephase(d.fbr, d.pot, G, 1, dt, st);
      // rows nt..fwd_kp-1 of the staged operand are read and masked off
      launch_gemm_probe((const double *)d.fbr, 2 * M, d.wP + (size_t)b * d.fwd_kp * d.fwd_np, d.fwd_kp, d.fwd_np,
                        (double *)(d.c0 + ((size_t)b * L + om) * M), 2 * M, 2 * M, nt, nl, st);
      QD_HIP(hipGetLastError());
    }
  }
  QD_HIP(hipMemsetAsync(d.fbr, 0, 16 * (size_t)M * fbr_slabs, st));
  d.ring = kRing;
  d.obs_ring = (double *)alloc(sizeof(double) * (size_t)kRing * nO * 5);
  d.acf_ring = (double2 *)alloc(sizeof(double2) * (size_t)kRing * nO);
  d.counters = (int *)alloc(4 * sizeof(int));
  const int cnt[4] = {sh.steps0, 0, 0, 0};
  QD_HIP(hipMemcpyAsync(d.counters, cnt, sizeof(cnt), hipMemcpyHostToDevice, st));
  d.flux_on = 0;
  d.units[KK_FLUX] = 0;
  if (sh.crp) {
    const qd_crp &c = in->crp;
    const int ng = c.n_gradient_points, half = ng / 2, ne = c.n_energies;
    d.flux_on = 1;
    d.flux_ns = c.n_dividing_surface;
    d.flux_ng = ng;
    d.flux_ne = ne;
    d.flux_T = kFluxWindow;
    d.flux_S = nO * nt * n1;
    d.units[KK_FLUX] = (d.flux_S + kBlock - 1) / kBlock;
    // restores the leapfrogged V(dt/2) on the stencil rows
    std::vector<double2> rph((size_t)ng * nt * n1);
    for (int g = 0; g < ng; g++)
      for (int k = 0; k < nt; k++)
        for (int i = 0; i < n1; i++) {
          const long idx = (long)k * M + (long)(d.flux_ns - half + g) * n1 + i;
          rph[((size_t)g * nt + k) * n1 + i] = phase(-0.5 * dt * in->pot[idx]);
        }
    d.flux_coef = upload(first_derivative_weights(ng, sh.dr2), st);
    d.flux_rph = upload(rph, st);
    const size_t S = d.flux_S;
    d.flux_W = (double2 *)alloc(16 * S * 2 * kFluxWindow);
    flux_A_ = (double2 *)alloc(16 * S * ne);
    flux_B_ = (double2 *)alloc(16 * S * ne);
    QD_HIP(hipMemsetAsync(flux_A_, 0, 16 * S * ne, st));
    QD_HIP(hipMemsetAsync(flux_B_, 0, 16 * S * ne, st));
    flux_ph_ = (double2 *)alloc(16 * (size_t)ne * kFluxWindow);
    flux_E_ = upload(std::vector<double>(c.energies, c.energies + ne), st);
    flux_eta_ = upload(std::vector<double>(c.eta_sq, c.eta_sq + ne), st);
    flux_out_ = (double *)alloc(sizeof(double) * ne);
    flux_window_start_ = sh.steps0;
  }
  QD_HIP(hipStreamSynchronize(st));   // host tables go out of scope
}

std::vector<double> first_derivative_weights(int n, double h) {
  // c_m = (-1)^(m+1) (k!)^2 / (m (k-m)! (k+m)!) / h, c_-m = -c_m, k = n / 2:
  // exact for polynomials of degree <= n - 1, as a product of ratios
  const int half = n / 2;
  std::vector<double> c(n, 0.0);
  for (int m = 1; m <= half; m++) {
    double r = 1.0;
    for (int q = 0; q < m; q++) r *= (double)(half - q) / (double)(half + q + 1);
    const double cm = ((m % 2) ? 1.0 : -1.0) * r / m / h;
    c[half + m] = cm;
    c[half - m] = -cm;
  }
  return c;
}

void DeviceJob::drain(int records_total) {
  const int nO = shape.nO;
  if (records_total <= records_done) return;
  QD_REQUIRE(records_total - records_done <= dev.ring, "observable ring overflow");
  const int n = records_total - records_done;
  obs.resize((size_t)records_total * nO * 5);
  acf.resize((size_t)records_total * 2);
  std::vector<double2> part((size_t)n * nO);
  // the new records fill ring slots from records_done % ring on: at most two
  // contiguous spans, one copy each per ring (not one per record)
  for (int r = records_done; r < records_total;) {
    const int slot = r % dev.ring;
    const int len = std::min(records_total - r, dev.ring - slot);
    QD_HIP(hipMemcpy(obs.data() + (size_t)r * nO * 5, dev.obs_ring + (size_t)slot * nO * 5,
                     sizeof(double) * nO * 5 * len, hipMemcpyDeviceToHost));
    QD_HIP(hipMemcpy(part.data() + (size_t)(r - records_done) * nO, dev.acf_ring + (size_t)slot * nO,
                     sizeof(double2) * nO * len, hipMemcpyDeviceToHost));
    r += len;
  }
  for (int q = 0; q < n; q++) {
    double re = 0.0, im = 0.0;
    for (int b = 0; b < nO; b++) {
      re += part[(size_t)q * nO + b].x;
      im += part[(size_t)q * nO + b].y;
    }
    acf[(size_t)(records_done + q) * 2] = re;
    acf[(size_t)(records_done + q) * 2 + 1] = im;
  }
  records_done = records_total;
}

void DeviceJob::copy_psi(double *dst, hipStream_t s) const {
  const size_t bytes = 16 * (size_t)shape.M * shape.nt * shape.nO;
  mem_->d2h(dst, dev.psi, bytes, s);
}

void DeviceJob::prephase_if_first(hipStream_t s) {
  if (steps != 0 || prephased) return;
  launch_prephase(dev.psi, dev.pot, (long)shape.M * shape.nt, shape.nO, shape.dt, s);
  QD_HIP(hipGetLastError());
  prephased = true;
}

void DeviceJob::flux_flush(hipStream_t s) {
  if (!dev.flux_on || flux_steps() == flux_window_start_) return;
  FluxWindow w;
  w.W = dev.flux_W;
  w.A = flux_A_;
  w.B = flux_B_;
  w.ph = flux_ph_;
  w.E = flux_E_;
  w.dt = shape.dt;
  w.S = dev.flux_S;
  w.ne = dev.flux_ne;
  w.T = dev.flux_T;
  w.s0 = flux_window_start_;
  w.n = flux_steps() - flux_window_start_;
  launch_flux_window(w, s);
  QD_HIP(hipGetLastError());
  flux_window_start_ = flux_steps();
}

void DeviceJob::flux_after_step(hipStream_t s) {
  if (dev.flux_on && flux_steps() - flux_window_start_ >= dev.flux_T) flux_flush(s);
}

int DeviceJob::steps_to_flux_flush() const {
  return dev.flux_on ? dev.flux_T - (flux_steps() - flux_window_start_) : std::numeric_limits<int>::max();
}

void DeviceJob::flux_to_crp(double *crp, hipStream_t s) {
  if (!dev.flux_on) return;
  QD_REQUIRE(!pending, "reaction probabilities need the last step finished");
  flux_flush(s);
  // P = Im(sum w conj(A) B) dr1 / (2 pi mu_r eta)
  const double scale = shape.dr1 / (2.0 * M_PI * shape.m2);
  launch_flux_finish(flux_A_, flux_B_, dev.w, flux_eta_, flux_out_, dev.flux_S, shape.n1, shape.nt, dev.flux_ne,
                     scale, s);
  QD_HIP(hipGetLastError());
  QD_HIP(hipMemcpyAsync(crp, flux_out_, sizeof(double) * dev.flux_ne, hipMemcpyDeviceToHost, s));
  QD_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
Engine::Engine(hipStream_t s) : stream_(s) {}

Engine::~Engine() {
  drop_graph();
  if (upload_done_) (void)hipEventDestroy(upload_done_);
  if (d_jobs_) (void)hipFree(d_jobs_);
  if (d_bases_) (void)hipFree(d_bases_);
}

void Engine::drop_graph() {
  for (int g = 0; g < 2; g++) {
    if (graph_exec_[g]) (void)hipGraphExecDestroy(graph_exec_[g]);
    if (graph_[g]) (void)hipGraphDestroy(graph_[g]);
    graph_exec_[g] = nullptr;
    graph_[g] = nullptr;
  }
}

// The descriptor tables go up with asynchronous copies on the engine's
// stream, ordered after the steps already enqueued there; the host copies
// stay untouched until the previous upload has passed (upload_done_), so the
// host does not wait for the stream's queued steps.
void Engine::set_jobs(const std::vector<DeviceJob *> &jobs) {
  drop_graph();
  jobs_ = jobs;
  const int n = (int)jobs.size();
  if (!upload_done_) QD_HIP(hipEventCreateWithFlags(&upload_done_, hipEventDisableTiming));
  else QD_HIP(hipEventSynchronize(upload_done_));
  if (n > n_alloc_jobs_) {
    QD_HIP(hipStreamSynchronize(stream_));   // the old tables may still be read by queued steps
    if (d_jobs_) QD_HIP(hipFree(d_jobs_));
    if (d_bases_) QD_HIP(hipFree(d_bases_));
    d_jobs_ = nullptr;
    d_bases_ = nullptr;
    const int cap = std::max(n, 8);
    QD_HIP(hipMalloc(&d_jobs_, sizeof(JobDev) * cap));
    QD_HIP(hipMalloc(&d_bases_, sizeof(int) * (KK_COUNT * (cap + 1) + 2 * KK_COUNT * cap)));
    n_alloc_jobs_ = cap;
  }
  if (n == 0) return;
  descs_.assign(n, JobDev{});
  // prefix tables of all jobs, then per job the single-job tables {0, units}
  bases_.assign((size_t)KK_COUNT * (n + 1) + 2 * KK_COUNT * n, 0);
  std::vector<JobDev> &descs = descs_;
  std::vector<int> &bases = bases_;
  lds_fftR_ = lds_fftr_ = 0;
  ct_n1_ = jobs[0]->dev.n1;
  ct_n2_ = jobs[0]->dev.n2;
  fwd_tiles_ = bwd_tiles_ = fwd_kp_ = bwd_kp_ = 0;
  nb_max_ = 1;
  mixed_ = jobs[0]->dev.mixed;
  fwd_waves_ = jobs[0]->dev.gemm_fwd_waves;
  bwd_waves_ = jobs[0]->dev.gemm_bwd_waves;
  any_flux_ = 0;
  any_acf_ = 0;
  fused_ = jobs[0]->dev.leg_step;
  split_ = jobs[0]->dev.leg_split;
  fft2_ = jobs[0]->dev.fft2;
  step_waves_ = jobs[0]->dev.step_waves;
  step_ttiles_ = step_ltiles_ = 0;
  lds_step_ = 0;
  for (int q = 0; q < n; q++) {
    const JobDev &d = jobs[q]->dev;
    QD_REQUIRE(d.leg_step == fused_ && d.step_waves == step_waves_ && d.leg_split == split_,
               "all jobs of one engine must share the Legendre step form");
    // the split kernel's tile counts address the tables: one set per engine
    QD_REQUIRE(!split_ || (d.sp_th == jobs[0]->dev.sp_th && d.sp_te == jobs[0]->dev.sp_te &&
                           d.sp_to == jobs[0]->dev.sp_to),
               "all jobs of one engine must share the parity-split tile counts");
    if (fused_) {
      step_ttiles_ = std::max(step_ttiles_, d.bwd_np / 16);
      step_ltiles_ = std::max(step_ltiles_, d.fwd_np / 16);
      lds_step_ = std::max(lds_step_, split_ ? leg_split_lds_bytes(d.sp_th, d.sp_te, d.sp_to, d.mixed)
                                             : leg_step_lds_bytes(d.bwd_np / 16, d.fwd_np / 16));
    }
    int *single = bases.data() + (size_t)KK_COUNT * (n + 1) + (size_t)2 * KK_COUNT * q;
    for (int k = 0; k < KK_COUNT; k++) single[2 * k + 1] = d.units[k];
    QD_REQUIRE(d.mixed == mixed_, "all jobs of one engine must share a precision");
    QD_REQUIRE(d.fft2 == fft2_ && d.mix_cols == jobs[0]->dev.mix_cols,
               "all jobs of one engine must share the FFT form and the r-pass tiling");
    QD_REQUIRE(d.gemm_fwd_waves == fwd_waves_ && d.gemm_bwd_waves == bwd_waves_,
               "all jobs of one engine must share the GEMM workgroup size");
    // the strips of a job (16 points when paired, 16 rows otherwise) follow
    // its tile counts, and one launch runs one kernel for every job
    QD_REQUIRE(d.leg_fwd_paired == jobs[0]->dev.leg_fwd_paired && d.leg_bwd_paired == jobs[0]->dev.leg_bwd_paired,
               "all jobs of one engine must share the Legendre kernel form (paired or one-row strips)");
    descs[q] = d;
    for (int k = 0; k < KK_COUNT; k++) bases[(size_t)k * (n + 1) + q + 1] = bases[(size_t)k * (n + 1) + q] + d.units[k];
    lds_fftR_ = std::max(lds_fftR_, (size_t)((d.nO + 1) * d.fft_rows * d.n1 + d.n1) * 16);
    lds_fftr_ = std::max(lds_fftr_, (size_t)(2 * d.fft_cols * d.n2 + d.n2) * 16);
    // compile-time kernels only when every job shares the length and tiling
    if (d.n1 != ct_n1_ || !fft_ct_supported(d.n1)) ct_n1_ = 0;
    if (d.n2 != ct_n2_ || !fft_ct_supported(d.n2) || d.fft_cols != fft_ct_cols(d.n2)) ct_n2_ = 0;
    fwd_tiles_ = std::max(fwd_tiles_, d.fwd_np / 16);
    bwd_tiles_ = std::max(bwd_tiles_, d.bwd_np / 16);
    fwd_kp_ = std::max(fwd_kp_, d.fwd_kp);
    bwd_kp_ = std::max(bwd_kp_, d.bwd_kp);
    nb_max_ = std::max(nb_max_, d.nO);
    any_flux_ |= d.flux_on;
    any_acf_ |= d.c0 != nullptr;
  }
  // the compile-time R passes take fft_ct_rows(n1, nb_max, inverse) rows per workgroup:
  // every job must have been tiled that way
  for (int q = 0; q < n && ct_n1_; q++)
    if (jobs[q]->dev.fft_rows != fft_ct_rows(ct_n1_, nb_max_, false) ||
        jobs[q]->dev.fft_rows_inv != fft_ct_rows(ct_n1_, nb_max_, true))
      ct_n1_ = 0;
  // the two-pass form has compile-time kernels only, instantiated for nb_max_
  QD_REQUIRE(!fft2_ || (ct_n1_ && ct_n2_ && jobs[0]->dev.mix_cols == fft_mix_cols(ct_n2_, nb_max_)),
             "two-pass FFT form without compile-time kernels for every job");
  for (int k = 0; k < KK_COUNT; k++) total_[k] = bases[(size_t)k * (n + 1) + n];
  QD_HIP(hipMemcpyAsync(d_jobs_, descs.data(), sizeof(JobDev) * n, hipMemcpyHostToDevice, stream_));
  QD_HIP(hipMemcpyAsync(d_bases_, bases.data(), sizeof(int) * bases.size(), hipMemcpyHostToDevice, stream_));
  QD_HIP(hipEventRecord(upload_done_, stream_));
}

StepArgs Engine::args(int kind, int gate) const {
  StepArgs a;
  a.jobs = d_jobs_;
  a.unit_base = d_bases_ + (size_t)kind * (jobs_.size() + 1);
  a.n_jobs = (int)jobs_.size();
  a.gate = gate;
  a.red_base = nullptr;
  a.red_units = 0;
  a.flux_base = nullptr;
  a.flux_units = 0;
  return a;
}

// the fused Legendre kernel of this engine's jobs (direct or parity-split
// form), fused steps or (finish) the backward transform alone
void Engine::launch_step_kernel(const StepArgs &a, int units, int finish, hipStream_t s) const {
  if (split_) {
    const JobDev &d = jobs_[0]->dev;
    launch_leg_split(a, units, d.sp_th, d.sp_te, d.sp_to, mixed_, lds_step_, finish, s);
  } else {
    launch_leg_step(a, units, step_ttiles_, step_ltiles_, lds_step_, finish, s);
  }
}

StepArgs Engine::args_one(int q, int kind) const {
  StepArgs a;
  a.jobs = d_jobs_ + q;
  a.unit_base = d_bases_ + (size_t)KK_COUNT * (jobs_.size() + 1) + (size_t)2 * KK_COUNT * q + 2 * kind;
  a.n_jobs = 1;
  a.gate = 0;
  a.red_base = nullptr;
  a.red_units = 0;
  a.flux_base = nullptr;
  a.flux_units = 0;
  return a;
}

void Engine::enqueue(hipEvent_t *ev) {
  hipStream_t s = stream_;
  // with ev: events ev[2k], ev[2k+1] bracket the launch of kernel kind k
  // (kinds fused into another kernel get none)
  auto begin = [&](int k) {
    if (ev) QD_HIP(hipEventRecord(ev[2 * k], s));
  };
  auto end = [&](int k) {
    if (ev) QD_HIP(hipEventRecord(ev[2 * k + 1], s));
  };
  if (fused_) {
    // the previous step's backward transform with this step's forward one;
    // then that previous step's flux values and record, for the jobs that
    // had one pending
    begin(QD_K_LEG_STEP);
    launch_step_kernel(args(KK_LEG_STEP), total_[KK_LEG_STEP], 0, s);
    end(QD_K_LEG_STEP);
    // the reduction and the flux surface values of that step ride on the
    // first FFT pass below
  } else {
    begin(QD_K_LEG_FWD);
    launch_leg_fwd(args(KK_LEG_FWD), total_[KK_LEG_FWD], fwd_tiles_, fwd_kp_, mixed_, fwd_waves_,
                   jobs_[0]->dev.leg_fwd_paired, s);
    end(QD_K_LEG_FWD);
  }
  if (fft2_) {
    // two-pass form: the r pass with the first half step, carrying the
    // previous step's reduction and flux values like the forward R pass
    // below (it writes only the r kinetic partial sums, which that
    // reduction does not read); then the R pass with the R kinetic factor
    begin(QD_K_FFT_r);
    StepArgs ar = args(KK_FFT_r);
    if (fused_) {
      ar.red_base = d_bases_ + (size_t)KK_REDUCE * (jobs_.size() + 1);
      ar.red_units = total_[KK_REDUCE];
      if (any_flux_) {
        ar.flux_base = d_bases_ + (size_t)KK_FLUX * (jobs_.size() + 1);
        ar.flux_units = total_[KK_FLUX];
      }
    }
    launch_fft_r_mix(ar, total_[KK_FFT_r], ct_n2_, nb_max_, s);
    end(QD_K_FFT_r);
    begin(QD_K_FFT_R_INV);
    launch_fft_R_kin(args(KK_FFT_RI), total_[KK_FFT_RI], any_acf_, ct_n1_, nb_max_, s);
    end(QD_K_FFT_R_INV);
  } else {
  begin(QD_K_FFT_R_FWD);
  {
    // fused sequence: the previous step's (gated) reduction as the leading
    // workgroups of the forward R pass, which reads and writes no partial
    // sum and no counter (one launch less per step)
    StepArgs ar = args(KK_FFT_R);
    if (fused_) {
      ar.red_base = d_bases_ + (size_t)KK_REDUCE * (jobs_.size() + 1);
      ar.red_units = total_[KK_REDUCE];
      if (any_flux_) {
        ar.flux_base = d_bases_ + (size_t)KK_FLUX * (jobs_.size() + 1);
        ar.flux_units = total_[KK_FLUX];
      }
    }
    if (ct_n1_) launch_fft_R_ct(ar, total_[KK_FFT_R], 0, 0, ct_n1_, nb_max_, s);
    else launch_fft_R_rt(ar, total_[KK_FFT_R], 0, 0, nb_max_, lds_fftR_, s);
  }
  end(QD_K_FFT_R_FWD);
  begin(QD_K_FFT_r);
  if (ct_n2_) launch_fft_r_ct(args(KK_FFT_r), total_[KK_FFT_r], ct_n2_, s);
  else launch_fft_r_rt(args(KK_FFT_r), total_[KK_FFT_r], lds_fftr_, s);
  end(QD_K_FFT_r);
  begin(QD_K_FFT_R_INV);
  if (ct_n1_) launch_fft_R_ct(args(KK_FFT_RI), total_[KK_FFT_RI], 1, any_acf_, ct_n1_, nb_max_, s);
  else launch_fft_R_rt(args(KK_FFT_RI), total_[KK_FFT_RI], 1, any_acf_, nb_max_, lds_fftR_, s);
  end(QD_K_FFT_R_INV);
  }
  if (!fused_) {
    begin(QD_K_LEG_BWD);
    launch_leg_bwd(args(KK_LEG_BWD), total_[KK_LEG_BWD], bwd_tiles_, bwd_kp_, mixed_, bwd_waves_,
                   jobs_[0]->dev.leg_bwd_paired, s);
    end(QD_K_LEG_BWD);
    if (any_flux_) {
      begin(QD_K_OBS);
      launch_flux(args(KK_FLUX), total_[KK_FLUX], s);
      end(QD_K_OBS);
    }
    begin(QD_K_REDUCE);
    launch_reduce(args(KK_REDUCE), total_[KK_REDUCE], s);
    end(QD_K_REDUCE);
  }
  QD_HIP(hipGetLastError());
}

void Engine::finish(int q) {
  if (jobs_.empty() || !fused_) return;
  hipStream_t s = stream_;
  // the fused kernel's backward-only form: the same workgroups and partial
  // sums as the fused steps, so a record does not depend on where a run was cut
  if (q < 0) {
    launch_step_kernel(args(KK_LEG_STEP), total_[KK_LEG_STEP], 1, s);
    if (any_flux_) launch_flux(args(KK_FLUX), total_[KK_FLUX], s);
    launch_reduce(args(KK_REDUCE), total_[KK_REDUCE], s);
  } else {
    const JobDev &d = jobs_[q]->dev;
    launch_step_kernel(args_one(q, KK_LEG_STEP), d.units[KK_LEG_STEP], 1, s);
    if (d.flux_on) launch_flux(args_one(q, KK_FLUX), d.units[KK_FLUX], s);
    launch_reduce(args_one(q, KK_REDUCE), d.units[KK_REDUCE], s);
  }
  QD_HIP(hipGetLastError());
}

unsigned Engine::timed_kinds() const {
  unsigned m = (fft2_ ? 0u : (1u << QD_K_FFT_R_FWD)) | (1u << QD_K_FFT_r) | (1u << QD_K_FFT_R_INV);
  m |= fused_ ? (1u << QD_K_LEG_STEP) : ((1u << QD_K_LEG_FWD) | (1u << QD_K_LEG_BWD) | (1u << QD_K_REDUCE));
  if (any_flux_ && !fused_) m |= 1u << QD_K_OBS;   // fused: carried by the first FFT pass
  return m;
}

void Engine::step(hipEvent_t *ev) {
  if (jobs_.empty()) return;
  enqueue(ev);
}

void Engine::step_graph(int k) {
  if (jobs_.empty()) return;
  QD_REQUIRE(k == 1 || k == kGraphSteps, "step graphs hold 1 or kGraphSteps steps");
  const int g = k == 1 ? 0 : 1;
  if (!graph_exec_[g]) {
    QD_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      for (int i = 0; i < k; i++) enqueue(nullptr);
    } catch (...) {
      hipGraph_t gr = nullptr;
      (void)hipStreamEndCapture(stream_, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    QD_HIP(hipStreamEndCapture(stream_, &graph_[g]));
    QD_HIP(hipGraphInstantiate(&graph_exec_[g], graph_[g], nullptr, nullptr, 0));
  }
  QD_HIP(hipGraphLaunch(graph_exec_[g], stream_));
}

}  // namespace qd
