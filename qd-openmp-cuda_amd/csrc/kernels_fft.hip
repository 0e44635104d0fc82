// kernels_fft.hip - FFT passes over the (R, r) plane of every FBR slab, for
// any length that factors into 2, 3, 5 (runtime-size path).  The common
// lengths use the compile-time kernels of kernels_fft_ct.hip instead.
//
// Replaces the batched cuFFT Z2Z plan over all l_max+1 slabs
// (wavepackets1device.cu:186-201; exec omegawavepacket.cu:242-269) with three
// passes that only touch the slabs that carry data (l >= omega):
//   R pass, forward  rotational . Coriolis . rotational (first half step), then
//                    the forward DFT along R (contiguous pencils)
//   r pass           forward DFT along r, kinetic phase exp(-i dt T) and the
//                    kinetic-energy partial sum, inverse DFT along r
//                    (_evolution_with_kinetic_ evolutionCUDAaux.cu:13-31;
//                     _psi_times_kinitic_energy_ :56-74)
//   R pass, inverse  inverse DFT along R and the 1/(n1 n2) scale
//                    (cublasZdscal, omegawavepacket.cu:256-267), then the second
//                    rotational . Coriolis . rotational half step with the Coriolis
//                    and rotational energies (cudaOpenMP2.cu:364-370)
// Forward uses exp(-2 pi i jk/n), inverse exp(+2 pi i jk/n), both unnormalised,
// like CUFFT_FORWARD / CUFFT_INVERSE.
#include <mutex>

#include "fft_common.h"
#include "step_reduce.h"
#include "flux_surface.h"
#include "host_internal.h"

namespace qd {

// One Stockham stage for one butterfly (runtime n, Ns).
template <bool INV, int R>
__device__ __forceinline__ void butterfly(const double2 *__restrict__ src, double2 *__restrict__ dst, int n, int Ns,
                                          int idx, int base, int estride, const double2 *__restrict__ tw) {
  const int nb = n / R;
  const int k = idx % Ns;
  const int tws = n / (Ns * R);
  double2 v[R];
#pragma unroll
  for (int q = 0; q < R; q++) v[q] = src[base + (idx + q * nb) * estride];
#pragma unroll
  for (int q = 1; q < R; q++) {
    double2 w = tw[q * k * tws];
    if (INV) w.y = -w.y;
    v[q] = cmul(v[q], w);
  }
  double2 X[R];
  dft_small<INV, R>(v, X);
  const int d = (idx / Ns) * Ns * R + k;
#pragma unroll
  for (int q = 0; q < R; q++) dst[base + (d + q * Ns) * estride] = X[q];
}

// In-LDS FFT of npen pencils of length n; pencil p element e lives at
// buf[p * pstride + e * estride].  Ping-pongs between A and B; returns the
// buffer holding the result.  PEN_FAST maps consecutive threads to
// consecutive pencils (for pencils interleaved in LDS).
template <bool INV, bool PEN_FAST>
__device__ double2 *fft_lds(double2 *A, double2 *B, int n, int npen, int pstride, int estride,
                            const double2 *tw) {
  double2 *src = A, *dst = B;
  for (int Ns = 1; Ns < n;) {
    const int R = next_radix(n / Ns);
    const int nb = n / R;
    const int total = npen * nb;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      const int p = PEN_FAST ? t % npen : t / nb;
      const int idx = PEN_FAST ? t / npen : t % nb;
      const int base = p * pstride;
      if (R == 4) butterfly<INV, 4>(src, dst, n, Ns, idx, base, estride, tw);
      else if (R == 2) butterfly<INV, 2>(src, dst, n, Ns, idx, base, estride, tw);
      else if (R == 3) butterfly<INV, 3>(src, dst, n, Ns, idx, base, estride, tw);
      else butterfly<INV, 5>(src, dst, n, Ns, idx, base, estride, tw);
    }
    __syncthreads();
    double2 *tmp = src; src = dst; dst = tmp;
    Ns *= R;
  }
  return src;
}

// ---------------------------------------------------------------------------
// R pass (runtime size): one workgroup owns slab l and fft_rows consecutive
// r-rows for every omega block present at l (omega_b <= min(K, l)); the
// tiles of all blocks sit in LDS so the Coriolis mix and the R transforms
// share them.
// ---------------------------------------------------------------------------
template <bool INV, int NB, bool ACF>
__global__ __launch_bounds__(kBlock) void k_fft_R(StepArgs a) {
  constexpr int NF = ACF ? 5 : 3;
  extern __shared__ double2 lds[];
  __shared__ double red[4 * NF * NB];
  if (!INV && reduce_carried(a, red)) return;   // leading workgroups: the previous step's reduction
  if (!INV && flux_carried(a)) return;          // then its flux surface values
  const int bid = blockIdx.x - (INV ? 0 : a.red_units + a.flux_units);
  const int job = locate_job(a.unit_base, a.n_jobs, bid);
  const JobDev &jb = a.jobs[job];
  const int u = bid - a.unit_base[job];
  const int n1 = jb.n1, P = INV ? jb.fft_rows_inv : jb.fft_rows, M = jb.M, L = jb.L;
  const long LM = (long)L * M;
  const int groups = jb.n2 / P;
  const int l = jb.omin + u / groups, g = u % groups;
  const int nb = min(jb.kmax, l) - jb.omin + 1;
  const int PN = P * n1;
  const QD_G double *V = jb.corV + (long)l * jb.nO * jb.nO;
  const QD_G double2 *c0 = ACF ? jb.c0 + (long)l * M + (long)g * PN : nullptr;
  double2 *X = lds, *S = lds + nb * PN, *tw = S + PN;
  for (int e = threadIdx.x; e < n1; e += kBlock) tw[e] = jb.twR[e];
  for (int b = 0; b < nb; b++) {
    const double2 *src = jb.fbr + ((long)b * L + l) * M + (long)g * PN;
    for (int e = threadIdx.x; e < PN; e += kBlock) X[b * PN + e] = src[e];
  }
  __syncthreads();
  FbrSums<NB, NF> sm;
  sm.zero();
  if (!INV) {
    for (int e = threadIdx.x; e < PN; e += kBlock) {
      double2 x[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) x[b] = (b < nb) ? X[b * PN + e] : make_double2(0.0, 0.0);
      coriolis_point<false, NB, NF>(jb, x, nb, l, e % n1, g * P + e / n1, V, sm, nullptr, LM,
                                    jb.corPhi + ((long)l * n1 + e % n1) * jb.nO);
#pragma unroll
      for (int b = 0; b < NB; b++)
        if (b < nb) X[b * PN + e] = x[b];
    }
    __syncthreads();
  }
  for (int b = 0; b < nb; b++) {
    double2 *res = fft_lds<INV, false>(X + b * PN, S, n1, P, n1, 1, tw);
    if (res != X + b * PN) {
      for (int e = threadIdx.x; e < PN; e += kBlock) X[b * PN + e] = res[e];
      __syncthreads();
    }
  }
  if (INV) {
    const double s = 1.0 / (double)M;
    for (int e = threadIdx.x; e < PN; e += kBlock) {
      double2 x[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) x[b] = (b < nb) ? cscale(X[b * PN + e], s) : make_double2(0.0, 0.0);
      coriolis_point<true, NB, NF>(jb, x, nb, l, e % n1, g * P + e / n1, V, sm, ACF ? c0 + e : nullptr, LM,
                                   jb.corPhi + ((long)l * n1 + e % n1) * jb.nO);
#pragma unroll
      for (int b = 0; b < NB; b++)
        if (b < nb) X[b * PN + e] = x[b];
    }
    __syncthreads();
  }
  for (int b = 0; b < nb; b++) {
    double2 *dst = jb.fbr + ((long)b * L + l) * M + (long)g * PN;
    for (int e = threadIdx.x; e < PN; e += kBlock) dst[e] = X[b * PN + e];
  }
  if (INV) {
    store_fbr_partials<NB, NF>(jb, u, sm, red);
    // fused Legendre step: this step's backward transform is now pending
    if (u == 0 && threadIdx.x == 0 && jb.leg_step) jb.counters[3] = 1;
  }
}

void launch_fft_R_rt(const StepArgs &a, int total_units, int inverse, int acf, int nb_max, size_t lds,
                     hipStream_t stream) {
  const dim3 g(total_units + (inverse ? 0 : a.red_units + a.flux_units)), t(kBlock);
#define QD_FFTR_CASE(NB)                                                                    \
  if (nb_max <= NB) {                                                                       \
    if (!inverse) hipLaunchKernelGGL((k_fft_R<false, NB, false>), g, t, lds, stream, a);    \
    else if (acf) hipLaunchKernelGGL((k_fft_R<true, NB, true>), g, t, lds, stream, a);      \
    else hipLaunchKernelGGL((k_fft_R<true, NB, false>), g, t, lds, stream, a);              \
    return;                                                                                 \
  }
  QD_FFTR_CASE(1)
  QD_FFTR_CASE(2)
  QD_FFTR_CASE(4)
  QD_FFTR_CASE(6)
  QD_FFTR_CASE(9)
  QD_FFTR_CASE(12)
  QD_FFTR_CASE(16)
  QD_FFTR_CASE(24)
#undef QD_FFTR_CASE
}

// ---------------------------------------------------------------------------
// r pass (runtime size): fft_cols consecutive R-columns per workgroup
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_fft_r(StepArgs a) {
  extern __shared__ double2 lds[];
  __shared__ double red[4];
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int u = blockIdx.x - a.unit_base[job];
  const int n1 = jb.n1, n2 = jb.n2, Q = jb.fft_cols, M = jb.M;
  int b, l, g;
  decode_slab_unit(jb, u, n1 / Q, b, l, g);
  double2 *base = jb.fbr + ((long)b * jb.L + l) * M + (long)g * Q;
  double2 *A = lds, *B = lds + Q * n2, *tw = lds + 2 * Q * n2;
  for (int e = threadIdx.x; e < n2; e += kBlock) tw[e] = jb.twr[e];
  for (int e = threadIdx.x; e < Q * n2; e += kBlock) A[e] = base[(long)(e / Q) * n1 + (e % Q)];
  __syncthreads();
  double2 *X = fft_lds<false, true>(A, B, n2, Q, 1, Q, tw);
  double2 *Y = (X == A) ? B : A;
  double ke = 0.0;
  for (int e = threadIdx.x; e < Q * n2; e += kBlock) {
    const int i = g * Q + (e % Q), j = e / Q;
    const double2 x = cmul(X[e], cmul(jb.kin1[i], jb.kin2[j]));
    X[e] = x;
    ke += cnorm(x) * (jb.kinE1[i] + jb.kinE2[j]);
  }
  __syncthreads();
  double2 *Z = fft_lds<true, true>(X, Y, n2, Q, 1, Q, tw);
  for (int e = threadIdx.x; e < Q * n2; e += kBlock) base[(long)(e / Q) * n1 + (e % Q)] = Z[e];
  ke = wave_sum64(ke);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ke;
  __syncthreads();
  if (threadIdx.x == 0) jb.part_ke[u] = (red[0] + red[1]) + (red[2] + red[3]);
}

void launch_fft_r_rt(const StepArgs &a, int total_units, size_t lds, hipStream_t stream) {
  hipLaunchKernelGGL(k_fft_r, dim3(total_units), dim3(kBlock), lds, stream, a);
}

// ---------------------------------------------------------------------------
// probe: in-place DFT of `count` contiguous pencils of length n (one per block)
// ---------------------------------------------------------------------------
template <bool INV>
__global__ __launch_bounds__(kBlock) void k_fft_probe(double2 *data, const double2 *twg, int n) {
  extern __shared__ double2 lds[];
  double2 *A = lds, *B = lds + n, *tw = lds + 2 * n;
  double2 *base = data + (long)blockIdx.x * n;
  for (int e = threadIdx.x; e < n; e += kBlock) { tw[e] = twg[e]; A[e] = base[e]; }
  __syncthreads();
  double2 *res = fft_lds<INV, false>(A, B, n, 1, n, 1, tw);
  for (int e = threadIdx.x; e < n; e += kBlock) base[e] = res[e];
}

void launch_fft_probe(double2 *data, const double2 *tw, int n, int count, int inverse, hipStream_t stream) {
  const size_t lds = 3 * (size_t)n * sizeof(double2);
  if (inverse) hipLaunchKernelGGL(k_fft_probe<true>, dim3(count), dim3(kBlock), lds, stream, data, tw, n);
  else hipLaunchKernelGGL(k_fft_probe<false>, dim3(count), dim3(kBlock), lds, stream, data, tw, n);
}

// ---------------------------------------------------------------------------
// Allocation-free batched DFT on caller device memory (qd_fft_pencils_noalloc).
// Its one scratch buffer, the twiddle table, is `scratch`: a file-scope
// device array of kFftMaxN entries (8 KB), which every GPU the library runs
// on holds until the process ends.  Host code never hands `scratch` itself to
// a kernel: the array's device address on the current GPU comes from
// hipGetSymbolAddress, looked up once per device (with that device current)
// and kept.  Every call on one GPU rewrites that GPU's table, so calls that
// overlap on one GPU (different streams or threads) race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double2 scratch[kFftMaxN];

constexpr int kMaxDevices = 64;

// tw[e] = exp(-2 pi i e / n), e < n
__global__ __launch_bounds__(kBlock) void k_twiddles(double2 *tw, int n) {
  for (int e = threadIdx.x; e < n; e += kBlock) {
    double sn, cs;
    sincospi(-2.0 * (double)e / (double)n, &sn, &cs);
    tw[e] = make_double2(cs, sn);
  }
}

// device address of `scratch` on the current device
static double2 *scratch_address() {
  static std::mutex mu;
  static double2 *addr[kMaxDevices] = {};
  int dev = 0;
  QD_HIP(hipGetDevice(&dev));
  QD_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    QD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QD_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the FFT scratch table");
    addr[dev] = (double2 *)p;
  }
  return addr[dev];
}

void fft_pencils_noalloc(double2 *data, int n, int count, int inverse, hipStream_t stream) {
  QD_REQUIRE(data != nullptr && count > 0 && n >= 2, "bad FFT arguments (data, n >= 2, count > 0)");
  if (n > kFftMaxN)
    throw Error(QD_ERR_CAPACITY, "n = " + std::to_string(n) + " exceeds the 512 twiddles of the FFT scratch table");
  int m = n;
  for (int p : {2, 3, 5})
    while (m % p == 0) m /= p;
  QD_REQUIRE(m == 1, "n must factor into 2, 3, 5");
  init_kernel_attributes();
  double2 *tw = scratch_address();
  hipLaunchKernelGGL(k_twiddles, dim3(1), dim3(kBlock), 0, stream, tw, n);
  launch_fft_probe(data, tw, n, count, inverse, stream);
  QD_HIP(hipGetLastError());
}

void init_attrs_fft() {
  const int lim = 160 * 1024;
#define QD_FFTR_ATTR(NB)                                                                                             \
  (void)hipFuncSetAttribute((const void *)k_fft_R<false, NB, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
  (void)hipFuncSetAttribute((const void *)k_fft_R<true, NB, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);  \
  (void)hipFuncSetAttribute((const void *)k_fft_R<true, NB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  QD_FFTR_ATTR(1) QD_FFTR_ATTR(2) QD_FFTR_ATTR(4) QD_FFTR_ATTR(6) QD_FFTR_ATTR(9) QD_FFTR_ATTR(12)
  QD_FFTR_ATTR(16) QD_FFTR_ATTR(24)
#undef QD_FFTR_ATTR
  (void)hipFuncSetAttribute((const void *)k_fft_r, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_fft_probe<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_fft_probe<true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
}

}  // namespace qd
