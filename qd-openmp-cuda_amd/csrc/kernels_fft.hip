// kernels_fft.hip - LDS-staged Stockham FFT passes over the (R, r) plane of
// every FBR slab.
//
// Replaces the batched cuFFT Z2Z plan over all l_max+1 slabs
// (wavepackets1device.cu:186-201; exec omegawavepacket.cu:242-269) with three
// passes that only touch the slabs that carry data (l >= omega):
//   k_fft_R<false>  forward DFT along R (contiguous pencils)
//   k_fft_r         forward DFT along r, kinetic phase exp(-i dt T) and the
//                   kinetic-energy partial sum, inverse DFT along r
//                   (_evolution_with_kinetic_ evolutionCUDAaux.cu:13-31;
//                    _psi_times_kinitic_energy_ :56-74)
//   k_fft_R<true>   inverse DFT along R and the 1/(n1 n2) scale
//                   (cublasZdscal, omegawavepacket.cu:256-267)
// Forward uses exp(-2 pi i jk/n), inverse exp(+2 pi i jk/n), both unnormalised,
// like CUFFT_FORWARD / CUFFT_INVERSE.  Lengths factor into radices 4, 2, 3, 5.
#include "qd_internal.h"
#include "qd_complex.h"

namespace qd {

struct FftPlan {
  int nst;
  int r[16];
};

__device__ __forceinline__ void make_plan(int n, FftPlan &p) {
  p.nst = 0;
  while (n % 4 == 0) { p.r[p.nst++] = 4; n /= 4; }
  while (n % 2 == 0) { p.r[p.nst++] = 2; n /= 2; }
  while (n % 3 == 0) { p.r[p.nst++] = 3; n /= 3; }
  while (n % 5 == 0) { p.r[p.nst++] = 5; n /= 5; }
}

template <bool INV>
__device__ __forceinline__ double2 rot_i(double2 a) {   // multiply by -i (forward) / +i (inverse)
  return INV ? cmul_pi(a) : cmul_mi(a);
}

// One Stockham stage for one butterfly.
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
  if (R == 2) {
    X[0] = cadd(v[0], v[1]);
    X[1] = csub(v[0], v[1]);
  } else if (R == 4) {
    const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const double2 a2 = cadd(v[1], v[3]), a3 = rot_i<INV>(csub(v[1], v[3]));
    X[0] = cadd(a0, a2);
    X[2] = csub(a0, a2);
    X[1] = cadd(a1, a3);
    X[3] = csub(a1, a3);
  } else if (R == 3) {
    const double c3 = -0.5, s3 = 0.86602540378443864676;
    const double2 s = cadd(v[1], v[2]);
    const double2 d = rot_i<INV>(csub(v[1], v[2]));
    const double2 t = make_double2(v[0].x + c3 * s.x, v[0].y + c3 * s.y);
    X[0] = cadd(v[0], s);
    X[1] = make_double2(t.x + s3 * d.x, t.y + s3 * d.y);
    X[2] = make_double2(t.x - s3 * d.x, t.y - s3 * d.y);
  } else if (R == 5) {
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    const double2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const double2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    X[0] = cadd(v[0], cadd(a1, a2));
    const double2 t1 = make_double2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    const double2 t2 = make_double2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    const double2 u1 = rot_i<INV>(make_double2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    const double2 u2 = rot_i<INV>(make_double2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    X[1] = cadd(t1, u1);
    X[4] = csub(t1, u1);
    X[2] = cadd(t2, u2);
    X[3] = csub(t2, u2);
  }
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
  FftPlan pl;
  make_plan(n, pl);
  double2 *src = A, *dst = B;
  int Ns = 1;
  for (int s = 0; s < pl.nst; s++) {
    const int R = pl.r[s];
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

// unit -> (block b, slab l, group g), units ordered by block, slab (l >= omega_b), group
__device__ __forceinline__ void decode_slab_unit(const JobDev &jb, int u, int groups, int &b, int &l, int &g) {
  b = 0;
  for (;;) {
    const int ub = (jb.L - (jb.omin + b)) * groups;
    if (u < ub || b == jb.nO - 1) break;
    u -= ub;
    b++;
  }
  l = jb.omin + b + u / groups;
  g = u % groups;
}

__device__ double wave_sum_fft(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------
// FFT along R: fft_rows consecutive r-rows (contiguous n1-pencils) per workgroup
// ---------------------------------------------------------------------------
template <bool INV>
__global__ __launch_bounds__(kBlock) void k_fft_R(StepArgs a) {
  extern __shared__ double2 lds[];
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int u = blockIdx.x - a.unit_base[job];
  const int n1 = jb.n1, P = jb.fft_rows, M = jb.M;
  int b, l, g;
  decode_slab_unit(jb, u, jb.n2 / P, b, l, g);
  double2 *base = jb.fbr + ((long)b * jb.L + l) * M + (long)g * P * n1;
  double2 *A = lds, *B = lds + P * n1, *tw = lds + 2 * P * n1;
  for (int e = threadIdx.x; e < n1; e += kBlock) tw[e] = jb.twR[e];
  for (int e = threadIdx.x; e < P * n1; e += kBlock) A[e] = base[e];
  __syncthreads();
  double2 *res = fft_lds<INV, false>(A, B, n1, P, n1, 1, tw);
  if (INV) {
    const double s = 1.0 / (double)M;
    for (int e = threadIdx.x; e < P * n1; e += kBlock) base[e] = cscale(res[e], s);
  } else {
    for (int e = threadIdx.x; e < P * n1; e += kBlock) base[e] = res[e];
  }
}

void launch_fft_R(const StepArgs &a, int total_units, int inverse, size_t lds, hipStream_t stream) {
  if (inverse) hipLaunchKernelGGL(k_fft_R<true>, dim3(total_units), dim3(kBlock), lds, stream, a);
  else hipLaunchKernelGGL(k_fft_R<false>, dim3(total_units), dim3(kBlock), lds, stream, a);
}

// ---------------------------------------------------------------------------
// FFT along r with the kinetic phase: fft_cols consecutive R-columns per workgroup
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
  ke = wave_sum_fft(ke);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ke;
  __syncthreads();
  if (threadIdx.x == 0) jb.part_ke[u] = (red[0] + red[1]) + (red[2] + red[3]);
}

void launch_fft_r(const StepArgs &a, int total_units, size_t lds, hipStream_t stream) {
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

void init_attrs_fft() {
  const int lim = 160 * 1024;
  (void)hipFuncSetAttribute((const void *)k_fft_R<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_fft_R<true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_fft_r, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_fft_probe<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_fft_probe<true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
}

}  // namespace qd
