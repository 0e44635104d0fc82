// kernels_gemm.hip - Legendre DVR <-> FBR transforms on the matrix cores.
//
// Reference: OmegaWavepacket::forward_legendre_transform / backward_...
// (omegawavepacket.cu:200-240) issue cublasZgemm with a complex B whose
// imaginary part is zero (tables built at omegawavepacket.cu:134-181).
// Viewing the interleaved complex (M x K) operand as a real (2M x K) matrix
// turns each transform into a real DGEMM with half the flops:
//   forward : FBR_b[:, omega_b..l_max] (2M x N_l) = psi_b (2M x nt) . wP_b (nt x N_l)
//   backward: psi_b (2M x nt) = FBR_b (2M x N_l) . P_b^T (N_l x nt)
// One wave computes a 16-row strip of the output for all columns with
// v_mfma_f64_16x16x4f64; the operands are issued as (B^T, A^T) so the 16x16
// accumulator tile holds C^T and every store instruction writes 16
// consecutive rows (128 B) per column.  B is staged once per workgroup in LDS.
//
// The mixed form splits both operands into FP32 hi/lo parts and accumulates
// hi*hi and hi*lo + lo*hi with v_mfma_f32_16x16x4f32 in FP32, flushing the FP32
// accumulators into FP64 every 16 k-terms.
#include "qd_internal.h"
#include "qd_complex.h"

namespace qd {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int lds_stride(int nt_tiles) { return nt_tiles * 16 + ((nt_tiles & 1) ? 0 : 16); }

// C[r, c] = sum_k A[r, k] Bs[k][c] for rows r0..r0+15, columns 0..N-1
template <int NT>
__device__ __forceinline__ void strip_f64(const double *__restrict__ A, long lda, int K, int ksteps,
                                          const double *__restrict__ Bs, int Ws, double *__restrict__ C,
                                          long ldc, int N, int r0, int rows) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const int r = r0 + lr;
  const bool rok = r < rows;
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  double an[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int k = 4 * q + lk;
    an[q] = (rok && k < K) ? A[(long)k * lda + r] : 0.0;
  }
  for (int s0 = 0; s0 < ksteps; s0 += 4) {
    double ac[4];
#pragma unroll
    for (int q = 0; q < 4; q++) ac[q] = an[q];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = 4 * (s0 + 4 + q) + lk;
      an[q] = (rok && k < K) ? A[(long)k * lda + r] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (s0 + q < ksteps) {
        const double *brow = Bs + (4 * (s0 + q) + lk) * Ws + lr;
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(brow[t * 16], ac[q], acc[t], 0, 0, 0);
      }
    }
  }
  // v_mfma_f64_16x16x4f64 accumulator layout: item it of lane (lr, lk) holds
  // D[lk + 4 it][lr], i.e. C[r0 + lr][t*16 + lk + 4 it]
  if (rok) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
#pragma unroll
      for (int it = 0; it < 4; it++) {
        const int c = t * 16 + lk + 4 * it;
        if (c < N) C[(long)c * ldc + r] = acc[t][it];
      }
    }
  }
}

// mixed precision: hi/lo split FP32 products, FP32 accumulation flushed to FP64 every 4 k-steps
template <int NT>
__device__ __forceinline__ void strip_mixed(const double *__restrict__ A, long lda, int K, int ksteps,
                                            const float *__restrict__ Bh, const float *__restrict__ Bl, int Ws,
                                            double *__restrict__ C, long ldc, int N, int r0, int rows) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const int r = r0 + lr;
  const bool rok = r < rows;
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  double an[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int k = 4 * q + lk;
    an[q] = (rok && k < K) ? A[(long)k * lda + r] : 0.0;
  }
  for (int s0 = 0; s0 < ksteps; s0 += 4) {
    float ah[4], al[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      ah[q] = (float)an[q];
      al[q] = (float)(an[q] - (double)ah[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = 4 * (s0 + 4 + q) + lk;
      an[q] = (rok && k < K) ? A[(long)k * lda + r] : 0.0;
    }
    v4f hh[NT], hl[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) { hh[t] = v4f{0.f, 0.f, 0.f, 0.f}; hl[t] = v4f{0.f, 0.f, 0.f, 0.f}; }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (s0 + q < ksteps) {
        const int off = (4 * (s0 + q) + lk) * Ws + lr;
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const float bh = Bh[off + t * 16], bl = Bl[off + t * 16];
          hh[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bh, ah[q], hh[t], 0, 0, 0);
          hl[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bl, ah[q], hl[t], 0, 0, 0);
          hl[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bh, al[q], hl[t], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
#pragma unroll
      for (int it = 0; it < 4; it++) acc[t][it] += (double)hh[t][it] + (double)hl[t][it];
    }
  }
  // v_mfma_f32_16x16x4f32 accumulator layout: item it of lane (lr, lk) holds
  // D[4 lk + it][lr], i.e. C[r0 + lr][t*16 + 4 lk + it]
  if (rok) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
#pragma unroll
      for (int it = 0; it < 4; it++) {
        const int c = t * 16 + 4 * lk + it;
        if (c < N) C[(long)c * ldc + r] = acc[t][it];
      }
    }
  }
}

template <int NT, bool FWD, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_leg(StepArgs a) {
  extern __shared__ double lds_d[];
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int u = blockIdx.x - a.unit_base[job];
  const int groups = FWD ? jb.gemm_fwd_groups : jb.gemm_bwd_groups;
  const int b = u / groups, g = u % groups;
  const int om = jb.omin + b, nl = jb.lmax - om + 1;
  const long G = (long)jb.M * jb.nt;
  const long slab0 = ((long)b * jb.L + om) * jb.M;
  const double *A = FWD ? (const double *)(jb.psi + b * G) : (const double *)(jb.fbr + slab0);
  double *C = FWD ? (double *)(jb.fbr + slab0) : (double *)(jb.psi + b * G);
  const int K = FWD ? jb.nt : nl;
  const int N = FWD ? nl : jb.nt;
  const int kp = FWD ? jb.fwd_kp : jb.bwd_kp;
  const int np = FWD ? jb.fwd_np : jb.bwd_np;
  constexpr int Ws = lds_stride(NT);
  constexpr int W = NT * 16;
  const long rows = 2L * jb.M;
  const int wave = threadIdx.x >> 6;
  const int S = jb.gemm_strips;
  if (!MIXED) {
    const double *Bg = (FWD ? jb.wP : jb.Pt) + (long)b * kp * np;
    double *Bs = lds_d;
    for (int e = threadIdx.x; e < kp * W; e += kBlock) {
      const int k = e / W, c = e % W;
      Bs[k * Ws + c] = c < np ? Bg[k * np + c] : 0.0;
    }
    __syncthreads();
    for (int s = 0; s < S; s++) {
      const long r0 = ((long)(g * S + s) * 4 + wave) * 16;
      if (r0 >= rows) break;
      strip_f64<NT>(A, rows, K, kp / 4, Bs, Ws, C, rows, N, (int)r0, (int)rows);
    }
  } else {
    const float *Bgh = (FWD ? jb.wP_hi : jb.Pt_hi) + (long)b * kp * np;
    const float *Bgl = (FWD ? jb.wP_lo : jb.Pt_lo) + (long)b * kp * np;
    float *Bh = (float *)lds_d, *Bl = Bh + kp * Ws;
    for (int e = threadIdx.x; e < kp * W; e += kBlock) {
      const int k = e / W, c = e % W;
      Bh[k * Ws + c] = c < np ? Bgh[k * np + c] : 0.f;
      Bl[k * Ws + c] = c < np ? Bgl[k * np + c] : 0.f;
    }
    __syncthreads();
    for (int s = 0; s < S; s++) {
      const long r0 = ((long)(g * S + s) * 4 + wave) * 16;
      if (r0 >= rows) break;
      strip_mixed<NT>(A, rows, K, kp / 4, Bh, Bl, Ws, C, rows, N, (int)r0, (int)rows);
    }
  }
}

template <bool FWD>
static void launch_leg(const StepArgs &a, int total_units, int nt_tiles, int kp_max, int mixed, hipStream_t s) {
  const dim3 g(total_units), t(kBlock);
#define QD_LEG_CASE(NT)                                                                              \
  case NT: {                                                                                         \
    const size_t lds = (size_t)kp_max * lds_stride(NT) * (mixed ? 2 * sizeof(float) : sizeof(double)); \
    if (mixed) hipLaunchKernelGGL((k_leg<NT, FWD, true>), g, t, lds, s, a);                          \
    else hipLaunchKernelGGL((k_leg<NT, FWD, false>), g, t, lds, s, a);                               \
    break;                                                                                           \
  }
  switch (nt_tiles) {
    QD_LEG_CASE(1)
    QD_LEG_CASE(2)
    QD_LEG_CASE(3)
    QD_LEG_CASE(4)
    QD_LEG_CASE(5)
    QD_LEG_CASE(6)
    QD_LEG_CASE(7)
    QD_LEG_CASE(8)
    QD_LEG_CASE(10)
    QD_LEG_CASE(12)
    QD_LEG_CASE(15)
    default: break;
  }
#undef QD_LEG_CASE
}

void launch_leg_fwd(const StepArgs &a, int total_units, int nt_tiles, int kp_max, int mixed, hipStream_t s) {
  launch_leg<true>(a, total_units, nt_tiles, kp_max, mixed, s);
}
void launch_leg_bwd(const StepArgs &a, int total_units, int nt_tiles, int kp_max, int mixed, hipStream_t s) {
  launch_leg<false>(a, total_units, nt_tiles, kp_max, mixed, s);
}

// probe: C (rows x n, ldc) = A (rows x k, lda) . B ([kp][np] row-major), one strip per wave
template <int NT>
__global__ __launch_bounds__(kBlock) void k_gemm_probe(const double *A, int lda, const double *Bg, int kp, int np,
                                                       double *C, int ldc, int rows, int K, int N) {
  extern __shared__ double Bs[];
  constexpr int Ws = lds_stride(NT), W = NT * 16;
  for (int e = threadIdx.x; e < kp * W; e += kBlock) {
    const int k = e / W, c = e % W;
    Bs[k * Ws + c] = c < np ? Bg[k * np + c] : 0.0;
  }
  __syncthreads();
  const int r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (r0 < rows) strip_f64<NT>(A, lda, K, kp / 4, Bs, Ws, C, ldc, N, r0, rows);
}

void launch_gemm_probe(const double *A, int lda, const double *B, int kp, int np, double *C, int ldc, int rows,
                       int k, int n, hipStream_t s) {
  const int nt = (np + 15) / 16;
  const dim3 g((rows + 63) / 64), t(kBlock);
#define QD_PROBE_CASE(NT)                                                                                  \
  case NT:                                                                                                 \
    hipLaunchKernelGGL(k_gemm_probe<NT>, g, t, (size_t)kp * lds_stride(NT) * sizeof(double), s, A, lda, B, \
                       kp, np, C, ldc, rows, k, n);                                                        \
    break;
  switch (nt) {
    QD_PROBE_CASE(1)
    QD_PROBE_CASE(2)
    QD_PROBE_CASE(3)
    QD_PROBE_CASE(4)
    QD_PROBE_CASE(5)
    QD_PROBE_CASE(6)
    QD_PROBE_CASE(7)
    QD_PROBE_CASE(8)
    default: break;
  }
#undef QD_PROBE_CASE
}

template <int NT>
static void set_leg_attrs(int lim) {
  (void)hipFuncSetAttribute((const void *)k_leg<NT, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg<NT, false, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg<NT, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg<NT, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
}

void init_attrs_gemm() {
  const int lim = 160 * 1024;
  set_leg_attrs<1>(lim); set_leg_attrs<2>(lim); set_leg_attrs<3>(lim); set_leg_attrs<4>(lim);
  set_leg_attrs<5>(lim); set_leg_attrs<6>(lim); set_leg_attrs<7>(lim); set_leg_attrs<8>(lim);
  set_leg_attrs<10>(lim); set_leg_attrs<12>(lim); set_leg_attrs<15>(lim);
  (void)hipFuncSetAttribute((const void *)k_gemm_probe<8>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
}

}  // namespace qd
