// kernels_step.hip - pointwise and reduction kernels of the split-operator step.
//
//   k_dvr     absorber + potential phase on the DVR grid
//             (reference: _dump_wavepacket_ evolutionCUDAaux.cu:243-252 and
//              _evolution_with_potential_ evolutionCUDAaux.cu:5-11; the
//              V(-dt/2) start phase of cudaOpenMP2.cu:344 when steps == 0)
//   k_obs     norm, potential energy and autocorrelation on the DVR grid
//             (omegawavepacket.cu:76-132)
//   k_reduce  fixed-order reduction of the per-workgroup partial sums into one
//             observable record per step; advances the device step counters
//             (replaces the host-synchronous cublasZdotc calls and
//              print_energies, cudaOpenMP2.cu:425-445)
//
// All reductions are two-level and in a fixed order, so an FP64 run is
// bitwise reproducible and a job's results do not depend on which other jobs
// share the launch.
#include "qd_internal.h"
#include "qd_complex.h"

namespace qd {

// ---------------------------------------------------------------------------
// block reductions (wavefront shuffles + LDS across the 4 waves)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sums NV values over the 256-thread block; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds /* >= 4*NV */) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const double s = wave_sum(v[q]);
    if (lane == 0) lds[wave * NV + q] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = (lds[q] + lds[NV + q]) + (lds[2 * NV + q] + lds[3 * NV + q]);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// DVR pass: psi_b <- d1 d2 exp(-i dt V) psi_b  (x exp(+i dt/2 V) on the first step)
// ---------------------------------------------------------------------------
constexpr int kDvrPerThread = 4;

__global__ __launch_bounds__(kBlock) void k_dvr(StepArgs a) {
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int u = blockIdx.x - a.unit_base[job];
  const long G = (long)jb.M * jb.nt;
  const bool first = jb.counters[0] == 0;
  const long base = (long)u * (kBlock * kDvrPerThread) + threadIdx.x;
#pragma unroll
  for (int q = 0; q < kDvrPerThread; q++) {
    const long idx = base + q * kBlock;
    if (idx < G) {
      double2 t = jb.vphase[idx];
      if (first) {
        double s, c;
        sincos(0.5 * jb.dt * jb.pot[idx], &s, &c);
        t = cmul(t, make_double2(c, s));
      }
      for (int b = 0; b < jb.nO; b++) {
        double2 *p = jb.psi + b * G + idx;
        *p = cmul(t, *p);
      }
    }
  }
}

void launch_dvr(const StepArgs &a, int total_units, hipStream_t stream) {
  hipLaunchKernelGGL(k_dvr, dim3(total_units), dim3(kBlock), 0, stream, a);
}

// ---------------------------------------------------------------------------
// DVR observables: per (block b, theta slice k, chunk) partial sums of
// |psi|^2, V |psi|^2 and conj(psi0) psi
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_obs(StepArgs a) {
  __shared__ double red[16];
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int u = blockIdx.x - a.unit_base[job];
  const int chunk = u % jb.obs_chunks;
  const int bk = u / jb.obs_chunks;             // b * nt + k
  const int k = bk % jb.nt;
  const int b = bk / jb.nt;
  const int M = jb.M;
  const int len = (M + jb.obs_chunks - 1) / jb.obs_chunks;
  const int m0 = chunk * len, m1 = min(M, m0 + len);
  const double2 *p = jb.psi + ((long)b * jb.nt + k) * M;
  const double *V = jb.pot + (long)k * M;
  const double2 *p0 = jb.psi0 ? jb.psi0 + ((long)b * jb.nt + k) * M : nullptr;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int m = m0 + threadIdx.x; m < m1; m += kBlock) {
    const double2 x = p[m];
    const double q = cnorm(x);
    v[0] += q;
    v[1] += q * V[m];
    if (p0) {
      const double2 y = p0[m];
      v[2] += y.x * x.x + y.y * x.y;
      v[3] += y.x * x.y - y.y * x.x;
    }
  }
  block_sum<4>(v, red);
  if (threadIdx.x == 0) {
    double *out = jb.part_obs + (long)u * 4;
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; out[3] = v[3];
  }
}

void launch_obs(const StepArgs &a, int total_units, hipStream_t stream) {
  hipLaunchKernelGGL(k_obs, dim3(total_units), dim3(kBlock), 0, stream, a);
}

// ---------------------------------------------------------------------------
// Final reduction: one workgroup per (job, omega block).  Each sums its
// block's partials in a fixed order and writes its part of the step record;
// the last workgroup of a job (ticket in counters[2]) advances the counters.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sum_contiguous(const double *src, long count, double *lds) {
  double v[1];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  long q = threadIdx.x;
  for (; q + 3 * kBlock < count; q += 4 * kBlock) {
    a0 += src[q];
    a1 += src[q + kBlock];
    a2 += src[q + 2 * kBlock];
    a3 += src[q + 3 * kBlock];
  }
  for (; q < count; q += kBlock) a0 += src[q];
  v[0] = (a0 + a1) + (a2 + a3);
  block_sum<1>(v, lds);
  return v[0];   // valid in thread 0
}

__global__ __launch_bounds__(kBlock) void k_reduce(StepArgs a) {
  __shared__ double red[16];   // block_sum<4> needs 4 waves x 4 values
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int b = blockIdx.x - a.unit_base[job];
  const int nO = jb.nO;
  const double area = jb.dr1 * jb.dr2;
  const int slot = jb.counters[1] % jb.ring;
  // kinetic energy: FFT_r units ordered by block, slab, column group
  const int gcols = jb.n1 / jb.fft_cols;
  long ke_off = 0;
  for (int q = 0; q < b; q++) ke_off += (long)(jb.L - (jb.omin + q)) * gcols;
  const long nke = (long)(jb.L - (jb.omin + b)) * gcols;
  const double ke = sum_contiguous(jb.part_ke + ke_off, nke, red);
  // Coriolis / rotational energies of the fused inverse-R pass: [b][2][units]
  const long ncor = jb.units[KK_FFT_R];
  const double ce = sum_contiguous(jb.part_cor + (2L * b) * ncor, ncor, red);
  const double re = sum_contiguous(jb.part_cor + (2L * b + 1) * ncor, ncor, red);
  // DVR partials: [b][k][chunk][4], weighted by w_k
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const long per_b = (long)jb.nt * jb.obs_chunks;
  for (long q = threadIdx.x; q < per_b; q += kBlock) {
    const double wk = jb.w[q / jb.obs_chunks];
    const double4 s = *(const double4 *)(jb.part_obs + ((long)b * per_b + q) * 4);
    v[0] += wk * s.x; v[1] += wk * s.y; v[2] += wk * s.z; v[3] += wk * s.w;
  }
  block_sum<4>(v, red);
  if (threadIdx.x == 0) {
    double *rec = jb.obs_ring + ((long)slot * nO + b) * 5;
    rec[0] = v[0] * area;
    rec[1] = v[1] * area;
    rec[2] = ke * area / (double)jb.M;
    rec[3] = re * area;
    rec[4] = ce * area;
    jb.acf_ring[(long)slot * nO + b] = make_double2(v[2] * area, v[3] * area);
    __threadfence();
    const int t = atomicAdd(jb.counters + 2, 1);
    if (t == nO - 1) {
      jb.counters[2] = 0;
      jb.counters[0] += 1;
      jb.counters[1] += 1;
      __threadfence();
    }
  }
}

void launch_reduce(const StepArgs &a, int total_units, hipStream_t stream) {
  hipLaunchKernelGGL(k_reduce, dim3(total_units), dim3(kBlock), 0, stream, a);
}

}  // namespace qd

namespace qd {

// plain device copy of n double2 values (HBM traffic calibration probe)
__global__ __launch_bounds__(kBlock) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n) {
  for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < n; q += (long)gridDim.x * kBlock) dst[q] = src[q];
}

void launch_copy(const double2 *src, double2 *dst, long n, hipStream_t stream) {
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(kBlock), 0, stream, src, dst, n);
}

// flux accumulation at the dividing surface (filled in by the CRP pass)
void launch_flux(const StepArgs &, int, hipStream_t) {}

}  // namespace qd
