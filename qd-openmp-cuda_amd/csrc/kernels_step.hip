// kernels_step.hip - pointwise and reduction kernels of the split-operator step.
//
//   k_dvr     absorber + potential phase on the DVR grid
//             (reference: _dump_wavepacket_ evolutionCUDAaux.cu:243-252 and
//              _evolution_with_potential_ evolutionCUDAaux.cu:5-11; the
//              V(-dt/2) start phase of cudaOpenMP2.cu:344 when steps == 0)
//   k_fbr     rotational(dt/4) . Coriolis(dt/2) . rotational(dt/4) on the FBR
//             grid for every omega block at once, with the Coriolis and
//             rotational energies of the second half
//             (reference: _evolution_with_rotational_ evolutionCUDAaux.cu:33-54,
//              _evolution_with_coriolis_ :178-209, _coriolis_matrices_production_
//              :211-241, OmegaWavepacket::calculate_rotational_energy...
//              omegawavepacket.cu:397-428)
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
// FBR pass: for slab l and grid point (i, j), with x_b the FBR value of block b,
//   y_b = rot2(l, j) * sum_b1 corF(l, i)[b][b1] x_b1
// where corF(l, i) = exp(-i dt/2 l(l+1) I1_i) * V exp(-i dt/2 e / (2 m1 R_i^2)) V^T
// is rotational(dt/4) . Coriolis(dt/2) . rotational(dt/4) restricted to the
// radial factor of the rotational phase (the phase is a scalar over omega).
// With energy: Coriolis energy on x (pre-update state; the scalar rotational
// phase cancels in <x_b|C|x_b1>) and rotational energy on y.
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(kBlock) void k_fbr(StepArgs a, int with_energy) {
  __shared__ double red[4 * 2 * NB];
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int u = blockIdx.x - a.unit_base[job];
  const int l = jb.omin + u / jb.fbr_chunks;
  const int m = (u % jb.fbr_chunks) * kBlock + threadIdx.x;
  const int nb = min(jb.kmax, l) - jb.omin + 1;   // blocks with omega <= min(K, l)
  const int M = jb.M, nO = jb.nO, L = jb.L;
  double ce[NB], re[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) { ce[b] = 0.0; re[b] = 0.0; }
  if (m < M) {
    const int i = m % jb.n1, j = m / jb.n1;
    double2 x[NB];
#pragma unroll
    for (int b = 0; b < NB; b++)
      x[b] = (b < nb) ? jb.fbr[((long)b * L + l) * M + m] : make_double2(0.0, 0.0);
    if (with_energy) {
      const double I1 = jb.I1[i];
#pragma unroll
      for (int b = 0; b < NB; b++) {
        if (b < nb) {
          const double2 c = jb.cmat[l * nO + b];
          double e = c.x * cnorm(x[b]);
          if (b + 1 < nb) e += 2.0 * c.y * cdot_re(x[b], x[b + 1]);
          ce[b] = e * I1;
        }
      }
    }
    const double2 *F = jb.corF + ((long)l * jb.n1 + i) * nO * nO;
    const double2 r2 = jb.rot2[(long)l * jb.n2 + j];
    const double wrot = (double)l * (l + 1) * (jb.I1[i] + jb.I2[j]);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      if (b < nb) {
        double2 y = make_double2(0.0, 0.0);
#pragma unroll
        for (int b1 = 0; b1 < NB; b1++)
          if (b1 < nb) y = cfma(F[b * nO + b1], x[b1], y);
        y = cmul(r2, y);
        jb.fbr[((long)b * L + l) * M + m] = y;
        re[b] = wrot * cnorm(y);
      }
    }
  }
  if (with_energy) {
    double v[2 * NB];
#pragma unroll
    for (int b = 0; b < NB; b++) { v[2 * b] = ce[b]; v[2 * b + 1] = re[b]; }
    block_sum<2 * NB>(v, red);
    if (threadIdx.x == 0) {
      double *out = jb.part_fbr + (long)u * nO * 2;
      for (int b = 0; b < nO; b++) {
        out[2 * b] = b < NB ? v[2 * b] : 0.0;
        out[2 * b + 1] = b < NB ? v[2 * b + 1] : 0.0;
      }
    }
  }
}

void launch_fbr(const StepArgs &a, int total_units, int with_energy, int nb_max, hipStream_t stream) {
  const dim3 g(total_units), t(kBlock);
  if (nb_max <= 1) hipLaunchKernelGGL(k_fbr<1>, g, t, 0, stream, a, with_energy);
  else if (nb_max <= 2) hipLaunchKernelGGL(k_fbr<2>, g, t, 0, stream, a, with_energy);
  else if (nb_max <= 4) hipLaunchKernelGGL(k_fbr<4>, g, t, 0, stream, a, with_energy);
  else if (nb_max <= 6) hipLaunchKernelGGL(k_fbr<6>, g, t, 0, stream, a, with_energy);
  else if (nb_max <= 9) hipLaunchKernelGGL(k_fbr<9>, g, t, 0, stream, a, with_energy);
  else if (nb_max <= 12) hipLaunchKernelGGL(k_fbr<12>, g, t, 0, stream, a, with_energy);
  else if (nb_max <= 16) hipLaunchKernelGGL(k_fbr<16>, g, t, 0, stream, a, with_energy);
  else hipLaunchKernelGGL(k_fbr<kMaxOmega>, g, t, 0, stream, a, with_energy);
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
// Final reduction: one workgroup per job.
// ---------------------------------------------------------------------------
__device__ double block_sum_strided(const double *src, long count, long stride, double *lds) {
  double v[1] = {0.0};
  for (long q = threadIdx.x; q < count; q += kBlock) v[0] += src[q * stride];
  block_sum<1>(v, lds);
  return v[0];   // valid in thread 0
}

__global__ __launch_bounds__(kBlock) void k_reduce(StepArgs a) {
  __shared__ double red[16];   // block_sum<4> needs 4 waves x 4 values
  __shared__ double acc[2];
  const int job = blockIdx.x;   // one unit per job
  const JobDev &jb = a.jobs[job];
  const int nO = jb.nO;
  const double area = jb.dr1 * jb.dr2;
  const int slot = jb.counters[1] % jb.ring;
  double *rec = jb.obs_ring + (long)slot * nO * 5;
  if (threadIdx.x == 0) { acc[0] = 0.0; acc[1] = 0.0; }
  // FBR-pass partials: [unit][b][2], units = fbr units of this job
  const long nfbr = jb.units[KK_FBR];
  // FFT_r partials: units ordered by block, slab, column group
  const int gcols = jb.n1 / jb.fft_cols;
  long ke_off = 0;
  for (int b = 0; b < nO; b++) {
    const int om = jb.omin + b;
    const long nke = (long)(jb.L - om) * gcols;
    const double ke = block_sum_strided(jb.part_ke + ke_off, nke, 1, red);
    ke_off += nke;
    const double ce = block_sum_strided(jb.part_fbr + 2 * b, nfbr, 2L * nO, red);
    const double re = block_sum_strided(jb.part_fbr + 2 * b + 1, nfbr, 2L * nO, red);
    // DVR partials: [b][k][chunk][4], weighted by w_k
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const long per_b = (long)jb.nt * jb.obs_chunks;
    for (long q = threadIdx.x; q < per_b; q += kBlock) {
      const double wk = jb.w[q / jb.obs_chunks];
      const double *s = jb.part_obs + ((long)b * per_b + q) * 4;
      v[0] += wk * s[0]; v[1] += wk * s[1]; v[2] += wk * s[2]; v[3] += wk * s[3];
    }
    block_sum<4>(v, red);
    if (threadIdx.x == 0) {
      rec[b * 5 + 0] = v[0] * area;
      rec[b * 5 + 1] = v[1] * area;
      rec[b * 5 + 2] = ke * area / (double)jb.M;
      rec[b * 5 + 3] = re * area;
      rec[b * 5 + 4] = ce * area;
      acc[0] += v[2] * area;
      acc[1] += v[3] * area;
    }
  }
  if (threadIdx.x == 0) {
    jb.acf_ring[slot] = make_double2(acc[0], acc[1]);
    jb.counters[0] += 1;
    jb.counters[1] += 1;
  }
}

void launch_reduce(const StepArgs &a, int n_jobs, hipStream_t stream) {
  hipLaunchKernelGGL(k_reduce, dim3(n_jobs), dim3(kBlock), 0, stream, a);
}

}  // namespace qd

namespace qd {

// flux accumulation at the dividing surface (filled in by the CRP pass)
void launch_flux(const StepArgs &, int, hipStream_t) {}

}  // namespace qd
